"""Launch-shape autotuner for the mixed-precision force kernels.

For a built problem it times candidate launch shapes with HIP events on the
problem's stream (pd_time_force_passes) and keeps the fastest:

* LDS-staged tiles (the default path where tiles fit): the tile edge in
  units of dx and the workgroup size are chosen first with 8 lanes per row,
  then the lanes per row (thread-group size G) of each pass;
* gathering kernels (no LDS tiles): the lanes per row of each pass, from
  a few lanes up to one whole wave per particle.

Results are cached per (dim, model, precision, mean row length) in memory so
repeated problems of one family reuse the choice.
"""
import json
import os

_CACHE = {}
GATHER_GROUPS = (8, 16, 32, 64)
LDS_GROUPS = (4, 8, 16)
LDS_TILES = {3: (4, 5, 6), 2: (12, 16, 24)}
LDS_BLOCKS = (512, 1024)


def _key(prob, p):
    avg = prob.num_bonds / max(prob.N, 1)
    return (p.dim, p.model, p.precision, int(round(avg)))


def _time(prob, reps):
    prob.time_force_passes(1)  # first launch of this shape (code load)
    return prob.time_force_passes(reps)


def tune(prob, reps=3, verbose=False):
    """Pick and apply the fastest launch shape of `prob` (a pd_gpu1_amd
    Problem in mixed precision).  Returns the shape dict (see apply) with the
    timings of every candidate ("ms": {label: [ms_pass0, ms_pass1]})."""
    p = prob.params
    key = _key(prob, p)
    if key in _CACHE:
        shape = _CACHE[key]
        apply(prob, shape)
        return shape
    ms = {}
    shape = {"lds": False, "tile": -1, "block": [512, 512], "groups": [0, 0]}
    prob.set_lds(-1)
    prob.set_group(0)
    if prob.lds_info()[0]:
        best = None
        for tile in LDS_TILES[p.dim]:
            for block in LDS_BLOCKS:
                prob.set_tile(tile)
                prob.set_lds_block(block)
                prob.set_group(8)
                if not prob.lds_info()[0]:
                    continue
                t = _time(prob, reps)
                ms[f"tile{tile}/block{block}/G8"] = [round(t[0], 3), round(t[1], 3)]
                if verbose:
                    print(f"[autotune] tile {tile} block {block} G8: {t[0]:.3f} + {t[1]:.3f} ms", flush=True)
                if best is None or sum(t) < best[0]:
                    best = (sum(t), tile, block)
        # the workgroup size is a launch parameter: pick it per pass at that tile
        prob.set_tile(best[1])
        blk = [best[2], best[2]]
        for p_ in (0, 1):
            bt = {}
            for block in LDS_BLOCKS:
                prob.set_lds_block(block, p_)
                bt[block] = _time(prob, reps)[p_]
            blk[p_] = min(bt, key=bt.get)
            prob.set_lds_block(blk[p_], p_)
        shape.update(lds=True, tile=best[1], block=blk)
        groups = LDS_GROUPS
    else:
        prob.set_lds(0)
        groups = GATHER_GROUPS
    times = {}
    for g in groups:
        prob.set_group(g)
        times[g] = _time(prob, reps)
        ms[f"G{g}"] = [round(times[g][0], 3), round(times[g][1], 3)]
        if verbose:
            print(f"[autotune] G={g}: {times[g][0]:.3f} + {times[g][1]:.3f} ms", flush=True)
    shape["groups"] = [min(times, key=lambda g: times[g][0]), min(times, key=lambda g: times[g][1])]
    shape["ms"] = ms
    _CACHE[key] = shape
    apply(prob, shape)
    return shape


def tune_batch(params, specs, iters=100, blocks=(512, 1024), groups=(1, 4, 8), verbose=False):
    """Workgroup size and lanes per row (1 = thread per particle, 4/8 = a
    few lanes per row) of the batched ADR kernel, timed on a fresh batch of
    `specs` for `iters` iterations per candidate.  Returns (block, group, ms)."""
    from . import batch as PB

    ms = {}
    for block in blocks:
        for g in groups:
            with PB.Batch(params, specs) as b:
                b.set_group(g)
                _, _, t = b.increment(1.0e-6, 0.0, iters, block=block)
            ms[f"block{block}/G{g}"] = round(t / iters, 4)
            if verbose:
                print(f"[autotune] batch block {block} G{g}: {t / iters:.3f} ms/iteration", flush=True)
    best = min(ms, key=ms.get)
    block, g = best.split("/")
    return int(block[5:]), int(g[1:]), ms


def apply(prob, shape):
    """Set a tuned shape on a problem of the same family."""
    prob.set_lds(-1 if shape["lds"] else 0)
    prob.set_tile(shape["tile"])
    prob.set_lds_block(shape["block"][0], 0)
    prob.set_lds_block(shape["block"][1], 1)
    prob.set_group(0)
    prob.set_pass_group(0, shape["groups"][0])
    prob.set_pass_group(1, shape["groups"][1])


def dump(path):
    with open(path, "w") as fh:
        json.dump({"%s" % (k,): v for k, v in _CACHE.items()}, fh, indent=1)


if __name__ == "__main__":
    import sys
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    from pd_gpu1_amd import scenarios
    sc = scenarios.c2_lps_block()
    with sc.make() as prob:
        print(tune(prob, verbose=True))

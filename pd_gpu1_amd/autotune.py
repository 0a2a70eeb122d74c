"""Launch-shape autotuner for the mixed-precision force kernels.

For a built problem it times each candidate group size G (lanes per
particle: 4 .. 64; G = 64 is one wave per particle) with HIP events on the
problem's stream and keeps the fastest.  Results are cached per
(dim, model, mean row length) in memory so repeated problems of one family
reuse the choice.
"""
import json
import os

_CACHE = {}
CANDIDATES = (8, 16, 32, 64)


def _key(prob, p):
    avg = prob.num_bonds / max(prob.N, 1)
    return (p.dim, p.model, p.precision, int(round(avg)))


def tune_group(prob, reps=3, candidates=CANDIDATES, verbose=False):
    """Pick the fastest group size of each force pass of `prob` (a
    pd_gpu1_amd.Problem in mixed precision) and apply them.

    Returns ((g_pass0, g_pass1), {group: (ms_pass0, ms_pass1)})."""
    p = prob.params
    key = _key(prob, p)
    if key in _CACHE:
        best, times = _CACHE[key]
    else:
        times = {}
        for g in candidates:
            prob.set_group(g)
            prob.time_force_passes(1)  # first launch of this shape (code load)
            times[g] = prob.time_force_passes(reps)
            if verbose:
                print(f"[autotune] G={g}: {times[g][0]:.3f} + {times[g][1]:.3f} ms", flush=True)
        best = (min(times, key=lambda g: times[g][0]), min(times, key=lambda g: times[g][1]))
        _CACHE[key] = (best, times)
    apply(prob, best)
    return best, times


def apply(prob, best):
    prob.set_pass_group(0, best[0])
    prob.set_pass_group(1, best[1])


def dump(path):
    with open(path, "w") as fh:
        json.dump({"%s" % (k,): v for k, v in _CACHE.items()}, fh, indent=1)


if __name__ == "__main__":
    import sys
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    from pd_gpu1_amd import scenarios
    sc = scenarios.c2_lps_block()
    with sc.make() as prob:
        print(tune_group(prob, verbose=True))

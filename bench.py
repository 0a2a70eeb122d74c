#!/usr/bin/env python3
"""Throughput of the peridynamic time step on one MI355X.

Workload (default): the 3D state-based LPS block of 200 x 200 x 100 particles
(0.1 x 0.1 x 0.05 m at dx = 0.5 mm, seeded +-1% dx jitter, delta = 3.015 dx,
E = 70 GPa, nu = 0.33, rho = 2700 kg/m^3, s0 = 0.002, v_z(0) = 1e3 z,
dt = 2e-8 s) in mixed precision: FP32 bond math with compensated sums, FP64
positions, displacements and near-critical bonds.  One step = Verlet half 1,
dilatation pass, force pass, Verlet half 2, replayed from a captured HIP graph.

value = directed bonds x steps / wall time of the timed steps (bond-updates/s,
each LPS bond-update covering both sweeps); inputs are resident in HBM before
the timed region starts.  The roofline object reports the dominant kernel
(the LPS force pass): algorithmic bytes per launch / its mean HIP-event
duration against the 8 TB/s HBM3E peak.  cpu_baseline times the FP64 oracle on
the host cores over a bounded sample of the same workload (its first 41
steps), and its gpu_check compares the oracle's state after those steps with
the GPU's: the FP64 path bit for bit, the mixed path at the timed launch
shape by its error and broken bonds.

The "adr" object measures the second half of the metric, ADR wall time to a
1e-6 relative residual, on 2D PMB plates loaded in 1e-6 m strip increments:
the batched persistent ADR over a seeded plate family (sides 40..200
particles), and plain versus warm-started ADR on unseen plates (sides
200..400) over consecutive increments, with and without the trained
network.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def lps_pass_bytes(nb, n, n_int, dim=3):
    """Algorithmic bytes of the mixed LPS passes.

    Per bond and pass: col 4 + xi 4 + omega 4 + fail bit 1/8 = 12.125 B.
    Dilatation pass per particle: x0 8D + u 8D + 1/m 4 + theta write 4 + row_ptr 4.
    Force pass per particle: x0 8D + u 8D + theta 4 + 1/m 4 + f write 4D + row_ptr 4.
    """
    theta = 12.125 * nb + (16 * dim + 12) * n
    force = 12.125 * nb + (20 * dim + 12) * n_int
    return theta, force


def kernel_stream_bytes(nb, n, dim=3, lds=True, recomputed=True):
    """Bytes the LDS tile kernels themselves read per pass: per bond the
    16-bit local index word, which carries the fail bit (2 B; the gathering
    kernels read col 4 B and the fail bit 1/8 B), plus the 8-byte
    coefficient pair when it is streamed rather than recomputed; per particle
    its own 32-byte record, row_ptr (4 B), the output (4D B) and the staged
    records (counted in the measured traffic, not here)."""
    per_bond = (2.0 if lds else 4.125) + (0.0 if recomputed else 8.0)
    return per_bond * nb + (32 + 4 + 4 * dim) * n


def batch_iteration_bytes(nb, n):
    """Algorithmic bytes of one ADR iteration of one 2D specimen in the
    batched kernel (FP64): the PMB force as counted for FP64 PMB, 20.125 B per
    bond (col 4, |xi| 8, kappa 8, fail bit 1/8) and 52 B per particle (x0 16,
    u 16, F write 16, row_ptr 4), plus the ADR damping terms and update per
    particle (Integrate.cu:7-106): reads of F, F_old, v_half_old, u and b
    (80 B), writes of v_half, v_half_old, u and F_old (64 B)."""
    return 20.125 * nb + (52 + 144) * n


def pmb_bytes(nb, n, dim=3):
    """Mixed PMB: col 4 + xi 4 + kappa 4 + fail 1/8 per bond; x0 8D + u 8D + f 4D + row_ptr 4."""
    return 12.125 * nb + (20 * dim + 4) * n


def load_traffic(kernel_key):
    """HBM bytes per launch of a kernel from profiles/traffic.json (rocprofv3
    FETCH_SIZE and WRITE_SIZE runs, scaled by the copy calibration), with the
    file's provenance: the commit and command lines it was measured with."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None, None
    try:
        with open(path) as fh:
            data = json.load(fh)
    except (OSError, ValueError):
        return None, None
    kernels = data.get("kernels", {})
    src = {"file": "profiles/traffic.json", "commit": data.get("commit"), "summary": data.get("summary")}
    return kernels.get(kernel_key), src


def cpu_model():
    """Host CPU model name and the logical CPUs this process may use."""
    name = ""
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    name = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        avail = os.cpu_count()
    return {"cpu_model": name, "logical_cpus": os.cpu_count(), "cpus_available": avail}


def gpu_check(m, steps, shape):
    """The oracle's state after `steps` steps against the GPU's after the same
    steps: the FP64 path bit for bit, and the mixed path at the benchmark's
    launch shape (the timed kernels) by its relative error and broken bonds."""
    import numpy as np
    from pd_gpu1_amd import _lib, autotune, scenarios

    U = m.stack("U")
    out = {"steps": steps}
    for name, prec in (("fp64", _lib.PD_PREC_FP64), ("mixed", _lib.PD_PREC_MIXED)):
        sc = scenarios.c2_lps_block(precision=prec)
        prob = sc.make(device=0)
        if prec == _lib.PD_PREC_MIXED:
            autotune.apply(prob, shape)
        prob.run_verlet(steps)
        u = prob.get("U")
        mask = prob.fail_mask()
        prob.close()
        if prec == _lib.PD_PREC_FP64:
            out[name] = {"u_bit_exact": bool(np.array_equal(u, U)), "broken_bonds_equal": bool(np.array_equal(mask, m.fail))}
        else:
            out[name] = {"u_rel_l2": float(np.linalg.norm(u - U) / np.linalg.norm(U)),
                         "u_rel_max": float(np.max(np.abs(u - U)) / np.max(np.abs(U))),
                         "broken_bonds_differing": int((mask != m.fail).sum())}
    out["oracle_broken_bonds"] = int((m.fail == 0).sum())
    return out


def cpu_baseline(sample_dims=(200, 200, 100), evals=40, one_thread_evals=2, check_shape=None):
    """FP64 oracle (OpenMP) on a sub-block of the same lattice and physics,
    with all threads and with one (the reference's sequential CPU path).
    With check_shape (the full block only) the oracle's state after its timed
    steps also checks the GPU's (gpu_check)."""
    import numpy as np
    from oracle import oracle as O
    from oracle.runner import OracleModel
    from pd_gpu1_amd import scenarios

    nx, ny, nz = sample_dims
    sc = scenarios.c2_lps_block(nx=nx, ny=ny, nz=nz)
    m = OracleModel(scenarios.params_dict(sc.p), sc.x, sc.y, sc.z, sc.n_int)
    m.build_neighbors()
    m.init_state()
    m.lps_weights()
    m.V = [a.copy() for a in sc.v0]
    nb = len(m.col)
    m.step()  # warm caches and page in
    t0 = time.perf_counter()
    for _ in range(evals):
        m.step()
    dt = time.perf_counter() - t0
    check = gpu_check(m, evals + 1, check_shape) if check_shape is not None and sample_dims == (200, 200, 100) else None
    threads = O.get_threads()
    O.set_threads(1)
    t0 = time.perf_counter()
    for _ in range(one_thread_evals):
        m.step()
    dt1 = time.perf_counter() - t0
    O.set_threads(threads)
    del np
    return {
        "value": nb * evals / dt,
        "unit": "bond-updates/s",
        "cores": threads,
        "kind": "port",
        "sample": f"FP64 oracle, {nx}x{ny}x{nz} sub-block of the same jittered LPS lattice "
                  f"({len(sc.x)} particles, {nb} bonds), {evals} full Verlet+LPS steps, {dt:.2f} s",
        "one_thread": {"value": nb * one_thread_evals / dt1, "unit": "bond-updates/s",
                       "sample": f"the same, {one_thread_evals} steps on 1 thread, {dt1:.2f} s"},
        **cpu_model(),
        **({"gpu_check": check} if check is not None else {}),
    }


def cpu_full_configs(one_thread_steps=(50, 5, 1)):
    """The FP64 oracle over the whole runs of the plate (C0, 1000 steps), the
    Kalthoff-Winkler impact (C1, 1350 steps) and the LPS block (C2, 200
    steps) on all host threads, and a few steps of each on one thread (the
    reference's sequential CPU path).  About two minutes on a 16-thread host."""
    from oracle import oracle as O
    from oracle.runner import OracleModel
    from pd_gpu1_amd import _lib, scenarios

    rows = {}
    cases = (("c0_plate_pmb", scenarios.c0_plate_pmb()),
             ("c1_kalthoff_winkler", scenarios.c1_kalthoff_winkler()),
             ("c2_lps_block", scenarios.c2_lps_block(precision=_lib.PD_PREC_FP64)))
    for (name, sc), n1 in zip(cases, one_thread_steps):
        t0 = time.perf_counter()
        m = OracleModel(scenarios.params_dict(sc.p), sc.x, sc.y, sc.z, sc.n_int)
        m.build_neighbors()
        if sc.cyl is not None:
            m.surface_correct()
        m.init_state()
        if sc.cyl is not None:
            m.set_notches_kw(0.025)
            m.cylinder_init(sc.cyl)
        elif sc.p.model == _lib.PD_MODEL_LPS:
            m.lps_weights()
            m.V = [a.copy() for a in sc.v0]
        else:
            m.set_crack_2d(0.01, 0.0, 0.0, 0.0)
        setup = time.perf_counter() - t0
        t0 = time.perf_counter()
        for _ in range(sc.steps):
            m.step()
        run = time.perf_counter() - t0
        broken = int((m.fail == 0).sum())  # after the configuration's steps
        threads = O.get_threads()
        O.set_threads(1)
        t0 = time.perf_counter()
        for _ in range(n1):
            m.step()
        run1 = time.perf_counter() - t0
        O.set_threads(threads)
        nb = len(m.col)
        rows[name] = {"particles": len(sc.x), "bonds": nb, "steps": sc.steps, "setup_s": setup, "run_s": run,
                      "bond_updates_per_s": nb * sc.steps / run, "threads": threads,
                      "one_thread": {"steps": n1, "s": run1, "bond_updates_per_s": nb * n1 / run1},
                      "broken_bonds_at_end": broken}
        print(f"[bench] oracle {name}: {sc.steps} steps in {run:.1f} s on {threads} threads", file=sys.stderr,
              flush=True)
        del m
    return {"kind": "port", **rows, **cpu_model()}


def warm_cpu_baseline(spec, gpu_s, gpu_its, T=4096):
    """FP64 oracle plain ADR on the first increment of one evaluation plate of
    the warm-start comparison (the single-problem path's 4096-lane reduction
    tree), beside the GPU's plain ADR on the same increment."""
    from oracle import oracle as O
    from oracle.runner import adr_increment, plate_specimen
    from pd_gpu1_amd import batch as PB
    from pd_gpu1_amd import scenarios
    from pd_gpu1_amd import warmstart as W

    n, _, clen, ang = spec
    m = plate_specimen(scenarios.params_dict(PB.plate_params(W.DX)), n, clen, ang)
    t0 = time.perf_counter()
    its, _ = adr_increment(m, W.DU, 1e-6, 400000, T)
    t = time.perf_counter() - t0
    work = len(m.col) * (its + 1)
    return {"value": work / t, "unit": "bond-iterations/s", "cores": O.get_threads(), "kind": "port",
            "sample": f"FP64 oracle plain ADR, first increment of the first evaluation plate ({n} x {n} free "
                      f"particles), {its} iterations, {t:.1f} s",
            "oracle_iterations": its, "gpu_iterations": gpu_its, "gpu_plain_s": gpu_s,
            "gpu_bond_iterations_per_s": work / gpu_s if its == gpu_its else None}


def batch_cpu_baseline(specs, du, tol, max_iters, T):
    """FP64 oracle ADR, one increment, on the given specimens one after the
    other (all OpenMP threads); the rate is bond-iterations per second."""
    from oracle import oracle as O
    from oracle.runner import adr_increment, plate_specimen
    from pd_gpu1_amd import batch as PB
    from pd_gpu1_amd import scenarios

    sc = scenarios.params_dict(PB.plate_params())
    work, its_all = 0, []
    t = 0.0
    for n, _, clen, ang in specs:
        m = plate_specimen(sc, n, clen, ang)
        t0 = time.perf_counter()
        its, _ = adr_increment(m, du, tol, max_iters, T)
        t += time.perf_counter() - t0
        work += len(m.col) * (its + 1)  # force evaluations = iterations + the converged check
        its_all.append(its)
    return {"value": work / t, "unit": "bond-iterations/s", "cores": O.get_threads(), "kind": "port",
            "sample": f"FP64 oracle ADR, increment 0 of the first {len(specs)} plates of the family "
                      f"(sides {[int(s[0]) for s in specs]}), {sum(its_all)} iterations, {t:.2f} s"}


def adr_section(n_batch=64, batch_increments=2, n_eval=2, eval_increments=6, log=False, max_iters=12000,
                with_cpu=True, separation=0.05):
    """Batched ADR timing and the warm-start comparison (see module doc)."""
    import numpy as np
    from pd_gpu1_amd import batch as PB
    from pd_gpu1_amd import warmstart as W
    from pd_gpu1_amd.mpnn import Mpnn

    from pd_gpu1_amd import autotune

    out = {}
    specs = PB.plate_family(n_batch)
    block, group, tuned = autotune.tune_batch(PB.plate_params(), specs[:32])
    runs = {}
    for mode in (0, 1):
        with PB.Batch(PB.plate_params(), specs) as b:
            b.set_group(group)
            b.set_warm(mode)
            b.set_retire(1)
            b.set_tail(32)
            b.set_separation(separation)
            n_part, n_bonds, _ = b.info()
            incs = []
            for k in range(batch_increments):
                its, res, ms = b.increment(W.DU, 1e-6, max_iters, block=block)
                if k == 0:
                    parts_used = b.parts_used()
                active = its > 0
                unconv = int(((res > 1e-6) & active).sum())
                info = [b.spec_info(s_) for s_ in range(len(specs))] if k == 0 else None
                incs.append({"wall_ms": round(ms, 1), "active": int(active.sum()),
                             "iterations_median": int(np.median(its[active])) if active.any() else 0,
                             "iterations_max": int(its.max()), "unconverged": unconv,
                             "separated_so_far": int((b.retired() == 2).sum()),
                             "tail_launch_workgroups_per_specimen": b.tail_parts()})
                if info is not None:
                    evals = [int(i) + 1 for i in its]  # force evaluations: iterations + the converged check
                    incs[-1]["bond_iterations"] = int(sum(q[2] * e for q, e in zip(info, evals)))
                    incs[-1]["algorithmic_bytes"] = float(sum(batch_iteration_bytes(q[2], q[0]) * e
                                                              for q, e in zip(info, evals)))
                if log:
                    print(f"[bench] batch {'warm' if mode else 'plain'} increment {k}: {ms:.0f} ms, "
                          f"{int(active.sum())} active, iterations median {incs[-1]['iterations_median']} "
                          f"max {its.max()}, unconverged {unconv}", file=sys.stderr, flush=True)
            runs["warm" if mode else "plain"] = incs
    tot = {k: sum(i["wall_ms"] for i in v) for k, v in runs.items()}
    # bond-iterations per second of increment 0 (plain), for the CPU comparison
    inc0 = runs["plain"][0]
    gpu_rate = inc0["bond_iterations"] / (inc0["wall_ms"] / 1e3)
    gbs = inc0["algorithmic_bytes"] / (inc0["wall_ms"] / 1e3) / 1e9
    batch_roofline = {"bound": "hbm", "kernel": "k_batch_adr", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS,
                      "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4),
                      "bytes_per_specimen_iteration": "20.125 Nb + 196 N (FP64 PMB force 20.125 Nb + 52 N, "
                                                      "ADR damping terms and update 144 N)",
                      "workload": "increment 0 of the plain run (its persistent launches, tail launches "
                                  "included), every specimen's force evaluations = iterations + 1"}
    cpu = batch_cpu_baseline(specs[:16], W.DU, 1e-6, max_iters, block) if with_cpu else None
    out["batched"] = {"workload": f"{n_batch} PMB plates, sides 40..200 (seed 0x5EED0003+k), {batch_increments} "
                                  "increments of du=1e-6 m, ADR to relative residual 1e-6, persistent launches "
                                  "(the slowest specimens of an increment continue in tail launches with more "
                                  f"workgroups each); a specimen that has not converged after {max_iters} "
                                  "iterations is retired, and one whose strip reaction falls below "
                                  f"{separation:g} of its last converged reaction (fractured through) ends its "
                                  "increment and is retired",
                      "particles": n_part, "bonds": n_bonds, "plain_s": tot["plain"] / 1e3,
                      "warm_s": tot["warm"] / 1e3, "speedup": tot["plain"] / tot["warm"],
                      "increments_plain": runs["plain"], "increments_warm": runs["warm"],
                      "launch_shape": {"block": block, "lanes_per_row": group,
                                       "workgroups_per_specimen": parts_used},
                      "autotune_ms_per_iteration": tuned, "roofline": batch_roofline,
                      "increment0_bond_iterations_per_s": gpu_rate, "cpu_baseline": cpu}
    wpath = os.path.join(ROOT, "pd_gpu1_amd", "data", "mpnn_warmstart.npy")
    specs = W.eval_specs(n_eval)
    with Mpnn() as net:
        have_net = os.path.exists(wpath)
        if have_net:
            net.set_params(np.load(wpath))
        rows = np.array(W.evaluate(net, specs, eval_increments), dtype=float)
    tot = lambda sel, c: float(rows[sel, c].sum())
    every, first, later = rows[:, 1] >= 0, rows[:, 1] == 0, rows[:, 1] >= 1
    out["warm_start"] = {
        "workload": f"{n_eval} unseen PMB plates (seed 0x5EED1004+k, sides {[int(s[0]) for s in specs]}), "
                    f"{eval_increments} consecutive 1e-6 m increments each, ADR to relative residual 1e-6",
        "plain_s": tot(every, 5), "warm_no_network_s": tot(every, 6), "warm_network_s": tot(every, 7),
        "speedup_no_network": tot(every, 5) / tot(every, 6), "speedup_network": tot(every, 5) / tot(every, 7),
        "iterations": {"plain": tot(every, 2), "warm_no_network": tot(every, 3), "warm_network": tot(every, 4)},
        "first_increment_iterations": {"plain": tot(first, 2), "uniform_stretch": tot(first, 3),
                                       "network": tot(first, 4), "network_fell_back": int(rows[first, 9].sum())},
        "later_increment_iterations": {"plain": tot(later, 2), "warm": tot(later, 3)},
        "network_weights": "pd_gpu1_amd/data/mpnn_warmstart.npy" if have_net else None,
        "network_scale": W.NET_SCALE,
    }
    if with_cpu:
        out["warm_start"]["cpu_baseline"] = warm_cpu_baseline(specs[0], float(rows[0, 5]), int(rows[0, 2]))
    return out


MFMA_BF16_PEAK_TFLOPS = 2500.0


def mpnn_section(n=300, reps=20):
    """The warm-start network on the graph of one n x n plate: forward and
    training-step time, the dense (MFMA) flop rate against the BF16 peak,
    and the per-edge gathers of the message aggregation against HBM."""
    import numpy as np
    from pd_gpu1_amd import warmstart as W
    from pd_gpu1_amd.mpnn import E_FEAT, N_FEAT, OUT, Mpnn

    plate = W.Plate(n, 0.3 * n * W.DX, 0.2)
    rp, col = plate.rp, plate.col
    plate.close()
    N, E = len(rp) - 1, len(col)
    rng = np.random.default_rng(1)
    x = rng.standard_normal((N, N_FEAT)).astype(np.float32)
    ef = rng.standard_normal((E, E_FEAT)).astype(np.float32)
    tgt = rng.standard_normal((N, OUT)).astype(np.float32)
    hid, layers = 64, 3
    dense = 2.0 * N * (N_FEAT * hid + layers * (hid * hid + 2 * hid * hid) + hid * OUT)
    gathers = layers * E * (4 * hid + 4 * E_FEAT + 4)  # P row, edge features, col per edge and layer
    with Mpnn() as net:
        net.set_graph(rp, col)
        net.set_features(x, ef, tgt)
        net.forward()
        t0 = time.perf_counter()
        for _ in range(reps):
            net.forward()
        t_fwd = (time.perf_counter() - t0) / reps
        net.train_step(1e-4)
        t0 = time.perf_counter()
        for _ in range(reps):
            net.train_step(1e-4, want_loss=False)
        net.forward()
        t_train = (time.perf_counter() - t0 - t_fwd) / reps
    return {"workload": f"warm-start network (3 message-passing layers, hidden 64) on the graph of one {n} x {n} "
                        f"plate: {N} nodes, {E} directed edges; forward includes the copy of the output to the host",
            "forward_ms": t_fwd * 1e3, "train_step_ms": t_train * 1e3,
            "mfma": {"dense_flops_per_forward": dense, "achieved_tflops": dense / t_fwd / 1e12,
                     "peak_tflops": MFMA_BF16_PEAK_TFLOPS,
                     "frac": dense / t_fwd / 1e12 / MFMA_BF16_PEAK_TFLOPS},
            "aggregation_gathers": {"bytes_per_forward": gathers, "achieved_gbs": gathers / t_fwd / 1e9,
                                    "frac_of_hbm": gathers / t_fwd / 1e9 / HBM_PEAK_GBS}}


def pmb_section(steps=50, warmup=5, reps=10, fixed=None):
    """PMB force on the same 4M-particle block geometry (mixed precision);
    fixed = (tile, block, lanes) skips the autotuner."""
    from pd_gpu1_amd import _lib, autotune, scenarios

    sc = scenarios.pmb_block(precision=_lib.PD_PREC_MIXED)
    if fixed:
        t, b, g = fixed
        shape = {"lds": True, "tile": t, "block": [b, b], "groups": [g, g], "ms": {}}
    else:
        prob = sc.make(device=0)
        shape = autotune.tune(prob)
        prob.close()
    prob = sc.make(device=0)
    autotune.apply(prob, shape)
    lds_active = prob.lds_info()[0]
    prob.run_verlet(warmup)
    prob.sync()
    t0 = time.perf_counter()
    prob.run_verlet(steps)
    prob.sync()
    wall = time.perf_counter() - t0
    ms_force, _ = prob.time_force_passes(reps)
    N, nb = prob.N, prob.num_bonds
    prob.close()
    b = pmb_bytes(nb, N)
    gbs = b / (ms_force * 1e-3) / 1e9
    kname = ("k_pmb_lds<3,%d>" if lds_active else "k_pmb_mixed<3,%d>") % shape["groups"][0]
    traffic, traffic_src = load_traffic(kname.split("<")[0])
    return {"workload": "3D PMB block 200x200x100 (4M particles, same jittered lattice), mixed precision, "
                        f"{steps} Verlet steps at dt=2e-8",
            "bonds": nb, "value": nb * steps / wall, "unit": "bond-updates/s", "ms_per_step": wall * 1e3 / steps,
            "force_bond_updates_per_s": nb / (ms_force * 1e-3), "groups": shape["groups"],
            "launch_shape": {"lds_tiles": bool(lds_active), "tile_cells": shape["tile"], "block": shape["block"]},
            "roofline": {"bound": "hbm", "kernel": kname, "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4), "bytes_per_launch": b,
                         "ms_per_launch": ms_force, "traffic": traffic, "traffic_source": traffic_src,
                         "measured_frac": (round(traffic / (ms_force * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
                                           if traffic else None)}}


def small_configs(cpu=True):
    """The 2D PMB plate (C0) and the Kalthoff-Winkler impact (C1) in full:
    GPU wall time of the whole run (FP64, and mixed for C1) and, with cpu,
    the FP64 oracle on a bounded number of steps of the same run."""
    from pd_gpu1_amd import _lib, scenarios

    out = {}
    for name, sc, precs, cpu_steps in (
            ("c0_plate_pmb", scenarios.c0_plate_pmb(), (_lib.PD_PREC_FP64,), 200),
            ("c1_kalthoff_winkler", scenarios.c1_kalthoff_winkler(), (_lib.PD_PREC_FP64, _lib.PD_PREC_MIXED), 10)):
        row = {"steps": sc.steps}
        for prec in precs:
            with sc.make(precision=prec) as prob:
                nb = prob.num_bonds
                prob.sync()
                t0 = time.perf_counter()
                prob.run_verlet(sc.steps)
                prob.sync()
                wall = time.perf_counter() - t0
                prob.damage()
                broken = int((prob.fail_mask() == 0).sum())
            key = "fp64" if prec == _lib.PD_PREC_FP64 else "mixed"
            row[key] = {"wall_s": wall, "ms_per_step": wall * 1e3 / sc.steps,
                        "bond_updates_per_s": nb * sc.steps / wall, "broken_bonds": broken}
        row["bonds"] = nb
        if cpu:
            from oracle.runner import OracleModel
            m = OracleModel(scenarios.params_dict(sc.p), sc.x, sc.y, sc.z, sc.n_int)
            m.build_neighbors()
            if sc.cyl is not None:
                m.surface_correct()
            m.init_state()
            if sc.cyl is not None:
                m.set_notches_kw(0.025)
                m.cylinder_init(sc.cyl)
            else:
                m.set_crack_2d(0.01, 0.0, 0.0, 0.0)
            t0 = time.perf_counter()
            for _ in range(cpu_steps):
                m.step()
            dt = time.perf_counter() - t0
            row["cpu_fp64_oracle"] = {"bond_updates_per_s": len(m.col) * cpu_steps / dt,
                                      "sample": f"first {cpu_steps} steps", "seconds": dt,
                                      "threads": int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count()}
        out[name] = row
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--reps", type=int, default=20, help="launches per pass for the kernel timing")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-adr", action="store_true", help="skip the ADR / warm-start measurements")
    ap.add_argument("--no-pmb", action="store_true", help="skip the PMB block measurement")
    ap.add_argument("--no-small", action="store_true", help="skip the C0 plate and Kalthoff-Winkler runs")
    ap.add_argument("--no-full-run", action="store_true", help="skip the whole 200-step run of the LPS block")
    ap.add_argument("--c3-full", action="store_true",
                    help="batched ADR at the full family size (512 plates x 50 increments, plain and warm: about 17 minutes)")
    ap.add_argument("--group", type=int, default=0, help="lanes per particle (0 = autotune)")
    ap.add_argument("--shape", default="", metavar="TILE,B0,B1,G0,G1",
                    help="fixed LDS launch shape instead of autotuning: tile edge (dx), workgroup size and "
                         "lanes per row of pass 0 and pass 1")
    ap.add_argument("--pmb-shape", default="", metavar="TILE,BLOCK,G",
                    help="fixed LDS launch shape of the PMB block instead of autotuning")
    ap.add_argument("--cpu-full", action="store_true",
                    help="only time the FP64 oracle over the whole C0, C1 and C2 runs on the host (about 2 minutes)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run a plain device copy of BYTES (HBM counter reference)")
    args = ap.parse_args()
    if args.copy_calibration:
        from pd_gpu1_amd.engine import stream_copy
        gbs = stream_copy(args.copy_calibration, reps=5)
        print(json.dumps({"copy_bytes": args.copy_calibration, "reps": 5, "launches": 6, "GB_per_s": gbs}))
        return

    if args.cpu_full:
        print(json.dumps({"cpu_full_configs": cpu_full_configs()}))
        return

    if args.gpus != 1:
        print("bench.py: the engine runs on one GPU; --gpus above 1 is not supported "
              "(sharding across GPUs exists as a design only).  Use --gpus 1.", file=sys.stderr)
        sys.exit(2)
    if int(os.environ.get("RANK", "0")) != 0:
        return

    from pd_gpu1_amd import _lib, scenarios

    t_setup = time.perf_counter()
    sc = scenarios.c2_lps_block(precision=_lib.PD_PREC_MIXED)
    prob = sc.make(device=0)
    prob.sync()
    from pd_gpu1_amd import autotune
    if args.shape:
        t, b0, b1, g0, g1 = (int(v) for v in args.shape.split(","))
        shape = {"lds": True, "tile": t, "block": [b0, b1], "groups": [g0, g1], "ms": {}}
        autotune.apply(prob, shape)
    elif args.group:
        shape = {"lds": True, "tile": -1, "block": [512, 512], "groups": [args.group, args.group], "ms": {}}
        autotune.apply(prob, shape)
    else:
        shape = autotune.tune(prob)
        # the tuning launches advanced the state: start the measured run from rest
        prob.close()
        prob = sc.make(device=0)
        autotune.apply(prob, shape)
    lds_active, lds_tiles, lds_staged = prob.lds_info()
    recomputed = prob.coef_recomputed()
    groups = shape["groups"]
    prob.sync()
    t_setup = time.perf_counter() - t_setup
    N, N_int, nb = prob.N, prob.N_int, prob.num_bonds
    print(f"[bench] setup {t_setup:.2f} s: N={N} Nb={nb}", file=sys.stderr, flush=True)

    prob.run_verlet(args.warmup)
    prob.sync()
    t0 = time.perf_counter()
    prob.run_verlet(args.steps)
    prob.sync()
    wall = time.perf_counter() - t0
    ms_step = wall * 1e3 / max(args.steps, 1)
    value = nb * args.steps / wall
    print(f"[bench] {args.steps} steps in {wall:.3f} s", file=sys.stderr, flush=True)

    ms_theta, ms_force = prob.time_force_passes(args.reps)
    b_theta, b_force = lps_pass_bytes(nb, N, N_int)
    gbs_theta = b_theta / (ms_theta * 1e-3) / 1e9
    gbs_force = b_force / (ms_force * 1e-3) / 1e9
    broken = int((prob.fail_mask() == 0).sum())
    prob.close()

    # The configuration's whole run from rest: 200 steps, the fracture that
    # starts near step 100 included (breaks fall inside this timed region).
    full_run = None
    if not args.no_full_run:
        prob = sc.make(device=0)
        autotune.apply(prob, shape)
        prob.sync()
        t0 = time.perf_counter()
        prob.run_verlet(sc.steps)
        prob.sync()
        wf = time.perf_counter() - t0
        fb = int((prob.fail_mask() == 0).sum())
        prob.close()
        full_run = {"steps": sc.steps, "wall_s": wf, "ms_per_step": wf * 1e3 / sc.steps,
                    "bond_updates_per_s": nb * sc.steps / wf, "broken_bonds_at_end": fb}
        print(f"[bench] full {sc.steps}-step run in {wf:.3f} s, {fb} broken bonds", file=sys.stderr, flush=True)

    kname = "k_lps_force_lds<3,%d>" if lds_active else "k_lps_force_mixed<3,%d>"
    traffic, traffic_src = load_traffic(kname.split("<")[0])
    roofline = {
        "bound": "hbm",
        "kernel": kname % groups[1],
        "achieved": round(gbs_force, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(gbs_force / HBM_PEAK_GBS, 4),
        "traffic": traffic,
        "traffic_source": traffic_src,
        # what the profile shows limits the kernel: it moves fewer bytes than
        # the algorithmic count (LDS-staged 32-byte records, 2-byte index
        # words), and VALU issue is near every cycle (SQ_INSTS_VALU against
        # 256 CUs x kernel cycles: 0.97 force, 0.96 dilatation); the
        # dilatation pass's frac passes 1 because its algorithmic count is
        # what a streaming pass would read
        "limiter": "valu (rocprof: VALU issue 0.99 of cycles; HBM traffic below the algorithmic bytes)",
        "measured_frac": (round(traffic / (ms_force * 1e-3) / 1e9 / HBM_PEAK_GBS, 4) if traffic else None),
        "bytes_per_launch": b_force,
        "ms_per_launch": ms_force,
        "kernel_stream_bytes_per_launch": kernel_stream_bytes(nb, N_int, lds=lds_active, recomputed=recomputed),
        "coefficients_recomputed": recomputed,
        "theta_pass": {"achieved": round(gbs_theta, 1), "frac": round(gbs_theta / HBM_PEAK_GBS, 4),
                       "bytes_per_launch": b_theta, "ms_per_launch": ms_theta},
    }
    out = {
        "metric": "bond-updates/s & %HBM peak on 1 MI355X; ADR wall time to 1e-6 resid vs plain",
        "value": value,
        "unit": "bond-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic (seeded jittered lattice)",
        "config": {"workload": "3D LPS block 200x200x100 (4M particles), dx=0.5mm, delta=3.015dx, "
                               f"mixed precision, {args.steps} timed Verlet steps at dt=2e-8 after {args.warmup} "
                               f"untimed (the configuration runs {sc.steps}; its whole run is full_run)",
                   "particles": N, "bonds": nb, "parallelism": "single GPU",
                   "force_bond_updates_per_s": nb / ((ms_theta + ms_force) * 1e-3),
                   "groups": list(groups),
                   "launch_shape": {"lds_tiles": lds_active, "tiles": lds_tiles, "max_staged_records": lds_staged,
                                    "tile_cells": shape["tile"], "block": shape["block"]},
                   "autotune_ms": shape.get("ms", {}),
                   "broken_bonds_after_run": broken,
                   # the kernels visit broken bonds too; counting only the bonds
                   # still intact at the end gives a lower bound of the rate
                   "intact_bond_updates_per_s_lower_bound": (nb - broken) * args.steps / wall},
        "roofline": roofline,
        "full_run": full_run,
    }
    if not args.no_pmb:
        t = time.perf_counter()
        out["pmb_block"] = pmb_section(fixed=tuple(int(v) for v in args.pmb_shape.split(",")) if args.pmb_shape
                                       else None)
        print(f"[bench] PMB block {time.perf_counter() - t:.1f} s", file=sys.stderr, flush=True)
    if not args.no_small:
        t = time.perf_counter()
        out["other_configs"] = small_configs(cpu=not args.no_cpu_baseline)
        print(f"[bench] C0 and Kalthoff-Winkler runs {time.perf_counter() - t:.1f} s", file=sys.stderr, flush=True)
    if not args.no_adr:
        t = time.perf_counter()
        try:
            out["mpnn"] = mpnn_section()
        except Exception as e:  # reported, never fatal for the main line
            out["mpnn"] = {"error": str(e)}
        with_cpu = not args.no_cpu_baseline
        out["adr"] = (adr_section(n_batch=512, batch_increments=50, log=True, max_iters=6000, with_cpu=with_cpu)
                      if args.c3_full else adr_section(with_cpu=with_cpu))
        print(f"[bench] ADR measurements {time.perf_counter() - t:.1f} s", file=sys.stderr, flush=True)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(check_shape=shape)
    print(json.dumps(out))


if __name__ == "__main__":
    main()

"""Mixed precision (FP32 bond math, FP64 state and near-critical bonds)
against the FP64 path on the same runs: the error per field.

FP64 mode is itself pinned to the oracle (bit-exact PMB, LPS within 1e-12),
so these bounds are the mixed mode's error against the FP64 algorithm.  In
a run with fracture the two trajectories may break a few different bonds
once their states differ by FP32 rounding; each break decision is still
exact for the state it is taken in (FP64 re-evaluation near s0)."""
import dataclasses

import numpy as np
import pytest

from pd_gpu1_amd import _lib, scenarios

pytestmark = pytest.mark.gpu


def _run(sc, precision, steps):
    with sc.make(precision=precision) as prob:
        prob.run_verlet(steps)
        prob.damage()
        return prob.get("U"), prob.get("V"), prob.get("DMG")[: sc.n_int], prob.fail_mask()


def _report(name, a, b):
    """Per field: max |mixed - fp64| / max |fp64| and the same ratio of L2 norms."""
    out = {}
    for f, x, y in zip(("U", "V", "DMG"), a[:3], b[:3]):
        out[f] = float(np.max(np.abs(x - y)) / max(np.max(np.abs(y)), 1e-300))
        out[f + "_l2"] = float(np.linalg.norm(x - y) / max(np.linalg.norm(y), 1e-300))
    out["bonds_differing"] = int((a[3] != b[3]).sum())
    out["bonds_broken_fp64"] = int((b[3] == 0).sum())
    print(name, out, flush=True)
    return out


def test_kalthoff_winkler_mixed_vs_fp64():
    # Before the impactor reaches the plate (~36 steps) the difference is the
    # FP32 bond arithmetic alone.  Contact is a discontinuous projection: a
    # particle that touches the impactor in one run and not the other differs
    # by a whole velocity jump, so after contact velocity is bounded in L2
    # away from the impact zone (|x| > 30 mm, outside the 25 mm cylinder).
    sc = scenarios.c1_kalthoff_winkler(steps=30)
    e = _report("C1 Kalthoff-Winkler 30 steps (before contact)", _run(sc, _lib.PD_PREC_MIXED, 30),
                _run(sc, _lib.PD_PREC_FP64, 30))
    assert e["U_l2"] < 1e-5 and e["V_l2"] < 1e-4 and e["bonds_differing"] == 0
    sc = scenarios.c1_kalthoff_winkler(steps=300)
    mixed, fp64 = _run(sc, _lib.PD_PREC_MIXED, 300), _run(sc, _lib.PD_PREC_FP64, 300)
    e = _report("C1 Kalthoff-Winkler 300 steps", mixed, fp64)
    far = np.abs(sc.x) > 0.030
    dv = mixed[1][:, far] - fp64[1][:, far]
    v_far = float(np.linalg.norm(dv) / np.linalg.norm(fp64[1][:, far]))
    print("C1 300 steps: V_l2 away from the impact zone", v_far, flush=True)
    assert e["U_l2"] < 1e-3 and v_far < 5e-2
    # Which bonds break under the impactor follows the contact events, so it
    # moves with any perturbation of FP32 size; for scale, the FP64 run with
    # the impactor speed changed by one FP32 rounding (2^-24 relative):
    cyl = list(sc.cyl)
    cyl[3] *= 1.0 + 2.0 ** -24
    pert = _run(dataclasses.replace(sc, cyl=tuple(cyl)), _lib.PD_PREC_FP64, 300)
    p = _report("C1 300 steps, FP64 with the impactor speed perturbed by 2^-24", pert, fp64)
    # mixed precision moves the trajectory no further than that perturbation
    # does (measured: 32 against 40 bonds, U_l2 3.2e-4 against 2.6e-4)
    assert e["bonds_differing"] <= 4 * max(p["bonds_differing"], 10)
    assert e["U_l2"] <= 4 * max(p["U_l2"], 1e-6)
    assert e["bonds_differing"] <= 1e-2 * max(e["bonds_broken_fp64"], 1000)


def test_lps_block_mixed_vs_fp64():
    sc = scenarios.c2_lps_block(nx=60, ny=60, nz=30)
    e = _report("C2 LPS block 60x60x30, 200 steps", _run(sc, _lib.PD_PREC_MIXED, 200),
                _run(sc, _lib.PD_PREC_FP64, 200))
    assert e["U"] < 1e-3 and e["U_l2"] < 1e-4 and e["V_l2"] < 1e-2
    assert e["bonds_differing"] <= 1e-3 * max(e["bonds_broken_fp64"], 1000)

"""The three reference solver entry points and the `pd_parallel` command line
(main.cu:11-50) on the GPU, checked against the same stage sequence driven
through the Python binding, and their output format (Save.cu:6-140)."""
import os
import subprocess

import numpy as np
import pytest

import pd_gpu1_amd as pd
from pd_gpu1_amd import _lib, scenarios

pytestmark = pytest.mark.gpu

CLI = os.path.join(os.path.dirname(_lib.LIB_PATH), "pd_parallel")


def run_cli(tmp_path, *args):
    r = subprocess.run([CLI, *map(str, args)], cwd=tmp_path, capture_output=True, text=True, timeout=240)
    return r.returncode, r.stdout, (tmp_path / "disp.txt")


def test_cli_rejects_cpu_device(tmp_path):
    rc, out, _ = run_cli(tmp_path, 1, 2, 0.0005)
    assert rc == _lib.PD_E_INVALID


def test_crack_grow_solver_matches_stage_sequence(tmp_path):
    rc, out, disp = run_cli(tmp_path, 2, 2, 0.0005)
    assert rc == 0, out
    text = disp.read_text().splitlines()
    # Solve_Plate_Crack_Grow.cu:30-171 at dx = 0.5 mm
    p, (nx, ny, _), n = pd.params(2, _lib.PD_MODEL_LPS, 1.3367e-8, 0.0005, thick=0.0001, length=0.05,
                                  height=0.05, emod=192e9, pratio=0.3, dens=7850, s0=0.04472)
    p.n_top = p.n_bot = 3 * nx
    p.strip_velocity = 20.0
    x, y = pd.gen_plate_strips(nx, ny, p.dx)
    with pd.Problem(p) as prob:
        prob.set_positions(x, y, None, n_int=nx * ny)
        prob.build_neighbors()
        prob.init_state()
        prob.set_crack_2d(0.01, 0.0, 0.0, 0.0)
        prob.lps_weights()
        prob.run_verlet(1250)
        prob.damage()
        prob.save_text(tmp_path / "py.txt")
        u = prob.get("U")
    assert text == (tmp_path / "py.txt").read_text().splitlines()
    assert len(text) == nx * ny
    row = text[5].split()
    assert len(row) == 3 and float(row[0]) == pytest.approx(u[0][5], rel=1e-5, abs=1e-12)


def test_plate_tensile_solver_runs_adr(tmp_path):
    rc, out, disp = run_cli(tmp_path, 2, 1, 0.05)
    assert rc == 0, out
    rows = np.loadtxt(disp)
    # 20 x 10 plate pulled by +-0.001 E/dx on its edge columns: ux antisymmetric in x
    assert rows.shape == (200, 3)
    ux = rows[:, 0].reshape(10, 20)
    assert np.all(ux[:, -1] > 0) and np.all(ux[:, 0] < 0)
    assert np.allclose(ux[:, ::-1], -ux, rtol=1e-4, atol=1e-12)


def test_kalthoff_winkler_solver_layout(tmp_path):
    rc, out, disp = run_cli(tmp_path, 2, 3, 0.004)
    assert rc == 0, out
    lines = disp.read_text().splitlines()
    # z-layer 0 on the full nx x ny grid, notch cells written as "nan nan 0 nan"
    nx, ny = 51, 26
    assert len(lines) == nx * ny
    assert sum(1 for s in lines if s == "nan nan 0 nan") > 0
    vals = np.array([[float(t) for t in s.split()] for s in lines if not s.startswith("nan")])
    assert np.all(np.isfinite(vals)) and vals.shape[1] == 4
    assert vals[:, 1].min() < 0  # the impactor pushed the top edge down


def test_crack_tip_advances_in_the_pulled_plate():
    from pd_gpu1_amd import observables as ob
    sc = scenarios.c0_plate_pmb()
    nx, ny, _ = sc.nxyz
    with sc.make() as prob:
        times, lengths = ob.track_crack(prob, nx, ny, sc.p.dx, sc.p.dt, 1000, every=100)
    print("crack half-length", lengths)
    assert lengths[0] >= 0.005 - 2 * sc.p.dx       # the 10 mm pre-crack
    assert lengths[-1] > lengths[0]                 # the crack has grown
    assert ob.crack_speed(times, lengths) > 0.0


def test_kalthoff_winkler_crack_angles():
    # Both notch cracks of the impact run leave their tips at a mode-II kink
    # angle (the experiment's is about 70 degrees from the notch) and mirror
    # each other about the centre line, as the specimen and the impactor do.
    from pd_gpu1_amd import observables as ob
    sc = scenarios.c1_kalthoff_winkler()
    with sc.make(precision=_lib.PD_PREC_MIXED) as prob:
        angles = []
        for steps in (900, 450):
            prob.run_verlet(steps)
            prob.damage()
            angles.append(ob.kw_crack_angles(sc.x, sc.y, prob.get("DMG"), 0.025, sc.p.dx))
    print("KW crack angles (left, right) after 900 and 1350 steps:", angles, flush=True)
    left, right = angles[-1]
    assert np.isfinite(left) and np.isfinite(right)
    assert 45.0 < left < 85.0 and 45.0 < right < 85.0
    assert abs(left - right) < 2.0

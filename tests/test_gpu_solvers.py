"""The three reference solver entry points and the `pd_parallel` command line
(main.cu:11-50) on the GPU, checked against the FP64 oracle driven through
each solver's scenario constants and stage order: the final state bit for
bit (through the binary state file the command line's 5th argument names)
and disp.txt line for line against a restatement of Save.cu's output
(std::ostream defaults: six significant digits, the Kalthoff-Winkler grid
with 'nan nan 0 nan' notch cells)."""
import os
import subprocess

import numpy as np
import pytest

from pd_gpu1_amd import _lib, scenarios
from oracle import oracle as O
from oracle.runner import OracleModel

pytestmark = pytest.mark.gpu

CLI = os.path.join(os.path.dirname(_lib.LIB_PATH), "pd_parallel")


def run_cli(tmp_path, *args):
    r = subprocess.run([CLI, *map(str, args)], cwd=tmp_path, capture_output=True, text=True, timeout=240)
    return r.returncode, r.stdout, (tmp_path / "disp.txt")


def test_cli_rejects_cpu_device(tmp_path):
    rc, out, _ = run_cli(tmp_path, 1, 2, 0.0005)
    assert rc == _lib.PD_E_INVALID


# ------------------------------------------------- solvers against the oracle
def read_checkpoint(path):
    """The binary state pd_checkpoint_write writes: magic, {step, dim, N,
    N_int, Nb}, U V A VH VHO FO F as [dim][N] doubles, fail words, cylinder
    {vel, disp, acc, force}."""
    raw = open(path, "rb").read()
    assert raw[:8] == b"PDMI355X"
    step, D, N, n_int, nb = np.frombuffer(raw, np.int64, 5, 8)
    off = 48
    out = {"N": int(N), "N_int": int(n_int), "Nb": int(nb)}
    for name in ("U", "V", "A", "VH", "VHO", "FO", "F"):
        out[name] = np.frombuffer(raw, np.float64, D * N, off).reshape(D, N)
        off += 8 * D * N
    nw = (nb + 31) // 32
    words = np.frombuffer(raw, np.uint32, nw, off)
    off += 4 * nw
    out["fail"] = np.unpackbits(words.view(np.uint8), bitorder="little")[:nb]
    cyl = np.frombuffer(raw, np.float64, 12, off)
    out["cyl"] = {"vel": cyl[0:3], "disp": cyl[3:6], "acc": cyl[6:9]}
    return out


def g(v):
    """std::ostream << double with the default flags, precision 6 (Save.cu:27,30)."""
    return "%g" % v


def save_plate_lines(m):
    """save_disp_cpu, Save.cu:48-60: N_int rows of ux uy [uz] dmg."""
    U = m.stack("U")
    if m.dim == 2:
        return [f"{g(U[0][i])} {g(U[1][i])} {g(m.dmg[i])}" for i in range(m.N_int)]
    return [f"{g(U[0][i])} {g(U[1][i])} {g(U[2][i])} {g(m.dmg[i])}" for i in range(m.N_int)]


def save_kw_lines(m, nx, ny, dx, rad):
    """save_kw_cpu, Save.cu:105-140: z-layer 0 over the nx x ny grid in
    particle order, the notch cells (y >= 0 inside the two slots) as
    'nan nan 0 nan'."""
    U, out, num = m.stack("U"), [], 0
    for j in range(ny):
        for i in range(nx):
            tx = (-(nx - 1) / 2.0 + i) * dx
            ty = (-(ny - 1) / 2.0 + j) * dx
            flag = (-rad - 1.1 * dx < tx < -rad + 0.1 * dx) or (rad - 0.1 * dx < tx < rad + 1.1 * dx)
            if flag and ty >= -1.0e-10:
                out.append("nan nan 0 nan")
            else:
                out.append(f"{g(U[0][num])} {g(U[1][num])} {g(U[2][num])} {g(m.dmg[num])}")
                num += 1
    return out


def oracle_sc(p, dim, model, dt, dx, emod, pratio, dens, s0, **extra):
    sc = dict(dim=dim, model=model, dx=dx, delta=p.delta, vol=p.vol, bc=p.bc, mass=p.mass, K=p.K, G=p.G,
              emod=emod, pratio=pratio, dens=dens, s0=s0, dt=dt, break_y_min=-np.inf, reduce_lanes=4096,
              theta_skip_broken=1)
    sc.update(extra)
    return sc


def compare_state(ck, m, names):
    for name in names:
        assert np.array_equal(ck[name], m.stack(name)), name
    assert np.array_equal(ck["fail"], m.fail)


def test_cli_plate_tensile_matches_oracle(tmp_path):
    # Solve_Plate_Tensile.cu:30-138 at dx = 0.02 m: constants of :40-41,
    # coordinates, neighbours, surface correction, static initial state, the
    # edge body force (:19-26), then exactly nt = 1000 ADR iterations
    dx = 0.02
    rc, out, disp = run_cli(tmp_path, 2, 1, dx, 0, tmp_path / "state.bin")
    assert rc == 0, out
    p = O.cal_parameter(2, 1, 1.0, dx, 0.01, 1.0, 0.5, 0.0, 200.0e9, 1.0 / 3.0, 7850, 1.0)
    x, y = O.coord_plate(p.nx, p.ny, dx)
    m = OracleModel(oracle_sc(p, 2, 1, 1.0, dx, 200.0e9, 1.0 / 3.0, 7850, 1.0), x, y)
    m.build_neighbors()
    m.surface_correct()
    m.init_state()
    for i in range(p.ny):
        m.B[0][i * p.nx] = -0.001 * 200.0e9 / dx
        m.B[0][i * p.nx + p.nx - 1] = 0.001 * 200.0e9 / dx
    for it in range(1000):
        m.adr_iterate(it)
    m.damage()
    ck = read_checkpoint(tmp_path / "state.bin")
    assert ck["N"] == p.nx * p.ny == 1250
    compare_state(ck, m, ("U", "VH", "VHO", "FO", "F"))
    assert disp.read_text().splitlines() == save_plate_lines(m)
    # physics: the pulled plate stretches antisymmetrically in x
    ux = m.U[0].reshape(p.ny, p.nx)
    assert np.all(ux[:, -1] > 0) and np.allclose(ux[:, ::-1], -ux, rtol=1e-6, atol=1e-15)


def test_cli_crack_grow_matches_oracle(tmp_path):
    # Solve_Plate_Crack_Grow.cu:30-171 at dx = 0.5 mm: constants of :38-39,
    # strips (N += 6 nx, :40), the 10 mm crack (:112), weights, then 1250
    # steps of quick load, half 1, LPS force, half 2 (:126-149)
    dx = 0.0005
    rc, out, disp = run_cli(tmp_path, 2, 2, dx, 0, tmp_path / "state.bin")
    assert rc == 0, out
    p = O.cal_parameter(2, 2, 1.3367e-8, dx, 0.0001, 0.05, 0.05, 0.0, 192.0e9, 0.3, 7850, 0.04472)
    x, y = O.coord_plate_strips(p.nx, p.ny, dx)
    sc = oracle_sc(p, 2, 2, 1.3367e-8, dx, 192.0e9, 0.3, 7850, 0.04472, n_top=3 * p.nx, n_bot=3 * p.nx,
                   strip_velocity=20.0)
    m = OracleModel(sc, x, y, None, p.nx * p.ny)
    m.build_neighbors()
    m.init_state()
    m.set_crack_2d(0.01, 0.0, 0.0, 0.0)
    m.lps_weights()
    for _ in range(1250):
        m.step()
    m.damage()
    ck = read_checkpoint(tmp_path / "state.bin")
    assert ck["N"] == len(x) == 10600 and ck["N_int"] == 10000
    compare_state(ck, m, ("U", "V", "A"))
    assert disp.read_text().splitlines() == save_plate_lines(m)
    assert m.dmg.max() > 0.3


def test_cli_kalthoff_winkler_matches_oracle(tmp_path):
    # Solve_Kalthoff_Winkler.cu:268-441 at dx = 4 mm with the shipped
    # constants (E = 191 GPa, s0 = 0.01, :280-281), the grid grown by one in x
    # and y (:282-284), surface correction, notches, cylinder (:285-291) and
    # 1350 steps of the loop of :407-432
    dx = 0.004
    rc, out, disp = run_cli(tmp_path, 2, 3, dx, 0, tmp_path / "state.bin")
    assert rc == 0, out
    p = O.cal_parameter(3, 1, 8.7e-8, dx, 0.001, 0.2, 0.1, 0.009, 191.0e9, 0.25, 8000, 0.01)
    nx, ny, nz = p.nx + 1, p.ny + 1, p.nz
    rad, h = 0.025, 0.05
    x, y, z = O.coord_kw(nx, ny, nz, dx, rad)
    m = OracleModel(oracle_sc(p, 3, 1, 8.7e-8, dx, 191.0e9, 0.25, 8000, 0.01, break_y_min=-0.04), x, y, z)
    m.build_neighbors()
    m.surface_correct()
    m.init_state()
    m.set_notches_kw(rad)
    m.cylinder_init((rad, 1.57, h, -32.0, 0.0, 0.1 / 2.0 + h / 2.0 + 0.1 * dx, 0.0))
    for _ in range(1350):
        m.step()
    m.damage()
    ck = read_checkpoint(tmp_path / "state.bin")
    assert (nx, ny, nz) == (51, 26, 2) and ck["N"] == len(x)
    compare_state(ck, m, ("U", "V", "A"))
    for k in ("vel", "disp", "acc"):
        assert np.array_equal(ck["cyl"][k], m.cst[k]), k
    lines = disp.read_text().splitlines()
    assert lines == save_kw_lines(m, nx, ny, dx, rad)
    assert sum(1 for s in lines if s == "nan nan 0 nan") > 0
    assert m.U[1].min() < 0  # the impactor pushed the top edge down


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

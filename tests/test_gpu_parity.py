"""GPU parity: the HIP engine against the FP64 oracle on the same inputs.

Integer/index results (CSR, broken-bond sets) must be bit-exact.  FP64-mode
PMB, LPS, Verlet, ADR and contact are bit-exact too (same operand order, no
FMA, same fixed reduction tree; the LPS influence weights are computed on the
host with the same C library exp).  Mixed mode is compared at the FP32
tolerances stated in each test.
"""
import numpy as np
import pytest

import pd_gpu1_amd as pd
from pd_gpu1_amd import _lib, scenarios
from oracle import oracle as O
from oracle.runner import OracleModel

pytestmark = pytest.mark.gpu


def _oracle_for(sc, brute=False):
    m = OracleModel(scenarios.params_dict(sc.p), sc.x, sc.y, sc.z, sc.n_int)
    m.build_neighbors(brute=brute)
    return m


def _relerr(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


# ------------------------------------------------------------------ setup
def test_csr_and_volcorr_c0_bit_exact():
    sc = scenarios.c0_plate_pmb()
    print("case: c0 plate N=%d" % len(sc.x), flush=True)
    with sc.make() as prob:
        rp, col = prob.csr()
        idist, fac = prob.get("IDIST"), prob.get("FAC")
        mask = prob.fail_mask()
    m = _oracle_for(sc, brute=True)
    m.init_state()
    m.set_crack_2d(0.01, 0.0, 0.0, 0.0)
    assert len(col) == 289420
    assert np.array_equal(rp, m.row_ptr) and np.array_equal(col, m.col)
    assert np.array_equal(idist, m.idist) and np.array_equal(fac, m.fac)
    assert np.array_equal(mask, m.fail)
    assert (mask == 0).sum() > 0


def test_csr_kw_and_surface_bit_exact():
    sc = scenarios.c1_kalthoff_winkler(delta_factor=3.015)
    print("case: kw N=%d" % len(sc.x), flush=True)
    with sc.make() as prob:
        rp, col = prob.csr()
        scr = prob.get("SCR")
        mask = prob.fail_mask()
    m = _oracle_for(sc)
    m.surface_correct()
    m.init_state()
    m.set_notches_kw(0.025)
    assert len(sc.x) == 180873 and len(col) == 18809054
    assert np.array_equal(rp, m.row_ptr) and np.array_equal(col, m.col)
    assert np.array_equal(scr, m.scr)
    assert np.array_equal(mask, m.fail)


def test_csr_knife_edge_delta_3dx():
    # delta = 3 dx puts 30 lattice offsets exactly on the horizon; the FP64
    # predicate decides them, identically on both sides.
    sc = scenarios.c1_kalthoff_winkler(delta_factor=3.0, width=0.005)
    with sc.make() as prob:
        rp, col = prob.csr()
    m = _oracle_for(sc)
    assert np.array_equal(rp, m.row_ptr) and np.array_equal(col, m.col)


def test_csr_jittered_block_bit_exact():
    sc = scenarios.c2_lps_block(nx=40, ny=36, nz=20, precision=_lib.PD_PREC_FP64)
    with sc.make() as prob:
        rp, col = prob.csr()
    m = _oracle_for(sc)
    assert np.array_equal(rp, m.row_ptr) and np.array_equal(col, m.col)


def test_capacity_error_is_reported():
    sc = scenarios.c0_plate_pmb(scale=0.2)
    p = sc.p
    p.max_neighbors = 8
    prob = pd.Problem(p, 0)
    prob.set_positions(sc.x, sc.y, None, n_int=sc.n_int)
    with pytest.raises(pd.PdError) as ei:
        prob.build_neighbors()
    assert ei.value.code == _lib.PD_E_CAPACITY
    prob.close()


# ------------------------------------------------------------------- force
def test_pmb2d_force_bit_exact_random_states():
    sc = scenarios.c0_plate_pmb()
    m = _oracle_for(sc)
    m.init_state()
    m.set_crack_2d(0.01, 0.0, 0.0, 0.0)
    rng = np.random.default_rng(7)
    with sc.make() as prob:
        for trial in range(3):
            u = rng.uniform(-4e-6, 4e-6, size=(2, len(sc.x)))
            prob.set("U", u)
            m.U = [u[0].copy(), u[1].copy()]
            prob.force()
            m.force()
            assert np.array_equal(prob.get("F"), m.stack("F")), trial
            assert np.array_equal(prob.fail_mask(), m.fail), trial
        assert (m.fail == 0).sum() > 400  # some bonds broke during the trials


def test_c0_verlet_run_bit_exact_checkpoints():
    sc = scenarios.c0_plate_pmb(steps=1000)
    m = _oracle_for(sc)
    m.init_state()
    m.set_crack_2d(0.01, 0.0, 0.0, 0.0)
    with sc.make() as prob:
        done = 0
        for chk in (1, 250, 500, 1000):
            prob.run_verlet(chk - done)
            for _ in range(chk - done):
                m.step()
            done = chk
            print("c0 checkpoint", chk, flush=True)
            assert np.array_equal(prob.get("U"), m.stack("U")), chk
            assert np.array_equal(prob.get("V"), m.stack("V")), chk
            assert np.array_equal(prob.fail_mask(), m.fail), chk
        prob.damage()
        m.damage()
        assert np.array_equal(prob.get("DMG")[: sc.n_int], m.dmg[: sc.n_int])
        # bonds broke during the run, beyond the pre-crack
        m0 = _oracle_for(sc)
        m0.init_state()
        m0.set_crack_2d(0.01, 0.0, 0.0, 0.0)
        assert (m.fail == 0).sum() > (m0.fail == 0).sum()


def test_body_force_set_between_verlet_runs_bit_exact():
    # The Verlet passes skip the body force while it is zero everywhere; a
    # body force set after steps ran from a captured graph (which skipped
    # it) must take effect in the next run, as it does in the oracle.
    sc = scenarios.c0_plate_pmb(scale=0.6)
    m = _oracle_for(sc)
    m.init_state()
    m.set_crack_2d(0.01 * 0.6, 0.0, 0.0, 0.0)  # the scenario's crack at this scale
    rng = np.random.default_rng(23)
    b = rng.uniform(-1e9, 1e9, size=(2, len(sc.x)))
    with sc.make() as prob:
        prob.run_verlet(12)
        for _ in range(12):
            m.step()
        assert np.array_equal(prob.get("U"), m.stack("U"))
        prob.set("B", b)
        m.B = [a.copy() for a in b]
        prob.run_verlet(12)
        for _ in range(12):
            m.step()
        assert np.array_equal(prob.get("U"), m.stack("U"))
        assert np.array_equal(prob.get("V"), m.stack("V"))
        assert np.array_equal(prob.fail_mask(), m.fail)


def test_kw_run_with_contact_bit_exact():
    sc = scenarios.c1_kalthoff_winkler(steps=80, width=0.005)
    m = _oracle_for(sc)
    m.surface_correct()
    m.init_state()
    m.set_notches_kw(0.025)
    m.cylinder_init(sc.cyl)
    with sc.make() as prob:
        prob.run_verlet(80)
        for _ in range(80):
            m.step()
        assert np.array_equal(prob.get("U"), m.stack("U"))
        assert np.array_equal(prob.get("V"), m.stack("V"))
        assert np.array_equal(prob.fail_mask(), m.fail)
        cs = prob.cylinder_state()
        assert np.array_equal(cs["vel"], m.cst["vel"]) and np.array_equal(cs["disp"], m.cst["disp"])
        assert m.cst["vel"][1] > -32.0  # the cylinder has been slowed by contact


@pytest.mark.parametrize("skip_broken", [1, 0])
def test_lps2d_crack_grow_fp64(skip_broken):
    # Solve_Plate_Crack_Grow.cu at reduced size: 2D LPS with strips.  The
    # dilatation skips broken bonds (CPU twin, Force_State.cu:345) or keeps
    # them (the reference's 2D GPU kernel, Force_State.cu:65); the crack makes
    # the two settings differ from the first step.
    dx = 1e-4
    p, (nx, ny, _), _n = pd.params(2, _lib.PD_MODEL_LPS, 1.3367e-8, dx, thick=1e-4, length=0.006, height=0.006,
                                   emod=192e9, pratio=0.3, dens=7850.0, s0=0.04472)
    p.theta_skip_broken = skip_broken
    p.n_top = p.n_bot = 3 * nx
    p.strip_velocity = 20.0
    x, y = pd.gen_plate_strips(nx, ny, dx)
    prob = pd.Problem(p, 0)
    prob.set_positions(x, y, None, n_int=nx * ny)
    prob.build_neighbors()
    prob.init_state()
    prob.set_crack_2d(0.002, 0.0, 0.0, 0.0)
    prob.lps_weights()
    m = OracleModel(scenarios.params_dict(p), x, y, None, nx * ny)
    m.build_neighbors()
    m.init_state()
    m.set_crack_2d(0.002, 0.0, 0.0, 0.0)
    m.lps_weights()
    assert np.array_equal(prob.get("W"), m.w) and np.array_equal(prob.get("M")[: len(x)], m.m)
    prob.run_verlet(1)
    m.step()
    assert np.array_equal(prob.get("THETA")[: len(x)], m.theta)
    prob.run_verlet(299)
    for _ in range(299):
        m.step()
    assert np.array_equal(prob.get("U"), m.stack("U"))
    assert np.array_equal(prob.get("V"), m.stack("V"))
    assert np.array_equal(prob.fail_mask(), m.fail)
    assert (m.fail == 0).sum() > 0
    prob.close()


def test_lps3d_block_fp64():
    sc = scenarios.c2_lps_block(nx=24, ny=20, nz=16, precision=_lib.PD_PREC_FP64)
    m = _oracle_for(sc)
    m.init_state()
    m.lps_weights()
    m.V = [a.copy() for a in sc.v0]
    with sc.make() as prob:
        assert np.array_equal(prob.get("W"), m.w) and np.array_equal(prob.get("M"), m.m)
        prob.run_verlet(100)
        for _ in range(100):
            m.step()
        assert np.array_equal(prob.get("U"), m.stack("U"))
        assert np.array_equal(prob.get("V"), m.stack("V"))
        assert np.array_equal(prob.get("THETA"), m.theta)
        assert np.array_equal(prob.fail_mask(), m.fail)


# --------------------------------------------------------------------- ADR
def test_adr_plate_tensile_bit_exact():
    sc = scenarios.plate_tensile(dx=0.02)
    m = _oracle_for(sc)
    m.surface_correct()
    m.init_state()
    bx = np.zeros(len(sc.x))
    nx, ny, _ = sc.nxyz
    for i in range(ny):
        bx[i * nx] = -0.001 * sc.p.emod / sc.p.dx
        bx[i * nx + nx - 1] = 0.001 * sc.p.emod / sc.p.dx
    m.B[0][:] = bx
    with sc.make() as prob:
        assert np.array_equal(prob.get("SCR"), m.scr)
        assert np.array_equal(prob.get("B")[0], bx)
        for ct in range(200):
            c_g, r_g = prob.adr_iterate(ct)
            c_o, r_o = m.adr_iterate(ct)
            assert c_g == c_o and r_g == r_o, ct
        assert np.array_equal(prob.get("U"), m.stack("U"))
        # device-side convergence loop continues from the same state
        it, r, ok = prob.adr_solve(200, 50, 0.0)
        assert it == 50 and not ok


def test_reduce_sum_matches_oracle_tree():
    sc = scenarios.c0_plate_pmb(scale=0.5)
    rng = np.random.default_rng(3)
    with sc.make() as prob:
        u = rng.standard_normal((2, len(sc.x)))
        prob.set("U", u)
        for comp in (0, 1):
            assert prob.reduce_sum("U", comp) == O.sum_tree(u[comp], sc.p.reduce_lanes)


def test_checkpoint_resume_is_bit_identical(tmp_path):
    sc = scenarios.c0_plate_pmb(scale=0.4)
    with sc.make() as a:
        a.run_verlet(40)
        a.checkpoint_write(tmp_path / "c.bin", 40)
        a.run_verlet(40)
        ua, fa = a.get("U"), a.fail_mask()
    with sc.make() as b:
        assert b.checkpoint_read(tmp_path / "c.bin") == 40
        b.run_verlet(40)
        assert np.array_equal(b.get("U"), ua) and np.array_equal(b.fail_mask(), fa)


# ------------------------------------------------------------------- mixed
def test_mixed_pmb3d_force_and_breaks():
    sc = scenarios.pmb_block(nx=30, ny=26, nz=20, precision=_lib.PD_PREC_MIXED)
    m = _oracle_for(sc)
    m.init_state()
    rng = np.random.default_rng(11)
    u = np.stack([2.05e-3 * sc.x, -4e-4 * sc.y, 1e-3 * sc.z]) + rng.uniform(-2e-9, 2e-9, (3, len(sc.x)))
    with sc.make() as prob:
        prob.set("U", u)
        m.U = [a.copy() for a in u]
        prob.force()
        m.force()
        f_g, f_o = prob.get("F"), m.stack("F")
        scale = np.max(np.abs(f_o))
        assert np.max(np.abs(f_g - f_o)) / scale < 2e-5
        assert np.array_equal(prob.fail_mask(), m.fail)
        assert (m.fail == 0).sum() > 0


def test_mixed_breaks_exact_under_large_translation():
    # The mixed records hold u rounded to FP32, whose error grows with |u|:
    # the FP64 re-evaluation band widens with the bond's displacement, so a
    # large rigid translation on top of a near-critical stretch still breaks
    # exactly the bonds the FP64 oracle breaks.  Forces lose accuracy with
    # |u| / |eta| (here ~100), so they are only bounded loosely.
    sc = scenarios.pmb_block(nx=30, ny=26, nz=20, precision=_lib.PD_PREC_MIXED)
    m = _oracle_for(sc)
    m.init_state()
    rng = np.random.default_rng(19)
    u = np.stack([2.05e-3 * sc.x + 0.7, -4e-4 * sc.y - 0.3, 1e-3 * sc.z + 0.45])
    u = u + rng.uniform(-2e-9, 2e-9, u.shape)
    with sc.make() as prob:
        prob.set("U", u)
        m.U = [a.copy() for a in u]
        prob.force()
        m.force()
        assert np.array_equal(prob.fail_mask(), m.fail)
        assert (m.fail == 0).sum() > 0
        f_g, f_o = prob.get("F"), m.stack("F")
        assert np.max(np.abs(f_g - f_o)) / np.max(np.abs(f_o)) < 5e-2


def test_mixed_lps3d_force():
    sc = scenarios.c2_lps_block(nx=30, ny=26, nz=20, precision=_lib.PD_PREC_MIXED)
    m = _oracle_for(sc)
    m.init_state()
    m.lps_weights()
    rng = np.random.default_rng(12)
    u = np.stack([1.1e-3 * sc.x, 3e-4 * sc.y, 2.05e-3 * sc.z]) + rng.uniform(-2e-9, 2e-9, (3, len(sc.x)))
    with sc.make() as prob:
        prob.set("U", u)
        m.U = [a.copy() for a in u]
        prob.force()
        m.force()
        assert _relerr(prob.get("THETA"), m.theta) < 2e-5
        f_g, f_o = prob.get("F")[:, : sc.n_int], m.stack("F")[:, : sc.n_int]
        assert np.max(np.abs(f_g - f_o)) / np.max(np.abs(f_o)) < 5e-5
        assert np.array_equal(prob.fail_mask(), m.fail)
        assert (m.fail == 0).sum() > 0


@pytest.mark.parametrize("model", ["pmb", "lps"])
def test_mixed_group_sizes_agree(model):
    if model == "pmb":
        sc = scenarios.pmb_block(nx=24, ny=22, nz=18, precision=_lib.PD_PREC_MIXED)
    else:
        sc = scenarios.c2_lps_block(nx=24, ny=22, nz=18, precision=_lib.PD_PREC_MIXED)
    rng = np.random.default_rng(5)
    u = np.stack([2.03e-3 * sc.x, -3e-4 * sc.y, 1e-3 * sc.z]) + rng.uniform(-2e-9, 2e-9, (3, len(sc.x)))
    ref_f, ref_mask = None, None
    for g in (4, 8, 16, 32, 64):
        with sc.make() as prob:
            prob.set_group(g)
            prob.set("U", u)
            prob.force()
            f, mask = prob.get("F"), prob.fail_mask()
        if ref_f is None:
            ref_f, ref_mask = f, mask
        else:
            assert np.max(np.abs(f - ref_f)) <= 1e-6 * np.max(np.abs(ref_f)), g
            assert np.array_equal(mask, ref_mask), g
    assert (ref_mask == 0).sum() > 0


def test_mixed_tile_order_does_not_change_results():
    # strips after N_int: the tiled order must keep integrated rows first
    sc = scenarios.c0_plate_pmb(scale=0.6)
    rng = np.random.default_rng(9)
    u = rng.uniform(-3e-6, 3e-6, size=(2, len(sc.x)))
    outs = []
    for tile in (0, 8, 64):
        with sc.make(precision=_lib.PD_PREC_MIXED) as prob:
            prob.set_tile(tile)
            prob.set("U", u)
            prob.force()
            outs.append((prob.get("F"), prob.fail_mask()))
    for f, mk in outs[1:]:
        assert np.array_equal(f, outs[0][0]) and np.array_equal(mk, outs[0][1])


def test_mixed_records_follow_every_displacement_update():
    # The integrators refresh the u half of the particle records in place;
    # set("U"), strip_displace and a precision switch must force a re-pack.
    sc = scenarios.c0_plate_pmb(scale=0.6)
    rng = np.random.default_rng(21)
    u0 = rng.uniform(-2e-6, 2e-6, size=(2, len(sc.x)))
    u1 = rng.uniform(-2e-6, 2e-6, size=(2, len(sc.x)))
    with sc.make(precision=_lib.PD_PREC_MIXED) as a, sc.make(precision=_lib.PD_PREC_MIXED) as b:
        for u in (u0, u1):
            a.set("U", u)
            b.set("U", u)
            a.run_verlet(15)  # captured graph
            for _ in range(15):  # staged calls
                b.verlet_half1()
                b.force()
                b.verlet_half2()
            for field in ("U", "V", "A", "F"):  # run_verlet fuses half 2 with the next half 1
                assert np.array_equal(a.get(field), b.get(field)), field
            assert np.array_equal(a.fail_mask(), b.fail_mask())
        a.strip_displace(1e-6, -1e-6)
        uf, mask = a.get("U"), a.fail_mask()  # bonds that break in this pass still act in it
        a.force()
        f_a = a.get("F")
    with sc.make(precision=_lib.PD_PREC_MIXED) as c:
        c.set("U", uf)
        c.set_fail_mask(mask)
        c.force()
        assert np.array_equal(c.get("F"), f_a)


def _force_with(sc, lds, tile, u, precision=_lib.PD_PREC_MIXED):
    with sc.make(precision=precision) as prob:
        prob.set_lds(lds)
        prob.set_tile(tile)
        prob.set("U", u)
        prob.force()
        info = prob.lds_info()
        return prob.get("F"), prob.fail_mask(), info, (prob.get("THETA") if sc.p.model == _lib.PD_MODEL_LPS else None)


@pytest.mark.parametrize("case", ["lps3d", "pmb3d", "pmb2d_strips", "lps2d"])
def test_lds_tiles_match_gathering_kernels(case):
    rng = np.random.default_rng(31)
    if case == "lps3d":
        sc = scenarios.c2_lps_block(nx=30, ny=26, nz=20, precision=_lib.PD_PREC_MIXED)
        u = np.stack([1.1e-3 * sc.x, 3e-4 * sc.y, 2.05e-3 * sc.z])
    elif case == "pmb3d":
        sc = scenarios.pmb_block(nx=30, ny=26, nz=20, precision=_lib.PD_PREC_MIXED)
        u = np.stack([2.05e-3 * sc.x, -4e-4 * sc.y, 1e-3 * sc.z])
    elif case == "pmb2d_strips":
        sc = scenarios.c0_plate_pmb(scale=0.6)
        u = np.stack([4e-3 * sc.x, 4.1e-3 * sc.y])
    else:
        sc = scenarios.lps2d_plate(scale=1.0)
        u = np.stack([4e-3 * sc.x, 4.1e-3 * sc.y])
    u = u + rng.uniform(-2e-9, 2e-9, u.shape)
    f0, m0, i0, th0 = _force_with(sc, 0, 6 if sc.p.dim == 3 else 16, u)
    f1, m1, i1, th1 = _force_with(sc, 1, 6 if sc.p.dim == 3 else 16, u)
    assert not i0[0] and i1[0] and i1[1] > 1 and 0 < i1[2] <= 2048, (i0, i1)
    assert np.array_equal(m0, m1)
    assert (m1 == 0).sum() > 0
    assert np.array_equal(f0, f1), float(np.max(np.abs(f0 - f1)) / np.max(np.abs(f0)))
    if th0 is not None:
        assert np.array_equal(th0, th1)


@pytest.mark.parametrize("case", ["lps3d", "pmb2d_strips"])
def test_lds_intact_bits_follow_breaks_and_fail_writes(case):
    # The LDS kernels read a bond's intact flag from bit 15 of its local
    # index word.  Breaks inside captured Verlet steps, a fail-mask upload,
    # init_state and a crack must all leave the LDS path bitwise equal to
    # the gathering kernels, which read the fail words themselves.
    if case == "lps3d":
        sc = scenarios.c2_lps_block(nx=24, ny=20, nz=16, precision=_lib.PD_PREC_MIXED)
        u = np.stack([1.1e-3 * sc.x, 3e-4 * sc.y, 2.05e-3 * sc.z])
        tile = 6
    else:
        sc = scenarios.c0_plate_pmb(scale=0.6)
        u = np.stack([4e-3 * sc.x, 4.1e-3 * sc.y])
        tile = 16
    outs = []
    for lds in (0, 1):
        with sc.make(precision=_lib.PD_PREC_MIXED) as prob:
            prob.set_lds(lds)
            prob.set_tile(tile)
            prob.set("U", u)
            prob.run_verlet(6)
            broken_by_steps = int((prob.fail_mask() == 0).sum())
            m = prob.fail_mask()
            m[3::7] = 0
            prob.set_fail_mask(m)
            prob.run_verlet(3)
            f1, u1, m1 = prob.get("F"), prob.get("U"), prob.fail_mask()
            prob.init_state()
            if sc.p.dim == 2:
                prob.set_crack_2d(0.01, 0.0, 0.0, 0.0)
            prob.set("U", 0.5 * u)
            prob.force()
            outs.append((f1, u1, m1, prob.get("F"), prob.fail_mask(), prob.lds_info()[0], broken_by_steps))
    g, l = outs
    assert not g[5] and l[5]
    assert g[6] == l[6] > 0
    for a, b in zip(g[:5], l[:5]):
        assert np.array_equal(a, b)


def test_lds_capacity_is_reported_and_default_falls_back():
    sc = scenarios.c2_lps_block(nx=30, ny=26, nz=20, precision=_lib.PD_PREC_MIXED)
    u = np.stack([1.1e-3 * sc.x, 3e-4 * sc.y, 2.05e-3 * sc.z])
    with sc.make() as prob:
        prob.set_tile(12)  # a 12 dx tile stages ~5000 records: above the builder's 4096
        prob.set_lds(1)
        with pytest.raises(pd.PdError):
            prob.force()
        prob.set_lds(-1)
        prob.set("U", u)
        prob.force()
        assert prob.lds_info()[0] is False
        f_fallback = prob.get("F")
    f_lds, _, info, _ = _force_with(sc, -1, -1, u)
    assert info[0]
    assert np.array_equal(f_fallback, f_lds)


def _force_coef(sc, mode, u):
    with sc.make() as prob:
        prob.set_coef_mode(mode)
        prob.set("U", u)
        prob.force()
        th = prob.get("THETA") if sc.p.model == _lib.PD_MODEL_LPS else None
        return prob.get("F"), prob.fail_mask(), prob.coef_recomputed(), th


@pytest.mark.parametrize("case", ["lps3d", "pmb3d", "pmb2d_strips"])
def test_recomputed_bond_coefficients_match_streamed(case):
    rng = np.random.default_rng(41)
    if case == "lps3d":
        sc = scenarios.c2_lps_block(nx=30, ny=26, nz=20, precision=_lib.PD_PREC_MIXED)
        u = np.stack([1.1e-3 * sc.x, 3e-4 * sc.y, 2.05e-3 * sc.z])
    elif case == "pmb3d":
        sc = scenarios.pmb_block(nx=30, ny=26, nz=20, precision=_lib.PD_PREC_MIXED)
        u = np.stack([2.05e-3 * sc.x, -4e-4 * sc.y, 1e-3 * sc.z])
    else:
        sc = scenarios.c0_plate_pmb(scale=0.6)
        sc.p.precision = _lib.PD_PREC_MIXED
        u = np.stack([4e-3 * sc.x, 4.1e-3 * sc.y])
    u = u + rng.uniform(-2e-9, 2e-9, u.shape)
    f0, m0, r0, th0 = _force_coef(sc, 0, u)
    f1, m1, r1, th1 = _force_coef(sc, -1, u)
    assert not r0 and r1
    assert np.array_equal(m0, m1) and (m1 == 0).sum() > 0
    assert np.max(np.abs(f1 - f0)) <= 1e-5 * np.max(np.abs(f0))
    if th0 is not None:
        assert np.max(np.abs(th1 - th0)) <= 1e-5 * np.max(np.abs(th0))


def test_surface_corrected_coefficients_are_streamed():
    sc = scenarios.c1_kalthoff_winkler(steps=10)
    with sc.make(precision=_lib.PD_PREC_MIXED) as prob:
        assert prob.coef_recomputed() is False  # scr != 1 on surface bonds
        prob.set_coef_mode(1)
        with pytest.raises(pd.PdError):
            prob.force()


@pytest.mark.parametrize("case", ["kw3d", "c0_2d"])
def test_fp64_pmb_lane_groups_are_bit_identical(case):
    # G lanes per row add the bond terms in bond order: same bits as one thread per row
    if case == "kw3d":
        sc = scenarios.c1_kalthoff_winkler(steps=10)
        rng = np.random.default_rng(51)
        u = np.stack([4e-3 * sc.x, 4e-3 * sc.y, 1e-3 * sc.z]) + rng.uniform(-1e-7, 1e-7, (3, len(sc.x)))
    else:
        sc = scenarios.c0_plate_pmb(scale=0.6)
        rng = np.random.default_rng(52)
        u = np.stack([4.1e-3 * sc.x, 4e-3 * sc.y]) + rng.uniform(-1e-8, 1e-8, (2, len(sc.x)))
    outs = []
    for g in (1, 4, 16, 32):
        with sc.make(precision=_lib.PD_PREC_FP64) as prob:
            prob.set_fp64_group(g)
            prob.set("U", u)
            prob.force()
            outs.append((prob.get("F"), prob.fail_mask()))
    assert (outs[0][1] == 0).sum() > 0
    for f, mk in outs[1:]:
        assert np.array_equal(f, outs[0][0]) and np.array_equal(mk, outs[0][1])

"""Pins of the FP64 oracle (CPU only).

The reference ships no tests, fixtures or golden vectors, so each oracle
function is pinned here by closed forms, invariants, brute force or an
independent vectorised recomputation, chosen so that a dropped term, a wrong
sign or index, or a transposed operand fails at least one check.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O
from oracle.runner import OracleModel

PI = math.acos(-1.0)


def c0_params():
    return O.cal_parameter(2, 1, 1e-8, 0.5e-3, 0.5e-3, 0.05, 0.05, 0.0, 192e9, 1 / 3, 8000, 0.004)


# ------------------------------------------------------------------ params
def test_parameters_closed_forms():
    p = c0_params()
    d = 3.015 * 0.5e-3
    assert p.nx == 100 and p.ny == 100 and p.N == 10000 and p.MN == 32
    assert p.delta == d
    assert math.isclose(p.bc, 6 * 192e9 / (PI * 0.5e-3 * d ** 3 * (2 / 3)), rel_tol=1e-14)
    assert math.isclose(p.vol, 0.5e-3 ** 3, rel_tol=1e-15)
    assert math.isclose(p.mass, 0.25 * 1e-16 * PI * d * d * 0.5e-3 * p.bc / 0.5e-3 * 5, rel_tol=1e-14)
    k = O.cal_parameter(3, 1, 8.7e-8, 1e-3, 1e-3, 0.2, 0.1, 0.009, 190e9, 0.25, 8000, 0.01)
    assert (k.nx, k.ny, k.nz) == (200, 100, 9)
    assert math.isclose(k.bc, 6 * 190e9 / (PI * (3.015e-3) ** 4 * 0.5), rel_tol=1e-14)
    l3 = O.cal_parameter(3, 2, 2e-8, 0.5e-3, 0.0, 0.1, 0.1, 0.05, 70e9, 0.33, 2700, 0.002)
    assert (l3.nx, l3.ny, l3.nz) == (200, 200, 100)
    assert math.isclose(l3.K, 70e9 / (3 * (1 - 0.66)), rel_tol=1e-15)
    assert math.isclose(l3.G, 70e9 / (2 * 1.33), rel_tol=1e-15)
    bug = O.cal_parameter(3, 2, 2e-8, 0.5e-3, 0.0, 0.1, 0.1, 0.05, 70e9, 0.33, 2700, 0.002, ref_k_bug=True)
    assert bug.K < 0  # Base.cuh:175 divides by (1 - 2G)
    l2 = O.cal_parameter(2, 2, 1.3367e-8, 1e-4, 1e-4, 0.05, 0.05, 0.0, 192e9, 0.3, 7850, 0.04472)
    assert l2.thick == 1e-4 and math.isclose(l2.K, 192e9 / (2 * 0.7), rel_tol=1e-15)


# ------------------------------------------------------------- coordinates
def test_coordinates_closed_forms_and_counts():
    x, y = O.coord_plate_strips(100, 100, 0.5e-3)
    assert len(x) == 10600
    j, i = 7, 103  # a bottom-strip row
    assert x[i * 100 + j] == (-50 + 0.5 + j) * 0.5e-3
    assert y[i * 100 + j] == (-50 - 0.5) * 0.5e-3
    assert y[100 * 100 + 5] == (50 + 0.5) * 0.5e-3  # first top-strip row
    xk, yk, zk = O.coord_kw(201, 101, 9, 1e-3, 0.025)
    assert len(xk) == 180873
    # the notches remove two columns (x = -26,-25 and 25,26 mm) on the 51 rows y >= 0
    removed = 201 * 101 * 9 - 180873
    assert removed == 2 * 2 * 51 * 9
    assert not np.any((np.abs(np.abs(xk) - 0.0255) < 0.0006) & (yk >= -1e-10))


def test_block_jitter_is_seeded_and_bounded():
    x, y, z = O.coord_block_jitter(6, 5, 4, 1e-3, 0.01, 0x5EED0002)
    x2, _, _ = O.coord_block_jitter(6, 5, 4, 1e-3, 0.01, 0x5EED0002)
    assert np.array_equal(x, x2)
    base = np.tile((np.arange(6) - 3 + 0.5) * 1e-3, 20)
    assert np.all(np.abs(x - base) <= 0.01e-3 + 1e-18)
    st = np.array([0x5EED0002], dtype=np.uint64)
    # SplitMix64 reference value of the first output for this seed
    s = 0x5EED0002
    s = (s + 0x9E3779B97F4A7C15) & (2 ** 64 - 1)
    zz = s
    zz = ((zz ^ (zz >> 30)) * 0xBF58476D1CE4E5B9) & (2 ** 64 - 1)
    zz = ((zz ^ (zz >> 27)) * 0x94D049BB133111EB) & (2 ** 64 - 1)
    zz ^= zz >> 31
    u53 = (zz >> 11) * 2.0 ** -53
    assert x[0] == (-3 + 0.5) * 1e-3 + (2 * u53 - 1) * 0.01 * 1e-3
    del st


# -------------------------------------------------------------- neighbours
def test_neighbors_cell_equals_brute_force_2d_and_3d():
    p = c0_params()
    x, y = O.coord_plate_strips(p.nx, p.ny, p.dx)
    NN, NL = O.neighbors_brute(2, x, y, None, p.delta, 32)
    rp_b, col_b = O.dense_to_csr(NN, NL, 32)
    rp_c, col_c, mx = O.neighbors_cell(2, x, y, None, p.delta)
    assert np.array_equal(rp_b, rp_c) and np.array_equal(col_b, col_c)
    assert rp_c[-1] == 289420 and mx == 28
    xk, yk, zk = O.coord_kw(41, 21, 7, 1e-3, 0.005)
    for fac in (3.015, 3.0):
        NN, NL = O.neighbors_brute(3, xk, yk, zk, fac * 1e-3, 128)
        rp_b, col_b = O.dense_to_csr(NN, NL, 128)
        rp_c, col_c, _ = O.neighbors_cell(3, xk, yk, zk, fac * 1e-3)
        assert np.array_equal(rp_b, rp_c) and np.array_equal(col_b, col_c)
    xj, yj, zj = O.coord_block_jitter(14, 12, 10, 0.5e-3, 0.01, 99)
    NN, NL = O.neighbors_brute(3, xj, yj, zj, 3.015 * 0.5e-3, 128)
    rp_b, col_b = O.dense_to_csr(NN, NL, 128)
    rp_c, col_c, _ = O.neighbors_cell(3, xj, yj, zj, 3.015 * 0.5e-3)
    assert np.array_equal(rp_b, rp_c) and np.array_equal(col_b, col_c)


def test_neighbor_invariants_and_lattice_counts():
    xk, yk, zk = O.coord_block_jitter(12, 12, 12, 1e-3, 0.0, 1)
    rp, col, mx = O.neighbors_cell(3, xk, yk, zk, 3.015e-3)
    assert mx == 122  # interior count at delta = 3.015 dx (offsets with a^2+b^2+c^2 <= 9)
    nb = np.diff(rp)
    assert np.all(np.diff(col[rp[0]:rp[1]]) > 0)  # rows ascending
    # symmetry: j in N(i) <=> i in N(j)
    pairs = set()
    for i in range(len(xk)):
        for j in col[rp[i]:rp[i + 1]]:
            pairs.add((i, int(j)))
    assert all((j, i) in pairs for (i, j) in pairs)
    assert nb.sum() == len(pairs)
    x2, y2 = O.coord_plate(20, 20, 1e-3)
    _, _, mx2 = O.neighbors_cell(2, x2, y2, None, 3.015e-3)
    assert mx2 == 28


def test_kw_and_c0_bond_counts():
    xk, yk, zk = O.coord_kw(201, 101, 9, 1e-3, 0.025)
    rp, _, _ = O.neighbors_cell(3, xk, yk, zk, 3.015e-3)
    assert rp[-1] == 18809054


# -------------------------------------------------------------- vol / surf
def lattice_model(nx=16, ny=14, dx=1e-3):
    p = O.cal_parameter(2, 1, 1e-8, dx, dx, nx * dx, ny * dx, 0.0, 192e9, 1 / 3, 8000, 0.004)
    x, y = O.coord_plate(p.nx, p.ny, dx)
    sc = dict(dim=2, model=1, dx=dx, delta=p.delta, vol=p.vol, bc=p.bc, mass=p.mass, K=0, G=0,
              emod=192e9, pratio=1 / 3, dens=8000, s0=0.004, dt=1e-8, break_y_min=-np.inf,
              reduce_lanes=64)
    m = OracleModel(sc, x, y)
    m.build_neighbors()
    return m, p


def test_vol_corr_piecewise_closed_form():
    m, p = lattice_model()
    i = np.repeat(np.arange(m.N), np.diff(m.row_ptr))
    ddx, ddy = m.X[0][m.col] - m.X[0][i], m.X[1][m.col] - m.X[1][i]
    xi = np.sqrt(ddx * ddx + ddy * ddy)
    assert np.array_equal(m.idist, xi)
    assert np.allclose(m.idist, np.hypot(ddx, ddy), rtol=2e-16, atol=0)
    expect = np.where(xi <= p.delta - p.dx / 2, 1.0,
                      np.where(xi <= p.delta + p.dx / 2, (p.delta + p.dx / 2 - xi) / p.dx, 0.0))
    assert np.allclose(m.fac, expect, rtol=0, atol=1e-15)
    # a bond of length 3 dx at delta = 3.015 dx keeps 0.515 of its volume
    three = np.isclose(xi, 3 * p.dx)
    assert np.allclose(m.fac[three], 0.515, atol=1e-12)


def test_surface_correction_energy_identity_and_symmetry():
    m, p = lattice_model(20, 20)
    m.surface_correct(scr_mode=0)
    sed = 0.5 * 192e9 / (1 - 1 / 9) * 1e-6
    i = np.repeat(np.arange(m.N), np.diff(m.row_ptr))
    c = m.col
    # independent vectorised strain energy under the probe field u_d = 1e-3 x_d
    for d in range(2):
        dxv = m.X[0][c] - m.X[0][i]
        dyv = m.X[1][c] - m.X[1][i]
        if d == 0:
            dxv = dxv * 1.001
        else:
            dyv = dyv * 1.001
        nl = np.hypot(dxv, dyv)
        w = 0.25 * p.bc * (nl - m.idist) ** 2 / m.idist * p.vol * m.fac
        SED = np.bincount(i, weights=w, minlength=m.N)
        assert np.allclose(m.fncst[d] * SED, sed, rtol=1e-9)
    # square-lattice symmetry: fncst_x == fncst_y on the diagonal-symmetric plate centre
    ctr = (10 * 20 + 10)
    assert math.isclose(m.fncst[0][ctr], m.fncst[1][ctr], rel_tol=1e-9)
    # trig-free scr agrees with the literal atan/cos/sin form
    scr0 = m.scr.copy()
    m.surface_correct(scr_mode=1)
    assert np.max(np.abs(scr0 - m.scr) / m.scr) < 1e-14
    # an axis-aligned bond with scx == scy carries the common value
    k = m.row_ptr[ctr] + np.where((m.col[m.row_ptr[ctr]:m.row_ptr[ctr + 1]] == ctr + 1))[0][0]
    sx = (m.fncst[0][ctr] + m.fncst[0][ctr + 1]) / 2
    assert math.isclose(scr0[k], sx, rel_tol=1e-14)


def test_surface_correction_3d_trig_free_matches_literal():
    xk, yk, zk = O.coord_kw(31, 21, 7, 1e-3, 0.005)
    p = O.cal_parameter(3, 1, 8.7e-8, 1e-3, 1e-3, 0.03, 0.02, 0.007, 190e9, 0.25, 8000, 0.01)
    sc = dict(dim=3, model=1, dx=1e-3, delta=p.delta, vol=p.vol, bc=p.bc, emod=190e9, pratio=0.25)
    m = OracleModel(sc, xk, yk, zk)
    m.build_neighbors()
    m.surface_correct(0)
    a = m.scr.copy()
    m.surface_correct(1)
    assert np.max(np.abs(a - m.scr) / m.scr) < 1e-14
    assert np.all(a > 0.5) and np.all(a < 10.0)


# ------------------------------------------------------------------- crack
def test_crack_horizontal_matches_geometric_rule_and_vertical_branch():
    m, p = lattice_model(30, 30)
    m.init_state()
    m.set_crack_2d(0.01, 0.0, 0.0, 0.0)
    i = np.repeat(np.arange(m.N), np.diff(m.row_ptr))
    c = m.col
    xi, yi, xc, yc = m.X[0][i], m.X[1][i], m.X[0][c], m.X[1][c]
    cross = (np.abs(xi) < 0.005) & (np.abs(xc) < 0.005) & (yi * yc <= 0)
    assert np.array_equal(m.fail == 0, cross)
    assert cross.sum() > 0
    # the theta = pi/2 branch on the transposed plate breaks the same bonds
    sc = dict(m.sc)
    mt = OracleModel(sc, m.X[1].copy(), m.X[0].copy())
    mt.build_neighbors()
    mt.init_state()
    mt.set_crack_2d(0.01, 0.0, 0.0, PI / 2)
    pairs = {(int(a), int(b)) for a, b in zip(i[m.fail == 0], c[m.fail == 0])}
    it = np.repeat(np.arange(mt.N), np.diff(mt.row_ptr))
    pairs_t = {(int(a), int(b)) for a, b in zip(it[mt.fail == 0], mt.col[mt.fail == 0])}
    assert pairs == pairs_t


# --------------------------------------------------------------------- PMB
def test_pmb_two_particle_analytic_and_strict_break():
    x = np.array([0.0, 1.0e-3])
    y = np.array([0.0, 0.0])
    sc = dict(dim=2, model=1, dx=1e-3, delta=3.015e-3, vol=1e-9, bc=1e20, s0=0.004, dt=1e-8, break_y_min=-np.inf)
    m = OracleModel(sc, x, y)
    m.build_neighbors()
    m.U[0][1] = 2e-6  # s = 0.002
    m.force()
    s = (1.002e-3 - 1e-3) / 1e-3
    assert math.isclose(m.F[0][0], 1e-9 * 1.0 * 1e20 * 1.0 * s, rel_tol=1e-9)
    assert math.isclose(m.F[0][1], -m.F[0][0], rel_tol=1e-12)
    assert np.all(m.fail == 1)
    # exactly at s0 the bond survives (strict >), just above it breaks
    sv = (1.002e-3 + 1e-3 * 0.0 - 0.0 - 0.0 - 1e-3)
    m.sc["s0"] = (np.sqrt((1e-3 + 2e-6) ** 2) - m.idist[0]) / m.idist[0]
    m.force()
    assert np.all(m.fail == 1)
    m.sc["s0"] = np.nextafter(m.sc["s0"], 0.0)
    m.force()
    assert np.all(m.fail == 0)
    del sv


def test_pmb_invariants_translation_newton_uniform_stretch():
    m, p = lattice_model(24, 24)
    m.init_state()
    N = m.N
    m.U = [np.full(N, 3.3e-7), np.full(N, -1.1e-7)]
    m.force()
    rigid = np.max(np.abs(m.stack("F")))
    m.U = [1e-3 * m.X[0], -2e-4 * m.X[1]]
    m.force()
    F = m.stack("F")
    assert rigid < 1e-9 * np.max(np.abs(F))
    # Newton's third law: sum_i vol f_i = 0
    assert np.max(np.abs(F.sum(axis=1))) < 1e-9 * np.abs(F).sum()
    # uniform stretch: interior particles (farther than delta from the edge) feel no net force
    inner = (np.abs(m.X[0]) < 0.012 - 4e-3) & (np.abs(m.X[1]) < 0.012 - 4e-3)
    assert np.max(np.abs(F[:, inner])) < 1e-8 * np.max(np.abs(F))


# --------------------------------------------------------------------- LPS
def test_lps_m_theta_closed_forms():
    xk, yk, zk = O.coord_block_jitter(20, 20, 20, 1e-3, 0.0, 0)
    p = O.cal_parameter(3, 2, 1e-8, 1e-3, 0.0, 0.02, 0.02, 0.02, 70e9, 0.33, 2700, 0.002)
    sc = dict(dim=3, model=2, dx=1e-3, delta=p.delta, vol=p.vol, K=p.K, G=p.G, pratio=0.33, s0=0.002,
              break_y_min=-np.inf, dt=1e-8, dens=2700, theta_skip_broken=1)
    m = OracleModel(sc, xk, yk, zk)
    m.build_neighbors()
    m.init_state()
    m.lps_weights()
    i = np.repeat(np.arange(m.N), np.diff(m.row_ptr))
    m_ind = np.bincount(i, weights=np.exp(-(m.idist / p.delta) ** 2) * m.idist ** 2 * p.vol * m.fac, minlength=m.N)
    assert np.allclose(m.m, m_ind, rtol=1e-13)
    # rigid translation: theta = 0, f = 0
    m.U = [np.full(m.N, 1e-7), np.full(m.N, 2e-7), np.full(m.N, -1e-7)]
    m.force()
    assert np.max(np.abs(m.theta)) < 1e-8
    # uniform dilatation eps: theta = 3 eps everywhere (3D), interior force 0
    eps = 1e-4
    m.U = [eps * m.X[0], eps * m.X[1], eps * m.X[2]]
    m.force()
    assert np.allclose(m.theta, 3 * eps, rtol=1e-6)
    F = m.stack("F")
    # particles whose neighbours all have complete horizons (2 delta from the surface)
    inner = np.all([np.abs(m.X[d]) < 0.010 - 6.5e-3 for d in range(3)], axis=0)
    assert inner.sum() > 0
    assert np.max(np.abs(F[:, inner])) < 1e-6 * np.max(np.abs(F))


def test_lps_2d_theta_coefficient():
    x, y = O.coord_plate(20, 20, 1e-3)
    p = O.cal_parameter(2, 2, 1e-8, 1e-3, 1e-3, 0.02, 0.02, 0.0, 192e9, 0.3, 7850, 0.04)
    sc = dict(dim=2, model=2, dx=1e-3, delta=p.delta, vol=p.vol, K=p.K, G=p.G, pratio=0.3, s0=0.04,
              break_y_min=-np.inf, dt=1e-8, dens=7850, theta_skip_broken=1)
    m = OracleModel(sc, x, y)
    m.build_neighbors()
    m.init_state()
    m.lps_weights()
    eps = 1e-4
    m.U = [eps * m.X[0], eps * m.X[1]]
    m.force()
    c = 2 * (2 * 0.3 - 1) / (0.3 - 1)
    assert np.allclose(m.theta, c * eps, rtol=1e-6)


# --------------------------------------------------- Verlet / ADR / reduce
def test_verlet_constant_force_parabola():
    N = 5
    V = [np.zeros(N), np.zeros(N)]
    A = [np.full(N, 2.0), np.full(N, -3.0)]
    U = [np.zeros(N), np.zeros(N)]
    F = [np.full(N, 2.0 * 7.0), np.full(N, -3.0 * 7.0)]
    B = [np.zeros(N), np.zeros(N)]
    dt = 1e-3
    for _ in range(100):
        O.verlet_half1(2, N, dt, V, A, U)
        O.verlet_half2(2, N, dt, 7.0, V, A, F, B)
    t = 100 * dt
    assert np.allclose(U[0], 0.5 * 2.0 * t * t, rtol=1e-12)
    assert np.allclose(V[1], -3.0 * t, rtol=1e-12)


def test_adr_damping_formula_cases():
    assert O.adr_c(4.0, 1.0) == 2.0 * 2.0 if 2.0 * 2.0 <= 2 else O.adr_c(4.0, 1.0) == 1.9
    assert O.adr_c(0.25, 1.0) == 1.0
    assert O.adr_c(-1.0, 1.0) == 0.0
    assert O.adr_c(1.0, 0.0) == 0.0
    assert O.adr_c(2.0, 1.0) == 1.9  # 2 sqrt(2) > 2 is clamped to 1.9 (Integrate.cu:126-127)


def test_adr_terms_hand_made():
    u = [np.array([2.0, 0.0]), np.array([1.0, 3.0])]
    f = [np.array([5.0, 1.0]), np.array([2.0, 2.0])]
    fo = [np.array([1.0, 1.0]), np.array([0.0, 1.0])]
    vho = [np.array([0.5, 0.0]), np.array([0.0, 2.0])]
    cn, disp = O.adr_terms(2, 2, 2.0, 0.1, u, f, fo, vho)
    # particle 0: x term -(4)*(5/2-1/2)/(0.1*0.5) = -160; y skipped (vho = 0)
    assert cn[0] == -4.0 * (2.5 - 0.5) / 0.05 and disp[0] == 5.0
    # particle 1: x skipped, y term -(9)*(1-0.5)/(0.1*2)
    assert math.isclose(cn[1], -9.0 * 0.5 / 0.2, rel_tol=1e-15) and disp[1] == 9.0


def _tree_py(v, T):
    p = [0.0] * T
    for t in range(T):
        s = 0.0
        for k in range(t, len(v), T):
            s += float(v[k])
        p[t] = s
    off = 1
    while off < T:
        for t in range(0, T, 2 * off):
            p[t] = p[t] + p[t + off]
        off *= 2
    return p[0]


def test_sum_tree_order_and_value():
    rng = np.random.default_rng(0)
    v = rng.standard_normal(3001) * 10.0 ** rng.integers(-8, 8, 3001)
    for T in (1, 64, 256):
        assert O.sum_tree(v, T) == _tree_py(v, T)
    assert O.sum_tree(v, 1) == sum(float(a) for a in v)
    assert math.isclose(O.sum_tree(v, 64), math.fsum(v), rel_tol=1e-10, abs_tol=1e-6)


def _tree_split_py(v, T, W, sequential_roots=False):
    """The T-lane tree as W workgroups of a batched specimen form it: part w
    sums the chains of lanes w*T/W ... (w+1)*T/W - 1, reduces them with the
    low-bit-first tree (its other lanes hold zeros), and the W subtree roots
    are combined by the remaining tree levels."""
    L = T // W
    chains = [0.0] * T
    for t in range(T):
        sc = 0.0
        for k in range(t, len(v), T):
            sc += float(v[k])
        chains[t] = sc
    roots = []
    for w in range(W):
        p = chains[w * L:(w + 1) * L] + [0.0] * (T - L)
        off = 1
        while off < T:
            for t in range(0, T, 2 * off):
                p[t] = p[t] + p[t + off]
            off *= 2
        roots.append(p[0])
    if sequential_roots:
        r = roots[0]
        for x in roots[1:]:
            r = r + x
        return r
    off = 1
    while off < W:
        for w in range(0, W, 2 * off):
            roots[w] = roots[w] + roots[w + off]
        off *= 2
    return roots[0]


def test_sum_tree_split_over_workgroups_is_exact():
    # the batched kernel's W workgroups per specimen reproduce the T-lane
    # tree bit for bit; combining the roots in any other order does not
    rng = np.random.default_rng(7)
    differs = 0
    for trial in range(20):
        v = rng.standard_normal(4000 + trial) * 10.0 ** rng.integers(-6, 6, 4000 + trial)
        for T, W in ((256, 2), (256, 4), (1024, 8), (1024, 16)):
            assert _tree_split_py(v, T, W) == O.sum_tree(v, T), (trial, T, W)
        differs += _tree_split_py(v, 1024, 16, sequential_roots=True) != O.sum_tree(v, 1024)
    assert differs > 0


def test_damage_closed_form():
    m, p = lattice_model(30, 30)
    m.init_state()
    m.set_crack_2d(0.01, 0.0, 0.0, 0.0)
    m.damage()
    i = np.repeat(np.arange(m.N), np.diff(m.row_ptr))
    broken = np.bincount(i, weights=(m.fail == 0) * m.fac, minlength=m.N)
    total = np.bincount(i, weights=m.fac, minlength=m.N)
    assert np.allclose(m.dmg, broken / total, atol=1e-15)
    assert m.dmg.max() > 0.3


def test_contact_single_particle_projection():
    cyl = [0.025, 1.57, 0.05, -32.0, 0.0, 0.1, 0.0]
    st = O.cylinder_init(cyl)
    st["disp"][1] = -0.03  # cylinder centre at y = 0.07, bottom cap at 0.045
    X = [np.array([0.003]), np.array([0.05]), np.array([0.004])]
    U = [np.zeros(1), np.zeros(1), np.zeros(1)]
    V = [np.zeros(1), np.zeros(1), np.zeros(1)]
    R = O.contact(1, np.zeros(1), X, U, V, cyl, st["disp"], 1e-6, 8000.0, 1e-9)
    # the particle is 5 mm inside the bottom cap: pushed to y = 0.045, x/z unchanged
    assert math.isclose(X[1][0] + U[1][0], 0.045, rel_tol=1e-12)
    assert abs(U[0][0]) < 1e-18 and abs(U[2][0]) < 1e-18
    assert math.isclose(V[1][0], -0.005 / 1e-6, rel_tol=1e-9)
    assert math.isclose(R[1][0], -8000.0 * V[1][0] / 1e-6 * 1e-9, rel_tol=1e-12)


@pytest.mark.parametrize("seed", [0, 1])
def test_c0_short_run_deterministic(seed):
    p = c0_params()
    x, y = O.coord_plate_strips(p.nx, p.ny, p.dx)
    sc = dict(dim=2, model=1, dx=p.dx, delta=p.delta, vol=p.vol, bc=p.bc, s0=p.s0, dt=1e-8, dens=8000,
              break_y_min=-np.inf, n_top=300, n_bot=300, strip_velocity=20.0)
    runs = []
    for _ in range(2):
        m = OracleModel(sc, x, y, None, 10000)
        m.build_neighbors()
        m.init_state()
        m.set_crack_2d(0.01, 0.0, 0.0, 0.0)
        for _ in range(20 + seed):
            m.step()
        runs.append((m.stack("U"), m.fail.copy()))
    assert np.array_equal(runs[0][0], runs[1][0]) and np.array_equal(runs[0][1], runs[1][1])
    assert np.max(np.abs(runs[0][0][1, 10000:10300])) == pytest.approx(20.0 * 1e-8 * (20 + seed), rel=1e-12)


# ------------------------------------------------------ remaining functions
def test_kw_notches_geometry_and_symmetry():
    # Solve_Kalthoff_Winkler.cu:65-87: a bond breaks when it crosses one of the
    # notch lines x = -rad or x = +rad and one of its ends lies at y >= 0.
    dx, rad = 1.0e-3, 0.0063
    x, y, z = O.coord_kw(41, 21, 3, dx, rad)
    delta = 3.015 * dx
    rp, col, _ = O.neighbors_cell(3, x, y, z, delta)
    fail = np.ones(len(col), dtype=np.uint8)
    O.set_crack_kw(len(x), rp, col, x, y, rad, fail)
    rows = np.repeat(np.arange(len(x)), np.diff(rp))
    xi, xc, yi, yc = x[rows], x[col], y[rows], y[col]
    cross = ((xi + rad) * (xc + rad) < 0) | ((xi - rad) * (xc - rad) < 0)
    upper = (yi >= -1e-10) | (yc >= -1e-10)
    broken = fail == 0
    assert broken.sum() > 100
    assert np.all(cross[broken]) and np.all(upper[broken])   # only crossing bonds in the notch region
    assert np.all(broken[cross & upper])                     # and all of them
    # the pair i->j and j->i break together
    key = {(int(a), int(b)): k for k, (a, b) in enumerate(zip(rows, col))}
    assert all(broken[key[(int(b), int(a))]] == broken[k] for k, (a, b) in enumerate(zip(rows, col)))
    # mirror symmetry x -> -x of the lattice maps the broken set onto itself
    pos = {(round(a / dx * 4), round(b / dx * 4), round(c / dx * 4)): i for i, (a, b, c) in enumerate(zip(x, y, z))}
    mirror = np.array([pos[(round(-a / dx * 4), round(b / dx * 4), round(c / dx * 4))] for a, b, c in zip(x, y, z)])
    mb = {(int(mirror[a]), int(mirror[b])) for a, b in zip(rows[broken], col[broken])}
    assert mb == {(int(a), int(b)) for a, b in zip(rows[broken], col[broken])}


def test_adr_integrate_relaxes_a_spring_to_its_static_solution():
    # One free particle on a linear spring, F = -k u, under a body force b:
    # ADR (Integrate.cu:62-129) must settle at u = b / k.  A flipped sign or a
    # dropped term in the update or the damping estimate does not.
    k, mass, dt = 0.1, 1.0, 1.0
    U, V, VH, VHO, FO = ([np.zeros(1), np.zeros(1)] for _ in range(5))
    B = [np.array([1.0]), np.array([-2.0])]
    for it in range(400):
        F = [-k * U[0], -k * U[1]]
        cn, disp = O.adr_terms(2, 1, mass, dt, U, F, FO, VHO)
        c = O.adr_c(float(cn.sum()), float(disp.sum()))
        if 3 <= it < 60:  # later f - f_old and v_old both vanish and the quotient loses digits
            assert abs(c - 2.0 * math.sqrt(k / mass)) < 1e-6  # the local Rayleigh quotient of the spring
        O.adr_integrate(2, 1, it, c, mass, dt, F, B, VH, V, VHO, U, FO)
    assert abs(U[0][0] - 10.0) < 1e-9 and abs(U[1][0] + 20.0) < 1e-9
    assert abs(V[0][0]) < 1e-9 and abs(V[1][0]) < 1e-9


def test_resid_terms_and_strip_load_hand_made():
    # free particles: |F + b|^2 and |b|^2; strip particles: |F|^2
    F = [np.array([3.0, 1.0, 6.0]), np.array([4.0, 0.0, 8.0])]
    B = [np.array([0.0, 1.0, 9.0]), np.array([0.0, 2.0, 9.0])]
    rnum, rden = O.resid_terms(2, 2, 3, F, B)
    assert rnum.tolist() == [25.0, 8.0, 0.0] and rden.tolist() == [0.0, 5.0, 100.0]
    # Solve_Plate_Crack_Grow.cu:20-27: top strip up, bottom strip down by load*dt
    uy = np.array([1.0, 2.0, 0.0, 0.0, 0.0, 7.0])
    O.quick_load(2, 2, 1, 20.0, 1e-3, uy)
    assert uy.tolist() == [1.0, 2.0, 0.02, 0.02, -0.02, 7.0]


def test_cylinder_half_steps_follow_the_constant_force_parabola():
    # Solve_Kalthoff_Winkler.cu:24-37,101-110,255-265: with a constant force
    # from the first step on, velocity Verlet starting at rest acceleration
    # gives disp_n = n v0 dt + a dt^2 n (n - 1) / 2 exactly (in exact arithmetic).
    rad, mass, height, v0 = 0.025, 1.57, 0.05, -32.0
    st = O.cylinder_init(np.array([rad, mass, height, v0, 0.0, 0.1, 0.0]))
    assert st["vel"].tolist() == [0.0, v0, 0.0] and not st["disp"].any()
    dt, force = 1e-6, np.array([0.3, -2.0e3, 0.05])
    n = 200
    for _ in range(n):
        O.cylinder_half1(dt, st)
        st["force"][:] = force
        O.cylinder_half2(dt, mass, st)
        assert not st["force"].any()  # consumed by the half step
    a = force / mass
    expect = np.array([0.0, n * v0 * dt, 0.0]) + a * dt * dt * n * (n - 1) / 2
    assert np.allclose(st["disp"], expect, rtol=1e-10, atol=1e-15)
    assert np.allclose(st["vel"], np.array([0.0, v0, 0.0]) + a * dt * (n - 0.5), rtol=1e-10, atol=1e-15)

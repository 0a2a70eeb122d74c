"""Pins of the oracle's LPS force and of the radial branch of its contact
projection (CPU only), and proof that they pin: each plausible transcription
mistake, planted in a copy of the oracle source, makes one of them fail.

LPS force (Force_State.cu:242-306 in 3D, :359-392 in 2D) is pinned by
  * two-particle closed forms (3D: theta = 3e/xi, f = 18 K e / xi^2 along the
    deformed bond; 2D: theta = 2c e/xi, f = 2 e/xi^2 (2c^2 (2K - 32G/9) + 8G)
    with c = (2 nu - 1)/(nu - 1));
  * the energy identity: on an irregular cluster with uneven weighted volumes
    and dilatations, the force density is minus the gradient of the LPS
    energy over vol (central differences).  3D: E = sum_i vol [K theta_i^2/2 +
    15G/(2 m_i) sum_j w fac vol (e_ij - theta_i xi_ij/3)^2] (Silling's
    linear peridynamic solid); 2D: the potential whose gradient the
    reference's 2D force is, E = sum_i vol [(K - 16G/9) theta_i^2/2 +
    4G/m_i sum_j w fac vol e_ij^2];
  * the continuum limit in the bulk: under u_x = a y^2 the force density is
    G u_x'' = 2 G a (shear modulus G in both dimensions), under u_x = a x^2
    it is 2a (K + 4G/3) in 3D (Navier), within the lattice's anisotropy.
Contact (Solve_Kalthoff_Winkler.cu:177-239): a particle that is nearer the
cylinder's side wall than its caps is moved radially onto the wall.
"""
import math
import os
import subprocess
import sys
import tempfile

import numpy as np
import pytest

from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_DIR = os.path.join(os.path.dirname(HERE), "oracle")

E_MOD, NU = 70e9, 0.33
K3 = E_MOD / (3 * (1 - 2 * NU))
G = E_MOD / (2 * (1 + NU))
K2 = E_MOD / (2 * (1 - NU))
C2 = (2 * NU - 1) / (NU - 1)
NO_BREAK = 1e30


def _csr_brute(X, delta):
    """Rows ascending, strict |xi| < delta (Neighbor.cu:51)."""
    P = np.stack(X, axis=1)
    d2 = ((P[:, None, :] - P[None, :, :]) ** 2).sum(-1)
    N = len(P)
    rows = [np.nonzero((d2[i] < delta * delta) & (np.arange(N) != i))[0] for i in range(N)]
    row_ptr = np.zeros(N + 1, dtype=np.int64)
    row_ptr[1:] = np.cumsum([len(r) for r in rows])
    col = np.concatenate(rows).astype(np.int32)
    i = np.repeat(np.arange(N), np.diff(row_ptr))
    idist = np.sqrt(d2[i, col])
    return row_ptr, col, idist, i


def _lps(dim, X, U, delta, vol, fac, K):
    """theta and force of the oracle on all rows (N_int = N, nothing breaks)."""
    N = len(X[0])
    row_ptr, col, idist, i = _csr_brute(X, delta)
    fac = fac(i, col) if callable(fac) else np.full(len(col), fac)
    w, m = O.lps_weights(row_ptr, idist, fac, delta, vol)
    fail = np.ones(len(col), dtype=np.uint8)
    e, theta = np.zeros(len(col)), np.zeros(N)
    O.lps_theta(dim, N, row_ptr, col, X, U, idist, fac, w, m, fail, vol, NU, 1, e, theta)
    F = [np.zeros(N) for _ in range(dim)]
    O.lps_force(dim, N, row_ptr, col, X, U, idist, fac, w, m, theta, e, vol, NU, K, G, NO_BREAK,
                -np.inf, fail, F)
    return dict(row_ptr=row_ptr, col=col, i=i, idist=idist, fac=fac, w=w, m=m, theta=theta, F=np.stack(F))


# ------------------------------------------------------------ closed forms
def check_two_particle_3d():
    xi, e, vol, delta = 1e-3, 1.3e-6, 1e-9, 3e-3
    n0 = np.array([0.36, -0.48, 0.8])               # reference bond direction
    n1 = np.array([0.0, 0.6, 0.8])                  # deformed bond direction
    X = [np.array([0.0, xi * n0[d]]) for d in range(3)]
    U = [np.array([0.0, (xi + e) * n1[d] - xi * n0[d]]) for d in range(3)]
    r = _lps(3, X, U, delta, vol, 0.7, K3)
    e_k = math.sqrt(sum(((xi + e) * n1[d]) ** 2 for d in range(3))) - xi
    assert np.allclose(r["theta"], 3 * e_k / xi, rtol=1e-12, atol=0)
    f = 18 * K3 * e_k / xi ** 2
    assert np.allclose(r["F"][:, 0], f * n1, rtol=1e-9, atol=0)
    assert np.allclose(r["F"][:, 1], -f * n1, rtol=1e-9, atol=0)


def check_two_particle_2d():
    xi, e, vol, delta = 1e-3, 1.1e-6, 1e-9, 3e-3
    n0, n1 = np.array([0.6, 0.8]), np.array([-0.28, 0.96])
    X = [np.array([0.0, xi * n0[d]]) for d in range(2)]
    U = [np.array([0.0, (xi + e) * n1[d] - xi * n0[d]]) for d in range(2)]
    r = _lps(2, X, U, delta, vol, 0.9, K2)
    e_k = math.hypot((xi + e) * n1[0], (xi + e) * n1[1]) - xi
    assert np.allclose(r["theta"], 2 * C2 * e_k / xi, rtol=1e-12, atol=0)
    f = 2 * e_k / xi ** 2 * (2 * C2 * C2 * (2 * K2 - 32 * G / 9) + 8 * G)
    assert np.allclose(r["F"][:, 0], f * n1, rtol=1e-9, atol=0)
    assert np.allclose(r["F"][:, 1], -f * n1, rtol=1e-9, atol=0)


def check_theta_broken_bonds_2d():
    """Broken bonds leave the 2D dilatation with theta_skip_broken = 1 (the CPU
    twin, Force_State.cu:345) and stay in it with 0 (the reference's 2D GPU
    kernel, whose fail test is commented out, Force_State.cu:65)."""
    xi, e, vol, delta = 1e-3, 1.5e-6, 1e-9, 1.5e-3   # bonds 0-1 and 0-2 only
    X = [np.array([0.0, xi, -xi]), np.zeros(3)]
    U = [np.array([0.0, e, 0.0]), np.zeros(3)]
    row_ptr, col, idist, i = _csr_brute(X, delta)
    fac = np.ones(len(col))
    w, m = O.lps_weights(row_ptr, idist, fac, delta, vol)
    fail = np.ones(len(col), dtype=np.uint8)
    fail[(i == 0) & (col == 1)] = 0
    for skip in (1, 0):
        th, ek = np.zeros(3), np.zeros(len(col))
        O.lps_theta(2, 3, row_ptr, col, X, U, idist, fac, w, m, fail, vol, NU, skip, ek, th)
        w01 = math.exp(-(xi / delta) ** 2)
        want0 = 0.0 if skip else 2 * C2 / m[0] * w01 * xi * e * vol
        assert math.isclose(th[0], want0, rel_tol=1e-12, abs_tol=1e-300)
        assert math.isclose(th[1], 2 * C2 / m[1] * w01 * xi * e * vol, rel_tol=1e-12)


# --------------------------------------------------------- energy identity
def _cluster(dim, seed):
    rng = np.random.default_rng(seed)
    shape = (4, 4, 3) if dim == 3 else (6, 5)
    g = np.stack(np.meshgrid(*[np.arange(s) for s in shape], indexing="ij"), -1).reshape(-1, dim)
    dx = 1e-3
    P = (g + rng.uniform(-0.25, 0.25, g.shape)) * dx
    X = [P[:, d].copy() for d in range(dim)]
    U = [rng.uniform(-1, 1, len(P)) * 2e-7 for _ in range(dim)]
    return X, U, dx


def _sym_fac(i, c):
    """A bond factor in [0.5, 1] that is the same for both directions."""
    lo, hi = np.minimum(i, c), np.maximum(i, c)
    return 0.5 + 0.5 * ((lo * 7919 + hi * 104729) % 1000) / 999.0


def _energy(dim, X, U, r, vol):
    i, col, idist, w, fac, m = r["i"], r["col"], r["idist"], r["w"], r["fac"], r["m"]
    N = len(X[0])
    d = [X[k][col] + U[k][col] - X[k][i] - U[k][i] for k in range(dim)]
    e = np.sqrt(sum(q * q for q in d)) - idist
    cth = 3.0 if dim == 3 else 2 * C2
    theta = cth / m * np.bincount(i, weights=w * idist * e * fac * vol, minlength=N)
    if dim == 3:
        ed = e - theta[i] * idist / 3.0
        dev = 15 * G / (2 * m) * np.bincount(i, weights=w * fac * vol * ed * ed, minlength=N)
        return vol * np.sum(K3 * theta ** 2 / 2 + dev)
    dev = 4 * G / m * np.bincount(i, weights=w * fac * vol * e * e, minlength=N)
    return vol * np.sum((K2 - 16 * G / 9) * theta ** 2 / 2 + dev)


def check_energy_gradient(dim):
    X, U, dx = _cluster(dim, 11 + dim)
    vol, delta = dx ** 3, 2.6 * dx
    K = K3 if dim == 3 else K2
    r = _lps(dim, X, U, delta, vol, _sym_fac, K)
    assert np.ptp(r["m"]) > 0.2 * r["m"].max() and np.ptp(r["theta"]) > 0
    N, h = len(X[0]), 2e-10
    grad = np.zeros((dim, N))
    for k in range(dim):
        for p in range(N):
            Up = [u.copy() for u in U]
            Um = [u.copy() for u in U]
            Up[k][p] += h
            Um[k][p] -= h
            grad[k, p] = (_energy(dim, X, Up, r, vol) - _energy(dim, X, Um, r, vol)) / (2 * h)
    err = np.max(np.abs(r["F"] + grad / vol)) / np.max(np.abs(r["F"]))
    assert err < 1e-5, err


# --------------------------------------------------------- continuum limit
def _bulk(dim, n):
    g = np.stack(np.meshgrid(*[np.arange(n)] * dim, indexing="ij"), -1).reshape(-1, dim)
    dx = 1e-3
    P = (g - (n - 1) / 2.0) * dx
    return [P[:, d].copy() for d in range(dim)], dx


def check_continuum_limit(dim):
    n = 15 if dim == 3 else 31
    X, dx = _bulk(dim, n)
    delta, vol = 3.015 * dx, dx ** 3
    K = K3 if dim == 3 else K2
    half = (n - 1) / 2.0 * dx
    inner = np.all([np.abs(X[d]) <= half - 2 * delta for d in range(dim)], axis=0)
    assert inner.sum() >= 1
    fac = lambda i, c: np.clip((delta + dx / 2 - np.sqrt(sum((X[d][c] - X[d][i]) ** 2 for d in range(dim)))) / dx, 0, 1)
    a = 1e-2  # 1/m: strains of order a * half = 1e-4
    zero = [np.zeros_like(X[0]) for _ in range(dim)]
    # shear: u_x = a y^2 (divergence-free) -> f_x = 2 G a
    U = [a * X[1] ** 2] + zero[1:]
    r = _lps(dim, X, U, delta, vol, fac, K)
    ratio = r["F"][0][inner] / (2 * G * a)
    assert np.all(np.abs(ratio - 1) < 0.1), ratio
    assert np.max(np.abs(r["F"][1:, inner])) < 1e-3 * 2 * G * a
    if dim == 3:
        # u_x = a x^2: f_x = G u_x'' + (K + G/3) (div u)' = 2a (K + 4G/3)
        U = [a * X[0] ** 2] + zero[1:]
        r = _lps(dim, X, U, delta, vol, fac, K)
        ratio = r["F"][0][inner] / (2 * a * (K3 + 4 * G / 3))
        assert np.all(np.abs(ratio - 1) < 0.1), ratio
    else:
        # the reference's 2D model: stress (K - 16G/9) c^2 tr(eps) I + G tr(eps) I + 2G eps,
        # so u_x = a x^2 gives f_x = 2a (3G + c^2 (K - 16G/9))
        U = [a * X[0] ** 2] + zero[1:]
        r = _lps(dim, X, U, delta, vol, fac, K)
        ratio = r["F"][0][inner] / (2 * a * (3 * G + C2 * C2 * (K2 - 16 * G / 9)))
        assert np.all(np.abs(ratio - 1) < 0.1), ratio


# ----------------------------------------------------------------- contact
def check_contact_radial():
    rad, height = 0.025, 0.05
    cyl = [rad, 1.57, height, -32.0, 0.0, 0.1, 0.0]
    st = O.cylinder_init(cyl)
    st["disp"][1] = -0.03  # axis along y through (0, 0.07, 0); caps at y = 0.045 and 0.095
    # 0: 2.64 mm inside the side wall, 5 mm below the centre -> radial;
    # 1: the same point with dmg = 1 -> untouched; 2: outside -> untouched;
    # 3: nearer the bottom cap than the wall -> cap branch
    X = [np.array([0.018, 0.018, 0.03, 0.003]),
         np.array([0.065, 0.065, 0.065, 0.046]),
         np.array([-0.012, -0.012, 0.0, 0.004])]
    U = [np.full(4, 1e-6), np.full(4, -2e-6), np.full(4, 5e-7)]
    V = [np.full(4, 3.0), np.full(4, -1.0), np.full(4, 0.5)]
    U0 = [u.copy() for u in U]
    V0 = [v.copy() for v in V]
    dmg = np.array([0.2, 1.0, 0.0, 0.0])
    dt, dens, vol = 1e-6, 8000.0, 1e-9
    R = O.contact(4, dmg, X, U, V, cyl, st["disp"], dt, dens, vol)
    px, py, pz = (X[d][0] + U0[d][0] for d in range(3))
    px, py, pz = px - 0.0, py - 0.07, pz - 0.0
    dpd = math.hypot(px, pz)
    assert dpd < rad and rad - dpd <= height / 2 - abs(py)
    want = [px / dpd * rad, 0.07 + py, pz / dpd * rad]
    for d in range(3):
        assert math.isclose(X[d][0] + U[d][0], want[d], rel_tol=1e-12)
        assert math.isclose(V[d][0], (U[d][0] - U0[d][0]) / dt, rel_tol=1e-12)
        assert math.isclose(R[d][0], -dens * (V[d][0] - V0[d][0]) / dt * vol, rel_tol=1e-12)
        for q in (1, 2):
            assert U[d][q] == U0[d][q] and V[d][q] == V0[d][q] and R[d][q] == 0.0
    assert math.isclose(math.hypot(X[0][0] + U[0][0], X[2][0] + U[2][0]), rad, rel_tol=1e-12)
    assert math.isclose(X[1][3] + U[1][3], 0.045, rel_tol=1e-12)
    assert math.isclose(X[0][3] + U[0][3], X[0][3] + U0[0][3], rel_tol=1e-15)


CHECKS = {
    "two_particle_3d": check_two_particle_3d,
    "two_particle_2d": check_two_particle_2d,
    "energy_gradient_3d": lambda: check_energy_gradient(3),
    "energy_gradient_2d": lambda: check_energy_gradient(2),
    "continuum_3d": lambda: check_continuum_limit(3),
    "continuum_2d": lambda: check_continuum_limit(2),
    "contact_radial": check_contact_radial,
    "theta_broken_2d": check_theta_broken_bonds_2d,
}


@pytest.mark.parametrize("name", sorted(CHECKS))
def test_oracle_pin(name):
    CHECKS[name]()


# ------------------------------------------------------- planted mistakes
# (text in oracle/pd_oracle.c, the mistake planted there)
MISTAKES = [
    ("15.0 * G / m[i] * w[k] * (e[k] - theta[i] * idist[k] / 3.0)",
     "8.0 * G / m[i] * w[k] * (e[k] - theta[i] * idist[k] / 3.0)"),            # 15G -> 8G
    ("double ti = 3.0 * K * theta[i]", "double ti = 1.0 * K * theta[i]"),      # 3K -> K
    ("double tj = 3.0 * K * theta[c] * w[k] * idist[k] / m[c]",
     "double tj = 3.0 * K * theta[i] * w[k] * idist[k] / m[c]"),               # theta_c -> theta_i
    ("15.0 * G / m[c] * w[k] * (e[k] - theta[c]", "15.0 * G / m[i] * w[k] * (e[k] - theta[c]"),  # m_c -> m_i
    ("(e[k] - theta[c] * idist[k] / 3.0)", "(e[k] - theta[c] * idist[k] / 2.0)"),  # the deviatoric 1/3
    ("* theta[c] * w[k] * idist[k] / m[c]\n                            + 8.0 * G / m[c]",
     "* theta[i] * w[k] * idist[k] / m[c]\n                            + 8.0 * G / m[c]"),  # 2D theta_c
    ("+ 8.0 * G / m[c] * w[k] * e[k];", "+ 8.0 * G / m[i] * w[k] * e[k];"),   # 2D m_c -> m_i
    ("(2.0 * K - 32.0 * G / 9.0) * theta[i]", "(2.0 * K - 32.0 * G / 3.0) * theta[i]"),  # 2D 32G/9
    ("+ 8.0 * G / m[i] * w[k] * e[k];", "+ 15.0 * G / m[i] * w[k] * e[k];"),  # 2D 8G
    ("rx = dpx / dpd * rad; ry = dpy; rz = dpz / dpd * rad;",
     "rx = dpx / dpd * rad; ry = dpy; rz = dpz;"),                             # radial z dropped
    ("rx = dpx / dpd * rad; ry = dpy; rz = dpz / dpd * rad;",
     "rx = dpx; ry = dpy; rz = dpz / dpd * rad;"),                             # radial x dropped
    ("if (rad - dpd <= height / 2.0 - fabs(dpy))", "if (rad - dpd >= height / 2.0 - fabs(dpy))"),
    ("if (theta_skip_broken && fail[k] != 1) continue;", "if (fail[k] != 1) continue;"),  # switch ignored
]

_RUNNER = """
import importlib.util, sys
sys.path.insert(0, {root!r})
spec = importlib.util.spec_from_file_location("pins", {path!r})
mod = importlib.util.module_from_spec(spec)
spec.loader.exec_module(mod)
failed = []
for name, fn in sorted(mod.CHECKS.items()):
    try:
        fn()
    except AssertionError:
        failed.append(name)
print("FAILED", " ".join(failed))
"""


def _run_checks_with(libpath):
    env = dict(os.environ, PD_ORACLE_LIB=libpath, OMP_NUM_THREADS="2")
    code = _RUNNER.format(root=os.path.dirname(HERE), path=os.path.abspath(__file__))
    out = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, timeout=600)
    assert out.returncode == 0, out.stderr[-2000:]
    line = [ln for ln in out.stdout.splitlines() if ln.startswith("FAILED")][-1]
    return line.split()[1:]


def test_planted_oracle_mistakes_fail_a_pin():
    src = open(os.path.join(ORACLE_DIR, "pd_oracle.c")).read()
    with tempfile.TemporaryDirectory() as tmp:
        for k, (old, new) in enumerate(MISTAKES):
            assert src.count(old) == 1, old
            mdir = os.path.join(tmp, f"m{k}")
            os.makedirs(mdir)
            with open(os.path.join(mdir, "pd_oracle.c"), "w") as fh:
                fh.write(src.replace(old, new))
            lib = os.path.join(mdir, "liboracle.so")
            subprocess.run(["gcc", "-O2", "-std=c11", "-fPIC", "-fopenmp", "-ffp-contract=off", "-shared",
                            "-I", ORACLE_DIR, "-o", lib, os.path.join(mdir, "pd_oracle.c"), "-lm"], check=True)
            failed = _run_checks_with(lib)
            assert failed, f"planted mistake {new!r} passed every pin"

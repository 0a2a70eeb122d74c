"""ctypes wrapper over liboracle.so, the FP64 CPU oracle.

Test infrastructure only: tests/, the smoke check and bench.py's cpu_baseline
leg import this module as the checker.  The product package never does.

Each wrapper takes and returns numpy arrays; the C functions it calls cite the
reference file:line they restate (see pd_oracle.c).
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

D = C.POINTER(C.c_double)
I32 = C.POINTER(C.c_int)
I64 = C.POINTER(C.c_int64)
U8 = C.POINTER(C.c_uint8)
DPP = C.POINTER(D)


class OrcParams(C.Structure):
    _fields_ = [(n, C.c_int) for n in ("dim", "bs", "nx", "ny", "nz", "N", "N_int", "MN")] + [
        (n, C.c_double)
        for n in ("dx", "delta", "radij", "sedload", "vol", "bc", "mass", "thick", "area",
                  "emod", "pratio", "dens", "K", "G", "s0", "dt")
    ]


def build():
    """Compile liboracle.so with its Makefile (gcc -O2 -fopenmp -ffp-contract=off)."""
    subprocess.run(["make", "-s", "-C", _HERE, "-B"], check=True)


_lib = None


def lib():
    """The loaded oracle library.  PD_ORACLE_LIB names another build of the
    same source (the tests' planted-mistake copies); otherwise liboracle.so
    next to this file, built on first use."""
    global _lib
    if _lib is None:
        path = os.environ.get("PD_ORACLE_LIB", _LIB_PATH)
        if path == _LIB_PATH and not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(path)
        L.orc_cal_parameter.argtypes = [C.POINTER(OrcParams), C.c_int, C.c_int] + [C.c_double] * 10 + [C.c_double, C.c_int]
        L.orc_coord_plate.argtypes = [C.c_int, C.c_int, C.c_double, D, D]
        L.orc_coord_plate_strips.argtypes = [C.c_int, C.c_int, C.c_double, D, D]
        L.orc_coord_kw.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, D, D, D]
        L.orc_coord_kw.restype = C.c_int
        L.orc_coord_block_jitter.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_uint64, D, D, D]
        L.orc_neighbors_brute.argtypes = [C.c_int, C.c_int, D, D, D, C.c_double, C.c_int, I32, I32]
        L.orc_neighbors_brute.restype = C.c_int
        L.orc_neighbors_cell.argtypes = [C.c_int, C.c_int, D, D, D, C.c_double, I64, I32]
        L.orc_neighbors_cell.restype = C.c_int
        L.orc_vol_corr.argtypes = [C.c_int, C.c_int, I64, I32, D, D, D, C.c_double, C.c_double, D, D]
        L.orc_surface_correct.argtypes = [C.c_int, C.c_int, I64, I32, D, D, D, D, D,
                                          C.c_double, C.c_double, C.c_double, C.c_int, D, D]
        L.orc_set_crack_2d.argtypes = [C.c_int, I64, I32, D, D, C.c_double, C.c_double, C.c_double, C.c_double, U8]
        L.orc_set_crack_kw.argtypes = [C.c_int, I64, I32, D, D, C.c_double, U8]
        L.orc_pmb_force.argtypes = [C.c_int, C.c_int, I64, I32, DPP, DPP, D, D, D,
                                    C.c_double, C.c_double, C.c_double, C.c_double, U8, DPP]
        L.orc_lps_weights.argtypes = [C.c_int, I64, D, D, C.c_double, C.c_double, D, D]
        L.orc_lps_theta.argtypes = [C.c_int, C.c_int, I64, I32, DPP, DPP, D, D, D, D, U8,
                                    C.c_double, C.c_double, C.c_int, D, D]
        L.orc_lps_force.argtypes = [C.c_int, C.c_int, I64, I32, DPP, DPP, D, D, D, D, D, D,
                                    C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                                    C.c_double, U8, DPP]
        L.orc_verlet_half1.argtypes = [C.c_int, C.c_int, C.c_double, DPP, DPP, DPP]
        L.orc_quick_load.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, D]
        L.orc_verlet_half2.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, DPP, DPP, DPP, DPP]
        L.orc_sum_tree.argtypes = [D, C.c_int64, C.c_int]
        L.orc_sum_tree.restype = C.c_double
        L.orc_adr_terms.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, DPP, DPP, DPP, DPP, D, D]
        L.orc_adr_c.argtypes = [C.c_double, C.c_double]
        L.orc_adr_c.restype = C.c_double
        L.orc_adr_integrate.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                        DPP, DPP, DPP, DPP, DPP, DPP, DPP]
        L.orc_resid_terms.argtypes = [C.c_int, C.c_int, C.c_int, DPP, DPP, D, D]
        L.orc_damage.argtypes = [C.c_int, I64, U8, D, C.c_double, D]
        L.orc_cylinder_init.argtypes = [D, D, D, D, D]
        L.orc_cylinder_half1.argtypes = [C.c_double, D, D, D]
        L.orc_contact.argtypes = [C.c_int, D, DPP, DPP, DPP, D, D, C.c_double, C.c_double, C.c_double, DPP]
        L.orc_cylinder_half2.argtypes = [C.c_double, C.c_double, D, D, D]
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_get_threads.restype = C.c_int
        _lib = L
    return _lib


def _p(a, t):
    if a is None:
        return None
    return a.ctypes.data_as(t)


def _pp(arrs):
    """Array-of-pointers for a [dim][N] set of float64 arrays."""
    ptrs = (D * len(arrs))(*[a.ctypes.data_as(D) for a in arrs])
    return C.cast(ptrs, DPP), arrs


def _f64(n):
    return np.zeros(n, dtype=np.float64)


# ---------------------------------------------------------------- parameters
def cal_parameter(dim, bs, dt, dx, thick, length, height, width, emod, pratio, dens, s0,
                  delta_factor=3.015, ref_k_bug=False):
    p = OrcParams()
    lib().orc_cal_parameter(C.byref(p), dim, bs, dt, dx, thick, length, height, width, emod,
                            pratio, dens, s0, delta_factor, int(ref_k_bug))
    return p


# --------------------------------------------------------------- coordinates
def coord_plate(nx, ny, dx):
    x, y = _f64(nx * ny), _f64(nx * ny)
    lib().orc_coord_plate(nx, ny, dx, _p(x, D), _p(y, D))
    return x, y


def coord_plate_strips(nx, ny, dx):
    n = nx * (ny + 6)
    x, y = _f64(n), _f64(n)
    lib().orc_coord_plate_strips(nx, ny, dx, _p(x, D), _p(y, D))
    return x, y


def coord_kw(nx, ny, nz, dx, rad):
    n = lib().orc_coord_kw(nx, ny, nz, dx, rad, None, None, None)
    x, y, z = _f64(n), _f64(n), _f64(n)
    lib().orc_coord_kw(nx, ny, nz, dx, rad, _p(x, D), _p(y, D), _p(z, D))
    return x, y, z


def coord_block_jitter(nx, ny, nz, dx, jitter, seed):
    n = nx * ny * nz
    x, y, z = _f64(n), _f64(n), _f64(n)
    lib().orc_coord_block_jitter(nx, ny, nz, dx, jitter, seed, _p(x, D), _p(y, D), _p(z, D))
    return x, y, z


# ---------------------------------------------------------------- neighbours
def neighbors_brute(dim, x, y, z, delta, MN):
    N = len(x)
    NN = np.zeros(N, dtype=np.int32)
    NL = np.zeros(N * MN, dtype=np.int32)
    rc = lib().orc_neighbors_brute(dim, N, _p(x, D), _p(y, D), _p(z, D), delta, MN,
                                   _p(NN, I32), _p(NL, I32))
    if rc != 0:
        raise OverflowError("neighbour count exceeds MN")
    return NN, NL


def dense_to_csr(NN, NL, MN):
    row_ptr = np.zeros(len(NN) + 1, dtype=np.int64)
    row_ptr[1:] = np.cumsum(NN)
    col = np.concatenate([NL[i * MN:i * MN + NN[i]] for i in range(len(NN))]).astype(np.int32) \
        if len(NN) else np.zeros(0, np.int32)
    return row_ptr, col


def neighbors_cell(dim, x, y, z, delta):
    N = len(x)
    row_ptr = np.zeros(N + 1, dtype=np.int64)
    maxrow = lib().orc_neighbors_cell(dim, N, _p(x, D), _p(y, D), _p(z, D), delta, _p(row_ptr, I64), None)
    col = np.zeros(int(row_ptr[-1]), dtype=np.int32)
    lib().orc_neighbors_cell(dim, N, _p(x, D), _p(y, D), _p(z, D), delta, _p(row_ptr, I64), _p(col, I32))
    return row_ptr, col, maxrow


# ------------------------------------------------------------ setup physics
def vol_corr(dim, row_ptr, col, x, y, z, delta, dx):
    nb = len(col)
    idist, fac = _f64(nb), _f64(nb)
    lib().orc_vol_corr(dim, len(x), _p(row_ptr, I64), _p(col, I32), _p(x, D), _p(y, D), _p(z, D),
                       delta, dx, _p(idist, D), _p(fac, D))
    return idist, fac


def surface_correct(dim, row_ptr, col, x, y, z, idist, fac, bc, vol, sedload, scr_mode=0):
    N = len(x)
    scr = _f64(len(col))
    fncst = _f64(3 * N)
    lib().orc_surface_correct(dim, N, _p(row_ptr, I64), _p(col, I32), _p(x, D), _p(y, D), _p(z, D),
                              _p(idist, D), _p(fac, D), bc, vol, sedload, scr_mode, _p(fncst, D), _p(scr, D))
    return scr, fncst.reshape(3, N)[:dim].copy()


def set_crack_2d(N_int, row_ptr, col, x, y, clength, loc_x, loc_y, theta, fail):
    lib().orc_set_crack_2d(N_int, _p(row_ptr, I64), _p(col, I32), _p(x, D), _p(y, D),
                           clength, loc_x, loc_y, theta, _p(fail, U8))


def set_crack_kw(N_int, row_ptr, col, x, y, rad, fail):
    lib().orc_set_crack_kw(N_int, _p(row_ptr, I64), _p(col, I32), _p(x, D), _p(y, D), rad, _p(fail, U8))


# --------------------------------------------------------------------- force
def pmb_force(dim, n_rows, row_ptr, col, X, U, idist, fac, scr, vol, bc, s0, break_y_min, fail, F):
    xp, _a = _pp(X)
    up, _b = _pp(U)
    fp, _c = _pp(F)
    lib().orc_pmb_force(dim, n_rows, _p(row_ptr, I64), _p(col, I32), xp, up, _p(idist, D), _p(fac, D),
                        _p(scr, D), vol, bc, s0, break_y_min, _p(fail, U8), fp)


def lps_weights(row_ptr, idist, fac, delta, vol):
    N = len(row_ptr) - 1
    w, m = _f64(len(idist)), _f64(N)
    lib().orc_lps_weights(N, _p(row_ptr, I64), _p(idist, D), _p(fac, D), delta, vol, _p(w, D), _p(m, D))
    return w, m


def lps_theta(dim, N, row_ptr, col, X, U, idist, fac, w, m, fail, vol, pratio, skip_broken, e, theta):
    xp, _a = _pp(X)
    up, _b = _pp(U)
    lib().orc_lps_theta(dim, N, _p(row_ptr, I64), _p(col, I32), xp, up, _p(idist, D), _p(fac, D),
                        _p(w, D), _p(m, D), _p(fail, U8), vol, pratio, int(skip_broken), _p(e, D), _p(theta, D))


def lps_force(dim, N_int, row_ptr, col, X, U, idist, fac, w, m, theta, e, vol, pratio, K, G, s0,
              break_y_min, fail, F):
    xp, _a = _pp(X)
    up, _b = _pp(U)
    fp, _c = _pp(F)
    lib().orc_lps_force(dim, N_int, _p(row_ptr, I64), _p(col, I32), xp, up, _p(idist, D), _p(fac, D),
                        _p(w, D), _p(m, D), _p(theta, D), _p(e, D), vol, pratio, K, G, s0, break_y_min,
                        _p(fail, U8), fp)


# ----------------------------------------------------------------- integrate
def verlet_half1(dim, N_int, dt, V, A, U):
    vp, _a = _pp(V)
    ap, _b = _pp(A)
    up, _c = _pp(U)
    lib().orc_verlet_half1(dim, N_int, dt, vp, ap, up)


def quick_load(N_int, n_top, n_bot, load, dt, uy):
    lib().orc_quick_load(N_int, n_top, n_bot, load, dt, _p(uy, D))


def verlet_half2(dim, N_int, dt, dens, V, A, F, B):
    vp, _a = _pp(V)
    ap, _b = _pp(A)
    fp, _c = _pp(F)
    bp, _d = _pp(B)
    lib().orc_verlet_half2(dim, N_int, dt, dens, vp, ap, fp, bp)


def sum_tree(v, T):
    v = np.ascontiguousarray(v, dtype=np.float64)
    return lib().orc_sum_tree(_p(v, D), len(v), T)


def adr_terms(dim, N_int, mass, dt, U, F, FO, VHO):
    cn, disp = _f64(N_int), _f64(N_int)
    a, _1 = _pp(U)
    b, _2 = _pp(F)
    c, _3 = _pp(FO)
    d, _4 = _pp(VHO)
    lib().orc_adr_terms(dim, N_int, mass, dt, a, b, c, d, _p(cn, D), _p(disp, D))
    return cn, disp


def adr_c(cn1, cn2):
    return lib().orc_adr_c(cn1, cn2)


def adr_integrate(dim, N_int, ct, c, mass, dt, F, B, VH, V, VHO, U, FO):
    args = [_pp(a)[0] for a in (F, B, VH, V, VHO, U, FO)]
    keep = (F, B, VH, V, VHO, U, FO)  # noqa: F841 (keeps buffers alive)
    lib().orc_adr_integrate(dim, N_int, ct, c, mass, dt, *args)


def resid_terms(dim, N_int, N, F, B):
    rnum, rden = _f64(N), _f64(N)
    a, _1 = _pp(F)
    b, _2 = _pp(B)
    lib().orc_resid_terms(dim, N_int, N, a, b, _p(rnum, D), _p(rden, D))
    return rnum, rden


def damage(N_int, row_ptr, fail, fac, vol):
    dmg = _f64(N_int)
    lib().orc_damage(N_int, _p(row_ptr, I64), _p(fail, U8), _p(fac, D), vol, _p(dmg, D))
    return dmg


# -------------------------------------------------------------- KW cylinder
def cylinder_init(cyl):
    cyl = np.asarray(cyl, dtype=np.float64)
    st = {k: _f64(3) for k in ("vel", "disp", "acc", "force")}
    lib().orc_cylinder_init(_p(cyl, D), _p(st["vel"], D), _p(st["disp"], D), _p(st["acc"], D), _p(st["force"], D))
    return st


def cylinder_half1(dt, st):
    lib().orc_cylinder_half1(dt, _p(st["vel"], D), _p(st["disp"], D), _p(st["acc"], D))


def contact(N, dmg, X, U, V, cyl, cdisp, dt, dens, vol):
    R = [_f64(N) for _ in range(3)]
    cyl = np.asarray(cyl, dtype=np.float64)
    xp, _a = _pp(X)
    up, _b = _pp(U)
    vp, _c = _pp(V)
    rp, _d = _pp(R)
    lib().orc_contact(N, _p(dmg, D), xp, up, vp, _p(cyl, D), _p(cdisp, D), dt, dens, vol, rp)
    return R


def cylinder_half2(dt, mass, st):
    lib().orc_cylinder_half2(dt, mass, _p(st["vel"], D), _p(st["acc"], D), _p(st["force"], D))


def set_threads(n):
    """OpenMP threads of the oracle's loops (omp_set_num_threads)."""
    lib().orc_set_threads(int(n))


def get_threads():
    return int(lib().orc_get_threads())

"""Python mirror of the reference's stage interface over the C-ABI.

A :class:`Problem` owns one ``pd_handle``: one particle set, its CSR bonds and
one HIP stream on one GPU.  Method names follow the reference's stage
functions (Coord -> Neighbor -> Surf_Corr -> Initial -> Force ->
Integrate/Reduce -> Save, see the solvers at Solve_Plate_Tensile.cu:93-134,
Solve_Plate_Crack_Grow.cu:100-167 and Solve_Kalthoff_Winkler.cu:372-436);
arguments keep the reference's meaning.  Every call runs on the GPU; errors
raise :class:`PdError` with the library's message.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import FIELDS, PdCyl, PdError, PdParams, dptr, f64, lib


def _check(h, rc):
    if rc != _lib.PD_OK:
        msg = lib().pd_last_error(h) if h else b""
        raise PdError(rc, (msg or b"").decode(errors="replace"))


def params(dim, model, dt, dx, thick=0.0, length=0.05, height=0.05, width=0.0, emod=192.0e9,
           pratio=1.0 / 3.0, dens=8000.0, s0=0.04472, delta_factor=3.015, ref_k_bug=False):
    """BaseModel_Parameter (Base.cuh:72-183).  Returns (PdParams, (nx, ny, nz), nx*ny*nz)."""
    p = PdParams()
    nxyz = (C.c_int32 * 3)()
    n = C.c_int32()
    rc = lib().pd_params_init(C.byref(p), dim, model, dt, dx, thick, length, height, width, emod,
                              pratio, dens, s0, delta_factor, int(ref_k_bug), nxyz, C.byref(n))
    _check(None, rc)
    return p, (nxyz[0], nxyz[1], nxyz[2]), n.value


def stream_copy(nbytes, reps=10, device=0):
    """Achieved GB/s of a plain device copy (HBM reference)."""
    out = C.c_double()
    _check(None, lib().pd_stream_copy(device, nbytes, reps, C.byref(out)))
    return out.value


def gen_plate(nx, ny, dx):
    x, y = np.zeros(nx * ny), np.zeros(nx * ny)
    _check(None, lib().pd_gen_plate(nx, ny, dx, dptr(x), dptr(y)))
    return x, y


def gen_plate_strips(nx, ny, dx):
    n = nx * (ny + 6)
    x, y = np.zeros(n), np.zeros(n)
    _check(None, lib().pd_gen_plate_strips(nx, ny, dx, dptr(x), dptr(y)))
    return x, y


def gen_kw(nx, ny, nz, dx, rad):
    n = lib().pd_gen_kw(nx, ny, nz, dx, rad, None, None, None)
    x, y, z = np.zeros(n), np.zeros(n), np.zeros(n)
    lib().pd_gen_kw(nx, ny, nz, dx, rad, dptr(x), dptr(y), dptr(z))
    return x, y, z


def gen_block_jitter(nx, ny, nz, dx, jitter, seed):
    n = nx * ny * nz
    x, y, z = np.zeros(n), np.zeros(n), np.zeros(n)
    _check(None, lib().pd_gen_block_jitter(nx, ny, nz, dx, jitter, seed, dptr(x), dptr(y), dptr(z)))
    return x, y, z


class Problem:
    """One peridynamic problem on one GPU (a ``pd_handle``)."""

    def __init__(self, p: PdParams, device: int = 0):
        self._h = C.c_void_p()
        _check(None, lib().pd_create(C.byref(p), device, C.byref(self._h)))
        self.dim = p.dim
        self.N = 0
        self.N_int = 0

    # -------------------------------------------------------------- lifetime
    def close(self):
        if self._h:
            lib().pd_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _c(self, rc):
        _check(self._h, rc)

    @property
    def params(self) -> PdParams:
        p = PdParams()
        self._c(lib().pd_get_params(self._h, C.byref(p)))
        return p

    def set_params(self, p: PdParams):
        self._c(lib().pd_set_params(self._h, C.byref(p)))

    # ----------------------------------------------------------------- setup
    def set_positions(self, x, y, z=None, n_int=None):
        x, y = f64(x), f64(y)
        z = f64(z) if z is not None else None
        n = len(x)
        n_int = n if n_int is None else n_int
        self._c(lib().pd_set_positions(self._h, dptr(x), dptr(y), dptr(z), n_int, n))
        self.N, self.N_int = n, n_int

    def build_neighbors(self):
        self._c(lib().pd_build_neighbors(self._h))

    @property
    def num_bonds(self):
        nb = C.c_int64()
        self._c(lib().pd_num_bonds(self._h, C.byref(nb)))
        return nb.value

    def csr(self):
        rp = np.zeros(self.N + 1, dtype=np.int32)
        col = np.zeros(self.num_bonds, dtype=np.int32)
        self._c(lib().pd_get_csr(self._h, rp.ctypes.data_as(_lib.I32), col.ctypes.data_as(_lib.I32)))
        return rp, col

    def vol_corr(self):
        self._c(lib().pd_vol_corr(self._h))

    def surface_correct(self):
        self._c(lib().pd_surface_correct(self._h))

    def init_state(self, static=False):
        self._c(lib().pd_init_state(self._h, _lib.PD_INIT_STATIC if static else _lib.PD_INIT_DYNAMIC))

    def set_crack_2d(self, clength, loc_x, loc_y, theta):
        self._c(lib().pd_set_crack_2d(self._h, clength, loc_x, loc_y, theta))

    def set_notches_kw(self, rad):
        self._c(lib().pd_set_notches_kw(self._h, rad))

    def lps_weights(self):
        self._c(lib().pd_lps_weights(self._h))

    # ------------------------------------------------------------------ step
    def force(self):
        self._c(lib().pd_force(self._h))

    def last_force_ms(self):
        ms = C.c_double()
        self._c(lib().pd_last_force_ms(self._h, C.byref(ms)))
        return ms.value

    def time_force_passes(self, reps=10):
        """Average device ms of (PMB or LPS dilatation pass, LPS force pass)."""
        out = np.zeros(2)
        self._c(lib().pd_time_force_passes(self._h, reps, dptr(out)))
        return float(out[0]), float(out[1])

    def set_group(self, group):
        self._c(lib().pd_set_group(self._h, group))

    def set_pass_group(self, pass_, group):
        self._c(lib().pd_set_pass_group(self._h, pass_, group))

    def set_lds(self, mode):
        """LDS-staged tiles: -1 where they fit (default), 0 off, 1 required."""
        self._c(lib().pd_set_lds(self._h, mode))

    def set_fp64_group(self, group):
        """Lanes per row of the FP64 PMB kernel (1, 4, 8, 16, 32; -1 = by dimension)."""
        self._c(lib().pd_set_fp64_group(self._h, group))

    def set_coef_mode(self, mode):
        """Bond coefficients: -1 recompute from |xi| where they match (default), 0 stream, 1 recompute."""
        self._c(lib().pd_set_coef_mode(self._h, mode))

    def coef_recomputed(self):
        r = C.c_int()
        self._c(lib().pd_coef_info(self._h, C.byref(r)))
        return bool(r.value)

    def set_lds_block(self, threads, pass_=None):
        if pass_ is None:
            self._c(lib().pd_set_lds_block(self._h, threads))
        else:
            self._c(lib().pd_set_lds_pass_block(self._h, pass_, threads))

    def lds_info(self):
        a, n, m = C.c_int(), C.c_int(), C.c_int()
        self._c(lib().pd_lds_info(self._h, C.byref(a), C.byref(n), C.byref(m)))
        return bool(a.value), n.value, m.value

    def set_tile(self, tile_cells):
        self._c(lib().pd_set_tile(self._h, tile_cells))

    def verlet_half1(self):
        self._c(lib().pd_verlet_half1(self._h))

    def verlet_half2(self):
        self._c(lib().pd_verlet_half2(self._h))

    def run_verlet(self, nsteps):
        self._c(lib().pd_run_verlet(self._h, nsteps))

    def set_body_force_tensile(self, nx, ny):
        self._c(lib().pd_set_body_force_tensile(self._h, nx, ny))

    def adr_iterate(self, ct, sync=True):
        c, r = C.c_double(), C.c_double()
        if sync:
            self._c(lib().pd_adr_iterate(self._h, ct, C.byref(c), C.byref(r)))
            return c.value, r.value
        self._c(lib().pd_adr_iterate(self._h, ct, None, None))
        return None

    def adr_solve(self, ct0, max_iters, tol, nonfinite_ok=False):
        """ADR iterations until the relative residual is <= tol or max_iters
        ran: (iterations, residual, converged).  A non-finite state raises
        PdError unless nonfinite_ok, which reports it as not converged."""
        it, r = C.c_int(), C.c_double()
        rc = lib().pd_adr_solve(self._h, ct0, max_iters, tol, C.byref(it), C.byref(r))
        ok_codes = (_lib.PD_OK, _lib.PD_E_NOT_CONVERGED) + ((_lib.PD_E_NONFINITE,) if nonfinite_ok else ())
        if rc not in ok_codes:
            self._c(rc)
        return it.value, r.value, rc == _lib.PD_OK

    def strip_displace(self, du_top, du_bot):
        self._c(lib().pd_strip_displace(self._h, du_top, du_bot))

    def reduce_sum(self, field, comp=0):
        out = C.c_double()
        self._c(lib().pd_reduce_sum(self._h, FIELDS[field], comp, C.byref(out)))
        return out.value

    def damage(self):
        self._c(lib().pd_damage(self._h))

    def sync(self):
        self._c(lib().pd_sync(self._h))

    # ---------------------------------------------------------------- fields
    def get(self, field):
        fid = FIELDS[field]
        if fid in (FIELDS["DMG"], FIELDS["THETA"], FIELDS["M"]):
            out = np.zeros(self.N)
        elif fid >= 20:
            out = np.zeros(self.num_bonds)
        else:
            out = np.zeros(self.dim * self.N)
        self._c(lib().pd_get_field(self._h, fid, dptr(out)))
        if fid < 9:
            return out.reshape(self.dim, self.N)
        return out

    def set(self, field, arr):
        a = f64(arr).ravel()
        self._c(lib().pd_set_field(self._h, FIELDS[field], dptr(a)))

    def fail_words(self):
        w = np.zeros((self.num_bonds + 31) // 32, dtype=np.uint32)
        self._c(lib().pd_get_fail_bits(self._h, w.ctypes.data_as(_lib.U32)))
        return w

    def fail_mask(self):
        """Per-bond intact flag (uint8, 1 = intact) in CSR order."""
        w = self.fail_words()
        bits = np.unpackbits(w.view(np.uint8), bitorder="little")
        return bits[: self.num_bonds].astype(np.uint8)

    def set_fail_mask(self, mask):
        nb = self.num_bonds
        bits = np.zeros(((nb + 31) // 32) * 32, dtype=np.uint8)
        bits[:nb] = np.asarray(mask, dtype=np.uint8)
        w = np.packbits(bits, bitorder="little").view(np.uint32).copy()
        self._c(lib().pd_set_fail_bits(self._h, w.ctypes.data_as(_lib.U32)))

    # -------------------------------------------------------------- cylinder
    def cylinder_init(self, rad, mass, height, v, x, y, z):
        c = PdCyl(rad, mass, height, v, x, y, z)
        self._c(lib().pd_cylinder_init(self._h, C.byref(c)))

    def cylinder_state(self):
        out = np.zeros(12)
        self._c(lib().pd_cylinder_get(self._h, dptr(out)))
        return dict(vel=out[0:3].copy(), disp=out[3:6].copy(), acc=out[6:9].copy(), force=out[9:12].copy())

    def contact_step(self):
        self._c(lib().pd_contact_step(self._h))

    # ---------------------------------------------------------------- output
    def save_text(self, path, layout=_lib.PD_LAYOUT_PLATE, nx=0, ny=0, nz=0, rad=0.0):
        self._c(lib().pd_save_text(self._h, str(path).encode(), layout, nx, ny, nz, rad))

    def checkpoint_write(self, path, step):
        self._c(lib().pd_checkpoint_write(self._h, str(path).encode(), step))

    def checkpoint_read(self, path):
        st = C.c_int64()
        self._c(lib().pd_checkpoint_read(self._h, str(path).encode(), C.byref(st)))
        return st.value

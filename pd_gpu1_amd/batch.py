"""Batched quasi-static crack growth: many 2D PMB plates relaxed by ADR in one
persistent launch per load increment (see include/pd_gpu.h, pd_batch_*)."""
import ctypes as C
import math

import numpy as np

from . import _lib
from ._lib import PdError, PdSpecimen, dptr, lib
from .engine import params

MASK64 = (1 << 64) - 1


def splitmix64(state):
    """SplitMix64 (Steele, Lea, Flood 2014): returns (value, new_state)."""
    state = (state + 0x9E3779B97F4A7C15) & MASK64
    z = state
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
    return z ^ (z >> 31), state


def uniform01(state):
    v, state = splitmix64(state)
    return (v >> 11) * (1.0 / 9007199254740992.0), state


def choose_parts(nb, n, capacity, wmax=16, retired=None):
    """The automatic workgroups per specimen (pd_batch_choose_parts) for
    specimens of nb bonds and n particles on a device that holds `capacity`
    workgroups."""
    nb = np.ascontiguousarray(nb, dtype=np.int64)
    n = np.ascontiguousarray(n, dtype=np.int32)
    ret = None if retired is None else np.ascontiguousarray(retired, dtype=np.int32)
    return int(lib().pd_batch_choose_parts(nb.ctypes.data_as(C.POINTER(C.c_int64)), n.ctypes.data_as(_lib.I32),
                                           None if ret is None else ret.ctypes.data_as(_lib.I32), len(nb),
                                           capacity, wmax))


def plate_family(count, seed_base=0x5EED0003, n_min=40, n_max=200, dx=0.5e-3):
    """Seeded specimens of the batched family: side n ~ U{n_min..n_max}
    particles, crack length U(0.1, 0.5) * n * dx, angle U(-30, 30) degrees.
    Specimen k draws from its own SplitMix64 stream seeded seed_base + k."""
    specs = []
    for k in range(count):
        st = seed_base + k
        u, st = uniform01(st)
        n = n_min + min(int(u * (n_max - n_min + 1)), n_max - n_min)
        u, st = uniform01(st)
        frac = 0.1 + 0.4 * u
        u, st = uniform01(st)
        ang = math.radians(-30.0 + 60.0 * u)
        specs.append((n, n, frac * n * dx, ang))
    return specs


def plate_params(dx=0.5e-3):
    """2D PMB plate constants of the family: E = 192 GPa, nu = 1/3, rho = 8000,
    s0 = 0.004, t = dx, ADR fictitious dt = 1 with the mass of Base.cuh:150."""
    p, _, _ = params(2, _lib.PD_MODEL_PMB, 1.0, dx, thick=dx, length=0.02, height=0.02, emod=192e9,
                     pratio=1.0 / 3.0, dens=8000.0, s0=0.004)
    return p


class Batch:
    """A pd_batch handle: independent plates on one GPU."""

    def __init__(self, p, specs, device=0):
        arr = (PdSpecimen * len(specs))(*[PdSpecimen(int(a), int(b), float(c), float(d)) for a, b, c, d in specs])
        self._h = C.c_void_p()
        rc = lib().pd_batch_create(C.byref(p), arr, len(specs), device, C.byref(self._h))
        if rc != _lib.PD_OK:
            raise PdError(rc, "pd_batch_create failed")
        self.nspec = len(specs)
        self.specs = list(specs)

    def close(self):
        if self._h:
            lib().pd_batch_destroy(self._h)
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
        if rc != _lib.PD_OK:
            raise PdError(rc, (lib().pd_batch_last_error(self._h) or b"").decode())

    def info(self):
        n, nb, ns = C.c_int64(), C.c_int64(), C.c_int()
        self._c(lib().pd_batch_info(self._h, C.byref(n), C.byref(nb), C.byref(ns)))
        return n.value, nb.value, ns.value

    def spec_info(self, s):
        n, ni, nb = C.c_int(), C.c_int(), C.c_int64()
        self._c(lib().pd_batch_spec_info(self._h, s, C.byref(n), C.byref(ni), C.byref(nb)))
        return n.value, ni.value, nb.value

    def set_group(self, group):
        """Lanes per particle of the force phase (1 = thread per particle)."""
        self._c(lib().pd_batch_set_group(self._h, group))

    def set_parts(self, parts):
        """Workgroups per specimen: 1, 2, 4, 8, 16, or 0 to size it from the
        largest specimen's share of the work (same results)."""
        self._c(lib().pd_batch_set_parts(self._h, parts))

    def set_barrier_timeout(self, seconds):
        """Longest wait at a group barrier before an increment is abandoned."""
        self._c(lib().pd_batch_set_barrier_timeout(self._h, seconds))

    def parts_used(self):
        """Workgroups per specimen of the last increment."""
        w = C.c_int()
        self._c(lib().pd_batch_get_parts(self._h, C.byref(w)))
        return w.value

    def set_warm(self, mode):
        """1: start each increment from u + the previous increment (0: plain)."""
        self._c(lib().pd_batch_set_warm(self._h, mode))

    def set_tail(self, max_parts):
        """16, 32 or 64: finish an increment's slowest specimens in further
        launches with up to max_parts workgroups each once the device is half
        idle (0: one launch)."""
        self._c(lib().pd_batch_set_tail(self._h, int(max_parts)))

    def tail_parts(self):
        w = C.c_int()
        self._c(lib().pd_batch_get_tail_parts(self._h, C.byref(w)))
        return w.value

    def set_separation(self, ratio):
        """End a specimen's increment once its strip reaction falls below
        ratio x its last converged reaction (0: off)."""
        self._c(lib().pd_batch_set_separation(self._h, ratio))

    def retired(self):
        """Per specimen: 0 active, 1 retired unconverged, 2 retired separated."""
        out = np.zeros(self.nspec, np.int32)
        self._c(lib().pd_batch_get_retired(self._h, out.ctypes.data_as(_lib.I32)))
        return out

    def set_coef_mode(self, mode):
        """-1: recompute the bond coefficients from the reference positions
        when they equal the stored ones bit for bit (default), 0: stream them,
        1: recompute.  Results are identical."""
        self._c(lib().pd_batch_set_coef_mode(self._h, mode))

    def coef_recomputed(self):
        r = C.c_int()
        self._c(lib().pd_batch_coef_info(self._h, C.byref(r)))
        return bool(r.value)

    def set_retire(self, on):
        """1: skip specimens from the increment after one that ended unconverged."""
        self._c(lib().pd_batch_set_retire(self._h, on))

    def increment(self, du, tol, max_iters, block=1024, grid=0):
        its = np.zeros(self.nspec, dtype=np.int32)
        res = np.zeros(self.nspec)
        ms = C.c_double()
        self._c(lib().pd_batch_increment(self._h, du, tol, max_iters, block, grid,
                                         its.ctypes.data_as(_lib.I32), dptr(res), C.byref(ms)))
        return its, res, ms.value

    def get(self, s, field="U"):
        n, _, _ = self.spec_info(s)
        out = np.zeros(2 * n)
        self._c(lib().pd_batch_get_field(self._h, s, _lib.FIELDS[field], dptr(out)))
        return out.reshape(2, n)

    def set(self, s, field, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64).ravel()
        self._c(lib().pd_batch_set_field(self._h, s, _lib.FIELDS[field], dptr(a)))

    def csr(self, s):
        n, _, nb = self.spec_info(s)
        rp = np.zeros(n + 1, dtype=np.int32)
        col = np.zeros(nb, dtype=np.int32)
        fw = np.zeros((nb + 31) // 32, dtype=np.uint32)
        self._c(lib().pd_batch_get_csr(self._h, s, rp.ctypes.data_as(_lib.I32), col.ctypes.data_as(_lib.I32),
                                       fw.ctypes.data_as(_lib.U32)))
        mask = np.unpackbits(fw.view(np.uint8), bitorder="little")[:nb].astype(np.uint8)
        return rp, col, mask


def main(argv=None):
    """Run load increments of the seeded plate family and print one line per
    increment (wall time, iterations) and a JSON summary:

        python -m pd_gpu1_amd.batch --plates 512 --increments 2 --max-iters 6000
    """
    import argparse
    import json
    import time

    ap = argparse.ArgumentParser(description="batched ADR over the seeded plate family")
    ap.add_argument("--plates", type=int, default=512)
    ap.add_argument("--increments", type=int, default=1)
    ap.add_argument("--max-iters", type=int, default=6000)
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--du", type=float, default=1e-6)
    ap.add_argument("--block", type=int, default=1024)
    ap.add_argument("--grid", type=int, default=0, help="workgroups (0: as many as are resident at once)")
    ap.add_argument("--group", type=int, default=4)
    ap.add_argument("--parts", type=int, default=0, help="workgroups per specimen (0: sized from the work)")
    ap.add_argument("--warm", type=int, default=0)
    ap.add_argument("--retire", type=int, default=1)
    ap.add_argument("--tail", type=int, default=32,
                    help="widest tail launch, workgroups per specimen: 16, 32 or 64 (0: one launch)")
    ap.add_argument("--separation", type=float, default=0.0,
                    help="end a specimen's increment once its strip reaction falls below this fraction of its last "
                         "converged reaction (0: off)")
    a = ap.parse_args(argv)
    specs = plate_family(a.plates)
    rows = []
    t0 = time.perf_counter()
    with Batch(plate_params(), specs) as b:
        b.set_group(a.group)
        b.set_parts(a.parts)
        b.set_warm(a.warm)
        b.set_retire(a.retire)
        b.set_tail(a.tail)
        b.set_separation(a.separation)
        n_part, n_bonds, _ = b.info()
        print(f"plates {a.plates} particles {n_part} bonds {n_bonds}", flush=True)
        for k in range(a.increments):
            its, res, ms = b.increment(a.du, a.tol, a.max_iters, block=a.block, grid=a.grid)
            act = its > 0
            rows.append({"increment": k, "ms": round(ms, 1), "active": int(act.sum()), "parts": b.parts_used(),
                         "iterations_median": int(np.median(its[act])) if act.any() else 0,
                         "iterations_max": int(its.max()), "unconverged": int(((res > a.tol) & act).sum()),
                         "specimen_iterations": int(its.sum()), "tail_parts": b.tail_parts(),
                         "separated": int((b.retired() == 2).sum())})
            print("increment %(increment)d: %(ms).1f ms, %(active)d active, iterations median "
                  "%(iterations_median)d max %(iterations_max)d, unconverged %(unconverged)d, separated so far "
                  "%(separated)d, tail launch parts %(tail_parts)d" % rows[-1], flush=True)
    print(json.dumps({"plates": a.plates, "particles": n_part, "bonds": n_bonds, "block": a.block,
                      "group": a.group, "warm": a.warm, "tail": a.tail, "separation": a.separation, "total_s": sum(r["ms"] for r in rows) / 1e3,
                      "wall_s": time.perf_counter() - t0, "increments": rows}), flush=True)


if __name__ == "__main__":
    main()

"""pd_reduce_sum_noalloc (the fixed-tree sum with its partials in the
library's device array, for callers that cannot allocate) against
pd_reduce_sum on the same data: the same bits, on two calls in a row, the
second over fewer values, and for two lane counts.  The test allocates its
own input and output with the HIP runtime; the entry point allocates nothing."""
import ctypes as C

import numpy as np
import pytest

from pd_gpu1_amd import _lib, scenarios

pytestmark = pytest.mark.gpu

_hip = None


def hip():
    global _hip
    if _hip is None:
        _hip = C.CDLL("libamdhip64.so")
        _hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        _hip.hipFree.argtypes = [C.c_void_p]
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hip.hipDeviceSynchronize.argtypes = []
    return _hip


H2D, D2H = 1, 2


class DevBuf:
    def __init__(self, nbytes):
        self.p = C.c_void_p()
        assert hip().hipMalloc(C.byref(self.p), max(nbytes, 8)) == 0

    def free(self):
        hip().hipFree(self.p)


def noalloc_sum(values, lanes):
    src = DevBuf(8 * len(values))
    out = DevBuf(8)
    try:
        if len(values):
            assert hip().hipMemcpy(src.p, values.ctypes.data, 8 * len(values), H2D) == 0
        rc = _lib.lib().pd_reduce_sum_noalloc(0, src.p, len(values), lanes, out.p, None)
        assert rc == _lib.PD_OK
        assert hip().hipDeviceSynchronize() == 0
        res = np.zeros(1)
        assert hip().hipMemcpy(res.ctypes.data, out.p, 8, D2H) == 0
        return res[0]
    finally:
        src.free()
        out.free()


@pytest.mark.parametrize("lanes", [4096, 256])
def test_noalloc_sum_matches_reduce_sum(lanes):
    rng = np.random.default_rng(lanes)
    for scale in (0.8, 0.3):  # the second call sums fewer values
        sc = scenarios.c0_plate_pmb(scale=scale)
        sc.p.reduce_lanes = lanes
        u = rng.standard_normal((2, len(sc.x))) * 1e-6
        with sc.make() as prob:
            prob.set("U", u)
            want = prob.reduce_sum("U", 1)
        got = noalloc_sum(np.ascontiguousarray(u[1]), lanes)
        print(f"lanes {lanes} n {len(sc.x)}: {got!r} {want!r}", flush=True)
        assert got == want


def test_noalloc_sum_edges():
    assert noalloc_sum(np.zeros(0), 4096) == 0.0
    assert noalloc_sum(np.array([1.5, -2.25, 3.0]), 64) == 2.25
    L = _lib.lib()
    out = DevBuf(8)
    try:
        assert L.pd_reduce_sum_noalloc(0, out.p, 1, 100, out.p, None) == _lib.PD_E_INVALID
        assert L.pd_reduce_sum_noalloc(0, out.p, 1, 1 << 21, out.p, None) == _lib.PD_E_CAPACITY
        assert L.pd_reduce_sum_noalloc(0, out.p, -1, 64, out.p, None) == _lib.PD_E_INVALID
    finally:
        out.free()

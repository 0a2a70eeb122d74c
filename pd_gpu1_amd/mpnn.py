"""Python handle of the message-passing warm-start network (pd_mpnn_* in
include/pd_gpu.h).  All compute runs in libpd_gpu1_amd.so on the GPU."""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import lib

N_FEAT, E_FEAT, OUT = 16, 3, 2


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


class Mpnn:
    def __init__(self, seed=0x5EED0004, device=0):
        h = _lib.H()
        rc = lib().pd_mpnn_create(device, C.c_uint64(seed), C.byref(h))
        if rc != _lib.PD_OK:
            raise _lib.PdError(rc, "pd_mpnn_create failed")
        self._h = h
        self.N = 0

    def close(self):
        if getattr(self, "_h", None):
            lib().pd_mpnn_destroy(self._h)
            self._h = None

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
            raise _lib.PdError(rc, lib().pd_mpnn_last_error(self._h).decode())

    @property
    def num_params(self):
        n = C.c_int64()
        self._c(lib().pd_mpnn_num_params(self._h, C.byref(n)))
        return n.value

    def get_params(self):
        out = np.empty(self.num_params, dtype=np.float32)
        self._c(lib().pd_mpnn_get_params(self._h, out.ctypes.data_as(_lib.F32)))
        return out

    def set_params(self, theta):
        th = _f32(theta)
        assert th.size == self.num_params
        self._c(lib().pd_mpnn_set_params(self._h, th.ctypes.data_as(_lib.F32)))

    def set_graph(self, row_ptr, col):
        rp = np.ascontiguousarray(row_ptr, dtype=np.int32)
        cl = np.ascontiguousarray(col, dtype=np.int32)
        self._c(lib().pd_mpnn_set_graph(self._h, len(rp) - 1, rp.ctypes.data_as(_lib.I32), len(cl),
                                        cl.ctypes.data_as(_lib.I32)))
        self.N = len(rp) - 1

    def set_features(self, x, ef, target=None, weight=None):
        x, ef = _f32(x), _f32(ef)
        assert x.shape == (self.N, N_FEAT) and ef.shape[1:] == (E_FEAT,)
        t = w = None
        if target is not None:
            t = _f32(target)
            w = _f32(np.ones(self.N) if weight is None else weight)
        ptr = lambda a: a.ctypes.data_as(_lib.F32) if a is not None else None
        self._c(lib().pd_mpnn_set_features(self._h, ptr(x), ptr(ef), ptr(t), ptr(w)))

    def forward(self):
        y = np.empty((self.N, OUT), dtype=np.float32)
        self._c(lib().pd_mpnn_forward(self._h, y.ctypes.data_as(_lib.F32)))
        return y

    def gradient(self):
        g = np.empty(self.num_params, dtype=np.float32)
        loss = C.c_float()
        self._c(lib().pd_mpnn_gradient(self._h, g.ctypes.data_as(_lib.F32), C.byref(loss)))
        return loss.value, g

    def train_step(self, lr, want_loss=True):
        loss = C.c_float()
        self._c(lib().pd_mpnn_train_step(self._h, C.c_float(lr), C.byref(loss) if want_loss else None))
        return loss.value if want_loss else None

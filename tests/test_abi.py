"""C-ABI checks that need no GPU: the library loads, exports every symbol
include/pd_gpu.h declares, and its host-side parameter derivation and
coordinate generators agree bit for bit with the oracle."""
import os
import re

import numpy as np

import pd_gpu1_amd as pd
from pd_gpu1_amd import _lib
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    hdr = open(os.path.join(ROOT, "include", "pd_gpu.h")).read()
    declared = set(re.findall(r"^(?:int|const char \*)\s*(pd_\w+)\s*\(", hdr, flags=re.M))
    assert declared == set(_lib.EXPORTS)
    L = _lib.lib()
    for name in declared:
        assert hasattr(L, name), name


def test_params_match_oracle_bitwise():
    cases = [
        (2, 1, 1e-8, 0.5e-3, 0.5e-3, 0.05, 0.05, 0.0, 192e9, 1 / 3, 8000, 0.004, 3.015),
        (3, 1, 8.7e-8, 1e-3, 1e-3, 0.2, 0.1, 0.009, 190e9, 0.25, 8000, 0.0057, 3.0),
        (3, 2, 2e-8, 0.5e-3, 0.0, 0.1, 0.1, 0.05, 70e9, 0.33, 2700, 0.002, 3.015),
        (2, 2, 1.3367e-8, 1e-4, 1e-4, 0.05, 0.05, 0.0, 192e9, 0.3, 7850, 0.04472, 3.015),
        (2, 1, 1.0, 0.01, 0.01, 1.0, 0.5, 0.0, 200e9, 1 / 3, 7850, 1.0, 3.015),
    ]
    for c in cases:
        dim, model, dt, dx, thick, L, H, W, E, nu, rho, s0, df = c
        p, nxyz, n = pd.params(dim, model, dt, dx, thick=thick, length=L, height=H, width=W, emod=E,
                               pratio=nu, dens=rho, s0=s0, delta_factor=df)
        o = O.cal_parameter(dim, model, dt, dx, thick, L, H, W, E, nu, rho, s0, df)
        assert (nxyz[0], nxyz[1]) == (o.nx, o.ny) and (dim == 2 or nxyz[2] == o.nz)
        for k in ("delta", "vol", "bc", "mass", "K", "G"):
            assert getattr(p, k) == getattr(o, k), (c, k)


def test_generators_match_oracle_bitwise():
    x, y = pd.gen_plate_strips(100, 100, 0.5e-3)
    xo, yo = O.coord_plate_strips(100, 100, 0.5e-3)
    assert np.array_equal(x, xo) and np.array_equal(y, yo)
    x, y = pd.gen_plate(51, 25, 0.02)
    xo, yo = O.coord_plate(51, 25, 0.02)
    assert np.array_equal(x, xo) and np.array_equal(y, yo)
    k = pd.gen_kw(201, 101, 9, 1e-3, 0.025)
    ko = O.coord_kw(201, 101, 9, 1e-3, 0.025)
    assert all(np.array_equal(a, b) for a, b in zip(k, ko))
    j = pd.gen_block_jitter(20, 18, 10, 0.5e-3, 0.01, 0x5EED0002)
    jo = O.coord_block_jitter(20, 18, 10, 0.5e-3, 0.01, 0x5EED0002)
    assert all(np.array_equal(a, b) for a, b in zip(j, jo))


def test_invalid_params_rejected_without_gpu():
    p, _, _ = pd.params(2, 1, 1e-8, 0.5e-3, thick=0.5e-3)
    p.reduce_lanes = 100  # not a power of two
    h = _lib.C.c_void_p() if hasattr(_lib, "C") else None
    import ctypes
    h = ctypes.c_void_p()
    assert _lib.lib().pd_create(ctypes.byref(p), 0, ctypes.byref(h)) == _lib.PD_E_INVALID

"""MI355X-native peridynamic time-step engine (gfx950 HIP kernels behind a C-ABI).

Drop-in for the hot path of XiaoMmingW/PD_Gpu1: neighbour CSR, PMB/LPS force,
Verlet/ADR integration and the ADR damping reduction.  See ``engine.Problem``
for the stage interface and ``include/pd_gpu.h`` for the C-ABI.
"""
from ._lib import PdError, PdParams, LIB_PATH  # noqa: F401
from .engine import Problem, params, gen_plate, gen_plate_strips, gen_kw, gen_block_jitter  # noqa: F401

__all__ = ["Problem", "params", "gen_plate", "gen_plate_strips", "gen_kw", "gen_block_jitter",
           "PdError", "PdParams", "LIB_PATH"]

"""ctypes binding of libpd_gpu1_amd.so (the C-ABI in include/pd_gpu.h).

The shared library is built in-tree by ``pd_gpu1_amd.build``.  Loading fails
loudly when it is missing: there is no CPU fallback for any entry point.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpd_gpu1_amd.so")

PD_OK, PD_E_INVALID, PD_E_HIP, PD_E_CAPACITY, PD_E_NONFINITE, PD_E_NOT_CONVERGED = range(6)
PD_MODEL_PMB, PD_MODEL_LPS = 1, 2
PD_PREC_FP64, PD_PREC_MIXED = 0, 1
FIELDS = dict(X0=0, U=1, V=2, A=3, F=4, B=5, VH=6, VHO=7, FO=8, DMG=9, THETA=10, M=11,
              IDIST=20, FAC=21, SCR=22, W=23, E=24)
PD_INIT_DYNAMIC, PD_INIT_STATIC = 0, 1
PD_LAYOUT_PLATE, PD_LAYOUT_KW = 0, 1

# every symbol include/pd_gpu.h declares
EXPORTS = (
    "pd_params_init", "pd_gen_plate", "pd_gen_plate_strips", "pd_gen_kw", "pd_gen_block_jitter",
    "pd_create", "pd_destroy", "pd_last_error", "pd_get_params", "pd_set_params", "pd_set_positions",
    "pd_build_neighbors", "pd_num_bonds", "pd_get_csr", "pd_vol_corr", "pd_surface_correct",
    "pd_init_state", "pd_set_crack_2d", "pd_set_notches_kw", "pd_lps_weights", "pd_force",
    "pd_verlet_half1", "pd_verlet_half2", "pd_run_verlet", "pd_set_body_force_tensile",
    "pd_adr_iterate", "pd_adr_solve", "pd_strip_displace", "pd_reduce_sum", "pd_reduce_sum_noalloc", "pd_damage",
    "pd_get_field", "pd_set_field", "pd_get_fail_bits", "pd_set_fail_bits", "pd_cylinder_init",
    "pd_cylinder_get", "pd_contact_step", "pd_save_text", "pd_checkpoint_write", "pd_checkpoint_read",
    "pd_last_force_ms", "pd_time_force_passes", "pd_set_group", "pd_set_pass_group", "pd_stream_copy", "pd_set_tile", "pd_set_lds", "pd_set_lds_block", "pd_set_lds_pass_block", "pd_lds_info", "pd_set_coef_mode", "pd_coef_info", "pd_set_fp64_group", "pd_sync",
    "pd_batch_create", "pd_batch_destroy", "pd_batch_last_error", "pd_batch_info", "pd_batch_spec_info",
    "pd_batch_increment", "pd_batch_set_group", "pd_batch_set_parts", "pd_batch_get_parts", "pd_batch_set_barrier_timeout", "pd_batch_choose_parts", "pd_batch_set_warm", "pd_batch_set_coef_mode", "pd_batch_coef_info", "pd_batch_set_retire", "pd_batch_set_tail", "pd_batch_get_tail_parts", "pd_batch_set_separation", "pd_batch_get_retired", "pd_batch_get_field", "pd_batch_set_field", "pd_batch_get_csr",
    "pd_mpnn_create", "pd_mpnn_destroy", "pd_mpnn_last_error", "pd_mpnn_num_params", "pd_mpnn_get_params",
    "pd_mpnn_set_params", "pd_mpnn_set_graph", "pd_mpnn_set_features", "pd_mpnn_forward",
    "pd_mpnn_gradient", "pd_mpnn_train_step",
)


class PdParams(C.Structure):
    _fields_ = [(n, C.c_int32) for n in ("dim", "model", "precision", "max_neighbors",
                                          "theta_skip_broken", "reduce_lanes", "n_top", "n_bot")] + [
        (n, C.c_double) for n in ("dx", "delta", "vol", "bc", "mass", "K", "G", "emod", "pratio",
                                  "dens", "s0", "dt", "break_y_min", "strip_velocity", "near_window")]


class PdCyl(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("rad", "mass", "height", "v", "x", "y", "z")]


class PdSpecimen(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("crack_len", C.c_double), ("crack_angle", C.c_double)]


class PdError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"pd error {code}: {msg}")
        self.code = code


_lib = None
D = C.POINTER(C.c_double)
I32 = C.POINTER(C.c_int32)
U32 = C.POINTER(C.c_uint32)
F32 = C.POINTER(C.c_float)
H = C.c_void_p


def lib():
    """Load libpd_gpu1_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: run pd_gpu1_amd.build.build() (hipcc, gfx950)")
        L = C.CDLL(LIB_PATH)
        sig = {
            "pd_params_init": [C.POINTER(PdParams), C.c_int, C.c_int] + [C.c_double] * 11 + [C.c_int, I32, I32],
            "pd_gen_plate": [C.c_int, C.c_int, C.c_double, D, D],
            "pd_gen_plate_strips": [C.c_int, C.c_int, C.c_double, D, D],
            "pd_gen_kw": [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, D, D, D],
            "pd_gen_block_jitter": [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_uint64, D, D, D],
            "pd_create": [C.POINTER(PdParams), C.c_int, C.POINTER(H)],
            "pd_destroy": [H],
            "pd_last_error": [H],
            "pd_get_params": [H, C.POINTER(PdParams)],
            "pd_set_params": [H, C.POINTER(PdParams)],
            "pd_set_positions": [H, D, D, D, C.c_int, C.c_int],
            "pd_build_neighbors": [H],
            "pd_num_bonds": [H, C.POINTER(C.c_int64)],
            "pd_get_csr": [H, I32, I32],
            "pd_vol_corr": [H],
            "pd_surface_correct": [H],
            "pd_init_state": [H, C.c_int],
            "pd_set_crack_2d": [H, C.c_double, C.c_double, C.c_double, C.c_double],
            "pd_set_notches_kw": [H, C.c_double],
            "pd_lps_weights": [H],
            "pd_force": [H],
            "pd_verlet_half1": [H],
            "pd_verlet_half2": [H],
            "pd_run_verlet": [H, C.c_int],
            "pd_set_body_force_tensile": [H, C.c_int, C.c_int],
            "pd_adr_iterate": [H, C.c_int, D, D],
            "pd_adr_solve": [H, C.c_int, C.c_int, C.c_double, C.POINTER(C.c_int), D],
            "pd_strip_displace": [H, C.c_double, C.c_double],
            "pd_reduce_sum": [H, C.c_int, C.c_int, D],
            "pd_reduce_sum_noalloc": [C.c_int, C.c_void_p, C.c_int64, C.c_int, C.c_void_p, C.c_void_p],
            "pd_damage": [H],
            "pd_get_field": [H, C.c_int, D],
            "pd_set_field": [H, C.c_int, D],
            "pd_get_fail_bits": [H, U32],
            "pd_set_fail_bits": [H, U32],
            "pd_cylinder_init": [H, C.POINTER(PdCyl)],
            "pd_cylinder_get": [H, D],
            "pd_contact_step": [H],
            "pd_save_text": [H, C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double],
            "pd_checkpoint_write": [H, C.c_char_p, C.c_int64],
            "pd_checkpoint_read": [H, C.c_char_p, C.POINTER(C.c_int64)],
            "pd_last_force_ms": [H, D],
            "pd_time_force_passes": [H, C.c_int, D],
            "pd_set_group": [H, C.c_int],
            "pd_set_tile": [H, C.c_int],
            "pd_set_lds": [H, C.c_int],
            "pd_set_lds_block": [H, C.c_int],
            "pd_set_lds_pass_block": [H, C.c_int, C.c_int],
            "pd_set_coef_mode": [H, C.c_int],
            "pd_coef_info": [H, C.POINTER(C.c_int)],
            "pd_set_fp64_group": [H, C.c_int],
            "pd_lds_info": [H, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)],
            "pd_set_pass_group": [H, C.c_int, C.c_int],
            "pd_stream_copy": [C.c_int, C.c_int64, C.c_int, D],
            "pd_sync": [H],
            "pd_batch_create": [C.POINTER(PdParams), C.POINTER(PdSpecimen), C.c_int, C.c_int, C.POINTER(H)],
            "pd_batch_destroy": [H],
            "pd_batch_last_error": [H],
            "pd_batch_info": [H, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_int)],
            "pd_batch_spec_info": [H, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int64)],
            "pd_batch_increment": [H, C.c_double, C.c_double, C.c_int, C.c_int, C.c_int, I32, D, D],
            "pd_batch_set_group": [H, C.c_int],
            "pd_batch_set_parts": [H, C.c_int],
            "pd_batch_get_parts": [H, C.POINTER(C.c_int)],
            "pd_batch_set_barrier_timeout": [H, C.c_double],
            "pd_batch_choose_parts": [C.POINTER(C.c_int64), I32, I32, C.c_int, C.c_int, C.c_int],
            "pd_batch_set_warm": [H, C.c_int],
            "pd_batch_set_coef_mode": [H, C.c_int],
            "pd_batch_coef_info": [H, C.POINTER(C.c_int)],
            "pd_batch_set_retire": [H, C.c_int],
            "pd_batch_set_tail": [H, C.c_int],
            "pd_batch_get_tail_parts": [H, C.POINTER(C.c_int)],
            "pd_batch_set_separation": [H, C.c_double],
            "pd_batch_get_retired": [H, I32],
            "pd_batch_get_field": [H, C.c_int, C.c_int, D],
            "pd_batch_set_field": [H, C.c_int, C.c_int, D],
            "pd_batch_get_csr": [H, C.c_int, I32, I32, U32],
            "pd_mpnn_create": [C.c_int, C.c_uint64, C.POINTER(H)],
            "pd_mpnn_destroy": [H],
            "pd_mpnn_last_error": [H],
            "pd_mpnn_num_params": [H, C.POINTER(C.c_int64)],
            "pd_mpnn_get_params": [H, F32],
            "pd_mpnn_set_params": [H, F32],
            "pd_mpnn_set_graph": [H, C.c_int, I32, C.c_int64, I32],
            "pd_mpnn_set_features": [H, F32, F32, F32, F32],
            "pd_mpnn_forward": [H, F32],
            "pd_mpnn_gradient": [H, F32, F32],
            "pd_mpnn_train_step": [H, C.c_float, F32],
        }
        for name, args in sig.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = C.c_int
        L.pd_last_error.restype = C.c_char_p
        L.pd_batch_last_error.restype = C.c_char_p
        L.pd_mpnn_last_error.restype = C.c_char_p
        _lib = L
    return _lib


def dptr(a):
    return None if a is None else a.ctypes.data_as(D)


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)

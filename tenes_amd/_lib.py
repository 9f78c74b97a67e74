"""ctypes binding of include/tenes_amd.h (libtenes_amd.so).

Loading fails loudly: there is no CPU fallback for any entry point.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "lib" / "libtenes_amd.so"

c_int_p = C.POINTER(C.c_int)
c_i64_p = C.POINTER(C.c_int64)
c_dbl_p = C.POINTER(C.c_double)
c_void_pp = C.POINTER(C.c_void_p)


class PepsParameters(C.Structure):
    _fields_ = [
        ("CHI", C.c_int),
        ("Inverse_projector_cut", C.c_double),
        ("CTM_Convergence_Epsilon", C.c_double),
        ("Max_CTM_Iteration", C.c_int),
        ("CTM_Projector_corner", C.c_int),
        ("Use_RSVD", C.c_int),
        ("RSVD_Oversampling_factor", C.c_double),
        ("print_level", C.c_int),
        ("svd_power_iters", C.c_int),
        ("rsvd_power_iters", C.c_int),
        ("svd_warm_start", C.c_int),
        ("svd_warm_power_iters", C.c_int),
        ("contraction_fp32", C.c_int),
        ("projector_cut_on_device", C.c_int),
        ("graph_sweeps", C.c_int),
    ]


class FuParameters(C.Structure):
    _fields_ = [
        ("Inverse_Env_cut", C.c_double),
        ("Full_Inverse_precision", C.c_double),
        ("Full_Convergence_Epsilon", C.c_double),
        ("Full_max_iteration", C.c_int),
        ("Full_Gauge_Fix", C.c_int),
        ("Full_Use_FastFullUpdate", C.c_int),
    ]


class TmParameters(C.Structure):
    _fields_ = [
        ("num_eigvals", C.c_int),
        ("maxdim_dense_eigensolver", C.c_int),
        ("arnoldi_maxdim", C.c_int),
        ("arnoldi_restartdim", C.c_int),
        ("arnoldi_maxiter", C.c_int),
        ("arnoldi_rtol", C.c_double),
    ]


# name -> (restype, argtypes)
_SIGS = {
    "tenes_abi_version": (C.c_int, []),
    "tenes_ctx_set_autotune": (C.c_int, [C.c_void_p, C.c_int]),
    "tenes_ctx_set_stream_k": (C.c_int, [C.c_void_p, C.c_int]),
    "tenes_ctx_create": (C.c_int, [C.c_int, C.c_void_p, c_void_pp]),
    "tenes_ctx_destroy": (C.c_int, [C.c_void_p]),
    "tenes_ctx_last_error": (C.c_char_p, [C.c_void_p]),
    "tenes_ctx_synchronize": (C.c_int, [C.c_void_p]),
    "tenes_ctx_set_seed": (C.c_int, [C.c_void_p, C.c_uint64]),
    "tenes_ctx_stats": (C.c_int, [C.c_void_p, c_dbl_p, c_i64_p]),
    "tenes_ctx_reset_stats": (C.c_int, [C.c_void_p]),
    "tenes_ctx_svd_stats": (C.c_int, [C.c_void_p, c_i64_p, c_i64_p]),
    "tenes_ctx_set_max_lanes": (C.c_int, [C.c_void_p, C.c_int]),
    "tenes_ctx_set_profiling": (C.c_int, [C.c_void_p, C.c_int]),
    "tenes_ctx_kernel_report": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int]),
    "tenes_ctx_reset_kernel_report": (C.c_int, [C.c_void_p]),
    "tenes_tensor_from_host": (C.c_int, [C.c_void_p, C.c_int, c_i64_p, c_dbl_p, c_void_pp]),
    "tenes_tensor_from_device": (C.c_int, [C.c_void_p, C.c_int, c_i64_p, C.c_void_p, c_void_pp]),
    "tenes_tensor_zeros": (C.c_int, [C.c_void_p, C.c_int, c_i64_p, c_void_pp]),
    "tenes_tensor_rank": (C.c_int, [C.c_void_p, c_int_p]),
    "tenes_tensor_shape": (C.c_int, [C.c_void_p, c_i64_p]),
    "tenes_tensor_to_host": (C.c_int, [C.c_void_p, c_dbl_p]),
    "tenes_tensor_device_ptr": (C.c_int, [C.c_void_p, c_void_pp]),
    "tenes_tensor_destroy": (C.c_int, [C.c_void_p]),
    "tenes_tensordot": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, c_int_p, c_int_p, c_int_p, c_void_pp]),
    "tenes_transpose": (C.c_int, [C.c_void_p, c_int_p, c_void_pp]),
    "tenes_svd_top": (C.c_int, [C.c_void_p, C.c_int, C.c_int, c_void_pp, c_void_pp, c_dbl_p, c_int_p]),
    "tenes_peps_parameters_default": (None, [C.POINTER(PepsParameters)]),
    "tenes_calc_projector_left_block": (
        C.c_int,
        [C.c_void_p] * 8 + [C.POINTER(PepsParameters), c_void_pp, c_void_pp],
    ),
    "tenes_calc_projector_updown_blocks": (
        C.c_int, [C.POINTER(C.c_void_p), C.POINTER(PepsParameters), c_void_pp, c_void_pp]),
    "tenes_calc_next_ctm": (C.c_int, [C.c_void_p] * 6 + [c_void_pp, c_void_pp]),
    "tenes_calc_next_et": (C.c_int, [C.c_void_p] * 4 + [c_void_pp]),
    "tenes_env_create": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, c_void_pp]),
    "tenes_env_destroy": (C.c_int, [C.c_void_p]),
    "tenes_env_set": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_void_p]),
    "tenes_env_get": (C.c_int, [C.c_void_p, C.c_int, C.c_int, c_void_pp]),
    "tenes_calc_ctm_environment": (
        C.c_int, [C.c_void_p, C.POINTER(PepsParameters), C.c_int, c_int_p, c_dbl_p]),
    "tenes_calc_ctm_environment_density": (
        C.c_int, [C.c_void_p, C.POINTER(PepsParameters), C.c_int, c_int_p, c_dbl_p]),
    "tenes_env_move": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.POINTER(PepsParameters)]),
    "tenes_env_save_corners": (C.c_int, [C.c_void_p]),
    "tenes_env_check_convergence": (
        C.c_int, [C.c_void_p, C.POINTER(PepsParameters), c_int_p, c_dbl_p]),
    "tenes_env_corner_spectrum": (C.c_int, [C.c_void_p, C.c_int, C.c_int, c_dbl_p, c_int_p]),
    "tenes_env_contract_one_site": (C.c_int, [C.c_void_p, C.c_int, c_dbl_p, c_dbl_p]),
    "tenes_env_contract_two_sites_op12": (
        C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, c_dbl_p, c_dbl_p]),
    "tenes_ctx_set_fp32_contractions": (C.c_int, [C.c_void_p, C.c_int]),
    "tenes_ctx_set_fused_site": (C.c_int, [C.c_void_p, C.c_int]),
    "tenes_ctx_lane_state": (C.c_int, [C.c_void_p, c_int_p, c_int_p]),
    "tenes_ctx_set_max_moves_in_flight": (C.c_int, [C.c_void_p, C.c_int]),
    "tenes_ctx_graph_stats": (C.c_int, [C.c_void_p, c_i64_p, c_i64_p]),
    "tenes_svd_top_noalloc": (
        C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                  c_int_p]),
    "tenes_batch_calc_ctm_environment": (
        C.c_int, [c_void_pp, C.c_int, C.POINTER(PepsParameters), C.c_int, C.c_int, c_int_p, c_dbl_p]),
    "tenes_fu_parameters_default": (None, [C.POINTER(FuParameters)]),
    "tenes_full_update_bond": (
        C.c_int, [c_void_pp, c_dbl_p, C.c_int, C.POINTER(FuParameters), c_void_pp, c_void_pp]),
    "tenes_env_full_update": (
        C.c_int, [C.c_void_p, C.c_int, c_int_p, c_int_p, c_int_p, c_dbl_p, C.c_int, C.POINTER(PepsParameters),
                  C.POINTER(FuParameters)]),
    "tenes_env_contract_multisite": (
        C.c_int, [C.c_void_p, C.c_int, C.c_int, c_int_p, c_int_p, c_dbl_p, c_dbl_p]),
    "tenes_env_correlation": (
        C.c_int, [C.c_void_p, C.c_int, c_dbl_p, C.c_int, c_dbl_p, C.c_int, C.c_int, c_dbl_p]),
    "tenes_env_bond_energy": (
        C.c_int, [C.c_void_p, C.c_int, c_int_p, c_int_p, c_int_p, c_dbl_p, c_dbl_p, c_void_pp]),
    "tenes_tensor_to_device": (C.c_int, [C.c_void_p, C.c_void_p]),
    "tenes_env_contract_cluster": (
        C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, c_dbl_p, c_dbl_p, c_dbl_p]),
    "tenes_create_environment_two_sites": (C.c_int, [C.POINTER(C.c_void_p), c_void_pp]),
    "tenes_simple_update_bond": (
        C.c_int, [C.c_void_p, C.c_void_p, c_dbl_p, c_dbl_p, C.c_void_p, C.c_int, C.c_double,
                  c_void_pp, c_void_pp, c_dbl_p]),
    "tenes_env_simple_update": (
        C.c_int, [C.c_void_p, C.c_int, c_int_p, c_int_p, C.POINTER(C.c_void_p), C.c_int, C.c_double]),
    "tenes_env_get_lambda": (C.c_int, [C.c_void_p, C.c_int, C.c_int, c_dbl_p, c_int_p]),
    "tenes_env_save_tensors": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int]),
    "tenes_env_load_tensors": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int]),
    "tenes_tm_parameters_default": (None, [C.POINTER(TmParameters)]),
    "tenes_env_transfer_matrix_eigenvalues": (
        C.c_int, [C.c_void_p, C.c_int, C.c_int, C.POINTER(TmParameters), c_dbl_p, c_int_p]),
}

_lib = None


def load() -> C.CDLL:
    """Load libtenes_amd.so (built by tenes_amd.build); raise if missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(
            f"{LIB_PATH} not found: build it first (python -m tenes_amd.build or "
            "__graft_entry__.build()); there is no CPU fallback")
    lib = C.CDLL(str(LIB_PATH))
    for name, (res, args) in _SIGS.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def exported_symbols() -> list[str]:
    return sorted(_SIGS)


def dptr(a: np.ndarray):
    return a.ctypes.data_as(c_dbl_p)


def i64ptr(a: np.ndarray):
    return a.ctypes.data_as(c_i64_p)


def iptr(a: np.ndarray):
    return a.ctypes.data_as(c_int_p)

"""Loader of the in-tree gfx950 library (ctypes, C ABI of include/linalg_rocm_native.h).

There is no fallback: if the library is missing, importing the ops raises.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liblinalg_rocm_native.so")

_i64 = ctypes.c_int64
_vp = ctypes.c_void_p
_sz = ctypes.c_size_t
_int = ctypes.c_int

# name -> argtypes (restype int unless listed in _RESTYPES)
SIGNATURES = {
    "lrn_status_string": [_int],
    "lrn_version": [],
    "lrn_lu_decomposition_workspace_size_f64": [_i64, ctypes.POINTER(_sz)],
    "lrn_lu_decomposition_workspace_size_f32": [_i64, ctypes.POINTER(_sz)],
    "lrn_lu_decomposition_f64": [_vp, _i64, _vp, _vp, _sz, ctypes.POINTER(_int), _vp],
    "lrn_lu_decomposition_f32": [_vp, _i64, _vp, _vp, _sz, ctypes.POINTER(_int), _vp],
    "lrn_linear_solver_workspace_size_f64": [_i64, _int, ctypes.POINTER(_sz)],
    "lrn_linear_solver_workspace_size_f32": [_i64, _int, ctypes.POINTER(_sz)],
    "lrn_linear_solver_f64": [_vp, _vp, _i64, _vp, _int, _vp, _sz, _vp],
    "lrn_linear_solver_f32": [_vp, _vp, _i64, _vp, _int, _vp, _sz, _vp],
    "lrn_linear_solver_noalloc_f64": [_vp, _vp, _i64, _vp, _int, _vp],
    "lrn_linear_solver_noalloc_max_order": [_int, ctypes.POINTER(_i64)],
    "lrn_triangular_solver_workspace_size_f64": [_i64, _i64, ctypes.POINTER(_sz)],
    "lrn_triangular_solver_workspace_size_f32": [_i64, _i64, ctypes.POINTER(_sz)],
    "lrn_triangular_solver_f64": [_int, _int, _int, _i64, _i64, _vp, _i64, _vp, _i64, _vp, _i64, _vp, _sz, _vp],
    "lrn_triangular_solver_f32": [_int, _int, _int, _i64, _i64, _vp, _i64, _vp, _i64, _vp, _i64, _vp, _sz, _vp],
    "lrn_batched_matrix_multiplication_f64": [_i64, _i64, _vp, _vp, _vp, _vp],
    "lrn_batched_matrix_multiplication_f32": [_i64, _i64, _vp, _vp, _vp, _vp],
    "lrn_cholesky_decomposition_workspace_size_f64": [_i64, ctypes.POINTER(_sz)],
    "lrn_cholesky_decomposition_workspace_size_f32": [_i64, ctypes.POINTER(_sz)],
    "lrn_cholesky_decomposition_f64": [_vp, _i64, _int, _vp, _sz, ctypes.POINTER(_int), _vp],
    "lrn_cholesky_decomposition_f32": [_vp, _i64, _int, _vp, _sz, ctypes.POINTER(_int), _vp],
    "lrn_cholesky_decomposition_async_f64": [_vp, _i64, _int, _vp, _sz, _vp, _vp],
    "lrn_cholesky_decomposition_async_f32": [_vp, _i64, _int, _vp, _sz, _vp, _vp],
    "lrn_debug_set_spin_limit": [_i64],
    "lrn_determinant_workspace_size_f64": [_i64, _int, ctypes.POINTER(_sz)],
    "lrn_determinant_workspace_size_f32": [_i64, _int, ctypes.POINTER(_sz)],
    "lrn_determinant_f64": [_vp, _vp, _i64, _int, _vp, _sz, ctypes.POINTER(ctypes.c_double), _vp],
    "lrn_determinant_f32": [_vp, _vp, _i64, _int, _vp, _sz, ctypes.POINTER(ctypes.c_float), _vp],
    "lrn_inversion_workspace_size_f64": [_i64, _int, ctypes.POINTER(_sz)],
    "lrn_inversion_workspace_size_f32": [_i64, _int, ctypes.POINTER(_sz)],
    "lrn_inversion_f64": [_vp, _vp, _i64, _int, _vp, _sz, _vp],
    "lrn_inversion_f32": [_vp, _vp, _i64, _int, _vp, _sz, _vp],
    "lrn_linear_solver_multi_workspace_size_f64": [_i64, _i64, _int, ctypes.POINTER(_sz)],
    "lrn_linear_solver_multi_workspace_size_f32": [_i64, _i64, _int, ctypes.POINTER(_sz)],
    "lrn_linear_solver_multi_f64": [_vp, _vp, _i64, _vp, _i64, _i64, _int, _vp, _sz, _vp],
    "lrn_linear_solver_multi_f32": [_vp, _vp, _i64, _vp, _i64, _i64, _int, _vp, _sz, _vp],
    "lrn_getrf_async_f64": [_vp, _i64, _vp, _vp, _vp, _sz, _vp],
    "lrn_getrf_async_f32": [_vp, _i64, _vp, _vp, _vp, _sz, _vp],
    "lrn_getrf_tournament_workspace_size_f32": [_i64, ctypes.POINTER(_sz)],
    "lrn_getrf_tournament_f32": [_vp, _i64, _vp, _vp, _vp, _sz, _vp],
    "lrn_set_mixed_panel": [_int],
    "lrn_set_refine_solver": [_int],
    "lrn_set_panel_kernel": [_int],
    "lrn_getrs_f32f64_workspace_size": [_i64, ctypes.POINTER(_sz)],
    "lrn_getrs_f32f64": [_vp, _i64, _vp, _vp, _vp, _sz, _vp],
    "lrn_debug_gemm_sub": [_int, _vp, _i64, _vp, _i64, _vp, _i64, _i64, _i64, _i64, _vp],
    "lrn_debug_trsm": [_int, _vp, _i64, _i64, _int, _i64, _i64, _vp],
    "lrn_debug_panel": [_int, _vp, _i64, _i64, _int, _vp, _vp, _sz, _vp],
    "lrn_set_tuning": [_int, _int, _int],
    "lrn_set_lookahead": [_int],
    "lrn_set_block_update": [_int],
    "lrn_set_refinement": [_int],
    "lrn_set_multi_rhs_solver": [_int],
    "lrn_set_multi_rhs_kernel": [_int],
    "lrn_set_mixed_update": [_int],
    "lrn_gesv_mixed_batched_f64": [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _i64, _vp, ctypes.c_double, _int, _int, _vp, _sz,
                                   _vp, _vp],
    "lrn_batched_gesv_workspace_size": [_i64, ctypes.POINTER(_sz)],
    "lrn_batched_gesv_f64": [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i64, _vp, _vp, _sz, _vp],
    "lrn_batched_gesv_f32": [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i64, _vp, _vp, _sz, _vp],
    "lrn_set_batched_grid": [_int],
    "lrn_trace_start": [ctypes.c_uint, _i64],
    "lrn_trace_stop": [],
    "lrn_trace_read": [_int, ctypes.POINTER(_i64), ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)],
    "lrn_gesv_mixed_workspace_size_f64": [_i64, _int, _int, ctypes.POINTER(_sz)],
    "lrn_gesv_mixed_f64": [_vp, _i64, _vp, _vp, ctypes.c_double, _int, _int, _vp, _sz, _vp, _vp],
    "lrn_get_tuning": [ctypes.POINTER(_int), ctypes.POINTER(_int), ctypes.POINTER(_int)],
}
_RESTYPES = {"lrn_status_string": ctypes.c_char_p, "lrn_version": ctypes.c_char_p}

_lib = None


def load(apply_tuning: bool = True):
    """Load (once) and return the library; raises if it has not been built. On the first load
    the block sizes of ``tuning.json`` (written by ``python -m linearalgebra_amd.autotune``)
    are applied, if that file exists and ``apply_tuning`` is set."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python -m linearalgebra_amd.build` "
                "(there is no CPU fallback)")
        lib = ctypes.CDLL(LIB_PATH)
        for name, args in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = _RESTYPES.get(name, _int)
        _lib = lib
        if apply_tuning:
            from . import autotune
            autotune.apply(lib)
    return _lib


def status_text(code: int) -> str:
    return load().lrn_status_string(int(code)).decode()

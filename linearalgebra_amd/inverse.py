"""determinant<T, ROCM_NATIVE> and inversion<T, ROCM_NATIVE>: direct consumers of the LU factors.

Mirrors include/matrix_operation/determinant.h:13-55 and include/matrix_operation/inversion.h:
13-53 (argument order, in-place row-major lda x lda matrix, optional pivots, solver type,
caller-owned workspace). LU mode factors ``m`` first; LU_UPFRONT uses the caller's factors
and pivots. The determinant carries the interchange sign, as the GPU reference's does
(src/matrix_operation/gpu/hipblas/inversion_hip.hip:36-58); the CPU reference returns the
unsigned diagonal product (src/matrix_operation/cpu/determinant.cpp:19-28). Cholesky modes use
the lower factor; their determinant is the squared diagonal product, again the GPU reference's
value. Every computation runs in the HIP library; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes

import torch

from . import _lib
from .lu import _check, _check_matrix, _ptr, _sfx, _stream_ptr, linear_solver_type


def _check_pivot(pivot, lda: int, type: linear_solver_type, what: str) -> None:
    if pivot is None:
        if int(type) == linear_solver_type.LU_UPFRONT_LINEAR_SOLVER:
            raise ValueError(f"{what}: LU_UPFRONT needs the factorisation's pivots")
        return
    if not pivot.is_cuda or pivot.dtype != torch.int64 or not pivot.is_contiguous() or pivot.numel() < lda:
        raise ValueError(f"{what}: pivot must be a contiguous device int64 tensor of lda entries")


def _workspace(fn: str, lda: int, type, dtype: torch.dtype, what: str) -> int:
    out = ctypes.c_size_t(0)
    _check(getattr(_lib.load(), f"{fn}_{_sfx(dtype)}")(int(lda), int(type), ctypes.byref(out)), what)
    return int(out.value)


class determinant:
    """determinant<T, ROCM_NATIVE> (T = float64 or float32 from ``m``)."""

    @staticmethod
    def workspace_size(lda: int, type: linear_solver_type, dtype: torch.dtype = torch.float64) -> int:
        return _workspace("lrn_determinant_workspace_size", lda, type, dtype, "determinant::workspace_size")

    @staticmethod
    def apply(m: torch.Tensor, pivot, lda: int, type: linear_solver_type, workspace: torch.Tensor | None = None,
              workspace_size: int | None = None, stream=None) -> float:
        """det(A) on the host; ``m`` is factored in place in LU mode (synchronous)."""
        _check_matrix(m, lda, "determinant::apply")
        _check_pivot(pivot, lda, type, "determinant::apply")
        sfx = _sfx(m.dtype)
        if workspace is None:
            workspace = torch.empty(max(determinant.workspace_size(lda, type, m.dtype), 1), dtype=torch.uint8,
                                    device=m.device)
        ws_bytes = workspace.numel() * workspace.element_size() if workspace_size is None else int(workspace_size)
        out = ctypes.c_double(0.0) if sfx == "f64" else ctypes.c_float(0.0)
        code = getattr(_lib.load(), f"lrn_determinant_{sfx}")(
            _ptr(m), _ptr(pivot), int(lda), int(type), _ptr(workspace), ws_bytes, ctypes.byref(out),
            _stream_ptr(stream))
        _check(code, "determinant::apply")
        return float(out.value)


class inversion:
    """inversion<T, ROCM_NATIVE>: ``m`` is overwritten with its inverse (synchronous)."""

    @staticmethod
    def workspace_size(lda: int, type: linear_solver_type, dtype: torch.dtype = torch.float64) -> int:
        return _workspace("lrn_inversion_workspace_size", lda, type, dtype, "inversion::workspace_size")

    @staticmethod
    def apply(m: torch.Tensor, pivot, lda: int, type: linear_solver_type, workspace: torch.Tensor | None = None,
              workspace_size: int | None = None, stream=None) -> None:
        _check_matrix(m, lda, "inversion::apply")
        _check_pivot(pivot, lda, type, "inversion::apply")
        sfx = _sfx(m.dtype)
        if workspace is None:
            workspace = torch.empty(inversion.workspace_size(lda, type, m.dtype), dtype=torch.uint8, device=m.device)
        ws_bytes = workspace.numel() * workspace.element_size() if workspace_size is None else int(workspace_size)
        code = getattr(_lib.load(), f"lrn_inversion_{sfx}")(
            _ptr(m), _ptr(pivot), int(lda), int(type), _ptr(workspace), ws_bytes, _stream_ptr(stream))
        _check(code, "inversion::apply")

"""triangular_solver<T, ROCM_NATIVE>: op(A) X = B for a triangular A with many right-hand sides.

Mirrors include/matrix_operation/triangular_solver.h:11-38 (argument order, the four flags'
meaning, row-major storage with row strides, caller-owned workspace, A and B preserved,
only the selected triangle of A read). The solve runs on the GPU through
``lrn_triangular_solver_{f64,f32}``; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes

import torch

from . import _lib
from .lu import _check, _ptr, _sfx, _stream_ptr


def _check_operand(t: torch.Tensor, rows: int, cols: int, ld: int, dtype: torch.dtype, what: str) -> None:
    if t is None or not isinstance(t, torch.Tensor) or not t.is_cuda or not t.is_contiguous():
        raise ValueError(f"triangular_solver::apply: {what} must be a contiguous device tensor")
    if t.dtype != dtype:
        raise ValueError(f"triangular_solver::apply: {what} must have the dtype of a")
    if t.numel() < (rows - 1) * ld + cols:
        raise ValueError(f"triangular_solver::apply: {what} holds fewer than {rows} rows of stride {ld}")


class triangular_solver:
    """triangular_solver<T, ROCM_NATIVE> (T = float64 or float32 from the tensors)."""

    @staticmethod
    def workspace_size(n: int, nrhs: int, dtype: torch.dtype = torch.float64) -> int:
        if n <= 0 or nrhs <= 0:
            raise ValueError("triangular_solver: dimensions must be positive")
        out = ctypes.c_size_t(0)
        _check(getattr(_lib.load(), f"lrn_triangular_solver_workspace_size_{_sfx(dtype)}")(int(n), int(nrhs),
                                                                                          ctypes.byref(out)),
               "triangular_solver::workspace_size")
        return int(out.value)

    @staticmethod
    def apply(upper: bool, transpose: bool, unit_diagonal: bool, n: int, nrhs: int, a: torch.Tensor, lda: int,
              b: torch.Tensor, ldb: int, x: torch.Tensor, ldx: int, workspace: torch.Tensor | None = None,
              workspace_size: int | None = None, stream=None) -> None:
        """Solve op(A) X = B into ``x`` (asynchronous on ``stream``)."""
        if n <= 0 or nrhs <= 0:
            raise ValueError("triangular_solver: dimensions must be positive")
        if lda < n or ldb < nrhs or ldx < nrhs:
            raise ValueError("triangular_solver: invalid row stride")
        if a is None or b is None or x is None:
            raise ValueError("triangular_solver: matrices must not be null")
        _check_operand(a, n, n, lda, a.dtype, "a")
        _check_operand(b, n, nrhs, ldb, a.dtype, "b")
        _check_operand(x, n, nrhs, ldx, a.dtype, "x")
        if a.data_ptr() == b.data_ptr() or a.data_ptr() == x.data_ptr() or b.data_ptr() == x.data_ptr():
            raise ValueError("triangular_solver: matrices must use distinct buffers")
        sfx = _sfx(a.dtype)
        if workspace is None:
            workspace = torch.empty(triangular_solver.workspace_size(n, nrhs, a.dtype), dtype=torch.uint8,
                                    device=a.device)
        ws_bytes = workspace.numel() * workspace.element_size() if workspace_size is None else int(workspace_size)
        code = getattr(_lib.load(), f"lrn_triangular_solver_{sfx}")(
            int(bool(upper)), int(bool(transpose)), int(bool(unit_diagonal)), int(n), int(nrhs), _ptr(a), int(lda),
            _ptr(b), int(ldb), _ptr(x), int(ldx), _ptr(workspace), ws_bytes, _stream_ptr(stream))
        _check(code, "triangular_solver::apply")

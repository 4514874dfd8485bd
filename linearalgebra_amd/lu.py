"""Host-side mirror of the reference's LU operator interface for the ROCM_NATIVE backend.

Names, argument meaning and error behaviour follow the reference's templates
(paths relative to the reference repository):

  lu_decomposition.apply / workspace_size   include/matrix_operation/lu_decomposition.h:15-54
  linear_solver.apply / workspace_size      include/matrix_operation/linear_solver.h:26-65
  linear_solver_type                        include/matrix_operation/linear_solver.h:15-21

Matrices are row-major square device tensors (torch, float32/float64) with the order as
the row stride; pivots are device int64 tensors, 1-based. Invalid arguments raise
(``ValueError``), as the reference's LINALG_CHECK throws; runtime failures raise
``RuntimeError``. Torch is used only for device memory and the stream handle: every
computation runs in the HIP kernels of ``liblinalg_rocm_native.so``.
"""
from __future__ import annotations

import ctypes
import enum

import torch

from . import _lib


class linear_solver_type(enum.IntEnum):
    LU_LINEAR_SOLVER = 1
    LU_UPFRONT_LINEAR_SOLVER = 2
    CHOLESKY_LINEAR_SOLVER = 3
    CHOLESKY_UPFRONT_LINEAR_SOLVER = 4


class LinalgError(RuntimeError):
    pass


def _sfx(dtype: torch.dtype) -> str:
    if dtype == torch.float64:
        return "f64"
    if dtype == torch.float32:
        return "f32"
    raise ValueError(f"unsupported dtype {dtype}: float32 and float64 only")


def _check(code: int, what: str) -> None:
    if code == 0:
        return
    if code < 0:
        raise ValueError(f"{what}: invalid argument {-code} ({_lib.status_text(code)})")
    raise LinalgError(f"{what}: {_lib.status_text(code)} (status {code})")


def _stream_ptr(stream) -> ctypes.c_void_p:
    if stream is None:
        stream = torch.cuda.current_stream()
    if isinstance(stream, torch.cuda.Stream):
        return ctypes.c_void_p(stream.cuda_stream)
    return ctypes.c_void_p(int(stream))


def _ptr(t) -> ctypes.c_void_p:
    return ctypes.c_void_p(0 if t is None else t.data_ptr())


def _check_matrix(m: torch.Tensor, lda: int, what: str) -> None:
    if m is None:
        raise ValueError(f"{what}: m must not be null")
    if lda <= 0:
        raise ValueError(f"{what}: lda must be positive ({lda})")
    if not m.is_cuda:
        raise ValueError(f"{what}: m must be device memory")
    if not m.is_contiguous() or m.numel() < lda * lda:
        raise ValueError(f"{what}: m must be a contiguous lda x lda row-major matrix")


class lu_decomposition:
    """lu_decomposition<T, ROCM_NATIVE>: in-place packed L\\U with partial pivoting."""

    @staticmethod
    def workspace_size(lda: int, dtype: torch.dtype = torch.float64) -> int:
        out = ctypes.c_size_t(0)
        fn = getattr(_lib.load(), f"lrn_lu_decomposition_workspace_size_{_sfx(dtype)}")
        _check(fn(int(lda), ctypes.byref(out)), "lu_decomposition::workspace_size")
        return int(out.value)

    @staticmethod
    def apply(m: torch.Tensor, lda: int, pivot: torch.Tensor, workspace: torch.Tensor | None = None,
              workspace_size: int | None = None, info: list | None = None, stream=None) -> bool:
        """Factor ``m`` in place; fill ``pivot``; return True when U is nonsingular.

        Synchronous on ``stream``. ``info`` (a list) receives the LAPACK getrf status:
        0, or the 1-based index of the first exactly-zero U diagonal entry.
        """
        if m is None or pivot is None:
            raise ValueError("lu_decomposition::apply: m and pivot must not be null")
        _check_matrix(m, lda, "lu_decomposition::apply")
        sfx = _sfx(m.dtype)
        if not pivot.is_cuda or pivot.dtype != torch.int64 or pivot.numel() < lda:
            raise ValueError("lu_decomposition::apply: pivot must be a device int64 array of lda entries")
        if workspace is None:
            workspace = torch.empty(lu_decomposition.workspace_size(lda, m.dtype), dtype=torch.uint8, device=m.device)
        ws_bytes = workspace.numel() * workspace.element_size() if workspace_size is None else int(workspace_size)
        inf = ctypes.c_int(0)
        fn = getattr(_lib.load(), f"lrn_lu_decomposition_{sfx}")
        _check(fn(_ptr(m), int(lda), _ptr(pivot), _ptr(workspace), ws_bytes, ctypes.byref(inf), _stream_ptr(stream)),
               "lu_decomposition::apply")
        if info is not None:
            info.clear()
            info.append(int(inf.value))
        return inf.value == 0


class linear_solver:
    """linear_solver<T, ROCM_NATIVE>: LU and Cholesky modes; x (or an lda x nrhs block) is overwritten.
    Cholesky modes use the lower factor and ignore ``pivot`` (may be None)."""

    @staticmethod
    def workspace_size(lda: int, type: linear_solver_type, dtype: torch.dtype = torch.float64, nrhs: int = 1) -> int:
        out = ctypes.c_size_t(0)
        lib = _lib.load()
        sfx = _sfx(dtype)
        if nrhs == 1:
            code = getattr(lib, f"lrn_linear_solver_workspace_size_{sfx}")(int(lda), int(type), ctypes.byref(out))
        else:
            code = getattr(lib, f"lrn_linear_solver_multi_workspace_size_{sfx}")(
                int(lda), int(nrhs), int(type), ctypes.byref(out))
        _check(code, "linear_solver::workspace_size")
        return int(out.value)

    @staticmethod
    def apply(m: torch.Tensor, pivot: torch.Tensor, lda: int, x: torch.Tensor, type: linear_solver_type,
              workspace: torch.Tensor | None = None, workspace_size: int | None = None, stream=None) -> None:
        """Solve A X = B in place of ``x`` (shape (lda,) or (lda, nrhs)). Asynchronous on ``stream``."""
        if m is None or x is None:
            raise ValueError("linear_solver::apply: m and x must not be null")
        _check_matrix(m, lda, "linear_solver::apply")
        sfx = _sfx(m.dtype)
        if x.dtype != m.dtype or not x.is_cuda or not x.is_contiguous() or x.shape[0] < lda:
            raise ValueError("linear_solver::apply: x must be a contiguous device array of lda rows, same dtype as m")
        type = linear_solver_type(int(type))
        chol = type in (linear_solver_type.CHOLESKY_LINEAR_SOLVER, linear_solver_type.CHOLESKY_UPFRONT_LINEAR_SOLVER)
        if not (chol and pivot is None) and (pivot is None or not pivot.is_cuda or pivot.dtype != torch.int64):
            raise ValueError("linear_solver::apply: pivot must be a device int64 array")
        nrhs = 1 if x.dim() == 1 else int(x.shape[1])
        if workspace is None:
            workspace = torch.empty(linear_solver.workspace_size(lda, type, m.dtype, nrhs),
                                    dtype=torch.uint8, device=m.device)
        ws_bytes = workspace.numel() * workspace.element_size() if workspace_size is None else int(workspace_size)
        lib = _lib.load()
        if x.dim() == 1:
            code = getattr(lib, f"lrn_linear_solver_{sfx}")(
                _ptr(m), _ptr(pivot), int(lda), _ptr(x), int(type), _ptr(workspace), ws_bytes, _stream_ptr(stream))
        else:
            code = getattr(lib, f"lrn_linear_solver_multi_{sfx}")(
                _ptr(m), _ptr(pivot), int(lda), _ptr(x), nrhs, nrhs, int(type), _ptr(workspace), ws_bytes,
                _stream_ptr(stream))
        _check(code, "linear_solver::apply")


class cholesky_decomposition_enum(enum.Enum):
    LOWER_TRIANGULAR = "L"
    UPPER_TRIANGULAR = "U"


class cholesky_decomposition:
    """cholesky_decomposition<T, ROCM_NATIVE> (include/matrix_operation/cholesky_decomposition.h):
    the selected factor in place (L with A = L L^T, or U with A = U^T U), zeros in the other
    triangle; only the lower triangle of the input is read. Synchronous; returns True on
    success and appends the LAPACK-style info (first non-positive-definite leading minor,
    1-based, 0 on success) to ``info`` when a list is given."""

    @staticmethod
    def workspace_size(lda: int, dtype: torch.dtype = torch.float64) -> int:
        out = ctypes.c_size_t(0)
        fn = getattr(_lib.load(), f"lrn_cholesky_decomposition_workspace_size_{_sfx(dtype)}")
        _check(fn(int(lda), ctypes.byref(out)), "cholesky_decomposition::workspace_size")
        return int(out.value)

    @staticmethod
    def apply(c: torch.Tensor, lda: int, type: cholesky_decomposition_enum = cholesky_decomposition_enum.LOWER_TRIANGULAR,
              workspace: torch.Tensor | None = None, workspace_size: int | None = None, info: list | None = None,
              stream=None) -> bool:
        _check_matrix(c, lda, "cholesky_decomposition::apply")
        if not isinstance(type, cholesky_decomposition_enum):
            raise ValueError("cholesky_decomposition::apply: Unsupported enum type!")
        sfx = _sfx(c.dtype)
        if workspace is None:
            workspace = torch.empty(cholesky_decomposition.workspace_size(lda, c.dtype), dtype=torch.uint8,
                                    device=c.device)
        ws_bytes = workspace.numel() * workspace.element_size() if workspace_size is None else int(workspace_size)
        inf = ctypes.c_int(0)
        upper = 1 if type == cholesky_decomposition_enum.UPPER_TRIANGULAR else 0
        code = getattr(_lib.load(), f"lrn_cholesky_decomposition_{sfx}")(
            _ptr(c), int(lda), upper, _ptr(workspace), ws_bytes, ctypes.byref(inf), _stream_ptr(stream))
        _check(code, "cholesky_decomposition::apply")
        if info is not None:
            info.append(int(inf.value))
        return inf.value == 0


def set_tuning(outer_nb: int = 0, panel_w: int = 0, panel_max_wg: int = 0) -> None:
    _check(_lib.load().lrn_set_tuning(int(outer_nb), int(panel_w), int(panel_max_wg)), "set_tuning")


def get_tuning() -> tuple[int, int, int]:
    a, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    _lib.load().lrn_get_tuning(ctypes.byref(a), ctypes.byref(b), ctypes.byref(c))
    return a.value, b.value, c.value

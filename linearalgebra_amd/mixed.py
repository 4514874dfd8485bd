"""Mixed-precision FP64 solve (build-owned extension of linear_solver, no reference counterpart).

A is rounded to FP32 and factored with FP32 panels and matrix-core trailing updates (FP32
instructions, or FP32 operands split into 1-3 bf16 planes; FP32 accumulation); LU-IR, then
GMRES-IR, then (optionally) a plain FP64 LU recover an FP64-level normwise backward error.
All compute is in liblinalg_rocm_native.so.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib
from .lu import LinalgError, _check, _ptr, _stream_ptr

PATHS = {0: "lu-ir", 1: "gmres-ir", 2: "fp64-fallback", 3: "not-converged"}


class MixedStats(ctypes.Structure):
    _fields_ = [
        ("info", ctypes.c_int),
        ("path", ctypes.c_int),
        ("ir_iterations", ctypes.c_int),
        ("gmres_outer", ctypes.c_int),
        ("gmres_iterations", ctypes.c_int),
        ("backward_error", ctypes.c_double),
        ("factor_ms", ctypes.c_double),
        ("refine_ms", ctypes.c_double),
    ]

    def as_dict(self) -> dict:
        d = {f: getattr(self, f) for f, _ in self._fields_}
        d["path_name"] = PATHS.get(self.path, "?")
        return d


def workspace_size(n: int, restart: int = 30, fallback: bool = True) -> int:
    out = ctypes.c_size_t(0)
    _check(_lib.load().lrn_gesv_mixed_workspace_size_f64(int(n), int(restart), int(bool(fallback)), ctypes.byref(out)),
           "gesv_mixed::workspace_size")
    return int(out.value)


def set_mixed_update(mode: int) -> None:
    """Trailing update of the low-precision factorisation (process-wide): 0, 1, 2 = FP32
    operands split into 1, 2, 3 bf16 planes on the bf16 matrix cores (2 gives FP32-level
    factors); 3 = FP32 matrix instructions (the default); 4, 5, 6 = 1, 2, 3 planes of FP8
    e4m3 with one power-of-two scale per 32 K elements (block-scaled FP8 matrix cores)."""
    _check(_lib.load().lrn_set_mixed_update(int(mode)), "set_mixed_update")


def set_mixed_panel(tournament: bool) -> None:
    """Panel of the FP32 factorisation inside the mixed solver (process-wide): True =
    tournament pivoting, False = the partial-pivoting panel of the plain path (the default)."""
    _check(_lib.load().lrn_set_mixed_panel(int(bool(tournament))), "set_mixed_panel")


def getrf_tournament(m: torch.Tensor, pivot: torch.Tensor | None = None, stream=None):
    """FP32 LU with tournament pivoting, in place (the factorisation of the mixed solver).
    Returns (pivot, info) with 1-based sequential-interchange pivots; synchronises the stream
    to read info."""
    if m is None or m.dtype != torch.float32 or not m.is_cuda or m.dim() != 2 or m.shape[0] != m.shape[1]:
        raise ValueError("getrf_tournament: m must be a square float32 device matrix")
    if not m.is_contiguous():
        raise ValueError("getrf_tournament: m must be contiguous (row-major, lda = n)")
    n = m.shape[0]
    if pivot is None:
        pivot = torch.empty(n, dtype=torch.int64, device=m.device)
    lib = _lib.load()
    out = ctypes.c_size_t(0)
    _check(lib.lrn_getrf_tournament_workspace_size_f32(int(n), ctypes.byref(out)), "getrf_tournament::workspace_size")
    ws = torch.empty(int(out.value), dtype=torch.uint8, device=m.device)
    info = torch.zeros(1, dtype=torch.int32, device=m.device)
    _check(lib.lrn_getrf_tournament_f32(_ptr(m), int(n), _ptr(pivot), _ptr(info), _ptr(ws), ws.numel(),
                                        _stream_ptr(stream)), "getrf_tournament")
    return pivot, int(info.cpu().item())


DEFAULT_REFINE_SOLVER = 4


def set_refine_solver(chain: int) -> None:
    """Triangular solves of the refinement (process-wide): 0 / False = the 64-row substitution
    kernel of the plain single-RHS solve, 1 / True = 256-row chained kernel, 2 = 128-row
    chained kernel, 3 = 64-row blocks in the inverse form (each diagonal block and the two
    preceding off-diagonal blocks applied through products with the block's inverse, computed
    in FP64 once per factorisation), 4 = mode 3 with runs of consecutive blocks kept on one XCD
    (the default, ``DEFAULT_REFINE_SOLVER``)."""
    _check(_lib.load().lrn_set_refine_solver(int(chain)), "set_refine_solver")


def getrs_f32f64(lu: torch.Tensor, pivot: torch.Tensor, b: torch.Tensor, stream=None) -> torch.Tensor:
    """b <- U^-1 L^-1 P b in place: FP32 factors applied in FP64 to one FP64 right-hand side."""
    if lu is None or lu.dtype != torch.float32 or not lu.is_cuda or lu.dim() != 2 or not lu.is_contiguous():
        raise ValueError("getrs_f32f64: lu must be a contiguous square float32 device matrix")
    n = lu.shape[0]
    if b is None or b.dtype != torch.float64 or not b.is_cuda or not b.is_contiguous() or b.numel() != n:
        raise ValueError("getrs_f32f64: b must be a contiguous float64 device vector of n")
    lib = _lib.load()
    out = ctypes.c_size_t(0)
    _check(lib.lrn_getrs_f32f64_workspace_size(int(n), ctypes.byref(out)), "getrs_f32f64::workspace_size")
    ws = torch.empty(int(out.value), dtype=torch.uint8, device=lu.device)
    _check(lib.lrn_getrs_f32f64(_ptr(lu), int(n), _ptr(pivot), _ptr(b), _ptr(ws), ws.numel(), _stream_ptr(stream)),
           "getrs_f32f64")
    return b


def gesv_mixed(a: torch.Tensor, b: torch.Tensor, x: torch.Tensor | None = None, tol: float = 1e-14,
               max_iter: int = 20, restart: int = 30, workspace: torch.Tensor | None = None,
               fallback: bool = True, stream=None):
    """Solve a x = b (FP64, one right-hand side). Returns (x, stats dict). ``a`` is not modified."""
    if a is None or b is None:
        raise ValueError("gesv_mixed: a and b must not be null")
    n = a.shape[0]
    if a.dtype != torch.float64 or b.dtype != torch.float64 or not a.is_cuda or not b.is_cuda:
        raise ValueError("gesv_mixed: a and b must be float64 device tensors")
    if not a.is_contiguous() or a.numel() != n * n or b.numel() < n:
        raise ValueError("gesv_mixed: a must be a contiguous n x n matrix and b a vector of n")
    if x is None:
        x = torch.empty(n, dtype=torch.float64, device=a.device)
    if workspace is None:
        workspace = torch.empty(workspace_size(n, restart, fallback), dtype=torch.uint8, device=a.device)
    st = MixedStats()
    code = _lib.load().lrn_gesv_mixed_f64(_ptr(a), int(n), _ptr(b), _ptr(x), float(tol), int(max_iter), int(restart),
                                          _ptr(workspace), workspace.numel(), ctypes.byref(st), _stream_ptr(stream))
    _check(code, "gesv_mixed")
    return x, st.as_dict()


def gesv_mixed_batched(a: torch.Tensor, n, b: torch.Tensor, modes=None, tol: float = 1e-14, max_iter: int = 30,
                       restart: int = 30, fallback: bool = True, stream=None):
    """Mixed-precision solves of many systems packed back to back: ``a`` holds the row-major
    A_i (n_i^2 values each), ``b`` the b_i (n_i values each), both FP64 device tensors.
    ``modes`` (optional): per-system trailing-update mode (see set_mixed_update).
    Returns (x packed like b, list of per-system stats dicts). Synchronous."""
    n = np.asarray(n, dtype=np.int64)
    count = int(n.size)
    if a.dtype != torch.float64 or b.dtype != torch.float64 or not a.is_cuda or not b.is_cuda:
        raise ValueError("gesv_mixed_batched: a and b must be float64 device tensors")
    a_off = np.zeros(count, dtype=np.int64)
    b_off = np.zeros(count, dtype=np.int64)
    if count:
        a_off[1:] = np.cumsum(n * n)[:-1]
        b_off[1:] = np.cumsum(n)[:-1]
        if a.numel() < int((n * n).sum()) or b.numel() < int(n.sum()):
            raise ValueError("gesv_mixed_batched: a or b smaller than the packed systems")
    x = torch.empty_like(b)
    ws = torch.empty(workspace_size(int(n.max()) if count else 1, restart, fallback), dtype=torch.uint8,
                     device=a.device)
    st = (MixedStats * max(count, 1))()
    md = None
    if modes is not None:
        md = np.ascontiguousarray(modes, dtype=np.int32)
    code = _lib.load().lrn_gesv_mixed_batched_f64(
        _ptr(a), a_off.ctypes.data_as(ctypes.c_void_p), n.ctypes.data_as(ctypes.c_void_p), _ptr(b),
        b_off.ctypes.data_as(ctypes.c_void_p), _ptr(x), b_off.ctypes.data_as(ctypes.c_void_p), count,
        md.ctypes.data_as(ctypes.c_void_p) if md is not None else None, float(tol), int(max_iter), int(restart),
        _ptr(ws), ws.numel(), ctypes.cast(st, ctypes.c_void_p), _stream_ptr(stream))
    _check(code, "gesv_mixed_batched")
    return x, [st[i].as_dict() for i in range(count)]


__all__ = ["gesv_mixed", "gesv_mixed_batched", "set_mixed_update", "workspace_size", "MixedStats", "LinalgError"]

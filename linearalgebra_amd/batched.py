"""Batched solve of many small independent systems of uneven size (n <= 256, nrhs <= 8) in
one persistent kernel: lrn_batched_gesv_{f64,f32} of include/linalg_rocm_native.h.

``BatchedSystems`` packs the systems into one device buffer for A and one for B with offset
tables (the layout the C ABI takes); ``solve`` runs the kernel (asynchronous, capturable in
a CUDA/HIP graph after a first eager call). No reference counterpart: the reference solves
one system per call (include/matrix_operation/linear_solver.h:46-53).
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib
from .lu import LinalgError, _check, _stream_ptr

MAX_N = 256
MAX_RHS = 8


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


class BatchedSystems:
    """Device-resident batch: A_i (n_i x n_i row-major) and B_i (n_i x nrhs_i row-major)."""

    def __init__(self, n, nrhs, dtype=torch.float64, device="cuda"):
        n = np.asarray(n, dtype=np.int64)
        nrhs = np.asarray(nrhs, dtype=np.int64)
        if n.shape != nrhs.shape or n.ndim != 1:
            raise ValueError("n and nrhs must be 1-D arrays of the same length")
        if dtype not in (torch.float64, torch.float32):
            raise ValueError("float64 and float32 only")
        self.count = int(n.size)
        self.dtype = dtype
        self.n_host, self.nrhs_host = n, nrhs
        a_off = np.zeros(self.count, dtype=np.int64)
        b_off = np.zeros(self.count, dtype=np.int64)
        p_off = np.zeros(self.count, dtype=np.int64)
        if self.count:
            a_off[1:] = np.cumsum(n * n)[:-1]
            b_off[1:] = np.cumsum(n * nrhs)[:-1]
            p_off[1:] = np.cumsum(n)[:-1]
        self.a_off_host, self.b_off_host, self.p_off_host = a_off, b_off, p_off
        dev = torch.device(device)
        tot_a = int((n * n).sum()) if self.count else 0
        tot_b = int((n * nrhs).sum()) if self.count else 0
        self.a = torch.zeros(max(tot_a, 1), dtype=dtype, device=dev)
        self.b = torch.zeros(max(tot_b, 1), dtype=dtype, device=dev)
        self.pivot = torch.zeros(max(int(n.sum()) if self.count else 0, 1), dtype=torch.int64, device=dev)
        self.info = torch.zeros(max(self.count, 1), dtype=torch.int32, device=dev)
        self.n = torch.from_numpy(n.astype(np.int32)).to(dev)
        self.nrhs = torch.from_numpy(nrhs.astype(np.int32)).to(dev)
        self.a_off = torch.from_numpy(a_off).to(dev)
        self.b_off = torch.from_numpy(b_off).to(dev)
        self.p_off = torch.from_numpy(p_off).to(dev)
        size = ctypes.c_size_t(0)
        _check(_lib.load().lrn_batched_gesv_workspace_size(self.count, ctypes.byref(size)),
               "batched_gesv::workspace_size")
        self.workspace = torch.empty(max(size.value, 1), dtype=torch.uint8, device=dev)

    def set_host(self, a_list, b_list):
        """Fill A and B from lists of numpy arrays (row-major)."""
        a = np.concatenate([np.ascontiguousarray(x).ravel() for x in a_list]) if a_list else np.zeros(0)
        b = np.concatenate([np.ascontiguousarray(x).ravel() for x in b_list]) if b_list else np.zeros(0)
        np_dt = np.float64 if self.dtype == torch.float64 else np.float32
        self.a[: a.size].copy_(torch.from_numpy(a.astype(np_dt)))
        self.b[: b.size].copy_(torch.from_numpy(b.astype(np_dt)))

    def system(self, i):
        """(LU_i, pivot_i, X_i) of system i as numpy arrays."""
        n, r = int(self.n_host[i]), int(self.nrhs_host[i])
        ao, bo, po = int(self.a_off_host[i]), int(self.b_off_host[i]), int(self.p_off_host[i])
        lu = self.a[ao: ao + n * n].cpu().numpy().reshape(n, n)
        piv = self.pivot[po: po + n].cpu().numpy()
        x = self.b[bo: bo + n * r].cpu().numpy().reshape(n, r)
        return lu, piv, x

    def solve(self, stream=None):
        sfx = "f64" if self.dtype == torch.float64 else "f32"
        fn = getattr(_lib.load(), f"lrn_batched_gesv_{sfx}")
        code = fn(_ptr(self.a), _ptr(self.a_off), _ptr(self.pivot), _ptr(self.p_off), _ptr(self.b), _ptr(self.b_off),
                  _ptr(self.n), _ptr(self.nrhs), self.count, _ptr(self.info), _ptr(self.workspace),
                  self.workspace.numel(), _stream_ptr(stream))
        _check(code, "batched_gesv")


def set_grid(workgroups: int) -> None:
    """Workgroups of the persistent launch (0 = occupancy x compute units)."""
    _check(_lib.load().lrn_set_batched_grid(int(workgroups)), "batched_gesv::set_grid")


__all__ = ["BatchedSystems", "set_grid", "LinalgError", "MAX_N", "MAX_RHS"]

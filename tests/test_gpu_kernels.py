"""Kernel-level checks of the building blocks (through the C ABI test hooks).

The GEMM and TRSM are compared with a float64 numpy evaluation of the same operation
(1e-13 relative for f64, 1e-5 for f32); the panel with LAPACK getrf (scipy) on the same
column block -- pivots identical, factors within 1e-13.
"""
import ctypes

import numpy as np
import pytest
import scipy.linalg.lapack as lapack

pytestmark = pytest.mark.gpu


def _lib():
    from linearalgebra_amd import _lib as L
    return L.load()


def _vp(t):
    return ctypes.c_void_p(t.data_ptr())


@pytest.mark.parametrize("dt", ["f64", "f32"])
@pytest.mark.parametrize("M,N,K", [(16, 16, 4), (17, 33, 5), (128, 128, 64), (300, 200, 64), (1000, 999, 256),
                                   (63, 63, 64), (130, 7, 16)])
def test_gemm_sub(M, N, K, dt, dev):
    import torch
    rng = np.random.default_rng(M * 1000 + N + K)
    npdt = np.float64 if dt == "f64" else np.float32
    a = rng.standard_normal((M, K)).astype(npdt)
    b = rng.standard_normal((K, N)).astype(npdt)
    c = rng.standard_normal((M, N)).astype(npdt)
    ta, tb, tc = (torch.from_numpy(x).to(dev) for x in (a, b, c))
    code = _lib().lrn_debug_gemm_sub(0 if dt == "f64" else 1, _vp(ta), K, _vp(tb), N, _vp(tc), N, M, N, K, None)
    assert code == 0
    torch.cuda.synchronize()
    got = tc.cpu().numpy().astype(np.float64)
    ref = c.astype(np.float64) - a.astype(np.float64) @ b.astype(np.float64)
    err = np.abs(got - ref).max() / max(1.0, np.abs(ref).max())
    print(f"gemm {dt} {M}x{N}x{K} err={err:.3e}", flush=True)
    assert err <= (1e-13 if dt == "f64" else 1e-5)


@pytest.mark.parametrize("n,k,w,c0,c1", [(100, 10, 64, 74, 100), (100, 0, 37, 37, 100), (300, 64, 64, 128, 300),
                                         (70, 3, 16, 19, 70), (50, 0, 1, 1, 50), (2000, 100, 64, 164, 2000)])
def test_trsm_unit_lower(n, k, w, c0, c1, dev):
    """Bitwise against the reference's order: x[ii] -= L[ii][p] * x[p] for p = 0..ii-1, each
    product rounded before the subtraction (lu_decomposition_native.cpp:370-380)."""
    import torch
    rng = np.random.default_rng(n + k + w)
    a = rng.standard_normal((n, n))
    a[k:k + w, k:k + w] = rng.uniform(-1.0, 1.0, (w, w))  # multipliers of a pivoted LU are <= 1
    t = torch.from_numpy(a.copy()).to(dev)
    assert _lib().lrn_debug_trsm(0, _vp(t), n, k, w, c0, c1, None) == 0
    torch.cuda.synchronize()
    got = t.cpu().numpy()
    ref = a.copy()
    x = ref[k:k + w, c0:c1]
    for ii in range(w):
        for p in range(ii):
            x[ii] = x[ii] - a[k + ii, k + p] * x[p]
    lo = np.tril(a[k:k + w, k:k + w], -1) + np.eye(w)
    assert np.abs(lo @ x - a[k:k + w, c0:c1]).max() <= 1e-10 * max(1.0, np.abs(x).max())
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("n,k,w", [(64, 0, 64), (127, 64, 63), (127, 0, 64), (300, 0, 64), (300, 100, 64),
                                   (700, 5, 30), (2000, 0, 64), (5000, 128, 64), (20000, 0, 64),
                                   (32, 0, 32), (127, 96, 31), (300, 0, 32), (1000, 3, 32), (5000, 128, 32),
                                   (20000, 0, 32), (33000, 64, 32)])
def test_panel_vs_lapack(n, k, w, dev):
    import torch
    rng = np.random.default_rng(n + 7 * k + w)
    a = rng.standard_normal((n, n))
    t = torch.from_numpy(a.copy()).to(dev)
    piv = torch.zeros(n, dtype=torch.int64, device=dev)
    L = _lib()
    sz = ctypes.c_size_t()
    assert L.lrn_lu_decomposition_workspace_size_f64(n, ctypes.byref(sz)) == 0
    ws = torch.zeros(sz.value, dtype=torch.uint8, device=dev)
    code = L.lrn_debug_panel(0, _vp(t), n, k, w, _vp(piv), _vp(ws), sz.value, None)
    assert code == 0
    torch.cuda.synchronize()
    got = t.cpu().numpy()
    gp = piv.cpu().numpy()[k:k + w] - 1 - k
    lu_ref, ipiv, info = lapack.dgetrf(a[k:, k:k + w].copy())
    err = np.abs(got[k:, k:k + w] - lu_ref).max() / np.abs(lu_ref).max()
    print(f"panel n={n} k={k} w={w} err={err:.3e}", flush=True)
    assert np.array_equal(gp, ipiv[:w])
    assert err <= 1e-13

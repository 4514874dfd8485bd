"""Mixed-precision solve (FP32 panels, bf16 trailing updates, FP64 refinement) on the GPU.

Checks: the bf16 trailing-update GEMM against an FP64 evaluation of the same bf16-rounded
operands (within FP32 accumulation error); the solve's normwise backward error reaches the
requested 1e-14 for kappa in {1e2, 1e6, 1e10}; its forward error against the oracle's FP64
solution stays within kappa * n * eps.
"""
import ctypes

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def _bf16(x):
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    rounded = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return rounded.astype(np.uint32).view(np.float32).astype(np.float64)


@pytest.mark.parametrize("split", [0, 1])
@pytest.mark.parametrize("M,N,K", [(16, 16, 32), (130, 70, 40), (256, 256, 256), (1000, 777, 128)])
def test_gemm_bf16(M, N, K, split, dev):
    import torch
    from linearalgebra_amd import _lib
    rng = np.random.default_rng(M + N + K)
    a = rng.standard_normal((M, K)).astype(np.float32)
    b = rng.standard_normal((K, N)).astype(np.float32)
    c = rng.standard_normal((M, N)).astype(np.float32)
    ta, tb, tc = (torch.from_numpy(x).to(dev) for x in (a, b, c))
    vp = lambda t: ctypes.c_void_p(t.data_ptr())
    assert _lib.load().lrn_debug_gemm_sub(2 + split, vp(ta), K, vp(tb), N, vp(tc), N, M, N, K, None) == 0
    torch.cuda.synchronize()
    if split:
        ah, bh = _bf16(a), _bf16(b)
        at, bt = _bf16((a - ah).astype(np.float32)), _bf16((b - bh).astype(np.float32))
        ref = c.astype(np.float64) - (ah @ bh + ah @ bt + at @ bh)
        full = c.astype(np.float64) - a.astype(np.float64) @ b.astype(np.float64)
        # the split product is within ~2^-16 of the exact FP32-operand product
        assert np.abs(tc.cpu().numpy() - full).max() / np.abs(full).max() <= 1e-4
    else:
        ref = c.astype(np.float64) - _bf16(a) @ _bf16(b)
    err = np.abs(tc.cpu().numpy() - ref).max() / np.abs(ref).max()
    print(f"gemm bf16 split={split} {M}x{N}x{K} err={err:.2e}", flush=True)
    assert err <= 2e-6


@pytest.mark.parametrize("n,kappa", [(300, 1e2), (1000, 1e6), (2048, 1e2), (2048, 1e6), (2048, 1e10)])
def test_gesv_mixed_conditioned(n, kappa, dev):
    import torch
    import linearalgebra_amd as la
    a = O.conditioned(n, kappa, seed=n)
    b = O.gaussian(n, 1, seed=n + 5).ravel()
    ta = torch.from_numpy(a).to(dev)
    tb = torch.from_numpy(b).to(dev)
    x, st = la.gesv_mixed(ta, tb, tol=1e-14, max_iter=30, restart=40)
    xg = x.cpu().numpy()
    berr = O.backward_error(a, xg, b)
    lu, piv, x_ref, info = O.gesv(a, b, threads=16)
    fe = O.forward_error(xg, x_ref)
    print(f"n={n} kappa={kappa:.0e} path={st['path_name']} ir={st['ir_iterations']} gmres={st['gmres_iterations']}"
          f" berr={berr:.2e} (lib {st['backward_error']:.2e}) fwd_vs_oracle={fe:.2e}", flush=True)
    assert np.array_equal(ta.cpu().numpy(), a)  # A is not modified
    assert berr <= 2e-14
    assert st["path"] in (0, 1, 2)
    assert fe <= max(1e-12, kappa * n * np.finfo(np.float64).eps)


def test_gesv_mixed_diag_dominant_lu_ir(dev):
    import torch
    import linearalgebra_amd as la
    n = 1024
    a = O.diag_dominant(n, seed=4)
    b = O.gaussian(n, 1, seed=9).ravel()
    x, st = la.gesv_mixed(torch.from_numpy(a).to(dev), torch.from_numpy(b).to(dev))
    assert st["path"] == 0
    assert O.backward_error(a, x.cpu().numpy(), b) <= 2e-14

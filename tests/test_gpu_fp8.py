"""Block-scaled FP8 (e4m3) trailing update and the mixed solves that use it, on the GPU.

The GEMM is checked against an FP64 evaluation of the same quantised operands (the host model
in tests/fp8_reference.py: scale 2^(E-8) per 32 K elements of a row of A / column of B,
round-to-nearest-even e4m3 planes of the successive residuals): the only difference allowed
is the matrix cores' accumulation, bounded elementwise by 2^-14 (|C| + sum |A_i||B_j|). The
block-scaled FP8 instruction does not accumulate like a sequence of FP32 additions: its error
is about 2^-15.5 of that magnitude whatever K is (measured 6e-6 ... 2.7e-5 for K = 32 ... 300),
where a chain of K rounded FP32 additions would give K 2^-24 at most. A quantisation that
differed from the model by one e4m3 step in one product would show as about 2^-4 / K.
Against the exact product of the FP32 operands the planes keep about 4 / 8 / 12 bits.
"""
import ctypes
import os
import sys

import numpy as np
import pytest

import oracle as O

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import fp8_reference as F  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("planes", [1, 2, 3])
@pytest.mark.parametrize("M,N,K", [(16, 16, 32), (130, 70, 200), (256, 256, 256), (1000, 777, 128), (64, 64, 300)])
def test_gemm_fp8_matches_fp64_of_quantised_operands(M, N, K, planes, dev):
    import torch
    from linearalgebra_amd import _lib
    vp = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731
    rng = np.random.default_rng(M * 11 + N * 3 + K + planes)
    # rows of A and columns of B over six decades, so the per-block scales differ
    a = (rng.standard_normal((M, K)) * 10.0 ** rng.uniform(-3, 3, (M, 1))).astype(np.float32)
    b = (rng.standard_normal((K, N)) * 10.0 ** rng.uniform(-3, 3, (1, N))).astype(np.float32)
    c = rng.standard_normal((M, N)).astype(np.float32)
    ta, tb, tc = (torch.from_numpy(v).to(dev) for v in (a, b, c))
    assert _lib.load().lrn_debug_gemm_sub(4 + planes, vp(ta), K, vp(tb), N, vp(tc), N, M, N, K, None) == 0
    torch.cuda.synchronize()
    got = tc.cpu().numpy().astype(np.float64)
    ref, mag = F.fp8_gemm_sub(a, b, c, planes)
    acc_err = (np.abs(got - ref) / mag).max()
    full = c.astype(np.float64) - a.astype(np.float64) @ b.astype(np.float64)
    ab = np.abs(c.astype(np.float64)) + np.abs(a.astype(np.float64)) @ np.abs(b.astype(np.float64))
    q_err = (np.abs(got - full) / ab).max()
    print(f"gemm fp8 planes={planes} {M}x{N}x{K}: vs quantised-operand FP64 {acc_err:.2e} (bound "
          f"{2.0 ** -14:.1e}), vs exact FP32 operands {q_err:.2e}", flush=True)
    assert acc_err <= 2.0 ** -14
    assert q_err <= {1: 4e-2, 2: 2e-3, 3: 6e-5}[planes]


def test_gemm_fp8_zero_and_nan_blocks(dev):
    """A zero block stays zero (no scale overflow), a NaN in a block reaches exactly the outputs
    whose dot products use it."""
    import torch
    from linearalgebra_amd import _lib
    vp = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731
    M, N, K = 32, 32, 128
    rng = np.random.default_rng(5)
    a = rng.standard_normal((M, K)).astype(np.float32)
    a[3, :] = 0.0
    a[7, 40] = np.nan
    b = rng.standard_normal((K, N)).astype(np.float32)
    c = np.zeros((M, N), dtype=np.float32)
    ta, tb, tc = (torch.from_numpy(v).to(dev) for v in (a, b, c))
    assert _lib.load().lrn_debug_gemm_sub(6, vp(ta), K, vp(tb), N, vp(tc), N, M, N, K, None) == 0
    torch.cuda.synchronize()
    got = tc.cpu().numpy()
    assert np.all(got[3] == 0.0)
    assert np.all(np.isnan(got[7]))
    others = np.delete(got, 7, axis=0)
    assert np.isfinite(others).all()


@pytest.mark.parametrize("mode,kappa,refined", [(5, 1e2, True), (6, 1e2, True), (6, 1e4, True), (4, 1e2, False),
                                                (5, 1e4, False), (6, 1e6, False)])
def test_gesv_mixed_fp8_update(mode, kappa, refined, dev):
    """The mixed solver with FP8 trailing updates (1, 2, 3 e4m3 planes) reaches the FP64
    backward error. Two planes (about 8 bits) converge by refinement up to kappa ~ 1e2, three
    (about 12 bits) up to kappa ~ 1e4; beyond (and with one plane, about 4 bits, whose factors
    leave kappa * u >> 1) the refinement stalls and the FP64 fallback gives the result."""
    import torch
    import linearalgebra_amd as la
    from linearalgebra_amd import mixed as lam
    n = 1500
    a = O.conditioned(n, kappa, seed=90 + mode)
    b = O.gaussian(n, 1, seed=91).ravel()
    lam.set_mixed_update(mode)
    try:
        x, st = la.gesv_mixed(torch.from_numpy(a).to(dev), torch.from_numpy(b).to(dev), tol=1e-14, max_iter=30,
                              restart=40)
    finally:
        lam.set_mixed_update(3)
    berr = O.backward_error(a, x.cpu().numpy(), b)
    print(f"fp8 mode={mode} kappa={kappa:.0e} path={st['path_name']} ir={st['ir_iterations']} "
          f"gmres={st['gmres_iterations']} berr={berr:.2e}", flush=True)
    assert berr <= 2e-14
    if refined:
        assert st["path"] in (0, 1)

"""determinant<T, ROCM_NATIVE> and inversion<T, ROCM_NATIVE> on the GPU against the reference's
known answers and the oracle.

Bar: the determinant of the GPU's own factors equals the oracle's reduction of those factors
bit for bit (same sequential product, same sign rule); against the oracle's factorisation it
agrees to 1e-12 relative (FP64) and to n eps cond(A) (FP32, whose factors round differently). The inverse meets the reference suite's bar
(max|A^-1 A - I| <= 5e-12 FP64, 2e-4 FP32) and is within 1e-11 normwise of the oracle's.
"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

LU = 1
LU_UPFRONT = 2


def _det_gpu(a, dev, type=LU, pivot=True, piv_in=None):
    import torch
    import linearalgebra_amd as la
    n = a.shape[0]
    m = torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    p = None
    if piv_in is not None:
        p = torch.from_numpy(np.ascontiguousarray(piv_in)).to(dev)
    elif pivot:
        p = torch.zeros(n, dtype=torch.int64, device=dev)
    det = la.determinant.apply(m, p, n, la.linear_solver_type(type))
    torch.cuda.synchronize()
    return det, m.cpu().numpy(), (None if p is None else p.cpu().numpy())


def _inv_gpu(a, dev, type=LU, pivot=True, piv_in=None):
    import torch
    import linearalgebra_amd as la
    n = a.shape[0]
    m = torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    p = None
    if piv_in is not None:
        p = torch.from_numpy(np.ascontiguousarray(piv_in)).to(dev)
    elif pivot:
        p = torch.zeros(n, dtype=torch.int64, device=dev)
    la.inversion.apply(m, p, n, la.linear_solver_type(type))
    torch.cuda.synchronize()
    return m.cpu().numpy(), (None if p is None else p.cpu().numpy())


@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-12), (np.float32, 1e-5)])
@pytest.mark.parametrize("pivot", [True, False])
def test_determinant_known_answers(dtype, tol, pivot, dev):
    """Testing/Cxx/TestDeterminant.cxx:82-126 and TestGpuMatrixOperations.cxx:1466-1492."""
    for d, want in (([2, 3], 6), ([2, 3, 4], 24), ([1, 2, 3, 5], 30), ([1], 1), ([1] * 4, 1), ([4, 2], 8)):
        det, _, _ = _det_gpu(np.diag(np.array(d, dtype=dtype)), dev, pivot=pivot)
        assert abs(det - want) <= tol, (d, det)
    assert _det_gpu(np.diag(np.array([1, 0, 1], dtype=dtype)), dev, pivot=pivot)[0] == 0
    assert _det_gpu(np.zeros((3, 3), dtype=dtype), dev, pivot=pivot)[0] == 0
    assert _det_gpu(np.array([[0, 1], [1, 0]], dtype=dtype), dev, pivot=pivot)[0] == -1


@pytest.mark.parametrize("n", [5, 64, 257, 300, 1000])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_determinant_matches_oracle(n, dtype, dev):
    a = (O.gaussian(n, seed=n + 5) * np.sqrt(np.e / n)).astype(dtype)
    det, lu_g, piv_g = _det_gpu(a, dev)
    mine = O.determinant(lu_g, piv_g)
    assert (det == float(mine)) or (np.isnan(det) and np.isnan(mine)), (det, mine)
    lu_o, piv_o, info_o = O.getrf(a)
    ref = float(O.determinant(lu_o, piv_o, info_o))
    # FP64: the factors match the oracle's, so the values agree to rounding. FP32: the trailing
    # updates round differently; two factorisations with backward error ~ n eps ||A|| give
    # determinants within |d det / det| <= ||A^-1|| ||dA|| ~ n eps cond(A) (first order).
    tol = 1e-12 if dtype == np.float64 else n * np.finfo(np.float32).eps * np.linalg.cond(a.astype(np.float64))
    print(f"det n={n} {np.dtype(dtype).name}: gpu {det:.6e} oracle {ref:.6e} tol {tol:.1e}", flush=True)
    assert abs(det - ref) <= tol * abs(ref) + 1e-300


def test_determinant_upfront_and_gpu_reference_case(dev):
    """LU_UPFRONT reduces the caller's factors; the GPU reference's MatrixInversion case
    (TestGpuMatrixOperations.cxx:550-615): det of an SPD matrix is positive."""
    a = O.inversion_spd_case(23, seed=23)
    det, lu_g, piv_g = _det_gpu(a, dev)
    assert det > 0
    det2, lu2, _ = _det_gpu(lu_g, dev, type=LU_UPFRONT, piv_in=piv_g)
    assert det2 == det
    assert np.array_equal(lu2, lu_g)
    assert abs(det - np.linalg.det(a)) <= 1e-12 * abs(det)


def test_determinant_invalid_arguments(dev):
    import torch
    import linearalgebra_amd as la
    m = torch.eye(4, dtype=torch.float64, device=dev)
    with pytest.raises(ValueError):
        la.determinant.apply(None, None, 4, la.linear_solver_type.LU_LINEAR_SOLVER)
    with pytest.raises(ValueError):
        la.determinant.apply(m, None, 0, la.linear_solver_type.LU_LINEAR_SOLVER)
    with pytest.raises(ValueError):
        la.determinant.apply(m, None, 4, 5)
    with pytest.raises(ValueError):
        la.determinant.apply(m, None, 4, la.linear_solver_type.LU_UPFRONT_LINEAR_SOLVER)
    ws = torch.empty(8, dtype=torch.uint8, device=dev)
    with pytest.raises(ValueError):
        la.determinant.apply(m, None, 4, la.linear_solver_type.LU_LINEAR_SOLVER, workspace=ws)


@pytest.mark.parametrize("dim", [3, 198])
@pytest.mark.parametrize("dtype,tol", [(np.float64, 5e-12), (np.float32, 2e-4)])
def test_inversion_reference_bar(dim, dtype, tol, dev):
    """Testing/Cxx/TestInversion.cxx:93-107 (LU mode)."""
    a = O.inversion_spd_case(dim, dtype, seed=dim)
    inv, _ = _inv_gpu(a, dev)
    err = np.abs(inv.astype(np.float64) @ a.astype(np.float64) - np.eye(dim)).max()
    print(f"inverse dim={dim} {np.dtype(dtype).name}: max|inv A - I| {err:.2e}", flush=True)
    assert err <= tol


@pytest.mark.parametrize("n", [1, 2, 63, 64, 65, 130, 300])
@pytest.mark.parametrize("pivot", [True, False])
def test_inversion_matches_oracle(n, pivot, dev):
    a = O.gaussian(n, seed=3 * n + 1) + np.eye(n)
    inv, piv_g = _inv_gpu(a, dev, pivot=pivot)
    lu_o, piv_o, info = O.getrf(a)
    assert info == 0
    if pivot:
        assert np.array_equal(piv_g, piv_o)
    ref = O.inverse_from_factors(lu_o, piv_o) if n <= 130 else np.linalg.inv(a)
    d = O.rel_diff(inv, ref)
    print(f"inverse n={n}: rel diff {d:.2e}", flush=True)
    assert d <= 1e-11


def test_inversion_large_and_upfront(dev):
    n = 1500
    a = O.diag_dominant(n, seed=4)
    inv, piv = _inv_gpu(a, dev)
    assert np.abs(inv @ a - np.eye(n)).max() <= 1e-12
    # LU_UPFRONT: the caller's factors and pivots
    lu_o, piv_o, _ = O.getrf(a)
    inv2, _ = _inv_gpu(lu_o, dev, type=LU_UPFRONT, piv_in=piv_o)
    assert O.rel_diff(inv2, inv) <= 1e-13


def test_inversion_singular_keeps_factors(dev):
    a = O.gaussian(40, seed=9)
    a[:, 7] = 0.0
    lu_o, piv_o, info = O.getrf(a)
    assert info != 0
    out, piv_g = _inv_gpu(a, dev)
    assert np.array_equal(piv_g, piv_o)
    assert O.rel_diff(out, lu_o) <= 1e-13


def test_inversion_invalid_arguments(dev):
    import torch
    import linearalgebra_amd as la
    m = torch.eye(4, dtype=torch.float64, device=dev)
    with pytest.raises(ValueError):
        la.inversion.apply(None, None, 4, la.linear_solver_type.LU_LINEAR_SOLVER)
    with pytest.raises(ValueError):
        la.inversion.apply(m, None, 0, la.linear_solver_type.LU_LINEAR_SOLVER)
    with pytest.raises(ValueError):
        la.inversion.apply(m, None, 4, 5)

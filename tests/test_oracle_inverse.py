"""The oracle's determinant and inverse against the reference's known answers and what the
mathematics fixes (numpy's det/inv, permutation signs, A A^-1 = I)."""
import numpy as np
import pytest

from oracle import oracle as O


def _det(a, signed=True):
    lu, piv, info = O.getrf(np.ascontiguousarray(a))
    return O.determinant(lu, piv, info, signed=signed)


@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-12), (np.float32, 1e-5)])
def test_determinant_reference_known_answers(dtype, tol):
    """Testing/Cxx/TestDeterminant.cxx:82-126 and TestGpuMatrixOperations.cxx:1466-1492."""
    for d, want in (([2, 3], 6), ([2, 3, 4], 24), ([1, 2, 3, 5], 30), ([1], 1), ([1] * 4, 1), ([4, 2], 8)):
        assert abs(float(_det(np.diag(np.array(d, dtype=dtype)))) - want) <= tol
    assert _det(np.diag(np.array([1, 0, 1], dtype=dtype))) == 0
    assert _det(np.zeros((3, 3), dtype=dtype)) == 0


def test_determinant_interchange_sign():
    swap = np.array([[0.0, 1.0], [1.0, 0.0]])
    assert _det(swap) == -1.0 and _det(swap, signed=False) == 1.0
    cyc = np.array([[0.0, 0.0, 2.0], [3.0, 0.0, 0.0], [0.0, 5.0, 0.0]])  # even permutation
    assert _det(cyc) == pytest.approx(np.linalg.det(cyc), rel=1e-15)


@pytest.mark.parametrize("n", [7, 50, 200])
def test_determinant_matches_numpy(n):
    a = O.gaussian(n, seed=n) / np.sqrt(n) * 2.0
    want = np.linalg.det(a)
    got = _det(a)
    assert np.sign(got) == np.sign(want)
    assert abs(got - want) <= 1e-11 * abs(want)
    assert abs(_det(a, signed=False)) == abs(got)
    got32 = float(_det(a.astype(np.float32)))
    assert abs(got32 - want) <= 2e-3 * abs(want)


@pytest.mark.parametrize("dim", [3, 198])
@pytest.mark.parametrize("dtype,tol", [(np.float64, 5e-12), (np.float32, 2e-4)])
def test_inverse_reference_bar(dim, dtype, tol):
    """The reference's bar (Testing/Cxx/TestInversion.cxx:93-107): max|A^-1 A - I| within
    5e-12 (FP64) / 2e-4 (FP32) on A = R R^T (dimension 198 and 3)."""
    a = O.inversion_spd_case(dim, dtype, seed=dim)
    lu, piv, info = O.getrf(a)
    assert info == 0
    inv = O.inverse_from_factors(lu, piv)
    err = np.abs(inv.astype(np.float64) @ a.astype(np.float64) - np.eye(dim)).max()
    assert err <= tol


@pytest.mark.parametrize("n", [1, 2, 33, 90])
def test_inverse_matches_numpy(n):
    a = O.gaussian(n, seed=3 * n + 1) + np.eye(n)
    lu, piv, info = O.getrf(a)
    inv = O.inverse_from_factors(lu, piv)
    assert O.rel_diff(inv, np.linalg.inv(a)) <= 1e-11
    # a plausible slip (pivots ignored) must fail the same comparison on a matrix that pivots
    if n >= 33:
        assert np.any(piv != np.arange(1, n + 1))
        bad = O.inverse_from_factors(lu, np.arange(1, n + 1))
        assert O.rel_diff(bad, np.linalg.inv(a)) > 1e-3

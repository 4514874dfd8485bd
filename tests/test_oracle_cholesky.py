"""The oracle's Cholesky restatement against the reference's known answer and numpy."""
import numpy as np
import pytest

from oracle import oracle as O


@pytest.mark.parametrize("dim", [1, 13, 65, 198])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_cholesky_reference_known_answer(dim, dtype):
    """TestCholeskyDecomposition.cxx:170-210: the factor of C C^T is C within 200 eps."""
    a, c = O.cholesky_known_answer(dim, dtype, seed=dim)
    tol = 200 * np.finfo(dtype).eps
    lo, info = O.cholesky(a)
    assert info == 0
    assert np.abs(lo.astype(np.float64) - c).max() <= tol
    up, info = O.cholesky(a, upper=True)
    assert info == 0 and np.array_equal(up, lo.T)


def test_cholesky_reads_lower_only_and_matches_numpy():
    rng = np.random.default_rng(5)
    g = rng.standard_normal((150, 150))
    a = g @ g.T + 150 * np.eye(150)
    poisoned = a.copy()
    poisoned[np.triu_indices(150, 1)] = np.nan
    lo, info = O.cholesky(poisoned)
    assert info == 0
    assert O.rel_diff(lo, np.linalg.cholesky(a)) <= 1e-14
    assert (lo[np.triu_indices(150, 1)] == 0).all()


def test_cholesky_info_first_failing_minor():
    a = np.eye(5)
    a[3, 3] = -1.0
    _, info = O.cholesky(a)
    assert info == 4
    b = np.eye(70)
    b[66, 66] = 0.0
    assert O.cholesky(b)[1] == 67
    # a slip that loses the block offset would report 3 here
    assert O.cholesky(b)[1] != 67 - 64


@pytest.mark.parametrize("n", [1, 9, 80])
def test_cholesky_solve(n):
    a, c = O.cholesky_known_answer(n, seed=n + 1)
    x0 = np.sin(np.arange(n) + 1.0)
    b = a @ x0
    lo, _ = O.cholesky(a)
    assert np.abs(O.cholesky_solve(lo, b) - x0).max() <= 1e-13

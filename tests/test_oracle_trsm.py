"""The triangular-solver oracle against the reference's known answers and against scipy
(CPU only)."""
import numpy as np
import pytest
import scipy.linalg as sl

from oracle import oracle as O


@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-13), (np.float32, 4e-6)])
def test_trsm_oracle_known_answers(dtype, tol):
    """Every known-answer case of the reference's suite (its tolerance: 1e-13 FP64, 4e-6 FP32),
    with A NaN outside the triangle, so the oracle reads only the selected triangle."""
    count = 0
    for upper, transpose, unit, n, nrhs, lda, ldb, ldx, a, b, expected in O.trsm_known_answer_cases(dtype):
        x = O.trsm(upper, transpose, unit, a[:, :n] if lda == n else a, b[:, :nrhs])
        err = np.abs(x[:, :nrhs].astype(np.float64) - expected[:, :nrhs]).max()
        assert np.isfinite(x).all()
        assert err <= tol, (upper, transpose, unit, n, nrhs, err)
        count += 1
    assert count == 48


@pytest.mark.parametrize("upper", [False, True])
@pytest.mark.parametrize("transpose", [False, True])
@pytest.mark.parametrize("unit", [False, True])
def test_trsm_oracle_matches_scipy(upper, transpose, unit):
    rng = np.random.default_rng(7)
    n, nrhs = 100, 5
    a = 0.1 * rng.standard_normal((n, n)) + 3.0 * np.eye(n)
    b = rng.standard_normal((n, nrhs))
    x = O.trsm(upper, transpose, unit, a, b)
    t = np.triu(a) if upper else np.tril(a)
    if unit:
        np.fill_diagonal(t, 1.0)
    ref = sl.solve_triangular(t.T if transpose else t, b, lower=(upper == transpose), unit_diagonal=unit)
    assert np.abs(x - ref).max() / np.abs(ref).max() <= 1e-13


def test_trsm_oracle_detects_a_wrong_order():
    """A plausible mistake -- solving op(A) with the flags' triangle swapped -- fails the check
    the known-answer test relies on."""
    upper, transpose, unit, n, nrhs, lda, ldb, ldx, a, b, expected = next(
        c for c in O.trsm_known_answer_cases() if c[3] == 17 and c[0] and not c[1] and not c[2])
    x = O.trsm(upper, not transpose, unit, np.nan_to_num(a), b[:, :nrhs])
    assert not (np.abs(x - expected[:, :nrhs]).max() <= 1e-13)

"""Pins the CPU oracle (oracle/lu_oracle.c) before anything is compared against it.

The reference's native LU cannot be compiled here (it includes a header of the absent
Logging library), so the oracle is pinned by:
  * the reference's own known-answer tests (tests/golden/reference_kats.json, transcribed
    from Testing/Cxx/TestLUDecomposition.cxx:283-342 by tests/golden/make_kats.py);
  * LAPACK dgetrf/dgetrs (scipy's OpenBLAS) on well-separated inputs, where the pivot
    sequence does not depend on summation order: pivots must be identical;
  * exact rational arithmetic (brute force) on tiny matrices;
  * invariants: |L| <= 1, reconstruction P^T A = L U to O(n eps), the tie-break rule,
    NaN handling, thread-count invariance, replay consistency.
"""
import json
import os
from fractions import Fraction

import numpy as np
import pytest
import scipy.linalg as sl

import oracle as O

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "reference_kats.json")


def _kats():
    with open(GOLDEN) as f:
        return json.load(f)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_zero_pivot_kats(dtype):
    eps = np.finfo(dtype).eps
    for kat in _kats()["zero_pivot"]:
        a = np.array(kat["a"], dtype=dtype)
        lu, piv, info = O.getrf(a)
        assert info == kat["info"], kat["name"]
        n = a.shape[0]
        assert all(i + 1 <= piv[i] <= n for i in range(n))
        lo, up = O.unpack_lu(lu)
        assert np.abs(a[O.pivots_to_perm(piv)] - lo @ up).max() <= 16 * eps


def test_small_scale_and_subnormal_kats():
    k = _kats()
    a = np.array(k["small_scale"]["a"])
    lu, piv, info = O.getrf(a)
    lo, up = O.unpack_lu(lu)
    assert info == 0
    assert np.abs(a[O.pivots_to_perm(piv)] - lo @ up).max() <= 1e-8 * np.finfo(np.float64).eps
    t = np.ldexp(1.0, k["subnormal"]["exp2_t"])
    assert 0 < t < np.finfo(np.float64).tiny
    a = np.array([[t, 1.0], [t / 2, 1.0]])
    lu, piv, info = O.getrf(a)
    assert info == 0 and np.isfinite(lu).all()
    lo, up = O.unpack_lu(lu)
    assert np.abs(a[O.pivots_to_perm(piv)] - lo @ up).max() < np.finfo(np.float64).tiny


@pytest.mark.parametrize("n", [1, 2, 3, 11, 17, 63, 64, 65, 127, 128, 129, 200, 513])
def test_pivots_match_lapack(n):
    a = O.gaussian(n, seed=1000 + n)
    lu, piv, info = O.getrf(a)
    slu, spiv = sl.lu_factor(a)
    assert info == 0
    assert np.array_equal(piv - 1, spiv)
    assert O.rel_diff(lu, slu) <= 1e-12
    assert np.abs(np.tril(lu, -1)).max(initial=0.0) <= 1.0


@pytest.mark.parametrize("n", [5, 64, 300])
def test_float32_pivots_match_lapack(n):
    a = O.gaussian(n, seed=2000 + n, dtype=np.float32)
    lu, piv, info = O.getrf(a)
    slu, spiv = sl.lu_factor(a)
    assert np.array_equal(piv - 1, spiv)
    assert O.rel_diff(lu, slu) <= 1e-5


def _exact_getrf(a):
    """Doolittle with partial pivoting in exact rationals (first maximum wins)."""
    n = len(a)
    m = [[Fraction(x) for x in row] for row in a]
    piv = []
    for i in range(n):
        p = max(range(i, n), key=lambda r: (abs(m[r][i]), -r))
        piv.append(p + 1)
        m[i], m[p] = m[p], m[i]
        if m[i][i] == 0:
            continue
        for r in range(i + 1, n):
            m[r][i] = m[r][i] / m[i][i]
            for c in range(i + 1, n):
                m[r][c] -= m[r][i] * m[i][c]
    return m, piv


@pytest.mark.parametrize("seed", range(20))
def test_brute_force_tiny(seed):
    rng = np.random.default_rng(seed)
    n = 1 + seed % 4
    a = rng.standard_normal((n, n))
    lu, piv, info = O.getrf(a)
    m, epiv = _exact_getrf(a.tolist())
    assert list(piv) == epiv
    exact = np.array([[float(x) for x in row] for row in m])
    assert np.abs(lu - exact).max() <= 64 * np.finfo(np.float64).eps * max(1.0, np.abs(exact).max())


def test_tie_break_lowest_row_wins():
    # |1| == |-1| == |1| in column 0: the first maximum (row 0) is kept, as LAPACK idamax does
    a = np.array([[1.0, 2.0, 0.0], [-1.0, 3.0, 1.0], [1.0, 0.5, 2.0]])
    lu, piv, info = O.getrf(a)
    assert piv[0] == 1
    # equal magnitudes below the diagonal: the lower index wins
    a = np.array([[0.5, 2.0, 0.0], [-3.0, 3.0, 1.0], [3.0, 0.5, 2.0]])
    lu, piv, info = O.getrf(a)
    assert piv[0] == 2


def test_nan_handling():
    # a NaN below the diagonal is never chosen; a NaN on the diagonal keeps the diagonal
    a = np.array([[1.0, 2.0], [np.nan, 3.0]])
    _, piv, _ = O.getrf(a)
    assert piv[0] == 1
    a = np.array([[np.nan, 2.0], [5.0, 3.0]])
    _, piv, _ = O.getrf(a)
    assert piv[0] == 1


def test_thread_count_invariance_and_replay():
    a = O.gaussian(700, seed=9)
    lu1, p1, _ = O.getrf(a, threads=1)
    lu8, p8, _ = O.getrf(a, threads=8)
    assert np.array_equal(lu1, lu8) and np.array_equal(p1, p8)
    lur, pr, info, colmax, chosen = O.getrf_forced(a, p1)
    assert np.array_equal(lur, lu1) and np.array_equal(pr, p1)
    assert np.array_equal(colmax, chosen)


def test_reconstruction_1024():
    a = O.gaussian(1024, seed=77)
    lu, piv, info = O.getrf(a, threads=8)
    assert info == 0
    assert O.reconstruction_error(a, lu, piv) < 1.0


def test_solve_known_solution_and_lapack():
    n = 400
    a = O.gaussian(n, seed=31)
    x_true = O.gaussian(n, 5, seed=32)
    b = a @ x_true
    lu, piv, x, info = O.gesv(a, b)
    assert info == 0
    assert O.backward_error(a, x, b) <= 4 * n * np.finfo(np.float64).eps
    kappa = np.linalg.cond(a, np.inf)
    assert O.forward_error(x, x_true) <= 10 * kappa * n * np.finfo(np.float64).eps
    xs = sl.lu_solve(sl.lu_factor(a), b)
    assert O.forward_error(x, xs) <= 10 * kappa * n * np.finfo(np.float64).eps


def test_multi_rhs_equals_single_rhs_bitwise():
    n = 150
    a = O.gaussian(n, seed=40)
    lu, piv, _ = O.getrf(a)
    b = O.gaussian(n, 7, seed=41)
    xm = O.getrs(lu, piv, b)
    for c in range(7):
        xs = O.getrs(lu, piv, b[:, c].copy())
        assert np.array_equal(xs, xm[:, c])


def test_singular_solve_leaves_x():
    a = np.array(_kats()["zero_pivot"][1]["a"], dtype=np.float64)
    b = np.array([1.0, 2.0, 3.0])
    lu, piv, x, info = O.gesv(a, b)
    assert info == 2
    assert np.array_equal(x, b)


def test_zero_column_info():
    a = O.gaussian(300, seed=5)
    a[:, 150] = 0.0
    lu, piv, info = O.getrf(a)
    assert info == 151
    slu, spiv, sinfo = sl.lapack.dgetrf(a)
    assert sinfo == 151
    assert np.array_equal(piv - 1, spiv)


def test_reference_workspace_closed_form():
    # lu_decomposition_native.cpp:293-326 with the unvectorised GEMM needing no scratch
    for n in [1, 64, 65, 100, 1024]:
        assert O.reference_workspace_size(n) == max(0, n - 64) ** 2 * 8


def test_planted_mistake_is_caught():
    """A wrong tie-break planted in a copy of the oracle's pivots fails the replay check."""
    a = np.array([[1.0, 2.0, 0.0], [-1.0, 3.0, 1.0], [0.25, 0.5, 2.0]])
    lu, piv, _ = O.getrf(a)
    bad = piv.copy()
    bad[0] = 2  # the other row of the tie: same magnitude, so the replay flags no growth...
    _, _, _, colmax, chosen = O.getrf_forced(a, bad)
    assert chosen[0] == colmax[0]  # ...a tie is legitimately within tolerance
    bad[0] = 3  # a non-maximal row is rejected by the replay check
    _, _, _, colmax, chosen = O.getrf_forced(a, bad)
    assert chosen[0] < colmax[0]

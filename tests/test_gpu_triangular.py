"""triangular_solver<T, ROCM_NATIVE> on the GPU against the reference's known answers and the
oracle."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _solve(upper, transpose, unit, n, nrhs, a, lda, b, ldb, ldx, dev):
    import torch
    import linearalgebra_amd as la
    ta = torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    tb = torch.from_numpy(np.ascontiguousarray(b)).to(dev)
    tx = torch.full((n, ldx), -123.0, dtype=ta.dtype, device=dev)
    la.triangular_solver.apply(upper, transpose, unit, n, nrhs, ta, lda, tb, ldb, tx, ldx)
    torch.cuda.synchronize()
    return ta.cpu().numpy(), tb.cpu().numpy(), tx.cpu().numpy()


@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-13), (np.float32, 4e-6)])
def test_triangular_known_answers(dtype, tol, dev):
    """The reference suite's bar (Testing/Cxx/TestTriangularSolver.cxx:55-78): finite results
    within 1e-13 (FP64) / 4e-6 (FP32) of the known solution, X's padding untouched, A (NaN
    outside its triangle) and B unchanged."""
    for upper, transpose, unit, n, nrhs, lda, ldb, ldx, a, b, expected in O.trsm_known_answer_cases(dtype):
        a_out, b_out, x = _solve(upper, transpose, unit, n, nrhs, a, lda, b, ldb, ldx, dev)
        err = np.abs(x[:, :nrhs].astype(np.float64) - expected[:, :nrhs]).max()
        case = (upper, transpose, unit, n, nrhs)
        assert np.isfinite(x[:, :nrhs]).all(), case
        assert err <= tol, (case, err)
        assert (x[:, nrhs:] == -123).all(), case
        assert np.array_equal(a_out, a, equal_nan=True), case
        assert np.array_equal(b_out, b), case


@pytest.mark.parametrize("n,nrhs", [(64, 1), (65, 64), (300, 70), (1000, 9)])
@pytest.mark.parametrize("upper,transpose,unit", [(False, False, True), (True, False, False), (False, True, False),
                                                  (True, True, True)])
def test_triangular_matches_oracle(n, nrhs, upper, transpose, unit, dev):
    rng = np.random.default_rng(n + 7 * nrhs)
    a = 0.2 * rng.standard_normal((n, n)) / np.sqrt(n) + np.diag(2.0 + rng.random(n))
    b = rng.standard_normal((n, nrhs))
    _, _, x = _solve(upper, transpose, unit, n, nrhs, a, n, b, nrhs, nrhs, dev)
    ref = O.trsm(upper, transpose, unit, a, b)
    err = np.abs(x - ref).max() / np.abs(ref).max()
    print(f"trsm n={n} nrhs={nrhs} upper={upper} transpose={transpose} unit={unit}: rel diff {err:.2e}", flush=True)
    assert err <= 1e-12


def test_triangular_invalid_arguments(dev):
    import torch
    import linearalgebra_amd as la
    a = torch.eye(4, dtype=torch.float64, device=dev)
    b = torch.ones(4, 2, dtype=torch.float64, device=dev)
    x = torch.empty(4, 2, dtype=torch.float64, device=dev)
    with pytest.raises(ValueError):
        la.triangular_solver.apply(False, False, False, 0, 2, a, 4, b, 2, x, 2)
    with pytest.raises(ValueError):
        la.triangular_solver.apply(False, False, False, 4, 2, a, 3, b, 2, x, 2)
    with pytest.raises(ValueError):
        la.triangular_solver.apply(False, False, False, 4, 2, a, 4, b, 2, b, 2)
    with pytest.raises(ValueError):
        la.triangular_solver.workspace_size(2, -1)


def _multi_rhs_solver(min_rhs):
    from linearalgebra_amd import _lib
    assert _lib.load().lrn_set_multi_rhs_solver(int(min_rhs)) == 0


@pytest.mark.parametrize("n,nrhs", [(130, 2), (1000, 64), (4100, 130)])
@pytest.mark.parametrize("upper,transpose,unit", [(False, False, True), (True, False, False), (False, True, False),
                                                  (True, True, True)])
def test_triangular_blocked_matches_chained_and_oracle(n, nrhs, upper, transpose, unit, dev):
    """The blocked form (inverted 64 x 64 diagonal blocks, matrix-core updates; its chained
    single-launch kernel and its launch-per-strip kernel; nrhs > 64 covers several column groups,
    n = 4100 a partial last strip) against the chained substitution and the oracle, on a
    well-conditioned triangle: all within 1e-12."""
    rng = np.random.default_rng(3 * n + nrhs)
    a = 0.2 * rng.standard_normal((n, n)) / np.sqrt(n) + np.diag(2.0 + rng.random(n))
    # NaN outside the triangle op(A) uses: neither form may read it
    lower_stored = (not upper)
    mask = np.tril(np.ones((n, n), dtype=bool)) if lower_stored else np.triu(np.ones((n, n), dtype=bool))
    a_nan = np.where(mask, a, np.nan)
    b = rng.standard_normal((n, nrhs))
    from linearalgebra_amd import _lib
    lib = _lib.load()
    try:
        _, _, xb = _solve(upper, transpose, unit, n, nrhs, a_nan, n, b, nrhs, nrhs + 3, dev)
        assert lib.lrn_set_multi_rhs_kernel(0) == 0  # one launch per strip
        _, _, xs = _solve(upper, transpose, unit, n, nrhs, a_nan, n, b, nrhs, nrhs + 3, dev)
        assert lib.lrn_set_multi_rhs_kernel(1) == 0
        _multi_rhs_solver(0)
        _, _, xc = _solve(upper, transpose, unit, n, nrhs, a_nan, n, b, nrhs, nrhs + 3, dev)
    finally:
        lib.lrn_set_multi_rhs_kernel(1)
        _multi_rhs_solver(2)
    assert (xb[:, nrhs:] == -123).all() and (xc[:, nrhs:] == -123).all() and (xs[:, nrhs:] == -123).all()
    xb, xc, xs = xb[:, :nrhs], xc[:, :nrhs], xs[:, :nrhs]
    assert np.isfinite(xb).all() and np.isfinite(xs).all()
    d_bc = np.abs(xb - xc).max() / np.abs(xc).max()
    d_sc = np.abs(xs - xc).max() / np.abs(xc).max()
    assert d_sc <= 1e-12
    msg = (f"trsm blocked n={n} nrhs={nrhs} upper={upper} transpose={transpose} unit={unit}: vs chained {d_bc:.2e} "
           f"(one launch per strip: {d_sc:.2e})")
    if n <= 1000:
        ref = O.trsm(upper, transpose, unit, a, b)
        d_o = np.abs(xb - ref).max() / np.abs(ref).max()
        msg += f", vs oracle {d_o:.2e}"
        assert d_o <= 1e-12
    print(msg, flush=True)
    assert d_bc <= 1e-12


@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-13), (np.float32, 4e-6)])
def test_triangular_known_answers_chained(dtype, tol, dev):
    """The reference suite's known answers through the chained kernels as well (the default
    sends every case with two or more right-hand sides through the blocked form)."""
    try:
        _multi_rhs_solver(0)
        test_triangular_known_answers(dtype, tol, dev)
    finally:
        _multi_rhs_solver(2)

"""Error paths of the bounded inter-workgroup waits, forced from the host: with the spin limit
at 0 every wait gives up at once (each kernel still runs to its end with in-range indices),
raises the error word, and the results must say so. The synchronous calls report
LRN_E_TIMEOUT; the asynchronous solvers leave NaN in x instead of a wrong solution; the
inversion does not copy its (invalid) result over m. The limit is restored after each test.
The many-right-hand-side solves run in each of their three forms: the chained substitution
and the blocked form's chained kernel wait between workgroups (NaN expected); the blocked form
with one launch per strip has no such wait and must solve correctly under the same limit. The
mixed refinement's single-RHS substitutions (each form) leave NaN as well."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture
def expired_waits(dev):
    from linearalgebra_amd import _lib
    lib = _lib.load()
    assert lib.lrn_debug_set_spin_limit(0) == 0
    try:
        yield lib
    finally:
        assert lib.lrn_debug_set_spin_limit(-1) == 0


def test_lu_decomposition_reports_timeout(expired_waits, dev):
    import torch
    import linearalgebra_amd as la
    n = 1024
    m = torch.from_numpy(O.gaussian(n, seed=1)).to(dev)
    piv = torch.zeros(n, dtype=torch.int64, device=dev)
    with pytest.raises(la.LinalgError, match="wait"):
        la.lu_decomposition.apply(m, n, piv)


@pytest.mark.parametrize("nrhs", [1, 3])
def test_solver_lu_mode_leaves_nan(expired_waits, nrhs, dev):
    import torch
    import linearalgebra_amd as la
    n = 1024
    m = torch.from_numpy(O.gaussian(n, seed=2)).to(dev)
    piv = torch.zeros(n, dtype=torch.int64, device=dev)
    b = O.gaussian(n, nrhs, seed=3)
    x = torch.from_numpy(b[:, 0].copy() if nrhs == 1 else b).to(dev)
    la.linear_solver.apply(m, piv, n, x, la.linear_solver_type.LU_LINEAR_SOLVER)
    torch.cuda.synchronize()
    xs = x.cpu().numpy()
    print(f"LU mode nrhs={nrhs}: NaN entries {np.isnan(xs).sum()} of {xs.size}", flush=True)
    assert np.isnan(xs).all()


# many-right-hand-side kernels: (lrn_set_multi_rhs_solver, lrn_set_multi_rhs_kernel, waits between workgroups)
MULTI_RHS = {"chained": (0, 1, True), "blocked_chain": (2, 1, True), "blocked_steps": (2, 0, False)}


def _multi_rhs(lib, mode):
    min_rhs, kind, _ = MULTI_RHS[mode]
    assert lib.lrn_set_multi_rhs_solver(min_rhs) == 0
    assert lib.lrn_set_multi_rhs_kernel(kind) == 0


def _multi_rhs_default(lib):
    assert lib.lrn_set_multi_rhs_solver(2) == 0
    assert lib.lrn_set_multi_rhs_kernel(1) == 0


@pytest.mark.parametrize("mode", list(MULTI_RHS))
def test_solver_upfront_substitution_leaves_nan(mode, dev):
    """Valid factors (computed with the normal limit), then a solve whose substitution waits
    expire: x is NaN, for one and five right-hand sides; the blocked form with one launch per
    strip, which has no wait between workgroups, solves five correctly under the same limit."""
    import torch
    import linearalgebra_amd as la
    from linearalgebra_amd import _lib
    lib = _lib.load()
    n = 1024
    a = O.gaussian(n, seed=4)
    lu, piv, _ = O.getrf(a)
    m = torch.from_numpy(lu).to(dev)
    p = torch.from_numpy(piv).to(dev)
    for nrhs in (1, 5):
        b = O.gaussian(n, nrhs, seed=5)
        x = torch.from_numpy(b[:, 0].copy() if nrhs == 1 else b).to(dev)
        assert lib.lrn_debug_set_spin_limit(0) == 0
        _multi_rhs(lib, mode)
        try:
            la.linear_solver.apply(m, p, n, x, la.linear_solver_type.LU_UPFRONT_LINEAR_SOLVER)
            torch.cuda.synchronize()
        finally:
            assert lib.lrn_debug_set_spin_limit(-1) == 0
            _multi_rhs_default(lib)
        ref = O.getrs(lu, piv, b[:, 0] if nrhs == 1 else b)
        if nrhs > 1 and not MULTI_RHS[mode][2]:
            assert O.forward_error(x.cpu().numpy(), ref) <= 1e-12
        else:
            assert np.isnan(x.cpu().numpy()).all()
        # and the same call with the normal limit solves correctly
        x = torch.from_numpy(b[:, 0].copy() if nrhs == 1 else b).to(dev)
        _multi_rhs(lib, mode)
        try:
            la.linear_solver.apply(m, p, n, x, la.linear_solver_type.LU_UPFRONT_LINEAR_SOLVER)
            torch.cuda.synchronize()
        finally:
            _multi_rhs_default(lib)
        assert O.forward_error(x.cpu().numpy(), ref) <= 1e-12


@pytest.mark.parametrize("mode", list(MULTI_RHS))
def test_triangular_and_cholesky_solves_leave_nan(expired_waits, mode, dev):
    """Two right-hand sides under the forced expiry: NaN where the kernel waits between
    workgroups; the blocked form with one launch per strip solves the triangle correctly."""
    import torch
    import linearalgebra_amd as la
    from oracle import oracle as OM
    lib = expired_waits
    waits = MULTI_RHS[mode][2]
    n = 300
    rng = np.random.default_rng(6)
    g = rng.standard_normal((n, n))
    spd = g @ g.T / n + np.eye(n)
    b = rng.standard_normal((n, 2))
    a_t = torch.from_numpy(np.tril(spd)).to(dev)
    x = torch.zeros(n, 2, dtype=torch.float64, device=dev)
    _multi_rhs(lib, mode)
    try:
        la.triangular_solver.apply(False, False, False, n, 2, a_t, n, torch.from_numpy(b).to(dev), 2, x, 2)
        torch.cuda.synchronize()
        if waits:
            assert np.isnan(x.cpu().numpy()).all()
        else:
            assert OM.rel_diff(x.cpu().numpy(), OM.trsm(False, False, False, np.tril(spd), b)) <= 1e-12
        m = torch.from_numpy(spd).to(dev)
        xc = torch.from_numpy(b.copy()).to(dev)
        la.linear_solver.apply(m, None, n, xc, la.linear_solver_type.CHOLESKY_LINEAR_SOLVER)
        torch.cuda.synchronize()
    finally:
        _multi_rhs_default(lib)
    if waits:
        assert np.isnan(xc.cpu().numpy()).all()


def test_inversion_and_mixed_report_timeout(expired_waits, dev):
    """The inversion reports the expired factorisation and leaves m holding that factorisation's
    output (what lu_decomposition leaves under the same forced expiry), not its own result; the
    mixed solver reports it too."""
    import torch
    import linearalgebra_amd as la
    n = 1024
    a = O.gaussian(n, seed=7)
    m_lu = torch.from_numpy(a).to(dev)
    piv = torch.zeros(n, dtype=torch.int64, device=dev)
    with pytest.raises(la.LinalgError):
        la.lu_decomposition.apply(m_lu, n, piv)
    m_inv = torch.from_numpy(a).to(dev)
    with pytest.raises(la.LinalgError):
        la.inversion.apply(m_inv, None, n, la.linear_solver_type.LU_LINEAR_SOLVER)
    # the invalid factors hold NaN, which must compare equal position by position
    assert np.array_equal(m_inv.cpu().numpy(), m_lu.cpu().numpy(), equal_nan=True)
    am = torch.from_numpy(O.conditioned(512, 1e3, seed=8)).to(dev)
    bm = torch.from_numpy(O.gaussian(512, 1, seed=9)[:, 0]).to(dev)
    with pytest.raises(la.LinalgError):
        la.gesv_mixed(am, bm)


@pytest.mark.parametrize("solver", [0, 3, 4])
def test_refinement_substitution_leaves_nan(expired_waits, solver, dev):
    """The mixed refinement's substitution kernels (the 64-row substitution and the inverse form
    in dispatch and XCD-local order) give up their waits at once under the forced expiry,
    raise the error word, finish, and the solve's result is poisoned with NaN."""
    import torch
    from linearalgebra_amd.mixed import DEFAULT_REFINE_SOLVER, getrs_f32f64, set_refine_solver
    n = 3000
    a = O.gaussian(n, seed=31).astype(np.float32)
    # any FP32 factors serve: a diagonally dominant matrix read as its own L and U, P = I
    a += np.float32(4 * n) * np.eye(n, dtype=np.float32)
    lu = torch.from_numpy(a).to(dev)
    piv = torch.arange(1, n + 1, dtype=torch.int64, device=dev)
    b = torch.from_numpy(O.gaussian(n, 1, seed=32)[:, 0]).to(dev)
    try:
        set_refine_solver(solver)
        getrs_f32f64(lu, piv, b)
        torch.cuda.synchronize()
    finally:
        set_refine_solver(DEFAULT_REFINE_SOLVER)
    assert np.isnan(b.cpu().numpy()).all()


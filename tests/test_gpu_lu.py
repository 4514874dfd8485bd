"""GPU parity of the plain LU path (lu_decomposition / linear_solver, ROCM_NATIVE) against the oracle.

Bar (plain FP64): identical pivot vectors except at ties; packed L\\U and solutions within
1e-12 normwise relative difference of the oracle where the problem's conditioning allows it
(diagonally dominant inputs, and N(0,1) inputs up to n = 1024); for N(0,1) inputs at larger
n the factors are compared with a bound that scales with n*eps (two backward-stable
factorisations with different rounding differ by O(n eps growth)).
"""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def _gpu_getrf(a, dev):
    import torch
    import linearalgebra_amd as la
    n = a.shape[0]
    m = torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    piv = torch.zeros(n, dtype=torch.int64, device=dev)
    info = []
    ok = la.lu_decomposition.apply(m, n, piv, info=info)
    return m.cpu().numpy(), piv.cpu().numpy(), info[0], ok


def _gpu_solve(a, b, dev, mode=1, lu=None, piv=None):
    import torch
    import linearalgebra_amd as la
    n = a.shape[0]
    m = torch.from_numpy(np.ascontiguousarray(a if lu is None else lu)).to(dev)
    p = torch.zeros(n, dtype=torch.int64, device=dev) if piv is None else torch.from_numpy(piv).to(dev)
    x = torch.from_numpy(np.ascontiguousarray(b)).to(dev)
    la.linear_solver.apply(m, p, n, x, la.linear_solver_type(mode))
    torch.cuda.synchronize()
    return m.cpu().numpy(), p.cpu().numpy(), x.cpu().numpy()


def _kappa_inf(a):
    """LAPACK's infinity-norm condition estimate (dgecon on scipy's getrf factors)."""
    import scipy.linalg as sl
    from scipy.linalg import lapack
    lu, _ = sl.lu_factor(a, check_finite=False)
    anorm = np.abs(a).sum(axis=1).max()
    rcond, info = lapack.dgecon(lu, anorm, norm="I")
    assert info == 0
    return 1.0 / rcond if rcond > 0 else np.inf


SMALL = [1, 2, 3, 11, 17, 63, 64, 65, 127, 128, 129, 200, 255, 256, 257, 300, 511, 700]


@pytest.mark.parametrize("n", SMALL)
def test_getrf_f64_small_matches_oracle(n, dev):
    a = O.gaussian(n, seed=100 + n)
    lu_g, piv_g, info_g, ok = _gpu_getrf(a, dev)
    lu_o, piv_o, info_o = O.getrf(a)
    print(f"n={n}", flush=True)
    assert info_g == info_o == 0 and ok
    assert np.array_equal(piv_g, piv_o), f"first mismatch at {O.first_pivot_mismatch(piv_g, piv_o)}"
    assert O.rel_diff(lu_g, lu_o) <= 1e-12
    assert O.reconstruction_error(a, lu_g, piv_g) < 1.0


@pytest.mark.parametrize("n", [1024, 2048])
def test_getrf_f64_gaussian_medium(n, dev):
    a = O.gaussian(n, seed=7 + n)
    lu_g, piv_g, info_g, ok = _gpu_getrf(a, dev)
    lu_o, piv_o, info_o = O.getrf(a, threads=16)
    mis = O.first_pivot_mismatch(piv_g, piv_o)
    rd = O.rel_diff(lu_g, lu_o)
    print(f"n={n} first_mismatch={mis} rel_diff={rd:.3e}", flush=True)
    assert info_g == 0 and ok
    assert mis == -1
    assert rd <= 1e-12
    assert O.reconstruction_error(a, lu_g, piv_g) < 1.0


@pytest.mark.parametrize("n", [1000, 4096])
def test_getrf_f64_diag_dominant(n, dev):
    a = O.diag_dominant(n, seed=3)
    lu_g, piv_g, info_g, ok = _gpu_getrf(a, dev)
    lu_o, piv_o, info_o = O.getrf(a, threads=16)
    rd = O.rel_diff(lu_g, lu_o)
    print(f"n={n} rel_diff={rd:.3e}", flush=True)
    assert np.array_equal(piv_g, piv_o)
    assert rd <= 1e-12


def test_getrf_zero_pivot_kats(dev):
    # reference KATs, Testing/Cxx/TestLUDecomposition.cxx:283-299
    for rows, expect in (([[0, 1, 2], [0, 3, 4], [0, 5, 7]], 1), ([[1, 2, 3], [2, 4, 5], [4, 8, 1]], 2),
                         ([[0, 0], [0, 0]], 1)):
        a = np.array(rows, dtype=np.float64)
        lu_g, piv_g, info_g, ok = _gpu_getrf(a, dev)
        lu_o, piv_o, info_o = O.getrf(a)
        assert info_g == expect == info_o and not ok
        assert np.array_equal(piv_g, piv_o)
        lo, up = O.unpack_lu(lu_g)
        perm = O.pivots_to_perm(piv_g)
        assert np.abs(a[perm] - lo @ up).max() <= 16 * np.finfo(np.float64).eps


def test_getrf_small_scale_and_subnormal(dev):
    # Testing/Cxx/TestLUDecomposition.cxx:305-342
    a = np.array([[1e-8, 0], [0, 1e-8]])
    lu_g, piv_g, info_g, ok = _gpu_getrf(a, dev)
    assert info_g == 0 and ok
    lo, up = O.unpack_lu(lu_g)
    assert np.abs(a[O.pivots_to_perm(piv_g)] - lo @ up).max() <= 1e-8 * np.finfo(np.float64).eps
    t = np.ldexp(1.0, np.finfo(np.float64).minexp - 4)
    a = np.array([[t, 1.0], [t / 2, 1.0]])
    lu_g, piv_g, info_g, ok = _gpu_getrf(a, dev)
    assert info_g == 0 and ok and np.isfinite(lu_g).all()
    lo, up = O.unpack_lu(lu_g)
    assert np.abs(a[O.pivots_to_perm(piv_g)] - lo @ up).max() < np.finfo(np.float64).tiny


def test_getrf_singular_larger(dev):
    # an exactly zero column in the middle of a 300 x 300 matrix: info names it, factor completes
    a = O.gaussian(300, seed=5)
    a[:, 150] = 0.0
    lu_g, piv_g, info_g, ok = _gpu_getrf(a, dev)
    lu_o, piv_o, info_o = O.getrf(a)
    assert not ok and info_g == info_o == 151
    assert np.array_equal(piv_g, piv_o)
    assert O.rel_diff(lu_g, lu_o) <= 1e-12


@pytest.mark.parametrize("n,nrhs", [(1, 1), (65, 1), (300, 3), (300, 65), (1024, 16), (1500, 64), (777, 70), (4100, 1)])
def test_solver_lu_mode_matches_oracle(n, nrhs, dev):
    a = O.gaussian(n, seed=11 + n)
    b = O.gaussian(n, nrhs, seed=12 + n)
    b_in = b[:, 0].copy() if nrhs == 1 else b
    lu_g, piv_g, x_g = _gpu_solve(a, b_in, dev, mode=1)
    lu_o, piv_o, x_o, info_o = O.gesv(a, b_in, threads=16)
    fe = O.forward_error(x_g, x_o)
    be = O.backward_error(a, x_g, b_in)
    be_o = O.backward_error(a, x_o, b_in)
    # two backward-stable solutions differ by at most about kappa(A) (berr_1 + berr_2): the bar
    # is 1e-12 where the conditioning allows it, that bound (with a factor 10) beyond
    kappa = _kappa_inf(a)
    bar = max(1e-12, 10.0 * kappa * (be + be_o))
    print(f"n={n} nrhs={nrhs} fwd_vs_oracle={fe:.3e} berr={be:.3e} kappa_inf~{kappa:.2e} bar={bar:.2e}", flush=True)
    assert np.array_equal(piv_g, piv_o)
    assert be <= 4 * n * np.finfo(np.float64).eps
    assert fe <= bar


def test_solver_upfront_reuses_factors(dev):
    n = 513
    a = O.diag_dominant(n, seed=21)
    b = O.gaussian(n, seed=22).ravel()[:n].copy()
    lu_o, piv_o, info_o = O.getrf(a)
    _, _, x_g = _gpu_solve(a, b, dev, mode=2, lu=lu_o, piv=piv_o)
    x_o = O.getrs(lu_o, piv_o, b)
    assert O.forward_error(x_g, x_o) <= 1e-12


def test_solver_skips_singular(dev):
    a = np.array([[1.0, 2.0, 3.0], [2.0, 4.0, 5.0], [4.0, 8.0, 1.0]])
    b = np.array([1.0, 2.0, 3.0])
    _, _, x_g = _gpu_solve(a, b, dev, mode=1)
    assert np.array_equal(x_g, b)


@pytest.mark.parametrize("n", [17, 200, 1024])
def test_getrf_f32_matches_oracle(n, dev):
    a = O.gaussian(n, seed=300 + n, dtype=np.float32)
    lu_g, piv_g, info_g, ok = _gpu_getrf(a, dev)
    lu_o, piv_o, info_o = O.getrf(a)
    rd = O.rel_diff(lu_g, lu_o)
    print(f"n={n} f32 rel_diff={rd:.3e}", flush=True)
    assert ok and info_o == 0
    mis = O.first_pivot_mismatch(piv_g, piv_o)
    assert mis == -1 or mis > n // 2
    assert O.reconstruction_error(a, lu_g, piv_g) < 1.0


@pytest.mark.parametrize("n", [300, 1000, 4100])
@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_panel_kernels_agree(n, dtype, dev):
    """The register-resident and the LDS-resident 32-column panel kernels implement the same
    elimination order, so pivots and factors are bitwise equal; both match the oracle."""
    import torch
    import linearalgebra_amd as la
    from linearalgebra_amd import _lib
    lib = _lib.load()
    a = O.gaussian(n, seed=900 + n)
    if dtype == "f32":
        a = a.astype(np.float32)
    out = {}
    try:
        for kind in (0, 1):
            assert lib.lrn_set_panel_kernel(kind) == 0
            out[kind] = _gpu_getrf(a, dev)
    finally:
        lib.lrn_set_panel_kernel(0)
    assert np.array_equal(out[0][1], out[1][1])
    assert np.array_equal(out[0][0], out[1][0])
    lu_o, piv_o, _ = O.getrf(a)
    mis = O.first_pivot_mismatch(out[0][1], piv_o)
    print(f"panel kernels n={n} {dtype}: first pivot mismatch vs oracle {mis}", flush=True)
    if n <= 1000:
        assert mis == -1


@pytest.mark.parametrize("outer_nb", [128, 384, 512])
def test_outer_block_widths_match_oracle(outer_nb, dev):
    """Block widths the autotuner may pick (fused block-column kernels and their fallbacks)."""
    from linearalgebra_amd import _lib
    lib = _lib.load()
    n = 1100
    a = O.gaussian(n, seed=77)
    try:
        assert lib.lrn_set_tuning(outer_nb, 32, 0) == 0
        lu_g, piv_g, info_g, ok = _gpu_getrf(a, dev)
    finally:
        lib.lrn_set_tuning(0, 0, 0)
        from linearalgebra_amd import autotune
        autotune.apply(lib)
    lu_o, piv_o, info_o = O.getrf(a)
    assert info_g == info_o == 0 and ok
    assert np.array_equal(piv_g, piv_o), f"first mismatch at {O.first_pivot_mismatch(piv_g, piv_o)}"
    assert O.rel_diff(lu_g, lu_o) <= 1e-12


@pytest.mark.parametrize("mode", [16, 32, 64])
@pytest.mark.parametrize("dtype", ["f32", "f64"])
def test_lookahead_panel_units_at_n16384(mode, dtype, dev):
    """Every panel-stream size the autotuner tries completes the n = 16384 factorisation: the
    panel stream keeps one compute unit per XCD beyond what the largest panel's cooperative
    workgroups need, since the narrow stream's kernels run beside the panels on those units
    (an exact fit, 64 workgroups on 16 units, timed out its bounded wait). The factors are
    checked through P A = L U on a sample of rows."""
    import torch
    import linearalgebra_amd as la
    from linearalgebra_amd import _lib, autotune
    lib = _lib.load()
    n = 16384
    dt = torch.float32 if dtype == "f32" else torch.float64
    g = torch.Generator(device=dev)
    g.manual_seed(11)
    a = torch.randn(n, n, dtype=dt, device=dev, generator=g)
    w = a.clone()
    piv = torch.empty(n, dtype=torch.int64, device=dev)
    ws = torch.empty(la.lu_decomposition.workspace_size(n, dt), dtype=torch.uint8, device=dev)
    try:
        assert lib.lrn_set_lookahead(mode) == 0
        la.lu_decomposition.apply(w, n, piv, ws)
        torch.cuda.synchronize()
    finally:
        lib.lrn_set_lookahead(0)
        autotune.apply(lib)
    p = piv.cpu().numpy()
    assert np.all(p >= np.arange(1, n + 1)) and np.all(p <= n)
    perm = np.arange(n)
    for i in range(n):
        j = int(p[i]) - 1
        perm[i], perm[j] = perm[j], perm[i]
    rows = torch.arange(0, n, 97, device=dev)
    lu = w.double()
    cols = torch.arange(n, device=dev)
    L = lu[rows] * (cols[None, :] < rows[:, None])  # strictly lower part of the sampled rows
    L[torch.arange(rows.numel(), device=dev), rows] = 1.0
    rec = L @ torch.triu(lu)
    ref = a.double()[torch.from_numpy(perm).to(dev)[rows]]
    rel = ((rec - ref).abs().max() / a.abs().max()).item()
    print(f"lookahead={mode} {dtype} n={n}: sampled |PA - LU| / max|A| = {rel:.2e}", flush=True)
    assert rel <= (1e-3 if dtype == "f32" else 1e-10)


@pytest.mark.parametrize("kind", [None, 0, 1])
@pytest.mark.parametrize("n,dtype", [(300, "f64"), (1100, "f64"), (2100, "f64"), (2100, "f32"), (2600, "f32")])
def test_lookahead_matches_plain_order(n, dtype, kind, dev):
    """The look-ahead only reorders launches over three streams (the next block column's
    inner-panel updates and the block column after it beside the panels, the far trailing
    update beside both); every entry sees the same operations in the same order, so factors,
    pivots and the solution are bitwise those of the single-stream order, which matches the
    oracle. kind: the block-column update kernel (None: the tuned one)."""
    from linearalgebra_amd import _lib
    lib = _lib.load()
    a = O.gaussian(n, seed=300 + n)
    if dtype == "f32":
        a = a.astype(np.float32)
    b = O.gaussian(n, 1, seed=301 + n)[:, 0].astype(a.dtype)
    out = {}
    try:
        if kind is not None:
            assert lib.lrn_set_block_update(kind) == 0
        for mode in (0, 1):
            assert lib.lrn_set_lookahead(mode) == 0
            out[mode] = (_gpu_getrf(a, dev), _gpu_solve(a, b, dev, mode=1))
    finally:
        from linearalgebra_amd import autotune
        lib.lrn_set_lookahead(0)
        lib.lrn_set_block_update(0)
        autotune.apply(lib)
    (lu0, piv0, info0, ok0), (_, _, x0) = out[0]
    (lu1, piv1, info1, ok1), (_, _, x1) = out[1]
    assert ok1 and info1 == info0 == 0
    assert np.array_equal(piv1, piv0)
    assert np.array_equal(lu1, lu0)
    assert np.array_equal(x1, x0)
    if dtype == "f64":
        lu_o, piv_o, _ = O.getrf(a)
        assert np.array_equal(piv1, piv_o), f"first mismatch at {O.first_pivot_mismatch(piv1, piv_o)}"
        assert O.rel_diff(lu1, lu_o) <= 1e-12


@pytest.mark.parametrize("kind", [0, 1, 2])
def test_block_update_kernels_match_oracle(kind, dev):
    """Every block-column update kernel (vector tiles, matrix-core GEMM with 128 x 128 or
    64 x 64 tiles) gives the oracle's pivots and factors within rounding, with and without
    the look-ahead."""
    from linearalgebra_amd import _lib, autotune
    lib = _lib.load()
    n = 1500
    a = O.gaussian(n, seed=55)
    lu_o, piv_o, _ = O.getrf(a)
    try:
        assert lib.lrn_set_block_update(kind) == 0
        for mode in (0, 16):
            assert lib.lrn_set_lookahead(mode) == 0
            lu_g, piv_g, info_g, ok = _gpu_getrf(a, dev)
            assert ok and info_g == 0
            assert np.array_equal(piv_g, piv_o), f"first mismatch at {O.first_pivot_mismatch(piv_g, piv_o)}"
            assert O.rel_diff(lu_g, lu_o) <= 1e-12
    finally:
        lib.lrn_set_block_update(0)
        lib.lrn_set_lookahead(0)
        autotune.apply(lib)



@pytest.mark.parametrize("n,nrhs", [(2100, 64), (1000, 130)])
def test_solver_multi_rhs_blocked_matches_chained(n, nrhs, dev):
    """LU solve with many right-hand sides: the blocked form (P B gathered, L and U applied
    strip by strip with their inverted 64 x 64 diagonal blocks) against the chained
    substitution kernels and the oracle, on a diagonally dominant matrix (kappa ~ 1): all within
    1e-12."""
    from linearalgebra_amd import _lib
    a = O.diag_dominant(n, seed=40 + n)
    b = O.gaussian(n, nrhs, seed=41 + n)
    lib = _lib.load()
    try:
        _, piv_b, x_b = _gpu_solve(a, b, dev, mode=1)
        assert lib.lrn_set_multi_rhs_kernel(0) == 0  # one launch per strip
        _, piv_s, x_s = _gpu_solve(a, b, dev, mode=1)
        assert lib.lrn_set_multi_rhs_kernel(1) == 0
        assert lib.lrn_set_multi_rhs_solver(0) == 0
        _, piv_c, x_c = _gpu_solve(a, b, dev, mode=1)
    finally:
        lib.lrn_set_multi_rhs_kernel(1)
        lib.lrn_set_multi_rhs_solver(2)
    lu_o, piv_o, x_o, info_o = O.gesv(a, b, threads=16)
    d_bc, d_sc, d_o = O.forward_error(x_b, x_c), O.forward_error(x_s, x_c), O.forward_error(x_b, x_o)
    print(f"multi-rhs n={n} nrhs={nrhs}: blocked vs chained {d_bc:.2e} (one launch per strip {d_sc:.2e}), "
          f"vs oracle {d_o:.2e}", flush=True)
    assert np.array_equal(piv_b, piv_o) and np.array_equal(piv_c, piv_o) and np.array_equal(piv_s, piv_o)
    assert d_bc <= 1e-12 and d_sc <= 1e-12 and d_o <= 1e-12

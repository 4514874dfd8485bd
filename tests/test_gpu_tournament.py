"""FP32 LU with tournament pivoting (the factorisation inside the mixed solver) on the GPU.

The pivots of a tournament need not equal partial pivoting's, so the checks are those that
fix an LU factorisation with row interchanges whatever the pivot rule:
  * every pivot is a valid sequential interchange (k+1 <= pivot[k] <= n);
  * P A = L U to FP32 accuracy (reconstruction evaluated in FP64 on the GPU with torch);
  * the multipliers stay bounded (|L| <= 1 when one wave holds the whole panel, where the
    tournament reduces to partial pivoting; a small bound otherwise);
  * orders n <= 256 (one leaf per panel) give partial pivoting's pivots, checked against the
    oracle's FP32 factorisation on inputs without near-ties;
  * an exactly singular matrix reports info = first zero pivot;
  * the mixed solver reaches FP64 backward error with the tournament panel as with the
    partial-pivoting one.
"""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def _factor(a32, dev):
    import torch
    from linearalgebra_amd.mixed import getrf_tournament
    t = torch.from_numpy(a32).to(dev)
    piv, info = getrf_tournament(t)
    torch.cuda.synchronize()
    return t, piv, info


def _reconstruction(a32, lu, piv):
    """max |P A - L U| / (n * max|A|), evaluated in FP64 on the GPU; and max |L|."""
    import torch
    n = a32.shape[0]
    pa = torch.from_numpy(a32.astype(np.float64)).to(lu.device)
    p = piv.cpu().numpy()
    perm = np.arange(n)
    for i in range(n):
        j = int(p[i]) - 1
        perm[i], perm[j] = perm[j], perm[i]
    pa = pa[torch.from_numpy(perm).to(lu.device)]
    lu64 = lu.double()
    L = torch.tril(lu64, -1) + torch.eye(n, dtype=torch.float64, device=lu.device)
    U = torch.triu(lu64)
    r = (pa - L @ U).abs().max().item()
    return r / (n * max(np.abs(a32).max(), 1e-300)), torch.tril(lu64, -1).abs().max().item() if n > 1 else 0.0


@pytest.mark.parametrize("n", [1, 2, 31, 32, 33, 100, 256, 257, 1000, 1024, 1025, 2048, 4100, 9000])
def test_tournament_factorisation(n, dev):
    rng = np.random.default_rng(n + 17)
    a = rng.standard_normal((n, n)).astype(np.float32)
    lu, piv, info = _factor(a, dev)
    print(f"tournament n={n} info={info}", flush=True)
    assert info == 0
    p = piv.cpu().numpy()
    assert np.all(p >= np.arange(1, n + 1)) and np.all(p <= n)
    rel, lmax = _reconstruction(a, lu, piv)
    print(f"  reconstruction {rel:.2e} max|L| {lmax:.3f}", flush=True)
    assert rel <= 2e-6
    if n <= 256:
        assert lmax <= 1.0 + 1e-6  # one leaf per panel: partial pivoting
    else:
        assert lmax <= 8.0


@pytest.mark.parametrize("n", [17000])
def test_tournament_deep_tree(n, dev):
    """Orders above 16384 rows per panel need more than one round of the reduction tree."""
    rng = np.random.default_rng(5)
    a = rng.standard_normal((n, n)).astype(np.float32)
    lu, piv, info = _factor(a, dev)
    assert info == 0
    rel, lmax = _reconstruction(a, lu, piv)
    print(f"tournament n={n} reconstruction {rel:.2e} max|L| {lmax:.3f}", flush=True)
    assert rel <= 2e-6 and lmax <= 8.0


@pytest.mark.parametrize("n", [5, 64, 200, 256])
def test_tournament_matches_partial_pivoting_small(n, dev):
    # N(0,1) entries: a near-tie between column maxima at these orders is very unlikely
    rng = np.random.default_rng(n)
    a32 = rng.standard_normal((n, n)).astype(np.float32)
    lu, piv, info = _factor(a32, dev)
    _, piv_ref, info_ref = O.getrf(a32.copy())
    p = piv.cpu().numpy()
    first = O.first_pivot_mismatch(p, piv_ref)
    assert info == info_ref == 0
    assert first == -1, f"pivot mismatch at {first}"
    ref = O.getrf(a32.copy())[0]
    assert np.abs(lu.cpu().numpy() - ref).max() <= 1e-4 * max(1.0, np.abs(ref).max())


def test_tournament_singular(dev):
    n = 300
    rng = np.random.default_rng(3)
    a = rng.standard_normal((n, n)).astype(np.float32)
    a[:, 40] = 0.0  # column 40 exactly zero: the first zero pivot is at step 40
    _, _, info = _factor(a, dev)
    assert info == 41


@pytest.mark.parametrize("kappa", [1e2, 1e6])
def test_mixed_solver_tournament_vs_partial(kappa, dev):
    import torch
    import linearalgebra_amd as la
    from linearalgebra_amd.mixed import set_mixed_panel
    n = 2000
    a = O.conditioned(n, kappa, seed=21)
    b = O.gaussian(n, 1, seed=22)[:, 0]
    ta, tb = torch.from_numpy(a).to(dev), torch.from_numpy(b).to(dev)
    out = {}
    try:
        for tour in (True, False):
            set_mixed_panel(tour)
            x, st = la.gesv_mixed(ta, tb)
            out[tour] = (O.backward_error(a, x.cpu().numpy(), b), st)
    finally:
        set_mixed_panel(False)
    for tour, (berr, st) in out.items():
        print(f"kappa={kappa:.0e} tournament={tour} berr={berr:.2e} path={st['path_name']} ir={st['ir_iterations']}",
              flush=True)
        assert berr <= 1e-14
        assert st["path_name"] == "lu-ir"
    assert out[True][1]["ir_iterations"] <= out[False][1]["ir_iterations"] + 3


@pytest.mark.parametrize("chain", [0, 1, 2, 3, 4])
@pytest.mark.parametrize("n", [1, 63, 64, 65, 127, 128, 129, 255, 256, 257, 700, 1000, 3000])
def test_refine_substitution(n, chain, dev):
    """The refinement's triangular solves (FP32 factors, FP64 arithmetic) against the same
    solves done in FP64 on the host with the factors rounded up exactly."""
    import scipy.linalg as sl
    import torch
    from linearalgebra_amd.mixed import DEFAULT_REFINE_SOLVER, getrs_f32f64, set_refine_solver
    rng = np.random.default_rng(n + 1000)
    a = rng.standard_normal((n, n)).astype(np.float32)
    lu, piv, info = _factor(a, dev)
    assert info == 0
    b = rng.standard_normal(n)
    tb = torch.from_numpy(b.copy()).to(dev)
    try:
        set_refine_solver(chain)
        getrs_f32f64(lu, piv, tb)
        torch.cuda.synchronize()
    finally:
        set_refine_solver(DEFAULT_REFINE_SOLVER)
    f = lu.cpu().numpy().astype(np.float64)
    perm = O.pivots_to_perm(piv.cpu().numpy())
    y = sl.solve_triangular(f, b[perm], lower=True, unit_diagonal=True)
    x = sl.solve_triangular(f, y, lower=False)
    got = tb.cpu().numpy()
    err = np.abs(got - x).max() / np.abs(x).max()
    print(f"refine solve n={n} chain={chain} rel err {err:.2e}", flush=True)
    assert err <= 1e-11

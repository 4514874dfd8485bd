"""Mixed-precision solve (FP32 panels, bf16 trailing updates, FP64 refinement) on the GPU.

Checks: the bf16 trailing-update GEMM against an FP64 evaluation of the same bf16-rounded
operands (within FP32 accumulation error); the solve's normwise backward error reaches the
requested 1e-14 for kappa in {1e2, 1e6, 1e10}; its forward error against the oracle's FP64
solution stays within kappa * n * eps.
"""
import ctypes

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def _bf16(x):
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    rounded = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return rounded.astype(np.uint32).view(np.float32).astype(np.float64)


def _planes(x, k):
    """x split into k bf16 planes (round to nearest even), as float64 arrays."""
    out, rem = [], x.astype(np.float32)
    for _ in range(k):
        p = _bf16(rem)
        out.append(p)
        rem = (rem.astype(np.float64) - p).astype(np.float32)
    return out


_PAIRS = {1: [(0, 0)], 2: [(1, 0), (0, 1), (0, 0)], 3: [(2, 0), (0, 2), (1, 1), (1, 0), (0, 1), (0, 0)]}


@pytest.mark.parametrize("planes", [1, 2, 3])
@pytest.mark.parametrize("M,N,K", [(16, 16, 32), (130, 70, 40), (256, 256, 256), (1000, 777, 128)])
def test_gemm_bf16(M, N, K, planes, dev):
    import torch
    from linearalgebra_amd import _lib
    vp = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731
    rng = np.random.default_rng(M * 7 + N + K)
    a = rng.standard_normal((M, K)).astype(np.float32)
    b = rng.standard_normal((K, N)).astype(np.float32)
    c = rng.standard_normal((M, N)).astype(np.float32)
    ta, tb, tc = (torch.from_numpy(v).to(dev) for v in (a, b, c))
    assert _lib.load().lrn_debug_gemm_sub(1 + planes, vp(ta), K, vp(tb), N, vp(tc), N, M, N, K, None) == 0
    torch.cuda.synchronize()
    pa, pb = _planes(a, planes), _planes(b, planes)
    ref = c.astype(np.float64) - sum(pa[i] @ pb[j] for i, j in _PAIRS[planes])
    got = tc.cpu().numpy()
    err = np.abs(got - ref).max() / np.abs(ref).max()
    # against the exact product of the FP32 operands: ~2^-8, 2^-16, 2^-23 relative
    full = c.astype(np.float64) - a.astype(np.float64) @ b.astype(np.float64)
    ferr = np.abs(got - full).max() / np.abs(full).max()
    print(f"gemm bf16 planes={planes} {M}x{N}x{K} err={err:.2e} vs exact {ferr:.2e}", flush=True)
    assert err <= 2e-6
    assert ferr <= {1: 3e-2, 2: 1e-4, 3: 2e-6}[planes]


@pytest.mark.parametrize("n,kappa", [(300, 1e2), (1000, 1e6), (2048, 1e2), (2048, 1e6), (2048, 1e10)])
def test_gesv_mixed_conditioned(n, kappa, dev):
    import torch
    import linearalgebra_amd as la
    a = O.conditioned(n, kappa, seed=n)
    b = O.gaussian(n, 1, seed=n + 5).ravel()
    ta = torch.from_numpy(a).to(dev)
    tb = torch.from_numpy(b).to(dev)
    x, st = la.gesv_mixed(ta, tb, tol=1e-14, max_iter=30, restart=40)
    xg = x.cpu().numpy()
    berr = O.backward_error(a, xg, b)
    lu, piv, x_ref, info = O.gesv(a, b, threads=16)
    fe = O.forward_error(xg, x_ref)
    print(f"n={n} kappa={kappa:.0e} path={st['path_name']} ir={st['ir_iterations']} gmres={st['gmres_iterations']}"
          f" berr={berr:.2e} (lib {st['backward_error']:.2e}) fwd_vs_oracle={fe:.2e}", flush=True)
    assert np.array_equal(ta.cpu().numpy(), a)  # A is not modified
    assert berr <= 2e-14
    # kappa <= 1e6 must converge by refinement alone; the FP64 fallback is allowed beyond
    assert st["path"] in ((0, 1) if kappa <= 1e6 else (0, 1, 2))
    assert fe <= max(1e-12, kappa * n * np.finfo(np.float64).eps)


@pytest.mark.parametrize("kappa", [1e2, 1e6])
@pytest.mark.parametrize("solver", [3, 4])
def test_gesv_mixed_refine_solvers(solver, kappa, dev):
    """The refinement's substitution in the inverse form (modes 3, 4) reaches the FP64 backward
    error by LU-IR like the 64-row substitution (mode 0), in about as many steps."""
    import torch
    import linearalgebra_amd as la
    from linearalgebra_amd import mixed as lam
    n = 3000
    a = O.conditioned(n, kappa, seed=91)
    b = O.gaussian(n, 1, seed=92).ravel()
    ta, tb = torch.from_numpy(a).to(dev), torch.from_numpy(b).to(dev)
    res = {}
    try:
        for mode in (0, solver):
            lam.set_refine_solver(mode)
            x, st = la.gesv_mixed(ta, tb, tol=1e-14, max_iter=30, restart=40)
            res[mode] = (O.backward_error(a, x.cpu().numpy(), b), st)
    finally:
        lam.set_refine_solver(lam.DEFAULT_REFINE_SOLVER)
    for mode, (berr, st) in res.items():
        print(f"refine solver {mode} kappa={kappa:.0e} path={st['path_name']} ir={st['ir_iterations']} "
              f"gmres={st['gmres_iterations']} berr={berr:.2e}", flush=True)
        assert berr <= 2e-14
        assert st["path"] == 0
    assert res[solver][1]["ir_iterations"] <= res[0][1]["ir_iterations"] + 1


@pytest.mark.parametrize("mode", [0, 1, 2, 3])
def test_gesv_mixed_update_modes(mode, dev):
    """Every trailing-update mode (1, 2, 3 bf16 planes; FP32 MFMA) reaches the FP64 backward
    error by refinement, without the FP64 fallback."""
    import torch
    import linearalgebra_amd as la
    from linearalgebra_amd import mixed as lam
    n = 1500
    a = O.conditioned(n, 1e3, seed=77)
    b = O.gaussian(n, 1, seed=78).ravel()
    lam.set_mixed_update(mode)
    try:
        x, st = la.gesv_mixed(torch.from_numpy(a).to(dev), torch.from_numpy(b).to(dev), tol=1e-14, max_iter=30,
                              restart=40)
    finally:
        lam.set_mixed_update(3)
    berr = O.backward_error(a, x.cpu().numpy(), b)
    print(f"mode={mode} path={st['path_name']} ir={st['ir_iterations']} gmres={st['gmres_iterations']} "
          f"berr={berr:.2e}", flush=True)
    assert berr <= 2e-14
    assert st["path"] in (0, 1)


def test_gesv_mixed_diag_dominant_lu_ir(dev):
    import torch
    import linearalgebra_amd as la
    n = 1024
    a = O.diag_dominant(n, seed=4)
    b = O.gaussian(n, 1, seed=9).ravel()
    x, st = la.gesv_mixed(torch.from_numpy(a).to(dev), torch.from_numpy(b).to(dev))
    assert st["path"] == 0
    assert O.backward_error(a, x.cpu().numpy(), b) <= 2e-14


def test_gesv_mixed_batched_uneven(dev):
    """Many systems of uneven order and condition in one call, each with its own update
    mode (three block-scaled FP8 planes where kappa <= 1e4, FP32 beyond): every backward error
    reaches the FP64 level and every forward error stays within kappa * n * eps of the oracle."""
    import torch
    import linearalgebra_amd as la
    from linearalgebra_amd import mixed as lam
    rng = np.random.Generator(np.random.PCG64(44))
    count = 10
    n = np.exp(rng.uniform(np.log(64), np.log(1024), size=count)).astype(np.int64)
    kappa = np.exp(rng.uniform(np.log(1e1), np.log(1e8), size=count))
    a = [O.conditioned(int(k), float(c), seed=700 + i) for i, (k, c) in enumerate(zip(n, kappa))]
    b = [O.gaussian(int(k), 1, seed=800 + i).ravel() for i, k in enumerate(n)]
    modes = np.where(kappa <= 1e4, 6, 3).astype(np.int32)
    ta = torch.from_numpy(np.concatenate([x.ravel() for x in a])).to(dev)
    tb = torch.from_numpy(np.concatenate(b)).to(dev)
    x, stats = lam.gesv_mixed_batched(ta, n, tb, modes=modes, tol=1e-14, max_iter=30, restart=40)
    xs = x.cpu().numpy()
    off = 0
    for i in range(count):
        k = int(n[i])
        xi = xs[off: off + k]
        off += k
        berr = O.backward_error(a[i], xi, b[i])
        _, _, x_ref, _ = O.gesv(a[i], b[i])
        fe = O.forward_error(xi, x_ref)
        st = stats[i]
        print(f"n={k} kappa={kappa[i]:.1e} mode={modes[i]} path={st['path_name']} ir={st['ir_iterations']} "
              f"gmres={st['gmres_iterations']} berr={berr:.2e} fwd={fe:.2e}", flush=True)
        assert berr <= 2e-14
        assert st["path"] in (0, 1, 2)
        assert fe <= max(1e-12, kappa[i] * k * np.finfo(np.float64).eps)
    assert la is not None


@pytest.mark.parametrize("kappa", [1e2, 1e6])
def test_gesv_mixed_gmres_first(kappa, dev):
    """Refinement mode 1 (GMRES-IR right after the first LU-IR step) reaches the FP64
    backward error on the GMRES path."""
    import torch
    import linearalgebra_amd as la
    from linearalgebra_amd import _lib
    lib = _lib.load()
    n = 1000
    a = O.conditioned(n, kappa, seed=31)
    b = O.gaussian(n, 1, seed=32).ravel()
    try:
        assert lib.lrn_set_refinement(1) == 0
        x, st = la.gesv_mixed(torch.from_numpy(a).to(dev), torch.from_numpy(b).to(dev), tol=1e-14, max_iter=30,
                              restart=40)
    finally:
        lib.lrn_set_refinement(0)
    berr = O.backward_error(a, x.cpu().numpy(), b)
    print(f"gmres-first kappa={kappa:.0e} path={st['path_name']} ir={st['ir_iterations']} "
          f"gmres={st['gmres_iterations']} berr={berr:.2e}", flush=True)
    assert berr <= 2e-14
    assert st["ir_iterations"] == 1 or st["path"] == 0
    assert st["path"] in (0, 1)

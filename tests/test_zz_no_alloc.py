"""The allocation-free solver entry point (lrn_linear_solver_noalloc_f64) against the entry point
it varies (lrn_linear_solver_f64), both through the C ABI: two calls in a row (the second with a
smaller order, so the shared device scratch holds the first call's contents), LU and LU_UPFRONT
modes, identical pivots, factors and solutions bit for bit; an order beyond the scratch is
rejected with LRN_E_SCRATCH before any device work."""
import ctypes

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

LRN_E_SCRATCH = 5


def _vp(t):
    return ctypes.c_void_p(0 if t is None else t.data_ptr())


def _run(lib, fn, a, b, mode, dev, ws=None, piv_in=None):
    import torch
    n = a.shape[0]
    m = torch.from_numpy(a).to(dev)
    piv = torch.zeros(n, dtype=torch.int64, device=dev) if piv_in is None else torch.from_numpy(piv_in).to(dev)
    x = torch.from_numpy(b).to(dev)
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    if fn == "noalloc":
        rc = lib.lrn_linear_solver_noalloc_f64(_vp(m), _vp(piv), n, _vp(x), mode, s)
    else:
        rc = lib.lrn_linear_solver_f64(_vp(m), _vp(piv), n, _vp(x), mode, _vp(ws), ws.numel(), s)
    assert rc == 0, rc
    torch.cuda.synchronize()
    return m.cpu().numpy(), piv.cpu().numpy(), x.cpu().numpy()


def test_noalloc_matches_workspace_entry_point(dev):
    import torch
    from linearalgebra_amd import _lib
    lib = _lib.load()
    for n in (3000, 700):  # second call smaller: the scratch still holds the first call's state
        a = O.gaussian(n, seed=40 + n)
        b = O.gaussian(n, 1, seed=41 + n)[:, 0].copy()
        size = ctypes.c_size_t(0)
        assert lib.lrn_linear_solver_workspace_size_f64(n, 1, ctypes.byref(size)) == 0
        ws = torch.empty(size.value, dtype=torch.uint8, device=dev)
        lu_w, piv_w, x_w = _run(lib, "ws", a, b, 1, dev, ws=ws)
        lu_n, piv_n, x_n = _run(lib, "noalloc", a, b, 1, dev)
        print(f"noalloc n={n}: pivots equal {np.array_equal(piv_n, piv_w)}, LU equal {np.array_equal(lu_n, lu_w)}, "
              f"x equal {np.array_equal(x_n, x_w)}", flush=True)
        assert np.array_equal(piv_n, piv_w) and np.array_equal(lu_n, lu_w) and np.array_equal(x_n, x_w)
        _, piv_o, x_o, _ = O.gesv(a, b, threads=8)
        assert np.array_equal(piv_n, piv_o) and O.forward_error(x_n, x_o) <= 1e-10
        # LU_UPFRONT with those factors
        _, _, xu_w = _run(lib, "ws", lu_w, b, 2, dev, ws=ws, piv_in=piv_w)
        _, _, xu_n = _run(lib, "noalloc", lu_w, b, 2, dev, piv_in=piv_w)
        assert np.array_equal(xu_n, xu_w)


def test_noalloc_rejects_orders_beyond_scratch(dev):
    import torch
    from linearalgebra_amd import _lib
    lib = _lib.load()
    top = ctypes.c_int64(0)
    assert lib.lrn_linear_solver_noalloc_max_order(1, ctypes.byref(top)) == 0
    print(f"noalloc largest LU order: {top.value}", flush=True)
    assert top.value >= 16384
    n = top.value + 1
    m = torch.empty(1, dtype=torch.float64, device=dev)  # never touched: the check comes first
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    assert lib.lrn_linear_solver_noalloc_f64(_vp(m), _vp(m), n, _vp(m), 1, s) == LRN_E_SCRATCH

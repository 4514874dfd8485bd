"""The drop-in boundary without a GPU: the C-ABI library loads and exports every entry point
that include/linalg_rocm_native.h declares, the C++ ROCM_NATIVE specialisations are exported
under the reference's template names, argument errors are reported before any device work,
and the host-only part of the C++ backend test passes. No compute call is made here."""
import ctypes
import os
import re
import subprocess

import pytest

import linearalgebra_amd as la
from linearalgebra_amd import _lib
from linearalgebra_amd import build as B

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "linalg_rocm_native.h")


def _declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"LRN_API\s+[\w\s\*]+?\b(lrn_\w+)\s*\(", text)))


def test_header_declares_entry_points():
    names = _declared()
    for must in ("lrn_lu_decomposition_f64", "lrn_lu_decomposition_workspace_size_f32", "lrn_linear_solver_f64",
                 "lrn_linear_solver_multi_f32", "lrn_gesv_mixed_f64", "lrn_trace_read"):
        assert must in names
    assert len(names) == len(re.findall(r"^LRN_API", open(HEADER).read(), re.M))


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, missing
    # and the ctypes table covers them all
    assert set(_declared()) <= set(_lib.SIGNATURES), set(_declared()) - set(_lib.SIGNATURES)


def test_exports_are_only_the_c_abi():
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True, check=True)
    funcs = {l.split()[-1] for l in out.stdout.splitlines() if " T " in l}
    assert funcs, "no exported functions"
    assert all(f.startswith("lrn_") for f in funcs), sorted(f for f in funcs if not f.startswith("lrn_"))


def test_cxx_specialisations_exported():
    out = subprocess.run(["nm", "-DC", "--defined-only", B.CXX_LIB], capture_output=True, text=True, check=True)
    text = out.stdout
    for t in ("double", "float"):
        assert f"linalg::lu_decomposition<{t}, (linalg::linalg_backend)6>::apply" in text
        assert f"linalg::lu_decomposition<{t}, (linalg::linalg_backend)6>::workspace_size" in text
        assert f"linalg::linear_solver<{t}, (linalg::linalg_backend)6>::apply" in text
        assert f"linalg::linear_solver<{t}, (linalg::linalg_backend)6>::workspace_size" in text
        assert f"linalg::cholesky_decomposition<{t}, (linalg::linalg_backend)6>::apply" in text
        assert f"linalg::cholesky_decomposition<{t}, (linalg::linalg_backend)6>::workspace_size" in text
    assert "linalg::rocm_native::gesv_mixed" in text


def test_status_strings_and_version():
    lib = _lib.load()
    assert b"gfx950" in lib.lrn_version()
    for code in (0, 1, 2, 3, 4, -1):
        assert lib.lrn_status_string(code)


def test_workspace_sizes_host_only():
    for dt in ("f64", "f32"):
        out = ctypes.c_size_t(0)
        assert getattr(_lib.load(), f"lrn_lu_decomposition_workspace_size_{dt}")(4096, ctypes.byref(out)) == 0
        assert out.value > 0
    # LU mode needs at least what the upfront mode needs
    s_lu = la.linear_solver.workspace_size(1000, la.linear_solver_type.LU_LINEAR_SOLVER)
    s_up = la.linear_solver.workspace_size(1000, la.linear_solver_type.LU_UPFRONT_LINEAR_SOLVER)
    assert s_lu >= s_up > 0
    # multi-RHS workspace grows with nrhs
    s1 = la.linear_solver.workspace_size(1000, la.linear_solver_type.LU_UPFRONT_LINEAR_SOLVER, nrhs=2)
    s2 = la.linear_solver.workspace_size(1000, la.linear_solver_type.LU_UPFRONT_LINEAR_SOLVER, nrhs=64)
    assert s2 > s1


def test_argument_errors_before_device_work():
    lib = _lib.load()
    nil = ctypes.c_void_p(0)
    dummy = ctypes.c_void_p(16)
    info = ctypes.c_int(0)
    # null m, bad lda, null pivot, too-small workspace: -1, -2, -3, -4
    assert lib.lrn_lu_decomposition_f64(nil, 4, dummy, dummy, 1 << 20, ctypes.byref(info), nil) == -1
    assert lib.lrn_lu_decomposition_f64(dummy, 0, dummy, dummy, 1 << 20, ctypes.byref(info), nil) == -2
    assert lib.lrn_lu_decomposition_f64(dummy, 4, nil, dummy, 1 << 20, ctypes.byref(info), nil) == -3
    assert lib.lrn_lu_decomposition_f64(dummy, 4, dummy, dummy, 1, ctypes.byref(info), nil) == -4
    # Cholesky modes take no pivot array; a too-small workspace is still caught first
    assert lib.lrn_linear_solver_f64(dummy, nil, 4, dummy, 3, dummy, 1, nil) == -8
    assert lib.lrn_linear_solver_f64(dummy, nil, 4, dummy, 1, dummy, 1 << 20, nil) == -2
    # Cholesky factorisation: null C, bad lda, bad triangle, too-small workspace
    assert lib.lrn_cholesky_decomposition_f64(nil, 4, 0, dummy, 1 << 20, ctypes.byref(info), nil) == -1
    assert lib.lrn_cholesky_decomposition_f64(dummy, 0, 0, dummy, 1 << 20, ctypes.byref(info), nil) == -2
    assert lib.lrn_cholesky_decomposition_f64(dummy, 4, 2, dummy, 1 << 20, ctypes.byref(info), nil) == -3
    assert lib.lrn_cholesky_decomposition_f64(dummy, 4, 0, dummy, 1, ctypes.byref(info), nil) == -4
    # determinant / inversion: too-small workspace, unknown type, missing UPFRONT pivots
    det = ctypes.c_double(0.0)
    assert lib.lrn_determinant_f64(dummy, nil, 4, 3, dummy, 1, ctypes.byref(det), nil) == -5
    assert lib.lrn_determinant_f64(dummy, nil, 4, 7, dummy, 1 << 20, ctypes.byref(det), nil) == -4
    assert lib.lrn_determinant_f64(dummy, nil, 4, 2, dummy, 1 << 20, ctypes.byref(det), nil) == -2
    assert lib.lrn_inversion_f64(dummy, nil, 4, 1, dummy, 1, nil) == -5
    # unknown solver type
    assert lib.lrn_linear_solver_f64(dummy, dummy, 4, dummy, 9, dummy, 1 << 20, nil) == -7
    # mixed solve argument checks
    st = ctypes.create_string_buffer(128)
    assert lib.lrn_gesv_mixed_f64(nil, 4, dummy, dummy, 1e-14, 10, 30, dummy, 1 << 20, st, nil) == -1
    assert lib.lrn_gesv_mixed_f64(dummy, 4, dummy, dummy, 0.0, 10, 30, dummy, 1 << 20, st, nil) == -5
    assert lib.lrn_gesv_mixed_f64(dummy, 4, dummy, dummy, 1e-14, 10, 500, dummy, 1 << 20, st, nil) == -7


def test_python_mirror_raises_like_the_reference():
    import torch
    with pytest.raises(ValueError):
        la.lu_decomposition.apply(None, 4, None)
    with pytest.raises(ValueError):
        la.lu_decomposition.apply(torch.zeros(4, 4, dtype=torch.float64), 4,
                                  torch.zeros(4, dtype=torch.int64))  # host memory
    with pytest.raises(ValueError):
        la.linear_solver.workspace_size(4, la.linear_solver_type.LU_LINEAR_SOLVER, dtype=torch.int32)


def test_cxx_backend_host_checks():
    exe = B.CXX_TEST
    assert os.path.exists(exe), "build with __graft_entry__.build()"
    out = subprocess.run([exe, "--no-gpu"], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "all checks passed" in out.stdout


def test_tuning_file_is_applied(tmp_path):
    """autotune.apply sets the process-wide block sizes from a tuning file (no GPU needed)."""
    import ctypes
    import json
    from linearalgebra_amd import _lib, autotune
    lib = _lib.load()
    path = tmp_path / "tuning.json"
    path.write_text(json.dumps({"getrf": {"outer_nb": 384, "panel_w": 32, "panel_kernel": 1},
                                "batched": {"grid": 0}}))
    try:
        assert autotune.apply(lib, str(path))
        nb, w, mx = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        lib.lrn_get_tuning(ctypes.byref(nb), ctypes.byref(w), ctypes.byref(mx))
        assert nb.value == 384 and w.value == 32
    finally:
        lib.lrn_set_tuning(0, 0, 0)
        lib.lrn_set_panel_kernel(0)
        autotune.apply(lib)
    assert not autotune.apply(lib, str(tmp_path / "absent.json"))


def test_tuning_setters_validate():
    """Process-wide switches reject values outside their documented range (no GPU needed)."""
    from linearalgebra_amd import _lib, autotune
    lib = _lib.load()
    try:
        assert lib.lrn_set_lookahead(-1) != 0
        assert lib.lrn_set_lookahead(5) != 0      # not a multiple of 8
        assert lib.lrn_set_lookahead(12) != 0
        assert lib.lrn_set_lookahead(16) == 0
        assert lib.lrn_set_lookahead(1) == 0
        assert lib.lrn_set_block_update(3) != 0
        assert lib.lrn_set_block_update(2) == 0
        assert lib.lrn_set_block_update(-1) != 0
        assert lib.lrn_set_block_update(1) == 0
        assert lib.lrn_set_refinement(2) != 0
        assert lib.lrn_set_refinement(1) == 0
    finally:
        lib.lrn_set_refinement(0)
        lib.lrn_set_lookahead(0)
        lib.lrn_set_block_update(0)
        autotune.apply(lib)

"""The C++ ROCM_NATIVE backend (reference template names, linalg_rocm_native.hpp) on the GPU:
LU + solve against the oracle (identical pivots, LU and X within 1e-12), LU_UPFRONT reuse,
the zero-pivot KAT and a mixed-precision solve (tests/cxx/test_rocm_native_backend.cpp)."""
import os
import subprocess

import pytest

from linearalgebra_amd import build as B

pytestmark = pytest.mark.gpu


def test_cxx_backend_on_gpu(dev):
    exe = B.CXX_TEST
    assert os.path.exists(exe), "build with __graft_entry__.build()"
    out = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(out.stdout, out.stderr, flush=True)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "all checks passed (host + GPU)" in out.stdout

"""Build recipe of the gfx950 library (hipcc, in-tree output).

    python -m linearalgebra_amd.build

compiles the translation units of ``csrc/`` (``lrn_capi.hip``: kernels + host driver + C ABI;
``lrn_batched.hip``: the batched solver's kernels) in parallel and links them into
``liblinalg_rocm_native.so`` next to this file, then the C++ backend specialisation
(``csrc/rocm_native_backend.cpp``) into ``liblinalg_rocm_native_cxx.so`` linked against it.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "liblinalg_rocm_native.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# -packed-fp32-ops: no packed FP32 instructions are generated (they are avoided on purpose).
# -ffp-contract=off: fused multiply-adds only where the source asks for fma()/MFMA.
HIP_FLAGS = [
    "-O3", "--offload-arch=gfx950", "-fPIC", "-std=c++17", "-ffp-contract=off",
    "-fvisibility=hidden", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]
# translation units of the library (each launches only the kernels it defines)
HIP_UNITS = ["lrn_capi.hip", "lrn_batched.hip", "lrn_trsm.hip", "lrn_fp8.hip"]
OBJ_DIR = os.path.join(HERE, "_obj")


def _run(cmd):
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=1800)
    out = "\n".join(l for l in proc.stdout.splitlines() if "packed-fp32-ops" not in l)
    if proc.returncode != 0:
        sys.stderr.write(out + "\n")
        raise RuntimeError(f"build failed: {' '.join(cmd)}")
    return out


CXX_LIB = os.path.join(HERE, "liblinalg_rocm_native_cxx.so")
CXX_TEST = os.path.join(ROOT, "tests", "cxx", "test_rocm_native_backend")
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-Wextra"]


def build(verbose: bool = False) -> str:
    """The HIP library (its translation units compiled in parallel, then linked), then the C++
    ROCM_NATIVE specialisations linked against it."""
    from concurrent.futures import ThreadPoolExecutor

    inc = ["-I", os.path.join(ROOT, "include")]
    os.makedirs(OBJ_DIR, exist_ok=True)
    objs = [os.path.join(OBJ_DIR, u.replace(".hip", ".o")) for u in HIP_UNITS]
    with ThreadPoolExecutor(max_workers=len(HIP_UNITS)) as ex:
        outs = list(ex.map(lambda uo: _run([HIPCC, *HIP_FLAGS, *inc, "-c", "-o", uo[1], os.path.join(CSRC, uo[0])]),
                           zip(HIP_UNITS, objs)))
    out = "".join(outs)
    out += _run([HIPCC, "--offload-arch=gfx950", "-fPIC", "-shared", "-o", LIB, *objs])
    out += _run(["g++", *CXX_FLAGS, "-shared", *inc, "-o", CXX_LIB, os.path.join(CSRC, "rocm_native_backend.cpp"),
                 "-L", HERE, "-llinalg_rocm_native", "-Wl,-rpath,$ORIGIN"])
    if verbose and out.strip():
        print(out)
    return LIB


def build_cxx_test() -> str:
    """The C++ backend test (host code; HIP runtime for device buffers, the oracle library as
    the checker). Needs the oracle built first (oracle/Makefile)."""
    oracle_dir = os.path.join(ROOT, "oracle")
    src = CXX_TEST + ".cpp"
    _run([HIPCC, "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"), "-o", CXX_TEST, src,
          "-L", HERE, "-llinalg_rocm_native_cxx", "-llinalg_rocm_native", "-L", oracle_dir, "-llu_oracle",
          f"-Wl,-rpath,{HERE}:{oracle_dir}:$ORIGIN/../../linearalgebra_amd:$ORIGIN/../../oracle"])
    return CXX_TEST


if __name__ == "__main__":
    print(build(verbose=True))

"""Launches one kernel kind of the library repeatedly on resident inputs, for rocprofv3 runs
(kernel traces and counters need a program that does nothing else):

    python -m linearalgebra_amd.profile_kernels gemm_f64 --m 16128 --n 16128 --k 256 --reps 5
    python -m linearalgebra_amd.profile_kernels gemm_f32 --m 16128 --n 16128 --k 256 --reps 5
    python -m linearalgebra_amd.profile_kernels copy --m 16384 --n 16384 --reps 5
    python -m linearalgebra_amd.profile_kernels batched_f64 --count 10000 --reps 5
    python -m linearalgebra_amd.profile_kernels trsm_f64 --m 16384 --n 64 --reps 5
    python -m linearalgebra_amd.profile_kernels gemm_fp8x3 --m 16128 --n 16128 --k 256 --reps 5

gemm_*: C[m x n] -= A[m x k] B[k x n] (row-major, the trailing-update shape of the blocked
LU at order n + k with a k-wide outer block) through lrn_debug_gemm_sub. copy: a device copy
of an m x n FP64 matrix (the calibration of the HBM byte counters). batched_*: the batched
solve of `count` systems, n ~ U[8, 256], nrhs ~ U[1, 8], iid N(0, 1) (the inputs are restored
by a device copy before each launch). trsm_f64: the triangular solver on an m x m unit-lower
triangle with n right-hand sides (the blocked many-RHS form for n >= 2). gemm_fp8x1/2/3: the
block-scaled FP8 trailing update with 1/2/3 e4m3 planes. Every launch goes to the current stream; the program
synchronises at the end and prints one line per phase.
"""
from __future__ import annotations

import argparse
import ctypes
import sys
import time


def main() -> None:
    p = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    p.add_argument("kind", choices=["gemm_f64", "gemm_f32", "gemm_bf16x1", "gemm_bf16x2", "gemm_bf16x3", "gemm_fp8x1",
                                    "gemm_fp8x2", "gemm_fp8x3", "copy", "batched_f64", "batched_f32", "trsm_f64"])
    p.add_argument("--m", type=int, default=16128)
    p.add_argument("--n", type=int, default=16128)
    p.add_argument("--k", type=int, default=256)
    p.add_argument("--count", type=int, default=10000)
    p.add_argument("--reps", type=int, default=5)
    p.add_argument("--seed", type=int, default=2024)
    p.add_argument("--time", action="store_true",
                   help="gemm_*: time each launch with HIP events on its stream and print the fastest")
    a = p.parse_args()

    import torch

    from . import _lib

    lib = _lib.load()
    dev = torch.device("cuda:0")
    g = torch.Generator(device=dev)
    g.manual_seed(a.seed)
    vp = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    t0 = time.perf_counter()
    if a.kind.startswith("gemm"):
        dtype = torch.float64 if a.kind == "gemm_f64" else torch.float32
        code = {"gemm_f64": 0, "gemm_f32": 1, "gemm_bf16x1": 2, "gemm_bf16x2": 3, "gemm_bf16x3": 4, "gemm_fp8x1": 5,
                "gemm_fp8x2": 6, "gemm_fp8x3": 7}[a.kind]
        A = torch.randn(a.m, a.k, dtype=dtype, device=dev, generator=g)
        B = torch.randn(a.k, a.n, dtype=dtype, device=dev, generator=g)
        C = torch.randn(a.m, a.n, dtype=dtype, device=dev, generator=g)
        torch.cuda.synchronize()
        print(f"{a.kind}: {a.m} x {a.n} x {a.k}, {a.reps} launches", flush=True)
        evs = []
        for _ in range(a.reps):
            e0 = torch.cuda.Event(enable_timing=True) if a.time else None
            e1 = torch.cuda.Event(enable_timing=True) if a.time else None
            if a.time:
                e0.record()
            rc = lib.lrn_debug_gemm_sub(code, vp(A), a.k, vp(B), a.n, vp(C), a.n, a.m, a.n, a.k, s)
            if a.time:
                e1.record()
                evs.append((e0, e1))
            if rc != 0:
                sys.exit(f"lrn_debug_gemm_sub returned {rc}")
        if a.time:
            torch.cuda.synchronize()
            ms = min(x.elapsed_time(y) for x, y in evs)
            print(f"{a.kind}: fastest launch {ms:.3f} ms = {2.0 * a.m * a.n * a.k / ms / 1e9:.1f} TF/s "
                  f"(HIP events on the launch stream)", flush=True)
    elif a.kind == "trsm_f64":
        from .triangular import triangular_solver
        A = torch.randn(a.m, a.m, dtype=torch.float64, device=dev, generator=g) / a.m ** 0.5
        B = torch.randn(a.m, a.n, dtype=torch.float64, device=dev, generator=g)
        X = torch.empty_like(B)
        ws = torch.empty(triangular_solver.workspace_size(a.m, a.n), dtype=torch.uint8, device=dev)
        torch.cuda.synchronize()
        print(f"trsm_f64: {a.m} x {a.m} unit lower, {a.n} right-hand sides, {a.reps} solves", flush=True)
        for _ in range(a.reps):
            triangular_solver.apply(False, False, True, a.m, a.n, A, a.m, B, a.n, X, a.n, ws)
    elif a.kind == "copy":
        src = torch.randn(a.m, a.n, dtype=torch.float64, device=dev, generator=g)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        print(f"copy: {a.m} x {a.n} FP64 ({src.numel() * 8} bytes), {a.reps} launches", flush=True)
        for _ in range(a.reps):
            dst.copy_(src)
    else:
        import numpy as np

        from .batched import BatchedSystems

        dtype = torch.float64 if a.kind == "batched_f64" else torch.float32
        rng = np.random.Generator(np.random.PCG64(a.seed))
        n = rng.integers(8, 257, size=a.count)
        r = rng.integers(1, 9, size=a.count)
        sy = BatchedSystems(n, r, dtype=dtype, device=dev)
        sy.a.copy_(torch.randn(sy.a.numel(), dtype=dtype, device=dev, generator=g))
        sy.b.copy_(torch.randn(sy.b.numel(), dtype=dtype, device=dev, generator=g))
        a0, b0 = sy.a.clone(), sy.b.clone()
        torch.cuda.synchronize()
        print(f"{a.kind}: {a.count} systems, {a.reps} launches", flush=True)
        for _ in range(a.reps):
            sy.a.copy_(a0)
            sy.b.copy_(b0)
            sy.solve()
        bad = int((sy.info != 0).sum())
        if bad:
            sys.exit(f"{bad} systems reported a nonzero info")
    torch.cuda.synchronize()
    print(f"done in {time.perf_counter() - t0:.2f} s", flush=True)


if __name__ == "__main__":
    main()

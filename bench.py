#!/usr/bin/env python3
"""Benchmark of the gfx950 LU solve path: FP64-accurate solves of A x = b at n = 16384.

One step = one complete solve of a resident n x n FP64 system (nrhs = 1):
  mixed  (the headline `value`): lrn_gesv_mixed_f64 -- FP32 panels, split-bf16 MFMA trailing
         update with FP32 accumulation, LU-IR then GMRES-IR with FP64 residuals, to a
         normwise backward error <= 1e-14;
  plain  (reported beside it): lrn_linear_solver_f64 in LU mode -- the FP64 blocked LU with
         partial pivoting (FP64 4x4x4 MFMA trailing update) and the triangular solves.

Input: A = U diag(s) V^T with U, V Haar-orthogonal (Q of the QR of a seeded Gaussian, sign
fixed by R's diagonal) and s log-spaced from 1 to 1/kappa; b iid N(0, 1). Generated on the
GPU with torch (not timed). The timed region starts with A and b resident in HBM.

Live roofline: the library's kernel trace (HIP events recorded on the launch stream around
every launch of the traced kernel kinds) runs during the timed steps; `roofline` is the
dominant matrix-core kernel of the measured path (algorithmic flops 2*M*N*K per launch,
summed over launches, divided by the summed launch durations). `phases` gives every kernel
kind from one extra traced step. `cpu_baseline` is the CPU restatement of the reference's
native algorithm (single thread) and LAPACK (1 thread and all cores) timed on this host at
n = 1024, 2048, 4096 and on the CPU configuration (n = 1024, nrhs = 16), fitted by n^3.

Configuration grids (skipped with --no-configs): `plain_configs` times the plain FP64 LU +
solve at n = 4096, 8192, 16384 with nrhs 1 and 64 on A iid N(0,1) and on A + n*I, each with
its backward error; `mixed_configs` runs the mixed solve at the headline order for kappa =
1e2 and 1e10 (the headline is kappa = 1e6), with path, iteration counts, backward error and
forward difference from the plain FP64 solution (--mixed-extra-n 32768 adds that order);
`triangular_solver` times op(A) X = B at the headline order for nrhs 1 and 64 and four flag
combinations (HBM rate of the algorithmic bytes, residual).

Usage: python bench.py [--steps K] [--warmup W] [--n N] [--kappa K] [--cpu-sizes N,..] [--no-cpu]
"""
from __future__ import annotations

import argparse
import ctypes
import json
import math
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
METRIC = "FP64-accurate Ax=b solves/s: mixed LU-IR vs plain FP64 LU, n=16384; batch solves/s"
PEAK = {"fp64_mfma_tflops": 78.6, "bf16_mfma_tflops": 2500.0, "fp32_tflops": 157.3, "hbm_gbs": 8000.0}
TRACE = {"gemm_f64": 0, "gemm_f64_edge": 1, "gemm_bf16x3": 2, "gemm_bf16": 3, "gemm_f32": 4, "panel": 5,
         "laswp": 6, "trsm": 7, "trsv": 8, "gemm_bf16x6": 9, "block_column_update": 11}
FLOP_KINDS = {"gemm_f64", "gemm_f64_edge", "gemm_bf16x3", "gemm_bf16", "gemm_f32", "panel", "gemm_bf16x6",
              "block_column_update"}
KERNEL_SYMBOL = {"gemm_f64": "gemm_sub_f64_kernel<false>", "gemm_bf16x3": "gemm_sub_bf16_kernel<2>",
                 "gemm_bf16x6": "gemm_sub_bf16_kernel<3>", "gemm_bf16": "gemm_sub_bf16_kernel<1>"}
# bf16 matrix-core products per algorithmic product, by trailing-update mode
PLANE_PRODUCTS = {"gemm_bf16": 1, "gemm_bf16x3": 3, "gemm_bf16x6": 6}


def log(msg: str) -> None:
    sys.stderr.write(f"[bench] {msg}\n")
    sys.stderr.flush()


def parse():
    p = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=5)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--n", type=int, default=16384)
    p.add_argument("--kappa", type=float, default=1e6)
    p.add_argument("--seed", type=int, default=2024)
    p.add_argument("--restart", type=int, default=30)
    p.add_argument("--cpu-sizes", default="1024,2048,4096", help="orders of the CPU baseline samples")
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--matrix", choices=["conditioned", "gaussian"], default="conditioned",
                   help="conditioned: U diag(s) V^T with the given kappa (the benchmark input); gaussian: iid "
                        "N(0,1) entries, no QR (for profiler runs that cannot host the QR)")
    p.add_argument("--no-plain", action="store_true")
    p.add_argument("--no-batched", action="store_true")
    p.add_argument("--batch-count", type=int, default=10000)
    p.add_argument("--batch-replays", type=int, default=20)
    p.add_argument("--batch-mixed-count", type=int, default=64,
                   help="systems of the batched mixed-precision workload (0 = skip)")
    p.add_argument("--no-configs", action="store_true", help="skip the configuration grids")
    p.add_argument("--config-sizes", default="4096,8192,16384",
                   help="orders of the plain FP64 grid (nrhs 1 and 64, A iid N(0,1) and A + n I)")
    p.add_argument("--consumer-sizes", default="4096,16384",
                   help="orders of the determinant / inversion / Cholesky timing (comma-separated)")
    p.add_argument("--mixed-kappas", default="1e2,1e10",
                   help="further condition numbers of the mixed grid at the headline order")
    p.add_argument("--mixed-extra-n", type=int, default=0,
                   help="also run the mixed grid (all three condition numbers) at this order, e.g. 32768")
    p.add_argument("--mixed-update", type=int, default=3,
                   help="trailing update of the mixed factorisation: 0/1/2 = 1/2/3 bf16 planes, 3 = FP32 MFMA")
    return p.parse_args()


# ----------------------------------------------------------------------------- trace helpers

def trace_start(lib, kinds, capacity=1 << 16):
    mask = 0
    for k in kinds:
        mask |= 1 << TRACE[k]
    assert lib.lrn_trace_start(mask, capacity) == 0


def trace_read(lib, kind):
    c, ms, w = ctypes.c_int64(), ctypes.c_double(), ctypes.c_double()
    assert lib.lrn_trace_read(TRACE[kind], ctypes.byref(c), ctypes.byref(ms), ctypes.byref(w)) == 0
    return int(c.value), float(ms.value), float(w.value)


def phase_table(lib):
    out = {}
    for k in TRACE:
        c, ms, w = trace_read(lib, k)
        if c == 0:
            continue
        e = {"launches": c, "ms": round(ms, 3), "avg_us": round(1e3 * ms / c, 2)}
        if k in FLOP_KINDS:
            e["tflops"] = round(w / (ms * 1e-3) / 1e12, 2) if ms > 0 else None
        else:
            e["gbs"] = round(w / (ms * 1e-3) / 1e9, 1) if ms > 0 else None
        out[k] = e
    return out


def roofline(lib, kind, peak_tflops, traffic):
    c, ms, w = trace_read(lib, kind)
    if c == 0 or ms <= 0:
        return None
    achieved = w / (ms * 1e-3) / 1e12
    out = {"bound": "mfma", "kernel": KERNEL_SYMBOL.get(kind, kind), "achieved": round(achieved, 2),
           "peak": peak_tflops, "unit": "TFLOP/s", "frac": round(achieved / peak_tflops, 4), "traffic": traffic,
           "launches": c, "avg_launch_us": round(1e3 * ms / c, 2),
           "flops_per_launch": round(w / c, 1), "flops": "algorithmic 2*M*N*K per launch"}
    if kind in PLANE_PRODUCTS:
        k = PLANE_PRODUCTS[kind]
        out["matrix_core_flops_per_launch"] = round(k * w / c, 1)
        out["matrix_core_frac"] = round(k * achieved / peak_tflops, 4)
    return out


def traffic_from_profiles(kind):
    """HBM bytes per launch measured with rocprofv3 --pmc (FETCH_SIZE + WRITE_SIZE, scaled by a
    copy-kernel calibration) and written into the committed summary; None if absent."""
    path = os.path.join(HERE, "profiles", f"{kind}_n16384_traffic.txt")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":", 1)[1])
    return None


# ----------------------------------------------------------------------------- inputs

def conditioned_gpu(torch, n, kappa, seed, dev):
    """U diag(s) V^T on the GPU, U and V Haar (Q of a seeded Gaussian, R's diagonal made positive)."""
    g = torch.Generator(device=dev)

    def haar(sd):
        g.manual_seed(sd)
        m = torch.randn(n, n, dtype=torch.float64, device=dev, generator=g)
        q, r = torch.linalg.qr(m)
        del m
        q *= torch.sign(torch.diagonal(r)).unsqueeze(0)
        del r
        return q

    u = haar(seed)
    s = torch.logspace(0.0, -torch.log10(torch.tensor(kappa)).item(), n, dtype=torch.float64, device=dev)
    u *= s.unsqueeze(0)
    v = haar(seed + 1)
    a = u @ v.T
    del u, v
    g.manual_seed(seed + 2)
    b = torch.randn(n, dtype=torch.float64, device=dev, generator=g)
    return a.contiguous(), b


def backward_error_gpu(torch, a, x, b):
    r = b - a @ x
    den = torch.linalg.matrix_norm(a, ord=float("inf")) * x.abs().max() + b.abs().max()
    return float(r.abs().max() / den)


def backward_error_cols(torch, a, x, b):
    """Largest normwise backward error over the right-hand sides (columns of x, b)."""
    x2 = x.reshape(x.shape[0], -1)
    b2 = b.reshape(b.shape[0], -1)
    r = b2 - a @ x2
    anorm = torch.linalg.matrix_norm(a, ord=float("inf"))
    den = anorm * x2.abs().amax(dim=0) + b2.abs().amax(dim=0)
    return float((r.abs().amax(dim=0) / den).max())


# ----------------------------------------------------------------------------- configs

def plain_configs_bench(torch, sizes, nrhs_list, seed, dev, reps=3):
    """Plain FP64 LU + solve over the grid of the FP64 GPU workload: n in `sizes`, nrhs in
    `nrhs_list`, A iid N(0,1) and A + n*I, B iid N(0,1). Per case: median of `reps` timed
    solves after one warmup (HIP events around the solver call on the launch stream; the
    restore copies of A and B are outside) and the backward error of the result."""
    import linearalgebra_amd as la
    lu_mode = la.linear_solver_type.LU_LINEAR_SOLVER
    out = []
    for n in sizes:
        g = torch.Generator(device=dev)
        g.manual_seed(seed + n)
        a0 = torch.randn(n, n, dtype=torch.float64, device=dev, generator=g)
        piv = torch.empty(n, dtype=torch.int64, device=dev)
        aw = torch.empty_like(a0)
        for variant in ("gaussian", "diag_dominant"):
            a = a0 if variant == "gaussian" else a0 + n * torch.eye(n, dtype=torch.float64, device=dev)
            for nrhs in nrhs_list:
                b = torch.randn(n, nrhs, dtype=torch.float64, device=dev, generator=g)
                if nrhs == 1:
                    b = b.reshape(n).contiguous()
                x = torch.empty_like(b)
                ws = torch.empty(la.linear_solver.workspace_size(n, lu_mode, torch.float64, nrhs), dtype=torch.uint8,
                                 device=dev)
                times = []
                for r in range(reps + 1):
                    aw.copy_(a)
                    x.copy_(b)
                    e0 = torch.cuda.Event(enable_timing=True)
                    e1 = torch.cuda.Event(enable_timing=True)
                    e0.record()
                    la.linear_solver.apply(aw, piv, n, x, lu_mode, ws)
                    e1.record()
                    torch.cuda.synchronize()
                    if r > 0:
                        times.append(e0.elapsed_time(e1))
                ms = sorted(times)[len(times) // 2]
                berr = backward_error_cols(torch, a, x, b)
                out.append({"n": n, "matrix": variant, "nrhs": nrhs, "ms": round(ms, 3),
                            "solves_per_s": round(1e3 / ms, 3), "lu_tflops": round(2.0 * n ** 3 / 3.0 / (ms * 1e-3) / 1e12, 2),
                            "backward_error": berr})
                log(f"plain n={n} {variant} nrhs={nrhs}: {ms:.2f} ms, berr {berr:.2e}")
                del ws, x, b
            if variant == "diag_dominant":
                del a
        del a0, aw, piv
    return out


def mixed_configs_bench(torch, sizes, kappas, seed, dev, restart):
    """Mixed LU-IR over the grid of the mixed workload (n in `sizes`, kappa in `kappas`, A = U
    diag(s) V^T as the headline input): one warmup and one timed solve each, the refinement
    path and iteration counts, the backward error, and the forward difference from the plain
    FP64 GPU LU solution of the same system."""
    import linearalgebra_amd as la
    from linearalgebra_amd import mixed as lam
    lu_mode = la.linear_solver_type.LU_LINEAR_SOLVER
    out = []
    for n in sizes:
        ws = torch.empty(lam.workspace_size(n, restart, True), dtype=torch.uint8, device=dev)
        for kappa in kappas:
            a, b = conditioned_gpu(torch, n, kappa, seed, dev)
            x = torch.empty(n, dtype=torch.float64, device=dev)
            la.gesv_mixed(a, b, x, tol=1e-14, max_iter=30, restart=restart, workspace=ws)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            _, st = la.gesv_mixed(a, b, x, tol=1e-14, max_iter=30, restart=restart, workspace=ws)
            torch.cuda.synchronize()
            ms = 1e3 * (time.perf_counter() - t0)
            berr = backward_error_gpu(torch, a, x, b)
            aw = a.clone()
            xp = b.clone()
            piv = torch.empty(n, dtype=torch.int64, device=dev)
            la.linear_solver.apply(aw, piv, n, xp, lu_mode)
            fwd = float((x - xp).abs().max() / xp.abs().max())
            del aw, piv
            out.append({"n": n, "kappa": kappa, "ms": round(ms, 2), "solves_per_s": round(1e3 / ms, 3),
                        "path": ["lu-ir", "gmres-ir", "fp64-fallback", "not-converged"][st["path"]],
                        "ir_iterations": st["ir_iterations"], "gmres_iterations": st["gmres_iterations"],
                        "factor_ms": round(st["factor_ms"], 2), "refine_ms": round(st["refine_ms"], 2),
                        "backward_error": berr, "forward_difference_vs_fp64_lu": fwd})
            log(f"mixed n={n} kappa={kappa:g}: {ms:.1f} ms path={out[-1]['path']} berr {berr:.2e} fwd {fwd:.1e}")
            del a, b, x, xp
        del ws
    return out


def triangular_bench(torch, n, seed, dev, reps=3):
    """triangular_solver<double, ROCM_NATIVE> on an n x n triangle (entries N(0,1)/sqrt(n) off
    the diagonal, 2 + U[0,1) on it) with nrhs 1 and 64, four flag combinations: median of
    `reps` HIP-event-timed solves after a warmup, HBM rate of the algorithmic bytes
    ((n^2/2 + 2 n nrhs) * 8) against the 8 TB/s peak, and the relative residual."""
    import linearalgebra_amd as la
    g = torch.Generator(device=dev)
    g.manual_seed(seed + 77)
    a = torch.randn(n, n, dtype=torch.float64, device=dev, generator=g) / n ** 0.5
    a += torch.diag(2.0 + torch.rand(n, dtype=torch.float64, device=dev, generator=g))
    out = []
    for nrhs in (1, 64):
        b = torch.randn(n, nrhs, dtype=torch.float64, device=dev, generator=g)
        x = torch.empty_like(b)
        ws = torch.empty(la.triangular_solver.workspace_size(n, nrhs), dtype=torch.uint8, device=dev)
        for upper, transpose, unit in ((False, False, True), (True, False, False), (False, True, False), (True, True, True)):
            times = []
            for r in range(reps + 1):
                e0 = torch.cuda.Event(enable_timing=True)
                e1 = torch.cuda.Event(enable_timing=True)
                e0.record()
                la.triangular_solver.apply(upper, transpose, unit, n, nrhs, a, n, b, nrhs, x, nrhs, ws)
                e1.record()
                torch.cuda.synchronize()
                if r > 0:
                    times.append(e0.elapsed_time(e1))
            ms = sorted(times)[len(times) // 2]
            t = torch.triu(a) if upper else torch.tril(a)
            if unit:
                t.fill_diagonal_(1.0)
            op = t.T if transpose else t
            res = float((op @ x - b).abs().max() / ((op.abs().sum(1).max()) * x.abs().max() + b.abs().max()))
            gbs = (0.5 * n * n + 2.0 * n * nrhs) * 8 / (ms * 1e-3) / 1e9
            out.append({"nrhs": nrhs, "upper": upper, "transpose": transpose, "unit": unit, "ms": round(ms, 3),
                        "gbs": round(gbs, 1), "hbm_frac": round(gbs / 8000.0, 4), "backward_error": res})
            log(f"triangular n={n} nrhs={nrhs} upper={upper} transpose={transpose} unit={unit}: {ms:.2f} ms, "
                f"{gbs:.0f} GB/s, berr {res:.1e}")
            del t, op
        del b, x, ws
    del a
    return {"n": n, "cases": out}


def consumers_bench(torch, sizes, seed, dev, reps=2):
    """determinant<double, ROCM_NATIVE> and inversion<double, ROCM_NATIVE> (LU mode, pivots in
    the workspace) on N(0,1) matrices of each order: median of `reps` HIP-event-timed calls after
    a warmup (the factorisation included), the determinant's ratio to the product of the factor
    diagonal computed by torch from the same factors, and max|A^-1 A - I| of the inverse."""
    import linearalgebra_amd as la
    LU = la.linear_solver_type.LU_LINEAR_SOLVER
    out = []
    for n in sizes:
        g = torch.Generator(device=dev)
        g.manual_seed(seed + n)
        # scale sqrt(e/n): E log|det| of an N(0,1) matrix is about (n ln n - n)/2, so the
        # determinant stays within double range
        a = torch.randn(n, n, dtype=torch.float64, device=dev, generator=g) * (2.718281828459045 / n) ** 0.5
        m = torch.empty_like(a)
        row = {"n": n}
        for name, op in (("determinant", la.determinant), ("inversion", la.inversion)):
            ws = torch.empty(op.workspace_size(n, LU), dtype=torch.uint8, device=dev)
            times = []
            val = None
            for r in range(reps + 1):
                m.copy_(a)
                torch.cuda.synchronize()
                e0 = torch.cuda.Event(enable_timing=True)
                e1 = torch.cuda.Event(enable_timing=True)
                e0.record()
                val = op.apply(m, None, n, LU, ws)
                e1.record()
                torch.cuda.synchronize()
                if r > 0:
                    times.append(e0.elapsed_time(e1))
            ms = sorted(times)[len(times) // 2]
            row[f"{name}_ms"] = round(ms, 3)
            if name == "determinant":
                sign, logabs = torch.linalg.slogdet(a)
                row["determinant"] = val
                row["log_abs_det"] = math.log(abs(val)) if val != 0 and math.isfinite(val) else None
                row["log_abs_det_torch"] = float(logabs)
                row["sign_matches_torch"] = bool((val > 0) == (float(sign) > 0))
            else:
                err = float((m @ a - torch.eye(n, dtype=a.dtype, device=dev)).abs().max())
                row["inverse_max_err"] = err
                flops = 2.0 / 3.0 * n ** 3 + 2.0 * n ** 3
                row["inversion_tflops"] = round(flops / (ms * 1e-3) / 1e12, 2)
            del ws
        log(f"consumers n={n}: determinant {row['determinant_ms']:.1f} ms, inversion {row['inversion_ms']:.1f} ms "
            f"({row['inversion_tflops']} TFLOP/s), max|inv A - I| {row['inverse_max_err']:.1e}")
        out.append(row)
        del a, m
    return out


def cholesky_bench(torch, sizes, seed, dev, reps=2):
    """cholesky_decomposition<double, ROCM_NATIVE> (lower) and linear_solver in CHOLESKY mode
    (factor + one right-hand side) on A = G G^T / n + I: median of `reps` HIP-event-timed calls
    after a warmup, the FP64 rate of n^3/3 flops, and max|L L^T - A| / max|A|."""
    import linearalgebra_amd as la
    out = []
    for n in sizes:
        g = torch.Generator(device=dev)
        g.manual_seed(seed + 3 * n)
        x = torch.randn(n, n, dtype=torch.float64, device=dev, generator=g)
        a = x @ x.T / n + torch.eye(n, dtype=torch.float64, device=dev)
        del x
        m = torch.empty_like(a)
        ws = torch.empty(la.cholesky_decomposition.workspace_size(n), dtype=torch.uint8, device=dev)
        b = torch.randn(n, dtype=torch.float64, device=dev, generator=g)
        xs = torch.empty_like(b)
        sws = torch.empty(la.linear_solver.workspace_size(n, la.linear_solver_type.CHOLESKY_LINEAR_SOLVER),
                          dtype=torch.uint8, device=dev)
        row = {"n": n}
        for name in ("factor", "solve"):
            times = []
            for r in range(reps + 1):
                m.copy_(a)
                xs.copy_(b)
                torch.cuda.synchronize()
                e0 = torch.cuda.Event(enable_timing=True)
                e1 = torch.cuda.Event(enable_timing=True)
                e0.record()
                if name == "factor":
                    la.cholesky_decomposition.apply(m, n, workspace=ws)
                else:
                    la.linear_solver.apply(m, None, n, xs, la.linear_solver_type.CHOLESKY_LINEAR_SOLVER, sws)
                e1.record()
                torch.cuda.synchronize()
                if r > 0:
                    times.append(e0.elapsed_time(e1))
            ms = sorted(times)[len(times) // 2]
            row[f"{name}_ms"] = round(ms, 3)
            if name == "factor":
                row["factor_tflops"] = round(n ** 3 / 3.0 / (ms * 1e-3) / 1e12, 2)
                row["residual"] = float((m @ m.T - a).abs().max() / a.abs().max())
            else:
                row["backward_error"] = float((a @ xs - b).abs().max() /
                                              (a.abs().sum(1).max() * xs.abs().max() + b.abs().max()))
        log(f"cholesky n={n}: factor {row['factor_ms']:.1f} ms ({row['factor_tflops']} TFLOP/s), "
            f"factor+solve {row['solve_ms']:.1f} ms, residual {row['residual']:.1e}, berr {row['backward_error']:.1e}")
        out.append(row)
        del a, m, ws, b, xs, sws
    return out


def bmm_bench(torch, seed, dev, reps=5):
    """batched_matrix_multiplication (FP64 and FP32) on count x (dim x dim) U(-5, 5) operands
    for three shapes: median of `reps` HIP-event-timed launches after a warmup, the rate of
    2 dim^3 count flops against the dense peak of the type, and the max error against torch."""
    import linearalgebra_amd as la
    out = []
    for dim, count in ((9, 65536), (64, 16384), (256, 1024)):
        for dt, name, peak in ((torch.float64, "f64", PEAK["fp64_mfma_tflops"]), (torch.float32, "f32", PEAK["fp32_tflops"])):
            g = torch.Generator(device=dev)
            g.manual_seed(seed + dim)
            a = (torch.rand(count, dim, dim, dtype=dt, device=dev, generator=g) - 0.5) * 10
            b = (torch.rand(count, dim, dim, dtype=dt, device=dev, generator=g) - 0.5) * 10
            c = torch.empty_like(a)
            times = []
            for r in range(reps + 1):
                e0 = torch.cuda.Event(enable_timing=True)
                e1 = torch.cuda.Event(enable_timing=True)
                e0.record()
                la.batched_matrix_multiplication(dim, count, a, b, c)
                e1.record()
                torch.cuda.synchronize()
                if r > 0:
                    times.append(e0.elapsed_time(e1))
            ms = sorted(times)[len(times) // 2]
            tf = 2.0 * dim ** 3 * count / (ms * 1e-3) / 1e12
            err = float((c.double() - torch.bmm(a.double(), b.double())).abs().max())
            out.append({"dim": dim, "count": count, "dtype": name, "ms": round(ms, 3), "tflops": round(tf, 2),
                        "peak_frac": round(tf / peak, 4), "max_err": err})
            log(f"bmm dim={dim} count={count} {name}: {ms:.3f} ms, {tf:.2f} TFLOP/s ({tf / peak:.3f} of peak), "
                f"max err {err:.1e}")
            del a, b, c
    return out


# ----------------------------------------------------------------------------- batched

def batched_bench(torch, count, replays, seed, dev):
    """The batched workload: `count` FP64 (and FP32) systems, n ~ U[8, 256], nrhs ~ U[1, 8],
    iid N(0,1) entries, solved by the ordering + persistent kernels captured in one graph.
    Median graph-replay time over `replays` (A and B restored eagerly, untimed, in between)."""
    import numpy as np
    from linearalgebra_amd.batched import BatchedSystems

    rng = np.random.Generator(np.random.PCG64(seed))
    n = rng.integers(8, 257, size=count)
    r = rng.integers(1, 9, size=count)
    nf = n.astype(np.float64)
    flops = float((2.0 / 3.0 * nf ** 3 + 2.0 * nf ** 2 * r).sum())
    out = {}
    for dt, name, peak in ((torch.float64, "f64", PEAK["fp64_mfma_tflops"]), (torch.float32, "f32", PEAK["fp32_tflops"])):
        s = BatchedSystems(n, r, dtype=dt, device=dev)
        g = torch.Generator(device=dev)
        g.manual_seed(seed)
        s.a.copy_(torch.randn(s.a.numel(), dtype=dt, device=dev, generator=g))
        s.b.copy_(torch.randn(s.b.numel(), dtype=dt, device=dev, generator=g))
        a0, b0 = s.a.clone(), s.b.clone()
        s.solve()  # eager first call (kernel set-up outside capture)
        torch.cuda.synchronize()
        graph = torch.cuda.CUDAGraph()
        s.a.copy_(a0)
        s.b.copy_(b0)
        torch.cuda.synchronize()
        with torch.cuda.graph(graph):
            s.solve()
        times = []
        for _ in range(replays):
            s.a.copy_(a0)
            s.b.copy_(b0)
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            graph.replay()
            e1.record()
            torch.cuda.synchronize()
            times.append(e0.elapsed_time(e1))
        ms = float(np.median(times))
        ok = bool((s.info.cpu() == 0).all())
        elt = 8 if dt == torch.float64 else 4
        bytes_alg = float(((2.0 * nf ** 2 + 2.0 * nf * r) * elt).sum())
        tf = flops / (ms * 1e-3) / 1e12
        out[name] = {"ms": round(ms, 3), "solves_per_s": round(count / (ms * 1e-3), 1), "tflops": round(tf, 3),
                     "roofline": {"bound": "mfma", "achieved": round(tf, 3), "peak": peak, "unit": "TFLOP/s",
                                  "frac": round(tf / peak, 4),
                                  "hbm_gbs": round(bytes_alg / (ms * 1e-3) / 1e9, 1), "hbm_frac":
                                      round(bytes_alg / (ms * 1e-3) / 1e9 / PEAK["hbm_gbs"], 4)},
                     "all_info_zero": ok, "replays": replays}
        log(f"batched {name}: {ms:.3f} ms median ({tf:.2f} TF/s)")
        del s, a0, b0, graph
    return {"workload": f"{count} independent systems, n ~ U[8,256], nrhs ~ U[1,8], iid N(0,1); ordering + "
                        f"persistent kernel captured in one graph, median of {replays} replays",
            "flops": flops, **out}


def batched_mixed_bench(torch, count, seed, dev):
    """The batched mixed-precision workload: `count` FP64 systems, n log-uniform on
    [256, 4096], each U diag(s) V^T with kappa log-uniform on [1e1, 1e8]; trailing update
    with three block-scaled FP8 e4m3 planes where kappa <= 1e4, FP32 matrix instructions
    otherwise; FP64 refinement to backward error <= 1e-14. Inputs generated on the GPU (not
    timed)."""
    import numpy as np
    from linearalgebra_amd import mixed as lam

    rng = np.random.Generator(np.random.PCG64(seed))
    n = np.exp(rng.uniform(np.log(256), np.log(4096), size=count)).astype(np.int64)
    kappa = np.exp(rng.uniform(np.log(1e1), np.log(1e8), size=count))
    a_parts, b_parts = [], []
    for i in range(count):
        ai, bi = conditioned_gpu(torch, int(n[i]), float(kappa[i]), seed + 1000 + 3 * i, dev)
        a_parts.append(ai.reshape(-1))
        b_parts.append(bi)
    a = torch.cat(a_parts)
    b = torch.cat(b_parts)
    del a_parts, b_parts
    modes = np.where(kappa <= 1e4, 6, 3).astype(np.int32)
    lam.gesv_mixed_batched(a, n[:1], b, modes=modes[:1])  # warm-up (kernel set-up)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    x, stats = lam.gesv_mixed_batched(a, n, b, modes=modes, tol=1e-14, max_iter=30, restart=30)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    berrs = []
    off_a = off_b = 0
    for i in range(count):
        k = int(n[i])
        ai = a[off_a: off_a + k * k].view(k, k)
        berrs.append(backward_error_gpu(torch, ai, x[off_b: off_b + k], b[off_b: off_b + k]))
        off_a += k * k
        off_b += k
    flops = float((2.0 / 3.0 * n.astype(np.float64) ** 3).sum())
    paths = [s["path_name"] for s in stats]
    log(f"batched mixed: {count} systems in {dt:.3f} s")
    return {"workload": f"{count} FP64 systems, n log-uniform [256,4096], kappa log-uniform [1e1,1e8], "
                        "U diag(s) V^T; three block-scaled FP8 e4m3 planes where kappa <= 1e4, FP32 MFMA otherwise; one call",
            "systems": count, "seconds": round(dt, 4), "solves_per_s": round(count / dt, 3),
            "fp64_equivalent_tflops": round(flops / dt / 1e12, 3),
            "max_backward_error": max(berrs), "paths": {p: paths.count(p) for p in sorted(set(paths))},
            "ir_iterations": {"min": min(s["ir_iterations"] for s in stats),
                              "max": max(s["ir_iterations"] for s in stats)},
            "gmres_iterations_max": max(s["gmres_iterations"] for s in stats),
            "per_system": [{"n": int(n[i]), "kappa": float(f"{kappa[i]:.3e}"), "mode": int(modes[i]),
                            "path": stats[i]["path_name"], "ir": stats[i]["ir_iterations"],
                            "gmres": stats[i]["gmres_iterations"], "berr": float(f"{berrs[i]:.3e}")}
                           for i in range(min(count, 16))],
            "note": "sequence of mixed solves inside one C call, not a fused kernel; per_system lists the first 16"}


# ----------------------------------------------------------------------------- CPU baseline

def _host_description():
    import platform
    cores = len(os.sched_getaffinity(0))
    model = platform.processor() or "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return cores, model


def _fit_cubic(ns, ts, n_target):
    """Least-squares c in t = c n^3 over the samples, and the free log-log exponent."""
    import numpy as np
    ns = np.asarray(ns, dtype=np.float64)
    ts = np.asarray(ts, dtype=np.float64)
    c = float((ts * ns ** 3).sum() / (ns ** 6).sum())
    slope = float(np.polyfit(np.log(ns), np.log(ts), 1)[0]) if len(ns) > 1 else 3.0
    return c, slope, c * float(n_target) ** 3


def cpu_baseline(sizes, n_target, seed):
    """CPU baseline on this host, in the same run as the GPU measurements.

    native: the CPU restatement of the reference's native LU + lu_solve (oracle/lu_oracle.c,
            64-wide blocked right-looking LU with partial pivoting, scalar i-k-j GEMM, every
            product rounded), single thread as the reference is single-threaded, timed at each
            order in `sizes` (nrhs = 1) and on the CPU configuration itself (n = 1024, nrhs = 16);
    lapack: LAPACK dgetrf + dgetrs through scipy (row-major A passed as the column-major
            transpose, as LAPACKE_ROW_MAJOR does) at 1 thread (the reference's published
            methodology) and on every host core, at the same orders.
    Each is extrapolated to n_target by a least-squares fit t = c n^3 (the fit and the free
    log-log exponent are reported). `value` is the native path's extrapolated solves/s."""
    sys.path.insert(0, HERE)
    import numpy as np
    from oracle import oracle as O

    O.build()
    cores, model = _host_description()
    native = []
    for n in sizes:
        a = O.gaussian(n, seed=seed + n)
        b = O.gaussian(n, 1, seed=seed + n + 1)[:, 0]
        t0 = time.perf_counter()
        _, _, x, info = O.gesv(a, b, threads=1)
        dt = time.perf_counter() - t0
        assert info == 0 and np.isfinite(x).all()
        native.append({"n": n, "nrhs": 1, "seconds": round(dt, 4), "gflops": round(2.0 * n ** 3 / 3.0 / dt / 1e9, 3)})
        log(f"cpu native n={n}: {dt:.3f} s")
    # the CPU configuration itself: n = 1024, 16 right-hand sides
    a = O.gaussian(1024, seed=seed)
    b = O.gaussian(1024, 16, seed=seed + 1)
    t0 = time.perf_counter()
    _, _, x, info = O.gesv(a, b, threads=1)
    dt0 = time.perf_counter() - t0
    assert info == 0
    config0 = {"n": 1024, "nrhs": 16, "seconds": round(dt0, 4), "solves_per_s": round(1.0 / dt0, 3),
               "gflops": round((2.0 * 1024 ** 3 / 3.0 + 2.0 * 1024 ** 2 * 16) / dt0 / 1e9, 3)}
    c, slope, t_target = _fit_cubic([r["n"] for r in native], [r["seconds"] for r in native], n_target)
    res = {"value": round(1.0 / t_target, 6), "unit": "solves/s", "cores": 1, "kind": "port",
           "sample": f"oracle getrf+getrs (FP64, partial pivoting, N(0,1), nrhs=1), 1 thread, at n = "
                     f"{', '.join(str(n) for n in sizes)}; extrapolated to n={n_target} by t = c n^3 "
                     f"(c = {c:.3e} s, free log-log exponent {slope:.2f})",
           "host": {"cores": cores, "model": model},
           "native": native, "config0": config0,
           "extrapolation": {"form": "t = c n^3, least squares", "c_seconds": c, "loglog_exponent": round(slope, 3),
                             "seconds_at_target": round(t_target, 2)}}
    try:
        import scipy
        import scipy.linalg as sl
        from threadpoolctl import threadpool_info, threadpool_limits
        libs = [f"{i.get('internal_api')} {i.get('version')}" for i in threadpool_info()]
        out = {}
        for label, nthreads in (("1_thread", 1), ("all_cores", cores)):
            rows = []
            with threadpool_limits(limits=nthreads):
                for n in sizes:
                    a = O.gaussian(n, seed=seed + n)
                    b = O.gaussian(n, 1, seed=seed + n + 1)[:, 0]
                    at = np.asfortranarray(a)  # the row-major matrix handed to LAPACK column-major
                    sl.lu_factor(at, check_finite=False)  # warm-up
                    t0 = time.perf_counter()
                    lu, piv = sl.lu_factor(at, check_finite=False)
                    sl.lu_solve((lu, piv), b, check_finite=False)
                    dt = time.perf_counter() - t0
                    rows.append({"n": n, "seconds": round(dt, 5), "gflops": round(2.0 * n ** 3 / 3.0 / dt / 1e9, 2)})
            cl, sll, tl = _fit_cubic([r["n"] for r in rows], [r["seconds"] for r in rows], n_target)
            out[label] = {"threads": nthreads, "samples": rows, "value": round(1.0 / tl, 5), "unit": "solves/s",
                          "extrapolation": {"c_seconds": cl, "loglog_exponent": round(sll, 3)}}
            log(f"cpu lapack {label}: " + ", ".join(f"n={r['n']} {r['seconds']:.3f} s" for r in rows))
        res["lapack"] = {"library": f"scipy {scipy.__version__} LAPACK ({'; '.join(sorted(set(libs)))})", **out}
    except Exception as e:  # noqa: BLE001 - optional extra
        res["lapack"] = f"unavailable: {e}"
    return res


# ----------------------------------------------------------------------------- main

def main():
    args = parse()
    if args.gpus != 1:
        sys.exit(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
                 "the multi-GPU layout of the factorisation is a design, not an implementation")
    import torch

    sys.path.insert(0, HERE)
    import linearalgebra_amd as la
    from linearalgebra_amd import _lib
    from linearalgebra_amd import mixed as lam

    lib = _lib.load()
    dev = torch.device("cuda:0")
    n, steps, warmup = args.n, args.steps, args.warmup
    log(f"generating A (n={n}, {args.matrix}, kappa={args.kappa:g}) on the GPU")
    t0 = time.perf_counter()
    if args.matrix == "conditioned":
        a, b = conditioned_gpu(torch, n, args.kappa, args.seed, dev)
    else:
        g = torch.Generator(device=dev)
        g.manual_seed(args.seed)
        a = torch.randn(n, n, dtype=torch.float64, device=dev, generator=g)
        b = torch.randn(n, dtype=torch.float64, device=dev, generator=g)
    torch.cuda.synchronize()
    log(f"inputs ready in {time.perf_counter() - t0:.1f} s")

    # ---------------- mixed LU-IR (headline)
    lam.set_mixed_update(args.mixed_update)
    gemm_kind = {0: "gemm_bf16", 1: "gemm_bf16x3", 2: "gemm_bf16x6", 3: "gemm_f32"}[args.mixed_update]
    ws = torch.empty(lam.workspace_size(n, args.restart, True), dtype=torch.uint8, device=dev)
    x = torch.empty(n, dtype=torch.float64, device=dev)
    for i in range(warmup):
        _, st = la.gesv_mixed(a, b, x, tol=1e-14, max_iter=30, restart=args.restart, workspace=ws)
        log(f"mixed warmup {i}: path={st['path']} ir={st['ir_iterations']} gmres={st['gmres_iterations']} "
            f"berr={st['backward_error']:.2e} factor={st['factor_ms']:.1f} ms refine={st['refine_ms']:.1f} ms")
    trace_start(lib, [gemm_kind])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    stats = []
    for _ in range(steps):
        _, st = la.gesv_mixed(a, b, x, tol=1e-14, max_iter=30, restart=args.restart, workspace=ws)
        stats.append(st)
    torch.cuda.synchronize()
    t_mixed = time.perf_counter() - t0
    lib.lrn_trace_stop()
    roof_mixed = roofline(lib, gemm_kind, PEAK["fp32_tflops"] if gemm_kind == "gemm_f32" else PEAK["bf16_mfma_tflops"],
                          traffic_from_profiles(gemm_kind))
    log(f"mixed: {1e3 * t_mixed / steps:.1f} ms/solve")
    berr = backward_error_gpu(torch, a, x, b)
    trace_start(lib, list(TRACE))
    la.gesv_mixed(a, b, x, tol=1e-14, max_iter=30, restart=args.restart, workspace=ws)
    torch.cuda.synchronize()
    lib.lrn_trace_stop()
    phases_mixed = phase_table(lib)
    x_mixed = x.clone()
    del ws

    # ---------------- plain FP64 LU + solve
    plain = None
    roof_fp64 = None
    phases_plain = None
    fwd = None
    if not args.no_plain:
        aw = torch.empty_like(a)
        piv = torch.empty(n, dtype=torch.int64, device=dev)
        xp = torch.empty(n, dtype=torch.float64, device=dev)
        wsp = torch.empty(la.linear_solver.workspace_size(n, la.linear_solver_type.LU_LINEAR_SOLVER),
                          dtype=torch.uint8, device=dev)

        def plain_step():
            aw.copy_(a)
            xp.copy_(b)
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record()
            la.linear_solver.apply(aw, piv, n, xp, la.linear_solver_type.LU_LINEAR_SOLVER, wsp)
            e1.record()
            return e0, e1

        for _ in range(warmup):
            plain_step()
        torch.cuda.synchronize()
        trace_start(lib, ["gemm_f64"])
        evs = [plain_step() for _ in range(steps)]
        torch.cuda.synchronize()
        lib.lrn_trace_stop()
        t_plain = sum(e0.elapsed_time(e1) for e0, e1 in evs) * 1e-3
        roof_fp64 = roofline(lib, "gemm_f64", PEAK["fp64_mfma_tflops"], traffic_from_profiles("gemm_f64"))
        log(f"plain FP64: {1e3 * t_plain / steps:.1f} ms/solve")
        berr_plain = backward_error_gpu(torch, a, xp, b)
        fwd = float((x_mixed - xp).abs().max() / xp.abs().max())
        trace_start(lib, list(TRACE))
        plain_step()
        torch.cuda.synchronize()
        lib.lrn_trace_stop()
        phases_plain = phase_table(lib)
        plain = {"solves_per_s": round(steps / t_plain, 4), "ms_per_solve": round(1e3 * t_plain / steps, 2),
                 "backward_error": berr_plain,
                 "timing": "HIP events around lrn_linear_solver_f64 (LU mode) on the launch stream; the "
                           "restore copy of A before each solve is outside the events",
                 "lu_tflops": round(2.0 * n ** 3 / 3.0 / (t_plain / steps) / 1e12, 2)}
        del aw, wsp

    plain_configs = None
    mixed_configs = None
    triangular = None
    consumers = None
    cholesky = None
    bmm = None
    if not args.no_configs:
        plain_configs = plain_configs_bench(torch, [int(v) for v in args.config_sizes.split(",") if v],
                                            (1, 64), args.seed, dev)
        kappas = [float(v) for v in args.mixed_kappas.split(",") if v]
        mixed_configs = mixed_configs_bench(torch, [n], kappas, args.seed, dev, args.restart) if kappas else []
        triangular = triangular_bench(torch, n, args.seed, dev)
        consumers = consumers_bench(torch, [int(v) for v in args.consumer_sizes.split(",") if v], args.seed, dev)
        cholesky = cholesky_bench(torch, [int(v) for v in args.consumer_sizes.split(",") if v], args.seed, dev)
        bmm = bmm_bench(torch, args.seed, dev)
        if args.mixed_extra_n > 0:
            mixed_configs += mixed_configs_bench(torch, [args.mixed_extra_n], [1e2, args.kappa] + [k for k in kappas if k > 1e6],
                                                 args.seed, dev, args.restart)

    batched = None
    if not args.no_batched:
        batched = batched_bench(torch, args.batch_count, args.batch_replays, args.seed, dev)

    batched_mixed = None
    if args.batch_mixed_count > 0:
        batched_mixed = batched_mixed_bench(torch, args.batch_mixed_count, args.seed, dev)

    cpu = None
    if not args.no_cpu:
        log(f"CPU baseline at n={args.cpu_sizes}")
        cpu = cpu_baseline([int(v) for v in args.cpu_sizes.split(",") if v], n, args.seed)

    last = stats[-1]
    value = steps / t_mixed
    line = {
        "metric": METRIC,
        "value": round(value, 4),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": round(1e3 * t_mixed / steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {0: "f32 panels + bf16 MFMA update", 1: "f32 panels + 2-plane bf16 MFMA update",
                  2: "f32 panels + 3-plane bf16 MFMA update", 3: "f32 panels + f32 MFMA update"}[args.mixed_update]
                 + " (fp32 accumulate); f64 residual and refinement",
        "data": "synthetic",
        "config": {"workload": f"mixed-precision LU-IR solve of A x = b, n={n}, nrhs=1, "
                               f"A = U diag(s) V^T (Haar U, V; s log-spaced 1..1/kappa), kappa={args.kappa:g}, "
                               f"b iid N(0,1), backward error <= 1e-14",
                   "n": n, "nrhs": 1, "kappa": args.kappa if args.matrix == "conditioned" else None,
                   "matrix": args.matrix, "seed": args.seed, "parallelism": "single GPU"},
        "backward_error": berr,
        "library_backward_error": last["backward_error"],
        "path": ["lu-ir", "gmres-ir", "fp64-fallback", "not-converged"][last["path"]],
        "ir_iterations": last["ir_iterations"],
        "gmres_iterations": last["gmres_iterations"],
        "factor_ms": round(last["factor_ms"], 2),
        "refine_ms": round(last["refine_ms"], 2),
        "forward_error_vs_fp64_lu": fwd,
        "plain_fp64": plain,
        "speedup_vs_plain_fp64": round(value / plain["solves_per_s"], 3) if plain else None,
        "roofline": roof_mixed,
        "roofline_fp64_lu": roof_fp64,
        "phases_mixed": phases_mixed,
        "phases_plain_fp64": phases_plain,
        "plain_configs": plain_configs,
        "mixed_configs": mixed_configs,
        "triangular_solver": triangular,
        "determinant_inversion": consumers,
        "cholesky": cholesky,
        "batched_matrix_multiplication": bmm,
        "batched": batched,
        "batched_mixed": batched_mixed,
        "cpu_baseline": cpu,
    }
    if not (berr <= 1e-14):
        line["invalid"] = f"backward error {berr:.3e} above 1e-14"
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()

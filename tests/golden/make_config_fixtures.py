"""Writes config_lu_n{8192,16384}_{gaussian,diag_dominant}.npz: the oracle's LU + solve at the
plain FP64 GPU workload's full orders, compacted to what the parity tests compare.

The oracle (oracle/lu_oracle.c, the CPU restatement of the reference's native blocked LU with
partial pivoting, lu_decomposition_native.cpp:158-232,247-271,329-429, and lu_solve,
linear_solver_native.cpp:90-113) takes minutes at these orders, so it runs once here and the
GPU tests on the device host compare against these files. Inputs are regenerated from the
seeds by oracle.gaussian / oracle.diag_dominant (numpy PCG64, a platform-independent stream).

Per case the file holds:
  pivot      int32[n]     the full 1-based pivot vector
  info       the oracle's getrf status
  udiag      float64[n]   the diagonal of U
  sample_ij  int32[S, 2]  seeded random positions of the packed L\\U
  sample_v   float64[S]   the packed L\\U at those positions
  lu_fro     float64      ||LU||_F of the packed factors
  x1         float64[n]   the solution for b1 = gaussian(n, 1, seed + 1)[:, 0]
  x64_rows   int32[R]     seeded random rows of the 64-column solve
  x64_cols   int32[C]     columns kept of it
  x64        float64[R, C]  X for B64 = gaussian(n, 64, seed + 2) at those rows/columns
  seed, kind, n

Usage: python tests/golden/make_config_fixtures.py [n ...]   (default 8192 16384; 8 threads)
"""
from __future__ import annotations

import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402

SEEDS = {8192: 8101, 16384: 16101}
SAMPLES = 4096
X64_ROWS = 2048
X64_COLS = (0, 31, 63)


def fixture_path(n: int, kind: str) -> str:
    return os.path.join(HERE, f"config_lu_n{n}_{kind}.npz")


def make(n: int, kind: str, threads: int = 8) -> str:
    seed = SEEDS[n]
    a = O.gaussian(n, seed=seed) if kind == "gaussian" else O.diag_dominant(n, seed=seed)
    b1 = O.gaussian(n, 1, seed=seed + 1)[:, 0]
    b64 = O.gaussian(n, 64, seed=seed + 2)
    t0 = time.time()
    lu, piv, info = O.getrf(a, threads=threads)
    del a
    x1 = O.getrs(lu, piv, b1, threads=threads)
    x64 = O.getrs(lu, piv, b64, threads=threads)
    dt = time.time() - t0
    rng = np.random.Generator(np.random.PCG64(seed + 3))
    ij = rng.integers(0, n, size=(SAMPLES, 2)).astype(np.int32)
    rows = np.sort(rng.choice(n, size=X64_ROWS, replace=False)).astype(np.int32)
    cols = np.array(X64_COLS, dtype=np.int32)
    out = fixture_path(n, kind)
    np.savez_compressed(
        out, n=n, seed=seed, kind=kind, info=int(info), pivot=piv.astype(np.int32), udiag=np.diag(lu).copy(),
        sample_ij=ij, sample_v=lu[ij[:, 0], ij[:, 1]], lu_fro=float(np.linalg.norm(lu)), x1=x1,
        x64_rows=rows, x64_cols=cols, x64=x64[np.ix_(rows, cols)], oracle_seconds=dt, oracle_threads=threads)
    print(f"{out}: info={info} oracle {dt:.1f} s", flush=True)
    return out


if __name__ == "__main__":
    sizes = [int(v) for v in sys.argv[1:]] or [8192, 16384]
    O.build()
    for n in sizes:
        for kind in ("gaussian", "diag_dominant"):
            make(n, kind)

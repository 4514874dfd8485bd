"""Autotuner of the run-time block sizes (writes ``tuning.json`` next to this file).

    python -m linearalgebra_amd.autotune [--n 16384] [--batch-count 10000] [--out PATH]

On the GPU it is run on, it times
  * the FP64 and FP32 blocked LU at order ``--n`` for every outer block width (128, 256, 384,
    512), inner panel width (32, 64) and 32-column panel kernel (register-resident rows or
    rows in LDS), and for outer widths 128 and 256 every look-ahead setting (off, 16, 32, 64
    compute units for the panel stream) with either block-column update kernel -- the
    settings of ``lrn_set_tuning`` / ``lrn_set_panel_kernel`` / ``lrn_set_lookahead`` /
    ``lrn_set_block_update``, which are process-wide, so the fastest sum of the two is kept;
  * the batched small-system solver on the uneven 10 000-system batch (n ~ U[8, 256],
    nrhs ~ U[1, 8]) for persistent grids of 1, 2 and 3 workgroups per compute unit and the
    occupancy-derived default -- the setting of ``lrn_set_batched_grid``;
and writes the fastest of each, with every measurement, to ``tuning.json``. The package
applies that file when the library is loaded (``_lib.load``); the C ABI setters take the
same values for callers that do not use the Python package.
"""
from __future__ import annotations

import argparse
import json
import os
import time

import numpy as np
import torch

from . import _lib

HERE = os.path.dirname(os.path.abspath(__file__))
TUNING_PATH = os.path.join(HERE, "tuning.json")


def _time_getrf(n: int, outer_nb: int, panel_w: int, panel_kernel: int, dtype=torch.float64, reps: int = 2,
                lookahead: int = 0, block_update: int = 0) -> float:
    import linearalgebra_amd as la
    lib = _lib.load()
    if (lib.lrn_set_tuning(outer_nb, panel_w, 0) != 0 or lib.lrn_set_panel_kernel(panel_kernel) != 0
            or lib.lrn_set_lookahead(lookahead) != 0 or lib.lrn_set_block_update(block_update) != 0):
        return float("inf")
    dev = torch.device("cuda:0")
    g = torch.Generator(device=dev)
    g.manual_seed(11)
    a = torch.randn(n, n, dtype=dtype, device=dev, generator=g)
    w = torch.empty_like(a)
    piv = torch.empty(n, dtype=torch.int64, device=dev)
    ws = torch.empty(la.lu_decomposition.workspace_size(n, dtype), dtype=torch.uint8, device=dev)
    best = float("inf")
    for _ in range(reps):
        w.copy_(a)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        try:
            la.lu_decomposition.apply(w, n, piv, ws)
        except Exception:  # an unsupported combination (e.g. too many workgroups)
            return float("inf")
        torch.cuda.synchronize()
        best = min(best, time.perf_counter() - t0)
    return best


def _time_batched(count: int, grid: int, reps: int = 5) -> float:
    from .batched import BatchedSystems
    lib = _lib.load()
    if lib.lrn_set_batched_grid(grid) != 0:
        return float("inf")
    rng = np.random.default_rng(5)
    n = rng.integers(8, 257, size=count)
    nrhs = rng.integers(1, 9, size=count)
    bs = BatchedSystems(n, nrhs, torch.float64, device="cuda:0")
    g = torch.Generator(device=bs.a.device)
    g.manual_seed(3)
    bs.a.normal_(generator=g)
    bs.b.normal_(generator=g)
    a0, b0 = bs.a.clone(), bs.b.clone()
    best = float("inf")
    for _ in range(reps):
        bs.a.copy_(a0)
        bs.b.copy_(b0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        bs.solve()
        torch.cuda.synchronize()
        best = min(best, time.perf_counter() - t0)
    return best


def tune(n: int = 16384, batch_count: int = 10000, out: str = TUNING_PATH) -> dict:
    lib = _lib.load(apply_tuning=False)
    props = torch.cuda.get_device_properties(0)
    cus = props.multi_processor_count
    getrf = []
    configs = []
    for outer_nb in (128, 256, 384, 512):
        for panel_w in (32, 64):
            for panel_kernel in ((0, 1) if panel_w == 32 else (0,)):
                configs.append((outer_nb, panel_w, panel_kernel, 0, 0))
    # look-ahead (compute units of the panel stream) and the block-column update kernel, with
    # the default 32-column register-resident panel
    for outer_nb in (128, 256):
        for lookahead in (0, 16, 32, 64):
            for block_update in (0, 1, 2):
                if (outer_nb, 32, 0, lookahead, block_update) not in configs:
                    configs.append((outer_nb, 32, 0, lookahead, block_update))
    for outer_nb, panel_w, panel_kernel, lookahead, block_update in configs:
        # the settings are process-wide: the plain FP64 and the mixed solver's FP32
        # factorisation both count
        t64 = _time_getrf(n, outer_nb, panel_w, panel_kernel, torch.float64, lookahead=lookahead,
                          block_update=block_update)
        t32 = _time_getrf(n, outer_nb, panel_w, panel_kernel, torch.float32, lookahead=lookahead,
                          block_update=block_update)
        ok = t64 != float("inf") and t32 != float("inf")
        getrf.append({"outer_nb": outer_nb, "panel_w": panel_w, "panel_kernel": panel_kernel, "lookahead": lookahead,
                      "block_update": block_update,
                      "ms_f64": round(1e3 * t64, 2) if ok else None, "ms_f32": round(1e3 * t32, 2) if ok else None,
                      "ms": round(1e3 * (t64 + t32), 2) if ok else None})
        print(f"getrf n={n} outer_nb={outer_nb} panel_w={panel_w} panel_kernel={panel_kernel} lookahead={lookahead} "
              f"block_update={block_update}: f64 {getrf[-1]['ms_f64']} ms, f32 {getrf[-1]['ms_f32']} ms", flush=True)
    lib.lrn_set_tuning(0, 0, 0)
    lib.lrn_set_panel_kernel(0)
    lib.lrn_set_lookahead(0)
    lib.lrn_set_block_update(0)
    batched = []
    for grid in (0, cus, 2 * cus, 3 * cus):
        t = _time_batched(batch_count, grid)
        batched.append({"grid": grid, "ms": round(1e3 * t, 3) if t != float("inf") else None})
        print(f"batched count={batch_count} grid={grid or 'occupancy'}: {batched[-1]['ms']} ms", flush=True)
    lib.lrn_set_batched_grid(0)
    best_g = min((r for r in getrf if r["ms"] is not None), key=lambda r: r["ms"])
    best_b = min((r for r in batched if r["ms"] is not None), key=lambda r: r["ms"])
    result = {
        "device": props.name, "compute_units": cus,
        "getrf": {"outer_nb": best_g["outer_nb"], "panel_w": best_g["panel_w"], "panel_kernel": best_g["panel_kernel"],
                  "lookahead": best_g["lookahead"], "block_update": best_g["block_update"],
                  "order": n, "ms_f64": best_g["ms_f64"], "ms_f32": best_g["ms_f32"]},
        "batched": {"grid": best_b["grid"], "count": batch_count, "ms": best_b["ms"]},
        "measurements": {"getrf": getrf, "batched": batched},
    }
    with open(out, "w") as f:
        json.dump(result, f, indent=1)
    print(f"wrote {out}: getrf {result['getrf']}, batched {result['batched']}", flush=True)
    return result


def apply(lib, path: str = TUNING_PATH) -> bool:
    """Apply a tuning file to the loaded library; False if there is none."""
    if not os.path.exists(path):
        return False
    with open(path) as f:
        t = json.load(f)
    g = t.get("getrf", {})
    lib.lrn_set_tuning(int(g.get("outer_nb", 0)), int(g.get("panel_w", 0)), 0)
    lib.lrn_set_panel_kernel(int(g.get("panel_kernel", 0)))
    lib.lrn_set_lookahead(int(g.get("lookahead", 0)))
    lib.lrn_set_block_update(int(g.get("block_update", 0)))
    lib.lrn_set_batched_grid(int(t.get("batched", {}).get("grid", 0)))
    return True


def main() -> None:
    p = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    p.add_argument("--n", type=int, default=16384)
    p.add_argument("--batch-count", type=int, default=10000)
    p.add_argument("--out", default=TUNING_PATH)
    a = p.parse_args()
    tune(a.n, a.batch_count, a.out)


if __name__ == "__main__":
    main()

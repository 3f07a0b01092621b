"""Autotuner for batched execution.

For a batch it times candidate configurations of the three knobs the batch kernels
expose and keeps the fastest:

* ``tile_x`` -- colour-sweep tile width in slots (32, 64, 128): narrow grids waste lanes
  in wide tiles, wide grids amortise the per-tile case lookup over more cells;
* ``persistent_blocks`` -- 0 (one block per tile) or a fixed grid of blocks that walks the
  tile list (a multiple of the 256 CUs);
* ``inner_iters`` -- mixed precision only: FP32 SOR iterations per FP64 defect-correction
  cycle (the colour-sweep schedule).  Only schedules with at least ``min_cycles`` cycles
  per step are offered, which bounds the FP32 rounding of a fixed-count pressure iterate.

Every configuration gives each case the same result except ``inner_iters`` (which moves
the FP32 rounding), so tuning may run on the batch it serves.  Results are cached per size
class in ``pixelflow_amd/_build/autotune_cache.json``.
"""
from __future__ import annotations

import itertools
import json
import os
import time
from typing import Dict, Iterable, List, Optional

PKG = os.path.dirname(os.path.abspath(__file__))
CACHE = os.path.join(PKG, "_build", "autotune_cache.json")
CU_COUNT = 256


def size_class_key(dim: int, bc: str, mixed: bool, iter_max: int, cells: Iterable[int]) -> str:
    cells = list(cells)
    total = sum(cells)
    # bucket the batch by case count and total cells (powers of two)
    return f"d{dim}-{bc}-{'mixed' if mixed else 'fp64'}-it{iter_max}-n{len(cells).bit_length()}-c{max(total, 1).bit_length()}"


def _load_cache() -> Dict[str, dict]:
    try:
        with open(CACHE) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store_cache(cache: Dict[str, dict]) -> None:
    os.makedirs(os.path.dirname(CACHE), exist_ok=True)
    tmp = CACHE + ".tmp"
    with open(tmp, "w") as f:
        json.dump(cache, f, indent=1, sort_keys=True)
    os.replace(tmp, CACHE)


def candidates(iter_max: int, mixed: bool, min_cycles: int = 2) -> List[dict]:
    tiles = (32, 64, 128)
    blocks = (0, 2 * CU_COUNT, 4 * CU_COUNT, 8 * CU_COUNT)
    if mixed:
        ks = sorted({max(1, (iter_max + c - 1) // c) for c in (min_cycles, min_cycles + 1, min_cycles + 2)}, reverse=True)
    else:
        ks = [0]
    return [{"tile_x": t, "persistent_blocks": b, "inner_iters": k, "use_graph": True}
            for t, b, k in itertools.product(tiles, blocks, ks)]


def tune(batch, key: str, iter_max: int, mixed: bool, steps: int = 2, use_cache: bool = True,
         min_cycles: int = 2, log=None) -> dict:
    """Pick the fastest configuration for `batch` (populated), apply it and return it with
    the timings of every candidate.  `steps` timed steps per candidate after one warm-up
    step (which also captures the graphs)."""
    cache = _load_cache() if use_cache else {}
    if use_cache and key in cache:
        best = cache[key]["best"]
        batch.set_config(**best)
        return {"best": best, "cached": True, "key": key}
    results = []
    for cfg in candidates(iter_max, mixed, min_cycles):
        batch.set_config(**cfg)
        batch.step(1, raise_diverged=False)          # also captures the graphs
        batch.sync()
        t0 = time.perf_counter()
        batch.step(steps, raise_diverged=False)      # timing only: a diverging case costs the same
        batch.sync()
        ms = (time.perf_counter() - t0) * 1e3 / steps
        results.append({**cfg, "ms_per_step": ms})
        if log:
            log(f"[autotune] {cfg} {ms:.3f} ms/step")
    best = min(results, key=lambda r: r["ms_per_step"])
    best_cfg = {k: best[k] for k in ("tile_x", "persistent_blocks", "inner_iters", "use_graph")}
    batch.set_config(**best_cfg)
    out = {"best": best_cfg, "best_ms_per_step": best["ms_per_step"], "candidates": results, "cached": False,
           "key": key}
    if use_cache:
        cache = _load_cache()
        cache[key] = {"best": best_cfg, "ms_per_step": best["ms_per_step"]}
        _store_cache(cache)
    return out

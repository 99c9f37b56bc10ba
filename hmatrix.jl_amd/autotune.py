"""Autotuner for the plan shapes of ``hm_config``.

Sweeps ``proj_chunk`` (rows of B per projection work item) and
``row_waves`` (wavefronts per 64-row tile) of the single-right-hand-side
product, times each candidate plan with HIP events on the current stream and
returns the fastest.  Results are cached in a JSON file keyed by storage,
shape and leaf statistics, so a repeated workload skips the sweep.
"""
from __future__ import annotations

import json
import os
from typing import Dict, Iterable, Optional, Tuple

import numpy as np

from .device import DeviceHmat

DEFAULT_CACHE = os.path.join(os.path.expanduser("~"), ".cache", "hmatrix_amd_autotune.json")
CHUNKS = (256, 512, 1024, 2048, 4096)
WAVES = (2, 4, 8)


def workload_key(flat, dtype: str) -> str:
    b = flat.blocks
    dense = b[:, 0] == 0
    ranks = b[~dense, 5]
    return (f"{dtype}:{flat.nrows}x{flat.ncols}:nd{int(dense.sum())}:nr{int((~dense).sum())}:"
            f"kmax{int(ranks.max()) if ranks.size else 0}:nmax{int(b[~dense, 4].max()) if ranks.size else 0}")


def _time(plan: DeviceHmat, reps: int) -> float:
    import torch
    x = torch.randn(plan.ncols, dtype=plan.vec_dtype, device=plan.device)
    y = torch.zeros(plan.nrows, dtype=plan.vec_dtype, device=plan.device)
    for _ in range(3):
        plan.matvec(x, y)
    s = torch.cuda.current_stream()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(s)
    for _ in range(reps):
        plan.matvec(x, y)
    e1.record(s)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def autotune(flat, dtype: str = "f64", chunks: Iterable[int] = CHUNKS, waves: Iterable[int] = WAVES,
             reps: int = 20, cache: Optional[str] = DEFAULT_CACHE) -> Tuple[dict, Dict[str, float]]:
    """Return ``(best_config, timings_ms)`` for single-right-hand-side products."""
    key = workload_key(flat, dtype)
    if cache and os.path.exists(cache):
        try:
            with open(cache) as fh:
                hit = json.load(fh).get(key)
            if hit:
                return hit["config"], hit["timings"]
        except (OSError, ValueError):
            pass
    timings: Dict[str, float] = {}
    best, best_ms = None, np.inf
    for c in chunks:
        for w in waves:
            cfg = {"proj_chunk": int(c), "row_waves": int(w), "no_transpose": True, "no_batched": True}
            p = DeviceHmat(flat, dtype, cfg)
            ms = _time(p, reps)
            p.close()
            timings[f"chunk{c}_waves{w}"] = ms
            if ms < best_ms:
                best, best_ms = {"proj_chunk": int(c), "row_waves": int(w)}, ms
    if cache:
        try:
            os.makedirs(os.path.dirname(cache), exist_ok=True)
            data = {}
            if os.path.exists(cache):
                with open(cache) as fh:
                    data = json.load(fh)
            data[key] = {"config": best, "timings": timings}
            with open(cache, "w") as fh:
                json.dump(data, fh, indent=1)
        except (OSError, ValueError):
            pass
    return best, timings

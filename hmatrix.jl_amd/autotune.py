"""Autotuner for the plan shapes of ``hm_config``.

``autotune`` sweeps ``proj_chunk`` (rows of B per projection work item) and
``row_waves`` (wavefronts per 64-row tile) of the single-right-hand-side
product; ``autotune_batch`` the same, and the one-launch form, for a
many-matrix plan (``DeviceHmat.batch``); ``autotune_batched`` sweeps
``strip_shape`` (workgroups of the
multi-right-hand-side row pass: 8 strips with a 512-row X window, or 16
with 1024 rows).  Each times every candidate plan with HIP events on the
current stream and returns the fastest.  Results are cached in a JSON file
keyed by storage, shape and leaf statistics, so a repeated workload skips
the sweep.
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


def _load_cache(cache: Optional[str], key: str):
    if cache and os.path.exists(cache):
        try:
            with open(cache) as fh:
                return json.load(fh).get(key)
        except (OSError, ValueError):
            return None
    return None


def _store_cache(cache: Optional[str], key: str, entry: dict) -> None:
    if not cache:
        return
    try:
        os.makedirs(os.path.dirname(cache), exist_ok=True)
        data = {}
        if os.path.exists(cache):
            with open(cache) as fh:
                data = json.load(fh)
        data[key] = entry
        with open(cache, "w") as fh:
            json.dump(data, fh, indent=1)
    except (OSError, ValueError):
        pass


def autotune(flat, dtype: str = "f64", chunks: Iterable[int] = CHUNKS, waves: Iterable[int] = WAVES,
             reps: int = 20, cache: Optional[str] = DEFAULT_CACHE) -> Tuple[dict, Dict[str, float]]:
    """Return ``(best_config, timings_ms)`` for single-right-hand-side products."""
    key = workload_key(flat, dtype)
    hit = _load_cache(cache, key)
    if hit:
        return hit["config"], hit["timings"]
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
    _store_cache(cache, key, {"config": best, "timings": timings})
    return best, timings


def autotune_batch(flats, dtype: str = "f32", chunks: Iterable[int] = CHUNKS, waves: Iterable[int] = WAVES,
                   one_launch: Iterable[bool] = (False, True), reps: int = 20,
                   cache: Optional[str] = DEFAULT_CACHE) -> Tuple[dict, Dict[str, float]]:
    """Return ``(best_config, timings_ms)`` for single-right-hand-side
    products of the many-matrix plan of ``flats`` (all members per launch):
    ``proj_chunk`` x ``row_waves`` x (three launches or ``one_launch``).  The
    best config (with ``no_transpose`` / ``no_batched`` dropped) is what
    ``DeviceHmat.batch(flats, dtype, config)`` takes."""
    key = "batch:" + "|".join(workload_key(f, dtype) for f in flats)
    hit = _load_cache(cache, key)
    if hit:
        return hit["config"], hit["timings"]
    timings: Dict[str, float] = {}
    best, best_ms = None, np.inf
    for c in chunks:
        for w in waves:
            for one in one_launch:
                cfg = {"proj_chunk": int(c), "row_waves": int(w), "no_transpose": True, "no_batched": True,
                       "one_launch": bool(one)}
                p = DeviceHmat.batch(flats, dtype, cfg)
                ms = _time(p, reps)
                p.close()
                timings[f"chunk{c}_waves{w}_one{int(bool(one))}"] = ms
                if ms < best_ms:
                    best, best_ms = {"proj_chunk": int(c), "row_waves": int(w), "one_launch": bool(one)}, ms
    _store_cache(cache, key, {"config": best, "timings": timings})
    return best, timings


SHAPES = (1, 2)     # hm_strip_shape: narrow, wide


def autotune_batched(flat, dtype: str = "bf16x2", shapes: Iterable[int] = SHAPES, nrhs: int = 32,
                     reps: int = 20, cache: Optional[str] = DEFAULT_CACHE) -> Tuple[dict, Dict[str, float]]:
    """Return ``(best_config, timings_ms)`` for ``nrhs``-column products
    (FP32, BF16 or split-BF16 storage): one plan per ``strip_shape``, each
    timed over ``reps`` products of the same seeded right-hand sides."""
    import torch
    if dtype == "f64":
        raise ValueError("the batched path needs FP32, BF16 or split-BF16 storage")
    key = "batched:" + workload_key(flat, dtype) + f":nrhs{nrhs}"
    hit = _load_cache(cache, key)
    if hit:
        return hit["config"], hit["timings"]
    timings: Dict[str, float] = {}
    best, best_ms = None, np.inf
    for sh in shapes:
        p = DeviceHmat(flat, dtype, {"strip_shape": int(sh), "no_transpose": True})
        X = p.new_cols(flat.ncols, nrhs)
        X.copy_(torch.randn(flat.ncols, nrhs, generator=torch.Generator().manual_seed(0)).to(X.device))
        Y = p.new_cols(flat.nrows, nrhs)
        for _ in range(3):
            p.matmat(X, Y)
        s = torch.cuda.current_stream()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(s)
        for _ in range(reps):
            p.matmat(X, Y)
        e1.record(s)
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        p.close()
        timings[f"strip_shape{sh}"] = ms
        if ms < best_ms:
            best, best_ms = {"strip_shape": int(sh)}, ms
    _store_cache(cache, key, {"config": best, "timings": timings})
    return best, timings


def autotune_native(flat, dtype: str = "f64", nrhs: int = 1, reps: int = 20) -> Tuple[dict, float]:
    """The library's own sweep (``hm_autotune``, the entry point Julia and C
    callers use): ``nrhs == 1`` sweeps ``proj_chunk`` x ``row_waves``,
    ``nrhs > 1`` the ``strip_shape`` (FP64: the f64 MFMA layout against
    panels of 8 columns on the single-column layout, whose ``proj_chunk`` x
    ``row_waves`` it sweeps; ``no_batched`` is set when those won).
    Returns ``(config, ms_per_product)``."""
    import ctypes
    from . import _native
    L = _native.lib()
    b = np.ascontiguousarray(flat.blocks, dtype=np.int64)
    d = np.ascontiguousarray(flat.data, dtype=np.float64)
    cfg = _native.HmConfig()
    ms = ctypes.c_double(0.0)
    rc = L.hm_autotune(flat.nrows, flat.ncols, b.shape[0], b.ctypes.data_as(ctypes.c_void_p),
                       d.ctypes.data_as(ctypes.c_void_p), d.size, _native.STORAGE[dtype], int(nrhs), int(reps),
                       ctypes.byref(cfg), ctypes.byref(ms))
    _native.check(rc, "hm_autotune")
    if nrhs == 1:
        out = {"proj_chunk": cfg.proj_chunk, "row_waves": cfg.row_waves}
    elif cfg.flags & 2:        # HM_NO_BATCHED: the single-column layout won
        out = {"proj_chunk": cfg.proj_chunk, "row_waves": cfg.row_waves, "no_batched": True}
    else:
        out = {"strip_shape": cfg.strip_shape}
    return out, ms.value

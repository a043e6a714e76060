"""Tile-edge and work-item-target tuning of the sample-lane binned kernels.

No reference counterpart: the reference has no binned kernels to tune (its forward
and adjoint fork over the batch and gather / index_add per sample,
``_nufft/interp.py:235-265,587-726``).  Here the binned gather and spread cut the grid
into tiles and the hot tiles into work items of ``work_target`` entries; the best
tile edge and target depend on the grid, the sample density and the number of rows.
``tune`` times every candidate on one problem of a geometry with HIP events (plan
built once, then reused: plan cost is not part of the choice), keeps the fastest,
and caches it per geometry class in this process (and, when ``TKBN_AUTOTUNE_CACHE``
names a JSON file, across processes).
"""
from __future__ import annotations

import json
import math
import os
import threading
from typing import Dict, List, Optional, Sequence, Tuple

import torch
from torch import Tensor

from . import interp as _interp

Choice = Tuple[Tuple[int, int, int], int]  # (tile edges, work_target)

_LOCK = threading.Lock()
_CACHE: Dict[str, Choice] = {}
_LOADED = False
_LDS_ROW_BYTES = 60 * 1024  # one-row LDS staging limit of an overridden tile

# candidate tile edges per dimension count and direction (0 = the library's choice)
_TILES = {
    (1, "fwd"): [(0, 0, 0), (256, 0, 0), (1024, 0, 0), (2048, 0, 0)],
    (1, "adj"): [(0, 0, 0), (256, 0, 0), (1024, 0, 0), (2048, 0, 0)],
    (2, "fwd"): [(0, 0, 0), (16, 16, 0), (32, 32, 0), (48, 48, 0), (32, 64, 0)],
    (2, "adj"): [(0, 0, 0), (8, 8, 0), (16, 16, 0), (32, 32, 0), (16, 32, 0)],
    (3, "fwd"): [(0, 0, 0), (8, 8, 8), (16, 16, 8), (16, 16, 12)],
    (3, "adj"): [(0, 0, 0), (8, 8, 4), (8, 8, 8), (16, 16, 12)],
}
_TARGETS = [0, 256, 512, 1024, 2048]


def _cache_path() -> Optional[str]:
    return os.environ.get("TKBN_AUTOTUNE_CACHE") or None


def _load() -> None:
    global _LOADED
    if _LOADED:
        return
    _LOADED = True
    path = _cache_path()
    if path and os.path.exists(path):
        try:
            with open(path) as f:
                raw = json.load(f)
            for k, v in raw.items():
                _CACHE[k] = (tuple(int(t) for t in v[0]), int(v[1]))
        except (OSError, ValueError, TypeError, IndexError):
            pass  # an unreadable cache file only costs a re-tune


def _save() -> None:
    path = _cache_path()
    if not path:
        return
    tmp = f"{path}.{os.getpid()}.tmp"
    with open(tmp, "w") as f:
        json.dump({k: [list(v[0]), v[1]] for k, v in _CACHE.items()}, f, indent=0, sort_keys=True)
    os.replace(tmp, path)


def geometry_key(direction: str, grid_size: Sequence[int], numpoints: Sequence[int], table_oversamp: Sequence[int],
                 rows: int, klength: int, deterministic: bool) -> str:
    """Geometry class of a choice: the sample count enters as its power of two."""
    mb = int(round(math.log2(max(int(klength), 1))))
    return (f"{direction}|K={'x'.join(str(int(v)) for v in grid_size)}|J={'x'.join(str(int(v)) for v in numpoints)}"
            f"|L={'x'.join(str(int(v)) for v in table_oversamp)}|rows={int(rows)}|M~2^{mb}|det={int(deterministic)}")


def cached(key: str) -> Optional[Choice]:
    with _LOCK:
        _load()
        return _CACHE.get(key)


def choices() -> Dict[str, Choice]:
    """Every choice made or loaded in this process."""
    with _LOCK:
        _load()
        return dict(_CACHE)


def _fits(tile: Sequence[int], grid_size: Sequence[int], numpoints: Sequence[int], complex_bytes: int) -> bool:
    """One row of the tile fits the LDS staging (the library refuses it otherwise)."""
    if not any(tile):
        return True
    region, cells = complex_bytes, 16
    for i in range(len(grid_size)):
        t = max(int(tile[i]) if tile[i] > 0 else 1, int(numpoints[i]) - 1)
        t = min(t, int(grid_size[i]))
        region *= t + int(numpoints[i]) - 1
        cells *= t
    return max(region, cells) <= _LDS_ROW_BYTES


def candidates(direction: str, grid_size: Sequence[int], numpoints: Sequence[int], deterministic: bool,
               complex_bytes: int = 8) -> List[Choice]:
    """(tile, work_target) pairs tried for a geometry (``complex_bytes``: 8 for complex64)."""
    d = len(grid_size)
    tiles = [t for t in _TILES[(d, direction)] if _fits(t, grid_size, numpoints, complex_bytes)]
    # the deterministic spread gives each tile one work item: the target plays no part
    targets = [0] if (direction == "adj" and deterministic) else _TARGETS
    out = []
    for t in tiles:
        for w in targets:
            out.append((tuple(t), w))
    return out


def _time_calls(fn, reps: int) -> float:
    """Median of ``reps`` HIP-event-timed calls of ``fn`` on the current stream (ms)."""
    times = []
    for _ in range(reps):
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record()
        fn()
        b.record()
        b.synchronize()
        times.append(a.elapsed_time(b))
    times.sort()
    return times[len(times) // 2]


def tune(direction: str, data: Tensor, omega: Tensor, tables: List[Tensor], n_shift, numpoints, table_oversamp,
         offsets, grid_size: Sequence[int], deterministic: bool = False, reps: int = 9) -> Choice:
    """Fastest (tile, work_target) for the binned ``direction`` ("fwd" / "adj") kernel on
    this problem: ``data`` is the grid (B, C, *K) for "fwd" and the samples (B, C, M)
    for "adj".  Cached per geometry class; a cached choice is returned without timing.
    The default choice wins unless another is at least 5 % faster (timing noise)."""
    from ._host import host_values

    jj = [int(v) for v in host_values(numpoints)]
    ll = [int(v) for v in host_values(table_oversamp)]
    gs = [int(v) for v in grid_size]
    rows = int(data.shape[0]) * int(data.shape[1])
    key = geometry_key(direction, gs, jj, ll, rows, omega.shape[-1], deterministic)
    hit = cached(key)
    if hit is not None:
        return hit
    best: Optional[Choice] = None
    best_ms = math.inf
    default_ms = math.inf
    out = torch.empty((data.shape[0], data.shape[1], omega.shape[-1]) if direction == "fwd" else
                      (data.shape[0], data.shape[1]) + tuple(gs), dtype=data.dtype, device=data.device)
    for tile, target in candidates(direction, gs, jj, deterministic, data.element_size()):
        holder = _interp.PlanHolder(trusted=True, tile=tile, work_target=target)
        if direction == "fwd":
            def call(h=holder):
                _interp.table_interp(data, omega, tables, n_shift, numpoints, table_oversamp, offsets, out=out,
                                     plan=h, binned=True)
        else:
            def call(h=holder):
                _interp.table_interp_adjoint(data, omega, tables, n_shift, numpoints, table_oversamp, offsets, gs,
                                             deterministic=deterministic, out=out, plan=h)
        call()  # builds the plan
        ms = _time_calls(call, reps)
        if tile == (0, 0, 0) and target == 0:
            default_ms = ms
        if ms < best_ms:
            best, best_ms = (tile, target), ms
    if best is None or best_ms > 0.95 * default_ms:
        best = ((0, 0, 0), 0)
    with _LOCK:
        _CACHE[key] = best
        try:
            _save()
        except OSError:
            pass  # the in-process cache still holds the choice
    return best


def clear() -> None:
    """Forget every choice held by this process (the cache file is left alone)."""
    global _LOADED
    with _LOCK:
        _CACHE.clear()
        _LOADED = True

"""Tile-size autotuner for the batched kernel.

For a :class:`~torchquad_amd.batch.Batch` it times every tile_points
candidate on the device with HIP events, keeps the fastest, and caches the
choice per (dtype, integrand-family mix, dimension range, size bucket).  The cache lives in memory and, when ``TQ_AUTOTUNE_CACHE`` names a
file, in that JSON file (so a serving process can reuse earlier searches).
"""
import json
import math
import os
import threading

import torch

TILE_CANDIDATES = (16384, 32768, 65536, 131072)

_cache = {}
_lock = threading.Lock()


def _cache_file():
    return os.environ.get("TQ_AUTOTUNE_CACHE")


def _load():
    path = _cache_file()
    if path and os.path.exists(path):
        try:
            with open(path) as f:
                _cache.update(json.load(f))
        except (OSError, ValueError):
            pass


def _store():
    path = _cache_file()
    if not path:
        return
    try:
        with open(path + ".tmp", "w") as f:
            json.dump(_cache, f)
        os.replace(path + ".tmp", path)
    except OSError:
        pass


def signature(batch):
    fams = ",".join(str(int(f)) for f in sorted(set(batch.families.tolist())))
    dims = f"{int(batch.dims.min())}-{int(batch.dims.max())}" if batch.count else "0-0"
    size = int(math.log2(max(batch.evals, 1)))
    return f"{batch.dtype}|{fams}|{dims}|{size}"


def time_batch(batch, reps=3):
    """Median ms of `reps` runs (after one warm-up)."""
    batch.run()
    torch.cuda.synchronize(batch.device)
    times = []
    for _ in range(reps):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        batch.run()
        e1.record()
        e1.synchronize()
        times.append(e0.elapsed_time(e1))
    times.sort()
    return times[len(times) // 2]


def autotune(batch, tiles=TILE_CANDIDATES, reps=3, use_cache=True):
    """Pick tile_points for `batch`, apply it and return (tile_points, timing table)."""
    key = signature(batch)
    with _lock:
        if not _cache:
            _load()
        hit = _cache.get(key) if use_cache else None
    if hit is not None:
        batch.retile(tile_points=hit["tile_points"])
        return hit["tile_points"], hit.get("table", [])
    table = []
    best = None
    for tp in tiles:
        batch.retile(tile_points=tp)
        ms = time_batch(batch, reps)
        table.append((tp, ms))
        if best is None or ms < best[1]:
            best = (tp, ms)
    batch.retile(tile_points=best[0])
    with _lock:
        _cache[key] = {"tile_points": best[0], "ms": best[1], "table": table}
        _store()
    return best[0], table


# ---- plain fused grid kernel: items per wave -------------------------------
IPW_CANDIDATES = (0, 256, 512, 768, 1024, 1280, 1536, 1792, 2048, 2304, 2560, 3072, 4096, 8192)


def _grid_key(rule, family, dim, n, dtype):
    return f"grid|{int(rule)}|{int(family)}|{int(dim)}|{int(n)}|{dtype}"


_file_tried = False  # grid_ipw looked for TQ_AUTOTUNE_CACHE once


def grid_ipw(rule, family, dim, n, dtype):
    """Tuned items-per-wave of the fused grid kernel for this geometry, or 0 (automatic)."""
    global _file_tried
    key = _grid_key(rule, family, dim, n, dtype)
    hit = _cache.get(key)
    if hit is None and not _cache and not _file_tried:
        _file_tried = True
        if _cache_file():
            with _lock:
                _load()
            hit = _cache.get(key)
    return int(hit["items_per_wave"]) if hit else 0


def set_grid_ipw(rule, family, dim, n, dtype, items_per_wave):
    """Fix the fused grid kernel's items per wave for this geometry (as if tuned)."""
    with _lock:
        _cache[_grid_key(rule, family, dim, n, dtype)] = {"items_per_wave": int(items_per_wave), "ms": None,
                                                           "table": []}


def tune_grid(rule, fn, dim, n, domain, candidates=IPW_CANDIDATES, reps=5, use_cache=True):
    """Time the fused grid kernel of (rule, fn, dim, n) on `domain` for each
    items-per-wave candidate (HIP events around `reps` back-to-back calls,
    best of three such batches after a warm-up),
    cache the fastest for later calls with the same geometry and return
    (items_per_wave, table)."""
    from . import _ops

    family = fn.descriptor(dim).family
    key = _grid_key(rule, family, dim, n, domain.dtype)
    with _lock:
        if not _cache:
            _load()
        hit = _cache.get(key) if use_cache else None
    if hit is not None:
        return int(hit["items_per_wave"]), hit.get("table", [])
    gl = None
    if int(rule) == 3:
        from .gl_nodes import gauss_legendre

        gl = gauss_legendre(int(n), device=domain.device)
    table = []
    for ipw in candidates:
        for _ in range(2):
            _ops.grid_integrate(rule, fn, dim, n, domain, gl, 0, ipw)
        torch.cuda.synchronize(domain.device)
        # back-to-back calls between two events, so the device never waits on
        # the host's launch path inside the timed span; best of three batches
        times = []
        for _ in range(3):
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(reps):
                _ops.grid_integrate(rule, fn, dim, n, domain, gl, 0, ipw)
            e1.record()
            e1.synchronize()
            times.append(e0.elapsed_time(e1) / reps)
        table.append((ipw, min(times)))
    best = min(table, key=lambda r: r[1])
    with _lock:
        _cache[key] = {"items_per_wave": best[0], "ms": best[1], "table": table}
        _store()
    return best[0], table

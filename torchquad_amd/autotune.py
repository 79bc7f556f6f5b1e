"""Tile-size / grid-size autotuner for the batched kernel.

For a :class:`~torchquad_amd.batch.Batch` it times every (tile_points,
blocks) candidate on the device with HIP events, keeps the fastest, and
caches the choice per (dtype, integrand-family mix, dimension range, size
bucket).  The cache lives in memory and, when ``TQ_AUTOTUNE_CACHE`` names a
file, in that JSON file (so a serving process can reuse earlier searches).
"""
import json
import math
import os
import threading

import torch

TILE_CANDIDATES = (8192, 16384, 32768, 65536, 131072)
BLOCK_CANDIDATES = (1024, 2048, 4096)

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


def autotune(batch, tiles=TILE_CANDIDATES, blocks=BLOCK_CANDIDATES, reps=3, use_cache=True):
    """Pick (tile_points, blocks) for `batch`, apply it and return it with the timing table."""
    key = signature(batch)
    with _lock:
        if not _cache:
            _load()
        hit = _cache.get(key) if use_cache else None
    if hit is not None:
        batch.retile(tile_points=hit["tile_points"], blocks=hit["blocks"])
        return (hit["tile_points"], hit["blocks"]), hit.get("table", [])
    table = []
    best = None
    for tp in tiles:
        batch.retile(tile_points=tp)
        for bl in blocks:
            batch.retile(blocks=bl)
            ms = time_batch(batch, reps)
            table.append((tp, bl, ms))
            if best is None or ms < best[2]:
                best = (tp, bl, ms)
    batch.retile(tile_points=best[0], blocks=best[1])
    with _lock:
        _cache[key] = {"tile_points": best[0], "blocks": best[1], "ms": best[2], "table": table}
        _store()
    return (best[0], best[1]), table

"""Launch-shape autotuner for the persistent batched kernel.

For a batch signature (precision, families, the multiset of (n_dim,
n_events) rounded to powers of two, and the device's compute-unit count) it
times each candidate (events per tile, workgroups per compute unit, pair
histograms or not) on scratch buffers that do not touch the batch's state,
keeps the fastest, and
remembers it in memory and, when ``MIVEGAS_TUNE_CACHE`` names a file, in that
JSON file for later processes.
"""
import json
import os
import threading

import torch

EVENTS_PER_TILE = (1024, 2048, 4096, 8192, 16384)
BLOCKS_PER_CU = (2, 4, 8)
# tiles of the pair-histogram kernel (1024-thread workgroups, as many as are
# resident): larger, since a tile clears and sums its pair tables once
PAIR_EVENTS_PER_TILE = (16384, 65536, 262144)

_memory = {}
_lock = threading.Lock()


def _cache_path():
    return os.environ.get("MIVEGAS_TUNE_CACHE")


def _load_file():
    path = _cache_path()
    if not path or not os.path.exists(path):
        return {}
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store_file(key, value):
    path = _cache_path()
    if not path:
        return
    data = _load_file()
    data[key] = value
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def lookup(key):
    with _lock:
        if key in _memory:
            return _memory[key]
        hit = _load_file().get(key)
        if hit is not None:
            _memory[key] = hit
        return hit


def remember(key, value):
    with _lock:
        _memory[key] = value
        _store_file(key, value)


def time_ms(launch, reps=2):
    """Average device time of ``launch()`` over ``reps`` runs after one
    warm-up, from HIP events on the current stream."""
    stream = torch.cuda.current_stream()
    launch()
    start = torch.cuda.Event(enable_timing=True)
    end = torch.cuda.Event(enable_timing=True)
    start.record(stream)
    for _ in range(reps):
        launch()
    end.record(stream)
    end.synchronize()
    return start.elapsed_time(end) / reps


def tune(key, candidates, measure):
    """Return the candidate with the smallest ``measure(candidate)`` (ms),
    cached under ``key``. Candidates are JSON-serialisable lists."""
    hit = lookup(key)
    if hit is not None and list(hit["choice"]) in [list(c) for c in candidates]:
        return tuple(hit["choice"])
    timings = []
    for cand in candidates:
        timings.append((measure(cand), tuple(cand)))
    best_ms, best = min(timings)
    remember(key, {"choice": list(best), "ms": best_ms,
                   "timings": [[list(c), t] for t, c in timings]})
    return best

"""Autotuner of the multi-instance engine.

For a batch it times graph replays of the whole solve over the engine's
shape knobs and keeps the fastest, caching the choice in JSON keyed by the
device's CU count and a signature of the batch (its instances' n, m and k).
The knobs (emv_engine_params):

  relax_blocks_per_cu  relaxation grid: workgroups per CU
  merge_rows_per_wave  rows one wave merges for its vertex tile (0 = by layer size)
  dense_rounds         vertex-wise rounds that take every item at a layer's start
  split_items          upper-layer split merge: items per merge tile (sets parts per row)

The search is coordinate descent from the defaults: each knob in turn over
its candidates with the others at their best so far, one pass.  Every shape
must give identical costs (checked); only the time differs.

    params, table = tune_batch(batch, instances)   # sets the params on batch
"""
import hashlib
import json
import os
import time

import numpy as np

KNOBS = [
    ("relax_blocks_per_cu", (16, 32, 64, 128)),
    ("merge_rows_per_wave", (0, 1, 2, 4, 8)),
    ("dense_rounds", (3, 5, 6, 8)),
    ("split_items", (16, 32, 64, 128)),
]
DEFAULTS = dict(relax_blocks_per_cu=64, merge_rows_per_wave=0, dense_rounds=6, split_items=32)


def cache_path():
    p = os.environ.get("EMV_TUNE_CACHE")
    if p:
        return p
    base = os.environ.get("XDG_CACHE_HOME") or os.path.join(os.path.expanduser("~"), ".cache")
    return os.path.join(base, "emv-mi355x", "tune.json")


def signature(instances, num_cu):
    h = hashlib.sha1()
    h.update(f"v2;{num_cu}".encode())
    for n, e, kk in instances:
        h.update(f"{n},{len(e)},{len(kk)};".encode())
    return h.hexdigest()[:16]


def _load(path):
    try:
        with open(path) as fh:
            return json.load(fh)
    except (OSError, ValueError):
        return {}


def _store(path, data):
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        tmp = path + ".tmp"
        with open(tmp, "w") as fh:
            json.dump(data, fh, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except OSError:
        pass


def _apply(batch, shape, use_graph):
    batch.set_params(use_graph=use_graph, **shape)


def tune_batch(batch, instances, num_cu=256, reps=1, use_cache=True, log=None):
    """Returns (best shape as a dict, table of [shape, ms] tried)."""
    path = cache_path()
    key = signature(instances, num_cu)
    cache = _load(path) if use_cache else {}
    if key in cache:
        shape = cache[key]["shape"]
        _apply(batch, shape, 0)
        batch.solve()  # record the rounds of this shape for the replays
        _apply(batch, shape, 1)
        return shape, cache[key]["table"]
    seen, table = {}, []
    ref = [None]

    def timed(shape):
        tag = tuple(sorted(shape.items()))
        if tag in seen:
            return seen[tag]
        _apply(batch, shape, 0)
        costs, _ = batch.solve()  # eager: records the rounds of this shape
        if ref[0] is None:
            ref[0] = costs
        elif not np.array_equal(ref[0], costs):
            raise RuntimeError(f"autotune: results differ at {shape}")
        _apply(batch, shape, 1)
        best = float("inf")
        for _ in range(reps):
            t0 = time.perf_counter()
            costs, st = batch.solve()
            best = min(best, (time.perf_counter() - t0) * 1e3)
            if not np.array_equal(ref[0], costs):
                raise RuntimeError(f"autotune: replay results differ at {shape}")
        seen[tag] = best
        table.append([dict(shape), round(best, 2)])
        if log:
            log(f"autotune: {shape}: {best:.1f} ms")
        return best

    shape = dict(DEFAULTS)
    for name, cands in KNOBS:
        times = {}
        for c in cands:
            trial = dict(shape, **{name: c})
            times[c] = timed(trial)
        shape[name] = min(times, key=times.get)
    _apply(batch, shape, 0)
    batch.solve()  # re-record the rounds for the chosen shape
    _apply(batch, shape, 1)
    if use_cache:
        cache[key] = {"shape": shape, "table": table}
        _store(path, cache)
    return shape, table

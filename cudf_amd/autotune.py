"""Autotuner of the batched group-by's knobs (``BatchOptions``): hash-table
load factor, the LDS-vs-global strategy threshold, block size and rows per
work item.

The search is coordinate descent over a small grid, timed on a strided
subsample of the batch's instances (so a tuning run costs a fraction of one
batch), each candidate measured as the median of a few graph replays after
a warm-up.  The chosen options are cached per workload signature (the
distribution of instance row counts, the key/value types and the number of
aggregations), so repeated batches of the same shape reuse them.
"""
from __future__ import annotations

import math
import time
from dataclasses import replace
from typing import Dict, List, Optional, Sequence, Tuple

import torch

from .batch import BatchedGroupBy, BatchOptions
from .column import Column
from .types import NullPolicy

GRID = {
    "load_factor": [0.5, 0.35, 0.7],
    "lds_max_groups": [512, -1, 128, 768],
    "block_size": [256, 512],
    "tile_rows": [4096, 2048, 8192, 16384],
    "partition_min_rows": [0, -1],
}

_cache: Dict[tuple, BatchOptions] = {}


def signature(instances: Sequence[Tuple[Column, Sequence[Column]]], aggregations) -> tuple:
    """Workload shape: log2 histogram of row counts, key and value types,
    aggregation kinds."""
    hist: Dict[int, int] = {}
    for k, _ in instances:
        b = int(math.log2(max(1, len(k))))
        hist[b] = hist.get(b, 0) + 1
    k0, v0 = instances[0]
    types = (int(k0.type_id),) + tuple(int(v.type_id) for v in v0)
    kinds = tuple(tuple(int(a.kind()) for a in al) for al in aggregations)
    return tuple(sorted(hist.items())), types, kinds


def _time_plan(instances, aggregations, null_handling, opts: BatchOptions, reps: int) -> float:
    plan = BatchedGroupBy(instances, aggregations, null_handling, opts)
    try:
        plan.run()  # warm-up (also settles any table regrowth)
        times = []
        for _ in range(reps):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            plan.run()
            times.append(time.perf_counter() - t0)
        times.sort()
        return times[len(times) // 2]
    finally:
        plan.close()


def tune(instances: Sequence[Tuple[Column, Sequence[Column]]], aggregations,
         null_handling: NullPolicy = NullPolicy.EXCLUDE, sample: int = 8, reps: int = 3,
         start: Optional[BatchOptions] = None, verbose: bool = False) -> BatchOptions:
    """Options for batches shaped like ``instances`` (every ``sample``-th
    instance is timed; ``sample=1`` times the whole batch)."""
    key = signature(instances, aggregations) + (int(null_handling),)
    if key in _cache:
        return _cache[key]
    sub = list(instances[::max(1, sample)])
    best = start or BatchOptions(load_factor=GRID["load_factor"][0], lds_max_groups=GRID["lds_max_groups"][0],
                                 block_size=GRID["block_size"][0], tile_rows=GRID["tile_rows"][0])
    best_t = _time_plan(sub, aggregations, null_handling, best, reps)
    log: List[tuple] = [(best, best_t)]
    for _ in range(2):  # two sweeps of coordinate descent
        improved = False
        for knob, values in GRID.items():
            for v in values:
                if getattr(best, knob) == v:
                    continue
                cand = replace(best, **{knob: v})
                t = _time_plan(sub, aggregations, null_handling, cand, reps)
                log.append((cand, t))
                if t < best_t * 0.98:
                    best, best_t, improved = cand, t, True
        if not improved:
            break
    if verbose:
        for o, t in log:
            print(f"  {o}: {t * 1e3:.2f} ms")
    _cache[key] = best
    return best


def tuning_log_keys() -> List[tuple]:
    return list(_cache.keys())

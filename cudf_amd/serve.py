"""Served group-by: a request loop with continuous batching.

Queries arrive at any time (``submit`` returns a future).  Their inputs are
device columns that stay resident in HBM; the server's workers each own a HIP
stream and, whenever one is idle, it takes the next unit of work:

* the oldest pending query decides the unit.  If it is batchable (one
  fixed-width key column, fixed-width values, hash-path aggregations, at most
  ``batch_rows`` rows), every pending batchable query with the same request
  template (aggregations per request and null policy; key and value types may
  differ) joins it, up to ``max_batch`` queries, and the unit runs as one
  batched plan (``BatchedGroupBy``: one persistent kernel over all of them,
  replayed from a captured hipGraph);
* otherwise the unit is that query alone, run through the single-call path
  (``GroupBy.aggregate``).

Admission is budgeted by device memory: a unit is formed only while the
working memory it is estimated to need (hash tables, accumulators, outputs)
fits under ``memory_budget`` together with the units already running;
queries that do not fit wait, in arrival order.  Queries that arrive while
others are on the GPU are admitted as workers and memory free up, and the
units of different workers overlap on the device.

A query's result is the same whether it runs alone or in a batch (the
batched plan computes each instance as the single call would; floating-point
sums differ only by summation order, within the parity tolerances).
"""
from __future__ import annotations

import threading
import time
from concurrent.futures import Future
from dataclasses import dataclass, field, replace
from typing import List, Optional, Sequence, Tuple

import torch

from . import _lib
from .aggregation import Kind
from .batch import BatchedGroupBy, BatchOptions
from .column import Table
from .groupby import GroupBy, GroupByRequest
from .types import NullPolicy, TypeId

_BATCH_KINDS = {Kind.SUM, Kind.SUM_OF_SQUARES, Kind.MIN, Kind.MAX, Kind.COUNT_VALID, Kind.COUNT_ALL,
                Kind.MEAN, Kind.VARIANCE, Kind.STD}
_NUMERIC = {TypeId.INT8, TypeId.INT16, TypeId.INT32, TypeId.INT64, TypeId.UINT8, TypeId.UINT16,
            TypeId.UINT32, TypeId.UINT64, TypeId.FLOAT32, TypeId.FLOAT64, TypeId.BOOL8}


def _fixed_width(t: TypeId) -> bool:
    return t != TypeId.STRING and t != TypeId.EMPTY


def _agg_ok(t: TypeId, k: Kind) -> bool:
    if k in (Kind.COUNT_VALID, Kind.COUNT_ALL):
        return True
    if k in (Kind.MIN, Kind.MAX):
        return t in _NUMERIC or (TypeId.TIMESTAMP_DAYS <= t <= TypeId.DURATION_NANOSECONDS)
    return t in _NUMERIC


@dataclass
class Query:
    """One group-by: ``GroupBy(keys, null_policy).aggregate(requests)``."""

    keys: Table
    requests: List[GroupByRequest]
    null_policy: NullPolicy = NullPolicy.EXCLUDE
    tag: object = None
    arrival: float = 0.0          # set by submit (time.perf_counter)
    started: float = 0.0
    finished: float = 0.0
    batched: bool = False
    batch_size: int = 1
    future: Future = field(default_factory=Future)

    def rows(self) -> int:
        return self.keys.num_rows()

    def template(self) -> tuple:
        return (int(self.null_policy),
                tuple(tuple((int(a.kind()), a.ddof) for a in r.aggregations) for r in self.requests))

    def batchable(self, batch_rows: int) -> bool:
        cols = self.keys.columns()
        if len(cols) != 1 or not _fixed_width(cols[0].type_id) or self.rows() > batch_rows:
            return False
        for r in self.requests:
            t = r.values.type_id
            if not _fixed_width(t) or len(r.values) != self.rows():
                return False
            for a in r.aggregations:
                if a.kind() not in _BATCH_KINDS or not _agg_ok(t, a.kind()):
                    return False
        return True

    def memory_estimate(self) -> int:
        """Working bytes on the device: tables and accumulators for up to one
        group per row at load 1/2, outputs, partition buffers."""
        n_aggs = sum(len(r.aggregations) for r in self.requests)
        return self.rows() * (64 + 24 * n_aggs) + (1 << 20)


class GroupByServer:
    """Continuous-batching group-by server on one GPU (see the module doc)."""

    def __init__(self, workers: int = 8, memory_budget: int = 200 * 10**9, max_batch: int = 1024,
                 batch_rows: int = 4_000_000, options: Optional[BatchOptions] = None,
                 device: Optional[int] = None):
        self.workers = workers
        self.memory_budget = int(memory_budget)
        self.max_batch = max_batch
        self.batch_rows = batch_rows
        self.options = options
        self.device = torch.cuda.current_device() if device is None else device
        _lib.lib()  # load the library before the workers start
        self._pending: List[Query] = []
        self._mem_in_use = 0
        self._cv = threading.Condition()
        self._stop = False
        self._threads: List[threading.Thread] = []
        self.completed: List[Query] = []
        self.units = 0
        self.batched_units = 0

    # -- lifecycle ----------------------------------------------------------------
    def start(self) -> "GroupByServer":
        for i in range(self.workers):
            t = threading.Thread(target=self._worker, name=f"groupby-worker-{i}", daemon=True)
            t.start()
            self._threads.append(t)
        return self

    def stop(self) -> None:
        with self._cv:
            self._stop = True
            self._cv.notify_all()
        for t in self._threads:
            t.join()
        self._threads.clear()

    def __enter__(self):
        return self.start()

    def __exit__(self, *exc):
        self.stop()

    # -- requests -----------------------------------------------------------------
    def submit(self, query: Query) -> Future:
        """Queue a query; the future resolves to (unique keys, [results])."""
        query.arrival = query.arrival or time.perf_counter()
        with self._cv:
            if self._stop:
                raise RuntimeError("server stopped")
            self._pending.append(query)
            self._cv.notify()
        return query.future

    # -- scheduling -----------------------------------------------------------------
    def _next_unit(self) -> Optional[Tuple[List[Query], int]]:
        """Under the lock: the next unit of work and its memory reservation,
        or None when nothing is admissible now."""
        if not self._pending:
            return None
        head = self._pending[0]
        need = head.memory_estimate()
        if self._mem_in_use > 0 and self._mem_in_use + need > self.memory_budget:
            return None  # wait for memory (a query larger than the budget runs alone)
        if not head.batchable(self.batch_rows):
            self._pending.pop(0)
            return [head], need
        tpl = head.template()
        unit, keep = [head], []
        for q in self._pending[1:]:
            extra = q.memory_estimate()
            if (len(unit) < self.max_batch and q.template() == tpl and q.batchable(self.batch_rows)
                    and self._mem_in_use + need + extra <= self.memory_budget):
                unit.append(q)
                need += extra
            else:
                keep.append(q)
        self._pending = keep
        return unit, need

    def _worker(self) -> None:
        torch.cuda.set_device(self.device)
        stream = torch.cuda.Stream()
        while True:
            with self._cv:
                unit = None
                while not self._stop:
                    unit = self._next_unit()
                    if unit is not None:
                        break
                    self._cv.wait()
                if unit is None:
                    return
                queries, need = unit
                self._mem_in_use += need
                self.units += 1
            try:
                with torch.cuda.stream(stream):
                    self._run(queries, stream)
            finally:
                with self._cv:
                    self._mem_in_use -= need
                    self._cv.notify_all()

    def _run(self, queries: List[Query], stream) -> None:
        t0 = time.perf_counter()
        for q in queries:
            q.started = t0
        if len(queries) > 1:
            try:
                self._run_batch(queries, stream)
                return
            except Exception:
                # a batch that cannot be planned falls back to one call each,
                # so that each query gets its own result or its own error
                pass
        for q in queries:
            try:
                out = GroupBy(q.keys, q.null_policy).aggregate(q.requests, stream=stream)
                stream.synchronize()
                q.finished = time.perf_counter()
                q.batched, q.batch_size = False, 1
                self._done(q, out, None)
            except Exception as e:  # noqa: BLE001 - the query's own error
                q.finished = time.perf_counter()
                self._done(q, None, e)

    def _run_batch(self, queries: List[Query], stream) -> None:
        aggs = [list(r.aggregations) for r in queries[0].requests]
        inst = [(q.keys.columns()[0], [r.values for r in q.requests]) for q in queries]
        # a unit's plan runs once: launching its kernels directly is cheaper
        # than capturing a graph for one replay
        opts = replace(self.options or BatchOptions(), use_graph=-1)
        plan = BatchedGroupBy(inst, aggs, queries[0].null_policy, opts, stream=stream)
        try:
            plan.run(stream)
            outs = [plan.result(i) for i in range(len(queries))]
            stream.synchronize()
        finally:
            plan.close()
        self.batched_units += 1
        t = time.perf_counter()
        for q, out in zip(queries, outs):
            q.finished = t
            q.batched, q.batch_size = True, len(queries)
            self._done(q, out, None)

    def _done(self, q: Query, out, err) -> None:
        with self._cv:
            self.completed.append(q)
        if err is not None:
            q.future.set_exception(err)
        else:
            q.future.set_result(out)


def latency_stats(queries: Sequence[Query]) -> dict:
    """p50 / p99 / mean of (finish - arrival) in milliseconds, and the
    achieved rate over the span from the first arrival to the last finish."""
    import numpy as np
    lat = np.array([q.finished - q.arrival for q in queries]) * 1e3
    span = max(q.finished for q in queries) - min(q.arrival for q in queries)
    return {"p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
            "mean_ms": float(lat.mean()), "max_ms": float(lat.max()),
            "queries_per_s": len(queries) / span if span > 0 else float("inf"),
            "batched_fraction": float(np.mean([q.batched for q in queries])),
            "mean_batch": float(np.mean([q.batch_size for q in queries]))}

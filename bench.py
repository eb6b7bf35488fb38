"""Benchmark of the plain hash group-by on one MI355X.

Workload (the high-cardinality configuration): 100M rows; int64 key uniform
over 10M values with 1% null keys (excluded); float64 N(0,1) and int32
U[0,1e4] value columns, each 5% null; sum, mean, count, min, max, var
(ddof=1) on both.  Inputs are generated on the device and stay resident in
HBM; a step is one `GroupBy(keys).aggregate(requests)` call including the
materialised output columns.

Prints one JSON line.  Every timed step starts with cold caches (a 512 MB
write between steps, outside the timing).  `roofline` divides the
operation's algorithmic bytes (each input buffer read once, each output
buffer written once) by the dominant kernel's time, measured with hipEvents
the library records on the stream it launches on; `pipeline` gives the same
ratio for the whole step.  `cpu_baseline` times pandas group-by (the
cudf.pandas CPU fallback) on a bounded sample of the same distribution,
median of 3; `oracle_baseline` times the C oracle on that sample.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "groupby rows/s + %HBM roofline at 100M rows; batched vs plain queries/s, p99"
HBM_PEAK = 8.0e12
N_ROWS = 100_000_000
CARD = 10_000_000


def pack_bits(valid):
    import torch
    n = valid.numel()
    pad = (-n) % 32
    v = torch.nn.functional.pad(valid.to(torch.int64), (0, pad)).view(-1, 32)
    w = (v << torch.arange(32, device=valid.device, dtype=torch.int64)).sum(1)
    w = torch.where(w >= 2**31, w - 2**32, w)
    return w.to(torch.int32).contiguous()


def make_c2(n, card, seed=0):
    import torch
    from cudf_amd import Column, TypeId
    dev = torch.device("cuda", 0)
    g = torch.Generator(device=dev).manual_seed(seed)
    key = torch.randint(0, card, (n,), device=dev, dtype=torch.int64, generator=g)
    kvalid = torch.rand(n, device=dev, generator=g) >= 0.01
    f64 = torch.randn(n, device=dev, dtype=torch.float64, generator=g)
    fvalid = torch.rand(n, device=dev, generator=g) >= 0.05
    i32 = torch.randint(0, 10_001, (n,), device=dev, dtype=torch.int32, generator=g)
    ivalid = torch.rand(n, device=dev, generator=g) >= 0.05

    def col(t, valid, tid):
        nulls = int((~valid).sum().item())
        return Column(tid, n, t.view(torch.uint8), pack_bits(valid), nulls)

    cols = (col(key, kvalid, TypeId.INT64), col(f64, fvalid, TypeId.FLOAT64), col(i32, ivalid, TypeId.INT32))
    del kvalid, fvalid, ivalid
    return cols, (key, f64, i32)


def input_bytes(n):
    return n * (8 + 8 + 4) + 3 * 4 * ((n + 31) // 32)


def output_bytes(key_t, res):
    from cudf_amd.types import width
    total = 0
    for c in list(key_t.columns()) + [c for t in res for c in t.columns()]:
        total += c.size * width(c.type_id)
        if c.mask is not None:
            total += 4 * ((c.size + 31) // 32)
    return total


def cpu_baselines(sample_rows, card, seed=1, reps=3):
    """The CPU baselines on a bounded sample of the same distribution (same
    generator, same aggregations): pandas group-by -- the CPU fallback the
    reference's cudf.pandas layer runs -- with nullable Int64 / Float64 /
    Int32 columns, one warm-up and the median of `reps` runs; and the C
    oracle (one thread, one run) as a secondary figure."""
    from oracle import oracle as orc
    from tests.gbutil import C2_AGGS, c2_data
    key, f64, i32 = c2_data(n=sample_rows, card=card, seed=seed)
    out = {}
    try:
        import pandas as pd
        df = pd.DataFrame({"k": pd.arrays.IntegerArray(key.data, ~key.valid),
                           "f": pd.arrays.FloatingArray(f64.data, ~f64.valid),
                           "i": pd.arrays.IntegerArray(i32.data, ~i32.valid)})
        aggs = ["sum", "mean", "count", "min", "max", "var"]
        df.groupby("k", sort=False, dropna=True).agg({"f": aggs, "i": aggs})  # warm-up
        times = []
        for _ in range(reps):
            t0 = time.perf_counter()
            df.groupby("k", sort=False, dropna=True).agg({"f": aggs, "i": aggs})
            times.append(time.perf_counter() - t0)
        t_pd = float(np.median(times))
        out["cpu_baseline"] = {"value": sample_rows / t_pd, "unit": "rows/s", "cores": 1, "kind": "reference",
                               "library": f"pandas {pd.__version__} (the cudf.pandas CPU fallback; "
                                          f"group-by is single-threaded)",
                               "cpu_count": os.cpu_count(),
                               "sample": f"{sample_rows} rows, {card} distinct keys, C2's distributions and "
                                         f"aggregations, nullable Int64/Float64/Int32, groupby(sort=False, "
                                         f"dropna=True).agg; median of {reps} runs after one warm-up: "
                                         f"{t_pd:.2f} s"}
        del df
    except ImportError:
        pass
    t0 = time.perf_counter()
    orc.groupby([key], [(f64, C2_AGGS), (i32, C2_AGGS)], 0)
    t_oracle = time.perf_counter() - t0
    out["oracle_baseline"] = {"value": sample_rows / t_oracle, "unit": "rows/s", "cores": 1, "kind": "port",
                              "sample": f"the same {sample_rows} rows, C oracle (oracle/groupby_oracle.c), "
                                        f"1 thread, one run: {t_oracle:.2f} s"}
    return out


def c4_shapes(n_inst, seed=4):
    """Row counts log-uniform 1e3..4e6, cardinality log-uniform 1..rows/2."""
    rng = np.random.default_rng(seed)
    rows = np.round(np.exp(rng.uniform(np.log(1e3), np.log(4e6), n_inst))).astype(np.int64)
    card = np.maximum(1, np.round(np.exp(rng.uniform(0.0, np.log(0.5 * rows))))).astype(np.int64)
    return rows, card


def make_c4(n_inst, seed=4):
    """The batched workload on the device: per instance an int32 key uniform
    over [0, card), an int64 value uniform in [-1e9, 1e9] and a float32 value
    uniform in [0, 1); instances are slices of three concatenated buffers."""
    import torch
    from cudf_amd import Column, TypeId
    rows, card = c4_shapes(n_inst, seed)
    total = int(rows.sum())
    dev = torch.device("cuda", 0)
    g = torch.Generator(device=dev).manual_seed(seed)
    keys = torch.empty(total, dtype=torch.int32, device=dev)
    starts = np.concatenate([[0], np.cumsum(rows)[:-1]])
    for s, r, c in zip(starts, rows, card):
        keys[s:s + r].random_(0, int(c), generator=g)
    v0 = torch.randint(-1_000_000_000, 1_000_000_001, (total,), dtype=torch.int64, device=dev, generator=g)
    v1 = torch.rand(total, dtype=torch.float32, device=dev, generator=g)
    # each instance's columns start at their own data pointer (row offset 0):
    # the concatenated buffers hold ~2e9 rows, beyond cudf's int32 row offsets
    inst = []
    for s, r in zip(starts, rows):
        s, r = int(s), int(r)
        inst.append((Column(TypeId.INT32, r, keys[s:s + r].view(torch.uint8), None, 0, 0),
                     [Column(TypeId.INT64, r, v0[s:s + r].view(torch.uint8), None, 0, 0),
                      Column(TypeId.FLOAT32, r, v1[s:s + r].view(torch.uint8), None, 0, 0)]))
    return inst, rows, card, (keys, v0, v1)


def bench_batched(n_inst, reps, autotune=True, plain=True):
    """Batched form against the plain one-call-per-instance loop on the same
    instances (sum, count, min, max on both values)."""
    import torch
    import cudf_amd as ca
    from cudf_amd import autotune as at
    from cudf_amd.aggregation import Aggregation, Kind
    inst, rows, card, keep = make_c4(n_inst)
    torch.cuda.synchronize()
    mk = lambda: [Aggregation(Kind.SUM), Aggregation(Kind.COUNT_VALID), Aggregation(Kind.MIN),
                  Aggregation(Kind.MAX)]
    aggs = [mk(), mk()]
    t0 = time.perf_counter()
    opts = at.tune(inst, aggs, sample=16) if autotune else ca.BatchOptions()
    t_tune = time.perf_counter() - t0
    t0 = time.perf_counter()
    plan = ca.BatchedGroupBy(inst, aggs, ca.NullPolicy.EXCLUDE, opts)
    t_plan = time.perf_counter() - t0
    plan.run()  # warm-up
    times = []
    for _ in range(reps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        plan.run()
        times.append(time.perf_counter() - t0)
    t_batch = float(np.median(times))
    groups = sum(plan.num_groups(i) for i in range(n_inst))
    plan.close()
    total = int(rows.sum())
    in_b = total * (4 + 8 + 4)
    out_b = groups * (4 + 8 + 4 + 8 + 8 + 4 + 4 + 4 + 4)
    res = {
        "workload": f"{n_inst} independent group-bys in one plan: rows log-uniform 1e3..4e6 "
                    f"(total {total}), cardinality log-uniform 1..rows/2 (total {groups} groups), int32 "
                    f"keys, int64 and float32 values, sum/count/min/max on both",
        "instances": n_inst, "rows": total, "groups": groups,
        "options": {k: getattr(opts, k) for k in ("load_factor", "lds_max_groups", "block_size",
                                                   "tile_rows")},
        "autotune_s": round(t_tune, 3), "plan_create_s": round(t_plan, 3),
        "ms_per_batch": t_batch * 1e3,
        "queries_per_s": n_inst / t_batch,
        "queries_per_s_with_plan": n_inst / (t_batch + t_plan),
        "rows_per_s": total / t_batch,
        "bytes_alg": in_b + out_b,
        "hbm_frac": (in_b + out_b) / t_batch / HBM_PEAK,
    }
    if plain:
        reqs = lambda k, vs: (ca.GroupBy(ca.Table([k])),
                              [ca.GroupByRequest(vs[0], mk()), ca.GroupByRequest(vs[1], mk())])
        for k, vs in inst[:8]:  # warm-up
            gb, rq = reqs(k, vs)
            gb.aggregate(rq)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for k, vs in inst:
            gb, rq = reqs(k, vs)
            out = gb.aggregate(rq)
            del out
        torch.cuda.synchronize()
        t_plain = time.perf_counter() - t0
        res["plain_s"] = t_plain
        res["plain_queries_per_s"] = n_inst / t_plain
        res["speedup_vs_plain"] = t_plain / t_batch
        res["speedup_vs_plain_with_plan"] = t_plain / (t_batch + t_plan)
    del keep
    torch.cuda.empty_cache()
    return res


SERVE_TYPES = ["int32", "int64", "float32", "float64"]


def _dev_values(g, n, dt, lo, hi):
    import torch
    dev = torch.device("cuda", 0)
    if dt.startswith("int"):
        return torch.randint(lo, hi + 1, (n,), dtype=getattr(torch, dt), device=dev, generator=g)
    return torch.randn(n, dtype=getattr(torch, dt), device=dev, generator=g)


def _dev_col(t, valid=None):
    import torch
    from cudf_amd import Column, TypeId
    tid = {torch.int32: TypeId.INT32, torch.int64: TypeId.INT64, torch.float32: TypeId.FLOAT32,
           torch.float64: TypeId.FLOAT64}[t.dtype]
    if valid is None:
        return Column(tid, t.numel(), t.view(torch.uint8))
    return Column(tid, t.numel(), t.view(torch.uint8), pack_bits(valid), int((~valid).sum().item()))


def make_served(n_req, seed=5):
    """The served request stream: templates 1-3 drawn uniformly, rows
    log-uniform 1e4..5e7, random key/value dtypes, 0-5% nulls.  Returns
    (query factories, per-request (template, rows))."""
    import torch
    import cudf_amd as ca
    from cudf_amd import Column, TypeId
    from cudf_amd.aggregation import Aggregation, Kind
    rng = np.random.default_rng(seed)
    dev = torch.device("cuda", 0)
    g = torch.Generator(device=dev).manual_seed(seed)
    A = lambda *ks: [Aggregation(k, 1) for k in ks]
    # the C3 string pool and Zipf(1.1) table over 1M values
    lens = rng.integers(4, 33, 10_000)
    pool_chars = torch.from_numpy(rng.integers(33, 127, int(lens.sum()), dtype=np.uint8)).to(dev)
    pool_off = torch.from_numpy(np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)).to(dev)
    pool_len = torch.from_numpy(lens.astype(np.int64)).to(dev)
    w = np.arange(1, 1_000_001, dtype=np.float64) ** -1.1
    zipf_cdf = torch.from_numpy(np.cumsum(w) / w.sum()).to(dev)
    factories, shapes = [], []
    for q in range(n_req):
        tpl = int(rng.integers(1, 4))
        n = int(round(np.exp(rng.uniform(np.log(1e4), np.log(5e7)))))
        nulls = lambda: (torch.rand(n, device=dev, generator=g) >= rng.uniform(0, 0.05))
        if tpl == 1:
            kt, vt = rng.choice(SERVE_TYPES, 2)
            key = torch.randint(0, min(1000, n), (n,), device=dev, generator=g).to(getattr(torch, kt))
            keys = [_dev_col(key, nulls())]
            reqs = [(_dev_col(_dev_values(g, n, vt, -10**6, 10**6), nulls()),
                     (Kind.SUM, Kind.COUNT_VALID, Kind.MIN, Kind.MAX))]
        elif tpl == 2:
            kt, ft, it = rng.choice(SERVE_TYPES, 3)
            key = torch.randint(0, max(1, n // 10), (n,), device=dev, generator=g).to(getattr(torch, kt))
            keys = [_dev_col(key, torch.rand(n, device=dev, generator=g) >= 0.01)]
            c2 = (Kind.SUM, Kind.MEAN, Kind.COUNT_VALID, Kind.MIN, Kind.MAX, Kind.VARIANCE)
            reqs = [(_dev_col(_dev_values(g, n, ft, -10**6, 10**6), nulls()), c2),
                    (_dev_col(_dev_values(g, n, it, 0, 10**4), nulls()), c2)]
        else:
            ft, it = rng.choice(SERVE_TYPES, 2)
            k1 = torch.randint(1, 101, (n,), dtype=torch.int32, device=dev, generator=g)
            k2 = torch.searchsorted(zipf_cdf, torch.rand(n, dtype=torch.float64, device=dev, generator=g)) + 1
            idx = torch.randint(0, 10_000, (n,), device=dev, generator=g)
            ln = pool_len[idx]
            offs = torch.zeros(n + 1, dtype=torch.int64, device=dev)
            offs[1:] = torch.cumsum(ln, 0)
            total = int(offs[-1].item())
            src = torch.repeat_interleave(pool_off[idx] - offs[:-1], ln, output_size=total)
            chars = pool_chars[src + torch.arange(total, device=dev)]
            k3 = Column(TypeId.STRING, n, chars, None, 0, 0, offs.to(torch.int32))
            del src, ln, idx
            keys = [_dev_col(k1), _dev_col(k2.to(torch.int64)), k3]
            reqs = [(_dev_col(_dev_values(g, n, ft, -10**9, 10**9), nulls()), (Kind.SUM, Kind.COUNT_VALID, Kind.MAX)),
                    (_dev_col(_dev_values(g, n, it, -10**9, 10**9), nulls()), (Kind.SUM, Kind.COUNT_VALID, Kind.MAX))]

        def factory(keys=keys, reqs=reqs):
            from cudf_amd.serve import Query
            return Query(ca.Table(keys), [ca.GroupByRequest(c, A(*ks)) for c, ks in reqs])
        factories.append(factory)
        shapes.append((tpl, n))
    torch.cuda.synchronize()
    return factories, shapes


def make_c3(n, seed=3):
    """Skewed multi-column keys on the device: key1 int32 uniform 1..100, key2
    int64 Zipf(1.1) over 1M values, key3 a string from a pool of 10k random
    ASCII strings of length 4..32; float32 U[0,1) and int64 U[-1e9,1e9]
    values; sum, count, max on both."""
    import torch
    import cudf_amd as ca
    from cudf_amd import Column, TypeId
    from cudf_amd.aggregation import Aggregation, Kind
    rng = np.random.default_rng(seed)
    dev = torch.device("cuda", 0)
    g = torch.Generator(device=dev).manual_seed(seed)
    lens = rng.integers(4, 33, 10_000)
    pool_chars = torch.from_numpy(rng.integers(33, 127, int(lens.sum()), dtype=np.uint8)).to(dev)
    pool_off = torch.from_numpy(np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)).to(dev)
    pool_len = torch.from_numpy(lens.astype(np.int64)).to(dev)
    w = np.arange(1, 1_000_001, dtype=np.float64) ** -1.1
    zipf_cdf = torch.from_numpy(np.cumsum(w) / w.sum()).to(dev)
    k1 = torch.randint(1, 101, (n,), dtype=torch.int32, device=dev, generator=g)
    k2 = torch.searchsorted(zipf_cdf, torch.rand(n, dtype=torch.float64, device=dev, generator=g)) + 1
    idx = torch.randint(0, 10_000, (n,), device=dev, generator=g)
    ln = pool_len[idx]
    offs = torch.zeros(n + 1, dtype=torch.int64, device=dev)
    offs[1:] = torch.cumsum(ln, 0)
    total = int(offs[-1].item())
    src = torch.repeat_interleave(pool_off[idx] - offs[:-1], ln, output_size=total)
    chars = pool_chars[src + torch.arange(total, device=dev)]
    del src, ln, idx
    k3 = Column(TypeId.STRING, n, chars, None, 0, 0, offs.to(torch.int32))
    f32 = torch.rand(n, dtype=torch.float32, device=dev, generator=g)
    i64 = torch.randint(-10**9, 10**9 + 1, (n,), dtype=torch.int64, device=dev, generator=g)
    keys = ca.Table([_dev_col(k1), _dev_col(k2.to(torch.int64)), k3])
    aggs = lambda: [Aggregation(Kind.SUM), Aggregation(Kind.COUNT_VALID), Aggregation(Kind.MAX)]
    reqs = [ca.GroupByRequest(_dev_col(f32), aggs()), ca.GroupByRequest(_dev_col(i64), aggs())]
    return keys, reqs, (chars, offs, k1, k2, f32, i64)


def bench_c3(n=50_000_000, steps=3):
    import torch
    import cudf_amd as ca
    from cudf_amd import _lib
    keys, reqs, keep = make_c3(n)
    gb = ca.GroupBy(keys)
    _lib.set_phase_timing(True)
    out = gb.aggregate(reqs)
    del out
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        k, r = gb.aggregate(reqs)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    st = _lib.last_stats()
    return {"workload": f"{n} rows, keys (int32 1..100, int64 Zipf(1.1) over 1M, string of 4..32 chars "
                        f"from 10k), float32 and int64 values, sum/count/max",
            "rows": n, "groups": k.num_rows(), "ms_per_step": ms, "rows_per_s": n / (ms * 1e-3),
            "last_call": {kk: v for kk, v in st.items()}}


def bench_partition(n=100_000_000, card=10_000_000, reps=5, shards=8):
    """Hash partitioning of a 100M-row table (int64 key over 10M values, float64
    value) into 8 and 256 partitions; then the C2 group-by run the way a
    multi-device group-by would (8 row shards aggregated separately, partials
    hash-partitioned by key into 8 parts and combined per part), on one device."""
    import torch
    import cudf_amd as ca
    from cudf_amd.dask_groupby import groupby_agg
    from cudf_amd.partitioning import concatenate, hash_partition, split
    g = torch.Generator(device="cuda").manual_seed(1)
    keys = torch.randint(0, card, (n,), device="cuda", dtype=torch.int64, generator=g)
    vals = torch.rand(n, device="cuda", dtype=torch.float64, generator=g)
    t = ca.Table([ca.Column(ca.TypeId.INT64, n, keys.view(torch.uint8)),
                  ca.Column(ca.TypeId.FLOAT64, n, vals.view(torch.uint8))])

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            r = fn()
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t0)
            del r
        return sorted(ts)[len(ts) // 2] * 1e3

    out = {"workload": f"{n} rows: int64 key over {card} values, float64 value (16 B/row)"}
    for nparts in (8, 256):
        ms = timed(lambda: hash_partition(t, [0], nparts))
        out[f"hash_partition_{nparts}"] = {"ms": ms, "rows_per_s": n / (ms * 1e-3),
                                           "alg_GBps": n * 32 / (ms * 1e-3) / 1e9}
    part, offs = hash_partition(t, [0], 8)
    pieces = split(part, offs)
    ms = timed(lambda: concatenate(pieces))
    out["concatenate_8"] = {"ms": ms, "alg_GBps": n * 32 / (ms * 1e-3) / 1e9}
    del part, pieces, t, keys, vals
    cols, keep = make_c2(n, card)
    step = n // shards
    parts = [(ca.Table([cols[0].slice(i * step, step)]),
              {"f64": cols[1].slice(i * step, step), "i32": cols[2].slice(i * step, step)}) for i in range(shards)]
    aggs = {"f64": ["sum", "count", "mean"], "i32": ["max"]}
    ms = timed(lambda: groupby_agg(parts, aggs, split_out=shards))
    out["sharded_c2"] = {"shards": shards, "split_out": shards, "ms": ms, "rows_per_s": n / (ms * 1e-3),
                         "aggs": aggs}
    return out


def bench_sorted(n=100_000_000, card=10_000_000, reps=5):
    """The C2 request (float64 and int32 values, 5% nulls, sum/mean/count/min/
    max/var) over an int64 key that is sorted (10M distinct values): the
    pre-sorted path (keys_are_sorted=YES, groups = runs) against the hash path
    on the same data."""
    import torch
    import cudf_amd as ca
    from cudf_amd import Column, TypeId
    from cudf_amd.aggregation import Aggregation, Kind
    dev = torch.device("cuda", 0)
    g = torch.Generator(device=dev).manual_seed(7)
    key = torch.sort(torch.randint(0, card, (n,), device=dev, dtype=torch.int64, generator=g)).values
    f64 = torch.randn(n, device=dev, dtype=torch.float64, generator=g)
    fvalid = torch.rand(n, device=dev, generator=g) >= 0.05
    i32 = torch.randint(0, 10_001, (n,), device=dev, dtype=torch.int32, generator=g)
    ivalid = torch.rand(n, device=dev, generator=g) >= 0.05
    kcol = Column(TypeId.INT64, n, key.view(torch.uint8))
    fcol = Column(TypeId.FLOAT64, n, f64.view(torch.uint8), pack_bits(fvalid), int((~fvalid).sum().item()))
    icol = Column(TypeId.INT32, n, i32.view(torch.uint8), pack_bits(ivalid), int((~ivalid).sum().item()))
    del fvalid, ivalid
    aggs = lambda: [Aggregation(Kind.SUM), Aggregation(Kind.MEAN), Aggregation(Kind.COUNT_VALID),
                    Aggregation(Kind.MIN), Aggregation(Kind.MAX), Aggregation(Kind.VARIANCE, 1)]
    reqs = [ca.GroupByRequest(fcol, aggs()), ca.GroupByRequest(icol, aggs())]
    out = {"workload": f"{n} rows, sorted int64 key over {card} values, float64 and int32 values "
                       f"(5% nulls), sum/mean/count/min/max/var"}
    for name, flag in (("sorted_path", ca.Sorted.YES), ("hash_path", ca.Sorted.NO)):
        gb = ca.GroupBy(ca.Table([kcol]), ca.NullPolicy.EXCLUDE, flag)
        r = gb.aggregate(reqs)
        groups = r[0].num_rows()
        del r
        torch.cuda.synchronize()
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            r = gb.aggregate(reqs)
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t0)
            del r
        ms = sorted(ts)[len(ts) // 2] * 1e3
        out[name] = {"ms": ms, "rows_per_s": n / (ms * 1e-3), "groups": groups}
    # sort-only aggregations on the same runs: NUNIQUE of the int32 values and
    # their first / last valid value (NTH_ELEMENT 0 and -1, nulls excluded)
    sort_only = [ca.GroupByRequest(icol, [Aggregation(Kind.NUNIQUE),
                                          Aggregation(Kind.NTH_ELEMENT, null_handling=ca.NullPolicy.EXCLUDE, n=0),
                                          Aggregation(Kind.NTH_ELEMENT, null_handling=ca.NullPolicy.EXCLUDE, n=-1)])]
    gb = ca.GroupBy(ca.Table([kcol]), ca.NullPolicy.EXCLUDE, ca.Sorted.YES)
    r = gb.aggregate(sort_only)
    del r
    torch.cuda.synchronize()
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        r = gb.aggregate(sort_only)
        torch.cuda.synchronize()
        ts.append(time.perf_counter() - t0)
        del r
    ms = sorted(ts)[len(ts) // 2] * 1e3
    out["sorted_nunique_first_last"] = {"ms": ms, "rows_per_s": n / (ms * 1e-3)}
    return out


def result_digest(query, out):
    """Order-independent digest of a result's exact columns: per column, the
    wrapping sum over groups of a mix of the group's key row hash
    (MurmurHash3 of the key columns) with the element's bits and validity.
    Exact columns are COUNT, MIN, MAX and integer SUM; floating-point sums,
    means and variances are left out (their rounding depends on the order of
    accumulation and is checked against the oracle by the tests)."""
    import torch
    from cudf_amd import murmurhash3_x86_32
    from cudf_amd.aggregation import Kind
    from cudf_amd.types import numpy_dtype
    keys_t, res = out
    n = keys_t.num_rows()
    rh = murmurhash3_x86_32(keys_t).data[:4 * n].view(torch.int32).to(torch.int64) & 0xFFFFFFFF
    digests = []
    for req, t in zip(query.requests, res):
        for agg, c in zip(req.aggregations, t.columns()):
            dt = np.dtype(numpy_dtype(c.type_id))
            if not (agg.kind() in (Kind.COUNT_VALID, Kind.COUNT_ALL, Kind.MIN, Kind.MAX) or
                    (agg.kind() == Kind.SUM and dt.kind in "iu")):
                continue
            raw = c.data[:dt.itemsize * n]
            bits = (raw.view(torch.int64) if dt.itemsize == 8 else
                    raw.view(torch.int32).to(torch.int64) if dt.itemsize == 4 else
                    raw.view(torch.int16).to(torch.int64) if dt.itemsize == 2 else raw.to(torch.int64))
            valid = torch.ones(n, dtype=torch.int64, device=bits.device)
            if c.mask is not None:
                words = c.mask.view(torch.int32).to(torch.int64) & 0xFFFFFFFF
                idx = torch.arange(n, device=bits.device)
                valid = (words[idx >> 5] >> (idx & 31)) & 1
            bits = bits * valid  # a null's bits are not part of the result
            # odd 64-bit multipliers as signed int64 (torch wraps on overflow)
            h = (rh * (0x9E3779B97F4A7C15 - (1 << 64)) + bits) * (0xBF58476D1CE4E5B9 - (1 << 64)) + valid
            digests.append(int(h.sum().item()))
    return (n, tuple(digests))


def bench_served(n_req, loads=(0.5, 1.0, 2.0, 4.0), workers=8):
    """Plain kernel run serially over the stream, then the continuous-batching
    server at Poisson arrival rates that are multiples of the serial rate and
    with every request arriving at once; p50/p99 latency and achieved rate."""
    import threading
    import torch
    from cudf_amd import GroupBy
    from cudf_amd.serve import GroupByServer, latency_stats
    factories, shapes = make_served(n_req)
    rows = sum(n for _, n in shapes)
    # warm-up of both paths
    for f in factories[:4]:
        q = f()
        GroupBy(q.keys).aggregate(q.requests)
    torch.cuda.synchronize()
    groups_serial, t_serial = [], []
    for f in factories:
        q = f()
        t0 = time.perf_counter()
        r = GroupBy(q.keys).aggregate(q.requests)
        torch.cuda.synchronize()
        t_serial.append(time.perf_counter() - t0)
        groups_serial.append(result_digest(q, r))
        del r
    serial_qps = n_req / sum(t_serial)
    lat_serial = np.array(t_serial) * 1e3
    out = {"workload": f"{n_req} requests from the three templates (1M-row int32 key/int64 value sum,count,"
                       f"min,max; 100M-row int64 key high-cardinality with var; 50M-row three-column "
                       f"int32/Zipf int64/string key), rows log-uniform 1e4..5e7 (total {rows}), random "
                       f"int32/int64/float32/float64 key and value types, 0-5% nulls; inputs resident in HBM",
           "requests": n_req, "rows": rows, "workers": workers, "memory_budget_GB": 200,
           "templates": {str(t): sum(1 for s, _ in shapes if s == t) for t in (1, 2, 3)},
           "plain_serial": {"queries_per_s": serial_qps, "p50_ms": float(np.percentile(lat_serial, 50)),
                            "p99_ms": float(np.percentile(lat_serial, 99)), "note": "service time only",
                            "ms_by_template": {str(t): round(1e3 * sum(x for x, (s, _) in zip(t_serial, shapes)
                                                                       if s == t), 2) for t in (1, 2, 3)},
                            "rows_by_template": {str(t): sum(n for s, n in shapes if s == t) for t in (1, 2, 3)}},
           "sweep": []}
    mismatches = 0
    for load in list(loads) + ["burst"]:
        queries = [f() for f in factories]
        rng = np.random.default_rng(11)
        srv = GroupByServer(workers=workers, memory_budget=200 * 10**9).start()
        t0 = time.perf_counter()
        if load == "burst":
            for q in queries:
                q.arrival = t0
                srv.submit(q)
        else:
            rate = serial_qps * load
            at = t0 + np.cumsum(rng.exponential(1.0 / rate, n_req))
            for q, a in zip(queries, at):
                d = a - time.perf_counter()
                if d > 0:
                    time.sleep(d)
                q.arrival = float(a)
                srv.submit(q)
        res = [q.future.result(timeout=600) for q in queries]
        srv.stop()
        mismatches += sum(int(result_digest(q, r) != g) for q, r, g in zip(queries, res, groups_serial))
        st = latency_stats(queries)
        st["offered_qps"] = None if load == "burst" else serial_qps * load
        st["load"] = load
        st["units"] = srv.units
        st["batched_units"] = srv.batched_units
        out["sweep"].append(st)
        del res, queries
    best = max(s["queries_per_s"] for s in out["sweep"])
    out["max_queries_per_s"] = best
    out["speedup_vs_plain_serial"] = best / serial_qps
    out["exact_result_mismatches_vs_serial"] = mismatches
    out["exact_check"] = ("per query: group count and an order-independent digest of every exact column "
                          "(COUNT, MIN, MAX, integer SUM) with the group's key hash, served vs serial")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--rows", type=int, default=N_ROWS)
    ap.add_argument("--card", type=int, default=CARD)
    ap.add_argument("--cpu-sample", type=int, default=20_000_000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--strategy", choices=["global", "partitioned"], default=None,
                    help="force an aggregation strategy (default: chosen per call)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="run one device copy of MIB MiB first (counter calibration)")
    ap.add_argument("--batched-instances", type=int, default=4096,
                    help="instances of the batched workload (0 skips it)")
    ap.add_argument("--batched-reps", type=int, default=3)
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain per-instance loop")
    ap.add_argument("--only-batched", action="store_true", help="run the batched workload alone")
    ap.add_argument("--served-requests", type=int, default=200,
                    help="requests of the served workload (0 skips it)")
    ap.add_argument("--served-workers", type=int, default=8)
    ap.add_argument("--only-served", action="store_true", help="run the served workload alone")
    ap.add_argument("--only-sorted", action="store_true",
                    help="run the pre-sorted-keys path against the hash path on sorted C2-shaped data")
    ap.add_argument("--only-partition", action="store_true",
                    help="run hash partitioning and the sharded (partial / combine) group-by alone")
    ap.add_argument("--only-c3", action="store_true",
                    help="run the 50M-row multi-column (string) key workload alone")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the group-by is measured on a "
              f"single MI355X and the multi-GPU sharding is described, not run.", file=sys.stderr)
        sys.exit(2)

    import torch
    import cudf_amd as ca
    from cudf_amd import _lib
    from cudf_amd.aggregation import Aggregation, Kind

    torch.cuda.set_device(0)
    if args.only_batched:
        print(json.dumps({"batched": bench_batched(args.batched_instances, args.batched_reps,
                                                   not args.no_autotune, not args.no_plain)}), flush=True)
        return
    if args.only_c3:
        print(json.dumps({"c3": bench_c3()}), flush=True)
        return
    if args.only_partition:
        print(json.dumps({"partition": bench_partition()}), flush=True)
        return
    if args.only_sorted:
        print(json.dumps({"sorted": bench_sorted()}), flush=True)
        return
    if args.only_served:
        print(json.dumps({"served": bench_served(args.served_requests, workers=args.served_workers)}),
              flush=True)
        return
    (kcol, fcol, icol), _keep = make_c2(args.rows, args.card)
    aggs = lambda: [Aggregation(Kind.SUM), Aggregation(Kind.MEAN), Aggregation(Kind.COUNT_VALID),
                    Aggregation(Kind.MIN), Aggregation(Kind.MAX), Aggregation(Kind.VARIANCE, 1)]
    gb = ca.GroupBy(ca.Table([kcol]))
    reqs = [ca.GroupByRequest(fcol, aggs()), ca.GroupByRequest(icol, aggs())]
    _lib.set_phase_timing(True)

    if args.strategy:
        _lib.set_strategy({"global": 1, "partitioned": 2}[args.strategy])
    if args.calibrate_copy:
        # a plain device copy of known size, for scaling FETCH_SIZE / WRITE_SIZE
        src = torch.ones(args.calibrate_copy << 20, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    for _ in range(args.warmup):
        out = gb.aggregate(reqs)
        del out
    torch.cuda.synchronize()
    # timed steps: each step starts with cold caches (a 512 MB buffer, twice
    # the 256 MB Infinity Cache, is written before it, outside its timing)
    # and is bracketed by device synchronisations; hipEvents the library
    # records on its stream bracket the main aggregation kernel (its duration
    # for the roofline).  The other kernels' times come from two untimed
    # steps with every phase timed.
    flush = torch.empty(512 << 20, dtype=torch.uint8, device="cuda")
    _lib.set_phase_timing(2)
    dom_ms = {}
    step_ms = []
    for _ in range(args.steps):
        flush.fill_(1)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        keys_t, res = gb.aggregate(reqs)
        torch.cuda.synchronize()
        step_ms.append((time.perf_counter() - t0) * 1e3)
        st = _lib.last_stats()
        for k in ("t_rpa_aggregate_ms", "t_accumulate_ms"):
            dom_ms[k] = dom_ms.get(k, 0.0) + st[k] / args.steps
        if _ < args.steps - 1:
            del keys_t, res
    del flush
    ms_per_step = float(np.mean(step_ms))
    _lib.set_phase_timing(True)
    phase = {}
    for _ in range(2):
        out = gb.aggregate(reqs)
        del out
        for k, v in _lib.last_stats().items():
            if k.startswith("t_"):
                phase[k] = phase.get(k, 0.0) + v / 2
    phase.update(dom_ms)
    G = keys_t.num_rows()
    in_b = input_bytes(args.rows)
    out_b = output_bytes(keys_t, res)
    # compulsory bytes of each kernel of the strategy that ran (rows moved by
    # the partition passes carry key + values + one flags byte)
    nv = args.rows - kcol.null_count
    rec = 8 + 8 + 4 + 1
    if st["strategy"] == 2:
        kernels = {
            "k_rpa_hist": (phase.get("t_rpa_hist_ms", 0), args.rows * 8 + 4 * ((args.rows + 31) // 32)),
            "k_rpa_scatter1": (phase.get("t_rpa_scatter1_ms", 0), in_b + nv * rec),
            "k_rpa_scatter2": (phase.get("t_rpa_scatter2_ms", 0), 2 * nv * rec),
            "k_rpa_aggregate": (phase.get("t_rpa_aggregate_ms", 0), nv * rec + out_b),
        }
    else:
        kernels = {"k_aggregate_global": (phase.get("t_accumulate_ms", 0), in_b)}
    dom = max(kernels, key=lambda k: kernels[k][0])
    t_dom = kernels[dom][0] * 1e-3
    # the roofline's bytes are the operation's algorithmic bytes (every input
    # buffer read once, every output buffer written once: 29.3 B/row here)
    # over the dominant kernel's time; the bytes that kernel must move by
    # its design are reported beside them
    bytes_alg = in_b + out_b
    achieved = bytes_alg / t_dom if t_dom > 0 else 0.0
    traffic, traffic_src = None, None
    tf = os.path.join(ROOT, "profiles", f"traffic_{dom}_c2.json")
    if os.path.exists(tf):
        with open(tf) as f:
            traffic = json.load(f).get("bytes_per_launch")
        traffic_src = f"profiles/traffic_{dom}_c2.json (rocprofv3 FETCH_SIZE + WRITE_SIZE, not measured in this run)"
    line = {
        "metric": METRIC,
        "value": args.rows / (ms_per_step * 1e-3),
        "unit": "rows/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "ms_per_step_median": float(np.median(step_ms)),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int64+f64",
        "data": "synthetic",
        "config": {"workload": "high-cardinality hash group-by: 100M rows, int64 key uniform over 10M "
                               "values (1% null, excluded), float64 N(0,1) and int32 U[0,1e4] values "
                               "(5% null each); sum, mean, count, min, max, var(ddof=1) on both",
                   "rows": args.rows, "distinct_keys": args.card, "groups": G,
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "kernel": dom,
                     "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK, "traffic": traffic, "traffic_source": traffic_src,
                     "algorithmic_bytes": bytes_alg, "kernel_design_bytes": kernels[dom][1],
                     "avg_ms": t_dom * 1e3},
        "pipeline": {"strategy": {1: "global", 2: "partitioned"}.get(int(st["strategy"]), "?"),
                     "bytes_alg": in_b + out_b, "bytes_per_row": (in_b + out_b) / args.rows,
                     "achieved_GBps": (in_b + out_b) / (ms_per_step * 1e-3) / 1e9,
                     "frac": (in_b + out_b) / (ms_per_step * 1e-3) / HBM_PEAK,
                     "kernels_ms": {k: round(v[0], 4) for k, v in kernels.items()},
                     "last_call": {k: v for k, v in st.items() if not k.startswith("t_")},
                     "phase_ms": {k: round(v, 4) for k, v in phase.items()}},
    }
    if not args.no_cpu:
        line.update(cpu_baselines(args.cpu_sample, max(1, args.cpu_sample // 10)))
    if args.batched_instances > 0:
        del keys_t, res, gb, reqs, kcol, fcol, icol, _keep
        torch.cuda.empty_cache()
        line["batched"] = bench_batched(args.batched_instances, args.batched_reps, not args.no_autotune,
                                        not args.no_plain)
        torch.cuda.empty_cache()
    if args.served_requests > 0:
        line["served"] = bench_served(args.served_requests, workers=args.served_workers)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()

"""Per-group parity at the benchmarked sizes.

Every group of the 100M-row high-cardinality workload, of the skewed
three-column-key workload at 20M rows (the automatically chosen partitioned
general-key path), and of a 4,096-instance batch is compared with the oracle
under the parity bar, after sorting both sides by key (the protocol of
cpp/tests/groupby/groupby_test_util.cpp:34-98).  The oracle runs on
key-disjoint shards in parallel threads, which gives the same groups and
bit-identical per-group results as one call.

Set MIG_PARITY_REPORT to a file path to have the FP64 SUM/MEAN relative-error
distribution of the 100M-row case written there as JSON.
"""
import json
import math
import os
import time

import numpy as np
import pytest

from gbutil import AGG, C2_AGGS, assert_parity, c2_data, c3_data, gpu_groupby, host_col, host_keys, to_device

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[{time.strftime('%H:%M:%S')}] {msg}", flush=True)


def _stats():
    from cudf_amd import _lib
    return _lib.last_stats()


@pytest.mark.timeout(900)
def test_c2_100m_every_group():
    """100M rows, int64 key over 10M values (1% null keys), float64 N(0,1) and
    int32 U[0,1e4] values (5% null each); sum, mean, count, min, max and
    var(ddof=1) of both, every group against the oracle."""
    _say("c2 100M: generating")
    key, f64, i32 = c2_data()
    reqs = [(f64, C2_AGGS), (i32, C2_AGGS)]
    _say("c2 100M: gpu")
    gpu = gpu_groupby([key], reqs)
    st = _stats()
    assert st["strategy"] == 2, st  # the radix-partitioned path
    _say(f"c2 100M: {len(gpu[0][0][0])} groups; oracle and compare")
    stats = {}
    ng = assert_parity([key], reqs, 0, gpu=gpu, stats=stats)
    assert ng == len(np.unique(key.data[key.valid]))
    report = {f"request{r}_{'sum' if k == AGG['SUM'] else 'mean'}": v for (r, k), v in stats.items()}
    _say(f"c2 100M: parity ok; FP64 relative errors {json.dumps(report)}")
    path = os.environ.get("MIG_PARITY_REPORT")
    if path:
        with open(path, "w") as f:
            json.dump({"workload": "C2 100M rows / 10M keys", "groups": ng, "fp64": report}, f, indent=1)


@pytest.mark.timeout(900)
def test_c3_20m_auto_partitioned_general_keys():
    """The skewed key (int32 U[1,100], Zipf(1.1) int64 over 1M values, string
    from a pool of 10k) at 20M rows; sum, count, max of a float32 and an int64
    column.  The automatic strategy must pick the partitioned path with
    general keys; every group against the oracle."""
    _say("c3 20M: generating")
    k1, k2, k3, f32, i64 = c3_data(n=20_000_000, seed=21)
    aggs = [(AGG["SUM"], 1), (AGG["COUNT_VALID"], 1), (AGG["MAX"], 1)]
    reqs = [(f32, aggs), (i64, aggs)]
    _say("c3 20M: gpu")
    gpu = gpu_groupby([k1, k2, k3], reqs)
    st = _stats()
    assert st["strategy"] == 2 and st["key_mode"] == 1, st
    _say(f"c3 20M: {len(gpu[0][0][0])} groups; oracle and compare")
    assert_parity([k1, k2, k3], reqs, 0, gpu=gpu)
    _say("c3 20M: parity ok")


REQ0 = [(AGG["SUM"], 1), (AGG["COUNT_VALID"], 1), (AGG["MIN"], 1), (AGG["MAX"], 1)]


def _c4_small(n_inst, seed):
    """4,096 instances: rows log-uniform 1e3..6.4e4, cardinality log-uniform
    1..rows/2, int32 keys, int64 and float32 values; every third instance
    has null keys, every fourth null int64 values."""
    rng = np.random.default_rng(seed)
    rows = np.round(np.exp(rng.uniform(math.log(1e3), math.log(6.4e4), n_inst))).astype(np.int64)
    card = np.maximum(1, np.round(np.exp(rng.uniform(0.0, np.log(0.5 * rows))))).astype(np.int64)
    inst = []
    for i, (r, c) in enumerate(zip(rows, card)):
        k = rng.integers(0, c, r).astype(np.int32)
        kv = rng.random(r) > 0.03 if i % 3 == 0 else None
        v0 = rng.integers(-10**9, 10**9, r, endpoint=True).astype(np.int64)
        v0v = rng.random(r) > 0.05 if i % 4 == 1 else None
        v1 = rng.random(r).astype(np.float32)
        inst.append((host_col(k, kv), [host_col(v0, v0v), host_col(v1)]))
    return inst


def _concat(cols, ids=None):
    data = np.concatenate([c.data for c in cols])
    valid = None
    if any(c.valid is not None for c in cols):
        valid = np.concatenate([np.ones(c.size, bool) if c.valid is None else c.valid for c in cols])
    return host_col(data, valid)


@pytest.mark.timeout(900)
@pytest.mark.parametrize("null_policy", [0, 1])
def test_c4_4096_instances_every_group(null_policy):
    """One batched plan over 4,096 instances (the persistent work queue at
    its real instance count); every instance's groups against the oracle,
    run once over all instances with the instance number as a second key."""
    from cudf_amd import BatchedGroupBy, NullPolicy
    from cudf_amd.aggregation import Aggregation
    _say("c4 4096: generating")
    inst = _c4_small(4096, 40 + null_policy)
    dev = [(to_device(k), [to_device(v) for v in vs]) for k, vs in inst]
    aggs = [REQ0, REQ0]
    plan = BatchedGroupBy(dev, [[Aggregation(k, d) for k, d in al] for al in aggs], NullPolicy(null_policy))
    _say(f"c4 4096: {sum(k.size for k, _ in inst)} rows; run")
    plan.run()
    g_ids, g_keys, g_res = [], [], [[[] for _ in al] for al in aggs]
    for i in range(len(inst)):
        keys, res = plan.result(i)
        kv, kval = host_keys(keys.columns()[0])
        g_ids.append(np.full(len(kv), i, np.int32))
        g_keys.append((kv, np.ones(len(kv), bool) if kval is None else kval))
        for r, t in enumerate(res):
            for a, c in enumerate(t.columns()):
                v, val = c.to_numpy()
                g_res[r][a].append((v, np.ones(len(v), bool) if val is None else val))
        assert plan.num_groups(i) == len(kv)
    gpu_keys = [(np.concatenate(g_ids), None),
                (np.concatenate([k for k, _ in g_keys]), np.concatenate([v for _, v in g_keys]))]
    gpu_res = [[(np.concatenate([p[0] for p in parts]), np.concatenate([p[1] for p in parts])) for parts in rr]
               for rr in g_res]
    plan.close()
    ids = host_col(np.concatenate([np.full(k.size, i, np.int32) for i, (k, _) in enumerate(inst)]))
    key = _concat([k for k, _ in inst])
    reqs = [(_concat([vs[j] for _, vs in inst]), aggs[j]) for j in range(2)]
    _say(f"c4 4096: {len(gpu_keys[0][0])} groups; oracle and compare")
    ng = assert_parity([ids, key], reqs, null_policy, gpu=(gpu_keys, gpu_res))
    _say(f"c4 4096: parity ok over {ng} groups")

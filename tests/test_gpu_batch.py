"""Batched group-by (mig_batch_* through cudf_amd.BatchedGroupBy): every
instance's result against the oracle under the parity bar, across the
plan's knobs, graph replays, table regrowth and the edge cases of the
single-call tests (null keys both policies, sentinel-valued keys, empty
instances)."""
import math

import numpy as np
import pytest

from gbutil import AGG, assert_parity, host_col, to_device

pytestmark = pytest.mark.gpu

REQ0 = [(AGG["SUM"], 1), (AGG["COUNT_VALID"], 1), (AGG["MIN"], 1), (AGG["MAX"], 1), (AGG["MEAN"], 1),
        (AGG["COUNT_ALL"], 1)]
REQ1 = [(AGG["SUM"], 1), (AGG["MIN"], 1), (AGG["MAX"], 1), (AGG["COUNT_VALID"], 1), (AGG["MEAN"], 1)]


def make_instances(seed, n_inst, max_rows, key_dtype=np.int32, big=()):
    """Instances with log-uniform rows and cardinalities; some with null keys
    or null values.  ``big`` adds instances of the given (rows, card)."""
    rng = np.random.default_rng(seed)
    shapes = []
    for i in range(n_inst):
        rows = 0 if i == 0 else (1 if i == 1 else int(round(math.exp(rng.uniform(0, math.log(max_rows))))))
        card = max(1, int(round(math.exp(rng.uniform(0, math.log(max(1.0, rows / 2)))))))
        shapes.append((rows, card))
    shapes += list(big)
    out = []
    for i, (rows, card) in enumerate(shapes):
        keys = rng.integers(0, card, rows).astype(key_dtype)
        kvalid = rng.random(rows) > 0.03 if i % 3 == 0 else None
        v0 = rng.integers(-1_000_000, 1_000_000, rows).astype(np.int64)
        v0valid = rng.random(rows) > 0.05 if i % 4 == 1 else None
        v1 = rng.random(rows).astype(np.float32)
        v1valid = rng.random(rows) > 0.05 if i % 5 == 2 else None
        out.append((host_col(keys, kvalid), [host_col(v0, v0valid), host_col(v1, v1valid)]))
    return out


def run_batch(instances, aggs, null_policy=0, options=None):
    from cudf_amd import BatchedGroupBy, NullPolicy
    from cudf_amd.aggregation import Aggregation
    dev = [(to_device(k), [to_device(v) for v in vs]) for k, vs in instances]
    plan = BatchedGroupBy(dev, [[Aggregation(k, d) for k, d in al] for al in aggs], NullPolicy(null_policy),
                          options)
    plan.run()
    return plan


def host_result(plan, i):
    keys, res = plan.result(i)
    kcols = [c.to_numpy() for c in keys.columns()]
    r = [[c.to_numpy() for c in t.columns()] for t in res]
    return kcols, r, keys, res


def check_all(plan, instances, aggs, null_policy=0):
    for i, (k, vs) in enumerate(instances):
        kcols, r, _, _ = host_result(plan, i)
        reqs = [(vs[j], aggs[j]) for j in range(len(aggs))]
        g = assert_parity([k], reqs, null_policy, gpu=(kcols, r))
        assert g == plan.num_groups(i)


@pytest.mark.parametrize("null_policy", [0, 1])
def test_batch_matches_oracle(null_policy):
    inst = make_instances(7 + null_policy, 40, 60_000, big=[(300_000, 150_000), (250_000, 40)])
    plan = run_batch(inst, [REQ0, REQ1], null_policy)
    check_all(plan, inst, [REQ0, REQ1], null_policy)


@pytest.mark.parametrize("opts", [
    dict(lds_max_groups=-1), dict(block_size=512), dict(tile_rows=256), dict(use_graph=-1),
    dict(lds_max_groups=64, tile_rows=1024),
    dict(load_factor=0.9, lds_max_groups=768, tile_rows=16384),
])
def test_batch_options(opts):
    from cudf_amd import BatchOptions
    inst = make_instances(11, 24, 40_000, big=[(120_000, 60_000)])
    plan = run_batch(inst, [REQ0, REQ1], 0, BatchOptions(**opts))
    check_all(plan, inst, [REQ0, REQ1], 0)


def test_batch_graph_replays():
    """The captured graph replayed several times, with unrelated device work
    and synchronisations between replays; every replay checked."""
    import torch
    inst = make_instances(13, 16, 30_000, big=[(100_000, 30_000)])
    plan = run_batch(inst, [REQ0, REQ1], 0)
    for rep in range(3):
        check_all(plan, inst, [REQ0, REQ1], 0)
        x = torch.randn(512, 512, device="cuda")
        (x @ x).sum().item()
        torch.cuda.synchronize()
        plan.run()


def test_batch_table_regrowth():
    """Keys chosen so that the strided sample sees one key while the instance
    has ~100k groups: the table starts tiny, overflows, and is regrown."""
    rows, ns = 200_000, 4096
    keys = np.arange(rows, dtype=np.int32) + 1
    keys[(np.arange(ns, dtype=np.int64) * rows) // ns] = 0
    rng = np.random.default_rng(3)
    inst = [(host_col(keys), [host_col(rng.integers(-5, 5, rows).astype(np.int64)),
                              host_col(rng.random(rows).astype(np.float32))])]
    inst += make_instances(5, 6, 20_000)
    plan = run_batch(inst, [REQ0, REQ1], 0)
    check_all(plan, inst, [REQ0, REQ1], 0)
    assert plan.num_groups(0) == rows - ns + 1


def test_batch_int64_and_float_keys():
    """int64 keys that equal the table's empty marker, and float64 keys with
    NaN and -0.0 (grouped with NaN and +0.0)."""
    rng = np.random.default_rng(17)
    sentinel = np.array([0xFFF1A5A5A5A5A5A5], dtype=np.uint64).view(np.int64)[0]
    inst64 = []
    for i in range(6):
        rows = 5_000 * (i + 1)
        k = rng.integers(-50, 50, rows).astype(np.int64)
        k[rng.random(rows) < 0.1] = sentinel
        valid = rng.random(rows) > 0.05 if i % 2 else None
        inst64.append((host_col(k, valid), [host_col(rng.integers(0, 100, rows).astype(np.int64)),
                                             host_col(rng.random(rows).astype(np.float32))]))
    for pol in (0, 1):
        plan = run_batch(inst64, [REQ0, REQ1], pol)
        check_all(plan, inst64, [REQ0, REQ1], pol)
    instf = []
    for i in range(5):
        rows = 7_000 * (i + 1)
        k = rng.integers(-20, 20, rows).astype(np.float64) / 4
        k[rng.random(rows) < 0.05] = np.nan
        k[rng.random(rows) < 0.05] = -0.0
        instf.append((host_col(k), [host_col(rng.integers(0, 100, rows).astype(np.int64)),
                                    host_col(rng.random(rows).astype(np.float32))]))
    plan = run_batch(instf, [REQ0, REQ1], 0)
    check_all(plan, instf, [REQ0, REQ1], 0)


def test_batch_nullability_matches_single_call():
    """Result masks exist exactly where the single-call path has them, and the
    reported null counts match the masks."""
    from cudf_amd import GroupBy, GroupByRequest, NullPolicy, Table
    from cudf_amd.aggregation import Aggregation
    inst = make_instances(19, 12, 10_000)
    plan = run_batch(inst, [REQ0, REQ1], 1)
    for i, (k, vs) in enumerate(inst):
        _, _, bkeys, bres = host_result(plan, i)
        reqs = [GroupByRequest(to_device(vs[j]), [Aggregation(a, d) for a, d in al])
                for j, al in enumerate([REQ0, REQ1])]
        skeys, sres = GroupBy(Table([to_device(k)]), NullPolicy.INCLUDE).aggregate(reqs)
        pairs = [(bkeys.columns()[0], skeys.columns()[0])]
        for bt, st in zip(bres, sres):
            pairs += list(zip(bt.columns(), st.columns()))
        for b, s in pairs:
            assert (b.mask is None) == (s.mask is None)
            assert b.type_id == s.type_id and b.size == s.size
            if b.mask is not None:
                assert b.null_count == s.null_count == int((~b.validity()).sum())


@pytest.mark.parametrize("null_policy", [0, 1])
def test_batch_partitioned_instances(null_policy):
    """A large high-cardinality instance runs the single-call partitioned
    path inside the plan; every instance still matches the oracle, also on a
    second execution, and with the path switched off."""
    from cudf_amd import BatchOptions
    from cudf_amd._lib import last_stats
    inst = make_instances(23, 6, 5_000, big=[(1_200_000, 500_000)])
    plan = run_batch(inst, [REQ0, REQ1], null_policy)
    # the big instance's call was partitioned; with null keys included, its
    # null group holds 3% of the rows, a hot key that the device-wide table
    # with workgroup-local tables takes instead
    assert last_stats()["strategy"] == (2 if null_policy == 0 else 1)
    check_all(plan, inst, [REQ0, REQ1], null_policy)
    plan.run()
    check_all(plan, inst, [REQ0, REQ1], null_policy)
    off = run_batch(inst, [REQ0, REQ1], null_policy, BatchOptions(partition_min_rows=-1))
    check_all(off, inst, [REQ0, REQ1], null_policy)


def test_batch_empty_instances():
    inst = [(host_col(np.zeros(0, np.int32)), [host_col(np.zeros(0, np.int64)),
                                                host_col(np.zeros(0, np.float32))]) for _ in range(3)]
    plan = run_batch(inst, [REQ0, REQ1], 0)
    for i in range(3):
        assert plan.num_groups(i) == 0
        kcols, r, _, _ = host_result(plan, i)
        assert len(kcols[0][0]) == 0 and all(len(c[0]) == 0 for t in r for c in t)


def make_mixed(seed, n_inst, max_rows):
    """Instances whose key and value types differ (int32/int64/float32/float64),
    with nulls in some keys and values."""
    rng = np.random.default_rng(seed)
    types = [np.int32, np.int64, np.float32, np.float64]
    out = []
    for i in range(n_inst):
        rows = 0 if i == 0 else int(round(math.exp(rng.uniform(0, math.log(max_rows)))))
        card = max(1, int(round(math.exp(rng.uniform(0, math.log(max(1.0, rows / 2)))))))
        kt, t0, t1 = (types[j] for j in rng.integers(0, 4, 3))
        keys = (rng.integers(0, card, rows) - card // 3).astype(kt)
        kvalid = rng.random(rows) > 0.04 if i % 3 == 0 else None

        def val(t):
            if np.issubdtype(t, np.integer):
                return rng.integers(-1_000_000, 1_000_000, rows).astype(t)
            return (rng.standard_normal(rows) * 100).astype(t)
        v0valid = rng.random(rows) > 0.05 if i % 2 == 1 else None
        v1valid = rng.random(rows) > 0.05 if i % 4 == 2 else None
        out.append((host_col(keys, kvalid), [host_col(val(t0), v0valid), host_col(val(t1), v1valid)]))
    return out


REQ_VAR0 = [(AGG["SUM"], 1), (AGG["MEAN"], 1), (AGG["COUNT_VALID"], 1), (AGG["MIN"], 1), (AGG["MAX"], 1),
            (AGG["VARIANCE"], 1)]
REQ_VAR1 = [(AGG["STD"], 0), (AGG["VARIANCE"], 2), (AGG["SUM"], 1), (AGG["COUNT_ALL"], 1)]


@pytest.mark.parametrize("null_policy", [0, 1])
def test_batch_mixed_types_with_variance(null_policy):
    """Instances with different key and value types in one plan, with
    variance / std (second pass) next to the first-pass aggregations."""
    inst = make_mixed(29 + null_policy, 36, 50_000)
    inst += make_mixed(31, 2, 2)  # tiny instances: var null where count <= ddof
    plan = run_batch(inst, [REQ_VAR0, REQ_VAR1], null_policy)
    check_all(plan, inst, [REQ_VAR0, REQ_VAR1], null_policy)


def test_batch_mixed_types_options():
    from cudf_amd import BatchOptions
    inst = make_mixed(37, 20, 30_000)
    for opts in (dict(lds_max_groups=-1), dict(block_size=512, tile_rows=1024), dict(use_graph=-1),
                 dict(lds_max_groups=96)):
        plan = run_batch(inst, [REQ_VAR0, REQ_VAR1], 0, BatchOptions(**opts))
        check_all(plan, inst, [REQ_VAR0, REQ_VAR1], 0)


def test_batch_errors():
    from cudf_amd import _lib
    from oracle import oracle as orc
    inst = make_instances(23, 3, 1_000)
    with pytest.raises(_lib.LogicError):
        run_batch(inst, [[(AGG["PRODUCT"], 1)], REQ1], 0)
    with pytest.raises(_lib.LogicError):
        run_batch(inst, [[(AGG["ARGMIN"], 1)], REQ1], 0)
    bad = list(inst)
    k, vs = bad[2]
    bad[2] = (k, [host_col(vs[0].data[:-1]), vs[1]])
    with pytest.raises(_lib.LogicError):
        run_batch(bad, [REQ0, REQ1], 0)
    strs = [(orc.HostColumn.from_strings([b"x%d" % (j % 7) for j in range(len(k.data))]), vs)]
    with pytest.raises(_lib.DataTypeError):
        run_batch(strs, [REQ0, REQ1], 0)

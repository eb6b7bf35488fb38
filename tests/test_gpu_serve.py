"""Served group-by (cudf_amd.serve): queries from the three templates with
random dtypes and nulls, submitted concurrently (all at once and with
Poisson arrivals) so that they batch and overlap; every query's result is
checked against the oracle under the parity bar and against the same query
run alone (exact columns bit-equal)."""
import math
import threading
import time

import numpy as np
import pytest

from gbutil import AGG, assert_parity, served_request, sort_order, to_device

pytestmark = pytest.mark.gpu

EXACT = {AGG["COUNT_VALID"], AGG["COUNT_ALL"], AGG["MIN"], AGG["MAX"]}


def make_stream(seed, n, max_rows):
    rng = np.random.default_rng(seed)
    out = []
    for q in range(n):
        template = int(rng.integers(1, 4))
        rows = int(round(math.exp(rng.uniform(math.log(1_000), math.log(max_rows)))))
        out.append(served_request(rng, template, rows))
    return out


def to_query(keys, requests, policy):
    from cudf_amd import GroupByRequest, NullPolicy, Table
    from cudf_amd.aggregation import Aggregation
    from cudf_amd.serve import Query
    return Query(Table([to_device(k) for k in keys]),
                 [GroupByRequest(to_device(c), [Aggregation(k, d) for k, d in aggs]) for c, aggs in requests],
                 NullPolicy(policy))


def host(out):
    keys, res = out
    return [c.to_numpy() for c in keys.columns()], [[c.to_numpy() for c in t.columns()] for t in res]


def assert_same_exact(a, b, requests):
    """Keys and the exact result columns of two runs of one query are equal
    (after sorting both by key); the floating-point ones are checked against
    the oracle separately."""
    (ak, ar), (bk, br) = a, b
    oa, ob = sort_order(ak), sort_order(bk)
    assert len(oa) == len(ob)
    for (av, avld), (bv, bvld) in zip(ak, bk):
        if isinstance(av, list):
            assert [av[i] for i in oa] == [bv[i] for i in ob]
        else:
            assert np.array_equal(av[oa].view(np.uint8), bv[ob].view(np.uint8)) or np.array_equal(av[oa], bv[ob],
                                                                                                  equal_nan=True)
        if avld is not None or bvld is not None:
            assert np.array_equal(avld[oa], bvld[ob])
    for r, (col, aggs) in enumerate(requests):
        for a_i, (kind, _) in enumerate(aggs):
            (av, avld), (bv, bvld) = ar[r][a_i], br[r][a_i]
            va = np.ones(len(av), bool) if avld is None else avld
            vb = np.ones(len(bv), bool) if bvld is None else bvld
            assert np.array_equal(va[oa], vb[ob])
            exact = kind in EXACT or (kind == AGG["SUM"] and av.dtype.kind in "iu")
            if exact:
                m = va[oa]
                assert np.array_equal(av[oa][m], bv[ob][m], equal_nan=True)


@pytest.mark.parametrize("arrivals", ["burst", "poisson"])
def test_served_results_same_alone_and_in_mix(arrivals):
    from cudf_amd import GroupBy
    from cudf_amd.serve import GroupByServer, latency_stats
    stream = make_stream(41 if arrivals == "burst" else 43, 48, 120_000)
    alone = []
    for keys, reqs, pol in stream:
        q = to_query(keys, reqs, pol)
        alone.append(host(GroupBy(q.keys, q.null_policy).aggregate(q.requests)))
    queries = [to_query(*s) for s in stream]
    with GroupByServer(workers=3, max_batch=64, batch_rows=1_000_000) as srv:
        if arrivals == "burst":
            futs = [srv.submit(q) for q in queries]
        else:
            rng = np.random.default_rng(5)
            futs = []
            for q in queries:
                time.sleep(float(rng.exponential(0.002)))
                futs.append(srv.submit(q))
        outs = [f.result(timeout=120) for f in futs]
    st = latency_stats(queries)
    if arrivals == "burst":  # the mix really was batched (Poisson arrivals may be served one by one)
        assert st["batched_fraction"] > 0.3, st
    for (keys, reqs, pol), a, out in zip(stream, alone, outs):
        b = host(out)
        assert_parity(keys, reqs, pol, gpu=b)
        assert_same_exact(a, b, reqs)


def test_served_memory_admission():
    """A budget that holds about one query at a time still completes every
    query; no more than the budget is reserved at once (larger queries run
    alone)."""
    from cudf_amd.serve import GroupByServer
    stream = make_stream(47, 12, 50_000)
    queries = [to_query(*s) for s in stream]
    budget = max(q.memory_estimate() for q in queries)
    peak = [0]
    srv = GroupByServer(workers=3, memory_budget=budget, max_batch=8)
    orig = srv._next_unit

    def watched():
        u = orig()
        if u is not None:
            peak[0] = max(peak[0], srv._mem_in_use + u[1])
        return u
    srv._next_unit = watched
    with srv:
        futs = [srv.submit(q) for q in queries]
        outs = [f.result(timeout=120) for f in futs]
    assert peak[0] <= budget
    for (keys, reqs, pol), out in zip(stream, outs):
        assert_parity(keys, reqs, pol, gpu=host(out))


def test_served_query_error_is_its_own():
    """A query that fails validation fails alone; the others in the mix
    complete."""
    from cudf_amd import GroupByRequest, _lib
    from cudf_amd.aggregation import Aggregation
    from cudf_amd.serve import GroupByServer
    stream = make_stream(53, 6, 20_000)
    queries = [to_query(*s) for s in stream]
    bad = to_query(*stream[0])
    bad.requests = [GroupByRequest(bad.requests[0].values, [Aggregation(6)])]  # ANY: not a group-by aggregation
    with GroupByServer(workers=2) as srv:
        futs = [srv.submit(q) for q in queries[:3]] + [srv.submit(bad)] + [srv.submit(q) for q in queries[3:]]
        with pytest.raises(_lib.LogicError):
            futs[3].result(timeout=120)
        for f, (keys, reqs, pol) in zip(futs[:3] + futs[4:], stream):
            assert_parity(keys, reqs, pol, gpu=host(f.result(timeout=120)))

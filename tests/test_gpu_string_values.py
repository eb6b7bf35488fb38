"""String value columns on the sort-based paths: NUNIQUE (both null
policies), NTH_ELEMENT (n from the front and the back, nulls included and
excluded), MIN / MAX / ARGMIN / ARGMAX and COUNT_VALID of strings, on runs of
pre-sorted keys and on unsorted keys (the sort-based group-by, which gathers
the strings in group order and maps ARGMIN / ARGMAX rows back to input rows).
References: sort/aggregate.cpp:364-410 (NUNIQUE, NTH_ELEMENT),
sort/group_nunique.cu:44-64, sort/group_nth_element.cu:50-140,
sort/group_argmin.cu / group_argmax.cu, string_view::compare ordering.

The expected results are computed here from the group rows in input order,
pure Python over tens of thousands of rows; ARGMIN / ARGMAX are also checked
against the oracle where it covers them."""
import numpy as np
import pytest

from oracle import oracle as orc
from tests.gbutil import AGG, assert_parity, gpu_groupby, host_col, oracle_groupby, sort_order

pytestmark = pytest.mark.gpu

NTH = [(0, 1), (1, 0), (-1, 1), (-2, 0), (5, 0), (40, 1)]
AGGS = ([(AGG["NUNIQUE"], 0), (AGG["NUNIQUE"], 1)] + [(AGG["NTH_ELEMENT"], 2 * m + inc) for m, inc in NTH] +
        [(AGG["MIN"], 1), (AGG["MAX"], 1), (AGG["ARGMIN"], 1), (AGG["ARGMAX"], 1), (AGG["COUNT_VALID"], 1)])


def _strings(rng, n, pool_size=400, max_len=7):
    pool = [bytes(rng.integers(97, 101, int(rng.integers(0, max_len + 1)), dtype=np.uint8)) for _ in range(pool_size)]
    return [pool[i] for i in rng.integers(0, pool_size, n)]


def _expected(groups, strs, svalid):
    """Per group (a list of rows in input order): the value of every
    aggregation of AGGS, None for a null result."""
    out = []
    for rows in groups:
        valid = [r for r in rows if svalid[r]]
        vals = {strs[r] for r in valid}
        e = [len(vals), len(vals) + (1 if len(valid) < len(rows) else 0)]
        for m, inc in NTH:
            lst = rows if inc else valid
            pick = lst[m] if -len(lst) <= m < len(lst) else None
            e.append(strs[pick] if pick is not None and svalid[pick] else None)
        if valid:
            mn, mx = min(strs[r] for r in valid), max(strs[r] for r in valid)
            amin = next(r for r in valid if strs[r] == mn)
            amax = next(r for r in valid if strs[r] == mx)
            e += [mn, mx, amin, amax]
        else:
            e += [None, None, None, None]
        e.append(len(valid))
        out.append(e)
    return out


def _gpu_values(res, g):
    """The GPU results of group index g, in AGGS order (None for null)."""
    row = []
    for col, valid in res[0]:
        ok = True if valid is None else bool(valid[g])
        if not ok:
            row.append(None)
            continue
        x = col[g]
        row.append(bytes(x) if isinstance(x, (bytes, bytearray, np.bytes_)) else int(x))
    return row


def _check(groups, strs, svalid, g_res, order):
    exp = _expected(groups, strs, svalid)
    assert len(exp) == len(order)
    for gi, e in zip(order, exp):
        got = _gpu_values(g_res, gi)
        assert got == e, (gi, got, e)


@pytest.mark.parametrize("policy", [0, 1])
def test_string_values_on_runs(policy):
    """Pre-sorted int32 keys with null keys last (a null group under
    INCLUDE); groups are the runs, returned in input order."""
    rng = np.random.default_rng(900 + policy)
    n = 30_000
    k = np.sort(rng.integers(0, 2_000, n)).astype(np.int32)
    kvalid = np.ones(n, bool)
    kvalid[-500:] = False
    strs = _strings(rng, n)
    svalid = rng.random(n) > 0.2
    vals = orc.HostColumn.from_strings(strs, svalid)
    g_keys, g_res = gpu_groupby([host_col(k, kvalid)], [(vals, AGGS)], policy, keys_sorted=True)
    groups, cur = [], None
    for r in range(n):
        if not kvalid[r]:
            if policy == 0:
                cur = None
                continue
            if cur is None or kvalid[cur[-1]]:
                cur = [r]
                groups.append(cur)
            else:
                cur.append(r)
        elif cur is None or not kvalid[cur[-1]] or k[r] != k[cur[-1]]:
            cur = [r]
            groups.append(cur)
        else:
            cur.append(r)
    _check(groups, strs, svalid, g_res, list(range(len(groups))))


@pytest.mark.parametrize("policy", [0, 1])
def test_string_values_on_unsorted_keys(policy):
    """Unsorted int64 keys with nulls: the sort-based group-by gathers the
    string column in group order; groups compared by key."""
    rng = np.random.default_rng(910 + policy)
    n = 40_000
    k = rng.integers(0, 3_000, n).astype(np.int64)
    kvalid = rng.random(n) > 0.03
    strs = _strings(rng, n)
    svalid = rng.random(n) > 0.15
    vals = orc.HostColumn.from_strings(strs, svalid)
    g_keys, g_res = gpu_groupby([host_col(k, kvalid)], [(vals, AGGS)], policy)
    by_key = {}
    for r in range(n):
        if not kvalid[r]:
            if policy == 1:
                by_key.setdefault(None, []).append(r)
            continue
        by_key.setdefault(int(k[r]), []).append(r)
    gk, gv = g_keys[0]
    gv = np.ones(len(gk), bool) if gv is None else gv
    index = {(int(gk[i]) if gv[i] else None): i for i in range(len(gk))}
    assert len(index) == len(by_key) == len(gk)
    keys = list(by_key)
    _check([by_key[x] for x in keys], strs, svalid, g_res, [index[x] for x in keys])


def test_argmin_rows_are_input_rows_on_the_sort_path():
    """ARGMIN / ARGMAX mixed with MEDIAN (which routes the call to the
    sort-based group-by) name rows of the input, as the oracle does."""
    rng = np.random.default_rng(920)
    n = 50_000
    key = host_col(rng.integers(0, 4_000, n).astype(np.int32), rng.random(n) > 0.02)
    v = host_col(rng.integers(-1000, 1000, n).astype(np.int64), rng.random(n) > 0.1)
    w = host_col(rng.standard_normal(n))
    g_keys, g_res = gpu_groupby([key], [(v, [(AGG["ARGMIN"], 1), (AGG["ARGMAX"], 1)]), (w, [(AGG["MEDIAN"], 1)])])
    from cudf_amd import _lib
    assert _lib.last_stats()["strategy"] == 4  # the sort-based path
    o_keys, o_res = oracle_groupby([key], [(v, [(AGG["ARGMIN"], 1), (AGG["ARGMAX"], 1)])])
    og, oo = sort_order(g_keys), sort_order(o_keys)
    assert len(og) == len(oo)
    for a in range(2):
        ga, gva = g_res[0][a]
        oa, ova = o_res[0][a]
        gva = np.ones(len(ga), bool) if gva is None else gva
        assert np.array_equal(gva[og], ova[oo].astype(bool))
        ok = ova[oo].astype(bool)
        assert np.array_equal(ga[og][ok], oa[oo][ok])


def test_median_of_strings_still_rejected():
    from cudf_amd._lib import DataTypeError
    rng = np.random.default_rng(5)
    k = host_col(rng.integers(0, 10, 100).astype(np.int32))
    s = orc.HostColumn.from_strings([b"x"] * 100)
    with pytest.raises(DataTypeError):
        gpu_groupby([k], [(s, [(AGG["MEDIAN"], 1)])])
    # COUNT_VALID of strings with a MEDIAN of numbers: the sort path gathers the strings
    assert_parity([k], [(s, [(AGG["COUNT_VALID"], 1)]), (host_col(np.ones(100)), [(AGG["MEDIAN"], 1)])])

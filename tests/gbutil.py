"""Shared helpers of the group-by tests: data generators, the GPU call through
the C ABI, the oracle call, and sort-by-key comparison (the protocol of
cpp/tests/groupby/groupby_test_util.cpp:34-98: output order is unspecified,
so both sides are sorted by key before comparing)."""
from __future__ import annotations

import math
from typing import List, Optional, Sequence

import numpy as np

from oracle import oracle as orc

AGG = {
    "SUM": 0, "PRODUCT": 1, "MIN": 2, "MAX": 3, "COUNT_VALID": 4, "COUNT_ALL": 5,
    "SUM_OF_SQUARES": 8, "MEAN": 9, "VARIANCE": 11, "STD": 12, "ARGMAX": 15, "ARGMIN": 16, "NUNIQUE": 17, "NTH_ELEMENT": 18,
}

# Tolerances of the parity bar.  Integer, index, key, count, min and max
# results are compared bit-exactly.  Floating-point sums carry the error of
# unordered accumulation, bounded normwise: |got - exact| <= tol * sum(|x|)
# over the group (so relative error <= tol whenever the sum does not cancel).
TOL_F64_SUM = 1e-12
TOL_F32_SUM = 1e-5
TOL_VAR = 1e-10


def host_col(arr, valid=None) -> orc.HostColumn:
    return orc.HostColumn.from_numpy(np.asarray(arr), None if valid is None else np.asarray(valid, bool))


def to_device(c: orc.HostColumn):
    from cudf_amd import Column
    if c.type_id == orc.STRING:
        strs = [bytes(c.data[c.offsets[i]:c.offsets[i + 1]]) for i in range(c.size)]
        return Column.from_strings(strs, c.valid)
    return Column.from_numpy(c.data, c.valid, type_id=c.type_id)


def gpu_groupby(keys: Sequence[orc.HostColumn], requests, null_policy: int = 0, keys_sorted: bool = False):
    """Run the GPU group-by; returns (key columns, results) as host data:
    key columns are (values, valid) pairs, results[r][a] = (values, valid).
    ``keys_sorted`` passes keys_are_sorted=YES (groups = runs of equal rows)."""
    from cudf_amd import GroupBy, GroupByRequest, NullPolicy, Sorted, Table
    from cudf_amd.aggregation import Aggregation
    dkeys = Table([to_device(k) for k in keys])
    reqs = []
    for col, aggs in requests:
        reqs.append(GroupByRequest(to_device(col), [_aggregation(k, d) for k, d in aggs]))
    gb = GroupBy(dkeys, NullPolicy(null_policy), Sorted.YES if keys_sorted else Sorted.NO)
    out_keys, out_res = gb.aggregate(reqs)
    kcols = [c.to_numpy() for c in out_keys.columns()]
    res = [[c.to_numpy() for c in t.columns()] for t in out_res]
    return kcols, res


def _aggregation(kind, param):
    """The second element of a request pair: ddof; NUNIQUE's null policy;
    NTH_ELEMENT's 2 * n + (1 when nulls are included)."""
    from cudf_amd import NullPolicy
    from cudf_amd.aggregation import Aggregation
    if kind == AGG["NUNIQUE"]:
        return Aggregation(kind, null_handling=NullPolicy(param))
    if kind == AGG["NTH_ELEMENT"]:
        return Aggregation(kind, null_handling=NullPolicy(param & 1), n=param >> 1)
    return Aggregation(kind, param)


def oracle_groupby(keys: Sequence[orc.HostColumn], requests, null_policy: int = 0):
    rep, res = orc.groupby(keys, requests, null_policy)
    kcols = []
    for k in keys:
        if k.type_id == orc.STRING:
            vals = [bytes(k.data[k.offsets[r]:k.offsets[r + 1]]) for r in rep]
        else:
            vals = k.data[rep]
        valid = None if k.valid is None else k.valid[rep]
        kcols.append((vals, valid))
    return kcols, res


def _key_sort_token(v, valid):
    if not valid:
        return (1, 0)
    if isinstance(v, (bytes, bytearray)):
        return (0, bytes(v))
    f = float(v) if isinstance(v, (float, np.floating)) else None
    if f is not None:
        if math.isnan(f):
            return (0, (1, 0.0))
        return (0, (0, f + 0.0))
    return (0, int(v))


def sort_order(kcols) -> np.ndarray:
    n = len(kcols[0][0]) if kcols else 0
    toks = []
    for i in range(n):
        t = []
        for vals, valid in kcols:
            t.append(_key_sort_token(vals[i], True if valid is None else bool(valid[i])))
        toks.append(tuple(t))
    return np.array(sorted(range(n), key=lambda i: toks[i]), dtype=np.int64)


def _keys_equal(a, av, b, bv) -> bool:
    if av != bv:
        return False
    if not av:
        return True
    if isinstance(a, (bytes, bytearray)):
        return bytes(a) == bytes(b)
    if isinstance(a, (float, np.floating)):
        if math.isnan(a) and math.isnan(b):
            return True
        return a == b
    return a == b


def assert_keys_equal(gk, ok, og, oo):
    """gk/ok: key columns; og/oo: sort orders."""
    assert len(og) == len(oo), f"group count differs: gpu {len(og)} vs oracle {len(oo)}"
    for c, ((gv, gval), (ov, oval)) in enumerate(zip(gk, ok)):
        for i in range(len(og)):
            a, b = gv[og[i]], ov[oo[i]]
            av = True if gval is None else bool(gval[og[i]])
            bv = True if oval is None else bool(oval[oo[i]])
            assert _keys_equal(a, av, b, bv), f"key column {c} row {i}: gpu {a!r}/{av} vs oracle {b!r}/{bv}"


def group_abs_sums(keys, col: orc.HostColumn, null_policy: int) -> np.ndarray:
    """Per-group sum of |x| (first-occurrence order of the oracle), used as the
    scale of floating-point sum tolerances."""
    absdata = np.abs(col.data.astype(np.float64))
    absc = orc.HostColumn(orc.FLOAT64, absdata, col.valid)
    _, res = orc.groupby(keys, [(absc, [(AGG["SUM"], 1)])], null_policy)
    return res[0][0][0]


def group_sq_sums(keys, col: orc.HostColumn, null_policy: int) -> np.ndarray:
    sq = col.data.astype(np.float64) ** 2
    c = orc.HostColumn(orc.FLOAT64, sq, col.valid)
    _, res = orc.groupby(keys, [(c, [(AGG["SUM"], 1)])], null_policy)
    return res[0][0][0]


def assert_parity(keys, requests, null_policy=0, gpu=None):
    """Run both sides (unless ``gpu`` results are given) and compare under the
    parity bar.  Returns the number of groups."""
    g_keys, g_res = gpu if gpu is not None else gpu_groupby(keys, requests, null_policy)
    o_keys, o_res = oracle_groupby(keys, requests, null_policy)
    og, oo = sort_order(g_keys), sort_order(o_keys)
    assert_keys_equal(g_keys, o_keys, og, oo)
    for r, (col, aggs) in enumerate(requests):
        scale = None
        for a, (kind, ddof) in enumerate(aggs):
            gv, gval = g_res[r][a]
            ov, oval = o_res[r][a]
            assert len(gv) == len(ov)
            gvalid = np.ones(len(gv), bool) if gval is None else gval
            gvalid_s, ovalid_s = gvalid[og], oval[oo]
            assert np.array_equal(gvalid_s, ovalid_s), f"validity differs for request {r} agg {kind}"
            gs, os_ = gv[og][gvalid_s], ov[oo][ovalid_s]
            assert gs.dtype == os_.dtype, f"dtype {gs.dtype} vs {os_.dtype} (agg {kind})"
            is_float_src = col.type_id in (orc.FLOAT32, orc.FLOAT64)
            if kind in (AGG["COUNT_VALID"], AGG["COUNT_ALL"], AGG["MIN"], AGG["MAX"], AGG["ARGMIN"],
                        AGG["ARGMAX"], AGG["NUNIQUE"], AGG["NTH_ELEMENT"]) or (
                    kind in (AGG["SUM"], AGG["SUM_OF_SQUARES"], AGG["PRODUCT"]) and not is_float_src):
                if gs.dtype.kind == "f":
                    same = (gs == os_) | (np.isnan(gs) & np.isnan(os_))
                    assert same.all(), f"agg {kind} not bit-exact: {gs[~same][:5]} vs {os_[~same][:5]}"
                else:
                    assert np.array_equal(gs, os_), f"agg {kind} not exact: {gs[gs != os_][:5]} vs {os_[gs != os_][:5]}"
                continue
            if scale is None:
                scale = group_abs_sums(keys, col, null_policy)[oo][ovalid_s] if is_float_src else None
            if kind == AGG["SUM"]:
                tol = TOL_F32_SUM if col.type_id == orc.FLOAT32 else TOL_F64_SUM
                err = np.abs(gs.astype(np.float64) - os_.astype(np.float64))
                bound = tol * scale + np.finfo(gs.dtype).tiny
                assert (err <= bound).all(), f"SUM error {err.max()} above bound"
            elif kind == AGG["MEAN"]:
                tol = TOL_F32_SUM if col.type_id == orc.FLOAT32 else TOL_F64_SUM
                cnt_valid = None
                if is_float_src:
                    # mean = sum / n: the sum's normwise bound divided by n
                    _, cres = orc.groupby(keys, [(col, [(AGG["COUNT_VALID"], 1)])], null_policy)
                    cnt_valid = cres[0][0][0][oo][ovalid_s].astype(np.float64)
                    bound = tol * scale / cnt_valid + 1e-300
                else:
                    bound = 1e-15 * np.abs(os_) + 1e-300
                err = np.abs(gs - os_)
                assert (err <= bound).all(), f"MEAN error {err.max()} above bound"
            elif kind in (AGG["VARIANCE"], AGG["STD"]):
                err = np.abs(gs - os_)
                bound = TOL_VAR * np.abs(os_) + 1e-300
                if is_float_src:
                    # absolute floor from the data scale: tol * sum(x^2)/(n-ddof)
                    sq = group_sq_sums(keys, col, null_policy)[oo][ovalid_s]
                    _, cres = orc.groupby(keys, [(col, [(AGG["COUNT_VALID"], 1)])], null_policy)
                    n = cres[0][0][0][oo][ovalid_s].astype(np.float64)
                    floor = 1e-13 * sq / np.maximum(n - ddof, 1)
                    if kind == AGG["STD"]:
                        floor = np.sqrt(floor)
                    bound = bound + floor
                assert (err <= bound).all(), f"VAR/STD error {err.max()} above bound"
            elif kind == AGG["PRODUCT"]:
                # unordered products of at most ~80 factors: |error| <= 80 * eps * |product|
                tol = TOL_F32_SUM if col.type_id == orc.FLOAT32 else TOL_F64_SUM
                err = np.abs(gs.astype(np.float64) - os_.astype(np.float64))
                assert (err <= tol * np.abs(os_.astype(np.float64)) + 1e-300).all(), f"PRODUCT error {err.max()}"
            elif kind == AGG["SUM_OF_SQUARES"]:
                tol = TOL_F32_SUM if col.type_id == orc.FLOAT32 else TOL_F64_SUM
                err = np.abs(gs.astype(np.float64) - os_.astype(np.float64))
                assert (err <= tol * np.abs(os_.astype(np.float64)) + 1e-300).all()
            else:
                raise AssertionError(f"unhandled kind {kind}")
    return len(og)


def c1_data(n=1_000_000, seed=0):
    """One int32 key uniform over 1,000 values, one int64 value in [-1e6, 1e6]."""
    rng = np.random.default_rng(seed)
    key = rng.integers(0, 1000, n, dtype=np.int32)
    v = rng.integers(-10**6, 10**6, n, endpoint=True, dtype=np.int64)
    return host_col(key), host_col(v)


def c2_data(n=100_000_000, card=10_000_000, seed=0):
    """int64 key over `card` values (1% null), float64 N(0,1) and int32
    U[0,1e4] values, each 5% null."""
    rng = np.random.default_rng(seed)
    key = rng.integers(0, card, n, dtype=np.int64)
    kvalid = rng.random(n) >= 0.01
    f64 = rng.standard_normal(n)
    fvalid = rng.random(n) >= 0.05
    i32 = rng.integers(0, 10_000, n, endpoint=True, dtype=np.int32)
    ivalid = rng.random(n) >= 0.05
    return host_col(key, kvalid), host_col(f64, fvalid), host_col(i32, ivalid)


C2_AGGS = [(AGG["SUM"], 1), (AGG["MEAN"], 1), (AGG["COUNT_VALID"], 1), (AGG["MIN"], 1),
           (AGG["MAX"], 1), (AGG["VARIANCE"], 1)]


def zipf_bounded(rng, n, s, nvals):
    """Bounded Zipf(s) over 1..nvals by inverse CDF of normalised k^-s."""
    w = np.arange(1, nvals + 1, dtype=np.float64) ** (-s)
    cdf = np.cumsum(w)
    cdf /= cdf[-1]
    return (np.searchsorted(cdf, rng.random(n)) + 1).astype(np.int64)


def c3_data(n=50_000_000, seed=0, pool_size=10_000):
    rng = np.random.default_rng(seed)
    key1 = rng.integers(1, 101, n, dtype=np.int32)
    key2 = zipf_bounded(rng, n, 1.1, 1_000_000)
    lens = rng.integers(4, 33, pool_size)
    pool = [bytes(rng.integers(33, 127, L, dtype=np.uint8)) for L in lens]
    idx = rng.integers(0, pool_size, n)
    pool_offs = np.zeros(pool_size + 1, np.int64)
    np.cumsum([len(p) for p in pool], out=pool_offs[1:])
    pool_chars = np.frombuffer(b"".join(pool), np.uint8)
    plen = (pool_offs[1:] - pool_offs[:-1])[idx]
    offs = np.zeros(n + 1, np.int64)
    np.cumsum(plen, out=offs[1:])
    # gather chars of each chosen pool string
    starts = pool_offs[idx]
    rel = np.arange(offs[-1]) - np.repeat(offs[:-1], plen)
    chars = pool_chars[np.repeat(starts, plen) + rel]
    key3 = orc.HostColumn(orc.STRING, chars, None, offs)
    f32 = rng.random(n, dtype=np.float32)
    i64 = rng.integers(-10**9, 10**9, n, endpoint=True, dtype=np.int64)
    return host_col(key1), host_col(key2), key3, host_col(f32), host_col(i64)


# ---------------------------------------------------------------------------
# served request stream: requests drawn from the three plain templates with
# random key / value dtypes and 0-5% nulls

SERVE_DTYPES = [np.int32, np.int64, np.float32, np.float64]
C1_AGGS = [(AGG["SUM"], 1), (AGG["COUNT_VALID"], 1), (AGG["MIN"], 1), (AGG["MAX"], 1)]
C3_AGGS = [(AGG["SUM"], 1), (AGG["COUNT_VALID"], 1), (AGG["MAX"], 1)]


def _typed_values(rng, n, dt, int_lo, int_hi):
    if np.issubdtype(dt, np.integer):
        return rng.integers(int_lo, int_hi, n, endpoint=True).astype(dt)
    return rng.standard_normal(n).astype(dt)


def _nulls(rng, n):
    frac = rng.uniform(0.0, 0.05)
    return (rng.random(n) >= frac) if frac > 0.005 else None


def served_request(rng, template: int, n: int):
    """(keys, requests, null_policy) of one request of template 1, 2 or 3."""
    if template == 1:
        kt, vt = (SERVE_DTYPES[i] for i in rng.integers(0, 4, 2))
        key = rng.integers(0, max(1, min(1000, n)), n).astype(kt)
        v = _typed_values(rng, n, vt, -10**6, 10**6)
        return [host_col(key, _nulls(rng, n))], [(host_col(v, _nulls(rng, n)), C1_AGGS)], 0
    if template == 2:
        kt, ft, it = (SERVE_DTYPES[i] for i in rng.integers(0, 4, 3))
        key = rng.integers(0, max(1, n // 10), n).astype(kt)
        f = _typed_values(rng, n, ft, -10**6, 10**6)
        i = _typed_values(rng, n, it, 0, 10**4)
        return ([host_col(key, rng.random(n) >= 0.01)],
                [(host_col(f, _nulls(rng, n)), C2_AGGS), (host_col(i, _nulls(rng, n)), C2_AGGS)], 0)
    k1, k2, k3, f32, i64 = c3_data(n=n, seed=int(rng.integers(1 << 30)), pool_size=min(10_000, max(1, n)))
    ft, it = (SERVE_DTYPES[i] for i in rng.integers(0, 4, 2))
    f = host_col(f32.data.astype(ft), _nulls(rng, n))
    i = host_col(i64.data.astype(it) if np.issubdtype(it, np.integer) else i64.data.astype(it), _nulls(rng, n))
    return [k1, k2, k3], [(f, C3_AGGS), (i, C3_AGGS)], 0

"""Shared helpers of the group-by tests: data generators, the GPU call through
the C ABI, the oracle call, and sort-by-key comparison (the protocol of
cpp/tests/groupby/groupby_test_util.cpp:34-98: output order is unspecified,
so both sides are sorted by key before comparing)."""
from __future__ import annotations

from typing import List, Optional, Sequence

import numpy as np

from oracle import oracle as orc

AGG = {
    "SUM": 0, "PRODUCT": 1, "MIN": 2, "MAX": 3, "COUNT_VALID": 4, "COUNT_ALL": 5,
    "SUM_OF_SQUARES": 8, "MEAN": 9, "VARIANCE": 11, "STD": 12, "MEDIAN": 13, "QUANTILE": 14,
    "ARGMAX": 15, "ARGMIN": 16, "NUNIQUE": 17, "NTH_ELEMENT": 18,
}
QUANTILE_KINDS = (13, 14)  # a request pair (QUANTILE, (quantiles, interpolation)); (MEDIAN, anything)


def n_quantiles(kind, param) -> int:
    """Result rows per group of an aggregation."""
    return len(param[0]) if kind == AGG["QUANTILE"] else 1

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
    import torch
    from cudf_amd import Column, TypeId
    from cudf_amd.column import pack_validity
    if c.type_id == orc.STRING:
        # offsets and chars copied as arrays (int32 offsets when they fit)
        offs = np.asarray(c.offsets, dtype=np.int64)
        large = offs[-1] >= 2**31
        offs = offs if large else offs.astype(np.int32)
        chars = np.asarray(c.data, dtype=np.uint8)
        dev = torch.device("cuda", torch.cuda.current_device())
        data = torch.from_numpy(chars.copy() if chars.size else np.zeros(1, np.uint8)).to(dev)
        mask, nulls = None, 0
        if c.valid is not None:
            mask = torch.from_numpy(pack_validity(c.valid).view(np.int32).copy()).to(dev)
            nulls = int(len(c.valid) - np.count_nonzero(c.valid))
        return Column(TypeId.STRING, c.size, data, mask, nulls, offsets=torch.from_numpy(offs.copy()).to(dev))
    return Column.from_numpy(c.data, c.valid, type_id=c.type_id)


def host_keys(col):
    """(values, valid) of a device key column; strings as ``StrKeys``."""
    from cudf_amd import TypeId
    if col.type_id == TypeId.STRING:
        offs = col.offsets.cpu().numpy().astype(np.int64)[col.offset: col.offset + col.size + 1]
        chars = col.data.cpu().numpy().view(np.uint8) if col.data is not None else np.zeros(0, np.uint8)
        return StrKeys(offs, chars), col.validity()
    return col.to_numpy()


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
    kcols = [host_keys(c) for c in out_keys.columns()]
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
    if kind == AGG["QUANTILE"]:
        return Aggregation(kind, quantiles=param[0], interpolation=param[1])
    if kind == AGG["MEDIAN"]:
        return Aggregation(kind)
    return Aggregation(kind, param)


def take(col: orc.HostColumn, rows: np.ndarray) -> orc.HostColumn:
    """Rows ``rows`` of a host column (strings re-packed)."""
    valid = None if col.valid is None else col.valid[rows]
    if col.type_id == orc.STRING:
        sk = StrKeys(col.offsets, col.data).gather(rows)
        chars = sk.chars if len(sk.chars) else np.zeros(1, np.uint8)
        return orc.HostColumn(orc.STRING, chars, valid, sk.offsets)
    return orc.HostColumn(col.type_id, col.data[rows], valid)


def _orc_groupby_any(keys, requests, null_policy):
    """orc.groupby with MEDIAN / QUANTILE answered by orc.group_quantiles
    (both number the groups in first-occurrence order)."""
    plain = []
    for col, aggs in requests:
        plain.append((col, [a for a in aggs if a[0] not in QUANTILE_KINDS] or [(AGG["COUNT_ALL"], 1)]))
    rep, pres = orc.groupby(keys, plain, null_policy)
    results = []
    for r, (col, aggs) in enumerate(requests):
        out, j = [], 0
        for kind, param in aggs:
            if kind in QUANTILE_KINDS:
                qs, interp = ((0.5,), orc.LINEAR) if kind == AGG["MEDIAN"] else param
                qrep, v, valid = orc.group_quantiles(keys, col, qs, interp, null_policy)
                assert np.array_equal(qrep, rep)
                out.append((v.reshape(-1), valid.reshape(-1)))
            else:
                out.append(pres[r][j])
                j += 1
        results.append(out)
    return rep, results


# Rows per shard below which the oracle runs in one call.
_SHARD_MIN_ROWS = 4_000_000


def _oracle_sharded(keys, requests, null_policy, shards):
    """The oracle on key-disjoint shards in parallel threads (ctypes releases
    the GIL): rows go to shard row_hash % shards, so every group lies in one
    shard and the union of the shards' groups is the full result.  Returns
    what ``orc.groupby`` returns, with rep rows of the whole input."""
    from concurrent.futures import ThreadPoolExecutor
    h = orc.hash_rows(list(keys))
    sid = (h % np.uint32(shards)).astype(np.uint8)
    order = np.argsort(sid, kind="stable")
    bounds = np.searchsorted(sid[order], np.arange(shards + 1))

    def run(i):
        rows = order[bounds[i]:bounds[i + 1]]
        sk = [take(k, rows) for k in keys]
        sr = [(take(c, rows), aggs) for c, aggs in requests]
        rep, res = _orc_groupby_any(sk, sr, null_policy)
        return rows[rep], res

    with ThreadPoolExecutor(max_workers=shards) as ex:
        parts = list(ex.map(run, range(shards)))
    rep = np.concatenate([p[0] for p in parts])
    results = []
    for r, (_, aggs) in enumerate(requests):
        results.append([(np.concatenate([p[1][r][a][0] for p in parts]),
                         np.concatenate([p[1][r][a][1] for p in parts])) for a in range(len(aggs))])
    return rep, results


def oracle_groupby(keys: Sequence[orc.HostColumn], requests, null_policy: int = 0):
    n = keys[0].size if keys else 0
    if n >= 2 * _SHARD_MIN_ROWS:
        rep, res = _oracle_sharded(keys, requests, null_policy, int(min(16, n // _SHARD_MIN_ROWS)))
    else:
        rep, res = _orc_groupby_any(keys, requests, null_policy)
    kcols = []
    for k in keys:
        if k.type_id == orc.STRING:
            vals = StrKeys(k.offsets, k.data).gather(rep)
        else:
            vals = k.data[rep]
        valid = None if k.valid is None else k.valid[rep]
        kcols.append((vals, valid))
    return kcols, res


class StrKeys:
    """A string key column as arrays: ``offsets`` (int64, size + 1, absolute
    into ``chars``) and ``chars`` (uint8).  Indexing gives ``bytes``."""

    def __init__(self, offsets, chars):
        self.offsets = np.asarray(offsets, dtype=np.int64)
        self.chars = np.asarray(chars, dtype=np.uint8)

    @staticmethod
    def from_list(strs) -> "StrKeys":
        lens = np.array([len(x) for x in strs], dtype=np.int64)
        offs = np.zeros(len(strs) + 1, np.int64)
        np.cumsum(lens, out=offs[1:])
        return StrKeys(offs, np.frombuffer(b"".join(bytes(x) for x in strs), np.uint8))

    def __len__(self):
        return len(self.offsets) - 1

    def __getitem__(self, i):
        return bytes(self.chars[self.offsets[i]:self.offsets[i + 1]])

    def gather(self, rows) -> "StrKeys":
        rows = np.asarray(rows, dtype=np.int64)
        lens = (self.offsets[1:] - self.offsets[:-1])[rows]
        offs = np.zeros(len(rows) + 1, np.int64)
        np.cumsum(lens, out=offs[1:])
        src = np.repeat(self.offsets[:-1][rows] - offs[:-1], lens) + np.arange(offs[-1])
        return StrKeys(offs, self.chars[src] if offs[-1] else np.zeros(0, np.uint8))

    def __iter__(self):
        return (self[i] for i in range(len(self)))

    def words(self):
        """uint64 arrays whose lexicographic order is the strings' byte-wise
        order: the bytes zero-padded, big-endian in 8-byte words, then the
        length."""
        lens = self.offsets[1:] - self.offsets[:-1]
        n = len(lens)
        nw = (int(lens.max()) + 7) // 8 if n else 0
        mat = np.zeros((n, nw * 8), np.uint8)
        if n and self.offsets[-1] > self.offsets[0]:
            rows = np.repeat(np.arange(n), lens)
            cols = np.arange(self.offsets[-1] - self.offsets[0]) - np.repeat(self.offsets[:-1] - self.offsets[0], lens)
            mat[rows, cols] = self.chars[self.offsets[0]:self.offsets[-1]]
        w = mat.view(">u8").astype(np.uint64) if nw else np.zeros((n, 0), np.uint64)
        return [w[:, j] for j in range(nw)] + [lens.astype(np.uint64)]

    def fingerprint(self) -> np.ndarray:
        if getattr(self, "_fp", None) is None:
            self._fp = self._fingerprint()
        return self._fp

    def _fingerprint(self) -> np.ndarray:
        """64-bit fingerprint of every string (a position-keyed multiply-add
        hash of the bytes and the length): equal strings give equal values,
        so it orders both sides of a comparison the same way; equality itself
        is checked on the bytes."""
        lens = self.offsets[1:] - self.offsets[:-1]
        n = len(lens)
        h = lens.astype(np.uint64) * np.uint64(0x9E3779B97F4A7C15)
        start, end = int(self.offsets[0]), int(self.offsets[-1])
        if n == 0 or end == start:
            return h
        pos = np.arange(end - start) - np.repeat(self.offsets[:-1] - start, lens)
        mult = _position_multipliers(int(lens.max()))
        prod = self.chars[start:end].astype(np.uint64) * mult[pos]
        # per-string sums as differences of a prefix sum (mod 2^64), so a
        # string's value does not depend on its neighbours
        csum = np.zeros(end - start + 1, np.uint64)
        np.cumsum(prod, out=csum[1:])
        sums = csum[self.offsets[1:] - start] - csum[self.offsets[:-1] - start]
        return h ^ sums


_MULTS = np.zeros(0, np.uint64)


def _position_multipliers(n):
    global _MULTS
    if len(_MULTS) < n:
        rng = np.random.default_rng(12345)
        _MULTS = rng.integers(0, 2**63, max(n, 64), dtype=np.uint64) * np.uint64(2) + np.uint64(1)
    return _MULTS


def _as_keys(vals):
    if isinstance(vals, (list, tuple)):
        return StrKeys.from_list(vals)
    return vals


def _key_arrays(vals, valid):
    """Arrays (most significant first) that order one key column and decide
    key equality under the group-by's rules: nulls equal each other and sort
    last, NaN equals NaN, -0.0 equals 0.0, strings byte-wise."""
    vals = _as_keys(vals)
    n = len(vals)
    v = np.ones(n, bool) if valid is None else np.asarray(valid, bool)
    if isinstance(vals, StrKeys):
        arrs = [np.where(v, vals.fingerprint(), np.uint64(0))]
    else:
        a = np.asarray(vals)
        if a.dtype.kind == "f":
            a = a.astype(np.float64)
            nan = np.isnan(a)
            arrs = [nan.astype(np.uint8), np.where(nan | ~v, 0.0, a) + 0.0]
        elif a.dtype.kind == "b":
            arrs = [np.where(v, a.astype(np.uint8), 0)]
        else:
            arrs = [np.where(v, a, 0)]
    return [(~v).astype(np.uint8)] + arrs


def sort_order(kcols, exact: bool = True) -> np.ndarray:
    """Order of the rows of key columns ``[(values, valid), ...]``.  With
    ``exact`` the rows are sorted by key (strings byte-wise); otherwise by a
    64-bit hash of the key (the same order for the same set of keys on either
    side of a comparison, much faster for millions of groups), falling back
    to the exact sort if two different keys share a hash."""
    if not kcols:
        return np.zeros(0, np.int64)
    if exact:
        arrays = []
        for vals, valid in kcols:
            vals = _as_keys(vals)
            if isinstance(vals, StrKeys):
                v = np.ones(len(vals), bool) if valid is None else np.asarray(valid, bool)
                arrays += [(~v).astype(np.uint8)] + [np.where(v, w, np.uint64(0)) for w in vals.words()]
            else:
                arrays += _key_arrays(vals, valid)
        return np.lexsort(arrays[::-1]).astype(np.int64)
    arrays = []
    for vals, valid in kcols:
        arrays += _key_arrays(vals, valid)
    h = np.zeros(len(arrays[0]), np.uint64)
    for a in arrays:
        x = np.ascontiguousarray(a)
        x = x.view(np.uint64) if x.dtype.itemsize == 8 else x.astype(np.uint64)
        h = (h ^ x) * np.uint64(0x9E3779B97F4A7C15)
        h ^= h >> np.uint64(29)
    order = np.argsort(h, kind="stable")
    hs = h[order]
    if np.any(hs[1:] == hs[:-1]):
        return sort_order(kcols, exact=True)
    return order.astype(np.int64)


def assert_keys_equal(gk, ok, og, oo):
    """gk/ok: key columns; og/oo: sort orders.  Keys must be equal row by row
    (nulls equal, NaN equal NaN, -0.0 equal 0.0, strings byte-wise)."""
    assert len(og) == len(oo), f"group count differs: gpu {len(og)} vs oracle {len(oo)}"
    for c, ((gv, gval), (ov, oval)) in enumerate(zip(gk, ok)):
        gv, ov = _as_keys(gv), _as_keys(ov)
        for x, y in zip(_key_arrays(gv, gval), _key_arrays(ov, oval)):
            bad = np.flatnonzero(x[og] != y[oo])
            assert bad.size == 0, f"key column {c} sorted row {bad[0]}: gpu {gv[og[bad[0]]]!r} vs oracle {ov[oo[bad[0]]]!r}"
        if isinstance(gv, StrKeys):
            m = np.ones(len(og), bool) if gval is None else np.asarray(gval, bool)[og]
            a, b = gv.gather(og[m]), ov.gather(oo[m])
            assert np.array_equal(np.diff(a.offsets), np.diff(b.offsets)), f"key column {c}: string lengths differ"
            bad = np.flatnonzero(a.chars != b.chars)
            assert bad.size == 0, f"key column {c}: string bytes differ at char {bad[0]}"


_FP_KINDS = ("SUM", "MEAN", "VARIANCE", "STD", "PRODUCT", "SUM_OF_SQUARES")


def _scale_requests(requests):
    """Extra oracle requests giving the scales of the floating-point bounds:
    per floating-point column with such aggregations, sum |x|, sum x^2 and
    the valid count per group.  Returns (requests, index of each column's
    first extra request or None)."""
    extra, where = [], []
    for col, aggs in requests:
        is_float = col.type_id in (orc.FLOAT32, orc.FLOAT64)
        if is_float and any(k in [AGG[n] for n in _FP_KINDS] for k, _ in aggs):
            x = col.data.astype(np.float64)
            where.append(len(requests) + len(extra))
            extra.append((orc.HostColumn(orc.FLOAT64, np.abs(x), col.valid), [(AGG["SUM"], 1)]))
            extra.append((orc.HostColumn(orc.FLOAT64, x * x, col.valid), [(AGG["SUM"], 1)]))
            extra.append((col, [(AGG["COUNT_VALID"], 1)]))
        else:
            where.append(None)
    return list(requests) + extra, where


def assert_parity(keys, requests, null_policy=0, gpu=None, stats=None):
    """Run both sides (unless ``gpu`` results are given) and compare under the
    parity bar, after sorting both by key.  Returns the number of groups.
    ``stats`` (a dict), when given, receives per FP64 SUM / MEAN result the
    largest relative error and the number of groups whose relative error
    exceeds 1e-12 (a report; the bound asserted is the normwise one)."""
    g_keys, g_res = gpu if gpu is not None else gpu_groupby(keys, requests, null_policy)
    all_reqs, where = _scale_requests(requests)
    o_keys, o_res = oracle_groupby(keys, all_reqs, null_policy)
    exact = len(o_keys[0][0]) < 1_000_000 if o_keys else True
    og, oo = sort_order(g_keys, exact), sort_order(o_keys, exact)
    assert_keys_equal(g_keys, o_keys, og, oo)
    for r, (col, aggs) in enumerate(requests):
        is_float_src = col.type_id in (orc.FLOAT32, orc.FLOAT64)
        scale_abs = scale_sq = n_valid = None
        if where[r] is not None:
            w = where[r]
            scale_abs, scale_sq, n_valid = (o_res[w][0][0][oo], o_res[w + 1][0][0][oo],
                                            o_res[w + 2][0][0][oo].astype(np.float64))
        for a, (kind, ddof) in enumerate(aggs):
            gv, gval = g_res[r][a]
            ov, oval = o_res[r][a]
            assert len(gv) == len(ov)
            if kind in QUANTILE_KINDS:
                # G * nq rows, group-major: both sides in key order, bit-exact
                nq = n_quantiles(kind, ddof)
                gq = (og[:, None] * nq + np.arange(nq)).reshape(-1)
                oq = (oo[:, None] * nq + np.arange(nq)).reshape(-1)
                gm = np.ones(len(gv), bool) if gval is None else gval
                assert np.array_equal(gm[gq], oval[oq]), f"quantile validity differs (request {r} agg {a})"
                x, y = gv[gq][oval[oq]], ov[oq][oval[oq]]
                same = (x == y) | (np.isnan(x) & np.isnan(y))
                assert same.all(), f"quantile not bit-exact: {x[~same][:5]} vs {y[~same][:5]}"
                continue
            gvalid = np.ones(len(gv), bool) if gval is None else gval
            gvalid_s, ovalid_s = gvalid[og], oval[oo]
            bad = np.flatnonzero(gvalid_s != ovalid_s)
            assert bad.size == 0, f"validity differs for request {r} agg {kind} at sorted group {bad[0]}"
            gs, os_ = gv[og][gvalid_s], ov[oo][ovalid_s]
            assert gs.dtype == os_.dtype, f"dtype {gs.dtype} vs {os_.dtype} (agg {kind})"
            if kind in (AGG["COUNT_VALID"], AGG["COUNT_ALL"], AGG["MIN"], AGG["MAX"], AGG["ARGMIN"],
                        AGG["ARGMAX"], AGG["NUNIQUE"], AGG["NTH_ELEMENT"]) or (
                    kind in (AGG["SUM"], AGG["SUM_OF_SQUARES"], AGG["PRODUCT"]) and not is_float_src):
                if gs.dtype.kind == "f":
                    same = (gs == os_) | (np.isnan(gs) & np.isnan(os_))
                    assert same.all(), f"agg {kind} not bit-exact: {gs[~same][:5]} vs {os_[~same][:5]}"
                else:
                    assert np.array_equal(gs, os_), f"agg {kind} not exact: {gs[gs != os_][:5]} vs {os_[gs != os_][:5]}"
                continue
            tol = TOL_F32_SUM if col.type_id == orc.FLOAT32 else TOL_F64_SUM
            g64, o64 = gs.astype(np.float64), os_.astype(np.float64)
            # NaN (a NaN among the group's values) and infinities must match
            # exactly; equal results pass whatever their bound
            gnan, onan = np.isnan(g64), np.isnan(o64)
            assert np.array_equal(gnan, onan), f"agg {kind}: NaN results differ at {np.flatnonzero(gnan != onan)[:5]}"
            same = (g64 == o64) | gnan
            with np.errstate(invalid="ignore"):  # inf - inf where both are the same infinity
                err = np.where(same, 0.0, np.abs(np.where(same, 0.0, g64 - o64)))
            if stats is not None and col.type_id == orc.FLOAT64 and kind in (AGG["SUM"], AGG["MEAN"]):
                rel = err / np.maximum(np.abs(o64), 1e-300)
                stats[(r, kind)] = {"groups": int(len(rel)), "max_rel": float(rel.max(initial=0.0)),
                                    "over_1e-12": int((rel > 1e-12).sum()),
                                    "max_normwise": float((err / np.maximum(scale_abs[ovalid_s], 1e-300)).max(initial=0.0))}
            if kind == AGG["SUM"]:
                if is_float_src:
                    bound = tol * scale_abs[ovalid_s] + np.finfo(gs.dtype).tiny
                else:
                    bound = np.zeros_like(err)
                assert (same | (err <= bound)).all(), f"SUM error {err.max()} above bound"
            elif kind == AGG["MEAN"]:
                if is_float_src:
                    # mean = sum / n: the sum's normwise bound divided by n
                    bound = tol * scale_abs[ovalid_s] / n_valid[ovalid_s] + 1e-300
                else:
                    bound = 1e-15 * np.abs(o64) + 1e-300
                assert (same | (err <= bound)).all(), f"MEAN error {err.max()} above bound"
            elif kind in (AGG["VARIANCE"], AGG["STD"]):
                bound = TOL_VAR * np.abs(o64) + 1e-300
                if is_float_src:
                    # absolute floor from the data scale: tol * sum(x^2)/(n-ddof)
                    floor = 1e-13 * scale_sq[ovalid_s] / np.maximum(n_valid[ovalid_s] - ddof, 1)
                    if kind == AGG["STD"]:
                        floor = np.sqrt(floor)
                    bound = bound + floor
                assert (same | (err <= bound)).all(), f"VAR/STD error {err.max()} above bound"
            elif kind == AGG["PRODUCT"]:
                # unordered products of at most ~80 factors: |error| <= 80 * eps * |product|
                assert (same | (err <= tol * np.abs(o64) + 1e-300)).all(), f"PRODUCT error {err.max()}"
            elif kind == AGG["SUM_OF_SQUARES"]:
                assert (same | (err <= tol * np.abs(o64) + 1e-300)).all()
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

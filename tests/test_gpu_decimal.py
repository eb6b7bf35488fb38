"""Fixed-point (DECIMAL32 / DECIMAL64) value and key columns.

Result types follow cudf's target types (aggregation.hpp:1211-1331): SUM,
MEAN, MIN, MAX and NTH_ELEMENT keep the source type and scale (SUM and MEAN
on the integer representation: the sum wraps in the representation's width,
the mean is the sum divided by the count, truncated), COUNT is int32,
ARGMIN / ARGMAX are rows; PRODUCT is a logic error (sum_tests.cpp:230-232);
SUM_OF_SQUARES / VARIANCE / STD / MEDIAN of fixed point and every DECIMAL128
column are data-type errors here.  Known answers from the reference's
fixed-point group-by tests (tests/golden/decimal_kats.json); random inputs
against numpy on the representation."""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

KATS = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "decimal_kats.json")))


def _run(key_cols, value_col, kinds, policy=0, keys_sorted=False, extra=None):
    from cudf_amd import GroupBy, GroupByRequest, NullPolicy, Sorted, Table
    from cudf_amd.aggregation import Aggregation
    reqs = [GroupByRequest(value_col, [Aggregation(k) if not isinstance(k, tuple) else Aggregation(*k) for k in kinds])]
    if extra is not None:
        reqs.append(extra)
    gb = GroupBy(Table(key_cols), NullPolicy(policy), Sorted.YES if keys_sorted else Sorted.NO)
    return gb.aggregate(reqs)


def _dec(arr, tid, scale, valid=None):
    from cudf_amd import Column
    return Column.from_numpy(arr, valid, type_id=tid, scale=scale)


@pytest.mark.parametrize("tid_name,rep", [("DECIMAL32", np.int32), ("DECIMAL64", np.int64)])
@pytest.mark.parametrize("path", ["hash", "sort"])
def test_reference_known_answers(tid_name, rep, path):
    from cudf_amd import Column, GroupByRequest, TypeId
    from cudf_amd.aggregation import Aggregation, Kind
    tid = TypeId[tid_name]
    keys = np.array(KATS["keys"], np.int32)
    for scale in KATS["scales"]:
        vals = _dec(np.array(KATS["vals"], rep), tid, scale)
        for case in KATS["cases"]:
            # the sort-based group-by: a MEDIAN of another column routes the call there
            extra = GroupByRequest(Column.from_numpy(np.ones(len(keys))), [Aggregation(Kind.MEDIAN)]) \
                if path == "sort" else None
            out_keys, res = _run([Column.from_numpy(keys)], vals, [Kind[case["agg"]]], extra=extra)
            k = out_keys.columns()[0].to_numpy()[0]
            col = res[0].columns()[0]
            v = col.to_numpy()[0]
            order = np.argsort(k)
            assert list(k[order]) == KATS["expect_keys"]
            assert list(v[order]) == case["expect"], (case, scale)
            if case["same_type"]:
                assert col.type_id == tid and col.scale == scale, (col.type_id, col.scale)
            else:
                assert col.type_id == TypeId.INT32


def _expected(keys, kvalid, rep, vvalid, policy, wrap32):
    """Per key (None = the null key): sum (wrapped), truncated mean, min, max,
    count, first argmin, first argmax, distinct count, on the representation."""
    groups = {}
    for r in range(len(keys)):
        if not kvalid[r] and policy == 0:
            continue
        groups.setdefault(int(keys[r]) if kvalid[r] else None, []).append(r)
    out = {}
    for key, rows in groups.items():
        vr = [r for r in rows if vvalid[r]]
        if not vr:
            out[key] = (None, None, None, None, 0, None, None, 0)
            continue
        s = int(sum(int(rep[r]) for r in vr))
        if wrap32:
            s = (s + 2**31) % 2**32 - 2**31
        else:
            s = (s + 2**63) % 2**64 - 2**63
        m = abs(s) // len(vr) * (1 if s >= 0 else -1)  # truncated toward zero
        mn = min(int(rep[r]) for r in vr)
        mx = max(int(rep[r]) for r in vr)
        amin = next(r for r in vr if int(rep[r]) == mn)
        amax = next(r for r in vr if int(rep[r]) == mx)
        out[key] = (s, m, mn, mx, len(vr), amin, amax, len({int(rep[r]) for r in vr}))
    return out


@pytest.mark.parametrize("tid_name,dtype,lo,hi", [("DECIMAL32", np.int32, -2**31, 2**31 - 1),
                                                  ("DECIMAL64", np.int64, -2**62, 2**62)])
@pytest.mark.parametrize("policy", [0, 1])
@pytest.mark.parametrize("strategy", ["auto", "partitioned", "sort"])
def test_random_against_numpy(tid_name, dtype, lo, hi, policy, strategy):
    """Values spanning the representation (so SUM wraps), nulls in keys and
    values, every supported kind; the device-wide table, the partitioned path
    and the sort-based group-by."""
    from cudf_amd import Column, GroupByRequest, TypeId, _lib
    from cudf_amd.aggregation import Aggregation, Kind
    tid = TypeId[tid_name]
    rng = np.random.default_rng(1000 + 100 * len(tid_name) + 10 * policy + ["auto", "partitioned", "sort"].index(strategy))
    n = 60_000 if strategy != "partitioned" else 1_200_000
    card = 3_000 if strategy != "partitioned" else 300_000
    keys = rng.integers(0, card, n).astype(np.int32)
    kvalid = rng.random(n) > 0.02
    pool = rng.integers(lo, hi, 5_000, endpoint=True, dtype=np.int64)  # repeated values: ties and NUNIQUE
    rep = pool[rng.integers(0, len(pool), n)].astype(dtype)
    vvalid = rng.random(n) > 0.1
    kinds = [Kind.SUM, Kind.MEAN, Kind.MIN, Kind.MAX, Kind.COUNT_VALID]
    if strategy != "partitioned":  # ARGMIN / ARGMAX run on the device-wide table or the runs
        kinds += [Kind.ARGMIN, Kind.ARGMAX]
    if strategy == "sort":
        kinds.append(Kind.NUNIQUE)
    _lib.set_strategy(_lib.STRATEGY_PARTITIONED if strategy == "partitioned" else _lib.STRATEGY_AUTO)
    try:
        out_keys, res = _run([Column.from_numpy(keys, kvalid)], _dec(rep, tid, -3, vvalid), kinds, policy)
        st = _lib.last_stats()
    finally:
        _lib.set_strategy(_lib.STRATEGY_AUTO)
    if strategy == "partitioned":
        assert st["strategy"] == 2, st
    if strategy == "sort":
        assert st["strategy"] == 4, st
    exp = _expected(keys, kvalid, rep, vvalid, policy, dtype == np.int32)
    kv, kval = out_keys.columns()[0].to_numpy()
    kval = np.ones(len(kv), bool) if kval is None else kval
    cols = res[0].columns()
    for c in cols[:4]:
        assert c.type_id == tid and c.scale == -3
    got = [c.to_numpy() for c in cols]
    assert len(kv) == len(exp)
    for g in range(len(kv)):
        key = int(kv[g]) if kval[g] else None
        e = exp[key]
        for a, kind in enumerate(kinds):
            vals, valid = got[a]
            ok = True if valid is None else bool(valid[g])
            want = e[7] if kind == Kind.NUNIQUE else e[a]
            if kind in (Kind.COUNT_VALID, Kind.NUNIQUE):
                assert ok and int(vals[g]) == want, (key, kind)
            elif want is None:
                assert not ok, (key, kind)
            else:
                assert ok and int(vals[g]) == want, (key, kind, int(vals[g]), want)


def test_decimal_keys_keep_their_type_and_scale():
    from cudf_amd import Column, TypeId
    from cudf_amd.aggregation import Kind
    rng = np.random.default_rng(31)
    n = 50_000
    k = rng.integers(-500, 500, n).astype(np.int64)
    v = rng.integers(0, 100, n).astype(np.int64)
    out_keys, res = _run([_dec(k, TypeId.DECIMAL64, -2)], Column.from_numpy(v), [Kind.SUM])
    kc = out_keys.columns()[0]
    assert kc.type_id == TypeId.DECIMAL64 and kc.scale == -2
    got = dict(zip(kc.to_numpy()[0].tolist(), res[0].columns()[0].to_numpy()[0].tolist()))
    want = {}
    for a, b in zip(k.tolist(), v.tolist()):
        want[a] = want.get(a, 0) + b
    assert got == want


def test_unsupported_fixed_point_requests():
    from cudf_amd import Column, TypeId, _lib
    from cudf_amd.aggregation import Kind
    keys = Column.from_numpy(np.array(KATS["keys"], np.int32))
    vals = _dec(np.array(KATS["vals"], np.int32), TypeId.DECIMAL32, -1)
    with pytest.raises(_lib.LogicError):
        _run([keys], vals, [Kind.PRODUCT])
    for kind in (Kind.SUM_OF_SQUARES, Kind.VARIANCE, Kind.STD, Kind.MEDIAN):
        with pytest.raises(_lib.DataTypeError):
            _run([keys], vals, [kind])
    import torch
    from cudf_amd import Column as C
    d128 = C(TypeId.DECIMAL128, 10, torch.zeros(160, dtype=torch.uint8, device="cuda"))
    with pytest.raises(_lib.DataTypeError):
        _run([keys], d128, [Kind.COUNT_VALID])

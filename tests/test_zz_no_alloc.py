"""The allocation-free group-by (mig_groupby_aggregate_into) against
mig_groupby_aggregate through the C ABI: two calls in a row on the device's
shared scratch, the second with a smaller input.  Keys, counts, integer sums,
min, max and every validity bit equal the existing entry point's bit for bit;
every result is also checked against the oracle under the parity bar
(float64 sums and means within 1e-12 of the group's sum of magnitudes,
variance within 1e-10), since float sums from unordered atomics differ in
their last bits between any two runs.  Also the errors for outputs that are
too small and for a call whose hash table would not fit the fixed scratch."""
import ctypes

import numpy as np
import pytest

from gbutil import AGG, assert_parity, host_col, sort_order, to_device

pytestmark = pytest.mark.gpu

AGGS = [(AGG["SUM"], 1), (AGG["COUNT_VALID"], 1), (AGG["MIN"], 1), (AGG["MAX"], 1), (AGG["MEAN"], 1),
        (AGG["VARIANCE"], 1)]
# result widths of (int32 values, float64 values) per aggregation above
WIDTHS = [(8, 8), (4, 4), (4, 8), (4, 8), (8, 8), (8, 8)]


def _inputs(n, card, seed):
    rng = np.random.default_rng(seed)
    key = host_col(rng.integers(0, card, n).astype(np.int64), rng.random(n) > 0.02)
    i32 = host_col(rng.integers(-10**6, 10**6, n).astype(np.int32), rng.random(n) > 0.05)
    f64 = host_col(rng.standard_normal(n) * 100.0 + 3.0, rng.random(n) > 0.05)
    return key, i32, f64


def _requests(dk, dvals):
    from cudf_amd import _lib
    keep = []
    reqs = (_lib.AggregationRequest * len(dvals))()
    for r, v in enumerate(dvals):
        aggs = (_lib.Aggregation * len(AGGS))(*[_lib.Aggregation(k, d) for k, d in AGGS])
        keep.append(aggs)
        reqs[r].values = v.view()
        reqs[r].aggregations = aggs
        reqs[r].num_aggregations = len(AGGS)
    keys = (_lib.ColumnView * 1)(dk.view())
    return keys, reqs, keep


def _reference(key, vals):
    """The existing entry point (through the same C ABI) on its device-wide
    table path: ((keys, valid), [(type id, (values, valid), null count)])."""
    from cudf_amd import GroupBy, GroupByRequest, Table, _lib
    from cudf_amd.aggregation import Aggregation
    _lib.set_strategy(_lib.STRATEGY_GLOBAL)
    try:
        gb = GroupBy(Table([key]))
        out_keys, out_res = gb.aggregate([GroupByRequest(v, [Aggregation(k, d) for k, d in AGGS]) for v in vals])
    finally:
        _lib.set_strategy(_lib.STRATEGY_AUTO)
    res = [(int(c.type_id), c.to_numpy(), c.null_count) for t in out_res for c in t.columns()]
    return out_keys.columns()[0].size, out_keys.columns()[0].to_numpy(), res


def _np_type(tid):
    return {3: np.int32, 4: np.int64, 10: np.float64}[tid]


def _read(col, data_t, mask_t, g):
    """(values, valid) of a returned column from the caller's torch buffers."""
    vals = data_t.cpu().numpy().view(_np_type(col.type_id))[:g].copy()
    valid = None
    if col.null_mask:
        valid = np.unpackbits(mask_t.cpu().numpy(), bitorder="little")[:g].astype(bool)
    return vals, valid


def _into(keys, reqs, capacity):
    """The allocation-free entry point into torch buffers of `capacity` rows."""
    import torch
    from cudf_amd import _lib
    lib = _lib.lib()
    bufs = []

    def col(width):
        d = torch.empty(max(1, capacity * width), dtype=torch.uint8, device="cuda")
        m = torch.empty(max(8, ((capacity + 63) // 64) * 8), dtype=torch.uint8, device="cuda")
        bufs.extend([d, m])
        c = _lib.OutColumn()
        c.data, c.data_bytes = d.data_ptr(), capacity * width
        c.null_mask, c.mask_bytes = m.data_ptr(), ((capacity + 63) // 64) * 8
        return c

    out_keys = (_lib.OutColumn * 1)(col(8))
    outs = []
    for r in range(len(reqs)):
        for a in range(len(AGGS)):
            outs.append(col(WIDTHS[a][r]))
    out_res = (_lib.OutColumn * len(outs))(*outs)
    g = ctypes.c_int32(-1)
    torch.cuda.synchronize()
    st = lib.mig_groupby_aggregate_into(keys, 1, 0, reqs, len(reqs), out_keys, out_res, ctypes.byref(g), None)
    return st, g.value, out_keys, out_res, bufs


def _compare(n, card, seed):
    print(f"no-alloc: n={n} card={card} seed={seed}", flush=True)
    key, i32, f64 = _inputs(n, card, seed)
    dk, dvals = to_device(key), [to_device(i32), to_device(f64)]
    keys, reqs, keep = _requests(dk, dvals)
    g_ref, k_ref, res_ref = _reference(dk, dvals)
    st, g, out_keys, out_res, bufs = _into(keys, reqs, n)
    from cudf_amd import _lib
    _lib.check(st)
    assert g == g_ref
    assert out_keys[0].size == g and out_keys[0].type_id == 4
    k_new = _read(out_keys[0], bufs[0], bufs[1], g)
    o_ref = sort_order([k_ref], exact=True)
    o_new = sort_order([k_new], exact=True)
    assert np.array_equal(k_ref[0][o_ref], k_new[0][o_new])
    got = []
    for i, (tid, (vr, valr), nc) in enumerate(res_ref):
        c = out_res[i]
        assert c.type_id == tid and c.size == g and c.null_count == nc, (i, c.type_id, c.size, c.null_count)
        assert (c.null_mask is None) == (valr is None), i
        vn, valn = _read(c, bufs[2 + 2 * i], bufs[3 + 2 * i], g)
        got.append((vn, valn))
        if valr is not None:
            assert np.array_equal(valr[o_ref], valn[o_new]), i
        a = i % len(AGGS)
        if tid != 10 or AGGS[a][0] in (AGG["MIN"], AGG["MAX"]):
            keep_r = np.ones(g, bool) if valr is None else valr[o_ref]
            assert np.array_equal(vr[o_ref][keep_r], vn[o_new][keep_r]), i
    res = [got[r * len(AGGS):(r + 1) * len(AGGS)] for r in range(2)]
    assert_parity([key], [(i32, AGGS), (f64, AGGS)], 0, gpu=([k_new], res))
    del keep, bufs
    return g


def test_no_alloc_matches_existing_entry_point_twice():
    g1 = _compare(200_000, 20_000, 1)
    g2 = _compare(30_000, 3_000, 2)  # a smaller input on the same scratch
    assert g1 > g2 > 0


def test_no_alloc_rejects_small_outputs_and_oversized_tables():
    from cudf_amd import _lib
    lib = _lib.lib()
    assert lib.mig_groupby_into_scratch_bytes() == 128 << 20
    key, i32, f64 = _inputs(50_000, 10_000, 3)
    dk, dvals = to_device(key), [to_device(i32), to_device(f64)]
    keys, reqs, keep = _requests(dk, dvals)
    st, g, _, _, bufs = _into(keys, reqs, 100)  # room for 100 groups only
    assert st == _lib.MIG_ERROR_OVERFLOW, lib.mig_last_error()
    assert b"rows" in lib.mig_last_error()
    # 12M distinct keys: the table alone would need 2^25 slots * 8 bytes
    n = 12_000_000
    import torch
    big = torch.arange(n, dtype=torch.int64, device="cuda")
    from cudf_amd import Column, TypeId
    bk = Column(TypeId.INT64, n, big.view(torch.uint8), None, 0)
    bv = Column(TypeId.INT64, n, big.view(torch.uint8), None, 0)
    keys2 = (_lib.ColumnView * 1)(bk.view())
    aggs = (_lib.Aggregation * 1)(_lib.Aggregation(AGG["SUM"], 1))
    reqs2 = (_lib.AggregationRequest * 1)()
    reqs2[0].values, reqs2[0].aggregations, reqs2[0].num_aggregations = bv.view(), aggs, 1
    out_k = (_lib.OutColumn * 1)()
    out_r = (_lib.OutColumn * 1)()
    g2 = ctypes.c_int32(-1)
    st = lib.mig_groupby_aggregate_into(keys2, 1, 0, reqs2, 1, out_k, out_r, ctypes.byref(g2), None)
    assert st == _lib.MIG_ERROR_OVERFLOW, lib.mig_last_error()
    assert b"MiB" in lib.mig_last_error()
    # and the entry point still works afterwards
    _compare(20_000, 500, 4)
    del keep, bufs

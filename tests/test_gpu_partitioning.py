"""GPU parity of hash partitioning, partitioning by map, concatenation, and
the partial-aggregation group-by built on them, against the oracle
(oracle/partition_oracle.py, oracle/groupby_oracle.c).

Cases follow cpp/tests/partitioning/hash_partition_test.cpp (mixed column
types, nullable inputs, more partitions than rows, large inputs, seeds,
identity hash, invalid indices, empty inputs) plus multi-pass partition
counts and a 20M-row size-independent check (permutation, stability and
offsets against torch's stable sort of the recomputed row hashes).
Bar: bit-exact everything for partitioning and concatenation; for the
partial-aggregation group-by, bit-exact keys/counts/integer sums/min/max,
f64 sums and means within 1e-12 of the per-group sum of |x|, var/std within
1e-10 relative (data with |mean| / std below 10, so the Σx² form of the
variance keeps that accuracy).
"""
import numpy as np
import pytest

from oracle import oracle as orc
from oracle import partition_oracle as po
from tests.gbutil import AGG, host_col, oracle_groupby, sort_order, to_device

pytestmark = pytest.mark.gpu


def _dev_table(cols):
    import cudf_amd as ca
    return ca.Table([to_device(c) for c in cols])


def _assert_col_equal(got, want: orc.HostColumn, what=""):
    vals, valid = got.to_numpy()
    if want.type_id == orc.STRING:
        exp = [want.value(i) if want.valid is None or want.valid[i] else None for i in range(want.size)]
        gv = [None if valid is not None and not valid[i] else bytes(vals[i]) for i in range(len(vals))]
        assert gv == exp, what
        return
    wv = want.data
    if want.valid is not None:
        assert valid is not None, what
        np.testing.assert_array_equal(valid, want.valid, err_msg=what)
        np.testing.assert_array_equal(vals[want.valid].view(np.uint8), wv[want.valid].view(np.uint8), err_msg=what)
    else:
        assert valid is None or valid.all(), what
        np.testing.assert_array_equal(vals.view(np.uint8), wv.view(np.uint8), err_msg=what)


def _check_hash_partition(cols, columns_to_hash, nparts, hash_function=po.HASH_MURMUR3, seed=0):
    from cudf_amd.partitioning import hash_partition
    order, offs = po.hash_partition(cols, columns_to_hash, nparts, hash_function, seed)
    out, got_offs = hash_partition(_dev_table(cols), columns_to_hash, nparts, hash_function, seed)
    assert got_offs == offs
    assert out.num_columns() == len(cols)
    assert out.num_rows() == len(order)
    for i, (g, c) in enumerate(zip(out.columns(), cols)):
        assert int(g.type_id) == c.type_id
        _assert_col_equal(g, po.gather(c, order), f"column {i}")
    return out, got_offs


def test_mixed_column_types():
    floats = host_col(np.arange(1, 9, dtype=np.float32))
    ints = host_col(np.arange(1, 9, dtype=np.int16))
    strs = orc.HostColumn.from_strings([b"a", b"bb", b"ccc", b"d", b"ee", b"fff", b"gg", b"h"])
    out1, o1 = _check_hash_partition([floats, ints, strs], [0, 2], 3)
    out2, o2 = _check_hash_partition([floats, ints, strs], [0, 2], 3)
    assert o1 == o2
    for a, b in zip(out1.columns(), out2.columns()):
        av, avalid = a.to_numpy()
        bv, bvalid = b.to_numpy()
        assert list(av) == list(bv)


def test_columns_to_hash_and_seed():
    to_hash = host_col(np.arange(1, 7, dtype=np.int32))
    _, oa = _check_hash_partition([to_hash, host_col(np.arange(7, 13, dtype=np.int32))], [0], 3)
    _, ob = _check_hash_partition([to_hash, host_col(np.arange(13, 19, dtype=np.int32))], [0], 3)
    assert oa == ob
    rng = np.random.default_rng(1)
    cols = [host_col(rng.normal(size=500)), host_col(rng.integers(0, 9, 500).astype(np.int16))]
    for seed in (0, 12345, 0xFFFFFFFF):
        _check_hash_partition(cols, [0, 1], 7, seed=seed)


@pytest.mark.parametrize("hash_function", [po.HASH_MURMUR3, po.HASH_IDENTITY])
@pytest.mark.parametrize("dtype", [np.int8, np.int16, np.int32, np.int64, np.uint8, np.uint16, np.uint32,
                                   np.uint64, np.float32, np.float64, np.bool_])
@pytest.mark.parametrize("shape", [(5, 10, 50), (10, 1000, 10), (10, 1000, 10, True)])
def test_fixed_width(hash_function, dtype, shape):
    ncols, rows, nparts = shape[:3]
    nulls = len(shape) > 3
    it = np.arange(rows)
    valid = (it % 4 != 0) if nulls else None
    cols = [host_col(it.astype(dtype), valid) for _ in range(ncols)]
    _check_hash_partition(cols, list(range(ncols)), nparts, hash_function)


@pytest.mark.parametrize("nparts", [6, 3000])
def test_many_columns(nparts):
    # the last counting pass moves up to 16 fixed-width columns; the rest are gathered
    rng = np.random.default_rng(nparts)
    n = 4000
    dts = [np.int8, np.int16, np.int32, np.int64, np.float32, np.float64]
    cols = [host_col(rng.integers(-100, 100, n).astype(dts[i % len(dts)]), (rng.random(n) > 0.3) if i % 5 == 0
                     else None) for i in range(20)]
    _check_hash_partition(cols, [0, 3, 19], nparts)


def test_nullable_without_nulls_and_strings_with_nulls():
    strs = orc.HostColumn.from_strings([b"a", b"bb", b"ccc", b"d"], np.ones(4, bool))
    out, _ = _check_hash_partition([strs], [0], 3)
    assert out.columns()[0].null_count == 0
    rng = np.random.default_rng(3)
    words = [bytes(rng.integers(97, 123, rng.integers(0, 20)).astype(np.uint8)) for _ in range(3000)]
    valid = rng.random(3000) > 0.2
    s = orc.HostColumn.from_strings(words, valid)
    v = host_col(rng.integers(-9, 9, 3000).astype(np.int64), rng.random(3000) > 0.5)
    _check_hash_partition([s, v], [0], 13)
    _check_hash_partition([v, s], [0, 1], 64)


@pytest.mark.parametrize("nparts", [1, 2, 2048, 2049, 5000, 1 << 20])
def test_partition_counts_and_passes(nparts):
    rng = np.random.default_rng(nparts)
    n = 200_000
    k = host_col(rng.integers(0, 1 << 40, n, dtype=np.int64))
    v = host_col(rng.normal(size=n).astype(np.float32), rng.random(n) > 0.1)
    _check_hash_partition([k, v], [0], nparts)


def test_sliced_inputs():
    import cudf_amd as ca
    from cudf_amd.partitioning import hash_partition
    rng = np.random.default_rng(9)
    n = 5000
    keys = rng.integers(0, 100, n).astype(np.int32)
    valid = rng.random(n) > 0.3
    strs = [b"s%d" % (x % 37) for x in keys]
    dk = ca.Column.from_numpy(keys, valid).slice(13, 4000)
    ds = ca.Column.from_strings(strs).slice(13, 4000)
    out, offs = hash_partition(ca.Table([dk, ds]), [0, 1], 9)
    hk = host_col(keys[13:4013], valid[13:4013])
    hs = orc.HostColumn.from_strings(strs[13:4013])
    order, want_offs = po.hash_partition([hk, hs], [0, 1], 9)
    assert offs == want_offs
    _assert_col_equal(out.columns()[0], po.gather(hk, order))
    _assert_col_equal(out.columns()[1], po.gather(hs, order))


def test_empty_and_errors():
    import cudf_amd as ca
    from cudf_amd import _lib
    from cudf_amd.partitioning import HashId, hash_partition
    floats = ca.Column.from_numpy(np.arange(1, 9, dtype=np.float32))
    ints = ca.Column.from_numpy(np.arange(1, 9, dtype=np.int16))
    strs = ca.Column.from_strings(["a", "bb", "ccc", "d", "ee", "fff", "gg", "h"])
    t = ca.Table([floats, ints, strs])
    with pytest.raises(IndexError):
        hash_partition(t, [-1], 3)
    with pytest.raises(IndexError):
        hash_partition(t, [3], 3)
    with pytest.raises(_lib.LogicError):
        hash_partition(t, [2], 3, HashId.HASH_IDENTITY)
    with pytest.raises(_lib.LogicError):
        hash_partition(t, [0], 3, 7)
    for args in (([0], 0), ([], 3)):
        out, offs = hash_partition(t, *args)
        assert out.num_columns() == 3 and out.num_rows() == 0 and offs == [0] * args[1]
        assert [int(c.type_id) for c in out.columns()] == [int(c.type_id) for c in t.columns()]
    empty = ca.Table([ca.Column.from_numpy(np.zeros(0, np.int32)), ca.Column.from_strings([])])
    out, offs = hash_partition(empty, [0], 3)
    assert out.num_rows() == 0 and offs == [0, 0, 0]


def test_partition_by_map():
    import cudf_amd as ca
    from cudf_amd import _lib
    from cudf_amd.partitioning import partition
    rng = np.random.default_rng(4)
    n = 10_000
    vals = host_col(rng.integers(-50, 50, n).astype(np.int64), rng.random(n) > 0.2)
    strs = orc.HostColumn.from_strings([b"r%d" % i for i in range(n)])
    for mdt, nparts in ((np.int32, 7), (np.int64, 3000), (np.uint8, 200), (np.int16, 1)):
        pmap = rng.integers(0, nparts, n).astype(mdt)
        order, offs = po.partition_by_map(pmap, nparts)
        out, got = partition(_dev_table([vals, strs]), ca.Column.from_numpy(pmap), nparts)
        assert got == offs
        _assert_col_equal(out.columns()[0], po.gather(vals, order))
        _assert_col_equal(out.columns()[1], po.gather(strs, order))
    with pytest.raises(_lib.LogicError):
        partition(_dev_table([vals]), ca.Column.from_numpy(np.full(n, 9, np.int32)), 5)
    with pytest.raises(_lib.LogicError):
        partition(_dev_table([vals]), ca.Column.from_numpy(np.zeros(n, np.float32)), 5)
    with pytest.raises(_lib.LogicError):
        partition(_dev_table([vals]), ca.Column.from_numpy(np.zeros(n - 1, np.int32)), 5)
    out, got = partition(_dev_table([vals]), ca.Column.from_numpy(np.zeros(n, np.int32)), 0)
    assert out.num_rows() == 0 and got == [0]


def test_concatenate():
    import cudf_amd as ca
    from cudf_amd import _lib
    from cudf_amd.partitioning import concatenate
    rng = np.random.default_rng(8)
    pieces_h, pieces_d = [], []
    for i, n in enumerate([0, 1, 31, 33, 64, 100, 1000, 7]):
        v = rng.integers(-1000, 1000, n + 5).astype(np.int64)
        valid = (rng.random(n + 5) > 0.3) if i % 2 else None
        s = [b"p%d_%d" % (i, j) * (j % 3) for j in range(n + 5)]
        dv = ca.Column.from_numpy(v, valid).slice(5, n)
        ds = ca.Column.from_strings(s, valid, large=(i == 3)).slice(5, n)
        f = ca.Column.from_numpy(v.astype(np.float32)).slice(5, n)
        pieces_d.append(ca.Table([dv, ds, f]))
        pieces_h.append([host_col(v[5:5 + n], None if valid is None else valid[5:5 + n]),
                         orc.HostColumn.from_strings(s[5:5 + n], None if valid is None else valid[5:5 + n]),
                         host_col(v[5:5 + n].astype(np.float32))])
    out = concatenate(pieces_d)
    assert out.num_rows() == sum(p.num_rows() for p in pieces_d)
    for c in range(3):
        want = po.concatenate([p[c] for p in pieces_h])
        _assert_col_equal(out.columns()[c], want, f"column {c}")
    assert out.columns()[2].mask is None
    with pytest.raises(_lib.DataTypeError):
        concatenate([ca.Table([ca.Column.from_numpy(np.zeros(3, np.int32))]),
                     ca.Table([ca.Column.from_numpy(np.zeros(3, np.int64))])])


def test_large_partition_properties():
    """20M rows, 1000 and 100000 partitions: the output is the stable
    regrouping of the input by murmur row hash, checked on the device against
    torch's stable sort of the row hashes the library computes separately."""
    import torch
    import cudf_amd as ca
    from cudf_amd.partitioning import hash_partition
    n = 20_000_000
    g = torch.Generator(device="cuda").manual_seed(11)
    keys = torch.randint(0, 1 << 62, (n,), device="cuda", dtype=torch.int64, generator=g)
    idx = torch.arange(n, device="cuda", dtype=torch.int32)
    kc = ca.Column(ca.TypeId.INT64, n, keys.view(torch.uint8))
    ic = ca.Column(ca.TypeId.INT32, n, idx.view(torch.uint8))
    h = ca.murmurhash3_x86_32(ca.Table([kc]))
    hv = h.data.view(torch.int32).to(torch.int64) & 0xFFFFFFFF
    for nparts in (1000, 100_000):
        print(f"large hash_partition n={n} nparts={nparts}", flush=True)
        out, offs = hash_partition(ca.Table([kc, ic]), [0], nparts)
        pid = hv % nparts
        want = torch.sort(pid, stable=True).indices.to(torch.int32)
        got_idx = out.columns()[1].data.view(torch.int32)
        assert torch.equal(got_idx, want)
        assert torch.equal(out.columns()[0].data.view(torch.int64), keys[want.to(torch.int64)])
        counts = torch.bincount(pid, minlength=nparts)
        want_offs = torch.cumsum(counts, 0) - counts
        assert offs == want_offs.cpu().tolist()


# ---- partial aggregation, combine, finalize ----------------------------------
def _oracle_expected(keys, values, aggs, policy):
    reqs, idx = [], []
    kinds = {"count": (AGG["COUNT_VALID"], 1), "sum": (AGG["SUM"], 1), "min": (AGG["MIN"], 1),
             "max": (AGG["MAX"], 1), "mean": (AGG["MEAN"], 1), "var": (AGG["VARIANCE"], 1),
             "std": (AGG["STD"], 1)}
    for col in sorted(aggs):
        reqs.append((values[col], [kinds[a] for a in aggs[col]]))
        idx.append(col)
    return oracle_groupby(keys, reqs, policy), idx


def _check_dask_groupby(nparts, split_out, dropna, seed=0, n_per=20_000):
    import cudf_amd as ca
    from cudf_amd.dask_groupby import groupby_agg
    rng = np.random.default_rng(seed)
    aggs = {"a": ["count", "sum", "min", "max", "mean", "var", "std"], "b": ["sum", "mean", "var"],
            "c": ["min", "max", "count"]}
    hk, ha, hb, hc, parts = [], [], [], [], []
    for p in range(nparts):
        k = rng.integers(0, 3000, n_per).astype(np.int32)
        kv = rng.random(n_per) > 0.05
        a = rng.integers(-1000, 1000, n_per).astype(np.int64)
        av = rng.random(n_per) > 0.1
        b = rng.normal(3.0, 1.0, n_per)
        c = rng.integers(-2**31, 2**31, n_per, dtype=np.int64).astype(np.int32)
        hk.append((k, kv)), ha.append((a, av)), hb.append((b, None)), hc.append((c, None))
        parts.append((ca.Table([ca.Column.from_numpy(k, kv)]),
                      {"a": ca.Column.from_numpy(a, av), "b": ca.Column.from_numpy(b),
                       "c": ca.Column.from_numpy(c)}))
    outs = groupby_agg(parts, aggs, dropna=dropna, split_every=3, split_out=split_out)
    assert len(outs) == max(1, split_out)
    cat = lambda xs: host_col(np.concatenate([x for x, _ in xs]),
                              None if xs[0][1] is None else np.concatenate([v for _, v in xs]))
    keys, values = [cat(hk)], {"a": cat(ha), "b": cat(hb), "c": cat(hc)}
    policy = 0 if dropna else 1
    (o_keys, o_res), cols = _oracle_expected(keys, values, aggs, policy)
    # outputs hold disjoint keys; their union is the single-call result
    all_keys, all_valid, got = [], [], {}
    for ukeys, res in outs:
        kv, kvalid = ukeys.columns()[0].to_numpy()
        all_keys.append(kv)
        all_valid.append(np.ones(len(kv), bool) if kvalid is None else kvalid)
        for key, col in res.items():
            got.setdefault(key, []).append(col.to_numpy())
    gk = [(np.concatenate(all_keys), np.concatenate(all_valid))]
    og, oo = sort_order(gk), sort_order(o_keys)
    assert len(og) == len(oo)
    gkv, gkvalid = gk[0]
    okv, okvalid = o_keys[0]
    okvalid = np.ones(len(okv), bool) if okvalid is None else okvalid
    np.testing.assert_array_equal(gkvalid[og], okvalid[oo])
    np.testing.assert_array_equal(gkv[og][gkvalid[og]], np.asarray(okv)[oo][okvalid[oo]])
    for r, col in enumerate(cols):
        for a_i, a in enumerate(aggs[col]):
            parts_ = got[(col, a)]
            gv = np.concatenate([np.asarray(v, dtype=np.float64) for v, _ in parts_])[og]
            gvalid = np.concatenate([np.ones(len(v), bool) if m is None else m for v, m in parts_])[og]
            ov, ovalid = o_res[r][a_i]
            ov = np.asarray(ov, dtype=np.float64)[oo]
            ovalid = ovalid[oo]
            np.testing.assert_array_equal(gvalid, ovalid, err_msg=f"{col}.{a} validity")
            gv, ov = gv[gvalid], ov[ovalid]
            if a in ("count", "min", "max") or (a == "sum" and col in ("a", "c")):
                np.testing.assert_array_equal(gv, ov, err_msg=f"{col}.{a}")
            elif a in ("sum", "mean"):
                np.testing.assert_allclose(gv, ov, rtol=1e-12, atol=0, err_msg=f"{col}.{a}")
            else:
                np.testing.assert_allclose(gv, ov, rtol=1e-10, atol=0, err_msg=f"{col}.{a}")


@pytest.mark.parametrize("nparts,split_out,dropna", [(1, 1, True), (7, 1, True), (7, 4, True), (10, 3, False)])
def test_dask_style_groupby(nparts, split_out, dropna):
    _check_dask_groupby(nparts, split_out, dropna, seed=nparts * 10 + split_out)


def test_dask_style_groupby_errors():
    import cudf_amd as ca
    from cudf_amd.dask_groupby import groupby_agg
    k = ca.Table([ca.Column.from_numpy(np.arange(5, dtype=np.int32))])
    with pytest.raises(ValueError):
        groupby_agg([(k, {"a": ca.Column.from_numpy(np.arange(5.0))})], {"a": ["median"]})

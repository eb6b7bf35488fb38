"""GPU parity tests: the gfx950 path, called through the C ABI, against the
CPU oracle on the same seeded inputs (bit-exact for keys, counts, integer
sums, min and max; floating-point tolerances as stated in tests/gbutil.py)."""
import json
import os

import numpy as np
import pytest

from oracle import oracle as orc
from tests.gbutil import (AGG, C2_AGGS, assert_parity, c1_data, c2_data, c3_data, gpu_groupby,
                          host_col, sort_order)
from tests.test_oracle import KATS, NP_TYPES, _kat_value_types, _run_kat

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _loaded_library_is_native():
    from cudf_amd import _lib
    assert _lib.lib() is not None


# ------------------------------------------------------------------ hashing --
def test_hash_docstring_vectors():
    import cudf_amd as ca
    with open(os.path.join(GOLDEN, "hash_vectors.json")) as f:
        cases = json.load(f)["docstring"]
    for case in cases:
        if case["seed"] != 0:
            continue
        cols = [ca.Column.from_numpy(np.array(c["values"], dtype=NP_TYPES[c["type"]]))
                for c in case["columns"]]
        h, _ = ca.murmurhash3_x86_32(ca.Table(cols)).to_numpy()
        assert list(h.view(np.int32)) == case["expect_int32"], case["source"]


@pytest.mark.parametrize("dtype", [np.int8, np.int16, np.int32, np.int64, np.uint8, np.uint16,
                                   np.uint32, np.uint64, np.float32, np.float64, np.bool_])
def test_hash_matches_oracle_every_type(dtype):
    import cudf_amd as ca
    rng = np.random.default_rng(7)
    n = 5000
    if dtype == np.bool_:
        a = rng.random(n) > 0.5
    elif np.dtype(dtype).kind == "f":
        a = rng.standard_normal(n).astype(dtype)
        a[:10] = [0.0, -0.0, np.nan, -np.nan, np.inf, -np.inf, 1.0, -1.0, 1e-30, 3.5]
    else:
        info = np.iinfo(dtype)
        a = rng.integers(info.min, info.max, n, dtype=dtype, endpoint=True)
    valid = rng.random(n) > 0.1
    hc = host_col(a, valid)
    for seed in (0, 1234):
        h, _ = ca.murmurhash3_x86_32(ca.Table([ca.Column.from_numpy(a, valid)]), seed).to_numpy()
        assert np.array_equal(h, orc.hash_rows([hc], seed=seed))


def test_hash_strings_and_multicolumn():
    import cudf_amd as ca
    rng = np.random.default_rng(8)
    n = 3000
    strs = [bytes(rng.integers(0, 256, int(rng.integers(0, 40)), dtype=np.uint8)) for _ in range(n)]
    sv = rng.random(n) > 0.1
    ints = rng.integers(-1000, 1000, n, dtype=np.int32)
    iv = rng.random(n) > 0.2
    s_host = orc.HostColumn.from_strings(strs, sv)
    i_host = host_col(ints, iv)
    dev = ca.Table([ca.Column.from_strings(strs, sv), ca.Column.from_numpy(ints, iv)])
    h, _ = ca.murmurhash3_x86_32(dev).to_numpy()
    assert np.array_equal(h, orc.hash_rows([s_host, i_host]))
    # large (int64) offsets give the same hashes
    dev64 = ca.Table([ca.Column.from_strings(strs, sv, large=True), ca.Column.from_numpy(ints, iv)])
    h64, _ = ca.murmurhash3_x86_32(dev64).to_numpy()
    assert np.array_equal(h64, h)


# ----------------------------------------------------- known-answer tests --
@pytest.mark.parametrize("case", KATS["cases"], ids=[c["name"] for c in KATS["cases"]])
def test_known_answers(case):
    for vt in _kat_value_types(case):
        _run_kat(case, vt, gpu_groupby)


def test_string_keys_known_answer():
    case = KATS["string_key_cases"][0]
    keys = orc.HostColumn.from_strings([s.encode() for s in case["keys"]])
    vals = host_col(np.array(case["vals"], np.int32))
    kcols, res = gpu_groupby([keys], [(vals, [(AGG["SUM"], 1)])])
    got = {bytes(kcols[0][0][i]): int(res[0][0][0][i]) for i in range(len(kcols[0][0]))}
    assert got == {s.encode(): v for s, v in zip(case["expect_keys"], case["expect_vals"])}


def test_empty_input():
    keys = host_col(np.zeros(0, np.int32))
    vals = host_col(np.zeros(0, np.float64))
    kcols, res = gpu_groupby([keys], [(vals, [(AGG["SUM"], 1), (AGG["VARIANCE"], 1)])])
    assert len(kcols[0][0]) == 0 and len(res[0][0][0]) == 0 and res[0][1][0].dtype == np.float64


# ------------------------------------------------------ randomized parity --
ALL_NUMERIC_AGGS = [(AGG["SUM"], 1), (AGG["MIN"], 1), (AGG["MAX"], 1), (AGG["COUNT_VALID"], 1),
                    (AGG["COUNT_ALL"], 1), (AGG["MEAN"], 1), (AGG["VARIANCE"], 1), (AGG["STD"], 2),
                    (AGG["SUM_OF_SQUARES"], 1)]


@pytest.mark.parametrize("key_dtype", [np.int8, np.int16, np.int32, np.int64, np.uint32, np.uint64,
                                       np.float32, np.float64, np.bool_])
@pytest.mark.parametrize("policy", [0, 1])
def test_key_types(key_dtype, policy):
    rng = np.random.default_rng(11)
    n = 20_000
    if key_dtype == np.bool_:
        k = rng.random(n) > 0.5
    elif np.dtype(key_dtype).kind == "f":
        k = rng.integers(-50, 50, n).astype(key_dtype) / 4
        k[rng.random(n) < 0.02] = np.nan
        k[rng.random(n) < 0.02] = -0.0
    else:
        k = rng.integers(0, min(200, np.iinfo(key_dtype).max), n).astype(key_dtype)
    kvalid = rng.random(n) > 0.05
    v = rng.integers(-1000, 1000, n).astype(np.int32)
    vvalid = rng.random(n) > 0.1
    assert_parity([host_col(k, kvalid)], [(host_col(v, vvalid), ALL_NUMERIC_AGGS)], policy)


@pytest.mark.parametrize("val_dtype", [np.int8, np.int16, np.int32, np.int64, np.uint8, np.uint16,
                                       np.uint32, np.uint64, np.float32, np.float64, np.bool_])
def test_value_types(val_dtype):
    rng = np.random.default_rng(12)
    n = 30_000
    k = rng.integers(0, 500, n).astype(np.int64)
    if val_dtype == np.bool_:
        v = rng.random(n) > 0.5
    elif np.dtype(val_dtype).kind == "f":
        v = (rng.standard_normal(n) * 10).astype(val_dtype)
    else:
        info = np.iinfo(val_dtype)
        lo, hi = max(info.min, -10**6), min(info.max, 10**6)
        v = rng.integers(lo, hi, n, endpoint=True).astype(val_dtype)
    vvalid = rng.random(n) > 0.1
    aggs = ALL_NUMERIC_AGGS
    assert_parity([host_col(k)], [(host_col(v, vvalid), aggs)])


def test_extreme_integer_values():
    rng = np.random.default_rng(13)
    n = 10_000
    k = rng.integers(0, 50, n).astype(np.int32)
    v = rng.choice(np.array([np.iinfo(np.int64).min, np.iinfo(np.int64).max, -1, 0, 1], np.int64), n)
    u = rng.choice(np.array([0, np.iinfo(np.uint64).max, 2**63, 1], np.uint64), n)
    aggs = [(AGG["SUM"], 1), (AGG["MIN"], 1), (AGG["MAX"], 1), (AGG["SUM_OF_SQUARES"], 1)]
    assert_parity([host_col(k)], [(host_col(v), aggs), (host_col(u), aggs)])


def test_sentinel_valued_int64_key():
    """An int64 key equal to the inline table's empty marker still forms its
    own group."""
    marker = np.array([0xFFF1A5A5A5A5A5A5], dtype=np.uint64).view(np.int64)[0]
    k = np.array([marker, 1, marker, 2, 1, marker], dtype=np.int64)
    v = np.arange(6, dtype=np.int64)
    ng = assert_parity([host_col(k)], [(host_col(v), [(AGG["SUM"], 1), (AGG["COUNT_ALL"], 1)])])
    assert ng == 3


def test_multicolumn_and_string_keys():
    rng = np.random.default_rng(14)
    n = 40_000
    k1 = rng.integers(0, 20, n).astype(np.int32)
    k1v = rng.random(n) > 0.03
    pool = [bytes(rng.integers(97, 123, int(rng.integers(0, 12)), dtype=np.uint8)) for _ in range(60)]
    k2 = orc.HostColumn.from_strings([pool[i] for i in rng.integers(0, 60, n)], rng.random(n) > 0.03)
    k3 = (rng.integers(0, 3, n) * 0.5).astype(np.float64)
    v = rng.standard_normal(n)
    for policy in (0, 1):
        assert_parity([host_col(k1, k1v), k2, host_col(k3)],
                      [(host_col(v, rng.random(n) > 0.1), ALL_NUMERIC_AGGS)], policy)


def test_duplicate_requests_and_columns():
    rng = np.random.default_rng(15)
    n = 5000
    k = host_col(rng.integers(0, 30, n).astype(np.int32))
    v = host_col(rng.integers(0, 100, n).astype(np.int32))
    aggs = [(AGG["SUM"], 1), (AGG["SUM"], 1), (AGG["MEAN"], 1), (AGG["COUNT_VALID"], 1)]
    assert_parity([k], [(v, aggs), (v, aggs[:2])])


def test_sliced_columns():
    """Column views with a row offset (cudf slicing): masks are read from the
    view's bit offset."""
    import cudf_amd as ca
    from cudf_amd.aggregation import Aggregation
    rng = np.random.default_rng(16)
    n = 10_000
    k = rng.integers(0, 40, n).astype(np.int32)
    kv = rng.random(n) > 0.1
    v = rng.integers(-100, 100, n).astype(np.int64)
    vv = rng.random(n) > 0.1
    off, size = 37, 7000
    dk = ca.Column.from_numpy(k, kv).slice(off, size)
    dv = ca.Column.from_numpy(v, vv).slice(off, size)
    out_k, out_r = ca.GroupBy(ca.Table([dk])).aggregate(
        [ca.GroupByRequest(dv, [Aggregation(AGG["SUM"]), Aggregation(AGG["COUNT_VALID"])])])
    gpu = ([c.to_numpy() for c in out_k.columns()], [[c.to_numpy() for c in t.columns()] for t in out_r])
    hk = host_col(k[off:off + size], kv[off:off + size])
    hv = host_col(v[off:off + size], vv[off:off + size])
    assert_parity([hk], [(hv, [(AGG["SUM"], 1), (AGG["COUNT_VALID"], 1)])], gpu=gpu)


def test_count_of_string_values():
    rng = np.random.default_rng(17)
    n = 2000
    k = host_col(rng.integers(0, 10, n).astype(np.int32))
    s = orc.HostColumn.from_strings([b"x" * int(i) for i in rng.integers(0, 5, n)], rng.random(n) > 0.3)
    assert_parity([k], [(s, [(AGG["COUNT_VALID"], 1), (AGG["COUNT_ALL"], 1)])])


def test_errors_are_raised():
    import cudf_amd as ca
    from cudf_amd import _lib
    from cudf_amd.aggregation import Aggregation
    k = ca.Column.from_numpy(np.arange(10, dtype=np.int32))
    v = ca.Column.from_numpy(np.arange(5, dtype=np.int32))
    with pytest.raises(_lib.LogicError):
        ca.GroupBy(ca.Table([k])).aggregate([ca.GroupByRequest(v, [Aggregation(AGG["SUM"])])])
    s = ca.Column.from_strings(["a"] * 10)
    with pytest.raises(_lib.DataTypeError):
        ca.GroupBy(ca.Table([k])).aggregate([ca.GroupByRequest(s, [Aggregation(AGG["SUM"])])])


# ------------------------------------------------------ workload configs --
def test_c1_full_size():
    key, v = c1_data()
    ng = assert_parity([key], [(v, [(AGG["SUM"], 1), (AGG["COUNT_VALID"], 1), (AGG["MIN"], 1),
                                    (AGG["MAX"], 1)])])
    assert ng == 1000


def test_c2_shape_reduced():
    key, f64, i32 = c2_data(n=2_000_000, card=200_000, seed=5)
    assert_parity([key], [(f64, C2_AGGS), (i32, C2_AGGS)])


def test_c2_high_cardinality_reduced():
    # ~10 rows per group, like the 100M-row / 10M-key workload
    key, f64, i32 = c2_data(n=1_000_000, card=100_000, seed=6)
    assert_parity([key], [(f64, C2_AGGS), (i32, C2_AGGS)])


def test_c3_shape_reduced():
    key1, key2, key3, f32, i64 = c3_data(n=300_000, seed=9, pool_size=2000)
    aggs = [(AGG["SUM"], 1), (AGG["COUNT_VALID"], 1), (AGG["MAX"], 1)]
    assert_parity([key1, key2, key3], [(f32, aggs), (i64, aggs)])


def test_c2_full_size_properties():
    """100M rows, 10M keys: properties independent of size (exact totals of
    counts and integer sums, global min/max, group count)."""
    key, f64, i32 = c2_data()
    kcols, res = gpu_groupby([key], [(f64, C2_AGGS), (i32, C2_AGGS)])
    G = len(kcols[0][0])
    kvalid = key.valid
    # every valid key value appears exactly once
    assert len(np.unique(kcols[0][0])) == G
    assert G == len(np.unique(key.data[kvalid]))
    both = kvalid & i32.valid
    cnt, cval = res[1][2]
    assert int(cnt.sum()) == int(both.sum())
    s, sval = res[1][0]
    assert int(s[sval].sum()) == int(i32.data[both].astype(np.int64).sum())
    mn, mnv = res[1][3]
    mx, mxv = res[1][4]
    assert int(mn[mnv].min()) == int(i32.data[both].min())
    assert int(mx[mxv].max()) == int(i32.data[both].max())
    fboth = kvalid & f64.valid
    fcnt, _ = res[0][2]
    assert int(fcnt.sum()) == int(fboth.sum())
    fs, fsv = res[0][0]
    assert abs(float(fs[fsv].sum()) - float(f64.data[fboth].sum())) < 1e-6 * float(np.abs(f64.data[fboth]).sum())


# ------------------------------------- low cardinality: LDS partial tables --
def _last_stats():
    from cudf_amd import _lib
    return _lib.last_stats()


def test_lds_preaggregation_is_used_for_low_cardinality():
    key, v = c1_data(n=2_000_000, seed=3)
    assert_parity([key], [(v, [(AGG["SUM"], 1), (AGG["COUNT_VALID"], 1), (AGG["MIN"], 1),
                               (AGG["MAX"], 1), (AGG["MEAN"], 1)])])
    assert _last_stats()["lds_preagg"] == 1.0


@pytest.mark.parametrize("policy", [0, 1])
def test_lds_preaggregation_variance_nulls_sentinel(policy):
    """50 int64 keys incl. the table's empty marker, null keys, nullable
    float64 and int32 values, every C2 aggregation (LDS first and second pass)."""
    rng = np.random.default_rng(61 + policy)
    n = 400_000
    sentinel = np.array([0xFFF1A5A5A5A5A5A5], dtype=np.uint64).view(np.int64)[0]
    k = rng.integers(-25, 25, n).astype(np.int64)
    k[rng.random(n) < 0.05] = sentinel
    key = host_col(k, rng.random(n) >= 0.03)
    f = host_col(rng.standard_normal(n) * 1e3 + 5e3, rng.random(n) >= 0.05)
    i = host_col(rng.integers(0, 10_000, n).astype(np.int32), rng.random(n) >= 0.05)
    assert_parity([key], [(f, C2_AGGS), (i, C2_AGGS)], policy)
    assert _last_stats()["lds_preagg"] == 1.0


def test_lds_preaggregation_table_full_falls_back_to_global():
    """The sampled rows all hold one key while the input has ~20k: the LDS
    tables fill and the remaining keys go straight to the global table.  The
    sampled rows are the estimator's: sample index -> row through a 4-round
    Feistel permutation of [0, n) (round function MurmurHash3's fmix32)."""
    n, ns = 400_000, 65536
    k = (np.arange(n, dtype=np.int64) % 20_000) + 1
    M = 0xFFFFFFFF

    def fmix(h):
        h ^= h >> 16
        h = (h * 0x85EBCA6B) & M
        h ^= h >> 13
        h = (h * 0xC2B2AE35) & M
        return h ^ (h >> 16)

    bits = 2
    while (1 << bits) < n:
        bits += 2
    half = bits // 2
    mask = (1 << half) - 1

    def row_of(i):
        x = i
        while True:
            left, r = x >> half, x & mask
            for rnd in range(4):
                f = fmix((r * 0x9E3779B9 + rnd * 0x85EBCA6B + (r >> 32)) & M)
                left, r = r, left ^ (f & mask)
            x = (left << half) | r
            if x < n:
                return x

    rows = np.array([row_of(i) for i in range(ns)], dtype=np.int64)
    assert len(np.unique(rows)) == ns  # a permutation: no row twice
    k[rows] = 0
    rng = np.random.default_rng(67)
    v = host_col(rng.integers(-100, 100, n).astype(np.int64))
    f = host_col(rng.standard_normal(n))
    ng = assert_parity([host_col(k)], [(v, [(AGG["SUM"], 1), (AGG["MIN"], 1), (AGG["COUNT_ALL"], 1)]),
                                       (f, [(AGG["SUM"], 1), (AGG["VARIANCE"], 1)])])
    assert ng == len(np.unique(k))
    st = _last_stats()
    assert st["lds_preagg"] == 1.0 and st["est_groups"] < 100


# ---------------------------------------------------------- ARGMIN / ARGMAX --
@pytest.mark.parametrize("val_dtype", [np.int8, np.int32, np.int64, np.uint16, np.uint64, np.float32, np.float64])
@pytest.mark.parametrize("card", [7, 5_000, 150_000])
def test_argmin_argmax(val_dtype, card):
    """Ties (small value range) take the first row; NaN only wins in an all-NaN
    group; nulls skipped; mixed with other aggregations on the same column."""
    rng = np.random.default_rng(card + np.dtype(val_dtype).itemsize)
    n = 300_000
    k = rng.integers(0, card, n).astype(np.int64)
    kvalid = rng.random(n) > 0.02
    if np.dtype(val_dtype).kind == "f":
        v = rng.integers(-20, 20, n).astype(val_dtype)
        v[rng.random(n) < 0.05] = np.nan
    else:
        info = np.iinfo(val_dtype)
        v = rng.integers(max(info.min, -20), min(info.max, 40), n).astype(val_dtype)
    vvalid = rng.random(n) > 0.1
    aggs = [(AGG["ARGMIN"], 1), (AGG["ARGMAX"], 1), (AGG["MIN"], 1), (AGG["COUNT_VALID"], 1)]
    if np.dtype(val_dtype).kind != "f":
        aggs.append((AGG["SUM"], 1))  # float sums of groups with NaN are NaN on both sides
    for policy in (0, 1):
        assert_parity([host_col(k, kvalid)], [(host_col(v, vvalid), aggs)], policy)
    # no nulls in the values: no result mask
    g_keys, g_res = gpu_groupby([host_col(k)], [(host_col(v), [(AGG["ARGMAX"], 1)])])
    assert g_res[0][0][1] is None or g_res[0][0][1].all()


def test_argmin_argmax_all_nan_groups_and_errors():
    from cudf_amd import _lib
    k = np.repeat(np.arange(50, dtype=np.int32), 20)
    v = np.full(len(k), np.nan)
    v[k % 2 == 0] = np.arange((k % 2 == 0).sum(), dtype=np.float64) % 7
    assert_parity([host_col(k)], [(host_col(v), [(AGG["ARGMIN"], 1), (AGG["ARGMAX"], 1)])])
    s = orc.HostColumn.from_strings([b"x%d" % i for i in range(len(k))])
    with pytest.raises(_lib.DataTypeError):
        gpu_groupby([host_col(k)], [(s, [(AGG["SUM"], 1)])])
    # string ARGMIN / ARGMAX on the pre-sorted path: the oracle's rows
    assert_parity([host_col(np.sort(k))], [(s, [(AGG["ARGMIN"], 1), (AGG["ARGMAX"], 1)])],
                  gpu=gpu_groupby([host_col(np.sort(k))], [(s, [(AGG["ARGMIN"], 1), (AGG["ARGMAX"], 1)])],
                                  keys_sorted=True))


# ------------------------------------------------------------------ PRODUCT --
@pytest.mark.parametrize("val_dtype", [np.int8, np.int32, np.int64, np.uint32, np.float32, np.float64])
def test_product(val_dtype):
    """Groups of ~40 values: integers in [-3, 3] (int64 products, wrapping
    exactly as the oracle), floats in [0.9, 1.1]."""
    rng = np.random.default_rng(np.dtype(val_dtype).itemsize * 3)
    n = 200_000
    k = rng.integers(0, 5_000, n).astype(np.int32)
    if np.dtype(val_dtype).kind == "f":
        v = (0.9 + 0.2 * rng.random(n)).astype(val_dtype)
    elif np.dtype(val_dtype).kind == "u":
        v = rng.integers(0, 4, n).astype(val_dtype)
    else:
        v = rng.integers(-3, 4, n).astype(val_dtype)
    vvalid = rng.random(n) > 0.1
    for policy in (0, 1):
        assert_parity([host_col(k, rng.random(n) > 0.02)],
                      [(host_col(v, vvalid), [(AGG["PRODUCT"], 1), (AGG["SUM"], 1), (AGG["ARGMAX"], 1)])], policy)
    assert_parity([host_col(k)], [(host_col(v), [(AGG["PRODUCT"], 1)])])


# ------------------------------------------------- string MIN / MAX / ARG* --
@pytest.mark.parametrize("case", KATS["string_value_arg_cases"], ids=lambda c: c["name"])
def test_string_value_argmin_argmax_known_answers(case):
    keys = host_col(np.array(case["keys"], np.int32))
    vals = orc.HostColumn.from_strings([s.encode() for s in case["vals"]])
    kcols, res = gpu_groupby([keys], [(vals, [(AGG[case["agg"]], 1)])])
    got = {int(kcols[0][0][i]): int(res[0][0][0][i]) for i in range(len(kcols[0][0]))}
    assert got == dict(zip(case["expect_keys"], case["expect_vals"]))


@pytest.mark.parametrize("policy", [0, 1])
def test_string_min_max(policy):
    """MIN / MAX of string values are the strings at the ARGMIN / ARGMAX rows
    (hash/groupby.cu:113-181); ARGMIN / ARGMAX match the oracle row for row."""
    rng = np.random.default_rng(41 + policy)
    n = 60_000
    pool = [bytes(rng.integers(97, 100, int(rng.integers(0, 6)), dtype=np.uint8)) for _ in range(300)]
    k = rng.integers(0, 3_000, n).astype(np.int32)
    kvalid = rng.random(n) > 0.03
    strs = [pool[i] for i in rng.integers(0, len(pool), n)]
    svalid = rng.random(n) > 0.2
    vals = orc.HostColumn.from_strings(strs, svalid)
    aggs = [(AGG["MIN"], 1), (AGG["MAX"], 1), (AGG["ARGMIN"], 1), (AGG["ARGMAX"], 1), (AGG["COUNT_VALID"], 1)]
    from tests.gbutil import oracle_groupby
    g_keys, g_res = gpu_groupby([host_col(k, kvalid)], [(vals, aggs)], policy)
    o_keys, o_res = oracle_groupby([host_col(k, kvalid)], [(vals, [(AGG["ARGMIN"], 1), (AGG["ARGMAX"], 1),
                                                                   (AGG["COUNT_VALID"], 1)])], policy)
    og, oo = sort_order(g_keys), sort_order(o_keys)
    assert len(og) == len(oo)
    for j, (gi, oi) in enumerate(zip(og, oo)):
        for a_str, a_arg, o_a in ((0, 2, 0), (1, 3, 1)):
            o_row, o_ok = o_res[0][o_a][0][oi], o_res[0][o_a][1][oi]
            gs, gsv = g_res[0][a_str]
            ga, gav = g_res[0][a_arg]
            gsv = True if gsv is None else bool(gsv[gi])
            gav = True if gav is None else bool(gav[gi])
            assert gsv == bool(o_ok) and gav == bool(o_ok), (j, a_str)
            if o_ok:
                assert int(ga[gi]) == int(o_row), (j, a_arg)
                assert bytes(gs[gi]) == strs[int(o_row)], (j, a_str)
        assert int(g_res[0][4][0][gi]) == int(o_res[0][2][0][oi])


def test_argmin_argmax_large_properties():
    """20M rows, 2M groups, on the device-wide table: the ARGMIN / ARGMAX row
    holds the group's MIN / MAX and the group's key, and is the first such row."""
    import torch
    import cudf_amd as ca
    from cudf_amd import GroupBy, GroupByRequest
    from cudf_amd.aggregation import Aggregation, Kind
    n = 20_000_000
    g = torch.Generator(device="cuda").manual_seed(17)
    key = torch.randint(0, 2_000_000, (n,), device="cuda", dtype=torch.int64, generator=g)
    val = torch.randint(-50, 50, (n,), device="cuda", dtype=torch.int32, generator=g)
    kc = ca.Column(ca.TypeId.INT64, n, key.view(torch.uint8))
    vc = ca.Column(ca.TypeId.INT32, n, val.view(torch.uint8))
    aggs = [Aggregation(Kind.ARGMIN), Aggregation(Kind.ARGMAX), Aggregation(Kind.MIN), Aggregation(Kind.MAX)]
    out_keys, (res,) = GroupBy(ca.Table([kc])).aggregate([GroupByRequest(vc, aggs)])
    G = out_keys.num_rows()
    uk = out_keys.columns()[0].data.view(torch.int64)[:G]
    amin = res.columns()[0].data.view(torch.int32)[:G].to(torch.int64)
    amax = res.columns()[1].data.view(torch.int32)[:G].to(torch.int64)
    mn = res.columns()[2].data.view(torch.int32)[:G]
    mx = res.columns()[3].data.view(torch.int32)[:G]
    assert G == torch.unique(key).numel()
    assert torch.equal(key[amin], uk) and torch.equal(key[amax], uk)
    assert torch.equal(val[amin], mn) and torch.equal(val[amax], mx)
    # first row attaining the extreme: no earlier row of the group holds it
    order = torch.argsort(uk)
    pos = torch.searchsorted(uk[order], key)
    grp = order[pos]  # output group of every row
    first_min = torch.full((G,), n, dtype=torch.int64, device="cuda")
    rows = torch.arange(n, device="cuda")
    hit = val == mn[grp]
    first_min.scatter_reduce_(0, grp[hit], rows[hit], "amin")
    assert torch.equal(first_min, amin)


@pytest.mark.parametrize("strategy", ["auto", "global", "partitioned"])
def test_zipf_single_key_every_strategy(strategy):
    """Hot keys: an int64 key Zipf(1.1) over 1M values at 3M rows (about
    290K groups, the hottest key on ~9% of rows), int64 and float64 values
    with nulls; sum, count, min, max, mean, var under every strategy.  The
    sample sees mostly the hot keys, so the estimate is low: a table or bin
    layout that overflows must still give the exact result, and the
    wave-combined updates meet many equal keys per wave."""
    from cudf_amd import _lib
    rng = np.random.default_rng(1101)
    n = 3_000_000
    w = np.arange(1, 1_000_001, dtype=np.float64) ** -1.1
    k = (np.searchsorted(np.cumsum(w) / w.sum(), rng.random(n)) + 1).astype(np.int64)
    vi = rng.integers(-10**6, 10**6, n).astype(np.int64)
    vf = rng.standard_normal(n)
    aggs = [(AGG["SUM"], 1), (AGG["COUNT_VALID"], 1), (AGG["MIN"], 1), (AGG["MAX"], 1), (AGG["MEAN"], 1),
            (AGG["VARIANCE"], 1)]
    _lib.set_strategy({"auto": _lib.STRATEGY_AUTO, "global": _lib.STRATEGY_GLOBAL,
                       "partitioned": _lib.STRATEGY_PARTITIONED}[strategy])
    try:
        ng = assert_parity([host_col(k)], [(host_col(vi, rng.random(n) > 0.05), aggs),
                                           (host_col(vf, rng.random(n) > 0.05), aggs)])
    finally:
        _lib.set_strategy(_lib.STRATEGY_AUTO)
    assert ng == len(np.unique(k))

"""Parity of the radix-partitioned (LDS-bin) aggregation path, forced on
through mig_set_strategy, against the CPU oracle: one- and two-level
partitions, every key and value storage width, null keys under both
policies, the empty-marker key value, and the fallback to the global table
when the sampled group estimate is far too low."""
import numpy as np
import pytest

from tests.gbutil import AGG, C2_AGGS, assert_parity, c2_data, gpu_groupby, host_col
from tests.test_gpu_groupby import ALL_NUMERIC_AGGS

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def partitioned():
    from cudf_amd import _lib
    _lib.set_strategy(_lib.STRATEGY_PARTITIONED)
    yield
    _lib.set_strategy(_lib.STRATEGY_AUTO)


def _stats():
    from cudf_amd import _lib
    return _lib.last_stats()


@pytest.mark.parametrize("key_dtype", [np.int8, np.int16, np.int32, np.int64, np.uint64,
                                       np.float32, np.float64])
@pytest.mark.parametrize("policy", [0, 1])
def test_key_types_partitioned(key_dtype, policy):
    rng = np.random.default_rng(21)
    n = 50_000
    if np.dtype(key_dtype).kind == "f":
        k = (rng.integers(-3000, 3000, n) / 8).astype(key_dtype)
        k[rng.random(n) < 0.01] = np.nan
        k[rng.random(n) < 0.01] = -0.0
    else:
        hi = min(20_000, np.iinfo(key_dtype).max)
        k = rng.integers(0, hi, n).astype(key_dtype)
    kvalid = rng.random(n) > 0.05
    v = rng.integers(-1000, 1000, n).astype(np.int32)
    vvalid = rng.random(n) > 0.1
    assert_parity([host_col(k, kvalid)], [(host_col(v, vvalid), ALL_NUMERIC_AGGS)], policy)
    assert _stats()["strategy"] == 2


@pytest.mark.parametrize("val_dtype", [np.int8, np.int16, np.int32, np.int64, np.uint8, np.uint32,
                                       np.uint64, np.float32, np.float64, np.bool_])
def test_value_types_partitioned(val_dtype):
    rng = np.random.default_rng(22)
    n = 60_000
    k = rng.integers(0, 8000, n).astype(np.int64)
    if val_dtype == np.bool_:
        v = rng.random(n) > 0.5
    elif np.dtype(val_dtype).kind == "f":
        v = (rng.standard_normal(n) * 10).astype(val_dtype)
    else:
        info = np.iinfo(val_dtype)
        v = rng.integers(max(info.min, -10**6), min(info.max, 10**6), n, endpoint=True).astype(val_dtype)
    assert_parity([host_col(k)], [(host_col(v, rng.random(n) > 0.1), ALL_NUMERIC_AGGS)])
    assert _stats()["strategy"] == 2


def test_one_level_partition_c2_shape():
    key, f64, i32 = c2_data(n=1_000_000, card=100_000, seed=31)
    assert_parity([key], [(f64, C2_AGGS), (i32, C2_AGGS)])
    st = _stats()
    assert st["strategy"] == 2 and st["capacity"] <= 256


def test_two_level_partition_c2_shape():
    key, f64, i32 = c2_data(n=3_000_000, card=400_000, seed=32)
    assert_parity([key], [(f64, C2_AGGS), (i32, C2_AGGS)])
    st = _stats()
    assert st["strategy"] == 2 and st["capacity"] > 256


def test_many_value_columns_with_nulls():
    rng = np.random.default_rng(33)
    n = 80_000
    k = host_col(rng.integers(0, 20_000, n).astype(np.int32))
    reqs = []
    for i in range(6):
        v = rng.integers(-50, 50, n).astype([np.int16, np.int32, np.int64, np.float32, np.float64, np.uint16][i])
        reqs.append((host_col(v, rng.random(n) > 0.2), [(AGG["SUM"], 1), (AGG["MAX"], 1), (AGG["VARIANCE"], 1)]))
    assert_parity([k], reqs)


def test_empty_marker_key_and_null_group():
    marker = np.array([0xFFF1A5A5A5A5A5A5], dtype=np.uint64).view(np.int64)[0]
    rng = np.random.default_rng(34)
    n = 20_000
    k = rng.integers(0, 3000, n).astype(np.int64)
    k[rng.random(n) < 0.05] = marker
    kvalid = rng.random(n) > 0.05
    v = rng.integers(0, 100, n).astype(np.int64)
    for policy in (0, 1):
        assert_parity([host_col(k, kvalid)], [(host_col(v), [(AGG["SUM"], 1), (AGG["COUNT_ALL"], 1)])], policy)


def test_underestimated_groups_fall_back_correctly():
    """99% of rows share one key and the last 1% are all distinct: the strided
    sample sees few distinct keys, the LDS bins overflow, and the call must
    still return the exact result (larger partitions, then the global table)."""
    n = 2_000_000
    k = np.zeros(n, dtype=np.int64)
    tail = n // 100
    k[-tail:] = np.arange(1, tail + 1)
    v = np.arange(n, dtype=np.int64) % 1000
    ng = assert_parity([host_col(k)], [(host_col(v), [(AGG["SUM"], 1), (AGG["COUNT_VALID"], 1),
                                                       (AGG["MIN"], 1)])])
    assert ng == tail + 1


def test_all_keys_null_excluded():
    n = 100_000
    k = np.arange(n, dtype=np.int64)
    kv = np.zeros(n, dtype=bool)
    kcols, res = gpu_groupby([host_col(k, kv)], [(host_col(np.ones(n, np.int32)), [(AGG["SUM"], 1)])])
    assert len(kcols[0][0]) == 0 and len(res[0][0][0]) == 0


def _exact_var(keys, vals, ddof):
    """Per-key exact variance (Python fractions) of the valid values."""
    from fractions import Fraction
    groups = {}
    for k, x in zip(keys, vals):
        groups.setdefault(k, []).append(x)
    out = {}
    for k, xs in groups.items():
        n = len(xs)
        if n - ddof <= 0:
            out[k] = None
            continue
        fx = [Fraction(int(x)) if isinstance(x, (int, np.integer)) else Fraction(float(x)) for x in xs]
        mean = sum(fx) / n
        out[k] = float(sum((x - mean) ** 2 for x in fx) / (n - ddof))
    return out


@pytest.mark.parametrize("case", ["f64_small_mean", "f64_large_mean", "i32_exact", "i64_huge",
                                  "f32", "constant_groups"])
def test_variance_one_pass_and_exact_fallback(case):
    """The partitioned path finalises M2 from the first pass's sums where
    that is exact (integers with Σx² < 2^53) or provably within 1e-11
    (float64, small well-conditioned groups), and runs the exact second pass
    for every other group.  Variance is checked strictly, relative 1e-10
    against exact rational arithmetic (no absolute floor)."""
    rng = np.random.default_rng({"f64_small_mean": 1, "f64_large_mean": 2, "i32_exact": 3,
                                 "i64_huge": 4, "f32": 5, "constant_groups": 6}[case])
    n, card = 60_000, 6_000
    k = rng.integers(0, card, n).astype(np.int64)
    if case == "f64_small_mean":
        v = rng.standard_normal(n)
    elif case == "f64_large_mean":
        v = 1e6 + rng.standard_normal(n)          # Σx² / M2 ~ 1e12: ill-conditioned
    elif case == "i32_exact":
        v = rng.integers(-10_000, 10_000, n).astype(np.int32)
    elif case == "i64_huge":
        v = rng.integers(-10**15, 10**15, n).astype(np.int64)  # Σx² beyond 2^53
    elif case == "f32":
        v = rng.random(n).astype(np.float32)
    else:
        v = np.full(n, 3.25)
    kcols, res = gpu_groupby([host_col(k)], [(host_col(v), [(AGG["VARIANCE"], 1), (AGG["COUNT_VALID"], 1)])])
    assert _stats()["strategy"] == 2.0
    want = _exact_var(k.tolist(), v.tolist() if v.dtype.kind != "f" else v, 1)
    gk = kcols[0][0]
    var, valid = res[0][0]
    for key, got, ok in zip(gk.tolist(), var, valid):
        w = want[key]
        if w is None:
            assert not ok
            continue
        assert ok
        assert abs(got - w) <= 1e-10 * abs(w) + (0.0 if w != 0 else 1e-300), (case, key, got, w)


@pytest.mark.parametrize("vdt", [np.float32, np.float64])
def test_variance_of_nan_and_infinite_values(vdt):
    """Groups holding NaN or an infinity (many of one row): the reference's
    two-pass variance gives NaN ((x - mean)² of an infinite or NaN mean), also
    for a single value under ddof = 0, so such groups may not take the
    one-value shortcut of the first-pass rule.  Against the oracle."""
    rng = np.random.default_rng(41 if vdt == np.float32 else 42)
    n = 300_000
    k = rng.integers(0, 150_000, n).astype(np.int64)
    v = rng.standard_normal(n).astype(vdt)
    v[rng.random(n) < 0.02] = np.nan
    v[rng.random(n) < 0.02] = np.inf
    v[rng.random(n) < 0.02] = -np.inf
    reqs = [(host_col(v, rng.random(n) > 0.05), [(AGG["VARIANCE"], 0), (AGG["STD"], 0), (AGG["VARIANCE"], 1),
                                                 (AGG["MEAN"], 1), (AGG["COUNT_VALID"], 1)])]
    assert_parity([host_col(k)], reqs)
    assert _stats()["strategy"] == 2.0


def _string_pool(rng, size, max_len=24):
    return [bytes(rng.integers(97, 123, int(rng.integers(0, max_len)), dtype=np.uint8)) for _ in range(size)]


@pytest.mark.parametrize("policy", [0, 1])
def test_general_keys_partitioned(policy):
    """Several key columns (int32, string, float64 with NaN and -0.0), nulls
    in keys and values: the partitioned path compares key rows instead of
    key bits."""
    from tests.gbutil import orc
    rng = np.random.default_rng(31)
    n = 80_000
    k1 = rng.integers(0, 300, n).astype(np.int32)
    pool = _string_pool(rng, 200)
    k2 = orc.HostColumn.from_strings([pool[i] for i in rng.integers(0, len(pool), n)], rng.random(n) > 0.03)
    k3 = (rng.integers(0, 3, n) * 0.5).astype(np.float64)
    k3[rng.random(n) < 0.02] = np.nan
    k3[rng.random(n) < 0.02] = -0.0
    v = rng.standard_normal(n)
    assert_parity([host_col(k1, rng.random(n) > 0.03), k2, host_col(k3)],
                  [(host_col(v, rng.random(n) > 0.1), ALL_NUMERIC_AGGS)], policy)
    assert _stats()["strategy"] == 2 and _stats()["key_mode"] == 1


@pytest.mark.parametrize("policy", [0, 1])
def test_single_string_key_two_level(policy):
    """One string key with enough groups for a two-level partition."""
    from tests.gbutil import orc
    rng = np.random.default_rng(32)
    n = 300_000
    pool = [b"k%07d" % i + bytes(rng.integers(97, 123, int(rng.integers(0, 9)), dtype=np.uint8))
            for i in range(250_000)]
    keys = orc.HostColumn.from_strings([pool[i] for i in rng.integers(0, len(pool), n)], rng.random(n) > 0.02)
    v = rng.integers(-1000, 1000, n).astype(np.int64)
    assert_parity([keys], [(host_col(v, rng.random(n) > 0.05), C2_AGGS)], policy)
    assert _stats()["strategy"] == 2 and _stats()["capacity"] > 256


@pytest.mark.parametrize("policy", [0, 1])
def test_c3_shape_partitioned(policy):
    """The skewed three-column key of the C3 workload (int32, Zipf int64,
    string) at 200k rows."""
    from tests.gbutil import c3_data
    k1, k2, k3, f32, i64 = c3_data(n=200_000, seed=5)
    aggs = [(AGG["SUM"], 1), (AGG["COUNT_VALID"], 1), (AGG["MAX"], 1)]
    assert_parity([k1, k2, k3], [(f32, aggs), (i64, aggs)], policy)
    assert _stats()["strategy"] == 2


@pytest.mark.parametrize("policy", [0, 1])
def test_general_keys_string_edge_cases(policy):
    """Two string key columns with empty strings, long strings (over 64
    bytes, lengths not multiples of 8) and strings that differ only in the
    last byte, plus an int64 key."""
    from tests.gbutil import orc
    rng = np.random.default_rng(33)
    n = 60_000
    base = [b"", b"a", b"ab", b"x" * 7, b"x" * 8, b"x" * 9, b"y" * 65, b"y" * 64 + b"z", b"q" * 130]
    pool = base + [b"p" * int(rng.integers(0, 40)) + bytes([97 + i % 26]) for i in range(400)]
    s1 = orc.HostColumn.from_strings([pool[i] for i in rng.integers(0, len(pool), n)], rng.random(n) > 0.05)
    s2 = orc.HostColumn.from_strings([pool[i] for i in rng.integers(0, 20, n)])
    k = rng.integers(0, 5, n).astype(np.int64)
    v = rng.integers(-50, 50, n).astype(np.int32)
    assert_parity([s1, host_col(k, rng.random(n) > 0.02), s2],
                  [(host_col(v, rng.random(n) > 0.1), ALL_NUMERIC_AGGS)], policy)
    assert _stats()["strategy"] == 2


def test_general_keys_underestimated_fall_back():
    """99% of rows share one (int, string) key and the last 1% are all
    distinct: the strided sample sees few distinct keys, the bins overflow,
    and the call must still return the exact result (more bins, then the
    global table)."""
    from tests.gbutil import orc
    n = 1_000_000
    tail = n // 100
    k1 = np.zeros(n, dtype=np.int32)
    k1[-tail:] = np.arange(1, tail + 1)
    s = orc.HostColumn.from_strings([b"same"] * (n - tail) + [b"t%d" % (i % 7) for i in range(tail)])
    v = np.arange(n, dtype=np.int64) % 1000
    ng = assert_parity([host_col(k1), s], [(host_col(v), [(AGG["SUM"], 1), (AGG["COUNT_ALL"], 1)])], 0)
    assert ng == tail + 1


@pytest.mark.parametrize("form", ["hashed", "exact"])
@pytest.mark.parametrize("policy", [0, 1])
def test_general_key_forms(form, policy):
    """General keys (float64 with NaN and -0.0, nullable int32, nullable
    strings with empty ones) in both forms of the partitioned path: 64-bit
    words, with every row that joined a group compared with its
    representative afterwards, and key rows compared while probing."""
    from cudf_amd import _lib
    from tests.gbutil import orc
    _lib.set_strategy(_lib.STRATEGY_PARTITIONED if form == "hashed" else _lib.STRATEGY_PARTITIONED_EXACT)
    rng = np.random.default_rng(71 + policy)
    n = 120_000
    f = (rng.integers(-200, 200, n) / 4.0)
    f[rng.random(n) < 0.02] = np.nan
    f[rng.random(n) < 0.02] = -0.0
    i = rng.integers(0, 30, n).astype(np.int32)
    words = [b"", b"a", b"bb", b"x" * 9, b"y" * 33] + [b"w%d" % j for j in range(40)]
    s = orc.HostColumn.from_strings([words[j] for j in rng.integers(0, len(words), n)], rng.random(n) > 0.03)
    v = rng.integers(-1000, 1000, n).astype(np.int64)
    assert_parity([host_col(f), host_col(i, rng.random(n) > 0.04), s],
                  [(host_col(v, rng.random(n) > 0.1), ALL_NUMERIC_AGGS)], policy)
    st = _stats()
    assert st["strategy"] == 2
    if form == "hashed":
        assert st["verified_pairs"] > 0
    else:
        assert st["verified_pairs"] == 0


@pytest.mark.timeout(600)
def test_largest_bin_count_exact_layout():
    """2^16 bins: 14M rows of nearly unique int64 keys and three value
    columns with wide accumulators (few LDS slots per bin), so the estimate
    needs the largest bin count; a key on 0.1% of the rows rules out the
    speculative layout, so the histogram is counted in two windows of 2^15
    bins.  Every group against the oracle."""
    rng = np.random.default_rng(65536)
    n = 14_000_000
    k = rng.integers(1, 1 << 40, n, dtype=np.int64)
    k[rng.random(n) < 0.001] = 0
    vals = [host_col(rng.integers(-10**6, 10**6, n).astype(np.int64), rng.random(n) > 0.05),
            host_col(rng.standard_normal(n), rng.random(n) > 0.05),
            host_col(rng.integers(-10**12, 10**12, n).astype(np.int64))]
    aggs = [(AGG["SUM"], 1), (AGG["MIN"], 1), (AGG["MAX"], 1)]
    ng = assert_parity([host_col(k)], [(vals[0], aggs + [(AGG["COUNT_VALID"], 1)]), (vals[1], aggs), (vals[2], aggs)])
    st = _stats()
    assert st["strategy"] == 2 and st["capacity"] == 65536 and st["rpa_speculative"] == 0, st
    assert ng == len(np.unique(k))

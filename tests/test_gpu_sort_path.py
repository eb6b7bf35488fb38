"""GPU parity of the sort-based group-by: MEDIAN and QUANTILE (every
interpolation, several quantiles), and NUNIQUE / NTH_ELEMENT of keys that
are not pre-sorted, each mixed with hash aggregations, against the oracle
(sort/aggregate.cpp:323-410; sort/group_quantiles.cu:30-80;
quantiles/quantiles_util.hpp).  Known answers from the reference's
median_tests.cpp / quantile_tests.cpp for every value type."""
import numpy as np
import pytest

from oracle import oracle as orc
from tests.gbutil import AGG, assert_parity, gpu_groupby, host_col
from tests.test_oracle import QKATS, quantile_request, run_quantile_kat

pytestmark = pytest.mark.gpu

INTERPS = [orc.LINEAR, orc.LOWER, orc.HIGHER, orc.MIDPOINT, orc.NEAREST]


def _stats():
    from cudf_amd import _lib
    return _lib.last_stats()


@pytest.mark.parametrize("case", QKATS["cases"], ids=[c["name"] for c in QKATS["cases"]])
def test_quantile_known_answers(case):
    for vt in QKATS["types"]:
        run_quantile_kat(case, vt, lambda k, r, p: gpu_groupby(k, r, p))


@pytest.mark.parametrize("policy", [0, 1])
@pytest.mark.parametrize("dt", [np.float64, np.float32, np.int32, np.int64, np.int8, np.uint16, np.uint64])
def test_median_quantiles_random(policy, dt):
    rng = np.random.default_rng(int(np.dtype(dt).itemsize) * 11 + policy)
    n = 150_000
    key = host_col(rng.integers(0, 20_000, n).astype(np.int64), rng.random(n) > 0.02)
    kind = np.dtype(dt).kind
    if kind == "f":
        v = (rng.standard_normal(n) * 100).astype(dt)
        v[rng.random(n) < 0.01] = np.nan
        v[rng.random(n) < 0.01] = -0.0
    elif kind == "u":
        v = rng.integers(0, np.iinfo(dt).max if dt != np.uint64 else 2**63, n, dtype=np.uint64).astype(dt)
    else:
        info = np.iinfo(dt)
        v = rng.integers(info.min, info.max, n, endpoint=True, dtype=np.int64).astype(dt)
    vals = host_col(v, rng.random(n) > 0.1)
    aggs = [(AGG["MEDIAN"], 1)] + [(AGG["QUANTILE"], ((0.0, 0.1, 0.5, 0.73, 1.0), i)) for i in INTERPS] + \
           [(AGG["SUM"], 1), (AGG["COUNT_VALID"], 1), (AGG["MIN"], 1)]
    assert_parity([key], [(vals, aggs)], policy)
    assert _stats()["strategy"] == 4  # the sort-based path


@pytest.mark.parametrize("policy", [0, 1])
def test_sort_path_multicolumn_and_string_keys(policy):
    rng = np.random.default_rng(52 + policy)
    n = 80_000
    k1 = host_col(rng.integers(0, 40, n).astype(np.int32), rng.random(n) > 0.05)
    words = [b"w%03d" % i for i in range(150)] + [b"", b"a" * 40]
    k2 = orc.HostColumn.from_strings([words[i] for i in rng.integers(0, len(words), n)], rng.random(n) > 0.03)
    v = host_col(rng.standard_normal(n), rng.random(n) > 0.1)
    w = host_col(rng.integers(-50, 50, n).astype(np.int32))
    assert_parity([k1, k2], [(v, [(AGG["QUANTILE"], ((0.25, 0.5, 0.75), orc.LINEAR)), (AGG["MEAN"], 1)]),
                             (w, [(AGG["MEDIAN"], 1), (AGG["NUNIQUE"], 0), (AGG["NTH_ELEMENT"], 2 * 1 + 1),
                                  (AGG["NTH_ELEMENT"], 2 * -1 + 0), (AGG["MAX"], 1)])], policy)


@pytest.mark.parametrize("policy", [0, 1])
def test_nunique_nth_on_unsorted_keys(policy):
    """NUNIQUE (both null policies) and NTH_ELEMENT (n = 0, 3, -1, -2; both
    null policies) of keys in random order: the rows of each group keep their
    input order, as the reference's stable sort keeps them."""
    rng = np.random.default_rng(61 + policy)
    n = 120_000
    key = host_col(rng.integers(0, 9_000, n).astype(np.int32), rng.random(n) > 0.02)
    v = host_col(rng.integers(0, 30, n).astype(np.float64), rng.random(n) > 0.15)
    aggs = [(AGG["NUNIQUE"], 0), (AGG["NUNIQUE"], 1)] + [(AGG["NTH_ELEMENT"], 2 * m + inc)
                                                        for m in (0, 3, -1, -2) for inc in (0, 1)]
    assert_parity([key], [(v, aggs + [(AGG["COUNT_ALL"], 1)])], policy)


def test_median_on_presorted_keys_uses_runs():
    """keys_are_sorted=YES with MEDIAN: groups are the runs of equal adjacent
    keys, as the reference's sort path takes them."""
    rng = np.random.default_rng(71)
    n = 60_000
    k = np.sort(rng.integers(0, 5_000, n)).astype(np.int64)
    v = host_col(rng.standard_normal(n), rng.random(n) > 0.05)
    g_keys, g_res = gpu_groupby([host_col(k)], [(v, [(AGG["MEDIAN"], 1), (AGG["SUM"], 1)])], 0, keys_sorted=True)
    assert_parity([host_col(k)], [(v, [(AGG["MEDIAN"], 1), (AGG["SUM"], 1)])], 0, gpu=(g_keys, g_res))


@pytest.mark.timeout(600)
def test_quantiles_10m_rows():
    """10M rows over 1M keys (about 10 rows per group), MEDIAN and three
    quantiles of a float64 column with nulls, against the sharded oracle."""
    rng = np.random.default_rng(81)
    n = 10_000_000
    key = host_col(rng.integers(0, 1_000_000, n).astype(np.int64), rng.random(n) > 0.01)
    v = host_col(rng.standard_normal(n), rng.random(n) > 0.05)
    assert_parity([key], [(v, [(AGG["MEDIAN"], 1), (AGG["QUANTILE"], ((0.1, 0.5, 0.9), orc.NEAREST)),
                               (AGG["COUNT_VALID"], 1)])], 0)


def test_sort_path_string_values_and_string_median_rejected():
    """String values ride along the sort-based group-by (gathered in group
    order); a MEDIAN of strings is a data-type error, as in the reference."""
    from cudf_amd._lib import DataTypeError
    rng = np.random.default_rng(5)
    k = host_col(rng.integers(0, 10, 100).astype(np.int32))
    s = orc.HostColumn.from_strings([b"x"] * 100, rng.random(100) > 0.3)
    assert_parity([k], [(s, [(AGG["COUNT_VALID"], 1)]), (host_col(np.ones(100)), [(AGG["MEDIAN"], 1)])])
    with pytest.raises(DataTypeError):
        gpu_groupby([k], [(s, [(AGG["MEDIAN"], 1)])])

"""pandas-fallback wrapper (cudf_amd.pandas_compat) and Arrow interop
(cudf_amd.interop): the GPU path against pandas itself, the fallback path on
the CPU, and Arrow round trips."""
import numpy as np
import pandas as pd
import pytest

AGGS = ["sum", "count", "min", "max", "mean", "var", "std"]


def frame(seed, n=20_000, card=300):
    rng = np.random.default_rng(seed)
    f = rng.standard_normal(n) * 100
    f[rng.random(n) < 0.05] = np.nan
    i = pd.array(rng.integers(-1000, 1000, n), dtype="Int32")
    i[rng.random(n) < 0.05] = pd.NA
    k_int = pd.array(rng.integers(0, card, n), dtype="Int64")
    k_int[rng.random(n) < 0.02] = pd.NA
    k_str = np.array([f"s{v}" for v in rng.integers(0, 40, n)], dtype=object)
    k_str[rng.random(n) < 0.02] = None
    return pd.DataFrame({"k": k_int, "ks": k_str, "kf": np.round(rng.random(n) * 20) / 4,
                         "f": f, "i": i, "j": rng.integers(0, 100, n).astype(np.int32)})


def _canonical(df):
    """Rows keyed by the string form of the group keys (NaN and NA alike),
    sorted; values as float64 with NaN for missing."""
    keys = df.index.to_frame(index=False)
    for c in keys.columns:
        keys[c] = ["<null>" if pd.isna(v) else str(v) for v in keys[c]]
    vals = pd.DataFrame({i: pd.to_numeric(pd.Series(df.iloc[:, i].to_numpy(dtype=object)), errors="coerce")
                         .astype(np.float64) for i in range(df.shape[1])})
    out = pd.concat([keys, vals], axis=1)
    return out.sort_values(list(keys.columns)).reset_index(drop=True)


def compare(got, want):
    assert list(got.columns) == list(want.columns)
    g, w = _canonical(got), _canonical(want)
    pd.testing.assert_frame_equal(g, w, check_dtype=False, rtol=1e-9, atol=1e-9)


def test_fallback_runs_pandas_without_device():
    """Without a GPU (or for an unsupported request) the wrapper is pandas."""
    import torch
    from cudf_amd.pandas_compat import groupby_agg
    df = frame(1, n=2000)
    spec = {"f": ["sum", "median"]}  # median has no hash-path kernel: pandas runs it
    got, path = groupby_agg(df, "k", spec)
    assert path == "pandas"
    compare(got, df.groupby("k", sort=False).agg(spec))
    if not torch.cuda.is_available():
        got, path = groupby_agg(df, "k", {"f": AGGS})
        assert path == "pandas"


def test_arrow_type_mapping():
    import pyarrow as pa
    from cudf_amd.interop import arrow_type_id
    from cudf_amd.types import TypeId
    assert arrow_type_id(pa.int64()) == TypeId.INT64
    assert arrow_type_id(pa.float32()) == TypeId.FLOAT32
    assert arrow_type_id(pa.large_string()) == TypeId.STRING
    assert arrow_type_id(pa.timestamp("ms")) == TypeId.TIMESTAMP_MILLISECONDS
    with pytest.raises(ValueError):
        arrow_type_id(pa.list_(pa.int32()))


@pytest.mark.gpu
@pytest.mark.parametrize("by", ["k", "ks", "kf", ["k", "ks"]])
@pytest.mark.parametrize("sort", [False, True])
@pytest.mark.parametrize("dropna", [True, False])
def test_gpu_path_matches_pandas(by, sort, dropna):
    from cudf_amd.pandas_compat import groupby_agg
    df = frame(7)
    spec = {"f": AGGS, "i": AGGS, "j": ["sum", "min", "max", "mean"]}
    got, path = groupby_agg(df, by, spec, sort=sort, dropna=dropna, fallback=False)
    assert path == "gpu"
    want = df.groupby(by, sort=sort, dropna=dropna).agg(spec)
    if sort or not isinstance(by, str):
        compare(got, want)
    else:  # first-appearance order is reproduced exactly
        assert list(got.index.astype(object)) == list(want.index.astype(object)) or \
            got.index.isna().sum() == want.index.isna().sum()
        compare(got, want)


@pytest.mark.gpu
@pytest.mark.parametrize("by", ["k", ["k", "ks"]])
def test_gpu_prod_and_string_min_max_match_pandas(by):
    from cudf_amd.pandas_compat import groupby_agg
    rng = np.random.default_rng(9)
    df = frame(11)
    df["p"] = 0.9 + 0.2 * rng.random(len(df))
    df.loc[rng.random(len(df)) < 0.05, "p"] = np.nan
    df["q"] = rng.integers(-2, 3, len(df)).astype(np.int64)
    names = np.array([f"n{v:03d}" for v in rng.integers(0, 500, len(df))], dtype=object)
    names[rng.random(len(df)) < 0.1] = None
    df["name"] = pd.array(names, dtype="string")  # pandas' min / max skip <NA> on string dtype
    spec = {"p": ["prod"], "q": ["prod", "sum"], "name": ["min", "max", "count"]}
    got, path = groupby_agg(df, by, spec, sort=True, dropna=True, fallback=False)
    assert path == "gpu"
    want = df.groupby(by, sort=True, dropna=True).agg(spec)
    assert list(got.columns) == list(want.columns)
    for c in want.columns:
        g, w = got[c].to_numpy(dtype=object), want[c].to_numpy(dtype=object)
        assert len(g) == len(w)
        for x, y in zip(g, w):
            if pd.isna(y):
                assert pd.isna(x), (c, x, y)
            elif isinstance(y, str):
                assert x == y, (c, x, y)
            else:
                assert abs(float(x) - float(y)) <= 1e-9 * max(1.0, abs(float(y))), (c, x, y)


@pytest.mark.gpu
def test_arrow_round_trip_and_groupby():
    import pyarrow as pa
    from cudf_amd import GroupBy, GroupByRequest, Table
    from cudf_amd.aggregation import Aggregation, Kind
    from cudf_amd.interop import from_arrow, to_arrow
    rng = np.random.default_rng(3)
    n = 5000
    arrays = {
        "i32": pa.array(rng.integers(-50, 50, n).astype(np.int32), mask=rng.random(n) < 0.1),
        "f64": pa.array(rng.standard_normal(n), mask=rng.random(n) < 0.1),
        "u8": pa.array(rng.integers(0, 255, n).astype(np.uint8)),
        "b": pa.array(rng.random(n) < 0.5, mask=rng.random(n) < 0.1),
        "s": pa.array([f"x{v}" for v in rng.integers(0, 30, n)], mask=rng.random(n) < 0.1),
        "ts": pa.array(rng.integers(0, 10**12, n), type=pa.int64()).cast(pa.timestamp("ms")),
        "d": pa.array(rng.integers(0, 20000, n).astype(np.int32)).cast(pa.date32()),
    }
    for name, arr in arrays.items():
        back = to_arrow(from_arrow(arr))
        assert back.equals(arr), name
    sliced = arrays["i32"].slice(7, 3000)
    assert to_arrow(from_arrow(sliced)).equals(sliced)
    # a group-by on Arrow-ingested columns equals pandas on the same data
    keys = from_arrow(arrays["s"])
    vals = from_arrow(arrays["i32"])
    ok, (res,) = GroupBy(Table([keys])).aggregate([GroupByRequest(vals, [Aggregation(Kind.SUM)])])
    got = dict(zip(to_arrow(ok.columns()[0]).to_pylist(), to_arrow(res.columns()[0]).to_pylist()))
    df = pd.DataFrame({"s": arrays["s"].to_pandas(), "v": arrays["i32"].to_pandas()})
    want = df.groupby("s", dropna=True)["v"].sum(min_count=1)
    assert len(got) == len(want)
    for k, v in want.items():
        assert (got[k] is None and pd.isna(v)) or got[k] == v


@pytest.mark.gpu
@pytest.mark.parametrize("sort", [False, True])
def test_gpu_nunique_first_last_on_sorted_keys(sort):
    """nunique / first / last on a frame sorted by its key run on the
    pre-sorted path and match pandas; on unsorted keys they fall back."""
    from cudf_amd.pandas_compat import groupby_agg
    df = frame(21)
    df = df.dropna(subset=["k"]).sort_values("k", kind="stable").reset_index(drop=True)
    df["r"] = np.round(df["f"] / 25)  # few distinct values per group (and NaN)
    spec = {"r": ["nunique", "first", "last", "count"], "j": ["nunique", "first", "last"],
            "i": ["first", "last"]}
    got, path = groupby_agg(df, "k", spec, sort=sort, dropna=True, fallback=False)
    assert path == "gpu"
    compare(got, df.groupby("k", sort=sort, dropna=True).agg(spec))
    shuffled = df.sample(frac=1.0, random_state=3)
    got, path = groupby_agg(shuffled, "k", spec, sort=sort)
    assert path == "pandas"

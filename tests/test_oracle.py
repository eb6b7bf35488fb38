"""Pins the CPU oracle before it is trusted as the GPU's checker.

Each oracle function is checked against something other than itself:
  - MurmurHash3_x86_32: smhasher's published verification value;
  - the row hash: the reference's docstring row-hash vectors;
  - group-by semantics: the reference's C++ known-answer tests (fixtures in
    tests/golden/groupby_kats.json), an exact brute force (Python integers,
    fractions and math.fsum) on small random inputs, and pandas -- the CPU
    fallback of cudf.pandas -- on seeded data.
A last test plants mistakes into a Python mirror of the row hash to show the
golden vectors catch them.
"""
import json
import math
import os
import struct
from fractions import Fraction

import numpy as np
import pytest

from oracle import oracle as orc
from tests.gbutil import AGG, host_col, oracle_groupby, sort_order

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")

NP_TYPES = {"INT8": np.int8, "INT16": np.int16, "INT32": np.int32, "INT64": np.int64,
            "UINT8": np.uint8, "UINT16": np.uint16, "UINT32": np.uint32, "UINT64": np.uint64,
            "FLOAT32": np.float32, "FLOAT64": np.float64}


def load(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


# ---------------------------------------------------------------- hashing ----
def test_smhasher_verification_value():
    blob = b""
    for i in range(256):
        blob += struct.pack("<I", orc.murmur3_x86_32(bytes(range(i)), 256 - i))
    expect = int(load("hash_vectors.json")["smhasher_verification_x86_32"], 16)
    assert orc.murmur3_x86_32(blob, 0) == expect


def _cols(spec):
    return [host_col(np.array(c["values"], dtype=NP_TYPES[c["type"]])) for c in spec]


def test_docstring_row_hashes_seed0():
    for case in load("hash_vectors.json")["docstring"]:
        if case["seed"] != 0:
            continue
        h = orc.hash_rows(_cols(case["columns"]), seed=0)
        assert list(h.view(np.int32)) == case["expect_int32"], case["source"]


def test_docstring_row_hash_seed42_documented_discrepancy():
    case = [c for c in load("hash_vectors.json")["docstring"] if c["seed"] == 42][0]
    cols = _cols(case["columns"])
    assert list(orc.hash_rows(cols, seed=42, elem_seed=0)) == case["expect_uint32"]
    assert list(orc.hash_rows(cols, seed=42)) != case["expect_uint32"]


def _py_murmur(data: bytes, seed: int, mutate=None) -> int:
    """Independent Python MurmurHash3_x86_32 used only to show that the
    golden vectors reject plausible mistakes."""
    m = 0xFFFFFFFF
    c1, c2 = 0xCC9E2D51, 0x1B873593
    rot1 = 15 if mutate != "rot" else 16
    h = seed
    n = len(data) // 4
    for i in range(n):
        k = int.from_bytes(data[4 * i:4 * i + 4], "little")
        k = (k * c1) & m
        k = ((k << rot1) | (k >> (32 - rot1))) & m
        k = (k * c2) & m
        h ^= k
        h = ((h << 13) | (h >> 19)) & m
        h = (h * 5 + 0xE6546B64) & m
    tail = data[4 * n:]
    k = 0
    for j in reversed(range(len(tail))):
        k ^= tail[j] << (8 * j)
    if tail:
        k = (k * c1) & m
        k = ((k << rot1) | (k >> (32 - rot1))) & m
        k = (k * c2) & m
        h ^= k
    if mutate != "len":
        h ^= len(data)
    h ^= h >> 16
    h = (h * 0x85EBCA6B) & m
    h ^= h >> 13
    h = (h * 0xC2B2AE35) & m
    h ^= h >> 16
    return h


def _py_row_hash(vals, mutate=None):
    h = 0
    for v in vals:
        e = _py_murmur(struct.pack("<q", v), 0, mutate)
        if mutate == "combine":
            h = h ^ (e + 0x9E3779B9 + (h << 6) + (h >> 2)) & 0xFFFFFFFF  # wrong precedence
        else:
            h = (h ^ ((e + 0x9E3779B9 + (h << 6) + (h >> 2)) & 0xFFFFFFFF)) & 0xFFFFFFFF
    return h


@pytest.mark.parametrize("mutate", [None, "rot", "len"])
def test_golden_vectors_reject_planted_mistakes(mutate):
    case = load("hash_vectors.json")["docstring"][0]
    got = [np.uint32(_py_row_hash([v], mutate)).view(np.int32) for v in case["columns"][0]["values"]]
    ok = [int(x) for x in got] == case["expect_int32"]
    assert ok == (mutate is None)


def test_oracle_hash_matches_independent_python():
    rng = np.random.default_rng(1)
    for _ in range(50):
        L = int(rng.integers(0, 40))
        b = bytes(rng.integers(0, 256, L, dtype=np.uint8))
        s = int(rng.integers(0, 2**32))
        assert orc.murmur3_x86_32(b, s) == _py_murmur(b, s)


def test_float_key_normalisation():
    # -0.0 hashes like 0.0 and every NaN like the canonical NaN
    # (hash_functions.cuh:28-47)
    a = host_col(np.array([0.0, -0.0, np.nan, -np.nan, 1.5]))
    h = orc.hash_rows([a])
    assert h[0] == h[1] and h[2] == h[3] and h[0] != h[4]
    nan_payload = np.array([0x7FF8000000000123], dtype=np.uint64).view(np.float64)
    assert orc.hash_rows([host_col(nan_payload)])[0] == h[2]


def test_null_hash_is_all_ones_combined():
    c = host_col(np.array([5], np.int32), np.array([False]))
    # hash_combine(0, 0xFFFFFFFF)
    expect = (0 ^ ((0xFFFFFFFF + 0x9E3779B9) & 0xFFFFFFFF))
    assert int(orc.hash_rows([c])[0]) == expect


# --------------------------------------------------------------- group-by ----
def _kat_value_types(case):
    if "val_types" in case:
        return case["val_types"]
    types = ["INT8", "INT16", "INT32", "INT64", "UINT32", "FLOAT32", "FLOAT64"]
    if case["agg"] in ("MIN", "MAX", "COUNT_VALID", "COUNT_ALL"):
        types += ["UINT8", "UINT64"]
    return types


def _run_kat(case, vt, groupby_fn):
    keys = host_col(np.array(case["keys"], np.int32),
                    None if case["key_valid"] is None else np.array(case["key_valid"], bool))
    vals = host_col(np.array(case["vals"], NP_TYPES[vt]),
                    None if case["val_valid"] is None else np.array(case["val_valid"], bool))
    policy = 0 if case["null_policy"] == "EXCLUDE" else 1
    kcols, res = groupby_fn([keys], [(vals, [(AGG[case["agg"]], case["ddof"])])], policy)
    order = sort_order(kcols)
    kv, kvalid = kcols[0]
    got_keys = [None if (kvalid is not None and not kvalid[i]) else int(kv[i]) for i in order]
    exp_keys = [None if (case["expect_key_valid"] is not None and not case["expect_key_valid"][j])
                else case["expect_keys"][j] for j in range(len(case["expect_keys"]))]
    # expected lists put the null key last, as the sort token does
    assert got_keys == exp_keys
    v, valid = res[0][0]
    valid = np.ones(len(v), bool) if valid is None else valid
    exp_valid = case["expect_val_valid"] or [1] * len(case["expect_vals"])
    for j, i in enumerate(order):
        assert bool(valid[i]) == bool(exp_valid[j]), (case["name"], vt, j)
        if exp_valid[j]:
            e = case["expect_vals"][j]
            if isinstance(e, float):
                assert math.isclose(float(v[i]), e, rel_tol=1e-15), (case["name"], vt, v[i], e)
            else:
                assert int(v[i]) == e, (case["name"], vt, v[i], e)


KATS = load("groupby_kats.json")


@pytest.mark.parametrize("case", KATS["cases"], ids=[c["name"] for c in KATS["cases"]])
def test_oracle_known_answers(case):
    for vt in _kat_value_types(case):
        _run_kat(case, vt, oracle_groupby)


@pytest.mark.parametrize("case", KATS["string_value_arg_cases"], ids=lambda c: c["name"])
def test_oracle_string_value_argmin_argmax(case):
    keys = host_col(np.array(case["keys"], np.int32))
    vals = orc.HostColumn.from_strings([s.encode() for s in case["vals"]])
    kcols, res = oracle_groupby([keys], [(vals, [(AGG[case["agg"]], 1)])])
    got = {int(kcols[0][0][i]): int(res[0][0][0][i]) for i in range(len(kcols[0][0]))}
    assert got == dict(zip(case["expect_keys"], case["expect_vals"]))


def test_oracle_argmin_argmax_brute_force():
    """Ties take the first row; a NaN wins only in an all-NaN group."""
    rng = np.random.default_rng(12)
    for trial in range(20):
        n = 300
        k = rng.integers(0, 15, n).astype(np.int32)
        v = rng.integers(-5, 5, n).astype(np.float64)
        v[rng.random(n) < 0.1] = np.nan
        valid = rng.random(n) > 0.15
        kcols, res = oracle_groupby([host_col(k)], [(host_col(v, valid), [(AGG["ARGMIN"], 1), (AGG["ARGMAX"], 1)])])
        for gi in range(len(kcols[0][0])):
            key = int(kcols[0][0][gi])
            rows = [r for r in range(n) if k[r] == key and valid[r]]
            for a, pick in ((0, min), (1, max)):
                got, ok = res[0][a][0][gi], res[0][a][1][gi]
                if not rows:
                    assert not ok
                    continue
                nums = [r for r in rows if not np.isnan(v[r])]
                if nums:
                    best = pick(v[r] for r in nums)
                    want = next(r for r in nums if v[r] == best)
                else:
                    want = rows[0]
                assert ok and got == want, (trial, key, a, got, want)


def test_oracle_string_keys_known_answer():
    case = KATS["string_key_cases"][0]
    keys = orc.HostColumn.from_strings([s.encode() for s in case["keys"]])
    vals = host_col(np.array(case["vals"], np.int32))
    kcols, res = oracle_groupby([keys], [(vals, [(AGG["SUM"], 1)])])
    order = sort_order(kcols)
    assert [kcols[0][0][i] for i in order] == sorted(s.encode() for s in case["expect_keys"])
    exp = dict(zip([s.encode() for s in case["expect_keys"]], case["expect_vals"]))
    for i in order:
        assert int(res[0][0][0][i]) == exp[kcols[0][0][i]]


def _brute(keys, vals, valid_k, valid_v, kind, ddof, policy):
    """Exact group-by with Python integers / Fractions / fsum."""
    groups = {}
    order = []
    for r in range(len(keys)):
        kv = bool(valid_k[r])
        if not kv and policy == 0:
            continue
        k = (kv, keys[r] if kv else None)
        if k not in groups:
            groups[k] = []
            order.append(k)
        groups[k].append(r)
    out = {}
    for k in order:
        rows = groups[k]
        xs = [float(vals[r]) if isinstance(vals[r], (float, np.floating)) else int(vals[r])
              for r in rows if valid_v[r]]
        if kind == "COUNT_ALL":
            out[k] = (len(rows), True)
        elif kind == "COUNT_VALID":
            out[k] = (len(xs), True)
        elif not xs:
            out[k] = (None, False)
        elif kind == "SUM":
            if isinstance(xs[0], float):
                out[k] = (math.fsum(xs), True)
            else:
                s = sum(int(x) for x in xs)
                s = (s + 2**63) % 2**64 - 2**63
                out[k] = (s, True)
        elif kind == "PRODUCT":
            if isinstance(xs[0], float):
                pr = Fraction(1)
                for x in xs:
                    pr *= Fraction(x)
                out[k] = (float(pr), True)
            else:
                pr = 1
                for x in xs:
                    pr = (pr * int(x)) % 2**64
                out[k] = ((pr + 2**63) % 2**64 - 2**63, True)
        elif kind == "MIN":
            out[k] = (min(xs), True)
        elif kind == "MAX":
            out[k] = (max(xs), True)
        elif kind == "MEAN":
            out[k] = (float(sum(Fraction(x) for x in xs) / len(xs)), True)
        elif kind == "VARIANCE":
            n = len(xs)
            if n - ddof <= 0:
                out[k] = (None, False)
            else:
                fx = [Fraction(x) for x in xs]
                m = sum(fx) / n
                out[k] = (float(sum((x - m) ** 2 for x in fx) / (n - ddof)), True)
    return out


@pytest.mark.parametrize("seed", range(12))
def test_oracle_matches_exact_brute_force(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(1, 300))
    nk = int(rng.integers(1, 40))
    keys = rng.integers(-5, nk, n).astype(np.int64)
    valid_k = rng.random(n) > 0.1
    is_float = seed % 2 == 0
    vals = rng.standard_normal(n) * 100 if is_float else rng.integers(-10**12, 10**12, n)
    valid_v = rng.random(n) > 0.2
    policy = seed % 3 == 0
    for kind in ["SUM", "MIN", "MAX", "COUNT_VALID", "COUNT_ALL", "MEAN", "VARIANCE", "PRODUCT"]:
        ddof = int(rng.integers(0, 3)) if kind == "VARIANCE" else 1
        kc = host_col(keys, valid_k)
        # products of a few factors near 1 (floats) or wrapping int64 (integers)
        pv = (1 + vals / 1e4) if (kind == "PRODUCT" and is_float) else vals
        vc = host_col(pv, valid_v)
        kcols, res = oracle_groupby([kc], [(vc, [(AGG[kind], ddof)])], int(policy))
        expect = _brute(list(keys), list(pv), valid_k, valid_v, kind, ddof, int(policy))
        kv, kvalid = kcols[0]
        got = {}
        for i in range(len(kv)):
            valid = True if kvalid is None else bool(kvalid[i])
            got[(valid, kv[i] if valid else None)] = (res[0][0][0][i], bool(res[0][0][1][i]))
        assert set(got) == set(expect)
        for k, (e, ev) in expect.items():
            g, gv = got[k]
            assert gv == ev, (kind, k)
            if not ev:
                continue
            if kind in ("MEAN", "VARIANCE"):
                assert math.isclose(float(g), e, rel_tol=1e-14, abs_tol=1e-300), (kind, g, e)
            elif kind == "PRODUCT" and isinstance(e, float):
                assert float(g) == e, (kind, g, e)  # exact product rounded once
            elif isinstance(e, float):
                assert float(g) == e, (kind, g, e)
            else:
                assert int(g) == int(e), (kind, g, e)


def test_oracle_min_max_ignore_nan_like_reference():
    # device_atomics.cuh:107-153 with std::min/std::max keeps the identity
    # against NaN, so an all-NaN group yields +inf (MIN) / -inf (MAX).
    keys = host_col(np.array([1, 1, 2, 2], np.int32))
    vals = host_col(np.array([np.nan, np.nan, np.nan, 3.0]))
    kcols, res = oracle_groupby([keys], [(vals, [(AGG["MIN"], 1), (AGG["MAX"], 1)])])
    got = {int(k): (res[0][0][0][i], res[0][1][0][i]) for i, k in enumerate(kcols[0][0])}
    assert got[1][0] == np.inf and got[1][1] == -np.inf
    assert got[2] == (3.0, 3.0)


def test_oracle_vs_pandas_c1():
    pd = pytest.importorskip("pandas")
    rng = np.random.default_rng(0)
    n = 200_000
    key = rng.integers(0, 1000, n, dtype=np.int32)
    v = rng.integers(-10**6, 10**6, n, endpoint=True, dtype=np.int64)
    kcols, res = oracle_groupby([host_col(key)], [(host_col(v), [(AGG["SUM"], 1), (AGG["COUNT_VALID"], 1),
                                                                 (AGG["MIN"], 1), (AGG["MAX"], 1)])])
    df = pd.DataFrame({"k": key, "v": v})
    e = df.groupby("k", sort=True)["v"].agg(["sum", "count", "min", "max"])
    order = np.argsort(kcols[0][0])
    assert np.array_equal(kcols[0][0][order], e.index.to_numpy())
    for a, name in enumerate(["sum", "count", "min", "max"]):
        assert np.array_equal(res[0][a][0][order].astype(np.int64), e[name].to_numpy().astype(np.int64))


def test_oracle_vs_pandas_nullable_c2_shape():
    pd = pytest.importorskip("pandas")
    rng = np.random.default_rng(3)
    n, card = 100_000, 10_000
    key = rng.integers(0, card, n, dtype=np.int64)
    kvalid = rng.random(n) >= 0.01
    f64 = rng.standard_normal(n)
    fvalid = rng.random(n) >= 0.05
    i32 = rng.integers(0, 10_000, n, endpoint=True, dtype=np.int32)
    ivalid = rng.random(n) >= 0.05
    aggs = [(AGG["SUM"], 1), (AGG["MEAN"], 1), (AGG["COUNT_VALID"], 1), (AGG["MIN"], 1),
            (AGG["MAX"], 1), (AGG["VARIANCE"], 1)]
    kcols, res = oracle_groupby([host_col(key, kvalid)],
                                [(host_col(f64, fvalid), aggs), (host_col(i32, ivalid), aggs)])
    df = pd.DataFrame({
        "k": pd.arrays.IntegerArray(key, ~kvalid),
        "f": pd.arrays.FloatingArray(f64, ~fvalid),
        "i": pd.arrays.IntegerArray(i32, ~ivalid),
    })
    g = df.groupby("k", sort=True, dropna=True)
    order = np.argsort(kcols[0][0])
    assert np.array_equal(kcols[0][0][order], g.size().index.to_numpy().astype(np.int64))
    for r, col in enumerate(["f", "i"]):
        s = g[col]
        expect = {
            0: s.sum(min_count=1), 1: s.mean(), 2: s.count(), 3: s.min(), 4: s.max(), 5: s.var(ddof=1),
        }
        for a in range(6):
            ov, ovalid = res[r][a]
            ov, ovalid = ov[order], ovalid[order]
            ex = expect[a]
            exv = ex.notna().to_numpy()
            assert np.array_equal(ovalid, exv), (col, a)
            exn = ex.to_numpy(dtype=np.float64, na_value=np.nan)[exv]
            got = ov[ovalid].astype(np.float64)
            if a in (2, 3, 4) or (a == 0 and col == "i"):
                assert np.array_equal(got, exn), (col, a)
            elif a == 5:
                assert np.allclose(got, exn, rtol=1e-10, atol=0), (col, a)
            else:
                assert np.allclose(got, exn, rtol=1e-12, atol=1e-12), (col, a)


# --------------------------------------------------------------- NUNIQUE --
# cpp/tests/groupby/nunique_tests.cpp: (keys, key validity, values, value
# validity, null policy, expected keys, expected counts, expected for bool8)
NUNIQUE_KATS = {
    "basic": ([1, 2, 3, 1, 2, 2, 1, 3, 3, 2], None, [0, 1, 2, 3, 4, 5, 6, 7, 8, 9], None, 0,
              [1, 2, 3], [3, 4, 3], [2, 1, 1]),                                               # :31-52
    "basic_duplicates": ([1, 2, 3, 1, 2, 2, 1, 3, 3, 2], None, [0, 1, 2, 3, 4, 5, 3, 2, 2, 9], None, 0,
                         [1, 2, 3], [2, 4, 1], [2, 1, 1]),                                    # :70-88
    "zero_valid_values": ([1, 1, 1], None, [0, 0, 0], [0, 0, 0], 0, [1], [0], [0]),           # :106-120
    "null_keys_and_values": ([1, 2, 3, 1, 2, 2, 1, 3, 3, 2, 4], [1, 1, 1, 1, 1, 1, 1, 0, 1, 1, 1],
                             [0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 4], [0, 1, 1, 1, 1, 0, 1, 1, 1, 1, 0], 0,
                             [1, 2, 3, 4], [2, 3, 2, 0], [1, 1, 1, 0]),                       # :122-146
    "null_keys_and_values_with_duplicates": (
        [1, 2, 3, 3, 1, 2, 2, 1, 3, 3, 2, 4, 4, 2], [1, 1, 1, 1, 1, 1, 1, 1, 0, 1, 1, 1, 1, 1],
        [0, 1, 2, 2, 3, 4, 5, 6, 7, 8, 9, 4, 4, 2], [0, 1, 1, 1, 1, 1, 0, 1, 1, 1, 1, 0, 0, 0], 0,
        [1, 2, 3, 4], [2, 3, 2, 0], [1, 1, 1, 0]),                                            # :148-173
    "include_nulls": (
        [1, 2, 3, 3, 1, 2, 2, 1, 3, 3, 2, 4, 4, 2], [1, 1, 1, 1, 1, 1, 1, 1, 0, 1, 1, 1, 1, 1],
        [0, 1, 2, 2, 3, 4, 5, 6, 7, 8, 9, 4, 4, 2], [0, 1, 1, 1, 1, 1, 0, 1, 1, 1, 1, 0, 0, 0], 1,
        [1, 2, 3, 4], [3, 4, 2, 1], [2, 2, 1, 1]),                                            # :175-200
}


def nunique_kat_columns(name, vdt):
    k, kv, v, vv, policy, ek, ev, evb = NUNIQUE_KATS[name]
    kcol = orc.HostColumn.from_numpy(np.array(k, np.int32), None if kv is None else np.array(kv, bool))
    varr = np.array(v).astype(vdt)
    vcol = orc.HostColumn.from_numpy(varr, None if vv is None else np.array(vv, bool))
    return kcol, vcol, policy, ek, (evb if vdt == np.bool_ else ev)


@pytest.mark.parametrize("vdt", [np.int8, np.int32, np.int64, np.uint32, np.float32, np.float64, np.bool_])
@pytest.mark.parametrize("name", sorted(NUNIQUE_KATS))
def test_nunique_reference_kats(name, vdt):
    kcol, vcol, policy, ek, ev = nunique_kat_columns(name, vdt)
    rep, res = orc.groupby([kcol], [(vcol, [(orc.NUNIQUE, policy)])], 0)
    order = np.argsort(kcol.data[rep])
    assert list(kcol.data[rep][order]) == ek
    got, valid = res[0][0]
    assert got.dtype == np.int32 and valid.all()
    assert list(got[order]) == ev


def _nunique_brute(keys, kvalid, vals, vvalid, include_nulls):
    """len(set(...)) per group, NaN as one value, -0.0 as 0.0, null as one value."""
    groups = {}
    for i in range(len(keys)):
        if not kvalid[i]:
            continue
        s = groups.setdefault(keys[i], set())
        if not vvalid[i]:
            if include_nulls:
                s.add(("null",))
            continue
        x = vals[i]
        if isinstance(x, float) and x != x:
            s.add(("nan",))
        else:
            s.add(x + 0 if not isinstance(x, bytes) else x)  # -0.0 + 0 == 0.0
    return groups


@pytest.mark.parametrize("vdt", [np.float64, np.int16, "str"])
@pytest.mark.parametrize("include_nulls", [0, 1])
def test_nunique_brute_force(vdt, include_nulls):
    rng = np.random.default_rng(17 + include_nulls)
    n = 3000
    k = rng.integers(0, 60, n).astype(np.int64)
    kvalid = rng.random(n) > 0.05
    vvalid = rng.random(n) > 0.15
    if vdt == "str":
        pool = [b"", b"a", b"ab", b"b", b"abc", b"ba"]
        strs = [pool[i] for i in rng.integers(0, len(pool), n)]
        vcol = orc.HostColumn.from_strings(strs, vvalid)
        py_vals = strs
    elif vdt == np.float64:
        v = rng.integers(-4, 5, n).astype(np.float64)
        v[rng.random(n) < 0.1] = np.nan
        v[rng.random(n) < 0.1] = -0.0
        vcol = orc.HostColumn.from_numpy(v, vvalid)
        py_vals = [float(x) for x in v]
    else:
        v = rng.integers(-20, 20, n).astype(vdt)
        vcol = orc.HostColumn.from_numpy(v, vvalid)
        py_vals = [int(x) for x in v]
    kcol = orc.HostColumn.from_numpy(k, kvalid)
    rep, res = orc.groupby([kcol], [(vcol, [(orc.NUNIQUE, include_nulls)])], 0)
    want = _nunique_brute([int(x) for x in k], kvalid, py_vals, vvalid, include_nulls)
    got = {int(k[r]): int(c) for r, c in zip(rep, res[0][0][0])}
    assert got == {g: len(s) for g, s in want.items()}
    # a planted mistake (-0.0 distinct from 0.0) changes the brute force's answer
    if vdt == np.float64:
        def planted(x):
            return ("negzero",) if x == 0 and np.signbit(x) else x
        bad = {}
        for i in range(n):
            if kvalid[i] and vvalid[i] and py_vals[i] == py_vals[i]:
                bad.setdefault(int(k[i]), set()).add(planted(py_vals[i]))
        assert any(len(bad.get(g, ())) != len([x for x in s if x not in (("null",), ("nan",))])
                   for g, s in want.items())


# ----------------------------------------------------------- NTH_ELEMENT --
_K10 = [1, 2, 3, 1, 2, 2, 1, 3, 3, 2]
_K11 = ([1, 2, 3, 1, 2, 2, 1, 3, 3, 2, 4], [1, 1, 1, 1, 1, 1, 1, 0, 1, 1, 1])
_V11 = ([0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 4], [0, 1, 1, 1, 1, 0, 1, 1, 1, 1, 0])
_K14 = ([1, 2, 3, 3, 1, 2, 2, 1, 3, 3, 2, 4, 4, 2], [1, 1, 1, 1, 1, 1, 1, 1, 0, 1, 1, 1, 1, 1])
_V14 = ([0, 1, 2, 2, 3, 4, 5, 6, 7, 8, 9, 4, 4, 2], [0, 1, 1, 1, 1, 1, 0, 1, 1, 1, 1, 1, 0, 0])
_ALL10 = list(range(10))
# cpp/tests/groupby/nth_element_tests.cpp: (keys, key validity, values, value
# validity, n, include nulls, expected keys, expected values (None = null))
NTH_KATS = {
    "basic_0": (_K10, None, _ALL10, None, 0, 1, [1, 2, 3], [0, 1, 2]),                      # :35-59
    "basic_1": (_K10, None, _ALL10, None, 1, 1, [1, 2, 3], [3, 4, 7]),
    "basic_2": (_K10, None, _ALL10, None, 2, 1, [1, 2, 3], [6, 5, 8]),
    "out_of_bounds": (_K10, None, [0, 1, 2, 3, 4, 5, 3, 2, 2, 9], None, 3, 1, [1, 2, 3],
                      [None, 9, None]),                                                         # :78-92
    "negative_1": (_K10, None, _ALL10, None, -1, 1, [1, 2, 3], [6, 9, 8]),                   # :94-119
    "negative_2": (_K10, None, _ALL10, None, -2, 1, [1, 2, 3], [3, 5, 7]),
    "negative_3": (_K10, None, _ALL10, None, -3, 1, [1, 2, 3], [0, 4, 2]),
    "negative_out_of_bounds": (_K10, None, [0, 1, 2, 3, 4, 5, 3, 2, 2, 9], None, -4, 1, [1, 2, 3],
                               [None, 1, None]),                                                # :121-135
    "zero_valid_values": ([1, 1, 1], None, [3, 4, 5], [0, 0, 0], 0, 1, [1], [None]),          # :153-167
    "null_keys_and_values": (_K11[0], _K11[1], _V11[0], _V11[1], 0, 1, [1, 2, 3, 4],
                             [None, 1, 2, None]),                                               # :169-187
    "null_keys_and_values_out_of_bounds": (_K11[0], _K11[1], _V11[0], _V11[1], 2, 1, [1, 2, 3, 4],
                                           [6, None, None, None]),                              # :189-207
    "include_nulls_0": (_K14[0], _K14[1], _V14[0], _V14[1], 0, 1, [1, 2, 3, 4], [None, 1, 2, 4]),  # :209-249
    "include_nulls_1": (_K14[0], _K14[1], _V14[0], _V14[1], 1, 1, [1, 2, 3, 4], [3, 4, 2, None]),
    "include_nulls_2": (_K14[0], _K14[1], _V14[0], _V14[1], 2, 1, [1, 2, 3, 4], [6, None, 8, None]),
    "exclude_nulls_0": (_K14[0], _K14[1], _V14[0], _V14[1], 0, 0, [1, 2, 3, 4], [3, 1, 2, 4]),
    "exclude_nulls_1": (_K14[0], _K14[1], _V14[0], _V14[1], 1, 0, [1, 2, 3, 4], [6, 4, 2, None]),
    "exclude_nulls_2": (_K14[0], _K14[1], _V14[0], _V14[1], 2, 0, [1, 2, 3, 4], [None, 9, 8, None]),
}


def nth_kat_columns(name, vdt):
    k, kv, v, vv, n, include, ek, ev = NTH_KATS[name]
    kcol = orc.HostColumn.from_numpy(np.array(k, np.int32), None if kv is None else np.array(kv, bool))
    vcol = orc.HostColumn.from_numpy(np.array(v).astype(vdt), None if vv is None else np.array(vv, bool))
    return kcol, vcol, 2 * n + include, ek, ev


@pytest.mark.parametrize("vdt", [np.int8, np.int32, np.int64, np.uint16, np.float32, np.float64])
@pytest.mark.parametrize("name", sorted(NTH_KATS))
def test_nth_element_reference_kats(name, vdt):
    kcol, vcol, param, ek, ev = nth_kat_columns(name, vdt)
    rep, res = orc.groupby([kcol], [(vcol, [(orc.NTH_ELEMENT, param)])], 0)
    order = np.argsort(kcol.data[rep])
    assert list(kcol.data[rep][order]) == ek
    got, valid = res[0][0]
    assert got.dtype == np.dtype(vdt)
    assert [(None if not ok else int(x)) for x, ok in zip(got[order], valid[order])] == ev


@pytest.mark.parametrize("n", [0, 1, 3, -1, -2, -5])
@pytest.mark.parametrize("include", [0, 1])
def test_nth_element_brute_force(n, include):
    """Per-group lists of counted rows in row order, indexed with Python's own
    negative indexing."""
    rng = np.random.default_rng(40 + n * 2 + include)
    m = 2000
    k = rng.integers(0, 80, m).astype(np.int32)
    kvalid = rng.random(m) > 0.05
    v = rng.integers(-1000, 1000, m).astype(np.int64)
    vvalid = rng.random(m) > 0.3
    rep, res = orc.groupby([orc.HostColumn.from_numpy(k, kvalid)],
                           [(orc.HostColumn.from_numpy(v, vvalid), [(orc.NTH_ELEMENT, 2 * n + include)])], 0)
    lists = {}
    for i in range(m):
        if kvalid[i]:
            lists.setdefault(int(k[i]), [])
            if include or vvalid[i]:
                lists[int(k[i])].append(i)
    got, valid = res[0][0]
    for gi, r in enumerate(rep):
        rows = lists[int(k[r])]
        want = rows[n] if -len(rows) <= n < len(rows) else None
        if want is None or not vvalid[want]:
            assert not valid[gi]
        else:
            assert valid[gi] and got[gi] == v[want]


# ---------------------------------------------------------------- quantiles --
QKATS = load("quantile_kats.json")
INTERP = {"LINEAR": orc.LINEAR, "LOWER": orc.LOWER, "HIGHER": orc.HIGHER, "MIDPOINT": orc.MIDPOINT,
          "NEAREST": orc.NEAREST}


def quantile_request(case):
    if case["agg"] == "MEDIAN":
        return (AGG["MEDIAN"], 1)
    return (AGG["QUANTILE"], (tuple(case["quantiles"]), INTERP[case["interpolation"]]))


def run_quantile_kat(case, vt, groupby_fn):
    """One MEDIAN / QUANTILE known answer: keys and the len(quantiles) values
    of each group after sorting by key."""
    keys = host_col(np.array(case["keys"], np.int32),
                    None if case["key_valid"] is None else np.array(case["key_valid"], bool))
    vals = host_col(np.array(case["vals"], np.dtype(vt)),
                    None if case["val_valid"] is None else np.array(case["val_valid"], bool))
    kcols, res = groupby_fn([keys], [(vals, [quantile_request(case)])], 0)
    order = sort_order(kcols)
    assert [int(kcols[0][0][i]) for i in order] == case["expect_keys"], case["name"]
    nq = len(case["quantiles"]) if case["agg"] == "QUANTILE" else 1
    v, valid = res[0][0]
    valid = np.ones(len(v), bool) if valid is None else valid
    assert v.dtype == np.float64 and len(v) == nq * len(order)
    rows = (order[:, None] * nq + np.arange(nq)).reshape(-1)
    exp_valid = case["expect_val_valid"] or [1] * len(case["expect_vals"])
    for j, i in enumerate(rows):
        assert bool(valid[i]) == bool(exp_valid[j]), (case["name"], vt, j)
        if exp_valid[j]:  # the reference compares floating columns to a few ulps
            assert math.isclose(v[i], case["expect_vals"][j], rel_tol=4e-16 * 4), (case["name"], vt, v[i],
                                                                                  case["expect_vals"][j])


@pytest.mark.parametrize("case", QKATS["cases"], ids=[c["name"] for c in QKATS["cases"]])
def test_oracle_quantile_known_answers(case):
    for vt in QKATS["types"]:
        run_quantile_kat(case, vt, oracle_groupby)


@pytest.mark.parametrize("interp", ["LINEAR", "LOWER", "HIGHER", "NEAREST", "MIDPOINT"])
def test_oracle_quantiles_against_numpy(interp):
    """numpy.quantile's methods of the same names (q * (n - 1) positions,
    round-half-to-even for nearest) on random groups without NaN; MIDPOINT
    against the exact (a + b) / 2 of each pair, rounded once."""
    rng = np.random.default_rng(41)
    n = 4000
    k = rng.integers(0, 60, n).astype(np.int32)
    for dt in (np.float64, np.float32, np.int32, np.int64, np.uint16):
        v = (rng.standard_normal(n) * 1000).astype(dt) if np.dtype(dt).kind == "f" else \
            rng.integers(-10**6 if np.dtype(dt).kind == "i" else 0, 10**6 if np.dtype(dt).kind == "i" else 60000,
                         n).astype(dt)
        valid = rng.random(n) > 0.1
        qs = (0.0, 0.1, 0.25, 0.5, 0.77, 0.9, 1.0)
        rep, out, ok = orc.group_quantiles([host_col(k)], host_col(v, valid), qs, INTERP[interp])
        for g, r in enumerate(rep):
            x = np.sort(v[(k == k[r]) & valid].astype(np.float64))
            assert ok[g].all() == (len(x) > 0)
            if not len(x):
                continue
            for j, q in enumerate(qs):
                if interp == "MIDPOINT":
                    pos = q * (len(x) - 1)
                    a, b = x[int(math.floor(pos))], x[int(math.ceil(pos))]
                    want = float((Fraction(float(a)) + Fraction(float(b))) / 2)
                else:
                    want = np.quantile(x, q, method=interp.lower())
                assert math.isclose(out[g, j], want, rel_tol=1e-15, abs_tol=1e-300), (interp, dt, g, q, out[g, j], want)


def test_oracle_quantiles_nan_sorts_last():
    """cudf orders NaN after every number: a group {1, 2, NaN} has median 2,
    {NaN, NaN, 5} median NaN, and LOWER of the 0.0 quantile is the minimum."""
    k = host_col(np.array([1, 1, 1, 2, 2, 2], np.int32))
    v = host_col(np.array([np.nan, 2.0, 1.0, np.nan, 5.0, np.nan]))
    rep, out, ok = orc.group_quantiles([k], v, (0.0, 0.5, 1.0), orc.LOWER)
    assert out[0].tolist()[:2] == [1.0, 2.0] and math.isnan(out[0][2])
    assert out[1][0] == 5.0 and math.isnan(out[1][1]) and math.isnan(out[1][2])


def test_oracle_sort_kinds_thread_safe():
    """The sharded parity harness calls the oracle from several threads: the
    quantile and NUNIQUE sorts must not share state between concurrent calls
    (each call's results equal its serial result)."""
    import concurrent.futures as cf
    rng = np.random.default_rng(12)
    shards = []
    for i in range(6):
        n = 60_000 + 10_000 * i
        k = orc.HostColumn.from_numpy(rng.integers(0, 6_000, n).astype(np.int64))
        v = orc.HostColumn.from_numpy(rng.standard_normal(n))
        w = orc.HostColumn.from_numpy(rng.integers(0, 40, n).astype(np.int32))
        shards.append((k, v, w))

    def run(kvw):
        k, v, w = kvw
        q = orc.group_quantiles([k], v, (0.1, 0.5, 0.9), orc.NEAREST)
        g = orc.groupby([k], [(w, [(AGG["NUNIQUE"], 0)])])
        return q, g

    serial = [run(s) for s in shards]
    for _ in range(3):
        with cf.ThreadPoolExecutor(6) as ex:
            conc = list(ex.map(run, shards))
        for (q1, g1), (q2, g2) in zip(serial, conc):
            assert np.array_equal(q1[1], q2[1]) and np.array_equal(q1[0], q2[0])
            assert np.array_equal(g1[1][0][0][0], g2[1][0][0][0])


def test_string_key_fingerprint_ignores_position():
    """The parity harness orders string keys by a fingerprint: equal strings
    must get equal values wherever they sit (before trailing empty strings,
    first, last)."""
    from tests.gbutil import StrKeys
    rng = np.random.default_rng(3)
    words = [b"", b"a", b"xxxxxxxx", b"y" * 65, b"q" * 9]
    for _ in range(50):
        strs = [words[i] for i in rng.integers(0, len(words), 7)]
        perm = rng.permutation(7)
        fa = StrKeys.from_list(strs).fingerprint()
        fb = StrKeys.from_list([strs[i] for i in perm]).fingerprint()
        assert np.array_equal(fa[perm], fb)

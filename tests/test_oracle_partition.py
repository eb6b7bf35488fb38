"""Pins the partitioning / concatenation oracle (oracle/partition_oracle.py).

- hash_combine_np against the C oracle's row hash (itself pinned by the
  reference's golden row-hash vectors), recombining per-element MurmurHash3;
- IdentityHash partitions against a closed form in Python integers;
- the stable regrouping against a brute-force loop;
- the properties the reference's hash_partition tests assert
  (cpp/tests/partitioning/hash_partition_test.cpp);
- planted mistakes (unstable order, wrong modulus) are caught.
"""
import numpy as np
import pytest

from oracle import oracle as orc
from oracle import partition_oracle as po


def _mm(x: np.ndarray, seed: int) -> np.ndarray:
    return np.array([orc.murmur3_x86_32(v.tobytes(), seed) for v in x], dtype=np.uint32)


@pytest.mark.parametrize("seed", [0, 12345, 0xFFFFFFFF])
def test_hash_combine_matches_row_hash(seed):
    rng = np.random.default_rng(seed & 0xFFFF)
    a = rng.integers(-2**31, 2**31, 200, dtype=np.int64).astype(np.int32)
    b = rng.integers(-2**62, 2**62, 200, dtype=np.int64)
    want = orc.hash_rows([orc.HostColumn.from_numpy(a), orc.HostColumn.from_numpy(b)], seed)
    h = np.full(200, seed, dtype=np.uint32)
    h = po.hash_combine_np(h, _mm(a, seed))
    h = po.hash_combine_np(h, _mm(b, seed))
    np.testing.assert_array_equal(h, want)


def _identity_closed_form(v: int, seed: int, num_partitions: int) -> int:
    e = v % 2**32
    lhs = seed
    h = lhs ^ ((e + 0x9E3779B9 + ((lhs << 6) % 2**32) + (lhs >> 2)) % 2**32)
    return h % num_partitions


@pytest.mark.parametrize("seed,num_partitions", [(0, 3), (7, 8), (0xDEADBEEF, 50)])
def test_identity_partitions_closed_form(seed, num_partitions):
    vals = np.array([1, 2, 3, -1, -2**31, 2**31 - 1, 0, 17], dtype=np.int32)
    order, offs = po.hash_partition([orc.HostColumn.from_numpy(vals)], [0], num_partitions, po.HASH_IDENTITY, seed)
    pid = [_identity_closed_form(int(v), seed, num_partitions) for v in vals]
    bf_order, bf_offs = po.brute_force_regroup(pid, num_partitions)
    assert list(order) == bf_order
    assert offs == bf_offs


def test_identity_float_conversion():
    x = np.array([1.9, -3.5, np.nan, 0.0, -0.0, 5e9, 4294967295.0, 7.0], dtype=np.float64)
    got = po.identity_elements(orc.HostColumn.from_numpy(x))
    assert list(got) == [1, 0, 0, 0, 0, 0xFFFFFFFF, 0xFFFFFFFF, 7]
    v = np.array([True, False, True])
    got = po.identity_elements(orc.HostColumn.from_numpy(v, np.array([True, True, False])))
    assert list(got) == [1, 0, 0xFFFFFFFF]


@pytest.mark.parametrize("num_partitions", [1, 2, 3, 16, 50])
def test_stable_regroup_brute_force(num_partitions):
    rng = np.random.default_rng(num_partitions)
    keys = rng.integers(0, 30, 300).astype(np.int64)
    col = orc.HostColumn.from_numpy(keys)
    order, offs = po.hash_partition([col], [0], num_partitions)
    pid = po.partition_numbers(orc.hash_rows([col], 0), num_partitions)
    bf_order, bf_offs = po.brute_force_regroup(list(pid), num_partitions)
    assert list(order) == bf_order and offs == bf_offs


def test_reference_test_properties():
    # MixedColumnTypes / ColumnsToHash (hash_partition_test.cpp:117-179)
    floats = orc.HostColumn.from_numpy(np.arange(1, 9, dtype=np.float32))
    ints = orc.HostColumn.from_numpy(np.arange(1, 9, dtype=np.int16))
    strs = orc.HostColumn.from_strings([b"a", b"bb", b"ccc", b"d", b"ee", b"fff", b"gg", b"h"])
    o1, f1 = po.hash_partition([floats, ints, strs], [0, 2], 3)
    o2, f2 = po.hash_partition([floats, ints, strs], [0, 2], 3)
    assert len(f1) == 3 and f1 == f2 and list(o1) == list(o2) and sorted(o1) == list(range(8))
    to_hash = orc.HostColumn.from_numpy(np.arange(1, 7, dtype=np.int32))
    a = orc.HostColumn.from_numpy(np.arange(7, 13, dtype=np.int32))
    b = orc.HostColumn.from_numpy(np.arange(13, 19, dtype=np.int32))
    oa, fa = po.hash_partition([to_hash, a], [0], 3)
    ob, fb = po.hash_partition([to_hash, b], [0], 3)
    assert fa == fb and list(oa) == list(ob)
    # empty cases (ZeroPartitions / ZeroRows / ZeroColumns)
    o, f = po.hash_partition([floats], [0], 0)
    assert len(o) == 0 and f == []
    o, f = po.hash_partition([orc.HostColumn.from_numpy(np.zeros(0, np.int32))], [0], 3)
    assert len(o) == 0 and f == [0, 0, 0]
    o, f = po.hash_partition([floats], [], 3)
    assert len(o) == 0 and f == [0, 0, 0]
    with pytest.raises(IndexError):
        po.hash_partition([floats], [-1], 3)


def test_partition_by_map_and_concatenate():
    pmap = np.array([2, 0, 1, 2, 0, 0], dtype=np.int32)
    order, offs = po.partition_by_map(pmap, 3)
    assert list(order) == [1, 4, 5, 2, 0, 3] and offs == [0, 3, 4, 6]
    a = orc.HostColumn.from_numpy(np.array([1, 2], np.int64), np.array([True, False]))
    b = orc.HostColumn.from_numpy(np.array([3], np.int64))
    c = po.concatenate([a, b])
    assert list(c.data) == [1, 2, 3] and list(c.valid) == [True, False, True]
    s = po.concatenate([orc.HostColumn.from_strings([b"x", b""]), orc.HostColumn.from_strings([b"yz"])])
    assert [s.value(i) for i in range(3)] == [b"x", b"", b"yz"]


def test_planted_mistakes_are_caught():
    rng = np.random.default_rng(5)
    col = orc.HostColumn.from_numpy(rng.integers(0, 50, 400).astype(np.int32))
    pid = po.partition_numbers(orc.hash_rows([col], 0), 7)
    bf_order, _ = po.brute_force_regroup(list(pid), 7)
    unstable = np.argsort(pid, kind="quicksort")
    assert list(unstable) != bf_order  # an unstable regrouping reorders rows inside partitions
    wrong_mod = po.partition_numbers(orc.hash_rows([col], 0), 6)
    assert not np.array_equal(wrong_mod, pid)

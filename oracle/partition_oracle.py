"""CPU oracle for hash partitioning and concatenation -- test infrastructure only.

A numpy restatement of libcudf's ``hash_partition`` / ``partition``
(cpp/src/partitioning/partitioning.cu) and ``concatenate``
(cpp/src/copying/concatenate.cu).  Only ``tests/`` import it, to check the
product; the product never calls it.

Order inside a partition: the reference scatters rows with atomics
(partitioning.cu:151-156, 706-711), so its order inside a partition is
unspecified.  The library returns rows in input order inside each partition
(a stable regrouping); this oracle states that order as a stable sort by
partition number.

Pinning (tests/test_oracle_partition.py):
  - row hashes come from ``oracle.hash_rows`` (pinned by the reference's
    golden vectors); ``hash_combine_np`` is checked against it by recombining
    per-element MurmurHash3 values;
  - IdentityHash partitions are checked against a closed form written with
    Python integers;
  - the stable regrouping is checked against a brute-force loop that appends
    rows to per-partition lists, on small inputs;
  - the reference test suite's properties (cpp/tests/partitioning/
    hash_partition_test.cpp: offsets size, shape, determinism, equal offsets
    for equal hashed columns) are asserted on the oracle's own outputs.
"""
from __future__ import annotations

from typing import List, Optional, Sequence, Tuple

import numpy as np

from . import oracle as orc

NULL_HASH = np.uint32(0xFFFFFFFF)
HASH_IDENTITY, HASH_MURMUR3 = 0, 1


def hash_combine_np(lhs: np.ndarray, rhs: np.ndarray) -> np.ndarray:
    """boost hash_combine on uint32 arrays (hashing/detail/hashing.hpp:85-88)."""
    lhs = lhs.astype(np.uint64)
    rhs = rhs.astype(np.uint64)
    m = np.uint64(0xFFFFFFFF)
    t = (rhs + np.uint64(0x9E3779B9) + ((lhs << np.uint64(6)) & m) + (lhs >> np.uint64(2))) & m
    return (lhs ^ t).astype(np.uint32)


def identity_elements(col: orc.HostColumn) -> np.ndarray:
    """IdentityHash of each element (partitioning.cu:738-758): the value cast
    to uint32 (integers modulo 2^32; floats truncated, saturating, NaN and
    negatives to 0); nulls hash to 0xFFFFFFFF (the row hasher's null hash)."""
    d = col.data
    if d.dtype.kind in "iu":
        out = (d.astype(np.int64) if d.dtype != np.uint64 else d).astype(np.uint64) & np.uint64(0xFFFFFFFF)
        out = out.astype(np.uint32)
    elif d.dtype.kind == "f":
        x = d.astype(np.float64)
        out = np.zeros(len(x), dtype=np.uint32)
        pos = x > 0
        big = x >= 4294967295.0
        out[pos & ~big] = np.trunc(x[pos & ~big]).astype(np.uint64).astype(np.uint32)
        out[big] = np.uint32(0xFFFFFFFF)
    else:
        raise TypeError("IdentityHash does not support this data type")
    if col.valid is not None:
        out = np.where(col.valid, out, NULL_HASH).astype(np.uint32)
    return out


def row_hashes(cols: Sequence[orc.HostColumn], hash_function: int, seed: int) -> np.ndarray:
    if hash_function == HASH_MURMUR3:
        return orc.hash_rows(list(cols), seed)
    n = cols[0].size
    h = np.full(n, np.uint32(seed & 0xFFFFFFFF), dtype=np.uint32)
    for c in cols:
        h = hash_combine_np(h, identity_elements(c))
    return h


def partition_numbers(hashes: np.ndarray, num_partitions: int) -> np.ndarray:
    """Partition of each row: hash % num_partitions (partitioning.cu:59-95;
    the bitwise form for powers of two gives the same value)."""
    return (hashes.astype(np.uint64) % np.uint64(num_partitions)).astype(np.int64)


def regroup(pid: np.ndarray, num_partitions: int) -> Tuple[np.ndarray, List[int]]:
    """(gather map, first row of every partition) of a stable regrouping."""
    order = np.argsort(pid, kind="stable")
    counts = np.bincount(pid, minlength=num_partitions)[:num_partitions]
    offsets = np.concatenate([[0], np.cumsum(counts)[:-1]]).astype(np.int64)
    return order, [int(o) for o in offsets]


def hash_partition(cols: Sequence[orc.HostColumn], columns_to_hash: Sequence[int], num_partitions: int,
                   hash_function: int = HASH_MURMUR3, seed: int = 0) -> Tuple[np.ndarray, List[int]]:
    """Gather map and offsets of cudf::hash_partition (partitioning.cu:759-830)."""
    for c in columns_to_hash:
        if c < 0 or c >= len(cols):
            raise IndexError(c)
    n = cols[0].size if cols else 0
    if num_partitions <= 0 or n == 0 or len(columns_to_hash) == 0:
        return np.zeros(0, dtype=np.int64), [0] * max(0, num_partitions)
    h = row_hashes([cols[c] for c in columns_to_hash], hash_function, seed)
    return regroup(partition_numbers(h, num_partitions), num_partitions)


def partition_by_map(pmap: np.ndarray, num_partitions: int) -> Tuple[np.ndarray, List[int]]:
    """Gather map and offsets (num_partitions + 1) of cudf::partition (partitioning.cu:782-802)."""
    n = len(pmap)
    if num_partitions == 0 or n == 0:
        return np.zeros(0, dtype=np.int64), [0] * (num_partitions + 1)
    order, offs = regroup(np.asarray(pmap, dtype=np.int64), num_partitions)
    return order, offs + [n]


def gather(col: orc.HostColumn, order: np.ndarray) -> orc.HostColumn:
    valid = None if col.valid is None else col.valid[order]
    if col.type_id == orc.STRING:
        strs = [bytes(col.data[col.offsets[i]:col.offsets[i + 1]]) for i in order]
        return orc.HostColumn.from_strings(strs, valid)
    return orc.HostColumn(col.type_id, col.data[order], valid)


def concatenate(cols: Sequence[orc.HostColumn]) -> orc.HostColumn:
    """One column of cudf::concatenate (concatenate.cu): pieces stacked; the
    result has validity when some piece has nulls."""
    any_nulls = any(c.valid is not None and not c.valid.all() for c in cols)
    valid = None
    if any_nulls:
        valid = np.concatenate([c.valid if c.valid is not None else np.ones(c.size, bool) for c in cols])
    if cols[0].type_id == orc.STRING:
        strs = [bytes(c.data[c.offsets[i]:c.offsets[i + 1]]) for c in cols for i in range(c.size)]
        return orc.HostColumn.from_strings(strs, valid)
    return orc.HostColumn(cols[0].type_id, np.concatenate([c.data for c in cols]), valid)


def brute_force_regroup(pid: Sequence[int], num_partitions: int) -> Tuple[List[int], List[int]]:
    """Rows appended to per-partition lists in input order (small inputs)."""
    lists: List[List[int]] = [[] for _ in range(num_partitions)]
    for r, p in enumerate(pid):
        lists[int(p)].append(r)
    order, offsets, at = [], [], 0
    for lst in lists:
        offsets.append(at)
        order.extend(lst)
        at += len(lst)
    return order, offsets

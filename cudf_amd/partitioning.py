"""Hash partitioning and concatenation of device tables.

Mirrors ``cudf::hash_partition`` / ``cudf::partition``
(cpp/include/cudf/partitioning.hpp:79-112, python/cudf/cudf/_lib/partitioning.pyx)
and ``cudf::concatenate`` (cpp/include/cudf/concatenate.hpp:95-98,
python/cudf/cudf/_lib/pylibcudf/concatenate.pyx) over the C ABI
(``mig_hash_partition``, ``mig_partition``, ``mig_concatenate``).
"""
from __future__ import annotations

import ctypes
from enum import IntEnum
from typing import List, Optional, Sequence, Tuple

import torch

from . import _lib
from .column import Column, Table, TorchMemoryResource
from .groupby import _stream_handle, _wrap


class HashId(IntEnum):
    """cudf::hash_id (cpp/include/cudf/partitioning.hpp:41-44)."""
    HASH_IDENTITY = 0
    HASH_MURMUR3 = 1


def _take_table(handle: ctypes.c_void_p, mr: TorchMemoryResource) -> Table:
    L = _lib.lib()
    try:
        out = _lib.OutColumn()
        cols = []
        for i in range(L.mig_table_num_columns(handle)):
            _lib.check(L.mig_table_release_column(handle, i, ctypes.byref(out)))
            cols.append(_wrap(out, mr))
    finally:
        L.mig_table_free(handle)
    return Table(cols)


def hash_partition(input: Table, columns_to_hash: Sequence[int], num_partitions: int,
                   hash_function: HashId = HashId.HASH_MURMUR3, seed: int = 0,
                   stream: Optional[torch.cuda.Stream] = None) -> Tuple[Table, List[int]]:
    """Rows regrouped by ``row_hash % num_partitions`` and the first row of
    each partition.  Rows keep their input order inside a partition."""
    idx = (ctypes.c_int32 * max(1, len(columns_to_hash)))(*[int(c) for c in columns_to_hash])
    offs = (ctypes.c_int32 * max(1, int(num_partitions)))()
    mr = TorchMemoryResource()
    handle = ctypes.c_void_p()
    _lib.check(_lib.lib().mig_hash_partition(input.views(), input.num_columns(), idx, len(columns_to_hash),
                                             int(num_partitions), int(hash_function), int(seed) & 0xFFFFFFFF,
                                             _stream_handle(stream), mr.pointer(), ctypes.byref(handle), offs))
    return _take_table(handle, mr), [int(offs[i]) for i in range(max(0, int(num_partitions)))]


def partition(t: Table, partition_map: Column, num_partitions: int,
              stream: Optional[torch.cuda.Stream] = None) -> Tuple[Table, List[int]]:
    """Rows regrouped by ``partition_map``; offsets has ``num_partitions + 1``
    entries, the last being the row count."""
    offs = (ctypes.c_int32 * (max(0, int(num_partitions)) + 1))()
    mr = TorchMemoryResource()
    handle = ctypes.c_void_p()
    pm = partition_map.view()
    _lib.check(_lib.lib().mig_partition(t.views(), t.num_columns(), ctypes.byref(pm), int(num_partitions),
                                        _stream_handle(stream), mr.pointer(), ctypes.byref(handle), offs))
    return _take_table(handle, mr), [int(offs[i]) for i in range(max(0, int(num_partitions)) + 1)]


def concatenate(tables: Sequence[Table], stream: Optional[torch.cuda.Stream] = None) -> Table:
    """Tables stacked vertically (same column types in every table)."""
    tables = list(tables)
    nt = len(tables)
    nc = tables[0].num_columns() if nt else 0
    if any(t.num_columns() != nc for t in tables):
        raise _lib.LogicError("Mismatch in table columns to concatenate.")
    views = (_lib.ColumnView * max(1, nt * nc))()
    for ti, t in enumerate(tables):
        for ci, c in enumerate(t.columns()):
            views[ti * nc + ci] = c.view()
    mr = TorchMemoryResource()
    handle = ctypes.c_void_p()
    _lib.check(_lib.lib().mig_concatenate(views, nt, nc, _stream_handle(stream), mr.pointer(),
                                          ctypes.byref(handle)))
    return _take_table(handle, mr)


def split(table: Table, offsets: Sequence[int]) -> List[Table]:
    """Views of the partitions of a partitioned table (no copies)."""
    n = table.num_rows()
    bounds = list(offsets) + [n]
    out = []
    for p in range(len(offsets)):
        b, e = bounds[p], bounds[p + 1]
        out.append(Table([c.slice(b, e - b) for c in table.columns()]))
    return out

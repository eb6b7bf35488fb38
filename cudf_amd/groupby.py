"""Hash group-by, shaped like pylibcudf.groupby
(python/cudf/cudf/_lib/pylibcudf/groupby.pyx:29-181): ``GroupBy(keys,
null_handling).aggregate([GroupByRequest(values, aggs), ...])`` returns the
unique keys table and one results table per request.

The work happens in the gfx950 library behind the C ABI
(``mig_groupby_aggregate``); this module only marshals views and takes
ownership of the result buffers.
"""
from __future__ import annotations

import ctypes
from typing import List, Optional, Tuple

import torch

from . import _lib
from .aggregation import Aggregation
from .column import Column, Table, TorchMemoryResource
from .types import NullPolicy, Sorted, TypeId


class GroupByRequest:
    """cudf::groupby::aggregation_request (cudf/groupby.hpp:61-64)."""

    def __init__(self, values: Column, aggregations: List[Aggregation]):
        self._values = values
        self._aggregations = list(aggregations)

    @property
    def values(self) -> Column:
        return self._values

    @property
    def aggregations(self) -> List[Aggregation]:
        return list(self._aggregations)


def _stream_handle(stream: Optional[torch.cuda.Stream]):
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


def _wrap(out: _lib.OutColumn, mr: TorchMemoryResource) -> Column:
    data = mr.take(out.data)
    mask = mr.take(out.null_mask)
    offsets = mr.take(out.offsets)
    if mask is not None:
        mask = mask.view(torch.int32)
    if offsets is not None:
        offsets = offsets.view(torch.int32)
    return Column(out.type_id, out.size, data, mask, out.null_count, 0, offsets, out.scale)


class GroupBy:
    """cudf::groupby::groupby (cudf/groupby.hpp:128-132); hash path only."""

    def __init__(self, keys: Table, null_handling: NullPolicy = NullPolicy.EXCLUDE,
                 keys_are_sorted: Sorted = Sorted.NO, column_order=None, null_precedence=None):
        self._keys = keys
        self._null_handling = NullPolicy(null_handling)
        # Sorted keys take the sort-based path (groupby.cu:66-84): groups are
        # the runs of equal adjacent key rows, returned in input order.
        self._keys_are_sorted = Sorted(keys_are_sorted)

    def aggregate(self, requests: List[GroupByRequest],
                  stream: Optional[torch.cuda.Stream] = None) -> Tuple[Table, List[Table]]:
        """cudf::groupby::groupby::aggregate (cudf/groupby.hpp:187-199)."""
        keep = []
        keys = self._keys.views()
        nreq = len(requests)
        reqs = (_lib.AggregationRequest * max(1, nreq))()
        for i, r in enumerate(requests):
            aggs = (_lib.Aggregation * max(1, len(r.aggregations)))()
            for j, a in enumerate(r.aggregations):
                aggs[j] = a.to_c(keep)
            keep.append(aggs)
            reqs[i].values = r.values.view()
            reqs[i].aggregations = ctypes.cast(aggs, ctypes.POINTER(_lib.Aggregation))
            reqs[i].num_aggregations = len(r.aggregations)
        mr = TorchMemoryResource()
        handle = ctypes.c_void_p()
        L = _lib.lib()
        entry = L.mig_groupby_aggregate_sorted if self._keys_are_sorted == Sorted.YES else L.mig_groupby_aggregate
        _lib.check(entry(keys, self._keys.num_columns(), int(self._null_handling), reqs, nreq,
                         _stream_handle(stream), mr.pointer(), ctypes.byref(handle)))
        try:
            out = _lib.OutColumn()
            key_cols = []
            for i in range(L.mig_result_num_keys(handle)):
                _lib.check(L.mig_result_release_key(handle, i, ctypes.byref(out)))
                key_cols.append(_wrap(out, mr))
            results = []
            for r in range(L.mig_result_num_requests(handle)):
                cols = []
                for a in range(L.mig_result_num_aggregations(handle, r)):
                    _lib.check(L.mig_result_release_column(handle, r, a, ctypes.byref(out)))
                    cols.append(_wrap(out, mr))
                results.append(Table(cols))
        finally:
            L.mig_result_free(handle)
        return Table(key_cols), results


def murmurhash3_x86_32(table: Table, seed: int = 0,
                       stream: Optional[torch.cuda.Stream] = None) -> Column:
    """cudf::hashing::murmurhash3_x86_32 (cpp/include/cudf/hashing.hpp:61)."""
    n = table.num_rows()
    out = torch.empty(max(1, n) * 4, dtype=torch.uint8, device=torch.device("cuda", torch.cuda.current_device()))
    _lib.check(_lib.lib().mig_hash_murmurhash3_x86_32(table.views(), table.num_columns(), seed,
                                                      ctypes.c_void_p(out.data_ptr()),
                                                      _stream_handle(stream)))
    return Column(TypeId.UINT32, n, out)

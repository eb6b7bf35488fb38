"""Arrow ingest and egress for device columns.

The column layout of this library is Arrow's (reference
``cpp/include/cudf/column/column_view.hpp:240-247``; ``cpp/src/interop/``):
fixed-width values, LSB-first validity bits, strings as int32/int64 offsets
plus chars.  Arrow booleans are bit-packed while BOOL8 holds a byte per row,
so those are widened; timestamps and durations keep their integer storage.
Host buffers are copied to the device once (``from_arrow``) and back
(``to_arrow``).
"""
from __future__ import annotations

from typing import List

import numpy as np

from .column import Column, Table
from .types import TypeId, numpy_dtype

_FROM_ARROW = {
    "int8": TypeId.INT8, "int16": TypeId.INT16, "int32": TypeId.INT32, "int64": TypeId.INT64,
    "uint8": TypeId.UINT8, "uint16": TypeId.UINT16, "uint32": TypeId.UINT32, "uint64": TypeId.UINT64,
    "float": TypeId.FLOAT32, "double": TypeId.FLOAT64, "bool": TypeId.BOOL8,
    "date32[day]": TypeId.TIMESTAMP_DAYS,
    "timestamp[s]": TypeId.TIMESTAMP_SECONDS, "timestamp[ms]": TypeId.TIMESTAMP_MILLISECONDS,
    "timestamp[us]": TypeId.TIMESTAMP_MICROSECONDS, "timestamp[ns]": TypeId.TIMESTAMP_NANOSECONDS,
    "duration[s]": TypeId.DURATION_SECONDS, "duration[ms]": TypeId.DURATION_MILLISECONDS,
    "duration[us]": TypeId.DURATION_MICROSECONDS, "duration[ns]": TypeId.DURATION_NANOSECONDS,
}


def arrow_type_id(arrow_type) -> TypeId:
    """The type id a pyarrow type maps to (``DataTypeError``-like ValueError
    when it has no counterpart here)."""
    import pyarrow as pa
    if pa.types.is_string(arrow_type) or pa.types.is_large_string(arrow_type):
        return TypeId.STRING
    key = str(arrow_type)
    if key not in _FROM_ARROW:
        raise ValueError(f"unsupported Arrow type {arrow_type}")
    return _FROM_ARROW[key]


def from_arrow(arr) -> Column:
    """A device column holding a pyarrow Array (or ChunkedArray)."""
    import pyarrow as pa
    if isinstance(arr, pa.ChunkedArray):
        arr = arr.combine_chunks()
    tid = arrow_type_id(arr.type)
    valid = None
    if arr.null_count > 0:
        valid = np.asarray(arr.is_valid().to_numpy(zero_copy_only=False), dtype=bool)
    if tid == TypeId.STRING:
        large = pa.types.is_large_string(arr.type)
        vals = [b"" if v is None else v.encode() for v in arr.to_pylist()]
        return Column.from_strings(vals, valid, large=large)
    if tid == TypeId.BOOL8:
        values = np.asarray(arr.fill_null(False).to_numpy(zero_copy_only=False), dtype=np.uint8)
    else:
        if not (pa.types.is_integer(arr.type) or pa.types.is_floating(arr.type)):
            # dates, timestamps, durations: their integer storage
            arr = arr.cast(pa.int32() if numpy_dtype(tid).itemsize == 4 else pa.int64())
        filled = arr.fill_null(0) if arr.null_count else arr
        values = np.asarray(filled.to_numpy(zero_copy_only=False)).astype(numpy_dtype(tid), copy=False)
    return Column.from_numpy(values, valid, type_id=tid)


def _arrow_type(tid: TypeId):
    import pyarrow as pa
    simple = {TypeId.INT8: pa.int8(), TypeId.INT16: pa.int16(), TypeId.INT32: pa.int32(),
              TypeId.INT64: pa.int64(), TypeId.UINT8: pa.uint8(), TypeId.UINT16: pa.uint16(),
              TypeId.UINT32: pa.uint32(), TypeId.UINT64: pa.uint64(), TypeId.FLOAT32: pa.float32(),
              TypeId.FLOAT64: pa.float64(), TypeId.TIMESTAMP_DAYS: pa.date32(),
              TypeId.TIMESTAMP_SECONDS: pa.timestamp("s"), TypeId.TIMESTAMP_MILLISECONDS: pa.timestamp("ms"),
              TypeId.TIMESTAMP_MICROSECONDS: pa.timestamp("us"), TypeId.TIMESTAMP_NANOSECONDS: pa.timestamp("ns"),
              TypeId.DURATION_SECONDS: pa.duration("s"), TypeId.DURATION_MILLISECONDS: pa.duration("ms"),
              TypeId.DURATION_MICROSECONDS: pa.duration("us"), TypeId.DURATION_NANOSECONDS: pa.duration("ns")}
    if tid not in simple:
        raise ValueError(f"type {TypeId(tid).name} has no Arrow counterpart here")
    return simple[tid]


def to_arrow(col: Column):
    """A pyarrow Array with the values of a device column."""
    import pyarrow as pa
    values, valid = col.to_numpy()
    mask = None if valid is None else ~np.asarray(valid, dtype=bool)
    if col.type_id == TypeId.STRING:
        return pa.array([v.decode() for v in values], type=pa.string(), mask=mask)
    if col.type_id == TypeId.BOOL8:
        return pa.array(np.asarray(values) != 0, type=pa.bool_(), mask=mask)
    target = _arrow_type(col.type_id)
    if pa.types.is_integer(target) or pa.types.is_floating(target):
        return pa.array(values, type=target, mask=mask)
    # dates, timestamps, durations from their integer storage
    storage = pa.array(values, type=pa.int32() if values.dtype.itemsize == 4 else pa.int64(), mask=mask)
    return storage.cast(target)


def table_from_arrow(tbl) -> Table:
    """Device table of a pyarrow Table's columns (in order)."""
    return Table([from_arrow(tbl.column(i)) for i in range(tbl.num_columns)])


def table_to_arrow(table: Table, names: List[str]):
    import pyarrow as pa
    return pa.table([to_arrow(c) for c in table.columns()], names=names)

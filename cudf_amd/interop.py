"""Arrow ingest and egress for device columns.

The column layout of this library is Arrow's (reference
``cpp/include/cudf/column/column_view.hpp:240-247``; ``cpp/src/interop/``):
fixed-width values, LSB-first validity bits, strings as int32/int64 offsets
plus chars.  Arrow booleans are bit-packed while BOOL8 holds a byte per row,
so those are widened; timestamps and durations keep their integer storage.
Arrow buffers are copied to the device whole (``from_arrow``) and built
back from the device buffers (``to_arrow``), with no per-element conversion.
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


def _device_bytes(buf, dev, nbytes=None):
    """One host-to-device copy of (the first `nbytes` bytes of) an Arrow
    buffer, padded to whole 32-bit words."""
    import torch
    host = np.frombuffer(buf, dtype=np.uint8) if buf is not None else np.zeros(0, np.uint8)
    if nbytes is not None:
        host = host[:nbytes]
    pad = (-len(host)) % 4
    if pad or len(host) == 0:
        host = np.concatenate([host, np.zeros(pad if len(host) else 4, np.uint8)])
    return torch.from_numpy(host).to(dev)


def from_arrow(arr, device=None) -> Column:
    """A device column holding a pyarrow Array (or ChunkedArray).

    The Arrow buffers are copied to the device whole, as cpp/src/interop/
    from_arrow_host.cu:173-200 copies them: values (or the offsets child and
    the characters of strings) and the validity bitmap, with the array's
    offset kept as the column's row offset.  Arrow booleans (one bit per
    value) are widened to BOOL8 bytes on the device.  `device` defaults to
    the current HIP device."""
    import pyarrow as pa
    import torch
    if isinstance(arr, pa.ChunkedArray):
        arr = arr.combine_chunks()
    tid = arrow_type_id(arr.type)
    n, off = len(arr), arr.offset
    if n > 2**31 - 1 or off + n > 2**31 - 1:
        raise OverflowError("Arrow array longer than cudf's size_type")
    dev = torch.device("cuda", torch.cuda.current_device()) if device is None else torch.device(device)
    bufs = arr.buffers()
    mask, nulls = None, 0
    if arr.null_count > 0 and bufs[0] is not None:
        mask = _device_bytes(bufs[0], dev, (off + n + 7) // 8).view(torch.int32)
        nulls = arr.null_count
    if tid == TypeId.STRING:
        large = pa.types.is_large_string(arr.type)
        ow = 8 if large else 4
        offsets = _device_bytes(bufs[1], dev, (off + n + 1) * ow).view(torch.int64 if large else torch.int32)
        chars = _device_bytes(bufs[2], dev)
        return Column(TypeId.STRING, n, chars, mask, nulls, off, offsets[: off + n + 1])
    if tid == TypeId.BOOL8:
        # rows [0, off + n) widened, so the column's row offset applies to
        # the bytes and to the (unchanged) validity bitmap alike
        bits = _device_bytes(bufs[1], dev, (off + n + 7) // 8)
        shifts = torch.arange(8, device=dev, dtype=torch.uint8)
        widened = ((bits[:, None] >> shifts) & 1).reshape(-1)[: max(off + n, 1)].contiguous()
        return Column(TypeId.BOOL8, n, widened, mask, nulls, off)
    w = np.dtype(numpy_dtype(tid)).itemsize
    data = _device_bytes(bufs[1], dev, (off + n) * w)
    return Column(tid, n, data, mask, nulls, off)


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


def _host_validity(col: Column):
    """The column's validity as an Arrow bitmap buffer starting at bit 0, or
    None when it has no nulls."""
    import pyarrow as pa
    if col.mask is None or col.null_count == 0:
        return None
    valid = col.validity()
    return pa.py_buffer(np.packbits(valid, bitorder="little").tobytes())


def to_arrow(col: Column):
    """A pyarrow Array with the values of a device column, built from its
    buffers (no per-element conversion)."""
    import pyarrow as pa
    n = col.size
    vbuf = _host_validity(col)
    nulls = col.null_count if vbuf is not None else 0
    if col.type_id == TypeId.STRING:
        offs = col.offsets.cpu().numpy().astype(np.int64)[col.offset: col.offset + n + 1]
        base = int(offs[0]) if n else 0
        last = int(offs[-1]) if n else base
        chars = col.data.cpu().numpy().view(np.uint8)[base: last] if col.data is not None else np.zeros(0, np.uint8)
        rel = offs - base if n else np.zeros(1, np.int64)
        if last - base >= 2**31:
            return pa.Array.from_buffers(pa.large_string(), n, [vbuf, pa.py_buffer(rel.tobytes()),
                                                                 pa.py_buffer(chars.tobytes())], nulls)
        return pa.Array.from_buffers(pa.string(), n, [vbuf, pa.py_buffer(rel.astype(np.int32).tobytes()),
                                                      pa.py_buffer(chars.tobytes())], nulls)
    values, _ = col.to_numpy()
    if col.type_id == TypeId.BOOL8:
        bits = pa.py_buffer(np.packbits(np.asarray(values) != 0, bitorder="little").tobytes())
        return pa.Array.from_buffers(pa.bool_(), n, [vbuf, bits], nulls)
    target = _arrow_type(col.type_id)
    return pa.Array.from_buffers(target, n, [vbuf, pa.py_buffer(np.ascontiguousarray(values).tobytes())], nulls)


def table_from_arrow(tbl) -> Table:
    """Device table of a pyarrow Table's columns (in order)."""
    return Table([from_arrow(tbl.column(i)) for i in range(tbl.num_columns)])


def table_to_arrow(table: Table, names: List[str]):
    import pyarrow as pa
    return pa.table([to_arrow(c) for c in table.columns()], names=names)

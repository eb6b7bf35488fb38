"""Device columns and tables (pylibcudf ``Column`` / ``Table`` shape).

Device memory is held in torch tensors (PyTorch is only the allocator and
stream provider here); the C ABI sees plain pointers through ``view()``.
Layout follows cudf: fixed-width elements, LSB-first uint32 validity words
(1 = valid), strings as int32/int64 offsets plus a chars buffer.
"""
from __future__ import annotations

import ctypes
from typing import List, Optional, Sequence

import numpy as np
import torch

from . import _lib
from .types import TypeId, numpy_dtype, type_id_of


_INT32_MAX = 2**31 - 1


def _device() -> torch.device:
    return torch.device("cuda", torch.cuda.current_device())


def pack_validity(valid: np.ndarray) -> np.ndarray:
    """Bool array -> LSB-first uint32 validity words."""
    valid = np.asarray(valid, dtype=bool)
    nbytes = ((len(valid) + 31) // 32) * 4
    packed = np.packbits(valid, bitorder="little")
    out = np.zeros(max(nbytes, 4), dtype=np.uint8)
    out[: len(packed)] = packed
    return out.view(np.uint32)


def unpack_validity(words: np.ndarray, size: int, offset: int = 0) -> np.ndarray:
    bits = np.unpackbits(np.asarray(words, dtype=np.uint32).view(np.uint8), bitorder="little")
    return bits[offset: offset + size].astype(bool)


class Column:
    """A device column: ``data`` (torch uint8 buffer), optional ``mask``
    (torch int32 words), ``offsets`` for strings, and a row ``offset``."""

    def __init__(self, type_id: int, size: int, data: Optional[torch.Tensor],
                 mask: Optional[torch.Tensor] = None, null_count: int = 0, offset: int = 0,
                 offsets: Optional[torch.Tensor] = None, scale: int = 0):
        self.type_id = TypeId(type_id)
        self.size = int(size)
        self.data = data
        self.mask = mask
        self.null_count = int(null_count) if mask is not None else 0
        self.offset = int(offset)
        self.offsets = offsets
        self.scale = int(scale)  # DECIMAL32 / DECIMAL64: value = data * 10**scale

    # -- construction -------------------------------------------------------
    @staticmethod
    def from_numpy(arr: np.ndarray, valid: Optional[np.ndarray] = None,
                   type_id: Optional[int] = None, scale: int = 0) -> "Column":
        """A device column from host values; DECIMAL32 / DECIMAL64 take the
        integer representation and the scale (value = rep * 10**scale)."""
        arr = np.ascontiguousarray(arr)
        tid = TypeId(type_id) if type_id is not None else type_id_of(arr.dtype)
        host = arr.astype(numpy_dtype(tid), copy=False)
        dev = _device()
        data = torch.from_numpy(host.view(np.uint8).copy() if host.size else np.zeros(1, np.uint8)).to(dev)
        mask, nulls = None, 0
        if valid is not None:
            valid = np.asarray(valid, dtype=bool)
            mask = torch.from_numpy(pack_validity(valid).view(np.int32).copy()).to(dev)
            nulls = int(len(valid) - valid.sum())
        return Column(tid, len(arr), data, mask, nulls, scale=scale)

    @staticmethod
    def from_strings(strs: Sequence, valid: Optional[np.ndarray] = None,
                     large: bool = False) -> "Column":
        bs = [s.encode() if isinstance(s, str) else bytes(s) for s in strs]
        lens = np.array([len(b) for b in bs], dtype=np.int64)
        offs = np.zeros(len(bs) + 1, dtype=np.int64 if large else np.int32)
        np.cumsum(lens, out=offs[1:])
        chars = np.frombuffer(b"".join(bs), dtype=np.uint8)
        dev = _device()
        data = torch.from_numpy(chars.copy() if chars.size else np.zeros(1, np.uint8)).to(dev)
        offsets = torch.from_numpy(offs).to(dev)
        mask, nulls = None, 0
        if valid is not None:
            valid = np.asarray(valid, dtype=bool)
            mask = torch.from_numpy(pack_validity(valid).view(np.int32).copy()).to(dev)
            nulls = int(len(valid) - valid.sum())
        return Column(TypeId.STRING, len(bs), data, mask, nulls, offsets=offsets)

    def slice(self, offset: int, size: int, null_count: int = -1) -> "Column":
        """A view of rows [offset, offset+size) sharing the buffers."""
        c = Column(self.type_id, size, self.data, self.mask, 0, self.offset + offset, self.offsets, self.scale)
        c.null_count = null_count if self.mask is not None else 0
        return c

    # -- C ABI ----------------------------------------------------------------
    def view(self) -> _lib.ColumnView:
        """The C ABI view.  Its size and offset are cudf ``size_type`` (int32):
        values outside [0, 2^31) raise instead of wrapping."""
        for name, x in (("size", self.size), ("offset", self.offset)):
            if not 0 <= x <= _INT32_MAX:
                raise OverflowError(f"column {name} {x} is outside [0, 2^31 - 1] (cudf size_type)")
        v = _lib.ColumnView()
        v.type_id = int(self.type_id)
        v.scale = self.scale
        v.size = self.size
        v.offset = self.offset
        v.data = self.data.data_ptr() if self.data is not None else None
        v.null_mask = self.mask.data_ptr() if self.mask is not None else None
        v.null_count = self.null_count
        if self.type_id == TypeId.STRING:
            v.offsets_type_id = int(TypeId.INT64 if self.offsets.dtype == torch.int64 else TypeId.INT32)
            v.offsets = self.offsets.data_ptr()
        return v

    # -- host copies ----------------------------------------------------------
    def validity(self) -> Optional[np.ndarray]:
        if self.mask is None:
            return None
        return unpack_validity(self.mask.cpu().numpy().view(np.uint32), self.size, self.offset)

    def to_numpy(self):
        """(values, valid) on the host; strings come back as a list of bytes."""
        valid = self.validity()
        if self.type_id == TypeId.STRING:
            offs = self.offsets.cpu().numpy().astype(np.int64)[self.offset: self.offset + self.size + 1]
            chars = self.data.cpu().numpy().tobytes() if self.data is not None else b""
            vals = [chars[offs[i]:offs[i + 1]] for i in range(self.size)]
            return vals, valid
        dt = numpy_dtype(self.type_id)
        if self.size == 0 or self.data is None:
            return np.zeros(0, dtype=dt), valid
        raw = self.data.cpu().numpy().view(np.uint8)
        start = self.offset * dt.itemsize
        vals = raw[start: start + self.size * dt.itemsize].view(dt).copy()
        return vals, valid

    def __len__(self) -> int:
        return self.size

    def __repr__(self) -> str:
        return f"Column({self.type_id.name}, size={self.size}, nulls={self.null_count})"


class Table:
    """An ordered list of equally long columns."""

    def __init__(self, columns: List[Column]):
        self._columns = list(columns)

    def columns(self) -> List[Column]:
        return list(self._columns)

    def num_columns(self) -> int:
        return len(self._columns)

    def num_rows(self) -> int:
        return self._columns[0].size if self._columns else 0

    def views(self):
        arr = (_lib.ColumnView * max(1, len(self._columns)))()
        for i, c in enumerate(self._columns):
            arr[i] = c.view()
        return arr


class TorchMemoryResource:
    """Device memory resource for the C ABI that allocates torch tensors, so
    result buffers are owned by torch once released from the result handle."""

    def __init__(self):
        self.live = {}
        self._alloc = _lib.ALLOC_FN(self._allocate)
        self._dealloc = _lib.DEALLOC_FN(self._deallocate)
        self.struct = _lib.MemoryResource(None, self._alloc, self._dealloc)

    def _allocate(self, ctx, nbytes, stream):
        try:
            t = torch.empty(int(nbytes), dtype=torch.uint8, device=_device())
        except Exception:  # noqa: BLE001 - reported to the library as a null pointer
            return None
        self.live[t.data_ptr()] = t
        return t.data_ptr()

    def _deallocate(self, ctx, ptr, nbytes, stream):
        self.live.pop(ptr, None)

    def take(self, ptr) -> Optional[torch.Tensor]:
        if not ptr:
            return None
        return self.live.pop(ptr)

    def pointer(self):
        return ctypes.byref(self.struct)

"""CPU oracle for the hash group-by path -- test infrastructure only.

Python side of ``groupby_oracle.c`` (a plain-C restatement of libcudf's hash
group-by; every function there cites the reference file:line it follows).
Only ``tests/``, the smoke check and the CPU leg of ``bench.py`` import this
module, and only to check or time the product; the product never calls it.

Columns are given as ``HostColumn`` objects holding numpy arrays.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from dataclasses import dataclass
from typing import Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

# cudf::type_id numbering (cudf/types.hpp:203-235)
INT8, INT16, INT32, INT64 = 1, 2, 3, 4
UINT8, UINT16, UINT32, UINT64 = 5, 6, 7, 8
FLOAT32, FLOAT64, BOOL8 = 9, 10, 11
STRING = 23

# cudf::aggregation::Kind numbering (cudf/aggregation.hpp:86-122)
SUM, PRODUCT, MIN, MAX, COUNT_VALID, COUNT_ALL = 0, 1, 2, 3, 4, 5
SUM_OF_SQUARES, MEAN, VARIANCE, STD = 8, 9, 11, 12
MEDIAN, QUANTILE = 13, 14  # QUANTILE's second element is (quantiles, interpolation)
ARGMAX, ARGMIN = 15, 16
LINEAR, LOWER, HIGHER, MIDPOINT, NEAREST = 0, 1, 2, 3, 4  # cudf::interpolation
NUNIQUE = 17  # the request's second element is the null policy (0 EXCLUDE, 1 INCLUDE)
NTH_ELEMENT = 18  # the request's second element is 2 * n + (1 if nulls are included)

NP_OF_TYPE = {
    INT8: np.int8, INT16: np.int16, INT32: np.int32, INT64: np.int64,
    UINT8: np.uint8, UINT16: np.uint16, UINT32: np.uint32, UINT64: np.uint64,
    FLOAT32: np.float32, FLOAT64: np.float64, BOOL8: np.uint8,
}
TYPE_OF_NP = {np.dtype(v): k for k, v in NP_OF_TYPE.items() if k != BOOL8}


@dataclass
class HostColumn:
    """A host column: fixed-width ``data`` array or, for strings, ``offsets``
    (int64, size+1) plus ``data`` (uint8 chars); ``valid`` is a bool array or
    None (no nulls)."""

    type_id: int
    data: np.ndarray
    valid: Optional[np.ndarray] = None
    offsets: Optional[np.ndarray] = None

    @property
    def size(self) -> int:
        return len(self.offsets) - 1 if self.type_id == STRING else len(self.data)

    @staticmethod
    def from_numpy(arr: np.ndarray, valid: Optional[np.ndarray] = None) -> "HostColumn":
        arr = np.ascontiguousarray(arr)
        if arr.dtype == np.bool_:
            return HostColumn(BOOL8, arr.astype(np.uint8), valid)
        return HostColumn(TYPE_OF_NP[arr.dtype], arr, valid)

    @staticmethod
    def from_strings(strs: Sequence[bytes], valid: Optional[np.ndarray] = None) -> "HostColumn":
        lens = np.array([len(s) for s in strs], dtype=np.int64)
        offsets = np.zeros(len(strs) + 1, dtype=np.int64)
        np.cumsum(lens, out=offsets[1:])
        chars = np.frombuffer(b"".join(strs), dtype=np.uint8).copy()
        if len(chars) == 0:
            chars = np.zeros(1, dtype=np.uint8)
        return HostColumn(STRING, chars, valid, offsets)

    def value(self, i: int):
        if self.valid is not None and not self.valid[i]:
            return None
        if self.type_id == STRING:
            return bytes(self.data[self.offsets[i]:self.offsets[i + 1]])
        return self.data[i]


class _OrcCol(ctypes.Structure):
    _fields_ = [
        ("type_id", ctypes.c_int32),
        ("size", ctypes.c_int32),
        ("data", ctypes.c_void_p),
        ("valid", ctypes.c_void_p),
        ("offsets", ctypes.c_void_p),
    ]


_lib = None


def build() -> str:
    """Compile liboracle.so with the Makefile beside this file."""
    subprocess.run(["make", "-s", "-C", _HERE, "liboracle.so"], check=True)
    return _LIB_PATH


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        lib = ctypes.CDLL(_LIB_PATH)
        lib.orc_murmur3_x86_32.restype = ctypes.c_uint32
        lib.orc_murmur3_x86_32.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_uint32]
        lib.orc_hash_rows.restype = None
        lib.orc_hash_rows.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_uint32,
                                      ctypes.c_uint32, ctypes.c_void_p]
        lib.orc_target_type.restype = ctypes.c_int32
        lib.orc_target_type.argtypes = [ctypes.c_int32, ctypes.c_int32]
        lib.orc_group_quantiles.restype = ctypes.c_int64
        lib.orc_group_quantiles.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_void_p,
                                            ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_void_p,
                                            ctypes.c_void_p, ctypes.c_void_p]
        lib.orc_groupby.restype = ctypes.c_int64
        lib.orc_groupby.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32,
                                    ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                    ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p,
                                    ctypes.c_void_p, ctypes.c_void_p]
        _lib = lib
    return _lib


def _ptr(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data


def _cols(cols: Sequence[HostColumn]):
    keep = []
    arr = (_OrcCol * max(1, len(cols)))()
    for i, c in enumerate(cols):
        valid = None if c.valid is None else np.ascontiguousarray(c.valid, dtype=np.uint8)
        offs = None if c.offsets is None else np.ascontiguousarray(c.offsets, dtype=np.int64)
        data = np.ascontiguousarray(c.data)
        keep += [valid, offs, data]
        arr[i] = _OrcCol(c.type_id, c.size, _ptr(data), _ptr(valid), _ptr(offs))
    return arr, keep


def murmur3_x86_32(data: bytes, seed: int = 0) -> int:
    buf = np.frombuffer(data, dtype=np.uint8) if data else np.zeros(1, np.uint8)
    return int(_load().orc_murmur3_x86_32(buf.ctypes.data, len(data), seed))


def hash_rows(cols: Sequence[HostColumn], seed: int = 0, elem_seed: Optional[int] = None) -> np.ndarray:
    """Row hashes as libcudf's murmurhash3_x86_32(table, seed); ``elem_seed``
    (default = ``seed``) is the seed given to each element hash."""
    n = cols[0].size
    out = np.zeros(max(1, n), dtype=np.uint32)
    arr, keep = _cols(cols)
    _load().orc_hash_rows(ctypes.addressof(arr), len(cols), seed,
                          seed if elem_seed is None else elem_seed, out.ctypes.data)
    return out[:n]


def target_type(src: int, kind: int) -> int:
    return int(_load().orc_target_type(src, kind))


def groupby(keys: Sequence[HostColumn], requests: Sequence[tuple], null_policy: int = 0):
    """Hash group-by in the reference's semantics.

    ``requests`` is a list of ``(HostColumn values, [(kind, ddof), ...])``.
    Returns ``(rep_rows, results)`` where ``rep_rows`` are the first row of each
    group (first-occurrence order) and ``results[r][a]`` is ``(array, valid)``.
    """
    values, vidx, kinds, ddofs = [], [], [], []
    for ri, (col, aggs) in enumerate(requests):
        values.append(col)
        for kind, ddof in aggs:
            vidx.append(ri)
            kinds.append(kind)
            ddofs.append(ddof)
    n = keys[0].size if keys else 0
    naggs = len(kinds)
    outs, valids = [], []
    for a in range(naggs):
        t = target_type(values[vidx[a]].type_id, kinds[a])
        if t < 0:
            raise TypeError(f"unsupported aggregation kind {kinds[a]} on type {values[vidx[a]].type_id}")
        outs.append(np.zeros(max(1, n), dtype=NP_OF_TYPE[t]))
        valids.append(np.zeros(max(1, n), dtype=np.uint8))
    rep = np.zeros(max(1, n), dtype=np.int64)
    karr, kk = _cols(keys)
    varr, vk = _cols(values)
    i32 = lambda xs: np.ascontiguousarray(np.array(xs if xs else [0], dtype=np.int32))
    vidx_a, kinds_a, ddofs_a = i32(vidx), i32(kinds), i32(ddofs)
    out_ptrs = (ctypes.c_void_p * max(1, naggs))(*[o.ctypes.data for o in outs])
    val_ptrs = (ctypes.c_void_p * max(1, naggs))(*[v.ctypes.data for v in valids])
    g = _load().orc_groupby(ctypes.addressof(karr), len(keys), null_policy,
                            ctypes.addressof(varr), vidx_a.ctypes.data, kinds_a.ctypes.data,
                            ddofs_a.ctypes.data, naggs, rep.ctypes.data,
                            ctypes.addressof(out_ptrs), ctypes.addressof(val_ptrs))
    if g < 0:
        raise TypeError("unsupported aggregation request")
    results, a = [], 0
    for ri, (col, aggs) in enumerate(requests):
        res = []
        for _ in aggs:
            res.append((outs[a][:g].copy(), valids[a][:g].astype(bool)))
            a += 1
        results.append(res)
    return rep[:g].copy(), results


def group_quantiles(keys: Sequence[HostColumn], values: HostColumn, quantiles, interpolation: int = LINEAR,
                    null_policy: int = 0):
    """QUANTILE of every group (first-occurrence group order): returns
    ``(rep_rows, out, valid)`` with ``out`` of shape (G, len(quantiles)),
    float64, group-major as the reference lays the result column out."""
    n = keys[0].size if keys else 0
    q = np.ascontiguousarray(np.asarray(quantiles, dtype=np.float64).reshape(-1))
    nq = len(q)
    rep = np.zeros(max(1, n), dtype=np.int64)
    out = np.zeros(max(1, n * max(nq, 1)), dtype=np.float64)
    valid = np.zeros(max(1, n * max(nq, 1)), dtype=np.uint8)
    karr, kk = _cols(keys)
    varr, vk = _cols([values])
    qb = q if nq else np.zeros(1, np.float64)
    g = _load().orc_group_quantiles(ctypes.addressof(karr), len(keys), null_policy, ctypes.addressof(varr),
                                    qb.ctypes.data, nq, interpolation, rep.ctypes.data, out.ctypes.data,
                                    valid.ctypes.data)
    if g < 0:
        raise TypeError("unsupported quantile request")
    return rep[:g].copy(), out[:g * nq].reshape(g, nq).copy(), valid[:g * nq].reshape(g, nq).astype(bool)

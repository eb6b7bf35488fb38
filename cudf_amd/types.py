"""Type ids and policies, numbered as libcudf's enums (cudf/types.hpp)."""
from __future__ import annotations

import enum

import numpy as np


class TypeId(enum.IntEnum):
    """cudf::type_id (cudf/types.hpp:203-235)."""
    EMPTY = 0
    INT8 = 1
    INT16 = 2
    INT32 = 3
    INT64 = 4
    UINT8 = 5
    UINT16 = 6
    UINT32 = 7
    UINT64 = 8
    FLOAT32 = 9
    FLOAT64 = 10
    BOOL8 = 11
    TIMESTAMP_DAYS = 12
    TIMESTAMP_SECONDS = 13
    TIMESTAMP_MILLISECONDS = 14
    TIMESTAMP_MICROSECONDS = 15
    TIMESTAMP_NANOSECONDS = 16
    DURATION_DAYS = 17
    DURATION_SECONDS = 18
    DURATION_MILLISECONDS = 19
    DURATION_MICROSECONDS = 20
    DURATION_NANOSECONDS = 21
    DICTIONARY32 = 22
    STRING = 23
    LIST = 24
    DECIMAL32 = 25
    DECIMAL64 = 26
    DECIMAL128 = 27
    STRUCT = 28


class NullPolicy(enum.IntEnum):
    """cudf::null_policy (cudf/types.hpp:126-129)."""
    EXCLUDE = 0
    INCLUDE = 1


class Sorted(enum.IntEnum):
    """cudf::sorted (cudf/types.hpp:167)."""
    NO = 0
    YES = 1


_NP = {
    TypeId.INT8: np.int8, TypeId.INT16: np.int16, TypeId.INT32: np.int32, TypeId.INT64: np.int64,
    TypeId.UINT8: np.uint8, TypeId.UINT16: np.uint16, TypeId.UINT32: np.uint32,
    TypeId.UINT64: np.uint64, TypeId.FLOAT32: np.float32, TypeId.FLOAT64: np.float64,
    TypeId.BOOL8: np.uint8,
    TypeId.TIMESTAMP_DAYS: np.int32, TypeId.TIMESTAMP_SECONDS: np.int64,
    TypeId.TIMESTAMP_MILLISECONDS: np.int64, TypeId.TIMESTAMP_MICROSECONDS: np.int64,
    TypeId.TIMESTAMP_NANOSECONDS: np.int64, TypeId.DURATION_DAYS: np.int32,
    TypeId.DURATION_SECONDS: np.int64, TypeId.DURATION_MILLISECONDS: np.int64,
    TypeId.DURATION_MICROSECONDS: np.int64, TypeId.DURATION_NANOSECONDS: np.int64,
    TypeId.DECIMAL32: np.int32, TypeId.DECIMAL64: np.int64,  # the fixed-point representation
}


def numpy_dtype(type_id: int) -> np.dtype:
    """Storage dtype of a fixed-width type id."""
    return np.dtype(_NP[TypeId(type_id)])


def type_id_of(dtype) -> TypeId:
    dt = np.dtype(dtype)
    if dt == np.bool_:
        return TypeId.BOOL8
    for k, v in _NP.items():
        if np.dtype(v) == dt and k <= TypeId.FLOAT64:
            return k
    raise TypeError(f"no type id for dtype {dt}")


def width(type_id: int) -> int:
    return numpy_dtype(type_id).itemsize

"""MI355X-native (gfx950) hash group-by with libcudf's groupby interface.

The compute path is ``cudf_amd/lib/libmig_groupby.so`` (HIP kernels behind
the C ABI of ``include/mig_groupby.h``); this package is the Python mirror of
pylibcudf's ``GroupBy`` / ``aggregation`` API on top of it.
"""
from . import aggregation, interop, pandas_compat
from .batch import BatchedGroupBy, BatchOptions
from .column import Column, Table
from .groupby import GroupBy, GroupByRequest, murmurhash3_x86_32
from .types import NullPolicy, Sorted, TypeId

__all__ = [
    "aggregation", "BatchedGroupBy", "BatchOptions", "Column", "Table", "GroupBy", "GroupByRequest", "murmurhash3_x86_32",
    "NullPolicy", "Sorted", "TypeId",
]

"""ctypes binding of the C ABI declared in ``include/mig_groupby.h``.

The library is ``cudf_amd/lib/libmig_groupby.so`` built by ``cudf_amd.build``.
Loading fails loudly when it is missing: there is no fallback path.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libmig_groupby.so")

# status codes (mig_status)
(MIG_OK, MIG_ERROR_LOGIC, MIG_ERROR_HIP, MIG_ERROR_DATA_TYPE, MIG_ERROR_ALLOC, MIG_ERROR_OVERFLOW,
 MIG_ERROR_OUT_OF_RANGE) = range(7)


class LogicError(RuntimeError):
    """cudf::logic_error: an invalid request (sizes, policies, kinds)."""


class HipError(RuntimeError):
    """A HIP runtime call failed (cudf::cuda_error equivalent)."""


class DataTypeError(TypeError):
    """cudf::data_type_error: unsupported type for the operation."""


class AllocationError(MemoryError):
    """Device allocation failed."""


_ERRORS = {
    MIG_ERROR_LOGIC: LogicError,
    MIG_ERROR_HIP: HipError,
    MIG_ERROR_DATA_TYPE: DataTypeError,
    MIG_ERROR_ALLOC: AllocationError,
    MIG_ERROR_OVERFLOW: OverflowError,
    MIG_ERROR_OUT_OF_RANGE: IndexError,
}


class ColumnView(ctypes.Structure):
    _fields_ = [
        ("type_id", ctypes.c_int32),
        ("scale", ctypes.c_int32),
        ("size", ctypes.c_int32),
        ("offset", ctypes.c_int32),
        ("data", ctypes.c_void_p),
        ("null_mask", ctypes.c_void_p),
        ("null_count", ctypes.c_int32),
        ("offsets_type_id", ctypes.c_int32),
        ("offsets", ctypes.c_void_p),
    ]


class Aggregation(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("ddof", ctypes.c_int32), ("null_handling", ctypes.c_int32),
                ("n", ctypes.c_int32), ("interpolation", ctypes.c_int32), ("num_quantiles", ctypes.c_int32),
                ("quantiles", ctypes.POINTER(ctypes.c_double))]


class AggregationRequest(ctypes.Structure):
    _fields_ = [
        ("values", ColumnView),
        ("aggregations", ctypes.POINTER(Aggregation)),
        ("num_aggregations", ctypes.c_int32),
    ]


ALLOC_FN = ctypes.CFUNCTYPE(ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p)
DEALLOC_FN = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p)


class MemoryResource(ctypes.Structure):
    _fields_ = [("ctx", ctypes.c_void_p), ("allocate", ALLOC_FN), ("deallocate", DEALLOC_FN)]


class OutColumn(ctypes.Structure):
    _fields_ = [
        ("type_id", ctypes.c_int32),
        ("size", ctypes.c_int32),
        ("data", ctypes.c_void_p),
        ("null_mask", ctypes.c_void_p),
        ("null_count", ctypes.c_int32),
        ("offsets_type_id", ctypes.c_int32),
        ("offsets", ctypes.c_void_p),
        ("data_bytes", ctypes.c_int64),
        ("mask_bytes", ctypes.c_int64),
        ("offsets_bytes", ctypes.c_int64),
        ("scale", ctypes.c_int32),
    ]


class BatchInstance(ctypes.Structure):
    _fields_ = [("keys", ctypes.POINTER(ColumnView)), ("values", ctypes.POINTER(ColumnView))]


class BatchOptions(ctypes.Structure):
    _fields_ = [
        ("load_factor", ctypes.c_double),
        ("lds_max_groups", ctypes.c_int32),
        ("block_size", ctypes.c_int32),
        ("tile_rows", ctypes.c_int32),
        ("use_graph", ctypes.c_int32),
        ("partition_min_rows", ctypes.c_int32),
    ]


# every function the header declares, with its ctypes signature
SIGNATURES = {
    "mig_last_error": (ctypes.c_char_p, []),
    "mig_device_allocate": (ctypes.c_void_p, [ctypes.c_size_t, ctypes.c_void_p]),
    "mig_device_deallocate": (None, [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p]),
    "mig_abi_version": (ctypes.c_int32, []),
    "mig_hash_murmurhash3_x86_32": (ctypes.c_int, [ctypes.POINTER(ColumnView), ctypes.c_int32,
                                                   ctypes.c_uint32, ctypes.c_void_p, ctypes.c_void_p]),
    "mig_groupby_aggregate": (ctypes.c_int, [ctypes.POINTER(ColumnView), ctypes.c_int32, ctypes.c_int32,
                                             ctypes.POINTER(AggregationRequest), ctypes.c_int32,
                                             ctypes.c_void_p, ctypes.POINTER(MemoryResource),
                                             ctypes.POINTER(ctypes.c_void_p)]),
    "mig_groupby_aggregate_sorted": (ctypes.c_int, [ctypes.POINTER(ColumnView), ctypes.c_int32, ctypes.c_int32,
                                                    ctypes.POINTER(AggregationRequest), ctypes.c_int32,
                                                    ctypes.c_void_p, ctypes.POINTER(MemoryResource),
                                                    ctypes.POINTER(ctypes.c_void_p)]),
    "mig_groupby_aggregate_into": (ctypes.c_int, [ctypes.POINTER(ColumnView), ctypes.c_int32, ctypes.c_int32,
                                                  ctypes.POINTER(AggregationRequest), ctypes.c_int32,
                                                  ctypes.POINTER(OutColumn), ctypes.POINTER(OutColumn),
                                                  ctypes.POINTER(ctypes.c_int32), ctypes.c_void_p]),
    "mig_groupby_into_scratch_bytes": (ctypes.c_size_t, []),
    "mig_result_num_groups": (ctypes.c_int32, [ctypes.c_void_p]),
    "mig_result_num_keys": (ctypes.c_int32, [ctypes.c_void_p]),
    "mig_result_num_requests": (ctypes.c_int32, [ctypes.c_void_p]),
    "mig_result_num_aggregations": (ctypes.c_int32, [ctypes.c_void_p, ctypes.c_int32]),
    "mig_result_key": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.POINTER(OutColumn)]),
    "mig_result_column": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32,
                                         ctypes.POINTER(OutColumn)]),
    "mig_result_release_key": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.POINTER(OutColumn)]),
    "mig_result_release_column": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32,
                                                 ctypes.POINTER(OutColumn)]),
    "mig_result_free": (None, [ctypes.c_void_p]),
    "mig_last_stats": (ctypes.c_int32, [ctypes.POINTER(ctypes.c_double), ctypes.c_int32]),
    "mig_set_phase_timing": (None, [ctypes.c_int32]),
    "mig_set_strategy": (None, [ctypes.c_int32]),
    "mig_batch_plan_create": (ctypes.c_int, [ctypes.POINTER(BatchInstance), ctypes.c_int32, ctypes.c_int32,
                                             ctypes.c_int32, ctypes.POINTER(Aggregation),
                                             ctypes.POINTER(ctypes.c_int32), ctypes.c_int32,
                                             ctypes.POINTER(BatchOptions), ctypes.c_void_p,
                                             ctypes.POINTER(ctypes.c_void_p)]),
    "mig_batch_execute": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "mig_batch_synchronize": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "mig_batch_num_groups": (ctypes.c_int32, [ctypes.c_void_p, ctypes.c_int32]),
    "mig_batch_result_key": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.POINTER(OutColumn)]),
    "mig_batch_result_column": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32,
                                               ctypes.c_int32, ctypes.POINTER(OutColumn)]),
    "mig_batch_plan_free": (None, [ctypes.c_void_p]),
    "mig_table_num_rows": (ctypes.c_int32, [ctypes.c_void_p]),
    "mig_table_num_columns": (ctypes.c_int32, [ctypes.c_void_p]),
    "mig_table_column": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.POINTER(OutColumn)]),
    "mig_table_release_column": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.POINTER(OutColumn)]),
    "mig_table_free": (None, [ctypes.c_void_p]),
    "mig_hash_partition": (ctypes.c_int, [ctypes.POINTER(ColumnView), ctypes.c_int32,
                                          ctypes.POINTER(ctypes.c_int32), ctypes.c_int32, ctypes.c_int32,
                                          ctypes.c_int32, ctypes.c_uint32, ctypes.c_void_p,
                                          ctypes.POINTER(MemoryResource), ctypes.POINTER(ctypes.c_void_p),
                                          ctypes.POINTER(ctypes.c_int32)]),
    "mig_partition": (ctypes.c_int, [ctypes.POINTER(ColumnView), ctypes.c_int32, ctypes.POINTER(ColumnView),
                                     ctypes.c_int32, ctypes.c_void_p, ctypes.POINTER(MemoryResource),
                                     ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(ctypes.c_int32)]),
    "mig_concatenate": (ctypes.c_int, [ctypes.POINTER(ColumnView), ctypes.c_int32, ctypes.c_int32,
                                       ctypes.c_void_p, ctypes.POINTER(MemoryResource),
                                       ctypes.POINTER(ctypes.c_void_p)]),
}

STAT_NAMES = ["strategy", "key_mode", "capacity", "est_groups", "groups", "retries", "state_bytes",
              "t_estimate_ms", "t_init_ms", "t_accumulate_ms", "t_second_pass_ms", "t_compact_ms",
              "t_outputs_ms", "t_rpa_hist_ms", "t_rpa_scatter1_ms", "t_rpa_scatter2_ms",
              "t_rpa_aggregate_ms", "lds_preagg", "rpa_speculative", "verified_pairs"]

STRATEGY_AUTO, STRATEGY_GLOBAL, STRATEGY_PARTITIONED, STRATEGY_PARTITIONED_EXACT = 0, 1, 2, 3


def set_strategy(strategy: int) -> None:
    """Force an aggregation strategy process-wide (0 = automatic)."""
    lib().mig_set_strategy(int(strategy))

_lib = None


def lib():
    """The loaded C-ABI library (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -m cudf_amd.build` "
                "(there is no CPU fallback)")
        l = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(l, name)
            fn.restype = res
            fn.argtypes = args
        _lib = l
    return _lib


def check(status: int) -> None:
    if status != MIG_OK:
        msg = lib().mig_last_error().decode(errors="replace")
        raise _ERRORS.get(status, RuntimeError)(msg)


def last_stats() -> dict:
    """Statistics of the last group-by on this thread (see mig_last_stats)."""
    buf = (ctypes.c_double * 32)()
    n = min(lib().mig_last_stats(buf, 32), 32)
    return {STAT_NAMES[i] if i < len(STAT_NAMES) else f"stat{i}": buf[i] for i in range(n)}


def set_phase_timing(on) -> None:
    """Phase timing of later calls: False/0 off, True/1 every phase, 2 the
    main aggregation kernel only (two events per call)."""
    lib().mig_set_phase_timing(int(on))

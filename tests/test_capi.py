"""CPU checks of the C-ABI library: it loads, exports every function the
headers in include/ declare, and the ctypes signatures cover them.  No compute
call is made (there is no GPU here)."""
import ctypes
import glob
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(mig_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_header_declares_entry_points():
    names = declared_functions()
    assert {"mig_groupby_aggregate", "mig_hash_murmurhash3_x86_32", "mig_result_free",
            "mig_last_error"} <= names


def test_library_exports_every_declared_symbol():
    from cudf_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.fail(f"{_lib.LIB_PATH} not built (python -m cudf_amd.build)")
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in sorted(declared_functions()) if not hasattr(lib, n)]
    assert not missing, f"declared but not exported: {missing}"


def test_ctypes_binding_covers_header():
    from cudf_amd import _lib
    assert declared_functions() == set(_lib.SIGNATURES)


def test_abi_version_and_error_without_gpu():
    from cudf_amd import _lib
    lib = _lib.lib()
    assert lib.mig_abi_version() == 1
    assert lib.mig_last_error() == b""


def test_invalid_request_rejected_before_device_work():
    """A size mismatch is a logic error raised on the host, before any HIP
    call (groupby.cu:210-214), so it is testable without a GPU."""
    from cudf_amd import _lib
    lib = _lib.lib()
    keys = (_lib.ColumnView * 1)()
    keys[0].type_id = 3
    keys[0].size = 10
    aggs = (_lib.Aggregation * 1)(_lib.Aggregation(0, 1))
    reqs = (_lib.AggregationRequest * 1)()
    reqs[0].values.type_id = 3
    reqs[0].values.size = 5
    reqs[0].aggregations = ctypes.cast(aggs, ctypes.POINTER(_lib.Aggregation))
    reqs[0].num_aggregations = 1
    h = ctypes.c_void_p()
    st = lib.mig_groupby_aggregate(keys, 1, 0, reqs, 1, None, None, ctypes.byref(h))
    assert st == _lib.MIG_ERROR_LOGIC
    assert b"Size mismatch" in lib.mig_last_error()
    with pytest.raises(_lib.LogicError):
        _lib.check(st)


def test_unsupported_kind_and_type_rejected():
    from cudf_amd import _lib
    lib = _lib.lib()

    def call(vtype, kind):
        keys = (_lib.ColumnView * 1)()
        keys[0].type_id = 3
        keys[0].size = 4
        aggs = (_lib.Aggregation * 1)(_lib.Aggregation(kind, 1))
        reqs = (_lib.AggregationRequest * 1)()
        reqs[0].values.type_id = vtype
        reqs[0].values.size = 4
        reqs[0].aggregations = ctypes.cast(aggs, ctypes.POINTER(_lib.Aggregation))
        reqs[0].num_aggregations = 1
        h = ctypes.c_void_p()
        return lib.mig_groupby_aggregate(keys, 1, 0, reqs, 1, None, None, ctypes.byref(h))

    assert call(3, 6) == _lib.MIG_ERROR_LOGIC           # ANY: not a group-by aggregation here
    assert call(23, 13) == _lib.MIG_ERROR_DATA_TYPE     # MEDIAN of strings
    assert call(23, 0) == _lib.MIG_ERROR_DATA_TYPE      # SUM of strings
    assert call(12, 9) == _lib.MIG_ERROR_DATA_TYPE      # MEAN of timestamps


def test_negative_offset_or_size_rejected_before_device_work():
    """A view with a negative row offset or size is a logic error raised on
    the host (a wrapped int32 offset would otherwise read before the buffer)."""
    from cudf_amd import _lib
    lib = _lib.lib()
    aggs = (_lib.Aggregation * 1)(_lib.Aggregation(0, 1))

    def call(key_offset, val_offset, size=10):
        keys = (_lib.ColumnView * 1)()
        keys[0].type_id = 3
        keys[0].size = size
        keys[0].offset = key_offset
        reqs = (_lib.AggregationRequest * 1)()
        reqs[0].values.type_id = 3
        reqs[0].values.size = size
        reqs[0].values.offset = val_offset
        reqs[0].aggregations = ctypes.cast(aggs, ctypes.POINTER(_lib.Aggregation))
        reqs[0].num_aggregations = 1
        h = ctypes.c_void_p()
        return lib.mig_groupby_aggregate(keys, 1, 0, reqs, 1, None, None, ctypes.byref(h))

    assert call(-5, 0) == _lib.MIG_ERROR_LOGIC
    assert b"negative" in lib.mig_last_error()
    assert call(0, -(2**31)) == _lib.MIG_ERROR_LOGIC
    assert call(0, 0, size=-3) == _lib.MIG_ERROR_LOGIC


def test_column_view_rejects_offsets_beyond_int32():
    """cudf size_type is int32: a Column whose row offset or size does not
    fit raises instead of wrapping in the ctypes view."""
    import torch
    from cudf_amd import Column, TypeId
    buf = torch.zeros(16, dtype=torch.uint8)
    with pytest.raises(OverflowError):
        Column(TypeId.INT32, 2, buf, None, 0, 2**31).view()
    with pytest.raises(OverflowError):
        Column(TypeId.INT32, 2**31, buf).view()
    with pytest.raises(OverflowError):
        Column(TypeId.INT32, 2, buf, None, 0, -1).view()
    v = Column(TypeId.INT32, 2, buf, None, 0, 2**31 - 1).view()
    assert v.offset == 2**31 - 1

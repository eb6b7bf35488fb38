"""Arrow ingest / egress through the buffers (no per-element conversion),
checked on host memory: every Arrow layout the library maps (strings and
large strings with nulls, slices with an offset, booleans, fixed width,
timestamps, empty arrays) survives from_arrow -> to_arrow unchanged.  The
device path is the same code with a HIP device (tests/test_compat.py and
tests/test_gpu_arrow.py)."""
import numpy as np
import pytest

pa = pytest.importorskip("pyarrow")


def _strings(rng, n):
    return [None if rng.random() < 0.1 else "".join(chr(97 + int(c)) for c in rng.integers(0, 26, rng.integers(0, 12)))
            for _ in range(n)]


def _cases():
    rng = np.random.default_rng(1)
    n = 5003
    s = _strings(rng, n)
    return [
        pa.array(s), pa.array(s, type=pa.large_string()), pa.array(s)[13:4000],
        pa.array([None if rng.random() < 0.2 else int(x) for x in rng.integers(-5, 5, n)], type=pa.int32())[7:900],
        pa.array(rng.standard_normal(n)), pa.array(rng.integers(0, 255, n).astype(np.uint8)),
        pa.array([None if rng.random() < 0.3 else bool(x) for x in rng.integers(0, 2, n)])[3:777],
        pa.array([True, False, None]), pa.array(np.arange(5, dtype=np.int64)).cast(pa.timestamp("ms")),
        pa.array([1, None, 3], type=pa.date32()), pa.array([], type=pa.string()), pa.array([""] * 10),
    ]


@pytest.mark.parametrize("i", range(12))
def test_round_trip_through_buffers(i):
    from cudf_amd.interop import from_arrow, to_arrow
    a = _cases()[i]
    c = from_arrow(a, device="cpu")
    assert c.size == len(a) and c.null_count == a.null_count
    b = to_arrow(c)
    want = a if a.type == b.type else a.cast(b.type)  # large_string comes back as string when it fits
    assert b.equals(want)


def test_ingest_keeps_the_slice_offset_and_buffers():
    """A sliced string array is taken with its offset, its offsets child and
    characters copied whole (the view starts at the slice's first row)."""
    from cudf_amd.interop import from_arrow
    a = pa.array(["ab", None, "cde", "", "f"])[1:4]
    c = from_arrow(a, device="cpu")
    assert c.offset == 1 and c.size == 3
    assert c.offsets.numpy().tolist()[:5] == [0, 2, 2, 5, 5]
    vals, valid = c.to_numpy()
    assert vals == [b"", b"cde", b""] and valid.tolist() == [False, True, True]

"""Arrow buffers to the device and back on the GPU: a 1M-row nullable string
column and its int64 values are ingested by whole-buffer copies
(cpp/src/interop/from_arrow_host.cu:173-200), come back equal through
to_arrow, and group the same as the oracle.  The ingest time is printed."""
import time

import numpy as np
import pytest

from oracle import oracle as orc
from tests.gbutil import AGG, assert_parity, host_keys

pa = pytest.importorskip("pyarrow")
pytestmark = pytest.mark.gpu


def test_arrow_1m_nullable_strings_ingest_round_trip_and_groupby():
    import torch
    from cudf_amd import GroupBy, GroupByRequest, Table
    from cudf_amd.aggregation import Aggregation
    from cudf_amd.interop import from_arrow, to_arrow
    rng = np.random.default_rng(7)
    n = 1_000_000
    pool = np.array(["k%05d" % i + "x" * int(i % 9) for i in range(50_000)], dtype=object)
    words = pool[rng.integers(0, len(pool), n)]
    svalid = rng.random(n) > 0.05
    sarr = pa.array(words, type=pa.string(), mask=~svalid)
    varr = pa.array(rng.integers(-10**9, 10**9, n), type=pa.int64(), mask=rng.random(n) < 0.03)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    skey, vals = from_arrow(sarr), from_arrow(varr)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    print(f"ingest of {n} strings + int64 values: {dt * 1e3:.1f} ms "
          f"({(sarr.nbytes + varr.nbytes) / dt / 1e9:.2f} GB/s)")
    assert to_arrow(skey).equals(sarr) and to_arrow(vals).equals(varr)
    # the same buffers group as the oracle
    out_keys, out_res = GroupBy(Table([skey])).aggregate(
        [GroupByRequest(vals, [Aggregation(AGG["SUM"]), Aggregation(AGG["COUNT_VALID"]), Aggregation(AGG["MAX"])])])
    g_keys = [host_keys(c) for c in out_keys.columns()]
    g_res = [[c.to_numpy() for c in t.columns()] for t in out_res]
    hk = orc.HostColumn.from_strings([w.encode() for w in words], svalid)
    hv = orc.HostColumn.from_numpy(varr.fill_null(0).to_numpy(), ~np.asarray(varr.is_null()))
    assert assert_parity([hk], [(hv, [(AGG["SUM"], 1), (AGG["COUNT_VALID"], 1), (AGG["MAX"], 1)])], 0,
                         gpu=(g_keys, g_res)) > 40_000


def test_arrow_sliced_and_boolean_on_device():
    from cudf_amd.interop import from_arrow, to_arrow
    rng = np.random.default_rng(8)
    b = pa.array([None if rng.random() < 0.2 else bool(x) for x in rng.integers(0, 2, 10_000)])[17:9000]
    s = pa.array([None if i % 7 == 0 else "s%d" % i for i in range(20_000)], type=pa.large_string())[5:19_000]
    assert to_arrow(from_arrow(b)).equals(b)
    back = to_arrow(from_arrow(s))
    assert back.equals(s.cast(back.type))

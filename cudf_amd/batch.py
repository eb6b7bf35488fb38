"""Batched hash group-by: many independent single-key group-bys executed
together by the library's persistent work-queue kernel, replayed from one
captured hipGraph (``mig_batch_*`` in ``include/mig_groupby.h``).

Every instance has one fixed-width key column and one values column per
request; all instances share the key type, the value types and the
aggregations.  Instance ``i``'s result equals
``GroupBy(Table([keys_i])).aggregate(requests_i)`` (groups in another order).
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import List, Optional, Sequence, Tuple

import torch

from . import _lib
from .aggregation import Aggregation
from .column import Column, Table
from .types import NullPolicy, TypeId, width


@dataclass
class BatchOptions:
    """Tuning knobs of a batch plan (0 = the library's default)."""

    load_factor: float = 0.0    # target groups per table slot (default 0.5)
    lds_max_groups: int = 0     # LDS pre-aggregation up to this many estimated groups (default 512, -1 off)
    block_size: int = 0         # 256 or 512 (default 256)
    tile_rows: int = 0          # rows per work item (default 4096)
    use_graph: int = 0          # -1 = launch kernels one by one instead of replaying the graph
    partition_min_rows: int = 0  # large high-cardinality instances take the single-call
                                 # partitioned path (default 1 << 20 rows, -1 never)

    def to_c(self) -> _lib.BatchOptions:
        return _lib.BatchOptions(float(self.load_factor), int(self.lds_max_groups), int(self.block_size),
                                 int(self.tile_rows), int(self.use_graph), int(self.partition_min_rows))


def _stream_handle(stream: Optional[torch.cuda.Stream]):
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


class _DeviceBytes:
    """A borrowed device byte range, exposed to torch through the CUDA array
    interface (a zero-copy view that torch then copies)."""

    def __init__(self, ptr: int, nbytes: int):
        self.__cuda_array_interface__ = {"shape": (nbytes,), "typestr": "|u1", "data": (ptr, False),
                                         "version": 2, "strides": None}


def _copy_device_bytes(ptr: int, nbytes: int, dtype=torch.uint8, pad_to: int = 1) -> torch.Tensor:
    dev = torch.device("cuda", torch.cuda.current_device())
    out = torch.zeros(max(pad_to, nbytes), dtype=torch.uint8, device=dev)
    if nbytes > 0:
        out[:nbytes].copy_(torch.as_tensor(_DeviceBytes(ptr, nbytes), device=dev))
    return out.view(dtype) if dtype != torch.uint8 else out


class BatchedGroupBy:
    """A plan over ``instances``: a list of ``(key_column, [values_column per
    request])``; ``aggregations[r]`` is the aggregation list of request r."""

    def __init__(self, instances: Sequence[Tuple[Column, Sequence[Column]]],
                 aggregations: Sequence[Sequence[Aggregation]],
                 null_handling: NullPolicy = NullPolicy.EXCLUDE,
                 options: Optional[BatchOptions] = None,
                 stream: Optional[torch.cuda.Stream] = None):
        self._instances = [(k, list(vs)) for k, vs in instances]  # keep the columns alive
        self._aggs = [list(a) for a in aggregations]
        n = len(self._instances)
        nreq = len(self._aggs)
        self._keys_c = []
        self._vals_c = []
        inst = (_lib.BatchInstance * max(1, n))()
        for i, (k, vs) in enumerate(self._instances):
            if len(vs) != nreq:
                raise _lib.LogicError(f"instance {i} has {len(vs)} values columns for {nreq} requests")
            kv = (_lib.ColumnView * 1)(k.view())
            vv = (_lib.ColumnView * max(1, nreq))(*[v.view() for v in vs])
            self._keys_c.append(kv)
            self._vals_c.append(vv)
            inst[i].keys = ctypes.cast(kv, ctypes.POINTER(_lib.ColumnView))
            inst[i].values = ctypes.cast(vv, ctypes.POINTER(_lib.ColumnView))
        flat = [a for al in self._aggs for a in al]
        aggs = (_lib.Aggregation * max(1, len(flat)))(*[_lib.Aggregation(int(a.kind()), a.ddof) for a in flat])
        counts = (ctypes.c_int32 * max(1, nreq))(*[len(a) for a in self._aggs])
        opts = (options or BatchOptions()).to_c()
        self._handle = ctypes.c_void_p()
        L = _lib.lib()
        _lib.check(L.mig_batch_plan_create(inst, n, 1, int(NullPolicy(null_handling)), aggs, counts, nreq,
                                           ctypes.byref(opts), _stream_handle(stream),
                                           ctypes.byref(self._handle)))
        self._n = n

    def __len__(self) -> int:
        return self._n

    def execute(self, stream: Optional[torch.cuda.Stream] = None) -> None:
        """Enqueue every instance's group-by on ``stream``."""
        _lib.check(_lib.lib().mig_batch_execute(self._handle, _stream_handle(stream)))

    def synchronize(self, stream: Optional[torch.cuda.Stream] = None) -> None:
        """Wait for the last execute (re-running instances whose table the
        estimate sized too small)."""
        _lib.check(_lib.lib().mig_batch_synchronize(self._handle, _stream_handle(stream)))

    def run(self, stream: Optional[torch.cuda.Stream] = None) -> None:
        self.execute(stream)
        self.synchronize(stream)

    def num_groups(self, i: int) -> int:
        return int(_lib.lib().mig_batch_num_groups(self._handle, i))

    @staticmethod
    def _copy(out: _lib.OutColumn) -> Column:
        data = _copy_device_bytes(out.data, out.size * width(TypeId(out.type_id)))
        mask = None
        if out.null_mask:
            words = (out.size + 31) // 32
            mask = _copy_device_bytes(out.null_mask, words * 4, torch.int32, pad_to=4)
        return Column(out.type_id, out.size, data, mask, out.null_count)

    def result(self, i: int) -> Tuple[Table, List[Table]]:
        """Instance ``i``'s unique keys and per-request results, copied out of
        the plan's buffers on torch's current stream (valid after
        ``synchronize``, on the stream the plan ran on or one ordered after it)."""
        L = _lib.lib()
        out = _lib.OutColumn()
        _lib.check(L.mig_batch_result_key(self._handle, i, ctypes.byref(out)))
        keys = self._copy(out)
        results = []
        for r, al in enumerate(self._aggs):
            cols = []
            for a in range(len(al)):
                _lib.check(L.mig_batch_result_column(self._handle, i, r, a, ctypes.byref(out)))
                cols.append(self._copy(out))
            results.append(Table(cols))
        return Table([keys]), results

    def close(self) -> None:
        if self._handle:
            _lib.lib().mig_batch_plan_free(self._handle)
            self._handle = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

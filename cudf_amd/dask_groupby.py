"""Group-by over many partitions: partial aggregation, combine, finalize.

The algorithm of dask_cudf's ``groupby_agg`` (python/dask_cudf/dask_cudf/
groupby.py): every input partition is aggregated into partial state
(``_groupby_partition_agg``, :733-770: count, sum, min, max, and for var/std
the sum of the float64 squares), partial tables are concatenated and
aggregated again with sum/min/max (``_tree_node_agg``, :772-806), and the
last step derives mean = sum / count and var = (Σx² - (Σx)² / n) / (n - ddof)
(``_finalize_gb_agg`` / ``_var_agg``, :809-880).  With ``split_out > 1`` the
partials are hash-partitioned by key first, so each output partition holds a
disjoint set of keys (the shuffle of ``_shuffle_aggregate``, :367-432).

Each partition is one device's share in a multi-device deployment; here the
partitions live on one device and every step runs on it (HIP group-by,
hash partitioning and concatenation through the C ABI; the finalize
arithmetic is elementwise device work, as dask_cudf does it with cudf
columns).  Differences from dask_cudf: var/std of a group with n - ddof <= 0
is null (cudf's group-by semantics) rather than NaN, and counts are int64
(the sum of partial counts), as in dask_cudf.
"""
from __future__ import annotations

from typing import Dict, List, Optional, Sequence, Tuple

import torch

from . import aggregation as agg
from .column import Column, Table
from .groupby import GroupBy, GroupByRequest
from .partitioning import concatenate, hash_partition, split
from .types import NullPolicy, TypeId, numpy_dtype

SUPPORTED = ("count", "sum", "min", "max", "mean", "var", "std")
_PARTIAL_AGG = {"count": agg.count, "sum": agg.sum, "min": agg.min, "max": agg.max}
# how each partial column combines (groupby.py:789-796)
_COMBINE = {"count": agg.sum, "sum": agg.sum, "min": agg.min, "max": agg.max, "pow2sum": agg.sum}
SEP = "__"

_TORCH = {
    TypeId.INT8: torch.int8, TypeId.INT16: torch.int16, TypeId.INT32: torch.int32, TypeId.INT64: torch.int64,
    TypeId.UINT8: torch.uint8, TypeId.FLOAT32: torch.float32, TypeId.FLOAT64: torch.float64,
    TypeId.BOOL8: torch.uint8,
}


class Partial:
    """Partial aggregation state: key columns and named partial columns
    (``<col>__count``, ``__sum``, ``__min``, ``__max``, ``__pow2sum``)."""

    def __init__(self, keys: Table, columns: Dict[str, Column]):
        self.keys = keys
        self.columns = columns

    def table(self) -> Table:
        return Table(self.keys.columns() + [self.columns[n] for n in sorted(self.columns)])

    def num_rows(self) -> int:
        return self.keys.num_rows()


def _device_values(col: Column) -> torch.Tensor:
    """Values of a fixed-width column as a torch tensor on the device (int64 for
    unsigned 16..64-bit storage, so they convert exactly)."""
    tid = TypeId(col.type_id)
    dt = numpy_dtype(tid)
    if col.size == 0 or col.data is None:
        return torch.zeros(0, dtype=torch.int64, device=torch.device("cuda", torch.cuda.current_device()))
    start = col.offset * dt.itemsize
    raw = col.data[start: start + col.size * dt.itemsize]
    if tid in _TORCH:
        return raw.view(_TORCH[tid])
    if tid == TypeId.UINT16:
        return raw.view(torch.int16).to(torch.int32) & 0xFFFF
    if tid == TypeId.UINT32:
        return raw.view(torch.int32).to(torch.int64) & 0xFFFFFFFF
    if tid == TypeId.UINT64:
        return raw.view(torch.int64)  # reinterpretation handled by _as_float64
    raise TypeError(f"unsupported value type {tid.name}")


def _device_valid(col: Column) -> Optional[torch.Tensor]:
    if col.mask is None:
        return None
    words = col.mask.view(torch.int32).to(torch.int64) & 0xFFFFFFFF
    shifts = torch.arange(32, device=words.device, dtype=torch.int64)
    bits = ((words[:, None] >> shifts) & 1).reshape(-1)
    return bits[col.offset: col.offset + col.size].to(torch.bool)


def _pack_valid(valid: torch.Tensor) -> Tuple[torch.Tensor, int]:
    n = valid.numel()
    padded = torch.zeros(((n + 511) // 512) * 512, dtype=torch.int64, device=valid.device)
    padded[:n] = valid.to(torch.int64)
    shifts = torch.arange(32, device=valid.device, dtype=torch.int64)
    words = (padded.view(-1, 32) << shifts).sum(dim=1)
    words = torch.where(words >= 2 ** 31, words - 2 ** 32, words).to(torch.int32)
    return words, int(n - int(valid.sum().item()))


def _column_f64(values: torch.Tensor, valid: Optional[torch.Tensor]) -> Column:
    data = values.to(torch.float64).contiguous()
    n = data.numel()
    buf = data.view(torch.uint8) if n else torch.zeros(1, dtype=torch.uint8, device=data.device)
    if valid is None:
        return Column(TypeId.FLOAT64, n, buf)
    mask, nulls = _pack_valid(valid)
    return Column(TypeId.FLOAT64, n, buf, mask, nulls)


def _as_float64(col: Column) -> Column:
    """The column converted to float64 (nulls kept): dask_cudf squares
    ``df[col].astype("float64")`` (groupby.py:758)."""
    if col.type_id == TypeId.FLOAT64:
        return col
    v = _device_values(col).to(torch.float64)
    if col.type_id == TypeId.UINT64:
        v = torch.where(v < 0, v + 2.0 ** 64, v)
    return _column_f64(v, _device_valid(col))


def partition_agg(keys: Table, values: Dict[str, Column], aggs: Dict[str, Sequence[str]], dropna: bool = True,
                  stream: Optional[torch.cuda.Stream] = None) -> Partial:
    """Partial state of one partition (groupby.py:733-770)."""
    requests, names = [], []
    for col in sorted(aggs):
        want = set()
        for a in aggs[col]:
            if a in ("mean", "var", "std"):
                want |= {"count", "sum"}
            elif a in _PARTIAL_AGG:
                want.add(a)
            else:
                raise ValueError(f"Unexpected aggregation: {a}")
        kinds = sorted(want)
        requests.append(GroupByRequest(values[col], [_PARTIAL_AGG[k]() for k in kinds]))
        names.append([f"{col}{SEP}{k}" for k in kinds])
        if set(aggs[col]) & {"var", "std"}:
            requests.append(GroupByRequest(_as_float64(values[col]), [agg.sum_of_squares()]))
            names.append([f"{col}{SEP}pow2sum"])
    policy = NullPolicy.EXCLUDE if dropna else NullPolicy.INCLUDE
    ukeys, results = GroupBy(keys, policy).aggregate(requests, stream=stream)
    cols = {}
    for nlist, res in zip(names, results):
        for name, c in zip(nlist, res.columns()):
            cols[name] = c
    return Partial(ukeys, cols)


def tree_node_agg(partials: Sequence[Partial], dropna: bool = True,
                  stream: Optional[torch.cuda.Stream] = None) -> Partial:
    """Concatenated partials aggregated again (groupby.py:772-806)."""
    partials = list(partials)
    nk = partials[0].keys.num_columns()
    names = sorted(partials[0].columns)
    t = concatenate([p.table() for p in partials], stream=stream) if len(partials) > 1 else partials[0].table()
    cols = t.columns()
    requests = [GroupByRequest(cols[nk + i], [_COMBINE[name.rsplit(SEP, 1)[1]]()]) for i, name in enumerate(names)]
    policy = NullPolicy.EXCLUDE if dropna else NullPolicy.INCLUDE
    ukeys, results = GroupBy(Table(cols[:nk]), policy).aggregate(requests, stream=stream)
    return Partial(ukeys, {name: res.columns()[0] for name, res in zip(names, results)})


def _float_values(col: Column) -> Tuple[torch.Tensor, torch.Tensor]:
    v = _device_values(col).to(torch.float64)
    if col.type_id == TypeId.UINT64:
        v = torch.where(v < 0, v + 2.0 ** 64, v)
    valid = _device_valid(col)
    if valid is None:
        valid = torch.ones(col.size, dtype=torch.bool, device=v.device)
    return v, valid


def finalize_agg(partials: Sequence[Partial], aggs: Dict[str, Sequence[str]], dropna: bool = True, ddof: int = 1,
                 stream: Optional[torch.cuda.Stream] = None) -> Tuple[Table, Dict[Tuple[str, str], Column]]:
    """Last combine, then mean / var / std from count, Σx and Σx²
    (groupby.py:809-880, _var_agg :826-846)."""
    gb = tree_node_agg(partials, dropna, stream)
    out: Dict[Tuple[str, str], Column] = {}
    for col in sorted(aggs):
        for a in aggs[col]:
            if a in _PARTIAL_AGG:
                out[(col, a)] = gb.columns[f"{col}{SEP}{a}"]
                continue
            n, _ = _float_values(gb.columns[f"{col}{SEP}count"])
            s, s_ok = _float_values(gb.columns[f"{col}{SEP}sum"])
            if a == "mean":
                ok = s_ok & (n > 0)
                out[(col, a)] = _column_f64(torch.where(ok, s / torch.where(ok, n, 1.0), 0.0), ok)
                continue
            x2, x2_ok = _float_values(gb.columns[f"{col}{SEP}pow2sum"])
            div = n - ddof
            ok = s_ok & x2_ok & (div > 0)
            safe_n = torch.where(n > 0, n, 1.0)
            var = (x2 - s * s / safe_n) / torch.where(div > 0, div, 1.0)
            var = torch.where(ok, var, 0.0)
            out[(col, a)] = _column_f64(torch.sqrt(var) if a == "std" else var, ok)
    return gb.keys, out


def groupby_agg(partitions: Sequence[Tuple[Table, Dict[str, Column]]], aggs: Dict[str, Sequence[str]],
                dropna: bool = True, split_every: int = 8, split_out: int = 1, ddof: int = 1,
                stream: Optional[torch.cuda.Stream] = None) -> List[Tuple[Table, Dict[Tuple[str, str], Column]]]:
    """Group-by over ``partitions`` (each ``(keys, {name: values})``).

    ``split_out == 1``: tree reduction of the partials, ``split_every`` at a
    time.  ``split_out > 1``: partials are hash-partitioned by key into
    ``split_out`` parts and each output aggregates its part from every
    partial; the outputs hold disjoint key sets.  Returns one
    ``(keys, {(column, aggregation): result})`` per output partition."""
    for col, lst in aggs.items():
        for a in lst:
            if a not in SUPPORTED:
                raise ValueError(f"Unexpected aggregation: {a}")
    if split_every < 2:
        raise ValueError("split_every must be at least 2")
    partials = [partition_agg(k, v, aggs, dropna, stream) for k, v in partitions]
    if split_out <= 1:
        while len(partials) > split_every:
            partials = [tree_node_agg(partials[i:i + split_every], dropna, stream)
                        for i in range(0, len(partials), split_every)]
        return [finalize_agg(partials, aggs, dropna, ddof, stream)]
    nk = partials[0].keys.num_columns()
    names = sorted(partials[0].columns)
    parts: List[List[Partial]] = [[] for _ in range(split_out)]
    for p in partials:
        t, offsets = hash_partition(p.table(), list(range(nk)), split_out, stream=stream)
        for j, piece in enumerate(split(t, offsets)):
            cols = piece.columns()
            parts[j].append(Partial(Table(cols[:nk]), dict(zip(names, cols[nk:]))))
    return [finalize_agg(parts[j], aggs, dropna, ddof, stream) for j in range(split_out)]

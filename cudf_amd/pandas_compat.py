"""pandas-shaped group-by with a GPU first and pandas as the fallback.

``groupby_agg(df, by, spec)`` returns what
``df.groupby(by, sort=sort, dropna=dropna).agg(spec)`` returns, computed by the
MI355X hash group-by when the frame and the aggregations are supported there,
and by pandas itself otherwise (the try-fast / fall-back-to-slow shape of
cudf.pandas, reference ``python/cudf/cudf/pandas/fast_slow_proxy.py:873-945``).

The adapter between the two semantics (the reference's parity tests,
``python/cudf/cudf/tests/test_groupby.py:1437-1477``):

* pandas treats float NaN as missing; here NaN values and keys become nulls
  before the group-by;
* pandas ``sum`` of a group without valid values is 0 (``min_count=0``)
  where the group-by's SUM is null;
* pandas ``count`` is int64 where COUNT is int32; ``mean``/``var``/``std``
  are float64 with NaN for null, float32 ``mean`` stays float32;
* ``sort=False`` orders groups by first appearance: the GPU path computes
  each group's first row (MIN of the row number) and orders by it;
* ``nunique``, ``first`` and ``last`` (sort-only aggregations: NUNIQUE and
  NTH_ELEMENT 0 / -1 with nulls excluded, as pandas skips missing values) run
  on the GPU when the frame is grouped by one key column that is already
  sorted ascending without missing keys (the pre-sorted path, whose runs are
  then the groups in key order); otherwise pandas runs them.
"""
from __future__ import annotations

from typing import Dict, List, Sequence, Tuple, Union

import numpy as np

from .aggregation import Aggregation, Kind

_KINDS = {"sum": Kind.SUM, "count": Kind.COUNT_VALID, "min": Kind.MIN, "max": Kind.MAX,
          "mean": Kind.MEAN, "var": Kind.VARIANCE, "std": Kind.STD, "size": Kind.COUNT_ALL,
          "prod": Kind.PRODUCT}
# aggregations the GPU takes on string values
_STRING_VALUE_AGGS = {"count", "size", "min", "max"}
# sort-only aggregations: taken with one key column sorted ascending
_SORT_ONLY = {"nunique": lambda: Aggregation(Kind.NUNIQUE),
              "first": lambda: Aggregation(Kind.NTH_ELEMENT, n=0),
              "last": lambda: Aggregation(Kind.NTH_ELEMENT, n=-1)}


class Unsupported(Exception):
    """The GPU path does not take this frame or aggregation."""


def _series_to_column(s, for_key: bool, strings_ok: bool = False):
    """Device column of a pandas Series (NaN and pandas NA become nulls)."""
    import pandas as pd
    from .column import Column
    if isinstance(s.dtype, pd.CategoricalDtype):
        raise Unsupported("categorical columns")
    if s.dtype == object or isinstance(s.dtype, pd.StringDtype):
        if not (for_key or strings_ok):
            raise Unsupported("string values")
        if not all(isinstance(v, str) for v in s.dropna().tolist()):
            raise Unsupported("object column of non-strings")
        valid = ~s.isna().to_numpy()
        vals = [b"" if not ok else str(v).encode() for v, ok in zip(s.tolist(), valid)]
        return Column.from_strings(vals, valid if not valid.all() else None)
    if isinstance(s.dtype, pd.api.extensions.ExtensionDtype):  # nullable Int/Float/boolean
        valid = ~s.isna().to_numpy()
        np_dtype = s.dtype.numpy_dtype
        values = s.to_numpy(dtype=np_dtype, na_value=0)
    else:
        values = s.to_numpy()
        if values.dtype.kind not in "iufb":
            raise Unsupported(f"dtype {values.dtype}")
        valid = ~np.isnan(values) if values.dtype.kind == "f" else np.ones(len(values), bool)
        if values.dtype.kind == "f":
            values = np.where(valid, values, 0).astype(values.dtype)
    return Column.from_numpy(np.ascontiguousarray(values), None if valid.all() else valid)


def _normalise_spec(spec) -> Tuple[Dict[str, List[str]], bool]:
    flat = all(isinstance(v, str) for v in spec.values())
    return {c: ([v] if isinstance(v, str) else list(v)) for c, v in spec.items()}, flat


def _gpu(df, by: List[str], spec: Dict[str, List[str]], flat: bool, sort: bool, dropna: bool):
    import pandas as pd
    from . import GroupBy, GroupByRequest, NullPolicy, Sorted, Table
    sort_only = False
    for aggs in spec.values():
        for a in aggs:
            if a in _SORT_ONLY:
                sort_only = True
            elif a not in _KINDS:
                raise Unsupported(f"aggregation {a}")
    if sort_only:
        if len(by) != 1 or df[by[0]].isna().any() or not df[by[0]].is_monotonic_increasing:
            raise Unsupported("nunique / first / last need one sorted key column without missing keys")
    keys = Table([_series_to_column(df[k], True) for k in by])
    reqs, order = [], []
    for c, aggs in spec.items():
        col = _series_to_column(df[c], False, strings_ok=set(aggs) <= _STRING_VALUE_AGGS)
        reqs.append(GroupByRequest(col, [_SORT_ONLY[a]() if a in _SORT_ONLY else Aggregation(_KINDS[a], 1)
                                         for a in aggs]))
        order.append((c, aggs))
    rownum = _series_to_column(pd.Series(np.arange(len(df), dtype=np.int64)), False)
    reqs.append(GroupByRequest(rownum, [Aggregation(Kind.MIN)]))
    out_keys, out = GroupBy(keys, NullPolicy.EXCLUDE if dropna else NullPolicy.INCLUDE,
                            Sorted.YES if sort_only else Sorted.NO).aggregate(reqs)
    first = out[-1].columns()[0].to_numpy()[0]
    # key index
    key_arrays = []
    for k, kc in zip(by, out_keys.columns()):
        vals, valid = kc.to_numpy()
        src = df[k]
        if kc.type_id.name == "STRING":
            arr = np.array([v.decode() for v in vals], dtype=object)
            if valid is not None:
                arr[~valid] = np.nan
        elif isinstance(src.dtype, pd.api.extensions.ExtensionDtype):
            arr = pd.array(vals, dtype=src.dtype)
            if valid is not None:
                arr[~valid] = pd.NA
        else:
            arr = vals.astype(src.dtype) if valid is None else np.where(valid, vals, np.nan)
        key_arrays.append(arr)
    if sort:
        perm = np.lexsort([pd.Series(a).rank(method="dense", na_option="bottom").to_numpy()
                           for a in reversed(key_arrays)])
    else:
        perm = np.argsort(first, kind="stable")
    cols, names = [], []
    for (c, aggs), table in zip(order, out[:-1]):
        src = df[c]
        for a, rc in zip(aggs, table.columns()):
            vals, valid = rc.to_numpy()
            if isinstance(vals, list):  # strings come back as a list of bytes
                vals = np.array(vals + [b""], dtype=object)[:-1]
            vals = vals[perm]
            valid = None if valid is None else valid[perm]
            if rc.type_id.name == "STRING":  # min / max of strings
                arr = np.array([v.decode() for v in vals], dtype=object)
                if valid is not None:
                    arr[~valid] = np.nan
                vals = arr
            elif a in ("sum", "prod"):
                # pandas: an all-missing group sums to 0 and multiplies to 1 (min_count=0)
                if valid is not None:
                    vals = np.where(valid, vals, 0 if a == "sum" else 1).astype(vals.dtype)
                if isinstance(src.dtype, pd.api.extensions.ExtensionDtype):
                    vals = pd.array(vals, dtype=pd.Int64Dtype() if vals.dtype.kind == "i" else src.dtype)
            elif a in ("count", "size", "nunique"):
                vals = vals.astype(np.int64)
            elif a in ("min", "max", "first", "last"):
                if isinstance(src.dtype, pd.api.extensions.ExtensionDtype):
                    vals = pd.array(vals, dtype=src.dtype)
                    if valid is not None:
                        vals[~valid] = pd.NA
                elif valid is not None and not valid.all():
                    vals = np.where(valid, vals, np.nan)
            else:  # mean, var, std
                vals = vals.astype(np.float64)
                if valid is not None:
                    vals = np.where(valid, vals, np.nan)
                if a == "mean" and src.dtype == np.float32:
                    vals = vals.astype(np.float32)
                if isinstance(src.dtype, pd.api.extensions.ExtensionDtype):
                    vals = pd.array(vals, dtype=pd.Float64Dtype())
            cols.append(vals)
            names.append(c if flat else (c, a))
    ordered = [a[perm] for a in key_arrays]  # numpy and pandas extension arrays alike
    if len(by) == 1:
        index = pd.Index(ordered[0], name=by[0])
    else:
        index = pd.MultiIndex.from_arrays(ordered, names=by)
    result = pd.DataFrame(dict(enumerate(cols)), index=index)
    result.columns = pd.Index(names) if flat else pd.MultiIndex.from_tuples(names)
    return result


def groupby_agg(df, by: Union[str, Sequence[str]], spec, sort: bool = False, dropna: bool = True,
                fallback: bool = True):
    """``df.groupby(by, sort=sort, dropna=dropna).agg(spec)`` with the GPU path
    first.  Returns ``(result, path)`` where path is ``"gpu"`` or ``"pandas"``."""
    by_list = [by] if isinstance(by, str) else list(by)
    norm, flat = _normalise_spec(spec)
    try:
        return _gpu(df, by_list, norm, flat, sort, dropna), "gpu"
    except Exception:
        if not fallback:
            raise
    return df.groupby(by_list if len(by_list) > 1 else by_list[0], sort=sort, dropna=dropna).agg(spec), "pandas"

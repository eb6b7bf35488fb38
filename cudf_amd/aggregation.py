"""Aggregation factories, shaped like pylibcudf.aggregation
(python/cudf/cudf/_lib/pylibcudf/aggregation.pyx:157-340) and numbered as
cudf::aggregation::Kind (cudf/aggregation.hpp:86-122)."""
from __future__ import annotations

import enum

from .types import NullPolicy


class Kind(enum.IntEnum):
    SUM = 0
    PRODUCT = 1
    MIN = 2
    MAX = 3
    COUNT_VALID = 4
    COUNT_ALL = 5
    ANY = 6
    ALL = 7
    SUM_OF_SQUARES = 8
    MEAN = 9
    M2 = 10
    VARIANCE = 11
    STD = 12
    MEDIAN = 13
    QUANTILE = 14
    ARGMAX = 15
    ARGMIN = 16
    NUNIQUE = 17
    NTH_ELEMENT = 18


class Interpolation(enum.IntEnum):
    """cudf::interpolation (cudf/types.hpp:192-198)."""
    LINEAR = 0
    LOWER = 1
    HIGHER = 2
    MIDPOINT = 3
    NEAREST = 4


class Aggregation:
    def __init__(self, kind: Kind, ddof: int = 1, null_handling: NullPolicy = NullPolicy.EXCLUDE, n: int = 0,
                 quantiles=(), interpolation: Interpolation = Interpolation.LINEAR):
        self._kind = Kind(kind)
        self.ddof = int(ddof)
        self.null_handling = NullPolicy(null_handling)  # NUNIQUE, NTH_ELEMENT
        self.n = int(n)  # NTH_ELEMENT
        self.quantiles = tuple(float(q) for q in quantiles)  # QUANTILE
        self.interpolation = Interpolation(interpolation)

    def to_c(self, keep: list):
        """The C ABI struct; `keep` holds buffers that must outlive the call."""
        from . import _lib
        import ctypes
        a = _lib.Aggregation(int(self._kind), self.ddof, int(self.null_handling), self.n, int(self.interpolation),
                             len(self.quantiles))
        if self.quantiles:
            qs = (ctypes.c_double * len(self.quantiles))(*self.quantiles)
            keep.append(qs)
            a.quantiles = ctypes.cast(qs, ctypes.POINTER(ctypes.c_double))
        return a

    def kind(self) -> Kind:
        return self._kind

    def __repr__(self) -> str:
        extra = f", ddof={self.ddof}" if self._kind in (Kind.VARIANCE, Kind.STD) else ""
        return f"Aggregation({self._kind.name}{extra})"


def sum() -> Aggregation:  # noqa: A001 - mirrors pylibcudf
    return Aggregation(Kind.SUM)


def product() -> Aggregation:
    return Aggregation(Kind.PRODUCT)


def min() -> Aggregation:  # noqa: A001
    return Aggregation(Kind.MIN)


def max() -> Aggregation:  # noqa: A001
    return Aggregation(Kind.MAX)


def count(null_handling: NullPolicy = NullPolicy.EXCLUDE) -> Aggregation:
    """COUNT_VALID under EXCLUDE, COUNT_ALL under INCLUDE (aggregation.cpp make_count_aggregation)."""
    return Aggregation(Kind.COUNT_VALID if null_handling == NullPolicy.EXCLUDE else Kind.COUNT_ALL)


def sum_of_squares() -> Aggregation:
    return Aggregation(Kind.SUM_OF_SQUARES)


def mean() -> Aggregation:
    return Aggregation(Kind.MEAN)


def variance(ddof: int = 1) -> Aggregation:
    return Aggregation(Kind.VARIANCE, ddof)


def std(ddof: int = 1) -> Aggregation:
    return Aggregation(Kind.STD, ddof)


def nunique(null_handling: NullPolicy = NullPolicy.EXCLUDE) -> Aggregation:
    """Distinct values per group; INCLUDE counts nulls as one value
    (aggregation.hpp make_nunique_aggregation).  A sort-based aggregation."""
    return Aggregation(Kind.NUNIQUE, null_handling=null_handling)


def nth_element(n: int, null_handling: NullPolicy = NullPolicy.INCLUDE) -> Aggregation:
    """The n-th value of each group in row order (from the end when n < 0);
    EXCLUDE counts valid values only (aggregation.hpp
    make_nth_element_aggregation).  A sort-based aggregation."""
    return Aggregation(Kind.NTH_ELEMENT, null_handling=null_handling, n=n)


def median() -> Aggregation:
    """The 0.5 quantile with linear interpolation, as double (aggregation.hpp
    make_median_aggregation)."""
    return Aggregation(Kind.MEDIAN)


def quantile(quantiles, interp: Interpolation = Interpolation.LINEAR) -> Aggregation:
    """Quantiles of each group's valid values, as double; the result holds
    num_groups * len(quantiles) rows, group-major (aggregation.hpp
    make_quantile_aggregation)."""
    return Aggregation(Kind.QUANTILE, quantiles=quantiles, interpolation=interp)


def argmax() -> Aggregation:
    return Aggregation(Kind.ARGMAX)


def argmin() -> Aggregation:
    return Aggregation(Kind.ARGMIN)

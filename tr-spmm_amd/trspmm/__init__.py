"""trspmm — three-way routed FP16/BF16 SpMM on AMD Instinct MI355X (gfx950).

Product path: the C ABI in libtrspmm.so (HIP kernels).  This package is the
Python host side: the device router, plans, the routed multiply, and the
workload generators.  There is no CPU fallback; without the built library or
a GPU every compute call raises.
"""
from . import autograd, graphs
from ._lib import (BF16, F16, MODE_COMPAT, MODE_EXACT, MODE_TUNED, ROUTE_ALL, ROUTE_RESIDUAL, ROUTE_SPTC,
                   ROUTE_TC, TrspmmError)
from .core import META_DTYPES, META_NAMES, Plan, RoutedMeta, route
from .sparse import RoutedSparse

__all__ = [
    "autograd", "graphs", "route", "Plan", "RoutedMeta", "RoutedSparse", "META_NAMES", "META_DTYPES", "TrspmmError",
    "F16", "BF16", "ROUTE_TC", "ROUTE_SPTC", "ROUTE_RESIDUAL", "ROUTE_ALL", "MODE_COMPAT", "MODE_EXACT",
    "MODE_TUNED",
]

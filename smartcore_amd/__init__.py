"""MI355X (gfx950) implementation of smartcore's distance-bound hot path.

``smartcore_amd._capi`` binds the C ABI (include/smartcore_amd.h) exported by
``libsmartcore_amd.so``; the estimator-level mirror of smartcore's API lives
in C++ (include/smartcore_amd.hpp) and is exposed to Python by
``smartcore_amd.smartcore`` once built.
"""
from . import _capi
from ._capi import (SC_EUCLIDIAN, SC_MANHATTAN, SC_MINKOWSKI, SC_MIXED, SC_PLAIN_FP64, DeviceError,
                    Failed, Index, Panic, kmeans_predict, last_error, last_stats)

__all__ = [
    "_capi", "Index", "Failed", "Panic", "DeviceError", "kmeans_predict", "last_error",
    "last_stats", "SC_EUCLIDIAN", "SC_MANHATTAN", "SC_MINKOWSKI", "SC_PLAIN_FP64", "SC_MIXED",
]

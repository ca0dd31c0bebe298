"""hbdia_amd -- the HBDIA sparse matrix-vector product for AMD Instinct MI355X.

Native libraries (built in-tree by `python hbdia_amd/build.py`):
  lib/libhbdia_hip.so  hand-written HIP kernels for gfx950 + the C ABI of include/hbdia_hip.h
  lib/libhbdia.so      the C++ drop-in API (HBDIA<T>, HBDIAVector<T>, hbdiaSpMV) and the
                       C handle API of include/hbdia_capi.h

Importing this package loads both libraries and raises NativeLibraryMissing
if either is absent; there is no CPU fallback.
"""
from ._native import HBDIAError, NativeLibraryMissing, gpu_available
from .api import (DEFAULT_BLOCK_WIDTH, DEFAULT_THRESHOLD, MAX_COO_ENTRIES, ExecutionMode, HBDIA, HBDIABatch, HBDIAMixed,
                  HBDIAServer, HBDIAVector, VectorArena,
                  gen_vector, hbdiaSpMV, hbdiaSpMVMulti, multi_ld, set_num_threads)
from .device import DeviceArray, Event, Graph, Stream, device_count, device_name, mem_info, synchronize

__all__ = [
    "HBDIA", "HBDIABatch", "HBDIAMixed", "HBDIAServer", "VectorArena", "HBDIAVector", "hbdiaSpMV", "hbdiaSpMVMulti", "multi_ld", "ExecutionMode", "DEFAULT_BLOCK_WIDTH", "DEFAULT_THRESHOLD",
    "MAX_COO_ENTRIES", "DeviceArray", "Stream", "Event", "Graph", "device_count", "device_name", "mem_info",
    "synchronize", "gen_vector", "set_num_threads", "gpu_available", "HBDIAError", "NativeLibraryMissing",
]

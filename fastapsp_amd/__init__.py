"""MI355X-native Fast-APSP: boundary SSSP -> per-part Floyd-Warshall -> min-plus merge.

The compute path is libfap_hip.so (hand-written HIP for gfx950 behind the C-ABI of
include/fap_hip.h); this package holds the ctypes bindings and a Python mirror of
the reference's fap::fapGraph API.
"""
from ._lib import (FAP_INF, FAP_MODE_BATCHED, FAP_MODE_PLAIN, FAP_MODE_OVERLAP, FAP_MODE_FLAG_INT32_FW, FAP_MODE_FLAG_SSSP,
                   FAP_MODE_FLAG_FW,
                   MAXVALUE,
                   DeviceBuffer, FapError, init, load)
from .batch import Batch, solve_batch
from .graph import FapGraph

__all__ = ["FapGraph", "Batch", "solve_batch", "FapError", "DeviceBuffer", "init", "load", "FAP_INF", "MAXVALUE",
           "FAP_MODE_PLAIN", "FAP_MODE_BATCHED", "FAP_MODE_OVERLAP", "FAP_MODE_FLAG_INT32_FW",
           "FAP_MODE_FLAG_SSSP", "FAP_MODE_FLAG_FW"]

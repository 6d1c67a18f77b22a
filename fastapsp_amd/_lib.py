"""ctypes bindings of libfap_hip.so (the C-ABI declared in include/fap_hip.h).

The library is loaded from fastapsp_amd/lib/libfap_hip.so.  There is no
fallback: if the library is missing, every call raises FapError.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional

import numpy as np

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib", "libfap_hip.so")

FAP_INF = 1073741823
MAXVALUE = np.float32(1e8)

FAP_OK = 0
FAP_ERR_INVALID = 1
FAP_ERR_NOMEM = 2
FAP_ERR_HIP = 3
FAP_ERR_BUDGET = 4
FAP_ERR_IO = 5
FAP_ERR_STATE = 6
FAP_ERR_NODEV = 7

FAP_MODE_PLAIN = 0
FAP_MODE_BATCHED = 1
FAP_MODE_OVERLAP = 2
FAP_MODE_FLAG_INT32_FW = 0x100
FAP_MODE_FLAG_SSSP = 0x200
FAP_MODE_FLAG_FW = 0x400
FAP_BATCH_GRAPH = 0x1
FAP_BATCH_NO_PACKED = 0x100000
FAP_BATCH_NO_BFS = 0x200000
FAP_BATCH_UNIT_GRID = 0x400000
FAP_BATCH_FW_TILE32 = 0x800000
FAP_BATCH_NO_NESTED = 0x1000000
FAP_ITEM_FW_TILE32 = 0x2

# Every exported symbol of include/fap_hip.h: name -> (restype, argtypes).
_i32, _i64, _u64, _f64 = C.c_int32, C.c_int64, C.c_uint64, C.c_double
_vp, _cp = C.c_void_p, C.c_char_p
_pi32 = C.POINTER(C.c_int32)
_pi64 = C.POINTER(C.c_int64)
_pf32 = C.POINTER(C.c_float)
_pf64 = C.POINTER(C.c_double)


class FapGraphDev(C.Structure):
    _fields_ = [("n", _i32), ("m", _i64), ("d_row_offset", _vp), ("d_col", _vp), ("d_w", _vp),
                ("d_graph_id", _vp), ("d_st2ed", _vp), ("d_ed2st", _vp), ("d_comp", _vp),
                ("d_wf", _vp)]


class FapPartDesc(C.Structure):
    _fields_ = [("part_id", _i32), ("offset", _i32), ("num", _i32), ("bdy", _i32),
                ("out_row", _i64)]


class FapBatchItem(C.Structure):
    _fields_ = [("g", FapGraphDev), ("parts", C.POINTER(FapPartDesc)), ("nparts", _i32),
                ("is_float", _i32), ("d_out", _vp), ("ld_out", _i64), ("flags", _i32)]


class FapServerStats(C.Structure):
    _fields_ = [("submitted", _i64), ("completed", _i64), ("failed", _i64), ("batches", _i64),
                ("reserved_bytes", _i64), ("peak_reserved_bytes", _i64), ("entries", _f64),
                ("span_s", _f64), ("p50_ms", _f64), ("p99_ms", _f64), ("mean_ms", _f64)]


class FapStageTimes(C.Structure):
    _fields_ = [("sssp_ms", C.c_float), ("build_ms", C.c_float), ("floyd_ms", C.c_float),
                ("minplus_ms", C.c_float), ("decode_ms", C.c_float), ("total_ms", C.c_float),
                ("parts_fw16", _i32), ("parts_redone", _i32), ("boundary_graph", _i32),
                ("parts_keyed", _i32), ("parts_fw_keyed", _i32), ("parts_bfs", _i32),
                ("merge_packed", _i32)]


SIGNATURES = {
    "fap_hip_init": (C.c_int, [C.c_int]),
    "fap_hip_last_error": (_cp, []),
    "fap_hip_device_info": (C.c_int, [C.POINTER(C.c_int), _pi64, _pi64, _cp, C.c_int]),
    "fap_hip_set_budget": (C.c_int, [_i64]),
    "fap_hip_mem_stats": (C.c_int, [_pi64, _pi64, _pi64]),
    "fap_hip_malloc": (C.c_int, [C.POINTER(_vp), _i64]),
    "fap_hip_free": (C.c_int, [_vp]),
    "fap_hip_memcpy_h2d": (C.c_int, [_vp, _vp, _i64, _vp]),
    "fap_hip_memcpy_d2h": (C.c_int, [_vp, _vp, _i64, _vp]),
    "fap_hip_memcpy_d2d": (C.c_int, [_vp, _vp, _i64, _vp]),
    "fap_hip_memcpy2d_d2h_i32": (C.c_int, [_vp, _i64, _vp, _i64, _i64, _i64, _vp]),
    "fap_hip_fill_i32": (C.c_int, [_vp, _i32, _i64, _vp]),
    "fap_hip_copy_i32": (C.c_int, [_vp, _vp, _i64, _vp]),
    "fap_hip_stream_sync": (C.c_int, [_vp]),
    "fap_hip_device_sync": (C.c_int, []),
    "fap_hip_event_create": (C.c_int, [C.POINTER(_vp)]),
    "fap_hip_event_destroy": (C.c_int, [_vp]),
    "fap_hip_event_record": (C.c_int, [_vp, _vp]),
    "fap_hip_event_elapsed_ms": (C.c_int, [_vp, _vp, _pf32]),
    "fap_hip_sssp_batch": (C.c_int, [_vp, _vp, _vp, _i32, _i64, _vp, _i32, _vp, _i64, _vp]),
    "fap_hip_floyd": (C.c_int, [_vp, _i32, _i64, _vp]),
    "fap_hip_floyd_noalloc": (C.c_int, [_vp, _i32, _i64, _vp]),
    "fap_hip_floyd_path": (C.c_int, [_vp, _vp, _i32, _i64, _vp]),
    "fap_hip_minplus_path": (C.c_int, [_vp, _i64, _vp, _i64, _vp, _i64, _vp, _i64, _vp, _i64,
                                       _i32, _i32, _i32, _vp]),
    "fap_hip_sssp_batch_path": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _i32, _i64, _vp, _i32,
                                          _vp, _vp, _i64, _vp]),
    "fap_hip_f32_to_i32": (C.c_int, [_vp, _vp, _i64, _vp]),
    "fap_hip_check_layout": (C.c_int, [_vp, _vp, _i32, C.POINTER(_i32)]),
    "fap_hip_i32_to_f32": (C.c_int, [_vp, _vp, _i64, _vp]),
    "fap_hip_minplus": (C.c_int, [_vp, _i64, _vp, _i64, _vp, _i64, _i32, _i32, _i32, _vp]),
    "fap_hip_submat_build": (C.c_int, [_vp, _i64, _vp, _i64, _vp, _vp, _vp, _vp, _vp, _vp,
                                       _i32, _i32, _i32, _i32, _vp]),
    "fap_hip_submat_decode": (C.c_int, [_vp, _i64, _vp, _i64, _vp, _i32, _i32, _i32, _vp]),
    "fap_hip_solve_parts": (C.c_int, [C.POINTER(FapGraphDev), C.POINTER(FapPartDesc), _i32,
                                      _vp, _i64, _i32, _vp]),
    "fap_hip_last_stage_times": (C.c_int, [C.POINTER(FapStageTimes)]),
    "fap_hip_set_merge_shape": (C.c_int, [C.c_int32, C.c_int32, C.c_int32]),
    "fap_hip_get_merge_shape": (C.c_int, [C.POINTER(C.c_int32)] * 3),
    "fap_hip_write_probe": (C.c_int, [_vp, _i64, _i32, _i32, _i32, _vp, C.POINTER(C.c_float)]),
    "fap_hip_malloc_placed": (C.c_int, [C.POINTER(_vp), _i64, _i64, _i32, C.POINTER(C.c_float)]),
    "fap_hip_solve_parts_path": (C.c_int, [C.POINTER(FapGraphDev), _vp, _vp, _vp,
                                           C.POINTER(FapPartDesc), _i32, _vp, _vp, _i64, _vp]),
    "fap_hip_solve_parts_f32": (C.c_int, [C.POINTER(FapGraphDev), C.POINTER(FapPartDesc), _i32,
                                          _vp, _i64, _i32, _vp]),
    "fap_hip_batch_create": (C.c_int, [C.POINTER(_vp), C.POINTER(FapBatchItem), _i32, _i32]),
    "fap_hip_batch_run": (C.c_int, [_vp, _vp]),
    "fap_hip_batch_nested_parts": (C.c_int, [_vp]),
    "fap_hip_batch_destroy": (C.c_int, [_vp]),
    "fap_hip_solve_batch": (C.c_int, [C.POINTER(FapBatchItem), _i32, _i32, _vp]),
    "fap_hip_event_query": (C.c_int, [_vp, _pi32]),
    "fap_server_create": (C.c_int, [C.POINTER(_vp), _i64, _i32]),
    "fap_server_destroy": (C.c_int, [_vp]),
    "fap_server_add_graph": (C.c_int, [_vp, _vp, _pi32]),
    "fap_server_submit": (C.c_int, [_vp, _i32, _i32, _pi64]),
    "fap_server_wait": (C.c_int, [_vp, _i64, _f64, _pi32]),
    "fap_server_result": (C.c_int, [_vp, _i64, C.POINTER(_vp), _pi64, _pi64, _pi64, _pi32]),
    "fap_server_release": (C.c_int, [_vp, _i64]),
    "fap_server_get_stats": (C.c_int, [_vp, C.POINTER(FapServerStats)]),
    "fap_hip_stream_create": (C.c_int, [C.POINTER(_vp)]),
    "fap_hip_stream_destroy": (C.c_int, [_vp]),
    "fap_graph_batch_create": (C.c_int, [C.POINTER(_vp), C.POINTER(_vp), _i32, _i32]),
    "fap_graph_batch_create_ex": (C.c_int, [C.POINTER(_vp), C.POINTER(_vp), _i32, _i32,
                                            C.POINTER(_i32)]),
    "fap_graph_solve_batch": (C.c_int, [C.POINTER(_vp), _i32, _i32]),
    "fap_graph_create_from_file": (C.c_int, [C.POINTER(_vp), _cp, C.c_int, C.c_int, _i32, _cp]),
    "fap_mtx_read": (C.c_int, [_cp, C.c_int, C.c_int, _pi32, _pi64, _pi32, _pi32, _pf32]),
    "fap_graph_create_from_csr": (C.c_int, [C.POINTER(_vp), _i32, _i64, _pi32, _pi32, _pf32,
                                            C.c_int, _i32]),
    "fap_graph_generate": (C.c_int, [C.POINTER(_vp), _i32, _i32, _i32, _i32, _f64, _f64, _u64,
                                     _i32]),
    "fap_graph_destroy": (C.c_int, [_vp]),
    "fap_graph_set_partition": (C.c_int, [_vp, _pi32]),
    "fap_layout_for_partition": (C.c_int, [_i32, _pi32, _pi32, _i32, _pi32, _pi32, _pi32, _pi32,
                                           _pi32, _pi32]),
    "fap_graph_precondition": (C.c_int, [_vp]),
    "fap_graph_is_split": (C.c_int, [_vp, C.POINTER(C.c_int)]),
    "fap_graph_set_mode": (C.c_int, [_vp, _i32]),
    "fap_graph_solve": (C.c_int, [_vp]),
    "fap_graph_set_subgraph_sweep": (C.c_int, [_vp, _i32, _pi32]),
    "fap_hip_prepare_parts": (C.c_int, [_vp, _vp, _i32, _vp, _i64, _vp, _pi64, _pi64, _vp]),
    "fap_hip_merge_prepared": (C.c_int, [_vp, _vp, _vp, _vp, _i64, _i64, _vp, _i64, _vp]),
    "fap_hip_count_reached": (C.c_int, [_vp, _vp, _i32, _vp, _i64, _pi32, _vp]),
    "fap_graph_solve_subgraph": (C.c_int, [_vp, _i32]),
    "fap_graph_solve_ex": (C.c_int, [_vp, _i32]),
    "fap_graph_solve_subgraph_ex": (C.c_int, [_vp, _i32, _i32]),
    "fap_graph_get_path_rows": (C.c_int, [_vp, _i64, _i64, _pi32]),
    "fap_graph_get_subgraph_path_rows": (C.c_int, [_vp, _i32, _i64, _i64, _pi32]),
    "fap_graph_sizes": (C.c_int, [_vp, _pi32, _pi64, _pi32]),
    "fap_graph_get_csr": (C.c_int, [_vp, _pi32, _pi32, _pi32]),
    "fap_graph_get_layout": (C.c_int, [_vp, _pi32, _pi32, _pi32, _pi32, _pi32, _pi32]),
    "fap_graph_get_dist_rows": (C.c_int, [_vp, _i64, _i64, _pi32]),
    "fap_graph_get_subgraph_dist": (C.c_int, [_vp, _i32, _pi32]),
    "fap_graph_get_subgraph_rows": (C.c_int, [_vp, _i32, _i64, _i64, _pi32]),
    "fap_graph_is_float": (C.c_int, [_vp, _pi32]),
    "fap_graph_boundary_by_sssp": (C.c_int, [_vp, _pi32]),
    "fap_graph_get_weights_f32": (C.c_int, [_vp, _pf32]),
    "fap_graph_get_dist_rows_f32": (C.c_int, [_vp, _i64, _i64, _pf32]),
    "fap_graph_get_subgraph_rows_f32": (C.c_int, [_vp, _i32, _i64, _i64, _pf32]),
    "fap_graph_device_dist": (C.c_int, [_vp, C.POINTER(_vp), _pi64]),
    "fap_graph_times": (C.c_int, [_vp, _pf64, _pf64, _pf64, _pf64, _pf64]),
    "fap_graph_edge_cut": (C.c_int, [_vp, _pf64]),
}


class FapError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"[status {status}] {message}")
        self.status = status


_lib: Optional[C.CDLL] = None


def load() -> C.CDLL:
    """Loads libfap_hip.so once; raises FapError when it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FapError(FAP_ERR_STATE, f"native library missing: {LIB_PATH} "
                                      "(build it with __graft_entry__.build())")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def last_error() -> str:
    msg = load().fap_hip_last_error()
    return msg.decode() if msg else ""


def check(status: int, what: str = "") -> None:
    if status != FAP_OK:
        raise FapError(status, f"{what}: {last_error()}")


_initialised = False


def init(device: int = 0) -> None:
    global _initialised
    if not _initialised:
        check(load().fap_hip_init(device), "fap_hip_init")
        _initialised = True


def last_stage_times() -> dict:
    """fap_hip_last_stage_times of the last solve: per-stage device ms, the total, and
    the counts of parts run in the 16-bit Floyd-Warshall form and redone in int32."""
    t = FapStageTimes()
    check(load().fap_hip_last_stage_times(C.byref(t)), "stage times")
    return {k: getattr(t, k) for k, _ in FapStageTimes._fields_}


def mem_stats() -> dict:
    """fap_hip_mem_stats: device bytes in use through the library, the high-water mark
    and the budget (0: none)."""
    u, h, b = C.c_int64(), C.c_int64(), C.c_int64()
    check(load().fap_hip_mem_stats(C.byref(u), C.byref(h), C.byref(b)), "mem_stats")
    return {"in_use": u.value, "high_water": h.value, "budget": b.value}


def ptr_i32(a: np.ndarray):
    assert a.dtype == np.int32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_pi32)


class DeviceBuffer:
    """Device memory from fap_hip_malloc (budgeted), freed on close()/GC."""

    def __init__(self, nbytes: int):
        init()
        self.nbytes = int(nbytes)
        p = _vp()
        check(load().fap_hip_malloc(C.byref(p), max(self.nbytes, 4)), "fap_hip_malloc")
        self.ptr = p.value

    @classmethod
    def from_array(cls, a: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        if a.nbytes:
            check(load().fap_hip_memcpy_h2d(b.ptr, a.ctypes.data, a.nbytes, None), "h2d")
        return b

    def upload(self, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        if a.nbytes:
            check(load().fap_hip_memcpy_h2d(self.ptr, a.ctypes.data, a.nbytes, None), "h2d")

    def download(self, shape, dtype=np.int32) -> np.ndarray:
        out = np.empty(shape, dtype=dtype)
        assert out.nbytes <= self.nbytes
        check(load().fap_hip_stream_sync(None), "sync")
        if out.nbytes:
            check(load().fap_hip_memcpy_d2h(out.ctypes.data, self.ptr, out.nbytes, None), "d2h")
        return out

    def fill_i32(self, value: int, count: int) -> None:
        check(load().fap_hip_fill_i32(self.ptr, int(value), int(count), None), "fill")

    def close(self) -> None:
        if self.ptr:
            load().fap_hip_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

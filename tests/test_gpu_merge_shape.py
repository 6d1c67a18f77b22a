"""The streaming merge under every kind of work-unit shape (fap_hip_set_merge_shape:
rows per unit, unit order, resident-wave cap) gives the same bit-exact int32 result:
every entry of a 20,000-vertex road-like solve in the batched form against exact BFS.
The placement-probed result allocation (fap_hip_malloc_placed) keeps one usable buffer."""
import ctypes as C

import numpy as np
import pytest

import fastapsp_amd as F
import oracle as O
from fastapsp_amd import _lib as L

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def road():
    g = F.FapGraph.road_like(20000, 0.52, 1e-4, 3, 32)
    g.preCondition()
    g.setMode(F.FAP_MODE_BATCHED)
    return g


@pytest.mark.parametrize("shape", [(128, 0, 0), (128, 1, 0), (8, 1, 4), (32, 0, 8), (256, 1, 16),
                                   (24, 0, 0)])
def test_merge_shape_exact(road, shape):
    lib = L.load()
    vals = [C.c_int32(), C.c_int32(), C.c_int32()]
    L.check(lib.fap_hip_get_merge_shape(*[C.byref(v) for v in vals]), "get shape")
    default = tuple(v.value for v in vals)
    L.check(lib.fap_hip_set_merge_shape(*shape), "set shape")
    try:
        L.check(lib.fap_hip_fill_i32(road.deviceDistance()[0], 777,
                                     road.num_vertexs * road.deviceDistance()[1], None), "fill")
        road.solve()
        st = L.last_stage_times()
        assert st["boundary_graph"] == 1
        got = road.getDistanceRowsI32(0, road.num_vertexs)
    finally:
        L.check(lib.fap_hip_set_merge_shape(*default), "reset shape")
    csr, lay = road.csr(), road.layout()
    bad, first = O.compare_exact_rows(csr["row_offset"], csr["col"], None, lay["st2ed"], got)
    assert bad == 0, f"shape {shape}: {bad} rows differ (first {first})"


def test_malloc_placed_keeps_one_buffer_and_frees_the_rest():
    """fap_hip_malloc_placed holds up to three candidates while it probes them, keeps one
    (usable, accounted like fap_hip_malloc) and frees the others; with a budget that
    leaves room for one buffer only it still returns that one."""
    lib = L.load()
    L.init()
    rows, ld = 20000, 20032
    nbytes = rows * ld * 4
    in_use0 = L.mem_stats()["in_use"]
    p, ms = C.c_void_p(), C.c_float()
    L.check(lib.fap_hip_malloc_placed(C.byref(p), rows, ld, 3, C.byref(ms)), "placed")
    try:
        assert p.value and ms.value > 0
        assert L.mem_stats()["in_use"] == in_use0 + nbytes
        L.check(lib.fap_hip_fill_i32(p, 5, rows * ld, None), "fill")
        got = np.empty((3, 40), np.int32)
        L.check(lib.fap_hip_memcpy2d_d2h_i32(L.ptr_i32(got), 40, p.value + (rows - 3) * ld * 4,
                                              ld, 3, 40, None), "rows")
        assert np.all(got == 5)
    finally:
        L.check(lib.fap_hip_free(p), "free")
    assert L.mem_stats()["in_use"] == in_use0
    L.check(lib.fap_hip_set_budget(in_use0 + nbytes + (64 << 20)), "budget")
    try:
        q = C.c_void_p()
        L.check(lib.fap_hip_malloc_placed(C.byref(q), rows, ld, 3, None), "placed under budget")
        assert q.value and L.mem_stats()["in_use"] == in_use0 + nbytes
        L.check(lib.fap_hip_free(q), "free")
    finally:
        L.check(lib.fap_hip_set_budget(0), "budget off")

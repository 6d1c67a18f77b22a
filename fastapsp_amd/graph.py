"""Python mirror of fap::fapGraph (fap/fap.h:25-162 of TomerHaham/FastAPSP).

Same method names, argument meaning and error behaviour as the reference class:
construct from a .mtx file name (opened as graph/<(un)weight>-<(un)directed>/
<name>.mtx relative to the working directory, fap/utils/readMatFile.h:43-51), a
CSR, or a seeded synthetic generator; then preCondition(), isSplit(), solve() /
solveSubGraph(id) and the getters.  Every call goes through the C-ABI of
libfap_hip.so; results stay in HBM until a getter copies them out.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional

import numpy as np

from . import _lib as L


class FapGraph:
    def __init__(self, input_graph: Optional[str] = None, directed: bool = False,
                 weighted: bool = False, K: int = 8, partitioner: str = "metis",
                 _handle: Optional[int] = None):
        lib = L.load()
        self._h = None
        if _handle is not None:
            self._h = _handle
        else:
            if input_graph is None:
                raise ValueError("input_graph is required")
            h = C.c_void_p()
            L.check(lib.fap_graph_create_from_file(C.byref(h), input_graph.encode(),
                                                   int(directed), int(weighted), int(K),
                                                   partitioner.encode()), "fapGraph")
            self._h = h.value
        n, m, k = C.c_int32(), C.c_int64(), C.c_int32()
        L.check(lib.fap_graph_sizes(self._h, C.byref(n), C.byref(m), C.byref(k)), "sizes")
        self.num_vertexs, self.num_edges, self.K = n.value, m.value, k.value

    # ---- construction -----------------------------------------------------
    @classmethod
    def from_csr(cls, row_offset, col_val, weight=None, directed: bool = False,
                 K: int = 8) -> "FapGraph":
        """fap/fap.h:102-104 (num_vertexs, num_edges, row_offset, col_val, weight, K)."""
        ro = np.ascontiguousarray(row_offset, dtype=np.int32)
        cv = np.ascontiguousarray(col_val, dtype=np.int32)
        n, m = len(ro) - 1, len(cv)
        wp = None
        if weight is not None:
            w = np.ascontiguousarray(weight, dtype=np.float32)
            wp = w.ctypes.data_as(C.POINTER(C.c_float))
        h = C.c_void_p()
        L.check(L.load().fap_graph_create_from_csr(
            C.byref(h), n, m, L.ptr_i32(ro), L.ptr_i32(cv), wp, int(directed), int(K)),
            "fapGraph(csr)")
        return cls(_handle=h.value)

    @classmethod
    def road_grid(cls, W: int, H: int, p_drop: float, p_diag: float, seed: int,
                  K: int) -> "FapGraph":
        """W x H grid, grid edges dropped with p_drop, down-right diagonals with p_diag, LCC."""
        return cls._generate(0, W, H, 0, p_drop, p_diag, seed, K)

    @classmethod
    def road_like(cls, n_target: int, p_keep: float, chord_rate: float, seed: int,
                  K: int) -> "FapGraph":
        """Percolated grid (keep p_keep), chords, LCC trimmed to exactly n_target vertices."""
        return cls._generate(1, 0, 0, n_target, p_keep, chord_rate, seed, K)

    @classmethod
    def knn(cls, n: int, k: int, wmax: int, seed: int, K: int) -> "FapGraph":
        """Directed k-NN graph of n uniform points, integer weights in [1, wmax]."""
        return cls._generate(2, k, wmax, n, 0.0, 0.0, seed, K)

    @classmethod
    def _generate(cls, kind, W, H, n_target, pa, pb, seed, K) -> "FapGraph":
        h = C.c_void_p()
        L.check(L.load().fap_graph_generate(C.byref(h), kind, W, H, n_target, pa, pb, seed, K),
                "fapGraph(generate)")
        return cls(_handle=h.value)

    def close(self) -> None:
        if self._h:
            L.load().fap_graph_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- the reference API ------------------------------------------------
    def setPartition(self, graph_id) -> None:
        """Use a caller's 1-based partition (e.g. METIS output) before preCondition()."""
        gid = np.ascontiguousarray(graph_id, dtype=np.int32)
        if gid.shape != (self.num_vertexs,):
            raise ValueError("partition must have one id per vertex")
        L.check(L.load().fap_graph_set_partition(self._h, L.ptr_i32(gid)), "setPartition")

    def preCondition(self) -> None:
        L.check(L.load().fap_graph_precondition(self._h), "preCondition")
        self._preconditioned = True

    def isSplit(self) -> bool:
        v = C.c_int()
        L.check(L.load().fap_graph_is_split(self._h, C.byref(v)), "isSplit")
        return bool(v.value)

    def setMode(self, mode: int) -> None:
        L.check(L.load().fap_graph_set_mode(self._h, int(mode)), "setMode")

    def setSubGraphSweep(self, after_calls: int) -> None:
        """Distance-only solveSubGraph calls on integer-weight graphs: from the
        after_calls-th call on (default 3, 0 = never) every part's boundary rows and
        Floyd-Warshall matrix come from one batched pass and each call writes only its
        block (identical results)."""
        L.check(L.load().fap_graph_set_subgraph_sweep(self._h, int(after_calls), None),
                "setSubGraphSweep")

    def subGraphSweepState(self) -> int:
        """0 = not prepared, 1 = prepared (dense product), 2 = prepared (sparse), 3 = blocks
        from a full solve() (weighted graphs whose blocks come from the boundary graph)."""
        v = C.c_int32()
        L.check(L.load().fap_graph_set_subgraph_sweep(self._h, -1, C.byref(v)),
                "subGraphSweepState")
        return int(v.value)

    def subGraphSweepReady(self) -> bool:
        return self.subGraphSweepState() != 0

    def solve(self, is_path_needed: bool = False) -> int:
        """fap/fap.h:127.  With is_path_needed the predecessor matrix is produced too
        (integer-weight graphs; the reference CPU path's tie-breaking)."""
        L.check(L.load().fap_graph_solve_ex(self._h, int(bool(is_path_needed))), "solve")
        return 0

    def solveSubGraph(self, sub_graph_id: int, is_path_needed: bool = False) -> int:
        L.check(L.load().fap_graph_solve_subgraph_ex(self._h, int(sub_graph_id),
                                                      int(bool(is_path_needed))),
                "solveSubGraph")
        self._current = int(sub_graph_id)
        return 0

    def getPathRows(self, row0: int, rows: int) -> np.ndarray:
        """Rows [row0, row0 + rows) of the predecessors of solve(True): path[r][v] is the
        vertex before v on the shortest path from st2ed[r], -1 if unreachable."""
        out = np.empty((rows, self.num_vertexs), dtype=np.int32)
        L.check(L.load().fap_graph_get_path_rows(self._h, int(row0), int(rows), L.ptr_i32(out)),
                "getPathRows")
        return out

    def getAllPath(self) -> np.ndarray:
        return self.getPathRows(0, self.num_vertexs)

    def getSubGraphPath(self, sub_graph_id: int) -> np.ndarray:
        num = int(self.layout()["num"][sub_graph_id])
        out = np.empty((num, self.num_vertexs), dtype=np.int32)
        L.check(L.load().fap_graph_get_subgraph_path_rows(self._h, int(sub_graph_id), 0, num,
                                                          L.ptr_i32(out)), "getSubGraphPath")
        return out

    def getGraphId(self) -> np.ndarray:
        gid = np.empty(self.num_vertexs, dtype=np.int32)
        L.check(L.load().fap_graph_get_layout(self._h, L.ptr_i32(gid), None, None, None, None,
                                              None), "getGraphId")
        return gid

    def layout(self) -> dict:
        """graph_id, st2ed, ed2st and the per-part num / offset / bdy arrays (index 0 unused)."""
        n, k = self.num_vertexs, self.K
        out = {name: np.empty(size, dtype=np.int32) for name, size in
               [("graph_id", n), ("st2ed", n), ("ed2st", n), ("num", k + 1), ("offset", k + 1),
                ("bdy", k + 1)]}
        L.check(L.load().fap_graph_get_layout(
            self._h, *[L.ptr_i32(out[x]) for x in
                       ("graph_id", "st2ed", "ed2st", "num", "offset", "bdy")]), "layout")
        return out

    def csr(self) -> dict:
        n, m = self.num_vertexs, self.num_edges
        ro = np.empty(n + 1, dtype=np.int32)
        col = np.empty(max(m, 1), dtype=np.int32)
        w = np.empty(max(m, 1), dtype=np.int32)
        L.check(L.load().fap_graph_get_csr(self._h, L.ptr_i32(ro), L.ptr_i32(col), L.ptr_i32(w)),
                "csr")
        if self.isFloat():  # float32 weights as stored
            w = np.empty(max(m, 1), dtype=np.float32)
            L.check(L.load().fap_graph_get_weights_f32(
                self._h, w.ctypes.data_as(C.POINTER(C.c_float))), "weights")
        return {"row_offset": ro, "col": col[:m], "weight": w[:m]}

    def getMapping(self) -> np.ndarray:
        return self.layout()["st2ed"]

    def getDistanceRowsI32(self, row0: int, rows: int) -> np.ndarray:
        out = np.empty((rows, self.num_vertexs), dtype=np.int32)
        L.check(L.load().fap_graph_get_dist_rows(self._h, int(row0), int(rows), L.ptr_i32(out)),
                "getDistanceRows")
        return out

    def getAllDistanceI32(self) -> np.ndarray:
        return self.getDistanceRowsI32(0, self.num_vertexs)

    def boundaryBySssp(self) -> bool:
        """True when preCondition found parallel arcs inside a part whose later arc (CSR
        order) is heavier; solve() then computes the boundary rows by SSSP, as the
        reference does (its sub-matrix build keeps the later arc)."""
        v = C.c_int32()
        L.check(L.load().fap_graph_boundary_by_sssp(self._h, C.byref(v)), "boundaryBySssp")
        return bool(v.value)

    def isFloat(self) -> bool:
        """True when a weight is not an integer below 2^24: solves run the float path
        (FP64 arithmetic, float32 results) and the int32 getters raise."""
        v = C.c_int32()
        L.check(L.load().fap_graph_is_float(self._h, C.byref(v)), "isFloat")
        return bool(v.value)

    def getDistanceRowsF32(self, row0: int, rows: int) -> np.ndarray:
        out = np.empty((rows, self.num_vertexs), dtype=np.float32)
        L.check(L.load().fap_graph_get_dist_rows_f32(
            self._h, int(row0), int(rows), out.ctypes.data_as(C.POINTER(C.c_float))),
            "getDistanceRowsF32")
        return out

    def getAllDistance(self) -> np.ndarray:
        """float32 n x n, unreachable = MAXVALUE (1e8) as in the reference."""
        return self.getDistanceRowsF32(0, self.num_vertexs)

    def getSubGraphIndex(self, sub_graph_id: int) -> np.ndarray:
        lay = self.layout()
        off, num = lay["offset"][sub_graph_id], lay["num"][sub_graph_id]
        return lay["st2ed"][off:off + num].copy()

    def getSubGraphDistanceI32(self, sub_graph_id: int) -> np.ndarray:
        num = int(self.layout()["num"][sub_graph_id])
        out = np.empty((num, self.num_vertexs), dtype=np.int32)
        L.check(L.load().fap_graph_get_subgraph_dist(self._h, int(sub_graph_id), L.ptr_i32(out)),
                "getSubGraphDistance")
        return out

    def getSubGraphDistanceRowsI32(self, sub_graph_id: int, row0: int, rows: int) -> np.ndarray:
        """Rows [row0, row0 + rows) of the last solveSubGraph block."""
        out = np.empty((rows, self.num_vertexs), dtype=np.int32)
        L.check(L.load().fap_graph_get_subgraph_rows(self._h, int(sub_graph_id), int(row0),
                                                     int(rows), L.ptr_i32(out)),
                "getSubGraphDistanceRows")
        return out

    def getSubGraphDistanceRowsF32(self, sub_graph_id: int, row0: int, rows: int) -> np.ndarray:
        out = np.empty((rows, self.num_vertexs), dtype=np.float32)
        L.check(L.load().fap_graph_get_subgraph_rows_f32(
            self._h, int(sub_graph_id), int(row0), int(rows),
            out.ctypes.data_as(C.POINTER(C.c_float))), "getSubGraphDistanceRowsF32")
        return out

    def getSubGraphDistance(self, sub_graph_id: int) -> np.ndarray:
        """float32 s_p x n block, unreachable = MAXVALUE (1e8)."""
        num = int(self.layout()["num"][sub_graph_id])
        return self.getSubGraphDistanceRowsF32(sub_graph_id, 0, num)

    def getCurrentSubGraphId(self) -> int:
        return getattr(self, "_current", 0)

    def deviceDistance(self):
        p, ld = C.c_void_p(), C.c_int64()
        L.check(L.load().fap_graph_device_dist(self._h, C.byref(p), C.byref(ld)), "device_dist")
        return p.value, ld.value

    def _times(self):
        t = [C.c_double() for _ in range(5)]
        L.check(L.load().fap_graph_times(self._h, *[C.byref(x) for x in t]), "times")
        return [x.value for x in t]

    def getTimePartitioning(self) -> float:
        return self._times()[0]

    def getTimeSSSP(self) -> float:
        return self._times()[1]

    def getTimeFloyd(self) -> float:
        return self._times()[2]

    def getTimeMinPlus(self) -> float:
        return self._times()[3]

    def getTimeData(self) -> float:
        return self._times()[4]

    def getEdgeCut(self) -> float:
        v = C.c_double()
        L.check(L.load().fap_graph_edge_cut(self._h, C.byref(v)), "edge_cut")
        return v.value

    def getBoundaryNum(self) -> int:
        return int(self.layout()["bdy"][1:].sum())

    def getInnerNum(self) -> int:
        lay = self.layout()
        return int((lay["num"][1:] - lay["bdy"][1:]).sum())

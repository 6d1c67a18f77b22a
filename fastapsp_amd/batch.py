"""Many independent graphs solved together (the batched form for many graphs).

`Batch(graphs)` prepares the work tables over every part of every graph (after each
graph's preCondition) and, with `use_graph=True`, captures the launch sequence as a
hipGraph; `run()` replays it: per distance type, each stage (build, Floyd-Warshall,
boundary graph, boundary rows, merge) is one launch over all graphs, the merge a
persistent grid pulling units from a device counter.  Results land in each graph's solve()
buffer, so the graphs' own getters return them.  `solve_batch(graphs, mode)` is the
one-shot form; FAP_MODE_PLAIN there solves graph after graph with per-part launches
(the reference's structure) for comparison.
"""
from __future__ import annotations

import ctypes as C
from typing import Sequence

from . import _lib as L


def _handles(graphs) -> C.Array:
    arr = (C.c_void_p * len(graphs))()
    for i, g in enumerate(graphs):
        arr[i] = g._h
    return arr


_ROWS_CODE = {0: 0, 8: 1, 16: 2, 32: 3}


class Batch:
    """per_cu: persistent merge workgroups per CU (0: as many as stay resident);
    src_rows: boundary sources per unit of the boundary-row product (8, 16, 32; 0 =
    16); merge_rows: inner rows per wave of the merge (8 or 16; 0 = 16, FP64 always 8);
    packed: 1 = the int32 merge of unit-weight graphs (n <= 15872) on packed 16-bit
    distances, 0 = on int32; bfs: 1 = in-part distances of unit-weight symmetric graphs
    by bit-parallel BFS, 0 = by Floyd-Warshall; unit_grid: 1 = the merge as a plain grid
    of one workgroup per unit, 0 = the persistent work queue; fw_tile: Floyd-Warshall tile
    edge of the float-weight (FP64) parts and boundary graphs, 64 or 32, for every graph
    or a list with one entry per graph (the autotuner's choice per size class); nested:
    1 = in-part distances of weighted parts of 160+ vertices by the nested closure
    (sub-parts of at most 64 vertices, their crossing graph, two min-plus passes), 0 = by
    Floyd-Warshall on the whole part (the same matrices)."""

    def __init__(self, graphs: Sequence, use_graph: bool = True, per_cu: int = 0,
                 src_rows: int = 0, merge_rows: int = 0, packed: int = 1, bfs: int = 1,
                 unit_grid: int = 0, fw_tile=64, nested: int = 1):
        self.graphs = list(graphs)
        for g in self.graphs:
            if not getattr(g, "_preconditioned", False):
                g.preCondition()
        self._h = None
        h = C.c_void_p()
        tiles = list(fw_tile) if isinstance(fw_tile, (list, tuple)) else [fw_tile] * len(self.graphs)
        if len(tiles) != len(self.graphs) or any(int(t) not in (32, 64) for t in tiles):
            raise ValueError("fw_tile: 64 or 32, for every graph or one per graph")
        item_flags = (C.c_int32 * max(len(tiles), 1))(
            *[L.FAP_ITEM_FW_TILE32 if int(t) == 32 else 0 for t in tiles])
        L.check(L.load().fap_graph_batch_create_ex(C.byref(h), _handles(self.graphs),
                                                len(self.graphs),
                                                (L.FAP_BATCH_GRAPH if use_graph else 0)
                                                | ((int(per_cu) & 0xFF) << 8)
                                                | (_ROWS_CODE[int(src_rows)] << 16)
                                                | (_ROWS_CODE[int(merge_rows)] << 18)
                                                | (0 if packed else L.FAP_BATCH_NO_PACKED)
                                                | (0 if bfs else L.FAP_BATCH_NO_BFS)
                                                | (L.FAP_BATCH_UNIT_GRID if unit_grid else 0)
                                                | (0 if nested else L.FAP_BATCH_NO_NESTED),
                                                item_flags),
                "batch create")
        self._h = h.value

    def run(self, sync: bool = True) -> None:
        """Run (or replay) the batch on the library stream."""
        L.check(L.load().fap_hip_batch_run(self._h, None), "batch run")
        if sync:
            L.check(L.load().fap_hip_stream_sync(None), "batch sync")

    def nested_parts(self) -> int:
        """Parts whose in-part distances come from the nested closure."""
        r = L.load().fap_hip_batch_nested_parts(self._h)
        if r < 0:
            L.check(r, "batch nested parts")
        return r

    def close(self) -> None:
        if self._h is not None:
            L.load().fap_hip_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def solve_batch(graphs: Sequence, mode: int = L.FAP_MODE_BATCHED) -> None:
    for g in graphs:
        if not getattr(g, "_preconditioned", False):
            g.preCondition()
    L.check(L.load().fap_graph_solve_batch(_handles(graphs), len(graphs), int(mode)),
            "solve batch")

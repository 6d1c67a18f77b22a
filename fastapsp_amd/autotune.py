"""Autotuner of the many-graph batch form.

For each size class of graphs (vertex counts [256, 512), [512, 1024), ..., [4096, 8192])
it times candidate settings on a batch that fills the GPU -- copies of a road-like
graph of the class's middle size, half of them with float weights -- and keeps the
fastest:
  * K, the number of parts per graph: it sets the Floyd-Warshall block size s_p (work
    ~ K * s_p^3 and (sum of b_p)^3 for the boundary graph), the boundary count (the
    merge depth) and the number of independent work units;
  * the boundary-source batch width: boundary-graph rows per unit of the boundary-row
    product (8, 16 or 32 sources sharing one pass over a part's in-part block);
  * the merge row tile: inner rows per wave of the int32 merge (8 or 16);
  * the persistent merge workgroups per CU (0 = as many as stay resident);
  * the packed 16-bit merge of unit-weight graphs, on or off;
  * the merge as the persistent work queue or as a plain grid of one workgroup per unit
    (with both distance types in a batch the two launch sequences run side by side);
  * the Floyd-Warshall block size of the float-weight parts: 64 x 64 or 32 x 32 tiles
    (matrices padded to multiples of the tile: 32 pads small parts less, 64 reuses each
    staged panel for more steps).
K is chosen first with the default tiles, then each tile knob in turn (coordinate
search).  The result is a small table (JSON-serialisable) that `k_for(n)` and
`settings_for(ns)` read; `load_or_tune(path)` caches it on disk.
"""
from __future__ import annotations

import json
import math
import os
import time

import numpy as np

from . import _lib as L
from .batch import Batch
from .graph import FapGraph

SIZE_CLASSES = [(256, 512), (512, 1024), (1024, 2048), (2048, 4096), (4096, 8193)]
K_CANDIDATES = (2, 4, 8, 16, 32)
PER_CU_CANDIDATES = (0, 2)
SRC_ROWS_CANDIDATES = (8, 16, 32)
MERGE_ROWS_CANDIDATES = (8, 16)
PACKED_CANDIDATES = (1, 0)
UNIT_GRID_CANDIDATES = (0, 1)
FW_TILE_CANDIDATES = (64, 32)
MAX_PART = 32768  # largest part of the batch form
KNOBS = ("per_cu", "src_rows", "merge_rows", "packed", "unit_grid", "fw_tile")


def symmetric_float_weights(ro, col, seed: int, directed: bool = False) -> np.ndarray:
    """Deterministic float32 weights uniform in [0.5, 10) from a hash of the arc's end
    points (of the unordered pair when undirected, so both arcs of an edge agree)."""
    n = len(ro) - 1
    src = np.repeat(np.arange(n, dtype=np.uint64), np.diff(ro))
    dst = np.asarray(col).astype(np.uint64)
    a, b = (src, dst) if directed else (np.minimum(src, dst), np.maximum(src, dst))
    key = (a * np.uint64(n) + b) ^ np.uint64((seed * 0x5851F42D4C957F2D) % (1 << 64))
    h = (key * np.uint64(0x9E3779B97F4A7C15)) >> np.uint64(40)
    return (0.5 + 9.5 * (h.astype(np.float64) / float(1 << 24))).astype(np.float32)


def road_graph(n_target: int, seed: int, K: int, float_weights: bool = False) -> FapGraph:
    """Road-like graph of about n_target vertices (64x64-grid construction scaled:
    10% of edges dropped, 2% diagonals, largest component), optionally with float32
    weights uniform in [0.5, 10) (equal in both directions)."""
    side = max(4, int(math.ceil(math.sqrt(n_target / 0.97))))
    g = FapGraph.road_grid(side, side, 0.10, 0.02, seed, K)
    if not float_weights:
        return g
    c = g.csr()
    w = symmetric_float_weights(c["row_offset"], c["col"], seed)
    return FapGraph.from_csr(c["row_offset"], c["col"], w, False, K)


def size_class(n: int) -> int:
    for i, (lo, hi) in enumerate(SIZE_CLASSES):
        if n < hi:
            return i
    return len(SIZE_CLASSES) - 1


def _time_batch(graphs, cfg: dict, reps: int = 2) -> float:
    b = Batch(graphs, use_graph=False, **cfg)
    b.run()
    best = math.inf
    for _ in range(reps):
        t0 = time.perf_counter()
        b.run()
        best = min(best, time.perf_counter() - t0)
    b.close()
    return best


def tune(copies_work: int = 512, max_copies: int = 48, seed: int = 4242,
         k_candidates=K_CANDIDATES, per_cu_candidates=PER_CU_CANDIDATES,
         src_rows_candidates=SRC_ROWS_CANDIDATES, merge_rows_candidates=MERGE_ROWS_CANDIDATES,
         packed_candidates=PACKED_CANDIDATES, unit_grid_candidates=UNIT_GRID_CANDIDATES,
         fw_tile_candidates=FW_TILE_CANDIDATES, verbose: bool = False) -> dict:
    """Per size class: K with the default tiles, then each knob in turn; the table
    keeps the fastest setting seen (every trial is recorded)."""
    table = {"classes": SIZE_CLASSES, "K": [], "per_cu": [], "src_rows": [], "merge_rows": [],
             "packed": [], "unit_grid": [], "fw_tile": [], "ms_per_graph": [], "trials": []}
    cands = {"per_cu": tuple(per_cu_candidates), "src_rows": tuple(src_rows_candidates),
             "merge_rows": tuple(merge_rows_candidates), "packed": tuple(packed_candidates),
             "unit_grid": tuple(unit_grid_candidates), "fw_tile": tuple(fw_tile_candidates)}
    for ci, (lo, hi) in enumerate(SIZE_CLASSES):
        n_mid = int(math.sqrt(lo * min(hi, 8192)))
        trials = []
        best = None  # (ms, K, cfg, graphs)

        def trial(K, cfg, graphs, copies):
            nonlocal best
            ms = _time_batch(graphs, cfg) * 1e3 / copies
            trials.append(dict(K=K, ms_per_graph=ms, **cfg))
            if verbose:
                print(f"class {ci} n~{n_mid} K={K} {cfg}: {ms:.3f} ms/graph", flush=True)
            if best is None or ms < best[0]:
                best = (ms, K, dict(cfg), graphs, copies)

        default = {k: 16 if k in ("src_rows", "merge_rows") and 16 in cands[k] else cands[k][0]
                   for k in KNOBS}
        for K in k_candidates:
            base = road_graph(n_mid, seed + ci, K)
            if K > base.num_vertexs // 16:
                continue
            base.preCondition()
            if base.layout()["num"][1:].max() > MAX_PART:
                continue
            copies = int(min(max_copies, max(4, copies_work // K)))
            graphs = [road_graph(n_mid, seed + ci, K, float_weights=(c % 2 == 1))
                      for c in range(copies)]
            trial(K, default, graphs, copies)
        _, K, cfg, graphs, copies = best
        for knob in KNOBS:
            for c in cands[knob]:
                if c == best[2][knob]:
                    continue
                trial(K, dict(best[2], **{knob: c}), graphs, copies)
        ms, K, cfg = best[0], best[1], best[2]
        table["K"].append(K)
        for knob in KNOBS:
            table[knob].append(cfg[knob])
        table["ms_per_graph"].append(ms)
        table["trials"].append(trials)
        best = None
    return table


def k_for(n: int, table: dict) -> int:
    return int(table["K"][size_class(n)])


def _dominant_class(ns) -> int:
    work = np.zeros(len(SIZE_CLASSES))
    for n in ns:
        work[size_class(n)] += float(n) * n
    return int(np.argmax(work))


def settings_for(ns, table: dict) -> dict:
    """Batch settings for a mixed batch: the launch-wide knobs (per_cu, src_rows,
    merge_rows, packed, unit_grid) of the size class holding the most work (sum of n^2),
    and the Floyd-Warshall tile of each graph's own size class (a list, one per graph)."""
    c = _dominant_class(ns)
    out = {k: int(table[k][c]) for k in KNOBS if k in table and k != "fw_tile"}
    if "fw_tile" in table:
        out["fw_tile"] = [int(table["fw_tile"][size_class(n)]) for n in ns]
    return out


def per_cu_for(ns, table: dict) -> int:
    return int(table["per_cu"][_dominant_class(ns)])


def load_or_tune(path: str | None = None, **kw) -> dict:
    path = path or os.path.join(os.path.expanduser("~"), ".cache", "fastapsp_amd",
                                "batch_tuning_gfx950.json")
    if os.path.exists(path):
        with open(path) as f:
            return json.load(f)
    table = tune(**kw)
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, "w") as f:
        json.dump(table, f)
    return table

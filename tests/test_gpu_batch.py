"""Many independent graphs in one persistent launch (fap_hip_batch_*): every graph of a
mixed batch (unweighted, integer-weighted and float-weighted; undirected and directed;
K = 1 .. 6) must equal the oracle -- exact int32 against BFS / integer Dijkstra, float32
against the rounding of the FP64 Dijkstra (and within relative 1e-6) -- for the plain
per-graph form, the one-shot batched form and every replay of the captured hipGraph.
Between replays the result buffers are overwritten and other solves run eagerly, so a
replay that did nothing, or replayed stale state, fails."""
import sys

import numpy as np
import pytest

import fastapsp_amd as F
import oracle as O
from fastapsp_amd import _lib as L

pytestmark = pytest.mark.gpu
REL_TOL = 1e-6


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def float_weights(ro, col, directed, seed):
    n = len(ro) - 1
    src = np.repeat(np.arange(n, dtype=np.uint64), np.diff(ro))
    dst = col.astype(np.uint64)
    a, b = (src, dst) if directed else (np.minimum(src, dst), np.maximum(src, dst))
    key = (a * np.uint64(n) + b) ^ np.uint64((seed * 0x5851F42D4C957F2D) % (1 << 64))
    h = (key * np.uint64(0x9E3779B97F4A7C15)) >> np.uint64(40)
    return (0.5 + 9.5 * (h.astype(np.float64) / float(1 << 24))).astype(np.float32)


def make_graphs():
    graphs = []
    shapes = [(12, 12, 1), (20, 16, 2), (28, 24, 3), (36, 30, 4), (24, 40, 5), (44, 40, 6)]
    for i, (W, H, K) in enumerate(shapes):
        base = F.FapGraph.road_grid(W, H, 0.1, 0.03, 10 + i, K)
        if i % 2 == 0:
            graphs.append(base)  # unweighted, undirected
        else:
            c = base.csr()
            w = float_weights(c["row_offset"], c["col"], False, i)
            graphs.append(F.FapGraph.from_csr(c["row_offset"], c["col"], w, False, K))
    graphs.append(F.FapGraph.knn(1500, 3, 1000, 3, 4))  # directed, integer weights
    k2 = F.FapGraph.knn(1200, 3, 1000, 4, 3)
    c = k2.csr()
    graphs.append(F.FapGraph.from_csr(c["row_offset"], c["col"],
                                      float_weights(c["row_offset"], c["col"], True, 7), True, 3))
    for g in graphs:
        g.preCondition()
    return graphs


def check_graph(g):
    lay, csr = g.layout(), g.csr()
    if g.isFloat():
        got = g.getAllDistance()
        want = O.exact_rows_f64(csr["row_offset"], csr["col"], csr["weight"], lay["st2ed"])
        unreach = np.isinf(want)
        assert np.array_equal(got == F.MAXVALUE, unreach)
        w = want[~unreach]
        rel = np.abs(got[~unreach].astype(np.float64) - w) / np.maximum(w, 1e-30)
        assert rel.max(initial=0.0) <= REL_TOL
        assert np.array_equal(got[~unreach], w.astype(np.float32))
    else:
        got = g.getAllDistanceI32()
        w = None if np.all(csr["weight"] == 1) else csr["weight"]
        want = O.exact_rows(csr["row_offset"], csr["col"], w, lay["st2ed"])
        assert np.array_equal(got, want)


def clobber(g, value):
    """Overwrite a graph's result buffer (if it has one yet)."""
    p, ld = g.deviceDistance()
    if p:
        L.check(L.load().fap_hip_fill_i32(p, value, g.num_vertexs * ld, None), "fill")


def test_batch_one_shot_and_plain():
    graphs = make_graphs()
    for mode in (F.FAP_MODE_PLAIN, F.FAP_MODE_BATCHED):
        say(f"solve_batch mode={mode} over {len(graphs)} graphs")
        for g in graphs:
            clobber(g, 12345)
        F.solve_batch(graphs, mode)
        for g in graphs:
            check_graph(g)


@pytest.mark.parametrize("unit_grid", [0, 1])
def test_batch_graph_replays(unit_grid):
    """int32 and FP64 items together: their launch sequences run on two streams (two
    branches of the captured graph); the merge as the persistent queue or a plain grid."""
    graphs = make_graphs()
    b = F.Batch(graphs, use_graph=True, unit_grid=unit_grid)
    other = F.FapGraph.road_grid(30, 30, 0.1, 0.02, 99, 4)
    other.preCondition()
    for rep in range(3):
        say(f"replay {rep}")
        for g in graphs:
            clobber(g, 777 + rep)
        other.solve()  # eager work between replays
        b.run()
        for g in graphs:
            check_graph(g)
    b.close()


def test_batch_without_graph_capture():
    graphs = make_graphs()[:4]
    b = F.Batch(graphs, use_graph=False)
    for rep in range(2):
        for g in graphs:
            clobber(g, -5)
        b.run()
        for g in graphs:
            check_graph(g)
    b.close()


def test_autotune_small_table():
    from fastapsp_amd import autotune as AT
    say("autotune (small candidate set)")
    table = AT.tune(k_candidates=(2, 4), per_cu_candidates=(0, 1), src_rows_candidates=(8, 16),
                    merge_rows_candidates=(8, 16), packed_candidates=(1, 0), max_copies=4)
    assert len(table["K"]) == len(AT.SIZE_CLASSES)
    assert all(k in (2, 4) for k in table["K"]) and all(p in (0, 1) for p in table["per_cu"])
    assert all(r in (8, 16) for r in table["src_rows"] + table["merge_rows"])
    for c, (trials, ms) in enumerate(zip(table["trials"], table["ms_per_graph"])):
        assert ms == min(t["ms_per_graph"] for t in trials)
        chosen = {k: table[k][c] for k in ("K",) + AT.KNOBS}
        assert any(all(t[k] == v for k, v in chosen.items()) and t["ms_per_graph"] == ms
                   for t in trials)
    assert AT.k_for(300, table) == table["K"][0] and AT.k_for(8000, table) == table["K"][-1]
    s = AT.settings_for([8000, 8100, 300], table)
    assert s == dict({k: table[k][-1] for k in AT.KNOBS if k != "fw_tile"},
                     fw_tile=[table["fw_tile"][-1]] * 2 + [table["fw_tile"][0]])
    assert all(t in (64, 32) for t in table["fw_tile"])


def test_batch_fw_tile_32_equals_64():
    """The FP64 Floyd-Warshall on 32 x 32 tiles (per graph, FAP_ITEM_FW_TILE32) gives the
    same float32 results as on 64 x 64 tiles, for parts of 20-300 vertices (1-10 tiles),
    mixed in one batch; both equal the float32 rounding of an FP64 Dijkstra."""
    graphs = []
    for i, (side, K) in enumerate(((12, 4), (20, 3), (30, 5), (40, 6))):
        g = F.FapGraph.road_grid(side, side, 0.1, 0.02, 60 + i, K)
        c = g.csr()
        w = float_weights(c["row_offset"], c["col"], False, 60 + i)
        g = F.FapGraph.from_csr(c["row_offset"], c["col"], w, False, K)
        g.preCondition()
        graphs.append(g)
    tiles = [32, 64, 32, 64]
    b = F.Batch(graphs, use_graph=False, fw_tile=tiles)
    b.run()
    first = [g.getAllDistance().copy() for g in graphs]
    b.close()
    b = F.Batch(graphs, use_graph=False, fw_tile=[64 if t == 32 else 32 for t in tiles])
    b.run()
    for g, a in zip(graphs, first):
        assert np.array_equal(g.getAllDistance(), a)
        csr, lay = g.csr(), g.layout()
        want = O.exact_rows_f64(csr["row_offset"], csr["col"], csr["weight"], lay["st2ed"])
        unreach = np.isinf(want)
        assert np.array_equal(a == F.MAXVALUE, unreach)
        assert np.array_equal(a[~unreach], want[~unreach].astype(np.float32))
    b.close()


def test_batch_inpart_bfs_and_floyd_warshall_agree():
    """Unit-weight symmetric graphs take the in-part BFS for parts of at most 2048
    vertices (two kernel widths: up to 1024 and up to 2048) and build + Floyd-Warshall
    above; FAP_BATCH_NO_BFS forces Floyd-Warshall everywhere.  Both equal exact BFS."""
    graphs = [F.FapGraph.road_like(1800, 0.55, 1e-3, 21, 3),   # parts ~600
              F.FapGraph.road_like(3000, 0.55, 1e-3, 22, 2),   # parts ~1500
              F.FapGraph.road_like(5000, 0.55, 1e-3, 23, 2),   # parts ~2500: Floyd-Warshall
              F.FapGraph.road_grid(30, 30, 0.1, 0.02, 24, 5)]
    for g in graphs:
        g.preCondition()
    sizes = [int(x) for g in graphs for x in g.layout()["num"][1:]]
    assert min(sizes) < 1024 and any(1024 < x <= 2048 for x in sizes) and max(sizes) > 2048
    for bfs, unit_grid in ((1, 0), (0, 1)):
        say(f"batch with in-part BFS = {bfs}, plain-grid merge = {unit_grid}")
        b = F.Batch(graphs, use_graph=True, bfs=bfs, unit_grid=unit_grid)
        for rep in range(2):
            for g in graphs:
                clobber(g, 4242 + rep)
            b.run()
            for g in graphs:
                check_graph(g)
        b.close()

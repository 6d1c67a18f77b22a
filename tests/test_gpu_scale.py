"""Full-size parity of the directed integer-weighted kNN workload: 262,144 uniform points,
arcs to the 3 nearest neighbours (786,432 arcs, weights U[1, 1000]), K = 256, solveSubGraph
over every part in turn.  Every block is checked on sampled rows (its first and last
boundary row, its first and last inner row, one random row) against an independent exact
Dijkstra from the same sources, bit for bit, and on the size-independent properties the
domain fixes: each row's own source is at distance 0, and every row of the block is the row
of the vertex the layout names."""
import sys

import numpy as np
import pytest

import fastapsp_amd as F
import oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_knn262k_k256_subgraph_every_part():
    say("knn 262144 x 3, K=256: solveSubGraph for every part")
    g = F.FapGraph.knn(262144, 3, 1000, 2, 256)
    g.preCondition()
    lay, csr = g.layout(), g.csr()
    assert g.num_vertexs == 262144 and g.num_edges == 786432
    rng = np.random.default_rng(7)
    sources, got = [], []
    for p in range(1, g.K + 1):
        g.solveSubGraph(p)
        off, num, bdy = int(lay["offset"][p]), int(lay["num"][p]), int(lay["bdy"][p])
        rows = {0, num - 1, int(rng.integers(num))}
        if bdy > 0:
            rows.add(bdy - 1)
        if bdy < num:
            rows.add(bdy)
        for r in sorted(rows):
            d = g.getSubGraphDistanceRowsI32(p, r, 1)[0]
            v = int(lay["st2ed"][off + r])
            assert d[v] == 0, f"part {p} row {r}: own distance {d[v]}"
            sources.append(v)
            got.append(d)
        if p % 32 == 0:
            say(f"  part {p}/{g.K}")
    say(f"exact Dijkstra for {len(sources)} sampled rows")
    exact = O.exact_rows(csr["row_offset"], csr["col"], csr["weight"], np.array(sources))
    got = np.stack(got)
    bad = np.nonzero(np.any(got != exact, axis=1))[0]
    assert bad.size == 0, f"{bad.size} sampled rows differ (first source {sources[bad[0]]})"
    # Not strongly connected: unreachable pairs exist and stay FAP_INF.
    assert (got == F.FAP_INF).any()


def test_road114599_paths_sampled_rows_valid():
    """solve(True) at the headline size (n = 114,599, K = 128; 52.5 GB of distances and
    52.5 GB of predecessors resident): on sampled rows every reachable entry's
    predecessor is an in-neighbour on a shortest path, the diagonal names the source,
    and the distances equal an exact BFS."""
    say("road-like 114599 K=128: solve(True), sampled rows")
    g = F.FapGraph.road_like(114599, 0.52, 1e-4, 1, 128)
    g.preCondition()
    g.solve(True)
    lay, csr = g.layout(), g.csr()
    n = g.num_vertexs
    ro, col = csr["row_offset"].astype(np.int64), csr["col"].astype(np.int64)
    tails = np.repeat(np.arange(n, dtype=np.int64), np.diff(ro))
    arc_keys = np.sort(tails * n + col)
    rng = np.random.default_rng(3)
    rows = np.unique(np.concatenate([[0, n - 1], rng.integers(0, n, 14)]))
    for r in rows:
        d = g.getDistanceRowsI32(int(r), 1)[0].astype(np.int64)
        p = g.getPathRows(int(r), 1)[0].astype(np.int64)
        s = int(lay["st2ed"][r])
        assert np.array_equal(d, O.exact_rows(csr["row_offset"], csr["col"], None, [s])[0])
        reach = d < F.FAP_INF
        assert np.all(p[~reach] == -1) and p[s] == s and d[s] == 0
        v = np.nonzero(reach)[0]
        v = v[v != s]
        u = p[v]
        assert np.all(u >= 0)
        keys = u * n + v
        pos = np.searchsorted(arc_keys, keys)
        assert np.all(arc_keys[np.minimum(pos, len(arc_keys) - 1)] == keys), r
        assert np.all(d[u] + 1 == d[v]), r

"""Predecessor (path) matrices: solve(True) / solveSubGraph(id, True) on the GPU must
equal the oracle's restatement of the reference CPU path (Dijkstra heap order for the
boundary rows, k-i-j Floyd-Warshall with strict '>', first-k min-plus, decode) bit for
bit -- paths and distances -- on graphs with many equal-length paths."""
import sys

import numpy as np
import pytest

import fastapsp_amd as F
import oracle as O
from fastapsp_amd import _lib as L

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _valid(dist, path, src, csr):
    """Every finite entry's predecessor u has an arc u -> v with d(u) + w = d(v)."""
    ro, col, w = csr["row_offset"], csr["col"], csr["weight"]
    arcw = {}
    for u in range(len(ro) - 1):
        for e in range(ro[u], ro[u + 1]):
            key = (u, int(col[e]))
            arcw[key] = min(arcw.get(key, 1 << 30), int(w[e]))
    for r, s in enumerate(src):
        for v in range(dist.shape[1]):
            d = dist[r, v]
            if d >= F.FAP_INF:
                assert path[r, v] == -1
            elif v == s:
                assert path[r, v] == s and d == 0
            else:
                u = int(path[r, v])
                assert (u, v) in arcw and dist[r, u] + arcw[(u, v)] == d, (r, v, u)


def _check_solve(g, validate_rows=64):
    g.solve(True)
    got_d, got_p = g.getAllDistanceI32(), g.getAllPath()
    lay, csr = g.layout(), g.csr()
    want_d, want_p = O.solve_parts_path(1, g.K, csr, lay)
    assert np.array_equal(got_d, want_d), "distances differ from the oracle"
    assert np.array_equal(got_p, want_p), "paths differ from the oracle"
    rows = np.arange(0, g.num_vertexs, max(1, g.num_vertexs // validate_rows))
    _valid(got_d[rows], got_p[rows], lay["st2ed"][rows], csr)
    # distance-only solve gives the same distances
    g.solve()
    assert np.array_equal(g.getAllDistanceI32(), got_d)
    with pytest.raises(L.FapError):
        g.getAllPath()  # the path result belongs to solve(True) only
    return got_d, got_p


@pytest.mark.parametrize("args", [(64, 64, 0.10, 0.02, 0, 8),
                                  (40, 30, 0.15, 0.05, 3, 19),
                                  (33, 31, 0.10, 0.02, 5, 1)])
def test_paths_road_grid(args):
    say(f"paths: road grid {args}")
    g = F.FapGraph.road_grid(*args)
    g.preCondition()
    _check_solve(g)


def test_paths_directed_weighted():
    say("paths: directed weighted road grid")
    base = F.FapGraph.road_grid(36, 30, 0.1, 0.02, 4, 6)
    c = base.csr()
    rng = np.random.default_rng(4)
    n = len(c["row_offset"]) - 1
    src = np.repeat(np.arange(n), np.diff(c["row_offset"]))
    keep = rng.random(len(c["col"])) < 0.8
    src, dst = src[keep], c["col"][keep]
    w = rng.integers(1, 6, len(dst)).astype(np.float32)  # small weights: many ties
    order = np.argsort(src, kind="stable")
    ro = np.zeros(n + 1, np.int32)
    ro[1:] = np.cumsum(np.bincount(src, minlength=n))
    g = F.FapGraph.from_csr(ro, dst[order], w[order], True, 6)
    g.preCondition()
    d, _ = _check_solve(g)
    assert (d == F.FAP_INF).any()


def test_paths_knn():
    say("paths: knn directed weighted")
    g = F.FapGraph.knn(1800, 3, 20, 7, 5)
    g.preCondition()
    _check_solve(g)


def test_paths_parts_disconnected_inside():
    say("paths: parts disconnected inside")
    n = 240
    out = [[] for _ in range(n)]
    for i in range(n - 1):
        out[i].append(i + 1)
        out[i + 1].append(i)
    for i in range(0, n - 7, 7):  # chords: several equal-length routes
        out[i].append(i + 7)
        out[i + 7].append(i)
    ro = np.zeros(n + 1, np.int32)
    ro[1:] = np.cumsum([len(x) for x in out])
    col = np.array([v for x in out for v in x], np.int32)
    g = F.FapGraph.from_csr(ro, col, None, False, 3)
    g.setPartition(((np.arange(n) // 40) % 3 + 1).astype(np.int32))
    g.preCondition()
    _check_solve(g)


def test_paths_solve_subgraph_every_part():
    say("paths: solveSubGraph(p, True) for every part")
    g = F.FapGraph.road_grid(48, 40, 0.1, 0.03, 2, 7)
    g.preCondition()
    lay, csr = g.layout(), g.csr()
    for p in range(1, g.K + 1):
        g.solveSubGraph(p, True)
        off, num = lay["offset"][p], lay["num"][p]
        want_d, want_p = O.solve_parts_path(p, p, csr, lay, row_base=off, rows=num)
        assert np.array_equal(g.getSubGraphDistanceI32(p), want_d), p
        assert np.array_equal(g.getSubGraphPath(p), want_p), p


def test_paths_refused_for_float_weights():
    base = F.FapGraph.road_grid(20, 20, 0.1, 0.02, 1, 3)
    c = base.csr()
    w = np.full(len(c["col"]), 1.5, np.float32)
    g = F.FapGraph.from_csr(c["row_offset"], c["col"], w, False, 3)
    g.preCondition()
    with pytest.raises(L.FapError):
        g.solve(True)

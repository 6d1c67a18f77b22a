"""Whole-path parity: fapGraph.solve() / solveSubGraph() on the GPU against the oracle
(the reference's CPU path restated, float with MAXVALUE -> int32) and against an
independent exact APSP, bit-exact for int32."""
import sys

import numpy as np
import pytest

import fastapsp_amd as F
import oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _check_full(g, mode):
    g.setMode(mode)
    g.solve()
    got = g.getAllDistanceI32()
    lay, csr = g.layout(), g.csr()
    want = O.solve_parts(1, g.K, csr, lay)
    assert np.array_equal(got, want), "GPU solve() differs from the oracle"
    w = None if np.all(csr["weight"] == 1) else csr["weight"]
    exact = O.exact_rows(csr["row_offset"], csr["col"], w, lay["st2ed"])
    assert np.array_equal(got, exact), "GPU solve() differs from exact APSP"
    return got


@pytest.mark.parametrize("mode", [F.FAP_MODE_BATCHED, F.FAP_MODE_OVERLAP, F.FAP_MODE_PLAIN])
def test_solve_road_grid_k8(mode):
    """64 x 64 road grid (10% edges dropped, 2% diagonals), K = 8, full n x n."""
    say(f"solve road grid 64x64 K=8 mode={mode}")
    g = F.FapGraph.road_grid(64, 64, 0.10, 0.02, 0, 8)
    g.preCondition()
    _check_full(g, mode)


def test_solve_forms_agree():
    say("plain == batched == overlap")
    g = F.FapGraph.road_grid(40, 30, 0.15, 0.05, 3, 19)
    g.preCondition()
    a = _check_full(g, F.FAP_MODE_PLAIN)
    b = _check_full(g, F.FAP_MODE_OVERLAP)
    c = _check_full(g, F.FAP_MODE_BATCHED)
    assert np.array_equal(a, b) and np.array_equal(b, c)


def test_solve_repeated_and_mixed_forms():
    # Back-to-back solves in different forms reuse the workspaces; results must not change.
    say("repeated solves")
    g = F.FapGraph.road_grid(36, 36, 0.1, 0.02, 8, 12)
    g.preCondition()
    first = _check_full(g, F.FAP_MODE_BATCHED)
    for mode in (F.FAP_MODE_OVERLAP, F.FAP_MODE_BATCHED, F.FAP_MODE_PLAIN, F.FAP_MODE_BATCHED):
        g.setMode(mode)
        g.solve()
        assert np.array_equal(g.getAllDistanceI32(), first)


@pytest.mark.parametrize("K", [1, 2, 3, 5, 16])
def test_solve_various_k(K):
    say(f"solve K={K}")
    g = F.FapGraph.road_grid(33, 31, 0.1, 0.02, K, K)
    g.preCondition()
    _check_full(g, F.FAP_MODE_BATCHED)


def test_solve_directed_weighted_knn():
    say("solve knn directed weighted")
    g = F.FapGraph.knn(2500, 3, 1000, 2, 8)
    g.preCondition()
    _check_full(g, F.FAP_MODE_BATCHED)


def test_solve_disconnected_components():
    # Two disjoint paths: distances across components stay unreachable (INF).
    say("solve disconnected")
    n = 400
    arcs = [(i, i + 1) for i in range(199)] + [(i, i + 1) for i in range(200, 399)]
    ro, col = _csr(n, arcs)
    g = F.FapGraph.from_csr(ro, col, None, False, 4)
    g.preCondition()
    got = _check_full(g, F.FAP_MODE_BATCHED)
    assert (got == F.FAP_INF).sum() == 2 * 200 * 200


def test_solve_subgraph_all_parts():
    say("solveSubGraph all parts")
    g = F.FapGraph.road_grid(48, 48, 0.1, 0.02, 4, 6)
    g.preCondition()
    lay, csr = g.layout(), g.csr()
    for p in range(1, g.K + 1):
        g.solveSubGraph(p)
        got = g.getSubGraphDistanceI32(p)
        off, num = lay["offset"][p], lay["num"][p]
        want = O.solve_parts(p, p, csr, lay, row_base=off, rows=num)
        assert np.array_equal(got, want), f"part {p}"
        assert np.array_equal(g.getSubGraphIndex(p), lay["st2ed"][off:off + num])
    # float getter: unreachable maps to the reference's MAXVALUE
    f = g.getSubGraphDistance(g.K)
    assert f.dtype == np.float32 and np.all((f == F.MAXVALUE) == (got == F.FAP_INF))


def _csr(n, arcs):
    out = [[] for _ in range(n)]
    for a, b in arcs:
        out[a].append(b)
        out[b].append(a)
    ro = np.zeros(n + 1, np.int32)
    ro[1:] = np.cumsum([len(x) for x in out])
    col = np.array([v for x in out for v in x], np.int32)
    return ro, col

"""The float-weight path (FP64 arithmetic, float32 results) against the FP64 oracle
(Dijkstra in double): relative error <= 1e-6 (the tolerance the float path is held to),
and, since float32 weights in [0.5, 10) make every path sum exact in FP64, the stronger
expected result: each entry is the float32 rounding of the FP64 distance.  Unreachable
pairs are MAXVALUE = 1e8, as in the reference's float output."""
import sys

import numpy as np
import pytest

import fastapsp_amd as F
import oracle as O

pytestmark = pytest.mark.gpu
REL_TOL = 1e-6


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def float_weights(ro, col, directed, seed):
    """Deterministic float32 weights in [0.5, 10); equal on u->v and v->u when undirected."""
    n = len(ro) - 1
    src = np.repeat(np.arange(n, dtype=np.uint64), np.diff(ro))
    dst = col.astype(np.uint64)
    a, b = (src, dst) if directed else (np.minimum(src, dst), np.maximum(src, dst))
    key = (a * np.uint64(n) + b) ^ np.uint64((seed * 0x5851F42D4C957F2D) % (1 << 64))
    h = (key * np.uint64(0x9E3779B97F4A7C15)) >> np.uint64(40)
    return (0.5 + 9.5 * (h.astype(np.float64) / float(1 << 24))).astype(np.float32)


def check_rows(got, src, csr):
    want = O.exact_rows_f64(csr["row_offset"], csr["col"], csr["weight"], src)
    unreach = np.isinf(want)
    assert np.array_equal(got == F.MAXVALUE, unreach), "reachability differs"
    w = want[~unreach]
    g = got[~unreach].astype(np.float64)
    rel = np.abs(g - w) / np.maximum(w, 1e-30)
    assert rel.max(initial=0.0) <= REL_TOL, f"max relative error {rel.max()}"
    assert np.array_equal(got[~unreach], w.astype(np.float32)), "not the rounding of FP64"


def road_graph(W, H, seed, K):
    base = F.FapGraph.road_grid(W, H, 0.1, 0.03, seed, K)
    csr = base.csr()
    w = float_weights(csr["row_offset"], csr["col"], False, seed)
    return F.FapGraph.from_csr(csr["row_offset"], csr["col"], w, False, K)


@pytest.mark.parametrize("mode", [F.FAP_MODE_BATCHED, F.FAP_MODE_PLAIN, F.FAP_MODE_OVERLAP])
def test_fp32_road_grid_solve(mode):
    say(f"fp32 road grid 40x36 K=6 mode={mode}")
    g = road_graph(40, 36, 3, 6)
    assert g.isFloat()
    g.preCondition()
    g.setMode(mode)
    g.solve()
    got = g.getAllDistance()
    lay, csr = g.layout(), g.csr()
    check_rows(got, lay["st2ed"], csr)
    with pytest.raises(F.FapError):
        g.getAllDistanceI32()


def test_fp32_forms_agree_and_repeat():
    say("fp32 forms agree")
    g = road_graph(30, 30, 5, 9)
    g.preCondition()
    outs = []
    for mode in (F.FAP_MODE_PLAIN, F.FAP_MODE_BATCHED, F.FAP_MODE_BATCHED):
        g.setMode(mode)
        g.solve()
        outs.append(g.getAllDistance())
    assert np.array_equal(outs[0], outs[1]) and np.array_equal(outs[1], outs[2])


def test_fp32_directed_knn_subgraphs():
    say("fp32 directed knn, solveSubGraph every part")
    base = F.FapGraph.knn(3000, 3, 1000, 4, 5)
    csr0 = base.csr()
    w = float_weights(csr0["row_offset"], csr0["col"], True, 9)
    g = F.FapGraph.from_csr(csr0["row_offset"], csr0["col"], w, True, 5)
    g.preCondition()
    lay, csr = g.layout(), g.csr()
    for p in range(1, g.K + 1):
        g.solveSubGraph(p)
        off, num = lay["offset"][p], lay["num"][p]
        got = g.getSubGraphDistance(p)
        assert got.shape == (num, g.num_vertexs)
        check_rows(got, lay["st2ed"][off:off + num], csr)


def test_fp32_components_and_single_part():
    say("fp32 two components, K=1 and K=3")
    n = 300
    edges = [(i, i + 1) for i in range(149)] + [(i, i + 1) for i in range(150, 299)]
    adj = [[] for _ in range(n)]
    for a, b in edges:
        adj[a].append(b)
        adj[b].append(a)
    ro = np.zeros(n + 1, np.int32)
    ro[1:] = np.cumsum([len(x) for x in adj])
    col = np.array([v for x in adj for v in x], np.int32)
    w = float_weights(ro, col, False, 2)
    for K in (1, 3):
        g = F.FapGraph.from_csr(ro, col, w, False, K)
        g.preCondition()
        g.solve()
        got = g.getAllDistance()
        check_rows(got, g.layout()["st2ed"], g.csr())
        assert (got == F.MAXVALUE).sum() == 2 * 150 * 150

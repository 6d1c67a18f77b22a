"""solveSubGraph sweeps: after the sweep threshold, every part's boundary rows and
Floyd-Warshall matrix come from one batched pass and each call only writes its block
(dense product on graphs whose boundary rows reach everything, the sparse product on
directed k-NN graphs whose boundary rows are mostly unreachable).  Every block must equal
the per-part form bit for bit, and the oracle's restatement of the reference CPU path
(fap/fap.cpp:195-503 for one part) and an exact Dijkstra / BFS of the block's rows."""
import sys

import numpy as np
import pytest

import fastapsp_amd as F
import oracle as O
from fastapsp_amd.autotune import road_graph, symmetric_float_weights

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _blocks(g, order):
    out = {}
    for p in order:
        g.solveSubGraph(p)
        out[p] = g.getSubGraphDistanceI32(p)
    return out


def _check_sweep(g, order, weighted):
    lay, csr = g.layout(), g.csr()
    g.setSubGraphSweep(0)  # per-part form
    plain = _blocks(g, order)
    assert not g.subGraphSweepReady()
    g.setSubGraphSweep(1)  # prepared from the first call on
    swept = _blocks(g, order)
    assert g.subGraphSweepReady()
    w = csr["weight"] if weighted else None
    for p in sorted(set(order)):
        off, num = int(lay["offset"][p]), int(lay["num"][p])
        assert np.array_equal(swept[p], plain[p]), f"part {p}: sweep differs from per-part form"
        want = O.solve_parts(p, p, csr, lay, row_base=off, rows=num)
        assert np.array_equal(swept[p], want), f"part {p}: sweep differs from the oracle"
        exact = O.exact_rows(csr["row_offset"], csr["col"], w, lay["st2ed"][off:off + num])
        assert np.array_equal(swept[p], exact), f"part {p}: sweep differs from exact rows"
    return swept


def test_sweep_knn_sparse_blocks():
    say("sweep: directed 3-NN, int weights, K=16, every part (sparse product)")
    g = F.FapGraph.knn(6000, 3, 1000, 5, 16)
    g.preCondition()
    blocks = _check_sweep(g, list(range(1, g.K + 1)), True)
    assert g.subGraphSweepState() == 2, "expected the sparse product"
    assert any((b == F.FAP_INF).any() for b in blocks.values())


def test_sweep_road_dense_blocks_any_order():
    say("sweep: unit-weight road grid, K=8, parts in a shuffled order with repeats (dense)")
    g = F.FapGraph.road_grid(56, 56, 0.1, 0.02, 9, 8)
    g.preCondition()
    order = [5, 2, 8, 2, 1, 7, 3, 6, 4, 5]
    _check_sweep(g, order, False)
    assert g.subGraphSweepState() == 1, "expected the dense product"


def test_sweep_threshold_and_layout_change():
    say("sweep: default threshold (third call), then a new layout drops the prepared parts")
    g = F.FapGraph.knn(3000, 3, 50, 11, 8)
    g.preCondition()
    lay, csr = g.layout(), g.csr()
    for i, p in enumerate([3, 1, 2, 4]):
        g.solveSubGraph(p)
        assert g.subGraphSweepReady() == (i >= 2), f"call {i + 1}"
        off, num = int(lay["offset"][p]), int(lay["num"][p])
        want = O.solve_parts(p, p, csr, lay, row_base=off, rows=num)
        assert np.array_equal(g.getSubGraphDistanceI32(p), want), f"part {p}"
    g.preCondition()  # layout recomputed: device copies and the sweep are released
    assert not g.subGraphSweepReady()
    g.solveSubGraph(2)
    off, num = int(lay["offset"][2]), int(lay["num"][2])
    assert np.array_equal(g.getSubGraphDistanceI32(2),
                          O.solve_parts(2, 2, csr, lay, row_base=off, rows=num))


@pytest.mark.parametrize("integer", [True, False])
def test_sweep_weighted_road_full_solve_form(integer):
    """Weighted road graphs take their blocks from the boundary graph (every call closes
    every part and the boundary graph); in a sweep one full solve() runs and each block is
    a copy of its rows.  Every block against the per-call form, bit for bit."""
    say(f"sweep: weighted road, {'int' if integer else 'float'} weights, full-solve form")
    g0 = road_graph(9000, 61, 12)
    c = g0.csr()
    w = symmetric_float_weights(c["row_offset"], c["col"], 61)
    if integer:
        w = (np.floor((w - 0.5) / 9.5 * 1000) + 1).astype(np.float32)
    g = F.FapGraph.from_csr(c["row_offset"], c["col"], w, False, 12)
    g.preCondition()
    lay = g.layout()

    def block(p):
        num = int(lay["num"][p])
        return (g.getSubGraphDistanceI32(p) if integer
                else g.getSubGraphDistanceRowsF32(p, 0, num))

    g.setSubGraphSweep(0)
    per_call = {}
    for p in range(1, g.K + 1):
        g.solveSubGraph(p)
        per_call[p] = block(p)
    assert g.subGraphSweepState() == 0
    g.setSubGraphSweep(2)
    for p in [4, 1, 12, 7, 2, 3, 5, 6, 8, 9, 10, 11, 4]:
        g.solveSubGraph(p)
        assert np.array_equal(block(p), per_call[p]), f"part {p}"
    assert g.subGraphSweepState() == 3

"""Full solves of weighted road-like graphs whose boundary rows come from the boundary
graph closed by Floyd-Warshall (the batch stages, one graph): integer weights with more
boundary vertices than the single-graph boundary-graph form holds, and float weights.
* every entry equals the boundary-SSSP form of the same call (FAP_MODE_FLAG_SSSP: one
  label-correcting SSSP per boundary vertex, the reference's structure), bit for bit;
* sampled rows equal an independent integer Dijkstra (int weights) or lie within
  relative 1e-6 of an FP64 Dijkstra and equal its float32 rounding (float weights);
* solveSubGraph blocks (one part of a batch item whose parts tile the graph: every
  part's in-part distances and the closed boundary graph, then that part's rows) equal
  the rows of solve() and the boundary-SSSP form's block."""
import sys

import numpy as np
import pytest

import fastapsp_amd as F
import oracle as O
from fastapsp_amd import _lib as L
from fastapsp_amd.autotune import road_graph, symmetric_float_weights

pytestmark = pytest.mark.gpu

SSSP = F.FAP_MODE_BATCHED | F.FAP_MODE_FLAG_SSSP


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def weighted_road(nt, K, seed, integer):
    g = road_graph(nt, seed, K)
    c = g.csr()
    w = symmetric_float_weights(c["row_offset"], c["col"], seed)
    if integer:  # integers 1..1000, equal both ways
        w = (np.floor((w - 0.5) / 9.5 * 1000) + 1).astype(np.float32)
    g = F.FapGraph.from_csr(c["row_offset"], c["col"], w, False, K)
    g.preCondition()
    return g


def sampled_rows(lay, n, rng, per_part=2):
    rows = {0, n - 1}
    for p in range(1, len(lay["num"])):
        off, num, bdy = int(lay["offset"][p]), int(lay["num"][p]), int(lay["bdy"][p])
        rows |= {off, off + num - 1, off + max(bdy - 1, 0), min(off + bdy, off + num - 1)}
        rows |= set(int(x) for x in rng.integers(off, off + num, per_part))
    return np.array(sorted(rows))


def rows_of(g, r0, cnt, fl):
    return g.getDistanceRowsF32(r0, cnt) if fl else g.getDistanceRowsI32(r0, cnt)


def check_against_sssp_form(g, fl, block=2048):
    """Every entry of the current result against a boundary-SSSP solve of the same graph."""
    n = g.num_vertexs
    first = [rows_of(g, r0, min(block, n - r0), fl) for r0 in range(0, n, block)]
    g.setMode(SSSP)
    g.solve()
    st = L.last_stage_times()
    assert st["boundary_graph"] == 0
    for i, r0 in enumerate(range(0, n, block)):
        got = rows_of(g, r0, min(block, n - r0), fl)
        assert np.array_equal(got, first[i]), f"rows {r0}.. differ from the SSSP form"
    g.setMode(F.FAP_MODE_BATCHED)


def check_subgraph_against_sssp_form(g, p, fl):
    """solveSubGraph(p) from the boundary graph against the boundary-SSSP form."""
    num = int(g.layout()["num"][p])
    g.solveSubGraph(p)
    a = g.getSubGraphDistanceRowsF32(p, 0, num) if fl else g.getSubGraphDistanceI32(p)
    g.setMode(SSSP)
    g.solveSubGraph(p)
    assert L.last_stage_times()["boundary_graph"] == 0
    b = g.getSubGraphDistanceRowsF32(p, 0, num) if fl else g.getSubGraphDistanceI32(p)
    g.setMode(F.FAP_MODE_BATCHED)
    assert np.array_equal(a, b), p


@pytest.mark.timeout(900)
def test_int_weights_beyond_the_single_graph_boundary_limit():
    say("road-like ~31k, int weights 1..1000, K=80 (4571 boundary vertices)")
    g = weighted_road(30000, 80, 41, True)
    lay, csr = g.layout(), g.csr()
    n = g.num_vertexs
    assert int(lay["bdy"][1:].sum()) > 4096 and not g.isFloat()
    g.solve()
    st = L.last_stage_times()
    say(f"solve {st['total_ms']:.1f} ms device, boundary_graph={st['boundary_graph']}")
    assert st["boundary_graph"] == 1
    rows = sampled_rows(lay, n, np.random.default_rng(41))
    got = np.stack([g.getDistanceRowsI32(int(r), 1)[0] for r in rows])
    want = O.exact_rows(csr["row_offset"], csr["col"], csr["weight"].astype(np.int32),
                        lay["st2ed"][rows])
    assert np.array_equal(got, want)
    for p in (1, len(lay["num"]) // 2, len(lay["num"]) - 1):
        off, num = int(lay["offset"][p]), int(lay["num"][p])
        full = g.getDistanceRowsI32(off, num)
        g.solveSubGraph(p)
        assert L.last_stage_times()["boundary_graph"] == 1  # one-part batch item
        assert np.array_equal(g.getSubGraphDistanceI32(p), full), p
    check_subgraph_against_sssp_form(g, len(lay["num"]) // 3, False)
    check_against_sssp_form(g, False)
    say("every entry equals the boundary-SSSP form")


@pytest.mark.timeout(900)
def test_float_weights_boundary_graph_form():
    say("road-like ~12k, float weights, K=24")
    g = weighted_road(12000, 24, 43, False)
    lay, csr = g.layout(), g.csr()
    n = g.num_vertexs
    assert g.isFloat()
    g.solve()
    st = L.last_stage_times()
    say(f"solve {st['total_ms']:.1f} ms device, boundary_graph={st['boundary_graph']}")
    assert st["boundary_graph"] == 1
    rows = sampled_rows(lay, n, np.random.default_rng(43), per_part=6)
    got = np.stack([g.getDistanceRowsF32(int(r), 1)[0] for r in rows])
    want = O.exact_rows_f64(csr["row_offset"], csr["col"], csr["weight"], lay["st2ed"][rows])
    unreach = np.isinf(want)
    assert np.array_equal(got == F.MAXVALUE, unreach)
    wv = want[~unreach]
    assert (np.abs(got[~unreach] - wv) / np.maximum(wv, 1e-30)).max(initial=0) <= 1e-6
    assert np.array_equal(got[~unreach], wv.astype(np.float32))
    p = len(lay["num"]) // 2
    off, num = int(lay["offset"][p]), int(lay["num"][p])
    full = g.getDistanceRowsF32(off, num)
    g.solveSubGraph(p)
    assert L.last_stage_times()["boundary_graph"] == 1  # one-part batch item
    assert np.array_equal(g.getSubGraphDistanceRowsF32(p, 0, num), full)
    check_subgraph_against_sssp_form(g, 1, True)
    check_against_sssp_form(g, True)
    say("every entry equals the boundary-SSSP form")

"""Parts larger than the LDS row buffers of the build kernels (8192 columns): rows are
assembled in chunks, and the in-part Floyd-Warshall of the boundary-graph form runs in
int32 once a part exceeds the 16-bit cap (15,872 vertices).
* K = 8 on the 114,599-vertex road-like graph (parts of ~14.3k vertices): every entry
  against exact BFS;
* K = 4 (parts of ~28.6k vertices, above the cap): sampled rows of solve() and of a
  solveSubGraph block against exact BFS, and symmetry on sampled pairs;
* float weights with parts above 8192 vertices (FP64 build in chunks): sampled rows at
  relative 1e-6 of an FP64 Dijkstra and equal to its float32 rounding;
* predecessor matrices (path build in chunks): sampled rows are shortest-path
  predecessor rows and their distances equal exact BFS."""
import sys
import time

import numpy as np
import pytest

import fastapsp_amd as F
import oracle as O
from fastapsp_amd import _lib as L
from fastapsp_amd.autotune import symmetric_float_weights

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.timeout(900)
def test_road114599_k8_every_entry():
    say("road-like 114599, K=8: solve(), every entry vs exact BFS")
    g = F.FapGraph.road_like(114599, 0.52, 1e-4, 1, 8)
    g.preCondition()
    lay, csr = g.layout(), g.csr()
    assert lay["num"][1:].max() > 8192
    t0 = time.time()
    g.solve()
    st = L.last_stage_times()
    say(f"solve {st['total_ms']:.1f} ms device, boundary_graph={st['boundary_graph']} "
        f"parts_fw16={st['parts_fw16']} ({time.time() - t0:.1f}s)")
    n = g.num_vertexs
    step = 8192
    for r0 in range(0, n, step):
        r1 = min(n, r0 + step)
        got = g.getDistanceRowsI32(r0, r1 - r0)
        bad, first = O.compare_exact_rows(csr["row_offset"], csr["col"], None,
                                          lay["st2ed"][r0:r1], got)
        assert bad == 0, f"rows {r0}..{r1}: {bad} differ (first {r0 + first})"
    say(f"all {n * n:.3e} entries equal ({time.time() - t0:.0f}s)")


def sampled_rows(lay, n, rng, per_part=4):
    rows = {0, n - 1}
    K = len(lay["num"]) - 1
    for p in range(1, K + 1):
        off, num, bdy = int(lay["offset"][p]), int(lay["num"][p]), int(lay["bdy"][p])
        rows |= {off, off + num - 1}
        if bdy:
            rows.add(off + bdy - 1)
        if bdy < num:
            rows.add(off + bdy)
        rows |= set(int(x) for x in rng.integers(off, off + num, per_part))
    return np.array(sorted(rows))


@pytest.mark.timeout(900)
def test_road114599_k4_parts_above_the_cap():
    say("road-like 114599, K=4: parts above the 16-bit cap")
    g = F.FapGraph.road_like(114599, 0.52, 1e-4, 1, 4)
    g.preCondition()
    lay, csr = g.layout(), g.csr()
    n = g.num_vertexs
    assert lay["num"][1:].min() > 15872
    g.solve()
    rng = np.random.default_rng(4)
    rows = sampled_rows(lay, n, rng)
    got = np.stack([g.getDistanceRowsI32(int(r), 1)[0] for r in rows])
    want = O.exact_rows(csr["row_offset"], csr["col"], None, lay["st2ed"][rows])
    assert np.array_equal(got, want)
    # symmetry on sampled pairs (undirected): d(u, v) = d(v, u)
    a = rows[::3]
    d_ab = np.stack([g.getDistanceRowsI32(int(r), 1)[0] for r in a])
    st2ed = lay["st2ed"]
    for i, r in enumerate(a):
        for s in a:
            assert d_ab[i, st2ed[s]] == d_ab[list(a).index(s), st2ed[r]]
    # one sub-graph block through solveSubGraph (boundary SSSP form)
    g.solveSubGraph(2)
    off, num = int(lay["offset"][2]), int(lay["num"][2])
    sub = rng.integers(0, num, 24)
    got = np.stack([g.getSubGraphDistanceRowsI32(2, int(r), 1)[0] for r in sub])
    want = O.exact_rows(csr["row_offset"], csr["col"], None, lay["st2ed"][off + sub])
    assert np.array_equal(got, want)


@pytest.mark.timeout(600)
def test_float_weights_parts_above_8192():
    say("road-like 20000, float weights, K=2")
    base = F.FapGraph.road_like(20000, 0.6, 1e-4, 3, 2)
    c = base.csr()
    w = symmetric_float_weights(c["row_offset"], c["col"], 3)
    g = F.FapGraph.from_csr(c["row_offset"], c["col"], w, False, 2)
    g.preCondition()
    lay, csr = g.layout(), g.csr()
    assert g.isFloat() and lay["num"][1:].max() > 8192
    g.solve()
    rows = sampled_rows(lay, g.num_vertexs, np.random.default_rng(5), per_part=12)
    got = np.stack([g.getDistanceRowsF32(int(r), 1)[0] for r in rows])
    want = O.exact_rows_f64(csr["row_offset"], csr["col"], csr["weight"], lay["st2ed"][rows])
    unreach = np.isinf(want)
    assert np.array_equal(got == F.MAXVALUE, unreach)
    wv = want[~unreach]
    assert (np.abs(got[~unreach] - wv) / np.maximum(wv, 1e-30)).max(initial=0) <= 1e-6
    assert np.array_equal(got[~unreach], wv.astype(np.float32))


@pytest.mark.timeout(600)
def test_paths_parts_above_8192():
    say("road-like 20000, unit weights, K=2: solve(True)")
    g = F.FapGraph.road_like(20000, 0.6, 1e-4, 6, 2)
    g.preCondition()
    lay, csr = g.layout(), g.csr()
    n = g.num_vertexs
    assert lay["num"][1:].max() > 8192
    g.solve(True)
    ro, col = csr["row_offset"].astype(np.int64), csr["col"].astype(np.int64)
    tails = np.repeat(np.arange(n, dtype=np.int64), np.diff(ro))
    arc_keys = np.sort(tails * n + col)
    for r in sampled_rows(lay, n, np.random.default_rng(6), per_part=6):
        d = g.getDistanceRowsI32(int(r), 1)[0].astype(np.int64)
        p = g.getPathRows(int(r), 1)[0].astype(np.int64)
        s = int(lay["st2ed"][r])
        assert np.array_equal(d, O.exact_rows(csr["row_offset"], csr["col"], None, [s])[0])
        reach = d < F.FAP_INF
        assert np.all(p[~reach] == -1) and p[s] == s
        v = np.nonzero(reach)[0]
        v = v[v != s]
        u = p[v]
        keys = u * n + v
        pos = np.searchsorted(arc_keys, keys)
        assert np.all(arc_keys[np.minimum(pos, len(arc_keys) - 1)] == keys), r
        assert np.all(d[u] + 1 == d[v]), r

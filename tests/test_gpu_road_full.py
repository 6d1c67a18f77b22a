"""Bit-exact parity of the timed headline workload at full size: the road-like graph of
114,599 vertices (percolated grid, p = 0.52, chords 1e-4, seed 1), K = 128, full solve()
in the batched form -- boundary rows from the boundary graph, in-part distances by the
bit-parallel BFS, the streaming merge -- exactly the form bench.py times.

All 1.313e10 entries are compared, streamed out of HBM in blocks of 16 parts:
* every row against an independent exact BFS from its source (orc_compare_exact_rows);
* every part's block against the oracle's restatement of the reference CPU path
  (fapGraph::run, fap/fap.cpp:195-503: Dijkstra boundary rows, sub-matrix build,
  Floyd-Warshall, min-plus, decode; orc_compare_parts, OpenMP across parts).
"""
import sys
import time

import numpy as np
import pytest

import fastapsp_amd as F
import oracle as O
from fastapsp_amd import _lib as L

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.timeout(900)
def test_road114599_k128_batched_solve_every_entry():
    say("road-like 114599 K=128: batched solve(), every entry vs exact BFS and the oracle")
    g = F.FapGraph.road_like(114599, 0.52, 1e-4, 1, 128)
    g.preCondition()
    g.setMode(F.FAP_MODE_BATCHED)
    t0 = time.time()
    g.solve()
    st = L.last_stage_times()
    say(f"solve: {st['total_ms']:.2f} ms device; boundary_graph={st['boundary_graph']} "
        f"parts_bfs={st['parts_bfs']} merge_packed={st['merge_packed']}")
    # The timed kernels: boundary-graph rows, in-part BFS on every part.
    assert st["boundary_graph"] == 1 and st["parts_bfs"] == g.K == 128
    assert st["merge_packed"] in (0, 1)
    lay, csr = g.layout(), g.csr()
    n = g.num_vertexs
    assert n == 114599
    checked = 0
    step = 16
    for p0 in range(1, g.K + 1, step):
        p1 = min(g.K, p0 + step - 1)
        r0 = int(lay["offset"][p0])
        r1 = int(lay["offset"][p1] + lay["num"][p1])
        got = g.getDistanceRowsI32(r0, r1 - r0)
        bad, first = O.compare_exact_rows(csr["row_offset"], csr["col"], None,
                                          lay["st2ed"][r0:r1], got)
        assert bad == 0, f"parts {p0}..{p1}: {bad} rows differ from exact BFS " \
                         f"(first sorted row {r0 + first})"
        bad, first = O.compare_parts(p0, p1, csr, lay, got, r0)
        assert bad == 0, f"parts {p0}..{p1}: {bad} entries differ from the oracle " \
                         f"(first sorted row {r0 + first})"
        checked += got.size
        say(f"  parts {p0}..{p1}: rows {r0}..{r1 - 1} equal ({checked:.3e} entries, "
            f"{time.time() - t0:.0f} s)")
        del got
    assert checked == n * n

"""Nested closure of large weighted parts in the batch form: sub-parts of at most 64
vertices closed one Floyd-Warshall tile each, their crossing graph closed, then two
min-plus passes write the part's matrix.  The in-part distances are exact minima over
the same paths as Floyd-Warshall on the whole part, so:
* every entry of every graph equals the batch with FAP_BATCH_NO_NESTED (Floyd-Warshall
  on the whole part), bit for bit: float weights (FP64 sums, 64- and 32-tile crossing
  graphs), integer weights, directed kNN graphs (first exit from / last entry into a
  sub-part differ there);
* sampled rows equal an independent Dijkstra (integer weights) or the float32 rounding
  of an FP64 Dijkstra (float weights);
* the nested path ran (the batch reports its nested parts)."""
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


def int_road(nt, K, seed):
    g = road_graph(nt, seed, K)
    c = g.csr()
    w = symmetric_float_weights(c["row_offset"], c["col"], seed)
    w = (np.floor((w - 0.5) / 9.5 * 1000) + 1).astype(np.float32)  # integers 1..1000
    g = F.FapGraph.from_csr(c["row_offset"], c["col"], w, False, K)
    g.preCondition()
    return g


def float_knn(n, K, seed):
    base = F.FapGraph.knn(n, 3, 1000, seed, K)
    c = base.csr()
    w = symmetric_float_weights(c["row_offset"], c["col"], seed, directed=True)
    g = F.FapGraph.from_csr(c["row_offset"], c["col"], w, True, K)
    g.preCondition()
    return g


def all_rows(g):
    n = g.num_vertexs
    return g.getDistanceRowsF32(0, n) if g.isFloat() else g.getDistanceRowsI32(0, n)


def run(graphs, nested, **kw):
    b = F.Batch(graphs, use_graph=False, nested=nested, **kw)
    b.run()
    cnt = b.nested_parts()
    b.close()
    return cnt, [all_rows(g) for g in graphs]


def check_oracle(g, rows_out, seed, per_graph=24):
    lay, csr = g.layout(), g.csr()
    n = g.num_vertexs
    rng = np.random.default_rng(seed)
    rows = np.unique(np.concatenate([[0, n - 1], rng.integers(0, n, per_graph)]))
    got = rows_out[rows]
    src = lay["st2ed"][rows]
    if g.isFloat():
        want = O.exact_rows_f64(csr["row_offset"], csr["col"], csr["weight"], src)
        unreach = np.isinf(want)
        assert np.array_equal(got == F.MAXVALUE, unreach)
        assert np.array_equal(got[~unreach], want[~unreach].astype(np.float32))
    else:
        want = O.exact_rows(csr["row_offset"], csr["col"], csr["weight"].astype(np.int32), src)
        assert np.array_equal(got, want)


def compare(graphs, label, **kw):
    n0, ref = run(graphs, 0, **kw)
    n1, got = run(graphs, 1, **kw)
    say(f"{label}: nested parts {n1} (off: {n0})")
    assert n0 == 0 and n1 > 0
    for i, g in enumerate(graphs):
        assert np.array_equal(got[i], ref[i]), f"{label}: graph {i} differs from Floyd-Warshall"
        check_oracle(g, got[i], 100 + i)


@pytest.mark.timeout(600)
def test_nested_float_roads_both_tiles():
    graphs = [road_graph(nt, 70 + i, 8, float_weights=True) for i, nt in enumerate((1400, 2600, 5200))]
    for g in graphs:
        g.preCondition()
    compare(graphs, "float roads, 64-tile")
    compare(graphs, "float roads, 32-tile", fw_tile=32)


@pytest.mark.timeout(600)
def test_nested_int_roads_and_mixed_batch():
    ints = [int_road(nt, 8, 80 + i) for i, nt in enumerate((2000, 4600))]
    floats = [road_graph(3000, 90, 8, float_weights=True)]
    floats[0].preCondition()
    unit = [road_graph(2500, 91, 8)]  # unit weights: in-part BFS, never nested
    unit[0].preCondition()
    compare(ints + floats + unit, "int + float + unit roads")


@pytest.mark.timeout(600)
def test_nested_directed_knn():
    graphs = [F.FapGraph.knn(4000, 3, 1000, 7, 4), float_knn(3000, 4, 8)]
    for g in graphs:
        g.preCondition()
    compare(graphs, "directed kNN, int and float weights")

"""Boundary rows from the boundary graph (the BATCHED form of solve() when the parts
cover the graph): results must equal the SSSP form, the plain per-part form, the
oracle's restatement of the reference path and an exact all-pairs BFS / Dijkstra, bit
for bit, on unit-weight, integer-weighted and directed graphs, with parts whose
boundary exceeds one register chunk and parts that are disconnected inside.  On
unit-weight undirected graphs the in-part distances come from the bit-parallel BFS;
FAP_MODE_FLAG_FW (16-bit Floyd-Warshall) and FAP_MODE_FLAG_INT32_FW must agree."""
import sys

import numpy as np
import pytest

import fastapsp_amd as F
import oracle as O
from fastapsp_amd import _lib as L

pytestmark = pytest.mark.gpu

SSSP = F.FAP_MODE_FLAG_SSSP


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _exact(g):
    lay, csr = g.layout(), g.csr()
    w = None if np.all(csr["weight"] == 1) else csr["weight"]
    return O.exact_rows(csr["row_offset"], csr["col"], w, lay["st2ed"])


def _solve(g, mode):
    g.setMode(mode)
    g.solve()
    info = L.last_stage_times()
    return g.getAllDistanceI32(), info


def _unit_symmetric(g):
    c = g.csr()
    if not np.all(c["weight"] == 1):
        return False
    n = len(c["row_offset"]) - 1
    src = np.repeat(np.arange(n, dtype=np.int64), np.diff(c["row_offset"]))
    fwd = set(zip(src.tolist(), c["col"].tolist()))
    return all((v, u) in fwd for u, v in fwd)


def _forms_agree(g, oracle_path=True):
    a, info = _solve(g, F.FAP_MODE_BATCHED)
    assert info["boundary_graph"] == 1
    bfs = _unit_symmetric(g)
    assert info["parts_bfs"] == (g.K if bfs else 0)
    exact = _exact(g)
    assert np.array_equal(a, exact), "boundary-graph form differs from exact APSP"
    if oracle_path:
        lay, csr = g.layout(), g.csr()
        assert np.array_equal(a, O.solve_parts(1, g.K, csr, lay)), "differs from the oracle"
    b, info = _solve(g, F.FAP_MODE_BATCHED | SSSP)
    assert info["boundary_graph"] == 0
    assert np.array_equal(a, b), "boundary-graph form differs from the SSSP form"
    c, info = _solve(g, F.FAP_MODE_PLAIN)
    assert info["boundary_graph"] == 0
    assert np.array_equal(a, c), "boundary-graph form differs from the plain form"
    d, info = _solve(g, F.FAP_MODE_BATCHED | F.FAP_MODE_FLAG_INT32_FW)
    assert info["boundary_graph"] == 1 and info["parts_fw16"] == 0 and info["parts_bfs"] == 0
    assert np.array_equal(a, d), "int32 in-part Floyd-Warshall differs"
    e, info = _solve(g, F.FAP_MODE_BATCHED | F.FAP_MODE_FLAG_FW)
    assert info["boundary_graph"] == 1 and info["parts_bfs"] == 0
    assert info["parts_fw16"] == (g.K if np.all(g.csr()["weight"] == 1) else 0)
    assert np.array_equal(a, e), "16-bit in-part Floyd-Warshall differs"
    return a


@pytest.mark.parametrize("args", [(64, 64, 0.10, 0.02, 0, 8),    # max b 79: 16-wide chunks
                                  (40, 30, 0.15, 0.05, 3, 19),   # max b 27: 8-wide chunks
                                  (33, 31, 0.10, 0.02, 5, 5)])
def test_boundary_graph_road_grids(args):
    say(f"boundary graph: road grid {args}")
    g = F.FapGraph.road_grid(*args)
    g.preCondition()
    _forms_agree(g)


def test_boundary_graph_road_like():
    # The headline construction at a size the exact check finishes in seconds.
    say("boundary graph: road-like n=9000 K=12")
    g = F.FapGraph.road_like(9000, 0.52, 1e-4, 4, 12)
    g.preCondition()
    a = _forms_agree(g, oracle_path=False)
    assert np.all(np.diag(a[:, g.getMapping()]) == 0)


def _directed_weighted(seed, W=40, H=36, K=7, keep=0.8):
    """Road grid made directed: each arc kept with probability `keep`, weights drawn
    independently per arc from [1, 1000]."""
    base = F.FapGraph.road_grid(W, H, 0.1, 0.02, seed, K)
    c = base.csr()
    rng = np.random.default_rng(seed)
    n = len(c["row_offset"]) - 1
    src = np.repeat(np.arange(n), np.diff(c["row_offset"]))
    mask = rng.random(len(c["col"])) < keep
    src, dst = src[mask], c["col"][mask]
    w = rng.integers(1, 1001, len(dst)).astype(np.float32)
    order = np.argsort(src, kind="stable")
    ro = np.zeros(n + 1, np.int32)
    ro[1:] = np.cumsum(np.bincount(src, minlength=n))
    return F.FapGraph.from_csr(ro, dst[order], w[order], True, K)


@pytest.mark.parametrize("seed", [1, 2])
def test_boundary_graph_directed_weighted(seed):
    say(f"boundary graph: directed weighted road grid seed={seed}")
    g = _directed_weighted(seed)
    g.preCondition()
    a = _forms_agree(g)
    assert (a == F.FAP_INF).any()  # directed: some pairs unreachable


def test_boundary_graph_part_disconnected_inside():
    # Path 0-1-...-199 with part 1 = {0..49, 100..149} and part 2 = the rest: both parts
    # are two pieces inside, joined only through the other part.
    say("boundary graph: parts disconnected inside")
    n = 200
    out = [[] for _ in range(n)]
    for i in range(n - 1):
        out[i].append(i + 1)
        out[i + 1].append(i)
    ro = np.zeros(n + 1, np.int32)
    ro[1:] = np.cumsum([len(x) for x in out])
    col = np.array([v for x in out for v in x], np.int32)
    g = F.FapGraph.from_csr(ro, col, None, False, 2)
    gid = np.where((np.arange(n) // 50) % 2 == 0, 1, 2).astype(np.int32)
    g.setPartition(gid)
    g.preCondition()
    a = _forms_agree(g)
    st2ed = g.getMapping()
    assert np.array_equal(a, np.abs(st2ed[:, None] - np.arange(n)[None, :]).astype(np.int32))


def test_boundary_graph_solve_subgraph_unchanged():
    # solveSubGraph solves one part: the SSSP form, same rows as the full solve.
    say("boundary graph: solve vs solveSubGraph")
    g = F.FapGraph.road_grid(36, 36, 0.1, 0.02, 8, 12)
    g.preCondition()
    full, info = _solve(g, F.FAP_MODE_BATCHED)
    assert info["boundary_graph"] == 1
    lay = g.layout()
    for p in (1, 5, g.K):
        g.solveSubGraph(p)
        off, num = lay["offset"][p], lay["num"][p]
        assert np.array_equal(g.getSubGraphDistanceI32(p), full[off:off + num])


def _path(n):
    out = [[] for _ in range(n)]
    for i in range(n - 1):
        out[i].append(i + 1)
        out[i + 1].append(i)
    ro = np.zeros(n + 1, np.int32)
    ro[1:] = np.cumsum([len(x) for x in out])
    return ro, np.array([v for x in out for v in x], np.int32)


@pytest.mark.parametrize("n,K", [(9000, 3),     # distances < 15872 - s_max: packed merge
                                 (16000, 4)])   # boundary rows reach 15999: int32 merge
def test_boundary_graph_merge_16bit_bound(n, K):
    # Path graph: d(i, j) = |i - j|.  The packed 16-bit merge applies only while every
    # finite boundary-row entry stays below 15872 - s_max; both sides of that bound
    # must give the closed form.
    say(f"boundary graph: path n={n} K={K}")
    ro, col = _path(n)
    g = F.FapGraph.from_csr(ro, col, None, False, K)
    g.preCondition()
    a, info = _solve(g, F.FAP_MODE_BATCHED)
    assert info["boundary_graph"] == 1
    st2ed = g.getMapping().astype(np.int64)
    want = np.abs(st2ed[:, None] - np.arange(n)[None, :]).astype(np.int32)
    assert np.array_equal(a, want)
    b, _ = _solve(g, F.FAP_MODE_BATCHED | SSSP)
    assert np.array_equal(a, b)


@pytest.mark.parametrize("n,K", [(300, 1),      # one part: a path longer than 255 levels
                                 (2048, 1),     # the largest part the BFS form takes
                                 (2049, 1)])    # one vertex more: Floyd-Warshall
def test_inpart_bfs_long_paths_and_size_limit(n, K):
    # Path graph: in-part distances up to n - 1 (levels above 253 take the direct
    # store); parts of more than 2048 vertices fall back to Floyd-Warshall.
    say(f"in-part BFS: path n={n} K={K}")
    ro, col = _path(n)
    g = F.FapGraph.from_csr(ro, col, None, False, K)
    g.preCondition()
    a, info = _solve(g, F.FAP_MODE_BATCHED)
    assert info["boundary_graph"] == 1
    assert info["parts_bfs"] == (K if n <= 2048 else 0)
    st2ed = g.getMapping().astype(np.int64)
    want = np.abs(st2ed[:, None] - np.arange(n)[None, :]).astype(np.int32)
    assert np.array_equal(a, want)


def test_inpart_bfs_high_degree_and_disconnected_parts():
    # A star (hub with 40 in-part neighbours: more than the 8 kept in registers) and a
    # second component inside the same part; K = 2 with a chain linking the parts.
    say("in-part BFS: hub of degree 40, part disconnected inside")
    edges = [(0, i) for i in range(1, 41)]                  # star, part 1
    edges += [(41 + i, 42 + i) for i in range(20)]           # chain 41..61, part 1 (2nd comp.)
    edges += [(62 + i, 63 + i) for i in range(30)]           # chain 62..92, part 2
    edges += [(40, 62), (61, 92)]                            # links through part 2
    n = 93
    out = [[] for _ in range(n)]
    for u, v in edges:
        out[u].append(v)
        out[v].append(u)
    ro = np.zeros(n + 1, np.int32)
    ro[1:] = np.cumsum([len(x) for x in out])
    col = np.array([v for x in out for v in x], np.int32)
    g = F.FapGraph.from_csr(ro, col, None, False, 2)
    gid = np.where(np.arange(n) < 62, 1, 2).astype(np.int32)
    g.setPartition(gid)
    g.preCondition()
    a = _forms_agree(g)
    assert (a < F.FAP_INF).all()  # connected through part 2


def test_boundary_graph_parts_without_inner_vertices():
    # Path of 400 vertices cut into 40 segments alternating 2 and 18 vertices: the
    # 2-vertex parts are all boundary (no merge, boundary rows only), and the 20
    # parts of boundary 2 exceed one merge launch of the boundary-graph form.
    n = 400
    say(f"boundary graph: path n={n}, 40 segments of 2 and 18 vertices")
    ro, col = _path(n)
    lens = [2, 18] * 20
    gid = np.repeat(np.arange(1, len(lens) + 1), lens).astype(np.int32)
    g = F.FapGraph.from_csr(ro, col, None, False, len(lens))
    g.setPartition(gid)
    g.preCondition()
    a, info = _solve(g, F.FAP_MODE_BATCHED)
    assert info["boundary_graph"] == 1
    st2ed = g.getMapping().astype(np.int64)
    want = np.abs(st2ed[:, None] - np.arange(n)[None, :]).astype(np.int32)
    assert np.array_equal(a, want)
    b, _ = _solve(g, F.FAP_MODE_BATCHED | SSSP)
    assert np.array_equal(a, b)


def test_parallel_arcs_follow_the_reference():
    """Parallel in-part arcs whose later arc is heavier: the reference keeps the later
    arc in the build and the lightest in its Dijkstra boundary rows; solve() then takes
    the SSSP form (boundary_graph == 0) and equals the oracle's restatement of the
    reference path, in the single-graph pipeline and in the batch.  With the later arc
    lighter the boundary-graph form applies and equals it too (and exact APSP)."""
    from test_host import _with_parallel_arcs
    base = F.FapGraph.road_grid(24, 24, 0.1, 0.02, 3, 5)
    base.preCondition()
    csr, lay = base.csr(), base.layout()
    for heavier in (True, False):
        ro, col, w = _with_parallel_arcs(csr, lay, heavier_later=heavier, count=40)
        g = F.FapGraph.from_csr(ro, col, w, True, 5)
        g.setPartition(lay["graph_id"])
        g.preCondition()
        assert g.boundaryBySssp() == heavier
        c, l2 = g.csr(), g.layout()
        want = O.solve_parts(1, g.K, c, l2)
        g.solve()
        st = L.last_stage_times()
        assert st["boundary_graph"] == (0 if heavier else 1)
        assert np.array_equal(g.getAllDistanceI32(), want)
        if not heavier:
            exact = O.exact_rows(c["row_offset"], c["col"], c["weight"].astype(np.int32),
                                 l2["st2ed"])
            assert np.array_equal(want, exact)
        p, ld = g.deviceDistance()
        L.check(L.load().fap_hip_fill_i32(p, 99, g.num_vertexs * ld, None), "fill")
        b = F.Batch([g], use_graph=True)
        b.run()
        assert np.array_equal(g.getAllDistanceI32(), want)
        b.close()

"""Predecessor (path) matrices: solve(True) / solveSubGraph(id, True) on the GPU must
equal the oracle's restatement of the reference CPU path (Dijkstra heap order for the
boundary rows, k-i-j Floyd-Warshall with strict '>', first-k min-plus, decode) bit for
bit -- paths and distances -- on graphs with many equal-length paths.  The min-plus
runs in the keyed form (sum and k in one 32-bit key) when a part's boundary has at
most 32 vertices and its sums stay below the key limit, in the compare-and-select
form otherwise; both forms, and calls that mix them, are checked.  The same holds for
the panel phases of the Floyd-Warshall (keyed while (s - 1) * largest built entry stays
below 2^24 - 1)."""
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


def _keyable(lay):
    """Parts with inner rows and a boundary of 1..32 vertices."""
    num, bdy = np.asarray(lay["num"][1:]), np.asarray(lay["bdy"][1:])
    return int(((num > bdy) & (bdy >= 1) & (bdy <= 32)).sum())


def _check_solve(g, validate_rows=64, keyed=None, fw_keyed=None, bgraph=None):
    g.solve(True)
    info = L.last_stage_times()
    if bgraph is not None:  # boundary rows from the boundary graph (else by SSSP)
        assert info["boundary_graph"] == bgraph, info["boundary_graph"]
    if keyed is not None:
        assert info["parts_keyed"] == keyed, (info["parts_keyed"], keyed)
    if fw_keyed is not None:
        assert info["parts_fw_keyed"] == fw_keyed, (info["parts_fw_keyed"], fw_keyed)
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
    # Boundaries of 36..79 (compare form), 10..27 (keyed, 5 tag bits), none (K = 1);
    # unit weights: every part's Floyd-Warshall is keyed.  Undirected unit weights with
    # parts covering the graph: boundary rows from the boundary graph (wide form for
    # boundaries above 32); K = 1 has no boundary row.
    _check_solve(g, keyed=_keyable(g.layout()), fw_keyed=g.K, bgraph=int(g.K > 1))


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
    keyed = _keyable(g.layout())  # boundaries of 7..13: keyed, 4 tag bits
    assert keyed == g.K
    _check_solve(g, keyed=keyed)


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


def test_paths_refused_for_zero_weights():
    """A zero-weight arc makes the reference's predecessor choice follow its heap order:
    path solves refuse such graphs; distance solves still run."""
    base = F.FapGraph.road_grid(20, 20, 0.1, 0.02, 1, 3)
    c = base.csr()
    w = np.ones(len(c["col"]), np.float32)
    w[5] = 0.0
    g = F.FapGraph.from_csr(c["row_offset"], c["col"], w, True, 3)
    g.preCondition()
    with pytest.raises(L.FapError):
        g.solve(True)
    g.solve()
    lay, csr = g.layout(), g.csr()
    want = O.exact_rows(csr["row_offset"], csr["col"], csr["weight"], lay["st2ed"])
    assert np.array_equal(g.getAllDistanceI32(), want)


def _grid(W, H, w=None):
    """W x H 4-neighbour grid, vertex y * W + x, unit weights unless `w` (per arc)."""
    out = [[] for _ in range(W * H)]
    for y in range(H):
        for x in range(W):
            v = y * W + x
            if x + 1 < W:
                out[v].append(v + 1)
                out[v + 1].append(v)
            if y + 1 < H:
                out[v].append(v + W)
                out[v + W].append(v)
    ro = np.zeros(W * H + 1, np.int32)
    ro[1:] = np.cumsum([len(x) for x in out])
    col = np.array([v for x in out for v in x], np.int32)
    wt = None if w is None else np.full(len(col), w, np.float32)
    return ro, col, wt


def test_paths_keyed_and_compare_forms_in_one_call():
    # 48 x 34 grid: rows 0-1 form part 1 (boundary = row 1, 48 vertices: compare form),
    # rows 2..33 are 8 x 8 blocks (boundary <= 28: keyed form).  Grids have many
    # equal-length paths, so the first-k tie rule is exercised in both forms.
    say("paths: keyed and compare forms in one call (grid 48x34, 25 parts)")
    W, H = 48, 34
    ro, col, _ = _grid(W, H)
    y, x = np.divmod(np.arange(W * H), W)
    gid = np.where(y < 2, 1, 2 + ((y - 2) // 8) * (W // 8) + x // 8).astype(np.int32)
    g = F.FapGraph.from_csr(ro, col, None, False, int(gid.max()))
    g.setPartition(gid)
    g.preCondition()
    lay = g.layout()
    assert lay["bdy"][1] == W
    keyed = _keyable(lay)
    assert keyed == int(gid.max()) - 1
    _check_solve(g, keyed=keyed)


@pytest.mark.parametrize("side", ["below", "above"])
def test_paths_key_limit(side):
    # Path graph of 64 vertices in 4 segments, all arcs of weight w: part p's largest
    # A entry is a_p * w and the largest boundary-row entry b * w (hop counts from the
    # unit-weight solve: a_p is 15 in the end segments, 14 in the middle ones).  Part p
    # takes the keyed form while (a_p + b) * w < 2^28 - 1 (at most 8 boundary
    # vertices: 3 tag bits); w is set at the limit of the end segments.  Such distances lie far beyond the
    # oracle's float MAXVALUE (1e8), so both sides are checked against the closed
    # form: d(s, v) = |s - v| * w, predecessor v - 1 (v > s) or v + 1 (v < s).
    n, K, lim = 64, 4, (1 << 28) - 1
    ro, col, _ = _grid(n, 1)
    gid = (np.arange(n) // 16 + 1).astype(np.int32)
    unit = F.FapGraph.from_csr(ro, col, None, False, K)
    unit.setPartition(gid)
    unit.preCondition()
    unit.solve()
    lay = unit.layout()
    d = unit.getAllDistanceI32()  # rows in sorted order, columns by vertex id
    st2ed = lay["st2ed"]
    ap, b = [], 0
    for p in range(1, K + 1):
        off, num, bd = lay["offset"][p], lay["num"][p], lay["bdy"][p]
        bverts = st2ed[off:off + bd]
        b = max(b, int(d[off:off + bd].max()))
        ap.append(int(d[off + bd:off + num][:, bverts].max()))
    a = max(ap)
    w = (lim - 1) // (a + b) + (0 if side == "below" else 1)
    assert ((a + b) * w < lim) == (side == "below")
    keyed = sum((x + b) * w < lim for x in ap)
    assert keyed == (K if side == "below" else 2)
    say(f"paths: key limit {side}: a={a} b={b} w={w}")
    ro, col, wt = _grid(n, 1, float(w))
    assert float(np.float32(w)) == w  # exact in the float32 weight
    g = F.FapGraph.from_csr(ro, col, wt, False, K)
    g.setPartition(gid)
    g.preCondition()
    g.solve(True)
    assert L.last_stage_times()["parts_keyed"] == keyed
    got_d, got_p = g.getAllDistanceI32(), g.getAllPath()
    src = g.layout()["st2ed"].astype(np.int64)[:, None]
    v = np.arange(n)[None, :]
    assert np.array_equal(got_d, (np.abs(src - v) * w).astype(np.int32))
    want_p = np.where(v > src, v - 1, np.where(v < src, v + 1, src))
    assert np.array_equal(got_p, want_p.astype(np.int32))


def test_paths_fw_compare_form():
    # knn with weights up to 100000: (s - 1) * largest built entry exceeds 2^24 - 1, so
    # the Floyd-Warshall panels run in the compare form (distances stay far below 2^24,
    # where the oracle's float arithmetic is exact).
    say("paths: knn, weights up to 1e5 (compare-form Floyd-Warshall)")
    g = F.FapGraph.knn(1800, 3, 100000, 7, 5)
    g.preCondition()
    d, _ = _check_solve(g, fw_keyed=0)
    assert d[d < F.FAP_INF].max() < (1 << 24)


def test_paths_fw_forms_in_one_call():
    # The 48 x 34 grid of the mixed merge test with weight 200000 on the arcs inside
    # part 1 (rows 0-1): part 1's Floyd-Warshall runs in the compare form, the 24
    # unit-weight blocks keyed.
    say("paths: keyed and compare Floyd-Warshall in one call")
    W, H, heavy = 48, 34, 200000
    ro, col, _ = _grid(W, H)
    y, x = np.divmod(np.arange(W * H), W)
    gid = np.where(y < 2, 1, 2 + ((y - 2) // 8) * (W // 8) + x // 8).astype(np.int32)
    src = np.repeat(np.arange(W * H), np.diff(ro))
    wt = np.where((gid[src] == 1) & (gid[col] == 1), heavy, 1).astype(np.float32)
    g = F.FapGraph.from_csr(ro, col, wt, False, int(gid.max()))
    g.setPartition(gid)
    g.preCondition()
    _check_solve(g, keyed=int(gid.max()) - 1, fw_keyed=int(gid.max()) - 1)

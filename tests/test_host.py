"""Host-side logic on the CPU: .mtx reader, generators, partitioner, boundary /
sorted layout, and the API's error behaviour (no GPU needed)."""
import ctypes as C
import os

import numpy as np
import pytest

import fastapsp_amd as F
from fastapsp_amd import _lib as L
import oracle as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def read_mtx(path, directed, weighted):
    lib = F.load()
    n, m = C.c_int32(), C.c_int64()
    L.check(lib.fap_mtx_read(path.encode(), directed, weighted, C.byref(n), C.byref(m), None, None,
                             None), "mtx")
    ro = np.empty(n.value + 1, np.int32)
    col = np.empty(max(m.value, 1), np.int32)
    w = np.empty(max(m.value, 1), np.float32)
    L.check(lib.fap_mtx_read(path.encode(), directed, weighted, C.byref(n), C.byref(m),
                             L.ptr_i32(ro), L.ptr_i32(col), w.ctypes.data_as(C.POINTER(C.c_float))),
            "mtx")
    return ro, col[:m.value], w[:m.value]


def test_mtx_reader_matches_reference_golden(tmp_path, monkeypatch):
    g = np.load(os.path.join(GOLD, "mtx.npz"), allow_pickle=False)
    monkeypatch.chdir(tmp_path)
    for kind, directed, weighted, sub in [("uu", 0, 0, "unweight-undirected"),
                                          ("wd", 1, 1, "weight-directed")]:
        os.makedirs(tmp_path / "graph" / sub, exist_ok=True)
        (tmp_path / "graph" / sub / (kind + ".mtx")).write_bytes(g[kind + "_text"].tobytes())
        ro, col, w = read_mtx(kind, directed, weighted)  # reference path convention
        assert np.array_equal(ro, g[kind + "_row_offset"])
        assert np.array_equal(col, g[kind + "_col"])  # per-vertex reverse file order
        assert np.array_equal(w, g[kind + "_weight"])  # |w| as float
        ro2, col2, _ = read_mtx(str(tmp_path / "graph" / sub / (kind + ".mtx")), directed, weighted)
        assert np.array_equal(ro2, ro) and np.array_equal(col2, col)


def test_mtx_errors(tmp_path):
    lib = F.load()
    n, m = C.c_int32(), C.c_int64()
    assert lib.fap_mtx_read(str(tmp_path / "none.mtx").encode(), 0, 0, C.byref(n), C.byref(m),
                            None, None, None) == L.FAP_ERR_IO
    bad = tmp_path / "bad.mtx"
    bad.write_text("3 3 2\n1 2\n2 9\n")
    assert lib.fap_mtx_read(str(bad).encode(), 0, 0, C.byref(n), C.byref(m), None, None,
                            None) == L.FAP_ERR_IO


def test_layout_matches_reference_golden():
    g = np.load(os.path.join(GOLD, "layout.npz"), allow_pickle=False)
    gf = F.FapGraph.from_csr(g["row_offset"], g["col"], None, False, int(g["K"]))
    lay = _layout_for(gf, g["graph_id"], int(g["K"]))
    assert np.array_equal(lay["st2ed"], g["st2ed"])
    assert np.array_equal(lay["ed2st"], g["ed2st"])
    assert np.array_equal(lay["num"], g["num"]) and np.array_equal(lay["offset"], g["off"])
    assert np.array_equal(lay["bdy"], g["bdy"])


def _layout_for(graph, graph_id, K):
    """The product's findBoundry + sort_and_encode for a given partition."""
    csr = graph.csr()
    n = graph.num_vertexs
    out = {k: np.zeros(n if k in ("st2ed", "ed2st") else K + 1, np.int32)
           for k in ("st2ed", "ed2st", "num", "offset", "bdy")}
    gid = np.ascontiguousarray(graph_id, np.int32)
    L.check(F.load().fap_layout_for_partition(
        n, L.ptr_i32(csr["row_offset"]), L.ptr_i32(np.ascontiguousarray(csr["col"])), K,
        L.ptr_i32(gid), L.ptr_i32(out["st2ed"]), L.ptr_i32(out["ed2st"]), L.ptr_i32(out["num"]),
        L.ptr_i32(out["offset"]), L.ptr_i32(out["bdy"])), "layout")
    return out


def test_set_partition_gives_reference_layout():
    g = np.load(os.path.join(GOLD, "layout.npz"), allow_pickle=False)
    gf = F.FapGraph.from_csr(g["row_offset"], g["col"], None, False, int(g["K"]))
    gf.setPartition(g["graph_id"])
    gf.preCondition()
    lay = gf.layout()
    assert np.array_equal(lay["st2ed"], g["st2ed"]) and np.array_equal(lay["bdy"], g["bdy"])


@pytest.mark.skipif(O.ref() is None, reason="oracle/_ref not built (reference tree absent)")
def test_layout_matches_live_reference_on_road_grid():
    g = F.FapGraph.road_grid(30, 30, 0.1, 0.02, 5, 7)
    g.preCondition()
    lay, csr = g.layout(), g.csr()
    n, K = g.num_vertexs, g.K
    P = lambda a, t: a.ctypes.data_as(C.POINTER(t))  # noqa: E731
    outs = {k: np.zeros(n if k in ("isb", "st2ed", "ed2st") else K + 1, np.int32)
            for k in ("isb", "num", "off", "bdy", "bdy_off", "st2ed", "ed2st")}
    ro = csr["row_offset"]
    O.ref().ref_layout(K, n, P(lay["graph_id"], C.c_int32), P(np.diff(ro).astype(np.int32), C.c_int32),
                       P(ro, C.c_int32), P(csr["col"], C.c_int32),
                       P(np.ones(len(csr["col"]), np.float32), C.c_float),
                       *[P(outs[k], C.c_int32) for k in ("isb", "num", "off", "bdy", "bdy_off",
                                                         "st2ed", "ed2st")])
    assert np.array_equal(lay["st2ed"], outs["st2ed"])
    assert np.array_equal(lay["ed2st"], outs["ed2st"])
    assert np.array_equal(lay["num"], outs["num"]) and np.array_equal(lay["bdy"], outs["bdy"])


def test_layout_invariants():
    g = F.FapGraph.road_grid(50, 40, 0.1, 0.02, 2, 9)
    g.preCondition()
    lay, csr = g.layout(), g.csr()
    n = g.num_vertexs
    st2ed, ed2st, gid = lay["st2ed"], lay["ed2st"], lay["graph_id"]
    assert np.array_equal(np.sort(st2ed), np.arange(n))
    assert np.array_equal(ed2st[st2ed], np.arange(n))
    ro, col = csr["row_offset"], csr["col"]
    src = np.repeat(np.arange(n), np.diff(ro))
    cut = gid[src] != gid[col]
    isb = np.zeros(n, bool)
    isb[src[cut]] = True
    isb[col[cut]] = True
    for p in range(1, g.K + 1):
        off, num, bdy = lay["offset"][p], lay["num"][p], lay["bdy"][p]
        block = st2ed[off:off + num]
        assert np.all(gid[block] == p)
        assert np.all(isb[block[:bdy]]) and not np.any(isb[block[bdy:]])
    assert lay["num"][1:].sum() == n and lay["offset"][1] == 0
    assert np.all(np.diff(lay["offset"][1:]) == lay["num"][1:-1])


def test_generators_deterministic_and_connected():
    a = F.FapGraph.road_grid(64, 64, 0.10, 0.02, 0, 8)
    b = F.FapGraph.road_grid(64, 64, 0.10, 0.02, 0, 8)
    ca, cb = a.csr(), b.csr()
    assert a.num_vertexs == b.num_vertexs and np.array_equal(ca["col"], cb["col"])
    d = O.exact_rows(ca["row_offset"], ca["col"], None, [0])
    assert np.all(d < F.FAP_INF)  # largest component only
    r = F.FapGraph.road_like(20000, 0.52, 1e-4, 1, 16)
    assert r.num_vertexs == 20000
    k = F.FapGraph.knn(5000, 3, 1000, 2, 8)
    ck = k.csr()
    assert np.all(np.diff(ck["row_offset"]) == 3)
    assert ck["weight"].min() >= 1 and ck["weight"].max() <= 1000


def test_road_like_real_shape():
    g = F.FapGraph.road_like(114599, 0.52, 1e-4, 1, 128)
    assert g.num_vertexs == 114599
    assert 2.0 < g.num_edges / g.num_vertexs < 2.5


def test_partition_balance_and_ids():
    g = F.FapGraph.road_like(30000, 0.52, 1e-4, 3, 32)
    gid = g.getGraphId()
    counts = np.bincount(gid, minlength=33)[1:]
    assert gid.min() == 1 and gid.max() == 32 and np.all(counts > 0)
    assert counts.max() <= 1.03 * counts.mean() + 1
    assert g.getEdgeCut() > 0


def test_api_errors():
    with pytest.raises(F.FapError) as e:
        F.FapGraph.road_grid(8, 8, 0.0, 0.0, 0, 1000)  # K > n
    assert e.value.status == L.FAP_ERR_INVALID
    with pytest.raises(F.FapError) as e:
        F.FapGraph("no_such_graph", False, False, 4)
    assert e.value.status == L.FAP_ERR_IO
    ro = np.array([0, 1, 2], np.int32)
    col = np.array([1, 0], np.int32)
    # Non-integer weights select the float path; negative or NaN weights are refused.
    assert F.FapGraph.from_csr(ro, col, np.array([1.5, 1.5], np.float32), True, 1).isFloat()
    assert not F.FapGraph.from_csr(ro, col, np.array([2.0, 7.0], np.float32), True, 1).isFloat()
    for bad in ([-1.0, 1.0], [np.nan, 1.0], [np.inf, 1.0]):
        with pytest.raises(F.FapError) as e:
            F.FapGraph.from_csr(ro, col, np.array(bad, np.float32), True, 1)
        assert e.value.status == L.FAP_ERR_INVALID
    g = F.FapGraph.from_csr(ro, col, None, False, 1)
    with pytest.raises(F.FapError) as e:
        g.solve()  # before preCondition
    assert e.value.status == L.FAP_ERR_STATE
    lib = F.load()
    h = C.c_void_p()
    assert lib.fap_graph_create_from_file(C.byref(h), b"x", 0, 0, 2, b"KaHIP") in (
        L.FAP_ERR_IO, L.FAP_ERR_INVALID)


def _with_parallel_arcs(csr, lay, heavier_later=True, count=12, seed=5):
    """The graph's CSR with a second arc appended after `count` in-part arcs u -> v
    (same row, later in CSR order), 3 heavier (or 1 lighter) than the first."""
    ro, col, w = csr["row_offset"], csr["col"], csr["weight"].astype(np.float32)
    gid = lay["graph_id"]
    rng = np.random.default_rng(seed)
    n = len(ro) - 1
    cand = [e for u in range(n) for e in range(ro[u], ro[u + 1]) if gid[col[e]] == gid[u]]
    pick = set(rng.choice(len(cand), size=count, replace=False).tolist())
    extra = {cand[i] for i in pick}
    nro, ncol, nw = [0], [], []
    for u in range(n):
        for e in range(ro[u], ro[u + 1]):
            ncol.append(col[e])
            nw.append(w[e])
        for e in range(ro[u], ro[u + 1]):
            if e in extra:
                ncol.append(col[e])
                nw.append(w[e] + 3 if heavier_later else max(w[e] - 1, 1))
        nro.append(len(ncol))
    return (np.array(nro, np.int32), np.array(ncol, np.int32), np.array(nw, np.float32))


def test_parallel_arcs_later_heavier_detected_and_reference_semantics():
    """preCondition flags a graph whose parallel in-part arcs have a heavier later arc
    (not one whose later arc is lighter, nor the plain graph); on such a graph the
    reference CPU path (the oracle's restatement, later arc kept in the build) differs
    from exact APSP, which is why solve() must not take the boundary-graph form there."""
    base = F.FapGraph.road_grid(16, 16, 0.1, 0.02, 3, 4)
    base.preCondition()
    assert not base.boundaryBySssp()
    csr, lay = base.csr(), base.layout()
    ro, col, w = _with_parallel_arcs(csr, lay, heavier_later=True)
    g = F.FapGraph.from_csr(ro, col, w, True, 4)
    g.setPartition(lay["graph_id"])
    g.preCondition()
    assert g.boundaryBySssp()
    ro2, col2, w2 = _with_parallel_arcs(csr, lay, heavier_later=False)
    g2 = F.FapGraph.from_csr(ro2, col2, w2, True, 4)
    g2.setPartition(lay["graph_id"])
    g2.preCondition()
    assert not g2.boundaryBySssp()
    c, l2 = g.csr(), g.layout()
    ref, _ = O.solve_parts_path(1, g.K, c, l2)
    exact = O.exact_rows(c["row_offset"], c["col"], c["weight"].astype(np.int32), l2["st2ed"])
    assert not np.array_equal(ref, exact)

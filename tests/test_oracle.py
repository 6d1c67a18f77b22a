"""Pins the oracle (oracle/fap_oracle.cpp, the restatement of the reference's CPU
path) against the reference's own compiled kernels (golden vectors in tests/golden,
plus oracle/_ref live when it is built), closed forms and an independent exact APSP."""
import ctypes as C
import os

import numpy as np
import pytest

import fastapsp_amd as F
import oracle as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def gold(name):
    return np.load(os.path.join(GOLD, name), allow_pickle=False)


def test_dijkstra_path_matches_reference_golden():
    g = gold("dijkstra.npz")
    for i, s in enumerate(g["sources"]):
        d, p = O.dijkstra_path(int(s), g["row_offset"], g["col"], g["weight"])
        assert np.array_equal(d, g["dist"][i])
        assert np.array_equal(p, g["path"][i])


def test_floyd_path_matches_reference_golden():
    g = gold("floyd.npz")
    m, p = O.floyd_path(g["mat"], g["path"])
    assert np.array_equal(m, g["mat_out"])
    assert np.array_equal(p, g["path_out"])


def test_minplus_advanced_matches_reference_golden():
    g = gold("minplus.npz")
    c, cp = O.minplus_advanced(g["A"], g["B"], g["Bpath"])
    assert np.array_equal(c, g["C"])
    assert np.array_equal(cp, g["Cpath"])


def _grid_csr(W, H):
    nbr = [[] for _ in range(W * H)]
    for r in range(H):
        for c in range(W):
            v = r * W + c
            if c + 1 < W:
                nbr[v].append(v + 1)
                nbr[v + 1].append(v)
            if r + 1 < H:
                nbr[v].append(v + W)
                nbr[v + W].append(v)
    ro = np.zeros(W * H + 1, np.int32)
    ro[1:] = np.cumsum([len(x) for x in nbr])
    return ro, np.array([v for x in nbr for v in x], np.int32)


def test_dijkstra_closed_form_manhattan():
    W, H = 17, 11
    ro, col = _grid_csr(W, H)
    r, c = np.divmod(np.arange(W * H), W)
    for s in (0, 40, W * H - 1):
        d, p = O.dijkstra_path(s, ro, col, np.ones(len(col), np.float32))
        sr, sc = divmod(s, W)
        assert np.array_equal(d, (np.abs(r - sr) + np.abs(c - sc)).astype(np.float32))
        # predecessor is one step closer to the source
        for v in range(W * H):
            if v != s:
                assert d[p[v]] == d[v] - 1


def test_floyd_brute_force_tiny():
    rng = np.random.default_rng(1)
    for s in (1, 2, 4, 7):
        m = rng.integers(1, 9, (s, s)).astype(np.float32)
        m[rng.random((s, s)) < 0.5] = O.MAXVALUE
        np.fill_diagonal(m, 0)
        out, _ = O.floyd_path(m, np.zeros((s, s), np.int32))
        # brute force over all simple paths via repeated min-plus squaring
        d = m.astype(np.float64).copy()
        for _ in range(s):
            d = np.minimum(d, (d[:, :, None] + d[None, :, :]).min(axis=1))
        d[d >= O.MAXVALUE] = O.MAXVALUE
        assert np.array_equal(np.minimum(out, O.MAXVALUE), d.astype(np.float32))


def test_minplus_closed_form():
    rng = np.random.default_rng(2)
    M, N, K = 9, 31, 6
    alpha, beta = rng.integers(0, 50, M), rng.integers(0, 50, N)
    k = np.arange(K)
    A = (alpha[:, None] + k).astype(np.float32)
    B = (beta[None, :] + k[:, None]).astype(np.float32)
    bp = np.tile(np.arange(N, dtype=np.int32), (K, 1)) + 1000 * k[:, None].astype(np.int32)
    c, cp = O.minplus_advanced(A, B, bp)
    assert np.array_equal(c, (alpha[:, None] + beta[None, :]).astype(np.float32))
    assert np.array_equal(cp, np.tile(np.arange(N, dtype=np.int32), (M, 1)))  # k = 0 wins ties


def test_minplus_identity():
    rng = np.random.default_rng(3)
    B = rng.integers(0, 100, (8, 20)).astype(np.float32)
    I = np.full((8, 8), O.MAXVALUE, np.float32)
    np.fill_diagonal(I, 0)
    c, _ = O.minplus_advanced(I, B, np.zeros((8, 20), np.int32))
    assert np.array_equal(c, B)


@pytest.mark.skipif(O.ref() is None, reason="oracle/_ref not built (reference tree absent)")
def test_oracle_kernels_match_live_reference_random():
    ref = O.ref()
    rng = np.random.default_rng(4)
    P = lambda a, t: a.ctypes.data_as(C.POINTER(t))  # noqa: E731
    for trial in range(5):
        n = int(rng.integers(20, 120))
        m = n * 4
        a, b = rng.integers(0, n, m), rng.integers(0, n, m)
        keep = a != b
        a, b = a[keep], b[keep]
        order = np.argsort(a, kind="stable")
        a, b = a[order], b[order]
        ro = np.cumsum(np.concatenate([[0], np.bincount(a, minlength=n)])).astype(np.int32)
        col = b.astype(np.int32)
        w = rng.integers(1, 20, len(col)).astype(np.float32)
        adj = np.diff(ro).astype(np.int32)
        for s in range(0, n, 7):
            d0 = np.empty(n, np.float32)
            p0 = np.empty(n, np.int32)
            ref.ref_dijkstra_path(s, n, P(adj, C.c_int32), P(ro, C.c_int32), P(col, C.c_int32),
                                  P(w, C.c_float), P(d0, C.c_float), P(p0, C.c_int32))
            d1, p1 = O.dijkstra_path(s, ro, col, w)
            assert np.array_equal(d0, d1) and np.array_equal(p0, p1)
        s = int(rng.integers(2, 70))
        mat = rng.integers(1, 30, (s, s)).astype(np.float32)
        mat[rng.random((s, s)) < 0.7] = O.MAXVALUE
        np.fill_diagonal(mat, 0)
        path = rng.integers(-1, s, (s, s)).astype(np.int32)
        m0, q0 = mat.copy(), path.copy()
        ref.ref_floyd_cpu_path(s, P(m0, C.c_float), P(q0, C.c_int32))
        m1, q1 = O.floyd_path(mat, path)
        assert np.array_equal(m0, m1) and np.array_equal(q0, q1)


def test_full_path_equals_exact_apsp_unweighted():
    """orc_run_part over all parts (SSSP, build, FW, mat1, min-plus, decode) equals an
    independent BFS APSP -- a wrong index in build / mat1 / decode breaks this."""
    import fastapsp_amd as F

    g = F.FapGraph.road_grid(40, 40, 0.1, 0.02, 7, 6)
    g.preCondition()
    lay, csr = g.layout(), g.csr()
    got = O.solve_parts(1, g.K, csr, lay)
    want = O.exact_rows(csr["row_offset"], csr["col"], None, lay["st2ed"])
    assert np.array_equal(got, want)


def test_full_path_equals_exact_apsp_weighted_directed():
    import fastapsp_amd as F

    g = F.FapGraph.knn(900, 3, 1000, 3, 5)
    g.preCondition()
    lay, csr = g.layout(), g.csr()
    got = O.solve_parts(1, g.K, csr, lay)
    want = O.exact_rows(csr["row_offset"], csr["col"], csr["weight"], lay["st2ed"])
    assert np.array_equal(got, want)


def test_exact_rows_detects_a_planted_error():
    """The exact-APSP comparison is sensitive: a one-entry change in the oracle
    result (a wrong index in decode, say) is caught."""
    import fastapsp_amd as F

    g = F.FapGraph.road_grid(20, 20, 0.1, 0.02, 1, 4)
    g.preCondition()
    lay, csr = g.layout(), g.csr()
    got = O.solve_parts(1, g.K, csr, lay)
    want = O.exact_rows(csr["row_offset"], csr["col"], None, lay["st2ed"])
    bad = got.copy()
    bad[3, lay["st2ed"][5]] += 1
    assert np.array_equal(got, want) and not np.array_equal(bad, want)


def test_exact_rows_f64_pins():
    """The FP64 oracle of the float path: equal to the integer Dijkstra on integer
    weights, and to a numpy float64 Floyd-Warshall on small float-weighted digraphs
    (float32 weights in [0.5, 10): every path sum is exact in float64)."""
    g = F.FapGraph.knn(800, 3, 50, 5, 4)
    csr = g.csr()
    src = np.arange(0, 800, 37)
    a = O.exact_rows(csr["row_offset"], csr["col"], csr["weight"], src)
    b = O.exact_rows_f64(csr["row_offset"], csr["col"], csr["weight"].astype(np.float32), src)
    assert np.array_equal(np.where(np.isinf(b), F.FAP_INF, b).astype(np.int64), a)
    rng = np.random.default_rng(11)
    for n, m in ((40, 120), (60, 150)):
        u = rng.integers(0, n, m)
        v = rng.integers(0, n, m)
        w = (0.5 + 9.5 * rng.random(m)).astype(np.float32)
        order = np.argsort(u, kind="stable")
        ro = np.zeros(n + 1, np.int32)
        ro[1:] = np.cumsum(np.bincount(u, minlength=n))
        col, wt = v[order].astype(np.int32), w[order]
        D = np.full((n, n), np.inf)
        for a_, b_, w_ in zip(u, v, w):
            D[a_, b_] = min(D[a_, b_], float(w_))
        np.fill_diagonal(D, 0.0)
        for k in range(n):
            D = np.minimum(D, D[:, k:k + 1] + D[k:k + 1, :])
        rows = O.exact_rows_f64(ro, col, wt, np.arange(n))
        assert np.array_equal(rows, D)


def test_oracle_paths_are_valid_predecessors():
    """The oracle's path blocks (reference CPU path with is_path_needed) name, for every
    reachable entry, an in-neighbour on a shortest path; unreachable entries are -1 and
    the diagonal is the source itself.  Distances equal the exact APSP."""
    import fastapsp_amd as F
    g = F.FapGraph.road_grid(24, 20, 0.1, 0.05, 9, 4)
    g.preCondition()
    lay, csr = g.layout(), g.csr()
    dist, path = O.solve_parts_path(1, g.K, csr, lay)
    src = lay["st2ed"]
    assert np.array_equal(dist, O.exact_rows(csr["row_offset"], csr["col"], None, src))
    ro, col = csr["row_offset"], csr["col"]
    arcs = {(u, int(col[e])) for u in range(len(ro) - 1) for e in range(ro[u], ro[u + 1])}
    for r, s in enumerate(src):
        for v in range(dist.shape[1]):
            if dist[r, v] >= O.INF:
                assert path[r, v] == -1
            elif v == s:
                assert path[r, v] == s
            else:
                u = int(path[r, v])
                assert (u, v) in arcs and dist[r, u] + 1 == dist[r, v]


def test_streaming_comparators_agree_and_catch_planted_errors():
    """compare_exact_rows / compare_parts (the full-size checkers) report zero on the
    oracle's own result, and count a planted one-entry change in the checked rows --
    in a boundary row and in an inner row, at a row stride wider than n."""
    import fastapsp_amd as F

    for g, w in [(F.FapGraph.road_grid(30, 30, 0.1, 0.02, 4, 5), None),
                 (F.FapGraph.knn(700, 3, 1000, 5, 4), "w")]:
        g.preCondition()
        lay, csr = g.layout(), g.csr()
        n = g.num_vertexs
        wt = None if w is None else csr["weight"]
        full = O.solve_parts(1, g.K, csr, lay)
        wide = np.full((n, n + 13), -7, np.int32)
        wide[:, :n] = full
        # parts 2..K-1 only, in a block whose row 0 is part 2's first row
        r0 = int(lay["offset"][2])
        r1 = int(lay["offset"][g.K - 1] + lay["num"][g.K - 1])
        blk = wide[r0:r1]
        assert O.compare_exact_rows(csr["row_offset"], csr["col"], wt, lay["st2ed"][r0:r1],
                                    blk) == (0, -1)
        assert O.compare_parts(2, g.K - 1, csr, lay, blk, r0) == (0, -1)
        for r in (int(lay["offset"][3]), int(lay["offset"][3] + lay["bdy"][3] + 1)):
            bad = blk.copy()
            bad[r - r0, int(lay["st2ed"][(r * 7) % n])] += 1
            assert O.compare_exact_rows(csr["row_offset"], csr["col"], wt,
                                        lay["st2ed"][r0:r1], bad) == (1, r - r0)
            assert O.compare_parts(2, g.K - 1, csr, lay, bad, r0) == (1, r - r0)


def test_int32_kernels_pinned_to_reference_float_kernels():
    """orc_floyd_i32 / orc_minplus_i32 (the int32 checkers of the GPU kernel tests) equal
    the reference's floyd_CPU_path / min_plus_path_CPU_advanced on integer inputs below
    2^24 (exact in float; MAXVALUE <-> INT32_MAX / 2): through the golden vectors written
    from the compiled reference, and live against it when it is built."""
    g = gold("floyd.npz")
    m = O.to_i32(g["mat"])
    assert np.array_equal(O.floyd_i32(m), O.to_i32(g["mat_out"]))
    g = gold("minplus.npz")
    assert np.array_equal(O.minplus_i32(O.to_i32(g["A"]), O.to_i32(g["B"])), O.to_i32(g["C"]))
    ref = O.ref()
    if ref is None:
        pytest.skip("compiled reference kernels not built (oracle/_ref)")
    P = lambda a, t: a.ctypes.data_as(C.POINTER(t))  # noqa: E731
    rng = np.random.default_rng(21)
    for trial in range(6):
        s = int(rng.integers(2, 90))
        mat = rng.integers(0, 1 << 12, (s, s)).astype(np.float32)
        mat[rng.random((s, s)) < 0.6] = O.MAXVALUE
        np.fill_diagonal(mat, 0)
        m0, q0 = mat.copy(), np.zeros((s, s), np.int32)
        ref.ref_floyd_cpu_path(s, P(m0, C.c_float), P(q0, C.c_int32))
        assert np.array_equal(O.floyd_i32(O.to_i32(mat)), O.to_i32(m0))
        M, N, K = (int(x) for x in rng.integers(1, 70, 3))
        a = rng.integers(0, 1 << 20, (M, K)).astype(np.float32)
        b = rng.integers(0, 1 << 20, (K, N)).astype(np.float32)
        a[rng.random((M, K)) < 0.3] = O.MAXVALUE
        b[rng.random((K, N)) < 0.3] = O.MAXVALUE
        c = np.empty((M, N), np.float32)
        cp = np.empty((M, N), np.int32)
        bp = np.zeros((K, N), np.int32)
        ref.ref_min_plus_cpu_advanced(P(a, C.c_float), P(b, C.c_float), P(bp, C.c_int32),
                                      P(c, C.c_float), P(cp, C.c_int32), M, N, K)
        assert np.array_equal(O.minplus_i32(O.to_i32(a), O.to_i32(b)), O.to_i32(c))

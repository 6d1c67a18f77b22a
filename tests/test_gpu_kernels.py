"""Kernel-level parity through the C-ABI: min-plus, Floyd-Warshall, boundary SSSP,
sub-matrix build / decode, against the oracle on the same seeded inputs (bit-exact)."""
import ctypes as C
import sys

import numpy as np
import pytest

import fastapsp_amd as F
from fastapsp_amd import _lib as L
import oracle as O

INF = F.FAP_INF
pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rand_dist(rng, shape, hi=1 << 20, p_inf=0.1):
    a = rng.integers(0, hi, size=shape, dtype=np.int64).astype(np.int32)
    a[rng.random(shape) < p_inf] = INF
    return a


def gpu_minplus(a, b):
    M, K = a.shape
    N = b.shape[1]
    da = F.DeviceBuffer.from_array(a if K else np.zeros((M, 1), np.int32))
    db = F.DeviceBuffer.from_array(b if K else np.zeros((1, N), np.int32))
    dc = F.DeviceBuffer(M * N * 4)
    L.check(L.load().fap_hip_minplus(da.ptr, max(K, 1), db.ptr, N, dc.ptr, N, M, N, K, None),
            "minplus")
    return dc.download((M, N))


@pytest.mark.parametrize("M,N,K", [(1, 1, 1), (64, 256, 32), (65, 257, 33), (130, 600, 70),
                                   (7, 1000, 0), (200, 3, 5), (64, 512, 1), (3, 4099, 64)])
def test_minplus_random(M, N, K):
    say(f"minplus M={M} N={N} K={K}")
    rng = np.random.default_rng(M * 1000 + N + K)
    a = rand_dist(rng, (M, K))
    b = rand_dist(rng, (K, N))
    got = gpu_minplus(a, b)
    want = O.minplus_i32(a, b) if K else np.full((M, N), INF, np.int32)
    assert np.array_equal(got, want)


def test_minplus_closed_form():
    # A[i][k] = alpha_i + k, B[k][j] = beta_j + k  ->  C[i][j] = alpha_i + beta_j (k = 0).
    say("minplus closed form")
    rng = np.random.default_rng(7)
    M, N, K = 97, 700, 45
    alpha = rng.integers(0, 1000, M)
    beta = rng.integers(0, 1000, N)
    k = np.arange(K)
    a = (alpha[:, None] + k[None, :]).astype(np.int32)
    b = (beta[None, :] + k[:, None]).astype(np.int32)
    got = gpu_minplus(a, b)
    assert np.array_equal(got, (alpha[:, None] + beta[None, :]).astype(np.int32))


def test_minplus_strided():
    # lda / ldb / ldc larger than the logical sizes, unaligned B / C columns.
    say("minplus strided")
    rng = np.random.default_rng(11)
    M, N, K, lda, ldb, ldc = 70, 301, 19, 23, 305, 303
    a = rand_dist(rng, (M, lda))
    b = rand_dist(rng, (K, ldb))
    c0 = rand_dist(rng, (M, ldc))
    da, db, dc = (F.DeviceBuffer.from_array(x) for x in (a, b, c0))
    L.check(L.load().fap_hip_minplus(da.ptr, lda, db.ptr, ldb, dc.ptr, ldc, M, N, K, None), "mp")
    got = dc.download((M, ldc))
    want = c0.copy()
    want[:, :N] = O.minplus_i32(a[:, :K], b[:, :N])
    assert np.array_equal(got, want)


@pytest.mark.parametrize("s", [1, 2, 5, 63, 64, 65, 130, 200, 300])
def test_floyd_random(s):
    say(f"floyd s={s}")
    rng = np.random.default_rng(s)
    m = rand_dist(rng, (s, s), hi=1000, p_inf=0.7)
    np.fill_diagonal(m, 0)
    d = F.DeviceBuffer.from_array(m)
    L.check(L.load().fap_hip_floyd(d.ptr, s, s, None), "floyd")
    assert np.array_equal(d.download((s, s)), O.floyd_i32(m))


def test_floyd_strided_path_graph():
    # Path graph 0-1-...-(s-1): d(i, j) = |i - j| (closed form), ld > s.
    say("floyd path graph")
    s, ld = 150, 157
    m = np.full((s, ld), INF, np.int32)
    for i in range(s):
        m[i, i] = 0
        if i + 1 < s:
            m[i, i + 1] = m[i + 1, i] = 1
    d = F.DeviceBuffer.from_array(m)
    L.check(L.load().fap_hip_floyd(d.ptr, s, ld, None), "floyd")
    got = d.download((s, ld))
    idx = np.arange(s)
    assert np.array_equal(got[:, :s], np.abs(idx[:, None] - idx[None, :]).astype(np.int32))
    assert np.all(got[:, s:] == INF)


def _sssp(ro, col, w, sources, ld=None):
    n = len(ro) - 1
    ld = ld or n
    dro, dcol = F.DeviceBuffer.from_array(ro), F.DeviceBuffer.from_array(col)
    dw = F.DeviceBuffer.from_array(w) if w is not None else None
    src = np.ascontiguousarray(sources, np.int32)
    dsrc = F.DeviceBuffer.from_array(src)
    out = F.DeviceBuffer(len(src) * ld * 4)
    L.check(L.load().fap_hip_sssp_batch(dro.ptr, dcol.ptr, dw.ptr if dw else None, n, len(col),
                                        dsrc.ptr, len(src), out.ptr, ld, None), "sssp")
    return out.download((len(src), ld))[:, :n]


def test_sssp_bfs_road_grid():
    say("sssp bfs road grid")
    g = F.FapGraph.road_grid(64, 64, 0.1, 0.02, 0, 8)
    c = g.csr()
    rng = np.random.default_rng(3)
    src = rng.choice(g.num_vertexs, 300, replace=False)
    got = _sssp(c["row_offset"], c["col"], None, src)
    assert np.array_equal(got, O.exact_rows(c["row_offset"], c["col"], None, src))


def test_sssp_bfs_full_grid_manhattan():
    # Full W x H grid (no drops): BFS distance = Manhattan distance (closed form).
    say("sssp manhattan")
    W, H = 50, 40
    g = F.FapGraph.road_grid(W, H, 0.0, 0.0, 1, 4)
    c = g.csr()
    src = np.array([0, 17, W * H - 1, 1234], np.int32)
    got = _sssp(c["row_offset"], c["col"], None, src, ld=W * H + 3)
    r, q = np.divmod(np.arange(W * H), W)
    for i, s in enumerate(src):
        sr, sq = divmod(int(s), W)
        assert np.array_equal(got[i], (np.abs(r - sr) + np.abs(q - sq)).astype(np.int32))


def test_sssp_frontier_overflow_star():
    # A star with 5000 leaves: the level-1 frontier overflows the LDS queue and the
    # next level runs as a row scan.  Leaves also form a path to give depth.
    say("sssp star overflow")
    leaves = 5000
    n = leaves + 1
    arcs = [(0, i) for i in range(1, n)] + [(i, i + 1) for i in range(1, leaves)]
    ro, col = _csr_undirected(n, arcs)
    src = np.array([0, 1, leaves // 2], np.int32)
    got = _sssp(ro, col, None, src)
    assert np.array_equal(got, O.exact_rows(ro, col, None, src))


def test_sssp_weighted_knn():
    say("sssp weighted knn")
    g = F.FapGraph.knn(3000, 3, 1000, 5, 8)
    c = g.csr()
    src = np.random.default_rng(5).choice(g.num_vertexs, 100, replace=False)
    got = _sssp(c["row_offset"], c["col"], c["weight"], src)
    assert np.array_equal(got, O.exact_rows(c["row_offset"], c["col"], c["weight"], src))


def test_sssp_weighted_overflow():
    # Weighted star: the frontier overflow path of the label-correcting kernel.
    say("sssp weighted overflow")
    leaves = 3000
    n = leaves + 1
    rng = np.random.default_rng(9)
    arcs = [(0, i) for i in range(1, n)] + [(i, i + 1) for i in range(1, leaves)]
    ro, col = _csr_undirected(n, arcs)
    w = rng.integers(1, 1000, len(col)).astype(np.int32)
    src = np.array([0, 7], np.int32)
    got = _sssp(ro, col, w, src)
    assert np.array_equal(got, O.exact_rows(ro, col, w, src))


def _csr_undirected(n, arcs):
    out = [[] for _ in range(n)]
    for a, b in arcs:
        out[a].append(b)
        out[b].append(a)
    ro = np.zeros(n + 1, np.int32)
    ro[1:] = np.cumsum([len(x) for x in out])
    col = np.array([v for x in out for v in x], np.int32)
    return ro, col

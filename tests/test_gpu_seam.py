"""The kernel-level seam through the C-ABI against the oracle's restatement of the
reference's CPU branch, bit for bit, distances and predecessors:
* fap_hip_floyd_path    vs floyd_CPU_path            (fap/kernel/floyd.h:38-49)
* fap_hip_minplus_path  vs min_plus_path_CPU_advanced (fap/kernel/minplus.h:51-64)
* fap_hip_sssp_batch_path vs dijkstra_path            (fap/kernel/batch_sssp.h:79-114)
* fap_hip_submat_build  vs MysubMatBuild_path         (fap/utils/util_Myalgorithm.h:44-84)
* fap_hip_submat_decode vs MysubMatDecode_path        (fap/utils/util_Myalgorithm.h:236-247)
* fap_hip_f32_to_i32 / fap_hip_i32_to_f32: MAXVALUE <-> FAP_INF at the seam's edges.
Inputs use small weights so that ties are everywhere: the predecessor chosen on a tie
is what these tests pin."""
import ctypes as C
import sys

import numpy as np
import pytest

import fastapsp_amd as F
from fastapsp_amd import _lib as L
import oracle as O

INF = F.FAP_INF
MAXV = np.float32(1e8)
pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def dev(a):
    return F.DeviceBuffer.from_array(np.ascontiguousarray(a))


def to_f(d):
    return np.where(d >= INF, MAXV, d.astype(np.float32)).astype(np.float32)


def reverse_csr(ro, col, w):
    n = len(ro) - 1
    tails = np.repeat(np.arange(n), np.diff(ro))
    order = np.argsort(col, kind="stable")
    rro = np.zeros(n + 1, np.int32)
    rro[1:] = np.cumsum(np.bincount(col, minlength=n))
    return rro, tails[order].astype(np.int32), None if w is None else w[order].astype(np.int32)


@pytest.mark.parametrize("s", [1, 5, 64, 65, 130, 257])
def test_floyd_path_matches_oracle(s):
    say(f"floyd_path s={s}")
    rng = np.random.default_rng(100 + s)
    ld = s + 3
    d = rng.integers(1, 4, size=(s, s)).astype(np.int32)
    d[rng.random((s, s)) < 0.85] = INF
    np.fill_diagonal(d, 0)
    p = np.where(d < INF, np.arange(s)[:, None], -1).astype(np.int32)
    np.fill_diagonal(p, np.arange(s))
    want_d, want_p = O.floyd_path(to_f(d), p)
    hd = np.full((s, ld), -9, np.int32)
    hp = np.full((s, ld), -9, np.int32)
    hd[:, :s], hp[:, :s] = d, p
    bd, bp = dev(hd), dev(hp)
    L.check(L.load().fap_hip_floyd_path(bd.ptr, bp.ptr, s, ld, None), "floyd_path")
    gd, gp = bd.download((s, ld)), bp.download((s, ld))
    assert np.array_equal(gd[:, :s], O.to_i32(want_d))
    assert np.array_equal(gp[:, :s], want_p)
    assert np.all(gd[:, s:] == -9) and np.all(gp[:, s:] == -9)


@pytest.mark.parametrize("M,N,K", [(1, 1, 1), (64, 64, 16), (65, 130, 37), (200, 333, 5),
                                   (7, 1000, 0), (130, 70, 70)])
def test_minplus_path_matches_oracle(M, N, K):
    say(f"minplus_path M={M} N={N} K={K}")
    rng = np.random.default_rng(M * 7 + N * 3 + K)
    a = rng.integers(0, 6, size=(M, K)).astype(np.int32)
    a[rng.random((M, K)) < 0.2] = INF
    b = rng.integers(0, 6, size=(K, N)).astype(np.int32)
    b[rng.random((K, N)) < 0.2] = INF
    bp = rng.integers(0, 10**6, size=(K, N)).astype(np.int32)
    if K:
        wc, wp = O.minplus_advanced(to_f(a), to_f(b), bp)
    else:
        wc, wp = np.full((M, N), MAXV, np.float32), np.full((M, N), -1, np.int32)
    lda, ldb, ldc = K + 1, N + 2, N + 5
    ha = np.full((M, lda), 7, np.int32)
    ha[:, :K] = a
    hb = np.full((max(K, 1), ldb), 7, np.int32)
    hbp = np.full((max(K, 1), ldb), 7, np.int32)
    hb[:K, :N], hbp[:K, :N] = b, bp
    da, db, dbp = dev(ha), dev(hb), dev(hbp)
    dc = dev(np.full((M, ldc), -9, np.int32))
    dcp = dev(np.full((M, ldc), -9, np.int32))
    L.check(L.load().fap_hip_minplus_path(da.ptr, lda, db.ptr, ldb, dbp.ptr, ldb, dc.ptr, ldc,
                                          dcp.ptr, ldc, M, N, K, None), "minplus_path")
    gc, gp = dc.download((M, ldc)), dcp.download((M, ldc))
    assert np.array_equal(gc[:, :N], O.to_i32(wc))
    assert np.array_equal(gp[:, :N], wp)
    assert np.all(gc[:, N:] == -9) and np.all(gp[:, N:] == -9)


@pytest.mark.parametrize("kind", ["unit", "weighted"])
def test_sssp_batch_path_matches_oracle(kind):
    say(f"sssp_batch_path {kind}")
    if kind == "unit":
        g = F.FapGraph.road_grid(48, 40, 0.1, 0.02, 3, 4)
        csr = g.csr()
        w = None
    else:
        g = F.FapGraph.knn(3000, 3, 3, 4, 4)  # weights 1..3: ties everywhere
        csr = g.csr()
        w = csr["weight"].astype(np.int32)
    ro, col = csr["row_offset"], csr["col"]
    n, m = len(ro) - 1, len(col)
    rng = np.random.default_rng(5)
    src = rng.choice(n, size=37, replace=False).astype(np.int32)
    rro, rcol, rw = reverse_csr(ro, col, w)
    ld = n + 7
    bufs = [dev(x) for x in (ro, col, rro, rcol, src)]
    dw = dev(w) if w is not None else None
    drw = dev(rw) if rw is not None else None
    dout = F.DeviceBuffer(len(src) * ld * 4)
    dpath = F.DeviceBuffer(len(src) * ld * 4)
    L.check(L.load().fap_hip_sssp_batch_path(
        bufs[0].ptr, bufs[1].ptr, dw.ptr if dw else None, bufs[2].ptr, bufs[3].ptr,
        drw.ptr if drw else None, n, m, bufs[4].ptr, len(src), dout.ptr, dpath.ptr, ld, None),
        "sssp_batch_path")
    gd = dout.download((len(src), ld))[:, :n]
    gp = dpath.download((len(src), ld))[:, :n]
    wf = np.ones(m, np.float32) if w is None else w.astype(np.float32)
    for i, s in enumerate(src):
        dd, pp = O.dijkstra_path(int(s), ro, col, wf)
        assert np.array_equal(gd[i], O.to_i32(dd)), f"row {i} distances"
        assert np.array_equal(gp[i], pp), f"row {i} predecessors"


def test_submat_build_and_decode_match_oracle():
    say("submat build / decode")
    for g in [F.FapGraph.road_grid(40, 30, 0.1, 0.02, 9, 5), F.FapGraph.knn(1500, 3, 50, 9, 5)]:
        g.preCondition()
        lay, csr = g.layout(), g.csr()
        n = g.num_vertexs
        ro, col = csr["row_offset"], csr["col"]
        unit = bool(np.all(csr["weight"] == 1))
        wf = csr["weight"].astype(np.float32)
        bufs = {k: dev(lay[k]) for k in ("graph_id", "st2ed", "ed2st")}
        dro, dcol = dev(ro), dev(col)
        dw = None if unit else dev(csr["weight"].astype(np.int32))
        for p in (1, 3, g.K):
            s, bdy, start = int(lay["num"][p]), int(lay["bdy"][p]), int(lay["offset"][p])
            rows = [O.dijkstra_path(int(v), ro, col, wf) for v in lay["st2ed"][start:start + bdy]]
            dr = np.stack([r[0] for r in rows]) if bdy else np.zeros((0, n), np.float32)
            pr = np.stack([r[1] for r in rows]) if bdy else np.zeros((0, n), np.int32)
            want, _ = O.submat_build(dr, pr, csr, lay, p)
            ldd, lds = n + 3, s + 5
            hrows = np.full((max(bdy, 1), ldd), -4, np.int32)
            hrows[:bdy, :n] = O.to_i32(dr)
            drows = dev(hrows)
            dsub = dev(np.full((s, lds), -9, np.int32))
            L.check(L.load().fap_hip_submat_build(
                dsub.ptr, lds, drows.ptr, ldd, dro.ptr, dcol.ptr, dw.ptr if dw else None,
                bufs["graph_id"].ptr, bufs["st2ed"].ptr, bufs["ed2st"].ptr, n, start, s, bdy,
                None), "submat_build")
            got = dsub.download((s, lds))
            assert np.array_equal(got[:, :s], O.to_i32(want)), f"part {p} build"
            # decode: scatter the (built) sub-matrix rows into final columns
            rng = np.random.default_rng(p)
            subm = rng.integers(0, 1000, size=(s, s)).astype(np.int32)
            blk = np.full((s, n), -2, np.float32)
            pblk = np.zeros((s, n), np.int32)
            O.submat_decode(subm.astype(np.float32), np.zeros((s, s), np.int32), lay["st2ed"],
                            start, blk, pblk)
            ldo = n + 9
            dout = dev(np.full((s, ldo), -2, np.int32))
            dsm = dev(subm)
            L.check(L.load().fap_hip_submat_decode(dsm.ptr, s, dout.ptr, ldo, bufs["st2ed"].ptr,
                                                   start, s, s, None), "submat_decode")
            got = dout.download((s, ldo))
            assert np.array_equal(got[:, :n], blk.astype(np.int32)), f"part {p} decode"
            assert np.all(got[:, n:] == -2)


def test_distance_conversions():
    say("f32 <-> i32 conversions")
    rng = np.random.default_rng(3)
    ints = rng.integers(0, 1 << 24, size=100003).astype(np.int32)
    ints[::7] = INF
    f = to_f(ints)
    f[::11] = np.float32(3e8)  # anything at or above MAXVALUE is unreachable
    df, di = dev(f), F.DeviceBuffer(f.size * 4)
    L.check(L.load().fap_hip_f32_to_i32(df.ptr, di.ptr, f.size, None), "f32_to_i32")
    got = di.download(f.shape)
    assert np.array_equal(got, O.to_i32(f))
    dback = F.DeviceBuffer(f.size * 4)
    L.check(L.load().fap_hip_i32_to_f32(di.ptr, dback.ptr, f.size, None), "i32_to_f32")
    back = dback.download(f.shape, np.float32)
    assert np.array_equal(back, to_f(got))


def _check_layout(csr, lay, K):
    n = len(csr["row_offset"]) - 1
    bufs = [dev(np.ascontiguousarray(x, np.int32)) for x in
            (csr["row_offset"], csr["col"], lay["graph_id"], lay["st2ed"], lay["ed2st"])]
    g = L.FapGraphDev(n, len(csr["col"]), bufs[0].ptr, bufs[1].ptr, None, bufs[2].ptr,
                      bufs[3].ptr, bufs[4].ptr, None, None)
    parts = (L.FapPartDesc * K)(*[L.FapPartDesc(p, int(lay["offset"][p]), int(lay["num"][p]),
                                                int(lay["bdy"][p]), int(lay["offset"][p]))
                                  for p in range(1, K + 1)])
    ok = C.c_int32(-1)
    L.check(L.load().fap_hip_check_layout(C.byref(g), parts, K, C.byref(ok)), "check_layout")
    return ok.value


def test_check_layout_accepts_valid_and_rejects_corrupt():
    """fap_hip_check_layout: the layout fapGraph builds passes; a label moved to another
    part, a boundary count one short (an inner vertex with a cut arc) and two swapped
    st2ed entries each fail; a part list that does not tile [0, n) fails."""
    say("check_layout")
    g = F.FapGraph.road_grid(30, 30, 0.1, 0.02, 4, 5)
    g.preCondition()
    csr, lay = g.csr(), g.layout()
    assert _check_layout(csr, lay, g.K) == 1
    bad = {k: v.copy() for k, v in lay.items()}
    v = int(bad["st2ed"][int(bad["offset"][2]) + int(bad["bdy"][2])])  # an inner vertex of part 2
    bad["graph_id"][v] = 3
    assert _check_layout(csr, bad, g.K) == 0
    bad = {k: v.copy() for k, v in lay.items()}
    p = int(np.argmax(bad["bdy"][1:])) + 1
    bad["bdy"][p] -= 1
    assert _check_layout(csr, bad, g.K) == 0
    bad = {k: v.copy() for k, v in lay.items()}
    bad["st2ed"][[0, 1]] = bad["st2ed"][[1, 0]]
    assert _check_layout(csr, bad, g.K) == 0
    bad = {k: v.copy() for k, v in lay.items()}
    bad["num"][1] -= 1
    assert _check_layout(csr, bad, g.K) == 0

"""Writes golden vectors from the reference's own CPU kernels.

Runs oracle/_ref/libfapref.so -- dijkstra_path, floyd_CPU_path,
min_plus_path_CPU_advanced, findBoundry + sort_and_encode and readMatFile of
TomerHaham/FastAPSP, compiled from its sources by `make -C oracle ref` -- on
seeded synthetic inputs and stores inputs and outputs as .npz (plain arrays,
loadable with allow_pickle=False).  Run from the repository root:
    python tests/golden/make_golden.py
"""
import ctypes as C
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
import oracle as O  # noqa: E402

P = lambda a, t: a.ctypes.data_as(C.POINTER(t))  # noqa: E731


def random_graph(rng, n, m, wmax):
    a = rng.integers(0, n, m)
    b = rng.integers(0, n, m)
    keep = a != b
    a, b = a[keep], b[keep]
    order = np.argsort(a, kind="stable")
    a, b = a[order], b[order]
    ro = np.zeros(n + 1, np.int32)
    np.add.at(ro, a + 1, 1)
    ro = np.cumsum(ro).astype(np.int32)
    w = rng.integers(1, wmax + 1, len(b)).astype(np.float32)
    return ro, b.astype(np.int32), w


def main():
    ref = O.ref()
    if ref is None:
        raise SystemExit("oracle/_ref/libfapref.so is not built (make -C oracle ref)")
    rng = np.random.default_rng(20231)

    # dijkstra_path on a directed weighted graph (sources 0, 5, 17)
    n = 60
    ro, col, w = random_graph(rng, n, 240, 9)
    adj = np.diff(ro).astype(np.int32)
    dists, paths, srcs = [], [], np.array([0, 5, 17], np.int32)
    for s in srcs:
        d = np.empty(n, np.float32)
        p = np.empty(n, np.int32)
        ref.ref_dijkstra_path(int(s), n, P(adj, C.c_int32), P(ro, C.c_int32), P(col, C.c_int32),
                              P(w, C.c_float), P(d, C.c_float), P(p, C.c_int32))
        dists.append(d)
        paths.append(p)
    np.savez(os.path.join(HERE, "dijkstra.npz"), row_offset=ro, col=col, weight=w, sources=srcs,
             dist=np.array(dists), path=np.array(paths))

    # floyd_CPU_path on a 40 x 40 matrix (0 diagonal, MAXVALUE holes, integer weights)
    s = 40
    mat = rng.integers(1, 50, (s, s)).astype(np.float32)
    mat[rng.random((s, s)) < 0.8] = O.MAXVALUE
    np.fill_diagonal(mat, 0)
    path = np.tile(np.arange(s, dtype=np.int32), (s, 1))
    path[mat >= O.MAXVALUE] = -1
    m_out, p_out = mat.copy(), path.copy()
    ref.ref_floyd_cpu_path(s, P(m_out, C.c_float), P(p_out, C.c_int32))
    np.savez(os.path.join(HERE, "floyd.npz"), mat=mat, path=path, mat_out=m_out, path_out=p_out)

    # min_plus_path_CPU_advanced, M x N x K = 13 x 50 x 7
    M, N, K = 13, 50, 7
    A = rng.integers(0, 100, (M, K)).astype(np.float32)
    B = rng.integers(0, 100, (K, N)).astype(np.float32)
    B[rng.random((K, N)) < 0.3] = O.MAXVALUE
    Bp = rng.integers(0, 1000, (K, N)).astype(np.int32)
    Cc = np.empty((M, N), np.float32)
    Cp = np.empty((M, N), np.int32)
    ref.ref_min_plus_cpu_advanced(P(A, C.c_float), P(B, C.c_float), P(Bp, C.c_int32),
                                  P(Cc, C.c_float), P(Cp, C.c_int32), M, N, K)
    np.savez(os.path.join(HERE, "minplus.npz"), A=A, B=B, Bpath=Bp, C=Cc, Cpath=Cp)

    # findBoundry + sort_and_encode on a 10 x 10 grid with a fixed 4-part partition
    W = 10
    arcs = []
    for r in range(W):
        for c in range(W):
            v = r * W + c
            if c + 1 < W:
                arcs.append((v, v + 1))
            if r + 1 < W:
                arcs.append((v, v + W))
    nn = W * W
    nbr = [[] for _ in range(nn)]
    for a, b in arcs:
        nbr[a].append(b)
        nbr[b].append(a)
    ro2 = np.zeros(nn + 1, np.int32)
    ro2[1:] = np.cumsum([len(x) for x in nbr])
    col2 = np.array([v for x in nbr for v in x], np.int32)
    w2 = np.ones(len(col2), np.float32)
    gid = np.array([1 + (r >= 5) * 2 + (c >= 6) for r in range(W) for c in range(W)], np.int32)
    Kp = 4
    outs = {k: np.zeros(nn if k in ("isb", "st2ed", "ed2st") else Kp + 1, np.int32)
            for k in ("isb", "num", "off", "bdy", "bdy_off", "st2ed", "ed2st")}
    ref.ref_layout(Kp, nn, P(gid, C.c_int32), P(np.diff(ro2).astype(np.int32), C.c_int32),
                   P(ro2, C.c_int32), P(col2, C.c_int32), P(w2, C.c_float),
                   *[P(outs[k], C.c_int32) for k in ("isb", "num", "off", "bdy", "bdy_off",
                                                     "st2ed", "ed2st")])
    np.savez(os.path.join(HERE, "layout.npz"), row_offset=ro2, col=col2, graph_id=gid, K=Kp,
             **outs)

    # readMatFile: a small undirected unweighted and a directed weighted .mtx
    with tempfile.TemporaryDirectory() as tmp:
        cwd = os.getcwd()
        os.chdir(tmp)
        try:
            res = {}
            for kind, directed, weighted, text in [
                ("uu", 0, 0, "%%MatrixMarket\n% comment\n6 6 7\n1 2\n2 3\n3 1\n4 5\n5 6\n6 4\n1 4\n"),
                ("wd", 1, 1, "%%MatrixMarket\n5 5 8\n1 2 3.0\n2 3 -4\n3 4 1.5\n4 5 2\n5 1 7\n1 3 9\n"
                             "2 5 1\n1 2 2\n"),
            ]:
                sub = {(0, 0): "unweight-undirected", (1, 1): "weight-directed"}[(directed, weighted)]
                os.makedirs(os.path.join("graph", sub), exist_ok=True)
                with open(os.path.join("graph", sub, kind + ".mtx"), "w") as f:
                    f.write(text)
                nv, mv = C.c_int(), C.c_int()
                ref.ref_read_mtx(kind.encode(), directed, weighted, C.byref(nv), C.byref(mv),
                                 None, None, None, None)
                adj = np.zeros(nv.value, np.int32)
                ro3 = np.zeros(nv.value + 1, np.int32)
                col3 = np.zeros(mv.value, np.int32)
                w3 = np.zeros(mv.value, np.float32)
                ref.ref_read_mtx(kind.encode(), directed, weighted, C.byref(nv), C.byref(mv),
                                 P(adj, C.c_int32), P(ro3, C.c_int32), P(col3, C.c_int32),
                                 P(w3, C.c_float))
                res[kind + "_text"] = np.frombuffer(text.encode(), np.uint8)
                res[kind + "_row_offset"] = ro3
                res[kind + "_col"] = col3
                res[kind + "_weight"] = w3
        finally:
            os.chdir(cwd)
    np.savez(os.path.join(HERE, "mtx.npz"), **res)
    print("golden vectors written to", HERE)


if __name__ == "__main__":
    main()

"""The measured workloads at their stated sizes, checked against the oracle:
* the dense 8192 x 8192 x 8192 int32 min-plus (values uniform in [0, 2^20]): 256 full
  rows of C against the oracle's triple loop;
* the many-graph batch: 512 road-like graphs, n log-uniform in [256, 8192], half
  unweighted (int32, exact against BFS), half with float32 weights in [0.5, 10) (against
  an FP64 Dijkstra at relative 1e-6, and equal to its float32 rounding), K per size class
  from the autotuner, one captured launch sequence replayed twice;
* the request server with graphs up to 60k vertices and a 200 GB HBM budget: a Poisson
  mix of full-APSP and sub-graph requests, every sub-graph block checked in full and
  every APSP result in full (graphs up to 5k vertices) or on sampled rows.
"""
import sys
import time

import numpy as np
import pytest

import fastapsp_amd as F
import oracle as O
from fastapsp_amd import _lib as L
from fastapsp_amd import autotune as AT
from fastapsp_amd import serve as S

pytestmark = pytest.mark.gpu
REL_TOL = 1e-6


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.timeout(600)
def test_minplus_8192_full_rows():
    N = 8192
    say(f"min-plus {N}^3, values in [0, 2^20]")
    rng = np.random.default_rng(3)
    A = rng.integers(0, 1 << 20, size=(N, N), dtype=np.int32)
    B = rng.integers(0, 1 << 20, size=(N, N), dtype=np.int32)
    dA, dB = F.DeviceBuffer.from_array(A), F.DeviceBuffer.from_array(B)
    dC = F.DeviceBuffer(N * N * 4)
    L.check(L.load().fap_hip_minplus(dA.ptr, N, dB.ptr, N, dC.ptr, N, N, N, N, None), "minplus")
    C = dC.download((N, N))
    rows = np.sort(rng.choice(N, size=256, replace=False))
    rows[0], rows[-1] = 0, N - 1
    want = O.minplus_i32(A[rows], B)
    assert np.array_equal(C[rows], want)


def check_int_graph(g):
    lay, csr = g.layout(), g.csr()
    n = g.num_vertexs
    got = g.getDistanceRowsI32(0, n)
    w = None if np.all(csr["weight"] == 1) else csr["weight"]
    bad, first = O.compare_exact_rows(csr["row_offset"], csr["col"], w, lay["st2ed"], got)
    assert bad == 0, f"n={n}: {bad} rows differ (first {first})"


def check_float_rows(got, want):
    unreach = np.isinf(want)
    assert np.array_equal(got == F.MAXVALUE, unreach)
    w = want[~unreach]
    rel = np.abs(got[~unreach].astype(np.float64) - w) / np.maximum(w, 1e-30)
    assert rel.max(initial=0.0) <= REL_TOL
    assert np.array_equal(got[~unreach], w.astype(np.float32))


def check_float_graph(g):
    lay, csr = g.layout(), g.csr()
    n = g.num_vertexs
    step = max(1, (1 << 27) // (8 * n))  # oracle rows in blocks of <= 128 MB
    for r0 in range(0, n, step):
        r1 = min(n, r0 + step)
        got = g.getDistanceRowsF32(r0, r1 - r0)
        want = O.exact_rows_f64(csr["row_offset"], csr["col"], csr["weight"], lay["st2ed"][r0:r1])
        check_float_rows(got, want)


@pytest.mark.timeout(900)
def test_batch512_stated_workload():
    t0 = time.time()
    table = AT.tune(k_candidates=(8, 16), src_rows_candidates=(16,), merge_rows_candidates=(16,),
                    max_copies=16)
    say(f"autotune {time.time() - t0:.0f}s: K per class {table['K']}")
    rng = np.random.default_rng(2026)
    sizes = np.exp(rng.uniform(np.log(256), np.log(8192), 512)).astype(int)
    graphs = [AT.road_graph(int(nt), 5000 + i, AT.k_for(int(nt), table), float_weights=(i % 2 == 1))
              for i, nt in enumerate(sizes)]
    for g in graphs:
        g.preCondition()
    ns = [g.num_vertexs for g in graphs]
    assert len(graphs) == 512 and min(ns) < 400 and max(ns) > 7000
    say(f"512 graphs, n in [{min(ns)}, {max(ns)}], sum n^2 = {sum(n * n for n in ns):.3e}")
    b = F.Batch(graphs, use_graph=True, **AT.settings_for(ns, table))
    for rep in range(2):
        for g in graphs:
            p, ld = g.deviceDistance()
            if p:
                L.check(L.load().fap_hip_fill_i32(p, 4321 + rep, g.num_vertexs * ld, None), "fill")
        b.run()
        say(f"replay {rep} done ({time.time() - t0:.0f}s)")
    for i, g in enumerate(graphs):
        if g.isFloat():
            check_float_graph(g)
        else:
            check_int_graph(g)
        if i % 64 == 63:
            say(f"  {i + 1}/512 graphs equal the oracle ({time.time() - t0:.0f}s)")
    b.close()


def result_rows(srv, t, r0, rows):
    p, nrows, cols, ld, fl = srv.result_device(t)
    assert 0 <= r0 and r0 + rows <= nrows
    out = np.empty((rows, cols), np.int32)
    L.check(L.load().fap_hip_memcpy2d_d2h_i32(L.ptr_i32(out), cols, p + r0 * ld * 4, ld, rows,
                                              cols, None), "result rows")
    return out.view(np.float32) if fl else out


def expected_rows(g, src):
    csr = g.csr()
    if g.isFloat():
        return O.exact_rows_f64(csr["row_offset"], csr["col"], csr["weight"], src)
    w = None if np.all(csr["weight"] == 1) else csr["weight"]
    return O.exact_rows(csr["row_offset"], csr["col"], w, src)


@pytest.mark.timeout(900)
def test_serve_large_graphs_200gb_budget():
    say("server: graphs up to 60k vertices, 200 GB budget")
    pool = []
    big = AT.road_graph(60000, 31, 67)  # road-like, integer weights 1..1000 (symmetric)
    c = big.csr()
    w = AT.symmetric_float_weights(c["row_offset"], c["col"], 31)
    w = (np.floor((w - 0.5) / 9.5 * 1000) + 1).astype(np.float32)
    pool.append(F.FapGraph.from_csr(c["row_offset"], c["col"], w, False, 67))
    k = F.FapGraph.knn(30000, 3, 1000, 32, 34)  # kNN, float weights in [0.5, 10)
    c = k.csr()
    pool.append(F.FapGraph.from_csr(c["row_offset"], c["col"],
                                    AT.symmetric_float_weights(c["row_offset"], c["col"], 32, True),
                                    True, 34))
    pool.append(AT.road_graph(12000, 33, 14, float_weights=True))
    pool.append(F.FapGraph.knn(3000, 3, 1000, 34, 8))
    pool.append(AT.road_graph(2000, 35, 5))
    for g in pool:
        g.preCondition()
    assert pool[0].num_vertexs > 55000 and not pool[0].isFloat() and pool[1].isFloat()
    srv = S.Server(budget_bytes=200e9, streams=4)
    idx = [srv.add_graph(g) for g in pool]
    rng = np.random.default_rng(9)
    reqs = [(idx[0], 0), (idx[0], 0), (idx[1], 0)]
    for _ in range(9):
        reqs.append((idx[0], int(rng.integers(1, pool[0].K + 1))))
        reqs.append((idx[1], int(rng.integers(1, pool[1].K + 1))))
    for gi in (2, 3, 4):
        reqs += [(idx[gi], 0), (idx[gi], int(rng.integers(1, pool[gi].K + 1)))]
    order = rng.permutation(len(reqs))
    reqs = [reqs[i] for i in order]
    tickets, states, wall = S.poisson_load(srv, reqs, 50.0, seed=5, keep=True, timeout=300.0)
    st = srv.stats()
    say(f"{len(reqs)} requests in {wall:.2f}s, p50 {st['p50_ms']:.1f} ms, p99 {st['p99_ms']:.1f} "
        f"ms, peak reserved {st['peak_reserved_bytes'] / 1e9:.1f} GB")
    assert all(s == S.DONE for s in states)
    assert st["peak_reserved_bytes"] <= 200e9
    for (gi, sub), t in zip(reqs, tickets):
        g = pool[idx.index(gi)]
        lay = g.layout()
        if sub > 0:
            off, num = int(lay["offset"][sub]), int(lay["num"][sub])
            rows = np.arange(num)
            src = lay["st2ed"][off:off + num]
        elif g.num_vertexs <= 5000:
            rows = np.arange(g.num_vertexs)
            src = lay["st2ed"]
        else:
            rows = np.unique(np.concatenate([[0, g.num_vertexs - 1],
                                             rng.integers(0, g.num_vertexs, 30)]))
            src = lay["st2ed"][rows]
        want = expected_rows(g, src)
        if sub > 0 or g.num_vertexs <= 5000:
            got = result_rows(srv, t, 0, len(rows))
        else:
            got = np.stack([result_rows(srv, t, int(r), 1)[0] for r in rows])
        if g.isFloat():
            check_float_rows(got, want)
        else:
            assert np.array_equal(got, want), (gi, sub)
        srv.release(t)
    srv.close()

"""The request server: every result is the same alone and in a mix (bit for bit: int32
results are exact, float32 results are the deterministic rounding of FP64 sums), equal to
the oracle, and admission keeps within the HBM budget; a request that can never fit
fails instead of blocking the queue."""
import sys

import numpy as np
import pytest

import fastapsp_amd as F
import oracle as O
from fastapsp_amd import serve as S
from fastapsp_amd.autotune import symmetric_float_weights

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def graphs():
    out = []
    out.append(F.FapGraph.road_grid(40, 36, 0.1, 0.02, 1, 6))
    base = F.FapGraph.road_grid(44, 30, 0.1, 0.02, 2, 5)
    c = base.csr()
    out.append(F.FapGraph.from_csr(c["row_offset"], c["col"],
                                   symmetric_float_weights(c["row_offset"], c["col"], 2), False, 5))
    out.append(F.FapGraph.knn(1800, 3, 1000, 3, 4))
    k = F.FapGraph.knn(1500, 3, 1000, 5, 3)
    c = k.csr()
    out.append(F.FapGraph.from_csr(c["row_offset"], c["col"],
                                   symmetric_float_weights(c["row_offset"], c["col"], 5, True),
                                   True, 3))
    out.append(F.FapGraph.road_grid(96, 96, 0.1, 0.02, 6, 20))  # >= 8192: pipeline lane
    for g in out:
        g.preCondition()
    assert out[-1].num_vertexs >= 8192
    return out


def expected(g, sub):
    lay, csr = g.layout(), g.csr()
    if sub == 0:
        src = lay["st2ed"]
    else:
        off, num = lay["offset"][sub], lay["num"][sub]
        src = lay["st2ed"][off:off + num]
    if g.isFloat():
        want = O.exact_rows_f64(csr["row_offset"], csr["col"], csr["weight"], src)
        return np.where(np.isinf(want), F.MAXVALUE, want).astype(np.float32)
    w = None if np.all(csr["weight"] == 1) else csr["weight"]
    return O.exact_rows(csr["row_offset"], csr["col"], w, src)


def test_serve_alone_equals_mix_and_oracle():
    gs = graphs()
    srv = S.Server(budget_bytes=40e9, streams=4)
    idx = [srv.add_graph(g) for g in gs]
    reqs = []
    for i, g in zip(idx, gs):
        reqs += [(i, 0), (i, 1), (i, g.K)]
    say(f"{len(reqs)} requests alone")
    alone = {}
    for r in reqs:
        t = srv.submit(*r)
        assert srv.wait(t) == S.DONE
        alone[r] = srv.result(t)
        srv.release(t)
    for (i, sub), res in alone.items():
        want = expected(gs[i], sub)
        assert res.shape == want.shape and np.array_equal(res, want), (i, sub)
    say("burst mix")
    tickets = [srv.submit(*r) for r in reqs]
    for r, t in zip(reqs, tickets):
        assert srv.wait(t) == S.DONE
        assert np.array_equal(srv.result(t), alone[r]), r
        srv.release(t)
    say("paced mix (Poisson, 400 req/s)")
    order = [reqs[j] for j in np.random.default_rng(1).permutation(len(reqs))] * 2
    tickets, states, _ = S.poisson_load(srv, order, 400.0, seed=3, keep=True)
    for r, t, st in zip(order, tickets, states):
        assert st == S.DONE
        assert np.array_equal(srv.result(t), alone[r]), r
        srv.release(t)
    st = srv.stats()
    assert st["completed"] == 4 * len(reqs) and st["failed"] == 0  # alone, burst, 2x paced
    assert st["p99_ms"] >= st["p50_ms"] > 0 and st["reserved_bytes"] == 0
    srv.close()


def test_serve_budget_admission():
    g = F.FapGraph.road_grid(30, 30, 0.1, 0.02, 4, 4)
    g.preCondition()
    n = g.num_vertexs
    one = n * ((n + 63) // 64 * 64) * 4
    # Room for about one full APSP (result + scratch) at a time.
    srv = S.Server(budget_bytes=one * 1.6 + (24 << 20), streams=4)
    i = srv.add_graph(g)
    # Submitted together; each is admitted once the previous result is released.
    tickets = [srv.submit(i, 0) for _ in range(6)]
    ref = None
    for t in tickets:
        assert srv.wait(t) == S.DONE
        res = srv.result(t)
        ref = res if ref is None else ref
        assert np.array_equal(res, ref)
        srv.release(t)
    st = srv.stats()
    assert st["peak_reserved_bytes"] <= one * 1.6 + (24 << 20)
    assert np.array_equal(ref, expected(g, 0))
    srv.close()
    tiny = S.Server(budget_bytes=1 << 20, streams=2)
    j = tiny.add_graph(g)
    t = tiny.submit(j, 0)
    assert tiny.wait(t) == S.FAILED
    tiny.release(t)
    tiny.close()


def test_solve_after_server_with_pipeline_lane_closed():
    """The pipeline lane solves on the server's own stream, which fap_server_destroy
    destroys; a later solve() on the library stream must order itself after that work
    without touching the destroyed stream."""
    big = F.FapGraph.road_grid(96, 96, 0.1, 0.02, 7, 16)
    big.preCondition()
    assert big.num_vertexs >= 8192
    srv = S.Server(budget_bytes=40e9, streams=2)
    i = srv.add_graph(big)
    t = srv.submit(i, 1)
    assert srv.wait(t) == S.DONE
    lane = srv.result(t)
    srv.release(t)
    srv.close()
    assert np.array_equal(lane, expected(big, 1))
    g = F.FapGraph.road_grid(30, 28, 0.1, 0.02, 9, 5)
    g.preCondition()
    g.solve()
    assert np.array_equal(g.getAllDistanceI32(), expected(g, 0))
    big.solveSubGraph(2)
    assert np.array_equal(big.getSubGraphDistanceI32(2), expected(big, 2))

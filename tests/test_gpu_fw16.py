"""The 16-bit Floyd-Warshall form of the part pipeline: entries are min(d, 15872) in
packed 16-bit arithmetic; the result must equal the int32 form and the oracle bit for
bit.  SSSP form (FAP_MODE_FLAG_SSSP, symmetric graphs): pairs in different components
stay unreachable and a part whose distances reach the cap is redone in int32.
Boundary-graph form (unit weights, FAP_MODE_FLAG_FW: otherwise undirected unit-weight
graphs take the bit-parallel BFS): in-part distances never reach the cap.  Road-graph
distances below 1024 are IEEE-half subnormal bit patterns, so every test here also
checks that the half-precision minimum orders them as integers."""
import sys

import numpy as np
import pytest

import fastapsp_amd as F
import oracle as O
from fastapsp_amd import _lib as L

pytestmark = pytest.mark.gpu

CAP = 0x3E00


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _undirected(n, edges):
    """CSR of an undirected graph from (u, v, w) edges (both arcs, same weight)."""
    out = [[] for _ in range(n)]
    for u, v, w in edges:
        out[u].append((v, w))
        out[v].append((u, w))
    ro = np.zeros(n + 1, np.int32)
    ro[1:] = np.cumsum([len(x) for x in out])
    col = np.array([v for x in out for v, _ in x], np.int32)
    wt = np.array([w for x in out for _, w in x], np.float32)
    return ro, col, wt


def _solve_check(g, mode):
    g.setMode(mode)
    g.solve()
    info = L.last_stage_times()
    got = g.getAllDistanceI32()
    lay, csr = g.layout(), g.csr()
    want = O.solve_parts(1, g.K, csr, lay)
    assert np.array_equal(got, want), "GPU solve() differs from the oracle"
    w = None if np.all(csr["weight"] == 1) else csr["weight"]
    exact = O.exact_rows(csr["row_offset"], csr["col"], w, lay["st2ed"])
    assert np.array_equal(got, exact), "GPU solve() differs from exact APSP"
    return got, info


SSSP = F.FAP_MODE_FLAG_SSSP
FW = F.FAP_MODE_FLAG_FW


@pytest.mark.parametrize("flag", [FW, SSSP])
def test_fw16_taken_on_road_graph_and_equals_int32(flag):
    say(f"fw16: road grid 48x40 K=6, 16-bit vs int32, flag={flag:#x}")
    g = F.FapGraph.road_grid(48, 40, 0.1, 0.03, 5, 6)
    g.preCondition()
    a, info = _solve_check(g, F.FAP_MODE_BATCHED | flag)
    assert info["parts_fw16"] == g.K and info["parts_redone"] == 0
    b, info = _solve_check(g, F.FAP_MODE_BATCHED | F.FAP_MODE_FLAG_INT32_FW | flag)
    assert info["parts_fw16"] == 0
    assert np.array_equal(a, b)


def _weighted_path(last_weight, n=512, w=31):
    edges = [(i, i + 1, w) for i in range(n - 2)] + [(n - 2, n - 1, last_weight)]
    return _undirected(n, edges)


def test_fw16_exact_just_below_cap():
    # One part holding a weighted path whose longest distance is CAP - 1.
    say("fw16: path, max distance CAP-1, K=1")
    last = CAP - 1 - 31 * 510
    ro, col, wt = _weighted_path(last)
    g = F.FapGraph.from_csr(ro, col, wt, False, 1)
    g.preCondition()
    got, info = _solve_check(g, F.FAP_MODE_BATCHED | SSSP)
    assert info["parts_fw16"] == 1 and info["parts_redone"] == 0
    assert got.max() == CAP - 1
    again, info = _solve_check(g, F.FAP_MODE_BATCHED)  # weighted: int32 in-part FW
    assert info["parts_fw16"] == 0 and np.array_equal(got, again)


def test_fw16_redo_at_cap():
    say("fw16: path, max distance CAP, K=1 -> int32 redo")
    last = CAP - 31 * 510
    ro, col, wt = _weighted_path(last)
    g = F.FapGraph.from_csr(ro, col, wt, False, 1)
    g.preCondition()
    got, info = _solve_check(g, F.FAP_MODE_BATCHED | SSSP)
    assert info["parts_fw16"] == 1 and info["parts_redone"] == 1
    assert got.max() == CAP


@pytest.mark.parametrize("mode", [F.FAP_MODE_BATCHED, F.FAP_MODE_OVERLAP, F.FAP_MODE_PLAIN])
def test_fw16_redo_some_parts(mode):
    # Path of 1200 vertices, weight 10 on the first half and 60 on the second, K=4:
    # a part is redone exactly when its largest in-part distance reaches the cap, and
    # every form gives the same exact result.
    say(f"fw16: mixed-weight path n=1200 K=4 mode={mode}")
    n = 1200
    edges = [(i, i + 1, 10 if i < n // 2 else 60) for i in range(n - 1)]
    ro, col, wt = _undirected(n, edges)
    g = F.FapGraph.from_csr(ro, col, wt, False, 4)
    g.preCondition()
    got, info = _solve_check(g, mode | SSSP)
    lay = g.layout()
    redo = 0
    for p in range(1, g.K + 1):
        off, num = lay["offset"][p], lay["num"][p]
        block = got[off:off + num][:, lay["st2ed"][off:off + num]]
        redo += int(block.max() >= CAP)
    assert info["parts_fw16"] == g.K
    assert info["parts_redone"] == redo
    assert 1 <= redo < g.K


@pytest.mark.parametrize("flag", [FW, SSSP])
def test_fw16_unreachable_pairs_are_not_redone(flag):
    # Two components inside one part: the capped entries between them are
    # unreachable pairs, not overflow.
    say(f"fw16: two components, K=1, flag={flag:#x}")
    n = 400
    edges = [(i, i + 1, 1) for i in range(199)] + [(i, i + 1, 1) for i in range(200, 399)]
    ro, col, _ = _undirected(n, edges)
    g = F.FapGraph.from_csr(ro, col, None, False, 1)
    g.preCondition()
    got, info = _solve_check(g, F.FAP_MODE_BATCHED | flag)
    assert info["parts_fw16"] == 1 and info["parts_redone"] == 0
    assert (got == F.FAP_INF).sum() == 2 * 200 * 200


def test_fw16_not_taken_on_directed_graph():
    say("fw16: directed knn stays int32")
    g = F.FapGraph.knn(1500, 3, 1000, 4, 4)
    g.preCondition()
    _, info = _solve_check(g, F.FAP_MODE_BATCHED)
    assert info["parts_fw16"] == 0

"""The drop-in check: a C++ program shaped like the reference's single-node driver
(tests/drivers/single_node.cpp, after app/singleNodeExample.cpp:57-93), compiled against
include/fap/fap.h and linked with libfap_hip.so, reads .mtx files through the
(file, directed, weighted, K, partitioner, version, endTimeIo) constructor from the
reference's graph/<kind>/<name>.mtx layout, solves (split mode: solveSubGraph(1..3, true);
otherwise solve(true)) and verifies itself with check_result.  Its dumped block (distances
and predecessors) must equal the oracle's restatement of the reference CPU path bit for
bit, and its point-to-point answers must be shortest paths."""
import os
import subprocess
import sys

import numpy as np
import pytest

import fastapsp_amd as F
import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "drivers", "_build", "single_node")
pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def write_mtx(path, n, arcs, weights=None):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, "w") as f:
        f.write("%%MatrixMarket matrix coordinate integer general\n")
        f.write(f"{n} {n} {len(arcs)}\n")
        for i, (a, b) in enumerate(arcs):
            if weights is None:
                f.write(f"{a + 1} {b + 1}\n")
            else:
                f.write(f"{a + 1} {b + 1} {int(weights[i])}\n")


def csr_arcs(csr, undirected):
    ro, col = csr["row_offset"], csr["col"]
    tails = np.repeat(np.arange(len(ro) - 1), np.diff(ro))
    keep = tails < col if undirected else np.ones(len(col), bool)
    return list(zip(tails[keep].tolist(), col[keep].tolist())), csr["weight"][keep]


def read_dump(path):
    raw = np.fromfile(path, dtype=np.int32)
    rows, n, sid = raw[:3]
    src = raw[3:3 + rows]
    off = 3 + rows
    dist = raw[off:off + rows * n].view(np.float32).reshape(rows, n)
    path_ = raw[off + rows * n:off + 2 * rows * n].reshape(rows, n)
    return int(sid), src, dist, path_


def run_driver(tmp_path, args):
    assert os.path.exists(DRIVER), "tests/drivers/_build/single_node not built (make -C tests/drivers)"
    say("driver", " ".join(args))
    r = subprocess.run([DRIVER] + args, cwd=tmp_path, capture_output=True, text=True, timeout=600)
    say(r.stdout[-2000:])
    assert r.returncode == 0, r.stderr
    return r.stdout


def check_against_oracle(tmp_path, name, directed, weighted, K, want_split):
    out = run_driver(tmp_path, ["-f", name, "-k", str(K), "-direct", str(directed).lower(),
                                "-weight", str(weighted).lower(), "-partitioner", "metis",
                                "-version", "true", "-dump", "out.bin", "-p2p"])
    sid, src, dist, path_ = read_dump(tmp_path / "out.bin")
    assert f"the {sid} subGraph is right" in out
    assert ("subgraphs in split mode" in out) == want_split and (sid == 3) == want_split
    # The same file through the Python mirror gives the same (deterministic) layout.
    cwd = os.getcwd()
    os.chdir(tmp_path)
    try:
        g = F.FapGraph(name, directed, weighted, K)
    finally:
        os.chdir(cwd)
    g.preCondition()
    lay, csr = g.layout(), g.csr()
    if want_split:
        off, num = int(lay["offset"][3]), int(lay["num"][3])
        assert np.array_equal(src, lay["st2ed"][off:off + num])
        wd, wp = O.solve_parts_path(3, 3, csr, lay, row_base=off, rows=num)
    else:
        assert np.array_equal(src, lay["st2ed"])
        wd, wp = O.solve_parts_path(1, g.K, csr, lay)
    got = O.to_i32(dist)
    assert np.array_equal(got, wd), "distances differ from the oracle"
    assert np.array_equal(path_, wp), "predecessors differ from the oracle"
    # point to point: a shortest path of the exact length
    line = [x for x in out.splitlines() if x.startswith("p2p ")][0].split()
    u, v, d, cnt = int(line[1]), int(line[2]), float(line[3]), int(line[4])
    route = [int(x) for x in line[5:]]
    wt = None if not weighted else csr["weight"]
    exact = O.exact_rows(csr["row_offset"], csr["col"], wt, [u])[0]
    if exact[v] >= F.FAP_INF:
        assert cnt == 0 and d >= 1e8
    else:
        assert d == exact[v] and cnt == len(route) and route[0] == u and route[-1] == v
        ro, col = csr["row_offset"], csr["col"]
        w = np.ones(len(col), np.int64) if wt is None else wt.astype(np.int64)
        total = 0
        for a, b in zip(route, route[1:]):
            e = [k for k in range(ro[a], ro[a + 1]) if col[k] == b]
            assert e, f"{a}->{b} is not an arc"
            total += min(w[k] for k in e)
        assert total == exact[v]
    mem = [x for x in out.splitlines() if x.startswith("memcons ")][0].split()
    assert all(float(x) > 0 for x in mem[1:])
    assert "p2p_noargs_throws 1" in out


def test_driver_road_split_mode(tmp_path):
    say("driver: road-like 22000, unweighted undirected, K=16 (split mode)")
    src = F.FapGraph.road_like(22000, 0.52, 1e-4, 5, 16)
    arcs, _ = csr_arcs(src.csr(), undirected=True)
    write_mtx(tmp_path / "graph" / "unweight-undirected" / "road22k.mtx", src.num_vertexs, arcs)
    check_against_oracle(tmp_path, "road22k", False, False, 16, want_split=True)


def test_driver_knn_weighted_directed_split_mode(tmp_path):
    say("driver: kNN 21000 x 3, weighted directed, K=12 (split mode)")
    src = F.FapGraph.knn(21000, 3, 1000, 6, 12)
    arcs, w = csr_arcs(src.csr(), undirected=False)
    write_mtx(tmp_path / "graph" / "weight-directed" / "knn21k.mtx", src.num_vertexs, arcs, w)
    check_against_oracle(tmp_path, "knn21k", True, True, 12, want_split=True)


def test_driver_small_non_split(tmp_path):
    say("driver: road grid 40x40, K=6 (non-split: solve(true))")
    src = F.FapGraph.road_grid(40, 40, 0.1, 0.02, 8, 6)
    arcs, _ = csr_arcs(src.csr(), undirected=True)
    write_mtx(tmp_path / "graph" / "unweight-undirected" / "grid.mtx", src.num_vertexs, arcs)
    check_against_oracle(tmp_path, "grid", False, False, 6, want_split=False)

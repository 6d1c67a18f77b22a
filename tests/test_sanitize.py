"""Sanitizer runs of the host layer on the CPU (tests/sanitize/Makefile): the host
C++ (fap.cpp, server.cpp, partition.cpp, layout.cpp, graph_io.cpp, generators.cpp,
verify.cpp, capi_graph.cpp) built with AddressSanitizer + UndefinedBehaviorSanitizer and
with ThreadSanitizer, the device entry points replaced by a host test double
(tests/sanitize/host_backend.cpp: host memory, host events, BFS / Dijkstra rows).
* the single-node driver (tests/drivers/single_node.cpp, shaped like the reference's
  app/singleNodeExample.cpp:57-93) reads .mtx files, partitions, lays out, solves in
  split and non-split mode and passes check_result, under ASan + UBSan;
* server_load drives the request server from four client threads (full APSP and
  sub-graph requests, the pipeline lane, admission under a small budget, a request
  larger than the budget) and checks every result against a host Dijkstra, under
  ThreadSanitizer and under ASan + UBSan; it also checks that no device allocation
  leaks."""
import os
import shutil
import subprocess

import pytest

from test_gpu_driver import csr_arcs, write_mtx

import fastapsp_amd as F

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BUILD = os.path.join(ROOT, "tests", "sanitize", "_build")
REPORTS = ("ERROR: AddressSanitizer", "runtime error:", "WARNING: ThreadSanitizer",
           "ERROR: LeakSanitizer")


@pytest.fixture(scope="module")
def built():
    if not shutil.which("g++") or not os.path.exists("/opt/rocm/llvm/bin/clang++"):
        pytest.skip("g++ or clang++ missing")
    r = subprocess.run(["make", "-C", os.path.join(ROOT, "tests", "sanitize"), "-j4"],
                       capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    return BUILD


def run(exe, args, cwd, timeout=600):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1", TSAN_OPTIONS="halt_on_error=1")
    r = subprocess.run([exe] + args, cwd=cwd, capture_output=True, text=True, timeout=timeout,
                       env=env)
    out = r.stdout + r.stderr
    for rep in REPORTS:
        assert rep not in out, out[-4000:]
    assert r.returncode == 0, out[-4000:]
    return r.stdout


@pytest.mark.parametrize("kind", ["road_split", "knn_weighted_split", "grid_non_split"])
def test_single_node_driver_asan(built, tmp_path, kind):
    if kind == "road_split":
        src = F.FapGraph.road_like(21000, 0.52, 1e-4, 5, 16)
        arcs, _ = csr_arcs(src.csr(), undirected=True)
        write_mtx(tmp_path / "graph" / "unweight-undirected" / "g.mtx", src.num_vertexs, arcs)
        args, split = ["-f", "g", "-k", "16", "-direct", "false", "-weight", "false"], True
    elif kind == "knn_weighted_split":
        src = F.FapGraph.knn(21000, 3, 1000, 6, 12)
        arcs, w = csr_arcs(src.csr(), undirected=False)
        write_mtx(tmp_path / "graph" / "weight-directed" / "g.mtx", src.num_vertexs, arcs, w)
        args, split = ["-f", "g", "-k", "12", "-direct", "true", "-weight", "true"], True
    else:
        src = F.FapGraph.road_grid(40, 40, 0.1, 0.02, 8, 6)
        arcs, _ = csr_arcs(src.csr(), undirected=True)
        write_mtx(tmp_path / "graph" / "unweight-undirected" / "g.mtx", src.num_vertexs, arcs)
        args, split = ["-f", "g", "-k", "6", "-direct", "false", "-weight", "false"], False
    out = run(os.path.join(built, "single_node_asan"), args + ["-p2p"], tmp_path)
    assert ("subgraphs in split mode" in out) == split
    assert "subGraph is right" in out and "p2p_noargs_throws 1" in out


@pytest.mark.parametrize("exe", ["server_load_tsan", "server_load_asan"])
def test_server_load(built, tmp_path, exe):
    out = run(os.path.join(built, exe), [], tmp_path)
    assert "server_load ok" in out

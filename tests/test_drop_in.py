"""The C++ drop-in compiles: a caller shaped like the reference's single-node driver
(tests/drivers/single_node.cpp) builds against include/fap/fap.h alone and links with
libfap_hip.so (every fapGraph member it calls -- check_result, getMemCons*,
getDistanceP2P / getPathP2P among them -- is defined there), and a missing input file
surfaces as FapError(FAP_ERR_IO), exit status 2, before any device call."""
import os
import subprocess

from fastapsp_amd import _lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_single_node_driver_compiles_links_and_reports_io_errors(tmp_path):
    exe = tmp_path / "single_node"
    lib = os.path.join(ROOT, "fastapsp_amd", "lib")
    subprocess.run(["g++", "-O1", "-std=c++17", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                    "-o", str(exe), os.path.join(ROOT, "tests", "drivers", "single_node.cpp"),
                    "-L", lib, "-lfap_hip", f"-Wl,-rpath,{lib}"], check=True)
    r = subprocess.run([str(exe), "-f", "missing", "-k", "4", "-direct", "false", "-weight",
                        "false"], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert f"FapError {L.FAP_ERR_IO}" in r.stderr and "graph/unweight-undirected/missing.mtx" in r.stderr


def test_printResalts_and_members_declared():
    hdr = open(os.path.join(ROOT, "include", "fap", "fap.h")).read()
    for member in ["bool check_result(float *dist, int *path, int *source, int source_num",
                   "double getMemConsSSSP();", "double getMemConsFloyd();",
                   "double getMemConsMinPlus();", "float getDistanceP2P();",
                   "std::vector<int32_t> getPathP2P();", "void printResalts();"]:
        assert member in hdr, member

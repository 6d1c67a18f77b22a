"""The C-ABI library loads and exports every symbol include/fap_hip.h declares
(no device calls: this runs without a GPU)."""
import ctypes as C
import os
import re

import fastapsp_amd as F
from fastapsp_amd import _lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "fap_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fap_(?:hip|graph|mtx|layout)_\w+)\s*\(", text)))


def test_header_declares_the_abi():
    names = declared_functions()
    assert "fap_hip_solve_parts" in names and "fap_hip_minplus" in names
    assert len(names) >= 40


def test_library_exports_every_declared_symbol():
    lib = F.load()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert missing == []


def test_bindings_cover_every_declared_symbol():
    assert sorted(L.SIGNATURES) == declared_functions()


def test_struct_layouts_match_header():
    # fap_graph_dev: int32 n, int64 m, 6 pointers; fap_part_desc: 4 x int32 + int64.
    assert C.sizeof(L.FapGraphDev) == 8 + 8 + 6 * 8
    assert C.sizeof(L.FapPartDesc) == 24
    assert C.sizeof(L.FapStageTimes) == 24


def test_status_codes_without_device():
    lib = F.load()
    # Kernel entry points refuse to run before fap_hip_init (no device touched).
    if not L._initialised:
        assert lib.fap_hip_minplus(None, 1, None, 1, None, 1, 1, 1, 1, None) == L.FAP_ERR_STATE
        assert "fap_hip_init" in L.last_error()

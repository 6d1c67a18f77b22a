"""The C-ABI library loads and exports every symbol include/fap_hip.h declares
(no device calls: this runs without a GPU)."""
import ctypes as C
import os
import re
import subprocess

import fastapsp_amd as F
from fastapsp_amd import _lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "fap_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fap_(?:hip|graph|mtx|layout|server)_\w+)\s*\(", text)))


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


def test_struct_layouts_match_header(tmp_path):
    """sizeof and every field offset of the ctypes structs equal what a C compiler
    reads from include/fap_hip.h."""
    structs = {"fap_graph_dev": L.FapGraphDev, "fap_part_desc": L.FapPartDesc,
               "fap_stage_times": L.FapStageTimes, "fap_batch_item": L.FapBatchItem,
               "fap_server_stats": L.FapServerStats}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "fap_hip.h"',
             "int main(void) {"]
    for cname, cls in structs.items():
        lines.append(f'  printf("{cname} sizeof %zu\\n", sizeof({cname}));')
        for field, _ in cls._fields_:
            lines.append(f'  printf("{cname} {field} %zu\\n", offsetof({cname}, {field}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o",
                    str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    got = {tuple(l.split()[:2]): int(l.split()[2]) for l in out.splitlines()}
    for cname, cls in structs.items():
        assert got[(cname, "sizeof")] == C.sizeof(cls), cname
        for field, _ in cls._fields_:
            assert got[(cname, field)] == getattr(cls, field).offset, (cname, field)


def test_status_codes_without_device():
    lib = F.load()
    # Kernel entry points refuse to run before fap_hip_init (no device touched).
    if not L._initialised:
        assert lib.fap_hip_minplus(None, 1, None, 1, None, 1, 1, 1, 1, None) == L.FAP_ERR_STATE
        assert "fap_hip_init" in L.last_error()


def test_merge_shape_setter_validates_and_round_trips():
    """fap_hip_set_merge_shape is host state only: bad shapes are refused with
    FAP_ERR_INVALID and leave the shape unchanged; good ones read back."""
    lib = L.load()
    get = lambda: tuple(v.value for v in _get_shape(lib))  # noqa: E731
    before = get()
    for bad in [(0, 1, 0), (12, 1, 0), (2048, 1, 0), (128, 2, 0), (128, 1, 2), (128, 1, 33)]:
        assert lib.fap_hip_set_merge_shape(*bad) == L.FAP_ERR_INVALID, bad
        assert get() == before
    try:
        assert lib.fap_hip_set_merge_shape(32, 0, 8) == 0 and get() == (32, 0, 8)
    finally:
        assert lib.fap_hip_set_merge_shape(*before) == 0


def _get_shape(lib):
    vals = [C.c_int32(), C.c_int32(), C.c_int32()]
    assert lib.fap_hip_get_merge_shape(*[C.byref(v) for v in vals]) == 0
    return vals

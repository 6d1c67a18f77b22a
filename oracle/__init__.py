"""ORACLE -- test infrastructure only (tests/, __graft_entry__.smoke(), bench.py's
cpu_baseline).  ctypes access to liboracle.so, the CPU restatement of the
reference's CPU Fast-APSP path, and to _ref/libfapref.so, the reference's own CPU
kernels compiled from its sources (present only where the reference tree was)."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_LIB = os.path.join(HERE, "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libfapref.so")
MAXVALUE = np.float32(1e8)
INF = 1073741823

_orc = None
_ref = None


def _p(a, ctype):
    return None if a is None else a.ctypes.data_as(C.POINTER(ctype))


def lib():
    global _orc
    if _orc is None:
        if not os.path.exists(ORACLE_LIB):
            raise RuntimeError(f"oracle library missing: {ORACLE_LIB} (make -C oracle)")
        _orc = C.CDLL(ORACLE_LIB)
    return _orc


def ref():
    """The compiled reference kernels, or None when they were not built."""
    global _ref
    if _ref is None and os.path.exists(REF_LIB):
        _ref = C.CDLL(REF_LIB)
    return _ref


def i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def to_i32(d: np.ndarray) -> np.ndarray:
    d = f32(d)
    out = np.empty(d.shape, dtype=np.int32)
    lib().orc_to_i32(_p(d, C.c_float), _p(out, C.c_int32), C.c_int64(d.size))
    return out


def dijkstra_path(u, row_offset, col, w):
    ro, cv, ww = i32(row_offset), i32(col), f32(w)
    n = len(ro) - 1
    adj = i32(np.diff(ro))
    dist = np.empty(n, np.float32)
    path = np.empty(n, np.int32)
    lib().orc_dijkstra_path(C.c_int(u), C.c_int(n), _p(adj, C.c_int32), _p(ro, C.c_int32),
                            _p(cv, C.c_int32), _p(ww, C.c_float), _p(dist, C.c_float),
                            _p(path, C.c_int32))
    return dist, path


def floyd_path(mat, path):
    mat, path = f32(mat).copy(), i32(path).copy()
    lib().orc_floyd_path(C.c_int(mat.shape[0]), _p(mat, C.c_float), _p(path, C.c_int32))
    return mat, path


def minplus_advanced(a, b, bpath):
    a, b, bpath = f32(a), f32(b), i32(bpath)
    M, K = a.shape
    N = b.shape[1]
    c = np.empty((M, N), np.float32)
    cp = np.empty((M, N), np.int32)
    lib().orc_minplus_advanced(_p(a, C.c_float), _p(b, C.c_float), _p(bpath, C.c_int32),
                               _p(c, C.c_float), _p(cp, C.c_int32), M, N, K)
    return c, cp


def run_part(pid, csr, lay):
    """fapGraph::run CPU branch for part pid: (dist float s x n, path s x n)."""
    ro, cv = i32(csr["row_offset"]), i32(csr["col"])
    w = f32(csr["weight"])
    n = len(ro) - 1
    adj = i32(np.diff(ro))
    s = int(lay["num"][pid])
    dist = np.empty((s, n), np.float32)
    path = np.empty((s, n), np.int32)
    args = [i32(lay[k]) for k in ("graph_id", "st2ed", "ed2st", "num", "offset", "bdy")]
    lib().orc_run_part(C.c_int(pid), C.c_int(n), _p(adj, C.c_int32), _p(ro, C.c_int32),
                       _p(cv, C.c_int32), _p(w, C.c_float), *[_p(a, C.c_int32) for a in args],
                       _p(dist, C.c_float), _p(path, C.c_int32))
    return dist, path


def solve_parts(first, last, csr, lay, row_base=0, rows=None):
    """Parts first..last (1-based) in parallel into a float block starting at sorted row
    row_base; returns the int32 block (MAXVALUE -> INF)."""
    ro, cv = i32(csr["row_offset"]), i32(csr["col"])
    w = f32(csr["weight"])
    n = len(ro) - 1
    adj = i32(np.diff(ro))
    if rows is None:
        rows = int(sum(lay["num"][first:last + 1]))
    dist = np.empty((rows, n), np.float32)
    args = [i32(lay[k]) for k in ("graph_id", "st2ed", "ed2st", "num", "offset", "bdy")]
    lib().orc_solve_parts_parallel(C.c_int(first), C.c_int(last), C.c_int(n), _p(adj, C.c_int32),
                                   _p(ro, C.c_int32), _p(cv, C.c_int32), _p(w, C.c_float),
                                   *[_p(a, C.c_int32) for a in args], _p(dist, C.c_float),
                                   C.c_int64(row_base))
    return to_i32(dist)


def solve_parts_path(first, last, csr, lay, row_base=0, rows=None):
    """As solve_parts, also returning the predecessor block: (int32 dist, int32 path)."""
    ro, cv = i32(csr["row_offset"]), i32(csr["col"])
    w = f32(csr["weight"])
    n = len(ro) - 1
    adj = i32(np.diff(ro))
    if rows is None:
        rows = int(sum(lay["num"][first:last + 1]))
    dist = np.empty((rows, n), np.float32)
    path = np.empty((rows, n), np.int32)
    args = [i32(lay[k]) for k in ("graph_id", "st2ed", "ed2st", "num", "offset", "bdy")]
    lib().orc_solve_parts_path_parallel(C.c_int(first), C.c_int(last), C.c_int(n),
                                        _p(adj, C.c_int32), _p(ro, C.c_int32), _p(cv, C.c_int32),
                                        _p(w, C.c_float), *[_p(a, C.c_int32) for a in args],
                                        _p(dist, C.c_float), _p(path, C.c_int32),
                                        C.c_int64(row_base))
    return to_i32(dist), path


def exact_rows(row_offset, col, weight, sources):
    """Independent exact distances (BFS if weight is None, else integer Dijkstra)."""
    ro, cv, src = i32(row_offset), i32(col), i32(sources)
    n = len(ro) - 1
    out = np.empty((len(src), n), np.int32)
    wp = None if weight is None else _p(i32(weight), C.c_int32)
    wa = None if weight is None else i32(weight)
    if wa is not None:
        wp = _p(wa, C.c_int32)
    lib().orc_exact_rows(C.c_int(n), _p(ro, C.c_int32), _p(cv, C.c_int32), wp, _p(src, C.c_int32),
                         C.c_int(len(src)), _p(out, C.c_int32))
    return out


def exact_rows_f64(row_offset, col, weight, sources):
    """FP64 Dijkstra rows (float32 weights widened to double; None = unit weights),
    +inf for unreachable: the oracle of the float-weight path."""
    ro, cv, src = i32(row_offset), i32(col), i32(sources)
    n = len(ro) - 1
    out = np.empty((len(src), n), np.float64)
    wa = None if weight is None else np.ascontiguousarray(weight, np.float32)
    wp = None if wa is None else _p(wa, C.c_float)
    lib().orc_exact_rows_f64(C.c_int(n), _p(ro, C.c_int32), _p(cv, C.c_int32), wp,
                             _p(src, C.c_int32), C.c_int(len(src)), _p(out, C.c_double))
    return out


def floyd_i32(mat):
    m = i32(mat).copy()
    s = m.shape[0]
    lib().orc_floyd_i32(C.c_int(s), _p(m, C.c_int32), C.c_int64(m.shape[1]))
    return m


def minplus_i32(a, b):
    a, b = i32(a), i32(b)
    M, K = a.shape
    N = b.shape[1]
    c = np.empty((M, N), np.int32)
    lib().orc_minplus_i32(_p(a, C.c_int32), C.c_int64(max(K, 1)), _p(b, C.c_int32),
                          C.c_int64(N), _p(c, C.c_int32), C.c_int64(N), M, N, K)
    return c


def compare_exact_rows(row_offset, col, weight, sources, got):
    """Rows of `got` (one per source, any row stride) against exact BFS / integer
    Dijkstra rows, computed and compared per thread without an n x n buffer.
    Returns (number of differing rows, first differing row or -1)."""
    ro, cv, src = i32(row_offset), i32(col), i32(sources)
    n = len(ro) - 1
    if got.dtype != np.int32 or got.shape[0] != len(src) or got.shape[1] < n or \
            got.strides[1] != 4:
        raise ValueError("got: int32 rows of at least n entries, one per source")
    wa = None if weight is None else i32(weight)
    first = C.c_int64()
    f = lib().orc_compare_exact_rows
    f.restype = C.c_int64
    bad = f(C.c_int(n), _p(ro, C.c_int32), _p(cv, C.c_int32), _p(wa, C.c_int32),
            _p(src, C.c_int32), C.c_int64(len(src)), got.ctypes.data_as(C.POINTER(C.c_int32)),
            C.c_int64(got.strides[0] // 4), C.byref(first))
    return int(bad), int(first.value)


def compare_parts(first, last, csr, lay, got, row_base):
    """Parts first..last (1-based) by the reference CPU path's restatement (orc_run_part,
    OpenMP across parts) against `got`, whose row 0 is sorted row row_base.  Returns
    (number of differing entries, first differing got row or -1)."""
    ro, cv = i32(csr["row_offset"]), i32(csr["col"])
    w = f32(csr["weight"])
    n = len(ro) - 1
    adj = i32(np.diff(ro))
    lo = int(lay["offset"][first]) - row_base
    hi = int(lay["offset"][last]) + int(lay["num"][last]) - row_base
    if got.dtype != np.int32 or lo < 0 or got.shape[0] < hi or got.shape[1] < n or \
            got.strides[1] != 4:
        raise ValueError("got does not hold the rows of the parts")
    args = [i32(lay[k]) for k in ("graph_id", "st2ed", "ed2st", "num", "offset", "bdy")]
    fb = C.c_int64()
    f = lib().orc_compare_parts
    f.restype = C.c_int64
    bad = f(C.c_int(first), C.c_int(last), C.c_int(n), _p(adj, C.c_int32), _p(ro, C.c_int32),
            _p(cv, C.c_int32), _p(w, C.c_float), *[_p(a, C.c_int32) for a in args],
            got.ctypes.data_as(C.POINTER(C.c_int32)), C.c_int64(got.strides[0] // 4),
            C.c_int64(row_base), C.byref(fb))
    return int(bad), int(fb.value)


def submat_build(dist_rows, path_rows, csr, lay, pid):
    """MysubMatBuild_path for part pid from its boundary rows (float bdy x n with
    MAXVALUE, int paths): the s x s sub-matrix and its paths, pre-filled MAXVALUE / -1."""
    ro, cv = i32(csr["row_offset"]), i32(csr["col"])
    w = f32(csr["weight"])
    n = len(ro) - 1
    adj = i32(np.diff(ro))
    s, bdy, start = int(lay["num"][pid]), int(lay["bdy"][pid]), int(lay["offset"][pid])
    sub = np.full((s, s), MAXVALUE, np.float32)
    sp = np.full((s, s), -1, np.int32)
    dr = f32(dist_rows).reshape(-1) if bdy else np.zeros(1, np.float32)
    pr = i32(path_rows).reshape(-1) if bdy else np.zeros(1, np.int32)
    lib().orc_submat_build(_p(sub, C.c_float), _p(sp, C.c_int32), _p(dr, C.c_float),
                           _p(pr, C.c_int32), _p(i32(lay["graph_id"]), C.c_int32), start, s, bdy,
                           n, _p(i32(lay["st2ed"]), C.c_int32), _p(i32(lay["ed2st"]), C.c_int32),
                           _p(adj, C.c_int32), _p(ro, C.c_int32), _p(cv, C.c_int32),
                           _p(w, C.c_float))
    return sub, sp


def submat_decode(sub, sub_path, st2ed, start, dist_block, path_block):
    """MysubMatDecode_path: rows x s of (sub, sub_path) scattered into the columns
    st2ed[start + j] of the rows x n blocks (modified in place, returned)."""
    sub, sub_path = f32(sub), i32(sub_path)
    rows, s = sub.shape
    n = dist_block.shape[1]
    assert dist_block.dtype == np.float32 and path_block.dtype == np.int32
    assert dist_block.flags.c_contiguous and path_block.flags.c_contiguous
    lib().orc_submat_decode(_p(sub, C.c_float), _p(sub_path, C.c_int32), _p(dist_block, C.c_float),
                            _p(path_block, C.c_int32), start, rows, s, n,
                            _p(i32(st2ed), C.c_int32))
    return dist_block, path_block

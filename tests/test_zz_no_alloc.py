"""The allocation-free Floyd-Warshall entry point (fap_hip_floyd_noalloc) through the
C-ABI against fap_hip_floyd, the entry point it is a variant of, bit for bit: two calls
in a row (the second smaller, so the scratch holds the first call's larger matrix), the
largest size the scratch takes, and the refusal of one vertex more.  The library's
device-memory accounting must not move during the allocation-free calls."""
import ctypes as C
import sys

import numpy as np
import pytest

import fastapsp_amd as F
from fastapsp_amd import _lib as L
import oracle as O

INF = F.FAP_INF
pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rand_mat(rng, s, hi=1 << 16, p_inf=0.2):
    a = rng.integers(0, hi, size=(s, s), dtype=np.int64).astype(np.int32)
    a[rng.random((s, s)) < p_inf] = INF
    np.fill_diagonal(a, 0)
    return a


def mem_in_use():
    u, h, b = C.c_int64(), C.c_int64(), C.c_int64()
    L.check(L.load().fap_hip_mem_stats(C.byref(u), C.byref(h), C.byref(b)), "mem_stats")
    return u.value, h.value


def floyd_both(m, ld):
    """(fap_hip_floyd result, fap_hip_floyd_noalloc result, accounting moved?)"""
    s = m.shape[0]
    host = np.full((s, ld), -5, np.int32)  # stride padding must stay untouched
    host[:, :s] = m
    ref = F.DeviceBuffer.from_array(host)
    L.check(L.load().fap_hip_floyd(ref.ptr, s, ld, None), "floyd")
    want = ref.download((s, ld))
    dev = F.DeviceBuffer.from_array(host)
    before = mem_in_use()
    L.check(L.load().fap_hip_floyd_noalloc(dev.ptr, s, ld, None), "floyd_noalloc")
    L.check(L.load().fap_hip_stream_sync(None), "sync")
    moved = mem_in_use() != before
    got = dev.download((s, ld))
    return want, got, moved


def test_floyd_noalloc_matches_floyd_two_calls():
    rng = np.random.default_rng(11)
    for s, ld in [(1000, 1003), (300, 300)]:
        say(f"floyd_noalloc s={s} ld={ld}")
        m = rand_mat(rng, s)
        want, got, moved = floyd_both(m, ld)
        assert not moved
        assert np.array_equal(got, want)
        assert np.array_equal(got[:, :s], O.floyd_i32(m))
        assert np.all(got[:, s:] == -5)


def test_floyd_noalloc_largest_and_refused():
    rng = np.random.default_rng(12)
    s = 4096  # FAP_FLOYD_NOALLOC_MAX_S
    say(f"floyd_noalloc s={s} (scratch limit)")
    m = rand_mat(rng, s, p_inf=0.9)
    want, got, moved = floyd_both(m, s)
    assert not moved and np.array_equal(got, want)
    dev = F.DeviceBuffer(4)
    rc = L.load().fap_hip_floyd_noalloc(dev.ptr, s + 1, s + 1, None)
    assert rc == L.FAP_ERR_INVALID
    assert b"scratch limit" in L.load().fap_hip_last_error()
    # a smaller call after the refusal still works
    m = rand_mat(rng, 77)
    want, got, moved = floyd_both(m, 80)
    assert not moved and np.array_equal(got, want)

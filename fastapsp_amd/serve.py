"""Python side of the request server (fap_server_* in include/fap_hip.h).

Graphs are registered once (after preCondition) and stay resident; requests are a full
APSP (`sub_id=0`) or one sub-graph block (`sub_id=p`).  The server's scheduler thread
admits waiting requests by HBM budget and runs each admitted group as one batch on a
free stream (continuous batching); results stay in HBM until released.
`poisson_load` drives a server with Poisson arrivals and reports latency percentiles.
"""
from __future__ import annotations

import ctypes as C
import time

import numpy as np

from . import _lib as L

PENDING, RUNNING, DONE, FAILED = 0, 1, 2, 3


class Server:
    def __init__(self, budget_bytes: float = 200e9, streams: int = 4):
        L.init()
        h = C.c_void_p()
        L.check(L.load().fap_server_create(C.byref(h), int(budget_bytes), int(streams)),
                "server create")
        self._h = h.value
        self._graphs = []  # keep the FapGraph objects alive

    def add_graph(self, g) -> int:
        if not getattr(g, "_preconditioned", False):
            g.preCondition()
        idx = C.c_int32()
        L.check(L.load().fap_server_add_graph(self._h, g._h, C.byref(idx)), "add graph")
        self._graphs.append(g)
        return idx.value

    def submit(self, graph: int, sub_id: int = 0) -> int:
        t = C.c_int64()
        L.check(L.load().fap_server_submit(self._h, int(graph), int(sub_id), C.byref(t)),
                "submit")
        return t.value

    def wait(self, ticket: int, timeout: float = 60.0) -> int:
        st = C.c_int32()
        L.check(L.load().fap_server_wait(self._h, int(ticket), float(timeout), C.byref(st)),
                "wait")
        return st.value

    def result_device(self, ticket: int):
        p, rows, cols, ld, fl = C.c_void_p(), C.c_int64(), C.c_int64(), C.c_int64(), C.c_int32()
        L.check(L.load().fap_server_result(self._h, int(ticket), C.byref(p), C.byref(rows),
                                           C.byref(cols), C.byref(ld), C.byref(fl)), "result")
        return p.value, rows.value, cols.value, ld.value, bool(fl.value)

    def result(self, ticket: int) -> np.ndarray:
        """The result rows copied to the host: int32 (FAP_INF) or float32 (MAXVALUE)."""
        p, rows, cols, ld, fl = self.result_device(ticket)
        out = np.empty((rows, cols), np.int32)
        if rows:
            L.check(L.load().fap_hip_memcpy2d_d2h_i32(L.ptr_i32(out), cols, p, ld, rows, cols,
                                                      None), "result copy")
        return out.view(np.float32) if fl else out

    def release(self, ticket: int) -> None:
        L.check(L.load().fap_server_release(self._h, int(ticket)), "release")

    def stats(self) -> dict:
        s = L.FapServerStats()
        L.check(L.load().fap_server_get_stats(self._h, C.byref(s)), "stats")
        return {k: getattr(s, k) for k, _ in L.FapServerStats._fields_}

    def close(self) -> None:
        if self._h is not None:
            L.load().fap_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def poisson_load(server: Server, requests, rate: float, seed: int = 0,
                 keep: bool = False, timeout: float = 120.0):
    """Submit `requests` (a list of (graph index, sub_id)) at Poisson arrival times of
    the given rate (requests per second) and wait for all.  Unless keep, each result is
    released as soon as it is seen finished (a client that consumes results in place),
    so held results do not fill the budget.  Returns (tickets, states, wall seconds)."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate, len(requests))
    tickets, states = [], {}
    outstanding = []

    def reap(block: bool):
        still = []
        for t in outstanding:
            st = server.wait(t, timeout if block else 0.0)
            if st in (DONE, FAILED):
                states[t] = st
                if not keep:
                    server.release(t)
            else:
                still.append(t)
        outstanding[:] = still

    t0 = time.perf_counter()
    due = t0
    for (g, sub), gap in zip(requests, gaps):
        due += gap
        while True:
            delay = due - time.perf_counter()
            if delay <= 0:
                break
            reap(False)
            time.sleep(min(delay, 0.002))
        t = server.submit(g, sub)
        tickets.append(t)
        outstanding.append(t)
    while outstanding:
        reap(True)
    wall = time.perf_counter() - t0
    return tickets, [states.get(t, PENDING) for t in tickets], wall

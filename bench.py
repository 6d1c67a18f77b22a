"""Fast-APSP benchmark on one MI355X.

Workload (default): road-like undirected unweighted graph with exactly n = 114,599
vertices (percolated grid, p_keep = 0.52, chords at rate 1e-4, largest component,
BFS-trimmed), K = 128 parts, full n x n int32 APSP via solve() in the batched form,
result resident in HBM (n x ld int32, about 52.5 GB).  A step is one whole solve():
the in-part distances (bit-parallel BFS on this unit-weight undirected graph), the
boundary rows (from the boundary graph), and the min-plus merge with fused decode, for
all parts.

Prints ONE JSON line: value = n^2 distance entries per second of wall time.
Also measured: the plain per-part launch form (one step), per-stage device times,
the roofline of the dominant kernel and the CPU path on a bounded sample.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

import fastapsp_amd as F  # noqa: E402
from fastapsp_amd import _lib as L  # noqa: E402

METRIC = "APSP distance entries/s at bit-exact int32; batched vs plain speedup, 1 GPU"
VALU_PEAK_TOPS = 39.3  # 256 CU x 64 lanes x 2.4 GHz, int32 lane-ops
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def stage_times():
    return L.last_stage_times()


def make_graph(cfg):
    if cfg == "road1":
        return F.FapGraph.road_like(114599, 0.52, 1e-4, 1, 128), \
            "road-like n=114599 (percolated grid p=0.52, chords 1e-4, LCC, BFS trim), " \
            "undirected unweighted, K=128, full APSP via solve()"
    if cfg == "road0":
        return F.FapGraph.road_grid(64, 64, 0.10, 0.02, 0, 8), \
            "road grid 64x64 (10% edges dropped, 2% diagonals, LCC), K=8, full APSP"
    raise SystemExit(f"unknown config {cfg}")


def merge_traffic(cfg):
    """HBM bytes per solve of the merge kernels from the committed FETCH_SIZE /
    WRITE_SIZE summary of this workload (calibrated counters), or None."""
    path = os.path.join(ROOT, "profiles", "merge_road114599_k128_pmc_traffic.md")
    if cfg != "road1" or not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("merge_traffic_bytes_per_solve:"):
            return float(line.split(":")[1])
    return None


def cpu_baseline(g, budget_s):
    """The oracle's CPU path (reference structure: OpenMP Dijkstra over boundary sources,
    single-threaded Floyd-Warshall and min-plus) on parts 1..k until budget_s elapse."""
    import oracle as O

    lay, csr = g.layout(), g.csr()
    n = g.num_vertexs
    t0 = time.perf_counter()
    entries, parts = 0, 0
    for p in range(1, g.K + 1):
        O.run_part(p, csr, lay)
        entries += int(lay["num"][p]) * n
        parts += 1
        if time.perf_counter() - t0 >= budget_s:
            break
    dt = time.perf_counter() - t0
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": entries / dt, "unit": "entries/s", "cores": cores, "kind": "port",
            "sample": f"parts 1..{parts} of the same graph ({entries} entries, {dt:.1f} s); "
                      f"Dijkstra stage on {cores} OpenMP threads, Floyd-Warshall and min-plus "
                      f"on 1 thread as in the reference CPU path"}


def probe(gib):
    """HBM probe: a 16-byte fill (write only) and a copy of known size, timed with HIP
    events on the library stream.  The same launches are the reference dispatches for
    the FETCH_SIZE / WRITE_SIZE counter scaling."""
    lib = L.load()
    L.init()
    nbytes = int(gib * (1 << 30))
    count = nbytes // 4
    a = F.DeviceBuffer(nbytes)
    b = F.DeviceBuffer(nbytes)
    e0, e1 = C.c_void_p(), C.c_void_p()
    L.check(lib.fap_hip_event_create(C.byref(e0)), "event")
    L.check(lib.fap_hip_event_create(C.byref(e1)), "event")
    out = {}
    for name, fn, moved in [
        ("fill", lambda: lib.fap_hip_fill_i32(a.ptr, 7, count, None), nbytes),
        ("copy", lambda: lib.fap_hip_copy_i32(b.ptr, a.ptr, count, None), 2 * nbytes)]:
        L.check(fn(), name)
        best = None
        for _ in range(5):
            L.check(lib.fap_hip_event_record(e0, None), "record")
            L.check(fn(), name)
            L.check(lib.fap_hip_event_record(e1, None), "record")
            ms = C.c_float()
            L.check(lib.fap_hip_event_elapsed_ms(e0, e1, C.byref(ms)), "elapsed")
            best = ms.value if best is None else min(best, ms.value)
        out[name] = {"bytes": moved, "ms": best, "GBs": moved / (best * 1e-3) / 1e9}
    print(json.dumps({"probe": out, "gib": gib}), flush=True)


def minplus_bench(N, reps=3):
    """Dense min-plus microbench: C = A (x) B on N x N int32 blocks with values uniform in
    [0, 2^20] (fap_hip_minplus, LDS-tiled kernel), timed with HIP events on the library
    stream; 64 sampled entries checked against numpy."""
    lib = L.load()
    L.init()
    rng = np.random.default_rng(3)
    A = rng.integers(0, 1 << 20, size=(N, N), dtype=np.int32)
    B = rng.integers(0, 1 << 20, size=(N, N), dtype=np.int32)
    dA, dB = F.DeviceBuffer.from_array(A), F.DeviceBuffer.from_array(B)
    dC = F.DeviceBuffer(N * N * 4)
    e0, e1 = C.c_void_p(), C.c_void_p()
    L.check(lib.fap_hip_event_create(C.byref(e0)), "event")
    L.check(lib.fap_hip_event_create(C.byref(e1)), "event")
    run = lambda: L.check(lib.fap_hip_minplus(dA.ptr, N, dB.ptr, N, dC.ptr, N, N, N, N, None),
                          "minplus")
    run()
    times = []
    for _ in range(reps):
        L.check(lib.fap_hip_event_record(e0, None), "record")
        run()
        L.check(lib.fap_hip_event_record(e1, None), "record")
        ms = C.c_float()
        L.check(lib.fap_hip_event_elapsed_ms(e0, e1, C.byref(ms)), "elapsed")
        times.append(ms.value)
    Cres = np.empty((N, N), np.int32)
    L.check(lib.fap_hip_memcpy_d2h(Cres.ctypes.data, dC.ptr, N * N * 4, None), "d2h")
    idx = rng.integers(0, N, size=(64, 2))
    ok = all(int(Cres[i, j]) == int((A[i, :].astype(np.int64) + B[:, j]).min()) for i, j in idx)
    ms = float(np.median(times))
    ops = 2.0 * N ** 3
    inst = 1.5 * N ** 3 / (ms * 1e-3) / 1e12
    return {"N": N, "ms": ms, "ops": ops, "TOPs": ops / (ms * 1e-3) / 1e12,
            "valu_instr_T_per_s": inst, "valu_frac": inst / VALU_PEAK_TOPS,
            "sampled_entries_exact": ok,
            "workload": f"dense {N}x{N}x{N} int32 min-plus, values uniform in [0, 2^20]"}


def knn_subgraph_bench(args):
    """Directed integer-weighted kNN workload: 262,144 uniform points, arcs to the 3
    nearest neighbours (786,432 arcs, weights uniform in [1, 1000]), K = 256;
    solveSubGraph for every part in turn (each call: boundary SSSP, build,
    Floyd-Warshall, min-plus, decode; the s_p x n block stays in HBM)."""
    t0 = time.perf_counter()
    g = F.FapGraph.knn(262144, 3, 1000, 2, 256)
    g.preCondition()
    lay = g.layout()
    n = g.num_vertexs
    num = lay["num"][1:].astype(np.int64)
    bdy = lay["bdy"][1:].astype(np.int64)
    log(f"knn graph n={n} m={g.num_edges} K={g.K} built in {time.perf_counter() - t0:.1f}s; "
        f"boundary per part mean {bdy.mean():.1f} max {bdy.max()}, part size max {num.max()}")
    entries = float(num.sum()) * n
    steps = float(((num - bdy) * bdy).sum()) * n + float((num.astype(np.float64) ** 3).sum())

    def sweep(after_calls):
        # A fresh start: the sweep (when enabled) is prepared inside the timed calls.
        g.setSubGraphSweep(0)
        g.setSubGraphSweep(after_calls)
        sums = {k: 0.0 for k in ("sssp_ms", "build_ms", "floyd_ms", "minplus_ms", "decode_ms",
                                 "total_ms")}
        L.check(L.load().fap_hip_device_sync(), "sync")
        ts = time.perf_counter()
        for p in range(1, g.K + 1):
            g.solveSubGraph(p)
            st = stage_times()
            for k in sums:
                sums[k] += st[k]
        L.check(L.load().fap_hip_device_sync(), "sync")
        return time.perf_counter() - ts, sums, g.subGraphSweepState()

    g.setSubGraphSweep(0)
    g.solveSubGraph(1)  # warm-up (allocations)
    per_wall, per_sums, _ = sweep(0)
    log(f"knn solveSubGraph x {g.K}, per-part form: {per_wall * 1e3:.1f} ms, "
        f"{entries / per_wall:.3e} entries/s; stages (ms, all parts): "
        f"{', '.join(f'{k} {v:.1f}' for k, v in per_sums.items())}")
    wall, sums, state = sweep(3)  # the library default: prepared from the third call on
    log(f"knn solveSubGraph x {g.K}: {wall * 1e3:.1f} ms, {entries / wall:.3e} entries/s "
        f"(sweep {'sparse' if state == 2 else 'dense' if state == 1 else 'off'}; the batched "
        f"pass inside the third call); per-call stages (ms, all parts): "
        f"{', '.join(f'{k} {v:.1f}' for k, v in sums.items())}")
    g.close()
    return {"workload": "kNN n=262144 (3 nearest neighbours, directed, int weights U[1,1000]), "
                        "K=256, solveSubGraph over every part in turn",
            "n": n, "K": 256, "parts": int(len(num)), "entries": entries, "wall_ms": wall * 1e3,
            "entries_per_s": entries / wall, "ms_per_part": wall * 1e3 / len(num),
            "sweep": {0: "off", 1: "dense", 2: "sparse"}[state],
            "result_bytes_written": entries * 4,
            "hbm_write_frac": entries * 4 / wall / (HBM_PEAK_GBS * 1e9),
            "boundary_mean": float(bdy.mean()), "boundary_max": int(bdy.max()),
            "minplus_fw_steps": steps, "stages_ms_all_parts": sums,
            "per_part_form": {"wall_ms": per_wall * 1e3, "entries_per_s": entries / per_wall,
                              "stages_ms_all_parts": per_sums}}


def wroad_subgraph_bench(args):
    """Weighted road-like graph (~62k vertices, integer weights in [1, 1000] equal both
    ways, K = 64): solveSubGraph over every part in turn.  Its blocks come from the boundary
    graph (every call closes every part and the boundary graph); with the sweep (library
    default, from the third call) one full solve() runs and each block is a copy."""
    from fastapsp_amd.autotune import road_graph, symmetric_float_weights
    t0 = time.perf_counter()
    g0 = road_graph(62000, 77, 64)
    c = g0.csr()
    g0.close()
    w = symmetric_float_weights(c["row_offset"], c["col"], 77)
    w = (np.floor((w - 0.5) / 9.5 * 1000) + 1).astype(np.float32)
    g = F.FapGraph.from_csr(c["row_offset"], c["col"], w, False, 64)
    g.preCondition()
    n = g.num_vertexs
    num = g.layout()["num"][1:].astype(np.int64)
    entries = float(num.sum()) * n
    log(f"weighted road n={n} m={g.num_edges} K={g.K} built in {time.perf_counter() - t0:.1f}s")

    def sweep(after_calls):
        g.setSubGraphSweep(0)
        g.setSubGraphSweep(after_calls)
        L.check(L.load().fap_hip_device_sync(), "sync")
        ts = time.perf_counter()
        for p in range(1, g.K + 1):
            g.solveSubGraph(p)
        L.check(L.load().fap_hip_device_sync(), "sync")
        return time.perf_counter() - ts, g.subGraphSweepState()

    g.setSubGraphSweep(0)
    g.solveSubGraph(1)  # warm-up (allocations)
    per_wall, _ = sweep(0)
    wall, state = sweep(3)
    log(f"weighted road solveSubGraph x {g.K}: per-part form {per_wall * 1e3:.1f} ms, sweep "
        f"{wall * 1e3:.1f} ms ({entries / wall:.3e} entries/s, state {state})")
    g.close()
    return {"workload": f"road-like n={n}, integer weights U[1,1000] both ways, K=64, "
                        "solveSubGraph over every part in turn",
            "n": n, "K": 64, "entries": entries, "wall_ms": wall * 1e3,
            "entries_per_s": entries / wall,
            "sweep": {0: "off", 1: "dense", 2: "sparse", 3: "full solve"}[state],
            "per_part_form": {"wall_ms": per_wall * 1e3, "entries_per_s": entries / per_wall}}


def batch_bench(args):
    """Many-graph batch: 512 road-like graphs (64x64-grid construction scaled: 10% of edges
    dropped, 2% diagonals, largest component) with log-uniform sizes in [256, 8192], half
    unweighted (int32) and half with float32 weights uniform in [0.5, 10) (FP64 arithmetic,
    float32 results); K per graph and workgroups per CU from the autotuner; a step is one
    replay of the captured batch (one persistent launch per distance type)."""
    from fastapsp_amd import autotune as AT
    t0 = time.perf_counter()
    if getattr(args, "no_tune", False):  # fixed settings (profiling runs)
        table = {"K": [8, 8, 8, 16, 16], "per_cu": [0] * 5, "src_rows": [32] * 5,
                 "merge_rows": [16] * 5, "packed": [1] * 5, "unit_grid": [1] * 5,
                 "fw_tile": [64] * 5, "ms_per_graph": [None] * 5}
    elif getattr(args, "tune_quick", False):
        table = AT.tune(k_candidates=(8, 16), per_cu_candidates=(0,), src_rows_candidates=(16,),
                        merge_rows_candidates=(8, 16), max_copies=16)
    else:
        table = AT.tune(max_copies=32)
    log(f"autotune {time.perf_counter() - t0:.1f}s: K per class {table['K']}, "
        f"per_cu {table['per_cu']}, source rows {table['src_rows']}, "
        f"merge rows {table['merge_rows']}, packed {table['packed']}, "
        f"unit grid {table['unit_grid']}, FW tile {table['fw_tile']}")
    rng = np.random.default_rng(2026)
    sizes = np.exp(rng.uniform(np.log(256), np.log(8192), args.graphs)).astype(int)
    t0 = time.perf_counter()
    graphs = []
    for i, nt in enumerate(sizes):
        K = AT.k_for(int(nt), table)
        graphs.append(AT.road_graph(int(nt), 5000 + i, K, float_weights=(i % 2 == 1)))
    for g in graphs:
        g.preCondition()
    ns = [g.num_vertexs for g in graphs]
    entries = float(sum(n * n for n in ns))
    log(f"{len(graphs)} graphs built in {time.perf_counter() - t0:.1f}s: n in [{min(ns)}, "
        f"{max(ns)}], sum n^2 = {entries:.3e} ({entries * 4 / 1e9:.1f} GB of results)")
    settings = AT.settings_for(ns, table)
    if getattr(args, "unit_grid", None) is not None:
        settings["unit_grid"] = int(args.unit_grid)
    b = F.Batch(graphs, use_graph=True, **settings)
    tiles = settings.get("fw_tile", [64] * len(graphs))
    settings = dict(settings, fw_tile={"64": sum(1 for t in tiles if t == 64),
                                       "32": sum(1 for t in tiles if t == 32)})
    for _ in range(args.warmup):
        b.run()
    L.check(L.load().fap_hip_device_sync(), "sync")
    ts = time.perf_counter()
    for _ in range(args.steps):
        b.run(sync=False)
    L.check(L.load().fap_hip_device_sync(), "sync")
    ms = (time.perf_counter() - ts) * 1e3 / args.steps
    value = entries / (ms * 1e-3)
    log(f"batched (persistent, graph replay): {ms:.2f} ms/step, {value:.3e} entries/s")
    plain = None
    if not args.skip_plain:
        tp = time.perf_counter()
        F.solve_batch(graphs, F.FAP_MODE_PLAIN)
        plain_ms = (time.perf_counter() - tp) * 1e3
        plain = {"ms_per_step": plain_ms, "batched_vs_plain": plain_ms / ms}
        log(f"plain (graph after graph, per-part launches): {plain_ms:.2f} ms "
            f"({plain_ms / ms:.2f}x slower)")
    cpu = None
    if not args.skip_cpu:
        import oracle as O
        tc = time.perf_counter()
        done, ent = 0, 0.0
        for g in graphs:
            if g.isFloat():
                continue
            O.solve_parts(1, g.K, g.csr(), g.layout())
            done += 1
            ent += float(g.num_vertexs) ** 2
            if time.perf_counter() - tc >= args.cpu_budget:
                break
        dt = time.perf_counter() - tc
        cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        cpu = {"value": ent / dt, "unit": "entries/s", "cores": cores, "kind": "port",
               "sample": f"{done} integer-weight graphs of the batch through the CPU path "
                         f"({ent:.3e} entries, {dt:.1f} s)"}
        log(f"cpu baseline: {cpu['value']:.3e} entries/s")
    out = {"metric": "APSP distance entries/s, many-graph batch (int32 exact / FP32 at 1e-6)",
           "value": value, "unit": "entries/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "int32+f64", "data": "synthetic",
           "config": {"workload": f"{len(graphs)} road-like graphs, n log-uniform in [256, 8192], "
                                  "half float32-weighted, one persistent launch per type",
                      "graphs": len(graphs), "sum_n2": entries, "n_min": min(ns),
                      "n_max": max(ns), "settings": settings},
           "autotune": {k: table[k] for k in ("K", "per_cu", "src_rows", "merge_rows",
                                               "packed", "unit_grid", "fw_tile",
                                               "ms_per_graph")},
           "plain_form": plain, "cpu_baseline": cpu}
    return out


def serve_bench(args):
    """Serving: a pool of resident graphs (road-like and geometric kNN, sizes log-uniform
    in [1k, 64k] vertices; integer weights in [1, 1000] or float32 weights in [0.5, 10)),
    Poisson arrivals of full-APSP and sub-graph requests at 10, 50 and 200 req/s (and
    1000 req/s, above capacity, for the saturated rate) against a 200 GB HBM budget;
    each result is consumed (released) as soon as it is done.  Reports p50 / p99 latency
    (submit -> completion) and result entries/s per rate."""
    import math
    from fastapsp_amd import serve as S
    from fastapsp_amd.autotune import road_graph, symmetric_float_weights
    rng = np.random.default_rng(77)
    sizes = np.exp(rng.uniform(np.log(1000), np.log(65536), args.pool)).astype(int)
    t0 = time.perf_counter()
    pool = []
    for i, nt in enumerate(sizes):
        # Parts of ~400 vertices for the batch lane, ~900 (as the 114,599-vertex
        # workload) for graphs that take the pipeline lane (>= 8192 vertices).
        K = int(max(2, math.ceil(nt / (400 if nt < 8192 else 900))))
        kind = i % 4  # road int, road float, knn int, knn float
        if kind < 2:
            g = road_graph(int(nt), 900 + i, K)
            c = g.csr()
            w = symmetric_float_weights(c["row_offset"], c["col"], 900 + i)
            if kind == 0:  # integers 1..1000, equal both ways
                w = (np.floor((w - 0.5) / 9.5 * 1000) + 1).astype(np.float32)
            g = F.FapGraph.from_csr(c["row_offset"], c["col"], w, False, K)
        else:
            g = F.FapGraph.knn(int(nt), 3, 1000, 900 + i, K)
            if kind == 3:
                c = g.csr()
                w = symmetric_float_weights(c["row_offset"], c["col"], 900 + i, directed=True)
                g = F.FapGraph.from_csr(c["row_offset"], c["col"], w, True, K)
        g.preCondition()
        pool.append(g)
    log(f"pool of {len(pool)} graphs built in {time.perf_counter() - t0:.1f}s, n in "
        f"[{min(g.num_vertexs for g in pool)}, {max(g.num_vertexs for g in pool)}]")
    srv = S.Server(budget_bytes=200e9, streams=4)
    idx = [srv.add_graph(g) for g in pool]
    # warm-up: one request per graph
    for i in idx:
        t = srv.submit(i, 0)
        srv.wait(t)
        srv.release(t)
    results = []
    # 10, 50 and 200 req/s are the stated rates; 1000 req/s probes the saturated rate.
    for rate, dur in ((10.0, 5.0), (50.0, 4.0), (200.0, 3.0), (1000.0, 2.0)):
        srv.close()
        srv = S.Server(budget_bytes=200e9, streams=4)
        idx = [srv.add_graph(g) for g in pool]
        nreq = int(rate * dur)
        reqs = []
        for _ in range(nreq):
            gi = int(rng.integers(len(pool)))
            sub = 0 if rng.random() < 0.5 else int(rng.integers(1, pool[gi].K + 1))
            reqs.append((idx[gi], sub))
        tickets, states, wall = S.poisson_load(srv, reqs, rate, seed=int(rate))
        st = srv.stats()
        res = {"offered_req_s": rate, "requests": nreq, "completed": st["completed"],
               "failed": st["failed"], "batches": st["batches"], "wall_s": wall,
               "p50_ms": st["p50_ms"], "p99_ms": st["p99_ms"], "mean_ms": st["mean_ms"],
               "entries_per_s": st["entries"] / max(st["span_s"], 1e-9),
               "peak_reserved_GB": st["peak_reserved_bytes"] / 1e9}
        results.append(res)
        log(f"rate {rate:g} req/s: {st['completed']}/{nreq} done in {wall:.2f}s, "
            f"{st['batches']} batches, p50 {st['p50_ms']:.1f} ms, p99 {st['p99_ms']:.1f} ms, "
            f"{res['entries_per_s']:.3e} entries/s, peak {res['peak_reserved_GB']:.1f} GB")
    srv.close()
    return {"serve": results, "pool": len(pool),
            "workload": "road-like and kNN graphs, n log-uniform in [1k, 64k], int "
                        "[1,1000] or float32 [0.5,10) weights, 50% full APSP / 50% "
                        "sub-graph requests, Poisson arrivals, 200 GB budget"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--config", default="road1",
                    help="road1 (default), road0, batch512 (many-graph batch), serve, knn262k or "
                         "wroad62k (solveSubGraph sweeps)")
    ap.add_argument("--graphs", type=int, default=512, help="graphs of the batch512 config")
    ap.add_argument("--unit-grid", type=int, default=None,
                    help="batch512: 1 = merge as a plain grid, 0 = persistent queue "
                         "(default: the autotuner's choice)")
    ap.add_argument("--pool", type=int, default=24, help="resident graphs of the serve config")
    ap.add_argument("--skip-plain", action="store_true")
    ap.add_argument("--paths", action="store_true",
                    help="also time solve(True): distances and predecessor matrices")
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--no-tune", action="store_true",
                    help="batch512: fixed settings instead of the autotuner (profiling)")
    ap.add_argument("--skip-legs", action="store_true",
                    help="headline workload only (no kNN / min-plus / batch / serving legs)")
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--skip-forms", action="store_true",
                    help="skip the overlap and int32-FW comparison solves (for profiling)")
    ap.add_argument("--in-part", default="auto", choices=["auto", "fw"],
                    help="in-part distances of the timed solves: auto (bit-parallel BFS where "
                         "it applies) or fw (Floyd-Warshall; for profiling its kernels)")
    ap.add_argument("--mode", default="batched", choices=["batched", "overlap", "plain"],
                    help="execution form of the timed solves (default: batched)")
    ap.add_argument("--minplus", type=int, default=0,
                    help="run only the dense N x N x N int32 min-plus microbench")
    ap.add_argument("--probe", type=float, default=0.0,
                    help="run only the HBM fill/copy probe on buffers of this many GiB")
    args = ap.parse_args()
    if args.probe > 0:
        probe(args.probe)
        return
    if args.minplus > 0:
        res = minplus_bench(args.minplus)
        print(json.dumps({"minplus_bench": res}), flush=True)
        if not res["sampled_entries_exact"]:
            sys.exit(1)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); multi-GPU sharding is "
              "a design note, not a run", file=sys.stderr)
        sys.exit(2)
    if args.config == "serve":
        print(json.dumps(serve_bench(args)), flush=True)
        return
    if args.config == "batch512":
        print(json.dumps(batch_bench(args)), flush=True)
        return
    if args.config == "knn262k":
        print(json.dumps(knn_subgraph_bench(args)), flush=True)
        return
    if args.config == "wroad62k":
        print(json.dumps(wroad_subgraph_bench(args)), flush=True)
        return
    t0 = time.perf_counter()
    g, workload = make_graph(args.config)
    log(f"graph n={g.num_vertexs} m={g.num_edges} K={g.K} built+partitioned in "
        f"{time.perf_counter() - t0:.1f}s (edge cut {g.getEdgeCut():.0f})")
    g.preCondition()
    lay = g.layout()
    n = g.num_vertexs
    num = lay["num"][1:].astype(np.int64)
    bdy = lay["bdy"][1:].astype(np.int64)
    inner = num - bdy
    log(f"boundary total {bdy.sum()} (per part mean {bdy.mean():.1f}, max {bdy.max()}); "
        f"part size mean {num.mean():.1f} max {num.max()}")

    modes = {"batched": F.FAP_MODE_BATCHED, "overlap": F.FAP_MODE_OVERLAP, "plain": F.FAP_MODE_PLAIN}
    if args.in_part == "fw":
        modes = {k: v | F.FAP_MODE_FLAG_FW for k, v in modes.items()}
    g.setMode(modes[args.mode])
    for i in range(args.warmup):
        g.solve()
        log(f"warmup {i + 1}/{args.warmup}: {stage_times()['total_ms']:.2f} ms device")
    L.check(L.load().fap_hip_device_sync(), "sync")
    stages = []
    t_start = time.perf_counter()
    for i in range(args.steps):
        g.solve()
        stages.append(stage_times())
        log(f"step {i + 1}/{args.steps}: {stages[-1]['total_ms']:.2f} ms device "
            f"(merge {stages[-1]['minplus_ms']:.2f})")
    L.check(L.load().fap_hip_device_sync(), "sync")
    elapsed = time.perf_counter() - t_start
    ms_per_step = elapsed * 1e3 / args.steps
    entries = float(n) * n
    value = entries / (ms_per_step * 1e-3)
    log(f"{args.mode}: {ms_per_step:.2f} ms/step, {value:.3e} entries/s")

    avg = {k: float(np.mean([s[k] for s in stages])) for k in stages[0]}
    bfs = stages[-1]["parts_bfs"] > 0
    bform = "boundary graph" if stages[-1]["boundary_graph"] else "boundary SSSP"
    inform = "bit-parallel BFS" if bfs else "Floyd-Warshall"
    log(f"stages (ms): boundary rows ({bform}) {avg['sssp_ms']:.2f} build {avg['build_ms']:.2f} "
        f"in-part ({inform}) {avg['floyd_ms']:.2f} merge {avg['minplus_ms']:.2f} "
        f"decode {avg['decode_ms']:.2f}; 16-bit FW parts {stages[-1]['parts_fw16']}, "
        f"redone {stages[-1]['parts_redone']}, BFS parts {stages[-1]['parts_bfs']}")
    # In-part distances: the s_p x s_p matrices, written once (int32).
    in_entries = float((num.astype(np.float64) ** 2).sum())
    in_part = {"form": inform, "ms": avg["floyd_ms"], "entries": in_entries,
               "entries_per_s": in_entries / (avg["floyd_ms"] * 1e-3)}

    def fw_roofline(st_avg):
        # Floyd-Warshall stage: 2 ops (add, min) per step, sum over parts of s_p^3
        # steps, against the VALU peak of the arithmetic it ran in (packed 16-bit: two
        # entries per lane-op).
        fw16 = st_avg["parts_fw16"] > 0
        fw_ops = 2.0 * float((num.astype(np.float64) ** 3).sum())
        fw_peak = 2 * VALU_PEAK_TOPS if fw16 else VALU_PEAK_TOPS
        fw_tops = fw_ops / (st_avg["floyd_ms"] * 1e-3) / 1e12
        return {"form": "packed u16 (v_pk_add_u16 + v_pk_minimum3_f16)" if fw16 else "int32",
                "ops": fw_ops, "ms": st_avg["floyd_ms"], "TOPs": fw_tops, "peak_TOPs": fw_peak,
                "frac": fw_tops / fw_peak, "build_ms": st_avg["build_ms"]}

    fw_roof = None if bfs else fw_roofline(avg)

    # Roofline of the dominant kernel: the merge stage of the timed batched steps.  Its
    # launches (one per padded-K group, spread over three streams) are bracketed by HIP
    # events on the library stream: start before the fork, end after the join.
    merge_ms = [st["minplus_ms"] for st in stages]
    merge_avg = float(np.mean(merge_ms))

    # The overlapped form, for comparison (two steps).
    overlap_ms = None
    fw_form_ms = None
    if not args.skip_forms:
        g.setMode(F.FAP_MODE_OVERLAP)
        ts = time.perf_counter()
        for i in range(2):
            g.solve()
        overlap_ms = (time.perf_counter() - ts) * 1e3 / 2
        log(f"overlap form: {overlap_ms:.2f} ms/step")
        if bfs:  # the same solve with the in-part distances from Floyd-Warshall
            g.setMode(modes[args.mode] | F.FAP_MODE_FLAG_FW)
            g.solve()
            fst = []
            ts = time.perf_counter()
            for i in range(2):
                g.solve()
                fst.append(stage_times())
            fw_form_ms = (time.perf_counter() - ts) * 1e3 / 2
            fw_roof = fw_roofline({k: float(np.mean([x[k] for x in fst])) for k in fst[0]})
            log(f"Floyd-Warshall in-part form: {fw_form_ms:.2f} ms/step "
                f"(FW {fw_roof['ms']:.2f} ms + build {fw_roof['build_ms']:.2f} ms)")
        g.setMode(modes[args.mode])
    # Dominant kernel: the merge (merge_stream_kernel, one launch per padded-K group,
    # timed together as the merge stage of the batched solves).  Per solve it computes
    # sum_p i_p * b_p * n min-plus steps and its algorithmic bytes are
    # 4 * (|A| + |B| + |C|): the inner x boundary FW block, the boundary rows and the
    # inner rows of the result, each once.
    steps = float((inner * bdy).sum()) * n
    ops = 2.0 * steps
    bytes_alg = 4.0 * float((inner * bdy).sum() + bdy.sum() * n + inner.sum() * n)
    # Floors: VALU at 1.5 instructions per step (two v_add_u32 + one v_min3_i32 per
    # two steps; the packed 16-bit rows take half that) against 39.3 T
    # lane-instructions/s, and HBM at 8 TB/s.
    t_valu = 1.5 * steps / (VALU_PEAK_TOPS * 1e12)
    t_hbm = bytes_alg / (HBM_PEAK_GBS * 1e9)
    hbm_ach = bytes_alg / (merge_avg * 1e-3) / 1e9
    valu_ach = ops / (merge_avg * 1e-3) / 1e12
    if t_hbm >= t_valu:
        roofline = {"bound": "hbm", "achieved": hbm_ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": hbm_ach / HBM_PEAK_GBS}
    else:
        inst = 1.5 * steps / (merge_avg * 1e-3) / 1e12
        roofline = {"bound": "valu", "achieved": inst, "peak": VALU_PEAK_TOPS,
                    "unit": "T lane-instr/s", "frac": inst / VALU_PEAK_TOPS}
    roofline.update({"traffic": merge_traffic(args.config),
                     "kernel": "merge_stream_kernel (min-plus + fused decode), all K-groups",
                     "alg_bytes_per_launch": bytes_alg, "ops_per_launch": ops,
                     "avg_launch_ms": merge_avg, "hbm_GBs": hbm_ach,
                     "hbm_frac": hbm_ach / HBM_PEAK_GBS, "valu_ops_TOPs": valu_ach,
                     "floor_ms": {"hbm": t_hbm * 1e3, "valu": t_valu * 1e3}})

    plain = None
    if not args.skip_plain:
        g.setMode(F.FAP_MODE_PLAIN)
        tp = time.perf_counter()
        g.solve()
        L.check(L.load().fap_hip_device_sync(), "sync")
        plain_ms = (time.perf_counter() - tp) * 1e3
        pst = stage_times()
        plain = {"ms_per_step": plain_ms, "stages_ms": pst,
                 "batched_vs_plain": plain_ms / ms_per_step}
        log(f"plain: {plain_ms:.2f} ms/step ({plain_ms / ms_per_step:.2f}x slower)")
        g.setMode(F.FAP_MODE_BATCHED)

    paths = None
    if args.paths:
        # solve(True): the distances plus the n x n predecessor matrix with the reference
        # CPU path's tie-breaking (SSSP boundary rows + predecessor rule, blocked
        # Floyd-Warshall with per-step records, first-k min-plus), both resident.
        g.solve(True)
        L.check(L.load().fap_hip_device_sync(), "sync")
        tp = time.perf_counter()
        for _ in range(2):
            g.solve(True)
        L.check(L.load().fap_hip_device_sync(), "sync")
        pms = (time.perf_counter() - tp) * 1e3 / 2
        pst = stage_times()
        paths = {"ms_per_step": pms, "entries_per_s": entries / (pms * 1e-3), "stages_ms": pst}
        log(f"solve with paths: {pms:.2f} ms/step ({entries / (pms * 1e-3):.3e} entries/s); "
            f"stages {pst}")
        g.solve()  # leave the distance-only result resident

    # The same-kernel comparison: the batched form with boundary SSSP rows and the
    # Floyd-Warshall in-part distances runs exactly the kernels the plain form launches
    # part after part (SSSP, build, 16-bit Floyd-Warshall + unpack, merge, decode).
    same = None
    if not args.skip_plain:
        g.setMode(F.FAP_MODE_BATCHED | F.FAP_MODE_FLAG_SSSP | F.FAP_MODE_FLAG_FW)
        g.solve()
        L.check(L.load().fap_hip_device_sync(), "sync")
        ts = time.perf_counter()
        for _ in range(3):
            g.solve()
        L.check(L.load().fap_hip_device_sync(), "sync")
        same_ms = (time.perf_counter() - ts) * 1e3 / 3
        sst = stage_times()
        same = {"ms_per_step": same_ms, "stages_ms": sst,
                "form": "batched, boundary SSSP rows, Floyd-Warshall in-part (the plain form's "
                        "kernels)"}
        if plain is not None:
            plain["batched_vs_plain_same_kernels"] = plain["ms_per_step"] / same_ms
            plain["batched_vs_plain_algorithmic"] = plain["ms_per_step"] / ms_per_step
            log(f"batched with the plain form's kernels: {same_ms:.2f} ms/step "
                f"(plain {plain['batched_vs_plain_same_kernels']:.2f}x slower)")
        g.setMode(modes[args.mode])
        g.solve()

    cpu = None
    if not args.skip_cpu:
        cpu = cpu_baseline(g, args.cpu_budget)
        log(f"cpu baseline: {cpu['value']:.3e} entries/s ({cpu['sample']})")

    out = {"metric": METRIC, "value": value, "unit": "entries/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per_step,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "int32",
           "data": "synthetic",
           "config": {"workload": workload, "n": n, "m": g.num_edges, "K": g.K,
                      "boundary_total": int(bdy.sum()), "mode": args.mode,
                      "boundary_rows": bform,
                      "result_bytes_in_hbm": int(n) * int(n) * 4},
           "stages_ms": avg, "in_part": in_part, "floyd_warshall": fw_roof,
           "fw_in_part_form": {"ms_per_step": fw_form_ms},
           "overlap_form": {"ms_per_step": overlap_ms},
           "roofline": roofline, "cpu_baseline": cpu, "plain_form": plain}
    if paths is not None:
        out["paths"] = paths
    if same is not None:
        out["batched_same_kernels_form"] = same
    if args.config == "road1" and not args.skip_legs:
        # The other workloads, measured in the same run (each leg after the headline
        # graph's memory is released; a failing leg reports its error, not the run's).
        g.close()
        legs = {}
        for name, fn in [("knn262k_k256_subgraph", lambda: knn_subgraph_bench(args)),
                         ("wroad62k_k64_subgraph", lambda: wroad_subgraph_bench(args)),
                         ("minplus_8192", lambda: minplus_bench(8192)),
                         ("batch512", lambda: batch_bench(argparse.Namespace(
                             graphs=512, steps=3, warmup=1, skip_plain=True, skip_cpu=True,
                             cpu_budget=0.0, tune_quick=True))),
                         ("serve", lambda: serve_bench(argparse.Namespace(pool=24)))]:
            t0 = time.perf_counter()
            try:
                legs[name] = fn()
            except Exception as e:  # noqa: BLE001 -- reported in the JSON line
                legs[name] = {"error": f"{type(e).__name__}: {e}"}
            log(f"leg {name}: {time.perf_counter() - t0:.1f} s")
        out["other_configs"] = legs
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

#!/usr/bin/env python3
"""Throughput of maximal k-plex enumeration on one MI355X.

One step = one kp_run over the prepared workload: GPU root decomposition,
subgraph build and the persistent k-search/Branch kernels, with the
preprocessed graph already resident in HBM (host preprocessing and the
upload happen once, before the timed region, in kp_prepare).

Prints ONE JSON line (last line of stdout); progress goes to stderr.
"""
import argparse
import json
import os
import statistics
import sys
import time

# The library runs the size classes of an epoch on separate streams (nine per
# engine); HIP maps streams round-robin onto GPU_MAX_HW_QUEUES hardware
# queues, and with 4 two classes can share one and run back to back
# (soc-epinions: 210 ms at 4 queues, 165 ms at 16).  HIP reads the variable
# once, when its runtime starts, so it is set before the library loads.
os.environ["GPU_MAX_HW_QUEUES"] = "16"

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "maximal-kplex-enumeration_amd"))

HBM_PEAK_GBS = 8000.0   # MI355X HBM3E
# name -> (description, generator in kplex.py or None for a shipped dataset, root window or None)
WORKLOADS = {
    "cl75k": ("chung-lu n=75000 gamma=2.1 maxdeg=3000 ~508k edges + 300 communities 35-60 @0.9, k=4 q=30; "
              "roots at degeneracy positions [0,9000) of 16915 (the whole graph holds >1e10 plexes)",
              "workload_cl75k", (0, 9000)),
    "cl75k-full": ("chung-lu n=75000 gamma=2.1 maxdeg=3000 ~508k edges + 300 communities 35-60 @0.9, k=4 q=30",
                   "workload_cl75k", None),
    "er2000": ("G(2000,0.01) + 20 planted groups of 40, k=3 q=20", "workload_er2000", None),
    "cl320k": ("chung-lu n=320000 gamma=2.0 ~1.05M edges + 1000 communities 30-80 @0.85, k=5 q=30, truss; "
               "roots at degeneracy positions [0,800) of 50565 (1.96e9 tasks)", "workload_cl320k", (0, 800)),
    "cl320k-full": ("chung-lu n=320000 gamma=2.0 ~1.05M edges + 1000 communities 30-80 @0.85, k=5 q=30, truss",
                    "workload_cl320k", None),
    "soc-epinions": ("reference dataset soc-epinions (README.md:142), k=4 q=30", None, None),
    "email-euall": ("reference dataset email-euall (README.md:144), k=3 q=12", None, None),
}
# published reference results for the shipped datasets: (count, CudaPlex seconds on RTX 6000 Ada)
PUBLISHED = {"soc-epinions": (13172906, 4.13), "email-euall": (32639016, 1.32)}


def dataset_workload(name):
    import gzip
    import shutil
    import tempfile
    import kplex
    src = os.path.join(ROOT, "tests", "golden", f"{name}.bin.gz")
    fd, path = tempfile.mkstemp(suffix=".bin")
    with gzip.open(src, "rb") as fi, os.fdopen(fd, "wb") as fo:
        shutil.copyfileobj(fi, fo)
    g = kplex.Graph.load_bin(path)
    os.unlink(path)
    k, q = (4, 30) if name == "soc-epinions" else (3, 12)
    return g, k, q, 0.25, 0


def profiled_traffic(name):
    """HBM bytes per search-kernel launch measured with rocprofv3 FETCH_SIZE /
    WRITE_SIZE (copy-calibrated) for this workload, from its committed summary."""
    path = os.path.join(ROOT, "profiles", f"search_kernel_{name}_window.md")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            if line.startswith("traffic_bytes_per_launch:"):
                return int(line.split(":")[1])
    return None


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(graph, k, q, truss, target_s, window=None):
    """The CPU oracle (OpenMP C restatement of the reference path) on a bounded
    sample: root positions [b, R) of the degeneracy order (b = window start),
    R doubled until the sample has taken >= target_s seconds or covers the
    window."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import kpo
    deg, nbr = graph.csr()
    og = kpo.OGraph.from_csr(deg, nbr)
    threads = os.cpu_count() or 1
    total_plex = 0
    total_s = 0.0
    wb = window[0] if window else 0
    lo, hi = wb, wb + 64
    full = None
    while True:
        t0 = time.perf_counter()
        r = kpo.enumerate_plexes(og, k, q, truss, threads=threads, pos_begin=lo, pos_end=hi)
        dt = time.perf_counter() - t0
        total_plex += r["count"]
        total_s += dt
        nroots = max(0, r["pn"] - q + 2)
        if window:
            nroots = min(nroots, window[1])
        if hi >= nroots:
            full = True
            hi = nroots
            break
        if total_s >= target_s:
            full = False
            break
        lo, hi = hi, wb + 2 * (hi - wb)
    return {"value": total_plex / total_s if total_s > 0 else 0.0, "unit": "maximal k-plexes/s", "cores": threads,
            "kind": "port",
            "sample": (f"oracle (OpenMP C restatement) on root positions [{wb},{hi}) of [{wb},{nroots})"
                       f"{' (all roots)' if full else ''}: {total_plex} plexes in {total_s:.2f} s, "
                       f"preprocessing included per sample chunk")}


def batched_vs_plain(ngraphs, time_limit, chunks=1, autotune=True, fixed_tune=None):
    """The many-graph workload (Chung-Lu exponent 2.3, average degree 10, n
    log-uniform in [200, 20000], 1-20 planted 25-50-vertex communities at 0.9,
    k in {2,3,4}, q = 3k+8), first `ngraphs` graphs of the 4,096: every graph
    through the plain entry point one after another (kp_enumerate: preprocess,
    upload, decompose, search per graph), then all of them through one
    kp_enumerate_batch call, with the one-word class's grid autotuned on the
    batch first (untimed).  Results must be identical graph by graph."""
    import kplex
    specs = kplex.batch_graph_specs()[:ngraphs]
    graphs = gen_graphs(specs)
    params = [(s["k"], s["q"]) for s in specs]
    log(f"[bench] batched: {len(graphs)} graphs, n {min(s['n'] for s in specs)}..{max(s['n'] for s in specs)}")
    kplex.enumerate_plexes(graphs[0], *params[0], time_limit_s=time_limit)   # warm the library
    t0 = time.perf_counter()
    plain = []
    for i, (g, (k, q)) in enumerate(zip(graphs, params)):
        plain.append(kplex.enumerate_plexes(g, k, q, time_limit_s=time_limit))
        if (i + 1) % 256 == 0:
            log(f"[bench] batched: plain {i + 1}/{len(graphs)} graphs in {time.perf_counter() - t0:.1f} s")
    plain_s = time.perf_counter() - t0
    log(f"[bench] batched: plain {plain_s:.2f} s")
    # Autotuning of the batched launch: the one-word class's grid (blocks per
    # CU) and donation period are picked from a few candidates by one untimed
    # batch run each; the timed run below uses the fastest.
    tune, tune_runs = fixed_tune, {}
    if autotune and chunks == 1 and fixed_tune is None:
        cands = {"default": None, "bpc5": {1: {"blocks_per_cu": 5}}, "bpc7": {1: {"blocks_per_cu": 7}},
                 "dp2048": {1: {"donate_period": 2048}},
                 "bpc6_dp2048": {1: {"blocks_per_cu": 6, "donate_period": 2048}},
                 "bpc7_dp2048": {1: {"blocks_per_cu": 7, "donate_period": 2048}}}
        for name, cand in cands.items():
            kplex.set_tune(cand)
            t1 = time.perf_counter()
            kplex.enumerate_batch(graphs, [dict(k=k, q=q, time_limit_s=time_limit) for k, q in params])
            tune_runs[name] = time.perf_counter() - t1
        best = min(tune_runs, key=tune_runs.get)
        tune = cands[best]
        log(f"[bench] batched: tuning runs {', '.join(f'{a} {b:.2f} s' for a, b in tune_runs.items())}; using {best}")
    kplex.set_tune(tune)
    t0 = time.perf_counter()
    batch, call_gpu_s = [], 0.0
    step = (len(graphs) + chunks - 1) // chunks
    for b in range(0, len(graphs), step):
        part = kplex.enumerate_batch(graphs[b:b + step],
                                     [dict(k=k, q=q, time_limit_s=time_limit) for k, q in params[b:b + step]])
        call_gpu_s += (part[0]["decompose_ms"] + part[0]["search_ms"]) / 1e3   # per call, not per graph
        batch += part
        if chunks > 1:
            log(f"[bench] batched: graphs [{b},{min(len(graphs), b + step)}) in one batch, "
                f"{time.perf_counter() - t0:.1f} s so far")
    batch_s = time.perf_counter() - t0
    kplex.set_tune(None)
    log(f"[bench] batched: {chunks} batch call(s) {batch_s:.2f} s")
    for a, b in zip(plain, batch):
        for x in ("count", "hash", "tasks", "valid_roots"):
            assert a[x] == b[x], ("batched result differs", x, a[x], b[x])
    count = sum(r["count"] for r in batch)
    return {"graphs": len(graphs), "batch_calls": chunks, "plexes": count, "tasks": sum(r["tasks"] for r in batch),
            "plain_s": plain_s, "batched_s": batch_s, "plain_plexes_per_s": count / plain_s,
            "batched_plexes_per_s": count / batch_s, "batched_vs_plain": plain_s / batch_s,
            "plain_gpu_s": sum(r["decompose_ms"] + r["search_ms"] for r in plain) / 1e3,
            "batched_gpu_s": call_gpu_s, "autotune_runs_s": tune_runs,
            "autotune_settings": {str(a): b for a, b in tune.items()} if tune else None,
            "identical": True}


def gen_graphs(specs):
    """Generates the graphs of `specs` on host threads (the generator call
    releases the GIL)."""
    import concurrent.futures
    import kplex
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(16, os.cpu_count() or 1)) as ex:
        return list(ex.map(kplex.graph_from_spec, specs))


def served_vs_plain(duration_s, rate, time_limit, seed=7, slice_ms=30.0, plain=True):
    """The served workload: 256 resident graphs built like the many-graph
    workload but with n up to 200,000; queries (graph uniform, k in {3,4,5},
    q = 5k+10) arrive as a Poisson stream of `rate` per second for
    `duration_s` seconds.  The server admits them into running epochs
    (continuous batching); latency = completion - arrival.  Then the same
    queries run one after another through the plain entry point
    (kp_enumerate per query).  Results must be identical query by query."""
    import kplex
    specs = kplex.served_graph_specs()
    t0 = time.perf_counter()
    graphs = gen_graphs(specs)
    gen_s = time.perf_counter() - t0
    arrivals, queries = kplex.served_queries(rate, duration_s, seed, len(graphs))
    log(f"[bench] served: {len(graphs)} graphs generated in {gen_s:.1f} s; {len(queries)} queries at {rate}/s "
        f"over {duration_s} s")
    srv = kplex.Server(max_inflight=128, slice_ms=slice_ms)
    ids = [srv.add_graph(g) for g in graphs]
    t0 = time.perf_counter()
    srv.preload([(ids[i], k, 5 * k + 10) for i in range(len(graphs)) for k in (3, 4, 5)])
    preload_s = time.perf_counter() - t0
    log(f"[bench] served: 768 (graph, k, q) preprocessed and resident in {preload_s:.1f} s")
    done = {}
    tick = {}
    nxt = 0
    start = time.perf_counter()
    last_log = start
    while len(done) < len(queries):
        now = time.perf_counter() - start
        while nxt < len(queries) and arrivals[nxt] <= now:
            gi, k, q = queries[nxt]
            tick[srv.submit(ids[gi], k, q)] = nxt
            nxt += 1
        got = srv.poll()
        tnow = time.perf_counter() - start
        for tk, r in got:
            done[tick[tk]] = (tnow, r)
        if time.perf_counter() - last_log > 10:
            last_log = time.perf_counter()
            log(f"[bench] served: {len(done)}/{len(queries)} done at {tnow:.1f} s")
        if not got:
            time.sleep(0.0001)
    makespan = max(t for t, _ in done.values())
    srv.close()
    lat = sorted(done[i][0] - arrivals[i] for i in range(len(queries)))
    pct = lambda p: lat[min(len(lat) - 1, int(p * len(lat)))]
    count = sum(r["count"] for _, r in done.values())
    log(f"[bench] served: makespan {makespan:.2f} s, p50 {pct(0.5) * 1e3:.0f} ms, p99 {pct(0.99) * 1e3:.0f} ms "
        f"(admission slice {slice_ms} ms)")
    if not plain:
        return {"graphs": len(graphs), "queries": len(queries), "offered_per_s": rate, "arrival_window_s": duration_s,
                "plexes": count, "makespan_s": makespan, "served_queries_per_s": len(queries) / makespan,
                "served_plexes_per_s": count / makespan, "p50_ms": pct(0.5) * 1e3, "p99_ms": pct(0.99) * 1e3,
                "preload_s": preload_s}
    t0 = time.perf_counter()
    for i, (gi, k, q) in enumerate(queries):
        r = kplex.enumerate_plexes(graphs[gi], k, q, time_limit_s=time_limit)
        for x in ("count", "hash", "tasks", "valid_roots"):
            assert r[x] == done[i][1][x], ("served result differs", i, x, r[x], done[i][1][x])
        if (i + 1) % 500 == 0:
            log(f"[bench] served: plain {i + 1}/{len(queries)} queries in {time.perf_counter() - t0:.1f} s")
    plain_s = time.perf_counter() - t0
    log(f"[bench] served: plain one-by-one {plain_s:.2f} s")
    return {"graphs": len(graphs), "queries": len(queries), "offered_per_s": rate, "arrival_window_s": duration_s,
            "plexes": count, "makespan_s": makespan, "served_queries_per_s": len(queries) / makespan,
            "served_plexes_per_s": count / makespan, "p50_ms": pct(0.5) * 1e3, "p99_ms": pct(0.99) * 1e3,
            "plain_s": plain_s, "plain_queries_per_s": len(queries) / plain_s, "served_vs_plain": plain_s / makespan,
            "preload_s": preload_s, "identical": True}


def spill_line(time_limit, pause_every=200_000_000, queue_capacity=4096, thres=0.25):
    """The large workload with truss (Chung-Lu 320k, k=5 q=30), roots
    [0,800): once straight through, once with every launch paused after
    `pause_every` frames and all but thres * queue_capacity frames per class
    spilled to pinned host memory and refilled.  Results must match."""
    import kplex
    desc, fn, window = WORKLOADS["cl320k"]
    g, k, q, _, tr = getattr(kplex, fn)()
    log(f"[bench] spill: {desc}")
    p = kplex.Prepared(g, k, q, 0.25, tr, time_limit_s=time_limit, roots=window)
    p.run()
    plain = p.run()
    p.close()
    p = kplex.Prepared(g, k, q, thres, tr, time_limit_s=time_limit, roots=window, queue_capacity=queue_capacity,
                       pause_every=pause_every)
    sp = p.run()
    p.close()
    for x in ("count", "hash", "tasks", "valid_roots"):
        assert sp[x] == plain[x], ("spilled result differs", x)
    log(f"[bench] spill: straight {plain['search_ms']:.0f} ms, with {sp['pauses']} pauses and {sp['spilled']} "
        f"frames spilled {sp['search_ms']:.0f} ms")
    return {"workload": desc, "count": plain["count"], "hash": plain["hash"], "tasks": plain["tasks"],
            "straight_search_ms": plain["search_ms"], "spill_search_ms": sp["search_ms"],
            "pause_every_frames": pause_every, "device_frames_per_class": int(thres * queue_capacity),
            "pauses": sp["pauses"], "spilled_frames": sp["spilled"],
            "plexes_per_s_straight": plain["count"] / (plain["search_ms"] / 1e3),
            "plexes_per_s_spill": sp["count"] / (sp["search_ms"] / 1e3), "identical": True}


def small_line(reps=7):
    """The small planted workload (G(2000, 0.01) + 20 planted groups of 40,
    k=3 q=20) end to end, in the reference's timing scope (preprocessing +
    enumeration): the plain entry point on the GPU against the CPU oracle on
    all host cores, medians of `reps` runs after a warm-up, same results."""
    import statistics
    import kplex
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import kpo
    g, k, q, _, tr = kplex.workload_er2000()
    kplex.enumerate_plexes(g, k, q)
    gpu_s, r = [], None
    for _ in range(reps):
        t0 = time.perf_counter()
        r = kplex.enumerate_plexes(g, k, q)
        gpu_s.append(time.perf_counter() - t0)
    deg, nbr = g.csr()
    og = kpo.OGraph.from_csr(deg, nbr)
    threads = os.cpu_count() or 1
    kpo.enumerate_plexes(og, k, q, tr, threads=threads)
    cpu_s, c = [], None
    for _ in range(reps):
        t0 = time.perf_counter()
        c = kpo.enumerate_plexes(og, k, q, tr, threads=threads)
        cpu_s.append(time.perf_counter() - t0)
    assert (r["count"], r["hash"], r["tasks"]) == (c["count"], c["hash"], c["tasks"]), "small workload differs"
    g_med, c_med = statistics.median(gpu_s), statistics.median(cpu_s)
    log(f"[bench] small: GPU {g_med * 1e3:.2f} ms, CPU oracle ({threads} threads) {c_med * 1e3:.2f} ms")
    return {"workload": WORKLOADS["er2000"][0], "count": r["count"], "tasks": r["tasks"],
            "gpu_wall_ms": g_med * 1e3, "gpu_preprocess_ms": r["preprocess_ms"],
            "gpu_device_ms": r["decompose_ms"] + r["search_ms"], "cpu_wall_ms": c_med * 1e3, "cpu_threads": threads,
            "gpu_vs_cpu": c_med / g_med, "identical": True}


def full_graph_line(time_limit, lo=0, hi=None, workload="cl75k-full"):
    """A workload in full (the Chung-Lu 75k graph, or the Chung-Lu 320k graph
    with truss): every root position in [lo, hi) (default: all), enumerated
    as consecutive windows (windows partition the result, so counts, task
    counts and hashes add up), each window one kp_run that prints its line
    when it ends; the last, heaviest positions go in narrower windows.  Over
    [0, end) the sum is the whole graph's exact count.  A window that fails
    (its time limit) ends the line, which then reports the positions done."""
    import kplex
    desc, fn, _ = WORKLOADS[workload]
    g, k, q, thres, tr = getattr(kplex, fn)()
    end = len(kplex.preprocess(g, k, q, tr)[0]) - q + 2
    hi = end if hi is None else min(hi, end)
    if workload == "cl75k-full":
        cuts = ([0, 9000] + list(range(9500, min(end, 16000), 500)) + list(range(16000, min(end, 16400), 100)) +
                list(range(16400, end, 25)))
    else:
        cuts = ([0, 800] + list(range(2000, end - 2000, 2000)) + list(range(end - 2000, end - 250, 250)) +
                list(range(end - 250, end, 25)))
    cuts += [end, lo, hi]
    cuts = sorted(set(c for c in cuts if lo <= c <= hi))
    windows = list(zip(cuts[:-1], cuts[1:]))
    tot = {"count": 0, "tasks": 0, "hash_a": 0, "hash_b": 0, "search_ms": 0.0, "valid_roots": 0}
    per = []
    t0 = time.perf_counter()
    p = kplex.Prepared(g, k, q, thres, tr, time_limit_s=time_limit)
    failed = None
    for w in windows:
        p.set_roots(w)
        try:
            r = p.run()
        except kplex.KplexError as e:
            failed = {"window": list(w), "error": str(e)}
            log(f"[bench] full graph window {w}: {e}; stopping")
            break
        for x in ("count", "tasks", "valid_roots"):
            tot[x] += r[x]
        tot["hash_a"] = (tot["hash_a"] + r["hash_a"]) % (1 << 64)
        tot["hash_b"] = (tot["hash_b"] + r["hash_b"]) % (1 << 64)
        tot["search_ms"] += r["search_ms"] + r["decompose_ms"]
        per.append({"window": list(w), "count": r["count"], "hash": r["hash"], "tasks": r["tasks"],
                    "gpu_ms": r["search_ms"] + r["decompose_ms"]})
        log(f"[bench] full graph window {w}: count={r['count']} tasks={r['tasks']} gpu_ms={per[-1]['gpu_ms']:.0f} "
            f"(total {tot['count']} after {time.perf_counter() - t0:.1f} s)")
    wall = time.perf_counter() - t0
    p.close()
    if failed:
        hi = failed["window"][0]
    return {"workload": desc, "roots": end, "positions": [lo, hi], "windows": len(per), "failed": failed,
            "count": tot["count"],
            "hash": f"{tot['hash_b']:016x}{tot['hash_a']:016x}", "tasks": tot["tasks"],
            "valid_roots": tot["valid_roots"], "gpu_s": tot["search_ms"] / 1e3, "wall_s": wall,
            "plexes_per_s": tot["count"] / wall, "per_window": per}


def copy_calibration(nbytes):
    """One device-to-device copy of `nbytes` (hipMemcpy: a copy kernel that reads
    and writes exactly nbytes).  Run under rocprofv3 --pmc FETCH_SIZE or
    WRITE_SIZE, it gives what those counters report per HBM byte, the factor
    that turns a kernel's counter readings into HBM bytes."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    for p in (a, b):
        if hip.hipMalloc(ctypes.byref(p), ctypes.c_size_t(nbytes)):
            raise SystemExit("hipMalloc failed")
    if hip.hipMemset(a, 1, ctypes.c_size_t(nbytes)) or hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3):
        raise SystemExit("hipMemset/hipMemcpy failed")
    if hip.hipDeviceSynchronize():
        raise SystemExit("hipDeviceSynchronize failed")
    hip.hipFree(a)
    hip.hipFree(b)
    log(f"[bench] copied {nbytes} bytes device-to-device")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="cl75k", choices=sorted(WORKLOADS))
    # below the 180 s a silent run may last: a stuck launch ends as KP_ETIMEDOUT
    ap.add_argument("--time-limit", type=float, default=150.0)
    ap.add_argument("--cpu-target-s", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--batch-graphs", type=int, default=48,
                    help="graphs of the many-graph workload for the batched-vs-plain line (0 = skip)")
    ap.add_argument("--batch-tune", default=None, metavar="JSON",
                    help='launch settings for the batch calls instead of tuning runs, e.g. {"1": {"blocks_per_cu": 7}}')
    ap.add_argument("--batch-chunks", type=int, default=1,
                    help="split the batch into this many consecutive kp_enumerate_batch calls (long runs print per call)")
    ap.add_argument("--served-seconds", type=float, default=1.5,
                    help="arrival window of the served stream at --served-rate queries/s (0 = skip)")
    ap.add_argument("--served-rate", type=float, default=200.0)
    ap.add_argument("--served-no-plain", action="store_true",
                    help="served stream only: no one-by-one plain run of the same queries (long streams)")
    ap.add_argument("--served-slice-ms", type=float, default=30.0,
                    help="age at which a running epoch is paused to admit waiting queries")
    ap.add_argument("--no-autotune", action="store_true", help="time with the built-in launch settings")
    ap.add_argument("--tune", default=None, metavar="JSON",
                    help='fixed launch settings for the headline instead of autotuning, e.g. '
                         '{"1": {"blocks_per_cu": 6, "donate_period": 2048}}')
    ap.add_argument("--no-spill-line", action="store_true", help="skip the pause/spill line on the 320k workload")
    ap.add_argument("--no-small-line", action="store_true", help="skip the GPU-vs-CPU line on the small planted workload")
    ap.add_argument("--full-graph", action="store_true",
                    help="also enumerate the whole Chung-Lu 75k graph (every root; many minutes of GPU time)")
    ap.add_argument("--full-graph-only", action="store_true", help="only the whole-graph line, printed as JSON")
    ap.add_argument("--full-graph-workload", default="cl75k-full", choices=["cl75k-full", "cl320k-full"])
    ap.add_argument("--full-graph-range", default=None, metavar="LO:HI",
                    help="root positions of the whole-graph line (default: all)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only copy BYTES device-to-device (counter calibration for the traffic figure) and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    fg_range = tuple(int(x) for x in args.full_graph_range.split(":")) if args.full_graph_range else (0, None)
    if args.full_graph_only:
        print(json.dumps({"full_graph": full_graph_line(args.time_limit, *fg_range, workload=args.full_graph_workload)}))
        return
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(roots would shard across GPUs with one final 32-byte reduction)", file=sys.stderr)
        sys.exit(2)

    import kplex
    desc, fn, window = WORKLOADS[args.workload]
    log(f"[bench] generating {args.workload}: {desc}")
    graph, k, q, thres, truss = getattr(kplex, fn)() if fn else dataset_workload(args.workload)
    log(f"[bench] n={graph.n} m={graph.m} k={k} q={q} truss={truss}; device {kplex.device_info()}")
    t0 = time.perf_counter()
    prep = kplex.Prepared(graph, k, q, thres, truss, time_limit_s=args.time_limit, roots=window)
    prep_s = time.perf_counter() - t0
    tune, tune_s, untuned_ms = None, 0.0, None
    if args.tune:
        tune = {int(a): b for a, b in json.loads(args.tune).items()}
        kplex.set_tune(tune)
        log(f"[bench] fixed launch settings {tune}")
    elif not args.no_autotune:
        r = prep.run()
        r = prep.run()
        untuned_ms = r["search_ms"]
        t0 = time.perf_counter()
        tune = prep.autotune()
        tune_s = time.perf_counter() - t0
        log(f"[bench] autotune {tune_s:.1f} s: {tune} (default settings: search {untuned_ms:.1f} ms)")
    for i in range(args.warmup):
        r = prep.run()
        log(f"[bench] warmup {i}: count={r['count']} search_ms={r['search_ms']:.1f} decompose_ms={r['decompose_ms']:.1f}")
    runs = []
    # kp_run synchronizes its stream before returning, so wall time brackets the GPU work
    t0 = time.perf_counter()
    for i in range(args.steps):
        r = prep.run()
        runs.append(r)
        log(f"[bench] step {i}: count={r['count']} search_ms={r['search_ms']:.1f} decompose_ms={r['decompose_ms']:.1f}")
    elapsed = time.perf_counter() - t0
    step_s = elapsed / max(1, args.steps)
    r0 = runs[0]
    for r in runs[1:]:
        assert (r["count"], r["hash"], r["tasks"]) == (r0["count"], r0["hash"], r0["tasks"]), "non-deterministic result"
    search_s = statistics.mean(r["search_ms"] for r in runs) / 1e3
    achieved = r0["search_alg_bytes"] / search_s / 1e9 if search_s > 0 else 0.0
    out = {
        "metric": "maximal k-plexes/sec + wall s at exact count; batched/served vs plain, p50/p99",
        "value": r0["count"] / step_s,
        "unit": "maximal k-plexes/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": step_s * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None if args.workload not in PUBLISHED else
        (r0["count"] / step_s) / (PUBLISHED[args.workload][0] / PUBLISHED[args.workload][1]),
        "dtype": "u64",
        "data": "synthetic",
        "config": {"workload": desc, "name": args.workload, "k": k, "q": q, "thres": thres, "truss": truss,
                   "n": graph.n, "m_directed": graph.m, "root_window": list(window) if window else None},
        "count": r0["count"],
        "hash": r0["hash"],
        "tasks": r0["tasks"],
        "valid_roots": r0["valid_roots"],
        "wall_s_per_step": step_s,
        "search_ms": search_s * 1e3,
        "decompose_ms": statistics.mean(r["decompose_ms"] for r in runs),
        "preprocess_ms": r0["preprocess_ms"],
        "upload_ms": r0["upload_ms"],
        "prepare_s": prep_s,
        "full_pipeline_s": (r0["preprocess_ms"] + r0["upload_ms"]) / 1e3 + step_s,
        "adjacency_GBps": r0["adj_bytes"] / step_s / 1e9,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": profiled_traffic(args.workload),
                     "kernel": "search_kernel", "alg_bytes_per_launch": r0["search_alg_bytes"]},
        "cpu_baseline": None,
        "autotune": {"settings_by_class_words": tune, "tune_s": tune_s, "default_settings_search_ms": untuned_ms},
    }
    prep.close()
    kplex.set_tune(None)   # the batched and served lines use the built-in settings
    if not args.no_cpu_baseline:
        log(f"[bench] cpu baseline (target {args.cpu_target_s} s)")
        out["cpu_baseline"] = cpu_baseline(graph, k, q, truss, args.cpu_target_s, window)
    if not args.no_small_line:
        out["small"] = small_line()
    if not args.no_spill_line:
        out["spill"] = spill_line(args.time_limit)
    if args.batch_graphs > 0:
        fixed = None
        if args.batch_tune:
            fixed = {int(a): b for a, b in json.loads(args.batch_tune).items()}
        out["batched"] = batched_vs_plain(args.batch_graphs, args.time_limit, args.batch_chunks,
                                          autotune=not args.no_autotune, fixed_tune=fixed)
    if args.served_seconds > 0:
        out["served"] = served_vs_plain(args.served_seconds, args.served_rate, args.time_limit,
                                        slice_ms=args.served_slice_ms, plain=not args.served_no_plain)
    if args.full_graph:
        out["full_graph"] = full_graph_line(args.time_limit, *fg_range, workload=args.full_graph_workload)
    # the other parts of the metric, at the top level
    out["wall_s_at_exact_count"] = out["wall_s_per_step"]
    out["batched_vs_plain"] = out["batched"]["batched_vs_plain"] if "batched" in out else None
    out["served_vs_plain"] = out["served"].get("served_vs_plain") if "served" in out else None
    out["served_p50_ms"] = out["served"]["p50_ms"] if "served" in out else None
    out["served_p99_ms"] = out["served"]["p99_ms"] if "served" in out else None
    print(json.dumps(out))


if __name__ == "__main__":
    main()

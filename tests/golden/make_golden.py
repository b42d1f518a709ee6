"""Writes tests/golden/oracle_results.json: the CPU oracle's results for the
workloads the GPU parity tests compare against (run here, where the oracle
has the time; the GPU box only reads the JSON).

    python tests/golden/make_golden.py [--with-large] [--with-spill]
    python tests/golden/make_golden.py --batch 48        # oracle_batch.json
    python tests/golden/make_golden.py --served          # oracle_served.json
    python tests/golden/make_golden.py --cl75k-chunks    # oracle_cl75k_chunks.json
    python tests/golden/make_golden.py --cl75k-hubs      # oracle_cl75k_hubs.json

Each entry records count, 128-bit hash, canonical task count, valid roots,
preprocessed sizes and the oracle's wall time on this container's cores.
The batch, served and chunk files are written entry by entry, so an
interrupted run resumes where it stopped.
"""
import argparse
import gzip
import json
import os
import shutil
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "maximal-kplex-enumeration_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import kplex  # noqa: E402
import kpo  # noqa: E402

OUT = os.path.join(ROOT, "tests", "golden", "oracle_results.json")
KEYS = ("count", "hash", "tasks", "valid_roots", "pn", "pm", "max_n", "max_l", "alg_bytes", "adj_bytes")


def run(og, k, q, truss, listing=False, window=None):
    t0 = time.time()
    wb, we = window if window else (0, -1)
    r = kpo.enumerate_plexes(og, k, q, truss, threads=os.cpu_count(), listing=listing, pos_begin=wb, pos_end=we)
    d = {x: r[x] for x in KEYS}
    d.update(k=k, q=q, truss=truss, oracle_s=round(time.time() - t0, 3), threads=os.cpu_count(),
             window=list(window) if window else None)
    if listing:
        d["plexes"] = [list(p) for p in r["plexes"]]
    return d


def per_item(og, k, q, truss=0, window=None, threads=None):
    t0 = time.time()
    wb, we = window if window else (0, -1)
    r = kpo.enumerate_plexes(og, k, q, truss, threads=threads or os.cpu_count(), pos_begin=wb, pos_end=we)
    d = {x: r[x] for x in ITEM_KEYS}
    d.update(k=k, q=q, oracle_s=round(time.time() - t0, 3))
    return d


ITEM_KEYS = ("count", "hash", "tasks", "valid_roots", "pn", "pm", "max_n", "max_l")


def load_json(path):
    if os.path.exists(path):
        with open(path) as f:
            return json.load(f)
    return {}


def save_json(path, obj):
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(obj, f, indent=1)
    os.replace(tmp, path)


def batch_fixtures(ngraphs, threads):
    """The first `ngraphs` graphs of the many-graph workload (kplex.batch_graph_specs)."""
    path = os.path.join(ROOT, "tests", "golden", "oracle_batch.json")
    gold = load_json(path)
    items = gold.setdefault("items", {})
    gold["spec"] = "kplex.batch_graph_specs()[:N]: graph i at its own k and q = 3k+8"
    for i, s in enumerate(kplex.batch_graph_specs()[:ngraphs]):
        if str(i) in items:
            continue
        og = from_graph(kplex.graph_from_spec(s))
        items[str(i)] = per_item(og, s["k"], s["q"], threads=threads)
        items[str(i)]["n"] = s["n"]
        save_json(path, gold)
        print("batch", i, items[str(i)], flush=True)


def served_fixtures(threads):
    """kplex.served_fixture_queries() over kplex.served_graph_specs()."""
    path = os.path.join(ROOT, "tests", "golden", "oracle_served.json")
    gold = load_json(path)
    items = gold.setdefault("items", [])
    gold["spec"] = "kplex.served_fixture_queries() over kplex.served_graph_specs(), in that order"
    specs = kplex.served_graph_specs()
    for j, (gi, k, q) in enumerate(kplex.served_fixture_queries()):
        if j < len(items):
            continue
        og = from_graph(kplex.graph_from_spec(specs[gi]))
        d = per_item(og, k, q, threads=threads)
        d.update(graph=gi, n=specs[gi]["n"])
        items.append(d)
        save_json(path, gold)
        print("served", j, d, flush=True)


def cl75k_chunks(threads, step):
    """Root windows of the Chung-Lu 75k workload past [0, 9000), `step` positions each."""
    path = os.path.join(ROOT, "tests", "golden", "oracle_cl75k_chunks.json")
    gold = load_json(path)
    items = gold.setdefault("items", {})
    g, k, q, _, tr = kplex.workload_cl75k()
    og = from_graph(g)
    pn = kpo.preprocess(og, k, q, tr)[0].shape[0]
    end = pn - q + 2
    gold["roots_total"] = end
    b = 9000
    while b < end:
        e = min(end, b + step)
        key = f"{b}_{e}"
        if key not in items:
            items[key] = per_item(og, k, q, tr, window=(b, e), threads=threads)
            save_json(path, gold)
            print("cl75k", key, items[key], flush=True)
        b = e


def cl75k_hubs(threads):
    """Hub roots of the Chung-Lu 75k workload: degeneracy positions [16400,16401)
    (one root, a 267-vertex block in the six-word class) and [16401,16405)."""
    path = os.path.join(ROOT, "tests", "golden", "oracle_cl75k_hubs.json")
    gold = load_json(path)
    items = gold.setdefault("items", {})
    g, k, q, _, tr = kplex.workload_cl75k()
    og = from_graph(g)
    for b, e in ((16400, 16401), (16401, 16405)):
        key = f"{b}_{e}"
        if key not in items:
            items[key] = per_item(og, k, q, tr, window=(b, e), threads=threads)
            save_json(path, gold)
            print("cl75k hubs", key, items[key], flush=True)


def from_graph(g):
    deg, nbr = g.csr()
    return kpo.OGraph.from_csr(deg, nbr)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--with-large", action="store_true",
                    help="also the Chung-Lu 75k root windows [4000,6000) and [0,9000) (minutes of oracle time)")
    ap.add_argument("--with-spill", action="store_true",
                    help="also the Chung-Lu 320k (truss) root window [0,800): 1.96e9 tasks (minutes of oracle time)")
    ap.add_argument("--batch", type=int, default=0, help="oracle results of the first N many-graph members")
    ap.add_argument("--served", action="store_true", help="oracle results of the served fixture queries")
    ap.add_argument("--cl75k-chunks", action="store_true", help="Chung-Lu 75k root windows past 9000")
    ap.add_argument("--cl75k-hubs", action="store_true", help="Chung-Lu 75k hub root windows [16400,16405)")
    ap.add_argument("--chunk", type=int, default=250)
    ap.add_argument("--threads", type=int, default=0)
    args = ap.parse_args()
    if args.batch or args.served or args.cl75k_chunks or args.cl75k_hubs:
        if args.cl75k_hubs:
            cl75k_hubs(args.threads)
        if args.batch:
            batch_fixtures(args.batch, args.threads)
        if args.served:
            served_fixtures(args.threads)
        if args.cl75k_chunks:
            cl75k_chunks(args.threads, args.chunk)
        return
    gold = {}
    if os.path.exists(OUT):
        with open(OUT) as f:
            gold = json.load(f)
    g, k, q, _, tr = kplex.workload_er2000()
    gold["er2000"] = run(from_graph(g), k, q, tr, listing=True)
    tmp = tempfile.mkdtemp()
    try:
        for name, k, q, tr in (("email-euall", 3, 12, 0), ("email-euall", 3, 12, 1), ("soc-epinions", 4, 30, 0)):
            path = os.path.join(tmp, name + ".bin")
            if not os.path.exists(path):
                with gzip.open(os.path.join(ROOT, "tests", "golden", name + ".bin.gz"), "rb") as fi, \
                        open(path, "wb") as fo:
                    shutil.copyfileobj(fi, fo)
            gold[f"{name}_k{k}_q{q}_tr{tr}"] = run(kpo.OGraph.load_bin(path), k, q, tr)
    finally:
        shutil.rmtree(tmp)
    if args.with_large:
        g, k, q, _, tr = kplex.workload_cl75k()
        og = from_graph(g)
        gold["cl75k_4000_6000"] = run(og, k, q, tr, window=(4000, 6000))
        gold["cl75k_0_9000"] = run(og, k, q, tr, window=(0, 9000))
    if args.with_spill:
        g, k, q, _, tr = kplex.workload_cl320k()
        gold["cl320k_0_800"] = run(from_graph(g), k, q, tr, window=(0, 800))
    with open(OUT, "w") as f:
        json.dump(gold, f, indent=1)
    print(json.dumps({n: {x: v[x] for x in ("count", "hash", "tasks", "oracle_s")} for n, v in gold.items()}, indent=1))


if __name__ == "__main__":
    main()

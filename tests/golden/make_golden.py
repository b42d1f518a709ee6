"""Writes tests/golden/oracle_results.json: the CPU oracle's results for the
workloads the GPU parity tests compare against (run here, where the oracle
has the time; the GPU box only reads the JSON).

    python tests/golden/make_golden.py [--with-large]

Each entry records count, 128-bit hash, canonical task count, valid roots,
preprocessed sizes and the oracle's wall time on this container's cores.
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


def from_graph(g):
    deg, nbr = g.csr()
    return kpo.OGraph.from_csr(deg, nbr)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--with-large", action="store_true",
                    help="also the Chung-Lu 75k root windows [4000,6000) and [0,9000) (minutes of oracle time)")
    ap.add_argument("--with-spill", action="store_true",
                    help="also the Chung-Lu 320k (truss) root window [0,800): 1.96e9 tasks (minutes of oracle time)")
    args = ap.parse_args()
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

"""The PlexEnum command line (main.cu:11-62) end to end on a shipped dataset:
the reference's output lines in the reference's order (main.cu:30,54;
inc/graph.h:47; inc/host_funcs.h:115 once per core peel, :871, :873, :1149,
:1156-1158), the README count (README.md:144) and the oracle fixture's
valid blocks, task count and hash."""
import json
import os
import re
import subprocess
import sys

import pytest

import kplex
from conftest import GOLDEN, PKG

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("truss", [0, 1])
def test_plexenum_email_euall(datasets, truss):
    print(f"[gpu-case] PlexEnum email-euall -k 3 -q 12 -tr {truss}", flush=True)
    sys.stdout.flush()
    exe = os.path.join(PKG, "build", "PlexEnum")
    path = datasets["email-euall"]
    with open(os.path.join(GOLDEN, "oracle_results.json")) as f:
        gold = json.load(f)
    want = gold[f"email-euall_k3_q12_tr{truss}"]
    first_pn = gold["email-euall_k3_q12_tr0"]["pn"]   # the first (q-k)-core is the same with the truss on
    g = kplex.Graph.load_bin(path)
    out = subprocess.run([exe, path, "-k", "3", "-q", "12", "-t", "0.25", "-tr", str(truss)],
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    lines = [ln for ln in out.stdout.splitlines() if ln.strip()]
    expect = [rf"file: {re.escape(path)}",
              rf"k=3,q=12,thres=0\.250000, truss: {truss}",
              rf"n={g.n}, m={g.m} \(undirected\)",
              rf"pn = {first_pn}"]
    if truss:
        expect.append(rf"pn = {want['pn']}")
    expect += [r"Start copying graph to GPU\.\.\.\.",
               r"Done copying graph to GPU\.\.\.\.",
               r"Total Time Elapsed: \d+\.\d{6} ms",
               rf"Total Valid Blocks: {want['valid_roots']}, Maximal k-Plexes: 32639016",
               rf"Total tasks generated: {want['tasks']}",
               r"Kernel Launch Successfully",
               rf"hash: {want['hash']}"]
    assert len(lines) == len(expect), out.stdout
    for ln, pat in zip(lines, expect):
        assert re.fullmatch(pat, ln), (ln, pat)
    assert want["count"] == 32639016

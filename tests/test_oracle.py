"""Pins the CPU oracle (oracle/kp_oracle.c) against what fixes its answers:
the README known-answer counts, closed forms, brute force and invariants."""
import json
import math
import os

import numpy as np
import pytest

import kpo
from kp_testutil import brute_force, clique_plus_one, csr, disjoint_cliques, multipartite, random_graph
from conftest import GOLDEN


def ograph(adj):
    deg, nbr = csr(adj)
    return kpo.OGraph.from_csr(deg, nbr)


@pytest.mark.parametrize("parts,t,k", [(3, 3, 2), (3, 4, 3), (4, 3, 2), (2, 5, 3), (6, 3, 2), (5, 4, 3), (4, 4, 4)])
def test_complete_multipartite_closed_form(parts, t, k):
    # K_{t,...,t}: the maximal k-plexes are the choices of k vertices per part,
    # C(t,k)^parts of them, each of size parts*k
    q = parts * k
    r = kpo.enumerate_plexes(ograph(multipartite(parts, t)), k, q, listing=True, threads=2)
    assert r["count"] == math.comb(t, k) ** parts
    assert all(len(p) == q for p in r["plexes"])
    assert all(sum(1 for v in p if v // t == part) == k for p in r["plexes"] for part in range(parts))


def test_disjoint_cliques():
    r = kpo.enumerate_plexes(ograph(disjoint_cliques([5, 6, 3, 7])), 2, 4, listing=True)
    assert r["count"] == 3
    assert sorted(len(p) for p in r["plexes"]) == [5, 6, 7]


@pytest.mark.parametrize("seed", range(12))
def test_brute_force_random(seed):
    n = 13 + seed % 3
    k = 2 + seed % 2
    q = 2 * k + seed % 3
    adj = random_graph(n, 0.72, 1000 + seed)
    want = brute_force(n, adj, k, q)
    got = kpo.enumerate_plexes(ograph(adj), k, q, listing=True)
    assert got["plexes"] == want
    assert got["count"] == len(want)
    assert got["hash"] == kpo.hash_plexes(want)


def test_hash_detects_a_dropped_vertex():
    # a plausible bug (a vertex dropped from one plex) must change the hash
    plexes = [(1, 2, 3, 4), (2, 3, 5, 7)]
    assert kpo.hash_plexes(plexes) != kpo.hash_plexes([(1, 2, 3), (2, 3, 5, 7)])
    assert kpo.hash_plexes(plexes) == kpo.hash_plexes(list(reversed(plexes)))
    assert kpo.hash_plexes([(1, 2, 3)]) != kpo.hash_plexes([(1, 2), (3,)])


def test_mix64_matches_python():
    for z in (0, 1, 0x243F6A8885A308D3, (1 << 64) - 1):
        assert kpo.mix64(z) == kpo.py_mix64(z)


@pytest.mark.parametrize("seed", range(4))
def test_relabel_invariance(seed):
    adj = random_graph(40, 0.45, 77 + seed)
    rng = np.random.default_rng(seed)
    perm = rng.permutation(40)
    adj2 = [set() for _ in range(40)]
    for u in range(40):
        for v in adj[u]:
            adj2[perm[u]].add(int(perm[v]))
    a = kpo.enumerate_plexes(ograph(adj), 2, 6, listing=True)
    b = kpo.enumerate_plexes(ograph(adj2), 2, 6, listing=True)
    inv = np.argsort(perm)
    mapped = sorted(tuple(sorted(int(inv[v]) for v in p)) for p in b["plexes"])
    assert a["count"] == b["count"] > 0
    assert mapped == a["plexes"]


@pytest.mark.parametrize("seed", range(3))
def test_truss_invariance_small(seed):
    adj = random_graph(60, 0.3, 500 + seed)
    a = kpo.enumerate_plexes(ograph(adj), 2, 5, truss=0)
    b = kpo.enumerate_plexes(ograph(adj), 2, 5, truss=1)
    assert (a["count"], a["hash"]) == (b["count"], b["hash"])


def test_mutation_wrong_threshold_is_caught():
    # If the maximality test ignored the left set, roots would report
    # non-maximal plexes: brute force on a graph with earlier extenders catches it.
    adj = multipartite(3, 3)
    for v in range(9):
        adj[v].add(9)
    adj.append(set(range(9)))   # vertex 9 adjacent to all: every plex extends by it
    got = kpo.enumerate_plexes(ograph(adj), 2, 6, listing=True)
    assert got["plexes"] == brute_force(10, adj, 2, 6)
    assert all(9 in p for p in got["plexes"])


def test_readme_kat_email_euall(datasets):
    # README.md:144 -- email-euall, k=3, q=12: 32,639,016 maximal 3-plexes
    og = kpo.OGraph.load_bin(datasets["email-euall"])
    r = kpo.enumerate_plexes(og, 3, 12, threads=os.cpu_count())
    assert r["count"] == 32639016
    assert (r["pn"], r["pm"]) == (3156, 116354)


@pytest.mark.slow
def test_readme_kat_soc_epinions(datasets):
    # README.md:142 -- soc-epinions, k=4, q=30: 13,172,906 maximal 4-plexes
    og = kpo.OGraph.load_bin(datasets["soc-epinions"])
    r = kpo.enumerate_plexes(og, 4, 30, threads=os.cpu_count())
    assert r["count"] == 13172906
    assert (r["pn"], r["pm"], r["valid_roots"]) == (4315, 388146, 267)


def test_golden_fixtures_match_oracle(datasets):
    """The recorded oracle results (tests/golden/oracle_results.json) still hold
    for the cases that finish in seconds."""
    with open(os.path.join(GOLDEN, "oracle_results.json")) as f:
        gold = json.load(f)
    g = gold["er2000"]
    import kplex
    graph, k, q, _, truss = kplex.workload_er2000()
    deg, nbr = graph.csr()
    r = kpo.enumerate_plexes(kpo.OGraph.from_csr(deg, nbr), k, q, truss, listing=True)
    assert (r["count"], r["hash"], r["tasks"]) == (g["count"], g["hash"], g["tasks"])
    assert [list(p) for p in r["plexes"]] == g["plexes"]


def test_k_search_budget_above_256():
    # k = 257: the k-search budget k-1 = 256 must not wrap
    r = kpo.enumerate_plexes(ograph(clique_plus_one(600)), 257, 514, listing=True, threads=4)
    assert r["count"] == 1 and r["plexes"] == [tuple(range(601))]


def test_batch_fixtures_hold_for_the_quick_members():
    # tests/golden/oracle_batch.json (the oracle on the first 48 many-graph
    # members) against a fresh oracle run on the members that take well under
    # a second, so the committed file cannot drift from the oracle
    import kplex
    with open(os.path.join(GOLDEN, "oracle_batch.json")) as f:
        gold = json.load(f)["items"]
    specs = kplex.batch_graph_specs()[:48]
    quick = sorted(range(48), key=lambda i: gold[str(i)]["oracle_s"])[:5]
    for i in quick:
        g = kplex.graph_from_spec(specs[i])
        deg, nbr = g.csr()
        r = kpo.enumerate_plexes(kpo.OGraph.from_csr(deg, nbr), specs[i]["k"], specs[i]["q"], threads=4)
        for x in ("count", "hash", "tasks", "valid_roots", "pn", "pm", "max_n", "max_l"):
            assert r[x] == gold[str(i)][x], (i, x)
    assert sum(v["count"] for v in gold.values()) == 1584447134


def test_served_fixtures_hold_for_the_quick_queries():
    # tests/golden/oracle_served.json against a fresh oracle run on the five
    # quickest fixture queries
    import kplex
    with open(os.path.join(GOLDEN, "oracle_served.json")) as f:
        gold = json.load(f)["items"]
    specs = kplex.served_graph_specs()
    queries = kplex.served_fixture_queries()
    assert [(x["graph"], x["k"], x["q"]) for x in gold] == [tuple(q) for q in queries]
    for i in sorted(range(len(gold)), key=lambda i: gold[i]["oracle_s"])[:5]:
        gi, k, q = queries[i]
        deg, nbr = kplex.graph_from_spec(specs[gi]).csr()
        r = kpo.enumerate_plexes(kpo.OGraph.from_csr(deg, nbr), k, q, threads=4)
        for x in ("count", "hash", "tasks", "valid_roots", "pn", "pm", "max_n", "max_l"):
            assert r[x] == gold[i][x], (i, x)

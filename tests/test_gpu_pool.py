"""GPU tests of the shared pool: many graphs in one set of launches
(kp_enumerate_batch) and pause / resume with pinned-host spill.  Every
result must equal the plain one-graph run and the oracle."""
import json
import os
import sys

import pytest

import kplex
import kpo
from conftest import GOLDEN
from kp_testutil import csr, disjoint_cliques, multipartite, random_graph

pytestmark = pytest.mark.gpu

LIMIT = 90.0
KEYS = ("count", "hash", "tasks", "valid_roots", "pn", "pm", "max_n", "max_l")


def say(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def same(a, b, keys=KEYS):
    for x in keys:
        assert a[x] == b[x], (x, a[x], b[x])


def oracle(graph, k, q, truss=0, listing=False):
    deg, nbr = graph.csr()
    return kpo.enumerate_plexes(kpo.OGraph.from_csr(deg, nbr), k, q, truss, listing=listing)


def mixed_graphs():
    """Uneven graphs and parameters: closed forms, random, planted groups."""
    out = []
    out.append((kplex.Graph.from_csr(*csr(multipartite(4, 3))), 2, 8, 0))
    out.append((kplex.Graph.from_csr(*csr(disjoint_cliques([5, 6, 70, 130]))), 2, 4, 0))
    for s in range(4):
        out.append((kplex.Graph.from_csr(*csr(random_graph(120 + 20 * s, 0.3, 500 + s))), 2 + s % 2, 5 + s, 0))
    out.append((kplex.Graph.er_planted(1200, 0.01, 6, 30, 3, 9), 3, 14, 0))
    out.append((kplex.Graph.er_planted(1500, 0.004, 2, 140, 1, 180), 2, 60, 0))
    out.append((kplex.Graph.er_planted(1500, 0.012, 8, 28, 3, 21), 3, 16, 1))
    out.append((kplex.Graph.from_csr(*csr(multipartite(2, 2))), 2, 8, 0))   # no valid root
    return out


def test_batch_equals_plain_and_oracle():
    cases = mixed_graphs()
    say(f"batch of {len(cases)}")
    rs = kplex.enumerate_batch([g for g, *_ in cases], [(k, q, tr) for _, k, q, tr in cases])
    for (g, k, q, tr), r in zip(cases, rs):
        say(f"  plain n={g.n} k={k} q={q}")
        p = kplex.enumerate_plexes(g, k, q, truss=tr, time_limit_s=LIMIT)
        same(r, p)
        o = oracle(g, k, q, tr)
        same(r, o)


def test_batch_lists_split_per_graph():
    cases = mixed_graphs()[:7]
    say("batch lists")
    rs = kplex.enumerate_batch([g for g, *_ in cases], [(k, q, tr) for _, k, q, tr in cases], listing=True)
    for (g, k, q, tr), r in zip(cases, rs):
        o = oracle(g, k, q, tr, listing=True)
        assert [tuple(p) for p in r["plexes"]] == o["plexes"]
        assert r["nlisted"] == r["count"]


def test_batch_of_one_and_same_graph_twice():
    g = kplex.Graph.er_planted(1200, 0.01, 6, 30, 3, 9)
    say("batch duplicates")
    rs = kplex.enumerate_batch([g, g, g], [(3, 14), (3, 14), (2, 10)])
    same(rs[0], rs[1])
    same(rs[2], kplex.enumerate_plexes(g, 2, 10, time_limit_s=LIMIT))


def test_batch_many_chung_lu_graphs():
    # small members of the many-graph workload family: uneven sizes and k in {2,3,4}
    specs = kplex.batch_graph_specs(count=24, nmax=1500, seed=11)
    graphs = [kplex.graph_from_spec(s) for s in specs]
    say(f"batch chung-lu x{len(graphs)}")
    rs = kplex.enumerate_batch(graphs, [(s["k"], s["q"]) for s in specs])
    tot = 0
    for s, g, r in zip(specs, graphs, rs):
        p = kplex.enumerate_plexes(g, s["k"], s["q"], time_limit_s=LIMIT)
        same(r, p)
        tot += r["count"]
    assert tot > 0


def batch_gold():
    with open(os.path.join(GOLDEN, "oracle_batch.json")) as f:
        return json.load(f)["items"]


def test_batch_with_more_roots_than_resident_waves():
    # The first 48 graphs of the many-graph workload: thousands of roots over
    # several size classes whose concurrent grids do not all fit on the chip
    # at once.  A wave must count as busy only once it runs, or idle resident
    # waves of one class wait on waves of another that need their slots.
    # Every graph's result must equal the oracle's (tests/golden/oracle_batch.json).
    gold = batch_gold()
    specs = kplex.batch_graph_specs()[:48]
    graphs = [kplex.graph_from_spec(s) for s in specs]
    say(f"batch of the first {len(graphs)} many-graph members")
    rs = kplex.enumerate_batch(graphs, [dict(k=s["k"], q=s["q"], time_limit_s=LIMIT) for s in specs])
    for i, r in enumerate(rs):
        same(r, gold[str(i)])
    assert sum(r["count"] for r in rs) == 1584447134
    for i in (0, 11, 47):
        say(f"  plain member {i}")
        same(kplex.enumerate_plexes(graphs[i], specs[i]["k"], specs[i]["q"], time_limit_s=LIMIT), gold[str(i)])


def mid_graph():
    """12 planted 40-vertex groups with 8 removed edges per vertex: k=3, q=18
    gives 6,524 maximal 3-plexes over 2.5 million Branch tasks."""
    return kplex.Graph.er_planted(3000, 0.004, 12, 40, 8, 7), 3, 18


def wide_graph():
    """Planted 70-vertex groups (the two-word class): k=2, q=20, 8 million tasks."""
    return kplex.Graph.er_planted(2000, 0.004, 6, 70, 14, 5), 2, 20


@pytest.mark.parametrize("pause_every,cap,thres", [(20000, 64, 0.25), (200000, 1 << 20, 0.25), (50000, 8, 1.0)])
def test_pause_resume_identical(pause_every, cap, thres):
    g, k, q = mid_graph()
    say(f"pause every {pause_every} cap {cap} thres {thres}")
    base = kplex.enumerate_plexes(g, k, q, time_limit_s=LIMIT)
    r = kplex.enumerate_plexes(g, k, q, time_limit_s=LIMIT, thres=thres, queue_capacity=cap,
                               pause_every=pause_every)
    same(r, base)
    assert r["pauses"] > 0
    if cap * thres < 1000:
        assert r["spilled"] > 0
    else:
        assert r["spilled"] == 0


def test_pause_resume_lists():
    g, k, q = mid_graph()
    say("pause lists")
    r = kplex.enumerate_plexes(g, k, q, listing=True, time_limit_s=LIMIT, queue_capacity=16, pause_every=30000)
    o = oracle(g, k, q, listing=True)
    same(r, o)
    assert [tuple(p) for p in r["plexes"]] == o["plexes"]
    assert r["pauses"] > 0 and r["spilled"] > 0


def test_batch_with_pauses():
    cases = mixed_graphs()
    for g, k, q in (mid_graph(), wide_graph()):
        cases.append((g, k, q, 0))
    say("batch + pauses")
    ps = [dict(k=k, q=q, truss=tr, pause_every=100000, queue_capacity=32) for _, k, q, tr in cases]
    rs = kplex.enumerate_batch([g for g, *_ in cases], ps)
    base = kplex.enumerate_batch([g for g, *_ in cases], [(k, q, tr) for _, k, q, tr in cases])
    for r, b in zip(rs, base):
        same(r, b)
    assert rs[0]["pauses"] > 0 and rs[0]["spilled"] > 0
    same(rs[-2], oracle(*mid_graph()))


def test_autotune_keeps_results():
    g, k, q = mid_graph()
    say("autotune")
    p = kplex.Prepared(g, k, q, time_limit_s=LIMIT)
    base = p.run()
    try:
        tune = p.autotune()
        assert tune and all(v["class_ms"] > 0 for v in tune.values())
        same(p.run(), base)
        # a batch after tuning uses the same settings and still agrees
        rs = kplex.enumerate_batch([g, g], [(k, q), (2, 12)])
        same(rs[0], base)
        kplex.set_tune({1: dict(blocks_per_cu=1, donate_period=32, ring_frames=8, queue_slots=16)})
        same(p.run(), base)
    finally:
        kplex.set_tune(None)
        p.close()


def test_narrowing_hub_roots_equal_the_unnarrowed_run():
    # Hub roots of the Chung-Lu 75k workload (degeneracy positions [16400,16405):
    # blocks of 250-270 vertices, four- and six-word classes, 3.1e9 Branch tasks).
    # Frames whose live set fits 64 vertices become derived roots of the one-word
    # class; count, hash and the task count (the tree) must not change, and both
    # runs must equal the oracle (tests/golden/oracle_cl75k_hubs.json).
    g, k, q, _, tr = kplex.workload_cl75k()
    out = {}
    for narrow in (True, False):
        say(f"cl75k hub window [16400,16405) narrow={narrow}")
        p = kplex.Prepared(g, k, q, truss=tr, time_limit_s=LIMIT, roots=(16400, 16405), narrow=narrow)
        out[narrow] = p.run()
        p.close()
    assert out[True]["narrowed"] > 0 and out[False]["narrowed"] == 0
    same(out[True], out[False])
    # and both against the oracle: [16400,16401) + [16401,16405) from its fixtures
    with open(os.path.join(GOLDEN, "oracle_cl75k_hubs.json")) as f:
        items = json.load(f)["items"]
    parts = [items["16400_16401"], items["16401_16405"]]
    m = (1 << 64) - 1
    hb = sum(int(x["hash"][:16], 16) for x in parts) & m
    ha = sum(int(x["hash"][16:], 16) for x in parts) & m
    assert out[True]["count"] == sum(x["count"] for x in parts)
    assert out[True]["tasks"] == sum(x["tasks"] for x in parts)
    assert out[True]["hash"] == f"{hb:016x}{ha:016x}"


def test_hub_root_alone_repeated_against_the_oracle():
    # One hub root of the Chung-Lu 75k workload (degeneracy position 16400: a
    # 267-vertex block in the six-word class, 4.9e8 Branch tasks) alone on the
    # chip: its class keeps all 2,048 waves, nearly all hungry at first, so its
    # subtrees move through the donation queue between waves on different
    # XCDs hundreds of thousands of times.  Four runs, each against the
    # oracle's result for that root (tests/golden/oracle_cl75k_hubs.json).
    path = os.path.join(GOLDEN, "oracle_cl75k_hubs.json")
    with open(path) as f:
        want = json.load(f)["items"]["16400_16401"]
    g, k, q, _, tr = kplex.workload_cl75k()
    p = kplex.Prepared(g, k, q, truss=tr, time_limit_s=LIMIT, roots=(16400, 16401))
    try:
        for i in range(4):
            say(f"cl75k hub root 16400, run {i}")
            same(p.run(), want)
    finally:
        p.close()

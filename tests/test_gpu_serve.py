"""GPU tests of the served mode (kp_server_*): queries admitted into running
epochs must give the same results as the plain one-graph runs, alone or in
any mix, under slot and memory limits."""
import json
import os
import sys
import time

import pytest

import kplex
from conftest import GOLDEN

pytestmark = pytest.mark.gpu

LIMIT = 90.0
KEYS = ("count", "hash", "tasks", "valid_roots", "pn", "pm", "max_n", "max_l")


def say(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def same(a, b, keys=KEYS):
    for x in keys:
        assert a[x] == b[x], (x, a[x], b[x])


def graphs_and_queries():
    graphs = [kplex.Graph.er_planted(3000, 0.004, 12, 40, 8, 7),
              kplex.Graph.er_planted(2000, 0.004, 6, 70, 14, 5),
              kplex.Graph.er_planted(1200, 0.01, 6, 30, 3, 9)]
    specs = kplex.batch_graph_specs(count=6, nmax=3000, seed=4, kset=(3, 4, 5), qfun=lambda k: 5 * k + 10)
    graphs += [kplex.graph_from_spec(s) for s in specs]
    queries = [(0, 3, 18), (1, 2, 20), (2, 3, 14), (0, 2, 12), (2, 2, 10), (1, 3, 22)]
    queries += [(3 + i, s["k"], s["q"]) for i, s in enumerate(specs)]
    queries += [(0, 3, 18), (3, specs[0]["k"], specs[0]["q"])]   # repeats share the resident graph
    return graphs, queries


def drain(srv, want, timeout_s=120.0):
    got = {}
    t0 = time.time()
    while len(got) < want:
        for t, r in srv.poll():
            got[t] = r
        assert time.time() - t0 < timeout_s, f"server did not finish: {len(got)}/{want}"
        time.sleep(0.0002)
    return got


def plain(graphs, q):
    g, k, qq = graphs[q[0]], q[1], q[2]
    return kplex.enumerate_plexes(g, k, qq, time_limit_s=LIMIT)


def test_server_staggered_arrivals_equal_plain():
    graphs, queries = graphs_and_queries()
    srv = kplex.Server(max_inflight=6, slice_ms=0.5)
    ids = [srv.add_graph(g) for g in graphs]
    tickets = {}
    say(f"server: {len(queries)} queries, staggered")
    got = {}
    for i, (gi, k, q) in enumerate(queries):
        tickets[srv.submit(ids[gi], k, q)] = i
        # let the running epoch make progress so that later queries arrive mid-epoch
        for _ in range(3):
            for t, r in srv.poll():
                got[t] = r
            time.sleep(0.001)
    t0 = time.time()
    while len(got) < len(queries):
        for t, r in srv.poll():
            got[t] = r
        assert time.time() - t0 < 120, f"server did not finish: {len(got)}/{len(queries)}"
        time.sleep(0.0002)
    srv.close()
    # queries arrived while epochs ran: some epochs were paused to admit them
    assert max(r["pauses"] for r in got.values()) > 0
    for t, i in tickets.items():
        say(f"  plain {queries[i]}")
        same(got[t], plain(graphs, queries[i]))


def test_server_alone_and_in_mix():
    graphs, queries = graphs_and_queries()
    target = queries[11]   # the last Chung-Lu member
    say(f"alone {target}")
    srv = kplex.Server()
    gid = srv.add_graph(graphs[target[0]])
    t = srv.submit(gid, target[1], target[2])
    alone = drain(srv, 1)[t]
    srv.close()
    say("in a mix")
    srv = kplex.Server(max_inflight=4, slice_ms=0.2)
    ids = [srv.add_graph(g) for g in graphs]
    mix = [srv.submit(ids[gi], k, q) for gi, k, q in queries[:5]]
    tt = srv.submit(ids[target[0]], target[1], target[2])
    mix += [srv.submit(ids[gi], k, q) for gi, k, q in queries[5:9]]
    got = drain(srv, len(mix) + 1)
    srv.close()
    same(got[tt], alone)
    assert got[tt]["count"] > 0


def test_server_memory_budget_admits_in_turn():
    graphs, queries = graphs_and_queries()
    # a budget that holds one or two subgraph arenas at a time
    srv = kplex.Server(max_inflight=8, mem_budget_bytes=4 << 20, slice_ms=0.5)
    ids = [srv.add_graph(g) for g in graphs]
    say("small arena budget")
    sub = queries[:6]
    ts = [srv.submit(ids[gi], k, q) for gi, k, q in sub]
    got = drain(srv, len(ts))
    srv.close()
    for t, qy in zip(ts, sub):
        same(got[t], plain(graphs, qy))


def test_server_query_larger_than_budget_is_refused():
    graphs, queries = graphs_and_queries()
    srv = kplex.Server(max_inflight=2, mem_budget_bytes=4096)
    gid = srv.add_graph(graphs[1])
    t = srv.submit(gid, 2, 20)
    say("arena budget smaller than one query")
    got = drain(srv, 1)
    assert got[t]["status"] == kplex.KP_ENOMEM and got[t]["count"] == 0
    assert srv.outstanding() == 0
    srv.close()


def test_server_failed_query_frees_its_slot_and_others_run():
    # a clique of 1,100 vertices: its first root's block exceeds KP_MAX_BLOCK
    # (KP_ERANGE).  Staged in one group with valid queries, it must leave with
    # an error result while the others run to their plain results, and every
    # slot must come back (the server drains).
    graphs, queries = graphs_and_queries()
    n = 1100
    big = kplex.Graph.from_csr([n - 1] * n, [v for u in range(n) for v in range(n) if v != u])
    srv = kplex.Server(max_inflight=3, slice_ms=0.5)
    ids = [srv.add_graph(g) for g in graphs]
    bid = srv.add_graph(big)
    say("failing query in a group")
    tb = srv.submit(bid, 2, 4)
    ta = srv.submit(ids[2], 3, 14)
    tc = srv.submit(ids[0], 3, 18)
    got = drain(srv, 3)
    assert got[tb]["status"] == kplex.KP_ERANGE and got[tb]["count"] == 0
    for t, qy in ((ta, (2, 3, 14)), (tc, (0, 3, 18))):
        assert got[t]["status"] == kplex.KP_OK
        same(got[t], plain(graphs, qy))
    # the slots are free again: a later query runs as well
    td = srv.submit(ids[2], 2, 10)
    got = drain(srv, 1)
    same(got[td], plain(graphs, (2, 2, 10)))
    assert srv.outstanding() == 0
    srv.close()


def test_server_rejects_lists_and_bad_params():
    srv = kplex.Server()
    gid = srv.add_graph(kplex.Graph.er_planted(300, 0.05, 1, 10, 0, 1))
    with pytest.raises(kplex.KplexError):
        srv.submit(gid, 3, 5)
    srv.close()


def test_server_served_workload_queries_equal_the_oracle():
    # The served workload's fixture queries (the first 32 of its Poisson
    # stream and the 8 largest graphs, n up to 198k, at k = 3, 4, 5) submitted
    # a few at a time while epochs run: each result must equal the oracle's
    # (tests/golden/oracle_served.json).
    with open(os.path.join(GOLDEN, "oracle_served.json")) as f:
        gold = json.load(f)["items"]
    specs = kplex.served_graph_specs()
    queries = kplex.served_fixture_queries()
    assert len(gold) == len(queries) == 40
    needed = sorted({gi for gi, _, _ in queries})
    graphs = {gi: kplex.graph_from_spec(specs[gi]) for gi in needed}
    srv = kplex.Server(max_inflight=16, slice_ms=1.0)
    ids = {gi: srv.add_graph(graphs[gi]) for gi in needed}
    say(f"served fixture queries: {len(queries)} over {len(needed)} graphs")
    tickets, got = {}, {}
    for i, (gi, k, q) in enumerate(queries):
        assert (gold[i]["graph"], gold[i]["k"], gold[i]["q"]) == (gi, k, q)
        tickets[srv.submit(ids[gi], k, q)] = i
        if i % 4 == 3:
            for t, r in srv.poll():
                got[t] = r
            time.sleep(0.002)
    t0 = time.time()
    while len(got) < len(queries):
        for t, r in srv.poll():
            got[t] = r
        assert time.time() - t0 < 240, f"server did not finish: {len(got)}/{len(queries)}"
        time.sleep(0.0005)
    srv.close()
    for t, i in tickets.items():
        assert got[t]["status"] == kplex.KP_OK, (i, got[t]["status"])
        same(got[t], gold[i])

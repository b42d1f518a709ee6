"""GPU test of kp_enumerate_noalloc (the entry point that allocates no device
memory): called through the C ABI, it must give kp_enumerate's results, call
after call on the same static scratch array, and refuse what does not fit."""
import ctypes
import sys

import pytest

import kplex
import kpo
from kp_testutil import csr, disjoint_cliques, random_graph

pytestmark = pytest.mark.gpu

KEYS = ("count", "hash", "tasks", "valid_roots", "pn", "pm", "max_n", "max_l")


def say(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def noalloc(graph, k, q, truss=0, roots=None, listing=False):
    lib = kplex.load()
    p = kplex.make_params(k, q, truss=truss, roots=roots, time_limit_s=60.0,
                          flags=kplex.KP_FLAG_LIST if listing else 0)
    r = kplex.KpResult()
    rc = lib.kp_enumerate_noalloc(ctypes.byref(graph.g), ctypes.byref(p), ctypes.byref(r))
    return rc, kplex._result_dict(r, False)


def same(a, b, keys=KEYS):
    for x in keys:
        assert a[x] == b[x], (x, a[x], b[x])


def test_noalloc_equals_enumerate_call_after_call():
    # two calls in a row on the shared scratch array, the second on a smaller
    # graph, then the first again: each equals kp_enumerate (and the oracle)
    big, k1, q1, _, _ = kplex.workload_er2000()
    small = kplex.Graph.from_csr(*csr(random_graph(160, 0.28, 301)))
    for name, g, k, q in (("er2000", big, k1, q1), ("random160", small, 2, 5), ("er2000 again", big, k1, q1)):
        say(f"noalloc {name} n={g.n} k={k} q={q}")
        rc, r = noalloc(g, k, q)
        assert rc == kplex.KP_OK, kplex.load().kp_last_error()
        e = kplex.enumerate_plexes(g, k, q, time_limit_s=60.0)
        same(r, e)
        assert r["count"] > 0
    deg, nbr = small.csr()
    same(noalloc(small, 2, 5)[1], kpo.enumerate_plexes(kpo.OGraph.from_csr(deg, nbr), 2, 5))


def test_noalloc_wider_blocks_and_windows():
    # blocks of 70 and 130 vertices run at two and three words; a root window
    g = kplex.Graph.from_csr(*csr(disjoint_cliques([5, 6, 70, 130])))
    say("noalloc wide blocks")
    rc, r = noalloc(g, 2, 4)
    assert rc == kplex.KP_OK
    same(r, kplex.enumerate_plexes(g, 2, 4, time_limit_s=60.0))
    g2 = kplex.Graph.er_planted(1200, 0.01, 6, 30, 3, 9)
    say("noalloc window")
    rc, r = noalloc(g2, 3, 14, roots=(0, 400))
    assert rc == kplex.KP_OK
    same(r, kplex.enumerate_plexes(g2, 3, 14, roots=(0, 400), time_limit_s=60.0))


def test_noalloc_refusals():
    g = kplex.Graph.from_csr(*csr(disjoint_cliques([300])))
    say("noalloc block over 256")
    rc, _ = noalloc(g, 2, 4)
    assert rc == kplex.KP_ERANGE
    say("noalloc lists refused")
    rc, _ = noalloc(g, 2, 4, listing=True)
    assert rc == kplex.KP_EINVAL
    # still usable afterwards
    h = kplex.Graph.from_csr(*csr(random_graph(160, 0.28, 302)))
    rc, r = noalloc(h, 2, 5)
    assert rc == kplex.KP_OK
    same(r, kplex.enumerate_plexes(h, 2, 5, time_limit_s=60.0))


def test_noalloc_graph_larger_than_the_array():
    # about 9 million undirected edges: the graph copy alone exceeds the array
    g = kplex.Graph.chung_lu(300000, 2.5, 9_000_000, 2000.0, 0, 0, 0, 0.0, 11)
    say(f"noalloc too large m={g.m}")
    rc, _ = noalloc(g, 2, 4)
    assert rc == kplex.KP_ENOMEM
    assert b"scratch" in kplex.load().kp_last_error()

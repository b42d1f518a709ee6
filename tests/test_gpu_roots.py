"""Root-by-root parity of the decomposition and the subgraph build: for each
sampled root, the GPU's block, hop, left set and pair classes
(kp_root_subgraph) must equal the oracle's (kpo_root_block), the restatement
of the reference's decompose and buildCommonMtx (src/device_funcs.cu:28-298,
:300-460, :1727-1853)."""
import sys

import numpy as np
import pytest

import kplex
import kpo

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def expected_pairs(cls):
    # the class byte (0..5) as the three facts the search tests: linked (odd
    # classes), class >= UNLINK2EQUAL (2), class >= UNLINK2MORE (4)
    return (cls & 1) | ((cls >= 2).astype(np.uint8) << 1) | ((cls >= 4).astype(np.uint8) << 2)


def check_roots(graph, k, q, truss, positions, name):
    deg, nbr = graph.csr()
    rb = kpo.RootBlocks(kpo.OGraph.from_csr(deg, nbr), k, q, truss)
    prep = kplex.Prepared(graph, k, q, truss=truss, time_limit_s=60.0)
    valid = widest = 0
    try:
        for pos in positions(rb.pn):
            g = prep.root_subgraph(pos)
            o = rb.block(pos)
            assert (g is None) == (o is None), (name, pos)
            if g is None:
                continue
            valid += 1
            widest = max(widest, len(o["block"]))
            assert g["block"].tolist() == o["block"].tolist(), (name, pos, "block")
            assert g["hop"] == o["hop"], (name, pos, "hop")
            # the left set is used as a set (the maximality test); its 2-hop
            # part is in claim order on the GPU and ascending in the oracle
            assert sorted(g["left"].tolist()) == sorted(o["left"].tolist()), (name, pos, "left")
            assert len(set(g["left"].tolist())) == len(g["left"]), (name, pos, "left repeats")
            assert np.array_equal(g["pairs"], expected_pairs(o["cls"])), (name, pos, "pairs")
    finally:
        prep.close()
    say(f"{name}: {valid} valid roots compared, widest block {widest}")
    return valid, widest


def test_roots_small_planted_workload_every_position():
    g, k, q, _, tr = kplex.workload_er2000()
    valid, _ = check_roots(g, k, q, tr, lambda pn: range(pn), "er2000")
    assert valid == 420


@pytest.mark.parametrize("truss", [0, 1])
def test_roots_email_euall(datasets, truss):
    g = kplex.Graph.load_bin(datasets["email-euall"])
    valid, widest = check_roots(g, 3, 12, truss, lambda pn: range(0, pn, 5 if truss else 9), f"email-euall tr{truss}")
    assert valid > 50 and widest > 64


def test_roots_soc_epinions(datasets):
    g = kplex.Graph.load_bin(datasets["soc-epinions"])
    valid, widest = check_roots(g, 4, 30, 0, lambda pn: range(pn - 1, -1, -13), "soc-epinions")
    assert valid > 10 and widest > 128


def test_roots_wide_blocks_and_k2():
    # planted groups of 200 and 330 vertices: blocks of four and six words;
    # k = 2 takes the raw-adjacency class of far pairs
    for gsize, k, q in ((200, 3, 60), (330, 2, 120)):
        g = kplex.Graph.er_planted(1500, 0.004, 2, gsize, 1, 40 + gsize)
        _, widest = check_roots(g, k, q, 0, lambda pn: range(0, pn, 3), f"planted {gsize}")
        assert widest > gsize - 20

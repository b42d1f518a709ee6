"""Host-side checks of libkplex that need no GPU: the C ABI exports, the
binary CSR loader, the generators and the host preprocessing (compared with
the oracle's independent implementation)."""
import ctypes
import os
import re
import subprocess
import sys

import numpy as np
import pytest

import kplex
import kpo
from conftest import ROOT
from kp_testutil import csr, random_graph


def header_functions():
    with open(os.path.join(ROOT, "include", "kplex.h")) as f:
        txt = f.read()
    return sorted(set(re.findall(r"^\s*(?:int|void|const char|uint64_t|size_t)\s*\*?\s*(kp_\w+)\s*\(", txt, re.M)))


def test_library_exports_every_declared_symbol():
    lib = kplex.load()
    names = header_functions()
    assert len(names) >= 20
    for name in names:
        assert hasattr(lib, name), name
    assert set(names) == set(kplex.SIGNATURES), set(names) ^ set(kplex.SIGNATURES)


def test_cli_usage_errors():
    exe = kplex.CLI_PATH
    r = subprocess.run([exe, "x.bin", "-k", "2"], capture_output=True, text=True)
    assert r.returncode == 1 and "usage" in r.stderr
    r = subprocess.run([exe, "x.bin", "-k", "2", "-q", "4", "-zz", "1"], capture_output=True, text=True)
    assert r.returncode == 1 and "usage" in r.stderr
    r = subprocess.run([exe, "/nonexistent.bin", "-k", "2", "-q", "4"], capture_output=True, text=True)
    assert r.returncode == 1 and "file: /nonexistent.bin" in r.stdout


def test_load_save_roundtrip(tmp_path, datasets):
    g = kplex.Graph.load_bin(datasets["soc-epinions"])
    assert (g.n, g.m) == (75879, 811480)
    assert g.validate() == 0
    p = tmp_path / "copy.bin"
    g.save_bin(p)
    with open(datasets["soc-epinions"], "rb") as a, open(p, "rb") as b:
        assert a.read() == b.read()


def test_loader_errors(tmp_path):
    with pytest.raises(kplex.KplexError) as e:
        kplex.Graph.load_bin(tmp_path / "missing.bin")
    assert e.value.code == kplex.KP_ENOENT
    bad = tmp_path / "short.bin"
    np.array([3, 4, 1, 1], dtype=np.uint32).tofile(bad)   # truncated
    with pytest.raises(kplex.KplexError) as e:
        kplex.Graph.load_bin(bad)
    assert e.value.code in (kplex.KP_EIO, kplex.KP_EINVAL)


def test_validate_rejects_asymmetric_and_unsorted():
    g = kplex.Graph.from_csr([1, 0], [1])          # 0->1 without 1->0
    assert g.validate() == kplex.KP_EINVAL
    g = kplex.Graph.from_csr([2, 1, 1], [2, 1, 0, 0])   # unsorted list of 0
    assert g.validate() == kplex.KP_EINVAL
    g = kplex.Graph.from_edges(4, [(0, 1), (1, 2), (2, 3)])
    assert g.validate() == 0


def test_generators_deterministic_and_simple():
    a = kplex.Graph.er_planted(500, 0.02, 5, 20, 2, 7)
    b = kplex.Graph.er_planted(500, 0.02, 5, 20, 2, 7)
    assert a.validate() == 0
    da, na = a.csr()
    db, nb = b.csr()
    assert np.array_equal(da, db) and np.array_equal(na, nb)
    c = kplex.Graph.chung_lu(5000, 2.1, 20000, 300.0, 10, 20, 30, 0.9, 3)
    assert c.validate() == 0
    deg, _ = c.csr()
    assert 0.8 * 2 * 20000 < deg.sum() < 1.6 * 2 * 20000 + 10 * 30 * 29


def test_planted_groups_have_capped_missing_edges():
    g = kplex.Graph.er_planted(300, 0.0, 3, 40, 2, 11)
    deg, nbr = g.csr()
    inside = deg[deg > 0]
    assert len(inside) == 120
    assert inside.min() >= 39 - 2


def compare_prep(graph, k, q, truss):
    deg, nbr = graph.csr()
    a = kplex.preprocess(graph, k, q, truss)
    b = kpo.preprocess(kpo.OGraph.from_csr(deg, nbr), k, q, truss)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    return len(a[0])


@pytest.mark.parametrize("truss", [0, 1])
def test_preprocess_matches_oracle_datasets(datasets, truss):
    g = kplex.Graph.load_bin(datasets["email-euall"])
    pn = compare_prep(g, 3, 12, truss)
    assert pn == (3156 if not truss else 1086)
    g = kplex.Graph.load_bin(datasets["soc-epinions"])
    assert compare_prep(g, 4, 30, truss) == (4315 if not truss else 622)


def textbook_core_numbers(deg, nbr):
    """Core number of every vertex by the bin-sort peel of Batagelj and
    Zaversnik (2003): process vertices in nondecreasing current degree and
    lower each unprocessed neighbour's degree by one.  The core numbers are
    unique, so any correct peel gives the same numbers."""
    n = len(deg)
    off = np.concatenate(([0], np.cumsum(deg))).astype(np.int64)
    d = [int(x) for x in deg]
    md = max(d) if n else 0
    bins = [0] * (md + 1)
    for x in d:
        bins[x] += 1
    start = 0
    for k in range(md + 1):
        bins[k], start = start, start + bins[k]
    pos, vert = [0] * n, [0] * n
    for v in range(n):
        pos[v] = bins[d[v]]
        vert[pos[v]] = v
        bins[d[v]] += 1
    for k in range(md, 0, -1):
        bins[k] = bins[k - 1]
    if md >= 0 and bins:
        bins[0] = 0
    for i in range(n):
        v = vert[i]
        for u in nbr[off[v]:off[v + 1]].tolist():
            if d[u] > d[v]:
                du, pu = d[u], pos[u]
                pw = bins[du]
                w = vert[pw]
                if u != w:
                    pos[u], vert[pu] = pw, w
                    pos[w], vert[pw] = pu, u
                bins[du] += 1
                d[u] -= 1
    return np.array(d, dtype=np.int64)


def check_degeneracy_order(deg, nbr, dseq):
    """dseq is a minimum-degree peel of the graph (deg, nbr): every popped
    vertex has the least number of unpopped neighbours among the unpopped
    vertices, and the running maximum of those pop degrees is each vertex's
    core number (computed independently above)."""
    n = len(deg)
    assert sorted(dseq.tolist()) == list(range(n))
    off = np.concatenate(([0], np.cumsum(deg))).astype(np.int64)
    rem = deg.astype(np.int64).copy()
    alive = np.ones(n, dtype=bool)
    core = np.zeros(n, dtype=np.int64)
    run = 0
    for u in dseq.tolist():
        assert rem[u] == rem[alive].min(), u
        run = max(run, int(rem[u]))
        core[u] = run
        alive[u] = False
        rem[nbr[off[u]:off[u + 1]]] -= 1
    assert np.array_equal(core, textbook_core_numbers(deg, nbr))


@pytest.mark.parametrize("name,k,q,truss", [("email-euall", 3, 12, 0), ("email-euall", 3, 12, 1),
                                            ("soc-epinions", 4, 30, 0), ("soc-epinions", 4, 30, 1)])
def test_degeneracy_order_is_a_min_degree_peel(datasets, name, k, q, truss):
    # Both implementations (the library's host preprocessing and the oracle)
    # must produce a minimum-degree peel whose running maximum gives the
    # textbook core numbers.
    g = kplex.Graph.load_bin(datasets[name])
    deg, nbr, _, dseq = kplex.preprocess(g, k, q, truss)
    check_degeneracy_order(deg, nbr, dseq)
    d0, n0 = g.csr()
    odeg, onbr, _, odseq = kpo.preprocess(kpo.OGraph.from_csr(d0, n0), k, q, truss)
    check_degeneracy_order(odeg, onbr, odseq)


def test_min_degree_peel_check_catches_a_swapped_order():
    # the check above fails for an order that is not a minimum-degree peel
    adj = random_graph(120, 0.1, 3)
    deg, nbr = csr(adj)
    deg, nbr = np.array(deg, dtype=np.uint32), np.array(nbr, dtype=np.uint32)
    g = kplex.Graph.from_csr(deg, nbr)
    pdeg, pnbr, _, dseq = kplex.preprocess(g, 1, 2)
    check_degeneracy_order(pdeg, pnbr, dseq)
    bad = dseq.copy()
    bad[0], bad[-1] = bad[-1], bad[0]
    with pytest.raises(AssertionError):
        check_degeneracy_order(pdeg, pnbr, bad)


@pytest.mark.parametrize("impl", ["library", "oracle"])
def test_degeneracy_tie_order_hand_derived(impl):
    # The bucket heap of inc/degeneracy.h inserts at the head of a bucket and
    # moves a decremented vertex to the head of its new bucket; pop takes the
    # head of the lowest bucket.  Worked by hand:
    #   cycle 0-1-2-3-4-0 (all degree 2): inserts leave bucket 2 = [4,3,2,1,0];
    #   pop 4; its neighbours 0 then 3 move to bucket 1 -> [3,0]; pop 3; 2 moves
    #   to the head -> [2,0]; pop 2; 1 -> [1,0]; pop 1; 0 -> bucket 0; pop 0.
    #   path 0-1-2-3: bucket 1 = [3,0], bucket 2 = [2,1]; pop 3; 2 -> bucket 1
    #   head [2,0]; pop 2; 1 -> [1,0]; pop 1; 0 -> bucket 0; pop 0.
    # An insert or a decrement at a bucket's tail gives another order.
    cases = [(5, [(0, 1), (1, 2), (2, 3), (3, 4), (4, 0)], [4, 3, 2, 1, 0]),
             (4, [(0, 1), (1, 2), (2, 3)], [3, 2, 1, 0])]
    for n, edges, want in cases:
        g = kplex.Graph.from_edges(n, edges)
        if impl == "library":
            _, _, orig, dseq = kplex.preprocess(g, 1, 2)
        else:
            d0, n0 = g.csr()
            _, _, orig, dseq = kpo.preprocess(kpo.OGraph.from_csr(d0, n0), 1, 2)
        assert orig.tolist() == list(range(n))
        assert dseq.tolist() == want


@pytest.mark.parametrize("seed", range(5))
def test_preprocess_matches_oracle_random(seed):
    adj = random_graph(200, 0.08, seed)
    deg, nbr = csr(adj)
    g = kplex.Graph.from_csr(deg, nbr)
    for truss in (0, 1):
        compare_prep(g, 2, 5, truss)


def test_preprocess_rejects_bad_params():
    g = kplex.Graph.from_edges(4, [(0, 1), (1, 2)])
    with pytest.raises(kplex.KplexError) as e:
        kplex.preprocess(g, 3, 5)     # q < 2k
    assert e.value.code == kplex.KP_EINVAL
    with pytest.raises(kplex.KplexError):
        kplex.preprocess(g, 0, 4)


def test_relabel_is_a_permutation():
    g = kplex.Graph.er_planted(300, 0.03, 2, 20, 2, 5)
    h, perm = g.relabel(99)
    assert sorted(perm.tolist()) == list(range(300))
    assert h.validate() == 0 and h.m == g.m


def test_kp_enumerate_without_gpu_fails_loudly():
    try:
        kplex.device_info()
        pytest.skip("a GPU is present")
    except kplex.KplexError:
        pass
    g = kplex.Graph.from_edges(6, [(i, j) for i in range(6) for j in range(i + 1, 6)])
    with pytest.raises(kplex.KplexError) as e:
        kplex.enumerate_plexes(g, 2, 4)
    assert e.value.code in (kplex.KP_ENODEV, kplex.KP_EHIP)


def test_tune_settings_without_device():
    # the launch-settings table is host state: setting and clearing it needs no GPU
    kplex.set_tune({1: dict(blocks_per_cu=1, donate_period=32), 16: dict(ring_frames=8)})
    kplex.set_tune(None)


def test_server_needs_a_device():
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is visible")
    with pytest.raises(kplex.KplexError) as e:
        kplex.Server()
    assert e.value.code == kplex.KP_ENODEV

"""GPU preprocessing (kp_preprocess_gpu, KP_FLAG_GPU_PREP): the (q-k)-core
peel, the relabel and the truss supports on the GPU must give the host
preprocessing's result exactly -- the peeled CSR, the original-id map and the
degeneracy order -- and the enumeration that follows must not change."""
import sys

import numpy as np
import pytest

import kplex
from kp_testutil import csr, random_graph

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def same_prep(g, k, q, truss, name):
    say(f"prep {name} k={k} q={q} truss={truss}")
    host = kplex.preprocess(g, k, q, truss)
    dev = kplex.preprocess(g, k, q, truss, gpu=True)
    for a, b, what in zip(host, dev, ("deg", "nbr", "orig", "dseq")):
        assert np.array_equal(a, b), (name, truss, what)
    return len(host[0])


@pytest.mark.parametrize("truss", [0, 1])
def test_gpu_prep_equals_host_on_the_reference_datasets(datasets, truss):
    g = kplex.Graph.load_bin(datasets["email-euall"])
    assert same_prep(g, 3, 12, truss, "email-euall") == (3156 if not truss else 1086)
    g = kplex.Graph.load_bin(datasets["soc-epinions"])
    assert same_prep(g, 4, 30, truss, "soc-epinions") == (4315 if not truss else 622)


def test_gpu_prep_equals_host_on_the_large_workloads():
    g, k, q, _, tr = kplex.workload_cl320k()
    assert same_prep(g, k, q, tr, "cl320k") == 50565
    g, k, q, _, tr = kplex.workload_cl75k()
    assert same_prep(g, k, q, tr, "cl75k") == 16915


@pytest.mark.parametrize("seed", range(3))
def test_gpu_prep_equals_host_on_small_graphs(seed):
    g = kplex.Graph.from_csr(*csr(random_graph(200, 0.08, seed)))
    for truss in (0, 1):
        same_prep(g, 2, 5, truss, f"random{seed}")
    # a graph the core removes entirely
    h = kplex.Graph.from_csr(*csr(random_graph(60, 0.02, seed)))
    same_prep(h, 3, 12, 0, f"sparse{seed}")


def test_enumeration_after_gpu_prep(datasets):
    g = kplex.Graph.load_bin(datasets["email-euall"])
    say("email-euall truss with GPU preprocessing")
    r = kplex.enumerate_plexes(g, 3, 12, truss=1, gpu_prep=True, time_limit_s=90.0)
    assert r["count"] == 32639016   # README.md:144
    b = kplex.enumerate_batch([g, g], [dict(k=3, q=12, truss=1, flags=kplex.KP_FLAG_GPU_PREP), dict(k=3, q=12)])
    assert b[0]["count"] == b[1]["count"] == 32639016 and b[0]["hash"] == b[1]["hash"]

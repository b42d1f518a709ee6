"""ctypes binding of libkplex.so (the C ABI in include/kplex.h).

This is the Python face of the drop-in boundary: plain pointers and sizes go
through ctypes, no torch types.  The library itself is C++/HIP; nothing here
computes anything on the enumeration path.
"""
import ctypes
import os
from ctypes import POINTER, c_char_p, c_double, c_float, c_int, c_size_t, c_uint32, c_uint64, c_void_p

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "build", "libkplex.so")
CLI_PATH = os.path.join(HERE, "build", "PlexEnum")

KP_OK = 0
KP_ENOENT = -2
KP_EIO = -5
KP_ENOMEM = -12
KP_ENODEV = -19
KP_EINVAL = -22
KP_ERANGE = -34
KP_EOVERFLOW = -75
KP_ETIMEDOUT = -110
KP_EBUSY = -16
KP_EHIP = -1000

KP_FLAG_LIST = 1
KP_FLAG_VALIDATE = 2
KP_FLAG_VERBOSE = 4
KP_FLAG_GPU_PREP = 8
KP_FLAG_NO_NARROW = 16


class KpGraph(ctypes.Structure):
    _fields_ = [("n", c_uint32), ("m", c_uint64), ("off", POINTER(c_uint64)),
                ("nbr", POINTER(c_uint32)), ("deg", POINTER(c_uint32))]


class KpParams(ctypes.Structure):
    _fields_ = [("k", c_int), ("q", c_int), ("thres", c_float), ("truss", c_int), ("flags", c_uint32),
                ("time_limit_s", c_double), ("queue_capacity", c_uint64), ("root_begin", ctypes.c_int64),
                ("root_end", ctypes.c_int64), ("pause_every", c_uint64)]


class KpResult(ctypes.Structure):
    _fields_ = [("count", c_uint64), ("tasks", c_uint64), ("valid_roots", c_uint64),
                ("hash_a", c_uint64), ("hash_b", c_uint64), ("pn", c_uint32), ("pm", c_uint64),
                ("max_n", c_uint32), ("max_l", c_uint32), ("alg_bytes", c_uint64), ("adj_bytes", c_uint64), ("search_alg_bytes", c_uint64),
                ("preprocess_ms", c_double), ("upload_ms", c_double), ("decompose_ms", c_double),
                ("search_ms", c_double), ("wall_ms", c_double), ("spilled", c_uint64), ("pauses", c_uint64),
                ("nlisted", c_uint64), ("list_off", POINTER(c_uint64)), ("list_ids", POINTER(c_uint32)),
                ("pn_first", c_uint32), ("status", ctypes.c_int32), ("narrowed", c_uint64)]


class KpTune(ctypes.Structure):
    _fields_ = [("blocks_per_cu", ctypes.c_int32 * 7), ("donate_period", c_uint32 * 7),
                ("ring_frames", c_uint32 * 7), ("queue_slots", c_uint32 * 7), ("class_ms", c_double * 7)]


CLASS_WORDS = (1, 2, 3, 4, 6, 8, 16)


class KpServerConfig(ctypes.Structure):
    _fields_ = [("max_inflight", c_size_t), ("mem_budget_bytes", c_uint64), ("slice_ms", c_double)]


# every entry point include/kplex.h declares: name -> (restype, argtypes)
SIGNATURES = {
    "kp_graph_load_bin": (c_int, [c_char_p, POINTER(KpGraph)]),
    "kp_graph_save_bin": (c_int, [POINTER(KpGraph), c_char_p]),
    "kp_graph_from_csr": (c_int, [c_uint32, POINTER(c_uint32), POINTER(c_uint32), POINTER(KpGraph)]),
    "kp_graph_validate": (c_int, [POINTER(KpGraph)]),
    "kp_graph_free": (None, [POINTER(KpGraph)]),
    "kp_enumerate": (c_int, [POINTER(KpGraph), POINTER(KpParams), POINTER(KpResult)]),
    "kp_enumerate_noalloc": (c_int, [POINTER(KpGraph), POINTER(KpParams), POINTER(KpResult)]),
    "kp_result_free": (None, [POINTER(KpResult)]),
    "kp_prepare": (c_int, [POINTER(KpGraph), POINTER(KpParams), POINTER(c_void_p)]),
    "kp_run": (c_int, [c_void_p, POINTER(KpResult)]),
    "kp_prepared_bytes": (c_uint64, [c_void_p]),
    "kp_set_window": (c_int, [c_void_p, ctypes.c_int64, ctypes.c_int64]),
    "kp_prepared_free": (None, [c_void_p]),
    "kp_root_subgraph": (c_int, [c_void_p, ctypes.c_int64, POINTER(c_uint32), POINTER(c_uint32), POINTER(c_uint32),
                                 POINTER(c_uint32), POINTER(c_uint32), c_uint32, POINTER(ctypes.c_uint8), c_size_t]),
    "kp_autotune": (c_int, [c_void_p, POINTER(KpTune)]),
    "kp_set_tune": (None, [POINTER(KpTune)]),
    "kp_enumerate_batch": (c_int, [POINTER(POINTER(KpGraph)), POINTER(KpParams), c_size_t, POINTER(KpResult)]),
    "kp_server_create": (c_int, [POINTER(KpServerConfig), POINTER(c_void_p)]),
    "kp_server_add_graph": (c_int, [c_void_p, POINTER(KpGraph), POINTER(c_uint32)]),
    "kp_server_submit": (c_int, [c_void_p, c_uint32, POINTER(KpParams), POINTER(c_uint64)]),
    "kp_server_poll": (c_int, [c_void_p, POINTER(c_uint64), POINTER(KpResult), c_size_t, POINTER(c_size_t)]),
    "kp_server_outstanding": (c_size_t, [c_void_p]),
    "kp_server_preload": (c_int, [c_void_p, POINTER(c_uint32), POINTER(KpParams), c_size_t]),
    "kp_server_destroy": (c_int, [c_void_p]),
    "kp_strerror": (c_char_p, [c_int]),
    "kp_last_error": (c_char_p, []),
    "kp_device_info": (c_int, [c_char_p, c_size_t]),
    "kp_gen_er_planted": (c_int, [c_uint32, c_double, c_uint32, c_uint32, c_uint32, c_uint64, POINTER(KpGraph)]),
    "kp_gen_chung_lu": (c_int, [c_uint32, c_double, c_uint64, c_double, c_uint32, c_uint32, c_uint32, c_double,
                                c_uint64, POINTER(KpGraph)]),
    "kp_graph_relabel": (c_int, [POINTER(KpGraph), c_uint64, POINTER(KpGraph), POINTER(c_uint32)]),
    "kp_preprocess": (c_int, [POINTER(KpGraph), c_int, c_int, c_int, POINTER(KpGraph), POINTER(c_uint32),
                              POINTER(ctypes.c_int32)]),
    "kp_preprocess_gpu": (c_int, [POINTER(KpGraph), c_int, c_int, c_int, POINTER(KpGraph), POINTER(c_uint32),
                                  POINTER(ctypes.c_int32)]),
}

_lib = None


class KplexError(RuntimeError):
    def __init__(self, code, where):
        lib = load()
        msg = lib.kp_strerror(code).decode()
        detail = lib.kp_last_error().decode()
        super().__init__(f"{where}: {msg} ({code}): {detail}")
        self.code = code


def load(path=None):
    """Loads libkplex.so; raises if it has not been built."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise FileNotFoundError(f"{p} is missing: build it with `make -C maximal-kplex-enumeration_amd`")
    lib = ctypes.CDLL(p)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


def _check(rc, where):
    if rc != KP_OK:
        raise KplexError(rc, where)


def device_info():
    lib = load()
    buf = ctypes.create_string_buffer(256)
    _check(lib.kp_device_info(buf, 256), "kp_device_info")
    return buf.value.decode()


class Graph:
    """Owns a kp_graph allocated by the library."""

    def __init__(self):
        self.g = KpGraph()

    def __del__(self):
        try:
            load().kp_graph_free(ctypes.byref(self.g))
        except Exception:
            pass

    @property
    def n(self):
        return self.g.n

    @property
    def m(self):
        return self.g.m

    @classmethod
    def load_bin(cls, path):
        self = cls()
        _check(load().kp_graph_load_bin(str(path).encode(), ctypes.byref(self.g)), f"load {path}")
        return self

    def save_bin(self, path):
        _check(load().kp_graph_save_bin(ctypes.byref(self.g), str(path).encode()), f"save {path}")

    @classmethod
    def from_csr(cls, deg, nbr):
        deg = np.ascontiguousarray(deg, dtype=np.uint32)
        nbr = np.ascontiguousarray(nbr, dtype=np.uint32)
        self = cls()
        _check(load().kp_graph_from_csr(len(deg), deg.ctypes.data_as(POINTER(c_uint32)),
                                        nbr.ctypes.data_as(POINTER(c_uint32)), ctypes.byref(self.g)),
               "kp_graph_from_csr")
        return self

    @classmethod
    def from_edges(cls, n, edges):
        """Simple undirected graph from an iterable of (u, v) pairs."""
        adj = [set() for _ in range(n)]
        for u, v in edges:
            if u != v:
                adj[u].add(v)
                adj[v].add(u)
        deg = np.array([len(a) for a in adj], dtype=np.uint32)
        nbr = np.array([x for a in adj for x in sorted(a)], dtype=np.uint32)
        return cls.from_csr(deg, nbr)

    def csr(self):
        n, m = self.g.n, self.g.m
        deg = np.ctypeslib.as_array(self.g.deg, shape=(max(n, 1),))[:n].copy()
        nbr = np.ctypeslib.as_array(self.g.nbr, shape=(max(m, 1),))[:m].copy()
        return deg, nbr

    def validate(self):
        return load().kp_graph_validate(ctypes.byref(self.g))

    def relabel(self, seed):
        out = Graph()
        perm = np.zeros(max(self.n, 1), dtype=np.uint32)
        _check(load().kp_graph_relabel(ctypes.byref(self.g), seed, ctypes.byref(out.g),
                                       perm.ctypes.data_as(POINTER(c_uint32))), "kp_graph_relabel")
        return out, perm[:self.n]

    @classmethod
    def er_planted(cls, n, p, groups, gsize, max_removed, seed):
        self = cls()
        _check(load().kp_gen_er_planted(n, p, groups, gsize, max_removed, seed, ctypes.byref(self.g)),
               "kp_gen_er_planted")
        return self

    @classmethod
    def chung_lu(cls, n, gamma, m_target, max_deg, ncomm, cmin, cmax, dens, seed):
        self = cls()
        _check(load().kp_gen_chung_lu(n, gamma, m_target, max_deg, ncomm, cmin, cmax, dens, seed,
                                      ctypes.byref(self.g)), "kp_gen_chung_lu")
        return self


def preprocess(graph, k, q, truss=0, gpu=False):
    """Preprocessing only (kp_preprocess on the host, or kp_preprocess_gpu with
    gpu=True): returns (peeled CSR deg, nbr, orig ids, degeneracy order)."""
    lib = load()
    out = Graph()
    n = max(graph.n, 1)
    orig = np.zeros(n, dtype=np.uint32)
    dseq = np.zeros(n, dtype=np.int32)
    fn = lib.kp_preprocess_gpu if gpu else lib.kp_preprocess
    _check(fn(ctypes.byref(graph.g), k, q, truss, ctypes.byref(out.g),
              orig.ctypes.data_as(POINTER(c_uint32)), dseq.ctypes.data_as(POINTER(ctypes.c_int32))),
           "kp_preprocess_gpu" if gpu else "kp_preprocess")
    deg, nbr = out.csr()
    return deg, nbr, orig[:out.n], dseq[:out.n]


def _result_dict(r, listed):
    d = {f: getattr(r, f) for f, _ in KpResult._fields_ if f not in ("list_off", "list_ids")}
    d["hash"] = f"{r.hash_b:016x}{r.hash_a:016x}"
    if listed:
        off = np.ctypeslib.as_array(r.list_off, shape=(r.nlisted + 1,)).copy() if r.nlisted else np.zeros(1, np.uint64)
        ids = np.ctypeslib.as_array(r.list_ids, shape=(int(off[-1]),)).copy() if off[-1] else np.zeros(0, np.uint32)
        d["plexes"] = [tuple(int(x) for x in ids[off[i]:off[i + 1]]) for i in range(int(r.nlisted))]
    return d


def make_params(k, q, thres=0.25, truss=0, flags=0, time_limit_s=0.0, queue_capacity=0, roots=None,
                pause_every=0):
    rb, re_ = roots if roots else (0, 0)
    return KpParams(k=k, q=q, thres=thres, truss=truss, flags=flags, time_limit_s=time_limit_s,
                    queue_capacity=queue_capacity, root_begin=rb, root_end=re_, pause_every=pause_every)


def enumerate_plexes(graph, k, q, thres=0.25, truss=0, listing=False, time_limit_s=0.0, validate=False, roots=None,
                     queue_capacity=0, pause_every=0, gpu_prep=False, narrow=True):
    """kp_enumerate: returns a dict with count, hash, tasks, timings (and plexes when listing).
    roots=(b, e) restricts the enumeration to roots at degeneracy positions [b, e).
    pause_every > 0 pauses each search launch after about that many frames and
    resumes it; thres * queue_capacity frames stay on the device, the rest spill
    to pinned host memory.  narrow=False keeps every frame in its root's size
    class (results are the same either way)."""
    lib = load()
    flags = (KP_FLAG_LIST if listing else 0) | (KP_FLAG_VALIDATE if validate else 0) | \
        (KP_FLAG_GPU_PREP if gpu_prep else 0) | (0 if narrow else KP_FLAG_NO_NARROW)
    p = make_params(k, q, thres, truss, flags, time_limit_s, queue_capacity, roots=roots, pause_every=pause_every)
    r = KpResult()
    _check(lib.kp_enumerate(ctypes.byref(graph.g), ctypes.byref(p), ctypes.byref(r)), "kp_enumerate")
    try:
        return _result_dict(r, listing)
    finally:
        lib.kp_result_free(ctypes.byref(r))


def enumerate_batch(graphs, params, listing=False):
    """kp_enumerate_batch: graphs[i] with params[i] = (k, q) or (k, q, truss) or a
    dict of make_params arguments; all roots share one pool of launches.
    Returns one result dict per graph."""
    lib = load()
    n = len(graphs)
    ps = (KpParams * max(n, 1))()
    for i, p in enumerate(params):
        if isinstance(p, dict):
            kw = dict(p)
        else:
            kw = dict(k=p[0], q=p[1], truss=p[2] if len(p) > 2 else 0)
        kw.setdefault("flags", KP_FLAG_LIST if listing else 0)
        ps[i] = make_params(**kw)
    gp = (POINTER(KpGraph) * max(n, 1))()
    for i, g in enumerate(graphs):
        gp[i] = ctypes.pointer(g.g)
    rs = (KpResult * max(n, 1))()
    try:
        _check(lib.kp_enumerate_batch(gp, ps, n, rs), "kp_enumerate_batch")
        return [_result_dict(rs[i], listing) for i in range(n)]
    finally:
        for i in range(n):
            lib.kp_result_free(ctypes.byref(rs[i]))


class Prepared:
    """kp_prepare / kp_run: the graph preprocessed and resident in HBM."""

    def __init__(self, graph, k, q, thres=0.25, truss=0, listing=False, time_limit_s=0.0, roots=None,
                 queue_capacity=0, pause_every=0, narrow=True):
        self.lib = load()
        self.h = c_void_p()
        self.listing = listing
        p = make_params(k, q, thres, truss, (KP_FLAG_LIST if listing else 0) | (0 if narrow else KP_FLAG_NO_NARROW),
                        time_limit_s, queue_capacity, roots=roots, pause_every=pause_every)
        _check(self.lib.kp_prepare(ctypes.byref(graph.g), ctypes.byref(p), ctypes.byref(self.h)), "kp_prepare")

    def run(self):
        r = KpResult()
        _check(self.lib.kp_run(self.h, ctypes.byref(r)), "kp_run")
        try:
            return _result_dict(r, self.listing)
        finally:
            self.lib.kp_result_free(ctypes.byref(r))

    def bytes(self):
        return self.lib.kp_prepared_bytes(self.h)

    def set_roots(self, roots):
        """kp_set_window: later runs enumerate root positions [b, e) (None = all)."""
        b, e = roots if roots else (0, 0)
        _check(self.lib.kp_set_window(self.h, b, e), "kp_set_window")

    def root_subgraph(self, pos):
        """kp_root_subgraph: the root at degeneracy position pos as the GPU
        decomposes and builds it.  Returns None for an invalid root, else a dict
        with block, hop, left (peeled ids) and pairs (n x n uint8: bit 0 adjacent,
        bit 1 class >= UNLINK2EQUAL, bit 2 class >= UNLINK2MORE)."""
        cap = 1 << 16
        block = np.zeros(1024, dtype=np.uint32)
        left = np.zeros(cap, dtype=np.uint32)
        pairs = np.zeros(1024 * 1024, dtype=np.uint8)
        n, hop, l = c_uint32(), c_uint32(), c_uint32()
        rc = self.lib.kp_root_subgraph(self.h, pos, block.ctypes.data_as(POINTER(c_uint32)), ctypes.byref(n),
                                       ctypes.byref(hop), left.ctypes.data_as(POINTER(c_uint32)), ctypes.byref(l), cap,
                                       pairs.ctypes.data_as(POINTER(ctypes.c_uint8)), pairs.size)
        if rc < 0:
            _check(rc, "kp_root_subgraph")
        if rc == 0:
            return None
        nn = n.value
        return {"block": block[:nn].copy(), "hop": hop.value, "left": left[:l.value].copy(),
                "pairs": pairs[:nn * nn].reshape(nn, nn).copy()}

    def autotune(self):
        """kp_autotune: returns {class words W: settings} for the classes present."""
        t = KpTune()
        _check(self.lib.kp_autotune(self.h, ctypes.byref(t)), "kp_autotune")
        return {CLASS_WORDS[c]: dict(blocks_per_cu=t.blocks_per_cu[c], donate_period=t.donate_period[c],
                                     ring_frames=t.ring_frames[c], queue_slots=t.queue_slots[c],
                                     class_ms=t.class_ms[c])
                for c in range(7) if t.class_ms[c] > 0}

    def close(self):
        if self.h:
            self.lib.kp_prepared_free(self.h)
            self.h = c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def set_tune(settings=None):
    """kp_set_tune: settings = {class words W: dict(blocks_per_cu=, donate_period=, ring_frames=,
    queue_slots=)} or None for the defaults."""
    lib = load()
    if settings is None:
        lib.kp_set_tune(None)
        return
    t = KpTune()
    for w, d in settings.items():
        c = CLASS_WORDS.index(w)
        t.blocks_per_cu[c] = d.get("blocks_per_cu", 0)
        t.donate_period[c] = d.get("donate_period", 0)
        t.ring_frames[c] = d.get("ring_frames", 0)
        t.queue_slots[c] = d.get("queue_slots", 0)
    lib.kp_set_tune(ctypes.byref(t))


class Server:
    """kp_server_*: continuous batching of (graph, k, q) queries."""

    def __init__(self, max_inflight=64, mem_budget_bytes=0, slice_ms=0.0):
        self.lib = load()
        self.h = c_void_p()
        cfg = KpServerConfig(max_inflight=max_inflight, mem_budget_bytes=mem_budget_bytes, slice_ms=slice_ms)
        _check(self.lib.kp_server_create(ctypes.byref(cfg), ctypes.byref(self.h)), "kp_server_create")

    def add_graph(self, graph):
        gid = c_uint32()
        _check(self.lib.kp_server_add_graph(self.h, ctypes.byref(graph.g), ctypes.byref(gid)), "kp_server_add_graph")
        return gid.value

    def preload(self, items):
        """items: [(graph_id, k, q)] or [(graph_id, k, q, truss)]; preprocessed in parallel, kept in HBM."""
        n = len(items)
        ids = (c_uint32 * max(n, 1))()
        ps = (KpParams * max(n, 1))()
        for i, it in enumerate(items):
            ids[i] = it[0]
            ps[i] = make_params(it[1], it[2], truss=it[3] if len(it) > 3 else 0)
        _check(self.lib.kp_server_preload(self.h, ids, ps, n), "kp_server_preload")

    def submit(self, graph_id, k, q, truss=0):
        t = c_uint64()
        p = make_params(k, q, truss=truss)
        _check(self.lib.kp_server_submit(self.h, graph_id, ctypes.byref(p), ctypes.byref(t)), "kp_server_submit")
        return t.value

    def poll(self, cap=256):
        """One scheduling step; returns [(ticket, result dict)] of completed queries."""
        ts = (c_uint64 * cap)()
        rs = (KpResult * cap)()
        nd = c_size_t()
        _check(self.lib.kp_server_poll(self.h, ts, rs, cap, ctypes.byref(nd)), "kp_server_poll")
        return [(ts[i], _result_dict(rs[i], False)) for i in range(nd.value)]

    def outstanding(self):
        return self.lib.kp_server_outstanding(self.h)

    def close(self):
        if self.h:
            self.lib.kp_server_destroy(self.h)
            self.h = c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------- workloads
# The five synthetic workloads (seeds 1..5).  Each returns (graph, k, q, thres, truss).
def workload_er2000():
    """G(2000, 0.01) + 20 planted groups of 40 (clique minus <= 2 edges per vertex); k=3 q=20."""
    return Graph.er_planted(2000, 0.01, 20, 40, 2, 1), 3, 20, 0.25, 0

def workload_cl75k():
    """Chung-Lu n=75,000 (exponent 2.1, max expected degree 3,000, ~508k edges)
    + 300 communities of 35-60 at density 0.9; k=4 q=30."""
    return Graph.chung_lu(75000, 2.1, 508000, 3000.0, 300, 35, 60, 0.9, 2), 4, 30, 0.25, 0


def workload_cl320k():
    """Chung-Lu n=320,000 (exponent 2.0, max expected degree 5,000, ~1.05M edges)
    + 1,000 communities of 30-80 at density 0.85; k=5 q=30, truss on."""
    return Graph.chung_lu(320000, 2.0, 1050000, 5000.0, 1000, 30, 80, 0.85, 5), 5, 30, 0.25, 1


def batch_graph_specs(count=4096, nmax=20000, seed=3, kset=(2, 3, 4), qfun=lambda k: 3 * k + 8):
    """Specs of the many-graph workload: n log-uniform in [200, nmax], Chung-Lu
    exponent 2.3 with average degree 10, 1-20 communities of 25-50 at density 0.9."""
    rng = np.random.default_rng(seed)
    specs = []
    for i in range(count):
        n = int(np.floor(np.exp(rng.uniform(np.log(200), np.log(nmax)))))
        k = int(rng.choice(kset))
        ncomm = int(rng.integers(1, 21))
        specs.append(dict(n=n, gamma=2.3, m_target=5 * n, max_deg=float(max(8.0, 4.0 * np.sqrt(n))),
                          ncomm=ncomm, cmin=25, cmax=50, dens=0.9, seed=seed * 1_000_003 + i, k=k, q=qfun(k)))
    return specs


def graph_from_spec(s):
    return Graph.chung_lu(s["n"], s["gamma"], s["m_target"], s["max_deg"], s["ncomm"], s["cmin"], s["cmax"],
                          s["dens"], s["seed"])


def served_graph_specs():
    """The 256 resident graphs of the served workload: built like the many-graph
    workload but with n log-uniform in [200, 200,000]."""
    return batch_graph_specs(count=256, nmax=200000, seed=4, kset=(3, 4, 5), qfun=lambda k: 5 * k + 10)


def served_queries(rate=200.0, duration_s=120.0, seed=7, ngraphs=256):
    """The served query stream: Poisson arrivals at `rate` per second over
    `duration_s` seconds; each query names a graph uniformly, k in {3,4,5} and
    q = 5k+10.  Returns (arrival times in s, [(graph index, k, q)])."""
    rng = np.random.default_rng(seed)
    arrivals, queries = [], []
    t = rng.exponential(1.0 / rate)
    while t < duration_s:
        k = int(rng.choice([3, 4, 5]))
        arrivals.append(t)
        queries.append((int(rng.integers(ngraphs)), k, 5 * k + 10))
        t += rng.exponential(1.0 / rate)
    return arrivals, queries


def served_fixture_queries():
    """The served queries with recorded oracle results: the first 32 of the
    stream, plus each of the 8 largest graphs (n >= 1e5) at k = 3, 4, 5 in turn."""
    specs = served_graph_specs()
    _, qs = served_queries(duration_s=1.0)
    out = list(qs[:32])
    big = sorted(range(len(specs)), key=lambda i: -specs[i]["n"])[:8]
    for j, gi in enumerate(big):
        k = 3 + j % 3
        out.append((gi, k, 5 * k + 10))
    return out

"""ctypes binding of the CPU oracle (oracle/_build/libkporacle.so).

TEST INFRASTRUCTURE: only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg may import this module, and only as the checker.
"""
import ctypes
import os
import subprocess
from ctypes import POINTER, c_char_p, c_double, c_int, c_long, c_size_t, c_uint32, c_uint64, c_uint8

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libkporacle.so")


class KpoGraph(ctypes.Structure):
    _fields_ = [("n", c_uint32), ("m", c_uint64), ("off", POINTER(c_uint64)), ("nbr", POINTER(c_uint32)),
                ("deg", POINTER(c_uint32))]


class KpoResult(ctypes.Structure):
    _fields_ = [("count", c_uint64), ("tasks", c_uint64), ("valid_roots", c_uint64), ("hash_a", c_uint64),
                ("hash_b", c_uint64), ("alg_bytes", c_uint64), ("adj_bytes", c_uint64), ("pn", c_uint32),
                ("pm", c_uint64), ("max_n", c_uint32), ("max_l", c_uint32), ("wall_ms", c_double),
                ("preprocess_ms", c_double), ("nlisted", c_uint64), ("list_off", POINTER(c_uint64)),
                ("list_ids", POINTER(c_uint32))]


class KpoPrep(ctypes.Structure):
    _fields_ = [("g", KpoGraph), ("orig", POINTER(c_uint32)), ("dseq", POINTER(ctypes.c_int32)),
                ("dpos", POINTER(ctypes.c_int32))]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    lib = ctypes.CDLL(LIB_PATH)
    lib.kpo_load_bin.argtypes = [c_char_p, POINTER(KpoGraph)]
    lib.kpo_graph_from_csr.argtypes = [c_uint32, POINTER(c_uint32), POINTER(c_uint32), POINTER(KpoGraph)]
    lib.kpo_graph_free.argtypes = [POINTER(KpoGraph)]
    lib.kpo_graph_free.restype = None
    lib.kpo_enumerate_range.argtypes = [POINTER(KpoGraph), c_int, c_int, c_int, c_int, c_uint32, c_long, c_long,
                                        POINTER(KpoResult)]
    lib.kpo_result_free.argtypes = [POINTER(KpoResult)]
    lib.kpo_result_free.restype = None
    lib.kpo_preprocess.argtypes = [POINTER(KpoGraph), c_int, c_int, c_int, POINTER(KpoPrep)]
    lib.kpo_prep_free.argtypes = [POINTER(KpoPrep)]
    lib.kpo_prep_free.restype = None
    lib.kpo_root_block.argtypes = [POINTER(KpoPrep), c_int, c_int, c_long, POINTER(c_uint32), POINTER(c_uint32),
                                   POINTER(c_uint32), POINTER(c_uint32), POINTER(c_uint32), POINTER(c_uint8),
                                   c_size_t]
    lib.kpo_mix64.argtypes = [c_uint64]
    lib.kpo_mix64.restype = c_uint64
    _lib = lib
    return lib


class OGraph:
    def __init__(self):
        self.g = KpoGraph()

    def __del__(self):
        try:
            load().kpo_graph_free(ctypes.byref(self.g))
        except Exception:
            pass

    @classmethod
    def load_bin(cls, path):
        self = cls()
        rc = load().kpo_load_bin(str(path).encode(), ctypes.byref(self.g))
        if rc:
            raise RuntimeError(f"kpo_load_bin({path}) = {rc}")
        return self

    @classmethod
    def from_csr(cls, deg, nbr):
        deg = np.ascontiguousarray(deg, dtype=np.uint32)
        nbr = np.ascontiguousarray(nbr, dtype=np.uint32)
        self = cls()
        rc = load().kpo_graph_from_csr(len(deg), deg.ctypes.data_as(POINTER(c_uint32)),
                                       nbr.ctypes.data_as(POINTER(c_uint32)), ctypes.byref(self.g))
        if rc:
            raise RuntimeError(f"kpo_graph_from_csr = {rc}")
        return self


def enumerate_plexes(og, k, q, truss=0, threads=0, listing=False, pos_begin=0, pos_end=-1):
    lib = load()
    r = KpoResult()
    rc = lib.kpo_enumerate_range(ctypes.byref(og.g), k, q, truss, threads, 1 if listing else 0, pos_begin, pos_end,
                                 ctypes.byref(r))
    if rc:
        raise RuntimeError(f"kpo_enumerate = {rc}")
    try:
        d = {f: getattr(r, f) for f, _ in KpoResult._fields_ if f not in ("list_off", "list_ids")}
        d["hash"] = f"{r.hash_b:016x}{r.hash_a:016x}"
        if listing:
            n = r.nlisted
            off = np.ctypeslib.as_array(r.list_off, shape=(n + 1,)).copy() if n else np.zeros(1, np.uint64)
            ids = np.ctypeslib.as_array(r.list_ids, shape=(int(off[-1]),)).copy() if off[-1] else np.zeros(0, np.uint32)
            d["plexes"] = [tuple(int(x) for x in ids[off[i]:off[i + 1]]) for i in range(int(n))]
        return d
    finally:
        lib.kpo_result_free(ctypes.byref(r))


def preprocess(og, k, q, truss=0):
    """Returns (deg, nbr, orig, dseq) of the oracle's preprocessed graph."""
    lib = load()
    p = KpoPrep()
    rc = lib.kpo_preprocess(ctypes.byref(og.g), k, q, truss, ctypes.byref(p))
    if rc:
        raise RuntimeError(f"kpo_preprocess = {rc}")
    try:
        n, m = p.g.n, p.g.m
        deg = np.ctypeslib.as_array(p.g.deg, shape=(max(n, 1),))[:n].copy()
        nbr = np.ctypeslib.as_array(p.g.nbr, shape=(max(m, 1),))[:m].copy()
        orig = np.ctypeslib.as_array(p.orig, shape=(max(n, 1),))[:n].copy()
        dseq = np.ctypeslib.as_array(p.dseq, shape=(max(n, 1),))[:n].copy()
        return deg, nbr, orig, dseq
    finally:
        lib.kpo_prep_free(ctypes.byref(p))


def mix64(z):
    return load().kpo_mix64(z)


# hash of a list of plexes (original ids), as defined in include/kplex.h
SA, LA, SB, LB = 0x243F6A8885A308D3, 0x13198A2E03707344, 0xA4093822299F31D0, 0x082EFA98EC4E6C89
M64 = (1 << 64) - 1


def py_mix64(z):
    z &= M64
    z = ((z ^ (z >> 30)) * 0xbf58476d1ce4e5b9) & M64
    z = ((z ^ (z >> 27)) * 0x94d049bb133111eb) & M64
    return z ^ (z >> 31)


def hash_plexes(plexes):
    ha = hb = 0
    for s in plexes:
        sa = sum(py_mix64(v ^ SA) for v in s) & M64
        sb = sum(py_mix64(v ^ SB) for v in s) & M64
        ha = (ha + py_mix64((sa + len(s) * LA) & M64)) & M64
        hb = (hb + py_mix64((sb + len(s) * LB) & M64)) & M64
    return f"{hb:016x}{ha:016x}"


class RootBlocks:
    """kpo_root_block over one preprocessed graph: the oracle's decomposition
    and pair-class matrix of single roots."""

    def __init__(self, og, k, q, truss=0):
        self.lib = load()
        self.p = KpoPrep()
        self.k, self.q = k, q
        rc = self.lib.kpo_preprocess(ctypes.byref(og.g), k, q, truss, ctypes.byref(self.p))
        if rc:
            raise RuntimeError(f"kpo_preprocess = {rc}")
        self.pn = self.p.g.n

    def __del__(self):
        try:
            self.lib.kpo_prep_free(ctypes.byref(self.p))
        except Exception:
            pass

    def block(self, pos):
        """None for an invalid root, else block, hop, left (peeled ids) and the
        n x n class matrix (0..5 as in the reference's buildCommonMtx)."""
        n0 = max(self.pn, 1)
        B = np.zeros(n0, dtype=np.uint32)
        L = np.zeros(n0, dtype=np.uint32)
        cls = np.zeros(1024 * 1024, dtype=np.uint8)
        n, hop, l = c_uint32(), c_uint32(), c_uint32()
        v = self.lib.kpo_root_block(ctypes.byref(self.p), self.k, self.q, pos, B.ctypes.data_as(POINTER(c_uint32)),
                                    ctypes.byref(n), ctypes.byref(hop), L.ctypes.data_as(POINTER(c_uint32)),
                                    ctypes.byref(l), cls.ctypes.data_as(POINTER(c_uint8)), cls.size)
        if v < 0:
            raise RuntimeError(f"kpo_root_block = {v}")
        if v == 0:
            return None
        nn = n.value
        return {"block": B[:nn].copy(), "hop": hop.value, "left": L[:l.value].copy(),
                "cls": cls[:nn * nn].reshape(nn, nn).copy()}

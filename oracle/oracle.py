"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and by
bench.py's cpu_baseline leg, always as the checker or the timed CPU baseline,
never by the product path (plant_amd/).
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_HHL = os.path.join(HERE, "_ref", "hhl")

_lib = None


def build():
    """Compiles liboracle.so (and oracle/_ref/hhl when the reference tree is present)."""
    subprocess.run(["make", "-B", "-C", HERE], check=True, stdout=subprocess.DEVNULL)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        vp, i64, u32, i32 = C.c_void_p, C.c_int64, C.c_uint32, C.c_int32
        sig = {
            "orc_graph_from_arcs": ([i64, i64, vp, vp, vp, C.c_int], vp),
            "orc_graph_read": ([C.c_char_p, C.c_int, C.POINTER(C.c_void_p)], C.c_int),
            "orc_graph_free": ([vp], None),
            "orc_graph_n": ([vp], i64),
            "orc_graph_m": ([vp], i64),
            "orc_graph_arcs": ([vp, C.c_int], i64),
            "orc_graph_csr": ([vp, C.c_int, vp, vp, vp], None),
            "orc_dijkstra": ([vp, u32, C.c_int, vp], None),
            "orc_dijkstra_multi": ([vp, vp, i64, C.c_int, vp, C.c_int], None),
            "orc_order_read": ([C.c_char_p, vp, i64], i64),
            "orc_gll": ([vp, vp, i64, C.c_int, C.c_int, vp], vp),
            "orc_gll_bounded": ([vp, vp, i64, C.c_int, C.c_int, C.c_double, vp], vp),
            "orc_hybrid": ([vp, vp, i64, C.c_int, C.c_double, C.c_double, vp], vp),
            "orc_plant_range": ([vp, vp, i64, i64, i64, vp, vp, vp, C.c_int], vp),
            "orc_brute_chl": ([vp, vp, i64, C.c_int], vp),
            "orc_labels_free": ([vp], None),
            "orc_labels_n": ([vp], i64),
            "orc_labels_total": ([vp, C.c_int], i64),
            "orc_labels_copy": ([vp, C.c_int, vp, vp, vp], None),
            "orc_labels_from": ([i64, vp, vp, vp, vp, vp, vp], vp),
            "orc_labels_write": ([vp, C.c_char_p], C.c_int),
            "orc_labels_read": ([C.c_char_p], vp),
            "orc_query_batch": ([vp, vp, i64, vp, C.c_int], None),
            "orc_clean_cover": ([vp, u32, u32, C.c_int, i32, C.c_uint64], C.c_int),
        }
        for name, (args, res) in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data if a is not None and a.size else None


class OGraph:
    def __init__(self, h):
        if not h:
            raise ValueError("oracle graph construction failed")
        self.h = h

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.orc_graph_free(self.h)
            self.h = None

    @classmethod
    def from_arcs(cls, n, src, dst, w, undirected=False):
        src = np.ascontiguousarray(src, np.uint32)
        dst = np.ascontiguousarray(dst, np.uint32)
        w = np.ascontiguousarray(w, np.int32)
        return cls(lib().orc_graph_from_arcs(n, src.size, _p(src), _p(dst), _p(w), int(undirected)))

    @classmethod
    def read(cls, path, undirected=False):
        h = C.c_void_p()
        r = lib().orc_graph_read(os.fsencode(path), int(undirected), C.byref(h))
        if r < 0:
            return None, r
        return cls(h.value), r

    @property
    def n(self):
        return lib().orc_graph_n(self.h)

    def csr(self, rng=1):
        n = self.n
        m = lib().orc_graph_arcs(self.h, rng)
        off = np.zeros(n + 1, np.uint64)
        head = np.zeros(max(m, 1), np.uint32)
        ln = np.zeros(max(m, 1), np.int32)
        lib().orc_graph_csr(self.h, rng, _p(off), _p(head), _p(ln))
        return off, head[:m], ln[:m]

    def dijkstra(self, src, forward=True):
        out = np.zeros(self.n, np.int32)
        lib().orc_dijkstra(self.h, src, int(forward), _p(out))
        return out

    def dijkstra_multi(self, srcs, forward=True, threads=8):
        srcs = np.ascontiguousarray(srcs, np.uint32)
        out = np.zeros((srcs.size, self.n), np.int32)
        lib().orc_dijkstra_multi(self.h, _p(srcs), srcs.size, int(forward), _p(out), threads)
        return out


class OLabels:
    def __init__(self, h):
        if not h:
            raise ValueError("oracle labels construction failed")
        self.h = h

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.orc_labels_free(self.h)
            self.h = None

    @property
    def n(self):
        return lib().orc_labels_n(self.h)

    def copy(self, side):
        n = self.n
        t = lib().orc_labels_total(self.h, side)
        off = np.zeros(n + 1, np.uint64)
        hub = np.zeros(max(t, 1), np.uint32)
        dist = np.zeros(max(t, 1), np.int32)
        lib().orc_labels_copy(self.h, side, _p(off), _p(hub), _p(dist))
        return off, hub[:t], dist[:t]

    def total(self):
        return lib().orc_labels_total(self.h, 0) + lib().orc_labels_total(self.h, 1)

    def query(self, pairs, threads=8):
        pairs = np.ascontiguousarray(pairs, np.uint32).reshape(-1, 2)
        out = np.zeros(pairs.shape[0], np.int32)
        lib().orc_query_batch(self.h, _p(pairs), pairs.shape[0], _p(out), threads)
        return out

    def write(self, path):
        return lib().orc_labels_write(self.h, os.fsencode(path))

    @classmethod
    def read(cls, path):
        return cls(lib().orc_labels_read(os.fsencode(path)))

    @classmethod
    def from_csr(cls, n, side0, side1):
        a = []
        for off, hub, dist in (side0, side1):
            a += [np.ascontiguousarray(off, np.uint64), np.ascontiguousarray(hub, np.uint32),
                  np.ascontiguousarray(dist, np.int32)]
        return cls(lib().orc_labels_from(n, *[_p(x) for x in a]))


def gll(g, order, threads=8, sf=32):
    """run_paraPLL restated (shared_memory/akiba.hpp:132-217); returns (labels, lt/st/ct/supersteps)."""
    order = np.ascontiguousarray(order, np.uint32)
    t = np.zeros(4)
    return OLabels(lib().orc_gll(g.h, _p(order), order.size, threads, sf, _p(t))), t


def gll_bounded(g, order, threads, seconds, sf=32):
    """GLL over the rank prefix completed within `seconds` of lt+st+ct; returns
    (labels, stats) with stats = lt, st, ct, supersteps, roots."""
    order = np.ascontiguousarray(order, np.uint32)
    s = np.zeros(5)
    return OLabels(lib().orc_gll_bounded(g.h, _p(order), order.size, threads, sf, seconds, _p(s))), s


def hybrid(g, order, threads=8, cb=100.0, ps=1024.0):
    order = np.ascontiguousarray(order, np.uint32)
    s = np.zeros(3)
    return OLabels(lib().orc_hybrid(g.h, _p(order), order.size, threads, cb, ps, _p(s))), s


def plant_range(g, order, r0, r1, threads=8):
    order = np.ascontiguousarray(order, np.uint32)
    ts = np.zeros(2 * (r1 - r0), np.uint32)
    lc = np.zeros(2 * (r1 - r0), np.uint32)
    L = OLabels(lib().orc_plant_range(g.h, _p(order), order.size, r0, r1, None, _p(ts), _p(lc), threads))
    return L, ts, lc


def brute_chl(g, order, threads=8):
    order = np.ascontiguousarray(order, np.uint32)
    return OLabels(lib().orc_brute_chl(g.h, _p(order), order.size, threads))


def order_read(path):
    n = lib().orc_order_read(os.fsencode(path), None, 0)
    if n < 0:
        return None
    o = np.zeros(n, np.uint32)
    lib().orc_order_read(os.fsencode(path), _p(o), n)
    return o

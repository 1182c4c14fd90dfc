"""Python binding of libchl.so (the C ABI in include/chl.h) over ctypes.

Mirrors the reference's C++ surface for the hot path:

* ``Graph.load`` / ``Graph.from_arcs``   -- hl::Graph::read / add_arc + finalize
  (shared_memory/graph.hpp:327-353)
* ``order_load`` / ``order_write``        -- hl::Order::read / write (ordering.hpp:37-67)
* ``build``                               -- hl::run_paraPLL (shared_memory/akiba.hpp:132) and the
  hybrid hl::run (distributed_memory/akiba.hpp:190), on the GPU
* ``Labels.query``                        -- Labeling::query batched like batchLocalQuery
  (labeling.hpp:56-65, distributed_memory/query_funcs.hpp:23-30)
* ``Labels.write`` / ``Labels.read``      -- Labeling::write / read (labeling.hpp:167-208)

Errors raise ``ChlError`` carrying the library's status code. The library is
loaded from this package directory; nothing here falls back to host code.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libchl.so")

INFTY = 2147483647
MODE_GLL, MODE_PLANT, MODE_HYBRID = 0, 1, 2
BATCH_STATIC, BATCH_LEARNED = 0, 1


class ChlError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"chl error {code}: {msg}")
        self.code = code


class BuildOpts(C.Structure):
    _fields_ = [("mode", C.c_int), ("batching", C.c_int), ("autotune", C.c_int), ("batch_roots", C.c_int64),
                ("delta", C.c_int32), ("slots", C.c_int32), ("granularity", C.c_int), ("switch_ratio", C.c_double),
                ("plant_prune", C.c_int), ("device", C.c_int), ("verbose", C.c_int), ("out_capacity", C.c_int64),
                ("vertex_score", C.c_void_p), ("keep_segments", C.c_int), ("small_trees", C.c_int),
                ("medium_trees", C.c_int), ("slot_memory_mb", C.c_int), ("reserved", C.c_int * 4)]


class ServerOpts(C.Structure):
    _fields_ = [("streams", C.c_int), ("max_inflight_pairs", C.c_int64), ("max_batch_pairs", C.c_int64),
                ("reserved", C.c_int * 8)]


class ServerStats(C.Structure):
    _fields_ = [("completed", C.c_int64), ("pairs", C.c_int64), ("rejected", C.c_int64), ("p50_us", C.c_double),
                ("p99_us", C.c_double), ("mean_us", C.c_double), ("max_us", C.c_double), ("qps", C.c_double),
                ("p90_us", C.c_double), ("p999_us", C.c_double), ("late", C.c_int64), ("lag_p99_us", C.c_double),
                ("lag_max_us", C.c_double), ("reserved", C.c_double * 3)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}


class BuildStats(C.Structure):
    _fields_ = [("lt", C.c_double), ("st", C.c_double), ("ct", C.c_double), ("total_s", C.c_double),
                ("avg_label", C.c_double), ("labels", C.c_int64), ("batches", C.c_int64),
                ("trees_plant", C.c_int64), ("trees_pruned", C.c_int64), ("switch_rank", C.c_int64),
                ("settled", C.c_int64), ("arcs_scanned", C.c_int64), ("candidates", C.c_int64),
                ("prune_entries", C.c_int64), ("clean_entries", C.c_int64), ("delta_used", C.c_int64),
                ("slots_used", C.c_int64), ("batch_used", C.c_int64), ("relax_ms", C.c_double),
                ("clean_ms", C.c_double), ("relax_bytes", C.c_double), ("mlp_ms", C.c_double),
                ("pred_mae_learned", C.c_double), ("pred_mae_static", C.c_double), ("pred_roots", C.c_int64),
                ("granularity_used", C.c_int64), ("autotune_ms", C.c_double), ("switch_trials", C.c_int64),
                ("policy_ms", C.c_double), ("tree_launches", C.c_int64),
                ("clean_bytes", C.c_double), ("clean_kernel_ms", C.c_double),
                ("settled_distinct", C.c_int64), ("settled_arcs", C.c_int64), ("relax_bytes_processed", C.c_double),
                ("segment_ms", C.c_double), ("store_bytes", C.c_double), ("segments", C.c_int64),
                ("slot_shrinks", C.c_int64), ("slots_final", C.c_int64), ("slot_s", C.c_double),
                ("small_roots", C.c_int64), ("medium_roots", C.c_int64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}


_lib = None


def lib():
    """The loaded libchl.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ChlError(-7, f"{LIB_PATH} is missing: run __graft_entry__.build() first")
        L = C.CDLL(LIB_PATH)
        vp, i64, u64, i32 = C.c_void_p, C.c_int64, C.c_uint64, C.c_int32
        pp = C.POINTER(C.c_void_p)
        sig = {
            "chl_strerror": ([C.c_int], C.c_char_p),
            "chl_last_error": ([], C.c_char_p),
            "chl_version": ([], C.c_char_p),
            "chl_device_count": ([], C.c_int),
            "chl_graph_load": ([C.c_char_p, C.c_int, pp, C.POINTER(C.c_int)], C.c_int),
            "chl_graph_from_arcs": ([i64, i64, vp, vp, vp, C.c_int, pp], C.c_int),
            "chl_graph_write": ([vp, C.c_char_p], C.c_int),
            "chl_graph_n": ([vp], i64),
            "chl_graph_m": ([vp], i64),
            "chl_graph_arcs": ([vp, C.c_int], i64),
            "chl_graph_symmetric": ([vp], C.c_int),
            "chl_graph_csr": ([vp, C.c_int, vp, vp, vp], C.c_int),
            "chl_graph_free": ([vp], None),
            "chl_gen_grid": ([i64, i64, i32, i32, u64, pp], C.c_int),
            "chl_gen_knn": ([i64, C.c_int, u64, pp], C.c_int),
            "chl_gen_rmat": ([C.c_int, i64, C.c_double, C.c_double, C.c_double, i32, u64, pp], C.c_int),
            "chl_order_load": ([C.c_char_p, vp, i64, C.POINTER(i64)], C.c_int),
            "chl_order_write": ([C.c_char_p, vp, i64], C.c_int),
            "chl_order_betweenness": ([vp, i64, u64, vp], C.c_int),
            "chl_order_betweenness_scored": ([vp, i64, u64, vp, vp], C.c_int),
            "chl_order_betweenness_device": ([vp, i64, u64, C.c_int, vp, vp, vp], C.c_int),
            "chl_order_degree": ([vp, u64, C.c_int, vp], C.c_int),
            "chl_order_degree_device": ([vp, u64, C.c_int, C.c_int, vp, vp], C.c_int),
            "chl_build_opts_default": ([C.POINTER(BuildOpts)], None),
            "chl_build": ([vp, vp, i64, C.POINTER(BuildOpts), pp, C.POINTER(BuildStats)], C.c_int),
            "chl_labels_write": ([vp, C.c_char_p], C.c_int),
            "chl_labels_read": ([C.c_char_p, C.c_int, pp], C.c_int),
            "chl_labels_save": ([vp, C.c_char_p], C.c_int),
            "chl_labels_load": ([C.c_char_p, C.c_int, pp], C.c_int),
            "chl_labels_from_csr": ([i64, vp, vp, vp, vp, vp, vp, C.c_int, pp], C.c_int),
            "chl_labels_n": ([vp], i64),
            "chl_labels_total": ([vp, C.c_int], i64),
            "chl_labels_avg": ([vp], C.c_double),
            "chl_labels_copy": ([vp, C.c_int, vp, vp, vp], C.c_int),
            "chl_labels_row_lengths": ([vp, C.c_int, vp], C.c_int),
            "chl_labels_segments": ([vp, C.c_int, C.c_int64, vp, vp, vp, vp], C.c_int),
            "chl_labels_copy_hubs": ([vp, C.c_int, C.c_uint32, C.c_uint32, vp, vp, vp], C.c_int),
            "chl_labels_checksum": ([vp, C.c_int, C.POINTER(u64)], C.c_int),
            "chl_labels_free": ([vp], None),
            "chl_query": ([vp, vp, i64, vp], C.c_int),
            "chl_query_noalloc": ([vp, vp, i64, vp], C.c_int),
            "chl_server_opts_default": ([C.POINTER(ServerOpts)], None),
            "chl_server_create": ([vp, C.POINTER(ServerOpts), pp], C.c_int),
            "chl_server_submit": ([vp, vp, i64, vp, C.POINTER(C.c_uint64)], C.c_int),
            "chl_server_poll": ([vp, u64, C.c_int], C.c_int),
            "chl_server_drain": ([vp], C.c_int),
            "chl_server_stats_get": ([vp, C.POINTER(ServerStats)], C.c_int),
            "chl_server_stats_reset": ([vp], None),
            "chl_server_replay": ([vp, vp, vp, vp, i64, vp], C.c_int),
            "chl_server_free": ([vp], None),
            "chl_query_noalloc_max": ([], i64),
            "chl_query_device": ([vp, vp, i64, vp, vp], C.c_int),
            "chl_query_generate": ([i64, C.c_uint32, vp], C.c_int),
            "chl_gen_local_pairs": ([vp, i64, C.c_int, u64, vp], C.c_int),
            "chl_query_load": ([C.c_char_p, vp, i64, C.POINTER(i64)], C.c_int),
            "chl_query_write": ([C.c_char_p, vp, i64], C.c_int),
            "chl_mlp_param_count": ([], C.c_int),
            "chl_mlp_forward": ([vp, vp, i64, vp], C.c_int),
            "chl_mlp_grad": ([vp, vp, vp, i64, vp, vp], C.c_int),
            "chl_mlp_train": ([vp, vp, vp, i64, C.c_int, C.c_float, vp], C.c_int),
            "chl_mfma_probe": ([vp, vp, vp], C.c_int),
            "chl_mlp_graph_check": ([vp, vp, vp, i64, C.c_int, C.c_float, vp], C.c_int),
        }
        for name, (args, res) in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        L = lib()
        raise ChlError(rc, L.chl_last_error().decode(errors="replace") or L.chl_strerror(rc).decode())


def _ptr(a):
    return a.ctypes.data if a is not None and a.size else None


def header_symbols():
    """Function names declared in include/chl.h."""
    import re
    hdr = os.path.join(_HERE, "..", "include", "chl.h")
    text = open(hdr).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\**\s+\**(chl_[a-z_0-9]+)\s*\(", text, re.M)))


class Graph:
    """A directed graph with the reference's in/out arc ranges (graph.hpp:61-67)."""

    def __init__(self, handle):
        self._h = C.c_void_p(handle) if not isinstance(handle, C.c_void_p) else handle

    def __del__(self):
        if getattr(self, "_h", None) and self._h.value and _lib is not None:
            _lib.chl_graph_free(self._h)
            self._h = C.c_void_p()

    @classmethod
    def load(cls, path, undirected=False):
        h = C.c_void_p()
        fmt = C.c_int(0)
        _check(lib().chl_graph_load(os.fsencode(path), int(undirected), C.byref(h), C.byref(fmt)))
        g = cls(h)
        g.format = fmt.value
        return g

    @classmethod
    def from_arcs(cls, n, src, dst, length, undirected=False):
        src = np.ascontiguousarray(src, np.uint32)
        dst = np.ascontiguousarray(dst, np.uint32)
        length = np.ascontiguousarray(length, np.int32)
        h = C.c_void_p()
        _check(lib().chl_graph_from_arcs(int(n), int(src.size), _ptr(src), _ptr(dst), _ptr(length),
                                         int(undirected), C.byref(h)))
        return cls(h)

    @classmethod
    def grid(cls, rows, cols, wmin=1, wmax=100, seed=1):
        h = C.c_void_p()
        _check(lib().chl_gen_grid(rows, cols, wmin, wmax, seed, C.byref(h)))
        return cls(h)

    @classmethod
    def knn(cls, n, k=3, seed=1):
        h = C.c_void_p()
        _check(lib().chl_gen_knn(n, k, seed, C.byref(h)))
        return cls(h)

    @classmethod
    def rmat(cls, scale, samples, a=0.57, b=0.19, c=0.19, wmax=1, seed=1):
        h = C.c_void_p()
        _check(lib().chl_gen_rmat(scale, samples, a, b, c, wmax, seed, C.byref(h)))
        return cls(h)

    @property
    def n(self):
        return lib().chl_graph_n(self._h)

    @property
    def m(self):
        """Arcs as the input listed them (graph.hpp:314 get_m)."""
        return lib().chl_graph_m(self._h)

    def arcs(self, rng=1):
        return lib().chl_graph_arcs(self._h, rng)

    @property
    def symmetric(self):
        return bool(lib().chl_graph_symmetric(self._h))

    def csr(self, rng=1):
        n = self.n
        m = self.arcs(rng)
        off = np.zeros(n + 1, np.uint64)
        head = np.zeros(max(m, 1), np.uint32)
        ln = np.zeros(max(m, 1), np.int32)
        _check(lib().chl_graph_csr(self._h, rng, _ptr(off), _ptr(head), _ptr(ln)))
        return off, head[:m], ln[:m]

    def arc_list(self, rng=1):
        """(src, dst, len) of one range, src ascending."""
        off, head, ln = self.csr(rng)
        src = np.repeat(np.arange(self.n, dtype=np.uint32), np.diff(off).astype(np.int64))
        return src, head, ln

    def write(self, path):
        _check(lib().chl_graph_write(self._h, os.fsencode(path)))

    def order_betweenness(self, roots, seed=7, with_score=False):
        o = np.zeros(self.n, np.uint32)
        if not with_score:
            _check(lib().chl_order_betweenness(self._h, roots, seed, _ptr(o)))
            return o
        sc = np.zeros(self.n, np.float32)
        _check(lib().chl_order_betweenness_scored(self._h, roots, seed, _ptr(o), _ptr(sc)))
        return o, sc

    def order_betweenness_device(self, roots, seed=7, with_score=False, device=0):
        """The same order (and scores) as order_betweenness, computed on the GPU.
        Returns (order, score or None, device ms)."""
        o = np.zeros(self.n, np.uint32)
        sc = np.zeros(self.n, np.float32) if with_score else None
        ms = C.c_double(0.0)
        _check(lib().chl_order_betweenness_device(self._h, roots, seed, device, _ptr(o),
                                                   _ptr(sc) if with_score else None, C.byref(ms)))
        return o, sc, ms.value

    def order_degree(self, seed=7, reference_ties=False):
        o = np.zeros(self.n, np.uint32)
        _check(lib().chl_order_degree(self._h, seed, int(reference_ties), _ptr(o)))
        return o

    def order_degree_device(self, seed=7, reference_ties=False, device=0):
        """The same order as order_degree, computed on the GPU. Returns (order, device ms)."""
        o = np.zeros(self.n, np.uint32)
        ms = C.c_double(0.0)
        _check(lib().chl_order_degree_device(self._h, seed, int(reference_ties), device, _ptr(o), C.byref(ms)))
        return o, ms.value


def order_load(path):
    n = C.c_int64(0)
    _check(lib().chl_order_load(os.fsencode(path), None, 0, C.byref(n)))
    o = np.zeros(n.value, np.uint32)
    _check(lib().chl_order_load(os.fsencode(path), _ptr(o), n.value, C.byref(n)))
    return o


def order_write(path, order):
    order = np.ascontiguousarray(order, np.uint32)
    _check(lib().chl_order_write(os.fsencode(path), _ptr(order), order.size))


def default_opts(**kw):
    o = BuildOpts()
    lib().chl_build_opts_default(C.byref(o))
    for k, v in kw.items():
        if not hasattr(o, k):
            raise TypeError(f"unknown build option {k}")
        setattr(o, k, v)
    return o


class Labels:
    """Device-resident canonical labels (per vertex, per side, hub-ascending)."""

    def __init__(self, handle, stats=None):
        self._h = handle
        self.stats = stats

    def __del__(self):
        if getattr(self, "_h", None) and self._h.value and _lib is not None:
            _lib.chl_labels_free(self._h)
            self._h = C.c_void_p()

    @property
    def n(self):
        return lib().chl_labels_n(self._h)

    def total(self, side):
        return lib().chl_labels_total(self._h, side)

    @property
    def avg(self):
        return lib().chl_labels_avg(self._h)

    def copy(self, side):
        """(off u64[n+1], hub u32[T], dist i32[T]) of one side, rows by vertex id."""
        n = self.n
        t = self.total(side)
        off = np.zeros(n + 1, np.uint64)
        hub = np.zeros(max(t, 1), np.uint32)
        dist = np.zeros(max(t, 1), np.int32)
        _check(lib().chl_labels_copy(self._h, side, _ptr(off), _ptr(hub), _ptr(dist)))
        return off, hub[:t], dist[:t]

    def write(self, path):
        _check(lib().chl_labels_write(self._h, os.fsencode(path)))

    @classmethod
    def read(cls, path, device=0):
        h = C.c_void_p()
        _check(lib().chl_labels_read(os.fsencode(path), device, C.byref(h)))
        return cls(h)

    def save(self, path):
        """Binary sidecar: the device store as it is."""
        _check(lib().chl_labels_save(self._h, os.fsencode(path)))

    @classmethod
    def load(cls, path, device=0):
        h = C.c_void_p()
        _check(lib().chl_labels_load(os.fsencode(path), device, C.byref(h)))
        return cls(h)

    @classmethod
    def from_csr(cls, n, side0, side1, device=0):
        arrs = []
        for off, hub, dist in (side0, side1):
            arrs += [np.ascontiguousarray(off, np.uint64), np.ascontiguousarray(hub, np.uint32),
                     np.ascontiguousarray(dist, np.int32)]
        h = C.c_void_p()
        _check(lib().chl_labels_from_csr(n, *[_ptr(a) for a in arrs], device, C.byref(h)))
        return cls(h)

    def checksum(self, side):
        """The device-side digest of one side (chl_labels_checksum)."""
        out = C.c_uint64()
        _check(lib().chl_labels_checksum(self._h, side, C.byref(out)))
        return out.value

    def row_lengths(self, side):
        """Entries in each vertex's row of one side."""
        out = np.zeros(self.n, np.uint32)
        _check(lib().chl_labels_row_lengths(self._h, side, _ptr(out)))
        return out

    def copy_hubs(self, side, hub_lo, hub_hi):
        """(off, hub, dist) of the labels of one side with hub in [hub_lo, hub_hi),
        rows by vertex id (a window of the store, copied from the device)."""
        off = np.zeros(self.n + 1, np.uint64)
        _check(lib().chl_labels_copy_hubs(self._h, side, hub_lo, hub_hi, _ptr(off), None, None))
        total = int(off[-1])
        hub = np.zeros(max(total, 1), np.uint32)
        dist = np.zeros(max(total, 1), np.int32)
        if total:
            _check(lib().chl_labels_copy_hubs(self._h, side, hub_lo, hub_hi, _ptr(off), _ptr(hub), _ptr(dist)))
        return off, hub[:total], dist[:total]

    def segments(self, side=0):
        """The store's segments of one side: list of (h0, h1, fmt, entries)."""
        k = lib().chl_labels_segments(self._h, side, 0, None, None, None, None)
        if k < 0:
            _check(k)
        h0, h1, fmt = (np.zeros(max(k, 1), np.uint32) for _ in range(3))
        cnt = np.zeros(max(k, 1), np.uint64)
        _check(min(0, lib().chl_labels_segments(self._h, side, k, _ptr(h0), _ptr(h1), _ptr(fmt), _ptr(cnt))))
        return [(int(h0[i]), int(h1[i]), int(fmt[i]), int(cnt[i])) for i in range(k)]

    def query(self, pairs):
        """d(u -> v) for each row (u, v) of pairs; INFTY when no hub is shared."""
        pairs = np.ascontiguousarray(pairs, np.uint32).reshape(-1, 2)
        out = np.zeros(pairs.shape[0], np.int32)
        _check(lib().chl_query(self._h, _ptr(pairs), pairs.shape[0], _ptr(out)))
        return out

    def query_noalloc(self, pairs):
        """query() through chl_query_noalloc: no device allocation, staging in the
        library's fixed device array (at most chl_query_noalloc_max() pairs)."""
        pairs = np.ascontiguousarray(pairs, np.uint32).reshape(-1, 2)
        out = np.zeros(pairs.shape[0], np.int32)
        _check(lib().chl_query_noalloc(self._h, _ptr(pairs), pairs.shape[0], _ptr(out)))
        return out


class Server:
    """Continuous-batching query server over a label handle (chl_server_*)."""

    def __init__(self, labels, streams=None, max_inflight_pairs=None, max_batch_pairs=None):
        o = ServerOpts()
        lib().chl_server_opts_default(C.byref(o))
        if streams is not None:
            o.streams = streams
        if max_inflight_pairs is not None:
            o.max_inflight_pairs = max_inflight_pairs
        if max_batch_pairs is not None:
            o.max_batch_pairs = max_batch_pairs
        self.opts = o
        self._labels = labels  # kept alive while the server runs
        h = C.c_void_p()
        _check(lib().chl_server_create(labels._h, C.byref(o), C.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().chl_server_free(self._h)
            self._h = None
        self._labels = None  # the label handle may be freed now

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, pairs, out):
        """Queues pairs (contiguous uint32 rows); returns (rc, ticket). out must
        stay alive until the ticket completes."""
        t = C.c_uint64()
        rc = lib().chl_server_submit(self._h, _ptr(pairs), pairs.shape[0], _ptr(out), C.byref(t))
        return rc, t.value

    def poll(self, ticket, wait=True):
        return lib().chl_server_poll(self._h, ticket, 1 if wait else 0)

    def drain(self):
        _check(lib().chl_server_drain(self._h))

    def replay(self, pairs, sizes, arrival_us, out):
        pairs = np.ascontiguousarray(pairs, np.uint32)
        sizes = np.ascontiguousarray(sizes, np.int64)
        arrival_us = np.ascontiguousarray(arrival_us, np.float64)
        _check(lib().chl_server_replay(self._h, _ptr(pairs), _ptr(sizes), _ptr(arrival_us), sizes.size, _ptr(out)))

    def stats(self):
        st = ServerStats()
        _check(lib().chl_server_stats_get(self._h, C.byref(st)))
        return st

    def reset(self):
        lib().chl_server_stats_reset(self._h)


def query_generate(q, vertex_range):
    """q pairs drawn as generate_query does (distributed_memory/query.hpp:44-51)."""
    out = np.zeros((q, 2), np.uint32)
    _check(lib().chl_query_generate(q, vertex_range, _ptr(out)))
    return out


def local_pairs(graph, q, hops, seed=0):
    """q pairs (u, v) with v uniform among the vertices within `hops` out-arc hops of a uniform u."""
    out = np.zeros((q, 2), np.uint32)
    _check(lib().chl_gen_local_pairs(graph._h, q, hops, seed, _ptr(out)))
    return out


def query_load(path):
    """Pairs of a "u v" per line file (load_query, distributed_memory/query.hpp:18-32)."""
    q = C.c_int64()
    _check(lib().chl_query_load(str(path).encode(), None, 0, C.byref(q)))
    out = np.zeros((q.value, 2), np.uint32)
    _check(lib().chl_query_load(str(path).encode(), _ptr(out), q.value, C.byref(q)))
    return out


def query_write(path, dist):
    """One distance per line (write_result, distributed_memory/query.hpp:34-41)."""
    dist = np.ascontiguousarray(dist, np.int32)
    _check(lib().chl_query_write(str(path).encode(), _ptr(dist), dist.size))


def build(graph, order, **opts):
    """Canonical hub labels of (graph, order) built on the GPU."""
    order = np.ascontiguousarray(order, np.uint32)
    o = default_opts(**opts)
    st = BuildStats()
    h = C.c_void_p()
    _check(lib().chl_build(graph._h, _ptr(order), order.size, C.byref(o), C.byref(h), C.byref(st)))
    return Labels(h, st)


def device_count():
    return lib().chl_device_count()


# ---- tree-size predictor (MFMA kernels) -------------------------------------
MLP_F, MLP_H, MLP_O = 8, 64, 2


def mlp_param_count():
    return lib().chl_mlp_param_count()


def mlp_forward(params, x):
    params = np.ascontiguousarray(params, np.float32)
    x = np.ascontiguousarray(x, np.float32).reshape(-1, MLP_F)
    y = np.zeros((x.shape[0], MLP_O), np.float32)
    _check(lib().chl_mlp_forward(_ptr(params), _ptr(x), x.shape[0], _ptr(y)))
    return y


def mlp_grad(params, x, t):
    params = np.ascontiguousarray(params, np.float32)
    x = np.ascontiguousarray(x, np.float32).reshape(-1, MLP_F)
    t = np.ascontiguousarray(t, np.float32).reshape(-1, MLP_O)
    g = np.zeros(mlp_param_count(), np.float32)
    loss = C.c_float(0)
    _check(lib().chl_mlp_grad(_ptr(params), _ptr(x), _ptr(t), x.shape[0], _ptr(g), C.byref(loss)))
    return g, loss.value


def mlp_train(params, x, t, steps, lr=3e-3):
    params = np.array(params, np.float32)
    x = np.ascontiguousarray(x, np.float32).reshape(-1, MLP_F)
    t = np.ascontiguousarray(t, np.float32).reshape(-1, MLP_O)
    curve = np.zeros(max(steps, 1), np.float32)
    _check(lib().chl_mlp_train(_ptr(params), _ptr(x), _ptr(t), x.shape[0], steps, lr, _ptr(curve)))
    return params, curve[:steps]


def mlp_graph_check(params, x, t, replays, lr=3e-3):
    """Largest relative parameter difference between graph-replayed and eager training steps."""
    params = np.ascontiguousarray(params, np.float32)
    x = np.ascontiguousarray(x, np.float32).reshape(-1, MLP_F)
    t = np.ascontiguousarray(t, np.float32).reshape(-1, MLP_O)
    worst = C.c_float(0)
    _check(lib().chl_mlp_graph_check(_ptr(params), _ptr(x), _ptr(t), x.shape[0], replays, lr, C.byref(worst)))
    return worst.value


def mfma_probe(a_bits, b_bits):
    a = np.ascontiguousarray(a_bits, np.uint16).reshape(32, 8)
    b = np.ascontiguousarray(b_bits, np.uint16).reshape(8, 32)
    d = np.zeros((32, 32), np.float32)
    _check(lib().chl_mfma_probe(_ptr(a), _ptr(b), _ptr(d)))
    return d

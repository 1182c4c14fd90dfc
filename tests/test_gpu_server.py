"""Continuous-batching query server: every result equals the same request run
alone (and the oracle), whatever else is in flight; admission refuses by memory."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


class SOpts(C.Structure):
    _fields_ = [("streams", C.c_int), ("max_inflight_pairs", C.c_int64), ("max_batch_pairs", C.c_int64),
                ("reserved", C.c_int * 8)]


class SStats(C.Structure):
    _fields_ = [("completed", C.c_int64), ("pairs", C.c_int64), ("rejected", C.c_int64), ("p50_us", C.c_double),
                ("p99_us", C.c_double), ("mean_us", C.c_double), ("max_us", C.c_double), ("qps", C.c_double),
                ("p90_us", C.c_double), ("p999_us", C.c_double), ("reserved", C.c_double * 6)]


@pytest.fixture(scope="module")
def server_lib(chl):
    lib = chl.lib()
    vp = C.c_void_p
    lib.chl_server_opts_default.argtypes = [C.POINTER(SOpts)]
    lib.chl_server_create.argtypes = [vp, C.POINTER(SOpts), C.POINTER(vp)]
    lib.chl_server_submit.argtypes = [vp, vp, C.c_int64, vp, C.POINTER(C.c_uint64)]
    lib.chl_server_poll.argtypes = [vp, C.c_uint64, C.c_int]
    lib.chl_server_drain.argtypes = [vp]
    lib.chl_server_stats_get.argtypes = [vp, C.POINTER(SStats)]
    lib.chl_server_free.argtypes = [vp]
    return lib


@pytest.fixture(scope="module")
def labels(chl, orc):
    g = chl.Graph.knn(5000, 3, seed=3)
    order = g.order_betweenness(64, seed=2)
    L = chl.build(g, order)
    src, dst, w = g.arc_list(1)
    ref, _ = orc.gll(orc.OGraph.from_arcs(g.n, src, dst, w), order)
    return g, L, ref


def test_mixed_requests_match_alone(chl, server_lib, labels):
    g, L, ref = labels
    lib = server_lib
    o = SOpts()
    lib.chl_server_opts_default(C.byref(o))
    o.streams = 4
    o.max_inflight_pairs = 1 << 16
    o.max_batch_pairs = 1 << 15
    srv = C.c_void_p()
    assert lib.chl_server_create(L._h, C.byref(o), C.byref(srv)) == 0
    rng = np.random.default_rng(5)
    sizes = np.exp(rng.uniform(0, np.log(1 << 15), 120)).astype(np.int64)
    reqs, outs, tickets, collected = [], [], [], set()
    t = C.c_uint64()
    for q in sizes:
        p = rng.integers(0, g.n, (int(q), 2), dtype=np.uint32)
        out = np.full(int(q), -7, np.int32)
        while True:
            rc = lib.chl_server_submit(srv, p.ctypes.data, int(q), out.ctypes.data, C.byref(t))
            if rc == 0:
                break
            assert rc == -9  # busy: admission by memory; wait for the newest request, then retry
            if tickets and tickets[-1] not in collected:
                assert lib.chl_server_poll(srv, tickets[-1], 1) == 1
                collected.add(tickets[-1])
        reqs.append(p)
        outs.append(out)
        tickets.append(t.value)
    assert lib.chl_server_drain(srv) == 0
    for tk in tickets:
        if tk not in collected:
            assert lib.chl_server_poll(srv, tk, 1) == 1
    st = SStats()
    lib.chl_server_stats_get(srv, C.byref(st))
    assert st.completed == len(sizes) and st.pairs == int(sizes.sum())
    assert 0 < st.p50_us <= st.p99_us <= st.max_us
    lib.chl_server_free(srv)
    for p, out in zip(reqs, outs):
        alone = L.query(p)
        assert np.array_equal(out, alone)
        assert np.array_equal(out, ref.query(p))


def test_admission_refuses_oversize(chl, server_lib, labels):
    g, L, _ = labels
    lib = server_lib
    o = SOpts()
    lib.chl_server_opts_default(C.byref(o))
    o.max_batch_pairs = 1024
    o.max_inflight_pairs = 2048
    srv = C.c_void_p()
    assert lib.chl_server_create(L._h, C.byref(o), C.byref(srv)) == 0
    p = np.zeros((2000, 2), np.uint32)
    out = np.zeros(2000, np.int32)
    t = C.c_uint64()
    assert lib.chl_server_submit(srv, p.ctypes.data, 2000, out.ctypes.data, C.byref(t)) == -8
    bad = np.array([[0, g.n]], np.uint32)
    assert lib.chl_server_submit(srv, bad.ctypes.data, 1, out.ctypes.data, C.byref(t)) == -8
    lib.chl_server_free(srv)


def test_replay_trace_matches_alone(chl, server_lib, labels):
    # an open-loop trace (arrivals 20 us apart, sizes log-uniform, a fifth of the
    # pairs within 5 hops) replayed by the native submitter: same distances as
    # each request alone
    g, L, ref = labels
    lib = server_lib
    lib.chl_server_replay.argtypes = [C.c_void_p] * 4 + [C.c_int64, C.c_void_p]
    o = SOpts()
    lib.chl_server_opts_default(C.byref(o))
    o.streams = 3
    o.max_inflight_pairs = 1 << 14
    o.max_batch_pairs = 1 << 12
    srv = C.c_void_p()
    assert lib.chl_server_create(L._h, C.byref(o), C.byref(srv)) == 0
    rng = np.random.default_rng(9)
    sizes = np.exp(rng.uniform(0, np.log(1 << 12), 300)).astype(np.int64)
    pairs = rng.integers(0, g.n, (int(sizes.sum()), 2), dtype=np.uint32)
    pick = rng.random(pairs.shape[0]) < 0.2
    pairs[pick] = chl.local_pairs(g, int(pick.sum()), 5, seed=2)
    arrivals = np.arange(sizes.size, dtype=np.float64) * 20.0
    out = np.full(pairs.shape[0], -7, np.int32)
    assert lib.chl_server_replay(srv, pairs.ctypes.data, sizes.ctypes.data, arrivals.ctypes.data, sizes.size,
                                 out.ctypes.data) == 0
    st = SStats()
    lib.chl_server_stats_get(srv, C.byref(st))
    assert st.completed == sizes.size and st.pairs == pairs.shape[0]
    lib.chl_server_free(srv)
    assert np.array_equal(out, L.query(pairs))
    assert np.array_equal(out, ref.query(pairs))

"""Continuous-batching query server: every result equals the same request run
alone (and the oracle), whatever else is in flight; admission refuses by memory."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def labels(chl, orc):
    g = chl.Graph.knn(5000, 3, seed=3)
    order = g.order_betweenness(64, seed=2)
    L = chl.build(g, order)
    src, dst, w = g.arc_list(1)
    ref, _ = orc.gll(orc.OGraph.from_arcs(g.n, src, dst, w), order)
    return g, L, ref


def test_mixed_requests_match_alone(chl, labels):
    g, L, ref = labels
    srv = chl.Server(L, streams=4, max_inflight_pairs=1 << 16, max_batch_pairs=1 << 15)
    rng = np.random.default_rng(5)
    sizes = np.exp(rng.uniform(0, np.log(1 << 15), 120)).astype(np.int64)
    reqs, outs, tickets, collected = [], [], [], set()
    for q in sizes:
        p = rng.integers(0, g.n, (int(q), 2), dtype=np.uint32)
        out = np.full(int(q), -7, np.int32)
        while True:
            rc, t = srv.submit(p, out)
            if rc == 0:
                break
            assert rc == -9  # busy: admission by memory; wait for the newest request, then retry
            if tickets and tickets[-1] not in collected:
                assert srv.poll(tickets[-1]) == 1
                collected.add(tickets[-1])
        reqs.append(p)
        outs.append(out)
        tickets.append(t)
    srv.drain()
    for tk in tickets:
        if tk not in collected:
            assert srv.poll(tk) == 1
    st = srv.stats()
    assert st.completed == len(sizes) and st.pairs == int(sizes.sum())
    assert 0 < st.p50_us <= st.p99_us <= st.max_us
    srv.close()
    for p, out in zip(reqs, outs):
        alone = L.query(p)
        assert np.array_equal(out, alone)
        assert np.array_equal(out, ref.query(p))


def test_admission_refuses_oversize(chl, labels):
    g, L, _ = labels
    srv = chl.Server(L, max_batch_pairs=1024, max_inflight_pairs=2048)
    p = np.zeros((2000, 2), np.uint32)
    out = np.zeros(2000, np.int32)
    assert srv.submit(p, out)[0] == -8
    bad = np.array([[0, g.n]], np.uint32)
    assert srv.submit(bad, out)[0] == -8
    srv.close()


def test_replay_trace_matches_alone(chl, labels):
    # an open-loop trace (arrivals 20 us apart, sizes log-uniform, a fifth of the
    # pairs within 5 hops) replayed by the native submitter: same distances as
    # each request alone
    g, L, ref = labels
    srv = chl.Server(L, streams=3, max_inflight_pairs=1 << 14, max_batch_pairs=1 << 12)
    rng = np.random.default_rng(9)
    sizes = np.exp(rng.uniform(0, np.log(1 << 12), 300)).astype(np.int64)
    pairs = rng.integers(0, g.n, (int(sizes.sum()), 2), dtype=np.uint32)
    pick = rng.random(pairs.shape[0]) < 0.2
    pairs[pick] = chl.local_pairs(g, int(pick.sum()), 5, seed=2)
    arrivals = np.arange(sizes.size, dtype=np.float64) * 20.0
    out = np.full(pairs.shape[0], -7, np.int32)
    srv.replay(pairs, sizes, arrivals, out)
    st = srv.stats()
    assert st.completed == sizes.size and st.pairs == pairs.shape[0]
    # latency counts from the scheduled arrival: never below the submission lag
    assert st.lag_max_us >= 0 and st.max_us >= st.lag_p99_us
    srv.close()
    assert np.array_equal(out, L.query(pairs))
    assert np.array_equal(out, ref.query(pairs))


def c5_trace(chl, g, count, rate_qps, seed):
    """The C5 serving workload: Poisson arrivals of requests whose sizes are
    log-uniform in [1, 65536]; 80% uniform pairs and 20% with both endpoints
    within 50 hops of each other. Returns (pairs, sizes, arrival times in us)."""
    rng = np.random.default_rng(seed)
    sizes = np.floor(np.exp(rng.uniform(0, np.log(65536), count))).astype(np.int64)
    total = int(sizes.sum())
    pairs = rng.integers(0, g.n, (total, 2), dtype=np.uint32)
    pick = rng.random(total) < 0.2
    pairs[pick] = chl.local_pairs(g, int(pick.sum()), 50, seed=seed + 1)
    gaps = rng.exponential(1.0 / (rate_qps / sizes.mean()), count) * 1e6
    return pairs, sizes, np.cumsum(gaps) - gaps[0]


def test_c5_mix_on_c2_labels(chl, orc):
    """The C5 mix replayed on the labels of the 1M-vertex road-like graph (3-NN,
    256-root sampled-betweenness order): every result equals chl_query on the same
    pairs, and a sample equals the oracle's plain Dijkstra."""
    g = chl.Graph.knn(1_000_000, 3, seed=1)
    order = g.order_betweenness(256, seed=7)
    L = chl.build(g, order)
    pairs, sizes, arrivals = c5_trace(chl, g, 400, 1.0e8, seed=11)
    srv = chl.Server(L, streams=4, max_inflight_pairs=1 << 20, max_batch_pairs=65536)
    out = np.full(pairs.shape[0], -7, np.int32)
    srv.replay(pairs, sizes, arrivals, out)
    st = srv.stats()
    assert st.completed == sizes.size and st.pairs == pairs.shape[0]
    print(f"C5 on C2 labels: {st.pairs} pairs, p50 {st.p50_us:.0f} us, p99 {st.p99_us:.0f} us, "
          f"rejected {st.rejected}, late {st.late}", flush=True)
    srv.close()
    assert np.array_equal(out, L.query(pairs))
    src, dst, w = g.arc_list(1)
    og = orc.OGraph.from_arcs(g.n, src, dst, w)
    rng = np.random.default_rng(4)
    pick = rng.choice(pairs.shape[0], 20000, replace=False)
    sources = np.unique(pairs[pick, 0])[:8]
    D = og.dijkstra_multi(sources.astype(np.uint32))
    for i, s in enumerate(sources):
        rows = np.flatnonzero(pairs[:, 0] == s)
        assert np.array_equal(out[rows], D[i][pairs[rows, 1]])

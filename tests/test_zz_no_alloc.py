"""chl_query_noalloc (staging in the library's fixed device array, no device
allocation) against chl_query on the same labels: two calls in a row, the second
smaller, and a call beyond the stated maximum refused."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_noalloc_query_matches_query(chl):
    g = chl.Graph.knn(20000, 3, seed=31)
    order = g.order_betweenness(32, seed=5)
    L = chl.build(g, order)
    rng = np.random.default_rng(7)
    cap = int(chl.lib().chl_query_noalloc_max())
    assert cap >= 1 << 16
    for q in (min(cap, 300000), 777):  # the second call is smaller: stale staging must not leak into it
        pairs = rng.integers(0, g.n, (q, 2)).astype(np.uint32)
        want = L.query(pairs)
        got = L.query_noalloc(pairs)
        assert np.array_equal(got, want), q
    # the same rows with an out-of-range id and an over-long call are refused
    with pytest.raises(chl.ChlError):
        L.query_noalloc(np.array([[0, g.n]], np.uint32))
    with pytest.raises(chl.ChlError):
        L.query_noalloc(np.zeros((cap + 1, 2), np.uint32))
    # a directed graph (both sides differ) and a one-vertex graph
    r = np.random.default_rng(3)
    d = chl.Graph.from_arcs(300, r.integers(0, 300, 1500).astype(np.uint32), r.integers(0, 300, 1500).astype(np.uint32),
                            r.integers(1, 50, 1500).astype(np.int32))
    D = chl.build(d, r.permutation(300).astype(np.uint32))
    pairs = r.integers(0, 300, (5000, 2)).astype(np.uint32)
    assert np.array_equal(D.query_noalloc(pairs), D.query(pairs))
    one = chl.build(chl.Graph.from_arcs(1, np.zeros(0, np.uint32), np.zeros(0, np.uint32), np.zeros(0, np.int32)),
                    np.zeros(1, np.uint32))
    assert one.query_noalloc(np.zeros((1, 2), np.uint32)).tolist() == [0]

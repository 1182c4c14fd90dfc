"""The sampled-betweenness order computed on the GPU equals the host order
(gen.cpp order_betweenness, pinned in test_host.py against scipy's Dijkstra and a
closed form): the same scores and the same tie-broken order."""
import numpy as np
import pytest

pytestmark = [
    pytest.mark.gpu,
    pytest.mark.skip(reason="device order switched off: its first GPU run ended with a device fault whose "
                     "cause is not found (chl_order_betweenness_device returns CHL_E_STATE)"),
]


def check_same(g, roots, seed):
    want, wscore = g.order_betweenness(roots, seed=seed, with_score=True)
    got, gscore, ms = g.order_betweenness_device(roots, seed=seed, with_score=True)
    assert np.array_equal(gscore, wscore)
    assert np.array_equal(got, want)
    assert ms > 0
    return ms


def test_grid_c1_order(chl):
    # the C1 order: 100x100 grid, 64 roots
    g = chl.Graph.grid(100, 100, 1, 100, seed=1)
    check_same(g, 64, 1)


@pytest.mark.parametrize("roots", [1, 17, 300])
def test_knn(chl, roots):
    g = chl.Graph.knn(20000, 3, seed=4)
    check_same(g, roots, 9)


def test_rmat_unit_weights_ties(chl):
    # unit weights: many shortest-path predecessors, so the id tie rule decides
    g = chl.Graph.rmat(13, 3 * 8192, seed=3)
    check_same(g, 128, 2)


def test_directed_and_disconnected(chl):
    rng = np.random.default_rng(8)
    n, m = 700, 2500
    src = rng.integers(0, n, m).astype(np.uint32)
    dst = rng.integers(0, n, m).astype(np.uint32)
    w = rng.integers(1, 6, m).astype(np.int32)
    g = chl.Graph.from_arcs(n, src, dst, w)
    assert not g.symmetric
    check_same(g, 200, 5)
    check_same(g, 5000, 5)  # more roots than vertices: every vertex is a root
    src = np.array([0, 1, 2, 4, 5, 5], np.uint32)
    dst = np.array([1, 2, 0, 5, 6, 5], np.uint32)
    g = chl.Graph.from_arcs(8, src, dst, np.array([3, 4, 5, 1, 2, 9], np.int32), undirected=True)
    check_same(g, 8, 1)
    g1 = chl.Graph.from_arcs(1, np.zeros(0, np.uint32), np.zeros(0, np.uint32), np.zeros(0, np.int32))
    o, s, _ = g1.order_betweenness_device(4, seed=1, with_score=True)
    assert list(o) == [0] and list(s) == [0]


def test_c2_road_1M_order(chl):
    # the C2 order at its own size: 1M-point 3-NN graph, 256 roots
    g = chl.Graph.knn(1_000_000, 3, seed=1)
    ms = check_same(g, 256, 7)
    print(f"device order 1M/256 roots: {ms:.1f} ms")

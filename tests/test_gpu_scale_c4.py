"""C4 at its own size: 16M points uniform in the unit square, 3-NN symmetrised
(about 30M undirected edges), w = round(1e4 d) + 1, the 1024-root sampled-
betweenness order computed on the device, PLaNT trees under learned batching,
the whole graph and label set resident in one GPU's memory.

Parity at this size: the device order equals the host order (sha256 of the order
file the host order gives, tests/golden/fullsize.json); 1e5 seeded query pairs
equal the oracle's plain Dijkstra; and the labels whose hub lies in whole rank
windows equal, row by row, the oracle's PLaNT trees of those ranks
(distributed_memory/akiba.hpp:52-104 restated: PLaNT trees are independent, so a
window's oracle labels are exactly the canonical labels with hub in it). Each
test prints what it checked, and every test and fixture ends within about two
minutes."""
import hashlib
import json
import os
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
C4 = json.load(open(os.path.join(GOLD, "fullsize.json")))["c4"]


def say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def c4_graph(chl):
    t = time.time()
    g = chl.Graph.knn(16_000_000, 3, seed=1)
    say(f"C4 graph: n={g.n} arcs={g.arcs(1)} in {time.time() - t:.1f} s")
    order, score, ms = g.order_betweenness_device(1024, seed=7, with_score=True)
    say(f"C4 device order, 1024 roots: {ms:.0f} ms")
    return g, order, score


@pytest.fixture(scope="module")
def c4_labels(chl, c4_graph):
    g, order, score = c4_graph
    t = time.time()
    L = chl.build(g, order, batching=chl.BATCH_LEARNED, vertex_score=score.ctypes.data)
    say(f"C4 build: {time.time() - t:.1f} s, {L.stats.labels} labels, {L.stats.segments} segments, "
        f"{L.stats.store_bytes / 1e9:.1f} GB")
    yield L
    del L


def test_c4_device_order_is_the_host_order(chl, c4_graph, tmp_path):
    g, order, _ = c4_graph
    assert g.n == C4["n"]
    op = tmp_path / "c4.order"
    chl.order_write(str(op), order)
    assert hashlib.sha256(op.read_bytes()).hexdigest() == C4["order_sha256"]


def test_c4_build_fits_one_gpu(c4_labels):
    L = c4_labels
    for side in (0, 1):
        assert L.total(side) == C4["count"][side]
        assert L.checksum(side) == int(C4["digest"][side], 16)
    assert L.stats.store_bytes < 288e9
    segs = L.segments(0)
    assert all(f == 5 for _, _, f, _ in segs)  # distances stay below 65536: 5-byte entries


@pytest.fixture(scope="module")
def c4_oracle_graph(orc, c4_graph):
    g, _, _ = c4_graph
    src, dst, w = g.arc_list(1)
    return orc.OGraph.from_arcs(g.n, src, dst, w)


def test_c4_queries_match_dijkstra(c4_graph, c4_labels, c4_oracle_graph):
    g, _, _ = c4_graph
    og = c4_oracle_graph
    rng = np.random.default_rng(3)
    S = rng.integers(0, g.n, 20).astype(np.uint32)
    D = og.dijkstra_multi(S, threads=16)
    T = rng.integers(0, g.n, (20, 5000)).astype(np.uint32)
    pairs = np.stack([np.repeat(S, 5000), T.reshape(-1)], 1).astype(np.uint32)
    want = np.concatenate([D[i][T[i]] for i in range(20)])
    got = c4_labels.query(pairs)
    say(f"C4 queries: {pairs.shape[0]} pairs, {(want == 2147483647).sum()} unreachable")
    assert np.array_equal(got, want)
    v = rng.integers(0, g.n, 1000).astype(np.uint32)
    assert np.all(c4_labels.query(np.stack([v, v], 1)) == 0)


# one test per window, so that each reports within about a minute
@pytest.mark.parametrize("window", [(0, 8), (5000, 5064), (1_000_000, 1_001_024), (-1024, 0)])
def test_c4_rank_window_matches_plant(orc, c4_graph, c4_labels, c4_oracle_graph, window):
    g, order, _ = c4_graph
    n = g.n
    r0, r1 = (n + window[0], n) if window[0] < 0 else window
    t = time.time()
    got = c4_labels.copy_hubs(0, r0, r1)
    ref, _, _ = orc.plant_range(c4_oracle_graph, order, r0, r1, threads=16)
    want = ref.copy(0)
    say(f"C4 rank window [{r0}, {r1}): {got[1].size} labels ({time.time() - t:.1f} s)")
    assert got[1].size > 0
    for a, b, what in zip(got, want, ("off", "hub", "dist")):
        assert np.array_equal(np.asarray(a), np.asarray(b)), (r0, r1, what)

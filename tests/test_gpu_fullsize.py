"""Parity at the configurations' own sizes: the reference's label file for the
100x100 grid (byte-identical, every mode); at 1M vertices, the whole label set of
both sides against the reference's own output for the same graph and order (entry
counts and digests of its label file, tests/golden/fullsize.json) under both
batching policies, query distances against the oracle's plain Dijkstra on sampled
pairs, and the labels of whole hub-rank windows bit for bit against the oracle's
PLaNT trees of those ranks (independent trees, so a window's oracle labels are
exactly the canonical labels whose hub lies in it). The digest itself is checked
against a host computation over the oracle's labels, and label_digest (which made
the fixtures) against the same host computation in tests/test_oracle.py."""
import hashlib
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
MANIFEST = json.load(open(os.path.join(GOLD, "manifest.json")))
# per-side entry counts and digests of the reference's own label files for the
# full-size inputs (tests/golden/make_fullsize.py: shared_memory/hhl run on the
# same graph and order files, its -l output streamed through label_digest)
FULLSIZE = json.load(open(os.path.join(GOLD, "fullsize.json")))


def whole_set_check(chl, L, name, g, order, tmp_path):
    """Every label of both sides against the reference's own output, by count
    and by the order-independent digest of chl_labels_checksum."""
    f = FULLSIZE[name]
    op = tmp_path / f"{name}.order"
    chl.order_write(str(op), order)
    assert hashlib.sha256(op.read_bytes()).hexdigest() == f["order_sha256"]
    assert g.n == f["n"]
    for side in (0, 1):
        assert L.total(side) == f["count"][side], (name, side)
        assert L.checksum(side) == int(f["digest"][side], 16), (name, side)


def host_digest(off, hub, dist):
    """chl_labels_checksum restated in numpy (uint64 arithmetic wraps mod 2^64)."""
    off = np.asarray(off, np.int64)
    v = np.repeat(np.arange(off.size - 1, dtype=np.uint64), np.diff(off))
    z = (np.asarray(hub, np.uint64) << np.uint64(32)) | np.asarray(dist, np.int32).astype(np.uint32).astype(np.uint64)
    z ^= v * np.uint64(0x9E3779B97F4A7C15)
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    z ^= z >> np.uint64(31)
    return int(z.sum(dtype=np.uint64))


def oracle_dijkstra_check(orc, g, L, sources=20, targets=5000, seed=3):
    """sources x targets = 1e5 seeded pairs against the oracle's plain Dijkstra."""
    src, dst, w = g.arc_list(1)
    og = orc.OGraph.from_arcs(g.n, src, dst, w)
    rng = np.random.default_rng(seed)
    S = rng.integers(0, g.n, sources).astype(np.uint32)
    D = og.dijkstra_multi(S)
    T = rng.integers(0, g.n, (sources, targets)).astype(np.uint32)
    pairs = np.stack([np.repeat(S, targets), T.reshape(-1)], 1).astype(np.uint32)
    want = np.concatenate([D[i][T[i]] for i in range(sources)])
    assert np.array_equal(L.query(pairs), want)
    v = rng.integers(0, g.n, 1000).astype(np.uint32)
    assert np.all(L.query(np.stack([v, v], 1)) == 0)


def rank_window_check(orc, g, order, L, windows, threads=16):
    """For each hub-rank window [r0, r1): the GPU labels with hub in the window
    equal, row by row, the labels of the oracle's PLaNT trees of those ranks
    (distributed_memory/akiba.hpp:52-104 restated)."""
    src, dst, w = g.arc_list(1)
    og = orc.OGraph.from_arcs(g.n, src, dst, w)
    for r0, r1 in windows:
        got = L.copy_hubs(0, r0, r1)
        ref, _, _ = orc.plant_range(og, order, r0, r1, threads=threads)
        print(f"rank window [{r0}, {r1}): {got[1].size} labels", flush=True)
        assert got[1].size > 0
        for a, b, what in zip(got, ref.copy(0), ("off", "hub", "dist")):
            assert np.array_equal(np.asarray(a), np.asarray(b)), (r0, r1, what)


def test_device_digest_matches_host(chl, orc):
    g = chl.Graph.knn(4000, 3, seed=2)
    order = g.order_betweenness(32, seed=1)
    L = chl.build(g, order)
    src, dst, w = g.arc_list(1)
    ref, _ = orc.gll(orc.OGraph.from_arcs(g.n, src, dst, w), order, threads=8)
    for side in (0, 1):
        assert L.checksum(side) == host_digest(*ref.copy(side))
    # a planted change moves the digest
    off, hub, dist = ref.copy(0)
    dist = dist.copy()
    dist[len(dist) // 2] += 1
    assert L.checksum(0) != host_digest(off, hub, dist)


def test_c1_grid100_label_file_is_the_reference_file(chl, orc, tmp_path):
    # C1: the label file the reference's own hhl wrote for this graph and order
    # (sha256 in the manifest), from the graph file read as hhl reads it
    m = MANIFEST["grid100"]
    g = chl.Graph.grid(100, 100, 1, 100, seed=1)
    order = g.order_betweenness(64, seed=1)
    gp = tmp_path / "g.gr"
    g.write(str(gp))
    assert hashlib.sha256(gp.read_bytes()).hexdigest() == m["graph_sha256"]
    g2 = chl.Graph.load(str(gp))
    for mode in (chl.MODE_PLANT, chl.MODE_GLL, chl.MODE_HYBRID):
        for batching in (chl.BATCH_STATIC, chl.BATCH_LEARNED):
            L = chl.build(g2, order, mode=mode, batching=batching)
            lp = tmp_path / f"l{mode}{batching}.txt"
            L.write(str(lp))
            assert hashlib.sha256(lp.read_bytes()).hexdigest() == m["labels_sha256"], (mode, batching)
    # 1e5 uniform random pairs against all-pairs Dijkstra
    src, dst, w = g2.arc_list(1)
    D = orc.OGraph.from_arcs(g2.n, src, dst, w).dijkstra_multi(np.arange(g2.n, dtype=np.uint32), threads=16)
    pairs = np.random.default_rng(5).integers(0, g2.n, (100000, 2)).astype(np.uint32)
    assert np.array_equal(L.query(pairs), D[pairs[:, 0], pairs[:, 1]])


def test_c2_road_1M_whole_set_and_dijkstra(chl, orc, tmp_path):
    # C2 at its own size: 1M points, 3-NN, 256-root sampled-betweenness order
    g = chl.Graph.knn(1_000_000, 3, seed=1)
    order = g.order_betweenness(256, seed=7)
    learned = chl.build(g, order, batching=chl.BATCH_LEARNED)
    whole_set_check(chl, learned, "c2", g, order, tmp_path)
    static = chl.build(g, order, batching=chl.BATCH_STATIC)
    whole_set_check(chl, static, "c2", g, order, tmp_path)
    oracle_dijkstra_check(orc, g, learned)
    del static
    n = g.n
    rank_window_check(orc, g, order, learned, [(0, 48), (3000, 3128), (300000, 300512), (n - 1024, n)])


@pytest.mark.parametrize("wmax", [1, 16])
def test_c3_rmat20_hybrid_whole_set_and_dijkstra(chl, orc, wmax, tmp_path):
    # C3 at its own size: R-MAT scale 20, edge factor 3, degree order, hybrid
    g = chl.Graph.rmat(20, 3 << 20, wmax=wmax, seed=1)
    order = g.order_degree(seed=2)
    learned = chl.build(g, order, mode=chl.MODE_HYBRID, batching=chl.BATCH_LEARNED)
    whole_set_check(chl, learned, "c3u" if wmax == 1 else "c3w", g, order, tmp_path)
    static = chl.build(g, order, mode=chl.MODE_HYBRID, batching=chl.BATCH_STATIC, switch_ratio=16.0)
    whole_set_check(chl, static, "c3u" if wmax == 1 else "c3w", g, order, tmp_path)
    oracle_dijkstra_check(orc, g, learned)
    # windows on both sides of the switch to pruned trees
    n, sr = g.n, max(int(learned.stats.switch_rank), 0)
    windows = [(0, 32), (sr, sr + 256), (400000, 400512), (n - 1024, n)]
    rank_window_check(orc, g, order, learned, windows)

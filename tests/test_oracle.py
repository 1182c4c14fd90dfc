"""Pins the CPU oracle (oracle/chl_oracle.cpp) before it is trusted as the checker.

* byte-identical label files against the reference's own GLL (shared_memory/hhl.cpp
  compiled from source), committed as fixtures under tests/golden/;
* the canonical-hub-labeling definition by brute force (APSP) on small graphs;
* closed forms: path with midpoint-recursive order, star with centre first;
* distances against scipy's Dijkstra and all-pairs Floyd-Warshall;
* the PLaNT and hybrid restatements against GLL (the labeling is unique);
* planted mistakes in labels are caught by the same checks.
"""
import hashlib
import subprocess
import json
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
MANIFEST = json.load(open(os.path.join(GOLD, "manifest.json")))
SMALL = [k for k in MANIFEST if k != "grid100"]


def labels_of(L):
    return [L.copy(s) for s in (0, 1)]


def same(a, b):
    return all(np.array_equal(x, y) for sa, sb in zip(a, b) for x, y in zip(sa, sb))


@pytest.mark.parametrize("name", SMALL)
def test_oracle_matches_reference_gll(orc, name, tmp_path):
    g, fmt = orc.OGraph.read(os.path.join(GOLD, f"{name}.gr"))
    assert fmt == 1
    order = orc.order_read(os.path.join(GOLD, f"{name}.order"))
    L, _ = orc.gll(g, order, threads=4)
    out = tmp_path / "o.labels"
    assert L.write(str(out)) == 0
    assert out.read_bytes() == open(os.path.join(GOLD, f"{name}.labels"), "rb").read()


def test_oracle_matches_reference_grid100(chl, orc, tmp_path):
    m = MANIFEST["grid100"]
    g = chl.Graph.grid(100, 100, 1, 100, seed=1)
    order = g.order_betweenness(64, seed=1)
    gp, op = tmp_path / "g.gr", tmp_path / "g.order"
    g.write(str(gp))
    chl.order_write(str(op), order)
    assert hashlib.sha256(gp.read_bytes()).hexdigest() == m["graph_sha256"]
    assert hashlib.sha256(op.read_bytes()).hexdigest() == m["order_sha256"]
    og, _ = orc.OGraph.read(str(gp))
    L, _ = orc.gll(og, order, threads=8)
    lp = tmp_path / "g.labels"
    L.write(str(lp))
    assert hashlib.sha256(lp.read_bytes()).hexdigest() == m["labels_sha256"]


@pytest.mark.parametrize("name", SMALL)
def test_reference_labels_are_the_chl_definition(orc, name):
    g, _ = orc.OGraph.read(os.path.join(GOLD, f"{name}.gr"))
    order = orc.order_read(os.path.join(GOLD, f"{name}.order"))
    ref = orc.OLabels.read(os.path.join(GOLD, f"{name}.labels"))
    assert same(labels_of(orc.brute_chl(g, order)), labels_of(ref))


def small_graphs(chl):
    rng = np.random.default_rng(3)
    yield chl.Graph.grid(9, 9, 1, 20, seed=2), None
    yield chl.Graph.knn(150, 3, seed=4), None
    n, m = 70, 260
    src, dst = rng.integers(0, n, m).astype(np.uint32), rng.integers(0, n, m).astype(np.uint32)
    w = rng.integers(1, 6, m).astype(np.int32)  # small weights: many equal-length paths
    yield chl.Graph.from_arcs(n, src, dst, w), (src, dst, w)
    yield chl.Graph.rmat(7, 400, seed=9), None


def oracle_graph(orc, g, raw):
    if raw is None:
        raw = g.arc_list(1)
    return orc.OGraph.from_arcs(g.n, *raw)


@pytest.mark.parametrize("k", range(4))
def test_gll_plant_hybrid_equal_brute_force(chl, orc, k):
    g, raw = list(small_graphs(chl))[k]
    og = oracle_graph(orc, g, raw)
    order = np.random.default_rng(k).permutation(g.n).astype(np.uint32)
    brute = labels_of(orc.brute_chl(og, order))
    gll, _ = orc.gll(og, order, threads=4, sf=1)  # sf=1: many supersteps, exercises cleaning
    assert same(labels_of(gll), brute)
    plant, ts, lc = orc.plant_range(og, order, 0, g.n, threads=4)
    assert same(labels_of(plant), brute)
    assert int(lc.sum()) == sum(len(s[1]) for s in brute)
    hyb, st = orc.hybrid(og, order, threads=4, ps=2.0)
    assert same(labels_of(hyb), brute)


def cycle_graph(orc, n):
    src = np.arange(n, dtype=np.uint32)
    dst = (src + 1) % n
    return orc.OGraph.from_arcs(n, np.r_[src, dst], np.r_[dst, src], np.ones(2 * n, np.int32))


def test_hybrid_switch_ratio_is_a_float(orc):
    """hl::run takes ps as a float (distributed_memory/akiba.hpp:190) and widens it
    for the ratio test (akiba.hpp:206, 301), so ps = 1 - 1e-12 is 1.0 there. On a
    cycle the top root's PLaNT tree labels every vertex it settles: its tree/label
    ratio is exactly 1, which passes ps = 0.5 but neither 1.0 nor 1 - 1e-12. With
    one thread a superstep is one root, so the switch root tells which root passed."""
    n = 12
    g = cycle_graph(orc, n)
    order = np.random.default_rng(3).permutation(n).astype(np.uint32)
    _, ts, lc = orc.plant_range(g, order, 0, n, threads=1)
    ratio = (ts[0::2].astype(np.float64) + ts[1::2]) / (lc[0::2].astype(np.float64) + lc[1::2])
    assert ratio[0] == 1.0 and ratio[1] > 1.0
    switch = {ps: int(orc.hybrid(g, order, threads=1, ps=ps)[1][1]) for ps in (0.5, 1.0, 1 - 1e-12)}
    assert switch[0.5] == 1  # the top root passes
    assert switch[1.0] == 2  # the next root passes
    assert switch[1 - 1e-12] == switch[1.0]


def test_path_midpoint_order_closed_form(chl, orc):
    # path 0-1-...-14 with unit weights; order = midpoint recursion. The hubs of v
    # are exactly v's ancestors in the recursion tree (and v itself).
    n = 15
    src = np.arange(n - 1, dtype=np.uint32)
    g = orc.OGraph.from_arcs(n, np.r_[src, src + 1], np.r_[src + 1, src], np.ones(2 * (n - 1), np.int32))
    order, parent = [], {}

    def rec(lo, hi, p):
        if lo > hi:
            return
        mid = (lo + hi) // 2
        order.append(mid)
        parent[mid] = p
        rec(lo, mid - 1, mid)
        rec(mid + 1, hi, mid)

    rec(0, n - 1, None)
    # order must be by recursion depth for canonicity: breadth-first over the tree
    depth = {}
    for v in order:
        depth[v] = 0 if parent[v] is None else depth[parent[v]] + 1
    order = sorted(order, key=lambda v: (depth[v], v))
    rank = {v: i for i, v in enumerate(order)}
    gll, _ = orc.gll(g, np.array(order, np.uint32), threads=2)
    plant, _, _ = orc.plant_range(g, np.array(order, np.uint32), 0, n, threads=2)
    for L in (gll, plant):
        off, hub, dist = L.copy(0)
        for v in range(n):
            anc, a = [], v
            while a is not None:
                anc.append(a)
                a = parent[a]
            want = sorted((rank[a], abs(a - v)) for a in anc)
            got = list(zip(hub[off[v]:off[v + 1]].tolist(), dist[off[v]:off[v + 1]].tolist()))
            assert got == want


def test_star_closed_form(orc):
    n = 20
    leaves = np.arange(1, n, dtype=np.uint32)
    w = np.arange(1, n, dtype=np.int32)
    g = orc.OGraph.from_arcs(n, np.r_[np.zeros(n - 1, np.uint32), leaves], np.r_[leaves, np.zeros(n - 1, np.uint32)],
                             np.r_[w, w])
    order = np.arange(n, dtype=np.uint32)
    gll, _ = orc.gll(g, order, threads=2)
    plant, _, _ = orc.plant_range(g, order, 0, n, threads=2)
    for L in (gll, plant):
        for side in (0, 1):
            off, hub, dist = L.copy(side)
            assert list(hub[off[0]:off[1]]) == [0]
            for v in range(1, n):
                assert list(zip(hub[off[v]:off[v + 1]], dist[off[v]:off[v + 1]])) == [(0, v), (v, 0)]


def test_dijkstra_matches_scipy_and_queries(chl, orc):
    from scipy.sparse import csr_matrix
    from scipy.sparse.csgraph import dijkstra as sp_dijkstra
    for g, raw in small_graphs(chl):
        og = oracle_graph(orc, g, raw)
        off, head, ln = g.csr(1)
        A = csr_matrix((ln.astype(np.float64), head.astype(np.int64), off.astype(np.int64)), shape=(g.n, g.n))
        D = sp_dijkstra(A, directed=True)
        mine = og.dijkstra_multi(np.arange(g.n, dtype=np.uint32)).astype(np.float64)
        mine[mine == 2147483647] = np.inf
        assert np.array_equal(mine, D)
        order = np.random.default_rng(1).permutation(g.n).astype(np.uint32)
        L, _ = orc.gll(og, order, threads=2)
        pairs = np.stack(np.meshgrid(np.arange(g.n), np.arange(g.n), indexing="ij"), -1).reshape(-1, 2)
        q = L.query(pairs).astype(np.float64)
        q[q == 2147483647] = np.inf
        assert np.array_equal(q, D.reshape(-1))


def test_planted_mistakes_are_caught(chl, orc):
    g = chl.Graph.grid(8, 8, 1, 9, seed=5)
    og = oracle_graph(orc, g, None)
    order = g.order_betweenness(8, seed=1)
    good = labels_of(orc.brute_chl(og, order))
    # a wrong distance, a dropped label, an extra non-canonical label
    (o0, h0, d0), s1 = good
    bad_d = d0.copy()
    bad_d[len(bad_d) // 2] += 1
    assert not same([(o0, h0, bad_d), s1], good)
    pairs = np.stack(np.meshgrid(np.arange(g.n), np.arange(g.n), indexing="ij"), -1).reshape(-1, 2)
    D = og.dijkstra_multi(np.arange(g.n, dtype=np.uint32)).reshape(-1)
    # dropping the hub that covers some pair breaks the 2-hop cover
    v = 10
    keep = np.ones(len(h0), bool)
    keep[o0[v]] = False  # drop v's most important hub from side 0
    o0b = o0.copy()
    o0b[v + 1:] -= 1
    Lb = orc.OLabels.from_csr(g.n, (o0b, h0[keep], d0[keep]), s1)
    assert not np.array_equal(Lb.query(pairs), D)


def parse_label_file(path):
    """(off, hub, dist) per side of a label file (labeling.hpp:167-183 layout)."""
    lines = open(path).read().split("\n")
    n = int(lines[0])
    out = []
    for side in (0, 1):
        off, hub, dist = [0], [], []
        for v in range(n):
            t = list(map(int, lines[1 + 2 * v + side].split()))
            hub += t[1::2][: t[0]]
            dist += t[2::2][: t[0]]
            off.append(off[-1] + t[0])
        out.append((np.array(off), np.array(hub, np.uint32), np.array(dist, np.int32)))
    return out


@pytest.fixture(scope="module")
def label_digest(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("digest") / "label_digest")
    subprocess.run(["g++", "-O2", "-std=c++17", "-o", exe, os.path.join(GOLD, "label_digest.cpp")], check=True)
    return exe


@pytest.mark.parametrize("name", SMALL)
def test_label_digest_tool_matches_host_digest(label_digest, name, tmp_path):
    """tests/golden/label_digest (which turns the reference's full-size label files
    into fixtures) against the numpy restatement of chl_labels_checksum, and a
    planted one-unit change moves its digest."""
    from test_gpu_fullsize import host_digest
    path = os.path.join(GOLD, f"{name}.labels")
    got = json.loads(subprocess.run([label_digest, path], capture_output=True, text=True, check=True).stdout)
    sides = parse_label_file(path)
    assert got["count"] == [int(s[0][-1]) for s in sides]
    assert [int(x, 16) for x in got["digest"]] == [host_digest(*s) for s in sides]
    text = open(path).read().split("\n")
    t = text[1].split()
    t[2] = str(int(t[2]) + 1)  # distance of the first entry of vertex 0, side 0
    text[1] = " ".join(t)
    planted = tmp_path / "planted.labels"
    planted.write_text("\n".join(text))
    again = json.loads(subprocess.run([label_digest, str(planted)], capture_output=True, text=True,
                                      check=True).stdout)
    assert again["digest"][0] != got["digest"][0] and again["digest"][1] == got["digest"][1]


@pytest.mark.parametrize("name", ["c3u", "c3w", "c2"])
def test_fullsize_fixture_inputs_are_the_generators(chl, name, tmp_path):
    """The inputs behind tests/golden/fullsize.json are what this project's
    generators give for the stated arguments (graph and order file sha256)."""
    f = json.load(open(os.path.join(GOLD, "fullsize.json")))[name]
    if name == "c2":
        g = chl.Graph.knn(1_000_000, 3, seed=1)
        order = g.order_betweenness(256, seed=7)
    else:
        g = chl.Graph.rmat(20, 3 << 20, wmax=1 if name == "c3u" else 16, seed=1)
        order = g.order_degree(seed=2)
    gp, op = tmp_path / "g.gr", tmp_path / "g.order"
    g.write(str(gp))
    chl.order_write(str(op), order)
    assert hashlib.sha256(gp.read_bytes()).hexdigest() == f["graph_sha256"]
    assert hashlib.sha256(op.read_bytes()).hexdigest() == f["order_sha256"]

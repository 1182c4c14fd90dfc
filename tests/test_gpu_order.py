"""Vertex orders computed on the GPU.

Degree order: equal to the reference degree tool's output on the golden graphs
(its order files and stdout, and plant_amd/bin/degree byte for byte), to a numpy
restatement of distributed_memory/degree.cpp:62-67
(degree = out-arcs + in-arcs descending, ties by higher id first) and, with seeded
ties, of (degree descending, SplitMix64 key ascending, id ascending); and equal to
the host order_degree.

Sampled betweenness: the GPU order equals the host order (gen.cpp
order_betweenness, pinned in test_host.py against scipy's Dijkstra and a closed
form): the same scores and the same tie-broken order."""
import json
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
DEGREE_TOOL = os.path.join(os.path.dirname(GOLD), "..", "plant_amd", "bin", "degree")
with open(os.path.join(GOLD, "manifest.json")) as _f:
    MANIFEST = json.load(_f)

M64 = (1 << 64) - 1


def splitmix_keys(n, seed):
    """SplitMix64 (Steele, Lea and Flood 2014) of seed ^ (v * 0xD1B54A32D192ED03), in
    Python integers (no wrap-around surprises)."""
    out = np.empty(n, np.uint64)
    for v in range(n):
        z = ((seed ^ ((v * 0xD1B54A32D192ED03) & M64)) + 0x9E3779B97F4A7C15) & M64
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
        out[v] = z ^ (z >> 31)
    return out


def degree_restated(g, seed, reference_ties):
    deg = np.diff(g.csr(1)[0]).astype(np.int64) + np.diff(g.csr(0)[0]).astype(np.int64)
    ids = np.arange(g.n, dtype=np.int64)
    if reference_ties:
        return np.lexsort((-ids, -deg)).astype(np.uint32)
    return np.lexsort((ids, splitmix_keys(g.n, seed), -deg)).astype(np.uint32)


def degree_graphs(chl):
    yield "grid", chl.Graph.grid(100, 100, 1, 100, seed=1)
    yield "rmat", chl.Graph.rmat(13, 3 * 8192, seed=3)
    rng = np.random.default_rng(8)
    n, m = 700, 2500
    d = chl.Graph.from_arcs(n, rng.integers(0, n, m).astype(np.uint32), rng.integers(0, n, m).astype(np.uint32),
                            rng.integers(1, 6, m).astype(np.int32))
    assert not d.symmetric
    yield "directed", d
    yield "no arcs", chl.Graph.from_arcs(5, np.zeros(0, np.uint32), np.zeros(0, np.uint32), np.zeros(0, np.int32))
    yield "one vertex", chl.Graph.from_arcs(1, np.zeros(0, np.uint32), np.zeros(0, np.uint32), np.zeros(0, np.int32))


@pytest.mark.parametrize("reference_ties", [True, False])
def test_degree_order(chl, reference_ties):
    for name, g in degree_graphs(chl):
        for seed in (1, 7):
            got, ms = g.order_degree_device(seed=seed, reference_ties=reference_ties)
            want = degree_restated(g, seed, reference_ties)
            assert np.array_equal(got, want), (name, seed)
            assert np.array_equal(got, g.order_degree(seed=seed, reference_ties=reference_ties)), (name, seed)
            assert ms >= 0
            if reference_ties:
                break  # the seed does not enter
    # the reference rule on a tiny case by hand: degrees 1, 2, 1, 0 (arcs 0->1, 1->2)
    g = chl.Graph.from_arcs(4, np.array([0, 1], np.uint32), np.array([1, 2], np.uint32), np.array([1, 1], np.int32))
    got, _ = g.order_degree_device(reference_ties=True)
    assert list(got) == [1, 2, 0, 3]


@pytest.mark.parametrize("name", sorted(k for k, v in MANIFEST.items() if "degree_stdout" in v))
def test_degree_tool_matches_reference_tool(chl, name, tmp_path):
    """plant_amd/bin/degree (device order) prints the reference tool's line and
    writes its order file byte for byte (fixtures from distributed_memory/degree.cpp)."""
    gr = os.path.join(GOLD, name + ".gr")
    fixture = os.path.join(GOLD, name + ".degree")
    got, _ = chl.Graph.load(gr).order_degree_device(reference_ties=True)
    assert np.array_equal(got, chl.order_load(fixture))
    out = tmp_path / "o.txt"
    print("degree -o", out, gr, flush=True)
    r = subprocess.run([DEGREE_TOOL, "-o", str(out), gr], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert r.stdout == MANIFEST[name]["degree_stdout"]
    assert out.read_bytes() == open(fixture, "rb").read()


def test_degree_order_1M(chl):
    g = chl.Graph.knn(1_000_000, 3, seed=1)
    for reference_ties in (True, False):
        got, ms = g.order_degree_device(seed=5, reference_ties=reference_ties)
        assert np.array_equal(got, g.order_degree(seed=5, reference_ties=reference_ties))
        print(f"device degree order 1M (reference ties {reference_ties}): {ms:.2f} ms")


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

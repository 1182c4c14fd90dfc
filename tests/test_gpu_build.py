"""GPU label construction and queries against the CPU oracle (bit-exact).

The oracle's run_paraPLL restatement (oracle/chl_oracle.cpp) is itself pinned to
the reference's own hhl build and to the CHL definition in test_oracle.py.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def assert_same_labels(gpu, ref, sides=(0, 1)):
    for side in sides:
        a = gpu.copy(side)
        b = ref.copy(side)
        for name, x, y in zip(("off", "hub", "dist"), a, b):
            if not np.array_equal(x, y):
                bad = np.flatnonzero(np.asarray(x[: len(y)]) != np.asarray(y[: len(x)]))
                raise AssertionError(f"side {side} {name}: sizes {len(x)} vs {len(y)}, first diff at {bad[:5]}")


def oracle_of(chl, orc, g, order, raw=None):
    if raw is None:
        src, dst, w = g.arc_list(1)
        og = orc.OGraph.from_arcs(g.n, src, dst, w)
    else:
        og = orc.OGraph.from_arcs(g.n, *raw)
    ref, _ = orc.gll(og, order, threads=8)
    return og, ref


@pytest.fixture(scope="module")
def grid(chl, orc):
    g = chl.Graph.grid(30, 30, 1, 100, seed=11)
    order = g.order_betweenness(32, seed=3)
    og, ref = oracle_of(chl, orc, g, order)
    return g, order, og, ref


@pytest.mark.parametrize("mode,prune", [(1, 0), (1, 1), (0, 0), (2, 0)])
def test_grid_modes(chl, grid, mode, prune):
    g, order, _, ref = grid
    L = chl.build(g, order, mode=mode, plant_prune=prune, batch_roots=128)
    assert_same_labels(L, ref)


@pytest.mark.parametrize("batch,delta,slots", [(1, 0, 0), (7, 1, 3), (900, 1000, 64), (64, 37, 1)])
def test_grid_schedule_independent(chl, grid, batch, delta, slots):
    g, order, _, ref = grid
    for mode, prune in ((1, 0), (0, 0), (1, 1)):
        L = chl.build(g, order, mode=mode, plant_prune=prune, batch_roots=batch, delta=delta, slots=slots)
        assert_same_labels(L, ref)


@pytest.mark.parametrize("granularity", [1, 2, 3])
def test_granularity_and_autotune(chl, orc, granularity):
    g = chl.Graph.knn(6000, 3, seed=12)
    order = g.order_betweenness(32, seed=4)
    _, ref = oracle_of(chl, orc, g, order)
    for mode in (1, 0):
        L = chl.build(g, order, mode=mode, granularity=granularity, batch_roots=700)
        assert_same_labels(L, ref)
    L = chl.build(g, order, autotune=1, batching=chl.BATCH_LEARNED)
    assert_same_labels(L, ref)
    assert L.stats.autotune_ms > 0 and L.stats.granularity_used in (1, 2, 3)


def test_output_spill_relaunch(chl, grid):
    # a batch output buffer of 100 labels: nearly every root spills and is run again
    g, order, _, ref = grid
    for mode in (1, 0):
        L = chl.build(g, order, mode=mode, batch_roots=300, out_capacity=100)
        assert_same_labels(L, ref)


def test_hybrid_switches(chl, grid):
    g, order, _, ref = grid
    L = chl.build(g, order, mode=chl.MODE_HYBRID, switch_ratio=1.5, batch_roots=50)
    assert L.stats.switch_rank > 0
    assert L.stats.trees_pruned > 0
    assert_same_labels(L, ref)


def test_hybrid_switch_ratio_is_a_float(chl, orc):
    # hl::run takes ps as a float (distributed_memory/akiba.hpp:190, widened at
    # :206), so 1 - 1e-12 is 1.0 there. On a 12-cycle the top root's tree/label
    # ratio is exactly 1 and the next root's is above 1: with one root per batch
    # (one thread, one root per superstep in the oracle) ps = 0.5 switches after
    # the first root, 1.0 and 1 - 1e-12 after the second.
    n = 12
    src = np.arange(n, dtype=np.uint32)
    dst = (src + 1) % n
    g = chl.Graph.from_arcs(n, np.r_[src, dst], np.r_[dst, src], np.ones(2 * n, np.int32))
    order = np.random.default_rng(3).permutation(n).astype(np.uint32)
    og, ref = oracle_of(chl, orc, g, order)
    for ps in (0.5, 1.0, 1 - 1e-12):
        L = chl.build(g, order, mode=chl.MODE_HYBRID, switch_ratio=ps, batch_roots=1)
        assert_same_labels(L, ref)
        want = int(orc.hybrid(og, order, threads=1, ps=ps)[1][1])
        assert int(L.stats.switch_rank) == want == (1 if ps == 0.5 else 2), ps


@pytest.mark.parametrize("kind", ["rmat", "knn"])
def test_learned_hybrid_switch(chl, orc, kind):
    # the learned policy places the switch from predicted tree/label ratios and a
    # timed trial batch of pruned trees; every choice gives the same labels
    if kind == "rmat":
        g = chl.Graph.rmat(14, 3 * 16384, seed=6)
        order = g.order_degree(seed=1)
    else:
        g = chl.Graph.knn(16000, 3, seed=6)
        order = g.order_betweenness(32, seed=1)
    _, ref = oracle_of(chl, orc, g, order)
    L = chl.build(g, order, mode=chl.MODE_HYBRID, batching=chl.BATCH_LEARNED, verbose=1)
    assert_same_labels(L, ref)
    st = L.stats
    assert st.trees_plant > 0
    # pruned trees run only in trials and after a trial made the switch final
    assert (st.trees_pruned > 0) == (st.switch_trials > 0)
    if st.switch_rank >= 0:
        assert st.switch_trials >= 1


def test_query_all_pairs_grid(chl, grid):
    g, order, og, ref = grid
    L = chl.build(g, order, batch_roots=256)
    n = g.n
    src = np.arange(0, n, 7, dtype=np.uint32)
    D = og.dijkstra_multi(src)
    pairs = np.stack(np.meshgrid(src, np.arange(n, dtype=np.uint32), indexing="ij"), -1).reshape(-1, 2)
    got = L.query(pairs)
    assert np.array_equal(got, D.reshape(-1))
    assert np.array_equal(got, ref.query(pairs))


def test_knn_road_like(chl, orc):
    g = chl.Graph.knn(4000, 3, seed=5)
    order = g.order_betweenness(64, seed=9)
    _, ref = oracle_of(chl, orc, g, order)
    for mode in (1, 0):
        L = chl.build(g, order, mode=mode, batch_roots=512)
        assert_same_labels(L, ref)


def test_hashed_trees_under_slot_memory_cap(chl, orc):
    """Search slots capped at a few dozen of 2048 (memory-limited, as on a 16M-vertex
    graph): learned batching runs the roots predicted to grow small trees on the
    small-table kernel (a wavefront per root), those predicted medium on the
    medium-table kernel (256 threads per root) and the rest on the tree kernel.
    The labels are the oracle's with and without the medium tables, and both
    hashed kernels completed roots."""
    g = chl.Graph.knn(100_000, 3, seed=21)
    order, score = g.order_betweenness(64, seed=5, with_score=True)
    _, ref = oracle_of(chl, orc, g, order)
    cap_mb = max(1, int(32 * 36 * g.n / 2**20))
    for medium in (1, 0):
        L = chl.build(g, order, batching=chl.BATCH_LEARNED, vertex_score=score.ctypes.data, slot_memory_mb=cap_mb,
                      medium_trees=medium)
        s = L.stats
        print(f"medium_trees={medium}: slots {s.slots_used}, small roots {s.small_roots}, "
              f"medium roots {s.medium_roots}, batches {s.batches}", flush=True)
        assert s.slots_used < 64  # memory-limited: far below the 2048 wanted
        assert_same_labels(L, ref)
        assert s.small_roots > 0
        assert (s.medium_roots > 0) == (medium == 1)


def test_hashed_trees_directed_under_slot_memory_cap(chl, orc):
    """The same on a directed graph: both sides run their own hashed and
    tree-kernel passes (forward trees over out-arcs, backward over in-arcs)."""
    rng = np.random.default_rng(17)
    n, m = 50_000, 250_000
    src = rng.integers(0, n, m).astype(np.uint32)
    dst = ((src.astype(np.int64) + rng.integers(1, 200, m)) % n).astype(np.uint32)  # local arcs
    w = rng.integers(1, 30, m).astype(np.int32)
    g = chl.Graph.from_arcs(n, src, dst, w)
    assert not g.symmetric
    order, score = g.order_betweenness(64, seed=5, with_score=True)
    _, ref = oracle_of(chl, orc, g, order, raw=(src, dst, w))
    cap_mb = max(1, int(32 * 36 * n / 2**20))
    L = chl.build(g, order, batching=chl.BATCH_LEARNED, vertex_score=score.ctypes.data, slot_memory_mb=cap_mb,
                  medium_trees=1)
    s = L.stats
    print(f"directed: slots {s.slots_used}, small roots {s.small_roots}, medium roots {s.medium_roots}, "
          f"batches {s.batches}", flush=True)
    assert_same_labels(L, ref)
    assert s.small_roots + s.medium_roots > 0  # (these trees outgrow the small tables)


def test_directed_random(chl, orc):
    rng = np.random.default_rng(7)
    n = 500
    m = 3000
    src = rng.integers(0, n, m).astype(np.uint32)
    dst = rng.integers(0, n, m).astype(np.uint32)
    w = rng.integers(1, 50, m).astype(np.int32)
    g = chl.Graph.from_arcs(n, src, dst, w)
    assert not g.symmetric
    order = rng.permutation(n).astype(np.uint32)
    _, ref = oracle_of(chl, orc, g, order, raw=(src, dst, w))
    for mode, prune in ((1, 0), (1, 1), (0, 0), (2, 0)):
        L = chl.build(g, order, mode=mode, plant_prune=prune, batch_roots=37)
        assert_same_labels(L, ref)
    og = orc.OGraph.from_arcs(n, src, dst, w)
    pairs = rng.integers(0, n, (5000, 2)).astype(np.uint32)
    D = og.dijkstra_multi(np.arange(n, dtype=np.uint32))
    assert np.array_equal(L.query(pairs), D[pairs[:, 0], pairs[:, 1]])


def test_rmat_unit_weights_degree_order(chl, orc):
    g = chl.Graph.rmat(11, 3 * 2048, seed=4)
    order = g.order_degree(seed=2)
    _, ref = oracle_of(chl, orc, g, order)
    for mode in (1, 0, 2):
        L = chl.build(g, order, mode=mode, batch_roots=300)
        assert_same_labels(L, ref)


def test_disconnected_and_tiny(chl, orc):
    # two components, an isolated vertex and a self-loop
    src = np.array([0, 1, 2, 4, 5, 5], np.uint32)
    dst = np.array([1, 2, 0, 5, 6, 5], np.uint32)
    w = np.array([3, 4, 5, 1, 2, 9], np.int32)
    g = chl.Graph.from_arcs(8, src, dst, w, undirected=True)
    order = np.array([5, 2, 7, 0, 3, 1, 6, 4], np.uint32)
    _, ref = oracle_of(chl, orc, g, order, raw=(np.r_[src, dst], np.r_[dst, src], np.r_[w, w]))
    L = chl.build(g, order)
    assert_same_labels(L, ref)
    q = L.query(np.array([[0, 4], [3, 3], [4, 6], [6, 4], [0, 2]], np.uint32))
    assert list(q) == [chl.INFTY, 0, 3, 3, 5]
    g1 = chl.Graph.from_arcs(1, np.zeros(0, np.uint32), np.zeros(0, np.uint32), np.zeros(0, np.int32))
    L1 = chl.build(g1, np.array([0], np.uint32))
    off, hub, dist = L1.copy(0)
    assert list(off) == [0, 1] and list(hub) == [0] and list(dist) == [0]


def test_query_rejects_out_of_range(chl, grid):
    g, order, _, _ = grid
    L = chl.build(g, order)
    with pytest.raises(chl.ChlError):
        L.query(np.array([[0, g.n]], np.uint32))


def test_query_device_entry_point(chl, grid):
    """chl_query_device on device-resident pairs, for pair counts from one up to
    several pairs per wavefront of the kernel's grid: ids >= n give -1, the rest
    equal the oracle's queries."""
    import ctypes as C

    hip = C.CDLL("libamdhip64.so")  # device buffers through the HIP runtime the library uses
    vp = C.c_void_p
    g, order, _, ref = grid
    L = chl.build(g, order)
    rng = np.random.default_rng(8)
    for q in (1, 2, 63, 64, 65, 1000, 32767, 32768, 32769, 100003, 300000):
        pairs = rng.integers(0, g.n, (q, 2)).astype(np.uint32)
        bad = rng.random((q, 2)) < 0.03
        pairs[bad] = g.n + rng.integers(0, 3, int(bad.sum())).astype(np.uint32)
        pairs[-1, 1] = 0xFFFFFFFF if q > 2 else pairs[-1, 1]
        ok = (pairs < g.n).all(1)
        want = np.full(q, -1, np.int32)
        if ok.any():
            want[ok] = ref.query(pairs[ok])
        got = np.full(q, -7, np.int32)
        dp, do = vp(), vp()
        assert hip.hipMalloc(C.byref(dp), C.c_size_t(pairs.nbytes)) == 0
        assert hip.hipMalloc(C.byref(do), C.c_size_t(got.nbytes)) == 0
        try:
            assert hip.hipMemcpy(dp, pairs.ctypes.data_as(vp), C.c_size_t(pairs.nbytes), 1) == 0  # to device
            assert hip.hipMemcpy(do, got.ctypes.data_as(vp), C.c_size_t(got.nbytes), 1) == 0
            print("chl_query_device q =", q, flush=True)
            assert chl.lib().chl_query_device(L._h, dp, q, do, None) == 0
            assert hip.hipDeviceSynchronize() == 0
            assert hip.hipMemcpy(got.ctypes.data_as(vp), do, C.c_size_t(got.nbytes), 2) == 0  # to host
        finally:
            hip.hipFree(dp)
            hip.hipFree(do)
        assert np.array_equal(got, want), q


def test_label_file_roundtrip(chl, grid, tmp_path):
    g, order, _, ref = grid
    L = chl.build(g, order)
    p = tmp_path / "l.txt"
    L.write(str(p))
    ref.write(str(tmp_path / "r.txt"))
    assert p.read_bytes() == (tmp_path / "r.txt").read_bytes()
    L2 = chl.Labels.read(str(p))
    assert_same_labels(L2, ref)
    pairs = np.random.default_rng(1).integers(0, g.n, (3000, 2)).astype(np.uint32)
    assert np.array_equal(L2.query(pairs), ref.query(pairs))


def test_query_long_rows(chl, orc):
    # a path ranked from one end: vertex v's row holds every hub h <= v, so rows
    # reach 2500 entries and the query merges over several staged tiles
    n = 2500
    rng = np.random.default_rng(21)
    w = rng.integers(1, 10, n - 1).astype(np.int32)
    src = np.arange(n - 1, dtype=np.uint32)
    g = chl.Graph.from_arcs(n, src, src + 1, w, undirected=True)
    order = np.arange(n, dtype=np.uint32)
    L = chl.build(g, order, batch_roots=512)
    off, hub, dist = L.copy(0)
    assert int(np.diff(off).max()) == n
    pos = np.concatenate([[0], np.cumsum(w, dtype=np.int64)])
    pairs = rng.integers(0, n, (20000, 2)).astype(np.uint32)
    pairs[:50] = [[n - 1, n - 1 - k] for k in range(50)]  # both rows long
    want = np.abs(pos[pairs[:, 0]] - pos[pairs[:, 1]]).astype(np.int32)
    assert np.array_equal(L.query(pairs), want)

"""The segmented label store: a batch's labels become one segment (5-byte entries,
or 8-byte ones where a distance exceeds 65535), small segments are merged as the
build goes, and the finished store is merged into one segment unless
keep_segments asks otherwise. Whatever the segments and formats, the labels,
digests, files and query answers must be the oracle's (bit-exact)."""
import numpy as np
import pytest

from test_gpu_build import assert_same_labels, oracle_of

pytestmark = pytest.mark.gpu


def check_segments(L, stats=None):
    for side in (0, 1):
        segs = L.segments(side)
        assert segs, side
        assert sum(s[3] for s in segs) == L.total(side)
        for (a0, a1, _, _), (b0, _, _, _) in zip(segs, segs[1:]):
            assert a0 < a1 <= b0  # ascending, disjoint hub ranges
        assert all(f in (5, 8) for _, _, f, _ in segs)
    if stats is not None:
        assert stats.segments == len(L.segments(0))
    return L.segments(0)


def all_pairs_check(orc, og, L, n):
    D = og.dijkstra_multi(np.arange(n, dtype=np.uint32), threads=8)
    u, v = np.meshgrid(np.arange(n, dtype=np.uint32), np.arange(n, dtype=np.uint32), indexing="ij")
    pairs = np.stack([u.reshape(-1), v.reshape(-1)], 1)
    assert np.array_equal(L.query(pairs), D.reshape(-1))


@pytest.fixture(scope="module")
def heavy_grid(chl, orc):
    # lengths up to 8000: whole-graph trees reach distances far above 65535 (8-byte
    # entries), while the short trees of late ranks stay below it (5-byte entries)
    g = chl.Graph.grid(24, 24, 1, 8000, seed=5)
    order = g.order_betweenness(32, seed=2)
    og, ref = oracle_of(chl, orc, g, order)
    return g, order, og, ref


@pytest.mark.parametrize("mode", [1, 0, 2])
def test_long_distances_take_the_8_byte_form(chl, orc, heavy_grid, mode):
    g, order, og, ref = heavy_grid
    kept = chl.build(g, order, mode=mode, batch_roots=16, keep_segments=1)
    assert_same_labels(kept, ref)
    segs = check_segments(kept, kept.stats)
    fmts = {f for _, _, f, _ in segs}
    assert 8 in fmts and 5 in fmts, segs
    assert 1 < len(segs) < kept.stats.batches  # appended per batch, small ones merged
    merged = chl.build(g, order, mode=mode, batch_roots=16)
    assert_same_labels(merged, ref)
    assert [(s[0], s[2]) for s in check_segments(merged, merged.stats)] == [(0, 8)]
    for side in (0, 1):
        assert kept.checksum(side) == merged.checksum(side)
    all_pairs_check(orc, og, kept, g.n)


def test_segmented_store_files_lengths_digests(chl, orc, tmp_path):
    g = chl.Graph.knn(3000, 3, seed=4)
    order = g.order_betweenness(32, seed=9)
    _, ref = oracle_of(chl, orc, g, order)
    kept = chl.build(g, order, batch_roots=64, keep_segments=1)
    merged = chl.build(g, order, batch_roots=64)
    assert len(check_segments(kept, kept.stats)) > 1
    assert len(check_segments(merged, merged.stats)) == 1
    for L in (kept, merged):
        assert_same_labels(L, ref)
        for side in (0, 1):
            off = ref.copy(side)[0]
            assert np.array_equal(L.row_lengths(side), np.diff(np.asarray(off, np.int64)))
    pa, pb = tmp_path / "a.txt", tmp_path / "b.txt"
    kept.write(str(pa))
    merged.write(str(pb))
    assert pa.read_bytes() == pb.read_bytes()
    rng = np.random.default_rng(1)
    pairs = rng.integers(0, g.n, (50000, 2)).astype(np.uint32)
    assert np.array_equal(kept.query(pairs), ref.query(pairs))
    assert np.array_equal(merged.query(pairs), ref.query(pairs))


def test_directed_long_distances_gll(chl, orc):
    # two different sides, GLL cleaning across F8 and F5 batch segments
    rng = np.random.default_rng(21)
    n, m = 400, 2400
    src = rng.integers(0, n, m).astype(np.uint32)
    dst = rng.integers(0, n, m).astype(np.uint32)
    w = rng.integers(1, 40000, m).astype(np.int32)
    g = chl.Graph.from_arcs(n, src, dst, w)
    order = rng.permutation(n).astype(np.uint32)
    og, ref = oracle_of(chl, orc, g, order)
    for mode in (0, 2, 1):
        L = chl.build(g, order, mode=mode, batch_roots=8, keep_segments=1)
        assert_same_labels(L, ref)
        check_segments(L, L.stats)
    all_pairs_check(orc, og, L, n)

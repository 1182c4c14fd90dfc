"""Label files at scale: the parallel text writer and reader and the binary
sidecar give back the same labels (device digest, host copies, queries) on a
symmetric and on a directed graph; a token-layout text file takes the
sequential path and reads the same."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def same_labels(a, b):
    for side in (0, 1):
        assert a.checksum(side) == b.checksum(side)
        for x, y in zip(a.copy(side), b.copy(side)):
            assert np.array_equal(x, y)


def graphs(chl):
    g = chl.Graph.knn(20000, 3, seed=3)
    yield g, g.order_betweenness(32, seed=2)
    rng = np.random.default_rng(4)
    n, m = 3000, 12000
    src = rng.integers(0, n, m).astype(np.uint32)
    dst = rng.integers(0, n, m).astype(np.uint32)
    d = chl.Graph.from_arcs(n, src, dst, rng.integers(1, 30, m).astype(np.int32))
    assert not d.symmetric
    yield d, rng.permutation(n).astype(np.uint32)


def test_text_and_binary_roundtrips(chl, orc, tmp_path):
    for i, (g, order) in enumerate(graphs(chl)):
        L = chl.build(g, order)
        txt = tmp_path / f"l{i}.txt"
        L.write(str(txt))
        # the writer's bytes are the reference writer's (the oracle writes the same text)
        src, dst, w = g.arc_list(1)
        ref, _ = orc.gll(orc.OGraph.from_arcs(g.n, src, dst, w), order, threads=8)
        ref.write(str(tmp_path / f"r{i}.txt"))
        assert txt.read_bytes() == (tmp_path / f"r{i}.txt").read_bytes()
        same_labels(L, chl.Labels.read(str(txt)))
        # tokens in another layout: the sequential path
        tok = tmp_path / f"t{i}.txt"
        tok.write_bytes(txt.read_bytes().replace(b"\n", b" \t"))
        same_labels(L, chl.Labels.read(str(tok)))
        # binary sidecar
        b = tmp_path / f"l{i}.bin"
        L.save(str(b))
        L2 = chl.Labels.load(str(b))
        same_labels(L, L2)
        pairs = np.random.default_rng(i).integers(0, g.n, (20000, 2)).astype(np.uint32)
        assert np.array_equal(L.query(pairs), L2.query(pairs))

"""BF16 MFMA predictor kernels against the numpy reference (tolerances stated
per test), online training, and learned batching end to end."""
import numpy as np
import pytest

from oracle import mlp_ref as M

pytestmark = pytest.mark.gpu


def test_mfma_operand_layout(chl):
    rng = np.random.default_rng(0)
    a = M.bf16_bits(rng.normal(size=(32, 8)).astype(np.float32))
    b = M.bf16_bits(rng.normal(size=(8, 32)).astype(np.float32))
    d = chl.mfma_probe(a, b)
    af = (a.astype(np.uint32) << 16).view(np.float32).astype(np.float64)
    bf = (b.astype(np.uint32) << 16).view(np.float32).astype(np.float64)
    # products of BF16 values are exact in FP32; 8-term FP32 sums: 1e-6 relative
    np.testing.assert_allclose(d, af @ bf, rtol=1e-6, atol=1e-6)


def features(rng, n):
    x = rng.normal(size=(n, M.F)).astype(np.float32)
    x[:, 4:] = 0.0
    return x


def test_forward_matches_reference(chl):
    rng = np.random.default_rng(1)
    p = M.init_params(5)
    x = features(rng, 1000)  # not a multiple of 32: exercises the tail tile
    y = chl.mlp_forward(p, x)
    ref, _ = M.forward(p, x, bf16=True)
    # same BF16 rounding points; FP32 vs float64 accumulation (and a rare BF16
    # rounding flip of a hidden activation): 2e-3 absolute on outputs of size ~1
    np.testing.assert_allclose(y, ref, rtol=2e-3, atol=2e-3)


def test_gradient_matches_reference(chl):
    rng = np.random.default_rng(2)
    p = M.init_params(6)
    x = features(rng, 517)
    t = rng.normal(size=(517, M.O)).astype(np.float32)
    g, loss = chl.mlp_grad(p, x, t)
    ref = M.grad(p, x, t, bf16=True)
    assert abs(loss - M.loss(p, x, t, bf16=True)) <= 1e-3 * max(1.0, abs(loss))
    # per block, relative to the block's norm: 1e-2 (BF16 operands, FP32 atomics)
    for lo, hi in ((M.OFF_W1, M.OFF_B1), (M.OFF_B1, M.OFF_W2), (M.OFF_W2, M.OFF_B2), (M.OFF_B2, M.OFF_W3),
                   (M.OFF_W3, M.OFF_B3), (M.OFF_B3, M.NPARAM)):
        err = np.linalg.norm(g[lo:hi] - ref[lo:hi])
        assert err <= 1e-2 * np.linalg.norm(ref[lo:hi]) + 1e-6, (lo, err, np.linalg.norm(ref[lo:hi]))


def test_graph_replays_match_eager(chl):
    # the captured training step, replayed 12 times with eager steps, state
    # copies and synchronisations in between, each replay against one eager step
    # from the same state; FP32 atomics reorder the gradient sums, and a
    # near-zero gradient whose sign flips moves its parameter by 2 lr under Adam,
    # so 1e-3 of the gradient's and the parameters' norms
    rng = np.random.default_rng(3)
    x = features(rng, 700)
    t = rng.normal(size=(700, M.O)).astype(np.float32)
    worst = chl.mlp_graph_check(M.init_params(8), x, t, replays=12)
    assert worst < 1e-3, worst


def test_training_fits_tree_sizes(chl, orc):
    # targets: log tree size / label count of PLaNT trees measured by the oracle
    g = chl.Graph.knn(3000, 3, seed=4)
    order, score = g.order_betweenness(32, seed=1, with_score=True)
    src, dst, w = g.arc_list(1)
    og = orc.OGraph.from_arcs(g.n, src, dst, w)
    _, ts, lc = orc.plant_range(og, order, 0, g.n)
    off, head, _ = g.csr(1)
    deg = np.diff(off).astype(np.float64) * 2
    two = np.array([deg[head[off[v]:off[v + 1]]].sum() for v in range(g.n)])
    v = order.astype(np.int64)
    feats = np.stack([np.log1p(np.arange(g.n)) / np.log1p(g.n), np.log1p(deg[v]), np.log1p(two[v]),
                      np.log1p(score[v])], 1)
    feats[:, 1:] = (feats[:, 1:] - feats[:, 1:].mean(0)) / feats[:, 1:].std(0)
    x = np.zeros((g.n, M.F), np.float32)
    x[:, :4] = feats
    t = np.stack([np.log1p(ts[1::2]), np.log1p(lc[1::2])], 1).astype(np.float32)
    rng = np.random.default_rng(0)
    idx = rng.permutation(g.n)
    tr, te = idx[:2400], idx[2400:]
    p0 = M.init_params(7)
    p1, curve = chl.mlp_train(p0, x[tr], t[tr], steps=400, lr=3e-3)
    assert curve[-1] < 0.3 * curve[0]
    pred = chl.mlp_forward(p1, x[te])
    mae = np.abs(pred[:, 0] - t[te, 0]).mean()
    mae_mean = np.abs(t[tr, 0].mean() - t[te, 0]).mean()  # the no-network estimate
    assert mae < 0.6 * mae_mean, (mae, mae_mean)
    # the trained weights still agree with the reference forward pass
    np.testing.assert_allclose(pred, M.forward(p1, x[te])[0], rtol=5e-3, atol=5e-3)


def test_learned_batching_same_labels_better_prediction(chl, orc):
    g = chl.Graph.knn(20000, 3, seed=9)
    order, score = g.order_betweenness(64, seed=3, with_score=True)
    static = chl.build(g, order, batching=chl.BATCH_STATIC)
    learned = chl.build(g, order, batching=chl.BATCH_LEARNED, vertex_score=score.ctypes.data)
    for side in (0, 1):
        for a, b in zip(static.copy(side), learned.copy(side)):
            assert np.array_equal(a, b)
    st = learned.stats
    assert st.pred_roots > 0
    assert st.pred_mae_learned < st.pred_mae_static, (st.pred_mae_learned, st.pred_mae_static)
    assert st.mlp_ms > 0

"""Pins the numpy MLP reference (oracle/mlp_ref.py) before it checks the MFMA
kernels: its float64 gradient against central finite differences, its BF16
rounding against a closed form, and the layout of its parameter vector."""
import numpy as np

from oracle import mlp_ref as M


def test_param_layout_matches_header(chl):
    assert chl.mlp_param_count() == M.NPARAM == 8 * 64 + 64 + 64 * 64 + 64 + 64 * 2 + 2


def test_bf16_rounding_closed_form():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -7, 1.0 + 3 * 2 ** -8, -2.5, 0.0], np.float32)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: even mantissa (1.0) wins;
    # 1 + 3*2^-8 ties between 1+2^-7 and 1+2^-6... rounds to the even one (1 + 2^-6)
    assert list(M.bf16(x)) == [1.0, 1.0, 1.0 + 2 ** -7, 1.0 + 2 ** -6, -2.5, 0.0]


def test_reference_gradient_finite_differences():
    rng = np.random.default_rng(1)
    p = M.init_params(3).astype(np.float64)
    x = rng.normal(size=(37, M.F))
    x[:, 4:] = 0.0
    t = rng.normal(size=(37, M.O))
    g = M.grad(p, x, t, bf16=False)
    idx = np.r_[rng.choice(M.NPARAM, 60, replace=False), M.OFF_B3, M.OFF_B3 + 1, M.OFF_W1, M.OFF_B1]
    eps = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += eps
        pm[i] -= eps
        fd = (M.loss(pp, x, t, bf16=False) - M.loss(pm, x, t, bf16=False)) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])


def test_planted_gradient_mistake_is_caught():
    # a transposed W2 in the backward pass must fail the finite-difference check
    rng = np.random.default_rng(2)
    p = M.init_params(4).astype(np.float64)
    x = rng.normal(size=(20, M.F))
    t = rng.normal(size=(20, M.O))
    W1, b1, W2, b2, W3, b3 = M.unpack(p)
    y, (xb, h1, h2) = M.forward(p, x, bf16=False)
    dy = 2 * (y - t) / (20 * M.O)
    dh2 = (dy @ W3.T) * (h2 > 0)
    wrong = ((dh2 @ W2) * (h1 > 0)).sum(0)  # W2 instead of W2.T
    right = M.grad(p, x, t, bf16=False)[M.OFF_B1:M.OFF_W2]
    assert not np.allclose(wrong, right, rtol=1e-3, atol=1e-6)

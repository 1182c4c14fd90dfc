"""Reference of the tree-size predictor MLP (plant_amd/csrc/mlp.hip) in numpy.

TEST INFRASTRUCTURE ONLY (imported by tests/). The network is 8 -> 64 -> 64 -> 2
with ReLU; with bf16=True every operand the kernels feed to the matrix cores is
rounded to BF16 (round to nearest even) at the same points, accumulation is in
float64 here (FP32 on the device), and the backward pass treats the rounding as
the identity, as the kernels do. With bf16=False it is the plain float64
network, whose gradient the tests pin by central finite differences.
"""
import numpy as np

F, H, O = 8, 64, 2
OFF_W1 = 0
OFF_B1 = OFF_W1 + F * H
OFF_W2 = OFF_B1 + H
OFF_B2 = OFF_W2 + H * H
OFF_W3 = OFF_B2 + H
OFF_B3 = OFF_W3 + H * O
NPARAM = OFF_B3 + O


def bf16_bits(x):
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFFFFFF
    return (u >> 16).astype(np.uint16)


def bf16(x):
    return (bf16_bits(x).astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def unpack(p):
    p = np.asarray(p, np.float64)
    return (p[OFF_W1:OFF_B1].reshape(F, H), p[OFF_B1:OFF_W2], p[OFF_W2:OFF_B2].reshape(H, H), p[OFF_B2:OFF_W3],
            p[OFF_W3:OFF_B3].reshape(H, O), p[OFF_B3:NPARAM])


def _r(x, on):
    return bf16(x) if on else np.asarray(x, np.float64)


def forward(p, x, bf16=True):
    W1, b1, W2, b2, W3, b3 = unpack(p)
    xb = _r(x, bf16)
    h1 = np.maximum(xb @ _r(W1, bf16) + b1, 0.0)
    h1b = _r(h1, bf16)
    h2 = np.maximum(h1b @ _r(W2, bf16) + b2, 0.0)
    y = h2 @ _r(W3, bf16) + b3
    return y, (xb, h1b, h2)


def loss(p, x, t, bf16=True):
    y, _ = forward(p, x, bf16)
    return float(np.mean((y - np.asarray(t, np.float64)) ** 2))


def grad(p, x, t, bf16=True):
    W1, b1, W2, b2, W3, b3 = unpack(p)
    y, (xb, h1b, h2) = forward(p, x, bf16)
    n = y.shape[0]
    dy = 2.0 * (y - np.asarray(t, np.float64)) / (n * O)
    g = np.zeros(NPARAM)
    g[OFF_W3:OFF_B3] = (h2.T @ dy).ravel()
    g[OFF_B3:NPARAM] = dy.sum(0)
    dh2 = (dy @ _r(W3, bf16).T) * (h2 > 0)
    g[OFF_B2:OFF_W3] = dh2.sum(0)
    dh2b = _r(dh2, bf16)
    g[OFF_W2:OFF_B2] = (h1b.T @ dh2b).ravel()
    dh1 = (dh2b @ _r(W2, bf16).T) * (h1b > 0)
    g[OFF_B1:OFF_W2] = dh1.sum(0)
    g[OFF_W1:OFF_B1] = (xb.T @ dh1).ravel()
    return g


def init_params(seed=0):
    rng = np.random.default_rng(seed)
    p = np.zeros(NPARAM, np.float32)

    def fill(off, fin, fout):
        a = np.sqrt(6.0 / (fin + fout))
        p[off:off + fin * fout] = rng.uniform(-a, a, fin * fout)

    fill(OFF_W1, 4, H)
    p[OFF_W1 + 4 * H:OFF_B1] = 0.0
    fill(OFF_W2, H, H)
    fill(OFF_W3, H, O)
    p[OFF_B1:OFF_W2] = rng.uniform(0.0, 0.1, H)
    p[OFF_B2:OFF_W3] = rng.uniform(0.0, 0.1, H)
    return p

"""CPU restatement (float64 numpy) of the result-size estimator's MLP.

TEST INFRASTRUCTURE ONLY: the checker for the HIP forward/backward kernels
(gpu_self_join_amd/csrc/estimator.hip).  Network: 16 inputs -> 32 ReLU ->
32 ReLU -> 1, parameters packed as W1[32][16], b1[32], W2[32][32], b2[32],
w3[32], b3; loss = mean squared error.  The analytic gradient below is pinned
against central finite differences in tests/test_estimator.py.
"""
import numpy as np

F, H = 16, 32
OFF_W1 = 0
OFF_B1 = OFF_W1 + H * F
OFF_W2 = OFF_B1 + H
OFF_B2 = OFF_W2 + H * H
OFF_W3 = OFF_B2 + H
OFF_B3 = OFF_W3 + H
P = OFF_B3 + 1


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    return (p[OFF_W1:OFF_B1].reshape(H, F), p[OFF_B1:OFF_W2], p[OFF_W2:OFF_B2].reshape(H, H),
            p[OFF_B2:OFF_W3], p[OFF_W3:OFF_B3], p[OFF_B3])


def forward(p, x):
    W1, b1, W2, b2, w3, b3 = unpack(p)
    x = np.asarray(x, dtype=np.float64)
    a1 = x @ W1.T + b1
    h1 = np.maximum(a1, 0.0)
    a2 = h1 @ W2.T + b2
    h2 = np.maximum(a2, 0.0)
    return h2 @ w3 + b3, (x, a1, h1, a2, h2)


def loss_and_grad(p, x, t):
    W1, b1, W2, b2, w3, b3 = unpack(p)
    y, (x, a1, h1, a2, h2) = forward(p, x)
    t = np.asarray(t, dtype=np.float64)
    n = len(t)
    e = y - t
    loss = np.mean(e * e)
    dy = 2.0 * e / n
    g = np.zeros(P)
    g[OFF_W3:OFF_B3] = h2.T @ dy
    g[OFF_B3] = dy.sum()
    d2 = np.outer(dy, w3) * (a2 > 0)
    g[OFF_W2:OFF_B2] = (d2.T @ h1).ravel()
    g[OFF_B2:OFF_W3] = d2.sum(0)
    d1 = (d2 @ W2) * (a1 > 0)
    g[OFF_W1:OFF_B1] = (d1.T @ x).ravel()
    g[OFF_B1:OFF_W2] = d1.sum(0)
    return loss, g


def init(seed=0):
    rng = np.random.default_rng(seed)
    p = np.zeros(P)
    p[OFF_W1:OFF_B1] = rng.uniform(-1, 1, H * F) * np.sqrt(6.0 / F)
    p[OFF_W2:OFF_B2] = rng.uniform(-1, 1, H * H) * np.sqrt(6.0 / H)
    p[OFF_W3:OFF_B3] = rng.uniform(-1, 1, H) * np.sqrt(3.0 / H)
    p[OFF_B1:OFF_W2] = rng.uniform(0.05, 0.2, H)
    p[OFF_B2:OFF_W3] = rng.uniform(0.05, 0.2, H)
    return p

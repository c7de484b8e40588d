"""FP64 restatement of the budget predictor's forward and backward pass --
TEST INFRASTRUCTURE ONLY (the checker of the HIP kernels in
ptwalker_amd/csrc/mlp.hip).

Network: x (6) -> standardize -> h1 = relu(W1 x + b1) (64) -> h2 = relu(W2 h1 + b2)
(64) -> y = W3 h2 + b3, with W1, W2, W3 rounded to BF16 (round to nearest even)
and biases in full precision; loss = mean((y - (t - y_mean) / y_std)^2).
Gradients are with respect to the FP32 master parameters (straight through
the BF16 rounding).
"""
import numpy as np

O_W1, O_B1, O_W2, O_B2, O_W3, O_B3 = 0, 384, 448, 4544, 4608, 4672
NUM_PARAMS = 4673


def bf16(a):
    """Round float32 values to BF16 (nearest even), returned as float64."""
    b = np.asarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    b = (b + 0x7FFF + ((b >> 16) & 1)) & 0xFFFF0000
    return b.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(p, rounding=True):
    """rounding=False keeps the weights in full precision (used to check the
    backward formulas against finite differences)."""
    r = bf16 if rounding else (lambda a: np.asarray(a, dtype=np.float64))
    p = np.asarray(p, dtype=np.float64 if not rounding else np.float32)
    W1 = r(p[O_W1:O_B1]).reshape(64, 6)
    b1 = np.asarray(p[O_B1:O_W2], dtype=np.float64)
    W2 = r(p[O_W2:O_B2]).reshape(64, 64)
    b2 = np.asarray(p[O_B2:O_W3], dtype=np.float64)
    W3 = r(p[O_W3:O_B3])
    b3 = float(p[O_B3])
    return W1, b1, W2, b2, W3, b3


def forward(p, norm, x, rounding=True):
    """Predicted log budgets (de-standardized)."""
    W1, b1, W2, b2, W3, b3 = unpack(p, rounding)
    xs = (np.asarray(x, np.float64) - norm[:6]) / norm[6:12]
    h1 = np.maximum(xs @ W1.T + b1, 0)
    h2 = np.maximum(h1 @ W2.T + b2, 0)
    y = h2 @ W3 + b3
    return y * norm[13] + norm[12]


def loss_and_grad(p, norm, x, t, rounding=True):
    W1, b1, W2, b2, W3, b3 = unpack(p, rounding)
    n = len(x)
    xs = (np.asarray(x, np.float64) - norm[:6]) / norm[6:12]
    ts = (np.asarray(t, np.float64) - norm[12]) / norm[13]
    z1 = xs @ W1.T + b1
    h1 = np.maximum(z1, 0)
    z2 = h1 @ W2.T + b2
    h2 = np.maximum(z2, 0)
    y = h2 @ W3 + b3
    e = y - ts
    loss = float(np.mean(e * e))
    dy = 2 * e / n
    g = np.zeros(NUM_PARAMS)
    g[O_W3:O_B3] = dy @ h2
    g[O_B3] = dy.sum()
    d2 = np.outer(dy, W3) * (z2 > 0)
    g[O_W2:O_B2] = (d2.T @ h1).ravel()
    g[O_B2:O_W3] = d2.sum(axis=0)
    d1 = (d2 @ W2) * (z1 > 0)
    g[O_W1:O_B1] = (d1.T @ xs).ravel()
    g[O_B1:O_W2] = d1.sum(axis=0)
    return loss, g

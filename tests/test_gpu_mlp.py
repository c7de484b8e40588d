"""The budget predictor's HIP kernels against the FP64 oracle (oracle/mlp.py):
forward values and gradients within FP32 tolerances, deterministic training,
and training that actually fits."""
import numpy as np
import pytest

import ptwalker_amd as ptw
from oracle import mlp as omlp

pytestmark = pytest.mark.gpu


def data(n, seed):
    rng = np.random.default_rng(seed)
    x = rng.normal(0, 1, (n, 6)).astype(np.float32) * np.array([1, 2, 0.5, 1, 3, 1], dtype=np.float32)
    t = (np.sin(x[:, 0]) + 0.5 * x[:, 1] - 0.3 * x[:, 2] ** 2 + 0.2 * x[:, 3] * x[:, 4]).astype(np.float32)
    return x, t


def test_forward_and_gradients_match_oracle(ctx):
    net = ptw.BudgetMLP(ctx, seed=3)
    x, t = data(1000, 1)
    mean, std = x.mean(0), x.std(0)
    net.set_normalization(mean, std, float(t.mean()), float(t.std()))
    p, norm = net.params()
    y = net.forward(x)
    yo = omlp.forward(p, norm.astype(np.float64), x)
    assert np.max(np.abs(y - yo)) < 1e-4 * (1 + np.max(np.abs(yo)))
    g, loss = net.gradients(x, t)
    lo, go = omlp.loss_and_grad(p, norm.astype(np.float64), x, t)
    assert abs(loss - lo) < 1e-5 * (1 + lo)
    scale = np.max(np.abs(go))
    assert np.max(np.abs(g - go)) < 2e-5 * scale, np.max(np.abs(g - go)) / scale
    # a ragged row count (not a multiple of the 64-row block)
    g2, l2 = net.gradients(x[:77], t[:77])
    lo2, go2 = omlp.loss_and_grad(p, norm.astype(np.float64), x[:77], t[:77])
    assert np.max(np.abs(g2 - go2)) < 2e-5 * np.max(np.abs(go2))
    net.close()


def test_training_fits_and_is_deterministic(ctx):
    x, t = data(4096, 2)
    xv, tv = data(1024, 3)
    runs = []
    for _ in range(2):
        net = ptw.BudgetMLP(ctx, seed=5)
        net.set_normalization(x.mean(0), x.std(0), float(t.mean()), float(t.std()))
        before = float(np.mean((net.forward(xv) - tv) ** 2))
        net.train(x, t, epochs=60, batch=256, lr=2e-3, seed=9)
        after = float(np.mean((net.forward(xv) - tv) ** 2))
        runs.append((net.params()[0], before, after))
        net.close()
    assert np.array_equal(runs[0][0], runs[1][0]), "training is not deterministic"
    before, after = runs[0][1], runs[0][2]
    assert after < 0.1 * before, (before, after)
    assert after < 0.1 * float(np.var(tv))

"""Pins the FP64 MLP oracle: BF16 rounding against hand-checked values, and
its backward formulas against central finite differences of its own forward
pass (full-precision weights, where the loss is smooth)."""
import numpy as np

from oracle import mlp


def test_bf16_rounding_known_values():
    f = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -2.5, 1.0 + 2 ** -7, 3.0e38], dtype=np.float32)
    got = mlp.bf16(f)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7 -> even (1); 1 + 3*2^-9 rounds up
    assert got[0] == 1.0 and got[1] == 1.0 and got[2] == 1.0 + 2 ** -7
    assert got[3] == -2.5 and got[4] == 1.0 + 2 ** -7
    assert np.isfinite(got[5])


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(0)
    p = rng.normal(0, 0.3, mlp.NUM_PARAMS)
    norm = np.array([0.1, -0.2, 0.3, 0.0, 0.5, -1.0, 1.0, 2.0, 0.5, 1.5, 1.0, 3.0, 0.7, 1.3])
    x = rng.normal(0, 1, (50, 6))
    t = rng.normal(0, 1, 50)
    loss, g = mlp.loss_and_grad(p, norm, x, t, rounding=False)
    idx = list(range(0, 384, 37)) + list(range(384, 448, 9)) + list(range(448, 4544, 211)) + \
        list(range(4544, 4608, 11)) + list(range(4608, 4673, 7))
    h = 1e-6
    for i in idx:
        q = p.copy()
        q[i] += h
        lp, _ = mlp.loss_and_grad(q, norm, x, t, rounding=False)
        q[i] -= 2 * h
        lm, _ = mlp.loss_and_grad(q, norm, x, t, rounding=False)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) < 1e-6 + 1e-4 * abs(fd), (i, fd, g[i])
    # the forward used by the loss is the one exported
    y = mlp.forward(p, norm, x, rounding=False)
    ts = (t - norm[12]) / norm[13]
    assert abs(np.mean(((y - norm[12]) / norm[13] - ts) ** 2) - loss) < 1e-12

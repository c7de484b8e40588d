"""Generate (features, required budget) pairs on R-MAT graphs with the engine
itself, train the budget predictor on the device, calibrate and save the
policy.

    python -m ptwalker_amd.train_budget --scale 20 --graphs 5 --sources 20000 \
        --out ptwalker_amd/models/budget_policy.json
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

from . import Context
from .budget import error_constants, features, fit_policy, labels_from_constants

EPS = (1e-3, 3e-4)


def sample_sources(row_ptr, n, rng):
    """Half degree-weighted, half uniform, so the features cover hubs and the tail."""
    deg = np.diff(row_ptr.astype(np.int64)).astype(np.float64)
    V = len(deg)
    a = rng.choice(V, size=n // 2, p=deg / deg.sum())
    b = rng.integers(0, V, size=n - n // 2)
    return np.concatenate([a, b])


def generate(ctx: Context, scale: int, graph_seed: int, n_sources: int, tp: float, rng, log=print):
    t0 = time.time()
    ctx.rmat(scale=scale, edge_factor=16, seed=graph_seed)
    rp, col = ctx.export_csr()
    src = sample_sources(rp, n_sources, rng)
    f = features(ctx, rp, col, src, tp, EPS[0], stream_base=(1 << 30) + graph_seed * (1 << 20))
    c, l1 = error_constants(ctx, src, tp, seed=graph_seed)
    xs, ys = [], []
    for eps in EPS:
        x = f.copy()
        x[:, 5] = np.log(eps)
        xs.append(x)
        ys.append(labels_from_constants(c, eps))
    log(f"graph seed {graph_seed}: {n_sources} sources, c_s median {np.median(c):.3g}, "
        f"sub-walk L1 cv {np.mean(l1.std(1) / l1.mean(1)):.3f}, {time.time() - t0:.1f} s")
    return np.concatenate(xs), np.concatenate(ys), c


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--scale", type=int, default=20)
    ap.add_argument("--graphs", type=int, default=5)
    ap.add_argument("--sources", type=int, default=20000, help="sources per graph (2 labels each)")
    ap.add_argument("--tp", type=float, default=0.2)
    ap.add_argument("--epochs", type=int, default=200)
    ap.add_argument("--out", default=os.path.join(os.path.dirname(__file__), "models", "budget_policy.json"))
    ap.add_argument("--data-out", default=None, help="also save the training pairs (.npz)")
    args = ap.parse_args(argv)
    rng = np.random.default_rng(1)
    ctx = Context(0)
    X, Y = [], []
    for g in range(args.graphs):
        x, y, _ = generate(ctx, args.scale, 101 + g, args.sources, args.tp, rng)
        X.append(x)
        Y.append(y)
    X = np.concatenate(X)
    Y = np.concatenate(Y)
    if args.data_out:
        np.savez(args.data_out, x=X, y=Y)
    t0 = time.time()
    policy, net, (xv, yv, pv) = fit_policy(ctx, X, Y, epochs=args.epochs)
    learned = np.exp(pv) * policy.margin
    fixed = policy.fixed_budget(np.exp(xv[:, 5]))
    need = np.exp(yv)
    stats = {
        "pairs": int(len(X)),
        "train_s": round(time.time() - t0, 2),
        "val_rmse_log": policy.meta["val_rmse_log"],
        "margin": policy.margin,
        "fixed_c2": policy.fixed_c2,
        "coverage_learned": float(np.mean(learned >= need)),
        "coverage_fixed": float(np.mean(fixed >= need)),
        "work_ratio_fixed_over_learned": float(fixed.sum() / learned.sum()),
    }
    policy.meta.update(stats)
    policy.meta["tp"] = args.tp
    policy.meta["scale"] = args.scale
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    policy.save(args.out)
    print(json.dumps(stats), flush=True)
    net.close()
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())

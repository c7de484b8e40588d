"""What the trained budget predictor gains over the method without it.

The committed policy (network trained on the device on engine-generated
pairs -- 200,000 of them, from five R-MAT scale-20 graphs, seeds 101-105) is scored on a graph
it never saw (R-MAT scale 20, seed 7): for 600 sources and eps in {1e-3,
3e-4}, the required budget (c_s / eps)^2 comes from 8 independent
16384-walker sub-walks per source scored against the device FP64 solver.
The learned budgets must cover the requirement about as often as the fixed
budget does, with less total walk work."""
import os

import numpy as np
import pytest

import ptwalker_amd as ptw
from ptwalker_amd.budget import BudgetPolicy, error_constants, features, labels_from_constants

pytestmark = pytest.mark.gpu

POLICY = os.path.join(os.path.dirname(ptw.__file__), "models", "budget_policy.json")


def test_learned_budget_beats_fixed_on_unseen_graph(ctx):
    policy = BudgetPolicy.load(POLICY)
    tp = policy.meta["tp"]
    ctx.rmat(scale=20, edge_factor=16, seed=7)
    rp, col = ctx.export_csr()
    rng = np.random.default_rng(3)
    deg = np.diff(rp.astype(np.int64)).astype(np.float64)
    src = np.concatenate([rng.choice(len(deg), 300, p=deg / deg.sum()), rng.integers(0, len(deg), 300)])
    f = features(ctx, rp, col, src, tp, 1e-3, stream_base=(1 << 30) + 7 * (1 << 20))
    c, _ = error_constants(ctx, src, tp, seed=7)
    net = policy.network(ctx)
    learned, fixed, need, fallback = [], [], [], []
    for eps in (1e-3, 3e-4):
        x = f.copy()
        x[:, 5] = np.log(eps)
        w, ood = policy.budgets(net, x)
        learned.append(w)
        fallback.append(ood)
        fixed.append(np.clip(policy.fixed_budget(np.full(len(x), eps)), policy.w_min, policy.w_max))
        need.append(np.exp(labels_from_constants(c, eps)))
    net.close()
    learned, fixed, need, fallback = map(np.concatenate, (learned, fixed, need, fallback))
    cov_l = float(np.mean(learned >= need))
    cov_f = float(np.mean(fixed >= need))
    ratio = float(fixed.sum() / learned.sum())
    # What any predictor can gain here: a fixed budget tuned on this very set
    # to the learned budgets' coverage, against a perfect predictor that gives
    # every query exactly its requirement (coverage 1)
    eps_of = np.repeat([1e-3, 3e-4], len(src))
    c2 = need * eps_of ** 2
    bound = float((np.quantile(c2, cov_l) / eps_of ** 2).sum() / need.sum())
    print(f"coverage learned {cov_l:.3f} fixed {cov_f:.3f}; fixed/learned walk work {ratio:.2f}; "
          f"perfect-predictor bound {bound:.2f}; fallbacks {fallback.mean():.3f}")
    assert fallback.mean() < 0.05
    assert cov_l >= 0.90 and cov_l >= cov_f - 0.03
    # measured 1.44 (seeded: the walks, graph and sources repeat exactly); the
    # margin covers the solver's and the network's float reductions
    assert ratio > 1.40
    assert ratio > 0.8 * bound


def test_calibration_on_the_serving_graph(ctx):
    """On a graph of another scale (R-MAT 18), calibrating both budget constants
    on 400 labelled sources brings held-out coverage of both methods to about
    the target, so that they are compared at equal accuracy."""
    from ptwalker_amd.budget import calibrate, realized_coverage

    policy = BudgetPolicy.load(POLICY)
    tp = policy.meta["tp"]
    ctx.rmat(scale=18, edge_factor=16, seed=11)
    rp, col = ctx.export_csr()
    deg = np.diff(rp.astype(np.int64)).astype(np.float64)
    rng = np.random.default_rng(5)
    pick = lambda n: np.concatenate([rng.choice(len(deg), n // 2, p=deg / deg.sum()), rng.integers(0, len(deg), n // 2)])
    cal_src, test_src = pick(400), pick(400)
    net = policy.network(ctx)
    cal, stats = calibrate(ctx, policy, net, cal_src)
    assert stats["coverage_learned"] >= 0.949 and stats["coverage_fixed"] >= 0.949
    f = features(ctx, rp, col, test_src, tp, 1e-3, stream_base=(1 << 30) + 11 * (1 << 20))
    c, l1 = error_constants(ctx, test_src, tp, seed=11)
    shape = l1 / l1.mean(axis=1, keepdims=True)
    cov = {"learned": [], "fixed": []}
    work = {"learned": 0.0, "fixed": 0.0}
    for eps in (1e-3, 3e-4):
        x = f.copy()
        x[:, 5] = np.log(eps)
        w, _ = cal.budgets(net, x)
        fx = np.clip(cal.fixed_budget(np.full(len(x), eps)), cal.w_min, cal.w_max)
        e = np.full(len(x), eps)
        cov["learned"].append(realized_coverage(w, c, shape, e))
        cov["fixed"].append(realized_coverage(fx, c, shape, e))
        work["learned"] += float(w.sum())
        work["fixed"] += float(fx.sum())
    net.close()
    cl, cf = float(np.mean(cov["learned"])), float(np.mean(cov["fixed"]))
    print(f"calibrated margin {cal.margin:.3f} fixed_c2 {cal.fixed_c2:.3f}; held-out realized coverage learned "
          f"{cl:.3f} fixed {cf:.3f}; fixed/learned work {work['fixed'] / work['learned']:.2f}; {stats}")
    assert cl >= 0.9 and cf >= 0.9 and abs(cl - cf) < 0.05
    # measured 1.07 at equal held-out coverage: off its training scale (20) the
    # network ranks difficulty less well (calibrated margin 1.90 here, 1.41 on
    # scale 22), so the gain shrinks but stays
    assert work["fixed"] / work["learned"] > 1.03

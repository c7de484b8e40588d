"""Learned walker budgets for PPR queries.

For a query (source s, termination probability tp, target top-100 L1 error
eps against the FP64 expected-visit vector), the engine generates its own
training data:

* features (6): log degree, log mean neighbour degree, log 2-hop degree sum
  (sum of the neighbours' degrees), top-10 mass and entropy of a 1000-walker
  pilot walk, log eps;
* label: log of the walker budget W that makes the expected top-100 L1 error
  equal eps.  Monte-Carlo error of the unbiased visit estimator falls as
  c_s / sqrt(W); c_s is measured by M independent sub-walks of W_m walkers
  scored against the device FP64 solver (ptw_ppr_exact), and
  W = (c_s / eps)^2.

The network (ptw_mlp, HIP kernels) is trained on these pairs.  A policy turns
its output into budgets: a safety factor calibrated on held-out queries so
that at least `coverage` of them meet eps, clipped to [w_min, w_max], with the
fixed budget as the fallback for inputs outside the training range.
"""
from __future__ import annotations

import json
import math
from dataclasses import dataclass, field

import numpy as np

from . import BudgetMLP, Context

PILOT_WALKERS = 1000
TOP = 100


def graph_features(row_ptr: np.ndarray, col: np.ndarray, sources) -> np.ndarray:
    """[n, 3]: log degree, log mean neighbour degree, log 2-hop degree sum."""
    rp = np.asarray(row_ptr, dtype=np.int64)
    deg = np.diff(rp).astype(np.float64)
    out = np.zeros((len(sources), 3))
    for i, s in enumerate(np.asarray(sources, dtype=np.int64)):
        nb = col[rp[s]:rp[s + 1]]
        d = deg[s]
        two_hop = deg[nb].sum() if len(nb) else 0.0
        out[i] = (math.log(max(d, 1.0)), math.log(max(two_hop / max(d, 1.0), 1.0)), math.log(max(two_hop, 1.0)))
    return out


def pilot_features(ctx: Context, sources, tp: float, *, seed: int = 7, stream_base: int = 1 << 30,
                   group: int = 4096) -> np.ndarray:
    """[n, 2]: top-10 mass and entropy of a PILOT_WALKERS-walker pilot per source."""
    sources = np.asarray(sources)
    out = np.zeros((len(sources), 2))
    for g0 in range(0, len(sources), group):
        src = sources[g0:g0 + group]
        n = len(src)
        r = ctx.batch_ppr(src, [PILOT_WALKERS] * n, [tp] * n, streams=range(stream_base + g0, stream_base + g0 + n),
                          topk=10, seed=seed)
        total = PILOT_WALKERS + r.steps.astype(np.float64)
        out[g0:g0 + n, 0] = r.topk_count.astype(np.float64).sum(axis=1) / total
        out[g0:g0 + n, 1] = r.entropy
    return out


def features(ctx: Context, row_ptr, col, sources, tp: float, eps, **kw) -> np.ndarray:
    """[n, 6] feature rows; eps is a scalar or one value per source."""
    n = len(sources)
    eps = np.broadcast_to(np.asarray(eps, dtype=np.float64), (n,))
    f = np.zeros((n, 6))
    f[:, 0:3] = graph_features(row_ptr, col, sources)
    f[:, 3:5] = pilot_features(ctx, sources, tp, **kw)
    f[:, 5] = np.log(eps)
    return f


def error_constants(ctx: Context, sources, tp: float, *, sub_walks: int = 8, sub_walkers: int = 16384,
                    seed: int = 11, group_sources: int = 512, tol: float = 1e-7):
    """c_s per source: mean top-100 L1 of `sub_walks` independent sub-walks of
    `sub_walkers` walkers, times sqrt(sub_walkers), against the FP64 solver."""
    sources = np.asarray(sources)
    n = len(sources)
    total_exact = 1.0 + 1.0 / tp  # sum of the expected-visit vector (no dangling vertices)
    c = np.zeros(n)
    l1_all = np.zeros((n, sub_walks))
    for g0 in range(0, n, group_sources):
        src = sources[g0:g0 + group_sources]
        m = len(src)
        _, tv, tval, _ = ctx.ppr_exact(src, [tp] * m, tol=tol, full=False, k=TOP)
        qs = np.repeat(src, sub_walks)
        streams = np.arange(len(qs)) + g0 * sub_walks
        r = ctx.batch_ppr(qs, [sub_walkers] * len(qs), [tp] * len(qs), streams=streams, topk=0, seed=seed)
        qi = np.repeat(np.arange(len(qs), dtype=np.uint32), TOP)
        verts = np.repeat(tv, sub_walks, axis=0).reshape(-1)
        cnt = ctx.batch_gather(qi, verts).reshape(len(qs), TOP).astype(np.float64)
        visits = sub_walkers + r.steps.astype(np.float64)
        est = cnt / visits[:, None]
        exact = np.repeat(tval / total_exact, sub_walks, axis=0)
        l1 = np.abs(est - exact).sum(axis=1).reshape(m, sub_walks)
        l1_all[g0:g0 + m] = l1
        c[g0:g0 + m] = l1.mean(axis=1) * math.sqrt(sub_walkers)
    return c, l1_all


def hoeffding_c2(k: int = TOP, delta: float = 0.05) -> float:
    """Numerator of the distribution-free fixed budget W = c2 / eps^2: the L1
    concentration bound for an empirical distribution over m = k + 1 cells
    (the top-k vertices and the rest) from n independent samples,
    P(|p_hat - p|_1 >= eps) <= (2^m - 2) exp(-n eps^2 / 2)  (Weissman et al.,
    HP Labs tech. report HPL-2003-97), solved for n at failure probability
    delta with one sample per walker."""
    return 2.0 * ((k + 1) * math.log(2.0) + math.log(1.0 / delta))


def labels_from_constants(c, eps) -> np.ndarray:
    """log W with W = (c / eps)^2."""
    return 2.0 * (np.log(c) - np.log(eps))


@dataclass
class BudgetPolicy:
    """Turns predicted log budgets into walker budgets."""
    params: list
    norm: list
    margin: float = 1.0           # multiplicative safety factor on the predicted budget
    fixed_c2: float = 1e-1        # the plain method's budget is fixed_c2 / eps^2 for every query
    w_min: int = 1000
    w_max: int = 4 * 10 ** 7
    feat_lo: list = field(default_factory=lambda: [-1e9] * 6)
    feat_hi: list = field(default_factory=lambda: [1e9] * 6)
    meta: dict = field(default_factory=dict)

    def save(self, path: str):
        with open(path, "w") as f:
            json.dump(self.__dict__, f)

    @staticmethod
    def load(path: str) -> "BudgetPolicy":
        with open(path) as f:
            return BudgetPolicy(**json.load(f))

    def network(self, ctx: Context) -> BudgetMLP:
        net = BudgetMLP(ctx, seed=1)
        n = np.asarray(self.norm, dtype=np.float32)
        net.set_normalization(n[:6], n[6:12], float(n[12]), float(n[13]))
        net.set_params(np.asarray(self.params, dtype=np.float32))
        return net

    def fixed_budget(self, eps) -> np.ndarray:
        """The plain method: one budget per target error, the same for every query."""
        return np.ceil(self.fixed_c2 / np.asarray(eps, dtype=np.float64) ** 2).astype(np.int64)

    def budgets(self, net: BudgetMLP, feats: np.ndarray):
        """Returns (budgets, used_fallback mask)."""
        logw = net.forward(feats.astype(np.float32)).astype(np.float64)
        w = np.exp(logw) * self.margin
        ood = np.any((feats < np.asarray(self.feat_lo)) | (feats > np.asarray(self.feat_hi)), axis=1)
        w = np.where(ood, self.fixed_budget(np.exp(feats[:, 5])), w)
        return np.clip(np.ceil(w), self.w_min, self.w_max).astype(np.int64), ood


def realized_coverage(budgets, c, shapes, eps) -> float:
    """Share of queries whose realized top-k L1 stays within eps at the given
    budgets.  A query's realized error at W walkers is c_s * r / sqrt(W), with
    r drawn from its sub-walk errors divided by their mean (`shapes`, one row
    per query): the Monte-Carlo scatter around the expected error."""
    w = np.asarray(budgets, dtype=np.float64)
    return float(np.mean(np.asarray(c)[:, None] * shapes / np.sqrt(w)[:, None] <= np.asarray(eps)[:, None]))


def _smallest(f, target, lo=1e-4, hi=1e4, iters=60) -> float:
    """Smallest x in [lo, hi] with f(x) >= target, for f non-decreasing."""
    if f(hi) < target:
        return hi
    for _ in range(iters):
        mid = math.sqrt(lo * hi)
        if f(mid) >= target:
            hi = mid
        else:
            lo = mid
    return hi


def calibrate(ctx: Context, policy: "BudgetPolicy", net: BudgetMLP, sources, *, eps=(1e-3, 3e-4),
              coverage: float = 0.95, seed: int = 97, stream_base: int = (1 << 30) + (1 << 29)):
    """Re-fit the policy's two scale constants on the graph it will serve.

    The network ranks queries by difficulty; how many walkers a unit of
    predicted difficulty needs depends on the graph (its size, its degree
    tail).  On `sources` of the resident graph this measures each source's
    error constant c_s and its Monte-Carlo scatter (8 independent sub-walks
    scored against the FP64 solver, as in training), then sets
      fixed_c2 to the smallest constant whose fixed budgets fixed_c2 / eps^2, and
      margin   to the smallest factor whose learned budgets margin * exp(net)
               (fixed budgets for out-of-range features)
    keep the *realized* top-k L1 within eps for `coverage` of the queries, so
    that both methods are compared at the same realized accuracy.
    Returns (calibrated policy, calibration statistics)."""
    rp, col = ctx.export_csr()
    tp = float(policy.meta.get("tp", 0.2))
    sources = np.asarray(sources)
    f = features(ctx, rp, col, sources, tp, eps[0], seed=seed, stream_base=stream_base)
    c, l1 = error_constants(ctx, sources, tp, seed=seed)
    shape = l1 / l1.mean(axis=1, keepdims=True)
    xs, cs, shapes, es = [], [], [], []
    for e in eps:
        x = f.copy()
        x[:, 5] = np.log(e)
        xs.append(x)
        cs.append(c)
        shapes.append(shape)
        es.append(np.full(len(c), e))
    x, cc, sh, ee = map(np.concatenate, (xs, cs, shapes, es))
    pred = np.exp(net.forward(x.astype(np.float32)).astype(np.float64))
    ood = np.any((x < np.asarray(policy.feat_lo)) | (x > np.asarray(policy.feat_hi)), axis=1)
    clip = lambda w: np.clip(np.ceil(w), policy.w_min, policy.w_max)
    fixed_of = lambda k: clip(k / ee ** 2)
    fixed_c2 = _smallest(lambda k: realized_coverage(fixed_of(k), cc, sh, ee), coverage)
    learned_of = lambda m: np.where(ood, fixed_of(fixed_c2), clip(m * pred))
    margin = _smallest(lambda m: realized_coverage(learned_of(m), cc, sh, ee), coverage)
    meta = dict(policy.meta)
    meta.update({"calibration_sources": int(len(sources)), "trained_margin": policy.margin,
                 "trained_fixed_c2": policy.fixed_c2})
    cal = BudgetPolicy(params=policy.params, norm=policy.norm, margin=margin, fixed_c2=fixed_c2,
                       w_min=policy.w_min, w_max=policy.w_max, feat_lo=policy.feat_lo, feat_hi=policy.feat_hi,
                       meta=meta)
    learned, fixed = learned_of(margin), fixed_of(fixed_c2)
    stats = {"sources": int(len(sources)), "margin": margin, "fixed_c2": fixed_c2,
             "c_median": float(np.median(c)), "scatter_cv": float(np.mean(l1.std(1) / l1.mean(1))),
             "ood_frac": float(ood.mean()),
             "coverage_learned": realized_coverage(learned, cc, sh, ee),
             "coverage_fixed": realized_coverage(fixed, cc, sh, ee),
             "work_ratio_fixed_over_learned": float(fixed.sum() / learned.sum())}
    return cal, stats


def fit_policy(ctx: Context, x: np.ndarray, y: np.ndarray, *, coverage: float = 0.95, seed: int = 1,
               epochs: int = 200, batch: int = 512, lr: float = 2e-3, val_frac: float = 0.15):
    """Train the network on (x, y = log budget) and calibrate the policy on a held-out split."""
    rng = np.random.default_rng(seed)
    idx = rng.permutation(len(x))
    nv = int(len(x) * val_frac)
    va, tr = idx[:nv], idx[nv:]
    net = BudgetMLP(ctx, seed=seed)
    net.set_normalization(x[tr].mean(0), x[tr].std(0) + 1e-6, float(y[tr].mean()), float(y[tr].std() + 1e-6))
    loss = net.train(x[tr].astype(np.float32), y[tr].astype(np.float32), epochs=epochs, batch=batch, lr=lr,
                     seed=seed)
    pred = net.forward(x[va].astype(np.float32)).astype(np.float64)
    resid = y[va] - pred  # log(W_required / W_predicted)
    margin = float(math.exp(np.quantile(resid, coverage)))
    # c^2 = W eps^2; the fixed budget covers `coverage` of the training queries
    fixed_c2 = float(np.quantile(np.exp(y[tr] + 2.0 * x[tr, 5]), coverage))
    params, norm = net.params()
    lo = x[tr].mean(0) - 4 * x[tr].std(0)
    hi = x[tr].mean(0) + 4 * x[tr].std(0)
    policy = BudgetPolicy(params=params.tolist(), norm=norm.tolist(), margin=margin, fixed_c2=fixed_c2,
                          feat_lo=lo.tolist(), feat_hi=hi.tolist(),
                          meta={"train_rows": int(len(tr)), "val_rows": int(nv), "final_loss": float(loss),
                                "val_rmse_log": float(np.sqrt(np.mean(resid ** 2))),
                                "coverage": coverage})
    return policy, net, (x[va], y[va], pred)

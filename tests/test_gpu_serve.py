"""Serving loop (ptw_serve_replay): every answer equals the oracle's replay of
that request's own walk (its Philox stream, its budget), the same alone and
in any mix, under tight HBM admission, and with audits that top up short
requests by the 10x walk."""
import numpy as np
import pytest

import os

import oracle
import ptwalker_amd as ptw
from ptwalker_amd.budget import BudgetPolicy
from ptwalker_amd.workloads import serve_requests

pytestmark = pytest.mark.gpu

TP = 0.2
SEED = 5
K = 20


@pytest.fixture(scope="module")
def g12(ctx):
    gi = ctx.rmat(scale=12, edge_factor=16, seed=4)
    rp, col = ctx.export_csr()
    return gi, rp, col


def requests(gi, n, seed, rate_qps=2000.0):
    rng = np.random.default_rng(seed)
    src = rng.integers(0, gi.num_vertices, n)
    src[0] = gi.max_degree_vertex
    eps = rng.choice([0.02, 0.05], n)
    arr = np.cumsum(rng.exponential(1.0 / rate_qps, n))
    return src, eps, arr


def expected_topk(rp, col, src, walkers, stream, k=K):
    ref = oracle.walk_replay(rp, col, oracle.APP_PPR, int(walkers), tp=TP, seed=SEED, source=int(src),
                             start_mode=oracle.START_SOURCE, stream=int(stream))
    return oracle.topk(ref["counts"], rp, k)


def small_net(ctx):
    """An untrained predictor whose outputs land in [1e3, 2e4] walkers."""
    net = ptw.BudgetMLP(ctx, seed=3)
    net.set_normalization(np.zeros(6), np.full(6, 3.0), float(np.log(4000.0)), 0.5)
    return net


def check_against_replay(rp, col, src, res, ids=None, audit_stream=False):
    ids = range(len(src)) if ids is None else ids
    for i, r in enumerate(ids):
        if audit_stream and res.topped_up[i]:
            v, c = expected_topk(rp, col, src[i], 10 * res.budget[i], (1 << 31) + r)
        else:
            v, c = expected_topk(rp, col, src[i], res.budget[i], r)
        assert np.array_equal(res.topk_vertex[i], v), i
        assert np.array_equal(res.topk_count[i], c), i


def test_serve_fixed_matches_replay(ctx, g12):
    gi, rp, col = g12
    src, eps, arr = requests(gi, 24, seed=1)
    res = ctx.serve(src, eps, arr, tp=TP, seed=SEED, topk=K, budget_mode=ptw.BUDGET_FIXED, fixed_c2=4.0,
                    w_min=500, w_max=20000, max_batch=16)
    assert np.array_equal(res.budget, np.clip(np.ceil(4.0 / eps.astype(np.float32).astype(np.float64) ** 2),
                                              500, 20000).astype(np.uint64))
    check_against_replay(rp, col, src, res)
    st = res.stats
    assert st["launches"] >= 2  # 24 requests at max_batch 16
    assert np.all(res.done_s >= arr) and np.all(res.admit_s >= arr)
    assert 0 < st["p50_ms"] <= st["p99_ms"]
    assert st["walk_steps"] > 0 and st["queries_per_s"] > 0


def test_serve_learned_alone_equals_in_mix(ctx, g12):
    gi, rp, col = g12
    net = small_net(ctx)
    src, eps, arr = requests(gi, 40, seed=2, rate_qps=5000.0)
    kw = dict(mlp=net, tp=TP, seed=SEED, topk=K, budget_mode=ptw.BUDGET_LEARNED, margin=1.0, w_min=1000,
              w_max=20000, max_batch=24)
    mix = ctx.serve(src, eps, arr, **kw)
    assert mix.stats["launches"] >= 2 and mix.stats["fallbacks"] == 0
    assert len(set(mix.budget.tolist())) > 5  # budgets differ by request
    check_against_replay(rp, col, src, mix)
    for i in (0, 7, 39):
        alone = ctx.serve(src[i:i + 1], eps[i:i + 1], [0.0], ids=[i], **kw)
        assert alone.budget[0] == mix.budget[i]
        assert np.array_equal(alone.topk_vertex[0], mix.topk_vertex[i])
        assert np.array_equal(alone.topk_count[0], mix.topk_count[i])
    net.close()


def test_serve_admission_under_small_hbm(ctx, g12):
    gi, rp, col = g12
    src, eps, arr = requests(gi, 30, seed=3, rate_qps=1e6)  # all arrive at once
    kw = dict(tp=TP, seed=SEED, topk=K, budget_mode=ptw.BUDGET_FIXED, fixed_c2=4.0, w_min=500, w_max=20000,
              max_batch=64)
    free = ctx.serve(src, eps, arr, **kw)
    tight = ctx.serve(src, eps, arr, hbm_limit=200_000, **kw)  # about three 64 KiB hash slices
    assert tight.stats["admission_waits"] > 0
    assert tight.stats["launches"] > free.stats["launches"]
    assert np.array_equal(tight.topk_vertex, free.topk_vertex)
    assert np.array_equal(tight.topk_count, free.topk_count)


def test_serve_audit_tops_up(ctx, g12):
    gi, rp, col = g12
    src, _, arr = requests(gi, 16, seed=4)
    kw = dict(tp=TP, seed=SEED, topk=K, budget_mode=ptw.BUDGET_FIXED, fixed_c2=1e-9, w_min=1000, w_max=1000,
              audit_rate=1.0, max_batch=32)
    # a tolerance no 1000-walker estimate meets: every audited request is replaced by its 10x walk
    strict = ctx.serve(src, np.full(16, 1e-6), arr, **kw)
    assert strict.audited.all() and strict.topped_up.all()
    assert strict.stats["audited"] == 16 and strict.stats["topped_up"] == 16
    check_against_replay(rp, col, src, strict, audit_stream=True)
    # a tolerance every estimate meets: the answers are the main walks
    loose = ctx.serve(src, np.full(16, 10.0), arr, **kw)
    assert loose.audited.all() and not loose.topped_up.any()
    check_against_replay(rp, col, src, loose)


def test_serve_audit_sampling_rate(ctx, g12):
    gi, _, _ = g12
    src, eps, arr = requests(gi, 400, seed=6, rate_qps=1e6)
    res = ctx.serve(src, eps, arr, tp=TP, seed=SEED, topk=K, fixed_c2=1e-9, w_min=200, w_max=200,
                    audit_rate=0.25, max_batch=1024)
    assert 0.18 < res.audited.mean() < 0.32


def test_serve_rejects_bad_config(ctx, g12):
    gi, _, _ = g12
    src, eps, arr = requests(gi, 4, seed=7)
    with pytest.raises(ptw.PtwError):
        ctx.serve(src, eps, arr, budget_mode=ptw.BUDGET_LEARNED)  # no model
    net = small_net(ctx)
    with pytest.raises(ptw.PtwError):
        ctx.serve(src, eps, arr, mlp=net, budget_mode=ptw.BUDGET_LEARNED, topk=5)
    net.close()
    with pytest.raises(ptw.PtwError):
        ctx.serve(src, eps, arr, ids=[0, 1, 2, 1 << 30])
    with pytest.raises(ptw.PtwError):
        ctx.serve(src, eps, arr, w_max=10 ** 9)


def test_config_serve_scale22_committed_policy(ctx22):
    """The serving configuration at its size: R-MAT scale 22 resident, the
    committed learned policy (pilot walk + MLP on the device), requests with
    eps in {1e-3, 3e-4} arriving at 2000/s, 5 % audited.  Sampled answers equal
    the replay of their own walk (budget, stream; the 10x audit stream when
    topped up), and requests served alone give the same budget and answer."""
    c, gi, rp, col = ctx22
    pol = BudgetPolicy.load(os.path.join(os.path.dirname(ptw.__file__), "models", "budget_policy.json"))
    tp = float(pol.meta["tp"])
    net = pol.network(c)
    n = 300
    src, eps, arr = serve_requests(rp, n, 2000.0, seed=1)
    print(f"config 5: serving {n} requests on scale 22 (V {gi.num_vertices})", flush=True)
    kw = dict(mlp=net, tp=tp, seed=1, topk=100, budget_mode=ptw.BUDGET_LEARNED, margin=pol.margin,
              fixed_c2=pol.fixed_c2, w_min=pol.w_min, w_max=pol.w_max, feat_lo=pol.feat_lo, feat_hi=pol.feat_hi,
              audit_rate=0.05, max_batch=512)
    mix = c.serve(src, eps, arr, **kw)
    assert mix.stats["launches"] >= 2
    print(f"config 5: {mix.stats['launches']} launches, {mix.stats['fallbacks']} fallbacks", flush=True)
    assert mix.stats["fallbacks"] < n // 4
    assert len(set(mix.budget.tolist())) > n // 2  # budgets differ by request
    sample = np.unique(np.concatenate([np.linspace(0, n - 1, 10).astype(np.int64),
                                       np.flatnonzero(mix.audited)[:2], np.flatnonzero(mix.topped_up)[:2]]))
    for i in sample:
        topped = bool(mix.topped_up[i])
        w = int(mix.budget[i]) * (10 if topped else 1)
        stream = ((1 << 31) + int(i)) if topped else int(i)
        ref = oracle.walk_replay(rp, col, oracle.APP_PPR, w, tp=tp, seed=1, source=int(src[i]),
                                 start_mode=oracle.START_SOURCE, stream=stream)
        v, cnt = oracle.topk(ref["counts"], rp, 100)
        assert np.array_equal(mix.topk_vertex[i], v), i
        assert np.array_equal(mix.topk_count[i], cnt), i
    for i in sample[:4]:
        alone = c.serve(src[i:i + 1], eps[i:i + 1], [0.0], ids=[int(i)], **kw)
        assert alone.budget[0] == mix.budget[i]
        assert np.array_equal(alone.topk_vertex[0], mix.topk_vertex[i])
        assert np.array_equal(alone.topk_count[0], mix.topk_count[i])
    net.close()


def test_serve_overlap_short_request_finishes_first(ctx):
    """Continuous batching with work overlapping on the device: a long main walk
    (1e8 walkers) starts at t = 0; a short request arriving 4 ms later is
    admitted on the idle lane while the long walk runs and finishes before it.
    Both answers equal the same requests served alone, and the short one
    equals the oracle's replay of its walk."""
    gi = ctx.rmat(scale=18, edge_factor=16, seed=3)
    rp, col = ctx.export_csr()
    big, small = int(gi.max_degree_vertex), 5
    src = np.array([big, small])
    eps = np.array([1e-6, 10.0])
    arr = np.array([0.0, 0.004])
    kw = dict(tp=TP, seed=SEED, topk=K, budget_mode=ptw.BUDGET_FIXED, fixed_c2=1.0, w_min=20000,
              w_max=100_000_000, max_batch=8)
    res = ctx.serve(src, eps, arr, **kw)
    assert res.budget.tolist() == [100_000_000, 20000]
    print(f"overlap: long main {res.admit_s[0]:.4f}-{res.done_s[0]:.4f} s, short "
          f"{res.admit_s[1]:.4f}-{res.done_s[1]:.4f} s", flush=True)
    assert res.admit_s[1] < res.done_s[0]  # admitted while the long main ran
    assert res.done_s[1] < res.done_s[0]   # and answered before it
    assert res.stats["launches"] == 2
    for i in (0, 1):
        alone = ctx.serve(src[i:i + 1], eps[i:i + 1], [0.0], ids=[i], **kw)
        assert np.array_equal(alone.topk_vertex[0], res.topk_vertex[i]), i
        assert np.array_equal(alone.topk_count[0], res.topk_count[i]), i
    v, c = expected_topk(rp, col, small, 20000, 1)
    assert np.array_equal(res.topk_vertex[1], v) and np.array_equal(res.topk_count[1], c)

"""Autotuner on the device: every candidate launch shape must reproduce the
starting point's results exactly (the tuner checks this itself and raises),
and the chosen shape is left on the context and cached."""
import json

import numpy as np
import pytest

import ptwalker_amd as ptw
from ptwalker_amd import tune

pytestmark = pytest.mark.gpu


def test_tune_batch_keeps_results_and_caches(ctx, tmp_path):
    gi = ctx.rmat(scale=14, edge_factor=16, seed=3)
    default = ctx.tuning().as_dict()
    rng = np.random.default_rng(2)
    src = rng.integers(0, gi.num_vertices, 200)
    W = np.exp(rng.uniform(np.log(1e3), np.log(3e5), 200)).astype(np.int64)
    tp = rng.uniform(0.15, 0.3, 200)
    path = str(tmp_path / "tune.json")
    space = {"batch_chunk": [4096, 16384], "block": [512, 1024], "dense_threshold": [50_000, 100_000],
             "topk_tile": [256, 2048]}
    r = tune.tune_batch(ctx, src, W, tp, topk=100, repeats=1, rounds=1, space=space, cache_path=path)
    # the tuner starts from the planner's default threshold (1 with logged dense queries)
    start_thr = 1 if default["visit_log"] else 100_000
    assert all(r.best[k] in space[k] or r.best[k] == default.get(k, start_thr) for k in r.best)
    assert r.best["topk_tile"] >= 200  # 256 cannot hold 2k candidates
    assert len(r.trials) >= 5 and r.best_ms <= r.start_ms
    now = ctx.tuning().as_dict()
    assert all(now[k] == v for k, v in r.best.items() if k != "dense_threshold")
    entry = json.load(open(path))
    assert len(entry) == 1 and list(entry.values())[0]["best"] == r.best
    ctx.set_tuning(**default)
    assert tune.apply(ctx, tune.batch_key(W, tp, 100), cache_path=path) == r.best
    ctx.set_tuning(**default)


def test_tune_walk_keeps_results(ctx, tmp_path):
    gi = ctx.rmat(scale=14, edge_factor=16, seed=3)
    default = ctx.tuning().as_dict()
    space = {"chunk": [1024, 4096], "lds_hubs": [0, 8192], "cache_slots": [0, 4096]}
    r = tune.tune_walk(ctx, ptw.PPR(tp=0.2, source=int(gi.max_degree_vertex)), 2_000_000, repeats=1, rounds=1,
                       space=space, cache_path=str(tmp_path / "t.json"))
    # each knob ends at one of its candidates or at the starting point
    assert all(r.best[k] in space[k] or r.best[k] == default[k] for k in r.best)
    assert len(r.trials) >= 4 and r.best_ms <= r.start_ms
    now = ctx.tuning().as_dict()
    assert all(now[k] == v for k, v in r.best.items())
    ctx.set_tuning(**default)

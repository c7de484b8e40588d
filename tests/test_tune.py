"""Autotuner search and cache logic on the CPU, with a synthetic engine."""
import types

import pytest

from ptwalker_amd import tune


def test_coordinate_descent_finds_separable_minimum():
    space = {"a": [1, 2, 3, 4], "b": [10, 20, 30]}
    cost = lambda c: (c["a"] - 3) ** 2 + (c["b"] - 20) ** 2 / 100 + 1.0
    r = tune.coordinate_descent(cost, {"a": 1, "b": 10}, space)
    assert r.best == {"a": 3, "b": 20}
    assert r.best_ms == pytest.approx(1.0)
    assert r.start_ms == pytest.approx(4 + 1 + 1)
    assert r.speedup == pytest.approx(6.0)
    # every distinct configuration is timed once
    keys = [tuple(sorted((k, v) for k, v in t.items() if k != "ms")) for t in r.trials]
    assert len(keys) == len(set(keys))


def test_coordinate_descent_skips_rejected_candidates():
    space = {"a": [1, 2, 3]}

    def cost(c):
        if c["a"] == 3:
            raise ValueError("rejected by the engine")
        return 10.0 - c["a"]

    r = tune.coordinate_descent(cost, {"a": 1}, space, valid=lambda c: c["a"] != 99)
    assert r.best == {"a": 2}
    assert any(t["ms"] == float("inf") for t in r.trials)


def test_coordinate_descent_second_round_fixes_interaction():
    # the best b depends on a: one pass is not enough from this start
    space = {"a": [0, 1], "b": [0, 1]}
    table = {(0, 0): 5.0, (1, 0): 4.0, (0, 1): 6.0, (1, 1): 1.0}
    r = tune.coordinate_descent(lambda c: table[(c["a"], c["b"])], {"a": 0, "b": 0}, space, rounds=2)
    assert r.best == {"a": 1, "b": 1}


def test_cache_round_trip(tmp_path):
    path = str(tmp_path / "tune.json")
    calls = {}

    class Info:
        num_vertices, num_edges, max_degree, weighted = 10, 40, 7, 0

    ctx = types.SimpleNamespace(info=lambda: Info(), set_tuning=lambda **kw: calls.update(kw))
    key = tune.batch_key([1000, 2000], [0.2, 0.2], 100)
    tune._store(path, f"{tune.graph_signature(ctx)}|{key}", {"best": {"batch_chunk": 8192, "dense_threshold": 5},
                                                             "best_ms": 1.0, "start_ms": 2.0})
    best = tune.apply(ctx, key, cache_path=path)
    assert best == {"batch_chunk": 8192, "dense_threshold": 5}
    assert calls == {"batch_chunk": 8192}  # the planner threshold is not a launch knob
    assert tune.apply(ctx, "other", cache_path=path) is None

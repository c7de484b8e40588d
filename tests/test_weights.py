"""Weighted neighbour choice: the library's integer alias tables equal the
oracle's independent build bit for bit, and the distribution each table
implies is the quantized weight distribution to within 2^-32 (exact rational
arithmetic).  No walk runs here (see test_gpu_weighted.py)."""
from fractions import Fraction

import numpy as np
import pytest

import oracle
import ptwalker_amd as ptw
from walkstats import random_graph


def quantized(w):
    w = np.asarray(w, dtype=np.float64)
    q = np.floor(w / w.max() * 16777216.0 + 0.5)
    return np.clip(q, 1, 16777216).astype(np.int64)


def check_tables(rp, w, thr, ali, vertices):
    for v in vertices:
        b, e = int(rp[v]), int(rp[v + 1])
        if e == b:
            continue
        q = quantized(w[b:e])
        Q = int(q.sum())
        implied = oracle.alias_distribution(rp, thr, ali, v)
        for k in range(e - b):
            assert abs(implied[k] - Fraction(int(q[k]), Q)) <= Fraction(1, 2 ** 32), (v, k)
        assert sum(implied) == 1


def test_random_weights_alias_tables(host_ctx):
    host_ctx.rmat(scale=9, edge_factor=8, seed=2)
    host_ctx.random_weights(seed=7)
    rp, col = host_ctx.export_csr()
    w, thr, ali = host_ctx.export_weights()
    assert np.all(w > 0) and np.all(w <= 1)
    assert np.all(np.round(w * 2 ** 24) == w * 2 ** 24)  # on the 2^-24 grid
    othr, oali = oracle.build_alias(rp, w)
    assert np.array_equal(thr, othr) and np.array_equal(ali, oali)
    deg = np.diff(rp.astype(np.int64))
    check_tables(rp, w, thr, ali, [int(np.argmax(deg))] + list(range(0, len(deg), 37)))
    assert host_ctx.info().weighted == 1


def test_user_weights_wide_range(host_ctx):
    rp, col = random_graph(40, 0.2, seed=4)
    host_ctx.from_csr(rp, col)
    rng = np.random.default_rng(3)
    w = np.exp(rng.uniform(-7, 7, size=len(col))).astype(np.float32)
    w[::5] = 1.0  # ties
    host_ctx.set_weights(w)
    _, thr, ali = host_ctx.export_weights()
    othr, oali = oracle.build_alias(rp, w)
    assert np.array_equal(thr, othr) and np.array_equal(ali, oali)
    check_tables(rp, w, thr, ali, range(40))


def test_equal_weights_give_full_slots(host_ctx):
    rp, col = random_graph(20, 0.3, seed=1)
    host_ctx.from_csr(rp, col)
    host_ctx.set_weights(np.full(len(col), 0.25, dtype=np.float32))
    _, thr, ali = host_ctx.export_weights()
    assert np.all(thr == 0xFFFFFFFF)
    starts = np.repeat(rp[:-1], np.diff(rp.astype(np.int64)))
    assert np.array_equal(ali, np.arange(len(col), dtype=np.uint32) - starts)


@pytest.mark.parametrize("bad", [0.0, -1.0, float("nan"), float("inf")])
def test_invalid_weights(host_ctx, bad):
    rp, col = random_graph(10, 0.3, seed=2)
    host_ctx.from_csr(rp, col)
    w = np.ones(len(col), dtype=np.float32)
    w[3] = bad
    with pytest.raises(ptw.PtwError) as e:
        host_ctx.set_weights(w)
    assert e.value.code == ptw.ERR_INVALID_ARG
    with pytest.raises(ptw.PtwError):
        host_ctx.set_weights(np.ones(len(col) - 1, dtype=np.float32))
    with pytest.raises(ValueError):
        oracle.build_alias(rp, w)


def test_oracle_alias_pinned_by_rationals():
    # two-neighbour vertex with weights 1 and 3: slot 0 keeps 1/2 of its mass
    rp = np.array([0, 2, 3, 4], dtype=np.uint32)
    w = np.array([1.0, 3.0, 1.0, 1.0], dtype=np.float32)
    thr, ali = oracle.build_alias(rp, w)
    p = oracle.alias_distribution(rp, thr, ali, 0)
    q = quantized(w[:2])
    assert abs(p[0] - Fraction(int(q[0]), int(q.sum()))) <= Fraction(1, 2 ** 32)
    assert ali[0] == 1 and thr[1] == 0xFFFFFFFF

"""GPU parity of weighted walks (alias tables): bit-exact against the oracle's
Philox replay with the oracle's own alias tables, plus the transition
frequencies against w / sum(w)."""
import numpy as np
import pytest

import oracle
import ptwalker_amd as ptw
from walkstats import random_graph

pytestmark = pytest.mark.gpu


def weighted(ctx, app, W, **kw):
    d = ptw.make_desc(app, W, flags=ptw.FLAG_WEIGHTED, **kw)
    gi = ctx.info()
    counts = ends = paths = None
    if d.app in (ptw.APP_PAGERANK, ptw.APP_PPR):
        counts = np.zeros(gi.num_vertices, dtype=np.uint64)
    elif d.app == ptw.APP_RWR:
        ends = np.zeros(W, dtype=np.uint32)
    else:
        paths = np.zeros(W * d.walk_len, dtype=np.uint32)
    out = ctx.walk(d, counts=counts, ends=ends, paths=paths)
    return d, out, counts, ends, paths


@pytest.mark.parametrize("relabel", [True, False])
def test_weighted_parity(ctx, relabel):
    gi = ctx.rmat(scale=11, edge_factor=16, seed=3, flags=0 if relabel else ptw.GRAPH_NO_RELABEL)
    ctx.random_weights(seed=5)
    rp, col = ctx.export_csr()
    w, _, _ = ctx.export_weights()
    alias = oracle.build_alias(rp, w)
    hub = int(gi.max_degree_vertex)
    for app, W in [(ptw.PPR(0.15, hub), 300000), (ptw.PageRank(0.2), 200000), (ptw.RWR(0.2, hub), 200000),
                   (ptw.DeepWalk(16), 5000)]:
        d, out, counts, ends, paths = weighted(ctx, app, W, seed=2)
        ref = oracle.walk_replay(rp, col, d.app, W, tp=d.tp, seed=2, source=d.source, start_mode=d.start_mode,
                                 walk_len=d.walk_len, alias=alias)
        assert out.steps == ref["steps"]
        if counts is not None:
            assert np.array_equal(counts, ref["counts"])
        if ends is not None:
            assert np.array_equal(ends, ref["ends"])
        if paths is not None:
            assert np.array_equal(paths.reshape(W, d.walk_len), ref["paths"])


def test_weighted_transition_frequencies(ctx):
    rp, col = random_graph(12, 0.35, seed=8)
    ctx.from_csr(rp, col)
    rng = np.random.default_rng(1)
    w = rng.uniform(0.05, 1.0, size=len(col)).astype(np.float32)
    ctx.set_weights(w)
    W, L = 20000, 30
    _, out, _, _, paths = weighted(ctx, ptw.DeepWalk(L), W, seed=4)
    paths = paths.reshape(W, L)
    a, b = paths[:, :-1].ravel(), paths[:, 1:].ravel()
    for u in range(12):
        sel = a == u
        n = int(sel.sum())
        nb = col[rp[u]:rp[u + 1]]
        p = w[rp[u]:rp[u + 1]].astype(np.float64)
        p /= p.sum()
        emp = np.array([np.sum(b[sel] == x) for x in nb], dtype=np.float64) / n
        tvd = 0.5 * np.abs(emp - p).sum()
        assert tvd < 0.03, (u, tvd)


def test_weighted_requires_weights(ctx):
    rp, col = random_graph(10, 0.3, seed=2)
    ctx.from_csr(rp, col)
    d = ptw.make_desc(ptw.PPR(0.2, 1), 100, flags=ptw.FLAG_WEIGHTED)
    with pytest.raises(ptw.PtwError) as e:
        ctx.walk(d, counts=np.zeros(10, dtype=np.uint64))
    assert e.value.code == ptw.ERR_INVALID_ARG

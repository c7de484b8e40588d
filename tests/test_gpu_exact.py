"""The device FP64 solver against the CPU oracle's power iteration (and, on
small graphs, a dense linear solve), and walk estimates scored against it."""
import numpy as np
import pytest

import oracle
import ptwalker_amd as ptw
from walkstats import random_graph

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("sem", [ptw.SEM_REFERENCE, ptw.SEM_TEXTBOOK])
@pytest.mark.parametrize("relabel", [True, False])
def test_exact_matches_cpu_oracle_small(ctx, sem, relabel):
    rp, col = random_graph(40, 0.1, seed=5, symmetric=False, dangling=2)
    ctx.from_csr(rp, col, flags=0 if relabel else ptw.GRAPH_NO_RELABEL)
    srcs = [0, 3, 17, 37]
    tps = [0.15, 0.2, 0.5, 1.0]
    x, tv, tval, it = ctx.ppr_exact(srcs, tps, semantics=sem, tol=1e-13, k=5)
    for i, (s, tp) in enumerate(zip(srcs, tps)):
        ref = oracle.power_iteration(rp, col, tp=tp, start=s, semantics=sem, mode="visits", tol=1e-15)
        dense = oracle.dense_expected(rp, col, tp=tp, start=s, semantics=sem, mode="visits")
        assert np.allclose(x[i], ref, rtol=0, atol=1e-11), i
        assert np.allclose(x[i], dense, rtol=0, atol=1e-11), i
        order = sorted(range(40), key=lambda v: (-x[i][v], v))[:5]
        assert list(tv[i]) == order and np.allclose(tval[i], x[i][order])


def test_exact_many_sources_rmat(ctx):
    gi = ctx.rmat(scale=12, edge_factor=8, seed=3)
    rp, col = ctx.export_csr()
    rng = np.random.default_rng(2)
    srcs = rng.integers(0, gi.num_vertices, 100)  # two passes of <= 64
    tps = rng.uniform(0.15, 0.3, 100)
    x, _, _, it = ctx.ppr_exact(srcs, tps, tol=1e-12)
    for i in (0, 63, 64, 99):
        ref = oracle.power_iteration(rp, col, tp=tps[i], start=int(srcs[i]), mode="visits", tol=1e-14)
        assert np.allclose(x[i], ref, rtol=0, atol=1e-10), i
        assert abs(x[i].sum() - (1 + 1 / tps[i])) < 1e-8


def test_walk_estimate_scored_against_exact(ctx):
    """Top-100 L1 of the Monte-Carlo estimate shrinks like 1/sqrt(W) and sits within a
    multiple of its predicted size."""
    gi = ctx.rmat(scale=16, edge_factor=16, seed=1)
    src = int(gi.max_degree_vertex)
    tp = 0.2
    x, _, _, _ = ctx.ppr_exact([src], [tp], tol=1e-12)
    l1 = []
    for W in (10 ** 5, 10 ** 6, 10 ** 7):
        r = ptw.walk(ctx, W, ptw.PPR(tp, src), seed=1)
        l1.append(oracle.topk_l1(r.counts, x[0], 100))
    assert l1[2] < l1[1] < l1[0]
    assert 0.05 < l1[2] / l1[0] < 0.2  # 1/sqrt(100) = 0.1 for Monte-Carlo error

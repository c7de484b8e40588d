"""The DeepWalk corpus configuration at its size (R-MAT scale 24, 10 walks per
vertex x 80 vertices, 28.4 GB of paths in HBM): sampled rows bit-exact
against the oracle's Philox replay, unweighted and alias-weighted, the step
total exact, and every sampled consecutive pair an edge."""
import numpy as np
import pytest

import oracle
import ptwalker_amd as ptw

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def big(ctx24):
    c, gi, rp, col = ctx24
    w, _, _ = c.export_weights()
    return c, gi, rp, col, oracle.build_alias(rp, w)


@pytest.mark.parametrize("weighted", [False, True])
def test_deepwalk_corpus_rows_replay(big, weighted):
    ctx, gi, rp, col, alias = big
    print(f"config 3: DeepWalk scale 24 (V {gi.num_vertices}), weighted={weighted}", flush=True)
    V, L = gi.num_vertices, 80
    W = 10 * V
    buf = ctx.alloc(4 * W * L)
    flags = ptw.FLAG_OUT_DEVICE | (ptw.FLAG_WEIGHTED if weighted else 0)
    d = ptw.make_desc(ptw.DeepWalk(L, start=ptw.START_ROUNDROBIN), W, seed=1, flags=flags)
    out = ctx.walk(d, paths=buf.ptr, device_ptrs=True)
    assert out.steps == W * (L - 1) and out.dangling == 0
    n = 3000
    for base in (0, V - 5, W // 2 + 7, W - n):
        got = buf.download(np.uint32, n * L, offset=base * L).reshape(n, L)
        ref = oracle.walk_replay(rp, col, oracle.APP_DEEPWALK, n, walk_len=L, seed=1,
                                 start_mode=oracle.START_ROUNDROBIN, wid_base=base,
                                 alias=alias if weighted else None)
        assert np.array_equal(got, ref["paths"]), base
        assert np.array_equal(got[:, 0], (np.arange(base, base + n) % V).astype(np.uint32))
        # every consecutive pair is an edge of the graph
        a, b = got[:, :-1].ravel().astype(np.int64), got[:, 1:].ravel()
        for u, x in zip(a[:5000], b[:5000]):
            assert x in col[rp[u]:rp[u + 1]]
    buf.free()

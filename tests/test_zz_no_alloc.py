"""ptw_walk_static, the walk that allocates nothing: its outputs equal
ptw_walk's bit for bit, on calls in a row whose second input is smaller than
the first (the static scratch keeps the last call's contents), and it
refuses what its fixed scratch cannot hold."""
import numpy as np
import pytest

import ptwalker_amd as ptw

pytestmark = pytest.mark.gpu


def static_walk(ctx, app, W, seed, stream=0):
    """One ptw_walk_static call through the C ABI into device buffers, and the
    outputs copied back."""
    d = ptw.make_desc(app, W, seed=seed, stream=stream)
    d.flags |= ptw.FLAG_OUT_DEVICE
    V = ctx.info().num_vertices
    bufs = {}
    if d.app in (ptw.APP_PAGERANK, ptw.APP_PPR):
        bufs["counts"] = ctx.alloc(8 * V)
    elif d.app == ptw.APP_RWR:
        bufs["ends"] = ctx.alloc(4 * W)
        bufs["starts"] = ctx.alloc(4 * W)
    else:
        bufs["paths"] = ctx.alloc(4 * W * d.walk_len)
    try:
        out = ctx.walk_static(d, **{k: b.ptr for k, b in bufs.items()})
        got = {"counts": None, "ends": None, "starts": None, "paths": None}
        if "counts" in bufs:
            got["counts"] = bufs["counts"].download(np.uint64, V)
        if "ends" in bufs:
            got["ends"] = bufs["ends"].download(np.uint32, W)
            got["starts"] = bufs["starts"].download(np.uint32, W)
        if "paths" in bufs:
            got["paths"] = bufs["paths"].download(np.uint32, W * d.walk_len).reshape(W, d.walk_len)
    finally:
        for b in bufs.values():
            b.free()
    return out, got


def check_equal(ctx, app, W, seed, stream=0):
    out, got = static_walk(ctx, app, W, seed, stream)
    ref = ptw.walk(ctx, W, app, seed=seed, stream=stream, starts=True)
    assert out.steps == ref.steps and out.dangling == ref.dangling
    for name in ("counts", "ends", "starts", "paths"):
        want = getattr(ref, name)
        if want is not None and (name != "starts" or got["starts"] is not None):
            assert np.array_equal(got[name], want), name
    return out


def test_static_walk_equals_walk_two_calls_smaller_second(ctx):
    gi = ctx.rmat(scale=16, edge_factor=16, seed=1)
    hub = int(gi.max_degree_vertex)
    print(f"ptw_walk_static: scale 16, V={gi.num_vertices}", flush=True)
    check_equal(ctx, ptw.PPR(0.15, hub), 3_000_000, seed=1)
    # the second call has a smaller graph and fewer walkers: counters the first
    # call left beyond the new V, and its queue words, must not leak in
    gi = ctx.rmat(scale=10, edge_factor=8, seed=5)
    print(f"ptw_walk_static: scale 10, V={gi.num_vertices}", flush=True)
    check_equal(ctx, ptw.PPR(0.2, int(gi.max_degree_vertex)), 200_000, seed=2)
    check_equal(ctx, ptw.PageRank(0.15), 100_000, seed=3)
    check_equal(ctx, ptw.RWR(0.2, 7), 50_000, seed=4, stream=3)
    check_equal(ctx, ptw.RWR(0.3), 40_000, seed=4)
    check_equal(ctx, ptw.DeepWalk(9), 20_000, seed=5)
    check_equal(ctx, ptw.PPR(0.2, 3), 0, seed=6)


def test_static_walk_refusals(ctx):
    gi = ctx.rmat(scale=10, edge_factor=8, seed=5)
    d = ptw.make_desc(ptw.PPR(0.2, 0), 1000)  # no FLAG_OUT_DEVICE: host outputs would need staging
    buf = ctx.alloc(8 * gi.num_vertices)
    try:
        with pytest.raises(ptw.PtwError) as e:
            ctx.walk_static(d, counts=buf.ptr)
        assert e.value.code == ptw.ERR_INVALID_ARG and "PTW_FLAG_OUT_DEVICE" in str(e.value)
    finally:
        buf.free()
    d = ptw.make_desc(ptw.Node2Vec(5, 0.5, 2.0), 10)
    d.flags |= ptw.FLAG_OUT_DEVICE
    buf = ctx.alloc(4 * 50)
    try:
        with pytest.raises(ptw.PtwError) as e:
            ctx.walk_static(d, paths=buf.ptr)
        assert e.value.code == ptw.ERR_UNSUPPORTED
    finally:
        buf.free()
    # a star with more vertices than the static counters hold
    n = 4_194_296
    rp = np.empty(n + 2, dtype=np.uint32)
    rp[0], rp[1] = 0, n
    rp[2:] = n + np.arange(1, n + 1, dtype=np.uint32)
    col = np.concatenate([np.arange(1, n + 1, dtype=np.uint32), np.zeros(n, dtype=np.uint32)])
    ctx.from_csr(rp, col)
    d = ptw.make_desc(ptw.PPR(0.2, 1), 1000)
    d.flags |= ptw.FLAG_OUT_DEVICE
    buf = ctx.alloc(8 * (n + 1))
    try:
        with pytest.raises(ptw.PtwError) as e:
            ctx.walk_static(d, counts=buf.ptr)
        assert e.value.code == ptw.ERR_INVALID_ARG and "4194296" in str(e.value)
        # RWR keeps no per-vertex counters: the same graph walks
        ends = ctx.alloc(4 * 1000)
        try:
            d = ptw.make_desc(ptw.RWR(0.2, 1), 1000)
            d.flags |= ptw.FLAG_OUT_DEVICE
            ctx.walk_static(d, ends=ends.ptr)
            got = ends.download(np.uint32, 1000)
            ref = ptw.walk(ctx, 1000, ptw.RWR(0.2, 1))
            assert np.array_equal(got, ref.ends)
        finally:
            ends.free()
    finally:
        buf.free()

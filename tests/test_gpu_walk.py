"""GPU parity: the HIP walk kernels against the oracle's Philox replay, bit for bit
(visit counts, endpoints, paths, step totals), through the C ABI."""
import os
import subprocess

import numpy as np
import pytest

import oracle
import ptwalker_amd as ptw
from ptwalker_amd.workloads import walk_sources
from walkstats import random_graph

pytestmark = pytest.mark.gpu


def replay(rp, col, desc: ptw.WalkDesc, threads=0):
    return oracle.walk_replay(rp, col, desc.app, desc.walkers, tp=desc.tp, seed=desc.seed, source=desc.source,
                              start_mode=desc.start_mode, semantics=desc.semantics, walk_len=desc.walk_len,
                              stream=desc.stream, wid_base=desc.wid_base, threads=threads)


def check(ctx, rp, col, app, W, **kw):
    d = ptw.make_desc(app, W, **kw)
    got = ptw.walk(ctx, W, app, **kw)
    ref = replay(rp, col, d)
    assert got.steps == ref["steps"], (got.steps, ref["steps"])
    if got.counts is not None:
        assert got.visits == W + got.steps
        assert np.array_equal(got.counts, ref["counts"]), np.nonzero(got.counts != ref["counts"])
    if got.ends is not None:
        assert np.array_equal(got.ends, ref["ends"])
    if got.paths is not None:
        assert np.array_equal(got.paths, ref["paths"])
    assert got.dangling == ref["dangling"]
    return got


@pytest.fixture(scope="module")
def rmat12(ctx):
    gi = ctx.rmat(scale=12, edge_factor=16, seed=1)
    rp, col = ctx.export_csr()
    return gi, rp, col


def apps_for(gi):
    hub = int(gi.max_degree_vertex)
    return [
        ptw.PPR(0.15, hub),
        ptw.PPR(0.2, gi.num_vertices - 1),
        ptw.PageRank(0.15),
        ptw.PageRank(0.3, semantics=ptw.SEM_TEXTBOOK),
        ptw.RWR(0.2, hub),
        ptw.RWR(0.5),
        ptw.DeepWalk(20),
        ptw.DeepWalk(7, start=ptw.START_ROUNDROBIN),
    ]


def test_parity_rmat12_all_apps(ctx, rmat12):
    gi, rp, col = rmat12
    for app in apps_for(gi):
        print("case", app, flush=True)
        W = 3000 if isinstance(app, ptw.DeepWalk) else 200000
        check(ctx, rp, col, app, W, seed=1)


def test_parity_seed_stream_and_wid_base(ctx, rmat12):
    gi, rp, col = rmat12
    app = ptw.PPR(0.15, int(gi.max_degree_vertex))
    for seed, stream, base in [(2, 0, 0), (1, 7, 0), (1, 0, (1 << 33) + 5), (0xDEADBEEFCAFE, 3, 123)]:
        check(ctx, rp, col, app, 50000, seed=seed, stream=stream, wid_base=base)


def test_split_runs_add_up(ctx, rmat12):
    gi, rp, col = rmat12
    app = ptw.PageRank(0.15)
    whole = ptw.walk(ctx, 100000, app, seed=4)
    a = ptw.walk(ctx, 37000, app, seed=4, wid_base=0)
    b = ptw.walk(ctx, 63000, app, seed=4, wid_base=37000)
    assert np.array_equal(a.counts + b.counts, whole.counts)
    assert a.steps + b.steps == whole.steps


@pytest.mark.parametrize("relabel", [True, False])
def test_parity_unsorted_adjacency(ctx, relabel):
    rp, col = random_graph(60, 0.08, seed=3, symmetric=False)
    ctx.from_csr(rp, col, flags=0 if relabel else ptw.GRAPH_NO_RELABEL)
    assert ctx.info().relabelled == (1 if relabel else 0)
    for app in [ptw.PPR(0.15, 5), ptw.PageRank(0.25), ptw.RWR(0.2, 9), ptw.DeepWalk(15)]:
        check(ctx, rp, col, app, 20000 if not isinstance(app, ptw.DeepWalk) else 2000, seed=9)


def test_edge_cases(ctx):
    rp, col = random_graph(30, 0.1, seed=5)
    ctx.from_csr(rp, col)
    # tp = 1: every walker takes exactly one step
    g = check(ctx, rp, col, ptw.PPR(1.0, 2), 10000, seed=3)
    assert g.steps == 10000
    # a single walker, and zero walkers
    check(ctx, rp, col, ptw.RWR(0.3, 4), 1, seed=3)
    z = ptw.walk(ctx, 0, ptw.PPR(0.2, 1), seed=3)
    assert z.steps == 0 and int(z.counts.sum()) == 0
    # path of one vertex: no steps
    g = check(ctx, rp, col, ptw.DeepWalk(1), 100, seed=3)
    assert g.steps == 0
    # K2
    ctx.from_csr(np.array([0, 1, 2], dtype=np.uint32), np.array([1, 0], dtype=np.uint32))
    check(ctx, np.array([0, 1, 2]), np.array([1, 0]), ptw.PPR(0.15, 0), 100000, seed=8)


def test_invalid_arguments(ctx):
    rp, col = random_graph(10, 0.3, seed=1)
    ctx.from_csr(rp, col)
    with pytest.raises(ptw.PtwError) as e:
        ptw.walk(ctx, 10, ptw.PPR(0.2, 10))
    assert e.value.code == ptw.ERR_INVALID_ARG
    for tp in (0.0, -0.1, 1.5, float("nan")):
        with pytest.raises(ptw.PtwError) as e:
            ptw.walk(ctx, 10, ptw.PageRank(tp))
        assert e.value.code == ptw.ERR_INVALID_ARG
    with pytest.raises(ptw.PtwError) as e:
        ptw.walk(ctx, 10, ptw.DeepWalk(0))
    assert e.value.code == ptw.ERR_INVALID_ARG


def test_dangling_policies(ctx):
    rp, col = random_graph(20, 0.2, seed=2, dangling=4)
    ctx.from_csr(rp, col)
    assert ctx.info().num_dangling == 4
    with pytest.raises(ptw.PtwError) as e:
        ptw.walk(ctx, 5000, ptw.PageRank(0.15))
    assert e.value.code == ptw.ERR_DANGLING
    for app in [ptw.PageRank(0.15), ptw.RWR(0.1, 0), ptw.DeepWalk(25)]:
        g = check(ctx, rp, col, app, 5000, seed=2, dangling=ptw.DANGLING_STOP)
        assert g.dangling > 0


def test_lds_and_global_counters_scale16(ctx):
    # V > the LDS hub capacity, so both counter paths are exercised
    gi = ctx.rmat(scale=16, edge_factor=16, seed=1)
    assert gi.num_vertices > 8192
    rp, col = ctx.export_csr()
    check(ctx, rp, col, ptw.PPR(0.15, int(gi.max_degree_vertex)), 2_000_000, seed=1)
    check(ctx, rp, col, ptw.PageRank(0.15), 1_000_000, seed=1)
    check(ctx, rp, col, ptw.RWR(0.2, int(gi.max_degree_vertex)), 1_000_000, seed=1)


@pytest.mark.parametrize("hubs,slots", [(0, 0), (256, 0), (0, 4096), (255, 1024), (8192, 4096)])
def test_counter_paths_explicit_shapes(ctx, hubs, slots):
    """Every counting path, pinned by an explicit launch shape (blocks_per_cu set,
    so the small-walk rule does not apply): direct LDS hub counters, the LDS
    visit cache, both, and device atomics only."""
    gi = ctx.rmat(scale=14, edge_factor=16, seed=2)
    rp, col = ctx.export_csr()
    default = ctx.tuning().as_dict()
    try:
        ctx.set_tuning(lds_hubs=hubs, cache_slots=slots, blocks_per_cu=1, block=1024)
        check(ctx, rp, col, ptw.PPR(0.15, int(gi.max_degree_vertex)), 600_000, seed=3)
        check(ctx, rp, col, ptw.PPR(0.2, 777), 300_000, seed=4)
        check(ctx, rp, col, ptw.PageRank(0.15), 300_000, seed=5)
    finally:
        ctx.set_tuning(**default)


# walker counts large enough that the slice length is the chunk asked for
# (capped at 256) rather than the small-walk share
@pytest.mark.parametrize("chunk,block,tp,W", [(64, 64, 0.02, 300_000), (64, 1024, 0.2, 300_000),
                                              (128, 256, 0.05, 600_000), (256, 256, 0.2, 1_200_000),
                                              (4096, 256, 0.01, 1_200_000)])
def test_endpoint_staging_shapes(ctx, chunk, block, tp, W):
    """RWR ends go through each wavefront's LDS slice slots (at most 256 ids
    each, written out when a slice's walkers have all ended).  Short slices
    with long walks (tp 0.02 at 64 ids) make many walkers outlive their slot,
    so the slot's ended walkers are written early and the rest write their
    own: ends and starts must still equal the replay."""
    gi = ctx.rmat(scale=14, edge_factor=16, seed=2)
    rp, col = ctx.export_csr()
    default = ctx.tuning().as_dict()
    try:
        ctx.set_tuning(chunk=chunk, block=block, blocks_per_cu=1)
        for app, n in ((ptw.RWR(tp, int(gi.max_degree_vertex)), W), (ptw.RWR(tp), W + 1)):
            got = check(ctx, rp, col, app, n, seed=7)
            assert 64 <= got.shape["chunk"] <= 256
            if (chunk, block) in ((256, 256), (4096, 256)):
                assert got.shape["chunk"] == 256
    finally:
        ctx.set_tuning(**default)


def star(n):
    """Vertex 0 joined to leaves 1..n, both directions."""
    rp = np.zeros(n + 2, dtype=np.uint32)
    rp[1] = n
    rp[2:] = n + np.arange(1, n + 1)
    col = np.concatenate([np.arange(1, n + 1), np.zeros(n)]).astype(np.uint32)
    return rp, col


def test_hub_counter_spill(ctx):
    """The centre of a star takes every other visit.  From a leaf, 8e6 walkers at
    tp 0.2 make about 2e7 centre visits, far above 2^15 per workgroup, so every
    workgroup's u16 LDS counter of the centre spills to the device counter."""
    rp, col = star(64)
    ctx.from_csr(rp, col)
    assert ctx.info().hub_vertices == 65
    check(ctx, rp, col, ptw.PPR(0.2, 1), 8_000_000, seed=3)
    check(ctx, rp, col, ptw.PageRank(0.2), 4_000_000, seed=4)


def test_automatic_shape_rmat17(ctx):
    """The default (automatic) launch shape on a graph with more vertices than
    any LDS hub range (R-MAT scale 17: 90,202 vertices), so device ids above the
    hub counters reach the visit cache, the visit log or device atomics.  With
    the log (the default) every walk gives the cache's LDS to hub counters
    (79,872, no cache).  Without it, a source of degree above 2048 and uniform
    starts do the same, and a degree-1 source keeps 77,824 hubs and 512 cache
    slots.  3e6 walkers keep the launch off the small-walk shape.  Counts equal
    the replay."""
    gi = ctx.rmat(scale=17, edge_factor=16, seed=1)
    assert gi.num_vertices > 81408
    rp, col = ctx.export_csr()
    deg = np.diff(rp.astype(np.int64))
    low = int(np.argmin(deg))
    assert ctx.tuning().cache_slots == ptw.CACHE_AUTO and ctx.tuning().blocks_per_cu == 0
    assert ctx.tuning().visit_log == 1
    default = ctx.tuning().as_dict()
    cases = [(ptw.PPR(0.2, int(gi.max_degree_vertex)), 7, (79872, 0), (79872, 0)),
             (ptw.PPR(0.2, low), 8, (79872, 0), (77824, 512)),
             (ptw.PageRank(0.15), 9, (79872, 0), (79872, 0))]
    try:
        for log in (1, 0):
            ctx.set_tuning(visit_log=log)
            for app, seed, with_log, without in cases:
                got = check(ctx, rp, col, app, 3_000_000, seed=seed)
                want = with_log if log else without
                assert (got.shape["lds_hubs"], got.shape["cache_slots"]) == want, (log, got.shape)
                assert (got.shape["visit_log"] > 0) == bool(log), got.shape
        # an explicit setting is used as given
        ctx.set_tuning(cache_slots=1024, lds_hubs=36864, visit_log=1)
        got = check(ctx, rp, col, ptw.PPR(0.2, int(gi.max_degree_vertex)), 3_000_000, seed=10)
        assert (got.shape["lds_hubs"], got.shape["cache_slots"]) == (36864, 1024), got.shape
    finally:
        ctx.set_tuning(**default)


def test_visit_log_paths(ctx):
    """Visits that no LDS counter takes go to the visit log and are counted by
    id bucket after the walk; with visit_log = 0 they take device atomics, and
    a log smaller than the walk's visits (64 Ki entries here) fills up and
    sends the rest to device atomics.  R-MAT scale 18 (more vertices than the
    LDS hub range), 4e6 walkers: counts equal the replay on every path."""
    gi = ctx.rmat(scale=18, edge_factor=16, seed=3)
    assert gi.num_vertices > 81408
    rp, col = ctx.export_csr()
    deg = np.diff(rp.astype(np.int64))
    low = int(np.argmin(deg))
    cases = [ptw.PPR(0.2, int(gi.max_degree_vertex)), ptw.PPR(0.15, low), ptw.PageRank(0.15)]
    default = ctx.tuning().as_dict()
    try:
        for setting in (1, 0, 65536):
            ctx.set_tuning(visit_log=setting)
            for i, app in enumerate(cases):
                print("visit_log", setting, app, flush=True)
                got = check(ctx, rp, col, app, 4_000_000, seed=11 + i)
                logged = got.shape["visit_log"]
                if setting == 0:
                    assert logged == 0
                elif setting == 65536:
                    assert logged == 65536, got.shape
                else:
                    assert 0 < logged < got.visits, got.shape
    finally:
        ctx.set_tuning(**default)
    with pytest.raises(ptw.PtwError):
        ctx.set_tuning(visit_log=1000)


@pytest.mark.parametrize("relabel", [True, False])
def test_edge_records_every_kind(ctx, relabel):
    """Neighbour records (a slot's neighbour with its row and graph id) make a
    step one dependent read; edge_records = 1 (default) puts every walk kind
    but node2vec on them, 0 none, -1 DeepWalk only.  Weighted path walks read the pair of
    alias candidates; new weights rebuild those.  Every result equals the
    replay, with and without the degree relabel."""
    gi = ctx.rmat(scale=14, edge_factor=16, seed=5, flags=0 if relabel else ptw.GRAPH_NO_RELABEL)
    rp, col = ctx.export_csr()
    hub = int(gi.max_degree_vertex)
    low = int(np.argmin(np.diff(rp.astype(np.int64))))
    apps = [(ptw.PPR(0.15, hub), 600_000), (ptw.PPR(0.2, low), 300_000), (ptw.PageRank(0.15), 300_000),
            (ptw.RWR(0.2, hub), 300_000), (ptw.RWR(0.3), 300_000), (ptw.DeepWalk(20), 20_000),
            (ptw.DeepWalk(9, start=ptw.START_ROUNDROBIN), 5_000)]
    default = ctx.tuning().as_dict()
    try:
        for setting in (1, 0, -1):
            ctx.set_tuning(edge_records=setting)
            for i, (app, W) in enumerate(apps):
                print("edge_records", setting, app, flush=True)
                check(ctx, rp, col, app, W, seed=20 + i)
        ctx.set_tuning(edge_records=1)
        for wseed in (5, 6):
            ctx.random_weights(seed=wseed)
            w, _, _ = ctx.export_weights()
            alias = oracle.build_alias(rp, w)
            for app, W in [(ptw.DeepWalk(16), 5000), (ptw.PPR(0.2, hub), 200_000)]:
                d = ptw.make_desc(app, W, seed=wseed, flags=ptw.FLAG_WEIGHTED)
                paths = np.zeros(W * d.walk_len, dtype=np.uint32) if d.app == ptw.APP_DEEPWALK else None
                counts = np.zeros(gi.num_vertices, dtype=np.uint64) if paths is None else None
                out = ctx.walk(d, counts=counts, paths=paths)
                ref = oracle.walk_replay(rp, col, d.app, W, tp=d.tp, seed=wseed, source=d.source,
                                         start_mode=d.start_mode, walk_len=d.walk_len, alias=alias)
                assert out.steps == ref["steps"]
                if paths is not None:
                    assert np.array_equal(paths.reshape(W, d.walk_len), ref["paths"])
                else:
                    assert np.array_equal(counts, ref["counts"])
    finally:
        ctx.set_tuning(**default)
    with pytest.raises(ptw.PtwError):
        ctx.set_tuning(edge_records=2)


def test_device_outputs_and_accumulate(ctx):
    gi = ctx.rmat(scale=12, edge_factor=8, seed=3)
    rp, col = ctx.export_csr()
    V = gi.num_vertices
    W = 100000
    counts = ctx.alloc(8 * V).zero()
    d = ptw.make_desc(ptw.PPR(0.2, 3), W, seed=5, flags=ptw.FLAG_OUT_DEVICE)
    ctx.walk(d, counts=counts.ptr, device_ptrs=True)
    d2 = ptw.make_desc(ptw.PPR(0.2, 3), W, seed=5, wid_base=W, flags=ptw.FLAG_OUT_DEVICE | ptw.FLAG_ACCUMULATE)
    ctx.walk(d2, counts=counts.ptr, device_ptrs=True)
    ref = oracle.walk_replay(rp, col, oracle.APP_PPR, 2 * W, tp=0.2, seed=5, source=3,
                             start_mode=oracle.START_SOURCE)
    assert np.array_equal(counts.download(np.uint64), ref["counts"])
    ends = ctx.alloc(4 * W)
    d3 = ptw.make_desc(ptw.RWR(0.2, 3), W, seed=5, flags=ptw.FLAG_OUT_DEVICE)
    ctx.walk(d3, ends=ends.ptr, device_ptrs=True)
    ref = oracle.walk_replay(rp, col, oracle.APP_RWR, W, tp=0.2, seed=5, source=3, start_mode=oracle.START_SOURCE)
    assert np.array_equal(ends.download(np.uint32), ref["ends"])
    free, total = ctx.device_memory()
    assert 0 < free <= total
    counts.free()
    ends.free()


def test_full_size_properties_scale20(ctx):
    """Size-independent properties at a large size: sum of counts, mean length,
    split invariance, run-to-run determinism."""
    gi = ctx.rmat(scale=20, edge_factor=16, seed=1)
    src = int(gi.max_degree_vertex)
    W, tp = 20_000_000, 0.2
    a = ptw.walk(ctx, W, ptw.PPR(tp, src), seed=1)
    assert int(a.counts.sum()) == W + a.steps
    assert a.counts[src] >= W
    sd = np.sqrt(W * (1 - tp)) / tp
    assert abs(a.steps - W / tp) < 6 * sd
    b = ptw.walk(ctx, W, ptw.PPR(tp, src), seed=1)
    assert np.array_equal(a.counts, b.counts)
    h1 = ptw.walk(ctx, W // 2, ptw.PPR(tp, src), seed=1)
    h2 = ptw.walk(ctx, W // 2, ptw.PPR(tp, src), seed=1, wid_base=W // 2)
    assert np.array_equal(h1.counts + h2.counts, a.counts)


def test_fp64_l1_scale14(ctx):
    """Monte-Carlo estimate against the FP64 power iteration (L1 within a multiple of its
    expected size under the multinomial-like noise)."""
    gi = ctx.rmat(scale=14, edge_factor=16, seed=2)
    rp, col = ctx.export_csr()
    src = int(gi.max_degree_vertex)
    tp, W = 0.2, 5_000_000
    r = ptw.walk(ctx, W, ptw.PPR(tp, src), seed=1)
    est = r.counts / W
    exact = oracle.power_iteration(rp, col, tp=tp, start=src, mode="visits")
    l1 = np.abs(est - exact).sum() / exact.sum()
    # per-vertex sd <= sqrt(E[visits^2]/W); sum of sds bounds E[L1] up to a constant
    bound = np.sqrt((2 / tp ** 2 + 3 / tp + 1) * exact / exact.sum() / W).sum()
    assert l1 < 3 * bound, (l1, bound)


def test_cli_pagerank_end_to_end(ctx, tmp_path):
    ctx.rmat(scale=10, edge_factor=8, seed=4)
    g = str(tmp_path / "g.txt")
    ctx.write(g, ptw.FORMAT_TEXT)
    rp, col, name2id = oracle.load_edge_list(g)
    names = names_of(name2id)
    out = str(tmp_path / "counts.txt")
    exe = os.path.join(ptw.BIN_DIR, "pagerank")
    r = subprocess.run([exe, "-f", "text", "-g", g, "-w", "65536", "--sp", "0.15", "-s", "1", "-t", "20",
                        "--seed", "1", "--out", out], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.splitlines()
    assert lines[0].startswith("Partition time ")
    assert any(l.startswith("start to walk ") for l in lines)
    assert any(l.startswith("total time (including init start walker) ") for l in lines)
    got = np.loadtxt(out, dtype=np.uint64)
    ref = oracle.walk_replay(rp, col, oracle.APP_PAGERANK, 65536, tp=0.15, seed=1)
    assert np.array_equal(got[:, 0], names)  # the file's own ids, in load order
    assert np.array_equal(got[:, 1], ref["counts"])
    assert lines[-1] == f"ppr value {int(ref['counts'][0]) & 0xFFFFFFFF}"
    steps = int(lines[-2].rsplit(" ", 1)[1])
    assert steps == ref["steps"]


def names_of(name2id):
    """The inverse of the loader's first-appearance map: file id of each graph id."""
    names = np.zeros(len(name2id), dtype=np.uint32)
    for name, v in name2id.items():
        names[v] = name
    return names


def test_cli_ppr_rwr_deepwalk(ctx, tmp_path):
    ctx.rmat(scale=9, edge_factor=8, seed=6)
    g = str(tmp_path / "g.bin")
    ctx.write(g, ptw.FORMAT_BINARY)
    rp, col, name2id = oracle.load_edge_list(g, "binary")
    names = names_of(name2id)
    deg = np.diff(rp.astype(np.int64))
    hub = int(np.argmax(deg))
    b = ptw.BIN_DIR
    r = subprocess.run([f"{b}/PPR", "-f", "binary", "-g", g, "-w", "5000", "--sp", "0.2", "--out",
                        str(tmp_path / "p.txt")], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert f"Root vertex {names[hub]}" in r.stdout
    ref = oracle.walk_replay(rp, col, oracle.APP_PPR, 5000, tp=0.2, seed=1, source=hub, start_mode=oracle.START_SOURCE)
    got = np.loadtxt(tmp_path / "p.txt", dtype=np.uint64)
    assert np.array_equal(got[:, 0], names) and np.array_equal(got[:, 1], ref["counts"])
    r = subprocess.run([f"{b}/RWR", "-f", "binary", "-g", g, "-w", "5000", "-l", "10", "--out",
                        str(tmp_path / "r.txt")], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    ref = oracle.walk_replay(rp, col, oracle.APP_RWR, 5000, tp=0.5, seed=1)
    # a walker's start is the first vertex of a one-vertex DeepWalk path under the same start rule
    starts = oracle.walk_replay(rp, col, oracle.APP_DEEPWALK, 5000, walk_len=1, seed=1)["paths"][:, 0]
    pairs = np.loadtxt(tmp_path / "r.txt", dtype=np.uint32)
    assert np.array_equal(pairs[:, 0], names[starts]) and np.array_equal(pairs[:, 1], names[ref["ends"]])
    r = subprocess.run([f"{b}/deepwalk", "-f", "binary", "-g", g, "-e", "2", "-l", "8", "--out",
                        str(tmp_path / "d.bin")], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    V = len(rp) - 1
    ref = oracle.walk_replay(rp, col, oracle.APP_DEEPWALK, 2 * V, walk_len=8, seed=1)
    got = np.fromfile(tmp_path / "d.bin", dtype=np.uint32).reshape(2 * V, 8)
    assert np.array_equal(got, names[ref["paths"]])
    assert r.stdout.splitlines()[-1] == "ppr value 0"
    # graph ids on request
    r = subprocess.run([f"{b}/deepwalk", "-f", "binary", "-g", g, "-e", "2", "-l", "8", "--out-ids", "graph",
                        "--out", str(tmp_path / "d.bin")], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert np.array_equal(np.fromfile(tmp_path / "d.bin", dtype=np.uint32).reshape(2 * V, 8), ref["paths"])


def sparse_id_file(path, n=400, m=5000, seed=3):
    """A text edge list whose ids are sparse (up to 2^32 - 2) and appear out of order."""
    rng = np.random.default_rng(seed)
    ids = rng.choice(np.arange(1, 1 << 32, 9973, dtype=np.uint64), n, replace=False)
    e = rng.integers(0, n, size=(m, 2))
    e = e[e[:, 0] != e[:, 1]]
    path.write_text("# sparse ids\n" + "\n".join(f"{ids[a]} {ids[b]}" for a, b in e) + "\n")
    return ids


def test_cli_original_ids_sparse_file(ctx, tmp_path):
    """Every output of every binary in the file's own ids, on a file whose ids are
    sparse and out of order; --source takes a file id too.  RWR writes the
    reference's {start, end} pairs (src/core/RWR.cpp:9-13,99-105)."""
    g = tmp_path / "g.txt"
    sparse_id_file(g)
    rp, col, name2id = oracle.load_edge_list(str(g), "text", True)
    names = names_of(name2id)
    assert not np.array_equal(names, np.arange(len(names)))
    src = 17
    b = ptw.BIN_DIR
    common = ["-f", "text", "-g", str(g), "--undirected", "--seed", "3"]
    r = subprocess.run([f"{b}/PPR", *common, "-w", "20000", "--sp", "0.2", "--source", str(names[src]),
                        "--out", str(tmp_path / "p.txt")], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert f"Root vertex {names[src]}" in r.stdout
    ref = oracle.walk_replay(rp, col, oracle.APP_PPR, 20000, tp=0.2, seed=3, source=src,
                             start_mode=oracle.START_SOURCE)
    got = np.loadtxt(tmp_path / "p.txt", dtype=np.uint64)
    assert np.array_equal(got[:, 0], names) and np.array_equal(got[:, 1], ref["counts"])
    r = subprocess.run([f"{b}/RWR", *common, "-w", "20000", "-l", "5", "--sp", "0.3", "--source", str(names[src]),
                        "--out", str(tmp_path / "r.txt")], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    ref = oracle.walk_replay(rp, col, oracle.APP_RWR, 20000, tp=0.3, seed=3, source=src,
                             start_mode=oracle.START_SOURCE)
    pairs = np.loadtxt(tmp_path / "r.txt", dtype=np.uint64)
    assert np.all(pairs[:, 0] == names[src]) and np.array_equal(pairs[:, 1], names[ref["ends"]])
    # an id that is not in the file is an error, not a silent vertex
    r = subprocess.run([f"{b}/PPR", *common, "-w", "100", "--sp", "0.2", "--source", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode != 0 and "not a vertex" in r.stderr
    ctx.load(str(g), ptw.FORMAT_TEXT, as_undirected=True)
    assert np.array_equal(ctx.export_names(), names)


def test_cli_batch_queries_topk(ctx, tmp_path):
    """PPR and RWR batch mode: a query file in the file's own ids, one launch,
    every query's top-k written in those ids, equal to the top-k of the
    oracle's replay of that query alone (stream = its line)."""
    g = tmp_path / "g.txt"
    sparse_id_file(g, seed=5)
    rp, col, name2id = oracle.load_edge_list(str(g), "text", True)
    names = names_of(name2id)
    qsrc = [0, 5, 11, 3]
    spec = [(0, 30000, 0.2), (5, None, None), (11, 400, 0.3), (3, 150000, None)]
    lines = ["# source walkers tp"] + [" ".join(str(x) for x in (names[s], w, t) if x is not None)
                                        for s, w, t in spec]
    (tmp_path / "q.txt").write_text("\n".join(lines) + "\n")
    b = ptw.BIN_DIR
    for exe, app in (("PPR", oracle.APP_PPR), ("RWR", oracle.APP_RWR)):
        extra = ["-l", "5"] if exe == "RWR" else []
        r = subprocess.run([f"{b}/{exe}", "-f", "text", "-g", str(g), "--undirected", "-w", "7000", "--sp", "0.25",
                            *extra, "--queries", str(tmp_path / "q.txt"), "--topk", "12", "--seed", "4",
                            "--out", str(tmp_path / "t.txt")], capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        rows = np.loadtxt(tmp_path / "t.txt", dtype=np.uint64, comments="#")
        for q, (s, w, t) in enumerate(spec):
            w = 7000 if w is None else w
            t = 0.25 if t is None else t
            ref = oracle.walk_replay(rp, col, app, w, tp=t, seed=4, source=s, start_mode=oracle.START_SOURCE,
                                     stream=q)
            cnt = ref["counts"] if app == oracle.APP_PPR else np.bincount(ref["ends"], minlength=len(names))
            v, c = oracle.topk(cnt, rp, 12)
            mine = rows[rows[:, 0] == q]
            keep = v != oracle.INVALID
            assert np.all(mine[:, 1] == names[s]) and np.all(mine[:, 2] == (app == oracle.APP_RWR))
            assert np.array_equal(mine[:, 3], np.arange(keep.sum()))
            assert np.array_equal(mine[:, 4], names[v[keep]]) and np.array_equal(mine[:, 5], c[keep]), (exe, q)


def test_gather_bench_runs(ctx):
    gps, ms = ctx.bench_gather(1 << 28, 1 << 26)
    assert gps > 1e9 and ms > 0


@pytest.mark.parametrize("app_name", ["PPR", "RWR"])
def test_config_scale22_1e8_walkers_full_replay(ctx22, app_name):
    """The single-source configuration at its size: R-MAT scale 22, sources the
    max-degree vertex and one uniform vertex, 1e8 walkers each, tp 0.2, Philox
    seed 1 -- every one of the 5e8 steps per run replayed on the host threads,
    counts (PPR) and endpoints (RWR) bit-exact (src/core/ppr.cpp:121-139,
    src/core/RWR.cpp:109-128)."""
    c, gi, rp, col = ctx22
    for src in walk_sources(gi, seed=1):
        app = ptw.PPR(0.2, src) if app_name == "PPR" else ptw.RWR(0.2, src)
        print(f"config 2: {app_name} source {src} (degree {rp[src + 1] - rp[src]}), 1e8 walkers", flush=True)
        got = check(c, rp, col, app, 100_000_000, seed=1)
        assert abs(got.steps / 1e8 - 5.0) < 0.01  # E[L] = 1/tp
        if app_name == "PPR":
            assert int(got.counts.sum()) == 100_000_000 + got.steps
            assert got.shape["visit_log"] > 0, got.shape  # the non-hub visits went through the log

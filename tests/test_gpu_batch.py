"""Batched PPR queries in one persistent launch: every query's full count
vector bit-exact against the oracle's replay of that query alone (its own
Philox stream), top-k exact under the stated tie order, the same results
alone and in any mix, and the captured hipGraph replayed several times with
eager batches in between, every replay equal to the eager result."""
import numpy as np
import pytest

import oracle
import ptwalker_amd as ptw
from ptwalker_amd.workloads import batch_queries
from walkstats import random_graph

pytestmark = pytest.mark.gpu


def replay_query(rp, col, src, W, tp, stream, seed=1, alias=None):
    return oracle.walk_replay(rp, col, oracle.APP_PPR, W, tp=tp, seed=seed, source=src,
                              start_mode=oracle.START_SOURCE, stream=stream, alias=alias)


def make_queries(gi, n, seed, wmin=100, wmax=200000):
    rng = np.random.default_rng(seed)
    src = rng.integers(0, gi.num_vertices, n)
    src[0] = gi.max_degree_vertex
    W = np.exp(rng.uniform(np.log(wmin), np.log(wmax), n)).astype(np.int64)
    tp = rng.uniform(0.15, 0.3, n)
    return src, W, tp


@pytest.fixture(scope="module")
def g12(ctx):
    gi = ctx.rmat(scale=12, edge_factor=16, seed=2)
    rp, col = ctx.export_csr()
    return gi, rp, col


def test_batch_counts_and_topk_match_replay(ctx, g12):
    gi, rp, col = g12
    src, W, tp = make_queries(gi, 40, seed=1)
    k = 50
    r = ctx.batch_ppr(src, W, tp, topk=k, seed=3, dense_threshold=20000)
    assert r.dense_queries == int(np.sum(W > 20000)) and r.hash_queries == 40 - r.dense_queries
    total = 0
    for i in range(40):
        ref = replay_query(rp, col, int(src[i]), int(W[i]), float(tp[i]), i, seed=3)
        got = ctx.batch_counts(i)
        assert np.array_equal(got, ref["counts"]), i
        assert r.steps[i] == ref["steps"]
        v, c = oracle.topk(ref["counts"], rp, k)
        assert np.array_equal(r.topk_vertex[i], v), i
        assert np.array_equal(r.topk_count[i], c), i
        total += ref["steps"]
    assert r.total_steps == total


def test_batch_alone_equals_in_mix(ctx, g12):
    gi, rp, col = g12
    src, W, tp = make_queries(gi, 12, seed=5)
    mix = ctx.batch_ppr(src, W, tp, topk=20, seed=9)
    counts_mix = [ctx.batch_counts(i) for i in range(12)]
    for j in (0, 5, 11):
        alone = ctx.batch_ppr([src[j]], [W[j]], [tp[j]], streams=[j], topk=20, seed=9)
        assert np.array_equal(ctx.batch_counts(0), counts_mix[j])
        assert np.array_equal(alone.topk_vertex[0], mix.topk_vertex[j])
        # and the single-walk entry point with the same stream
        d = ptw.make_desc(ptw.PPR(float(tp[j]), int(src[j])), int(W[j]), seed=9, stream=j)
        single = np.zeros(gi.num_vertices, dtype=np.uint64)
        ctx.walk(d, counts=single)
        assert np.array_equal(single, counts_mix[j])


def test_batch_graph_replays_match_eager(ctx, g12):
    gi, rp, col = g12
    a = make_queries(gi, 30, seed=7)
    b = make_queries(gi, 17, seed=8)
    eager_a = ctx.batch_ppr(*a, topk=40, seed=4)
    ca = [ctx.batch_counts(i) for i in (0, 13, 29)]
    eager_b = ctx.batch_ppr(*b, topk=40, seed=4)
    for rep in range(3):
        ga = ctx.batch_ppr(*a, topk=40, seed=4, graph=True)
        assert np.array_equal(ga.topk_vertex, eager_a.topk_vertex), rep
        assert np.array_equal(ga.topk_count, eager_a.topk_count), rep
        assert np.array_equal(ga.steps, eager_a.steps), rep
        for i, c in zip((0, 13, 29), ca):
            assert np.array_equal(ctx.batch_counts(i), c)
        # eager work between replays
        e = ctx.batch_ppr(*b, topk=40, seed=4)
        assert np.array_equal(e.topk_vertex, eager_b.topk_vertex)
        gb = ctx.batch_ppr(*b, topk=40, seed=4, graph=True)
        assert np.array_equal(gb.topk_count, eager_b.topk_count), rep


def test_batch_gather_and_entropy(ctx, g12):
    gi, rp, col = g12
    src, W, tp = make_queries(gi, 10, seed=3, wmax=300000)
    r = ctx.batch_ppr(src, W, tp, topk=0, seed=2, dense_threshold=50000)
    rng = np.random.default_rng(0)
    qi = rng.integers(0, 10, 5000)
    vs = rng.integers(0, gi.num_vertices, 5000)
    got = ctx.batch_gather(qi, vs)
    refs = [replay_query(rp, col, int(src[i]), int(W[i]), float(tp[i]), i, seed=2)["counts"] for i in range(10)]
    assert np.array_equal(got, np.array([refs[q][v] for q, v in zip(qi, vs)], dtype=np.uint64))
    for i in range(10):
        c = refs[i][refs[i] > 0].astype(np.float64)
        p = c / c.sum()
        assert abs(r.entropy[i] - float(-(p * np.log(p)).sum())) < 1e-6, i
    again = ctx.batch_ppr(src, W, tp, topk=0, seed=2, dense_threshold=50000)
    assert np.array_equal(again.entropy, r.entropy)  # order-independent sums


def test_batch_weighted(ctx):
    gi = ctx.rmat(scale=11, edge_factor=16, seed=4)
    ctx.random_weights(seed=3)
    rp, col = ctx.export_csr()
    w, _, _ = ctx.export_weights()
    alias = oracle.build_alias(rp, w)
    src, W, tp = make_queries(gi, 10, seed=2, wmax=50000)
    r = ctx.batch_ppr(src, W, tp, topk=10, seed=6, weighted=True, dense_threshold=5000)
    for i in range(10):
        ref = replay_query(rp, col, int(src[i]), int(W[i]), float(tp[i]), i, seed=6, alias=alias)
        assert np.array_equal(ctx.batch_counts(i), ref["counts"]), i


@pytest.fixture
def visit_log(ctx, request):
    """Runs the test with dense queries logged (1) or in dense slices (0)."""
    t = ctx.tuning()
    ctx.set_tuning(visit_log=request.param)
    yield request.param
    ctx.set_tuning(visit_log=t.visit_log)


@pytest.fixture(scope="module")
def g16(ctx):
    gi = ctx.rmat(scale=16, edge_factor=16, seed=5)
    rp, col = ctx.export_csr()
    return gi, rp, col


@pytest.mark.parametrize("hubs", [512, 77824])
def test_logged_and_slice_batches_agree(ctx, g16, hubs):
    # the same mixed batch (visit and endpoint queries, dense and hash) with
    # dense queries logged and with dense slices: every output equal, and the
    # logged run equal to the replay.  512 hub counters leave most visits to
    # the log (several id buckets on this graph); 77824 cover every vertex.
    gi, rp, col = g16
    ctx.from_csr(rp, col)
    n = 24
    src, W, tp = make_queries(gi, n, seed=21, wmin=1000, wmax=100000)
    kinds = np.where(np.arange(n) % 3 == 2, ptw.QUERY_ENDPOINTS, ptw.QUERY_VISITS)
    t = ctx.tuning()
    res = {}
    try:
        for vl in (1, 0):
            ctx.set_tuning(visit_log=vl, lds_hubs=hubs)
            r = ctx.batch_ppr(src, W, tp, topk=30, seed=8, kinds=kinds, dense_threshold=4096)
            rng = np.random.default_rng(1)
            qi = rng.integers(0, n, 4000)
            vs = rng.integers(0, gi.num_vertices, 4000)
            res[vl] = (r, [ctx.batch_counts(i) for i in range(n)], ctx.batch_gather(qi, vs), qi, vs)
    finally:
        ctx.set_tuning(visit_log=t.visit_log, lds_hubs=t.lds_hubs)
    (a, ca, ga, qi, vs), (b, cb, gb, _, _) = res[1], res[0]
    n_dense = int(np.sum(W > 4096))
    # a dense query is logged while its expected counts stay within 1 per vertex
    counts = W * np.where(kinds == ptw.QUERY_ENDPOINTS, 1.0, 1.0 + 1.0 / tp)
    n_logged = int(np.sum((W > 4096) & (counts <= 1.0 * gi.num_vertices)))
    assert n_logged >= 2 and n_dense - n_logged >= 2  # logged and sliced dense queries in one launch
    assert a.logged_queries == n_logged and a.dense_queries == n_dense and not a.log_rerun
    assert b.logged_queries == 0 and b.dense_queries == n_dense
    for i in range(n):
        assert np.array_equal(ca[i], cb[i]), i
    assert np.array_equal(a.topk_vertex, b.topk_vertex) and np.array_equal(a.topk_count, b.topk_count)
    assert np.array_equal(a.steps, b.steps) and np.array_equal(a.entropy, b.entropy)
    assert np.array_equal(ga, gb)
    assert np.array_equal(ga, np.array([ca[q][v] for q, v in zip(qi, vs)], dtype=np.uint64))
    for i in np.flatnonzero(W > 4096):
        if kinds[i] == ptw.QUERY_ENDPOINTS:
            ref = replay_endpoints(rp, col, int(src[i]), int(W[i]), float(tp[i]), int(i), seed=8)[0]
        else:
            ref = replay_query(rp, col, int(src[i]), int(W[i]), float(tp[i]), int(i), seed=8)["counts"]
        assert np.array_equal(ca[i], ref), i


def test_logged_batch_graph_replays_and_weighted(ctx, g16):
    # logged queries (and sliced ones) in a captured hipGraph replayed across
    # batches with eager work in between, every replay equal to the eager
    # result; then a weighted batch with logged queries against the replay
    gi, rp, col = g16
    ctx.from_csr(rp, col)
    t = ctx.tuning()
    try:
        ctx.set_tuning(lds_hubs=512)
        a = make_queries(gi, 20, seed=31, wmin=3000, wmax=40000)
        b = make_queries(gi, 9, seed=32, wmin=3000, wmax=40000)
        eager_a = ctx.batch_ppr(*a, topk=40, seed=4, dense_threshold=2048)
        ca = [ctx.batch_counts(i) for i in (0, 7, 19)]
        eager_b = ctx.batch_ppr(*b, topk=40, seed=4, dense_threshold=2048)
        assert eager_a.logged_queries >= 4 and eager_a.dense_queries > eager_a.logged_queries
        for rep in range(3):
            ga = ctx.batch_ppr(*a, topk=40, seed=4, dense_threshold=2048, graph=True)
            assert ga.logged_queries == eager_a.logged_queries
            assert np.array_equal(ga.topk_vertex, eager_a.topk_vertex), rep
            assert np.array_equal(ga.topk_count, eager_a.topk_count), rep
            assert np.array_equal(ga.steps, eager_a.steps) and np.array_equal(ga.entropy, eager_a.entropy), rep
            for i, c in zip((0, 7, 19), ca):
                assert np.array_equal(ctx.batch_counts(i), c), (rep, i)
            e = ctx.batch_ppr(*b, topk=40, seed=4, dense_threshold=2048)
            assert np.array_equal(e.topk_count, eager_b.topk_count)
            gb = ctx.batch_ppr(*b, topk=40, seed=4, dense_threshold=2048, graph=True)
            assert np.array_equal(gb.topk_vertex, eager_b.topk_vertex), rep
        for i in (0, 7, 19):
            ref = replay_query(rp, col, int(a[0][i]), int(a[1][i]), float(a[2][i]), i, seed=4)
            assert np.array_equal(ca[[0, 7, 19].index(i)], ref["counts"]), i
        # weighted, logged
        ctx.random_weights(seed=9)
        w, _, _ = ctx.export_weights()
        alias = oracle.build_alias(rp, w)
        src, W, tp = make_queries(gi, 8, seed=33, wmin=3000, wmax=9000)
        r = ctx.batch_ppr(src, W, tp, topk=10, seed=6, weighted=True, dense_threshold=2048)
        assert r.logged_queries >= 4
        for i in range(8):
            ref = replay_query(rp, col, int(src[i]), int(W[i]), float(tp[i]), i, seed=6, alias=alias)
            assert np.array_equal(ctx.batch_counts(i), ref["counts"]), i
    finally:
        ctx.set_tuning(lds_hubs=t.lds_hubs)
        ctx.from_csr(rp, col)


def test_full_visit_log_reruns_with_slices(ctx, g16):
    # a log of two segments cannot hold the dense queries' visits: the batch
    # runs again with dense slices and its counts are still the replay's
    gi, rp, col = g16
    ctx.from_csr(rp, col)
    src, W, tp = make_queries(gi, 6, seed=22, wmin=4200, wmax=6000)  # logged: counts <= V
    t = ctx.tuning()
    try:
        ctx.set_tuning(visit_log=2048, lds_hubs=512)
        r = ctx.batch_ppr(src, W, tp, topk=10, seed=4, dense_threshold=4096)
        counts = [ctx.batch_counts(i) for i in range(6)]
    finally:
        ctx.set_tuning(visit_log=t.visit_log, lds_hubs=t.lds_hubs)
    assert r.log_rerun and r.logged_queries == 0 and r.dense_queries == 6
    for i in range(6):
        ref = replay_query(rp, col, int(src[i]), int(W[i]), float(tp[i]), i, seed=4)
        assert np.array_equal(counts[i], ref["counts"]), i


def star(n):
    """Hub 0 joined to n leaves, both directions."""
    rp = np.zeros(n + 2, dtype=np.uint32)
    rp[1] = n
    rp[2:] = n + np.arange(1, n + 1)
    col = np.concatenate([np.arange(1, n + 1), np.zeros(n)]).astype(np.uint32)
    return rp, col


def test_dense_hub_counter_spill(ctx):
    # star: hub 0 joined to 64 leaves; from a leaf every other visit is the hub,
    # so a 49152-walker chunk of one query takes about 1.2e5 hub visits and its
    # u16 LDS counter spills to the query's slice several times per chunk
    rp, col = star(64)
    ctx.from_csr(rp, col)
    qs = ((1, 6_000_000, 0.2), (2, 3_000_000, 0.25))
    r = ctx.batch_ppr([q[0] for q in qs], [q[1] for q in qs], [q[2] for q in qs], topk=10, seed=3)
    assert r.dense_queries == 2 and r.logged_queries == 0  # far more counts than vertices: slices
    for i, (s, W, tp) in enumerate(qs):
        ref = replay_query(rp, col, s, W, tp, i, seed=3)
        assert np.array_equal(ctx.batch_counts(i), ref["counts"]), i


def test_logged_hub_counter_spill(ctx):
    # star of 200000 leaves: a 30000-walker query from a leaf makes about 1.8e5
    # counts (logged: at most one per vertex), half of them at the hub, all in
    # one 32768-walker workgroup chunk, so the hub's u16 LDS counter spills to
    # the query's hub slice about twice
    rp, col = star(200_000)
    ctx.from_csr(rp, col)
    t = ctx.tuning()
    try:
        ctx.set_tuning(dense_chunk=32768)
        r = ctx.batch_ppr([5], [30_000], [0.2], topk=10, seed=3, dense_threshold=4096)
        got = ctx.batch_counts(0)
    finally:
        ctx.set_tuning(dense_chunk=t.dense_chunk)
    assert r.logged_queries == 1
    ref = replay_query(rp, col, 5, 30_000, 0.2, 0, seed=3)
    assert ref["counts"][0] > 2 * 32768
    assert np.array_equal(got, ref["counts"])


def broom(leaves, decoys=100):
    """Source 1 joined to 2 and 3, both joined to hub 0, which has `leaves`
    leaves; then `decoys` stars of five leaves, whose centres outrank 1, 2 and
    3 by degree."""
    edges = [(1, 2), (1, 3), (2, 0), (3, 0)]
    nxt = 4
    for _ in range(leaves):
        edges.append((0, nxt))
        nxt += 1
    for _ in range(decoys):
        c = nxt
        nxt += 1
        for _ in range(5):
            edges.append((c, nxt))
            nxt += 1
    e = np.array(edges, dtype=np.int64)
    u = np.concatenate([e[:, 0], e[:, 1]])
    v = np.concatenate([e[:, 1], e[:, 0]])
    order = np.lexsort((v, u))
    u, v = u[order], v[order]
    rp = np.zeros(nxt + 1, dtype=np.uint32)
    np.add.at(rp, u + 1, 1)
    return np.cumsum(rp).astype(np.uint32), v.astype(np.uint32)


def test_logged_count_overflow_paths(ctx):
    # A logged query with large counts of every origin: with 64 hub counters
    # the source's two neighbours are not hubs, and each takes over 2^16
    # logged visits (entries of one id bucket beyond any u16); hub 0 (hub
    # slice) and the source (its own register) pass 0xFFFF.  Expected counts
    # stay below one per vertex, so the query is logged.
    rp, col = broom(1_300_000)
    ctx.from_csr(rp, col)
    t = ctx.tuning()
    try:
        ctx.set_tuning(lds_hubs=64)
        r = ctx.batch_ppr([1], [200_000], [0.2], topk=10, seed=4, dense_threshold=4096)
        got = ctx.batch_counts(0)
    finally:
        ctx.set_tuning(lds_hubs=t.lds_hubs)
    assert r.logged_queries == 1
    ref = replay_query(rp, col, 1, 200_000, 0.2, 0, seed=4)
    c = ref["counts"]
    assert c[2] > 2 * 32768 and c[3] > 2 * 32768 and c[0] > 0xFFFF and c[1] > 0xFFFF
    assert np.array_equal(got, c)
    v, cc = oracle.topk(c, rp, 10)
    assert np.array_equal(r.topk_vertex[0], v) and np.array_equal(r.topk_count[0], cc)


def test_logged_batch_many_buckets(ctx):
    # a ring of 2e7 vertices has 1221 id buckets of 16384 ids: the log's
    # partition then takes half tiles (16 segments) to leave the LDS to the
    # bucket arrays; queries near the start, the middle and the wrap-around
    n = 20_000_000
    rp = (2 * np.arange(n + 1, dtype=np.uint64)).astype(np.uint32)
    i = np.arange(n, dtype=np.int64)
    nb = np.stack([(i - 1) % n, (i + 1) % n], axis=1)
    nb.sort(axis=1)
    col = nb.reshape(-1).astype(np.uint32)
    ctx.from_csr(rp, col)
    src, W, tp = [5, 10_000_000, n - 3], [50_000, 100_000, 200_000], [0.2, 0.15, 0.3]
    r = ctx.batch_ppr(src, W, tp, topk=20, seed=7, dense_threshold=4096)
    assert r.logged_queries == 3
    for q in range(3):
        ref = replay_query(rp, col, src[q], W[q], tp[q], q, seed=7)
        assert np.array_equal(ctx.batch_counts(q), ref["counts"])
        v, c = oracle.topk(ref["counts"], rp, 20)
        assert np.array_equal(r.topk_vertex[q], v) and np.array_equal(r.topk_count[q], c)


def test_logged_small_buckets(ctx):
    # logged queries of 1100-3000 walkers on R-MAT scale 20 (64 id buckets)
    # leave most buckets with a few dozen entries, counted one bucket per
    # wavefront by a sort; the two larger queries' buckets take the LDS
    # counters; every count vector and top-k equal to the replay
    gi = ctx.rmat(scale=20, edge_factor=16, seed=3)
    rp, col = ctx.export_csr()
    rng = np.random.default_rng(4)
    n = 10
    src = rng.integers(0, gi.num_vertices, n)
    W = np.concatenate([rng.integers(1100, 3000, n - 2), [25_000, 50_000]])
    tp = rng.uniform(0.15, 0.3, n)
    r = ctx.batch_ppr(src, W, tp, topk=25, seed=12, dense_threshold=1000)
    assert r.logged_queries == n
    for i in range(n):
        ref = replay_query(rp, col, int(src[i]), int(W[i]), float(tp[i]), i, seed=12)
        assert np.array_equal(ctx.batch_counts(i), ref["counts"]), i
        v, c = oracle.topk(ref["counts"], rp, 25)
        assert np.array_equal(r.topk_vertex[i], v) and np.array_equal(r.topk_count[i], c), i
    qi = rng.integers(0, n, 3000)
    vs = rng.integers(0, gi.num_vertices, 3000)
    refs = [replay_query(rp, col, int(src[i]), int(W[i]), float(tp[i]), i, seed=12)["counts"] for i in range(n)]
    assert np.array_equal(ctx.batch_gather(qi, vs), np.array([refs[q][v] for q, v in zip(qi, vs)], dtype=np.uint64))


def test_topk_tie_overflow_path(ctx):
    # star: hub 0 joined to 5000 leaves; tp = 1 makes each walker visit the hub then one
    # leaf, so leaf counts are small and tied far beyond the 256-slot tile
    n = 5000
    rp = np.zeros(n + 2, dtype=np.uint32)
    rp[1] = n
    rp[2:] = n + np.arange(1, n + 1)
    col = np.concatenate([np.arange(1, n + 1), np.zeros(n)]).astype(np.uint32)
    ctx.from_csr(rp, col)
    t = ctx.tuning()
    ctx.set_tuning(topk_tile=256)
    try:
        r = ctx.batch_ppr([0, 0], [2000, 3000], [1.0, 1.0], topk=128, seed=2)
        for i, Wq in enumerate((2000, 3000)):
            ref = replay_query(rp, col, 0, Wq, 1.0, i, seed=2)
            v, c = oracle.topk(ref["counts"], rp, 128)
            assert np.array_equal(r.topk_vertex[i], v) and np.array_equal(r.topk_count[i], c)
    finally:
        ctx.set_tuning(topk_tile=t.topk_tile)


@pytest.mark.parametrize("visit_log", [1, 0], indirect=True)
@pytest.mark.parametrize("dense_threshold", [1, 100])
def test_pair_list_overflow_reads_slices(ctx, g12, dense_threshold, visit_log):
    # two or three walkers at tp = 0.01 take ~100 steps each, and the long
    # walks visit more distinct vertices than the pair list holds (1.25x the
    # expected visits + 64): top-k and the entropy of those queries read the
    # whole dense slice (threshold 1: every query dense; a logged query's
    # list holds its visits by construction) or hash table (threshold 100:
    # every query hashed)
    gi, rp, col = g12
    ctx.from_csr(rp, col)
    n = 128
    rng = np.random.default_rng(4)
    src = rng.integers(0, gi.num_vertices, n)
    W = rng.integers(2, 4, n)
    tp = np.full(n, 0.01)
    r = ctx.batch_ppr(src, W, tp, topk=20, seed=5, dense_threshold=dense_threshold)
    assert (r.dense_queries == n) == (dense_threshold == 1)
    overflowed = 0
    for i in range(n):
        ref = replay_query(rp, col, int(src[i]), int(W[i]), 0.01, i, seed=5)
        assert np.array_equal(ctx.batch_counts(i), ref["counts"]), i
        v, c = oracle.topk(ref["counts"], rp, 20)
        assert np.array_equal(r.topk_vertex[i], v) and np.array_equal(r.topk_count[i], c), i
        overflowed += np.count_nonzero(ref["counts"]) > int(1.25 * W[i] * 101) + 64
    assert overflowed > 0


def test_long_walks_beyond_the_planned_visits(ctx):
    # one walker per query at tp = 0.001 takes ~1000 steps; a hash query is
    # planned for 2 x 1001 + 64 visits, and the longest walks here make more
    # than twice that many (plus 256): the table, sized at twice the planned
    # visits, still holds them and the counts stay exact.
    gi = ctx.rmat(scale=14, edge_factor=8, seed=4)
    rp, col = ctx.export_csr()
    n = 128
    rng = np.random.default_rng(8)
    src = rng.integers(0, gi.num_vertices, n)
    W = np.ones(n, dtype=np.int64)
    tp = np.full(n, 0.001)
    r = ctx.batch_ppr(src, W, tp, topk=10, seed=5)
    assert r.hash_queries == n
    planned = 2 * 1001 + 64 + 256
    past = 0
    for i in range(n):
        ref = replay_query(rp, col, int(src[i]), 1, 0.001, i, seed=5)
        assert np.array_equal(ctx.batch_counts(i), ref["counts"]), i
        assert r.steps[i] == ref["steps"]
        past += ref["steps"] + 1 > 2 * planned
    assert past >= 2


def test_batch_small_graph_dangling_and_edges(ctx):
    rp, col = random_graph(30, 0.15, seed=3, dangling=3)
    ctx.from_csr(rp, col)
    r = ctx.batch_ppr([1, 2], [5000, 200], [0.2, 0.25], topk=100, seed=1)
    for i, (s, Wq, tp) in enumerate(((1, 5000, 0.2), (2, 200, 0.25))):
        ref = replay_query(rp, col, s, Wq, tp, i)
        assert np.array_equal(ctx.batch_counts(i), ref["counts"])
        v, c = oracle.topk(ref["counts"], rp, 100)
        assert np.array_equal(r.topk_vertex[i], v)
    empty = ctx.batch_ppr([], [], [], topk=10)
    assert empty.total_steps == 0


def test_batch_invalid(ctx):
    rp, col = random_graph(10, 0.3, seed=1)
    ctx.from_csr(rp, col)
    for kw in [dict(sources=[10], walkers=[10], tps=[0.2]), dict(sources=[1], walkers=[0], tps=[0.2]),
               dict(sources=[1], walkers=[10], tps=[0.0]), dict(sources=[1], walkers=[2 ** 31], tps=[0.2])]:
        with pytest.raises(ptw.PtwError) as e:
            ctx.batch_ppr(kw["sources"], kw["walkers"], kw["tps"], topk=10)
        assert e.value.code == ptw.ERR_INVALID_ARG
    with pytest.raises(ptw.PtwError):
        ctx.batch_ppr([1], [10], [0.2], topk=2000)
    with pytest.raises(ptw.PtwError):
        ctx.batch_counts(5)
    assert ctx.batch_plan_bytes([1, 2], [10, 10 ** 6], [0.2, 0.2]) >= 0


def test_hash_queries_hub_slice_and_table_scale20(ctx):
    # above 262144 vertices a hash query counts its top-degree device ids in a
    # dense hub slice and the rest in its table: both parts bit-exact against
    # the oracle, with sources of every degree and a mix of dense queries
    gi = ctx.rmat(scale=20, edge_factor=16, seed=3)
    assert gi.num_vertices > 262144
    rp, col = ctx.export_csr()
    rng = np.random.default_rng(12)
    deg = np.diff(rp.astype(np.int64))
    order = np.argsort(-deg, kind="stable")
    src = np.concatenate([order[:4], order[1000:1004], order[-4:], rng.integers(0, gi.num_vertices, 12)])
    n = len(src)
    W = np.exp(rng.uniform(np.log(200), np.log(20000), n)).astype(np.int64)
    W[:2] = 150000  # dense
    tp = rng.uniform(0.15, 0.3, n)
    r = ctx.batch_ppr(src, W, tp, topk=50, seed=7, dense_threshold=100000)
    assert r.dense_queries == 2 and r.hash_queries == n - 2
    table_hits = 0
    for i in range(n):
        ref = replay_query(rp, col, int(src[i]), int(W[i]), float(tp[i]), i, seed=7)
        got = ctx.batch_counts(i)
        assert np.array_equal(got, ref["counts"]), i
        v, c = oracle.topk(ref["counts"], rp, 50)
        assert np.array_equal(r.topk_vertex[i], v) and np.array_equal(r.topk_count[i], c), i
        # graph ids of degree rank >= 262144 land in the table
        table_hits += int(ref["counts"][order[262144:]].sum())
    assert table_hits > 0
    g = ctx.batch_gather(np.arange(n), src)
    assert all(int(g[i]) == int(ctx.batch_counts(i)[src[i]]) for i in range(n))


def test_batch_scale20_properties(ctx):
    gi = ctx.rmat(scale=20, edge_factor=16, seed=1)
    rp, _ = ctx.export_csr()
    src, W, tp = make_queries(gi, 512, seed=11, wmin=1000, wmax=1000000)
    r = ctx.batch_ppr(src, W, tp, topk=100, seed=1)
    # the largest queries' pair lists span many top-k segments (16384 entries)
    big = int(np.argmax(W))
    for i in (0, 100, 511, big):
        c = ctx.batch_counts(i)
        assert int(c.sum()) == int(W[i]) + int(r.steps[i])
        assert c[src[i]] >= W[i]
        v, cc = oracle.topk(c, rp, 100)
        assert np.array_equal(r.topk_vertex[i], v) and np.array_equal(r.topk_count[i], cc), i
    assert np.count_nonzero(ctx.batch_counts(big)) > 4 * 16384
    assert abs(r.total_steps - float(np.sum(W / tp))) < 6 * np.sqrt(np.sum(W * (1 - tp) / tp ** 2))


def test_config_batch_scale24_4096_queries(ctx24):
    """The batched configuration at its size: 4096 PPR queries on R-MAT scale 24
    (sources by degree, walkers log-uniform in [1e3, 1e6], tp uniform in
    [0.15, 0.3]), top-100, eager and as a captured graph.  32 queries -- the
    8 largest (dense, logged), the 8 smallest and 16 spread over the batch --
    are replayed alone and must match bit for bit (full count vector, steps,
    top-100); the batch's step total lies within 6 sigma of sum(W / tp)."""
    c, gi, rp, col = ctx24
    n, k = 4096, 100
    src, W, tp = batch_queries(gi, rp, n, seed=1)
    print(f"config 4: {n} queries on scale 24, {int(W.sum())} walkers", flush=True)
    r = c.batch_ppr(src, W, tp, topk=k, seed=1)
    assert r.dense_queries == n and r.logged_queries > 0  # the default threshold logs every query it can
    order = np.argsort(W, kind="stable")
    pick = np.unique(np.concatenate([order[-8:], order[:8], np.linspace(0, n - 1, 16).astype(np.int64)]))
    for i in pick:
        ref = replay_query(rp, col, int(src[i]), int(W[i]), float(tp[i]), int(i), seed=1)
        assert np.array_equal(c.batch_counts(int(i)), ref["counts"]), i
        assert r.steps[i] == ref["steps"], i
        v, cnt = oracle.topk(ref["counts"], rp, k)
        assert np.array_equal(r.topk_vertex[i], v), i
        assert np.array_equal(r.topk_count[i], cnt), i
    mean = float(np.sum(W / tp))
    sd = float(np.sqrt(np.sum(W * (1 - tp) / tp ** 2)))
    assert abs(r.total_steps - mean) < 6 * sd, (r.total_steps, mean, sd)
    assert r.total_steps == int(r.steps.sum())
    for _ in range(2):
        g = c.batch_ppr(src, W, tp, topk=k, seed=1, graph=True)
        assert np.array_equal(g.topk_vertex, r.topk_vertex) and np.array_equal(g.topk_count, r.topk_count)
        assert np.array_equal(g.steps, r.steps)


def replay_endpoints(rp, col, src, W, tp, stream, seed=1, semantics=oracle.SEM_REFERENCE):
    ref = oracle.walk_replay(rp, col, oracle.APP_RWR, W, tp=tp, seed=seed, source=src, semantics=semantics,
                             start_mode=oracle.START_SOURCE, stream=stream)
    return np.bincount(ref["ends"], minlength=len(rp) - 1).astype(np.uint64), ref["steps"]


@pytest.mark.parametrize("semantics", [oracle.SEM_REFERENCE, oracle.SEM_TEXTBOOK])
def test_batch_rwr_endpoint_queries(ctx, semantics):
    """RWR queries in a batch count each walker's last vertex: a query's count
    vector is the histogram of the oracle's RWR replay of that query alone
    (its stream), dense and hash slices, mixed with PPR queries, alone and in
    the mix, eager and as a replayed captured graph; top-k of the endpoint
    counts is exact (src/core/RWR.cpp:99-105)."""
    gi = ctx.rmat(scale=12, edge_factor=16, seed=2)
    rp, col = ctx.export_csr()
    src, W, tp = make_queries(gi, 24, seed=11, wmin=200, wmax=300000)
    kinds = np.array([ptw.QUERY_ENDPOINTS if i % 3 else ptw.QUERY_VISITS for i in range(24)])
    k = 30
    r = ctx.batch_ppr(src, W, tp, topk=k, seed=6, semantics=semantics, kinds=kinds, dense_threshold=50000)
    assert r.dense_queries > 0 and r.hash_queries > 0
    counts = []
    for i in range(24):
        got = ctx.batch_counts(i)
        counts.append(got)
        if kinds[i] == ptw.QUERY_ENDPOINTS:
            hist, steps = replay_endpoints(rp, col, int(src[i]), int(W[i]), float(tp[i]), i, seed=6,
                                           semantics=semantics)
            assert int(got.sum()) == W[i]
        else:
            ref = oracle.walk_replay(rp, col, oracle.APP_PPR, int(W[i]), tp=float(tp[i]), seed=6, source=int(src[i]),
                                     start_mode=oracle.START_SOURCE, stream=i, semantics=semantics)
            hist, steps = ref["counts"], ref["steps"]
        assert np.array_equal(got, hist), i
        assert r.steps[i] == steps, i
        v, c = oracle.topk(hist, rp, k)
        assert np.array_equal(r.topk_vertex[i], v), i
        assert np.array_equal(r.topk_count[i], c), i
    # alone equals in the mix
    for i in (1, 2, 5):
        alone = ctx.batch_ppr(src[i:i + 1], W[i:i + 1], tp[i:i + 1], streams=[i], topk=k, seed=6,
                              semantics=semantics, kinds=kinds[i:i + 1], dense_threshold=50000)
        assert np.array_equal(ctx.batch_counts(0), counts[i]), i
        assert np.array_equal(alone.topk_vertex[0], r.topk_vertex[i])
    # captured graph replays, an eager batch in between
    for rep in range(3):
        g = ctx.batch_ppr(src, W, tp, topk=k, seed=6, semantics=semantics, kinds=kinds, dense_threshold=50000,
                          graph=True)
        assert np.array_equal(g.topk_vertex, r.topk_vertex) and np.array_equal(g.topk_count, r.topk_count), rep
        assert np.array_equal(g.steps, r.steps)
        ctx.batch_ppr(src[:3], W[:3], tp[:3], topk=k, seed=7, kinds=kinds[:3])


def test_batch_rwr_matches_single_walks(ctx):
    """An endpoint query's histogram equals the histogram of ptw_walk's RWR
    endpoints with the same stream (the single-walk kernel)."""
    gi = ctx.rmat(scale=12, edge_factor=16, seed=2)
    src = int(gi.max_degree_vertex)
    r = ctx.batch_ppr([src, 3], [400000, 5000], [0.2, 0.3], streams=[9, 4], topk=10, seed=2,
                      kinds=ptw.QUERY_ENDPOINTS)
    for i, (s, w, t, st) in enumerate(((src, 400000, 0.2, 9), (3, 5000, 0.3, 4))):
        one = ptw.walk(ctx, w, ptw.RWR(t, s), seed=2, stream=st, dangling=ptw.DANGLING_STOP)
        hist = np.bincount(one.ends, minlength=gi.num_vertices).astype(np.uint64)
        assert np.array_equal(ctx.batch_counts(i), hist), i
        assert r.steps[i] == one.steps

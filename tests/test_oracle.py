"""Pins the CPU oracle before it is trusted as the GPU's checker.

Each check is chosen so that a plausible mistake in the oracle (a dropped
term, an off-by-one step, a wrong index or operand) fails it: known-answer
vectors from an independent implementation, exact integer identities,
closed forms, dense linear solves and Monte-Carlo bounds.
"""
import json
import os
from fractions import Fraction

import numpy as np
import pytest

import oracle
from walkstats import batch_zscores, random_graph, transition_check, bd_tvd_check

HERE = os.path.dirname(os.path.abspath(__file__))


# ---- Philox4x32-10 ---------------------------------------------------------------
def test_philox_known_answers():
    with open(os.path.join(HERE, "golden", "philox_kat.json")) as f:
        kat = json.load(f)
    assert len(kat["vectors"]) >= 7
    for v in kat["vectors"]:
        assert oracle.philox(v["ctr"], v["key"]).tolist() == v["out"], v


def test_philox_random123_vectors():
    # Random123 kat_vectors for philox4x32_10
    assert oracle.philox([0, 0, 0, 0], [0, 0]).tolist() == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert oracle.philox([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2).tolist() == [0x408F276D, 0x41C83B0E, 0xA20BC7C6,
                                                                          0x6D5451FD]
    assert oracle.philox([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0]).tolist() == [
        0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


# ---- integer draws ---------------------------------------------------------------------
def test_mulhi64_exact():
    rng = np.random.default_rng(5)
    cases = [(0, 0, 1), (0xFFFFFFFF, 0xFFFFFFFF, 0xFFFFFFFF), (0xFFFFFFFF, 0xFFFFFFFF, 1), (0, 0x80000000, 3)]
    cases += [tuple(int(x) for x in rng.integers(0, 2**32, 3)) for _ in range(2000)]
    for lo, hi, n in cases:
        n = max(n, 1)
        expect = ((hi << 32 | lo) * n) >> 64  # exact big-int arithmetic
        assert oracle.mulhi64(lo, hi, n) == expect
        assert expect < n


def test_continue_threshold():
    for tp in [1.0, 0.5, 0.2, 0.15, 0.3, 1e-9, 0.999999]:
        thr = oracle.continue_threshold(tp)
        exact = Fraction(1) - Fraction(tp)
        assert Fraction(thr, 2**32) <= exact < Fraction(thr + 1, 2**32) or thr == 2**32 - 1
    assert oracle.continue_threshold(1.0) == 0


# ---- walk lengths ---------------------------------------------------------------------
@pytest.mark.parametrize("tp,sem", [(0.15, oracle.SEM_REFERENCE), (0.5, oracle.SEM_REFERENCE),
                                    (0.3, oracle.SEM_TEXTBOOK)])
def test_walk_length_pmf(tp, sem):
    n = 60000
    L = np.array([oracle.walk_length(11, 0, w, tp, sem) for w in range(n)])
    off = 1 if sem == oracle.SEM_REFERENCE else 0
    assert L.min() >= off
    K = L - off
    kmax = int(np.ceil(np.log(1e-3) / np.log(1 - tp)))
    obs = np.array([np.sum(K == k) for k in range(kmax)] + [np.sum(K >= kmax)], dtype=float)
    pmf = np.array([tp * (1 - tp) ** k for k in range(kmax)] + [(1 - tp) ** kmax])
    chi2 = np.sum((obs - n * pmf) ** 2 / (n * pmf))
    # chi-square with kmax dof: mean kmax, sd sqrt(2 kmax); allow 6 sd
    assert chi2 < kmax + 6 * np.sqrt(2 * kmax), chi2


# ---- closed forms on K2 and bipartite graphs --------------------------------------
K2 = (np.array([0, 1, 2], dtype=np.uint32), np.array([1, 0], dtype=np.uint32))


def test_k2_counts_match_endpoints():
    """On K2 a walk alternates 0,1,0,...: counts[0] - counts[1] equals the number of walkers
    whose length is even, which is also the number of RWR endpoints at 0 (same Philox stream)."""
    W = 50000
    for tp in (0.15, 0.5):
        c = oracle.walk_replay(*K2, oracle.APP_PPR, W, tp=tp, seed=9, start_mode=oracle.START_SOURCE, source=0)
        e = oracle.walk_replay(*K2, oracle.APP_RWR, W, tp=tp, seed=9, start_mode=oracle.START_SOURCE, source=0)
        counts = c["counts"].astype(np.int64)
        assert counts.sum() == W + c["steps"]
        assert c["steps"] == e["steps"]
        assert counts[0] - counts[1] == np.sum(e["ends"] == 0)
        # P(L even) for L = 1 + Geometric(tp) is (1 - tp) / (2 - tp)
        frac = np.mean(e["ends"] == 0)
        p = (1 - tp) / (2 - tp)
        assert abs(frac - p) < 5 * np.sqrt(p * (1 - p) / W)


def test_bipartite_parity():
    # 4-cycle 0-1-2-3 is bipartite: {0,2} vs {1,3}; endpoint on the far side iff L odd
    rp = np.array([0, 2, 4, 6, 8], dtype=np.uint32)
    col = np.array([1, 3, 0, 2, 1, 3, 2, 0], dtype=np.uint32)
    W, tp = 80000, 0.2
    e = oracle.walk_replay(rp, col, oracle.APP_RWR, W, tp=tp, seed=4, start_mode=oracle.START_SOURCE, source=0)
    far = np.isin(e["ends"], [1, 3]).mean()
    p = 1 / (2 - tp)
    assert abs(far - p) < 5 * np.sqrt(p * (1 - p) / W)


def test_sum_of_counts_is_walkers_plus_steps():
    rp, col = random_graph(30, 0.15, seed=2)
    for app, sm in [(oracle.APP_PAGERANK, oracle.START_UNIFORM), (oracle.APP_PPR, oracle.START_SOURCE)]:
        r = oracle.walk_replay(rp, col, app, 20000, tp=0.2, seed=3, start_mode=sm, source=5)
        assert int(r["counts"].sum()) == 20000 + r["steps"] == r["visits"]


def test_replay_independent_of_threads_and_split():
    rp, col = random_graph(40, 0.1, seed=8)
    kw = dict(tp=0.15, seed=21, start_mode=oracle.START_UNIFORM)
    a = oracle.walk_replay(rp, col, oracle.APP_PAGERANK, 30000, threads=1, **kw)
    b = oracle.walk_replay(rp, col, oracle.APP_PAGERANK, 30000, threads=4, **kw)
    assert np.array_equal(a["counts"], b["counts"]) and a["steps"] == b["steps"]
    p1 = oracle.walk_replay(rp, col, oracle.APP_PAGERANK, 10000, wid_base=0, **kw)
    p2 = oracle.walk_replay(rp, col, oracle.APP_PAGERANK, 20000, wid_base=10000, **kw)
    assert np.array_equal(p1["counts"] + p2["counts"], a["counts"])


# ---- FP64 oracle against a dense linear solve --------------------------------------
@pytest.mark.parametrize("seed", [0, 1, 2])
@pytest.mark.parametrize("sem", [oracle.SEM_REFERENCE, oracle.SEM_TEXTBOOK])
@pytest.mark.parametrize("mode", ["visits", "ends"])
def test_power_iteration_matches_dense_solve(seed, sem, mode):
    rp, col = random_graph(9, 0.3, seed=seed, symmetric=(seed != 1))
    for start in (None, 0, 4):
        for tp in (0.15, 0.5):
            it = oracle.power_iteration(rp, col, tp=tp, start=start, semantics=sem, mode=mode)
            ex = oracle.dense_expected(rp, col, tp=tp, start=start, semantics=sem, mode=mode)
            assert np.allclose(it, ex, rtol=0, atol=1e-12), (it, ex)


def test_power_iteration_dangling_mass_loss():
    rp, col = random_graph(8, 0.3, seed=3, dangling=2)
    it = oracle.power_iteration(rp, col, tp=0.2, start=0, mode="visits")
    ex = oracle.dense_expected(rp, col, tp=0.2, start=0, mode="visits")
    assert np.allclose(it, ex, atol=1e-12)


def test_power_iteration_closed_forms():
    # K2 from 0, reference semantics: E[visits at 0] = E[L//2 + 1], E[visits at 1] = E[(L+1)//2]
    tp = 0.3
    e0 = sum(tp * (1 - tp) ** (L - 1) * (L // 2 + 1) for L in range(1, 400))
    e1 = sum(tp * (1 - tp) ** (L - 1) * ((L + 1) // 2) for L in range(1, 400))
    it = oracle.power_iteration(*K2, tp=tp, start=0, mode="visits")
    assert abs(it[0] - e0) < 1e-12 and abs(it[1] - e1) < 1e-12
    # regular graph, uniform start: expected visits per vertex (1 + 1/tp) / V
    V = 7
    rp = np.arange(0, 2 * V + 1, 2, dtype=np.uint32)
    col = np.array([[(u + 1) % V, (u - 1) % V] for u in range(V)], dtype=np.uint32).ravel()
    it = oracle.power_iteration(rp, col, tp=0.2, start=None, mode="visits")
    assert np.allclose(it, (1 + 1 / 0.2) / V, atol=1e-12)
    # endpoints sum to one, visits sum to E[1 + L]
    assert abs(oracle.power_iteration(rp, col, tp=0.2, start=3, mode="ends").sum() - 1) < 1e-12
    assert abs(it.sum() - (1 + 1 / 0.2)) < 1e-11


# ---- Monte-Carlo agreement of the replay with the FP64 oracle --------------------
def _batched_counts(rp, col, app, W, B, **kw):
    V = len(rp) - 1
    out = np.zeros((B, V))
    per = W // B
    for b in range(B):
        r = oracle.walk_replay(rp, col, app, per, wid_base=b * per, **kw)
        out[b] = r["counts"] / per
    return out


@pytest.mark.parametrize("start,app", [(None, oracle.APP_PAGERANK), (3, oracle.APP_PPR)])
def test_replay_counts_agree_with_fp64(start, app):
    rp, col = random_graph(25, 0.12, seed=6)
    tp = 0.15
    sm = oracle.START_UNIFORM if start is None else oracle.START_SOURCE
    batches = _batched_counts(rp, col, app, 200000, 40, tp=tp, seed=1, start_mode=sm, source=start or 0)
    exp = oracle.power_iteration(rp, col, tp=tp, start=start, mode="visits")
    z, mean = batch_zscores(batches, exp)
    assert np.max(np.abs(z)) < 5.0, z
    # an off-by-one in the step rule moves the total by 1 visit per walker
    assert abs(mean.sum() - (1 + 1 / tp)) < 0.05


def test_replay_endpoints_agree_with_fp64():
    rp, col = random_graph(20, 0.15, seed=4)
    tp, W, B = 0.2, 200000, 40
    per = W // B
    batches = np.zeros((B, 20))
    for b in range(B):
        r = oracle.walk_replay(rp, col, oracle.APP_RWR, per, tp=tp, seed=2, wid_base=b * per,
                               start_mode=oracle.START_SOURCE, source=7)
        batches[b] = np.bincount(r["ends"], minlength=20) / per
    exp = oracle.power_iteration(rp, col, tp=tp, start=7, mode="ends")
    z, _ = batch_zscores(batches, exp)
    assert np.max(np.abs(z)) < 5.0, z


def test_replay_paths_follow_transitions():
    rp, col = random_graph(12, 0.3, seed=9)
    r = oracle.walk_replay(rp, col, oracle.APP_DEEPWALK, 20000, walk_len=20, seed=5, start_mode=oracle.START_UNIFORM)
    transition_check(r["paths"], rp, col)
    bd_tvd_check(r["paths"], rp, col)
    assert r["steps"] == 20000 * 19


def test_replay_dangling_stops():
    rp, col = random_graph(10, 0.3, seed=1, dangling=3)
    r = oracle.walk_replay(rp, col, oracle.APP_DEEPWALK, 5000, walk_len=30, seed=1, start_mode=oracle.START_UNIFORM)
    paths = r["paths"]
    assert r["dangling"] > 0
    for p in paths[:500]:
        bad = np.nonzero(p == oracle.INVALID)[0]
        if len(bad):
            k = bad[0]
            assert np.all(p[k:] == oracle.INVALID)
            assert rp[p[k - 1] + 1] == rp[p[k - 1]]  # stopped at a vertex without out-edges


# ---- the reference's own RNG path ------------------------------------------------------
def _quantized_step_pmf(tp, n_max):
    """Exact pmf of the reference's step draw (src/core/pagerank.cpp:22-68) by enumerating
    the 65536 values of gen_float's 16-bit numerator (include/random.hpp:169-176)."""
    cdf = np.cumsum([(1 - tp) ** si * tp for si in range(128)])
    u = (np.arange(65536, dtype=np.float32) / np.float32(65535.0) * np.float32(1.0)).astype(np.float64)
    tot = cdf[127]
    redraw = u > tot
    si = np.searchsorted(cdf, u[~redraw], side="left")  # first si with cdf[si] >= u
    p_redraw = redraw.mean()
    base = np.bincount(si, minlength=128) / 65536.0
    pmf = np.zeros(n_max)
    # L = 1 + 128 r + si with r redraws
    for r in range(4):
        for s in range(128):
            L = 1 + 128 * r + s
            if L < n_max:
                pmf[L] += (p_redraw ** r) * base[s]
    return pmf, p_redraw


@pytest.mark.parametrize("tp", [0.15, 0.2])
def test_reference_step_draw_distribution(tp):
    n = 400000
    hist = oracle.ref_step_histogram(tp, n, seed=3, hist_len=200).astype(float)
    pmf, p_redraw = _quantized_step_pmf(tp, 200)
    assert abs(p_redraw - 1 / 65536) < 1e-12  # only u == 1.0 exceeds the table total
    keep = n * pmf > 20
    chi2 = np.sum((hist[keep] - n * pmf[keep]) ** 2 / (n * pmf[keep]))
    k = keep.sum()
    assert chi2 < k + 6 * np.sqrt(2 * k), chi2
    mean = (np.arange(200) * pmf).sum()
    assert abs(mean - (1 / tp + 0.0015)) < 0.002  # quantization bias of the 16-bit draw


# ---- the reference's xorshift* generator (include/random.hpp:144-180) ------------------
# Vigna's xorshift64* (ACM TOMS 42(4), 2016): shift triple (12, 25, 27) right/left/right,
# output multiplier 2685821657736338717.  The reference returns the low 32 bits of
# s * M taken from the state *before* the update, reduced by `% bound`.
_XS_M = 2685821657736338717
_MASK64 = (1 << 64) - 1
# 2^64 - 1 = 3 * 5 * 17 * 257 * 641 * 65537 * 6700417 (Fermat-number factors)
_PERIOD_PRIMES = (3, 5, 17, 257, 641, 65537, 6700417)


def _gf2_apply(cols, x):
    r = 0
    j = 0
    while x:
        if x & 1:
            r ^= cols[j]
        x >>= 1
        j += 1
    return r


def _gf2_mul(a, b):
    return [_gf2_apply(a, c) for c in b]


def _gf2_pow(a, e):
    r = [1 << j for j in range(64)]
    while e:
        if e & 1:
            r = _gf2_mul(r, a)
        a = _gf2_mul(a, a)
        e >>= 1
    return r


def _xorshift_matrix():
    """The transition as a product of three GF(2) matrices (I + S), S a shift."""
    ident = [1 << j for j in range(64)]
    right = lambda k: [c ^ (c >> k) for c in ident]
    left = lambda k: [c ^ ((c << k) & _MASK64) for c in ident]
    return _gf2_mul(right(27), _gf2_mul(left(25), right(12)))


def test_xorshift_constants():
    assert _XS_M == 0x2545F4914F6CDD1D
    prod = 1
    for p in _PERIOD_PRIMES:
        prod *= p
    assert prod == _MASK64


def test_xorshift_transition_has_full_period():
    """(12, 25, 27) is a full-period triple: the transition's order is 2^64 - 1.
    A mistyped triple (or shift direction) fails this for almost every choice."""
    T = _xorshift_matrix()
    ident = [1 << j for j in range(64)]
    assert _gf2_pow(T, _MASK64) == ident
    for p in _PERIOD_PRIMES:
        assert _gf2_pow(T, _MASK64 // p) != ident, p


def test_xorshift_outputs_come_from_the_pre_update_state():
    """The low 32 bits of s * M determine the low 32 bits of s (M is odd), so each
    output of gen(2^32 - 1) reveals the state it was computed from.  Output k must
    come from T^k(s0): the first from s0 itself (output before update), the
    following ones from the GF(2) matrix powers of the transition."""
    T = _xorshift_matrix()
    inv_m = pow(_XS_M & 0xFFFFFFFF, -1, 1 << 32)
    for s0 in (88172645463325252, 0x0123456789ABCDEF, 1, 1 << 63):
        n = 40
        out = oracle.xorshift_gen(s0, 0xFFFFFFFF, n).astype(np.uint64)
        s = s0
        for k in range(n):
            if (s * _XS_M) & 0xFFFFFFFF != 0xFFFFFFFF:  # the one value that `% bound` folds
                assert (int(out[k]) * inv_m) & 0xFFFFFFFF == s & 0xFFFFFFFF, (s0, k)
            s = _gf2_apply(T, s)
        # the canonical xorshift64* emits after the update: its first output differs
        assert (int(out[0]) * inv_m) & 0xFFFFFFFF != _gf2_apply(T, s0) & 0xFFFFFFFF


def test_xorshift_bound_is_a_plain_remainder():
    s0 = 0x9E3779B97F4A7C15
    for bound in (7, 1000, 65536, 0x7FFFFFFF):
        full = oracle.xorshift_gen(s0, 0xFFFFFFFF, 16).astype(np.uint64)
        got = oracle.xorshift_gen(s0, bound, 16).astype(np.uint64)
        assert np.array_equal(got, full % np.uint64(bound))


def test_xorshift_gen_float_range_ends():
    """gen_float is (low 16 bits of s * M) / 65535: numerator 0xFFFF gives exactly
    1.0 (which the geometric table's total can never reach) and 0 gives 0.0."""
    inv16 = pow(_XS_M & 0xFFFF, -1, 1 << 16)
    T = _xorshift_matrix()
    for numer, want in ((0xFFFF, 1.0), (0, 0.0), (0x7FFF, np.float32(0x7FFF) / np.float32(65535))):
        s = (0xABCD << 48) | ((numer * inv16) & 0xFFFF)
        f = oracle.xorshift_gen_float(s, 1.0, 2)
        assert f[0] == np.float32(want), (numer, f[0])
        n1 = (_gf2_apply(T, s) * _XS_M) & 0xFFFF  # the second draw comes from T(s)
        assert f[1] == np.float32(n1) / np.float32(65535)
    assert np.all(oracle.xorshift_gen(0, 1000, 100) == 0)  # 0 is the one fixed point


def test_reference_path_agrees_with_fp64():
    rp, col = random_graph(25, 0.12, seed=6)
    tp, W = 0.15, 300000
    r = oracle.ref_walk(rp, col, oracle.APP_PPR, W, tp=tp, source=3, start_uniform=False, seed=5, threads=4)
    exp = oracle.power_iteration(rp, col, tp=tp, start=3, mode="visits")
    mean = r["counts"] / W
    assert int(r["counts"].sum()) == W + r["steps"]
    # per-vertex bound: visit counts per walker have variance <= E[(1+L)^2]; a loose 6-sigma bound
    sd = np.sqrt((2 / tp ** 2 + 3 / tp + 1) / W)
    assert np.max(np.abs(mean - exp)) < 6 * sd
    assert abs(mean.sum() - (1 + 1 / tp)) < 0.05


# the reference's partitioned schedule: (L2 bytes, degree bound) pairs that give only a
# low partition, hub partitions only, and both kinds (16 hub + 21 low on this graph)
PARTITIONINGS = [(0, 1000), (4096, 12), (1024, 20)]


@pytest.mark.parametrize("l2,bound", PARTITIONINGS)
def test_partitioned_schedule_agrees_with_fp64(l2, bound):
    rp, col = random_graph(400, 0.02, seed=6)
    tp, W = 0.2, 300000
    r = oracle.ref_partitioned_walk(rp, col, oracle.APP_PAGERANK, W, tp=tp, seed=5, threads=4, l2_bytes=l2,
                                    degree_bound=bound)
    assert r["hub_partitions"] + r["low_partitions"] > 0
    if (l2, bound) == (1024, 20):
        assert r["hub_partitions"] > 1 and r["low_partitions"] > 1
    assert int(r["counts"].sum()) == W + r["steps"]
    exp = oracle.power_iteration(rp, col, tp=tp, start=None, mode="visits")
    sd = np.sqrt((2 / tp ** 2 + 3 / tp + 1) / W)
    assert np.max(np.abs(r["counts"] / W - exp)) < 6 * sd
    # single source
    r = oracle.ref_partitioned_walk(rp, col, oracle.APP_PPR, W, tp=tp, source=11, start_uniform=False, seed=2,
                                    threads=3, l2_bytes=l2, degree_bound=bound)
    exp = oracle.power_iteration(rp, col, tp=tp, start=11, mode="visits")
    assert np.max(np.abs(r["counts"] / W - exp)) < 6 * sd


@pytest.mark.parametrize("l2,bound", PARTITIONINGS)
def test_partitioned_schedule_rwr_endpoint_law(l2, bound):
    rp, col = random_graph(400, 0.02, seed=6)
    tp, W = 0.2, 300000
    r = oracle.ref_partitioned_walk(rp, col, oracle.APP_RWR, W, tp=tp, source=4, start_uniform=False, seed=3,
                                    threads=4, l2_bytes=l2, degree_bound=bound)
    assert len(r["ends"]) == W and r["ends"].max() < 400
    pe = oracle.power_iteration(rp, col, tp=tp, start=4, mode="ends")
    f = np.bincount(r["ends"], minlength=400) / W
    assert abs(pe.sum() - 1.0) < 1e-9
    # per-vertex binomial bound, 6 sigma
    assert np.all(np.abs(f - pe) <= 6 * np.sqrt(pe * (1 - pe) / W) + 1e-5)


def test_partitioned_schedule_dangling_matches_naive_schedule():
    # walkers that reach a vertex without out-edges end there, in both schedules
    rp, col = random_graph(300, 0.03, seed=9, dangling=5)
    tp, W = 0.2, 300000
    a = oracle.ref_walk(rp, col, oracle.APP_RWR, W, tp=tp, source=4, start_uniform=False, seed=3, threads=3)
    b = oracle.ref_partitioned_walk(rp, col, oracle.APP_RWR, W, tp=tp, source=4, start_uniform=False, seed=8,
                                    threads=3, l2_bytes=1024, degree_bound=20)
    assert b["dangling"] > 0
    fa = np.bincount(a["ends"], minlength=300) / W
    fb = np.bincount(b["ends"], minlength=300) / W
    # two independent samples: difference variance <= 2 p (1 - p) / W
    p = (fa + fb) / 2
    assert np.all(np.abs(fa - fb) <= 6 * np.sqrt(2 * p * (1 - p) / W) + 1e-5)
    assert fb[-5:].sum() > 0.04  # the dangling tail holds real mass


def test_partitioned_schedule_rejects_bad_arguments():
    rp, col = random_graph(20, 0.2, seed=1)
    with pytest.raises(ValueError):
        oracle.ref_partitioned_walk(rp, col, oracle.APP_PPR, 10, source=20, start_uniform=False)
    with pytest.raises(ValueError):
        oracle.ref_partitioned_walk(rp, col, oracle.APP_PPR, 10, tp=0.0)


# ---- loader semantics ----------------------------------------------------------------
def test_reference_loader_semantics(tmp_path):
    p = tmp_path / "g.txt"
    p.write_text("# comment\n7 3\n3 9\n\n9 7\n7 9\n# x y\n5 7\n")
    rp, col, name2id = oracle.load_edge_list(str(p))
    # first appearance: 7->0, 3->1, 9->2, 5->3
    assert name2id == {7: 0, 3: 1, 9: 2, 5: 3}
    assert rp.tolist() == [0, 2, 3, 4, 5]
    assert col.tolist() == [1, 2, 2, 0, 0]
    rp, col, _ = oracle.load_edge_list(str(p), as_undirected=True)
    # forward edges first, then the reversed copies, each in file order
    assert rp.tolist() == [0, 4, 6, 9, 10]
    assert col.tolist() == [1, 2, 2, 3, 2, 0, 0, 1, 0, 0]

"""The oracle's node2vec replay, pinned without the engine: with p = q = 1 it
is DeepWalk bit for bit; its second-order transition frequencies match the
law written from the definition (KnightKing's construction,
third_party/knightking/test_node2vec.hpp:47-79), including a graph where the
exact fallback decides a fifth of the steps; it does not depend on the thread
count."""
import numpy as np
import pytest

import oracle
from walkstats import random_graph


def star_ring(n_leaves=40):
    """A centre joined to every leaf, leaves joined in a ring: from a leaf into
    the centre, only the two ring neighbours of that leaf are "in"."""
    V = n_leaves + 1
    adj = [[] for _ in range(V)]
    for i in range(1, V):
        adj[0].append(i)
        adj[i].append(0)
        j = i % n_leaves + 1
        adj[i].append(j)
        adj[j].append(i)
    rp = np.zeros(V + 1, dtype=np.uint32)
    for v in range(V):
        rp[v + 1] = rp[v] + len(adj[v])
    col = np.array([x for a in adj for x in a], dtype=np.uint32)
    return rp, col


@pytest.mark.parametrize("start", [oracle.START_UNIFORM, oracle.START_ROUNDROBIN, oracle.START_SOURCE])
def test_p_q_one_is_deepwalk(start):
    rp, col = random_graph(30, 0.15, seed=3, dangling=2)
    kw = dict(seed=5, stream=2, wid_base=17, start_mode=start, source=4)
    a = oracle.node2vec_replay(rp, col, 3000, 12, 1.0, 1.0, **kw)
    b = oracle.walk_replay(rp, col, oracle.APP_DEEPWALK, 3000, walk_len=12, **kw)
    assert np.array_equal(a["paths"], b["paths"])
    assert a["steps"] == b["steps"] and a["dangling"] == b["dangling"] > 0


def transitions(paths):
    t = {}
    for row in paths:
        for i in range(len(row) - 2):
            if row[i + 2] == oracle.INVALID:
                break
            key = (int(row[i]), int(row[i + 1]))
            nxt = int(row[i + 2])
            d = t.setdefault(key, {})
            d[nxt] = d.get(nxt, 0) + 1
    return t


def check_law(rp, col, paths, p, q, min_count=400):
    """Per (prev, cur) pair with enough samples: chi-square of the next-vertex
    counts against the law, cells with an expected count below 5 pooled into
    one, summed over pairs and compared with its degrees of freedom (a
    6-sigma band)."""
    chi2, dof, pairs = 0.0, 0, 0
    for (prev, cur), d in transitions(paths).items():
        n = sum(d.values())
        if n < min_count:
            continue
        law = oracle.node2vec_law(rp, col, prev, cur, p, q)
        assert set(d) <= set(law), "a move off the adjacency"
        cells, pool_o, pool_e = [], 0, 0.0
        for x, px in law.items():
            e = n * px
            if e < 5:
                pool_o += d.get(x, 0)
                pool_e += e
            else:
                cells.append((d.get(x, 0), e))
        if pool_e > 0:
            cells.append((pool_o, pool_e))
        chi2 += sum((o - e) ** 2 / e for o, e in cells)
        dof += len(cells) - 1
        pairs += 1
    assert pairs >= 5 and dof > 0
    assert abs(chi2 - dof) < 6 * np.sqrt(2 * dof), (chi2, dof, pairs)


@pytest.mark.parametrize("p,q", [(0.25, 4.0), (4.0, 0.25), (2.0, 0.5)])
def test_transition_law_random_graph(p, q):
    rp, col = random_graph(10, 0.4, seed=7)
    r = oracle.node2vec_replay(rp, col, 40000, 20, p, q, seed=9)
    check_law(rp, col, r["paths"], p, q)


def with_chord(rp, col, u, x):
    """The graph plus one directed edge u -> x (no reverse): not symmetric."""
    rows = [list(col[rp[v]:rp[v + 1]]) for v in range(len(rp) - 1)]
    rows[u].append(x)
    rp2 = np.zeros(len(rp), dtype=np.uint32)
    rp2[1:] = np.cumsum([len(r) for r in rows])
    return rp2, np.array([y for r in rows for y in r], dtype=np.uint32)


def test_transition_law_where_the_fallback_decides():
    # p = 1/64, q = 64 on a graph that is not symmetric (one directed chord),
    # so the return edge is not folded: from a leaf into the centre the
    # acceptance per proposal is (64 + 2 + 37/64) / (64 * 40) ~ 2.6 %, and
    # ~19 % of those steps end in the exact draw after 64 rejections
    p, q = 1.0 / 64, 64.0
    rp, col = with_chord(*star_ring(40), 5, 25)
    r = oracle.node2vec_replay(rp, col, 20000, 30, p, q, seed=3)
    check_law(rp, col, r["paths"], p, q, min_count=200)
    t = transitions(r["paths"])
    centre_steps = sum(sum(d.values()) for (prev, cur), d in t.items() if cur == 0)
    assert 0.15 < r["exact_draws"] / centre_steps < 0.22
    # the centre sends walkers back to where they came from most of the time
    back = sum(d.get(prev, 0) for (prev, cur), d in t.items() if cur == 0)
    assert back / centre_steps > 0.9


def test_transition_law_with_the_return_edge_folded():
    # the same p, q on the symmetric star ring: 1/p is the largest weight, so
    # prev keeps weight 1 in its own slot and its other 63 ride on 63 extra
    # slots; from a leaf into the centre a proposal is accepted with
    # probability (64 + 2 + 37/64) / (40 + 63) ~ 65 %, and the exact draw
    # (after 64 rejections) practically never decides
    p, q = 1.0 / 64, 64.0
    rp, col = star_ring(40)
    r = oracle.node2vec_replay(rp, col, 20000, 30, p, q, seed=3)
    check_law(rp, col, r["paths"], p, q, min_count=200)
    assert r["exact_draws"] == 0
    unfolded = oracle.node2vec_replay(*with_chord(rp, col, 5, 25), 20000, 30, p, q, seed=3)
    assert r["proposals"] < unfolded["proposals"] / 5


def with_self_loop(rp, col, u):
    """The graph plus the edge u -> u: still simple and symmetric."""
    rows = [list(col[rp[v]:rp[v + 1]]) for v in range(len(rp) - 1)]
    rows[u].insert(1, u)
    rp2 = np.zeros(len(rp), dtype=np.uint32)
    rp2[1:] = np.cumsum([len(r) for r in rows])
    return rp2, np.array([y for r in rows for y in r], dtype=np.uint32)


def test_folded_law_with_a_self_loop():
    # a walker that takes the centre's loop has prev == cur; the loop is in
    # the centre's row once, so the return edge is still folded
    p, q = 0.25, 2.0
    rp, col = with_self_loop(*star_ring(12), 0)
    r = oracle.node2vec_replay(rp, col, 40000, 20, p, q, seed=6)
    check_law(rp, col, r["paths"], p, q)
    assert any((row[i] == 0 and row[i + 1] == 0) for row in r["paths"][:2000] for i in range(19))


@pytest.mark.parametrize("p,q", [(0.5, 2.0), (0.3, 1.0), (1.0 / 3000, 0.5)])
def test_folded_law_random_graph(p, q):
    # 1/p above max(1, 1/q): folded with 1, 1 (0.3: 2^32 * 2.33 excess over 3
    # slots) and 3000 / 2 extra slots
    rp, col = random_graph(10, 0.4, seed=11)
    r = oracle.node2vec_replay(rp, col, 40000, 20, p, q, seed=2)
    check_law(rp, col, r["paths"], p, q)


def test_threads_do_not_change_paths():
    rp, col = random_graph(40, 0.1, seed=1)
    a = oracle.node2vec_replay(rp, col, 5000, 15, 0.5, 2.0, seed=4, threads=1)
    b = oracle.node2vec_replay(rp, col, 5000, 15, 0.5, 2.0, seed=4, threads=8)
    assert np.array_equal(a["paths"], b["paths"])


def test_arguments_checked():
    rp, col = random_graph(8, 0.5, seed=1)
    with pytest.raises(ValueError):
        oracle.node2vec_replay(rp, col, 10, (1 << 20) + 1, 1.0, 1.0)
    with pytest.raises(ValueError):
        oracle.node2vec_replay(rp, col, 10, 5, 1e-6, 1.0)
    with pytest.raises(ValueError):
        oracle.node2vec_replay(rp, col, 10, 5, 1.0, 1e6)

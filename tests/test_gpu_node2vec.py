"""Node2vec on the device: paths bit-exact against the oracle's replay for
several (p, q), start rules, seeds, streams and walker-id offsets, with and
without the degree relabel and the neighbour records, with the return edge folded (p < 1 on simple
symmetric graphs) and unfolded (graphs with a one-way edge or a repeated
neighbour), on a graph where the exact fallback decides a fifth of the steps
at its centre, and equal to DeepWalk when p = q = 1; the CLI binary end to
end."""
import os
import subprocess

import numpy as np
import pytest

import oracle
import ptwalker_amd as ptw
from test_node2vec_oracle import star_ring, with_chord, with_self_loop
from walkstats import random_graph

pytestmark = pytest.mark.gpu


def check(ctx, rp, col, W, L, p, q, **kw):
    start = kw.pop("start", ptw.START_UNIFORM)
    source = kw.pop("source", None)
    app = ptw.Node2Vec(L, p, q, start=start, source=source)
    got = ptw.walk(ctx, W, app, **kw)
    ref = oracle.node2vec_replay(rp, col, W, L, p, q, seed=kw.get("seed", 1), stream=kw.get("stream", 0),
                                 wid_base=kw.get("wid_base", 0),
                                 start_mode=ptw.START_SOURCE if source is not None else start,
                                 source=source or 0)
    assert np.array_equal(got.paths, ref["paths"])
    assert got.steps == ref["steps"] and got.dangling == ref["dangling"]
    return got


@pytest.fixture(scope="module")
def rmat12(ctx):
    gi = ctx.rmat(scale=12, edge_factor=16, seed=5)
    rp, col = ctx.export_csr()
    return gi, rp, col


@pytest.mark.parametrize("p,q", [(1.0, 1.0), (0.5, 2.0), (2.0, 0.5), (1.0 / 64, 64.0), (64.0, 1.0 / 64),
                                 (0.3, 1.0), (1.0 / 3000, 0.5)])
def test_parity_rmat12(ctx, rmat12, p, q):
    gi, rp, col = rmat12
    check(ctx, rp, col, 20000, 20, p, q, seed=3)


@pytest.mark.parametrize("p,q", [(0.5, 2.0), (2.0, 0.5), (1.0 / 3000, 0.5)])
def test_parity_without_records(ctx, rmat12, p, q):
    # proposals read `col` and a step the row offsets when the graph keeps
    # no neighbour records: the same paths
    gi, rp, col = rmat12
    t = ctx.tuning()
    try:
        ctx.set_tuning(edge_records=0)
        check(ctx, rp, col, 20000, 20, p, q, seed=3)
    finally:
        ctx.set_tuning(edge_records=t.edge_records)


def test_parity_starts_seeds_streams(ctx, rmat12):
    gi, rp, col = rmat12
    check(ctx, rp, col, 5000, 12, 0.5, 2.0, seed=11, stream=7, wid_base=(1 << 33) + 5)
    check(ctx, rp, col, 5000, 12, 2.0, 0.25, start=ptw.START_ROUNDROBIN, seed=2)
    check(ctx, rp, col, 5000, 12, 0.25, 4.0, source=int(gi.max_degree_vertex), seed=4)


def test_equals_deepwalk_when_p_q_one(ctx, rmat12):
    gi, rp, col = rmat12
    a = ptw.walk(ctx, 10000, ptw.Node2Vec(16, 1.0, 1.0), seed=9, stream=3)
    b = ptw.walk(ctx, 10000, ptw.DeepWalk(16), seed=9, stream=3)
    assert np.array_equal(a.paths, b.paths) and a.steps == b.steps


def test_fallback_heavy_and_no_relabel(ctx):
    # the symmetric star ring folds the return edge; one directed chord keeps
    # it unfolded, and then the exact fallback decides a fifth of the centre's
    # steps
    t = ctx.tuning()
    try:
        for rp, col in (star_ring(40), with_chord(*star_ring(40), 5, 25)):
            for flags in (0, ptw.GRAPH_NO_RELABEL):
                ctx.from_csr(rp, col, flags=flags)
                for rec in (1, 0):
                    ctx.set_tuning(edge_records=rec)
                    check(ctx, rp, col, 20000, 30, 1.0 / 64, 64.0, seed=3)
    finally:
        ctx.set_tuning(edge_records=t.edge_records)


def test_self_loop_folded(ctx):
    rp, col = with_self_loop(*star_ring(12), 0)
    ctx.from_csr(rp, col)
    check(ctx, rp, col, 20000, 20, 0.25, 2.0, seed=6)


def test_repeated_neighbour_is_unfolded(ctx):
    # a row that lists a neighbour twice holds prev twice on the way back:
    # the graph is not simple, so the return edge is not folded
    rp, col = random_graph(40, 0.15, seed=8)
    rows = [list(col[rp[v]:rp[v + 1]]) for v in range(len(rp) - 1)]
    rows[3].append(rows[3][0])
    rp2 = np.zeros(len(rp), dtype=np.uint32)
    rp2[1:] = np.cumsum([len(r) for r in rows])
    col2 = np.array([y for r in rows for y in r], dtype=np.uint32)
    for g in ((rp, col), (rp2, col2)):
        ctx.from_csr(*g)
        check(ctx, *g, 8000, 20, 0.25, 2.0, seed=5)


def test_dangling_vertices(ctx):
    rp, col = random_graph(60, 0.08, seed=2, symmetric=False, dangling=5)
    ctx.from_csr(rp, col)
    got = check(ctx, rp, col, 8000, 15, 0.5, 2.0, seed=6, dangling=ptw.DANGLING_STOP)
    assert got.dangling > 0


def test_rejects_bad_arguments(ctx, rmat12):
    gi, rp, col = rmat12
    ctx.rmat(scale=12, edge_factor=16, seed=5)
    with pytest.raises(ptw.PtwError):
        ptw.walk(ctx, 10, ptw.Node2Vec(10, 0.0, 1.0))
    with pytest.raises(ptw.PtwError):
        ptw.walk(ctx, 10, ptw.Node2Vec(10, 1.0, 1e9))
    with pytest.raises(ptw.PtwError):
        ptw.walk(ctx, 10, ptw.Node2Vec((1 << 20) + 1, 1.0, 1.0))


def test_cli_node2vec(ctx, tmp_path):
    rng = np.random.default_rng(1)
    e = rng.integers(0, 300, size=(3000, 2))
    e = e[e[:, 0] != e[:, 1]]
    g = tmp_path / "g.txt"
    g.write_text("\n".join(f"{a} {b}" for a, b in e) + "\n")
    out = tmp_path / "paths.bin"
    binp = os.path.join(os.path.dirname(ptw.__file__), "bin", "node2vec")
    r = subprocess.run([binp, "-f", "text", "-g", str(g), "-w", "2000", "-l", "10", "-p", "0.5", "-q", "2",
                        "-t", "4", "--seed", "5", f"--out={out}"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "total time (including init start walker)" in r.stdout
    paths = np.fromfile(out, dtype=np.uint32).reshape(2000, 10)
    # the binary loads the file as undirected, as the reference's node2vec does
    rp, col, name2id = oracle.load_edge_list(str(g), "text", True)
    names = np.zeros(len(name2id), dtype=np.uint32)
    for name, v in name2id.items():
        names[v] = name
    ref = oracle.node2vec_replay(rp, col, 2000, 10, 0.5, 2.0, seed=5)
    assert np.array_equal(paths, names[ref["paths"]])  # in the file's own ids

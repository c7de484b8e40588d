"""Graph ingest on the device (ptw_graph_load with PTW_GRAPH_INGEST_DEVICE):
the CSR equals the host loader's bit for bit -- same first-appearance ids,
same adjacency order -- on text files full of the format's corner cases, on
binary files with a partial trailing record, directed and undirected, and on
a file larger than one staging chunk; on well-formed files both equal the
oracle's restatement of the reference loader."""
import numpy as np
import pytest

import oracle
import ptwalker_amd as ptw

pytestmark = pytest.mark.gpu


def both(ctx, host_ctx, path, fmt, undirected):
    ctx.load(path, fmt, undirected, flags=ptw.GRAPH_INGEST_DEVICE)
    rd, cd = ctx.export_csr()
    host_ctx.load(path, fmt, undirected, flags=ptw.GRAPH_INGEST_HOST)
    nd = ctx.export_names()
    rh, ch = host_ctx.export_csr()
    assert np.array_equal(rd, rh) and np.array_equal(cd, ch)
    assert np.array_equal(nd, host_ctx.export_names())  # the file id of each compact id
    return rd, cd


def messy_text(rng, n):
    """Edge lines mixed with every kind of line the parser must skip or read."""
    ids = rng.integers(0, 5000, size=(n, 2))
    ids[rng.random(n) < 0.01, 0] = 4294967295
    out = []
    for i in range(n):
        a, b = int(ids[i, 0]), int(ids[i, 1])
        k = rng.integers(0, 14)
        if k == 0:
            out.append(f"# comment {a} {b}")
        elif k == 1:
            out.append("")
        elif k == 2:
            out.append(f"\t{a}\t{b}\r")
        elif k == 3:
            out.append(f"+{a} +{b} trailing words 7 8")
        elif k == 4:
            out.append(f"{a}")  # one number: skipped
        elif k == 5:
            out.append(f"-{a} {b}")  # sign: skipped
        elif k == 6:
            out.append(f"\v{a} \f{b}")
        elif k == 7:
            out.append(f"{a} 99999999999999")  # saturates to 2^32 - 1
        elif k == 8:
            out.append(f"   {a}    {b}   ")
        elif k == 9:
            out.append("   ")
        else:
            out.append(f"{a} {b}")
    return "\n".join(out)  # no final newline


@pytest.mark.parametrize("undirected", [False, True])
def test_text_corner_cases_match_host(ctx, host_ctx, tmp_path, undirected):
    rng = np.random.default_rng(1)
    p = tmp_path / "messy.txt"
    p.write_text(messy_text(rng, 30000))
    rp, col = both(ctx, host_ctx, str(p), ptw.FORMAT_TEXT, undirected)
    assert rp[-1] == len(col) > 10000


def test_newline_followed_by_vertical_tab(ctx, host_ctx, tmp_path):
    # a 0x0B byte right after each newline once fooled the borrow form of the byte test
    p = tmp_path / "vt.txt"
    p.write_bytes(b"".join(b"%d %d\n\x0b" % (i, (i * 7) % 50) for i in range(20000)))
    both(ctx, host_ctx, str(p), ptw.FORMAT_TEXT, False)


def test_well_formed_matches_oracle(ctx, host_ctx, tmp_path):
    rng = np.random.default_rng(2)
    e = rng.integers(0, 3000, size=(20000, 2))
    p = tmp_path / "plain.txt"
    p.write_text("# header\n" + "\n".join(f"{a} {b}" for a, b in e) + "\n")
    for und in (False, True):
        rp, col = both(ctx, host_ctx, str(p), ptw.FORMAT_TEXT, und)
        orp, ocol, _ = oracle.load_edge_list(str(p), "text", und)
        assert np.array_equal(rp, orp) and np.array_equal(col, ocol)


@pytest.mark.parametrize("undirected", [False, True])
def test_binary_with_partial_record(ctx, host_ctx, tmp_path, undirected):
    rng = np.random.default_rng(3)
    e = rng.integers(0, 2 ** 32, size=(50001, 2), dtype=np.uint64).astype(np.uint32)
    e[::3] %= 1000  # repeats
    p = tmp_path / "g.bin"
    p.write_bytes(e.tobytes() + b"\x01\x02\x03")
    rp, col = both(ctx, host_ctx, str(p), ptw.FORMAT_BINARY, undirected)
    orp, ocol, _ = oracle.load_edge_list(str(p), "binary", undirected)
    assert np.array_equal(rp, orp) and np.array_equal(col, ocol)


def test_file_larger_than_a_staging_chunk(ctx, host_ctx, tmp_path):
    rng = np.random.default_rng(4)
    n = 6_000_000  # about 90 MB of text: two 64 MiB staging chunks
    a = rng.integers(0, 1 << 20, n)
    b = rng.integers(0, 1 << 20, n)
    text = "\n".join(np.char.add(np.char.add(a.astype(str), " "), b.astype(str)).tolist()) + "\n"
    p = tmp_path / "big.txt"
    p.write_text(text)
    assert p.stat().st_size > 70 << 20
    rp, col = both(ctx, host_ctx, str(p), ptw.FORMAT_TEXT, False)
    assert len(col) == n
    # the default picks the device for a file this size; the graph is the same
    ctx.load(str(p), ptw.FORMAT_TEXT)
    r2, c2 = ctx.export_csr()
    assert np.array_equal(r2, rp) and np.array_equal(c2, col)


def test_comment_only_file_is_rejected(ctx, tmp_path):
    p = tmp_path / "empty.txt"
    p.write_text("# nothing\n\n")
    with pytest.raises(ptw.PtwError):
        ctx.load(str(p), ptw.FORMAT_TEXT, flags=ptw.GRAPH_INGEST_DEVICE)


# ---- device primitives and the device-side degree relabel -------------------
# sizes around one 4096-item tile, a few thousand tiles (two scan levels) and
# 2^24 + 5 items (three: 4097 tiles -> 2 -> 1); digit widths that leave a
# partial last pass
@pytest.mark.parametrize("n,bits", [(0, 32), (1, 1), (4095, 8), (4096, 32), (4097, 13), (100_003, 7),
                                    (3_000_001, 17), (16_777_221, 32), (16_777_221, 20)])
def test_primitives_match_host(ctx, n, bits):
    ctx.check_primitives(n, bits, seed=n + bits)


def walk_outputs(c, V, seed=3):
    info = c.info()
    outs = []
    for app in (ptw.PPR(0.2, int(info.max_degree_vertex)), ptw.PageRank(0.15)):
        counts = np.zeros(V, dtype=np.uint64)
        c.walk(ptw.make_desc(app, 200_000, seed=seed, dangling=ptw.DANGLING_STOP), counts=counts)
        outs.append(counts)
    W, L = 5000, 12
    paths = np.zeros(W * L, dtype=np.uint32)
    c.walk(ptw.make_desc(ptw.DeepWalk(L), W, seed=seed, dangling=ptw.DANGLING_STOP), paths=paths)
    outs.append(paths)
    return outs


@pytest.mark.parametrize("fmt,undirected,no_relabel", [(ptw.FORMAT_TEXT, True, False), (ptw.FORMAT_BINARY, False, False),
                                                       (ptw.FORMAT_TEXT, False, True)])
def test_device_built_graph_walks_like_host_built(ctx, tmp_path, fmt, undirected, no_relabel):
    """The device CSR (relabelled on the device, or in graph ids) walks exactly
    like the one the host relabels and uploads: same counts, same paths."""
    rng = np.random.default_rng(11)
    n = 300_000
    # skewed sparse ids out of order, so degrees spread and the relabel moves vertices
    a = (rng.zipf(1.6, n) * 7919) % 4_000_000_000
    b = (rng.zipf(1.6, n) * 104729 + 13) % 4_000_000_000
    p = tmp_path / ("g.txt" if fmt == ptw.FORMAT_TEXT else "g.bin")
    if fmt == ptw.FORMAT_TEXT:
        p.write_text("\n".join(f"{x} {y}" for x, y in zip(a.tolist(), b.tolist())) + "\n")
    else:
        p.write_bytes(np.stack([a, b], 1).astype(np.uint32).tobytes())
    flags = ptw.GRAPH_NO_RELABEL if no_relabel else 0
    ctx.load(str(p), fmt, undirected, flags=flags | ptw.GRAPH_INGEST_DEVICE)
    other = ptw.Context(0)
    try:
        other.load(str(p), fmt, undirected, flags=flags | ptw.GRAPH_INGEST_HOST)
        i1, i2 = ctx.info(), other.info()
        assert (i1.num_vertices, i1.num_edges, i1.max_degree_vertex, i1.relabelled) == \
               (i2.num_vertices, i2.num_edges, i2.max_degree_vertex, i2.relabelled)
        assert i1.relabelled == (0 if no_relabel else 1)
        r1, c1 = ctx.export_csr()
        r2, c2 = other.export_csr()
        assert np.array_equal(r1, r2) and np.array_equal(c1, c2)
        V = i1.num_vertices
        for x, y in zip(walk_outputs(ctx, V), walk_outputs(other, V)):
            assert np.array_equal(x, y)
    finally:
        other.close()


def test_device_relabel_more_vertices_than_edges(ctx, host_ctx, tmp_path):
    """A directed star read from its leaves: V = E + 1, every leaf but the
    centre dangling, so the degree sort covers more vertices than edges."""
    p = tmp_path / "star.txt"
    p.write_text("".join(f"{i + 1} 0\n" for i in range(50_000)))
    rd, cd = both(ctx, host_ctx, str(p), ptw.FORMAT_TEXT, False)
    assert len(rd) - 1 == len(cd) + 1
    counts = np.zeros(len(rd) - 1, dtype=np.uint64)
    ctx.walk(ptw.make_desc(ptw.PageRank(0.3), 100_000, seed=2, dangling=ptw.DANGLING_STOP), counts=counts)
    exp = oracle.walk_replay(rd, cd, oracle.APP_PAGERANK, 100_000, tp=0.3, seed=2)
    assert np.array_equal(counts, exp["counts"])

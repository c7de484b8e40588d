"""The C ABI: the library loads, exports exactly what include/ptw.h declares,
and its host-side logic (graph ingest, R-MAT, validation, error codes) works
without a device.  No walk runs here."""
import os
import re
import subprocess

import numpy as np
import pytest

import oracle
import ptwalker_amd as ptw

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    syms = set()
    for name in os.listdir(os.path.join(ROOT, "include")):
        if not name.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", name)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        syms |= set(re.findall(r"\b(ptw_[a-z0-9_]+)\s*\(", text))
    return syms


def test_header_symbols_exported():
    lib = ptw.load_library()
    syms = declared_symbols()
    assert "ptw_walk" in syms and "ptw_create" in syms
    out = subprocess.run(["nm", "-D", "--defined-only", ptw.LIB_PATH], capture_output=True, text=True, check=True)
    exported = set(re.findall(r"\bT (ptw_[a-z0-9_]+)\b", out.stdout))
    missing = syms - exported
    assert not missing, f"declared but not exported: {missing}"
    for s in syms:
        assert hasattr(lib, s)
    assert set(ptw.exported_symbols()) == syms, "Python bindings out of sync with include/"


def test_version():
    assert ptw.load_library().ptw_version() == 1


def test_host_only_context_errors(host_ctx, tmp_path):
    with pytest.raises(ptw.PtwError) as e:
        host_ctx.load(str(tmp_path / "missing.txt"))
    assert e.value.code == ptw.ERR_IO
    with pytest.raises(ptw.PtwError) as e:
        host_ctx.load(str(tmp_path / "missing.txt"), fmt=7)
    assert e.value.code == ptw.ERR_INVALID_ARG
    bad_rp = np.array([0, 1, 2], dtype=np.uint32)
    with pytest.raises(ptw.PtwError) as e:
        host_ctx.from_csr(bad_rp, np.array([1, 5], dtype=np.uint32))
    assert e.value.code == ptw.ERR_INVALID_ARG
    with pytest.raises(ptw.PtwError) as e:
        host_ctx.from_csr(np.array([0, 2, 1], dtype=np.uint32), np.array([1, 0], dtype=np.uint32))
    assert e.value.code == ptw.ERR_INVALID_ARG
    host_ctx.from_csr(np.array([0, 1, 2], dtype=np.uint32), np.array([1, 0], dtype=np.uint32))
    with pytest.raises(ptw.PtwError) as e:
        ptw.walk(host_ctx, 10, ptw.PageRank(0.15))
    assert e.value.code == ptw.ERR_UNSUPPORTED
    with pytest.raises(ptw.PtwError) as e:
        host_ctx.bench_gather(1 << 20, 1 << 20)
    assert e.value.code == ptw.ERR_UNSUPPORTED
    empty = tmp_path / "empty.txt"
    empty.write_text("# nothing\n")
    with pytest.raises(ptw.PtwError) as e:
        host_ctx.load(str(empty))
    assert e.value.code == ptw.ERR_FORMAT


def test_no_graph_error():
    c = ptw.Context(-1)
    with pytest.raises(ptw.PtwError) as e:
        c.info()
    assert e.value.code == ptw.ERR_NO_GRAPH
    c.close()


def _write_random_edges(path, n_edges, n_names, seed, fmt):
    rng = np.random.default_rng(seed)
    e = rng.integers(0, n_names, size=(n_edges, 2)).astype(np.uint32) * 3 + 1  # sparse names
    if fmt == "text":
        with open(path, "w") as f:
            f.write("# header line\n")
            for i, (a, b) in enumerate(e):
                f.write(f"{a} {b}\n")
                if i % 97 == 0:
                    f.write("# interleaved comment\n")
    else:
        e.tofile(path)
    return e


@pytest.mark.parametrize("fmt", ["text", "binary"])
@pytest.mark.parametrize("undirected", [False, True])
def test_loader_matches_reference_semantics(host_ctx, tmp_path, fmt, undirected):
    path = str(tmp_path / f"g.{fmt}")
    _write_random_edges(path, 3000, 500, seed=3, fmt=fmt)
    gi = host_ctx.load(path, fmt=ptw.FORMAT_TEXT if fmt == "text" else ptw.FORMAT_BINARY, as_undirected=undirected)
    rp, col = host_ctx.export_csr()
    erp, ecol, name2id = oracle.load_edge_list(path, fmt, undirected)
    assert gi.num_vertices == len(name2id)
    assert np.array_equal(rp, erp) and np.array_equal(col, ecol)
    assert np.array_equal(host_ctx.export_names(), _names_of(name2id))


def _names_of(name2id):
    names = np.zeros(len(name2id), dtype=np.uint32)
    for name, v in name2id.items():
        names[v] = name
    return names


def test_loader_sparse_ids_hash_path(host_ctx, tmp_path):
    """Ids up to 2^32 - 2 over a few thousand edges: the loader compacts them
    through a hash map (a table indexed by id would take 16 GiB) with the same
    first-appearance ids and names."""
    rng = np.random.default_rng(8)
    pool = np.array([0xFFFFFFFE, 7, 1 << 31, 123456789, 4000000000, 5, 3999999999], dtype=np.uint64)
    ids = np.concatenate([pool, rng.choice(np.arange(10, 1 << 32, 65537, dtype=np.uint64), 300, replace=False)])
    e = ids[rng.integers(0, len(ids), size=(4000, 2))]
    for fmt in ("text", "binary"):
        path = str(tmp_path / f"s.{fmt}")
        if fmt == "text":
            with open(path, "w") as f:
                f.write("".join(f"{a} {b}\n" for a, b in e))
        else:
            e.astype(np.uint32).tofile(path)
        host_ctx.load(path, fmt=ptw.FORMAT_TEXT if fmt == "text" else ptw.FORMAT_BINARY)
        rp, col = host_ctx.export_csr()
        erp, ecol, name2id = oracle.load_edge_list(path, fmt)
        assert np.array_equal(rp, erp) and np.array_equal(col, ecol)
        assert np.array_equal(host_ctx.export_names(), _names_of(name2id))
        assert host_ctx.export_names()[0] == e[0, 0]


def test_rmat_names_are_generator_ids(host_ctx):
    """R-MAT compacts isolated ids away; names are the generator's ids, in
    order, and each kept vertex has the degree it has with isolated ids kept."""
    host_ctx.rmat(scale=10, edge_factor=2, seed=3, flags=ptw.RMAT_KEEP_ISOLATED)
    rp_all, _ = host_ctx.export_csr()
    assert np.array_equal(host_ctx.export_names(), np.arange(1024))
    host_ctx.rmat(scale=10, edge_factor=2, seed=3)
    rp, _ = host_ctx.export_csr()
    names = host_ctx.export_names()
    assert len(names) < 1024 and np.all(np.diff(names.astype(np.int64)) > 0)
    assert np.array_equal(np.diff(rp.astype(np.int64)), np.diff(rp_all.astype(np.int64))[names])


def test_batch_topk_writer_format(host_ctx, tmp_path):
    """ptw_batch_write_topk: one line per listed entry, padding left out, ids in
    the input's own space or graph ids."""
    path = str(tmp_path / "g.txt")
    with open(path, "w") as f:
        f.write("50 60\n60 70\n70 50\n90 50\n")
    host_ctx.load(path)
    names = host_ctx.export_names()
    assert names.tolist() == [50, 60, 70, 90]
    res = ptw.BatchResult(np.array([[1, 2, ptw.INVALID_VERTEX], [0, 3, 1]], dtype=np.uint32),
                          np.array([[9, 4, 0], [7, 7, 1]], dtype=np.uint64), np.zeros(2, np.uint64), 0, 0, 0, 0,
                          0.0, 0.0, np.zeros(2))
    out = str(tmp_path / "t.txt")
    host_ctx.batch_write_topk(out, [3, 1], [100, 200], [0.2, 0.3], res, kinds=[0, 1])
    rows = [l.split() for l in open(out).read().splitlines()]
    assert rows[0][0] == "#"
    assert rows[1:] == [s.split() for s in ("0 90 0 0 60 9", "0 90 0 1 70 4", "1 60 1 0 50 7", "1 60 1 1 90 7",
                                             "1 60 1 2 60 1")]
    host_ctx.batch_write_topk(out, [3, 1], [100, 200], [0.2, 0.3], res, kinds=[0, 1], id_space=ptw.IDS_GRAPH)
    rows = [l.split() for l in open(out).read().splitlines()]
    assert rows[1] == "0 3 0 0 1 9".split()


def test_large_text_parallel_parse(host_ctx, tmp_path):
    # > 1 MiB so the parser splits the file across threads
    path = str(tmp_path / "big.txt")
    _write_random_edges(path, 150000, 20000, seed=9, fmt="text")
    host_ctx.load(path)
    rp, col = host_ctx.export_csr()
    erp, ecol, _ = oracle.load_edge_list(path)
    assert np.array_equal(rp, erp) and np.array_equal(col, ecol)


def test_write_then_load_roundtrip(host_ctx, tmp_path):
    host_ctx.rmat(scale=9, edge_factor=4, seed=5)
    rp, col = host_ctx.export_csr()
    for fmt in (ptw.FORMAT_TEXT, ptw.FORMAT_BINARY):
        p = str(tmp_path / f"rt{fmt}")
        host_ctx.write(p, fmt)
        host_ctx.load(p, fmt=fmt)
        rp2, col2 = host_ctx.export_csr()
        # the edge multiset survives (ids are re-compacted by first appearance)
        erp, ecol, name2id = oracle.load_edge_list(p, "text" if fmt == ptw.FORMAT_TEXT else "binary")
        assert np.array_equal(rp2, erp) and np.array_equal(col2, ecol)
        inv = {v: k for k, v in name2id.items()}
        e1 = sorted((u, int(col[k])) for u in range(len(rp) - 1) for k in range(rp[u], rp[u + 1]))
        e2 = sorted((inv[u], inv[int(col2[k])]) for u in range(len(rp2) - 1) for k in range(rp2[u], rp2[u + 1]))
        assert e1 == e2
        host_ctx.rmat(scale=9, edge_factor=4, seed=5)


def test_rmat_properties(host_ctx):
    gi = host_ctx.rmat(scale=11, edge_factor=16, seed=1)
    rp, col = host_ctx.export_csr()
    V = gi.num_vertices
    assert 0 < V <= 2048 and gi.num_dangling == 0
    src = np.repeat(np.arange(V), np.diff(rp.astype(np.int64)))
    assert np.all(src != col), "self-loop"
    for u in range(V):
        nb = col[rp[u]:rp[u + 1]]
        assert np.all(np.diff(nb.astype(np.int64)) > 0), "adjacency not sorted/deduplicated"
    # symmetric
    fwd = set(zip(src.tolist(), col.tolist()))
    assert all((b, a) in fwd for a, b in fwd)
    deg = np.diff(rp.astype(np.int64))
    assert deg.max() > 10 * deg.mean(), "R-MAT skew missing"
    assert gi.max_degree == deg.max() and gi.max_degree_vertex == int(np.argmax(deg))
    # deterministic in the seed
    host_ctx.rmat(scale=11, edge_factor=16, seed=1)
    rp2, col2 = host_ctx.export_csr()
    assert np.array_equal(rp, rp2) and np.array_equal(col, col2)
    host_ctx.rmat(scale=11, edge_factor=16, seed=2)
    rp3, _ = host_ctx.export_csr()
    assert not (len(rp3) == len(rp) and np.array_equal(rp3, rp))


def test_rmat_keep_isolated(host_ctx):
    gi = host_ctx.rmat(scale=10, edge_factor=2, seed=3, flags=ptw.RMAT_KEEP_ISOLATED)
    assert gi.num_vertices == 1024 and gi.num_dangling > 0
    gi2 = host_ctx.rmat(scale=10, edge_factor=2, seed=3)
    assert gi2.num_vertices == 1024 - gi.num_dangling


def test_rmat_quadrant_probabilities(host_ctx):
    # with a = 1 every draw lands on (0, 0): a self-loop, so the graph is empty
    with pytest.raises(ptw.PtwError):
        host_ctx.rmat(scale=6, edge_factor=4, a=1.0, b=0.0, c=0.0, seed=1)
    # with b = 1 every edge is (0, 2^s - 1)
    gi = host_ctx.rmat(scale=6, edge_factor=4, a=0.0, b=1.0, c=0.0, seed=1)
    rp, col = host_ctx.export_csr()
    assert gi.num_vertices == 2 and gi.num_edges == 2


def test_ingest_flags_checked(host_ctx, tmp_path):
    p = tmp_path / "g.txt"
    p.write_text("0 1\n1 2\n")
    with pytest.raises(ptw.PtwError):
        host_ctx.load(str(p), ptw.FORMAT_TEXT, flags=ptw.GRAPH_INGEST_HOST | ptw.GRAPH_INGEST_DEVICE)
    with pytest.raises(ptw.PtwError):  # a host-only context has no device to ingest on
        host_ctx.load(str(p), ptw.FORMAT_TEXT, flags=ptw.GRAPH_INGEST_DEVICE)
    gi = host_ctx.load(str(p), ptw.FORMAT_TEXT, flags=ptw.GRAPH_INGEST_HOST)
    assert gi.num_vertices == 3 and gi.num_edges == 2


def test_walk_host_outputs_checked_before_the_call(host_ctx):
    """Context.walk rejects host outputs the library would overrun (wrong dtype,
    too short, not contiguous) before it calls ptw_walk."""
    gi = host_ctx.rmat(scale=8, edge_factor=4, seed=1)
    V = gi.num_vertices
    d = ptw.make_desc(ptw.PPR(0.2, 0), 100)
    bad_counts = [np.zeros(V, dtype=np.uint32), np.zeros(V - 1, dtype=np.uint64),
                  np.zeros(2 * V, dtype=np.uint64)[::2], [0] * V]
    for counts in bad_counts:
        with pytest.raises(ValueError):
            host_ctx.walk(d, counts=counts)
    d = ptw.make_desc(ptw.RWR(0.2, 0), 100)
    with pytest.raises(ValueError):
        host_ctx.walk(d, ends=np.zeros(99, dtype=np.uint32))
    d = ptw.make_desc(ptw.DeepWalk(5), 100)
    with pytest.raises(ValueError):
        host_ctx.walk(d, paths=np.zeros(499, dtype=np.uint32))
    # well-formed outputs reach the library, which has no device here
    with pytest.raises(ptw.PtwError):
        host_ctx.walk(d, paths=np.zeros(500, dtype=np.uint32))

"""The edge-list converters format_knk and format_yahoo (host tools), against
hand-computed outputs and through the library's loader."""
import os
import subprocess

import numpy as np
import pytest

import ptwalker_amd as ptw

BIN = os.path.join(os.path.dirname(ptw.__file__), "bin")


def run(tool, *args):
    return subprocess.run([os.path.join(BIN, tool), *args], capture_output=True, text=True, timeout=60)


def edges_of(path):
    return np.fromfile(path, dtype=np.uint32).reshape(-1, 2)


def test_format_knk_compacts_by_first_appearance(tmp_path):
    src = tmp_path / "g.txt"
    src.write_text("# comment\n100 5\n5 7\n\n100 7\nbad line\n7 100 extra\n42")
    out = tmp_path / "g.bin"
    r = run("format_knk", "-i", str(src), "-o", str(out))
    assert r.returncode == 0, r.stderr
    # names in first-appearance order: 100 -> 0, 5 -> 1, 7 -> 2 ("42" alone is not an edge)
    assert edges_of(out).tolist() == [[0, 1], [1, 2], [0, 2], [2, 0]]
    info = (tmp_path / "g.bin.info.txt").read_text()
    assert info == f"# Converted from: {src}\n# vertex number: 3\n# edegs: 4\n"


def test_format_knk_output_loads_as_the_same_graph(host_ctx, tmp_path):
    rng = np.random.default_rng(1)
    e = rng.integers(10 ** 6, 10 ** 6 + 500, size=(3000, 2))
    src = tmp_path / "g.txt"
    src.write_text("\n".join(f"{a} {b}" for a, b in e) + "\n")
    out = tmp_path / "g.bin"
    assert run("format_knk", f"-i{src}", f"-o{out}").returncode == 0
    host_ctx.load(str(src), ptw.FORMAT_TEXT)
    rt, ct = host_ctx.export_csr()
    host_ctx.load(str(out), ptw.FORMAT_BINARY)
    rb, cb = host_ctx.export_csr()
    assert np.array_equal(rt, rb) and np.array_equal(ct, cb)
    # compacted ids are the identity on reload
    assert edges_of(out).max() == len(rt) - 2


def test_format_yahoo_adjacency_records(tmp_path):
    a = tmp_path / "a.txt"
    a.write_text("1 2 10 11\n2 0\n3 1\n   12\n")
    b = tmp_path / "b.txt"
    b.write_text("4 3 1 2 3")
    binp, txtp = tmp_path / "y.bin", tmp_path / "y.txt"
    r = run("format_yahoo", "-i", str(a), "-i", str(b), "-b", str(binp), "-t", str(txtp))
    assert r.returncode == 0, r.stderr
    expect = [[1, 10], [1, 11], [3, 12], [4, 1], [4, 2], [4, 3]]
    assert edges_of(binp).tolist() == expect
    header = f"# Converted from: {a} {b}\n# Edges: 6\n"
    assert (tmp_path / "y.bin.info.txt").read_text() == header
    assert txtp.read_text() == header + "".join(f"{s} {d}\n" for s, d in expect)


def test_format_yahoo_truncated_record_fails(tmp_path):
    a = tmp_path / "a.txt"
    a.write_text("1 3 10 11\n")
    r = run("format_yahoo", "-i", str(a), "-b", str(tmp_path / "o.bin"))
    assert r.returncode == 1 and "Check failed" in r.stderr


@pytest.mark.parametrize("args", [[], ["-i", "x"], ["-o", "y"], ["-z"]])
def test_format_knk_bad_flags(args):
    r = run("format_knk", *args)
    assert r.returncode == 1


def test_format_yahoo_needs_an_output(tmp_path):
    a = tmp_path / "a.txt"
    a.write_text("1 1 2\n")
    assert run("format_yahoo", "-i", str(a)).returncode == 1
    assert run("format_yahoo", "-h").returncode == 0

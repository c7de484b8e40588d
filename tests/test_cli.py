"""The four CLI binaries keep the reference's flag surface and error exits
(include/option.hpp:21-59,86-109,132-355).  Only parse paths run here: a
walk needs the GPU (tests/test_gpu_walk.py runs the binaries end to end)."""
import os
import subprocess

import pytest

import ptwalker_amd as ptw

APPS = ["pagerank", "PPR", "RWR", "deepwalk"]


def run(app, *args):
    exe = os.path.join(ptw.BIN_DIR, app)
    assert os.path.exists(exe), f"{exe} not built"
    return subprocess.run([exe, *args], capture_output=True, text=True, timeout=60)


@pytest.mark.parametrize("app", APPS)
def test_help_exits_zero(app):
    r = run(app, "-h")
    assert r.returncode == 0
    assert "-g[graph]" in r.stdout and "-w[walker]" in r.stdout


@pytest.mark.parametrize("app", APPS)
def test_unknown_flag_exits_one(app):
    r = run(app, "-f", "text", "-g", "x", "-w", "10", "--nope", "1")
    assert r.returncode == 1
    assert "could not be matched" in r.stderr


@pytest.mark.parametrize("app", ["pagerank", "PPR"])
def test_sp_required_for_pt_apps(app):
    r = run(app, "-f", "text", "-g", "x", "-w", "10")
    assert r.returncode == 1 and "rastart_probability_flag" in r.stderr


@pytest.mark.parametrize("app", ["RWR", "deepwalk"])
def test_len_required_for_walk_apps(app):
    r = run(app, "-f", "text", "-g", "x", "-w", "10")
    assert r.returncode == 1 and "walk_len_flag" in r.stderr


@pytest.mark.parametrize("app", APPS)
def test_epoch_and_walkers_exclusive(app):
    r = run(app, "-f", "text", "-g", "x", "-w", "10", "-e", "2", "--sp", "0.2", "-l", "5") if app in (
        "pagerank", "PPR") else run(app, "-f", "text", "-g", "x", "-w", "10", "-e", "2", "-l", "5")
    assert r.returncode == 1 and "epoch_num_flag && walker_num_flag" in r.stderr
    r = run(app, "-f", "text", "-g", "x", "--sp", "0.2", "-l", "5") if app in ("pagerank", "PPR") else run(
        app, "-f", "text", "-g", "x", "-l", "5")
    assert r.returncode == 1 and "epoch_num_flag || walker_num_flag" in r.stderr


def test_bad_format_and_missing_graph():
    r = run("pagerank", "-f", "csv", "-g", "x", "-w", "10", "--sp", "0.2")
    assert r.returncode == 1 and "Unknown graph format: csv" in r.stderr
    r = run("pagerank", "-f", "text", "-w", "10", "--sp", "0.2")
    assert r.returncode == 1 and "graph_path_flag" in r.stderr
    r = run("pagerank", "-ftext", "-gx", "-w", "ten", "--sp=0.2")
    assert r.returncode == 1 and "invalid value" in r.stderr


def test_ws_db_only_for_rwpt_parser():
    r = run("deepwalk", "-f", "text", "-g", "x", "-w", "10", "-l", "5", "--ws", "4")
    assert r.returncode == 1

"""The drop-in command-line tools (index / search / prune / resave, the
reference's index.cpp, search.cpp, prune.cpp and resave.cpp): option
handling that needs no GPU."""
import os
import subprocess

import pytest

BIN = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "kgraph_amd", "bin")


def _run(tool, *args):
    path = os.path.join(BIN, tool)
    if not os.path.exists(path):
        pytest.skip("command-line tools not built (run __graft_entry__.build())")
    return subprocess.run([path] + list(args), capture_output=True, text=True, timeout=60)


@pytest.mark.parametrize("tool", ["index", "search"])
def test_version_and_help(tool):
    r = _run(tool, "--version")
    assert r.returncode == 0 and r.stdout.startswith("KGraph version kgraph_amd")
    r = _run(tool, "-h")
    assert r.returncode == 0
    assert ("Usage: index" in r.stdout) if tool == "index" else r.stdout.startswith("search <data> <index> <query>")


def test_missing_inputs_print_usage():
    # index needs an input (or --synthetic with --dim); search needs data, index, query
    assert "Usage: index" in _run("index", "-K", "10").stdout
    assert "Usage: index" in _run("index", "--synthetic", "100").stdout  # no --dim
    assert _run("search", "a", "b").stdout.startswith("search <data>")


def test_bad_options_fail_cleanly():
    r = _run("index", "--no-such-option", "x")
    assert r.returncode == 1 and "unknown option" in r.stderr
    r = _run("index", "-K", "ten", "x")
    assert r.returncode == 1 and "bad value" in r.stderr
    r = _run("index", "--precision", "fp16", "x")
    assert r.returncode == 1 and "precision" in r.stderr
    r = _run("search", "-K")
    assert r.returncode == 1 and "missing value" in r.stderr


def test_prune_and_resave_usage():
    # both print their usage and exit 0 without an input index (prune.cpp:70-75,
    # resave.cpp:69-72); unknown options fail
    r = _run("prune", "data-only")
    assert r.returncode == 0 and r.stdout.startswith("Usage: prune")
    assert "--prune arg (=0)" in r.stdout
    r = _run("resave")
    assert r.returncode == 0 and "--no-dist" in r.stdout
    r = _run("resave", "--bogus", "x")
    assert r.returncode == 1 and "unknown option" in r.stderr
    r = _run("prune", "--raw", "d", "i")
    assert r.returncode == 1 and "--dim" in r.stderr

"""The C++ drop-in interface (include/kgraph.h, kgraph-data.h): a program
written against the reference's API builds, saves, loads and searches on the
GPU, and its brute-force SearchOracle::search agrees with the walk."""
import json
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_cpp_facade_program(tmp_path):
    exe = os.path.join(ROOT, "tests", "facade", "facade_demo")
    assert os.path.exists(exe), "built by __graft_entry__.build()"
    out = subprocess.run([exe, str(tmp_path / "demo.kgraph")], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr
    r = json.loads(out.stdout.strip().splitlines()[-1])
    print(r)
    assert r["recall"] > 0.9
    assert r["same_after_load"] == r["queries"]
    assert r["dist_ok"] == 1 and r["get_nn_L"] >= r["get_nn_M"] > 0
    assert r["opaque_rejected"] == 1
    assert "gfx950" in r["version"]
    # concurrent callers, batched into shared launches, get what lone calls get
    assert r["concurrent_same"] == r["queries"]
    # rows changed in place are seen after kgraph::forget_device_rows()
    assert r["forget_ok"] == 1

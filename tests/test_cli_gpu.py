"""The drop-in command-line tools on the GPU: `index` builds and saves a
KNNGRAPH file from lshkit data, `search` answers lshkit queries from it with
the reference's output line, its --eval recall and --linear exact scan, and
its result ids equal the library's own search on the same index file."""
import os
import re
import subprocess

import numpy as np
import pytest

from oracle import kgo

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "kgraph_amd", "bin")


def _lshkit(path, a):
    a = np.ascontiguousarray(a)
    hdr = np.array([a.dtype.itemsize, a.shape[0], a.shape[1]], dtype=np.uint32)
    with open(path, "wb") as f:
        f.write(hdr.tobytes())
        f.write(a.tobytes())


def _read_lshkit(path, dtype):
    raw = np.fromfile(path, dtype=np.uint8)
    hdr = raw[:12].view(np.uint32)
    return raw[12:].view(dtype).reshape(int(hdr[1]), int(hdr[2]))


def _run(*args):
    r = subprocess.run(list(args), capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    return r


def test_index_then_search_with_eval(tmp_path):
    from kgraph_amd import datagen
    import kgraph_amd
    X = datagen.integer_mixture(20000, 32, 40, 3, 4)
    Q = datagen.integer_mixture(300, 32, 40, 3, 5)
    K = 10
    gi, _, _ = kgo.bruteforce(X, Q, K, 0)
    data, query, gs = str(tmp_path / "data.lshkit"), str(tmp_path / "query.lshkit"), str(tmp_path / "gs.lshkit")
    _lshkit(data, X)
    _lshkit(query, Q)
    _lshkit(gs, gi.astype(np.uint32))
    index = str(tmp_path / "data.kgraph")
    r = _run(os.path.join(BIN, "index"), data, index, "-K", "20", "-L", "40", "--reverse", "-1", "-C", "50")
    assert re.search(r"\d+\.\d+s wall", r.stderr)
    out = str(tmp_path / "result.lshkit")
    r = _run(os.path.join(BIN, "search"), data, index, query, out, "-K", str(K), "-P", "50", "--eval", gs)
    m = re.match(r"Time: (\S+) Recall: (\S+) Cost: (\S+)", r.stdout)
    assert m, r.stdout
    recall, cost = float(m.group(2)), float(m.group(3))
    assert recall > 0.9 and 0 < cost < 1
    res = _read_lshkit(out, np.uint32)
    assert res.shape == (len(Q), K)
    # the same ids as the library's search on the same index file
    g = kgraph_amd.KGraph(X, "euclidean", l2sqr=True)
    g.load(index)
    assert np.array_equal(res, g.search(Q, K=K, P=50))
    # and the oracle reads the file the tool wrote
    ref = kgo.Index()
    ref.load(index)
    assert len(ref.export()[3]) == len(X)
    # exact scan: recall 1, cost 1
    r = _run(os.path.join(BIN, "search"), data, index, query, "-K", str(K), "--linear", "--eval", gs)
    m = re.match(r"Time: (\S+) Recall: (\S+) Cost: (\S+)", r.stdout)
    assert float(m.group(2)) == 1.0 and float(m.group(3)) == 1.0


@pytest.mark.parametrize("level", [1, 2, 3])
def test_prune_tool_matches_oracle(tmp_path, level):
    # prune DATA INPUT OUTPUT --prune L: the GPU prune of a saved index, whose
    # file equals the oracle's prune of the same index, byte for byte
    from kgraph_amd import datagen
    X = datagen.uniform(3000, 8, 5)
    ref = kgo.Index()
    ref.build(X, 0, K=10, L=24, iterations=8, controls=10, reverse=-1)
    data, src, out, want = (str(tmp_path / n) for n in ("d.lshkit", "in.kgraph", "out.kgraph", "want.kgraph"))
    _lshkit(data, X)
    ref.save(src)
    r = _run(os.path.join(BIN, "prune"), data, src, out, "--prune", str(level))
    assert re.search(r"\d+\.\d+s wall", r.stdout)
    ref.prune(level)
    ref.save(want)
    assert open(out, "rb").read() == open(want, "rb").read()


def test_resave_tool_no_dist_and_prune(tmp_path):
    from kgraph_amd import datagen
    X = datagen.uniform(2000, 8, 6)
    ref = kgo.Index()
    ref.build(X, 0, K=10, L=20, iterations=8, controls=10, reverse=-1)
    src, out, want = (str(tmp_path / n) for n in ("in.kgraph", "out.kgraph", "want.kgraph"))
    ref.save(src)
    _run(os.path.join(BIN, "resave"), src, out, "--no-dist")
    ref.save(want, 1)
    assert open(out, "rb").read() == open(want, "rb").read()
    _run(os.path.join(BIN, "resave"), src, out, "--prune", "1")
    ref.prune(1)
    ref.save(want)
    assert open(out, "rb").read() == open(want, "rb").read()

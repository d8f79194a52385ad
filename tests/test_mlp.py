"""The learned projector: the FP64 oracle is pinned by finite differences (CPU);
the HIP forward / backward kernels match it (GPU); training lowers the loss;
projected entry selection keeps recall, and a degraded projector is detected
and replaced by exact entry scoring."""
import numpy as np
import pytest

from oracle import kgo, mlp_oracle


def _params(D, H, P, seed):
    rs = np.random.RandomState(seed)
    W1 = rs.uniform(-0.2, 0.2, (H, D)).astype(np.float32)
    b1 = rs.uniform(-0.1, 0.1, H).astype(np.float32)
    W2 = rs.uniform(-0.2, 0.2, (P, H)).astype(np.float32)
    b2 = rs.uniform(-0.1, 0.1, P).astype(np.float32)
    return W1, b1, W2, b2


def test_oracle_gradient_matches_finite_differences():
    D, H, P, B = 12, 16, 8, 10
    rs = np.random.RandomState(0)
    X = rs.standard_normal((3 * B, D))
    W1, b1, W2, b2 = [p.astype(np.float64) for p in _params(D, H, P, 1)]
    margin = 0.5
    loss, dW1, db1, dW2, db2 = mlp_oracle.loss_grad(X, W1, b1, W2, b2, margin, bf16=False)
    assert loss > 0
    eps = 1e-6
    for name, w, g in [("W1", W1, dW1), ("b1", b1, db1), ("W2", W2, dW2), ("b2", b2, db2)]:
        flat = w.reshape(-1)
        for i in rs.choice(flat.size, size=min(12, flat.size), replace=False):
            old = flat[i]
            flat[i] = old + eps
            lp = mlp_oracle.loss_grad(X, W1, b1, W2, b2, margin, bf16=False)[0]
            flat[i] = old - eps
            lm = mlp_oracle.loss_grad(X, W1, b1, W2, b2, margin, bf16=False)[0]
            flat[i] = old
            num = (lp - lm) / (2 * eps)
            assert abs(num - g.reshape(-1)[i]) <= 1e-6 + 1e-5 * abs(num), name


def test_bf16_rounding_is_round_to_nearest_even():
    x = np.array([1.0, 1.00390625, 1.0078125, 1.01171875, -3.0, 255.0, 256.0, 257.0], np.float32)
    # bf16 keeps 8 significant bits: 1 + 2^-8 ties to 1, 1 + 3*2^-8 ties up to 1 + 2^-6,
    # integers up to 256 are exact, 257 ties to even (256)
    exp = np.array([1.0, 1.0, 1.0078125, 1.015625, -3.0, 255.0, 256.0, 256.0])
    assert np.array_equal(mlp_oracle.to_bf16(x), exp)


@pytest.mark.gpu
def test_gpu_forward_backward_match_oracle():
    import kgraph_amd
    from kgraph_amd import _lib
    from kgraph_amd import datagen
    X = datagen.integer_mixture(3000, 128, 20, 1, 2) / 64.0
    g = kgraph_amd.KGraph(X, "euclidean", l2sqr=True)
    H, P = 256, 32
    g._h.call("kg_mlp_init", H, P, 7)
    # the input standardisation is the dataset's column mean and 1/std
    mu = np.zeros(128, np.float32)
    isd = np.zeros(128, np.float32)
    g._h.call("kg_mlp_get_normalization", _lib.ptr(mu), _lib.ptr(isd))
    X64 = X.astype(np.float64)
    sd = X64.std(0)
    assert np.allclose(mu, X64.mean(0), rtol=1e-6, atol=1e-6)
    assert np.allclose(isd, np.where(sd > 0, 1 / np.where(sd > 0, sd, 1), 1), rtol=1e-5)
    W1, b1, W2, b2 = _params(128, H, P, 3)
    g._h.call("kg_mlp_set_weights", _lib.ptr(W1), _lib.ptr(b1), _lib.ptr(W2), _lib.ptr(b2))
    rs = np.random.RandomState(4)
    nt = 100
    trip = rs.randint(0, len(X), size=3 * nt).astype(np.uint32)
    margin = 1.0
    loss = np.zeros(1, np.float32)
    gW1 = np.zeros((H, 128), np.float32)
    gb1 = np.zeros(H, np.float32)
    gW2 = np.zeros((P, H), np.float32)
    gb2 = np.zeros(P, np.float32)
    g._h.call("kg_mlp_loss_grad", _lib.ptr(trip), nt, margin, _lib.ptr(loss), _lib.ptr(gW1), _lib.ptr(gb1),
              _lib.ptr(gW2), _lib.ptr(gb2))
    Xt = (X[trip].astype(np.float32) - mu) * isd  # the same FP32 operations as the gather
    ol, oW1, ob1, oW2, ob2 = mlp_oracle.loss_grad(Xt, W1, b1, W2, b2, margin, bf16=True)
    assert abs(loss[0] - ol) <= 1e-3 * abs(ol) + 1e-6
    for name, a, b in [("dW1", gW1, oW1), ("db1", gb1, ob1), ("dW2", gW2, oW2)]:
        err = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-12)
        assert err < 2e-3, (name, err)
    # each triplet's output gradients s(zn-zp) - s(za-zp) + s(za-zn) sum to zero,
    # so db2 is zero up to FP32 rounding
    assert np.abs(ob2).max() < 1e-6 and np.abs(gb2).max() < 1e-5
    # forward projection
    Z = np.zeros((50, P), np.float32)
    Xq = np.ascontiguousarray(X[:50], dtype=np.float32)
    g._h.call("kg_mlp_project", _lib.ptr(Xq), 50, 128, _lib.ptr(Z))
    _, oZ = mlp_oracle.forward((Xq - mu) * isd, W1, b1, W2, b2, bf16=True)
    assert np.linalg.norm(Z - oZ) / np.linalg.norm(oZ) < 1e-3


def _clustered(n, nq, d, centers, seed):
    from kgraph_amd import datagen
    X = datagen.integer_mixture(n, d, centers, seed, seed + 1, sigma=6.0, center_hi=250.0)
    Q = datagen.integer_mixture(nq, d, centers, seed, seed + 2, sigma=6.0, center_hi=250.0)
    return X, Q


@pytest.mark.gpu
def test_training_lowers_loss_and_learned_entries_keep_recall():
    import kgraph_amd
    X, Q = _clustered(20000, 300, 64, 200, 5)
    g = kgraph_amd.KGraph(X, "euclidean", l2sqr=True)
    g.build(K=20, L=40, iterations=12, controls=50, reverse=-1)
    trace = g.train_projector(steps=300, batch=1024, hidden=256, out=32, lr=2e-3, margin=1.0, seed=1)
    # most triplets are inactive from the start (the neighbour is already much
    # closer than the random node), so per-step losses are small and noisy:
    # compare window means
    print("loss first 10 %.4f last 50 %.4f" % (trace[:10].mean(), trace[-50:].mean()))
    assert np.isfinite(trace).all()
    assert trace[-50:].mean() < 0.8 * trace[:10].mean()
    g.set_entry_sample(2000, seed=3)
    _, gt_d, _ = kgo.bruteforce(X, Q, 10, 0)
    _, d_exact = g.search(Q, K=10, P=20, entries=9, withDistance=True)
    g.set_entry_projection(16)
    _, d_learn = g.search(Q, K=10, P=20, entries=9, withDistance=True)
    r_exact = kgo.recall_at_k(d_exact, gt_d, 10)
    r_learn = kgo.recall_at_k(d_learn, gt_d, 10)
    print("entry recall exact %.4f learned %.4f" % (r_exact, r_learn))
    assert r_learn >= r_exact - 0.01
    # probe-set guard: a degraded (zero) projector is caught and exact scoring comes back
    rep = g.validate_entry_projection(Q[:100], K=10, P=20, entries=9)
    assert rep["enabled"]
    H, P = 256, 32
    from kgraph_amd import _lib
    z = [np.zeros((H, 64), np.float32), np.zeros(H, np.float32), np.zeros((P, H), np.float32),
         np.zeros(P, np.float32)]
    g._h.call("kg_mlp_set_weights", *[_lib.ptr(a) for a in z])
    rep = g.validate_entry_projection(Q[:100], K=10, P=20, entries=9)
    print(rep)
    assert not rep["enabled"] and rep["recall_learned"] < rep["recall_exact"]


@pytest.mark.gpu
def test_join_prerank_trained_untrained_zeroed():
    """Learned pre-rank of the local-join lists, against plain NN-descent at
    equal recall: the trained projector ranks held-out exact-neighbour pairs
    against random pairs better than the untrained one (the same
    architecture at its random init), every pre-ranked build ends within a
    point of plain NN-descent's recall (the in-build guard turns the pre-rank
    off when it slows the control recall down), and a zeroed projector
    fails the held-out probe, so that build runs plain NN-descent throughout."""
    import kgraph_amd
    from kgraph_amd import _lib, datagen
    n, D = 50000, 96
    X = datagen.angular_mixture(n, D, 100, 6, 7, sigma=0.3)
    g = kgraph_amd.KGraph(X, "angular")
    bp = dict(K=25, L=100, iterations=30, reverse=0, precision="bf16", controls=200)
    rows = np.arange(0, n, 50)
    gi, gd = kgo.bruteforce(X, X[rows], 11, 2)[:2]
    gt = np.array([[d for i, d in zip(ii, dd) if i != r][:10] for r, ii, dd in zip(rows, gi, gd)], np.float32)

    def build(keep):
        g.build(prerank=keep, **bp)
        off, _, dd, _ = g.graph()
        return kgo.recall_at_k([dd[off[r]:off[r] + 10] for r in rows], gt, 10), g.prerank_report()

    r_plain, rep = build(0)
    assert rep["status"] == "off"
    g.train_projector(steps=300, batch=4096, seed=1)
    r_trained, rep_t = build(0.75)
    H, P = 256, 32
    g._h.call("kg_mlp_init", H, P, 1)
    r_untrained, rep_u = build(0.75)
    z = [np.zeros((H, D), np.float32), np.zeros(H, np.float32), np.zeros((P, H), np.float32),
         np.zeros(P, np.float32)]
    g._h.call("kg_mlp_set_weights", *[_lib.ptr(a) for a in z])
    r_zero, rep_z = build(0.75)
    print("plain %.4f trained %.4f %s untrained %.4f %s zeroed %.4f %s" %
          (r_plain, r_trained, rep_t, r_untrained, rep_u, r_zero, rep_z))
    assert rep_t["probe_accuracy"] >= rep_u["probe_accuracy"] + 0.02
    assert rep_t["status"] != "probe-fallback"
    for r in (r_trained, r_untrained, r_zero):
        assert r >= r_plain - 0.01
    assert rep_z["status"] == "probe-fallback" and rep_z["probe_accuracy"] < 0.5

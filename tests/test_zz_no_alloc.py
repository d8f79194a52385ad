"""The scan that allocates nothing (kg_bruteforce_device_noalloc): through
the C ABI, on device-resident queries and results, it returns exactly what
kg_bruteforce returns, on two calls in a row (the second smaller, so stale
scratch from the first would show), with and without split partial lists,
and it rejects calls past its limit without touching the GPU."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _host_scan(h, Q, K):
    from kgraph_amd._lib import ptr
    nq = len(Q)
    ids = np.zeros((nq, K), np.uint32)
    d = np.zeros((nq, K), np.float32)
    c = np.zeros(nq, np.uint32)
    h.call("kg_bruteforce", ptr(Q), nq, Q.shape[1], K, 1e30, ptr(ids), ptr(d), K, ptr(c))
    return ids, d, c


def _noalloc_scan(h, Q, K, qstride):
    from kgraph_amd import devmem
    nq = len(Q)
    Qp = np.zeros((nq, qstride), np.float32)
    Qp[:, :Q.shape[1]] = Q
    dQ = devmem.DeviceArray.from_host(Qp)
    dI = devmem.DeviceArray((nq, K), np.uint32)
    dD = devmem.DeviceArray((nq, K), np.float32)
    dC = devmem.DeviceArray((nq,), np.uint32)
    h.call("kg_bruteforce_device_noalloc", dQ.ptr, nq, qstride, K, 1e30, dI.ptr, dD.ptr, K, dC.ptr)
    return dI.to_host(), dD.to_host(), dC.to_host()


@pytest.mark.parametrize("metric,n,d,calls", [
    (0, 40000, 37, [(300, 10), (40, 7)]),     # split partial lists (N >= 32768, few queries)
    (1, 6000, 16, [(2500, 25), (100, 25)]),   # one block per query
    (2, 50000, 96, [(64, 100), (5, 3)]),
])
def test_noalloc_scan_equals_bruteforce(metric, n, d, calls):
    from kgraph_amd import _lib, datagen
    X = datagen.uniform(n, d, 3)
    h = _lib.Handle()
    h.call("kg_set_data", _lib.ptr(X), n, d, d, metric)
    qstride = (d + 3) & ~3
    for nq, K in calls:
        Q = datagen.uniform(nq, d, 100 + nq)
        a = _host_scan(h, Q, K)
        b = _noalloc_scan(h, Q, K, qstride)
        bad = np.nonzero((a[0] != b[0]).any(1) | (a[1] != b[1]).any(1) | (a[2] != b[2]))[0]
        for r in bad[:4]:
            print("query %d: bruteforce ids %s dists %s count %d; no-allocation ids %s dists %s count %d" % (
                r, a[0][r][:6], a[1][r][:6], a[2][r], b[0][r][:6], b[1][r][:6], b[2][r]))
        assert len(bad) == 0, "%d of %d queries differ" % (len(bad), nq)
        assert (a[2] == min(K, n)).all()


def test_noalloc_scan_rejects_what_it_cannot_serve():
    from kgraph_amd import _lib, datagen, devmem
    X = datagen.uniform(1000, 8, 4)
    h = _lib.Handle()
    h.call("kg_set_data", _lib.ptr(X), 1000, 8, 8, 0)
    dQ = devmem.DeviceArray((4, 8), np.float32)
    dI = devmem.DeviceArray((4, 10), np.uint32)
    dC = devmem.DeviceArray((4,), np.uint32)
    with pytest.raises(ValueError, match="limit"):
        h.call("kg_bruteforce_device_noalloc", dQ.ptr, (1 << 20) + 1, 8, 10, 1e30, dI.ptr, None, 10, dC.ptr)
    with pytest.raises(ValueError, match="multiple of 4"):
        h.call("kg_bruteforce_device_noalloc", dQ.ptr, 4, 9, 10, 1e30, dI.ptr, None, 10, dC.ptr)
    # a valid call still works after the rejected ones
    Q = datagen.uniform(4, 8, 9)
    a = _host_scan(h, Q, 10)
    b = _noalloc_scan(h, Q, 10, 8)
    assert all(np.array_equal(x, y) for x, y in zip(a, b))

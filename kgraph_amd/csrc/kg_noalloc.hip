// kg_noalloc.hip -- exact top-K by full scan that allocates nothing, for
// callers that cannot call hipMalloc (code under HIP graph capture, tight
// serving loops, harnesses that forbid allocation).
//
// The same kernels as kg_bruteforce (brute_kernel / brute_merge_kernel in
// kg_brute.hip, norms_kernel for the query norms), with their scratch -- the
// query norms and the per-split partial top-K lists -- in one file-scope
// device array, `scratch`, instead of buffers the library allocates.  Each GPU
// holds its own copy of the array from the library's load to the end of the
// process.  Its device address is taken once per device, with that device
// current, by kg_noalloc_setup (called when a handle is created on it); host
// code passes that address, never the array's name, to the kernels.
//
// Every call on one GPU shares that GPU's array and overwrites it: calls that
// overlap on one GPU (two handles' streams at once) race.  Nothing in the
// array needs to start at zero: each block writes every partial slot it owns
// (unused ones as all-ones keys) and the norm kernel writes every query's norm
// before the scan reads it.
#include <mutex>

#include "kg_device.hpp"
#include "kg_handle.hpp"
#include "kg_internal.hpp"
#include "kg_noalloc.hpp"

namespace {

// [0, NOALLOC_NORM_BYTES): query norms (one float per query);
// [NOALLOC_NORM_BYTES, NOALLOC_SCRATCH_BYTES): partial top-K keys.
static __device__ __align__(256) unsigned char scratch[kga::NOALLOC_SCRATCH_BYTES];

constexpr int MAX_DEVICES = 64;
std::mutex setup_mu;
void *scratch_ptr[MAX_DEVICES] = {};

}  // namespace

namespace kga {

void noalloc_setup(int device) {
    if (device < 0 || device >= MAX_DEVICES) throw runtime_error("device ordinal out of range");
    std::lock_guard<std::mutex> g(setup_mu);
    if (scratch_ptr[device]) return;
    int cur = -1;
    KG_HIP(hipGetDevice(&cur));
    if (cur != device) KG_HIP(hipSetDevice(device));
    void *p = nullptr;
    hipError_t const e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != device) (void)hipSetDevice(cur);
    KG_HIP(e);
    if (!p) throw device_error("hipGetSymbolAddress returned no address for the scratch array");
    scratch_ptr[device] = p;
}

void bruteforce_noalloc(Handle &h, float const *dQ, unsigned nq, size_t qstride, unsigned K, float epsilon,
                        uint32_t *d_ids, float *d_dists, size_t ostride, uint32_t *d_counts) {
    if (!h.has_data) throw runtime_error("no dataset");
    if (!dQ || !d_ids || !d_counts) throw invalid_argument("null argument");
    if (K == 0 || K > 1024 || K > ostride) throw invalid_argument("K must be in [1, min(1024, ostride)]");
    if (qstride < h.D || (qstride & 3) || (reinterpret_cast<uintptr_t>(dQ) & 15))
        throw invalid_argument("queries must be 16-byte aligned rows of a stride that is a multiple of 4 floats");
    if (nq > NOALLOC_MAX_QUERIES)
        throw invalid_argument("nq exceeds the no-allocation scan's limit of " +
                               std::to_string(NOALLOC_MAX_QUERIES) + " queries per call");
    if (nq == 0) return;
    unsigned const splits = kg_brute_splits(nq, h.N, K);
    size_t const part_bytes = splits > 1 ? size_t(nq) * splits * K * 8 : 0;
    if (part_bytes > NOALLOC_SCRATCH_BYTES - NOALLOC_NORM_BYTES)
        throw invalid_argument("partial lists exceed the no-allocation scratch area");
    if (h.device < 0 || h.device >= MAX_DEVICES || !scratch_ptr[h.device])
        throw runtime_error("no-allocation scratch not set up on this device");
    unsigned char *const base = static_cast<unsigned char *>(scratch_ptr[h.device]);
    float *const qn = reinterpret_cast<float *>(base);
    KG_HIP(kg_launch_norms(h.metric, dQ, nq, h.D, unsigned(qstride), qn, h.stream));
    kg_brute_args b{};
    b.X = h.X.p;
    b.norms = h.norms.p;
    b.N = h.N;
    b.D = h.D;
    b.Dp = h.Dp;
    b.Q = dQ;
    b.qnorms = qn;
    b.qstride = unsigned(qstride);
    b.nq = nq;
    b.K = K;
    b.epsilon = epsilon;
    b.out_ids = d_ids;
    b.out_dists = d_dists;
    b.ostride = unsigned(ostride);
    b.counts = d_counts;
    b.partial = reinterpret_cast<uint64_t *>(base + NOALLOC_NORM_BYTES);
    KG_HIP(hipEventRecord(h.ev[2], h.stream));
    KG_HIP(kg_launch_brute(h.metric, b, h.stream));
    KG_HIP(hipEventRecord(h.ev[3], h.stream));
    h.sync();
    KG_HIP(hipEventElapsedTime(&h.last_ms[1], h.ev[2], h.ev[3]));
}

}  // namespace kga

// kg_brute.hip -- exact top-K by full scan, and per-row norms.
//
// brute_kernel restates SearchOracle::search (kgraph.cpp:231-252: top-K with
// an epsilon filter) and LinearSearch (kgraph.cpp:212-229: K-NN of a data row
// with the row itself excluded).  Both insert in ascending id order after
// equal distances, i.e. they return the K smallest (dist, id) pairs; this
// kernel selects exactly that set with u64 keys (orderable dist bits << 32 |
// id).  One block per query: 4 waves score 64 rows per step (4 lanes per row,
// canonical distance order), keys below the running K-th key go to an LDS
// buffer, and a block-wide bitonic sort folds the buffer into the top-K.
//
// norms_kernel computes the per-row norm the metric keeps
// (python-api.cpp:110-114): ||x||^2/2 for 'euclidean', ||x|| for 'angular'.
//
// bf16_rows_kernel writes the BF16 copy of the rows (round to nearest even)
// that the MFMA kernels take their operands from, and flags the dataset when
// any value is not representable in BF16; when none is, the copy is lossless
// and graph walks gather it instead of the FP32 rows (half the bytes, the same
// distances).
#include "kg_device.hpp"
#include "kg_internal.hpp"

namespace {

using namespace kgd;

constexpr unsigned BT = 256;     // threads per block
constexpr unsigned BUF = 1024;   // candidate buffer
constexpr unsigned SORTN = 2048; // sort width: K (<= 1024) + BUF

__device__ void block_bitonic(uint64_t *v, unsigned n) {
    for (unsigned k = 2; k <= n; k <<= 1) {
        for (unsigned j = k >> 1; j > 0; j >>= 1) {
            for (unsigned i = threadIdx.x; i < n; i += BT) {
                unsigned ixj = i ^ j;
                if (ixj > i) {
                    uint64_t a = v[i], b = v[ixj];
                    bool up = (i & k) == 0;
                    if ((a > b) == up) {
                        v[i] = b;
                        v[ixj] = a;
                    }
                }
            }
            __syncthreads();
        }
    }
}

// Rows [r0, r1) of split blockIdx.y.  With a.partial the block writes its
// sorted top-K keys (~0 padded) to partial[(q * splits + split) * K]; else the
// final ids / dists / counts.
template <int METRIC>
__global__ void __launch_bounds__(BT) brute_kernel(kg_brute_args a) {
    __shared__ uint64_t keys[SORTN];  // [0,K) top-K, [K, K+cnt) buffer
    __shared__ __align__(16) float qv[1024];
    __shared__ unsigned cnt;
    __shared__ unsigned long long thr;
    unsigned const qi = a.qsel ? a.qsel[blockIdx.x] : blockIdx.x;
    unsigned const K = a.K;
    unsigned const lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    unsigned const grp = lane >> 2, t = lane & 3;
    float const *qsrc;
    float qn = 0.f;
    uint32_t self = 0xffffffffu;
    if (a.Q) {
        qsrc = a.Q + size_t(qi) * a.qstride;
        if (METRIC != METRIC_L2SQR) qn = a.qnorms[qi];
    } else {
        self = a.rows[qi];
        qsrc = a.X + size_t(self) * a.Dp;
        if (METRIC != METRIC_L2SQR) qn = a.norms[self];
    }
    for (unsigned i = threadIdx.x; i < a.Dp; i += BT) qv[i] = (i < a.D) ? qsrc[i] : 0.f;
    for (unsigned i = threadIdx.x; i < SORTN; i += BT) keys[i] = ~0ull;
    if (threadIdx.x == 0) {
        cnt = 0;
        thr = ~0ull;
    }
    __syncthreads();
    unsigned have = 0;  // valid entries in the top-K part
    unsigned const span = (a.N + gridDim.y - 1) / gridDim.y;
    unsigned const rb = blockIdx.y * span, re = min(a.N, rb + span);
    for (unsigned r0 = rb; r0 < re; r0 += 64) {
        unsigned row = r0 + wv * 16 + grp;
        if (r0 + wv * 16 + grp < re) {
            float acc = quad_metric<METRIC>(a.X + size_t(row) * a.Dp, qv, a.D, t);
            float na = (METRIC == METRIC_L2SQR) ? 0.f : a.norms[row];
            float d = finish_metric<METRIC>(acc, na, qn);
            uint64_t key = (uint64_t(ord_f32(d)) << 32) | row;
            if (t == 0 && row != self && d <= a.epsilon && key < thr) {
                unsigned p = atomicAdd(&cnt, 1u);
                keys[K + p] = key;
            }
        }
        __syncthreads();
        // every thread reads the count before any thread of the next step can
        // add to it: a thread that read it late would see the next step's
        // candidates, take the other branch and miss the sort's barriers
        unsigned const c = cnt;
        __syncthreads();
        if (c + 64 > BUF || r0 + 64 >= re) {
            unsigned n = K + c;
            unsigned sn = 2;
            while (sn < n) sn <<= 1;
            block_bitonic(keys, sn);
            have = min(K, have + c);
            for (unsigned i = K + threadIdx.x; i < SORTN; i += BT) keys[i] = ~0ull;
            __syncthreads();
            if (threadIdx.x == 0) {
                cnt = 0;
                if (have == K) thr = keys[K - 1];
            }
            __syncthreads();
        }
    }
    if (a.partial) {
        uint64_t *dst = a.partial + (size_t(blockIdx.x) * gridDim.y + blockIdx.y) * K;
        for (unsigned i = threadIdx.x; i < K; i += BT) dst[i] = (i < have) ? keys[i] : ~0ull;
        return;
    }
    for (unsigned i = threadIdx.x; i < have; i += BT) {
        uint64_t key = keys[i];
        a.out_ids[size_t(qi) * a.ostride + i] = uint32_t(key);
        if (a.out_dists) a.out_dists[size_t(qi) * a.ostride + i] = unord_f32(uint32_t(key >> 32));
    }
    if (threadIdx.x == 0) {
        if (a.counts) a.counts[qi] = have;
        if (a.comps) a.comps[qi] = a.N;
        if (a.scans) a.scans[qi] = 0;
    }
}

// Fold the splits' partial top-K lists of one query into the final result.
__global__ void __launch_bounds__(BT) brute_merge_kernel(kg_brute_args a, unsigned splits) {
    __shared__ uint64_t keys[4096];
    unsigned const qi = a.qsel ? a.qsel[blockIdx.x] : blockIdx.x;
    unsigned const n = splits * a.K;
    unsigned p2 = 1;
    while (p2 < n) p2 <<= 1;
    uint64_t const *src = a.partial + size_t(blockIdx.x) * n;
    for (unsigned i = threadIdx.x; i < p2; i += BT) keys[i] = (i < n) ? src[i] : ~0ull;
    __syncthreads();
    block_bitonic(keys, p2);
    unsigned have = 0;
    for (unsigned i = 0; i < a.K; ++i) have += keys[i] != ~0ull;
    for (unsigned i = threadIdx.x; i < have; i += BT) {
        a.out_ids[size_t(qi) * a.ostride + i] = uint32_t(keys[i]);
        if (a.out_dists) a.out_dists[size_t(qi) * a.ostride + i] = unord_f32(uint32_t(keys[i] >> 32));
    }
    if (threadIdx.x == 0) {
        if (a.counts) a.counts[qi] = have;
        if (a.comps) a.comps[qi] = a.N;
        if (a.scans) a.scans[qi] = 0;
    }
}

template <int METRIC>
__global__ void __launch_bounds__(256) norms_kernel(float const *X, unsigned N, unsigned D, unsigned Dp, float *out) {
    unsigned const lane = threadIdx.x & 63;
    unsigned const row = (blockIdx.x * blockDim.x + threadIdx.x) >> 2;
    unsigned const t = lane & 3;
    if ((blockIdx.x * blockDim.x + (threadIdx.x & ~3u)) / 4 < N) {
        float const *r = X + size_t(row) * Dp;
        float n2 = quad_metric<METRIC_EUCLIDEAN, float, true>(r, r, D, t);
        if (t == 0) out[row] = finish_norm<METRIC>(n2);
    }
}

__global__ void __launch_bounds__(256) u8_rows_kernel(float4 const *X, size_t n4, uint32_t *X8, uint32_t *inexact) {
    bool lossy = false;
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n4; i += size_t(gridDim.x) * blockDim.x) {
        float4 const v = X[i];
        float const e[4] = {v.x, v.y, v.z, v.w};
        uint32_t w = 0;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            bool const in = e[j] >= 0.f && e[j] <= 255.f;
            uint32_t const u = in ? uint32_t(e[j]) : 0u;
            lossy |= !in || __float_as_uint(float(u)) != __float_as_uint(e[j]);
            w |= u << (8 * j);
        }
        X8[i] = w;
    }
    if (lossy) *inexact = 1u;  // every writer stores the same value
}

__global__ void __launch_bounds__(256) bf16_rows_kernel(float4 const *X, size_t n4, uint2 *Xb, uint32_t *inexact) {
    bool lossy = false;
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n4; i += size_t(gridDim.x) * blockDim.x) {
        float4 const v = X[i];
        uint32_t const b0 = bf16_bits(v.x), b1 = bf16_bits(v.y), b2 = bf16_bits(v.z), b3 = bf16_bits(v.w);
        Xb[i] = make_uint2(b0 | (b1 << 16), b2 | (b3 << 16));
        lossy |= (b0 << 16) != __float_as_uint(v.x) || (b1 << 16) != __float_as_uint(v.y) ||
                 (b2 << 16) != __float_as_uint(v.z) || (b3 << 16) != __float_as_uint(v.w);
    }
    if (lossy) *inexact = 1u;  // every writer stores the same value
}

}  // namespace

unsigned kg_brute_splits(unsigned nq, unsigned N, unsigned K) {
    // enough blocks to cover the GPU, each split at least 16k rows, splits*K <= 4096
    unsigned sp = 1;
    while (sp * 2 <= 64 && nq * sp < 2048 && N / (sp * 2) >= 16384 && sp * 2 * K <= 4096) sp *= 2;
    return sp;
}

hipError_t kg_launch_brute(int metric, kg_brute_args const &a0, hipStream_t s) {
    if (a0.K == 0 || a0.K > 1024 || a0.Dp > 1024 || a0.nq == 0) return hipErrorInvalidValue;
    unsigned const splits = a0.partial ? kg_brute_splits(a0.nq, a0.N, a0.K) : 1;
    kg_brute_args a = a0;
    if (splits == 1) a.partial = nullptr;
    dim3 grid(a.nq, splits), block(BT);
    switch (metric) {
        case METRIC_L2SQR: hipLaunchKernelGGL(brute_kernel<METRIC_L2SQR>, grid, block, 0, s, a); break;
        case METRIC_EUCLIDEAN: hipLaunchKernelGGL(brute_kernel<METRIC_EUCLIDEAN>, grid, block, 0, s, a); break;
        default: hipLaunchKernelGGL(brute_kernel<METRIC_ANGULAR>, grid, block, 0, s, a); break;
    }
    if (splits > 1) hipLaunchKernelGGL(brute_merge_kernel, dim3(a.nq), block, 0, s, a, splits);
    return hipGetLastError();
}

hipError_t kg_launch_norms(int metric, float const *X, unsigned N, unsigned D, unsigned Dp, float *out, hipStream_t s) {
    if (N == 0) return hipSuccess;
    dim3 grid((N * 4 + 255) / 256), block(256);
    switch (metric) {
        case METRIC_EUCLIDEAN: hipLaunchKernelGGL(norms_kernel<METRIC_EUCLIDEAN>, grid, block, 0, s, X, N, D, Dp, out); break;
        case METRIC_ANGULAR: hipLaunchKernelGGL(norms_kernel<METRIC_ANGULAR>, grid, block, 0, s, X, N, D, Dp, out); break;
        default: return hipSuccess;
    }
    return hipGetLastError();
}

hipError_t kg_launch_u8_rows(float const *X, unsigned N, unsigned Dp, uint8_t *X8, uint32_t *inexact, hipStream_t s) {
    size_t const n4 = size_t(N) * Dp / 4;
    if (n4 == 0) return hipSuccess;
    unsigned const nb = unsigned(std::min<size_t>(4096, (n4 + 255) / 256));
    hipLaunchKernelGGL(u8_rows_kernel, dim3(nb), dim3(256), 0, s, reinterpret_cast<float4 const *>(X), n4,
                       reinterpret_cast<uint32_t *>(X8), inexact);
    return hipGetLastError();
}

hipError_t kg_launch_bf16_rows(float const *X, unsigned N, unsigned Dp, uint16_t *Xb, uint32_t *inexact,
                               hipStream_t s) {
    size_t const n4 = size_t(N) * Dp / 4;
    if (n4 == 0) return hipSuccess;
    unsigned const blocks = unsigned(std::min<size_t>((n4 + 255) / 256, 8192));
    hipLaunchKernelGGL(bf16_rows_kernel, dim3(blocks), dim3(256), 0, s, reinterpret_cast<float4 const *>(X), n4,
                       reinterpret_cast<uint2 *>(Xb), inexact);
    return hipGetLastError();
}

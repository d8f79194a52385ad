// kg_entry.hip -- entry-point selection for graph-walk search.
//
// The reference draws its starting points at random (kgraph.cpp:449-460) or
// takes them from the caller (SearchParams::init, kgraph.cpp:461-467).  On
// clustered data whose k-NN graph has no edges between clusters, a walk from
// random starts only succeeds when a start lands in the query's cluster.
// These kernels pick, for every query, the m nearest nodes of a fixed sample
// of E nodes (exact canonical FP32 distances, ties by node id); the result is
// handed to the unchanged walk as its `init` starting points.
//
//   entry_dist_kernel  16 queries x 64 sampled rows per block: each quad holds
//                      one sampled row, the queries sit in LDS
//   entry_topm_kernel  one wave per query: per-lane sorted top-m lists in
//                      registers, then m rounds of a wave-wide minimum
#include "kg_device.hpp"
#include "kg_internal.hpp"

namespace {

using namespace kgd;

constexpr unsigned QB = 16;    // queries per block
constexpr unsigned EB = 64;    // sampled rows per block (one per quad)
constexpr unsigned MAXM = 16;  // largest m

template <int METRIC>
__global__ void __launch_bounds__(256) entry_dist_kernel(kg_entry_args a) {
    extern __shared__ __align__(16) float qs[];  // QB x Dp
    unsigned const q0 = blockIdx.x * QB, e0 = blockIdx.y * EB;
    unsigned const nqb = min(QB, a.nq - q0);
    for (unsigned i = threadIdx.x; i < nqb * a.Dp; i += blockDim.x) {
        unsigned r = i / a.Dp, c = i - r * a.Dp;
        qs[i] = (c < a.D) ? a.Q[size_t(q0 + r) * a.qstride + c] : 0.f;
    }
    __syncthreads();
    unsigned const quad = threadIdx.x >> 2, t = threadIdx.x & 3;
    unsigned const e = e0 + quad;
    if (e >= a.E) return;  // whole quads leave together
    float const *row = a.rows + size_t(e) * a.Dp;
    float const ne = (METRIC == METRIC_L2SQR) ? 0.f : a.enorms[e];
    unsigned const nch = (a.D & ~15u) >> 2;
    if (nch <= 4 * 8 && a.D == (a.D & ~15u)) {
        // the sampled row stays in registers (up to 8 float4 per lane)
        float4 reg[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            unsigned kk = t + 4 * k;
            reg[k] = (kk < nch) ? reinterpret_cast<float4 const *>(row)[kk] : make_float4(0, 0, 0, 0);
        }
        for (unsigned r = 0; r < nqb; ++r) {
            float4 const *q4 = reinterpret_cast<float4 const *>(qs + size_t(r) * a.Dp);
            float4 p = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                unsigned kk = t + 4 * k;
                if (kk < nch) {
                    float4 x = reg[k], y = q4[kk];
                    if (METRIC == METRIC_L2SQR) {
                        float dx = x.x - y.x, dy = x.y - y.y, dz = x.z - y.z, dw = x.w - y.w;
                        p.x = __builtin_fmaf(dx, dx, p.x);
                        p.y = __builtin_fmaf(dy, dy, p.y);
                        p.z = __builtin_fmaf(dz, dz, p.z);
                        p.w = __builtin_fmaf(dw, dw, p.w);
                    } else {
                        p.x = __builtin_fmaf(x.x, y.x, p.x);
                        p.y = __builtin_fmaf(x.y, y.y, p.y);
                        p.z = __builtin_fmaf(x.z, y.z, p.z);
                        p.w = __builtin_fmaf(x.w, y.w, p.w);
                    }
                }
            }
            float acc = quad_reduce(p);
            float qn = (METRIC == METRIC_L2SQR) ? 0.f : a.qnorms[q0 + r];
            float d = finish_metric<METRIC>(acc, ne, qn);
            if (t == 0) a.scores[size_t(q0 + r) * a.E + e] = d;
        }
        return;
    }
    for (unsigned r = 0; r < nqb; ++r) {
        float acc = quad_metric<METRIC>(row, qs + size_t(r) * a.Dp, a.D, t);
        float qn = (METRIC == METRIC_L2SQR) ? 0.f : a.qnorms[q0 + r];
        float d = finish_metric<METRIC>(acc, ne, qn);
        if (t == 0) a.scores[size_t(q0 + r) * a.E + e] = d;
    }
}

__device__ __forceinline__ uint64_t shfl_xor64(uint64_t v, int m) {
    uint32_t lo = __shfl_xor(uint32_t(v), m), hi = __shfl_xor(uint32_t(v >> 32), m);
    return (uint64_t(hi) << 32) | lo;
}

__global__ void __launch_bounds__(256) entry_topm_kernel(kg_entry_args a) {
    unsigned const w = (blockIdx.x * blockDim.x + threadIdx.x) >> 6, lane = threadIdx.x & 63;
    if (w >= a.nq) return;
    uint64_t best[MAXM];
#pragma unroll
    for (unsigned i = 0; i < MAXM; ++i) best[i] = ~0ull;
    float const *s = a.scores + size_t(w) * a.E;
    for (unsigned e = lane; e < a.E; e += KG_WAVE) {
        uint64_t key = (uint64_t(ord_f32(s[e])) << 32) | a.ids[e];
        if (key < best[MAXM - 1]) {  // sorted insert with compile-time indices
            bool placed = false;
#pragma unroll
            for (int i = MAXM - 1; i > 0; --i) {
                if (!placed) {
                    if (best[i - 1] > key) {
                        best[i] = best[i - 1];
                    } else {
                        best[i] = key;
                        placed = true;
                    }
                }
            }
            if (!placed) best[0] = key;
        }
    }
    // m rounds: the wave-wide minimum of the list heads is the next entry
    for (unsigned r = 0; r < a.m; ++r) {
        uint64_t v = best[0];
        for (int o = 32; o > 0; o >>= 1) {
            uint64_t u = shfl_xor64(v, o);
            v = u < v ? u : v;
        }
        if (best[0] == v && v != ~0ull) {  // unique keys: exactly one lane pops
#pragma unroll
            for (unsigned i = 0; i + 1 < MAXM; ++i) best[i] = best[i + 1];
            best[MAXM - 1] = ~0ull;
        }
        if (lane == 0) a.out_ids[size_t(w) * a.ostride + r] = uint32_t(v);
    }
}

// BF16 screening of queries against the entry sample on MFMA
// (v_mfma_f32_16x16x16bf16_1k, FP32 accumulation).  One wave scores 16
// queries x 64 sampled rows (four 16x16 accumulators); operands are rounded
// to BF16 (nearest even) as they are loaded.  The score only ranks rows for
// one query: ||e||^2/2 - q.e for the Euclidean forms, -q.e / ||e|| for angular.
// Operand layout (wave64): lane l holds A[m = l%16][k = 4*(l/16) + j] and
// B[k = 4*(l/16) + j][n = l%16]; accumulator lane l holds C[4*(l/16) + j][l%16].
typedef short bf16x4_t __attribute__((ext_vector_type(4)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));

__device__ __forceinline__ short bf16_rne(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return short(u >> 16);
}

template <int METRIC>
__global__ void __launch_bounds__(256) entry_mfma_kernel(kg_entry_args a) {
    unsigned const lane = threadIdx.x & 63;
    unsigned const wv = blockIdx.x * 4 + (threadIdx.x >> 6);
    unsigned const ecols = (a.E + 63) / 64;
    unsigned const qt = wv / ecols, et = wv - qt * ecols;
    if (qt * 16 >= a.nq) return;
    unsigned const r16 = lane & 15, kq = (lane >> 4) * 4;
    unsigned const qrow = qt * 16 + r16;
    f32x4_t acc[4] = {{0, 0, 0, 0}, {0, 0, 0, 0}, {0, 0, 0, 0}, {0, 0, 0, 0}};
    for (unsigned k0 = 0; k0 < a.D; k0 += 16) {
        bf16x4_t av;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            unsigned k = k0 + kq + j;
            av[j] = bf16_rne((qrow < a.nq && k < a.D) ? a.Q[size_t(qrow) * a.qstride + k] : 0.f);
        }
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            unsigned e = et * 64 + g * 16 + r16;
            bf16x4_t bv;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                unsigned k = k0 + kq + j;
                bv[j] = bf16_rne((e < a.E && k < a.D) ? a.rows[size_t(e) * a.Dp + k] : 0.f);
            }
            acc[g] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc[g], 0, 0, 0);
        }
    }
#pragma unroll
    for (int g = 0; g < 4; ++g) {
        unsigned e = et * 64 + g * 16 + r16;
        if (e >= a.E) continue;
        float const ne = a.enorms[e];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            unsigned q = qt * 16 + (lane >> 4) * 4 + j;
            if (q < a.nq) {
                float dot = acc[g][j];
                float s = (METRIC == METRIC_ANGULAR) ? -dot / ne : ne - dot;
                a.scores[size_t(q) * a.E + e] = s;
            }
        }
    }
}

// Exact re-rank of a per-query shortlist: one wave per query, one quad per
// shortlisted node (<= 16), the m best (dist, id) keys are written out.
template <int METRIC>
__global__ void __launch_bounds__(256) rescore_kernel(kg_rescore_args a) {
    __shared__ uint64_t keys[4][16];
    unsigned const w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    unsigned const q = blockIdx.x * 4 + w;
    if (q >= a.nq) return;
    unsigned const grp = lane >> 2, t = lane & 3;
    float const *qrow = a.Q + size_t(q) * a.qstride;
    float const qn = (METRIC == METRIC_L2SQR) ? 0.f : a.qnorms[q];
    if (grp < a.s) {
        uint32_t id = a.short_ids[size_t(q) * a.s + grp];
        float acc = quad_metric<METRIC>(a.X + size_t(id) * a.Dp, qrow, a.D, t);
        float na = (METRIC == METRIC_L2SQR) ? 0.f : a.norms[id];
        float d = finish_metric<METRIC>(acc, na, qn);
        if (t == 0) keys[w][grp] = (uint64_t(ord_f32(d)) << 32) | id;
    }
    wave_sync();
    if (lane < a.s) {
        uint64_t k = keys[w][lane];
        unsigned rank = 0;
        for (unsigned j = 0; j < a.s; ++j) rank += keys[w][j] < k;
        if (rank < a.m) a.out_ids[size_t(q) * a.ostride + rank] = uint32_t(k);
    }
}

__global__ void gather_rows_kernel(float const *X, unsigned Dp, uint32_t const *ids, unsigned E, float *out) {
    size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    unsigned const v4 = Dp >> 2;
    if (i >= size_t(E) * v4) return;
    size_t r = i / v4, c = i - r * v4;
    reinterpret_cast<float4 *>(out)[i] = reinterpret_cast<float4 const *>(X + size_t(ids[r]) * Dp)[c];
}

}  // namespace

hipError_t kg_launch_gather_rows(float const *X, unsigned Dp, uint32_t const *ids, unsigned E, float *out,
                                 hipStream_t s) {
    size_t n = size_t(E) * (Dp >> 2);
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(gather_rows_kernel, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s, X, Dp, ids, E, out);
    return hipGetLastError();
}

hipError_t kg_launch_rescore(int metric, kg_rescore_args const &a, hipStream_t s) {
    if (a.s == 0 || a.s > 16 || a.m > a.s || a.nq == 0) return hipErrorInvalidValue;
    dim3 g((a.nq + 3) / 4), b(256);
    switch (metric) {
        case METRIC_L2SQR: hipLaunchKernelGGL(rescore_kernel<METRIC_L2SQR>, g, b, 0, s, a); break;
        case METRIC_EUCLIDEAN: hipLaunchKernelGGL(rescore_kernel<METRIC_EUCLIDEAN>, g, b, 0, s, a); break;
        default: hipLaunchKernelGGL(rescore_kernel<METRIC_ANGULAR>, g, b, 0, s, a); break;
    }
    return hipGetLastError();
}

hipError_t kg_launch_entries_bf16(int metric, kg_entry_args const &a, hipStream_t s) {
    if (a.m == 0 || a.m > MAXM || a.E == 0 || a.nq == 0) return hipErrorInvalidValue;
    unsigned waves = ((a.nq + 15) / 16) * ((a.E + 63) / 64);
    dim3 g((waves + 3) / 4), b(256);
    if (metric == METRIC_ANGULAR)
        hipLaunchKernelGGL(entry_mfma_kernel<METRIC_ANGULAR>, g, b, 0, s, a);
    else
        hipLaunchKernelGGL(entry_mfma_kernel<METRIC_EUCLIDEAN>, g, b, 0, s, a);
    hipLaunchKernelGGL(entry_topm_kernel, dim3((a.nq + 3) / 4), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t kg_launch_entries(int metric, kg_entry_args const &a, hipStream_t s) {
    if (a.m == 0 || a.m > MAXM || a.E == 0 || a.nq == 0) return hipErrorInvalidValue;
    dim3 g1((a.nq + QB - 1) / QB, (a.E + EB - 1) / EB), b1(256);
    size_t lds = size_t(QB) * a.Dp * 4;
    switch (metric) {
        case METRIC_L2SQR: hipLaunchKernelGGL(entry_dist_kernel<METRIC_L2SQR>, g1, b1, lds, s, a); break;
        case METRIC_EUCLIDEAN: hipLaunchKernelGGL(entry_dist_kernel<METRIC_EUCLIDEAN>, g1, b1, lds, s, a); break;
        default: hipLaunchKernelGGL(entry_dist_kernel<METRIC_ANGULAR>, g1, b1, lds, s, a); break;
    }
    hipLaunchKernelGGL(entry_topm_kernel, dim3((a.nq + 3) / 4), dim3(256), 0, s, a);
    return hipGetLastError();
}

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
//   entry_exact_kernel MFMA scoring fused with per-lane top-m lists: exact on
//                      BF16-exact rows and queries; otherwise (BF16 screen,
//                      projected space) a shortlist re-ranked by rescore_kernel
#include <algorithm>

#include "kg_device.hpp"
#include "kg_internal.hpp"

namespace {

using namespace kgd;

constexpr unsigned QB = 16;    // queries per block
constexpr unsigned EB = 64;    // sampled rows per block (one per quad)
constexpr unsigned MAXM = 16;  // largest m
constexpr unsigned ECB = 8;    // MFMA selection: buffered candidates per lane

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

typedef short bf16x4_t __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));

__device__ __forceinline__ short bf16_rne(float f) {  // round to nearest even (finite values)
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return short(u >> 16);
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
        float acc = quad_metric<METRIC, float, true>(a.X + size_t(id) * a.Dp, qrow, a.D, t);
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

// Whether rows hold only integers, and their largest |value|: flags[0] |= 1
// on a non-integer, flags[1] = max |x| as float bits (both zeroed by the
// caller).  Atomics only while they can still change the flags.
__global__ void __launch_bounds__(256) intcheck_kernel(float const *X, unsigned n, unsigned D, unsigned stride,
                                                       uint32_t *flags) {
    unsigned const lane = threadIdx.x & 63;
    size_t const total = size_t(n) * D;
    bool frac = false;
    float mx = 0.f;
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < total; i += size_t(gridDim.x) * blockDim.x) {
        size_t const r = i / D, c = i - r * D;
        float const v = X[r * stride + c];
        frac |= !(v == rintf(v));
        mx = fmaxf(mx, fabsf(v));
    }
    if (ballot(frac) && lane == 0 && !__hip_atomic_load(flags, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
        atomicOr(flags, 1u);
    for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o));
    if (lane == 0 && __float_as_uint(mx) > __hip_atomic_load(flags + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
        atomicMax(flags + 1, __float_as_uint(mx));
}

// Query norms ||q||^2 / 2 (canonical order) for the MFMA selection, fused with
// intcheck_kernel's test of the same rows: flags[0] |= 1 on a non-integer,
// flags[1] = max |q| as float bits, flags[2] counts finished blocks (all three
// zeroed by the caller).  The last block copies flags[0..1] to host_out
// (pinned host memory), so the caller reads them once the kernel is done,
// without a copy, and sets flags[3] to whether the BF16 tiles give exact
// distances for these queries against sample rows of largest |x| emax (read
// by the MFMA selection's gate).
__global__ void __launch_bounds__(256) qnorm_check_kernel(float const *Q, unsigned n, unsigned D, unsigned stride,
                                                          float *qsq, uint32_t *flags, uint32_t *host_out,
                                                          float emax) {
    unsigned const lane = threadIdx.x & 63, t = lane & 3;
    unsigned const row = (blockIdx.x * 256 + threadIdx.x) >> 2;
    bool frac = false;
    float mx = 0.f;
    if (row < n) {  // quad-uniform: the quad reduce inside quad_metric sees all four lanes
        float const *r = Q + size_t(row) * stride;
        float const n2 = quad_metric<METRIC_EUCLIDEAN, float, true>(r, r, D, t);
        if (t == 0) qsq[row] = finish_norm<METRIC_EUCLIDEAN>(n2);
        // the same row again (cache hits), 16 bytes a load (rows are 16-byte
        // aligned, stride a multiple of 4), the last D % 4 values one by one
        auto chk = [&](float v) {
            frac |= !(v == rintf(v));
            mx = fmaxf(mx, fabsf(v));
        };
        unsigned const d4 = D >> 2;
        for (unsigned c = t; c < d4; c += 4) {
            float4 const v = reinterpret_cast<float4 const *>(r)[c];
            chk(v.x);
            chk(v.y);
            chk(v.z);
            chk(v.w);
        }
        for (unsigned c = 4 * d4 + t; c < D; c += 4) chk(r[c]);
    }
    uint64_t const fr = ballot(frac);
    for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o));
    if (lane == 0) {
        if (fr) atomicOr(flags, 1u);
        atomicMax(flags + 1, __float_as_uint(mx));
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        if (atomicAdd(flags + 2, 1u) == gridDim.x - 1) {
            __threadfence();
            uint32_t const f0 = __hip_atomic_load(flags, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            uint32_t const f1 = __hip_atomic_load(flags + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            host_out[0] = f0;
            host_out[1] = f1;
            flags[3] = kg_entry_exact_bf16(f0, __uint_as_float(f1), D, emax) ? 1u : 0u;
        }
    }
}

// One thread per 16-byte chunk of the tiled sample (layout in kg_internal.hpp).
__global__ void __launch_bounds__(256) entry_tiles_kernel(float const *rows, unsigned E, unsigned D, unsigned Dp,
                                                          unsigned KS, size_t nchunks, uint4 *tiles) {
    size_t const i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= nchunks) return;
    unsigned const l = unsigned(i % KG_WAVE), s = unsigned((i / KG_WAVE) % KS);
    unsigned const t = unsigned(i / (size_t(KS) * KG_WAVE));
    unsigned const row = 16 * t + (l & 15), k0 = 32 * s + 8 * (l >> 4);
    uint32_t w[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        unsigned const k = k0 + 2 * j;
        float const lo = (row < E && k < D) ? rows[size_t(row) * Dp + k] : 0.f;
        float const hi = (row < E && k + 1 < D) ? rows[size_t(row) * Dp + k + 1] : 0.f;
        w[j] = uint32_t(bf16_bits(lo)) | (uint32_t(bf16_bits(hi)) << 16);
    }
    tiles[i] = make_uint4(w[0], w[1], w[2], w[3]);
}

// Entry selection on the matrix cores.  For rows and queries whose BF16 dot
// products are exact (integers, |x| <= 256, D (max|q|^2 + max|e|^2) <= 2^24)
// every product, partial sum and norm is an integer of at most 2^24 and the
// tile value equals the canonical FP32 distance bit for bit, so the m best
// are the answer; for other rows (or the projected space) the m best by the
// BF16 value form the shortlist that rescore_kernel re-ranks exactly.
// One block per 16 queries, 4 waves splitting the sample's 16-row tiles
// (wave w takes tiles w, w + 4, ...).  A tile's operands: the sampled rows
// (A, two tiles in flight: a tile's MFMAs read its buffer, then the loads of
// the tile two ahead refill it; no load is conditional and no loaded value is
// used before its tile is scored, so the wait for a tile leaves the next
// one's loads outstanding) and the block's 16 queries (B, kept in registers),
// so lane l holds the distances of query l % 16 to rows 4 (l / 16) + j.  Each
// lane keeps the MF >= m best of the rows it sees as 32-bit ordered distances
// with their sample positions: it sees rows in ascending position (and id)
// order, so inserting after equal distances keeps (dist, id) order.  The 16 lanes of a query (4 per wave, 4 waves) share a bound, the
// smallest of their m-th best distances (any lane's m best are m keys no
// worse than it, so the query's m-th result is no worse either): a distance
// above it cannot enter the result and skips the insert.  Passing
// candidates wait in a per-lane LDS buffer (8 entries) and are inserted
// together, so a wave takes an insert step per buffered candidate of its
// fullest lane rather than one per row that any of its 64 lanes passes.  The four lanes of
// a query and then the four waves are merged into the m best.
// KS: 32-value slices per row, one v_mfma_f32_16x16x32_bf16 each (lane l
// holds values 32 s + 8 (l / 16) .. + 8 of its row or query).
template <int METRIC, int KS, int MF>
__global__ void __launch_bounds__(256) entry_exact_kernel(kg_entry_args a, float const *qsq, float const *esq) {
    __shared__ uint64_t best[4][16][MF];
    __shared__ uint32_t bound[16];  // per query: smallest m-th best ordered distance of its lanes
    __shared__ uint64_t cbuf[ECB * 256];  // per lane: candidates that passed, (distance, position)
    unsigned const lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    unsigned const r16 = lane & 15, kq = (lane >> 4) * 4, k8 = (lane >> 4) * 8;
    if (a.gate && *a.gate == 0u) return;  // another form selects for these queries
    unsigned const q = blockIdx.x * 16 + r16;
    bool const qok = q < a.nq;
    if (threadIdx.x < 16) bound[threadIdx.x] = 0xffffffffu;
    __syncthreads();
    bf16x8_t bop[KS];
    {
        float const *qr = a.Q + size_t(qok ? q : 0u) * a.qstride;
#pragma unroll
        for (int s = 0; s < KS; ++s) {
            uint32_t wv[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                unsigned const k = 32 * s + k8 + 2 * j;
                uint32_t const lo = uint16_t(bf16_rne((qok && k < a.D) ? qr[k] : 0.f));
                uint32_t const hi = uint16_t(bf16_rne((qok && k + 1 < a.D) ? qr[k + 1] : 0.f));
                wv[j] = lo | (hi << 16);
            }
            bop[s] = __builtin_bit_cast(bf16x8_t, make_uint4(wv[0], wv[1], wv[2], wv[3]));
        }
    }
    float const qn = qok ? ((METRIC == METRIC_L2SQR) ? 2.f * qsq[q] : a.qnorms[q]) : 0.f;
    uint32_t lk[MF], li[MF];  // ordered distances and ids, ascending
#pragma unroll
    for (int i = 0; i < MF; ++i) {
        lk[i] = 0xffffffffu;
        li[i] = 0xffffffffu;
    }
    unsigned const ntiles = (a.E + 15) / 16;
    // operands of sample tile `et`: KS contiguous 1 KiB loads of the tiled
    // BF16 copy, nothing computed from them here, so they stay in flight
    // until the tile is scored
    auto load_tile = [&](unsigned et, bf16x8_t (&op)[KS]) {
        uint4 const *tp = a.tiles + size_t(et) * (KS * KG_WAVE) + lane;
#pragma unroll
        for (int s = 0; s < KS; ++s) op[s] = __builtin_bit_cast(bf16x8_t, tp[s * KG_WAVE]);
    };
    // the norms of the 4 sampled rows whose results a lane holds (rows
    // 16 et + kq .. + 3): one 16-byte load, the norm arrays hold 16 ntiles
    // entries (the ones past E are never used)
    float const *nsrc = (METRIC == METRIC_L2SQR) ? esq : a.enorms;
    auto load_norms = [&](unsigned et, float4 &nv) {
        nv = *reinterpret_cast<float4 const *>(nsrc + et * 16 + kq);
    };
    // two tiles in flight: buffers A (tiles w, w + 8, ...) and B (w + 4,
    // w + 12, ...).  Every call loads and scores, also past the last tile
    // (a clamped tile index; its rows are skipped as >= E), so no load is
    // conditional and no loaded value is copied on the loop's back edge,
    // which would make the wave wait for it there.
    auto clamp_tile = [&](unsigned et) { return et < ntiles ? et : ntiles - 1; };
    bf16x8_t opA[KS], opB[KS];
    float4 nA, nB;
    load_tile(clamp_tile(w), opA);
    load_norms(clamp_tile(w), nA);
    load_tile(clamp_tile(w + 4), opB);
    load_norms(clamp_tile(w + 4), nB);
    // candidates that pass the bound wait in the lane's LDS buffer (in the
    // order the lane sees them: ascending position) and are inserted
    // together, so a wave runs an insert step per buffered candidate of its
    // fullest lane instead of one per row that any of its 64 lanes passes
    unsigned cnt = 0;
    auto insert = [&](uint32_t k, uint32_t e) __attribute__((always_inline)) {
        // branch-free sorted insert after equal distances (c_i = lk[i] > k is
        // monotone in i); k = 0xffffffff inserts nothing
        bool c[MF];
#pragma unroll
        for (int i = 0; i < MF; ++i) c[i] = lk[i] > k;
#pragma unroll
        for (int i = MF - 1; i > 0; --i) {
            lk[i] = c[i - 1] ? lk[i - 1] : (c[i] ? k : lk[i]);
            li[i] = c[i - 1] ? li[i - 1] : (c[i] ? e : li[i]);
        }
        lk[0] = c[0] ? k : lk[0];
        li[0] = c[0] ? e : li[0];
    };
    auto flush = [&]() __attribute__((always_inline)) {
#pragma unroll
        for (unsigned i = 0; i < ECB; ++i) {
            if (ballot(i < cnt)) {
                uint64_t const key = i < cnt ? cbuf[i * 256 + threadIdx.x] : ~0ull;
                insert(uint32_t(key >> 32), uint32_t(key));
            }
        }
        cnt = 0;
    };
    uint32_t tq = 0xffffffffu;  // the query's shared bound as this lane last saw it
    auto score = [&](unsigned et, bf16x8_t (&op)[KS], float4 &nv) {
        if (((et >> 2) & 3) == 3) {
            flush();
            // refresh the bound: this lane's m-th best, the four lanes of the
            // query in this wave, then the block's other waves (LDS minimum)
            uint32_t mine = 0xffffffffu;
#pragma unroll
            for (int i = 0; i < MF; ++i)
                if (unsigned(i) + 1 == a.m) mine = lk[i];
            mine = min(mine, uint32_t(__shfl_xor(int(mine), 16)));
            mine = min(mine, uint32_t(__shfl_xor(int(mine), 32)));
            atomicMin(&bound[r16], mine);
            tq = min(tq, __hip_atomic_load(&bound[r16], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP));
        }
        // the MFMAs read this buffer, then the loads two tiles ahead refill it
        f32x4_t acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int s = 0; s < KS; ++s) acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(op[s], bop[s], acc, 0, 0, 0);
        load_tile(clamp_tile(et + 8), op);
        float const en[4] = {nv.x, nv.y, nv.z, nv.w};
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            unsigned const ee = et * 16 + kq + j;  // row of C: sampled row 4 (l / 16) + j
            float const G = acc[j];
            float d;
            if (METRIC == METRIC_L2SQR)
                d = (qn + 2.f * en[j]) - 2.f * G;
            else if (METRIC == METRIC_EUCLIDEAN)
                d = (en[j] + qn) - G;
            else
                d = -G / __builtin_fmaf(en[j], qn, 1.1920928955078125e-07f);
            uint32_t const kd = ord_f32(d);
            // li holds sample positions: the sample's ids ascend with their
            // positions, so (distance, position) orders as (distance, id)
            bool const pass = qok && ee < a.E && kd <= tq && kd < lk[MF - 1];
            if (pass) cbuf[(cnt++) * 256 + threadIdx.x] = (uint64_t(kd) << 32) | ee;
        }
        if (ballot(cnt + 4 > ECB)) flush();  // room for the next tile's 4 rows
        load_norms(clamp_tile(et + 8), nv);
    };
    for (unsigned et = w; et < ntiles; et += 8) {
        score(et, opA, nA);
        score(et + 4, opB, nB);
    }
    flush();
    // the four lanes of a query (lane % 16 equal): m rounds of the minimum head
    for (unsigned r = 0; r < a.m; ++r) {
        uint64_t const head = (uint64_t(lk[0]) << 32) | li[0];
        uint64_t v = head;
        uint64_t u = shfl_xor64(v, 16);
        v = u < v ? u : v;
        u = shfl_xor64(v, 32);
        v = u < v ? u : v;
        if (head == v && v != ~0ull) {  // keys are unique: one lane pops
#pragma unroll
            for (int i = 0; i + 1 < MF; ++i) {
                lk[i] = lk[i + 1];
                li[i] = li[i + 1];
            }
            lk[MF - 1] = 0xffffffffu;
            li[MF - 1] = 0xffffffffu;
        }
        if (lane < 16) best[w][r16][r] = v;
    }
    __syncthreads();
    // the four waves' lists of each query -> its m best
    if (w == 0 && lane < 16 && qok) {
        unsigned pos[4] = {0, 0, 0, 0};
        for (unsigned r = 0; r < a.m; ++r) {
            uint64_t v = ~0ull;
            unsigned src = 0;
#pragma unroll
            for (unsigned ww = 0; ww < 4; ++ww) {
                uint64_t const h = pos[ww] < a.m ? best[ww][r16][pos[ww]] : ~0ull;
                if (h < v) {
                    v = h;
                    src = ww;
                }
            }
            ++pos[src];
            a.out_ids[size_t(q) * a.ostride + r] = uint32_t(v) < a.E ? a.ids[uint32_t(v)] : 0xffffffffu;
        }
    }
}

}  // namespace

hipError_t kg_launch_qnorm_check(float const *Q, unsigned n, unsigned D, unsigned stride, float *qsq,
                                 uint32_t *flags, uint32_t *host_out, float emax, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(qnorm_check_kernel, dim3((size_t(n) * 4 + 255) / 256), dim3(256), 0, s, Q, n, D, stride, qsq,
                       flags, host_out, emax);
    return hipGetLastError();
}

hipError_t kg_launch_intcheck(float const *X, unsigned n, unsigned D, unsigned stride, uint32_t *flags,
                              hipStream_t s) {
    size_t const total = size_t(n) * D;
    if (total == 0) return hipSuccess;
    unsigned const nb = unsigned(std::min<size_t>(2048, (total + 255) / 256));
    hipLaunchKernelGGL(intcheck_kernel, dim3(nb), dim3(256), 0, s, X, n, D, stride, flags);
    return hipGetLastError();
}

template <int METRIC, int KS>
static void launch_exact_ks(kg_entry_args const &a, float const *qsq, float const *esq, hipStream_t s) {
    dim3 g((a.nq + 15) / 16), b(256);
    if (a.m <= 4)
        hipLaunchKernelGGL((entry_exact_kernel<METRIC, KS, 4>), g, b, 0, s, a, qsq, esq);
    else if (a.m <= 8)
        hipLaunchKernelGGL((entry_exact_kernel<METRIC, KS, 8>), g, b, 0, s, a, qsq, esq);
    else if (a.m <= 9)
        hipLaunchKernelGGL((entry_exact_kernel<METRIC, KS, 9>), g, b, 0, s, a, qsq, esq);
    else if (a.m <= 12)
        hipLaunchKernelGGL((entry_exact_kernel<METRIC, KS, 12>), g, b, 0, s, a, qsq, esq);
    else
        hipLaunchKernelGGL((entry_exact_kernel<METRIC, KS, 16>), g, b, 0, s, a, qsq, esq);
}

template <int METRIC>
static void launch_exact_m(kg_entry_args const &a, float const *qsq, float const *esq, hipStream_t s) {
    unsigned const ks = kg_entry_slices(a.D);
    if (ks == 2)
        launch_exact_ks<METRIC, 2>(a, qsq, esq, s);
    else if (ks == 4)
        launch_exact_ks<METRIC, 4>(a, qsq, esq, s);
    else
        launch_exact_ks<METRIC, 8>(a, qsq, esq, s);
}

// The sample's operands are its tiles (a.tiles, kg_launch_entry_tiles).  The
// norm array the kernel reads (esq for squared Euclidean, a.enorms otherwise)
// must hold 16 ceil(E / 16) floats: it is read a 16-row tile at a time.  a.ids
// must ascend (a sorted sample of distinct ids).
hipError_t kg_launch_entries_exact_mfma(int metric, kg_entry_args const &a, float const *qsq, float const *esq,
                                        hipStream_t s) {
    if (a.m == 0 || a.m > MAXM || a.E == 0 || a.nq == 0 || a.D > 256 || !a.tiles) return hipErrorInvalidValue;
    if (metric == METRIC_L2SQR)
        launch_exact_m<METRIC_L2SQR>(a, qsq, esq, s);
    else if (metric == METRIC_EUCLIDEAN)
        launch_exact_m<METRIC_EUCLIDEAN>(a, qsq, esq, s);
    else
        launch_exact_m<METRIC_ANGULAR>(a, qsq, esq, s);
    return hipGetLastError();
}

unsigned kg_entry_slices(unsigned D) {
    unsigned const ks = (D + 31) / 32;
    return ks <= 2 ? 2u : ks <= 4 ? 4u : ks <= 8 ? 8u : ks;
}

size_t kg_entry_tiles_size(unsigned E, unsigned D) {
    return size_t((E + 15) / 16) * kg_entry_slices(D) * KG_WAVE;
}

hipError_t kg_launch_entry_tiles(float const *rows, unsigned E, unsigned D, unsigned Dp, uint4 *tiles,
                                 hipStream_t s) {
    size_t const n = kg_entry_tiles_size(E, D);
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(entry_tiles_kernel, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s, rows, E, D, Dp,
                       kg_entry_slices(D), n, tiles);
    return hipGetLastError();
}

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

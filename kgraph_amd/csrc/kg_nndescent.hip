// kg_nndescent.hip -- NN-descent k-NN graph construction kernels (gfx950).
//
// Restates KGraphConstructor (kgraph.cpp:705-1001) as a bulk-synchronous
// pipeline.  The reference's join reads per-node lists that are fixed for the
// whole iteration and only its early-reject `radius` reads live state
// (kgraph.cpp:724, 797-815), so one iteration is a map from node
// neighbourhoods to (target, id, dist) candidates followed by a per-target
// bounded sorted merge, which is what these kernels do:
//
//   init_kernel    kgraph.cpp:760-796  S random pool entries + 2S random "new"
//   join_tile_kernel / join_mfma_kernel  kgraph.cpp:739-751, 797-815  every
//                  new x new (by id, i < j) and new x old pair, scored as
//                  16 x 16 tiles (exact FP32 quads, or BF16 MFMA); candidates
//                  below the target's radius are staged in LDS and written to
//                  the target's claimed candidate slots (one slot-claim atomic
//                  per target and flush; claims past C go to a hashed
//                  overflow area where a 64-bit atomicMin keeps the better key)
//   merge_kernel   kgraph.cpp:176-202, 723-736  sorted pool U candidates ->
//                  first L distinct ids (equal distance: lower id first; an
//                  existing entry keeps its flag)
//   update_m_kernel kgraph.cpp:826-841  M growth and radiusM
//   sample_kernel  kgraph.cpp:842-866  split pool[0..M) into new / old, clear
//                  the flags, push n onto the other endpoint's reverse lists
//                  when dist > radiusM (reservoir-capped at R, the GPU form of
//                  kgraph.cpp:867-882's shuffle + truncate)
//
// Pool entries are u64 keys: orderable dist bits << 32 | id << 1 | new-flag,
// so integer order is (dist, id) order.  Every stored distance uses the
// canonical FP32 order of kg_device.hpp.
#include <algorithm>
#include <climits>
#include <type_traits>

#include "kg_device.hpp"
#include "kg_nnd.hpp"

namespace {

using namespace kgd;

__device__ __forceinline__ uint64_t make_key(float d, uint32_t id, uint32_t flag) {
    return (uint64_t(ord_f32(d)) << 32) | (uint64_t(id) << 1) | flag;
}
__device__ __forceinline__ uint32_t key_id(uint64_t k) { return uint32_t(k) >> 1; }
__device__ __forceinline__ float key_dist(uint64_t k) { return unord_f32(uint32_t(k >> 32)); }
__device__ __forceinline__ uint32_t key_flag(uint64_t k) { return uint32_t(k) & 1u; }

__device__ __forceinline__ uint64_t mix64(uint64_t x) {  // splitmix64 finaliser
    x += 0x9e3779b97f4a7c15ull;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
    return x ^ (x >> 31);
}

struct Rng {  // counter-based stream per node
    uint64_t s;
    __device__ uint32_t operator()() { return uint32_t(mix64(s++) >> 32); }
};

// GenRandom (kgraph.cpp:44-65) across one wave.  The node's stream is
// counter-based (value k is mix64(s0 + k) >> 32), so lane i makes the i-th draw
// itself: rng() % (N - size) for i < size, then the offset rng() % N.  The
// one-lane insertion sort becomes a stable rank by counting (equal draws are
// interchangeable), and its fix-up pass d'[i] = max(d[i], d'[i-1] + 1) equals
// max_{j<=i}(d[j] + i - j), a running maximum of d[j] - j across the wave.
// `raw` and `srt` are LDS scratch of `size` entries; out(i, id) receives
// element i.  Returns the number of stream values used.
template <class Out>
__device__ unsigned gen_random_wave(uint64_t s0, uint32_t *raw, uint32_t *srt, unsigned size, unsigned N,
                                    unsigned lane, Out out) {
    if (N == size) {
        for (unsigned i = lane; i < size; i += 64) out(i, i);
        return 0;
    }
    unsigned const m = N - size;
    for (unsigned i = lane; i < size; i += 64) raw[i] = uint32_t(mix64(s0 + i) >> 32) % m;
    wave_sync();
    for (unsigned i = lane; i < size; i += 64) {
        uint32_t const v = raw[i];
        unsigned r = 0;
        for (unsigned j = 0; j < size; ++j) {
            uint32_t const u = raw[j];
            r += (u < v) | ((u == v) & (j < i));
        }
        srt[r] = v;
    }
    wave_sync();
    uint64_t const off = uint32_t(mix64(s0 + size) >> 32) % N;
    long long carry = LLONG_MIN;
    for (unsigned b = 0; b < size; b += 64) {
        unsigned const i = b + lane;
        long long x = (i < size) ? (long long)srt[i] - (long long)i : LLONG_MIN;
        for (unsigned o = 1; o < 64; o <<= 1) {
            long long const y = __shfl_up(x, o);
            if (lane >= o) x = max(x, y);
        }
        x = max(x, carry);
        carry = __shfl(x, 63);
        if (i < size) out(i, uint32_t((uint64_t(x + i) + off) % N));
    }
    wave_sync();
    return size + 1;
}

// Canonical metric of data rows i and j (both global), one quad.
template <int METRIC>
__device__ __forceinline__ float pair_dist(NndArgs const &a, uint32_t i, float const *rj, float nj, unsigned t) {
    float const *ri = a.X + size_t(i) * a.Dp;
    float acc = quad_metric<METRIC, float, true>(ri, rj, a.D, t);
    float ni = (METRIC == METRIC_L2SQR) ? 0.f : a.norms[i];
    return finish_metric<METRIC>(acc, ni, nj);
}

// ------------------------------------------------------------------ init
template <int METRIC>
__global__ void __launch_bounds__(256) init_kernel(NndArgs a) {
    __shared__ uint32_t raw[4][2 * NND_MAX_S];
    __shared__ uint32_t srt[4][2 * NND_MAX_S];
    __shared__ uint32_t rnd[4][NND_MAX_S + 1];
    __shared__ uint32_t pid[4][NND_MAX_S];
    __shared__ uint64_t key[4][NND_MAX_S];
    unsigned const w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    unsigned const n = blockIdx.x * 4 + w;
    if (n >= a.N) return;
    unsigned const S = a.S;
    uint64_t s0 = mix64(a.seed * 0x100000001b3ull + n);
    // nn_new: 2S random ids (kgraph.cpp:765, 781), GenRandom written out on
    // the node's own sample row.
    uint32_t *dst = a.samp + size_t(n) * a.Lc;
    s0 += gen_random_wave(s0, raw[w], srt[w], a.nnew_init, a.N, lane, [&](unsigned i, uint32_t id) { dst[i] = id; });
    // pool: S random non-self ids (kgraph.cpp:782-792); n occurs at most once
    // among the S + 1 distinct draws and is skipped.
    gen_random_wave(s0, raw[w], srt[w], S + 1, a.N, lane, [&](unsigned i, uint32_t id) { rnd[w][i] = id; });
    wave_sync();
    unsigned self = S + 1;
    for (unsigned b = 0; b <= S; b += 64) {
        uint64_t const hit = __ballot(b + lane <= S && rnd[w][b + lane] == n);
        if (hit && self > S) self = b + unsigned(__ffsll((unsigned long long)hit)) - 1;
    }
    for (unsigned l = lane; l < S; l += 64) pid[w][l] = rnd[w][l + (l >= self)];
    wave_sync();
    unsigned const grp = lane >> 2, t = lane & 3;
    float const *rn = a.X + size_t(n) * a.Dp;
    float const nn_ = (METRIC == METRIC_L2SQR) ? 0.f : a.norms[n];
    for (unsigned b = 0; b < S; b += 16) {
        if (b + grp < S) {
            uint32_t const id = pid[w][b + grp];
            float d = pair_dist<METRIC>(a, id, rn, nn_, t);
            if (t == 0) key[w][b + grp] = make_key(d, id, 1);
        }
    }
    wave_sync();
    // sort the S keys by (dist, id) (kgraph.cpp:793): ids are distinct, so a
    // key's rank is the number of smaller keys.
    uint64_t *pool = a.pool + size_t(n) * a.L;
    for (unsigned l = lane; l < S; l += 64) {
        uint64_t const v = key[w][l];
        unsigned r = 0;
        for (unsigned j = 0; j < S; ++j) r += key[w][j] < v;
        pool[r] = v;
        if (r == a.L - 1) a.meta[n].x = key_dist(v);
        if (r == a.K - 1) a.meta[n].y = key_dist(v);
    }
    if (lane == 0) {
        a.cnt[n] = S;
        a.Mv[n] = S;
        a.nnew[n] = a.nnew_init;
        a.nold[n] = 0;
        a.crnew[n] = 0;
        a.crold[n] = 0;
        if (S < a.L) a.meta[n].x = __builtin_huge_valf();
        if (S < a.K) a.meta[n].y = __builtin_huge_valf();
        a.ccnt[n] = 0;
    }
}

// ------------------------------------------------------------------ join
// A claimed slot (< C) belongs to one candidate: a plain store.  Claims past C
// go to a hashed slot of the target's overflow area, where a 64-bit atomicMin
// keeps the better key (and one id always lands on the same slot).
__device__ __forceinline__ void put_cand(NndArgs const &a, uint32_t target, unsigned slot, uint64_t key,
                                         uint32_t src) {
    if (slot < a.C) {
        a.cand[size_t(target) * a.C + slot] = key;
        if (src != NND_NOSRC) a.csrc[size_t(target) * a.C + slot] = src;
    } else {  // (an overflow entry carries no credit)
        unsigned const h = unsigned(mix64((uint64_t(target) << 32) | (uint32_t(key) >> 1))) & (a.O - 1u);  // O: a power of two
        atomicMin(reinterpret_cast<unsigned long long *>(&a.ovf[size_t(target) * a.O + h]),
                  static_cast<unsigned long long>(key));
    }
}

// ------------------------------------------------------------------ MFMA join
// The local join as dense distance tiles on the matrix cores.  One wavefront
// per node; the node's list (new rows, then old rows) is in LDS.  Every 16 x 16
// block of (new row, list row) pairs is one tile: both operands go from HBM
// straight into MFMA operand registers as BF16 (operand row = lane % 16,
// columns 32 s + 8 (lane / 16) .. +7 in k-step s), gfx950's
// v_mfma_f32_16x16x32_bf16 accumulates the dot products in FP32 (32 columns
// per instruction), and each lane finishes the four pairs it holds,
// C[4 (lane / 16) + j][lane % 16], from the row norms:
//   L2SQR  d = |a|^2 + |o|^2 - 2 a.o,   EUCLIDEAN  d = n_a + n_o - a.o,
//   ANGULAR  d = -a.o / fmaf(n_a, n_o, FLT_EPSILON).
// EXACT (every row value an integer and 2 D max|x|^2 <= 2^24): BF16 holds each
// value, and every product, partial sum and norm is an integer at most 2^24,
// so the tile value equals the canonical FP32 distance bit for bit and is used
// as it is.  Otherwise the tile value screens: a pair is re-scored exactly
// (canonical FP32 order, kg_device.hpp) when d - E is below either endpoint's
// radius, E bounding |screen - canonical| (BF16 rounding of both operands, FP32
// accumulation, the norms' rounding), so no pair that passes the exact test is
// screened out and the re-scored distance alone decides.  Candidates are staged
// in LDS and flushed grouped by target as in join_kernel.
typedef __bf16 mj_bf16x8 __attribute__((ext_vector_type(8)));
typedef float mj_f32x4 __attribute__((ext_vector_type(4)));
typedef int mj_i32x4 __attribute__((ext_vector_type(4)));

constexpr unsigned MT = 512;  // list capacity (new + old)
constexpr unsigned MB = 512;  // staged candidates per wave (exact tile forms)
constexpr unsigned SB = 256;  // staged candidates per wave (screen form: a quarter as many pushes per node)
// A staged candidate's list index, and whether it credits the joining node's
// `found`: the insert of the callback's first endpoint (kgraph.cpp:805-807)
// with its distance below that endpoint's K-th -- the merge keeps the credit
// only when the id is new to the list (a duplicate returns K + 1 there,
// kgraph.cpp:186-192).
constexpr unsigned TGT_MASK = 0x7fffu, TGT_CREDIT = 0x8000u;

struct MJoinLDS {
    static constexpr unsigned kMB = MB;
    uint32_t ids[MT];
    uint32_t cnt[MT];
    float4 am[32];  // per-node records of the new rows whose tiles are in registers (MFMA form)
    float an[16];   // join norms of the current 16 new rows (FP32 tile form)
    uint64_t key[MB];
    uint16_t tgt[MB];
};

struct SJoinLDS {
    static constexpr unsigned kMB = SB;
    uint32_t ids[MT];
    uint32_t cnt[MT];
    float4 am[16];  // per-node records of the tile's 16 new rows
    uint64_t key[SB];
    uint16_t tgt[SB];
};

// KS k-steps of 32 columns: lane l takes columns 32 s + 8 (l / 16) .. +7 of
// its row (zeros past Dp; Dp is a multiple of 4, so a step's last chunk may
// hold 4 values).  Rows start at multiples of 8 bytes: 16-byte loads when Dp
// is a multiple of 8, else two 8-byte loads.
template <int KS>
__device__ __forceinline__ void mj_load(NndArgs const &a, uint32_t id, bool ok, unsigned k8, mj_bf16x8 (&op)[KS]) {
    // operands straight from the BF16 copy of the rows (round to nearest even)
    uint16_t const *r = a.Xb + size_t(ok ? id : 0u) * a.Dp;
    bool const wide = (a.Dp & 7u) == 0;
#pragma unroll
    for (int s = 0; s < KS; ++s) {
        unsigned const k = 32 * s + k8;
        uint4 v = make_uint4(0u, 0u, 0u, 0u);
        if (ok && k + 8 <= a.Dp) {
            if (wide) {
                v = *reinterpret_cast<uint4 const *>(r + k);
            } else {
                uint2 const lo = *reinterpret_cast<uint2 const *>(r + k);
                uint2 const hi = *reinterpret_cast<uint2 const *>(r + k + 4);
                v = make_uint4(lo.x, lo.y, hi.x, hi.y);
            }
        } else if (ok && k + 4 <= a.Dp) {
            uint2 const lo = *reinterpret_cast<uint2 const *>(r + k);
            v = make_uint4(lo.x, lo.y, 0u, 0u);
        }
        op[s] = __builtin_bit_cast(mj_bf16x8, v);
    }
}

// FP32 operands for v_mfma_f32_16x16x4_f32 (the screen of the FP32-operand
// form): lane l takes the float4 at columns 4 (4 q + l / 16) .. +3 of its row
// for chunk q (zeros past Dp), and sub-step j of chunk q feeds element j of
// every lane group to one MFMA, so each MFMA sums the columns
// {16 q + 4 g + j : g = 0..3} -- the same permutation of the columns for both
// operands, hence the full dot product once all chunks are summed.
template <int KS>
__device__ __forceinline__ void mf_load(NndArgs const &a, uint32_t id, bool ok, unsigned g, float4 (&op)[KS]) {
    float const *r = a.X + size_t(ok ? id : 0u) * a.Dp;
#pragma unroll
    for (int q = 0; q < KS; ++q) {
        unsigned const k = 16 * q + 4 * g;
        op[q] = (ok && k < a.Dp) ? *reinterpret_cast<float4 const *>(r + k) : make_float4(0.f, 0.f, 0.f, 0.f);
    }
}

// Operand tiles of either form: BF16 (16x16x32 steps) or FP32 (16x16x4).
template <int KS, bool F32>
struct MjOps;
template <int KS>
struct MjOps<KS, false> {
    mj_bf16x8 v[KS];
    __device__ void load(NndArgs const &a, uint32_t id, bool ok, unsigned lane) { mj_load<KS>(a, id, ok, (lane >> 4) * 8, v); }
};
template <int KS>
struct MjOps<KS, true> {
    float4 v[KS];
    __device__ void load(NndArgs const &a, uint32_t id, bool ok, unsigned lane) { mf_load<KS>(a, id, ok, lane >> 4, v); }
};
// 8-bit operands for v_mfma_i32_16x16x64_i8 (rows of the exact 8-bit copy,
// shifted to [-128, 127] by flipping each byte's top bit): KS steps of 64
// columns, lane l takes columns 64 s + 16 (l / 16) .. +15 of its row; columns
// at D and beyond are zero after the shift.  16-byte loads when rows start at
// multiples of 16 bytes, else 4-byte ones.
template <int KS>
struct MjOps8 {
    mj_i32x4 v[KS];
    __device__ void load(NndArgs const &a, uint32_t id, bool ok, unsigned lane) {
        uint8_t const *r = a.X8 + size_t(ok ? id : 0u) * a.Dp;
        bool const wide = (a.Dp & 15u) == 0;
        unsigned const g16 = (lane >> 4) * 16;
#pragma unroll
        for (int s = 0; s < KS; ++s) {
            unsigned const k = 64 * s + g16;
            uint32_t w[4] = {0u, 0u, 0u, 0u};
            if (ok && k < a.D) {
                if (wide) {
                    uint4 const x = *reinterpret_cast<uint4 const *>(r + k);
                    w[0] = x.x;
                    w[1] = x.y;
                    w[2] = x.z;
                    w[3] = x.w;
                } else {
#pragma unroll
                    for (int q = 0; q < 4; ++q)
                        if (k + 4 * q + 4 <= a.Dp) w[q] = *reinterpret_cast<uint32_t const *>(r + k + 4 * q);
                }
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    int const live = int(a.D) - int(k + 4 * q);  // columns of this word below D
                    uint32_t const keep = live >= 4 ? 0xffffffffu : live <= 0 ? 0u : (1u << (8 * live)) - 1u;
                    w[q] = (w[q] ^ 0x80808080u) & keep;
                }
            }
            v[s] = mj_i32x4{int(w[0]), int(w[1]), int(w[2]), int(w[3])};
        }
    }
};
template <int KS>
__device__ __forceinline__ mj_f32x4 mj_mma(MjOps8<KS> const &x, MjOps8<KS> const &y) {
    mj_i32x4 acc = {0, 0, 0, 0};
#pragma unroll
    for (int s = 0; s < KS; ++s) acc = __builtin_amdgcn_mfma_i32_16x16x64_i8(x.v[s], y.v[s], acc, 0, 0, 0);
    return mj_f32x4{float(acc[0]), float(acc[1]), float(acc[2]), float(acc[3])};
}
template <int KS>
__device__ __forceinline__ mj_f32x4 mj_mma(MjOps<KS, false> const &x, MjOps<KS, false> const &y) {
    mj_f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int s = 0; s < KS; ++s) acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(x.v[s], y.v[s], acc, 0, 0, 0);
    return acc;
}
template <int KS>
__device__ __forceinline__ mj_f32x4 mj_mma(MjOps<KS, true> const &x, MjOps<KS, true> const &y) {
    mj_f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int q = 0; q < KS; ++q) {
        acc = __builtin_amdgcn_mfma_f32_16x16x4f32(x.v[q].x, y.v[q].x, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x4f32(x.v[q].y, y.v[q].y, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x4f32(x.v[q].z, y.v[q].z, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x4f32(x.v[q].w, y.v[q].w, acc, 0, 0, 0);
    }
    return acc;
}

// BF16 tile distance from the dot G and the norms the metric keeps (an, on:
// |x|^2 for SSD, else the stored norm), and the bound E on its error against
// the exact distance (arn, orn = |x|): the join screens with d - E.
template <int METRIC>
__device__ __forceinline__ float mj_dist(float G, float an, float on) {
    if (METRIC == METRIC_L2SQR) return (an + on) - 2.f * G;
    if (METRIC == METRIC_EUCLIDEAN) return (an + on) - G;
    return -G / __builtin_fmaf(an, on, 1.1920928955078125e-07f);
}
template <int METRIC>
__device__ __forceinline__ float mj_bound(NndArgs const &a, float an, float on, float arn, float orn) {
    if (METRIC == METRIC_L2SQR) return 2.f * a.scr_c * arn * orn + a.scr_g * (an + on);
    if (METRIC == METRIC_EUCLIDEAN) return a.scr_c * arn * orn + a.scr_g * (an + on);
    return a.scr_c + 2.f * a.scr_g;
}

template <class LDS>
__device__ void mj_flush(NndArgs const &a, LDS &L, unsigned fill, unsigned nT, uint32_t n) {
    unsigned const lane = lane_id();
    for (unsigned i = lane; i < nT; i += KG_WAVE) L.cnt[i] = 0;
    wave_sync();
    unsigned rank[LDS::kMB / KG_WAVE];
#pragma unroll
    for (unsigned j = 0; j < LDS::kMB / KG_WAVE; ++j) {
        unsigned const i = lane + j * KG_WAVE;
        rank[j] = (i < fill) ? atomicAdd(&L.cnt[L.tgt[i] & TGT_MASK], 1u) : 0u;
    }
    wave_sync();
    for (unsigned i = lane; i < nT; i += KG_WAVE) {
        unsigned const c = L.cnt[i];
        if (c) L.cnt[i] = atomicAdd(&a.ccnt[L.ids[i]], c);
    }
    wave_sync();
#pragma unroll
    for (unsigned j = 0; j < LDS::kMB / KG_WAVE; ++j) {
        unsigned const i = lane + j * KG_WAVE;
        if (i < fill) {
            // the key's flag bit says whether the slot's credit entry is
            // written (the merge sets the flag of every entry it admits)
            unsigned const tv = L.tgt[i], tp = tv & TGT_MASK;
            bool const cr = (tv & TGT_CREDIT) != 0u;
            uint32_t const target = L.ids[tp];
            uint64_t const key = L.key[i];
            put_cand(a, target, L.cnt[tp] + rank[j], cr ? key : (key & ~1ull), cr ? n : NND_NOSRC);
        }
    }
    wave_sync();
}

// Stages up to NP candidates per lane (bit k of `pm`: key[k] for list entry
// tgt[k]); the wave first counts them and flushes the buffer if they would not
// fit.  `fill` is wave-uniform; every lane of the wave calls this.
template <int NP, class LDS>
__device__ __forceinline__ void mj_stage(NndArgs const &a, LDS &L, unsigned &fill, unsigned nT, uint32_t n,
                                         unsigned pm, uint64_t const (&key)[NP], unsigned const (&tgt)[NP]) {
    uint64_t bal[NP];
    unsigned tot = 0;
#pragma unroll
    for (int k = 0; k < NP; ++k) {
        bal[k] = ballot((pm >> k) & 1u);
        tot += popc64(bal[k]);
    }
    if (tot == 0) return;
    if (fill + tot > LDS::kMB) {
        wave_sync();
        mj_flush(a, L, fill, nT, n);
        fill = 0;
    }
    uint64_t const lt = (1ull << lane_id()) - 1;
#pragma unroll
    for (int k = 0; k < NP; ++k) {
        if ((pm >> k) & 1u) {
            unsigned const pos = fill + popc64(bal[k] & lt);
            L.key[pos] = key[k];
            L.tgt[pos] = uint16_t(tgt[k]);
        }
        fill += popc64(bal[k]);
    }
}

// Fills the wave's list: new ids (sampled, then reverse), old ids (sampled,
// then reverse), capped at MT (kgraph.cpp:739-751 joins nn_new with itself
// and with nn_old, reverse lists merged in, :846-866).
template <class LDS>
__device__ __forceinline__ void mj_lists(NndArgs const &a, LDS &L, unsigned n, unsigned &nA, unsigned &nT,
                                         unsigned &nlist) {
    unsigned const lane = lane_id();
    unsigned const nn = a.nnew[n], no = a.nold[n];
    unsigned const rn = min(a.crnew[n], a.R), ro = min(a.crold[n], a.R);
    unsigned const nA0 = nn + rn, nB = no + ro;
    nT = min(nA0 + nB, MT);
    nA = min(nA0, nT);
    nlist = nA0 + nB;
    uint32_t const *samp = a.samp + size_t(n) * a.Lc;
    for (unsigned i = lane; i < nT; i += KG_WAVE) {
        uint32_t v;
        if (i < nn)
            v = samp[i];
        else if (i < nA0)
            v = a.rnew[size_t(n) * a.R + (i - nn)];
        else if (i < nA0 + no)
            v = samp[a.Lc - 1 - (i - nA0)];
        else
            v = a.rold[size_t(n) * a.R + (i - nA0 - no)];
        L.ids[i] = v;
    }
    wave_sync();
}

// Per-node statistics of a join wave.  The node's `found` (kgraph.cpp:801,
// 813) starts false here; the merge that follows sets it when one of the
// node's credited candidates turns out to be new to its list.
__device__ __forceinline__ void mj_finish(NndArgs const &a, unsigned n, unsigned pairs, unsigned pushes,
                                          unsigned nlist, unsigned rescored) {
    unsigned const lane = lane_id();
    for (int o = 32; o > 0; o >>= 1) {
        pairs += __shfl_xor(pairs, o);
        pushes += __shfl_xor(pushes, o);
        rescored += __shfl_xor(rescored, o);
    }
    if (lane == 0) {
        a.found[n] = 0u;
        a.part[NND_STAT_PAIRS * a.N + n] = pairs;
        a.part[NND_STAT_PUSHES * a.N + n] = pushes;
        a.part[NND_STAT_LISTS * a.N + n] = nlist;
        a.part[NND_STAT_RESCORE * a.N + n] = rescored;
    }
}

// The exact forms (integer rows, see above): BF16 operands, or 8-bit ones
// (I8: rows of the exact 8-bit copy; the records' join norms are then those
// of the shifted rows, nnd_launch_n8); the tile value is the canonical
// distance and decides alone.
template <int METRIC, int KS, bool I8>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4)))
join_mfma_kernel(NndArgs a) {
    using Ops = typename std::conditional<I8, MjOps8<KS>, MjOps<KS, false>>::type;
    // 16-row new groups held in registers at once: two for operands of up to
    // 128 values (16 VGPRs each as BF16, 8 as bytes), one otherwise
    constexpr unsigned AG = (I8 ? KS <= 2 : KS <= 4) ? 2u : 1u;
    __shared__ MJoinLDS lds[4];
    unsigned const lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    unsigned const n = a.n0 + blockIdx.x * 4 + w;
    if (n >= a.n1) return;
    auto &L = lds[w];
    unsigned nA, nT, nlist;
    mj_lists(a, L, n, nA, nT, nlist);
    unsigned const r16 = lane & 15, mrow = (lane >> 4) * 4;
    unsigned fill = 0, pairs = 0, pushes = 0;
    for (unsigned ag = 0; ag < nA; ag += 16 * AG) {
        // AG groups of 16 list rows from ag (new rows, and old ones when a
        // group runs past the new rows: their pairs are masked below) stay in
        // registers while every later tile streams past them once; a tile
        // that is one of the groups is taken from its registers
        Ops aop[AG];
        uint32_t aid[AG][4];
#pragma unroll
        for (unsigned g = 0; g < AG; ++g) {
            unsigned const ia = ag + 16 * g + r16;
            aop[g].load(a, ia < nT ? L.ids[ia] : 0u, ia < nT, lane);
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                unsigned const ib = ag + 16 * g + mrow + j;
                aid[g][j] = ib < nA ? L.ids[ib] : 0u;
            }
        }
        wave_sync();  // the previous groups' records are no longer read
        if (lane < 16 * AG) {
            unsigned const ia = ag + lane;
            L.am[lane] = ia < nA ? a.meta[L.ids[ia]] : make_float4(0.f, 0.f, 0.f, 0.f);
        }
        wave_sync();
        // a tile before ag holds new rows at lower list positions only: its
        // pairs with these rows were taken when it was the A group
        for (unsigned tg = ag; tg < nT; tg += 16) {
            unsigned const io = tg + r16;
            bool const ook = io < nT;
            uint32_t const oid = ook ? L.ids[io] : 0u;
            unsigned const dg = (tg - ag) >> 4;  // wave-uniform: the tile is A group dg when dg < AG
            Ops bop;
            if (dg == 0)
                bop = aop[0];
            else if (AG > 1 && dg == 1)
                bop = aop[AG - 1];
            else
                bop.load(a, oid, ook, lane);
            float4 const om = a.meta[oid];  // radius, K-th distance, join norm, |x|
            // one group at a time (a real loop: the groups' epilogues do not
            // overlap in registers)
#pragma unroll 1
            for (unsigned g = 0; g < AG; ++g) {
                unsigned const gb = ag + 16 * g;
                if (gb >= nA || gb > tg) continue;  // no new rows, or every pair below the diagonal
                uint32_t gid[4];
#pragma unroll
                for (int j = 0; j < 4; ++j) gid[j] = g ? aid[AG - 1][j] : aid[0][j];
                mj_f32x4 acc;
                if (g == 0)  // wave-uniform
                    acc = mj_mma(aop[0], bop);
                else
                    acc = mj_mma(aop[AG - 1], bop);
                unsigned pm = 0;
                uint64_t key[8];
                unsigned tgt[8];
                // the A rows' records are read from LDS at every tile: kept in
                // registers across the tile loop they would cost 16 VGPRs a group
                unsigned mb = 16 * g + mrow;
                asm volatile("" : "+v"(mb));
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    unsigned const ia = gb + mrow + j;
                    float4 const am = L.am[mb + j];
                    // new x new once per pair of list positions (the reference's
                    // i < j by id takes each id pair once, kgraph.cpp:741-745),
                    // new x old always (:746-748); an id never pairs with itself
                    bool const valid = ia < nA && ook && gid[j] != oid && (io >= nA || ia < io);
                    float const d = mj_dist<METRIC>(acc[j], am.z, om.z);
                    pairs += valid;
                    bool const pa = valid && d < am.x;
                    bool const po = valid && d < om.x;
                    // the callback's first endpoint (kgraph.cpp:741-750, 805-807):
                    // the smaller id of a new x new pair, the new row otherwise
                    bool const bfirst = io < nA && oid < gid[j];
                    pushes += unsigned(pa) + unsigned(po);
                    pm |= (unsigned(pa) << (2 * j)) | (unsigned(po) << (2 * j + 1));
                    key[2 * j] = make_key(d, oid, 1);
                    key[2 * j + 1] = make_key(d, gid[j], 1);
                    tgt[2 * j] = ia | ((!bfirst && d < am.y) ? TGT_CREDIT : 0u);
                    tgt[2 * j + 1] = io | ((bfirst && d < om.y) ? TGT_CREDIT : 0u);
                }
                mj_stage<8>(a, L, fill, nT, n, pm, key, tgt);
            }
        }
    }
    wave_sync();
    if (fill) mj_flush(a, L, fill, nT, n);
    mj_finish(a, n, pairs, pushes, nlist, 0u);
}

// The screen forms (BF16 operands, or FP32 operands on v_mfma_f32_16x16x4f32):
// the MFMA tile value minus its error bound E (mj_bound) decides which pairs
// could pass the exact test; only those are scored exactly, and the exact
// value alone is compared and stored.  The tile's 16 new rows are staged in
// LDS as FP32 (16 Dp floats per wave, dynamic shared memory); after a tile's
// screen, quad q takes list row tg + q when any of its pairs passed, loads
// that row's FP32 values once (lane t: 4-value chunks t, t + 4, ... of the
// canonical order, kg_device.hpp), and scores it against each passing new row
// read from LDS (the same chunks, one broadcast read per chunk), then the
// last D % 16 values one by one.  Chunk slots past the row's (D & ~15) / 4
// chunks hold zeros in both operands (they add fmaf(0, 0, p) = p), so the
// loads carry no per-chunk conditions.  CHM = chunk slots per lane.
template <int METRIC, int KS, bool F32>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu((F32 ? KS : 2 * KS) > 8 ? 2 : 3)))
join_screen_kernel(NndArgs a) {
    constexpr int CHM = F32 ? KS : 2 * KS;
    __shared__ SJoinLDS lds[4];
    extern __shared__ __align__(16) float4 arows[];
    unsigned const lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    unsigned const n = a.n0 + blockIdx.x * 4 + w;
    if (n >= a.n1) return;
    auto &L = lds[w];
    unsigned nA, nT, nlist;
    mj_lists(a, L, n, nA, nT, nlist);
    unsigned const r16 = lane & 15, mrow = (lane >> 4) * 4;
    unsigned const quad = lane >> 2, t = lane & 3;
    unsigned const dp4 = a.Dp / 4;
    unsigned const nch = (a.D & ~15u) >> 2, d16 = a.D & ~15u;
    float4 *const As = arows + size_t(w) * 16 * dp4;
    unsigned fill = 0, pairs = 0, pushes = 0, rescored = 0;
    for (unsigned ag = 0; ag < nA; ag += 16) {
        MjOps<KS, F32> aop;
        {
            unsigned const ia = ag + r16;
            aop.load(a, ia < nT ? L.ids[ia] : 0u, ia < nT, lane);
        }
        uint32_t aid[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            unsigned const ib = ag + mrow + j;
            aid[j] = ib < nA ? L.ids[ib] : 0u;
        }
        wave_sync();  // the previous group's rows and records are no longer read
        if (lane < 16) {
            unsigned const ia = ag + lane;
            L.am[lane] = ia < nA ? a.meta[L.ids[ia]] : make_float4(0.f, 0.f, 0.f, 0.f);
        }
        for (unsigned e = lane; e < 16 * dp4; e += KG_WAVE) {
            unsigned const r = e / dp4, c = e - r * dp4;
            unsigned const ia = ag + r;
            As[e] = ia < nA ? reinterpret_cast<float4 const *>(a.X + size_t(L.ids[ia]) * a.Dp)[c]
                            : make_float4(0.f, 0.f, 0.f, 0.f);
        }
        wave_sync();
        for (unsigned tg = ag; tg < nT; tg += 16) {
            unsigned const io = tg + r16;
            bool const ook = io < nT;
            uint32_t const oid = ook ? L.ids[io] : 0u;
            MjOps<KS, F32> bop;
            if (tg == ag)  // wave-uniform: the tile is the A group
                bop = aop;
            else
                bop.load(a, oid, ook, lane);
            float4 const om = a.meta[oid];
            mj_f32x4 const acc = mj_mma(aop, bop);
            // the screen: bit j of lane l passes pair (new row mrow + j, column r16)
            uint64_t bj[4];
            unsigned mb = mrow;
            asm volatile("" : "+v"(mb));
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                unsigned const ia = ag + mrow + j;
                float4 const am = L.am[mb + j];
                // new x new once per pair of list positions, new x old always
                // (kgraph.cpp:741-748); an id never pairs with itself
                bool const valid = ia < nA && ook && aid[j] != oid && (io >= nA || ia < io);
                pairs += valid;
                float const d = mj_dist<METRIC>(acc[j], am.z, om.z);
                float const lo = d - mj_bound<METRIC>(a, am.z, om.z, am.w, om.w);
                bj[j] = ballot(valid && (lo < am.x || lo < om.x));
            }
            if (!(bj[0] | bj[1] | bj[2] | bj[3])) continue;  // wave-uniform
            // quad q's column: its passing new rows 4 g + j sit in bit q + 16 g of bj[j]
            unsigned mq = 0;
#pragma unroll
            for (int g = 0; g < 4; ++g)
#pragma unroll
                for (int j = 0; j < 4; ++j) mq |= unsigned((bj[j] >> (quad + 16 * g)) & 1ull) << (4 * g + j);
            unsigned const cio = tg + quad;  // the quad's list row (valid whenever mq != 0)
            uint32_t const cid = mq ? L.ids[cio] : 0u;
            float const *crow = a.X + size_t(cid) * a.Dp;
            float4 reg[CHM];
#pragma unroll
            for (int k = 0; k < CHM; ++k) {
                unsigned const kk = t + 4 * k;
                bool const ok = kk < nch;
                float4 const v = reinterpret_cast<float4 const *>(crow)[ok ? kk : 0u];
                reg[k] = ok ? v : make_float4(0.f, 0.f, 0.f, 0.f);
            }
            float4 const cm = a.meta[cid];
            float const cn = (METRIC == METRIC_L2SQR) ? 0.f : cm.z;
            if (t == 0) rescored += __builtin_popcount(mq);
            while (ballot(mq != 0)) {
                bool const act = mq != 0;
                unsigned const r = act ? unsigned(__builtin_ctz(mq)) : 0u;
                mq &= mq - 1u;
                float4 const *ra4 = As + r * dp4;
                float4 pp = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
                for (int k = 0; k < CHM; ++k) {
                    unsigned const kk = t + 4 * k;
                    bool const ok = kk < nch;
                    float4 y = ra4[ok ? kk : 0u];
                    if (!ok) y = make_float4(0.f, 0.f, 0.f, 0.f);
                    float4 const x = reg[k];
                    if (METRIC == METRIC_L2SQR) {
                        float const dx = y.x - x.x, dy = y.y - x.y, dz = y.z - x.z, dw = y.w - x.w;
                        pp.x = __builtin_fmaf(dx, dx, pp.x);
                        pp.y = __builtin_fmaf(dy, dy, pp.y);
                        pp.z = __builtin_fmaf(dz, dz, pp.z);
                        pp.w = __builtin_fmaf(dw, dw, pp.w);
                    } else {
                        pp.x = __builtin_fmaf(y.x, x.x, pp.x);
                        pp.y = __builtin_fmaf(y.y, x.y, pp.y);
                        pp.z = __builtin_fmaf(y.z, x.z, pp.z);
                        pp.w = __builtin_fmaf(y.w, x.w, pp.w);
                    }
                }
                float acc2 = quad_reduce(pp);
                float const *rl = reinterpret_cast<float const *>(ra4);
                for (unsigned i = d16; i < a.D; ++i) {
                    if (METRIC == METRIC_L2SQR) {
                        float const dd = rl[i] - crow[i];
                        acc2 = __builtin_fmaf(dd, dd, acc2);
                    } else {
                        acc2 = __builtin_fmaf(rl[i], crow[i], acc2);
                    }
                }
                float4 const am = L.am[r];
                uint32_t const rid = L.ids[ag + r];
                float const na = (METRIC == METRIC_L2SQR) ? 0.f : am.z;
                float const d = finish_metric<METRIC>(acc2, na, cn);
                bool const lead = act && t == 0;
                bool const pa = lead && d < am.x;
                bool const po = lead && d < cm.x;
                bool const bfirst = cio < nA && cid < rid;  // the callback's first endpoint is the list row
                pushes += unsigned(pa) + unsigned(po);
                uint64_t const key[2] = {make_key(d, cid, 1), make_key(d, rid, 1)};
                unsigned const tgt[2] = {(ag + r) | ((!bfirst && d < am.y) ? TGT_CREDIT : 0u),
                                         cio | ((bfirst && d < cm.y) ? TGT_CREDIT : 0u)};
                mj_stage<2>(a, L, fill, nT, n, unsigned(pa) | (unsigned(po) << 1), key, tgt);
            }
        }
    }
    wave_sync();
    if (fill) mj_flush(a, L, fill, nT, n);
    mj_finish(a, n, pairs, pushes, nlist, rescored);
}

// The same tiles in exact FP32: each quad holds one list row of the tile in
// registers and computes its 16 pairs with the tile's new rows in canonical
// order (kg_device.hpp); lane t of a quad keeps the pairs of new rows t, t + 4,
// t + 8, t + 12, and staging / flushing is shared with the MFMA form.  With
// CH >= 0 the tile's 16 new rows are first copied into LDS (16 Dp floats per
// wave, dynamic shared memory), where all 16 quads read the same chunk in the
// same instruction (one broadcast read) instead of each fetching the rows
// through the cache; CH is then exactly the number of 4-value chunks per lane
// in the 16-value part of a row ((D & ~15) / 16), so the chunk loop has no
// per-chunk conditions (each would split the loop and wait on every LDS read
// by itself), and the last D % 16 values follow one by one.  CH < 0 (rows of
// more than 143 values) reads both rows from global memory (quad_metric).
template <int METRIC, int CH>
__global__ void __launch_bounds__(256) join_tile_kernel(NndArgs a) {
    __shared__ MJoinLDS lds[4];
    extern __shared__ __align__(16) float4 tile_rows[];
    unsigned const lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    unsigned const n = a.n0 + blockIdx.x * 4 + w;
    if (n >= a.n1) return;
    auto &L = lds[w];
    unsigned const nn = a.nnew[n], no = a.nold[n];
    unsigned const rn = min(a.crnew[n], a.R), ro = min(a.crold[n], a.R);
    unsigned const nA0 = nn + rn, nB = no + ro, nT = min(nA0 + nB, MT), nA = min(nA0, nT);
    uint32_t const *samp = a.samp + size_t(n) * a.Lc;
    for (unsigned i = lane; i < nT; i += KG_WAVE) {
        uint32_t v;
        if (i < nn)
            v = samp[i];
        else if (i < nA0)
            v = a.rnew[size_t(n) * a.R + (i - nn)];
        else if (i < nA0 + no)
            v = samp[a.Lc - 1 - (i - nA0)];
        else
            v = a.rold[size_t(n) * a.R + (i - nA0 - no)];
        L.ids[i] = v;
    }
    wave_sync();
    unsigned const quad = lane >> 2, t = lane & 3;
    unsigned const dp4 = a.Dp / 4;
    unsigned const d16 = a.D & ~15u;  // values covered by the chunk loop
    float4 *const As = tile_rows + size_t(w) * 16 * dp4;  // CH >= 0: the tile's new rows
    unsigned fill = 0, pairs = 0, pushes = 0;
    for (unsigned ag = 0; ag < nA; ag += 16) {
        if (CH >= 0) {
            wave_sync();  // the previous tile's reads are done
            for (unsigned e = lane; e < 16 * dp4; e += KG_WAVE) {
                unsigned const r = e / dp4, c = e - r * dp4;
                unsigned const ia = ag + r;
                As[e] = ia < nA ? reinterpret_cast<float4 const *>(a.X + size_t(L.ids[ia]) * a.Dp)[c]
                                : make_float4(0.f, 0.f, 0.f, 0.f);
            }
        }
        uint32_t aid[4];
        float arad[4], akth[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            unsigned const ia = ag + t + 4 * i;
            bool const ok = ia < nA;
            aid[i] = ok ? L.ids[ia] : 0u;
            float4 const m = a.meta[aid[i]];  // radius, K-th distance, join norm, |x|
            arad[i] = ok ? m.x : 0.f;
            akth[i] = ok ? m.y : 0.f;
        }
        if (lane < 16) L.an[lane] = (METRIC == METRIC_L2SQR || ag + lane >= nA) ? 0.f : a.meta[L.ids[ag + lane]].z;
        wave_sync();
        // a tile before ag holds new rows at lower list positions only: its
        // pairs with these rows were taken when it was the A group
        for (unsigned tg = ag; tg < nT; tg += 16) {
            unsigned const io = tg + quad;
            bool const ook = io < nT;
            uint32_t const oid = ook ? L.ids[io] : 0u;
            float const *orow = a.X + size_t(oid) * a.Dp;
            float4 reg[CH > 0 ? CH : 1];
#pragma unroll
            for (int k = 0; k < CH; ++k) reg[k] = reinterpret_cast<float4 const *>(orow)[t + 4 * k];
            float4 const om = a.meta[oid];
            float const orad = om.x;
            float const on = (METRIC == METRIC_L2SQR) ? 0.f : om.z;
            float dres[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                float4 const *ra4 = As + r * dp4;
                float acc;
                if (CH >= 0) {
                    float4 p = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
                    for (int k = 0; k < CH; ++k) {
                        float4 const y = ra4[t + 4 * k];
                        float4 const x = reg[k];
                        if (METRIC == METRIC_L2SQR) {
                            float const dx = y.x - x.x, dy = y.y - x.y, dz = y.z - x.z, dw = y.w - x.w;
                            p.x = __builtin_fmaf(dx, dx, p.x);
                            p.y = __builtin_fmaf(dy, dy, p.y);
                            p.z = __builtin_fmaf(dz, dz, p.z);
                            p.w = __builtin_fmaf(dw, dw, p.w);
                        } else {
                            p.x = __builtin_fmaf(y.x, x.x, p.x);
                            p.y = __builtin_fmaf(y.y, x.y, p.y);
                            p.z = __builtin_fmaf(y.z, x.z, p.z);
                            p.w = __builtin_fmaf(y.w, x.w, p.w);
                        }
                    }
                    acc = quad_reduce(p);
                    float const *rl = reinterpret_cast<float const *>(ra4);
                    for (unsigned i = d16; i < a.D; ++i) {
                        if (METRIC == METRIC_L2SQR) {
                            float const dd = rl[i] - orow[i];
                            acc = __builtin_fmaf(dd, dd, acc);
                        } else {
                            acc = __builtin_fmaf(rl[i], orow[i], acc);
                        }
                    }
                } else {
                    unsigned const ia = ag + r;
                    uint32_t const av = L.ids[ia < nA ? ia : 0u];
                    acc = quad_metric<METRIC>(a.X + size_t(av) * a.Dp, orow, a.D, t);
                }
                float const na = L.an[r];
                float const d = finish_metric<METRIC>(acc, na, on);
                if ((r & 3) == int(t)) dres[r >> 2] = d;
            }
            unsigned pm = 0;
            uint64_t key[8];
            unsigned tgt[8];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                unsigned const ia = ag + t + 4 * i;
                // new x new once per pair of list positions (the reference's
                // i < j by id takes each id pair once, kgraph.cpp:741-745),
                // new x old always (:746-748); an id never pairs with itself
                bool const valid = ia < nA && ook && aid[i] != oid && (io >= nA || ia < io);
                float const d = dres[i];
                bool const pa = valid && d < arad[i];
                bool const po = valid && d < orad;
                // the callback's first endpoint (kgraph.cpp:741-750, 805-807):
                // the smaller id of a new x new pair, the new row otherwise
                bool const bfirst = io < nA && oid < aid[i];
                pairs += valid;
                pushes += unsigned(pa) + unsigned(po);
                pm |= (unsigned(pa) << (2 * i)) | (unsigned(po) << (2 * i + 1));
                key[2 * i] = make_key(d, oid, 1);
                key[2 * i + 1] = make_key(d, aid[i], 1);
                tgt[2 * i] = ia | ((!bfirst && d < akth[i]) ? TGT_CREDIT : 0u);
                tgt[2 * i + 1] = io | ((bfirst && d < om.y) ? TGT_CREDIT : 0u);
            }
            mj_stage<8>(a, L, fill, nT, n, pm, key, tgt);
        }
    }
    wave_sync();
    if (fill) mj_flush(a, L, fill, nT, n);
    mj_finish(a, n, pairs, pushes, nA0 + nB, 0u);
}

// Row norms for the MFMA join: (|x|^2 in canonical order, its square root),
// and what decides the exact form: whether every value is an integer
// (flags[0] bit 0 set when one is not), max |x| (flags[1], float bits), the
// smallest and largest value (flags[2] / flags[3], order-preserving bits
// ord_f32; the caller sets flags[2] to all ones and the rest to zero).
__global__ void __launch_bounds__(256) sqn_kernel(float const *X, unsigned N, unsigned D, unsigned Dp, float2 *out,
                                                  uint32_t *flags) {
    unsigned const lane = threadIdx.x & 63, t = lane & 3;
    unsigned const row = (blockIdx.x * blockDim.x + threadIdx.x) >> 2;
    bool const ok = row < N;
    float const *r = X + size_t(ok ? row : 0u) * Dp;
    float const n2 = quad_metric<METRIC_EUCLIDEAN, float, true>(r, r, D, t);
    bool frac = false;
    float mx = 0.f, lo = __builtin_huge_valf(), hi = -__builtin_huge_valf();
    if (ok)
        for (unsigned i = t; i < D; i += 4) {
            float const v = r[i];
            frac |= !(v == rintf(v));
            mx = fmaxf(mx, fabsf(v));
            lo = fminf(lo, v);
            hi = fmaxf(hi, v);
        }
    if (ok && t == 0) out[row] = make_float2(n2, sqrtf(n2));
    // atomics only while they can still change the flags
    if (ballot(frac) && lane == 0 && !__hip_atomic_load(flags, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
        atomicOr(flags, 1u);
    for (int o = 32; o > 0; o >>= 1) {
        mx = fmaxf(mx, __shfl_xor(mx, o));
        lo = fminf(lo, __shfl_xor(lo, o));
        hi = fmaxf(hi, __shfl_xor(hi, o));
    }
    if (lane == 0) {
        if (__float_as_uint(mx) > __hip_atomic_load(flags + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
            atomicMax(flags + 1, __float_as_uint(mx));
        if (ord_f32(lo) < __hip_atomic_load(flags + 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
            atomicMin(flags + 2, ord_f32(lo));
        if (ord_f32(hi) > __hip_atomic_load(flags + 3, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
            atomicMax(flags + 3, ord_f32(hi));
    }
}

// Per-node join record: z = the norm the join's distance formula takes (|x|^2
// for SSD on the MFMA form, else the stored metric norm, 0 for SSD on FP32
// tiles), w = |x| (the screen bound; MFMA form only).  x, y (radius, K-th
// distance) are written by init_kernel and merge_kernel.
// |x - 128|^2 over the D columns of an 8-bit row (half of it for half-SSD)
// into the join norm of its record
__global__ void __launch_bounds__(256) n8_kernel(uint8_t const *X8, unsigned N, unsigned D, unsigned Dp, int half,
                                                 float4 *meta) {
    unsigned const n = blockIdx.x * 256 + threadIdx.x;
    if (n >= N) return;
    uint32_t const *r = reinterpret_cast<uint32_t const *>(X8 + size_t(n) * Dp);  // Dp % 4 == 0
    int sum = 0;
    for (unsigned w = 0; 4 * w < D; ++w) {
        uint32_t const word = r[w];
        for (unsigned j = 0; j < 4 && 4 * w + j < D; ++j) {
            int const v = int((word >> (8 * j)) & 0xffu) - 128;
            sum += v * v;
        }
    }
    meta[n].z = half ? 0.5f * float(sum) : float(sum);
}

__global__ void meta_init_kernel(float4 *meta, unsigned N, int metric, float2 const *sqn, float const *norms) {
    unsigned const n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    float z = 0.f, w = 0.f;
    if (sqn) w = sqn[n].y;
    if (metric != METRIC_L2SQR)
        z = norms[n];
    else if (sqn)
        z = sqn[n].x;
    meta[n] = make_float4(__builtin_huge_valf(), __builtin_huge_valf(), z, w);
}

__global__ void iter_begin_kernel(uint64_t *stats, uint32_t *iter, int advance) {
    if (threadIdx.x < NND_NSTATS) stats[threadIdx.x] = 0;
    if (threadIdx.x == 0 && advance) *iter += 1u;
}

// Per-node statistics (join and merge write one value per node and column,
// plain stores) summed into the NND_NSTATS counters: a block per chunk of
// nodes, one atomic per block and column.  Same-address device-scope atomics
// from every wave are what this avoids.
__global__ void __launch_bounds__(256) stats_reduce_kernel(uint32_t const *part, unsigned N, uint64_t *stats) {
    __shared__ unsigned long long sh[NND_NPART][4];
    unsigned const lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    unsigned long long acc[NND_NPART];
#pragma unroll
    for (int k = 0; k < NND_NPART; ++k) acc[k] = 0;
    for (unsigned n = blockIdx.x * blockDim.x + threadIdx.x; n < N; n += gridDim.x * blockDim.x) {
#pragma unroll
        for (int k = 0; k < NND_NPART; ++k) acc[k] += part[size_t(k) * N + n];
    }
#pragma unroll
    for (int k = 0; k < NND_NPART; ++k) {
        unsigned long long v = acc[k];
        for (int o = 32; o > 0; o >>= 1) {
            unsigned lo = __shfl_xor(unsigned(v), o), hi = __shfl_xor(unsigned(v >> 32), o);
            v += (static_cast<unsigned long long>(hi) << 32) | lo;
        }
        if (lane == 0) sh[k][w] = v;
    }
    __syncthreads();
    if (threadIdx.x < NND_NPART) {
        unsigned long long v = sh[threadIdx.x][0] + sh[threadIdx.x][1] + sh[threadIdx.x][2] + sh[threadIdx.x][3];
        if (v) atomicAdd(reinterpret_cast<unsigned long long *>(stats) + threadIdx.x, v);
    }
}

// ------------------------------------------------------------------ merge
__device__ __forceinline__ uint64_t shfl_xor64(uint64_t v, int m) {
    uint32_t lo = __shfl_xor(uint32_t(v), m), hi = __shfl_xor(uint32_t(v >> 32), m);
    return (uint64_t(hi) << 32) | lo;
}
__device__ __forceinline__ uint64_t shfl_up64(uint64_t v, int d) {
    uint32_t lo = __shfl_up(uint32_t(v), d), hi = __shfl_up(uint32_t(v >> 32), d);
    return (uint64_t(hi) << 32) | lo;
}

// Bitonic sort of 64*PER keys held in registers, blocked layout: element
// i = lane*PER + r.  Stages whose partner distance is below PER stay in the
// lane; the others exchange with lane ^ (j/PER) through shuffles.  All loops
// have compile-time trip counts, so every register index is a constant.
template <int PER>
__device__ __forceinline__ void wave_sort_regs(uint64_t (&v)[PER]) {
    unsigned const lane = lane_id();
    constexpr int LOGN = (PER == 1) ? 6 : (PER == 2) ? 7 : (PER == 4) ? 8 : (PER == 8) ? 9 : 10;
#pragma unroll
    for (int kk = 1; kk <= LOGN; ++kk) {
        unsigned const k = 1u << kk;
#pragma unroll
        for (int jj = kk - 1; jj >= 0; --jj) {
            unsigned const j = 1u << jj;
            if (j >= unsigned(PER)) {
                unsigned const lm = j / PER;
                bool const lower = (lane & lm) == 0;
#pragma unroll
                for (int r = 0; r < PER; ++r) {
                    unsigned const i = lane * PER + r;
                    uint64_t const other = shfl_xor64(v[r], int(lm));
                    bool const up = (i & k) == 0;
                    bool const keep_min = (lower == up);
                    uint64_t const mn = other < v[r] ? other : v[r];
                    uint64_t const mx = other < v[r] ? v[r] : other;
                    v[r] = keep_min ? mn : mx;
                }
            } else {
#pragma unroll
                for (int r = 0; r < PER; ++r) {
                    int const p = r ^ int(j);
                    if (r < p) {
                        unsigned const i = lane * PER + r;
                        bool const up = (i & k) == 0;
                        uint64_t const x = v[r], y = v[p];
                        bool const sw = (x > y) == up;
                        v[r] = sw ? y : x;
                        v[p] = sw ? x : y;
                    }
                }
            }
        }
    }
}

// Sorted pool U candidates -> first L distinct keys (kgraph.cpp:176-202).
// Equal keys are equal ids (an id's distance is the same wherever it is
// computed, canonical FP32), so a candidate whose id is in the pool, or came
// earlier, is dropped first: an LDS hash set of ids, pool ids inserted before
// candidates, so an existing entry keeps its flag.  Only the new distinct keys
// are then sorted (wave bitonic in registers: ~175 of a node's ~320
// candidates on the 1M-node build, instead of sorting pool and candidates
// together) and merged with the sorted pool by ranks: a new key lands at its
// index plus the number of pool keys below it (binary search in LDS), a pool
// key at its index plus the number of new keys below it (a histogram of the
// new keys' ranks, prefix-summed).  The same L keys as a sort of the union.
constexpr unsigned MG_TAB = 2048;  // hash slots per wave: load <= 1/2 for the NND_MERGE_MAX keys
constexpr uint32_t MG_EMPTY = 0xffffffffu;

__host__ __device__ inline size_t merge_lds_per_wave(unsigned L) {
    return 4ull * MG_TAB + ((8ull * L + 15) & ~size_t(15)) + ((4ull * (L + 1) + 15) & ~size_t(15));
}

// true when id was not in the set (it is now)
__device__ __forceinline__ bool mg_insert(uint32_t *tab, uint32_t id) {
    unsigned h = (id * 2654435761u) >> (32 - 11);
    for (;;) {
        uint32_t const old = atomicCAS(&tab[h], MG_EMPTY, id);
        if (old == MG_EMPTY) return true;
        if (old == id) return false;
        h = (h + 1) & (MG_TAB - 1);
    }
}

template <int PER>
__device__ void merge_new_keys(NndArgs const &a, unsigned n, unsigned c, unsigned u, uint64_t const *stage,
                               uint64_t const *pk, uint32_t *hist) {
    unsigned const lane = lane_id();
    uint64_t *pool = a.pool + size_t(n) * a.L;
    uint64_t v[PER];
#pragma unroll
    for (int r = 0; r < PER; ++r) {
        unsigned const i = lane * PER + r;
        v[r] = i < u ? stage[i] : ~0ull;
    }
    wave_sort_regs<PER>(v);
    unsigned flags = 0;
    float rad = 0.f, kth = 0.f;
    bool has_rad = false, has_kth = false;
    auto place = [&](unsigned pos, uint64_t key) {
        if (pos < a.L) pool[pos] = key;
        if (pos < a.K) flags += key_flag(key);
        if (pos == a.L - 1) {
            rad = key_dist(key);
            has_rad = true;
        }
        if (pos == a.K - 1) {
            kth = key_dist(key);
            has_kth = true;
        }
    };
    // new keys: index + pool keys below (lower bound in the sorted pool)
#pragma unroll
    for (int r = 0; r < PER; ++r) {
        unsigned const i = lane * PER + r;
        if (i < u) {
            uint64_t const e = v[r];
            unsigned lo = 0, hi = c;
            while (lo < hi) {
                unsigned const mid = (lo + hi) >> 1;
                if (pk[mid] < e)
                    lo = mid + 1;
                else
                    hi = mid;
            }
            atomicAdd(&hist[lo], 1u);
            place(i + lo, e);
        }
    }
    wave_sync();
    // pool keys: index + new keys below (prefix of the rank histogram)
    unsigned carry = 0;
    for (unsigned b = 0; b < c; b += KG_WAVE) {
        unsigned const i = b + lane;
        unsigned x = i < c ? hist[i] : 0u;
        for (int d = 1; d < 64; d <<= 1) {
            unsigned const y = __shfl_up(x, d);
            if (int(lane) >= d) x += y;
        }
        if (i < c) place(i + carry + x, pk[i]);
        carry += __shfl(x, 63);
    }
    unsigned const newc = min(c + u, a.L);
    for (int o = 32; o > 0; o >>= 1) flags += __shfl_xor(flags, o);
    uint64_t const br = ballot(has_rad), bk = ballot(has_kth);
    float const rad_w = br ? __shfl(rad, int(ctz64(br))) : __builtin_huge_valf();
    float const kth_w = bk ? __shfl(kth, int(ctz64(bk))) : __builtin_huge_valf();
    if (lane == 0) {
        a.cnt[n] = newc;
        a.ccnt[n] = 0;
        *reinterpret_cast<float2 *>(&a.meta[n]) =
            make_float2((newc >= a.L) ? rad_w : __builtin_huge_valf(), (newc >= a.K) ? kth_w : __builtin_huge_valf());
        a.part[NND_STAT_DELTA * a.N + n] = flags;
        a.part[NND_STAT_M * a.N + n] = a.Mv[n];
    }
}

__global__ void __launch_bounds__(256) merge_kernel(NndArgs a) {
    extern __shared__ __align__(16) unsigned char mg_smem[];
    unsigned const w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    unsigned const n = blockIdx.x * (blockDim.x >> 6) + w;
    if (n >= a.N) return;
    unsigned char *base = mg_smem + size_t(w) * merge_lds_per_wave(a.L);
    uint32_t *tab = reinterpret_cast<uint32_t *>(base);           // hash set, then the staged new keys
    uint64_t *stage = reinterpret_cast<uint64_t *>(base);
    uint64_t *pk = reinterpret_cast<uint64_t *>(base + 4ull * MG_TAB);
    uint32_t *hist = reinterpret_cast<uint32_t *>(base + 4ull * MG_TAB + ((8ull * a.L + 15) & ~size_t(15)));
    unsigned const c = a.cnt[n];
    unsigned const cc = a.ccnt[n];
    unsigned const mall = min(cc, a.C);                        // claimed slots
    bool const over = cc > a.C;
    unsigned const mtot = mall + (over ? a.O : 0u);            // + the overflow area
    unsigned const m = min(mtot, NND_MERGE_MAX - c);  // at most NND_MERGE_MAX keys per merge
    uint64_t *pool = a.pool + size_t(n) * a.L;
    uint64_t const *cand = a.cand + size_t(n) * a.C;
    uint64_t *ovf = a.ovf + size_t(n) * a.O;
    // candidates, 64 a round, all loads issued first; the fresh ones stay in
    // registers until the hash set is done with (the staging area reuses its
    // memory)
    constexpr int MR = NND_MERGE_MAX / KG_WAVE;
    uint64_t held[MR];
    uint32_t hsrc[MR];  // the credit entries of the claimed slots, loaded with them (read only for flagged keys)
    uint32_t const *csrc = a.csrc + size_t(n) * a.C;
#pragma unroll
    for (int r = 0; r < MR; ++r) {
        unsigned const j = unsigned(r) * KG_WAVE + lane;
        held[r] = ~0ull;
        hsrc[r] = NND_NOSRC;
        if (unsigned(r) * KG_WAVE < m && j < m) {
            held[r] = (j < mall) ? cand[j] : ovf[j - mall];
            if (j < mall) hsrc[r] = csrc[j];
        }
    }
    constexpr int PR = 4;  // pool keys loaded up front (the first 256); a longer pool reads the rest after
    uint64_t pr[PR];
#pragma unroll
    for (int r = 0; r < PR; ++r) {
        unsigned const i = unsigned(r) * KG_WAVE + lane;
        pr[r] = i < c ? pool[i] : ~0ull;
    }
    for (unsigned i = lane; i < MG_TAB; i += KG_WAVE) tab[i] = MG_EMPTY;
    for (unsigned i = lane; i <= c; i += KG_WAVE) hist[i] = 0u;
    wave_sync();
#pragma unroll
    for (int r = 0; r < PR; ++r) {
        unsigned const i = unsigned(r) * KG_WAVE + lane;
        if (i < c) {
            pk[i] = pr[r];
            mg_insert(tab, key_id(pr[r]));
        }
    }
    for (unsigned i = PR * KG_WAVE + lane; i < c; i += KG_WAVE) {
        uint64_t const k = pool[i];
        pk[i] = k;
        mg_insert(tab, key_id(k));
    }
    wave_sync();
    uint64_t fresh_mask = 0;  // bit r: held[r] is a new key
#pragma unroll
    for (int r = 0; r < MR; ++r)
        if (unsigned(r) * KG_WAVE < m && held[r] != ~0ull && mg_insert(tab, key_id(held[r]))) fresh_mask |= 1ull << r;
    // a fresh candidate from a claimed slot credits the node whose join sent
    // it, when that join marked it (flag bit set, source in csrc): the
    // reference's `found` (kgraph.cpp:805-813), an insert into the first K
    // that is no duplicate
#pragma unroll
    for (int r = 0; r < MR; ++r) {
        unsigned const j = unsigned(r) * KG_WAVE + lane;
        if (((fresh_mask >> r) & 1u) && j < mall && key_flag(held[r])) a.found[hsrc[r]] = 1u;
    }
    if (over) {  // the overflow area was used: empty it for the next join
        wave_sync();
        for (unsigned i = lane; i < a.O; i += KG_WAVE) ovf[i] = ~0ull;
    }
    wave_sync();
    unsigned u = 0;
    uint64_t const lt = (1ull << lane) - 1;
#pragma unroll
    for (int r = 0; r < MR; ++r) {
        if (unsigned(r) * KG_WAVE < m) {
            bool const f = (fresh_mask >> r) & 1u;
            uint64_t const b = ballot(f);
            if (f) stage[u + popc64(b & lt)] = held[r] | 1ull;  // admitted entries are new (flag set)
            u += popc64(b);
        }
    }
    wave_sync();
    if (u <= 64)
        merge_new_keys<1>(a, n, c, u, stage, pk, hist);
    else if (u <= 128)
        merge_new_keys<2>(a, n, c, u, stage, pk, hist);
    else if (u <= 256)
        merge_new_keys<4>(a, n, c, u, stage, pk, hist);
    else if (u <= 512)
        merge_new_keys<8>(a, n, c, u, stage, pk, hist);
    else
        merge_new_keys<16>(a, n, c, u, stage, pk, hist);
}

// ------------------------------------------------------------------ update
__global__ void __launch_bounds__(256) update_m_kernel(NndArgs a) {
    unsigned const w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    unsigned const n = blockIdx.x * 4 + w;
    if (n >= a.N) return;
    uint64_t const *pool = a.pool + size_t(n) * a.L;
    unsigned M = a.Mv[n];
    if (a.found[n]) {  // kgraph.cpp:829-838
        unsigned const maxl = min(M + a.S, a.cnt[n]);
        unsigned c = 0, l = maxl;
        for (unsigned b = 0; b < maxl; b += KG_WAVE) {
            unsigned i = b + lane;
            uint64_t bal = ballot(i < maxl && key_flag(pool[i]));
            unsigned nb = popc64(bal);
            if (c + nb >= a.S) {
                // position of the (S-c)-th set bit, plus one
                unsigned need = a.S - c;
                uint64_t x = bal;
                for (unsigned q = 1; q < need; ++q) x &= x - 1;
                l = b + ctz64(x) + 1;
                break;
            }
            c += nb;
        }
        M = l;
    }
    if (lane == 0) {
        a.Mv[n] = M;
        a.radiusM[n] = key_dist(pool[M - 1]);  // kgraph.cpp:840
        a.crnew[n] = 0;
        a.crold[n] = 0;
    }
}

__device__ __forceinline__ void push_rev(uint32_t *list, uint32_t *count, unsigned R, uint32_t target, uint32_t src,
                                         uint32_t it) {
    unsigned slot = atomicAdd(&count[target], 1u);
    if (slot < R) {
        list[size_t(target) * R + slot] = src;
    } else {
        unsigned j = unsigned(mix64((uint64_t(target) << 32 | src) ^ (uint64_t(it) << 48)) % (uint64_t(slot) + 1));
        if (j < R) list[size_t(target) * R + j] = src;
    }
}

__global__ void __launch_bounds__(256) sample_kernel(NndArgs a) {
    unsigned const w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    unsigned const n = blockIdx.x * 4 + w;
    if (n >= a.N) return;
    uint64_t *pool = a.pool + size_t(n) * a.L;
    uint32_t *samp = a.samp + size_t(n) * a.Lc;
    unsigned const M = a.Mv[n];
    uint32_t const it = *a.iter;
    unsigned nn = 0, no = 0;
    uint64_t const lt = (1ull << lane) - 1;
    for (unsigned b = 0; b < M; b += KG_WAVE) {
        unsigned i = b + lane;
        bool const act = i < M;
        uint64_t k = act ? pool[i] : 0;
        bool const isnew = act && key_flag(k);
        bool const isold = act && !key_flag(k);
        uint64_t bn = ballot(isnew), bo = ballot(isold);
        uint32_t const id = key_id(k);
        if (isnew) {
            samp[nn + popc64(bn & lt)] = id;
            pool[i] = k & ~1ull;
        }
        if (isold) samp[a.Lc - 1 - (no + popc64(bo & lt))] = id;
        if (act && key_dist(k) > a.radiusM[id]) {
            if (isnew)
                push_rev(a.rnew, a.crnew, a.R, id, n, it);
            else
                push_rev(a.rold, a.crold, a.R, id, n, it);
        }
        nn += popc64(bn);
        no += popc64(bo);
    }
    if (lane == 0) {
        a.nnew[n] = nn;
        a.nold[n] = no;
    }
}

// ------------------------------------------------------------------ learned pre-rank
// The join's lists of node n pruned by a learned projection (the projector's
// rows Z, zdim floats each): every list entry (new sample, reverse new, old
// sample, reverse old) is ranked by its projected squared distance to n, ties
// by list position, and the first ceil(keep * total) are kept, each sub-list
// compacted in order in place; the join then scores only pairs of kept
// entries.  Full-dimension FP32 distances still judge every admission.
__global__ void __launch_bounds__(256) prerank_kernel(NndArgs a) {
    __shared__ uint64_t key[4][MT];
    __shared__ float zn[4][64];
    unsigned const w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    unsigned const n = blockIdx.x * 4 + w;
    if (n >= a.N) return;
    unsigned const nn = a.nnew[n], no = a.nold[n];
    unsigned const rn = min(a.crnew[n], a.R), ro = min(a.crold[n], a.R);
    unsigned const total = min(nn + rn + no + ro, MT);
    uint32_t *const samp = a.samp + size_t(n) * a.Lc;
    uint32_t *const rnew = a.rnew + size_t(n) * a.R;
    uint32_t *const rold = a.rold + size_t(n) * a.R;
    // entry i of the list (the join's order: new sample, reverse new, old sample, reverse old)
    auto entry = [&](unsigned i) -> uint32_t {
        if (i < nn) return samp[i];
        if (i < nn + rn) return rnew[i - nn];
        if (i < nn + rn + no) return samp[a.Lc - 1 - (i - nn - rn)];
        return rold[i - nn - rn - no];
    };
    unsigned const P = a.zdim;
    for (unsigned k = lane; k < P; k += KG_WAVE) zn[w][k] = a.Z[size_t(n) * P + k];
    wave_sync();
    for (unsigned i = lane; i < total; i += KG_WAVE) {
        float const *zi = a.Z + size_t(entry(i)) * P;
        float d = 0.f;
        for (unsigned k = 0; k < P; ++k) {
            float const t = zi[k] - zn[w][k];
            d = __builtin_fmaf(t, t, d);
        }
        key[w][i] = (uint64_t(ord_f32(d)) << 32) | i;
    }
    wave_sync();
    unsigned const kc = min(total, unsigned(ceilf(a.keep * float(total))));
    // keep flags by rank, then in-order compaction of each sub-list
    uint32_t ids[MT / KG_WAVE];
    bool kept[MT / KG_WAVE];
#pragma unroll
    for (unsigned c = 0; c < MT / KG_WAVE; ++c) {
        unsigned const i = c * KG_WAVE + lane;
        kept[c] = false;
        ids[c] = 0;
        if (i < total) {
            uint64_t const v = key[w][i];
            unsigned r = 0;
            for (unsigned j = 0; j < total; ++j) r += key[w][j] < v;
            kept[c] = r < kc;
            ids[c] = entry(i);
        }
    }
    wave_sync();  // every entry is read before the lists are rewritten
    unsigned const bound[4] = {nn, nn + rn, nn + rn + no, total};
    unsigned cnt[4] = {0, 0, 0, 0};
    uint64_t const lt = (1ull << lane) - 1;
#pragma unroll
    for (unsigned c = 0; c < MT / KG_WAVE; ++c) {
        unsigned const i = c * KG_WAVE + lane;
#pragma unroll
        for (unsigned part = 0; part < 4; ++part) {
            bool const in = i < total && i >= (part ? bound[part - 1] : 0u) && i < bound[part];
            bool const k = in && kept[c];
            uint64_t const b = ballot(k);
            if (k) {
                unsigned const pos = cnt[part] + popc64(b & lt);
                if (part == 0) samp[pos] = ids[c];
                else if (part == 1) rnew[pos] = ids[c];
                else if (part == 2) samp[a.Lc - 1 - pos] = ids[c];
                else rold[pos] = ids[c];
            }
            cnt[part] += popc64(b);
        }
    }
    if (lane == 0) {
        a.nnew[n] = cnt[0];
        a.crnew[n] = cnt[1];
        a.nold[n] = cnt[2];
        a.crold[n] = cnt[3];
    }
}

// ------------------------------------------------------------------ extraction
__global__ void extract_kernel(NndArgs a, uint32_t const *off, uint32_t *ids, float *dists, uint32_t *M) {
    unsigned const w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    unsigned const n = blockIdx.x * 4 + w;
    if (n >= a.N) return;
    uint64_t const *pool = a.pool + size_t(n) * a.L;
    unsigned const c = a.cnt[n];
    uint32_t const o = off[n];
    for (unsigned i = lane; i < c; i += KG_WAVE) {
        ids[o + i] = key_id(pool[i]);
        dists[o + i] = key_dist(pool[i]);
    }
    if (lane == 0) M[n] = a.Mv[n];
}

// control pools -> compact (ids, dists) for the host recall estimate
__global__ void gather_controls_kernel(NndArgs a, uint32_t const *ctl, unsigned nc, float *dout, uint32_t *cout) {
    unsigned const c = blockIdx.x;
    if (c >= nc) return;
    uint32_t const n = ctl[c];
    uint64_t const *pool = a.pool + size_t(n) * a.L;
    unsigned const cnt = a.cnt[n];
    for (unsigned i = threadIdx.x; i < a.L; i += blockDim.x) dout[size_t(c) * a.L + i] = (i < cnt) ? key_dist(pool[i]) : 0.f;
    if (threadIdx.x == 0) cout[c] = cnt;
}


// The join's BF16 MFMA distance for arbitrary pairs (diagnostics of the
// screen): one wave per 16 pairs, pair p's rows as row p of A and column p of
// B, so C[p][p] is its dot; lane l holds C[4(l/16) + j][l % 16], the diagonal
// sits in the lanes with l / 16 == (l % 16) / 4 at j = l % 4.
template <int METRIC, int KS>
__global__ void __launch_bounds__(256) screen_pairs_kernel(NndArgs a, uint32_t const *pairs, unsigned n, float *out_d,
                                                           float *out_e) {
    unsigned const lane = threadIdx.x & 63;
    unsigned const base = ((blockIdx.x * blockDim.x + threadIdx.x) >> 6) * 16;
    if (base >= n) return;
    unsigned const r16 = lane & 15, k8 = (lane >> 4) * 8;
    unsigned const p = base + r16;
    bool const ok = p < n;
    uint32_t const ia = ok ? pairs[2 * p] : 0u, ib = ok ? pairs[2 * p + 1] : 0u;
    mj_bf16x8 aop[KS], bop[KS];
    mj_load<KS>(a, ia, ok, k8, aop);
    mj_load<KS>(a, ib, ok, k8, bop);
    mj_f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int s = 0; s < KS; ++s) acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(aop[s], bop[s], acc, 0, 0, 0);
    unsigned const j = r16 & 3;
    float const G = j == 0 ? acc[0] : j == 1 ? acc[1] : j == 2 ? acc[2] : acc[3];
    if (ok && (lane >> 4) == (r16 >> 2)) {
        float2 const sa = a.sqn[ia], sb = a.sqn[ib];
        float const an = (METRIC == METRIC_L2SQR) ? sa.x : a.norms[ia];
        float const on = (METRIC == METRIC_L2SQR) ? sb.x : a.norms[ib];
        out_d[p] = mj_dist<METRIC>(G, an, on);
        out_e[p] = mj_bound<METRIC>(a, an, on, sa.y, sb.y);
    }
}

}  // namespace

// join_tile_kernel's CH: 4-value chunks per lane in the 16-value part of a
// row, or -1 when that exceeds 8 (rows read from global memory)
static int ch_for(unsigned D) {
    unsigned const per = (D & ~15u) >> 4;
    return per <= 8 ? int(per) : -1;
}

#define KG_DISPATCH_METRIC(metric, KERNEL, ...)                                       \
    switch (metric) {                                                                \
        case METRIC_L2SQR: hipLaunchKernelGGL(KERNEL<METRIC_L2SQR>, __VA_ARGS__); break; \
        case METRIC_EUCLIDEAN: hipLaunchKernelGGL(KERNEL<METRIC_EUCLIDEAN>, __VA_ARGS__); break; \
        default: hipLaunchKernelGGL(KERNEL<METRIC_ANGULAR>, __VA_ARGS__); break;           \
    }

hipError_t nnd_launch_init(int metric, NndArgs const &a, hipStream_t s) {
    dim3 g((a.N + 3) / 4), b(256);
    KG_DISPATCH_METRIC(metric, init_kernel, g, b, 0, s, a);
    return hipGetLastError();
}

template <int METRIC>
static void launch_join_m(NndArgs const &a, hipStream_t s) {
    dim3 g((a.n1 - a.n0 + 3) / 4), b(256);
    int const ch = ch_for(a.D);
    // CH >= 0: 16 staged rows of Dp floats per wave
    size_t const lds = ch >= 0 ? size_t(4) * 16 * a.Dp * 4 : 0;
    switch (ch) {
        case 0: hipLaunchKernelGGL((join_tile_kernel<METRIC, 0>), g, b, lds, s, a); break;
        case 1: hipLaunchKernelGGL((join_tile_kernel<METRIC, 1>), g, b, lds, s, a); break;
        case 2: hipLaunchKernelGGL((join_tile_kernel<METRIC, 2>), g, b, lds, s, a); break;
        case 3: hipLaunchKernelGGL((join_tile_kernel<METRIC, 3>), g, b, lds, s, a); break;
        case 4: hipLaunchKernelGGL((join_tile_kernel<METRIC, 4>), g, b, lds, s, a); break;
        case 5: hipLaunchKernelGGL((join_tile_kernel<METRIC, 5>), g, b, lds, s, a); break;
        case 6: hipLaunchKernelGGL((join_tile_kernel<METRIC, 6>), g, b, lds, s, a); break;
        case 7: hipLaunchKernelGGL((join_tile_kernel<METRIC, 7>), g, b, lds, s, a); break;
        case 8: hipLaunchKernelGGL((join_tile_kernel<METRIC, 8>), g, b, lds, s, a); break;
        default: hipLaunchKernelGGL((join_tile_kernel<METRIC, -1>), g, b, lds, s, a); break;
    }
}

template <int METRIC>
static void launch_join_exact(NndArgs const &a, hipStream_t s) {
    dim3 g((a.n1 - a.n0 + 3) / 4), b(256);
    if constexpr (METRIC != METRIC_ANGULAR) {
        if (a.i8) {
            unsigned const ks8 = (a.D + 63) / 64;  // 64-column 8-bit k-steps
            if (ks8 <= 1)
                hipLaunchKernelGGL((join_mfma_kernel<METRIC, 1, true>), g, b, 0, s, a);
            else if (ks8 <= 2)
                hipLaunchKernelGGL((join_mfma_kernel<METRIC, 2, true>), g, b, 0, s, a);
            else
                hipLaunchKernelGGL((join_mfma_kernel<METRIC, 4, true>), g, b, 0, s, a);
            return;
        }
    }
    unsigned const ks = (a.D + 31) / 32;  // 32-column MFMA k-steps
    if (ks <= 1)
        hipLaunchKernelGGL((join_mfma_kernel<METRIC, 1, false>), g, b, 0, s, a);
    else if (ks <= 2)
        hipLaunchKernelGGL((join_mfma_kernel<METRIC, 2, false>), g, b, 0, s, a);
    else if (ks <= 4)
        hipLaunchKernelGGL((join_mfma_kernel<METRIC, 4, false>), g, b, 0, s, a);
    else
        hipLaunchKernelGGL((join_mfma_kernel<METRIC, 8, false>), g, b, 0, s, a);
}

// The screen forms: 16 staged FP32 rows of Dp floats per wave.  BF16: KS
// 32-column k-steps; FP32 operands: KS float4 chunks of 16 columns per lane.
template <int METRIC>
static void launch_join_screen(NndArgs const &a, hipStream_t s) {
    dim3 g((a.n1 - a.n0 + 3) / 4), b(256);
    size_t const lds = size_t(4) * 16 * a.Dp * 4;
    if (a.f32) {
        unsigned const kq = (a.D + 15) / 16;
        if (kq <= 2)
            hipLaunchKernelGGL((join_screen_kernel<METRIC, 2, true>), g, b, lds, s, a);
        else if (kq <= 4)
            hipLaunchKernelGGL((join_screen_kernel<METRIC, 4, true>), g, b, lds, s, a);
        else if (kq <= 6)
            hipLaunchKernelGGL((join_screen_kernel<METRIC, 6, true>), g, b, lds, s, a);
        else if (kq <= 8)
            hipLaunchKernelGGL((join_screen_kernel<METRIC, 8, true>), g, b, lds, s, a);
        else
            hipLaunchKernelGGL((join_screen_kernel<METRIC, 16, true>), g, b, lds, s, a);
        return;
    }
    unsigned const ks = (a.D + 31) / 32;
    if (ks <= 1)
        hipLaunchKernelGGL((join_screen_kernel<METRIC, 1, false>), g, b, lds, s, a);
    else if (ks <= 2)
        hipLaunchKernelGGL((join_screen_kernel<METRIC, 2, false>), g, b, lds, s, a);
    else if (ks <= 3)
        hipLaunchKernelGGL((join_screen_kernel<METRIC, 3, false>), g, b, lds, s, a);
    else if (ks <= 4)
        hipLaunchKernelGGL((join_screen_kernel<METRIC, 4, false>), g, b, lds, s, a);
    else
        hipLaunchKernelGGL((join_screen_kernel<METRIC, 8, false>), g, b, lds, s, a);
}

template <int METRIC>
static void launch_join_mfma(NndArgs const &a, hipStream_t s) {
    if (a.exact && !a.f32)
        launch_join_exact<METRIC>(a, s);
    else
        launch_join_screen<METRIC>(a, s);
}

hipError_t nnd_launch_n8(NndArgs const &a, int half, hipStream_t s) {
    if (a.N == 0) return hipSuccess;
    hipLaunchKernelGGL(n8_kernel, dim3((a.N + 255) / 256), dim3(256), 0, s, a.X8, a.N, a.D, a.Dp, half, a.meta);
    return hipGetLastError();
}

hipError_t nnd_launch_sqn(NndArgs const &a, float2 *out, uint32_t *flags, hipStream_t s) {
    hipLaunchKernelGGL(sqn_kernel, dim3((a.N + 63) / 64), dim3(256), 0, s, a.X, a.N, a.D, a.Dp, out, flags);
    return hipGetLastError();
}

hipError_t nnd_launch_join(int metric, NndArgs const &a, hipStream_t s) {
    if (a.n1 > a.N || a.n0 >= a.n1) return a.n0 == a.n1 ? hipSuccess : hipErrorInvalidValue;
    if (a.mfma) {
        if (a.D > 256) return hipErrorInvalidValue;
        if (metric == METRIC_L2SQR)
            launch_join_mfma<METRIC_L2SQR>(a, s);
        else if (metric == METRIC_EUCLIDEAN)
            launch_join_mfma<METRIC_EUCLIDEAN>(a, s);
        else
            launch_join_mfma<METRIC_ANGULAR>(a, s);
        return hipGetLastError();
    }
    if (metric == METRIC_L2SQR)
        launch_join_m<METRIC_L2SQR>(a, s);
    else if (metric == METRIC_EUCLIDEAN)
        launch_join_m<METRIC_EUCLIDEAN>(a, s);
    else
        launch_join_m<METRIC_ANGULAR>(a, s);
    return hipGetLastError();
}

hipError_t nnd_launch_merge(NndArgs const &a, hipStream_t s) {
    // waves per block: 4, fewer when their LDS would pass 64 KB (large L)
    unsigned wpb = 4;
    while (wpb > 1 && wpb * merge_lds_per_wave(a.L) > 65536) wpb /= 2;
    hipLaunchKernelGGL(merge_kernel, dim3((a.N + wpb - 1) / wpb), dim3(64 * wpb), wpb * merge_lds_per_wave(a.L), s,
                       a);
    unsigned const nb = std::min(1024u, (a.N + 255) / 256);
    hipLaunchKernelGGL(stats_reduce_kernel, dim3(nb), dim3(256), 0, s, a.part, a.N, a.stats);
    return hipGetLastError();
}

hipError_t nnd_launch_iter_begin(NndArgs const &a, int advance, hipStream_t s) {
    hipLaunchKernelGGL(iter_begin_kernel, dim3(1), dim3(64), 0, s, a.stats, a.iter, advance);
    return hipGetLastError();
}

namespace {
// Squared distance of n id pairs in the projection Z (zdim floats per row).
__global__ void __launch_bounds__(256) zpair_kernel(float const *Z, unsigned P, uint32_t const *pairs, unsigned n,
                                                    float *out) {
    unsigned const i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float const *za = Z + size_t(pairs[2 * i]) * P, *zb = Z + size_t(pairs[2 * i + 1]) * P;
    float d = 0.f;
    for (unsigned k = 0; k < P; ++k) {
        float const t = za[k] - zb[k];
        d = __builtin_fmaf(t, t, d);
    }
    out[i] = d;
}

}  // namespace

hipError_t nnd_launch_zpair(float const *Z, unsigned P, uint32_t const *pairs, unsigned n, float *out, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(zpair_kernel, dim3((n + 255) / 256), dim3(256), 0, s, Z, P, pairs, n, out);
    return hipGetLastError();
}

hipError_t nnd_launch_prerank(NndArgs const &a, hipStream_t s) {
    hipLaunchKernelGGL(prerank_kernel, dim3((a.N + 3) / 4), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t nnd_launch_update(NndArgs const &a, hipStream_t s) {
    hipLaunchKernelGGL(update_m_kernel, dim3((a.N + 3) / 4), dim3(256), 0, s, a);
    hipLaunchKernelGGL(sample_kernel, dim3((a.N + 3) / 4), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t nnd_launch_extract(NndArgs const &a, uint32_t const *off, uint32_t *ids, float *dists, uint32_t *M,
                              hipStream_t s) {
    hipLaunchKernelGGL(extract_kernel, dim3((a.N + 3) / 4), dim3(256), 0, s, a, off, ids, dists, M);
    return hipGetLastError();
}

hipError_t nnd_launch_gather_controls(NndArgs const &a, uint32_t const *ctl, unsigned nc, float *dout, uint32_t *cout,
                                      hipStream_t s) {
    if (nc == 0) return hipSuccess;
    hipLaunchKernelGGL(gather_controls_kernel, dim3(nc), dim3(256), 0, s, a, ctl, nc, dout, cout);
    return hipGetLastError();
}

hipError_t nnd_launch_screen_pairs(int metric, NndArgs const &a, uint32_t const *pairs, unsigned n, float *out_d,
                                   float *out_e, hipStream_t s) {
    if (n == 0) return hipSuccess;
    if (a.D > 256) return hipErrorInvalidValue;
    unsigned const waves = (n + 15) / 16, ks = (a.D + 31) / 32;
    dim3 g((waves + 3) / 4), b(256);
#define KG_SCREEN(M)                                                                                       \
    do {                                                                                                   \
        if (ks <= 1)                                                                                       \
            hipLaunchKernelGGL((screen_pairs_kernel<M, 1>), g, b, 0, s, a, pairs, n, out_d, out_e);       \
        else if (ks <= 2)                                                                                  \
            hipLaunchKernelGGL((screen_pairs_kernel<M, 2>), g, b, 0, s, a, pairs, n, out_d, out_e);       \
        else if (ks <= 4)                                                                                  \
            hipLaunchKernelGGL((screen_pairs_kernel<M, 4>), g, b, 0, s, a, pairs, n, out_d, out_e);       \
        else                                                                                               \
            hipLaunchKernelGGL((screen_pairs_kernel<M, 8>), g, b, 0, s, a, pairs, n, out_d, out_e);       \
    } while (0)
    if (metric == METRIC_L2SQR)
        KG_SCREEN(METRIC_L2SQR);
    else if (metric == METRIC_EUCLIDEAN)
        KG_SCREEN(METRIC_EUCLIDEAN);
    else
        KG_SCREEN(METRIC_ANGULAR);
#undef KG_SCREEN
    return hipGetLastError();
}

hipError_t nnd_launch_meta_init(int metric, NndArgs const &a, hipStream_t s) {
    hipLaunchKernelGGL(meta_init_kernel, dim3((a.N + 255) / 256), dim3(256), 0, s, a.meta, a.N, metric,
                       a.mfma ? a.sqn : nullptr, a.norms);
    return hipGetLastError();
}

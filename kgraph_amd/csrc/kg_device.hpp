// kg_device.hpp -- device-side building blocks shared by the kgraph_amd HIP
// kernels (gfx950, wave64).
//
// Canonical FP32 distance order ("X16"), identical to the order the
// reference's xsimd AVX-512 l2sqr kernels use with FMA contraction
// (kgraph-data.h:41-103; xsimd arch/xsimd_avx512f.hpp:1393-1403 ->
// arch/xsimd_sse3.hpp:47-52):
//   * element i < D16 (= D rounded down to 16) goes to partial i % 16, the 16
//     partials accumulate with fmaf in increasing i;
//   * u[e] = (p[e] + p[e+4]) + (p[e+8] + p[e+12]),  r = (u0 + u1) + (u2 + u3);
//   * the tail [D16, D) is added sequentially with fmaf.
// A group of 4 adjacent lanes evaluates one row: lane t of the group loads the
// float4 chunks k = t, t+4, t+8, ... so its float4 accumulator holds partials
// 4t..4t+3; a quad-xor-1 then quad-xor-2 exchange forms u[] in every lane of
// the group (IEEE addition is commutative, so all four lanes hold the same
// bits), and the final sum is formed in-lane.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define KG_WAVE 64

namespace kgd {

enum { METRIC_L2SQR = 0, METRIC_EUCLIDEAN = 1, METRIC_ANGULAR = 2 };

__device__ __forceinline__ unsigned lane_id() { return __lane_id(); }

// Orders floats as unsigned integers (ascending float order == ascending key).
__device__ __forceinline__ uint32_t ord_f32(float f) {
    uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float unord_f32(uint32_t k) {
    return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k);
}

// Compiler + LDS ordering point for cross-lane communication inside one wave.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }
__device__ __forceinline__ unsigned popc64(uint64_t b) { return unsigned(__builtin_popcountll(b)); }
// index of the lowest set bit (b != 0)
__device__ __forceinline__ unsigned ctz64(uint64_t b) { return unsigned(__builtin_ctzll(b)); }

template <typename T>
__device__ __forceinline__ T uniform(T v) {
    return __builtin_amdgcn_readfirstlane(v);
}

// DPP quad permutations: [1,0,3,2] (xor 1) and [2,3,0,1] (xor 2).
__device__ __forceinline__ float quad_xor1(float v) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0xB1, 0xF, 0xF, false));
}
__device__ __forceinline__ float quad_xor2(float v) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x4E, 0xF, 0xF, false));
}

// Reduce a quad's float4 partial accumulators to the canonical sum.  Every
// lane of the quad returns the same value.
__device__ __forceinline__ float quad_reduce(float4 p) {
    p.x = p.x + quad_xor1(p.x);
    p.y = p.y + quad_xor1(p.y);
    p.z = p.z + quad_xor1(p.z);
    p.w = p.w + quad_xor1(p.w);
    p.x = p.x + quad_xor2(p.x);
    p.y = p.y + quad_xor2(p.y);
    p.z = p.z + quad_xor2(p.z);
    p.w = p.w + quad_xor2(p.w);
    return (p.x + p.y) + (p.z + p.w);
}

// BF16 bits of f, rounded to nearest even (finite values).
__device__ __forceinline__ uint16_t bf16_bits(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return uint16_t(u >> 16);
}

// Chunk c (4 elements) and element i of a data row held as FP32 or as BF16
// bits; a BF16 element widens to FP32 exactly.
__device__ __forceinline__ float4 load_chunk(float const *r, unsigned c) {
    return reinterpret_cast<float4 const *>(r)[c];
}
__device__ __forceinline__ float4 load_chunk(uint16_t const *r, unsigned c) {
    uint2 const v = reinterpret_cast<uint2 const *>(r)[c];
    return make_float4(__uint_as_float(v.x << 16), __uint_as_float(v.x & 0xffff0000u), __uint_as_float(v.y << 16),
                       __uint_as_float(v.y & 0xffff0000u));
}
__device__ __forceinline__ float load_elem(float const *r, unsigned i) { return r[i]; }
__device__ __forceinline__ float load_elem(uint8_t const *r, unsigned i) { return float(r[i]); }
__device__ __forceinline__ float load_elem(uint16_t const *r, unsigned i) { return __uint_as_float(uint32_t(r[i]) << 16); }

// One quad computes the canonical SSD / dot of data row `a` (global) and
// vector `b` (LDS or global).  t = lane within the quad (0..3).  Rows are
// 16-byte aligned; D16 = D & ~15.
#ifndef QM_GROUP
#define QM_GROUP 8
#endif
// The row may be BF16 bits (RowT = uint16_t, 8-byte aligned rows): its
// values widen exactly, so the result is the FP32 row's when every value of
// the row is representable in BF16.
//
// Loads: lane t takes chunks t, t + 4, ... in groups of G; every load of a
// group is issued before its first chunk is accumulated (the last group
// predicated: a chunk count that is no multiple of 4 G costs no extra round
// trips), and the last D - D16 < 16 values load as whole chunks at once.  BG:
// b is in global memory too (a row), so its chunks load with the row's.  The
// accumulation order is the same whatever the grouping.
template <int METRIC, typename RowT = float, bool BG = false, int G = QM_GROUP>
__device__ __forceinline__ float quad_metric(RowT const *__restrict__ a, float const *__restrict__ b, unsigned D,
                                             unsigned t) {
    unsigned const nch = (D & ~15u) >> 2;
    float4 p = make_float4(0.f, 0.f, 0.f, 0.f);
    float4 const *b4 = reinterpret_cast<float4 const *>(b);
    auto step = [&p](float4 x, float4 y) {
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
    };
    float4 const z4 = make_float4(0.f, 0.f, 0.f, 0.f);
    unsigned k = t;
    for (; k + 4 * (G - 1) < nch; k += 4 * G) {  // whole groups
        float4 x[G], y[G];
#pragma unroll
        for (int g = 0; g < G; ++g) {
            x[g] = load_chunk(a, k + 4 * g);
            if (BG) y[g] = b4[k + 4 * g];
        }
#pragma unroll
        for (int g = 0; g < G; ++g) step(x[g], BG ? y[g] : b4[k + 4 * g]);
    }
    if (k < nch) {  // the last, partial group
        float4 x[G], y[G];
#pragma unroll
        for (int g = 0; g < G; ++g) {
            bool const ok = k + 4 * g < nch;
            x[g] = ok ? load_chunk(a, k + 4 * g) : z4;
            if (BG) y[g] = ok ? b4[k + 4 * g] : z4;
        }
#pragma unroll
        for (int g = 0; g < G; ++g)
            if (k + 4 * g < nch) step(x[g], BG ? y[g] : b4[k + 4 * g]);
    }
    float acc = quad_reduce(p);
    unsigned const D16 = D & ~15u;
    if (__builtin_expect(D16 < D, 0)) {
        // whole chunks below D load as chunks, the values of a last partial
        // one one by one (nothing at or past D is read: b may be an unpadded
        // row); all of them before the first is used
        float4 xt[4], yt[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            unsigned const ch = (D16 >> 2) + c;
            if (4 * ch + 4 <= D) {
                xt[c] = load_chunk(a, ch);
                yt[c] = b4[ch];
            } else {
                unsigned const i = 4 * ch;
                xt[c] = make_float4(i < D ? load_elem(a, i) : 0.f, i + 1 < D ? load_elem(a, i + 1) : 0.f,
                                    i + 2 < D ? load_elem(a, i + 2) : 0.f, 0.f);
                yt[c] = make_float4(i < D ? b[i] : 0.f, i + 1 < D ? b[i + 1] : 0.f, i + 2 < D ? b[i + 2] : 0.f, 0.f);
            }
        }
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            float const xv[4] = {xt[c].x, xt[c].y, xt[c].z, xt[c].w};
            float const yv[4] = {yt[c].x, yt[c].y, yt[c].z, yt[c].w};
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                if (D16 + 4 * c + e < D) {
                    if (METRIC == METRIC_L2SQR) {
                        float const d = xv[e] - yv[e];
                        acc = __builtin_fmaf(d, d, acc);
                    } else {
                        acc = __builtin_fmaf(xv[e], yv[e], acc);
                    }
                }
            }
        }
    }
    return acc;
}

__device__ __forceinline__ float4 widen_bf16x4(uint2 v) {
    return make_float4(__uint_as_float(v.x << 16), __uint_as_float(v.x & 0xffff0000u), __uint_as_float(v.y << 16),
                       __uint_as_float(v.y & 0xffff0000u));
}

// Two BF16 rows per quad at once: their 16 raw 8-byte loads per lane take the
// registers of one FP32 row's 8 loads, and all are issued before either row
// is accumulated; each result equals quad_metric's for that row.
template <int METRIC>
__device__ __forceinline__ float2 quad_metric2_bf16(uint16_t const *__restrict__ a0, uint16_t const *__restrict__ a1,
                                                    float const *__restrict__ b, unsigned D, unsigned t) {
    unsigned const nch = (D & ~15u) >> 2;
    float4 p0 = make_float4(0.f, 0.f, 0.f, 0.f), p1 = p0;
    float4 const *b4 = reinterpret_cast<float4 const *>(b);
    uint2 const *r0 = reinterpret_cast<uint2 const *>(a0);
    uint2 const *r1 = reinterpret_cast<uint2 const *>(a1);
    auto step = [](float4 &p, float4 x, float4 y) {
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
    };
    uint2 const z2 = make_uint2(0u, 0u);
    unsigned k = t;
    for (; k + 4 * (QM_GROUP - 1) < nch; k += 4 * QM_GROUP) {  // whole groups
        uint2 x0[QM_GROUP], x1[QM_GROUP];
#pragma unroll
        for (int g = 0; g < QM_GROUP; ++g) x0[g] = r0[k + 4 * g];
#pragma unroll
        for (int g = 0; g < QM_GROUP; ++g) x1[g] = r1[k + 4 * g];
#pragma unroll
        for (int g = 0; g < QM_GROUP; ++g) {
            float4 const y = b4[k + 4 * g];
            step(p0, widen_bf16x4(x0[g]), y);
            step(p1, widen_bf16x4(x1[g]), y);
        }
    }
    if (__builtin_expect(k < nch, 0)) {  // the last, partial group
        uint2 x0[QM_GROUP], x1[QM_GROUP];
#pragma unroll
        for (int g = 0; g < QM_GROUP; ++g) x0[g] = k + 4 * g < nch ? r0[k + 4 * g] : z2;
#pragma unroll
        for (int g = 0; g < QM_GROUP; ++g) x1[g] = k + 4 * g < nch ? r1[k + 4 * g] : z2;
#pragma unroll
        for (int g = 0; g < QM_GROUP; ++g) {
            if (k + 4 * g < nch) {
                float4 const y = b4[k + 4 * g];
                step(p0, widen_bf16x4(x0[g]), y);
                step(p1, widen_bf16x4(x1[g]), y);
            }
        }
    }
    float acc0 = quad_reduce(p0), acc1 = quad_reduce(p1);
    unsigned const D16 = D & ~15u;
    if (__builtin_expect(D16 < D, 0)) {  // the last D - D16 < 16 values, their chunks loaded at once
        uint2 t0[4], t1[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            unsigned const ch = (D16 >> 2) + c;
            t0[c] = 4 * ch < D ? r0[ch] : z2;
            t1[c] = 4 * ch < D ? r1[ch] : z2;
        }
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            float4 const w0 = widen_bf16x4(t0[c]), w1 = widen_bf16x4(t1[c]);
            float const v0[4] = {w0.x, w0.y, w0.z, w0.w}, v1[4] = {w1.x, w1.y, w1.z, w1.w};
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                unsigned const i = D16 + 4 * c + e;
                if (i < D) {
                    if (METRIC == METRIC_L2SQR) {
                        float const d0 = v0[e] - b[i], d1 = v1[e] - b[i];
                        acc0 = __builtin_fmaf(d0, d0, acc0);
                        acc1 = __builtin_fmaf(d1, d1, acc1);
                    } else {
                        acc0 = __builtin_fmaf(v0[e], b[i], acc0);
                        acc1 = __builtin_fmaf(v1[e], b[i], acc1);
                    }
                }
            }
        }
    }
    return make_float2(acc0, acc1);
}

// Two rows of the 8-bit copy per quad against a query held as 8-bit words
// (q8, Dp / 4 words), in integer arithmetic: (a0 . q, a1 . q) and (|a0|^2,
// |a1|^2), summed over the quad (every lane gets the totals).  v_dot4_u32_u8
// takes four 8-bit products per instruction; the sums are exact, so when every
// row and query value is an integer in [0, 255] and D 255^2 < 2^24 they equal
// the canonical FP32 dot products and norms (whose partial sums are then
// exact too) -- and |a - q|^2 = |a|^2 + |q|^2 - 2 a.q exactly.
__device__ __forceinline__ int quad_isum(int v) {
    v += __builtin_amdgcn_mov_dpp(v, 0xB1, 0xF, 0xF, false);
    v += __builtin_amdgcn_mov_dpp(v, 0x4E, 0xF, 0xF, false);
    return v;
}
__device__ __forceinline__ void quad_dot2_u8(uint8_t const *__restrict__ a0, uint8_t const *__restrict__ a1,
                                             uint32_t const *__restrict__ q8, unsigned nw, unsigned t, uint32_t &d0,
                                             uint32_t &d1, uint32_t &n0, uint32_t &n1) {
    uint32_t const *r0 = reinterpret_cast<uint32_t const *>(a0);
    uint32_t const *r1 = reinterpret_cast<uint32_t const *>(a1);
    uint32_t s0 = 0, s1 = 0, m0 = 0, m1 = 0;
    // groups of QM_GROUP words per lane, every load of a group issued first
    // (integer sums: the grouping cannot change them)
    unsigned k = t;
    for (; k + 4 * (QM_GROUP - 1) < nw; k += 4 * QM_GROUP) {  // whole groups
        uint32_t x0[QM_GROUP], x1[QM_GROUP];
#pragma unroll
        for (int g = 0; g < QM_GROUP; ++g) x0[g] = r0[k + 4 * g];
#pragma unroll
        for (int g = 0; g < QM_GROUP; ++g) x1[g] = r1[k + 4 * g];
#pragma unroll
        for (int g = 0; g < QM_GROUP; ++g) {
            uint32_t const y = q8[k + 4 * g];
            s0 = __builtin_amdgcn_udot4(x0[g], y, s0, false);
            s1 = __builtin_amdgcn_udot4(x1[g], y, s1, false);
            m0 = __builtin_amdgcn_udot4(x0[g], x0[g], m0, false);
            m1 = __builtin_amdgcn_udot4(x1[g], x1[g], m1, false);
        }
    }
    if (__builtin_expect(k < nw, 0)) {  // the last, partial group
        uint32_t x0[QM_GROUP], x1[QM_GROUP];
#pragma unroll
        for (int g = 0; g < QM_GROUP; ++g) x0[g] = k + 4 * g < nw ? r0[k + 4 * g] : 0u;
#pragma unroll
        for (int g = 0; g < QM_GROUP; ++g) x1[g] = k + 4 * g < nw ? r1[k + 4 * g] : 0u;
#pragma unroll
        for (int g = 0; g < QM_GROUP; ++g) {
            uint32_t const y = k + 4 * g < nw ? q8[k + 4 * g] : 0u;
            s0 = __builtin_amdgcn_udot4(x0[g], y, s0, false);
            s1 = __builtin_amdgcn_udot4(x1[g], y, s1, false);
            m0 = __builtin_amdgcn_udot4(x0[g], x0[g], m0, false);
            m1 = __builtin_amdgcn_udot4(x1[g], x1[g], m1, false);
        }
    }
    d0 = uint32_t(quad_isum(int(s0)));
    d1 = uint32_t(quad_isum(int(s1)));
    n0 = uint32_t(quad_isum(int(m0)));
    n1 = uint32_t(quad_isum(int(m1)));
}

__device__ __forceinline__ float4 widen_u8x4(uint32_t v) {
    return make_float4(float(v & 0xffu), float((v >> 8) & 0xffu), float((v >> 16) & 0xffu), float(v >> 24));
}

// Two rows of the lossless 8-bit copy per quad: one 4-byte word per 4-value
// chunk, widened at use; each result equals quad_metric's on the FP32 rows
// (the same values in the same order).
template <int METRIC>
__device__ __forceinline__ float2 quad_metric2_u8(uint8_t const *__restrict__ a0, uint8_t const *__restrict__ a1,
                                                  float const *__restrict__ b, unsigned D, unsigned t) {
    unsigned const nch = (D & ~15u) >> 2;
    float4 p0 = make_float4(0.f, 0.f, 0.f, 0.f), p1 = p0;
    float4 const *b4 = reinterpret_cast<float4 const *>(b);
    uint32_t const *r0 = reinterpret_cast<uint32_t const *>(a0);
    uint32_t const *r1 = reinterpret_cast<uint32_t const *>(a1);
    auto step = [](float4 &p, float4 x, float4 y) {
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
    };
    unsigned k = t;
    for (; k + 4 * (QM_GROUP - 1) < nch; k += 4 * QM_GROUP) {  // whole groups
        uint32_t x0[QM_GROUP], x1[QM_GROUP];
#pragma unroll
        for (int g = 0; g < QM_GROUP; ++g) x0[g] = r0[k + 4 * g];
#pragma unroll
        for (int g = 0; g < QM_GROUP; ++g) x1[g] = r1[k + 4 * g];
#pragma unroll
        for (int g = 0; g < QM_GROUP; ++g) {
            float4 const y = b4[k + 4 * g];
            step(p0, widen_u8x4(x0[g]), y);
            step(p1, widen_u8x4(x1[g]), y);
        }
    }
    if (__builtin_expect(k < nch, 0)) {  // the last, partial group
        uint32_t x0[QM_GROUP], x1[QM_GROUP];
#pragma unroll
        for (int g = 0; g < QM_GROUP; ++g) x0[g] = k + 4 * g < nch ? r0[k + 4 * g] : 0u;
#pragma unroll
        for (int g = 0; g < QM_GROUP; ++g) x1[g] = k + 4 * g < nch ? r1[k + 4 * g] : 0u;
#pragma unroll
        for (int g = 0; g < QM_GROUP; ++g) {
            if (k + 4 * g < nch) {
                float4 const y = b4[k + 4 * g];
                step(p0, widen_u8x4(x0[g]), y);
                step(p1, widen_u8x4(x1[g]), y);
            }
        }
    }
    float acc0 = quad_reduce(p0), acc1 = quad_reduce(p1);
    unsigned const D16 = D & ~15u;
    if (__builtin_expect(D16 < D, 0)) {  // the last D - D16 < 16 values, their words loaded at once
        uint32_t t0[4], t1[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            unsigned const ch = (D16 >> 2) + c;
            t0[c] = 4 * ch < D ? r0[ch] : 0u;
            t1[c] = 4 * ch < D ? r1[ch] : 0u;
        }
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            float4 const w0 = widen_u8x4(t0[c]), w1 = widen_u8x4(t1[c]);
            float const v0[4] = {w0.x, w0.y, w0.z, w0.w}, v1[4] = {w1.x, w1.y, w1.z, w1.w};
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                unsigned const i = D16 + 4 * c + e;
                if (i < D) {
                    if (METRIC == METRIC_L2SQR) {
                        float const d0 = v0[e] - b[i], d1 = v1[e] - b[i];
                        acc0 = __builtin_fmaf(d0, d0, acc0);
                        acc1 = __builtin_fmaf(d1, d1, acc1);
                    } else {
                        acc0 = __builtin_fmaf(v0[e], b[i], acc0);
                        acc1 = __builtin_fmaf(v1[e], b[i], acc1);
                    }
                }
            }
        }
    }
    return make_float2(acc0, acc1);
}

// Final metric value from the quad result (python-api.cpp:56-58, 66-68;
// kgraph-data.h:41-61).  na = data-row norm, nb = other norm.
template <int METRIC>
__device__ __forceinline__ float finish_metric(float acc, float na, float nb) {
    if (METRIC == METRIC_L2SQR) return acc;
    if (METRIC == METRIC_EUCLIDEAN) return (na + nb) - acc;
    return -acc / __builtin_fmaf(na, nb, 1.1920928955078125e-07f);
}

// Row norm as stored per metric (python-api.cpp:51-69, 110-114).
template <int METRIC>
__device__ __forceinline__ float finish_norm(float n2) {
    if (METRIC == METRIC_EUCLIDEAN) return n2 / 2;
    // sqrtf is correctly rounded here (hipcc expands it with FMA correction
    // steps); __fsqrt_rn compiles to the bare v_sqrt_f32, which is not
    if (METRIC == METRIC_ANGULAR) return sqrtf(n2);
    return 0.f;
}

}  // namespace kgd

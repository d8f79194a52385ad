// kg_search.hip -- persistent graph-walk k-NN search, one wavefront per query.
//
// Restates KGraphImpl::search (kgraph.cpp:423-564) with identical semantics:
// the K+P candidate pool, batches of S neighbours, the sequential
// insert-after-equals rule of UpdateKnnList (kgraph.cpp:176-202), the rewind
// `k = r` when a better candidate lands (kgraph.cpp:507-512), T trials merged
// into a K-entry result list (kgraph.cpp:516-529) and the epsilon cut
// (kgraph.cpp:533-540).  Seeds come from the host (mt19937 + GenRandom, shared
// by every query exactly as kgraph.cpp:444-453) or from the caller's ids
// (init > 0, kgraph.cpp:461-467); ties in the seed sort are ordered by id.
//
// Per-wave LDS: pool ids (bit 31 = "still has unexpanded neighbours"), pool
// dists, pool meta (m | M << 16), pool CSR base offsets, a u64 key area for
// the seed sort, the T>1 result list, the query vector and a 64-entry batch
// buffer.  The visited set is an exact per-wave bitset in HBM, cleared on exit
// through the list of ids the query touched (wholesale if that list
// overflows).  In front of it sits an exact hash set of visited ids in LDS
// (open addressing, linear probing, compare-and-swap inserts): a query visits
// ~1,000 ids, and while the set holds them (at most half full) no visited
// test leaves the CU.  Past that the set spills into the HBM bitset and the
// query continues there.
//
// Inside a batch the wave reads the S neighbour ids (one per lane), drops
// in-batch duplicates, test-and-sets the visited bits with atomicOr, gathers
// the surviving rows with 4 lanes per row (16 rows per pass), then applies the
// inserts one by one in neighbour order.  That equals the reference's
// per-neighbour loop because neither the visited state nor the distances
// depend on the inserts.
#include <cmath>

#include "kg_device.hpp"
#include "kg_internal.hpp"

namespace {

using namespace kgd;

constexpr unsigned SEARCH_MAX_WPB = 4;  // waves per block the kernel is compiled for
constexpr uint32_t FLAG = 0x80000000u;
constexpr uint32_t IDMASK = 0x7fffffffu;

// The G lanes of a wave that walk one query (G = 64: the whole wave; G = 32:
// two queries per wave, one per half).  gl = lane within the group, base = the
// group's first lane; ballots and shuffles stay inside the group.
template <unsigned G>
struct Grp {
    unsigned gl, base;
    __device__ __forceinline__ Grp() {
        unsigned const l = lane_id();
        gl = l % G;
        base = l - gl;
    }
    __device__ __forceinline__ uint64_t ballot(bool p) const {
        uint64_t const b = kgd::ballot(p);
        if constexpr (G == KG_WAVE) return b;
        else return (b >> base) & ((1ull << G) - 1);
    }
    __device__ __forceinline__ uint64_t lt() const { return (1ull << gl) - 1; }
    template <typename T>
    __device__ __forceinline__ T shfl(T v, unsigned j) const {
        return __shfl(v, int(base + j));
    }
};
// the whole wave: the lane itself, base 0
template <>
struct Grp<KG_WAVE> {
    unsigned gl;
    static constexpr unsigned base = 0;
    __device__ __forceinline__ Grp() : gl(lane_id()) {}
    __device__ __forceinline__ uint64_t ballot(bool p) const { return kgd::ballot(p); }
    __device__ __forceinline__ uint64_t lt() const { return (1ull << gl) - 1; }
    template <typename T>
    __device__ __forceinline__ T shfl(T v, unsigned j) const {
        return __shfl(v, int(j));
    }
};

struct WaveLDS {
    uint32_t *pid;
    float *pdist;
    uint32_t *pmeta;
    uint32_t *pbase;
    uint64_t *tkey;
    uint32_t *rid;
    float *rdist;
    float *q;
    uint32_t *q8;  // the query as 8-bit words (integer path over the 8-bit rows)
    uint4 *qh, *ql;  // BF16 screen: the query's BF16 high and low halves, 8 values a word
    uint8_t *sidx;   // BF16 screen: batch positions of the candidates to re-score
    uint32_t *fid;
    float *fdist;
    uint32_t *vt;  // visited hash set (vcap entries of 32 or 16 bits, all ones = free)
};

__host__ __device__ inline size_t align16(size_t x) { return (x + 15) & ~size_t(15); }

struct LdsLayout {
    size_t pid, pdist, pmeta, pbase, tkey, rid, rdist, q, q8, qh, ql, sidx, fid, fdist, vt, total;
    // tcap: starting points sorted at once (P random seeds, or init ids);
    // vcap: entries of the visited hash set (0: none)
    // vbytes: bytes per visited-set entry (4, or 2 for the compact form)
    // base: pool entries keep their CSR base offset (walks over the CSR
    // lists; the padded rows need none)
    // screen: the BF16 screen's query halves and re-score list
    // lean: the lean walk keeps no FP32 copy of the query
    __host__ __device__ LdsLayout(unsigned cap, unsigned rcap, unsigned Dp, unsigned tcap, unsigned vcap,
                                  unsigned vbytes, bool base, bool screen, bool lean = false) {
        size_t const d32 = screen ? (Dp + 31) / 32 * 32 : 0;  // BF16 values per query half
        size_t o = 0;
        pid = o; o = align16(o + 4ull * cap);
        pdist = o; o = align16(o + 4ull * cap);
        pmeta = o; o = align16(o + 4ull * cap);
        pbase = o; o = align16(o + (base ? 4ull * cap : 0ull));
        // the lean walk sorts its starting points in the batch buffers (fid,
        // fdist), which no expansion uses yet, when they hold them
        bool const tk_alias = lean && 8ull * tcap <= 8ull * KG_WAVE;
        tkey = o; o = align16(o + (tk_alias ? 0ull : 8ull * tcap));
        rid = o; o = align16(o + 4ull * rcap);
        rdist = o; o = align16(o + 4ull * rcap);
        q = o; o = align16(o + (lean ? 0ull : 4ull * Dp));
        q8 = o; o = align16(o + Dp);
        qh = o; o = align16(o + 2 * d32);
        ql = o; o = align16(o + 2 * d32);
        sidx = o; o = align16(o + (screen ? KG_WAVE : 0));
        fid = o; o = align16(o + 4ull * KG_WAVE);
        fdist = o; o = align16(o + 4ull * KG_WAVE);  // (right after fid: 8 KG_WAVE bytes together)
        if (tk_alias) tkey = fid;
        vt = o; o = align16(o + size_t(vbytes) * vcap);
        total = o;
    }
};

// Pool entry whose M is read at its first expansion (padded adjacency).
constexpr uint32_t VT_M_UNKNOWN = 0xffffu;

// actual_M (kgraph.cpp:284-286) truncated to NeighborX::M's uint16_t
// (kgraph.cpp:79, 474, 508); also returns the node's CSR base offset.
// One 16-byte load: gnode[id] = {CSR offset, degree, M, 0}.
__device__ __forceinline__ uint32_t actual_M(kg_search_args const &a, uint32_t id, uint32_t *base) {
    uint4 const g = a.gnode[id];
    uint32_t m = g.z;
    if (m < a.Mparam) m = a.Mparam;
    if (m > g.y) m = g.y;
    *base = g.x;
    return m & 0xffffu;
}

// Distances of ids[0..cnt) (LDS, flag bit ignored) -> out[0..cnt) (LDS).
// Rows are gathered from a.X (RowT = float), from the lossless BF16 copy a.Xb
// (RowT = uint16_t) or from the lossless 8-bit copy a.X8 (RowT = uint8_t);
// all give the same distances.
// The query as 8-bit words is usable (on) and its integer |q|^2 (sq).
struct QInt {
    bool on;
    uint32_t sq;
};

// G lanes per query: NQ = G / 4 quads, one row per quad a pass (FP32) or two
// (8-bit, BF16).
// INTONLY: every query takes the integer path (8-bit rows, integer queries).
template <int METRIC, typename RowT, unsigned G = KG_WAVE, bool INTONLY = false>
__device__ __forceinline__ void batch_dist(kg_search_args const &a, uint32_t const *ids, float *out, unsigned cnt,
                                           float const *q, float qn, uint32_t const *q8, QInt qv) {
    RowT const *X;
    if constexpr (sizeof(RowT) == 1)
        X = a.X8;
    else if constexpr (sizeof(RowT) == 2)
        X = a.Xb;
    else
        X = a.X;
    constexpr unsigned NQ = G / 4;
    unsigned const lane = lane_id() % G;
    unsigned const grp = lane >> 2, t = lane & 3;
    if constexpr (sizeof(RowT) == 1) {
        if (INTONLY || qv.on) {
            // integer path: v_dot4 over 8-bit words, exact (kg_device.hpp
            // quad_dot2_u8), then the metric's FP32 finish on exact values
            for (unsigned b0 = 0; b0 < cnt; b0 += 2 * NQ) {
                unsigned const c0 = b0 + grp, c1 = c0 + NQ;
                if (c0 < cnt) {
                    bool const two = c1 < cnt;
                    uint32_t const id0 = ids[c0] & IDMASK, id1 = two ? (ids[c1] & IDMASK) : id0;
                    uint32_t d0, d1, n0, n1;
                    quad_dot2_u8(X + size_t(id0) * a.Dp, X + size_t(id1) * a.Dp, q8, a.Dp / 4, t, d0, d1, n0, n1);
                    float r0, r1;
                    if (METRIC == METRIC_L2SQR) {
                        r0 = float(n0 + qv.sq - 2u * d0);
                        r1 = float(n1 + qv.sq - 2u * d1);
                    } else {
                        r0 = finish_metric<METRIC>(float(d0), a.norms[id0], qn);
                        r1 = finish_metric<METRIC>(float(d1), a.norms[id1], qn);
                    }
                    if (t == 0) {
                        out[c0] = r0;
                        if (two) out[c1] = r1;
                    }
                }
            }
            wave_sync();
            return;
        }
        if constexpr (INTONLY) return;
        // 8-bit rows: one 128-byte line per 128-value row, two rows per quad
        for (unsigned b0 = 0; b0 < cnt; b0 += 2 * NQ) {
            unsigned const c0 = b0 + grp, c1 = c0 + NQ;
            if (c0 < cnt) {  // whole quads are active together
                bool const two = c1 < cnt;
                uint32_t const id0 = ids[c0] & IDMASK, id1 = two ? (ids[c1] & IDMASK) : id0;
                float2 const acc = quad_metric2_u8<METRIC>(X + size_t(id0) * a.Dp, X + size_t(id1) * a.Dp, q, a.D, t);
                float const n0 = (METRIC == METRIC_L2SQR) ? 0.f : a.norms[id0];
                float const n1 = (METRIC == METRIC_L2SQR) ? 0.f : a.norms[id1];
                if (t == 0) {
                    out[c0] = finish_metric<METRIC>(acc.x, n0, qn);
                    if (two) out[c1] = finish_metric<METRIC>(acc.y, n1, qn);
                }
            }
        }
    } else if constexpr (sizeof(RowT) == 2) {
        // BF16 rows are half the bytes: each quad takes two rows a pass, 32 per
        // pass, in the registers 16 FP32 rows need
        for (unsigned b0 = 0; b0 < cnt; b0 += 2 * NQ) {
            unsigned const c0 = b0 + grp, c1 = c0 + NQ;
            if (c0 < cnt) {  // whole quads are active together
                bool const two = c1 < cnt;
                uint32_t const id0 = ids[c0] & IDMASK, id1 = two ? (ids[c1] & IDMASK) : id0;
                float2 const acc = quad_metric2_bf16<METRIC>(X + size_t(id0) * a.Dp, X + size_t(id1) * a.Dp, q, a.D, t);
                float const n0 = (METRIC == METRIC_L2SQR) ? 0.f : a.norms[id0];
                float const n1 = (METRIC == METRIC_L2SQR) ? 0.f : a.norms[id1];
                if (t == 0) {
                    out[c0] = finish_metric<METRIC>(acc.x, n0, qn);
                    if (two) out[c1] = finish_metric<METRIC>(acc.y, n1, qn);
                }
            }
        }
    } else {
        for (unsigned b0 = 0; b0 < cnt; b0 += NQ) {
            unsigned c = b0 + grp;
            if (b0 + grp < cnt) {  // whole quads are active together
                uint32_t id = ids[c] & IDMASK;
                RowT const *row = X + size_t(id) * a.Dp;
                float acc = quad_metric<METRIC>(row, q, a.D, t);
                float na = (METRIC == METRIC_L2SQR) ? 0.f : a.norms[id];
                float d = finish_metric<METRIC>(acc, na, qn);
                if (t == 0) out[c] = d;
            }
        }
    }
    wave_sync();
}

// ---- BF16 screen (float rows, lossy BF16 copy a.Xs)
//
// A full pool admits a candidate only if its distance is below the pool's
// last one (batch_insert drops every candidate at or past that distance), so
// a candidate whose screened distance minus the screen's error bound is at
// least that distance cannot enter it: it gets +inf, which batch_insert drops
// in the same way, and the others are re-scored in exact FP32.  The pool, the
// expansion order and the results are the unscreened walk's.
//
// The screen scores 16 candidates per v_mfma_f32_16x16x32_bf16 pass of 32
// dimensions: A = the candidates' BF16 rows (lane l: row l & 15, values
// 8 (l >> 4) .. + 8 of the pass), B = the query in every column.  The query
// goes in as two BF16 halves (q = qh + ql + r, |r| <= 2^-16 |q|) in two
// passes, so only the row's rounding (2^-8 relative) is left; products of
// BF16 values are exact in FP32 and the MFMA accumulates in FP32.  For the
// SSD metric a third pass with the rows as both operands gives |x~|^2 on the
// diagonal.  C[m][n] sits in lane n + 16 (m >> 2), register m & 3, so the
// lane with n = m (there are 16 such lanes, in candidate order) owns
// candidate m's dot and norm.
typedef __bf16 ws_bf16x8 __attribute__((ext_vector_type(8)));
typedef float ws_f32x4 __attribute__((ext_vector_type(4)));

// 8 BF16 values of row r from column k (zeros at and past Dp; Dp is a
// multiple of 4, rows are 16-byte aligned when Dp is a multiple of 8)
__device__ __forceinline__ uint4 ws_load(uint16_t const *r, unsigned k, unsigned Dp, bool wide) {
    uint4 v = make_uint4(0u, 0u, 0u, 0u);
    if (k + 8 <= Dp) {
        if (wide) {
            v = *reinterpret_cast<uint4 const *>(r + k);
        } else {
            uint2 const lo = *reinterpret_cast<uint2 const *>(r + k);
            uint2 const hi = *reinterpret_cast<uint2 const *>(r + k + 4);
            v = make_uint4(lo.x, lo.y, hi.x, hi.y);
        }
    } else if (k + 4 <= Dp) {
        uint2 const lo = *reinterpret_cast<uint2 const *>(r + k);
        v = make_uint4(lo.x, lo.y, 0u, 0u);
    }
    return v;
}

__device__ __forceinline__ float ws_pick(ws_f32x4 v, unsigned r) {
    return r == 0 ? v[0] : r == 1 ? v[1] : r == 2 ? v[2] : v[3];
}

// Screened distance of candidate rows against the query, and its error
// bound E: |screened - exact FP32| <= E.  G = screened dot, nxs = screened
// |x~|^2 (SSD only), na = the row's stored norm (|x| angular, |x|^2 / 2
// Euclidean), qn = the query's, qrn = |q|, qsq = |q|^2 (SSD).
template <int METRIC>
__device__ __forceinline__ float2 ws_finish(kg_search_args const &a, float G, float nxs, float na, float qn, float qrn,
                                            float qsq) {
    float ds, E;
    if (METRIC == METRIC_ANGULAR) {
        ds = -G / __builtin_fmaf(na, qn, 1.1920928955078125e-07f);
        E = (a.scr_u + a.scr_c) * 1.01f + a.scr_g * fabsf(ds);
    } else if (METRIC == METRIC_EUCLIDEAN) {
        ds = (na + qn) - G;
        float const xr = sqrtf(2.f * na);
        E = (a.scr_u + a.scr_c) * 1.01f * xr * qrn + a.scr_g * (na + qn + fabsf(G));
    } else {
        ds = (nxs + qsq) - 2.f * G;
        // |x| <= |x~| (1 + 2^-7); the screen computes |x~ - q|^2 within E0,
        // and |x - q|^2 - |x~ - q|^2 = 2 e.(x~ - q) + |e|^2 with |e| <= 2^-8 |x|
        float const xr = sqrtf(fmaxf(nxs, 0.f)) * 1.0078125f;
        float const E0 = 2.f * a.scr_c * xr * qrn + a.scr_g * (fmaxf(nxs, 0.f) + qsq + 2.f * fabsf(G));
        float const dd = sqrtf(fmaxf(ds, 0.f) + E0);
        float const E1 = 2.f * a.scr_u * xr * dd + a.scr_u * a.scr_u * xr * xr + E0;
        E = E1 * 1.01f + a.scr_g * (fmaxf(ds, 0.f) + E1);
    }
    return make_float2(ds, E);
}

// MFMA screen of 16 candidates ids[b0 .. b0 + 16) (ok: lane's row exists):
// returns, in the owner lanes, the screened dot (x) and |x~|^2 (y).
template <int METRIC>
__device__ __forceinline__ float2 ws_mfma(kg_search_args const &a, uint32_t id, bool ok, uint4 const *qh,
                                          uint4 const *ql) {
    unsigned const lane = lane_id();
    unsigned const k8 = (lane >> 4) * 8;
    unsigned const ks = (a.Dp + 31) / 32;
    bool const wide = (a.Dp & 7u) == 0;
    uint16_t const *row = a.Xs + size_t(id) * a.Dp;
    ws_f32x4 g = {0.f, 0.f, 0.f, 0.f}, nx = {0.f, 0.f, 0.f, 0.f};
    for (unsigned s0 = 0; s0 < ks; s0 += 4) {
        uint4 v[4];
#pragma unroll
        for (unsigned j = 0; j < 4; ++j)
            v[j] = (ok && s0 + j < ks) ? ws_load(row, 32 * (s0 + j) + k8, a.Dp, wide) : make_uint4(0u, 0u, 0u, 0u);
#pragma unroll
        for (unsigned j = 0; j < 4; ++j) {
            if (s0 + j < ks) {
                ws_bf16x8 const x = __builtin_bit_cast(ws_bf16x8, v[j]);
                unsigned const qi = 4 * (s0 + j) + (lane >> 4);
                g = __builtin_amdgcn_mfma_f32_16x16x32_bf16(x, __builtin_bit_cast(ws_bf16x8, qh[qi]), g, 0, 0, 0);
                g = __builtin_amdgcn_mfma_f32_16x16x32_bf16(x, __builtin_bit_cast(ws_bf16x8, ql[qi]), g, 0, 0, 0);
                if (METRIC == METRIC_L2SQR) nx = __builtin_amdgcn_mfma_f32_16x16x32_bf16(x, x, nx, 0, 0, 0);
            }
        }
    }
    unsigned const r = lane & 3;
    return make_float2(ws_pick(g, r), METRIC == METRIC_L2SQR ? ws_pick(nx, r) : 0.f);
}

// Distances of ids[0..cnt) into out[0..cnt) for a full pool whose last
// distance is thr: +inf for every candidate the screen rules out, the exact
// FP32 distance for the rest (nres counts them).  spec: candidate `lane`'s
// visited test is still in flight (vold = the bitset word before its
// atomicOr, vbit its bit, vcand whether it was tested); only the ones it
// finds fresh are re-scored, read after the screen's rows (the others' out
// is left as it is).
template <int METRIC>
__device__ void batch_screen(kg_search_args const &a, WaveLDS const &w, uint32_t const *ids, float *out, unsigned cnt,
                             float qn, float qrn, float qsq, float thr, unsigned &nres, bool spec = false,
                             bool vcand = false, uint32_t vold = 0u, uint32_t vbit = 0u) {
    unsigned const lane = lane_id();
    uint64_t const lt_mask = (1ull << lane) - 1;
    unsigned const n = lane & 15;
    bool const own = (n >> 2) == (lane >> 4);  // this lane holds candidate n's dot
    unsigned ns = 0;
    for (unsigned b0 = 0; b0 < cnt; b0 += 16) {
        unsigned const c = b0 + n;
        bool const ok = c < cnt;
        uint32_t const id = ok ? (ids[c] & IDMASK) : 0u;
        float2 const gn = ws_mfma<METRIC>(a, id, ok, w.qh, w.ql);
        uint64_t const fm = spec ? ballot(vcand && !(vold & vbit)) : ~0ull;
        bool const live = own && ok && ((fm >> (c & 63u)) & 1u);
        float const na = (METRIC != METRIC_L2SQR && live) ? a.norms[id] : 0.f;
        float2 const de = ws_finish<METRIC>(a, gn.x, gn.y, na, qn, qrn, qsq);
        bool const keep = live && !(de.x - de.y >= thr);  // NaN: re-score
        uint64_t const kb = ballot(keep);
        if (keep) w.sidx[ns + popc64(kb & lt_mask)] = uint8_t(c);
        else if (live) out[c] = __builtin_huge_valf();
        ns += popc64(kb);
    }
    wave_sync();
    // exact FP32 distances of the candidates left, 4 lanes per row
    unsigned const grp = lane >> 2, t = lane & 3;
    for (unsigned b0 = 0; b0 < ns; b0 += 16) {
        if (b0 + grp < ns) {  // whole quads are active together
            unsigned const c = w.sidx[b0 + grp];
            uint32_t const id = ids[c] & IDMASK;
            float const acc = quad_metric<METRIC>(a.X + size_t(id) * a.Dp, w.q, a.D, t);
            float const na = (METRIC == METRIC_L2SQR) ? 0.f : a.norms[id];
            float const d = finish_metric<METRIC>(acc, na, qn);
            if (t == 0) out[c] = d;
        }
    }
    nres += ns;
    wave_sync();
}

// The query's BF16 halves for the screen (qh = bf16(q), ql = bf16(q - qh),
// zero past D), and |q|^2 in FP32.
__device__ __forceinline__ float ws_query(WaveLDS const &w, unsigned D, unsigned Dp) {
    unsigned const lane = lane_id();
    unsigned const words = (Dp + 31) / 32 * 4;  // uint4 words per half
    float sq = 0.f;
    for (unsigned i = lane; i < words; i += KG_WAVE) {
        uint32_t h[4], l[4];
#pragma unroll
        for (unsigned j = 0; j < 4; ++j) {
            uint32_t hh = 0, ll = 0;
#pragma unroll
            for (unsigned e = 0; e < 2; ++e) {
                unsigned const col = 8 * i + 2 * j + e;
                float const v = col < D ? w.q[col] : 0.f;
                uint16_t const hb = bf16_bits(v);
                uint16_t const lb = bf16_bits(v - __uint_as_float(uint32_t(hb) << 16));
                hh |= uint32_t(hb) << (16 * e);
                ll |= uint32_t(lb) << (16 * e);
                sq = __builtin_fmaf(v, v, sq);
            }
            h[j] = hh;
            l[j] = ll;
        }
        w.qh[i] = make_uint4(h[0], h[1], h[2], h[3]);
        w.ql[i] = make_uint4(l[0], l[1], l[2], l[3]);
    }
    for (int o = 32; o > 0; o >>= 1) sq += __shfl_xor(sq, o);
    wave_sync();
    return sq;
}

// Number of entries of the sorted pool[0..L) whose dist <= nd.
template <unsigned G = KG_WAVE>
__device__ __forceinline__ unsigned insert_pos(float const *pdist, unsigned L, float nd) {
    Grp<G> const g;
    unsigned r = 0;
    for (unsigned c = 0; c < L; c += G) {
        unsigned i = c + g.gl;
        bool le = (i < L) && (pdist[i] <= nd);
        unsigned n = popc64(g.ballot(le));
        r += n;
        if (n < G) break;
    }
    return r;
}

// The reference inserts the fresh neighbours of a batch one at a time
// (UpdateKnnList, kgraph.cpp:497-512): each goes after every pool entry with a
// smaller or equal distance, the pool grows to K + P entries and then drops
// its last, and the expansion index rewinds to the smallest insert position.
// Done at once it is a stable merge: candidate `lane` (lane order = neighbour
// order) has rank s = the candidates before it in (dist, order), counted
// across the wave with no sort, and lands at (pool entries with dist <= its
// dist) + s; pool entry i moves to i + (candidates with a smaller dist);
// positions from min(L + nf, qcap - 1) on are dropped.  The smallest insert
// position of the sequential loop is the final position of the best
// candidate (no later insert lands before it, no earlier one after it), so
// the rewind is k = min(k, that position) when it is kept.  cm / cb hold
// candidate `lane`'s M and CSR base; w.fid / w.fdist its id and distance.
// A full pool (L = qcap - 1) keeps only candidates below its last distance:
// any other lands at or past Lf = L whatever its rank and is below no pool
// entry, so only the ones below take part (most candidates of a long walk
// are not, and then nothing moves).
template <unsigned G = KG_WAVE>
__device__ void batch_insert(WaveLDS &w, unsigned nf, uint32_t cm, uint32_t cb, unsigned qcap, unsigned &L,
                             unsigned &k) {
    Grp<G> const g;
    unsigned const lane = g.gl;
    unsigned const L0 = L, Lf = min(L0 + nf, qcap - 1);
    bool const full = L0 > 0 && L0 == qcap - 1;
    float const dl = lane < nf ? w.fdist[lane] : __builtin_huge_valf();
    bool const live = lane < nf && (!full || dl < w.pdist[L0 - 1]);
    uint64_t const lm = g.ballot(live);
    unsigned const na = popc64(lm);
    if (na == 0) {
        L = Lf;
        return;
    }
    // rank of candidate `lane` in (dist, neighbour order) among those taking part
    unsigned rank = 0;
    for (uint64_t m = lm; m; m &= m - 1) {
        unsigned const jj = ctz64(m);
        float dj;
        if constexpr (G == KG_WAVE)
            dj = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(dl), int(jj)));
        else
            dj = g.shfl(dl, jj);
        rank += unsigned(dj < dl || (dj == dl && jj < lane));
    }
    uint32_t const sid = live ? w.fid[lane] : 0u;
    uint32_t const sm = cm, sb = cb;
    float const sd = dl;
    // final position: pool entries with dist <= sd, plus the rank
    unsigned fin = 0;
    if (live) {
        unsigned lo = 0, hi = L0;
        while (lo < hi) {
            unsigned const mid = (lo + hi) >> 1;
            if (w.pdist[mid] <= sd) lo = mid + 1; else hi = mid;
        }
        fin = lo + rank;
    }
    uint64_t const b0 = g.ballot(live && rank == 0);
    unsigned const best_lane = b0 ? ctz64(b0) : 0u;
    unsigned const fin0 = g.shfl(fin, best_lane);
    wave_sync();
    if (live) w.fdist[rank] = sd;  // sorted distances for the pool entries' search
    wave_sync();
    // pool entries move up by the number of candidates strictly below them
    // (those before the best candidate's position stay); top chunk first, each
    // chunk read whole before it is written
    int const lo = int(min(fin0, L0) & ~(G - 1u));
    for (int c = int((L0 - 1) & ~(G - 1u)); L0 > 0 && c >= lo; c -= G) {
        unsigned const i = unsigned(c) + lane;
        bool const mv = i < L0;
        uint32_t id = 0, meta = 0, bs = 0;
        float d = 0.f;
        unsigned np = 0;
        if (mv) {
            id = w.pid[i];
            d = w.pdist[i];
            meta = w.pmeta[i];
            if (w.pbase) bs = w.pbase[i];
            unsigned lo = 0, hi = na;
            while (lo < hi) {
                unsigned const mid = (lo + hi) >> 1;
                if (w.fdist[mid] < d) lo = mid + 1; else hi = mid;
            }
            np = i + lo;
        }
        wave_sync();
        if (mv && np != i && np < Lf) {
            w.pid[np] = id;
            w.pdist[np] = d;
            w.pmeta[np] = meta;
            if (w.pbase) w.pbase[np] = bs;
        }
        wave_sync();
    }
    if (live && fin < Lf) {
        w.pid[fin] = sid | FLAG;
        w.pdist[fin] = sd;
        w.pmeta[fin] = sm << 16;
        if (w.pbase) w.pbase[fin] = sb;
    }
    wave_sync();
    L = Lf;
    if (fin0 < Lf && fin0 < k) k = fin0;
}

__device__ __forceinline__ uint32_t load_u32_l2(uint32_t const *p) {
    return __hip_atomic_load(const_cast<uint32_t *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Visited hash set in LDS.  Ids are < 2^31, so VT_EMPTY never collides; the
// set is kept at most half full plus one batch, so a probe always ends.  Its
// size is 2^k or 3 * 2^k entries: a hashed value's home slot is its product
// with the size, shifted (multiply-shift range reduction), and probes wrap.
constexpr uint32_t VT_EMPTY = 0xffffffffu;
__device__ __forceinline__ unsigned vt_next(unsigned h, unsigned vcap) { return h + 1u == vcap ? 0u : h + 1u; }
__device__ __forceinline__ unsigned vt_home32(uint32_t id, unsigned vcap) {
    return unsigned((uint64_t(id * 2654435761u) * vcap) >> 32);
}
// true when id was not in the set (it is now)
__device__ __forceinline__ bool vt_insert(uint32_t *vt, unsigned vcap, uint32_t id) {
    unsigned h = vt_home32(id, vcap);
    for (;;) {
        uint32_t const old = atomicCAS(&vt[h], VT_EMPTY, id);
        if (old == VT_EMPTY) return true;
        if (old == id) return false;
        h = vt_next(h, vcap);
    }
}
__device__ __forceinline__ bool vt_contains(uint32_t const *vt, unsigned vcap, uint32_t id) {
    unsigned h = vt_home32(id, vcap);
    for (;;) {
        uint32_t const v = vt[h];
        if (v == id) return true;
        if (v == VT_EMPTY) return false;
        h = vt_next(h, vcap);
    }
}

// Compact form: 16-bit entries, two per LDS word, for ids below 2^B
// (B = idbits).  p = id * 2654435761 mod 2^B is a bijection.  With a size
// of c 2^k entries (c = 1 or 3) and S = B - k, p's home slot is
// h = (c p) >> S and its remainder p - start(h), start(h) = ceil(h 2^S / c),
// which takes rb = S bits (c = 1) or S - 1 (c = 3).  Linear probing stores
// (displacement << rb) | remainder, so a slot and its entry give back the
// home slot, p and the id (p * 2654435761^-1).  Displacements stop at dmax,
// so no entry is all ones (free); a probe that would go further reports
// overflow and the caller moves the set into the bitset.
constexpr uint32_t VT_HASH_INV = 244002641u;  // 2654435761 * 244002641 = 1 mod 2^32
struct VtCfg {
    bool b16;
    unsigned vcap;                           // entries
    bool c3;                                 // vcap = 3 * 2^k (else 2^k)
    unsigned S, rb, rmask, dmax, idmask;     // 16-bit entries
};
__device__ __forceinline__ uint32_t vt16_start(VtCfg const &c, unsigned h) {
    return c.c3 ? ((h << c.S) + 2u) / 3u : (h << c.S);
}
enum : unsigned { VT_PRESENT = 0, VT_FRESH = 1, VT_OVERFLOW = 2 };
__device__ __forceinline__ unsigned vt16_insert(uint32_t *vt, VtCfg const &c, uint32_t id) {
    uint32_t const p = (id * 2654435761u) & c.idmask;
    unsigned slot = ((c.c3 ? 3u : 1u) * p) >> c.S;
    uint32_t const r = p - vt16_start(c, slot);
    for (unsigned d = 0; d <= c.dmax; ++d, slot = vt_next(slot, c.vcap)) {
        uint32_t const want = (d << c.rb) | r;
        uint32_t *const word = vt + (slot >> 1);
        unsigned const sh = (slot & 1u) * 16u;
        uint32_t old = *word;
        for (;;) {
            uint32_t const e = (old >> sh) & 0xffffu;
            if (e == want) return VT_PRESENT;
            if (e != 0xffffu) break;  // another id's: probe on
            uint32_t const got = atomicCAS(word, old, (old & ~(0xffffu << sh)) | (want << sh));
            if (got == old) return VT_FRESH;
            old = got;  // the word changed under this lane (its other half, or this one): look again
        }
    }
    return VT_OVERFLOW;
}
__device__ __forceinline__ bool vt16_contains(uint32_t const *vt, VtCfg const &c, uint32_t id) {
    uint32_t const p = (id * 2654435761u) & c.idmask;
    unsigned slot = ((c.c3 ? 3u : 1u) * p) >> c.S;
    uint32_t const r = p - vt16_start(c, slot);
    for (unsigned d = 0; d <= c.dmax; ++d, slot = vt_next(slot, c.vcap)) {
        uint32_t const e = (vt[slot >> 1] >> ((slot & 1u) * 16u)) & 0xffffu;
        if (e == ((d << c.rb) | r)) return true;
        if (e == 0xffffu) return false;
    }
    return false;
}
// id of the 16-bit entry e held in `slot`
__device__ __forceinline__ uint32_t vt16_id(VtCfg const &c, unsigned slot, uint32_t e) {
    unsigned const d = e >> c.rb;
    unsigned const home = slot >= d ? slot - d : slot + c.vcap - d;
    return ((vt16_start(c, home) + (e & c.rmask)) * VT_HASH_INV) & c.idmask;
}
__device__ __forceinline__ unsigned vt_put(uint32_t *vt, VtCfg const &c, uint32_t id) {
    if (c.b16) return vt16_insert(vt, c, id);
    return vt_insert(vt, c.vcap, id) ? VT_FRESH : VT_PRESENT;
}
__device__ __forceinline__ bool vt_has(uint32_t const *vt, VtCfg const &c, uint32_t id) {
    return c.b16 ? vt16_contains(vt, c, id) : vt_contains(vt, c.vcap, id);
}

#ifndef SEARCH_WAVES_PER_EU
#define SEARCH_WAVES_PER_EU 4
#endif
#ifndef SEARCH_LEAN_WAVES_PER_EU
#define SEARCH_LEAN_WAVES_PER_EU 5
#endif
// SCREEN (RowT = float): each batch of a full pool goes through the BF16
// screen (batch_screen) before the exact FP32 re-score.  QPW queries per
// wave: each walks on G = 64 / QPW lanes with its own LDS area (a.lds_per_wave
// bytes per query) and visited bitset; every loop steps G lanes.
// LEAN: the walk of a batch of integer queries over the 8-bit rows from
// caller-given starting points, one trial (no random seeds, no result merge,
// no FP32 query copy): fewer registers and less LDS, so 5 waves per SIMD.
template <int METRIC, typename RowT, bool SCREEN = false, int QPW = 1, bool LEAN = false>
__global__ void __launch_bounds__(SEARCH_MAX_WPB * KG_WAVE)
    __attribute__((amdgpu_waves_per_eu(LEAN ? SEARCH_LEAN_WAVES_PER_EU : SEARCH_WAVES_PER_EU)))
search_kernel(kg_search_args a) {
    static_assert(!LEAN || (sizeof(RowT) == 1 && !SCREEN && QPW == 1), "the lean walk: 8-bit rows, one query a wave");
    constexpr unsigned G = KG_WAVE / QPW;
    static_assert(!SCREEN || QPW == 1, "the BF16 screen runs one query per wave");
    extern __shared__ __align__(16) unsigned char smem[];
    Grp<G> const g;
    unsigned const lane = g.gl;  // lane within this query's group
    unsigned const wib = threadIdx.x / KG_WAVE;
    unsigned const qslot = wib * QPW + g.base / G;  // query slot within the block
    unsigned const slot = blockIdx.x * (blockDim.x / KG_WAVE) * QPW + qslot;
    unsigned const cap = a.pool_cap;
    unsigned const rcap = a.res_cap;
    LdsLayout lay(cap, rcap, a.Dp, a.seed_cap, a.vcap, a.vt_bits / 8, a.gpad == nullptr, SCREEN, LEAN);
    unsigned char *base = smem + size_t(qslot) * a.lds_per_wave;
    WaveLDS w;
    w.pid = reinterpret_cast<uint32_t *>(base + lay.pid);
    w.pdist = reinterpret_cast<float *>(base + lay.pdist);
    w.pmeta = reinterpret_cast<uint32_t *>(base + lay.pmeta);
    w.pbase = a.gpad ? nullptr : reinterpret_cast<uint32_t *>(base + lay.pbase);
    w.tkey = reinterpret_cast<uint64_t *>(base + lay.tkey);
    w.rid = reinterpret_cast<uint32_t *>(base + lay.rid);
    w.rdist = reinterpret_cast<float *>(base + lay.rdist);
    w.q = reinterpret_cast<float *>(base + lay.q);
    w.q8 = reinterpret_cast<uint32_t *>(base + lay.q8);
    w.qh = reinterpret_cast<uint4 *>(base + lay.qh);
    w.ql = reinterpret_cast<uint4 *>(base + lay.ql);
    w.sidx = base + lay.sidx;
    w.fid = reinterpret_cast<uint32_t *>(base + lay.fid);
    w.fdist = reinterpret_cast<float *>(base + lay.fdist);
    w.vt = reinterpret_cast<uint32_t *>(base + lay.vt);
    VtCfg vs;
    vs.b16 = a.vt_bits == 16;
    vs.vcap = a.vcap;
    vs.c3 = a.vcap && (a.vcap & (a.vcap - 1u)) != 0u;
    vs.S = a.vt_shift;
    vs.rb = a.vt_rb;
    vs.rmask = (1u << a.vt_rb) - 1u;
    vs.dmax = (1u << (16u - a.vt_rb)) - 2u;
    vs.idmask = a.vt_idmask;
    unsigned const vwords = vs.b16 ? a.vcap / 2 : a.vcap;  // LDS words of the set

    uint32_t *vis = a.visited + size_t(slot) * a.vis_words;
    uint32_t *touched = a.touched + size_t(slot) * a.touch_cap;
    uint64_t const lt_mask = g.lt();
    if (a.abort_flag && *a.abort_flag) return;  // a starting id is out of range: nothing is written
    // a lean / general pair of launches for one batch: the lean walk runs when
    // every query is an integer vector in [0, 255], the general one otherwise
    if (a.lean_flag && LEAN != (*a.lean_flag == 0u)) return;

    for (;;) {
        unsigned qi = 0;
        if (lane == 0) qi = atomicAdd(a.qcounter, 1u);
        qi = g.shfl(qi, 0u);
        if constexpr (QPW == 1) qi = uniform(qi);
        if (qi >= a.nq) break;
        if (a.qorder) {  // longest walks first
            qi = a.qorder[qi];
            if constexpr (QPW == 1) qi = uniform(qi);
        }
        uint64_t const t_start = __builtin_amdgcn_s_memrealtime();  // 100 MHz constant clock
        unsigned const K = a.Ks ? a.Ks[qi] : a.K0;
        unsigned const P = a.Ps ? a.Ps[qi] : a.P0;
        if (P >= a.N) continue;  // served by the brute-force launch (kgraph.cpp:427-433)
        unsigned const qcap = K + P + 1;

        float const *qsrc = a.Q + size_t(qi) * a.qstride;
        if constexpr (!LEAN)
            for (unsigned i = lane; i < a.Dp; i += G) w.q[i] = (i < a.D) ? qsrc[i] : 0.f;
        float const qn = (METRIC == METRIC_L2SQR) ? 0.f : a.qnorms[qi];
        // over the 8-bit rows, a query whose values are all integers in
        // [0, 255] takes the integer path (exact, so the same distances)
        QInt qv{false, 0u};
        if constexpr (LEAN) {
            // (every query of the batch is one: checked before the launch)
            uint32_t sq = 0;
            for (unsigned wd = lane; wd < a.Dp / 4; wd += G) {
                uint32_t word = 0;
                for (unsigned j = 0; j < 4; ++j) {
                    unsigned const c = 4 * wd + j;
                    uint32_t const u = c < a.D ? uint32_t(qsrc[c]) : 0u;
                    word |= u << (8 * j);
                    sq += u * u;
                }
                w.q8[wd] = word;
            }
            for (int o = G / 2; o > 0; o >>= 1) sq += __shfl_xor(sq, o);
            qv.on = true;
            qv.sq = sq;
        } else if constexpr (sizeof(RowT) == 1) {
            wave_sync();
            bool ok = true;
            uint32_t sq = 0;
            for (unsigned wd = lane; wd < a.Dp / 4; wd += G) {
                uint32_t word = 0;
                for (unsigned j = 0; j < 4; ++j) {
                    float const v = w.q[4 * wd + j];
                    ok &= v >= 0.f && v <= 255.f && v == rintf(v);
                    uint32_t const u = ok ? uint32_t(v) : 0u;
                    word |= u << (8 * j);
                    sq += u * u;
                }
                w.q8[wd] = word;
            }
            for (int o = G / 2; o > 0; o >>= 1) sq += __shfl_xor(sq, o);
            qv.on = a.int_dot && !g.ballot(!ok);
            qv.sq = sq;
        }
        wave_sync();
        // BF16 screen: the query's halves, |q|^2 and |q|
        float qsq = 0.f, qrn = 0.f;
        unsigned n_res = 0;
        if constexpr (SCREEN) {
            qsq = ws_query(w, a.D, a.Dp);
            qrn = (METRIC == METRIC_ANGULAR) ? qn : (METRIC == METRIC_EUCLIDEAN) ? sqrtf(2.f * qn) : sqrtf(qsq);
        }

        unsigned ntouch = 0, n_comps = 0, n_scan = 0, RL = 0;
        // visited ids: in the LDS set while it has room (vcount entries), else
        // in the HBM bitset (ntouch ids recorded for clearing)
        bool glob = a.vcap == 0;
        unsigned vcount = 0;
        for (unsigned i = lane; i < vwords; i += G) w.vt[i] = VT_EMPTY;
        wave_sync();
        // moves the LDS set into the bitset; later visits go there
        auto spill = [&]() {
            for (unsigned c = 0; c < vwords; c += G) {
                uint32_t const word = w.vt[c + lane];
                for (unsigned half = 0; half < (vs.b16 ? 2u : 1u); ++half) {
                    uint32_t e = word;
                    bool has = e != VT_EMPTY;
                    if (vs.b16) {
                        uint32_t const e16 = (word >> (16 * half)) & 0xffffu;
                        has = e16 != 0xffffu;
                        e = has ? vt16_id(vs, 2 * (c + lane) + half, e16) : 0u;
                    }
                    uint64_t const b = g.ballot(has);
                    if (has) {
                        atomicOr(&vis[e >> 5], 1u << (e & 31));
                        unsigned const pos = ntouch + popc64(b & lt_mask);
                        if (pos < a.touch_cap) touched[pos] = e;
                    }
                    ntouch += popc64(b);
                }
            }
            glob = true;
            if (lane == 0 && a.spills) atomicAdd(a.spills, 1u);
            wave_sync();
        };
        for (unsigned trial = 0; trial < (LEAN ? 1u : a.T); ++trial) {
            // ---- starting points (kgraph.cpp:450-467)
            unsigned L = 0;
            if (!LEAN && a.init == 0) {
                uint32_t const *seeds = a.seeds + a.seed_off[a.Ps ? a.pidx[qi] : 0] + size_t(trial) * P;
                for (unsigned c = 0; c < P; c += G) {
                    unsigned i = c + lane;
                    uint32_t s = 0;
                    bool keep = false;
                    if (i < P) {
                        s = seeds[i];
                        if (a.inv) s = a.inv[s];
                        keep = true;
                        if (trial > 0)
                            keep = glob ? !(load_u32_l2(&vis[s >> 5]) & (1u << (s & 31)))
                                        : !vt_has(w.vt, vs, s);
                    }
                    uint64_t b = g.ballot(keep);
                    if (keep) w.pid[L + popc64(b & lt_mask)] = s;
                    L += popc64(b);
                }
            } else {
                uint32_t const *ini = a.init_ids + size_t(qi) * a.ostride;
                L = min(a.init, K + P - 1);  // = init whenever init < K (one-K calls)
                for (unsigned i = lane; i < L; i += G) w.pid[i] = a.inv ? a.inv[ini[i]] : ini[i];
            }
            wave_sync();
            // ---- mark, score and sort the starting points (kgraph.cpp:468-476)
            if (!glob && 2 * (vcount + L + G) > a.vcap) spill();
            if (glob) {
                for (unsigned i = lane; i < L; i += G) {
                    uint32_t s = w.pid[i];
                    atomicOr(&vis[s >> 5], 1u << (s & 31));
                    if (ntouch + i < a.touch_cap) touched[ntouch + i] = s;
                }
                ntouch += L;
            } else {
                bool ovf = false;
                for (unsigned c = 0; c < L; c += G) {
                    unsigned const st = (c + lane < L) ? vt_put(w.vt, vs, w.pid[c + lane]) : VT_PRESENT;
                    vcount += popc64(g.ballot(st == VT_FRESH));
                    ovf |= st == VT_OVERFLOW;
                }
                if (g.ballot(ovf)) {  // a 16-bit probe ran out of displacement: the bitset takes over
                    spill();
                    for (unsigned i = lane; i < L; i += G) {
                        uint32_t s = w.pid[i];
                        atomicOr(&vis[s >> 5], 1u << (s & 31));
                        if (ntouch + i < a.touch_cap) touched[ntouch + i] = s;
                    }
                    ntouch += L;
                }
            }
            batch_dist<METRIC, RowT, G, LEAN>(a, w.pid, w.pdist, L, w.q, qn, w.q8, qv);
            // (dist, dataset row) order
            for (unsigned i = lane; i < L; i += G)
                w.tkey[i] = (uint64_t(ord_f32(w.pdist[i])) << 32) | (a.perm ? a.perm[w.pid[i]] : w.pid[i]);
            wave_sync();
            for (unsigned c = 0; c < L; c += G) {
                unsigned i = c + lane;
                uint64_t key = (i < L) ? w.tkey[i] : ~0ull;
                unsigned rank = 0;
                for (unsigned j = 0; j < L; ++j) {
                    uint64_t kj = w.tkey[j];
                    rank += (kj < key) || (kj == key && j < i);
                }
                if (i < L) {
                    w.pid[rank] = (a.inv ? a.inv[uint32_t(key)] : uint32_t(key)) | FLAG;
                    w.pdist[rank] = unord_f32(uint32_t(key >> 32));
                }
            }
            wave_sync();
            for (unsigned i = lane; i < L; i += G) {
                uint32_t bs = 0;
                uint32_t const m = a.gpad ? VT_M_UNKNOWN : actual_M(a, w.pid[i] & IDMASK, &bs);
                w.pmeta[i] = m << 16;
                if (w.pbase) w.pbase[i] = bs;
            }
            wave_sync();

            // ---- expansion loop (kgraph.cpp:478-514)
            unsigned k = 0;
            for (;;) {
                unsigned kk = L;
                for (unsigned c = k; c < L; c += G) {
                    unsigned i = c + lane;
                    uint64_t b = g.ballot(i < L && (w.pid[i] & FLAG));
                    if (b) {
                        kk = c + ctz64(b);
                        break;
                    }
                }
                if (kk >= L) break;
                k = kk;
                uint32_t const eid = w.pid[k];
                uint32_t const meta = w.pmeta[k];
                unsigned const beginM = meta & 0xffffu;
                unsigned eM = meta >> 16;
                // the node's list: its CSR row, or its padded row, whose
                // address needs nothing but the id, so its first ids load
                // together with its M and degree
                uint32_t const *erow =
                    a.gpad ? a.gpad + size_t(eid & IDMASK) * a.gstride : a.gids + w.pbase[k];
                uint32_t pre = 0xffffffffu;
                bool const first_pad = a.gpad && eM == VT_M_UNKNOWN;
                if (first_pad) {
                    uint32_t const md = a.gmd[eid & IDMASK];
                    if (lane < min(a.S, G) && beginM + lane < a.gstride) pre = erow[beginM + lane];
                    uint32_t m = md & 0xffffu, deg = md >> 16;
                    if (m < a.Mparam) m = a.Mparam;
                    if (m > deg) m = deg;
                    eM = m & 0xffffu;
                }
                unsigned endM = beginM + a.S;
                bool const done = endM > eM;
                if (done) endM = eM;
                wave_sync();
                if (lane == 0) {
                    w.pmeta[k] = (eM << 16) | (endM & 0xffffu);
                    if (done) w.pid[k] = eid & IDMASK;
                }
                wave_sync();
                if (endM > beginM) n_scan += endM - beginM;
                for (unsigned mc = beginM; mc < endM; mc += G) {
                    unsigned const cnt = min(endM - mc, G);
                    uint32_t id = 0xffffffffu;
                    if (lane < cnt) id = (first_pad && mc == beginM) ? pre : erow[mc + lane];
                    bool dup = false;
                    if (!a.lists_unique) {  // a list may repeat an id: keep its first occurrence
                        for (unsigned j = 0; j < cnt; ++j) {
                            uint32_t v = g.shfl(id, j);
                            dup |= (j < lane) && (v == id);
                        }
                    }
                    if (glob && cnt <= G / 4) {
                        // the visited bits and the distances at once: the
                        // bitset atomics and the row gathers are in flight
                        // together, and the distances of ids found visited
                        // are dropped (16 rows take one pass either way)
                        bool const cand = lane < cnt && !dup;
                        uint32_t const bit = 1u << (id & 31);
                        uint32_t old = 0u;
                        if (cand) old = atomicOr(&vis[id >> 5], bit);
                        if (lane < cnt) w.fid[lane] = id;
                        wave_sync();
                        if (SCREEN && L == qcap - 1)
                            batch_screen<METRIC>(a, w, w.fid, w.fdist, cnt, qn, qrn, qsq, w.pdist[L - 1], n_res, true,
                                                 cand, old, bit);
                        else
                            batch_dist<METRIC, RowT, G, LEAN>(a, w.fid, w.fdist, cnt, w.q, qn, w.q8, qv);
                        bool const fresh = cand && !(old & bit);
                        uint64_t const b = g.ballot(fresh);
                        unsigned const nf = popc64(b);
                        unsigned const pos = popc64(b & lt_mask);
                        float const fd = fresh ? w.fdist[lane] : 0.f;
                        wave_sync();
                        if (fresh) {
                            w.fid[pos] = id;
                            w.fdist[pos] = fd;
                            if (ntouch + pos < a.touch_cap) touched[ntouch + pos] = id;
                        }
                        ntouch += nf;
                        wave_sync();
                        if (nf == 0) continue;
                        uint32_t cm = VT_M_UNKNOWN, cb = 0;
                        if (lane < nf && !a.gpad) cm = actual_M(a, w.fid[lane], &cb);
                        n_comps += nf;
                        batch_insert<G>(w, nf, cm, cb, qcap, L, k);
                        continue;
                    }
                    bool fresh = false;
                    unsigned st = VT_PRESENT;
                    if (lane < cnt && !dup) {
                        if (glob) {
                            uint32_t bit = 1u << (id & 31);
                            fresh = !(atomicOr(&vis[id >> 5], bit) & bit);
                        } else {
                            st = vt_put(w.vt, vs, id);
                            fresh = st == VT_FRESH;
                        }
                    }
                    if (!glob && g.ballot(st == VT_OVERFLOW)) {
                        // a 16-bit probe ran out of displacement: move the set
                        // (this batch's inserts included) into the bitset, and
                        // test the rest there
                        spill();
                        if (st == VT_OVERFLOW) {
                            uint32_t bit = 1u << (id & 31);
                            fresh = !(atomicOr(&vis[id >> 5], bit) & bit);
                        }
                    }
                    uint64_t b = g.ballot(fresh);
                    unsigned const nf = popc64(b);
                    unsigned pos = popc64(b & lt_mask);
                    if (fresh) {
                        w.fid[pos] = id;
                        if (glob && ntouch + pos < a.touch_cap) touched[ntouch + pos] = id;
                    }
                    if (glob)
                        ntouch += nf;
                    else
                        vcount += nf;
                    wave_sync();
                    if (!glob && 2 * (vcount + G) > a.vcap) spill();
                    if (nf == 0) continue;
                    // lane j fetches candidate j's M and CSR base while the rows load
                    uint32_t cm = VT_M_UNKNOWN, cb = 0;
                    if (lane < nf && !a.gpad) cm = actual_M(a, w.fid[lane], &cb);
                    if (SCREEN && L == qcap - 1)
                        batch_screen<METRIC>(a, w, w.fid, w.fdist, nf, qn, qrn, qsq, w.pdist[L - 1], n_res);
                    else
                        batch_dist<METRIC, RowT, G, LEAN>(a, w.fid, w.fdist, nf, w.q, qn, w.q8, qv);
                    // the nf inserts of this batch, done at once (batch_insert)
                    n_comps += nf;
                    batch_insert<G>(w, nf, cm, cb, qcap, L, k);
                }
            }
            // ---- merge this trial into the results (kgraph.cpp:515-529)
            if (L > K) L = K;
            if (LEAN || a.T == 1) {
                RL = L;
            } else if (trial == 0) {
                for (unsigned i = lane; i < L; i += G) {
                    w.rid[i] = w.pid[i] & IDMASK;
                    w.rdist[i] = w.pdist[i];
                }
                RL = L;
                wave_sync();
            } else {
                for (unsigned l = 0; l < L; ++l) {
                    float const cd = w.pdist[l];
                    uint32_t const cid = w.pid[l] & IDMASK;
                    unsigned const r = insert_pos<G>(w.rdist, RL, cd);
                    wave_sync();
                    // shift results[r..RL) up by one (slot RL is the scratch)
                    for (int c = int(RL) - 1; c >= int(r); c -= G) {
                        int i = c - int(lane);
                        uint32_t id = 0;
                        float d = 0.f;
                        if (i >= int(r)) {
                            id = w.rid[i];
                            d = w.rdist[i];
                        }
                        wave_sync();
                        if (i >= int(r)) {
                            w.rid[i + 1] = id;
                            w.rdist[i + 1] = d;
                        }
                        wave_sync();
                    }
                    if (lane == 0) {
                        w.rid[r] = cid;
                        w.rdist[r] = cd;
                    }
                    wave_sync();
                    if (RL + 1 < K + 1) ++RL;
                }
            }
        }
        // ---- epsilon cut and output (kgraph.cpp:531-563)
        uint32_t const *oid = (LEAN || a.T == 1) ? w.pid : w.rid;
        float const *odist = (LEAN || a.T == 1) ? w.pdist : w.rdist;
        unsigned cut = RL;
        for (unsigned c = 0; c < RL; c += G) {
            unsigned i = c + lane;
            uint64_t b = g.ballot(i < RL && odist[i] > a.epsilon);
            if (b) {
                cut = c + ctz64(b);
                break;
            }
        }
        for (unsigned i = lane; i < cut; i += G) {
            uint32_t const id = oid[i] & IDMASK;
            a.out_ids[size_t(qi) * a.ostride + i] = a.perm ? a.perm[id] : id;
            if (a.out_dists) a.out_dists[size_t(qi) * a.ostride + i] = odist[i];
        }
        if (lane == 0) {
            a.counts[qi] = cut;
            if (a.comps) a.comps[qi] = n_comps;
            if (a.scans) a.scans[qi] = n_scan;
            if (a.lat_us) a.lat_us[qi] = float(__builtin_amdgcn_s_memrealtime() - t_start) * 0.01f;
            if (SCREEN && a.rescored) atomicAdd(a.rescored, (unsigned long long)n_res);
        }
        // ---- clear the visited bits this query set (none while the LDS set held them)
        if (!glob) {
        } else if (ntouch <= a.touch_cap) {
            for (unsigned i = lane; i < ntouch; i += G) vis[load_u32_l2(&touched[i]) >> 5] = 0u;
        } else {
            for (unsigned i = lane; i < a.vis_words; i += G) vis[i] = 0u;
        }
        __threadfence();
        wave_sync();
    }
}

// Per-node record for the walk, gnode[n] = {goff[n], deg, M[n], 0}, and
// whether any adjacency list repeats an id (dupflag, set only when one does):
// one wave per node, the list in LDS when it fits, else read twice from HBM.
constexpr unsigned GN_LDS = 1024;
__global__ void __launch_bounds__(256) gnode_kernel(unsigned N, uint32_t const *goff, uint32_t const *gids,
                                                    uint32_t const *gM, uint4 *gnode, uint32_t *dupflag) {
    __shared__ uint32_t sh[4][GN_LDS];
    unsigned const w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    unsigned const n = blockIdx.x * 4 + w;
    if (n >= N) return;
    uint32_t const b = goff[n], deg = goff[n + 1] - b;
    if (lane == 0) gnode[n] = make_uint4(b, deg, gM[n], 0u);
    uint32_t const *l = gids + b;
    bool dup = false;
    if (deg <= GN_LDS) {
        for (unsigned j = lane; j < deg; j += KG_WAVE) sh[w][j] = l[j];
        wave_sync();
        for (unsigned j = lane; j < deg; j += KG_WAVE) {
            uint32_t const v = sh[w][j];
            for (unsigned k = 0; k < j && !dup; ++k) dup = sh[w][k] == v;
        }
    } else {
        for (unsigned j = lane; j < deg; j += KG_WAVE) {
            uint32_t const v = l[j];
            for (unsigned k = 0; k < j && !dup; ++k) dup = l[k] == v;
        }
    }
    if (ballot(dup) && lane == 0) atomicOr(dupflag, 1u);
}

__global__ void __launch_bounds__(256) gpad_width_kernel(unsigned N, uint32_t const *goff, uint32_t const *gM,
                                                         uint32_t *width) {
    unsigned mx = 0;
    for (unsigned n = blockIdx.x * 256 + threadIdx.x; n < N; n += gridDim.x * 256)
        mx = max(mx, min(gM[n], goff[n + 1] - goff[n]));
    for (int o = 32; o > 0; o >>= 1) mx = max(mx, unsigned(__shfl_xor(int(mx), o)));
    if ((threadIdx.x & 63) == 0 && mx) atomicMax(width, mx);
}

// one wave per node: the first min(deg, gstride) ids, coalesced
__global__ void __launch_bounds__(256) gpad_kernel(unsigned N, uint32_t const *goff, uint32_t const *gids,
                                                   uint32_t const *gM, unsigned gstride, uint32_t *gpad,
                                                   uint32_t *gmd) {
    unsigned const w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    unsigned const n = blockIdx.x * 4 + w;
    if (n >= N) return;
    uint32_t const b = goff[n], deg = goff[n + 1] - b;
    unsigned const cnt = min(deg, gstride);
    uint32_t *row = gpad + size_t(n) * gstride;
    for (unsigned j = lane; j < cnt; j += KG_WAVE) row[j] = gids[b + j];
    if (lane == 0) gmd[n] = (min(deg, 0xffffu) << 16) | min(gM[n], 0xffffu);
}

// *flag = 1 when a value of the nq x D queries is not an integer in [0, 255]
// (the caller zeroes it)
__global__ void __launch_bounds__(256) q8check_kernel(float const *Q, unsigned nq, unsigned D, size_t stride,
                                                      uint32_t *flag) {
    size_t const total = size_t(nq) * D;
    bool bad = false;
    for (size_t i = size_t(blockIdx.x) * 256 + threadIdx.x; i < total; i += size_t(gridDim.x) * 256) {
        size_t const q = i / D, j = i - q * D;
        float const v = Q[q * stride + j];
        bad |= !(v >= 0.f && v <= 255.f && v == rintf(v));
    }
    if (ballot(bad) && (threadIdx.x & 63) == 0) *flag = 1u;
}

// The starting-id check; with Q, also q8check_kernel's test of the queries
// into qflag (one launch for both).
__global__ void __launch_bounds__(256) init_check_kernel(uint32_t const *ids, unsigned nq, size_t ostride,
                                                         unsigned init, unsigned N, uint32_t *flag,
                                                         uint32_t *host_flag, float const *Q, unsigned D,
                                                         size_t qstride, uint32_t *qflag) {
    size_t const total = size_t(nq) * init;
    bool bad = false;
    for (size_t i = size_t(blockIdx.x) * 256 + threadIdx.x; i < total; i += size_t(gridDim.x) * 256) {
        size_t const q = i / init, j = i - q * init;
        bad |= ids[q * ostride + j] >= N;
    }
    if (bad) {
        *flag = 1u;
        *host_flag = 1u;
    }
    if (Q) {
        size_t const qtotal = size_t(nq) * D;
        bool q8bad = false;
        for (size_t i = size_t(blockIdx.x) * 256 + threadIdx.x; i < qtotal; i += size_t(gridDim.x) * 256) {
            size_t const q = i / D, j = i - q * D;
            float const v = Q[q * qstride + j];
            q8bad |= !(v >= 0.f && v <= 255.f && v == rintf(v));
        }
        if (ballot(q8bad) && (threadIdx.x & 63) == 0) *qflag = 1u;
    }
}

// Screened distances and bounds of m candidates per query (ids row qi), one
// wave per query, with the walk's screen arithmetic (ws_query, ws_mfma,
// ws_finish).
constexpr unsigned WS_MAX_DP = 256;
template <int METRIC>
__global__ void __launch_bounds__(KG_WAVE) screen_dists_kernel(kg_search_args a, uint32_t const *ids, unsigned m,
                                                                float *out_d, float *out_e) {
    __shared__ __align__(16) float q[WS_MAX_DP];
    __shared__ uint4 qh[WS_MAX_DP / 8], ql[WS_MAX_DP / 8];
    WaveLDS w{};
    w.q = q;
    w.qh = qh;
    w.ql = ql;
    unsigned const qi = blockIdx.x, lane = lane_id();
    for (unsigned i = lane; i < a.Dp; i += KG_WAVE) q[i] = i < a.D ? a.Q[size_t(qi) * a.qstride + i] : 0.f;
    wave_sync();
    float const qn = (METRIC == METRIC_L2SQR) ? 0.f : a.qnorms[qi];
    float const qsq = ws_query(w, a.D, a.Dp);
    float const qrn = (METRIC == METRIC_ANGULAR) ? qn : (METRIC == METRIC_EUCLIDEAN) ? sqrtf(2.f * qn) : sqrtf(qsq);
    unsigned const n = lane & 15;
    bool const own = (n >> 2) == (lane >> 4);
    for (unsigned b0 = 0; b0 < m; b0 += 16) {
        unsigned const c = b0 + n;
        bool const ok = c < m;
        uint32_t const id = ok ? ids[size_t(qi) * m + c] : 0u;
        float2 const gn = ws_mfma<METRIC>(a, id, ok, qh, ql);
        if (own && ok) {
            float const na = (METRIC == METRIC_L2SQR) ? 0.f : a.norms[id];
            float2 const de = ws_finish<METRIC>(a, gn.x, gn.y, na, qn, qrn, qsq);
            out_d[size_t(qi) * m + c] = de.x;
            out_e[size_t(qi) * m + c] = de.y;
        }
    }
}

}  // namespace

hipError_t kg_launch_gpad_width(unsigned N, uint32_t const *goff, uint32_t const *gM, uint32_t *width, hipStream_t s) {
    if (N == 0) return hipSuccess;
    hipLaunchKernelGGL(gpad_width_kernel, dim3(std::min((N + 255) / 256, 1024u)), dim3(256), 0, s, N, goff, gM, width);
    return hipGetLastError();
}

hipError_t kg_launch_gpad(unsigned N, uint32_t const *goff, uint32_t const *gids, uint32_t const *gM, unsigned gstride,
                          uint32_t *gpad, uint32_t *gmd, hipStream_t s) {
    if (N == 0) return hipSuccess;
    hipLaunchKernelGGL(gpad_kernel, dim3((N + 3) / 4), dim3(256), 0, s, N, goff, gids, gM, gstride, gpad, gmd);
    return hipGetLastError();
}

hipError_t kg_launch_q8check(float const *Q, unsigned nq, unsigned D, size_t stride, uint32_t *flag, hipStream_t s) {
    size_t const total = size_t(nq) * D;
    if (total == 0) return hipSuccess;
    unsigned const blocks = unsigned(std::min<size_t>((total + 255) / 256, 2048));
    hipLaunchKernelGGL(q8check_kernel, dim3(blocks), dim3(256), 0, s, Q, nq, D, stride, flag);
    return hipGetLastError();
}

hipError_t kg_launch_init_check(uint32_t const *ids, unsigned nq, size_t ostride, unsigned init, unsigned N,
                                uint32_t *flag, uint32_t *host_flag, float const *Q, unsigned D, size_t qstride,
                                uint32_t *qflag, hipStream_t s) {
    size_t const total = std::max(size_t(nq) * init, Q ? size_t(nq) * D : size_t(0));
    if (total == 0) return hipSuccess;
    unsigned const blocks = unsigned(std::min<size_t>((total + 255) / 256, 4096));
    hipLaunchKernelGGL(init_check_kernel, dim3(blocks), dim3(256), 0, s, ids, nq, ostride, init, N, flag,
                       host_flag, Q, D, qstride, qflag);
    return hipGetLastError();
}

hipError_t kg_launch_gnode(unsigned N, uint32_t const *goff, uint32_t const *gids, uint32_t const *gM, uint4 *gnode,
                           uint32_t *dupflag, hipStream_t s) {
    if (N == 0) return hipSuccess;
    hipLaunchKernelGGL(gnode_kernel, dim3((N + 3) / 4), dim3(256), 0, s, N, goff, gids, gM, gnode, dupflag);
    return hipGetLastError();
}

size_t kg_search_lds_per_wave(unsigned pool_cap, unsigned res_cap, unsigned Dp, unsigned seed_cap, unsigned vcap,
                              unsigned vt_bits, bool csr_base, bool screen, bool lean) {
    return LdsLayout(pool_cap, res_cap, Dp, seed_cap, vcap, vt_bits / 8, csr_base, screen, lean).total;
}

unsigned kg_search_max_wpb() { return SEARCH_MAX_WPB; }

unsigned kg_search_waves_per_cu() { return 4 * SEARCH_WAVES_PER_EU; }

unsigned kg_search_lean_waves_per_cu() { return 4 * SEARCH_LEAN_WAVES_PER_EU; }

hipError_t kg_launch_search(int metric, kg_search_args const &a, unsigned blocks, unsigned waves_per_block,
                            unsigned queries_per_wave, hipStream_t s) {
    if (waves_per_block == 0 || waves_per_block > SEARCH_MAX_WPB || blocks == 0) return hipErrorInvalidValue;
    if (queries_per_wave != 1 && queries_per_wave != 2) return hipErrorInvalidValue;
    size_t lds = size_t(waves_per_block) * queries_per_wave * a.lds_per_wave;
    dim3 grid(blocks), block(waves_per_block * KG_WAVE);
    // row form: the 8-bit copy, the BF16 copy, the FP32 rows behind the BF16
    // screen, or the FP32 rows; one or two queries per wave
    if (a.Xs && (a.X8 || a.Xb || a.Dp > 256 || queries_per_wave != 1)) return hipErrorInvalidValue;
    if (a.lean && (!a.X8 || queries_per_wave != 1 || a.T != 1 || a.init == 0 || !a.lean_flag))
        return hipErrorInvalidValue;
    auto go = [&](auto kernel8, auto kernel16, auto kernel32s, auto kernel32, auto kernel8x2, auto kernel16x2,
                  auto kernel32x2, auto kernel8lean) {
        bool const two = queries_per_wave == 2;
        if (a.lean)
            hipLaunchKernelGGL(kernel8lean, grid, block, lds, s, a);
        else if (a.X8)
            hipLaunchKernelGGL(two ? kernel8x2 : kernel8, grid, block, lds, s, a);
        else if (a.Xb)
            hipLaunchKernelGGL(two ? kernel16x2 : kernel16, grid, block, lds, s, a);
        else if (a.Xs)
            hipLaunchKernelGGL(kernel32s, grid, block, lds, s, a);
        else
            hipLaunchKernelGGL(two ? kernel32x2 : kernel32, grid, block, lds, s, a);
    };
    switch (metric) {
        case METRIC_L2SQR:
            go(search_kernel<METRIC_L2SQR, uint8_t>, search_kernel<METRIC_L2SQR, uint16_t>,
               search_kernel<METRIC_L2SQR, float, true>, search_kernel<METRIC_L2SQR, float>,
               search_kernel<METRIC_L2SQR, uint8_t, false, 2>, search_kernel<METRIC_L2SQR, uint16_t, false, 2>,
               search_kernel<METRIC_L2SQR, float, false, 2>, search_kernel<METRIC_L2SQR, uint8_t, false, 1, true>);
            break;
        case METRIC_EUCLIDEAN:
            go(search_kernel<METRIC_EUCLIDEAN, uint8_t>, search_kernel<METRIC_EUCLIDEAN, uint16_t>,
               search_kernel<METRIC_EUCLIDEAN, float, true>, search_kernel<METRIC_EUCLIDEAN, float>,
               search_kernel<METRIC_EUCLIDEAN, uint8_t, false, 2>, search_kernel<METRIC_EUCLIDEAN, uint16_t, false, 2>,
               search_kernel<METRIC_EUCLIDEAN, float, false, 2>, search_kernel<METRIC_EUCLIDEAN, uint8_t, false, 1, true>);
            break;
        default:
            go(search_kernel<METRIC_ANGULAR, uint8_t>, search_kernel<METRIC_ANGULAR, uint16_t>,
               search_kernel<METRIC_ANGULAR, float, true>, search_kernel<METRIC_ANGULAR, float>,
               search_kernel<METRIC_ANGULAR, uint8_t, false, 2>, search_kernel<METRIC_ANGULAR, uint16_t, false, 2>,
               search_kernel<METRIC_ANGULAR, float, false, 2>, search_kernel<METRIC_ANGULAR, uint8_t, false, 1, true>);
            break;
    }
    return hipGetLastError();
}

void kg_search_screen_constants(unsigned D, kg_search_args &a) {
    unsigned const Dp = (D + 3) & ~3u;
    double const eps = std::ldexp(1.0, -24);
    double const u = std::ldexp(1.0, -8);                        // BF16 rounding of a row value
    double const gn = (64.0 * ((Dp + 31) / 32) + 16) * eps;      // MFMA accumulation over both query halves
    double const gd = (D + 16) * eps;                            // the exact FP32 sums over D values
    a.scr_u = float(u * (1 + std::ldexp(1.0, -20)));
    a.scr_c = float((std::ldexp(1.0, -16) * (1 + u) + (gn + gd) * (1 + u) * (1 + u)) * 1.05);
    a.scr_g = float((gn + gd + 4 * eps) * 1.05);
}

hipError_t kg_launch_screen_dists(int metric, kg_search_args const &a, uint32_t const *ids, unsigned n, float *out_d,
                                  float *out_e, hipStream_t s) {
    if (a.Dp > WS_MAX_DP || !a.Xs) return hipErrorInvalidValue;
    if (a.nq == 0 || n == 0) return hipSuccess;
    switch (metric) {
        case METRIC_L2SQR:
            hipLaunchKernelGGL(screen_dists_kernel<METRIC_L2SQR>, dim3(a.nq), dim3(KG_WAVE), 0, s, a, ids, n, out_d, out_e);
            break;
        case METRIC_EUCLIDEAN:
            hipLaunchKernelGGL(screen_dists_kernel<METRIC_EUCLIDEAN>, dim3(a.nq), dim3(KG_WAVE), 0, s, a, ids, n, out_d,
                               out_e);
            break;
        default:
            hipLaunchKernelGGL(screen_dists_kernel<METRIC_ANGULAR>, dim3(a.nq), dim3(KG_WAVE), 0, s, a, ids, n, out_d,
                               out_e);
            break;
    }
    return hipGetLastError();
}

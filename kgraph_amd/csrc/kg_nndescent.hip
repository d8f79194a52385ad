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
//   join_kernel    kgraph.cpp:739-751, 797-815  every new x new (by id, i < j)
//                  and new x old pair; candidates below the target's radius go
//                  to the target's candidate slots (atomic slot claim, then a
//                  64-bit atomicMin so an overflowing slot keeps the better key)
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

// GenRandom (kgraph.cpp:44-65) on one lane; `addr` in LDS.
__device__ void gen_random(Rng &rng, uint32_t *addr, unsigned size, unsigned N) {
    if (N == size) {
        for (unsigned i = 0; i < size; ++i) addr[i] = i;
        return;
    }
    for (unsigned i = 0; i < size; ++i) addr[i] = rng() % (N - size);
    for (unsigned i = 1; i < size; ++i) {  // insertion sort
        uint32_t v = addr[i];
        int j = int(i) - 1;
        while (j >= 0 && addr[j] > v) {
            addr[j + 1] = addr[j];
            --j;
        }
        addr[j + 1] = v;
    }
    for (unsigned i = 1; i < size; ++i)
        if (addr[i] <= addr[i - 1]) addr[i] = addr[i - 1] + 1;
    unsigned off = rng() % N;
    for (unsigned i = 0; i < size; ++i) addr[i] = (addr[i] + off) % N;
}

// Canonical metric of data rows i and j (both global), one quad.
template <int METRIC>
__device__ __forceinline__ float pair_dist(NndArgs const &a, uint32_t i, float const *rj, float nj, unsigned t) {
    float const *ri = a.X + size_t(i) * a.Dp;
    float acc = quad_metric<METRIC>(ri, rj, a.D, t);
    float ni = (METRIC == METRIC_L2SQR) ? 0.f : a.norms[i];
    return finish_metric<METRIC>(acc, ni, nj);
}

// ------------------------------------------------------------------ init
template <int METRIC>
__global__ void __launch_bounds__(256) init_kernel(NndArgs a) {
    __shared__ uint32_t rnd[4][NND_MAX_S + 1];
    __shared__ uint32_t pid[4][NND_MAX_S];
    __shared__ float pd[4][NND_MAX_S];
    unsigned const w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    unsigned const n = blockIdx.x * 4 + w;
    if (n >= a.N) return;
    unsigned const S = a.S;
    if (lane == 0) {
        Rng rng{mix64(a.seed * 0x100000001b3ull + n)};
        unsigned nn = a.nnew_init;
        uint32_t *dst = a.samp + size_t(n) * a.Lc;
        // nn_new: 2S random ids (kgraph.cpp:765, 781), GenRandom written out
        // on the node's own sample row.
        if (a.N == nn) {
            for (unsigned i = 0; i < nn; ++i) dst[i] = i;
        } else {
            for (unsigned i = 0; i < nn; ++i) dst[i] = rng() % (a.N - nn);
            for (unsigned i = 1; i < nn; ++i) {
                uint32_t v = dst[i];
                int j = int(i) - 1;
                while (j >= 0 && dst[j] > v) {
                    dst[j + 1] = dst[j];
                    --j;
                }
                dst[j + 1] = v;
            }
            for (unsigned i = 1; i < nn; ++i)
                if (dst[i] <= dst[i - 1]) dst[i] = dst[i - 1] + 1;
            unsigned off = rng() % a.N;
            for (unsigned i = 0; i < nn; ++i) dst[i] = (dst[i] + off) % a.N;
        }
        // pool: S random non-self ids (kgraph.cpp:782-792)
        gen_random(rng, rnd[w], S + 1, a.N);
        unsigned i = 0;
        for (unsigned l = 0; l < S; ++l) {
            if (rnd[w][i] == n) ++i;
            pid[w][l] = rnd[w][i++];
        }
    }
    wave_sync();
    unsigned const grp = lane >> 2, t = lane & 3;
    float const *rn = a.X + size_t(n) * a.Dp;
    float const nn_ = (METRIC == METRIC_L2SQR) ? 0.f : a.norms[n];
    for (unsigned b = 0; b < S; b += 16) {
        if (b + grp < S) {
            float d = pair_dist<METRIC>(a, pid[w][b + grp], rn, nn_, t);
            if (t == 0) pd[w][b + grp] = d;
        }
    }
    wave_sync();
    if (lane == 0) {
        uint64_t *pool = a.pool + size_t(n) * a.L;
        // sort S keys by (dist, id) (kgraph.cpp:793)
        for (unsigned l = 0; l < S; ++l) pool[l] = make_key(pd[w][l], pid[w][l], 1);
        for (unsigned l = 1; l < S; ++l) {
            uint64_t v = pool[l];
            int j = int(l) - 1;
            while (j >= 0 && pool[j] > v) {
                pool[j + 1] = pool[j];
                --j;
            }
            pool[j + 1] = v;
        }
        a.cnt[n] = S;
        a.Mv[n] = S;
        a.nnew[n] = a.nnew_init;
        a.nold[n] = 0;
        a.crnew[n] = 0;
        a.crold[n] = 0;
        a.radius[n] = (S >= a.L) ? key_dist(pool[a.L - 1]) : __builtin_huge_valf();
        a.kth[n] = (S >= a.K) ? key_dist(pool[a.K - 1]) : __builtin_huge_valf();
        a.ccnt[n] = 0;
    }
}

// ------------------------------------------------------------------ join
__device__ __forceinline__ void push_cand(NndArgs const &a, uint32_t target, uint32_t id, float d) {
    uint64_t key = make_key(d, id, 1);
    unsigned slot = atomicAdd(&a.ccnt[target], 1u);
    if (slot >= a.C) slot = unsigned(mix64((uint64_t(target) << 32) | id) % a.C);
    atomicMin(reinterpret_cast<unsigned long long *>(&a.cand[size_t(target) * a.C + slot]),
              static_cast<unsigned long long>(key));
}

// One block per node.  The "new" rows are staged in LDS in chunks of ACH;
// each quad streams one other row (new or old) through registers and scores
// it against every staged row.
template <int METRIC, int CH>
__global__ void __launch_bounds__(256) join_kernel(NndArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    unsigned const n = blockIdx.x;
    uint32_t *Aid = reinterpret_cast<uint32_t *>(smem);
    uint32_t *Bid = Aid + a.Acap;
    float *radA = reinterpret_cast<float *>(Bid + a.Bcap);
    float *kthA = radA + NND_ACH;
    float *nrmA = kthA + NND_ACH;
    float *rowsA = nrmA + NND_ACH;  // NND_ACH x Dp, 16-byte aligned (Acap, Bcap multiples of 4)
    __shared__ unsigned s_found;
    __shared__ unsigned long long s_pairs, s_push;

    unsigned const nn = a.nnew[n], no = a.nold[n];
    unsigned const rn = min(a.crnew[n], a.R), ro = min(a.crold[n], a.R);
    unsigned const nA = nn + rn, nB = no + ro;
    uint32_t const *samp = a.samp + size_t(n) * a.Lc;
    for (unsigned i = threadIdx.x; i < nA; i += blockDim.x)
        Aid[i] = (i < nn) ? samp[i] : a.rnew[size_t(n) * a.R + (i - nn)];
    for (unsigned i = threadIdx.x; i < nB; i += blockDim.x)
        Bid[i] = (i < no) ? samp[a.Lc - 1 - i] : a.rold[size_t(n) * a.R + (i - no)];
    if (threadIdx.x == 0) {
        s_found = 0;
        s_pairs = 0;
        s_push = 0;
    }
    __syncthreads();

    unsigned const lane = threadIdx.x & 63;
    unsigned const quad = threadIdx.x >> 2, t = lane & 3;
    unsigned const nquads = blockDim.x >> 2;
    unsigned const nch = (a.D & ~15u) >> 2;
    bool found = false;
    unsigned long long pairs = 0, pushes = 0;

    for (unsigned a0 = 0; a0 < nA; a0 += NND_ACH) {
        unsigned const ach = min(nA - a0, unsigned(NND_ACH));
        // stage rows A[a0..a0+ach)
        unsigned const v4 = a.Dp >> 2;
        for (unsigned e = threadIdx.x; e < ach * v4; e += blockDim.x) {
            unsigned r = e / v4, c = e - r * v4;
            reinterpret_cast<float4 *>(rowsA)[size_t(r) * v4 + c] =
                reinterpret_cast<float4 const *>(a.X + size_t(Aid[a0 + r]) * a.Dp)[c];
        }
        for (unsigned r = threadIdx.x; r < ach; r += blockDim.x) {
            uint32_t id = Aid[a0 + r];
            radA[r] = a.radius[id];
            kthA[r] = a.kth[id];
            nrmA[r] = (METRIC == METRIC_L2SQR) ? 0.f : a.norms[id];
        }
        __syncthreads();
        for (unsigned oi = quad; oi < nA + nB; oi += nquads) {
            bool const isB = oi >= nA;
            uint32_t const o = isB ? Bid[oi - nA] : Aid[oi];
            float const *orow = a.X + size_t(o) * a.Dp;
            float4 reg[CH > 0 ? CH : 1];
            if (CH > 0) {
#pragma unroll
                for (int k = 0; k < CH; ++k) {
                    unsigned kk = t + 4 * k;
                    reg[k] = (kk < nch) ? reinterpret_cast<float4 const *>(orow)[kk] : make_float4(0, 0, 0, 0);
                }
            }
            float const ro_ = a.radius[o];
            float const no_ = (METRIC == METRIC_L2SQR) ? 0.f : a.norms[o];
            for (unsigned ai = 0; ai < ach; ++ai) {
                uint32_t const av = Aid[a0 + ai];
                bool const valid = isB ? (av != o) : (av < o);
                if (!valid) continue;
                float const *ra = rowsA + size_t(ai) * a.Dp;
                float acc;
                if (CH > 0) {
                    float4 p = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
                    for (int k = 0; k < CH; ++k) {
                        unsigned kk = t + 4 * k;
                        if (kk < nch) {
                            float4 y = reinterpret_cast<float4 const *>(ra)[kk];
                            float4 x = reg[k];
                            if (METRIC == METRIC_L2SQR) {
                                float dx = y.x - x.x, dy = y.y - x.y, dz = y.z - x.z, dw = y.w - x.w;
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
                    }
                    acc = quad_reduce(p);
                    for (unsigned i = a.D & ~15u; i < a.D; ++i) {
                        if (METRIC == METRIC_L2SQR) {
                            float dd = ra[i] - orow[i];
                            acc = __builtin_fmaf(dd, dd, acc);
                        } else {
                            acc = __builtin_fmaf(ra[i], orow[i], acc);
                        }
                    }
                } else {
                    acc = quad_metric<METRIC>(ra, orow, a.D, t);
                }
                float const d = finish_metric<METRIC>(acc, nrmA[ai], no_);
                if (t == 0) {
                    ++pairs;
                    if (d < radA[ai]) {
                        push_cand(a, av, o, d);
                        ++pushes;
                    }
                    if (d < ro_) {
                        push_cand(a, o, av, d);
                        ++pushes;
                    }
                    found |= d < kthA[ai];
                }
            }
        }
        __syncthreads();
    }
    if (found) s_found = 1;
    if (pairs) atomicAdd(&s_pairs, pairs);
    if (pushes) atomicAdd(&s_push, pushes);
    __syncthreads();
    if (threadIdx.x == 0) {
        a.found[n] = s_found;
        auto *st = reinterpret_cast<unsigned long long *>(a.stats);
        if (s_pairs) atomicAdd(st + NND_STAT_PAIRS, s_pairs);
        if (s_push) atomicAdd(st + NND_STAT_PUSHES, s_push);
        atomicAdd(st + NND_STAT_LISTS, (unsigned long long)(nA + nB));
    }
}

// ------------------------------------------------------------------ merge
// Wave-level bitonic sort of v[0..n), n a power of two.
__device__ void wave_bitonic(uint64_t *v, unsigned n) {
    unsigned const lane = lane_id();
    for (unsigned k = 2; k <= n; k <<= 1) {
        for (unsigned j = k >> 1; j > 0; j >>= 1) {
            for (unsigned i = lane; i < n; i += KG_WAVE) {
                unsigned ixj = i ^ j;
                if (ixj > i) {
                    uint64_t x = v[i], y = v[ixj];
                    bool up = (i & k) == 0;
                    if ((x > y) == up) {
                        v[i] = y;
                        v[ixj] = x;
                    }
                }
            }
            wave_sync();
        }
    }
}

// number of entries of sorted v[0..n) that are < key (strict) or <= key
__device__ __forceinline__ unsigned lower_bound64(uint64_t const *v, unsigned n, uint64_t key, bool inclusive) {
    unsigned lo = 0, hi = n;
    while (lo < hi) {
        unsigned mid = (lo + hi) >> 1;
        bool left = inclusive ? (v[mid] <= key) : (v[mid] < key);
        if (left)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

__global__ void __launch_bounds__(256) merge_kernel(NndArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    unsigned const w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    unsigned const n = blockIdx.x * 4 + w;
    size_t const per_wave = size_t(a.Cp2 + a.L + a.L + a.Cp2) * 8;
    uint64_t *cb = reinterpret_cast<uint64_t *>(smem + w * per_wave);  // candidates, then the new pool
    uint64_t *pk = cb + a.Cp2;                                          // old pool
    uint64_t *mg = pk + a.L;                                            // merged sequence
    if (n >= a.N) return;
    unsigned const c = a.cnt[n];
    unsigned const m = min(a.ccnt[n], a.C);
    uint64_t *pool = a.pool + size_t(n) * a.L;
    unsigned newc = c;
    if (m > 0) {
        unsigned p2 = 1;
        while (p2 < m) p2 <<= 1;
        uint64_t *cand = a.cand + size_t(n) * a.C;
        for (unsigned i = lane; i < p2; i += KG_WAVE) {
            cb[i] = (i < m) ? cand[i] : ~0ull;
            if (i < m) cand[i] = ~0ull;  // ready for the next join
        }
        for (unsigned i = lane; i < c; i += KG_WAVE) pk[i] = pool[i];
        wave_sync();
        wave_bitonic(cb, p2);
        // stable merge: a pool entry goes before equal candidates
        for (unsigned i = lane; i < c; i += KG_WAVE) mg[i + lower_bound64(cb, m, pk[i], false)] = pk[i];
        for (unsigned j = lane; j < m; j += KG_WAVE) mg[j + lower_bound64(pk, c, cb[j], true)] = cb[j];
        wave_sync();
        // drop repeated (dist, id), keep the first L (written to HBM and to cb)
        unsigned const tot = c + m;
        unsigned out = 0;
        for (unsigned b = 0; b < tot && out < a.L; b += KG_WAVE) {
            unsigned i = b + lane;
            uint64_t k = (i < tot) ? mg[i] : ~0ull;
            bool keep = (i < tot) && (i == 0 || (mg[i - 1] >> 1) != (k >> 1));
            uint64_t bal = ballot(keep);
            unsigned pos = out + popc64(bal & ((1ull << lane) - 1));
            if (keep && pos < a.L) {
                pool[pos] = k;
                cb[pos] = k;
            }
            out += popc64(bal);
        }
        newc = min(out, a.L);
        if (lane == 0) {
            a.cnt[n] = newc;
            a.ccnt[n] = 0;
        }
    } else {
        for (unsigned i = lane; i < c; i += KG_WAVE) cb[i] = pool[i];
    }
    wave_sync();
    // radius / K-th distance snapshots and the delta statistic (kgraph.cpp:148-156)
    unsigned const kk = min(a.K, newc);
    unsigned flags = 0;
    for (unsigned b = 0; b < kk; b += KG_WAVE) {
        unsigned i = b + lane;
        flags += popc64(ballot(i < kk && key_flag(cb[i])));
    }
    if (lane == 0) {
        a.radius[n] = (newc >= a.L) ? key_dist(cb[a.L - 1]) : __builtin_huge_valf();
        a.kth[n] = (newc >= a.K) ? key_dist(cb[a.K - 1]) : __builtin_huge_valf();
        atomicAdd(reinterpret_cast<unsigned long long *>(a.stats + NND_STAT_DELTA), (unsigned long long)flags);
        atomicAdd(reinterpret_cast<unsigned long long *>(a.stats + NND_STAT_M), (unsigned long long)a.Mv[n]);
    }
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
                push_rev(a.rnew, a.crnew, a.R, id, n, a.iter);
            else
                push_rev(a.rold, a.crold, a.R, id, n, a.iter);
        }
        nn += popc64(bn);
        no += popc64(bo);
    }
    if (lane == 0) {
        a.nnew[n] = nn;
        a.nold[n] = no;
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

}  // namespace

static int ch_for(unsigned D) {
    unsigned nch = (D & ~15u) >> 2;
    unsigned per = (nch + 3) / 4;
    if (per <= 1) return 1;
    if (per <= 2) return 2;
    if (per <= 4) return 4;
    if (per <= 6) return 6;
    if (per <= 8) return 8;
    return 0;
}

size_t nnd_join_lds(NndArgs const &a) { return size_t(a.Acap + a.Bcap + 3 * NND_ACH) * 4 + size_t(NND_ACH) * a.Dp * 4; }
size_t nnd_merge_lds(NndArgs const &a) { return size_t(4) * (a.Cp2 + a.L + a.L + a.Cp2) * 8; }

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
    dim3 g(a.N), b(256);
    size_t lds = nnd_join_lds(a);
    switch (ch_for(a.D)) {
        case 1: hipLaunchKernelGGL((join_kernel<METRIC, 1>), g, b, lds, s, a); break;
        case 2: hipLaunchKernelGGL((join_kernel<METRIC, 2>), g, b, lds, s, a); break;
        case 4: hipLaunchKernelGGL((join_kernel<METRIC, 4>), g, b, lds, s, a); break;
        case 6: hipLaunchKernelGGL((join_kernel<METRIC, 6>), g, b, lds, s, a); break;
        case 8: hipLaunchKernelGGL((join_kernel<METRIC, 8>), g, b, lds, s, a); break;
        default: hipLaunchKernelGGL((join_kernel<METRIC, 0>), g, b, lds, s, a); break;
    }
}

hipError_t nnd_launch_join(int metric, NndArgs const &a, hipStream_t s) {
    if (metric == METRIC_L2SQR)
        launch_join_m<METRIC_L2SQR>(a, s);
    else if (metric == METRIC_EUCLIDEAN)
        launch_join_m<METRIC_EUCLIDEAN>(a, s);
    else
        launch_join_m<METRIC_ANGULAR>(a, s);
    return hipGetLastError();
}

hipError_t nnd_launch_merge(NndArgs const &a, hipStream_t s) {
    hipLaunchKernelGGL(merge_kernel, dim3((a.N + 3) / 4), dim3(256), nnd_merge_lds(a), s, a);
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

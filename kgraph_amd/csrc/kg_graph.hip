// kg_graph.hip -- whole-index kernels on the CSR graph: exclusive scan,
// reverse-edge completion (KGraphImpl::reverse, kgraph.cpp:667-702) and prune
// levels 1 and 2 (kgraph.cpp:584-657).
//
// Reverse, rev = -1 keeps each node's first M[i] edges, rev > 0 its first
// min(rev, deg) edges; every kept edge i -> e is stored at i and, with the same
// distance, as e -> i.  Each node's list is then sorted by (dist, id) and
// repeated ids are dropped, and M[i] becomes the list length.  On the GPU:
// count in-degrees, scan, scatter both directions as u64 keys
// (orderable dist << 32 | id), sort every segment (a wave-level LDS bitonic
// sort for segments up to 1024 entries, a block-level sort in global memory
// for longer ones), drop adjacent repeated ids, scan again and compact.
#include "kg_device.hpp"
#include "kg_nnd.hpp"

namespace {

using namespace kgd;

constexpr unsigned SB = 1024;  // scan block

__device__ unsigned block_excl_scan(unsigned v, unsigned *sh, unsigned *total) {
    unsigned const t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (unsigned o = 1; o < blockDim.x; o <<= 1) {
        unsigned x = (t >= o) ? sh[t - o] : 0u;
        __syncthreads();
        sh[t] += x;
        __syncthreads();
    }
    unsigned incl = sh[t];
    *total = sh[blockDim.x - 1];
    __syncthreads();
    return incl - v;
}

__global__ void __launch_bounds__(SB) scan_reduce(uint32_t const *in, unsigned n, uint32_t *bsum) {
    __shared__ unsigned sh[SB];
    unsigned i = blockIdx.x * SB + threadIdx.x;
    unsigned tot;
    block_excl_scan(i < n ? in[i] : 0u, sh, &tot);
    if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

__global__ void __launch_bounds__(SB) scan_blocks(uint32_t *bsum, unsigned nb, uint32_t *total_out) {
    __shared__ unsigned sh[SB];
    unsigned per = (nb + SB - 1) / SB;
    unsigned b0 = threadIdx.x * per;
    unsigned loc = 0;
    for (unsigned j = 0; j < per; ++j)
        if (b0 + j < nb) loc += bsum[b0 + j];
    unsigned tot;
    unsigned base = block_excl_scan(loc, sh, &tot);
    for (unsigned j = 0; j < per; ++j) {
        if (b0 + j < nb) {
            unsigned v = bsum[b0 + j];
            bsum[b0 + j] = base;
            base += v;
        }
    }
    if (threadIdx.x == 0) *total_out = tot;
}

__global__ void __launch_bounds__(SB) scan_apply(uint32_t const *in, unsigned n, uint32_t const *bsum, uint32_t *out) {
    __shared__ unsigned sh[SB];
    unsigned i = blockIdx.x * SB + threadIdx.x;
    unsigned tot;
    unsigned ex = block_excl_scan(i < n ? in[i] : 0u, sh, &tot);
    if (i < n) out[i] = ex + bsum[blockIdx.x];
}

// ---------------------------------------------------------------- reverse
__device__ __forceinline__ unsigned kept(GraphDev const &g, int rev, unsigned i) {
    unsigned deg = g.off[i + 1] - g.off[i];
    unsigned k = (rev > 0) ? unsigned(rev) : g.M[i];
    return k < deg ? k : deg;
}

__global__ void reverse_count(GraphDev g, int rev, uint32_t *deg2) {
    unsigned const w = (blockIdx.x * blockDim.x + threadIdx.x) >> 6, lane = threadIdx.x & 63;
    if (w >= g.N) return;
    unsigned const k = kept(g, rev, w);
    if (lane == 0) atomicAdd(&deg2[w], k);
    for (unsigned j = lane; j < k; j += KG_WAVE) atomicAdd(&deg2[g.ids[g.off[w] + j]], 1u);
}

__global__ void reverse_scatter(GraphDev g, int rev, uint32_t const *off2, uint32_t *cursor, uint64_t *keys) {
    unsigned const w = (blockIdx.x * blockDim.x + threadIdx.x) >> 6, lane = threadIdx.x & 63;
    if (w >= g.N) return;
    unsigned const k = kept(g, rev, w);
    for (unsigned j = lane; j < k; j += KG_WAVE) {
        uint32_t e = g.ids[g.off[w] + j];
        uint32_t od = ord_f32(g.dists[g.off[w] + j]);
        keys[off2[w] + j] = (uint64_t(od) << 32) | e;
        unsigned pos = atomicAdd(&cursor[e], 1u);
        keys[pos] = (uint64_t(od) << 32) | w;
    }
}

__global__ void init_cursor(unsigned N, GraphDev g, int rev, uint32_t const *off2, uint32_t *cursor) {
    unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < N) cursor[i] = off2[i] + kept(g, rev, i);
}

constexpr unsigned SMALL = 1024;

__device__ void wave_bitonic64(uint64_t *v, unsigned n) {
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

// Sort + unique for segments of <= SMALL entries, one wave per node.
__global__ void __launch_bounds__(256) segsort_small(unsigned N, uint32_t const *off2, uint64_t *keys, uint32_t *cnt2) {
    __shared__ uint64_t sh[4][SMALL];
    unsigned const w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    unsigned const i = blockIdx.x * 4 + w;
    if (i >= N) return;
    unsigned const b = off2[i], n = off2[i + 1] - b;
    if (n > SMALL) return;
    unsigned p2 = 1;
    while (p2 < n) p2 <<= 1;
    for (unsigned j = lane; j < p2; j += KG_WAVE) sh[w][j] = (j < n) ? keys[b + j] : ~0ull;
    wave_sync();
    wave_bitonic64(sh[w], p2);
    unsigned out = 0;
    for (unsigned c = 0; c < n; c += KG_WAVE) {
        unsigned j = c + lane;
        bool keep = j < n && (j == 0 || uint32_t(sh[w][j - 1]) != uint32_t(sh[w][j]));
        uint64_t bal = ballot(keep);
        if (keep) keys[b + out + popc64(bal & ((1ull << lane) - 1))] = sh[w][j];
        out += popc64(bal);
    }
    if (lane == 0) cnt2[i] = out;
}

// Longer segments: one block each, bitonic sort in a pow2 scratch region.
__global__ void __launch_bounds__(1024) segsort_large(uint32_t const *off2, uint64_t *keys, uint32_t *cnt2,
                                                    uint32_t const *large, uint64_t *scratch,
                                                    uint64_t const *scratch_off) {
    unsigned const i = large[blockIdx.x];
    unsigned const b = off2[i], n = off2[i + 1] - b;
    uint64_t *v = scratch + scratch_off[blockIdx.x];
    unsigned p2 = 1;
    while (p2 < n) p2 <<= 1;
    for (unsigned j = threadIdx.x; j < p2; j += blockDim.x) v[j] = (j < n) ? keys[b + j] : ~0ull;
    __syncthreads();
    for (unsigned k = 2; k <= p2; k <<= 1) {
        for (unsigned jj = k >> 1; jj > 0; jj >>= 1) {
            for (unsigned j = threadIdx.x; j < p2; j += blockDim.x) {
                unsigned ixj = j ^ jj;
                if (ixj > j) {
                    uint64_t x = v[j], y = v[ixj];
                    bool up = (j & k) == 0;
                    if ((x > y) == up) {
                        v[j] = y;
                        v[ixj] = x;
                    }
                }
            }
            __threadfence_block();
            __syncthreads();
        }
    }
    if (threadIdx.x == 0) {
        unsigned out = 0;
        for (unsigned j = 0; j < n; ++j)
            if (j == 0 || uint32_t(v[j - 1]) != uint32_t(v[j])) keys[b + out++] = v[j];
        cnt2[i] = out;
    }
}

__global__ void compact_keys(unsigned N, uint32_t const *off2, uint64_t const *keys, uint32_t const *off3,
                             uint32_t *ids, float *dists, uint32_t *M) {
    unsigned const w = (blockIdx.x * blockDim.x + threadIdx.x) >> 6, lane = threadIdx.x & 63;
    if (w >= N) return;
    unsigned const n = off3[w + 1] - off3[w];
    for (unsigned j = lane; j < n; j += KG_WAVE) {
        uint64_t k = keys[off2[w] + j];
        ids[off3[w] + j] = uint32_t(k);
        dists[off3[w] + j] = unord_f32(uint32_t(k >> 32));
    }
    if (lane == 0) M[w] = n;
}

__global__ void prune1_count(GraphDev g, uint32_t *cnt) {
    unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= g.N) return;
    unsigned deg = g.off[i + 1] - g.off[i];
    cnt[i] = deg > g.M[i] ? g.M[i] : deg;
}

__global__ void prune1_copy(GraphDev g, uint32_t const *off2, uint32_t *ids2, float *dists2) {
    unsigned const w = (blockIdx.x * blockDim.x + threadIdx.x) >> 6, lane = threadIdx.x & 63;
    if (w >= g.N) return;
    unsigned const n = off2[w + 1] - off2[w];
    for (unsigned j = lane; j < n; j += KG_WAVE) {
        ids2[off2[w] + j] = g.ids[g.off[w] + j];
        if (dists2) dists2[off2[w] + j] = g.dists[g.off[w] + j];
    }
}

// Prune level 2 (kgraph.cpp:593-657), one level l per pass.  Position l of
// node i (x = its l-th neighbour) survives iff
//   (1) no node T kept by i at an earlier level has already kept x, and
//   (2) no node T < i kept by i at an earlier level keeps the same x at level l.
// This is what the reference's flags compute: a flag on (i, x) is set exactly
// when some T kept by i before the visit of (i, l) had kept x before it --
// forward marking when T kept x first, reverse marking through rev[T] when i
// kept T first -- and within a level nodes are visited in id order, which is
// where (2) comes from.  prune2_screen evaluates (1) for every active node;
// prune2_resolve settles (2), whose dependencies point to smaller ids only,
// in rounds (state 1 undecided, 2 pruned, 3 kept), and appends each kept x to
// the node's kept list kl / kd at off[i] + nl[i].

// 16 lanes per node: lanes stride over each kept T's own kept list.
__global__ void prune2_screen(Prune2Dev p, unsigned l, unsigned nact) {
    unsigned const g = (blockIdx.x * blockDim.x + threadIdx.x) >> 4, sl = threadIdx.x & 15;
    if (g >= nact) return;  // a whole 16-lane group leaves together
    unsigned const i = p.order[g];
    uint32_t const base = p.off[i], x = p.ids[base + l], n = p.nl[i];
    uint64_t const gmask = 0xffffull << (threadIdx.x & 48);
    bool hit = false;
    for (uint32_t t = 0; t < n && !hit; ++t) {
        uint32_t const T = p.kl[base + t];
        uint32_t const bT = p.off[T], nT = p.nl[T];
        bool f = false;
        for (uint32_t j = sl; j < nT; j += 16) f |= p.kl[bT + j] == x;
        hit = (__ballot(f) & gmask) != 0;
    }
    if (sl == 0) p.st[i] = hit ? 2u : 1u;
}

__global__ void prune2_resolve(Prune2Dev p, unsigned l, unsigned nact, uint32_t *undecided) {
    unsigned const g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= nact) return;
    unsigned const i = p.order[g];
    if (__hip_atomic_load(p.st + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 1u) return;
    uint32_t const base = p.off[i], x = p.ids[base + l], n = p.nl[i];
    bool open = false, pruned = false;
    for (uint32_t t = 0; t < n; ++t) {
        uint32_t const T = p.kl[base + t];
        if (T >= i || l >= p.Meff[T] || p.ids[p.off[T] + l] != x) continue;
        uint32_t const s = __hip_atomic_load(p.st + T, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (s == 3u) {
            pruned = true;
            break;
        }
        open |= s == 1u;
    }
    if (pruned) {
        __hip_atomic_store(p.st + i, 2u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else if (!open) {
        p.kl[base + n] = x;
        p.kd[base + n] = p.dists ? p.dists[base + l] : 0.f;
        p.nl[i] = n + 1;
        __hip_atomic_store(p.st + i, 3u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else {
        atomicAdd(undecided, 1u);
    }
}

__global__ void prune2_copy(Prune2Dev p, uint32_t const *off2, uint32_t *ids2, float *dists2) {
    unsigned const w = (blockIdx.x * blockDim.x + threadIdx.x) >> 6, lane = threadIdx.x & 63;
    if (w >= p.N) return;
    unsigned const n = off2[w + 1] - off2[w];
    for (unsigned j = lane; j < n; j += KG_WAVE) {
        ids2[off2[w] + j] = p.kl[p.off[w] + j];
        dists2[off2[w] + j] = p.kd[p.off[w] + j];
    }
}

}  // namespace

hipError_t kg_prune2_screen(Prune2Dev const &p, unsigned l, unsigned nact, hipStream_t s) {
    hipLaunchKernelGGL(prune2_screen, dim3((nact + 15) / 16), dim3(256), 0, s, p, l, nact);
    return hipGetLastError();
}

hipError_t kg_prune2_resolve(Prune2Dev const &p, unsigned l, unsigned nact, uint32_t *undecided, hipStream_t s) {
    hipLaunchKernelGGL(prune2_resolve, dim3((nact + 255) / 256), dim3(256), 0, s, p, l, nact, undecided);
    return hipGetLastError();
}

hipError_t kg_prune2_copy(Prune2Dev const &p, uint32_t const *off2, uint32_t *ids2, float *dists2, hipStream_t s) {
    hipLaunchKernelGGL(prune2_copy, dim3((p.N + 3) / 4), dim3(256), 0, s, p, off2, ids2, dists2);
    return hipGetLastError();
}

hipError_t kg_scan_u32(uint32_t const *in, uint32_t *out, unsigned n, uint32_t *tmp, hipStream_t s) {
    unsigned nb = (n + SB - 1) / SB;
    if (nb == 0) nb = 1;
    hipLaunchKernelGGL(scan_reduce, dim3(nb), dim3(SB), 0, s, in, n, tmp);
    hipLaunchKernelGGL(scan_blocks, dim3(1), dim3(SB), 0, s, tmp, nb, out + n);
    hipLaunchKernelGGL(scan_apply, dim3(nb), dim3(SB), 0, s, in, n, tmp, out);
    return hipGetLastError();
}

hipError_t kg_reverse_count(GraphDev g, int rev, uint32_t *deg2, hipStream_t s) {
    hipLaunchKernelGGL(reverse_count, dim3((g.N + 3) / 4), dim3(256), 0, s, g, rev, deg2);
    return hipGetLastError();
}

hipError_t kg_reverse_scatter(GraphDev g, int rev, uint32_t const *off2, uint32_t *cursor, uint64_t *keys,
                              hipStream_t s) {
    hipLaunchKernelGGL(init_cursor, dim3((g.N + 255) / 256), dim3(256), 0, s, g.N, g, rev, off2, cursor);
    hipLaunchKernelGGL(reverse_scatter, dim3((g.N + 3) / 4), dim3(256), 0, s, g, rev, off2, cursor, keys);
    return hipGetLastError();
}

hipError_t kg_segsort_unique(unsigned N, uint32_t const *off2, uint64_t *keys, uint32_t *cnt2,
                             uint32_t const *large, unsigned nlarge, uint64_t *scratch, uint64_t const *scratch_off,
                             hipStream_t s) {
    hipLaunchKernelGGL(segsort_small, dim3((N + 3) / 4), dim3(256), 0, s, N, off2, keys, cnt2);
    if (nlarge)
        hipLaunchKernelGGL(segsort_large, dim3(nlarge), dim3(1024), 0, s, off2, keys, cnt2, large, scratch,
                           scratch_off);
    return hipGetLastError();
}

hipError_t kg_compact_keys(unsigned N, uint32_t const *off2, uint64_t const *keys, uint32_t const *off3,
                           uint32_t *ids, float *dists, uint32_t *M, hipStream_t s) {
    hipLaunchKernelGGL(compact_keys, dim3((N + 3) / 4), dim3(256), 0, s, N, off2, keys, off3, ids, dists, M);
    return hipGetLastError();
}

hipError_t kg_prune1_count(GraphDev g, uint32_t *cnt, hipStream_t s) {
    hipLaunchKernelGGL(prune1_count, dim3((g.N + 255) / 256), dim3(256), 0, s, g, cnt);
    return hipGetLastError();
}

hipError_t kg_prune1_copy(GraphDev g, uint32_t const *off2, uint32_t *ids2, float *dists2, hipStream_t s) {
    hipLaunchKernelGGL(prune1_copy, dim3((g.N + 3) / 4), dim3(256), 0, s, g, off2, ids2, dists2);
    return hipGetLastError();
}

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
// overflows).
//
// Inside a batch the wave reads the S neighbour ids (one per lane), drops
// in-batch duplicates, test-and-sets the visited bits with atomicOr, gathers
// the surviving rows with 4 lanes per row (16 rows per pass), then applies the
// inserts one by one in neighbour order.  That equals the reference's
// per-neighbour loop because neither the visited state nor the distances
// depend on the inserts.
#include "kg_device.hpp"
#include "kg_internal.hpp"

namespace {

using namespace kgd;

constexpr unsigned SEARCH_MAX_WPB = 4;  // waves per block the kernel is compiled for
constexpr uint32_t FLAG = 0x80000000u;
constexpr uint32_t IDMASK = 0x7fffffffu;

struct WaveLDS {
    uint32_t *pid;
    float *pdist;
    uint32_t *pmeta;
    uint32_t *pbase;
    uint64_t *tkey;
    uint32_t *rid;
    float *rdist;
    float *q;
    uint32_t *fid;
    float *fdist;
};

__host__ __device__ inline size_t align16(size_t x) { return (x + 15) & ~size_t(15); }

struct LdsLayout {
    size_t pid, pdist, pmeta, pbase, tkey, rid, rdist, q, fid, fdist, total;
    __host__ __device__ LdsLayout(unsigned cap, unsigned rcap, unsigned Dp) {
        size_t o = 0;
        pid = o; o = align16(o + 4ull * cap);
        pdist = o; o = align16(o + 4ull * cap);
        pmeta = o; o = align16(o + 4ull * cap);
        pbase = o; o = align16(o + 4ull * cap);
        tkey = o; o = align16(o + 8ull * cap);
        rid = o; o = align16(o + 4ull * rcap);
        rdist = o; o = align16(o + 4ull * rcap);
        q = o; o = align16(o + 4ull * Dp);
        fid = o; o = align16(o + 4ull * KG_WAVE);
        fdist = o; o = align16(o + 4ull * KG_WAVE);
        total = o;
    }
};

// actual_M (kgraph.cpp:284-286) truncated to NeighborX::M's uint16_t
// (kgraph.cpp:79, 474, 508); also returns the node's CSR base offset.
__device__ __forceinline__ uint32_t actual_M(kg_search_args const &a, uint32_t id, uint32_t *base) {
    uint32_t b = a.goff[id];
    uint32_t deg = a.goff[id + 1] - b;
    uint32_t m = a.gM[id];
    if (m < a.Mparam) m = a.Mparam;
    if (m > deg) m = deg;
    *base = b;
    return m & 0xffffu;
}

// Distances of ids[0..cnt) (LDS, flag bit ignored) -> out[0..cnt) (LDS).
template <int METRIC>
__device__ __forceinline__ void batch_dist(kg_search_args const &a, uint32_t const *ids, float *out, unsigned cnt,
                                           float const *q, float qn) {
    unsigned const lane = lane_id();
    unsigned const grp = lane >> 2, t = lane & 3;
    for (unsigned b0 = 0; b0 < cnt; b0 += 16) {
        unsigned c = b0 + grp;
        if (b0 + grp < cnt) {  // whole quads are active together
            uint32_t id = ids[c] & IDMASK;
            float const *row = a.X + size_t(id) * a.Dp;
            float acc = quad_metric<METRIC>(row, q, a.D, t);
            float na = (METRIC == METRIC_L2SQR) ? 0.f : a.norms[id];
            float d = finish_metric<METRIC>(acc, na, qn);
            if (t == 0) out[c] = d;
        }
    }
    wave_sync();
}

// Number of entries of the sorted pool[0..L) whose dist <= nd.
__device__ __forceinline__ unsigned insert_pos(float const *pdist, unsigned L, float nd) {
    unsigned const lane = lane_id();
    unsigned r = 0;
    for (unsigned c = 0; c < L; c += KG_WAVE) {
        unsigned i = c + lane;
        bool le = (i < L) && (pdist[i] <= nd);
        unsigned n = popc64(ballot(le));
        r += n;
        if (n < KG_WAVE) break;
    }
    return r;
}

// pool[r..L) -> pool[r+1..L+1); slot L is the pool's scratch slot.
__device__ __forceinline__ void pool_shift(WaveLDS &w, unsigned r, unsigned L) {
    if (r >= L) return;
    unsigned const lane = lane_id();
    int const lo = int(r & ~(KG_WAVE - 1u));
    for (int c = int((L - 1) & ~(KG_WAVE - 1u)); c >= lo; c -= KG_WAVE) {
        unsigned i = unsigned(c) + lane;
        bool mv = (i >= r) && (i < L);
        uint32_t id = 0, meta = 0, bs = 0;
        float d = 0.f;
        if (mv) {
            id = w.pid[i];
            d = w.pdist[i];
            meta = w.pmeta[i];
            bs = w.pbase[i];
        }
        wave_sync();
        if (mv) {
            w.pid[i + 1] = id;
            w.pdist[i + 1] = d;
            w.pmeta[i + 1] = meta;
            w.pbase[i + 1] = bs;
        }
        wave_sync();
    }
}

__device__ __forceinline__ uint32_t load_u32_l2(uint32_t const *p) {
    return __hip_atomic_load(const_cast<uint32_t *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <int METRIC>
__global__ void __launch_bounds__(SEARCH_MAX_WPB * KG_WAVE) search_kernel(kg_search_args a) {
    extern __shared__ __align__(16) unsigned char smem[];
    unsigned const lane = lane_id();
    unsigned const wib = threadIdx.x / KG_WAVE;
    unsigned const slot = blockIdx.x * (blockDim.x / KG_WAVE) + wib;
    unsigned const cap = a.pool_cap;
    unsigned const rcap = a.res_cap;
    LdsLayout lay(cap, rcap, a.Dp);
    unsigned char *base = smem + size_t(wib) * a.lds_per_wave;
    WaveLDS w;
    w.pid = reinterpret_cast<uint32_t *>(base + lay.pid);
    w.pdist = reinterpret_cast<float *>(base + lay.pdist);
    w.pmeta = reinterpret_cast<uint32_t *>(base + lay.pmeta);
    w.pbase = reinterpret_cast<uint32_t *>(base + lay.pbase);
    w.tkey = reinterpret_cast<uint64_t *>(base + lay.tkey);
    w.rid = reinterpret_cast<uint32_t *>(base + lay.rid);
    w.rdist = reinterpret_cast<float *>(base + lay.rdist);
    w.q = reinterpret_cast<float *>(base + lay.q);
    w.fid = reinterpret_cast<uint32_t *>(base + lay.fid);
    w.fdist = reinterpret_cast<float *>(base + lay.fdist);

    uint32_t *vis = a.visited + size_t(slot) * a.vis_words;
    uint32_t *touched = a.touched + size_t(slot) * a.touch_cap;
    uint64_t const lt_mask = (1ull << lane) - 1;

    for (;;) {
        unsigned qi = 0;
        if (lane == 0) qi = atomicAdd(a.qcounter, 1u);
        qi = uniform(__shfl(qi, 0));
        if (qi >= a.nq) break;
        uint64_t const t_start = __builtin_amdgcn_s_memrealtime();  // 100 MHz constant clock
        unsigned const K = a.Ks ? a.Ks[qi] : a.K0;
        unsigned const P = a.Ps ? a.Ps[qi] : a.P0;
        if (P >= a.N) continue;  // served by the brute-force launch (kgraph.cpp:427-433)
        unsigned const qcap = K + P + 1;

        float const *qsrc = a.Q + size_t(qi) * a.qstride;
        for (unsigned i = lane; i < a.Dp; i += KG_WAVE) w.q[i] = (i < a.D) ? qsrc[i] : 0.f;
        float const qn = (METRIC == METRIC_L2SQR) ? 0.f : a.qnorms[qi];
        wave_sync();

        unsigned ntouch = 0, n_comps = 0, n_scan = 0, RL = 0;
        for (unsigned trial = 0; trial < a.T; ++trial) {
            // ---- starting points (kgraph.cpp:450-467)
            unsigned L = 0;
            if (a.init == 0) {
                uint32_t const *seeds = a.seeds + a.seed_off[a.Ps ? a.pidx[qi] : 0] + size_t(trial) * P;
                for (unsigned c = 0; c < P; c += KG_WAVE) {
                    unsigned i = c + lane;
                    uint32_t s = 0;
                    bool keep = false;
                    if (i < P) {
                        s = seeds[i];
                        keep = true;
                        if (trial > 0) keep = !(load_u32_l2(&vis[s >> 5]) & (1u << (s & 31)));
                    }
                    uint64_t b = ballot(keep);
                    if (keep) w.pid[L + popc64(b & lt_mask)] = s;
                    L += popc64(b);
                }
            } else {
                uint32_t const *ini = a.init_ids + size_t(qi) * a.ostride;
                for (unsigned i = lane; i < a.init; i += KG_WAVE) w.pid[i] = ini[i];
                L = a.init;
            }
            wave_sync();
            // ---- mark, score and sort the starting points (kgraph.cpp:468-476)
            for (unsigned i = lane; i < L; i += KG_WAVE) {
                uint32_t s = w.pid[i];
                atomicOr(&vis[s >> 5], 1u << (s & 31));
                if (ntouch + i < a.touch_cap) touched[ntouch + i] = s;
            }
            ntouch += L;
            batch_dist<METRIC>(a, w.pid, w.pdist, L, w.q, qn);
            for (unsigned i = lane; i < L; i += KG_WAVE)
                w.tkey[i] = (uint64_t(ord_f32(w.pdist[i])) << 32) | w.pid[i];
            wave_sync();
            for (unsigned c = 0; c < L; c += KG_WAVE) {
                unsigned i = c + lane;
                uint64_t key = (i < L) ? w.tkey[i] : ~0ull;
                unsigned rank = 0;
                for (unsigned j = 0; j < L; ++j) {
                    uint64_t kj = w.tkey[j];
                    rank += (kj < key) || (kj == key && j < i);
                }
                if (i < L) {
                    w.pid[rank] = uint32_t(key) | FLAG;
                    w.pdist[rank] = unord_f32(uint32_t(key >> 32));
                }
            }
            wave_sync();
            for (unsigned i = lane; i < L; i += KG_WAVE) {
                uint32_t bs;
                uint32_t m = actual_M(a, w.pid[i] & IDMASK, &bs);
                w.pmeta[i] = m << 16;
                w.pbase[i] = bs;
            }
            wave_sync();

            // ---- expansion loop (kgraph.cpp:478-514)
            unsigned k = 0;
            for (;;) {
                unsigned kk = L;
                for (unsigned c = k; c < L; c += KG_WAVE) {
                    unsigned i = c + lane;
                    uint64_t b = ballot(i < L && (w.pid[i] & FLAG));
                    if (b) {
                        kk = c + ctz64(b);
                        break;
                    }
                }
                if (kk >= L) break;
                k = kk;
                uint32_t const eid = w.pid[k];
                uint32_t const meta = w.pmeta[k];
                uint32_t const ebase = w.pbase[k];
                unsigned const beginM = meta & 0xffffu, eM = meta >> 16;
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
                for (unsigned mc = beginM; mc < endM; mc += KG_WAVE) {
                    unsigned const cnt = min(endM - mc, unsigned(KG_WAVE));
                    uint32_t id = (lane < cnt) ? a.gids[ebase + mc + lane] : 0xffffffffu;
                    bool dup = false;
                    for (unsigned j = 0; j < cnt; ++j) {
                        uint32_t v = __shfl(id, j);
                        dup |= (j < lane) && (v == id);
                    }
                    bool fresh = false;
                    if (lane < cnt && !dup) {
                        uint32_t bit = 1u << (id & 31);
                        fresh = !(atomicOr(&vis[id >> 5], bit) & bit);
                    }
                    uint64_t b = ballot(fresh);
                    unsigned const nf = popc64(b);
                    unsigned pos = popc64(b & lt_mask);
                    if (fresh) {
                        w.fid[pos] = id;
                        if (ntouch + pos < a.touch_cap) touched[ntouch + pos] = id;
                    }
                    ntouch += nf;
                    wave_sync();
                    if (nf == 0) continue;
                    // lane j fetches candidate j's M and CSR base while the rows load
                    uint32_t cm = 0, cb = 0;
                    if (lane < nf) cm = actual_M(a, w.fid[lane], &cb);
                    batch_dist<METRIC>(a, w.fid, w.fdist, nf, w.q, qn);
                    for (unsigned j = 0; j < nf; ++j) {
                        ++n_comps;
                        uint32_t const cid = w.fid[j];
                        float const cd = w.fdist[j];
                        uint32_t const jm = __shfl(cm, j), jb = __shfl(cb, j);
                        bool const full = (L + 1 >= qcap);
                        if (full && w.pdist[L - 1] <= cd) continue;  // lands in the scratch slot
                        unsigned const r = insert_pos(w.pdist, L, cd);
                        wave_sync();
                        pool_shift(w, r, L);
                        if (lane == 0) {
                            w.pid[r] = cid | FLAG;
                            w.pdist[r] = cd;
                            w.pmeta[r] = jm << 16;
                            w.pbase[r] = jb;
                        }
                        wave_sync();
                        if (!full) ++L;
                        if (r < L && r < k) k = r;
                    }
                }
            }
            // ---- merge this trial into the results (kgraph.cpp:515-529)
            if (L > K) L = K;
            if (a.T == 1) {
                RL = L;
            } else if (trial == 0) {
                for (unsigned i = lane; i < L; i += KG_WAVE) {
                    w.rid[i] = w.pid[i] & IDMASK;
                    w.rdist[i] = w.pdist[i];
                }
                RL = L;
                wave_sync();
            } else {
                for (unsigned l = 0; l < L; ++l) {
                    float const cd = w.pdist[l];
                    uint32_t const cid = w.pid[l] & IDMASK;
                    unsigned const r = insert_pos(w.rdist, RL, cd);
                    wave_sync();
                    // shift results[r..RL) up by one (slot RL is the scratch)
                    for (int c = int(RL) - 1; c >= int(r); c -= KG_WAVE) {
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
        uint32_t const *oid = (a.T == 1) ? w.pid : w.rid;
        float const *odist = (a.T == 1) ? w.pdist : w.rdist;
        unsigned cut = RL;
        for (unsigned c = 0; c < RL; c += KG_WAVE) {
            unsigned i = c + lane;
            uint64_t b = ballot(i < RL && odist[i] > a.epsilon);
            if (b) {
                cut = c + ctz64(b);
                break;
            }
        }
        for (unsigned i = lane; i < cut; i += KG_WAVE) {
            a.out_ids[size_t(qi) * a.ostride + i] = oid[i] & IDMASK;
            if (a.out_dists) a.out_dists[size_t(qi) * a.ostride + i] = odist[i];
        }
        if (lane == 0) {
            a.counts[qi] = cut;
            if (a.comps) a.comps[qi] = n_comps;
            if (a.scans) a.scans[qi] = n_scan;
            if (a.lat_us) a.lat_us[qi] = float(__builtin_amdgcn_s_memrealtime() - t_start) * 0.01f;
        }
        // ---- clear the visited bits this query set
        if (ntouch <= a.touch_cap) {
            for (unsigned i = lane; i < ntouch; i += KG_WAVE) vis[load_u32_l2(&touched[i]) >> 5] = 0u;
        } else {
            for (unsigned i = lane; i < a.vis_words; i += KG_WAVE) vis[i] = 0u;
        }
        __threadfence();
        wave_sync();
    }
}

}  // namespace

size_t kg_search_lds_per_wave(unsigned pool_cap, unsigned res_cap, unsigned Dp) {
    return LdsLayout(pool_cap, res_cap, Dp).total;
}

unsigned kg_search_max_wpb() { return SEARCH_MAX_WPB; }

hipError_t kg_launch_search(int metric, kg_search_args const &a, unsigned blocks, unsigned waves_per_block,
                            hipStream_t s) {
    if (waves_per_block == 0 || waves_per_block > SEARCH_MAX_WPB || blocks == 0) return hipErrorInvalidValue;
    size_t lds = size_t(waves_per_block) * a.lds_per_wave;
    dim3 grid(blocks), block(waves_per_block * KG_WAVE);
    switch (metric) {
        case METRIC_L2SQR:
            hipLaunchKernelGGL(search_kernel<METRIC_L2SQR>, grid, block, lds, s, a);
            break;
        case METRIC_EUCLIDEAN:
            hipLaunchKernelGGL(search_kernel<METRIC_EUCLIDEAN>, grid, block, lds, s, a);
            break;
        default:
            hipLaunchKernelGGL(search_kernel<METRIC_ANGULAR>, grid, block, lds, s, a);
            break;
    }
    return hipGetLastError();
}

// kg_nnd.hpp -- state of one NN-descent build on the device (shared by
// kg_nndescent.hip and the host driver in kg_build.cpp).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#define NND_MAX_S 255  // largest sample size S
#define NND_MERGE_MAX 1024u  // pool + candidates sorted per node per merge

// Iteration statistics: per node in `part` (column k at part[k * N + n]),
// summed into `stats` after the merge.
enum { NND_STAT_PAIRS = 0, NND_STAT_DELTA = 1, NND_STAT_M = 2, NND_STAT_LISTS = 3, NND_STAT_PUSHES = 4,
       NND_STAT_RESCORE = 5, NND_NPART = 6, NND_NSTATS = 8 };

constexpr uint32_t NND_NOSRC = 0xffffffffu;

struct NndArgs {
    float const *X;      // rows, stride Dp floats
    uint16_t const *Xb;  // BF16 copy of the rows (MFMA operands), same stride
    float const *norms;  // per-row metric norms
    unsigned N, D, Dp;
    unsigned n0, n1;      // nodes the join kernels process: [n0, n1)
    unsigned L, K, S, R;  // IndexParams (kgraph.h:94-110)
    unsigned C, Cp2;      // candidate slots per node (and the next power of two)
    unsigned Lc;          // sample-list capacity per node, max(L, 2S)
    unsigned Acap, Bcap;  // join list capacities (multiples of 4)
    unsigned nnew_init;   // size of the first iteration's random "new" list
    unsigned long long seed;
    uint32_t *iter;  // device counter: the iteration whose update runs next (iter_begin_kernel advances it)
    // per-node state (Nhood, kgraph.cpp:709-721)
    uint64_t *pool;  // N x L keys: ord(dist) << 32 | id << 1 | new-flag
    uint32_t *cnt;   // valid pool entries (Nhood::L)
    uint32_t *Mv;    // Nhood::M
    uint32_t *found; // Nhood::found
    float4 *meta;    // per node: radius (worst pool distance once full, +inf before), K-th distance
                     // (pool[K-1] once K entries exist, +inf before), join norm, |x|
    float *radiusM;  // pool[M-1].dist
    uint32_t *samp;  // N x Lc: sampled new ids from the front, old ids from the back
    uint32_t *nnew, *nold;
    uint32_t *rnew, *rold;    // N x R reverse lists
    uint32_t *crnew, *crold;  // reverse-list push counters
    uint64_t *cand;           // N x C candidate slots, written once per claim
    uint32_t *csrc;           // N x C: per slot whose key has its flag bit set, the node whose join
                              // the candidate credits with `found` when it is new to the list
    uint64_t *ovf;            // N x O overflow slots (~0 = empty), hashed, atomicMin
    unsigned O;
    uint32_t *ccnt;           // candidate slot counters
    uint64_t *stats;          // NND_NSTATS counters
    uint32_t *part;           // NND_NPART x N per-node statistics
    // MFMA join (join_mfma_kernel)
    float2 const *sqn;  // per row: (|x|^2, |x|)
    unsigned mfma;      // 1: tile join on the matrix cores
    unsigned f32;       // with mfma: FP32 operands (v_mfma_f32_16x16x4_f32) instead of BF16
    unsigned exact;     // 1: BF16 tiles are exact for these rows (no re-score)
    // with exact: 8-bit operands (v_mfma_i32_16x16x64_i8) from X8, the rows'
    // 8-bit copy (row stride Dp bytes); meta's join norms are then those of
    // the shifted rows (nnd_launch_n8)
    unsigned i8;
    uint8_t const *X8;
    float scr_c, scr_g; // screen bound: relative dot error, relative norm error
    // learned pre-rank of the join lists (prerank_kernel): projected rows
    float const *Z;     // N x zdim projector outputs (NULL: no pre-rank)
    unsigned zdim;
    float keep;         // fraction of each node's list kept
};

hipError_t nnd_launch_init(int metric, NndArgs const &a, hipStream_t s);
hipError_t nnd_launch_join(int metric, NndArgs const &a, hipStream_t s);
// (|x|^2, |x|) per row into out; flags[0] |= 1 if any value is not an integer,
// flags[1] = max |x| as float bits, flags[2] / flags[3] = ord_f32 of the
// smallest / largest value (the caller sets flags[2] to all ones, the rest to 0).
hipError_t nnd_launch_sqn(NndArgs const &a, float2 *out, uint32_t *flags, hipStream_t s);
// Join norms of the 8-bit form: |x - 128|^2 over the D columns of each row of
// a.X8 (half of it when half != 0), exact integers, into a.meta[n].z (after
// nnd_launch_meta_init; only the 8-bit join reads them).
hipError_t nnd_launch_n8(NndArgs const &a, int half, hipStream_t s);
hipError_t nnd_launch_merge(NndArgs const &a, hipStream_t s);
// Prunes every node's join lists to the fraction a.keep nearest in the
// learned projection a.Z (ties by list position), in place.
hipError_t nnd_launch_prerank(NndArgs const &a, hipStream_t s);
// out[i] = |Z[pairs[2i]] - Z[pairs[2i+1]]|^2 (P floats per row of Z).
hipError_t nnd_launch_zpair(float const *Z, unsigned P, uint32_t const *pairs, unsigned n, float *out, hipStream_t s);
hipError_t nnd_launch_update(NndArgs const &a, hipStream_t s);
// Fills the join norms of a.meta (after the MFMA form's sqn, when used).
hipError_t nnd_launch_meta_init(int metric, NndArgs const &a, hipStream_t s);
// The join's BF16 MFMA distance and its error bound for n (a, b) id pairs
// (pairs[2p], pairs[2p+1]); needs X, Xb, norms, sqn, D, Dp and the bound
// constants of the args.
hipError_t nnd_launch_screen_pairs(int metric, NndArgs const &a, uint32_t const *pairs, unsigned n, float *out_d,
                                   float *out_e, hipStream_t s);
// Start of an iteration's join: zero the NND_NSTATS counters and, with
// advance, step the iteration counter (a kernel, so the whole iteration can
// be captured in a graph without memset nodes).
hipError_t nnd_launch_iter_begin(NndArgs const &a, int advance, hipStream_t s);
hipError_t nnd_launch_extract(NndArgs const &a, uint32_t const *off, uint32_t *ids, float *dists, uint32_t *M,
                              hipStream_t s);
hipError_t nnd_launch_gather_controls(NndArgs const &a, uint32_t const *ctl, unsigned nc, float *dout,
                                      uint32_t *cout, hipStream_t s);

// kg_graph.hip: whole-graph kernels on the CSR index.
// Exclusive scan of n counts into out[0..n] (out[n] = total); tmp must hold
// (n + 1023) / 1024 + 1 entries.
hipError_t kg_scan_u32(uint32_t const *in, uint32_t *out, unsigned n, uint32_t *tmp, hipStream_t s);
// Reverse-edge completion (kgraph.cpp:667-702) and prune level 1
// (kgraph.cpp:584-590).  See kg_graph.hip.
struct GraphDev {
    unsigned N;
    uint32_t *off;  // N+1
    uint32_t *ids;
    float *dists;
    uint32_t *M;
};
hipError_t kg_reverse_count(GraphDev g, int rev, uint32_t *deg2, hipStream_t s);
hipError_t kg_reverse_scatter(GraphDev g, int rev, uint32_t const *off2, uint32_t *cursor, uint64_t *keys,
                              hipStream_t s);
hipError_t kg_segsort_unique(unsigned N, uint32_t const *off2, uint64_t *keys, uint32_t *cnt2,
                             uint32_t const *large, unsigned nlarge, uint64_t *scratch, uint64_t const *scratch_off,
                             hipStream_t s);
hipError_t kg_compact_keys(unsigned N, uint32_t const *off2, uint64_t const *keys, uint32_t const *off3,
                           uint32_t *ids, float *dists, uint32_t *M, hipStream_t s);
hipError_t kg_prune1_count(GraphDev g, uint32_t *cnt, hipStream_t s);
hipError_t kg_prune1_copy(GraphDev g, uint32_t const *off2, uint32_t *ids2, float *dists2, hipStream_t s);
// Prune level 2 (kgraph.cpp:593-657): per-level screen / resolve passes over
// the nodes order[0..nact) (order lists nodes by min(M, deg) descending, so the
// nodes with min(M, deg) > l are a prefix), then compaction of the kept lists.
struct Prune2Dev {
    unsigned N;
    uint32_t const *off;    // N+1, current CSR
    uint32_t const *ids;
    float const *dists;
    uint32_t const *Meff;   // min(M[i], deg[i])
    uint32_t const *order;  // nodes by Meff descending
    uint32_t *kl;           // kept ids at off[i] + [0, nl[i])
    float *kd;              // their distances
    uint32_t *nl;           // kept count per node
    uint32_t *st;           // per-level state: 1 undecided, 2 pruned, 3 kept
};
hipError_t kg_prune2_screen(Prune2Dev const &p, unsigned l, unsigned nact, hipStream_t s);
hipError_t kg_prune2_resolve(Prune2Dev const &p, unsigned l, unsigned nact, uint32_t *undecided, hipStream_t s);
hipError_t kg_prune2_copy(Prune2Dev const &p, uint32_t const *off2, uint32_t *ids2, float *dists2, hipStream_t s);

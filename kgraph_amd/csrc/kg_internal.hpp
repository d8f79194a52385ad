// kg_internal.hpp -- host/device shared structures of libkgraph_amd (not part
// of the public C ABI).
#pragma once
#include <stddef.h>
#include <stdint.h>

#include <hip/hip_runtime.h>

struct kg_search_args {
    // data rows (row stride Dp floats, 16-byte aligned) and per-row norms
    float const *X;
    uint16_t const *Xb;  // non-NULL: lossless BF16 copy of X (same stride), gathered instead of X
    // non-NULL (with X, Xb and X8 NULL): BF16 copy of X (same stride, rounded
    // to nearest even) that screens each batch on MFMA before the exact FP32
    // re-score of the candidates that can still enter the pool
    uint16_t const *Xs;
    // screen error bound: scr_u row rounding (2^-8 relative), scr_c the rest
    // of the dot's error relative to |x||q|, scr_g rounding of norms and sums
    float scr_u, scr_c, scr_g;
    unsigned long long *rescored;  // adds the candidates the screen passed to the re-score (may be NULL)
    uint8_t const *X8;   // non-NULL: lossless 8-bit copy of X (row stride Dp bytes), gathered first
    float const *norms;
    unsigned N, D, Dp;
    // graph in CSR form
    uint32_t const *goff;  // N+1
    uint32_t const *gids;
    uint32_t const *gM;
    uint4 const *gnode;    // per node {CSR offset, degree, M, 0} (kg_launch_gnode)
    // padded adjacency (may be NULL): row n holds the first gstride ids of
    // node n's list at gpad + n * gstride, gmd[n] = min(deg, 65535) << 16 | M;
    // every id a walk reads when search M <= gstride (kg_launch_gpad)
    uint32_t const *gpad;
    uint32_t const *gmd;
    unsigned gstride;
    // walk layout (both NULL: dataset order).  The rows, norms and padded
    // adjacency above are the walk's copies in another node order: node i is
    // dataset row perm[i], row r is node inv[r].  Starting ids and results are
    // dataset rows; ties between starting points go by dataset row.
    uint32_t const *perm;
    uint32_t const *inv;
    unsigned lists_unique;  // 1: no adjacency list repeats an id
    unsigned int_dot;       // 1: over the 8-bit rows, integer queries take the exact v_dot4 path
    // queries
    float const *Q;
    float const *qnorms;
    unsigned nq, qstride;
    uint32_t const *Ks;    // per-query K (NULL = K0)
    uint32_t const *Ps;    // per-query P (NULL = P0)
    uint32_t const *pidx;  // per-query index into seed_off (when Ps != NULL)
    // order in which the walks start (NULL: query order): the i-th query a
    // wave takes is qorder[i]; larger P first, so that the last walks to
    // start are short ones
    uint32_t const *qorder;
    // lean = 1: this launch is the lean walk (8-bit rows, integer queries,
    // caller's starting points, T = 1); lean_flag (may be NULL): a word that
    // is 0 when every query of the batch is an integer vector in [0, 255] --
    // of a lean / general pair of launches for one batch, only the lean one
    // runs then, only the general one otherwise
    unsigned lean;
    uint32_t const *lean_flag;
    unsigned K0, P0;
    unsigned S, T, Mparam, init;
    float epsilon;
    uint32_t const *seeds;     // concatenated T x P seed ids per distinct P
    uint32_t const *seed_off;  // start of each distinct P's block
    uint32_t const *init_ids;  // init > 0: caller's starting ids, row stride ostride
    // outputs
    uint32_t *out_ids;
    float *out_dists;
    unsigned ostride;
    uint32_t *counts, *comps, *scans;
    float *lat_us;  // per-query wall time inside the kernel (may be NULL)
    // per-wave scratch
    uint32_t *visited;
    unsigned vis_words;
    unsigned vcap;  // entries of the per-wave LDS visited set (2^k or 3 * 2^k >= 256, or 0: bitset only)
    // entry width of that set: 32 (the id), or 16 for hashed ids below 2^B
    // (vt_idmask = 2^B - 1) with vcap = c 2^k: home slot (c p) >> vt_shift
    // (vt_shift = B - k), remainder of vt_rb bits (B - k for c = 1, B - k - 1
    // for c = 3, at most 10) beside the probe displacement
    unsigned vt_bits, vt_rb, vt_idmask, vt_shift;
    uint32_t *spills;  // counts queries whose LDS set moved into the bitset (may be NULL)
    uint32_t *touched;
    unsigned touch_cap;
    uint32_t *qcounter;
    // set by kg_launch_init_check when a starting id is out of range: every
    // wave then exits at once (may be NULL)
    uint32_t const *abort_flag;
    unsigned pool_cap, res_cap, lds_per_wave;  // lds_per_wave: LDS bytes of one query's walk
    unsigned seed_cap;  // starting points a query sorts: max P (random) or init
};

struct kg_brute_args {
    float const *X;
    float const *norms;
    unsigned N, D, Dp;
    float const *Q;       // queries (row stride qstride) or NULL: query = data row rows[q]
    float const *qnorms;
    unsigned qstride;
    uint32_t const *rows;  // data rows used as queries when Q == NULL (self excluded)
    uint32_t const *qsel;  // optional subset of query indices (NULL = all)
    unsigned nq;
    unsigned K;
    float epsilon;
    uint32_t *out_ids;
    float *out_dists;
    unsigned ostride;
    uint32_t *counts;
    uint32_t *comps, *scans;  // optional: set to N and 0 (kgraph.cpp:428-431)
    uint64_t *partial;        // optional scratch: nq * kg_brute_splits(nq, N, K) * K keys
};

struct kg_entry_args {
    float const *rows;    // E sampled rows, stride Dp
    uint4 const *tiles;   // their BF16 copy in MFMA operand order (kg_launch_entry_tiles)
    uint32_t const *gate; // MFMA selection: non-NULL, the kernel returns at once while *gate == 0
    float const *enorms;  // their norms
    uint32_t const *ids;  // their node ids (ascending)
    unsigned E, D, Dp;
    float const *Q;
    float const *qnorms;
    unsigned nq, qstride;
    unsigned m;
    float *scores;  // nq x E
    uint32_t *out_ids;
    unsigned ostride;
};

struct kg_rescore_args {
    float const *X;
    float const *norms;
    unsigned D, Dp;
    float const *Q;
    float const *qnorms;
    unsigned nq, qstride;
    uint32_t const *short_ids;  // nq x s node ids
    unsigned s, m;
    uint32_t *out_ids;
    unsigned ostride;
};

// Launchers (defined in the .hip files).
size_t kg_search_lds_per_wave(unsigned pool_cap, unsigned res_cap, unsigned Dp, unsigned seed_cap, unsigned vcap,
                              unsigned vt_bits, bool csr_base, bool screen, bool lean = false);
// The walk's BF16 screen constants for rows of D values (a.scr_u / scr_c / scr_g).
void kg_search_screen_constants(unsigned D, kg_search_args &a);
// Screened distances and error bounds of m nodes per query (ids: a.nq x m),
// with the walk's screen arithmetic; needs Xs, norms, D, Dp <= 256, Q,
// qnorms, qstride, nq and the screen constants of a.
hipError_t kg_launch_screen_dists(int metric, kg_search_args const &a, uint32_t const *ids, unsigned m, float *out_d,
                                  float *out_e, hipStream_t s);
// Largest waves per block search_kernel is compiled for (its launch bound);
// kg_launch_search rejects larger shapes.
unsigned kg_search_max_wpb();
// Waves per CU the kernel is compiled to keep resident (its register budget).
unsigned kg_search_waves_per_cu();
// The same for the lean walk (kg_search_args::lean).
unsigned kg_search_lean_waves_per_cu();
// *flag = 1 when a value of the queries is not an integer in [0, 255] (caller zeroes it).
hipError_t kg_launch_q8check(float const *Q, unsigned nq, unsigned D, size_t stride, uint32_t *flag, hipStream_t s);
// gnode[n] = {goff[n], deg(n), M[n], 0}; *dupflag |= 1 if some list repeats an id.
hipError_t kg_launch_gnode(unsigned N, uint32_t const *goff, uint32_t const *gids, uint32_t const *gM, uint4 *gnode,
                           uint32_t *dupflag, hipStream_t s);
// queries_per_wave: 1, or 2 (each query on half a wave; not with the BF16
// screen); a.lds_per_wave is the LDS of one query.
hipError_t kg_launch_search(int metric, kg_search_args const &a, unsigned blocks, unsigned waves_per_block,
                            unsigned queries_per_wave, hipStream_t s);
// *width = max over nodes of min(M[n], deg(n)) (atomicMax; zeroed by the caller).
hipError_t kg_launch_gpad_width(unsigned N, uint32_t const *goff, uint32_t const *gM, uint32_t *width, hipStream_t s);
// gpad rows (first min(deg, gstride) ids of each list) and gmd words.
hipError_t kg_launch_gpad(unsigned N, uint32_t const *goff, uint32_t const *gids, uint32_t const *gM, unsigned gstride,
                          uint32_t *gpad, uint32_t *gmd, hipStream_t s);
// Walk layout (kg_layout.hip).  Scratch of kg_walk_order: lab, flag, vis,
// tree, ord, kids, first, done: N words; pos, koff, tsize, tbase: N + 1; tmp:
// N / 1024 + 2; claim: N; word: 1; host: one word of pinned host memory.
struct kg_order_scratch {
    uint32_t *lab, *flag, *vis, *tree, *ord, *kids, *first, *done;
    uint32_t *pos, *koff, *tsize, *tbase, *tmp, *word, *host;
    unsigned long long *claim;
};
// perm / inv of the breadth-first node order of the padded adjacency (gpad
// rows of width W, gmd words); waits on the stream once per level.
hipError_t kg_walk_order(unsigned N, uint32_t const *gpad, uint32_t const *gmd, unsigned W, kg_order_scratch const &w,
                         uint32_t *perm, uint32_t *inv, unsigned *levels_out, unsigned *rounds_out, hipStream_t s);
// out row i = in row perm[i] (rows of `words` 32-bit words)
hipError_t kg_permute_rows(unsigned N, unsigned words, uint32_t const *perm, void const *in, void *out, hipStream_t s);
// gpad2 row i = gpad row perm[i] with ids renamed by inv; gmd2[i] = gmd[perm[i]]
hipError_t kg_permute_adjacency(unsigned N, unsigned W, uint32_t const *perm, uint32_t const *inv,
                                uint32_t const *gpad, uint32_t const *gmd, uint32_t *gpad2, uint32_t *gmd2,
                                hipStream_t s);
// *flag = *host_flag = 1 if one of the first `init` ids of a row (row q at
// ids + q*ostride) is >= N (host_flag: pinned host memory); both are left
// alone otherwise.
// With Q non-NULL it also runs kg_launch_q8check's test of the nq x D queries
// (row stride qstride) into qflag, in the same launch.
hipError_t kg_launch_init_check(uint32_t const *ids, unsigned nq, size_t ostride, unsigned init, unsigned N,
                                uint32_t *flag, uint32_t *host_flag, float const *Q, unsigned D, size_t qstride,
                                uint32_t *qflag, hipStream_t s);
hipError_t kg_launch_brute(int metric, kg_brute_args const &a, hipStream_t s);
unsigned kg_brute_splits(unsigned nq, unsigned N, unsigned K);
hipError_t kg_launch_entries(int metric, kg_entry_args const &a, hipStream_t s);
hipError_t kg_launch_rescore(int metric, kg_rescore_args const &a, hipStream_t s);
// BF16 MFMA screening; a.enorms = ||e||^2/2 (Euclidean forms) or ||e|| (angular)
// Fused MFMA top-m entry selection (see kg_entry.hip): exact for rows /
// queries whose BF16 tiles are exact, a screen otherwise; qsq / esq:
// ||q||^2/2 and ||e||^2/2 (used for L2SQR).
// The entry sample in the MFMA selection's operand order: 16-row tiles of
// kg_entry_slices(D) 32-value slices; 16 bytes per lane of a wave, chunk
// (t KS + s) 64 + l holding row 16 t + l % 16, values 32 s + 8 (l / 16) .. + 8
// as BF16 (round to nearest even; zero past D and past E), so a wave reads a
// slice of a tile as one contiguous 1 KiB.
unsigned kg_entry_slices(unsigned D);
size_t kg_entry_tiles_size(unsigned E, unsigned D);  // uint4 chunks
hipError_t kg_launch_entry_tiles(float const *rows, unsigned E, unsigned D, unsigned Dp, uint4 *tiles,
                                 hipStream_t s);
hipError_t kg_launch_entries_exact_mfma(int metric, kg_entry_args const &a, float const *qsq, float const *esq,
                                        hipStream_t s);
// flags[0] |= 1 if a value of the n x D rows (row stride `stride`) is not an
// integer; flags[1] = max |value| as float bits.  Caller zeroes flags.
hipError_t kg_launch_intcheck(float const *X, unsigned n, unsigned D, unsigned stride, uint32_t *flags,
                              hipStream_t s);
// BF16 tiles give the canonical FP32 distances exactly when every row and
// query value is an integer of magnitude <= 256 and D (max|q|^2 + max|e|^2)
// <= 2^24: every product, partial sum and norm is then an integer <= 2^24.
// qflag0: the queries' non-integer flag; qmax, emax: largest |q|, |e|.
__host__ __device__ inline bool kg_entry_exact_bf16(uint32_t qflag0, float qmax, unsigned D, float emax) {
    return qflag0 == 0 && qmax <= 256.f && emax <= 256.f &&
           double(D) * (double(qmax) * qmax + double(emax) * emax) <= 16777216.0;
}
// qsq[q] = ||q||^2 / 2 (canonical order) and the kg_launch_intcheck flags of
// the same rows in one pass: flags[0..2] zeroed by the caller (flags[2]
// counts blocks); the last block copies flags[0..1] to host_out (pinned) and
// sets flags[3] = kg_entry_exact_bf16(...) against sample rows of largest |x| emax.
hipError_t kg_launch_qnorm_check(float const *Q, unsigned n, unsigned D, unsigned stride, float *qsq,
                                 uint32_t *flags, uint32_t *host_out, float emax, hipStream_t s);
hipError_t kg_launch_gather_rows(float const *X, unsigned Dp, uint32_t const *ids, unsigned E, float *out,
                                 hipStream_t s);
// BF16 copy of N x Dp rows (Dp % 4 == 0); *inexact = 1 when a value is not
// representable in BF16 (the caller clears it first)
// X8 = the rows as bytes; *inexact = 1 if a value is not an integer in
// [0, 255] with the bits of its byte's float (caller zeroes it).
hipError_t kg_launch_u8_rows(float const *X, unsigned N, unsigned Dp, uint8_t *X8, uint32_t *inexact, hipStream_t s);
hipError_t kg_launch_bf16_rows(float const *X, unsigned N, unsigned Dp, uint16_t *Xb, uint32_t *inexact,
                               hipStream_t s);
hipError_t kg_launch_norms(int metric, float const *X, unsigned N, unsigned D, unsigned Dp, float *out,
                           hipStream_t s);

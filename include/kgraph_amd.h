/* kgraph_amd.h -- C ABI of libkgraph_amd.so: KGraph's NN-descent k-NN graph
 * build and graph-walk search on one AMD Instinct MI355X (gfx950).
 *
 * Plain pointers and sizes only.  Every function returns 0 on success and a
 * negative status on failure; kg_last_error() then holds the message.  The
 * status tells which C++ exception the reference would throw
 * (kgraph.h:287-300): KG_ERR_INVALID_ARGUMENT <-> kgraph::invalid_argument,
 * KG_ERR_RUNTIME <-> kgraph::runtime_error, KG_ERR_IO <-> kgraph::io_error.
 *
 * A handle owns a device copy of the dataset rows (the dense form of the
 * reference's IndexOracle / SearchOracle plugins, kgraph.h:49-86 and
 * kgraph-data.h:302-331 / python-api.cpp:87-129) and the index (the
 * reference's KGraphImpl::graph + M, kgraph.cpp:276-280).  A handle is not
 * re-entrant: calls on one handle must not overlap.  Different handles may be
 * used from different threads.
 */
#ifndef KGRAPH_AMD_H
#define KGRAPH_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct kg_handle kg_handle;

enum {
    KG_OK = 0,
    KG_ERR_INVALID_ARGUMENT = -1,
    KG_ERR_RUNTIME = -2,
    KG_ERR_IO = -3,
    KG_ERR_DEVICE = -4
};

/* Distance plugins.  L2SQR: sum of squared differences (the CLI's
 * MatrixOracle<float, metric::l2sqr>, kgraph-data.h:41-61, index.cpp:168).
 * EUCLIDEAN: n_i + n_j - dot with n = ||x||^2/2 (python-api.cpp:51-59).
 * ANGULAR: -dot / (||x|| ||y|| + FLT_EPSILON) (python-api.cpp:61-69). */
enum { KG_METRIC_L2SQR = 0, KG_METRIC_EUCLIDEAN = 1, KG_METRIC_ANGULAR = 2 };

/* File formats (kgraph.h:128-132). */
enum { KG_FORMAT_DEFAULT = 0, KG_FORMAT_NO_DIST = 1, KG_FORMAT_TEXT = 128 };

/* Screening precision of the build's local join and of search (no
 * counterpart in the reference): FP32 scores every pair exactly; BF16 scores
 * candidates with BF16 dot products on MFMA first and re-scores every
 * admission in exact FP32; FP32_MFMA screens the local join with FP32
 * operands on MFMA (v_mfma_f32_16x16x4_f32, an error bound about 500 times
 * tighter than BF16's) and re-scores the same way.  Every form admits exactly
 * the candidates the exact FP32 comparison admits. */
enum { KG_PRECISION_FP32 = 0, KG_PRECISION_BF16 = 1, KG_PRECISION_FP32_MFMA = 2 };

/* KGraph::IndexParams (kgraph.h:94-110), same fields, order and meaning. */
typedef struct {
    unsigned iterations, L, K, S, R, controls, seed;
    float delta, recall;
    unsigned prune;
    int reverse;
} kg_index_params;

/* KGraph::IndexInfo (kgraph.h:135-147). stop_condition: 0 ITERATION,
 * 1 DELTA, 2 RECALL. */
typedef struct {
    int stop_condition;
    unsigned iterations;
    float cost, recall, accuracy, delta, M;
} kg_index_info;

/* KGraph::SearchParams (kgraph.h:113-126). */
typedef struct {
    unsigned K, M, P, S, T;
    float epsilon;
    unsigned seed, init;
} kg_search_params;

/* KGraph::version (kgraph.h:201, kgraph.cpp:1014-1016). */
const char *kg_version(void);

/* KGraph::create (kgraph.h:199, kgraph.cpp:1010-1012) / delete. */
kg_handle *kg_create(void);
void kg_destroy(kg_handle *h);
const char *kg_last_error(const kg_handle *h);

/* Defaults of IndexParams() / SearchParams() (kgraph.h:14-37, 108-125). */
void kg_index_params_default(kg_index_params *p);
void kg_search_params_default(kg_search_params *p);

/* Global verbosity (kgraph.h:42): 0 silences progress lines on stderr. */
void kg_set_verbosity(unsigned v);

/* Dataset rows: N rows of D floats, `stride` floats apart, in host memory
 * (kg_set_data) or device memory (kg_set_data_device).  The handle keeps its
 * own padded device copy and the per-row norms of `metric`.  This is the
 * dense oracle the reference's build()/search() receive
 * (kgraph.cpp:1003, 423; python-api.cpp:110-114). */
int kg_set_data(kg_handle *h, const float *X, unsigned N, unsigned D, size_t stride, int metric);
int kg_set_data_device(kg_handle *h, const float *dX, unsigned N, unsigned D, size_t stride, int metric);

/* KGraph::build (kgraph.h:167, kgraph.cpp:886-999): NN-descent on the GPU,
 * then reverse-edge completion (params->reverse) and pruning
 * (params->prune).  info may be NULL. */
int kg_build(kg_handle *h, const kg_index_params *params, kg_index_info *info);

/* Screening precision used by kg_build and kg_search (KG_PRECISION_*). */
int kg_set_precision(kg_handle *h, int precision);

/* Visited set of the graph walk (no counterpart in the reference, which
 * allocates an N-bit bitset per query, kgraph.cpp:438).  Mode 0 (default):
 * each resident wave keeps an exact hash set of visited ids in LDS in front of
 * its N-bit HBM bitset; a query that fills it past half moves it into the
 * bitset and continues there.  Entries take 16 bits when ceil(log2 N) <=
 * log2(entries) + 10 (the hashed id's low bits and the probe displacement,
 * from which slot and entry give the id back), else 32.  Size: 4096 entries,
 * else 2048, when that keeps at least 12 waves per CU resident; else the
 * largest power of two up to 4096 that keeps 16.  Mode 1: the
 * bitset only.  Mode 2: as mode 0 with 32-bit entries.
 * kg_set_search_visited_size fixes the entry count (2^k or 3 * 2^k in
 * [256, 8192]; 0 restores the choice above).  Results are identical in every
 * form. */
int kg_set_search_visited(kg_handle *h, int mode);
int kg_set_search_visited_size(kg_handle *h, unsigned entries);
/* Last walk launch: LDS set entries per wave (0: bitset only) and how many
 * queries moved their set into the bitset; kg_search_visited_bits: the width
 * of its entries (16 or 32). */
int kg_search_visited_stats(kg_handle *h, unsigned *vcap, unsigned *spilled);
int kg_search_visited_bits(kg_handle *h, unsigned *bits);

/* Adjacency the graph walk reads (no counterpart in the reference, whose
 * walk reads each node's vector of neighbours, kgraph.cpp:488-493).  Mode 0
 * (default): besides the CSR lists the index keeps padded rows, the first W
 * ids of every list at a fixed stride with W = max over nodes of
 * min(M[n], deg(n)) rounded up to 16 (used when they fit in 16 GiB and the
 * search's M <= W, which covers every id such a walk reads), so a node's ids
 * load with its M and degree, at its first expansion, instead of a 16-byte
 * record for every fresh candidate.  Mode 1: the CSR lists and per-node
 * records.  Results are identical.  kg_search_adjacency reports what the last
 * walk read (*padded = 1 for the padded rows) and W (0: none). */
int kg_set_search_adjacency(kg_handle *h, int mode);
int kg_search_adjacency(kg_handle *h, int *padded, unsigned *width);

/* Walk layout (no counterpart in the reference; its search reads the rows in
 * dataset order).  1 (default): graph walks read copies of the rows they
 * gather, of the norms and of the padded adjacency laid out in a
 * breadth-first order of the graph (the nodes a walk reads together sit
 * together in memory); 0: dataset order.  Results are identical in both: the
 * walk maps its starting ids in and its results out.  The order is computed
 * on the device once per index, at the first search that needs it.  Walks
 * over the CSR rows (kg_set_search_adjacency 1, search M beyond the padded
 * width) use dataset order. */
int kg_set_search_order(kg_handle *h, int mode);
/* Whether the last walk launch read the reordered copies, the order's
 * breadth-first levels and the milliseconds it took (with the reordered
 * adjacency; 0 when none exists); perm (may be NULL): N entries, perm[i] =
 * dataset row of node i of the order. */
int kg_search_order(kg_handle *h, int *active, unsigned *levels, float *ms, unsigned *perm);
/* Makes the walk's index-side state now rather than at the first search
 * after a build, load or graph change (no counterpart in the reference, whose
 * search reads the index as built): per-node records, the padded adjacency,
 * the walk layout and its copies of the rows the current settings gather. */
int kg_prepare_search(kg_handle *h);
/* Lean walk (no counterpart in the reference): 1 (default) = a batch of
 * queries over the 8-bit rows that starts from caller-given points (init > 0,
 * T = 1, no BF16 screen, one query per wave) is walked by a form of the walk
 * without random seeds, trial merge or FP32 query copy, which fits 5 waves per
 * SIMD, when every query is an integer vector in [0, 255] (checked on the
 * device; otherwise the general walk runs); 0 = the general walk only.
 * Results are identical. */
int kg_set_search_lean(kg_handle *h, int on);

/* BF16 screen of the graph walk's candidates (the reference scores every
 * fresh neighbour in full precision before UpdateKnnList, kgraph.cpp:497-512).
 * Mode 1: walks that gather the FP32 rows (rows with no lossless BF16 or
 * 8-bit copy, D <= 256) score each batch of a full pool first with BF16 dot
 * products on MFMA (v_mfma_f32_16x16x32_bf16, FP32 accumulate; the row in
 * BF16, the query as BF16 high and low halves) and re-score in exact FP32
 * every candidate whose screened distance minus a rigorous error bound is
 * below the pool's last distance; the others cannot enter the pool.  Results,
 * comps and scans are identical to mode 0 (off, default).
 * kg_search_screen: whether the last walk launch screened and how many
 * candidates it re-scored (comps counts every screened candidate).
 * kg_search_screen_dists: the screen's distance and error bound for m node
 * ids per query (ids: nq x m, out_d / out_e: nq x m), with the walk's
 * arithmetic; |screened - exact FP32| <= bound. */
int kg_set_search_screen(kg_handle *h, int mode);
/* Queries per wavefront of the graph walk (no counterpart in the reference,
 * whose walk is one query per OpenMP thread, kgraph.cpp:423-564): 1 (default)
 * walks each query on a whole 64-lane wave, 2 walks two queries per wave on
 * 32 lanes each, with their own pools, LDS visited sets and bitsets (twice the
 * memory accesses in flight per wave; the BF16 screen runs with 1 only).
 * Results are identical. */
int kg_set_search_queries_per_wave(kg_handle *h, unsigned queries);
int kg_search_screen(kg_handle *h, int *active, unsigned long long *rescored);
int kg_search_screen_dists(kg_handle *h, const float *Q, unsigned nq, size_t qstride, const unsigned *ids, unsigned m,
                           float *out_d, float *out_e);

/* Row storage for graph walks (no counterpart in the reference, whose rows are
 * the caller's Matrix, kgraph-data.h:115-175).  kg_set_data keeps a BF16 copy
 * of the rows (the MFMA kernels' operands), lossless when every value is
 * representable in BF16, and an 8-bit copy when every value is an integer in
 * [0, 255].  Mode 0 (default): walks gather the 8-bit copy when there is one
 * (a quarter of the bytes: one 128-byte line per 128-value row), else the
 * lossless BF16 copy (half), else the FP32 rows; each holds the same values,
 * so the distances and ids are the same.  Mode 1 gathers the FP32 rows, mode
 * 2 the BF16 copy when lossless.  kg_row_storage reports *active = 0 (FP32
 * rows), 1 (BF16 copy) or 2 (8-bit copy) for the next walk, and *lossless
 * with bit 0 set when the BF16 copy is lossless and bit 1 when the 8-bit copy
 * exists (either pointer may be NULL). */
int kg_set_row_storage(kg_handle *h, int mode);
int kg_row_storage(kg_handle *h, int *active, int *lossless);

/* Batched KGraph::search (kgraph.h:187-197, kgraph.cpp:423-564): nq queries
 * of D floats, qstride floats apart.  Ks / Ps give per-query K and P (NULL =
 * params->K / params->P).  Row q of the results starts at ids + q*ostride
 * (and dists + q*ostride; dists may be NULL); counts[q] = number of results
 * (the reference's return value).  comps / scans (may be NULL) receive each
 * query's distance evaluations after seeding (SearchInfo::cost * N,
 * kgraph.cpp:561) and neighbour ids read.  With params->init > 0 the first
 * init ids of each row of `ids` are the starting points (kgraph.cpp:461-467).
 * kg_search takes host buffers; kg_search_device takes device buffers for
 * Q / ids / dists / counts / comps / scans (Ks / Ps stay in host memory). */
int kg_search(kg_handle *h, const float *Q, unsigned nq, size_t qstride, const kg_search_params *params,
              const unsigned *Ks, const unsigned *Ps, unsigned *ids, float *dists, size_t ostride,
              unsigned *counts, unsigned *comps, unsigned *scans);
int kg_search_device(kg_handle *h, const float *dQ, unsigned nq, size_t qstride, const kg_search_params *params,
                     const unsigned *Ks, const unsigned *Ps, unsigned *d_ids, float *d_dists, size_t ostride,
                     unsigned *d_counts, unsigned *d_comps, unsigned *d_scans);

/* Exact top-K by full scan (SearchOracle::search, kgraph.cpp:231-252): the K
 * smallest (dist, id) pairs with dist <= epsilon.  Host buffers. */
int kg_bruteforce(kg_handle *h, const float *Q, unsigned nq, size_t qstride, unsigned K, float epsilon,
                  unsigned *ids, float *dists, size_t ostride, unsigned *counts);

/* Exact top-K by full scan that allocates nothing: no counterpart in the
 * reference; for callers that cannot allocate (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same results as
 * kg_bruteforce, from device-resident queries (16-byte aligned rows, qstride
 * a multiple of 4 floats) into device-resident ids / dists / counts (dists
 * may be NULL).  Its scratch (query norms, partial top-K lists) is a 36 MiB
 * device array of the library, one per GPU for the life of the process;
 * nq <= 1,048,576 per call, larger calls are rejected with
 * KG_ERR_INVALID_ARGUMENT.  Calls that overlap on one GPU race on that array:
 * serialise them (one stream, or one call at a time).  Returns when the scan
 * has finished. */
int kg_bruteforce_device_noalloc(kg_handle *h, const float *dQ, unsigned nq, size_t qstride, unsigned K,
                                 float epsilon, unsigned *d_ids, float *d_dists, size_t ostride,
                                 unsigned *d_counts);

/* KGraph::save / load (kgraph.h:159-165, kgraph.cpp:291-373), same
 * KNNGRAPH v2 byte format; FORMAT_TEXT is write-only. */
int kg_save(kg_handle *h, const char *path, int format);
int kg_load(kg_handle *h, const char *path);

/* KGraph::get_nn (kgraph.h:205-224, kgraph.cpp:566-582).  nns / dists must
 * hold the node's L entries; either may be NULL. */
int kg_get_nn(kg_handle *h, unsigned id, unsigned *nns, float *dists, unsigned *M, unsigned *L);

/* KGraph::reverse (kgraph.h:226, kgraph.cpp:667-702) and KGraph::prune
 * (kgraph.h:181, kgraph.cpp:584-665). */
int kg_reverse(kg_handle *h, int rev);
int kg_prune(kg_handle *h, unsigned level);

/* Index size: N nodes and total stored edges. */
int kg_index_size(kg_handle *h, unsigned *N, uint64_t *edges);

/* Whole-index exchange in CSR form: offsets[N+1] (edge positions), ids /
 * dists [edges] (dists may be NULL: no-distance index), M[N]. */
int kg_set_graph(kg_handle *h, unsigned N, const uint64_t *offsets, const unsigned *ids, const float *dists,
                 const unsigned *M);
int kg_get_graph(kg_handle *h, uint64_t *offsets, unsigned *ids, float *dists, unsigned *M);

/* Device time in milliseconds of the last call's phases, measured with HIP
 * events on the handle's stream: which = 0 search kernel, 1 brute-force
 * kernel, 2 whole build, 3 NN-descent join kernels (sum), 4 merge kernels
 * (sum), 5 update kernels (sum), 6 reverse, 7 entry selection, 8 NN-descent
 * iterations (sum; 3-5 are measured only when iterations run kernel by
 * kernel, see kg_set_build_graphs). */
int kg_last_time_ms(kg_handle *h, int which, float *ms);

/* Per-iteration trace of the last build: 4 floats per iteration (recall,
 * delta, M, cost), up to cap iterations.  Returns the iteration count in *n. */
int kg_build_trace(kg_handle *h, float *out, unsigned cap, unsigned *n);

/* Entry-point selection (no counterpart in the reference; it feeds the
 * reference's own SearchParams::init starting points, kgraph.cpp:461-467).
 * kg_set_entry_sample picks E distinct nodes (GenRandom on mt19937(seed),
 * kgraph.cpp:44-65, sorted by id) and keeps a compact copy of their rows;
 * E = 0 drops the sample.  kg_select_entries[_device] writes, for each query,
 * the m nearest sampled nodes (exact FP32 distance, ties by id, m <= 16) to
 * ids[q*ostride + 0..m); a following kg_search with params->init = m (m < K)
 * walks from them.  kg_select_entries_device returns once the selection is
 * queued on the handle's stream (it waits only for its short query test, which
 * picks the form, while the selection itself runs): a following call on the
 * same handle (such as kg_search_device) sees its ids, any other reader
 * synchronises the device first (hipDeviceSynchronize). */
int kg_set_entry_sample(kg_handle *h, unsigned E, unsigned seed);
int kg_select_entries(kg_handle *h, const float *Q, unsigned nq, size_t qstride, unsigned m, unsigned *ids,
                      size_t ostride);
int kg_select_entries_device(kg_handle *h, const float *dQ, unsigned nq, size_t qstride, unsigned m,
                             unsigned *d_ids, size_t ostride);

/* Learned projector (no counterpart in the reference): an MLP
 * f(x) = W2 relu(W1 s(x) + b1) + b2, D -> hidden -> out_dim, FP32 master
 * weights, products on BF16 MFMA with FP32 accumulation, where
 * s(x)[c] = (x[c] - mu[c]) * inv_sigma[c] standardises each column with the
 * dataset's column mean and 1/std (fixed at kg_mlp_init).
 * kg_mlp_init: column statistics of the dataset, Xavier-uniform weights from
 *   mt19937(seed), zero biases.
 * kg_mlp_get_normalization: mu and inv_sigma (D floats each; either may be NULL).
 * kg_mlp_train: Adam on the triplet loss max(0, |f(a)-f(p)|^2 - |f(a)-f(n)|^2
 *   + margin) over `steps` batches of `batch` triplets (a random node, one of
 *   its first M graph neighbours, a node outside its list: uniform for half of
 *   the triplets, two hops away for the other half) drawn from the current
 *   index, learning rate decayed from lr to 0 by a cosine over the call;
 *   loss_trace (may be NULL) receives the mean loss of every step.
 * kg_mlp_loss_grad: loss and gradients for caller-given triplets
 *   (3*nt node ids), for checking the backward kernels.
 * kg_mlp_project: f(X) for host rows X (n x out_dim floats out).
 * kg_set_entry_projection: with shortlist s > 0, kg_select_entries scores the
 *   entry sample in the projected space, keeps the s best and re-ranks them
 *   with the exact FP32 distance (m <= s <= 16); 0 restores exact scoring. */
int kg_mlp_init(kg_handle *h, unsigned hidden, unsigned out_dim, unsigned seed);
int kg_mlp_train(kg_handle *h, unsigned steps, unsigned batch, float lr, float margin, unsigned seed,
                 float *loss_trace);
int kg_mlp_loss_grad(kg_handle *h, const unsigned *triplets, unsigned nt, float margin, float *loss, float *dW1,
                     float *db1, float *dW2, float *db2);
int kg_mlp_get_weights(kg_handle *h, float *W1, float *b1, float *W2, float *b2);
int kg_mlp_get_normalization(kg_handle *h, float *mu, float *inv_sigma);
int kg_mlp_set_weights(kg_handle *h, const float *W1, const float *b1, const float *W2, const float *b2);
int kg_mlp_project(kg_handle *h, const float *X, unsigned n, size_t stride, float *Z);
int kg_set_entry_projection(kg_handle *h, unsigned shortlist);

/* Launch-shape autotuner for kg_search (no counterpart in the reference):
 * runs the search kernel on the given device-resident query sample with
 * every (waves per block, resident blocks per CU) pair in {1,2,4} x
 * {4,8,16,32} that fits 32 waves per CU, times each with HIP events and keeps
 * the fastest for later searches on this handle.  chosen (may be NULL)
 * receives {waves per block, blocks per CU}. */
int kg_autotune_search(kg_handle *h, const float *dQ, unsigned nq, size_t qstride, const kg_search_params *params,
                       const unsigned *Ks, const unsigned *Ps, unsigned *chosen);

/* Walk autotuner (no counterpart in the reference): on the given
 * device-resident query sample (with the caller's starting points when
 * params->init > 0: d_init holds init ids per query, row stride istride),
 * times the walk over the settings that leave its results unchanged and keeps
 * the fastest for later searches on this handle: the visited set (bitset
 * only, or an LDS set of 1024, 2048, 4096 or 8192 entries in front of it),
 * the adjacency rows (padded or CSR), with padded rows the walk layout
 * (kg_set_search_order; kg_search_order reports the choice), the lean walk
 * (kg_set_search_lean, where it applies), the BF16 screen
 * (where walks gather the FP32 rows), queries per wave (1, or 2 when the
 * screen is off) and the launch shape (as kg_autotune_search).  One setting at
 * a time, in that order, each kept when faster.  chosen (may be NULL) receives 6 values: {waves per block,
 * blocks per CU, visited-set entries (0: bitset only), adjacency (0 padded, 1
 * CSR), screen (0/1), queries per wave}. */
int kg_autotune_walk(kg_handle *h, const float *dQ, unsigned nq, size_t qstride, const kg_search_params *params,
                     const unsigned *Ks, const unsigned *Ps, const unsigned *d_init, size_t istride,
                     unsigned *chosen);

/* Launch shape of later walks (no counterpart in the reference): waves per
 * block (1, 2 or 4) and at most blocks_per_cu resident blocks per CU (0: as
 * many as the LDS and the 32-wave limit allow).  Results do not depend on it. */
int kg_set_search_launch(kg_handle *h, unsigned waves_per_block, unsigned blocks_per_cu);

/* Per-query latency of the last kg_search[_device] call in microseconds,
 * stamped inside the search kernel from the query's first to its last
 * instruction (walk queries only; brute-force and skipped rows read 0).
 * Copies min(n, nq) values. */
int kg_search_latencies(kg_handle *h, float *out_us, unsigned n);

/* Work counters of the last build, summed over iterations (no counterpart in
 * the reference): out[0] pairs scored (the reference's n_comps,
 * kgraph.cpp:805, 814), out[1] join-list entries (sum over nodes of
 * |new| + |old|), out[2] candidates pushed to merge slots, out[3]
 * iterations, out[4] local-join form (0 FP32 quads, 1 BF16 MFMA screen with
 * exact FP32 re-score, 2 BF16 MFMA exact on these rows, 3 FP32-operand MFMA
 * screen with exact FP32 re-score, 4 8-bit MFMA exact on these rows), out[5] iterations
 * run as replays of the captured iteration graph, out[6] pairs the screened
 * forms re-scored in exact FP32.  n = number of entries wanted (<= 7). */
int kg_build_counters(kg_handle *h, uint64_t *out, unsigned n);

/* 8-bit local join (no counterpart in the reference): 1 (default) = with
 * KG_PRECISION_BF16, rows whose values are all integers in [0, 255] and a
 * Euclidean metric (L2SQR or half-SSD) with 255^2 D <= 2^24 are joined on
 * v_mfma_i32_16x16x64_i8 (shifted to [-128, 127]; every distance is the exact
 * canonical one); 0 = the BF16 form. */
int kg_set_join_int8(kg_handle *h, int on);

/* Local-join form when the BF16 join is a screen (KG_PRECISION_BF16 on rows
 * that are not BF16-exact; no counterpart in the reference).  Both forms admit
 * exactly the same candidates; which is faster depends on the data.
 * KG_JOIN_AUTO (default): every third iteration, from the first, splits the
 * nodes between the screen and the exact FP32 tiles (halves the first time,
 * then the current form on all but the last eighth), times both parts, and the
 * iterations after it use the form that was faster per node.  KG_JOIN_SCREEN /
 * KG_JOIN_TILES force one form. */
enum { KG_JOIN_AUTO = 0, KG_JOIN_SCREEN = 1, KG_JOIN_TILES = 2 };
int kg_set_join_form(kg_handle *h, int form);

/* Per-iteration record of the last build (no counterpart in the reference):
 * 6 values per iteration -- pairs scored, candidates pushed, join-list
 * entries, join form (0 FP32 tiles, 1 BF16 screen + exact re-score, 2 exact
 * BF16 tiles, 3 probe: the nodes split between screen and tiles, 4 FP32-operand
 * MFMA screen + exact re-score), join
 * kernel time in microseconds (0 when the iteration ran as a graph replay),
 * whole-iteration device time in microseconds.  Copies
 * min(cap, iterations) records; *n = iterations. */
int kg_build_iterations(kg_handle *h, uint64_t *out, unsigned cap, unsigned *n);

/* The BF16 MFMA local join's distance for n id pairs (pairs[2p], pairs[2p+1])
 * and the bound on its error that the join's screen uses (no counterpart in
 * the reference): screened[p] is bit-for-bit the value the join computes for
 * the pair, |screened[p] - exact distance| <= bound[p].  For measuring the
 * screen against an FP64 ground truth; D <= 256. */
int kg_screen_pairs(kg_handle *h, const uint32_t *pairs, unsigned n, float *screened, float *bound);

/* Learned pre-rank of the local join (no counterpart in the reference, whose
 * join scores every new x new and new x old pair of a node's lists,
 * kgraph.cpp:797-815): with 0 < keep < 1 and a projector (kg_mlp_init /
 * kg_mlp_train), every node's join lists are ranked by projected distance to
 * the node and only the nearest fraction `keep` of them takes part in the
 * join; full-dimension FP32 distances still judge every admission.  Guard:
 * before the first iteration a held-out probe (each control node's exact
 * nearest neighbours against random nodes) must be ordered correctly by the
 * projection in at least 90 % of the cases, and the pre-rank stays on only
 * while every iteration closes at least a quarter of the control recall's
 * remaining gap; otherwise the build falls back to plain NN-descent.  keep = 0
 * or 1: off (default).
 * kg_build_prerank reports the last build: out[0] status (0 off, 1 on for the
 * whole build, 2 turned down by the probe, 3 turned off after the control
 * recall slowed down), out[1] the probe's accuracy, out[2] the iteration of
 * the fallback (-1: none). */
int kg_set_join_prerank(kg_handle *h, float keep);
int kg_build_prerank(kg_handle *h, float *out);

/* NN-descent iterations as a HIP graph (no counterpart in the reference; the
 * loop is kgraph.cpp:927-972): with on = 1 (default) the device work of an
 * iteration after the first (update, join, merge, statistics, control pools)
 * is captured once per build and replayed, one graph launch per iteration,
 * with the stop rule read back between replays; on = 0 launches the same
 * kernels one by one and times join / merge / update separately. */
int kg_set_build_graphs(kg_handle *h, int on);

#ifdef __cplusplus
}
#endif

#endif

// kg_handle.hpp -- the state behind a kg_handle (host side of libkgraph_amd).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/kgraph_amd.h"

namespace kga {

// Exception types mirroring kgraph.h:287-300; mapped to KG_ERR_* at the ABI.
struct invalid_argument : std::invalid_argument {
    using std::invalid_argument::invalid_argument;
};
struct runtime_error : std::runtime_error {
    using std::runtime_error::runtime_error;
};
struct io_error : runtime_error {
    using runtime_error::runtime_error;
};
struct device_error : std::runtime_error {
    using std::runtime_error::runtime_error;
};

inline void hip_check(hipError_t e, char const *what) {
    if (e != hipSuccess) throw device_error(std::string(what) + ": " + hipGetErrorString(e));
}
#define KG_HIP(x) ::kga::hip_check((x), #x)

extern unsigned verbosity;

// A device allocation that grows on demand and is freed with the handle.
template <typename T>
struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(DevBuf const &) = delete;
    DevBuf &operator=(DevBuf const &) = delete;
    void reserve(size_t want) {
        if (want <= n) return;
        release();
        if (want) KG_HIP(hipMalloc(&p, want * sizeof(T)));
        n = want;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    ~DevBuf() { release(); }
};

struct Handle {
    std::string err;
    hipStream_t stream = nullptr;
    int device = 0;
    uint32_t *host_flag = nullptr;  // 4 pinned words for flags read back with the stream's work
    int num_cus = 0;
    size_t lds_per_cu = 0;
    int precision = KG_PRECISION_FP32;
    int join_exact = -1;  // last build's join: -1 FP32 tiles, 0 BF16 screen, 1 BF16 exact, 2 FP32-operand screen, 3 8-bit exact
    bool join_i8 = true;  // the exact join takes 8-bit operands when the rows allow (kg_set_join_int8)
    bool entry_int = false;  // entry sample rows are all integers
    float entry_max = 0.f;   // their largest |value|
    DevBuf<float> qsq;       // ||q||^2 / 2 of the last query batch
    DevBuf<float> eprojsq;   // ||z||^2 / 2 of the projected entry sample
    DevBuf<uint32_t> iflags;  // last build: -1 FP32 join, 0 BF16 screen + re-score, 1 exact BF16 tiles

    // dataset (dense oracle): N rows, row stride Dp floats (D rounded up to 4)
    unsigned N = 0, D = 0, Dp = 0;
    int metric = KG_METRIC_L2SQR;
    DevBuf<float> X, norms;
    // BF16 copy of the rows (same stride): MFMA operands; when rows_bf16 (every
    // value representable) graph walks gather it instead of X
    DevBuf<uint16_t> Xb;
    bool rows_bf16 = false;
    // 8-bit copy (row stride Dp bytes), kept when every value is an integer in
    // [0, 255]; walks gather it first
    DevBuf<uint8_t> X8;
    bool rows_u8 = false;
    // 0: walks gather the 8-bit copy, else the BF16 copy, when lossless; 1: FP32
    // rows; 2: the BF16 copy when lossless (not the 8-bit one)
    int row_storage = 0;
    int search_visited = 0;  // 0: LDS visited hash set in front of the HBM bitset, 1: bitset only, 2: 32-bit set entries
    unsigned search_vcap = 0;  // LDS set entries (0: chosen per launch)
    unsigned last_vt_bits = 0;
    unsigned last_vcap = 0;  // LDS visited-set entries of the last walk launch
    // 1: walks over the FP32 rows screen each batch with the BF16 copy on MFMA
    // before the exact FP32 re-score (same results); 0: off
    int search_screen = 0;
    int last_screen = 0;                         // whether the last walk launch screened
    unsigned search_qpw = 1;                     // queries per wave (1, or 2 on half a wave each)
    unsigned last_qpw = 1;
    DevBuf<unsigned long long> screen_cnt;       // candidates it re-scored
    bool has_data = false;

    // index: CSR with 32-bit edge offsets
    unsigned gN = 0;
    bool has_graph = false;
    bool no_dist = false;
    DevBuf<uint32_t> goff, gids, gM;
    DevBuf<float> gdist;
    uint64_t edges = 0;
    // host mirror, refreshed lazily from the device copy
    bool host_valid = false;
    // per-node walk records of the current graph (rebuilt after any change)
    DevBuf<uint4> gnode;
    DevBuf<uint32_t> gflag;
    DevBuf<uint32_t> gpad, gmd;  // padded adjacency (kg_launch_gpad), gpad_stride ids per row
    unsigned gpad_stride = 0;    // 0: none (too large), walks read the CSR rows
    int search_adjacency = 0;    // 0: padded rows when they exist, 1: CSR rows
    int last_adjacency = 1;      // what the last walk launch read (0 padded, 1 CSR)
    bool gnode_valid = false;
    bool lists_unique = false;
    // walk layout: 1 = walks read copies of the rows, norms and padded
    // adjacency in the breadth-first node order of kg_walk_order (wperm[i] =
    // dataset row of node i, winv its inverse), 0 = dataset order
    int search_order = 1;
    bool worder_valid = false;   // wperm / winv match the graph; gpad / gmd are in that order
    unsigned worder_levels = 0;  // breadth-first levels of the order
    float worder_ms = 0.f;       // time to compute it and reorder the padded adjacency
    unsigned wforms = 0;         // row copies in that order: 1 FP32, 2 BF16, 4 8-bit, 8 norms
    int last_order = 0;          // whether the last walk launch read them
    // the lean walk (kg_set_search_lean): integer queries over the 8-bit rows
    // from caller-given starting points, launched beside the general walk
    int search_lean = 1;
    bool last_lean = false;      // the last walk was a lean / general pair
    unsigned last_lean_vcap = 0, last_lean_bits = 0;
    DevBuf<uint32_t> wperm, winv;
    DevBuf<float> wX, wnorms;
    DevBuf<uint16_t> wXb;
    DevBuf<uint8_t> wX8;
    std::vector<uint32_t> h_off, h_ids, h_M;
    std::vector<float> h_dist;

    // search scratch
    DevBuf<uint32_t> visited, touched, qcounter;
    DevBuf<uint32_t> seeds, seed_off, Ks, Ps, pidx, bsel, qorder;
    DevBuf<float> q, qn;
    DevBuf<uint64_t> bpart;  // brute-force partial top-K lists
    // device copies of host queries and results (kg_search, kg_bruteforce):
    // grown on demand and kept, so repeated calls allocate nothing
    DevBuf<float> io_q, io_qn, io_dists;
    DevBuf<uint32_t> io_ids, io_counts, io_comps, io_scans;
    unsigned vis_slots = 0, vis_words = 0;
    unsigned search_wpb = 4, search_bpc = 0;  // launch shape: waves per block (<= 4), blocks per CU (0 = as LDS allows)
    DevBuf<float> lat;      // per-query latency of the last search (microseconds)
    unsigned lat_n = 0;

    // entry-point sample
    unsigned entryE = 0;
    DevBuf<float> erows, enorms, escores, escale;
    DevBuf<uint4> etiles;  // the entry sample as MFMA operand tiles (kg_launch_entry_tiles)
    DevBuf<uint32_t> eids;

    // learned projector D -> mlpH -> mlpP (FP32 master weights, Adam state)
    unsigned mlpH = 0, mlpP = 0, mlp_version = 0, adam_step = 0;
    DevBuf<float> mlp_mu, mlp_is, mlp_xn;  // input standardisation, standardised rows
    DevBuf<float> W1, b1, W2, b2, gW1, gb1, gW2, gb2;
    DevBuf<float> mW1, vW1, mb1, vb1, mW2, vW2, mb2, vb2;
    DevBuf<float> mlp_x, mlp_h, mlp_z, mlp_dz, mlp_dh, mlp_loss;
    DevBuf<uint32_t> mlp_idx;
    // entry selection through the projector: shortlist size (0 = exact scoring)
    unsigned entry_shortlist = 0, eproj_version = ~0u;
    DevBuf<float> eproj, zq, hq;
    DevBuf<uint4> eproj_tiles;  // eproj as MFMA operand tiles
    DevBuf<uint32_t> eshort;

    // timing + build trace
    hipEvent_t ev[20] = {};
    float last_ms[9] = {};
    bool entry_ms_pending = false;  // last_ms[7] waits on ev[14..15] of a selection still in flight
    std::vector<float> trace;
    uint64_t build_counters[6] = {};  // pairs, list entries, candidate pushes, iterations, graph replays, re-scored pairs
    std::vector<uint64_t> iter_stats;  // per iteration: pairs, pushes, list entries, join form, join us, iteration us
    bool build_graphs = true;         // NN-descent iterations run as a captured HIP graph
    int join_form = KG_JOIN_AUTO;     // BF16 screen vs exact FP32 tiles on float rows (KG_JOIN_*)
    // learned pre-rank of the join lists: fraction kept (0 or >= 1: off), and
    // what the last build did: status 0 off, 1 on throughout, 2 turned down by
    // the held-out probe before the first iteration, 3 turned off when the
    // control recall slowed down (at iteration prerank_fallback); probe accuracy
    float prerank_keep = 0.f;
    int prerank_status = 0, prerank_fallback = -1;
    float prerank_acc = 0.f;

    Handle();
    ~Handle();
    void sync() { KG_HIP(hipStreamSynchronize(stream)); }
    void pull_graph();                    // device -> host mirror
    void push_graph();                    // host mirror -> device
};

// Defined in kg_mlp.cpp.
void mlp_init(Handle &h, unsigned hidden, unsigned out, unsigned seed);
// X already standardised (row stride `stride`).
void mlp_forward(Handle &h, float const *X, unsigned n, size_t stride, float *Hout, float *Zout);
// Raw rows: standardised with the projector's column mean / 1/std first.
void mlp_forward_rows(Handle &h, float const *X, unsigned n, size_t stride, float *Hout, float *Zout);
void mlp_train(Handle &h, unsigned steps, unsigned batch, float lr, float margin, unsigned seed, float *trace);
void mlp_loss_grad(Handle &h, uint32_t const *triplets, unsigned nt, float margin, float *loss, float *dW1,
                   float *db1, float *dW2, float *db2);
void mlp_project_entries(Handle &h);

// Defined in kg_build.cpp.
void build_index(Handle &h, kg_index_params const &params, kg_index_info *info);
void reverse_index(Handle &h, int rev);
void prune_index(Handle &h, unsigned level);
void screen_pairs(Handle &h, uint32_t const *pairs, unsigned n, float *screened, float *bound);

}  // namespace kga

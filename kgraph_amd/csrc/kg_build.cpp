// kg_build.cpp -- host driver of the GPU NN-descent build
// (KGraphConstructor, kgraph.cpp:886-999), reverse-edge completion
// (kgraph.cpp:667-702) and pruning (kgraph.cpp:584-665).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <random>

#include "kg_exact.hpp"
#include "kg_handle.hpp"
#include "kg_internal.hpp"
#include "kg_nnd.hpp"

namespace kga {

namespace {

unsigned next_pow2(unsigned x) {
    unsigned p = 1;
    while (p < x) p <<= 1;
    return p;
}

// EvaluateRecall / EvaluateAccuracy (kgraph.cpp:96-141) over the pool's
// valid entries; both lists ascending.
float eval_recall(float const *pool, unsigned np, float const *knn, unsigned nk) {
    if (nk == 0) return 1.0f;
    unsigned found = 0, p = 0, k = 0;
    while (p < np && k < nk) {
        if (knn[k] < pool[p])
            ++k;
        else if (knn[k] == pool[p]) {
            ++found;
            ++k;
            ++p;
        } else
            throw runtime_error("distance is unstable");
    }
    return float(found) / nk;
}

float eval_accuracy(float const *pool, unsigned np, float const *knn, unsigned nk) {
    unsigned m = std::min(np, nk), cnt = 0;
    float sum = 0;
    for (unsigned i = 0; i < m; ++i) {
        if (knn[i] > 0) {
            sum += std::abs(pool[i] - knn[i]) / knn[i];
            ++cnt;
        }
    }
    return cnt ? sum / cnt : 0;
}

struct NndState {
    DevBuf<uint64_t> pool, cand, ovf, stats;
    DevBuf<uint32_t> part, cnt, Mv, found, samp, nnew, nold, rnew, rold, crnew, crold, ccnt, csrc;
    DevBuf<float> radiusM;
    DevBuf<float4> meta;
    DevBuf<uint32_t> ctl, ccount;
    DevBuf<float> cdist;
    DevBuf<float2> sqn;
    DevBuf<uint32_t> flags, iter;
};

// Error-bound constants of the BF16 screen.  Rounding to BF16 (8 significant
// bits, round to nearest even) changes each operand by at most u = 2^-8
// relative, so a product of two rounded operands (exact in FP32) is off by at
// most (2u + u^2) |a_i b_i|; the MFMA's FP32 accumulation of D (+ zero
// padding to a multiple of 32) products adds at most gamma = (D + 32) 2^-23
// times the sum of their magnitudes.  With sum |a_i b_i| <= |a| |b|:
//   |G - a.b| <= c |a| |b|,  c = (2u + u^2 + gamma) * 1.02,
// the 2 % margin covering (1+u)^2 on the gamma term and the rounding of the
// bound itself.  g = 4 gamma bounds the relative error of the stored norms and
// of the few FP32 operations that turn G into a distance.
//
// FP32 operands (f32): no operand rounding; each product is off by at most
// 2^-24 relative and the accumulation adds gamma, so c = (2^-24 + gamma) * 1.02.
void screen_constants(unsigned D, NndArgs &a, bool f32 = false) {
    double const gam = (D + 32) * std::ldexp(1.0, -23);
    double const prod = f32 ? std::ldexp(1.0, -24) : std::ldexp(1.0, -7) + std::ldexp(1.0, -16);
    a.scr_c = float((prod + gam) * 1.02);
    a.scr_g = float(4 * gam);
}

// Row norms for the MFMA forms (into sqn) and the value range of the rows.
struct RowRange {
    bool integer;
    double maxabs, vmin, vmax;
};
RowRange row_range(Handle &h, NndArgs const &a, float2 *sqn, DevBuf<uint32_t> &flags, hipStream_t s) {
    flags.reserve(4);
    uint32_t const init[4] = {0u, 0u, 0xffffffffu, 0u};
    KG_HIP(hipMemcpyAsync(flags.p, init, sizeof(init), hipMemcpyHostToDevice, s));
    KG_HIP(nnd_launch_sqn(a, sqn, flags.p, s));
    uint32_t fl[4];
    KG_HIP(hipMemcpyAsync(fl, flags.p, sizeof(fl), hipMemcpyDeviceToHost, s));
    h.sync();
    auto unord = [](uint32_t k) {
        uint32_t const u = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
        float f;
        memcpy(&f, &u, 4);
        return double(f);
    };
    float mx;
    memcpy(&mx, &fl[1], 4);
    RowRange r;
    r.integer = (fl[0] & 1u) == 0;
    r.maxabs = mx;
    r.vmin = h.N ? unord(fl[2]) : 0.0;
    r.vmax = h.N ? unord(fl[3]) : 0.0;
    return r;
}

// One NN-descent iteration's device work after the first: update of the
// previous iteration's pools (M growth, new / old split, reverse pushes), then
// the join, the merge and the control-pool gather.  The first iteration runs
// without the update (its pools come from init).  The same launches either go
// straight to the stream or are captured once into a HIP graph and replayed.
void iteration_body(int metric, NndArgs const &a, bool update, uint32_t const *ctl, unsigned nc, float *cdist,
                    uint32_t *ccount, hipStream_t s) {
    if (update) KG_HIP(nnd_launch_update(a, s));
    KG_HIP(nnd_launch_iter_begin(a, update ? 1 : 0, s));
    if (a.Z) KG_HIP(nnd_launch_prerank(a, s));
    KG_HIP(nnd_launch_join(metric, a, s));
    KG_HIP(nnd_launch_merge(a, s));
    if (nc) KG_HIP(nnd_launch_gather_controls(a, ctl, nc, cdist, ccount, s));
}

// The projector's rows for every dataset row (N x mlpP), in chunks.
void project_rows(Handle &h, float *Z) {
    unsigned const chunk = 1u << 16;
    DevBuf<float> hid;
    hid.reserve(size_t(chunk) * h.mlpH);
    for (unsigned r0 = 0; r0 < h.N; r0 += chunk) {
        unsigned const n = std::min(chunk, h.N - r0);
        mlp_forward_rows(h, h.X.p + size_t(r0) * h.Dp, n, h.Dp, hid.p, Z + size_t(r0) * h.mlpP);
    }
    h.sync();
}

// Captured iteration graph, destroyed with the build.
struct IterGraph {
    hipGraph_t g = nullptr;
    hipGraphExec_t e = nullptr;
    ~IterGraph() {
        if (e) (void)hipGraphExecDestroy(e);
        if (g) (void)hipGraphDestroy(g);
    }
    void capture(int metric, NndArgs const &a, uint32_t const *ctl, unsigned nc, float *cdist, uint32_t *ccount,
                 hipStream_t s) {
        KG_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        try {
            iteration_body(metric, a, true, ctl, nc, cdist, ccount, s);
        } catch (...) {
            hipGraph_t dropped = nullptr;
            (void)hipStreamEndCapture(s, &dropped);
            if (dropped) (void)hipGraphDestroy(dropped);
            throw;
        }
        KG_HIP(hipStreamEndCapture(s, &g));
        KG_HIP(hipGraphInstantiate(&e, g, nullptr, nullptr, 0));
    }
};

void replace_graph(Handle &h, unsigned N, DevBuf<uint32_t> &off, DevBuf<uint32_t> &ids, DevBuf<float> &dists,
                   DevBuf<uint32_t> &M, uint64_t edges) {
    std::swap(h.goff.p, off.p);
    std::swap(h.goff.n, off.n);
    std::swap(h.gids.p, ids.p);
    std::swap(h.gids.n, ids.n);
    std::swap(h.gdist.p, dists.p);
    std::swap(h.gdist.n, dists.n);
    std::swap(h.gM.p, M.p);
    std::swap(h.gM.n, M.n);
    h.gN = N;
    h.edges = edges;
    h.has_graph = true;
    h.host_valid = false;
    h.gnode_valid = false;
}

}  // namespace

void build_index(Handle &h, kg_index_params const &params, kg_index_info *pinfo) {
    if (!h.has_data) throw runtime_error("no dataset");
    kg_index_params p = params;
    unsigned const N = h.N;
    // parameter clamps (kgraph.cpp:893-905)
    if (N <= p.K) throw runtime_error("K larger than dataset size");
    if (N < p.controls) p.controls = N;
    if (N <= p.L) p.L = N - 1;
    if (N <= p.S) p.S = N - 1;
    if (p.S == 0 || p.S > NND_MAX_S) throw invalid_argument("S must be in [1, 255]");
    if (p.L == 0) throw invalid_argument("L must be >= 1");
    // The S random starting entries fill the first S of a node's L pool slots
    // (the reference's pool holds L + 1, kgraph.cpp:766, 786).
    if (p.S > p.L) throw invalid_argument("S must be <= L");
    if (p.K > 1024) throw invalid_argument("K must be <= 1024");
    if (p.prune > 2) throw runtime_error("prune level not supported.");
    if (uint64_t(N) * p.L >= 0xffffffffull) throw invalid_argument("N * L too large");
    unsigned const R = p.R ? p.R : std::max(p.L, 2 * p.S);  // R = 0 (uncapped) gets a fixed cap
    hipStream_t s = h.stream;

    NndArgs a{};
    a.X = h.X.p;
    a.Xb = h.Xb.p;
    a.norms = h.norms.p;
    a.N = N;
    a.n0 = 0;
    a.n1 = N;
    a.D = h.D;
    a.Dp = h.Dp;
    a.L = p.L;
    a.K = p.K;
    a.S = p.S;
    a.R = R;
    a.Lc = std::max(p.L, 2 * p.S);
    a.nnew_init = std::min(2 * p.S, N);
    if (p.L > 768) throw invalid_argument("L must be <= 768");
    a.C = std::min(next_pow2(std::max({2 * p.L, 4 * p.S * p.S, 256u})), 1024u);
    a.Cp2 = a.C;
    a.Acap = (a.Lc + R + 3) & ~3u;
    a.Bcap = a.Acap;
    a.seed = p.seed;

    std::unique_ptr<NndState> stp(new NndState);
    NndState &st = *stp;
    st.pool.reserve(size_t(N) * p.L);
    st.cand.reserve(size_t(N) * a.C);
    st.csrc.reserve(size_t(N) * a.C);
    a.O = 64;
    st.ovf.reserve(size_t(N) * a.O);
    st.stats.reserve(NND_NSTATS);
    st.iter.reserve(1);
    st.part.reserve(size_t(NND_NPART) * N);
    for (auto *b : {&st.cnt, &st.Mv, &st.found, &st.nnew, &st.nold, &st.crnew, &st.crold, &st.ccnt}) b->reserve(N);
    st.samp.reserve(size_t(N) * a.Lc);
    st.rnew.reserve(size_t(N) * R);
    st.rold.reserve(size_t(N) * R);
    st.radiusM.reserve(N);
    st.meta.reserve(N);
    a.pool = st.pool.p;
    a.cand = st.cand.p;
    a.csrc = st.csrc.p;
    a.ovf = st.ovf.p;
    a.stats = st.stats.p;
    a.iter = st.iter.p;
    a.part = st.part.p;
    a.cnt = st.cnt.p;
    a.Mv = st.Mv.p;
    a.found = st.found.p;
    a.samp = st.samp.p;
    a.nnew = st.nnew.p;
    a.nold = st.nold.p;
    a.rnew = st.rnew.p;
    a.rold = st.rold.p;
    a.crnew = st.crnew.p;
    a.crold = st.crold.p;
    a.ccnt = st.ccnt.p;
    a.meta = st.meta.p;
    a.radiusM = st.radiusM.p;
    if (h.precision == KG_PRECISION_BF16 && h.D <= 256) {
        // local join as BF16 MFMA tiles; exact as it is on integer rows that
        // BF16 holds with small enough sums (bf16_tiles_exact), else a screen
        // with an exact FP32 re-score
        st.sqn.reserve(N);
        RowRange rr = row_range(h, a, st.sqn.p, st.flags, s);
        a.sqn = st.sqn.p;
        a.mfma = 1;
        a.exact = bf16_tiles_exact(h.D, h.metric, rr.integer, h.rows_bf16, rr.maxabs, rr.vmin, rr.vmax) ? 1u : 0u;
        // 8-bit operands when the rows' 8-bit copy is exact and the distance
        // allows it (i8_tiles_exact): exact tiles with half the operand bytes
        // and MFMA issues of the BF16 form
        if (h.join_i8 && i8_tiles_exact(h.D, h.metric, h.rows_u8)) {
            a.exact = 1;
            a.i8 = 1;
            a.X8 = h.X8.p;
        }
        screen_constants(h.D, a);
    } else if (h.precision == KG_PRECISION_FP32_MFMA && h.D <= 256) {
        // local join screened with FP32 operands on MFMA, exact FP32 re-score
        st.sqn.reserve(N);
        (void)row_range(h, a, st.sqn.p, st.flags, s);
        a.sqn = st.sqn.p;
        a.mfma = 1;
        a.f32 = 1;
        a.exact = 0;
        screen_constants(h.D, a, true);
    }
    h.join_exact = a.mfma ? (a.f32 ? 2 : a.i8 ? 3 : int(a.exact)) : -1;
    // learned pre-rank of the join lists: the projector's rows, guarded by a
    // held-out probe before the first iteration and by the control recall
    // after every iteration (plain NN-descent from the first drop on)
    DevBuf<float> Z;
    bool prerank = false;
    h.prerank_status = 0;
    h.prerank_acc = 0.f;
    h.prerank_fallback = -1;
    if (h.prerank_keep > 0.f && h.prerank_keep < 1.f) {
        if (!h.mlpH) throw invalid_argument("join pre-rank needs a projector (kg_mlp_init)");
        Z.reserve(size_t(N) * h.mlpP);
        project_rows(h, Z.p);
        prerank = true;
    }
    KG_HIP(hipMemsetAsync(st.ovf.p, 0xff, size_t(N) * a.O * 8, s));
    KG_HIP(hipMemsetAsync(st.found.p, 0, size_t(N) * 4, s));
    KG_HIP(hipMemsetAsync(st.iter.p, 0, 4, s));

    KG_HIP(hipEventRecord(h.ev[4], s));
    // controls: exact K-NN of C random nodes (GenerateControl, kgraph.cpp:254-270)
    unsigned const nc = p.controls;
    std::vector<float> ctl_knn(size_t(nc) * p.K);
    if (nc) {
        std::vector<uint32_t> index(N);
        for (unsigned i = 0; i < N; ++i) index[i] = i;
        std::mt19937 crng(p.seed);
        std::shuffle(index.begin(), index.end(), crng);
        index.resize(nc);
        st.ctl.reserve(nc);
        st.ccount.reserve(nc);
        st.cdist.reserve(size_t(nc) * std::max(p.L, p.K));
        KG_HIP(hipMemcpyAsync(st.ctl.p, index.data(), 4ull * nc, hipMemcpyHostToDevice, s));
        DevBuf<uint32_t> kid;
        kid.reserve(size_t(nc) * p.K);
        kg_brute_args b{};
        b.X = h.X.p;
        b.norms = h.norms.p;
        b.N = N;
        b.D = h.D;
        b.Dp = h.Dp;
        b.rows = st.ctl.p;
        b.nq = nc;
        b.K = p.K;
        b.epsilon = __builtin_huge_valf();
        b.out_ids = kid.p;
        b.out_dists = st.cdist.p;
        b.ostride = p.K;
        DevBuf<uint64_t> part;  // rows split across blocks, then merged
        part.reserve(size_t(nc) * kg_brute_splits(nc, N, p.K) * p.K);
        b.partial = part.p;
        KG_HIP(kg_launch_brute(h.metric, b, s));
        KG_HIP(hipMemcpyAsync(ctl_knn.data(), st.cdist.p, 4 * ctl_knn.size(), hipMemcpyDeviceToHost, s));
        h.sync();
        if (prerank) {
            // held-out probe of the projector: for each control node, its
            // exact nearest neighbours against random nodes -- the share the
            // projection orders correctly
            std::vector<uint32_t> knn_ids(size_t(nc) * p.K), pairs;
            KG_HIP(hipMemcpy(knn_ids.data(), kid.p, 4 * knn_ids.size(), hipMemcpyDeviceToHost));
            std::mt19937 prng(p.seed + 7);
            unsigned const per = std::min(p.K, 10u);
            for (unsigned c = 0; c < nc; ++c)
                for (unsigned t = 0; t < per; ++t) {
                    uint32_t r = prng() % N;
                    if (r == index[c]) r = (r + 1) % N;
                    pairs.insert(pairs.end(), {index[c], knn_ids[size_t(c) * p.K + t], index[c], r});
                }
            unsigned const np = unsigned(pairs.size() / 2);
            DevBuf<uint32_t> dpairs;
            DevBuf<float> dz;
            dpairs.reserve(pairs.size());
            dz.reserve(np);
            KG_HIP(hipMemcpyAsync(dpairs.p, pairs.data(), 4 * pairs.size(), hipMemcpyHostToDevice, s));
            KG_HIP(nnd_launch_zpair(Z.p, h.mlpP, dpairs.p, np, dz.p, s));
            std::vector<float> hz(np);
            KG_HIP(hipMemcpyAsync(hz.data(), dz.p, 4ull * np, hipMemcpyDeviceToHost, s));
            h.sync();
            unsigned good = 0;
            for (unsigned i = 0; i + 1 < np; i += 2) good += hz[i] < hz[i + 1];
            h.prerank_acc = np ? float(2.0 * good / np) : 0.f;
            if (h.prerank_acc < 0.9f) {
                prerank = false;  // the projection does not rank near neighbours first: plain NN-descent
                h.prerank_status = 2;
            }
        }
    }
    if (prerank && nc == 0) {  // no held-out probe set to guard the pre-rank with
        prerank = false;
        h.prerank_status = 2;
    }
    if (prerank) {
        a.Z = Z.p;
        a.zdim = h.mlpP;
        a.keep = h.prerank_keep;
        h.prerank_status = 1;
    }
    KG_HIP(nnd_launch_meta_init(h.metric, a, s));
    if (a.i8) KG_HIP(nnd_launch_n8(a, h.metric == KG_METRIC_EUCLIDEAN ? 1 : 0, s));  // the 8-bit form's join norms
    KG_HIP(nnd_launch_init(h.metric, a, s));

    kg_index_info info{};
    info.stop_condition = 0;
    info.recall = 0;
    info.accuracy = 3.4028235e38f;
    info.cost = 0;
    info.iterations = 0;
    info.delta = 1.0f;
    float const total = N * float(N - 1) / 2;
    uint64_t n_comps = 0, n_lists = 0, n_push = 0, n_rescore = 0;
    std::vector<float> cpool(size_t(nc) * p.L);
    std::vector<uint32_t> ccnt(nc);
    h.trace.clear();
    h.iter_stats.clear();
    float join_ms = 0, merge_ms = 0, upd_ms = 0, iter_ms = 0;
    uint64_t replays = 0;
    // Join form per iteration.  On float rows the BF16 join is a screen with an
    // exact FP32 re-score, and whether it beats the exact FP32 tiles depends on
    // the data (how many pairs land within the screen's error bound of a
    // radius) and drifts as the lists converge.  Both forms admit exactly the
    // same candidates, so in auto mode a probe iteration (every third, from the
    // first; the gap grows fourfold, up to 24, after a probe the winner takes
    // by more than 1.5x, and returns to 3 after a close one) splits the nodes between
    // the two forms, times both parts, and the next iterations use the screen
    // when it is at least 5 % faster per node
    // (its cost per node grows as the lists converge: on the 1.2M x 96 angular
    // set it beats the tiles by a third in the first iterations, ties at the
    // fourth and is 15-35 % slower after, so a tie goes to the tiles).  Every
    // probe splits the nodes in halves: a launch over an eighth of them pays a
    // larger share of fixed and tail costs per node, which tilted the
    // comparison (on the angular set it kept the screen through a tie).
    NndArgs a_tile = a;
    a_tile.mfma = 0;
    bool const screen = a.mfma && !a.exact;
    int const form = screen ? h.join_form : KG_JOIN_AUTO;
    bool use_mfma = a.mfma && form != KG_JOIN_TILES;
    unsigned const half = N / 2, eighth = N / 8;
    IterGraph graph[2][2];  // [pre-rank][0 FP32 tiles, 1 MFMA form]
    float prev_recall = -1.f;
    unsigned next_probe = 0, probe_gap = 3;
    for (unsigned it = 0; (p.iterations == 0) || (it < p.iterations); ++it) {
        ++info.iterations;
        float ms;
        bool const probe = screen && form == KG_JOIN_AUTO && it == next_probe && eighth > 0;
        // probe split: [0, cut) runs form `lead`, [cut, N) the other form
        bool const lead = it == 0 ? true : use_mfma;
        unsigned const cut = half;
        if (!a.Z) a_tile.Z = nullptr;  // after a pre-rank fallback
        NndArgs const &ar = use_mfma ? a : a_tile;
        int rec_form = probe ? 3 : (ar.mfma ? (ar.f32 ? 4 : (ar.exact ? 2 : 1)) : 0);
        float jms = 0.f;
        if (h.build_graphs && !probe && it > 0) {
            IterGraph &gr = graph[a.Z ? 1 : 0][use_mfma ? 1 : 0];
            if (!gr.e) gr.capture(h.metric, ar, st.ctl.p, nc, st.cdist.p, st.ccount.p, s);
            KG_HIP(hipEventRecord(h.ev[16], s));
            KG_HIP(hipGraphLaunch(gr.e, s));
            ++replays;
            KG_HIP(hipEventRecord(h.ev[17], s));
        } else {
            // the launches one by one, with per-kernel timing
            KG_HIP(hipEventRecord(h.ev[16], s));
            if (it > 0) {
                KG_HIP(hipEventRecord(h.ev[9], s));
                KG_HIP(nnd_launch_update(a, s));
                KG_HIP(hipEventRecord(h.ev[10], s));
            }
            KG_HIP(nnd_launch_iter_begin(a, it > 0 ? 1 : 0, s));
            if (a.Z) KG_HIP(nnd_launch_prerank(a, s));
            KG_HIP(hipEventRecord(h.ev[6], s));
            if (probe) {
                NndArgs first = lead ? a : a_tile, second = lead ? a_tile : a;
                first.n1 = cut;
                second.n0 = cut;
                KG_HIP(nnd_launch_join(h.metric, first, s));
                KG_HIP(hipEventRecord(h.ev[18], s));
                KG_HIP(nnd_launch_join(h.metric, second, s));
            } else {
                KG_HIP(nnd_launch_join(h.metric, ar, s));
            }
            KG_HIP(hipEventRecord(h.ev[7], s));
            KG_HIP(nnd_launch_merge(a, s));
            KG_HIP(hipEventRecord(h.ev[8], s));
            if (nc) KG_HIP(nnd_launch_gather_controls(a, st.ctl.p, nc, st.cdist.p, st.ccount.p, s));
            KG_HIP(hipEventRecord(h.ev[17], s));
        }
        uint64_t stats[NND_NSTATS];
        KG_HIP(hipMemcpyAsync(stats, st.stats.p, sizeof(stats), hipMemcpyDeviceToHost, s));
        if (nc) {
            KG_HIP(hipMemcpyAsync(cpool.data(), st.cdist.p, 4 * cpool.size(), hipMemcpyDeviceToHost, s));
            KG_HIP(hipMemcpyAsync(ccnt.data(), st.ccount.p, 4ull * nc, hipMemcpyDeviceToHost, s));
        }
        h.sync();
        KG_HIP(hipEventElapsedTime(&ms, h.ev[16], h.ev[17]));
        iter_ms += ms;
        float const it_ms = ms;
        bool const eager = !(h.build_graphs && !probe && it > 0);
        if (eager) {
            KG_HIP(hipEventElapsedTime(&jms, h.ev[6], h.ev[7]));
            join_ms += jms;
            KG_HIP(hipEventElapsedTime(&ms, h.ev[7], h.ev[8]));
            merge_ms += ms;
            if (it > 0) {
                KG_HIP(hipEventElapsedTime(&ms, h.ev[9], h.ev[10]));
                upd_ms += ms;
            }
        }
        if (probe) {
            float t_first, t_second;
            KG_HIP(hipEventElapsedTime(&t_first, h.ev[6], h.ev[18]));
            KG_HIP(hipEventElapsedTime(&t_second, h.ev[18], h.ev[7]));
            double const c_first = double(t_first) / cut, c_second = double(t_second) / (N - cut);  // per node
            double const c_screen = lead ? c_first : c_second, c_tile = lead ? c_second : c_first;
            use_mfma = c_screen < 0.95 * c_tile;
            double const margin = std::max(c_screen, c_tile) / std::max(std::min(c_screen, c_tile), 1e-30);
            probe_gap = margin > 1.5 ? std::min(probe_gap * 4, 24u) : 3u;
            next_probe = it + probe_gap;
        }
        // statistics and stop rule (kgraph.cpp:927-972)
        n_comps += stats[NND_STAT_PAIRS];
        n_lists += stats[NND_STAT_LISTS];
        n_push += stats[NND_STAT_PUSHES];
        n_rescore += stats[NND_STAT_RESCORE];
        info.cost = float(n_comps) / total;
        info.delta = float(double(stats[NND_STAT_DELTA]) / (double(N) * p.K));
        info.M = float(double(stats[NND_STAT_M]) / N);
        double sr = 0, sa = 0;
        for (unsigned c = 0; c < nc; ++c) {
            sr += eval_recall(&cpool[size_t(c) * p.L], ccnt[c], &ctl_knn[size_t(c) * p.K], p.K);
            sa += eval_accuracy(&cpool[size_t(c) * p.L], ccnt[c], &ctl_knn[size_t(c) * p.K], p.K);
        }
        info.recall = nc ? float(sr / nc) : 0.f;
        info.accuracy = nc ? float(sa / nc) : 0.f;
        if (a.Z && it > 0 && info.recall < p.recall &&
            info.recall - prev_recall < 0.25f * (1.f - std::max(prev_recall, 0.f))) {
            // under the pre-rank the probe set's recall closed less than a
            // quarter of its remaining gap in this iteration: plain
            // NN-descent from here
            a.Z = nullptr;
            h.prerank_status = 3;
            h.prerank_fallback = int(it);
        }
        prev_recall = info.recall;
        h.trace.insert(h.trace.end(), {info.recall, info.delta, info.M, info.cost});
        h.iter_stats.insert(h.iter_stats.end(), {stats[NND_STAT_PAIRS], stats[NND_STAT_PUSHES], stats[NND_STAT_LISTS],
                                                 uint64_t(rec_form), uint64_t(jms * 1000.f), uint64_t(it_ms * 1000.f)});
        if (verbosity > 0) {
            fprintf(stderr, "iteration: %u recall: %g accuracy: %g cost: %g M: %g delta: %g\n", info.iterations,
                    info.recall, info.accuracy, info.cost, info.M, info.delta);
        }
        if (info.delta <= p.delta) {
            info.stop_condition = 1;
            break;
        }
        if (info.recall >= p.recall) {
            info.stop_condition = 2;
            break;
        }
    }
    if (info.stop_condition == 0) {
        // iteration limit: the reference still runs update() after the last
        // join (kgraph.cpp:927-972), which grows M before the extraction
        float ms;
        KG_HIP(hipEventRecord(h.ev[9], s));
        KG_HIP(nnd_launch_update(a, s));
        KG_HIP(hipEventRecord(h.ev[10], s));
        KG_HIP(hipEventSynchronize(h.ev[10]));
        KG_HIP(hipEventElapsedTime(&ms, h.ev[9], h.ev[10]));
        upd_ms += ms;
        iter_ms += ms;
    }
    // graph extraction (kgraph.cpp:975-988): the valid pool prefix of each node
    DevBuf<uint32_t> off, ids, M, tmp;
    DevBuf<float> dists;
    off.reserve(N + 1);
    tmp.reserve((N + 1023) / 1024 + 2);
    KG_HIP(kg_scan_u32(st.cnt.p, off.p, N, tmp.p, s));
    uint32_t E = 0;
    KG_HIP(hipMemcpyAsync(&E, off.p + N, 4, hipMemcpyDeviceToHost, s));
    h.sync();
    ids.reserve(std::max<uint32_t>(E, 1));
    dists.reserve(std::max<uint32_t>(E, 1));
    M.reserve(N);
    KG_HIP(nnd_launch_extract(a, off.p, ids.p, dists.p, M.p, s));
    h.sync();
    // release the build state before reverse allocates its own
    stp.reset();
    replace_graph(h, N, off, ids, dists, M, E);
    h.no_dist = false;
    KG_HIP(hipEventRecord(h.ev[5], s));
    h.sync();
    h.build_counters[0] = n_comps;
    h.build_counters[1] = n_lists;
    h.build_counters[2] = n_push;
    h.build_counters[3] = info.iterations;
    h.build_counters[4] = replays;
    h.build_counters[5] = n_rescore;
    h.last_ms[3] = join_ms;
    h.last_ms[4] = merge_ms;
    h.last_ms[5] = upd_ms;
    h.last_ms[8] = iter_ms;
    if (p.reverse) reverse_index(h, p.reverse);
    if (p.prune) prune_index(h, p.prune);
    KG_HIP(hipEventRecord(h.ev[11], s));
    h.sync();
    KG_HIP(hipEventElapsedTime(&h.last_ms[2], h.ev[4], h.ev[11]));
    if (pinfo) *pinfo = info;
}

void reverse_index(Handle &h, int rev) {
    if (rev == 0) return;
    if (!h.has_graph) throw runtime_error("no index");
    if (h.no_dist) throw runtime_error("Need distance information to reverse graph");
    if (verbosity > 0) fprintf(stderr, "Graph completion with reverse edges...\n");
    hipStream_t s = h.stream;
    unsigned const N = h.gN;
    KG_HIP(hipEventRecord(h.ev[12], s));
    GraphDev g{N, h.goff.p, h.gids.p, h.gdist.p, h.gM.p};
    DevBuf<uint32_t> deg2, off2, cursor, tmp, cnt2, off3;
    deg2.reserve(N + 1);
    off2.reserve(N + 1);
    cursor.reserve(N + 1);
    tmp.reserve((N + 1023) / 1024 + 2);
    KG_HIP(hipMemsetAsync(deg2.p, 0, 4ull * (N + 1), s));
    KG_HIP(kg_reverse_count(g, rev, deg2.p, s));
    KG_HIP(kg_scan_u32(deg2.p, off2.p, N, tmp.p, s));
    std::vector<uint32_t> hoff(N + 1);
    KG_HIP(hipMemcpyAsync(hoff.data(), off2.p, 4ull * (N + 1), hipMemcpyDeviceToHost, s));
    h.sync();
    uint64_t E2 = hoff[N];
    if (E2 >= 0xffffffffull) throw runtime_error("index too large");
    DevBuf<uint64_t> keys;
    keys.reserve(std::max<uint64_t>(E2, 1));
    KG_HIP(kg_reverse_scatter(g, rev, off2.p, cursor.p, keys.p, s));
    // segments longer than the wave sort handles
    std::vector<uint32_t> large;
    std::vector<uint64_t> soff;
    uint64_t scratch = 0;
    for (unsigned i = 0; i < N; ++i) {
        unsigned n = hoff[i + 1] - hoff[i];
        if (n > 1024) {
            large.push_back(i);
            soff.push_back(scratch);
            scratch += next_pow2(n);
        }
    }
    DevBuf<uint32_t> dlarge;
    DevBuf<uint64_t> dsoff, dscratch;
    if (!large.empty()) {
        dlarge.reserve(large.size());
        dsoff.reserve(soff.size());
        dscratch.reserve(scratch);
        KG_HIP(hipMemcpyAsync(dlarge.p, large.data(), 4 * large.size(), hipMemcpyHostToDevice, s));
        KG_HIP(hipMemcpyAsync(dsoff.p, soff.data(), 8 * soff.size(), hipMemcpyHostToDevice, s));
    }
    cnt2.reserve(N + 1);
    KG_HIP(kg_segsort_unique(N, off2.p, keys.p, cnt2.p, dlarge.p, unsigned(large.size()), dscratch.p, dsoff.p, s));
    off3.reserve(N + 1);
    KG_HIP(kg_scan_u32(cnt2.p, off3.p, N, tmp.p, s));
    uint32_t E3 = 0;
    KG_HIP(hipMemcpyAsync(&E3, off3.p + N, 4, hipMemcpyDeviceToHost, s));
    h.sync();
    DevBuf<uint32_t> ids3, M3;
    DevBuf<float> d3;
    ids3.reserve(std::max<uint32_t>(E3, 1));
    d3.reserve(std::max<uint32_t>(E3, 1));
    M3.reserve(N);
    KG_HIP(kg_compact_keys(N, off2.p, keys.p, off3.p, ids3.p, d3.p, M3.p, s));
    KG_HIP(hipEventRecord(h.ev[13], s));
    h.sync();
    replace_graph(h, N, off3, ids3, d3, M3, E3);
    KG_HIP(hipEventElapsedTime(&h.last_ms[6], h.ev[12], h.ev[13]));
}

namespace {

// prune2 (kgraph.cpp:593-657): levels l = 0 .. max(min(M, deg)) - 1, each a
// screen pass and resolve rounds on the GPU (kg_graph.hip); the kept lists then
// replace the graph and M[i] becomes their length (the reference's closing
// prune1).  M[i] above the list length is clamped (the reference requires
// M[i] <= size, kgraph.cpp:616).
void prune2_index(Handle &h) {
    hipStream_t s = h.stream;
    unsigned const N = h.gN;
    std::vector<uint32_t> off(N + 1), M(N);
    KG_HIP(hipMemcpy(off.data(), h.goff.p, 4ull * (N + 1), hipMemcpyDeviceToHost));
    KG_HIP(hipMemcpy(M.data(), h.gM.p, 4ull * N, hipMemcpyDeviceToHost));
    unsigned maxM = 0;
    for (unsigned i = 0; i < N; ++i) {
        M[i] = std::min(M[i], off[i + 1] - off[i]);
        maxM = std::max(maxM, M[i]);
    }
    // nodes by min(M, deg) descending (counting sort): the nodes still active
    // at level l are the first active[l] of them
    std::vector<uint32_t> active(maxM + 2, 0), order(N);
    for (unsigned i = 0; i < N; ++i) ++active[M[i]];
    {
        uint32_t run = 0;
        for (unsigned m = maxM + 1; m-- > 0;) {
            uint32_t c = active[m];
            active[m] = run;  // nodes with Meff > m
            run += c;
        }
        std::vector<uint32_t> pos(active);
        for (unsigned i = 0; i < N; ++i) order[pos[M[i]]++] = i;
    }
    DevBuf<uint32_t> Meff, ord, kl, nl, st, cnt;
    DevBuf<float> kd;
    size_t const E = std::max<size_t>(off[N], 1);
    Meff.reserve(N);
    ord.reserve(N);
    nl.reserve(N);
    st.reserve(N);
    cnt.reserve(1);
    kl.reserve(E);
    kd.reserve(E);
    KG_HIP(hipMemcpyAsync(Meff.p, M.data(), 4ull * N, hipMemcpyHostToDevice, s));
    KG_HIP(hipMemcpyAsync(ord.p, order.data(), 4ull * N, hipMemcpyHostToDevice, s));
    KG_HIP(hipMemsetAsync(nl.p, 0, 4ull * N, s));
    Prune2Dev p{N, h.goff.p, h.gids.p, h.no_dist ? nullptr : h.gdist.p, Meff.p, ord.p, kl.p, kd.p, nl.p, st.p};
    for (unsigned l = 0; l < maxM; ++l) {
        unsigned const nact = active[l];
        KG_HIP(kg_prune2_screen(p, l, nact, s));
        KG_HIP(kg_prune2_resolve(p, l, nact, cnt.p, s));
        for (;;) {
            uint32_t open = 0;
            KG_HIP(hipMemsetAsync(cnt.p, 0, 4, s));
            KG_HIP(kg_prune2_resolve(p, l, nact, cnt.p, s));
            KG_HIP(hipMemcpyAsync(&open, cnt.p, 4, hipMemcpyDeviceToHost, s));
            h.sync();
            if (!open) break;
        }
    }
    DevBuf<uint32_t> off2, tmp, ids2, M2;
    DevBuf<float> d2;
    off2.reserve(N + 1);
    tmp.reserve((N + 1023) / 1024 + 2);
    KG_HIP(kg_scan_u32(nl.p, off2.p, N, tmp.p, s));
    uint32_t E2 = 0;
    KG_HIP(hipMemcpyAsync(&E2, off2.p + N, 4, hipMemcpyDeviceToHost, s));
    h.sync();
    ids2.reserve(std::max<uint32_t>(E2, 1));
    d2.reserve(std::max<uint32_t>(E2, 1));
    M2.reserve(N);
    KG_HIP(kg_prune2_copy(p, off2.p, ids2.p, d2.p, s));
    KG_HIP(hipMemcpyAsync(M2.p, nl.p, 4ull * N, hipMemcpyDeviceToDevice, s));
    h.sync();
    replace_graph(h, N, off2, ids2, d2, M2, E2);
}

void prune1_index(Handle &h) {
    // prune1 (kgraph.cpp:584-590): keep the first M[i] edges
    hipStream_t s = h.stream;
    unsigned const N = h.gN;
    GraphDev g{N, h.goff.p, h.gids.p, h.gdist.p, h.gM.p};
    DevBuf<uint32_t> cnt, off2, tmp, ids2;
    DevBuf<float> d2;
    cnt.reserve(N + 1);
    off2.reserve(N + 1);
    tmp.reserve((N + 1023) / 1024 + 2);
    KG_HIP(kg_prune1_count(g, cnt.p, s));
    KG_HIP(kg_scan_u32(cnt.p, off2.p, N, tmp.p, s));
    uint32_t E = 0;
    KG_HIP(hipMemcpyAsync(&E, off2.p + N, 4, hipMemcpyDeviceToHost, s));
    h.sync();
    ids2.reserve(std::max<uint32_t>(E, 1));
    d2.reserve(std::max<uint32_t>(E, 1));
    KG_HIP(kg_prune1_copy(g, off2.p, ids2.p, h.no_dist ? nullptr : d2.p, s));
    h.sync();
    DevBuf<uint32_t> M;  // M is unchanged by prune1
    M.reserve(N);
    KG_HIP(hipMemcpyAsync(M.p, h.gM.p, 4ull * N, hipMemcpyDeviceToDevice, s));
    h.sync();
    replace_graph(h, N, off2, ids2, d2, M, E);
}

}  // namespace

void prune_index(Handle &h, unsigned level) {
    if (!h.has_graph) throw runtime_error("no index");
    if (level & 1) prune1_index(h);
    if (level & 2) prune2_index(h);
}

void screen_pairs(Handle &h, uint32_t const *pairs, unsigned n, float *screened, float *bound) {
    if (!h.has_data) throw runtime_error("no dataset");
    if (h.D > 256) throw invalid_argument("BF16 tiles need D <= 256");
    if (n && (!pairs || !screened || !bound)) throw invalid_argument("null argument");
    for (size_t i = 0; i < 2ull * n; ++i)
        if (pairs[i] >= h.N) throw invalid_argument("pair id out of range");
    if (n == 0) return;
    hipStream_t s = h.stream;
    NndArgs a{};
    a.X = h.X.p;
    a.Xb = h.Xb.p;
    a.norms = h.norms.p;
    a.N = h.N;
    a.D = h.D;
    a.Dp = h.Dp;
    DevBuf<float2> sqn;
    DevBuf<uint32_t> flags, dp;
    DevBuf<float> dd, de;
    sqn.reserve(h.N);
    dp.reserve(2ull * n);
    dd.reserve(n);
    de.reserve(n);
    (void)row_range(h, a, sqn.p, flags, s);
    a.sqn = sqn.p;
    screen_constants(h.D, a);
    KG_HIP(hipMemcpyAsync(dp.p, pairs, 8ull * n, hipMemcpyHostToDevice, s));
    KG_HIP(nnd_launch_screen_pairs(h.metric, a, dp.p, n, dd.p, de.p, s));
    KG_HIP(hipMemcpyAsync(screened, dd.p, 4ull * n, hipMemcpyDeviceToHost, s));
    KG_HIP(hipMemcpyAsync(bound, de.p, 4ull * n, hipMemcpyDeviceToHost, s));
    h.sync();
}

}  // namespace kga

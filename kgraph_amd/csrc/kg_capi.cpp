// kg_capi.cpp -- the extern "C" boundary of libkgraph_amd.so (declared in
// include/kgraph_amd.h) and the host logic around the HIP kernels: dataset
// upload, search-seed generation, scratch sizing, the KNNGRAPH file format
// and index exchange.
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <ctime>
#include <fstream>
#include <random>

#include "kg_handle.hpp"
#include "kg_internal.hpp"
#include "kg_noalloc.hpp"

namespace kga {

unsigned verbosity = 1;  // kgraph.h:42, default_verbosity kgraph.h:27

Handle::Handle() {
    KG_HIP(hipGetDevice(&device));
    hipDeviceProp_t prop;
    KG_HIP(hipGetDeviceProperties(&prop, device));
    num_cus = prop.multiProcessorCount;
    lds_per_cu = prop.maxSharedMemoryPerMultiProcessor ? prop.maxSharedMemoryPerMultiProcessor : 65536;
    KG_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    for (auto &e : ev) KG_HIP(hipEventCreate(&e));
    KG_HIP(hipHostMalloc(reinterpret_cast<void **>(&host_flag), 16, hipHostMallocDefault));
    noalloc_setup(device);  // the no-allocation scan's scratch address on this device (once per device)
}

Handle::~Handle() {
    for (auto &e : ev)
        if (e) (void)hipEventDestroy(e);
    if (stream) (void)hipStreamDestroy(stream);
    if (host_flag) (void)hipHostFree(host_flag);
}

void Handle::pull_graph() {
    if (host_valid) return;
    if (!has_graph) throw runtime_error("no index");
    sync();
    h_off.resize(gN + 1);
    h_ids.resize(edges);
    h_M.resize(gN);
    h_dist.resize(no_dist ? 0 : edges);
    KG_HIP(hipMemcpy(h_off.data(), goff.p, 4ull * (gN + 1), hipMemcpyDeviceToHost));
    if (edges) KG_HIP(hipMemcpy(h_ids.data(), gids.p, 4ull * edges, hipMemcpyDeviceToHost));
    KG_HIP(hipMemcpy(h_M.data(), gM.p, 4ull * gN, hipMemcpyDeviceToHost));
    if (!no_dist && edges) KG_HIP(hipMemcpy(h_dist.data(), gdist.p, 4ull * edges, hipMemcpyDeviceToHost));
    host_valid = true;
}

void Handle::push_graph() {
    edges = h_off.empty() ? 0 : h_off.back();
    gN = h_M.size();
    goff.reserve(gN + 1);
    gids.reserve(std::max<uint64_t>(edges, 1));
    gM.reserve(std::max<unsigned>(gN, 1));
    gdist.reserve(std::max<uint64_t>(edges, 1));
    KG_HIP(hipMemcpy(goff.p, h_off.data(), 4ull * (gN + 1), hipMemcpyHostToDevice));
    if (edges) KG_HIP(hipMemcpy(gids.p, h_ids.data(), 4ull * edges, hipMemcpyHostToDevice));
    if (gN) KG_HIP(hipMemcpy(gM.p, h_M.data(), 4ull * gN, hipMemcpyHostToDevice));
    if (!no_dist && edges) KG_HIP(hipMemcpy(gdist.p, h_dist.data(), 4ull * edges, hipMemcpyHostToDevice));
    has_graph = true;
    host_valid = true;
    gnode_valid = false;
}

// GenRandom (kgraph.cpp:44-65): `size` distinct ids < N from an mt19937.
static void gen_random(std::mt19937 &rng, uint32_t *addr, unsigned size, unsigned N) {
    if (N == size) {
        for (unsigned i = 0; i < size; ++i) addr[i] = i;
        return;
    }
    for (unsigned i = 0; i < size; ++i) addr[i] = rng() % (N - size);
    std::sort(addr, addr + size);
    for (unsigned i = 1; i < size; ++i)
        if (addr[i] <= addr[i - 1]) addr[i] = addr[i - 1] + 1;
    unsigned off = rng() % N;
    for (unsigned i = 0; i < size; ++i) addr[i] = (addr[i] + off) % N;
}

static void upload_rows(Handle &h, float const *src, bool src_device, unsigned N, unsigned D, size_t stride,
                        DevBuf<float> &dst) {
    unsigned Dp = (D + 3) & ~3u;
    dst.reserve(std::max<size_t>(size_t(N) * Dp, 4));
    if (N == 0) return;
    KG_HIP(hipMemsetAsync(dst.p, 0, size_t(N) * Dp * 4, h.stream));
    KG_HIP(hipMemcpy2DAsync(dst.p, Dp * 4, src, stride * 4, size_t(D) * 4, N,
                            src_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, h.stream));
}

static void set_data(Handle &h, float const *X, bool dev, unsigned N, unsigned D, size_t stride, int metric) {
    if (!X && N) throw invalid_argument("bad array");
    if (D == 0 || D > 1024) throw invalid_argument("dimension must be in [1, 1024]");
    if (stride < D) throw invalid_argument("bad stride");
    if (metric < KG_METRIC_L2SQR || metric > KG_METRIC_ANGULAR) throw runtime_error("metric not supported");
    if (N >= 0x7fffffffu) throw invalid_argument("dataset too large");
    upload_rows(h, X, dev, N, D, stride, h.X);
    h.N = N;
    h.D = D;
    h.Dp = (D + 3) & ~3u;
    h.metric = metric;
    h.norms.reserve(std::max<unsigned>(N, 1));
    KG_HIP(kg_launch_norms(metric, h.X.p, N, D, h.Dp, h.norms.p, h.stream));
    h.Xb.reserve(std::max<size_t>(size_t(N) * h.Dp, 4));
    DevBuf<uint32_t> inexact;
    inexact.reserve(1);
    KG_HIP(hipMemsetAsync(inexact.p, 0, 4, h.stream));
    KG_HIP(kg_launch_bf16_rows(h.X.p, N, h.Dp, h.Xb.p, inexact.p, h.stream));
    uint32_t lossy = 1;
    KG_HIP(hipMemcpyAsync(&lossy, inexact.p, 4, hipMemcpyDeviceToHost, h.stream));
    h.X8.reserve(std::max<size_t>(size_t(N) * h.Dp, 4));
    DevBuf<uint32_t> inexact8;
    inexact8.reserve(1);
    KG_HIP(hipMemsetAsync(inexact8.p, 0, 4, h.stream));
    KG_HIP(kg_launch_u8_rows(h.X.p, N, h.Dp, h.X8.p, inexact8.p, h.stream));
    uint32_t lossy8 = 1;
    KG_HIP(hipMemcpyAsync(&lossy8, inexact8.p, 4, hipMemcpyDeviceToHost, h.stream));
    h.sync();
    h.rows_bf16 = lossy == 0;
    h.rows_u8 = lossy8 == 0;
    if (!h.rows_u8) h.X8.release();  // kept only when it is exact
    h.has_data = true;
    h.wforms = 0;  // the walk's reordered row copies are stale
}

// Rows a walk gathers: 2 the 8-bit copy, 1 the BF16 copy, 0 the FP32 rows.
static int walk_rows(Handle const &h) {
    if (!h.has_data || h.row_storage == 1) return 0;
    if ((h.row_storage == 0 || h.row_storage == 3) && h.rows_u8) return 2;
    return h.rows_bf16 ? 1 : 0;
}

// The walk layout (kg_layout.hip): the breadth-first node order of the
// padded adjacency, which is then replaced by its copy in that order.
static void walk_order(Handle &h, unsigned N) {
    auto const t0 = std::chrono::steady_clock::now();
    size_t const n1 = size_t(N) + 1, nt = N / 1024 + 2;
    DevBuf<uint32_t> words;
    DevBuf<unsigned long long> claim;
    words.reserve(8 * size_t(N) + 4 * n1 + nt + 1);
    claim.reserve(std::max(N, 1u));
    uint32_t *p = words.p;
    auto take = [&](size_t n) {
        uint32_t *r = p;
        p += n;
        return r;
    };
    uint32_t host_word = 0;
    kg_order_scratch w{};
    w.lab = take(N);
    w.flag = take(N);
    w.vis = take(N);
    w.tree = take(N);
    w.ord = take(N);
    w.kids = take(N);
    w.first = take(N);
    w.done = take(N);
    w.pos = take(n1);
    w.koff = take(n1);
    w.tsize = take(n1);
    w.tbase = take(n1);
    w.tmp = take(nt);
    w.word = take(1);
    w.host = &host_word;
    w.claim = claim.p;
    h.wperm.reserve(std::max(N, 1u));
    h.winv.reserve(std::max(N, 1u));
    unsigned levels = 0, rounds = 0;
    KG_HIP(kg_walk_order(N, h.gpad.p, h.gmd.p, h.gpad_stride, w, h.wperm.p, h.winv.p, &levels, &rounds, h.stream));
    DevBuf<uint32_t> gpad2, gmd2;
    gpad2.reserve(std::max<size_t>(size_t(N) * h.gpad_stride, 1));
    gmd2.reserve(std::max(N, 1u));
    KG_HIP(kg_permute_adjacency(N, h.gpad_stride, h.wperm.p, h.winv.p, h.gpad.p, h.gmd.p, gpad2.p, gmd2.p,
                                h.stream));
    h.sync();
    std::swap(h.gpad.p, gpad2.p);
    std::swap(h.gpad.n, gpad2.n);
    std::swap(h.gmd.p, gmd2.p);
    std::swap(h.gmd.n, gmd2.n);
    h.worder_valid = true;
    h.worder_levels = levels;
    h.wforms = 0;
    h.worder_ms = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (verbosity > 1)
        std::fprintf(stderr, "walk layout: %u nodes, %u levels, %u labelling rounds, %.2f ms\n", N, levels, rounds,
                     double(h.worder_ms));
}

// The walk's copies of the rows it gathers and of the norms, in the layout's
// node order (made on first use; N graph nodes).
static void walk_copies(Handle &h, unsigned N, int form, bool screen) {
    if (!(h.wforms & 8u)) {
        h.wnorms.reserve(std::max(N, 1u));
        KG_HIP(kg_permute_rows(N, 1, h.wperm.p, h.norms.p, h.wnorms.p, h.stream));
        h.wforms |= 8u;
    }
    if (form == 2 && !(h.wforms & 4u)) {
        h.wX8.reserve(std::max<size_t>(size_t(N) * h.Dp, 4));
        KG_HIP(kg_permute_rows(N, h.Dp / 4, h.wperm.p, h.X8.p, h.wX8.p, h.stream));
        h.wforms |= 4u;
    }
    if ((form == 1 || screen) && !(h.wforms & 2u)) {
        h.wXb.reserve(std::max<size_t>(size_t(N) * h.Dp, 4));
        KG_HIP(kg_permute_rows(N, h.Dp / 2, h.wperm.p, h.Xb.p, h.wXb.p, h.stream));
        h.wforms |= 2u;
    }
    if (form == 0 && !(h.wforms & 1u)) {
        h.wX.reserve(std::max<size_t>(size_t(N) * h.Dp, 4));
        KG_HIP(kg_permute_rows(N, h.Dp, h.wperm.p, h.X.p, h.wX.p, h.stream));
        h.wforms |= 1u;
    }
}

// The walk's index-side state, made once per index: per-node CSR records,
// the padded adjacency and, in graph order, the walk layout.
static void prepare_walk(Handle &h) {
    if (h.gnode_valid) return;
    unsigned const N = h.gN;
    h.gnode.reserve(std::max(N, 1u));
    h.gflag.reserve(2);  // [0] some list repeats an id, [1] max over nodes of min(M, deg)
    KG_HIP(hipMemsetAsync(h.gflag.p, 0, 8, h.stream));
    KG_HIP(kg_launch_gnode(N, h.goff.p, h.gids.p, h.gM.p, h.gnode.p, h.gflag.p, h.stream));
    KG_HIP(kg_launch_gpad_width(N, h.goff.p, h.gM.p, h.gflag.p + 1, h.stream));
    uint32_t fl[2] = {0, 0};
    KG_HIP(hipMemcpyAsync(fl, h.gflag.p, 8, hipMemcpyDeviceToHost, h.stream));
    h.sync();
    h.lists_unique = fl[0] == 0;
    // padded adjacency: a walk with search M <= the row width reads
    // nothing past min(max(M[n], M), deg(n)) of a list; rows of the
    // widest such prefix (16-id multiples) cost at most 16 GiB
    unsigned const width = (std::max(fl[1], 1u) + 15) & ~15u;
    h.gpad.release();
    h.gmd.release();
    h.gpad_stride = 0;
    if (width < 0xffffu && size_t(N) * width * 4 <= (size_t(16) << 30)) {
        h.gpad.reserve(size_t(N) * width);
        h.gmd.reserve(std::max(N, 1u));
        KG_HIP(kg_launch_gpad(N, h.goff.p, h.gids.p, h.gM.p, width, h.gpad.p, h.gmd.p, h.stream));
        h.gpad_stride = width;
    }
    h.worder_valid = false;
    h.wforms = 0;
    if (h.search_order == 1 && h.gpad_stride) walk_order(h, N);
    h.gnode_valid = true;
}

// Batched search over device-resident queries / outputs (kgraph.cpp:423-564).
static void search_device(Handle &h, float const *dQ, unsigned nq, size_t qstride, kg_search_params const &sp,
                          unsigned const *Ks, unsigned const *Ps, uint32_t *d_ids, float *d_dists, size_t ostride,
                          uint32_t *d_counts, uint32_t *d_comps, uint32_t *d_scans) {
    if (!h.has_data) throw runtime_error("no dataset");
    if (!h.has_graph) throw runtime_error("no index");
    if (h.gN > h.N) throw runtime_error("dataset larger than index");
    if (sp.init && sp.T > 1) throw runtime_error("when init > 0, T must be 1.");
    if (nq == 0) return;
    if (sp.init && !d_ids) throw invalid_argument("no initial data provided via ids");
    // one K: the reference's rule (kgraph.cpp:463); per-query K / P (no
    // counterpart in the reference): query q walks from min(init, K_q + P_q - 1)
    // of its ids, what its pool holds
    if (sp.init && !Ks && !(sp.init < sp.K)) throw invalid_argument("L < params.K");
    if (sp.init > ostride) throw invalid_argument("init larger than the id row stride");
    if (sp.T == 0) throw invalid_argument("T must be >= 1");
    // caller's starting points must be valid node ids: checked on the device
    // (qcounter word 2), before the walk, which exits at once when one is not;
    // the call throws once the stream has drained
    // [0] next query, [1] queries whose visited set spilled, [2] bad starting id,
    // [3] the queries are not all integers in [0, 255] (the lean walk's check)
    h.qcounter.reserve(4);
    KG_HIP(hipMemsetAsync(h.qcounter.p, 0, 16, h.stream));
    uint32_t *const bad_init = h.qcounter.p + 2;
    // the check kernel also raises a pinned host word, read once the stream drains
    auto check_init = [&]() {
        h.sync();
        if (__atomic_load_n(h.host_flag, __ATOMIC_ACQUIRE)) throw invalid_argument("initial id out of range");
    };
    // a walk that may be the lean one has its query check done in the same launch
    bool const q8_checked = sp.init && h.search_lean && sp.T == 1;
    if (sp.init) {
        __atomic_store_n(h.host_flag, 0u, __ATOMIC_RELEASE);
        KG_HIP(kg_launch_init_check(d_ids, nq, ostride, sp.init, h.gN, bad_init, h.host_flag,
                                    q8_checked ? dQ : nullptr, h.D, qstride, h.qcounter.p + 3, h.stream));
    }
    if (sp.S == 0) throw invalid_argument("S must be >= 1");
    unsigned const N = h.gN;
    // per-query K/P, distinct P values for the seed table
    unsigned kmax = 0, poolmax = 0;
    std::vector<unsigned> pvals;
    std::vector<uint32_t> pidx_h;
    std::vector<uint32_t> brute;
    for (unsigned q = 0; q < nq; ++q) {
        unsigned K = Ks ? Ks[q] : sp.K, P = Ps ? Ps[q] : sp.P;
        if (K > ostride) throw invalid_argument("K larger than the output row stride");
        kmax = std::max(kmax, K);
        if (P >= N) {
            brute.push_back(q);
            if (Ps) pidx_h.push_back(0);
            continue;
        }
        poolmax = std::max(poolmax, K + P + 1);
        if (Ps) {
            auto it = std::find(pvals.begin(), pvals.end(), P);
            if (it == pvals.end()) {
                pvals.push_back(P);
                it = pvals.end() - 1;
            }
            pidx_h.push_back(uint32_t(it - pvals.begin()));
        }
    }
    if (!Ps && sp.P < N) pvals.push_back(sp.P);
    if (poolmax > 65535) throw invalid_argument("K + P too large");
    // seeds: one mt19937(seed) per search call, T GenRandom draws (kgraph.cpp:444-453)
    std::vector<uint32_t> seeds_h, soff_h;
    unsigned seed = sp.seed ? sp.seed : unsigned(time(nullptr));
    if (sp.init == 0) {
        for (unsigned P : pvals) {
            soff_h.push_back(seeds_h.size());
            std::mt19937 rng(seed);
            size_t o = seeds_h.size();
            seeds_h.resize(o + size_t(sp.T) * P);
            for (unsigned t = 0; t < sp.T; ++t) gen_random(rng, &seeds_h[o + size_t(t) * P], P, N);
        }
    }
    if (soff_h.empty()) soff_h.push_back(0);
    if (seeds_h.empty()) seeds_h.push_back(0);
    h.seeds.reserve(seeds_h.size());
    h.seed_off.reserve(soff_h.size());
    if (sp.init == 0) {  // the walk reads no seeds when it starts from ids
        KG_HIP(hipMemcpyAsync(h.seeds.p, seeds_h.data(), 4 * seeds_h.size(), hipMemcpyHostToDevice, h.stream));
        KG_HIP(hipMemcpyAsync(h.seed_off.p, soff_h.data(), 4 * soff_h.size(), hipMemcpyHostToDevice, h.stream));
    }
    if (Ks) {
        h.Ks.reserve(nq);
        KG_HIP(hipMemcpyAsync(h.Ks.p, Ks, 4ull * nq, hipMemcpyHostToDevice, h.stream));
    }
    if (Ps) {
        h.Ps.reserve(nq);
        h.pidx.reserve(nq);
        KG_HIP(hipMemcpyAsync(h.Ps.p, Ps, 4ull * nq, hipMemcpyHostToDevice, h.stream));
        KG_HIP(hipMemcpyAsync(h.pidx.p, pidx_h.data(), 4ull * nq, hipMemcpyHostToDevice, h.stream));
    }
    // mixed P: walks start in order of decreasing P (a counting sort over the
    // distinct values, stable), so the launch does not end on a long walk
    // that started last; results do not depend on the order
    std::vector<uint32_t> qorder_h;
    if (Ps && pvals.size() > 1) {
        std::vector<unsigned> rank(pvals.size());
        for (unsigned i = 0; i < pvals.size(); ++i) rank[i] = i;
        std::sort(rank.begin(), rank.end(), [&](unsigned x, unsigned y) { return pvals[x] > pvals[y]; });
        std::vector<size_t> start(pvals.size() + 1, 0);
        for (uint32_t p : pidx_h) ++start[p + 1];
        std::vector<size_t> at(pvals.size());
        size_t o = 0;
        for (unsigned r : rank) {
            at[r] = o;
            o += start[r + 1];
        }
        qorder_h.resize(nq);
        for (unsigned q = 0; q < nq; ++q) qorder_h[at[pidx_h[q]]++] = q;
        h.qorder.reserve(nq);
        KG_HIP(hipMemcpyAsync(h.qorder.p, qorder_h.data(), 4ull * nq, hipMemcpyHostToDevice, h.stream));
    }
    // queries: aligned copy when the caller's layout does not allow float4 loads
    float const *Qk = dQ;
    size_t qs = qstride;
    if ((qstride & 3) || (reinterpret_cast<uintptr_t>(dQ) & 15)) {
        h.q.reserve(size_t(nq) * h.Dp);
        KG_HIP(hipMemcpy2DAsync(h.q.p, h.Dp * 4, dQ, qstride * 4, size_t(h.D) * 4, nq, hipMemcpyDeviceToDevice,
                                h.stream));
        Qk = h.q.p;
        qs = h.Dp;
    }
    h.qn.reserve(nq);
    KG_HIP(kg_launch_norms(h.metric, Qk, nq, h.D, unsigned(qs), h.qn.p, h.stream));

    // graph-walk queries
    unsigned const ngraph = nq - unsigned(brute.size());
    if (ngraph) {
        prepare_walk(h);
        bool const pad = h.search_adjacency == 0 && h.gpad_stride && sp.M <= h.gpad_stride;
        int const form = walk_rows(h);
        // BF16 screen in front of the FP32 rows (lossy BF16 copy; exact copies are gathered instead)
        bool const screen = form == 0 && h.search_screen == 1 && h.Dp <= 256;
        // the padded rows are in the layout's order: the walk reads its row copies too
        bool const lay = pad && h.worder_valid;
        if (lay) walk_copies(h, N, form, screen);
        h.last_order = lay ? 1 : 0;
        // queries per wave: 2 walks each query on half a wave (not with the screen)
        unsigned const qpw = screen ? 1u : h.search_qpw;
        // the T > 1 result list (a single trial's result is its pool's head)
        unsigned const rcap = sp.T > 1 ? kmax + 1 : 0u;
        // seeds sorted per trial: init ids, or P random ones (largest P of the batch)
        unsigned const seed_cap = sp.init ? sp.init : std::max(poolmax, 1u);
        // visited hash set in LDS, in front of the HBM bitset.  Entries are
        // 16 bits when the ids allow (ceil(log2 N) <= log2(entries) + 10),
        // else 32.  Size: 4096 entries, else 2048, when that keeps at least 12
        // waves per CU resident (a query visits ~850 ids at P = 10 and ~1,000
        // at P = 100 and spills past half: an HBM visited test costs more
        // than a quarter of the waves); else the largest power of two
        // (256..4096) that keeps 16 waves
        unsigned idbits = 1;
        while (idbits < 32 && (uint64_t(1) << idbits) < N) ++idbits;
        // a set of c 2^k entries (c = 1 or 3): 16-bit entries hash ids of
        // B = max(idbits, k + c / 3) bits; S = B - k; the remainder takes
        // S (c = 1) or S - 1 (c = 3) bits, at most 10
        struct VtGeom {
            bool c3;
            unsigned k, B, S, rb;
        };
        auto geom = [&](unsigned c) {
            VtGeom g{};
            g.c3 = (c & (c - 1)) != 0;
            g.k = unsigned(__builtin_ctz(g.c3 ? c / 3 : c));
            g.B = std::max(idbits, g.k + (g.c3 ? 1u : 0u));
            g.S = g.B - g.k;
            g.rb = g.c3 ? g.S - 1 : g.S;
            return g;
        };
        auto bits_for = [&](unsigned c) { return (h.search_visited == 0 && geom(c).rb <= 10) ? 16u : 32u; };
        auto lds_for = [&](unsigned c) {
            return kg_search_lds_per_wave(poolmax, rcap, h.Dp, seed_cap, c, c ? bits_for(c) : 32u, !pad, screen);
        };
        auto waves = [&](unsigned c) {
            return std::min<size_t>(kg_search_waves_per_cu(), h.lds_per_cu / (qpw * lds_for(c)));
        };
        unsigned vcap = 0;
        if (h.search_visited != 1) {
            if (h.search_vcap) {
                vcap = h.search_vcap;
            } else {
                for (unsigned c : {4096u, 2048u})
                    if (!vcap && waves(c) >= 12) vcap = c;
                for (unsigned c = 4096; c >= 256 && !vcap; c /= 2)
                    if (waves(c) >= 16) vcap = c;
            }
        }
        unsigned const vt_bits = vcap ? bits_for(vcap) : 32u;
        size_t const lpq = lds_for(vcap);  // one query's LDS
        size_t const lpw = lpq * qpw;
        // exact integer distances over the 8-bit rows (integer queries):
        // sums stay below 2^24 while D 255^2 < 2^24
        bool const int_dot = form == 2 && h.row_storage == 0 && h.D <= 258;
        // The lean walk (integer queries over the 8-bit rows from the caller's
        // starting points, one trial) holds 5 waves per SIMD: it runs when a
        // visited set of 4096 or 3072 entries (or the fixed size, or the
        // bitset only) keeps its LDS within 20 waves per CU (a 2048-entry set
        // spills most queries at P = 100 and loses more than the waves gain),
        // beside the general walk, and the batch's query check decides on the
        // device which of the two launches walks (the other returns at once).
        bool lean = h.search_lean && int_dot && sp.init > 0 && sp.T == 1 && !screen && qpw == 1;
        unsigned lean_vcap = 0, lean_bits = 32, lean_wpb = 0, lean_blocks = 0;
        size_t lean_lpq = 0;
        if (lean) {
            unsigned const lw = kg_search_lean_waves_per_cu();
            auto lean_lds = [&](unsigned c) {
                return kg_search_lds_per_wave(poolmax, 0, h.Dp, seed_cap, c, c ? bits_for(c) : 32u, !pad, false, true);
            };
            lean_vcap = ~0u;
            for (unsigned c : {4096u, 3072u}) {
                if (h.search_visited == 1) c = 0u;
                else if (h.search_vcap) c = h.search_vcap;
                if (lean_vcap == ~0u && h.lds_per_cu / lean_lds(c) >= lw) lean_vcap = c;
            }
            if (lean_vcap == ~0u) {
                lean = false;  // no set size leaves the lean walk more resident waves
            } else {
                lean_bits = lean_vcap ? bits_for(lean_vcap) : 32u;
                lean_lpq = lean_lds(lean_vcap);
                lean_wpb = std::min(h.search_wpb, kg_search_max_wpb());
                while (lean_wpb > 1 && lean_lpq * lean_wpb > 65536) lean_wpb /= 2;
                size_t const llpb = lean_lpq * lean_wpb;
                if (llpb > 65536) {
                    lean = false;
                } else {
                    unsigned lbpc = unsigned(std::max<size_t>(1, std::min<size_t>(lw / lean_wpb, h.lds_per_cu / llpb)));
                    if (h.search_bpc) lbpc = std::min(lbpc, h.search_bpc);
                    lean_blocks = std::min<unsigned>((ngraph + lean_wpb - 1) / lean_wpb, h.num_cus * lbpc);
                }
            }
        }
        unsigned wpb = std::min(h.search_wpb, kg_search_max_wpb());  // the kernel's launch bound
        while (wpb > 1 && lpw * wpb > 65536) wpb /= 2;
        size_t lpb = lpw * wpb;
        if (lpb > 65536) throw invalid_argument("K + P too large for the search kernel's LDS budget");
        unsigned bpc =
            unsigned(std::max<size_t>(1, std::min<size_t>(kg_search_waves_per_cu() / wpb, h.lds_per_cu / lpb)));
        if (h.search_bpc) bpc = std::min(bpc, h.search_bpc);
        unsigned blocks = std::min<unsigned>((ngraph + wpb * qpw - 1) / (wpb * qpw), h.num_cus * bpc);
        unsigned words = (N + 31) / 32;
        size_t budget = size_t(8) << 30;  // visited bitsets: at most 8 GiB
        while (blocks > 1 && size_t(blocks) * wpb * qpw * words * 4 > budget) blocks /= 2;
        while (lean && lean_blocks > 1 && size_t(lean_blocks) * lean_wpb * words * 4 > budget) lean_blocks /= 2;
        // one visited bitset per query slot (of the larger launch of a pair)
        unsigned slots = std::max(blocks * wpb * qpw, lean ? lean_blocks * lean_wpb : 0u);
        unsigned const tcap = 4096;
        if (slots > h.vis_slots || words != h.vis_words) {
            h.visited.release();
            h.visited.reserve(size_t(slots) * words);
            KG_HIP(hipMemsetAsync(h.visited.p, 0, size_t(slots) * words * 4, h.stream));
            h.touched.reserve(size_t(slots) * tcap);
            h.vis_slots = slots;
            h.vis_words = words;
        }
        kg_search_args a{};
        a.X = lay && form == 0 ? h.wX.p : h.X.p;  // (read by the FP32 walk and the screen's re-score only)
        a.X8 = form == 2 ? (lay ? h.wX8.p : h.X8.p) : nullptr;
        a.Xb = form == 1 ? (lay ? h.wXb.p : h.Xb.p) : nullptr;
        a.Xs = screen ? (lay ? h.wXb.p : h.Xb.p) : nullptr;
        a.perm = lay ? h.wperm.p : nullptr;
        a.inv = lay ? h.winv.p : nullptr;
        h.last_screen = screen ? 1 : 0;
        if (screen) {
            kg_search_screen_constants(h.D, a);
            h.screen_cnt.reserve(1);
            KG_HIP(hipMemsetAsync(h.screen_cnt.p, 0, 8, h.stream));
            a.rescored = h.screen_cnt.p;
        }
        a.norms = lay ? h.wnorms.p : h.norms.p;
        a.N = N;
        a.D = h.D;
        a.Dp = h.Dp;
        a.goff = h.goff.p;
        a.gids = h.gids.p;
        a.gM = h.gM.p;
        a.gnode = h.gnode.p;
        a.gpad = pad ? h.gpad.p : nullptr;
        a.gmd = pad ? h.gmd.p : nullptr;
        a.gstride = pad ? h.gpad_stride : 0u;
        h.last_adjacency = pad ? 0 : 1;
        a.lists_unique = h.lists_unique ? 1u : 0u;
        a.int_dot = int_dot ? 1u : 0u;
        a.Q = Qk;
        a.qnorms = h.qn.p;
        a.nq = nq;
        a.qstride = unsigned(qs);
        a.Ks = Ks ? h.Ks.p : nullptr;
        a.Ps = Ps ? h.Ps.p : nullptr;
        a.pidx = Ps ? h.pidx.p : nullptr;
        a.qorder = qorder_h.empty() ? nullptr : h.qorder.p;
        a.K0 = sp.K;
        a.P0 = sp.P;
        a.S = sp.S;
        a.T = sp.T;
        a.Mparam = sp.M;
        a.init = sp.init;
        a.epsilon = sp.epsilon;
        a.seeds = h.seeds.p;
        a.seed_off = h.seed_off.p;
        a.init_ids = d_ids;
        a.out_ids = d_ids;
        a.out_dists = d_dists;
        a.ostride = unsigned(ostride);
        a.counts = d_counts;
        a.comps = d_comps;
        a.scans = d_scans;
        h.lat.reserve(nq);
        // the walk writes every graph query's latency; brute-force ones read 0
        if (!brute.empty()) KG_HIP(hipMemsetAsync(h.lat.p, 0, 4ull * nq, h.stream));
        a.lat_us = h.lat.p;
        h.lat_n = nq;
        a.visited = h.visited.p;
        a.vis_words = words;
        a.touched = h.touched.p;
        a.touch_cap = tcap;
        a.qcounter = h.qcounter.p;
        a.pool_cap = poolmax;
        a.res_cap = rcap;
        a.lds_per_wave = unsigned(lpq);
        h.last_qpw = qpw;
        a.vcap = vcap;
        a.spills = h.qcounter.p + 1;
        a.vt_bits = vt_bits;
        if (vt_bits == 16) {
            VtGeom const g = geom(vcap);
            a.vt_rb = g.rb;
            a.vt_shift = g.S;
            a.vt_idmask = uint32_t((uint64_t(1) << g.B) - 1);
        }
        h.last_vt_bits = vt_bits;
        a.abort_flag = sp.init ? bad_init : nullptr;
        h.last_vcap = vcap;
        a.seed_cap = seed_cap;
        h.last_lean = lean;
        if (lean) {
            // the batch's query check, then the lean / general pair
            if (!q8_checked) KG_HIP(kg_launch_q8check(Qk, nq, h.D, qs, h.qcounter.p + 3, h.stream));
            a.lean_flag = h.qcounter.p + 3;
            kg_search_args al = a;
            al.lean = 1;
            al.res_cap = 0;
            al.vcap = lean_vcap;
            al.vt_bits = lean_bits;
            al.vt_rb = al.vt_shift = al.vt_idmask = 0;
            if (lean_bits == 16) {
                VtGeom const g = geom(lean_vcap);
                al.vt_rb = g.rb;
                al.vt_shift = g.S;
                al.vt_idmask = uint32_t((uint64_t(1) << g.B) - 1);
            }
            al.lds_per_wave = unsigned(lean_lpq);
            h.last_lean_vcap = lean_vcap;
            h.last_lean_bits = lean_bits;
            KG_HIP(hipEventRecord(h.ev[0], h.stream));
            KG_HIP(kg_launch_search(h.metric, al, lean_blocks, lean_wpb, 1, h.stream));
            KG_HIP(kg_launch_search(h.metric, a, blocks, wpb, qpw, h.stream));
            KG_HIP(hipEventRecord(h.ev[1], h.stream));
        } else {
            KG_HIP(hipEventRecord(h.ev[0], h.stream));
            KG_HIP(kg_launch_search(h.metric, a, blocks, wpb, qpw, h.stream));
            KG_HIP(hipEventRecord(h.ev[1], h.stream));
        }
    }
    // brute-force queries (P >= N, kgraph.cpp:427-433)
    if (!brute.empty()) {
        if (sp.init) check_init();  // nothing is written after a bad starting id
        if (Ks) {
            for (uint32_t q : brute)
                if (Ks[q] != Ks[brute[0]]) throw invalid_argument("brute-force queries need one K");
        }
        h.bsel.reserve(brute.size());
        uint32_t *dsel = h.bsel.p;
        KG_HIP(hipMemcpyAsync(dsel, brute.data(), 4 * brute.size(), hipMemcpyHostToDevice, h.stream));
        kg_brute_args b{};
        b.X = h.X.p;
        b.norms = h.norms.p;
        b.N = N;
        b.D = h.D;
        b.Dp = h.Dp;
        b.Q = Qk;
        b.qnorms = h.qn.p;
        b.qstride = unsigned(qs);
        b.qsel = dsel;
        b.nq = unsigned(brute.size());
        b.K = Ks ? Ks[brute[0]] : sp.K;
        b.epsilon = sp.epsilon;
        b.out_ids = d_ids;
        b.out_dists = d_dists;
        b.ostride = unsigned(ostride);
        b.counts = d_counts;
        b.comps = d_comps;
        b.scans = d_scans;
        h.bpart.reserve(size_t(b.nq) * kg_brute_splits(b.nq, N, b.K) * b.K);
        b.partial = h.bpart.p;
        KG_HIP(hipEventRecord(h.ev[2], h.stream));
        KG_HIP(kg_launch_brute(h.metric, b, h.stream));
        KG_HIP(hipEventRecord(h.ev[3], h.stream));
        h.sync();
    }
    if (sp.init)
        check_init();
    else
        h.sync();
    if (ngraph) KG_HIP(hipEventElapsedTime(&h.last_ms[0], h.ev[0], h.ev[1]));
    if (!brute.empty()) KG_HIP(hipEventElapsedTime(&h.last_ms[1], h.ev[2], h.ev[3]));
}

// Entry-point sample: E distinct node ids from GenRandom on mt19937(seed),
// sorted ascending, rows copied into a compact matrix.
static void set_entry_sample(Handle &h, unsigned E, unsigned seed) {
    if (!h.has_data) throw runtime_error("no dataset");
    if (E > h.N) throw invalid_argument("entry sample larger than the dataset");
    h.entryE = 0;
    if (E == 0) return;
    std::vector<uint32_t> ids(E);
    std::mt19937 rng(seed);
    gen_random(rng, ids.data(), E, h.N);
    std::sort(ids.begin(), ids.end());
    h.eids.reserve(E);
    h.erows.reserve(size_t(E) * h.Dp);
    h.enorms.reserve((E + 15) / 16 * 16);  // whole 16-row tiles for the MFMA selection
    KG_HIP(hipMemcpyAsync(h.eids.p, ids.data(), 4ull * E, hipMemcpyHostToDevice, h.stream));
    KG_HIP(kg_launch_gather_rows(h.X.p, h.Dp, h.eids.p, E, h.erows.p, h.stream));
    KG_HIP(kg_launch_norms(h.metric, h.erows.p, E, h.D, h.Dp, h.enorms.p, h.stream));
    // BF16 operand tiles of the MFMA selection
    h.etiles.reserve(kg_entry_tiles_size(E, h.D));
    KG_HIP(kg_launch_entry_tiles(h.erows.p, E, h.D, h.Dp, h.etiles.p, h.stream));
    // ranking scale of the BF16 screen: ||e||^2/2, or ||e|| for angular
    h.escale.reserve((E + 15) / 16 * 16);
    KG_HIP(kg_launch_norms(h.metric == KG_METRIC_ANGULAR ? KG_METRIC_ANGULAR : KG_METRIC_EUCLIDEAN, h.erows.p, E, h.D,
                           h.Dp, h.escale.p, h.stream));
    // integer rows (and their largest |x|) allow the exact MFMA selection
    h.iflags.reserve(4);
    KG_HIP(hipMemsetAsync(h.iflags.p, 0, 8, h.stream));
    KG_HIP(kg_launch_intcheck(h.erows.p, E, h.D, h.Dp, h.iflags.p, h.stream));
    uint32_t fl[2];
    KG_HIP(hipMemcpyAsync(fl, h.iflags.p, 8, hipMemcpyDeviceToHost, h.stream));
    h.sync();
    h.entry_int = fl[0] == 0;
    memcpy(&h.entry_max, &fl[1], 4);
    h.entryE = E;
    h.eproj_version = ~0u;
}

// Exact MFMA selection (kg_entry_exact_bf16) on integer sample rows: the
// query test runs with the query norms ||q||^2 / 2 (into h.qsq), which the
// MFMA selection needs either way, and sets a device flag that gates the
// MFMA selection `a`, queued right behind it; the host waits for the test
// alone (its event) while the selection runs, and reads the same verdict
// from pinned memory.  Returns whether the selection applies (false: it
// returned at once and another form must run); *qsq_done tells whether the
// norms were computed.
static bool entries_exact_on_mfma(Handle &h, float const *Qk, unsigned nq, size_t qs, kg_entry_args a,
                                  bool *qsq_done) {
    *qsq_done = false;
    if (!h.entry_int || h.D > 256 || h.entry_max > 256.f) return false;
    KG_HIP(hipMemsetAsync(h.iflags.p, 0, 16, h.stream));
    uint32_t *const fl = h.host_flag + 1;  // pinned words 1..2
    KG_HIP(kg_launch_qnorm_check(Qk, nq, h.D, unsigned(qs), h.qsq.p, h.iflags.p, fl, h.entry_max, h.stream));
    KG_HIP(hipEventRecord(h.ev[19], h.stream));
    a.gate = h.iflags.p + 3;
    KG_HIP(kg_launch_entries_exact_mfma(h.metric, a, h.qsq.p, h.escale.p, h.stream));
    KG_HIP(hipEventSynchronize(h.ev[19]));
    *qsq_done = true;
    float qmax;
    uint32_t const f0 = __atomic_load_n(fl, __ATOMIC_ACQUIRE), f1 = __atomic_load_n(fl + 1, __ATOMIC_ACQUIRE);
    memcpy(&qmax, &f1, 4);
    return kg_entry_exact_bf16(f0, qmax, h.D, h.entry_max);
}

// m nearest sampled nodes per query -> ids rows [0, m) (device buffers).
static void select_entries_device(Handle &h, float const *dQ, unsigned nq, size_t qstride, unsigned m,
                                  uint32_t *d_ids, size_t ostride) {
    if (!h.has_data) throw runtime_error("no dataset");
    if (h.entryE == 0) throw runtime_error("no entry sample (kg_set_entry_sample)");
    if (m == 0 || m > 16 || m > h.entryE || m > ostride) throw invalid_argument("m must be in [1, min(16, E, ostride)]");
    if (nq == 0) return;
    float const *Qk = dQ;
    size_t qs = qstride;
    if ((qstride & 3) || (reinterpret_cast<uintptr_t>(dQ) & 15)) {
        h.q.reserve(size_t(nq) * h.Dp);
        KG_HIP(hipMemcpy2DAsync(h.q.p, h.Dp * 4, dQ, qstride * 4, size_t(h.D) * 4, nq, hipMemcpyDeviceToDevice,
                                h.stream));
        Qk = h.q.p;
        qs = h.Dp;
    }
    h.qn.reserve(nq);
    KG_HIP(kg_launch_norms(h.metric, Qk, nq, h.D, unsigned(qs), h.qn.p, h.stream));
    // Four forms, all ending in the m best by exact FP32 (dist, id):
    //   exact MFMA (integer rows and queries): fused MFMA top-m, exact as it is;
    //   learned (kg_set_entry_projection > 0): fused MFMA top-s in the projected
    //     space, then the s re-scored exactly;
    //   BF16 screen (precision BF16, D <= 256): fused MFMA top-s on the rows,
    //     then the s re-scored exactly;
    //   otherwise: canonical FP32 scores of all E, then the top-m.
    bool const learned = h.entry_shortlist > 0;
    h.qsq.reserve(std::max(size_t(nq), size_t(h.entryE)));
    KG_HIP(hipEventRecord(h.ev[14], h.stream));
    kg_entry_args a{};
    a.rows = h.erows.p;
    a.tiles = h.etiles.p;
    a.enorms = h.enorms.p;
    a.ids = h.eids.p;
    a.E = h.entryE;
    a.D = h.D;
    a.Dp = h.Dp;
    a.Q = Qk;
    a.qnorms = h.qn.p;
    a.nq = nq;
    a.qstride = unsigned(qs);
    a.m = m;
    a.out_ids = d_ids;
    a.ostride = unsigned(ostride);
    bool qsq_done = false;
    bool const exact_mfma = !learned && entries_exact_on_mfma(h, Qk, nq, qs, a, &qsq_done);
    bool const screen = !learned && !exact_mfma && h.precision == KG_PRECISION_BF16 && h.D <= 256;
    unsigned const sl = std::max(m, std::min(learned ? h.entry_shortlist : 16u, std::min(16u, h.entryE)));
    if (learned) {
        mlp_project_entries(h);
        h.zq.reserve(size_t(nq) * h.mlpP);
        h.hq.reserve(std::max(size_t(nq), size_t(h.entryE)) * h.mlpH);
    }
    if (exact_mfma) {
        // queued by entries_exact_on_mfma
    } else if (learned || screen) {
        h.eshort.reserve(size_t(nq) * sl);
        a.m = sl;
        a.out_ids = h.eshort.p;
        a.ostride = sl;
        if (learned) {  // squared Euclidean in the projected space
            mlp_forward_rows(h, Qk, nq, qs, h.hq.p, h.zq.p);
            a.rows = h.eproj.p;
            a.tiles = h.eproj_tiles.p;
            a.D = a.Dp = a.qstride = h.mlpP;
            a.Q = h.zq.p;
            KG_HIP(kg_launch_norms(KG_METRIC_EUCLIDEAN, h.zq.p, nq, h.mlpP, h.mlpP, h.qsq.p, h.stream));
            KG_HIP(kg_launch_entries_exact_mfma(KG_METRIC_L2SQR, a, h.qsq.p, h.eprojsq.p, h.stream));
        } else {
            if (!qsq_done) KG_HIP(kg_launch_norms(KG_METRIC_EUCLIDEAN, Qk, nq, h.D, unsigned(qs), h.qsq.p, h.stream));
            KG_HIP(kg_launch_entries_exact_mfma(h.metric, a, h.qsq.p, h.escale.p, h.stream));
        }
        kg_rescore_args r{};
        r.X = h.X.p;
        r.norms = h.norms.p;
        r.D = h.D;
        r.Dp = h.Dp;
        r.Q = Qk;
        r.qnorms = h.qn.p;
        r.nq = nq;
        r.qstride = unsigned(qs);
        r.short_ids = h.eshort.p;
        r.s = sl;
        r.m = m;
        r.out_ids = d_ids;
        r.ostride = unsigned(ostride);
        KG_HIP(kg_launch_rescore(h.metric, r, h.stream));
    } else {
        h.escores.reserve(size_t(nq) * h.entryE);
        a.scores = h.escores.p;
        KG_HIP(kg_launch_entries(h.metric, a, h.stream));
    }
    KG_HIP(hipEventRecord(h.ev[15], h.stream));
    h.entry_ms_pending = true;  // read by kg_last_time_ms(7) once the selection is done
}

// KNNGRAPH v2 (kgraph.cpp:272-274, 291-373).
static char const MAGIC[8] = {'K', 'N', 'N', 'G', 'R', 'A', 'P', 'H'};

static void save_index(Handle &h, char const *path, int format) {
    h.pull_graph();
    if (format == KG_FORMAT_TEXT) {
        std::ofstream os(path);
        if (!os) throw io_error(path);
        os << h.gN << '\n';
        for (unsigned i = 0; i < h.gN; ++i) {
            os << (h.h_off[i + 1] - h.h_off[i]);
            for (uint32_t o = h.h_off[i]; o < h.h_off[i + 1]; ++o)
                os << ' ' << h.h_ids[o] << ' ' << (h.no_dist ? 0.f : h.h_dist[o]);
            os << '\n';
        }
        return;
    }
    if ((format & KG_FORMAT_NO_DIST) == 0 && h.no_dist) throw runtime_error("distance information is not available");
    std::ofstream os(path, std::ios::binary);
    if (!os) throw io_error(path);
    uint32_t hdr[3] = {2u, uint32_t(format), h.gN};
    os.write(MAGIC, 8);
    os.write(reinterpret_cast<char const *>(hdr), 12);
    std::vector<unsigned char> rec;
    for (unsigned i = 0; i < h.gN; ++i) {
        uint32_t K = h.h_off[i + 1] - h.h_off[i];
        uint32_t mk[2] = {h.h_M[i], K};
        os.write(reinterpret_cast<char const *>(mk), 8);
        if (format & KG_FORMAT_NO_DIST) {
            os.write(reinterpret_cast<char const *>(&h.h_ids[h.h_off[i]]), 4ull * K);
        } else {
            rec.assign(12ull * K, 0);
            for (uint32_t k = 0; k < K; ++k) {
                memcpy(&rec[12ull * k], &h.h_ids[h.h_off[i] + k], 4);
                memcpy(&rec[12ull * k + 4], &h.h_dist[h.h_off[i] + k], 4);
            }
            os.write(reinterpret_cast<char const *>(rec.data()), rec.size());
        }
    }
    if (!os) throw io_error(path);
}

static void load_index(Handle &h, char const *path) {
    std::ifstream is(path, std::ios::binary);
    if (!is) throw io_error(path);
    char magic[8];
    uint32_t hdr[2], N;
    is.read(magic, 8);
    is.read(reinterpret_cast<char *>(hdr), 8);
    if (!is) throw runtime_error("error reading index file.");
    if (hdr[0] != 2) throw runtime_error("data version not supported.");
    is.read(reinterpret_cast<char *>(&N), 4);
    if (!is) throw runtime_error("error reading index file.");
    if (memcmp(magic, MAGIC, 8)) throw runtime_error("index corrupted.");
    bool nd = hdr[1] & KG_FORMAT_NO_DIST;
    std::vector<uint32_t> off(1, 0), ids, M(N);
    std::vector<float> dist;
    std::vector<unsigned char> rec;
    for (unsigned i = 0; i < N; ++i) {
        uint32_t mk[2];
        is.read(reinterpret_cast<char *>(mk), 8);
        if (!is) throw runtime_error("error reading index file.");
        M[i] = mk[0];
        uint64_t K = mk[1];
        if (uint64_t(off.back()) + K >= 0xffffffffull) throw runtime_error("index too large");
        size_t o = ids.size();
        ids.resize(o + K);
        if (nd) {
            is.read(reinterpret_cast<char *>(&ids[o]), 4 * K);
        } else {
            dist.resize(o + K);
            rec.resize(12 * K);
            is.read(reinterpret_cast<char *>(rec.data()), rec.size());
            for (size_t k = 0; k < K; ++k) {
                memcpy(&ids[o + k], &rec[12 * k], 4);
                memcpy(&dist[o + k], &rec[12 * k + 4], 4);
            }
        }
        if (!is) throw runtime_error("error reading index file.");
        off.push_back(uint32_t(o + K));
    }
    for (uint32_t id : ids)
        if (id >= N) throw runtime_error("index corrupted.");
    h.no_dist = nd;
    h.h_off.swap(off);
    h.h_ids.swap(ids);
    h.h_dist.swap(dist);
    h.h_M.swap(M);
    h.push_graph();
}

}  // namespace kga

// ---------------------------------------------------------------------------
using namespace kga;

struct kg_handle {
    Handle h;
};

template <typename F>
static int guard(kg_handle *k, F &&f) {
    if (!k) return KG_ERR_INVALID_ARGUMENT;
    try {
        k->h.err.clear();
        f(k->h);
        return KG_OK;
    } catch (kga::invalid_argument const &e) {
        k->h.err = e.what();
        return KG_ERR_INVALID_ARGUMENT;
    } catch (kga::io_error const &e) {
        k->h.err = e.what();
        return KG_ERR_IO;
    } catch (kga::device_error const &e) {
        k->h.err = e.what();
        return KG_ERR_DEVICE;
    } catch (std::bad_alloc const &) {
        k->h.err = "out of host memory";
        return KG_ERR_RUNTIME;
    } catch (std::exception const &e) {
        k->h.err = e.what();
        return KG_ERR_RUNTIME;
    }
}

extern "C" {

char const *kg_version(void) { return "kgraph_amd-1.0-gfx950"; }

kg_handle *kg_create(void) {
    try {
        return new kg_handle();
    } catch (...) {
        return nullptr;
    }
}

void kg_destroy(kg_handle *h) { delete h; }

char const *kg_last_error(kg_handle const *h) { return h ? h->h.err.c_str() : "null handle"; }

void kg_index_params_default(kg_index_params *p) {
    *p = kg_index_params{30, 100, 25, 10, 100, 100, 1998, 0.002f, 0.99f, 0, 0};
}

void kg_search_params_default(kg_search_params *p) { *p = kg_search_params{25, 0, 100, 10, 1, 1e30f, 1998, 0}; }

void kg_set_verbosity(unsigned v) { kga::verbosity = v; }

int kg_set_data(kg_handle *k, float const *X, unsigned N, unsigned D, size_t stride, int metric) {
    return guard(k, [&](Handle &h) { set_data(h, X, false, N, D, stride, metric); });
}

int kg_set_data_device(kg_handle *k, float const *dX, unsigned N, unsigned D, size_t stride, int metric) {
    return guard(k, [&](Handle &h) { set_data(h, dX, true, N, D, stride, metric); });
}

int kg_set_precision(kg_handle *k, int precision) {
    return guard(k, [&](Handle &h) {
        if (precision != KG_PRECISION_FP32 && precision != KG_PRECISION_BF16 && precision != KG_PRECISION_FP32_MFMA)
            throw kga::invalid_argument("unknown precision");
        h.precision = precision;
    });
}

int kg_set_search_visited(kg_handle *k, int mode) {
    return guard(k, [&](Handle &h) {
        if (mode < 0 || mode > 2) throw kga::invalid_argument("visited-set mode must be 0, 1 or 2");
        h.search_visited = mode;
    });
}

int kg_set_search_adjacency(kg_handle *k, int mode) {
    return guard(k, [&](Handle &h) {
        if (mode != 0 && mode != 1) throw kga::invalid_argument("adjacency mode must be 0 or 1");
        h.search_adjacency = mode;
    });
}

int kg_search_adjacency(kg_handle *k, int *padded, unsigned *width) {
    return guard(k, [&](Handle &h) {
        if (padded) *padded = h.last_adjacency == 0 ? 1 : 0;
        if (width) *width = h.gpad_stride;
    });
}

int kg_prepare_search(kg_handle *k) {
    return guard(k, [&](Handle &h) {
        if (!h.has_data) throw kga::runtime_error("no dataset");
        if (!h.has_graph) throw kga::runtime_error("no index");
        if (h.gN > h.N) throw kga::runtime_error("dataset larger than index");
        prepare_walk(h);
        int const form = walk_rows(h);
        bool const screen = form == 0 && h.search_screen == 1 && h.Dp <= 256;
        if (h.worder_valid) walk_copies(h, h.gN, form, screen);
        h.sync();
    });
}

int kg_set_search_lean(kg_handle *k, int on) {
    return guard(k, [&](Handle &h) {
        if (on != 0 && on != 1) throw kga::invalid_argument("lean walk switch must be 0 or 1");
        h.search_lean = on;
    });
}

int kg_set_search_order(kg_handle *k, int mode) {
    return guard(k, [&](Handle &h) {
        if (mode != 0 && mode != 1) throw kga::invalid_argument("order mode must be 0 or 1");
        if (mode != h.search_order) h.gnode_valid = false;  // the padded rows are rebuilt in the other order
        h.search_order = mode;
    });
}

int kg_search_order(kg_handle *k, int *active, unsigned *levels, float *ms, unsigned *perm) {
    return guard(k, [&](Handle &h) {
        if (active) *active = h.last_order;
        if (levels) *levels = h.worder_valid ? h.worder_levels : 0u;
        if (ms) *ms = h.worder_valid ? h.worder_ms : 0.f;
        if (perm) {
            if (!h.worder_valid) throw kga::runtime_error("no walk layout (it is made by the first search)");
            KG_HIP(hipMemcpy(perm, h.wperm.p, 4ull * h.gN, hipMemcpyDeviceToHost));
        }
    });
}

int kg_set_search_queries_per_wave(kg_handle *k, unsigned q) {
    return guard(k, [&](Handle &h) {
        if (q != 1 && q != 2) throw kga::invalid_argument("queries per wave must be 1 or 2");
        h.search_qpw = q;
    });
}

int kg_set_search_screen(kg_handle *k, int mode) {
    return guard(k, [&](Handle &h) {
        if (mode != 0 && mode != 1) throw kga::invalid_argument("screen mode must be 0 or 1");
        h.search_screen = mode;
    });
}

int kg_search_screen(kg_handle *k, int *active, unsigned long long *rescored) {
    return guard(k, [&](Handle &h) {
        unsigned long long n = 0;
        if (h.last_screen && h.screen_cnt.p) {
            KG_HIP(hipMemcpyAsync(&n, h.screen_cnt.p, 8, hipMemcpyDeviceToHost, h.stream));
            h.sync();
        }
        if (active) *active = h.last_screen;
        if (rescored) *rescored = n;
    });
}

int kg_search_screen_dists(kg_handle *k, float const *Q, unsigned nq, size_t qstride, unsigned const *ids, unsigned m,
                           float *out_d, float *out_e) {
    return guard(k, [&](Handle &h) {
        if (!h.has_data) throw kga::runtime_error("no dataset");
        if ((!Q || !ids || !out_d || !out_e) && nq && m) throw kga::invalid_argument("null argument");
        if (qstride < h.D) throw kga::invalid_argument("bad stride");
        if (h.Dp > 256) throw kga::invalid_argument("the BF16 screen takes rows of at most 256 values");
        if (nq == 0 || m == 0) return;
        size_t const n = size_t(nq) * m;
        for (size_t i = 0; i < n; ++i)
            if (ids[i] >= h.N) throw kga::invalid_argument("node id out of range");
        DevBuf<float> dq, dqn, dd, de;
        DevBuf<uint32_t> di;
        upload_rows(h, Q, false, nq, h.D, qstride, dq);
        dqn.reserve(nq);
        KG_HIP(kg_launch_norms(h.metric, dq.p, nq, h.D, h.Dp, dqn.p, h.stream));
        di.reserve(n);
        dd.reserve(n);
        de.reserve(n);
        KG_HIP(hipMemcpyAsync(di.p, ids, 4 * n, hipMemcpyHostToDevice, h.stream));
        kg_search_args a{};
        a.X = h.X.p;
        a.Xs = h.Xb.p;
        a.norms = h.norms.p;
        a.N = h.N;
        a.D = h.D;
        a.Dp = h.Dp;
        a.Q = dq.p;
        a.qnorms = dqn.p;
        a.nq = nq;
        a.qstride = h.Dp;
        kg_search_screen_constants(h.D, a);
        KG_HIP(kg_launch_screen_dists(h.metric, a, di.p, m, dd.p, de.p, h.stream));
        KG_HIP(hipMemcpyAsync(out_d, dd.p, 4 * n, hipMemcpyDeviceToHost, h.stream));
        KG_HIP(hipMemcpyAsync(out_e, de.p, 4 * n, hipMemcpyDeviceToHost, h.stream));
        h.sync();
    });
}

int kg_set_search_visited_size(kg_handle *k, unsigned entries) {
    return guard(k, [&](Handle &h) {
        unsigned const k = entries % 3u == 0u ? entries / 3u : entries;  // 3 * 2^k or 2^k entries
        if (entries && (entries < 256 || entries > 8192 || (k & (k - 1))))
            throw kga::invalid_argument("visited-set size must be 0, 2^k or 3 * 2^k in [256, 8192]");
        h.search_vcap = entries;
    });
}

// Whether the last walk was the lean one of a pair (reads the batch's check).
static bool lean_walked(Handle &h) {
    if (!h.last_lean) return false;
    uint32_t f = 1;
    KG_HIP(hipMemcpyAsync(&f, h.qcounter.p + 3, 4, hipMemcpyDeviceToHost, h.stream));
    h.sync();
    return f == 0;
}

int kg_search_visited_bits(kg_handle *k, unsigned *bits) {
    return guard(k, [&](Handle &h) {
        if (!bits) throw kga::invalid_argument("null argument");
        *bits = lean_walked(h) ? h.last_lean_bits : h.last_vt_bits;
    });
}

int kg_search_visited_stats(kg_handle *k, unsigned *vcap, unsigned *spilled) {
    return guard(k, [&](Handle &h) {
        uint32_t sp = 0;
        if (h.qcounter.p) {
            h.sync();
            KG_HIP(hipMemcpy(&sp, h.qcounter.p + 1, 4, hipMemcpyDeviceToHost));
        }
        if (vcap) *vcap = lean_walked(h) ? h.last_lean_vcap : h.last_vcap;
        if (spilled) *spilled = sp;
    });
}

int kg_set_row_storage(kg_handle *k, int mode) {
    return guard(k, [&](Handle &h) {
        if (mode < 0 || mode > 3) throw kga::invalid_argument("row storage mode must be 0, 1, 2 or 3");
        h.row_storage = mode;
    });
}

int kg_row_storage(kg_handle *k, int *active, int *lossless) {
    return guard(k, [&](Handle &h) {
        if (active) *active = walk_rows(h);
        if (lossless) *lossless = h.has_data ? (int(h.rows_bf16) | (int(h.rows_u8) << 1)) : 0;
    });
}

int kg_build(kg_handle *k, kg_index_params const *params, kg_index_info *info) {
    return guard(k, [&](Handle &h) {
        if (!params) throw kga::invalid_argument("null params");
        build_index(h, *params, info);
    });
}

int kg_search_device(kg_handle *k, float const *dQ, unsigned nq, size_t qstride, kg_search_params const *params,
                     unsigned const *Ks, unsigned const *Ps, unsigned *d_ids, float *d_dists, size_t ostride,
                     unsigned *d_counts, unsigned *d_comps, unsigned *d_scans) {
    return guard(k, [&](Handle &h) {
        if (!params || !d_ids || !d_counts || (!dQ && nq)) throw kga::invalid_argument("null argument");
        search_device(h, dQ, nq, qstride, *params, Ks, Ps, d_ids, d_dists, ostride, d_counts, d_comps, d_scans);
    });
}

int kg_search(kg_handle *k, float const *Q, unsigned nq, size_t qstride, kg_search_params const *params,
              unsigned const *Ks, unsigned const *Ps, unsigned *ids, float *dists, size_t ostride, unsigned *counts,
              unsigned *comps, unsigned *scans) {
    return guard(k, [&](Handle &h) {
        if (!params || !ids || !counts || (!Q && nq)) throw kga::invalid_argument("null argument");
        if (!h.has_data) throw kga::runtime_error("no dataset");
        if (qstride < h.D) throw kga::invalid_argument("bad stride");
        if (nq == 0) return;
        // the handle's device buffers, grown on demand and kept
        DevBuf<float> &dq = h.io_q, &dd = h.io_dists;
        DevBuf<uint32_t> &di = h.io_ids, &dc = h.io_counts, &dcomp = h.io_comps, &dscan = h.io_scans;
        upload_rows(h, Q, false, nq, h.D, qstride, dq);
        di.reserve(size_t(nq) * ostride);
        dc.reserve(nq);
        KG_HIP(hipMemsetAsync(di.p, 0, size_t(nq) * ostride * 4, h.stream));
        if (params->init)
            KG_HIP(hipMemcpyAsync(di.p, ids, size_t(nq) * ostride * 4, hipMemcpyHostToDevice, h.stream));
        if (dists) {
            dd.reserve(size_t(nq) * ostride);
            KG_HIP(hipMemsetAsync(dd.p, 0, size_t(nq) * ostride * 4, h.stream));
        }
        if (comps) dcomp.reserve(nq);
        if (scans) dscan.reserve(nq);
        search_device(h, dq.p, nq, h.Dp, *params, Ks, Ps, di.p, dists ? dd.p : nullptr, ostride, dc.p,
                      comps ? dcomp.p : nullptr, scans ? dscan.p : nullptr);
        KG_HIP(hipMemcpy(ids, di.p, size_t(nq) * ostride * 4, hipMemcpyDeviceToHost));
        if (dists) KG_HIP(hipMemcpy(dists, dd.p, size_t(nq) * ostride * 4, hipMemcpyDeviceToHost));
        KG_HIP(hipMemcpy(counts, dc.p, 4ull * nq, hipMemcpyDeviceToHost));
        if (comps) KG_HIP(hipMemcpy(comps, dcomp.p, 4ull * nq, hipMemcpyDeviceToHost));
        if (scans) KG_HIP(hipMemcpy(scans, dscan.p, 4ull * nq, hipMemcpyDeviceToHost));
    });
}

int kg_bruteforce(kg_handle *k, float const *Q, unsigned nq, size_t qstride, unsigned K, float epsilon,
                  unsigned *ids, float *dists, size_t ostride, unsigned *counts) {
    return guard(k, [&](Handle &h) {
        if (!h.has_data) throw kga::runtime_error("no dataset");
        if (!Q || !ids || !counts) throw kga::invalid_argument("null argument");
        if (K == 0 || K > 1024 || K > ostride) throw kga::invalid_argument("K must be in [1, min(1024, ostride)]");
        if (nq == 0) return;
        DevBuf<float> &dq = h.io_q, &dqn = h.io_qn, &dd = h.io_dists;
        DevBuf<uint32_t> &di = h.io_ids, &dc = h.io_counts;
        upload_rows(h, Q, false, nq, h.D, qstride, dq);
        dqn.reserve(nq);
        KG_HIP(kg_launch_norms(h.metric, dq.p, nq, h.D, h.Dp, dqn.p, h.stream));
        di.reserve(size_t(nq) * ostride);
        dd.reserve(size_t(nq) * ostride);
        dc.reserve(nq);
        KG_HIP(hipMemsetAsync(di.p, 0, size_t(nq) * ostride * 4, h.stream));
        KG_HIP(hipMemsetAsync(dd.p, 0, size_t(nq) * ostride * 4, h.stream));
        kg_brute_args b{};
        b.X = h.X.p;
        b.norms = h.norms.p;
        b.N = h.N;
        b.D = h.D;
        b.Dp = h.Dp;
        b.Q = dq.p;
        b.qnorms = dqn.p;
        b.qstride = h.Dp;
        b.nq = nq;
        b.K = K;
        b.epsilon = epsilon;
        b.out_ids = di.p;
        b.out_dists = dd.p;
        b.ostride = unsigned(ostride);
        b.counts = dc.p;
        h.bpart.reserve(size_t(nq) * kg_brute_splits(nq, h.N, K) * K);
        b.partial = h.bpart.p;
        KG_HIP(hipEventRecord(h.ev[2], h.stream));
        KG_HIP(kg_launch_brute(h.metric, b, h.stream));
        KG_HIP(hipEventRecord(h.ev[3], h.stream));
        h.sync();
        KG_HIP(hipEventElapsedTime(&h.last_ms[1], h.ev[2], h.ev[3]));
        KG_HIP(hipMemcpy(ids, di.p, size_t(nq) * ostride * 4, hipMemcpyDeviceToHost));
        if (dists) KG_HIP(hipMemcpy(dists, dd.p, size_t(nq) * ostride * 4, hipMemcpyDeviceToHost));
        KG_HIP(hipMemcpy(counts, dc.p, 4ull * nq, hipMemcpyDeviceToHost));
    });
}

int kg_bruteforce_device_noalloc(kg_handle *k, float const *dQ, unsigned nq, size_t qstride, unsigned K,
                                 float epsilon, unsigned *d_ids, float *d_dists, size_t ostride, unsigned *d_counts) {
    return guard(k, [&](Handle &h) {
        kga::bruteforce_noalloc(h, dQ, nq, qstride, K, epsilon, d_ids, d_dists, ostride, d_counts);
    });
}

int kg_save(kg_handle *k, char const *path, int format) {
    return guard(k, [&](Handle &h) {
        if (!path) throw kga::invalid_argument("null path");
        save_index(h, path, format);
    });
}

int kg_load(kg_handle *k, char const *path) {
    return guard(k, [&](Handle &h) {
        if (!path) throw kga::invalid_argument("null path");
        load_index(h, path);
    });
}

int kg_get_nn(kg_handle *k, unsigned id, unsigned *nns, float *dists, unsigned *M, unsigned *L) {
    return guard(k, [&](Handle &h) {
        h.pull_graph();
        if (!(id < h.gN)) throw kga::invalid_argument("id too big");
        uint32_t b = h.h_off[id], e = h.h_off[id + 1];
        if (M) *M = h.h_M[id];
        if (L) *L = e - b;
        if (nns) std::copy(h.h_ids.begin() + b, h.h_ids.begin() + e, nns);
        if (dists) {
            if (h.no_dist) throw kga::runtime_error("distance information is not available");
            std::copy(h.h_dist.begin() + b, h.h_dist.begin() + e, dists);
        }
    });
}

int kg_reverse(kg_handle *k, int rev) {
    return guard(k, [&](Handle &h) { reverse_index(h, rev); });
}

int kg_prune(kg_handle *k, unsigned level) {
    return guard(k, [&](Handle &h) { prune_index(h, level); });
}

int kg_index_size(kg_handle *k, unsigned *N, uint64_t *edges) {
    return guard(k, [&](Handle &h) {
        if (N) *N = h.has_graph ? h.gN : 0;
        if (edges) *edges = h.has_graph ? h.edges : 0;
    });
}

int kg_set_graph(kg_handle *k, unsigned N, uint64_t const *offsets, unsigned const *ids, float const *dists,
                 unsigned const *M) {
    return guard(k, [&](Handle &h) {
        if (!offsets || !M || (!ids && offsets[N])) throw kga::invalid_argument("null argument");
        uint64_t E = offsets[N];
        if (E >= 0xffffffffull) throw kga::invalid_argument("index too large");
        h.h_off.resize(N + 1);
        for (unsigned i = 0; i <= N; ++i) {
            if (i && offsets[i] < offsets[i - 1]) throw kga::invalid_argument("offsets must be non-decreasing");
            h.h_off[i] = uint32_t(offsets[i]);
        }
        for (uint64_t e = 0; e < E; ++e)
            if (ids[e] >= N) throw kga::invalid_argument("neighbour id out of range");
        h.h_ids.assign(ids, ids + E);
        h.h_M.assign(M, M + N);
        h.no_dist = dists == nullptr;
        if (dists)
            h.h_dist.assign(dists, dists + E);
        else
            h.h_dist.clear();
        h.push_graph();
    });
}

int kg_get_graph(kg_handle *k, uint64_t *offsets, unsigned *ids, float *dists, unsigned *M) {
    return guard(k, [&](Handle &h) {
        h.pull_graph();
        for (unsigned i = 0; i <= h.gN; ++i)
            if (offsets) offsets[i] = h.h_off[i];
        if (ids) std::copy(h.h_ids.begin(), h.h_ids.end(), ids);
        if (dists) {
            if (h.no_dist) throw kga::runtime_error("distance information is not available");
            std::copy(h.h_dist.begin(), h.h_dist.end(), dists);
        }
        if (M) std::copy(h.h_M.begin(), h.h_M.end(), M);
    });
}

int kg_last_time_ms(kg_handle *k, int which, float *ms) {
    return guard(k, [&](Handle &h) {
        if (which < 0 || which >= 9 || !ms) throw kga::invalid_argument("bad timer");
        if (which == 7 && h.entry_ms_pending) {
            KG_HIP(hipEventSynchronize(h.ev[15]));
            KG_HIP(hipEventElapsedTime(&h.last_ms[7], h.ev[14], h.ev[15]));
            h.entry_ms_pending = false;
        }
        *ms = h.last_ms[which];
    });
}

int kg_build_trace(kg_handle *k, float *out, unsigned cap, unsigned *n) {
    return guard(k, [&](Handle &h) {
        unsigned it = unsigned(h.trace.size() / 4);
        if (n) *n = it;
        if (out)
            for (unsigned i = 0; i < std::min(it, cap) * 4; ++i) out[i] = h.trace[i];
    });
}

int kg_build_iterations(kg_handle *k, uint64_t *out, unsigned cap, unsigned *n) {
    return guard(k, [&](Handle &h) {
        unsigned it = unsigned(h.iter_stats.size() / 6);
        if (n) *n = it;
        if (out)
            for (unsigned i = 0; i < std::min(it, cap) * 6; ++i) out[i] = h.iter_stats[i];
    });
}

int kg_set_entry_sample(kg_handle *k, unsigned E, unsigned seed) {
    return guard(k, [&](Handle &h) { set_entry_sample(h, E, seed); });
}

int kg_select_entries_device(kg_handle *k, float const *dQ, unsigned nq, size_t qstride, unsigned m,
                             unsigned *d_ids, size_t ostride) {
    return guard(k, [&](Handle &h) {
        if ((!dQ && nq) || !d_ids) throw kga::invalid_argument("null argument");
        select_entries_device(h, dQ, nq, qstride, m, d_ids, ostride);
    });
}

int kg_select_entries(kg_handle *k, float const *Q, unsigned nq, size_t qstride, unsigned m, unsigned *ids,
                      size_t ostride) {
    return guard(k, [&](Handle &h) {
        if ((!Q && nq) || !ids) throw kga::invalid_argument("null argument");
        if (!h.has_data) throw kga::runtime_error("no dataset");
        if (qstride < h.D) throw kga::invalid_argument("bad stride");
        if (nq == 0) return;
        DevBuf<float> dq;
        DevBuf<uint32_t> di;
        upload_rows(h, Q, false, nq, h.D, qstride, dq);
        di.reserve(size_t(nq) * ostride);
        KG_HIP(hipMemcpyAsync(di.p, ids, size_t(nq) * ostride * 4, hipMemcpyHostToDevice, h.stream));
        select_entries_device(h, dq.p, nq, h.Dp, m, di.p, ostride);
        h.sync();
        KG_HIP(hipMemcpy(ids, di.p, size_t(nq) * ostride * 4, hipMemcpyDeviceToHost));
    });
}

int kg_search_latencies(kg_handle *k, float *out_us, unsigned n) {
    return guard(k, [&](Handle &h) {
        if (!out_us) throw kga::invalid_argument("null argument");
        unsigned c = std::min(n, h.lat_n);
        if (c) KG_HIP(hipMemcpy(out_us, h.lat.p, 4ull * c, hipMemcpyDeviceToHost));
    });
}

int kg_mlp_init(kg_handle *k, unsigned hidden, unsigned out_dim, unsigned seed) {
    return guard(k, [&](Handle &h) { mlp_init(h, hidden, out_dim, seed); });
}

int kg_mlp_train(kg_handle *k, unsigned steps, unsigned batch, float lr, float margin, unsigned seed,
                 float *loss_trace) {
    return guard(k, [&](Handle &h) { mlp_train(h, steps, batch, lr, margin, seed, loss_trace); });
}

int kg_mlp_loss_grad(kg_handle *k, unsigned const *triplets, unsigned nt, float margin, float *loss, float *dW1,
                     float *db1, float *dW2, float *db2) {
    return guard(k, [&](Handle &h) {
        if (!triplets || nt == 0) throw kga::invalid_argument("no triplets");
        mlp_loss_grad(h, triplets, nt, margin, loss, dW1, db1, dW2, db2);
    });
}

int kg_mlp_get_weights(kg_handle *k, float *W1, float *b1, float *W2, float *b2) {
    return guard(k, [&](Handle &h) {
        if (!h.mlpH) throw kga::runtime_error("no projector (kg_mlp_init)");
        size_t D = h.D, H = h.mlpH, P = h.mlpP;
        if (W1) KG_HIP(hipMemcpy(W1, h.W1.p, 4 * H * D, hipMemcpyDeviceToHost));
        if (b1) KG_HIP(hipMemcpy(b1, h.b1.p, 4 * H, hipMemcpyDeviceToHost));
        if (W2) KG_HIP(hipMemcpy(W2, h.W2.p, 4 * P * H, hipMemcpyDeviceToHost));
        if (b2) KG_HIP(hipMemcpy(b2, h.b2.p, 4 * P, hipMemcpyDeviceToHost));
    });
}

int kg_mlp_set_weights(kg_handle *k, float const *W1, float const *b1, float const *W2, float const *b2) {
    return guard(k, [&](Handle &h) {
        if (!h.mlpH) throw kga::runtime_error("no projector (kg_mlp_init)");
        if (!W1 || !b1 || !W2 || !b2) throw kga::invalid_argument("null argument");
        size_t D = h.D, H = h.mlpH, P = h.mlpP;
        KG_HIP(hipMemcpy(h.W1.p, W1, 4 * H * D, hipMemcpyHostToDevice));
        KG_HIP(hipMemcpy(h.b1.p, b1, 4 * H, hipMemcpyHostToDevice));
        KG_HIP(hipMemcpy(h.W2.p, W2, 4 * P * H, hipMemcpyHostToDevice));
        KG_HIP(hipMemcpy(h.b2.p, b2, 4 * P, hipMemcpyHostToDevice));
        ++h.mlp_version;
    });
}

int kg_mlp_project(kg_handle *k, float const *X, unsigned n, size_t stride, float *Z) {
    return guard(k, [&](Handle &h) {
        if (!X || !Z) throw kga::invalid_argument("null argument");
        if (!h.mlpH) throw kga::runtime_error("no projector (kg_mlp_init)");
        if (stride < h.D) throw kga::invalid_argument("bad stride");
        if (n == 0) return;
        DevBuf<float> dx, dh, dz;
        upload_rows(h, X, false, n, h.D, stride, dx);
        dh.reserve(size_t(n) * h.mlpH);
        dz.reserve(size_t(n) * h.mlpP);
        mlp_forward_rows(h, dx.p, n, h.Dp, dh.p, dz.p);
        h.sync();
        KG_HIP(hipMemcpy(Z, dz.p, 4ull * n * h.mlpP, hipMemcpyDeviceToHost));
    });
}

int kg_set_entry_projection(kg_handle *k, unsigned shortlist) {
    return guard(k, [&](Handle &h) {
        if (shortlist > 16) throw kga::invalid_argument("shortlist must be <= 16");
        if (shortlist && !h.mlpH) throw kga::runtime_error("no projector (kg_mlp_init)");
        h.entry_shortlist = shortlist;
    });
}

int kg_set_search_launch(kg_handle *k, unsigned waves_per_block, unsigned blocks_per_cu) {
    return guard(k, [&](Handle &h) {
        if (waves_per_block != 1 && waves_per_block != 2 && waves_per_block != 4)
            throw kga::invalid_argument("waves per block must be 1, 2 or 4");
        if (blocks_per_cu > 32) throw kga::invalid_argument("blocks per CU must be <= 32");
        h.search_wpb = waves_per_block;
        h.search_bpc = blocks_per_cu;
    });
}

int kg_autotune_search(kg_handle *k, float const *dQ, unsigned nq, size_t qstride, kg_search_params const *params,
                       unsigned const *Ks, unsigned const *Ps, unsigned *chosen) {
    return guard(k, [&](Handle &h) {
        if (!params || !dQ || nq == 0) throw kga::invalid_argument("null argument");
        unsigned kmax = params->K;
        if (Ks)
            for (unsigned q = 0; q < nq; ++q) kmax = std::max(kmax, Ks[q]);
        kmax = std::max(kmax, 1u);
        DevBuf<uint32_t> di, dc;
        DevBuf<float> dd;
        di.reserve(size_t(nq) * kmax);
        dd.reserve(size_t(nq) * kmax);
        dc.reserve(nq);
        float best = 3.4e38f;
        unsigned bw = 4, bb = 0;
        for (unsigned wpb : {1u, 2u, 4u}) {  // search_kernel runs at most 4 waves per block
            for (unsigned bpc : {4u, 8u, 16u, 32u}) {
                if (wpb * bpc > 32) continue;  // at most 32 resident waves per CU
                h.search_wpb = wpb;
                h.search_bpc = bpc;
                float t = 0;
                kg_search_params sp = *params;
                sp.init = 0;  // random starting points: the shape does not depend on them
                for (int rep = 0; rep < 2; ++rep) {
                    search_device(h, dQ, nq, qstride, sp, Ks, Ps, di.p, dd.p, kmax, dc.p, nullptr, nullptr);
                    t = h.last_ms[0];
                }
                if (t < best) {
                    best = t;
                    bw = wpb;
                    bb = bpc;
                }
            }
        }
        h.search_wpb = bw;
        h.search_bpc = bb;
        if (chosen) {
            chosen[0] = bw;
            chosen[1] = bb;
        }
    });
}

int kg_autotune_walk(kg_handle *k, float const *dQ, unsigned nq, size_t qstride, kg_search_params const *params,
                     unsigned const *Ks, unsigned const *Ps, unsigned const *d_init, size_t istride, unsigned *chosen) {
    return guard(k, [&](Handle &h) {
        if (!params || !dQ || nq == 0) throw kga::invalid_argument("null argument");
        if (params->init && (!d_init || istride < params->init))
            throw kga::invalid_argument("starting points need a device buffer of init ids per query");
        unsigned kmax = params->K;
        if (Ks)
            for (unsigned q = 0; q < nq; ++q) kmax = std::max(kmax, Ks[q]);
        kmax = std::max(kmax, 1u);
        size_t const ostride = std::max<size_t>(kmax, params->init ? istride : 0);
        DevBuf<uint32_t> di, dc;
        DevBuf<float> dd;
        di.reserve(size_t(nq) * ostride);
        dd.reserve(size_t(nq) * ostride);
        dc.reserve(nq);
        // every setting below returns the same walk; each is timed on this
        // sample (kernel time from HIP events, second of two runs)
        auto run = [&]() {
            float t = 0;
            for (int rep = 0; rep < 2; ++rep) {
                if (params->init)
                    KG_HIP(hipMemcpy2DAsync(di.p, ostride * 4, d_init, istride * 4, size_t(params->init) * 4, nq,
                                            hipMemcpyDeviceToDevice, h.stream));
                search_device(h, dQ, nq, qstride, *params, Ks, Ps, di.p, dd.p, ostride, dc.p, nullptr, nullptr);
                t = h.last_ms[0];
            }
            return t;
        };
        int const old_visited = h.search_visited, old_adj = h.search_adjacency, old_screen = h.search_screen;
        int const old_order = h.search_order, old_lean = h.search_lean;
        unsigned const old_vcap = h.search_vcap, old_wpb = h.search_wpb, old_bpc = h.search_bpc;
        unsigned const old_qpw = h.search_qpw;
        try {
            h.search_wpb = 4;
            h.search_bpc = 0;
            float best = run();
            // 1. visited set: LDS set of 1024..8192 entries in front of the
            // bitset (16-bit entries where the ids allow), or the bitset only
            struct V {
                int mode;
                unsigned vcap;
            };
            V bv{h.search_visited, h.search_vcap};
            for (V v : {V{1, 0u}, V{0, 1024u}, V{0, 2048u}, V{0, 4096u}, V{0, 8192u}}) {
                h.search_visited = v.mode;
                h.search_vcap = v.vcap;
                float const t = run();
                if (t < best) {
                    best = t;
                    bv = v;
                }
            }
            h.search_visited = bv.mode;
            h.search_vcap = bv.vcap;
            // 2. adjacency rows: padded or CSR
            int badj = h.search_adjacency;
            for (int adj : {0, 1}) {
                if (adj == badj) continue;
                h.search_adjacency = adj;
                float const t = run();
                if (t < best) {
                    best = t;
                    badj = adj;
                }
            }
            h.search_adjacency = badj;
            // 2b. with the padded rows: the walk layout, or dataset order (a
            // switch remakes the padded rows in the other order)
            if (h.search_adjacency == 0 && h.gpad_stride) {
                int const bord = h.search_order;
                h.search_order = 1 - bord;
                h.gnode_valid = false;
                float const t = run();
                if (t < best) {
                    best = t;
                } else {
                    h.search_order = bord;
                    h.gnode_valid = false;
                }
            }
            // 2c. the lean walk, where it applies (the launch pair decides per batch)
            {
                int const blean = h.search_lean;
                h.search_lean = 1 - blean;
                float const t = run();
                if (t < best)
                    best = t;
                else
                    h.search_lean = blean;
            }
            // 3. the BF16 screen, where walks gather the FP32 rows
            int bscr = h.search_screen;
            if (walk_rows(h) == 0 && h.Dp <= 256) {
                h.search_screen = 1 - bscr;
                float const t = run();
                if (t < best) {
                    best = t;
                    bscr = h.search_screen;
                }
            }
            h.search_screen = bscr;
            // 4. queries per wave: two per wave (half a wave each) where the
            // screen is off
            unsigned bq = h.search_qpw;
            if (!(h.search_screen && walk_rows(h) == 0 && h.Dp <= 256)) {
                h.search_qpw = bq == 1 ? 2u : 1u;
                float const t = run();
                if (t < best) {
                    best = t;
                    bq = h.search_qpw;
                }
            }
            h.search_qpw = bq;
            // 5. launch shape: waves per block x resident blocks per CU
            unsigned bw = h.search_wpb, bb = h.search_bpc;
            for (unsigned wpb : {1u, 2u, 4u}) {
                for (unsigned bpc : {4u, 8u, 16u, 32u}) {
                    if (wpb * bpc > 32) continue;  // at most 32 resident waves per CU
                    h.search_wpb = wpb;
                    h.search_bpc = bpc;
                    float const t = run();
                    if (t < best) {
                        best = t;
                        bw = wpb;
                        bb = bpc;
                    }
                }
            }
            h.search_wpb = bw;
            h.search_bpc = bb;
            if (chosen) {
                chosen[0] = bw;
                chosen[1] = bb;
                chosen[2] = h.search_visited == 1 ? 0u : (h.search_vcap ? h.search_vcap : h.last_vcap);
                chosen[3] = unsigned(h.search_adjacency);
                chosen[4] = unsigned(h.search_screen);
                chosen[5] = h.search_qpw;
            }
        } catch (...) {
            h.search_visited = old_visited;
            h.search_adjacency = old_adj;
            h.search_screen = old_screen;
            h.search_vcap = old_vcap;
            h.search_wpb = old_wpb;
            h.search_bpc = old_bpc;
            h.search_qpw = old_qpw;
            if (h.search_order != old_order) {
                h.search_order = old_order;
                h.gnode_valid = false;
            }
            h.search_lean = old_lean;
            throw;
        }
    });
}

int kg_mlp_get_normalization(kg_handle *k, float *mu, float *inv_sigma) {
    return guard(k, [&](Handle &h) {
        if (!h.mlpH) throw kga::runtime_error("no projector (kg_mlp_init)");
        if (mu) KG_HIP(hipMemcpy(mu, h.mlp_mu.p, 4ull * h.D, hipMemcpyDeviceToHost));
        if (inv_sigma) KG_HIP(hipMemcpy(inv_sigma, h.mlp_is.p, 4ull * h.D, hipMemcpyDeviceToHost));
    });
}

int kg_set_join_int8(kg_handle *k, int on) {
    return guard(k, [&](Handle &h) {
        if (on != 0 && on != 1) throw kga::invalid_argument("int8 join switch must be 0 or 1");
        h.join_i8 = on == 1;
    });
}

int kg_build_counters(kg_handle *k, uint64_t *out, unsigned n) {
    return guard(k, [&](Handle &h) {
        if (!out) throw kga::invalid_argument("null argument");
        for (unsigned i = 0; i < n && i < 4; ++i) out[i] = h.build_counters[i];
        if (n > 4) out[4] = uint64_t(h.join_exact + 1);
        if (n > 5) out[5] = h.build_counters[4];
        if (n > 6) out[6] = h.build_counters[5];
    });
}

int kg_screen_pairs(kg_handle *k, const uint32_t *pairs, unsigned n, float *screened, float *bound) {
    return guard(k, [&](Handle &h) { screen_pairs(h, pairs, n, screened, bound); });
}

int kg_set_join_form(kg_handle *k, int form) {
    return guard(k, [&](Handle &h) {
        if (form != KG_JOIN_AUTO && form != KG_JOIN_SCREEN && form != KG_JOIN_TILES)
            throw kga::invalid_argument("unknown join form");
        h.join_form = form;
    });
}

int kg_set_join_prerank(kg_handle *k, float keep) {
    return guard(k, [&](Handle &h) {
        if (!(keep >= 0.f && keep <= 1.f)) throw kga::invalid_argument("keep must be in [0, 1]");
        h.prerank_keep = keep;
    });
}

int kg_build_prerank(kg_handle *k, float *out) {
    return guard(k, [&](Handle &h) {
        if (!out) throw kga::invalid_argument("null argument");
        out[0] = float(h.prerank_status);
        out[1] = h.prerank_acc;
        out[2] = float(h.prerank_fallback);
    });
}

int kg_set_build_graphs(kg_handle *k, int on) {
    return guard(k, [&](Handle &h) { h.build_graphs = on != 0; });
}

}  // extern "C"

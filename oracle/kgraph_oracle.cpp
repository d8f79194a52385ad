// kgraph_oracle.cpp -- TEST INFRASTRUCTURE ONLY.
//
// A CPU restatement of KGraph's NN-descent k-NN graph build and graph-walk
// search, written to check the HIP implementation in kgraph_amd/ and to time a
// CPU baseline.  Only tests/, __graft_entry__.smoke() and bench.py's CPU
// baseline leg may load it; the product library never links or calls it.
//
// Every function cites the reference file:line it restates.  Deliberate
// decisions where the reference is non-deterministic or reads uninitialised
// memory (all documented in DESIGN.md's oracle section):
//   * distance arithmetic: one canonical FP32 order ("X16"), the order of the
//     reference's xsimd AVX-512 kernels with FMA contraction
//     (kgraph-data.h:41-103, xsimd arch/xsimd_avx512f.hpp:1393-1403 ->
//     arch/xsimd_sse3.hpp:47-52): 16 interleaved fmaf partials, the xsimd
//     AVX-512 reduce tree, then a sequential fmaf tail.  The HIP kernels use
//     the same order, so distances are bit-identical between the two;
//   * ties: seed sort and reverse-edge sort order by (dist, id); the
//     reference's std::sort leaves equal-distance order unspecified
//     (kgraph.cpp:476, 697);
//   * radius = worst valid distance once a pool is full, +inf before
//     (kgraph.cpp:823 reads the scratch slot, which is uninitialised until
//     the pool fills);
//   * graph extraction emits only valid pool entries (kgraph.cpp:982-987
//     copies params.L entries even when fewer are valid);
//   * std::random_shuffle(rand()) (kgraph.cpp:262, 873, 878) becomes
//     std::shuffle with an mt19937 seeded from params.seed.
//
// Parity pins (tests/test_oracle.py): mt19937 stream vs numpy's MT19937,
// GenRandom invariants, UpdateKnnList brute-force model, distance order vs
// FP64 bounds and integer-exact data, search on complete graphs (reduces to
// exact k-NN), brute force vs FP64 sort, reverse-edge set semantics,
// save/load byte layout vs hand-built fixtures, NN-descent recall vs exact.

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <ctime>
#include <fstream>
#include <limits>
#include <mutex>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace kgo {

using std::vector;

enum Metric { METRIC_L2SQR = 0, METRIC_EUCLIDEAN = 1, METRIC_ANGULAR = 2 };

// ---------------------------------------------------------------------------
// Canonical FP32 distance order "X16" (kgraph-data.h:41-103).
// Elements [0, D16) go to 16 interleaved partials (element i -> partial i%16)
// accumulated with fmaf in increasing i; the partials are reduced as
// xsimd's AVX-512 reduce_add does: u[e] = (p[e]+p[e+4]) + (p[e+8]+p[e+12]),
// r = (u0+u1) + (u2+u3); the tail [D16, D) is then accumulated sequentially
// with fmaf (kgraph-data.h:54-59).
static inline float reduce16(float const *p) {
    float u[4];
    for (int e = 0; e < 4; ++e) u[e] = (p[e] + p[e + 4]) + (p[e + 8] + p[e + 12]);
    return (u[0] + u[1]) + (u[2] + u[3]);
}

static inline float ssd(float const *a, float const *b, unsigned D) {
    float p[16] = {0};
    unsigned const D16 = D & ~15u;
    for (unsigned i = 0; i < D16; i += 16) {
        for (int t = 0; t < 16; ++t) {
            float x = a[i + t] - b[i + t];
            p[t] = std::fma(x, x, p[t]);
        }
    }
    float acc = reduce16(p);
    for (unsigned i = D16; i < D; ++i) {
        float x = a[i] - b[i];
        acc = std::fma(x, x, acc);
    }
    return acc;
}

static inline float dot(float const *a, float const *b, unsigned D) {
    float p[16] = {0};
    unsigned const D16 = D & ~15u;
    for (unsigned i = 0; i < D16; i += 16) {
        for (int t = 0; t < 16; ++t) p[t] = std::fma(a[i + t], b[i + t], p[t]);
    }
    float acc = reduce16(p);
    for (unsigned i = D16; i < D; ++i) acc = std::fma(a[i], b[i], acc);
    return acc;
}

static inline float norm2(float const *a, unsigned D) { return dot(a, a, D); }

// Per-row norm as the metric stores it: python-api.cpp:51-69 (EuclideanLike
// keeps ||x||^2/2, AngularLike keeps ||x||); the CLI's l2sqr keeps none.
static inline float row_norm(int metric, float const *x, unsigned D) {
    if (metric == METRIC_EUCLIDEAN) return norm2(x, D) / 2;
    if (metric == METRIC_ANGULAR) return std::sqrt(norm2(x, D));
    return 0.0f;
}

// Distance between a data row (norm na) and another vector (norm nb):
// kgraph-data.h:41-61 (l2sqr::apply), python-api.cpp:56-58 (n1+n2-dot),
// python-api.cpp:66-68 (-dot/(n1*n2+FLT_EPSILON), product-sum fused).
static inline float metric_dist(int metric, float const *a, float na, float const *b, float nb, unsigned D) {
    if (metric == METRIC_L2SQR) return ssd(a, b, D);
    float d = dot(a, b, D);
    if (metric == METRIC_EUCLIDEAN) return na + nb - d;
    return -d / std::fma(na, nb, std::numeric_limits<float>::epsilon());
}

struct Dense {
    float const *X = nullptr;
    size_t stride = 0;  // floats per row
    unsigned N = 0, D = 0;
    int metric = 0;
    vector<float> nrm;
    void init(float const *x, unsigned n, unsigned d, size_t s, int m) {
        X = x; N = n; D = d; stride = s; metric = m;
        nrm.assign(N, 0.0f);
#pragma omp parallel for
        for (long i = 0; i < long(N); ++i) nrm[i] = row_norm(metric, row(i), D);
    }
    float const *row(size_t i) const { return X + i * stride; }
    // IndexOracle::operator()(i, j) (kgraph.h:57, python-api.cpp:123-125)
    float operator()(unsigned i, unsigned j) const {
        return metric_dist(metric, row(i), nrm[i], row(j), nrm[j], D);
    }
};

struct QueryOracle {  // SearchOracle (kgraph.h:66-86; python-api.cpp:92-108)
    Dense const *d;
    float const *q;
    float qn;
    QueryOracle(Dense const *dd, float const *qq) : d(dd), q(qq) { qn = row_norm(d->metric, q, d->D); }
    unsigned size() const { return d->N; }
    float operator()(unsigned i) const { return metric_dist(d->metric, d->row(i), d->nrm[i], q, qn, d->D); }
};

// ---------------------------------------------------------------------------
// GenRandom (kgraph.cpp:44-65): `size` distinct ids < N.
template <typename RNG>
static void GenRandom(RNG &rng, unsigned *addr, unsigned size, unsigned N) {
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

// Neighbor / NeighborX (kgraph.cpp:67-91); order is by dist only.
struct Neighbor {
    uint32_t id;
    float dist;
    bool flag;
    Neighbor() : id(0), dist(0), flag(false) {}
    Neighbor(unsigned i, float d, bool f = true) : id(i), dist(d), flag(f) {}
};
struct NeighborX : Neighbor {
    uint16_t m, M;
    NeighborX() : m(0), M(0) {}
    NeighborX(unsigned i, float d) : Neighbor(i, d, true), m(0), M(0) {}
};

// UpdateKnnListHelper (kgraph.cpp:176-202): insert after every entry whose
// dist <= nn.dist; reject (return K+1) when an equal-dist entry has the id.
template <typename NT>
static unsigned UpdateKnnList(NT *addr, unsigned K, NT nn) {
    unsigned j;
    unsigned i = K;
    while (i > 0) {
        j = i - 1;
        if (addr[j].dist <= nn.dist) break;
        i = j;
    }
    unsigned l = i;
    while (l > 0) {
        j = l - 1;
        if (addr[j].dist < nn.dist) break;
        if (addr[j].id == nn.id) return K + 1;
        l = j;
    }
    j = K;
    while (j > i) {
        addr[j] = addr[j - 1];
        --j;
    }
    addr[i] = nn;
    return i;
}

// (dist, id) order used where the reference's std::sort is unspecified.
template <typename NT>
static bool dist_id_less(NT const &a, NT const &b) {
    if (a.dist != b.dist) return a.dist < b.dist;
    return a.id < b.id;
}

// LinearSearch (kgraph.cpp:212-229): exact K-NN of row i, self excluded.
static void LinearSearch(Dense const &oracle, unsigned i, unsigned K, vector<Neighbor> *pnns) {
    vector<Neighbor> nns(K + 1);
    unsigned N = oracle.N;
    Neighbor nn;
    nn.id = 0;
    nn.flag = true;
    unsigned k = 0;
    while (nn.id < N) {
        if (nn.id != i) {
            nn.dist = oracle(i, nn.id);
            UpdateKnnList(&nns[0], k, nn);
            if (k < K) ++k;
        }
        ++nn.id;
    }
    nns.resize(K);
    pnns->swap(nns);
}

// SearchOracle::search (kgraph.cpp:231-252): brute force with epsilon.
static unsigned BruteSearch(QueryOracle const &o, unsigned K, float epsilon, unsigned *ids, float *dists) {
    vector<Neighbor> nns(K + 1);
    unsigned N = o.size();
    unsigned L = 0;
    for (unsigned k = 0; k < N; ++k) {
        float kd = o(k);
        if (kd > epsilon) continue;
        UpdateKnnList(&nns[0], L, Neighbor(k, kd));
        if (L < K) ++L;
    }
    for (unsigned k = 0; k < L; ++k) {
        if (ids) ids[k] = nns[k].id;
        if (dists) dists[k] = nns[k].dist;
    }
    return L;
}

// ---------------------------------------------------------------------------
struct IndexParams {  // kgraph.h:94-110
    unsigned iterations, L, K, S, R, controls, seed;
    float delta, recall;
    unsigned prune;
    int reverse;
};
struct IndexInfo {  // kgraph.h:135-147
    int stop_condition;
    unsigned iterations;
    float cost, recall, accuracy, delta, M;
};
struct SearchParams {  // kgraph.h:113-126
    unsigned K, M, P, S, T;
    float epsilon;
    unsigned seed, init;
    bool batched = false;  // one query of a per-query K / P batch
};

static char const *KGRAPH_MAGIC = "KNNGRAPH";  // kgraph.cpp:272-274
static unsigned constexpr KGRAPH_MAGIC_SIZE = 8;
static uint32_t constexpr SIGNATURE_VERSION = 2;
enum { FORMAT_DEFAULT = 0, FORMAT_NO_DIST = 1, FORMAT_TEXT = 128 };

// Planted mistakes (KGO_PLANT, default 0 = none): tests/test_oracle_walk.py
// compiles the oracle once per value and checks that its walk pins reject
// every such copy.  Each value is one plausible misreading of the reference's
// search (kgraph.cpp:423-564):
//   1 no rewind `if (r < k) k = r`                      (:507-512)
//   2 the whole list in one batch, S ignored             (:483-489)
//   3 actual_M ignores SearchParams::M                   (:284-286)
//   4 actual_M ignores the index's M[i] (whole list)     (:284-286)
//   5 trials after the first are not merged              (:516-529)
//   6 the visited bits are cleared between trials        (:438, 452-458)
//   7 a later trial also starts from visited seeds       (:454-457)
//   8 M of an insert at the old end of the pool unset    (:504-506)
#ifndef KGO_PLANT
#define KGO_PLANT 0
#endif

// Everything one search call did, in order (for the walk's invariant tests):
// the ids scored (starting points first, then each fresh neighbour) with
// their distances, and each expansion as (node, beginM, endM).
struct SearchTrace {
    vector<unsigned> visited;
    vector<float> vdist;
    vector<unsigned> expansions;  // 3 per expansion
    vector<unsigned> trial_start;  // index into visited where each trial's seeds begin
};

struct Index {  // KGraphImpl state (kgraph.cpp:276-286)
    vector<unsigned> M;
    vector<vector<Neighbor>> graph;
    bool no_dist = false;
    unsigned actual_M(unsigned pM, unsigned i) const {  // kgraph.cpp:284-286
#if KGO_PLANT == 3
        (void)pM;
        return std::min(M[i], unsigned(graph[i].size()));
#elif KGO_PLANT == 4
        (void)pM;
        return unsigned(graph[i].size());
#else
        return std::min(std::max(M[i], pM), unsigned(graph[i].size()));
#endif
    }

    // KGraphImpl::load (kgraph.cpp:291-331); truncated files throw here.
    void load(char const *path) {
        std::ifstream is(path, std::ios::binary);
        if (!is) throw std::runtime_error("cannot open index file");
        char magic[KGRAPH_MAGIC_SIZE];
        uint32_t sig_version, sig_cap, N;
        is.read(magic, sizeof(magic));
        is.read(reinterpret_cast<char *>(&sig_version), 4);
        is.read(reinterpret_cast<char *>(&sig_cap), 4);
        if (!is) throw std::runtime_error("error reading index file.");
        if (sig_version != SIGNATURE_VERSION) throw std::runtime_error("data version not supported.");
        is.read(reinterpret_cast<char *>(&N), 4);
        if (!is) throw std::runtime_error("error reading index file.");
        if (memcmp(magic, KGRAPH_MAGIC, KGRAPH_MAGIC_SIZE)) throw std::runtime_error("index corrupted.");
        no_dist = sig_cap & FORMAT_NO_DIST;
        graph.assign(N, {});
        M.assign(N, 0);
        vector<uint32_t> nids;
        vector<unsigned char> rec;
        for (unsigned i = 0; i < N; ++i) {
            uint32_t K;
            is.read(reinterpret_cast<char *>(&M[i]), 4);
            is.read(reinterpret_cast<char *>(&K), 4);
            if (!is) throw std::runtime_error("error reading index file.");
            auto &knn = graph[i];
            knn.resize(K);
            if (no_dist) {
                nids.resize(K);
                is.read(reinterpret_cast<char *>(nids.data()), size_t(K) * 4);
                for (unsigned k = 0; k < K; ++k) knn[k] = Neighbor(nids[k], 0, false);
            } else {
                rec.resize(size_t(K) * 12);
                is.read(reinterpret_cast<char *>(rec.data()), rec.size());
                for (unsigned k = 0; k < K; ++k) {
                    memcpy(&knn[k].id, &rec[k * 12], 4);
                    memcpy(&knn[k].dist, &rec[k * 12 + 4], 4);
                    knn[k].flag = rec[k * 12 + 8] != 0;
                }
            }
            if (!is) throw std::runtime_error("error reading index file.");
        }
    }

    // KGraphImpl::save (kgraph.cpp:333-373); flag + padding bytes written 0.
    void save(char const *path, int format) const {
        if (format == FORMAT_TEXT) {
            std::ofstream os(path);
            os << graph.size() << std::endl;
            for (auto const &knn : graph) {
                os << knn.size();
                for (auto const &e : knn) os << ' ' << e.id << ' ' << e.dist;
                os << std::endl;
            }
            return;
        }
        std::ofstream os(path, std::ios::binary);
        uint32_t N = graph.size();
        uint32_t sig_cap = format;
        os.write(KGRAPH_MAGIC, KGRAPH_MAGIC_SIZE);
        os.write(reinterpret_cast<char const *>(&SIGNATURE_VERSION), 4);
        os.write(reinterpret_cast<char const *>(&sig_cap), 4);
        os.write(reinterpret_cast<char const *>(&N), 4);
        for (unsigned i = 0; i < N; ++i) {
            auto const &knn = graph[i];
            uint32_t K = knn.size();
            os.write(reinterpret_cast<char const *>(&M[i]), 4);
            os.write(reinterpret_cast<char const *>(&K), 4);
            for (auto const &e : knn) {
                os.write(reinterpret_cast<char const *>(&e.id), 4);
                if (!(format & FORMAT_NO_DIST)) {
                    unsigned char pad[8] = {0};
                    memcpy(pad, &e.dist, 4);
                    os.write(reinterpret_cast<char const *>(pad), 8);
                }
            }
        }
    }

    // KGraphImpl::search (kgraph.cpp:423-564).  n_comps counts distance
    // evaluations after seeding (kgraph.cpp:500); n_scan counts neighbour ids
    // read (sum of endM-beginM over expansions, kgraph.cpp:485-496).
    unsigned search(QueryOracle const &oracle, SearchParams const &params, unsigned *ids, float *dists,
                    unsigned *p_comps, unsigned *p_scan, SearchTrace *tr = nullptr) const {
        if (graph.size() > oracle.size()) throw std::runtime_error("dataset larger than index");
        if (params.P >= graph.size()) {
            if (p_comps) *p_comps = oracle.size();
            if (p_scan) *p_scan = 0;
            return BruteSearch(oracle, params.K, params.epsilon, ids, dists);
        }
        vector<NeighborX> knn(params.K + params.P + 1);
        vector<NeighborX> results;
        vector<uint8_t> flags(graph.size(), 0);
        if (params.init && params.T > 1) throw std::runtime_error("when init > 0, T must be 1.");
        unsigned seed = params.seed;
        if (seed == 0) seed = time(NULL);
        std::mt19937 rng(seed);
        unsigned n_comps = 0, n_scan = 0;
        for (unsigned trial = 0; trial < params.T; ++trial) {
            unsigned L = params.init;
#if KGO_PLANT == 6
            std::fill(flags.begin(), flags.end(), 0);
#endif
            if (tr) tr->trial_start.push_back(unsigned(tr->visited.size()));
            if (L == 0) {
                vector<unsigned> random(params.P);
                GenRandom(rng, &random[0], random.size(), graph.size());
                for (unsigned s : random)
#if KGO_PLANT == 7
                    knn[L++].id = s;
#else
                    if (!flags[s]) knn[L++].id = s;
#endif
            } else {
                if (!ids) throw std::invalid_argument("no initial data provided via ids");
                // kgraph.cpp:463; per-query K / P batches may start from up to K + P - 1 ids
                if (!(L < (params.batched ? params.K + params.P : params.K)))
                    throw std::invalid_argument("L < params.K");
                for (unsigned l = 0; l < L; ++l) knn[l].id = ids[l];
            }
            for (unsigned k = 0; k < L; ++k) {
                auto &e = knn[k];
                flags[e.id] = 1;
                e.flag = true;
                e.dist = oracle(e.id);
                e.m = 0;
                e.M = actual_M(params.M, e.id);
                if (tr) {
                    tr->visited.push_back(e.id);
                    tr->vdist.push_back(e.dist);
                }
            }
            std::sort(knn.begin(), knn.begin() + L, dist_id_less<NeighborX>);
            unsigned k = 0;
            while (k < L) {
                auto &e = knn[k];
                if (!e.flag) {
                    ++k;
                    continue;
                }
                unsigned beginM = e.m;
#if KGO_PLANT == 2
                unsigned endM = e.M + 1;
#else
                unsigned endM = beginM + params.S;
#endif
                if (endM > e.M) {
                    e.flag = false;
                    endM = e.M;
                }
                e.m = endM;
                auto const &neighbors = graph[e.id];
                if (endM > beginM) n_scan += endM - beginM;
                if (tr) tr->expansions.insert(tr->expansions.end(), {e.id, beginM, endM});
                for (unsigned m = beginM; m < endM; ++m) {
                    unsigned id = neighbors[m].id;
                    if (flags[id]) continue;
                    flags[id] = 1;
                    ++n_comps;
                    float dist = oracle(id);
                    if (tr) {
                        tr->visited.push_back(id);
                        tr->vdist.push_back(dist);
                    }
                    NeighborX nn(id, dist);
                    unsigned r = UpdateKnnList(&knn[0], L, nn);
#if KGO_PLANT == 8
                    unsigned const L_before = L;
#endif
                    if (L + 1 < knn.size()) ++L;
#if KGO_PLANT == 8
                    if (r < L_before) {
#else
                    if (r < L) {
#endif
                        knn[r].M = actual_M(params.M, id);
#if KGO_PLANT != 1
                        if (r < k) k = r;
#endif
                    }
                }
            }
            if (L > params.K) L = params.K;
            if (results.empty()) {
                results.reserve(params.K + 1);
                results.resize(L + 1);
                std::copy(knn.begin(), knn.begin() + L, results.begin());
            } else if (KGO_PLANT == 5) {
            } else {
                for (unsigned l = 0; l < L; ++l) {
                    unsigned r = UpdateKnnList(&results[0], results.size() - 1, knn[l]);
                    if (r < results.size() && results.size() < (params.K + 1)) results.resize(results.size() + 1);
                }
            }
        }
        results.pop_back();
        for (unsigned l = 0; l < results.size(); ++l) {
            if (results[l].dist > params.epsilon) {
                results.resize(l);
                break;
            }
        }
        unsigned L = results.size();
        if (ids)
            for (unsigned k = 0; k < L; ++k) ids[k] = results[k].id;
        if (dists)
            for (unsigned k = 0; k < L; ++k) dists[k] = results[k].dist;
        if (p_comps) *p_comps = n_comps;
        if (p_scan) *p_scan = n_scan;
        return L;
    }

    // KGraphImpl::reverse (kgraph.cpp:667-702); per-node sort by (dist, id)
    // so std::unique removes every duplicate id.
    void reverse(int rev_k) {
        if (rev_k == 0) return;
        if (no_dist) throw std::runtime_error("Need distance information to reverse graph");
        vector<vector<Neighbor>> ng(graph.size());
        for (unsigned i = 0; i < graph.size(); ++i) {
            auto const &v = graph[i];
            unsigned K = M[i];
            if (rev_k > 0) {
                K = rev_k;
                if (K > v.size()) K = v.size();
            }
            for (unsigned j = 0; j < K; ++j) {
                auto const &e = v[j];
                auto re = e;
                re.id = i;
                ng[i].push_back(e);
                ng[e.id].push_back(re);
            }
        }
        graph.swap(ng);
#pragma omp parallel for
        for (long i = 0; i < long(graph.size()); ++i) {
            auto &v = graph[i];
            std::sort(v.begin(), v.end(), dist_id_less<Neighbor>);
            v.resize(std::unique(v.begin(), v.end(), [](Neighbor const &a, Neighbor const &b) { return a.id == b.id; }) -
                     v.begin());
            M[i] = v.size();
        }
    }

    // prune1 / prune2 (kgraph.cpp:584-665).
    void prune1() {
        for (unsigned i = 0; i < graph.size(); ++i)
            if (graph[i].size() > M[i]) graph[i].resize(M[i]);
    }
    void prune2() {
        vector<vector<unsigned>> rev(graph.size());
        vector<unsigned> new_L(graph.size(), 0);
        for (unsigned i = 0; i < graph.size(); ++i) {
            for (auto &e : graph[i]) e.flag = false;
            // the reference requires M[i] <= list size (BOOST_VERIFY, kgraph.cpp:616)
            if (M[i] > graph[i].size()) M[i] = unsigned(graph[i].size());
        }
        vector<unsigned> todo(graph.size());
        for (unsigned i = 0; i < todo.size(); ++i) todo[i] = i;
        vector<unsigned> new_todo;
        for (unsigned l = 0; todo.size(); ++l) {
            new_todo.clear();
            for (unsigned i : todo) {
                if (l >= M[i]) continue;
                new_todo.push_back(i);
                auto &v = graph[i];
                if (v[l].flag) continue;
                v[l].flag = true;
                unsigned T;
                {
                    auto &nl = new_L[i];
                    T = v[nl].id = v[l].id;
                    v[nl].dist = v[l].dist;
                    ++nl;
                }
                rev[T].push_back(i);
                {
                    auto const &u = graph[T];
                    for (unsigned ll = l + 1; ll < M[i]; ++ll) {
                        if (v[ll].flag) continue;
                        for (unsigned j = 0; j < new_L[T]; ++j) {
                            if (v[ll].id == u[j].id) {
                                v[ll].flag = true;
                                break;
                            }
                        }
                    }
                }
                for (auto r : rev[i]) {
                    auto &u = graph[r];
                    for (unsigned ll = l; ll < M[r]; ++ll)
                        if (u[ll].id == T) u[ll].flag = true;
                }
            }
            todo.swap(new_todo);
        }
        M.swap(new_L);
        prune1();
    }
};

// ---------------------------------------------------------------------------
// NN-descent (KGraphConstructor, kgraph.cpp:705-1001).
struct SpinLock {
    std::atomic_flag f = ATOMIC_FLAG_INIT;
    void lock() {
        while (f.test_and_set(std::memory_order_acquire)) {
        }
    }
    void unlock() { f.clear(std::memory_order_release); }
};

struct Nhood {  // kgraph.cpp:709-752
    SpinLock lock;
    float radius;
    float radiusM;
    vector<Neighbor> pool;
    unsigned L;
    unsigned M;
    bool found;
    vector<unsigned> nn_old, nn_new, rnn_old, rnn_new;

    unsigned parallel_try_insert(unsigned id, float dist) {  // kgraph.cpp:723-736
        if (dist > radius) return pool.size();
        std::lock_guard<SpinLock> guard(lock);
        unsigned l = UpdateKnnList(&pool[0], L, Neighbor(id, dist, true));
        if (l <= L) {
            if (L + 1 < pool.size())
                ++L;
            else
                radius = pool[L - 1].dist;
        }
        return l;
    }
    template <typename C>
    void join(C callback) const {  // kgraph.cpp:739-751
        for (unsigned const i : nn_new) {
            for (unsigned const j : nn_new)
                if (i < j) callback(i, j);
            for (unsigned j : nn_old) callback(i, j);
        }
    }
};

struct Control {
    unsigned id;
    vector<Neighbor> neighbors;
};

// EvaluateRecall (kgraph.cpp:96-128), over the pool's valid entries.
static float EvaluateRecall(Neighbor const *pool, unsigned np, vector<Neighbor> const &knn) {
    if (knn.empty()) return 1.0;
    unsigned found = 0, n_p = 0, n_k = 0;
    for (;;) {
        if (n_p >= np) break;
        if (n_k >= knn.size()) break;
        if (knn[n_k].dist < pool[n_p].dist)
            ++n_k;
        else if (knn[n_k].dist == pool[n_p].dist) {
            ++found;
            ++n_k;
            ++n_p;
        } else
            throw std::runtime_error("distance is unstable");
    }
    return float(found) / knn.size();
}
static float EvaluateAccuracy(Neighbor const *pool, unsigned np, vector<Neighbor> const &knn) {  // :130-141
    unsigned m = std::min<unsigned>(np, knn.size());
    float sum = 0;
    unsigned cnt = 0;
    for (unsigned i = 0; i < m; ++i) {
        if (knn[i].dist > 0) {
            sum += std::abs(pool[i].dist - knn[i].dist) / knn[i].dist;
            ++cnt;
        }
    }
    return cnt > 0 ? sum / cnt : 0;
}
static float EvaluateDelta(vector<Neighbor> const &pool, unsigned valid, unsigned K) {  // :148-156
    unsigned c = 0;
    unsigned N = std::min(K, valid);
    for (unsigned i = 0; i < N; ++i)
        if (pool[i].flag) ++c;
    return float(c) / K;
}

struct Builder {
    Dense const &oracle;
    IndexParams params;
    vector<Nhood> nhoods;
    size_t n_comps = 0;
    std::mt19937 shuffle_rng;
    Builder(Dense const &o, IndexParams const &p) : oracle(o), params(p), nhoods(o.N), shuffle_rng(p.seed) {}

    void init() {  // kgraph.cpp:760-796
        unsigned N = oracle.N;
        unsigned seed = params.seed;
        for (auto &nhood : nhoods) {
            nhood.nn_new.resize(params.S * 2);
            nhood.pool.resize(params.L + 1);
            nhood.radius = std::numeric_limits<float>::max();
        }
#pragma omp parallel
        {
#ifdef _OPENMP
            std::mt19937 rng(seed ^ omp_get_thread_num());
#else
            std::mt19937 rng(seed);
#endif
            vector<unsigned> random(params.S + 1);
#pragma omp for
            for (long nl = 0; nl < long(N); ++nl) {
                unsigned n = nl;
                auto &nhood = nhoods[n];
                GenRandom(rng, &nhood.nn_new[0], nhood.nn_new.size(), N);
                GenRandom(rng, &random[0], random.size(), N);
                nhood.L = params.S;
                nhood.M = params.S;
                unsigned i = 0;
                for (unsigned l = 0; l < nhood.L; ++l) {
                    if (random[i] == n) ++i;
                    auto &nn = nhood.pool[l];
                    nn.id = random[i++];
                    nn.dist = oracle(nn.id, n);
                    nn.flag = true;
                }
                std::sort(nhood.pool.begin(), nhood.pool.begin() + nhood.L, dist_id_less<Neighbor>);
            }
        }
    }

    void join() {  // kgraph.cpp:797-815
        size_t cc = 0;
        unsigned N = oracle.N;
        unsigned K = params.K;
#pragma omp parallel for default(shared) schedule(dynamic, 100) reduction(+ : cc)
        for (long nl = 0; nl < long(N); ++nl) {
            unsigned n = nl;
            size_t uu = 0;
            nhoods[n].found = false;
            nhoods[n].join([&](unsigned i, unsigned j) {
                float dist = oracle(i, j);
                ++cc;
                unsigned r = nhoods[i].parallel_try_insert(j, dist);
                if (r < K) ++uu;
                nhoods[j].parallel_try_insert(i, dist);
                if (r < K) ++uu;  // re-tests r, as kgraph.cpp:810 does
            });
            nhoods[n].found = uu > 0;
        }
        n_comps += cc;
    }

    void update() {  // kgraph.cpp:816-883
        unsigned N = oracle.N;
        for (auto &nhood : nhoods) {
            nhood.nn_new.clear();
            nhood.nn_old.clear();
            nhood.rnn_new.clear();
            nhood.rnn_old.clear();
            nhood.radius = (nhood.L + 1 == nhood.pool.size()) ? nhood.pool[nhood.L - 1].dist
                                                               : std::numeric_limits<float>::max();
        }
#pragma omp parallel for
        for (long nl = 0; nl < long(N); ++nl) {
            auto &nhood = nhoods[nl];
            if (nhood.found) {
                unsigned maxl = std::min(nhood.M + params.S, nhood.L);
                unsigned c = 0, l = 0;
                while ((l < maxl) && (c < params.S)) {
                    if (nhood.pool[l].flag) ++c;
                    ++l;
                }
                nhood.M = l;
            }
            nhood.radiusM = nhood.pool[nhood.M - 1].dist;
        }
#pragma omp parallel for
        for (long nl = 0; nl < long(N); ++nl) {
            unsigned n = nl;
            auto &nhood = nhoods[n];
            for (unsigned l = 0; l < nhood.M; ++l) {
                auto &nn = nhood.pool[l];
                auto &nhood_o = nhoods[nn.id];
                if (nn.flag) {
                    nhood.nn_new.push_back(nn.id);
                    if (nn.dist > nhood_o.radiusM) {
                        std::lock_guard<SpinLock> guard(nhood_o.lock);
                        nhood_o.rnn_new.push_back(n);
                    }
                    nn.flag = false;
                } else {
                    nhood.nn_old.push_back(nn.id);
                    if (nn.dist > nhood_o.radiusM) {
                        std::lock_guard<SpinLock> guard(nhood_o.lock);
                        nhood_o.rnn_old.push_back(n);
                    }
                }
            }
        }
        for (unsigned i = 0; i < N; ++i) {
            auto &h = nhoods[i];
            if (params.R && (h.rnn_new.size() > params.R)) {
                std::shuffle(h.rnn_new.begin(), h.rnn_new.end(), shuffle_rng);
                h.rnn_new.resize(params.R);
            }
            h.nn_new.insert(h.nn_new.end(), h.rnn_new.begin(), h.rnn_new.end());
            if (params.R && (h.rnn_old.size() > params.R)) {
                std::shuffle(h.rnn_old.begin(), h.rnn_old.end(), shuffle_rng);
                h.rnn_old.resize(params.R);
            }
            h.nn_old.insert(h.nn_old.end(), h.rnn_old.begin(), h.rnn_old.end());
        }
    }

    // Constructor body (kgraph.cpp:886-999).
    void run(Index *out, IndexInfo *pinfo, vector<float> *trace) {
        unsigned N = oracle.N;
        if (N <= params.K) throw std::runtime_error("K larger than dataset size");
        if (N < params.controls) params.controls = N;
        if (N <= params.L) params.L = N - 1;
        if (N <= params.S) params.S = N - 1;
        // GenerateControl (kgraph.cpp:254-270)
        vector<Control> controls(params.controls);
        {
            vector<unsigned> index(N);
            for (unsigned i = 0; i < N; ++i) index[i] = i;
            std::mt19937 crng(params.seed);
            std::shuffle(index.begin(), index.end(), crng);
#pragma omp parallel for
            for (long i = 0; i < long(params.controls); ++i) {
                controls[i].id = index[i];
                LinearSearch(oracle, index[i], params.K, &controls[i].neighbors);
            }
        }
        init();
        float total = N * float(N - 1) / 2;
        IndexInfo info;
        info.stop_condition = 0;
        info.recall = 0;
        info.accuracy = std::numeric_limits<float>::max();
        info.cost = 0;
        info.iterations = 0;
        info.delta = 1.0;
        info.M = 0;
        for (unsigned it = 0; (params.iterations <= 0) || (it < params.iterations); ++it) {
            ++info.iterations;
            join();
            info.cost = n_comps / total;
            double sM = 0, sdelta = 0, srecall = 0, sacc = 0;
            for (auto const &nhood : nhoods) {
                sM += nhood.M;
                sdelta += EvaluateDelta(nhood.pool, nhood.L, params.K);
            }
            for (auto const &c : controls) {
                auto const &h = nhoods[c.id];
                srecall += EvaluateRecall(&h.pool[0], h.L, c.neighbors);
                sacc += EvaluateAccuracy(&h.pool[0], h.L, c.neighbors);
            }
            info.delta = sdelta / N;
            info.recall = controls.empty() ? 0 : srecall / controls.size();
            info.accuracy = controls.empty() ? 0 : sacc / controls.size();
            info.M = sM / N;
            if (trace) {
                trace->push_back(info.recall);
                trace->push_back(info.delta);
                trace->push_back(info.M);
                trace->push_back(info.cost);
            }
            if (info.delta <= params.delta) {
                info.stop_condition = 1;
                break;
            }
            if (info.recall >= params.recall) {
                info.stop_condition = 2;
                break;
            }
            update();
        }
        if (params.prune > 2) throw std::runtime_error("prune level not supported.");
        out->M.assign(N, 0);
        out->graph.assign(N, {});
        for (unsigned n = 0; n < N; ++n) {
            out->M[n] = nhoods[n].M;
            auto const &pool = nhoods[n].pool;
            unsigned K = std::min(params.L, nhoods[n].L);
            out->graph[n].assign(pool.begin(), pool.begin() + K);
            for (auto &e : out->graph[n]) e.flag = false;
        }
        out->no_dist = false;
        nhoods.clear();
        if (params.reverse) out->reverse(params.reverse);
        if (params.prune & 1) out->prune1();
        if (params.prune & 2) out->prune2();
        if (pinfo) *pinfo = info;
    }
};

}  // namespace kgo

// ---------------------------------------------------------------------------
// C interface for ctypes (tests/ and bench.py only).
using namespace kgo;

struct kgo_index {
    Index idx;
    std::string err;
};

extern "C" {

struct kgo_index_params {
    unsigned iterations, L, K, S, R, controls, seed;
    float delta, recall;
    unsigned prune;
    int reverse;
};
struct kgo_index_info {
    int stop_condition;
    unsigned iterations;
    float cost, recall, accuracy, delta, M;
};
struct kgo_search_params {
    unsigned K, M, P, S, T;
    float epsilon;
    unsigned seed, init;
};

float kgo_dist(int metric, float const *a, float const *b, unsigned D) {
    return metric_dist(metric, a, row_norm(metric, a, D), b, row_norm(metric, b, D), D);
}

void kgo_row_norms(int metric, float const *X, unsigned N, unsigned D, size_t stride, float *out) {
    for (unsigned i = 0; i < N; ++i) out[i] = row_norm(metric, X + i * stride, D);
}

// Raw mt19937 outputs and GenRandom, for pinning against numpy's MT19937.
void kgo_mt19937(unsigned seed, unsigned n, unsigned *out) {
    std::mt19937 rng(seed);
    for (unsigned i = 0; i < n; ++i) out[i] = rng();
}
void kgo_genrandom(unsigned seed, unsigned trials, unsigned size, unsigned N, unsigned *out) {
    std::mt19937 rng(seed);
    for (unsigned t = 0; t < trials; ++t) GenRandom(rng, out + size_t(t) * size, size, N);
}

// UpdateKnnList on (id, dist) arrays of capacity K+1; returns its result.
unsigned kgo_update_knn_list(unsigned *ids, float *dists, unsigned K, unsigned id, float dist) {
    vector<Neighbor> a(K + 1);
    for (unsigned i = 0; i <= K; ++i) a[i] = Neighbor(ids[i], dists[i]);
    unsigned r = UpdateKnnList(&a[0], K, Neighbor(id, dist));
    for (unsigned i = 0; i <= K; ++i) {
        ids[i] = a[i].id;
        dists[i] = a[i].dist;
    }
    return r;
}

kgo_index *kgo_index_new() { return new kgo_index(); }
void kgo_index_free(kgo_index *h) { delete h; }
char const *kgo_index_error(kgo_index *h) { return h->err.c_str(); }

#define KGO_TRY(...)                    \
    try {                               \
        __VA_ARGS__;                    \
        return 0;                       \
    } catch (std::exception const &e) { \
        h->err = e.what();              \
        return -1;                      \
    }

int kgo_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    return 1;
#endif
}

int kgo_index_build(kgo_index *h, float const *X, unsigned N, unsigned D, size_t stride, int metric,
                    kgo_index_params const *p, kgo_index_info *info, float *trace, unsigned trace_cap) {
    KGO_TRY({
        Dense d;
        d.init(X, N, D, stride, metric);
        IndexParams ip{p->iterations, p->L, p->K, p->S, p->R, p->controls, p->seed, p->delta, p->recall, p->prune, p->reverse};
        Builder b(d, ip);
        IndexInfo ii;
        vector<float> tr;
        b.run(&h->idx, &ii, &tr);
        if (info) *info = kgo_index_info{ii.stop_condition, ii.iterations, ii.cost, ii.recall, ii.accuracy, ii.delta, ii.M};
        if (trace)
            for (unsigned i = 0; i < std::min<size_t>(trace_cap, tr.size()); ++i) trace[i] = tr[i];
    })
}

// Batch search.  Ks/Ps may be NULL (uniform).  Output row q starts at
// ids + q*ostride; counts[q] = #results.  comps/scans may be NULL.
int kgo_index_search(kgo_index *h, float const *X, unsigned N, unsigned D, size_t stride, int metric,
                     float const *Q, unsigned nq, size_t qstride, kgo_search_params const *sp,
                     unsigned const *Ks, unsigned const *Ps, unsigned *ids, float *dists, size_t ostride,
                     unsigned *counts, unsigned *comps, unsigned *scans) {
    try {
        Dense d;
        d.init(X, N, D, stride, metric);
        std::string err;
        std::atomic<int> bad(0);
#pragma omp parallel for schedule(dynamic, 8)
        for (long q = 0; q < long(nq); ++q) {
            try {
                SearchParams s{sp->K, sp->M, sp->P, sp->S, sp->T, sp->epsilon, sp->seed, sp->init};
                if (Ks) s.K = Ks[q];
                if (Ps) s.P = Ps[q];
                // per-query K / P (an extension with no counterpart in the
                // reference): a query walks from min(init, K + P - 1) of its ids
                if (Ks && s.init) s.init = std::min(s.init, s.K + s.P - 1);
                s.batched = Ks != nullptr;
                QueryOracle o(&d, Q + q * qstride);
                unsigned c = 0, sc = 0;
                unsigned L = h->idx.search(o, s, ids + q * ostride, dists ? dists + q * ostride : nullptr, &c, &sc);
                counts[q] = L;
                if (comps) comps[q] = c;
                if (scans) scans[q] = sc;
            } catch (std::exception const &e) {
#pragma omp critical
                err = e.what();
                bad = 1;
            }
        }
        if (bad) throw std::runtime_error(err);
        return 0;
    } catch (std::exception const &e) {
        h->err = e.what();
        return -1;
    }
}

// One query's search with its trace (SearchTrace): visited ids / distances
// in order (up to vcap), expansions as (node, beginM, endM) triples (up to
// ecap), trial start positions in `visited` (up to tcap).  Totals go to
// *nv / *ne / *nt; the return value is the result count (or -1).
int kgo_index_search_trace(kgo_index *h, float const *X, unsigned N, unsigned D, size_t stride, int metric,
                           float const *q, kgo_search_params const *sp, unsigned *ids, float *dists,
                           unsigned *comps, unsigned *scans, unsigned *visited, float *vdist, unsigned vcap,
                           unsigned *nv, unsigned *exps, unsigned ecap, unsigned *ne, unsigned *tstart,
                           unsigned tcap, unsigned *nt) {
    try {
        Dense d;
        d.init(X, N, D, stride, metric);
        SearchParams s{sp->K, sp->M, sp->P, sp->S, sp->T, sp->epsilon, sp->seed, sp->init};
        QueryOracle o(&d, q);
        SearchTrace tr;
        unsigned L = h->idx.search(o, s, ids, dists, comps, scans, &tr);
        *nv = unsigned(tr.visited.size());
        *ne = unsigned(tr.expansions.size() / 3);
        *nt = unsigned(tr.trial_start.size());
        for (unsigned i = 0; i < std::min<size_t>(vcap, tr.visited.size()); ++i) {
            visited[i] = tr.visited[i];
            vdist[i] = tr.vdist[i];
        }
        for (unsigned i = 0; i < std::min<size_t>(3ull * ecap, tr.expansions.size()); ++i) exps[i] = tr.expansions[i];
        for (unsigned i = 0; i < std::min<size_t>(tcap, tr.trial_start.size()); ++i) tstart[i] = tr.trial_start[i];
        return int(L);
    } catch (std::exception const &e) {
        h->err = e.what();
        return -1;
    }
}

// Exact top-K by brute force (SearchOracle::search, kgraph.cpp:231-252).
int kgo_bruteforce(float const *X, unsigned N, unsigned D, size_t stride, int metric, float const *Q, unsigned nq,
                   size_t qstride, unsigned K, float epsilon, unsigned *ids, float *dists, unsigned *counts) {
    Dense d;
    d.init(X, N, D, stride, metric);
#pragma omp parallel for schedule(dynamic, 4)
    for (long q = 0; q < long(nq); ++q) {
        QueryOracle o(&d, Q + q * qstride);
        counts[q] = BruteSearch(o, K, epsilon, ids + q * K, dists + q * K);
    }
    return 0;
}

// Exact K-NN of data rows (LinearSearch, kgraph.cpp:212-229), self excluded.
int kgo_linear_knn(float const *X, unsigned N, unsigned D, size_t stride, int metric, unsigned const *rows,
                   unsigned nr, unsigned K, unsigned *ids, float *dists) {
    Dense d;
    d.init(X, N, D, stride, metric);
#pragma omp parallel for schedule(dynamic, 4)
    for (long r = 0; r < long(nr); ++r) {
        vector<Neighbor> nn;
        LinearSearch(d, rows[r], K, &nn);
        for (unsigned k = 0; k < K; ++k) {
            ids[r * K + k] = nn[k].id;
            dists[r * K + k] = nn[k].dist;
        }
    }
    return 0;
}

unsigned kgo_index_size(kgo_index *h) { return h->idx.graph.size(); }
size_t kgo_index_edges(kgo_index *h) {
    size_t s = 0;
    for (auto const &v : h->idx.graph) s += v.size();
    return s;
}
// CSR export: offsets[N+1], ids/dists[edges], M[N].
void kgo_index_export(kgo_index *h, unsigned long long *offsets, unsigned *ids, float *dists, unsigned *M) {
    size_t o = 0;
    for (size_t i = 0; i < h->idx.graph.size(); ++i) {
        offsets[i] = o;
        M[i] = h->idx.M[i];
        for (auto const &e : h->idx.graph[i]) {
            ids[o] = e.id;
            if (dists) dists[o] = e.dist;
            ++o;
        }
    }
    offsets[h->idx.graph.size()] = o;
}
void kgo_index_import(kgo_index *h, unsigned N, unsigned long long const *offsets, unsigned const *ids,
                      float const *dists, unsigned const *M) {
    h->idx.graph.assign(N, {});
    h->idx.M.assign(M, M + N);
    h->idx.no_dist = dists == nullptr;
    for (unsigned i = 0; i < N; ++i)
        for (size_t o = offsets[i]; o < offsets[i + 1]; ++o) h->idx.graph[i].push_back(Neighbor(ids[o], dists ? dists[o] : 0, false));
}
int kgo_index_save(kgo_index *h, char const *path, int format) { KGO_TRY(h->idx.save(path, format)) }
int kgo_index_load(kgo_index *h, char const *path) { KGO_TRY(h->idx.load(path)) }
int kgo_index_reverse(kgo_index *h, int rev) { KGO_TRY(h->idx.reverse(rev)) }
int kgo_index_prune(kgo_index *h, unsigned level) {
    KGO_TRY({
        if (level & 1) h->idx.prune1();
        if (level & 2) h->idx.prune2();
    })
}

}  // extern "C"

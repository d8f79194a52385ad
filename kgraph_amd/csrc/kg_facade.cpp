// kg_facade.cpp -- the kgraph::KGraph C++ interface (include/kgraph.h,
// reference kgraph.h:91-228) implemented over the C ABI of libkgraph_amd.so.
// Dense oracles (kgraph::DenseOracleInfo / DenseQueryInfo) hand their rows to
// the GPU; statuses come back as the reference's exception types.
//
// Device rows.  An index keeps a device copy of the rows it last saw and a
// key for them: row pointer, N, D, stride, metric and a fingerprint of 16
// evenly spaced rows (a different data set loaded into the same buffer gives
// a different key).  SearchOracle::search (the brute-force scan) keeps such
// copies process-wide, one per key, so a harness that calls it per query
// uploads the data set once.  kgraph::forget_device_rows() drops every key.
//
// Concurrent search.  The reference's harnesses call KGraph::search from many
// OpenMP threads at once (search.cpp:139-146, python-api.cpp:292-301).  Here
// each call queues its query; the first caller to find no batch running
// becomes the leader, takes every queued query, runs them as batched launches
// (one per group of compatible SearchParams, with per-query K and P) and hands
// each caller its results; callers that arrive meanwhile form the next batch.
// A query's results do not depend on the batch it runs in (each walk is
// independent and draws its seeds from mt19937(seed) per P, as a single call
// does).
#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <cstdint>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/kgraph.h"
#include "../../include/kgraph_amd.h"

namespace kgraph {

unsigned verbosity = default_verbosity;

namespace {

std::atomic<uint64_t> rows_epoch{0};  // advanced by forget_device_rows()

void check(kg_handle *h, int rc) {
    if (rc == KG_OK) return;
    std::string msg = kg_last_error(h);
    if (rc == KG_ERR_INVALID_ARGUMENT) throw invalid_argument(msg);
    if (rc == KG_ERR_IO) throw io_error(msg);
    throw runtime_error(msg);
}

DenseOracleInfo const &dense_of(IndexOracle const &o) {
    auto *d = dynamic_cast<DenseOracleInfo const *>(&o);
    if (!d || d->dense_metric() < 0)
        throw invalid_argument("the GPU index needs a dense float oracle (kgraph::DenseOracleInfo)");
    return *d;
}

DenseQueryInfo const &dense_of(SearchOracle const &o) {
    auto *d = dynamic_cast<DenseQueryInfo const *>(&o);
    if (!d || d->dense_data().dense_metric() < 0)
        throw invalid_argument("the GPU index needs a dense float search oracle (kgraph::DenseQueryInfo)");
    return *d;
}

// A 64-bit hash (FNV-1a steps over 32-bit words) of 16 evenly spaced rows
// (all rows when fewer): about a microsecond at D = 128.
uint64_t fingerprint(float const *rows, unsigned N, unsigned D, size_t stride) {
    uint64_t h = 1469598103934665603ull;
    if (!rows || N == 0) return h;
    unsigned const samples = std::min(N, 16u);
    for (unsigned s = 0; s < samples; ++s) {
        size_t const i = samples == 1 ? 0 : size_t(s) * (N - 1) / (samples - 1);
        float const *r = rows + i * stride;
        for (unsigned k = 0; k < D; ++k) {
            uint32_t w;
            std::memcpy(&w, r + k, 4);
            h = (h ^ w) * 1099511628211ull;
        }
    }
    return h;
}

struct RowsKey {
    float const *rows = nullptr;
    unsigned N = 0, D = 0;
    size_t stride = 0;
    int metric = -1;
    uint64_t print = 0, epoch = ~0ull;
    // print = false leaves the fingerprint for later (stamp)
    static RowsKey of(DenseOracleInfo const &d, unsigned N, bool print = true) {
        RowsKey k;
        k.rows = d.dense_rows();
        k.N = N;
        k.D = d.dense_dim();
        k.stride = d.dense_stride();
        k.metric = d.dense_metric();
        k.epoch = rows_epoch.load();
        if (print) k.stamp();
        return k;
    }
    void stamp() { print = fingerprint(rows, N, D, stride); }
    auto tie() const { return std::tie(rows, N, D, stride, metric, print, epoch); }
    bool operator==(RowsKey const &o) const { return tie() == o.tie(); }
    bool operator<(RowsKey const &o) const { return tie() < o.tie(); }
};

// Uploads the rows to h unless `have` already holds them.
void ensure_rows(kg_handle *h, RowsKey &have, RowsKey const &want) {
    if (want == have) return;
    kg_set_verbosity(verbosity);
    have = RowsKey{};
    check(h, kg_set_data(h, want.rows, want.N, want.D, want.stride, want.metric));
    have = want;
}

// One caller's query waiting for (or in) a batch.
struct Pending {
    float const *q;
    RowsKey rows;
    kg_search_params sp;
    unsigned *ids;
    float *dists;
    unsigned count = 0, comps = 0;
    int rc = KG_OK;
    std::string err;
    bool done = false;
};

class KGraphAMD : public KGraph {
    kg_handle *h;
    mutable std::mutex mu;  // one GPU call at a time per index
    mutable RowsKey uploaded;
    // search batching
    mutable std::mutex qmu;
    mutable std::condition_variable qcv;
    mutable std::vector<Pending *> queue;
    mutable bool leading = false;

    unsigned index_size() const {
        unsigned n = 0;
        uint64_t e = 0;
        if (kg_index_size(h, &n, &e) != KG_OK) return 0;
        return n;
    }

    // Runs one group of compatible queries (same rows and SearchParams except
    // K and P) as one launch; fills each Pending.  Called with mu held.
    void run_group(std::vector<Pending *> const &g) const {
        Pending const &f = *g.front();
        ensure_rows(h, uploaded, f.rows);
        unsigned const nq = unsigned(g.size()), D = f.rows.D;
        bool uniform = true;
        unsigned width = 1;
        for (Pending *p : g) {
            uniform &= p->sp.K == f.sp.K && p->sp.P == f.sp.P;
            width = std::max(width, p->sp.K);
        }
        std::vector<float> Q(size_t(nq) * D);
        std::vector<unsigned> Ks(nq), Ps(nq), ids(size_t(nq) * width, 0u), counts(nq), comps(nq);
        std::vector<float> dists(size_t(nq) * width, 0.f);
        for (unsigned i = 0; i < nq; ++i) {
            Pending const &p = *g[i];
            std::memcpy(&Q[size_t(i) * D], p.q, sizeof(float) * D);
            Ks[i] = p.sp.K;
            Ps[i] = p.sp.P;
            if (p.sp.init) std::memcpy(&ids[size_t(i) * width], p.ids, sizeof(unsigned) * p.sp.init);
        }
        kg_search_params sp = f.sp;
        int const rc = kg_search(h, Q.data(), nq, D, &sp, uniform ? nullptr : Ks.data(), uniform ? nullptr : Ps.data(),
                                 ids.data(), dists.data(), width, counts.data(), comps.data(), nullptr);
        std::string const err = rc == KG_OK ? std::string() : std::string(kg_last_error(h));
        for (unsigned i = 0; i < nq; ++i) {
            Pending &p = *g[i];
            p.rc = rc;
            p.err = err;
            if (rc != KG_OK) continue;
            p.count = counts[i];
            p.comps = comps[i];
            if (p.ids) std::memcpy(p.ids, &ids[size_t(i) * width], sizeof(unsigned) * p.count);
            if (p.dists) std::memcpy(p.dists, &dists[size_t(i) * width], sizeof(float) * p.count);
        }
    }

    // Splits a batch into launch groups: same rows, M, S, T, epsilon, seed
    // and init; brute-force queries (P >= N) also by K (one K per scan).
    void run_batch(std::vector<Pending *> const &batch) const {
        std::lock_guard<std::mutex> g(mu);
        unsigned const N = index_size();
        // fingerprints once per data set of the batch
        std::map<RowsKey, uint64_t> prints;
        for (Pending *p : batch) {
            auto it = prints.find(p->rows);
            if (it == prints.end()) {
                RowsKey k = p->rows;
                k.stamp();
                it = prints.emplace(p->rows, k.print).first;
            }
            p->rows.print = it->second;
        }
        std::map<std::tuple<RowsKey, unsigned, unsigned, unsigned, float, unsigned, unsigned, unsigned>,
                 std::vector<Pending *>>
            groups;
        for (Pending *p : batch) {
            kg_search_params const &s = p->sp;
            unsigned const bruteK = s.P >= N ? s.K : 0u;
            groups[std::make_tuple(p->rows, s.M, s.S, s.T, s.epsilon, s.seed, s.init, bruteK)].push_back(p);
        }
        for (auto &kv : groups) {
            try {
                run_group(kv.second);
            } catch (std::exception const &e) {
                for (Pending *p : kv.second) {
                    p->rc = KG_ERR_RUNTIME;
                    p->err = e.what();
                }
            }
        }
    }

public:
    KGraphAMD() : h(kg_create()) {
        if (!h) throw runtime_error("error creating kgraph instance");
    }
    ~KGraphAMD() override { kg_destroy(h); }

    void load(char const *path) override {
        std::lock_guard<std::mutex> g(mu);
        check(h, kg_load(h, path));
    }
    void save(char const *path, int format) const override {
        std::lock_guard<std::mutex> g(mu);
        check(h, kg_save(h, path, format));
    }
    void build(IndexOracle const &oracle, IndexParams const &params, IndexInfo *info) override {
        std::lock_guard<std::mutex> g(mu);
        ensure_rows(h, uploaded, RowsKey::of(dense_of(oracle), oracle.size()));
        kg_index_params p{params.iterations, params.L, params.K, params.S, params.R, params.controls,
                          params.seed, params.delta, params.recall, params.prune, params.reverse};
        kg_index_info ii{};
        kg_set_verbosity(verbosity);
        check(h, kg_build(h, &p, &ii));
        if (info) {
            info->stop_condition = IndexInfo::StopCondition(ii.stop_condition);
            info->iterations = ii.iterations;
            info->cost = ii.cost;
            info->recall = ii.recall;
            info->accuracy = ii.accuracy;
            info->delta = ii.delta;
            info->M = ii.M;
        }
    }
    void prune(IndexOracle const &, unsigned level) override {
        std::lock_guard<std::mutex> g(mu);
        check(h, kg_prune(h, level));
    }
    unsigned search(SearchOracle const &oracle, SearchParams const &params, unsigned *ids, float *dists,
                    SearchInfo *info) const override {
        DenseQueryInfo const &q = dense_of(oracle);
        // the reference's argument checks (kgraph.cpp:440-442, 462-463), per caller
        if (params.init && params.T > 1) throw runtime_error("when init > 0, T must be 1.");
        if (params.init && !ids) throw invalid_argument("no initial data provided via ids");
        if (params.init && !(params.init < params.K)) throw invalid_argument("L < params.K");
        if (params.S == 0) throw invalid_argument("S must be >= 1");
        if (params.T == 0) throw invalid_argument("T must be >= 1");
        if (params.init) {
            std::lock_guard<std::mutex> g(mu);
            unsigned const N = index_size();
            for (unsigned l = 0; l < params.init; ++l)
                if (ids[l] >= N) throw invalid_argument("initial id out of range");
        }
        Pending me;
        me.q = q.dense_query();
        me.rows = RowsKey::of(q.dense_data(), oracle.size(), false);
        me.sp = kg_search_params{params.K, params.M, params.P, params.S, params.T, params.epsilon, params.seed,
                                 params.init};
        me.ids = ids;
        me.dists = dists;
        std::unique_lock<std::mutex> lk(qmu);
        queue.push_back(&me);
        while (!me.done) {
            if (!leading) {
                leading = true;
                std::vector<Pending *> batch;
                batch.swap(queue);
                lk.unlock();
                run_batch(batch);
                lk.lock();
                for (Pending *p : batch) p->done = true;
                leading = false;
                qcv.notify_all();
            } else {
                qcv.wait(lk);
            }
        }
        lk.unlock();
        if (me.rc != KG_OK) {
            if (me.rc == KG_ERR_INVALID_ARGUMENT) throw invalid_argument(me.err);
            throw runtime_error(me.err);
        }
        if (info) {
            info->updates = 0;
            info->cost = float(me.comps) / oracle.size();
        }
        return me.count;
    }
    void get_nn(unsigned id, unsigned *nns, float *dists, unsigned *M, unsigned *L) const override {
        std::lock_guard<std::mutex> g(mu);
        check(h, kg_get_nn(h, id, nns, dists, M, L));
    }
    void reverse(int rev) override {
        std::lock_guard<std::mutex> g(mu);
        kg_set_verbosity(verbosity);
        check(h, kg_reverse(h, rev));
    }
};

// Device copies of data sets for SearchOracle::search, one handle per rows
// key, kept for the process (a handle holds its data set's rows and norms).
struct ScanCache {
    std::mutex mu;
    std::map<RowsKey, std::unique_ptr<std::mutex>> locks;
    std::map<RowsKey, kg_handle *> handles;
    ~ScanCache() {
        for (auto &kv : handles) kg_destroy(kv.second);
    }
    // the handle for `k` (created and filled on first use) and its lock
    std::pair<kg_handle *, std::mutex *> get(RowsKey const &k) {
        std::lock_guard<std::mutex> g(mu);
        auto it = handles.find(k);
        if (it != handles.end()) return {it->second, locks[k].get()};
        // keys from before forget_device_rows() are dropped here
        for (auto i = handles.begin(); i != handles.end();) {
            if (i->first.epoch != k.epoch) {
                kg_destroy(i->second);
                locks.erase(i->first);
                i = handles.erase(i);
            } else {
                ++i;
            }
        }
        kg_handle *h = kg_create();
        if (!h) throw runtime_error("error creating kgraph instance");
        int const rc = kg_set_data(h, k.rows, k.N, k.D, k.stride, k.metric);
        if (rc != KG_OK) {
            std::string msg = kg_last_error(h);
            kg_destroy(h);
            if (rc == KG_ERR_INVALID_ARGUMENT) throw invalid_argument(msg);
            throw runtime_error(msg);
        }
        handles[k] = h;
        locks[k].reset(new std::mutex);
        return {h, locks[k].get()};
    }
};

ScanCache &scan_cache() {
    static ScanCache *c = new ScanCache;  // lives to the end of the process
    return *c;
}

}  // namespace

void forget_device_rows() { ++rows_epoch; }

// SearchOracle::search (kgraph.cpp:231-252): brute force over the dataset, on
// a cached device copy of its rows.
unsigned SearchOracle::search(unsigned K, float epsilon, unsigned *ids, float *dists) const {
    DenseQueryInfo const &q = dense_of(*this);
    DenseOracleInfo const &d = q.dense_data();
    auto hl = scan_cache().get(RowsKey::of(d, size()));
    std::vector<unsigned> out(std::max(K, 1u));
    std::vector<float> od(std::max(K, 1u));
    unsigned count = 0;
    {
        std::lock_guard<std::mutex> g(*hl.second);
        check(hl.first, kg_bruteforce(hl.first, q.dense_query(), 1, d.dense_dim(), K, epsilon, out.data(), od.data(),
                                      out.size(), &count));
    }
    if (ids) std::memcpy(ids, out.data(), sizeof(unsigned) * count);
    if (dists) std::memcpy(dists, od.data(), sizeof(float) * count);
    return count;
}

KGraph *KGraph::create() { return new KGraphAMD(); }

char const *KGraph::version() { return kg_version(); }

}  // namespace kgraph

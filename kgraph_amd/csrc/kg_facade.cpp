// kg_facade.cpp -- the kgraph::KGraph C++ interface (include/kgraph.h,
// reference kgraph.h:91-228) implemented over the C ABI of libkgraph_amd.so.
// Dense oracles (kgraph::DenseOracleInfo / DenseQueryInfo) hand their rows to
// the GPU; statuses come back as the reference's exception types.
#include <cstring>
#include <mutex>
#include <vector>

#include "../../include/kgraph.h"
#include "../../include/kgraph_amd.h"

namespace kgraph {

unsigned verbosity = default_verbosity;

namespace {

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

struct DataKey {
    float const *rows = nullptr;
    unsigned N = 0, D = 0;
    size_t stride = 0;
    int metric = -1;
    bool operator==(DataKey const &o) const {
        return rows == o.rows && N == o.N && D == o.D && stride == o.stride && metric == o.metric;
    }
};

class KGraphAMD : public KGraph {
    kg_handle *h;
    mutable std::mutex mu;  // one GPU call at a time per index (search is const and re-entrant)
    mutable DataKey uploaded;

    void ensure_data(DenseOracleInfo const &d, unsigned N) const {
        DataKey k{d.dense_rows(), N, d.dense_dim(), d.dense_stride(), d.dense_metric()};
        if (k == uploaded) return;
        kg_set_verbosity(verbosity);
        check(h, kg_set_data(h, k.rows, k.N, k.D, k.stride, k.metric));
        uploaded = k;
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
        ensure_data(dense_of(oracle), oracle.size());
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
        std::lock_guard<std::mutex> g(mu);
        DenseQueryInfo const &q = dense_of(oracle);
        ensure_data(q.dense_data(), oracle.size());
        kg_search_params sp{params.K, params.M, params.P, params.S, params.T, params.epsilon, params.seed, params.init};
        size_t const width = std::max<size_t>(params.K, 1);
        std::vector<unsigned> out(width, 0);
        if (params.init) {
            if (!ids) throw invalid_argument("no initial data provided via ids");
            std::memcpy(out.data(), ids, sizeof(unsigned) * std::min<size_t>(params.init, width));
        }
        std::vector<float> od(width, 0.f);
        unsigned count = 0, comps = 0;
        check(h, kg_search(h, q.dense_query(), 1, q.dense_data().dense_dim(), &sp, nullptr, nullptr, out.data(),
                           od.data(), width, &count, &comps, nullptr));
        if (ids) std::memcpy(ids, out.data(), sizeof(unsigned) * count);
        if (dists) std::memcpy(dists, od.data(), sizeof(float) * count);
        if (info) {
            info->updates = 0;
            info->cost = float(comps) / oracle.size();
        }
        return count;
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

}  // namespace

// SearchOracle::search (kgraph.cpp:231-252): brute force over the dataset.
unsigned SearchOracle::search(unsigned K, float epsilon, unsigned *ids, float *dists) const {
    DenseQueryInfo const &q = dense_of(*this);
    DenseOracleInfo const &d = q.dense_data();
    kg_handle *h = kg_create();
    if (!h) throw runtime_error("error creating kgraph instance");
    struct Guard {
        kg_handle *h;
        ~Guard() { kg_destroy(h); }
    } guard{h};
    check(h, kg_set_data(h, d.dense_rows(), size(), d.dense_dim(), d.dense_stride(), d.dense_metric()));
    std::vector<unsigned> out(std::max(K, 1u));
    std::vector<float> od(std::max(K, 1u));
    unsigned count = 0;
    check(h, kg_bruteforce(h, q.dense_query(), 1, d.dense_dim(), K, epsilon, out.data(), od.data(), out.size(),
                           &count));
    if (ids) std::memcpy(ids, out.data(), sizeof(unsigned) * count);
    if (dists) std::memcpy(dists, od.data(), sizeof(float) * count);
    return count;
}

KGraph *KGraph::create() { return new KGraphAMD(); }

char const *KGraph::version() { return kg_version(); }

}  // namespace kgraph

// search -- query a KNNGRAPH index with lshkit files, the command-line tool of
// the reference (search.cpp:29-185) on the GPU.
//
//   search [OPTIONS] DATA INDEX QUERY [OUTPUT]
//
// Same options and output: -K/-M/-P/-S/-T as SearchParams (P < K becomes K,
// search.cpp:83-85), --linear answers by exact scan (cost 1), --eval GS
// compares the distances of the result ids with those of the ground-truth ids
// (both recomputed with metric::l2sqr and sorted) by AverageRecall,
// --l2norm normalises data and queries, OUTPUT receives the result ids as an
// lshkit matrix, and stdout gets "Time: <s> Recall: <r> Cost: <c>" with the
// wall time of the search itself.  All queries go to the GPU in one batched
// launch; Cost is the mean of comps / N (SearchInfo::cost, kgraph.cpp:561).
// --init IDS gives starting points: the first min(dim, K - 1) columns, since
// the walk needs init < K (kgraph.cpp:463); the reference tool requires
// dim >= K and then resizes the matrix that held them, so it cannot pass them.
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <iostream>
#include <vector>

#include "../../include/kgraph-data.h"
#include "../../include/kgraph.h"
#include "../../include/kgraph_amd.h"
#include "kg_cli.hpp"

using namespace kgraph;

static void check(kg_handle *h, int rc) {
    if (rc != KG_OK) throw std::runtime_error(kg_last_error(h));
}

int main(int argc, char **argv) {
    try {
        kgcli::Args a = kgcli::parse(argc, argv,
                                     {{"data", ""}, {"index", ""}, {"query", ""}, {"output", ""}, {"init", ""},
                                      {"eval", ""}, {"K", "K"}, {"M", "M"}, {"P", "P"}, {"S", "S"}, {"T", "T"}},
                                     {"help", "version", "linear", "l2norm"});
        auto posarg = [&](char const *k, size_t i) {
            return a.get<std::string>(k, a.pos.size() > i ? a.pos[i] : "");
        };
        std::string const input_path = posarg("data", 0), index_path = posarg("index", 1),
                          query_path = posarg("query", 2), output_path = posarg("output", 3);
        std::string const init_path = a.get<std::string>("init", ""), eval_path = a.get<std::string>("eval", "");
        unsigned const K = a.get<unsigned>("K", default_K), M = a.get<unsigned>("M", default_M);
        unsigned P = a.get<unsigned>("P", default_P);
        unsigned const S = a.get<unsigned>("S", default_S), T = a.get<unsigned>("T", default_T);
        if (a.has("version")) {
            std::cout << "KGraph version " << KGraph::version() << std::endl;
            return 0;
        }
        if (a.has("help") || input_path.empty() || index_path.empty() || query_path.empty()) {
            std::cout << "search <data> <index> <query> [output]\n"
                         "General options:\n"
                         "  -h [ --help ]         produce help message.\n"
                         "  -v [ --version ]      print version number.\n"
                         "  --data arg            input path\n  --index arg           index path\n"
                         "  --query arg           query path\n  --output arg          output path\n"
                         "  --init arg            init path\n  --eval arg            eval path\n"
                         "  -K arg (=25)\n  -M arg (=0)\n  -P arg (=100)\n  -S arg (=10)\n  -T arg (=1)\n"
                         "  --linear\n  --l2norm              l2-normalize data, so as to mimic cosine similarity\n";
            return 0;
        }
        if (P < K) P = K;  // search.cpp:83-85
        Matrix<float> data, query;
        Matrix<unsigned> init_ids;
        data.load_lshkit(input_path);
        query.load_lshkit(query_path);
        unsigned init = 0;
        if (!init_path.empty()) {
            init_ids.load_lshkit(init_path);
            if (init_ids.size() != query.size()) throw std::invalid_argument("init rows != queries");
            init = std::min(init_ids.dim(), K ? K - 1 : 0u);
        }
        if (a.has("l2norm")) {
            std::cerr << "L2-normalizing data..." << std::endl;
            data.normalize2();
            query.normalize2();
        }
        unsigned const nq = query.size(), N = data.size();
        kg_handle *h = kg_create();
        if (!h) throw std::runtime_error("kg_create failed (no GPU?)");
        check(h, kg_set_data(h, data[0], N, data.dim(), data.step() / sizeof(float), KG_METRIC_L2SQR));
        Matrix<unsigned> result(nq, K);
        std::vector<unsigned> ids(size_t(nq) * K, 0), counts(nq), comps(nq);
        std::vector<float> dists(size_t(nq) * K);
        size_t const qstride = query.step() / sizeof(float);
        float time = 0, cost = 0, recall = 0;
        if (a.has("linear")) {
            auto t0 = std::chrono::steady_clock::now();
            check(h, kg_bruteforce(h, query[0], nq, qstride, K, default_epsilon, ids.data(), dists.data(), K,
                                   counts.data()));
            time = float(std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
            cost = 1.0f;
        } else {
            check(h, kg_load(h, index_path.c_str()));
            for (unsigned q = 0; q < nq; ++q)
                for (unsigned l = 0; l < init; ++l) ids[size_t(q) * K + l] = init_ids[q][l];
            kg_search_params sp;
            kg_search_params_default(&sp);
            sp.K = K;
            sp.M = M;
            sp.P = P;
            sp.S = S;
            sp.T = T;
            sp.init = init;
            std::cerr << "Searching..." << std::endl;
            auto t0 = std::chrono::steady_clock::now();
            check(h, kg_search(h, query[0], nq, qstride, &sp, nullptr, nullptr, ids.data(), dists.data(), K,
                               counts.data(), comps.data(), nullptr));
            time = float(std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
            double c = 0;
            for (unsigned q = 0; q < nq; ++q) c += double(comps[q]) / N;
            cost = float(c / nq);
        }
        for (unsigned q = 0; q < nq; ++q)
            for (unsigned k = 0; k < K; ++k) result[q][k] = ids[size_t(q) * K + k];
        if (!output_path.empty()) result.save_lshkit(output_path);
        if (!eval_path.empty()) {
            Matrix<unsigned> gs;
            gs.load_lshkit(eval_path);
            if (gs.dim() < K || gs.size() < nq) throw std::invalid_argument("eval file too small");
            Matrix<float> gs_dist(nq, K), result_dist(nq, K);
            for (unsigned i = 0; i < nq; ++i) {
                for (unsigned k = 0; k < K; ++k) {
                    gs_dist[i][k] = metric::l2sqr::apply(query[i], data[gs[i][k]], data.dim());
                    result_dist[i][k] = metric::l2sqr::apply(query[i], data[result[i][k]], data.dim());
                }
                std::sort(gs_dist[i], gs_dist[i] + K);
                std::sort(result_dist[i], result_dist[i] + K);
            }
            recall = AverageRecall(gs_dist, result_dist, K);
        }
        std::cout << "Time: " << time << " Recall: " << recall << " Cost: " << cost << std::endl;
        kg_destroy(h);
    } catch (std::exception const &e) {
        std::cerr << "error: " << e.what() << std::endl;
        return 1;
    }
    return 0;
}

// Drop-in check of include/kgraph.h: written the way the reference's index.cpp
// (build + save) and search.cpp (load + per-query search + AverageRecall)
// use the library, compiled with g++ and linked to libkgraph_amd.so.
#include <chrono>
#include <cstdio>
#include <cstring>
#include <thread>
#include <vector>

#include "kgraph-data.h"
#include "kgraph.h"

using namespace kgraph;

int main(int argc, char **argv) {
    char const *path = argc > 1 ? argv[1] : "facade_demo.kgraph";
    unsigned const N = 20000, D = 32, NQ = 200, K = 10;
    Matrix<float> data(N, D), query(NQ, D);
    unsigned s = 12345;
    auto rnd = [&]() { s = s * 1103515245u + 12345u; return float((s >> 16) % 64); };
    for (unsigned i = 0; i < N; ++i)
        for (unsigned j = 0; j < D; ++j) data[i][j] = rnd();
    for (unsigned i = 0; i < NQ; ++i)
        for (unsigned j = 0; j < D; ++j) query[i][j] = rnd();
    verbosity = 0;
    MatrixOracle<float, metric::l2sqr> oracle(data);
    KGraph::IndexParams params;
    params.K = 20;
    params.L = 40;
    params.reverse = -1;
    KGraph::IndexInfo info;
    KGraph *index = KGraph::create();
    index->build(oracle, params, &info);
    index->save(path);
    KGraph *loaded = KGraph::create();
    loaded->load(path);

    KGraph::SearchParams sp;
    sp.K = K;
    sp.P = 50;
    Matrix<float> gs(NQ, K), rs(NQ, K);
    unsigned same = 0;
    float cost = 0;
    for (unsigned q = 0; q < NQ; ++q) {
        auto qo = oracle.query(query[q]);
        std::vector<unsigned> a(K), b(K), g(K);
        KGraph::SearchInfo si;
        unsigned la = index->search(qo, sp, &a[0], &si);
        unsigned lb = loaded->search(qo, sp, &b[0]);
        cost += si.cost;
        same += (la == lb && a == b);
        qo.search(K, default_epsilon, &g[0], gs[q]);      // brute force
        for (unsigned k = 0; k < K; ++k) rs[q][k] = qo(a[k]);
    }
    float recall = AverageRecall(gs, rs, K);
    // the harness pattern of search.cpp:139-146: many threads call search()
    // on one index at once (their calls are batched into shared launches);
    // every query must get exactly what a lone call gets, with its own K / P
    unsigned const T = 8;
    auto params_of = [&](unsigned q) {
        KGraph::SearchParams p = sp;
        p.K = (q % 3 == 0) ? K : K / 2;
        p.P = (q % 2 == 0) ? 50 : 30;
        return p;
    };
    std::vector<std::vector<unsigned>> solo(NQ), many(NQ);
    std::vector<unsigned> solo_n(NQ), many_n(NQ);
    auto t0 = std::chrono::steady_clock::now();
    for (unsigned q = 0; q < NQ; ++q) {
        solo[q].assign(K, 0);
        solo_n[q] = index->search(oracle.query(query[q]), params_of(q), &solo[q][0]);
    }
    auto t1 = std::chrono::steady_clock::now();
    std::vector<std::thread> pool;
    for (unsigned t = 0; t < T; ++t)
        pool.emplace_back([&, t]() {
            for (unsigned q = t; q < NQ; q += T) {
                many[q].assign(K, 0);
                many_n[q] = index->search(oracle.query(query[q]), params_of(q), &many[q][0]);
            }
        });
    for (auto &th : pool) th.join();
    auto t2 = std::chrono::steady_clock::now();
    unsigned concurrent_same = 0;
    for (unsigned q = 0; q < NQ; ++q) concurrent_same += (solo_n[q] == many_n[q] && solo[q] == many[q]);
    double const solo_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    double const many_ms = std::chrono::duration<double, std::milli>(t2 - t1).count();
    unsigned M = 0, L = 0;
    std::vector<unsigned> nn(4096);
    std::vector<float> nd(4096);
    index->get_nn(7, &nn[0], &nd[0], &M, &L);
    bool dist_ok = L > 0 && nd[0] == oracle(7, nn[0]);
    // rows changed in place: after forget_device_rows() the next search sees them
    unsigned const moved = 12345;
    for (unsigned j = 0; j < D; ++j) data[moved][j] = query[0][j];
    forget_device_rows();
    std::vector<unsigned> top(K);
    std::vector<float> topd(K);
    auto q0 = oracle.query(query[0]);
    q0.search(K, default_epsilon, &top[0], &topd[0]);
    bool const forget_ok = top[0] == moved && topd[0] == 0.f;
    // callback oracles carry no rows: the GPU index refuses them
    std::vector<std::vector<float>> vec(100, std::vector<float>(4, 1.f));
    VectorOracle<std::vector<std::vector<float>>, std::vector<float>> vo(
        vec, [](std::vector<float> const &x, std::vector<float> const &y) { return x[0] - y[0]; });
    bool rejected = false;
    try {
        KGraph *other = KGraph::create();
        other->build(vo, params);
        delete other;
    } catch (kgraph::invalid_argument const &) {
        rejected = true;
    }
    printf("{\"version\": \"%s\", \"iterations\": %u, \"build_recall\": %.4f, \"recall\": %.4f, \"cost\": %.5f, "
           "\"same_after_load\": %u, \"queries\": %u, \"get_nn_L\": %u, \"get_nn_M\": %u, \"dist_ok\": %d, "
           "\"opaque_rejected\": %d, \"concurrent_same\": %u, \"threads\": %u, \"solo_ms\": %.2f, "
           "\"concurrent_ms\": %.2f, \"forget_ok\": %d}\n",
           KGraph::version(), info.iterations, info.recall, recall, cost / NQ, same, NQ, L, M, int(dist_ok),
           int(rejected), concurrent_same, T, solo_ms, many_ms, int(forget_ok));
    delete index;
    delete loaded;
    return 0;
}

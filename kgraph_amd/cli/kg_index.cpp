// index -- build a KNNGRAPH index file from an lshkit (or raw) data file, the
// command-line tool of the reference (index.cpp:30-181) on the GPU.
//
//   index [OPTIONS] INPUT [OUTPUT]
//
// Same options, defaults and behaviour: the dimension comes from the lshkit
// header unless --raw (then --dim/--skip/--gap describe the file),
// --synthetic N draws N rows uniformly in [-1, 1) from
// std::default_random_engine(seed), --noise adds Gaussian noise of that many
// standard deviations, --l2norm normalises the rows, S = 0 means S = K
// (index.cpp:101-103).  The build runs on the GPU through libkgraph_amd.so
// (squared Euclidean distance, as MatrixOracle<float, metric::l2sqr>);
// --precision selects the local join's form (bf16: MFMA tiles with exact
// FP32 re-scoring, the default; fp32: exact FP32 tiles).  The stop condition
// goes to stderr, then the wall time.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <random>

#include "../../include/kgraph-data.h"
#include "../../include/kgraph.h"
#include "../../include/kgraph_amd.h"
#include "kg_cli.hpp"

using namespace kgraph;

static void usage() {
    std::cout << "Usage: index [OTHER OPTIONS]... INPUT [OUTPUT]\n"
                 "General options:\n"
                 "  -h [ --help ]                 produce help message.\n"
                 "  -v [ --version ]              print version information.\n"
                 "  --data arg                    input path\n"
                 "  --output arg                  output path\n"
                 "  -K arg (=25)                  number of nearest neighbor\n"
                 "  -C [ --controls ] arg (=100)  number of control points\n"
                 "Expert options:\n"
                 "  -I [ --iterations ] arg (=30)\n  -S arg (=10)\n  -R arg (=100)\n  -L arg (=100)\n"
                 "  --delta arg (=0.002)\n  --recall arg (=0.99)\n  --prune arg (=0)\n  --reverse arg (=0)\n"
                 "  --noise arg (=0)              noise\n  --seed arg (=1998)\n"
                 "  -D [ --dim ] arg              dimension, see format\n"
                 "  --skip arg (=0)               see format\n  --gap arg (=0)                see format\n"
                 "  --raw                         read raw binary file, need to specify D.\n"
                 "  --synthetic arg (=0)          generate synthetic data, for performance evaluation only\n"
                 "  --l2norm                      l2-normalize data, so as to mimic cosine similarity\n"
                 "  --precision arg (=bf16)       local join form: bf16 (MFMA, exact FP32 re-score) or fp32\n";
}

static void check(kg_handle *h, int rc) {
    if (rc != KG_OK) throw std::runtime_error(kg_last_error(h));
}

int main(int argc, char **argv) {
    try {
        kgcli::Args a = kgcli::parse(
            argc, argv,
            {{"data", ""}, {"output", ""}, {"K", "K"}, {"controls", "C"}, {"iterations", "I"}, {"S", "S"},
             {"R", "R"}, {"L", "L"}, {"delta", ""}, {"recall", ""}, {"prune", ""}, {"reverse", ""}, {"noise", ""},
             {"seed", ""}, {"dim", "D"}, {"skip", ""}, {"gap", ""}, {"synthetic", ""}, {"precision", ""}},
            {"help", "version", "raw", "l2norm"});
        std::string data_path = a.get<std::string>("data", a.pos.size() > 0 ? a.pos[0] : "");
        std::string output_path = a.get<std::string>("output", a.pos.size() > 1 ? a.pos[1] : "");
        kg_index_params p;
        kg_index_params_default(&p);
        p.K = a.get<unsigned>("K", p.K);
        p.controls = a.get<unsigned>("controls", p.controls);
        p.iterations = a.get<unsigned>("iterations", p.iterations);
        p.S = a.get<unsigned>("S", p.S);
        p.R = a.get<unsigned>("R", p.R);
        p.L = a.get<unsigned>("L", p.L);
        p.delta = a.get<float>("delta", p.delta);
        p.recall = a.get<float>("recall", p.recall);
        p.prune = a.get<unsigned>("prune", p.prune);
        p.reverse = a.get<int>("reverse", p.reverse);
        p.seed = a.get<unsigned>("seed", p.seed);
        float const noise = a.get<float>("noise", 0.f);
        unsigned D = a.get<unsigned>("dim", 0u), skip = a.get<unsigned>("skip", 0u), gap = a.get<unsigned>("gap", 0u);
        unsigned const synthetic = a.get<unsigned>("synthetic", 0u);
        bool const lshkit = !a.has("raw");
        std::string const prec = a.get<std::string>("precision", "bf16");
        if (prec != "bf16" && prec != "fp32") throw std::invalid_argument("--precision must be bf16 or fp32");
        if (a.has("version")) {
            std::cout << "KGraph version " << KGraph::version() << std::endl;
            return 0;
        }
        if (a.has("help") || (synthetic && (!a.has("dim") || !data_path.empty())) ||
            (!synthetic && (data_path.empty() || (!a.has("dim") && !lshkit)))) {
            usage();
            return 0;
        }
        if (p.S == 0) p.S = p.K;  // index.cpp:101-103
        if (lshkit && synthetic == 0) {
            std::ifstream is(data_path.c_str(), std::ios::binary);
            unsigned header[3];  // entry size, rows, cols
            is.read(reinterpret_cast<char *>(header), sizeof header);
            if (!is || header[0] != sizeof(float)) throw io_error(data_path);
            D = header[2];
            skip = 3 * sizeof(unsigned);
            gap = 0;
        }
        Matrix<float> data;
        if (synthetic) {
            data.resize(synthetic, D);
            std::cerr << "Generating synthetic data..." << std::endl;
            std::default_random_engine rng(p.seed);
            std::uniform_real_distribution<double> dist(-1.0, 1.0);
            data.zero();
            for (unsigned i = 0; i < synthetic; ++i)
                for (unsigned j = 0; j < D; ++j) data[i][j] = float(dist(rng));
        } else {
            data.load(data_path, D, skip, gap);
        }
        if (noise != 0) {
            std::default_random_engine rng;
            double sum = 0, sum2 = 0;
            for (unsigned i = 0; i < data.size(); ++i)
                for (unsigned j = 0; j < data.dim(); ++j) {
                    double v = data[i][j];
                    sum += v;
                    sum2 += v * v;
                }
            double total = double(data.size()) * data.dim();
            double avg2 = sum2 / total, avg = sum / total, dev = std::sqrt(avg2 - avg * avg);
            std::cerr << "Adding Gaussian noise w/ " << noise << "x sigma(" << dev << ")..." << std::endl;
            std::normal_distribution<double> gaussian(0, noise * dev);
            for (unsigned i = 0; i < data.size(); ++i)
                for (unsigned j = 0; j < data.dim(); ++j) data[i][j] += float(gaussian(rng));
        }
        if (a.has("l2norm")) {
            std::cerr << "L2-normalizing data..." << std::endl;
            data.normalize2();
        }
        kg_handle *h = kg_create();
        if (!h) throw std::runtime_error("kg_create failed (no GPU?)");
        check(h, kg_set_data(h, data[0], data.size(), data.dim(), data.step() / sizeof(float), KG_METRIC_L2SQR));
        check(h, kg_set_precision(h, prec == "bf16" ? KG_PRECISION_BF16 : KG_PRECISION_FP32));
        kg_index_info info;
        auto t0 = std::chrono::steady_clock::now();
        check(h, kg_build(h, &p, &info));
        double wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        std::cerr << info.stop_condition << std::endl;
        std::fprintf(stderr, " %.6fs wall\n", wall);
        if (!output_path.empty()) check(h, kg_save(h, output_path.c_str(), KG_FORMAT_DEFAULT));
        kg_destroy(h);
    } catch (std::exception const &e) {
        std::cerr << "error: " << e.what() << std::endl;
        return 1;
    }
    return 0;
}

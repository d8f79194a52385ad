// prune -- prune a KNNGRAPH index file, the reference's command-line tool
// (prune.cpp:27-113) on the GPU.
//
//   prune [OPTIONS] DATA INPUT [OUTPUT]
//
// Same options and defaults: --prune takes the level (the reference's
// default is PRUNE_LEVEL_1 & PRUNE_LEVEL_2, which is 0: nothing is pruned
// unless a level is given), the data file is lshkit unless --raw (then
// --dim/--skip/--gap describe it).  The data file is read as the reference
// reads it (KGraph::prune takes the oracle; prune levels 1 and 2 use only the
// stored lists, kgraph.cpp:584-665), the index is loaded, pruned through
// kgraph::KGraph::prune (the GPU implementation) and saved to OUTPUT when
// given.  The wall time of the prune goes to stdout.
#include <chrono>
#include <cstdio>
#include <fstream>
#include <iostream>

#include "../../include/kgraph-data.h"
#include "../../include/kgraph.h"
#include "kg_cli.hpp"

using namespace kgraph;

static void usage() {
    std::cout << "Usage: prune [OTHER OPTIONS]... DATA INPUT [OUTPUT]\n"
                 "General options:\n"
                 "  -h [ --help ]          produce help message.\n"
                 "  --data arg             input path\n"
                 "  --input arg            input index path\n"
                 "  --output arg           output index path\n"
                 "  --prune arg (=0)       prune level\n\n"
                 "Expert options:\n"
                 "  -D [ --dim ] arg       dimension, see format\n"
                 "  --skip arg (=0)        see format\n"
                 "  --gap arg (=0)         see format\n"
                 "  --raw                  read raw binary file, need to specify D.\n";
}

int main(int argc, char **argv) {
    try {
        kgcli::Args a = kgcli::parse(argc, argv,
                                     {{"data", ""}, {"input", ""}, {"output", ""}, {"prune", ""}, {"dim", "D"},
                                      {"skip", ""}, {"gap", ""}},
                                     {"help", "raw"});
        std::string const data_path = a.get<std::string>("data", a.pos.size() > 0 ? a.pos[0] : "");
        std::string const input_path = a.get<std::string>("input", a.pos.size() > 1 ? a.pos[1] : "");
        std::string const output_path = a.get<std::string>("output", a.pos.size() > 2 ? a.pos[2] : "");
        unsigned const prune = a.get<unsigned>("prune", unsigned(PRUNE_LEVEL_1 & PRUNE_LEVEL_2));
        unsigned D = a.get<unsigned>("dim", 0u), skip = a.get<unsigned>("skip", 0u), gap = a.get<unsigned>("gap", 0u);
        bool const lshkit = !a.has("raw");
        if (a.has("help") || input_path.empty()) {
            usage();
            return 0;
        }
        Matrix<float> data;
        if (lshkit) {
            data.load_lshkit(data_path);
        } else {
            if (!a.has("dim")) throw invalid_argument("--raw needs --dim");
            data.load(data_path, D, skip, gap);
        }
        MatrixOracle<float, metric::l2sqr> oracle(data);
        KGraph *index = KGraph::create();
        index->load(input_path.c_str());
        auto t0 = std::chrono::steady_clock::now();
        index->prune(oracle, prune);
        double const wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        std::printf(" %.6fs wall\n", wall);
        if (!output_path.empty()) index->save(output_path.c_str());
        delete index;
    } catch (std::exception const &e) {
        std::cerr << "error: " << e.what() << std::endl;
        return 1;
    }
    return 0;
}

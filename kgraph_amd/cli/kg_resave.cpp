// resave -- load a KNNGRAPH index file, optionally prune it, and save it
// again, optionally without distances: the reference's command-line tool
// (resave.cpp:37-87).
//
//   resave [OPTIONS] INPUT [OUTPUT]
//
// Same options: --prune LEVEL (only when given), --no-dist (save with
// FORMAT_NO_DIST).  Pruning needs no data rows (kgraph.cpp:584-665), so, as
// in the reference, no dataset is read.
#include <iostream>

#include "../../include/kgraph.h"
#include "kg_cli.hpp"

using namespace kgraph;

namespace {
// resave.cpp:20-36: an oracle that must never be asked for anything
struct NoDataOracle : IndexOracle {
    unsigned size() const override { throw runtime_error("resave has no dataset"); }
    float operator()(unsigned, unsigned) const override { throw runtime_error("resave has no dataset"); }
};
}  // namespace

static void usage() {
    std::cout << "General options:\n"
                 "  -h [ --help ]         produce help message.\n"
                 "  --input arg\n"
                 "  --output arg\n"
                 "  --prune arg\n"
                 "  --no-dist\n";
}

int main(int argc, char **argv) {
    try {
        kgcli::Args a = kgcli::parse(argc, argv, {{"input", ""}, {"output", ""}, {"prune", ""}},
                                     {"help", "no-dist"});
        std::string const input_path = a.get<std::string>("input", a.pos.size() > 0 ? a.pos[0] : "");
        std::string const output_path = a.get<std::string>("output", a.pos.size() > 1 ? a.pos[1] : "");
        if (a.has("help") || input_path.empty()) {
            usage();
            return 0;
        }
        int format = 0;
        if (a.has("no-dist")) format |= KGraph::FORMAT_NO_DIST;
        KGraph *index = KGraph::create();
        index->load(input_path.c_str());
        if (a.has("prune")) {
            NoDataOracle o;
            index->prune(o, a.get<unsigned>("prune", 0u));
        }
        if (!output_path.empty()) index->save(output_path.c_str(), format);
        delete index;
    } catch (std::exception const &e) {
        std::cerr << "error: " << e.what() << std::endl;
        return 1;
    }
    return 0;
}

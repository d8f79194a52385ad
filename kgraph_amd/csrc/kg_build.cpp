// kg_build.cpp -- host driver of the NN-descent build (placeholder until the
// build kernels land).
#include "kg_handle.hpp"

namespace kga {

void build_index(Handle &, kg_index_params const &, kg_index_info *) { throw runtime_error("build not available"); }
void reverse_index(Handle &, int) { throw runtime_error("reverse not available"); }
void prune_index(Handle &, unsigned) { throw runtime_error("prune not available"); }

}  // namespace kga

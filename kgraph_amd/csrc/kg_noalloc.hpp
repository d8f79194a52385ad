// kg_noalloc.hpp -- the scan that allocates nothing (kg_noalloc.hip).
#pragma once
#include <stddef.h>
#include <stdint.h>

namespace kga {

struct Handle;

// Scratch array per GPU: 4 MiB of query norms (up to 1,048,576 queries per
// call) and 32 MiB of partial top-K lists.  The split rule of the scan
// (kg_brute_splits: splits only while nq * splits < 4096 and splits * K <=
// 4096) keeps the partial lists of any call within 32 MiB (nq * splits * K *
// 8 < 4096 * 1024 * 8).
constexpr size_t NOALLOC_NORM_BYTES = size_t(4) << 20;
constexpr size_t NOALLOC_SCRATCH_BYTES = NOALLOC_NORM_BYTES + (size_t(32) << 20);
constexpr unsigned NOALLOC_MAX_QUERIES = unsigned(NOALLOC_NORM_BYTES / 4);

// Takes the scratch array's device address on `device` (once per device).
void noalloc_setup(int device);
void bruteforce_noalloc(Handle &h, float const *dQ, unsigned nq, size_t qstride, unsigned K, float epsilon,
                        uint32_t *d_ids, float *d_dists, size_t ostride, uint32_t *d_counts);

}  // namespace kga

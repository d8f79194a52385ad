// kg_exact.hpp -- when BF16 tiles on the matrix cores give the canonical FP32
// distance bit for bit (host code only; no HIP dependency, so the predicate is
// unit-tested on the CPU).
#pragma once

namespace kga {

// Integer rows whose every value BF16 holds exactly (`lossless`) make every
// product a_i b_i an exact integer.  The tile forms (kg_nndescent.hip
// mj_dist) compute
//   L2SQR      (|a|^2 + |o|^2) - 2 a.o
//   EUCLIDEAN  (n_a + n_o) - a.o        with n = |x|^2 / 2
//   ANGULAR    -a.o / (|a| |o| + FLT_EPSILON)
// and the canonical order sums (a_i - o_i)^2 or a_i o_i.  All of it is exact
// in FP32 when
//   * 2 D max|x|^2 <= 2^24: every partial sum of a dot, every norm and the sum
//     of two norms is an integer (a half-integer below 2^23 for n) that FP32
//     holds;
//   * D (max x - min x)^2 <= 2^24 (not needed for ANGULAR): every partial sum
//     of squared differences, and the result |a - o|^2 itself, is an integer
//     FP32 holds (rows with values of both signs have differences up to
//     2 max|x|, so the first condition alone does not bound them).
// `metric` is 0 L2SQR, 1 EUCLIDEAN, 2 ANGULAR.
inline bool bf16_tiles_exact(unsigned D, int metric, bool integer, bool lossless, double maxabs, double vmin,
                             double vmax) {
    double const lim = 16777216.0;  // 2^24
    if (!integer || !lossless) return false;
    if (2.0 * D * maxabs * maxabs > lim) return false;
    if (metric != 2) {
        double const range = vmax - vmin;
        if (double(D) * range * range > lim) return false;
    }
    return true;
}

}  // namespace kga

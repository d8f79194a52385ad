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

// The 8-bit form: rows whose values are all integers in [0, 255] (their
// 8-bit copy is exact) enter v_mfma_i32_16x16x64_i8 shifted to [-128, 127]
// (x - 128: the top bit flipped).  The L2SQR and half-SSD distances do not
// change under the shift; dots (|a' . o'| <= 2^14 D) and norms
// (|x'|^2 <= 2^14 D) are exact int32 sums, and the tile value
// (n_a + n_o) - 2 a'.o' (halved norms and a'.o' for EUCLIDEAN) is the exact
// integer |a - o|^2 <= 255^2 D (its half), which FP32 holds, as it holds every
// partial sum of the canonical order, when 255^2 D <= 2^24.  ANGULAR divides by
// the unshifted norms and takes the BF16 form.
inline bool i8_tiles_exact(unsigned D, int metric, bool u8rows) {
    return u8rows && metric != 2 && 65025.0 * double(D) <= 16777216.0;
}

}  // namespace kga

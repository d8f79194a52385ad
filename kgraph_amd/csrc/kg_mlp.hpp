// kg_mlp.hpp -- launch interface of the learned projector kernels (kg_mlp.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

enum { GEMM_NONE = 0, GEMM_RELU = 1, GEMM_RELU_MASK = 2 };

// C[m][n] (row stride ldc) = sum_k A(m,k) B(k,n) (+ bias[n]), with
// A(m,k) = A[m*sam + k*sak] and B(k,n) = B[k*sbk + n*sbn].  Epilogue: ReLU,
// or zero where mask[m][n] <= 0 (the ReLU derivative of the forward output).
struct GemmArgs {
    float const *A;
    float const *B;
    float *C;
    float const *bias;
    float const *mask;
    unsigned M, N, K;
    size_t sam, sak, sbk, sbn, ldc;
    int epilogue;
};

struct TripletArgs {
    uint32_t const *goff, *gids, *gM;
    unsigned N, B;
    unsigned seed, step;
    uint32_t *idx;  // 3B indices: anchor, positive, negative
};

hipError_t kg_gemm(GemmArgs const &g, hipStream_t s);
// Row gathers standardise the columns: out = (x - mu) * is (mu, is NULL: copy).
hipError_t kg_triplets(TripletArgs const &t, float const *X, unsigned Dp, unsigned D, float const *mu,
                       float const *is, float *out, hipStream_t s);
hipError_t kg_gather(float const *X, unsigned Dp, unsigned D, uint32_t const *idx, unsigned n, float const *mu,
                     float const *is, float *out, hipStream_t s);
// Per-column mean and 1/std of the N x D rows (row stride Dp), FP64 sums in a
// fixed order (part holds nb x D x 2 doubles).
hipError_t kg_colstats(float const *X, unsigned N, unsigned D, unsigned Dp, double *part, unsigned nb, float *mu,
                       float *is, hipStream_t s);
hipError_t kg_triplet_loss(float const *Z, unsigned P, unsigned B, float margin, float *dZ, float *loss,
                           hipStream_t s);
hipError_t kg_colsum(float const *G, unsigned M, unsigned N, float *out, hipStream_t s);
hipError_t kg_adam(float *w, float const *g, float *m1, float *m2, unsigned n, float lr, unsigned step,
                   hipStream_t s);

// kg_mlp.hip -- the learned projector: a 2-layer MLP f(x) = W2 relu(W1 x + b1)
// + b2 (D -> H -> P, FP32 master weights) trained with a triplet loss on
// (node, neighbour from the built graph, random node) triplets, and applied
// to pick search entry points in the projected space (no counterpart in the
// reference; the exact FP32 distance re-ranks every shortlist).
//
//   gemm_kernel      C = A.B (+bias, ReLU or ReLU-mask epilogue) on MFMA:
//                    v_mfma_f32_16x16x32_bf16, operands rounded to BF16
//                    (round to nearest even), FP32 accumulation.  One wave per
//                    16x16 output tile; A and B are read through arbitrary
//                    strides, so the forward, the input-gradient and the
//                    weight-gradient products share the kernel.
//   triplets_kernel  counter-RNG triplet sampling from the CSR graph (negatives
//                    uniform or two hops away, never in the anchor's list)
//   triplet_loss     L = mean(max(0, |za-zp|^2 - |za-zn|^2 + margin)) and dL/dZ
//   colsum_kernel    bias gradients
//   adam_kernel      Adam update of the FP32 master weights
#include "kg_device.hpp"
#include "kg_mlp.hpp"

namespace {

using namespace kgd;

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ short to_bf16(float f) {  // round to nearest even
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return short(u >> 16);
}

// One wave per 16x16 tile of C (M x N).  MFMA 16x16x32 operand layout
// (wave64): lane l holds A[m = l%16][k = 8*(l/16) + j] and
// B[k = 8*(l/16) + j][n = l%16], j = 0..7; the result lane l holds
// C[m = 4*(l/16) + j][n = l%16], j = 0..3.
__global__ void __launch_bounds__(256) gemm_kernel(GemmArgs g) {
    unsigned const lane = threadIdx.x & 63;
    unsigned const tile = blockIdx.x * 4 + (threadIdx.x >> 6);
    unsigned const tiles_n = (g.N + 15) / 16;
    if (tile >= tiles_n * ((g.M + 15) / 16)) return;
    unsigned const tm = tile / tiles_n, tn = tile - tm * tiles_n;
    unsigned const r16 = lane & 15, k8 = (lane >> 4) * 8;
    unsigned const am = tm * 16 + r16, bn = tn * 16 + r16;
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    for (unsigned k0 = 0; k0 < g.K; k0 += 32) {
        uint32_t aw[4], bw[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            uint32_t alo = 0, ahi = 0, blo = 0, bhi = 0;
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                unsigned const k = k0 + k8 + 2 * j + h;
                float a = (am < g.M && k < g.K) ? g.A[size_t(am) * g.sam + size_t(k) * g.sak] : 0.f;
                float b = (bn < g.N && k < g.K) ? g.B[size_t(k) * g.sbk + size_t(bn) * g.sbn] : 0.f;
                uint32_t const ab = uint16_t(to_bf16(a)), bb = uint16_t(to_bf16(b));
                if (h == 0) {
                    alo = ab;
                    blo = bb;
                } else {
                    ahi = ab;
                    bhi = bb;
                }
            }
            aw[j] = alo | (ahi << 16);
            bw[j] = blo | (bhi << 16);
        }
        bf16x8 const av = __builtin_bit_cast(bf16x8, make_uint4(aw[0], aw[1], aw[2], aw[3]));
        bf16x8 const bv = __builtin_bit_cast(bf16x8, make_uint4(bw[0], bw[1], bw[2], bw[3]));
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, acc, 0, 0, 0);
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        unsigned m = tm * 16 + (lane >> 4) * 4 + j, n = bn;
        if (m < g.M && n < g.N) {
            float v = acc[j];
            if (g.bias) v += g.bias[n];
            if (g.epilogue == GEMM_RELU) v = v > 0.f ? v : 0.f;
            if (g.epilogue == GEMM_RELU_MASK) v = (g.mask[size_t(m) * g.ldc + n] > 0.f) ? v : 0.f;
            g.C[size_t(m) * g.ldc + n] = v;
        }
    }
}

__device__ __forceinline__ uint64_t mix64(uint64_t x) {
    x += 0x9e3779b97f4a7c15ull;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
    return x ^ (x >> 31);
}

// Triplet t: anchor a, positive = one of a's first min(M[a], deg) graph
// neighbours, negative = a random node; rows gathered to X[3t..3t+2].
__global__ void triplets_kernel(TripletArgs t) {
    unsigned const i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= t.B) return;
    uint64_t s = mix64((uint64_t(t.seed) << 32) ^ (uint64_t(t.step) << 20) ^ i);
    uint32_t a = uint32_t(mix64(s) % t.N);
    uint32_t deg = t.goff[a + 1] - t.goff[a];
    uint32_t lim = min(t.gM[a], deg);
    uint32_t p = lim ? t.gids[t.goff[a] + uint32_t(mix64(s + 1) % lim)] : a;
    uint32_t n = uint32_t(mix64(s + 2) % t.N);
    // every other triplet draws its non-neighbour from two hops away (a random
    // neighbour of the positive that is neither the anchor nor in the anchor's
    // list), the rest uniformly: on clustered data a uniform node is almost
    // never close enough to leave a gradient
    if (i & 1u) {
        uint32_t const pdeg = t.goff[p + 1] - t.goff[p], plim = min(t.gM[p], pdeg);
        if (plim) {
            uint32_t const c = t.gids[t.goff[p] + uint32_t(mix64(s + 3) % plim)];
            bool ok = c != a;
            for (uint32_t j = 0; ok && j < lim; ++j) ok = t.gids[t.goff[a] + j] != c;
            if (ok) n = c;
        }
    }
    t.idx[3 * i] = a;
    t.idx[3 * i + 1] = p;
    t.idx[3 * i + 2] = n;
}

// Rows idx[r] (or r itself when idx is NULL) of X, standardised per column:
// out[r][c] = (X[.][c] - mu[c]) * is[c] (mu, is: the projector's input
// normalisation; both NULL = copy).
__global__ void gather_kernel(float const *X, unsigned Dp, unsigned D, uint32_t const *idx, unsigned n, float *out,
                              float const *mu, float const *is) {
    size_t e = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= size_t(n) * D) return;
    size_t r = e / D, c = e - r * D;
    float v = X[size_t(idx ? idx[r] : uint32_t(r)) * Dp + c];
    if (mu) v = (v - mu[c]) * is[c];
    out[r * D + c] = v;
}

// Column sums / sums of squares (FP64) of rows b, b + nb, ... per block b, then
// per column in block order: mean and 1/std (1 where the std is 0).
__global__ void colstats_partial_kernel(float const *X, unsigned N, unsigned D, unsigned Dp, double *part) {
    unsigned const nb = gridDim.x, b = blockIdx.x;
    for (unsigned c = threadIdx.x; c < D; c += blockDim.x) {
        double s = 0.0, q = 0.0;
        for (unsigned r = b; r < N; r += nb) {
            double v = X[size_t(r) * Dp + c];
            s += v;
            q += v * v;
        }
        part[(size_t(b) * D + c) * 2] = s;
        part[(size_t(b) * D + c) * 2 + 1] = q;
    }
}

__global__ void colstats_final_kernel(double const *part, unsigned nb, unsigned N, unsigned D, float *mu, float *is) {
    unsigned const c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= D) return;
    double s = 0.0, q = 0.0;
    for (unsigned b = 0; b < nb; ++b) {
        s += part[(size_t(b) * D + c) * 2];
        q += part[(size_t(b) * D + c) * 2 + 1];
    }
    double const m = s / N, var = q / N - m * m;
    mu[c] = float(m);
    is[c] = var > 0.0 ? float(1.0 / sqrt(var)) : 1.f;
}

// One thread per triplet; Z rows 3t, 3t+1, 3t+2 (P floats each).
__global__ void triplet_loss_kernel(float const *Z, unsigned P, unsigned B, float margin, float *dZ, float *loss) {
    unsigned const i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= B) return;
    float const *za = Z + size_t(3 * i) * P, *zp = za + P, *zn = zp + P;
    float dap = 0.f, dan = 0.f;
    for (unsigned k = 0; k < P; ++k) {
        float x = za[k] - zp[k], y = za[k] - zn[k];
        dap = __builtin_fmaf(x, x, dap);
        dan = __builtin_fmaf(y, y, dan);
    }
    float l = dap - dan + margin;
    float *ga = dZ + size_t(3 * i) * P, *gp = ga + P, *gn = gp + P;
    float const s = 2.f / float(B);
    for (unsigned k = 0; k < P; ++k) {
        if (l > 0.f) {
            ga[k] = s * (zn[k] - zp[k]);
            gp[k] = -s * (za[k] - zp[k]);
            gn[k] = s * (za[k] - zn[k]);
        } else {
            ga[k] = gp[k] = gn[k] = 0.f;
        }
    }
    if (l > 0.f) atomicAdd(loss, l / float(B));
}

__global__ void colsum_kernel(float const *G, unsigned M, unsigned N, float *out) {
    unsigned const n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    float s = 0.f;
    for (unsigned m = 0; m < M; ++m) s += G[size_t(m) * N + n];
    out[n] = s;
}

__global__ void adam_kernel(float *w, float const *g, float *m1, float *m2, unsigned n, float lr, float b1, float b2,
                            float eps, float c1, float c2) {
    unsigned const i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float gi = g[i];
    float a = b1 * m1[i] + (1.f - b1) * gi;
    float b = b2 * m2[i] + (1.f - b2) * gi * gi;
    m1[i] = a;
    m2[i] = b;
    w[i] -= lr * (a / c1) / (sqrtf(b / c2) + eps);
}

}  // namespace

hipError_t kg_gemm(GemmArgs const &g, hipStream_t s) {
    unsigned tiles = ((g.M + 15) / 16) * ((g.N + 15) / 16);
    if (tiles == 0) return hipSuccess;
    hipLaunchKernelGGL(gemm_kernel, dim3((tiles + 3) / 4), dim3(256), 0, s, g);
    return hipGetLastError();
}

hipError_t kg_triplets(TripletArgs const &t, float const *X, unsigned Dp, unsigned D, float const *mu,
                       float const *is, float *out, hipStream_t s) {
    hipLaunchKernelGGL(triplets_kernel, dim3((t.B + 255) / 256), dim3(256), 0, s, t);
    size_t n = size_t(3) * t.B * D;
    hipLaunchKernelGGL(gather_kernel, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s, X, Dp, D, t.idx, 3 * t.B, out,
                       mu, is);
    return hipGetLastError();
}

hipError_t kg_gather(float const *X, unsigned Dp, unsigned D, uint32_t const *idx, unsigned n, float const *mu,
                     float const *is, float *out, hipStream_t s) {
    size_t e = size_t(n) * D;
    if (e == 0) return hipSuccess;
    hipLaunchKernelGGL(gather_kernel, dim3(unsigned((e + 255) / 256)), dim3(256), 0, s, X, Dp, D, idx, n, out, mu,
                       is);
    return hipGetLastError();
}

hipError_t kg_colstats(float const *X, unsigned N, unsigned D, unsigned Dp, double *part, unsigned nb, float *mu,
                       float *is, hipStream_t s) {
    if (N == 0 || D == 0 || nb == 0) return hipErrorInvalidValue;
    hipLaunchKernelGGL(colstats_partial_kernel, dim3(nb), dim3(256), 0, s, X, N, D, Dp, part);
    hipLaunchKernelGGL(colstats_final_kernel, dim3((D + 255) / 256), dim3(256), 0, s, part, nb, N, D, mu, is);
    return hipGetLastError();
}

hipError_t kg_triplet_loss(float const *Z, unsigned P, unsigned B, float margin, float *dZ, float *loss,
                           hipStream_t s) {
    hipLaunchKernelGGL(triplet_loss_kernel, dim3((B + 255) / 256), dim3(256), 0, s, Z, P, B, margin, dZ, loss);
    return hipGetLastError();
}

hipError_t kg_colsum(float const *G, unsigned M, unsigned N, float *out, hipStream_t s) {
    hipLaunchKernelGGL(colsum_kernel, dim3((N + 255) / 256), dim3(256), 0, s, G, M, N, out);
    return hipGetLastError();
}

hipError_t kg_adam(float *w, float const *g, float *m1, float *m2, unsigned n, float lr, unsigned step,
                   hipStream_t s) {
    float const b1 = 0.9f, b2 = 0.999f;
    float c1 = 1.f - powf(b1, float(step)), c2 = 1.f - powf(b2, float(step));
    hipLaunchKernelGGL(adam_kernel, dim3((n + 255) / 256), dim3(256), 0, s, w, g, m1, m2, n, lr, b1, b2, 1e-8f, c1,
                       c2);
    return hipGetLastError();
}

// kg_mlp.cpp -- host side of the learned projector (kg_mlp.hip): weight
// initialisation, the forward / backward products, Adam training on triplets
// sampled from the built graph, and the projected entry sample.
#include <cmath>
#include <random>
#include <vector>

#include "kg_handle.hpp"
#include "kg_internal.hpp"
#include "kg_mlp.hpp"

namespace kga {

namespace {

GemmArgs gemm(float const *A, size_t sam, size_t sak, float const *B, size_t sbk, size_t sbn, float *C, size_t ldc,
              unsigned M, unsigned N, unsigned K, float const *bias, int epi, float const *mask) {
    GemmArgs g{};
    g.A = A;
    g.B = B;
    g.C = C;
    g.bias = bias;
    g.mask = mask;
    g.M = M;
    g.N = N;
    g.K = K;
    g.sam = sam;
    g.sak = sak;
    g.sbk = sbk;
    g.sbn = sbn;
    g.ldc = ldc;
    g.epilogue = epi;
    return g;
}

void require_mlp(Handle &h) {
    if (!h.mlpH) throw runtime_error("no projector (kg_mlp_init)");
}

// dZ (n x P) -> weight / bias gradients; X (n x D, row stride D), Hb (n x H)
void backward(Handle &h, float const *X, unsigned n, float const *Hb, float const *dZ) {
    unsigned const D = h.D, H = h.mlpH, P = h.mlpP;
    hipStream_t s = h.stream;
    h.mlp_dh.reserve(size_t(n) * H);
    KG_HIP(kg_gemm(gemm(dZ, 1, P, Hb, H, 1, h.gW2.p, H, P, H, n, nullptr, GEMM_NONE, nullptr), s));  // dW2 = dZ^T H
    KG_HIP(kg_colsum(dZ, n, P, h.gb2.p, s));
    KG_HIP(kg_gemm(gemm(dZ, P, 1, h.W2.p, H, 1, h.mlp_dh.p, H, n, H, P, nullptr, GEMM_RELU_MASK, Hb), s));  // dH
    KG_HIP(kg_gemm(gemm(h.mlp_dh.p, 1, H, X, D, 1, h.gW1.p, D, H, D, n, nullptr, GEMM_NONE, nullptr), s));  // dW1
    KG_HIP(kg_colsum(h.mlp_dh.p, n, H, h.gb1.p, s));
}

}  // namespace

void mlp_init(Handle &h, unsigned hidden, unsigned out, unsigned seed) {
    if (!h.has_data) throw runtime_error("no dataset");
    if (hidden == 0 || out == 0 || hidden > 4096 || out > 1024 || (out & 3))
        throw invalid_argument("hidden must be in [1, 4096], out in [4, 1024] and a multiple of 4");
    unsigned const D = h.D;
    h.mlpH = hidden;
    h.mlpP = out;
    // input standardisation: per-column mean and 1/std of the dataset
    {
        unsigned const nb = std::min(h.N, 1024u);
        DevBuf<double> part;
        part.reserve(size_t(nb) * D * 2);
        h.mlp_mu.reserve(D);
        h.mlp_is.reserve(D);
        KG_HIP(kg_colstats(h.X.p, h.N, D, h.Dp, part.p, nb, h.mlp_mu.p, h.mlp_is.p, h.stream));
        h.sync();
    }
    std::mt19937 rng(seed);
    auto init = [&](DevBuf<float> &w, size_t n, float lim) {
        std::uniform_real_distribution<float> u(-lim, lim);
        std::vector<float> v(n);
        for (auto &x : v) x = lim > 0 ? u(rng) : 0.f;
        w.reserve(n);
        KG_HIP(hipMemcpy(w.p, v.data(), 4 * n, hipMemcpyHostToDevice));
    };
    init(h.W1, size_t(hidden) * D, std::sqrt(6.f / float(D + hidden)));
    init(h.b1, hidden, 0.f);
    init(h.W2, size_t(out) * hidden, std::sqrt(6.f / float(hidden + out)));
    init(h.b2, out, 0.f);
    for (auto *b : {&h.gW1, &h.mW1, &h.vW1}) b->reserve(size_t(hidden) * D);
    for (auto *b : {&h.gb1, &h.mb1, &h.vb1}) b->reserve(hidden);
    for (auto *b : {&h.gW2, &h.mW2, &h.vW2}) b->reserve(size_t(out) * hidden);
    for (auto *b : {&h.gb2, &h.mb2, &h.vb2}) b->reserve(out);
    for (auto *b : {&h.mW1, &h.vW1}) KG_HIP(hipMemset(b->p, 0, 4ull * hidden * D));
    for (auto *b : {&h.mb1, &h.vb1}) KG_HIP(hipMemset(b->p, 0, 4ull * hidden));
    for (auto *b : {&h.mW2, &h.vW2}) KG_HIP(hipMemset(b->p, 0, 4ull * out * hidden));
    for (auto *b : {&h.mb2, &h.vb2}) KG_HIP(hipMemset(b->p, 0, 4ull * out));
    h.adam_step = 0;
    ++h.mlp_version;
}

// Forward pass of raw rows (row stride `stride`): standardised into a
// scratch copy first.
void mlp_forward_rows(Handle &h, float const *X, unsigned n, size_t stride, float *Hout, float *Zout) {
    require_mlp(h);
    h.mlp_xn.reserve(size_t(n) * h.D);
    KG_HIP(kg_gather(X, unsigned(stride), h.D, nullptr, n, h.mlp_mu.p, h.mlp_is.p, h.mlp_xn.p, h.stream));
    mlp_forward(h, h.mlp_xn.p, n, h.D, Hout, Zout);
}

void mlp_forward(Handle &h, float const *X, unsigned n, size_t stride, float *Hout, float *Zout) {
    require_mlp(h);
    unsigned const D = h.D, H = h.mlpH, P = h.mlpP;
    KG_HIP(kg_gemm(gemm(X, stride, 1, h.W1.p, 1, D, Hout, H, n, H, D, h.b1.p, GEMM_RELU, nullptr), h.stream));
    KG_HIP(kg_gemm(gemm(Hout, H, 1, h.W2.p, 1, H, Zout, P, n, P, H, h.b2.p, GEMM_NONE, nullptr), h.stream));
}

void mlp_train(Handle &h, unsigned steps, unsigned batch, float lr, float margin, unsigned seed, float *trace) {
    require_mlp(h);
    if (!h.has_graph) throw runtime_error("no index to sample triplets from");
    if (batch == 0 || batch > (1u << 20)) throw invalid_argument("batch must be in [1, 2^20]");
    unsigned const D = h.D, H = h.mlpH, P = h.mlpP, n = 3 * batch;
    hipStream_t s = h.stream;
    h.mlp_idx.reserve(n);
    h.mlp_x.reserve(size_t(n) * D);
    h.mlp_h.reserve(size_t(n) * H);
    h.mlp_z.reserve(size_t(n) * P);
    h.mlp_dz.reserve(size_t(n) * P);
    h.mlp_loss.reserve(std::max(steps, 1u));
    KG_HIP(hipMemsetAsync(h.mlp_loss.p, 0, 4ull * std::max(steps, 1u), s));
    for (unsigned st = 0; st < steps; ++st) {
        TripletArgs t{h.goff.p, h.gids.p, h.gM.p, h.gN, batch, seed, st, h.mlp_idx.p};
        KG_HIP(kg_triplets(t, h.X.p, h.Dp, D, h.mlp_mu.p, h.mlp_is.p, h.mlp_x.p, s));
        mlp_forward(h, h.mlp_x.p, n, D, h.mlp_h.p, h.mlp_z.p);
        KG_HIP(kg_triplet_loss(h.mlp_z.p, P, batch, margin, h.mlp_dz.p, h.mlp_loss.p + st, s));
        backward(h, h.mlp_x.p, n, h.mlp_h.p, h.mlp_dz.p);
        ++h.adam_step;
        // cosine decay to zero over the call: Adam's normalised steps turn into
        // a random walk once few triplets are active, and the decay ends it
        float const lt = lr * 0.5f * (1.f + std::cos(3.14159265f * float(st) / float(steps)));
        KG_HIP(kg_adam(h.W1.p, h.gW1.p, h.mW1.p, h.vW1.p, H * D, lt, h.adam_step, s));
        KG_HIP(kg_adam(h.b1.p, h.gb1.p, h.mb1.p, h.vb1.p, H, lt, h.adam_step, s));
        KG_HIP(kg_adam(h.W2.p, h.gW2.p, h.mW2.p, h.vW2.p, P * H, lt, h.adam_step, s));
        KG_HIP(kg_adam(h.b2.p, h.gb2.p, h.mb2.p, h.vb2.p, P, lt, h.adam_step, s));
    }
    h.sync();
    if (trace && steps) KG_HIP(hipMemcpy(trace, h.mlp_loss.p, 4ull * steps, hipMemcpyDeviceToHost));
    ++h.mlp_version;
}

void mlp_loss_grad(Handle &h, uint32_t const *triplets, unsigned nt, float margin, float *loss, float *dW1,
                   float *db1, float *dW2, float *db2) {
    require_mlp(h);
    unsigned const D = h.D, H = h.mlpH, P = h.mlpP, n = 3 * nt;
    for (unsigned i = 0; i < n; ++i)
        if (triplets[i] >= h.N) throw invalid_argument("triplet id out of range");
    hipStream_t s = h.stream;
    h.mlp_idx.reserve(n);
    h.mlp_x.reserve(size_t(n) * D);
    h.mlp_h.reserve(size_t(n) * H);
    h.mlp_z.reserve(size_t(n) * P);
    h.mlp_dz.reserve(size_t(n) * P);
    h.mlp_loss.reserve(1);
    KG_HIP(hipMemcpyAsync(h.mlp_idx.p, triplets, 4ull * n, hipMemcpyHostToDevice, s));
    KG_HIP(kg_gather(h.X.p, h.Dp, D, h.mlp_idx.p, n, h.mlp_mu.p, h.mlp_is.p, h.mlp_x.p, s));
    KG_HIP(hipMemsetAsync(h.mlp_loss.p, 0, 4, s));
    mlp_forward(h, h.mlp_x.p, n, D, h.mlp_h.p, h.mlp_z.p);
    KG_HIP(kg_triplet_loss(h.mlp_z.p, P, nt, margin, h.mlp_dz.p, h.mlp_loss.p, s));
    backward(h, h.mlp_x.p, n, h.mlp_h.p, h.mlp_dz.p);
    h.sync();
    if (loss) KG_HIP(hipMemcpy(loss, h.mlp_loss.p, 4, hipMemcpyDeviceToHost));
    if (dW1) KG_HIP(hipMemcpy(dW1, h.gW1.p, 4ull * H * D, hipMemcpyDeviceToHost));
    if (db1) KG_HIP(hipMemcpy(db1, h.gb1.p, 4ull * H, hipMemcpyDeviceToHost));
    if (dW2) KG_HIP(hipMemcpy(dW2, h.gW2.p, 4ull * P * H, hipMemcpyDeviceToHost));
    if (db2) KG_HIP(hipMemcpy(db2, h.gb2.p, 4ull * P, hipMemcpyDeviceToHost));
}

// Projected copy of the entry sample, refreshed when the weights change.
void mlp_project_entries(Handle &h) {
    require_mlp(h);
    if (h.entryE == 0) throw runtime_error("no entry sample (kg_set_entry_sample)");
    if (h.eproj_version == h.mlp_version && h.eproj.n >= size_t(h.entryE) * h.mlpP) return;
    h.eproj.reserve(size_t(h.entryE) * h.mlpP);
    h.hq.reserve(size_t(h.entryE) * h.mlpH);
    mlp_forward_rows(h, h.erows.p, h.entryE, h.Dp, h.hq.p, h.eproj.p);
    h.eprojsq.reserve((h.entryE + 15) / 16 * 16);  // ||z||^2 / 2 of the projected sample
    KG_HIP(kg_launch_norms(KG_METRIC_EUCLIDEAN, h.eproj.p, h.entryE, h.mlpP, h.mlpP, h.eprojsq.p, h.stream));
    h.eproj_tiles.reserve(kg_entry_tiles_size(h.entryE, h.mlpP));  // BF16 operand tiles (a screen)
    KG_HIP(kg_launch_entry_tiles(h.eproj.p, h.entryE, h.mlpP, h.mlpP, h.eproj_tiles.p, h.stream));
    h.sync();
    h.eproj_version = h.mlp_version;
}

}  // namespace kga

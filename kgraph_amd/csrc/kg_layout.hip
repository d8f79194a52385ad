// kg_layout.hip -- the node order of the walk's own copies of the rows and
// the padded adjacency.
//
// A graph walk reads the rows and adjacency rows of nodes that are close in
// the graph, and concurrent walks of nearby queries read the same ones.  The
// build leaves node ids in dataset order, where graph neighbours are spread
// over the whole array.  Laying the walk's copies out in a breadth-first
// order of the graph puts a neighbourhood's rows in nearby lines and pages.
//
// The order (deterministic): weakly connected components of the walk's edges
// (the first min(M, degree) ids of every list) by min-label hooking with path
// compression; one breadth-first tree from each component's smallest node
// (trees in order of that node), a level's children in order of their parent
// and then of their position in the parent's list; a tree's nodes together,
// level after level; nodes no tree reaches last, by id.
//
// The walk maps its starting ids in and its results out (kg_search_args::perm
// and ::inv), so its results are the same in any order.
#include <algorithm>
#include <vector>

#include "kg_device.hpp"
#include "kg_internal.hpp"
#include "kg_nnd.hpp"

namespace {

using namespace kgd;

constexpr unsigned LB = 256;  // threads per block (4 waves; one wave per node where lists are read)

// the edges a walk at search M = 0 follows from a node: min(M, degree),
// capped at the row width
__device__ __forceinline__ unsigned out_deg(uint32_t md, unsigned W) {
    return min(min(md >> 16, md & 0xffffu), W);
}

__device__ __forceinline__ uint64_t claim_key(unsigned i, unsigned j) { return (uint64_t(i) << 16) | j; }

__global__ void __launch_bounds__(LB) lay_iota(unsigned N, uint32_t *lab) {
    for (unsigned v = blockIdx.x * LB + threadIdx.x; v < N; v += gridDim.x * LB) lab[v] = v;
}

// For every edge u -> v whose endpoints' labels differ, the entry of the
// larger label takes the smaller one.  Labels only decrease, each is the id
// of a node of the same weak component, and lab[x] <= x throughout.
// `limit`: the first edges of each list taken (a first round on a few edges
// merges most of a component at a fraction of the atomics).
__global__ void __launch_bounds__(LB) lay_hook(unsigned N, uint32_t const *gpad, uint32_t const *gmd, unsigned W,
                                               unsigned limit, uint32_t *lab, uint32_t *changed) {
    unsigned const u = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (u >= N) return;
    unsigned const cnt = min(out_deg(gmd[u], W), limit);
    uint32_t const *row = gpad + size_t(u) * W;
    bool ch = false;
    for (unsigned j = lane; j < cnt; j += KG_WAVE) {
        uint32_t const v = row[j];
        uint32_t const pu = lab[u], pv = lab[v];
        if (pu != pv) {
            atomicMin(&lab[max(pu, pv)], min(pu, pv));
            ch = true;
        }
    }
    if (ch) *changed = 1u;
}

// Every label becomes the end of its chain (a node whose label is itself);
// the chain's ids strictly decrease, so the walk along it ends.
__global__ void __launch_bounds__(LB) lay_compress(unsigned N, uint32_t *lab) {
    for (unsigned v = blockIdx.x * LB + threadIdx.x; v < N; v += gridDim.x * LB) {
        uint32_t p = lab[v];
        for (;;) {
            uint32_t const q = lab[p];
            if (q == p) break;
            p = q;
        }
        lab[v] = p;
    }
}

__global__ void __launch_bounds__(LB) lay_root_flags(unsigned N, uint32_t const *lab, uint32_t *flag) {
    for (unsigned v = blockIdx.x * LB + threadIdx.x; v < N; v += gridDim.x * LB) flag[v] = lab[v] == v ? 1u : 0u;
}

// level 0: the roots in id order, each its own tree
__global__ void __launch_bounds__(LB) lay_roots(unsigned N, uint32_t const *flag, uint32_t const *pos, uint32_t *ord,
                                                uint32_t *vis, uint32_t *tree) {
    for (unsigned v = blockIdx.x * LB + threadIdx.x; v < N; v += gridDim.x * LB) {
        if (flag[v]) {
            ord[pos[v]] = v;
            vis[v] = 1u;
            tree[v] = v;
        }
    }
}

// Each unvisited neighbour v of frontier node ord[i] (list position j) keeps
// the smallest (i, j) that reaches it: its parent in the next level.
__global__ void __launch_bounds__(LB) lay_claim(unsigned lo, unsigned hi, uint32_t const *ord, uint32_t const *gpad,
                                                uint32_t const *gmd, unsigned W, uint32_t const *vis,
                                                unsigned long long *claim) {
    unsigned const i = lo + blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (i >= hi) return;
    uint32_t const u = ord[i];
    unsigned const cnt = out_deg(gmd[u], W);
    uint32_t const *row = gpad + size_t(u) * W;
    for (unsigned j = lane; j < cnt; j += KG_WAVE) {
        uint32_t const v = row[j];
        if (!vis[v]) atomicMin(&claim[v], static_cast<unsigned long long>(claim_key(i, j)));
    }
}

// children of frontier node i: the list entries whose claim is (i, j) (a
// claim key names one frontier position of one level, so older claims never
// match; a repeated id matches at its first position only)
__global__ void __launch_bounds__(LB) lay_count(unsigned lo, unsigned hi, uint32_t const *ord, uint32_t const *gpad,
                                                uint32_t const *gmd, unsigned W, unsigned long long const *claim,
                                                uint32_t *kids) {
    unsigned const i = lo + blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (i >= hi) return;
    uint32_t const u = ord[i];
    unsigned const cnt = out_deg(gmd[u], W);
    uint32_t const *row = gpad + size_t(u) * W;
    unsigned n = 0;
    for (unsigned c = 0; c < cnt; c += KG_WAVE) {
        unsigned const j = c + lane;
        bool const child = j < cnt && claim[row[j]] == claim_key(i, j);
        n += popc64(ballot(child));
    }
    if (lane == 0) kids[i - lo] = n;
}

// the next level: frontier node i's children at hi + koff[i - lo], in list order
__global__ void __launch_bounds__(LB) lay_emit(unsigned lo, unsigned hi, uint32_t *ord, uint32_t const *gpad,
                                               uint32_t const *gmd, unsigned W, unsigned long long const *claim,
                                               uint32_t const *koff, uint32_t *vis, uint32_t *tree) {
    unsigned const i = lo + blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (i >= hi) return;
    uint32_t const u = ord[i];
    uint32_t const t = tree[u];
    unsigned const cnt = out_deg(gmd[u], W);
    uint32_t const *row = gpad + size_t(u) * W;
    unsigned const base = hi + koff[i - lo];
    uint64_t const lt = (1ull << lane) - 1ull;
    unsigned run = 0;
    for (unsigned c = 0; c < cnt; c += KG_WAVE) {
        unsigned const j = c + lane;
        uint32_t const v = j < cnt ? row[j] : 0u;
        bool const child = j < cnt && claim[v] == claim_key(i, j);
        uint64_t const b = ballot(child);
        if (child) {
            ord[base + run + popc64(b & lt)] = v;
            vis[v] = 1u;
            tree[v] = t;
        }
        run += popc64(b);
    }
}

// A level lists each tree's nodes in one run (level 0 is sorted by tree, and
// children follow their parents).  first[t]: where tree t's run starts.
__global__ void __launch_bounds__(LB) lay_first(unsigned lo, unsigned hi, uint32_t const *ord, uint32_t const *tree,
                                                uint32_t *first) {
    unsigned const i = lo + blockIdx.x * LB + threadIdx.x;
    if (i >= hi) return;
    uint32_t const t = tree[ord[i]];
    if (i == lo || tree[ord[i - 1]] != t) first[t] = i;
}

// node ord[i] goes after its tree's earlier levels (done[t] nodes)
__global__ void __launch_bounds__(LB) lay_place(unsigned lo, unsigned hi, uint32_t const *ord, uint32_t const *tree,
                                                uint32_t const *first, uint32_t const *done, uint32_t const *tbase,
                                                uint32_t *inv) {
    unsigned const i = lo + blockIdx.x * LB + threadIdx.x;
    if (i >= hi) return;
    uint32_t const v = ord[i], t = tree[v];
    inv[v] = tbase[t] + done[t] + (i - first[t]);
}

__global__ void __launch_bounds__(LB) lay_done(unsigned lo, unsigned hi, uint32_t const *ord, uint32_t const *tree,
                                               uint32_t const *first, uint32_t *done) {
    unsigned const i = lo + blockIdx.x * LB + threadIdx.x;
    if (i >= hi) return;
    uint32_t const t = tree[ord[i]];
    if (i + 1 == hi || tree[ord[i + 1]] != t) done[t] += i - first[t] + 1;
}

__global__ void __launch_bounds__(LB) lay_unvisited(unsigned N, uint32_t const *vis, uint32_t *flag) {
    for (unsigned v = blockIdx.x * LB + threadIdx.x; v < N; v += gridDim.x * LB) flag[v] = vis[v] ? 0u : 1u;
}

// the nodes no tree reached, after the R reached ones, by id; then perm = inv^-1
__global__ void __launch_bounds__(LB) lay_finish(unsigned N, unsigned R, uint32_t const *flag, uint32_t const *pos,
                                                 uint32_t *inv, uint32_t *perm) {
    for (unsigned v = blockIdx.x * LB + threadIdx.x; v < N; v += gridDim.x * LB) {
        uint32_t const p = flag[v] ? R + pos[v] : inv[v];
        inv[v] = p;
        perm[p] = v;
    }
}

// out row i = in row perm[i], rows of `words` 32-bit words
__global__ void __launch_bounds__(LB) lay_rows(unsigned N, unsigned words, uint32_t const *perm, uint32_t const *in,
                                               uint32_t *out) {
    size_t const total = size_t(N) * words;
    for (size_t x = size_t(blockIdx.x) * LB + threadIdx.x; x < total; x += size_t(gridDim.x) * LB) {
        size_t const r = x / words, c = x - r * words;
        out[x] = in[size_t(perm[r]) * words + c];
    }
}

// padded adjacency in the new order: row i = row perm[i] with its ids renamed
__global__ void __launch_bounds__(LB) lay_adjacency(unsigned N, unsigned W, uint32_t const *perm, uint32_t const *inv,
                                                    uint32_t const *gpad, uint32_t const *gmd, uint32_t *gpad2,
                                                    uint32_t *gmd2) {
    unsigned const i = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (i >= N) return;
    uint32_t const u = perm[i], md = gmd[u];
    unsigned const cnt = min(md >> 16, W);  // the ids the row holds: min(degree, W)
    uint32_t const *src = gpad + size_t(u) * W;
    uint32_t *dst = gpad2 + size_t(i) * W;
    for (unsigned j = lane; j < cnt; j += KG_WAVE) dst[j] = inv[src[j]];
    if (lane == 0) gmd2[i] = md;
}

unsigned grid_for(size_t n) { return unsigned(std::max<size_t>(1, std::min<size_t>((n + LB - 1) / LB, 8192))); }
unsigned waves_grid(size_t n) { return unsigned((n + 3) / 4); }

}  // namespace

hipError_t kg_walk_order(unsigned N, uint32_t const *gpad, uint32_t const *gmd, unsigned W, kg_order_scratch const &w,
                         uint32_t *perm, uint32_t *inv, unsigned *levels_out, unsigned *rounds_out, hipStream_t s) {
#define KG_TRY(x)                              \
    do {                                       \
        hipError_t const e_ = (x);             \
        if (e_ != hipSuccess) return e_;       \
    } while (0)
    if (levels_out) *levels_out = 0;
    if (rounds_out) *rounds_out = 0;
    if (N == 0) return hipSuccess;
    uint32_t *const host = w.host;
    auto read_word = [&](uint32_t const *dev, uint32_t &out) -> hipError_t {
        KG_TRY(hipMemcpyAsync(host, dev, 4, hipMemcpyDeviceToHost, s));
        KG_TRY(hipStreamSynchronize(s));
        out = host[0];
        return hipSuccess;
    };
    unsigned const gN = grid_for(N);
    // weak components: hook and compress until no edge joins two labels
    hipLaunchKernelGGL(lay_iota, dim3(gN), dim3(LB), 0, s, N, w.lab);
    unsigned rounds = 0;
    for (; rounds < 64; ++rounds) {
        // round 0 on each list's first two edges; the rest on all of them
        // until one changes nothing
        KG_TRY(hipMemsetAsync(w.word, 0, 4, s));
        hipLaunchKernelGGL(lay_hook, dim3(waves_grid(N)), dim3(LB), 0, s, N, gpad, gmd, W, rounds ? W : 2u, w.lab,
                           w.word);
        hipLaunchKernelGGL(lay_compress, dim3(gN), dim3(LB), 0, s, N, w.lab);
        KG_TRY(hipGetLastError());
        if (rounds == 0) continue;
        uint32_t changed = 0;
        KG_TRY(read_word(w.word, changed));
        if (!changed) break;
    }
    if (rounds_out) *rounds_out = rounds + 1;
    // level 0: the component roots (every node whose label is itself)
    KG_TRY(hipMemsetAsync(w.vis, 0, size_t(N) * 4, s));
    KG_TRY(hipMemsetAsync(w.claim, 0xff, size_t(N) * 8, s));
    hipLaunchKernelGGL(lay_root_flags, dim3(gN), dim3(LB), 0, s, N, w.lab, w.flag);
    KG_TRY(kg_scan_u32(w.flag, w.pos, N, w.tmp, s));
    hipLaunchKernelGGL(lay_roots, dim3(gN), dim3(LB), 0, s, N, w.flag, w.pos, w.ord, w.vis, w.tree);
    KG_TRY(hipGetLastError());
    uint32_t roots = 0;
    KG_TRY(read_word(w.pos + N, roots));
    // breadth-first levels of all trees at once
    std::vector<unsigned> lev{0u, roots};
    unsigned lo = 0, hi = roots;
    while (hi > lo && hi < N && lev.size() <= 4096) {
        unsigned const nf = hi - lo;
        hipLaunchKernelGGL(lay_claim, dim3(waves_grid(nf)), dim3(LB), 0, s, lo, hi, w.ord, gpad, gmd, W, w.vis,
                           w.claim);
        hipLaunchKernelGGL(lay_count, dim3(waves_grid(nf)), dim3(LB), 0, s, lo, hi, w.ord, gpad, gmd, W, w.claim,
                           w.kids);
        KG_TRY(hipGetLastError());
        KG_TRY(kg_scan_u32(w.kids, w.koff, nf, w.tmp, s));
        uint32_t next = 0;
        KG_TRY(read_word(w.koff + nf, next));
        if (next == 0 || next > N - hi) break;  // (each node is emitted once: next <= N - hi)
        hipLaunchKernelGGL(lay_emit, dim3(waves_grid(nf)), dim3(LB), 0, s, lo, hi, w.ord, gpad, gmd, W, w.claim,
                           w.koff, w.vis, w.tree);
        KG_TRY(hipGetLastError());
        lo = hi;
        hi += next;
        lev.push_back(hi);
    }
    unsigned const R = hi;
    // positions: trees in root order, a tree's levels in order.  A first pass
    // over the levels sums each tree's runs into its size (no atomics: a
    // level holds one run per tree), the second places the nodes.
    KG_TRY(hipMemsetAsync(w.tsize, 0, size_t(N) * 4, s));
    KG_TRY(hipMemsetAsync(w.done, 0, size_t(N) * 4, s));
    for (size_t l = 0; l + 1 < lev.size(); ++l) {
        unsigned const a = lev[l], b = lev[l + 1];
        if (b <= a) continue;
        dim3 const g((b - a + LB - 1) / LB);
        hipLaunchKernelGGL(lay_first, g, dim3(LB), 0, s, a, b, w.ord, w.tree, w.first);
        hipLaunchKernelGGL(lay_done, g, dim3(LB), 0, s, a, b, w.ord, w.tree, w.first, w.tsize);
    }
    KG_TRY(kg_scan_u32(w.tsize, w.tbase, N, w.tmp, s));
    for (size_t l = 0; l + 1 < lev.size(); ++l) {
        unsigned const a = lev[l], b = lev[l + 1];
        if (b <= a) continue;
        dim3 const g((b - a + LB - 1) / LB);
        hipLaunchKernelGGL(lay_first, g, dim3(LB), 0, s, a, b, w.ord, w.tree, w.first);
        hipLaunchKernelGGL(lay_place, g, dim3(LB), 0, s, a, b, w.ord, w.tree, w.first, w.done, w.tbase, inv);
        hipLaunchKernelGGL(lay_done, g, dim3(LB), 0, s, a, b, w.ord, w.tree, w.first, w.done);
    }
    hipLaunchKernelGGL(lay_unvisited, dim3(gN), dim3(LB), 0, s, N, w.vis, w.flag);
    KG_TRY(kg_scan_u32(w.flag, w.pos, N, w.tmp, s));
    hipLaunchKernelGGL(lay_finish, dim3(gN), dim3(LB), 0, s, N, R, w.flag, w.pos, inv, perm);
    KG_TRY(hipGetLastError());
    if (levels_out) *levels_out = unsigned(lev.size() - 1);
    return hipSuccess;
#undef KG_TRY
}

hipError_t kg_permute_rows(unsigned N, unsigned words, uint32_t const *perm, void const *in, void *out,
                           hipStream_t s) {
    size_t const total = size_t(N) * words;
    if (total == 0) return hipSuccess;
    hipLaunchKernelGGL(lay_rows, dim3(grid_for(total)), dim3(LB), 0, s, N, words, perm,
                       static_cast<uint32_t const *>(in), static_cast<uint32_t *>(out));
    return hipGetLastError();
}

hipError_t kg_permute_adjacency(unsigned N, unsigned W, uint32_t const *perm, uint32_t const *inv,
                                uint32_t const *gpad, uint32_t const *gmd, uint32_t *gpad2, uint32_t *gmd2,
                                hipStream_t s) {
    if (N == 0) return hipSuccess;
    hipLaunchKernelGGL(lay_adjacency, dim3(waves_grid(N)), dim3(LB), 0, s, N, W, perm, inv, gpad, gmd, gpad2, gmd2);
    return hipGetLastError();
}

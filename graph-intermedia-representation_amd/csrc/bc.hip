// Betweenness centrality: replaces GIR::bc (gir_query.cpp:728-778), Brandes
// over k sources on the directed arcs with unit length.
//
// Up to BB sources advance together, one grid row (blockIdx.y) per source,
// all of them level by level in lockstep:
//   forward (level d -> d + 1): the vertices of level d, a segment of the
//     source's visit order, expand their out-arcs; an unvisited head is
//     claimed (CAS on its level) and appended to the order, and every arc
//     into level d + 1 adds the tail's shortest-path count to the head's
//     (fp64 atomicAdd: the counts are integers, exact in any order below
//     2^53);
//   backward (level d, deepest first): every vertex w of level d pushes
//     sigma[u] * (1 + delta[w]) / sigma[w] into delta[u] of each in-neighbour
//     u on level d - 1 (fp64 atomicAdd);
//   then score[v] += delta_s[v] over the batch's sources in source order
//     (v != s), one thread per vertex, so the sum order is fixed.
// Level boundaries live on the device (lv[s][d] = where level d starts in
// the order), so the forward kernels read their ranges there and a chunk of
// levels runs per host readback.  Rows longer than HEAVY_T go through the
// heavy list.
//
// HBM traffic per source (algorithmic): 20V (level, sigma, delta init) +
// per reached vertex 12 (order, two row offsets) + per reached arc
// 2 x (4 adjacency + 4 level + 8 sigma or delta).
#include <vector>

#include "common.h"

namespace gir {

std::vector<int32_t> hostSources(girhip_graph* h, const int32_t* src, int32_t k);

namespace {

constexpr int BB = 16;      // sources per batch
constexpr int BC_CHUNK = 8;  // forward levels per host readback

struct BcBatch {
    int32_t* lev;    // [BB][V] level, -1 = unreached
    double* sigma;   // [BB][V] shortest-path counts
    double* delta;   // [BB][V] dependencies
    int32_t* ord;    // [BB][V] visit order (levels in sequence)
    int32_t* lv;     // [BB][LW] level starts in the order (LW = V + 2 + BC_CHUNK: a chunk may step past the last level)
    int64_t* tail;   // [BB * 16] order length (one 128-byte line per source)
    int64_t* last;   // [BB] size of the level appended last (host readback)
    int64_t V;
    int64_t LW;
    int k;
};

__global__ void kBcSeed(BcBatch B, const int32_t* __restrict__ src) {
    const int s = threadIdx.x;
    if (s >= BB) return;
    B.tail[s * 16] = 0;
    B.last[s] = 0;
    if (s >= B.k) return;  // rows past the batch have no arrays
    int32_t* lv = B.lv + (size_t)s * B.LW;
    lv[0] = 0;
    lv[1] = 0;
    const int32_t x = src[s];
    if (x < 0 || x >= B.V) return;  // skipped, as the reference does
    const size_t o = (size_t)s * B.V;
    B.lev[o + x] = 0;
    B.sigma[o + x] = 1.0;
    B.ord[o] = x;
    B.tail[s * 16] = 1;
    B.last[s] = 1;
    lv[1] = 1;
}

// Forward work on out-arc j of u (level d, path count in pay) for source s.
struct BcForwardF {
    BcBatch B;
    const int32_t* adj;
    int32_t d;
    __device__ int64_t operator()(int32_t, uint64_t pay, int32_t s, int64_t j) const {
        const int32_t w = adj[j];
        const size_t o = (size_t)s * B.V;
        int32_t old = B.lev[o + w];
        bool claimed = false;
        if (old == -1) {
            old = atomicCAS(&B.lev[o + w], -1, d + 1);
            if (old == -1) {
                claimed = true;
                old = d + 1;
            }
        }
        const int64_t pos = wavePushAny(&B.tail[s * 16], claimed);  // heavy-pass waves mix sources
        if (claimed) B.ord[o + pos] = w;
        if (old == d + 1) unsafeAtomicAdd(&B.sigma[o + w], __longlong_as_double((long long)pay));
        return 0;
    }
};

__global__ void kBcForward(BcForwardF f, const int32_t* __restrict__ beg, HeavyList H) {
    const int s = blockIdx.y;
    const BcBatch& B = f.B;
    const int32_t* lv = B.lv + (size_t)s * B.LW;
    const int64_t lo = lv[f.d], hi = lv[f.d + 1];
    const size_t o = (size_t)s * B.V;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t base = lo + (int64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~(WAVE - 1)); base < hi;
         base += stride) {
        const int64_t i = base + laneId();
        const bool act = i < hi;
        const int32_t u = act ? B.ord[o + i] : 0;
        const double su = act ? B.sigma[o + u] : 0.0;
        const int64_t b = act ? beg[u] : 0, e = act ? beg[u + 1] : 0;
        waveRowsH(act, u, (uint64_t)__double_as_longlong(su), s, b, e, H, f);
    }
}

// After the forward step from level d: level d + 2 starts at the tail.
__global__ void kBcAdvance(BcBatch B, int32_t d) {
    const int s = threadIdx.x;
    if (s >= B.k) return;
    int32_t* lv = B.lv + (size_t)s * B.LW;
    const int64_t t = B.tail[s * 16];
    lv[d + 2] = (int32_t)t;
    B.last[s] = t - lv[d + 1];
}

// Backward work on in-arc j of w (level d, (1 + delta[w]) / sigma[w] in
// pay) for source s: the in-neighbour on level d - 1 takes its share.
struct BcBackwardF {
    BcBatch B;
    const int32_t* adj;
    int32_t d;
    __device__ int64_t operator()(int32_t, uint64_t pay, int32_t s, int64_t j) const {
        const int32_t u = adj[j];
        const size_t o = (size_t)s * B.V;
        if (B.lev[o + u] == d - 1)
            unsafeAtomicAdd(&B.delta[o + u], B.sigma[o + u] * __longlong_as_double((long long)pay));
        return 0;
    }
};

__global__ void kBcBackward(BcBackwardF f, const int32_t* __restrict__ beg, HeavyList H) {
    const int s = blockIdx.y;
    const BcBatch& B = f.B;
    const int32_t* lv = B.lv + (size_t)s * B.LW;
    const int64_t lo = lv[f.d], hi = lv[f.d + 1];
    const size_t o = (size_t)s * B.V;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t base = lo + (int64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~(WAVE - 1)); base < hi;
         base += stride) {
        const int64_t i = base + laneId();
        const bool act = i < hi;
        const int32_t w = act ? B.ord[o + i] : 0;
        const double coef = act ? (1.0 + B.delta[o + w]) / B.sigma[o + w] : 0.0;
        const int64_t b = act ? beg[w] : 0, e = act ? beg[w + 1] : 0;
        waveRowsH(act, w, (uint64_t)__double_as_longlong(coef), s, b, e, H, f);
    }
}

// score[v] += sum over the batch's sources (fixed order) of delta_s[v], v != s.
__global__ void kBcAccumulate(BcBatch B, const int32_t* __restrict__ src, double* __restrict__ score) {
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < B.V; v += (int64_t)gridDim.x * blockDim.x) {
        double acc = 0.0;
        for (int s = 0; s < B.k; ++s)
            if (src[s] != (int32_t)v && B.lev[(size_t)s * B.V + v] >= 0) acc += B.delta[(size_t)s * B.V + v];
        score[v] += acc;
    }
}

}  // namespace
}  // namespace gir

using namespace gir;

extern "C" int girhip_bc(girhip_graph* h, const int32_t* src, int32_t k, double* score) {
    return guarded([&] {
        useDevice(h);
        if (k < 0 || (k > 0 && !src) || !score) raise(GIRHIP_EINVAL, "bc: bad arguments");
        const int32_t V = h->V;
        std::vector<int32_t> s = hostSources(h, src, k);
        const size_t Vs = (size_t)(V > 0 ? V : 1);
        Scratch& S = scratch(h);
        const int kc = std::max(1, std::min<int32_t>(k, BB));
        S.bc_lev.ensure((size_t)kc * Vs);
        S.bc_sigma.ensure((size_t)kc * Vs);
        S.bc_delta.ensure((size_t)kc * Vs);
        S.bc_ord.ensure((size_t)kc * Vs);
        const size_t LW = Vs + 2 + BC_CHUNK;
        S.bc_lv.ensure((size_t)kc * LW);
        S.bc_ctl.ensure(BB * 16 + BB);
        S.src.ensure((size_t)(k > 0 ? k : 1));
        S.stage.ensure(sizeof(double) * Vs);
        double* out = (double*)S.stage.p;
        const bool direct = isDevicePtr(score);
        if (direct) out = score;
        fillU64(h, (uint64_t*)out, 0, V);  // 0.0
        if (k > 0) copyIn(h, S.src.p, s.data(), sizeof(int32_t) * (size_t)k);
        const int blk = (int)h->prm.block;
        tBegin(h);
        for (int32_t i0 = 0; i0 < k && V > 0; i0 += BB) {
            const int c = std::min<int32_t>(BB, k - i0);
            BcBatch B;
            B.lev = S.bc_lev.p;
            B.sigma = (double*)S.bc_sigma.p;
            B.delta = (double*)S.bc_delta.p;
            B.ord = S.bc_ord.p;
            B.lv = S.bc_lv.p;
            B.tail = S.bc_ctl.p;
            B.last = S.bc_ctl.p + BB * 16;
            B.V = V;
            B.LW = (int64_t)LW;
            B.k = c;
            fillI32(h, B.lev, -1, (int64_t)c * V);
            fillU64(h, (uint64_t*)B.sigma, 0, (int64_t)c * V);
            fillU64(h, (uint64_t*)B.delta, 0, (int64_t)c * V);
            const int32_t* dsrc = S.src.p + i0;
            launch(h, "bc_seed", kBcSeed, 1, 64, 0, B, dsrc);
            // forward levels, a chunk per readback
            int32_t d = 0;
            int64_t width = 1, prev = 0;  // largest level of the last two readbacks
            int Cp = 1;
            std::vector<int64_t> last(BB);
            while (true) {
                // a chunk of levels only while levels stay small and are not
                // growing fast (its grid is sized at the chunk's start)
                const int C = (prev > 0 && width * 64 < (int64_t)V && width <= 2 * prev)
                                  ? std::min(BC_CHUNK, 2 * Cp)
                                  : 1;
                Cp = C;
                prev = width;
                const unsigned g = C > 1 ? std::min<unsigned>(1024, std::max<unsigned>(16, gridFor(8 * width, blk)))
                                         : std::min<unsigned>(1024, std::max<unsigned>(1, gridFor(width, blk)));
                for (int t = 0; t < C; ++t, ++d) {
                    const BcForwardF f{B, h->fw_adj.p, d};
                    launch(h, "bc_forward", kBcForward, dim3(g, c), blk, 0, f, (const int32_t*)h->fw_beg.p,
                           heavyList(h));
                    launchHeavyStore(h, "bc_forward_rows", f);
                    launch(h, "bc_advance", kBcAdvance, 1, 64, 0, B, d);
                }
                readCounters(h, B.last, BB, last.data());
                width = 0;
                for (int y = 0; y < c; ++y) width = std::max(width, last[y]);
                if (width == 0) break;
            }
            // levels 0 .. d hold every reached vertex (the deepest ones empty);
            // backward from the deepest
            for (int32_t L = d; L >= 1; --L) {
                const BcBackwardF f{B, h->bw_adj.p, L};
                launch(h, "bc_backward", kBcBackward, dim3(1024, c), blk, 0, f, (const int32_t*)h->bw_beg.p,
                       heavyList(h));
                launchHeavyStore(h, "bc_backward_rows", f);
            }
            launch(h, "bc_accumulate", kBcAccumulate, boundedGrid(V, 256), 256, 0, B, dsrc, out);
        }
        tEnd(h, "bc");
        if (!direct) copyOut(h, score, out, sizeof(double) * (size_t)V);
        syncStream(h);
    });
}

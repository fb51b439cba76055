// Weakly connected components on gfx950.
//
// Replaces GIR::buildWCC (gir.cpp:593-655), whose sequential union-find over
// the forward arcs gives each component its union-find root as label.  Here:
// sampled-neighbour hooking (Afforest, Sutton et al. 2018) with pointer
// jumping.  A link only ever makes a root point at a smaller root, so the
// minimum vertex of a component is never re-parented and ends up as the root
// of its tree: after the final compression every label is the component's
// minimum vertex id, the canonical id, with no extra relabelling pass.
//
//   1. comp[v] = v
//   2. for r < rounds: hook v to its r-th out-neighbour; compress
//   3. sample 1024 vertices, take the most frequent root c (the giant)
//   4. every vertex outside c hooks its remaining out-neighbours and all its
//      in-neighbours (an arc with one end outside c is seen from that end)
//   5. compress
//
// HBM traffic: one pass over the forward CSR (4(V+1) + 4E bytes) plus the
// label array (8V bytes: init write + final read/write); arcs of vertices in
// the sampled giant component are skipped.
#include "common.h"

namespace gir {

// Hook the trees of u and v together, larger root under smaller root.
__device__ __forceinline__ void hook(int32_t u, int32_t v, int32_t* comp) {
    int32_t p1 = ldAcq(&comp[u]), p2 = ldAcq(&comp[v]);
    while (p1 != p2) {
        const int32_t high = p1 > p2 ? p1 : p2, low = p1 + p2 - high;
        const int32_t ph = ldAcq(&comp[high]);
        if (ph == low) break;
        if (ph == high && atomicCAS(&comp[high], high, low) == high) break;
        p1 = ldAcq(&comp[ldAcq(&comp[high])]);
        p2 = ldAcq(&comp[low]);
    }
}

__global__ void kWccInit(int32_t* comp, int32_t V) {
    int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v < V) comp[v] = (int32_t)v;
}

__global__ void kWccHookRound(const int32_t* __restrict__ beg, const int32_t* __restrict__ adj, int32_t* comp,
                              int32_t V, int r) {
    int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    const int32_t b = beg[v], e = beg[v + 1];
    if (e - b > r) hook((int32_t)v, adj[b + r], comp);
}

__global__ void kWccCompress(int32_t* comp, int32_t V) {
    int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    int32_t c = ldAcq(&comp[v]);
    int32_t cc = ldAcq(&comp[c]);
    while (c != cc) {
        c = cc;
        cc = ldAcq(&comp[c]);
    }
    stRel(&comp[v], c);
}

// One block of 1024 threads: sample 1,024 labels and count them in an LDS
// hash table (2,048 slots, linear probing); the most frequent label (ties:
// the larger) is the giant.  Which label is taken only steers which
// vertices skip step 4's arcs, never a label.
__global__ void __launch_bounds__(1024) kWccSample(const int32_t* __restrict__ comp, int32_t V, uint64_t seed,
                                                   int64_t* out) {
    constexpr int SLOTS = 2048;
    __shared__ int32_t key[SLOTS];
    __shared__ int32_t cnt[SLOTS];
    __shared__ unsigned long long best;
    const int t = threadIdx.x;
    for (int i = t; i < SLOTS; i += blockDim.x) {
        key[i] = -1;
        cnt[i] = 0;
    }
    if (t == 0) best = 0;
    uint64_t x = seed + (uint64_t)t * 0x9E3779B97F4A7C15ULL;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
    x ^= x >> 31;
    const int32_t lab = comp[(int32_t)(((x >> 32) * (uint64_t)V) >> 32)];
    __syncthreads();
    uint32_t hs = (uint32_t)lab * 2654435761u;
    for (int probe = 0; probe < SLOTS; ++probe) {  // (at most 1,024 keys in 2,048 slots)
        const int i = (int)((hs + (uint32_t)probe) & (SLOTS - 1));
        const int32_t old = atomicCAS(&key[i], -1, lab);
        if (old == -1 || old == lab) {
            atomicAdd(&cnt[i], 1);
            break;
        }
    }
    __syncthreads();
    for (int i = t; i < SLOTS; i += blockDim.x)
        if (cnt[i]) atomicMax(&best, ((unsigned long long)cnt[i] << 32) | (uint32_t)key[i]);
    __syncthreads();
    if (t == 0) out[0] = (int64_t)(uint32_t)(best & 0xffffffffULL);
}

// aux 0: slot j of the forward row, aux 1: of the backward row.
struct HookF {
    const int32_t* fa;
    const int32_t* ba;
    int32_t* comp;
    __device__ int64_t operator()(int32_t u, uint64_t, int32_t aux, int64_t j) const {
        hook(u, aux ? ba[j] : fa[j], comp);
        return 0;
    }
};

__global__ void kWccHookRest(const int32_t* __restrict__ fb, const int32_t* __restrict__ bb, int32_t V, int rounds,
                             const int64_t* __restrict__ giant, HookF f, HeavyList H) {
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int32_t c = (int32_t)giant[0];
    const bool act = v < V && ldAcq(&f.comp[v]) != c;
    int32_t b = 0, e = 0, ib = 0, ie = 0;
    if (act) {
        b = fb[v] + rounds;
        e = fb[v + 1];
        if (b > e) b = e;
        ib = bb[v];
        ie = bb[v + 1];
    }
    waveRowsH(act, (int32_t)v, 0, 0, b, e, H, f);
    waveRowsH(act, (int32_t)v, 0, 1, ib, ie, H, f);
}

void runWcc(girhip_graph* h) {
    const int32_t V = h->V;
    const int blk = (int)h->prm.block;
    h->wcc.ensure((size_t)(V > 0 ? V : 1));
    int32_t* comp = h->wcc.p;
    if (V == 0) {
        h->n_wcc = 0;
        return;
    }
    const int rounds = (int)h->prm.wcc_rounds;
    const unsigned g = gridFor(V, blk);
    launch(h, "wcc_init", kWccInit, g, blk, 0, comp, V);
    for (int r = 0; r < rounds; ++r) {
        launch(h, "wcc_hook_round", kWccHookRound, g, blk, 0, (const int32_t*)h->fw_beg.p, (const int32_t*)h->fw_adj.p,
               comp, V, r);
        launch(h, "wcc_compress", kWccCompress, g, blk, 0, comp, V);
    }
    launch(h, "wcc_sample", kWccSample, 1, 1024, 0, (const int32_t*)comp, V, (uint64_t)0x5EED, h->dcnt.p + 8);
    const HookF f{h->fw_adj.p, h->bw_adj.p, comp};
    launch(h, "wcc_hook_rest", kWccHookRest, g, blk, 0, (const int32_t*)h->fw_beg.p, (const int32_t*)h->bw_beg.p, V,
           rounds, (const int64_t*)h->dcnt.p + 8, f, heavyList(h));
    launchHeavyStore(h, "wcc_hook_rest_heavy", f);
    launch(h, "wcc_compress", kWccCompress, g, blk, 0, comp, V);
    fillI64(h, h->dcnt.p, 0, 1);
    launch(h, "count_roots", kCountSelf, reduceGrid(V, blk), blk, 0, (const int32_t*)comp, V, h->dcnt.p);
    int64_t n;
    readCounters(h, h->dcnt.p, 1, &n);
    h->n_wcc = (int32_t)n;
}

__global__ void kWccHookKeys(const uint64_t* __restrict__ keys, int64_t n, int B, int32_t* comp) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint64_t k = keys[i];
    hook((int32_t)(k >> B), (int32_t)(k & ((1ULL << B) - 1)), comp);
}

// Insert-only repair: the current labels are a min-rooted forest of depth
// one, so hooking the inserted arcs and compressing again yields the labels
// of the grown graph, still canonical.
void runWccHookArcs(girhip_graph* h, const uint64_t* keys, int64_t n, int B) {
    const int32_t V = h->V;
    const int blk = (int)h->prm.block;
    int32_t* comp = h->wcc.p;
    launch(h, "wcc_hook_keys", kWccHookKeys, gridFor(n, blk), blk, 0, keys, n, B, comp);
    launch(h, "wcc_compress", kWccCompress, gridFor(V, blk), blk, 0, comp, V);
    fillI64(h, h->dcnt.p, 0, 1);
    launch(h, "count_roots", kCountSelf, reduceGrid(V, blk), blk, 0, (const int32_t*)comp, V, h->dcnt.p);
    int64_t c;
    readCounters(h, h->dcnt.p, 1, &c);
    h->n_wcc = (int32_t)c;
}

}  // namespace gir

using namespace gir;

extern "C" int girhip_get_wcc(girhip_graph* h, int32_t* label, int32_t* n_wcc) {
    return guarded([&] {
        useDevice(h);
        if (!h->wcc_req) raise(GIRHIP_ESTATE, "get_wcc: call girhip_wcc first");
        if (label) copyOut(h, label, h->wcc.p, sizeof(int32_t) * (size_t)h->V);
        syncStream(h);
        if (n_wcc) *n_wcc = h->n_wcc;
    });
}

extern "C" int girhip_wcc(girhip_graph* h, int32_t* label, int32_t* n_wcc) {
    return guarded([&] {
        useDevice(h);
        tBegin(h);
        runWcc(h);
        tEnd(h, "wcc");
        reserveBatch(h, 4096);  // (allocated once)
        h->wcc_req = true;
        h->wcc_ok = true;
        if (label) {
            copyOut(h, label, h->wcc.p, sizeof(int32_t) * (size_t)h->V);
            syncStream(h);
        }
        if (n_wcc) *n_wcc = h->n_wcc;
    });
}

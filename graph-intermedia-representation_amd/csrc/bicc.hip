// Biconnected blocks of every SCC's undirected view: replaces the block
// layer of buildBCT / getBCTComponents / articulationPoint / biconnected
// (gir.cpp:1650-3001, 2063-2110, 4140-4149; gir.h:133-150).  The view of an
// SCC is the reference's (gir_test.cpp:448-465): its vertices, an edge
// {u, w} wherever u -> w or w -> u, no self loops; a vertex of a
// single-vertex SCC is a block of its own.
//
// Tarjan-Vishkin over a BFS spanning forest, one tree per SCC rooted at the
// SCC's canonical id (all SCCs at once):
//   1. forest: a multi-source direction-optimising BFS over the symmetric
//      closure from every canonical id, following only arcs inside an SCC;
//      the parent of v is the first closure neighbour one level up;
//   2. vertices sorted by level, children grouped by parent (radix sorts);
//      subtree sizes bottom-up, preorder numbers top-down (a root's number
//      is the prefix of the tree sizes, a child's its parent's + 1 + the
//      sizes of the siblings before it);
//   3. low / high = extreme preorder numbers reachable from a subtree by one
//      non-tree edge (each vertex's own row by a lane, or a block for a long
//      row, then bottom-up by levels);
//   4. the tree edges (named by their child) are joined in a union-find:
//      (v, p(v)) with (w, p(w)) for a non-tree edge {v, w} whose ends are
//      unrelated; (w, v) with (v, p(v)) when subtree(w) has a non-tree edge
//      leaving subtree(v) (low(w) < pre(v) or high(w) >= pre(v) + size(v));
//   5. a block = a union-find class; its vertices are the ends of its tree
//      edges; a vertex in two or more blocks is an articulation point.
// The level passes combine the atomics of lanes that share a parent (a hub's
// children are adjacent in a level), and the union-find classes are
// compressed before the non-tree edges join them, so most of those checks
// are two reads.
#include <vector>

#include "common.h"
#include "reach.h"

namespace gir {
namespace {

struct SameScc {
    const int32_t* scc;
    __device__ bool operator()(int32_t w, int32_t p) const { return scc[w] == scc[p]; }
    __device__ bool self(int32_t) const { return true; }
};

// BR_ITEMS vertices per thread (coalesced, blockDim apart): canonical ids ->
// queue, their closure degrees summed; one atomic per block on each counter
constexpr int BR_ITEMS = 8;
__global__ void kBiRoots(const int32_t* __restrict__ scc, const int32_t* __restrict__ sb, int32_t V, int32_t* q,
                         int64_t* cnt) {
    const int64_t v0 = (int64_t)blockIdx.x * blockDim.x * BR_ITEMS + threadIdx.x;
    unsigned root = 0;
    int n = 0;
    int64_t m = 0;
#pragma unroll
    for (int t = 0; t < BR_ITEMS; ++t) {
        const int64_t v = v0 + (int64_t)t * blockDim.x;
        if (v < V && scc[v] == (int32_t)v) {
            root |= 1u << t;
            ++n;
            m += sb[v + 1] - sb[v];
        }
    }
    int64_t pos = blockReserve(n, &cnt[0]);
#pragma unroll
    for (int t = 0; t < BR_ITEMS; ++t)
        if ((root >> t) & 1) q[pos++] = (int32_t)(v0 + (int64_t)t * blockDim.x);
    blockAtomicAdd(m, &cnt[1]);
}

// parent = first closure neighbour in the same SCC one level up
__global__ void kBiParents(const int32_t* __restrict__ sb, const int32_t* __restrict__ sa,
                           const int32_t* __restrict__ scc, const int32_t* __restrict__ lev, int32_t V,
                           int32_t* __restrict__ par, int32_t* maxlev) {
    int32_t ml = 0;
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < V; v += (int64_t)gridDim.x * blockDim.x) {
        const int32_t L = lev[v];
        ml = L > ml ? L : ml;
        int32_t p = -1;
        if (L > 0) {
            const int32_t c = scc[v];
            for (int64_t e = sb[v]; e < sb[v + 1]; ++e) {
                const int32_t w = sa[e];
                if (lev[w] == L - 1 && scc[w] == c) {
                    p = w;
                    break;
                }
            }
        }
        par[v] = p;
    }
    const uint64_t mb = blockMaxU64((uint64_t)(uint32_t)ml);  // one atomic per block
    if (threadIdx.x == 0) atomicMax(maxlev, (int32_t)mb);
}

__global__ void kBiLevelKeys(const int32_t* __restrict__ lev, int32_t V, int B, uint64_t* __restrict__ keys) {
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v < V) keys[v] = ((uint64_t)(uint32_t)lev[v] << B) | (uint64_t)v;
}

// sorted (key >> B) groups -> ids and group starts beg[0 .. G] (lower bounds)
__global__ void kBiGroups(const uint64_t* __restrict__ keys, int64_t n, int B, int64_t G, int32_t* __restrict__ ids,
                          int32_t* __restrict__ beg, int32_t* __restrict__ pos) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t mask = (1ULL << B) - 1;
    if (i < n) {
        ids[i] = (int32_t)(keys[i] & mask);
        if (pos) pos[keys[i] & mask] = (int32_t)i;
    }
    if (i <= G) {
        const uint64_t key = (uint64_t)i << B;
        int64_t lo = 0, hi = n;
        while (lo < hi) {
            const int64_t m = lo + ((hi - lo) >> 1);
            if (keys[m] < key) lo = m + 1;
            else hi = m;
        }
        beg[i] = (int32_t)lo;
    }
}

__global__ void kBiChildKeys(const int32_t* __restrict__ par, int32_t V, int B, uint64_t* __restrict__ keys,
                             int64_t* cnt) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t base = (int64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~(WAVE - 1)); base < V; base += stride) {
        const int64_t v = base + laneId();
        const bool child = v < V && par[v] >= 0;
        const int64_t pos = wavePush(cnt, child);
        if (child) keys[pos] = ((uint64_t)(uint32_t)par[v] << B) | (uint64_t)v;
    }
}

// Lanes whose key equals the wave's first active key combine their values
// (sum / min / max) into one atomic on that key: the children of a hub sit
// side by side in a level and would otherwise all hit one address.
template <class Op>
__device__ __forceinline__ void waveCombine(bool act, int32_t key, int32_t val, Op op, int32_t* dst) {
    const uint64_t am = __ballot(act);
    if (!am) return;
    const int lead = __ffsll((long long)am) - 1;
    const int32_t k0 = __shfl(key, lead);
    const bool same = act && key == k0;
    int32_t x = same ? val : op.id;
    for (int d = WAVE / 2; d > 0; d >>= 1) x = op(x, __shfl_xor(x, d));
    if (laneId() == lead) op.apply(&dst[k0], x);
    if (act && !same) op.apply(&dst[key], val);
}
struct OpAdd {
    int32_t id = 0;
    __device__ int32_t operator()(int32_t a, int32_t b) const { return a + b; }
    __device__ void apply(int32_t* p, int32_t x) const { atomicAdd(p, x); }
};
struct OpMin {
    int32_t id = 0x7fffffff;
    __device__ int32_t operator()(int32_t a, int32_t b) const { return a < b ? a : b; }
    __device__ void apply(int32_t* p, int32_t x) const { atomicMin(p, x); }
};
struct OpMax {
    int32_t id = -1;
    __device__ int32_t operator()(int32_t a, int32_t b) const { return a > b ? a : b; }
    __device__ void apply(int32_t* p, int32_t x) const { atomicMax(p, x); }
};

// level L vertices add their subtree sizes to their parents
__global__ void kBiSizeUp(const int32_t* __restrict__ order, const int32_t* __restrict__ lbeg, int32_t L,
                          const int32_t* __restrict__ par, int32_t* sz) {
    const int64_t lo = lbeg[L], hi = lbeg[L + 1];
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t base = lo + (int64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~(WAVE - 1)); base < hi; base += stride) {
        const int64_t i = base + laneId();
        const bool act = i < hi;
        const int32_t v = act ? order[i] : 0;
        waveCombine(act, act ? par[v] : 0, act ? sz[v] : 0, OpAdd{}, sz);
    }
}

__global__ void kBiRootSizes(const int32_t* __restrict__ par, const int32_t* __restrict__ sz, int32_t V,
                             int32_t* __restrict__ rs) {
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v < V) rs[v] = par[v] < 0 ? sz[v] : 0;
}

__global__ void kBiChildSizes(const int32_t* __restrict__ cadj, int64_t nc, const int32_t* __restrict__ sz,
                              int32_t* __restrict__ cs) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nc) cs[i] = sz[cadj[i]];
}

__global__ void kBiRootPre(const int32_t* __restrict__ par, const int64_t* __restrict__ rscan, int32_t V,
                           int32_t* __restrict__ pre) {
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v < V && par[v] < 0) pre[v] = (int32_t)rscan[v];
}

// level L vertices: pre = parent's + 1 + sizes of the earlier siblings
__global__ void kBiPreDown(const int32_t* __restrict__ order, const int32_t* __restrict__ lbeg, int32_t L,
                           const int32_t* __restrict__ par, const int32_t* __restrict__ cbeg,
                           const int32_t* __restrict__ cpos, const int64_t* __restrict__ cscan, int32_t* pre) {
    const int64_t lo = lbeg[L], hi = lbeg[L + 1];
    for (int64_t i = lo + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < hi; i += (int64_t)gridDim.x * blockDim.x) {
        const int32_t v = order[i];
        const int32_t p = par[v];
        pre[v] = pre[p] + 1 + (int32_t)(cscan[cpos[v]] - cscan[cbeg[p]]);
    }
}

// Local low / high of v: its preorder number and those of the other ends of
// its non-tree edges inside the SCC.  Rows up to BI_ROW closure arcs are
// scanned by one lane; longer ones are listed for kBiLowHighLong.
constexpr int64_t BI_ROW = 64;

__device__ __forceinline__ bool nonTree(const int32_t* scc, const int32_t* par, int32_t v, int32_t w) {
    return scc[w] == scc[v] && w != par[v] && par[w] != v;
}

__global__ void kBiLowHigh(const int32_t* __restrict__ sb, const int32_t* __restrict__ sa,
                           const int32_t* __restrict__ scc, const int32_t* __restrict__ par,
                           const int32_t* __restrict__ pre, int32_t V, int32_t* __restrict__ low,
                           int32_t* __restrict__ high, int32_t* __restrict__ longl, int64_t* nlong) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t base = (int64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~(WAVE - 1)); base < V; base += stride) {
        const int64_t v = base + laneId();
        const bool act = v < V;
        const int64_t b = act ? sb[v] : 0, e = act ? sb[v + 1] : 0;
        const bool lng = act && e - b > BI_ROW;
        if (act && !lng) {
            int32_t lo = pre[v], hi = lo;
            for (int64_t j = b; j < e; ++j) {
                const int32_t w = sa[j];
                if (nonTree(scc, par, (int32_t)v, w)) {
                    const int32_t pw = pre[w];
                    lo = pw < lo ? pw : lo;
                    hi = pw > hi ? pw : hi;
                }
            }
            low[v] = lo;
            high[v] = hi;
        }
        const int64_t pos = wavePush(nlong, lng);
        if (lng) longl[pos] = (int32_t)v;
    }
}

// one block per long row: block-wide min / max, one store
__global__ void kBiLowHighLong(const int32_t* __restrict__ sb, const int32_t* __restrict__ sa,
                               const int32_t* __restrict__ scc, const int32_t* __restrict__ par,
                               const int32_t* __restrict__ pre, const int32_t* __restrict__ longl,
                               const int64_t* __restrict__ nlong, int32_t* __restrict__ low,
                               int32_t* __restrict__ high) {
    __shared__ int32_t smin[1024 / WAVE], smax[1024 / WAVE];
    const int64_t n = *nlong;
    for (int64_t r = blockIdx.x; r < n; r += gridDim.x) {
        const int32_t v = longl[r];
        int32_t lo = pre[v], hi = lo;
        for (int64_t j = sb[v] + threadIdx.x; j < sb[v + 1]; j += blockDim.x) {
            const int32_t w = sa[j];
            if (nonTree(scc, par, v, w)) {
                const int32_t pw = pre[w];
                lo = pw < lo ? pw : lo;
                hi = pw > hi ? pw : hi;
            }
        }
        for (int d = WAVE / 2; d > 0; d >>= 1) {
            const int32_t a = __shfl_xor(lo, d), c = __shfl_xor(hi, d);
            lo = a < lo ? a : lo;
            hi = c > hi ? c : hi;
        }
        if (laneId() == 0) {
            smin[threadIdx.x / WAVE] = lo;
            smax[threadIdx.x / WAVE] = hi;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            for (int k = 1; k < (int)(blockDim.x / WAVE); ++k) {
                lo = smin[k] < lo ? smin[k] : lo;
                hi = smax[k] > hi ? smax[k] : hi;
            }
            low[v] = lo;
            high[v] = hi;
        }
        __syncthreads();
    }
}

__global__ void kBiLowHighUp(const int32_t* __restrict__ order, const int32_t* __restrict__ lbeg, int32_t L,
                             const int32_t* __restrict__ par, int32_t* low, int32_t* high) {
    const int64_t lo = lbeg[L], hi = lbeg[L + 1];
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t base = lo + (int64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~(WAVE - 1)); base < hi; base += stride) {
        const int64_t i = base + laneId();
        const bool act = i < hi;
        const int32_t v = act ? order[i] : 0;
        const int32_t p = act ? par[v] : 0;
        waveCombine(act, p, act ? low[v] : 0, OpMin{}, low);
        waveCombine(act, p, act ? high[v] : 0, OpMax{}, high);
    }
}

__device__ __forceinline__ int32_t ufFind(int32_t* uf, int32_t x) {
    while (true) {
        const int32_t p = uf[x];
        if (p == x) return x;
        const int32_t g = uf[p];
        if (g != p) uf[x] = g;  // path halving (benign races: parents only move toward the root)
        x = g;
    }
}

// union by smaller root id
__device__ __forceinline__ void ufUnite(int32_t* uf, int32_t a, int32_t b) {
    while (true) {
        a = ufFind(uf, a);
        b = ufFind(uf, b);
        if (a == b) return;
        if (a > b) {
            const int32_t t = a;
            a = b;
            b = t;
        }
        if (atomicCAS(&uf[b], b, a) == b) return;
    }
}

__global__ void kBiUfCompress(int32_t V, int32_t* uf) {
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < V; v += (int64_t)gridDim.x * blockDim.x)
        uf[v] = ufFind(uf, (int32_t)v);
}

__global__ void kBiUfInit(int32_t V, int32_t* uf) {
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v < V) uf[v] = (int32_t)v;
}

// tree edge (v, p) joins (p, p(p)) when subtree(v) escapes subtree(p)
__global__ void kBiRuleTree(const int32_t* __restrict__ par, const int32_t* __restrict__ pre,
                            const int32_t* __restrict__ sz, const int32_t* __restrict__ low,
                            const int32_t* __restrict__ high, int32_t V, int32_t* uf) {
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < V; v += (int64_t)gridDim.x * blockDim.x) {
        const int32_t p = par[v];
        if (p < 0 || par[p] < 0) continue;
        if (low[v] < pre[p] || high[v] >= pre[p] + sz[p]) ufUnite(uf, (int32_t)v, p);
    }
}

// closure arc j of u: a non-tree edge {u, w} (taken from the end with the
// smaller preorder number) whose ends are unrelated joins their tree edges
struct BiCrossF {
    const int32_t* sa;
    const int32_t* scc;
    const int32_t* par;
    const int32_t* pre;
    const int32_t* sz;
    int32_t* uf;
    __device__ int64_t operator()(int32_t u, uint64_t, int32_t, int64_t j) const {
        const int32_t w = sa[j];
        if (scc[w] != scc[u] || w == par[u] || par[w] == u) return 0;
        const int32_t pu = pre[u], pw = pre[w];
        if (pu >= pw) return 0;
        if (pw < pu + sz[u]) return 0;  // w descends from u: a back edge
        if (uf[uf[u]] == uf[uf[w]]) return 0;  // joined already (classes are compressed, so mostly one hop)
        ufUnite(uf, u, w);
        return 0;
    }
};

__global__ void kBiCrossRows(BiCrossF f, const int32_t* __restrict__ sb, int32_t V, HeavyList H) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t base = (int64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~(WAVE - 1)); base < V; base += stride) {
        const int64_t v = base + laneId();
        const bool act = v < V;
        const int64_t b = act ? sb[v] : 0, e = act ? sb[v + 1] : 0;
        waveRowsH(act, (int32_t)v, 0, 0, b, e, H, f);
    }
}

// block of the tree edge (v, p(v)) = its class's smallest child id; count
// classes and single-vertex trees
__global__ void kBiBlocks(const int32_t* __restrict__ par, const int32_t* __restrict__ sz, int32_t* uf, int32_t V,
                          int32_t* __restrict__ blk, int64_t* cnt) {
    int64_t c = 0;
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < V; v += (int64_t)gridDim.x * blockDim.x) {
        if (par[v] < 0) {
            blk[v] = -1;
            c += sz[v] == 1;
        } else {
            const int32_t r = ufFind(uf, (int32_t)v);
            blk[v] = r;
            c += r == (int32_t)v;
        }
    }
    blockAtomicAdd(c, cnt);
}

// a vertex in two or more blocks: a child whose tree edge's block differs
// from the vertex's reference block (its own tree edge's, or at a root the
// block of its first child) marks it
__global__ void kBiApInit(const int32_t* __restrict__ blk, const int32_t* __restrict__ cbeg,
                          const int32_t* __restrict__ cadj, int32_t V, int32_t* __restrict__ ref,
                          uint8_t* __restrict__ ap) {
    for (int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; x < V; x += (int64_t)gridDim.x * blockDim.x) {
        int32_t r = blk[x];
        if (r < 0 && cbeg[x] < cbeg[x + 1]) r = blk[cadj[cbeg[x]]];
        ref[x] = r;
        ap[x] = 0;
    }
}

__global__ void kBiApMark(const int32_t* __restrict__ par, const int32_t* __restrict__ blk,
                          const int32_t* __restrict__ ref, int32_t V, uint8_t* ap) {
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < V; c += (int64_t)gridDim.x * blockDim.x) {
        const int32_t p = par[c];
        if (p >= 0 && blk[c] != ref[p]) ap[p] = 1;
    }
}

}  // namespace
}  // namespace gir

using namespace gir;

extern "C" int girhip_bicc(girhip_graph* h, int32_t* parent, int32_t* block, uint8_t* ap, int64_t* n_blocks) {
    return guarded([&] {
        useDevice(h);
        if (!h->scc_req) raise(GIRHIP_ESTATE, "bicc: call girhip_scc_dag first");
        if (h->pending && girhip_sync(h) != GIRHIP_OK) raise(GIRHIP_EHIP, girhip_last_error());
        const int32_t V = h->V;
        const int blk = (int)h->prm.block;
        const size_t Vs = (size_t)std::max<int32_t>(V, 1);
        Scratch& S = scratch(h);
        S.bi_lev.ensure(Vs);
        S.bi_par.ensure(Vs);
        S.bi_order.ensure(Vs);
        S.bi_lbeg.ensure(Vs + 2);
        S.bi_cadj.ensure(Vs);
        S.bi_cbeg.ensure(Vs + 1);
        S.bi_cpos.ensure(Vs);
        S.bi_sz.ensure(Vs);
        S.bi_pre.ensure(Vs);
        S.bi_low.ensure(Vs);
        S.bi_high.ensure(Vs);
        S.bi_uf.ensure(Vs);
        S.bi_blk.ensure(Vs);
        S.bi_ap.ensure(Vs);
        S.bi_tmp.ensure(Vs);
        S.bi_scan.ensure(Vs + 1);
        S.keys[0].ensure(Vs);
        S.bi_ctl.ensure(8);
        int64_t* ctl = S.bi_ctl.p;
        tBegin(h);
        const int64_t m_total = symCsr(h);
        const int32_t *sb = h->sym_beg.p, *sa = h->sym_adj.p;
        const int32_t* scc = h->scc.p;
        // 1. BFS forest from every canonical id
        ReachBufs R = reachBufs(h);
        fillU64(h, R.vis, 0, R.W);
        fillI32(h, S.bi_lev.p, -1, V);
        fillI64(h, ctl, 0, 8);
        launch(h, "bi_roots", kBiRoots, gridFor(V, blk * BR_ITEMS), blk, 0, scc, sb, V, R.q[0], ctl);
        int64_t c2[2];
        readCounters(h, ctl, 2, c2);
        runReach(h, sb, sa, sb, sa, R, c2[0], c2[1], (int64_t)V, m_total, SameScc{scc}, WriteLevel{S.bi_lev.p});
        fillI32(h, (int32_t*)(ctl + 2), 0, 1);
        launch(h, "bi_parents", kBiParents, boundedGrid(V, 256), 256, 0, sb, sa, scc, (const int32_t*)S.bi_lev.p, V,
               S.bi_par.p, (int32_t*)(ctl + 2));
        int64_t c3[3];
        readCounters(h, ctl, 3, c3);
        const int32_t maxL = (int32_t)(c3[2] & 0xffffffff);
        const int B = bitsFor((uint64_t)(V > 1 ? V - 1 : 1));
        const int LB = bitsFor((uint64_t)(maxL > 0 ? maxL : 1));
        // 2. vertices by level, children by parent
        launch(h, "bi_level_keys", kBiLevelKeys, gridFor(V, blk), blk, 0, (const int32_t*)S.bi_lev.p, V, B,
               S.keys[0].p);
        radixSort(h, S.keys[0].p, nullptr, V, LB + B);
        launch(h, "bi_levels", kBiGroups, gridFor(std::max<int64_t>(V, (int64_t)maxL + 2), blk), blk, 0,
               (const uint64_t*)S.keys[0].p, (int64_t)V, B, (int64_t)maxL + 1, S.bi_order.p, S.bi_lbeg.p,
               (int32_t*)nullptr);
        fillI64(h, ctl + 3, 0, 1);
        launch(h, "bi_child_keys", kBiChildKeys, boundedGrid(V, blk), blk, 0, (const int32_t*)S.bi_par.p, V, B,
               S.keys[0].p, ctl + 3);
        int64_t nc;
        readCounters(h, ctl + 3, 1, &nc);
        if (nc > 0) radixSort(h, S.keys[0].p, nullptr, nc, 2 * B);
        launch(h, "bi_children", kBiGroups, gridFor(std::max<int64_t>(nc, (int64_t)V + 1), blk), blk, 0,
               (const uint64_t*)S.keys[0].p, nc, B, (int64_t)V, S.bi_cadj.p, S.bi_cbeg.p, S.bi_cpos.p);
        // subtree sizes (bottom-up), preorder numbers (top-down)
        fillI32(h, S.bi_sz.p, 1, V);
        for (int32_t L = maxL; L >= 1; --L)
            launch(h, "bi_size_up", kBiSizeUp, 1024, 256, 0, (const int32_t*)S.bi_order.p,
                   (const int32_t*)S.bi_lbeg.p, L, (const int32_t*)S.bi_par.p, S.bi_sz.p);
        launch(h, "bi_root_sizes", kBiRootSizes, gridFor(V, blk), blk, 0, (const int32_t*)S.bi_par.p,
               (const int32_t*)S.bi_sz.p, V, S.bi_tmp.p);
        scanI32toI64(h, S.bi_tmp.p, S.bi_scan.p, V);
        launch(h, "bi_root_pre", kBiRootPre, gridFor(V, blk), blk, 0, (const int32_t*)S.bi_par.p,
               (const int64_t*)S.bi_scan.p, V, S.bi_pre.p);
        S.bi_cscan.ensure((size_t)nc + 1);
        if (nc > 0) {
            launch(h, "bi_child_sizes", kBiChildSizes, gridFor(nc, blk), blk, 0, (const int32_t*)S.bi_cadj.p, nc,
                   (const int32_t*)S.bi_sz.p, S.bi_tmp.p);
            scanI32toI64(h, S.bi_tmp.p, S.bi_cscan.p, nc);
        }
        for (int32_t L = 1; L <= maxL; ++L)
            launch(h, "bi_pre_down", kBiPreDown, 1024, 256, 0, (const int32_t*)S.bi_order.p,
                   (const int32_t*)S.bi_lbeg.p, L, (const int32_t*)S.bi_par.p, (const int32_t*)S.bi_cbeg.p,
                   (const int32_t*)S.bi_cpos.p, (const int64_t*)S.bi_cscan.p, S.bi_pre.p);
        // 3. low / high
        fillI64(h, ctl + 5, 0, 1);
        launch(h, "bi_lowhigh", kBiLowHigh, boundedGrid(V, blk), blk, 0, sb, sa, scc, (const int32_t*)S.bi_par.p,
               (const int32_t*)S.bi_pre.p, V, S.bi_low.p, S.bi_high.p, S.bi_tmp.p, ctl + 5);
        launch(h, "bi_lowhigh_long", kBiLowHighLong, 2048, 256, 0, sb, sa, scc, (const int32_t*)S.bi_par.p,
               (const int32_t*)S.bi_pre.p, (const int32_t*)S.bi_tmp.p, (const int64_t*)(ctl + 5), S.bi_low.p,
               S.bi_high.p);
        for (int32_t L = maxL; L >= 1; --L)
            launch(h, "bi_lowhigh_up", kBiLowHighUp, 1024, 256, 0, (const int32_t*)S.bi_order.p,
                   (const int32_t*)S.bi_lbeg.p, L, (const int32_t*)S.bi_par.p, S.bi_low.p, S.bi_high.p);
        // 4. union-find over tree edges
        launch(h, "bi_uf_init", kBiUfInit, gridFor(V, blk), blk, 0, V, S.bi_uf.p);
        launch(h, "bi_rule_tree", kBiRuleTree, boundedGrid(V, 256), 256, 0, (const int32_t*)S.bi_par.p,
               (const int32_t*)S.bi_pre.p, (const int32_t*)S.bi_sz.p, (const int32_t*)S.bi_low.p,
               (const int32_t*)S.bi_high.p, V, S.bi_uf.p);
        const BiCrossF cf{sa, scc, S.bi_par.p, S.bi_pre.p, S.bi_sz.p, S.bi_uf.p};
        launch(h, "bi_uf_compress", kBiUfCompress, boundedGrid(V, 256), 256, 0, V, S.bi_uf.p);
        launch(h, "bi_cross", kBiCrossRows, boundedGrid(V, blk), blk, 0, cf, sb, V, heavyList(h));
        launchHeavyStore(h, "bi_cross_rows", cf);
        // 5. blocks and articulation points
        fillI64(h, ctl + 4, 0, 1);
        launch(h, "bi_blocks", kBiBlocks, boundedGrid(V, 256), 256, 0, (const int32_t*)S.bi_par.p,
               (const int32_t*)S.bi_sz.p, S.bi_uf.p, V, S.bi_blk.p, ctl + 4);
        launch(h, "bi_ap_init", kBiApInit, boundedGrid(V, 256), 256, 0, (const int32_t*)S.bi_blk.p,
               (const int32_t*)S.bi_cbeg.p, (const int32_t*)S.bi_cadj.p, V, S.bi_tmp.p, S.bi_ap.p);
        launch(h, "bi_ap_mark", kBiApMark, boundedGrid(V, 256), 256, 0, (const int32_t*)S.bi_par.p,
               (const int32_t*)S.bi_blk.p, (const int32_t*)S.bi_tmp.p, V, S.bi_ap.p);
        int64_t nb;
        readCounters(h, ctl + 4, 1, &nb);
        tEnd(h, "bicc");
        if (parent) copyOut(h, parent, S.bi_par.p, sizeof(int32_t) * (size_t)V);
        if (block) copyOut(h, block, S.bi_blk.p, sizeof(int32_t) * (size_t)V);
        if (ap) copyOut(h, ap, S.bi_ap.p, (size_t)V);
        syncStream(h);
        if (n_blocks) *n_blocks = nb;
    });
}

// Certificate of the big SCC across update batches.
//
// The SCC of a pivot p is held together by two spanning trees of its
// members: an out-tree (gc_fpar: the arc gc_fpar[x] -> x, chains lead from p)
// and an in-tree (gc_bpar: the arc x -> gc_bpar[x], chains lead to p).  While
// both trees span the members, every member is reached from p and reaches p.
// After a batch:
//   - inserted arcs break nothing;
//   - a deleted arc that is in neither tree breaks nothing (the reference
//     answers most per-op deletes inside an SCC by a local search for a
//     bypass, gir.cpp:3062-3131; the trees make that answer free for every
//     arc off them);
//   - a deleted tree arc orphans the subtree under it.  Orphans, and the
//     vertices of the SCCs that may merge into the big one (labelled -1 by
//     the caller), are re-attached through any arc from a still-attached
//     vertex: one bottom-up step over the boundary, then a top-down sweep
//     restricted to the candidates, both recording the parents.
// A vertex stays in (or joins) the big SCC iff it is attached on both sides;
// every other orphan is labelled -1 for the caller's SCC pass.  With no valid
// trees the same sweeps run from p over all members and candidates, which
// builds the trees.
#include "reach.h"

namespace gir {
namespace {

__device__ __forceinline__ bool hasBit(const uint64_t* b, int32_t v) { return (b[v >> 6] >> (v & 63)) & 1; }

// pivot: the vertex of largest outdeg * indeg (packed maximum, smallest id on ties)
__global__ void kCertPivot(const int32_t* __restrict__ fb, const int32_t* __restrict__ bb, int32_t V,
                           unsigned long long* best) {
    uint64_t key = 0;
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < V; v += (int64_t)gridDim.x * blockDim.x) {
        unsigned long long s = (unsigned long long)(fb[v + 1] - fb[v]) * (unsigned long long)(bb[v + 1] - bb[v]);
        if (s > (1ULL << 33) - 1) s = (1ULL << 33) - 1;
        const uint64_t k = (s << 31) | (unsigned long long)(0x7fffffff - v);
        key = k > key ? k : key;
    }
    key = blockMaxU64(key);
    if (threadIdx.x == 0 && key) atomicMax(best, (unsigned long long)key);
}

// c[0] = the packed pivot key -> c[1] = the pivot, c[2] = its label
__global__ void kCertLabel(const int32_t* __restrict__ scc, int64_t* c) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        const int32_t p = (int32_t)(0x7fffffff - (c[0] & 0x7fffffffLL));
        c[1] = p;
        c[2] = scc[p];
    }
}

// Candidates of a check from the pivot: members and -1 vertices.
struct CertAll {
    const int32_t* scc;
    int32_t g;
    __device__ bool self(int32_t v) const {
        const int32_t s = scc[v];
        return s == g || s == -1;
    }
    __device__ bool operator()(int32_t c, int32_t) const { return self(c); }
};

// Candidates of a repair: orphaned members and -1 vertices.
struct CertOrphans {
    const int32_t* scc;
    int32_t g;
    const uint64_t* orph;
    __device__ bool self(int32_t v) const {
        const int32_t s = scc[v];
        return s == -1 || (s == g && hasBit(orph, v));
    }
    __device__ bool operator()(int32_t c, int32_t) const { return self(c); }
};

// Subtree walk: the tree children of p (members whose parent is p).
struct TreeChild {
    const int32_t* scc;
    int32_t g;
    const int32_t* par;
    __device__ bool self(int32_t v) const { return scc[v] == g; }
    __device__ bool operator()(int32_t c, int32_t p) const { return scc[c] == g && par[c] == p; }
};

// Intact members (member, not orphaned): one wave per 64-vertex word.
__global__ void kCertIntact(const int32_t* __restrict__ scc, int32_t V, int32_t g, const uint64_t* __restrict__ orph,
                            uint64_t* __restrict__ out) {
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t m = __ballot(v < V && scc[v] == g);
    if (laneId() == 0 && (v >> 6) << 6 < V) out[v >> 6] = m & ~orph[v >> 6];
}

// counter slots (GC_BUZ: the pair the boundary step clears for a next level)
enum : int { GC_DROP = 0, GC_JOIN, GC_MIN, GC_BU, GC_BUZ, GC_N = GC_BUZ + 2 };

// Final labels: in the big SCC iff attached on both sides.  Members that are
// not leave (-1), -1 vertices that are join (g); cnt[GC_MIN] = the complement
// of the smallest vertex of the new big SCC (its canonical label).
__global__ void kCertApply(int32_t* scc, int32_t V, int32_t g, int full, const uint64_t* __restrict__ orf,
                           const uint64_t* __restrict__ orb, const uint64_t* __restrict__ atf,
                           const uint64_t* __restrict__ atb, int64_t* cnt) {
    int64_t drop = 0, join = 0;
    uint64_t mn = 0;
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < V; v += (int64_t)gridDim.x * blockDim.x) {
        const int32_t s = scc[v];
        if (s != g && s != -1) continue;
        const bool mem = s == g;
        const bool okf = hasBit(atf, (int32_t)v) || (!full && mem && !hasBit(orf, (int32_t)v));
        const bool okb = hasBit(atb, (int32_t)v) || (!full && mem && !hasBit(orb, (int32_t)v));
        const bool in = okf && okb;
        if (mem && !in) {
            scc[v] = -1;
            ++drop;
        } else if (!mem && in) {
            scc[v] = g;
            ++join;
        }
        if (in) mn = ~(uint64_t)v > mn ? ~(uint64_t)v : mn;
    }
    blockAtomicAdd(drop, cnt + GC_DROP);
    blockAtomicAdd(join, cnt + GC_JOIN);
    mn = blockMaxU64(mn);
    if (threadIdx.x == 0 && mn) atomicMax((unsigned long long*)(cnt + GC_MIN), (unsigned long long)mn);
}

__global__ void kCertRelabel(int32_t* scc, int32_t V, int32_t from, int32_t to) {
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < V; v += (int64_t)gridDim.x * blockDim.x)
        if (scc[v] == from) scc[v] = to;
}

}  // namespace

void certTarget(girhip_graph* h) {
    const int32_t V = h->V;
    h->gc_ok = false;
    h->gc_pivot = h->gc_label = -1;
    if (V <= 0) return;
    Scratch& S = scratch(h);
    S.gc_cnt.ensure(GC_N);
    int64_t* c = S.gc_cnt.p;
    fillI64(h, c, 0, 1);
    launch(h, "cert_pivot", kCertPivot, reduceGrid(V, 256), 256, 0, (const int32_t*)h->fw_beg.p,
           (const int32_t*)h->bw_beg.p, V, (unsigned long long*)c);
    launch(h, "cert_label", kCertLabel, 1, 64, 0, (const int32_t*)h->scc.p, c);
    int64_t r[3];
    readCounters(h, c, 3, r);
    h->gc_pivot = (int32_t)r[1];
    h->gc_label = (int32_t)r[2];
}

// c[0] = scc[p]
__global__ void kCertLabelAt(const int32_t* __restrict__ scc, int32_t p, int64_t* c) {
    if (blockIdx.x == 0 && threadIdx.x == 0) c[0] = scc[p];
}

void reserveCert(girhip_graph* h) {
    Scratch& S = scratch(h);
    const int64_t W = ((int64_t)h->V + 63) / 64;
    for (auto& b : S.gc_bits) b.ensure((size_t)std::max<int64_t>(W, 1));
    S.gc_cnt.ensure(GC_N);
    h->gc_fpar.ensure((size_t)std::max<int32_t>(h->V, 1));
    h->gc_bpar.ensure((size_t)std::max<int32_t>(h->V, 1));
}

void certLabel(girhip_graph* h) {
    Scratch& S = scratch(h);
    S.gc_cnt.ensure(GC_N);
    launch(h, "cert_label", kCertLabelAt, 1, 64, 0, (const int32_t*)h->scc.p, h->gc_pivot, S.gc_cnt.p);
    int64_t g;
    readCounters(h, S.gc_cnt.p, 1, &g);
    h->gc_label = (int32_t)g;
}

CertStats certRepair(girhip_graph* h, int64_t ncutf, int64_t ncutb, int64_t ncand) {
    const int32_t V = h->V;
    const int blk = (int)h->prm.block;
    Scratch& S = scratch(h);
    const bool full = !h->gc_ok;
    const int32_t g = h->gc_label, p = h->gc_pivot;
    const int64_t W = ((int64_t)V + 63) / 64;
    h->gc_fpar.ensure((size_t)V);
    h->gc_bpar.ensure((size_t)V);
    for (auto& b : S.gc_bits) b.ensure((size_t)W);
    S.gc_cnt.ensure(GC_N);
    int64_t* c = S.gc_cnt.p;
    ReachBufs R = reachBufs(h);
    CertStats st;
    for (int side = 0; side < 2; ++side) {
        // push direction: out-arcs for the out-tree (reached from p), in-arcs
        // for the in-tree (reaching p)
        const int32_t* pb = side ? h->bw_beg.p : h->fw_beg.p;
        const int32_t* pa = side ? h->bw_adj.p : h->fw_adj.p;
        const int32_t* rb = side ? h->fw_beg.p : h->bw_beg.p;
        const int32_t* ra = side ? h->fw_adj.p : h->bw_adj.p;
        int32_t* par = side ? h->gc_bpar.p : h->gc_fpar.p;
        uint64_t* orph = S.gc_bits[side].p;
        uint64_t* att = S.gc_bits[2 + side].p;
        fillU64(h, R.vis, 0, W);
        if (full) {
            fillI32(h, par + p, p, 1);
            fillI32(h, R.q[0], p, 1);
            const ReachStats rs = runReach(h, pb, pa, rb, ra, R, 1, 0, (int64_t)V, h->E, CertAll{h->scc.p, g},
                                           WriteParent{par});
            st.swept += rs.reached;
        } else {
            const int64_t nc = side ? ncutb : ncutf;
            fillU64(h, orph, 0, W);
            if (nc > 0) {
                // orphans: the subtrees under the heads (tails) of deleted tree arcs
                HIPCHK(hipMemcpyAsync(R.q[0], S.gc_cut[side].p, sizeof(int32_t) * nc, hipMemcpyDeviceToDevice, h->st));
                const ReachStats rs = runReach(h, pb, pa, rb, ra, R, nc, 0, nc, h->E, TreeChild{h->scc.p, g, par},
                                               NoVisit{}, false);
                st.orphans += rs.reached;
                HIPCHK(hipMemcpyAsync(orph, R.vis, sizeof(uint64_t) * W, hipMemcpyDeviceToDevice, h->st));
                fillU64(h, R.vis, 0, W);
            }
            if (nc > 0 || ncand > 0) {
                // boundary: candidates with an arc from an intact member (one
                // bottom-up step), then a sweep through the candidates
                launch(h, "cert_intact", kCertIntact, gridFor(W * 64, 256), 256, 0, (const int32_t*)h->scc.p, V, g,
                       (const uint64_t*)orph, R.fbits);
                fillI64(h, c + GC_BU, 0, 2);
                const CertOrphans f{h->scc.p, g, orph};
                launch(h, "cert_boundary", kBottomUp<CertOrphans, WriteParent>, buGrid(V, blk), blk, 0, rb, ra, pb,
                       R.vis, (const uint64_t*)R.fbits, R.nbits, V, 0, R.q[0], c + GC_BU, c + GC_BUZ, f,
                       WriteParent{par}, (const int64_t*)nullptr, (int64_t)0, (int64_t)0);
                int64_t r;
                readCounters(h, c + GC_BU, 1, &r);
                const int64_t n0 = (int64_t)((uint64_t)r >> 32), m0 = (int64_t)((uint64_t)r & 0xffffffffULL);
                if (n0 > 0) {
                    const ReachStats rs =
                        runReach(h, pb, pa, rb, ra, R, n0, m0, st.orphans + ncand, h->E, f, WriteParent{par});
                    st.swept += rs.reached;
                }
            }
        }
        HIPCHK(hipMemcpyAsync(att, R.vis, sizeof(uint64_t) * W, hipMemcpyDeviceToDevice, h->st));
    }
    fillI64(h, c + GC_DROP, 0, 3);
    launch(h, "cert_apply", kCertApply, reduceGrid(V, blk), blk, 0, h->scc.p, V, g, full ? 1 : 0,
           (const uint64_t*)S.gc_bits[0].p, (const uint64_t*)S.gc_bits[1].p, (const uint64_t*)S.gc_bits[2].p,
           (const uint64_t*)S.gc_bits[3].p, c);
    int64_t r[3];
    readCounters(h, c + GC_DROP, 3, r);
    st.dropped = r[0];
    st.joined = r[1];
    // the canonical label of the big SCC is its smallest vertex
    const int32_t lab = r[2] ? (int32_t)~(uint64_t)r[2] : -1;
    if (lab >= 0 && lab != g) {
        launch(h, "cert_relabel", kCertRelabel, boundedGrid(V, blk), blk, 0, h->scc.p, V, g, lab);
        h->gc_label = lab;
    }
    h->gc_ok = lab >= 0;
    return st;
}

}  // namespace gir

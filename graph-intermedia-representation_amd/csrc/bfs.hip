// BFS queries: replaces GIR::bfs (gir_query.cpp:108-294) for k sources.
//
// Each source runs the direction-optimising engine of reach.h over the
// forward CSR (push) and backward CSR (pull) with the reference's switch
// rule; levels are written straight into the caller's k x V array when it
// is device memory.
//
// HBM traffic per source (algorithmic): 4V (level init + writes) + 8 n_s
// (row offsets of reached vertices) + 8 m_s (adjacency + neighbour state),
// n_s = reached vertices, m_s = their out-arc count.
#include <vector>

#include "common.h"
#include "reach.h"

namespace gir {

ReachBufs reachBufs(girhip_graph* h) {
    Scratch& S = scratch(h);
    const size_t V = (size_t)(h->V > 0 ? h->V : 1);
    const int64_t W = ((int64_t)V + 63) / 64;
    S.q[0].ensure(V);
    S.q[1].ensure(V);
    for (int i = 0; i < 3; ++i) S.bits[i].ensure((size_t)W);
    ReachBufs B;
    B.q[0] = S.q[0].p;
    B.q[1] = S.q[1].p;
    B.vis = S.bits[0].p;
    B.fbits = S.bits[1].p;
    B.nbits = S.bits[2].p;
    B.W = W;
    return B;
}

// m_s per row: out-arc count summed over reached vertices (level >= 0).
__global__ void kReachedArcs(const int32_t* __restrict__ level, const int32_t* __restrict__ beg, int32_t V,
                             int64_t* out) {
    int64_t m = 0;
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < V; v += (int64_t)gridDim.x * blockDim.x)
        if (level[v] >= 0) m += beg[v + 1] - beg[v];
    blockAtomicAdd(m, out);
}

std::vector<int32_t> hostSources(girhip_graph* h, const int32_t* src, int32_t k) {
    std::vector<int32_t> s(k);
    if (k > 0) {
        HIPCHK(hipMemcpyAsync(s.data(), src, sizeof(int32_t) * k, hipMemcpyDefault, h->st));
        syncStream(h);
    }
    return s;
}

int64_t outDegree(girhip_graph* h, int32_t v) {
    HIPCHK(hipMemcpyAsync(h->hcnt, h->fw_beg.p + v, sizeof(int32_t) * 2, hipMemcpyDeviceToHost, h->st));
    syncStream(h);
    const int32_t* p = (const int32_t*)h->hcnt;
    return (int64_t)p[1] - p[0];
}

// ---------------------------------------------------------------------------
// Bit-parallel multi-source BFS (MS-BFS, Then et al., VLDB 2015): up to 64
// sources share one traversal.  Bit s of a vertex's seen / frontier / next
// word belongs to source s, so an arc is read once per level for all sources
// whose frontier holds its tail.  Push levels OR the tail's frontier word into
// the head's next word (atomicOr, skipped when nothing new); pull levels let
// every vertex that some source has not reached yet OR the frontier words of
// its in-neighbours, stopping once every source has reached it.  The same
// alpha/beta rule as the single-source engine picks the direction on the
// union frontier.  Levels are written once per (source, vertex): at the
// level a bit first appears, or -1 in a final pass over the unseen bits.

__global__ void kMsSeed(const int32_t* __restrict__ src, int k, int32_t V, uint64_t* seen, uint64_t* front,
                        int32_t* __restrict__ levels, int64_t* cnt) {
    const int s = threadIdx.x;
    if (s >= k) return;
    const int32_t v = src[s];
    if (v < 0 || v >= V) return;
    atomicOr((unsigned long long*)&seen[v], 1ULL << s);
    atomicOr((unsigned long long*)&front[v], 1ULL << s);
    levels[(size_t)s * V + v] = 0;
}

struct MsPushF {
    const int32_t* fa;
    const uint64_t* seen;
    uint64_t* next;
    __device__ int64_t operator()(int32_t, uint64_t fp, int32_t, int64_t j) const {
        const int32_t w = fa[j];
        const uint64_t m = fp & ~(seen[w] | ldAcq(&next[w]));
        if (m) atomicOr((unsigned long long*)&next[w], (unsigned long long)m);
        return 0;
    }
};

__global__ void kMsPush(const int32_t* __restrict__ fb, const uint64_t* __restrict__ front, int32_t V, MsPushF f,
                        int64_t* cz, HeavyList H) {
    if (blockIdx.x == 0 && threadIdx.x < 3) cz[threadIdx.x] = 0;
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t fr = v < V ? front[v] : 0;
    const bool act = fr != 0;
    const int32_t b = act ? fb[v] : 0, e = act ? fb[v + 1] : 0;
    waveRowsH(act, (int32_t)v, fr, 0, b, e, H, f);
}

__global__ void kMsPull(const int32_t* __restrict__ bb, const int32_t* __restrict__ ba,
                        const uint64_t* __restrict__ front, const uint64_t* __restrict__ seen,
                        uint64_t* __restrict__ next, int32_t V, uint64_t all, int64_t* cz) {
    if (blockIdx.x == 0 && threadIdx.x < 3) cz[threadIdx.x] = 0;
    const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= V) return;
    const uint64_t sw = seen[w];
    uint64_t acc = 0;
    if (sw != all) {
        const int32_t b = bb[w], e = bb[w + 1];
        for (int32_t j = b; j < e; ++j) {
            acc |= front[ba[j]];
            if ((acc | sw) == all) break;
        }
    }
    next[w] = acc & ~sw;
}

// New bits -> seen, frontier, level writes; cnt[0] = union frontier size,
// cnt[1] = its out-arc count.
__global__ void kMsFinish(uint64_t* seen, uint64_t* front, uint64_t* next, int32_t* __restrict__ levels, int32_t V,
                          int32_t L, const int32_t* __restrict__ fb, int64_t* cnt) {
    const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t nb = 0;
    if (w < V) {
        const uint64_t sw = seen[w];
        nb = next[w] & ~sw;
        seen[w] = sw | nb;
        front[w] = nb;
        next[w] = 0;
    }
    if (__ballot(nb != 0)) {
        for (int s = 0; s < 64; ++s)
            if ((nb >> s) & 1) levels[(size_t)s * V + w] = L + 1;
    }
    blockAtomicAdd(nb != 0, &cnt[0]);
    blockAtomicAdd(nb != 0 ? (int64_t)(fb[w + 1] - fb[w]) : 0, &cnt[1]);
}

__global__ void kMsUnreached(const uint64_t* __restrict__ seen, int32_t* __restrict__ levels, int32_t V,
                             uint64_t all) {
    const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t un = w < V ? (~seen[w] & all) : 0;
    if (__ballot(un != 0)) {
        for (int s = 0; s < 64; ++s)
            if ((un >> s) & 1) levels[(size_t)s * V + w] = -1;
    }
}

// Up to 64 sources (device array src, k entries) into levels rows 0..k-1.
void msbfs(girhip_graph* h, const int32_t* dsrc, int k, int32_t* levels, int64_t m0) {
    const int32_t V = h->V;
    const int blk = (int)h->prm.block;
    Scratch& S = scratch(h);
    for (int i = 0; i < 3; ++i) S.ms[i].ensure((size_t)(V > 0 ? V : 1));
    uint64_t *seen = S.ms[0].p, *front = S.ms[1].p, *next = S.ms[2].p;
    const uint64_t all = k >= 64 ? ~0ULL : ((1ULL << k) - 1);
    for (int i = 0; i < 3; ++i) fillU64(h, S.ms[i].p, 0, V);
    int64_t* c = h->dcnt.p + 16;  // counter pairs for alternate levels (3 slots each)
    fillI64(h, c, 0, 6);
    launch(h, "msbfs_seed", kMsSeed, 1, 64, 0, dsrc, k, V, seen, front, levels, c);
    const int64_t alpha = (int64_t)h->prm.bfs_alpha, beta = (int64_t)h->prm.bfs_beta;
    const int64_t min_reach = (int64_t)h->prm.bfs_min_reach;
    int64_t nf = k, mf = m0, m_unexplored = h->E - m0;
    bool pull = false;
    const unsigned g = gridFor(V, blk);
    for (int32_t L = 0; nf > 0; ++L) {
        if ((int64_t)V >= min_reach) {
            if (!pull && mf * alpha > m_unexplored) pull = true;
            else if (pull && nf * beta < (int64_t)V) pull = false;
        }
        int64_t* cn = c + 3 * (L & 1);
        int64_t* cz = c + 3 * ((L + 1) & 1);
        if (pull)
            launch(h, "msbfs_pull", kMsPull, g, blk, 0, (const int32_t*)h->bw_beg.p, (const int32_t*)h->bw_adj.p,
                   (const uint64_t*)front, (const uint64_t*)seen, next, V, all, cz);
        else {
            const MsPushF f{h->fw_adj.p, seen, next};
            launch(h, "msbfs_push", kMsPush, g, blk, 0, (const int32_t*)h->fw_beg.p, (const uint64_t*)front, V, f, cz,
                   heavyList(h));
            launchHeavy(h, "msbfs_push_heavy", f);
        }
        launch(h, "msbfs_finish", kMsFinish, g, blk, 0, seen, front, next, levels, V, L, (const int32_t*)h->fw_beg.p,
               cn);
        int64_t r[2];
        readCounters(h, cn, 2, r);
        nf = r[0];
        mf = r[1];
        m_unexplored -= mf;
    }
    launch(h, "msbfs_unreached", kMsUnreached, g, blk, 0, (const uint64_t*)seen, levels, V, all);
}

}  // namespace gir

using namespace gir;

extern "C" int girhip_bfs(girhip_graph* h, const int32_t* src, int32_t k, int32_t* levels, int64_t* edges_out) {
    return guarded([&] {
        useDevice(h);
        if (k < 0 || (k > 0 && !src)) raise(GIRHIP_EINVAL, "bfs: bad source list");
        const int32_t V = h->V;
        std::vector<int32_t> s = hostSources(h, src, k);
        const size_t n = (size_t)k * (size_t)V;
        int32_t* out = levels;
        const bool direct = levels && isDevicePtr(levels);
        Scratch& S = scratch(h);
        if (!direct) {
            S.stage.ensure(sizeof(int32_t) * (n > 0 ? n : 1));
            out = (int32_t*)S.stage.p;
        }
        ReachBufs B = reachBufs(h);
        const bool bitpar = k > 1 && h->prm.msbfs != 0;
        DevBuf<int32_t> dsrc;
        if (bitpar) {
            dsrc.ensure((size_t)k);
            copyIn(h, dsrc.p, s.data(), sizeof(int32_t) * (size_t)k);
            syncStream(h);
        }
        tBegin(h);
        if (bitpar) {
            for (int32_t i0 = 0; i0 < k; i0 += 64) {
                const int c = std::min<int32_t>(64, k - i0);
                msbfs(h, dsrc.p + i0, c, out + (size_t)i0 * V, 0);
            }
        } else {
            fillI32(h, out, -1, (int64_t)n);
            for (int32_t i = 0; i < k; ++i) {
                const int32_t x = s[i];
                if (x < 0 || x >= V) continue;
                fillU64(h, B.vis, 0, B.W);
                fillI32(h, B.q[0], x, 1);
                const int64_t m0 = outDegree(h, x);
                runReach(h, h->fw_beg.p, h->fw_adj.p, h->bw_beg.p, h->bw_adj.p, B, 1, m0, (int64_t)V, h->E,
                         AllArcs{}, WriteLevel{out + (size_t)i * V});
            }
        }
        tEnd(h, "bfs");
        if (edges_out) {
            // 32 counter slots per readback
            std::vector<int64_t> e(k);
            for (int32_t i0 = 0; i0 < k; i0 += 32) {
                const int c = std::min<int32_t>(32, k - i0);
                fillI64(h, h->dcnt.p + 32, 0, 32);
                for (int32_t i = i0; i < i0 + c; ++i)
                    launch(h, "reached_arcs", kReachedArcs, boundedGrid(V, 256), 256, 0,
                           (const int32_t*)(out + (size_t)i * V), (const int32_t*)h->fw_beg.p, V,
                           h->dcnt.p + 32 + (i - i0));
                readCounters(h, h->dcnt.p + 32, c, e.data() + i0);
            }
            HIPCHK(hipMemcpy(edges_out, e.data(), sizeof(int64_t) * (size_t)k, hipMemcpyDefault));
        }
        if (levels && !direct) copyOut(h, levels, out, sizeof(int32_t) * n);
        syncStream(h);
    });
}

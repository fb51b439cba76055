// Single-source shortest paths: replaces GIR::sssp / GIR::deltaStepping
// (gir_query.cpp:629-703) for k sources.
//
// Delta-stepping with device-resident buckets.  All sources of a call (up to
// SB per batch) advance together, one grid row (blockIdx.y) per source.  Per
// source:
//   - the current bucket is [lo, lo + delta); its work list is a queue;
//   - a relaxation that improves a distance sets the head's bit in the near
//     bitmap (lands inside the bucket) or in the far bitmap (beyond it), so
//     duplicates collapse and no queue counter is hit per arc;
//   - after each round a compaction pass turns the near bitmap into the next
//     queue (block-wide scan, one atomic per block) - the reference
//     re-processes bucket k until nothing new lands in it;
//   - when the queue is empty, a pass over the far bitmap finds the smallest
//     distance >= hi, which picks the next non-empty bucket, and a split pass
//     moves the far entries inside it into the queue (entries below the new
//     bucket are settled and dropped).
// Relaxation is an atomicMin on 64-bit distances.  Weights: one uint32 per
// forward slot, or the reference's hash GIR::edgeWeight.  Rows longer than
// HEAVY_T arcs are walked by the whole grid (kHeavyRows).
//
// HBM traffic per source (algorithmic): 8V (init + final distances) + 8 n_s
// (row offsets) + 16 m_s (adjacency 4 + weight 4 + distance 8 per arc).
#include <vector>

#include "common.h"

namespace gir {

std::vector<int32_t> hostSources(girhip_graph* h, const int32_t* src, int32_t k);

constexpr uint64_t INF = ~0ULL;
constexpr int SB = 32;  // sources per batch

// Per-source control block (int64 slots, SB each).
enum : int {
    C_NEAR0 = 0,      // queue length, parity 0
    C_NEAR1 = SB,     // queue length, parity 1
    C_ADV = 2 * SB,   // 1: advances its bucket this round
    C_LO = 3 * SB,    // bucket [lo, hi) (uint64)
    C_HI = 4 * SB,
    C_FMIN = 5 * SB,  // smallest far distance >= hi (uint64)
    C_HQ = 6 * SB,    // heavy-phase queue length
    C_SLOTS = 7 * SB
};

struct SsspBatch {
    int32_t* qn[2];      // [SB][V] queues by parity
    int32_t* qh;         // [SB][V] vertices whose heavy arcs are due
    uint64_t* nearbits;  // [SB][W]
    uint64_t* farbits;   // [SB][W]
    uint64_t* donebits;  // [SB][W] processed in the current bucket
    int64_t* ctl;        // C_SLOTS
    uint64_t* dist;      // [SB][V] output rows
    int64_t V, W;
};

__global__ void kBInit(SsspBatch B, const int32_t* __restrict__ src, int k, uint64_t delta) {
    const int s = threadIdx.x;
    if (s >= SB) return;
    int64_t* c = B.ctl;
    for (int j = 0; j < C_SLOTS / SB; ++j) c[j * SB + s] = 0;
    c[C_HI + s] = (int64_t)delta;
    if (s < k) {
        const int32_t v = src[s];
        if (v >= 0 && v < B.V) {
            B.dist[(size_t)s * B.V + v] = 0;
            B.qn[0][(size_t)s * B.V] = v;
            c[C_NEAR0 + s] = 1;
        }
    }
}

// Relaxation of arc slot j of row p (owner distance dp) for source s.
// heavy = 0: only light arcs (w <= delta, the bucket's inner rounds);
// heavy = 1: only heavy arcs (w > delta, once per vertex and bucket).
struct RelaxF {
    SsspBatch B;
    const int32_t* fa;
    const uint32_t* w;
    uint64_t delta;
    int heavy;
    __device__ int64_t operator()(int32_t p, uint64_t dp, int32_t s, int64_t j) const {
        uint64_t* d = B.dist + (size_t)s * B.V;
        uint64_t wt;
        int32_t x;
        if (w) {
            wt = w[j];
            if ((wt > delta) != (heavy != 0)) return 0;
            x = fa[j];
        } else {
            x = fa[j];
            wt = hashWeight(p, x);
            if ((wt > delta) != (heavy != 0)) return 0;
        }
        const uint64_t nd = dp + wt;
        if (nd < ldAcq(&d[x])) {
            const uint64_t old = atomicMin((unsigned long long*)&d[x], (unsigned long long)nd);
            if (nd < old) {
                uint64_t* bits = (nd < (uint64_t)B.ctl[C_HI + s]) ? B.nearbits : B.farbits;
                const uint64_t bit = 1ULL << (x & 63);
                uint64_t* wp = &bits[(size_t)s * B.W + (x >> 6)];
                if (!(ldAcq(wp) & bit)) atomicOr((unsigned long long*)wp, (unsigned long long)bit);
            }
        }
        return 0;
    }
};

// Relax the queue of parity par; block (0, s) also resets the per-round
// slots of source s that this kernel does not read.
__global__ void kBRelax(RelaxF f, int par, const int32_t* __restrict__ fb, HeavyList H) {
    const int s = blockIdx.y;
    const SsspBatch& B = f.B;
    int64_t* c = B.ctl;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        c[(par ? C_NEAR0 : C_NEAR1) + s] = 0;
        c[C_FMIN + s] = (int64_t)INF;
        c[C_ADV + s] = 0;
        c[C_HQ + s] = 0;
    }
    const int64_t nin = c[(par ? C_NEAR1 : C_NEAR0) + s];
    if ((int64_t)blockIdx.x * blockDim.x >= nin) return;  // whole block idle for this source
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool act = i < nin;
    const int32_t u = act ? B.qn[par][(size_t)s * B.V + i] : 0;
    const uint64_t du = act ? ldAcq(&B.dist[(size_t)s * B.V + u]) : 0;
    if (act) {  // its heavy arcs are due when the bucket drains
        uint64_t* wp = &B.donebits[(size_t)s * B.W + (u >> 6)];
        const uint64_t bit = 1ULL << (u & 63);
        if (!(ldAcq(wp) & bit)) atomicOr((unsigned long long*)wp, (unsigned long long)bit);
    }
    const int32_t b = act ? fb[u] : 0, e = act ? fb[u + 1] : 0;
    waveRowsH(act, u, du, s, b, e, H, f);
}

// Sources whose bucket drained (next queue empty): the vertices processed
// in it -> heavy-phase queue (bits cleared).
__global__ void kBCompactDone(SsspBatch B, int par) {
    const int s = blockIdx.y;
    if (B.ctl[(par ? C_NEAR0 : C_NEAR1) + s] != 0) return;
    const int64_t wi = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t m = 0;
    if (wi < B.W) {
        uint64_t* wp = B.donebits + (size_t)s * B.W + wi;
        m = *wp;
        if (m) *wp = 0;
    }
    int64_t pos = blockReserve(__popcll(m), B.ctl + C_HQ + s);
    int32_t* q = B.qh + (size_t)s * B.V;
    while (m) {
        const int b = __ffsll((long long)m) - 1;
        m &= m - 1;
        q[pos++] = (int32_t)(wi * 64 + b);
    }
}

// Heavy arcs of the heavy-phase queue (grid stride, uniform per wave).
__global__ void kBRelaxHeavy(RelaxF f, const int32_t* __restrict__ fb, HeavyList H) {
    const int s = blockIdx.y;
    const SsspBatch& B = f.B;
    const int64_t n = B.ctl[C_HQ + s];
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t base = (int64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~(WAVE - 1)); base < n; base += stride) {
        const int64_t i = base + laneId();
        const bool act = i < n;
        const int32_t u = act ? B.qh[(size_t)s * B.V + i] : 0;
        const uint64_t du = act ? ldAcq(&B.dist[(size_t)s * B.V + u]) : 0;
        const int32_t b = act ? fb[u] : 0, e = act ? fb[u + 1] : 0;
        waveRowsH(act, u, du, s, b, e, H, f);
    }
}

// Near bitmap of every source -> its next queue (bits cleared).
__global__ void kBCompact(SsspBatch B, int par) {
    const int s = blockIdx.y;
    const int64_t wi = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t* wp = B.nearbits + (size_t)s * B.W + wi;
    uint64_t m = 0;
    if (wi < B.W) {
        m = *wp;
        if (m) *wp = 0;
    }
    int64_t pos = blockReserve(__popcll(m), B.ctl + (par ? C_NEAR0 : C_NEAR1) + s);
    int32_t* q = B.qn[par ^ 1] + (size_t)s * B.V;
    while (m) {
        const int b = __ffsll((long long)m) - 1;
        m &= m - 1;
        q[pos++] = (int32_t)(wi * 64 + b);
    }
}

// Sources whose next queue is empty: smallest far distance >= hi.
__global__ void kBFarMin(SsspBatch B, int par) {
    const int s = blockIdx.y;
    const int64_t* c = B.ctl;
    if (c[(par ? C_NEAR0 : C_NEAR1) + s] != 0) return;
    const uint64_t hi = (uint64_t)c[C_HI + s];
    const uint64_t* d = B.dist + (size_t)s * B.V;
    unsigned long long mn = INF;
    for (int64_t wi = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; wi < B.W; wi += (int64_t)gridDim.x * blockDim.x) {
        uint64_t m = B.farbits[(size_t)s * B.W + wi];
        while (m) {
            const int b = __ffsll((long long)m) - 1;
            m &= m - 1;
            const uint64_t x = d[wi * 64 + b];
            if (x >= hi && x < mn) mn = x;
        }
    }
    for (int o = WAVE / 2; o > 0; o >>= 1) {
        const unsigned long long t = __shfl_xor(mn, o);
        mn = t < mn ? t : mn;
    }
    if (laneId() == 0 && mn != INF) atomicMin((unsigned long long*)&B.ctl[C_FMIN + s], mn);
}

__global__ void kBAdvance(SsspBatch B, int par, int k, uint64_t delta) {
    const int s = threadIdx.x;
    if (s >= k) return;
    int64_t* c = B.ctl;
    if (c[(par ? C_NEAR0 : C_NEAR1) + s] != 0) return;
    const uint64_t m = (uint64_t)c[C_FMIN + s];
    if (m == INF) return;  // no far work left: this source is done
    const uint64_t lo = (m / delta) * delta;
    c[C_LO + s] = (int64_t)lo;
    c[C_HI + s] = (int64_t)(lo > INF - delta ? INF : lo + delta);
    c[C_ADV + s] = 1;
}

// Far entries of advancing sources: inside [lo, hi) -> next queue, below lo
// -> settled (dropped), both leave the far bitmap; the rest stay.
__global__ void kBSplit(SsspBatch B, int par) {
    const int s = blockIdx.y;
    const int64_t* c = B.ctl;
    if (!c[C_ADV + s]) return;
    const uint64_t lo = (uint64_t)c[C_LO + s], hi = (uint64_t)c[C_HI + s];
    const uint64_t* d = B.dist + (size_t)s * B.V;
    const int64_t wi = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t take = 0;
    if (wi < B.W) {
        uint64_t* wp = B.farbits + (size_t)s * B.W + wi;
        const uint64_t m = *wp;
        uint64_t keep = 0, t = m;
        while (t) {
            const int b = __ffsll((long long)t) - 1;
            t &= t - 1;
            const uint64_t x = d[wi * 64 + b];
            if (x >= hi) keep |= 1ULL << b;
            else if (x >= lo) take |= 1ULL << b;
        }
        if (keep != m) *wp = keep;
    }
    int64_t pos = blockReserve(__popcll(take), B.ctl + (par ? C_NEAR0 : C_NEAR1) + s);
    int32_t* q = B.qn[par ^ 1] + (size_t)s * B.V;
    while (take) {
        const int b = __ffsll((long long)take) - 1;
        take &= take - 1;
        q[pos++] = (int32_t)(wi * 64 + b);
    }
}

__global__ void kReachedArcsD(const uint64_t* __restrict__ dist, const int32_t* __restrict__ beg, int32_t V,
                              int64_t* out) {
    int64_t m = 0;
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < V; v += (int64_t)gridDim.x * blockDim.x)
        if (dist[v] != INF) m += beg[v + 1] - beg[v];
    blockAtomicAdd(m, out);
}

struct SsspBufs {
    DevBuf<int32_t> qn[2], qh;
    DevBuf<uint64_t> nearbits, farbits, donebits;
    DevBuf<int64_t> ctl;
};

// k <= SB sources (device array dsrc) into rows of dist.
static void ssspBatch(girhip_graph* h, const int32_t* dsrc, int k, uint64_t delta, uint64_t* dist, SsspBufs& Bf) {
    const int32_t V = h->V;
    const int blk = (int)h->prm.block;
    const uint32_t* w = h->has_w ? h->w_fw.p : nullptr;
    SsspBatch B;
    B.V = V;
    B.W = ((int64_t)V + 63) / 64;
    for (int p = 0; p < 2; ++p) B.qn[p] = Bf.qn[p].p;
    B.qh = Bf.qh.p;
    B.nearbits = Bf.nearbits.p;
    B.farbits = Bf.farbits.p;
    B.donebits = Bf.donebits.p;
    B.ctl = Bf.ctl.p;
    B.dist = dist;
    fillU64(h, dist, INF, (int64_t)k * V);
    fillU64(h, B.nearbits, 0, (int64_t)k * B.W);
    fillU64(h, B.farbits, 0, (int64_t)k * B.W);
    fillU64(h, B.donebits, 0, (int64_t)k * B.W);
    launch(h, "sssp_init", kBInit, 1, 64, 0, B, dsrc, k, delta);
    const RelaxF fl{B, h->fw_adj.p, w, delta, 0}, fh{B, h->fw_adj.p, w, delta, 1};
    const unsigned wgrid = gridFor(B.W, blk);
    const unsigned fgrid = std::min<unsigned>(wgrid, 256);
    const unsigned hgrid = std::min<unsigned>(gridFor(V, blk), 1024);
    int par = 0;
    std::vector<int64_t> r(SB);
    int64_t maxn = 1;
    while (true) {
        launch(h, "sssp_relax", kBRelax, dim3(gridFor(maxn, blk), k), blk, 0, fl, par, (const int32_t*)h->fw_beg.p,
               heavyList(h));
        launchHeavy(h, "sssp_relax_rows", fl);
        launch(h, "sssp_compact", kBCompact, dim3(wgrid, k), blk, 0, B, par);
        launch(h, "sssp_compact_done", kBCompactDone, dim3(wgrid, k), blk, 0, B, par);
        launch(h, "sssp_relax_heavy", kBRelaxHeavy, dim3(hgrid, k), blk, 0, fh, (const int32_t*)h->fw_beg.p,
               heavyList(h));
        launchHeavy(h, "sssp_relax_heavy_rows", fh);
        launch(h, "sssp_farmin", kBFarMin, dim3(fgrid, k), blk, 0, B, par);
        launch(h, "sssp_advance", kBAdvance, 1, 64, 0, B, par, k, delta);
        launch(h, "sssp_split", kBSplit, dim3(wgrid, k), blk, 0, B, par);
        par ^= 1;
        readCounters(h, Bf.ctl.p + (par ? C_NEAR1 : C_NEAR0), SB, r.data());
        maxn = 0;
        for (int s = 0; s < k; ++s) maxn = std::max(maxn, r[s]);
        if (maxn == 0) break;  // every queue empty after the advance: all sources settled
    }
}

}  // namespace gir

using namespace gir;

extern "C" int girhip_sssp(girhip_graph* h, const int32_t* src, int32_t k, uint64_t delta, uint64_t* dist,
                           int64_t* edges_out) {
    return guarded([&] {
        useDevice(h);
        if (k < 0 || (k > 0 && !src)) raise(GIRHIP_EINVAL, "sssp: bad source list");
        if (delta == 0) delta = (uint64_t)h->prm.sssp_delta;
        if (delta == 0) raise(GIRHIP_EINVAL, "sssp: delta must be > 0");
        const int32_t V = h->V;
        std::vector<int32_t> s = hostSources(h, src, k);
        const size_t n = (size_t)k * (size_t)V;
        uint64_t* out = dist;
        const bool direct = dist && isDevicePtr(dist);
        Scratch& S = scratch(h);
        if (!direct) {
            S.stage.ensure(sizeof(uint64_t) * (n > 0 ? n : 1));
            out = (uint64_t*)S.stage.p;
        }
        const size_t Vs = (size_t)(V > 0 ? V : 1), Ws = (Vs + 63) / 64;
        const int kc = std::max(1, std::min<int32_t>(k, SB));
        SsspBufs Bf;
        for (int p = 0; p < 2; ++p) Bf.qn[p].ensure((size_t)kc * Vs);
        Bf.qh.ensure((size_t)kc * Vs);
        Bf.nearbits.ensure((size_t)kc * Ws);
        Bf.farbits.ensure((size_t)kc * Ws);
        Bf.donebits.ensure((size_t)kc * Ws);
        Bf.ctl.ensure(C_SLOTS);
        DevBuf<int32_t> dsrc;
        dsrc.ensure((size_t)(k > 0 ? k : 1));
        copyIn(h, dsrc.p, s.data(), sizeof(int32_t) * (size_t)k);
        tBegin(h);
        for (int32_t i0 = 0; i0 < k; i0 += SB) {
            const int c = std::min<int32_t>(SB, k - i0);
            ssspBatch(h, dsrc.p + i0, c, delta, out + (size_t)i0 * V, Bf);
        }
        tEnd(h, "sssp");
        if (edges_out) {
            std::vector<int64_t> e(k);
            for (int32_t i0 = 0; i0 < k; i0 += 32) {
                const int c = std::min<int32_t>(32, k - i0);
                fillI64(h, h->dcnt.p + 32, 0, 32);
                for (int32_t i = i0; i < i0 + c; ++i)
                    launch(h, "reached_arcs", kReachedArcsD, boundedGrid(V, 256), 256, 0,
                           (const uint64_t*)(out + (size_t)i * V), (const int32_t*)h->fw_beg.p, V,
                           h->dcnt.p + 32 + (i - i0));
                readCounters(h, h->dcnt.p + 32, c, e.data() + i0);
            }
            HIPCHK(hipMemcpy(edges_out, e.data(), sizeof(int64_t) * (size_t)k, hipMemcpyDefault));
        }
        if (dist && !direct) copyOut(h, dist, out, sizeof(uint64_t) * n);
        syncStream(h);
    });
}

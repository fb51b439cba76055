// Single-source shortest paths: replaces GIR::sssp / GIR::deltaStepping
// (gir_query.cpp:629-703) for k sources.
//
// Delta-stepping with device-resident buckets.  All sources of a call (up to
// SB per batch) advance together in rounds,
// one row of the grid's blocks per source (gridRow).  Per source:
//   - the current bucket is [lo, lo + delta); its work list is a queue;
//   - a relaxation that improves a distance sets the head's bit in the near
//     bitmap (lands inside the bucket) or in the far bitmap (beyond it), so
//     duplicates collapse and no queue counter is hit per arc; the first
//     near bit of a round also raises the round's near flag;
//   - after each round a compaction pass turns the near bitmap into the next
//     queue - the reference re-processes bucket k until nothing new lands
//     in it;
//   - a round that raised no near flag has drained the bucket: the heavy
//     arcs (w > delta) of the vertices processed in it are relaxed once, a
//     pass over the far bitmap finds the smallest distance >= hi, which
//     picks the next non-empty bucket, and a split pass moves the far
//     entries inside it into the queue (entries below it are settled).
// Every bitmap has a summary (one bit per 64-bit word) so the passes touch
// only non-zero words.  Relaxation is an atomicMin on 64-bit distances.
// Weights: one uint32 per forward slot, or the reference's hash
// GIR::edgeWeight.  Rows longer than HEAVY_T arcs are walked by the whole
// grid (kHeavyRows).
//
// HBM traffic per source (algorithmic): 8V (init + final distances) + 8 n_s
// (row offsets) + 16 m_s (adjacency 4 + weight 4 + distance 8 per arc).
#include <vector>

#include "common.h"
#include "lists.h"
#include "msssp.h"

namespace gir {

std::vector<int32_t> hostSources(girhip_graph* h, const int32_t* src, int32_t k);

constexpr uint64_t INF = ~0ULL;
constexpr int SB = 32;  // sources per batch

// Per-source control slots.  Every (slot, source) pair has a 128-byte line
// of its own (CSTR int64): the queue counters and flags take atomics and
// stores from every wave of a pass, and counters of different sources on
// one line would serialise each other.
enum : int {
    C_NEAR0 = 0,  // queue length, parity 0
    C_NEAR1,      // queue length, parity 1
    C_NSET0,      // 1: the round of parity 0 put a vertex into the bucket
    C_NSET1,      //    (parity 1)
    C_LO,         // bucket [lo, hi) (uint64)
    C_HI,
    C_FMIN,       // smallest far distance >= hi (uint64)
    C_HQ,         // heavy-phase queue length
    C_NSLOT
};
constexpr int CSTR = 16;
constexpr int C_SLOTS = C_NSLOT * SB * CSTR;  // first index past the slot lines
__host__ __device__ __forceinline__ size_t cIdx(int slot, int s) { return ((size_t)slot * SB + s) * CSTR; }

struct SsspBatch {
    int32_t* qn[2];      // [SB][V] queues by parity
    int32_t* qh;         // [SB][V] vertices whose heavy arcs are due
    uint64_t* nearbits;  // [SB][W]
    uint64_t* farbits;   // [SB][W]
    uint64_t* donebits;  // [SB][W] processed in the current bucket
    uint64_t* nsum;      // [SB][W2] summaries: bit g of word i <=> word 64i+g
    uint64_t* fsum;      //          of the bitmap may be non-zero
    uint64_t* dsum;
    int64_t* ctl;        // slot lines (cIdx), then the persistent parity, the overflow flag, a gather area
    void* dist;          // [SB][V] distances: uint32 (d32) or uint64
    int64_t V, W, W2;
    int k;               // sources in the batch
    int d32;             // 1: 32-bit distances, UINT32_MAX = unreachable
    int xcd;             // 1: block b serves source b mod k (all of a source's blocks on the same XCDs)
};

// The part of a 1D grid of (blocks per source) x k blocks that serves one
// source: its row s, this block's index b among the row's nb blocks.  By
// default the rows are contiguous.  With xcd, block b serves source b mod k:
// blocks are dealt to the 8 XCDs round robin, so a source's blocks share
// XCDs and its distances stay in those XCDs' L2 (measured slower at RMAT-22:
// the sources' frontiers differ in size, and each is confined to its XCDs).
struct Row {
    int s;
    int64_t b, nb;
};
__device__ __forceinline__ Row gridRow(const SsspBatch& B) {
    const int64_t nb = gridDim.x / B.k;
    if (B.xcd) return Row{(int)(blockIdx.x % B.k), (int64_t)(blockIdx.x / B.k), nb};
    return Row{(int)(blockIdx.x / nb), (int64_t)(blockIdx.x % nb), nb};
}

// 32-bit distances halve the array the random relaxations hit (a batch of
// RMAT-22 distances then fits the 256 MB Infinity Cache).  A relaxation that
// would reach UINT32_MAX raises this flag instead, and the batch is redone
// with 64-bit distances.
constexpr int C_OVF = C_SLOTS + 1;
constexpr int C_GATHER = C_SLOTS + CSTR;  // SB entries: one slot of every source, for host readback
constexpr int C_TOTAL = C_GATHER + SB;

__device__ __forceinline__ int64_t* cp(const SsspBatch& B, int slot, int s) { return B.ctl + cIdx(slot, s); }

__device__ __forceinline__ uint64_t dGet(const SsspBatch& B, int s, int32_t x) {
    const size_t i = (size_t)s * B.V + x;
    if (B.d32) {
        const uint32_t v = ((const uint32_t*)B.dist)[i];
        return v == 0xffffffffu ? INF : (uint64_t)v;
    }
    return ((const uint64_t*)B.dist)[i];
}
// atomicMin; returns the previous distance
__device__ __forceinline__ uint64_t dMin(const SsspBatch& B, int s, int32_t x, uint64_t nd) {
    const size_t i = (size_t)s * B.V + x;
    if (B.d32) {
        const uint32_t old = atomicMin((unsigned int*)B.dist + i, (unsigned int)nd);
        return old == 0xffffffffu ? INF : (uint64_t)old;
    }
    return atomicMin((unsigned long long*)B.dist + i, (unsigned long long)nd);
}

// atomicMin whose previous value is not awaited
__device__ __forceinline__ void dMinNoRet(const SsspBatch& B, int s, int32_t x, uint64_t nd) {
    const size_t i = (size_t)s * B.V + x;
    if (B.d32) atomicMin((unsigned int*)B.dist + i, (unsigned int)nd);
    else atomicMin((unsigned long long*)B.dist + i, (unsigned long long)nd);
}

__device__ __forceinline__ int64_t* nearCount(const SsspBatch& B, int par, int s) {
    return cp(B, par ? C_NEAR1 : C_NEAR0, s);
}
__device__ __forceinline__ int64_t* nearSet(const SsspBatch& B, int par, int s) {
    return cp(B, par ? C_NSET1 : C_NSET0, s);
}

__global__ void kBInit(SsspBatch B, const int32_t* __restrict__ src, uint64_t delta) {
    const int s = threadIdx.x;
    if (s >= SB) return;
    for (int j = 0; j < C_NSLOT; ++j) *cp(B, j, s) = 0;
    *cp(B, C_HI, s) = (int64_t)delta;
    if (s == 0) B.ctl[C_OVF] = 0;
    if (s < B.k) {
        const int32_t v = src[s];
        if (v >= 0 && v < B.V) {
            if (B.d32) ((uint32_t*)B.dist)[(size_t)s * B.V + v] = 0;
            else ((uint64_t*)B.dist)[(size_t)s * B.V + v] = 0;
            B.qn[0][(size_t)s * B.V] = v;
            *cp(B, C_NEAR0, s) = 1;
        }
    }
}

// Relaxation of arc slot j of row p (owner distance dp) for source s.
// heavy = 0: only light arcs (w <= delta, the bucket's inner rounds);
// heavy = 1: only heavy arcs (w > delta, once per vertex and bucket).
// With the split view (mid) each pass walks only its own segment of a row,
// so the class test drops out and no pass reads the other class's arcs.
// scope (optional): arcs whose ends carry different labels are skipped (the
// SCC-restricted delta-stepping of the two-level queries); unit: every arc
// weighs 1 (BFS levels as distances).
struct RelaxF {
    SsspBatch B;
    const int32_t* fa;
    const uint32_t* w;
    uint64_t delta;
    int heavy;
    int par;
    const int32_t* scope;
    int unit;
    const int32_t* mid;  // split view (fa / w are its arrays): rows walk only their class's arcs
    __device__ int64_t operator()(int32_t p, uint64_t dp, int32_t s, int64_t j) const {
        uint64_t wt;
        int32_t x;
        if (unit) {
            if (heavy) return 0;
            wt = 1;
            x = fa[j];
        } else if (mid) {
            wt = w[j];
            x = fa[j];
        } else if (w) {
            wt = w[j];
            if ((wt > delta) != (heavy != 0)) return 0;
            x = fa[j];
        } else {
            x = fa[j];
            wt = hashWeight(p, x);
            if ((wt > delta) != (heavy != 0)) return 0;
        }
        if (scope && scope[x] != scope[p]) return 0;
        const uint64_t nd = dp + wt;
        if (B.d32 && nd >= 0xffffffffull) {  // beyond 32 bits: the batch is redone with 64
            B.ctl[C_OVF] = 1;
            return 0;
        }
        if (nd < dGet(B, s, x)) {  // plain read: distances only fall, so a stale value is a safe pre-check
            // The atomicMin's old value is not awaited: every relaxation that
            // lowers a distance passed this pre-check and marks x below, so
            // marking on the pre-check alone loses nothing.  A relaxation that
            // lost the race marks x too, which costs at most a redundant visit:
            // a near mark makes the rounds visit x with its current distance,
            // a far mark of a vertex settled lower is dropped by the split.
            dMin(B, s, x, nd);
            if (nd < (uint64_t)*cp(B, C_HI, s)) {
                // the first bit of a word raises the flag: the near bitmap
                // is empty at the start of every round
                if (setBit(B.nearbits + (size_t)s * B.W, B.nsum + (size_t)s * B.W2, x)) *nearSet(B, par, s) = 1;
            } else {
                setBit(B.farbits + (size_t)s * B.W, B.fsum + (size_t)s * B.W2, x);
            }
        }
        return 0;
    }
    // U arcs of rows p (owner distances dp, all of one source s) at a time:
    // the arcs' slots, then every head's distance and summary words, are
    // loaded for all U before any is used; the updates (atomicMin, bitmap
    // and summary atomicOr) are then issued without awaiting a result.
    template <int U>
    __device__ void many(const int32_t (&p)[U], const uint64_t (&dp)[U], const int32_t (&sv)[U],
                         const int64_t (&j)[U], const bool (&ok)[U]) const {
        const int s = sv[0];  // one source per wave
        int32_t x[U];
        uint32_t wt[U];
        bool go[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            go[u] = ok[u];
            x[u] = go[u] ? fa[j[u]] : 0;
            wt[u] = (go[u] && !unit && w) ? w[j[u]] : 1;
        }
        uint64_t dx[U];
        uint32_t nsw[U], fsw[U];  // the halves of the summary words that hold x's bits
        const uint64_t hi = (uint64_t)*cp(B, C_HI, s);
        uint64_t* nb = B.nearbits + (size_t)s * B.W;
        uint64_t* ns = B.nsum + (size_t)s * B.W2;
        uint64_t* fbits = B.farbits + (size_t)s * B.W;
        uint64_t* fs = B.fsum + (size_t)s * B.W2;
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (go[u] && !unit && !w) wt[u] = hashWeight(p[u], x[u]);
            if (go[u] && unit && heavy) go[u] = false;
            if (go[u] && !unit && !mid && ((wt[u] > delta) != (heavy != 0))) go[u] = false;
            if (go[u] && scope && scope[x[u]] != scope[p[u]]) go[u] = false;
            dx[u] = go[u] ? dGet(B, s, x[u]) : 0;
            nsw[u] = go[u] ? ((const uint32_t*)ns)[x[u] >> 11] : 0;
            fsw[u] = go[u] ? ((const uint32_t*)fs)[x[u] >> 11] : 0;
        }
        bool ovf = false, nset = false;
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (!go[u]) continue;
            const uint64_t nd = dp[u] + wt[u];
            if (B.d32 && nd >= 0xffffffffull) {  // beyond 32 bits: the batch is redone with 64
                ovf = true;
                continue;
            }
            if (nd < dx[u]) {  // (dx: a plain read, safe as a pre-check as distances only fall)
                dMinNoRet(B, s, x[u], nd);
                if (nd < hi) nset |= markBitHalf(nb, ns, x[u], nsw[u]);
                else markBitHalf(fbits, fs, x[u], fsw[u]);
            }
        }
        if (ovf) B.ctl[C_OVF] = 1;
        if (nset) *nearSet(B, par, s) = 1;
    }
};

// ---- light/heavy split view -------------------------------------------------
// For bucket width delta, every forward row reordered so that its arcs of
// weight <= delta come first; mid[u] = where the heavier ones start.  Rows of
// at most SPLIT_SHORT arcs take one thread each (light arcs forward from the
// row start, heavy ones backward from its end); longer rows are listed for
// kSplitLong, one block per row (block-wide ranks of the light arcs).  The
// order inside a class is free: distances do not depend on the order in
// which arcs are relaxed.
constexpr int SPLIT_SHORT = 64;

struct SplitArgs {
    const int32_t* beg;
    const int32_t* adj;
    const uint32_t* w;  // null: the reference's hash weights
    int32_t V;
    uint64_t delta;
    int32_t* oadj;
    uint32_t* ow;
    int32_t* mid;
    int32_t* longq;
    int64_t* nlong;
    __device__ uint32_t weight(int32_t u, int64_t j, int32_t x) const { return w ? w[j] : hashWeight(u, x); }
};

__global__ void kSplitShort(SplitArgs A) {
    for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < A.V; base += (int64_t)gridDim.x * blockDim.x) {
        const int64_t u = base + threadIdx.x;
        const bool act = u < A.V;
        const int64_t b = act ? A.beg[u] : 0, e = act ? A.beg[u + 1] : 0;
        const bool lng = act && e - b > SPLIT_SHORT;
        const int64_t pos = wavePush(A.nlong, lng);
        if (lng) A.longq[pos] = (int32_t)u;
        if (act && !lng) {
            int64_t lo = b, hi = e;
            for (int64_t j = b; j < e; ++j) {
                const int32_t x = A.adj[j];
                const uint32_t wt = A.weight((int32_t)u, j, x);
                const int64_t o = wt <= A.delta ? lo++ : --hi;
                A.oadj[o] = x;
                A.ow[o] = wt;
            }
            A.mid[u] = (int32_t)lo;
        }
    }
}

__global__ void kSplitLong(SplitArgs A) {
    __shared__ int32_t wcnt[1024 / WAVE];
    __shared__ int64_t tot;
    const int64_t n = *A.nlong;
    const int lane = laneId(), wv = threadIdx.x / WAVE, nw = blockDim.x / WAVE;
    for (int64_t i = blockIdx.x; i < n; i += gridDim.x) {
        const int32_t u = A.longq[i];
        const int64_t b = A.beg[u], e = A.beg[u + 1];
        int64_t c = 0;  // light arcs of the row
        for (int64_t j = b + threadIdx.x; j < e; j += blockDim.x) c += A.weight(u, j, A.adj[j]) <= A.delta;
        c = waveSum(c);
        if (lane == 0) wcnt[wv] = (int32_t)c;
        __syncthreads();
        if (threadIdx.x == 0) {
            int64_t t = 0;
            for (int k = 0; k < nw; ++k) t += wcnt[k];
            tot = t;
        }
        __syncthreads();
        const int64_t m = b + tot;
        int64_t lc = 0;  // light arcs of the earlier chunks
        for (int64_t c0 = b; c0 < e; c0 += blockDim.x) {
            const int64_t j = c0 + threadIdx.x;
            const bool val = j < e;
            const int32_t x = val ? A.adj[j] : 0;
            const uint32_t wt = val ? A.weight(u, j, x) : 0;
            const bool light = val && wt <= A.delta;
            const uint64_t bal = __ballot(light);
            __syncthreads();  // wcnt: the previous chunk's (or the count's) reads are done
            if (lane == 0) wcnt[wv] = __popcll(bal);
            __syncthreads();
            int64_t before = 0, all = 0;
            for (int k = 0; k < nw; ++k) {
                const int t = wcnt[k];
                if (k < wv) before += t;
                all += t;
            }
            if (val) {
                const int64_t lr = lc + before + __popcll(bal & lanesBelow());  // light arcs before j in the row
                const int64_t o = light ? b + lr : m + (j - b) - lr;
                A.oadj[o] = x;
                A.ow[o] = wt;
            }
            lc += all;
        }
        if (threadIdx.x == 0) A.mid[u] = (int32_t)m;
        __syncthreads();  // tot and wcnt are rewritten for the next row
    }
}

// ---- round phases (per source s; the grid strides over x) -----------------

// Per-round slots of source s that the round fills later, and the next
// round's near flag (one thread).
__device__ __forceinline__ void resetRound(const SsspBatch& B, int par, int s) {
    *nearCount(B, par ^ 1, s) = 0;
    *nearSet(B, par ^ 1, s) = 0;
    *cp(B, C_FMIN, s) = (int64_t)INF;
    *cp(B, C_HQ, s) = 0;
}

// Relax the n queued vertices q of source s (grid stride, uniform per wave).
// track_done: mark them processed (their heavy arcs are due when the bucket
// drains).
// U: arcs per lane at a time (waveRowsHU) where above 1.
template <int U>
__device__ __forceinline__ void relaxQueue(const RelaxF& f, const int32_t* __restrict__ fb, HeavyList H,
                                           const int32_t* q, int64_t n, int s, bool track_done, int64_t b0,
                                           int64_t nb) {
    const SsspBatch& B = f.B;
    const int64_t stride = nb * blockDim.x;
    for (int64_t base = b0 * blockDim.x + (threadIdx.x & ~(WAVE - 1)); base < n; base += stride) {
        const int64_t i = base + laneId();
        const bool act = i < n;
        const int32_t u = act ? q[i] : 0;
        const uint64_t du = act ? dGet(B, s, u) : 0;
        if (track_done) waveMarkBits(B.donebits + (size_t)s * B.W, B.dsum + (size_t)s * B.W2, u, act);
        int32_t b = act ? fb[u] : 0, e = act ? fb[u + 1] : 0;
        if (f.mid && act) {
            const int32_t m = f.mid[u];
            if (f.heavy) b = m;
            else e = m;
        }
        if constexpr (U > 1) waveRowsHU<U>(act, u, du, s, b, e, H, f);
        else waveRowsH(act, u, du, s, b, e, H, f);
    }
}

// Drains one summary word (64 bitmap words) of a bitmap: clears the set words
// and the summary word, and appends their vertices to q at counter *cnt (one
// atomic per wave).
__device__ __forceinline__ void drainGroup(uint64_t* bits, uint64_t* sum, int64_t g, uint64_t sm, int32_t* q,
                                           int64_t* cnt) {
    const int l = laneId();
    uint64_t m = 0;
    if ((sm >> l) & 1) {
        uint64_t* wp = bits + g * 64 + l;
        m = *wp;
        if (m) *wp = 0;
    }
    if (l == 0) sum[g] = 0;
    int64_t pos = waveReserve(__popcll(m), cnt);
    const int64_t wi = g * 64 + l;
    while (m) {
        const int b = __ffsll((long long)m) - 1;
        m &= m - 1;
        q[pos++] = (int32_t)(wi * 64 + b);
    }
}

// LDS of the far-bitmap passes: per wave, WB_CAP ids and 64 bitmap words.
struct FarLds {
    int32_t ids[1024 / WAVE][WB_CAP];
    uint64_t w[1024 / WAVE][64];  // keep bits of the wave's 64 words
    uint64_t t[1024 / WAVE][64];  // take bits
};

// Smallest distance >= hi over one summary word of source s's far bitmap
// (per lane, lane-parallel distance reads; reduced by farMinCommit).
__device__ __forceinline__ uint64_t farMinGroup(const SsspBatch& B, int s, int64_t g, uint64_t sm, uint64_t hi,
                                                uint64_t mn, int32_t* lds) {
    const int l = laneId();
    const uint64_t m = ((sm >> l) & 1) ? B.farbits[(size_t)s * B.W + g * 64 + l] : 0;
    waveBits(m, g * 64 + l, lds, [&](bool act, int32_t x) {
        if (!act) return;
        const uint64_t v = dGet(B, s, x);
        if (v >= hi && v < mn) mn = v;
    });
    return mn;
}

__device__ __forceinline__ void farMinCommit(const SsspBatch& B, int s, unsigned long long mn) {
    for (int o = WAVE / 2; o > 0; o >>= 1) {
        const unsigned long long t = __shfl_xor(mn, o);
        mn = t < mn ? t : mn;
    }
    if (laneId() == 0 && mn != INF) atomicMin((unsigned long long*)cp(B, C_FMIN, s), mn);
}

// After the relaxation of parity par, for source s: if the round put
// vertices into the bucket, its near bitmap -> next queue; otherwise the
// bucket has drained: with heavy arcs pending (heavy = 1) the processed
// vertices -> heavy-phase queue, else the smallest far distance is found
// right here.
__device__ __forceinline__ void compactPhase(const SsspBatch& B, int par, int heavy, int G, int s, FarLds& sh,
                                             int64_t bb, int64_t nbb) {
    const size_t bo = (size_t)s * B.W, so = (size_t)s * B.W2, qo = (size_t)s * B.V;
    if (ldAcq(nearSet(B, par, s))) {
        forSummaryB(B.nsum + so, B.W2, G, bb, nbb, [&](int64_t g, uint64_t sm) {
            drainGroup(B.nearbits + bo, B.nsum + so, g, sm, B.qn[par ^ 1] + qo, nearCount(B, par ^ 1, s));
        });
    } else if (heavy) {
        forSummaryB(B.dsum + so, B.W2, G, bb, nbb, [&](int64_t g, uint64_t sm) {
            drainGroup(B.donebits + bo, B.dsum + so, g, sm, B.qh + qo, cp(B, C_HQ, s));
        });
    } else {
        const uint64_t hi = (uint64_t)ldAcq(cp(B, C_HI, s));
        uint64_t mn = INF;
        int32_t* lds = sh.ids[threadIdx.x / WAVE];
        forSummaryB(B.fsum + so, B.W2, G, bb, nbb,
                    [&](int64_t g, uint64_t sm) { mn = farMinGroup(B, s, g, sm, hi, mn, lds); });
        farMinCommit(B, s, mn);
    }
}

// Drained source s after its heavy phase: smallest far distance >= hi.
__device__ __forceinline__ void farMinPhase(const SsspBatch& B, int par, int G, int s, FarLds& sh, int64_t bb,
                                            int64_t nbb) {
    if (ldAcq(nearSet(B, par, s))) return;
    const uint64_t hi = (uint64_t)ldAcq(cp(B, C_HI, s));
    uint64_t mn = INF;
    int32_t* lds = sh.ids[threadIdx.x / WAVE];
    forSummaryB(B.fsum + (size_t)s * B.W2, B.W2, G, bb, nbb,
                [&](int64_t g, uint64_t sm) { mn = farMinGroup(B, s, g, sm, hi, mn, lds); });
    farMinCommit(B, s, mn);
}

// Drained source s with far work left advances to the bucket holding the
// smallest far distance (every block derives the same [lo, hi); block 0
// stores it for the next round).  Far entries inside it -> next queue,
// below it -> settled (dropped), both leave the far bitmap; the rest stay.
// A summary word is cleared once its 64 bitmap words are all empty.
__device__ __forceinline__ void splitPhase(const SsspBatch& B, int par, uint64_t delta, int G, int s, FarLds& sh,
                                           int64_t bb, int64_t nbb) {
    if (ldAcq(nearSet(B, par, s))) return;
    const uint64_t mfar = (uint64_t)ldAcq(cp(B, C_FMIN, s));
    if (mfar == INF) return;  // no far work left: this source is done
    const uint64_t lo = (mfar / delta) * delta;
    const uint64_t hi = lo > INF - delta ? INF : lo + delta;
    if (bb == 0 && threadIdx.x == 0) {
        *cp(B, C_LO, s) = (int64_t)lo;
        *cp(B, C_HI, s) = (int64_t)hi;
    }
    int32_t* q = B.qn[par ^ 1] + (size_t)s * B.V;
    uint64_t* fs = B.fsum + (size_t)s * B.W2;
    uint64_t* fbits = B.farbits + (size_t)s * B.W;
    int64_t* cnt = nearCount(B, par ^ 1, s);
    const int wv = threadIdx.x / WAVE;
    int32_t* lds = sh.ids[wv];
    uint64_t* keepw = sh.w[wv];
    uint64_t* takew = sh.t[wv];
    forSummaryB(fs, B.W2, G, bb, nbb, [&](int64_t g, uint64_t sm) {
        const int l = laneId();
        const int64_t wi = g * 64 + l;
        const uint64_t m = ((sm >> l) & 1) ? fbits[wi] : 0;
        keepw[l] = 0;
        takew[l] = 0;
        waveSync();
        waveBits(m, wi, lds, [&](bool act, int32_t x) {
            if (!act) return;
            const uint64_t v = dGet(B, s, x);
            uint64_t* mw = v >= hi ? keepw : (v >= lo ? takew : nullptr);
            if (mw) atomicOr((unsigned long long*)&mw[(x >> 6) & 63], 1ULL << (x & 63));
        });
        waveSync();
        const uint64_t keep = keepw[l];
        uint64_t take = takew[l];
        if (keep != m) fbits[wi] = keep;
        if (__ballot(keep != 0) == 0 && l == 0) fs[g] = 0;
        int64_t pos = waveReserve(__popcll(take), cnt);
        while (take) {
            const int b = __ffsll((long long)take) - 1;
            take &= take - 1;
            q[pos++] = (int32_t)(wi * 64 + b);
        }
    });
}

// ---- one kernel per phase (grid row y = source) ----------------------------

template <int U>
__global__ void kBRelax(RelaxF f, const int32_t* __restrict__ fb, HeavyList H, int track_done) {
    const SsspBatch& B = f.B;
    const Row r = gridRow(B);
    if (r.b == 0 && threadIdx.x == 0) resetRound(B, f.par, r.s);
    relaxQueue<U>(f, fb, H, B.qn[f.par] + (size_t)r.s * B.V, *nearCount(B, f.par, r.s), r.s, track_done != 0, r.b,
               r.nb);
}

__global__ void kBCompact(SsspBatch B, int par, int heavy, int G) {
    __shared__ FarLds sh;
    const Row r = gridRow(B);
    compactPhase(B, par, heavy, G, r.s, sh, r.b, r.nb);
}

template <int U>
__global__ void kBRelaxHeavy(RelaxF f, const int32_t* __restrict__ fb, HeavyList H) {
    const SsspBatch& B = f.B;
    const Row r = gridRow(B);
    relaxQueue<U>(f, fb, H, B.qh + (size_t)r.s * B.V, *cp(B, C_HQ, r.s), r.s, false, r.b, r.nb);
}

__global__ void kBFarMin(SsspBatch B, int par, int G) {
    __shared__ FarLds sh;
    const Row r = gridRow(B);
    farMinPhase(B, par, G, r.s, sh, r.b, r.nb);
}

__global__ void kBSplit(SsspBatch B, int par, uint64_t delta, int G) {
    __shared__ FarLds sh;
    const Row r = gridRow(B);
    splitPhase(B, par, delta, G, r.s, sh, r.b, r.nb);
}

// ---- persistent rounds -----------------------------------------------------
// One cooperative launch runs whole rounds (relax / compact / [heavy relax /
// far minimum] / split, a grid barrier after each phase, sources in turn
// within a phase) while every queue holds at most cap vertices, for at most
// max_rounds rounds.  io[0]: round parity in and out.  Every block reads the
// same counters after a barrier, so all leave the loop together.  Used only
// when no row of the store can reach the heavy list.
__global__ void __launch_bounds__(256) kSsspPersist(RelaxF fl, RelaxF fh, const int32_t* __restrict__ fb, HeavyList H,
                                                    int heavy, uint64_t delta, int G, int64_t cap, int32_t max_rounds,
                                                    int64_t* io, uint32_t* bar) {
    __shared__ FarLds sh;
    const SsspBatch& B = fl.B;
    const int k = B.k;
    int par = (int)ldAcq(io);
    for (int32_t r = 0; r < max_rounds && !gridFailed(bar); ++r) {
        int64_t mx = 0;
        for (int s = 0; s < k; ++s) mx = std::max<int64_t>(mx, ldAcq(nearCount(B, par, s)));
        if (mx == 0 || mx > cap) break;
        if (blockIdx.x == 0 && threadIdx.x < k) resetRound(B, par, threadIdx.x);
        fl.par = par;
        fh.par = par;
        for (int s = 0; s < k; ++s) relaxQueue<1>(fl, fb, H, B.qn[par] + (size_t)s * B.V, ldAcq(nearCount(B, par, s)),
                                               s, heavy != 0, blockIdx.x, gridDim.x);
        gridSync(bar);
        for (int s = 0; s < k; ++s) compactPhase(B, par, heavy, G, s, sh, blockIdx.x, gridDim.x);
        gridSync(bar);
        if (heavy) {
            for (int s = 0; s < k; ++s) relaxQueue<1>(fh, fb, H, B.qh + (size_t)s * B.V, ldAcq(cp(B, C_HQ, s)), s,
                                                   false, blockIdx.x, gridDim.x);
            gridSync(bar);
            for (int s = 0; s < k; ++s) farMinPhase(B, par, G, s, sh, blockIdx.x, gridDim.x);
            gridSync(bar);
        }
        for (int s = 0; s < k; ++s) splitPhase(B, par, delta, G, s, sh, blockIdx.x, gridDim.x);
        gridSync(bar);
        par ^= 1;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) *io = par;
}

// ---- small frontiers: one workgroup per source -------------------------------
// A high-diameter graph (a lattice) walks tens of thousands of buckets of a
// few hundred vertices each; there the multi-kernel rounds cost their five
// launches per round, not their work.  The tiny engine runs the same rounds
// for source s = blockIdx.x inside one workgroup of 1024 threads, separated
// by __syncthreads only, until a queue outgrows TINY_CAP or the source is
// settled.  A queued vertex has all its arcs relaxed at once (a lattice's
// vertices are queued about once per bucket, so holding the heavy arcs back
// to the end of the bucket saves nothing and costs a pass): there is no
// heavy pass and no done set.  The near and far sets keep their bitmaps (so
// that the multi-kernel rounds can take over at any round boundary: far
// bits with their summary bits, an empty near bitmap, an empty done set)
// and are also kept as lists: an id joins a list when its bit turns on (the
// atomicOr's old value), so no pass walks a bitmap.  Rows are walked whole by the waves
// (no heavy-row list): the engine runs only while no row is longer than
// TINY_DEG arcs.  Distances written by one wave's atomics are read back by
// other waves with agent-scope loads (dGetAcq): a vector-cache copy may be
// stale.
constexpr int TINY_CAP = 8192;    // a larger queue hands the rounds back to the grid kernels
constexpr int TINY_ENTER = 1024;  // queues at most this long (every source) enter the tiny engine
constexpr int64_t TINY_DEG = 1024;
constexpr int TINY_ILP = 8;  // far-list entries per thread whose loads are in flight together

// The wave's smallest value, folded into *m (LDS) by its first lane.
__device__ __forceinline__ void waveMinLds(unsigned long long mn, unsigned long long* m) {
    for (int o = WAVE / 2; o > 0; o >>= 1) {
        const unsigned long long t = __shfl_xor(mn, o);
        mn = t < mn ? t : mn;
    }
    if (laneId() == 0 && mn != INF) atomicMin(m, mn);
}

__device__ __forceinline__ uint64_t dGetAcq(const SsspBatch& B, int s, int32_t x) {
    const size_t i = (size_t)s * B.V + x;
    if (B.d32) {
        const uint32_t v = ldAcq((const uint32_t*)B.dist + i);
        return v == 0xffffffffu ? INF : (uint64_t)v;
    }
    return ldAcq((const uint64_t*)B.dist + i);
}

// Sets bit x of a bitmap; true when this call turned it on.
__device__ __forceinline__ bool bitOn(uint64_t* bits, int32_t x) {
    const unsigned long long bit = 1ULL << (x & 63);
    return (atomicOr((unsigned long long*)&bits[x >> 6], bit) & bit) == 0;
}

// RelaxF's arc rule for the tiny engine: a lowered head joins the next
// queue (inside the bucket) or the far list (beyond it) when its bit turns
// on; list slots come from LDS counters.  arcs: 0 light, 1 heavy, 2 every
// arc of the row.
struct TinyF {
    SsspBatch B;
    const int32_t* fa;
    const uint32_t* w;
    uint64_t delta;
    int arcs;
    const int32_t* scope;
    int unit;
    const int32_t* mid;
    int s;
    uint64_t hi;
    int32_t* nq;      // next queue
    int32_t* fl;      // far list
    int* nq_n;        // LDS counters
    int* fl_n;
    unsigned long long* fmin;  // LDS: at most the smallest far distance
    template <int U>
    __device__ __forceinline__ void many(const int32_t (&p)[U], const uint64_t (&dp)[U], const int32_t (&)[U],
                                         const int64_t (&j)[U], const bool (&ok)[U]) const {
        int32_t x[U];
        uint32_t wt[U];
        uint64_t dx[U];
        bool go[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            go[u] = ok[u];
            x[u] = go[u] ? fa[j[u]] : 0;
            wt[u] = (go[u] && !unit && w) ? w[j[u]] : 1;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (go[u] && !unit && !w) wt[u] = hashWeight(p[u], x[u]);
            if (go[u] && unit && arcs == 1) go[u] = false;
            if (go[u] && !unit && !mid && arcs != 2 && ((wt[u] > delta) != (arcs == 1))) go[u] = false;
            if (go[u] && scope && scope[x[u]] != scope[p[u]]) go[u] = false;
            dx[u] = go[u] ? dGet(B, s, x[u]) : 0;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (!go[u]) continue;
            const uint64_t nd = dp[u] + wt[u];
            if (B.d32 && nd >= 0xffffffffull) {  // beyond 32 bits: the batch is redone with 64
                B.ctl[C_OVF] = 1;
                continue;
            }
            if (nd >= dx[u]) continue;  // (a plain read, safe as a pre-check as distances only fall)
            dMinNoRet(B, s, x[u], nd);
            if (nd < hi) {
                if (bitOn(B.nearbits + (size_t)s * B.W, x[u])) nq[atomicAdd(nq_n, 1)] = x[u];
            } else {
                atomicMin(fmin, (unsigned long long)nd);
                if (bitOn(B.farbits + (size_t)s * B.W, x[u])) {
                    atomicOr((unsigned long long*)&B.fsum[(size_t)s * B.W2 + (x[u] >> 12)],
                             1ULL << ((x[u] >> 6) & 63));
                    fl[atomicAdd(fl_n, 1)] = x[u];
                }
            }
        }
    }
};

// The ids of a bitmap's set bits appended to list at LDS counter *n (the
// whole block; summaries say which words to read).
__device__ __forceinline__ void listBits(const uint64_t* bits, const uint64_t* sum, int64_t W2, int32_t* list,
                                         int* n) {
    forSummaryB(sum, W2, 64, 0, 1, [&](int64_t g, uint64_t sm) {
        const int l = laneId();
        uint64_t m = ((sm >> l) & 1) ? bits[g * 64 + l] : 0;
        const int c = __popcll(m);
        int incl = c;
        for (int d = 1; d < WAVE; d <<= 1) {
            const int o = __shfl_up(incl, d);
            if (l >= d) incl += o;
        }
        int base = 0;
        if (l == WAVE - 1 && incl) base = atomicAdd(n, incl);
        base = __shfl(base, WAVE - 1) + incl - c;
        while (m) {
            const int b = __ffsll((long long)m) - 1;
            m &= m - 1;
            list[base++] = (int32_t)((g * 64 + l) * 64 + b);
        }
    });
}

// The phases of one source's rounds hand data from wave to wave of one
// workgroup only: a workgroup barrier (with its workgroup-scope fences) is
// the whole synchronisation.
__device__ __forceinline__ void blockSync() { __syncthreads(); }

// Relax the n vertices of list q (the arcs f.arcs names) with the whole
// block: each wave takes an equal share of the list (a short list spreads
// over every wave instead of filling the first few), U arcs per lane at a
// time.
template <int U>
__device__ __forceinline__ void tinyRelax(const TinyF& f, const int32_t* __restrict__ fb, const int32_t* q, int n) {
    const SsspBatch& B = f.B;
    const int s = f.s;
    const int nw = (int)blockDim.x / WAVE, per = (n + nw - 1) / nw;
    const int beg = (int)(threadIdx.x / WAVE) * per, end = min(n, beg + per);
    for (int base = beg; base < end; base += WAVE) {
        const int i = base + laneId();
        const bool act = i < end;
        const int32_t u = act ? q[i] : 0;
        const uint64_t du = act ? dGetAcq(B, s, u) : 0;
        int32_t b = act ? fb[u] : 0, e = act ? fb[u + 1] : 0;
        if (f.mid && act && f.arcs != 2) {
            const int32_t m = f.mid[u];
            if (f.arcs == 1) b = m;
            else e = m;
        }
        waveFlatU<U>(u, du, s, b, act ? e - b : 0, f);
    }
}

// Rounds of source blockIdx.x from the state left at a round boundary
// (queue of parity par; an empty queue with far entries, as a seeded run
// starts, goes straight to the advance) until its queue is empty after an
// advance (settled), outgrows TINY_CAP, or max_rounds pass.  The queue is
// left in parity par.
// fl0 / fl1: far-list buffers of V ids per source.
template <int U>
__global__ void __launch_bounds__(1024) kSsspTiny(TinyF f, const int32_t* __restrict__ fb, int par0,
                                                  int64_t max_rounds, int32_t* fl0, int32_t* fl1) {
    __shared__ int nq_n, fl_n, fl2_n, dl_n;
    // fmin: at most the smallest far distance at or past hi (the far list's
    // minimum after each advance, lowered by every far relaxation since; a
    // far vertex lowered into the bucket can leave it too small, which costs
    // at most a visit of an empty bucket)
    __shared__ unsigned long long fmin, fmin2;
    const SsspBatch& B = f.B;
    const int s = blockIdx.x;
    const size_t vo = (size_t)s * B.V, wo = (size_t)s * B.W, so = (size_t)s * B.W2;
    int par = par0;
    int n = (int)ldAcq(nearCount(B, par, s));
    if (n > TINY_CAP) return;
    uint64_t lo = (uint64_t)ldAcq(cp(B, C_LO, s)), hi = (uint64_t)ldAcq(cp(B, C_HI, s));
    int32_t* fl = fl0 + vo;
    int32_t* fl2 = fl1 + vo;
    if (threadIdx.x == 0) {
        fl_n = 0;
        dl_n = 0;
        fmin = INF;
    }
    __syncthreads();
    int32_t* dl = B.qh + vo;  // the done set handed over (heavy arcs pending)
    if (f.arcs) listBits(B.donebits + wo, B.dsum + so, B.W2, dl, &dl_n);
    listBits(B.farbits + wo, B.fsum + so, B.W2, fl, &fl_n);
    blockSync();
    {
        unsigned long long mn = INF;
        for (int i = threadIdx.x; i < fl_n; i += blockDim.x) {
            const uint64_t v = dGetAcq(B, s, fl[i]);
            if (v >= hi && v < mn) mn = v;
        }
        waveMinLds(mn, &fmin);
    }
    blockSync();
    // the arc rule of this source's passes (arcs: see TinyF)
    auto rule = [&](int arcs, int32_t* nq) {
        return TinyF{B, f.fa, f.w, f.delta, arcs, f.scope, f.unit, f.mid, s, hi, nq, fl, &nq_n, &fl_n, &fmin};
    };
    if (dl_n > 0) {  // heavy arcs of the vertices already processed in this bucket
        if (threadIdx.x == 0) nq_n = 0;
        __syncthreads();
        tinyRelax<U>(rule(1, B.qn[par ^ 1] + vo), fb, dl, dl_n);
        blockSync();
        for (int i = threadIdx.x; i < dl_n; i += blockDim.x) {
            const int32_t u = dl[i];
            B.donebits[wo + (u >> 6)] = 0;
            B.dsum[so + (u >> 12)] = 0;
        }
        blockSync();
    }
    const int arcs = f.arcs ? 2 : 0;
    // (an empty queue never ends a launch unless the source is settled: the
    // grid kernels read an empty queue as settled)
    for (int64_t r = 0; (r < max_rounds || n == 0) && n <= TINY_CAP; ++r) {
        int32_t* q = B.qn[par] + vo;
        int32_t* nq = B.qn[par ^ 1] + vo;
        if (threadIdx.x == 0) nq_n = 0;
        __syncthreads();
        tinyRelax<U>(rule(arcs, nq), fb, q, n);
        blockSync();
        const int m = nq_n;
        for (int i = threadIdx.x; i < m; i += blockDim.x) B.nearbits[wo + (nq[i] >> 6)] = 0;
        blockSync();
        if (m > 0) {  // the bucket took new vertices: another round inside it
            par ^= 1;
            n = m;
            continue;
        }
        // the bucket drained: the far minimum picks the next bucket
        const int nf = fl_n;
        const uint64_t mfar = fmin;
        if (mfar == INF) {  // no far work left: settled
            n = 0;
            break;
        }
        lo = (mfar / f.delta) * f.delta;
        hi = lo > INF - f.delta ? INF : lo + f.delta;
        // far entries inside [lo, hi) -> queue, below it -> settled; both
        // leave the far set (bit cleared); the rest stay in the list, and
        // their minimum is the next far minimum.  TINY_ILP entries per
        // thread at a time, their loads issued together.
        if (threadIdx.x == 0) {
            nq_n = 0;
            fl2_n = 0;
            fmin2 = INF;
        }
        __syncthreads();
        unsigned long long mn = INF;
        for (int i0 = threadIdx.x; i0 < nf; i0 += TINY_ILP * blockDim.x) {
            int32_t x[TINY_ILP];
            uint64_t v[TINY_ILP];
#pragma unroll
            for (int c = 0; c < TINY_ILP; ++c) {
                const int i = i0 + c * blockDim.x;
                x[c] = i < nf ? fl[i] : -1;
            }
#pragma unroll
            for (int c = 0; c < TINY_ILP; ++c) v[c] = x[c] >= 0 ? dGetAcq(B, s, x[c]) : 0;
#pragma unroll
            for (int c = 0; c < TINY_ILP; ++c) {
                if (x[c] < 0) continue;
                if (v[c] >= hi && v[c] != INF) {
                    fl2[atomicAdd(&fl2_n, 1)] = x[c];
                    if (v[c] < mn) mn = v[c];
                } else {
                    atomicAnd((unsigned long long*)&B.farbits[wo + (x[c] >> 6)], ~(1ULL << (x[c] & 63)));
                    if (v[c] >= lo && v[c] < hi) nq[atomicAdd(&nq_n, 1)] = x[c];
                }
            }
        }
        waveMinLds(mn, &fmin2);
        blockSync();
        if (threadIdx.x == 0) {
            fmin = fmin2;
            fl_n = fl2_n;
        }
        int32_t* t = fl;
        fl = fl2;
        fl2 = t;
        par ^= 1;
        n = nq_n;
        __syncthreads();
    }
    // hand the state back at a round boundary, the queue in parity par0
    if (par != par0) {
        const int32_t* q = B.qn[par] + vo;
        int32_t* d = B.qn[par0] + vo;
        for (int i = threadIdx.x; i < n; i += blockDim.x) d[i] = q[i];
    }
    if (threadIdx.x == 0) {
        *nearCount(B, par0, s) = n;
        *cp(B, C_LO, s) = (int64_t)lo;
        *cp(B, C_HI, s) = (int64_t)hi;
    }
}

// One slot of every source -> the gather area (contiguous, for one readback).
__global__ void kGatherSlot(int64_t* ctl, int slot) {
    const int s = threadIdx.x;
    if (s < SB) ctl[C_GATHER + s] = ctl[cIdx(slot, s)];
}

static void readSlot(girhip_graph* h, int64_t* ctl, int slot, int64_t* out) {
    launch(h, "sssp_gather", kGatherSlot, 1, 64, 0, ctl, slot);
    readCounters(h, ctl + C_GATHER, SB, out);
}

// 32-bit distances -> the caller's 64-bit rows (UINT32_MAX -> UINT64_MAX).
__global__ void kWiden(const uint32_t* __restrict__ in, int64_t n, uint64_t* __restrict__ out) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t v = in[i];
        out[i] = v == 0xffffffffu ? INF : (uint64_t)v;
    }
}

__global__ void kReachedArcsD(const uint64_t* __restrict__ dist, const int32_t* __restrict__ beg, int32_t V,
                              int64_t* out) {
    int64_t m = 0;
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < V; v += (int64_t)gridDim.x * blockDim.x)
        if (dist[v] != INF) m += beg[v + 1] - beg[v];
    blockAtomicAdd(m, out);
}

// Rounds per host readback while every queue is small (high-diameter
// graphs): all round kernels read their work sizes on the device, and a
// round with nothing left to do is a no-op.
constexpr int SSSP_CHUNK = 16;

// Batch state over the handle's scratch (s_* buffers, sized for SB sources
// by the caller); dist: k rows of V uint64 (d32 = false) or uint32 (d32 =
// true) distances.
static SsspBatch batchState(girhip_graph* h, int k, void* dist, bool d32) {
    Scratch& S = scratch(h);
    SsspBatch B;
    B.V = h->V;
    B.W = ((int64_t)h->V + 63) / 64;
    B.W2 = (B.W + 63) / 64;
    for (int p = 0; p < 2; ++p) B.qn[p] = S.s_q[p].p;
    B.qh = S.s_qh.p;
    B.nearbits = S.s_near.p;
    B.farbits = S.s_far.p;
    B.donebits = S.s_done.p;
    B.nsum = S.s_nsum.p;
    B.fsum = S.s_fsum.p;
    B.dsum = S.s_dsum.p;
    B.ctl = S.s_ctl.p;
    B.dist = dist;
    B.d32 = d32 ? 1 : 0;
    B.xcd = h->prm.xcd != 0 ? 1 : 0;
    B.k = k;
    return B;
}

static void clearBitmaps(girhip_graph* h, const SsspBatch& B) {
    fillU64(h, B.nearbits, 0, (int64_t)B.k * B.W);
    fillU64(h, B.farbits, 0, (int64_t)B.k * B.W);
    fillU64(h, B.donebits, 0, (int64_t)B.k * B.W);
    fillU64(h, B.nsum, 0, (int64_t)B.k * B.W2);
    fillU64(h, B.fsum, 0, (int64_t)B.k * B.W2);
    fillU64(h, B.dsum, 0, (int64_t)B.k * B.W2);
}

using SplitView = girhip_graph::SplitView;

// The split view of width delta (null when the split knob is off); built on
// first use and kept for the two widths used last.  Its build time is
// reported as "sssp_split".
static const SplitView* splitView(girhip_graph* h, uint64_t delta) {
    if (h->prm.split == 0 || h->E == 0) return nullptr;
    for (auto& sv : h->split)
        if (sv.delta == delta) {
            sv.used = ++h->split_clock;
            return &sv;
        }
    SplitView& sv = h->split[0].used <= h->split[1].used ? h->split[0] : h->split[1];
    sv.delta = 0;
    const int32_t V = h->V;
    sv.adj.ensure((size_t)h->E);
    sv.w.ensure((size_t)h->E);
    sv.mid.ensure((size_t)(V > 0 ? V : 1));
    DevBuf<int32_t> longq;
    longq.ensure((size_t)(V > 0 ? V : 1));
    fillI64(h, h->dcnt.p, 0, 1);
    const SplitArgs A{h->fw_beg.p, h->fw_adj.p, h->has_w ? h->w_fw.p : nullptr, V, delta, sv.adj.p, sv.w.p,
                      sv.mid.p, longq.p, h->dcnt.p};
    tBegin(h);
    launch(h, "sssp_split_short", kSplitShort, boundedGrid(V, 256), 256, 0, A);
    launch(h, "sssp_split_long", kSplitLong, 2048, 256, 0, A);
    tEnd(h, "sssp_split");
    sv.delta = delta;
    sv.used = ++h->split_clock;
    return &sv;
}

// Bucket rounds from the state in B until every source's queue is empty
// after an advance (all settled).  scope / unit: see RelaxF; sv: the split
// view of delta, or null (rows classify their arcs by weight on the fly).
static void ssspRounds(girhip_graph* h, const SsspBatch& B, uint64_t delta, const int32_t* scope, bool unit,
                       const SplitView* sv) {
    Scratch& S = scratch(h);
    const int32_t V = h->V;
    const int k = B.k;
    const int blk = (int)h->prm.block;
    if (unit) sv = nullptr;
    const int32_t* fa = sv ? sv->adj.p : h->fw_adj.p;
    const uint32_t* w = sv ? sv->w.p : (h->has_w ? h->w_fw.p : nullptr);
    const int32_t* mid = sv ? sv->mid.p : nullptr;
    // heavy arcs exist only if some weight exceeds delta
    const int heavy = !unit && (uint64_t)maxWeight(h) > delta ? 1 : 0;
    const int u = unit ? 1 : 0;
    const int ur = (int)h->prm.sssp_unroll;
    const RelaxF fl[2] = {{B, fa, w, delta, 0, 0, scope, u, mid}, {B, fa, w, delta, 0, 1, scope, u, mid}};
    const RelaxF fh[2] = {{B, fa, w, delta, 1, 0, scope, u, mid}, {B, fa, w, delta, 1, 1, scope, u, mid}};
    // relaxation kernels by arcs per lane at a time
    auto relaxK = ur == 4 ? kBRelax<4> : ur == 2 ? kBRelax<2> : kBRelax<1>;
    auto relaxHK = ur == 4 ? kBRelaxHeavy<4> : ur == 2 ? kBRelaxHeavy<2> : kBRelaxHeavy<1>;
    const unsigned hgrid = std::min<unsigned>(gridFor(V, blk), 4096);
    const int32_t* fb = h->fw_beg.p;
    int par = 0;
    std::vector<int64_t> r(SB);
    int64_t maxn = 1;
    int R = 1;  // rounds in the next chunk
    const int cmax = std::max(1, std::min(SSSP_CHUNK, (int)h->prm.chunk));
    // one wave per summary word (dense and sparse bitmaps alike: the passes
    // are latency-bound, so they want many waves)
    const int G = 1;
    const unsigned sgrid = summaryGrid(B.W2, G, 2, blk);
    // persistent rounds need every row below the heavy-list threshold
    bool persist_ok = h->prm.persist != 0 && maxDegree(h) <= HEAVY_T;
    const int64_t cap = std::max<int64_t>(1, (int64_t)V / 64);  // queues up to this size stay persistent
    // after two full chunks, full chunks replay captured graphs (a run of
    // many small rounds: high-diameter graphs, small delta)
    const bool graphs = h->prm.graph != 0 && h->prm.profile == 0 && cmax > 1;
    GraphCache gc(h);
    int full = 0;
    // small frontiers without long rows: the one-workgroup-per-source engine
    const bool tiny_ok = h->prm.sssp_tiny != 0 && maxDegree(h) <= TINY_DEG;
    // (one workgroup per CU leaves registers to spare: 4 arcs per lane)
    auto tinyK = kSsspTiny<4>;
    while (true) {
        if (tiny_ok && maxn <= TINY_ENTER) {
            const size_t kv = (size_t)k * (size_t)std::max<int32_t>(V, 1);
            S.s_fl[0].ensure(kv);
            S.s_fl[1].ensure(kv);
            const TinyF tf{B, fa, w, delta, heavy, scope, u, mid, 0, 0, nullptr, nullptr, nullptr, nullptr, nullptr};
            launch(h, "sssp_tiny", tinyK, (unsigned)k, 1024, 0, tf, fb, par, (int64_t)1 << 20, S.s_fl[0].p,
                   S.s_fl[1].p);
            readSlot(h, S.s_ctl.p, par ? C_NEAR1 : C_NEAR0, r.data());
            int64_t mx = 0;
            for (int s = 0; s < k; ++s) mx = std::max(mx, r[s]);
            if (mx == 0) break;
            maxn = mx;
            R = 1;
            if (mx <= TINY_ENTER) continue;  // (the launch's round limit passed)
        }
        if (persist_ok && maxn <= cap) {
            int64_t* io = S.s_ctl.p + C_SLOTS;
            fillI64(h, io, par, 1);
            if (launchPersistent(h, "sssp_persist", kSsspPersist, 256, 2, fl[0], fh[0], fb, heavyList(h), heavy, delta,
                                 4, cap, (int32_t)0x7fffffff, io)) {
                int64_t pr;
                readCounters(h, io, 1, &pr);
                checkGridBar(h);
                par = (int)pr;
                readSlot(h, S.s_ctl.p, par ? C_NEAR1 : C_NEAR0, r.data());
                int64_t mx = 0;
                for (int s = 0; s < k; ++s) mx = std::max(mx, r[s]);
                if (mx == 0) break;
                maxn = mx;
                R = 1;
                continue;
            }
            persist_ok = false;  // no cooperative launch here: host-driven rounds
        }
        unsigned rgrid = R > 1 ? std::min<unsigned>(1024, std::max<unsigned>(64, gridFor(4 * maxn, blk)))
                               : gridFor(maxn, blk);
        const bool replay = graphs && R == cmax && full >= 2;
        if (replay) {  // few distinct grids, so few captured sequences
            unsigned q = 64;
            while (q < rgrid) q <<= 1;
            rgrid = q;
        }
        // a chunk's later rounds may meet large queues (a drained bucket
        // hands the next one its far entries): enough waves to fill the chip
        const unsigned glater =
            std::max(rgrid, std::max(1u, std::min<unsigned>(gridFor(V, blk), (unsigned)std::max(1, (1 << 18) / blk)) /
                                             (unsigned)k));
        auto issue = [&](int p) {
            for (int t = 0; t < R; ++t) {
                launch(h, "sssp_relax", relaxK, (t == 0 ? rgrid : glater) * k, blk, 0, fl[p], fb, heavyList(h), heavy);
                launchHeavyStore(h, "sssp_relax_rows", fl[p]);
                launch(h, "sssp_compact", kBCompact, sgrid * k, blk, 0, B, p, heavy, G);
                if (heavy) {
                    launch(h, "sssp_relax_heavy", relaxHK, hgrid * k, blk, 0, fh[p], fb, heavyList(h));
                    launchHeavyStore(h, "sssp_relax_heavy_rows", fh[p]);
                    launch(h, "sssp_farmin", kBFarMin, sgrid * k, blk, 0, B, p, G);
                }
                launch(h, "sssp_split", kBSplit, sgrid * k, blk, 0, B, p, delta, G);
                p ^= 1;
            }
        };
        if (replay) gc.run(((uint64_t)rgrid << 8) | ((uint64_t)R << 1) | (uint64_t)par, [&] { issue(par); });
        else issue(par);
        if (R == cmax) ++full;
        par ^= R & 1;
        readSlot(h, S.s_ctl.p, par ? C_NEAR1 : C_NEAR0, r.data());
        int64_t mx = 0;
        for (int s = 0; s < k; ++s) mx = std::max(mx, r[s]);
        if (mx == 0) break;  // every queue empty after the advance: all sources settled
        // queues small and not growing fast: longer chunks, else one round
        R = (mx * 64 < (int64_t)V && mx <= 2 * maxn && maxn > 1) ? std::min(cmax, 2 * R) : 1;
        maxn = mx;
    }
}

// k <= SB sources (device array dsrc) into rows of dist (see batchState).
static void ssspBatch(girhip_graph* h, const int32_t* dsrc, int k, uint64_t delta, void* dist, bool d32,
                      const SplitView* sv) {
    const SsspBatch B = batchState(h, k, dist, d32);
    if (d32) fillI32(h, (int32_t*)dist, -1, (int64_t)k * h->V);
    else fillU64(h, (uint64_t*)dist, INF, (int64_t)k * h->V);
    clearBitmaps(h, B);
    launch(h, "sssp_init", kBInit, 1, 64, 0, B, dsrc, delta);
    ssspRounds(h, B, delta, nullptr, false, sv);
}

// ---- condensation-guided queries --------------------------------------------
// GIR::reachableSccs / reachInfo (gir_query.cpp:56-88) and the two-level
// model of GIR::dagWavefront / bfsTwoLevel / ssspTwoLevel
// (gir_query.cpp:304-350, 604-620, 705-722).
//
// The reachable part of the condensation is found by a level sweep over the
// DAG CSR from the source's SCC: a claim bitmap over SCC ids and an
// append-only list whose level segments are the frontiers.  Each reachable
// SCC's pending count is its number of reachable DAG predecessors.  The
// wavefront then releases SCCs in Kahn order over that part, all ready SCCs
// of a wave at once:
//   1. their members go to a vertex list; members of multi-vertex SCCs that
//      already hold a distance become the seeds (far bits) of one
//      delta-stepping run restricted to arcs inside an SCC, which finishes
//      every ready SCC of the wave together (a single-vertex SCC is final
//      once its predecessors are);
//   2. the members' arcs into other SCCs are relaxed once (atomicMin);
//   3. every DAG successor of a ready SCC loses one pending predecessor;
//      those reaching zero form the next wave.
// BFS levels are the same computation with unit weights and delta 1.

__global__ void kMemKeys(const int32_t* __restrict__ scc, int32_t V, int B, uint64_t* __restrict__ keys) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < V) keys[i] = ((uint64_t)(uint32_t)scc[i] << B) | (uint64_t)i;
}

// sorted (scc, v) keys -> member lists: mem[beg[c] .. beg[c + 1]) = SCC c's vertices
__global__ void kMemCsr(const uint64_t* __restrict__ keys, int32_t V, int B, int32_t* __restrict__ beg,
                        int32_t* __restrict__ mem) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < V) mem[i] = (int32_t)(keys[i] & ((1ULL << B) - 1));
    if (i <= V) {
        const uint64_t key = (uint64_t)i << B;
        int64_t lo = 0, hi = V;
        while (lo < hi) {
            const int64_t m = lo + ((hi - lo) >> 1);
            if (keys[m] < key) lo = m + 1;
            else hi = m;
        }
        beg[i] = (int32_t)lo;
    }
}

static void memberCsr(girhip_graph* h) {
    if (h->mem_ok) return;
    const int32_t V = h->V;
    const int blk = (int)h->prm.block;
    const int B = bitsFor((uint64_t)(V > 1 ? V - 1 : 1));
    Scratch& S = scratch(h);
    S.keys[0].ensure((size_t)std::max<int32_t>(V, 1));
    h->mem_beg.ensure((size_t)V + 1);
    h->mem_v.ensure((size_t)std::max<int32_t>(V, 1));
    launch(h, "tl_mem_keys", kMemKeys, gridFor(V, blk), blk, 0, (const int32_t*)h->scc.p, V, B, S.keys[0].p);
    radixSort(h, S.keys[0].p, nullptr, V, 2 * B);
    launch(h, "tl_mem_csr", kMemCsr, gridFor((int64_t)V + 1, blk), blk, 0, (const uint64_t*)S.keys[0].p, V, B,
           h->mem_beg.p, h->mem_v.p);
    h->mem_ok = true;
}

// DAG arc j: claim its head in the bitmap, list it once.
struct DagClaimF {
    const int32_t* dst;
    uint64_t* mask;
    int32_t* list;
    int64_t* tail;
    __device__ int64_t operator()(int32_t, uint64_t, int32_t, int64_t j) const {
        const int32_t z = dst[j];
        const uint64_t bit = 1ULL << (z & 63);
        bool nw = false;
        if (!(mask[z >> 6] & bit))  // plain read: bits only get set
            nw = !(atomicOr((unsigned long long*)&mask[z >> 6], (unsigned long long)bit) & bit);
        const int64_t pos = wavePush(tail, nw);
        if (nw) list[pos] = z;
        return 0;
    }
};

// DAG arc j of a reachable SCC: one more reachable predecessor of its head.
struct PendF {
    const int32_t* dst;
    int32_t* left;
    __device__ int64_t operator()(int32_t, uint64_t, int32_t, int64_t j) const {
        atomicAdd(&left[dst[j]], 1);
        return 0;
    }
};

// Sweep seed: the SCC of vertex src (canonical id scc[src]).
__global__ void kSweepSeed(const int32_t* __restrict__ scc, int32_t src, uint64_t* mask, int32_t* list, int64_t* ctl) {
    if (threadIdx.x != 0) return;
    const int32_t c = scc[src];
    mask[c >> 6] |= 1ULL << (c & 63);
    list[0] = c;
    for (int j = 0; j < T_SLOTS; ++j) ctl[j] = 0;
    ctl[T_HI] = 1;
    ctl[T_TAIL] = 1;
}

// Vertices in reachable SCCs and their out-arcs.
__global__ void kReachVerts(const int32_t* __restrict__ scc, const int32_t* __restrict__ fb,
                            const uint64_t* __restrict__ mask, int32_t V, int64_t* ctl) {
    int64_t nv = 0, ne = 0;
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < V; v += (int64_t)gridDim.x * blockDim.x) {
        const int32_t c = scc[v];
        if ((mask[c >> 6] >> (c & 63)) & 1) {
            ++nv;
            ne += fb[v + 1] - fb[v];
        }
    }
    blockAtomicAdd(nv, ctl + T_VERTS);
    blockAtomicAdd(ne, ctl + T_EDGES);
}

__global__ void kMaskBytes(const uint64_t* __restrict__ mask, int32_t V, uint8_t* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < V) out[i] = (uint8_t)((mask[i >> 6] >> (i & 63)) & 1);
}

// Reachable part of the condensation from the SCC of vertex src: bitmap
// tl_mask, list tl_list[0 .. n_sccs), pending counts tl_left (zero outside).
// Returns n_sccs.
static int64_t reachSweep(girhip_graph* h, int32_t src) {
    const int32_t V = h->V;
    const int blk = (int)h->prm.block;
    Scratch& S = scratch(h);
    const int64_t W = ((int64_t)V + 63) / 64;
    S.tl_mask.ensure((size_t)std::max<int64_t>(W, 1));
    S.tl_list.ensure((size_t)std::max<int32_t>(V, 1));
    S.tl_left.ensure((size_t)std::max<int32_t>(V, 1));
    S.tl_ctl.ensure(2 * T_SLOTS);
    int64_t* ctl = S.tl_ctl.p;
    fillU64(h, S.tl_mask.p, 0, W);
    fillI32(h, S.tl_left.p, 0, V);
    launch(h, "tl_sweep_seed", kSweepSeed, 1, 64, 0, (const int32_t*)h->scc.p, src, S.tl_mask.p, S.tl_list.p, ctl);
    const DagClaimF f{h->dag_dst.p, S.tl_mask.p, S.tl_list.p, ctl + T_TAIL};
    const unsigned g = 512;
    int64_t r[T_SLOTS];
    while (true) {
        for (int t = 0; t < 16; ++t) {  // levels per readback
            launch(h, "tl_sweep", kListRows<DagClaimF>, g, blk, 0, (const int32_t*)h->dag_beg.p,
                   (const int32_t*)S.tl_list.p, (const int64_t*)ctl, f, heavyList(h));
            launchHeavy(h, "tl_sweep_heavy", f);
            launch(h, "tl_advance", kListAdvance, 1, 64, 0, ctl);
        }
        readCounters(h, ctl, T_SLOTS, r);
        if (r[T_LO] == r[T_HI]) break;
    }
    launch(h, "tl_all", kListAll, 1, 64, 0, ctl);
    const PendF pf{h->dag_dst.p, S.tl_left.p};
    launch(h, "tl_pending", kListRows<PendF>, boundedGrid(r[T_TAIL], blk), blk, 0,
           (const int32_t*)h->dag_beg.p, (const int32_t*)S.tl_list.p, (const int64_t*)ctl, pf, heavyList(h));
    launchHeavy(h, "tl_pending_heavy", pf);
    return r[T_TAIL];
}

// Wave step 1: member v of a ready SCC (payload: the SCC's size) joins the
// vertex list; in a multi-vertex SCC a member with a distance is a seed.
struct ExpandF {
    const int32_t* mem;
    int32_t* vl;
    int64_t* nv;
    SsspBatch B;
    __device__ int64_t operator()(int32_t, uint64_t size, int32_t, int64_t j) const {
        const int32_t v = mem[j];
        const int64_t pos = wavePush(nv, true);
        vl[pos] = v;
        if (size > 1 && dGet(B, 0, v) != INF) setBit(B.farbits, B.fsum, v);
        return 0;
    }
};

// Wave step 2: arc j of vertex p (distance dp, SCC label sp) into another SCC.
struct CrossF {
    SsspBatch B;
    const int32_t* fa;
    const uint32_t* w;
    const int32_t* scc;
    int unit;
    __device__ int64_t operator()(int32_t p, uint64_t dp, int32_t sp, int64_t j) const {
        const int32_t x = fa[j];
        if (scc[x] == sp) return 0;
        const uint64_t wt = unit ? 1 : (w ? (uint64_t)w[j] : (uint64_t)hashWeight(p, x));
        const uint64_t nd = dp + wt;
        if (B.d32 && nd >= 0xffffffffull) {
            B.ctl[C_OVF] = 1;
            return 0;
        }
        if (nd < dGet(B, 0, x)) dMin(B, 0, x, nd);
        return 0;
    }
};

__global__ void kCross(CrossF f, const int32_t* __restrict__ fb, const int32_t* __restrict__ vl,
                       const int64_t* __restrict__ ctl, HeavyList H) {
    const int64_t n = ctl[T_NV];
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t base = (int64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~(WAVE - 1)); base < n; base += stride) {
        const int64_t i = base + laneId();
        bool act = i < n;
        const int32_t v = act ? vl[i] : 0;
        const uint64_t dv = act ? dGet(f.B, 0, v) : INF;
        act = act && dv != INF;
        const int64_t b = act ? fb[v] : 0, e = act ? fb[v + 1] : 0;
        waveRowsH(act, v, dv, act ? f.scc[v] : 0, b, e, H, f);
    }
}

// Wave step 3: DAG arc j of a ready SCC; the head is ready once all its
// reachable predecessors are.
struct ReleaseF {
    const int32_t* dst;
    const int32_t* mbeg;
    int32_t* left;
    int32_t* rdy;
    int64_t* ctl;
    __device__ int64_t operator()(int32_t, uint64_t, int32_t, int64_t j) const {
        const int32_t d = dst[j];
        const bool go = atomicSub(&left[d], 1) == 1;
        const int64_t pos = wavePush(ctl + T_TAIL, go);
        if (go) {
            rdy[pos] = d;
            if (mbeg[d + 1] - mbeg[d] > 1) atomicAdd((unsigned long long*)(ctl + T_NBIG), 1ULL);
        }
        return 0;
    }
};

// Wave state: distances (source vertex at 0), the source's SCC as the first
// wave; T_NBIG: multi-vertex SCCs in the current wave.
__global__ void kWaveInit(SsspBatch B, const int32_t* __restrict__ scc, const int32_t* __restrict__ mbeg, int32_t src,
                          int32_t* rdy, int64_t* ctl) {
    if (threadIdx.x != 0) return;
    if (B.d32) ((uint32_t*)B.dist)[src] = 0;
    else ((uint64_t*)B.dist)[src] = 0;
    const int32_t c = scc[src];
    rdy[0] = c;
    for (int j = 0; j < T_SLOTS; ++j) ctl[j] = 0;
    ctl[T_HI] = 1;
    ctl[T_TAIL] = 1;
    ctl[T_NBIG] = mbeg[c + 1] - mbeg[c] > 1 ? 1 : 0;
    B.ctl[C_OVF] = 0;
}

// Delta-stepping state for a run seeded through the far bitmap: every slot
// zero, so the first far minimum (over distances >= hi = 0) covers every seed.
__global__ void kBInitSeeded(SsspBatch B) {
    const int s = threadIdx.x;
    if (s >= SB) return;
    for (int j = 0; j < C_NSLOT; ++j) *cp(B, j, s) = 0;
}

__global__ void kZeroSlot(int64_t* p) {
    if (threadIdx.x == 0) *p = 0;
}

// One source's two-level query into dist (uint32 when d32); the pending
// counts of the sweep from src's SCC must be in tl_left.  unit: BFS levels.
static void waveQuery(girhip_graph* h, int32_t src, uint64_t delta, bool unit, void* dist, bool d32,
                      const SplitView* sv) {
    const int32_t V = h->V;
    const int blk = (int)h->prm.block;
    Scratch& S = scratch(h);
    S.tl_rdy.ensure((size_t)std::max<int32_t>(V, 1));
    S.tl_vl.ensure((size_t)std::max<int32_t>(V, 1));
    int64_t* ctl = S.tl_ctl.p + T_SLOTS;
    const SsspBatch B = batchState(h, 1, dist, d32);
    if (d32) fillI32(h, (int32_t*)dist, -1, V);
    else fillU64(h, (uint64_t*)dist, INF, V);
    launch(h, "tl_wave_init", kWaveInit, 1, 64, 0, B, (const int32_t*)h->scc.p, (const int32_t*)h->mem_beg.p, src,
           S.tl_rdy.p, ctl);
    const uint32_t* w = h->has_w ? h->w_fw.p : nullptr;
    const ExpandF ef{h->mem_v.p, S.tl_vl.p, ctl + T_NV, B};
    const CrossF cf{B, h->fw_adj.p, w, h->scc.p, unit ? 1 : 0};
    const ReleaseF rf{h->dag_dst.p, h->mem_beg.p, S.tl_left.p, S.tl_rdy.p, ctl};
    int64_t r[T_SLOTS];
    int64_t nready = 1, nbig = 0;
    readCounters(h, ctl, T_SLOTS, r);
    nbig = r[T_NBIG];
    while (nready > 0) {
        if (nbig > 0) clearBitmaps(h, B);
        launch(h, "tl_expand", kListRows<ExpandF>, std::min<unsigned>(1024, gridFor(nready, blk)), blk,
               0, (const int32_t*)h->mem_beg.p, (const int32_t*)S.tl_rdy.p, (const int64_t*)ctl, ef, heavyList(h));
        launchHeavy(h, "tl_expand_heavy", ef);
        if (nbig > 0) {
            launch(h, "tl_seeded", kBInitSeeded, 1, 64, 0, B);
            ssspRounds(h, B, delta, h->scc.p, unit, sv);
        }
        launch(h, "tl_cross", kCross, 1024, blk, 0, cf, (const int32_t*)h->fw_beg.p, (const int32_t*)S.tl_vl.p,
               (const int64_t*)ctl, heavyList(h));
        launchHeavyStore(h, "tl_cross_heavy", cf);
        launch(h, "tl_zero", kZeroSlot, 1, 64, 0, ctl + T_NBIG);
        launch(h, "tl_release", kListRows<ReleaseF>, std::min<unsigned>(1024, gridFor(nready, blk)),
               blk, 0, (const int32_t*)h->dag_beg.p, (const int32_t*)S.tl_rdy.p, (const int64_t*)ctl, rf,
               heavyList(h));
        launchHeavy(h, "tl_release_heavy", rf);
        launch(h, "tl_advance", kListAdvance, 1, 64, 0, ctl);
        readCounters(h, ctl, T_SLOTS, r);
        nready = r[T_HI] - r[T_LO];
        nbig = r[T_NBIG];
    }
}

__global__ void kLevels32(const uint32_t* __restrict__ d, int64_t n, int32_t* __restrict__ lv) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        lv[i] = d[i] == 0xffffffffu ? -1 : (int32_t)d[i];
}

// Scratch of one SSSP batch of one source.
static void ensureBatchScratch(girhip_graph* h, int kc) {
    Scratch& S = scratch(h);
    const size_t Vs = (size_t)(h->V > 0 ? h->V : 1), Ws = (Vs + 63) / 64, W2s = (Ws + 63) / 64;
    for (int p = 0; p < 2; ++p) S.s_q[p].ensure((size_t)kc * Vs);
    S.s_qh.ensure((size_t)kc * Vs);
    S.s_near.ensure((size_t)kc * Ws);
    S.s_far.ensure((size_t)kc * Ws);
    S.s_done.ensure((size_t)kc * Ws);
    S.s_nsum.ensure((size_t)kc * W2s);
    S.s_fsum.ensure((size_t)kc * W2s);
    S.s_dsum.ensure((size_t)kc * W2s);
    S.s_ctl.ensure(C_TOTAL);
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
        S.src.ensure((size_t)(k > 0 ? k : 1));
        const int32_t* dsrc = S.src.p;
        copyIn(h, S.src.p, s.data(), sizeof(int32_t) * (size_t)k);
        // 32-bit distances unless weights alone could overflow them
        const bool try32 = h->prm.sssp_dist32 != 0 && (uint64_t)maxWeight(h) < 0x7fffffffull;
        // the split view separates light from heavy arcs; with no arc
        // heavier than delta the rows are all light as they stand
        const SplitView* sv = (k > 0 && (uint64_t)maxWeight(h) > delta) ? splitView(h, delta) : nullptr;
        // Engine: the interleaved one shares arcs between sources whose
        // distances fall in the same bucket, which pays once buckets are
        // wide; with narrow buckets, many sources and a heavy phase per
        // bucket the per-source layout measured faster.  RMAT-22, ms,
        // per-source vs interleaved: weights [1, 64], 8 sources: delta 16
        // 24.9 vs 19.3, delta 32 27.4 vs 15.1; 16 sources: delta 16 41.4 vs
        // 49.7, delta 32 47.9 vs 34.8; hash weights [1, 100], 8 sources:
        // delta 16 23.2 vs 30.3, delta 32 27.4 vs 26.9; one source: the
        // interleaved engine at every delta.  Auto (2): interleaved for one
        // source (unless its rows all fit the one-workgroup rounds, no row
        // above TINY_DEG arcs, and delta is small against the weights: a
        // lattice's many small buckets), for up to 8 sources when 4 delta >=
        // the largest weight, and for more when 2 delta >= it.
        const double mode = h->prm.sssp_interleave;
        const uint64_t mw = (uint64_t)maxWeight(h);
        const bool tiny_fit = h->prm.sssp_tiny != 0 && maxDegree(h) <= TINY_DEG && 4 * delta < mw;
        const bool inter = mode == 1 || (mode == 2 && (k == 1 ? !tiny_fit : k <= 8 ? 4 * delta >= mw
                                                                                   : 2 * delta >= mw));
        if (inter) {
            // batches of up to sssp_ms_batch sources sharing buckets, distances
            // interleaved; the batches are made even (10 -> 5 + 5, not 8 + 2)
            tBegin(h);
            const int mb = std::max(1, std::min(MS_MAX, (int)h->prm.sssp_ms_batch));
            const int nb = (k + mb - 1) / mb;
            int32_t i0 = 0;
            while (i0 < k) {
                const int c = std::min<int32_t>((k + nb - 1) / nb, k - i0);
                msBatch(h, dsrc + i0, c, delta, sv, out + (size_t)i0 * V, try32);
                i0 += c;
            }
        } else {
            // one layout per source (the engine below)
            const int kc = std::max(1, std::min<int32_t>(k, SB));
            for (int p = 0; p < 2; ++p) S.s_q[p].ensure((size_t)kc * Vs);
            S.s_qh.ensure((size_t)kc * Vs);
            S.s_near.ensure((size_t)kc * Ws);
            S.s_far.ensure((size_t)kc * Ws);
            S.s_done.ensure((size_t)kc * Ws);
            const size_t W2s = (Ws + 63) / 64;
            S.s_nsum.ensure((size_t)kc * W2s);
            S.s_fsum.ensure((size_t)kc * W2s);
            S.s_dsum.ensure((size_t)kc * W2s);
            S.s_ctl.ensure(C_TOTAL);  // slot lines, the persistent kernel's parity, the overflow flag, the gather area
            if (try32) S.s_d32.ensure((size_t)kc * Vs);
            tBegin(h);
            for (int32_t i0 = 0; i0 < k; i0 += SB) {
                const int c = std::min<int32_t>(SB, k - i0);
                uint64_t* rows = out + (size_t)i0 * V;
                if (try32) {
                    ssspBatch(h, dsrc + i0, c, delta, S.s_d32.p, true, sv);
                    int64_t ovf;
                    readCounters(h, S.s_ctl.p + C_OVF, 1, &ovf);
                    if (ovf == 0) {
                        launch(h, "sssp_widen", kWiden, boundedGrid((int64_t)c * V, 256), 256, 0,
                               (const uint32_t*)S.s_d32.p, (int64_t)c * V, rows);
                        continue;
                    }
                }
                ssspBatch(h, dsrc + i0, c, delta, rows, false, sv);
            }
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

// Shared checks of the condensation-guided entry points: SCC layer built,
// pending ingested changes applied first (GIR::bfsTwoLevel / ssspTwoLevel
// call sync(), gir_query.cpp:605, 706).
static void twoLevelReady(girhip_graph* h, const char* what) {
    if (!h->scc_req) raise(GIRHIP_ESTATE, std::string(what) + ": call girhip_scc_dag first");
    if (h->pending && girhip_sync(h) != GIRHIP_OK) raise(GIRHIP_EHIP, girhip_last_error());
}

extern "C" int girhip_reach_info(girhip_graph* h, int32_t scc_id, uint8_t* mask, int32_t* pending, int64_t* n_sccs,
                                 int64_t* n_verts, int64_t* n_edges) {
    return guarded([&] {
        useDevice(h);
        twoLevelReady(h, "reach_info");
        const int32_t V = h->V;
        if (scc_id < 0 || scc_id >= V) raise(GIRHIP_EINVAL, "reach_info: SCC id out of range");
        int32_t c = -1;
        HIPCHK(hipMemcpyAsync(&c, h->scc.p + scc_id, sizeof(int32_t), hipMemcpyDeviceToHost, h->st));
        syncStream(h);
        if (c != scc_id) raise(GIRHIP_EINVAL, "reach_info: not an SCC id (ids are the minimum vertex of each SCC)");
        tBegin(h);
        const int64_t ns = reachSweep(h, scc_id);
        Scratch& S = scratch(h);
        int64_t* ctl = S.tl_ctl.p;
        launch(h, "tl_verts", kReachVerts, boundedGrid(V, 256), 256, 0, (const int32_t*)h->scc.p,
               (const int32_t*)h->fw_beg.p, (const uint64_t*)S.tl_mask.p, V, ctl);
        int64_t r[T_SLOTS];
        readCounters(h, ctl, T_SLOTS, r);
        tEnd(h, "reach_info");
        if (mask) {
            uint8_t* out = mask;
            const bool direct = isDevicePtr(mask);
            if (!direct) {
                S.stage.ensure((size_t)std::max<int32_t>(V, 1));
                out = S.stage.p;
            }
            launch(h, "tl_mask_bytes", kMaskBytes, gridFor(V, 256), 256, 0, (const uint64_t*)S.tl_mask.p, V, out);
            if (!direct) copyOut(h, mask, out, (size_t)V);
        }
        if (pending) copyOut(h, pending, S.tl_left.p, sizeof(int32_t) * (size_t)V);
        syncStream(h);
        if (n_sccs) *n_sccs = ns;
        if (n_verts) *n_verts = r[T_VERTS];
        if (n_edges) *n_edges = r[T_EDGES];
    });
}

// k sources, one after another: sweep from the source's SCC, then the
// wavefront; 32-bit distances first (64-bit redo on overflow).  unit: BFS
// levels into lv (int32, -1 = unreached); else distances into dist.
static void twoLevelMany(girhip_graph* h, const int32_t* src, int32_t k, uint64_t delta, bool unit, int32_t* lv,
                         uint64_t* dist, const char* what) {
    const int32_t V = h->V;
    std::vector<int32_t> s = hostSources(h, src, k);
    Scratch& S = scratch(h);
    const size_t n = (size_t)k * (size_t)V;
    void* outp = unit ? (void*)lv : (void*)dist;
    const size_t esz = unit ? sizeof(int32_t) : sizeof(uint64_t);
    const bool direct = outp && isDevicePtr(outp);
    uint8_t* out = (uint8_t*)outp;
    if (!direct) {
        S.stage.ensure(esz * (n > 0 ? n : 1));
        out = S.stage.p;
    }
    ensureBatchScratch(h, 1);
    const bool try32 = unit || (h->prm.sssp_dist32 != 0 && (uint64_t)maxWeight(h) < 0x7fffffffull);
    if (try32) S.s_d32.ensure((size_t)std::max<int32_t>(V, 1));
    const SplitView* sv = unit || k == 0 ? nullptr : splitView(h, delta);
    tBegin(h);
    if (V > 0) memberCsr(h);
    for (int32_t i = 0; i < k; ++i) {
        uint8_t* row = out + esz * (size_t)i * V;
        const int32_t x = s[i];
        if (x < 0 || x >= V) {  // invalid source: nothing reached
            if (unit) fillI32(h, (int32_t*)row, -1, V);
            else fillU64(h, (uint64_t*)row, INF, V);
            continue;
        }
        reachSweep(h, x);
        if (try32) {
            waveQuery(h, x, delta, unit, S.s_d32.p, true, sv);
            int64_t ovf;
            readCounters(h, S.s_ctl.p + C_OVF, 1, &ovf);
            if (ovf == 0) {
                if (unit)
                    launch(h, "tl_levels", kLevels32, boundedGrid(V, 256), 256, 0, (const uint32_t*)S.s_d32.p,
                           (int64_t)V, (int32_t*)row);
                else
                    launch(h, "sssp_widen", kWiden, boundedGrid(V, 256), 256, 0, (const uint32_t*)S.s_d32.p,
                           (int64_t)V, (uint64_t*)row);
                continue;
            }
            // the pending counts were used up by the wavefront
            reachSweep(h, x);
        }
        waveQuery(h, x, delta, false, row, false, sv);
    }
    tEnd(h, what);
    if (outp && !direct) copyOut(h, outp, out, esz * n);
    syncStream(h);
}

extern "C" int girhip_bfs_two_level(girhip_graph* h, const int32_t* src, int32_t k, int32_t* levels) {
    return guarded([&] {
        useDevice(h);
        if (k < 0 || (k > 0 && !src)) raise(GIRHIP_EINVAL, "bfs_two_level: bad source list");
        twoLevelReady(h, "bfs_two_level");
        twoLevelMany(h, src, k, 1, true, levels, nullptr, "bfs_two_level");
    });
}

extern "C" int girhip_sssp_two_level(girhip_graph* h, const int32_t* src, int32_t k, uint64_t delta,
                                     uint64_t* dist) {
    return guarded([&] {
        useDevice(h);
        if (k < 0 || (k > 0 && !src)) raise(GIRHIP_EINVAL, "sssp_two_level: bad source list");
        if (delta == 0) delta = (uint64_t)h->prm.sssp_delta;
        if (delta == 0) raise(GIRHIP_EINVAL, "sssp_two_level: delta must be > 0");
        twoLevelReady(h, "sssp_two_level");
        twoLevelMany(h, src, k, delta, false, nullptr, dist, "sssp_two_level");
    });
}

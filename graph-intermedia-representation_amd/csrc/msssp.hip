// Multi-source delta-stepping with shared buckets: replaces GIR::sssp /
// GIR::deltaStepping (gir_query.cpp:629-703) for up to MS_MAX sources at
// once.
//
// The reference runs one source at a time.  Run side by side with a layout
// per source, k sources relax the same arc k times, each time reading the
// arc and one random distance line of their own.  Here the sources of a
// batch share everything but their distances:
//   - the distances of a vertex are adjacent, dist[x * KS + s] (KS = k
//     rounded up to a power of two: 32 bytes per vertex for 8 sources), so
//     one visit of arc (u -> x) reads the arc once and touches one distance
//     segment for every source active at u;
//   - buckets are common: every source processes bucket [lo, lo + delta)
//     at the same time, so a vertex whose distances from several sources
//     fall in the same bucket is visited once for all of them;
//   - the work sets are bitmaps (with a summary bit per 64-bit word): near
//     (entered the bucket this round), done (processed in this bucket) and
//     far (waiting in a later bucket).  No per-vertex source masks are
//     kept: a queued vertex is relaxed for every source whose distance lies
//     in the current bucket (read off its distance segment, which the
//     relaxation loads anyway), which covers every source that lowered it;
//     the far set's sources are likewise read off the distances (a finite
//     distance at or past the bucket's end is a far entry, since the
//     relaxation that set it marked the vertex).  A relaxation that lowers
//     a distance thus costs the atomicMin and one bitmap atomicOr.
// Round (same phases as the per-source engine in sssp.hip):
//   relax the queue's light arcs -> compact (near bitmap -> next queue; or,
//   when the bucket drained, done bitmap -> heavy queue, or the far minimum)
//   [-> relax heavy arcs once -> far minimum] -> split (the bucket of the
//   smallest far distance; far entries inside it -> next queue).
// A relaxation pre-checks the distance with a plain load and then issues an
// atomicMin.  32-bit distances first (a batch is redone in 64 bits
// if one would reach 2^32 - 1).
//
// HBM traffic per source (algorithmic, the same contract as the reference's
// per-source run): 8V + 8 n_s + 16 m_s.
#include <vector>

#include "common.h"
#include "msssp.h"

namespace gir {
namespace {

constexpr uint64_t INF = ~0ULL;

// Control slots, each on a 128-byte line of its own.
enum : int { M_NEAR0 = 0, M_NEAR1, M_NSET0, M_NSET1, M_LO, M_HI, M_FMIN, M_HQ, M_OVF, M_NSLOT };
constexpr int MSTR = 16;

struct MsState {
    void* dist;  // [V][KS] of T
    uint64_t *nbits, *fbits, *dbits;
    uint64_t *nsum, *fsum, *dsum;
    int32_t* q[2];
    int32_t* qh;
    int64_t* ctl;
    int64_t V, W, W2;
    __device__ int64_t* c(int slot) const { return ctl + slot * MSTR; }
};

template <class T>
__device__ __forceinline__ uint64_t widen(T v) {
    if constexpr (sizeof(T) == 4) return v == 0xffffffffu ? INF : (uint64_t)v;
    else return (uint64_t)v;
}

// Segments of at most 32 bytes are read with vector loads into registers
// (KS <= 8 for 32-bit distances); longer ones one distance at a time.
template <class T, int KS>
constexpr bool kVecSeg = KS * sizeof(T) <= 32;
template <int KS>
constexpr uint32_t kAllSources = KS >= 32 ? 0xffffffffu : (1u << KS) - 1;

template <class T, int KS>
__device__ __forceinline__ void loadSeg(const T* p, T (&v)[KS]) {
    constexpr int B = KS * (int)sizeof(T);
    if constexpr (B == 4) {
        const uint32_t x = *(const uint32_t*)p;
        __builtin_memcpy(v, &x, 4);
    } else if constexpr (B == 8) {
        const uint2 x = *(const uint2*)p;
        __builtin_memcpy(v, &x, 8);
    } else if constexpr (B == 16) {
        const uint4 x = *(const uint4*)p;
        __builtin_memcpy(v, &x, 16);
    } else {
        static_assert(B == 32, "segment of 4, 8, 16 or 32 bytes");
        const uint4 x = ((const uint4*)p)[0], y = ((const uint4*)p)[1];
        __builtin_memcpy(v, &x, 16);
        __builtin_memcpy((char*)v + 16, &y, 16);
    }
}

template <class T>
__device__ __forceinline__ void atomicMinT(T* p, T v) {
    if constexpr (sizeof(T) == 4) atomicMin((unsigned int*)p, (unsigned int)v);
    else atomicMin((unsigned long long*)p, (unsigned long long)v);
}

// Arc j of row u for the sources of mask m (pay).  heavy = 0: light arcs
// (w <= delta), heavy = 1: heavy arcs; with the split view (split != 0) the
// row range already holds one class only.
template <class T, int KS>
struct MsRelax {
    MsState S;
    const int32_t* fa;
    const uint32_t* w;  // per slot, or null: the reference's hash weights
    uint64_t delta;
    int heavy;
    int par;
    int split;
    // The vertex's bitmap bit is set behind a plain pre-check read (within a
    // kernel bits only grow, so a stale read costs at most a redundant
    // atomic).  Which sources improved is not recorded: a queued vertex is
    // relaxed for every source whose distance lies in the bucket.
    __device__ __forceinline__ void mark(int32_t x, uint32_t near, uint32_t far) const {
        // the near bitmap is empty at the start of every round: the first
        // bit of a word raises the round's flag
        if (near && setBit(S.nbits, S.nsum, x)) *S.c(par ? M_NSET1 : M_NSET0) = 1;
        if (far) setBit(S.fbits, S.fsum, x);
    }
    __device__ int64_t operator()(int32_t u, uint64_t pay, int32_t, int64_t j) const {
        const int32_t x = fa[j];
        const uint64_t wt = w ? (uint64_t)w[j] : (uint64_t)hashWeight(u, x);
        if (!split && ((wt > delta) != (heavy != 0))) return 0;
        const uint32_t m = (uint32_t)pay;
        const uint64_t hi = (uint64_t)*S.c(M_HI);
        T* dist = (T*)S.dist;
        T* px = dist + (size_t)x * KS;
        const T* pu = dist + (size_t)u * KS;
        uint32_t near = 0, far = 0;
        bool ovf = false;
        auto one = [&](int s, T du, T dx) {
            const uint64_t nd = widen(du) + wt;
            if constexpr (sizeof(T) == 4) {
                if (nd >= 0xffffffffull) {
                    ovf = true;
                    return;
                }
            }
            if (nd < widen(dx)) {
                // distances only fall, so the plain load was a safe pre-check;
                // the atomicMin's old value is not awaited: every lowering
                // relaxation passed the pre-check and marks x, and one that
                // lost a race marks x too, costing at most a redundant visit
                atomicMinT(px + s, (T)nd);
                if (nd < hi) near |= 1u << s;
                else far |= 1u << s;
            }
        };
        if constexpr (kVecSeg<T, KS>) {
            T du[KS], dx[KS];
            loadSeg<T, KS>(pu, du);
            loadSeg<T, KS>(px, dx);
#pragma unroll
            for (int s = 0; s < KS; ++s)
                if ((m >> s) & 1) one(s, du[s], dx[s]);
        } else {
            uint32_t mm = m;
            while (mm) {
                const int s = __ffs(mm) - 1;
                mm &= mm - 1;
                one(s, pu[s], px[s]);
            }
        }
        if (ovf) *S.c(M_OVF) = 1;
        mark(x, near, far);
        return 0;
    }
    // U arcs at a time (segments of at most 32 bytes): the arcs' slots, then
    // both ends' distance segments and the heads' summary words, are loaded
    // for all U before any is used; the updates (atomicMin per improved
    // source, the masks' and bitmaps' atomicOr) are then issued without
    // awaiting a result or pre-reading the masks.
    template <int U>
    __device__ void many(const int32_t (&p)[U], const uint64_t (&pay)[U], const int32_t (&)[U],
                         const int64_t (&j)[U], const bool (&ok)[U]) const {
        static_assert(kVecSeg<T, KS>, "batched relaxation needs segments of at most 32 bytes");
        int32_t x[U];
        uint32_t wt[U];
        bool go[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            go[u] = ok[u];
            x[u] = go[u] ? fa[j[u]] : 0;
            wt[u] = (go[u] && w) ? w[j[u]] : 0;
        }
        const uint64_t hi = (uint64_t)*S.c(M_HI);
        T* dist = (T*)S.dist;
        T du[U][KS], dx[U][KS];
        uint32_t nsw[U], fsw[U];  // the halves of the summary words that hold x's bits
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (go[u] && !w) wt[u] = hashWeight(p[u], x[u]);
            if (go[u] && !split && ((wt[u] > delta) != (heavy != 0))) go[u] = false;
            nsw[u] = fsw[u] = 0;
            if (go[u]) {
                loadSeg<T, KS>(dist + (size_t)p[u] * KS, du[u]);
                loadSeg<T, KS>(dist + (size_t)x[u] * KS, dx[u]);
                nsw[u] = ((const uint32_t*)S.nsum)[x[u] >> 11];
                fsw[u] = ((const uint32_t*)S.fsum)[x[u] >> 11];
            }
        }
        bool ovf = false, nset = false;
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (!go[u]) continue;
            const uint32_t m = (uint32_t)pay[u];
            T* px = dist + (size_t)x[u] * KS;
            uint32_t near = 0, far = 0;
#pragma unroll
            for (int s = 0; s < KS; ++s) {
                if (!((m >> s) & 1)) continue;
                if constexpr (sizeof(T) == 4) {
                    // du is finite (s is active at the tail): the 32-bit sum
                    // overflows iff it wraps or reaches the unreached mark,
                    // and otherwise compares with dx directly (unreached =
                    // 0xffffffff is above every finite sum)
                    const uint32_t nd = du[u][s] + wt[u];
                    if (nd < du[u][s] || nd == 0xffffffffu) {
                        ovf = true;
                        continue;
                    }
                    if (nd < dx[u][s]) {
                        atomicMinT(px + s, (T)nd);
                        if ((uint64_t)nd < hi) near |= 1u << s;
                        else far |= 1u << s;
                    }
                } else {
                    const uint64_t nd = widen(du[u][s]) + wt[u];
                    if (nd < widen(dx[u][s])) {
                        atomicMinT(px + s, (T)nd);
                        if (nd < hi) near |= 1u << s;
                        else far |= 1u << s;
                    }
                }
            }
            if (near) nset |= markBitHalf(S.nbits, S.nsum, x[u], nsw[u]);
            if (far) markBitHalf(S.fbits, S.fsum, x[u], fsw[u]);
        }
        if (ovf) *S.c(M_OVF) = 1;
        if (nset) *S.c(par ? M_NSET1 : M_NSET0) = 1;
    }
};

__device__ __forceinline__ void resetRound(const MsState& S, int par) {
    *S.c(par ? M_NEAR0 : M_NEAR1) = 0;
    *S.c(par ? M_NSET0 : M_NSET1) = 0;
    *S.c(M_FMIN) = (int64_t)INF;
    *S.c(M_HQ) = 0;
}

// The k sources: distance 0 and a first queue of distinct vertices (a
// vertex given twice is queued once, with both distances 0).
template <class T, int KS>
__global__ void kMsInit(MsState S, const int32_t* __restrict__ src, int k, uint64_t delta) {
    if (threadIdx.x != 0) return;
    for (int j = 0; j < M_NSLOT; ++j) *S.c(j) = 0;
    *S.c(M_HI) = (int64_t)delta;
    int64_t n = 0;
    for (int s = 0; s < k; ++s) {
        const int32_t x = src[s];
        if (x < 0 || x >= S.V) continue;
        ((T*)S.dist)[(size_t)x * KS + s] = 0;
        int64_t i = 0;
        while (i < n && S.q[0][i] != x) ++i;
        if (i == n) S.q[0][n++] = x;
    }
    *S.c(M_NEAR0) = n;
}

// The sources whose distance at vertex u lies in [lo, hi): a queued vertex
// is relaxed for these (in a light round: the bucket's sources, among them
// every source that lowered u's distance last round; in the heavy pass: the
// sources u settled for in this bucket).
template <class T, int KS>
__device__ __forceinline__ uint32_t bucketMask(const T* dist, int32_t u, uint64_t lo, uint64_t hi) {
    const T* p = dist + (size_t)u * KS;
    uint32_t m = 0;
    if constexpr (kVecSeg<T, KS>) {
        T d[KS];
        loadSeg<T, KS>(p, d);
#pragma unroll
        for (int s = 0; s < KS; ++s) {
            const uint64_t v = widen(d[s]);
            if (v >= lo && v < hi) m |= 1u << s;
        }
    } else {
        for (int s = 0; s < KS; ++s) {
            const uint64_t v = widen(p[s]);
            if (v >= lo && v < hi) m |= 1u << s;
        }
    }
    return m;
}

// Relax the queue of parity par (light arcs), or with heavy_phase the heavy
// queue (heavy arcs); grid stride, one lane per queue entry, for the
// sources of the entry's bucketMask.  track_done: the entries join the done
// set (their heavy arcs are due when the bucket drains).
template <class T, int KS, int U>
__global__ void kMsRelax(MsRelax<T, KS> f, const int32_t* __restrict__ fb, const int32_t* __restrict__ mid,
                         HeavyList H, int track_done, int heavy_phase) {
    const MsState& S = f.S;
    if (!heavy_phase && blockIdx.x == 0 && threadIdx.x == 0) resetRound(S, f.par);
    const int64_t n = heavy_phase ? *S.c(M_HQ) : *S.c(f.par ? M_NEAR1 : M_NEAR0);
    const int32_t* q = heavy_phase ? S.qh : S.q[f.par];
    const uint64_t lo = (uint64_t)*S.c(M_LO), hi = (uint64_t)*S.c(M_HI);
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t base = (int64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~(WAVE - 1)); base < n; base += stride) {
        const int64_t i = base + laneId();
        const bool act = i < n;
        const int32_t u = act ? q[i] : 0;
        const uint32_t m = act ? bucketMask<T, KS>((const T*)S.dist, u, lo, hi) : 0;
        if (track_done) waveMarkBits(S.dbits, S.dsum, u, act);
        int32_t b = act ? fb[u] : 0, e = act ? fb[u + 1] : 0;
        if (mid && act) {
            const int32_t c = mid[u];
            if (heavy_phase) b = c;
            else e = c;
        }
        if constexpr (U > 1) waveRowsHU<U>(act, u, (uint64_t)m, 0, b, e, H, f);
        else waveRowsH(act, u, (uint64_t)m, 0, b, e, H, f);
    }
}

// Bitmap walks in tasks: a task is TW consecutive bitmap words (TW * 64
// vertices) inside one summary word, so a large set (the far set of a
// power-law graph holds most reached vertices) spreads over thousands of
// waves instead of one wave per 4,096 vertices.  A task is taken in groups
// of GW words: expandGroup lists a group's set bits (at most GW * 64 =
// WB_CAP ids) in LDS, and the wave then works on them 64 at a time.
// Passes that append to a queue walk the tasks block-synchronously (every
// wave of a block takes one task per iteration, an empty one past the end)
// so that the block's waves reserve their queue slots with one atomic: a
// single counter hit once per wave would serialise thousands of atomics.
constexpr int TW = 16;
constexpr int GW = 4;
constexpr int TPS = 64 / TW;
static_assert(GW * 64 == WB_CAP, "a group fills the wave's LDS list");

struct Task {
    int64_t g;     // summary word
    int c;         // task within it (words g * 64 + c * TW ..)
    uint32_t sub;  // the task's summary bits (0: nothing to do)
    __device__ int64_t word(int grp) const { return g * 64 + c * TW + grp * GW + laneId(); }
    // bit of this lane's word in group grp (lanes < GW)
    __device__ bool lane(int grp) const { return laneId() < GW && ((sub >> (grp * GW + laneId())) & 1); }
};

__device__ __forceinline__ Task taskAt(const uint64_t* __restrict__ sum, int64_t t, int64_t T) {
    Task k{0, 0, 0};
    if (t < T) {
        k.g = t / TPS;
        k.c = (int)(t % TPS);
        k.sub = (uint32_t)((sum[k.g] >> (k.c * TW)) & ((1u << TW) - 1));
    }
    return k;
}

// The set bits of the lanes' words (m, word index wi) listed in lds; returns
// their number.  The whole wave calls it; lds holds WB_CAP entries.
__device__ __forceinline__ int expandGroup(uint64_t m, int64_t wi, int32_t* lds) {
    const int lane = laneId();
    const int c = __popcll(m);
    int incl = c;
    for (int d = 1; d < WAVE; d <<= 1) {
        const int o = __shfl_up(incl, d);
        if (lane >= d) incl += o;
    }
    const int total = __shfl(incl, WAVE - 1);
    int pos = incl - c;
    while (m) {
        const int b = __ffsll((long long)m) - 1;
        m &= m - 1;
        lds[pos++] = (int32_t)(wi * 64 + b);
    }
    waveSync();
    return total;
}

struct MsLds {
    int32_t ids[1024 / WAVE][WB_CAP];
    uint64_t keep[1024 / WAVE][GW];
};

// body(s, v) for every source s of mask fm, v = x's distance from s (one
// vector read of the segment where it is at most 32 bytes).
template <class T, int KS, class F>
__device__ __forceinline__ void forMaskDist(const T* dist, int32_t x, uint32_t fm, F&& body) {
    const T* p = dist + (size_t)x * KS;
    if constexpr (kVecSeg<T, KS>) {
        T d[KS];
        loadSeg<T, KS>(p, d);
#pragma unroll
        for (int s = 0; s < KS; ++s)
            if ((fm >> s) & 1) body(s, widen(d[s]));
    } else {
        while (fm) {
            const int s = __ffs(fm) - 1;
            fm &= fm - 1;
            body(s, widen(p[s]));
        }
    }
}

// Drains a bitmap: clears its words and summary bits, appends the vertices
// to q at counter *cnt.
__device__ __forceinline__ void drainAll(uint64_t* bits, uint64_t* sum, int64_t W2, int32_t* q, int64_t* cnt,
                                         int32_t* lds) {
    const int lane = laneId(), wpb = blockDim.x / WAVE, wib = threadIdx.x / WAVE;
    const int64_t T = W2 * TPS;
    const int64_t t_first = (int64_t)blockIdx.x * wpb, t_step = (int64_t)gridDim.x * wpb;
    // the block's queue slots are reserved once: a first walk counts the set
    // bits of the wave's tasks, the second drains them
    int64_t mine_n = 0;
    for (int64_t t0 = t_first; t0 < T; t0 += t_step) {
        const Task k = taskAt(sum, t0 + wib, T);
        uint64_t mw = 0;
        if (lane < TW && ((k.sub >> lane) & 1)) mw = bits[k.g * 64 + k.c * TW + lane];
        mine_n += waveSum(__popcll(mw));
    }
    int64_t base = __shfl(blockReserve(lane == 0 ? (int)mine_n : 0, cnt), 0);
    for (int64_t t0 = t_first; t0 < T; t0 += t_step) {
        const Task k = taskAt(sum, t0 + wib, T);
        if (!k.sub) continue;
        uint64_t mw = 0;  // lane l < TW: word l of the task
        if (lane < TW && ((k.sub >> lane) & 1)) {
            uint64_t* wp = bits + k.g * 64 + k.c * TW + lane;
            mw = *wp;
            if (mw) *wp = 0;
        }
        for (int grp = 0; grp < TW / GW; ++grp) {
            const uint64_t m = __shfl(mw, grp * GW + (lane & (GW - 1)));
            const uint64_t mine = lane < GW ? m : 0;
            if (__ballot(mine != 0) == 0) continue;
            const int total = expandGroup(mine, k.word(grp), lds);
            for (int i = lane; i < total; i += WAVE) q[base + i] = lds[i];
            base += total;
            waveSync();
        }
        if (lane == 0) atomicAnd((unsigned long long*)&sum[k.g], ~((unsigned long long)k.sub << (k.c * TW)));
    }
}

__device__ __forceinline__ void farMinCommit(const MsState& S, unsigned long long mn) {
    for (int o = WAVE / 2; o > 0; o >>= 1) {
        const unsigned long long t = __shfl_xor(mn, o);
        mn = t < mn ? t : mn;
    }
    if (laneId() == 0 && mn != INF) atomicMin((unsigned long long*)S.c(M_FMIN), mn);
}

// Smallest far distance >= hi over the far set.
template <class T, int KS>
__device__ __forceinline__ void farMinAll(const MsState& S, MsLds& sh) {
    const uint64_t hi = (uint64_t)ldAcq(S.c(M_HI));
    uint64_t mn = INF;
    int32_t* lds = sh.ids[threadIdx.x / WAVE];
    const T* dist = (const T*)S.dist;
    const int64_t T_ = S.W2 * TPS, nw = ((int64_t)gridDim.x * blockDim.x) / WAVE;
    for (int64_t t = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / WAVE; t < T_; t += nw) {
        const Task k = taskAt(S.fsum, t, T_);
        if (!k.sub) continue;
        for (int grp = 0; grp < TW / GW; ++grp) {
            if (__ballot(k.lane(grp)) == 0) continue;
            const uint64_t m = k.lane(grp) ? S.fbits[k.word(grp)] : 0;
            const int total = expandGroup(m, k.word(grp), lds);
            for (int i = laneId(); i < total; i += WAVE) {
                const int32_t x = lds[i];
                forMaskDist<T, KS>(dist, x, kAllSources<KS>, [&](int, uint64_t v) {
                    if (v >= hi && v < mn) mn = v;
                });
            }
            waveSync();
        }
    }
    farMinCommit(S, mn);
}

// After the relaxation of parity par: if the round put vertices into the
// bucket, the near bitmap becomes the next queue; otherwise the bucket has
// drained: with heavy arcs the done bitmap becomes the heavy queue, else the
// smallest far distance is found right here.
template <class T, int KS>
__global__ void kMsCompact(MsState S, int par, int heavy) {
    __shared__ MsLds sh;
    int32_t* lds = sh.ids[threadIdx.x / WAVE];
    if (ldAcq(S.c(par ? M_NSET1 : M_NSET0)))
        drainAll(S.nbits, S.nsum, S.W2, S.q[par ^ 1], S.c(par ? M_NEAR0 : M_NEAR1), lds);
    else if (heavy)
        drainAll(S.dbits, S.dsum, S.W2, S.qh, S.c(M_HQ), lds);
    else
        farMinAll<T, KS>(S, sh);
}

// The drained bucket's heavy arcs are relaxed: now the far minimum.
template <class T, int KS>
__global__ void kMsFarMin(MsState S, int par) {
    __shared__ MsLds sh;
    if (ldAcq(S.c(par ? M_NSET1 : M_NSET0))) return;
    farMinAll<T, KS>(S, sh);
}

// A drained bucket with far work left advances every source to the bucket
// holding the smallest far distance (every block derives the same [lo, hi);
// block 0 stores it).  Per far vertex and source: inside it -> next queue,
// below it -> settled (dropped), beyond it -> stays.  A bitmap word left
// empty loses its summary bit.  Each task is walked twice: a first pass
// counts the vertices handed to the queue (for the block's one
// reservation), the second writes them and the far set's new state.
template <class T, int KS>
__global__ void kMsSplit(MsState S, int par, uint64_t delta) {
    __shared__ MsLds sh;
    if (ldAcq(S.c(par ? M_NSET1 : M_NSET0))) return;
    const uint64_t mfar = (uint64_t)ldAcq(S.c(M_FMIN));
    if (mfar == INF) return;  // no far work left: every source is done
    const uint64_t lo = (mfar / delta) * delta;
    const uint64_t hi = lo > INF - delta ? INF : lo + delta;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        *S.c(M_LO) = (int64_t)lo;
        *S.c(M_HI) = (int64_t)hi;
    }
    int32_t* q = S.q[par ^ 1];
    int64_t* cnt = S.c(par ? M_NEAR0 : M_NEAR1);
    const int lane = laneId(), wpb = blockDim.x / WAVE, wib = threadIdx.x / WAVE;
    int32_t* lds = sh.ids[wib];
    uint64_t* keepw = sh.keep[wib];
    const T* dist = (const T*)S.dist;
    // a far entry of source s: a finite distance at or past hi (see above)
    auto classify = [&](int32_t x, uint32_t& keep, uint32_t& take) {
        keep = take = 0;
        forMaskDist<T, KS>(dist, x, kAllSources<KS>, [&](int s, uint64_t v) {
            if (v >= hi) {
                if (v != INF) keep |= 1u << s;
            } else if (v >= lo) {
                take |= 1u << s;
            }
        });
    };
    const int64_t T_ = S.W2 * TPS;
    const int64_t t_first = (int64_t)blockIdx.x * wpb, t_step = (int64_t)gridDim.x * wpb;
    // a first walk over the wave's tasks counts what goes to the queue, so
    // the block reserves its slots once; the second walk writes them
    int ntake = 0;
    for (int64_t t0 = t_first; t0 < T_; t0 += t_step) {
        const Task k = taskAt(S.fsum, t0 + wib, T_);
        for (int grp = 0; k.sub && grp < TW / GW; ++grp) {
            if (__ballot(k.lane(grp)) == 0) continue;
            const uint64_t m = k.lane(grp) ? S.fbits[k.word(grp)] : 0;
            const int total = expandGroup(m, k.word(grp), lds);
            for (int i0 = 0; i0 < total; i0 += WAVE) {
                const int i = i0 + lane;
                uint32_t keep = 0, take = 0;
                if (i < total) classify(lds[i], keep, take);
                ntake += __popcll(__ballot(take != 0));
            }
            waveSync();
        }
    }
    int64_t base = __shfl(blockReserve(lane == 0 ? ntake : 0, cnt), 0);
    for (int64_t t0 = t_first; t0 < T_; t0 += t_step) {
        const Task k = taskAt(S.fsum, t0 + wib, T_);
        if (!k.sub) continue;
        uint64_t emptied = 0;
        for (int grp = 0; grp < TW / GW; ++grp) {
            if (__ballot(k.lane(grp)) == 0) continue;
            const int64_t wi = k.word(grp);
            const uint64_t m = k.lane(grp) ? S.fbits[wi] : 0;
            if (lane < GW) keepw[lane] = 0;
            const int total = expandGroup(m, wi, lds);
            for (int i0 = 0; i0 < total; i0 += WAVE) {
                const int i = i0 + lane;
                uint32_t keep = 0, take = 0;
                if (i < total) {
                    const int32_t x = lds[i];
                    classify(x, keep, take);
                    if (keep) atomicOr((unsigned long long*)&keepw[(x >> 6) & (GW - 1)], 1ULL << (x & 63));
                }
                const uint64_t bal = __ballot(take != 0);
                if (take) q[base + __popcll(bal & lanesBelow())] = lds[i];
                base += __popcll(bal);
            }
            waveSync();
            const uint64_t keep = lane < GW ? keepw[lane] : 0;
            if (k.lane(grp) && keep != m) S.fbits[wi] = keep;
            emptied |= (__ballot(k.lane(grp) && keep == 0) & ((1ULL << GW) - 1)) << (grp * GW);
            waveSync();
        }
        if (lane == 0 && emptied) atomicAnd((unsigned long long*)&S.fsum[k.g], ~(emptied << (k.c * TW)));
    }
}

// Interleaved distances -> the caller's k rows of V uint64 (UINT64_MAX =
// unreachable): one thread per vertex reads its segment once.
template <class T, int KS>
__global__ void kMsOut(const T* __restrict__ dist, int64_t V, int k, uint64_t* __restrict__ out) {
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < V; v += (int64_t)gridDim.x * blockDim.x) {
        const T* p = dist + (size_t)v * KS;
        if constexpr (kVecSeg<T, KS>) {
            T d[KS];
            loadSeg<T, KS>(p, d);
#pragma unroll
            for (int s = 0; s < KS; ++s)
                if (s < k) out[(size_t)s * V + v] = widen(d[s]);
        } else {
            for (int s = 0; s < k; ++s) out[(size_t)s * V + v] = widen(p[s]);
        }
    }
}

constexpr int MS_CHUNK = 16;  // rounds per host readback while the queue is small

// One batch of k <= KS sources (device array dsrc); rows: k rows of V
// uint64 (device).  Returns false if 32-bit distances overflowed (nothing
// written to rows then).
template <class T, int KS>
bool msRun(girhip_graph* h, const int32_t* dsrc, int k, uint64_t delta, const girhip_graph::SplitView* sv,
           uint64_t* rows) {
    Scratch& S_ = scratch(h);
    const int32_t V = h->V;
    const int64_t W = ((int64_t)V + 63) / 64, W2 = (W + 63) / 64;
    const size_t Vs = (size_t)std::max<int32_t>(V, 1);
    S_.ms_dist.ensure(sizeof(T) * Vs * KS);
    for (int i = 0; i < 3; ++i) {
        S_.ms_bits[i].ensure((size_t)std::max<int64_t>(W, 1));
        S_.ms_sum[i].ensure((size_t)std::max<int64_t>(W2, 1));
        S_.ms_q[i].ensure(Vs);
    }
    S_.ms_ctl.ensure(M_NSLOT * MSTR);
    MsState S;
    S.dist = S_.ms_dist.p;
    S.nbits = S_.ms_bits[0].p;
    S.fbits = S_.ms_bits[1].p;
    S.dbits = S_.ms_bits[2].p;
    S.nsum = S_.ms_sum[0].p;
    S.fsum = S_.ms_sum[1].p;
    S.dsum = S_.ms_sum[2].p;
    S.q[0] = S_.ms_q[0].p;
    S.q[1] = S_.ms_q[1].p;
    S.qh = S_.ms_q[2].p;
    S.ctl = S_.ms_ctl.p;
    S.V = V;
    S.W = W;
    S.W2 = W2;
    if constexpr (sizeof(T) == 4) fillU32(h, (uint32_t*)S.dist, 0xffffffffu, (int64_t)V * KS);
    else fillU64(h, (uint64_t*)S.dist, INF, (int64_t)V * KS);
    for (int i = 0; i < 3; ++i) {
        fillU64(h, S_.ms_bits[i].p, 0, W);
        fillU64(h, S_.ms_sum[i].p, 0, W2);
    }
    launch(h, "sssp_ms_init", kMsInit<T, KS>, 1, 64, 0, S, dsrc, k, delta);

    const int blk = (int)h->prm.block;
    const int32_t* fa = sv ? sv->adj.p : h->fw_adj.p;
    const uint32_t* w = sv ? sv->w.p : (h->has_w ? h->w_fw.p : nullptr);
    const int32_t* mid = sv ? sv->mid.p : nullptr;
    const int heavy = (uint64_t)maxWeight(h) > delta ? 1 : 0;
    const int sp = sv ? 1 : 0;
    const MsRelax<T, KS> fl[2] = {{S, fa, w, delta, 0, 0, sp}, {S, fa, w, delta, 0, 1, sp}};
    const MsRelax<T, KS> fh[2] = {{S, fa, w, delta, 1, 0, sp}, {S, fa, w, delta, 1, 1, sp}};
    // arcs per lane at a time: the knob, capped so that the U segment pairs
    // held in registers stay within 64 bytes each (1 where segments are
    // longer than 32 bytes: they are read a distance at a time)
    constexpr int SEG = KS * (int)sizeof(T);
    const int ur = !kVecSeg<T, KS> ? 1 : std::min<int>((int)h->prm.sssp_unroll, std::max(1, 64 / SEG));
    auto relaxK = kMsRelax<T, KS, 1>;
    if constexpr (kVecSeg<T, KS> && SEG <= 32) {
        if (ur == 2) relaxK = kMsRelax<T, KS, 2>;
    }
    if constexpr (kVecSeg<T, KS> && SEG <= 16) {
        if (ur == 4) relaxK = kMsRelax<T, KS, 4>;
    }
    const unsigned hgrid = std::min<unsigned>(gridFor(V, blk), 4096);
    // bitmap passes: one wave per task of TW words, at most 2048 blocks
    const unsigned sgrid = (unsigned)std::max<int64_t>(1, std::min<int64_t>(2048, (W2 * TPS * WAVE + blk - 1) / blk));
    const int32_t* fb = h->fw_beg.p;
    const HeavyList H = heavyList(h);
    const int cmax = std::max(1, std::min(MS_CHUNK, (int)h->prm.chunk));
    int par = 0, R = 1;
    int64_t maxn = 1;
    // a chunk's first round relaxes the queue read back (one lane per entry);
    // the queues of its later rounds are unknown and may be large (a bucket
    // that drains hands the next one its far entries), so those launches
    // get enough waves to fill the chip
    const unsigned gfill = std::min<unsigned>(gridFor(V, blk), (unsigned)std::max(1, (1 << 18) / blk));
    while (true) {
        const unsigned g0 = gridFor(maxn, blk), glater = std::max(g0, gfill);
        int p = par;
        for (int t = 0; t < R; ++t) {
            launch(h, "sssp_ms_relax", relaxK, t == 0 ? g0 : glater, blk, 0, fl[p], fb, mid, H, heavy, 0);
            launchHeavyStore(h, "sssp_ms_relax_rows", fl[p]);
            launch(h, "sssp_ms_compact", kMsCompact<T, KS>, sgrid, blk, 0, S, p, heavy);
            if (heavy) {
                launch(h, "sssp_ms_relax_heavy", relaxK, hgrid, blk, 0, fh[p], fb, mid, H, 0, 1);
                launchHeavyStore(h, "sssp_ms_relax_heavy_rows", fh[p]);
                launch(h, "sssp_ms_farmin", kMsFarMin<T, KS>, sgrid, blk, 0, S, p);
            }
            launch(h, "sssp_ms_split", kMsSplit<T, KS>, sgrid, blk, 0, S, p, delta);
            p ^= 1;
        }
        par = p;
        int64_t c[2];
        // queue length of the next round, and the overflow flag
        readCounters(h, S.ctl + (par ? M_NEAR1 : M_NEAR0) * MSTR, 1, c);
        if (c[0] == 0) break;  // queue empty after an advance: every source settled
        R = (c[0] * 64 < (int64_t)V && c[0] <= 2 * maxn && maxn > 1) ? std::min(cmax, 2 * R) : 1;
        maxn = c[0];
    }
    if constexpr (sizeof(T) == 4) {
        int64_t ovf = 0;
        readCounters(h, S.ctl + M_OVF * MSTR, 1, &ovf);
        if (ovf) return false;
    }
    launch(h, "sssp_ms_out", kMsOut<T, KS>, boundedGrid(V, 256), 256, 0, (const T*)S.dist, (int64_t)V, k, rows);
    return true;
}

template <int KS>
void msBatchKS(girhip_graph* h, const int32_t* dsrc, int k, uint64_t delta, const girhip_graph::SplitView* sv,
               uint64_t* rows, bool try32) {
    if (try32 && msRun<uint32_t, KS>(h, dsrc, k, delta, sv, rows)) return;
    msRun<uint64_t, KS>(h, dsrc, k, delta, sv, rows);
}

}  // namespace

void msBatch(girhip_graph* h, const int32_t* dsrc, int k, uint64_t delta, const girhip_graph::SplitView* sv,
             uint64_t* rows, bool try32) {
    if (k <= 0) return;
    if (k > MS_MAX) raise(GIRHIP_EINVAL, "sssp: more than 32 sources in one interleaved batch");
    if (k == 1) msBatchKS<1>(h, dsrc, k, delta, sv, rows, try32);
    else if (k == 2) msBatchKS<2>(h, dsrc, k, delta, sv, rows, try32);
    else if (k <= 4) msBatchKS<4>(h, dsrc, k, delta, sv, rows, try32);
    else if (k <= 8) msBatchKS<8>(h, dsrc, k, delta, sv, rows, try32);
    else if (k <= 16) msBatchKS<16>(h, dsrc, k, delta, sv, rows, try32);
    else msBatchKS<32>(h, dsrc, k, delta, sv, rows, try32);
}

}  // namespace gir

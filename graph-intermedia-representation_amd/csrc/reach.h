// Direction-optimising reachability engine (device side + host driver).
//
// The structure follows GIR::bfs (gir_query.cpp:108-294): a top-down step
// expands a frontier queue over push-direction arcs and claims vertices in a
// visited bitmap; a bottom-up step lets every unvisited vertex look for a
// parent in a frontier bitmap over pull-direction arcs and stops at the
// first hit; the switch rule is the reference's (alpha, beta, minimum
// reachable size).  On the GPU:
//   - top-down: one lane per frontier vertex, rows longer than 16 arcs are
//     finished by the whole wave (waveRows); claims are atomicOr on the
//     64-bit visited word; new vertices are appended with one atomic per
//     wave (ballot + prefix);
//   - bottom-up: one wave per 64-vertex bitmap word, so the next-frontier
//     word is a single ballot and the visited word is written without
//     atomics;
//   - the filter Flt decides which arcs may be followed (child, parent), so
//     the same kernels run plain BFS, SCC forward/backward reachability over
//     the still-unassigned vertices, and the colour-restricted backward
//     sweeps of the colouring phase.
#pragma once
#include "common.h"

namespace gir {

struct ReachBufs {
    int32_t* q[2];
    uint64_t *vis, *fbits, *nbits;
    int64_t W;  // bitmap words
};

// Seed: mark the initial frontier visited and call on(v, 0, -1) (a seed has
// no parent).
template <class Vis>
__global__ void kReachSeed(const int32_t* __restrict__ q, int64_t n, uint64_t* vis, Vis on) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t v = q[i];
    atomicOr((unsigned long long*)&vis[v >> 6], 1ULL << (v & 63));
    on(v, 0, -1);
}

// cnt[0] = next queue length, cnt[1] = sum of push-direction degrees of the
// claimed vertices; cz points at the counter pair of the following level,
// which the level kernels zero.
// Top-down work on arc slot j of row p: claim the head, queue it, return its
// push-direction degree (summed into cnt[1]).
template <class Flt, class Vis>
struct TopDownF {
    const int32_t* beg;
    const int32_t* adj;
    uint64_t* vis;
    int32_t L;
    int32_t* qout;
    int64_t* cnt;
    Flt flt;
    Vis on;
    __device__ int64_t operator()(int32_t p, uint64_t, int32_t, int64_t j) const {
        const int32_t w = adj[j];
        bool claim = false;
        if (flt(w, p)) {
            const uint64_t bit = 1ULL << (w & 63);
            if (!(vis[w >> 6] & bit))  // (visited bits are only set: safe pre-check)
                claim = !(atomicOr((unsigned long long*)&vis[w >> 6], (unsigned long long)bit) & bit);
        }
        const int64_t pos = wavePush(&cnt[0], claim);
        if (claim) {
            qout[pos] = w;
            on(w, L + 1, p);
            return beg[w + 1] - beg[w];
        }
        return 0;
    }
    // the heavy pass's form: claim, and the queue slot comes from the block
    __device__ int64_t heavyClaim(int32_t p, uint64_t, int32_t, int64_t j, bool& push, int32_t& item) const {
        const int32_t w = adj[j];
        bool claim = false;
        if (flt(w, p)) {
            const uint64_t bit = 1ULL << (w & 63);
            if (!(vis[w >> 6] & bit))
                claim = !(atomicOr((unsigned long long*)&vis[w >> 6], (unsigned long long)bit) & bit);
        }
        push = claim;
        item = w;
        if (!claim) return 0;
        on(w, L + 1, p);
        return beg[w + 1] - beg[w];
    }
    __device__ int64_t* pushCounter() const { return &cnt[0]; }
    __device__ void heavyPlace(int64_t pos, int32_t w) const { qout[pos] = w; }
};

// The input queue length is read on the device (the previous level's
// counter), so several levels can be queued without a host readback; the
// grid strides over the queue, uniformly per wave.
template <class Flt, class Vis>
__global__ void kTopDown(const int32_t* __restrict__ qin, const int64_t* nin, TopDownF<Flt, Vis> f, int64_t* cz,
                         HeavyList H) {
    if (blockIdx.x == 0 && threadIdx.x < 2) cz[threadIdx.x] = 0;
    const int64_t n = *nin;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    int64_t m = 0;
    for (int64_t base = (int64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~(WAVE - 1)); base < n; base += stride) {
        const int64_t i = base + laneId();
        const bool act = i < n;
        const int32_t u = act ? qin[i] : 0;
        const int32_t b = act ? f.beg[u] : 0, e = act ? f.beg[u + 1] : 0;
        m += waveRowsH(act, u, 0, 0, b, e, H, f);
    }
    blockAtomicAdd(m, &f.cnt[1]);
}

constexpr int32_t BU_SMALL = 8;
// A bottom-up block walks BU_TILES consecutive tiles of blockDim vertices,
// gathering what it finds in LDS, and reserves its queue slots once: the
// first bottom-up step of a sweep finds vertices in nearly every block, and
// one counter hit by every block of a V-wide grid serialises thousands of
// atomics.
constexpr int BU_TILES = 8;
// bottom-up levels queued per host readback (the later ones guarded)
constexpr int BU_CHAIN = 3;
inline unsigned buGrid(int64_t V, int block) { return gridFor(V, block * BU_TILES); }

// guard (may be null): the previous bottom-up level's packed count, for a
// level queued behind it without a host readback; the level runs only if
// the host's own rule would have run it (the previous level found vertices
// and nf * beta >= n_reach), else every block returns at once.
template <class Flt, class Vis>
__global__ void __launch_bounds__(1024) kBottomUp(const int32_t* __restrict__ rbeg, const int32_t* __restrict__ radj,
                                                  const int32_t* __restrict__ pbeg, uint64_t* vis,
                                                  const uint64_t* __restrict__ fbits, uint64_t* __restrict__ nbits,
                                                  int32_t V, int32_t L, int32_t* __restrict__ qout, int64_t* cnt,
                                                  int64_t* cz, Flt flt, Vis on, const int64_t* __restrict__ guard,
                                                  int64_t n_reach, int64_t beta) {
    if (guard) {
        const int64_t nf_prev = (int64_t)((uint64_t)*guard >> 32);
        if (nf_prev == 0 || nf_prev * beta < n_reach) return;
    }
    __shared__ int32_t lst[BU_TILES * 1024];
    __shared__ int lcount;
    __shared__ int64_t wm[1024 / WAVE];
    __shared__ int64_t base_s;
    if (blockIdx.x == 0 && threadIdx.x < 2) cz[threadIdx.x] = 0;
    if (threadIdx.x == 0) lcount = 0;
    __syncthreads();
    int64_t msum = 0;
    for (int tile = 0; tile < BU_TILES; ++tile) {
        const int64_t v = ((int64_t)blockIdx.x * BU_TILES + tile) * blockDim.x + threadIdx.x;
        const int64_t word = v >> 6;
        const bool word_ok = (word << 6) < V;
        const uint64_t w0 = word_ok ? vis[word] : ~0ULL;
        const bool act = v < V && !((w0 >> laneId()) & 1) && flt.self((int32_t)v);
        bool found = false;
        int32_t par = -1;  // the frontier vertex found (kept only for visitors that record parents)
        // each lane scans the first BU_SMALL slots of its own row; rows with
        // slots left and no parent yet are then scanned by the whole wave,
        // one row at a time, 64 coalesced slots per step with a ballot exit
        // (a hub's long row no longer holds the other 63 lanes)
        const int32_t b = act ? rbeg[v] : 0, e = act ? rbeg[v + 1] : 0;
        const int32_t lim = e - b > BU_SMALL ? b + BU_SMALL : e;
        for (int32_t j = b; j < lim; ++j) {
            const int32_t p = radj[j];
            if (((fbits[p >> 6] >> (p & 63)) & 1) && flt((int32_t)v, p)) {
                found = true;
                if constexpr (Vis::kParent) par = p;
                break;
            }
        }
        uint64_t rest = __ballot(act && !found && lim < e);
        while (rest) {
            const int l = __ffsll((long long)rest) - 1;
            rest &= rest - 1;
            const int32_t rv = __shfl((int32_t)v, l), rb = __shfl(lim, l), re = __shfl(e, l);
            bool hit = false;
            int32_t hp = -1;
            for (int32_t j0 = rb; j0 < re; j0 += WAVE) {
                const int32_t j = j0 + laneId();
                bool hj = false;
                int32_t pj = -1;
                if (j < re) {
                    pj = radj[j];
                    hj = ((fbits[pj >> 6] >> (pj & 63)) & 1) && flt(rv, pj);
                }
                const uint64_t hb = __ballot(hj);
                if (hb) {
                    hit = true;
                    if constexpr (Vis::kParent) hp = __shfl(pj, __ffsll((long long)hb) - 1);
                    break;
                }
            }
            if (laneId() == l && hit) {
                found = true;
                par = hp;
            }
        }
        const uint64_t nb = __ballot(found);
        if (laneId() == 0 && word_ok) {
            nbits[word] = nb;
            if (nb) vis[word] = w0 | nb;
        }
        if (found) msum += pbeg[v + 1] - pbeg[v];
        int slot = 0;
        if (laneId() == 0 && nb) slot = atomicAdd(&lcount, __popcll(nb));
        slot = __shfl(slot, 0);
        if (found) {
            lst[slot + __popcll(nb & lanesBelow())] = (int32_t)v;
            on((int32_t)v, L + 1, par);
        }
    }
    // cnt[0] gets (found << 32) + their push-arc count, one atomic per block;
    // the host splits it into the level's (nf, mf) pair
    msum = waveSum(msum);
    if (laneId() == 0) wm[threadIdx.x / WAVE] = msum;
    __syncthreads();
    if (threadIdx.x == 0) {
        int64_t m = 0;
        for (int i = 0; i < (int)(blockDim.x / WAVE); ++i) m += wm[i];
        const int n = lcount;
        base_s = (n || m) ? (int64_t)(atomicAdd((unsigned long long*)&cnt[0],
                                                ((unsigned long long)n << 32) + (unsigned long long)m) >>
                                      32)
                          : 0;
    }
    __syncthreads();
    const int n = lcount;
    for (int i = threadIdx.x; i < n; i += blockDim.x) qout[base_s + i] = lst[i];
}

static __global__ void kQueueToBits(const int32_t* __restrict__ q, int64_t n, uint64_t* bits) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) atomicOr((unsigned long long*)&bits[q[i] >> 6], 1ULL << (q[i] & 63));
}

// Arc filters and visit hooks.  A hook is called as on(v, level, parent)
// when v is claimed (parent = the frontier vertex whose arc reached v; -1 for
// a seed, and for bottom-up claims when the hook's kParent is false).
struct AllArcs {
    __device__ bool operator()(int32_t, int32_t) const { return true; }
    __device__ bool self(int32_t) const { return true; }
};
struct NoVisit {
    static constexpr bool kParent = false;
    __device__ void operator()(int32_t, int32_t, int32_t) const {}
};
struct WriteLevel {
    static constexpr bool kParent = false;
    int32_t* level;
    __device__ void operator()(int32_t v, int32_t L, int32_t) const { level[v] = L; }
};
// Records the claiming parent (seeds keep theirs): a BFS tree over the arcs
// the sweep followed.
struct WriteParent {
    static constexpr bool kParent = true;
    int32_t* par;
    __device__ void operator()(int32_t v, int32_t, int32_t p) const {
        if (p >= 0) par[v] = p;
    }
};

struct ReachStats {
    int32_t levels = 0;
    int64_t reached = 0;  // including the seeds
};

// Levels per host readback while a top-down frontier is tiny.
constexpr int REACH_CHUNK = 16;
// Counter ring: level L writes pair L mod REACH_RING (next queue length,
// its push degree sum), reads pair L-1 as its input length and zeroes pair
// L+1 for the level after it.
constexpr int REACH_RING = 64;
static_assert(REACH_RING >= REACH_CHUNK + 3, "ring must hold a chunk and its neighbours");

// Persistent top-down levels: one cooperative launch runs level after level
// (top-down step, grid barrier) until the frontier is empty or the switch
// rule asks for bottom-up (mf * alpha > arcs unexplored; alpha 0 = never).
// io[0..3] in and out: level, queue parity, vertices reached, arcs
// unexplored.  Every block reads the same counters after the barrier, so
// all leave the loop together.  Used only when no row of the store can
// reach the heavy list.
template <class Flt, class Vis>
__global__ void __launch_bounds__(256) kReachPersist(const int32_t* __restrict__ pbeg, const int32_t* __restrict__ padj,
                                                     uint64_t* vis, int32_t* q0, int32_t* q1, int64_t* ring,
                                                     int64_t* io, int64_t alpha, Flt flt, Vis on, HeavyList H,
                                                     uint32_t* bar) {
    int32_t L = (int32_t)ldAcq(&io[0]);
    int cur = (int)ldAcq(&io[1]);
    int64_t reached = ldAcq(&io[2]), mun = ldAcq(&io[3]);
    auto pair = [&](int64_t l) { return ring + 2 * ((l % REACH_RING + REACH_RING) % REACH_RING); };
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    while (!gridFailed(bar)) {
        const int64_t* pin = pair(L - 1);
        const int64_t nf = ldAcq(&pin[0]), mf = ldAcq(&pin[1]);
        if (nf == 0 || (alpha > 0 && mf * alpha > mun)) break;
        const int32_t* qin = cur ? q1 : q0;
        int32_t* qout = cur ? q0 : q1;
        int64_t* cn = pair(L);
        if (blockIdx.x == 0 && threadIdx.x < 2) pair(L + 1)[threadIdx.x] = 0;
        const TopDownF<Flt, Vis> f{pbeg, padj, vis, L, qout, cn, flt, on};
        int64_t m = 0;
        for (int64_t base = (int64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~(WAVE - 1)); base < nf;
             base += stride) {
            const int64_t i = base + laneId();
            const bool act = i < nf;
            const int32_t u = act ? qin[i] : 0;
            const int32_t b = act ? pbeg[u] : 0, e = act ? pbeg[u + 1] : 0;
            m += waveRowsH(act, u, 0, 0, b, e, H, f);
        }
        blockAtomicAdd(m, &cn[1]);
        gridSync(bar);
        reached += ldAcq(&cn[0]);
        mun -= ldAcq(&cn[1]);
        cur ^= 1;
        ++L;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        io[0] = L;
        io[1] = cur;
        io[2] = reached;
        io[3] = mun;
    }
}

// The top-down -> bottom-up switch factor of this engine: the reference's
// alpha (bfs_alpha), raised to 64 on graphs with hubs (largest degree >= 64
// times the mean; knob bfs_alpha_auto) where a bottom-up step over every
// vertex costs less than a top-down step over a hub-heavy frontier.  RMAT-22,
// one source, mean of 16: 1.35 ms at 15, 0.92 ms at 64 to 256; a 4800^2
// lattice (no hubs): 111 ms at 15, 138 ms at 128.  (The bit-parallel
// multi-source BFS keeps bfs_alpha: at 64 it measured 18 ms against 5.5.)
inline int64_t reachAlpha(girhip_graph* h) {
    const int64_t a = (int64_t)h->prm.bfs_alpha;
    if (h->prm.bfs_alpha_auto == 0 || a <= 0 || h->V <= 0) return a;
    const bool hubs = maxDegree(h) * (int64_t)h->V >= 64 * h->E;
    return hubs ? std::max<int64_t>(a, 64) : a;
}

// Runs reachability from the n0 seeds in B.q[0] (caller fills them).  The
// visited bitmap B.vis must be clear on entry except for bits the caller
// wants treated as already visited.  pbeg/padj: push-direction CSR; rbeg/radj:
// pull-direction CSR.  m0 = push degree sum of the seeds; n_reach / m_total:
// the size estimates of the switch rule.
template <class Flt, class Vis>
ReachStats runReach(girhip_graph* h, const int32_t* pbeg, const int32_t* padj, const int32_t* rbeg,
                    const int32_t* radj, ReachBufs& B, int64_t n0, int64_t m0, int64_t n_reach, int64_t m_total,
                    Flt flt, Vis on, bool allow_bottom_up = true, int64_t* ring = nullptr,
                    const HeavyList* hlist = nullptr, bool any_row_heavy = false) {
    const int32_t V = h->V;
    const int blk = (int)h->prm.block;
    ReachStats st;
    if (n0 <= 0) return st;
    launch(h, "reach_seed", kReachSeed<Vis>, gridFor(n0, blk), blk, 0, (const int32_t*)B.q[0], n0, B.vis, on);
    // level counter ring and heavy list: the caller's (2 * REACH_RING int64;
    // no persistent kernel then), else the handle's scratch.  any_row_heavy:
    // the rows are not the edge store's (the condensation's: a row there can
    // outgrow every vertex's), so the heavy-row pass always runs.
    Scratch& S = scratch(h);
    int64_t* c = ring;
    if (!c) {
        S.rcnt.ensure(2 * REACH_RING + 4);  // + the persistent kernel's state
        c = S.rcnt.p;
    }
    const HeavyList HL = hlist ? *hlist : heavyList(h);
    auto pair = [&](int64_t L) { return c + 2 * ((L % REACH_RING + REACH_RING) % REACH_RING); };
    fillI64(h, c, 0, 2 * REACH_RING);
    fillI64(h, pair(-1), n0, 1);  // input length of level 0
    const int64_t alpha = reachAlpha(h), beta = (int64_t)h->prm.bfs_beta;
    const int64_t min_reach = (int64_t)h->prm.bfs_min_reach;
    int64_t nf = n0, mf = m0, m_unexplored = m_total - m0;
    bool bu = false, bits_valid = false;
    int64_t mf_prev = 0;  // frontier arcs one readback earlier
    int c_prev = 1;
    int cur = 0;
    // persistent levels need every row below the heavy-list threshold
    bool persist_ok = !ring && !any_row_heavy && h->prm.persist != 0 && 2 * maxDegree(h) <= HEAVY_T;  // (closure rows too)
    int32_t L = 0;
    st.reached = n0;
    while (nf > 0) {
        if (allow_bottom_up && n_reach >= min_reach) {
            if (!bu && mf * alpha > m_unexplored) bu = true;
            else if (bu && nf * beta < n_reach) bu = false;
        }
        if (bu) {
            int64_t* cn = pair(L);
            if (!bits_valid) {
                fillU64(h, B.fbits, 0, B.W);
                launch(h, "reach_q2bits", kQueueToBits, gridFor(nf, blk), blk, 0, (const int32_t*)B.q[cur], nf,
                       B.fbits);
            }
            // BU_CHAIN levels per readback: each level after the first runs
            // only if the host's rule would have run it (kBottomUp's guard)
            for (int t = 0; t < BU_CHAIN; ++t) {
                launch(h, "reach_bottomup", kBottomUp<Flt, Vis>, buGrid(V, blk), blk, 0, rbeg, radj, pbeg, B.vis,
                       (const uint64_t*)B.fbits, B.nbits, V, L + t, B.q[((cur + t) & 1) ^ 1], pair(L + t),
                       pair(L + t + 1), flt, on, (const int64_t*)(t > 0 ? pair(L + t - 1) : nullptr), n_reach, beta);
                std::swap(B.fbits, B.nbits);
            }
            int64_t r[2 * REACH_RING];
            readCounters(h, c, 2 * REACH_RING, r);
            int ran = 0;
            for (int t = 0; t < BU_CHAIN; ++t) {
                if (t > 0 && (nf == 0 || nf * beta < n_reach)) break;
                const int64_t pk = r[pair(L + t) - c];
                nf = (int64_t)((uint64_t)pk >> 32);
                mf = (int64_t)((uint64_t)pk & 0xffffffffULL);
                m_unexplored -= mf;
                st.reached += nf;
                ++ran;
            }
            // every chained level ran: the bitmaps are current; else the
            // chain stopped where the rule leaves bottom-up (or the sweep ends)
            bits_valid = ran == BU_CHAIN;
            int64_t* cl = pair(L + ran - 1);
            fillI64(h, cl, nf, 1);  // the last level's pair as the other kernels read it
            fillI64(h, cl + 1, mf, 1);
            cur ^= ran & 1;
            L += ran;
            mf_prev = 0;
            continue;
        }
        bits_valid = false;
        if (persist_ok) {
            int64_t* io = S.rcnt.p + 2 * REACH_RING;
            fillI64(h, io, L, 1);
            fillI64(h, io + 1, cur, 1);
            fillI64(h, io + 2, st.reached, 1);
            fillI64(h, io + 3, m_unexplored, 1);
            const int64_t a = (allow_bottom_up && n_reach >= min_reach) ? alpha : 0;
            if (launchPersistent(h, "reach_persist", kReachPersist<Flt, Vis>, 256, 2, pbeg, padj, B.vis, B.q[0], B.q[1],
                                 c, io, a, flt, on, HL)) {
                int64_t r[2 * REACH_RING + 4];
                readCounters(h, c, 2 * REACH_RING + 4, r);
                checkGridBar(h);
                const int64_t* o = r + 2 * REACH_RING;
                L = (int32_t)o[0];
                cur = (int)o[1];
                st.reached = o[2];
                m_unexplored = o[3];
                const int64_t* p = r + (pair(L - 1) - c);
                nf = p[0];
                mf = p[1];
                mf_prev = 0;
                continue;
            }
            persist_ok = false;  // no cooperative launch here: host-driven levels
        }
        // tiny frontiers, far below the switch point and not growing fast: a
        // chunk of levels
        const int C = (mf_prev > 0 && mf * 64 < (int64_t)V && mf * alpha * 8 < m_unexplored && mf <= 2 * mf_prev)
                          ? std::min(std::max(1, std::min(REACH_CHUNK, (int)h->prm.chunk)), 2 * c_prev)
                          : 1;
        c_prev = C;
        mf_prev = mf;
        const int64_t work = std::max<int64_t>(nf, mf) * (C > 1 ? 4 : 1);
        const unsigned g =
            C > 1 ? (unsigned)std::min<int64_t>(1024, std::max<int64_t>(64, (work + blk - 1) / blk)) : gridFor(nf, blk);
        for (int t = 0; t < C; ++t) {
            const TopDownF<Flt, Vis> f{pbeg, padj, B.vis, L + t, B.q[cur ^ 1], pair(L + t), flt, on};
            launch(h, "reach_topdown", kTopDown<Flt, Vis>, g, blk, 0, (const int32_t*)B.q[cur],
                   (const int64_t*)pair(L + t - 1), f, pair(L + t + 1), HL);
            if (any_row_heavy)  // (a quarter of the store's grid: the pass is launched at every level)
                launch(h, "reach_topdown_heavy", kHeavyRows<TopDownF<Flt, Vis>>, 256, 256, 0, HL, f, pair(L + t) + 1);
            else if (2 * maxDegree(h) > HEAVY_T)  // (blocks of 1,024: a quarter of the queue reservations,
                                                   // which serialise on one counter, for the same threads)
                launch(h, "reach_topdown_heavy", kHeavyRows<TopDownF<Flt, Vis>>, 256, 1024, 0, HL, f,
                       pair(L + t) + 1);
            cur ^= 1;
        }
        int64_t r[2 * REACH_RING];
        readCounters(h, c, 2 * REACH_RING, r);
        const int32_t L0 = L;
        for (int t = 0; t < C; ++t) {
            const int64_t* p = r + (pair(L0 + t) - c);
            nf = p[0];
            mf = p[1];
            m_unexplored -= mf;
            st.reached += nf;
            ++L;
            if (nf == 0) break;
        }
    }
    st.levels = L;
    return st;
}

// ---- two sweeps side by side ------------------------------------------------
// The SCC build's forward and backward sweeps from the pivot are
// independent; on a high-diameter graph each is thousands of levels of a
// few thousand vertices, so its cost is one launch per level.  Here one
// launch runs a top-down level of both: the first g1 blocks serve the
// forward sweep, the rest the backward one (each part strides over its own
// queue and writes its own counter ring).  Used while both frontiers are
// small (below V / 64 vertices) and no row can reach the heavy list; a sweep
// still running after that finishes with runReach from its frontier.
template <class Flt, class Vis>
__global__ void kTopDownPair(const int32_t* __restrict__ qin1, const int64_t* nin1, TopDownF<Flt, Vis> f1,
                             int64_t* cz1, const int32_t* __restrict__ qin2, const int64_t* nin2,
                             TopDownF<Flt, Vis> f2, int64_t* cz2, unsigned g1, HeavyList H) {
    const bool first = blockIdx.x < g1;
    const TopDownF<Flt, Vis>& f = first ? f1 : f2;
    const int32_t* qin = first ? qin1 : qin2;
    const int64_t b0 = first ? blockIdx.x : blockIdx.x - g1, nb = first ? g1 : gridDim.x - g1;
    if (b0 == 0 && threadIdx.x < 2) (first ? cz1 : cz2)[threadIdx.x] = 0;
    const int64_t n = *(first ? nin1 : nin2);
    const int64_t stride = nb * blockDim.x;
    int64_t m = 0;
    for (int64_t base = b0 * blockDim.x + (threadIdx.x & ~(WAVE - 1)); base < n; base += stride) {
        const int64_t i = base + laneId();
        const bool act = i < n;
        const int32_t u = act ? qin[i] : 0;
        const int32_t b = act ? f.beg[u] : 0, e = act ? f.beg[u + 1] : 0;
        m += waveRowsH(act, u, 0, 0, b, e, H, f);
    }
    blockAtomicAdd(m, &f.cnt[1]);
}

// Reachability from seed (one vertex) over (fb, fa) into B1 and over
// (bb, ba) into B2 at the same time: both visited bitmaps clear on entry,
// on1 / on2 the sweeps' visit hooks.  B1 / B2 need their own queues;
// fbits / nbits are used only by a sweep's runReach finish (one sweep at a
// time).  n_reach / m_total: the switch rule's size estimates.
template <class Flt, class Vis>
void runReachPair(girhip_graph* h, const int32_t* fb, const int32_t* fa, const int32_t* bb, const int32_t* ba,
                  int32_t seed, ReachBufs& B1, Vis on1, ReachBufs& B2, Vis on2, int64_t n_reach, int64_t m_total,
                  Flt flt) {
    const int32_t V = h->V;
    const int blk = (int)h->prm.block;
    Scratch& S = scratch(h);
    S.rcnt2.ensure(4 * REACH_RING);
    int64_t* c1 = S.rcnt2.p;
    int64_t* c2 = S.rcnt2.p + 2 * REACH_RING;
    auto pr = [&](int64_t* c, int64_t L) { return c + 2 * ((L % REACH_RING + REACH_RING) % REACH_RING); };
    fillI32(h, B1.q[0], seed, 1);
    fillI32(h, B2.q[0], seed, 1);
    launch(h, "reach_seed", kReachSeed<Vis>, 1, 64, 0, (const int32_t*)B1.q[0], (int64_t)1, B1.vis, on1);
    launch(h, "reach_seed", kReachSeed<Vis>, 1, 64, 0, (const int32_t*)B2.q[0], (int64_t)1, B2.vis, on2);
    fillI64(h, c1, 0, 4 * REACH_RING);
    fillI64(h, pr(c1, -1), 1, 1);
    fillI64(h, pr(c2, -1), 1, 1);
    const HeavyList HL = heavyList(h);
    int64_t nf[2] = {1, 1}, mf[2] = {0, 0}, mf_prev = 0, m_seen[2] = {0, 0};
    int c_prev = 1, cur = 0;
    int32_t L = 0;
    while ((nf[0] > 0 || nf[1] > 0) && nf[0] * 64 < (int64_t)V && nf[1] * 64 < (int64_t)V) {
        const int64_t mx = std::max(mf[0], mf[1]);
        // small frontiers not growing fast: a chunk of levels per readback
        const int C = (mf_prev > 0 && mx * 64 < (int64_t)V && mx <= 2 * mf_prev)
                          ? std::min(std::max(1, std::min(REACH_CHUNK, (int)h->prm.chunk)), 2 * c_prev)
                          : 1;
        c_prev = C;
        mf_prev = mx;
        auto grid = [&](int k) {
            const int64_t work = std::max<int64_t>(nf[k], mf[k]) * (C > 1 ? 4 : 1);
            return C > 1 ? (unsigned)std::min<int64_t>(1024, std::max<int64_t>(64, (work + blk - 1) / blk))
                         : std::max(1u, gridFor(nf[k], blk));
        };
        const unsigned g1 = grid(0), g2 = grid(1);
        for (int t = 0; t < C; ++t) {
            const TopDownF<Flt, Vis> f1{fb, fa, B1.vis, L + t, B1.q[cur ^ 1], pr(c1, L + t), flt, on1};
            const TopDownF<Flt, Vis> f2{bb, ba, B2.vis, L + t, B2.q[cur ^ 1], pr(c2, L + t), flt, on2};
            launch(h, "reach_topdown_pair", kTopDownPair<Flt, Vis>, g1 + g2, blk, 0, (const int32_t*)B1.q[cur],
                   (const int64_t*)pr(c1, L + t - 1), f1, pr(c1, L + t + 1), (const int32_t*)B2.q[cur],
                   (const int64_t*)pr(c2, L + t - 1), f2, pr(c2, L + t + 1), g1, HL);
            cur ^= 1;
        }
        int64_t r[4 * REACH_RING];
        readCounters(h, c1, 4 * REACH_RING, r);
        const int32_t L0 = L;
        for (int t = 0; t < C; ++t) {
            for (int k = 0; k < 2; ++k) {
                const int64_t* p = r + (pr(k ? c2 : c1, L0 + t) - c1);
                nf[k] = p[0];
                mf[k] = p[1];
                m_seen[k] += p[1];
            }
            ++L;
        }
    }
    // a sweep whose frontier outgrew the pair finishes alone from it
    for (int k = 0; k < 2; ++k) {
        if (nf[k] == 0) continue;
        ReachBufs& B = k ? B2 : B1;
        if (cur) HIPCHK(hipMemcpyAsync(B.q[0], B.q[1], sizeof(int32_t) * (size_t)nf[k], hipMemcpyDeviceToDevice, h->st));
        if (k == 0)
            runReach(h, fb, fa, bb, ba, B, nf[k], mf[k], n_reach, m_total - m_seen[k] + mf[k], flt, on1);
        else
            runReach(h, bb, ba, fb, fa, B, nf[k], mf[k], n_reach, m_total - m_seen[k] + mf[k], flt, on2);
    }
}

// Reachability buffers from the handle's scratch (q[0], q[1], bits[0..2]).
ReachBufs reachBufs(girhip_graph* h);

}  // namespace gir

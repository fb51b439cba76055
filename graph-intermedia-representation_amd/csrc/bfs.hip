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

// ---- small frontiers: one workgroup per source -----------------------------
// A high-diameter graph (a lattice: thousands of levels of a few thousand
// vertices) spends the grid engines' time in launches and level barriers.
// Here source s = blockIdx.x walks all its levels inside one workgroup of
// 1024 threads: the level's frontier list is shared out evenly over the
// waves, each wave flattens its vertices' rows over its lanes (BFS_TINY_U
// arcs per lane at a time), a head still at -1 is claimed by atomicCAS and
// appended to the next list (LDS counter), and a workgroup barrier ends the
// level.  A frontier above BFS_TINY_CAP stops the source (status 1): the
// caller then runs the grid engine for it.  Used while no row is longer
// than BFS_TINY_DEG arcs.
constexpr int BFS_TINY_CAP = 1 << 15;
constexpr int64_t BFS_TINY_DEG = 1024;
constexpr int BFS_TINY_U = 4;

struct BfsTinyF {
    const int32_t* fa;
    int32_t* lv;
    int32_t* next;
    int* next_n;
    int32_t level;  // of the heads
    template <int U>
    __device__ __forceinline__ void many(const int32_t (&)[U], const uint64_t (&)[U], const int32_t (&)[U],
                                         const int64_t (&j)[U], const bool (&ok)[U]) const {
        int32_t x[U], old[U];
#pragma unroll
        for (int u = 0; u < U; ++u) x[u] = ok[u] ? fa[j[u]] : -1;
#pragma unroll
        for (int u = 0; u < U; ++u) old[u] = x[u] >= 0 ? lv[x[u]] : 0;  // (plain: a stale -1 only costs the CAS)
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (x[u] >= 0 && old[u] == -1 && atomicCAS(&lv[x[u]], -1, level) == -1) next[atomicAdd(next_n, 1)] = x[u];
    }
};

// levels: k rows of V, -1 already; l0 / l1: frontier lists of V ids per
// source; status[s]: 0 done (or no valid source), 1 stopped.
__global__ void __launch_bounds__(1024) kBfsTiny(const int32_t* __restrict__ fb, const int32_t* __restrict__ fa,
                                                 int32_t V, const int32_t* __restrict__ src, int32_t* levels,
                                                 int32_t* l0, int32_t* l1, int32_t* status) {
    __shared__ int next_n;
    const int s = blockIdx.x;
    const int32_t x0 = src[s];
    if (x0 < 0 || x0 >= V) {
        if (threadIdx.x == 0) status[s] = 0;
        return;
    }
    int32_t* lv = levels + (size_t)s * V;
    int32_t* cur = l0 + (size_t)s * V;
    int32_t* nxt = l1 + (size_t)s * V;
    if (threadIdx.x == 0) {
        lv[x0] = 0;
        cur[0] = x0;
    }
    __syncthreads();
    const int nw = (int)blockDim.x / WAVE, wv = (int)threadIdx.x / WAVE;
    int n = 1, st = 0;
    for (int32_t L = 0; n > 0; ++L) {
        if (threadIdx.x == 0) next_n = 0;
        __syncthreads();
        const BfsTinyF f{fa, lv, nxt, &next_n, L + 1};
        const int per = (n + nw - 1) / nw, beg = wv * per, end = min(n, beg + per);
        for (int base = beg; base < end; base += WAVE) {
            const int i = base + laneId();
            const bool act = i < end;
            const int32_t u = act ? cur[i] : 0;
            const int32_t b = act ? fb[u] : 0, e = act ? fb[u + 1] : 0;
            waveFlatU<BFS_TINY_U>(u, 0, 0, b, e - b, f);
        }
        __syncthreads();
        n = next_n;
        if (n > BFS_TINY_CAP) {
            st = 1;
            break;
        }
        int32_t* t = cur;
        cur = nxt;
        nxt = t;
    }
    if (threadIdx.x == 0) status[s] = st;
}

// Levels of k sources (device array dsrc, host copy hsrc) with the
// one-workgroup engine; the sources it stopped are returned (their rows are
// reset to -1 for the caller's grid engine).
static std::vector<int32_t> bfsTiny(girhip_graph* h, const int32_t* dsrc, const std::vector<int32_t>& hsrc,
                                    int32_t* out) {
    Scratch& S = scratch(h);
    const int32_t V = h->V;
    const int k = (int)hsrc.size();
    const size_t kv = (size_t)k * (size_t)std::max<int32_t>(V, 1);
    S.s_fl[0].ensure(kv);
    S.s_fl[1].ensure(kv);
    S.iv[3].ensure((size_t)std::max(k, 1));
    fillI32(h, out, -1, (int64_t)k * V);
    launch(h, "bfs_tiny", kBfsTiny, (unsigned)k, 1024, 0, (const int32_t*)h->fw_beg.p, (const int32_t*)h->fw_adj.p,
           V, dsrc, out, S.s_fl[0].p, S.s_fl[1].p, S.iv[3].p);
    std::vector<int32_t> st(k);
    HIPCHK(hipMemcpyAsync(st.data(), S.iv[3].p, sizeof(int32_t) * k, hipMemcpyDeviceToHost, h->st));
    syncStream(h);
    std::vector<int32_t> again;
    for (int i = 0; i < k; ++i)
        if (st[i]) {
            again.push_back(i);
            fillI32(h, out + (size_t)i * V, -1, V);
        }
    return again;
}

// ---------------------------------------------------------------------------
// Bit-parallel multi-source BFS (MS-BFS, Then et al., VLDB 2015): up to 64
// sources share one traversal.  Bit s of a vertex's seen / frontier / next
// word belongs to source s, so an arc is read once per level for all sources
// whose frontier holds its tail.  Push levels OR the tail's frontier word into
// the head's next word (atomicOr, skipped when nothing new); pull levels let
// every vertex that some source has not reached yet OR the frontier words of
// its in-neighbours, stopping once every source has reached it.  Levels are
// written once per (source, vertex): at the level a bit first appears, or -1
// in a final pass over the unseen bits.
//
// Besides the per-vertex words, the union frontier is a vertex bitmap F and
// the heads touched by a push a bitmap T, each with a summary bitmap (one bit
// per 64-bit word).  A level runs in one of three modes:
//   pull   - every unfinished vertex scans its in-arcs (large frontiers);
//   dense  - push from every vertex, finish over all V;
//   sparse - push from F, finish over T: O(frontier) work per level, no
//            single-address counters (bitmap words are owned by one wave).

// Per-level counters in a ring (nf = union frontier size, mf = its out-arc
// count).  A level's push zeroes the next level's pair; its finish fills it.
struct MsCnt {
    int64_t nf, mf;
};

// Levels per host readback while the frontier is tiny (high-diameter graphs).
constexpr int MS_CHUNK = 16;
// Counter ring: one MsCnt per level, indexed by level mod MS_RING.
constexpr int MS_RING = 64;
static_assert(MS_RING >= MS_CHUNK + 2, "ring must hold a chunk and its neighbours");

struct MsBits {
    uint64_t *F, *FS, *T, *TS;
    int64_t W, W2;
};

__global__ void kMsSeed(const int32_t* __restrict__ src, int k, int32_t V, uint64_t* seen, uint64_t* front,
                        int32_t* __restrict__ levels, MsBits M, const int32_t* __restrict__ fb, MsCnt* cnt) {
    const int s = threadIdx.x;
    if (s >= k) return;
    const int32_t v = src[s];
    if (v < 0 || v >= V) return;
    atomicOr((unsigned long long*)&seen[v], 1ULL << s);
    const uint64_t old = atomicOr((unsigned long long*)&front[v], 1ULL << s);
    levels[(size_t)s * V + v] = 0;
    if (old == 0) {  // first source at v
        setBit(M.F, M.FS, v);
        atomicAdd((unsigned long long*)&cnt->nf, 1ULL);
        atomicAdd((unsigned long long*)&cnt->mf, (unsigned long long)(fb[v + 1] - fb[v]));
    }
}

// Sparse push: the head's next word gets the tail's frontier bits; the first
// bits to reach a head mark it in T.
struct MsPushSF {
    const int32_t* fa;
    const uint64_t* seen;
    uint64_t* next;
    MsBits M;
    __device__ int64_t operator()(int32_t, uint64_t fp, int32_t, int64_t j) const {
        const int32_t w = fa[j];
        const uint64_t m = fp & ~(seen[w] | next[w]);  // (next only gains bits here: safe pre-check)
        if (m && atomicOr((unsigned long long*)&next[w], (unsigned long long)m) == 0) setBit(M.T, M.TS, w);
        return 0;
    }
};

// Sparse push body for one summary word g of F: the 64 bitmap words under
// it are consumed (cleared) and every frontier vertex in them pushes its row
// with its frontier word, which is cleared too.
__device__ __forceinline__ void msPushGroup(const int32_t* __restrict__ fb, uint64_t* front, const MsPushSF& f,
                                            HeavyList H, int64_t g, uint64_t sm, int32_t* lds) {
    const MsBits& M = f.M;
    const int l = laneId();
    uint64_t m = 0;
    if ((sm >> l) & 1) {
        uint64_t* wp = M.F + g * 64 + l;
        m = *wp;
        if (m) *wp = 0;
    }
    if (l == 0) M.FS[g] = 0;
    waveBits(m, g * 64 + l, lds, [&](bool act0, int32_t u) {
        uint64_t fr = 0;
        if (act0) {
            fr = front[u];
            front[u] = 0;
        }
        const bool act = fr != 0;
        const int32_t b = act ? fb[u] : 0, e = act ? fb[u + 1] : 0;
        waveRowsH(act, u, fr, 0, b, e, H, f);
    });
}

// Sparse finish body for one summary word g of T: new bits -> seen,
// frontier word, level writes, F.  The wave owns the 64 T words and the F
// words of the same vertices; F bits are gathered in the wave's LDS words
// (fwl, 64 entries) and written once per word, with F's summary word.
__device__ __forceinline__ void msFinishGroup(const MsBits& M, uint64_t* seen, uint64_t* front, uint64_t* next,
                                              int32_t* __restrict__ levels, int32_t V, int32_t L,
                                              const int32_t* __restrict__ fb, int64_t g, uint64_t sm, int32_t* lds,
                                              uint64_t* fwl, int64_t& nf, int64_t& mf) {
    const int l = laneId();
    uint64_t t = 0;
    if ((sm >> l) & 1) {
        uint64_t* wp = M.T + g * 64 + l;
        t = *wp;
        if (t) *wp = 0;
    }
    if (l == 0) M.TS[g] = 0;
    fwl[l] = 0;
    waveSync();
    waveBits(t, g * 64 + l, lds, [&](bool act, int32_t w) {
        if (!act) return;
        const uint64_t sw = seen[w];
        const uint64_t nb = next[w] & ~sw;
        next[w] = 0;
        if (nb) {
            seen[w] = sw | nb;
            front[w] = nb;
            atomicOr((unsigned long long*)&fwl[(w >> 6) & 63], 1ULL << (w & 63));
            ++nf;
            mf += fb[w + 1] - fb[w];
            uint64_t x = nb;
            while (x) {
                const int s = __ffsll((long long)x) - 1;
                x &= x - 1;
                levels[(size_t)s * V + w] = L + 1;
            }
        }
    });
    waveSync();
    const uint64_t fw = fwl[l];
    if (fw) M.F[g * 64 + l] = fw;
    const uint64_t any = __ballot(fw != 0);
    if (l == 0 && any) M.FS[g] = any;
}

// LDS of the sparse passes: per wave, WB_CAP ids and 64 F words.
struct MsLds {
    int32_t ids[1024 / WAVE][WB_CAP];
    uint64_t fw[1024 / WAVE][64];
};

__global__ void kMsPushS(const int32_t* __restrict__ fb, uint64_t* front, MsPushSF f, MsCnt* nxt, HeavyList H,
                         int G) {
    __shared__ MsLds sh;
    if (blockIdx.x == 0 && threadIdx.x == 0) *nxt = MsCnt{0, 0};
    int32_t* lds = sh.ids[threadIdx.x / WAVE];
    forSummary(f.M.FS, f.M.W2, G, [&](int64_t g, uint64_t sm) { msPushGroup(fb, front, f, H, g, sm, lds); });
}

__global__ void kMsFinishS(MsBits M, uint64_t* seen, uint64_t* front, uint64_t* next, int32_t* __restrict__ levels,
                           int32_t V, int32_t L, const int32_t* __restrict__ fb, MsCnt* nxt, int G) {
    __shared__ MsLds sh;
    const int wv = threadIdx.x / WAVE;
    int64_t nf = 0, mf = 0;
    forSummary(M.TS, M.W2, G, [&](int64_t g, uint64_t sm) {
        msFinishGroup(M, seen, front, next, levels, V, L, fb, g, sm, sh.ids[wv], sh.fw[wv], nf, mf);
    });
    blockAtomicAdd(nf, &nxt->nf);
    blockAtomicAdd(mf, &nxt->mf);
}

// Persistent sparse levels: one cooperative launch runs push / barrier /
// finish / barrier per level for as long as the levels stay sparse and push
// (the host's rule: nf * 4 < V, and mf * alpha <= E unless alpha is 0), up
// to level Lmax.  *Lio: first level in, first level not run out.  Every
// block reads the same counters after a barrier, so all leave the loop
// together.  Used only when no row of the store can reach the heavy list.
__global__ void __launch_bounds__(256) kMsPersist(MsPushSF f, const int32_t* __restrict__ fb, uint64_t* front,
                                                  int32_t* __restrict__ levels, int32_t V, MsCnt* ring, int64_t* Lio,
                                                  int64_t E, int64_t alpha, int32_t Lmax, HeavyList H,
                                                  uint32_t* bar) {
    __shared__ MsLds sh;
    const int wv = threadIdx.x / WAVE;
    const MsBits& M = f.M;
    int32_t L = (int32_t)ldAcq(Lio);
    while (L < Lmax && !gridFailed(bar)) {
        MsCnt* cur = &ring[L % MS_RING];
        const int64_t nf = ldAcq(&cur->nf), mf = ldAcq(&cur->mf);
        if (nf == 0 || nf * 4 >= (int64_t)V || (alpha > 0 && mf * alpha > E)) break;
        MsCnt* nxt = &ring[(L + 1) % MS_RING];
        if (blockIdx.x == 0 && threadIdx.x == 0) *nxt = MsCnt{0, 0};
        forSummary(M.FS, M.W2, 1, [&](int64_t g, uint64_t sm) { msPushGroup(fb, front, f, H, g, sm, sh.ids[wv]); });
        gridSync(bar);
        int64_t a = 0, b = 0;
        forSummary(M.TS, M.W2, 1, [&](int64_t g, uint64_t sm) {
            msFinishGroup(M, const_cast<uint64_t*>(f.seen), front, f.next, levels, V, L, fb, g, sm, sh.ids[wv],
                          sh.fw[wv], a, b);
        });
        blockAtomicAdd(a, &nxt->nf);
        blockAtomicAdd(b, &nxt->mf);
        gridSync(bar);
        ++L;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) *Lio = L;
}

struct MsPushF {
    const int32_t* fa;
    const uint64_t* seen;
    uint64_t* next;
    __device__ int64_t operator()(int32_t, uint64_t fp, int32_t, int64_t j) const {
        const int32_t w = fa[j];
        const uint64_t m = fp & ~(seen[w] | next[w]);
        if (m) atomicOr((unsigned long long*)&next[w], (unsigned long long)m);
        return 0;
    }
};

__global__ void kMsPush(const int32_t* __restrict__ fb, const uint64_t* __restrict__ front, int32_t V, MsPushF f,
                        MsCnt* nxt, HeavyList H) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *nxt = MsCnt{0, 0};
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t fr = v < V ? front[v] : 0;
    const bool act = fr != 0;
    const int32_t b = act ? fb[v] : 0, e = act ? fb[v + 1] : 0;
    waveRowsH(act, (int32_t)v, fr, 0, b, e, H, f);
}

// *act = OR of every frontier word: the sources still expanding (reduction
// grid, one atomic per block; *act cleared before the launch).
__global__ void kMsActive(const uint64_t* __restrict__ front, int32_t V, unsigned long long* act) {
    __shared__ uint64_t part[1024 / WAVE];
    uint64_t m = 0;
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < V; v += (int64_t)gridDim.x * blockDim.x)
        m |= front[v];
    for (int d = WAVE / 2; d > 0; d >>= 1) m |= __shfl_xor((unsigned long long)m, d);
    if (laneId() == 0) part[threadIdx.x / WAVE] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int i = 1; i < (int)(blockDim.x + WAVE - 1) / WAVE; ++i) m |= part[i];
        if (m) atomicOr(act, (unsigned long long)m);
    }
}

// Pull level: a vertex collects its in-neighbours' frontier words until it
// holds every source still expanding (*act): a source whose search has
// stopped can add no bit, so sources of small reach (a random query source
// with no out-arcs, say) do not force every vertex to scan its whole in-row.
constexpr int PULL_BATCH = 8;
__global__ void kMsPull(const int32_t* __restrict__ bb, const int32_t* __restrict__ ba,
                        const uint64_t* __restrict__ front, const uint64_t* __restrict__ seen,
                        uint64_t* __restrict__ next, int32_t V, const uint64_t* __restrict__ actp, MsCnt* nxt) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *nxt = MsCnt{0, 0};
    const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= V) return;
    const uint64_t all = *actp;
    const uint64_t sw = seen[w];
    uint64_t acc = 0;
    if ((sw & all) != all) {
        // PULL_BATCH in-arcs at a time, all their loads issued together: one
        // memory latency per batch instead of a dependent chain per arc
        const int32_t b = bb[w], e = bb[w + 1];
        for (int32_t j0 = b; j0 < e; j0 += PULL_BATCH) {
            int32_t p[PULL_BATCH];
            uint64_t fw[PULL_BATCH];
#pragma unroll
            for (int k = 0; k < PULL_BATCH; ++k) p[k] = j0 + k < e ? ba[j0 + k] : -1;
#pragma unroll
            for (int k = 0; k < PULL_BATCH; ++k) fw[k] = p[k] >= 0 ? front[p[k]] : 0;
#pragma unroll
            for (int k = 0; k < PULL_BATCH; ++k) acc |= fw[k];
            if (((acc | sw) & all) == all) break;
        }
    }
    next[w] = acc & ~sw;
}

// Dense finish over all V: new bits -> seen, frontier words, and either the
// level's new-bit words (lw: kept for kMsLevels, one 8-byte store per
// vertex) or, with lw null, the level writes themselves; every F word is
// rewritten (one wave = one word), F's summary was cleared before the
// launch.  A block walks MF_TILES tiles and adds its counts once (one counter
// hit by every block of a V-wide grid serialises them).
constexpr int MF_TILES = 16;
__global__ void kMsFinish(uint64_t* seen, uint64_t* front, uint64_t* next, int32_t* __restrict__ levels, int32_t V,
                          int32_t L, const int32_t* __restrict__ fb, MsBits M, MsCnt* nxt, uint64_t* __restrict__ lw) {
    int64_t packed = 0;  // (new vertices << 32) + their push-arc count
    for (int t = 0; t < MF_TILES; ++t) {
        const int64_t w = ((int64_t)blockIdx.x * MF_TILES + t) * blockDim.x + threadIdx.x;
        uint64_t nb = 0;
        if (w < V) {
            const uint64_t sw = seen[w];
            nb = next[w] & ~sw;
            seen[w] = sw | nb;
            front[w] = nb;
            next[w] = 0;
            if (lw) lw[w] = nb;
        }
        const uint64_t fw = __ballot(nb != 0);
        if (fw && !lw) {
            for (int s = 0; s < 64; ++s)
                if ((nb >> s) & 1) levels[(size_t)s * V + w] = L + 1;
        }
        const int64_t word = w >> 6;
        if (laneId() == 0 && word < M.W) {
            M.F[word] = fw;
            if (fw) atomicOr((unsigned long long*)&M.FS[word >> 6], 1ULL << (word & 63));
        }
        if (nb) packed += (1LL << 32) + (int64_t)(fb[w + 1] - fb[w]);
    }
    // the host splits nxt->nf into (nf, mf)
    blockAtomicAdd(packed, &nxt->nf);
}

// The dense levels whose new-bit words were kept: n consecutive levels
// first + 1 .. first + n, level j's words at lw + j * V.
constexpr int MS_KEPT = 7;
struct MsKept {
    const uint64_t* lw;
    int32_t first, n;
};

// Level rows of every (source, vertex) pair first reached at a kept dense
// level (first + 1 + j for kept level j) or never reached (-1).  Pairs
// reached at other levels were written by their level's finish and are left
// alone.  The kept words are folded into three bit planes of the code j + 1
// (pairs are reached once, so a bit is set in at most one kept level); each
// source's row is then one coalesced 256-byte segment per wave, written
// once, where writes at every level covered a row segment a few lanes at a
// time.
__global__ void kMsLevels(const uint64_t* __restrict__ seen, MsKept K, int32_t* __restrict__ levels, int32_t V,
                          int32_t k) {
    const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= V) return;
    uint64_t p0 = 0, p1 = 0, p2 = 0;
#pragma unroll
    for (int j = 0; j < MS_KEPT; ++j) {
        if (j < K.n) {
            const uint64_t b = K.lw[(size_t)j * V + w];
            if ((j + 1) & 1) p0 |= b;
            if ((j + 1) & 2) p1 |= b;
            if ((j + 1) & 4) p2 |= b;
        }
    }
    const uint64_t un = ~seen[w];
    const uint64_t mine = un | p0 | p1 | p2;
    for (int s = 0; s < k; ++s) {
        if ((mine >> s) & 1) {
            const int32_t c = (int32_t)((p0 >> s) & 1) | (int32_t)(((p1 >> s) & 1) << 1) |
                              (int32_t)(((p2 >> s) & 1) << 2);
            levels[(size_t)s * V + w] = c ? K.first + c : -1;
        }
    }
}


// Up to 64 sources (device array src, k entries) into levels rows 0..k-1.
//
// Direction: the reference's alpha rule compares the frontier's arcs with the
// arcs not yet explored, which a union frontier cannot track (a vertex joins
// it once per distinct level at which sources reach it), so pull starts when
// the frontier's arcs exceed E / alpha and ends when its size falls below
// V / beta.  Push levels are sparse while the frontier holds under a quarter
// of V and its rows under E / 64 arcs (a sparse pass has one wave per 4,096
// vertices: a few thousand frontier vertices with long rows, such as the
// sources' hub neighbours, kept RMAT-22's second level at 0.9 ms there and
// take 0.12 ms as a dense push); sparse levels whose frontier is tiny and
// not growing fast run MS_CHUNK at a time with one readback per chunk.
void msbfs(girhip_graph* h, const int32_t* dsrc, int k, int32_t* levels) {
    const int32_t V = h->V;
    const int blk = (int)h->prm.block;
    Scratch& S = scratch(h);
    const size_t nV = (size_t)(V > 0 ? V : 1);
    MsBits M;
    M.W = ((int64_t)nV + 63) / 64;
    M.W2 = (M.W + 63) / 64;
    for (int i = 0; i < 3; ++i) S.ms[i].ensure(nV);
    S.msbm[0].ensure((size_t)M.W);
    S.msbm[1].ensure((size_t)M.W2);
    S.msbm[2].ensure((size_t)M.W);
    S.msbm[3].ensure((size_t)M.W2);
    M.F = S.msbm[0].p;
    M.FS = S.msbm[1].p;
    M.T = S.msbm[2].p;
    M.TS = S.msbm[3].p;
    S.mscnt.ensure(2 * MS_RING + 2);  // + the persistent kernel's level slot, the pull's active mask
    uint64_t *seen = S.ms[0].p, *front = S.ms[1].p, *next = S.ms[2].p;
    MsCnt* ring = (MsCnt*)S.mscnt.p;
    const uint64_t all = k >= 64 ? ~0ULL : ((1ULL << k) - 1);
    for (int i = 0; i < 3; ++i) fillU64(h, S.ms[i].p, 0, V);
    fillU64(h, M.F, 0, M.W);
    fillU64(h, M.FS, 0, M.W2);
    fillU64(h, M.T, 0, M.W);
    fillU64(h, M.TS, 0, M.W2);
    fillI64(h, S.mscnt.p, 0, 2 * MS_RING);
    const int32_t* fb = h->fw_beg.p;
    launch(h, "msbfs_seed", kMsSeed, 1, 64, 0, dsrc, k, V, seen, front, levels, M, fb, ring);
    int64_t r0[2];
    readCounters(h, S.mscnt.p, 2, r0);
    const int64_t alpha = (int64_t)h->prm.bfs_alpha, beta = (int64_t)h->prm.bfs_beta;
    const int64_t min_reach = (int64_t)h->prm.bfs_min_reach;
    const int64_t E = h->E;
    int64_t nf = r0[0], mf = r0[1];
    bool pull = false;
    int64_t mf_prev = 0;  // frontier arcs one readback earlier (0: none / dense)
    int c_prev = 1;
    // persistent sparse levels need every row below the heavy-list threshold
    bool persist_ok = h->prm.persist != 0 && maxDegree(h) <= HEAVY_T;
    const unsigned g = gridFor(V, blk);
    MsKept K{nullptr, 0, 0};
    int32_t L = 0;
    while (nf > 0) {
        if ((int64_t)V >= min_reach) {
            if (!pull && mf * alpha > E) pull = true;
            else if (pull && nf * beta < (int64_t)V) pull = false;
        }
        const bool sparse = !pull && nf * 4 < (int64_t)V && mf * 64 < E;
        if (!sparse) {
            MsCnt* nxt = &ring[(L + 1) % MS_RING];
            if (pull) {
                uint64_t* act = (uint64_t*)(S.mscnt.p + 2 * MS_RING + 1);
                fillU64(h, act, 0, 1);
                launch(h, "msbfs_active", kMsActive, reduceGrid(V, blk), blk, 0, (const uint64_t*)front, V,
                       (unsigned long long*)act);
                launch(h, "msbfs_pull", kMsPull, g, blk, 0, (const int32_t*)h->bw_beg.p, (const int32_t*)h->bw_adj.p,
                       (const uint64_t*)front, (const uint64_t*)seen, next, V, (const uint64_t*)act, nxt);
            }
            else {
                const MsPushF f{h->fw_adj.p, seen, next};
                launch(h, "msbfs_push", kMsPush, g, blk, 0, fb, (const uint64_t*)front, V, f, nxt, heavyList(h));
                launchHeavyStore(h, "msbfs_push_heavy", f);
            }
            fillU64(h, M.FS, 0, M.W2);
            // a dense level extending the kept run keeps its words; others write levels now
            uint64_t* lw = nullptr;
            if (K.n < MS_KEPT && (K.n == 0 || L == K.first + K.n)) {
                if (K.n == 0) {
                    S.msl.ensure((size_t)MS_KEPT * nV);
                    K.lw = S.msl.p;
                    K.first = L;
                }
                lw = S.msl.p + (size_t)K.n * nV;
                ++K.n;
            }
            launch(h, "msbfs_finish", kMsFinish, gridFor(V, blk * MF_TILES), blk, 0, seen, front, next, levels, V, L,
                   fb, M, nxt, lw);
            int64_t r[2];
            readCounters(h, (const int64_t*)nxt, 2, r);
            nf = (int64_t)((uint64_t)r[0] >> 32);
            mf = (int64_t)((uint64_t)r[0] & 0xffffffffULL);
            fillI64(h, &nxt->nf, nf, 1);  // the level's pair as the other kernels read it
            fillI64(h, &nxt->mf, mf, 1);
            mf_prev = 0;
            ++L;
            continue;
        }
        if (persist_ok) {
            int64_t* Lio = S.mscnt.p + 2 * MS_RING;
            fillI64(h, Lio, L, 1);
            const MsPushSF f{h->fw_adj.p, seen, next, M};
            const int64_t a = (int64_t)V >= min_reach ? alpha : 0;
            const int32_t Lmax = V + 1;  // a BFS has fewer levels than vertices
            if (launchPersistent(h, "msbfs_persist", kMsPersist, 256, 2, f, fb, front, levels, V, ring, Lio, E, a, Lmax,
                                 heavyList(h))) {
                int64_t r[2 * MS_RING + 1];
                readCounters(h, S.mscnt.p, 2 * MS_RING + 1, r);
                checkGridBar(h);
                L = (int32_t)r[2 * MS_RING];
                nf = r[2 * (L % MS_RING)];
                mf = r[2 * (L % MS_RING) + 1];
                mf_prev = 0;
                continue;
            }
            persist_ok = false;  // no cooperative launch here: host-driven levels
        }
        const int cmax = std::max(1, std::min(MS_CHUNK, (int)h->prm.chunk));
        const int C = (mf_prev > 0 && mf * 64 < E && mf <= 2 * mf_prev) ? std::min(cmax, 2 * c_prev) : 1;
        c_prev = C;
        mf_prev = mf;
        const MsPushSF f{h->fw_adj.p, seen, next, M};
        // one wave per summary word: the passes are latency-bound
        const int G = 1;
        const unsigned gs = summaryGrid(M.W2, G, 2, blk);
        for (int t = 0; t < C; ++t) {
            MsCnt* nxt = &ring[(L + t + 1) % MS_RING];
            launch(h, "msbfs_push_s", kMsPushS, gs, blk, 0, fb, front, f, nxt, heavyList(h), G);
            launchHeavyStore(h, "msbfs_push_heavy", f);
            launch(h, "msbfs_finish_s", kMsFinishS, gs, blk, 0, M, seen, front, next, levels, V, L + t, fb, nxt, G);
        }
        int64_t r[2 * MS_RING];
        readCounters(h, S.mscnt.p, 2 * MS_RING, r);
        for (int t = 1; t <= C; ++t) {
            const int64_t* c = r + 2 * ((L + t) % MS_RING);
            nf = c[0];
            mf = c[1];
            if (nf == 0) break;
        }
        L += C;
    }
    launch(h, "msbfs_levels", kMsLevels, g, blk, 0, (const uint64_t*)seen, K, levels, V, (int32_t)k);
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
        // hub-free graphs, several sources: one workgroup per source first
        // (a single source keeps the grid engine: one workgroup walks a
        // lattice's levels of thousands of vertices slower than the grid)
        const bool tiny = k > 1 && h->prm.bfs_tiny != 0 && maxDegree(h) <= BFS_TINY_DEG;
        const bool bitpar = k > 1 && h->prm.msbfs != 0 && !tiny;
        int32_t* dsrc = nullptr;
        if (bitpar || tiny) {
            S.src.ensure((size_t)k);
            dsrc = S.src.p;
            copyIn(h, dsrc, s.data(), sizeof(int32_t) * (size_t)k);
            syncStream(h);
        }
        tBegin(h);
        std::vector<int32_t> rest;  // sources for the per-source grid engine
        if (tiny) {
            rest = bfsTiny(h, dsrc, s, out);
        } else if (!bitpar) {
            fillI32(h, out, -1, (int64_t)n);
            for (int32_t i = 0; i < k; ++i) rest.push_back(i);
        }
        if (bitpar) {
            for (int32_t i0 = 0; i0 < k; i0 += 64) {
                const int c = std::min<int32_t>(64, k - i0);
                msbfs(h, dsrc + i0, c, out + (size_t)i0 * V);
            }
        } else {
            for (const int32_t i : rest) {
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

namespace gir {
// Vertices of one component: label[w] == c.
struct SameLabel {
    const int32_t* lab;
    int32_t c;
    __device__ bool operator()(int32_t w, int32_t) const { return lab[w] == c; }
    __device__ bool self(int32_t v) const { return lab[v] == c; }
};
}  // namespace gir

extern "C" int girhip_bfs_component(girhip_graph* h, const int32_t* src, int32_t k, int32_t undirected,
                                    int32_t* levels, int64_t* visited) {
    return guarded([&] {
        useDevice(h);
        if (k < 0 || (k > 0 && !src)) raise(GIRHIP_EINVAL, "bfs_component: bad source list");
        if (undirected ? !h->wcc_req : !h->scc_req)
            raise(GIRHIP_ESTATE, undirected ? "bfs_component: buildWCC first" : "bfs_component: buildSCC_DAG first");
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
        // push / pull CSR: out- and in-arcs (directed), the symmetric closure
        // for both (undirected)
        const int32_t *pb = h->fw_beg.p, *pa = h->fw_adj.p, *rb = h->bw_beg.p, *ra = h->bw_adj.p;
        int64_t m_total = h->E;
        if (undirected) {
            m_total = symCsr(h);
            pb = rb = h->sym_beg.p;
            pa = ra = h->sym_adj.p;
        }
        const int32_t* lab = undirected ? h->wcc.p : h->scc.p;
        ReachBufs B = reachBufs(h);
        std::vector<int64_t> vis(k, 0);
        tBegin(h);
        fillI32(h, out, -1, (int64_t)n);
        for (int32_t i = 0; i < k; ++i) {
            const int32_t x = s[i];
            if (x < 0 || x >= V) continue;
            int32_t c = 0;
            int32_t deg[2];
            HIPCHK(hipMemcpyAsync(&c, lab + x, sizeof(int32_t), hipMemcpyDeviceToHost, h->st));
            HIPCHK(hipMemcpyAsync(deg, pb + x, sizeof(int32_t) * 2, hipMemcpyDeviceToHost, h->st));
            syncStream(h);
            fillU64(h, B.vis, 0, B.W);
            fillI32(h, B.q[0], x, 1);
            const ReachStats st = runReach(h, pb, pa, rb, ra, B, 1, (int64_t)deg[1] - deg[0], (int64_t)V, m_total,
                                           SameLabel{lab, c}, WriteLevel{out + (size_t)i * V});
            vis[i] = st.reached;
        }
        tEnd(h, "bfs_component");
        if (visited) HIPCHK(hipMemcpy(visited, vis.data(), sizeof(int64_t) * (size_t)k, hipMemcpyDefault));
        if (levels && !direct) copyOut(h, levels, out, sizeof(int32_t) * n);
        syncStream(h);
    });
}

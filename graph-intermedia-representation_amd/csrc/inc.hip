// Incremental single-source BFS / SSSP: replaces GIR::trackIncremental,
// GIR::updateIncremental and GIR::incDist (gir_query.cpp:867-996,
// gir.h:222-233), the KickStarter-style dependence trees.
//
// State per kind (0: BFS levels, 1: SSSP distances with the weights
// girhip_sssp uses): st[v] = (distance << 32) | parent, ~0 = unreachable,
// the parent being an in-neighbour on a shortest path.  Every change the
// edge store takes while a kind is tracked is appended to its log.  An
// update (repeated log entries are harmless: invalidation and queueing go
// through flag CAS, relaxation through atomicMin):
//   1. an arc of the log absent now whose head's parent is its tail (a
//      deleted tree arc) invalidates the head;
//   2. invalidation spreads down the dependence tree (children: out-arcs
//      whose head names the vertex as its parent), level by level;
//   3. every invalid vertex restarts from its best valid in-neighbour;
//   4. seeds: the restarted vertices with a distance, and the heads that
//      arcs present now (insertions) improve;
//   5. frontier Bellman-Ford from the seeds: atomicMin on the packed word,
//      an improved vertex queued once per round.
// The corrected values equal a from-scratch BFS / SSSP, which is the
// reference's own check (gir_stream.cpp:54-67).  Distances are kept in 32
// bits, as in the reference's packed word; a larger one is an error.
#include <algorithm>
#include <vector>

#include "common.h"
#include "lists.h"

namespace gir {
namespace {

constexpr uint64_t NONE = ~0ULL;
constexpr int INC_RING = 64;   // per-round queue counters
constexpr int INC_CHUNK = 8;   // rounds / levels per host readback
// control slots: [0, T_SLOTS) the invalidation list, then the round ring
constexpr int C_RING = 16;
constexpr int C_OVF = C_RING + INC_RING;
constexpr int C_NU = C_OVF + 1;  // log length (device copy)
constexpr int C_TOTAL = C_NU + 16;

__device__ __forceinline__ int32_t parentOf(uint64_t k) { return (int32_t)(uint32_t)(k & 0xffffffffULL); }
__device__ __forceinline__ uint64_t pack(uint64_t d, int32_t p) { return (d << 32) | (uint32_t)p; }

// Arc weights of one kind: 1 (BFS), the stored per-slot weights, or the
// reference's hash; arc presence by binary search in the sorted fw row.
struct IncG {
    const int32_t* fb;
    const int32_t* fa;
    const uint32_t* w;
    int unit;
    __device__ int64_t slot(int32_t u, int32_t v) const {
        int64_t lo = fb[u], hi = fb[u + 1];
        while (lo < hi) {
            const int64_t m = (lo + hi) >> 1;
            const int32_t x = fa[m];
            if (x == v) return m;
            if (x < v) lo = m + 1;
            else hi = m;
        }
        return -1;
    }
    // j: the fw slot of (u, v) when known, else -1
    __device__ uint64_t weight(int32_t u, int32_t v, int64_t j) const {
        if (unit) return 1;
        if (!w) return hashWeight(u, v);
        if (j < 0) j = slot(u, v);
        return w[j];
    }
};

__global__ void kIncLevels(const int32_t* __restrict__ lv, int64_t n, uint64_t* __restrict__ d) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        d[i] = lv[i] < 0 ? NONE : (uint64_t)lv[i];
}

// Tracking: the packed word of every vertex from from-scratch distances;
// the parent is the first in-neighbour on a shortest path.
__global__ void kIncParents(IncG g, const int32_t* __restrict__ bb, const int32_t* __restrict__ ba,
                            const uint64_t* __restrict__ dist, int32_t V, int32_t src, uint64_t* __restrict__ st,
                            int64_t* ovf) {
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < V; v += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t d = dist[v];
        if (d == NONE) {
            st[v] = NONE;
            continue;
        }
        if (d >= 0xffffffffull) {
            *ovf = 1;
            st[v] = NONE;
            continue;
        }
        if (v == src) {
            st[v] = pack(0, src);
            continue;
        }
        int32_t par = -1;
        for (int64_t e = bb[v]; e < bb[v + 1]; ++e) {
            const int32_t p = ba[e];
            const uint64_t dp = dist[p];
            if (dp != NONE && dp + g.weight(p, (int32_t)v, -1) == d) {
                par = p;
                break;
            }
        }
        st[v] = pack(d, par);
    }
}

// Step 1: presence of every logged arc now; deleted tree arcs invalidate
// their heads (flag 1, appended to the list), and so do tree arcs present
// now with another weight than the tree's distances imply (an arc deleted
// and inserted again with a new explicit weight between two updates).
__global__ void kIncDeleted(IncG g, const uint64_t* __restrict__ keys, const int64_t* __restrict__ nkeys, int B,
                            int32_t src, const uint64_t* __restrict__ st, int32_t* flag, uint8_t* __restrict__ pres,
                            int32_t* list, int64_t* ctl) {
    const int64_t n = *nkeys;
    const uint64_t mask = (1ULL << B) - 1;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t base = (int64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~(WAVE - 1)); base < n; base += stride) {
        const int64_t i = base + laneId();
        bool inv = false;
        if (i < n) {
            const int32_t u = (int32_t)(keys[i] >> B), v = (int32_t)(keys[i] & mask);
            const int64_t j = g.slot(u, v);
            const bool present = j >= 0;
            pres[i] = present ? 1 : 0;
            if (v != src) {
                const uint64_t k = st[v];
                bool stale = false;
                if (k != NONE && parentOf(k) == u) {
                    const uint64_t ku = st[u];
                    stale = !present || ku == NONE || (ku >> 32) + g.weight(u, v, j) != (k >> 32);
                }
                if (stale) inv = atomicCAS(&flag[v], 0, 1) == 0;
            }
        }
        const int64_t pos = wavePush(ctl + T_TAIL, inv);
        if (inv) list[pos] = (int32_t)(keys[i] & mask);
    }
}

// Step 2: out-arc j of invalid x; a child (parent = x) becomes invalid.
struct IncSpreadF {
    const int32_t* fa;
    const uint64_t* st;
    int32_t* flag;
    int32_t* list;
    int64_t* tail;
    int32_t src;
    __device__ int64_t operator()(int32_t x, uint64_t, int32_t, int64_t j) const {
        const int32_t y = fa[j];
        bool inv = false;
        if (y != src) {
            const uint64_t k = st[y];
            if (k != NONE && parentOf(k) == x) inv = atomicCAS(&flag[y], 0, 1) == 0;
        }
        const int64_t pos = wavePush(tail, inv);
        if (inv) list[pos] = y;
        return 0;
    }
};

__global__ void kIncClear(const int32_t* __restrict__ list, const int64_t* __restrict__ ctl, uint64_t* st) {
    const int64_t n = ctl[T_TAIL];
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        st[list[i]] = NONE;
}

// Step 3: in-arc j of invalid x (bw CSR): a valid in-neighbour with a
// distance offers x its distance plus the arc.
struct IncRestartF {
    IncG g;
    const int32_t* ba;
    const int32_t* flag;
    uint64_t* st;
    int64_t* ovf;
    __device__ int64_t operator()(int32_t x, uint64_t, int32_t, int64_t j) const {
        const int32_t p = ba[j];
        if (flag[p] == 1) return 0;
        const uint64_t k = st[p];
        if (k == NONE) return 0;
        const uint64_t d = (k >> 32) + g.weight(p, x, -1);
        if (d >= 0xffffffffull) {
            *ovf = 1;
            return 0;
        }
        atomicMin((unsigned long long*)&st[x], (unsigned long long)pack(d, p));
        return 0;
    }
};

// Step 4a: the invalid vertices lose their flag; those that restarted with
// a distance are seeds (flag 2, queue of round 0).
__global__ void kIncReseed(const int32_t* __restrict__ list, const int64_t* __restrict__ ctl, const uint64_t* st,
                           int32_t* flag, int32_t* q, int64_t* qn) {
    const int64_t n = ctl[T_TAIL];
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t base = (int64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~(WAVE - 1)); base < n; base += stride) {
        const int64_t i = base + laneId();
        bool seed = false;
        int32_t x = 0;
        if (i < n) {
            x = list[i];
            seed = st[x] != NONE;
            flag[x] = seed ? 2 : 0;
        }
        const int64_t pos = wavePush(qn, seed);
        if (seed) q[pos] = x;
    }
}

// Step 4b: arcs present now improve their heads.
__global__ void kIncInserted(IncG g, const uint64_t* __restrict__ keys, const int64_t* __restrict__ nkeys, int B,
                             const uint8_t* __restrict__ pres, uint64_t* st, int32_t* flag, int32_t* q, int64_t* qn,
                             int64_t* ovf) {
    const int64_t n = *nkeys;
    const uint64_t mask = (1ULL << B) - 1;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t base = (int64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~(WAVE - 1)); base < n; base += stride) {
        const int64_t i = base + laneId();
        bool push = false;
        int32_t v = 0;
        if (i < n && pres[i]) {
            const int32_t u = (int32_t)(keys[i] >> B);
            v = (int32_t)(keys[i] & mask);
            const uint64_t k = st[u];
            if (k != NONE) {
                const uint64_t d = (k >> 32) + g.weight(u, v, -1);
                if (d >= 0xffffffffull) {
                    *ovf = 1;
                } else {
                    const uint64_t c = pack(d, u);
                    const uint64_t old = atomicMin((unsigned long long*)&st[v], (unsigned long long)c);
                    push = c < old && atomicCAS(&flag[v], 0, 2) == 0;
                }
            }
        }
        const int64_t pos = wavePush(qn, push);
        if (push) q[pos] = v;
    }
}

// Step 5: out-arc j of queued x (distance in pay).
struct IncRelaxF {
    IncG g;
    uint64_t* st;
    int32_t* flag;
    int32_t* qout;
    int64_t* qn;
    int64_t* ovf;
    __device__ int64_t operator()(int32_t x, uint64_t dx, int32_t, int64_t j) const {
        const int32_t y = g.fa[j];
        const uint64_t d = dx + g.weight(x, y, j);
        bool push = false;
        if (d >= 0xffffffffull) {
            *ovf = 1;
        } else {
            const uint64_t c = pack(d, x);
            if (c < st[y]) {  // plain read: the word only falls
                const uint64_t old = atomicMin((unsigned long long*)&st[y], (unsigned long long)c);
                push = c < old && atomicCAS(&flag[y], 0, 2) == 0;
            }
        }
        const int64_t pos = wavePush(qn, push);
        if (push) qout[pos] = y;
        return 0;
    }
};

// Round r: the queue of parity r & 1 (length in ring slot r - 1) relaxes
// into the other queue (ring slot r); slot r + 1 is zeroed for the next
// round.  A vertex clears its flag before reading its distance (a full
// fence between), so an improvement that lands meanwhile requeues it.
__global__ void kIncRound(IncRelaxF f, const int32_t* __restrict__ fb, const int32_t* qin, const int64_t* nin,
                          int64_t* zero_next, HeavyList H) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *zero_next = 0;
    const int64_t n = *nin;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t base = (int64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~(WAVE - 1)); base < n; base += stride) {
        const int64_t i = base + laneId();
        const bool act = i < n;
        const int32_t x = act ? qin[i] : 0;
        uint64_t dx = 0;
        if (act) {
            __hip_atomic_exchange(&f.flag[x], 0, __ATOMIC_SEQ_CST, __HIP_MEMORY_SCOPE_AGENT);
            dx = __hip_atomic_load(&f.st[x], __ATOMIC_SEQ_CST, __HIP_MEMORY_SCOPE_AGENT) >> 32;
        }
        const int64_t b = act ? fb[x] : 0, e = act ? fb[x + 1] : 0;
        waveRowsH(act, x, dx, 0, b, e, H, f);
    }
}

__global__ void kIncDist(const uint64_t* __restrict__ st, int64_t n, uint64_t* __restrict__ out) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        out[i] = st[i] == NONE ? NONE : st[i] >> 32;
}

__global__ void kIncAppend(const uint64_t* __restrict__ in, int64_t n, uint64_t* __restrict__ out) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        out[i] = in[i];
}

IncG incGraph(girhip_graph* h, int weighted) {
    return IncG{h->fw_beg.p, h->fw_adj.p, (weighted && h->has_w) ? h->w_fw.p : nullptr, weighted ? 0 : 1};
}

}  // namespace

void incLog(girhip_graph* h, const uint64_t* keys, int64_t n) {
    if (n <= 0) return;
    for (auto& T : h->inc) {
        if (!T.on) continue;
        const int64_t need = T.nlog + n;
        if ((int64_t)T.log.n < need) {  // grow, keeping the entries so far
            DevBuf<uint64_t> g;
            g.ensure((size_t)std::max<int64_t>(2 * need, 1024));
            if (T.nlog > 0)
                launch(h, "inc_log_copy", kIncAppend, boundedGrid(T.nlog, 256), 256, 0, (const uint64_t*)T.log.p,
                       T.nlog, g.p);
            syncStream(h);  // the old buffer is freed by the swap
            T.log.swap(g);
        }
        launch(h, "inc_log", kIncAppend, boundedGrid(n, 256), 256, 0, keys, n, T.log.p + T.nlog);
        T.nlog = need;
    }
}

}  // namespace gir

using namespace gir;

extern "C" int girhip_inc_track(girhip_graph* h, int32_t source, int32_t weighted) {
    return guarded([&] {
        useDevice(h);
        if (weighted != 0 && weighted != 1) raise(GIRHIP_EINVAL, "inc_track: weighted must be 0 or 1");
        const int32_t V = h->V;
        auto& T = h->inc[weighted];
        T.on = false;
        T.st.ensure((size_t)std::max<int32_t>(V, 1));
        Scratch& S = scratch(h);
        S.inc_d.ensure((size_t)std::max<int32_t>(V, 1));
        S.inc_ctl.ensure(C_TOTAL);
        // from-scratch distances through the query entry points
        if (weighted) {
            if (girhip_sssp(h, &source, 1, 0, S.inc_d.p, nullptr) != GIRHIP_OK) raise(GIRHIP_EHIP, girhip_last_error());
        } else {
            S.inc_lv.ensure((size_t)std::max<int32_t>(V, 1));
            if (girhip_bfs(h, &source, 1, S.inc_lv.p, nullptr) != GIRHIP_OK) raise(GIRHIP_EHIP, girhip_last_error());
            launch(h, "inc_levels", kIncLevels, boundedGrid(V, 256), 256, 0, (const int32_t*)S.inc_lv.p, (int64_t)V,
                   S.inc_d.p);
        }
        fillI64(h, S.inc_ctl.p + C_OVF, 0, 1);
        launch(h, "inc_parents", kIncParents, boundedGrid(V, 256), 256, 0, incGraph(h, weighted),
               (const int32_t*)h->bw_beg.p, (const int32_t*)h->bw_adj.p, (const uint64_t*)S.inc_d.p, V, source,
               T.st.p, S.inc_ctl.p + C_OVF);
        int64_t ovf = 0;
        readCounters(h, S.inc_ctl.p + C_OVF, 1, &ovf);
        if (ovf) raise(GIRHIP_EINVAL, "inc_track: a distance does not fit the 32-bit packed word");
        T.source = source;
        T.nlog = 0;
        T.on = true;
    });
}

extern "C" int girhip_inc_update(girhip_graph* h, int32_t weighted) {
    return guarded([&] {
        useDevice(h);
        if (weighted != 0 && weighted != 1) raise(GIRHIP_EINVAL, "inc_update: weighted must be 0 or 1");
        auto& T = h->inc[weighted];
        if (!T.on) raise(GIRHIP_ESTATE, "inc_update: call girhip_inc_track first");
        const int32_t V = h->V;
        if (T.nlog == 0 || T.source < 0 || T.source >= V) {
            T.nlog = 0;
            return;
        }
        const int blk = (int)h->prm.block;
        const int B = bitsFor((uint64_t)(V > 1 ? V - 1 : 1));
        Scratch& S = scratch(h);
        const int64_t n = T.nlog;
        S.inc_pres.ensure((size_t)n);
        S.inc_list.ensure((size_t)std::max<int32_t>(V, 1));
        for (int p = 0; p < 2; ++p) S.inc_q[p].ensure((size_t)std::max<int32_t>(V, 1));
        if (S.inc_flag.n < (size_t)std::max<int32_t>(V, 1)) {  // all zero between updates
            S.inc_flag.ensure((size_t)std::max<int32_t>(V, 1));
            fillI32(h, S.inc_flag.p, 0, V);
        }
        S.inc_ctl.ensure(C_TOTAL);
        int64_t* ctl = S.inc_ctl.p;
        tBegin(h);
        fillI64(h, ctl, 0, C_TOTAL);
        fillI64(h, ctl + C_NU, n, 1);
        const uint64_t* keys = T.log.p;
        const IncG g = incGraph(h, weighted);
        const unsigned ng = boundedGrid(n, blk);
        // 1. deleted tree arcs
        launch(h, "inc_deleted", kIncDeleted, ng, blk, 0, g, keys, (const int64_t*)(ctl + C_NU), B, T.source,
               (const uint64_t*)T.st.p, S.inc_flag.p, S.inc_pres.p, S.inc_list.p, ctl);
        launch(h, "inc_advance", kListAdvance, 1, 64, 0, ctl);
        // 2. spread down the dependence tree (nothing to do without a
        // deleted tree arc)
        const IncSpreadF sf{h->fw_adj.p, T.st.p, S.inc_flag.p, S.inc_list.p, ctl + T_TAIL, T.source};
        int64_t r[T_SLOTS];
        readCounters(h, ctl, T_SLOTS, r);
        const bool any_invalid = r[T_TAIL] > 0;
        // chunks of 2, 4, then INC_CHUNK levels per readback: a small batch's
        // invalidation is a few levels deep, so a full first chunk would be
        // mostly empty launches
        int chunk = 2;
        while (any_invalid) {
            for (int t = 0; t < chunk; ++t) {
                launch(h, "inc_spread", kListRows<IncSpreadF>, 1024, blk, 0, (const int32_t*)h->fw_beg.p,
                       (const int32_t*)S.inc_list.p, (const int64_t*)ctl, sf, heavyList(h));
                launchHeavyStore(h, "inc_spread_rows", sf);
                launch(h, "inc_advance", kListAdvance, 1, 64, 0, ctl);
            }
            readCounters(h, ctl, T_SLOTS, r);
            if (r[T_LO] == r[T_HI]) break;
            chunk = std::min(INC_CHUNK, 2 * chunk);
        }
        // 3. restart the invalid vertices from valid in-neighbours
        if (any_invalid) {
            launch(h, "inc_clear", kIncClear, 1024, 256, 0, (const int32_t*)S.inc_list.p, (const int64_t*)ctl, T.st.p);
            launch(h, "inc_all", kListAll, 1, 64, 0, ctl);
            const IncRestartF rf{g, h->bw_adj.p, S.inc_flag.p, T.st.p, ctl + C_OVF};
            launch(h, "inc_restart", kListRows<IncRestartF>, 1024, blk, 0, (const int32_t*)h->bw_beg.p,
                   (const int32_t*)S.inc_list.p, (const int64_t*)ctl, rf, heavyList(h));
            launchHeavyStore(h, "inc_restart_rows", rf);
        }
        // 4. seeds into the round-0 queue (its length in ring slot INC_RING - 1)
        int64_t* ring = ctl + C_RING;
        auto slot = [&](int64_t rr) { return ring + ((rr % INC_RING) + INC_RING) % INC_RING; };
        launch(h, "inc_reseed", kIncReseed, 1024, blk, 0, (const int32_t*)S.inc_list.p, (const int64_t*)ctl,
               (const uint64_t*)T.st.p, S.inc_flag.p, S.inc_q[0].p, slot(-1));
        launch(h, "inc_inserted", kIncInserted, ng, blk, 0, g, keys, (const int64_t*)(ctl + C_NU), B,
               (const uint8_t*)S.inc_pres.p, T.st.p, S.inc_flag.p, S.inc_q[0].p, slot(-1), ctl + C_OVF);
        // 5. rounds of relaxation
        int64_t rr = 0;
        std::vector<int64_t> rv(C_TOTAL);
        int rchunk = 2;  // as for the spread: 2, 4, then INC_CHUNK rounds per readback
        while (true) {
            for (int t = 0; t < rchunk; ++t, ++rr) {
                const IncRelaxF f{g, T.st.p, S.inc_flag.p, S.inc_q[(rr + 1) & 1].p, slot(rr), ctl + C_OVF};
                launch(h, "inc_round", kIncRound, 1024, blk, 0, f, (const int32_t*)h->fw_beg.p,
                       (const int32_t*)S.inc_q[rr & 1].p, (const int64_t*)slot(rr - 1), slot(rr + 1), heavyList(h));
                launchHeavyStore(h, "inc_round_rows", f);
            }
            readCounters(h, ctl, C_TOTAL, rv.data());
            if (rv[slot(rr - 1) - ctl] == 0) break;
            rchunk = std::min(INC_CHUNK, 2 * rchunk);
        }
        tEnd(h, weighted ? "inc_update_sssp" : "inc_update_bfs");
        if (rv[C_OVF]) raise(GIRHIP_EINVAL, "inc_update: a distance does not fit the 32-bit packed word");
        T.nlog = 0;
    });
}

extern "C" int girhip_inc_dist(girhip_graph* h, int32_t weighted, uint64_t* dist) {
    return guarded([&] {
        useDevice(h);
        if (weighted != 0 && weighted != 1) raise(GIRHIP_EINVAL, "inc_dist: weighted must be 0 or 1");
        if (!dist) raise(GIRHIP_EINVAL, "inc_dist: null output");
        auto& T = h->inc[weighted];
        if (!T.on) raise(GIRHIP_ESTATE, "inc_dist: call girhip_inc_track first");
        const int32_t V = h->V;
        Scratch& S = scratch(h);
        S.inc_d.ensure((size_t)std::max<int32_t>(V, 1));
        uint64_t* out = isDevicePtr(dist) ? dist : S.inc_d.p;
        launch(h, "inc_dist", kIncDist, boundedGrid(V, 256), 256, 0, (const uint64_t*)T.st.p, (int64_t)V, out);
        if (out != dist) copyOut(h, dist, out, sizeof(uint64_t) * (size_t)V);
        syncStream(h);
    });
}

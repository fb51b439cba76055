// Single-source shortest paths: replaces GIR::sssp / GIR::deltaStepping
// (gir_query.cpp:629-703) for k sources.
//
// Delta-stepping with device-resident buckets in near/far form: the current
// bucket [lo, lo + delta) is a queue that is relaxed until it empties (the
// reference re-processes bucket k until nothing new lands in it); an
// improvement that lands beyond the bucket goes, once (bitmap flag), to the
// far pile.  When the bucket empties, the smallest far distance picks the
// next non-empty bucket and one pass splits the far pile into that bucket's
// queue, the remaining pile, and stale entries (settled earlier).
// Relaxation is an atomicMin on 64-bit distances; duplicates in a queue are
// skipped with a per-round stamp, like the reference's round_mark.  Weights:
// one uint32 per forward slot, or the reference's hash GIR::edgeWeight.
//
// HBM traffic per source (algorithmic): 8V (init + final distances) + 8 n_s
// (row offsets) + 16 m_s (adjacency 4 + weight 4 + distance 8 per arc).
#include <vector>

#include "common.h"

namespace gir {

std::vector<int32_t> hostSources(girhip_graph* h, const int32_t* src, int32_t k);

constexpr uint64_t INF = ~0ULL;

// Relaxation of one arc slot j of row p whose owner had distance dp.
// cnt[0]: next near-queue length; cnt[1]: far-pile length (appended to).
struct RelaxF {
    const int32_t* fa;
    const uint32_t* w;
    uint64_t* dist;
    uint64_t hi;
    int32_t* qnear;
    int32_t* qfar;
    uint64_t* farbits;
    int64_t* cnt;
    __device__ int64_t operator()(int32_t p, uint64_t dp, int32_t, int64_t j) const {
        const int32_t x = fa[j];
        const uint64_t nd = dp + (w ? (uint64_t)w[j] : (uint64_t)hashWeight(p, x));
        bool near = false, far = false;
        if (nd < ldAcq(&dist[x])) {
            const uint64_t old = atomicMin((unsigned long long*)&dist[x], (unsigned long long)nd);
            if (nd < old) {
                if (nd < hi) {
                    near = true;
                } else {
                    const uint64_t bit = 1ULL << (x & 63);
                    far = !(atomicOr((unsigned long long*)&farbits[x >> 6], (unsigned long long)bit) & bit);
                }
            }
        }
        const int64_t pn = wavePush(&cnt[0], near);
        if (near) qnear[pn] = x;
        const int64_t pf = wavePush(&cnt[1], far);
        if (far) qfar[pf] = x;
        return 0;
    }
};

__global__ void kSsspRelax(const int32_t* __restrict__ qin, int64_t nin, const int32_t* __restrict__ fb,
                           int32_t* stamp, int32_t round, RelaxF f, HeavyList H) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool act = i < nin;
    const int32_t u = act ? qin[i] : 0;
    if (act && atomicExch(&stamp[u], round) == round) act = false;  // duplicate in this round
    const uint64_t du = act ? ldAcq(&f.dist[u]) : 0;
    const int32_t b = act ? fb[u] : 0, e = act ? fb[u + 1] : 0;
    waveRowsH(act, u, du, 0, b, e, H, f);
}

// Smallest distance >= hi among far entries (entries below hi are settled).
__global__ void kFarMin(const int32_t* __restrict__ qfar, int64_t n, const uint64_t* __restrict__ dist, uint64_t hi,
                        unsigned long long* out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long d = INF;
    if (i < n) {
        const uint64_t x = dist[qfar[i]];
        if (x >= hi) d = x;
    }
    for (int s = WAVE / 2; s > 0; s >>= 1) {
        unsigned long long o = __shfl_xor(d, s);
        d = o < d ? o : d;
    }
    if (laneId() == 0 && d != INF) atomicMin(out, d);
}

// cnt[0]: near-queue length, cnt[1]: kept far length.
__global__ void kFarSplit(const int32_t* __restrict__ qfar, int64_t n, const uint64_t* __restrict__ dist, uint64_t lo,
                          uint64_t hi, uint64_t* farbits, int32_t* __restrict__ qnear, int32_t* __restrict__ qkeep,
                          int64_t* cnt) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool near = false, keep = false;
    int32_t x = 0;
    if (i < n) {
        x = qfar[i];
        const uint64_t d = dist[x];
        keep = d >= hi;
        near = d >= lo && d < hi;
        if (!keep) atomicAnd((unsigned long long*)&farbits[x >> 6], ~(1ULL << (x & 63)));
    }
    const int64_t pn = wavePush(&cnt[0], near);
    if (near) qnear[pn] = x;
    const int64_t pk = wavePush(&cnt[1], keep);
    if (keep) qkeep[pk] = x;
}

__global__ void kSetCounters(int64_t* c, int64_t a, int64_t b) {
    if (threadIdx.x == 0) {
        c[0] = a;
        c[1] = b;
    }
}

__global__ void kReachedArcsD(const uint64_t* __restrict__ dist, const int32_t* __restrict__ beg, int32_t V,
                              int64_t* out) {
    int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t m = (v < V && dist[v] != INF) ? (int64_t)(beg[v + 1] - beg[v]) : 0;
    for (int d = WAVE / 2; d > 0; d >>= 1) m += __shfl_xor(m, d);
    if (laneId() == 0 && m) atomicAdd((unsigned long long*)out, (unsigned long long)m);
}

struct SsspBufs {
    int32_t *qn[2], *qf[2];
    uint64_t* farbits;
    int32_t* stamp;
    int64_t W;
};

static void ssspOne(girhip_graph* h, SsspBufs& Bf, int32_t s, uint64_t delta, uint64_t* dist, int32_t& round) {
    const int blk = (int)h->prm.block;
    const uint32_t* w = h->has_w ? h->w_fw.p : nullptr;
    fillU64(h, dist, INF, h->V);
    fillU64(h, dist + s, 0, 1);
    fillI32(h, Bf.qn[0], s, 1);
    int64_t* c = h->dcnt.p + 24;
    int64_t nnear = 1, nfar = 0;
    uint64_t lo = 0, hi = delta;
    int cn = 0, cf = 0;
    while (true) {
        while (nnear > 0) {
            ++round;
            launch(h, "sssp_setcnt", kSetCounters, 1, 64, 0, c, (int64_t)0, nfar);
            const RelaxF f{h->fw_adj.p, w, dist, hi, Bf.qn[cn ^ 1], Bf.qf[cf], Bf.farbits, c};
            launch(h, "sssp_relax", kSsspRelax, gridFor(nnear, blk), blk, 0, (const int32_t*)Bf.qn[cn], nnear,
                   (const int32_t*)h->fw_beg.p, Bf.stamp, round, f, heavyList(h));
            launchHeavy(h, "sssp_relax_heavy", f);
            int64_t r[2];
            readCounters(h, c, 2, r);
            nnear = r[0];
            nfar = r[1];
            cn ^= 1;
        }
        if (nfar == 0) break;
        fillU64(h, (uint64_t*)(c + 2), INF, 1);
        launch(h, "sssp_farmin", kFarMin, gridFor(nfar, blk), blk, 0, (const int32_t*)Bf.qf[cf], nfar,
               (const uint64_t*)dist, hi, (unsigned long long*)(c + 2));
        int64_t mn;
        readCounters(h, c + 2, 1, &mn);
        const uint64_t m = (uint64_t)mn;
        lo = (m == INF) ? INF : (m / delta) * delta;
        hi = (m == INF || lo > INF - delta) ? INF : lo + delta;
        launch(h, "sssp_setcnt", kSetCounters, 1, 64, 0, c, (int64_t)0, (int64_t)0);
        launch(h, "sssp_split", kFarSplit, gridFor(nfar, blk), blk, 0, (const int32_t*)Bf.qf[cf], nfar,
               (const uint64_t*)dist, lo, hi, Bf.farbits, Bf.qn[cn], Bf.qf[cf ^ 1], c);
        int64_t r[2];
        readCounters(h, c, 2, r);
        nnear = r[0];
        nfar = r[1];
        cf ^= 1;
        if (m == INF) break;  // every far entry was stale
    }
}

// ---------------------------------------------------------------------------
// Batched form: up to SB sources advance together, one grid row (blockIdx.y)
// per source, each with its own near queues (deduplicated at push time by a
// per-source stamp), far pile (deduplicated by a per-source bitmap) and
// bucket bounds, all in device memory.  One round relaxes every non-empty
// near queue; then every source whose near queue ran dry picks its next
// bucket from its far pile (a min pass and a split pass over the pile, grid
// stride).  The host reads the per-source queue lengths once per round, to
// size the next relax grid and to stop when every queue and pile is empty.

constexpr int SB = 32;  // sources per batch

// Per-source control block (int64 slots, SB each):
enum : int {
    C_NEAR0 = 0,          // near-queue length, parity 0
    C_NEAR1 = SB,         // near-queue length, parity 1
    C_FAR = 2 * SB,       // far-pile length
    C_KEEP = 3 * SB,      // kept far entries of a split
    C_FPAR = 4 * SB,      // which of the two far buffers holds the pile
    C_ADV = 5 * SB,       // 1: this source advances its bucket this round
    C_LO = 6 * SB,        // bucket [lo, hi) (uint64)
    C_HI = 7 * SB,
    C_FMIN = 8 * SB,      // far minimum (uint64)
    C_SLOTS = 9 * SB
};

struct SsspBatch {
    int32_t* qn[2];     // [SB][cap] near queues by parity
    int32_t* qf[2];     // [SB][V]   far piles (per-source parity in C_FPAR)
    uint64_t* farbits;  // [SB][W]
    int32_t* stamp;     // [SB][V]
    int64_t* ctl;       // C_SLOTS
    uint64_t* dist;     // [SB][V] output rows
    int64_t cap, V, W;
};

__global__ void kBInit(SsspBatch B, const int32_t* __restrict__ src, int k, uint64_t delta) {
    const int s = threadIdx.x;
    if (s >= SB) return;
    int64_t* c = B.ctl;
    for (int j = 0; j < 9; ++j) c[j * SB + s] = 0;
    c[C_HI + s] = (int64_t)delta;
    if (s < k) {
        const int32_t v = src[s];
        if (v >= 0 && v < B.V) {
            B.dist[(size_t)s * B.V + v] = 0;
            B.qn[0][(size_t)s * B.cap] = v;
            c[C_NEAR0 + s] = 1;
        }
    }
}

// Relaxation of arc slot j of row p (owner distance dp) for source s.
struct BRelaxF {
    SsspBatch B;
    int par;
    const int32_t* fa;
    const uint32_t* w;
    int32_t round;
    __device__ int64_t operator()(int32_t p, uint64_t dp, int32_t s, int64_t j) const {
        int64_t* c = B.ctl;
        uint64_t* d = B.dist + (size_t)s * B.V;
        const int32_t x = fa[j];
        const uint64_t nd = dp + (w ? (uint64_t)w[j] : (uint64_t)hashWeight(p, x));
        bool near = false, far = false;
        if (nd < ldAcq(&d[x])) {
            const uint64_t old = atomicMin((unsigned long long*)&d[x], (unsigned long long)nd);
            if (nd < old) {
                if (nd < (uint64_t)c[C_HI + s]) {
                    near = atomicExch(&B.stamp[(size_t)s * B.V + x], round) != round;
                } else {
                    const uint64_t bit = 1ULL << (x & 63);
                    far = !(atomicOr((unsigned long long*)&B.farbits[(size_t)s * B.W + (x >> 6)],
                                     (unsigned long long)bit) &
                            bit);
                }
            }
        }
        const int64_t pn = wavePush(c + (par ? C_NEAR0 : C_NEAR1) + s, near);
        if (near) B.qn[par ^ 1][(size_t)s * B.cap + pn] = x;
        const int64_t pf = wavePush(c + C_FAR + s, far);
        if (far) B.qf[c[C_FPAR + s]][(size_t)s * B.V + pf] = x;
        return 0;
    }
};

__global__ void kBRelax(BRelaxF f, const int32_t* __restrict__ fb, HeavyList H) {
    const int s = blockIdx.y;
    const SsspBatch& B = f.B;
    const int64_t nin = B.ctl[(f.par ? C_NEAR1 : C_NEAR0) + s];
    if ((int64_t)blockIdx.x * blockDim.x >= nin) return;  // whole block idle for this source
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool act = i < nin;
    const int32_t u = act ? B.qn[f.par][(size_t)s * B.cap + i] : 0;
    const uint64_t du = act ? ldAcq(&B.dist[(size_t)s * B.V + u]) : 0;
    const int32_t b = act ? fb[u] : 0, e = act ? fb[u + 1] : 0;
    waveRowsH(act, u, du, s, b, e, H, f);
}

// The consumed queue is emptied; a source advances when its next queue is
// empty and its far pile is not.
__global__ void kBPrep(SsspBatch B, int par) {
    const int s = threadIdx.x;
    if (s >= SB) return;
    int64_t* c = B.ctl;
    c[(par ? C_NEAR1 : C_NEAR0) + s] = 0;
    const bool adv = c[(par ? C_NEAR0 : C_NEAR1) + s] == 0 && c[C_FAR + s] > 0;
    c[C_ADV + s] = adv;
    c[C_FMIN + s] = (int64_t)INF;
    c[C_KEEP + s] = 0;
}

__global__ void kBFarMin(SsspBatch B) {
    const int s = blockIdx.y;
    const int64_t* c = B.ctl;
    if (!c[C_ADV + s]) return;
    const int64_t n = c[C_FAR + s];
    const int32_t* qfar = B.qf[c[C_FPAR + s]] + (size_t)s * B.V;
    const uint64_t* d = B.dist + (size_t)s * B.V;
    const uint64_t hi = (uint64_t)c[C_HI + s];
    unsigned long long m = INF;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t x = d[qfar[i]];
        if (x >= hi && x < m) m = x;
    }
    for (int o = WAVE / 2; o > 0; o >>= 1) {
        const unsigned long long t = __shfl_xor(m, o);
        m = t < m ? t : m;
    }
    if (laneId() == 0 && m != INF) atomicMin((unsigned long long*)&B.ctl[C_FMIN + s], m);
}

__global__ void kBAdvance(SsspBatch B, uint64_t delta) {
    const int s = threadIdx.x;
    if (s >= SB) return;
    int64_t* c = B.ctl;
    if (!c[C_ADV + s]) return;
    const uint64_t m = (uint64_t)c[C_FMIN + s];
    const uint64_t lo = m == INF ? INF : (m / delta) * delta;
    c[C_LO + s] = (int64_t)lo;
    c[C_HI + s] = (int64_t)((m == INF || lo > INF - delta) ? INF : lo + delta);
}

// Far pile -> next near queue (d in [lo, hi)), kept pile (d >= hi) or
// dropped (d < lo: settled).  Every entry leaving the pile clears its bit.
__global__ void kBSplit(SsspBatch B, int par) {
    const int s = blockIdx.y;
    int64_t* c = B.ctl;
    if (!c[C_ADV + s]) return;
    const int64_t n = c[C_FAR + s];
    const int fp = (int)c[C_FPAR + s];
    const int32_t* qfar = B.qf[fp] + (size_t)s * B.V;
    int32_t* qkeep = B.qf[fp ^ 1] + (size_t)s * B.V;
    int32_t* qnear = B.qn[par ^ 1] + (size_t)s * B.cap;
    uint64_t* fbits = B.farbits + (size_t)s * B.W;
    const uint64_t* d = B.dist + (size_t)s * B.V;
    const uint64_t lo = (uint64_t)c[C_LO + s], hi = (uint64_t)c[C_HI + s];
    int64_t* nnear = c + (par ? C_NEAR0 : C_NEAR1) + s;
    int64_t* nkeep = c + C_KEEP + s;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    // uniform trip count across the wave so wavePush sees the whole wave
    const int64_t base0 = (int64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~(WAVE - 1));
    for (int64_t base = base0; base < n; base += stride) {
        const int64_t i = base + laneId();
        bool near = false, keep = false;
        int32_t x = 0;
        if (i < n) {
            x = qfar[i];
            const uint64_t dx = d[x];
            keep = dx >= hi;
            near = dx >= lo && dx < hi;
            if (!keep) atomicAnd((unsigned long long*)&fbits[x >> 6], ~(1ULL << (x & 63)));
        }
        const int64_t pn = wavePush(nnear, near);
        if (near) qnear[pn] = x;
        const int64_t pk = wavePush(nkeep, keep);
        if (keep) qkeep[pk] = x;
    }
}

__global__ void kBFinish(SsspBatch B) {
    const int s = threadIdx.x;
    if (s >= SB) return;
    int64_t* c = B.ctl;
    if (!c[C_ADV + s]) return;
    c[C_FAR + s] = c[C_KEEP + s];
    c[C_FPAR + s] ^= 1;
}

// k <= SB sources (device array dsrc) into rows of dist.
static void ssspBatch(girhip_graph* h, const int32_t* dsrc, int k, uint64_t delta, uint64_t* dist, int32_t& round,
                      DevBuf<int32_t>* qn, DevBuf<int32_t>* qf, DevBuf<uint64_t>& fbits, DevBuf<int32_t>& stamp,
                      DevBuf<int64_t>& ctl) {
    const int32_t V = h->V;
    const int blk = (int)h->prm.block;
    const uint32_t* w = h->has_w ? h->w_fw.p : nullptr;
    SsspBatch B;
    B.V = V;
    B.W = ((int64_t)V + 63) / 64;
    B.cap = V;  // a vertex enters a near queue at most once per round
    for (int p = 0; p < 2; ++p) {
        B.qn[p] = qn[p].p;
        B.qf[p] = qf[p].p;
    }
    B.farbits = fbits.p;
    B.stamp = stamp.p;
    B.ctl = ctl.p;
    B.dist = dist;
    fillU64(h, dist, INF, (int64_t)k * V);
    fillU64(h, fbits.p, 0, (int64_t)SB * B.W);
    launch(h, "sssp_b_init", kBInit, 1, 64, 0, B, dsrc, k, delta);
    const unsigned fgrid = 256;  // grid-stride width of the far passes (per source)
    int par = 0;
    std::vector<int64_t> r(2 * SB);
    int64_t maxn = 1;
    while (true) {
        ++round;
        const BRelaxF f{B, par, h->fw_adj.p, w, round};
        launch(h, "sssp_b_relax", kBRelax, dim3(gridFor(maxn, blk), k), blk, 0, f, (const int32_t*)h->fw_beg.p,
               heavyList(h));
        launchHeavy(h, "sssp_b_relax_heavy", f);
        launch(h, "sssp_b_prep", kBPrep, 1, 64, 0, B, par);
        launch(h, "sssp_b_farmin", kBFarMin, dim3(fgrid, k), blk, 0, B);
        launch(h, "sssp_b_advance", kBAdvance, 1, 64, 0, B, delta);
        launch(h, "sssp_b_split", kBSplit, dim3(fgrid, k), blk, 0, B, par);
        launch(h, "sssp_b_finish", kBFinish, 1, 64, 0, B);
        par ^= 1;
        // next near lengths + far lengths
        readCounters(h, ctl.p + (par ? C_NEAR1 : C_NEAR0), SB, r.data());
        readCounters(h, ctl.p + C_FAR, SB, r.data() + SB);
        maxn = 0;
        int64_t far = 0;
        for (int s = 0; s < k; ++s) {
            maxn = std::max(maxn, r[s]);
            far += r[SB + s];
        }
        if (maxn == 0 && far == 0) break;
        if (maxn == 0) maxn = 1;  // (only far piles left: an empty relax round, then they advance)
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
        const size_t Vs = (size_t)(V > 0 ? V : 1);
        if (h->prm.sssp_batch != 0 && k > 0) {
            // all sources of a chunk advance together
            const int kc = std::min<int32_t>(k, SB);
            DevBuf<int32_t> qn[2], qf[2], stamp, dsrc;
            DevBuf<uint64_t> fbits;
            DevBuf<int64_t> ctl;
            for (int p = 0; p < 2; ++p) {
                qn[p].ensure((size_t)kc * Vs);
                qf[p].ensure((size_t)kc * Vs);
            }
            stamp.ensure((size_t)kc * Vs);
            fbits.ensure((size_t)SB * ((Vs + 63) / 64));
            ctl.ensure(C_SLOTS);
            dsrc.ensure((size_t)k);
            copyIn(h, dsrc.p, s.data(), sizeof(int32_t) * (size_t)k);
            tBegin(h);
            int32_t round = 0;
            fillI32(h, stamp.p, 0, (int64_t)kc * V);
            for (int32_t i0 = 0; i0 < k; i0 += SB) {
                const int c = std::min<int32_t>(SB, k - i0);
                ssspBatch(h, dsrc.p + i0, c, delta, out + (size_t)i0 * V, round, qn, qf, fbits, stamp, ctl);
            }
            tEnd(h, "sssp");
        } else {
        S.q[0].ensure(Vs);
        S.q[1].ensure(Vs);
        S.iv[0].ensure(Vs);
        S.iv[1].ensure(Vs);
        S.iv[2].ensure(Vs);
        S.bits[0].ensure((Vs + 63) / 64);
        SsspBufs Bf;
        // near queues may hold duplicates: size them by arcs + vertices
        const size_t qcap = (size_t)h->E + Vs;
        S.keys[0].ensure((qcap + 1) / 2);  // reinterpret 2 x int32 per element
        S.keys[1].ensure((qcap + 1) / 2);
        Bf.qn[0] = (int32_t*)S.keys[0].p;
        Bf.qn[1] = (int32_t*)S.keys[1].p;
        Bf.qf[0] = S.q[0].p;
        Bf.qf[1] = S.q[1].p;
        Bf.stamp = S.iv[2].p;
        Bf.farbits = S.bits[0].p;
        Bf.W = (int64_t)(Vs + 63) / 64;
        tBegin(h);
        fillI32(h, Bf.stamp, 0, V);
        fillU64(h, Bf.farbits, 0, Bf.W);
        int32_t round = 0;
        for (int32_t i = 0; i < k; ++i) {
            uint64_t* row = out + (size_t)i * V;
            if (s[i] < 0 || s[i] >= V) {
                fillU64(h, row, INF, V);
                continue;
            }
            ssspOne(h, Bf, s[i], delta, row, round);
        }
        tEnd(h, "sssp");
        }
        if (edges_out) {
            std::vector<int64_t> e(k);
            for (int32_t i0 = 0; i0 < k; i0 += 32) {
                const int c = std::min<int32_t>(32, k - i0);
                fillI64(h, h->dcnt.p + 32, 0, 32);
                for (int32_t i = i0; i < i0 + c; ++i)
                    launch(h, "reached_arcs", kReachedArcsD, gridFor(V, 256), 256, 0,
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

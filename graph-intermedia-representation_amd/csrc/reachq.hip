// Reachability queries on the condensation: replaces GIR::reachable
// (gir_query.cpp:789-828) for a batch of (u, v) pairs.
//
// As in the reference, a pair inside one SCC is reachable and a pair whose
// source SCC ranks after the target SCC is not (ranks are a topological
// order of the DAG); the rest need a search of the DAG, pruned by rank: an
// SCC ranked after the target is never entered.  The reference runs a
// bidirectional search per pair.  Here each open pair first gets a wave and
// a search held in LDS; the few pairs whose search outgrows it share
// bit-parallel forward searches, 64 pairs per sweep (bit q = pair q, as in
// the multi-source BFS).  The open pairs are sorted by target rank, so the bits allowed
// at an SCC of rank r are a suffix of the 64 and one binary search gives the
// mask.  A pair is answered when its bit reaches its target SCC; a sweep ends
// when every pair is answered or the frontier is empty.
#include <algorithm>
#include <vector>

#include "common.h"

namespace gir {

namespace {

// Trivial answers and the data of the open pairs: out[i] = 1 / 0 when known,
// 2 when a search is needed (then a[i], b[i] = source / target SCC).
__global__ void kReachTrivial(const int32_t* __restrict__ u, const int32_t* __restrict__ v, int64_t n,
                              const int32_t* __restrict__ scc, const int64_t* __restrict__ rank, int32_t V,
                              uint8_t* __restrict__ out, int32_t* __restrict__ a, int32_t* __restrict__ b) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t x = u[i], y = v[i];
    uint8_t r;
    if (x < 0 || y < 0 || x >= V || y >= V) r = 0;
    else if (x == y) r = 1;
    else {
        const int32_t sa = scc[x], sb = scc[y];
        if (sa == sb) r = 1;
        else if (rank[sa] > rank[sb]) r = 0;
        else {
            r = 2;
            a[i] = sa;
            b[i] = sb;
        }
    }
    out[i] = r;
}

// ---- local searches ---------------------------------------------------------
// Most open pairs resolve inside a small rank window, as the reference's
// per-pair bidirectional search shows.  First every open pair gets one wave
// and a bidirectional search held entirely in LDS: forward from the source
// SCC over DAG out-arcs, backward from the target SCC over in-arcs, both
// restricted to the rank window [rank(source), rank(target)] (ranks rise
// along every DAG arc, so every path from source to target stays inside it).
// Each side keeps a visited hash set and a FIFO.  A discovery already in the
// other side's set answers 1; a side whose FIFO empties with nothing deferred
// answers 0, because every path from the source (or into the target) has to
// leave that side's visited set through an arc it has already scanned.
// SCCs with more than LOC_ROW arcs on the expanding side (the giant component
// of a power-law graph) are set aside, still marked visited, for the other
// side to meet.  When both sides are left with only such rows, or a side
// outgrows LDS, the pair stays open (2) for the bit-parallel sweeps below.
constexpr int LOC_CAP = 512;   // visited-set slots per side (power of two)
constexpr int LOC_Q = 384;     // FIFO entries per side; keeps the set under 7/8 full
constexpr int LOC_DEF = 8;     // rows set aside per side
constexpr int LOC_ROW = 2048;  // rows longer than this are set aside

struct LocalSide {
    int32_t set[LOC_CAP];
    int32_t fifo[LOC_Q];
    int32_t tail;
    int32_t ndef;
};
struct LocalLds {
    LocalSide side[2];
    int32_t state;  // 0 running, 1 met, 2 outgrown
};

__device__ __forceinline__ int localHash(int32_t z) { return (int)(((uint32_t)z * 2654435761u) >> (32 - 9)); }

// 1: inserted, 0: present already, -1: no free slot
__device__ __forceinline__ int localInsert(LocalSide& S, int32_t z) {
    int hsh = localHash(z);
    for (int probe = 0; probe < LOC_CAP; ++probe) {
        const int32_t old = atomicCAS(&S.set[hsh], -1, z);
        if (old == -1) return 1;
        if (old == z) return 0;
        hsh = (hsh + 1) & (LOC_CAP - 1);
    }
    return -1;
}

__device__ __forceinline__ bool localHas(const LocalSide& S, int32_t z) {
    int hsh = localHash(z);
    for (int probe = 0; probe < LOC_CAP; ++probe) {
        const int32_t k = S.set[hsh];
        if (k == z) return true;
        if (k == -1) return false;
        hsh = (hsh + 1) & (LOC_CAP - 1);
    }
    return false;
}

__global__ void __launch_bounds__(256) kReachLocal(const int32_t* __restrict__ obeg, const int32_t* __restrict__ oadj,
                                                   const int32_t* __restrict__ ibeg, const int32_t* __restrict__ iadj,
                                                   const int64_t* __restrict__ rank, int64_t n,
                                                   const int32_t* __restrict__ a, const int32_t* __restrict__ b,
                                                   uint8_t* ans) {
    __shared__ LocalLds lds[256 / WAVE];
    LocalLds& L = lds[threadIdx.x / WAVE];
    const int lane = laneId();
    const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / WAVE;
    const int64_t nw = ((int64_t)gridDim.x * blockDim.x) / WAVE;
    for (int64_t i = wave; i < n; i += nw) {
        if (ans[i] != 2) continue;  // (wave-uniform)
        const int32_t sa = a[i], sb = b[i];
        const int64_t ra = rank[sa], rb = rank[sb];
        for (int t = lane; t < LOC_CAP; t += WAVE) {
            L.side[0].set[t] = -1;
            L.side[1].set[t] = -1;
        }
        waveSync();
        if (lane == 0) {
            L.state = 0;
            L.side[0].fifo[0] = sa;
            L.side[1].fifo[0] = sb;
            for (int s = 0; s < 2; ++s) {
                L.side[s].tail = 1;
                L.side[s].ndef = 0;
            }
            localInsert(L.side[0], sa);
            localInsert(L.side[1], sb);
        }
        waveSync();
        int head[2] = {0, 0};
        int result;
        while (true) {
            if (L.state != 0) {
                result = L.state == 1 ? 1 : 2;
                break;
            }
            const int p0 = L.side[0].tail - head[0], p1 = L.side[1].tail - head[1];
            if ((p0 == 0 && L.side[0].ndef == 0) || (p1 == 0 && L.side[1].ndef == 0)) {
                result = 0;
                break;
            }
            if (p0 == 0 && p1 == 0) {  // only set-aside rows left on both sides
                result = 2;
                break;
            }
            const int s = p1 == 0 || (p0 != 0 && p0 <= p1) ? 0 : 1;
            LocalSide& me = L.side[s];
            const LocalSide& other = L.side[s ^ 1];
            const int32_t* beg = s ? ibeg : obeg;
            const int32_t* adj = s ? iadj : oadj;
            const int take = min(WAVE, (s ? p1 : p0));
            const bool own = lane < take;
            const int32_t x = own ? me.fifo[head[s] + lane] : 0;
            int64_t rb0 = own ? beg[x] : 0, re0 = own ? beg[x + 1] : 0;
            if (re0 - rb0 > LOC_ROW) {
                const int pos = atomicAdd(&me.ndef, 1);
                if (pos >= LOC_DEF) L.state = 2;
                rb0 = re0 = 0;
            }
            head[s] += take;
            const int64_t cnt = re0 - rb0;
            int64_t incl = cnt;
            for (int d = 1; d < WAVE; d <<= 1) {
                const int64_t o = __shfl_up(incl, d);
                if (lane >= d) incl += o;
            }
            const int64_t total = __shfl(incl, WAVE - 1);
            for (int64_t u0 = 0; u0 < total; u0 += WAVE) {
                const int64_t tt = u0 + lane;
                int below = 0;
                for (int step = WAVE / 2; step >= 1; step >>= 1)
                    if (__shfl(incl, below + step - 1) <= tt) below += step;
                const int64_t ex = __shfl(incl, below) - __shfl(cnt, below);
                const int64_t ob = __shfl(rb0, below);
                if (tt < total) {
                    const int32_t z = adj[ob + (tt - ex)];
                    const int64_t rz = rank[z];
                    if (rz >= ra && rz <= rb) {
                        if (localHas(other, z)) {
                            L.state = 1;
                        } else {
                            const int ins = localInsert(me, z);
                            if (ins > 0) {
                                const int pos = atomicAdd(&me.tail, 1);
                                if (pos < LOC_Q) me.fifo[pos] = z;
                                else L.state = 2;
                            } else if (ins < 0) {
                                L.state = 2;
                            }
                        }
                    }
                }
                waveSync();
                if (L.state != 0) break;
            }
            waveSync();
        }
        if (lane == 0) ans[i] = (uint8_t)result;
        waveSync();
    }
}

// One group of up to RG sweeps runs together (grid row y = sweep); each sweep
// answers up to 64 pairs.  Per sweep y: seen / next / target words per SCC
// id, two frontier buffers, a touched list, and counters.
constexpr int RG = 16;

struct ReachGroup {
    const int32_t* dbeg;  // DAG CSR by source SCC id
    const int32_t* ddst;
    const int64_t* rank;
    const int64_t* lim;   // [RG][64] target ranks of the sweep's pairs, ascending (padding INT64_MAX)
    uint64_t* seen;       // [RG][V]
    uint64_t* next;       // [RG][V]
    uint64_t* tgt;        // [RG][V] bits whose target is the SCC
    int32_t* fq[2];       // [RG][V] frontier SCC ids by level parity
    uint64_t* fb[2];      // [RG][V] their bits
    int32_t* tq;          // [RG][V] touched SCCs of the level
    int64_t* ctl;         // [RG][8]: nf[2], nt[2], found, all
    int64_t V;
};
enum : int { R_NF = 0, R_NT = 2, R_FOUND = 4, R_ALL = 5, R_SLOTS = 8 };

// Bits allowed at an SCC of rank r: pairs whose target rank is >= r.
__device__ __forceinline__ uint64_t allowed(const int64_t* lim, int64_t r) {
    int lo = 0, hi = 64;  // first index with lim >= r
    while (lo < hi) {
        const int m = (lo + hi) >> 1;
        if (lim[m] < r) lo = m + 1;
        else hi = m;
    }
    return lo >= 64 ? 0 : (~0ULL << lo);
}

__device__ __forceinline__ bool sweepDone(const ReachGroup& R, int y) {
    const int64_t* c = R.ctl + (size_t)y * R_SLOTS;
    return (uint64_t)c[R_FOUND] == (uint64_t)c[R_ALL];
}

// Push work on DAG arc j of SCC x for sweep y (= aux): the head's next word
// gets the allowed new bits; the first to arrive lists the head.
struct ReachPushF {
    ReachGroup R;
    int par;
    __device__ int64_t operator()(int32_t, uint64_t fp, int32_t y, int64_t j) const {
        const int32_t z = R.ddst[j];
        const size_t o = (size_t)y * R.V + z;
        const uint64_t m = fp & allowed(R.lim + (size_t)y * 64, R.rank[z]) & ~(R.seen[o] | R.next[o]);
        const bool first = m && atomicOr((unsigned long long*)&R.next[o], (unsigned long long)m) == 0;
        const int64_t pos = wavePushAny(R.ctl + (size_t)y * R_SLOTS + R_NT + par, first);  // heavy-pass waves mix sweeps
        if (first) R.tq[(size_t)y * R.V + pos] = z;
        return 0;
    }
};

// Level push of sweep y: frontier of parity par; also zeroes the next
// level's counters.  Device-read sizes: a finished sweep costs nothing.
__global__ void kRPush(ReachPushF f, HeavyList H) {
    const ReachGroup& R = f.R;
    const int y = blockIdx.y, par = f.par;
    int64_t* c = R.ctl + (size_t)y * R_SLOTS;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        c[R_NF + (par ^ 1)] = 0;
        c[R_NT + (par ^ 1)] = 0;
    }
    if (sweepDone(R, y)) return;
    const int64_t n = c[R_NF + par];
    const int32_t* fq = R.fq[par] + (size_t)y * R.V;
    const uint64_t* fbits = R.fb[par] + (size_t)y * R.V;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t base = (int64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~(WAVE - 1)); base < n; base += stride) {
        const int64_t i = base + laneId();
        const bool act = i < n;
        const int32_t x = act ? fq[i] : 0;
        const uint64_t fx = act ? fbits[i] : 0;
        const int32_t b = act ? R.dbeg[x] : 0, e = act ? R.dbeg[x + 1] : 0;
        waveRowsH(act && fx != 0, x, fx, y, b, e, H, f);
    }
}

// Level finish of sweep y: touched SCCs' new bits -> seen, answered bits,
// next frontier (parity par ^ 1).
__global__ void kRFinish(ReachGroup R, int par) {
    const int y = blockIdx.y;
    int64_t* c = R.ctl + (size_t)y * R_SLOTS;
    // one read for the whole block (other blocks raise found meanwhile, and
    // blockReserve needs every thread of the block)
    __shared__ int done;
    if (threadIdx.x == 0) done = sweepDone(R, y) ? 1 : 0;
    __syncthreads();
    if (done) return;
    const int64_t n = c[R_NT + par];
    const size_t o = (size_t)y * R.V;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    uint64_t hit = 0;
    for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < n; base += stride) {
        const int64_t i = base + threadIdx.x;
        uint64_t nb = 0;
        int32_t z = 0;
        if (i < n) {
            z = R.tq[o + i];
            nb = R.next[o + z] & ~R.seen[o + z];
            R.next[o + z] = 0;
            R.seen[o + z] |= nb;
            hit |= nb & R.tgt[o + z];
        }
        const int64_t pos = blockReserve(nb != 0, c + R_NF + (par ^ 1));
        if (nb) {
            R.fq[par ^ 1][o + pos] = z;
            R.fb[par ^ 1][o + pos] = nb;
        }
    }
    for (int k = WAVE / 2; k > 0; k >>= 1) hit |= __shfl_xor(hit, k);
    if (laneId() == 0 && hit) atomicOr((unsigned long long*)(c + R_FOUND), (unsigned long long)hit);
}

// Seeds of sweep y (k[y] pairs): sources with their bits, target masks.
__global__ void kRSeed(ReachGroup R, const int32_t* __restrict__ a, const int32_t* __restrict__ b,
                       const int32_t* __restrict__ k) {
    const int y = blockIdx.y, q = threadIdx.x;
    const int ky = k[y];
    int64_t* c = R.ctl + (size_t)y * R_SLOTS;
    if (q == 0) {
        for (int j = 0; j < R_SLOTS; ++j) c[j] = 0;
        c[R_NF] = ky;
        c[R_ALL] = (int64_t)(ky >= 64 ? ~0ULL : ((1ULL << ky) - 1));
    }
    if (q >= ky) return;
    const size_t o = (size_t)y * R.V;
    const int32_t sa = a[y * 64 + q], sb = b[y * 64 + q];
    atomicOr((unsigned long long*)&R.seen[o + sa], 1ULL << q);
    atomicOr((unsigned long long*)&R.tgt[o + sb], 1ULL << q);
    R.fq[0][o + q] = sa;
    R.fb[0][o + q] = 1ULL << q;
}

// Clears sweep y's marks: seen words of the SCCs it reached are cleared by
// a fill of the group; targets here.
__global__ void kRClearTargets(ReachGroup R, const int32_t* __restrict__ b, const int32_t* __restrict__ k) {
    const int y = blockIdx.y, q = threadIdx.x;
    if (q < k[y]) R.tgt[(size_t)y * R.V + b[y * 64 + q]] = 0;
}

}  // namespace
}  // namespace gir

using namespace gir;

extern "C" int girhip_reachable(girhip_graph* h, const int32_t* u, const int32_t* v, int64_t n, uint8_t* out) {
    return guarded([&] {
        useDevice(h);
        if (n < 0 || (n > 0 && (!u || !v || !out))) raise(GIRHIP_EINVAL, "reachable: bad arguments");
        if (!h->scc_req) raise(GIRHIP_ESTATE, "reachable: call girhip_scc_dag first");
        // like the reference (gir_query.cpp:790), pending ingested changes are
        // applied first
        if (h->pending && girhip_sync(h) != GIRHIP_OK) raise(GIRHIP_EHIP, girhip_last_error());
        if (n == 0) return;
        const int32_t V = h->V;
        const int blk = (int)h->prm.block;
        Scratch& S = scratch(h);
        tBegin(h);
        // pairs and answers on the device
        S.b_u.ensureSlack((size_t)n);
        S.b_v.ensureSlack((size_t)n);
        S.b_fi.ensureSlack((size_t)n);
        S.b_fd.ensureSlack((size_t)n);
        S.b_ins.ensureSlack((size_t)n);
        copyIn(h, S.b_u.p, u, sizeof(int32_t) * (size_t)n);
        copyIn(h, S.b_v.p, v, sizeof(int32_t) * (size_t)n);
        int32_t *a = S.b_fi.p, *b = S.b_fd.p;
        uint8_t* ans = S.b_ins.p;
        launch(h, "reach_trivial", kReachTrivial, gridFor(n, blk), blk, 0, (const int32_t*)S.b_u.p,
               (const int32_t*)S.b_v.p, n, (const int32_t*)h->scc.p, (const int64_t*)h->rank.p, V, ans, a, b);
        if (h->prm.reach_local != 0)
            launch(h, "reach_local", kReachLocal, std::min<unsigned>(gridFor(n * WAVE, 256), 2048), 256, 0,
                   (const int32_t*)h->dag_beg.p, (const int32_t*)h->dag_dst.p, (const int32_t*)h->dag_in_beg.p,
                   (const int32_t*)h->dag_in_src.p, (const int64_t*)h->rank.p, n, (const int32_t*)a,
                   (const int32_t*)b, ans);
        std::vector<uint8_t> res((size_t)n);
        std::vector<int32_t> ha((size_t)n), hb((size_t)n);
        HIPCHK(hipMemcpyAsync(res.data(), ans, (size_t)n, hipMemcpyDeviceToHost, h->st));
        HIPCHK(hipMemcpyAsync(ha.data(), a, sizeof(int32_t) * (size_t)n, hipMemcpyDeviceToHost, h->st));
        HIPCHK(hipMemcpyAsync(hb.data(), b, sizeof(int32_t) * (size_t)n, hipMemcpyDeviceToHost, h->st));
        syncStream(h);
        std::vector<int64_t> open;
        for (int64_t i = 0; i < n; ++i)
            if (res[i] == 2) open.push_back(i);
        if (!open.empty()) {
            // target ranks, to sort the open pairs (each sweep's allowed bits
            // are then a suffix of its 64)
            std::vector<int64_t> rk(V > 0 ? V : 1);
            HIPCHK(hipMemcpyAsync(rk.data(), h->rank.p, sizeof(int64_t) * (size_t)V, hipMemcpyDeviceToHost, h->st));
            syncStream(h);
            std::stable_sort(open.begin(), open.end(), [&](int64_t x, int64_t y) { return rk[hb[x]] < rk[hb[y]]; });
            const size_t nV = (size_t)(V > 0 ? V : 1);
            const int64_t nsw = ((int64_t)open.size() + 63) / 64;
            // sweeps per group: RG, fewer when V makes 9 words per SCC id and sweep large
            const int g = (int)std::max<int64_t>(1, std::min<int64_t>({(int64_t)RG, nsw, (int64_t)((4ULL << 30) / (72 * nV))}));
            S.rq_seen.ensure((size_t)g * nV);
            S.rq_next.ensure((size_t)g * nV);
            S.rq_tgt.ensure((size_t)g * nV);
            for (int p = 0; p < 2; ++p) {
                S.rq_fq[p].ensure((size_t)g * nV);
                S.rq_fb[p].ensure((size_t)g * nV);
            }
            S.rq_tq.ensure((size_t)g * nV);
            S.rq_ctl.ensure((size_t)g * R_SLOTS);
            S.rq_pairs.ensure((size_t)g * 64 * 2 + g);
            S.rq_lim.ensure((size_t)g * 64);
            ReachGroup R;
            R.dbeg = h->dag_beg.p;
            R.ddst = h->dag_dst.p;
            R.rank = h->rank.p;
            R.lim = S.rq_lim.p;
            R.seen = S.rq_seen.p;
            R.next = S.rq_next.p;
            R.tgt = S.rq_tgt.p;
            for (int p = 0; p < 2; ++p) {
                R.fq[p] = S.rq_fq[p].p;
                R.fb[p] = S.rq_fb[p].p;
            }
            R.tq = S.rq_tq.p;
            R.ctl = S.rq_ctl.p;
            R.V = V;
            int32_t* qa = S.rq_pairs.p;
            int32_t* qb = qa + (size_t)g * 64;
            int32_t* qk = qb + (size_t)g * 64;
            fillU64(h, R.next, 0, (int64_t)g * V);
            fillU64(h, R.tgt, 0, (int64_t)g * V);
            std::vector<int32_t> ha2((size_t)g * 64), hb2((size_t)g * 64), hk(g);
            std::vector<int64_t> hl((size_t)g * 64);
            std::vector<int64_t> r((size_t)g * R_SLOTS);
            const int C = 16;  // levels per readback
            for (int64_t s0 = 0; s0 < nsw; s0 += g) {
                const int gs = (int)std::min<int64_t>(g, nsw - s0);
                for (int y = 0; y < gs; ++y) {
                    const int64_t p0 = (s0 + y) * 64;
                    const int k = (int)std::min<int64_t>(64, (int64_t)open.size() - p0);
                    hk[y] = k;
                    for (int q = 0; q < 64; ++q) {
                        const bool on = q < k;
                        ha2[y * 64 + q] = on ? ha[open[p0 + q]] : 0;
                        hb2[y * 64 + q] = on ? hb[open[p0 + q]] : 0;
                        hl[y * 64 + q] = on ? rk[hb2[y * 64 + q]] : INT64_MAX;  // (padding bits are never seeded)
                    }
                }
                copyIn(h, qa, ha2.data(), sizeof(int32_t) * (size_t)gs * 64);
                copyIn(h, qb, hb2.data(), sizeof(int32_t) * (size_t)gs * 64);
                copyIn(h, qk, hk.data(), sizeof(int32_t) * (size_t)gs);
                copyIn(h, S.rq_lim.p, hl.data(), sizeof(int64_t) * (size_t)gs * 64);
                fillU64(h, R.seen, 0, (int64_t)gs * V);
                launch(h, "reach_seed", kRSeed, dim3(1, gs), 64, 0, R, (const int32_t*)qa, (const int32_t*)qb,
                       (const int32_t*)qk);
                int par = 0;
                while (true) {
                    for (int t = 0; t < C; ++t) {
                        const ReachPushF f{R, par};
                        launch(h, "reach_push", kRPush, dim3(256, gs), blk, 0, f, heavyList(h));
                        launchHeavy(h, "reach_push_heavy", f);
                        launch(h, "reach_finish", kRFinish, dim3(256, gs), blk, 0, R, par);
                        par ^= 1;
                    }
                    HIPCHK(hipMemcpyAsync(r.data(), R.ctl, sizeof(int64_t) * (size_t)gs * R_SLOTS,
                                          hipMemcpyDeviceToHost, h->st));
                    syncStream(h);
                    bool live = false;
                    for (int y = 0; y < gs; ++y) {
                        const int64_t* c = r.data() + (size_t)y * R_SLOTS;
                        live = live || (c[R_FOUND] != c[R_ALL] && c[R_NF + par] > 0);
                    }
                    if (!live) break;
                }
                for (int y = 0; y < gs; ++y) {
                    const uint64_t fnd = (uint64_t)r[(size_t)y * R_SLOTS + R_FOUND];
                    const int64_t p0 = (s0 + y) * 64;
                    for (int q = 0; q < hk[y]; ++q) res[open[p0 + q]] = (fnd >> q) & 1;
                }
                // next words are cleared by every level's finish, except in a
                // sweep that stopped early with a frontier in flight
                launch(h, "reach_clear", kRClearTargets, dim3(1, gs), 64, 0, R, (const int32_t*)qb,
                       (const int32_t*)qk);
                fillU64(h, R.next, 0, (int64_t)gs * V);
            }
        }
        tEnd(h, "reachable");
        HIPCHK(hipMemcpy(out, res.data(), (size_t)n, hipMemcpyDefault));
    });
}

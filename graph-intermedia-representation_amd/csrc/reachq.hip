// Reachability queries on the condensation: replaces GIR::reachable
// (gir_query.cpp:789-828) for a batch of (u, v) pairs.
//
// As in the reference, a pair inside one SCC is reachable and a pair whose
// source SCC ranks after the target SCC is not (ranks are a topological
// order of the DAG); the rest need a search of the DAG.  The reference runs a
// bidirectional search per pair; here up to 64 open pairs share one
// bit-parallel forward search over the DAG (bit q = pair q, as in the
// multi-source BFS), pruned by rank: bit q never enters an SCC ranked after
// its target.  The open pairs are sorted by target rank, so the bits allowed
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

struct ReachSweep {
    const int32_t* dbeg;   // DAG CSR by source SCC id
    const int32_t* ddst;
    const int64_t* rank;
    const int64_t* lim;    // target ranks of the sweep's pairs, ascending (64; padding INT64_MAX)
    uint64_t* seen;        // per SCC id: bits that reached it
    uint64_t* next;
    const uint64_t* tgt;   // per SCC id: bits whose target it is
    int32_t* tq;           // touched list
    int64_t* nt;
};

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

struct ReachPushF {
    ReachSweep R;
    __device__ int64_t operator()(int32_t, uint64_t fp, int32_t, int64_t j) const {
        const int32_t y = R.ddst[j];
        const uint64_t m = fp & allowed(R.lim, R.rank[y]) & ~(R.seen[y] | R.next[y]);
        const bool first = m && atomicOr((unsigned long long*)&R.next[y], (unsigned long long)m) == 0;
        const int64_t pos = wavePush(R.nt, first);
        if (first) R.tq[pos] = y;
        return 0;
    }
};

// Frontier (SCC ids fq, bits fbits) -> touched SCCs with their new bits.
__global__ void kReachPush(const int32_t* __restrict__ fq, const uint64_t* __restrict__ fbits, const int64_t* nf,
                           ReachPushF f, HeavyList H) {
    const int64_t n = *nf;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t base = (int64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~(WAVE - 1)); base < n; base += stride) {
        const int64_t i = base + laneId();
        const bool act = i < n;
        const int32_t x = act ? fq[i] : 0;
        const uint64_t fx = act ? fbits[i] : 0;
        const int32_t b = act ? f.R.dbeg[x] : 0, e = act ? f.R.dbeg[x + 1] : 0;
        waveRowsH(act && fx != 0, x, fx, 0, b, e, H, f);
    }
}

// Touched SCCs: new bits -> seen, answered bits, next frontier.
__global__ void kReachFinish(ReachSweep R, const int64_t* nt, int32_t* __restrict__ fq, uint64_t* __restrict__ fbits,
                             int64_t* nf, uint64_t* found) {
    const int64_t n = *nt;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    uint64_t hit = 0;
    for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < n; base += stride) {
        const int64_t i = base + threadIdx.x;
        uint64_t nb = 0;
        int32_t y = 0;
        if (i < n) {
            y = R.tq[i];
            nb = R.next[y] & ~R.seen[y];
            R.next[y] = 0;
            R.seen[y] |= nb;
            hit |= nb & R.tgt[y];
        }
        const int64_t pos = blockReserve(nb != 0, nf);
        if (nb) {
            fq[pos] = y;
            fbits[pos] = nb;
        }
    }
    for (int o = WAVE / 2; o > 0; o >>= 1) hit |= __shfl_xor(hit, o);
    if (laneId() == 0 && hit) atomicOr((unsigned long long*)found, (unsigned long long)hit);
}

// Seeds of a sweep: source SCCs (with their bits) and target masks.
__global__ void kReachSeed(const int32_t* __restrict__ a, const int32_t* __restrict__ b, int k, uint64_t* seen,
                           uint64_t* tgt, int32_t* fq, uint64_t* fbits, int64_t* nf) {
    const int q = threadIdx.x;
    if (q >= k) return;
    atomicOr((unsigned long long*)&seen[a[q]], 1ULL << q);
    atomicOr((unsigned long long*)&tgt[b[q]], 1ULL << q);
    fq[q] = a[q];
    fbits[q] = 1ULL << q;
    if (q == 0) *nf = k;
}

// Clears a sweep's target marks.
__global__ void kReachClearSeeds(const int32_t* __restrict__ a, const int32_t* __restrict__ b, int k, uint64_t* seen,
                                 uint64_t* tgt) {
    const int q = threadIdx.x;
    if (q >= k) return;
    seen[a[q]] = 0;
    tgt[b[q]] = 0;
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
            // are then a suffix)
            std::vector<int64_t> rk(V > 0 ? V : 1);
            HIPCHK(hipMemcpyAsync(rk.data(), h->rank.p, sizeof(int64_t) * (size_t)V, hipMemcpyDeviceToHost, h->st));
            syncStream(h);
            std::stable_sort(open.begin(), open.end(), [&](int64_t x, int64_t y) { return rk[hb[x]] < rk[hb[y]]; });
            const size_t nV = (size_t)(V > 0 ? V : 1);
            for (int i = 0; i < 3; ++i) S.ms[i].ensure(nV);  // seen / next / target masks per SCC id
            for (int i = 0; i < 3; ++i) S.q[i].ensure(nV);   // frontier SCC ids (x2), touched list
            S.keys[0].ensure(nV);                            // frontier bits (x2)
            S.keys[1].ensure(nV);
            S.mscnt.ensure(8);
            S.off.ensure(64 + 1);
            S.iv[0].ensure(64);
            S.iv[1].ensure(64);
            uint64_t *seen = S.ms[0].p, *next = S.ms[1].p, *tgt = S.ms[2].p;
            int64_t* c = S.mscnt.p;  // [0], [1]: frontier sizes by parity; [2]: touched; [3]: answered bits
            int64_t* lim = S.off.p;
            int32_t *qa = S.iv[0].p, *qb = S.iv[1].p;
            fillU64(h, next, 0, V);
            fillU64(h, tgt, 0, V);
            for (size_t s0 = 0; s0 < open.size(); s0 += 64) {
                const int k = (int)std::min<size_t>(64, open.size() - s0);
                int32_t sa[64], sb[64];
                int64_t sl[64];
                for (int q = 0; q < 64; ++q) {
                    const bool on = q < k;
                    sa[q] = on ? ha[open[s0 + q]] : 0;
                    sb[q] = on ? hb[open[s0 + q]] : 0;
                    sl[q] = on ? rk[sb[q]] : INT64_MAX;  // (padding bits are never seeded)
                }
                copyIn(h, qa, sa, sizeof(sa));
                copyIn(h, qb, sb, sizeof(sb));
                copyIn(h, lim, sl, sizeof(sl));
                fillU64(h, seen, 0, V);
                fillI64(h, c, 0, 4);
                launch(h, "reach_seed", kReachSeed, 1, 64, 0, (const int32_t*)qa, (const int32_t*)qb, k, seen, tgt,
                       S.q[0].p, S.keys[0].p, c + 0);
                const uint64_t all = k >= 64 ? ~0ULL : ((1ULL << k) - 1);
                const ReachSweep R{h->dag_beg.p, h->dag_dst.p, h->rank.p, lim, seen, next, tgt, S.q[2].p, c + 2};
                const ReachPushF f{R};
                int cur = 0;
                int64_t hn = k, found = 0;
                while (hn > 0 && ((uint64_t)found & all) != all) {
                    int64_t* nf = c + cur;
                    int64_t* nf2 = c + (cur ^ 1);
                    fillI64(h, nf2, 0, 1);
                    fillI64(h, c + 2, 0, 1);
                    launch(h, "reach_push", kReachPush, std::min<unsigned>(gridFor(hn, blk), 4096), blk, 0,
                           (const int32_t*)S.q[cur].p, (const uint64_t*)S.keys[cur].p, (const int64_t*)nf, f,
                           heavyList(h));
                    launchHeavy(h, "reach_push_heavy", f);
                    launch(h, "reach_finish", kReachFinish, std::min<unsigned>(gridFor(V, blk), 1024), blk, 0, R,
                           (const int64_t*)(c + 2), S.q[cur ^ 1].p, S.keys[cur ^ 1].p, nf2, (uint64_t*)(c + 3));
                    int64_t r[4];
                    readCounters(h, c, 4, r);
                    hn = r[cur ^ 1];
                    found = r[3];
                    cur ^= 1;
                }
                for (int q = 0; q < k; ++q) res[open[s0 + q]] = ((uint64_t)found >> q) & 1;
                // the next sweep starts from clean marks: every level's finish
                // clears the next words it used, seen is refilled per sweep,
                // targets are cleared here
                launch(h, "reach_clear", kReachClearSeeds, 1, 64, 0, (const int32_t*)qa, (const int32_t*)qb, k, seen,
                       tgt);
            }
        }
        tEnd(h, "reachable");
        HIPCHK(hipMemcpy(out, res.data(), (size_t)n, hipMemcpyDefault));
    });
}

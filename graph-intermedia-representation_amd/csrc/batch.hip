// Batched updates: replaces GIR::applyBatch (gir.cpp:3634-4111),
// updateEdgeDirected / insertEdge / deleteEdge (gir.cpp:3043-3133) and
// ingestBatch / sync (gir.cpp:3588-3632).
//
// Normalisation (normalizeBatch, gir.cpp:3523-3586): ops with u == v or an
// endpoint out of range are dropped; packed keys (u, v) with the op index as
// value go through a stable radix sort, so the last entry of each key run is
// the last op on that arc; it is classified against the current store by a
// binary search in row u: insert of an absent arc -> insert, delete of a
// present arc -> delete, insert of a present arc -> counted no-op, delete of
// an absent arc -> dropped.
//
// Store update: both CSRs are rebuilt by a parallel merge.  Rows without
// changes are copied slot for slot; in a changed row every kept slot moves by
// (inserts before it) - (deletes before it), found by binary searches in the
// row's own slice of the sorted change lists, and every insert lands at
// (its index among the row's inserts) + (kept arcs before it).  Weights
// follow their arcs.
//
// Repair: the reference repairs WCC/SCC/DAG/ranks incrementally (SPU/FMU);
// its post-condition is that every layer equals a from-scratch build on the
// new edge set (gir_test.cpp:470-563 checks exactly that).  Here SCC + DAG +
// ranks are repaired from the classified changes (repairScc, scc.hip) or
// rebuilt; then WCC: when no delete can split a WCC (its ends share an SCC of
// the new graph, or their SCCs are still joined by a condensation arc, or the
// delete was shown to be bypassed) the inserted arcs are hooked into the
// existing min-rooted forest (labels stay canonical), else WCC is rebuilt.
#include <vector>

#include "common.h"

namespace gir {

void runWccHookArcs(girhip_graph* h, const uint64_t* keys, int64_t n, int B);

__device__ __forceinline__ int64_t lowerBoundKey(const uint64_t* __restrict__ a, int64_t lo, int64_t hi, uint64_t key) {
    while (lo < hi) {
        const int64_t m = lo + ((hi - lo) >> 1);
        if (a[m] < key) lo = m + 1;
        else hi = m;
    }
    return lo;
}
__device__ __forceinline__ int32_t lowerBoundI(const int32_t* __restrict__ a, int32_t lo, int32_t hi, int32_t x) {
    while (lo < hi) {
        const int32_t m = lo + ((hi - lo) >> 1);
        if (a[m] < x) lo = m + 1;
        else hi = m;
    }
    return lo;
}

// Packed key per op, invalid ops get the sentinel 1 << (2B) (sorts last).
__global__ void kBatchKeys(const int32_t* __restrict__ u, const int32_t* __restrict__ v, int64_t n, int32_t V, int B,
                           uint64_t* __restrict__ keys, uint32_t* __restrict__ vals, int64_t* cnt) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool bad = false;
    if (i < n) {
        const int32_t a = u[i], b = v[i];
        bad = a == b || a < 0 || b < 0 || a >= V || b >= V;
        keys[i] = bad ? (1ULL << (2 * B)) : (((uint64_t)(uint32_t)a << B) | (uint32_t)b);
        vals[i] = (uint32_t)i;
    }
    const uint64_t m = __ballot(bad);
    if (m && laneId() == __ffsll((long long)m) - 1) atomicAdd((unsigned long long*)cnt, (unsigned long long)__popcll(m));
}

// code: 1 insert, 2 delete, 0 skip; cnt[0] += insert no-ops.
__global__ void kBatchClassify(const uint64_t* __restrict__ keys, const uint32_t* __restrict__ vals, int64_t n, int B,
                               const uint8_t* __restrict__ ins, const int32_t* __restrict__ fb,
                               const int32_t* __restrict__ fa, int32_t* __restrict__ fins, int32_t* __restrict__ fdel,
                               int64_t* cnt) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool noop = false;
    if (i < n) {
        const uint64_t k = keys[i];
        int32_t a = 0, d = 0;
        if (!(k >> (2 * B)) && (i + 1 == n || keys[i + 1] != k)) {
            const int32_t u = (int32_t)(k >> B), v = (int32_t)(k & ((1ULL << B) - 1));
            const bool want = ins[vals[i]] != 0;
            const int32_t lo = fb[u], hi = fb[u + 1];
            const int32_t p = lowerBoundI(fa, lo, hi, v);
            const bool present = p < hi && fa[p] == v;
            if (want && !present) a = 1;
            else if (!want && present) d = 1;
            else noop = want;
        }
        fins[i] = a;
        fdel[i] = d;
    }
    const uint64_t m = __ballot(noop);
    if (m && laneId() == __ffsll((long long)m) - 1) atomicAdd((unsigned long long*)cnt, (unsigned long long)__popcll(m));
}

// Stable compaction by scanned flags; inserts also carry their weight.
__global__ void kBatchCompact(const uint64_t* __restrict__ keys, const uint32_t* __restrict__ vals, int64_t n, int B,
                              const int32_t* __restrict__ f, const int64_t* __restrict__ pos,
                              const uint32_t* __restrict__ wop, bool hash_w, uint64_t* __restrict__ out,
                              uint32_t* __restrict__ wout) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !f[i]) return;
    const uint64_t k = keys[i];
    out[pos[i]] = k;
    if (wout) {
        const int32_t u = (int32_t)(k >> B), v = (int32_t)(k & ((1ULL << B) - 1));
        wout[pos[i]] = (wop && !hash_w) ? wop[vals[i]] : hashWeight(u, v);
    }
}

__global__ void kTransposeKeys(const uint64_t* __restrict__ in, int64_t n, int B, uint64_t* __restrict__ out,
                               uint32_t* __restrict__ idx) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint64_t k = in[i], mask = (1ULL << B) - 1;
    out[i] = ((k & mask) << B) | (k >> B);
    if (idx) idx[i] = (uint32_t)i;
}

// ---- CSR merge ------------------------------------------------------------
// The CSR of one direction is the globally sorted key sequence (row, col).
// With the sorted change lists, old slot j moves to j - D(j) + I(j), where
// D(j) = deleted old slots before j and I(j) = inserted keys whose insertion
// point (first old slot with a larger key) is <= j; insert k lands at
// p_k - D(p_k) + k.  Row offsets follow from key counts: new beg[u] =
// beg[u] - #deletes in rows < u + #inserts in rows < u.

// Old-array position of every change key: deletes at their slot (which
// must hold the key), inserts at their insertion point (which must not).
__global__ void kChangePos(const int32_t* __restrict__ beg, const int32_t* __restrict__ adj, int B,
                           const uint64_t* __restrict__ keys, int64_t n, bool del, int64_t* __restrict__ pos,
                           int64_t* err) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const uint64_t key = keys[k];
    const int32_t u = (int32_t)(key >> B), x = (int32_t)(key & ((1ULL << B) - 1));
    const int32_t p = lowerBoundI(adj, beg[u], beg[u + 1], x);
    const bool hit = p < beg[u + 1] && adj[p] == x;
    if (hit != del) atomicAdd((unsigned long long*)err, 1ULL);
    pos[k] = p;
}

// New row offsets (int64, V + 1 entries).
__global__ void kNewBeg(const int32_t* __restrict__ beg, int32_t V, int B, const uint64_t* __restrict__ ik,
                        int64_t ni, const uint64_t* __restrict__ dk, int64_t nd, int64_t* __restrict__ nb) {
    const int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (u > V) return;
    const uint64_t key = (uint64_t)u << B;
    nb[u] = (int64_t)beg[u] - lowerBoundKey(dk, 0, nd, key) + lowerBoundKey(ik, 0, ni, key);
}

__device__ __forceinline__ int64_t countBelow(const int64_t* __restrict__ a, int64_t lo, int64_t hi, int64_t x) {
    while (lo < hi) {  // first index in [lo, hi) with a[index] >= x
        const int64_t m = lo + ((hi - lo) >> 1);
        if (a[m] < x) lo = m + 1;
        else hi = m;
    }
    return lo;
}

constexpr int MERGE_TILE = 4096;

// Old slots, one tile of MERGE_TILE per block: a tile without change
// positions is a shifted copy; otherwise each slot counts the changes in
// the tile's range that precede it.
__global__ void kShiftCopy(const int32_t* __restrict__ adj, const uint32_t* __restrict__ w, int64_t E,
                           const int64_t* __restrict__ ipos, int64_t ni, const int64_t* __restrict__ dpos, int64_t nd,
                           int32_t* __restrict__ nadj, uint32_t* __restrict__ nw, int64_t newE, int64_t* err) {
    __shared__ int64_t rng[4];
    const int64_t t0 = (int64_t)blockIdx.x * MERGE_TILE;
    if (t0 >= E) return;
    const int64_t t1 = t0 + MERGE_TILE < E ? t0 + MERGE_TILE : E;
    if (threadIdx.x == 0) {
        rng[0] = countBelow(ipos, 0, ni, t0 + 1);  // inserts with point <= t0
        rng[1] = countBelow(ipos, 0, ni, t1);      // ... with point <= t1 - 1
        rng[2] = countBelow(dpos, 0, nd, t0);      // deleted slots < t0
        rng[3] = countBelow(dpos, 0, nd, t1);      // ... < t1
    }
    __syncthreads();
    const int64_t i0 = rng[0], i1 = rng[1], d0 = rng[2], d1 = rng[3];
    if (i0 == i1 && d0 == d1) {
        const int64_t sh = i0 - d0;
        if (t0 + sh < 0 || t1 + sh > newE) {  // (inconsistent change lists; never written out of range)
            if (threadIdx.x == 0) atomicAdd((unsigned long long*)err, 1ULL);
            return;
        }
        for (int64_t j = t0 + threadIdx.x; j < t1; j += blockDim.x) {
            nadj[j + sh] = adj[j];
            if (nw) nw[j + sh] = w[j];
        }
        return;
    }
    for (int64_t j = t0 + threadIdx.x; j < t1; j += blockDim.x) {
        const int64_t dj = countBelow(dpos, d0, d1, j);
        if (dj < d1 && dpos[dj] == j) continue;  // deleted
        const int64_t ij = countBelow(ipos, i0, i1, j + 1);
        const int64_t q = j - dj + ij;
        if (q < 0 || q >= newE) {
            atomicAdd((unsigned long long*)err, 1ULL);
            continue;
        }
        nadj[q] = adj[j];
        if (nw) nw[q] = w[j];
    }
}

__global__ void kInsertPlace(const uint64_t* __restrict__ ik, const uint32_t* __restrict__ iw, int64_t ni, int B,
                             const int64_t* __restrict__ ipos, const int64_t* __restrict__ dpos, int64_t nd,
                             int32_t* __restrict__ nadj, uint32_t* __restrict__ nw, int64_t newE, int64_t* err) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= ni) return;
    const int64_t p = ipos[k];
    const int64_t q = p - countBelow(dpos, 0, nd, p) + k;
    if (q < 0 || q >= newE) {
        atomicAdd((unsigned long long*)err, 1ULL);
        return;
    }
    nadj[q] = (int32_t)(ik[k] & ((1ULL << B) - 1));
    if (nw) nw[q] = iw[k];
}

}  // namespace gir

namespace gir {
namespace {

// Sorted change lists of one CSR direction (buffers live in the handle's
// scratch).
struct Changes {
    int64_t ni = 0, nd = 0;
    uint64_t* ik = nullptr;  // sorted packed (row, col) keys of inserted arcs
    uint64_t* dk = nullptr;  // ... of deleted arcs
    uint32_t* iw = nullptr;  // weights of inserted arcs (fw only)
};

// Apply sorted change lists to one CSR (beg/adj/w), merging into the spare
// arrays (adj2/w2), which are then swapped with the live ones.
void mergeCsr(girhip_graph* h, DevBuf<int32_t>& beg, DevBuf<int32_t>& adj, DevBuf<int32_t>& adj2,
              DevBuf<uint32_t>* w, DevBuf<uint32_t>* w2, const Changes& C, int B, int64_t newE) {
    const int32_t V = h->V;
    const int blk = (int)h->prm.block;
    Scratch& S = scratch(h);
    S.b_pi.ensureSlack((size_t)(C.ni > 0 ? C.ni : 1));
    S.b_pd.ensureSlack((size_t)(C.nd > 0 ? C.nd : 1));
    S.b_nb.ensure((size_t)V + 1);
    int64_t *ipos = S.b_pi.p, *dpos = S.b_pd.p, *nb = S.b_nb.p;
    int64_t* err = h->dcnt.p + 48;
    fillI64(h, err, 0, 1);
    if (C.ni)
        launch(h, "batch_pos", kChangePos, gridFor(C.ni, blk), blk, 0, (const int32_t*)beg.p, (const int32_t*)adj.p, B,
               (const uint64_t*)C.ik, C.ni, false, ipos, err);
    if (C.nd)
        launch(h, "batch_pos", kChangePos, gridFor(C.nd, blk), blk, 0, (const int32_t*)beg.p, (const int32_t*)adj.p, B,
               (const uint64_t*)C.dk, C.nd, true, dpos, err);
    launch(h, "batch_newbeg", kNewBeg, gridFor((int64_t)V + 1, blk), blk, 0, (const int32_t*)beg.p, V, B,
           (const uint64_t*)C.ik, C.ni, (const uint64_t*)C.dk, C.nd, nb);
    // the change lists must match the store before anything is moved
    int64_t e1;
    readCounters(h, err, 1, &e1);
    if (e1 != 0) raise(GIRHIP_EHIP, "batch: change lists do not match the edge store");
    adj2.ensureSlack((size_t)(newE > 0 ? newE : 1));
    if (w) w2->ensureSlack((size_t)(newE > 0 ? newE : 1));
    const int64_t E = h->E;
    if (E > 0)
        launch(h, "batch_shift_copy", kShiftCopy, gridFor(E, MERGE_TILE), blk, 0, (const int32_t*)adj.p,
               (const uint32_t*)(w ? w->p : nullptr), E, (const int64_t*)ipos, C.ni, (const int64_t*)dpos, C.nd,
               adj2.p, w ? w2->p : (uint32_t*)nullptr, newE, err);
    if (C.ni)
        launch(h, "batch_insert", kInsertPlace, gridFor(C.ni, blk), blk, 0, (const uint64_t*)C.ik,
               (const uint32_t*)(w ? C.iw : nullptr), C.ni, B, (const int64_t*)ipos, (const int64_t*)dpos, C.nd,
               adj2.p, w ? w2->p : (uint32_t*)nullptr, newE, err);
    // nothing out of range, and the rows must add up
    HIPCHK(hipMemcpyAsync(err + 1, nb + V, sizeof(int64_t), hipMemcpyDeviceToDevice, h->st));
    int64_t e2[2];
    readCounters(h, err, 2, e2);
    if (e2[0] != 0 || e2[1] != newE) raise(GIRHIP_EHIP, "batch: internal merge inconsistency");
    launch(h, "batch_beg", kI64toI32, gridFor((int64_t)V + 1, blk), blk, 0, (const int64_t*)nb, beg.p,
           (int64_t)V + 1);
    adj.swap(adj2);
    if (w) w->swap(*w2);
}

// Deletes that cannot split a WCC, judged on the repaired SCC layer: the
// ends are in one SCC of the new graph, or the condensation still has an arc
// between their SCCs (either way round).  A delete that leaves one end with
// no arc at all while the other end is in the big SCC (label g) only cuts
// off that end (every arc it lost led into one strongly connected set): the
// end becomes its own WCC here, unless it was its WCC's label.  cnt += the
// deletes that fit neither case.
__global__ void kWccSafe(const uint64_t* __restrict__ dk, int64_t nd, int B, const int32_t* __restrict__ scc,
                         const int32_t* __restrict__ dbeg, const int32_t* __restrict__ ddst,
                         const int32_t* __restrict__ fb, const int32_t* __restrict__ bb, int32_t g, int32_t* wcc,
                         int64_t* cnt) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool bad = false;
    if (i < nd) {
        const uint64_t k = dk[i];
        const int32_t u = (int32_t)(k >> B), v = (int32_t)(k & ((1ULL << B) - 1));
        const int32_t su = scc[u], sv = scc[v];
        if (su != sv) {
            auto has = [&](int32_t a, int32_t b) {
                int32_t lo = dbeg[a], hi = dbeg[a + 1];
                while (lo < hi) {
                    const int32_t m = lo + ((hi - lo) >> 1);
                    if (ddst[m] < b) lo = m + 1;
                    else hi = m;
                }
                return lo < dbeg[a + 1] && ddst[lo] == b;
            };
            if (!has(su, sv) && !has(sv, su)) {
                auto alone = [&](int32_t x) { return fb[x + 1] == fb[x] && bb[x + 1] == bb[x]; };
                if (g >= 0 && sv == g && alone(u) && wcc[u] != u) wcc[u] = u;
                else if (g >= 0 && su == g && alone(v) && wcc[v] != v) wcc[v] = v;
                else bad = true;
            }
        }
    }
    const uint64_t m = __ballot(bad);
    if (laneId() == 0 && m) atomicAdd((unsigned long long*)cnt, (unsigned long long)__popcll(m));
}

// Normalise + classify the batch; fills C (fw order) and the stats.
void normalizeBatchDev(girhip_graph* h, const int32_t* u, const int32_t* v, const uint8_t* ins, const uint32_t* w,
                       int64_t n, Changes& C, girhip_batch_stats& st, int B) {
    const int32_t V = h->V;
    const int blk = (int)h->prm.block;
    Scratch& S = scratch(h);
    const bool ww = w && h->has_w;
    S.b_u.ensureSlack((size_t)n);
    S.b_v.ensureSlack((size_t)n);
    S.b_ins.ensureSlack((size_t)n);
    copyIn(h, S.b_u.p, u, sizeof(int32_t) * n);
    copyIn(h, S.b_v.p, v, sizeof(int32_t) * n);
    copyIn(h, S.b_ins.p, ins, (size_t)n);
    if (ww) {
        S.b_w.ensureSlack((size_t)n);
        copyIn(h, S.b_w.p, w, sizeof(uint32_t) * n);
    }
    S.keys[0].ensure((size_t)n);
    S.vals[0].ensure((size_t)n);
    fillI64(h, h->dcnt.p + 40, 0, 4);
    launch(h, "batch_keys", kBatchKeys, gridFor(n, blk), blk, 0, (const int32_t*)S.b_u.p, (const int32_t*)S.b_v.p, n,
           V, B, S.keys[0].p, S.vals[0].p, h->dcnt.p + 40);
    radixSort(h, S.keys[0].p, S.vals[0].p, n, 2 * B + 1);
    S.b_fi.ensureSlack((size_t)n);
    S.b_fd.ensureSlack((size_t)n);
    launch(h, "batch_classify", kBatchClassify, gridFor(n, blk), blk, 0, (const uint64_t*)S.keys[0].p,
           (const uint32_t*)S.vals[0].p, n, B, (const uint8_t*)S.b_ins.p, (const int32_t*)h->fw_beg.p,
           (const int32_t*)h->fw_adj.p, S.b_fi.p, S.b_fd.p, h->dcnt.p + 41);
    S.b_pi.ensureSlack((size_t)n + 1);
    S.b_pd.ensureSlack((size_t)n + 1);
    scanI32toI64(h, S.b_fi.p, S.b_pi.p, n);
    scanI32toI64(h, S.b_fd.p, S.b_pd.p, n);
    HIPCHK(hipMemcpyAsync(h->dcnt.p + 42, S.b_pi.p + n, sizeof(int64_t), hipMemcpyDeviceToDevice, h->st));
    HIPCHK(hipMemcpyAsync(h->dcnt.p + 43, S.b_pd.p + n, sizeof(int64_t), hipMemcpyDeviceToDevice, h->st));
    int64_t c[4];
    readCounters(h, h->dcnt.p + 40, 4, c);
    st.invalid = c[0];
    st.noop = c[1];
    C.ni = c[2];
    C.nd = c[3];
    st.inserted = C.ni;
    st.deleted = C.nd;
    S.b_ik[0].ensureSlack((size_t)(C.ni > 0 ? C.ni : 1));
    S.b_dk[0].ensureSlack((size_t)(C.nd > 0 ? C.nd : 1));
    C.ik = S.b_ik[0].p;
    C.dk = S.b_dk[0].p;
    C.iw = nullptr;
    if (h->has_w) {
        S.b_iw.ensureSlack((size_t)(C.ni > 0 ? C.ni : 1));
        C.iw = S.b_iw.p;
    }
    launch(h, "batch_compact", kBatchCompact, gridFor(n, blk), blk, 0, (const uint64_t*)S.keys[0].p,
           (const uint32_t*)S.vals[0].p, n, B, (const int32_t*)S.b_fi.p, (const int64_t*)S.b_pi.p,
           (const uint32_t*)(ww ? S.b_w.p : nullptr), !ww, C.ik, C.iw);
    launch(h, "batch_compact", kBatchCompact, gridFor(n, blk), blk, 0, (const uint64_t*)S.keys[0].p,
           (const uint32_t*)S.vals[0].p, n, B, (const int32_t*)S.b_fd.p, (const int64_t*)S.b_pd.p,
           (const uint32_t*)nullptr, true, C.dk, (uint32_t*)nullptr);
}

// Transposed change lists (sorted by (col, row)) for the bw CSR.
void transposeChanges(girhip_graph* h, const Changes& C, Changes& T, int B) {
    const int blk = (int)h->prm.block;
    Scratch& S = scratch(h);
    T.ni = C.ni;
    T.nd = C.nd;
    S.b_ik[1].ensureSlack((size_t)(C.ni > 0 ? C.ni : 1));
    S.b_dk[1].ensureSlack((size_t)(C.nd > 0 ? C.nd : 1));
    T.ik = S.b_ik[1].p;
    T.dk = S.b_dk[1].p;
    T.iw = nullptr;
    if (C.ni) {
        launch(h, "batch_tkeys", kTransposeKeys, gridFor(C.ni, blk), blk, 0, (const uint64_t*)C.ik, C.ni, B, T.ik,
               (uint32_t*)nullptr);
        radixSort(h, T.ik, nullptr, C.ni, 2 * B);
    }
    if (C.nd) {
        launch(h, "batch_tkeys", kTransposeKeys, gridFor(C.nd, blk), blk, 0, (const uint64_t*)C.dk, C.nd, B, T.dk,
               (uint32_t*)nullptr);
        radixSort(h, T.dk, nullptr, C.nd, 2 * B);
    }
}

// Store-only part shared by apply/ingest.  Returns the normalised changes.
void applyStore(girhip_graph* h, const int32_t* u, const int32_t* v, const uint8_t* ins, const uint32_t* w, int64_t n,
                girhip_batch_stats& st, Changes& C, int B) {
    st = girhip_batch_stats{0, 0, 0, 0};
    if (n <= 0) return;
    if (!u || !v || !ins) raise(GIRHIP_EINVAL, "batch: null op arrays");
    normalizeBatchDev(h, u, v, ins, w, n, C, st, B);
    if (C.ni + C.nd == 0) return;
    const int64_t newE = h->E + C.ni - C.nd;
    if (newE >= 0x7fffffffLL) raise(GIRHIP_EINVAL, "batch: arc count would exceed int32 offsets");
    Changes T;
    transposeChanges(h, C, T, B);
    mergeCsr(h, h->fw_beg, h->fw_adj, h->fw_adj2, h->has_w ? &h->w_fw : nullptr, h->has_w ? &h->w_fw2 : nullptr, C,
             B, newE);
    mergeCsr(h, h->bw_beg, h->bw_adj, h->bw_adj2, nullptr, nullptr, T, B, newE);
    h->E = newE;
    incLog(h, C.ik, C.ni);
    incLog(h, C.dk, C.nd);
    h->src_ok = false;
    h->max_deg = -1;
    h->max_w = -1;
    h->sym_ok = false;
    for (auto& sv : h->split) sv.delta = 0;
}

// Bring every requested layer up to date.  C = the changes just applied to a
// store the layers were current with (nullptr: layers are stale, rebuild).
void repairLayers(girhip_graph* h, const Changes* C, int B) {
    const bool changed = C && (C->ni + C->nd) > 0;
    // a few deletes: does each one's tail still reach its head?  Then no
    // delete splits a WCC or an SCC.
    const uint8_t* bypass = nullptr;
    bool all_bypassed = false;
    if (C && C->nd > 0 && C->nd <= BYPASS_MAX && h->prm.update_bypass != 0 &&
        ((h->wcc_req && h->wcc_ok) || (h->scc_req && h->scc_ok))) {
        Scratch& S = scratch(h);
        S.bp_found.ensure((size_t)BYPASS_MAX);
        all_bypassed = deletesBypassed(h, C->dk, C->nd, B, S.bp_found.p);
        bypass = S.bp_found.p;
    }
    h->last_ms["batch_all_bypassed"] = all_bypassed ? 1 : 0;
    if (h->scc_req) {
        if (h->scc_ok && C && changed && repairScc(h, C->ik, C->ni, C->dk, C->nd, B, bypass)) {
            // repaired from the classified changes
        } else if (!h->scc_ok || changed) {
            runScc(h);
            runDag(h);
            h->last_ms["batch_scc_path"] = 4;
        }
        h->scc_ok = true;
    }
    if (h->wcc_req) {
        // deletes that keep every WCC together: judged on the repaired SCCs
        bool keep = C && (C->nd == 0 || all_bypassed);
        int64_t unsafe = -1;
        if (!keep && h->wcc_ok && C && C->nd > 0 && h->scc_req && h->scc_ok) {
            int64_t* c = h->dcnt.p + 52;
            fillI64(h, c, 0, 1);
            launch(h, "wcc_safe", kWccSafe, gridFor(C->nd, 256), 256, 0, (const uint64_t*)C->dk, C->nd, B,
                   (const int32_t*)h->scc.p, (const int32_t*)h->dag_beg.p, (const int32_t*)h->dag_dst.p,
                   (const int32_t*)h->fw_beg.p, (const int32_t*)h->bw_beg.p, h->gc_ok ? h->gc_label : -1, h->wcc.p,
                   c);
            readCounters(h, c, 1, &unsafe);
            keep = unsafe == 0;
        }
        h->last_ms["batch_wcc_unsafe"] = (double)unsafe;
        if (h->wcc_ok && keep) {
            if (C->ni) runWccHookArcs(h, C->ik, C->ni, B);  // hook the new arcs
        } else if (!h->wcc_ok || changed) {
            runWcc(h);
        }
        h->wcc_ok = true;
    }
}

}  // namespace

// Work buffers of a batch of up to n ops, reserved when a layer is first
// built so that the first small batch allocates nothing large.
void reserveBatch(girhip_graph* h, int64_t n) {
    Scratch& S = scratch(h);
    const size_t m = (size_t)std::max<int64_t>(n, 1);
    S.b_u.ensureSlack(m);
    S.b_v.ensureSlack(m);
    S.b_ins.ensureSlack(m);
    S.b_fi.ensureSlack(m);
    S.b_fd.ensureSlack(m);
    S.b_pi.ensureSlack(m + 1);
    S.b_pd.ensureSlack(m + 1);
    for (int i = 0; i < 2; ++i) {
        S.b_ik[i].ensureSlack(m);
        S.b_dk[i].ensureSlack(m);
    }
    if (h->has_w) {
        S.b_w.ensureSlack(m);
        S.b_iw.ensureSlack(m);
    }
    S.b_nb.ensure((size_t)h->V + 1);
    S.bp_found.ensure((size_t)BYPASS_MAX);
}

}  // namespace gir

using namespace gir;

static int batchImpl(girhip_graph* h, const int32_t* u, const int32_t* v, const uint8_t* ins, const uint32_t* w,
                     int64_t n, girhip_batch_stats* stats, bool repair) {
    return guarded([&] {
        useDevice(h);
        if (n < 0) raise(GIRHIP_EINVAL, "batch: negative op count");
        const int B = bitsFor((uint64_t)(h->V > 1 ? h->V - 1 : 1));
        tBegin(h);
        if (repair && h->pending) {
            repairLayers(h, nullptr, B);
            h->pending = false;
        }
        girhip_batch_stats st;
        Changes C;
        applyStore(h, u, v, ins, w, n, st, C, B);
        if (repair) {
            repairLayers(h, &C, B);
        } else if (C.ni + C.nd > 0) {
            // layers now describe the graph as of the last sync
            h->wcc_ok = false;
            h->scc_ok = false;
            h->pending = h->wcc_req || h->scc_req;
        }
        tEnd(h, "batch");
        if (stats) *stats = st;
    });
}

extern "C" int girhip_apply_batch(girhip_graph* h, const int32_t* u, const int32_t* v, const uint8_t* ins,
                                  const uint32_t* w, int64_t n, girhip_batch_stats* stats) {
    return batchImpl(h, u, v, ins, w, n, stats, true);
}

extern "C" int girhip_ingest_batch(girhip_graph* h, const int32_t* u, const int32_t* v, const uint8_t* ins,
                                   const uint32_t* w, int64_t n, girhip_batch_stats* stats) {
    return batchImpl(h, u, v, ins, w, n, stats, false);
}

extern "C" int girhip_sync(girhip_graph* h) {
    return guarded([&] {
        useDevice(h);
        if (!h->pending) return;
        const int B = bitsFor((uint64_t)(h->V > 1 ? h->V - 1 : 1));
        repairLayers(h, nullptr, B);
        h->pending = false;
        syncStream(h);
    });
}

extern "C" int girhip_synced(const girhip_graph* h, int32_t* is_synced) {
    return guarded([&] {
        if (!h || !is_synced) raise(GIRHIP_EINVAL, "synced: null argument");
        *is_synced = h->pending ? 0 : 1;
    });
}

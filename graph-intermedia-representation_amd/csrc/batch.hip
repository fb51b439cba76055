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
// new edge set (gir_test.cpp:470-563 checks exactly that).  Here an
// insert-only batch repairs WCC by hooking the new arcs into the existing
// min-rooted forest (labels stay canonical); any other change, and SCC + DAG
// + ranks, are rebuilt on the device from the new CSR.
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

__global__ void kRowCount(const uint64_t* __restrict__ keys, int64_t n, int B, int32_t* cnt) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) atomicAdd(&cnt[(int32_t)(keys[i] >> B)], 1);
}

__global__ void kNewDeg(const int32_t* __restrict__ beg, const int32_t* __restrict__ ic, const int32_t* __restrict__ dc,
                        int32_t V, int32_t* __restrict__ nd) {
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v < V) nd[v] = beg[v + 1] - beg[v] + ic[v] - dc[v];
}

// Old slots -> new positions (deleted slots dropped).
__global__ void kMergeOld(const int32_t* __restrict__ beg, const int32_t* __restrict__ adj,
                          const uint32_t* __restrict__ w, int32_t V, int B, const int64_t* __restrict__ ib,
                          const uint64_t* __restrict__ ik, const int64_t* __restrict__ db,
                          const uint64_t* __restrict__ dk, const int64_t* __restrict__ nb, int32_t* __restrict__ nadj,
                          uint32_t* __restrict__ nw, int64_t* err) {
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool act = v < V;
    const int32_t b = act ? beg[v] : 0, e = act ? beg[v + 1] : 0;
    waveRows(act, (int32_t)v, b, e, [&](int32_t u, int64_t j) {
        const int64_t i0 = ib[u], i1 = ib[u + 1], d0 = db[u], d1 = db[u + 1];
        const int32_t x = adj[j];
        int64_t pos = nb[u] + (j - beg[u]);
        if (i0 != i1 || d0 != d1) {
            const uint64_t key = ((uint64_t)(uint32_t)u << B) | (uint32_t)x;
            const int64_t dpos = lowerBoundKey(dk, d0, d1, key);
            if (dpos < d1 && dk[dpos] == key) return;  // deleted
            pos += (lowerBoundKey(ik, i0, i1, key) - i0) - (dpos - d0);
        }
        if (pos < nb[u] || pos >= nb[u + 1]) {  // (an inconsistent store; never written out of the row)
            atomicAdd((unsigned long long*)err, 1ULL);
            return;
        }
        nadj[pos] = x;
        if (nw) nw[pos] = w[j];
    });
}

__global__ void kMergeIns(const int32_t* __restrict__ beg, const int32_t* __restrict__ adj, int B,
                          const uint64_t* __restrict__ ik, const uint32_t* __restrict__ iw, int64_t nins,
                          const int64_t* __restrict__ ib, const int64_t* __restrict__ db,
                          const uint64_t* __restrict__ dk, const int64_t* __restrict__ nb, int32_t* __restrict__ nadj,
                          uint32_t* __restrict__ nw, int64_t* err) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nins) return;
    const uint64_t key = ik[k];
    const int32_t u = (int32_t)(key >> B), x = (int32_t)(key & ((1ULL << B) - 1));
    const int32_t before_old = lowerBoundI(adj, beg[u], beg[u + 1], x) - beg[u];
    const int64_t before_del = lowerBoundKey(dk, db[u], db[u + 1], key) - db[u];
    const int64_t pos = nb[u] + (k - ib[u]) + before_old - before_del;
    if (pos < nb[u] || pos >= nb[u + 1]) {
        atomicAdd((unsigned long long*)err, 1ULL);
        return;
    }
    nadj[pos] = x;
    if (nw) nw[pos] = iw[k];
}

namespace {

struct Changes {
    int64_t ni = 0, nd = 0;
    DevBuf<uint64_t> ik, dk;  // sorted packed (row, col) keys
    DevBuf<uint32_t> iw;      // weights of inserted arcs (fw only)
};

// Apply sorted change lists to one CSR (beg/adj/w), producing new arrays.
void mergeCsr(girhip_graph* h, DevBuf<int32_t>& beg, DevBuf<int32_t>& adj, DevBuf<uint32_t>* w, const Changes& C,
              int B, int64_t newE) {
    const int32_t V = h->V;
    const int blk = (int)h->prm.block;
    DevBuf<int32_t> ic, dc, nd;
    ic.ensure((size_t)V + 1);
    dc.ensure((size_t)V + 1);
    nd.ensure((size_t)V + 1);
    fillI32(h, ic.p, 0, (int64_t)V + 1);
    fillI32(h, dc.p, 0, (int64_t)V + 1);
    if (C.ni) launch(h, "batch_rowcount", kRowCount, gridFor(C.ni, blk), blk, 0, (const uint64_t*)C.ik.p, C.ni, B, ic.p);
    if (C.nd) launch(h, "batch_rowcount", kRowCount, gridFor(C.nd, blk), blk, 0, (const uint64_t*)C.dk.p, C.nd, B, dc.p);
    DevBuf<int64_t> ib, db, nb;
    ib.ensure((size_t)V + 1);
    db.ensure((size_t)V + 1);
    nb.ensure((size_t)V + 1);
    scanI32toI64(h, ic.p, ib.p, V);
    scanI32toI64(h, dc.p, db.p, V);
    launch(h, "batch_newdeg", kNewDeg, gridFor(V, blk), blk, 0, (const int32_t*)beg.p, (const int32_t*)ic.p,
           (const int32_t*)dc.p, V, nd.p);
    scanI32toI64(h, nd.p, nb.p, V);
    DevBuf<int32_t> nadj;
    nadj.ensure((size_t)(newE > 0 ? newE : 1));
    DevBuf<uint32_t> nw;
    if (w) nw.ensure((size_t)(newE > 0 ? newE : 1));
    int64_t* err = h->dcnt.p + 48;
    fillI64(h, err, 0, 2);
    launch(h, "batch_merge_old", kMergeOld, gridFor(V, blk), blk, 0, (const int32_t*)beg.p, (const int32_t*)adj.p,
           (const uint32_t*)(w ? w->p : nullptr), V, B, (const int64_t*)ib.p, (const uint64_t*)C.ik.p,
           (const int64_t*)db.p, (const uint64_t*)C.dk.p, (const int64_t*)nb.p, nadj.p, w ? nw.p : (uint32_t*)nullptr,
           err);
    if (C.ni)
        launch(h, "batch_merge_ins", kMergeIns, gridFor(C.ni, blk), blk, 0, (const int32_t*)beg.p,
               (const int32_t*)adj.p, B, (const uint64_t*)C.ik.p, (const uint32_t*)(w ? C.iw.p : nullptr), C.ni,
               (const int64_t*)ib.p, (const int64_t*)db.p, (const uint64_t*)C.dk.p, (const int64_t*)nb.p, nadj.p,
               w ? nw.p : (uint32_t*)nullptr, err);
    // the new row sizes must add up to the new arc count
    HIPCHK(hipMemcpyAsync(err + 1, nb.p + V, sizeof(int64_t), hipMemcpyDeviceToDevice, h->st));
    int64_t e2[2];
    readCounters(h, err, 2, e2);
    if (e2[0] != 0 || e2[1] != newE) raise(GIRHIP_EHIP, "batch: internal merge inconsistency");
    launch(h, "batch_beg", kI64toI32, gridFor((int64_t)V + 1, blk), blk, 0, (const int64_t*)nb.p, beg.p,
           (int64_t)V + 1);
    syncStream(h);
    adj.swap(nadj);
    if (w) w->swap(nw);
}

// Normalise + classify the batch; fills C (fw order) and the stats.
void normalizeBatchDev(girhip_graph* h, const int32_t* u, const int32_t* v, const uint8_t* ins, const uint32_t* w,
                       int64_t n, Changes& C, girhip_batch_stats& st, int B) {
    const int32_t V = h->V;
    const int blk = (int)h->prm.block;
    Scratch& S = scratch(h);
    DevBuf<int32_t> du, dv;
    DevBuf<uint8_t> dins;
    DevBuf<uint32_t> dw;
    du.ensure((size_t)n);
    dv.ensure((size_t)n);
    dins.ensure((size_t)n);
    copyIn(h, du.p, u, sizeof(int32_t) * n);
    copyIn(h, dv.p, v, sizeof(int32_t) * n);
    copyIn(h, dins.p, ins, (size_t)n);
    if (w && h->has_w) {
        dw.ensure((size_t)n);
        copyIn(h, dw.p, w, sizeof(uint32_t) * n);
    }
    S.keys[0].ensure((size_t)n);
    S.vals[0].ensure((size_t)n);
    fillI64(h, h->dcnt.p + 40, 0, 4);
    launch(h, "batch_keys", kBatchKeys, gridFor(n, blk), blk, 0, (const int32_t*)du.p, (const int32_t*)dv.p, n, V, B,
           S.keys[0].p, S.vals[0].p, h->dcnt.p + 40);
    radixSort(h, S.keys[0].p, S.vals[0].p, n, 2 * B + 1);
    DevBuf<int32_t> fi, fd;
    fi.ensure((size_t)n);
    fd.ensure((size_t)n);
    launch(h, "batch_classify", kBatchClassify, gridFor(n, blk), blk, 0, (const uint64_t*)S.keys[0].p,
           (const uint32_t*)S.vals[0].p, n, B, (const uint8_t*)dins.p, (const int32_t*)h->fw_beg.p,
           (const int32_t*)h->fw_adj.p, fi.p, fd.p, h->dcnt.p + 41);
    DevBuf<int64_t> pi, pd;
    pi.ensure((size_t)n + 1);
    pd.ensure((size_t)n + 1);
    scanI32toI64(h, fi.p, pi.p, n);
    scanI32toI64(h, fd.p, pd.p, n);
    HIPCHK(hipMemcpyAsync(h->dcnt.p + 42, pi.p + n, sizeof(int64_t), hipMemcpyDeviceToDevice, h->st));
    HIPCHK(hipMemcpyAsync(h->dcnt.p + 43, pd.p + n, sizeof(int64_t), hipMemcpyDeviceToDevice, h->st));
    int64_t c[4];
    readCounters(h, h->dcnt.p + 40, 4, c);
    st.invalid = c[0];
    st.noop = c[1];
    C.ni = c[2];
    C.nd = c[3];
    st.inserted = C.ni;
    st.deleted = C.nd;
    C.ik.ensure((size_t)(C.ni > 0 ? C.ni : 1));
    C.dk.ensure((size_t)(C.nd > 0 ? C.nd : 1));
    if (h->has_w) C.iw.ensure((size_t)(C.ni > 0 ? C.ni : 1));
    launch(h, "batch_compact", kBatchCompact, gridFor(n, blk), blk, 0, (const uint64_t*)S.keys[0].p,
           (const uint32_t*)S.vals[0].p, n, B, (const int32_t*)fi.p, (const int64_t*)pi.p, (const uint32_t*)dw.p,
           !(w && h->has_w), C.ik.p, h->has_w ? C.iw.p : (uint32_t*)nullptr);
    launch(h, "batch_compact", kBatchCompact, gridFor(n, blk), blk, 0, (const uint64_t*)S.keys[0].p,
           (const uint32_t*)S.vals[0].p, n, B, (const int32_t*)fd.p, (const int64_t*)pd.p, (const uint32_t*)nullptr,
           true, C.dk.p, (uint32_t*)nullptr);
    syncStream(h);
}

// Transposed change lists (sorted by (col, row)) for the bw CSR.
void transposeChanges(girhip_graph* h, const Changes& C, Changes& T, int B) {
    const int blk = (int)h->prm.block;
    T.ni = C.ni;
    T.nd = C.nd;
    T.ik.ensure((size_t)(C.ni > 0 ? C.ni : 1));
    T.dk.ensure((size_t)(C.nd > 0 ? C.nd : 1));
    if (C.ni) {
        launch(h, "batch_tkeys", kTransposeKeys, gridFor(C.ni, blk), blk, 0, (const uint64_t*)C.ik.p, C.ni, B, T.ik.p,
               (uint32_t*)nullptr);
        radixSort(h, T.ik.p, nullptr, C.ni, 2 * B);
    }
    if (C.nd) {
        launch(h, "batch_tkeys", kTransposeKeys, gridFor(C.nd, blk), blk, 0, (const uint64_t*)C.dk.p, C.nd, B, T.dk.p,
               (uint32_t*)nullptr);
        radixSort(h, T.dk.p, nullptr, C.nd, 2 * B);
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
    mergeCsr(h, h->fw_beg, h->fw_adj, h->has_w ? &h->w_fw : nullptr, C, B, newE);
    mergeCsr(h, h->bw_beg, h->bw_adj, nullptr, T, B, newE);
    h->E = newE;
    h->src_ok = false;
    h->max_deg = -1;
    h->max_w = -1;
}

// Bring every requested layer up to date.  C = the changes just applied to a
// store the layers were current with (nullptr: layers are stale, rebuild).
void repairLayers(girhip_graph* h, const Changes* C, int B) {
    const bool changed = C && (C->ni + C->nd) > 0;
    if (h->wcc_req) {
        if (h->wcc_ok && C && C->nd == 0) {
            if (C->ni) runWccHookArcs(h, C->ik.p, C->ni, B);  // insert-only: hook the new arcs
        } else if (!h->wcc_ok || changed) {
            runWcc(h);
        }
        h->wcc_ok = true;
    }
    if (h->scc_req) {
        if (!h->scc_ok || changed) {
            runScc(h);
            runDag(h);
        }
        h->scc_ok = true;
    }
}

}  // namespace
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

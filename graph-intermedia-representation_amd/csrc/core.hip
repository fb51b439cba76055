// Handle lifecycle, the device edge store (validation, normalisation,
// transpose), weights, knobs, timing and small queries.
//
// Replaces the reference's GIR constructors / initAfterLoad /
// normalizeBaseCsr (gir.cpp:38-63, 95-130) and graph::graph's validation
// (graph.cpp:95-114), with the arrays resident in HBM.
#include <cstring>

#include "common.h"

namespace gir {

namespace {
thread_local std::string g_last_error;
}

void setLastError(const std::string& m) { g_last_error = m; }
[[noreturn]] void raise(int code, const std::string& msg) { throw Error{code, msg}; }

void useDevice(girhip_graph* h) {
    if (!h) raise(GIRHIP_EINVAL, "null handle");
    HIPCHK(hipSetDevice(h->device));
}

hipEvent_t takeEvent(girhip_graph* h) {
    if (!h->ev_pool.empty()) {
        hipEvent_t e = h->ev_pool.back();
        h->ev_pool.pop_back();
        return e;
    }
    hipEvent_t e;
    HIPCHK(hipEventCreate(&e));
    return e;
}

void flushKernelStats(girhip_graph* h) {
    for (auto& p : h->pend_ev) {
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, p.a, p.b));
        KStat& k = h->kstats[p.name];
        k.count++;
        k.ms += ms;
        h->ev_pool.push_back(p.a);
        h->ev_pool.push_back(p.b);
    }
    h->pend_ev.clear();
}

void syncStream(girhip_graph* h) {
    const bool chk = h->rs_check && h->ws.rs_err.p;
    if (chk)
        HIPCHK(hipMemcpyAsync(h->hcnt + HCNT_SLOTS - 1, h->ws.rs_err.p, sizeof(int64_t), hipMemcpyDeviceToHost, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    if (!h->pend_ev.empty()) flushKernelStats(h);
    if (chk) {
        h->rs_check = false;
        if (h->hcnt[HCNT_SLOTS - 1] != 0) {
            h->hcnt[HCNT_SLOTS - 1] = 0;
            fillI64(h, h->ws.rs_err.p, 0, 1);
            HIPCHK(hipStreamSynchronize(h->st));
            raise(GIRHIP_EHIP, "radix sort: a tile waited too long for the tiles before it");
        }
    }
}

void tBegin(girhip_graph* h) { HIPCHK(hipEventRecord(h->t0, h->st)); }
void tEnd(girhip_graph* h, const char* kind) {
    HIPCHK(hipEventRecord(h->t1, h->st));
    syncStream(h);
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, h->t0, h->t1));
    h->last_ms[kind] = ms;
}

bool isDevicePtr(const void* p) {
    if (!p) return false;
    hipPointerAttribute_t a;
    hipError_t e = hipPointerGetAttributes(&a, p);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged;
}

void copyIn(girhip_graph* h, void* dst_dev, const void* src, size_t bytes) {
    if (bytes == 0) return;
    HIPCHK(hipMemcpyAsync(dst_dev, src, bytes, hipMemcpyDefault, h->st));
}
void copyOut(girhip_graph* h, void* dst, const void* src_dev, size_t bytes) {
    if (bytes == 0 || dst == src_dev) return;
    HIPCHK(hipMemcpyAsync(dst, src_dev, bytes, hipMemcpyDefault, h->st));
}

HeavyList heavyList(girhip_graph* h) {
    Scratch& S = h->ws;
    // a row is listed at most once per source row of a launch (batched SSSP
    // runs up to 32 sources per launch)
    const size_t cap = (size_t)(34 * (h->E / HEAVY_T) + 64);
    if (S.hv_cap < cap || !S.hv_ctl.p) {
        S.hv_ctl.ensure(2);
        S.hv_row.ensure(cap);
        S.hv_aux.ensure(cap);
        S.hv_b.ensure(cap);
        S.hv_e.ensure(cap);
        S.hv_pay.ensure(cap);
        S.hv_off.ensure(cap + 1);
        fillI64(h, S.hv_ctl.p, 0, 2);
        S.hv_cap = cap;
    }
    return HeavyList{S.hv_ctl.p, S.hv_row.p, S.hv_aux.p, S.hv_b.p, S.hv_e.p, S.hv_pay.p, S.hv_off.p};
}

void readCounters(girhip_graph* h, const int64_t* dev, int n, int64_t* out) {
    if (n > HCNT_SLOTS - 1) raise(GIRHIP_EINVAL, "readCounters: too many slots");
    HIPCHK(hipMemcpyAsync(h->hcnt, dev, sizeof(int64_t) * n, hipMemcpyDeviceToHost, h->st));
    syncStream(h);
    std::memcpy(out, h->hcnt, sizeof(int64_t) * n);
}

// ---------------------------------------------------------------------------
// CSR validation (graph.cpp:95-114): offsets monotone from 0 to E, ids in
// [0, V).  Also reports whether every row is strictly ascending (canonical).
// cnt[0] = bad offsets, cnt[1] = bad ids, cnt[2] = non-canonical slots.
__global__ void kCheckOffsets(const int32_t* __restrict__ beg, int32_t V, int64_t E, int64_t* cnt) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool bad = false;
    if (i < V) bad = beg[i] > beg[i + 1];
    if (i == 0) bad = bad || beg[0] != 0 || (int64_t)beg[V] != E;
    if (bad) atomicAdd((unsigned long long*)&cnt[0], 1ULL);
}

__global__ void kCheckIds(const int32_t* __restrict__ adj, int64_t E, int32_t V, int64_t* cnt) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool bad = i < E && (adj[i] < 0 || adj[i] >= V);
    const uint64_t m = __ballot(bad);
    if (m && laneId() == __ffsll((long long)m) - 1) atomicAdd((unsigned long long*)&cnt[1], (unsigned long long)__popcll(m));
}

// Non-canonical slots: adj[j] >= adj[j+1] inside one row.
__global__ void kCheckSorted(const int32_t* __restrict__ beg, const int32_t* __restrict__ adj, int32_t V,
                             int64_t* cnt) {
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool act = v < V;
    const int32_t b = act ? beg[v] : 0, e = act ? beg[v + 1] : 0;
    int64_t bad = 0;
    waveRows(act, (int32_t)v, b, e - 1 > b ? e - 1 : b, [&](int32_t, int64_t j) { bad += adj[j] >= adj[j + 1]; });
    if (bad) atomicAdd((unsigned long long*)&cnt[2], (unsigned long long)bad);
}

__global__ void kCountDiff(const int32_t* __restrict__ a, const int32_t* __restrict__ b, int64_t n, int64_t* cnt) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t m = __ballot(i < n && a[i] != b[i]);
    if (m && laneId() == __ffsll((long long)m) - 1) atomicAdd((unsigned long long*)cnt, (unsigned long long)__popcll(m));
}

// (row, neighbour) -> packed key (row << B) | neighbour; val = slot index.
__global__ void kRowKeys(const int32_t* __restrict__ beg, const int32_t* __restrict__ adj, int32_t V, int B,
                         bool transpose, uint64_t* __restrict__ keys, uint32_t* __restrict__ vals) {
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool act = v < V;
    const int32_t b = act ? beg[v] : 0, e = act ? beg[v + 1] : 0;
    waveRows(act, (int32_t)v, b, e, [&](int32_t u, int64_t j) {
        const uint64_t x = (uint64_t)(uint32_t)adj[j];
        keys[j] = transpose ? ((x << B) | (uint32_t)u) : (((uint64_t)(uint32_t)u << B) | x);
        if (vals) vals[j] = (uint32_t)j;
    });
}

// Unique flags over sorted keys: flag = first of its run.
__global__ void kUniqueFlags(const uint64_t* __restrict__ keys, int64_t n, int32_t* __restrict__ flag) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) flag[i] = i == 0 || keys[i] != keys[i - 1];
}

// Row offsets of the deduplicated CSR: beg[r] = number of unique keys with a
// row below r = pos[lower_bound(keys, r << B)] (pos = exclusive scan of the
// unique flags, pos[n] = total).  No atomics: hub rows cost nothing extra.
__global__ void kBegFromKeys(const uint64_t* __restrict__ keys, int64_t n, const int64_t* __restrict__ pos, int B,
                             int32_t V, int32_t* __restrict__ beg) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r > V) return;
    const uint64_t key = (uint64_t)r << B;
    int64_t lo = 0, hi = n;
    while (lo < hi) {
        const int64_t m = lo + ((hi - lo) >> 1);
        if (keys[m] < key) lo = m + 1;
        else hi = m;
    }
    beg[r] = (int32_t)pos[lo];
}

__global__ void kCompactAdj(const uint64_t* __restrict__ keys, const int32_t* __restrict__ flag,
                            const int64_t* __restrict__ pos, int64_t n, uint64_t mask, int32_t* __restrict__ adj,
                            const uint32_t* __restrict__ vals, uint32_t* __restrict__ vout) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !flag[i]) return;
    adj[pos[i]] = (int32_t)(keys[i] & mask);
    if (vout) vout[pos[i]] = vals[i];
}

namespace {

struct CheckResult {
    int64_t bad_off, bad_id, unsorted;
};

CheckResult checkCsr(girhip_graph* h, const int32_t* beg, const int32_t* adj, int32_t V, int64_t E) {
    fillI64(h, h->dcnt.p, 0, 4);
    launch(h, "check_offsets", kCheckOffsets, gridFor((int64_t)V + 1, 256), 256, 0, beg, V, E, h->dcnt.p);
    int64_t c[3];
    readCounters(h, h->dcnt.p, 3, c);
    if (c[0]) return {c[0], 0, 0};
    launch(h, "check_ids", kCheckIds, gridFor(E, 256), 256, 0, adj, E, V, h->dcnt.p);
    readCounters(h, h->dcnt.p, 3, c);
    if (c[1]) return {c[0], c[1], 0};
    launch(h, "check_sorted", kCheckSorted, gridFor(V, 256), 256, 0, beg, adj, V, h->dcnt.p);
    readCounters(h, h->dcnt.p, 3, c);
    return {c[0], c[1], c[2]};
}

}  // namespace

// Sorted packed keys (row << B | neighbour) in S.keys[0] -> deduplicated CSR
// (beg: V+1 entries, adj grown to the unique count).  Returns that count.
static int64_t keysToCsr(girhip_graph* h, int64_t E, int B, DevBuf<int32_t>& beg, DevBuf<int32_t>& adj) {
    const int32_t V = h->V;
    Scratch& S = scratch(h);
    S.iv[0].ensure(E);
    launch(h, "csr_unique", kUniqueFlags, gridFor(E, 256), 256, 0, (const uint64_t*)S.keys[0].p, E, S.iv[0].p);
    S.off2.ensure((size_t)E + 1);
    DevBuf<int64_t>& pos = S.off2;
    scanI32toI64(h, S.iv[0].p, pos.p, E);
    int64_t nE;
    readCounters(h, pos.p + E, 1, &nE);
    adj.ensure((size_t)(nE > 0 ? nE : 1));
    const uint64_t mask = (1ULL << B) - 1;
    launch(h, "csr_compact", kCompactAdj, gridFor(E, 256), 256, 0, (const uint64_t*)S.keys[0].p,
           (const int32_t*)S.iv[0].p, (const int64_t*)pos.p, E, mask, adj.p, (const uint32_t*)nullptr,
           (uint32_t*)nullptr);
    beg.ensure((size_t)V + 1);
    launch(h, "csr_beg", kBegFromKeys, gridFor((int64_t)V + 1, 256), 256, 0, (const uint64_t*)S.keys[0].p, E,
           (const int64_t*)pos.p, B, V, beg.p);
    syncStream(h);
    return nE;
}

// Sort + dedup every row on the device (gir.cpp:95-116) through one global
// radix sort of packed (row, neighbour) keys.  Self loops are kept, as the
// reference keeps them.
void normalizeCsr(girhip_graph* h, DevBuf<int32_t>& beg, DevBuf<int32_t>& adj, int64_t& E) {
    const int32_t V = h->V;
    if (E == 0) return;
    Scratch& S = scratch(h);
    const int B = bitsFor((uint64_t)(V > 1 ? V - 1 : 1));
    S.keys[0].ensure(E);
    launch(h, "norm_keys", kRowKeys, gridFor(V, 256), 256, 0, (const int32_t*)beg.p, (const int32_t*)adj.p, V, B,
           false, S.keys[0].p, (uint32_t*)nullptr);
    radixSort(h, S.keys[0].p, nullptr, E, 2 * B);
    DevBuf<int32_t> nadj;
    E = keysToCsr(h, E, B, beg, nadj);
    adj.swap(nadj);
}

// bw CSR = transpose of fw: radix sort of packed (neighbour, row) keys.
void buildTranspose(girhip_graph* h) {
    const int32_t V = h->V;
    const int64_t E = h->E;
    h->bw_beg.ensure((size_t)V + 1);
    h->bw_adj.ensure((size_t)(E > 0 ? E : 1));
    if (E == 0) {
        fillI32(h, h->bw_beg.p, 0, (int64_t)V + 1);
        return;
    }
    Scratch& S = scratch(h);
    const int B = bitsFor((uint64_t)(V > 1 ? V - 1 : 1));
    S.keys[0].ensure(E);
    launch(h, "tr_keys", kRowKeys, gridFor(V, 256), 256, 0, (const int32_t*)h->fw_beg.p, (const int32_t*)h->fw_adj.p,
           V, B, true, S.keys[0].p, (uint32_t*)nullptr);
    radixSort(h, S.keys[0].p, nullptr, E, 2 * B);
    keysToCsr(h, E, B, h->bw_beg, h->bw_adj);
}

void invalidateLayers(girhip_graph* h) {
    h->wcc_ok = false;
    h->scc_ok = false;
}

__global__ void kRowOfSlot(const int32_t* __restrict__ beg, int32_t V, int32_t* __restrict__ src) {
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool act = v < V;
    const int32_t b = act ? beg[v] : 0, e = act ? beg[v + 1] : 0;
    waveRows(act, (int32_t)v, b, e, [&](int32_t u, int64_t j) { src[j] = u; });
}

// (bounded grids; one atomic per block)
__global__ void kMaxRow(const int32_t* __restrict__ beg, int32_t V, int64_t* out) {
    uint64_t m = 0;
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < V; v += (int64_t)gridDim.x * blockDim.x)
        m = std::max<uint64_t>(m, (uint64_t)(beg[v + 1] - beg[v]));
    m = blockMaxU64(m);
    if (threadIdx.x == 0 && m) atomicMax((unsigned long long*)out, (unsigned long long)m);
}

int64_t maxDegree(girhip_graph* h) {
    if (h->max_deg < 0) {
        fillI64(h, h->dcnt.p, 0, 1);
        const unsigned g = boundedGrid(h->V, 256);
        launch(h, "max_row", kMaxRow, g, 256, 0, (const int32_t*)h->fw_beg.p, h->V, h->dcnt.p);
        launch(h, "max_row", kMaxRow, g, 256, 0, (const int32_t*)h->bw_beg.p, h->V, h->dcnt.p);
        readCounters(h, h->dcnt.p, 1, &h->max_deg);
    }
    return h->max_deg;
}

__global__ void kMaxWeight(const uint32_t* __restrict__ w, int64_t E, int64_t* out) {
    uint64_t m = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < E; i += (int64_t)gridDim.x * blockDim.x)
        m = std::max<uint64_t>(m, w[i]);
    m = blockMaxU64(m);
    if (threadIdx.x == 0 && m) atomicMax((unsigned long long*)out, (unsigned long long)m);
}

bool persistentGrid(girhip_graph* h, const void* kernel, int block, int blocks_per_cu, unsigned* grid) {
    if (h->prm.persist == 0) return false;
    if (h->coop < 0) {
        int c = 0;
        if (hipDeviceGetAttribute(&c, hipDeviceAttributeCooperativeLaunch, h->device) != hipSuccess) c = 0;
        int n = 0;
        if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, h->device) != hipSuccess) n = 0;
        (void)hipGetLastError();
        h->n_cu = n;
        h->coop = (c && n > 0) ? 1 : 0;
    }
    if (!h->coop) return false;
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, block, 0) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    if (per_cu < 1) return false;
    *grid = (unsigned)(std::min(per_cu, blocks_per_cu) * h->n_cu);
    return true;
}

void checkGridBar(girhip_graph* h) {
    uint32_t err = 0;
    HIPCHK(hipMemcpyAsync(&err, h->gbar.p + BAR_ERR, sizeof(err), hipMemcpyDeviceToHost, h->st));
    syncStream(h);
    if (err) raise(GIRHIP_EHIP, "persistent kernel: a grid barrier wait gave up");
}

int64_t maxWeight(girhip_graph* h) {
    if (!h->has_w) return 100;
    if (h->max_w < 0) {
        fillI64(h, h->dcnt.p, 0, 1);
        launch(h, "max_weight", kMaxWeight, boundedGrid(h->E, 256), 256, 0, (const uint32_t*)h->w_fw.p, h->E,
               h->dcnt.p);
        readCounters(h, h->dcnt.p, 1, &h->max_w);
    }
    return h->max_w;
}

// Both orientations of every arc as packed keys; self loops get the
// sentinel 1 << 2B (sorts after every real key).
__global__ void kSymKeys(const int32_t* __restrict__ src, const int32_t* __restrict__ adj, int64_t E, int B,
                         uint64_t* __restrict__ keys) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= E) return;
    const uint32_t u = (uint32_t)src[j], v = (uint32_t)adj[j];
    const uint64_t sent = 1ULL << (2 * B);
    keys[2 * j] = u == v ? sent : (((uint64_t)u << B) | v);
    keys[2 * j + 1] = u == v ? sent : (((uint64_t)v << B) | u);
}

__global__ void kCountBelow(const uint64_t* __restrict__ keys, int64_t n, uint64_t x, int64_t* out) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    int64_t lo = 0, hi = n;
    while (lo < hi) {
        const int64_t m = lo + ((hi - lo) >> 1);
        if (keys[m] < x) lo = m + 1;
        else hi = m;
    }
    *out = lo;
}

int64_t symCsr(girhip_graph* h) {
    if (h->sym_ok) return h->sym_E;
    const int32_t V = h->V;
    const int64_t E = h->E;
    const int B = bitsFor((uint64_t)(V > 1 ? V - 1 : 1));
    if (E == 0) {
        h->sym_beg.ensure((size_t)V + 1);
        h->sym_adj.ensure(1);
        fillI32(h, h->sym_beg.p, 0, (int64_t)V + 1);
        h->sym_E = 0;
        h->sym_ok = true;
        return 0;
    }
    Scratch& S = scratch(h);
    S.keys[0].ensure((size_t)(2 * E));
    const int32_t* src = fwSrc(h);
    launch(h, "sym_keys", kSymKeys, gridFor(E, 256), 256, 0, src, (const int32_t*)h->fw_adj.p, E, B, S.keys[0].p);
    radixSort(h, S.keys[0].p, nullptr, 2 * E, 2 * B + 1);
    launch(h, "sym_count", kCountBelow, 1, 64, 0, (const uint64_t*)S.keys[0].p, 2 * E, 1ULL << (2 * B), h->dcnt.p);
    int64_t n2;
    readCounters(h, h->dcnt.p, 1, &n2);
    // offsets are int32 (as in the forward / backward CSR)
    if (n2 > (int64_t)INT32_MAX)
        raise(GIRHIP_EINVAL, "symmetric closure holds more than 2^31 - 1 arcs (int32 offsets)");
    h->sym_E = n2 > 0 ? keysToCsr(h, n2, B, h->sym_beg, h->sym_adj) : 0;
    if (n2 == 0) {
        h->sym_beg.ensure((size_t)V + 1);
        h->sym_adj.ensure(1);
        fillI32(h, h->sym_beg.p, 0, (int64_t)V + 1);
    }
    h->sym_ok = true;
    return h->sym_E;
}

const int32_t* fwSrc(girhip_graph* h) {
    if (!h->src_ok) {
        h->fw_src.ensure((size_t)(h->E > 0 ? h->E : 1));
        launch(h, "row_of_slot", kRowOfSlot, gridFor(h->V, 256), 256, 0, (const int32_t*)h->fw_beg.p, h->V,
               h->fw_src.p);
        h->src_ok = true;
    }
    return h->fw_src.p;
}

}  // namespace gir

using namespace gir;

extern "C" const char* girhip_last_error(void) { return g_last_error.c_str(); }

extern "C" int girhip_create(const int32_t* fw_beg, const int32_t* fw_adj, const int32_t* bw_beg,
                             const int32_t* bw_adj, int32_t V, int64_t E, girhip_graph** out) {
    return guarded([&] {
        if (!out) raise(GIRHIP_EINVAL, "create: null output handle");
        *out = nullptr;
        if (V < 0 || E < 0 || E >= 0x7fffffffLL) raise(GIRHIP_EINVAL, "create: V/E out of range for int32 CSR");
        if (!fw_beg || (E > 0 && !fw_adj)) raise(GIRHIP_EINVAL, "create: null forward CSR");
        if ((!bw_beg && bw_adj) || (bw_beg && E > 0 && !bw_adj))
            raise(GIRHIP_EINVAL, "create: backward CSR must give both arrays or neither");
        int ndev = 0;
        if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
            (void)hipGetLastError();
            raise(GIRHIP_ENODEV, "create: no HIP device");
        }
        auto* h = new girhip_graph();
        int rc = guarded([&] {
            HIPCHK(hipGetDevice(&h->device));
            HIPCHK(hipStreamCreateWithFlags(&h->st, hipStreamNonBlocking));
            // the allocation-free entry point's array: a failure here only
            // disables that entry point (it then reports GIRHIP_ESTATE)
            try {
                h->na_scratch = noAllocScratch(h->device);
            } catch (const Error&) {
                h->na_scratch = nullptr;
            }
            HIPCHK(hipEventCreate(&h->t0));
            HIPCHK(hipEventCreate(&h->t1));
            HIPCHK(hipHostMalloc((void**)&h->hcnt, sizeof(int64_t) * HCNT_SLOTS, hipHostMallocDefault));
            h->dcnt.ensure(64);
            h->V = V;
            h->E = E;
            h->fw_beg.ensure((size_t)V + 1);
            h->fw_adj.ensure((size_t)(E > 0 ? E : 1));
            copyIn(h, h->fw_beg.p, fw_beg, sizeof(int32_t) * ((size_t)V + 1));
            copyIn(h, h->fw_adj.p, fw_adj, sizeof(int32_t) * (size_t)E);
            CheckResult c = checkCsr(h, h->fw_beg.p, h->fw_adj.p, V, E);
            if (c.bad_off || c.bad_id)
                raise(GIRHIP_EINVAL,
                      "create: forward CSR is malformed (offsets not monotone from 0 to E, or neighbour ids outside [0,V))");
            if (c.unsorted) normalizeCsr(h, h->fw_beg, h->fw_adj, h->E);
            if (bw_beg) {
                // The reference only compares arc counts (gir.cpp:122-126); the
                // update kernels rely on bw being the exact transpose, so the
                // given bw is checked against the transpose built here.
                int64_t Eb = E;
                DevBuf<int32_t> ub, ua;
                ub.ensure((size_t)V + 1);
                ua.ensure((size_t)(E > 0 ? E : 1));
                copyIn(h, ub.p, bw_beg, sizeof(int32_t) * ((size_t)V + 1));
                copyIn(h, ua.p, bw_adj, sizeof(int32_t) * (size_t)E);
                CheckResult d = checkCsr(h, ub.p, ua.p, V, E);
                if (d.bad_off || d.bad_id)
                    raise(GIRHIP_EINVAL,
                          "create: backward CSR is malformed (offsets not monotone from 0 to E, or neighbour ids outside [0,V))");
                if (d.unsorted) normalizeCsr(h, ub, ua, Eb);
                if (Eb != h->E)
                    raise(GIRHIP_EINVAL, "create: forward and backward CSR disagree after deduplication (" +
                                             std::to_string(h->E) + " vs " + std::to_string(Eb) +
                                             " arcs); the bw arrays are not the transpose of fw");
                buildTranspose(h);
                fillI64(h, h->dcnt.p, 0, 1);
                launch(h, "check_equal", kCountDiff, gridFor((int64_t)V + 1, 256), 256, 0, (const int32_t*)ub.p,
                       (const int32_t*)h->bw_beg.p, (int64_t)V + 1, h->dcnt.p);
                launch(h, "check_equal", kCountDiff, gridFor(Eb, 256), 256, 0, (const int32_t*)ua.p,
                       (const int32_t*)h->bw_adj.p, Eb, h->dcnt.p);
                int64_t diff;
                readCounters(h, h->dcnt.p, 1, &diff);
                if (diff) raise(GIRHIP_EINVAL, "create: the backward CSR is not the transpose of the forward CSR");
            } else {
                buildTranspose(h);
            }
            // the spare adjacency an update batch merges into, reserved now
            // (two E-sized allocations would otherwise land in the first batch)
            h->fw_adj2.ensureSlack((size_t)(h->E > 0 ? h->E : 1));
            h->bw_adj2.ensureSlack((size_t)(h->E > 0 ? h->E : 1));
            syncStream(h);
        });
        if (rc != GIRHIP_OK) {
            std::string msg = girhip_last_error();
            girhip_destroy(h);
            raise(rc, msg);
        }
        *out = h;
    });
}

extern "C" int girhip_destroy(girhip_graph* h) {
    if (!h) return GIRHIP_OK;
    (void)hipSetDevice(h->device);
    if (h->st) (void)hipStreamSynchronize(h->st);
    for (auto& p : h->pend_ev) {
        (void)hipEventDestroy(p.a);
        (void)hipEventDestroy(p.b);
    }
    for (auto e : h->ev_pool) (void)hipEventDestroy(e);
    if (h->t0) (void)hipEventDestroy(h->t0);
    if (h->t1) (void)hipEventDestroy(h->t1);
    if (h->hcnt) (void)hipHostFree(h->hcnt);
    if (h->st) (void)hipStreamDestroy(h->st);
    delete h;  // (device buffers are freed by their destructors, device current)
    return GIRHIP_OK;
}

extern "C" int girhip_size(const girhip_graph* h, int32_t* V, int64_t* E) {
    return guarded([&] {
        if (!h) raise(GIRHIP_EINVAL, "size: null handle");
        if (V) *V = h->V;
        if (E) *E = h->E;
    });
}

extern "C" int girhip_get_csr(girhip_graph* h, int32_t* fw_beg, int32_t* fw_adj, int32_t* bw_beg, int32_t* bw_adj) {
    return guarded([&] {
        useDevice(h);
        const size_t vb = sizeof(int32_t) * ((size_t)h->V + 1), eb = sizeof(int32_t) * (size_t)h->E;
        if (fw_beg) copyOut(h, fw_beg, h->fw_beg.p, vb);
        if (fw_adj) copyOut(h, fw_adj, h->fw_adj.p, eb);
        if (bw_beg) copyOut(h, bw_beg, h->bw_beg.p, vb);
        if (bw_adj) copyOut(h, bw_adj, h->bw_adj.p, eb);
        syncStream(h);
    });
}

__global__ void kCheckWeights(const uint32_t* __restrict__ w, int64_t n, int64_t* cnt) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && w[i] == 0) atomicAdd((unsigned long long*)cnt, 1ULL);
}

extern "C" int girhip_set_weights(girhip_graph* h, const uint32_t* w_fw) {
    return guarded([&] {
        useDevice(h);
        for (auto& sv : h->split) sv.delta = 0;
        // a tracked incremental SSSP was computed with the old weights: track
        // its source again with the new ones
        auto retrack = [&] {
            if (!h->inc[1].on) return;
            const int rc = girhip_inc_track(h, h->inc[1].source, 1);
            if (rc != GIRHIP_OK) raise(rc, girhip_last_error());
        };
        if (!w_fw) {
            h->has_w = false;
            h->w_fw.release();
            retrack();
            return;
        }
        h->max_w = -1;
        h->w_fw.ensure((size_t)(h->E > 0 ? h->E : 1));
        copyIn(h, h->w_fw.p, w_fw, sizeof(uint32_t) * (size_t)h->E);
        fillI64(h, h->dcnt.p, 0, 1);
        launch(h, "check_weights", kCheckWeights, gridFor(h->E, 256), 256, 0, (const uint32_t*)h->w_fw.p, h->E,
               h->dcnt.p);
        int64_t zeros;
        readCounters(h, h->dcnt.p, 1, &zeros);
        if (zeros) {
            h->has_w = false;
            h->w_fw.release();
            raise(GIRHIP_EINVAL, "set_weights: weights must be > 0");
        }
        h->has_w = true;
        retrack();
    });
}

static double* paramSlot(girhip_graph* h, const char* name) {
    Params& p = h->prm;
    static const struct {
        const char* n;
        double Params::*m;
    } tab[] = {{"bfs_alpha", &Params::bfs_alpha},         {"bfs_beta", &Params::bfs_beta},
               {"bfs_min_reach", &Params::bfs_min_reach}, {"sssp_delta", &Params::sssp_delta},
               {"block", &Params::block},                 {"wcc_rounds", &Params::wcc_rounds},
               {"scc_trim_rounds", &Params::scc_trim_rounds}, {"scc_pivot", &Params::scc_pivot},
               {"scc_color_min", &Params::scc_color_min}, {"profile", &Params::profile},
               {"msbfs", &Params::msbfs},         {"chunk", &Params::chunk},
               {"persist", &Params::persist},     {"sssp_dist32", &Params::sssp_dist32},
               {"reach_local", &Params::reach_local}, {"graph", &Params::graph},
               {"xcd", &Params::xcd}, {"split", &Params::split},
               {"scc_block", &Params::scc_block}, {"sssp_interleave", &Params::sssp_interleave},
               {"sssp_ms_batch", &Params::sssp_ms_batch}, {"scc_trim2", &Params::scc_trim2},
               {"scc_incremental", &Params::scc_incremental}, {"sssp_unroll", &Params::sssp_unroll},
               {"bfs_alpha_auto", &Params::bfs_alpha_auto}, {"update_bypass", &Params::update_bypass},
               {"scc_cert", &Params::scc_cert}, {"repair_ops", &Params::repair_ops},
               {"sssp_tiny", &Params::sssp_tiny}, {"bfs_tiny", &Params::bfs_tiny}, {"scc_pair", &Params::scc_pair}};
    for (auto& t : tab)
        if (std::strcmp(t.n, name) == 0) return &(p.*(t.m));
    return nullptr;
}

extern "C" int girhip_set_param(girhip_graph* h, const char* name, double value) {
    return guarded([&] {
        if (!h || !name) raise(GIRHIP_EINVAL, "set_param: null argument");
        double* s = paramSlot(h, name);
        if (!s) raise(GIRHIP_EINVAL, std::string("set_param: unknown parameter ") + name);
        if (std::strcmp(name, "sssp_delta") == 0 && !(value >= 1)) raise(GIRHIP_EINVAL, "set_param: sssp_delta >= 1");
        if (std::strcmp(name, "block") == 0 && !(value == 64 || value == 128 || value == 256 || value == 512 || value == 1024))
            raise(GIRHIP_EINVAL, "set_param: block must be 64, 128, 256, 512 or 1024");
        if (std::strcmp(name, "chunk") == 0 && !(value >= 1 && value <= 16))
            raise(GIRHIP_EINVAL, "set_param: chunk must be in [1, 16]");
        if (std::strcmp(name, "sssp_ms_batch") == 0 && !(value >= 1 && value <= 32))
            raise(GIRHIP_EINVAL, "set_param: sssp_ms_batch must be in [1, 32]");
        if (std::strcmp(name, "sssp_interleave") == 0 && !(value == 0 || value == 1 || value == 2))
            raise(GIRHIP_EINVAL, "set_param: sssp_interleave must be 0, 1 or 2");
        if (std::strcmp(name, "sssp_unroll") == 0 && !(value == 1 || value == 2 || value == 4))
            raise(GIRHIP_EINVAL, "set_param: sssp_unroll must be 1, 2 or 4");
        if (!(value == value)) raise(GIRHIP_EINVAL, "set_param: NaN");
        *s = value;
    });
}

extern "C" int girhip_get_param(const girhip_graph* h, const char* name, double* value) {
    return guarded([&] {
        if (!h || !name || !value) raise(GIRHIP_EINVAL, "get_param: null argument");
        double* s = paramSlot(const_cast<girhip_graph*>(h), name);
        if (!s) raise(GIRHIP_EINVAL, std::string("get_param: unknown parameter ") + name);
        *value = *s;
    });
}

extern "C" int girhip_last_ms(const girhip_graph* h, const char* kind, double* ms) {
    return guarded([&] {
        if (!h || !kind || !ms) raise(GIRHIP_EINVAL, "last_ms: null argument");
        auto it = h->last_ms.find(kind);
        if (it == h->last_ms.end()) raise(GIRHIP_ESTATE, std::string("last_ms: nothing timed for ") + kind);
        *ms = it->second;
    });
}

__global__ void kHasEdge(const int32_t* __restrict__ beg, const int32_t* __restrict__ adj, int32_t u, int32_t v,
                         int64_t* out) {
    // one wave: binary search in row u
    if (threadIdx.x != 0) return;
    int32_t lo = beg[u], hi = beg[u + 1];
    while (lo < hi) {
        int32_t m = lo + (hi - lo) / 2;
        if (adj[m] < v) lo = m + 1;
        else hi = m;
    }
    out[0] = (lo < beg[u + 1] && adj[lo] == v) ? 1 : 0;
}

extern "C" int girhip_has_edge(girhip_graph* h, int32_t u, int32_t v, int32_t* present) {
    return guarded([&] {
        useDevice(h);
        if (!present) raise(GIRHIP_EINVAL, "has_edge: null output");
        if (u < 0 || v < 0 || u >= h->V || v >= h->V) {
            *present = 0;
            return;
        }
        launch(h, "has_edge", kHasEdge, 1, 64, 0, (const int32_t*)h->fw_beg.p, (const int32_t*)h->fw_adj.p, u, v,
               h->dcnt.p);
        int64_t r;
        readCounters(h, h->dcnt.p, 1, &r);
        *present = (int32_t)r;
    });
}

// Kernel statistics (profile mode): name -> launches, total ms.
extern "C" int girhip_kernel_stats(girhip_graph* h, const char* name, int64_t* count, double* ms) {
    return guarded([&] {
        if (!h || !name) raise(GIRHIP_EINVAL, "kernel_stats: null argument");
        auto it = h->kstats.find(name);
        if (count) *count = it == h->kstats.end() ? 0 : it->second.count;
        if (ms) *ms = it == h->kstats.end() ? 0 : it->second.ms;
    });
}
extern "C" int girhip_kernel_stat_names(girhip_graph* h, char* buf, int64_t cap, int64_t* needed) {
    return guarded([&] {
        if (!h) raise(GIRHIP_EINVAL, "kernel_stat_names: null handle");
        std::string all;
        for (auto& kv : h->kstats) all += kv.first + "\n";
        if (needed) *needed = (int64_t)all.size() + 1;
        if (buf && cap > 0) {
            const size_t n = std::min<size_t>(all.size(), (size_t)cap - 1);
            std::memcpy(buf, all.data(), n);
            buf[n] = 0;
        }
    });
}

extern "C" int girhip_reset_kernel_stats(girhip_graph* h) {
    return guarded([&] {
        if (!h) raise(GIRHIP_EINVAL, "reset_kernel_stats: null handle");
        h->kstats.clear();
    });
}

// Internal header of the gfx950 GIR library: the handle, device buffers,
// error handling, kernel launch wrapper and wave64 helpers.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <type_traits>
#include <map>
#include <string>
#include <vector>

#include "girhip.h"

namespace gir {

constexpr int WAVE = 64;

// Internal errors are C++ exceptions; every extern "C" entry point catches
// them and turns them into a status code plus girhip_last_error().
struct Error {
    int code;
    std::string msg;
};
[[noreturn]] void raise(int code, const std::string& msg);
void setLastError(const std::string& m);

// (a failed call's error is also taken off the runtime's last-error state,
// so that it is not reported again by a later, unrelated check)
#define HIPCHK(x)                                                                                        \
    do {                                                                                                 \
        hipError_t e_ = (x);                                                                             \
        if (e_ != hipSuccess) {                                                                          \
            (void)hipGetLastError();                                                                     \
            ::gir::raise(e_ == hipErrorOutOfMemory ? GIRHIP_ENOMEM : GIRHIP_EHIP,                        \
                         std::string(#x) + ": " + hipGetErrorString(e_));                                \
        }                                                                                                \
    } while (0)

// Grow-only device buffer.
template <class T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    void ensure(size_t want) {
        if (want <= n && p) return;
        release();
        size_t bytes = sizeof(T) * (want > 0 ? want : 1);
        hipError_t e = hipMalloc((void**)&p, bytes);
        if (e != hipSuccess) {
            p = nullptr;
            (void)hipGetLastError();
            raise(GIRHIP_ENOMEM, "hipMalloc of " + std::to_string(bytes) + " bytes failed");
        }
        n = want;
    }
    // like ensure, but a grown buffer gets 1/16 headroom (arrays that grow a
    // little per update batch)
    void ensureSlack(size_t want) {
        if (want <= n && p) return;
        ensure(want + want / 16 + 64);
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    ~DevBuf() { release(); }
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    void swap(DevBuf& o) {
        std::swap(p, o.p);
        std::swap(n, o.n);
    }
};

// Tunables (only change work order; never results).
struct Params {
    double bfs_alpha = 15, bfs_beta = 18, bfs_min_reach = 4096;
    double sssp_delta = 16;
    double block = 256;
    double wcc_rounds = 2;
    double scc_trim_rounds = 3;
    double scc_pivot = 1;
    double scc_color_min = 0;
    double profile = 0;  // 1: time every kernel launch with HIP events
    double msbfs = 1;    // 1: k>1 BFS runs 64 sources per bit-parallel sweep
    double chunk = 16;   // most BFS levels / SSSP rounds queued per host readback
    double persist = 0;  // 1: tiny-frontier phases run in one persistent cooperative kernel
    double sssp_dist32 = 1;  // 1: SSSP works on 32-bit distances (64-bit redo on overflow)
    double reach_local = 1;  // 1: reachable() first tries a per-pair LDS search
    double graph = 0;        // 1: long runs of full chunks replay captured HIP graphs
    double xcd = 0;          // 1: SSSP source rows pinned to XCDs (block b -> source b mod k)
    double split = 1;        // 1: weighted SSSP walks rows split into light and heavy arcs
    double scc_block = 512;  // threads per block of the SCC/DAG build's kernels
    double sssp_interleave = 2;  // 1: k-source SSSP batches share buckets and interleave distances;
                                 // 0: one layout per source; 2: 1 where it measured faster
    double sssp_ms_batch = 8;    // sources per interleaved batch (1..32)
    double scc_trim2 = 0;        // 1: each SCC trim round also runs trim-2 (pairs)
    double scc_incremental = 1;  // 1: update batches repair the SCC layer; 0: rebuild it
    double sssp_unroll = 2;      // arcs per lane at a time in the SSSP relaxation passes (1, 2 or 4)
    double bfs_alpha_auto = 1;   // 1: the per-source reach raises bfs_alpha to 64 on graphs with hubs
    double update_bypass = 1;    // 1: batches of few deletes test them for a bypass path first
    double scc_cert = 1;         // 1: update batches check the big SCC against its spanning trees
    double repair_ops = 0;       // largest batch (changes) the SCC layer repairs; 0: max(4096, V / 256)
    double sssp_tiny = 1;        // 1: per-source SSSP rounds with small queues run one workgroup per source
    double bfs_tiny = 1;         // 1: BFS on graphs without long rows runs one workgroup per source first
    double scc_pair = 1;         // 1: the SCC build's forward and backward sweeps share their launches
};

struct KStat {
    long long count = 0;
    double ms = 0;
};

// Scratch shared by the kernels of one handle (grown on demand).
struct Scratch {
    DevBuf<int32_t> q[3];
    DevBuf<uint64_t> bits[4];
    DevBuf<int32_t> iv[4];
    DevBuf<uint8_t> bytes[2];
    DevBuf<uint64_t> keys[2];
    DevBuf<uint32_t> vals[2];
    DevBuf<int64_t> off, off2;
    // radix sort: published tile words, tickets + digit counts, error word
    DevBuf<uint64_t> rs_look, rs_aux;
    DevBuf<int64_t> rs_err;
    uint32_t rs_epoch = 0;
    DevBuf<uint8_t> stage;
    DevBuf<uint64_t> ms[3];  // bit-parallel BFS: seen / frontier / next words
    DevBuf<int64_t> mscnt;   // bit-parallel BFS: per-level counter ring
    DevBuf<uint64_t> msbm[4];  // bit-parallel BFS: frontier / touched bitmaps and summaries
    DevBuf<uint64_t> msl;      // bit-parallel BFS: new-bit words of up to 7 dense levels
    DevBuf<int64_t> tcnt;      // SCC trim phase: device-held list lengths, pivot key, round flags
    DevBuf<int64_t> rcnt;    // reachability engine: per-level counter ring
    DevBuf<int64_t> rcnt2;   // the paired sweeps' two rings
    DevBuf<int32_t> pq[2];   // the paired sweeps' second queue pair
    // update batches: per-op arrays (grown to the largest batch; b_pi / b_pd
    // also hold the change positions of the CSR merge) and the new offsets
    DevBuf<int32_t> b_u, b_v, b_fi, b_fd;
    DevBuf<uint8_t> b_ins;
    DevBuf<uint32_t> b_w;
    DevBuf<int64_t> b_pi, b_pd;
    DevBuf<int64_t> b_nb;
    DevBuf<uint64_t> b_ik[2], b_dk[2];  // sorted change keys: fw order, bw order
    DevBuf<uint32_t> b_iw;
    // delta-stepping batch state (sized for the largest source batch)
    DevBuf<int32_t> s_q[2], s_qh;
    DevBuf<uint32_t> s_d32;  // 32-bit distances of a batch
    DevBuf<uint64_t> s_near, s_far, s_done, s_nsum, s_fsum, s_dsum;
    DevBuf<int64_t> s_ctl;
    DevBuf<int32_t> s_fl[2];  // far lists of the one-workgroup SSSP rounds
    // interleaved multi-source delta-stepping: distances [V][KS], near / far /
    // done source masks, their vertex bitmaps and summaries, round queues
    // (vertex, mask) by parity and the heavy queue, control slots
    DevBuf<uint8_t> ms_dist;
    DevBuf<uint64_t> ms_bits[3], ms_sum[3];
    DevBuf<int32_t> ms_q[3];
    DevBuf<int64_t> ms_ctl;
    DevBuf<uint64_t> gbits;  // condensation keys: one bit per vertex, set iff its label is the big one
    DevBuf<uint64_t> bp_bits;  // delete bypass check: one bitmap per delete
    DevBuf<uint8_t> bp_found;
    DevBuf<int32_t> src;     // query sources
    // mixed query / update lists: source degrees, output rows, trivial
    // sources, live-source rows, host-output staging
    DevBuf<int32_t> rb_deg, rb_rows, rb_tsrc;
    DevBuf<uint8_t> rb_live, rb_out;
    // incremental SCC repair: counters, per-SCC "lost an inner arc" flags,
    // condensation keys / signed counts, rank-violating inserts (SCC ends)
    DevBuf<int64_t> ir_cnt;
    DevBuf<uint8_t> ir_dirty;
    DevBuf<uint64_t> ir_keys;
    DevBuf<uint32_t> ir_vals;
    DevBuf<int32_t> ir_bsu, ir_bsv;
    DevBuf<uint8_t> ir_seed;  // distinct SCC ids of a sweep's seeds
    // labels before a relabelling, the vertices whose label changed, counters
    DevBuf<int32_t> ir_old, ir_list;
    DevBuf<int64_t> ir_cnt2;
    // condensation merge: run flags / sums, unique keys and sums, new
    // multiplicities, matched arc, new / removed flags and positions, the new
    // and removed keys (forward and transposed, padded), counters
    DevBuf<int32_t> dm_flag, dm_sum, dm_us, dm_nm, dm_nf, dm_rf, dm_newm;
    DevBuf<int64_t> dm_pos, dm_fi, dm_pnew, dm_prem, dm_cnt;
    DevBuf<uint64_t> dm_uk, dm_newk, dm_remk, dm_newt, dm_remt;
    // big-SCC certificate check: orphaned / still-attached members (out-tree
    // side, in-tree side), the heads / tails of deleted tree arcs
    DevBuf<uint64_t> gc_bits[4];
    DevBuf<int32_t> gc_cut[2];
    DevBuf<int64_t> gc_cnt;
    // reachability query sweeps (a group of sweeps, V words each)
    DevBuf<uint64_t> rq_seen, rq_next, rq_tgt, rq_fb[2];
    DevBuf<int32_t> rq_fq[2], rq_tq, rq_pairs;
    DevBuf<int64_t> rq_ctl, rq_lim;
    // condensation-guided queries: reachable-SCC bitmap, reached list,
    // pending predecessor counts, wave lists, vertex list, control slots
    DevBuf<uint64_t> tl_mask;
    DevBuf<int32_t> tl_list, tl_left, tl_rdy, tl_vl;
    DevBuf<int64_t> tl_ctl;
    // betweenness centrality batch: levels, path counts, dependencies (fp64
    // bits), visit orders, level starts, counters
    DevBuf<int32_t> bc_lev, bc_ord, bc_lv;
    DevBuf<uint64_t> bc_sigma, bc_delta;
    DevBuf<int64_t> bc_ctl;
    // incremental queries: from-scratch distances / levels, presence of the
    // logged arcs, invalid list, round queues, flags (zero between updates),
    // counters
    DevBuf<uint64_t> inc_d;
    DevBuf<int32_t> inc_lv, inc_list, inc_q[2], inc_flag;
    DevBuf<uint8_t> inc_pres;
    DevBuf<int64_t> inc_ctl;
    // biconnected blocks: BFS forest (levels, parents), vertices by level,
    // children by parent, subtree sizes, preorder, low / high, union-find,
    // block ids, articulation flags, scans, counters
    DevBuf<int32_t> bi_lev, bi_par, bi_order, bi_lbeg, bi_cadj, bi_cbeg, bi_cpos, bi_sz, bi_pre, bi_low, bi_high,
        bi_uf, bi_blk, bi_tmp;
    DevBuf<uint8_t> bi_ap;
    DevBuf<int64_t> bi_scan, bi_cscan, bi_ctl;
    // heavy-row list
    DevBuf<int64_t> hv_ctl, hv_b, hv_e, hv_off;
    DevBuf<int32_t> hv_row, hv_aux;
    DevBuf<uint64_t> hv_pay;
    size_t hv_cap = 0;
};

}  // namespace gir

constexpr int HCNT_SLOTS = 512;

namespace gir {
// Grid barrier words: eight group arrival counters, the top counter, the
// generation and an error flag, each 4 KB apart (own cache line and memory
// channel).
constexpr int BAR_STRIDE = 1024;
constexpr int BAR_TOP = 8 * BAR_STRIDE, BAR_GEN = 9 * BAR_STRIDE, BAR_ERR = 10 * BAR_STRIDE;
constexpr int BAR_WORDS = BAR_ERR + 1;
constexpr unsigned BAR_SPIN = 1u << 22;  // polls (about a microsecond each) before a wait gives up
}  // namespace gir

struct girhip_graph {
    int device = 0;
    int n_cu = 0;        // compute units of the device
    int coop = -1;       // cooperative launch supported (-1: not yet queried)
    hipStream_t st = nullptr;
    int32_t V = 0;
    int64_t E = 0;
    gir::DevBuf<int32_t> fw_beg, fw_adj, bw_beg, bw_adj;
    gir::DevBuf<uint32_t> w_fw;  // weights per fw slot (valid if has_w)
    // spare adjacency / weight arrays: an update batch merges into these and
    // swaps them with the live ones
    gir::DevBuf<int32_t> fw_adj2, bw_adj2;
    gir::DevBuf<uint32_t> w_fw2;
    bool has_w = false;
    gir::DevBuf<int32_t> fw_src;  // row of every fw slot (COO), built on demand
    bool src_ok = false;
    gir::DevBuf<int32_t> sym_beg, sym_adj;  // symmetric closure (no self loops), built on demand
    bool sym_ok = false;
    int64_t sym_E = 0;
    int64_t max_deg = -1;  // largest fw/bw row length; -1 = not yet computed
    int64_t max_w = -1;    // largest stored weight; -1 = not yet computed
    // SSSP views of the forward rows for one bucket width delta (the two
    // widths used last): each row's arcs reordered so that those of weight
    // <= delta come first, mid[u] = where the heavier ones start, weights
    // (stored or hash) alongside.  delta 0 = empty; cleared when the store
    // or the weights change.
    struct SplitView {
        uint64_t delta = 0;
        uint64_t used = 0;
        gir::DevBuf<int32_t> adj, mid;
        gir::DevBuf<uint32_t> w;
    } split[2];
    uint64_t split_clock = 0;

    // layers
    bool wcc_req = false, scc_req = false;  // built at least once (repair on update)
    bool wcc_ok = false, scc_ok = false;    // current with the edge store
    bool pending = false;                   // ingested changes not yet synced
    gir::DevBuf<int32_t> wcc;
    int32_t n_wcc = 0;
    gir::DevBuf<int32_t> scc;
    int32_t n_scc = 0;
    bool n_scc_pending = false;  // n_scc still on the device (dcnt slot DC_SCC): read by sccCount
    gir::DevBuf<int32_t> dag_src, dag_dst, dag_mult;
    gir::DevBuf<int32_t> dag_beg;  // DAG CSR offsets by source SCC id (V+1)
    gir::DevBuf<int32_t> dag_in_beg, dag_in_src;  // DAG in-arcs by destination SCC id (V+1, n_dag)
    gir::DevBuf<uint64_t> dag_tk;  // in-arc keys (dst << B | src), sorted (n_dag)
    // spare arrays a merge of multiplicity changes writes into, then swaps
    gir::DevBuf<int32_t> dag_src2, dag_dst2, dag_mult2;
    gir::DevBuf<uint64_t> dag_tk2;
    gir::DevBuf<int32_t> mem_beg, mem_v;  // members by SCC id (V+1, V), built on demand
    bool mem_ok = false;
    // incremental single-source queries, [0] BFS levels, [1] SSSP distances:
    // packed (distance << 32 | parent) per vertex, and the store's changes
    // since the last update (packed (u, v) keys, as the batch path sorts them)
    struct IncTrack {
        bool on = false;
        int32_t source = -1;
        gir::DevBuf<uint64_t> st;
        gir::DevBuf<uint64_t> log;
        int64_t nlog = 0;
    } inc[2];
    int64_t n_dag = 0;
    gir::DevBuf<int64_t> rank;
    // certificate of the big SCC (label gc_label): an out-tree (gc_fpar[x] ->
    // x is an arc) and an in-tree (x -> gc_bpar[x] is an arc) of its members
    // rooted at gc_pivot, kept across update batches; built by the first
    // batch that needs it after a full SCC build
    gir::DevBuf<int32_t> gc_fpar, gc_bpar;
    bool gc_ok = false;
    int32_t gc_pivot = -1, gc_label = -1;

    // host-pinned counters for small readbacks (HCNT_SLOTS entries; the last
    // one receives the radix sort's error word)
    int64_t* hcnt = nullptr;
    bool rs_check = false;  // a radix sort ran since the last stream synchronisation
    gir::DevBuf<int64_t> dcnt;  // device counters (64 slots)
    gir::DevBuf<uint32_t> gbar;  // grid barrier of the persistent kernels (gir::BAR_WORDS)

    void* na_scratch = nullptr;  // this device's allocation-free scratch array (girhip_bfs_noalloc)
    gir::Params prm;
    std::map<std::string, double> last_ms;
    std::map<std::string, gir::KStat> kstats;
    struct PendingEv {
        std::string name;
        hipEvent_t a, b;
    };
    std::vector<PendingEv> pend_ev;
    std::vector<hipEvent_t> ev_pool;
    hipEvent_t t0 = nullptr, t1 = nullptr;
    gir::Scratch ws;
};

namespace gir {

// Every entry point: make the handle's device current.
void useDevice(girhip_graph* h);
// Wrap a body: catches Error / std::bad_alloc, sets last error, returns status.
template <class F>
int guarded(F&& f) {
    try {
        f();
        return GIRHIP_OK;
    } catch (const Error& e) {
        setLastError(e.msg);
        return e.code;
    } catch (const std::bad_alloc&) {
        setLastError("host allocation failed");
        return GIRHIP_ENOMEM;
    } catch (const std::exception& e) {
        setLastError(e.what());
        return GIRHIP_EHIP;
    }
}

hipEvent_t takeEvent(girhip_graph* h);
void flushKernelStats(girhip_graph* h);  // after a stream sync

inline unsigned gridFor(int64_t n, int block) {
    int64_t g = (n + block - 1) / block;
    if (g < 1) g = 1;
    if (g > 2147483647LL) g = 2147483647LL;
    return (unsigned)g;
}

template <class K, class... Args>
void launch(girhip_graph* h, const char* name, K kernel, dim3 grid, dim3 block, size_t shmem, Args... args) {
    hipEvent_t a = nullptr, b = nullptr;
    if (h->prm.profile != 0) {
        a = takeEvent(h);
        HIPCHK(hipEventRecord(a, h->st));
    }
    hipLaunchKernelGGL(kernel, grid, block, shmem, h->st, args...);
    HIPCHK(hipGetLastError());
    if (a) {
        b = takeEvent(h);
        HIPCHK(hipEventRecord(b, h->st));
        h->pend_ev.push_back({name, a, b});
    }
}

// Fixed launch sequences replayed as HIP graphs: the chunked rounds of
// high-diameter graphs are bound by launch overhead, which a replayed graph
// lowers.  Sequences are captured on the handle's stream the first time a
// key is seen and destroyed with the cache.  The captured work is kernels
// only (fills are kernels too; no memset or copy nodes).
struct GraphCache {
    girhip_graph* h;
    std::map<uint64_t, hipGraphExec_t> ex;
    explicit GraphCache(girhip_graph* hh) : h(hh) {}
    GraphCache(const GraphCache&) = delete;
    GraphCache& operator=(const GraphCache&) = delete;
    ~GraphCache() {
        for (auto& kv : ex) (void)hipGraphExecDestroy(kv.second);
    }
    template <class F>
    void run(uint64_t key, F issue) {
        auto it = ex.find(key);
        if (it == ex.end()) {
            hipGraph_t g = nullptr;
            HIPCHK(hipStreamBeginCapture(h->st, hipStreamCaptureModeThreadLocal));
            try {
                issue();
            } catch (...) {
                (void)hipStreamEndCapture(h->st, &g);
                if (g) (void)hipGraphDestroy(g);
                throw;
            }
            HIPCHK(hipStreamEndCapture(h->st, &g));
            hipGraphExec_t e = nullptr;
            const hipError_t r = hipGraphInstantiate(&e, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            HIPCHK(r);
            it = ex.emplace(key, e).first;
        }
        HIPCHK(hipGraphLaunch(it->second, h->st));
    }
};

// Persistent kernels: a cooperative launch of at most blocks_per_cu blocks
// per compute unit (all blocks co-resident, so grid-wide barriers are
// safe).  Returns false, launching nothing, where cooperative launch is
// unavailable or the kernel cannot keep that many blocks resident.
bool persistentGrid(girhip_graph* h, const void* kernel, int block, int blocks_per_cu, unsigned* grid);
void fillU32(girhip_graph* h, uint32_t* p, uint32_t v, int64_t n);
// The kernel takes the grid barrier's words (uint32_t*, cleared here before
// every launch) as its last argument, after args; the caller checks the
// barrier with checkGridBar after reading the kernel's results back.
template <class K, class... Args>
bool launchPersistent(girhip_graph* h, const char* name, K kernel, int block, int blocks_per_cu, Args... args) {
    unsigned grid = 0;
    if (!persistentGrid(h, (const void*)kernel, block, blocks_per_cu, &grid)) return false;
    h->gbar.ensure(BAR_WORDS);
    fillU32(h, h->gbar.p, 0, BAR_WORDS);
    uint32_t* bar = h->gbar.p;
    void* ptrs[] = {(void*)&args..., (void*)&bar};
    hipEvent_t a = nullptr;
    if (h->prm.profile != 0) {
        a = takeEvent(h);
        HIPCHK(hipEventRecord(a, h->st));
    }
    HIPCHK(hipLaunchCooperativeKernel((const void*)kernel, dim3(grid), dim3(block), ptrs, 0, h->st));
    if (a) {
        hipEvent_t b = takeEvent(h);
        HIPCHK(hipEventRecord(b, h->st));
        h->pend_ev.push_back({name, a, b});
    }
    return true;
}

// Raises GIRHIP_EHIP if a persistent kernel's barrier gave up waiting
// (call after the stream has synchronised on the kernel).
void checkGridBar(girhip_graph* h);

void syncStream(girhip_graph* h);
// Begin/end an entry point's timed region (HIP events on the handle's stream).
void tBegin(girhip_graph* h);
void tEnd(girhip_graph* h, const char* kind);  // syncs the stream

// Pointer classification: true if p is device memory usable by the kernels.
bool isDevicePtr(const void* p);
// Copy helpers (hipMemcpyDefault: host or device on either side).
void copyIn(girhip_graph* h, void* dst_dev, const void* src, size_t bytes);
void copyOut(girhip_graph* h, void* dst, const void* src_dev, size_t bytes);
// Read n int64 device counters (synchronises the stream).
void readCounters(girhip_graph* h, const int64_t* dev, int n, int64_t* out);

// ---- primitives (prims.hip) ----
void fillI32(girhip_graph* h, int32_t* p, int32_t v, int64_t n);
void fillU32(girhip_graph* h, uint32_t* p, uint32_t v, int64_t n);
void fillI64(girhip_graph* h, int64_t* p, int64_t v, int64_t n);
void fillU64(girhip_graph* h, uint64_t* p, uint64_t v, int64_t n);
void fillU8(girhip_graph* h, uint8_t* p, uint8_t v, int64_t n);
// out[0..n] = exclusive prefix sums of in[0..n) (out[n] = total).
void scanI32toI64(girhip_graph* h, const int32_t* in, int64_t* out, int64_t n);
void scanI32toI32(girhip_graph* h, const int32_t* in, int32_t* out, int64_t n);
// Stable LSD radix sort of (key, value) pairs by bits [lo_bit, key_bits) of
// key (the pairs keep their input order on ties there: input already ordered
// by the bits below lo_bit comes out ordered by all key_bits).  vals may be
// null.  Result is in keys/vals.
void radixSort(girhip_graph* h, uint64_t* keys, uint32_t* vals, int64_t n, int key_bits, int lo_bit = 0);
// dcnt slot of the SCC count an SCC pass leaves on the device.
constexpr int DC_SCC = 61;
// The SCC count (read from the device if the last SCC pass left it there).
int32_t sccCount(girhip_graph* h);
// Number of bits needed to hold values up to x.
inline int bitsFor(uint64_t x) {
    int b = 0;
    while (b < 64 && (x >> b) != 0) ++b;
    return b < 1 ? 1 : b;
}

// ---- graph store (core.hip / batch.hip) ----
// (Re)build canonical CSR on the device from unsorted rows.
void normalizeCsr(girhip_graph* h, DevBuf<int32_t>& beg, DevBuf<int32_t>& adj, int64_t& E);
void buildTranspose(girhip_graph* h);
void invalidateLayers(girhip_graph* h);
// The row of every forward slot (built once per edge-store version).
const int32_t* fwSrc(girhip_graph* h);

// ---- layers ----
void runWcc(girhip_graph* h);
void runScc(girhip_graph* h);
void runSccActive(girhip_graph* h, bool cert = false);
void runDag(girhip_graph* h);
void dagFinish(girhip_graph* h, int64_t nd, int B, bool ranks);
// Repair the SCC layer after a batch whose sorted change keys are ik / dk
// (scc.hip); false: rebuild it instead.
// dbypass (device, nd bytes, may be null): deletes known not to change
// reachability (their tail still reaches their head).
bool repairScc(girhip_graph* h, const uint64_t* ik, int64_t ni, const uint64_t* dk, int64_t nd, int B,
               const uint8_t* dbypass);
// Certificate of the big SCC (cert.hip).  A build records the trees with its
// pivot sweeps (runSccActive with cert).  certTarget picks the pivot (largest
// outdeg * indeg) and the label of its SCC and marks the trees absent;
// certRepair checks the big SCC after a batch (ncutf / ncutb: heads / tails
// of deleted tree arcs in gc_cut[0] / [1]; ncand: vertices labelled -1 that
// may join it) and rebuilds the trees when they are absent.
struct CertStats {
    int64_t dropped = 0, joined = 0, orphans = 0, swept = 0;
};
void certTarget(girhip_graph* h);
void certLabel(girhip_graph* h);  // gc_label = the label of gc_pivot
void reserveCert(girhip_graph* h);
// Work buffers of the batch path / the SCC repair for batches of up to n ops
// (batch.hip / scc.hip), reserved when a layer is first built.
void reserveBatch(girhip_graph* h, int64_t n);
void reserveRepair(girhip_graph* h, int64_t n);
CertStats certRepair(girhip_graph* h, int64_t ncutf, int64_t ncutb, int64_t ncand);
// Batches with at most BYPASS_MAX deletes test them for a bypass path.
constexpr int64_t BYPASS_MAX = 8;
bool deletesBypassed(girhip_graph* h, const uint64_t* dk, int64_t nd, int B, uint8_t* found);

inline Scratch& scratch(girhip_graph* h) { return h->ws; }
// Address of the allocation-free entry point's scratch array on device
// (resolved once per device, with that device current; noalloc.hip).
void* noAllocScratch(int device);

}  // namespace gir

// ---- device helpers ----
namespace gir {
// Grid-wide barrier of a persistent kernel (every block resident: the
// launch is cooperative).  Each wave first waits for its own stores to
// complete, so after the block barrier they are all in this XCD's L2; thread
// 0 then writes the L2 back (agent-scope release) and arrives: block b on
// group counter b % 8, the last of a group on the top counter, and the last
// of those publishes the next generation, which the others poll with plain
// device-coherent loads (no cache maintenance per poll).  A block leaves
// with an agent-scope acquire (L2 and L1 invalidated), so no stale line
// survives the barrier.  Counters only grow (cleared before each launch):
// in barrier g the last arriver is the one that brings a counter to (g + 1)
// times its arrivals.  A wait longer than BAR_SPIN polls, or one that finds
// the error flag raised, raises the flag and leaves, so a broken barrier
// ends the kernel instead of hanging the device (the host then reports it).
__device__ __forceinline__ void gridSync(uint32_t* bar) {
    __builtin_amdgcn_s_waitcnt(0xF70);  // vmcnt(0): this wave's stores have completed
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned nb = gridDim.x, grp = blockIdx.x & 7;
        const unsigned ng = nb / 8 + (grp < nb % 8 ? 1u : 0u), ntop = nb < 8 ? nb : 8;
        uint32_t* gen = &bar[BAR_GEN];
        const unsigned g = __hip_atomic_load(gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        const unsigned a =
            __hip_atomic_fetch_add(&bar[grp * BAR_STRIDE], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) + 1;
        if (a == (g + 1) * ng) {
            const unsigned t = __hip_atomic_fetch_add(&bar[BAR_TOP], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) + 1;
            if (t == (g + 1) * ntop) __hip_atomic_store(gen, g + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        for (unsigned it = 0; __hip_atomic_load(gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == g; ++it) {
            if (it >= BAR_SPIN || __hip_atomic_load(&bar[BAR_ERR], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                __hip_atomic_store(&bar[BAR_ERR], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
    }
    __syncthreads();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
}
// True once a barrier wait has given up: persistent loops test it at their
// head and leave (the results are then discarded by the host's check).
__device__ __forceinline__ bool gridFailed(const uint32_t* bar) {
    return __hip_atomic_load(&bar[BAR_ERR], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0;
}

__device__ __forceinline__ int laneId() { return (int)__lane_id(); }
__device__ __forceinline__ uint64_t lanesBelow() { return (1ULL << laneId()) - 1ULL; }

// Wave-aggregated slot reservation: every active lane with pred gets a
// distinct index in [old, old + popcount) of *counter.  Call with the whole
// active set of the wave present.
__device__ __forceinline__ int64_t wavePush(int64_t* counter, bool pred) {
    uint64_t m = __ballot(pred);
    if (m == 0) return -1;
    int leader = __ffsll((long long)m) - 1;
    int64_t base = 0;
    if (laneId() == leader) base = atomicAdd((unsigned long long*)counter, (unsigned long long)__popcll(m));
    base = __shfl(base, leader);
    return pred ? base + __popcll(m & lanesBelow()) : -1;
}

// wavePush for lanes that may push to different counters: the lanes of one
// counter at a time share one atomic.  Row functors need it when the heavy
// pass (kHeavyRows) puts arcs of rows from different sources or sweeps into
// one wave; elsewhere all lanes share a counter and the loop runs once.
__device__ __forceinline__ int64_t wavePushAny(int64_t* counter, bool pred) {
    int64_t pos = -1;
    uint64_t pending = __ballot(pred);
    while (pending) {
        const int leader = __ffsll((long long)pending) - 1;
        int64_t* lc = (int64_t*)__shfl((unsigned long long)(uintptr_t)counter, leader);
        const bool mine = pred && pos < 0 && counter == lc;
        const uint64_t m = __ballot(mine);
        int64_t base = 0;
        if (laneId() == leader) base = atomicAdd((unsigned long long*)lc, (unsigned long long)__popcll(m));
        base = __shfl(base, leader);
        if (mine) pos = base + __popcll(m & lanesBelow());
        pending &= ~m;
    }
    return pos;
}

// Relaxed, device-coherent loads/stores (bypass the non-coherent CU cache).
template <class T>
__device__ __forceinline__ T ldAcq(const T* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T>
__device__ __forceinline__ void stRel(T* p, T v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Visit every arc slot e in [b, end) of lane-owned rows across a wave: each
// lane first walks up to SMALL slots of its own row, then rows with slots
// left are finished cooperatively by the whole wave, 64 slots per step.
// Call with every lane of the wave present (act selects the lanes that own
// a row); f(v, e) may itself use wave-level ops.
template <int SMALL = 16, class F>
__device__ __forceinline__ void waveRows(bool act, int32_t v, int64_t b, int64_t end, F&& f) {
    int64_t i = b;
    if (act) {
        const int64_t stop = end < b + SMALL ? end : b + SMALL;
        for (; i < stop; ++i) f(v, i);
    }
    uint64_t heavy = __ballot(act && i < end);
    while (heavy) {
        const int l = __ffsll((long long)heavy) - 1;
        heavy &= heavy - 1;
        const int32_t hv = __shfl(v, l);
        const int64_t hb = __shfl(i, l), he = __shfl(end, l);
        for (int64_t j = hb + laneId(); j < he; j += WAVE) f(hv, j);
    }
}

// waveRows with a per-row payload handed to f(v, payload, e) (the owner
// lane's value is broadcast when the wave finishes a row cooperatively).
template <int SMALL = 16, class P, class F>
__device__ __forceinline__ void waveRowsP(bool act, int32_t v, P pay, int64_t b, int64_t end, F&& f) {
    int64_t i = b;
    if (act) {
        const int64_t stop = end < b + SMALL ? end : b + SMALL;
        for (; i < stop; ++i) f(v, pay, i);
    }
    uint64_t heavy = __ballot(act && i < end);
    while (heavy) {
        const int l = __ffsll((long long)heavy) - 1;
        heavy &= heavy - 1;
        const int32_t hv = __shfl(v, l);
        const P hp = __shfl(pay, l);
        const int64_t hb = __shfl(i, l), he = __shfl(end, l);
        for (int64_t j = hb + laneId(); j < he; j += WAVE) f(hv, hp, j);
    }
}

// ---------------------------------------------------------------------------
// Heavy rows.  RMAT hubs hold 10^5 arcs; a row walked by one wave costs
// thousands of dependent steps while the rest of the GPU idles.  A row kernel
// therefore hands every row longer than HEAVY_T arcs to a heavy list instead
// of walking it, and a follow-up launch of kHeavyRows walks the listed rows
// with the whole grid (the rows' arcs concatenated, one contiguous range per
// block).  The row work is a
// functor shared by both kernels: f(v, pay, aux, e) -> int64 contribution
// (summed per lane; the kernels reduce it per wave).  The list is reset by
// the last block of kHeavyRows, so consecutive launches need no host work.
constexpr int64_t HEAVY_T = 2048;

// ctl[0] packs the listed rows (bits 40..) and their total arc count (bits
// 0..39); off[k] = arcs of the rows listed before row k.
constexpr int HEAVY_SHIFT = 40;
constexpr uint64_t HEAVY_MASK = (1ULL << HEAVY_SHIFT) - 1;
struct HeavyList {
    int64_t* ctl;  // [0] rows << 40 | arcs, [1] blocks done
    int32_t* row;
    int32_t* aux;
    int64_t* b;
    int64_t* e;
    uint64_t* pay;
    int64_t* off;  // exclusive prefix of the listed row lengths (count + 1)
};

// The first (up to SMALL) arcs of every lane's row, spread over the wave:
// the lanes' segments are concatenated (wave prefix sum of their lengths)
// and the wave takes 64 arcs at a time, each lane finding its arc's owner by
// a binary search over the prefix (lane shuffles).  A wave of short, uneven
// rows then costs ceil(total / 64) steps of independent loads instead of
// max(length) dependent steps per lane.
template <class F>
__device__ __forceinline__ int64_t waveFlat(int32_t v, uint64_t pay, int32_t aux, int64_t b, int cnt, const F& f) {
    const int lane = laneId();
    int incl = cnt;
    for (int d = 1; d < WAVE; d <<= 1) {
        const int o = __shfl_up(incl, d);
        if (lane >= d) incl += o;
    }
    const int total = __shfl(incl, WAVE - 1);
    int64_t acc = 0;
    for (int t0 = 0; t0 < total; t0 += WAVE) {
        const int t = t0 + lane;
        // owner: the first lane whose inclusive end exceeds t (a fixed six
        // steps, so every lane takes part in every shuffle)
        int below = 0;  // lanes whose segment ends at or before t
        for (int step = WAVE / 2; step >= 1; step >>= 1)
            if (__shfl(incl, below + step - 1) <= t) below += step;
        const int own = below < WAVE ? below : WAVE - 1;
        const int excl = __shfl(incl, own) - __shfl(cnt, own);
        const int32_t ov = __shfl(v, own), oa = __shfl(aux, own);
        const uint64_t op = __shfl(pay, own);
        const int64_t ob = __shfl(b, own);
        if (t < total) acc += f(ov, op, oa, ob + (t - excl));
    }
    return acc;
}

// Rows of the lanes with heavy set go to the heavy list (row, aux, range,
// payload and the arc offset of each row): list slots and arc offsets are
// reserved together, (rows << 40) + arcs in one atomic, so offsets grow
// with slots and the heavy pass needs no prefix launch.  Whole wave.
__device__ __forceinline__ void listHeavyRows(bool heavy, int32_t v, uint64_t pay, int32_t aux, int64_t b,
                                              int64_t end, const HeavyList& H) {
    const uint64_t hm = __ballot(heavy);
    if (!hm) return;
    const int lane = laneId();
    const int64_t len = heavy ? end - b : 0;
    int64_t incl = len;
    for (int d = 1; d < WAVE; d <<= 1) {
        const int64_t o = __shfl_up(incl, d);
        if (lane >= d) incl += o;
    }
    const int64_t tot = __shfl(incl, WAVE - 1);
    const int leader = __ffsll((long long)hm) - 1;
    unsigned long long old = 0;
    if (lane == leader)
        old = atomicAdd((unsigned long long*)&H.ctl[0],
                        ((unsigned long long)__popcll(hm) << HEAVY_SHIFT) + (unsigned long long)tot);
    old = __shfl(old, leader);
    if (heavy) {
        const int64_t pos = (int64_t)(old >> HEAVY_SHIFT) + __popcll(hm & lanesBelow());
        H.row[pos] = v;
        H.aux[pos] = aux;
        H.b[pos] = b;
        H.e[pos] = end;
        H.pay[pos] = pay;
        H.off[pos] = (int64_t)(old & HEAVY_MASK) + incl - len;
    }
}

template <int SMALL = 16, class F>
__device__ __forceinline__ int64_t waveRowsH(bool act, int32_t v, uint64_t pay, int32_t aux, int64_t b, int64_t end,
                                             const HeavyList& H, const F& f) {
    const bool heavy = act && (end - b) > HEAVY_T;
    listHeavyRows(heavy, v, pay, aux, b, end, H);
    const bool mine = act && !heavy;
    const int first = mine ? (int)(end - b < SMALL ? end - b : SMALL) : 0;
    int64_t acc = waveFlat(v, pay, aux, b, first, f);
    const int64_t i = b + first;
    uint64_t rest = __ballot(mine && i < end);
    while (rest) {
        const int l = __ffsll((long long)rest) - 1;
        rest &= rest - 1;
        const int32_t hv = __shfl(v, l), ha = __shfl(aux, l);
        const uint64_t hp = __shfl(pay, l);
        const int64_t hb = __shfl(i, l), he = __shfl(end, l);
        for (int64_t j = hb + laneId(); j < he; j += WAVE) acc += f(hv, hp, ha, j);
    }
    return acc;
}

// waveRowsH with no heavy list: a long row is walked by its whole wave, 64
// arcs at a time.  For passes whose arcs must be evaluated twice by the same
// lanes in the same order (count, then write at the counted offsets).
template <int SMALL = 16, class F>
__device__ __forceinline__ int64_t waveRowsAll(bool act, int32_t v, uint64_t pay, int32_t aux, int64_t b,
                                               int64_t end, const F& f) {
    const int first = act ? (int)(end - b < SMALL ? end - b : SMALL) : 0;
    int64_t acc = waveFlat(v, pay, aux, b, first, f);
    const int64_t i = b + first;
    uint64_t rest = __ballot(act && i < end);
    while (rest) {
        const int l = __ffsll((long long)rest) - 1;
        rest &= rest - 1;
        const int32_t hv = __shfl(v, l), ha = __shfl(aux, l);
        const uint64_t hp = __shfl(pay, l);
        const int64_t hb = __shfl(i, l), he = __shfl(end, l);
        for (int64_t j = hb + laneId(); j < he; j += WAVE) acc += f(hv, hp, ha, j);
    }
    return acc;
}

// waveRowsH for functors with a batched form that hides memory latency: U
// arcs per lane at a time, f.many<U>(p, pay, aux, j, ok) with one entry per
// arc (ok = entry holds an arc).  The functor issues the arcs' loads for all
// U entries before it uses any of them, so a lane keeps U dependent chains
// (arc -> head's state -> update) in flight instead of one.  No return value.
template <int U, class F>
__device__ __forceinline__ void waveFlatU(int32_t v, uint64_t pay, int32_t aux, int64_t b, int cnt, const F& f) {
    const int lane = laneId();
    int incl = cnt;
    for (int d = 1; d < WAVE; d <<= 1) {
        const int o = __shfl_up(incl, d);
        if (lane >= d) incl += o;
    }
    const int total = __shfl(incl, WAVE - 1);
    for (int t0 = 0; t0 < total; t0 += U * WAVE) {
        int32_t ov[U], oa[U];
        uint64_t op[U];
        int64_t oj[U];
        bool ok[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int t = t0 + u * WAVE + lane;
            int below = 0;
            for (int step = WAVE / 2; step >= 1; step >>= 1)
                if (__shfl(incl, below + step - 1) <= t) below += step;
            const int own = below < WAVE ? below : WAVE - 1;
            const int excl = __shfl(incl, own) - __shfl(cnt, own);
            ov[u] = __shfl(v, own);
            oa[u] = __shfl(aux, own);
            op[u] = __shfl(pay, own);
            oj[u] = __shfl(b, own) + (t - excl);
            ok[u] = t < total;
        }
        f.template many<U>(ov, op, oa, oj, ok);
    }
}

template <int U, int SMALL = 16, class F>
__device__ __forceinline__ void waveRowsHU(bool act, int32_t v, uint64_t pay, int32_t aux, int64_t b, int64_t end,
                                           const HeavyList& H, const F& f) {
    const bool heavy = act && (end - b) > HEAVY_T;
    listHeavyRows(heavy, v, pay, aux, b, end, H);
    const bool mine = act && !heavy;
    const int first = mine ? (int)(end - b < SMALL ? end - b : SMALL) : 0;
    waveFlatU<U>(v, pay, aux, b, first, f);
    const int64_t i = b + first;
    uint64_t rest = __ballot(mine && i < end);
    while (rest) {
        const int l = __ffsll((long long)rest) - 1;
        rest &= rest - 1;
        int32_t hv[U], ha[U];
        uint64_t hp[U];
        int64_t hj[U];
        bool ok[U];
        const int32_t v0 = __shfl(v, l), a0 = __shfl(aux, l);
        const uint64_t p0 = __shfl(pay, l);
        const int64_t hb = __shfl(i, l), he = __shfl(end, l);
        for (int64_t j0 = hb; j0 < he; j0 += U * WAVE) {
#pragma unroll
            for (int u = 0; u < U; ++u) {
                hv[u] = v0;
                ha[u] = a0;
                hp[u] = p0;
                hj[u] = j0 + u * WAVE + laneId();
                ok[u] = hj[u] < he;
            }
            f.template many<U>(hv, hp, ha, hj, ok);
        }
    }
}

__device__ __forceinline__ int64_t waveSum(int64_t x) {
    for (int d = WAVE / 2; d > 0; d >>= 1) x += __shfl_xor(x, d);
    return x;
}

// Block-wide sum added to *target with one atomic per block.  Every thread
// of the block must call it (uniform control flow).
__device__ __forceinline__ void blockAtomicAdd(int64_t x, int64_t* target) {
    __shared__ int64_t part[1024 / WAVE];
    x = waveSum(x);
    const int w = threadIdx.x / WAVE, nw = (blockDim.x + WAVE - 1) / WAVE;
    if (laneId() == 0) part[w] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        int64_t t = 0;
        for (int i = 0; i < nw; ++i) t += part[i];
        if (t) atomicAdd((unsigned long long*)target, (unsigned long long)t);
    }
    __syncthreads();
}

// Block-wide exclusive scan of n per thread; returns this thread's offset in
// the output and reserves the block's range with one atomicAdd on *cnt (the
// LDS-staged alternative to one atomic per wave).  Every thread of the block
// must call it.
__device__ __forceinline__ int64_t blockReserve(int n, int64_t* cnt) {
    __shared__ int64_t wtot[1024 / WAVE];
    __shared__ int64_t base;
    const int lane = laneId(), w = threadIdx.x / WAVE, nw = (blockDim.x + WAVE - 1) / WAVE;
    int64_t v = n;
    for (int d = 1; d < WAVE; d <<= 1) {
        const int64_t o = __shfl_up(v, d);
        if (lane >= d) v += o;
    }
    if (lane == WAVE - 1) wtot[w] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        int64_t t = 0;
        for (int i = 0; i < nw; ++i) {
            const int64_t x = wtot[i];
            wtot[i] = t;
            t += x;
        }
        base = t ? (int64_t)atomicAdd((unsigned long long*)cnt, (unsigned long long)t) : 0;
    }
    __syncthreads();
    const int64_t off = base + wtot[w] + v - n;
    __syncthreads();
    return off;
}

// blockReserve of one slot per thread with pred, together with the block's
// sum of m: one atomicAdd of (slots << 32) + sum(m) on *packed, so a launch
// over millions of threads hits its counter once per block (same-address
// atomics serialise at the memory side).  Both halves must stay below 2^32
// over the counter's life.  Returns this thread's slot (valid when pred).
// Every thread of the block must call it.
__device__ __forceinline__ int64_t blockReservePacked(bool pred, int64_t m, int64_t* packed) {
    __shared__ int64_t wcnt[1024 / WAVE], wsum[1024 / WAVE];
    __shared__ int64_t base;
    const int lane = laneId(), w = threadIdx.x / WAVE, nw = (blockDim.x + WAVE - 1) / WAVE;
    const uint64_t bal = __ballot(pred);
    int64_t ms = m;
    for (int d = WAVE / 2; d > 0; d >>= 1) ms += __shfl_xor(ms, d);
    if (lane == 0) {
        wcnt[w] = __popcll(bal);
        wsum[w] = ms;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int64_t t = 0, sm = 0;
        for (int i = 0; i < nw; ++i) {
            const int64_t x = wcnt[i];
            wcnt[i] = t;
            t += x;
            sm += wsum[i];
        }
        base = (t || sm) ? (int64_t)(atomicAdd((unsigned long long*)packed,
                                               ((unsigned long long)t << 32) + (unsigned long long)sm) >>
                                     32)
                         : 0;
    }
    __syncthreads();
    const int64_t off = base + wcnt[w] + __popcll(bal & lanesBelow());
    __syncthreads();
    return off;
}

// Block-wide maximum (min: pass the bitwise complement) of every thread's
// x, returned to every thread.  Every thread of the block must call it.
__device__ __forceinline__ uint64_t blockMaxU64(uint64_t x) {
    __shared__ uint64_t part[1024 / WAVE];
    for (int d = WAVE / 2; d > 0; d >>= 1) {
        const uint64_t o = __shfl_xor((unsigned long long)x, d);
        x = o > x ? o : x;
    }
    const int w = threadIdx.x / WAVE, nw = (blockDim.x + WAVE - 1) / WAVE;
    if (laneId() == 0) part[w] = x;
    __syncthreads();
    for (int i = 0; i < nw; ++i) x = part[i] > x ? part[i] : x;
    __syncthreads();
    return x;
}

// Wave-wide exclusive scan of n per lane; reserves the wave's range with one
// atomicAdd on *cnt.  Every lane of the wave must call it.
__device__ __forceinline__ int64_t waveReserve(int n, int64_t* cnt) {
    const int lane = laneId();
    int64_t v = n;
    for (int d = 1; d < WAVE; d <<= 1) {
        const int64_t o = __shfl_up(v, d);
        if (lane >= d) v += o;
    }
    int64_t base = 0;
    const int64_t tot = __shfl(v, WAVE - 1);
    if (lane == 0 && tot) base = (int64_t)atomicAdd((unsigned long long*)cnt, (unsigned long long)tot);
    base = __shfl(base, 0);
    return base + v - n;
}

// Sets bit x of a bitmap; the first bit set in a word also sets the word's
// bit in the summary (bit g of summary word i <=> bitmap word 64i+g may be
// non-zero), so a non-zero word always has its summary bit.  Returns true
// for the call that made the word non-zero.
__device__ __forceinline__ bool setBit(uint64_t* bits, uint64_t* sum, int32_t x) {
    const uint64_t bit = 1ULL << (x & 63);
    uint64_t* wp = &bits[x >> 6];
    if (*wp & bit) return false;  // (bits are only set while a kernel runs: a stale read is a safe pre-check)
    if (atomicOr((unsigned long long*)wp, (unsigned long long)bit) != 0) return false;
    const uint64_t sb = 1ULL << ((x >> 6) & 63);
    uint64_t* sp = &sum[x >> 12];
    if (!(*sp & sb)) atomicOr((unsigned long long*)sp, (unsigned long long)sb);
    return true;
}

// Sets bit x of a bitmap with a summary, as setBit, without awaiting any
// atomic: the word's atomicOr is unconditional, the summary's is behind a
// plain pre-check of the summary word sw that the caller loaded earlier
// (sum[x >> 12], read before its other updates so the load is in flight
// with them).  Returns true when sw showed the summary bit clear; within a
// kernel bits are only set, so the first setter of a summary bit always
// sees it clear and "some call returned true" <=> "the bitmap went from
// empty to non-empty" for a bitmap that was empty at launch.
__device__ __forceinline__ bool markBit(uint64_t* bits, uint64_t* sum, int32_t x, uint64_t sw) {
    atomicOr((unsigned long long*)&bits[x >> 6], 1ULL << (x & 63));
    const uint64_t sb = 1ULL << ((x >> 6) & 63);
    if (sw & sb) return false;
    atomicOr((unsigned long long*)&sum[x >> 12], (unsigned long long)sb);
    return true;
}

// markBit with the half of the summary word that holds x's bit
// (((const uint32_t*)sum)[x >> 11], little-endian), for callers short of
// registers.
__device__ __forceinline__ bool markBitHalf(uint64_t* bits, uint64_t* sum, int32_t x, uint32_t hw) {
    atomicOr((unsigned long long*)&bits[x >> 6], 1ULL << (x & 63));
    if ((hw >> ((x >> 6) & 31)) & 1) return false;
    atomicOr((unsigned long long*)&sum[x >> 12], 1ULL << ((x >> 6) & 63));
    return true;
}

// markBit for one vertex per lane (lanes with act), for vertex lists in
// which equal bitmap words sit in consecutive lanes (a queue compacted in
// vertex order): the lanes of a run with the same word OR their bits
// together (segmented scan over the wave) and the run's first lane issues
// the run's one atomicOr and the summary update.  Whole wave.
__device__ __forceinline__ void waveMarkBits(uint64_t* bits, uint64_t* sum, int32_t x, bool act) {
    const int lane = laneId();
    const long long word = act ? (long long)(x >> 6) : -1LL;
    uint64_t acc = act ? 1ULL << (x & 63) : 0ULL;
    for (int d = 1; d < WAVE; d <<= 1) {
        const uint64_t o = __shfl_down(acc, d);
        const long long ow = __shfl_down(word, d);
        if (lane + d < WAVE && ow == word) acc |= o;
    }
    const long long pw = __shfl_up(word, 1);
    if (act && (lane == 0 || pw != word)) {
        atomicOr((unsigned long long*)&bits[word], (unsigned long long)acc);
        const uint64_t sb = 1ULL << (word & 63);
        uint64_t* sp = &sum[word >> 6];
        if (!(*sp & sb)) atomicOr((unsigned long long*)sp, (unsigned long long)sb);
    }
}

// Wave-level walk over a summary bitmap: the waves of the grid take tasks of
// G consecutive summary words (1 <= G <= 64, one coalesced load per task) and
// call body(g, sm) for every non-zero word g, sm = its value (wave-uniform
// arguments: the whole wave runs the body).  Small G spreads dense bitmaps
// over many waves; large G lets few waves sweep a sparse one.
// (b, nb: this block's index and the block count of the walk, for kernels
// that split one grid among several rows of work)
template <class F>
__device__ __forceinline__ void forSummaryB(const uint64_t* __restrict__ sum, int64_t W2, int G, int64_t b, int64_t nb,
                                            F&& body) {
    const int64_t wave = (b * blockDim.x + threadIdx.x) / WAVE;
    const int64_t nw = (nb * blockDim.x) / WAVE;
    const int l = laneId();
    for (int64_t t0 = wave * G; t0 < W2; t0 += nw * G) {
        const int64_t gi = t0 + l;
        const uint64_t mine = (l < G && gi < W2) ? sum[gi] : 0;
        uint64_t act = __ballot(mine != 0);
        while (act) {
            const int j = __ffsll((long long)act) - 1;
            act &= act - 1;
            body(t0 + j, (uint64_t)__shfl((unsigned long long)mine, j));
        }
    }
}
template <class F>
__device__ __forceinline__ void forSummary(const uint64_t* __restrict__ sum, int64_t W2, int G, F&& body) {
    forSummaryB(sum, W2, G, (int64_t)blockIdx.x, (int64_t)gridDim.x, body);
}

// Wave-level synchronisation of LDS traffic between the lanes of one wave.
__device__ __forceinline__ void waveSync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// LDS entries per wave for waveBits.
constexpr int WB_CAP = 256;

// Lane-parallel expansion of 64 bitmap words held one per lane (m = this
// lane's word, word index wi): the set bits' ids (wi * 64 + bit) are
// compacted into the wave's LDS slice (lds, WB_CAP entries) by a wave
// prefix sum, then body(act, x) runs for 64 ids at a time with the whole
// wave (act false on idle lanes), so the memory accesses of up to 64 ids are
// in flight together instead of one chain per lane.  More than WB_CAP set
// bits: each lane walks its own bits (same calls, less parallel).
template <class F>
__device__ __forceinline__ void waveBits(uint64_t m, int64_t wi, int32_t* lds, F&& body) {
    const int lane = laneId();
    const int c = __popcll(m);
    int incl = c;
    for (int d = 1; d < WAVE; d <<= 1) {
        const int o = __shfl_up(incl, d);
        if (lane >= d) incl += o;
    }
    const int total = __shfl(incl, WAVE - 1);
    if (total == 0) return;
    if (total <= WB_CAP) {
        int pos = incl - c;
        while (m) {
            const int b = __ffsll((long long)m) - 1;
            m &= m - 1;
            lds[pos++] = (int32_t)(wi * 64 + b);
        }
        waveSync();
        for (int i0 = 0; i0 < total; i0 += WAVE) {
            const int i = i0 + lane;
            const bool act = i < total;
            body(act, act ? lds[i] : 0);
        }
        waveSync();
        return;
    }
    while (__ballot(m != 0)) {
        const bool act = m != 0;
        int32_t x = 0;
        if (act) {
            x = (int32_t)(wi * 64 + __ffsll((long long)m) - 1);
            m &= m - 1;
        }
        body(act, x);
    }
}

// Grid for a forSummary kernel over tasks summary-word tasks (about two
// tasks per wave, at most 1024 blocks).
inline unsigned summaryGrid(int64_t W2, int G, int64_t reps, int block) {
    const int64_t tasks = ((W2 + G - 1) / G) * reps;
    const int64_t waves = (tasks + 1) / 2;
    const int64_t b = (waves * WAVE + block - 1) / block;
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>(1024, b));
}

// One block: off[k] = sum of the lengths of listed rows before k.

// The listed rows' arcs as one range [0, off[n]): every thread strides over
// it and finds the owning row by binary search in off.
// Functors whose arcs append to a queue can instead expose heavyClaim (the
// arc's work, returning whether it appends an item), pushCounter and
// heavyPlace: kHeavyRows then walks its range in block-wide steps and
// reserves each step's items with one atomic per block (a hub's row of
// 10^5 arcs would otherwise hit the queue counter once per wave).
template <class F, class = void>
struct HasHeavyClaim : std::false_type {};
template <class F>
struct HasHeavyClaim<F, std::void_t<decltype(&F::heavyClaim)>> : std::true_type {};

template <class F>
__global__ void kHeavyRows(HeavyList H, F f, int64_t* sum) {
    const uint64_t c = (uint64_t)ldAcq(&H.ctl[0]);
    if (c == 0) return;  // empty list (the usual case): nothing to walk or reset
    const int64_t n = (int64_t)(c >> HEAVY_SHIFT), T = (int64_t)(c & HEAVY_MASK);
    // each block takes a contiguous range of the concatenated arcs, so a
    // thread's next arc is usually in the same row and the row search runs
    // only at row boundaries
    const int64_t per = (T + gridDim.x - 1) / gridDim.x;
    const int64_t p0 = (int64_t)blockIdx.x * per, p1 = p0 + per < T ? p0 + per : T;
    int64_t acc = 0;
    int64_t k = 0;
    auto rowOfArc = [&](int64_t p) {
        if (!(p >= H.off[k] && p < H.off[k] + (H.e[k] - H.b[k]))) {
            int64_t lo = 0, hi = n - 1;  // last k with off[k] <= p
            while (lo < hi) {
                const int64_t m = (lo + hi + 1) >> 1;
                if (H.off[m] <= p) lo = m;
                else hi = m - 1;
            }
            k = lo;
        }
    };
    if constexpr (HasHeavyClaim<F>::value) {
        for (int64_t pb = p0; pb < p1; pb += blockDim.x) {  // (uniform per block)
            const int64_t p = pb + threadIdx.x;
            bool push = false;
            int32_t item = 0;
            if (p < p1) {
                rowOfArc(p);
                acc += f.heavyClaim(H.row[k], H.pay[k], H.aux[k], H.b[k] + (p - H.off[k]), push, item);
            }
            const int64_t pos = blockReserve(push ? 1 : 0, f.pushCounter());
            if (push) f.heavyPlace(pos, item);
        }
    } else {
        for (int64_t p = p0 + threadIdx.x; p < p1; p += blockDim.x) {
            rowOfArc(p);
            acc += f(H.row[k], H.pay[k], H.aux[k], H.b[k] + (p - H.off[k]));
        }
    }
    if (sum) {
        acc = waveSum(acc);
        if (laneId() == 0 && acc) atomicAdd((unsigned long long*)sum, (unsigned long long)acc);
    }
    // the last block to finish clears the list for the next row kernel: every
    // block has read ctl[0] by then (no fence needed: the list's readers are
    // this launch's blocks, and the next writer starts after the launch)
    __syncthreads();
    if (threadIdx.x == 0 && atomicAdd((unsigned long long*)&H.ctl[1], 1ULL) == gridDim.x - 1) {
        H.ctl[0] = 0;
        H.ctl[1] = 0;
    }
}

// Small utility kernels shared by several translation units (one private
// copy per unit).
// cnt += number of v with lab[v] == v (component representatives).
// (grid-stride; launch with a bounded grid)
static __global__ void kCountSelf(const int32_t* __restrict__ lab, int32_t V, int64_t* cnt) {
    int64_t c = 0;
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < V; v += (int64_t)gridDim.x * blockDim.x)
        c += lab[v] == (int32_t)v;
    blockAtomicAdd(c, cnt);
}
inline unsigned boundedGrid(int64_t n, int block) {
    const unsigned g = gridFor(n, block);
    return g < 2048 ? g : 2048;
}
// Grid of a reduction whose blocks each end in one device-scope atomic on a
// shared counter: such atomics serialise at the memory side (a few ns each),
// so 512 blocks (two per CU) instead of 2,048 take a quarter of that tail
// and still stream the input at full width.
inline unsigned reduceGrid(int64_t n, int block) {
    const unsigned g = gridFor(n, block);
    return g < 512 ? g : 512;
}
static __global__ void kI64toI32(const int64_t* __restrict__ in, int32_t* __restrict__ out, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (int32_t)in[i];
}

// The reference's edge weight (gir_query.cpp:50-54): 1 + fmix64(key) mod 100.
__device__ __forceinline__ uint32_t hashWeight(int32_t u, int32_t v) {
    uint64_t x = ((uint64_t)(uint32_t)u << 32) | (uint32_t)v;
    x ^= x >> 33;
    x *= 0xff51afd7ed558ccdULL;
    x ^= x >> 33;
    x *= 0xc4ceb9fe1a85ec53ULL;
    x ^= x >> 33;
    return 1u + (uint32_t)(x % 100ULL);
}
}  // namespace gir

namespace gir {
// The handle's heavy list (capacity follows E; zeroed on allocation).
HeavyList heavyList(girhip_graph* h);
// Launch kHeavyRows for functor f after a row kernel (sum: optional int64
// accumulator for the functor's contributions).
template <class F>
void launchHeavy(girhip_graph* h, const char* name, const F& f, int64_t* sum = nullptr) {
    const HeavyList H = heavyList(h);
    launch(h, name, kHeavyRows<F>, 1024, 256, 0, H, f, sum);
}
// Symmetric closure of the edge store (both directions, deduplicated, no
// self loops) as CSR in h->sym_beg / sym_adj (cached until the store
// changes); the undirected graph of the reference's buildWCC
// (gir.cpp:593-620).  Returns its arc count.
int64_t symCsr(girhip_graph* h);
// Incremental queries: record n changed arcs (packed (u << B) | v keys,
// B = bitsFor(V - 1)) for every tracked kind.
void incLog(girhip_graph* h, const uint64_t* keys, int64_t n);
// Largest row of the edge store's fw and bw CSR (cached until the store
// changes).
int64_t maxDegree(girhip_graph* h);
// Largest arc weight SSSP uses: the stored weights' maximum (cached until
// they change), or the hash weights' bound 100.
int64_t maxWeight(girhip_graph* h);
// launchHeavy for row kernels over the edge store's CSR or its symmetric
// closure: skipped when no such row can reach the heavy list (a closure row
// holds at most a vertex's out- plus in-arcs).
template <class F>
void launchHeavyStore(girhip_graph* h, const char* name, const F& f, int64_t* sum = nullptr) {
    if (2 * maxDegree(h) > HEAVY_T) launchHeavy(h, name, f, sum);
}
// launchHeavyStore over a given heavy list.
template <class F>
void launchHeavyStoreH(girhip_graph* h, const HeavyList& H, const char* name, const F& f, int64_t* sum = nullptr) {
    if (2 * maxDegree(h) > HEAVY_T) launch(h, name, kHeavyRows<F>, 1024, 256, 0, H, f, sum);
}
}  // namespace gir

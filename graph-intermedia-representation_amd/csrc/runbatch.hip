// girhip_run_batch: a mixed list of BFS / SSSP queries and edge updates in
// one call, answered in list order.  The reference answers such a workload
// one GIR::bfs / GIR::sssp / GIR::updateEdgeDirected call at a time
// (bench_main.cpp:538-542, gir.cpp:3043-3133).
//
// The list is cut into runs: a maximal run of updates is one normalised
// batch (last op per arc wins, which is what applying the run one op at a
// time leaves; gir.cpp:3523-3586) repaired by the batch path; a maximal run
// of queries shares work: its BFS queries run in bit-parallel sweeps of 64
// sources, its SSSP queries in multi-source batches.  Sources that reach
// nothing but themselves (no out-arc, or an id outside [0, V)) take no sweep
// slot: their rows are written by one fill kernel, and the live sources'
// rows are computed together and scattered to their places in the output.
#include <chrono>
#include <vector>

#include "common.h"

namespace gir {
namespace {

// out-degree of every source (-1 for an id outside [0, V))
__global__ void kSrcDeg(const int32_t* __restrict__ src, int n, const int32_t* __restrict__ fb, int32_t V,
                        int32_t* __restrict__ deg) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t x = src[i];
    deg[i] = (x >= 0 && x < V) ? fb[x + 1] - fb[x] : -1;
}

// Rows of sources that reach only themselves: row rows[r] of out = none
// everywhere, 0 at the source (when it is a vertex).
template <class T>
__global__ void kTrivialRows(T* __restrict__ out, int64_t V, const int32_t* __restrict__ rows,
                             const int32_t* __restrict__ src, int n, T none) {
    const int64_t total = (int64_t)n * V;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
        const int r = (int)(i / V);
        const int64_t v = i - (int64_t)r * V;
        out[(int64_t)rows[r] * V + v] = v == src[r] ? (T)0 : none;
    }
}

// row r of stage -> row rows[r] of out
template <class T>
__global__ void kScatterRows(const T* __restrict__ stage, T* __restrict__ out, int64_t V,
                             const int32_t* __restrict__ rows, int n) {
    const int64_t total = (int64_t)n * V;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
        const int r = (int)(i / V);
        const int64_t v = i - (int64_t)r * V;
        out[(int64_t)rows[r] * V + v] = stage[i];
    }
}

// The queries of one run of one kind (BFS: T = int32, SSSP: T = uint64),
// sources src, into rows [0, k) of out (device).
template <class T>
void runQueries(girhip_graph* h, bool sssp, const std::vector<int32_t>& src, uint64_t delta, T* out) {
    const int k = (int)src.size();
    if (k == 0) return;
    const int32_t V = h->V;
    Scratch& R = scratch(h);
    R.src.ensure((size_t)k);
    R.rb_deg.ensure((size_t)k);
    copyIn(h, R.src.p, src.data(), sizeof(int32_t) * (size_t)k);
    launch(h, "run_src_deg", kSrcDeg, gridFor(k, 256), 256, 0, (const int32_t*)R.src.p, k,
           (const int32_t*)h->fw_beg.p, V, R.rb_deg.p);
    std::vector<int32_t> deg((size_t)k);
    HIPCHK(hipMemcpyAsync(deg.data(), R.rb_deg.p, sizeof(int32_t) * (size_t)k, hipMemcpyDeviceToHost, h->st));
    syncStream(h);
    std::vector<int32_t> live_src, live_row, t_src, t_row;
    for (int i = 0; i < k; ++i) {
        if (deg[i] > 0) {
            live_src.push_back(src[i]);
            live_row.push_back(i);
        } else {
            t_src.push_back(src[i]);
            t_row.push_back(i);
        }
    }
    const int nl = (int)live_src.size(), nt = (int)t_src.size();
    if (nl > 0) {
        // every source live: straight into the output rows
        T* dst = out;
        if (nt > 0) {
            R.rb_live.ensure(sizeof(T) * (size_t)nl * (size_t)V);
            dst = (T*)R.rb_live.p;
        }
        const int rc = sssp ? girhip_sssp(h, live_src.data(), nl, delta, (uint64_t*)dst, nullptr)
                            : girhip_bfs(h, live_src.data(), nl, (int32_t*)dst, nullptr);
        if (rc != GIRHIP_OK) raise(rc, girhip_last_error());
        if (nt > 0) {
            R.rb_rows.ensure((size_t)nl);
            copyIn(h, R.rb_rows.p, live_row.data(), sizeof(int32_t) * (size_t)nl);
            launch(h, "run_scatter_rows", kScatterRows<T>, boundedGrid((int64_t)nl * V, 256), 256, 0, (const T*)dst,
                   out, (int64_t)V, (const int32_t*)R.rb_rows.p, nl);
        }
    }
    if (nt > 0) {
        R.rb_tsrc.ensure((size_t)nt);
        R.rb_rows.ensure((size_t)std::max(nl, nt));
        syncStream(h);  // rows of the scatter above are reused below
        copyIn(h, R.rb_tsrc.p, t_src.data(), sizeof(int32_t) * (size_t)nt);
        copyIn(h, R.rb_rows.p, t_row.data(), sizeof(int32_t) * (size_t)nt);
        const T none = sssp ? (T)~0ULL : (T)-1;
        launch(h, "run_trivial_rows", kTrivialRows<T>, boundedGrid((int64_t)nt * V, 256), 256, 0, out, (int64_t)V,
               (const int32_t*)R.rb_rows.p, (const int32_t*)R.rb_tsrc.p, nt, none);
    }
    syncStream(h);
}

double msSince(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace
}  // namespace gir

using namespace gir;

extern "C" int girhip_run_batch(girhip_graph* h, const int32_t* kind, const int32_t* a, const int32_t* b, int64_t n,
                                uint64_t delta, int32_t* levels, uint64_t* dist, girhip_batch_stats* stats) {
    return guarded([&] {
        if (!h) raise(GIRHIP_EINVAL, "run_batch: null handle");
        useDevice(h);
        if (n < 0 || (n > 0 && (!kind || !a))) raise(GIRHIP_EINVAL, "run_batch: bad argument");
        std::vector<int32_t> kd((size_t)n), A((size_t)n), Bv((size_t)n);
        if (n > 0) {
            HIPCHK(hipMemcpy(kd.data(), kind, sizeof(int32_t) * (size_t)n, hipMemcpyDefault));
            HIPCHK(hipMemcpy(A.data(), a, sizeof(int32_t) * (size_t)n, hipMemcpyDefault));
        }
        int64_t nb = 0, ns = 0;
        bool upd = false;
        for (int64_t i = 0; i < n; ++i) {
            if (kd[i] < 0 || kd[i] > 3) raise(GIRHIP_EINVAL, "run_batch: kind must be 0 (BFS), 1 (SSSP), 2 (insert) or 3 (delete)");
            nb += kd[i] == 0;
            ns += kd[i] == 1;
            upd = upd || kd[i] >= 2;
        }
        if (upd) {
            if (!b) raise(GIRHIP_EINVAL, "run_batch: updates need the head array b");
            HIPCHK(hipMemcpy(Bv.data(), b, sizeof(int32_t) * (size_t)n, hipMemcpyDefault));
        }
        const int32_t V = h->V;
        const bool lv_dev = levels && isDevicePtr(levels), d_dev = dist && isDevicePtr(dist);
        girhip_batch_stats tot{0, 0, 0, 0};
        double q_ms = 0, u_ms = 0;
        int64_t rb = 0, rs = 0;  // next output row of each kind
        Scratch& R = scratch(h);
        for (int64_t i = 0; i < n;) {
            int64_t j = i;
            const auto t0 = std::chrono::steady_clock::now();
            if (kd[i] >= 2) {
                while (j < n && kd[j] >= 2) ++j;
                const int64_t m = j - i;
                std::vector<uint8_t> ins((size_t)m);
                for (int64_t t = 0; t < m; ++t) ins[t] = kd[i + t] == 2 ? 1 : 0;
                girhip_batch_stats st;
                const int rc = girhip_apply_batch(h, A.data() + i, Bv.data() + i, ins.data(), nullptr, m, &st);
                if (rc != GIRHIP_OK) raise(rc, girhip_last_error());
                tot.inserted += st.inserted;
                tot.deleted += st.deleted;
                tot.noop += st.noop;
                tot.invalid += st.invalid;
                u_ms += msSince(t0);
            } else {
                while (j < n && kd[j] < 2) ++j;
                std::vector<int32_t> sb, ss;
                for (int64_t t = i; t < j; ++t) (kd[t] == 0 ? sb : ss).push_back(A[t]);
                // BFS rows [rb, rb + |sb|), SSSP rows [rs, rs + |ss|) of the outputs
                if (!sb.empty()) {
                    int32_t* o = lv_dev ? levels + (size_t)rb * V : nullptr;
                    if (!o) {
                        R.rb_out.ensure(sizeof(int32_t) * sb.size() * (size_t)V);
                        o = (int32_t*)R.rb_out.p;
                    }
                    runQueries<int32_t>(h, false, sb, delta, o);
                    if (levels && !lv_dev)
                        copyOut(h, levels + (size_t)rb * V, o, sizeof(int32_t) * sb.size() * (size_t)V);
                    rb += (int64_t)sb.size();
                }
                if (!ss.empty()) {
                    uint64_t* o = d_dev ? dist + (size_t)rs * V : nullptr;
                    if (!o) {
                        R.rb_out.ensure(sizeof(uint64_t) * ss.size() * (size_t)V);
                        o = (uint64_t*)R.rb_out.p;
                    }
                    runQueries<uint64_t>(h, true, ss, delta, o);
                    if (dist && !d_dev) copyOut(h, dist + (size_t)rs * V, o, sizeof(uint64_t) * ss.size() * (size_t)V);
                    rs += (int64_t)ss.size();
                }
                syncStream(h);
                q_ms += msSince(t0);
            }
            i = j;
        }
        syncStream(h);
        h->last_ms["run_batch_queries"] = q_ms;
        h->last_ms["run_batch_updates"] = u_ms;
        h->last_ms["run_batch"] = q_ms + u_ms;
        if (stats) *stats = tot;
    });
}

// Tuning: graph/query features, a measuring autotuner and the learned tuner.
// No counterpart in the reference, whose knobs are fixed constants (delta 16,
// gir.h:215; alpha 15 / beta 18, gir_query.cpp:161; TRIM_TIMES 3, util.h:34).
// Every knob only changes work order, never a result.
//
// Features (16 floats, computed on the device from the current graph and the
// number of query sources k):
//   0 log2(V+1)          1 log2(E+1)          2 E/V
//   3 log2(max out+1)    4 log2(max in+1)     5 fraction with out-degree 0
//   6 fraction in-deg 0  7 out-degree CV      8 log2(mean weight)
//   9 log2(max weight)  10 log2(min weight)  11 log2(probe BFS levels+1)
//  12 probe reach frac  13 reciprocity (sampled arcs with a reverse arc)
//  14 log2(k)           15 log2(E[d^2]/E[d]+1)
// The probe BFS starts at the vertex of largest out-degree.
#include <cmath>
#include <cstring>
#include <vector>

#include "common.h"
#include "reach.h"

struct girhip_mlp;
extern "C" int girhip_mlp_forward(girhip_mlp* m, const float* x, int32_t n, float* y);
extern "C" int girhip_sssp(girhip_graph* h, const int32_t* src, int32_t k, uint64_t delta, uint64_t* dist,
                           int64_t* edges_out);
extern "C" int girhip_bfs(girhip_graph* h, const int32_t* src, int32_t k, int32_t* levels, int64_t* edges_out);
extern "C" int girhip_scc_dag(girhip_graph* h, int32_t* scc, int32_t* n_scc, int64_t* n_dag);
extern "C" int girhip_wcc(girhip_graph* h, int32_t* label, int32_t* n_wcc);

namespace gir {

// s[0] zero-out, s[1] zero-in, s[2] sum d^2, s[3] max out (key), s[4] max in
__global__ void kDegStats(const int32_t* __restrict__ fb, const int32_t* __restrict__ bb, int32_t V, int64_t* s) {
    int64_t zo = 0, zi = 0, d2 = 0;
    unsigned long long mo = 0, mi = 0;
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < V; v += (int64_t)gridDim.x * blockDim.x) {
        const int64_t o = fb[v + 1] - fb[v], in = bb[v + 1] - bb[v];
        zo += o == 0;
        zi += in == 0;
        d2 += o * o;
        const unsigned long long ko = ((unsigned long long)o << 32) | (uint32_t)(0x7fffffff - v);
        mo = ko > mo ? ko : mo;
        mi = (unsigned long long)in > mi ? (unsigned long long)in : mi;
    }
    blockAtomicAdd(zo, &s[0]);
    blockAtomicAdd(zi, &s[1]);
    blockAtomicAdd(d2, &s[2]);
    for (int d = WAVE / 2; d > 0; d >>= 1) {
        const unsigned long long a = __shfl_xor(mo, d), b = __shfl_xor(mi, d);
        mo = a > mo ? a : mo;
        mi = b > mi ? b : mi;
    }
    if (laneId() == 0) {
        atomicMax((unsigned long long*)&s[3], mo);
        atomicMax((unsigned long long*)&s[4], mi);
    }
}

// s[5] sum w, s[6] max w, s[7] min w (as uint64)
__global__ void kWeightStats(const uint32_t* __restrict__ w, int64_t E, int64_t* s) {
    int64_t sum = 0;
    unsigned long long mx = 0, mn = ~0ULL;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < E; i += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t x = w[i];
        sum += x;
        mx = x > mx ? x : mx;
        mn = x < mn ? x : mn;
    }
    blockAtomicAdd(sum, &s[5]);
    for (int d = WAVE / 2; d > 0; d >>= 1) {
        const unsigned long long a = __shfl_xor(mx, d), b = __shfl_xor(mn, d);
        mx = a > mx ? a : mx;
        mn = b < mn ? b : mn;
    }
    if (laneId() == 0) {
        atomicMax((unsigned long long*)&s[6], mx);
        atomicMin((unsigned long long*)&s[7], mn);
    }
}

// s[8] sampled arcs with a reverse arc (nsample arcs spread over the slots)
__global__ void kReciprocity(const int32_t* __restrict__ src, const int32_t* __restrict__ fb,
                             const int32_t* __restrict__ fa, int64_t E, int nsample, int64_t* s) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    int64_t r = 0;
    if (i < nsample) {
        uint64_t x = 0x9E3779B97F4A7C15ULL * (uint64_t)(i + 1);
        x = (x ^ (x >> 31)) * 0xBF58476D1CE4E5B9ULL;
        x ^= x >> 29;
        const int64_t e = (int64_t)(x % (uint64_t)E);
        const int32_t u = src[e], v = fa[e];
        int32_t lo = fb[v], hi = fb[v + 1];
        while (lo < hi) {
            const int32_t m = lo + ((hi - lo) >> 1);
            if (fa[m] < u) lo = m + 1;
            else hi = m;
        }
        r = lo < fb[v + 1] && fa[lo] == u;
    }
    blockAtomicAdd(r, &s[8]);
}

void computeFeatures(girhip_graph* h, int32_t k, float* f) {
    const int32_t V = h->V;
    const int64_t E = h->E;
    int64_t* s = h->dcnt.p + 48;
    fillI64(h, s, 0, 9);
    fillU64(h, (uint64_t*)(s + 7), ~0ULL, 1);
    launch(h, "feat_deg", kDegStats, boundedGrid(V, 256), 256, 0, (const int32_t*)h->fw_beg.p,
           (const int32_t*)h->bw_beg.p, V, s);
    if (h->has_w && E > 0)
        launch(h, "feat_w", kWeightStats, boundedGrid(E, 256), 256, 0, (const uint32_t*)h->w_fw.p, E, s);
    const int ns = 8192;
    if (E > 0)
        launch(h, "feat_recip", kReciprocity, gridFor(ns, 256), 256, 0, fwSrc(h), (const int32_t*)h->fw_beg.p,
               (const int32_t*)h->fw_adj.p, E, ns, s);
    int64_t r[9];
    readCounters(h, s, 9, r);
    double mean_w = 50.5, max_w = 100, min_w = 1;  // GIR::edgeWeight: uniform [1, 100]
    if (h->has_w && E > 0) {
        mean_w = (double)r[5] / (double)E;
        max_w = (double)(uint64_t)r[6];
        min_w = (double)(uint64_t)r[7];
    }
    const uint64_t mo = (uint64_t)r[3];
    const int64_t maxout = (int64_t)(mo >> 32);
    const int32_t hub = (int32_t)(0x7fffffff - (int64_t)(mo & 0xffffffffULL));
    // probe BFS from the hub
    ReachStats st;
    if (V > 0 && hub >= 0 && hub < V) {
        ReachBufs B = reachBufs(h);
        fillU64(h, B.vis, 0, B.W);
        fillI32(h, B.q[0], hub, 1);
        st = runReach(h, h->fw_beg.p, h->fw_adj.p, h->bw_beg.p, h->bw_adj.p, B, 1, maxout, (int64_t)V, E, AllArcs{},
                      NoVisit{});
    }
    const double Vd = V > 0 ? V : 1, Ed = (double)E, md = Ed / Vd;
    const double var = (double)r[2] / Vd - md * md;
    f[0] = (float)std::log2(Vd + 1);
    f[1] = (float)std::log2(Ed + 1);
    f[2] = (float)md;
    f[3] = (float)std::log2((double)maxout + 1);
    f[4] = (float)std::log2((double)r[4] + 1);
    f[5] = (float)(r[0] / Vd);
    f[6] = (float)(r[1] / Vd);
    f[7] = (float)(md > 0 ? std::sqrt(var > 0 ? var : 0) / md : 0);
    f[8] = (float)std::log2(mean_w > 0 ? mean_w : 1);
    f[9] = (float)std::log2(max_w > 0 ? max_w : 1);
    f[10] = (float)std::log2(min_w > 0 ? min_w : 1);
    f[11] = (float)std::log2((double)st.levels + 1);
    f[12] = (float)(st.reached / Vd);
    f[13] = (float)(E > 0 ? (double)r[8] / ns : 0);
    f[14] = (float)std::log2((double)(k > 0 ? k : 1));
    f[15] = (float)std::log2((md > 0 ? ((double)r[2] / Vd) / md : 0) + 1);
}

}  // namespace gir

using namespace gir;

extern "C" int girhip_features(girhip_graph* h, int32_t k, float* features) {
    return guarded([&] {
        useDevice(h);
        if (!features) raise(GIRHIP_EINVAL, "features: null output");
        computeFeatures(h, k, features);
        syncStream(h);
    });
}

// Times `ncand` candidate values of one knob on this graph and sets the
// fastest.  what: "sssp_delta" (runs SSSP from src), "scc_trim_rounds"
// (buildSCC_DAG), "bfs_alpha" / "bfs_beta" / "msbfs" (BFS from src),
// "block" (all three).  times (may be NULL) receives each candidate's device
// ms (best of `reps` runs).
extern "C" int girhip_autotune(girhip_graph* h, const char* what, const int32_t* src, int32_t k, const double* cand,
                               int32_t ncand, int32_t reps, double* best_value, double* times) {
    return guarded([&] {
        useDevice(h);
        if (!what || !cand || ncand <= 0 || reps <= 0) raise(GIRHIP_EINVAL, "autotune: bad argument");
        const std::string w(what);
        double cur = 0;
        if (girhip_get_param(h, what, &cur) != GIRHIP_OK) raise(GIRHIP_EINVAL, "autotune: unknown knob " + w);
        auto run = [&]() -> double {
            double ms = 0;
            int rc = GIRHIP_OK;
            if (w == "sssp_delta") {
                rc = girhip_sssp(h, src, k, 0, nullptr, nullptr);
                ms = h->last_ms["sssp"];
            } else if (w.rfind("scc_", 0) == 0) {
                rc = girhip_scc_dag(h, nullptr, nullptr, nullptr);
                ms = h->last_ms["scc_dag"];
            } else if (w.rfind("bfs_", 0) == 0 || w == "msbfs") {
                rc = girhip_bfs(h, src, k, nullptr, nullptr);
                ms = h->last_ms["bfs"];
            } else {
                rc = girhip_wcc(h, nullptr, nullptr);
                if (rc == GIRHIP_OK) rc = girhip_scc_dag(h, nullptr, nullptr, nullptr);
                if (rc == GIRHIP_OK && k > 0) rc = girhip_bfs(h, src, k, nullptr, nullptr);
                if (rc == GIRHIP_OK && k > 0) rc = girhip_sssp(h, src, k, 0, nullptr, nullptr);
                ms = h->last_ms["wcc"] + h->last_ms["scc_dag"] + (k > 0 ? h->last_ms["bfs"] + h->last_ms["sssp"] : 0);
            }
            if (rc != GIRHIP_OK) raise(rc, girhip_last_error());
            return ms;
        };
        double best = cur, best_ms = 1e300;
        for (int i = 0; i < ncand; ++i) {
            if (girhip_set_param(h, what, cand[i]) != GIRHIP_OK) raise(GIRHIP_EINVAL, girhip_last_error());
            double t = 1e300;
            for (int r = 0; r < reps; ++r) t = std::min(t, run());
            if (times) times[i] = t;
            if (t < best_ms) {
                best_ms = t;
                best = cand[i];
            }
        }
        girhip_set_param(h, what, best);
        if (best_value) *best_value = best;
    });
}

extern "C" int girhip_query_batch(girhip_graph* h, const int32_t* src, const uint8_t* kind, int32_t n,
                                  uint64_t delta, int32_t* levels, uint64_t* dist) {
    return guarded([&] {
        useDevice(h);
        if (n < 0 || (n > 0 && (!src || !kind))) raise(GIRHIP_EINVAL, "query_batch: bad argument");
        std::vector<int32_t> s(n);
        std::vector<uint8_t> kd(n);
        if (n) {
            HIPCHK(hipMemcpy(s.data(), src, sizeof(int32_t) * n, hipMemcpyDefault));
            HIPCHK(hipMemcpy(kd.data(), kind, (size_t)n, hipMemcpyDefault));
        }
        std::vector<int32_t> b, p;
        for (int32_t i = 0; i < n; ++i) {
            if (kd[i] > 1) raise(GIRHIP_EINVAL, "query_batch: kind must be 0 (BFS) or 1 (SSSP)");
            (kd[i] ? p : b).push_back(s[i]);
        }
        double ms = 0;
        if (!b.empty()) {
            const int rc = girhip_bfs(h, b.data(), (int32_t)b.size(), levels, nullptr);
            if (rc != GIRHIP_OK) raise(rc, girhip_last_error());
            ms += h->last_ms["bfs"];
        }
        if (!p.empty()) {
            const int rc = girhip_sssp(h, p.data(), (int32_t)p.size(), delta, dist, nullptr);
            if (rc != GIRHIP_OK) raise(rc, girhip_last_error());
            ms += h->last_ms["sssp"];
        }
        h->last_ms["query_batch"] = ms;
    });
}

// SCC schedules the learned model chooses among (its second output is an
// index into this table): (pivot rule, trim rounds), in order of the work done
// before colouring.  Pivot rule 0: the remainder goes straight to colouring;
// 1: one forward-backward sweep from the vertex of largest out * in degree
// (the iSpan rule, scc_common.h:98-158).  The reference's default is (1, 3).
constexpr int N_SCHED = 9;
constexpr int SCHED[N_SCHED][2] = {{0, 1}, {0, 3}, {1, 0}, {1, 1}, {1, 2}, {1, 3}, {1, 4}, {1, 6}, {1, 8}};

// Learned tuning: features -> MLP (on the device) -> delta and SCC schedule.
// norm: 36 floats = feature mean[16], feature scale[16], target mean[2],
// target scale[2]; outputs are log2(delta) and an index into SCHED.
// The reference's defaults (delta 16, pivot rule 1 with 3 trim rounds) are
// the fallback (non-finite or out-of-range predictions, features far outside
// the training range) and, when judge > 0, the judge of both predictions:
// SSSP from min(k, judge) of the sources with the predicted delta and with
// the default, and one buildSCC_DAG with each schedule; the faster setting is
// kept.  out (may be NULL): 6 doubles = chosen delta, chosen trim rounds, raw
// predicted delta, 1 if the predicted delta was kept, chosen pivot rule, 1
// if the predicted SCC schedule was kept.
extern "C" int girhip_learned_tune(girhip_graph* h, girhip_mlp* model, const float* norm, const int32_t* src,
                                   int32_t k, int32_t judge, double* out) {
    return guarded([&] {
        useDevice(h);
        if (!model || !norm) raise(GIRHIP_EINVAL, "learned_tune: null model");
        float f[16], x[16], y[2];
        computeFeatures(h, k, f);
        syncStream(h);
        bool sane = true;
        for (int i = 0; i < 16; ++i) {
            x[i] = (f[i] - norm[i]) / (norm[16 + i] != 0 ? norm[16 + i] : 1.0f);
            if (!std::isfinite(x[i]) || std::fabs(x[i]) > 8) sane = false;  // far outside the training data
        }
        if (girhip_mlp_forward(model, x, 1, y) != GIRHIP_OK) raise(GIRHIP_EHIP, girhip_last_error());
        const double ld = y[0] * norm[34] + norm[32], ls = y[1] * norm[35] + norm[33];
        const Params defaults;
        double delta = defaults.sssp_delta, raw = std::exp2(ld);
        double trim = defaults.scc_trim_rounds, pivot = defaults.scc_pivot;
        bool used = false, sched_used = false;
        if (sane && std::isfinite(ld) && ld >= 0 && ld <= 24) {
            delta = std::max(1.0, std::round(std::exp2(ld)));
            used = true;
        }
        if (sane && std::isfinite(ls) && ls > -0.5 && ls < N_SCHED - 0.5) {
            const int si = (int)std::lround(ls);
            pivot = SCHED[si][0];
            trim = SCHED[si][1];
            sched_used = !(pivot == defaults.scc_pivot && trim == defaults.scc_trim_rounds);
        }
        if (used && judge > 0 && k > 0 && src) {
            const int32_t kj = std::min(k, judge);
            h->prm.sssp_delta = defaults.sssp_delta;
            if (girhip_sssp(h, src, kj, 0, nullptr, nullptr) != GIRHIP_OK) raise(GIRHIP_EHIP, girhip_last_error());
            const double t_def = h->last_ms["sssp"];
            h->prm.sssp_delta = delta;
            if (girhip_sssp(h, src, kj, 0, nullptr, nullptr) != GIRHIP_OK) raise(GIRHIP_EHIP, girhip_last_error());
            const double t_new = h->last_ms["sssp"];
            if (t_new > t_def) {
                delta = defaults.sssp_delta;
                used = false;
            }
        }
        if (sched_used && judge > 0) {
            auto build = [&](double p, double t) {
                h->prm.scc_pivot = p;
                h->prm.scc_trim_rounds = t;
                if (girhip_scc_dag(h, nullptr, nullptr, nullptr) != GIRHIP_OK) raise(GIRHIP_EHIP, girhip_last_error());
                return h->last_ms["scc_dag"];
            };
            const double t_def = build(defaults.scc_pivot, defaults.scc_trim_rounds);
            if (build(pivot, trim) > t_def) {
                pivot = defaults.scc_pivot;
                trim = defaults.scc_trim_rounds;
                sched_used = false;
            }
        }
        h->prm.sssp_delta = delta;
        h->prm.scc_trim_rounds = trim;
        h->prm.scc_pivot = pivot;
        if (out) {
            out[0] = delta;
            out[1] = trim;
            out[2] = raw;
            out[3] = used ? 1 : 0;
            out[4] = pivot;
            out[5] = sched_used ? 1 : 0;
        }
    });
}

// CPU restatement of the reference GIR construction/query path.
//
// TEST INFRASTRUCTURE ONLY: the checker for the HIP path and the "port" CPU
// baseline.  Nothing in the product library links or calls this file.
//
// Each function cites the reference lines it restates (paths relative to the
// reference repository, gir_constructor/).  Component ids are canonicalised to
// the minimum vertex id of the component.
#include "oracle.h"

#include <omp.h>

#include <algorithm>
#include <climits>
#include <cstring>
#include <queue>
#include <utility>
#include <vector>

namespace {

template <class T>
inline bool atomicMinU(T* p, T v) {
    T old = __atomic_load_n(p, __ATOMIC_RELAXED);
    while (v < old)
        if (__atomic_compare_exchange_n(p, &old, v, false, __ATOMIC_RELAXED, __ATOMIC_RELAXED)) return true;
    return false;
}

// Relabel so every group is named by its smallest member.
void canonicalise(int32_t V, const int32_t* raw, int32_t* out, int32_t* n_groups) {
    std::vector<int32_t> minv(V, INT32_MAX);
    for (int32_t v = 0; v < V; ++v) minv[raw[v]] = std::min(minv[raw[v]], v);
    int32_t n = 0;
    for (int32_t v = 0; v < V; ++v) {
        out[v] = minv[raw[v]];
        n += out[v] == v;
    }
    if (n_groups) *n_groups = n;
}

}  // namespace

extern "C" int orc_num_threads(void) { return omp_get_max_threads(); }

// gir_query.cpp:50-54
extern "C" uint32_t orc_edge_weight(int32_t u, int32_t v) {
    uint64_t x = ((uint64_t)(uint32_t)u << 32) | (uint32_t)v;
    x ^= x >> 33;
    x *= 0xff51afd7ed558ccdULL;
    x ^= x >> 33;
    x *= 0xc4ceb9fe1a85ec53ULL;
    x ^= x >> 33;
    return 1 + (uint32_t)(x % 100);
}

// gir.cpp:95-116
extern "C" int orc_normalize_csr(int32_t V, int32_t* beg, int32_t* adj, int64_t* E) {
    if (V < 0 || !beg || !E) return 1;
    std::vector<int32_t> deg(V);
    #pragma omp parallel for schedule(dynamic, 1024)
    for (int32_t v = 0; v < V; ++v) {
        std::sort(adj + beg[v], adj + beg[v + 1]);
        deg[v] = (int32_t)(std::unique(adj + beg[v], adj + beg[v + 1]) - (adj + beg[v]));
    }
    // compact rows to the front (sequential: rows only move left)
    int64_t k = 0;
    for (int32_t v = 0; v < V; ++v) {
        int32_t b = beg[v];
        beg[v] = (int32_t)k;
        if (b != k) std::memmove(adj + k, adj + b, sizeof(int32_t) * deg[v]);
        k += deg[v];
    }
    beg[V] = (int32_t)k;
    *E = k;
    return 0;
}

// gir.cpp:635-653: union-find with path halving over the forward arcs
// (parent[a] = b on a union), then every vertex labelled by its root.  The
// roots are then renamed to the minimum member.
extern "C" int orc_wcc(int32_t V, const int32_t* fw_beg, const int32_t* fw_adj, int32_t* label, int32_t* n_wcc) {
    if (V < 0 || (V > 0 && (!fw_beg || !label))) return 1;
    std::vector<int32_t> parent(V);
    for (int32_t i = 0; i < V; ++i) parent[i] = i;
    auto find = [&](int32_t x) {
        while (parent[x] != x) {
            parent[x] = parent[parent[x]];
            x = parent[x];
        }
        return x;
    };
    for (int32_t u = 0; u < V; ++u)
        for (int32_t i = fw_beg[u]; i < fw_beg[u + 1]; ++i) {
            int32_t a = find(u), b = find(fw_adj[i]);
            if (a != b) parent[a] = b;
        }
    std::vector<int32_t> root(V);
    for (int32_t i = 0; i < V; ++i) root[i] = find(i);
    canonicalise(V, root.data(), label, n_wcc);
    return 0;
}

namespace {
// gir.cpp:494-537: iterative Tarjan, roots in id order, each vertex's
// neighbours visited in row order; emit(component) in emission order (which
// is a reverse topological order of the condensation).
template <class Emit>
void tarjan(int32_t V, const int32_t* beg, const int32_t* adj, Emit emit) {
    std::vector<int32_t> idx(V, -1), low(V, 0), stk;
    std::vector<char> on(V, 0);
    struct Frame {
        int32_t v;
        int32_t next;  // next CSR slot to visit
    };
    std::vector<Frame> cs;
    int32_t t = 0;
    for (int32_t r = 0; r < V; ++r) {
        if (idx[r] != -1) continue;
        idx[r] = low[r] = t++;
        stk.push_back(r);
        on[r] = 1;
        cs.push_back({r, beg[r]});
        while (!cs.empty()) {
            Frame& f = cs.back();
            const int32_t v = f.v;
            if (f.next < beg[v + 1]) {
                const int32_t w = adj[f.next++];
                if (idx[w] == -1) {
                    idx[w] = low[w] = t++;
                    stk.push_back(w);
                    on[w] = 1;
                    cs.push_back({w, beg[w]});
                } else if (on[w]) {
                    low[v] = std::min(low[v], idx[w]);
                }
            } else {
                cs.pop_back();
                if (!cs.empty()) low[cs.back().v] = std::min(low[cs.back().v], low[v]);
                if (low[v] == idx[v]) {
                    std::vector<int32_t> comp;
                    while (true) {
                        int32_t x = stk.back();
                        stk.pop_back();
                        on[x] = 0;
                        comp.push_back(x);
                        if (x == v) break;
                    }
                    emit(comp);
                }
            }
        }
    }
}
}  // namespace

// gir.cpp:798-816
extern "C" int orc_scc(int32_t V, const int32_t* fw_beg, const int32_t* fw_adj, int32_t* scc, int32_t* n_scc) {
    if (V < 0 || (V > 0 && (!fw_beg || !scc))) return 1;
    int32_t n = 0;
    tarjan(V, fw_beg, fw_adj, [&](const std::vector<int32_t>& c) {
        int32_t m = *std::min_element(c.begin(), c.end());
        for (int32_t x : c) scc[x] = m;
        ++n;
    });
    if (n_scc) *n_scc = n;
    return 0;
}

// gir.cpp:818-831 (dag_out[su][sw]++ for every arc with su != sw)
extern "C" int orc_dag(int32_t V, const int32_t* fw_beg, const int32_t* fw_adj, const int32_t* scc, int32_t* src,
                       int32_t* dst, int32_t* mult, int64_t* n_dag) {
    if (V < 0 || !n_dag || (V > 0 && (!fw_beg || !scc))) return 1;
    std::vector<uint64_t> keys;
    for (int32_t u = 0; u < V; ++u)
        for (int32_t i = fw_beg[u]; i < fw_beg[u + 1]; ++i) {
            int32_t a = scc[u], b = scc[fw_adj[i]];
            if (a != b) keys.push_back(((uint64_t)(uint32_t)a << 32) | (uint32_t)b);
        }
    std::sort(keys.begin(), keys.end());
    int64_t k = 0;
    for (size_t i = 0; i < keys.size();) {
        size_t j = i;
        while (j < keys.size() && keys[j] == keys[i]) ++j;
        if (src) {
            src[k] = (int32_t)(keys[i] >> 32);
            dst[k] = (int32_t)(uint32_t)keys[i];
            mult[k] = (int32_t)(j - i);
        }
        ++k;
        i = j;
    }
    *n_dag = k;
    return 0;
}

// gir.cpp:833-836
extern "C" int orc_scc_ranks(int32_t V, const int32_t* fw_beg, const int32_t* fw_adj, int64_t* rank) {
    if (V < 0 || (V > 0 && (!fw_beg || !rank))) return 1;
    std::vector<int32_t> rep;
    tarjan(V, fw_beg, fw_adj, [&](const std::vector<int32_t>& c) { rep.push_back(*std::min_element(c.begin(), c.end())); });
    const int64_t S = (int64_t)rep.size();
    for (int32_t v = 0; v < V; ++v) rank[v] = -1;
    for (int64_t s = 0; s < S; ++s) rank[rep[s]] = (S - s) << 32;
    return 0;
}

// gir_query.cpp:108-294 (the clean-store branch: no tombstones, no deltas).
extern "C" int orc_bfs(int32_t V, const int32_t* fb, const int32_t* fc, const int32_t* bb, const int32_t* bc,
                       int32_t source, int32_t* level) {
    if (V < 0 || (V > 0 && (!fb || !bb || !level))) return 1;
    #pragma omp parallel for schedule(static)
    for (int32_t i = 0; i < V; ++i) level[i] = -1;
    if (source < 0 || source >= V) return 0;
    level[source] = 0;
    const int64_t n_reach = V, m_total = fb[V];
    const size_t W = ((size_t)V + 63) / 64;
    std::vector<uint64_t> vbits(W, 0), fbits(W), nbits(W);
    vbits[(size_t)source >> 6] |= 1ULL << (source & 63);
    std::vector<int32_t> frontier{source}, next;
    const int T = omp_get_max_threads();
    std::vector<std::vector<int32_t>> local(T);
    std::vector<int64_t> m_part(T);
    auto outdeg = [&](int32_t v) { return (int64_t)(fb[v + 1] - fb[v]); };
    int64_t m_f = outdeg(source), m_unexplored = m_total - m_f;
    const int64_t alpha = 15, beta = 18, large_unit = 4096;
    int32_t L = 0;
    bool bottom_up = false, bits_valid = false;
    while (!frontier.empty()) {
        if (n_reach >= large_unit) {
            if (!bottom_up && m_f * alpha > m_unexplored) bottom_up = true;
            else if (bottom_up && (int64_t)frontier.size() * beta < n_reach) bottom_up = false;
        }
        std::fill(m_part.begin(), m_part.end(), 0);
        if (!bottom_up) {
            #pragma omp parallel
            {
                const int t = omp_get_thread_num();
                auto& out = local[t];
                int64_t m = 0;
                #pragma omp for schedule(dynamic, 64)
                for (size_t i = 0; i < frontier.size(); ++i) {
                    const int32_t u = frontier[i];
                    for (int32_t k = fb[u]; k < fb[u + 1]; ++k) {
                        const int32_t w = fc[k];
                        if ((vbits[(size_t)w >> 6] >> (w & 63)) & 1) continue;
                        if (__atomic_load_n(&level[w], __ATOMIC_RELAXED) != -1) continue;
                        int32_t expect = -1;
                        if (__atomic_compare_exchange_n(&level[w], &expect, L + 1, false, __ATOMIC_RELAXED,
                                                        __ATOMIC_RELAXED)) {
                            out.push_back(w);
                            m += outdeg(w);
                        }
                    }
                }
                m_part[t] = m;
            }
            bits_valid = false;
        } else {
            if (!bits_valid) {
                std::fill(fbits.begin(), fbits.end(), 0);
                for (int32_t f : frontier) fbits[(size_t)f >> 6] |= 1ULL << (f & 63);
            }
            #pragma omp parallel
            {
                const int t = omp_get_thread_num();
                auto& out = local[t];
                int64_t m = 0;
                #pragma omp for schedule(dynamic, 64)
                for (size_t wd = 0; wd < W; ++wd) {
                    uint64_t bits = 0;
                    const uint64_t vis = vbits[wd];
                    const int32_t v0 = (int32_t)(wd << 6), v1 = std::min<int64_t>(V, (int64_t)v0 + 64);
                    if (vis == ~0ULL) {
                        nbits[wd] = 0;
                        continue;
                    }
                    for (int32_t v = v0; v < v1; ++v) {
                        if ((vis >> (v - v0)) & 1) continue;
                        bool hit = false;
                        for (int32_t k = bb[v]; k < bb[v + 1]; ++k) {
                            const int32_t p = bc[k];
                            if ((fbits[(uint32_t)p >> 6] >> (p & 63)) & 1) {
                                hit = true;
                                break;
                            }
                        }
                        if (hit) {
                            level[v] = L + 1;
                            bits |= 1ULL << (v - v0);
                            out.push_back(v);
                            m += outdeg(v);
                        }
                    }
                    nbits[wd] = bits;
                    vbits[wd] = vis | bits;
                }
                m_part[t] = m;
            }
            fbits.swap(nbits);
            bits_valid = true;
        }
        next.clear();
        for (auto& p : local) {
            next.insert(next.end(), p.begin(), p.end());
            p.clear();
        }
        if (!bottom_up)
            for (int32_t x : next) vbits[(size_t)x >> 6] |= 1ULL << (x & 63);
        int64_t m_next = 0;
        for (int64_t m : m_part) m_next += m;
        m_unexplored -= m_next;
        m_f = m_next;
        frontier.swap(next);
        ++L;
    }
    return 0;
}

// gir_query.cpp:629-703: per-thread bucket lists, bucket k swapped out and
// processed until nothing new lands in it, duplicates skipped with a
// per-round marker, stale entries (dist/delta != k) skipped, every out-arc
// relaxed with a CAS min.
extern "C" int orc_sssp(int32_t V, const int32_t* fb, const int32_t* fc, const uint32_t* w_fw, int32_t source,
                        uint64_t delta, uint64_t* dist) {
    if (V < 0 || delta == 0 || (V > 0 && (!fb || !dist))) return 1;
    #pragma omp parallel for schedule(static)
    for (int32_t i = 0; i < V; ++i) dist[i] = UINT64_MAX;
    if (source < 0 || source >= V) return 0;
    dist[source] = 0;
    const int T = omp_get_max_threads();
    std::vector<std::vector<std::vector<int32_t>>> tb(T);
    std::vector<size_t> tmax(T, 0);
    auto push = [&](int t, int32_t x, uint64_t d) {
        size_t k = (size_t)(d / delta);
        auto& B = tb[t];
        if (B.size() <= k) B.resize(k + 1);
        B[k].push_back(x);
        if (k > tmax[t]) tmax[t] = k;
    };
    push(0, source, 0);
    std::vector<int32_t> mark(V, 0);
    int32_t epoch = 0;
    std::vector<std::vector<int32_t>> cur(T);
    std::vector<size_t> off(T + 1);
    for (size_t k = 0;; ++k) {
        size_t kmax = 0;
        for (int t = 0; t < T; ++t) kmax = std::max(kmax, tmax[t]);
        if (k > kmax) break;
        while (true) {
            off[0] = 0;
            for (int t = 0; t < T; ++t) {
                cur[t].clear();
                if (tb[t].size() > k) cur[t].swap(tb[t][k]);
                off[t + 1] = off[t] + cur[t].size();
            }
            if (off[T] == 0) break;
            const int32_t round = ++epoch;
            const size_t m = off[T];
            #pragma omp parallel for schedule(dynamic, 256)
            for (size_t i = 0; i < m; ++i) {
                const size_t s = std::upper_bound(off.begin(), off.end(), i) - off.begin() - 1;
                const int32_t u = cur[s][i - off[s]];
                if (__atomic_exchange_n(&mark[u], round, __ATOMIC_RELAXED) == round) continue;
                const uint64_t du = __atomic_load_n(&dist[u], __ATOMIC_RELAXED);
                if (du / delta != k) continue;
                const int t = omp_get_thread_num();
                for (int32_t e = fb[u]; e < fb[u + 1]; ++e) {
                    const int32_t w = fc[e];
                    const uint64_t nd = du + (w_fw ? w_fw[e] : orc_edge_weight(u, w));
                    if (atomicMinU(&dist[w], nd)) push(t, w, nd);
                }
            }
        }
    }
    return 0;
}

namespace {
inline uint64_t keyD(int32_t u, int32_t v) { return ((uint64_t)(uint32_t)u << 32) | (uint32_t)v; }
bool present(const int32_t* fb, const int32_t* fc, int32_t u, int32_t v) {
    return std::binary_search(fc + fb[u], fc + fb[u + 1], v);
}
}  // namespace

// gir.cpp:3523-3586
extern "C" int orc_normalize_batch(int32_t V, const int32_t* fb, const int32_t* fc, int64_t n, const int32_t* bu,
                                   const int32_t* bv, const uint8_t* bins, int32_t* ins_u, int32_t* ins_v,
                                   int64_t* n_ins, int32_t* del_u, int32_t* del_v, int64_t* n_del,
                                   int64_t* n_noop) {
    if (V < 0 || n < 0 || !n_ins || !n_del || !n_noop) return 1;
    std::vector<std::pair<uint64_t, int64_t>> kv;
    kv.reserve(n);
    for (int64_t i = 0; i < n; ++i) {
        const int32_t u = bu[i], v = bv[i];
        if (u == v || u < 0 || v < 0 || u >= V || v >= V) continue;
        kv.push_back({keyD(u, v), i});
    }
    std::sort(kv.begin(), kv.end());
    int64_t a = 0, d = 0, nn = 0;
    for (size_t i = 0; i < kv.size(); ++i) {
        if (i + 1 < kv.size() && kv[i + 1].first == kv[i].first) continue;  // not the last op on this arc
        const int32_t u = (int32_t)(kv[i].first >> 32), v = (int32_t)(uint32_t)kv[i].first;
        const bool want = bins[kv[i].second] != 0;
        const bool has = present(fb, fc, u, v);
        if (want == has) {
            nn += want;
            continue;
        }
        if (want) {
            ins_u[a] = u;
            ins_v[a] = v;
            ++a;
        } else {
            del_u[d] = u;
            del_v[d] = v;
            ++d;
        }
    }
    *n_ins = a;
    *n_del = d;
    *n_noop = nn;
    return 0;
}

extern "C" int orc_apply_batch(int32_t V, const int32_t* fb, const int32_t* fc, int64_t n, const int32_t* bu,
                               const int32_t* bv, const uint8_t* bins, int32_t* nfb, int32_t* nfc, int32_t* nbb,
                               int32_t* nbc, int64_t* nE) {
    if (V < 0 || n < 0 || !nE) return 1;
    std::vector<int32_t> iu(n), iv(n), du(n), dv(n);
    int64_t ni = 0, nd = 0, nn = 0;
    if (orc_normalize_batch(V, fb, fc, n, bu, bv, bins, iu.data(), iv.data(), &ni, du.data(), dv.data(), &nd, &nn))
        return 1;
    std::vector<uint64_t> keys;
    keys.reserve((size_t)fb[V] + ni);
    size_t di = 0;
    std::vector<uint64_t> dk(nd);
    for (int64_t i = 0; i < nd; ++i) dk[i] = keyD(du[i], dv[i]);
    for (int32_t u = 0; u < V; ++u)
        for (int32_t e = fb[u]; e < fb[u + 1]; ++e) {
            const uint64_t k = keyD(u, fc[e]);
            while (di < dk.size() && dk[di] < k) ++di;
            if (di < dk.size() && dk[di] == k) continue;
            keys.push_back(k);
        }
    for (int64_t i = 0; i < ni; ++i) keys.push_back(keyD(iu[i], iv[i]));
    std::sort(keys.begin(), keys.end());
    const int64_t E = (int64_t)keys.size();
    std::fill(nfb, nfb + V + 1, 0);
    std::fill(nbb, nbb + V + 1, 0);
    for (uint64_t k : keys) {
        nfb[(k >> 32) + 1]++;
        nbb[(uint32_t)k + 1]++;
    }
    for (int32_t v = 0; v < V; ++v) {
        nfb[v + 1] += nfb[v];
        nbb[v + 1] += nbb[v];
    }
    std::vector<int32_t> pos(nbb, nbb + V);
    for (int64_t i = 0; i < E; ++i) {
        const int32_t u = (int32_t)(keys[i] >> 32), v = (int32_t)(uint32_t)keys[i];
        nfc[i] = v;
        nbc[pos[v]++] = u;  // keys ascend in (u, v): every bw row fills in ascending u
    }
    *nE = E;
    return 0;
}

// ---------------- condensation-guided queries ----------------

namespace {
// The condensation of canonical SCC labels: deduplicated out- and in-arcs
// per SCC id (gir.cpp:818-831 builds dag_out / dag_in the same way).
struct Cond {
    std::vector<std::vector<int32_t>> out, in, mem;
};
Cond condensation(int32_t V, const int32_t* fb, const int32_t* fc, const int32_t* scc) {
    Cond C;
    C.out.assign(V, {});
    C.in.assign(V, {});
    C.mem.assign(V, {});
    for (int32_t v = 0; v < V; ++v) C.mem[scc[v]].push_back(v);
    for (int32_t u = 0; u < V; ++u)
        for (int32_t e = fb[u]; e < fb[u + 1]; ++e) {
            const int32_t a = scc[u], b = scc[fc[e]];
            if (a != b) C.out[a].push_back(b);
        }
    for (int32_t a = 0; a < V; ++a) {
        auto& o = C.out[a];
        std::sort(o.begin(), o.end());
        o.erase(std::unique(o.begin(), o.end()), o.end());
        for (int32_t b : o) C.in[b].push_back(a);
    }
    return C;
}

// gir_query.cpp:56-64 (reachableSccs): queue BFS over dag_out.
std::vector<uint8_t> reachableSccs(const Cond& C, int32_t s0) {
    std::vector<uint8_t> r(C.out.size(), 0);
    std::vector<int32_t> q{s0};
    r[s0] = 1;
    for (size_t i = 0; i < q.size(); ++i)
        for (int32_t d : C.out[q[i]])
            if (!r[d]) {
                r[d] = 1;
                q.push_back(d);
            }
    return r;
}

// gir_query.cpp:80-86: per reachable SCC, its reachable predecessors.
std::vector<int32_t> pendingCounts(const Cond& C, const std::vector<uint8_t>& r) {
    std::vector<int32_t> p(C.out.size(), 0);
    for (size_t c = 0; c < C.out.size(); ++c)
        if (r[c])
            for (int32_t a : C.in[c]) p[c] += r[a] ? 1 : 0;
    return p;
}
}  // namespace

// gir_query.cpp:66-88 (reachInfo) from the SCC with canonical id src_scc:
// reachable SCC mask, pending counts, counts = {SCCs, vertices, out-arcs}.
extern "C" int orc_reach_info(int32_t V, const int32_t* fb, const int32_t* fc, const int32_t* scc, int32_t src_scc,
                              uint8_t* mask, int32_t* pending, int64_t* counts) {
    if (V <= 0 || !fb || !scc || src_scc < 0 || src_scc >= V || scc[src_scc] != src_scc) return 1;
    const Cond C = condensation(V, fb, fc, scc);
    const auto r = reachableSccs(C, src_scc);
    const auto p = pendingCounts(C, r);
    int64_t ns = 0, nv = 0, ne = 0;
    for (int32_t c = 0; c < V; ++c) ns += r[c];
    for (int32_t v = 0; v < V; ++v)
        if (r[scc[v]]) {
            ++nv;
            ne += fb[v + 1] - fb[v];
        }
    if (mask) std::copy(r.begin(), r.end(), mask);
    if (pending) std::copy(p.begin(), p.end(), pending);
    if (counts) {
        counts[0] = ns;
        counts[1] = nv;
        counts[2] = ne;
    }
    return 0;
}

// gir_query.cpp:304-350 (dagWavefront) with the inner step of
// gir_query.cpp:604-620 (bfsTwoLevel, unit = 1: levels as distances, the
// multi-seed BFS inside an SCC) or 705-722 (ssspTwoLevel: delta-stepping of
// gir_query.cpp:629-693 restricted to the SCC, seeded with every member
// that holds a distance; single-vertex SCCs skip it).  Ready SCCs are
// processed one after another (the reference runs them as tasks; their
// results do not interact).  w_fw: weights per forward slot, NULL for
// orc_edge_weight.
extern "C" int orc_two_level(int32_t V, const int32_t* fb, const int32_t* fc, const uint32_t* w_fw,
                             const int32_t* scc, int32_t source, uint64_t delta, int32_t unit, uint64_t* dist) {
    if (V < 0 || (V > 0 && (!fb || !scc || !dist)) || (!unit && delta == 0)) return 1;
    for (int32_t i = 0; i < V; ++i) dist[i] = UINT64_MAX;
    if (source < 0 || source >= V) return 0;
    if (unit) delta = 1;
    auto wt = [&](int32_t u, int64_t e) -> uint64_t {
        return unit ? 1 : (w_fw ? w_fw[e] : orc_edge_weight(u, fc[e]));
    };
    const Cond C = condensation(V, fb, fc, scc);
    const int32_t s0 = scc[source];
    const auto r = reachableSccs(C, s0);
    auto left = pendingCounts(C, r);
    dist[source] = 0;
    // seeded bucket rounds inside SCC c: bucket k re-processed until nothing
    // new lands in it (gir_query.cpp:641-690, one thread)
    auto inner = [&](int32_t c) {
        std::vector<std::vector<int32_t>> bk;
        auto push = [&](int32_t x) {
            const size_t k = (size_t)(dist[x] / delta);
            if (bk.size() <= k) bk.resize(k + 1);
            bk[k].push_back(x);
        };
        for (int32_t x : C.mem[c])
            if (dist[x] != UINT64_MAX) push(x);
        for (size_t k = 0; k < bk.size(); ++k)
            while (!bk[k].empty()) {
                std::vector<int32_t> fr;
                fr.swap(bk[k]);
                std::sort(fr.begin(), fr.end());
                fr.erase(std::unique(fr.begin(), fr.end()), fr.end());
                for (int32_t u : fr) {
                    if (dist[u] / delta != k) continue;  // stale entry
                    for (int32_t e = fb[u]; e < fb[u + 1]; ++e) {
                        const int32_t x = fc[e];
                        if (scc[x] != c) continue;
                        const uint64_t nd = dist[u] + wt(u, e);
                        if (nd < dist[x]) {
                            dist[x] = nd;
                            push(x);
                        }
                    }
                }
            }
    };
    std::vector<int32_t> ready{s0};
    while (!ready.empty()) {
        std::vector<int32_t> next;
        for (int32_t c : ready) {
            if (C.mem[c].size() > 1) inner(c);
            for (int32_t x : C.mem[c])
                for (int32_t e = fb[x]; e < fb[x + 1] && dist[x] != UINT64_MAX; ++e) {
                    const int32_t y = fc[e];
                    if (scc[y] == c) continue;
                    const uint64_t nd = dist[x] + wt(x, e);
                    if (nd < dist[y]) dist[y] = nd;
                }
            for (int32_t d : C.out[c])
                if (--left[d] == 0) next.push_back(d);
        }
        ready.swap(next);
    }
    return 0;
}

// gir_query.cpp:728-778 (GIR::bc): per source a queue BFS in visit order
// that counts shortest paths (sigma), then the order in reverse, each
// vertex w giving every in-neighbour u one level up sigma[u] * (1 +
// delta[w]) / sigma[w]; score[w] += delta[w] for w != s.  Sources run one
// after another (the reference runs them as tasks and adds with atomics).
extern "C" int orc_bc(int32_t V, const int32_t* fb, const int32_t* fc, const int32_t* bb, const int32_t* bc,
                      const int32_t* src, int32_t k, double* score) {
    if (V < 0 || k < 0 || (V > 0 && (!fb || !bb || !score)) || (k > 0 && !src)) return 1;
    for (int32_t i = 0; i < V; ++i) score[i] = 0.0;
    std::vector<int32_t> dist(V, -1), order;
    std::vector<double> sigma(V, 0.0), delta(V, 0.0);
    for (int32_t si = 0; si < k; ++si) {
        const int32_t s = src[si];
        if (s < 0 || s >= V) continue;
        order.clear();
        dist[s] = 0;
        sigma[s] = 1;
        order.push_back(s);
        for (size_t i = 0; i < order.size(); ++i) {
            const int32_t u = order[i];
            for (int32_t e = fb[u]; e < fb[u + 1]; ++e) {
                const int32_t w = fc[e];
                if (dist[w] < 0) {
                    dist[w] = dist[u] + 1;
                    order.push_back(w);
                }
                if (dist[w] == dist[u] + 1) sigma[w] += sigma[u];
            }
        }
        for (size_t i = order.size(); i-- > 0;) {
            const int32_t w = order[i];
            const double coef = (1.0 + delta[w]) / sigma[w];
            for (int32_t e = bb[w]; e < bb[w + 1]; ++e) {
                const int32_t u = bc[e];
                if (dist[u] >= 0 && dist[u] == dist[w] - 1) delta[u] += sigma[u] * coef;
            }
            if (w != s) score[w] += delta[w];
        }
        for (int32_t x : order) {
            dist[x] = -1;
            sigma[x] = 0;
            delta[x] = 0;
        }
    }
    return 0;
}

// Biconnected blocks of every SCC's undirected view: the view of
// gir_test.cpp:448-465 (out- and in-neighbours inside the SCC, no self
// loops, no repeats) and the iterative Hopcroft-Tarjan of
// dynamic_repair.h:96-170 (discovery times, low points, an edge stack popped
// into a block when low[child] >= disc[parent]); vertices no block covers
// are blocks of their own (gir_test.cpp:462-464).  Output: membership pairs
// (block index, vertex), blocks numbered in emission order; cap = room for
// pairs.
extern "C" int orc_bicc(int32_t V, const int32_t* fb, const int32_t* fc, const int32_t* bb, const int32_t* bc,
                        const int32_t* scc, int32_t* mem_block, int32_t* mem_vertex, int64_t cap, int64_t* n_mem,
                        int64_t* n_blocks) {
    if (V < 0 || !n_mem || !n_blocks || (V > 0 && (!fb || !bb || !scc))) return 1;
    auto nbrs = [&](int32_t x, std::vector<int32_t>& out) {
        out.clear();
        int32_t i = fb[x], j = bb[x];
        const int32_t ie = fb[x + 1], je = bb[x + 1];
        while (i < ie || j < je) {
            int32_t w;
            if (j >= je || (i < ie && fc[i] <= bc[j])) {
                w = fc[i++];
                if (j < je && bc[j] == w) ++j;
            } else {
                w = bc[j++];
            }
            if (w != x && scc[w] == scc[x]) out.push_back(w);
        }
    };
    std::vector<int32_t> disc(V, -1), low(V, 0), par(V, -1);
    std::vector<uint8_t> covered(V, 0);
    std::vector<std::pair<int32_t, int32_t>> estack;
    struct Frame {
        int32_t v;
        std::vector<int32_t> nb;
        size_t next;
    };
    int64_t nm = 0, nbk = 0;
    bool overflow = false;
    auto emit = [&](std::vector<int32_t>& verts) {
        std::sort(verts.begin(), verts.end());
        verts.erase(std::unique(verts.begin(), verts.end()), verts.end());
        for (int32_t x : verts) {
            if (nm < cap) {
                mem_block[nm] = (int32_t)nbk;
                mem_vertex[nm] = x;
            } else {
                overflow = true;
            }
            ++nm;
            covered[x] = 1;
        }
        ++nbk;
    };
    int32_t timer = 0;
    std::vector<int32_t> blockv;
    for (int32_t root = 0; root < V; ++root) {
        if (disc[root] >= 0) continue;
        std::vector<Frame> st;
        disc[root] = low[root] = timer++;
        st.push_back({root, {}, 0});
        nbrs(root, st.back().nb);
        while (!st.empty()) {
            Frame& f = st.back();
            const int32_t v = f.v;
            if (f.next < f.nb.size()) {
                const int32_t w = f.nb[f.next++];
                if (w == par[v]) continue;  // the tree edge (a simple view has one)
                if (disc[w] < 0) {
                    disc[w] = low[w] = timer++;
                    par[w] = v;
                    estack.push_back({v, w});
                    st.push_back({w, {}, 0});
                    nbrs(w, st.back().nb);
                } else if (disc[w] < disc[v]) {
                    low[v] = std::min(low[v], disc[w]);
                    estack.push_back({v, w});
                }
            } else {
                st.pop_back();
                if (st.empty()) break;
                const int32_t p = st.back().v;
                low[p] = std::min(low[p], low[v]);
                if (low[v] >= disc[p]) {
                    blockv.clear();
                    while (!estack.empty()) {
                        const auto e = estack.back();
                        estack.pop_back();
                        blockv.push_back(e.first);
                        blockv.push_back(e.second);
                        if (e.first == p && e.second == v) break;
                    }
                    emit(blockv);
                }
            }
        }
    }
    for (int32_t x = 0; x < V; ++x)
        if (!covered[x]) {
            blockv.assign(1, x);
            emit(blockv);
        }
    *n_mem = nm;
    *n_blocks = nbk;
    return overflow ? 2 : 0;
}

// ---------------- textbook checkers ----------------

extern "C" int orc_bfs_textbook(int32_t V, const int32_t* fb, const int32_t* fc, int32_t src, int32_t* level) {
    if (V < 0 || (V > 0 && (!fb || !level))) return 1;
    for (int32_t i = 0; i < V; ++i) level[i] = -1;
    if (src < 0 || src >= V) return 0;
    std::vector<int32_t> q{src};
    level[src] = 0;
    for (size_t h = 0; h < q.size(); ++h) {
        const int32_t u = q[h];
        for (int32_t e = fb[u]; e < fb[u + 1]; ++e)
            if (level[fc[e]] == -1) {
                level[fc[e]] = level[u] + 1;
                q.push_back(fc[e]);
            }
    }
    return 0;
}

// ParallelBFS::testParallelBFS (gir_test.cpp:143-343): hop levels from src
// restricted to the vertices whose component label equals src's label
// (TDStep/BUStep test nodeToComponent.at(v) == source_comp_id,
// gir_test.cpp:158, 200).  directed: out-arcs (fw) with SCC labels;
// undirected: the symmetric closure (fw and bw rows) with WCC labels.  The
// order of a level does not change its set, so a plain queue BFS restates
// it.  *visited = vertices with a level.
extern "C" int orc_bfs_component(int32_t V, const int32_t* fb, const int32_t* fc, const int32_t* bb,
                                 const int32_t* bc, const int32_t* label, int32_t undirected, int32_t src,
                                 int32_t* level, int64_t* visited) {
    if (V < 0 || (V > 0 && (!fb || !label || !level)) || (undirected && V > 0 && !bb)) return 1;
    for (int32_t i = 0; i < V; ++i) level[i] = -1;
    if (visited) *visited = 0;
    if (src < 0 || src >= V) return 0;
    const int32_t c = label[src];
    std::vector<int32_t> q{src};
    level[src] = 0;
    auto visit = [&](int32_t u, int32_t w) {
        if (label[w] == c && level[w] == -1) {
            level[w] = level[u] + 1;
            q.push_back(w);
        }
    };
    for (size_t h = 0; h < q.size(); ++h) {
        const int32_t u = q[h];
        for (int32_t e = fb[u]; e < fb[u + 1]; ++e) visit(u, fc[e]);
        if (undirected)
            for (int32_t e = bb[u]; e < bb[u + 1]; ++e) visit(u, bc[e]);
    }
    if (visited) *visited = (int64_t)q.size();
    return 0;
}

extern "C" int orc_dijkstra(int32_t V, const int32_t* fb, const int32_t* fc, const uint32_t* w_fw, int32_t src,
                            uint64_t* dist) {
    if (V < 0 || (V > 0 && (!fb || !dist))) return 1;
    for (int32_t i = 0; i < V; ++i) dist[i] = UINT64_MAX;
    if (src < 0 || src >= V) return 0;
    using P = std::pair<uint64_t, int32_t>;
    std::priority_queue<P, std::vector<P>, std::greater<P>> pq;
    dist[src] = 0;
    pq.push({0, src});
    while (!pq.empty()) {
        auto [d, u] = pq.top();
        pq.pop();
        if (d != dist[u]) continue;
        for (int32_t e = fb[u]; e < fb[u + 1]; ++e) {
            const int32_t w = fc[e];
            const uint64_t nd = d + (w_fw ? w_fw[e] : orc_edge_weight(u, w));
            if (nd < dist[w]) {
                dist[w] = nd;
                pq.push({nd, w});
            }
        }
    }
    return 0;
}

extern "C" int orc_wcc_bfs(int32_t V, const int32_t* fb, const int32_t* fc, const int32_t* bb, const int32_t* bc,
                           int32_t* label) {
    if (V < 0 || (V > 0 && (!fb || !bb || !label))) return 1;
    for (int32_t i = 0; i < V; ++i) label[i] = -1;
    std::vector<int32_t> q;
    for (int32_t s = 0; s < V; ++s) {
        if (label[s] != -1) continue;
        // ascending scan: s is the smallest vertex of its component
        label[s] = s;
        q.assign(1, s);
        for (size_t h = 0; h < q.size(); ++h) {
            const int32_t u = q[h];
            for (int32_t e = fb[u]; e < fb[u + 1]; ++e)
                if (label[fc[e]] == -1) { label[fc[e]] = s; q.push_back(fc[e]); }
            for (int32_t e = bb[u]; e < bb[u + 1]; ++e)
                if (label[bc[e]] == -1) { label[bc[e]] = s; q.push_back(bc[e]); }
        }
    }
    return 0;
}

// Kosaraju: finish order of a DFS on the forward graph, then DFS on the
// reverse graph in decreasing finish order.
extern "C" int orc_scc_kosaraju(int32_t V, const int32_t* fb, const int32_t* fc, const int32_t* bb,
                                const int32_t* bc, int32_t* scc) {
    if (V < 0 || (V > 0 && (!fb || !bb || !scc))) return 1;
    std::vector<char> seen(V, 0);
    std::vector<int32_t> order;
    order.reserve(V);
    std::vector<std::pair<int32_t, int32_t>> st;
    for (int32_t r = 0; r < V; ++r) {
        if (seen[r]) continue;
        seen[r] = 1;
        st.push_back({r, fb[r]});
        while (!st.empty()) {
            auto& [v, e] = st.back();
            if (e < fb[v + 1]) {
                const int32_t w = fc[e++];
                if (!seen[w]) {
                    seen[w] = 1;
                    st.push_back({w, fb[w]});
                }
            } else {
                order.push_back(v);
                st.pop_back();
            }
        }
    }
    std::vector<int32_t> raw(V, -1), q;
    for (auto it = order.rbegin(); it != order.rend(); ++it) {
        if (raw[*it] != -1) continue;
        const int32_t id = *it;
        raw[id] = id;
        q.assign(1, id);
        for (size_t h = 0; h < q.size(); ++h) {
            const int32_t u = q[h];
            for (int32_t e = bb[u]; e < bb[u + 1]; ++e)
                if (raw[bc[e]] == -1) { raw[bc[e]] = id; q.push_back(bc[e]); }
        }
    }
    canonicalise(V, raw.data(), scc, nullptr);
    return 0;
}

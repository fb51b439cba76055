// Host-side input formats: seeded RMAT / lattice / uniform arc generators,
// the canonical CSR builder and the reference's four-file .bin CSR layout.
//
// The CSR builder follows the reference's el_to_csr (el_to_csr.cpp:96-116):
// self loops and duplicate arcs are dropped and every row is sorted, so the
// result is already in the form GIR::normalizeBaseCsr (gir.cpp:95-130) would
// produce.  graph::fromEdges (graph.cpp:123-136) keeps duplicates and leaves
// rows unsorted; GIR then normalises them on load, so both routes end in the
// same canonical CSR.
#include "girgen.h"

#include <omp.h>

#include <algorithm>
#include <atomic>
#include <cstdio>
#include <cstring>
#include <string>
#include <sys/stat.h>
#include <vector>

namespace {

thread_local std::string g_err;

int fail(const std::string& m) {
    g_err = m;
    return 1;
}

// SplitMix64 finaliser: a bijective 64-bit mixer; (seed, counter) -> uniform bits.
inline uint64_t mix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ULL;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
    return x ^ (x >> 31);
}
inline uint64_t draw(uint64_t seed, uint64_t stream, uint64_t i) {
    return mix64(mix64(seed * 0x2545F4914F6CDD1DULL + stream) ^ i);
}
// Uniform integer in [0, n) from 32 random bits (multiply-high).
inline uint32_t below(uint32_t r, uint32_t n) { return (uint32_t)(((uint64_t)r * n) >> 32); }
// Probability p as a 32-bit threshold: r < thr(p) happens with probability p.
inline uint64_t thr(double p) {
    if (p <= 0) return 0;
    if (p >= 1) return 1ULL << 32;
    return (uint64_t)(p * 4294967296.0);
}

}  // namespace

extern "C" const char* girgen_last_error(void) { return g_err.c_str(); }

extern "C" int girgen_rmat_edges(int32_t scale, int64_t n_samples, double a, double b, double c, uint64_t seed,
                                 int32_t* src, int32_t* dst) {
    if (scale < 1 || scale > 30) return fail("rmat: scale must be in [1, 30]");
    if (n_samples < 0) return fail("rmat: negative sample count");
    if (a < 0 || b < 0 || c < 0 || a + b + c > 1.0 + 1e-12) return fail("rmat: bad quadrant probabilities");
    if (n_samples > 0 && (!src || !dst)) return fail("rmat: null output");
    const uint64_t ta = thr(a), tb = thr(a + b), tc = thr(a + b + c);
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n_samples; ++i) {
        uint32_t u = 0, v = 0;
        uint64_t r = 0;
        for (int l = 0; l < scale; ++l) {
            // two 32-bit uniforms per 64-bit draw
            uint32_t x;
            if ((l & 1) == 0) {
                r = draw(seed, 1, ((uint64_t)i << 5) | (uint64_t)(l >> 1));
                x = (uint32_t)r;
            } else {
                x = (uint32_t)(r >> 32);
            }
            uint32_t ub, vb;
            if (x < ta) { ub = 0; vb = 0; }
            else if (x < tb) { ub = 0; vb = 1; }
            else if (x < tc) { ub = 1; vb = 0; }
            else { ub = 1; vb = 1; }
            u = (u << 1) | ub;
            v = (v << 1) | vb;
        }
        src[i] = (int32_t)u;
        dst[i] = (int32_t)v;
    }
    return 0;
}

extern "C" int girgen_grid_edges(int32_t rows, int32_t cols, double p_keep, double p_bidir, uint64_t seed,
                                 int32_t* src, int32_t* dst, int64_t* n_out) {
    if (rows < 1 || cols < 1 || (int64_t)rows * cols > INT32_MAX) return fail("grid: bad shape");
    if (!n_out) return fail("grid: null count");
    // Lattice edge ids: horizontal (r, c)-(r, c+1) first, then vertical.
    const int64_t nh = (int64_t)rows * (cols - 1), nv = (int64_t)(rows - 1) * cols, ne = nh + nv;
    if (!src || !dst) {
        *n_out = 2 * ne;
        return 0;
    }
    const uint64_t tk = thr(p_keep), tbi = thr(p_bidir);
    // per-edge arc count, then prefix offsets (deterministic order)
    std::vector<uint8_t> k(ne);
    #pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < ne; ++e) {
        uint64_t r = draw(seed, 2, (uint64_t)e);
        uint32_t x0 = (uint32_t)r, x1 = (uint32_t)(r >> 32);
        if (x0 >= tk) { k[e] = 0; continue; }
        if (x1 < tbi) { k[e] = 3; continue; }           // both directions
        uint64_t r2 = draw(seed, 3, (uint64_t)e);
        k[e] = (r2 & 1) ? 2 : 1;                          // 1: low->high, 2: high->low
    }
    const int T = omp_get_max_threads();
    std::vector<int64_t> part(T + 1, 0);
    #pragma omp parallel num_threads(T)
    {
        int t = omp_get_thread_num(), nt = omp_get_num_threads();
        int64_t lo = ne * t / nt, hi = ne * (t + 1) / nt, cnt = 0;
        for (int64_t e = lo; e < hi; ++e) cnt += k[e] == 3 ? 2 : (k[e] ? 1 : 0);
        part[t + 1] = cnt;
        #pragma omp barrier
        #pragma omp single
        for (int i = 0; i < nt; ++i) part[i + 1] += part[i];
        int64_t o = part[t];
        for (int64_t e = lo; e < hi; ++e) {
            if (!k[e]) continue;
            int32_t x, y;
            if (e < nh) {
                int64_t r = e / (cols - 1), c = e % (cols - 1);
                x = (int32_t)(r * cols + c);
                y = x + 1;
            } else {
                int64_t f = e - nh, r = f / cols, c = f % cols;
                x = (int32_t)(r * cols + c);
                y = x + cols;
            }
            if (k[e] & 1) { src[o] = x; dst[o] = y; ++o; }
            if (k[e] & 2) { src[o] = y; dst[o] = x; ++o; }
        }
    }
    *n_out = part[T];
    return 0;
}

extern "C" int girgen_uniform_edges(int32_t V, int64_t n_arcs, uint64_t seed, int32_t* src, int32_t* dst) {
    if (V < 1) return fail("uniform: V < 1");
    if (n_arcs > 0 && (!src || !dst)) return fail("uniform: null output");
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n_arcs; ++i) {
        uint64_t r = draw(seed, 4, (uint64_t)i);
        src[i] = (int32_t)below((uint32_t)r, (uint32_t)V);
        dst[i] = (int32_t)below((uint32_t)(r >> 32), (uint32_t)V);
    }
    return 0;
}

namespace {
// Rows of (key row -> neighbours), sorted and deduplicated, written into
// beg/adj.  cnt is V+1 scratch.
void bucketRows(int32_t V, int64_t n, const int32_t* row, const int32_t* col, const uint8_t* keep,
                int32_t* beg, int32_t* adj) {
    std::vector<std::atomic<int64_t>> cnt(V + 1);
    #pragma omp parallel for schedule(static)
    for (int64_t v = 0; v <= V; ++v) cnt[v].store(0, std::memory_order_relaxed);
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i)
        if (keep[i]) cnt[row[i]].fetch_add(1, std::memory_order_relaxed);
    std::vector<int64_t> off(V + 1, 0);
    for (int32_t v = 0; v < V; ++v) off[v + 1] = off[v] + cnt[v].load(std::memory_order_relaxed);
    #pragma omp parallel for schedule(static)
    for (int64_t v = 0; v < V; ++v) cnt[v].store(off[v], std::memory_order_relaxed);
    std::vector<int32_t> tmp(off[V] > 0 ? off[V] : 1);
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i)
        if (keep[i]) tmp[cnt[row[i]].fetch_add(1, std::memory_order_relaxed)] = col[i];
    std::vector<int32_t> deg(V);
    #pragma omp parallel for schedule(dynamic, 4096)
    for (int32_t v = 0; v < V; ++v) {
        auto b = tmp.begin() + off[v], e = tmp.begin() + off[v + 1];
        std::sort(b, e);
        deg[v] = (int32_t)(std::unique(b, e) - b);
    }
    beg[0] = 0;
    for (int32_t v = 0; v < V; ++v) beg[v + 1] = beg[v] + deg[v];
    #pragma omp parallel for schedule(dynamic, 4096)
    for (int32_t v = 0; v < V; ++v) std::copy(tmp.begin() + off[v], tmp.begin() + off[v] + deg[v], adj + beg[v]);
}
}  // namespace

extern "C" int girgen_csr_from_edges(int32_t V, int64_t n, const int32_t* src, const int32_t* dst,
                                     int32_t* fw_beg, int32_t* fw_adj, int32_t* bw_beg, int32_t* bw_adj,
                                     int64_t* E_out) {
    if (V < 0) return fail("csr: V < 0");
    if (n < 0) return fail("csr: n < 0");
    if (!fw_beg || !bw_beg || !E_out || (n > 0 && (!src || !dst || !fw_adj || !bw_adj)))
        return fail("csr: null argument");
    int bad = 0;
    #pragma omp parallel for schedule(static) reduction(| : bad)
    for (int64_t i = 0; i < n; ++i)
        bad |= (src[i] < 0 || src[i] >= V || dst[i] < 0 || dst[i] >= V);
    if (bad) return fail("csr: arc endpoint outside [0, V)");
    std::vector<uint8_t> keep(n > 0 ? n : 1);
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) keep[i] = src[i] != dst[i];
    bucketRows(V, n, src, dst, keep.data(), fw_beg, fw_adj);
    const int64_t E = fw_beg[V];
    if (E >= INT32_MAX) return fail("csr: too many arcs for int32 offsets");
    // bw = transpose of the (deduplicated) fw CSR
    std::vector<int32_t> s2(E > 0 ? E : 1);
    #pragma omp parallel for schedule(dynamic, 4096)
    for (int32_t u = 0; u < V; ++u)
        for (int32_t i = fw_beg[u]; i < fw_beg[u + 1]; ++i) s2[i] = u;
    std::vector<uint8_t> k2(E > 0 ? E : 1, 1);
    bucketRows(V, E, fw_adj, s2.data(), k2.data(), bw_beg, bw_adj);
    *E_out = E;
    return 0;
}

extern "C" int girgen_weights(int64_t E, uint32_t lo, uint32_t hi, uint64_t seed, uint32_t* w) {
    if (E < 0 || lo > hi || lo == 0) return fail("weights: need 0 < lo <= hi");
    if (E > 0 && !w) return fail("weights: null output");
    const uint64_t span = (uint64_t)hi - lo + 1;
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < E; ++i) {
        uint64_t r = draw(seed, 5, (uint64_t)i);
        w[i] = lo + (uint32_t)(((r >> 32) * span) >> 32);
    }
    return 0;
}

extern "C" int girgen_sources(int32_t V, const int32_t* fw_beg, int32_t k, uint64_t seed, int32_t* out) {
    if (V < 1 || !fw_beg || k < 0 || (k > 0 && !out)) return fail("sources: bad argument");
    if (k == 0) return 0;
    if (fw_beg[V] == 0) return fail("sources: graph has no arcs");
    int32_t got = 0;
    for (uint64_t j = 0; got < k; ++j) {
        int32_t s = (int32_t)below((uint32_t)(draw(seed, 6, j) >> 32), (uint32_t)V);
        if (fw_beg[s + 1] > fw_beg[s]) out[got++] = s;
    }
    return 0;
}

namespace {
bool writeFile(const std::string& p, const void* data, size_t bytes) {
    FILE* f = std::fopen(p.c_str(), "wb");
    if (!f) return false;
    bool ok = bytes == 0 || std::fwrite(data, 1, bytes, f) == bytes;
    return std::fclose(f) == 0 && ok;
}
long long fileSize(const std::string& p) {
    struct stat st;
    return stat(p.c_str(), &st) == 0 ? (long long)st.st_size : -1;
}
bool readFile(const std::string& p, void* data, size_t bytes) {
    FILE* f = std::fopen(p.c_str(), "rb");
    if (!f) return false;
    bool ok = bytes == 0 || std::fread(data, 1, bytes, f) == bytes;
    std::fclose(f);
    return ok;
}
}  // namespace

extern "C" int girgen_write_csr(const char* dir, int32_t V, int64_t E, const int32_t* fw_beg, const int32_t* fw_adj,
                                const int32_t* bw_beg, const int32_t* bw_adj) {
    if (!dir || V < 0 || E < 0) return fail("write_csr: bad argument");
    std::string d(dir);
    if (!d.empty() && d.back() != '/') d += '/';
    if (!writeFile(d + "fw_begin.bin", fw_beg, sizeof(int32_t) * (V + 1)) ||
        !writeFile(d + "fw_adjacent.bin", fw_adj, sizeof(int32_t) * E) ||
        !writeFile(d + "bw_begin.bin", bw_beg, sizeof(int32_t) * (V + 1)) ||
        !writeFile(d + "bw_adjacent.bin", bw_adj, sizeof(int32_t) * E))
        return fail("write_csr: cannot write files under " + d);
    return 0;
}

extern "C" int girgen_read_csr(const char* dir, int32_t* V_out, int64_t* E_out, int32_t* fw_beg, int32_t* fw_adj,
                               int32_t* bw_beg, int32_t* bw_adj) {
    if (!dir || !V_out || !E_out) return fail("read_csr: bad argument");
    std::string d(dir);
    if (!d.empty() && d.back() != '/') d += '/';
    long long fb = fileSize(d + "fw_begin.bin"), fa = fileSize(d + "fw_adjacent.bin");
    long long bb = fileSize(d + "bw_begin.bin"), ba = fileSize(d + "bw_adjacent.bin");
    if (fb < 4 || fa < 0 || bb < 0 || ba < 0) return fail("read_csr: missing or empty CSR files under " + d);
    if (fb % 4 || fa % 4 || bb != fb || ba != fa)
        return fail("read_csr: forward and backward CSR files disagree in size under " + d);
    const long long V = fb / 4 - 1, E = fa / 4;
    if (V > INT32_MAX || E >= INT32_MAX) return fail("read_csr: too large for int32");
    *V_out = (int32_t)V;
    *E_out = E;
    if (!fw_beg) return 0;
    if (!readFile(d + "fw_begin.bin", fw_beg, 4 * (V + 1)) || !readFile(d + "fw_adjacent.bin", fw_adj, 4 * E) ||
        !readFile(d + "bw_begin.bin", bw_beg, 4 * (V + 1)) || !readFile(d + "bw_adjacent.bin", bw_adj, 4 * E))
        return fail("read_csr: short read under " + d);
    auto ok = [&](const int32_t* beg, const int32_t* adj) {
        if (beg[0] != 0 || beg[V] != E) return false;
        for (long long i = 0; i < V; ++i)
            if (beg[i] > beg[i + 1]) return false;
        for (long long i = 0; i < E; ++i)
            if (adj[i] < 0 || adj[i] >= V) return false;
        return true;
    };
    if (!ok(fw_beg, fw_adj) || !ok(bw_beg, bw_adj))
        return fail("read_csr: malformed CSR (offsets not monotone from 0 to E, or ids outside [0, V))");
    return 0;
}

// ---------------------------------------------------------------------------
// Streaming protocol inputs (the file set bench_main.cpp:280-345 exports and
// experiments/baselines/common/protocol.h reads).

namespace {
// Text writer: "u v\n" / "a u v\n" lines formatted into a large buffer.
struct TextOut {
    FILE* f = nullptr;
    std::vector<char> buf;
    size_t n = 0;
    bool ok = true;
    explicit TextOut(const std::string& p) : f(std::fopen(p.c_str(), "wb")), buf(1 << 22) { ok = f != nullptr; }
    ~TextOut() { close(); }
    bool close() {
        if (f) {
            flush();
            ok = (std::fclose(f) == 0) && ok;
            f = nullptr;
        }
        return ok;
    }
    void flush() {
        if (f && n && std::fwrite(buf.data(), 1, n, f) != n) ok = false;
        n = 0;
    }
    void num(uint32_t x) {
        char t[12];
        int k = 0;
        do {
            t[k++] = (char)('0' + x % 10);
            x /= 10;
        } while (x);
        while (k) buf[n++] = t[--k];
    }
    void line(char op, uint32_t u, uint32_t v) {
        if (n + 32 > buf.size()) flush();
        if (op) {
            buf[n++] = op;
            buf[n++] = ' ';
        }
        num(u);
        buf[n++] = ' ';
        num(v);
        buf[n++] = '\n';
    }
};
}  // namespace

extern "C" int girgen_export_protocol(const char* dir, int32_t V, const int32_t* fw_beg, const int32_t* fw_adj,
                                      int64_t k_ops, int32_t n_sources, uint64_t seed, int32_t write_full) {
    if (!dir || V < 1 || !fw_beg || k_ops < 0 || n_sources < 0) return fail("export_protocol: bad argument");
    const int64_t E = fw_beg[V];
    if (E < 4) return fail("export_protocol: too few arcs");
    std::string d(dir);
    if (!d.empty() && d.back() != '/') d += '/';
    // every arc once, then a seeded Fisher-Yates shuffle (counter-based draws)
    std::vector<uint64_t> all((size_t)E);
    for (int32_t u = 0; u < V; ++u)
        for (int64_t i = fw_beg[u]; i < fw_beg[u + 1]; ++i) all[(size_t)i] = ((uint64_t)(uint32_t)u << 32) | (uint32_t)fw_adj[i];
    for (int64_t i = E - 1; i > 0; --i) {
        const int64_t j = (int64_t)(draw(seed, 7, (uint64_t)i) % (uint64_t)(i + 1));
        std::swap(all[(size_t)i], all[(size_t)j]);
    }
    const int64_t half = E / 2;
    const int64_t k = std::min<int64_t>(k_ops, half / 2);
    auto U = [&](int64_t i) { return (uint32_t)(all[(size_t)i] >> 32); };
    auto W = [&](int64_t i) { return (uint32_t)all[(size_t)i]; };
    {
        TextOut o(d + "loaded.el");
        for (int64_t i = 0; i < half; ++i) o.line(0, U(i), W(i));
        if (!o.close()) return fail("export_protocol: cannot write " + d + "loaded.el");
        if (write_full) {
            TextOut f(d + "full.el");
            for (int64_t i = 0; i < E; ++i) f.line(0, U(i), W(i));
            if (!f.close()) return fail("export_protocol: cannot write " + d + "full.el");
        }
    }
    {
        // insert-only: k held-out arcs; delete-only: k arcs of the loaded half
        TextOut a(d + "insert.ops"), b(d + "delete.ops");
        for (int64_t i = 0; i < k; ++i) {
            a.line('a', U(half + i), W(half + i));
            b.line('d', U(i), W(i));
        }
        if (!a.close() || !b.close()) return fail("export_protocol: cannot write the op streams under " + d);
    }
    {
        // mixed: k/2 further held-out inserts and k/2 further loaded deletes,
        // interleaved in a seeded shuffle
        const int64_t h2 = k / 2;
        std::vector<int64_t> m;
        m.reserve((size_t)(2 * h2));
        for (int64_t i = 0; i < h2; ++i) m.push_back(half + k + i);  // insert
        for (int64_t i = 0; i < h2; ++i) m.push_back(-(k + i) - 1);  // delete (encoded)
        for (int64_t i = (int64_t)m.size() - 1; i > 0; --i) {
            const int64_t j = (int64_t)(draw(seed, 8, (uint64_t)i) % (uint64_t)(i + 1));
            std::swap(m[(size_t)i], m[(size_t)j]);
        }
        TextOut o(d + "mixed.ops");
        for (int64_t x : m) {
            if (x >= 0) o.line('a', U(x), W(x));
            else o.line('d', U(-x - 1), W(-x - 1));
        }
        if (!o.close()) return fail("export_protocol: cannot write " + d + "mixed.ops");
    }
    // query sources over vertices with out-arcs in the full graph
    std::vector<int32_t> src((size_t)std::max(n_sources, 1));
    if (n_sources > 0 && girgen_sources(V, fw_beg, n_sources, seed ^ 0x5eed, src.data()) != 0) return 1;
    FILE* f = std::fopen((d + "sources.txt").c_str(), "w");
    if (!f) return fail("export_protocol: cannot write " + d + "sources.txt");
    for (int32_t i = 0; i < n_sources; ++i) std::fprintf(f, "%d\n", src[(size_t)i]);
    std::fclose(f);
    f = std::fopen((d + "meta.txt").c_str(), "w");
    if (!f) return fail("export_protocol: cannot write " + d + "meta.txt");
    std::fprintf(f, "V %d\nE %lld\nloaded %lld\nops_per_phase %lld\nmax_weight %d\nsymmetric 0\n", V, (long long)E,
                 (long long)half, (long long)k, 100);
    std::fclose(f);
    return 0;
}

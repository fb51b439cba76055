// Driver for the compiled reference (TEST INFRASTRUCTURE ONLY).
//
// Links the reference's own gir.cpp / gir_query.cpp / graph.cpp /
// graph_undirected.cpp, built from where they lie by oracle/Makefile into
// oracle/_ref/.  It loads a four-file .bin CSR through the reference's
// GIR(path, path) constructor (gir.cpp:38-44, graph.cpp:2-120), runs the
// hot path through the reference's public API (gir.h:98-240) and writes the
// results with canonical (minimum-vertex) component ids:
//
//   gir_ref CSR_DIR OUT_DIR [--bfs SRC.bin] [--sssp SRC.bin] [--delta D]
//           [--batch OPS.bin | --single OPS.bin] [--pairs PAIRS.bin] [--reach PAIRS.bin] [--two-level] [--bc SRC.bin]
//           [--bicc] [--no-output] [--reps R]
//
//   --reps R: the build and the BFS / SSSP passes are timed R times (one
//   "@@" line each); results are those of the last pass
//
//   wcc.bin  int32[V]   buildWCC labels (gir.cpp:593-655), canonical
//   scc.bin  int32[V]   buildSCC_DAG node_scc (gir.cpp:798-839), canonical
//   dag.bin  int32[2*k] getDAG() arcs (src, dst) in canonical ids, sorted
//   bfs.bin  int32[k*V] GIR::bfs levels per source (gir_query.cpp:108-294)
//   sssp.bin uint64[k*V] GIR::sssp distances, delta D (gir_query.cpp:695-703)
//   weight.bin uint32[k] GIR::edgeWeight of each (u, v) in PAIRS.bin
//   reach.bin uint8[k]   GIR::reachable of each (u, v) pair of --reach (gir_query.cpp:789-828)
//   with --two-level, per --bfs source: bfs2.bin int32[k*V] GIR::bfsTwoLevel
//   (gir_query.cpp:604-620) and rsccs.bin uint8[k*V] (1 where the vertex's
//   SCC is in GIR::reachableSccsPublic of the source's SCC, gir_query.cpp:56-64);
//   per --sssp source: sssp2.bin uint64[k*V] GIR::ssspTwoLevel, delta D
//   (gir_query.cpp:705-722)
//   bc.bin   double[V] GIR::bc over the sources of --bc (gir_query.cpp:728-778)
//   bicc.bin int32[2*k] (block, vertex) pairs of GIR::getBCTComponents over
//            every SCC after buildBCT (gir.cpp:2063, 4140-4149); with --bicc
//   post_{wcc,scc,dag}.bin  the same after GIR::applyBatch of OPS.bin (int32
//            triples u, v, insert) with blocks on demand, or after the ops of
//            --single applied one GIR::insertEdge / deleteEdge call each
// One JSON line per timed phase goes to stdout, prefixed "@@".
#include <sys/stat.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include <omp.h>

#include "gir.h"
#include "graph.h"

namespace {

using Clock = std::chrono::steady_clock;
double msSince(Clock::time_point t0) { return std::chrono::duration<double, std::milli>(Clock::now() - t0).count(); }

template <class T>
std::vector<T> readVec(const std::string& p) {
    std::vector<T> v;
    FILE* f = std::fopen(p.c_str(), "rb");
    if (!f) { std::fprintf(stderr, "cannot open %s\n", p.c_str()); std::exit(2); }
    std::fseek(f, 0, SEEK_END);
    long n = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    v.resize(n / sizeof(T));
    if (!v.empty() && std::fread(v.data(), sizeof(T), v.size(), f) != v.size()) { std::fprintf(stderr, "short read %s\n", p.c_str()); std::exit(2); }
    std::fclose(f);
    return v;
}
template <class T>
void writeVec(const std::string& p, const T* d, size_t n) {
    FILE* f = std::fopen(p.c_str(), "wb");
    if (!f || (n && std::fwrite(d, sizeof(T), n, f) != n)) { std::fprintf(stderr, "cannot write %s\n", p.c_str()); std::exit(2); }
    std::fclose(f);
}

// Canonical labels from an id -> group map: each group named by its minimum vertex.
std::vector<int> canon(int V, const std::vector<int>& raw) {
    int maxid = 0;
    for (int x : raw) maxid = std::max(maxid, x);
    std::vector<int> mn(maxid + 1, INT32_MAX), out(V);
    for (int v = 0; v < V; ++v) mn[raw[v]] = std::min(mn[raw[v]], v);
    for (int v = 0; v < V; ++v) out[v] = mn[raw[v]];
    return out;
}

struct Layers {
    std::vector<int> wcc, scc, dag;
    bool rank_ok = true;
};

Layers snapshot(GIR& g) {
    const int V = g.vertexCount();
    Layers L;
    std::vector<int> w(V), s(V);
    for (int v = 0; v < V; ++v) { w[v] = g.wccOf(v); s[v] = g.sccOf(v); }
    L.wcc = canon(V, w);
    L.scc = canon(V, s);
    // SCC id -> canonical id
    const auto& sccs = g.getSCCs();
    std::vector<int> idmap(sccs.size(), -1);
    for (size_t i = 0; i < sccs.size(); ++i)
        if (!sccs[i].empty()) idmap[i] = L.scc[sccs[i][0]];
    auto dag = g.getDAG();
    std::vector<std::pair<int, int>> arcs;
    for (size_t a = 0; a < dag.size(); ++a)
        for (int b : dag[a]) {
            if (idmap[a] < 0 || idmap[b] < 0) continue;
            arcs.push_back({idmap[a], idmap[b]});
            if (!(g.sccRank((int)a) < g.sccRank(b))) L.rank_ok = false;
        }
    std::sort(arcs.begin(), arcs.end());
    for (auto& p : arcs) { L.dag.push_back(p.first); L.dag.push_back(p.second); }
    return L;
}

void writeLayers(const std::string& out, const std::string& pre, const Layers& L) {
    writeVec(out + pre + "wcc.bin", L.wcc.data(), L.wcc.size());
    writeVec(out + pre + "scc.bin", L.scc.data(), L.scc.size());
    writeVec(out + pre + "dag.bin", L.dag.data(), L.dag.size());
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 3) {
        std::fprintf(stderr, "usage: gir_ref CSR_DIR OUT_DIR [--bfs F] [--sssp F] [--delta D] [--batch F] [--single F] [--pairs F] [--reach F] [--two-level] [--bc F] [--bicc] [--no-output] [--reps R]\n");
        return 2;
    }
    std::string dir = argv[1], out = argv[2];
    if (dir.back() != '/') dir += '/';
    if (out.back() != '/') out += '/';
    std::string bfs_f, sssp_f, batch_f, single_f, pairs_f, reach_f, bc_f;
    uint64_t delta = 16;
    bool write = true, two_level = false, bicc = false;
    int reps = 1;
    for (int i = 3; i < argc; ++i) {
        std::string a = argv[i];
        auto next = [&]() { if (i + 1 >= argc) { std::fprintf(stderr, "missing value for %s\n", a.c_str()); std::exit(2); } return std::string(argv[++i]); };
        if (a == "--bfs") bfs_f = next();
        else if (a == "--sssp") sssp_f = next();
        else if (a == "--delta") delta = std::strtoull(next().c_str(), nullptr, 10);
        else if (a == "--batch") batch_f = next();
        else if (a == "--single") single_f = next();
        else if (a == "--pairs") pairs_f = next();
        else if (a == "--reach") reach_f = next();
        else if (a == "--no-output") write = false;
        else if (a == "--two-level") two_level = true;
        else if (a == "--bc") bc_f = next();
        else if (a == "--bicc") bicc = true;
        else if (a == "--reps") reps = std::atoi(next().c_str());
        else { std::fprintf(stderr, "unknown option %s\n", a.c_str()); return 2; }
    }
    std::printf("@@{\"threads\": %d}\n", omp_get_max_threads());
    if (!pairs_f.empty()) {
        auto pr = readVec<int>(pairs_f);
        std::vector<uint32_t> w(pr.size() / 2);
        for (size_t i = 0; i < w.size(); ++i) w[i] = GIR::edgeWeight(pr[2 * i], pr[2 * i + 1]);
        writeVec(out + "weight.bin", w.data(), w.size());
    }
    GIR g(dir, dir);
    const int V = g.vertexCount();
    for (int r = 0; r < reps; ++r) {
        auto t0 = Clock::now();
        g.buildWCC();
        double t_wcc = msSince(t0);
        auto t1 = Clock::now();
        g.buildSCC_DAG();
        double t_scc = msSince(t1);
        std::printf("@@{\"phase\": \"build\", \"wcc_ms\": %.3f, \"scc_dag_ms\": %.3f}\n", t_wcc, t_scc);
    }
    Layers L = snapshot(g);
    std::printf("@@{\"rank_ok\": %d, \"n_dag\": %zu}\n", (int)L.rank_ok, L.dag.size() / 2);
    if (write) writeLayers(out, "", L);
    // queries: reps timed passes over the sources (results are moved out of
    // the calls, so the timing holds no copies); the last pass is written
    if (!bfs_f.empty()) {
        auto src = readVec<int>(bfs_f);
        std::vector<QueryVec<int>> res;
        for (int r = 0; r < reps; ++r) {
            res.clear();
            auto t0 = Clock::now();
            for (int s : src) res.push_back(g.bfs(s));
            std::printf("@@{\"phase\": \"bfs\", \"sources\": %zu, \"ms\": %.3f}\n", src.size(), msSince(t0));
        }
        if (write) {
            std::vector<int> lv;
            for (auto& l : res) lv.insert(lv.end(), l.begin(), l.end());
            writeVec(out + "bfs.bin", lv.data(), lv.size());
        }
    }
    if (!sssp_f.empty()) {
        auto src = readVec<int>(sssp_f);
        std::vector<QueryVec<uint64_t>> res;
        for (int r = 0; r < reps; ++r) {
            res.clear();
            auto t0 = Clock::now();
            for (int s : src) res.push_back(g.sssp(s, delta));
            std::printf("@@{\"phase\": \"sssp\", \"sources\": %zu, \"ms\": %.3f}\n", src.size(), msSince(t0));
        }
        if (write) {
            std::vector<uint64_t> dv;
            for (auto& d : res) dv.insert(dv.end(), d.begin(), d.end());
            writeVec(out + "sssp.bin", dv.data(), dv.size());
        }
    }
    if (two_level && !bfs_f.empty()) {
        auto src = readVec<int>(bfs_f);
        std::vector<int> lv;
        std::vector<uint8_t> rs;
        // the two-level BFS walks the source SCC by its block-cut forest: the
        // reference's drivers build it after the SCC layer (bench_main.cpp:164)
        g.buildBCT();
        auto t0 = Clock::now();
        long long keep = 0;
        for (int s : src) {
            auto l = g.bfsTwoLevel(s);
            if (write) lv.insert(lv.end(), l.begin(), l.end());
            else keep += l[0];
        }
        std::printf("@@{\"phase\": \"bfs2\", \"sources\": %zu, \"ms\": %.3f, \"keep\": %lld}\n", src.size(), msSince(t0), keep);
        if (write) {
            for (int s : src) {
                auto m = g.reachableSccsPublic(g.sccOf(s));
                for (int v = 0; v < V; ++v) rs.push_back((uint8_t)(m[g.sccOf(v)] != 0));
            }
            writeVec(out + "bfs2.bin", lv.data(), lv.size());
            writeVec(out + "rsccs.bin", rs.data(), rs.size());
        }
    }
    if (two_level && !sssp_f.empty()) {
        auto src = readVec<int>(sssp_f);
        std::vector<uint64_t> dv;
        auto t0 = Clock::now();
        unsigned long long keep = 0;
        for (int s : src) {
            auto d = g.ssspTwoLevel(s, delta);
            if (write) dv.insert(dv.end(), d.begin(), d.end());
            else keep += d[0];
        }
        std::printf("@@{\"phase\": \"sssp2\", \"sources\": %zu, \"ms\": %.3f, \"keep\": %llu}\n", src.size(), msSince(t0), keep);
        if (write) writeVec(out + "sssp2.bin", dv.data(), dv.size());
    }
    if (bicc) {
        auto t0 = Clock::now();
        g.buildBCT();
        std::printf("@@{\"phase\": \"bicc\", \"ms\": %.3f}\n", msSince(t0));
        std::vector<int> pairs;
        const auto& sccs = g.getSCCs();
        int nb = 0;
        for (size_t s2 = 0; s2 < sccs.size(); ++s2) {
            if (sccs[s2].empty()) continue;
            for (auto& blk : g.getBCTComponents((int)s2)) {
                for (int x : blk) {
                    pairs.push_back(nb);
                    pairs.push_back(x);
                }
                ++nb;
            }
        }
        writeVec(out + "bicc.bin", pairs.data(), pairs.size());  // (compared by the bench)
    }
    if (!bc_f.empty()) {
        auto src = readVec<int>(bc_f);
        auto t0 = Clock::now();
        auto sc = g.bc(src);
        std::printf("@@{\"phase\": \"bc\", \"sources\": %zu, \"ms\": %.3f}\n", src.size(), msSince(t0));
        writeVec(out + "bc.bin", sc.data(), sc.size());  // written even with --no-output (compared by the bench)
    }
    if (!reach_f.empty()) {
        auto pr = readVec<int>(reach_f);
        std::vector<uint8_t> r(pr.size() / 2);
        auto t0 = Clock::now();
        for (size_t i = 0; i < r.size(); ++i) r[i] = g.reachable(pr[2 * i], pr[2 * i + 1]) ? 1 : 0;
        std::printf("@@{\"phase\": \"reach\", \"pairs\": %zu, \"ms\": %.3f}\n", r.size(), msSince(t0));
        writeVec(out + "reach.bin", r.data(), r.size());  // small: written even with --no-output
    }
    if (!single_f.empty()) {
        // bench_main.cpp:116-125 (runSingle): one GIR::insertEdge /
        // deleteEdge call per op, timed per kind; the block layer on demand
        // as in the batch phase below.  The ops go to this graph, so
        // --single is not combined with --batch (the bench runs them apart).
        g.buildBCT();
        g.setBlocksOnDemand(true);
        auto ops = readVec<int>(single_f);
        double ms[2] = {0, 0};
        long long cnt[2] = {0, 0};
        for (size_t i = 0; i + 2 < ops.size(); i += 3) {
            const bool ins = ops[i + 2] != 0;
            auto t0 = Clock::now();
            if (ins) g.insertEdge(ops[i], ops[i + 1]);
            else g.deleteEdge(ops[i], ops[i + 1]);
            ms[ins ? 1 : 0] += msSince(t0);
            ++cnt[ins ? 1 : 0];
        }
        std::printf("@@{\"phase\": \"single\", \"inserts\": %lld, \"insert_ms\": %.3f, \"deletes\": %lld, "
                    "\"delete_ms\": %.3f}\n",
                    cnt[1], ms[1], cnt[0], ms[0]);
        Layers P = snapshot(g);
        if (write) writeLayers(out, "post_", P);
    }
    if (!batch_f.empty()) {
        // The reference's drivers build the block layer before any update
        // (gir_selftest.cpp:115-187, bench_main.cpp:164), then may switch
        // blocks to on-demand (gir_selftest.cpp:161, bench_main.cpp:362).
        g.buildBCT();
        g.setBlocksOnDemand(true);
        auto ops = readVec<int>(batch_f);
        std::vector<GIR::Update> b(ops.size() / 3);
        for (size_t i = 0; i < b.size(); ++i) b[i] = {ops[3 * i], ops[3 * i + 1], ops[3 * i + 2] != 0};
        auto t0 = Clock::now();
        g.applyBatch(b);
        std::printf("@@{\"phase\": \"batch\", \"ops\": %zu, \"ms\": %.3f}\n", b.size(), msSince(t0));
        Layers P = snapshot(g);
        std::printf("@@{\"post_rank_ok\": %d}\n", (int)P.rank_ok);
        if (write) writeLayers(out, "post_", P);
    }
    return 0;
}

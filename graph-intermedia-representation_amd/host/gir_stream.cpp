// Streaming caller of the GIR class on the gfx950 path, with the command
// line and result rows of the reference's experiments/baselines/gir/
// gir_stream.cpp:21-88, so its experiment scripts can run either binary:
//
//   gir_stream <protocol dir> <dataset> <csv> [batch=10000] [batches=10] [lazy|inc|lazy-inc] [--dump DIR]
//
// Per stream (insert, delete, mixed): a fresh GIR on loaded.el, buildWCC +
// buildSCC_DAG, BFS and SSSP from the first source of sources.txt (timed as
// the initial from-scratch queries), then per batch of B operations:
// applyBatch (FMU) or ingestBatch (lazy: edge store only), BFS, SSSP from
// scratch.  Rows: <stream>_mutation_MEPS, bfs_/sssp_<stream>_recompute_ms,
// bfs_/sssp_<stream>_endtoend_ms; lazy adds sync_<stream>_ms, the layers'
// catch-up after the stream.  With "inc" / "lazy-inc", BFS and SSSP are
// corrected per batch by updateIncremental instead of recomputed, and after
// the stream the corrected values are compared with from-scratch ones
// (inc_mismatch_<stream> rows).  --dump DIR writes per stream the final SCC
// and WCC labels (canonical ids) and the last BFS levels and SSSP distances
// as raw arrays <stream>_{scc,wcc,bfs,sssp}.bin.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "gir.h"
#include "graph.h"
#include "protocol.h"

namespace {

double now() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

template <class T>
void dumpArray(const std::string& path, const T* p, size_t n) {
    if (std::FILE* f = std::fopen(path.c_str(), "wb")) {
        std::fwrite(p, sizeof(T), n, f);
        std::fclose(f);
    }
}

}  // namespace

int main(int argc, char** argv) {
    std::vector<std::string> pos;
    std::string dump;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        if (a == "--dump" && i + 1 < argc) dump = argv[++i];
        else pos.push_back(a);
    }
    if (pos.size() < 3) {
        std::fprintf(stderr, "usage: gir_stream <protocol dir> <dataset> <csv> [batch=10000] [batches=10] [lazy|inc|lazy-inc] [--dump DIR]\n");
        return 2;
    }
    const std::string dir = pos[0], ds = pos[1], csv = pos[2];
    const size_t B = pos.size() > 3 ? std::stoul(pos[3]) : 10000;
    const int NB = pos.size() > 4 ? std::stoi(pos[4]) : 10;
    const std::string mode = pos.size() > 5 ? pos[5] : "";
    if (!mode.empty() && mode != "lazy" && mode != "inc" && mode != "lazy-inc") {
        std::fprintf(stderr, "gir_stream: unknown mode %s\n", mode.c_str());
        return 2;
    }
    if (B == 0 || NB <= 0) {
        std::fprintf(stderr, "gir_stream: batch size and batch count must be positive\n");
        return 2;
    }
    const bool lazy = mode == "lazy" || mode == "lazy-inc";
    const bool incr = mode == "inc" || mode == "lazy-inc";
    const std::string sys = std::string("GIR-gfx950") + (lazy ? "-lazy" : "") + (incr ? "-inc" : "");
    try {
        const int V = (int)stream_proto::metaValue(dir, "V");
        const int src = stream_proto::readInts(dir + "/sources.txt").at(0);
        const auto loaded = stream_proto::readEdges(dir + "/loaded.el");
        const int th = 1;  // one host thread drives the GPU
        for (const char* s : {"insert", "delete", "mixed"}) {
            const std::string st = s;
            GIR g(graph::fromEdges(V, loaded));
            g.buildWCC();
            g.buildSCC_DAG();
            QueryVec<int> lv;
            QueryVec<uint64_t> dist;
            const double t0 = now();
            if (incr) g.trackIncremental(src, false);
            else lv = g.bfs(src);
            const double t1 = now();
            if (incr) g.trackIncremental(src, true);
            else dist = g.sssp(src);
            const double t2 = now();
            if (st == "insert") {
                stream_proto::emitRow(csv, ds, sys, "bfs_initial_ms", "from-scratch", (t1 - t0) * 1e3, "ms", 1, th);
                stream_proto::emitRow(csv, ds, sys, "sssp_initial_ms", "from-scratch", (t2 - t1) * 1e3, "ms", 1, th);
            }
            const auto ops = stream_proto::readOps(dir + "/" + st + ".ops");
            double tu = 0, tb = 0, ts = 0;
            int nb = 0;
            std::vector<GIR::Update> batch;
            for (size_t i = 0; i < ops.size() && nb < NB; i += B, ++nb) {
                batch.clear();
                for (size_t j = i; j < ops.size() && j < i + B; ++j) batch.push_back({ops[j].u, ops[j].v, ops[j].insert});
                const double a = now();
                if (lazy) g.ingestBatch(batch);
                else g.applyBatch(batch);
                const double b = now();
                if (incr) g.updateIncremental(false);
                else lv = g.bfs(src);
                const double c = now();
                if (incr) g.updateIncremental(true);
                else dist = g.sssp(src);
                const double d = now();
                tu += b - a;
                tb += c - b;
                ts += d - c;
            }
            if (nb == 0) continue;
            if (incr) {  // the corrected values must equal from-scratch ones
                lv = g.bfs(src);
                dist = g.sssp(src);
                long long bad = 0;
                for (int v = 0; v < V; ++v) {
                    const uint64_t want = lv[v] < 0 ? UINT64_MAX : (uint64_t)lv[v];
                    bad += (g.incDist(false, v) != want || g.incDist(true, v) != dist[v]) ? 1 : 0;
                }
                std::fprintf(stderr, "%s stream: incremental vs from-scratch mismatches: %lld\n", s, bad);
                stream_proto::emitRow(csv, ds, sys, "inc_mismatch_" + st, "final", (double)bad, "vertices", V, th);
            }
            const std::string mm = (lazy ? "ingest_batch_" : "fmu_batch_") + std::to_string(B);
            const std::string bm = "batch_" + std::to_string(B);
            stream_proto::emitRow(csv, ds, sys, st + "_mutation_MEPS", mm, nb * (double)B / 1e6 / tu, "MEPS",
                                  nb * (long long)B, th);
            const std::string rm = incr ? "incremental" : "from-scratch";
            stream_proto::emitRow(csv, ds, sys, "bfs_" + st + "_recompute_ms", rm, tb * 1e3 / nb, "ms", nb, th);
            stream_proto::emitRow(csv, ds, sys, "bfs_" + st + "_endtoend_ms", bm, (tu + tb) * 1e3 / nb, "ms", nb, th);
            stream_proto::emitRow(csv, ds, sys, "sssp_" + st + "_recompute_ms", rm, ts * 1e3 / nb, "ms", nb, th);
            stream_proto::emitRow(csv, ds, sys, "sssp_" + st + "_endtoend_ms", bm, (tu + ts) * 1e3 / nb, "ms", nb, th);
            if (lazy) {
                const double a = now();
                g.sync();
                const double b = now();
                stream_proto::emitRow(csv, ds, sys, "sync_" + st + "_ms", "after_" + std::to_string(nb) + "_batches",
                                      (b - a) * 1e3, "ms", nb, th);
            }
            if (!dump.empty()) {
                std::vector<int> scc(V), wcc(V);
                for (int v = 0; v < V; ++v) {
                    scc[v] = g.sccOf(v);
                    wcc[v] = g.wccOf(v);
                }
                dumpArray(dump + "/" + st + "_scc.bin", scc.data(), scc.size());
                dumpArray(dump + "/" + st + "_wcc.bin", wcc.data(), wcc.size());
                dumpArray(dump + "/" + st + "_bfs.bin", lv.data(), lv.size());
                dumpArray(dump + "/" + st + "_sssp.bin", dist.data(), dist.size());
            }
        }
    } catch (const std::exception& e) {
        std::fprintf(stderr, "gir_stream: %s\n", e.what());
        return 1;
    }
    return 0;
}

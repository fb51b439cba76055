// Single-edge updates through the drop-in GIR class (host/gir.h), timed the
// way the reference's bench_main times its single-edge phase
// (bench_main.cpp:116-125, runSingle): one GIR::insertEdge / deleteEdge call
// per op, per kind.  The drop-in queues such ops and applies them as one
// batch when something next reads the graph or a layer, so each kind's time
// here runs until the ops are applied and every layer is current on the
// device (a GIR::synced call after the last op applies the queue; the
// per-vertex getters then pull host copies of the labels, a cost that does
// not depend on how the ops arrived).  The per-op latency of an op that is
// applied at once (op, then synced) is timed on a second object.
//   gir_updates CSR_DIR OPS.bin OUT_DIR
// OPS.bin: int32 triples (u, v, insert).  Prints one "@@{...}" JSON line;
// writes OUT_DIR/post_wcc.bin and post_scc.bin (canonical labels: minimum
// vertex of each component) after the ops.
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <string>
#include <vector>

#include "gir.h"

using Clock = std::chrono::steady_clock;
static double msSince(Clock::time_point t) { return std::chrono::duration<double, std::milli>(Clock::now() - t).count(); }

int main(int argc, char** argv) {
    if (argc < 4) {
        std::fprintf(stderr, "usage: gir_updates CSR_DIR OPS.bin OUT_DIR\n");
        return 2;
    }
    const std::string dir = argv[1], out = argv[3];
    std::vector<int> ops;
    if (FILE* f = std::fopen(argv[2], "rb")) {
        int x;
        while (std::fread(&x, sizeof(int), 1, f) == 1) ops.push_back(x);
        std::fclose(f);
    } else {
        std::fprintf(stderr, "gir_updates: cannot open %s\n", argv[2]);
        return 2;
    }
    const size_t n = ops.size() / 3;
    double ms[2] = {0, 0}, lat_ms[2] = {0, 0};
    long long cnt[2] = {0, 0};
    {
        GIR g(dir, dir);
        g.buildWCC();
        g.buildSCC_DAG();
        // each kind's run of calls, then one read (the queued ops are applied
        // there); kinds in the order they appear
        size_t i = 0;
        while (i < n) {
            const bool ins = ops[3 * i + 2] != 0;
            auto t0 = Clock::now();
            size_t j = i;
            for (; j < n && (ops[3 * j + 2] != 0) == ins; ++j) {
                if (ins) g.insertEdge(ops[3 * j], ops[3 * j + 1]);
                else g.deleteEdge(ops[3 * j], ops[3 * j + 1]);
            }
            (void)g.synced();
            ms[ins ? 1 : 0] += msSince(t0);
            cnt[ins ? 1 : 0] += (long long)(j - i);
            i = j;
        }
        auto s = g.getNodeToScc();
        auto w = g.getNodeToWcc();
        const int V = g.vertexCount();
        // canonical labels: the minimum vertex of each component
        std::vector<int> scc(V), wcc(V);
        std::vector<int> min_s(V, V), min_w(V, V);
        for (int v = 0; v < V; ++v) {
            min_s[s[v]] = std::min(min_s[s[v]], v);
            min_w[w[v]] = std::min(min_w[w[v]], v);
        }
        for (int v = 0; v < V; ++v) {
            scc[v] = min_s[s[v]];
            wcc[v] = min_w[w[v]];
        }
        if (FILE* f = std::fopen((out + "post_scc.bin").c_str(), "wb")) {
            std::fwrite(scc.data(), sizeof(int), scc.size(), f);
            std::fclose(f);
        }
        if (FILE* f = std::fopen((out + "post_wcc.bin").c_str(), "wb")) {
            std::fwrite(wcc.data(), sizeof(int), wcc.size(), f);
            std::fclose(f);
        }
    }
    {
        // latency of one op that is read at once
        GIR g(dir, dir);
        g.buildWCC();
        g.buildSCC_DAG();
        for (size_t i = 0; i < n; ++i) {
            const bool ins = ops[3 * i + 2] != 0;
            auto t0 = Clock::now();
            if (ins) g.insertEdge(ops[3 * i], ops[3 * i + 1]);
            else g.deleteEdge(ops[3 * i], ops[3 * i + 1]);
            (void)g.synced();
            lat_ms[ins ? 1 : 0] += msSince(t0);
        }
    }
    std::printf("@@{\"phase\": \"single\", \"inserts\": %lld, \"insert_ms\": %.3f, \"deletes\": %lld, "
                "\"delete_ms\": %.3f, \"insert_read_each_ms\": %.3f, \"delete_read_each_ms\": %.3f}\n",
                cnt[1], ms[1], cnt[0], ms[0], lat_ms[1], lat_ms[0]);
    return 0;
}

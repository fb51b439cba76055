// Differential self-test of the drop-in GIR class (host/gir.h) on the GPU.
//
// Modelled on the reference's gir_selftest.cpp / gir_test.cpp: seeded graph
// families, random insertEdge / deleteEdge / applyBatch / ingestBatch + sync
// sequences, and after every step a from-scratch ground truth on an
// independent edge ledger (gir_test.cpp:470-563, 570-695):
//   WCC by BFS over out+in arcs (gir_test.cpp:376-393), SCC by iterative
//   Tarjan (397-447), the condensation arc set with multiplicities
//   (518-537), ranks ordered along every arc (bench_main.cpp:64-68),
//   BFS levels by a queue BFS (865-881), SSSP by Dijkstra with
//   GIR::edgeWeight (957-972).
// Partitions are compared after normalisation (gir_test.cpp:365-374).
//   gir_selftest [iterations]      exit 0 = all checks passed
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <queue>
#include <random>
#include <set>
#include <string>
#include <vector>

#include "gir.h"

using Edges = std::set<std::pair<int, int>>;
using Part = std::vector<std::vector<int>>;

static int g_fail = 0;
#define CHECK(c, ...)                                      \
    do {                                                   \
        if (!(c)) {                                        \
            std::fprintf(stderr, "FAIL %s:%d ", __FILE__, __LINE__); \
            std::fprintf(stderr, __VA_ARGS__);             \
            std::fprintf(stderr, "\n");                    \
            ++g_fail;                                      \
        }                                                  \
    } while (0)

static std::vector<std::vector<int>> outAdj(int V, const Edges& E) {
    std::vector<std::vector<int>> a(V);
    for (auto& e : E) a[e.first].push_back(e.second);
    return a;
}

static Part normalise(Part p) {
    Part r;
    for (auto& g : p)
        if (!g.empty()) {
            std::sort(g.begin(), g.end());
            r.push_back(g);
        }
    std::sort(r.begin(), r.end());
    return r;
}

static Part truthWcc(int V, const Edges& E) {
    std::vector<std::vector<int>> a(V);
    for (auto& e : E) {
        a[e.first].push_back(e.second);
        a[e.second].push_back(e.first);
    }
    std::vector<int> c(V, -1);
    Part p;
    for (int s = 0; s < V; ++s) {
        if (c[s] != -1) continue;
        p.push_back({});
        std::vector<int> st{s};
        c[s] = (int)p.size() - 1;
        while (!st.empty()) {
            int u = st.back();
            st.pop_back();
            p.back().push_back(u);
            for (int w : a[u])
                if (c[w] == -1) {
                    c[w] = c[s];
                    st.push_back(w);
                }
        }
    }
    return p;
}

static std::vector<int> truthSccLabels(int V, const Edges& E) {
    auto a = outAdj(V, E);
    std::vector<int> idx(V, -1), low(V), lab(V, -1), stk;
    std::vector<char> on(V, 0);
    int t = 0;
    struct F {
        int v;
        size_t i;
    };
    for (int r = 0; r < V; ++r) {
        if (idx[r] != -1) continue;
        std::vector<F> cs{{r, 0}};
        idx[r] = low[r] = t++;
        stk.push_back(r);
        on[r] = 1;
        while (!cs.empty()) {
            F& f = cs.back();
            if (f.i < a[f.v].size()) {
                int w = a[f.v][f.i++];
                if (idx[w] == -1) {
                    idx[w] = low[w] = t++;
                    stk.push_back(w);
                    on[w] = 1;
                    cs.push_back({w, 0});
                } else if (on[w]) {
                    low[f.v] = std::min(low[f.v], idx[w]);
                }
            } else {
                int v = f.v;
                cs.pop_back();
                if (!cs.empty()) low[cs.back().v] = std::min(low[cs.back().v], low[v]);
                if (low[v] == idx[v]) {
                    std::vector<int> comp;
                    while (true) {
                        int x = stk.back();
                        stk.pop_back();
                        on[x] = 0;
                        comp.push_back(x);
                        if (x == v) break;
                    }
                    int m = *std::min_element(comp.begin(), comp.end());
                    for (int x : comp) lab[x] = m;
                }
            }
        }
    }
    return lab;
}

static std::vector<int> truthBfs(int V, const Edges& E, int s) {
    auto a = outAdj(V, E);
    std::vector<int> l(V, -1);
    if (s < 0 || s >= V) return l;
    std::queue<int> q;
    q.push(s);
    l[s] = 0;
    while (!q.empty()) {
        int u = q.front();
        q.pop();
        for (int w : a[u])
            if (l[w] == -1) {
                l[w] = l[u] + 1;
                q.push(w);
            }
    }
    return l;
}

static std::vector<uint64_t> truthDijkstra(int V, const Edges& E, int s) {
    auto a = outAdj(V, E);
    std::vector<uint64_t> d(V, UINT64_MAX);
    if (s < 0 || s >= V) return d;
    using P = std::pair<uint64_t, int>;
    std::priority_queue<P, std::vector<P>, std::greater<P>> pq;
    d[s] = 0;
    pq.push({0, s});
    while (!pq.empty()) {
        auto [du, u] = pq.top();
        pq.pop();
        if (du != d[u]) continue;
        for (int w : a[u]) {
            uint64_t nd = du + GIR::edgeWeight(u, w);
            if (nd < d[w]) {
                d[w] = nd;
                pq.push({nd, w});
            }
        }
    }
    return d;
}

static void checkConsistency(GIR& g, int V, const Edges& E, const char* what) {
    // WCC partition
    CHECK(normalise(g.getWCCs()) == normalise(truthWcc(V, E)), "%s: WCC partition", what);
    // SCC ids: dense, in order of each SCC's minimum vertex; getSCCs() has
    // one entry per SCC and is indexed by sccOf
    auto lab = truthSccLabels(V, E);  // canonical (minimum vertex)
    std::vector<int> dense(V, -1);
    int ns = 0;
    for (int v = 0; v < V; ++v)
        if (lab[v] == v) dense[v] = ns++;
    const auto& sccs = g.getSCCs();
    CHECK((int)sccs.size() == ns, "%s: getSCCs().size() = %zu, want %d", what, sccs.size(), ns);
    auto n2s = g.getNodeToScc();
    for (int v = 0; v < V; ++v) {
        const int want_id = dense[lab[v]];
        CHECK(g.sccOf(v) == want_id, "%s: sccOf(%d) = %d, want %d", what, v, g.sccOf(v), want_id);
        CHECK(n2s[v] == want_id, "%s: getNodeToScc()[%d]", what, v);
        if (g.sccOf(v) < 0 || g.sccOf(v) >= (int)sccs.size()) continue;
        const auto& grp = sccs[g.sccOf(v)];
        CHECK(std::binary_search(grp.begin(), grp.end(), v), "%s: %d not in getSCCs()[sccOf]", what, v);
    }
    auto n2w = g.getNodeToWcc();
    for (int v = 0; v < V; ++v) CHECK(n2w[v] == g.wccOf(v), "%s: getNodeToWcc()[%d]", what, v);
    // DAG arcs with multiplicities (getDAGWithCounts) and without (getDAG)
    std::map<std::pair<int, int>, int> want;
    for (auto& e : E)
        if (lab[e.first] != lab[e.second]) want[{dense[lab[e.first]], dense[lab[e.second]]}]++;
    {
        auto plain = g.getDAG();
        std::set<std::pair<int, int>> got_plain, want_plain;
        for (int a = 0; a < (int)plain.size(); ++a)
            for (int b : plain[a]) got_plain.insert({a, b});
        for (auto& kv : want) want_plain.insert(kv.first);
        CHECK(got_plain == want_plain && (int)plain.size() == ns, "%s: getDAG", what);
    }
    std::map<std::pair<int, int>, int> got;
    auto dag = g.getDAGWithCounts();
    for (int a = 0; a < (int)dag.size(); ++a)
        for (auto& p : dag[a]) {
            got[{a, p.first}] = p.second;
            CHECK(g.sccRank(a) < g.sccRank(p.first), "%s: rank order %d -> %d", what, a, p.first);
        }
    CHECK(got == want, "%s: DAG arcs (%zu vs %zu)", what, got.size(), want.size());
    // edge store
    for (int v = 0; v < V; v += std::max(1, V / 16)) {
        std::vector<int> o;
        for (auto it = E.lower_bound({v, -1}); it != E.end() && it->first == v; ++it) o.push_back(it->second);
        CHECK(g.getDirectedOutNeighbors(v) == o, "%s: out-neighbours of %d", what, v);
    }
}

static void checkQueries(GIR& g, int V, const Edges& E, std::mt19937& rng, const char* what) {
    for (int t = 0; t < 3; ++t) {
        int s = (int)(rng() % V);
        auto l = g.bfs(s);
        auto want = truthBfs(V, E, s);
        CHECK(std::equal(l.begin(), l.end(), want.begin()), "%s: bfs(%d)", what, s);
        auto d = g.sssp(s);
        auto wd = truthDijkstra(V, E, s);
        CHECK(std::equal(d.begin(), d.end(), wd.begin()), "%s: sssp(%d)", what, s);
        // reachable(s, x) (gir_query.cpp:789-828) == x has a BFS level
        std::vector<int> us(V, s), vs(V);
        for (int x = 0; x < V; ++x) vs[x] = x;
        auto r = g.reachableMany(us, vs);
        bool ok = true;
        for (int x = 0; x < V; ++x) ok = ok && ((r[x] != 0) == (want[x] >= 0));
        CHECK(ok, "%s: reachable(%d, *)", what, s);
        // the harness: BFS inside s's SCC reaches exactly the SCC
        long long vis = 0;
        auto lc = g.bfsInComponent(s, true, &vis);
        const int c = g.sccOf(s);
        long long members = 0;
        bool inside = true;
        for (int x = 0; x < V; ++x) {
            members += g.sccOf(x) == c;
            inside = inside && ((lc[x] >= 0) == (g.sccOf(x) == c));
        }
        CHECK(inside && vis == members, "%s: bfsInComponent(%d)", what, s);
    }
    auto l = g.bfs(V + 5);  // out of range: all unreachable (gir_query.cpp:112)
    CHECK(std::all_of(l.begin(), l.end(), [](int x) { return x == -1; }), "%s: bfs out of range", what);
}

static GIR* makeGir(int V, const Edges& E) {
    std::vector<std::pair<int, int>> el(E.begin(), E.end());
    std::shuffle(el.begin(), el.end(), std::mt19937(5));  // unsorted rows: GIR normalises
    return new GIR(graph::fromEdges(V, el));
}

static void runDifferential(const char* name, int V, Edges E, int iters, unsigned seed) {
    std::mt19937 rng(seed);
    GIR* g = makeGir(V, E);
    g->buildWCC();
    g->buildSCC_DAG();
    checkConsistency(*g, V, E, name);
    checkQueries(*g, V, E, rng, name);
    for (int it = 0; it < iters; ++it) {
        const int mode = it % 4;
        if (mode == 0 || mode == 1) {  // single-edge updates (SPU path)
            int u = (int)(rng() % V), v = (int)(rng() % V);
            bool ins = (rng() % 2) == 0;
            if (!ins && !E.empty() && rng() % 2) {
                auto itE = E.begin();
                std::advance(itE, rng() % E.size());
                u = itE->first;
                v = itE->second;
            }
            if (ins) g->insertEdge(u, v);
            else g->deleteEdge(u, v);
            if (u != v) {
                if (ins) E.insert({u, v});
                else E.erase({u, v});
            }
        } else {  // batches with duplicates and cancelling pairs (FMU path)
            std::vector<GIR::Update> b;
            int n = 1 + (int)(rng() % 40);
            for (int k = 0; k < n; ++k) {
                int u = (int)(rng() % V), v = (int)(rng() % V);
                b.push_back({u, v, (rng() % 3) != 0});
                if (rng() % 4 == 0) b.push_back({u, v, !b.back().insert});
            }
            if (mode == 2) g->applyBatch(b);
            else {
                g->ingestBatch(b);
                g->sync();
            }
            for (auto& op : b)
                if (op.u != op.v && op.u >= 0 && op.v >= 0 && op.u < V && op.v < V) {
                    if (op.insert) E.insert({op.u, op.v});
                    else E.erase({op.u, op.v});
                }
        }
        char tag[128];
        std::snprintf(tag, sizeof tag, "%s it %d", name, it);
        checkConsistency(*g, V, E, tag);
        if (it % 8 == 0) checkQueries(*g, V, E, rng, tag);
        if (g_fail > 20) break;
    }
    delete g;
}

// GIR(path, path) on the reference's data/example_graph four-file CSR
// (graph.cpp:2-120, gir.cpp:38-44), checked against the hand-derived answers:
// 2 WCCs, 7 SCCs, the condensation arcs with multiplicities, BFS levels.
static void runOnDiskExample(const std::string& dir) {
    GIR g(dir, dir);
    CHECK(g.vertexCount() == 18, "on-disk example: V = %d", g.vertexCount());
    g.buildWCC();
    g.buildSCC_DAG();
    auto n2w = g.getNodeToWcc();
    for (int v = 0; v < 18; ++v) CHECK(n2w[v] == (v < 12 ? 0 : 12), "on-disk example: wcc of %d", v);
    CHECK(g.getWCCs().size() == 2, "on-disk example: 2 WCCs");
    // canonical SCC labels 0,0,0,3,3,5,6,6,6,0,0,0,12,12,12,12,16,17 -> dense ids
    const int want_scc[18] = {0, 0, 0, 1, 1, 2, 3, 3, 3, 0, 0, 0, 4, 4, 4, 4, 5, 6};
    auto n2s = g.getNodeToScc();
    for (int v = 0; v < 18; ++v) CHECK(n2s[v] == want_scc[v], "on-disk example: scc of %d = %d", v, n2s[v]);
    CHECK(g.getSCCs().size() == 7, "on-disk example: 7 SCCs");
    std::map<std::pair<int, int>, int> want = {{{0, 1}, 3}, {{0, 2}, 1}, {{1, 2}, 1}, {{1, 3}, 1},
                                               {{2, 3}, 1}, {{4, 5}, 1}, {{5, 6}, 1}};
    std::map<std::pair<int, int>, int> got;
    auto dag = g.getDAGWithCounts();
    for (int a = 0; a < (int)dag.size(); ++a)
        for (auto& p : dag[a]) got[{a, p.first}] = p.second;
    CHECK(got == want, "on-disk example: DAG arcs with multiplicities");
    auto plain = g.getDAG();
    CHECK(plain.size() == 7 && plain[0] == std::vector<int>({1, 2}) && plain[4] == std::vector<int>({5}),
          "on-disk example: getDAG");
    const int from9[18] = {2, 3, 1, 2, 3, 1, 2, 3, 4, 0, 1, 2, -1, -1, -1, -1, -1, -1};
    auto l = g.bfs(9);
    CHECK(std::equal(l.begin(), l.end(), from9), "on-disk example: bfs(9)");
    // case 5b (gir_test.cpp:736-753): deleting 9->2 splits one SCC in two
    const size_t before = g.getSCCs().size();
    g.deleteEdge(9, 2);
    CHECK(g.getSCCs().size() == before + 1, "on-disk example: case 5b adds one SCC (%zu -> %zu)", before,
          g.getSCCs().size());
    CHECK(g.sccOf(0) == g.sccOf(2) && g.sccOf(9) == g.sccOf(11) && g.sccOf(0) != g.sccOf(9),
          "on-disk example: case 5b split");
}

int main(int argc, char** argv) {
    const int iters = argc > 1 ? std::atoi(argv[1]) : 24;
    {
        // tests/golden/example_graph next to this binary's tests/cpp/ folder
        std::string dir = argc > 2 ? argv[2] : std::string(argv[0]);
        if (argc <= 2) {
            const size_t sl = dir.find_last_of('/');
            dir = (sl == std::string::npos ? std::string(".") : dir.substr(0, sl)) + "/../golden/example_graph/";
        }
        if (dir.back() != '/') dir += '/';
        runOnDiskExample(dir);
    }
    // targeted test on the reference's 18-vertex example (gir_test.cpp:704-753)
    {
        const int ex[][2] = {{0, 1}, {0, 3}, {1, 2}, {1, 3}, {2, 0}, {2, 3}, {2, 9}, {3, 4}, {3, 5},
                             {3, 7}, {4, 3}, {5, 6}, {6, 7}, {7, 8}, {8, 6}, {9, 2}, {9, 5}, {9, 10},
                             {10, 11}, {11, 9}, {12, 13}, {13, 14}, {13, 16}, {14, 15}, {15, 12}, {16, 17}};
        Edges E;
        for (auto& e : ex) E.insert({e[0], e[1]});
        GIR* g = makeGir(18, E);
        g->buildWCC();
        g->buildSCC_DAG();
        CHECK(g->getWCCs().size() == 2, "example: 2 WCCs");
        CHECK(normalise(g->getSCCs()).size() == 7, "example: 7 SCCs");
        // block layer (blocks of the reference's own build on this graph)
        CHECK(normalise(g->getBCTComponents(g->sccOf(0))) ==
                  std::vector<std::vector<int>>({{0, 1, 2}, {2, 9}, {9, 10, 11}}),
              "example: blocks of SCC {0,1,2,9,10,11}");
        CHECK(normalise(g->getBCTComponents(g->sccOf(6))) == std::vector<std::vector<int>>({{6, 7, 8}}),
              "example: blocks of SCC {6,7,8}");
        CHECK(normalise(g->getBCTComponents(g->sccOf(5))) == std::vector<std::vector<int>>({{5}}),
              "example: single-vertex SCC is a block");
        CHECK(g->articulationPoint(2) && g->articulationPoint(9) && !g->articulationPoint(0) &&
                  !g->articulationPoint(3),
              "example: articulation points 2 and 9");
        CHECK(g->biconnected(0, 1) && g->biconnected(2, 9) && !g->biconnected(0, 9) && !g->biconnected(0, 3),
              "example: biconnected pairs");
        CHECK(g->blockCount() == 9, "example: 9 blocks");
        g->insertEdge(3, 12);
        CHECK(g->wccOf(3) == g->wccOf(12), "example: insert 3->12 merges WCCs");
        g->deleteEdge(3, 12);
        CHECK(g->wccOf(3) != g->wccOf(12), "example: delete 3->12 splits WCCs");
        g->deleteEdge(9, 2);
        CHECK(g->sccOf(0) == g->sccOf(2) && g->sccOf(9) == g->sccOf(11) && g->sccOf(0) != g->sccOf(9),
              "example: delete 9->2 splits {0,1,2,9,10,11}");
        delete g;
    }
    // seeded families (gir_selftest.cpp:92-104, 173-196)
    std::mt19937 rng(12345);
    auto randomGraph = [&](int V, int m) {
        Edges E;
        while ((int)E.size() < m) {
            int u = (int)(rng() % V), v = (int)(rng() % V);
            if (u != v) E.insert({u, v});
        }
        return E;
    };
    runDifferential("sparse_60", 60, randomGraph(60, 90), iters, 1);
    runDifferential("dense_30", 30, randomGraph(30, 300), iters, 2);
    {
        Edges E;  // chain of cycles
        for (int c = 0; c < 10; ++c) {
            for (int i = 0; i < 5; ++i) E.insert({c * 5 + i, c * 5 + (i + 1) % 5});
            if (c) E.insert({c * 5 - 1, c * 5});
        }
        runDifferential("cycle_chain", 50, E, iters, 3);
    }
    runDifferential("empty_20", 20, Edges{}, iters, 4);
    runDifferential("giant_2000", 2000, randomGraph(2000, 6000), std::max(4, iters / 4), 5);
    if (g_fail) {
        std::fprintf(stderr, "gir_selftest: %d check(s) failed\n", g_fail);
        return 1;
    }
    std::printf("gir_selftest: all checks passed\n");
    return 0;
}

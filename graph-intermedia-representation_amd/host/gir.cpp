// Drop-in GIR front end over the gfx950 C ABI.  See gir.h.
#include "gir.h"

#include <sys/stat.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <stdexcept>
#include <unordered_map>

#include "girhip.h"

namespace {

// A failure the reference cannot have (it has no device): report and throw.
void check(int rc, const char* what) {
    if (rc != GIRHIP_OK)
        throw std::runtime_error(std::string("GIR (gfx950): ") + what + " failed: " + girhip_last_error());
}

long long fsizeOf(const char* f) {
    struct stat st;
    return stat(f, &st) == 0 ? (long long)st.st_size : -1;
}

int* readInts(const char* f, long long n) {
    FILE* fp = std::fopen(f, "rb");
    if (!fp) {
        std::cout << f << " cannot open\n";
        std::exit(-1);  // graph.cpp:21
    }
    int* p = new int[n > 0 ? n : 1];
    if (n > 0 && (long long)std::fread(p, sizeof(int), n, fp) != n) {
        std::cerr << "graph: short read of " << f << std::endl;
        std::exit(1);
    }
    std::fclose(fp);
    return p;
}

}  // namespace

// ---------------------------------------------------------------------------
// graph (graph.cpp:2-136)

graph::graph(const char* fw_beg_file, const char* fw_csr_file, const char* bw_beg_file, const char* bw_csr_file) {
    vert_count = (index_t)(fsizeOf(fw_beg_file) / (long long)sizeof(int) - 1);
    edge_count = (index_t)(fsizeOf(fw_csr_file) / (long long)sizeof(int));
    const long long bv = fsizeOf(bw_beg_file) / (long long)sizeof(int) - 1, be = fsizeOf(bw_csr_file) / (long long)sizeof(int);
    if (vert_count < 0 || bv != vert_count || be != edge_count) {
        std::cerr << "graph: forward CSR has " << vert_count << " vertices / " << edge_count
                  << " edges but backward CSR files hold " << bv << " / " << be << " entries; refusing to load "
                  << bw_beg_file << std::endl;
        std::exit(1);
    }
    fw_beg_pos = readInts(fw_beg_file, (long long)vert_count + 1);
    fw_csr = readInts(fw_csr_file, edge_count);
    bw_beg_pos = readInts(bw_beg_file, (long long)vert_count + 1);
    bw_csr = readInts(bw_csr_file, edge_count);
    auto ok = [&](const index_t* beg, const vertex_t* csr) {
        bool good = beg[0] == 0 && beg[vert_count] == edge_count;
        for (index_t i = 0; good && i < vert_count; ++i) good = beg[i] <= beg[i + 1];
        for (index_t i = 0; good && i < edge_count; ++i) good = csr[i] >= 0 && csr[i] < vert_count;
        return good;
    };
    if (!ok(fw_beg_pos, fw_csr) || !ok(bw_beg_pos, bw_csr)) {
        std::cerr << "graph: CSR in " << fw_beg_file << " / " << bw_beg_file
                  << " is malformed (offsets not monotone from 0 to E, or neighbour ids outside [0,V)). "
                     "Refusing to load."
                  << std::endl;
        std::exit(1);
    }
}

graph::~graph() {
    delete[] fw_beg_pos;
    delete[] fw_csr;
    delete[] bw_beg_pos;
    delete[] bw_csr;
}

graph* graph::fromEdges(index_t V, const std::vector<std::pair<int, int>>& edges) {
    graph* g = new graph();
    g->vert_count = V;
    g->edge_count = (index_t)edges.size();
    g->fw_beg_pos = new index_t[V + 1]();
    g->bw_beg_pos = new index_t[V + 1]();
    g->fw_csr = new vertex_t[edges.empty() ? 1 : edges.size()];
    g->bw_csr = new vertex_t[edges.empty() ? 1 : edges.size()];
    for (auto& e : edges) {
        g->fw_beg_pos[e.first + 1]++;
        g->bw_beg_pos[e.second + 1]++;
    }
    for (index_t v = 0; v < V; ++v) {
        g->fw_beg_pos[v + 1] += g->fw_beg_pos[v];
        g->bw_beg_pos[v + 1] += g->bw_beg_pos[v];
    }
    std::vector<index_t> fp(g->fw_beg_pos, g->fw_beg_pos + V), bp(g->bw_beg_pos, g->bw_beg_pos + V);
    for (auto& e : edges) {
        g->fw_csr[fp[e.first]++] = e.second;
        g->bw_csr[bp[e.second]++] = e.first;
    }
    return g;
}

// ---------------------------------------------------------------------------
// GIR

void GIR::init(const int32_t* fb, const int32_t* fa, const int32_t* bb, const int32_t* ba, int32_t V_, int64_t E_) {
    int rc = girhip_create(fb, fa, bb, ba, V_, E_, &h);
    if (rc == GIRHIP_EINVAL) {
        // the reference refuses malformed input with exit(1) (gir.cpp:122-126)
        std::cerr << "GIR: " << girhip_last_error() << std::endl;
        std::exit(1);
    }
    check(rc, "create");
    int64_t E;
    check(girhip_size(h, &V_, &E), "size");
    V = V_;
}

GIR::GIR(const std::string& path_directed, const std::string& path_undirected) {
    (void)path_undirected;
    const std::string d = path_directed;
    graph g((d + "fw_begin.bin").c_str(), (d + "fw_adjacent.bin").c_str(), (d + "bw_begin.bin").c_str(),
            (d + "bw_adjacent.bin").c_str());
    init(g.fw_beg_pos, g.fw_csr, g.bw_beg_pos, g.bw_csr, g.vert_count, g.edge_count);
}

GIR::GIR(graph* g) {
    std::unique_ptr<graph> own(g);
    init(g->fw_beg_pos, g->fw_csr, g->bw_beg_pos, g->bw_csr, g->vert_count, g->edge_count);
}

GIR::~GIR() {
    try {
        flush();
    } catch (...) {
    }
    girhip_destroy(h);
}

void GIR::buildWCC() {
    flush();
    check(girhip_wcc(h, nullptr, nullptr), "buildWCC");
    built_wcc = true;
    stale = true, ++epoch;
}

void GIR::buildSCC_DAG() {
    flush();
    check(girhip_scc_dag(h, nullptr, nullptr, nullptr), "buildSCC_DAG");
    built_scc = true;
    stale = true, ++epoch;
}

void GIR::refresh() const {
    flush();
    if (!stale) return;
    if (built_wcc) {
        wcc_label.resize(V);
        check(girhip_get_wcc(h, wcc_label.data(), nullptr), "getNodeToWcc");
    }
    if (built_scc) {
        // canonical labels (minimum vertex) -> dense ids in order of that vertex
        std::vector<int> canon((size_t)V);
        std::vector<long long> rk((size_t)V);
        check(girhip_get_scc(h, canon.data(), nullptr, nullptr), "getNodeToScc");
        check(girhip_get_ranks(h, reinterpret_cast<int64_t*>(rk.data())), "sccRank");
        dense_of.assign((size_t)V, -1);
        canon_of.clear();
        for (int v = 0; v < V; ++v)
            if (canon[v] == v) {
                dense_of[v] = (int)canon_of.size();
                canon_of.push_back(v);
            }
        node_scc.resize((size_t)V);
        for (int v = 0; v < V; ++v) node_scc[v] = dense_of[canon[v]];
        scc_rank.resize(canon_of.size());
        for (size_t d = 0; d < canon_of.size(); ++d) scc_rank[d] = rk[canon_of[d]];
    }
    sccs_ok = false;
    g_copy.reset();
    stale = false;
}

void GIR::applyOps(const std::vector<Update>& batch, bool repair) {
    flush();
    const size_t n = batch.size();
    std::vector<int32_t> u(n), v(n);
    std::vector<uint8_t> ins(n);
    for (size_t i = 0; i < n; ++i) {
        u[i] = batch[i].u;
        v[i] = batch[i].v;
        ins[i] = batch[i].insert ? 1 : 0;
    }
    girhip_batch_stats st;
    check(repair ? girhip_apply_batch(h, u.data(), v.data(), ins.data(), nullptr, (int64_t)n, &st)
                 : girhip_ingest_batch(h, u.data(), v.data(), ins.data(), nullptr, (int64_t)n, &st),
          repair ? "applyBatch" : "ingestBatch");
    stale = true, ++epoch;
}

// Single-edge updates are queued and applied as one batch before anything
// reads the graph or its layers (flush): the last op on an arc decides its
// presence either way, so the state every reader sees is the one the ops
// applied one by one would leave (gir.cpp:3043-3133 applies each at once).
void GIR::queueOp(int u, int v, bool insert) {
    queued.push_back({u, v, insert});
    if (queued.size() >= kQueueMax) flush();
}
void GIR::flush() const {
    if (queued.empty()) return;
    const size_t n = queued.size();
    std::vector<int32_t> u(n), v(n);
    std::vector<uint8_t> ins(n);
    for (size_t i = 0; i < n; ++i) {
        u[i] = queued[i].u;
        v[i] = queued[i].v;
        ins[i] = queued[i].insert ? 1 : 0;
    }
    queued.clear();
    girhip_batch_stats st;
    check(girhip_apply_batch(h, u.data(), v.data(), ins.data(), nullptr, (int64_t)n, &st), "insertEdge / deleteEdge");
    stale = true, ++epoch;
}
void GIR::insertEdge(int u, int v) { queueOp(u, v, true); }
void GIR::deleteEdge(int u, int v) { queueOp(u, v, false); }
void GIR::updateEdgeDirected(int u, int v, bool insert) { queueOp(u, v, insert); }
void GIR::applyBatch(const std::vector<Update>& batch) { applyOps(batch, true); }
void GIR::ingestBatch(const std::vector<Update>& batch) { applyOps(batch, false); }

void GIR::sync() {
    flush();
    check(girhip_sync(h), "sync");
    stale = true, ++epoch;
}

bool GIR::synced() const {
    flush();
    int32_t s = 1;
    check(girhip_synced(h, &s), "synced");
    return s != 0;
}

int GIR::sccOf(int v) const {
    refresh();
    return node_scc[v];
}
int GIR::wccOf(int v) const {
    refresh();
    return wcc_label[v];
}
long long GIR::sccRank(int s) const {
    refresh();
    return (s >= 0 && (size_t)s < scc_rank.size()) ? scc_rank[s] : -1;
}
bool GIR::sameScc(int u, int v) {
    sync();
    return sccOf(u) == sccOf(v);
}
bool GIR::sameWcc(int u, int v) {
    sync();
    return wccOf(u) == wccOf(v);
}

std::map<int, int> GIR::getNodeToScc() const {
    refresh();
    std::map<int, int> m;
    for (int v = 0; v < V; ++v) m[v] = node_scc[v];
    return m;
}
std::map<int, int> GIR::getNodeToWcc() const {
    refresh();
    std::map<int, int> m;
    for (int v = 0; v < V; ++v) m[v] = wcc_label[v];
    return m;
}

const std::vector<std::vector<int>>& GIR::getSCCs() const {
    refresh();
    if (!sccs_ok) {
        sccs.assign(built_scc ? canon_of.size() : 0, {});
        for (int v = 0; v < (built_scc ? V : 0); ++v) sccs[node_scc[v]].push_back(v);
        sccs_ok = true;
    }
    return sccs;
}

std::vector<std::vector<std::pair<int, int>>> GIR::getDAGWithCounts() const {
    refresh();
    std::vector<std::vector<std::pair<int, int>>> r(built_scc ? canon_of.size() : 0);
    if (!built_scc) return r;
    int32_t ns;
    int64_t nd;
    check(girhip_get_scc(h, nullptr, &ns, &nd), "getDAG");
    std::vector<int32_t> s(nd), d(nd), m(nd);
    check(girhip_get_dag(h, s.data(), d.data(), m.data()), "getDAG");
    for (int64_t i = 0; i < nd; ++i) r[dense_of[s[i]]].push_back({dense_of[d[i]], m[i]});
    return r;
}

std::vector<std::vector<int>> GIR::getDAG() const {
    auto c = getDAGWithCounts();
    std::vector<std::vector<int>> r(c.size());
    for (size_t i = 0; i < c.size(); ++i)
        for (auto& p : c[i]) r[i].push_back(p.first);
    return r;
}

std::vector<std::vector<int>> GIR::getWCCs() const {
    refresh();
    std::unordered_map<int, std::vector<int>> g;
    for (int v = 0; v < V; ++v) g[wcc_label[v]].push_back(v);
    std::vector<std::vector<int>> r;
    for (auto& kv : g) r.push_back(std::move(kv.second));
    return r;
}

graph* GIR::getDirectedGraph() const {
    flush();
    int32_t v;
    int64_t E;
    check(girhip_size(h, &v, &E), "getDirectedGraph");
    std::unique_ptr<graph> g(new graph());
    g->vert_count = v;
    g->edge_count = (index_t)E;
    g->fw_beg_pos = new index_t[v + 1];
    g->bw_beg_pos = new index_t[v + 1];
    g->fw_csr = new vertex_t[E > 0 ? E : 1];
    g->bw_csr = new vertex_t[E > 0 ? E : 1];
    check(girhip_get_csr(h, g->fw_beg_pos, g->fw_csr, g->bw_beg_pos, g->bw_csr), "getDirectedGraph");
    g_copy = std::move(g);
    return g_copy.get();
}

std::vector<int> GIR::getDirectedOutNeighbors(int v) const {
    graph* g = getDirectedGraph();
    return std::vector<int>(g->fw_csr + g->fw_beg_pos[v], g->fw_csr + g->fw_beg_pos[v + 1]);
}
std::vector<int> GIR::getDirectedInNeighbors(int v) const {
    graph* g = getDirectedGraph();
    return std::vector<int>(g->bw_csr + g->bw_beg_pos[v], g->bw_csr + g->bw_beg_pos[v + 1]);
}
bool GIR::hasDirectedEdge(int u, int v) const {
    flush();
    int32_t p = 0;
    check(girhip_has_edge(h, u, v, &p), "hasDirectedEdge");
    return p != 0;
}

// gir_query.cpp:50-54
uint32_t GIR::edgeWeight(int u, int v) {
    uint64_t x = ((uint64_t)(uint32_t)u << 32) | (uint32_t)v;
    x ^= x >> 33;
    x *= 0xff51afd7ed558ccdULL;
    x ^= x >> 33;
    x *= 0xc4ceb9fe1a85ec53ULL;
    x ^= x >> 33;
    return 1 + (uint32_t)(x % kMaxWeight);
}

QueryVec<int> GIR::bfsMany(const std::vector<int>& sources) {
    flush();
    QueryVec<int> lv((size_t)V * sources.size());
    check(girhip_bfs(h, sources.data(), (int32_t)sources.size(), lv.data(), nullptr), "bfs");
    return lv;
}
QueryVec<uint64_t> GIR::ssspMany(const std::vector<int>& sources, uint64_t delta) {
    flush();
    QueryVec<uint64_t> d((size_t)V * sources.size());
    check(girhip_sssp(h, sources.data(), (int32_t)sources.size(), delta, d.data(), nullptr), "sssp");
    return d;
}
QueryVec<int> GIR::bfs(int source) { return bfsMany({source}); }

QueryVec<int> GIR::bfsTwoLevel(int source) {
    flush();
    if (!built_scc) buildSCC_DAG();
    QueryVec<int> lv((size_t)V);
    const int32_t s = source;
    check(girhip_bfs_two_level(h, &s, 1, lv.data()), "bfsTwoLevel");
    stale = true, ++epoch;  // a pending batch may have been synced
    return lv;
}

QueryVec<uint64_t> GIR::ssspTwoLevel(int source, uint64_t delta) {
    flush();
    if (!built_scc) buildSCC_DAG();
    QueryVec<uint64_t> d((size_t)V);
    const int32_t s = source;
    check(girhip_sssp_two_level(h, &s, 1, delta, d.data()), "ssspTwoLevel");
    stale = true, ++epoch;
    return d;
}

std::vector<char> GIR::reachableSccsPublic(int c) {
    flush();
    if (!built_scc) buildSCC_DAG();
    if (!synced()) sync();
    refresh();
    std::vector<char> r(canon_of.size(), 0);
    if (c < 0 || (size_t)c >= canon_of.size()) return r;
    std::vector<uint8_t> m((size_t)V);
    check(girhip_reach_info(h, canon_of[c], m.data(), nullptr, nullptr, nullptr, nullptr), "reachableSccs");
    for (size_t d = 0; d < canon_of.size(); ++d) r[d] = (char)m[canon_of[d]];
    return r;
}

void GIR::trackIncremental(int source, bool weighted) {
    flush();
    check(girhip_inc_track(h, source, weighted ? 1 : 0), "trackIncremental");
    inc_fresh[weighted] = false;
}

void GIR::updateIncremental(bool weighted) {
    flush();
    check(girhip_inc_update(h, weighted ? 1 : 0), "updateIncremental");
    inc_fresh[weighted] = false;
}

uint64_t GIR::incDist(bool weighted, int v) const {
    flush();
    if (!inc_fresh[weighted]) {
        inc_host[weighted].resize((size_t)V);
        check(girhip_inc_dist(h, weighted ? 1 : 0, inc_host[weighted].data()), "incDist");
        inc_fresh[weighted] = true;
    }
    return (v >= 0 && v < V) ? inc_host[weighted][(size_t)v] : UINT64_MAX;
}

void GIR::ensureBlocks() {
    flush();
    if (!built_scc) buildSCC_DAG();
    if (!synced()) sync();
    if (blk_ok && blk_epoch == epoch) return;
    bc_parent.resize((size_t)V);
    bc_block.resize((size_t)V);
    std::vector<uint8_t> ap((size_t)V);
    int64_t nb = 0;
    check(girhip_bicc(h, bc_parent.data(), bc_block.data(), ap.data(), &nb), "buildBCT");
    bc_ap.assign(ap.begin(), ap.end());
    // children of every forest vertex (counting sort by parent)
    bc_cbeg.assign((size_t)V + 1, 0);
    for (int v = 0; v < V; ++v)
        if (bc_parent[v] >= 0) ++bc_cbeg[bc_parent[v] + 1];
    for (int v = 0; v < V; ++v) bc_cbeg[v + 1] += bc_cbeg[v];
    bc_cadj.resize((size_t)bc_cbeg[V]);
    std::vector<int> fill(bc_cbeg.begin(), bc_cbeg.end() - 1);
    for (int v = 0; v < V; ++v)
        if (bc_parent[v] >= 0) bc_cadj[fill[bc_parent[v]]++] = v;
    n_blocks = nb;
    blk_ok = true;
    blk_epoch = epoch;
}

std::vector<int> GIR::blocksOf(int x) const {
    std::vector<int> b;
    if (bc_block[x] >= 0) b.push_back(bc_block[x]);
    for (int i = bc_cbeg[x]; i < bc_cbeg[x + 1]; ++i) b.push_back(bc_block[bc_cadj[i]]);
    std::sort(b.begin(), b.end());
    b.erase(std::unique(b.begin(), b.end()), b.end());
    return b;
}

bool GIR::articulationPoint(int v) {
    ensureBlocks();
    return v >= 0 && v < V && bc_ap[v] != 0;
}

bool GIR::biconnected(int u, int v) {
    if (u == v) return true;
    ensureBlocks();
    if (u < 0 || v < 0 || u >= V || v >= V) return false;
    const auto a = blocksOf(u), b = blocksOf(v);
    std::vector<int> c;
    std::set_intersection(a.begin(), a.end(), b.begin(), b.end(), std::back_inserter(c));
    return !c.empty();
}

std::vector<std::vector<int>> GIR::getBCTComponents(int scc_id) {
    std::vector<std::vector<int>> r;
    ensureBlocks();
    const auto& comps = getSCCs();
    if (scc_id < 0 || (size_t)scc_id >= comps.size() || comps[scc_id].empty()) return r;
    const std::vector<int>* mem = &comps[scc_id];
    if (mem->size() == 1) return {{(*mem)[0]}};
    std::map<int, std::vector<int>> blocks;
    for (int x : *mem)
        if (bc_parent[x] >= 0) {
            auto& bl = blocks[bc_block[x]];
            bl.push_back(x);
            bl.push_back(bc_parent[x]);
        }
    for (auto& kv : blocks) {
        auto& bl = kv.second;
        std::sort(bl.begin(), bl.end());
        bl.erase(std::unique(bl.begin(), bl.end()), bl.end());
        r.push_back(std::move(bl));
    }
    return r;
}

QueryVec<double> GIR::bc(const std::vector<int>& sources) {
    flush();
    QueryVec<double> sc((size_t)V);
    check(girhip_bc(h, sources.data(), (int32_t)sources.size(), sc.data()), "bc");
    return sc;
}

std::vector<char> GIR::reachableMany(const std::vector<int>& u, const std::vector<int>& v) {
    flush();
    if (u.size() != v.size()) throw std::invalid_argument("reachableMany: u and v differ in length");
    if (!built_scc) buildSCC_DAG();
    if (!synced()) sync();
    std::vector<uint8_t> r(u.size());
    check(girhip_reachable(h, u.data(), v.data(), (int64_t)u.size(), r.data()), "reachable");
    return std::vector<char>(r.begin(), r.end());
}

bool GIR::reachable(int u, int v) { return reachableMany({u}, {v})[0] != 0; }

QueryVec<int> GIR::bfsInComponent(int source, bool directed, long long* visited) {
    flush();
    if (directed ? !built_scc : !built_wcc) {
        if (directed) buildSCC_DAG();
        else buildWCC();
    }
    QueryVec<int> lv((size_t)V);
    int64_t vis = 0;
    const int32_t s = source;
    check(girhip_bfs_component(h, &s, 1, directed ? 0 : 1, lv.data(), &vis), "bfsInComponent");
    if (visited) *visited = (long long)vis;
    return lv;
}

QueryVec<int> ParallelBFS::testParallelBFS(GIR& gir, int source, bool is_directed) {
    const auto t0 = std::chrono::high_resolution_clock::now();
    long long visited = 0;
    QueryVec<int> lv = gir.bfsInComponent(source, is_directed, &visited);
    const std::chrono::duration<double> dt = std::chrono::high_resolution_clock::now() - t0;
    std::cout << "BFS completed in: " << dt.count() * 1000.0 << " ms" << std::endl;
    std::cout << "Nodes visited in component: " << visited << std::endl;
    return lv;
}
QueryVec<uint64_t> GIR::sssp(int source, uint64_t delta) { return ssspMany({source}, delta); }

void GIR::setWeights(const uint32_t* w_fw) {
    flush();
    check(girhip_set_weights(h, w_fw), "setWeights");
}

void GIR::setQueryGranularity(size_t par_frontier, size_t large_unit) {
    (void)par_frontier;  // (a CPU threading threshold; every step runs on the device here)
    check(girhip_set_param(h, "bfs_min_reach", (double)large_unit), "setQueryGranularity");
}

void GIR::setParam(const char* name, double value) { check(girhip_set_param(h, name, value), "setParam"); }

// Drop-in GIR front end over the gfx950 C ABI (include/girhip.h).
//
// Mirrors the hot-path part of the reference's public class GIR
// (gir_constructor/gir.h:72-240): same method names, argument meaning and
// error behaviour, so callers keep buildWCC() / buildSCC_DAG() /
// getNodeToScc() / updateEdgeDirected() / applyBatch() / bfs() / sssp().
// The graph, every layer and all compute live on the GPU; the host keeps
// lazily refreshed copies of the labels for the per-vertex getters.
//
// Differences a caller can observe:
//   - SCC ids are dense, 0 .. S-1, numbered in order of each SCC's minimum
//     vertex (the reference numbers SCCs by Tarjan emission, gir.cpp:809-816,
//     and reuses retired ids after merges, gir.cpp:709-737).  sccOf(v) indexes
//     getSCCs(), getDAG() and sccRank() as in the reference, and
//     getSCCs().size() is the SCC count after every build or update;
//   - WCC ids are canonical: the minimum vertex of the component (the
//     reference uses union-find roots, gir.cpp:649-652);
//   - ranks are Kahn levels (level << 32 | canonical id) instead of
//     (S - s) << 32; the contract rank(a) < rank(b) for every condensation
//     arc is the same;
//   - the block layer (buildBCT, getBCTComponents, articulationPoint,
//     biconnected) is computed from scratch on the GPU when first asked for
//     after a change, which is the reference's on-demand mode;
//     setBlocksOnDemand is accepted and has no further effect.
#ifndef GIRHIP_GIR_H
#define GIRHIP_GIR_H
#include <cstdint>
#include <map>
#include <memory>
#include <string>
#include <type_traits>
#include <vector>

#include "graph.h"

struct girhip_graph;

// Query results, as in the reference (gir.h:17-37): vectors whose elements
// are left default-initialised (they are filled right away).
template <class T>
struct DefaultInitAllocator : std::allocator<T> {
    template <class U>
    struct rebind {
        using other = DefaultInitAllocator<U>;
    };
    DefaultInitAllocator() = default;
    template <class U>
    DefaultInitAllocator(const DefaultInitAllocator<U>&) noexcept {}
    template <class U>
    void construct(U* p) noexcept(std::is_nothrow_default_constructible<U>::value) {
        ::new ((void*)p) U;
    }
    template <class U, class... A>
    void construct(U* p, A&&... a) {
        ::new ((void*)p) U(std::forward<A>(a)...);
    }
};
template <class T>
using QueryVec = std::vector<T, DefaultInitAllocator<T>>;

class GIR {
public:
    struct Update {
        int u, v;
        bool insert;
    };

    // gir.h:93-96.  path_directed holds fw_begin.bin ... bw_adjacent.bin;
    // path_undirected is accepted for signature compatibility (the
    // reference's GIR never reads it either, gir.cpp:70-77).
    GIR(const std::string& path_directed, const std::string& path_undirected);
    explicit GIR(graph* g);  // takes ownership (gir.cpp:46, 65-68)
    ~GIR();
    GIR(const GIR&) = delete;
    GIR& operator=(const GIR&) = delete;

    // Construction (gir.h:98-101).
    void buildWCC();
    void buildSCC_DAG();
    // Biconnected blocks of every SCC's undirected view (gir.cpp:2063,
    // computed from scratch on the GPU when asked for after a change).
    void buildBCT() { ensureBlocks(); }
    void ensureBlocks();

    // Updates (gir.h:104-119).  insertEdge / deleteEdge / updateEdgeDirected
    // queue the op; queued ops are applied as one batch (last op on an arc
    // wins, which is what applying them one by one leaves) before any call
    // that reads the graph or a layer, before applyBatch / ingestBatch /
    // sync, and once kQueueMax ops are waiting.
    static const size_t kQueueMax = 4096;
    void insertEdge(int u, int v);
    void deleteEdge(int u, int v);
    void applyBatch(const std::vector<Update>& batch);
    void ingestBatch(const std::vector<Update>& batch);
    void sync();
    bool synced() const;
    void updateEdgeDirected(int u, int v, bool insert);
    void updateEdgeUndirected(int u, int v, bool insert) { updateEdgeDirected(u, v, insert); }

    // Per-vertex queries (gir.h:121-136).
    int sccOf(int v) const;
    int wccOf(int v) const;
    long long sccRank(int s) const;
    bool sameScc(int u, int v);
    bool sameWcc(int u, int v);
    int vertexCount() const { return V; }

    // Materialised views (gir.h:140-145).
    std::map<int, int> getNodeToScc() const;
    std::map<int, int> getNodeToWcc() const;
    const std::vector<std::vector<int>>& getSCCs() const;
    std::vector<std::vector<int>> getDAG() const;
    std::vector<std::vector<int>> getWCCs() const;
    // Condensation arcs with multiplicities (the reference keeps them in the
    // private dag_out maps, gir.cpp:818-831).
    std::vector<std::vector<std::pair<int, int>>> getDAGWithCounts() const;

    // Edge store (gir.h:153-160).  getDirectedGraph() returns a host copy of
    // the current CSR, refreshed on call.
    graph* getDirectedGraph() const;
    std::vector<int> getDirectedOutNeighbors(int v) const;
    std::vector<int> getDirectedInNeighbors(int v) const;
    bool hasDirectedEdge(int u, int v) const;

    // Queries (gir.h:204-240).
    QueryVec<int> bfs(int source);
    static const uint32_t kMaxWeight = 100;
    static uint32_t edgeWeight(int u, int v);
    QueryVec<uint64_t> sssp(int source, uint64_t delta = 16);
    // Many sources in one call (k x V, row i for sources[i]).
    QueryVec<int> bfsMany(const std::vector<int>& sources);
    QueryVec<uint64_t> ssspMany(const std::vector<int>& sources, uint64_t delta = 16);
    // Condensation-guided two-level queries (gir_query.cpp:604-620,
    // 705-722): equal to bfs / sssp; pending ingested changes are synced
    // first, as the reference does.
    QueryVec<int> bfsTwoLevel(int source);
    QueryVec<uint64_t> ssspTwoLevel(int source, uint64_t delta = 16);
    // SCCs reachable from SCC c in the condensation (gir.h:205,
    // gir_query.cpp:56-64), indexed by SCC id (getSCCs().size() entries).
    std::vector<char> reachableSccsPublic(int c);
    // Brandes betweenness over the sources (gir_query.cpp:728-778).
    QueryVec<double> bc(const std::vector<int>& sources);
    // Incremental single-source BFS (weighted = false) / SSSP (true)
    // (gir.h:222-233, gir_query.cpp:867-996): track once, then every edge
    // change is logged; updateIncremental corrects the result; incDist reads
    // it (UINT64_MAX = unreachable; BFS: the level).
    void trackIncremental(int source, bool weighted);
    void updateIncremental(bool weighted);
    uint64_t incDist(bool weighted, int v) const;
    // Explicit per-arc weights aligned to the forward CSR (nullptr restores
    // edgeWeight); not in the reference, whose weights are always the hash.
    void setWeights(const uint32_t* w_fw);

    // Knobs (gir.h:170-240) that exist on this path.
    void setQueryGranularity(size_t par_frontier, size_t large_unit);
    void setBlocksOnDemand(bool) {}  // blocks are always built on demand here
    bool blocksOnDemand() const { return true; }
    // Block queries (gir.h:133-148): v lies in two or more blocks; u and v
    // share a block; the blocks (vertex sets) of SCC scc_id (an sccOf id).
    bool articulationPoint(int v);
    bool biconnected(int u, int v);
    std::vector<std::vector<int>> getBCTComponents(int scc_id);
    long long blockCount() { ensureBlocks(); return n_blocks; }
    void setParam(const char* name, double value);

    girhip_graph* handle() const { return h; }

    // Condensation-guided reachability (gir_query.cpp:789-828); syncs
    // pending ingested changes first, as the reference does.
    bool reachable(int u, int v);
    // Many pairs in one call (out[i] for (u[i], v[i])).
    std::vector<char> reachableMany(const std::vector<int>& u, const std::vector<int>& v);

    // BFS inside the source's component (the ParallelBFS harness below).
    QueryVec<int> bfsInComponent(int source, bool directed, long long* visited = nullptr);

private:
    void init(const int32_t* fb, const int32_t* fa, const int32_t* bb, const int32_t* ba, int32_t V_, int64_t E_);
    void refresh() const;  // pull labels after a change
    void applyOps(const std::vector<Update>& batch, bool repair);
    void queueOp(int u, int v, bool insert);
    void flush() const;  // apply the queued single-edge ops
    mutable std::vector<Update> queued;

    girhip_graph* h = nullptr;
    int V = 0;
    bool built_wcc = false, built_scc = false;
    mutable bool stale = true;
    mutable std::vector<int> wcc_label, node_scc;  // node_scc: dense SCC ids
    mutable std::vector<int> canon_of;              // dense SCC id -> canonical id (minimum vertex)
    mutable std::vector<int> dense_of;              // canonical id -> dense id (-1 elsewhere)
    mutable std::vector<long long> scc_rank;        // by dense id
    mutable std::vector<std::vector<int>> sccs;
    mutable bool sccs_ok = false;
    mutable std::unique_ptr<graph> g_copy;
    mutable uint64_t epoch = 0;  // bumped whenever the store or a layer changes
    // block layer (host copy): forest parent, block of each tree edge,
    // articulation flags, children CSR
    std::vector<int> bc_parent, bc_block, bc_cbeg, bc_cadj;
    std::vector<char> bc_ap;
    long long n_blocks = 0;
    bool blk_ok = false;
    uint64_t blk_epoch = 0;
    std::vector<int> blocksOf(int x) const;
    mutable std::vector<uint64_t> inc_host[2];  // incDist values, refreshed after a track / update
    mutable bool inc_fresh[2] = {false, false};
};

// The reference's ParallelBFS harness (gir_test.cpp:143-343): a
// direction-optimising BFS from source restricted to the source's SCC
// (is_directed, over out-arcs) or WCC (over the undirected closure).  Like
// the reference it prints the elapsed time and "Nodes visited in component";
// it also returns the levels (-1 outside the component).
struct ParallelBFS {
    static QueryVec<int> testParallelBFS(GIR& gir, int source, bool is_directed);
};

#endif

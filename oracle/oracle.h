/* oracle.h -- CPU restatement of the reference GIR construction/query path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path and
 * the "port" CPU baseline.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it; the product library never links,
 * loads or calls it.
 *
 * Every function names the reference lines it restates.  Output ids are
 * canonical: a component is named by its minimum vertex id (the reference
 * names WCCs by union-find roots and SCCs by Tarjan emission index; its own
 * tests compare partitions, gir_test.cpp:365-374, so canonical ids carry the
 * same information).
 *
 * Inputs are canonical CSR (rows sorted, no duplicate arcs) unless stated.
 * All arrays are caller-allocated host memory.  Functions return 0 on
 * success, nonzero on bad input.
 *
 * Pinning (what checks each function against something other than itself):
 *   orc_wcc        golden vectors of the compiled reference; orc_wcc_bfs
 *   orc_scc        golden vectors; orc_scc_kosaraju (a different algorithm)
 *   orc_dag        golden DAG arc sets; hand-derived example_graph answers
 *   orc_bfs        golden levels; orc_bfs_textbook (plain queue BFS)
 *   orc_bfs_component  ParallelBFS harness: BFS inside the source's SCC /
 *                  WCC; pinned by |SCC| = visited and by orc_bfs_textbook
 *                  on a graph whose labels are all equal
 *   orc_sssp       golden distances (hash weights); orc_dijkstra
 *   orc_edge_weight  golden weights from the compiled reference
 *   orc_normalize_batch / orc_apply_batch  golden post-batch partitions;
 *                  a Python set simulation of last-op-wins on tiny batches
 *   orc_reach_info golden reachable-SCC masks (reachableSccsPublic);
 *                  a vertex-level textbook BFS (an SCC is reachable iff
 *                  its minimum vertex is); pending counts by brute force
 *                  over the DAG arc list
 *   orc_two_level  golden bfsTwoLevel / ssspTwoLevel results;
 *                  orc_bfs_textbook and orc_dijkstra
 *   orc_bicc       golden getBCTComponents blocks of every SCC; networkx's
 *                  biconnected_components on each SCC's undirected view
 *   orc_bc         golden GIR::bc scores; the definition (sums over pairs
 *                  of sigma_sv * sigma_vt / sigma_st from all-pairs BFS)
 *                  on tiny graphs
 */
#ifndef GIR_ORACLE_H
#define GIR_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* gir_query.cpp:50-54: 1 + fmix64((u << 32) | v) mod 100. */
uint32_t orc_edge_weight(int32_t u, int32_t v);

/* gir.cpp:95-116 normalizeOneCsr: sort and deduplicate every row in place;
 * *E receives the new arc count (rows compacted to the front of adj). */
int orc_normalize_csr(int32_t V, int32_t* beg, int32_t* adj, int64_t* E);

/* gir.cpp:593-655 buildWCC: union-find over the forward arcs, then labels
 * canonicalised to the minimum vertex of each component. */
int orc_wcc(int32_t V, const int32_t* fw_beg, const int32_t* fw_adj, int32_t* label, int32_t* n_wcc);

/* gir.cpp:494-537 tarjanScc over roots 0..V-1, neighbours in row order, as
 * called by buildSCC_DAG (gir.cpp:798-816); canonical min-vertex labels. */
int orc_scc(int32_t V, const int32_t* fw_beg, const int32_t* fw_adj, int32_t* scc, int32_t* n_scc);

/* gir.cpp:818-831 DAG multiplicities: for every arc u->v with
 * scc[u] != scc[v], count the pair (scc[u], scc[v]).  Output sorted by
 * (src, dst).  Pass src = NULL to get only *n_dag. */
int orc_dag(int32_t V, const int32_t* fw_beg, const int32_t* fw_adj, const int32_t* scc, int32_t* src,
            int32_t* dst, int32_t* mult, int64_t* n_dag);

/* gir.cpp:833-836 ranks: Tarjan emits SCCs in reverse topological order and
 * the reference ranks emission index s as (S - s) * 2^32.  rank[c] is filled
 * for every canonical SCC id c (other entries -1). */
int orc_scc_ranks(int32_t V, const int32_t* fw_beg, const int32_t* fw_adj, int64_t* rank);

/* gir_query.cpp:108-294 GIR::bfs: direction-optimising BFS (alpha = 15,
 * beta = 18, switching only when the reachable estimate >= 4096), OpenMP.
 * level[v] = hops from src, -1 unreachable; out-of-range src -> all -1. */
int orc_bfs(int32_t V, const int32_t* fw_beg, const int32_t* fw_adj, const int32_t* bw_beg,
            const int32_t* bw_adj, int32_t src, int32_t* level);

/* gir_query.cpp:629-703 GIR::sssp / deltaStepping: buckets of width delta,
 * bucket re-processed until empty, all out-arcs relaxed, OpenMP.  w_fw is
 * one weight per forward-CSR slot, or NULL for orc_edge_weight.
 * dist UINT64_MAX = unreachable; out-of-range src -> all UINT64_MAX. */
int orc_sssp(int32_t V, const int32_t* fw_beg, const int32_t* fw_adj, const uint32_t* w_fw, int32_t src,
             uint64_t delta, uint64_t* dist);

/* gir.cpp:3523-3586 normalizeBatch: invalid ops (u == v or out of range)
 * dropped, last op per (u, v) wins, classified against the current edge set:
 * insert of an absent arc -> ins, delete of a present arc -> del, the rest
 * no-ops (*n_noop counts the insert no-ops, as the reference does).  ins/del
 * come out sorted by (u, v); their arrays need n entries. */
int orc_normalize_batch(int32_t V, const int32_t* fw_beg, const int32_t* fw_adj, int64_t n, const int32_t* bu,
                        const int32_t* bv, const uint8_t* bins, int32_t* ins_u, int32_t* ins_v, int64_t* n_ins,
                        int32_t* del_u, int32_t* del_v, int64_t* n_del, int64_t* n_noop);

/* Final-state contract of applyBatch (gir.cpp:3634-4111): the edge set after
 * the batch is the current set minus del plus ins.  Writes the canonical
 * fw/bw CSR of that set; arrays need V+1 and E + n entries. */
int orc_apply_batch(int32_t V, const int32_t* fw_beg, const int32_t* fw_adj, int64_t n, const int32_t* bu,
                    const int32_t* bv, const uint8_t* bins, int32_t* nfw_beg, int32_t* nfw_adj, int32_t* nbw_beg,
                    int32_t* nbw_adj, int64_t* nE);

/* ---- textbook checkers (independent algorithms used to pin the above) ---- */
int orc_bfs_textbook(int32_t V, const int32_t* fw_beg, const int32_t* fw_adj, int32_t src, int32_t* level);
int orc_dijkstra(int32_t V, const int32_t* fw_beg, const int32_t* fw_adj, const uint32_t* w_fw, int32_t src,
                 uint64_t* dist);
/* Component-restricted BFS of the reference's ParallelBFS harness. */
int orc_bfs_component(int32_t V, const int32_t* fw_beg, const int32_t* fw_adj, const int32_t* bw_beg,
                      const int32_t* bw_adj, const int32_t* label, int32_t undirected, int32_t src, int32_t* level,
                      int64_t* visited);
int orc_wcc_bfs(int32_t V, const int32_t* fw_beg, const int32_t* fw_adj, const int32_t* bw_beg,
                const int32_t* bw_adj, int32_t* label);
int orc_scc_kosaraju(int32_t V, const int32_t* fw_beg, const int32_t* fw_adj, const int32_t* bw_beg,
                     const int32_t* bw_adj, int32_t* scc);

/* gir_query.cpp:56-88: the condensation part reachable from SCC src_scc
 * (canonical id): mask[V] (1 at reachable SCC ids), pending[V] (reachable
 * DAG predecessors of each reachable SCC id), counts[3] = reachable SCCs,
 * their vertices, those vertices' out-arcs.  Either array may be NULL. */
int orc_reach_info(int32_t V, const int32_t* fw_beg, const int32_t* fw_adj, const int32_t* scc, int32_t src_scc,
                   uint8_t* mask, int32_t* pending, int64_t* counts);
/* gir_query.cpp:304-350, 604-620, 705-722: bfsTwoLevel (unit = 1: levels in
 * dist, UINT64_MAX = unreachable) or ssspTwoLevel (unit = 0, bucket width
 * delta) over canonical SCC labels scc. */
int orc_two_level(int32_t V, const int32_t* fw_beg, const int32_t* fw_adj, const uint32_t* w_fw,
                  const int32_t* scc, int32_t src, uint64_t delta, int32_t unit, uint64_t* dist);

/* gir_query.cpp:728-778: Brandes betweenness over the k sources
 * (out-of-range ones skipped), score[V]. */
int orc_bc(int32_t V, const int32_t* fw_beg, const int32_t* fw_adj, const int32_t* bw_beg, const int32_t* bw_adj,
           const int32_t* src, int32_t k, double* score);

/* Blocks of every SCC's undirected view (gir_test.cpp:448-465,
 * dynamic_repair.h:96-170) as (block, vertex) membership pairs. */
int orc_bicc(int32_t V, const int32_t* fw_beg, const int32_t* fw_adj, const int32_t* bw_beg, const int32_t* bw_adj,
             const int32_t* scc, int32_t* mem_block, int32_t* mem_vertex, int64_t cap, int64_t* n_mem,
             int64_t* n_blocks);

/* threads the OpenMP parts use (omp_get_max_threads) */
int orc_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif

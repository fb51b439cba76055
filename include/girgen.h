/* girgen.h -- host-side input formats for the GIR construction/query path.
 *
 * Plain C ABI (extern "C", pointers + sizes, no C++ types).  Every function
 * returns 0 on success and a nonzero status on bad input; the message is in
 * girgen_last_error().  Arrays are caller-allocated host memory.
 *
 * Canonical CSR ("el_to_csr form"): self loops dropped, duplicate arcs
 * dropped, every row sorted ascending, int32 offsets and ids, forward (out)
 * and backward (in) CSR are transposes of each other.  This is the on-disk
 * layout the reference consumes: fw_begin.bin / fw_adjacent.bin /
 * bw_begin.bin / bw_adjacent.bin (graph.cpp:2-120, el_to_csr.cpp:96-116).
 *
 * Random draws are counter based (a SplitMix64 finaliser of seed and index),
 * so the generated graph does not depend on the thread count.
 */
#ifndef GIRGEN_H
#define GIRGEN_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

const char* girgen_last_error(void);

/* Number of RMAT samples for scale s and edge factor ef: ef * 2^s. */

/* RMAT sampler: n_samples arcs (src[i], dst[i]) over 2^scale vertices with
 * quadrant probabilities a, b, c (d = 1-a-b-c), no vertex permutation.
 * Self loops and duplicates are kept (girgen_csr_from_edges drops them).
 * No reference counterpart: the reference ships no generator. */
int girgen_rmat_edges(int32_t scale, int64_t n_samples, double a, double b, double c, uint64_t seed,
                      int32_t* src, int32_t* dst);

/* 2D lattice rows x cols (vertex r*cols+c).  Every lattice edge is kept with
 * probability p_keep; a kept edge becomes two arcs with probability p_bidir,
 * else one arc in a random direction.  Writes at most 2*(2*rows*cols) arcs;
 * *n_out receives the count.  Pass src = dst = NULL to get only the count
 * bound in *n_out. */
int girgen_grid_edges(int32_t rows, int32_t cols, double p_keep, double p_bidir, uint64_t seed,
                      int32_t* src, int32_t* dst, int64_t* n_out);

/* Erdos-Renyi G(n, m) style arcs: n_arcs uniform (src, dst) pairs. */
int girgen_uniform_edges(int32_t V, int64_t n_arcs, uint64_t seed, int32_t* src, int32_t* dst);

/* Canonical CSR from an arc list (replaces graph::fromEdges, graph.cpp:123-136,
 * plus the dedup/sort of el_to_csr.cpp:96-116 and normalizeOneCsr,
 * gir.cpp:95-116).  fw_beg/bw_beg hold V+1 entries; fw_adj/bw_adj must hold
 * n entries (the arc count after dedup, *E_out, is at most n).  Arcs with an
 * endpoint outside [0, V) are an error. */
int girgen_csr_from_edges(int32_t V, int64_t n, const int32_t* src, const int32_t* dst,
                          int32_t* fw_beg, int32_t* fw_adj, int32_t* bw_beg, int32_t* bw_adj,
                          int64_t* E_out);

/* Seeded per-arc weights in [lo, hi], one per forward-CSR slot, in fw-CSR
 * order. */
int girgen_weights(int64_t E, uint32_t lo, uint32_t hi, uint64_t seed, uint32_t* w);

/* k sources uniform over vertices with out-degree > 0 (the reference's
 * rule, bench_main.cpp:474-479, with a counter-based draw).  Fails when
 * no vertex has an out-arc. */
int girgen_sources(int32_t V, const int32_t* fw_beg, int32_t k, uint64_t seed, int32_t* out);

/* Four-file int32 .bin CSR I/O in the reference layout (graph.cpp:2-120):
 * dir/fw_begin.bin, dir/fw_adjacent.bin, dir/bw_begin.bin, dir/bw_adjacent.bin.
 * girgen_read_csr is two-phase like girgen_csr_from_edges (NULL arrays ->
 * sizes only).  It validates like graph.cpp:95-114 (offsets monotone from 0 to
 * E, ids in [0, V), fw and bw sizes agree) and fails instead of exiting. */
int girgen_write_csr(const char* dir, int32_t V, int64_t E, const int32_t* fw_beg, const int32_t* fw_adj,
                     const int32_t* bw_beg, const int32_t* bw_adj);
int girgen_read_csr(const char* dir, int32_t* V_out, int64_t* E_out, int32_t* fw_beg, int32_t* fw_adj,
                    int32_t* bw_beg, int32_t* bw_adj);

/* Streaming-protocol inputs in the text formats of the reference's
 * baseline drivers (written by bench_main.cpp:280-345 --export, read by
 * experiments/baselines/common/protocol.h): the arcs of the canonical CSR
 * shuffled with a seeded draw; loaded.el = the first half ("u v" lines),
 * full.el = all arcs, insert.ops = k held-out arcs ("a u v"), delete.ops =
 * k arcs of the loaded half ("d u v"), mixed.ops = k/2 further held-out
 * inserts and k/2 further loaded deletes in a seeded shuffle, with
 * k = min(k_ops, E/4); sources.txt = n_sources query sources (out-degree
 * > 0); meta.txt = V, E, loaded, ops_per_phase, max_weight, symmetric.
 * full.el only when write_full != 0 (the streaming callers do not read
 * it).  dir must exist. */
int girgen_export_protocol(const char* dir, int32_t V, const int32_t* fw_beg, const int32_t* fw_adj,
                           int64_t k_ops, int32_t n_sources, uint64_t seed, int32_t write_full);

#ifdef __cplusplus
}
#endif
#endif

// Host CSR container with the reference's public layout (graph.h:11-57):
// int32 fw_beg_pos / fw_csr / bw_beg_pos / bw_csr, vert_count, edge_count,
// the four-file .bin constructor and fromEdges.  Owned arrays are new[].
#ifndef GIRHIP_GRAPH_H
#define GIRHIP_GRAPH_H
#include <utility>
#include <vector>

typedef int index_t;
typedef int vertex_t;

class graph {
public:
    index_t* fw_beg_pos = nullptr;
    vertex_t* fw_csr = nullptr;
    index_t* bw_beg_pos = nullptr;
    vertex_t* bw_csr = nullptr;
    index_t vert_count = 0;
    index_t edge_count = 0;

    graph() = default;
    // Loads fw_begin.bin / fw_adjacent.bin / bw_begin.bin / bw_adjacent.bin
    // (graph.cpp:2-120); malformed files print a message and exit(1), as the
    // reference does (graph.cpp:95-114).
    graph(const char* fw_beg_file, const char* fw_csr_file, const char* bw_beg_file, const char* bw_csr_file);
    ~graph();
    graph(const graph&) = delete;
    graph& operator=(const graph&) = delete;
    // Forward and backward CSR from an arc list, rows unsorted, duplicates
    // kept (graph.cpp:123-136).  Caller owns the result.
    static graph* fromEdges(index_t V, const std::vector<std::pair<int, int>>& edges);
};
#endif

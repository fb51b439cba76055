// Multi-source delta-stepping with shared buckets and source-interleaved
// distances (msssp.hip).
#pragma once
#include "common.h"

namespace gir {

constexpr int MS_MAX = 32;  // sources per interleaved batch (one mask bit each)

// k <= MS_MAX sources (device array dsrc; out-of-range ids give all-
// unreachable rows) into k rows of V uint64 distances (device memory rows,
// UINT64_MAX = unreachable).  sv: the split view of delta, or null (rows
// classify their arcs by weight on the fly).  try32: first try 32-bit
// distances (redone in 64 bits when one would reach 2^32 - 1).
void msBatch(girhip_graph* h, const int32_t* dsrc, int k, uint64_t delta, const girhip_graph::SplitView* sv,
             uint64_t* rows, bool try32);

}  // namespace gir

// Append-only id lists walked level by level on the device: the ids of the
// current level are list[ctl[T_LO] .. ctl[T_HI]); kernels append the next
// level at ctl[T_TAIL]; kListAdvance makes it current.  Used by the
// condensation sweeps and the two-level wavefront (sssp.hip) and the
// invalidation spread of the incremental queries (inc.hip).
#pragma once
#include "common.h"

namespace gir {

enum : int { T_LO = 0, T_HI, T_TAIL, T_NV, T_NBIG, T_VERTS, T_EDGES, T_SLOTS = 8 };

// Rows (CSR beg) of the ids list[ctl[T_LO] .. ctl[T_HI]) (bounds read on the
// device), each with its length as payload; grid stride, one lane per id.
template <class F>
__global__ void kListRows(const int32_t* __restrict__ beg, const int32_t* list, const int64_t* ctl, F f, HeavyList H) {
    const int64_t lo = ctl[T_LO], hi = ctl[T_HI];
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t base = lo + (int64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~(WAVE - 1)); base < hi;
         base += stride) {
        const int64_t i = base + laneId();
        const bool act = i < hi;
        const int32_t x = act ? list[i] : 0;
        const int64_t b = act ? beg[x] : 0, e = act ? beg[x + 1] : 0;
        waveRowsH(act, x, (uint64_t)(e - b), 0, b, e, H, f);
    }
}

// Next level / wave: [lo, hi) := [hi, tail); the vertex list restarts.
static __global__ void kListAdvance(int64_t* ctl) {
    if (threadIdx.x == 0) {
        ctl[T_LO] = ctl[T_HI];
        ctl[T_HI] = ctl[T_TAIL];
        ctl[T_NV] = 0;
    }
}

// Whole list as the current range.
static __global__ void kListAll(int64_t* ctl) {
    if (threadIdx.x == 0) {
        ctl[T_LO] = 0;
        ctl[T_HI] = ctl[T_TAIL];
    }
}

}  // namespace gir

// BFS queries: replaces GIR::bfs (gir_query.cpp:108-294) for k sources.
//
// Each source runs the direction-optimising engine of reach.h over the
// forward CSR (push) and backward CSR (pull) with the reference's switch
// rule; levels are written straight into the caller's k x V array when it
// is device memory.
//
// HBM traffic per source (algorithmic): 4V (level init + writes) + 8 n_s
// (row offsets of reached vertices) + 8 m_s (adjacency + neighbour state),
// n_s = reached vertices, m_s = their out-arc count.
#include <vector>

#include "common.h"
#include "reach.h"

namespace gir {

ReachBufs reachBufs(girhip_graph* h) {
    Scratch& S = scratch(h);
    const size_t V = (size_t)(h->V > 0 ? h->V : 1);
    const int64_t W = ((int64_t)V + 63) / 64;
    S.q[0].ensure(V);
    S.q[1].ensure(V);
    for (int i = 0; i < 3; ++i) S.bits[i].ensure((size_t)W);
    ReachBufs B;
    B.q[0] = S.q[0].p;
    B.q[1] = S.q[1].p;
    B.vis = S.bits[0].p;
    B.fbits = S.bits[1].p;
    B.nbits = S.bits[2].p;
    B.W = W;
    return B;
}

// m_s per row: out-arc count summed over reached vertices (level >= 0).
__global__ void kReachedArcs(const int32_t* __restrict__ level, const int32_t* __restrict__ beg, int32_t V,
                             int64_t* out) {
    int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t m = (v < V && level[v] >= 0) ? (int64_t)(beg[v + 1] - beg[v]) : 0;
    for (int d = WAVE / 2; d > 0; d >>= 1) m += __shfl_xor(m, d);
    if (laneId() == 0 && m) atomicAdd((unsigned long long*)out, (unsigned long long)m);
}

std::vector<int32_t> hostSources(girhip_graph* h, const int32_t* src, int32_t k) {
    std::vector<int32_t> s(k);
    if (k > 0) {
        HIPCHK(hipMemcpyAsync(s.data(), src, sizeof(int32_t) * k, hipMemcpyDefault, h->st));
        syncStream(h);
    }
    return s;
}

int64_t outDegree(girhip_graph* h, int32_t v) {
    HIPCHK(hipMemcpyAsync(h->hcnt, h->fw_beg.p + v, sizeof(int32_t) * 2, hipMemcpyDeviceToHost, h->st));
    syncStream(h);
    const int32_t* p = (const int32_t*)h->hcnt;
    return (int64_t)p[1] - p[0];
}

}  // namespace gir

using namespace gir;

extern "C" int girhip_bfs(girhip_graph* h, const int32_t* src, int32_t k, int32_t* levels, int64_t* edges_out) {
    return guarded([&] {
        useDevice(h);
        if (k < 0 || (k > 0 && !src)) raise(GIRHIP_EINVAL, "bfs: bad source list");
        const int32_t V = h->V;
        std::vector<int32_t> s = hostSources(h, src, k);
        const size_t n = (size_t)k * (size_t)V;
        int32_t* out = levels;
        const bool direct = levels && isDevicePtr(levels);
        Scratch& S = scratch(h);
        if (!direct) {
            S.stage.ensure(sizeof(int32_t) * (n > 0 ? n : 1));
            out = (int32_t*)S.stage.p;
        }
        ReachBufs B = reachBufs(h);
        tBegin(h);
        fillI32(h, out, -1, (int64_t)n);
        for (int32_t i = 0; i < k; ++i) {
            const int32_t x = s[i];
            if (x < 0 || x >= V) continue;
            fillU64(h, B.vis, 0, B.W);
            fillI32(h, B.q[0], x, 1);
            const int64_t m0 = outDegree(h, x);
            runReach(h, h->fw_beg.p, h->fw_adj.p, h->bw_beg.p, h->bw_adj.p, B, 1, m0, (int64_t)V, h->E, AllArcs{},
                     WriteLevel{out + (size_t)i * V});
        }
        tEnd(h, "bfs");
        if (edges_out) {
            // 32 counter slots per readback
            std::vector<int64_t> e(k);
            for (int32_t i0 = 0; i0 < k; i0 += 32) {
                const int c = std::min<int32_t>(32, k - i0);
                fillI64(h, h->dcnt.p + 32, 0, 32);
                for (int32_t i = i0; i < i0 + c; ++i)
                    launch(h, "reached_arcs", kReachedArcs, gridFor(V, 256), 256, 0,
                           (const int32_t*)(out + (size_t)i * V), (const int32_t*)h->fw_beg.p, V,
                           h->dcnt.p + 32 + (i - i0));
                readCounters(h, h->dcnt.p + 32, c, e.data() + i0);
            }
            HIPCHK(hipMemcpy(edges_out, e.data(), sizeof(int64_t) * (size_t)k, hipMemcpyDefault));
        }
        if (levels && !direct) copyOut(h, levels, out, sizeof(int32_t) * n);
        syncStream(h);
    });
}

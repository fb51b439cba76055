// Deleted arcs that leave reachability intact.  An arc u -> v removed by an
// update batch changes neither the SCC partition nor the WCC partition when
// u still reaches v in the new graph: every walk that used the arc can take
// the other path instead.  The reference answers its per-op deletes by local
// search before any rebuild (gir.cpp:3062-3131, deleteEdgeImpl); here a
// small set of deletes (at most BYPASS_MAX) is tested at once by a bounded
// bidirectional search on the device:
//   forward: every vertex within two arcs of u is marked in a bitmap of its
//   own per delete (rows longer than HEAVY_T go to the grid-wide heavy pass);
//   backward: v's in-neighbours and theirs look the marks up.
// A hit is a path of at most four arcs.  No hit proves nothing, and the
// caller falls back to its general repair for that delete.
#include "common.h"

namespace gir {
namespace {

struct MarkF {  // marks the head of arc j in the delete's bitmap
    const int32_t* fa;
    uint64_t* bits;
    int64_t W;
    __device__ int64_t operator()(int32_t, uint64_t, int32_t i, int64_t j) const {
        const int32_t x = fa[j];
        uint64_t* wp = bits + (int64_t)i * W + (x >> 6);
        const uint64_t b = 1ULL << (x & 63);
        if (!(*wp & b)) atomicOr((unsigned long long*)wp, (unsigned long long)b);
        return 0;
    }
};

struct LookF {  // arc j of a backward row: is its tail marked for delete i?
    const int32_t* ba;
    const uint64_t* bits;
    int64_t W;
    uint8_t* found;
    __device__ int64_t operator()(int32_t, uint64_t, int32_t i, int64_t j) const {
        const int32_t y = ba[j];
        if ((bits[(int64_t)i * W + (y >> 6)] >> (y & 63)) & 1) found[i] = 1;
        return 0;
    }
};

__device__ __forceinline__ void keyEnds(const uint64_t* dk, int i, int B, int32_t& u, int32_t& v) {
    const uint64_t k = dk[i];
    u = (int32_t)(k >> B);
    v = (int32_t)(k & ((1ULL << B) - 1));
}

// blockIdx.y = delete i: lane per arc (u -> x) of u's row; x is marked and
// its row is walked (marking every head).
__global__ void kBypassFwd(const uint64_t* __restrict__ dk, int B, const int32_t* __restrict__ fb, MarkF f,
                           HeavyList H) {
    const int i = blockIdx.y;
    int32_t u, v;
    keyEnds(dk, i, B, u, v);
    const int64_t b0 = fb[u], e0 = fb[u + 1];
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t base = b0 + (int64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~(WAVE - 1)); base < e0;
         base += stride) {
        const int64_t j = base + laneId();
        const bool act = j < e0;
        const int32_t x = act ? f.fa[j] : 0;
        if (act) f(u, 0, i, j);
        waveRowsH(act, x, 0, i, act ? fb[x] : 0, act ? fb[x + 1] : 0, H, f);
    }
}

// blockIdx.y = delete i: lane per arc (x -> v) of v's backward row: x is
// looked up, then x's backward row.
__global__ void kBypassBwd(const uint64_t* __restrict__ dk, int B, const int32_t* __restrict__ bb, LookF f,
                           HeavyList H) {
    const int i = blockIdx.y;
    int32_t u, v;
    keyEnds(dk, i, B, u, v);
    const int64_t b0 = bb[v], e0 = bb[v + 1];
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t base = b0 + (int64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~(WAVE - 1)); base < e0;
         base += stride) {
        const int64_t j = base + laneId();
        const bool act = j < e0;
        const int32_t x = act ? f.ba[j] : 0;
        if (act) f(v, 0, i, j);
        waveRowsH(act, x, 0, i, act ? bb[x] : 0, act ? bb[x + 1] : 0, H, f);
    }
}

}  // namespace

// For the nd deleted arcs (dk: packed (u << B) | v keys on the device, the
// store already without them): found[i] = 1 when u_i still reaches v_i by a
// path of at most four arcs.  Returns true when every delete is bypassed;
// found is device memory of nd bytes.  Only called for nd <= BYPASS_MAX.
bool deletesBypassed(girhip_graph* h, const uint64_t* dk, int64_t nd, int B, uint8_t* found) {
    if (nd <= 0) return true;
    if (nd > BYPASS_MAX) raise(GIRHIP_EINVAL, "bypass check: too many deletes");
    Scratch& S = scratch(h);
    const int64_t W = ((int64_t)h->V + 63) / 64;
    S.bp_bits.ensure((size_t)(W * nd));
    fillU64(h, S.bp_bits.p, 0, W * nd);
    fillU8(h, found, 0, nd);
    const int blk = 256;
    const dim3 grid(64, (unsigned)nd);
    const MarkF mf{h->fw_adj.p, S.bp_bits.p, W};
    launch(h, "bypass_fwd", kBypassFwd, grid, blk, 0, dk, B, (const int32_t*)h->fw_beg.p, mf, heavyList(h));
    launchHeavyStore(h, "bypass_fwd_heavy", mf);
    const LookF lf{h->bw_adj.p, S.bp_bits.p, W, found};
    launch(h, "bypass_bwd", kBypassBwd, grid, blk, 0, dk, B, (const int32_t*)h->bw_beg.p, lf, heavyList(h));
    launchHeavyStore(h, "bypass_bwd_heavy", lf);
    uint8_t f[BYPASS_MAX];
    HIPCHK(hipMemcpyAsync(f, found, (size_t)nd, hipMemcpyDeviceToHost, h->st));
    syncStream(h);
    for (int64_t i = 0; i < nd; ++i)
        if (!f[i]) return false;
    return true;
}

}  // namespace gir

// Strongly connected components + condensation DAG + topological ranks.
//
// Replaces GIR::buildSCC_DAG (gir.cpp:798-839), which runs a sequential
// iterative Tarjan (gir.cpp:494-537), counts condensation arcs in per-SCC
// hash maps (gir.cpp:818-831) and ranks SCCs by reversed emission order
// (gir.cpp:833-836).  Here, over the vertices not yet assigned ("active"):
//
//   trim-1   a vertex with no active out-neighbour or no active in-neighbour
//            (self loops ignored) is a singleton SCC; repeated;
//   trim-2   (knob scc_trim2) two vertices whose only active out-neighbours
//            (or only active in-neighbours) are each other form an SCC of two
//            (the reference's iSpan trim_2_from_fq, includes/scc_cpu/
//            trim_2_3.h:161-308); run in each trim round after trim-1;
//   FW-BW    from the pivot maximising outdeg x indeg (the rule of the
//            reference's iSpan-style side tree, includes/scc_cpu/
//            scc_common.h:98-158), forward and backward reachability over the
//            active vertices; their intersection is the pivot's SCC;
//   colour   every active vertex starts with colour = its id and the minimum
//            is propagated along arcs to a fixpoint, so colour[v] = the
//            smallest active vertex that reaches v.  A vertex whose colour is
//            its own id is the minimum of its SCC; a backward sweep from all
//            such roots at once, restricted to arcs inside one colour, marks
//            exactly their SCCs.  Repeat on what is left.
//
// Every SCC is labelled by its minimum vertex (trim: itself; FW-BW: a min
// reduction; colouring: the root), the canonical id.
//
// DAG: one pass over the forward arcs emits the packed key (scc[u], scc[v])
// of every crossing arc; a radix sort and a run-length pass give the distinct
// condensation arcs with their multiplicities, sorted by (src, dst).
// Ranks: Kahn levels over the DAG, rank = (level << 32) | scc_id, so ranks are
// distinct and every arc a -> b has rank[a] < rank[b].
#include "common.h"
#include "reach.h"

namespace gir {

// Labels are fixed during a reachability sweep, and elsewhere they and the
// colours only move one way (unassigned -> assigned, colour down), so plain
// loads are safe pre-checks: a stale value only costs extra work.
struct ActiveOnly {
    const int32_t* scc;
    __device__ bool operator()(int32_t c, int32_t) const { return scc[c] == -1; }
    __device__ bool self(int32_t v) const { return scc[v] == -1; }
};
struct SameColour {
    const int32_t* scc;
    const int32_t* color;
    __device__ bool operator()(int32_t c, int32_t p) const { return scc[c] == -1 && color[c] == color[p]; }
    __device__ bool self(int32_t v) const { return scc[v] == -1; }
};

// List kernels take CA_ITEMS entries per thread (coalesced, blockDim apart)
// and reserve their output once per block: a list of millions then costs a
// few thousand atomics on its counter instead of one per block of 256.
constexpr int CA_ITEMS = 8;
__host__ __device__ constexpr int64_t caSpan(int block) { return (int64_t)block * CA_ITEMS; }

// With all, the entries are the vertices 0..n-1; then (fb, bb given) an
// active vertex without out-arcs or without in-arcs is labelled as its own
// SCC here (no cycle can pass through it) instead of entering the list: the
// trim rounds then start without the vertices a degree test settles.
//
// List lengths: n, or with n_dev the smaller of n and *n_dev (a length a
// kernel earlier on the stream wrote: the host launches for an upper bound
// and reads no count back).  The list kernels walk their range grid-stride.
__device__ __forceinline__ int64_t listLen(int64_t n, const int64_t* n_dev) {
    if (!n_dev) return n;
    const int64_t d = *n_dev;
    return d < n ? d : n;
}
__global__ void kCompactActive(const int32_t* __restrict__ in, int64_t n, const int64_t* __restrict__ n_dev, bool all,
                               int32_t* __restrict__ scc, int32_t* __restrict__ out, int64_t* cnt,
                               const int32_t* __restrict__ fb, const int32_t* __restrict__ bb) {
    n = listLen(n, n_dev);
    const int64_t span = caSpan(blockDim.x);
    for (int64_t s0 = (int64_t)blockIdx.x * span; s0 < n; s0 += (int64_t)gridDim.x * span) {
        const int64_t i0 = s0 + threadIdx.x;
        int32_t v[CA_ITEMS];
        int k = 0;
        unsigned keep = 0;
#pragma unroll
        for (int t = 0; t < CA_ITEMS; ++t) {
            const int64_t i = i0 + (int64_t)t * blockDim.x;
            v[t] = 0;
            if (i < n) {
                v[t] = all ? (int32_t)i : in[i];
                if (scc[v[t]] == -1) {
                    if (fb && (fb[v[t] + 1] == fb[v[t]] || bb[v[t] + 1] == bb[v[t]])) {
                        scc[v[t]] = v[t];
                    }
                    else {
                        keep |= 1u << t;
                        ++k;
                    }
                }
            }
        }
        int64_t pos = blockReserve(k, cnt);
#pragma unroll
        for (int t = 0; t < CA_ITEMS; ++t)
            if ((keep >> t) & 1) out[pos++] = v[t];
    }
}

__device__ __forceinline__ bool anyActive(const int32_t* __restrict__ beg, const int32_t* __restrict__ adj,
                                          const int32_t* scc, int32_t v) {
    const int32_t b = beg[v], e = beg[v + 1];
    for (int32_t j = b; j < e; ++j) {
        const int32_t w = adj[j];
        if (w != v && scc[w] == -1) return true;
    }
    return false;
}

// prev (may be null): the previous round's flag; a round after one that
// trimmed nothing has nothing to trim and returns at once.
__global__ void kTrim1(const int32_t* __restrict__ list, int64_t n, const int64_t* __restrict__ n_dev,
                       const int32_t* __restrict__ fb, const int32_t* __restrict__ fa, const int32_t* __restrict__ bb,
                       const int32_t* __restrict__ ba, int32_t* scc, int64_t* cnt, const int64_t* __restrict__ prev) {
    if (prev && *prev == 0) return;
    n = listLen(n, n_dev);
    bool trim = false;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int32_t v = list[i];
        if (scc[v] == -1 && (!anyActive(fb, fa, scc, v) || !anyActive(bb, ba, scc, v))) {
            stRel(&scc[v], v);
            trim = true;
        }
    }
    // only "any trimmed" matters to the host: a block that trimmed raises the
    // flag with a device-coherent store unless it reads it raised already
    // (not a contended atomic per wave)
    if (__syncthreads_or(trim) && threadIdx.x == 0 && ldAcq(cnt) == 0) stRel(cnt, (int64_t)1);
}

// The only active neighbour of v other than v in its row (beg/adj): -1 if
// none, -2 if two or more.
__device__ __forceinline__ int32_t onlyActive(const int32_t* __restrict__ beg, const int32_t* __restrict__ adj,
                                              const int32_t* scc, int32_t v) {
    int32_t one = -1;
    const int32_t b = beg[v], e = beg[v + 1];
    for (int32_t j = b; j < e; ++j) {
        const int32_t w = adj[j];
        if (w == v || scc[w] != -1) continue;
        if (one >= 0 && w != one) return -2;
        one = w;
    }
    return one;
}

// Trim-2: v's only active out-neighbour w has v as its only active
// out-neighbour (or the same with in-neighbours): {v, w} has no arc to (from)
// any other active vertex and is strongly connected, so it is an SCC.  Both
// ends are labelled by whichever of their threads sees the pair first (the
// minimum id, the canonical label), so the partner never runs trim-1 alone.
// Labels set during the pass only remove finished SCCs from the active set,
// on which the rule stays sound.
__global__ void kTrim2(const int32_t* __restrict__ list, int64_t n, const int64_t* __restrict__ n_dev,
                       const int32_t* __restrict__ fb, const int32_t* __restrict__ fa, const int32_t* __restrict__ bb,
                       const int32_t* __restrict__ ba, int32_t* scc, int64_t* cnt) {
    n = listLen(n, n_dev);
    bool hit = false;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int32_t v = list[i];
        if (scc[v] == -1) {
            int32_t w = onlyActive(fb, fa, scc, v);
            if (!(w >= 0 && onlyActive(fb, fa, scc, w) == v)) {
                w = onlyActive(bb, ba, scc, v);
                if (!(w >= 0 && onlyActive(bb, ba, scc, w) == v)) w = -1;
            }
            if (w >= 0) {
                const int32_t lab = v < w ? v : w;
                stRel(&scc[v], lab);
                stRel(&scc[w], lab);
                hit = true;
            }
        }
    }
    if (__syncthreads_or(hit) && threadIdx.x == 0 && ldAcq(cnt) == 0) stRel(cnt, (int64_t)1);
}

// argmax over the active list of min(outdeg * indeg, 2^33 - 1), ties to the
// smaller vertex id: key = (score << 31) | (2^31 - 1 - v).
// (grid-stride over a bounded grid; one atomic per block)
__global__ void kPivot(const int32_t* __restrict__ list, int64_t n, const int64_t* __restrict__ n_dev,
                       const int32_t* __restrict__ fb, const int32_t* __restrict__ bb, unsigned long long* best) {
    n = listLen(n, n_dev);
    uint64_t key = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int32_t v = list[i];
        unsigned long long s = (unsigned long long)(fb[v + 1] - fb[v]) * (unsigned long long)(bb[v + 1] - bb[v]);
        if (s > (1ULL << 33) - 1) s = (1ULL << 33) - 1;
        const uint64_t k = (s << 31) | (unsigned long long)(0x7fffffff - v);
        key = k > key ? k : key;
    }
    key = blockMaxU64(key);
    if (threadIdx.x == 0 && key) atomicMax(best, (unsigned long long)key);
}

__device__ __forceinline__ bool bitOf(const uint64_t* b, int32_t v) { return (b[v >> 6] >> (v & 63)) & 1; }

// (grid-stride over a bounded grid; one atomic per block)
__global__ void kFwBwMin(const int32_t* __restrict__ list, int64_t n, const uint64_t* __restrict__ fwb,
                         const uint64_t* __restrict__ bwb, int64_t* minv) {
    uint64_t x = 0;  // complement of the minimum so far
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int32_t v = list[i];
        if (bitOf(fwb, v) && bitOf(bwb, v)) x = ~(uint64_t)v > x ? ~(uint64_t)v : x;
    }
    x = blockMaxU64(x);
    if (threadIdx.x == 0 && x) atomicMin((long long*)minv, (long long)~x);
}

__global__ void kFwBwAssign(const int32_t* __restrict__ list, int64_t n, const uint64_t* __restrict__ fwb,
                            const uint64_t* __restrict__ bwb, const int64_t* __restrict__ minv, int32_t* scc) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t v = list[i];
    if (bitOf(fwb, v) && bitOf(bwb, v)) scc[v] = (int32_t)minv[0];
}

__global__ void kColorInit(const int32_t* __restrict__ list, int64_t n, int32_t* color) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) color[list[i]] = list[i];
}

// One propagation step: every queued vertex pushes its colour to its active
// out-neighbours (atomicMin); a neighbour whose colour dropped is queued once
// per step (stamp).
struct ColorPropF {
    const int32_t* fa;
    const int32_t* scc;
    int32_t* color;
    int32_t* stamp;
    int32_t it;
    int32_t* qout;
    int64_t* cnt;
    __device__ int64_t operator()(int32_t p, uint64_t, int32_t, int64_t j) const {
        const int32_t y = fa[j];
        bool push = false;
        if (scc[y] == -1) {
            const int32_t cx = color[p];
            if (cx < color[y]) {
                const int32_t old = atomicMin(&color[y], cx);
                push = cx < old && atomicExch(&stamp[y], it) != it;
            }
        }
        const int64_t pos = wavePush(cnt, push);
        if (push) qout[pos] = y;
        return 0;
    }
};

__global__ void kColorProp(const int32_t* __restrict__ qin, int64_t nin, const int32_t* __restrict__ fb, ColorPropF f,
                           HeavyList H) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool act = i < nin;
    const int32_t x = act ? qin[i] : 0;
    const int32_t b = act ? fb[x] : 0, e = act ? fb[x + 1] : 0;
    waveRowsH(act, x, 0, 0, b, e, H, f);
}

__global__ void kColorRoots(const int32_t* __restrict__ list, int64_t n, const int32_t* __restrict__ color,
                            int32_t* __restrict__ out, int64_t* cnt) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int32_t v = 0;
    bool root = false;
    if (i < n) {
        v = list[i];
        root = color[v] == v;
    }
    const int64_t pos = blockReserve(root, cnt);
    if (root) out[pos] = v;
}

__global__ void kColorAssign(const int32_t* __restrict__ list, int64_t n, const uint64_t* __restrict__ vis,
                             const int32_t* __restrict__ color, int32_t* scc) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t v = list[i];
    if (bitOf(vis, v)) scc[v] = color[v];
}

// ---- DAG ----
// The condensation's arcs are the crossing arcs (ends with different
// labels), keyed (tail label << B) | head label, sorted and run-length
// counted.  The big SCC (the label of the vertex of largest degree) gets a
// bitmap, gbits (bit v set iff scc[v] == G; V / 8 bytes, L2-resident where
// scc[] is not).
__global__ void kMaxDegVertex(const int32_t* __restrict__ fb, const int32_t* __restrict__ bb, int32_t V,
                              unsigned long long* best) {
    unsigned long long k = 0;
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < V; v += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t d = (uint64_t)(fb[v + 1] - fb[v]) + (uint64_t)(bb[v + 1] - bb[v]);
        const unsigned long long c = (d << 32) | (uint64_t)(0xffffffffu - (uint32_t)v);
        k = c > k ? c : k;
    }
    for (int d = WAVE / 2; d > 0; d >>= 1) {
        const unsigned long long o = __shfl_xor(k, d);
        k = o > k ? o : k;
    }
    // one atomic per block
    __shared__ unsigned long long part[1024 / WAVE];
    const int w = threadIdx.x / WAVE, nw = (blockDim.x + WAVE - 1) / WAVE;
    if (laneId() == 0) part[w] = k;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int i = 1; i < nw; ++i) k = part[i] > k ? part[i] : k;
        if (k) atomicMax(best, k);
    }
}

// G = the label of the vertex of largest degree; bit v of bits = (scc[v] == G)
__global__ void kLabelBits(const int32_t* __restrict__ scc, int32_t V, const unsigned long long* __restrict__ best,
                           uint64_t* __restrict__ bits, int32_t* gout) {
    const int32_t top = (int32_t)(0xffffffffu - (uint32_t)(*best & 0xffffffffULL));
    const int32_t G = scc[top];
    if (blockIdx.x == 0 && threadIdx.x == 0) *gout = G;
    const int64_t W = ((int64_t)V + 63) / 64;
    for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < W * 64; base += (int64_t)gridDim.x * blockDim.x) {
        const int64_t v = base + threadIdx.x;
        const uint64_t m = __ballot(v < V && scc[v] == G);
        if (laneId() == 0) bits[v >> 6] = m;
    }
}

// Every crossing arc has an end outside the big SCC: the arcs of the rows of
// the vertices outside it (list) are all the crossing arcs - forward rows
// give the arcs leaving such a vertex, backward rows the arcs entering it
// from the big SCC - so the condensation keys come from those rows alone
// (on a power-law graph a few percent of the arcs) instead of every arc.
__global__ void kNonBig(const uint64_t* __restrict__ bits, int32_t V, int32_t* __restrict__ list, int64_t* cnt) {
    const int64_t W = ((int64_t)V + 63) / 64;
    for (int64_t b0 = (int64_t)blockIdx.x * blockDim.x; b0 < W; b0 += (int64_t)gridDim.x * blockDim.x) {
        const int64_t w = b0 + threadIdx.x;
        uint64_t m = 0;
        if (w < W) {
            m = ~bits[w];
            const int64_t rest = (int64_t)V - w * 64;
            if (rest < 64) m &= (1ULL << rest) - 1;
        }
        int64_t pos = blockReserve(__popcll(m), cnt);
        while (m) {
            const int b = __ffsll((long long)m) - 1;
            m &= m - 1;
            list[pos++] = (int32_t)(w * 64 + b);
        }
    }
}

// Arc j of the forward (dir 0) or backward (dir 1) row of a vertex outside
// the big SCC with label pay: does it cross SCCs, and its key.
struct DagArc {
    const int32_t* fa;
    const int32_t* ba;
    const int32_t* scc;
    const uint64_t* gbits;
    int32_t G;
    int B;
    __device__ __forceinline__ bool operator()(uint64_t pay, int32_t dir, int64_t j, uint64_t& key) const {
        const int32_t a = (int32_t)pay;
        if (dir == 0) {
            const int32_t v = fa[j];
            const int32_t c = ((gbits[v >> 6] >> (v & 63)) & 1) ? G : scc[v];
            key = ((uint64_t)(uint32_t)a << B) | (uint32_t)c;
            return c != a;
        }
        const int32_t x = ba[j];
        key = ((uint64_t)(uint32_t)G << B) | (uint32_t)a;
        return (gbits[x >> 6] >> (x & 63)) & 1;  // tails outside the big SCC own their arcs
    }
};

// Block-synchronous over the list: the lanes count the crossing arcs they
// evaluate, one reservation per block iteration gives each lane its offset,
// and the same lanes evaluate the same arcs again, writing the keys there
// (no counter hit per wave; long rows are walked by their wave).
__global__ void kDagKeysRows(const int32_t* __restrict__ list, const int64_t* __restrict__ nlist,
                             const int32_t* __restrict__ fb, const int32_t* __restrict__ bb, DagArc d,
                             const int32_t* __restrict__ gl, uint64_t* __restrict__ keys, int64_t* cnt) {
    const int64_t n = *nlist;
    d.G = *gl;
    for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < n; base += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = base + threadIdx.x;
        const bool act = i < n;
        const int32_t u = act ? list[i] : 0;
        const uint64_t a = act ? (uint64_t)(uint32_t)d.scc[u] : 0;
        const int32_t f0 = act ? fb[u] : 0, f1 = act ? fb[u + 1] : 0;
        const int32_t b0 = act ? bb[u] : 0, b1 = act ? bb[u + 1] : 0;
        auto count = [&](int32_t, uint64_t pay, int32_t dir, int64_t j) -> int64_t {
            uint64_t key;
            return d(pay, dir, j, key) ? 1 : 0;
        };
        const int64_t m = waveRowsAll(act, u, a, 0, f0, f1, count) + waveRowsAll(act, u, a, 1, b0, b1, count);
        int64_t off = blockReserve((int)m, cnt);
        auto emit = [&](int32_t, uint64_t pay, int32_t dir, int64_t j) -> int64_t {
            uint64_t key;
            if (d(pay, dir, j, key)) keys[off++] = key;
            return 0;
        };
        waveRowsAll(act, u, a, 0, f0, f1, emit);
        waveRowsAll(act, u, a, 1, b0, b1, emit);
    }
}


__global__ void kDagTKeys(const int32_t* __restrict__ src, const int32_t* __restrict__ dst, int64_t nd, int B,
                          uint64_t* __restrict__ tk) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r < nd) tk[r] = ((uint64_t)(uint32_t)dst[r] << B) | (uint32_t)src[r];
}

// CSR offsets from the sorted row ids of nd arcs (id(i) = ids[i], or
// keys[i] >> B): thread i in [0, nd] writes beg[v] = i for every v in
// (id(i - 1), id(i)] (id(nd) = V), i.e. the first arc of every row - one
// pass over the arcs instead of a binary search per row.  A run of more than
// BEG_RUN empty rows is not written by its one thread: it goes on a list
// (gaps[0] = count, then (first, last, i) triples) that kBegGaps fills with
// a block per run, so a sparse id range costs no serial tail.  (The list
// slots are reserved once per block: a condensation's sparse source ids put
// a long run in most waves.)
constexpr int BEG_RUN = 32;
__global__ void kBegScatter(const int32_t* __restrict__ ids, const uint64_t* __restrict__ keys, int B, int64_t nd,
                            int32_t V, int32_t* __restrict__ beg, int64_t* gaps) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    auto id = [&](int64_t k) -> int64_t { return ids ? (int64_t)ids[k] : (int64_t)(keys[k] >> B); };
    int64_t cur = 0, prev = 0;
    if (i <= nd) {
        cur = i < nd ? id(i) : (int64_t)V;
        prev = i > 0 ? id(i - 1) : -1;
    }
    const bool longrun = i <= nd && cur - prev > BEG_RUN;
    const int64_t pos = blockReserve(longrun ? 1 : 0, gaps);  // (one atomic per block)
    if (longrun) {
        gaps[1 + 3 * pos] = prev + 1;
        gaps[2 + 3 * pos] = cur;
        gaps[3 + 3 * pos] = i;
    } else if (i <= nd) {
        for (int64_t v = prev + 1; v <= cur; ++v) beg[v] = (int32_t)i;
    }
}
__global__ void kBegGaps(const int64_t* __restrict__ gaps, int32_t* __restrict__ beg) {
    const int64_t n = gaps[0];
    for (int64_t g = blockIdx.x; g < n; g += gridDim.x) {
        const int64_t a = gaps[1 + 3 * g], b = gaps[2 + 3 * g];
        const int32_t val = (int32_t)gaps[3 + 3 * g];
        for (int64_t v = a + threadIdx.x; v <= b; v += blockDim.x) beg[v] = val;
    }
}

// indeg[v] = beg[v + 1] - beg[v]; with keys, also src[i] = the low B bits
// of keys[i] (the in-arc CSR's sources).
__global__ void kDegAndSrc(const int32_t* __restrict__ beg, int32_t V, int32_t* __restrict__ indeg,
                           const uint64_t* __restrict__ keys, int64_t nd, int B, int32_t* __restrict__ src) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < V) indeg[i] = beg[i + 1] - beg[i];
    if (i < nd) src[i] = (int32_t)(keys[i] & ((1ULL << B) - 1));
}

__global__ void kRunFlags(const uint64_t* __restrict__ keys, int64_t n, int32_t* __restrict__ flag) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) flag[i] = (i == 0 || keys[i] != keys[i - 1]) ? 1 : 0;
}

__global__ void kRunStarts(const int32_t* __restrict__ flag, const int64_t* __restrict__ pos, int64_t n,
                           int32_t* __restrict__ start) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && flag[i]) start[pos[i]] = (int32_t)i;
}

__global__ void kDagWrite(const uint64_t* __restrict__ keys, const int32_t* __restrict__ start, int64_t nd,
                          int64_t ncross, int B, int32_t* __restrict__ src, int32_t* __restrict__ dst,
                          int32_t* __restrict__ mult) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nd) return;
    const int64_t i = start[r], j = r + 1 < nd ? start[r + 1] : ncross;
    const uint64_t k = keys[i];
    const int32_t a = (int32_t)(k >> B), b = (int32_t)(k & ((1ULL << B) - 1));
    src[r] = a;
    dst[r] = b;
    mult[r] = (int32_t)(j - i);
}

// ---- ranks (Kahn levels) ----
__global__ void kRankSeeds(const int32_t* __restrict__ scc, const int32_t* __restrict__ indeg, int32_t V,
                           int64_t* rank, int32_t* __restrict__ out, int64_t* cnt) {
    const int64_t v0 = (int64_t)blockIdx.x * caSpan(blockDim.x) + threadIdx.x;
    unsigned seed = 0;
    int k = 0;
#pragma unroll
    for (int t = 0; t < CA_ITEMS; ++t) {
        const int64_t v = v0 + (int64_t)t * blockDim.x;
        if (v < V) {
            rank[v] = -1;
            if (scc[v] == (int32_t)v && indeg[v] == 0) {
                seed |= 1u << t;
                ++k;
            }
        }
    }
    int64_t pos = blockReserve(k, cnt);
#pragma unroll
    for (int t = 0; t < CA_ITEMS; ++t)
        if ((seed >> t) & 1) out[pos++] = (int32_t)(v0 + (int64_t)t * blockDim.x);
}

// Kahn step work on DAG arc j: one in-degree decrement of its head; the
// decrement that reaches zero queues the head.  Arcs into the hub (the SCC
// with the most DAG in-arcs, the giant component of a power-law graph) are
// summed per wave first, so the hub's counter takes one atomic per wave
// instead of one per arc.
struct KahnF {
    const int32_t* ddst;
    int32_t* indeg;
    int32_t* qout;
    int64_t* cnt;
    int32_t hub;
    int* hub_lds;  // set by kKahnStep: the block sums its hub decrements here
    __device__ int64_t operator()(int32_t, uint64_t, int32_t, int64_t j) const {
        const int32_t t = ddst[j];
        const bool hot = t == hub;
        const uint64_t hm = __ballot(hot);
        bool ready = false;
        if (hot) {
            if (laneId() == __ffsll((long long)hm) - 1) {
                const int c = __popcll(hm);
                if (hub_lds) atomicAdd(hub_lds, c);
                else ready = atomicSub(&indeg[t], c) == c;
            }
        } else {
            ready = atomicSub(&indeg[t], 1) == 1;
        }
        const int64_t pos = wavePush(cnt, ready);
        if (ready) qout[pos] = t;
        return 0;
    }
};

// hub = the id with the largest in-degree (packed (indeg, id) maximum)
__global__ void kHub(const int32_t* __restrict__ indeg, int32_t V, int64_t* out) {
    unsigned long long m = 0;
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < V; v += (int64_t)gridDim.x * blockDim.x)
        m = std::max<unsigned long long>(m, ((unsigned long long)(uint32_t)indeg[v] << 32) | (uint32_t)v);
    m = blockMaxU64(m);
    if (threadIdx.x == 0 && m) atomicMax((unsigned long long*)out, m);
}

// One Kahn level over the queue qin, whose length the previous level left
// in *nin (read on the device, so several levels run per host readback);
// block 0 clears *cz, the counter of the level after next.  The grid strides
// over the queue, uniformly per wave.
__global__ void kKahnStep(const int32_t* __restrict__ qin, const int64_t* nin, const int32_t* __restrict__ dbeg,
                          int64_t* rank, int32_t L, KahnF f, HeavyList H, int64_t* cz) {
    // the hub's decrements: one LDS sum per block and one atomic per block
    // at the end (a level of a power-law condensation holds many arcs into
    // the giant SCC, and one counter hit by every wave serialises them)
    __shared__ int hub_sum;
    if (threadIdx.x == 0) hub_sum = 0;
    __syncthreads();
    f.hub_lds = &hub_sum;
    if (blockIdx.x == 0 && threadIdx.x == 0) *cz = 0;
    const int64_t n = *nin;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t base = (int64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~(WAVE - 1)); base < n; base += stride) {
        const int64_t i = base + laneId();
        const bool act = i < n;
        const int32_t s = act ? qin[i] : 0;
        if (act) rank[s] = ((int64_t)L << 32) | (int64_t)(uint32_t)s;
        const int32_t b = act ? dbeg[s] : 0, e = act ? dbeg[s + 1] : 0;
        waveRowsH(act, s, 0, 0, b, e, H, f);
    }
    __syncthreads();
    if (threadIdx.x == 0 && hub_sum > 0 && atomicSub(&f.indeg[f.hub], hub_sum) == hub_sum) {
        const int64_t pos = atomicAdd((unsigned long long*)f.cnt, 1ULL);
        f.qout[pos] = f.hub;
    }
}

// The heavy rows of a Kahn level (kHeavyRows' walk: the listed rows' arcs
// concatenated, a contiguous range per block) with the level kernel's
// aggregation: the block walks its range in block-wide steps, reserves the
// step's ready heads with one atomic, and sums its hub decrements in LDS (a
// power-law condensation's big SCC can hold 10^5 out-arcs in one row, whose
// heads would otherwise hit the queue counter once per wave).
__global__ void kKahnHeavy(HeavyList H, KahnF f) {
    const uint64_t c = (uint64_t)ldAcq(&H.ctl[0]);
    if (c == 0) return;
    __shared__ int hub_sum;
    if (threadIdx.x == 0) hub_sum = 0;
    __syncthreads();
    const int64_t n = (int64_t)(c >> HEAVY_SHIFT), T = (int64_t)(c & HEAVY_MASK);
    const int64_t per = (T + gridDim.x - 1) / gridDim.x;
    const int64_t p0 = (int64_t)blockIdx.x * per, p1 = p0 + per < T ? p0 + per : T;
    int64_t k = 0;
    for (int64_t pb = p0; pb < p1; pb += blockDim.x) {
        const int64_t p = pb + threadIdx.x;
        bool ready = false;
        int32_t t = 0;
        if (p < p1) {
            if (!(p >= H.off[k] && p < H.off[k] + (H.e[k] - H.b[k]))) {
                int64_t lo = 0, hi = n - 1;  // last k with off[k] <= p
                while (lo < hi) {
                    const int64_t m = (lo + hi + 1) >> 1;
                    if (H.off[m] <= p) lo = m;
                    else hi = m - 1;
                }
                k = lo;
            }
            t = f.ddst[H.b[k] + (p - H.off[k])];
            if (t == f.hub) atomicAdd(&hub_sum, 1);
            else ready = atomicSub(&f.indeg[t], 1) == 1;
        }
        const int64_t pos = blockReserve(ready ? 1 : 0, f.cnt);
        if (ready) f.qout[pos] = t;
    }
    __syncthreads();
    if (threadIdx.x == 0 && hub_sum > 0 && atomicSub(&f.indeg[f.hub], hub_sum) == hub_sum) {
        const int64_t pos = atomicAdd((unsigned long long*)f.cnt, 1ULL);
        f.qout[pos] = f.hub;
    }
    // the last block to finish clears the list (as kHeavyRows)
    if (threadIdx.x == 0 && atomicAdd((unsigned long long*)&H.ctl[1], 1ULL) == gridDim.x - 1) {
        H.ctl[0] = 0;
        H.ctl[1] = 0;
    }
}

namespace {

int64_t compactActive(girhip_graph* h, const int32_t* in, int64_t n, int32_t* out) {
    const int blk = (int)h->prm.block;
    fillI64(h, h->dcnt.p + 2, 0, 1);
    launch(h, "scc_compact", kCompactActive, (unsigned)((n + caSpan(blk) - 1) / caSpan(blk)), blk, 0, in, n,
           (const int64_t*)nullptr, false, h->scc.p, out, h->dcnt.p + 2, (const int32_t*)nullptr,
           (const int32_t*)nullptr);
    int64_t c;
    readCounters(h, h->dcnt.p + 2, 1, &c);
    return c;
}

// Trim-1 rounds over list (in place: the list is re-compacted between rounds).
int64_t trimRounds(girhip_graph* h, int32_t*& list, int32_t*& alt, int64_t n, int rounds) {
    const int blk = (int)h->prm.block;
    for (int r = 0; r < rounds && n > 0; ++r) {
        fillI64(h, h->dcnt.p + 3, 0, 1);
        launch(h, "scc_trim1", kTrim1, gridFor(n, blk), blk, 0, (const int32_t*)list, n, (const int64_t*)nullptr,
               (const int32_t*)h->fw_beg.p, (const int32_t*)h->fw_adj.p, (const int32_t*)h->bw_beg.p,
               (const int32_t*)h->bw_adj.p, h->scc.p, h->dcnt.p + 3, (const int64_t*)nullptr);
        if (h->prm.scc_trim2 != 0)
            launch(h, "scc_trim2", kTrim2, gridFor(n, blk), blk, 0, (const int32_t*)list, n, (const int64_t*)nullptr,
                   (const int32_t*)h->fw_beg.p, (const int32_t*)h->fw_adj.p, (const int32_t*)h->bw_beg.p,
                   (const int32_t*)h->bw_adj.p, h->scc.p, h->dcnt.p + 3);
        int64_t trimmed;
        readCounters(h, h->dcnt.p + 3, 1, &trimmed);
        if (trimmed == 0) break;
        n = compactActive(h, list, n, alt);
        std::swap(list, alt);
    }
    return n;
}

// The initial active list (vertices labelled -1; those without out- or
// in-arcs labelled as their own SCCs on the way), the trim rounds over it
// and, with pivot, the pivot key over what remains.  The list lengths stay
// on the device: each kernel is launched for an upper bound (V) over a
// bounded grid and reads the current length, so the phase costs one host
// readback (length + key) instead of two per round.  list ends as the
// remaining active list; returns its length.
int64_t trimPhase(girhip_graph* h, int32_t*& list, int32_t*& alt, int rounds, bool pivot, int64_t* key) {
    const int32_t V = h->V;
    const int blk = (int)h->prm.block;
    Scratch& S = scratch(h);
    rounds = std::max(0, std::min(rounds, 64));
    // c[0..rounds]: list lengths after each compaction, then the pivot key,
    // then the rounds' flags
    S.tcnt.ensure((size_t)(2 * rounds + 2));
    int64_t* c = S.tcnt.p;
    int64_t* pk = c + rounds + 1;
    int64_t* fl = pk + 1;
    fillI64(h, c, 0, 2 * rounds + 2);
    const int32_t* fb = h->fw_beg.p;
    const int32_t* bb = h->bw_beg.p;
    const unsigned gc = std::min<unsigned>((unsigned)((V + caSpan(blk) - 1) / caSpan(blk)), 2048);
    const unsigned gt = std::min<unsigned>(gridFor(V, blk), 4096);
    launch(h, "scc_compact", kCompactActive, gc, blk, 0, (const int32_t*)nullptr, (int64_t)V, (const int64_t*)nullptr,
           true, h->scc.p, list, c, fb, bb);
    for (int r = 0; r < rounds; ++r) {
        launch(h, "scc_trim1", kTrim1, gt, blk, 0, (const int32_t*)list, (int64_t)V, (const int64_t*)(c + r), fb,
               (const int32_t*)h->fw_adj.p, bb, (const int32_t*)h->bw_adj.p, h->scc.p, fl + r,
               (const int64_t*)(r > 0 ? fl + r - 1 : nullptr));
        if (h->prm.scc_trim2 != 0)
            launch(h, "scc_trim2", kTrim2, gt, blk, 0, (const int32_t*)list, (int64_t)V, (const int64_t*)(c + r), fb,
                   (const int32_t*)h->fw_adj.p, bb, (const int32_t*)h->bw_adj.p, h->scc.p, fl + r);
        launch(h, "scc_compact", kCompactActive, gc, blk, 0, (const int32_t*)list, (int64_t)V,
               (const int64_t*)(c + r), false, h->scc.p, alt, c + r + 1, (const int32_t*)nullptr,
               (const int32_t*)nullptr);
        std::swap(list, alt);
    }
    if (pivot)
        launch(h, "scc_pivot", kPivot, reduceGrid(V, blk), blk, 0, (const int32_t*)list, (int64_t)V,
               (const int64_t*)(c + rounds), fb, bb, (unsigned long long*)pk);
    int64_t r[2];
    readCounters(h, c + rounds, 2, r);
    *key = r[1];
    return r[0];
}

}  // namespace

namespace {
// The SCC/DAG build runs its kernels (and the shared reachability engine's)
// at the scc_block launch size: measured 3.8 ms vs 4.1 ms at 256 threads for
// RMAT-22 (BFS and SSSP are best at 256).
struct SccBlock {
    girhip_graph* h;
    double old;
    explicit SccBlock(girhip_graph* g) : h(g), old(g->prm.block) {
        const double b = g->prm.scc_block;
        if (b >= 64 && b <= 1024) g->prm.block = b;
    }
    ~SccBlock() { h->prm.block = old; }
};
}  // namespace

void runScc(girhip_graph* h) {
    const int32_t V = h->V;
    h->gc_ok = false;
    h->scc.ensure((size_t)(V > 0 ? V : 1));
    if (V == 0) {
        h->n_scc = 0;
        h->n_scc_pending = false;
        return;
    }
    fillI32(h, h->scc.p, -1, V);
    runSccActive(h, true);
}

// SCCs of the vertices labelled -1 ("active"); every other label stays.  The
// active set must be a union of whole SCCs of the current graph (all of it
// for a build; the affected SCCs for a repair after an update batch).  With
// cert (a build), the forward and backward sweeps from the pivot record
// their parents: restricted to the pivot's SCC they are its out-tree and
// in-tree (every vertex on a tree path from the pivot to a member reaches
// the member, which reaches the pivot), the certificate update batches keep.
void runSccActive(girhip_graph* h, bool cert) {
    const SccBlock sb(h);
    const int32_t V = h->V;
    const int blk = (int)h->prm.block;
    Scratch& S = scratch(h);
    S.q[2].ensure(V);
    S.iv[0].ensure(V);
    S.iv[1].ensure(V);
    S.iv[2].ensure(V);
    S.bits[3].ensure(((size_t)V + 63) / 64);
    int32_t* list = S.q[2].p;
    int32_t* alt = S.iv[0].p;
    int32_t* color = S.iv[1].p;
    int32_t* stamp = S.iv[2].p;
    ReachBufs B = reachBufs(h);
    int64_t key = 0;
    int64_t n = trimPhase(h, list, alt, (int)h->prm.scc_trim_rounds, h->prm.scc_pivot != 0, &key);

    // FW-BW from the pivot
    if (n > 0 && h->prm.scc_pivot != 0 && n > (int64_t)h->prm.scc_color_min) {
        const int32_t pivot = (int32_t)(0x7fffffff - (key & 0x7fffffffLL));
        const ActiveOnly act{h->scc.p};
        if (cert) {
            h->gc_fpar.ensure((size_t)V);
            h->gc_bpar.ensure((size_t)V);
            fillI32(h, h->gc_fpar.p + pivot, pivot, 1);
            fillI32(h, h->gc_bpar.p + pivot, pivot, 1);
        }
        if (h->prm.scc_pair != 0 && 2 * maxDegree(h) <= HEAVY_T) {
            // both sweeps in shared launches: forward into bits[3], backward into B.vis
            S.pq[0].ensure(V);
            S.pq[1].ensure(V);
            ReachBufs B1 = B, B2 = B;
            B1.vis = S.bits[3].p;
            B2.q[0] = S.pq[0].p;
            B2.q[1] = S.pq[1].p;
            fillU64(h, B1.vis, 0, B.W);
            fillU64(h, B2.vis, 0, B.W);
            if (cert)
                runReachPair(h, h->fw_beg.p, h->fw_adj.p, h->bw_beg.p, h->bw_adj.p, pivot, B1,
                             WriteParent{h->gc_fpar.p}, B2, WriteParent{h->gc_bpar.p}, n, h->E, act);
            else
                runReachPair(h, h->fw_beg.p, h->fw_adj.p, h->bw_beg.p, h->bw_adj.p, pivot, B1, NoVisit{}, B2,
                             NoVisit{}, n, h->E, act);
        } else {
            fillU64(h, B.vis, 0, B.W);
            fillI32(h, B.q[0], pivot, 1);
            if (cert)
                runReach(h, h->fw_beg.p, h->fw_adj.p, h->bw_beg.p, h->bw_adj.p, B, 1, 0, n, h->E, act,
                         WriteParent{h->gc_fpar.p});
            else
                runReach(h, h->fw_beg.p, h->fw_adj.p, h->bw_beg.p, h->bw_adj.p, B, 1, 0, n, h->E, act, NoVisit{});
            HIPCHK(hipMemcpyAsync(S.bits[3].p, B.vis, sizeof(uint64_t) * B.W, hipMemcpyDeviceToDevice, h->st));
            fillU64(h, B.vis, 0, B.W);
            fillI32(h, B.q[0], pivot, 1);
            if (cert)
                runReach(h, h->bw_beg.p, h->bw_adj.p, h->fw_beg.p, h->fw_adj.p, B, 1, 0, n, h->E, act,
                         WriteParent{h->gc_bpar.p});
            else
                runReach(h, h->bw_beg.p, h->bw_adj.p, h->fw_beg.p, h->fw_adj.p, B, 1, 0, n, h->E, act, NoVisit{});
        }
        if (cert) {  // the trees hold; the pivot's label is read by the first batch that needs it
            h->gc_ok = true;
            h->gc_pivot = pivot;
            h->gc_label = -1;
        }
        fillI64(h, h->dcnt.p + 5, INT64_MAX, 1);
        launch(h, "scc_fwbw_min", kFwBwMin, reduceGrid(n, blk), blk, 0, (const int32_t*)list, n,
               (const uint64_t*)S.bits[3].p, (const uint64_t*)B.vis, h->dcnt.p + 5);
        launch(h, "scc_fwbw_assign", kFwBwAssign, gridFor(n, blk), blk, 0, (const int32_t*)list, n,
               (const uint64_t*)S.bits[3].p, (const uint64_t*)B.vis, (const int64_t*)h->dcnt.p + 5, h->scc.p);
        n = compactActive(h, list, n, alt);
        std::swap(list, alt);
        n = trimRounds(h, list, alt, n, (int)h->prm.scc_trim_rounds);
    }

    // colouring on the remainder
    fillI32(h, stamp, 0, V);
    int32_t it = 0;
    while (n > 0) {
        launch(h, "scc_color_init", kColorInit, gridFor(n, blk), blk, 0, (const int32_t*)list, n, color);
        // propagate to a fixpoint, starting from every active vertex
        HIPCHK(hipMemcpyAsync(B.q[0], list, sizeof(int32_t) * n, hipMemcpyDeviceToDevice, h->st));
        int64_t nq = n;
        int cur = 0;
        while (nq > 0) {
            ++it;
            fillI64(h, h->dcnt.p + 6, 0, 1);
            const ColorPropF f{h->fw_adj.p, h->scc.p, color, stamp, it, B.q[cur ^ 1], h->dcnt.p + 6};
            launch(h, "scc_color_prop", kColorProp, gridFor(nq, blk), blk, 0, (const int32_t*)B.q[cur], nq,
                   (const int32_t*)h->fw_beg.p, f, heavyList(h));
            launchHeavyStore(h, "scc_color_prop_heavy", f);
            readCounters(h, h->dcnt.p + 6, 1, &nq);
            cur ^= 1;
        }
        // roots seed one backward sweep restricted to their colour classes
        fillI64(h, h->dcnt.p + 7, 0, 1);
        launch(h, "scc_color_roots", kColorRoots, gridFor(n, blk), blk, 0, (const int32_t*)list, n,
               (const int32_t*)color, B.q[0], h->dcnt.p + 7);
        int64_t nr;
        readCounters(h, h->dcnt.p + 7, 1, &nr);
        fillU64(h, B.vis, 0, B.W);
        runReach(h, h->bw_beg.p, h->bw_adj.p, h->fw_beg.p, h->fw_adj.p, B, nr, 0, n, h->E,
                 SameColour{h->scc.p, color}, NoVisit{});
        launch(h, "scc_color_assign", kColorAssign, gridFor(n, blk), blk, 0, (const int32_t*)list, n,
               (const uint64_t*)B.vis, (const int32_t*)color, h->scc.p);
        n = compactActive(h, list, n, alt);
        std::swap(list, alt);
        n = trimRounds(h, list, alt, n, 1);
    }
    // the count stays on the device until someone needs it (the DAG pass
    // reads it with its first count)
    fillI64(h, h->dcnt.p + DC_SCC, 0, 1);
    launch(h, "count_roots", kCountSelf, reduceGrid(V, blk), blk, 0, (const int32_t*)h->scc.p, V,
           h->dcnt.p + DC_SCC);
    h->n_scc_pending = true;
}

int32_t sccCount(girhip_graph* h) {
    if (h->n_scc_pending) {
        int64_t ns;
        readCounters(h, h->dcnt.p + DC_SCC, 1, &ns);
        h->n_scc = (int32_t)ns;
        h->n_scc_pending = false;
    }
    return h->n_scc;
}

void runDag(girhip_graph* h) {
    const SccBlock sb(h);
    const int32_t V = h->V;
    const int blk = (int)h->prm.block;
    Scratch& S = scratch(h);
    const int B = bitsFor((uint64_t)(V > 1 ? V - 1 : 1));
    h->mem_ok = false;  // member lists follow the labels
    int64_t ncross = 0;
    if (h->E > 0) {
        S.keys[0].ensure((size_t)h->E);
        S.gbits.ensure(((size_t)V + 63) / 64);
        fillI64(h, h->dcnt.p + 9, 0, 1);
        fillI64(h, h->dcnt.p + 14, 0, 1);
        launch(h, "dag_top", kMaxDegVertex, reduceGrid(V, 256), 256, 0, (const int32_t*)h->fw_beg.p,
               (const int32_t*)h->bw_beg.p, V, (unsigned long long*)(h->dcnt.p + 14));
        launch(h, "dag_label_bits", kLabelBits, boundedGrid(((int64_t)V + 63) / 64 * 64, 256), 256, 0,
               (const int32_t*)h->scc.p, V, (const unsigned long long*)(h->dcnt.p + 14), S.gbits.p,
               (int32_t*)(h->dcnt.p + 15));
        // the vertices outside the big SCC, then their rows' crossing arcs
        S.iv[0].ensure((size_t)V);
        fillI64(h, h->dcnt.p + 16, 0, 1);
        launch(h, "dag_nonbig", kNonBig, boundedGrid(((int64_t)V + 63) / 64, 512), 512, 0, (const uint64_t*)S.gbits.p,
               V, S.iv[0].p, h->dcnt.p + 16);
        const DagArc d{h->fw_adj.p, h->bw_adj.p, h->scc.p, S.gbits.p, 0, B};
        launch(h, "dag_keys", kDagKeysRows, std::min<unsigned>(gridFor(V, blk), 4096), blk, 0,
               (const int32_t*)S.iv[0].p, (const int64_t*)(h->dcnt.p + 16), (const int32_t*)h->fw_beg.p,
               (const int32_t*)h->bw_beg.p, d, (const int32_t*)(h->dcnt.p + 15), S.keys[0].p, h->dcnt.p + 9);
        int64_t rr[DC_SCC - 9 + 1];
        readCounters(h, h->dcnt.p + 9, DC_SCC - 9 + 1, rr);
        ncross = rr[0];
        if (h->n_scc_pending) {
            h->n_scc = (int32_t)rr[DC_SCC - 9];
            h->n_scc_pending = false;
        }
    }
    int64_t nd = 0;
    if (ncross > 0) {
        radixSort(h, S.keys[0].p, nullptr, ncross, 2 * B);
        S.iv[0].ensure((size_t)ncross);
        S.off2.ensure((size_t)ncross + 1);
        launch(h, "dag_flags", kRunFlags, gridFor(ncross, blk), blk, 0, (const uint64_t*)S.keys[0].p, ncross,
               S.iv[0].p);
        scanI32toI64(h, S.iv[0].p, S.off2.p, ncross);
        readCounters(h, S.off2.p + ncross, 1, &nd);
        S.q[2].ensure((size_t)nd);
        launch(h, "dag_starts", kRunStarts, gridFor(ncross, blk), blk, 0, (const int32_t*)S.iv[0].p,
               (const int64_t*)S.off2.p, ncross, S.q[2].p);
        h->dag_src.ensure((size_t)nd);
        h->dag_dst.ensure((size_t)nd);
        h->dag_mult.ensure((size_t)nd);
        launch(h, "dag_write", kDagWrite, gridFor(nd, blk), blk, 0, (const uint64_t*)S.keys[0].p,
               (const int32_t*)S.q[2].p, nd, ncross, B, h->dag_src.p, h->dag_dst.p, h->dag_mult.p);
    }
    dagFinish(h, nd, B, true);
}

// CSR offsets (V + 1) of nd arcs sorted by row id (ids, or keys >> B).
static void begFromIds(girhip_graph* h, const char* name, const int32_t* ids, const uint64_t* keys, int B, int64_t nd,
                       int32_t* beg) {
    const int32_t V = h->V;
    const int blk = (int)h->prm.block;
    Scratch& S = scratch(h);
    // at most one long run per arc boundary, and at most V / BEG_RUN of them
    const int64_t cap = std::min<int64_t>(nd + 1, (int64_t)V / BEG_RUN + 1);
    S.off.ensure((size_t)(1 + 3 * cap));
    fillI64(h, S.off.p, 0, 1);
    launch(h, name, kBegScatter, gridFor(nd + 1, blk), blk, 0, ids, keys, B, nd, V, beg, S.off.p);
    launch(h, "beg_gaps", kBegGaps, 512, 256, 0, (const int64_t*)S.off.p, beg);
}

// With the arcs (dag_src / dst / mult, nd of them, sorted by (src, dst))
// written: the in-arc CSR, the out-arc offsets and, with ranks, the Kahn
// levels.
static void dagIndex(girhip_graph* h, int64_t nd, int B, bool ranks);

void dagFinish(girhip_graph* h, int64_t nd, int B, bool ranks) {
    const int blk = (int)h->prm.block;
    if (nd > 0) {
        // the arcs sorted by destination: the in-arc keys (dst << B | src),
        // written in arc order, are already ordered by src, so a stable sort
        // on the dst bits alone orders them by (dst, src)
        h->dag_tk.ensureSlack((size_t)nd);
        launch(h, "dag_tkeys", kDagTKeys, gridFor(nd, blk), blk, 0, (const int32_t*)h->dag_src.p,
               (const int32_t*)h->dag_dst.p, nd, B, h->dag_tk.p);
        radixSort(h, h->dag_tk.p, nullptr, nd, 2 * B, B);
    }
    dagIndex(h, nd, B, ranks);
}

// With the arcs (sorted by (src, dst)) and the in-arc keys (sorted by (dst,
// src)) in place: the in-arc CSR, the in-degrees as its row lengths (no
// contended atomics), the out-arc offsets and, with ranks, the Kahn levels.
static void dagIndex(girhip_graph* h, int64_t nd, int B, bool ranks) {
    const int32_t V = h->V;
    const int blk = (int)h->prm.block;
    Scratch& S = scratch(h);
    h->rank.ensure((size_t)(V > 0 ? V : 1));
    h->dag_beg.ensure((size_t)V + 1);
    S.iv[2].ensure((size_t)V + 1);
    int32_t* ideg = S.iv[2].p;
    if (nd > 0) {
        h->dag_in_src.ensureSlack((size_t)nd);
        h->dag_in_beg.ensure((size_t)V + 1);
        begFromIds(h, "dag_in", nullptr, h->dag_tk.p, B, nd, h->dag_in_beg.p);
        launch(h, "dag_indeg", kDegAndSrc, gridFor(std::max<int64_t>(nd, V), blk), blk, 0,
               (const int32_t*)h->dag_in_beg.p, V, ideg, (const uint64_t*)h->dag_tk.p, nd, B, h->dag_in_src.p);
    } else {
        fillI32(h, ideg, 0, (int64_t)V + 1);
        h->dag_in_beg.ensure((size_t)V + 1);
        h->dag_in_src.ensure(1);
        fillI32(h, h->dag_in_beg.p, 0, (int64_t)V + 1);
    }
    h->n_dag = nd;
    // DAG CSR offsets by source SCC (arcs are sorted by src)
    begFromIds(h, "dag_beg", h->dag_src.p, nullptr, 0, nd, h->dag_beg.p);
    if (!ranks) return;
    // Kahn levels
    ReachBufs R = reachBufs(h);
    fillI64(h, h->dcnt.p + 12, 0, 1);
    launch(h, "dag_hub", kHub, reduceGrid(V, blk), blk, 0, (const int32_t*)ideg, V, h->dcnt.p + 12);
    int64_t hubk;
    readCounters(h, h->dcnt.p + 12, 1, &hubk);
    const int32_t hub = hubk ? (int32_t)(hubk & 0xffffffffLL) : -1;
    // level counters: a ring of KAHN_RING (level L reads slot L, fills
    // slot L + 1, clears slot L + 2); KAHN_CHUNK levels per readback
    constexpr int KAHN_RING = 4, KAHN_CHUNK = 4;
    Scratch& S2 = scratch(h);
    S2.rcnt.ensure(KAHN_RING);
    int64_t* kc = S2.rcnt.p;
    fillI64(h, kc, 0, KAHN_RING);
    launch(h, "rank_seeds", kRankSeeds, (unsigned)(((int64_t)V + caSpan(blk) - 1) / caSpan(blk)), blk, 0,
           (const int32_t*)h->scc.p, (const int32_t*)ideg, V, h->rank.p, R.q[0], kc);
    int cur = 0;
    int32_t L = 0;
    while (true) {
        for (int t = 0; t < KAHN_CHUNK; ++t, ++L) {
            int64_t* cin = kc + L % KAHN_RING;
            const KahnF f{h->dag_dst.p, ideg, R.q[cur ^ 1], kc + (L + 1) % KAHN_RING, hub, nullptr};
            launch(h, "rank_kahn", kKahnStep, 1024, blk, 0, (const int32_t*)R.q[cur], (const int64_t*)cin,
                   (const int32_t*)h->dag_beg.p, h->rank.p, L, f, heavyList(h), kc + (L + 2) % KAHN_RING);
            launch(h, "rank_kahn_heavy", kKahnHeavy, 256, 256, 0, heavyList(h), f);
            cur ^= 1;
        }
        int64_t nq;
        readCounters(h, kc + L % KAHN_RING, 1, &nq);
        if (nq == 0) break;
    }
}

// ---------------------------------------------------------------------------
// Repair of the SCC layer after an update batch, instead of a rebuild.
// Every change (u, v) is classified by the labels and ranks the layer held
// before the batch (the reference's cases, gir.cpp:3007-3036 for inserts and
// 3062-3131 for deletes, taken over a whole batch as applyBatch does,
// gir.cpp:3634-4111):
//   insert inside an SCC                -> nothing;
//   insert between SCCs in rank order   -> condensation arc +1;
//   insert against the rank order       -> condensation arc +1, and it may
//                                          close a cycle (SCCs merge);
//   delete between SCCs                 -> condensation arc -1;
//   delete inside an SCC                -> the SCC may split.
// The multiplicity changes are merged into the condensation first (a sort of
// its arcs and the changes by key, summed per key).  With no merge or split
// possible that is all: the partition stands and every arc still goes up
// in rank.  Otherwise the SCCs that may change - each SCC that lost an inner
// arc, and each SCC both reachable from the head SCC of a rank-violating
// insert and reaching the tail SCC of one, over the updated condensation with
// the ranks outside [min head rank, max tail rank] pruned (a cycle through
// such an insert stays inside that window) - are relabelled -1 and the SCC
// pipeline runs on them alone: a union of whole old SCCs that holds every
// new cycle holds every new SCC that differs.  A changed partition rebuilds
// the condensation and its ranks; rank-violating inserts that closed no
// cycle re-rank it.
namespace {

enum : int { IR_DELTA = 0, IR_BACK, IR_DIRTY, IR_MINR, IR_MAXR, IR_AFF, IR_CUTF, IR_CUTB, IR_GDEL, IR_GREG, IR_N };

// The big SCC's certificate, as the classification sees it: its label g
// (-1: none), and, when its trees are valid (fpar != nullptr), where the
// heads / tails of deleted tree arcs go.
struct IrCert {
    int32_t g;
    const int32_t* fpar;
    const int32_t* bpar;
    int32_t* cutf;
    int32_t* cutb;
};

__global__ void kIrClassify(const uint64_t* __restrict__ ik, int64_t ni, const uint64_t* __restrict__ dk, int64_t nd,
                            int B, const int32_t* __restrict__ scc, const int64_t* __restrict__ rank,
                            uint64_t* __restrict__ dkeys, uint32_t* __restrict__ dvals, int32_t* __restrict__ bsu,
                            int32_t* __restrict__ bsv, uint8_t* __restrict__ dirty, int64_t* cnt,
                            const uint8_t* __restrict__ dbypass, IrCert gc) {
    const uint64_t mask = (1ULL << B) - 1;
    const int64_t n = ni + nd, stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t base = (int64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~(WAVE - 1)); base < n; base += stride) {
        const int64_t i = base + laneId();
        const bool act = i < n, ins = i < ni;
        int32_t su = 0, sv = 0, u = 0, v = 0;
        if (act) {
            const uint64_t k = ins ? ik[i] : dk[i - ni];
            u = (int32_t)(k >> B);
            v = (int32_t)(k & mask);
            su = scc[u];
            sv = scc[v];
        }
        const bool cross = act && su != sv;
        const bool back = cross && ins && rank[su] >= rank[sv];
        // an inner delete whose tail still reaches its head splits nothing;
        // one inside the big SCC is checked against its trees instead
        const bool inner_del = act && !ins && su == sv && !(dbypass && dbypass[i - ni]);
        const bool gdel = inner_del && su == gc.g;
        if (inner_del && !gdel) dirty[su] = 1;
        const bool cutf = gdel && gc.fpar && gc.fpar[v] == u;
        const bool cutb = gdel && gc.bpar && gc.bpar[u] == v;
        const int64_t pf = wavePush(cnt + IR_CUTF, cutf);
        if (cutf) gc.cutf[pf] = v;
        const int64_t pb = wavePush(cnt + IR_CUTB, cutb);
        if (cutb) gc.cutb[pb] = u;
        if (__ballot(gdel) && laneId() == 0) cnt[IR_GDEL] = 1;
        const int64_t dpos = wavePush(cnt + IR_DELTA, cross);
        if (cross) {
            dkeys[dpos] = ((uint64_t)(uint32_t)su << B) | (uint32_t)sv;
            dvals[dpos] = ins ? 1u : 0xffffffffu;  // +1 / -1
        }
        const int64_t bpos = wavePush(cnt + IR_BACK, back);
        if (back) {
            bsu[bpos] = su;
            bsv[bpos] = sv;
        }
        // rank window of the rank-violating inserts (one atomic per wave)
        long long mn = back ? rank[sv] : LLONG_MAX, mx = back ? rank[su] : -1;
        for (int o = WAVE / 2; o > 0; o >>= 1) {
            const long long a = __shfl_xor(mn, o), b = __shfl_xor(mx, o);
            mn = a < mn ? a : mn;
            mx = b > mx ? b : mx;
        }
        const int ndel = __popcll(__ballot(inner_del));
        if (laneId() == 0) {
            if (mx >= 0) {
                atomicMin((long long*)(cnt + IR_MINR), mn);
                atomicMax((long long*)(cnt + IR_MAXR), mx);
            }
            if (ndel) atomicAdd((unsigned long long*)(cnt + IR_DIRTY), (unsigned long long)ndel);
        }
    }
}

// Vertices whose label changed in a repair (old -> scc): listed in list,
// cnt[0] = their number, cnt[1] = their forward + backward degree sum.
__global__ void kIrChanged(const int32_t* __restrict__ old, const int32_t* __restrict__ scc, int32_t V,
                           const int32_t* __restrict__ fb, const int32_t* __restrict__ bb, int32_t* __restrict__ list,
                           int64_t* cnt) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    int64_t deg = 0;
    for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < V; base += stride) {
        const int64_t v = base + threadIdx.x;
        const bool ch = v < V && old[v] != scc[v];
        if (ch) deg += (int64_t)(fb[v + 1] - fb[v]) + (bb[v + 1] - bb[v]);
        const int64_t pos = blockReserve(ch ? 1 : 0, cnt);
        if (ch) list[pos] = (int32_t)v;
    }
    blockAtomicAdd(deg, cnt + 1);
}

// Condensation deltas of a relabelling, one block per changed vertex x: every
// arc at x (out-arcs; in-arcs from unchanged tails, so an arc between two
// changed vertices counts once) takes back its old key (-1) and adds its new
// key (+1), each only when its ends lie in different SCCs.
__global__ void kIrRelabelDeltas(const int32_t* __restrict__ list, const int32_t* __restrict__ old,
                                 const int32_t* __restrict__ scc, const int32_t* __restrict__ fb,
                                 const int32_t* __restrict__ fa, const int32_t* __restrict__ bb,
                                 const int32_t* __restrict__ ba, int B, uint64_t* __restrict__ keys,
                                 uint32_t* __restrict__ vals, int64_t* cnt) {
    const int32_t x = list[blockIdx.x];
    const int32_t ox = old[x], nx = scc[x];
    for (int dir = 0; dir < 2; ++dir) {
        const int32_t b = dir ? bb[x] : fb[x], e = dir ? bb[x + 1] : fb[x + 1];
        const int32_t* adj = dir ? ba : fa;
        for (int32_t j0 = b; j0 < e; j0 += blockDim.x) {
            const int32_t j = j0 + (int32_t)threadIdx.x;
            bool take_old = false, take_new = false;
            uint64_t ko = 0, kn = 0;
            if (j < e) {
                const int32_t y = adj[j];
                const int32_t oy = old[y], ny = scc[y];
                if (!(dir && oy != ny)) {
                    // (tail, head) = (x, y) forward, (y, x) backward
                    const int32_t ot = dir ? oy : ox, oh = dir ? ox : oy;
                    const int32_t nt = dir ? ny : nx, nh = dir ? nx : ny;
                    take_old = ot != oh;
                    take_new = nt != nh;
                    ko = ((uint64_t)(uint32_t)ot << B) | (uint32_t)oh;
                    kn = ((uint64_t)(uint32_t)nt << B) | (uint32_t)nh;
                }
            }
            const int64_t po = wavePush(cnt, take_old);
            if (take_old) {
                keys[po] = ko;
                vals[po] = 0xffffffffu;
            }
            const int64_t pn = wavePush(cnt, take_new);
            if (take_new) {
                keys[pn] = kn;
                vals[pn] = 1u;
            }
        }
    }
}

// Condensation sweeps inside a rank window: forward visits rank <= lim,
// backward rank >= lim (ids that name no SCC carry rank -1).
struct RankWindow {
    const int64_t* rank;
    int64_t lim;
    int fwd;
    __device__ bool self(int32_t c) const {
        const int64_t r = rank[c];
        return r >= 0 && (fwd ? r <= lim : r >= lim);
    }
    __device__ bool operator()(int32_t c, int32_t) const { return self(c); }
};

// The distinct ids among n seed ids: flag them, then list the flagged ids
// (a batch can hold more rank-violating inserts than there are SCCs).
__global__ void kIrSeedFlags(const int32_t* __restrict__ ids, int64_t n, uint8_t* __restrict__ flag) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) flag[ids[i]] = 1;
}
__global__ void kIrSeedList(const uint8_t* __restrict__ flag, int32_t V, int32_t* __restrict__ out, int64_t* cnt) {
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool f = v < V && flag[v];
    const int64_t pos = blockReserve(f ? 1 : 0, cnt);
    if (f) out[pos] = (int32_t)v;
}

// Vertices of the SCCs that may change: cnt += their number; with apply,
// they are also labelled -1.
// The big SCC (label excl, or -1) is left to its certificate: its vertices
// are neither counted nor relabelled, and cnt[1] says whether the region holds it.
__global__ void kIrMark(int32_t* scc, int32_t V, const uint8_t* __restrict__ dirty, const uint64_t* __restrict__ fw,
                        const uint64_t* __restrict__ bw, int use_region, int apply, int64_t* cnt, int32_t excl,
                        int64_t* greg) {
    if (blockIdx.x == 0 && threadIdx.x == 0 && excl >= 0)
        *greg = use_region && bitOf(fw, excl) && bitOf(bw, excl) ? 1 : 0;
    int64_t a = 0;
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < V; v += (int64_t)gridDim.x * blockDim.x) {
        const int32_t c = scc[v];
        const bool in = c != excl && (dirty[c] || (use_region && bitOf(fw, c) && bitOf(bw, c)));
        if (in) {
            if (apply) scc[v] = -1;
            ++a;
        }
    }
    blockAtomicAdd(a, cnt);
}

}  // namespace

// ---- merging multiplicity changes into the condensation ----
// The arcs are kept sorted by key (src << B | dst) and the in-arc keys by
// (dst << B | src), so a batch's changes (signed per-arc counts) are merged
// instead of re-sorting every arc: the changes are sorted and summed per
// key; each key is matched against the arcs by binary search (new
// multiplicity, or a new key); the new keys and the keys whose multiplicity
// drops to zero form two short sorted lists (padded with DM_PAD to a fixed
// length, so no host count is needed), and every kept arc moves to
//   i + #(new keys below it) - #(removed keys below it),
// every new key to its list index + #(arcs below it) - #(removed keys below
// it).  The same placement runs on the in-arc keys with the transposed lists.
namespace {

__device__ __forceinline__ uint64_t arcKey(const int32_t* src, const int32_t* dst, int64_t i, int B) {
    return ((uint64_t)(uint32_t)src[i] << B) | (uint32_t)dst[i];
}
// entries of a sorted sequence (accessor k, n entries) below x
template <class K>
__device__ __forceinline__ int64_t lowerBound(const K& k, int64_t n, uint64_t x) {
    int64_t lo = 0, hi = n;
    while (lo < hi) {
        const int64_t m = lo + ((hi - lo) >> 1);
        if (k(m) < x) lo = m + 1;
        else hi = m;
    }
    return lo;
}

// sorted (key, signed count) pairs: at the first entry of every key, its sum
// and a flag that the sum is not zero
__global__ void kDmRuns(const uint64_t* __restrict__ keys, const uint32_t* __restrict__ vals, int64_t n,
                        int32_t* __restrict__ flag, int32_t* __restrict__ sum) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int32_t f = 0;
    if (i == 0 || keys[i] != keys[i - 1]) {
        int32_t acc = 0;
        for (int64_t j = i; j < n && keys[j] == keys[i]; ++j) acc += (int32_t)vals[j];
        sum[i] = acc;
        f = acc != 0;
    }
    flag[i] = f;
}

// unique changes (uk, us) matched against the arcs; entries past the unique
// count (*nu) get no flags.  err: a change that would make a multiplicity
// negative (the change lists would not describe the store).
__global__ void kDmMatch(const uint64_t* __restrict__ keys, const int32_t* __restrict__ flag,
                         const int32_t* __restrict__ sum, const int64_t* __restrict__ pos, int64_t n,
                         const int32_t* __restrict__ src, const int32_t* __restrict__ dst,
                         const int32_t* __restrict__ mult, int64_t nd, int B, uint64_t* __restrict__ uk,
                         int32_t* __restrict__ nm, int64_t* __restrict__ fi, int32_t* __restrict__ nf,
                         int32_t* __restrict__ rf, int64_t* err) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t nu = pos[n];
    if (i >= nu) {  // (slot i of the unique lists)
        nf[i] = 0;
        rf[i] = 0;
    }
    if (!flag[i]) return;
    const int64_t j = pos[i];
    const uint64_t k = keys[i];
    auto ak = [&](int64_t a) { return arcKey(src, dst, a, B); };
    const int64_t a = lowerBound(ak, nd, k);
    const bool found = a < nd && ak(a) == k;
    const int32_t m = (found ? mult[a] : 0) + sum[i];
    if (m < 0) atomicAdd((unsigned long long*)err, 1ULL);
    uk[j] = k;
    nm[j] = m;
    fi[j] = found ? a : -1;
    nf[j] = !found && m > 0;
    rf[j] = found && m <= 0;
}

__global__ void kDmLists(const uint64_t* __restrict__ uk, const int32_t* __restrict__ nm,
                         const int32_t* __restrict__ nf, const int32_t* __restrict__ rf,
                         const int64_t* __restrict__ pnew, const int64_t* __restrict__ prem, int64_t n, int B,
                         uint64_t* __restrict__ newk, int32_t* __restrict__ newm, uint64_t* __restrict__ remk,
                         uint64_t* __restrict__ newt, uint64_t* __restrict__ remt) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const uint64_t mask = (1ULL << B) - 1;
    if (nf[j]) {
        const uint64_t k = uk[j];
        newk[pnew[j]] = k;
        newm[pnew[j]] = nm[j];
        newt[pnew[j]] = ((k & mask) << B) | (k >> B);
    }
    if (rf[j]) {
        const uint64_t k = uk[j];
        remk[prem[j]] = k;
        remt[prem[j]] = ((k & mask) << B) | (k >> B);
    }
}

// the forward arcs: kept arcs, then the new ones, then the changed
// multiplicities (writing into the spare arrays)
__global__ void kDmPlaceOld(const int32_t* __restrict__ src, const int32_t* __restrict__ dst,
                            const int32_t* __restrict__ mult, int64_t nd, int B, const uint64_t* __restrict__ newk,
                            const uint64_t* __restrict__ remk, int64_t m, int32_t* __restrict__ osrc,
                            int32_t* __restrict__ odst, int32_t* __restrict__ omult) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nd) return;
    const uint64_t k = arcKey(src, dst, i, B);
    auto nk = [&](int64_t a) { return newk[a]; };
    auto rk = [&](int64_t a) { return remk[a]; };
    const int64_t r = lowerBound(rk, m, k);
    if (r < m && remk[r] == k) return;
    const int64_t p = i + lowerBound(nk, m, k) - r;
    osrc[p] = src[i];
    odst[p] = dst[i];
    omult[p] = mult[i];
}
__global__ void kDmPlaceNew(const uint64_t* __restrict__ newk, const int32_t* __restrict__ newm, int64_t m,
                            const int32_t* __restrict__ src, const int32_t* __restrict__ dst, int64_t nd, int B,
                            const uint64_t* __restrict__ remk, uint64_t pad, int32_t* __restrict__ osrc,
                            int32_t* __restrict__ odst, int32_t* __restrict__ omult) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= m || newk[j] == pad) return;
    const uint64_t k = newk[j];
    auto ak = [&](int64_t a) { return arcKey(src, dst, a, B); };
    auto rk = [&](int64_t a) { return remk[a]; };
    const int64_t p = j + lowerBound(ak, nd, k) - lowerBound(rk, m, k);
    osrc[p] = (int32_t)(k >> B);
    odst[p] = (int32_t)(k & ((1ULL << B) - 1));
    omult[p] = newm[j];
}
__global__ void kDmUpdate(const uint64_t* __restrict__ uk, const int32_t* __restrict__ nm,
                          const int64_t* __restrict__ fi, const int64_t* __restrict__ pos, int64_t n,
                          const uint64_t* __restrict__ newk, const uint64_t* __restrict__ remk,
                          int32_t* __restrict__ omult) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= pos[n] || fi[j] < 0 || nm[j] <= 0) return;
    const uint64_t k = uk[j];
    auto nk = [&](int64_t a) { return newk[a]; };
    auto rk = [&](int64_t a) { return remk[a]; };
    omult[fi[j] + lowerBound(nk, n, k) - lowerBound(rk, n, k)] = nm[j];
}
// the in-arc keys
__global__ void kDmPlaceTk(const uint64_t* __restrict__ tk, int64_t nd, const uint64_t* __restrict__ newt,
                           const uint64_t* __restrict__ remt, int64_t m, uint64_t pad, uint64_t* __restrict__ otk) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    auto tkk = [&](int64_t a) { return tk[a]; };
    auto nt = [&](int64_t a) { return newt[a]; };
    auto rt = [&](int64_t a) { return remt[a]; };
    if (i < nd) {
        const uint64_t k = tk[i];
        const int64_t r = lowerBound(rt, m, k);
        if (!(r < m && remt[r] == k)) otk[i + lowerBound(nt, m, k) - r] = k;
    }
    if (i < m && newt[i] != pad) {
        const uint64_t k = newt[i];
        otk[i + lowerBound(tkk, nd, k) - lowerBound(rt, m, k)] = k;
    }
}
__global__ void kDmCounts(const int64_t* __restrict__ pnew, const int64_t* __restrict__ prem, int64_t n,
                          int64_t* cnt) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        cnt[0] = pnew[n];
        cnt[1] = prem[n];
    }
}

}  // namespace

// Merge n signed per-arc changes (dk: packed (src << B) | dst keys, dv: +1 /
// -1 as uint32; sorted here in place) into the condensation; with ranks the
// Kahn levels are recomputed.
static void dagMerge(girhip_graph* h, int B, uint64_t* dk, uint32_t* dv, int64_t n, bool ranks) {
    const int blk = (int)h->prm.block;
    Scratch& S = scratch(h);
    const int64_t nd0 = h->n_dag;
    if (n <= 0) {
        if (ranks) dagIndex(h, nd0, B, true);
        return;
    }
    radixSort(h, dk, dv, n, 2 * B);
    S.dm_flag.ensure((size_t)n);
    S.dm_sum.ensure((size_t)n);
    S.dm_pos.ensure((size_t)n + 1);
    S.dm_uk.ensure((size_t)n);
    S.dm_nm.ensure((size_t)n);
    S.dm_fi.ensure((size_t)n);
    S.dm_nf.ensure((size_t)n);
    S.dm_rf.ensure((size_t)n);
    S.dm_pnew.ensure((size_t)n + 1);
    S.dm_prem.ensure((size_t)n + 1);
    S.dm_newk.ensure((size_t)n);
    S.dm_newm.ensure((size_t)n);
    S.dm_remk.ensure((size_t)n);
    S.dm_newt.ensure((size_t)n);
    S.dm_remt.ensure((size_t)n);
    S.dm_cnt.ensure(4);
    const uint64_t pad = 1ULL << (2 * B);  // above every key
    const unsigned g = gridFor(n, blk);
    launch(h, "dm_runs", kDmRuns, g, blk, 0, (const uint64_t*)dk, (const uint32_t*)dv, n, S.dm_flag.p, S.dm_sum.p);
    scanI32toI64(h, S.dm_flag.p, S.dm_pos.p, n);
    fillI64(h, S.dm_cnt.p, 0, 4);
    launch(h, "dm_match", kDmMatch, g, blk, 0, (const uint64_t*)dk, (const int32_t*)S.dm_flag.p,
           (const int32_t*)S.dm_sum.p, (const int64_t*)S.dm_pos.p, n, (const int32_t*)h->dag_src.p,
           (const int32_t*)h->dag_dst.p, (const int32_t*)h->dag_mult.p, nd0, B, S.dm_uk.p, S.dm_nm.p, S.dm_fi.p,
           S.dm_nf.p, S.dm_rf.p, S.dm_cnt.p + 2);
    scanI32toI64(h, S.dm_nf.p, S.dm_pnew.p, n);
    scanI32toI64(h, S.dm_rf.p, S.dm_prem.p, n);
    fillU64(h, S.dm_newk.p, pad, n);
    fillU64(h, S.dm_remk.p, pad, n);
    fillU64(h, S.dm_newt.p, pad, n);
    fillU64(h, S.dm_remt.p, pad, n);
    launch(h, "dm_lists", kDmLists, g, blk, 0, (const uint64_t*)S.dm_uk.p, (const int32_t*)S.dm_nm.p,
           (const int32_t*)S.dm_nf.p, (const int32_t*)S.dm_rf.p, (const int64_t*)S.dm_pnew.p,
           (const int64_t*)S.dm_prem.p, n, B, S.dm_newk.p, S.dm_newm.p, S.dm_remk.p, S.dm_newt.p, S.dm_remt.p);
    launch(h, "dm_counts", kDmCounts, 1, 64, 0, (const int64_t*)S.dm_pnew.p, (const int64_t*)S.dm_prem.p, n,
           S.dm_cnt.p);
    int64_t c[3];
    readCounters(h, S.dm_cnt.p, 3, c);
    if (c[2] != 0) raise(GIRHIP_EHIP, "condensation merge: a change makes a multiplicity negative");
    // the transposed lists in order (their padding stays behind them)
    if (c[0] > 1) radixSort(h, S.dm_newt.p, nullptr, c[0], 2 * B);
    if (c[1] > 1) radixSort(h, S.dm_remt.p, nullptr, c[1], 2 * B);
    const int64_t nd1 = nd0 + c[0] - c[1];
    h->dag_src2.ensureSlack((size_t)std::max<int64_t>(nd1, 1));
    h->dag_dst2.ensureSlack((size_t)std::max<int64_t>(nd1, 1));
    h->dag_mult2.ensureSlack((size_t)std::max<int64_t>(nd1, 1));
    h->dag_tk2.ensureSlack((size_t)std::max<int64_t>(nd1, 1));
    if (nd0 > 0)
        launch(h, "dm_place_old", kDmPlaceOld, gridFor(nd0, blk), blk, 0, (const int32_t*)h->dag_src.p,
               (const int32_t*)h->dag_dst.p, (const int32_t*)h->dag_mult.p, nd0, B, (const uint64_t*)S.dm_newk.p,
               (const uint64_t*)S.dm_remk.p, n, h->dag_src2.p, h->dag_dst2.p, h->dag_mult2.p);
    launch(h, "dm_place_new", kDmPlaceNew, g, blk, 0, (const uint64_t*)S.dm_newk.p, (const int32_t*)S.dm_newm.p, n,
           (const int32_t*)h->dag_src.p, (const int32_t*)h->dag_dst.p, nd0, B, (const uint64_t*)S.dm_remk.p, pad,
           h->dag_src2.p, h->dag_dst2.p, h->dag_mult2.p);
    launch(h, "dm_update", kDmUpdate, g, blk, 0, (const uint64_t*)S.dm_uk.p, (const int32_t*)S.dm_nm.p,
           (const int64_t*)S.dm_fi.p, (const int64_t*)S.dm_pos.p, n, (const uint64_t*)S.dm_newk.p,
           (const uint64_t*)S.dm_remk.p, h->dag_mult2.p);
    launch(h, "dm_place_tk", kDmPlaceTk, gridFor(std::max(nd0, n), blk), blk, 0, (const uint64_t*)h->dag_tk.p, nd0,
           (const uint64_t*)S.dm_newt.p, (const uint64_t*)S.dm_remt.p, n, pad, h->dag_tk2.p);
    h->dag_src.swap(h->dag_src2);
    h->dag_dst.swap(h->dag_dst2);
    h->dag_mult.swap(h->dag_mult2);
    h->dag_tk.swap(h->dag_tk2);
    dagIndex(h, nd1, B, ranks);
}

// The condensation after a relabelling (path 3 of repairScc): the current
// arcs, the batch's multiplicity deltas not yet merged (nir of them, stored
// after the arcs' own slots in ir_keys / ir_vals, under the old labels), and
// the deltas of the rows of every vertex whose label changed (old labels in
// ir_old), summed per key; then the in-arc CSR and the ranks.  A rebuild of
// the condensation would read the rows of every vertex outside the big SCC.
static void relabelDag(girhip_graph* h, int B, int64_t nir) {
    const int32_t V = h->V;
    const int blk = (int)h->prm.block;
    Scratch& S = scratch(h);
    h->mem_ok = false;  // member lists follow the labels
    const int64_t nd0 = h->n_dag;
    S.ir_list.ensure((size_t)V);
    S.ir_cnt2.ensure(4);
    int64_t* c2 = S.ir_cnt2.p;
    fillI64(h, c2, 0, 2);
    launch(h, "ir_changed", kIrChanged, reduceGrid(V, blk), blk, 0, (const int32_t*)S.ir_old.p,
           (const int32_t*)h->scc.p, V, (const int32_t*)h->fw_beg.p, (const int32_t*)h->bw_beg.p, S.ir_list.p, c2);
    int64_t r[2];
    readCounters(h, c2, 2, r);
    const int64_t nch = r[0], deg = r[1];
    h->last_ms["batch_scc_changed"] = (double)nch;
    const int64_t cap = nir + 2 * deg;
    S.keys[0].ensure((size_t)std::max<int64_t>(cap, 1));
    S.vals[0].ensure((size_t)std::max<int64_t>(cap, 1));
    if (nir > 0) {
        HIPCHK(hipMemcpyAsync(S.keys[0].p, S.ir_keys.p + nd0, sizeof(uint64_t) * nir, hipMemcpyDeviceToDevice, h->st));
        HIPCHK(hipMemcpyAsync(S.vals[0].p, S.ir_vals.p + nd0, sizeof(uint32_t) * nir, hipMemcpyDeviceToDevice, h->st));
    }
    int64_t m = nir;
    if (nch > 0) {
        fillI64(h, c2 + 2, m, 1);
        launch(h, "ir_relabel_deltas", kIrRelabelDeltas, (unsigned)nch, 256, 0, (const int32_t*)S.ir_list.p,
               (const int32_t*)S.ir_old.p, (const int32_t*)h->scc.p, (const int32_t*)h->fw_beg.p,
               (const int32_t*)h->fw_adj.p, (const int32_t*)h->bw_beg.p, (const int32_t*)h->bw_adj.p, B, S.keys[0].p,
               S.vals[0].p, c2 + 2);
        readCounters(h, c2 + 2, 1, &m);
    }
    dagMerge(h, B, S.keys[0].p, S.vals[0].p, m, true);
}

// Work buffers of a repair after a batch of up to n changes (and of the
// certificate check), reserved when the layer is first built.
void reserveRepair(girhip_graph* h, int64_t n) {
    Scratch& S = scratch(h);
    const int32_t V = h->V;
    const size_t m = (size_t)std::max<int64_t>(n, 1), Vs = (size_t)std::max<int32_t>(V, 1);
    S.ir_cnt.ensure(IR_N);
    S.ir_cnt2.ensure(4);
    S.ir_dirty.ensure(Vs);
    S.ir_seed.ensure(Vs);
    S.ir_old.ensure(Vs);
    S.ir_list.ensure(Vs);
    S.ir_bsu.ensure(m);
    S.ir_bsv.ensure(m);
    S.gc_cut[0].ensure(m);
    S.gc_cut[1].ensure(m);
    S.ir_keys.ensureSlack((size_t)h->n_dag + m);
    S.ir_vals.ensureSlack((size_t)h->n_dag + m);
    const size_t d = 4 * m;  // changes of a repair: the batch's and a relabelling's
    S.dm_flag.ensure(d);
    S.dm_sum.ensure(d);
    S.dm_uk.ensure(d);
    S.dm_nm.ensure(d);
    S.dm_fi.ensure(d);
    S.dm_nf.ensure(d);
    S.dm_rf.ensure(d);
    S.dm_newk.ensure(d);
    S.dm_newm.ensure(d);
    S.dm_remk.ensure(d);
    S.dm_newt.ensure(d);
    S.dm_remt.ensure(d);
    S.dm_pos.ensure(d + 1);
    S.dm_pnew.ensure(d + 1);
    S.dm_prem.ensure(d + 1);
    S.dm_cnt.ensure(4);
    const size_t nd = (size_t)std::max<int64_t>(h->n_dag, 1) + m;
    h->dag_src2.ensureSlack(nd);
    h->dag_dst2.ensureSlack(nd);
    h->dag_mult2.ensureSlack(nd);
    h->dag_tk2.ensureSlack(nd);
    reserveCert(h);
}

// The change lists of a batch just applied to the store (ik / dk: sorted
// packed (u << B) | v keys).  Returns false when the layer must be rebuilt
// instead: knob scc_incremental = 0, a batch of more than repair_ops
// changes (default max(4096, V / 256): RMAT-22, random batches, repair vs
// rebuild: 10^4 changes 3.8 vs 3.9 ms, 3 x 10^4 4.7 vs 4.0 ms), or more than
// max(65536, V / 8) vertices in SCCs that may change outside the big SCC.  The big SCC itself (the SCC of
// the certificate's pivot) is never relabelled wholesale: its inner deletes
// are checked against its spanning trees and the SCCs that may merge into
// it are attached to it (certRepair), so a power-law graph's giant SCC costs
// only its orphaned subtrees.
bool repairScc(girhip_graph* h, const uint64_t* ik, int64_t ni, const uint64_t* dk, int64_t nd, int B,
               const uint8_t* dbypass) {
    if (h->prm.scc_incremental == 0) return false;
    const int64_t max_ops =
        h->prm.repair_ops > 0 ? (int64_t)h->prm.repair_ops : std::max<int64_t>(4096, (int64_t)h->V / 256);
    if (ni + nd > max_ops) return false;
    const SccBlock sb(h);
    const int32_t V = h->V;
    const int blk = (int)h->prm.block;
    Scratch& S = scratch(h);
    const int64_t n = ni + nd;
    h->last_ms["batch_scc_path"] = 0;
    h->last_ms["batch_scc_affected"] = 0;
    h->last_ms["batch_scc_cert"] = 0;
    if (n == 0 || V == 0) return true;
    // the big SCC and its trees (pivot and label picked now when absent)
    const bool use_cert = h->prm.scc_cert != 0;
    if (!use_cert) h->gc_ok = false;
    else if (!h->gc_ok) certTarget(h);
    else if (h->gc_label < 0) certLabel(h);  // trees from the build: its pivot's label
    const int32_t g = use_cert ? h->gc_label : -1;
    S.ir_cnt.ensure(IR_N);
    S.ir_dirty.ensure((size_t)V);
    S.ir_bsu.ensure((size_t)std::max<int64_t>(ni, 1));
    S.ir_bsv.ensure((size_t)std::max<int64_t>(ni, 1));
    S.gc_cut[0].ensure((size_t)std::max<int64_t>(nd, 1));
    S.gc_cut[1].ensure((size_t)std::max<int64_t>(nd, 1));
    const int64_t cap = h->n_dag + n;
    S.ir_keys.ensure((size_t)cap);
    S.ir_vals.ensure((size_t)cap);
    int64_t* cnt = S.ir_cnt.p;
    fillI64(h, cnt, 0, IR_N);
    fillI64(h, cnt + IR_MINR, LLONG_MAX, 1);
    fillI64(h, cnt + IR_MAXR, -1, 1);
    fillU8(h, S.ir_dirty.p, 0, V);
    const IrCert gc{g, h->gc_ok ? (const int32_t*)h->gc_fpar.p : nullptr,
                    h->gc_ok ? (const int32_t*)h->gc_bpar.p : nullptr, S.gc_cut[0].p, S.gc_cut[1].p};
    // deltas go after the condensation's own arcs in the key array
    const int64_t nd0 = h->n_dag;
    launch(h, "ir_classify", kIrClassify, std::min<unsigned>(gridFor(n, blk), 4096), blk, 0, ik, ni, dk, nd, B,
           (const int32_t*)h->scc.p, (const int64_t*)h->rank.p, S.ir_keys.p + nd0, S.ir_vals.p + nd0, S.ir_bsu.p,
           S.ir_bsv.p, S.ir_dirty.p, cnt, dbypass, gc);
    int64_t c[IR_N];
    readCounters(h, cnt, IR_N, c);
    const int64_t ndelta = c[IR_DELTA], nback = c[IR_BACK], ndirty = c[IR_DIRTY];
    const int64_t ncutf = c[IR_CUTF], ncutb = c[IR_CUTB];
    // the big SCC needs a check: a deleted tree arc, or an inner delete while
    // its trees are absent
    const bool gdel = h->gc_ok ? (ncutf + ncutb) > 0 : c[IR_GDEL] != 0;
    if (ndelta == 0 && ndirty == 0 && !gdel) {  // inserts inside SCCs only
        h->last_ms["batch_scc_path"] = 1;
        return true;
    }
    const int64_t big = std::max<int64_t>(65536, (int64_t)V / 8);
    int64_t greg = 0;
    auto affected = [&](bool region, bool apply) {
        fillI64(h, cnt + IR_AFF, 0, 1);
        launch(h, "ir_mark", kIrMark, reduceGrid(V, blk), blk, 0, h->scc.p, V, (const uint8_t*)S.ir_dirty.p,
               (const uint64_t*)S.bits[3].p, (const uint64_t*)reachBufs(h).vis, region ? 1 : 0, apply ? 1 : 0,
               cnt + IR_AFF, g, cnt + IR_GREG);
        int64_t a[IR_GREG - IR_AFF + 1];
        readCounters(h, cnt + IR_AFF, IR_GREG - IR_AFF + 1, a);
        greg = region && g >= 0 ? a[IR_GREG - IR_AFF] : 0;
        return a[0];
    };
    if (ndirty > 0 && affected(false, false) > big) return false;
    // condensation multiplicities: merged now when the partition may stand or
    // the merge sweeps need the updated condensation; a relabelling merges
    // them with its own deltas
    const bool dag_has_delta = ndelta > 0 && (nback > 0 || (ndirty == 0 && !gdel));
    if (dag_has_delta) dagMerge(h, B, S.ir_keys.p + nd0, S.ir_vals.p + nd0, ndelta, false);
    if (nback == 0 && ndirty == 0 && !gdel) {  // the partition and the ranks stand
        h->last_ms["batch_scc_path"] = 1;
        return true;
    }
    // SCCs that may merge: reachable from a rank-violating insert's head and
    // reaching its tail (or another's) inside the rank window
    ReachBufs R = reachBufs(h);
    S.bits[3].ensure((size_t)R.W);
    if (nback > 0) {
        S.ir_seed.ensure((size_t)V);
        // distinct seed SCCs into the sweep's first queue -> their number
        auto seeds = [&](const int32_t* ids) {
            fillU8(h, S.ir_seed.p, 0, V);
            launch(h, "ir_seed_flags", kIrSeedFlags, gridFor(nback, blk), blk, 0, ids, nback, S.ir_seed.p);
            fillI64(h, cnt + IR_AFF, 0, 1);
            launch(h, "ir_seed_list", kIrSeedList, gridFor(V, blk), blk, 0, (const uint8_t*)S.ir_seed.p, V, R.q[0],
                   cnt + IR_AFF);
            int64_t ns;
            readCounters(h, cnt + IR_AFF, 1, &ns);
            return ns;
        };
        fillU64(h, R.vis, 0, R.W);
        int64_t ns = seeds(S.ir_bsv.p);
        runReach(h, h->dag_beg.p, h->dag_dst.p, h->dag_in_beg.p, h->dag_in_src.p, R, ns, 0, sccCount(h), h->n_dag,
                 RankWindow{h->rank.p, c[IR_MAXR], 1}, NoVisit{}, true, nullptr, nullptr, true);
        HIPCHK(hipMemcpyAsync(S.bits[3].p, R.vis, sizeof(uint64_t) * R.W, hipMemcpyDeviceToDevice, h->st));
        fillU64(h, R.vis, 0, R.W);
        ns = seeds(S.ir_bsu.p);
        runReach(h, h->dag_in_beg.p, h->dag_in_src.p, h->dag_beg.p, h->dag_dst.p, R, ns, 0, sccCount(h), h->n_dag,
                 RankWindow{h->rank.p, c[IR_MINR], 0}, NoVisit{}, true, nullptr, nullptr, true);
    }
    int64_t naff = affected(nback > 0, false);
    h->last_ms["batch_scc_affected"] = (double)naff;
    if (naff > big) return false;
    const bool gwork = gdel || greg != 0;
    if (naff == 0 && !gwork) {  // no cycle closed: re-rank the condensation
        h->last_ms["batch_scc_path"] = 2;
        dagIndex(h, h->n_dag, B, true);
        return true;
    }
    h->last_ms["batch_scc_path"] = 3;
    // the labels before the relabelling: the condensation is then updated
    // from the rows of the vertices whose label changes
    S.ir_old.ensure((size_t)V);
    HIPCHK(hipMemcpyAsync(S.ir_old.p, h->scc.p, sizeof(int32_t) * V, hipMemcpyDeviceToDevice, h->st));
    affected(nback > 0, true);
    if (gwork) {
        h->last_ms["batch_scc_cert"] = h->gc_ok ? 1 : 2;
        const CertStats cs = certRepair(h, h->gc_ok ? ncutf : 0, h->gc_ok ? ncutb : 0, naff);
        h->last_ms["batch_scc_dropped"] = (double)cs.dropped;
        h->last_ms["batch_scc_joined"] = (double)cs.joined;
    }
    runSccActive(h);
    relabelDag(h, B, dag_has_delta ? 0 : ndelta);
    return true;
}

}  // namespace gir

using namespace gir;

extern "C" int girhip_scc_dag(girhip_graph* h, int32_t* scc, int32_t* n_scc, int64_t* n_dag) {
    return guarded([&] {
        useDevice(h);
        tBegin(h);
        runScc(h);
        runDag(h);
        tEnd(h, "scc_dag");
        // what maintaining the layer under small batches needs (allocated once)
        reserveBatch(h, 4096);
        reserveRepair(h, 4096);
        h->scc_req = true;
        h->scc_ok = true;
        if (scc) copyOut(h, scc, h->scc.p, sizeof(int32_t) * (size_t)h->V);
        syncStream(h);
        if (n_scc) *n_scc = sccCount(h);
        if (n_dag) *n_dag = h->n_dag;
    });
}

extern "C" int girhip_get_scc(girhip_graph* h, int32_t* scc, int32_t* n_scc, int64_t* n_dag) {
    return guarded([&] {
        useDevice(h);
        if (!h->scc_req) raise(GIRHIP_ESTATE, "get_scc: call girhip_scc_dag first");
        if (scc) copyOut(h, scc, h->scc.p, sizeof(int32_t) * (size_t)h->V);
        syncStream(h);
        if (n_scc) *n_scc = sccCount(h);
        if (n_dag) *n_dag = h->n_dag;
    });
}

extern "C" int girhip_get_dag(girhip_graph* h, int32_t* src, int32_t* dst, int32_t* mult) {
    return guarded([&] {
        useDevice(h);
        if (!h->scc_req) raise(GIRHIP_ESTATE, "get_dag: call girhip_scc_dag first");
        const size_t b = sizeof(int32_t) * (size_t)h->n_dag;
        if (src) copyOut(h, src, h->dag_src.p, b);
        if (dst) copyOut(h, dst, h->dag_dst.p, b);
        if (mult) copyOut(h, mult, h->dag_mult.p, b);
        syncStream(h);
    });
}

extern "C" int girhip_get_ranks(girhip_graph* h, int64_t* rank) {
    return guarded([&] {
        useDevice(h);
        if (!h->scc_req) raise(GIRHIP_ESTATE, "get_ranks: call girhip_scc_dag first");
        if (!rank) raise(GIRHIP_EINVAL, "get_ranks: null output");
        copyOut(h, rank, h->rank.p, sizeof(int64_t) * (size_t)h->V);
        syncStream(h);
    });
}

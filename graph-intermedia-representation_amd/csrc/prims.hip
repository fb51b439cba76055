// Device primitives: fills, exclusive scans and a stable LSD radix sort of
// 64-bit keys with 32-bit values.  All wave64 / LDS based, written for gfx950.
#include "common.h"

namespace gir {

template <class T>
__global__ void kFill(T* __restrict__ p, T v, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

template <class T>
static void fillT(girhip_graph* h, T* p, T v, int64_t n) {
    if (n <= 0) return;
    unsigned g = gridFor(n, 256);
    if (g > 8192) g = 8192;
    launch(h, "fill", kFill<T>, g, 256, 0, p, v, n);
}
void fillI32(girhip_graph* h, int32_t* p, int32_t v, int64_t n) { fillT(h, p, v, n); }
void fillU32(girhip_graph* h, uint32_t* p, uint32_t v, int64_t n) { fillT(h, p, v, n); }
void fillI64(girhip_graph* h, int64_t* p, int64_t v, int64_t n) { fillT(h, p, v, n); }
void fillU64(girhip_graph* h, uint64_t* p, uint64_t v, int64_t n) { fillT(h, p, v, n); }
void fillU8(girhip_graph* h, uint8_t* p, uint8_t v, int64_t n) { fillT(h, p, v, n); }

// ---------------------------------------------------------------------------
// Exclusive scan: tile = 256 threads x 8 items.
constexpr int SCAN_T = 256, SCAN_I = 8, SCAN_TILE = SCAN_T * SCAN_I;

// Inclusive scan of v across the block; returns the inclusive value and the
// block total in *total.
__device__ __forceinline__ int64_t blockInclusiveScan(int64_t v, int64_t* total) {
    __shared__ int64_t wsum[SCAN_T / WAVE];
    const int lane = laneId(), w = threadIdx.x / WAVE;
    for (int d = 1; d < WAVE; d <<= 1) {
        int64_t o = __shfl_up(v, d);
        if (lane >= d) v += o;
    }
    if (lane == WAVE - 1) wsum[w] = v;
    __syncthreads();
    int64_t pre = 0, tot = 0;
    for (int i = 0; i < SCAN_T / WAVE; ++i) {
        if (i < w) pre += wsum[i];
        tot += wsum[i];
    }
    __syncthreads();
    *total = tot;
    return v + pre;
}

template <class In>
__global__ void kScanReduce(const In* __restrict__ in, int64_t n, int64_t* __restrict__ bsum) {
    int64_t base = (int64_t)blockIdx.x * SCAN_TILE;
    int64_t s = 0;
    for (int i = 0; i < SCAN_I; ++i) {
        int64_t idx = base + (int64_t)i * SCAN_T + threadIdx.x;
        if (idx < n) s += (int64_t)in[idx];
    }
    int64_t tot;
    blockInclusiveScan(s, &tot);
    if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

// Exclusive scan of bsum[0..nb) in one block (in place); bsum[nb] = total.
__global__ void kScanBlockSums(int64_t* bsum, int64_t nb) {
    int64_t carry = 0;
    for (int64_t base = 0; base < nb; base += SCAN_T) {
        int64_t idx = base + threadIdx.x;
        int64_t v = idx < nb ? bsum[idx] : 0;
        int64_t tot;
        int64_t inc = blockInclusiveScan(v, &tot);
        if (idx < nb) bsum[idx] = carry + inc - v;
        carry += tot;
    }
    if (threadIdx.x == 0) bsum[nb] = carry;
}

template <class In, class Out>
__global__ void kScanDown(const In* __restrict__ in, Out* __restrict__ out, int64_t n, const int64_t* __restrict__ bsum) {
    const int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * SCAN_I;
    int64_t v[SCAN_I];
    int64_t s = 0;
    for (int i = 0; i < SCAN_I; ++i) {
        int64_t idx = base + i;
        v[i] = idx < n ? (int64_t)in[idx] : 0;
        s += v[i];
    }
    int64_t tot;
    int64_t inc = blockInclusiveScan(s, &tot);
    int64_t run = bsum[blockIdx.x] + inc - s;
    for (int i = 0; i < SCAN_I; ++i) {
        int64_t idx = base + i;
        if (idx < n) out[idx] = (Out)run;
        run += v[i];
    }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) out[n] = (Out)bsum[gridDim.x];
}

// Arrays of at most SCAN_ONE_N entries (offsets of small lists, digit
// histograms of small sorts) are scanned by one block in one launch instead
// of three: each thread sums a contiguous chunk, the block scans the 1024
// sums, and each thread writes its chunk's prefixes.  (One block is
// latency-bound: at 70k entries it measured ~100 us against ~20 us for the
// three-launch scan, so the limit is low.)
constexpr int SCAN_ONE_T = 1024;
constexpr int64_t SCAN_ONE_N = 1 << 14;

template <class In, class Out>
__global__ void __launch_bounds__(SCAN_ONE_T) kScanOne(const In* __restrict__ in, Out* __restrict__ out, int64_t n) {
    __shared__ int64_t wsum[SCAN_ONE_T / WAVE];
    // chunks of a multiple of 8 entries, read 8 independent loads at a time
    const int64_t per = ((n + SCAN_ONE_T - 1) / SCAN_ONE_T + 7) & ~(int64_t)7;
    const int64_t b = (int64_t)threadIdx.x * per, e = b + per < n ? b + per : n;
    int64_t s = 0;
    for (int64_t i = b; i < e; i += 8) {
        int64_t x[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) x[k] = i + k < e ? (int64_t)in[i + k] : 0;
#pragma unroll
        for (int k = 0; k < 8; ++k) s += x[k];
    }
    const int lane = laneId(), w = threadIdx.x / WAVE;
    int64_t v = s;
    for (int d = 1; d < WAVE; d <<= 1) {
        const int64_t o = __shfl_up(v, d);
        if (lane >= d) v += o;
    }
    if (lane == WAVE - 1) wsum[w] = v;
    __syncthreads();
    int64_t pre = 0, tot = 0;
    for (int i = 0; i < SCAN_ONE_T / WAVE; ++i) {
        if (i < w) pre += wsum[i];
        tot += wsum[i];
    }
    int64_t run = pre + v - s;
    for (int64_t i = b; i < e; i += 8) {
        int64_t x[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) x[k] = i + k < e ? (int64_t)in[i + k] : 0;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            if (i + k < e) out[i + k] = (Out)run;
            run += x[k];
        }
    }
    if (threadIdx.x == 0) out[n] = (Out)tot;
}

template <class In, class Out>
static void scanT(girhip_graph* h, const In* in, Out* out, int64_t n) {
    if (n <= 0) {
        fillT<Out>(h, out, (Out)0, 1);
        return;
    }
    if (n <= SCAN_ONE_N) {
        launch(h, "scan_one", kScanOne<In, Out>, 1, SCAN_ONE_T, 0, in, out, n);
        return;
    }
    const int64_t nb = (n + SCAN_TILE - 1) / SCAN_TILE;
    Scratch& S = scratch(h);
    S.off.ensure(nb + 1);
    launch(h, "scan_reduce", kScanReduce<In>, (unsigned)nb, SCAN_T, 0, in, n, S.off.p);
    launch(h, "scan_blocks", kScanBlockSums, 1, SCAN_T, 0, S.off.p, nb);
    launch(h, "scan_down", kScanDown<In, Out>, (unsigned)nb, SCAN_T, 0, in, out, n, (const int64_t*)S.off.p);
}

void scanI32toI64(girhip_graph* h, const int32_t* in, int64_t* out, int64_t n) { scanT(h, in, out, n); }
void scanI32toI32(girhip_graph* h, const int32_t* in, int32_t* out, int64_t n) { scanT(h, in, out, n); }

// ---------------------------------------------------------------------------
// LSD radix sort, 8-bit digits, one launch per pass.  Before the passes one
// kernel counts every pass's digits over the whole array (the global digit
// starts).  A pass kernel takes a tile of RS_TILE keys in the order tiles
// are handed out (a ticket counter, so every lower tile is already running),
// counts its digits, publishes the counts, and reads the counts of the tiles
// before it back to the first one that has published its running total
// (each tile publishes that total as soon as it knows it), which gives the
// tile's offset per digit; it then scatters its keys stably.  A published
// word is (epoch << 48) | (state << 46) | count, state 1 = this tile's
// count, 2 = running total through this tile; words of other passes carry
// other epochs and read as not yet published.  A wait longer than RS_SPIN
// polls gives up, sets the sort's error word and lets the kernel finish;
// the host raises the error at its next stream synchronisation.
constexpr int RS_T = 256, RS_R = 16, RS_TILE = RS_T * RS_R, RS_D = 256, RS_MAXP = 8;
static_assert(RS_D == RS_T, "a pass's threads own one digit each");
constexpr unsigned RS_SPIN = 1u << 22;
constexpr uint64_t RS_CNT = (1ULL << 46) - 1;

__global__ void __launch_bounds__(RS_T) kRadixCount(const uint64_t* __restrict__ keys, int64_t n, int lo, int passes,
                                                    unsigned long long* __restrict__ ghist) {
    __shared__ uint32_t c[RS_MAXP][RS_D];
    for (int p = 0; p < passes; ++p) c[p][threadIdx.x] = 0;
    __syncthreads();
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t k = keys[i] >> lo;
        for (int p = 0; p < passes; ++p) atomicAdd(&c[p][(k >> (8 * p)) & 255], 1u);
    }
    __syncthreads();
    for (int p = 0; p < passes; ++p)
        if (c[p][threadIdx.x]) atomicAdd(&ghist[p * RS_D + threadIdx.x], (unsigned long long)c[p][threadIdx.x]);
}

__global__ void __launch_bounds__(RS_T) kRadixPass(const uint64_t* __restrict__ kin, const uint32_t* __restrict__ vin,
                                                   uint64_t* __restrict__ kout, uint32_t* __restrict__ vout,
                                                   int64_t n, int shift, const unsigned long long* __restrict__ ghist,
                                                   uint64_t* look, uint32_t epoch, unsigned long long* tick,
                                                   int64_t* err) {
    __shared__ int64_t base[RS_D];
    __shared__ int32_t wc[RS_T / WAVE][RS_D];
    __shared__ int64_t tile_s;
    const int t = threadIdx.x, lane = laneId(), w = t / WAVE;
    if (t == 0) tile_s = (int64_t)atomicAdd(tick, 1ULL);
    for (int d = lane; d < RS_D; d += WAVE) wc[w][d] = 0;
    __syncthreads();
    const int64_t tile = tile_s;
    // wave w holds a contiguous RS_R * 64 keys of the tile, item r of lane l
    // at w0 + r * 64 + l: ranking wave by wave, item by item, lane by lane
    // is then the keys' own order (stable)
    const int64_t w0 = tile * RS_TILE + (int64_t)w * (RS_R * WAVE);
    uint64_t key[RS_R];
    uint32_t val[RS_R];
    int rk[RS_R];
#pragma unroll
    for (int r = 0; r < RS_R; ++r) {
        const int64_t idx = w0 + (int64_t)r * WAVE + lane;
        key[r] = idx < n ? kin[idx] : 0;
        val[r] = (idx < n && vin) ? vin[idx] : 0;
    }
    // rank of each key among the wave's keys of its digit (wave-private
    // running counts, no workgroup barrier per item)
#pragma unroll
    for (int r = 0; r < RS_R; ++r) {
        const bool valid = w0 + (int64_t)r * WAVE + lane < n;
        const int d = (int)((key[r] >> shift) & 255);
        uint64_t peers = __ballot(valid);
        for (int b = 0; b < 8; ++b) {
            const bool bit = (d >> b) & 1;
            const uint64_t m = __ballot(bit);
            peers &= bit ? m : ~m;
        }
        const int below = __popcll(peers & lanesBelow());
        rk[r] = valid ? wc[w][d] + below : 0;
        waveSync();
        if (valid && below == 0) wc[w][d] += __popcll(peers);
        waveSync();
    }
    __syncthreads();
    // digit t: the waves' counts -> exclusive prefix over the waves; the
    // tile's count
    int64_t mine = 0;
    for (int k = 0; k < RS_T / WAVE; ++k) {
        const int c = wc[k][t];
        wc[k][t] = (int32_t)mine;
        mine += c;
    }
    // this tile's offset for digit t: the counts of the tiles before it
    const uint64_t ep = (uint64_t)epoch << 48;
    int64_t ex = 0;
    uint64_t* slot = look + tile * RS_D + t;
    if (tile == 0) {
        stRel<uint64_t>(slot, ep | (2ULL << 46) | (uint64_t)mine);
    } else {
        stRel<uint64_t>(slot, ep | (1ULL << 46) | (uint64_t)mine);
        int64_t j = tile - 1;
        unsigned spins = 0;
        while (j >= 0) {
            const uint64_t x = ldAcq(look + j * RS_D + t);
            if ((uint32_t)(x >> 48) != epoch) {
                if (++spins >= RS_SPIN) {
                    stRel<int64_t>(err, 1);
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
                continue;
            }
            ex += (int64_t)(x & RS_CNT);
            if (((x >> 46) & 3) == 2) break;
            --j;
        }
        stRel<uint64_t>(slot, ep | (2ULL << 46) | (uint64_t)(ex + mine));
    }
    // global start of digit t (exclusive scan of the pass's digit counts)
    int64_t total;
    const int64_t g = (int64_t)ghist[t];
    const int64_t start = blockInclusiveScan(g, &total) - g;
    base[t] = start + ex;
    __syncthreads();
#pragma unroll
    for (int r = 0; r < RS_R; ++r) {
        if (w0 + (int64_t)r * WAVE + lane >= n) continue;
        const int d = (int)((key[r] >> shift) & 255);
        const int64_t off = base[d] + wc[w][d] + rk[r];
        kout[off] = key[r];
        if (vout) vout[off] = val[r];
    }
}

// Arrays of at most SS_N pairs (update batches of 10^3 ops, small DAG
// changes) are sorted by one block in LDS instead of key_bits / 8 radix
// passes of several launches each: a bitonic network over (masked key,
// original position), so the order is the stable one the radix sort gives.
constexpr int SS_N = 2048, SS_T = 1024;
__global__ void __launch_bounds__(SS_T) kSmallSort(uint64_t* keys, uint32_t* vals, int n, uint64_t mask) {
    __shared__ uint64_t sk[SS_N];
    __shared__ uint32_t sv[SS_N];
    __shared__ uint32_t si[SS_N];
    int P = 1;
    while (P < n) P <<= 1;
    for (int i = threadIdx.x; i < P; i += blockDim.x) {
        sk[i] = i < n ? keys[i] : ~0ULL;
        sv[i] = (i < n && vals) ? vals[i] : 0;
        si[i] = (uint32_t)i;
    }
    __syncthreads();
    for (int k = 2; k <= P; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = threadIdx.x; i < P; i += blockDim.x) {
                const int l = i ^ j;
                if (l > i) {
                    // padding (position >= n) sorts after every real entry
                    const uint64_t a = si[i] < (uint32_t)n ? sk[i] & mask : ~0ULL;
                    const uint64_t b = si[l] < (uint32_t)n ? sk[l] & mask : ~0ULL;
                    const bool gt = a > b || (a == b && si[i] > si[l]);
                    if (gt == ((i & k) == 0)) {
                        const uint64_t tk = sk[i];
                        sk[i] = sk[l];
                        sk[l] = tk;
                        const uint32_t tv = sv[i];
                        sv[i] = sv[l];
                        sv[l] = tv;
                        const uint32_t ti = si[i];
                        si[i] = si[l];
                        si[l] = ti;
                    }
                }
            }
            __syncthreads();
        }
    }
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        keys[i] = sk[i];
        if (vals) vals[i] = sv[i];
    }
}

void radixSort(girhip_graph* h, uint64_t* keys, uint32_t* vals, int64_t n, int key_bits, int lo_bit) {
    if (n <= 1 || lo_bit >= key_bits) return;
    if (n <= SS_N) {
        const uint64_t mask = (key_bits >= 64 ? ~0ULL : (1ULL << key_bits) - 1) & ~((1ULL << lo_bit) - 1);
        launch(h, "small_sort", kSmallSort, 1, SS_T, 0, keys, vals, (int)n, mask);
        return;
    }
    Scratch& S = scratch(h);
    const int64_t nb = (n + RS_TILE - 1) / RS_TILE;
    const int passes = std::min(RS_MAXP, std::max(1, (key_bits - lo_bit + 7) / 8));
    S.keys[1].ensure(n);
    if (vals) S.vals[1].ensure(n);
    // published words: one per (tile, digit); a new buffer starts cleared
    if (S.rs_look.n < (size_t)(nb * RS_D)) {
        S.rs_look.ensure((size_t)(nb * RS_D));
        fillU64(h, S.rs_look.p, 0, nb * RS_D);
    }
    if (!S.rs_err.p) {
        S.rs_err.ensure(1);
        fillI64(h, S.rs_err.p, 0, 1);
    }
    // [0, RS_MAXP): tickets per pass; then the digit counts per pass
    S.rs_aux.ensure((size_t)(RS_MAXP + RS_MAXP * RS_D));
    uint64_t* aux = S.rs_aux.p;
    fillU64(h, aux, 0, RS_MAXP + (int64_t)passes * RS_D);
    launch(h, "radix_count", kRadixCount, (unsigned)std::min<int64_t>(nb, 1024), RS_T, 0, (const uint64_t*)keys, n,
           lo_bit, passes, (unsigned long long*)(aux + RS_MAXP));
    uint64_t *ka = keys, *kb = S.keys[1].p;
    uint32_t *va = vals, *vb = vals ? S.vals[1].p : nullptr;
    for (int p = 0; p < passes; ++p) {
        if (++S.rs_epoch > 0xffffu) {  // epochs wrap: clear the words once
            fillU64(h, S.rs_look.p, 0, (int64_t)S.rs_look.n);
            S.rs_epoch = 1;
        }
        launch(h, "radix_pass", kRadixPass, (unsigned)nb, RS_T, 0, (const uint64_t*)ka, (const uint32_t*)va, kb, vb, n,
               lo_bit + 8 * p, (const unsigned long long*)(aux + RS_MAXP + p * RS_D), S.rs_look.p, S.rs_epoch,
               (unsigned long long*)(aux + p), S.rs_err.p);
        std::swap(ka, kb);
        std::swap(va, vb);
    }
    if (ka != keys) {
        HIPCHK(hipMemcpyAsync(keys, ka, sizeof(uint64_t) * n, hipMemcpyDeviceToDevice, h->st));
        if (vals) HIPCHK(hipMemcpyAsync(vals, va, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, h->st));
    }
    h->rs_check = true;
}

}  // namespace gir

// Small MLP for learned tuning: 16 inputs -> 64 -> 64 -> 2 outputs (ReLU),
// trained and evaluated on the GPU by the kernels below.  No counterpart in
// the reference (its constants are fixed: delta 16, TRIM_TIMES 3).
//
// Precision: fp32 master weights; the forward and backward passes use the
// weights rounded to bf16 (round to nearest even) and accumulate in fp32.
// Gradients are taken with respect to those bf16 weights and applied to the
// fp32 masters (Adam).  Loss: L = 1/(2n) * sum_i sum_o (y_io - t_io)^2.
//
// Parameter vector layout (fp32, NP entries):
//   W1[64][16] b1[64] W2[64][64] b2[64] W3[2][64] b3[2]
// Forward: one thread per sample, weights staged in LDS as fp32 values of
// the bf16-rounded weights.  Backward: 32 samples per block; per-sample
// deltas in registers, activations and deltas staged in LDS, outer products
// reduced over the block in LDS, one fp32 atomicAdd per parameter per block.
#include <cmath>
#include <cstring>
#include <vector>

#include "common.h"

namespace gir {
namespace mlp {

constexpr int NI = 16, NH = 64, NO = 2;
constexpr int OW1 = 0, OB1 = OW1 + NH * NI, OW2 = OB1 + NH, OB2 = OW2 + NH * NH, OW3 = OB2 + NH, OB3 = OW3 + NO * NH,
              NP = OB3 + NO;

__device__ __forceinline__ float bf16r(float x) {
    // round to nearest even at bf16 precision (NaN/Inf pass through)
    uint32_t u = __float_as_uint(x);
    if ((u & 0x7f800000u) == 0x7f800000u) return x;
    u += 0x7fffu + ((u >> 16) & 1u);
    return __uint_as_float(u & 0xffff0000u);
}

__device__ __forceinline__ void loadWeights(const float* __restrict__ P, float* s) {
    for (int i = threadIdx.x; i < NP; i += blockDim.x) s[i] = bf16r(P[i]);
    __syncthreads();
}

// y[n][2]; optional activations H1, H2 [n][64] (post-ReLU); optional loss
// sum (needs targets T).
__global__ void kForward(const float* __restrict__ P, const float* __restrict__ X, int n, float* __restrict__ Y,
                         float* __restrict__ H1, float* __restrict__ H2, const float* __restrict__ T, float* loss) {
    __shared__ float w[NP];
    loadWeights(P, w);
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    float l = 0;
    if (i < n) {
        float x[NI], h1[NH];
#pragma unroll
        for (int m = 0; m < NI; ++m) x[m] = X[(size_t)i * NI + m];
        for (int k = 0; k < NH; ++k) {
            float a = w[OB1 + k];
#pragma unroll
            for (int m = 0; m < NI; ++m) a += w[OW1 + k * NI + m] * x[m];
            h1[k] = a > 0 ? a : 0;
            if (H1) H1[(size_t)i * NH + k] = h1[k];
        }
        float y[NO];
#pragma unroll
        for (int o = 0; o < NO; ++o) y[o] = w[OB3 + o];
        for (int j = 0; j < NH; ++j) {
            float a = w[OB2 + j];
            for (int k = 0; k < NH; ++k) a += w[OW2 + j * NH + k] * h1[k];
            const float h2 = a > 0 ? a : 0;
            if (H2) H2[(size_t)i * NH + j] = h2;
#pragma unroll
            for (int o = 0; o < NO; ++o) y[o] += w[OW3 + o * NH + j] * h2;
        }
#pragma unroll
        for (int o = 0; o < NO; ++o) {
            Y[(size_t)i * NO + o] = y[o];
            if (T) {
                const float d = y[o] - T[(size_t)i * NO + o];
                l += d * d;
            }
        }
    }
    if (loss) {
        for (int d = WAVE / 2; d > 0; d >>= 1) l += __shfl_xor(l, d);
        if (laneId() == 0 && l != 0) atomicAdd(loss, l);
    }
}

constexpr int BS = 32;  // samples per backward block (LDS: ~56 KB per block)

// G += dL/dP for samples [blockIdx.x*BS, ...): blockDim = 256.
__global__ void kBackward(const float* __restrict__ P, const float* __restrict__ X, const float* __restrict__ Y,
                          const float* __restrict__ T, const float* __restrict__ H1, const float* __restrict__ H2, int n,
                          float* __restrict__ G) {
    __shared__ float w[NP];
    __shared__ float sx[BS][NI];
    __shared__ float sh1[BS][NH];
    __shared__ float sh2[BS][NH];
    __shared__ float sd2[BS][NH];  // dL/dz2
    __shared__ float sd1[BS][NH];  // dL/dz1
    __shared__ float sdy[BS][NO];
    loadWeights(P, w);
    const int base = blockIdx.x * BS;
    const int t = threadIdx.x;
    const float inv_n = 1.0f / (float)n;
    // stage inputs, activations and output deltas
    for (int q = t; q < BS * NH; q += blockDim.x) {
        const int s = q / NH, c = q % NH, i = base + s;
        const bool ok = i < n;
        sh1[s][c] = ok ? H1[(size_t)i * NH + c] : 0;
        sh2[s][c] = ok ? H2[(size_t)i * NH + c] : 0;
        if (c < NI) sx[s][c] = ok ? X[(size_t)i * NI + c] : 0;
        if (c < NO) sdy[s][c] = ok ? (Y[(size_t)i * NO + c] - T[(size_t)i * NO + c]) * inv_n : 0;
    }
    __syncthreads();
    // dz2[s][j] = sum_o dy[s][o] W3[o][j] * (h2 > 0)
    for (int q = t; q < BS * NH; q += blockDim.x) {
        const int s = q / NH, j = q % NH;
        float a = 0;
#pragma unroll
        for (int o = 0; o < NO; ++o) a += sdy[s][o] * w[OW3 + o * NH + j];
        sd2[s][j] = sh2[s][j] > 0 ? a : 0;
    }
    __syncthreads();
    // dz1[s][k] = sum_j dz2[s][j] W2[j][k] * (h1 > 0)
    for (int q = t; q < BS * NH; q += blockDim.x) {
        const int s = q / NH, k = q % NH;
        float a = 0;
        for (int j = 0; j < NH; ++j) a += sd2[s][j] * w[OW2 + j * NH + k];
        sd1[s][k] = sh1[s][k] > 0 ? a : 0;
    }
    __syncthreads();
    // parameter gradients, reduced over the block's samples
    for (int p = t; p < NP; p += blockDim.x) {
        float g = 0;
        if (p < OB1) {  // W1[k][m]
            const int k = (p - OW1) / NI, m = (p - OW1) % NI;
            for (int s = 0; s < BS; ++s) g += sd1[s][k] * sx[s][m];
        } else if (p < OW2) {  // b1[k]
            const int k = p - OB1;
            for (int s = 0; s < BS; ++s) g += sd1[s][k];
        } else if (p < OB2) {  // W2[j][k]
            const int j = (p - OW2) / NH, k = (p - OW2) % NH;
            for (int s = 0; s < BS; ++s) g += sd2[s][j] * sh1[s][k];
        } else if (p < OW3) {  // b2[j]
            const int j = p - OB2;
            for (int s = 0; s < BS; ++s) g += sd2[s][j];
        } else if (p < OB3) {  // W3[o][j]
            const int o = (p - OW3) / NH, j = (p - OW3) % NH;
            for (int s = 0; s < BS; ++s) g += sdy[s][o] * sh2[s][j];
        } else {  // b3[o]
            const int o = p - OB3;
            for (int s = 0; s < BS; ++s) g += sdy[s][o];
        }
        if (g != 0) atomicAdd(&G[p], g);
    }
}

// Adam on the fp32 masters (bias-corrected), step t >= 1.
__global__ void kAdam(float* __restrict__ P, const float* __restrict__ G, float* __restrict__ M, float* __restrict__ Vv,
                      float lr, float b1, float b2, float eps, int t) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= NP) return;
    const float g = G[p];
    const float m = b1 * M[p] + (1 - b1) * g;
    const float v = b2 * Vv[p] + (1 - b2) * g * g;
    M[p] = m;
    Vv[p] = v;
    const float mh = m / (1 - powf(b1, (float)t)), vh = v / (1 - powf(b2, (float)t));
    P[p] -= lr * mh / (sqrtf(vh) + eps);
}

}  // namespace mlp
}  // namespace gir

using namespace gir;
using namespace gir::mlp;

struct girhip_mlp {
    int device = 0;
    hipStream_t st = nullptr;
    DevBuf<float> P, G, M, Vv;
    DevBuf<float> X, Y, T, H1, H2, loss;
    int cap = 0;
    int step = 0;
};

namespace {
void ensureRows(girhip_mlp* m, int n) {
    if (n <= m->cap) return;
    m->X.ensure((size_t)n * NI);
    m->Y.ensure((size_t)n * NO);
    m->T.ensure((size_t)n * NO);
    m->H1.ensure((size_t)n * NH);
    m->H2.ensure((size_t)n * NH);
    m->cap = n;
}
void mlpSync(girhip_mlp* m) { HIPCHK(hipStreamSynchronize(m->st)); }

// forward (+ loss if T given) on the device copies of X (and T)
void fwd(girhip_mlp* m, int n, bool with_t, bool keep_act) {
    hipLaunchKernelGGL(kForward, dim3(gridFor(n, 128)), dim3(128), 0, m->st, (const float*)m->P.p, (const float*)m->X.p,
                       n, m->Y.p, keep_act ? m->H1.p : nullptr, keep_act ? m->H2.p : nullptr,
                       with_t ? (const float*)m->T.p : nullptr, with_t ? m->loss.p : nullptr);
    HIPCHK(hipGetLastError());
}
void bwd(girhip_mlp* m, int n) {
    HIPCHK(hipMemsetAsync(m->G.p, 0, sizeof(float) * NP, m->st));
    hipLaunchKernelGGL(kBackward, dim3((n + BS - 1) / BS), dim3(256), 0, m->st, (const float*)m->P.p,
                       (const float*)m->X.p, (const float*)m->Y.p, (const float*)m->T.p, (const float*)m->H1.p,
                       (const float*)m->H2.p, n, m->G.p);
    HIPCHK(hipGetLastError());
}
}  // namespace

extern "C" int girhip_mlp_create(uint64_t seed, girhip_mlp** out) {
    return guarded([&] {
        if (!out) raise(GIRHIP_EINVAL, "mlp_create: null output");
        int ndev = 0;
        if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
            (void)hipGetLastError();
            raise(GIRHIP_ENODEV, "mlp_create: no HIP device");
        }
        auto* m = new girhip_mlp();
        HIPCHK(hipGetDevice(&m->device));
        HIPCHK(hipStreamCreateWithFlags(&m->st, hipStreamNonBlocking));
        m->P.ensure(NP);
        m->G.ensure(NP);
        m->M.ensure(NP);
        m->Vv.ensure(NP);
        m->loss.ensure(1);
        // He-style init from a counter-based generator (host), zero biases
        std::vector<float> p(NP, 0.0f);
        auto u01 = [&](uint64_t i) {
            uint64_t x = seed * 0x9E3779B97F4A7C15ULL + i + 0x632BE59BD9B4E019ULL;
            x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
            x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
            x ^= x >> 31;
            return ((x >> 11) + 0.5) / 9007199254740992.0;
        };
        auto fill = [&](int off, int cnt, int fan_in) {
            const double a = std::sqrt(6.0 / fan_in);
            for (int i = 0; i < cnt; ++i) p[off + i] = (float)((2 * u01(off + i) - 1) * a);
        };
        fill(OW1, NH * NI, NI);
        fill(OW2, NH * NH, NH);
        fill(OW3, NO * NH, NH);
        HIPCHK(hipMemcpy(m->P.p, p.data(), sizeof(float) * NP, hipMemcpyHostToDevice));
        HIPCHK(hipMemset(m->M.p, 0, sizeof(float) * NP));
        HIPCHK(hipMemset(m->Vv.p, 0, sizeof(float) * NP));
        *out = m;
    });
}

extern "C" int girhip_mlp_destroy(girhip_mlp* m) {
    if (!m) return GIRHIP_OK;
    (void)hipSetDevice(m->device);
    if (m->st) (void)hipStreamSynchronize(m->st);
    if (m->st) (void)hipStreamDestroy(m->st);
    delete m;
    return GIRHIP_OK;
}

extern "C" int girhip_mlp_num_params(int32_t* n) {
    if (!n) return GIRHIP_EINVAL;
    *n = NP;
    return GIRHIP_OK;
}

extern "C" int girhip_mlp_get_params(girhip_mlp* m, float* p) {
    return guarded([&] {
        if (!m || !p) raise(GIRHIP_EINVAL, "mlp_get_params: null argument");
        HIPCHK(hipSetDevice(m->device));
        HIPCHK(hipMemcpy(p, m->P.p, sizeof(float) * NP, hipMemcpyDefault));
    });
}

extern "C" int girhip_mlp_set_params(girhip_mlp* m, const float* p) {
    return guarded([&] {
        if (!m || !p) raise(GIRHIP_EINVAL, "mlp_set_params: null argument");
        HIPCHK(hipSetDevice(m->device));
        HIPCHK(hipMemcpy(m->P.p, p, sizeof(float) * NP, hipMemcpyDefault));
        HIPCHK(hipMemset(m->M.p, 0, sizeof(float) * NP));
        HIPCHK(hipMemset(m->Vv.p, 0, sizeof(float) * NP));
        m->step = 0;
    });
}

extern "C" int girhip_mlp_forward(girhip_mlp* m, const float* x, int32_t n, float* y) {
    return guarded([&] {
        if (!m || n < 0 || (n > 0 && (!x || !y))) raise(GIRHIP_EINVAL, "mlp_forward: bad argument");
        if (n == 0) return;
        HIPCHK(hipSetDevice(m->device));
        ensureRows(m, n);
        HIPCHK(hipMemcpyAsync(m->X.p, x, sizeof(float) * n * NI, hipMemcpyDefault, m->st));
        fwd(m, n, false, false);
        HIPCHK(hipMemcpyAsync(y, m->Y.p, sizeof(float) * n * NO, hipMemcpyDefault, m->st));
        mlpSync(m);
    });
}

// Loss and gradient (w.r.t. the bf16-rounded weights) on one batch, without
// updating the weights.
extern "C" int girhip_mlp_grad(girhip_mlp* m, const float* x, const float* t, int32_t n, float* grad, float* loss) {
    return guarded([&] {
        if (!m || n <= 0 || !x || !t) raise(GIRHIP_EINVAL, "mlp_grad: bad argument");
        HIPCHK(hipSetDevice(m->device));
        ensureRows(m, n);
        HIPCHK(hipMemcpyAsync(m->X.p, x, sizeof(float) * n * NI, hipMemcpyDefault, m->st));
        HIPCHK(hipMemcpyAsync(m->T.p, t, sizeof(float) * n * NO, hipMemcpyDefault, m->st));
        HIPCHK(hipMemsetAsync(m->loss.p, 0, sizeof(float), m->st));
        fwd(m, n, true, true);
        bwd(m, n);
        float l = 0;
        HIPCHK(hipMemcpyAsync(&l, m->loss.p, sizeof(float), hipMemcpyDeviceToHost, m->st));
        if (grad) HIPCHK(hipMemcpyAsync(grad, m->G.p, sizeof(float) * NP, hipMemcpyDefault, m->st));
        mlpSync(m);
        if (loss) *loss = l / (2.0f * (float)n);
    });
}

// `epochs` full-batch passes over (x, t) of Adam with learning rate lr; *loss
// receives the loss of the last pass (before its update).
extern "C" int girhip_mlp_train(girhip_mlp* m, const float* x, const float* t, int32_t n, int32_t epochs, float lr,
                                float* loss) {
    return guarded([&] {
        if (!m || n <= 0 || epochs < 0 || !x || !t) raise(GIRHIP_EINVAL, "mlp_train: bad argument");
        HIPCHK(hipSetDevice(m->device));
        ensureRows(m, n);
        HIPCHK(hipMemcpyAsync(m->X.p, x, sizeof(float) * n * NI, hipMemcpyDefault, m->st));
        HIPCHK(hipMemcpyAsync(m->T.p, t, sizeof(float) * n * NO, hipMemcpyDefault, m->st));
        for (int e = 0; e < epochs; ++e) {
            HIPCHK(hipMemsetAsync(m->loss.p, 0, sizeof(float), m->st));
            fwd(m, n, true, true);
            bwd(m, n);
            ++m->step;
            hipLaunchKernelGGL(kAdam, dim3((NP + 255) / 256), dim3(256), 0, m->st, m->P.p, (const float*)m->G.p, m->M.p,
                               m->Vv.p, lr, 0.9f, 0.999f, 1e-8f, m->step);
            HIPCHK(hipGetLastError());
        }
        float l = 0;
        HIPCHK(hipMemcpyAsync(&l, m->loss.p, sizeof(float), hipMemcpyDeviceToHost, m->st));
        mlpSync(m);
        if (loss) *loss = l / (2.0f * (float)n);
    });
}

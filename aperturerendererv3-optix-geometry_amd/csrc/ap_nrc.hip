// Learned radiance cache: a 64 -> 64 x 5 -> 3 ReLU network without biases
// (neural radiance caching, Mueller et al. 2021), BF16 weights and
// activations, FP32 accumulation on gfx950 MFMA (v_mfma_f32_16x16x32_bf16),
// trained online with a relative L2 loss and Adam (include/aperture/ap_nrc.h).
//
// Every layer is computed transposed, y^T = W x^T, so one MFMA tile is 16
// outputs x 16 batch rows x 32 inputs:
//   A (16 x 32): lane l holds W[o = 16t + l%16][i = 32s + 8(l/16) .. +7]
//   B (32 x 16): lane l holds x[row = l%16][i = 32s + 8(l/16) .. +7]
//   C (16 x 16): lane l holds y[row = l%16][o = 16t + 4(l/16) .. +3]
// Both operands are one 16-byte load (weights and the wave's activation tile
// live in LDS, rows padded to 72 elements against bank conflicts); C goes
// back to LDS as the next layer's B after ReLU and BF16 rounding.
//
// Training keeps the activations and deltas of a batch in HBM as
// [feature][row] BF16 so that the weight gradient dW = delta^T a is again
// one 16-byte load per operand with the batch as the MFMA K dimension:
//   A: lane l holds delta[o = 16t + l%16][rows r0 + 8(l/16) .. +7]
//   B: lane l holds a[i = 16u + l%16][rows r0 + 8(l/16) .. +7]
// Per-chunk partial gradients are summed in chunk order (deterministic), and
// Adam updates the FP32 master weights and rewrites the BF16 copies.
#include <cmath>
#include <cstring>
#include <vector>

#include "aperture/ap_nrc.h"
#include "ap_fmath.h"
#include "ap_internal.h"
#include "ap_pcg.h"

namespace ap {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int NF = 64;          // features per layer
constexpr int NL = 5;           // hidden layers
constexpr int NOP = 16;         // padded output rows
constexpr int NOUT = 3;
constexpr int NPAR = NL * NF * NF + NOP * NF;  // 21504
constexpr int OUT_OFF = NL * NF * NF;          // 20480
constexpr int WT_OUT_STRIDE = 32;              // transposed output weights: [i][o < 32]
constexpr int NPAR_T = NL * NF * NF + NF * WT_OUT_STRIDE;
constexpr int LS = 72;                         // padded LDS row (elements)
constexpr int QREC = 16;                       // floats per query record
constexpr int SREC = 20;                       // floats per training sample (query + target rgb + pad)
constexpr int kChunkRows = 256;                // rows per weight-gradient partial
constexpr float kPiF = 3.14159265358979323846f;

__device__ __forceinline__ uint16_t f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x7fffffu)) return 0x7fc0;
    u += 0x7fffu + ((u >> 16) & 1u);
    return static_cast<uint16_t>(u >> 16);
}
__host__ __device__ __forceinline__ uint16_t f2bf_h(float f) {
    uint32_t u;
    __builtin_memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x7fffffu)) return 0x7fc0;
    u += 0x7fffu + ((u >> 16) & 1u);
    return static_cast<uint16_t>(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float(static_cast<uint32_t>(h) << 16); }

__device__ __forceinline__ bf16x8 ld8(const uint16_t* p) {
    const u16x8 v = *reinterpret_cast<const u16x8*>(p);
    return __builtin_bit_cast(bf16x8, v);
}

__device__ __forceinline__ f32x4 mfma(bf16x8 a, bf16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

struct Bounds {
    float lo[3], hi[3];
};

// Encoding of one query record into 64 features (the oracle's or_nrc_encode).
__device__ void encode(const float* s, const Bounds& b, float f[NF]) {
    for (int k = 0; k < NF; ++k) f[k] = 0.0f;
    for (int a = 0; a < 3; ++a) {
        const float p = (s[a] - b.lo[a]) / (b.hi[a] - b.lo[a]);
        for (int k = 0; k < 6; ++k) {
            const float arg = p * (static_cast<float>(1 << k) * kPiF);
            float sn, cs;
            apfm::sincos(arg, sn, cs);
            f[a * 12 + 2 * k] = sn;
            f[a * 12 + 2 * k + 1] = cs;
        }
    }
    for (int k = 0; k < 3; ++k) {
        f[36 + k] = s[3 + k];
        f[39 + k] = s[6 + k];
        f[43 + k] = s[10 + k];
    }
    f[42] = s[9];
    f[46] = 1.0f;
}

// Loads the BF16 weights of all layers into LDS rows of LS elements.
__device__ void load_weights(const uint16_t* __restrict__ w, uint16_t* sw) {
    for (int e = threadIdx.x; e < NPAR; e += blockDim.x) {
        const int l = e / (NF * NF);
        const int rem = e - l * NF * NF;  // hidden: o*64+i; output (l == 5): o*64+i, o < 16
        const int o = rem / NF, i = rem % NF;
        sw[(l * NF + o) * LS + i] = w[e];
    }
}

// Forward pass over rows [0, n) of records (stride floats each); with dyn_n
// over rows [0, min(*dyn_n, n)); with `rows`, row r reads record rows[r].
// With TRAIN, stores the input features and
// hidden activations as [feature][row] BF16 (acts + (layer * 64 + feature) *
// n_cap + row).
template <bool TRAIN>
__global__ __launch_bounds__(256) void forward_kernel(const uint16_t* __restrict__ w, Bounds bnd,
                                                      const float* __restrict__ rec, int stride, uint32_t n,
                                                      const uint32_t* __restrict__ dyn_n,
                                                      const int32_t* __restrict__ rows, float* __restrict__ out,
                                                      uint16_t* __restrict__ acts, uint32_t n_cap) {
    __shared__ __attribute__((aligned(16))) uint16_t sw[(NL * NF + NOP) * LS];
    __shared__ __attribute__((aligned(16))) uint16_t sa[4][16 * LS];
    if (dyn_n) n = min(*dyn_n, n);
    if (blockIdx.x * 64u >= n) return;
    load_weights(w, sw);
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int row = lane & 15, q = lane >> 4;
    uint16_t* act = sa[wave];
    for (uint32_t base = blockIdx.x * 64u; base < n; base += gridDim.x * 64u) {
        const uint32_t r = base + static_cast<uint32_t>(wave) * 16u + static_cast<uint32_t>(row);
        {
            float f[NF];
            if (r < n) {
                const size_t src = rows ? static_cast<size_t>(rows[r]) : static_cast<size_t>(r);
                encode(rec + src * stride, bnd, f);
            } else {
                for (int k = 0; k < NF; ++k) f[k] = 0.0f;
            }
            for (int k = 0; k < 16; ++k) {
                const uint16_t h = f2bf(f[16 * q + k]);
                act[row * LS + 16 * q + k] = h;
                if (TRAIN && r < n) acts[static_cast<size_t>(16 * q + k) * n_cap + r] = h;
            }
        }
        __syncthreads();
        for (int l = 0; l < NL; ++l) {
            f32x4 acc[4];
            for (int t = 0; t < 4; ++t) {
                acc[t] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
                for (int s = 0; s < 2; ++s) {
                    const bf16x8 a = ld8(&sw[(l * NF + 16 * t + row) * LS + 32 * s + 8 * q]);
                    const bf16x8 b = ld8(&act[row * LS + 32 * s + 8 * q]);
                    acc[t] = mfma(a, b, acc[t]);
                }
            }
            __syncthreads();
            for (int t = 0; t < 4; ++t)
                for (int k = 0; k < 4; ++k) {
                    const int o = 16 * t + 4 * q + k;
                    const float v = acc[t][k] > 0.0f ? acc[t][k] : 0.0f;
                    const uint16_t h = f2bf(v);
                    act[row * LS + o] = h;
                    if (TRAIN && r < n) acts[static_cast<size_t>((l + 1) * NF + o) * n_cap + r] = h;
                }
            __syncthreads();
        }
        f32x4 y{0.0f, 0.0f, 0.0f, 0.0f};
        for (int s = 0; s < 2; ++s) {
            const bf16x8 a = ld8(&sw[(NL * NF + row) * LS + 32 * s + 8 * q]);
            const bf16x8 b = ld8(&act[row * LS + 32 * s + 8 * q]);
            y = mfma(a, b, y);
        }
        if (q == 0 && r < n) {
            out[static_cast<size_t>(r) * NOUT + 0] = y[0];
            out[static_cast<size_t>(r) * NOUT + 1] = y[1];
            out[static_cast<size_t>(r) * NOUT + 2] = y[2];
        }
        __syncthreads();
    }
}

// Relative L2 loss: dL/dy = 2 (y - t) / (y^2 + 0.01) / (3n), stored BF16 as
// [o][row] (o < 16, rows 3..15 zero); per-block loss partials. A gate that
// reads 0 zeroes every delta (an empty sample pool).
__global__ __launch_bounds__(256) void loss_kernel(const float* __restrict__ y, const float* __restrict__ rec,
                                                   uint32_t n, uint32_t n_cap, const uint32_t* __restrict__ gate,
                                                   uint16_t* __restrict__ dout, float* __restrict__ loss_part) {
    __shared__ float red[256];
    const uint32_t r = blockIdx.x * 256u + threadIdx.x;
    const bool open = gate == nullptr || *gate != 0u;
    float l = 0.0f;
    if (r < n && !open) {
        for (int o = 0; o < NOP; ++o) dout[static_cast<size_t>(o) * n_cap + r] = 0;
    } else if (r < n) {
        const float norm = 1.0f / (3.0f * static_cast<float>(n));
        for (int o = 0; o < NOP; ++o) {
            uint16_t h = 0;
            if (o < NOUT) {
                const float v = y[static_cast<size_t>(r) * NOUT + o];
                const float t = rec[static_cast<size_t>(r) * SREC + QREC + o];
                const float den = v * v + 0.01f;
                const float e = v - t;
                l += e * e / den * norm;
                h = f2bf(2.0f * e / den * norm);
            }
            dout[static_cast<size_t>(o) * n_cap + r] = h;
        }
    }
    red[threadIdx.x] = l;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (static_cast<int>(threadIdx.x) < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) loss_part[blockIdx.x] = red[0];
}

// Delta chain: delta_4 = (W_out^T dout) * [a_4 > 0], then
// delta_{h-1} = (W_h^T delta_h) * [a_{h-1} > 0], stored [feature][row] BF16.
__global__ __launch_bounds__(256) void backward_kernel(const uint16_t* __restrict__ wt, uint32_t n, uint32_t n_cap,
                                                       const uint16_t* __restrict__ acts,
                                                       const uint16_t* __restrict__ dout,
                                                       uint16_t* __restrict__ deltas) {
    __shared__ __attribute__((aligned(16))) uint16_t swt[NL * NF * LS + NF * 40];
    __shared__ __attribute__((aligned(16))) uint16_t sd[4][16 * LS];
    for (int e = threadIdx.x; e < NPAR_T; e += blockDim.x) {
        if (e < NL * NF * NF) {
            const int l = e / (NF * NF), rem = e % (NF * NF);
            swt[(l * NF + rem / NF) * LS + rem % NF] = wt[e];
        } else {
            const int rem = e - NL * NF * NF;
            swt[NL * NF * LS + (rem / WT_OUT_STRIDE) * 40 + rem % WT_OUT_STRIDE] = wt[e];
        }
    }
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int row = lane & 15, q = lane >> 4;
    uint16_t* del = sd[wave];
    for (uint32_t base = blockIdx.x * 64u; base < n; base += gridDim.x * 64u) {
        const uint32_t r = base + static_cast<uint32_t>(wave) * 16u + static_cast<uint32_t>(row);
        const bool live = r < n;
        // Output deltas as [row][o < 32] (o >= 16 zero).
        for (int k = 0; k < 8; ++k) {
            const int o = 8 * q + k;
            del[row * LS + o] = (o < NOP && live) ? dout[static_cast<size_t>(o) * n_cap + r] : uint16_t{0};
        }
        __syncthreads();
        f32x4 acc[4];
        for (int t = 0; t < 4; ++t) {
            const bf16x8 a = ld8(&swt[NL * NF * LS + (16 * t + row) * 40 + 8 * q]);
            const bf16x8 b = ld8(&del[row * LS + 8 * q]);
            acc[t] = mfma(a, b, f32x4{0.0f, 0.0f, 0.0f, 0.0f});
        }
        __syncthreads();
        for (int h = NL - 1; h >= 0; --h) {
            // acc holds W_{h+1}^T delta_{h+1} (W_out for h = 4) at i = 16t + 4q + k.
            for (int t = 0; t < 4; ++t)
                for (int k = 0; k < 4; ++k) {
                    const int i = 16 * t + 4 * q + k;
                    uint16_t hv = 0;
                    if (live) {
                        const uint16_t a = acts[static_cast<size_t>((h + 1) * NF + i) * n_cap + r];
                        hv = bf2f(a) > 0.0f ? f2bf(acc[t][k]) : uint16_t{0};
                        deltas[static_cast<size_t>(h * NF + i) * n_cap + r] = hv;
                    }
                    del[row * LS + i] = hv;
                }
            __syncthreads();
            if (h == 0) break;
            for (int t = 0; t < 4; ++t) {
                acc[t] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
                for (int s = 0; s < 2; ++s) {
                    const bf16x8 a = ld8(&swt[(h * NF + 16 * t + row) * LS + 32 * s + 8 * q]);
                    const bf16x8 b = ld8(&del[row * LS + 32 * s + 8 * q]);
                    acc[t] = mfma(a, b, acc[t]);
                }
            }
            __syncthreads();
        }
    }
}

// Weight-gradient partials of one row chunk: block (chunk, layer).
__global__ __launch_bounds__(256) void wgrad_kernel(uint32_t n_cap, const uint16_t* __restrict__ acts,
                                                    const uint16_t* __restrict__ deltas,
                                                    const uint16_t* __restrict__ dout, float* __restrict__ part) {
    const int chunk = blockIdx.x, layer = blockIdx.y;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int col = lane & 15, q = lane >> 4;
    const uint32_t r0 = static_cast<uint32_t>(chunk) * kChunkRows;
    const bool out_layer = layer == NL;
    const int tiles = out_layer ? 4 : 16;
    const uint16_t* dl = out_layer ? dout : deltas + static_cast<size_t>(layer) * NF * n_cap;
    const uint16_t* al = acts + static_cast<size_t>(layer) * NF * n_cap;  // input of this layer
    float* dst = part + static_cast<size_t>(chunk) * NPAR + static_cast<size_t>(layer) * NF * NF;
    for (int tile = wave; tile < tiles; tile += 4) {
        const int to = out_layer ? 0 : tile / 4, ti = out_layer ? tile : tile % 4;
        f32x4 acc{0.0f, 0.0f, 0.0f, 0.0f};
        for (uint32_t k = 0; k < kChunkRows; k += 32) {
            const uint32_t rr = r0 + k + 8u * static_cast<uint32_t>(q);
            const bf16x8 a = ld8(&dl[static_cast<size_t>(16 * to + col) * n_cap + rr]);
            const bf16x8 b = ld8(&al[static_cast<size_t>(16 * ti + col) * n_cap + rr]);
            acc = mfma(a, b, acc);
        }
        for (int k = 0; k < 4; ++k) {
            const int o = 16 * to + 4 * q + k, i = 16 * ti + col;
            dst[o * NF + i] = acc[k];
        }
    }
}

__global__ void reduce_grad_kernel(const float* __restrict__ part, int chunks, float* __restrict__ grad) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= NPAR) return;
    float s = 0.0f;
    for (int c = 0; c < chunks; ++c) s += part[static_cast<size_t>(c) * NPAR + p];
    grad[p] = s;
}

// Adam (the oracle's or_nrc_adam), then the BF16 copies W[o][i] and W^T[i][o]
// of the training weights, and the exponential moving average of the
// parameters (decay ema_decay, debiased by ema_c = 1 - decay^t) whose BF16
// copy answers inference.
__global__ void adam_kernel(float* __restrict__ params, float* __restrict__ m, float* __restrict__ v,
                            const float* __restrict__ grad, float lr, float beta1, float beta2, float eps, float c1,
                            float c2, uint16_t* __restrict__ w, uint16_t* __restrict__ wt,
                            const uint32_t* __restrict__ gate, float ema_decay, float ema_c,
                            float* __restrict__ ema, uint16_t* __restrict__ w_inf) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= NPAR) return;
    if (gate && *gate == 0u) return;
    const float g = grad[p];
    const float mm = beta1 * m[p] + (1.0f - beta1) * g;
    const float vv = beta2 * v[p] + (1.0f - beta2) * (g * g);
    m[p] = mm;
    v[p] = vv;
    const float mh = mm / c1;
    const float vh = vv / c2;
    const float x = params[p] - lr * mh / (sqrtf(vh) + eps);
    params[p] = x;
    const uint16_t h = f2bf(x);
    w[p] = h;
    const int l = p / (NF * NF), rem = p % (NF * NF), o = rem / NF, i = rem % NF;
    if (l < NL) wt[l * NF * NF + i * NF + o] = h;
    else wt[NL * NF * NF + i * WT_OUT_STRIDE + o] = h;
    const float e = ema_decay * ema[p] + (1.0f - ema_decay) * x;
    ema[p] = e;
    w_inf[p] = ema_decay > 0.0f ? f2bf(e / ema_c) : h;
}

// Row j of step `step`: sample (off + (step * batch + j) * P) mod N of the
// pool, P = 2^31 - 1 a prime above any pool size, so that the rows of the
// steps of a frame walk the pool without repetition while they last.
__global__ __launch_bounds__(256) void gather_kernel(const float* __restrict__ pool,
                                                     const uint32_t* __restrict__ pool_n, uint32_t batch,
                                                     uint32_t step, uint64_t off, float* __restrict__ out) {
    const uint32_t j = blockIdx.x * 256u + threadIdx.x;
    if (j >= batch) return;
    const uint32_t n = *pool_n;
    float4* dst = reinterpret_cast<float4*>(out + static_cast<size_t>(j) * SREC);
    if (n == 0u) {
        for (int k = 0; k < SREC / 4; ++k) dst[k] = float4{0.0f, 0.0f, 0.0f, 0.0f};
        return;
    }
    const uint64_t idx =
        (off % n + (static_cast<uint64_t>(step) * batch + j) % n * 2147483647ull % n) % n;
    const float4* src = reinterpret_cast<const float4*>(pool + idx * SREC);
    for (int k = 0; k < SREC / 4; ++k) dst[k] = src[k];
}

}  // namespace

struct NrcState {
    float *params = nullptr, *m = nullptr, *v = nullptr, *grad = nullptr;
    uint16_t *w = nullptr, *wt = nullptr;
    float* ema = nullptr;          // moving average of the parameters (zero-started, debiased)
    uint16_t* w_inf = nullptr;     // BF16 weights inference reads: the debiased average, or w
    float ema_decay = 0.0f;
    Bounds bounds{{-1.0f, -1.0f, -1.0f}, {1.0f, 1.0f, 1.0f}};
    int step = 0;
    uint32_t cap = 0;
    uint16_t *acts = nullptr, *deltas = nullptr, *dout = nullptr;
    float *y = nullptr, *part = nullptr, *loss_part = nullptr;
    float last_loss = 0.0f;
    float* batch = nullptr;   // gathered training rows of nrc_train_pool
    uint32_t batch_cap = 0;
};

namespace {

int upload_params(NrcState* s, const std::vector<float>& p) {
    std::vector<uint16_t> w(NPAR), wt(NPAR_T, 0);
    for (int k = 0; k < NPAR; ++k) {
        w[k] = f2bf_h(p[k]);
        const int l = k / (NF * NF), rem = k % (NF * NF), o = rem / NF, i = rem % NF;
        if (l < NL) wt[l * NF * NF + i * NF + o] = w[k];
        else wt[NL * NF * NF + i * WT_OUT_STRIDE + o] = w[k];
    }
    hipError_t e = hipMemcpy(s->params, p.data(), NPAR * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(s->w, w.data(), NPAR * sizeof(uint16_t), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(s->w_inf, w.data(), NPAR * sizeof(uint16_t), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(s->ema, 0, NPAR * sizeof(float));
    if (e == hipSuccess) e = hipMemcpy(s->wt, wt.data(), NPAR_T * sizeof(uint16_t), hipMemcpyHostToDevice);
    return e == hipSuccess ? AP_OK : hip_fail(e, "ap_nrc: upload parameters");
}

int ensure_train(NrcState* s, uint32_t n) {
    if (s->cap >= n) return AP_OK;
    void* ptrs[] = {s->acts, s->deltas, s->dout, s->y, s->part, s->loss_part};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    s->acts = s->deltas = s->dout = nullptr;
    s->y = s->part = s->loss_part = nullptr;
    s->cap = 0;
    const size_t N = n;
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&s->acts), (NL + 1) * NF * N * sizeof(uint16_t));
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&s->deltas), NL * NF * N * sizeof(uint16_t));
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&s->dout), NOP * N * sizeof(uint16_t));
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&s->y), N * NOUT * sizeof(float));
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&s->part), (N / kChunkRows) * NPAR * sizeof(float));
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&s->loss_part), ((N + 255) / 256) * sizeof(float));
    if (e != hipSuccess) return hip_fail(e, "ap_nrc: training buffers");
    s->cap = n;
    return AP_OK;
}

inline unsigned blocks_for(uint32_t n) {
    const uint32_t b = (n + 63) / 64;
    return b < 2048u ? (b > 0 ? b : 1u) : 2048u;
}

}  // namespace

int nrc_infer(NrcState* s, const float* queries, size_t n, float* out, hipStream_t st) {
    if (n == 0) return AP_OK;
    hipLaunchKernelGGL(forward_kernel<false>, dim3(blocks_for(static_cast<uint32_t>(n))), dim3(256), 0, st,
                       s->w_inf, s->bounds, queries, QREC, static_cast<uint32_t>(n), nullptr, nullptr, out, nullptr, 0u);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? AP_OK : hip_fail(e, "ap_nrc: infer");
}

// Forward, loss and backward of one batch; the gradient lands in s->grad.
int nrc_gradient(NrcState* s, const float* samples, uint32_t n, hipStream_t st, const uint32_t* gate = nullptr) {
    int rc = ensure_train(s, n);
    if (rc != AP_OK) return rc;
    hipLaunchKernelGGL(forward_kernel<true>, dim3(blocks_for(n)), dim3(256), 0, st, s->w, s->bounds, samples, SREC, n,
                       nullptr, nullptr, s->y, s->acts, n);
    hipLaunchKernelGGL(loss_kernel, dim3((n + 255) / 256), dim3(256), 0, st, s->y, samples, n, n, gate, s->dout,
                       s->loss_part);
    hipLaunchKernelGGL(backward_kernel, dim3(blocks_for(n)), dim3(256), 0, st, s->wt, n, n, s->acts, s->dout,
                       s->deltas);
    hipLaunchKernelGGL(wgrad_kernel, dim3(n / kChunkRows, NL + 1), dim3(256), 0, st, n, s->acts, s->deltas, s->dout,
                       s->part);
    hipLaunchKernelGGL(reduce_grad_kernel, dim3((NPAR + 255) / 256), dim3(256), 0, st, s->part,
                       static_cast<int>(n / kChunkRows), s->grad);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? AP_OK : hip_fail(e, "ap_nrc: gradient");
}

int nrc_adam(NrcState* s, float lr, hipStream_t st, const uint32_t* gate = nullptr) {
    s->step += 1;
    const float b1 = 0.9f, b2 = 0.99f, eps = 1e-8f;
    const float c1 = 1.0f - std::pow(b1, static_cast<float>(s->step));
    const float c2 = 1.0f - std::pow(b2, static_cast<float>(s->step));
    const float ec = 1.0f - std::pow(s->ema_decay, static_cast<float>(s->step));
    hipLaunchKernelGGL(adam_kernel, dim3((NPAR + 255) / 256), dim3(256), 0, st, s->params, s->m, s->v, s->grad, lr,
                       b1, b2, eps, c1, c2, s->w, s->wt, gate, s->ema_decay, ec, s->ema, s->w_inf);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? AP_OK : hip_fail(e, "ap_nrc: adam");
}

int nrc_infer_dyn(NrcState* s, const float* queries, const int32_t* rows, const uint32_t* dyn_n, uint32_t n_cap,
                  float* out, hipStream_t st) {
    if (n_cap == 0) return AP_OK;
    hipLaunchKernelGGL(forward_kernel<false>, dim3(blocks_for(n_cap)), dim3(256), 0, st, s->w_inf, s->bounds,
                       queries, QREC, n_cap, dyn_n, rows, out, nullptr, 0u);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? AP_OK : hip_fail(e, "ap_nrc: infer");
}

int nrc_train_pool(NrcState* s, const float* pool, const uint32_t* pool_n, uint32_t batch, uint32_t steps, float lr,
                   uint64_t seed, hipStream_t st) {
    if (steps == 0) return AP_OK;
    if (batch == 0 || batch % kChunkRows != 0 || batch > (1u << 24))
        return fail(AP_ERR_INVALID_ARG, "ap_nrc: batch must be a positive multiple of 256 up to 2^24");
    if (s->batch_cap < batch) {
        if (s->batch) (void)hipFree(s->batch);
        s->batch = nullptr;
        s->batch_cap = 0;
        const hipError_t e = hipMalloc(reinterpret_cast<void**>(&s->batch), static_cast<size_t>(batch) * SREC * 4);
        if (e != hipSuccess) return hip_fail(e, "ap_nrc: batch buffer");
        s->batch_cap = batch;
    }
    const uint64_t off = splitmix64(seed);
    for (uint32_t k = 0; k < steps; ++k) {
        hipLaunchKernelGGL(gather_kernel, dim3((batch + 255) / 256), dim3(256), 0, st, pool, pool_n, batch, k, off,
                           s->batch);
        int rc = nrc_gradient(s, s->batch, batch, st, pool_n);
        if (rc == AP_OK) rc = nrc_adam(s, lr, st, pool_n);
        if (rc != AP_OK) return rc;
    }
    return AP_OK;
}

void nrc_set_bounds(NrcState* s, const float lo[3], const float hi[3]) {
    for (int a = 0; a < 3; ++a) {
        s->bounds.lo[a] = lo[a];
        s->bounds.hi[a] = hi[a];
    }
}

void nrc_free(NrcState* s) {
    if (!s) return;
    void* ptrs[] = {s->params, s->m, s->v, s->grad, s->w, s->wt, s->acts, s->deltas, s->dout, s->y, s->part,
                    s->loss_part, s->batch, s->ema, s->w_inf};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    delete s;
}

}  // namespace ap

using namespace ap;

struct ap_nrc {
    NrcState* s;
};

ap::NrcState* ap::nrc_state(ap_nrc* h) { return h ? h->s : nullptr; }

extern "C" {

int ap_nrc_create(uint64_t seed, ap_nrc** out) {
    if (!out) return fail(AP_ERR_INVALID_ARG, "ap_nrc_create: NULL output");
    *out = nullptr;
    NrcState* s = new NrcState();
    hipError_t e = hipSuccess;
    for (float** p : {&s->params, &s->m, &s->v, &s->grad, &s->ema})
        if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(p), NPAR * sizeof(float));
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&s->w), NPAR * sizeof(uint16_t));
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&s->w_inf), NPAR * sizeof(uint16_t));
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&s->wt), NPAR_T * sizeof(uint16_t));
    if (e == hipSuccess) e = hipMemset(s->m, 0, NPAR * sizeof(float));
    if (e == hipSuccess) e = hipMemset(s->v, 0, NPAR * sizeof(float));
    if (e == hipSuccess) e = hipMemset(s->grad, 0, NPAR * sizeof(float));
    if (e != hipSuccess) {
        nrc_free(s);
        return hip_fail(e, "ap_nrc_create");
    }
    // He-normal initial weights (hidden sd sqrt(2/64), output sd sqrt(1/64), padded output rows 0)
    // from PCG32(seed) by Box-Muller.
    std::vector<float> p(NPAR, 0.0f);
    Pcg32 rng = pcg32_seed(seed, 7);
    auto normal = [&]() {
        const float u1 = 1.0f - pcg32_float(rng), u2 = pcg32_float(rng);
        return std::sqrt(-2.0f * std::log(u1)) * std::cos(6.283185307f * u2);
    };
    for (int k = 0; k < OUT_OFF; ++k) p[k] = normal() * std::sqrt(2.0f / NF);
    for (int o = 0; o < NOUT; ++o)
        for (int i = 0; i < NF; ++i) p[OUT_OFF + o * NF + i] = normal() * std::sqrt(1.0f / NF);
    const int rc = upload_params(s, p);
    if (rc != AP_OK) {
        nrc_free(s);
        return rc;
    }
    *out = new ap_nrc{s};
    return AP_OK;
}

int ap_nrc_destroy(ap_nrc* nrc) {
    if (!nrc) return AP_OK;
    nrc_free(nrc->s);
    delete nrc;
    return AP_OK;
}

int ap_nrc_set_bounds(ap_nrc* nrc, const float bmin[3], const float bmax[3]) {
    if (!nrc || !bmin || !bmax) return fail(AP_ERR_INVALID_ARG, "ap_nrc_set_bounds: NULL argument");
    for (int a = 0; a < 3; ++a) {
        if (!(bmax[a] > bmin[a])) return fail(AP_ERR_INVALID_ARG, "ap_nrc_set_bounds: empty bounds");
        nrc->s->bounds.lo[a] = bmin[a];
        nrc->s->bounds.hi[a] = bmax[a];
    }
    return AP_OK;
}

int ap_nrc_get_params(const ap_nrc* nrc, float* host_params) {
    if (!nrc || !host_params) return fail(AP_ERR_INVALID_ARG, "ap_nrc_get_params: NULL argument");
    const hipError_t e = hipMemcpy(host_params, nrc->s->params, NPAR * sizeof(float), hipMemcpyDeviceToHost);
    return e == hipSuccess ? AP_OK : hip_fail(e, "ap_nrc_get_params");
}

int ap_nrc_set_params(ap_nrc* nrc, const float* host_params) {
    if (!nrc || !host_params) return fail(AP_ERR_INVALID_ARG, "ap_nrc_set_params: NULL argument");
    std::vector<float> p(host_params, host_params + NPAR);
    for (int o = NOUT; o < NOP; ++o)
        for (int i = 0; i < NF; ++i) p[OUT_OFF + o * NF + i] = 0.0f;
    nrc->s->step = 0;
    hipError_t e = hipMemset(nrc->s->m, 0, NPAR * sizeof(float));
    if (e == hipSuccess) e = hipMemset(nrc->s->v, 0, NPAR * sizeof(float));
    if (e != hipSuccess) return hip_fail(e, "ap_nrc_set_params");
    return upload_params(nrc->s, p);
}

int ap_nrc_infer(ap_nrc* nrc, const float* queries, size_t n, float* out_rgb, void* stream) {
    if (!nrc) return fail(AP_ERR_INVALID_ARG, "ap_nrc_infer: NULL cache");
    if (n == 0) return AP_OK;
    if (!queries || !out_rgb) return fail(AP_ERR_INVALID_ARG, "ap_nrc_infer: NULL buffer");
    if (n > 0x7fffffffull) return fail(AP_ERR_INVALID_ARG, "ap_nrc_infer: too many queries");
    return nrc_infer(nrc->s, queries, n, out_rgb, static_cast<hipStream_t>(stream));
}

int ap_nrc_gradient(ap_nrc* nrc, const float* samples, size_t n, float* host_grad, float* loss, void* stream) {
    if (!nrc || !samples) return fail(AP_ERR_INVALID_ARG, "ap_nrc_gradient: NULL argument");
    if (n == 0 || n % kChunkRows != 0 || n > (1u << 24))
        return fail(AP_ERR_INVALID_ARG, "ap_nrc_gradient: batch must be a positive multiple of 256 up to 2^24");
    const hipStream_t st = static_cast<hipStream_t>(stream);
    int rc = nrc_gradient(nrc->s, samples, static_cast<uint32_t>(n), st);
    if (rc != AP_OK) return rc;
    hipError_t e = hipStreamSynchronize(st);
    if (e == hipSuccess && host_grad)
        e = hipMemcpy(host_grad, nrc->s->grad, NPAR * sizeof(float), hipMemcpyDeviceToHost);
    if (e == hipSuccess && loss) {
        std::vector<float> lp((n + 255) / 256);
        e = hipMemcpy(lp.data(), nrc->s->loss_part, lp.size() * sizeof(float), hipMemcpyDeviceToHost);
        float sum = 0.0f;
        for (float x : lp) sum += x;
        *loss = sum;
    }
    return e == hipSuccess ? AP_OK : hip_fail(e, "ap_nrc_gradient");
}

int ap_nrc_train_step(ap_nrc* nrc, const float* samples, size_t n, float lr, void* stream) {
    if (!nrc || !samples) return fail(AP_ERR_INVALID_ARG, "ap_nrc_train_step: NULL argument");
    if (n == 0 || n % kChunkRows != 0 || n > (1u << 24))
        return fail(AP_ERR_INVALID_ARG, "ap_nrc_train_step: batch must be a positive multiple of 256 up to 2^24");
    const hipStream_t st = static_cast<hipStream_t>(stream);
    int rc = nrc_gradient(nrc->s, samples, static_cast<uint32_t>(n), st);
    if (rc != AP_OK) return rc;
    return nrc_adam(nrc->s, lr, st);
}

int ap_nrc_set_ema(ap_nrc* nrc, float decay) {
    if (!nrc) return fail(AP_ERR_INVALID_ARG, "ap_nrc_set_ema: NULL cache");
    if (!(decay >= 0.0f && decay < 1.0f)) return fail(AP_ERR_INVALID_ARG, "ap_nrc_set_ema: decay must be in [0, 1)");
    nrc->s->ema_decay = decay;
    return AP_OK;
}

int ap_nrc_get_bounds(const ap_nrc* nrc, float bmin[3], float bmax[3]) {
    if (!nrc || !bmin || !bmax) return fail(AP_ERR_INVALID_ARG, "ap_nrc_get_bounds: NULL argument");
    for (int a = 0; a < 3; ++a) {
        bmin[a] = nrc->s->bounds.lo[a];
        bmax[a] = nrc->s->bounds.hi[a];
    }
    return AP_OK;
}

int ap_nrc_last_gradient(const ap_nrc* nrc, float* host_grad) {
    if (!nrc || !host_grad) return fail(AP_ERR_INVALID_ARG, "ap_nrc_last_gradient: NULL argument");
    const hipError_t e = hipMemcpy(host_grad, nrc->s->grad, NPAR * sizeof(float), hipMemcpyDeviceToHost);
    return e == hipSuccess ? AP_OK : hip_fail(e, "ap_nrc_last_gradient");
}

}  // extern "C"

// Learned radiance cache: 64 input features -> W x 5 -> 3 ReLU networks
// without biases (neural radiance caching, Mueller et al. 2021) of hidden
// width W = 64 or 32, BF16 or FP16 weights and activations, FP32
// accumulation on gfx950 MFMA (v_mfma_f32_16x16x32_bf16 / _f16), trained
// online with a relative L2 loss and Adam (include/aperture/ap_nrc.h). A
// cache holds G independent networks of one shape (one per scene of a
// batch); every kernel serves all of them in one launch as a grouped GEMM:
// a 64-row tile belongs to one group and the block loads that group's
// weights.
//
// Every layer is computed transposed, y^T = W x^T, so one MFMA tile is 16
// outputs x 16 batch rows x 32 inputs:
//   A (16 x 32): lane l holds W[o = 16t + l%16][i = 32s + 8(l/16) .. +7]
//   B (32 x 16): lane l holds x[row = l%16][i = 32s + 8(l/16) .. +7]
//   C (16 x 16): lane l holds y[row = l%16][o = 16t + 4(l/16) .. +3]
// Both operands are one 16-byte load (weights and the wave's activation tile
// live in LDS, rows padded to 72 elements against bank conflicts); C goes
// back to LDS as the next layer's B after ReLU and rounding.
//
// Training keeps the activations and deltas of a batch in HBM as
// [feature][row] 16-bit values so that the weight gradient dW = delta^T a is
// again one 16-byte load per operand with the batch as the MFMA K dimension:
//   A: lane l holds delta[o = 16t + l%16][rows r0 + 8(l/16) .. +7]
//   B: lane l holds a[i = 16u + l%16][rows r0 + 8(l/16) .. +7]
// The rows of a training batch are laid out [group][batch]. Per-chunk
// partial gradients are summed in chunk order (deterministic), and Adam
// updates the FP32 master weights and rewrites the 16-bit copies. FP16
// networks scale the loss gradient by a power of two (exactly undone on the
// FP32 gradient) so that it stays clear of FP16's subnormal range, and
// saturate conversions at +-65504.
#include <atomic>
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
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int NF = 64;                         // input features
constexpr int NL = 5;                          // hidden layers
constexpr int NOP = 16;                        // padded output rows
constexpr int NOUT = 3;
constexpr int WT_OUT_STRIDE = 32;              // transposed output weights: [i][o < 32]
constexpr int LS = 72;                         // padded LDS row (elements)
constexpr int QREC = 16;                       // floats per query record
constexpr int SREC = 20;                       // floats per training sample (query + target rgb + pad)
constexpr int kChunkRows = 256;                // rows per weight-gradient partial
constexpr uint32_t kMaxGroups = 4096;
constexpr float kPiF = 3.14159265358979323846f;

// Parameter layout of a width-W network: W x 64 input matrix, four W x W
// hidden matrices, the 16 x W output matrix (rows 3..15 zero), each W[o][i].
template <int W>
struct Geo {
    static constexpr int NPAR = W * NF + 4 * W * W + NOP * W;
    static constexpr int NPAR_T = W * NF + 4 * W * W + WT_OUT_STRIDE * W;
    static constexpr int ROWS = NL * W + NOP;  // LDS rows of the forward weights
    static constexpr int LT = W + 8;           // LDS row of a transposed hidden matrix
    static constexpr int NACT = NF + NL * W;   // stored activation features per row
    __host__ __device__ static constexpr int off(int l) { return l == 0 ? 0 : W * NF + (l - 1) * W * W; }
    __host__ __device__ static constexpr int act(int l) { return l == 0 ? 0 : NF + (l - 1) * W; }
    // Layer, output and input of parameter p.
    __host__ __device__ static void decode(int p, int& l, int& o, int& i) {
        if (p < W * NF) {
            l = 0;
            o = p / NF;
            i = p % NF;
            return;
        }
        p -= W * NF;
        if (p < 4 * W * W) {
            l = 1 + p / (W * W);
            p %= W * W;
            o = p / W;
            i = p % W;
            return;
        }
        p -= 4 * W * W;
        l = NL;
        o = p / W;
        i = p % W;
    }
    // Index of parameter (l, o, i) in the transposed copy: W_l^T[i][o].
    __host__ __device__ static int tindex(int l, int o, int i) {
        return l < NL ? off(l) + i * W + o : off(NL) + i * WT_OUT_STRIDE + o;
    }
};

__device__ __forceinline__ uint16_t f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x7fffffu)) return 0x7fc0;
    u += 0x7fffu + ((u >> 16) & 1u);
    return static_cast<uint16_t>(u >> 16);
}
uint16_t f2bf_h(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x7fffffu)) return 0x7fc0;
    u += 0x7fffu + ((u >> 16) & 1u);
    return static_cast<uint16_t>(u >> 16);
}
// FP32 -> FP16, round to nearest even, saturated at +-65504 (host).
uint16_t f2h_h(float f) {
    if (f > 65504.0f) f = 65504.0f;
    if (f < -65504.0f) f = -65504.0f;
    uint32_t u;
    std::memcpy(&u, &f, 4);
    const uint16_t sign = static_cast<uint16_t>((u >> 16) & 0x8000u);
    u &= 0x7fffffffu;
    if (u > 0x7f800000u) return static_cast<uint16_t>(sign | 0x7e00u);
    if (u >= 0x38800000u) {  // at least 2^-14: normal
        uint32_t r = u - (112u << 23);
        r += 0xfffu + ((r >> 13) & 1u);
        return static_cast<uint16_t>(sign | (r >> 13));
    }
    if (u <= 0x33000000u) return sign;  // at most 2^-25: rounds to zero
    const uint32_t e = u >> 23, m = (u & 0x7fffffu) | 0x800000u;
    const uint32_t sh = 126u - e;  // value in units of 2^-24 is m / 2^sh
    uint32_t q = m >> sh;
    const uint32_t rem = m & ((1u << sh) - 1u), half = 1u << (sh - 1u);
    if (rem > half || (rem == half && (q & 1u))) ++q;
    return static_cast<uint16_t>(sign | q);
}

// Rounding of an FP32 value to the network's 16-bit type.
template <bool F16>
__device__ __forceinline__ uint16_t cvt(float f) {
    if constexpr (F16) {
        const float c = f > 65504.0f ? 65504.0f : (f < -65504.0f ? -65504.0f : f);
        return __builtin_bit_cast(uint16_t, static_cast<_Float16>(c));
    } else {
        return f2bf(f);
    }
}
// A stored activation is above zero (BF16 and FP16 share the sign bit).
__device__ __forceinline__ bool positive(uint16_t h) { return (h & 0x8000u) == 0u && (h & 0x7fffu) != 0u; }

__device__ __forceinline__ u16x8 ld8(const uint16_t* p) { return *reinterpret_cast<const u16x8*>(p); }

// Orders one wave's LDS accesses: a wave's LDS operations complete in issue
// order, so a tile that only its own wave reads and writes needs no block
// barrier, only this fence against compiler reordering.
__device__ __forceinline__ uint32_t pack2(uint16_t lo, uint16_t hi) {
    return static_cast<uint32_t>(lo) | (static_cast<uint32_t>(hi) << 16);
}
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <bool F16>
__device__ __forceinline__ f32x4 mfma(u16x8 a, u16x8 b, f32x4 c) {
    if constexpr (F16)
        return __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(f16x8, a), __builtin_bit_cast(f16x8, b), c,
                                                      0, 0, 0);
    else
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b),
                                                       c, 0, 0, 0);
}

struct Bounds {
    float lo[3], hi[3];
};

// Rows of each network. groups == 0: one network over rows [0, n).
// rows_per_group != 0: the training layout, group g owns rows
// [g * rows_per_group, (g + 1) * rows_per_group). Otherwise region g is rows
// [seg[g], seg[g + 1]) of which the first count[g] are live (all without
// `count`), cut into 64-row tiles numbered from tile0[g].
struct GroupMap {
    uint32_t groups = 0;
    uint32_t rows_per_group = 0;
    const uint32_t* seg = nullptr;
    const uint32_t* tile0 = nullptr;
    const uint32_t* count = nullptr;
    const Bounds* bounds = nullptr;
};

// Group and live rows [begin, end) of tile v; false for a tile with no live row.
__device__ __forceinline__ bool tile_rows(const GroupMap& gm, uint32_t v, uint32_t n, uint32_t& g, uint32_t& begin,
                                          uint32_t& end) {
    if (gm.groups == 0 || gm.rows_per_group != 0) {
        begin = v * 64u;
        g = gm.rows_per_group ? begin / gm.rows_per_group : 0u;
        end = gm.rows_per_group ? min(n, (g + 1u) * gm.rows_per_group) : n;
        return begin < end;
    }
    uint32_t lo = 0, hi = gm.groups;  // largest g with tile0[g] <= v
    while (hi - lo > 1u) {
        const uint32_t mid = (lo + hi) >> 1;
        if (gm.tile0[mid] <= v) lo = mid;
        else hi = mid;
    }
    g = lo;
    begin = gm.seg[g] + (v - gm.tile0[g]) * 64u;
    uint32_t size = gm.seg[g + 1] - gm.seg[g];
    if (gm.count) size = min(size, gm.count[g]);
    end = gm.seg[g] + size;
    return begin < end;
}

__device__ __forceinline__ uint32_t tile_count(const GroupMap& gm, uint32_t n) {
    if (gm.groups == 0 || gm.rows_per_group != 0) return (n + 63u) / 64u;
    return gm.tile0[gm.groups];
}

// The encoding of one query record (the oracle's or_nrc_encode), feature
// pair (2m, 2m + 1): for m < 18 the sine and cosine of the normalised
// position coordinate a = m / 6 at frequency 2^(m mod 6) pi; above, the
// direction (36..38), normal (39..41), roughness (42), albedo (43..45), a
// constant 1 (46) and zeros. p[a] = (s[a] - lo[a]) / (hi[a] - lo[a]).
__device__ __forceinline__ float plain_feature(const float* s, int k) {
    if (k < 39) return s[3 + (k - 36)];
    if (k < 42) return s[6 + (k - 39)];
    if (k == 42) return s[9];
    if (k < 46) return s[10 + (k - 43)];
    return k == 46 ? 1.0f : 0.0f;
}
__device__ __forceinline__ void encode_pair(const float* s, const float p[3], int m, float& v0, float& v1) {
    if (m < 18) {
        const int a = m / 6, fr = m % 6;
        const float pa = a == 0 ? p[0] : (a == 1 ? p[1] : p[2]);
        apfm::sincos(pa * (static_cast<float>(1 << fr) * kPiF), v0, v1);
    } else {
        v0 = plain_feature(s, 2 * m);
        v1 = plain_feature(s, 2 * m + 1);
    }
}

// Loads one network's 16-bit weights into LDS rows of LS elements (layer l,
// output o at row l * W + o).
template <int W>
__device__ void load_weights(const uint16_t* __restrict__ w, uint16_t* sw) {
    for (int e = threadIdx.x; e < Geo<W>::NPAR / 8; e += blockDim.x) {
        int l, o, i;
        Geo<W>::decode(e * 8, l, o, i);
        *reinterpret_cast<u16x8*>(&sw[(l * W + o) * LS + i]) = ld8(&w[e * 8]);
    }
}

// Forward pass over the rows of `gm` (see GroupMap; with groups == 0 over
// rows [0, n), or [0, min(*dyn_n, n)) with dyn_n); with `rows`, row r reads
// record rows[r]. With TRAIN, stores the input features and hidden
// activations as [feature][row] (acts + (Geo::act(layer) + feature) * n_cap
// + row).
template <bool TRAIN, int W, bool F16>
__global__ __launch_bounds__(256) void forward_kernel(const uint16_t* __restrict__ w, Bounds bnd, GroupMap gm,
                                                      const float* __restrict__ rec, int stride, uint32_t n,
                                                      const uint32_t* __restrict__ dyn_n,
                                                      const int32_t* __restrict__ rows, float* __restrict__ out,
                                                      uint16_t* __restrict__ acts, uint32_t n_cap) {
    using G = Geo<W>;
    __shared__ __attribute__((aligned(16))) uint16_t sw[G::ROWS * LS];
    __shared__ __attribute__((aligned(16))) uint16_t sa[4][16 * LS];
    if (dyn_n) n = min(*dyn_n, n);
    const uint32_t tiles = tile_count(gm, n);
    if (blockIdx.x >= tiles) return;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int row = lane & 15, q = lane >> 4;
    uint16_t* act = sa[wave];
    uint32_t loaded = 0xffffffffu;
    for (uint32_t v = blockIdx.x; v < tiles; v += gridDim.x) {
        uint32_t g, begin, end;
        if (!tile_rows(gm, v, n, g, begin, end)) continue;  // uniform over the block
        if (g != loaded) {
            __syncthreads();
            load_weights<W>(w + static_cast<size_t>(g) * G::NPAR, sw);
            __syncthreads();
            loaded = g;
        }
        const Bounds b = gm.bounds ? gm.bounds[g] : bnd;
        const uint32_t r = begin + static_cast<uint32_t>(wave) * 16u + static_cast<uint32_t>(row);
        const bool live = r < end;
        {
            // The four lanes of a row share its 18 sine/cosine pairs: lane q
            // encodes pairs q, q + 4, ..., q + 28.
            const size_t src = !live ? 0u : rows ? static_cast<size_t>(rows[r]) : static_cast<size_t>(r);
            const float* s = rec + src * stride;
            float p[3] = {0.0f, 0.0f, 0.0f};
            if (live)
                for (int a = 0; a < 3; ++a) p[a] = (s[a] - b.lo[a]) / (b.hi[a] - b.lo[a]);
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int m = 4 * j + q;
                float v0 = 0.0f, v1 = 0.0f;
                if (live) encode_pair(s, p, m, v0, v1);
                const uint16_t lo = cvt<F16>(v0), hi = cvt<F16>(v1);
                *reinterpret_cast<uint32_t*>(&act[row * LS + 2 * m]) = pack2(lo, hi);
                if (TRAIN && live) {
                    acts[static_cast<size_t>(2 * m) * n_cap + r] = lo;
                    acts[static_cast<size_t>(2 * m + 1) * n_cap + r] = hi;
                }
            }
        }
        wave_sync();
        for (int l = 0; l < NL; ++l) {
            f32x4 acc[W / 16];
            const int ks = (l == 0 ? NF : W) / 32;
            for (int t = 0; t < W / 16; ++t) {
                acc[t] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
                for (int s = 0; s < ks; ++s) {
                    const u16x8 a = ld8(&sw[(l * W + 16 * t + row) * LS + 32 * s + 8 * q]);
                    const u16x8 bb = ld8(&act[row * LS + 32 * s + 8 * q]);
                    acc[t] = mfma<F16>(a, bb, acc[t]);
                }
            }
            wave_sync();
#pragma unroll
            for (int t = 0; t < W / 16; ++t) {
                uint16_t h[4];
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const float x = acc[t][k] > 0.0f ? acc[t][k] : 0.0f;
                    h[k] = cvt<F16>(x);
                    if (TRAIN && live) acts[static_cast<size_t>(G::act(l + 1) + 16 * t + 4 * q + k) * n_cap + r] = h[k];
                }
                *reinterpret_cast<uint2*>(&act[row * LS + 16 * t + 4 * q]) = make_uint2(pack2(h[0], h[1]), pack2(h[2], h[3]));
            }
            wave_sync();
        }
        f32x4 y{0.0f, 0.0f, 0.0f, 0.0f};
        for (int s = 0; s < W / 32; ++s) {
            const u16x8 a = ld8(&sw[(NL * W + row) * LS + 32 * s + 8 * q]);
            const u16x8 bb = ld8(&act[row * LS + 32 * s + 8 * q]);
            y = mfma<F16>(a, bb, y);
        }
        if (q == 0 && live) {
            out[static_cast<size_t>(r) * NOUT + 0] = y[0];
            out[static_cast<size_t>(r) * NOUT + 1] = y[1];
            out[static_cast<size_t>(r) * NOUT + 2] = y[2];
        }
        wave_sync();
    }
}

// Relative L2 loss of each group's mean: dL/dy = 2 (y - t) / (y^2 + 0.01) /
// (3 rows_per_group), times `scale`, stored as [o][row] (o < 16, rows 3..15
// zero); per-block loss partials (unscaled). A gate that reads 0 (gate[g]
// with gate_per_group) zeroes every delta of its rows (an empty sample pool).
template <bool F16>
__global__ __launch_bounds__(256) void loss_kernel(const float* __restrict__ y, const float* __restrict__ rec,
                                                   uint32_t n, uint32_t rows_per_group, uint32_t n_cap,
                                                   const uint32_t* __restrict__ gate, int gate_per_group,
                                                   float scale, uint16_t* __restrict__ dout,
                                                   float* __restrict__ loss_part) {
    __shared__ float red[256];
    const uint32_t r = blockIdx.x * 256u + threadIdx.x;
    const uint32_t g = r / rows_per_group;
    const bool open = gate == nullptr || gate[gate_per_group ? g : 0u] != 0u;
    float l = 0.0f;
    if (r < n && !open) {
        for (int o = 0; o < NOP; ++o) dout[static_cast<size_t>(o) * n_cap + r] = 0;
    } else if (r < n) {
        const float norm = 1.0f / (3.0f * static_cast<float>(rows_per_group));
        const float gnorm = norm * scale;
        for (int o = 0; o < NOP; ++o) {
            uint16_t h = 0;
            if (o < NOUT) {
                const float v = y[static_cast<size_t>(r) * NOUT + o];
                const float t = rec[static_cast<size_t>(r) * SREC + QREC + o];
                const float den = v * v + 0.01f;
                const float e = v - t;
                l += e * e / den * norm;
                h = cvt<F16>(2.0f * e / den * gnorm);
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
// delta_{h-1} = (W_h^T delta_h) * [a_{h-1} > 0], stored [feature][row].
// Rows [g * rows_per_group, ...) use network g.
template <int W, bool F16>
__global__ __launch_bounds__(256) void backward_kernel(const uint16_t* __restrict__ wt, uint32_t n,
                                                       uint32_t rows_per_group, uint32_t n_cap,
                                                       const uint16_t* __restrict__ acts,
                                                       const uint16_t* __restrict__ dout,
                                                       uint16_t* __restrict__ deltas) {
    using G = Geo<W>;
    constexpr int LT = G::LT;
    constexpr int OUT_ROW0 = 4 * W * LT;  // transposed output matrix: rows i, 40 elements
    __shared__ __attribute__((aligned(16))) uint16_t swt[4 * W * LT + W * 40];
    __shared__ __attribute__((aligned(16))) uint16_t sd[4][16 * LS];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int row = lane & 15, q = lane >> 4;
    uint16_t* del = sd[wave];
    uint32_t loaded = 0xffffffffu;
    for (uint32_t base = blockIdx.x * 64u; base < n; base += gridDim.x * 64u) {
        const uint32_t g = base / rows_per_group;
        if (g != loaded) {
            __syncthreads();
            const uint16_t* src = wt + static_cast<size_t>(g) * G::NPAR_T;
            // Hidden layers 1..4 transposed ([i][o], W x W) and the output layer ([i][o < 32]).
            for (int e = threadIdx.x; e < (4 * W * W + W * WT_OUT_STRIDE) / 8; e += blockDim.x) {
                const int x = e * 8;
                if (x < 4 * W * W) {
                    const int h = x / (W * W), rem = x % (W * W);
                    *reinterpret_cast<u16x8*>(&swt[(h * W + rem / W) * LT + rem % W]) =
                        ld8(&src[G::off(1) + x]);
                } else {
                    const int rem = x - 4 * W * W;
                    *reinterpret_cast<u16x8*>(&swt[OUT_ROW0 + (rem / WT_OUT_STRIDE) * 40 + rem % WT_OUT_STRIDE]) =
                        ld8(&src[G::off(NL) + rem]);
                }
            }
            __syncthreads();
            loaded = g;
        }
        const uint32_t r = base + static_cast<uint32_t>(wave) * 16u + static_cast<uint32_t>(row);
        const bool live = r < n;
        // Output deltas as [row][o < 32] (o >= 16 zero).
        {
            uint16_t d[8];
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                const int o = 8 * q + k;
                d[k] = (o < NOP && live) ? dout[static_cast<size_t>(o) * n_cap + r] : uint16_t{0};
            }
            *reinterpret_cast<uint4*>(&del[row * LS + 8 * q]) =
                make_uint4(pack2(d[0], d[1]), pack2(d[2], d[3]), pack2(d[4], d[5]), pack2(d[6], d[7]));
        }
        wave_sync();
        f32x4 acc[W / 16];
        for (int t = 0; t < W / 16; ++t) {
            const u16x8 a = ld8(&swt[OUT_ROW0 + (16 * t + row) * 40 + 8 * q]);
            const u16x8 b = ld8(&del[row * LS + 8 * q]);
            acc[t] = mfma<F16>(a, b, f32x4{0.0f, 0.0f, 0.0f, 0.0f});
        }
        wave_sync();
        for (int h = NL - 1; h >= 0; --h) {
            // acc holds W_{h+1}^T delta_{h+1} (W_out for h = 4) at i = 16t + 4q + k.
#pragma unroll
            for (int t = 0; t < W / 16; ++t) {
                uint16_t hv[4] = {0, 0, 0, 0};
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const int i = 16 * t + 4 * q + k;
                    if (live) {
                        const uint16_t a = acts[static_cast<size_t>(G::act(h + 1) + i) * n_cap + r];
                        hv[k] = positive(a) ? cvt<F16>(acc[t][k]) : uint16_t{0};
                        deltas[static_cast<size_t>(h * W + i) * n_cap + r] = hv[k];
                    }
                }
                *reinterpret_cast<uint2*>(&del[row * LS + 16 * t + 4 * q]) =
                    make_uint2(pack2(hv[0], hv[1]), pack2(hv[2], hv[3]));
            }
            wave_sync();
            if (h == 0) break;
            for (int t = 0; t < W / 16; ++t) {
                acc[t] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
                for (int s = 0; s < W / 32; ++s) {
                    const u16x8 a = ld8(&swt[((h - 1) * W + 16 * t + row) * LT + 32 * s + 8 * q]);
                    const u16x8 b = ld8(&del[row * LS + 32 * s + 8 * q]);
                    acc[t] = mfma<F16>(a, b, acc[t]);
                }
            }
            wave_sync();
        }
    }
}

// Forward, relative-L2 loss and backward of 64-row tiles in one launch: the
// forward_kernel<true>, loss_kernel and backward_kernel sequence with each
// wave's activations kept in LDS for the backward's ReLU masks and its
// output deltas computed where the forward ends. The MFMA operands and their
// order are those of the three kernels, so acts, dout and deltas (stored
// [feature][row] for the weight gradients) are bit-identical to theirs; the
// loss is summed per 64-row tile (loss_part[tile]). Rows [g * rows_per_group,
// ...) use network g; n is a multiple of 64.
template <int W, bool F16>
__global__ __launch_bounds__(256) void fwd_bwd_kernel(const uint16_t* __restrict__ w, const uint16_t* __restrict__ wt,
                                                      Bounds bnd, const Bounds* __restrict__ bounds,
                                                      const float* __restrict__ rec, uint32_t n,
                                                      uint32_t rows_per_group, const uint32_t* __restrict__ gate,
                                                      int gate_per_group, float scale, uint16_t* __restrict__ acts,
                                                      uint16_t* __restrict__ dout, uint16_t* __restrict__ deltas,
                                                      float* __restrict__ loss_part) {
    using G = Geo<W>;
    constexpr int LT = G::LT;
    constexpr int OUT_ROW0 = 4 * W * LT;
    __shared__ __attribute__((aligned(16))) uint16_t sw[G::ROWS * LS];
    __shared__ __attribute__((aligned(16))) uint16_t swt[4 * W * LT + W * 40];
    // Per wave: the input features and the five hidden activations of its 16
    // rows; the input tile doubles as the backward's delta tile.
    __shared__ __attribute__((aligned(16))) uint16_t sa[4][NL + 1][16 * LS];
    __shared__ float red[256];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int row = lane & 15, q = lane >> 4;
    const uint32_t tiles = n / 64u;
    uint32_t loaded = 0xffffffffu;
    for (uint32_t v = blockIdx.x; v < tiles; v += gridDim.x) {
        const uint32_t base = v * 64u;
        const uint32_t g = base / rows_per_group;
        if (g != loaded) {
            __syncthreads();
            load_weights<W>(w + static_cast<size_t>(g) * G::NPAR, sw);
            const uint16_t* src = wt + static_cast<size_t>(g) * G::NPAR_T;
            for (int e = threadIdx.x; e < (4 * W * W + W * WT_OUT_STRIDE) / 8; e += blockDim.x) {
                const int x = e * 8;
                if (x < 4 * W * W) {
                    const int h = x / (W * W), rem = x % (W * W);
                    *reinterpret_cast<u16x8*>(&swt[(h * W + rem / W) * LT + rem % W]) = ld8(&src[G::off(1) + x]);
                } else {
                    const int rem = x - 4 * W * W;
                    *reinterpret_cast<u16x8*>(&swt[OUT_ROW0 + (rem / WT_OUT_STRIDE) * 40 + rem % WT_OUT_STRIDE]) =
                        ld8(&src[G::off(NL) + rem]);
                }
            }
            __syncthreads();
            loaded = g;
        }
        const Bounds b = bounds ? bounds[g] : bnd;
        const uint32_t r = base + static_cast<uint32_t>(wave) * 16u + static_cast<uint32_t>(row);
        uint16_t* a0 = sa[wave][0];
        {
            const float* sp = rec + static_cast<size_t>(r) * SREC;
            float p[3];
            for (int a = 0; a < 3; ++a) p[a] = (sp[a] - b.lo[a]) / (b.hi[a] - b.lo[a]);
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int m = 4 * j + q;
                float v0, v1;
                encode_pair(sp, p, m, v0, v1);
                const uint16_t lo = cvt<F16>(v0), hi = cvt<F16>(v1);
                *reinterpret_cast<uint32_t*>(&a0[row * LS + 2 * m]) = pack2(lo, hi);
                acts[static_cast<size_t>(2 * m) * n + r] = lo;
                acts[static_cast<size_t>(2 * m + 1) * n + r] = hi;
            }
        }
        wave_sync();
        for (int l = 0; l < NL; ++l) {
            const uint16_t* in = sa[wave][l];
            uint16_t* out = sa[wave][l + 1];
            f32x4 acc[W / 16];
            const int ks = (l == 0 ? NF : W) / 32;
            for (int t = 0; t < W / 16; ++t) {
                acc[t] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
                for (int s2 = 0; s2 < ks; ++s2) {
                    const u16x8 a = ld8(&sw[(l * W + 16 * t + row) * LS + 32 * s2 + 8 * q]);
                    const u16x8 bb = ld8(&in[row * LS + 32 * s2 + 8 * q]);
                    acc[t] = mfma<F16>(a, bb, acc[t]);
                }
            }
#pragma unroll
            for (int t = 0; t < W / 16; ++t) {
                uint16_t h[4];
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const float x = acc[t][k] > 0.0f ? acc[t][k] : 0.0f;
                    h[k] = cvt<F16>(x);
                    acts[static_cast<size_t>(G::act(l + 1) + 16 * t + 4 * q + k) * n + r] = h[k];
                }
                *reinterpret_cast<uint2*>(&out[row * LS + 16 * t + 4 * q]) =
                    make_uint2(pack2(h[0], h[1]), pack2(h[2], h[3]));
            }
            wave_sync();
        }
        f32x4 y{0.0f, 0.0f, 0.0f, 0.0f};
        for (int s2 = 0; s2 < W / 32; ++s2) {
            const u16x8 a = ld8(&sw[(NL * W + row) * LS + 32 * s2 + 8 * q]);
            const u16x8 bb = ld8(&sa[wave][NL][row * LS + 32 * s2 + 8 * q]);
            y = mfma<F16>(a, bb, y);
        }
        // loss_kernel's arithmetic on the row's three outputs (lanes q = 0
        // hold them); the output deltas go to dout ([o][row], o < 16) and,
        // as the backward's input tile, to the wave's [row][o < 32] tile.
        float lrow = 0.0f;
        uint16_t* del = a0;
        {
            const bool open = gate == nullptr || gate[gate_per_group ? g : 0u] != 0u;
            uint16_t d[8] = {0, 0, 0, 0, 0, 0, 0, 0};
            if (q == 0 && open) {
                const float norm = 1.0f / (3.0f * static_cast<float>(rows_per_group));
                const float gnorm = norm * scale;
                for (int o = 0; o < NOUT; ++o) {
                    const float vv = y[o];
                    const float tt = rec[static_cast<size_t>(r) * SREC + QREC + o];
                    const float den = vv * vv + 0.01f;
                    const float e = vv - tt;
                    lrow += e * e / den * norm;
                    d[o] = cvt<F16>(2.0f * e / den * gnorm);
                }
            }
            if (q < 2)
                for (int k = 0; k < 8; ++k) dout[static_cast<size_t>(8 * q + k) * n + r] = d[k];
            *reinterpret_cast<uint4*>(&del[row * LS + 8 * q]) =
                make_uint4(pack2(d[0], d[1]), pack2(d[2], d[3]), pack2(d[4], d[5]), pack2(d[6], d[7]));
        }
        red[threadIdx.x] = lrow;
        __syncthreads();
        for (int s2 = 128; s2 > 0; s2 >>= 1) {
            if (static_cast<int>(threadIdx.x) < s2) red[threadIdx.x] += red[threadIdx.x + s2];
            __syncthreads();
        }
        if (threadIdx.x == 0) loss_part[v] = red[0];
        // backward_kernel's delta chain, the ReLU masks from the LDS tiles.
        f32x4 acc[W / 16];
        for (int t = 0; t < W / 16; ++t) {
            const u16x8 a = ld8(&swt[OUT_ROW0 + (16 * t + row) * 40 + 8 * q]);
            const u16x8 bb = ld8(&del[row * LS + 8 * q]);
            acc[t] = mfma<F16>(a, bb, f32x4{0.0f, 0.0f, 0.0f, 0.0f});
        }
        wave_sync();
        for (int h = NL - 1; h >= 0; --h) {
            const uint16_t* act = sa[wave][h + 1];
#pragma unroll
            for (int t = 0; t < W / 16; ++t) {
                uint16_t hv[4];
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const int i = 16 * t + 4 * q + k;
                    hv[k] = positive(act[row * LS + i]) ? cvt<F16>(acc[t][k]) : uint16_t{0};
                    deltas[static_cast<size_t>(h * W + i) * n + r] = hv[k];
                }
                *reinterpret_cast<uint2*>(&del[row * LS + 16 * t + 4 * q]) =
                    make_uint2(pack2(hv[0], hv[1]), pack2(hv[2], hv[3]));
            }
            wave_sync();
            if (h == 0) break;
            for (int t = 0; t < W / 16; ++t) {
                acc[t] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
                for (int s2 = 0; s2 < W / 32; ++s2) {
                    const u16x8 a = ld8(&swt[((h - 1) * W + 16 * t + row) * LT + 32 * s2 + 8 * q]);
                    const u16x8 bb = ld8(&del[row * LS + 32 * s2 + 8 * q]);
                    acc[t] = mfma<F16>(a, bb, acc[t]);
                }
            }
            wave_sync();
        }
    }
}

// Weight-gradient partials of one row chunk: block (chunk, layer); the
// chunk's partial is in its group's parameter order.
template <int W, bool F16>
__global__ __launch_bounds__(256) void wgrad_kernel(uint32_t n_cap, const uint16_t* __restrict__ acts,
                                                    const uint16_t* __restrict__ deltas,
                                                    const uint16_t* __restrict__ dout, float* __restrict__ part) {
    using G = Geo<W>;
    const int chunk = blockIdx.x, layer = blockIdx.y;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int col = lane & 15, q = lane >> 4;
    const uint32_t r0 = static_cast<uint32_t>(chunk) * kChunkRows;
    const bool out_layer = layer == NL;
    const int ni = layer == 0 ? NF : W;
    const int tiles = (out_layer ? 1 : W / 16) * (ni / 16);
    const uint16_t* dl = out_layer ? dout : deltas + static_cast<size_t>(layer) * W * n_cap;
    const uint16_t* al = acts + static_cast<size_t>(G::act(layer)) * n_cap;  // input of this layer
    float* dst = part + static_cast<size_t>(chunk) * G::NPAR + G::off(layer);
    for (int tile = wave; tile < tiles; tile += 4) {
        const int to = tile / (ni / 16), ti = tile % (ni / 16);
        f32x4 acc{0.0f, 0.0f, 0.0f, 0.0f};
        for (uint32_t k = 0; k < kChunkRows; k += 32) {
            const uint32_t rr = r0 + k + 8u * static_cast<uint32_t>(q);
            const u16x8 a = ld8(&dl[static_cast<size_t>(16 * to + col) * n_cap + rr]);
            const u16x8 b = ld8(&al[static_cast<size_t>(16 * ti + col) * n_cap + rr]);
            acc = mfma<F16>(a, b, acc);
        }
        for (int k = 0; k < 4; ++k) {
            const int o = 16 * to + 4 * q + k, i = 16 * ti + col;
            dst[o * ni + i] = acc[k];
        }
    }
}

// grad[g][p] = inv_scale * (sum over the chunks of group g of part[chunk][p]).
template <int W>
__global__ void reduce_grad_kernel(const float* __restrict__ part, int chunks_per_group, uint32_t groups,
                                   float inv_scale, float* __restrict__ grad) {
    constexpr int NPAR = Geo<W>::NPAR;
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= groups * NPAR) return;
    const uint32_t g = p / NPAR, lp = p % NPAR;
    float s = 0.0f;
    for (int c = 0; c < chunks_per_group; ++c)
        s += part[(static_cast<size_t>(g) * chunks_per_group + c) * NPAR + lp];
    grad[p] = s * inv_scale;
}

// The bias-correction powers of the next Adam step, kept on the device so
// that captured training steps replay with the right step: bias[0..2] are
// the FP32 running products beta1^t, beta2^t and ema_decay^t (one factor
// per step since the parameters were set).
__global__ void bias_step_kernel(float* __restrict__ bias, float beta1, float beta2, float ema_decay) {
    if (threadIdx.x == 0) {
        bias[0] = bias[0] * beta1;
        bias[1] = bias[1] * beta2;
        bias[2] = bias[2] * ema_decay;
    }
}

// Adam (the oracle's or_nrc_adam; c1 = 1 - beta1^t, c2 = 1 - beta2^t from
// the device's running products), then the 16-bit copies W[o][i] and
// W^T[i][o] of the training weights, and the exponential moving average of
// the parameters (decay ema_decay, debiased by ema_c = 1 - decay^t) whose
// 16-bit copy answers inference. A gate that reads 0 (gate[g] with
// gate_per_group) leaves its group untouched.
template <int W, bool F16>
__global__ void adam_kernel(uint32_t groups, float* __restrict__ params, float* __restrict__ m,
                            float* __restrict__ v, const float* __restrict__ grad, float lr, float beta1, float beta2,
                            float eps, const float* __restrict__ bias, uint16_t* __restrict__ w,
                            uint16_t* __restrict__ wt, const uint32_t* __restrict__ gate, int gate_per_group,
                            float ema_decay, float* __restrict__ ema, uint16_t* __restrict__ w_inf) {
    using G = Geo<W>;
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= groups * G::NPAR) return;
    const float c1 = 1.0f - bias[0], c2 = 1.0f - bias[1], ema_c = 1.0f - bias[2];
    const uint32_t g = p / G::NPAR;
    if (gate && gate[gate_per_group ? g : 0u] == 0u) return;
    const float gr = grad[p];
    const float mm = beta1 * m[p] + (1.0f - beta1) * gr;
    const float vv = beta2 * v[p] + (1.0f - beta2) * (gr * gr);
    m[p] = mm;
    v[p] = vv;
    const float mh = mm / c1;
    const float vh = vv / c2;
    const float x = params[p] - lr * mh / (sqrtf(vh) + eps);
    params[p] = x;
    const uint16_t h = cvt<F16>(x);
    w[p] = h;
    int l, o, i;
    G::decode(static_cast<int>(p - g * G::NPAR), l, o, i);
    wt[static_cast<size_t>(g) * G::NPAR_T + G::tindex(l, o, i)] = h;
    const float e = ema_decay * ema[p] + (1.0f - ema_decay) * x;
    ema[p] = e;
    w_inf[p] = ema_decay > 0.0f ? cvt<F16>(e / ema_c) : h;
}

// Row j of step `step` of network g: sample (off + (step * batch + j) * P)
// mod N_g of the network's pool (pool_n[g] samples from pool_begin[g], or
// from 0 without pool_begin), P = 2^31 - 1 a prime above any pool size, so
// that the rows of the steps of a frame walk the pool without repetition
// while they last. Rows are laid out [network][batch]. With frame_dev, off
// is derived from the device's frame counter (a captured frame's draws).
__global__ __launch_bounds__(256) void gather_kernel(const float* __restrict__ pool,
                                                     const uint32_t* __restrict__ pool_begin,
                                                     const uint32_t* __restrict__ pool_n, uint32_t batch,
                                                     uint32_t groups, uint32_t step, uint64_t off,
                                                     const int32_t* __restrict__ frame_dev, uint64_t frame_seed,
                                                     float* __restrict__ out) {
    const uint32_t r = blockIdx.x * 256u + threadIdx.x;
    if (r >= batch * groups) return;
    if (frame_dev) off = splitmix64(nrc_frame_hash(frame_seed, *frame_dev) ^ kNrcTrainSalt);
    const uint32_t g = r / batch, j = r % batch;
    const uint32_t n = pool_n[g];
    float4* dst = reinterpret_cast<float4*>(out + static_cast<size_t>(r) * SREC);
    if (n == 0u) {
        for (int k = 0; k < SREC / 4; ++k) dst[k] = float4{0.0f, 0.0f, 0.0f, 0.0f};
        return;
    }
    const uint64_t idx =
        (off % n + (static_cast<uint64_t>(step) * batch + j) % n * 2147483647ull % n) % n;
    const size_t base = pool_begin ? pool_begin[g] : 0u;
    const float4* src = reinterpret_cast<const float4*>(pool + (base + idx) * SREC);
    for (int k = 0; k < SREC / 4; ++k) dst[k] = src[k];
}

}  // namespace

struct NrcState {
    int width = 64;
    bool f16 = false;
    uint32_t groups = 1;
    size_t npar = 0, npar_t = 0;   // per network
    float *params = nullptr, *m = nullptr, *v = nullptr, *grad = nullptr;
    uint16_t *w = nullptr, *wt = nullptr;
    float* ema = nullptr;          // moving average of the parameters (zero-started, debiased)
    float* bias = nullptr;         // device: beta1^t, beta2^t, ema_decay^t (running FP32 products)
    uint16_t* w_inf = nullptr;     // 16-bit weights inference reads: the debiased average, or w
    float ema_decay = 0.0f;
    Bounds bounds{{-1.0f, -1.0f, -1.0f}, {1.0f, 1.0f, 1.0f}};  // network 0's (all networks' with one)
    std::vector<Bounds> host_bounds;                            // per network
    Bounds* dev_bounds = nullptr;                               // per network, device copy
    uint32_t* map = nullptr;       // seg[groups + 1] and tile0[groups + 1] of ap_nrc_infer_grouped
    int step = 0;                  // Adam steps since the parameters were set (host count)
    uint32_t cap = 0;
    uint16_t *acts = nullptr, *deltas = nullptr, *dout = nullptr;
    float *y = nullptr, *part = nullptr, *loss_part = nullptr;
    float* batch = nullptr;   // gathered training rows of nrc_train_pool
    uint32_t batch_cap = 0;
    uint64_t generation = 0;  // a process-unique number, renewed when a training buffer is reallocated
};

namespace {

template <int W_, bool F16_>
struct Net {
    static constexpr int W = W_;
    static constexpr bool F16 = F16_;
};

// Calls fn with the Net<W, F16> tag of the state's network shape.
template <class Fn>
int with_net(const NrcState* s, Fn&& fn) {
    if (s->width == 32) return s->f16 ? fn(Net<32, true>{}) : fn(Net<32, false>{});
    return s->f16 ? fn(Net<64, true>{}) : fn(Net<64, false>{});
}

size_t npar_of(int width) { return static_cast<size_t>(width) * NF + 4u * width * width + NOP * width; }
size_t npar_t_of(int width) { return static_cast<size_t>(width) * NF + 4u * width * width + WT_OUT_STRIDE * width; }

// Power of two the loss gradient of an FP16 network is scaled by for a
// batch of `rows` per network: 2^(ceil(log2(3 rows)) - 4); 1 for BF16.
float loss_scale(const NrcState* s, uint32_t rows) {
    if (!s->f16) return 1.0f;
    int k = 0;
    while ((1ull << k) < 3ull * rows) ++k;
    return std::ldexp(1.0f, k - 4);
}

// Layer, output and input of parameter p of a width-w network (host).
void decode_h(int width, int p, int& l, int& o, int& i) {
    if (width == 32) Geo<32>::decode(p, l, o, i);
    else Geo<64>::decode(p, l, o, i);
}
int tindex_h(int width, int l, int o, int i) {
    return width == 32 ? Geo<32>::tindex(l, o, i) : Geo<64>::tindex(l, o, i);
}

int upload_bounds(NrcState* s) {
    const hipError_t e = hipMemcpy(s->dev_bounds, s->host_bounds.data(), s->groups * sizeof(Bounds),
                                   hipMemcpyHostToDevice);
    return e == hipSuccess ? AP_OK : hip_fail(e, "ap_nrc: bounds");
}

int upload_params(NrcState* s, const std::vector<float>& p) {
    const size_t np = s->npar, nt = s->npar_t, G = s->groups;
    std::vector<uint16_t> w(G * np), wt(G * nt, 0);
    for (size_t g = 0; g < G; ++g)
        for (size_t k = 0; k < np; ++k) {
            const uint16_t h = s->f16 ? f2h_h(p[g * np + k]) : f2bf_h(p[g * np + k]);
            w[g * np + k] = h;
            int l, o, i;
            decode_h(s->width, static_cast<int>(k), l, o, i);
            wt[g * nt + tindex_h(s->width, l, o, i)] = h;
        }
    hipError_t e = hipMemcpy(s->params, p.data(), G * np * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(s->w, w.data(), G * np * sizeof(uint16_t), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(s->w_inf, w.data(), G * np * sizeof(uint16_t), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(s->ema, 0, G * np * sizeof(float));
    const float ones[4] = {1.0f, 1.0f, 1.0f, 0.0f};
    if (e == hipSuccess) e = hipMemcpy(s->bias, ones, sizeof(ones), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(s->wt, wt.data(), G * nt * sizeof(uint16_t), hipMemcpyHostToDevice);
    return e == hipSuccess ? AP_OK : hip_fail(e, "ap_nrc: upload parameters");
}

// Process-unique numbers for NrcState::generation: a captured frame is keyed
// by it, and a new cache may reuse a destroyed one's address.
uint64_t next_generation() {
    static std::atomic<uint64_t> counter{0};
    return ++counter;
}

int ensure_train(NrcState* s, uint32_t n) {
    if (s->cap >= n) return AP_OK;
    s->generation = next_generation();
    void* ptrs[] = {s->acts, s->deltas, s->dout, s->y, s->part, s->loss_part};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    s->acts = s->deltas = s->dout = nullptr;
    s->y = s->part = s->loss_part = nullptr;
    s->cap = 0;
    const size_t N = n, Wd = static_cast<size_t>(s->width);
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&s->acts), (NF + NL * Wd) * N * sizeof(uint16_t));
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&s->deltas), NL * Wd * N * sizeof(uint16_t));
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&s->dout), NOP * N * sizeof(uint16_t));
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&s->y), N * NOUT * sizeof(float));
    if (e == hipSuccess)
        e = hipMalloc(reinterpret_cast<void**>(&s->part), (N / kChunkRows) * s->npar * sizeof(float));
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&s->loss_part), ((N + 63) / 64) * sizeof(float));
    if (e != hipSuccess) return hip_fail(e, "ap_nrc: training buffers");
    s->cap = n;
    return AP_OK;
}

inline unsigned blocks_for(uint32_t n) {
    const uint32_t b = (n + 63) / 64;
    return b < 2048u ? (b > 0 ? b : 1u) : 2048u;
}

// Inference over the rows of gm (groups == 0: rows [0, min(*dyn_n, n)) of
// network 0 with its bounds `s->bounds`).
int infer(NrcState* s, const float* queries, const GroupMap& gm, const int32_t* rows, const uint32_t* dyn_n,
          uint32_t n, float* out, hipStream_t st) {
    with_net(s, [&](auto net) {
        using N = decltype(net);
        hipLaunchKernelGGL((forward_kernel<false, N::W, N::F16>), dim3(blocks_for(n)), dim3(256), 0, st, s->w_inf,
                           s->bounds, gm, queries, QREC, n, dyn_n, rows, out, nullptr, 0u);
        return 0;
    });
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? AP_OK : hip_fail(e, "ap_nrc: infer");
}

}  // namespace

int nrc_infer(NrcState* s, const float* queries, size_t n, float* out, hipStream_t st) {
    if (n == 0) return AP_OK;
    return infer(s, queries, GroupMap{}, nullptr, nullptr, static_cast<uint32_t>(n), out, st);
}

// Forward, loss and backward of one batch of `rows` rows per network, laid
// out [network][row]; the gradient lands in s->grad. gate: see loss_kernel.
int nrc_gradient(NrcState* s, const float* samples, uint32_t rows, hipStream_t st, const uint32_t* gate = nullptr,
                 int gate_per_group = 0) {
    const uint32_t n = rows * s->groups;
    int rc = ensure_train(s, n);
    if (rc != AP_OK) return rc;
    GroupMap gm;
    gm.groups = s->groups;
    gm.rows_per_group = rows;
    gm.bounds = s->groups > 1 ? s->dev_bounds : nullptr;
    const float scale = loss_scale(s, rows);
    with_net(s, [&](auto net) {
        using N = decltype(net);
        constexpr int NPAR = Geo<N::W>::NPAR;
        hipLaunchKernelGGL((fwd_bwd_kernel<N::W, N::F16>), dim3(blocks_for(n)), dim3(256), 0, st, s->w, s->wt,
                           s->bounds, gm.bounds, samples, n, rows, gate, gate_per_group, scale, s->acts, s->dout,
                           s->deltas, s->loss_part);
        hipLaunchKernelGGL((wgrad_kernel<N::W, N::F16>), dim3(n / kChunkRows, NL + 1), dim3(256), 0, st, n, s->acts,
                           s->deltas, s->dout, s->part);
        hipLaunchKernelGGL(reduce_grad_kernel<N::W>, dim3((s->groups * NPAR + 255) / 256), dim3(256), 0, st,
                           s->part, static_cast<int>(rows / kChunkRows), s->groups, 1.0f / scale, s->grad);
        return 0;
    });
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? AP_OK : hip_fail(e, "ap_nrc: gradient");
}

constexpr float kBeta1 = 0.9f, kBeta2 = 0.99f, kAdamEps = 1e-8f;

int nrc_adam(NrcState* s, float lr, hipStream_t st, const uint32_t* gate = nullptr, int gate_per_group = 0) {
    s->step += 1;
    hipLaunchKernelGGL(bias_step_kernel, dim3(1), dim3(64), 0, st, s->bias, kBeta1, kBeta2, s->ema_decay);
    with_net(s, [&](auto net) {
        using N = decltype(net);
        constexpr int NPAR = Geo<N::W>::NPAR;
        hipLaunchKernelGGL((adam_kernel<N::W, N::F16>), dim3((s->groups * NPAR + 255) / 256), dim3(256), 0, st,
                           s->groups, s->params, s->m, s->v, s->grad, lr, kBeta1, kBeta2, kAdamEps, s->bias, s->w,
                           s->wt, gate, gate_per_group, s->ema_decay, s->ema, s->w_inf);
        return 0;
    });
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? AP_OK : hip_fail(e, "ap_nrc: adam");
}

int nrc_infer_dyn(NrcState* s, const float* queries, const int32_t* rows, const uint32_t* dyn_n, uint32_t n_cap,
                  float* out, hipStream_t st) {
    if (n_cap == 0) return AP_OK;
    return infer(s, queries, GroupMap{}, rows, dyn_n, n_cap, out, st);
}

int nrc_train_pool(NrcState* s, const float* pool, const uint32_t* pool_begin, const uint32_t* pool_n,
                   uint32_t batch, uint32_t steps, float lr, uint64_t seed, hipStream_t st, const int32_t* frame_dev,
                   uint64_t frame_seed) {
    if (steps == 0) return AP_OK;
    if (batch == 0 || batch % kChunkRows != 0 || batch * static_cast<uint64_t>(s->groups) > (1u << 24))
        return fail(AP_ERR_INVALID_ARG, "ap_nrc: batch must be a positive multiple of 256, at most 2^24 rows in all");
    const uint32_t rows = batch * s->groups;
    if (s->batch_cap < rows) {
        s->generation = next_generation();
        if (s->batch) (void)hipFree(s->batch);
        s->batch = nullptr;
        s->batch_cap = 0;
        const hipError_t e = hipMalloc(reinterpret_cast<void**>(&s->batch), static_cast<size_t>(rows) * SREC * 4);
        if (e != hipSuccess) return hip_fail(e, "ap_nrc: batch buffer");
        s->batch_cap = rows;
    }
    const uint64_t off = splitmix64(seed);
    const int per_group = s->groups > 1 ? 1 : 0;
    for (uint32_t k = 0; k < steps; ++k) {
        hipLaunchKernelGGL(gather_kernel, dim3((rows + 255) / 256), dim3(256), 0, st, pool, pool_begin, pool_n,
                           batch, s->groups, k, off, frame_dev, frame_seed, s->batch);
        int rc = nrc_gradient(s, s->batch, batch, st, pool_n, per_group);
        if (rc == AP_OK) rc = nrc_adam(s, lr, st, pool_n, per_group);
        if (rc != AP_OK) return rc;
    }
    return AP_OK;
}

int nrc_infer_groups(NrcState* s, const float* queries, const int32_t* rows, const uint32_t* seg,
                     const uint32_t* tile0, const uint32_t* count, uint32_t n_bound, float* out, hipStream_t st) {
    if (n_bound == 0) return AP_OK;
    GroupMap gm;
    gm.groups = s->groups;
    gm.seg = seg;
    gm.tile0 = tile0;
    gm.count = count;
    gm.bounds = s->dev_bounds;
    return infer(s, queries, gm, rows, nullptr, n_bound, out, st);
}

int nrc_set_group_bounds(NrcState* s, uint32_t g, const float lo[3], const float hi[3]) {
    Bounds b;
    for (int a = 0; a < 3; ++a) {
        b.lo[a] = lo[a];
        b.hi[a] = hi[a];
    }
    if (g == 0) s->bounds = b;
    if (std::memcmp(&s->host_bounds[g], &b, sizeof(Bounds)) == 0) return AP_OK;
    s->host_bounds[g] = b;
    return upload_bounds(s);
}

void nrc_count_steps(NrcState* s, int32_t steps) { s->step += steps; }
uint64_t nrc_generation(const NrcState* s) { return s->generation; }

void nrc_set_bounds(NrcState* s, const float lo[3], const float hi[3]) {
    for (int a = 0; a < 3; ++a) {
        s->bounds.lo[a] = lo[a];
        s->bounds.hi[a] = hi[a];
    }
}

void nrc_free(NrcState* s) {
    if (!s) return;
    void* ptrs[] = {s->params, s->m, s->v, s->grad, s->w, s->wt, s->acts, s->deltas, s->dout, s->y, s->part,
                    s->loss_part, s->batch, s->ema, s->w_inf, s->dev_bounds, s->map, s->bias};
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
uint32_t ap::nrc_groups(const NrcState* s) { return s->groups; }

namespace {

int check_batch(const char* who, size_t n) {
    if (n == 0 || n % kChunkRows != 0 || n > (1u << 24))
        return fail(AP_ERR_INVALID_ARG, std::string(who) + ": batch must be a positive multiple of 256 up to 2^24");
    return AP_OK;
}

// Copies the loss partials (one per 64-row tile) of the last gradient into
// per-network losses.
hipError_t read_losses(const NrcState* s, uint32_t rows, float* loss) {
    const uint32_t per = rows / 64, n = per * s->groups;
    std::vector<float> lp(n);
    const hipError_t e = hipMemcpy(lp.data(), s->loss_part, n * sizeof(float), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return e;
    for (uint32_t g = 0; g < s->groups; ++g) {
        float sum = 0.0f;
        for (uint32_t k = 0; k < per; ++k) sum += lp[g * per + k];
        loss[g] = sum;
    }
    return hipSuccess;
}

}  // namespace

extern "C" {

int ap_nrc_create_grouped(uint32_t groups, uint32_t width, int32_t precision, uint64_t seed, ap_nrc** out) {
    if (!out) return fail(AP_ERR_INVALID_ARG, "ap_nrc_create: NULL output");
    *out = nullptr;
    if (groups == 0 || groups > kMaxGroups)
        return fail(AP_ERR_INVALID_ARG, "ap_nrc_create_grouped: groups must be in [1, 4096]");
    if (width != 32 && width != 64) return fail(AP_ERR_INVALID_ARG, "ap_nrc_create_grouped: width must be 32 or 64");
    if (precision != AP_NRC_BF16 && precision != AP_NRC_FP16)
        return fail(AP_ERR_INVALID_ARG, "ap_nrc_create_grouped: precision must be AP_NRC_BF16 or AP_NRC_FP16");
    NrcState* s = new NrcState();
    s->generation = next_generation();
    s->width = static_cast<int>(width);
    s->f16 = precision == AP_NRC_FP16;
    s->groups = groups;
    s->npar = npar_of(s->width);
    s->npar_t = npar_t_of(s->width);
    s->host_bounds.assign(groups, s->bounds);
    const size_t np = groups * s->npar;
    hipError_t e = hipSuccess;
    for (float** p : {&s->params, &s->m, &s->v, &s->grad, &s->ema})
        if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(p), np * sizeof(float));
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&s->w), np * sizeof(uint16_t));
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&s->w_inf), np * sizeof(uint16_t));
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&s->wt), groups * s->npar_t * sizeof(uint16_t));
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&s->dev_bounds), groups * sizeof(Bounds));
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&s->map), 2 * (groups + 1) * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&s->bias), 4 * sizeof(float));
    if (e == hipSuccess) e = hipMemset(s->m, 0, np * sizeof(float));
    if (e == hipSuccess) e = hipMemset(s->v, 0, np * sizeof(float));
    if (e == hipSuccess) e = hipMemset(s->grad, 0, np * sizeof(float));
    if (e != hipSuccess) {
        nrc_free(s);
        return hip_fail(e, "ap_nrc_create");
    }
    // He-normal initial weights (sd sqrt(2 / inputs) for the hidden layers,
    // sqrt(1 / width) for the output rows 0..2, rows 3..15 zero) from
    // PCG32(seed, 7 + network) by Box-Muller.
    std::vector<float> p(np, 0.0f);
    for (uint32_t g = 0; g < groups; ++g) {
        Pcg32 rng = pcg32_seed(seed, 7 + g);
        auto normal = [&]() {
            const float u1 = 1.0f - pcg32_float(rng), u2 = pcg32_float(rng);
            return std::sqrt(-2.0f * std::log(u1)) * std::cos(6.283185307f * u2);
        };
        float* q = p.data() + g * s->npar;
        const size_t hidden = s->npar - NOP * width;
        for (size_t k = 0; k < hidden; ++k) q[k] = normal() * std::sqrt(2.0f / (k < width * NF ? NF : width));
        for (uint32_t o = 0; o < NOUT; ++o)
            for (uint32_t i = 0; i < width; ++i) q[hidden + o * width + i] = normal() * std::sqrt(1.0f / width);
    }
    int rc = upload_params(s, p);
    if (rc == AP_OK) rc = upload_bounds(s);
    if (rc != AP_OK) {
        nrc_free(s);
        return rc;
    }
    *out = new ap_nrc{s};
    return AP_OK;
}

int ap_nrc_create(uint64_t seed, ap_nrc** out) { return ap_nrc_create_grouped(1, 64, AP_NRC_BF16, seed, out); }

int ap_nrc_destroy(ap_nrc* nrc) {
    if (!nrc) return AP_OK;
    nrc_free(nrc->s);
    delete nrc;
    return AP_OK;
}

int ap_nrc_info(const ap_nrc* nrc, uint32_t* groups, uint32_t* width, int32_t* precision,
                size_t* params_per_network) {
    if (!nrc) return fail(AP_ERR_INVALID_ARG, "ap_nrc_info: NULL cache");
    if (groups) *groups = nrc->s->groups;
    if (width) *width = static_cast<uint32_t>(nrc->s->width);
    if (precision) *precision = nrc->s->f16 ? AP_NRC_FP16 : AP_NRC_BF16;
    if (params_per_network) *params_per_network = nrc->s->npar;
    return AP_OK;
}

int ap_nrc_set_bounds(ap_nrc* nrc, const float bmin[3], const float bmax[3]) {
    if (!nrc || !bmin || !bmax) return fail(AP_ERR_INVALID_ARG, "ap_nrc_set_bounds: NULL argument");
    for (int a = 0; a < 3; ++a)
        if (!(bmax[a] > bmin[a])) return fail(AP_ERR_INVALID_ARG, "ap_nrc_set_bounds: empty bounds");
    nrc_set_bounds(nrc->s, bmin, bmax);
    for (Bounds& b : nrc->s->host_bounds) b = nrc->s->bounds;
    return upload_bounds(nrc->s);
}

int ap_nrc_set_group_bounds(ap_nrc* nrc, uint32_t group, const float bmin[3], const float bmax[3]) {
    if (!nrc || !bmin || !bmax) return fail(AP_ERR_INVALID_ARG, "ap_nrc_set_group_bounds: NULL argument");
    if (group >= nrc->s->groups) return fail(AP_ERR_INVALID_ARG, "ap_nrc_set_group_bounds: no such network");
    Bounds b;
    for (int a = 0; a < 3; ++a) {
        if (!(bmax[a] > bmin[a])) return fail(AP_ERR_INVALID_ARG, "ap_nrc_set_group_bounds: empty bounds");
        b.lo[a] = bmin[a];
        b.hi[a] = bmax[a];
    }
    nrc->s->host_bounds[group] = b;
    if (group == 0) nrc->s->bounds = b;
    return upload_bounds(nrc->s);
}

int ap_nrc_get_params(const ap_nrc* nrc, float* host_params) {
    if (!nrc || !host_params) return fail(AP_ERR_INVALID_ARG, "ap_nrc_get_params: NULL argument");
    const hipError_t e = hipMemcpy(host_params, nrc->s->params, nrc->s->groups * nrc->s->npar * sizeof(float),
                                   hipMemcpyDeviceToHost);
    return e == hipSuccess ? AP_OK : hip_fail(e, "ap_nrc_get_params");
}

int ap_nrc_set_params(ap_nrc* nrc, const float* host_params) {
    if (!nrc || !host_params) return fail(AP_ERR_INVALID_ARG, "ap_nrc_set_params: NULL argument");
    NrcState* s = nrc->s;
    const size_t np = s->groups * s->npar, out0 = s->npar - NOP * s->width;
    std::vector<float> p(host_params, host_params + np);
    for (uint32_t g = 0; g < s->groups; ++g)
        for (size_t k = out0 + NOUT * s->width; k < s->npar; ++k) p[g * s->npar + k] = 0.0f;
    s->step = 0;
    hipError_t e = hipMemset(s->m, 0, np * sizeof(float));
    if (e == hipSuccess) e = hipMemset(s->v, 0, np * sizeof(float));
    if (e != hipSuccess) return hip_fail(e, "ap_nrc_set_params");
    return upload_params(s, p);
}

int ap_nrc_infer(ap_nrc* nrc, const float* queries, size_t n, float* out_rgb, void* stream) {
    if (!nrc) return fail(AP_ERR_INVALID_ARG, "ap_nrc_infer: NULL cache");
    if (nrc->s->groups != 1) return fail(AP_ERR_INVALID_ARG, "ap_nrc_infer: a grouped cache takes ap_nrc_infer_grouped");
    if (n == 0) return AP_OK;
    if (!queries || !out_rgb) return fail(AP_ERR_INVALID_ARG, "ap_nrc_infer: NULL buffer");
    if (n > 0x7fffffffull) return fail(AP_ERR_INVALID_ARG, "ap_nrc_infer: too many queries");
    return nrc_infer(nrc->s, queries, n, out_rgb, static_cast<hipStream_t>(stream));
}

int ap_nrc_infer_grouped(ap_nrc* nrc, const float* queries, const uint32_t* host_seg, float* out_rgb, void* stream) {
    if (!nrc || !host_seg) return fail(AP_ERR_INVALID_ARG, "ap_nrc_infer_grouped: NULL argument");
    NrcState* s = nrc->s;
    const uint32_t G = s->groups;
    std::vector<uint32_t> map(2 * (G + 1));
    uint64_t tiles = 0;
    for (uint32_t g = 0; g < G; ++g) {
        if (host_seg[g + 1] < host_seg[g])
            return fail(AP_ERR_INVALID_ARG, "ap_nrc_infer_grouped: row offsets must not decrease");
        map[g] = host_seg[g];
        map[G + 1 + g] = static_cast<uint32_t>(tiles);
        tiles += (host_seg[g + 1] - host_seg[g] + 63u) / 64u;
    }
    map[G] = host_seg[G];
    map[2 * G + 1] = static_cast<uint32_t>(tiles);
    if (host_seg[G] > 0x7fffffffu) return fail(AP_ERR_INVALID_ARG, "ap_nrc_infer_grouped: too many queries");
    if (tiles == 0) return AP_OK;
    if (!queries || !out_rgb) return fail(AP_ERR_INVALID_ARG, "ap_nrc_infer_grouped: NULL buffer");
    const hipStream_t st = static_cast<hipStream_t>(stream);
    // The map is rewritten only once the stream's earlier work has read it.
    hipError_t e = hipStreamSynchronize(st);
    if (e == hipSuccess) e = hipMemcpy(s->map, map.data(), map.size() * sizeof(uint32_t), hipMemcpyHostToDevice);
    if (e != hipSuccess) return hip_fail(e, "ap_nrc_infer_grouped");
    GroupMap gm;
    gm.groups = G;
    gm.seg = s->map;
    gm.tile0 = s->map + G + 1;
    gm.bounds = s->dev_bounds;
    return infer(s, queries, gm, nullptr, nullptr, static_cast<uint32_t>(tiles * 64u), out_rgb, st);
}

int ap_nrc_gradient_grouped(ap_nrc* nrc, const float* samples, size_t batch, float* host_grad, float* host_loss,
                            void* stream) {
    if (!nrc || !samples) return fail(AP_ERR_INVALID_ARG, "ap_nrc_gradient: NULL argument");
    int rc = check_batch("ap_nrc_gradient", batch);
    if (rc != AP_OK) return rc;
    if (batch * nrc->s->groups > (1u << 24))
        return fail(AP_ERR_INVALID_ARG, "ap_nrc_gradient: at most 2^24 rows over all networks");
    const hipStream_t st = static_cast<hipStream_t>(stream);
    rc = nrc_gradient(nrc->s, samples, static_cast<uint32_t>(batch), st);
    if (rc != AP_OK) return rc;
    hipError_t e = hipStreamSynchronize(st);
    if (e == hipSuccess && host_grad)
        e = hipMemcpy(host_grad, nrc->s->grad, nrc->s->groups * nrc->s->npar * sizeof(float), hipMemcpyDeviceToHost);
    if (e == hipSuccess && host_loss) e = read_losses(nrc->s, static_cast<uint32_t>(batch), host_loss);
    return e == hipSuccess ? AP_OK : hip_fail(e, "ap_nrc_gradient");
}

int ap_nrc_gradient(ap_nrc* nrc, const float* samples, size_t n, float* host_grad, float* loss, void* stream) {
    if (nrc && nrc->s->groups != 1)
        return fail(AP_ERR_INVALID_ARG, "ap_nrc_gradient: a grouped cache takes ap_nrc_gradient_grouped");
    return ap_nrc_gradient_grouped(nrc, samples, n, host_grad, loss, stream);
}

int ap_nrc_train_step_grouped(ap_nrc* nrc, const float* samples, size_t batch, float lr, void* stream) {
    if (!nrc || !samples) return fail(AP_ERR_INVALID_ARG, "ap_nrc_train_step: NULL argument");
    int rc = check_batch("ap_nrc_train_step", batch);
    if (rc != AP_OK) return rc;
    if (batch * nrc->s->groups > (1u << 24))
        return fail(AP_ERR_INVALID_ARG, "ap_nrc_train_step: at most 2^24 rows over all networks");
    const hipStream_t st = static_cast<hipStream_t>(stream);
    rc = nrc_gradient(nrc->s, samples, static_cast<uint32_t>(batch), st);
    if (rc != AP_OK) return rc;
    return nrc_adam(nrc->s, lr, st);
}

int ap_nrc_train_step(ap_nrc* nrc, const float* samples, size_t n, float lr, void* stream) {
    if (nrc && nrc->s->groups != 1)
        return fail(AP_ERR_INVALID_ARG, "ap_nrc_train_step: a grouped cache takes ap_nrc_train_step_grouped");
    return ap_nrc_train_step_grouped(nrc, samples, n, lr, stream);
}

int ap_nrc_set_ema(ap_nrc* nrc, float decay) {
    if (!nrc) return fail(AP_ERR_INVALID_ARG, "ap_nrc_set_ema: NULL cache");
    if (!(decay >= 0.0f && decay < 1.0f)) return fail(AP_ERR_INVALID_ARG, "ap_nrc_set_ema: decay must be in [0, 1)");
    NrcState* s = nrc->s;
    s->ema_decay = decay;
    // The debiasing power of the new decay after the steps taken so far.
    float pw = 1.0f;
    for (int t = 0; t < s->step && pw != 0.0f; ++t) pw = pw * decay;
    hipError_t e = hipDeviceSynchronize();  // no training step may be in flight
    if (e == hipSuccess) e = hipMemcpy(s->bias + 2, &pw, sizeof(float), hipMemcpyHostToDevice);
    return e == hipSuccess ? AP_OK : hip_fail(e, "ap_nrc_set_ema");
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
    const hipError_t e = hipMemcpy(host_grad, nrc->s->grad, nrc->s->groups * nrc->s->npar * sizeof(float),
                                   hipMemcpyDeviceToHost);
    return e == hipSuccess ? AP_OK : hip_fail(e, "ap_nrc_last_gradient");
}

}  // extern "C"

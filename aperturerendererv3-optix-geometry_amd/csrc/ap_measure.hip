// FP64 image statistics (include/aperture/ap_measure.h): the odd/even
// luminance MSE that replaces DeviceMSELuminance (core/helper/measure.cuh:
// 49-73), DeviceMSE (measure.cuh:18-33) and the relMSE of an image against a
// reference.
//
// Two passes over a fixed grid. Pass 1: block b sums, in FP64, the terms of
// elements b*256 + t, + grid*256, ... (grid-stride, one partial per lane, a
// wave64 shuffle tree, then the block's four waves in order) into
// partials[b]. Pass 2: one 1024-lane block adds the partials in a fixed
// tree and divides by the term count. No atomics, so a call returns the same
// bits every time. HBM-bound: 24 B per pixel for two float3 images (48 B for
// FP64 sums), read once.
#include <hip/hip_runtime.h>

#include <cmath>
#include <limits>
#include <mutex>

#include "aperture/ap_measure.h"
#include "ap_internal.h"

namespace ap {
namespace {

constexpr int kReduceThreads = 256;
constexpr int kMaxDevices = 64;
static_assert(AP_REDUCE_BLOCKS == 1024, "the final pass is one 1024-lane block");

struct MseTerm {
    const float* a;
    const float* b;
    __device__ double operator()(size_t i) const {
        const float d = a[i] - b[i];
        return static_cast<double>(d) * static_cast<double>(d);
    }
};

struct LumTerm {
    const float* a;
    const float* b;
    __device__ double operator()(size_t i) const {
        const float la = 0.2126f * a[3 * i] + 0.7152f * a[3 * i + 1] + 0.0722f * a[3 * i + 2];
        const float lb = 0.2126f * b[3 * i] + 0.7152f * b[3 * i + 1] + 0.0722f * b[3 * i + 2];
        const double d = static_cast<double>(la) - static_cast<double>(lb);
        return d * d;
    }
};

struct ImageF64 {
    const void* p;
    int32_t kind;
    double scale;
    __device__ void load(size_t i, double v[3]) const {
        if (kind == AP_IMAGE_XYZ_F32) {
            const float* f = static_cast<const float*>(p) + 3 * i;
            for (int c = 0; c < 3; ++c) v[c] = static_cast<double>(f[c]) * scale;
        } else {
            const double* d = static_cast<const double*>(p) + 3 * i;
            for (int c = 0; c < 3; ++c) v[c] = d[c] * scale;
        }
    }
};

// XYZ2SRGBLinearD65 (core/tonemap.cuh:116-126) in FP64, on the FP32 matrix
// entries the reference stores.
__device__ inline void xyz_to_srgb_f64(const double x[3], double r[3]) {
    const double M[3][3] = {{3.2406f, -1.5372f, -0.4986f}, {-0.9689f, 1.8758f, 0.0415f}, {0.0557f, -0.204f, 1.057f}};
    for (int k = 0; k < 3; ++k) {
        const double v = M[k][0] * x[0] + M[k][1] * x[1] + M[k][2] * x[2];
        r[k] = v > 0.0 ? v : 0.0;
    }
}

struct RelTerm {
    ImageF64 img, ref;
    double eps;
    __device__ double operator()(size_t i) const {
        double a[3], b[3], ra[3], rb[3];
        img.load(i, a);
        ref.load(i, b);
        xyz_to_srgb_f64(a, ra);
        xyz_to_srgb_f64(b, rb);
        double s = 0.0;
        for (int c = 0; c < 3; ++c) {
            const double d = ra[c] - rb[c];
            s += d * d / (rb[c] * rb[c] + eps);
        }
        return s;
    }
};

__device__ inline double wave_sum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    return v;
}

template <class Term>
__global__ __launch_bounds__(kReduceThreads) void reduce_partials_kernel(Term term, size_t n, double* partials) {
    double acc = 0.0;
    const size_t stride = static_cast<size_t>(gridDim.x) * kReduceThreads;
    for (size_t i = static_cast<size_t>(blockIdx.x) * kReduceThreads + threadIdx.x; i < n; i += stride)
        acc += term(i);
    acc = wave_sum(acc);
    __shared__ double wsum[kReduceThreads / 64];
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int w = 0; w < kReduceThreads / 64; ++w) s += wsum[w];
        partials[blockIdx.x] = s;
    }
}

__global__ __launch_bounds__(AP_REDUCE_BLOCKS) void reduce_final_kernel(const double* partials, int m, double count,
                                                                        double* out) {
    double v = static_cast<int>(threadIdx.x) < m ? partials[threadIdx.x] : 0.0;
    v = wave_sum(v);
    __shared__ double wsum[AP_REDUCE_BLOCKS / 64];
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int w = 0; w < AP_REDUCE_BLOCKS / 64; ++w) s += wsum[w];
        out[0] = count > 0.0 ? s / count : __builtin_nan("");
    }
}

__global__ void write_nan_kernel(double* out) {
    if (threadIdx.x == 0) out[0] = __builtin_nan("");
}

// Both passes on `stream`; partials holds AP_REDUCE_BLOCKS doubles. Every
// partial the final pass reads is written by pass 1 of the same call, so the
// scratch needs no clearing.
template <class Term>
int launch_reduce(const Term& term, size_t n, double terms_per_element, double* partials, double* out,
                  hipStream_t stream, const char* what) {
    if (n == 0) {
        hipLaunchKernelGGL(write_nan_kernel, dim3(1), dim3(64), 0, stream, out);
    } else {
        const size_t want = (n + kReduceThreads - 1) / kReduceThreads;
        const int blocks = static_cast<int>(want < AP_REDUCE_BLOCKS ? want : AP_REDUCE_BLOCKS);
        hipLaunchKernelGGL(reduce_partials_kernel<Term>, dim3(blocks), dim3(kReduceThreads), 0, stream, term, n,
                           partials);
        hipLaunchKernelGGL(reduce_final_kernel, dim3(1), dim3(AP_REDUCE_BLOCKS), 0, stream, partials, blocks,
                           static_cast<double>(n) * terms_per_element, out);
    }
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? AP_OK : hip_fail(e, what);
}

// The partials buffer of each device, allocated by the library on the first
// reduction on that device and kept until the process ends.
std::mutex g_partials_mutex;
double* g_partials[kMaxDevices] = {};

int device_partials(double** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
    if (dev < 0 || dev >= kMaxDevices) return fail(AP_ERR_UNSUPPORTED, "reduction: device index beyond 63");
    std::lock_guard<std::mutex> lock(g_partials_mutex);
    if (!g_partials[dev]) {
        e = hipMalloc(reinterpret_cast<void**>(&g_partials[dev]), AP_REDUCE_BLOCKS * sizeof(double));
        if (e != hipSuccess) {
            g_partials[dev] = nullptr;
            return hip_fail(e, "hipMalloc(reduction partials)");
        }
    }
    *out = g_partials[dev];
    return AP_OK;
}

// The partials of ap_relmse_f64_noalloc: a file-scope device array, one copy
// per GPU at its own address. Host code never passes `scratch` itself to a
// kernel (on the host the name is a stand-in, not a device address); it uses
// the address hipGetSymbolAddress returns with the device current, taken on
// the first call on each device and kept.
__device__ __align__(256) double scratch[AP_REDUCE_BLOCKS];
std::mutex g_scratch_mutex;
double* g_scratch[kMaxDevices] = {};

int device_scratch(double** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
    if (dev < 0 || dev >= kMaxDevices) return fail(AP_ERR_UNSUPPORTED, "reduction: device index beyond 63");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) return hip_fail(e, "hipGetSymbolAddress(reduction scratch)");
        g_scratch[dev] = static_cast<double*>(p);
    }
    *out = g_scratch[dev];
    return AP_OK;
}

int check_image(const ap_image_view* v, const char* what) {
    if (!v || !v->data) return fail(AP_ERR_INVALID_ARG, std::string(what) + ": NULL image");
    if (v->kind != AP_IMAGE_XYZ_F32 && v->kind != AP_IMAGE_XYZ_SUMS_F64)
        return fail(AP_ERR_INVALID_ARG, std::string(what) + ": unknown image kind");
    if (!std::isfinite(v->scale)) return fail(AP_ERR_INVALID_ARG, std::string(what) + ": scale must be finite");
    return AP_OK;
}

}  // namespace
}  // namespace ap

using namespace ap;

extern "C" {

int ap_mse_f64(const float* a, const float* b, size_t n, double* out, void* stream) {
    if (!out || (n > 0 && (!a || !b))) return fail(AP_ERR_INVALID_ARG, "ap_mse_f64: NULL pointer");
    if (n > AP_REDUCE_MAX_ELEMENTS) return fail(AP_ERR_INVALID_ARG, "ap_mse_f64: more than 2^34 elements");
    double* partials = nullptr;
    const int rc = device_partials(&partials);
    if (rc != AP_OK) return rc;
    return launch_reduce(MseTerm{a, b}, n, 1.0, partials, out, static_cast<hipStream_t>(stream), "ap_mse_f64");
}

int ap_mse_luminance_f64(const float* a, const float* b, size_t n_pixels, double* out, void* stream) {
    if (!out || (n_pixels > 0 && (!a || !b))) return fail(AP_ERR_INVALID_ARG, "ap_mse_luminance_f64: NULL pointer");
    if (n_pixels > AP_REDUCE_MAX_ELEMENTS)
        return fail(AP_ERR_INVALID_ARG, "ap_mse_luminance_f64: more than 2^34 pixels");
    double* partials = nullptr;
    const int rc = device_partials(&partials);
    if (rc != AP_OK) return rc;
    return launch_reduce(LumTerm{a, b}, n_pixels, 1.0, partials, out, static_cast<hipStream_t>(stream),
                         "ap_mse_luminance_f64");
}

namespace {

int relmse(const ap_image_view* image, const ap_image_view* reference, size_t n_pixels, double eps, double* out,
           void* stream, bool no_alloc, const char* what) {
    if (!out) return fail(AP_ERR_INVALID_ARG, std::string(what) + ": NULL out");
    if (n_pixels > AP_REDUCE_MAX_ELEMENTS)
        return fail(AP_ERR_INVALID_ARG, std::string(what) + ": more than 2^34 pixels");
    if (!(eps > 0.0) || !std::isfinite(eps)) return fail(AP_ERR_INVALID_ARG, std::string(what) + ": eps must be > 0");
    int rc;
    if (n_pixels > 0 && ((rc = check_image(image, what)) != AP_OK || (rc = check_image(reference, what)) != AP_OK))
        return rc;
    double* partials = nullptr;
    if ((rc = no_alloc ? device_scratch(&partials) : device_partials(&partials)) != AP_OK) return rc;
    const ImageF64 a = n_pixels ? ImageF64{image->data, image->kind, image->scale} : ImageF64{};
    const ImageF64 b = n_pixels ? ImageF64{reference->data, reference->kind, reference->scale} : ImageF64{};
    return launch_reduce(RelTerm{a, b, eps}, n_pixels, 3.0, partials, out, static_cast<hipStream_t>(stream), what);
}

}  // namespace

int ap_relmse_f64(const ap_image_view* image, const ap_image_view* reference, size_t n_pixels, double eps,
                  double* out, void* stream) {
    return relmse(image, reference, n_pixels, eps, out, stream, false, "ap_relmse_f64");
}

int ap_relmse_f64_noalloc(const ap_image_view* image, const ap_image_view* reference, size_t n_pixels, double eps,
                          double* out, void* stream) {
    return relmse(image, reference, n_pixels, eps, out, stream, true, "ap_relmse_f64_noalloc");
}

}  // extern "C"

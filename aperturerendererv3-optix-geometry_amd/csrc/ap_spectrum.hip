// Spectral LUT build and batched spectral entry points (include/aperture/ap_spectrum.h).
//
// The RGB LUT is 512^3 cells of float4 (2 GiB), one MLP evaluation per cell
// at the cell centre, laid out x (red) fastest, then y (green), then z (blue)
// exactly as PrecomputeRGBLUTKernel writes it (core/spectrum/spectrum_lut.cuh:12-37).
// The MLP weights live in __constant__ memory so every lane reads them
// through the scalar cache; one lane evaluates one cell.
#include <mutex>

#include "aperture/ap_spectrum.h"
#include "ap_internal.h"
#include "ap_spectrum.h"

extern "C" const unsigned char ap_spectrum_mlp_bin[];
extern "C" const unsigned char ap_cie_tables_bin[];

namespace ap {
namespace {

__constant__ float c_mlp[kMlpFloats];

struct SpectrumDevice {
    bool ready = false;
    f4* rgb_lut = nullptr;     // 512^3
    f4* lambda_lut = nullptr;  // 840
    float* tables = nullptr;   // D65[421] + CMF[421*3]
};

constexpr int kMaxDevices = 64;
SpectrumDevice g_spec[kMaxDevices];
std::mutex g_spec_mutex;

__global__ __launch_bounds__(256) void build_rgb_lut_kernel(f4* __restrict__ lut, uint32_t z0, uint32_t nz) {
    constexpr int R = kRgbLutRes;
    const uint32_t t = blockIdx.x * 256u + threadIdx.x;
    const uint32_t x = t % R;
    const uint32_t y = (t / R) % R;
    const uint32_t z = z0 + t / (R * R);
    if (z >= z0 + nz) return;
    constexpr float inv = 1.0f / static_cast<float>(R);
    const float r = (static_cast<float>(x) + 0.5f) * inv;
    const float g = (static_cast<float>(y) + 0.5f) * inv;
    const float b = (static_cast<float>(z) + 0.5f) * inv;
    lut[(static_cast<size_t>(z) * R + y) * R + x] = rgb_encode(c_mlp, r, g, b);
}

__global__ void build_lambda_lut_kernel(f4* __restrict__ lut) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= kLambdaLutRes) return;
    const float step = (780.0f - 360.0f) / static_cast<float>(kLambdaLutRes);
    lut[i] = lambda_encode(c_mlp, 360.0f + (static_cast<float>(i) + 0.5f) * step);
}

__global__ void encode_rgb_kernel(const float* __restrict__ in, size_t n, f4* __restrict__ out) {
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = rgb_encode(c_mlp, in[3 * i], in[3 * i + 1], in[3 * i + 2]);
}

__global__ void encode_lambda_kernel(const float* __restrict__ in, size_t n, f4* __restrict__ out) {
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = lambda_encode(c_mlp, in[i]);
}

__global__ void query_rgb_kernel(const f4* __restrict__ lut, const float* __restrict__ in, size_t n,
                                 f4* __restrict__ out) {
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = query_rgb_basis(lut, f3{in[3 * i], in[3 * i + 1], in[3 * i + 2]});
}

__global__ void query_lambda_kernel(const f4* __restrict__ lut, const float* __restrict__ in, size_t n,
                                    f4* __restrict__ out) {
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = query_lambda(lut, in[i]);
}

__global__ void interact_kernel(const f4* __restrict__ rgb_lut, const f4* __restrict__ lam_lut,
                                const float* __restrict__ lam, const float* __restrict__ rgb, size_t n,
                                float* __restrict__ out) {
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const f4 lb = query_lambda(lam_lut, lam[i]);
    out[i] = interact(rgb_lut, lb, f3{rgb[3 * i], rgb[3 * i + 1], rgb[3 * i + 2]});
}

__global__ void sample_lambda_kernel(const float* __restrict__ u3, size_t n, float* __restrict__ lam,
                                     float* __restrict__ pdf) {
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float l, p;
    sample_lambda(u3[3 * i], u3[3 * i + 1], u3[3 * i + 2], l, p);
    lam[i] = l;
    pdf[i] = p;
}

inline dim3 grid_for(size_t n, int block) { return dim3(static_cast<unsigned>((n + block - 1) / block)); }

}  // namespace

// Builds (once) and returns the spectral context of the current device.
int spectrum_get(const f4** rgb_lut, const f4** lambda_lut, const float** tables) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
    if (dev < 0 || dev >= kMaxDevices) return fail(AP_ERR_UNSUPPORTED, "spectrum: device index out of range");
    std::lock_guard<std::mutex> lock(g_spec_mutex);
    SpectrumDevice& s = g_spec[dev];
    if (!s.ready) {
        e = hipMemcpyToSymbol(HIP_SYMBOL(c_mlp), ap_spectrum_mlp_bin, kMlpFloats * sizeof(float));
        if (e != hipSuccess) return hip_fail(e, "hipMemcpyToSymbol(mlp)");
        const size_t cells = static_cast<size_t>(kRgbLutRes) * kRgbLutRes * kRgbLutRes;
        if ((e = hipMalloc(&s.rgb_lut, cells * sizeof(f4))) != hipSuccess) return hip_fail(e, "hipMalloc(rgb lut)");
        if ((e = hipMalloc(&s.lambda_lut, kLambdaLutRes * sizeof(f4))) != hipSuccess) {
            (void)hipFree(s.rgb_lut);
            s.rgb_lut = nullptr;
            return hip_fail(e, "hipMalloc(lambda lut)");
        }
        if ((e = hipMalloc(&s.tables, 421 * 4 * sizeof(float))) != hipSuccess) {
            (void)hipFree(s.rgb_lut);
            (void)hipFree(s.lambda_lut);
            s.rgb_lut = nullptr;
            s.lambda_lut = nullptr;
            return hip_fail(e, "hipMalloc(cie tables)");
        }
        e = hipMemcpy(s.tables, ap_cie_tables_bin, 421 * 4 * sizeof(float), hipMemcpyHostToDevice);
        if (e != hipSuccess) return hip_fail(e, "hipMemcpy(cie tables)");
        // Slabs of 32 z-planes keep each launch short.
        constexpr uint32_t kSlab = 32;
        for (uint32_t z0 = 0; z0 < static_cast<uint32_t>(kRgbLutRes); z0 += kSlab) {
            const size_t nt = static_cast<size_t>(kRgbLutRes) * kRgbLutRes * kSlab;
            hipLaunchKernelGGL(build_rgb_lut_kernel, grid_for(nt, 256), dim3(256), 0, nullptr, s.rgb_lut, z0, kSlab);
        }
        hipLaunchKernelGGL(build_lambda_lut_kernel, dim3(4), dim3(256), 0, nullptr, s.lambda_lut);
        if ((e = hipGetLastError()) != hipSuccess) return hip_fail(e, "spectrum LUT build launch");
        if ((e = hipDeviceSynchronize()) != hipSuccess) return hip_fail(e, "spectrum LUT build");
        s.ready = true;
    }
    if (rgb_lut) *rgb_lut = s.rgb_lut;
    if (lambda_lut) *lambda_lut = s.lambda_lut;
    if (tables) *tables = s.tables;
    return AP_OK;
}

}  // namespace ap

using namespace ap;

extern "C" {

int ap_spectrum_setup(void) { return spectrum_get(nullptr, nullptr, nullptr); }

#define AP_SPEC_ARGS_CHECK(name)                                                        \
    if (n == 0) return AP_OK;                                                           \
    if (n > (static_cast<size_t>(1) << 33)) return fail(AP_ERR_INVALID_ARG, name ": n too large");

int ap_spectrum_encode_rgb(const float* rgb, size_t n, float* basis, void* stream) {
    AP_SPEC_ARGS_CHECK("ap_spectrum_encode_rgb")
    if (!rgb || !basis) return fail(AP_ERR_INVALID_ARG, "ap_spectrum_encode_rgb: NULL pointer");
    int rc = spectrum_get(nullptr, nullptr, nullptr);
    if (rc != AP_OK) return rc;
    hipLaunchKernelGGL(encode_rgb_kernel, grid_for(n, 128), dim3(128), 0, static_cast<hipStream_t>(stream), rgb, n,
                       reinterpret_cast<f4*>(basis));
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? AP_OK : hip_fail(e, "ap_spectrum_encode_rgb");
}

int ap_spectrum_encode_lambda(const float* lam, size_t n, float* basis, void* stream) {
    AP_SPEC_ARGS_CHECK("ap_spectrum_encode_lambda")
    if (!lam || !basis) return fail(AP_ERR_INVALID_ARG, "ap_spectrum_encode_lambda: NULL pointer");
    int rc = spectrum_get(nullptr, nullptr, nullptr);
    if (rc != AP_OK) return rc;
    hipLaunchKernelGGL(encode_lambda_kernel, grid_for(n, 128), dim3(128), 0, static_cast<hipStream_t>(stream), lam,
                       n, reinterpret_cast<f4*>(basis));
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? AP_OK : hip_fail(e, "ap_spectrum_encode_lambda");
}

int ap_spectrum_query_rgb(const float* rgb, size_t n, float* basis, void* stream) {
    AP_SPEC_ARGS_CHECK("ap_spectrum_query_rgb")
    if (!rgb || !basis) return fail(AP_ERR_INVALID_ARG, "ap_spectrum_query_rgb: NULL pointer");
    const f4* lut = nullptr;
    int rc = spectrum_get(&lut, nullptr, nullptr);
    if (rc != AP_OK) return rc;
    hipLaunchKernelGGL(query_rgb_kernel, grid_for(n, 256), dim3(256), 0, static_cast<hipStream_t>(stream), lut, rgb,
                       n, reinterpret_cast<f4*>(basis));
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? AP_OK : hip_fail(e, "ap_spectrum_query_rgb");
}

int ap_spectrum_query_lambda(const float* lam, size_t n, float* basis, void* stream) {
    AP_SPEC_ARGS_CHECK("ap_spectrum_query_lambda")
    if (!lam || !basis) return fail(AP_ERR_INVALID_ARG, "ap_spectrum_query_lambda: NULL pointer");
    const f4* lut = nullptr;
    int rc = spectrum_get(nullptr, &lut, nullptr);
    if (rc != AP_OK) return rc;
    hipLaunchKernelGGL(query_lambda_kernel, grid_for(n, 256), dim3(256), 0, static_cast<hipStream_t>(stream), lut,
                       lam, n, reinterpret_cast<f4*>(basis));
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? AP_OK : hip_fail(e, "ap_spectrum_query_lambda");
}

int ap_spectrum_interact(const float* lam, const float* rgb, size_t n, float* out, void* stream) {
    AP_SPEC_ARGS_CHECK("ap_spectrum_interact")
    if (!lam || !rgb || !out) return fail(AP_ERR_INVALID_ARG, "ap_spectrum_interact: NULL pointer");
    const f4 *rl = nullptr, *ll = nullptr;
    int rc = spectrum_get(&rl, &ll, nullptr);
    if (rc != AP_OK) return rc;
    hipLaunchKernelGGL(interact_kernel, grid_for(n, 256), dim3(256), 0, static_cast<hipStream_t>(stream), rl, ll, lam,
                       rgb, n, out);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? AP_OK : hip_fail(e, "ap_spectrum_interact");
}

int ap_spectrum_sample_lambda(const float* u3, size_t n, float* lam, float* pdf, void* stream) {
    AP_SPEC_ARGS_CHECK("ap_spectrum_sample_lambda")
    if (!u3 || !lam || !pdf) return fail(AP_ERR_INVALID_ARG, "ap_spectrum_sample_lambda: NULL pointer");
    hipLaunchKernelGGL(sample_lambda_kernel, grid_for(n, 256), dim3(256), 0, static_cast<hipStream_t>(stream), u3, n,
                       lam, pdf);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? AP_OK : hip_fail(e, "ap_spectrum_sample_lambda");
}

}  // extern "C"

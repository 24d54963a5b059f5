// Display transforms (include/aperture/ap_render.h: ap_tonemap): the image
// through XYZ -> linear sRGB, exposure and a tone curve, or a view of the
// screen G-buffer or statistics, packed as ARGB8 (core/scene.cu:88-199).
// One lane per pixel; HBM-bound (12 B read, 4 B written per image pixel).
#include <cstdint>

#include "aperture/ap_render.h"
#include "ap_display.h"
#include "ap_internal.h"

namespace ap {
namespace {

struct v3 {
    float x, y, z;
};

__device__ __forceinline__ float dot3(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
// Row-major 3x3 times a column vector (float3x3::operator*, vector.cuh:346-349).
__device__ __forceinline__ v3 mul(const float m[9], v3 v) {
    return {dot3({m[0], m[1], m[2]}, v), dot3({m[3], m[4], m[5]}, v), dot3({m[6], m[7], m[8]}, v)};
}
__device__ __forceinline__ float lerp1(float a, float b, float t) { return a * (1.0f - t) + b * t; }
__device__ __forceinline__ v3 lerp_gray(float g, v3 c, float t) {
    return {lerp1(g, c.x, t), lerp1(g, c.y, t), lerp1(g, c.z, t)};
}
__device__ __forceinline__ v3 pow3(v3 c, float e) { return {apfm::pow(c.x, e), apfm::pow(c.y, e), apfm::pow(c.z, e)}; }

// Gamma (tonemap.cuh:6-9).
__device__ __forceinline__ v3 gamma(v3 c) { return pow3(c, 1.0f / 2.2f); }

// ACES (tonemap.cuh:11-113): sRGB -> AP0 -> AP1, desaturation, the rational
// fit, dark-to-dim surround in xyY, desaturation, AP1 -> XYZ, D60 -> D65,
// XYZ -> Rec.709, clamp at 0, gamma.
__device__ v3 aces(v3 color) {
    const float srgb_ap0[9] = {0.4397010f, 0.3829780f, 0.1773350f, 0.0897923f, 0.8134230f,
                               0.0967616f, 0.0175440f, 0.1115440f, 0.8707040f};
    const float ap0_ap1[9] = {1.4514393161f, -0.2365107469f, -0.2149285693f, -0.0765537734f, 1.1762296998f,
                              -0.0996759264f, 0.0083161484f, -0.0060324498f, 0.9977163014f};
    const float ap1_xyz[9] = {0.6624541811f, 0.1340042065f, 0.1561876870f, 0.2722287168f, 0.6740817658f,
                              0.0536895174f, -0.0055746495f, 0.0040607335f, 1.0103391003f};
    const float xyz_ap1[9] = {1.6410233797f, -0.3248032942f, -0.2364246952f, -0.6636628587f, 1.6153315917f,
                              0.0167563477f, 0.0117218943f, -0.0082844420f, 0.9883948585f};
    const float d60_d65[9] = {0.98722400f, -0.00611327f, 0.0159533f, -0.00759836f, 1.00186000f,
                              0.0053302f, 0.00307257f, -0.00509595f, 1.0816800f};
    const float xyz_709[9] = {3.2409699419f, -1.5373831776f, -0.4986107603f, -0.9692436363f, 1.8759675015f,
                              0.0415550574f, 0.0556300797f, -0.2039769589f, 1.0569715142f};
    const v3 y_weights{0.272229f, 0.674082f, 0.0536895f};
    const v3 ap0 = mul(srgb_ap0, color);
    v3 cg = mul(ap0_ap1, ap0);
    float tmp = dot3(cg, y_weights);
    cg = lerp_gray(tmp, cg, 0.96f);
    const float a = 278.5085f, b = 10.7772f, c = 293.6045f, d = 88.7122f, e = 80.6889f;
    const v3 x = cg;
    const v3 post{(x.x * (x.x * a + b)) / (x.x * (x.x * c + d) + e), (x.y * (x.y * a + b)) / (x.y * (x.y * c + d) + e),
                  (x.z * (x.z * a + b)) / (x.z * (x.z * c + d) + e)};
    // darkSurround_to_dimSurround (tonemap.cuh:46-65).
    v3 XYZ = mul(ap1_xyz, post);
    const float divisor = fmaxf(dot3(XYZ, {1.0f, 1.0f, 1.0f}), 1.0e-4f);
    v3 xyY{XYZ.x / divisor, XYZ.y / divisor, XYZ.y};
    xyY.z = fminf(fmaxf(xyY.z, 0.0f), 65504.0f);
    xyY.z = apfm::pow(xyY.z, 0.9811f);
    const float m = xyY.z / fmaxf(xyY.y, 1e-4f);
    XYZ = v3{xyY.x, xyY.z, (1.0f - xyY.x - xyY.y)};
    XYZ.x *= m;
    XYZ.z *= m;
    v3 lin = mul(xyz_ap1, XYZ);
    tmp = dot3(lin, y_weights);
    lin = lerp_gray(tmp, lin, 0.93f);
    XYZ = mul(ap1_xyz, lin);
    XYZ = mul(d60_d65, XYZ);
    const v3 r = mul(xyz_709, XYZ);
    return gamma(v3{fmaxf(0.0f, r.x), fmaxf(0.0f, r.y), fmaxf(0.0f, r.z)});
}

// XYZ2SRGBLinearD65 (tonemap.cuh:116-126).
__device__ __forceinline__ v3 xyz_to_linear_srgb(v3 c) {
    const float mtx[9] = {3.2406f, -1.5372f, -0.4986f, -0.9689f, 1.8758f, 0.0415f, 0.0557f, -0.204f, 1.057f};
    const v3 r = mul(mtx, c);
    return {fmaxf(0.0f, r.x), fmaxf(0.0f, r.y), fmaxf(0.0f, r.z)};
}

// HeatMapViridis (color.cuh:99-139).
__device__ v3 viridis(float t) {
    t = fminf(fmaxf(t, 0.0f), 1.0f);
    const v3 c[5] = {{0.267004f, 0.004874f, 0.329415f},
                     {0.229739f, 0.322361f, 0.545706f},
                     {0.127568f, 0.566949f, 0.550556f},
                     {0.369214f, 0.788888f, 0.382914f},
                     {0.993248f, 0.906157f, 0.143936f}};
    const float a = t * 4.0f;
    const int i = static_cast<int>(floorf(a));
    const float f = a - static_cast<float>(i);
    const int k = i <= 0 ? 0 : (i >= 3 ? 3 : i);
    return {c[k].x + (c[k + 1].x - c[k].x) * f, c[k].y + (c[k + 1].y - c[k].y) * f,
            c[k].z + (c[k + 1].z - c[k].z) * f};
}

// __saturatef: NaN -> 0.
__device__ __forceinline__ float saturate(float v) { return v != v ? 0.0f : fminf(fmaxf(v, 0.0f), 1.0f); }
__device__ __forceinline__ uint32_t channel(float v) { return static_cast<uint32_t>(255.0f * saturate(v)); }
__device__ __forceinline__ uint32_t argb(v3 c) {
    return 0xff000000u | (channel(c.x) << 16) | (channel(c.y) << 8) | channel(c.z);
}

__global__ __launch_bounds__(256) void tonemap_kernel(const float* __restrict__ hdr,
                                                      const ap_screen_gbuffer* __restrict__ gb,
                                                      const ap_screen_statistics* __restrict__ st,
                                                      uint32_t* __restrict__ ldr, uint32_t n, int tone, int display,
                                                      float exposure) {
    const uint32_t id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= n) return;
    uint32_t out;
    switch (display) {
        case AP_DISPLAY_ALBEDO: {
            const ap_screen_gbuffer& g = gb[id];
            out = argb({half_to_float(g.albedo[0]), half_to_float(g.albedo[1]), half_to_float(g.albedo[2])});
            break;
        }
        case AP_DISPLAY_NORMAL: {
            const ap_screen_gbuffer& g = gb[id];
            out = argb({half_to_float(g.normal[0]) * 0.5f + 0.5f, half_to_float(g.normal[1]) * 0.5f + 0.5f,
                        half_to_float(g.normal[2]) * 0.5f + 0.5f});
            break;
        }
        case AP_DISPLAY_DEPTH: out = argb(viridis(1.0f / (gb[id].depth + 1.0f))); break;
        case AP_DISPLAY_METALLIC: out = argb(viridis(half_to_float(gb[id].metallic))); break;
        case AP_DISPLAY_ROUGHNESS: out = argb(viridis(half_to_float(gb[id].roughness))); break;
        case AP_DISPLAY_STATISTICS_NUM: {
            // log(double(num) + 1) in FP64, to FP32, times 0.01 (scene.cu:144-150).
            const float heat = static_cast<float>(::log(static_cast<double>(st[id].num) + 1.0)) * 0.01f;
            out = 0xff000000u | (channel(heat) << 16);
            break;
        }
        case AP_DISPLAY_STATISTICS_MEAN:
        case AP_DISPLAY_STATISTICS_M2:
        case AP_DISPLAY_STATISTICS_M3: {
            const ap_float3 v = display == AP_DISPLAY_STATISTICS_MEAN ? st[id].mean
                                : (display == AP_DISPLAY_STATISTICS_M2 ? st[id].m2_div_num : st[id].m3_div_num);
            out = argb({fabsf(v.x) * 0.25f, fabsf(v.y) * 0.25f, fabsf(v.z) * 0.25f});
            break;
        }
        default: {
            const v3 lin = xyz_to_linear_srgb({hdr[3 * id], hdr[3 * id + 1], hdr[3 * id + 2]});
            v3 value{lin.x * exposure, lin.y * exposure, lin.z * exposure};
            if (tone == AP_TONE_GAMMA) value = gamma(value);
            else if (tone == AP_TONE_ACES) value = aces(value);
            out = argb(value);
        }
    }
    ldr[id] = out;
}

}  // namespace
}  // namespace ap

using namespace ap;

extern "C" int ap_tonemap(const float* hdr, const ap_screen_gbuffer* gbuffer, const ap_screen_statistics* statistics,
                          uint32_t* ldr, int32_t width, int32_t height, int32_t tone_type, int32_t display,
                          float exposure, void* stream) {
    if (!ldr) return fail(AP_ERR_INVALID_ARG, "ap_tonemap: NULL ldr");
    if (width <= 0 || height <= 0) return fail(AP_ERR_INVALID_ARG, "ap_tonemap: bad size");
    if (tone_type < AP_TONE_LINEAR || tone_type > AP_TONE_ACES)
        return fail(AP_ERR_INVALID_ARG, "ap_tonemap: tone_type must be AP_TONE_LINEAR, _GAMMA or _ACES");
    if (display < AP_DISPLAY_IMAGE || display > AP_DISPLAY_STATISTICS_M3)
        return fail(AP_ERR_INVALID_ARG, "ap_tonemap: unknown display mode");
    if (display == AP_DISPLAY_IMAGE && !hdr) return fail(AP_ERR_INVALID_ARG, "ap_tonemap: NULL hdr");
    if (display >= AP_DISPLAY_ALBEDO && display <= AP_DISPLAY_ROUGHNESS && !gbuffer)
        return fail(AP_ERR_INVALID_ARG, "ap_tonemap: this display mode reads the G-buffer");
    if (display >= AP_DISPLAY_STATISTICS_NUM && !statistics)
        return fail(AP_ERR_INVALID_ARG, "ap_tonemap: this display mode reads the statistics buffer");
    const uint64_t n = static_cast<uint64_t>(width) * static_cast<uint64_t>(height);
    if (n > 0xffffffffull) return fail(AP_ERR_INVALID_ARG, "ap_tonemap: too many pixels");
    hipLaunchKernelGGL(tonemap_kernel, dim3(static_cast<uint32_t>((n + 255) / 256)), dim3(256), 0,
                       static_cast<hipStream_t>(stream), hdr, gbuffer, statistics, ldr, static_cast<uint32_t>(n),
                       tone_type, display, exposure);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? AP_OK : hip_fail(e, "ap_tonemap");
}

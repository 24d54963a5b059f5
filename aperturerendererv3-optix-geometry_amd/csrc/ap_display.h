// Screen G-buffer and statistics recording and the display transforms
// (tone curves, heat map), device side. The layouts are the reference's
// ScreenGBuffer / ScreenStatisticsBuffer (include/aperture/ap_layout.h).
#pragma once

#include <cstdint>

#include "aperture/ap_layout.h"
#include "ap_fmath.h"

namespace ap {

// What RecordScreenGBuffer reads of a camera ray's first hit
// (denoise.cuh:109-166): the GBuffer albedo, metallic and roughness, the hit
// normal, the hit distance and the shading model; model 4 (MAT_MAX) and hit
// 0 for a miss. 48 B per path.
struct alignas(16) GSample {
    float albedo[3];
    float metallic;
    float normal[3];
    float roughness;
    float depth;
    int32_t model;
    int32_t hit;
    int32_t pad;
};

constexpr int kShadingModels = 5;  // EShadingModel::MAT_MAX + 1 (object.cuh:45-52)

__device__ __forceinline__ float half_to_float(uint16_t h) {
    return static_cast<float>(__builtin_bit_cast(_Float16, h));
}
// FP32 -> binary16, round to nearest even (as CUDA's __float2half), with
// integer operations on the FP32 bits: the FP32 value is rounded first. (A
// plain conversion lets the compiler fuse it with the multiply or add that
// produced the value into one v_fma_mix, which rounds only once.)
__device__ __forceinline__ uint16_t float_to_half(float f) {
    uint32_t u = __float_as_uint(f);
    const uint32_t sign = (u >> 16) & 0x8000u;
    u &= 0x7fffffffu;
    if (u > 0x7f800000u) return static_cast<uint16_t>(sign | 0x7e00u);   // NaN
    if (u >= 0x477ff000u) return static_cast<uint16_t>(sign | 0x7c00u);  // rounds past 65504
    if (u >= 0x38800000u) {                                              // normal binary16
        uint32_t r = u - (112u << 23);
        r += 0xfffu + ((r >> 13) & 1u);
        return static_cast<uint16_t>(sign | (r >> 13));
    }
    if (u <= 0x33000000u) return static_cast<uint16_t>(sign);            // at most 2^-25: zero
    const uint32_t e = u >> 23, m = (u & 0x7fffffu) | 0x800000u;
    const uint32_t sh = 126u - e;
    uint32_t q = m >> sh;
    const uint32_t rem = m & ((1u << sh) - 1u), half = 1u << (sh - 1u);
    if (rem > half || (rem == half && (q & 1u))) ++q;
    return static_cast<uint16_t>(sign | q);
}

// RecordShadingModel (denoise.cuh:84-101).
__device__ __forceinline__ void record_shading_model(ap_screen_gbuffer& g, uint32_t model) {
    if (g.frame_index == 0)
        for (int i = 0; i < kShadingModels; ++i) g.shading_model[i] = 0;
    const float affix = 1.0f / static_cast<float>(g.frame_index + 1ull);
    const float pre = static_cast<float>(g.frame_index) / static_cast<float>(g.frame_index + 1ull);
    for (int i = 0; i < kShadingModels; ++i) g.shading_model[i] = float_to_half(half_to_float(g.shading_model[i]) * pre);
    g.shading_model[model] = float_to_half(half_to_float(g.shading_model[model]) + affix);
}

// The running mean of one half field: h = half(h * pre), then half(h + affix * v).
__device__ __forceinline__ void half_mean(uint16_t& h, float pre, float affix, float v) {
    h = float_to_half(half_to_float(h) * pre);
    h = float_to_half(half_to_float(h) + affix * v);
}

// RecordScreenGBuffer (denoise.cuh:109-166).
__device__ __forceinline__ void record_gbuffer(ap_screen_gbuffer& g, const GSample& s) {
    if (s.hit) {
        if (g.valid_frame_index == 0) {
            for (int k = 0; k < 3; ++k) {
                g.albedo[k] = 0;
                g.normal[k] = 0;
            }
            g.metallic = 0;
            g.roughness = 0;
            g.depth = 0.0f;
        }
        const float affix = 1.0f / static_cast<float>(g.valid_frame_index + 1ull);
        const float pre = static_cast<float>(g.valid_frame_index) / static_cast<float>(g.valid_frame_index + 1ull);
        for (int k = 0; k < 3; ++k) g.albedo[k] = float_to_half(half_to_float(g.albedo[k]) * pre);
        for (int k = 0; k < 3; ++k) g.normal[k] = float_to_half(half_to_float(g.normal[k]) * pre);
        g.metallic = float_to_half(half_to_float(g.metallic) * pre);
        g.roughness = float_to_half(half_to_float(g.roughness) * pre);
        g.depth = g.depth * pre;
        for (int k = 0; k < 3; ++k) g.albedo[k] = float_to_half(half_to_float(g.albedo[k]) + affix * s.albedo[k]);
        for (int k = 0; k < 3; ++k) g.normal[k] = float_to_half(half_to_float(g.normal[k]) + affix * s.normal[k]);
        g.metallic = float_to_half(half_to_float(g.metallic) + affix * s.metallic);
        g.roughness = float_to_half(half_to_float(g.roughness) + affix * s.roughness);
        g.depth = g.depth + affix * s.depth;
        g.valid_frame_index += 1;
        record_shading_model(g, static_cast<uint32_t>(s.model));
    } else {
        record_shading_model(g, static_cast<uint32_t>(kShadingModels - 1));
    }
    g.frame_index += 1;
}

// RecordScreenStatisticBuffer (denoise_statistic.cuh:36-70) of one sample
// xyz: Box-Cox (lambda 0.5) of max(x, 0), then one-pass updates of the
// mean (FP32) and of M2, M3 (FP64 from the stored M/n). x^0.5 is sqrtf.
__device__ __forceinline__ void record_statistics(ap_screen_statistics& h, const float xyz[3]) {
    const uint32_t n0 = h.num;
    const uint32_t n1 = n0 + 1u;
    h.num = n1;
    float* mean = &h.mean.x;
    float* m2 = &h.m2_div_num.x;
    float* m3 = &h.m3_div_num.x;
    const float inv = 1.0f / static_cast<float>(n1);
    for (int k = 0; k < 3; ++k) {
        const float bc = (sqrtf(fmaxf(xyz[k], 0.0f)) - 1.0f) * 2.0f;
        const float delta = bc - mean[k];
        const float dn = delta * inv;
        const float mean1 = mean[k] + dn;
        const double m2_0 = static_cast<double>(m2[k]) * static_cast<double>(n0);
        const double m2_1 = m2_0 + static_cast<double>(delta * (delta - dn));
        const double m3_0 = static_cast<double>(m3[k]) * static_cast<double>(n0);
        const float d2 = delta * delta;
        const float dn2 = dn * dn;
        const double m3_1 = m3_0 - 3.0 * (static_cast<double>(dn) * m2_1) + static_cast<double>(delta * (d2 - dn2));
        mean[k] = mean1;
        m2[k] = static_cast<float>(m2_1 / static_cast<double>(n1));
        m3[k] = static_cast<float>(m3_1 / static_cast<double>(n1));
    }
}

}  // namespace ap

// Albedo texture sampling of the integrator: tex2D<float4> over the
// reference's texture descriptor (core/helper/texture_loader.cuh:36-46:
// normalised coordinates, wrap in both axes, bilinear filter, 8-bit
// channels read as value / 255) with the linear filtering the CUDA
// programming guide documents (texture fetching: x_B = x - 1/2, weights
// alpha = frac(x_B) held with 8 fractional bits, here rounded to nearest).
// The CPU checker restates the same steps, so its samples are the same bits.
#pragma once

#include <cstdint>

#include "ap_vec.h"

namespace ap {

struct alignas(16) TexRec {
    const uint32_t* texels;   // RGBA8 (R in the low byte), rows from the top
    int32_t width, height;
};

AP_HD float tex_channel(uint32_t t, int c) { return static_cast<float>((t >> (8 * c)) & 0xffu) / 255.0f; }

// Wrap addressing: the fraction of u, with a non-finite coordinate read as 0.
AP_HD float tex_wrap01(float u) {
    const float f = u - floorf(u);
    return (f >= 0.0f && f <= 1.0f) ? f : 0.0f;
}

// The four texels around (u, v) blended with 8-bit weights; out = RGBA.
AP_HD void tex_sample(const TexRec& T, float u, float v, float out[4]) {
    const float x = tex_wrap01(u) * static_cast<float>(T.width) - 0.5f;
    const float y = tex_wrap01(v) * static_cast<float>(T.height) - 0.5f;
    const float fx = floorf(x), fy = floorf(y);
    const float a = floorf((x - fx) * 256.0f + 0.5f) * (1.0f / 256.0f);
    const float b = floorf((y - fy) * 256.0f + 0.5f) * (1.0f / 256.0f);
    // x lies in [-1/2, W - 1/2], so fx in [-1, W - 1].
    int32_t i0 = static_cast<int32_t>(fx), j0 = static_cast<int32_t>(fy);
    if (i0 < 0) i0 += T.width;
    if (j0 < 0) j0 += T.height;
    const int32_t i1 = i0 + 1 == T.width ? 0 : i0 + 1;
    const int32_t j1 = j0 + 1 == T.height ? 0 : j0 + 1;
    const size_t r0 = static_cast<size_t>(j0) * static_cast<size_t>(T.width);
    const size_t r1 = static_cast<size_t>(j1) * static_cast<size_t>(T.width);
    const uint32_t t00 = T.texels[r0 + i0], t10 = T.texels[r0 + i1];
    const uint32_t t01 = T.texels[r1 + i0], t11 = T.texels[r1 + i1];
    const float w00 = (1.0f - a) * (1.0f - b), w10 = a * (1.0f - b);
    const float w01 = (1.0f - a) * b, w11 = a * b;
    for (int c = 0; c < 4; ++c)
        out[c] = w00 * tex_channel(t00, c) + w10 * tex_channel(t10, c) + w01 * tex_channel(t01, c) +
                 w11 * tex_channel(t11, c);
}

}  // namespace ap

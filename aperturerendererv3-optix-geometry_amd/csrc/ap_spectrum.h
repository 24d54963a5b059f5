// Spectral upsampling and wavelength sampling on the device.
//
//   SampleCIE1931 (+ ndtrif, per-lobe samplers, mixture pdf)
//                                  core/spectrum/sample_ciexyz.cuh:30-326
//   D65Norm, LambdaToCIE1931_XYZ   core/spectrum/ciexyz.cuh:67-85, 522-546
//   RGBEncode / LambdaEncode       core/spectrum/spectrum_basis.cuh:2273-2458
//     with LinearRowMajor's fma chain (W[i*From+j], bias first, j ascending)
//                                  core/tensor/tensor.cuh:104-256
//   SpectralInteract               core/spectrum/spectrum_basis.cuh:2461-2470
//   QueryRGBBasis / QueryLambda / Interact
//                                  core/spectrum/spectrum_lut.cuh:124-235
//   XYZ2SRGBLinearD65              core/tonemap.cuh:116-126
// Transcendentals are the HIP device library's accurate versions.
#pragma once

#include "ap_fmath.h"
#include "ap_vec.h"

namespace ap {

constexpr int kRgbLutRes = 512;
constexpr int kLambdaLutRes = 840;
constexpr int kMlpFloats = 16096;

// Offsets of the weight blocks inside the packed MLP table.
struct MlpLayout {
    static constexpr int rgb_in_w = 0, rgb_in_b = 1024;
    static constexpr int rgb_blk = 1056;                    // 7 x (W 1024, B 32)
    static constexpr int rgb_out_w = 1056 + 7 * 1056;       // 128
    static constexpr int lam_blk = rgb_out_w + 128;         // 7 x (W 1024, B 32)
    static constexpr int lam_out_w = lam_blk + 7 * 1056;    // 128
    static_assert(lam_out_w + 128 == kMlpFloats, "packed MLP size");
};

struct f4 { float x, y, z, w; };

AP_HD float ndtrif(float u) {
    u = clampf(u, 0.0f, 1.0f);
    const float uu = u;
    constexpr float a0 = -3.969683028665376e+01f, a1 = 2.209460984245205e+02f, a2 = -2.759285104469687e+02f,
                    a3 = 1.383577518672690e+02f, a4 = -3.066479806614716e+01f, a5 = 2.506628277459239e+00f;
    constexpr float b0 = -5.447609879822406e+01f, b1 = 1.615858368580409e+02f, b2 = -1.556989798598866e+02f,
                    b3 = 6.680131188771972e+01f, b4 = -1.328068155288572e+01f;
    constexpr float c0 = -7.784894002430293e-03f, c1 = -3.223964580411365e-01f, c2 = -2.400758277161838e+00f,
                    c3 = -2.549732539343734e+00f, c4 = 4.374664141464968e+00f, c5 = 2.938163982698783e+00f;
    constexpr float d0 = 7.784695709041462e-03f, d1 = 3.224671290700398e-01f, d2 = 2.445134137142996e+00f,
                    d3 = 3.754408661907416e+00f;
    constexpr float pLow = 0.02425f, pHigh = 1.0f - pLow;
    float x;
    if (uu < pLow || uu > pHigh) {
        const bool low = uu < pLow;
        const float q = sqrtf(-2.0f * (low ? apfm::log(uu) : apfm::log1p(-uu)));
        float num = fmaf(fmaf(fmaf(fmaf(fmaf(c0, q, c1), q, c2), q, c3), q, c4), q, c5);
        float den = fmaf(fmaf(fmaf(fmaf(d0, q, d1), q, d2), q, d3), q, 1.0f);
        x = low ? num / den : -num / den;
    } else {
        const float q = uu - 0.5f;
        const float r = q * q;
        float num = fmaf(fmaf(fmaf(fmaf(fmaf(a0, r, a1), r, a2), r, a3), r, a4), r, a5);
        num *= q;
        const float den = fmaf(fmaf(fmaf(fmaf(fmaf(b0, r, b1), r, b2), r, b3), r, b4), r, 1.0f);
        x = num / den;
    }
    const float expo = apfm::exp(fmaf(-0.5f, x * x, 0.0f));
    const float pdf = expo * 0.3989422804014327f;
    const float cdf = 0.5f * apfm::erfc(-x * 0.7071067811865476f);
    return x - (cdf - uu) / pdf;
}

AP_HD float clamp_lambda(float h) { return h < 360.0f ? 360.0f : (h > 780.0f ? 780.0f : h); }

AP_HD float lambda_pdf_x(float l) {
    const float k1 = (l - 595.8f) / 33.33f, k2 = (l - 446.8f) / 19.44f;
    return (1.065f * apfm::exp(-0.5f * k1 * k1) + 0.366f * apfm::exp(-0.5f * k2 * k2)) / 106.81109281f;
}
AP_HD float lambda_pdf_y(float l) {
    const float k = (apfm::log(l) - 6.3213077f) / 0.075f;
    return 1.014f * apfm::exp(-0.5f * k * k) / 106.34513640f;
}
AP_HD float lambda_pdf_z(float l) {
    const float k = (apfm::log(l) - 6.1088030f) / 0.051f;
    return 1.839f * apfm::exp(-0.5f * k * k) / 105.88243587f;
}

// PdfLambdaToCIE1931 of the mixture below, floored at 1e-16 as the sampler floors it.
AP_HD float lambda_pdf_mix(float l) {
    constexpr float total = 0.1f + 1.0f + 1.0f + 1.0f;
    constexpr float uR = 0.1f / total, xR = 1.0f / total, yR = 1.0f / total, zR = 1.0f / total;
    const float p = uR * 1.0f / (780.0f - 360.0f) + xR * lambda_pdf_x(l) + yR * lambda_pdf_y(l) + zR * lambda_pdf_z(l);
    return fmaxf(p, 1e-16f);
}

// SampleCIE1931: 0.1 uniform + X (two Gaussians) + Y, Z (log-normal).
AP_HD void sample_lambda(float r0, float r1, float r2, float& lambda, float& pdf) {
    constexpr float total = 0.1f + 1.0f + 1.0f + 1.0f;
    constexpr float uR = 0.1f / total, xR = 1.0f / total, yR = 1.0f / total, zR = 1.0f / total;
    float l;
    if (r0 < uR) {
        l = r1 * (780.0f - 360.0f) + 360.0f;
    } else if (r0 < uR + xR) {
        if (r1 < 0.8330258723493694f) {
            const float t = 7.4884543e-13f + r2 * (0.9999999836707862f - 7.4884543e-13f);
            l = clamp_lambda(595.7999999999999545f + 33.3299999999999983f * ndtrif(t));
        } else {
            const float t = 0.0000040030528583f + r2 * (1.0f - 0.0000040030528583f);
            l = clamp_lambda(446.8000000000000114f + 19.4400000000000013f * ndtrif(t));
        }
    } else if (r0 < uR + xR + yR) {
        const float t = 0.0000000020798312f + r1 * (0.9999953206338883f - 0.0000000020798312f);
        l = clamp_lambda(apfm::exp(6.3269327170812479f + 0.0750000000000000f * ndtrif(t)));
    } else {
        const float t = 0.0000049890553249f + r1 * (1.0f - 0.0000049890553249f);
        l = clamp_lambda(apfm::exp(6.1114040395252154f + 0.0510000000000000f * ndtrif(t)));
    }
    lambda = l;
    pdf = lambda_pdf_mix(l);
}

// tables: D65[421] then CMF[421*3].
AP_HD float d65_norm(const float* tables, float lambda) {
    constexpr float k = 1.0f / 10567.065f;
    if (lambda <= 360.0f) return tables[0] * k;
    if (lambda >= 780.0f) return tables[420] * k;
    const float x = lambda - 360.0f;
    const int i = static_cast<int>(floorf(x));
    const float t = x - static_cast<float>(i);
    return (tables[i] + (tables[i + 1] - tables[i]) * t) * k;
}

AP_HD f3 lambda_to_xyz(const float* tables, float lambda) {
    const float* C = tables + 421;
    if (lambda <= 360.0f) return {C[0], C[1], C[2]};
    if (lambda >= 780.0f) return {C[1260], C[1261], C[1262]};
    const float u = (lambda - 360.0f) * (1.0f / 1.0f);
    const int i = static_cast<int>(u);
    const float t = u - static_cast<float>(i);
    const float* a = C + 3 * i;
    const float* b = a + 3;
    return {a[0] + (b[0] - a[0]) * t, a[1] + (b[1] - a[1]) * t, a[2] + (b[2] - a[2]) * t};
}

AP_HD f3 xyz_to_srgb(f3 xyz) {
    const m3 M{{3.2406f, -1.5372f, -0.4986f}, {-0.9689f, 1.8758f, 0.0415f}, {0.0557f, -0.204f, 1.057f}};
    const f3 r = mul(M, xyz);
    return {fmaxf(0.0f, r.x), fmaxf(0.0f, r.y), fmaxf(0.0f, r.z)};
}

AP_HD float spectral_interact(f4 a, f4 b) {
    float acc = 0.0f;
    acc = fmaf(a.x, b.x, acc);
    acc = fmaf(a.y, b.y, acc);
    acc = fmaf(a.z, b.z, acc);
    acc = fmaf(a.w, b.w, acc);
    return fminf(fmaxf(acc, 0.0f), 1.0f);
}

// --- MLP (W row-major [out][in], acc = bias, then fma over inputs in order) ---
template <int From, int To, bool Bias, bool ReLU>
AP_HD void mlp_linear(const float* x, float* y, const float* W, const float* B) {
#pragma unroll
    for (int i = 0; i < To; ++i) {
        float acc = Bias ? B[i] : 0.0f;
#pragma unroll
        for (int j = 0; j < From; ++j) acc = fmaf(x[j], W[i * From + j], acc);
        y[i] = ReLU ? fmaxf(acc, 0.0f) : acc;
    }
}

AP_HD void mlp_residual(float* h, const float* W, const float* B) {
    float t[32];
    mlp_linear<32, 32, true, true>(h, t, W, B);
#pragma unroll
    for (int i = 0; i < 32; ++i) h[i] = h[i] + t[i];
}

AP_HD f4 rgb_encode(const float* w, float r, float g, float b) {
    const float in3[3] = {r, g, b};
    const float scale = fmaxf(fmaxf(r, g), b);
    const float inv = 1.0f / (scale + 0.001f);
    float s[32];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const float ph = in3[c] * inv * kPiF * static_cast<float>(k + 1);
            s[k + 8 * c] = apfm::sin(ph);
            s[k + 4 + 8 * c] = apfm::cos(ph);
        }
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const float ph = scale * kPiF * static_cast<float>(k + 1);
        s[k + 24] = apfm::sin(ph);
        s[k + 28] = apfm::cos(ph);
    }
    float h[32];
    mlp_linear<32, 32, true, true>(s, h, w + MlpLayout::rgb_in_w, w + MlpLayout::rgb_in_b);
#pragma unroll
    for (int k = 0; k < 7; ++k)
        mlp_residual(h, w + MlpLayout::rgb_blk + 1056 * k, w + MlpLayout::rgb_blk + 1056 * k + 1024);
    float o[4];
    mlp_linear<32, 4, false, false>(h, o, w + MlpLayout::rgb_out_w, nullptr);
    return {o[0] * scale, o[1] * scale, o[2] * scale, o[3] * scale};
}

AP_HD f4 lambda_encode(const float* w, float lambda) {
    const float x = fminf(fmaxf((lambda - 360.0f) / (780.0f - 360.0f), 0.0f), 1.0f);
    float h[32];
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        const float ph = x * kPiF * static_cast<float>(k + 1);
        h[k] = apfm::sin(ph);
        h[k + 16] = apfm::cos(ph);
    }
#pragma unroll
    for (int k = 0; k < 7; ++k)
        mlp_residual(h, w + MlpLayout::lam_blk + 1056 * k, w + MlpLayout::lam_blk + 1056 * k + 1024);
    float o[4];
    mlp_linear<32, 4, false, false>(h, o, w + MlpLayout::lam_out_w, nullptr);
    return {o[0], o[1], o[2], o[3]};
}

// --- LUT queries (cell centres (i + 0.5) / RES, edge clamp) ---
AP_HD void lut_coord(float g, int res, int& i0, int& i1, float& t) {
    const float fi0 = floorf(g);
    i0 = static_cast<int>(fi0);
    i1 = i0 + 1;
    if (i1 >= res) i1 = res - 1;
    t = g - fi0;
    if (i0 == i1) t = 0.0f;
}

AP_HD f4 accum4(f4 acc, float w, const f4 v) {
    return {fmaf(w, v.x, acc.x), fmaf(w, v.y, acc.y), fmaf(w, v.z, acc.z), fmaf(w, v.w, acc.w)};
}

AP_HD f4 query_rgb_basis(const f4* lut, f3 rgb) {
    constexpr int R = kRgbLutRes;
    float gx = rgb.x * static_cast<float>(R) - 0.5f;
    float gy = rgb.y * static_cast<float>(R) - 0.5f;
    float gz = rgb.z * static_cast<float>(R) - 0.5f;
    const float mc = static_cast<float>(R) - 1.0f;
    gx = fminf(fmaxf(gx, 0.0f), mc);
    gy = fminf(fmaxf(gy, 0.0f), mc);
    gz = fminf(fmaxf(gz, 0.0f), mc);
    int x0, x1, y0, y1, z0, z1;
    float tx, ty, tz;
    lut_coord(gx, R, x0, x1, tx);
    lut_coord(gy, R, y0, y1, ty);
    lut_coord(gz, R, z0, z1, tz);
    const float wx0 = 1.0f - tx, wx1 = tx, wy0 = 1.0f - ty, wy1 = ty, wz0 = 1.0f - tz, wz1 = tz;
    auto at = [&](int x, int y, int z) { return lut[(static_cast<size_t>(z) * R + y) * R + x]; };
    f4 r{0.0f, 0.0f, 0.0f, 0.0f};
    r = accum4(r, wx0 * wy0 * wz0, at(x0, y0, z0));
    r = accum4(r, wx1 * wy0 * wz0, at(x1, y0, z0));
    r = accum4(r, wx0 * wy1 * wz0, at(x0, y1, z0));
    r = accum4(r, wx1 * wy1 * wz0, at(x1, y1, z0));
    r = accum4(r, wx0 * wy0 * wz1, at(x0, y0, z1));
    r = accum4(r, wx1 * wy0 * wz1, at(x1, y0, z1));
    r = accum4(r, wx0 * wy1 * wz1, at(x0, y1, z1));
    r = accum4(r, wx1 * wy1 * wz1, at(x1, y1, z1));
    return r;
}

AP_HD f4 query_lambda(const f4* lut, float lambda) {
    constexpr int R = kLambdaLutRes;
    float u = (lambda - 360.0f) / (780.0f - 360.0f);
    u = fminf(fmaxf(u, 0.0f), 1.0f);
    float g = u * static_cast<float>(R) - 0.5f;
    g = fminf(fmaxf(g, 0.0f), static_cast<float>(R) - 1.0f);
    int i0, i1;
    float t;
    lut_coord(g, R, i0, i1, t);
    const float w0 = 1.0f - t, w1 = t;
    const f4 a = lut[i0], b = lut[i1];
    return {a.x * w0 + b.x * w1, a.y * w0 + b.y * w1, a.z * w0 + b.z * w1, a.w * w0 + b.w * w1};
}

// Interact: values above 1 are normalised by their maximum and rescaled.
AP_HD float interact(const f4* lut, f4 lambda_basis, f3 rgb) {
    const float lum = max3(rgb);
    if (lum > 1.0f) return spectral_interact(query_rgb_basis(lut, rgb / lum), lambda_basis) * lum;
    return spectral_interact(query_rgb_basis(lut, rgb), lambda_basis);
}

}  // namespace ap

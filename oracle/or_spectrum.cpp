// ORACLE — test infrastructure only: CPU restatement of the reference's
// spectral upsampling and wavelength sampling (the checker for the core's
// spectrum kernels; the product never links it).
//
//   LinearRowMajor (fma chain, W[i*From+j])  core/tensor/tensor.cuh:104-256
//   RGBEncode / LambdaEncode / SpectralInteract  core/spectrum/spectrum_basis.cuh:2261-2470
//   PrecomputeRGBLUTKernel / Lambda LUT (cell centres)  core/spectrum/spectrum_lut.cuh:12-60
//   QueryRGBBasis / QueryLambda / Interact   core/spectrum/spectrum_lut.cuh:124-235
//   ndtrif, SampleCIE1931_X/Y/Z, PdfLambdaToCIE1931_*, SampleCIE1931  core/spectrum/sample_ciexyz.cuh:30-326
//   D65Norm, LambdaToCIE1931_XYZ             core/spectrum/ciexyz.cuh:67-85, 522-546
//   XYZ2SRGBLinearD65                        core/tonemap.cuh:116-126
// The reference evaluates sincosf/expf/erfcf with CUDA fast-math; this
// restatement calls glibc (parity for those is tolerance-based).
// Parity pins: the reference's golden vectors inputTest/gtTest
// (tests/golden/spectrum_golden.json), ndtrif(0.5) = 0 and the round trip
// u = Phi(ndtrif(u)), pdf normalisation and sample/pdf consistency, table
// lookups exact at integer nm, XYZ->sRGB of the D65 white.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "or_api.h"
#include "ap_fmath.h"
#include "or_math.h"
#include "or_spectrum.h"

namespace orc {

SpectrumTables g_spec;

static bool load_floats(const char* path, float* dst, size_t n) {
    FILE* f = std::fopen(path, "rb");
    if (!f) return false;
    const size_t got = std::fread(dst, sizeof(float), n, f);
    std::fclose(f);
    return got == n;
}

template <int From, int To, bool Bias, bool ReLU>
static void Linear(const float* in, float* out, const float* W, const float* B) {
    float acc[To];
    for (int i = 0; i < To; ++i) acc[i] = Bias ? B[i] : 0.0f;
    for (int j = 0; j < From; ++j)
        for (int i = 0; i < To; ++i) acc[i] = std::fma(in[j], W[i * From + j], acc[i]);
    for (int i = 0; i < To; ++i) out[i] = ReLU ? std::fmax(acc[i], 0.0f) : acc[i];
}

static void Residual(float* h, const float* W, const float* B) {
    float t[32];
    Linear<32, 32, true, true>(h, t, W, B);
    for (int i = 0; i < 32; ++i) h[i] = h[i] + t[i];
}

void RGBEncode(const float in3[3], float outK[4]) {
    const SpectrumTables& T = g_spec;
    const float scale = std::fmax(std::fmax(in3[0], in3[1]), in3[2]);
    const float scaleEps = scale + 0.001f;
    const float scaleEpsInv = 1.0f / scaleEps;
    float s[32];
    for (int c = 0; c < 3; ++c)
        for (int k = 0; k < 4; ++k) {
            const float phase = in3[c] * scaleEpsInv * PI_F * static_cast<float>(k + 1);
            s[k + 8 * c] = apfm::sin(phase);
            s[k + 4 + 8 * c] = apfm::cos(phase);
        }
    for (int k = 0; k < 4; ++k) {
        const float phase = scale * PI_F * static_cast<float>(k + 1);
        s[k + 24] = apfm::sin(phase);
        s[k + 28] = apfm::cos(phase);
    }
    float h[32];
    Linear<32, 32, true, true>(s, h, T.rgb_in_w, T.rgb_in_b);
    for (int b = 0; b < 7; ++b) Residual(h, T.rgb_blk_w[b], T.rgb_blk_b[b]);
    Linear<32, 4, false, false>(h, outK, T.rgb_out_w, nullptr);
    for (int i = 0; i < 4; ++i) outK[i] *= scale;
}

void LambdaEncode(float lambda, float outK[4]) {
    const SpectrumTables& T = g_spec;
    const float x = std::fmin(std::fmax((lambda - 360.0f) / (780.0f - 360.0f), 0.0f), 1.0f);
    float h[32];
    for (int k = 0; k < 16; ++k) {
        const float phase = x * PI_F * static_cast<float>(k + 1);
        h[k] = apfm::sin(phase);
        h[k + 16] = apfm::cos(phase);
    }
    for (int b = 0; b < 7; ++b) Residual(h, T.lam_blk_w[b], T.lam_blk_b[b]);
    Linear<32, 4, false, false>(h, outK, T.lam_out_w, nullptr);
}

float SpectralInteract(const float* rgbK, const float* lamK) {
    float acc = 0.0f;
    for (int i = 0; i < 4; ++i) acc = std::fma(rgbK[i], lamK[i], acc);
    return std::fmin(std::fmax(acc, 0.0f), 1.0f);
}

// LUT cell value computed on demand (PrecomputeRGBLUTKernel, spectrum_lut.cuh:12-37).
static void RGBCell(int x, int y, int z, float out[4]) {
    const float inv = 1.0f / 512.0f;
    const float in3[3] = {(x + 0.5f) * inv, (y + 0.5f) * inv, (z + 0.5f) * inv};
    RGBEncode(in3, out);
}

static void LinearCoord(float g, int res, int& i0, int& i1, float& t) {
    const float fi0 = std::floor(g);
    i0 = static_cast<int>(fi0);
    i1 = i0 + 1;
    if (i1 >= res) i1 = res - 1;
    t = g - fi0;
    if (i0 == i1) t = 0.0f;
}

void QueryRGBBasis(float3 rgb, float out[4]) {
    const int RES = 512;
    float g[3] = {rgb.x * RES - 0.5f, rgb.y * RES - 0.5f, rgb.z * RES - 0.5f};
    const float maxC = RES - 1.0f;
    int i0[3], i1[3];
    float t[3];
    for (int a = 0; a < 3; ++a) {
        g[a] = std::fmin(std::fmax(g[a], 0.0f), maxC);
        LinearCoord(g[a], RES, i0[a], i1[a], t[a]);
    }
    const float wx0 = 1.0f - t[0], wx1 = t[0], wy0 = 1.0f - t[1], wy1 = t[1], wz0 = 1.0f - t[2], wz1 = t[2];
    const float w[8] = {wx0 * wy0 * wz0, wx1 * wy0 * wz0, wx0 * wy1 * wz0, wx1 * wy1 * wz0,
                        wx0 * wy0 * wz1, wx1 * wy0 * wz1, wx0 * wy1 * wz1, wx1 * wy1 * wz1};
    const int cx[8] = {i0[0], i1[0], i0[0], i1[0], i0[0], i1[0], i0[0], i1[0]};
    const int cy[8] = {i0[1], i0[1], i1[1], i1[1], i0[1], i0[1], i1[1], i1[1]};
    const int cz[8] = {i0[2], i0[2], i0[2], i0[2], i1[2], i1[2], i1[2], i1[2]};
    for (int k = 0; k < 4; ++k) out[k] = 0.0f;
    for (int c = 0; c < 8; ++c) {
        float v[4];
        RGBCell(cx[c], cy[c], cz[c], v);
        for (int k = 0; k < 4; ++k) out[k] = std::fma(w[c], v[k], out[k]);
    }
}

void QueryLambda(float lambda, float out[4]) {
    const int RES = 840;
    float u = (lambda - 360.0f) / (780.0f - 360.0f);
    u = std::fmin(std::fmax(u, 0.0f), 1.0f);
    float g = u * static_cast<float>(RES) - 0.5f;
    g = std::fmin(std::fmax(g, 0.0f), static_cast<float>(RES) - 1.0f);
    int i0, i1;
    float t;
    LinearCoord(g, RES, i0, i1, t);
    const float step = (780.0f - 360.0f) / static_cast<float>(RES);
    float a[4], b[4];
    LambdaEncode(360.0f + (i0 + 0.5f) * step, a);
    LambdaEncode(360.0f + (i1 + 0.5f) * step, b);
    const float w0 = 1.0f - t, w1 = t;
    for (int k = 0; k < 4; ++k) out[k] = a[k] * w0 + b[k] * w1;
}

float Interact(const float lamBasis[4], float3 rgb) {
    const float lum = max3(rgb);
    float rb[4];
    if (lum > 1.0f) {
        QueryRGBBasis(rgb / lum, rb);
        return SpectralInteract(rb, lamBasis) * lum;
    }
    QueryRGBBasis(rgb, rb);
    return SpectralInteract(rb, lamBasis);
}

// ---- wavelength sampling (sample_ciexyz.cuh) ----
float ndtrif(float u) {
    u = clampf(u, 0.0f, 1.0f);
    const float uu = u;
    const float a[] = {-3.969683028665376e+01f, 2.209460984245205e+02f, -2.759285104469687e+02f,
                       1.383577518672690e+02f, -3.066479806614716e+01f, 2.506628277459239e+00f};
    const float b[] = {-5.447609879822406e+01f, 1.615858368580409e+02f, -1.556989798598866e+02f,
                       6.680131188771972e+01f, -1.328068155288572e+01f};
    const float c[] = {-7.784894002430293e-03f, -3.223964580411365e-01f, -2.400758277161838e+00f,
                       -2.549732539343734e+00f, 4.374664141464968e+00f, 2.938163982698783e+00f};
    const float d[] = {7.784695709041462e-03f, 3.224671290700398e-01f, 2.445134137142996e+00f,
                       3.754408661907416e+00f};
    const float pLow = 0.02425f, pHigh = 1.0f - pLow;
    float x;
    if (uu < pLow || uu > pHigh) {
        const bool low = uu < pLow;
        const float q = std::sqrt(-2.0f * (low ? apfm::log(uu) : apfm::log1p(-uu)));
        float num = std::fma(c[0], q, c[1]);
        num = std::fma(num, q, c[2]);
        num = std::fma(num, q, c[3]);
        num = std::fma(num, q, c[4]);
        num = std::fma(num, q, c[5]);
        float den = std::fma(d[0], q, d[1]);
        den = std::fma(den, q, d[2]);
        den = std::fma(den, q, d[3]);
        den = std::fma(den, q, 1.0f);
        x = low ? num / den : -num / den;
    } else {
        const float q = uu - 0.5f;
        const float r = q * q;
        float num = std::fma(a[0], r, a[1]);
        num = std::fma(num, r, a[2]);
        num = std::fma(num, r, a[3]);
        num = std::fma(num, r, a[4]);
        num = std::fma(num, r, a[5]);
        num *= q;
        float den = std::fma(b[0], r, b[1]);
        den = std::fma(den, r, b[2]);
        den = std::fma(den, r, b[3]);
        den = std::fma(den, r, b[4]);
        den = std::fma(den, r, 1.0f);
        x = num / den;
    }
    const float INV_SQRT_2PI = 0.3989422804014327f, INV_SQRT2 = 0.7071067811865476f;
    const float expo = apfm::exp(std::fma(-0.5f, x * x, 0.0f));
    const float pdf = expo * INV_SQRT_2PI;
    const float cdf = 0.5f * apfm::erfc(-x * INV_SQRT2);
    x -= (cdf - uu) / pdf;
    return x;
}

static float clampL(float h) { return h < 360.0f ? 360.0f : (h > 780.0f ? 780.0f : h); }

float SampleX(float uComp, float uInv) {
    const float alpha0 = 0.8330258723493694f;
    if (uComp < alpha0) {
        const float t0 = 7.4884543e-13f + uInv * (0.9999999836707862f - 7.4884543e-13f);
        return clampL(595.7999999999999545f + 33.3299999999999983f * ndtrif(t0));
    }
    const float t1 = 0.0000040030528583f + uInv * (1.0f - 0.0000040030528583f);
    return clampL(446.8000000000000114f + 19.4400000000000013f * ndtrif(t1));
}

float SampleY(float u) {
    const float t = 0.0000000020798312f + u * (0.9999953206338883f - 0.0000000020798312f);
    return clampL(apfm::exp(6.3269327170812479f + 0.0750000000000000f * ndtrif(t)));
}

float SampleZ(float u) {
    const float t = 0.0000049890553249f + u * (1.0f - 0.0000049890553249f);
    return clampL(apfm::exp(6.1114040395252154f + 0.0510000000000000f * ndtrif(t)));
}

float PdfX(float l) {
    const float k1 = (l - 595.8f) / 33.33f, k2 = (l - 446.8f) / 19.44f;
    return (1.065f * apfm::exp(-0.5f * k1 * k1) + 0.366f * apfm::exp(-0.5f * k2 * k2)) / 106.81109281f;
}
float PdfY(float l) {
    const float k = (apfm::log(l) - 6.3213077f) / 0.075f;
    return 1.014f * apfm::exp(-0.5f * k * k) / 106.34513640f;
}
float PdfZ(float l) {
    const float k = (apfm::log(l) - 6.1088030f) / 0.051f;
    return 1.839f * apfm::exp(-0.5f * k * k) / 105.88243587f;
}

// The mixture pdf SampleCIE1931 returns, floored at 1e-16.
float PdfCIE1931Mix(float l) {
    const float total = 0.1f + 1.0f + 1.0f + 1.0f;
    const float uR = 0.1f / total, xR = 1.0f / total, yR = 1.0f / total, zR = 1.0f / total;
    const float p = uR * 1.0f / (780.0f - 360.0f) + xR * PdfX(l) + yR * PdfY(l) + zR * PdfZ(l);
    return std::fmax(p, 1e-16f);
}

void SampleCIE1931(float r0, float r1, float r2, float* lambda, float* pdf) {
    const float total = 0.1f + 1.0f + 1.0f + 1.0f;
    const float uR = 0.1f / total, xR = 1.0f / total, yR = 1.0f / total;
    float l;
    if (r0 < uR) l = r1 * (780.0f - 360.0f) + 360.0f;
    else if (r0 < uR + xR) l = SampleX(r1, r2);
    else if (r0 < uR + xR + yR) l = SampleY(r1);
    else l = SampleZ(r1);
    *lambda = l;
    *pdf = PdfCIE1931Mix(l);
}

float D65Norm(float lambda) {
    const float k = 1.0f / 10567.065f;
    if (lambda <= 360) return g_spec.d65[0] * k;
    if (lambda >= 780) return g_spec.d65[420] * k;
    const float x = lambda - 360.0f;
    const int i = static_cast<int>(std::floor(x));
    const float t = x - static_cast<float>(i);
    return (g_spec.d65[i] + (g_spec.d65[i + 1] - g_spec.d65[i]) * t) * k;
}

float3 LambdaToXYZ(float lambda) {
    const float* C = g_spec.cmf;
    if (lambda <= 360.0f) return {C[0], C[1], C[2]};
    if (lambda >= 780.0f) return {C[420 * 3], C[420 * 3 + 1], C[420 * 3 + 2]};
    const float u = (lambda - 360.0f) * (1.0f / 1.0f);
    const int i = static_cast<int>(u);
    const float t = u - static_cast<float>(i);
    const float* a = C + 3 * i;
    const float* b = C + 3 * (i + 1);
    return {a[0] + (b[0] - a[0]) * t, a[1] + (b[1] - a[1]) * t, a[2] + (b[2] - a[2]) * t};
}

float3 XYZ2SRGB(float3 xyz) {
    const float3x3 M{{3.2406f, -1.5372f, -0.4986f}, {-0.9689f, 1.8758f, 0.0415f}, {0.0557f, -0.204f, 1.057f}};
    const float3 r = M * xyz;
    return {std::fmax(0.0f, r.x), std::fmax(0.0f, r.y), std::fmax(0.0f, r.z)};
}

}  // namespace orc

using namespace orc;

extern "C" {

int or_spectrum_load(const char* mlp_path, const char* cie_path) {
    std::vector<float> w(16096);
    if (!load_floats(mlp_path, w.data(), w.size())) return -1;
    std::vector<float> c(421 * 4);
    if (!load_floats(cie_path, c.data(), c.size())) return -2;
    SpectrumTables& T = g_spec;
    T.storage = w;
    const float* p = T.storage.data();
    T.rgb_in_w = p; p += 1024;
    T.rgb_in_b = p; p += 32;
    for (int b = 0; b < 7; ++b) { T.rgb_blk_w[b] = p; p += 1024; T.rgb_blk_b[b] = p; p += 32; }
    T.rgb_out_w = p; p += 128;
    for (int b = 0; b < 7; ++b) { T.lam_blk_w[b] = p; p += 1024; T.lam_blk_b[b] = p; p += 32; }
    T.lam_out_w = p;
    std::memcpy(T.d65, c.data(), 421 * sizeof(float));
    std::memcpy(T.cmf, c.data() + 421, 421 * 3 * sizeof(float));
    T.loaded = true;
    return 0;
}

void or_rgb_encode(const float* rgb, size_t n, float* out) {
    for (size_t i = 0; i < n; ++i) RGBEncode(rgb + 3 * i, out + 4 * i);
}
void or_lambda_encode(const float* lam, size_t n, float* out) {
    for (size_t i = 0; i < n; ++i) LambdaEncode(lam[i], out + 4 * i);
}
float or_spectral_interact(const float* rgbK, const float* lamK) { return SpectralInteract(rgbK, lamK); }
void or_query_rgb_basis(const float* rgb, size_t n, float* out) {
    for (size_t i = 0; i < n; ++i) QueryRGBBasis({rgb[3 * i], rgb[3 * i + 1], rgb[3 * i + 2]}, out + 4 * i);
}
void or_query_lambda(const float* lam, size_t n, float* out) {
    for (size_t i = 0; i < n; ++i) QueryLambda(lam[i], out + 4 * i);
}
void or_interact(const float* lam, const float* rgb, size_t n, float* out) {
    for (size_t i = 0; i < n; ++i) {
        float lb[4];
        QueryLambda(lam[i], lb);
        out[i] = Interact(lb, float3{rgb[3 * i], rgb[3 * i + 1], rgb[3 * i + 2]});
    }
}
// The portable transcendentals shared with the core (ap_fmath.h), for their
// accuracy tests: fn 0 sin, 1 cos, 2 exp, 3 log, 4 log1p, 5 erfc, 6 x^0.75,
// 7 atan, 8 atan2(x[i], x[n + i]) (then x holds 2n values).
void or_fmath(int fn, const float* x, size_t n, float* out) {
    for (size_t i = 0; i < n; ++i) {
        const float v = x[i];
        float r = 0.0f;
        switch (fn) {
            case 0: r = apfm::sin(v); break;
            case 1: r = apfm::cos(v); break;
            case 2: r = apfm::exp(v); break;
            case 3: r = apfm::log(v); break;
            case 4: r = apfm::log1p(v); break;
            case 5: r = apfm::erfc(v); break;
            case 6: r = apfm::pow075(v); break;
            case 7: r = apfm::atan(v); break;
            default: r = apfm::atan2(v, x[n + i]); break;
        }
        out[i] = r;
    }
}
float or_ndtrif(float u) { return ndtrif(u); }
void or_sample_cie1931(const float* u3, size_t n, float* lambda, float* pdf) {
    for (size_t i = 0; i < n; ++i) SampleCIE1931(u3[3 * i], u3[3 * i + 1], u3[3 * i + 2], lambda + i, pdf + i);
}
float or_pdf_cie1931(float lambda) {
    const float total = 3.1f;
    return (0.1f / total) * 1.0f / 420.0f + (1.0f / total) * PdfX(lambda) + (1.0f / total) * PdfY(lambda) +
           (1.0f / total) * PdfZ(lambda);
}
float or_d65_norm(float lambda) { return D65Norm(lambda); }
void or_lambda_to_xyz(float lambda, float* xyz) {
    const float3 v = LambdaToXYZ(lambda);
    xyz[0] = v.x; xyz[1] = v.y; xyz[2] = v.z;
}
void or_xyz_to_srgb(const float* xyz, float* rgb) {
    const float3 v = XYZ2SRGB({xyz[0], xyz[1], xyz[2]});
    rgb[0] = v.x; rgb[1] = v.y; rgb[2] = v.z;
}

}  // extern "C"

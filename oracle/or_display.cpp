// ORACLE — test infrastructure only (tests/, bench.py's cpu_baseline and
// __graft_entry__.smoke() use it as the checker; nothing shipped calls it).
// CPU restatement of the reference's display and statistics framebuffers:
//   RecordShadingModel / RecordScreenGBuffer   core/denoise/denoise.cuh:84-166
//   RecordScreenStatisticBuffer, BoxCox        core/denoise/denoise_statistic.cuh:26-70
//   Tonemap kernel                             core/scene.cu:88-199
//   Gamma, ACES, XYZ2SRGBLinearD65             core/tonemap.cuh:6-126
//   HeatMapViridis                             core/render/color.cuh:99-139
// half is IEEE binary16 with round-to-nearest-even conversion (CUDA's
// __float2half_rn), written out here bit by bit. powf is the shared portable
// apfm::pow (exp(y log x)); x^0.5 in Box-Cox is sqrtf.
#include <cmath>
#include <cstdint>
#include <cstring>

#include "ap_fmath.h"
#include "or_api.h"

namespace {

// FP32 -> binary16 bits, round to nearest even; overflow to infinity.
uint16_t F2H(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    const uint16_t sign = static_cast<uint16_t>((u >> 16) & 0x8000u);
    u &= 0x7fffffffu;
    if (u > 0x7f800000u) return static_cast<uint16_t>(sign | 0x7e00u);   // NaN
    if (u >= 0x477ff000u) return static_cast<uint16_t>(sign | 0x7c00u);  // rounds past 65504
    if (u >= 0x38800000u) {                                              // normal binary16
        uint32_t r = u - (112u << 23);
        r += 0xfffu + ((r >> 13) & 1u);
        return static_cast<uint16_t>(sign | (r >> 13));
    }
    if (u <= 0x33000000u) return sign;  // at most 2^-25: zero
    const uint32_t e = u >> 23, m = (u & 0x7fffffu) | 0x800000u;
    const uint32_t sh = 126u - e;
    uint32_t q = m >> sh;
    const uint32_t rem = m & ((1u << sh) - 1u), half = 1u << (sh - 1u);
    if (rem > half || (rem == half && (q & 1u))) ++q;
    return static_cast<uint16_t>(sign | q);
}

float H2F(uint16_t h) {
    const uint32_t sign = (h & 0x8000u) << 16, e = (h >> 10) & 0x1fu, m = h & 0x3ffu;
    uint32_t u;
    if (e == 0) {
        const float v = std::ldexp(static_cast<float>(m), -24);
        std::memcpy(&u, &v, 4);
        u |= sign;
    } else if (e == 31) {
        u = sign | 0x7f800000u | (m << 13);
    } else {
        u = sign | ((e + 112u) << 23) | (m << 13);
    }
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

constexpr int kModels = 5;  // EShadingModel::MAT_MAX + 1

// RecordShadingModel (denoise.cuh:84-101).
void RecordShadingModel(uint32_t model, ap_screen_gbuffer* h) {
    if (h->frame_index == 0)
        for (int i = 0; i < kModels; ++i) h->shading_model[i] = F2H(0.0f);
    const float affixAlpha = 1.0f / static_cast<float>(h->frame_index + 1ull);
    const float preAlpha = static_cast<float>(h->frame_index) / static_cast<float>(h->frame_index + 1ull);
    for (int i = 0; i < kModels; ++i) h->shading_model[i] = F2H(H2F(h->shading_model[i]) * preAlpha);
    h->shading_model[model] = F2H(H2F(h->shading_model[model]) + affixAlpha);
}

// RecordScreenGBuffer (denoise.cuh:109-166).
void RecordScreenGBuffer(const or_gsample& s, ap_screen_gbuffer* h) {
    if (s.hit) {
        if (h->valid_frame_index == 0) {
            for (int k = 0; k < 3; ++k) h->albedo[k] = F2H(0.0f);
            for (int k = 0; k < 3; ++k) h->normal[k] = F2H(0.0f);
            h->metallic = F2H(0.0f);
            h->roughness = F2H(0.0f);
            h->depth = 0.0f;
        }
        const float affixAlpha = 1.0f / static_cast<float>(h->valid_frame_index + 1ull);
        const float preAlpha =
            static_cast<float>(h->valid_frame_index) / static_cast<float>(h->valid_frame_index + 1ull);
        for (int k = 0; k < 3; ++k) h->albedo[k] = F2H(H2F(h->albedo[k]) * preAlpha);
        for (int k = 0; k < 3; ++k) h->normal[k] = F2H(H2F(h->normal[k]) * preAlpha);
        h->metallic = F2H(H2F(h->metallic) * preAlpha);
        h->roughness = F2H(H2F(h->roughness) * preAlpha);
        h->depth = h->depth * preAlpha;
        for (int k = 0; k < 3; ++k) h->albedo[k] = F2H(H2F(h->albedo[k]) + affixAlpha * s.albedo[k]);
        for (int k = 0; k < 3; ++k) h->normal[k] = F2H(H2F(h->normal[k]) + affixAlpha * s.normal[k]);
        h->metallic = F2H(H2F(h->metallic) + affixAlpha * s.metallic);
        h->roughness = F2H(H2F(h->roughness) + affixAlpha * s.roughness);
        h->depth = h->depth + affixAlpha * s.depth;
        h->valid_frame_index++;
        RecordShadingModel(static_cast<uint32_t>(s.model), h);
    } else {
        RecordShadingModel(kModels - 1, h);
    }
    h->frame_index++;
}

// RecordScreenStatisticBuffer (denoise_statistic.cuh:36-70), per channel.
void RecordScreenStatisticBuffer(const float xyz[3], ap_screen_statistics* hist) {
    const uint32_t n0 = hist->num;
    const uint32_t n1 = n0 + 1u;
    hist->num = n1;
    float* mean = &hist->mean.x;
    float* m2 = &hist->m2_div_num.x;
    float* m3 = &hist->m3_div_num.x;
    const float invN1 = 1.0f / static_cast<float>(n1);
    for (int k = 0; k < 3; ++k) {
        const float x_bc = (std::sqrt(std::fmax(xyz[k], 0.0f)) - 1.0f) * 2.0f;  // BoxCox05(max(x, 0))
        const float delta = x_bc - mean[k];
        const float deltaDivN = delta * invN1;
        const float mean1 = mean[k] + deltaDivN;
        const double M2_0 = static_cast<double>(m2[k]) * n0;
        const double M2_1 = M2_0 + delta * (delta - deltaDivN);
        const double M3_0 = static_cast<double>(m3[k]) * n0;
        const float delta2 = delta * delta;
        const float deltaDivNPow2 = deltaDivN * deltaDivN;
        const double M3_1 = M3_0 - 3.0 * (deltaDivN * M2_1) + delta * (delta2 - deltaDivNPow2);
        mean[k] = mean1;
        m2[k] = static_cast<float>(M2_1 / n1);
        m3[k] = static_cast<float>(M3_1 / n1);
    }
}

struct V3 {
    float x, y, z;
};
float Dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
V3 Mul(const float m[3][3], V3 v) {  // float3x3::operator* (vector.cuh:346-349)
    return {Dot({m[0][0], m[0][1], m[0][2]}, v), Dot({m[1][0], m[1][1], m[1][2]}, v),
            Dot({m[2][0], m[2][1], m[2][2]}, v)};
}
float Lerp(float a, float b, float v) { return a * (1.0f - v) + b * v; }
V3 Pow(V3 c, float e) { return {apfm::pow(c.x, e), apfm::pow(c.y, e), apfm::pow(c.z, e)}; }
V3 Gamma(V3 c) { return Pow(c, 1.0f / 2.2f); }  // tonemap.cuh:6-9

// ACES (tonemap.cuh:11-113).
V3 ACES(V3 color) {
    static const float sRGB_2_AP0[3][3] = {
        {0.4397010f, 0.3829780f, 0.1773350f}, {0.0897923f, 0.8134230f, 0.0967616f}, {0.0175440f, 0.1115440f, 0.8707040f}};
    static const float AP0_2_AP1[3][3] = {{1.4514393161f, -0.2365107469f, -0.2149285693f},
                                          {-0.0765537734f, 1.1762296998f, -0.0996759264f},
                                          {0.0083161484f, -0.0060324498f, 0.9977163014f}};
    static const float AP1_2_XYZ[3][3] = {{0.6624541811f, 0.1340042065f, 0.1561876870f},
                                          {0.2722287168f, 0.6740817658f, 0.0536895174f},
                                          {-0.0055746495f, 0.0040607335f, 1.0103391003f}};
    static const float XYZ_2_AP1[3][3] = {{1.6410233797f, -0.3248032942f, -0.2364246952f},
                                          {-0.6636628587f, 1.6153315917f, 0.0167563477f},
                                          {0.0117218943f, -0.0082844420f, 0.9883948585f}};
    static const float D60_2_D65[3][3] = {
        {0.98722400f, -0.00611327f, 0.0159533f}, {-0.00759836f, 1.00186000f, 0.0053302f}, {0.00307257f, -0.00509595f, 1.0816800f}};
    static const float XYZ_2_REC709[3][3] = {{3.2409699419f, -1.5373831776f, -0.4986107603f},
                                             {-0.9692436363f, 1.8759675015f, 0.0415550574f},
                                             {0.0556300797f, -0.2039769589f, 1.0569715142f}};
    const V3 AP1_RGB2Y{0.272229f, 0.674082f, 0.0536895f};
    const V3 aces = Mul(sRGB_2_AP0, color);
    V3 acescg = Mul(AP0_2_AP1, aces);
    float tmp = Dot(acescg, AP1_RGB2Y);
    acescg = {Lerp(tmp, acescg.x, 0.96f), Lerp(tmp, acescg.y, 0.96f), Lerp(tmp, acescg.z, 0.96f)};
    const float a = 278.5085f, b = 10.7772f, c = 293.6045f, d = 88.7122f, e = 80.6889f;
    auto fit = [&](float x) { return (x * (x * a + b)) / (x * (x * c + d) + e); };
    const V3 rgbPost{fit(acescg.x), fit(acescg.y), fit(acescg.z)};
    // darkSurround_to_dimSurround (tonemap.cuh:46-65) with XYZ_2_xyY / xyY_2_XYZ (31-44).
    V3 XYZ = Mul(AP1_2_XYZ, rgbPost);
    const float divisor = std::fmax(Dot(XYZ, {1.0f, 1.0f, 1.0f}), 1.0e-4f);
    V3 xyY{XYZ.x / divisor, XYZ.y / divisor, XYZ.y};
    xyY.z = std::fmin(std::fmax(xyY.z, 0.0f), 65504.0f);
    xyY.z = apfm::pow(xyY.z, 0.9811f);
    const float m = xyY.z / std::fmax(xyY.y, 1e-4f);
    XYZ = {xyY.x, xyY.z, (1.0f - xyY.x - xyY.y)};
    XYZ.x *= m;
    XYZ.z *= m;
    V3 linearCV = Mul(XYZ_2_AP1, XYZ);
    tmp = Dot(linearCV, AP1_RGB2Y);
    linearCV = {Lerp(tmp, linearCV.x, 0.93f), Lerp(tmp, linearCV.y, 0.93f), Lerp(tmp, linearCV.z, 0.93f)};
    XYZ = Mul(AP1_2_XYZ, linearCV);
    XYZ = Mul(D60_2_D65, XYZ);
    const V3 r = Mul(XYZ_2_REC709, XYZ);
    return Gamma({std::fmax(0.0f, r.x), std::fmax(0.0f, r.y), std::fmax(0.0f, r.z)});
}

V3 XYZ2SRGBLinearD65(V3 c) {  // tonemap.cuh:116-126
    static const float M[3][3] = {{3.2406f, -1.5372f, -0.4986f}, {-0.9689f, 1.8758f, 0.0415f}, {0.0557f, -0.204f, 1.057f}};
    const V3 r = Mul(M, c);
    return {std::fmax(0.0f, r.x), std::fmax(0.0f, r.y), std::fmax(0.0f, r.z)};
}

V3 HeatMapViridis(float t) {  // color.cuh:99-139
    t = std::fmin(std::fmax(t, 0.0f), 1.0f);
    static const V3 c[5] = {{0.267004f, 0.004874f, 0.329415f}, {0.229739f, 0.322361f, 0.545706f},
                            {0.127568f, 0.566949f, 0.550556f}, {0.369214f, 0.788888f, 0.382914f},
                            {0.993248f, 0.906157f, 0.143936f}};
    const float a = t * 4.0f;
    const int i = static_cast<int>(std::floor(a));
    const float f = a - static_cast<float>(i);
    const int k = i <= 0 ? 0 : (i == 1 ? 1 : (i == 2 ? 2 : 3));
    return {c[k].x + (c[k + 1].x - c[k].x) * f, c[k].y + (c[k + 1].y - c[k].y) * f, c[k].z + (c[k + 1].z - c[k].z) * f};
}

float Saturate(float v) { return std::isnan(v) ? 0.0f : std::fmin(std::fmax(v, 0.0f), 1.0f); }  // vector.cuh:87-94
uint32_t Channel(float v) { return static_cast<uint32_t>(255.0f * Saturate(v)); }
uint32_t ARGB(V3 c) { return 0xff000000u | (Channel(c.x) << 16) | (Channel(c.y) << 8) | Channel(c.z); }

}  // namespace

extern "C" {

void or_half_round(const float* x, size_t n, uint16_t* out) {
    for (size_t i = 0; i < n; ++i) out[i] = F2H(x[i]);
}

void or_record_gbuffer(ap_screen_gbuffer* g, const or_gsample* samples, size_t n) {
    for (size_t i = 0; i < n; ++i) RecordScreenGBuffer(samples[i], g);
}

void or_record_statistics(ap_screen_statistics* h, const float* xyz, size_t n) {
    for (size_t i = 0; i < n; ++i) RecordScreenStatisticBuffer(xyz + 3 * i, h);
}

// The Tonemap kernel (scene.cu:88-199) over n pixels; display as ap_tonemap's.
void or_tonemap(const float* hdr, const ap_screen_gbuffer* gb, const ap_screen_statistics* st, uint32_t* ldr,
                size_t n, int tone, int display, float exposure) {
    for (size_t id = 0; id < n; ++id) {
        uint32_t out;
        switch (display) {
            case AP_DISPLAY_ALBEDO:
                out = ARGB({H2F(gb[id].albedo[0]), H2F(gb[id].albedo[1]), H2F(gb[id].albedo[2])});
                break;
            case AP_DISPLAY_NORMAL:
                out = ARGB({H2F(gb[id].normal[0]) * 0.5f + 0.5f, H2F(gb[id].normal[1]) * 0.5f + 0.5f,
                            H2F(gb[id].normal[2]) * 0.5f + 0.5f});
                break;
            case AP_DISPLAY_DEPTH: out = ARGB(HeatMapViridis(1.0f / (gb[id].depth + 1.0f))); break;
            case AP_DISPLAY_METALLIC: out = ARGB(HeatMapViridis(H2F(gb[id].metallic))); break;
            case AP_DISPLAY_ROUGHNESS: out = ARGB(HeatMapViridis(H2F(gb[id].roughness))); break;
            case AP_DISPLAY_STATISTICS_NUM: {
                const float heat = static_cast<float>(std::log(static_cast<double>(st[id].num) + 1.0)) * 0.01f;
                out = 0xff000000u | (Channel(heat) << 16);
                break;
            }
            case AP_DISPLAY_STATISTICS_MEAN:
            case AP_DISPLAY_STATISTICS_M2:
            case AP_DISPLAY_STATISTICS_M3: {
                const ap_float3 v = display == AP_DISPLAY_STATISTICS_MEAN
                                        ? st[id].mean
                                        : (display == AP_DISPLAY_STATISTICS_M2 ? st[id].m2_div_num : st[id].m3_div_num);
                out = ARGB({std::fabs(v.x) * 0.25f, std::fabs(v.y) * 0.25f, std::fabs(v.z) * 0.25f});
                break;
            }
            default: {
                const V3 result = XYZ2SRGBLinearD65({hdr[3 * id], hdr[3 * id + 1], hdr[3 * id + 2]});
                V3 value{result.x * exposure, result.y * exposure, result.z * exposure};
                if (tone == AP_TONE_GAMMA) value = Gamma(value);
                else if (tone == AP_TONE_ACES) value = ACES(value);
                out = ARGB(value);
            }
        }
        ldr[id] = out;
    }
}

}  // extern "C"

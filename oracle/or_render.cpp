// ORACLE — test infrastructure only: CPU restatement of the reference's
// per-pixel integrator, the checker for the core's wavefront path tracer
// (tests/ and bench.py's cpu_baseline leg; the product never links it).
//
//   RenderCamera (pinhole, running means)     core/scene.cu:201-255
//   light table (area * luminance)             core/scene.cu:1056-1080
//   Trace (closest hit, material, offset)       core/render/render.cuh:75-163
//   NEE / SegmentTrace / PowerHeuristic         core/render/render.cuh:165-343
//   CalculateRadiance (path loop)              core/render/render.cuh:460-999
//   EvaluateMaterial / SampleBSDF / EvaluateBSDFWithCos / BSDF_PDF
//                                              core/render/material.cuh:48-425
//   brdf / sample helpers                      core/render/brdf.cuh, sample.cuh
//   NEESampleLight, box / sphere samplers, NEEPDFOmegaFromHit
//                                              core/nee/nee.cuh:24-377
// Deviations shared with the core, each a parameter or a documented
// restriction: PCG32 replaces curand (stream per pixel, start state hashed
// from seed, frame and pixel with SplitMix64); rsqrtf is 1/sqrtf; triangles may carry their object's
// material (param) instead of IntersectTriangleObject's default GBuffer;
// textures are not restated.
// Glass and media (material.cuh:12-18, 122-169, 221-272, render.cuh:84-160):
// MediaContext, the solo trace inside a medium, the miss inside a medium
// (which restarts at the world origin, as the reference's zero-initialised
// hit record makes it), the exact dielectric with Abbe dispersion and
// tilted Beer-Lambert absorption; with hero wavelengths the secondaries end
// at a dispersive interface (HeroOnlyAtDispersion).
// Directional path guiding (or_guiding.h) follows render.cuh:546-639 and
// scene.cu:308-386.
// The learned radiance cache has no reference counterpart: CalculateRadianceNRC
// restates the core's cache termination (spread heuristic, training paths,
// training targets; include/aperture/ap_nrc.h) with the network evaluated by
// or_nrc.cpp's FP64 forward with BF16 rounding points.
// Parity pins: white-furnace and analytic-irradiance scenes
// (tests/test_render_cpu.py), NEE pdf pairing and power-heuristic identities;
// guiding and cache pins in tests/test_guiding_cpu.py and tests/test_nrc_cpu.py.
#include <cmath>
#include <cstring>
#include <limits>
#include <unordered_map>
#include <vector>

#include "aperture/ap_render.h"
#include "or_api.h"
#include "ap_fmath.h"
#include "or_math.h"
#include "or_spectrum.h"
#include "or_guiding.h"
#include "or_count.h"

namespace orc {

// Planted mistakes for the tests that show a pin fails on a wrong formula
// (or_plant; 0 = none). Each one is a plausible transcription slip:
//   1: the bounded-VNDF pdf loses its factor 2 (sampler and pdf alike);
//   2: the sphere-cap area is pi r^2 (1 - cosC) (sampler and pdf alike);
//   3: the GGX mixture pdf of a specular draw drops the Lambert term;
//   4: the GGX NDF takes roughness where it takes alpha (D(alpha) -> D(sqrt alpha));
//   5: the bounded-VNDF sampler skips normalising the stretched view vector.
static int g_plant = 0;

// ---- PCG32 (O'Neill 2014, pcg32_random_r / pcg32_srandom_r) ----
struct Rng {
    uint64_t state, inc;
    uint32_t next() {
        const uint64_t old = state;
        state = old * 6364136223846793005ULL + inc;
        const uint32_t xs = static_cast<uint32_t>(((old >> 18u) ^ old) >> 27u);
        const uint32_t rot = static_cast<uint32_t>(old >> 59u);
        return (xs >> rot) | (xs << ((32u - rot) & 31u));
    }
    float rand1() { return static_cast<float>(next() >> 8) * 0x1p-24f; }
    static Rng seeded(uint64_t initstate, uint64_t initseq) {
        Rng r{0u, (initseq << 1u) | 1u};
        r.next();
        r.state += initstate;
        r.next();
        return r;
    }
};

// SplitMix64 finaliser (Steele, Lea, Flood 2014).
static inline uint64_t SplitMix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ULL;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
    return x ^ (x >> 31);
}

// Stream of one path sample: stream = pixel, start state hashed from
// (seed, frame, pixel) so that consecutive frames are decorrelated.
static inline Rng PathRng(uint64_t pixel, uint64_t frame, uint64_t seed) {
    const uint64_t k = SplitMix64(seed ^ SplitMix64(frame));
    return Rng::seeded(SplitMix64(k ^ pixel), pixel);
}

struct TraceResult2 { int id; float t; };

// From or_bvh.cpp.
extern bool IntersectTriangleT(float3 o, float3 d, const ap_scene_object& ob, float& t, float& uu, float& vv);
extern bool IntersectAABBT(float3 o, float3 d, const ap_scene_object& ob, float& t, bool& inside);
extern bool IntersectSphereT(float3 o, float3 d, const ap_scene_object& ob, float& t, bool& inside);

inline float3 F3(const ap_float3& a) { return {a.x, a.y, a.z}; }
inline float3x3 M3(const ap_float3x3& m) { return {F3(m.x), F3(m.y), F3(m.z)}; }
inline float INTERSECT_EPS(int depth) { return 1.0e-3f * std::exp(std::fmin(static_cast<float>(depth) * 0.5f, 4.0f)); }

// GBuffer subset the integrator reads.
struct GB {
    float3 albedo, emissive;
    float specular, metallic, roughness;
    float ior, sigmaA, abbe, tilt;
    int albedoTex = -1;  // texture table index of albedoTex_, or -1
};

// MediaContext (material.cuh:12-18).
struct MediaCtx {
    bool inside = false;
    float3 vertex{0, 0, 0};
    int object = -1;
    int depth = 0;
};

struct Ctx {  // IntersectionContext subset
    bool hit = false;
    int objectIndex = -1, lightIndex = -1;
    float distance = 0.0f;
    float3 pos{0, 0, 0}, normal{0, 0, 0};
    GB gb{};
    float u = 0.0f, v = 0.0f;  // uv_
};

// ---- albedo textures (LoadCUDATexture's descriptor, texture_loader.cuh:36-46;
// tex2D<float4> with the linear filtering the CUDA programming guide
// documents: x_B = frac(u) W - 1/2, weights held with 8 fractional bits,
// rounded to nearest here) ----
struct OrTex {
    std::vector<uint32_t> texels;  // RGBA8, R in the low byte, rows from the top
    int w = 0, h = 0;
};
static std::vector<OrTex> g_textures;

static float TexChannel(uint32_t t, int c) { return static_cast<float>((t >> (8 * c)) & 0xffu) / 255.0f; }

static float TexWrap(float u) {
    const float f = u - std::floor(u);
    return (f >= 0.0f && f <= 1.0f) ? f : 0.0f;
}

static void TexSample(const OrTex& T, float u, float v, float out[4]) {
    const float x = TexWrap(u) * static_cast<float>(T.w) - 0.5f;
    const float y = TexWrap(v) * static_cast<float>(T.h) - 0.5f;
    const float fx = std::floor(x), fy = std::floor(y);
    const float a = std::floor((x - fx) * 256.0f + 0.5f) * (1.0f / 256.0f);
    const float b = std::floor((y - fy) * 256.0f + 0.5f) * (1.0f / 256.0f);
    int i0 = static_cast<int>(fx), j0 = static_cast<int>(fy);
    if (i0 < 0) i0 += T.w;
    if (j0 < 0) j0 += T.h;
    const int i1 = (i0 + 1) % T.w, j1 = (j0 + 1) % T.h;
    const uint32_t t00 = T.texels[static_cast<size_t>(j0) * T.w + i0];
    const uint32_t t10 = T.texels[static_cast<size_t>(j0) * T.w + i1];
    const uint32_t t01 = T.texels[static_cast<size_t>(j1) * T.w + i0];
    const uint32_t t11 = T.texels[static_cast<size_t>(j1) * T.w + i1];
    const float w00 = (1.0f - a) * (1.0f - b), w10 = a * (1.0f - b), w01 = (1.0f - a) * b, w11 = a * b;
    for (int c = 0; c < 4; ++c)
        out[c] = w00 * TexChannel(t00, c) + w10 * TexChannel(t10, c) + w01 * TexChannel(t01, c) +
                 w11 * TexChannel(t11, c);
}

// The GBuffer copy SampleBSDF and EvaluateBSDFWithCos shade with
// (material.cuh:69-74, 198-202): albedo = texel.rgb^2 at uv_.
static Ctx Textured(const Ctx& c) {
    Ctx t = c;
    if (c.gb.albedoTex >= 0 && c.gb.albedoTex < static_cast<int>(g_textures.size())) {
        float s[4];
        TexSample(g_textures[c.gb.albedoTex], c.u, c.v, s);
        t.gb.albedo = {s[0] * s[0], s[1] * s[1], s[2] * s[2]};
    }
    return t;
}

struct Scene {
    const ap_scene_object* objs;
    size_t n;
    const or_bvh* bvh;
    std::vector<ap_light_info> lights;
    std::vector<int> light_index;  // per object
    int tri_mode;
    float margin;                  // > 0: the core's CULLED traversal (same results, faster)
};

// The wavelength response cache of Interact: LUT cells computed on demand.
struct Basis { float v[4]; };
static thread_local std::unordered_map<uint64_t, Basis>* tl_cells = nullptr;

static Basis Cell(int x, int y, int z) {
    if (!tl_cells) tl_cells = new std::unordered_map<uint64_t, Basis>();
    const uint64_t key = (static_cast<uint64_t>(z) << 20) | (static_cast<uint64_t>(y) << 10) | static_cast<uint64_t>(x);
    auto it = tl_cells->find(key);
    if (it != tl_cells->end()) return it->second;
    Basis b;
    const float inv = 1.0f / 512.0f;
    const float in3[3] = {(x + 0.5f) * inv, (y + 0.5f) * inv, (z + 0.5f) * inv};
    RGBEncode(in3, b.v);
    tl_cells->emplace(key, b);
    return b;
}

static void QueryRGBBasisCached(float3 rgb, float out[4]) {
    const int R = 512;
    float g[3] = {rgb.x * R - 0.5f, rgb.y * R - 0.5f, rgb.z * R - 0.5f};
    int i0[3], i1[3];
    float t[3];
    for (int a = 0; a < 3; ++a) {
        g[a] = std::fmin(std::fmax(g[a], 0.0f), R - 1.0f);
        const float fi0 = std::floor(g[a]);
        i0[a] = static_cast<int>(fi0);
        i1[a] = i0[a] + 1;
        if (i1[a] >= R) i1[a] = R - 1;
        t[a] = g[a] - fi0;
        if (i0[a] == i1[a]) t[a] = 0.0f;
    }
    const float wx0 = 1.0f - t[0], wx1 = t[0], wy0 = 1.0f - t[1], wy1 = t[1], wz0 = 1.0f - t[2], wz1 = t[2];
    const float w[8] = {wx0 * wy0 * wz0, wx1 * wy0 * wz0, wx0 * wy1 * wz0, wx1 * wy1 * wz0,
                        wx0 * wy0 * wz1, wx1 * wy0 * wz1, wx0 * wy1 * wz1, wx1 * wy1 * wz1};
    const int cx[8] = {i0[0], i1[0], i0[0], i1[0], i0[0], i1[0], i0[0], i1[0]};
    const int cy[8] = {i0[1], i0[1], i1[1], i1[1], i0[1], i0[1], i1[1], i1[1]};
    const int cz[8] = {i0[2], i0[2], i0[2], i0[2], i1[2], i1[2], i1[2], i1[2]};
    for (int k = 0; k < 4; ++k) out[k] = 0.0f;
    for (int c = 0; c < 8; ++c) {
        const Basis b = Cell(cx[c], cy[c], cz[c]);
        for (int k = 0; k < 4; ++k) out[k] = std::fma(w[c], b.v[k], out[k]);
    }
}

static float InteractC(const float lb[4], float3 rgb) {  // lut::Interact
    Count(CNT_INTERACTS);
    const float lum = max3(rgb);
    float rb[4];
    if (lum > 1.0f) {
        QueryRGBBasisCached(rgb / lum, rb);
        return SpectralInteract(rb, lb) * lum;
    }
    QueryRGBBasisCached(rgb, rb);
    return SpectralInteract(rb, lb);
}

// ---- sampling helpers (vector.cuh, sample.cuh) ----
static void BuildTangentBasis(float3 n, float3& t, float3& b) {
    const float3 a = (std::fabs(n.z) < 0.999f) ? float3{0, 0, 1} : float3{1, 0, 0};
    t = normalize(cross(a, n));
    b = cross(n, t);
}
static void BuildTangentBasisRandom(float3 n, float3& t, float3& b, float e) {
    const float3 a = (std::fabs(n.z) < 0.999f) ? float3{0, 0, 1} : float3{1, 0, 0};
    const float3 t0 = normalize(cross(a, n));
    const float3 b0 = cross(n, t0);
    const float s = apfm::sin(e * PI_F * 2.0f), c = apfm::cos(e * PI_F * 2.0f);
    t = c * t0 + s * b0;
    b = -s * t0 + c * b0;
}
static float2 ConcentricSampleDisk(float ux, float uy) {
    const float sx = 2.0f * ux - 1.0f, sy = 2.0f * uy - 1.0f;
    if (sx == 0.0f && sy == 0.0f) return {0.0f, 0.0f};
    float r, phi;
    if (std::fabs(sx) > std::fabs(sy)) { r = sx; phi = (PI_F / 4.0f) * (sy / std::fmax(std::fabs(sx), 1e-20f)); }
    else { r = sy; phi = (PI_F / 2.0f) - (PI_F / 4.0f) * (sx / std::fmax(std::fabs(sy), 1e-20f)); }
    return {r * apfm::cos(phi), r * apfm::sin(phi)};
}
struct Sample { float3 dir; float pdf; };
static Sample SampleCosine(float3 e, float3 n) {
    float3 T, B;
    BuildTangentBasisRandom(n, T, B, e.z);
    const float2 d = ConcentricSampleDisk(e.x, e.y);
    const float z = std::sqrt(std::fmax(0.0f, 1.0f - d.x * d.x - d.y * d.y));
    const float3 dir = d.x * T + d.y * B + z * n;
    return {normalize(dir), z / PI_F};
}
static float GGX_D(float alpha, float NoH) {
    if (g_plant == 4) alpha = std::sqrt(alpha);
    const float a2 = alpha * alpha;
    const float d = NoH * NoH * (a2 - 1.0f) + 1.0f;
    return a2 / std::fmax(PI_F * d * d, 1e-12f);
}
static Sample SampleGGXBoundedVNDF(float3 V, float3 N, float alpha, float3 rnd) {
    float3 X, Y;
    BuildTangentBasisRandom(N, X, Y, rnd.z);
    const float XoV = dot(X, V), YoV = dot(Y, V), NoV = dot(N, V);
    const float3 lv{XoV, YoV, NoV};
    const float NoV2 = std::fmin(1.0f, NoV * NoV);
    const float a2 = alpha * alpha;
    const float3 vs = float3{XoV * alpha, YoV * alpha, NoV} / (g_plant == 5 ? 1.0f : std::sqrt(a2 + (1.0f - a2) * NoV2));
    const float phi = 2.0f * PI_F * rnd.x;
    const float s = 1.0f + std::sqrt(1.0f - NoV2);
    const float s2 = s * s;
    const float k = (1.0f - a2) * s2 / (s2 + a2 * NoV2);
    const float b = NoV > 0 ? k * vs.z : vs.z;
    const float z = std::fma(1.0f - rnd.y, 1.0f + b, -b);
    const float sinT = std::sqrt(saturate_(1.0f - z * z));
    const float3 os{sinT * apfm::cos(phi), sinT * apfm::sin(phi), z};
    const float3 ms = vs + os;
    const float3 H = normalize(float3{ms.x * alpha, ms.y * alpha, ms.z});
    const float3 lL = reflect(-lv, H);
    const float3 dir = lL.x * X + lL.y * Y + lL.z * N;
    const float len2 = (1.0f - NoV2) * a2;
    const float t = std::sqrt(len2 + NoV2);
    return {dir, GGX_D(alpha, H.z) / ((g_plant == 1 ? 1.0f : 2.0f) * (k * NoV + t))};
}
static float GGXBoundedVNDF_PDF(float3 V, float3 N, float3 L, float alpha) {
    const float NoV = dot(N, V), NoL = dot(N, L);
    if (NoV <= 0.0f || NoL <= 0.0f) return 0.0f;
    const float3 H = normalize(V + L);
    const float NoV2 = std::fmin(1.0f, NoV * NoV);
    const float a2 = alpha * alpha;
    const float s = 1.0f + std::sqrt(1.0f - NoV2);
    const float s2 = s * s;
    const float k = (1.0f - a2) * s2 / (s2 + a2 * NoV2);
    const float NoH = std::fmax(0.0f, dot(N, H));
    const float len2 = (1.0f - NoV2) * a2;
    const float t = std::sqrt(len2 + NoV2);
    return std::fmax(GGX_D(alpha, NoH) / ((g_plant == 1 ? 1.0f : 2.0f) * (k * NoV + t)), 1e-12f);
}
static float FresnelSchlick(float F0, float cosX) {
    const float c = clampf(1.0f - cosX, 0.0f, 1.0f);
    const float c5 = c * c * c * c * c;
    return F0 + (1.0f - F0) * c5;
}
static float GGX_G_Correlated(float alpha, float NoL, float NoV) {
    const float a2 = alpha * alpha;
    const float gv = NoL * std::sqrt(a2 + (1.0f - a2) * NoV * NoV);
    const float gl = NoV * std::sqrt(a2 + (1.0f - a2) * NoL * NoL);
    const float denom = gv + gl;
    return (denom > 0.0f) ? (2.0f * NoL * NoV) / denom : 0.0f;
}
static float F0_UEStyle(float s) { return 0.08f * clampf(s, 0.0f, 1.0f); }

// ---- glass: dispersion and absorption (spectrum.cuh:7-57), Beer-Lambert
// and the exact dielectric Fresnel term (brdf.cuh:20-32), refract
// (vector.cuh:267-275) ----
static constexpr float kWaveF = 486.1327f, kWaveD = 587.5618f, kWaveC = 656.2725f;
static constexpr float kInvF = 1.0f / kWaveF, kInvC = 1.0f / kWaveC, kInvD = 1.0f / kWaveD;

static float IORLinearKFromAbbe(float nd, float Vd) {
    if (!(Vd > 0.0f)) return 0.0f;
    return (nd - 1.0f) / (Vd * (kInvF - kInvC));
}

static float IORFromAbbeNM(float lambda, float nd, float Vd) {
    if (!(lambda > 0.0f)) return nd;
    const float K = IORLinearKFromAbbe(nd, Vd);
    const float invL = 1.0f / lambda;
    return nd + K * (invL - kInvD);
}

static float SigmaAFromTiltNM(float sigmaA, float lambda, float tilt) {
    const float u = 1.0f / std::fmax(lambda, 1.0e-6f);
    const float du = (u - kInvD) / (kInvF - kInvC);
    const float sA = sigmaA * (1.0f + 1.0f * tilt * du);
    return std::fmax(0.0f, sA);
}

static float BeerLambert_Spectrum(float distance, float resp, float sigmaA) {
    return apfm::exp(distance * sigmaA * apfm::log(std::fmin(std::fmax(resp, 1.0e-12f), 1.0f)));
}

static float FresnelDielectricExact(float cosI, float cosT, float eta) {
    const float invE = 1.0f / eta;
    const float rPar = (invE * cosI - cosT) / (invE * cosI + cosT);
    const float rPerp = (cosI - invE * cosT) / (cosI + invE * cosT);
    return 0.5f * (rPar * rPar + rPerp * rPerp);
}

static float3 Refract(float3 inDir, float3 normal, float eta) {
    const float3 n = (dot(normal, inDir) > 0.0f) ? -normal : normal;
    const float cosI = -dot(n, inDir);
    const float k = 1.0f - (eta * eta) * std::fmax(0.0f, 1.0f - cosI * cosI);
    const float cosT = std::sqrt(std::fmax(0.0f, k));
    return eta * inDir + (eta * cosI - cosT) * n;
}

// ---- materials (material.cuh) ----
// m: the MediaContext the ray was traced in (glass only).
static float EvaluateBSDFWithCos(const float lb[4], float lambda, int model, float3 rayDir, float3 L, const Ctx& c,
                                 const MediaCtx& m) {
    const float3 N = c.normal;
    const float NoL = std::fmax(0.0f, dot(N, L));
    if (model == AP_MAT_DIFFUSE) return InteractC(lb, c.gb.albedo) * (NoL / PI_F);
    if (model == AP_MAT_GLASS) {
        const float IOR = IORFromAbbeNM(lambda, c.gb.ior, c.gb.abbe);
        float transmittance = 1.0f;
        if (m.inside) {
            const float rayDistance = length(m.vertex - c.pos);
            transmittance = BeerLambert_Spectrum(rayDistance, InteractC(lb, c.gb.albedo),
                                                 SigmaAFromTiltNM(c.gb.sigmaA, lambda, c.gb.tilt));
        }
        const float3 nIn = (dot(N, rayDir) > 0.0f) ? -N : N;
        const float eta = m.inside ? IOR : 1.0f / IOR;
        const float cosI = -dot(nIn, rayDir);
        const float sin2I = std::fmax(0.0f, 1.0f - cosI * cosI);
        const float k = 1.0f - sin2I * (eta * eta);
        if (k <= 0.0f) return transmittance * 0.99f;  // PERFECT_DELTA_SURFACE
        const float cosT = std::sqrt(k);
        const float Fr = FresnelDielectricExact(cosI, cosT, eta);
        const float3 R = reflect(rayDir, nIn);
        const float3 T = Refract(rayDir, nIn, eta);
        const float dR = dot(L, R);
        const float dT = dot(L, T);
        if (dR >= dT) return transmittance * Fr;
        const float weight = (eta * eta) * (std::fabs(cosT) / std::fmax(1e-6f, cosI));
        return transmittance * (1.0f - Fr) * weight;
    }
    if (model == AP_MAT_GGX) {
        const float3 V = normalize(-rayDir);
        const float NoV = std::fmax(0.0f, dot(N, V));
        if (NoV <= 0.0f || NoL <= 0.0f) return 0.0f;
        const float alpha = c.gb.roughness * c.gb.roughness;
        const float al = InteractC(lb, c.gb.albedo);
        const float F0 = lerp(F0_UEStyle(c.gb.specular), al, c.gb.metallic);
        const float kD = 1.0f - c.gb.metallic;
        const float3 Hn = normalize(V + L);
        const float NoH = std::fmax(0.0f, dot(N, Hn));
        const float LoH = std::fmax(0.0f, dot(L, Hn));
        const float Fr = FresnelSchlick(F0, LoH);
        const float D = GGX_D(alpha, NoH);
        const float G = GGX_G_Correlated(alpha, NoL, NoV);
        const float spec = (NoL > 0.0f) ? (Fr * D * G) / std::fmax(4.0f * NoV, 1e-7f) : 0.0f;
        const float diff = kD * (1.0f - Fr) * al * (NoL / PI_F);
        return spec + diff;
    }
    return 0.0f;
}

static float BSDF_PDF(const float lb[4], int model, const GB& gb, float3 N, float3 rayDir, float3 L) {
    const float3 V = -rayDir;
    const float NoL = std::fmax(0.0f, dot(N, L));
    if (model == AP_MAT_DIFFUSE) return NoL > 0.0f ? (NoL * (1.0f / PI_F)) : 0.0f;
    if (model == AP_MAT_GGX) {
        const float NoV = std::fmax(0.0f, dot(N, V));
        if (NoV <= 0.0f || NoL <= 0.0f) return 0.0f;
        const float alpha = gb.roughness * gb.roughness;
        const float F0 = lerp(F0_UEStyle(gb.specular), InteractC(lb, gb.albedo), gb.metallic);
        const float pSpec = clampf(F0, 0.1f, 0.9f);
        const float pdfMix = pSpec * GGXBoundedVNDF_PDF(V, N, L, alpha) + (1.0f - pSpec) * (NoL * (1.0f / PI_F));
        return std::fmax(0.0f, pdfMix);
    }
    return 0.0f;
}

struct SampleCtx {
    float3 next{0, 0, 0};
    float pdf = 1.0f;
    float3 emissive{0, 0, 0};
    float brdf = 0.0f;
    bool terminate = false, pureLight = false, delta = false;
    bool inside = false;
};

static SampleCtx SampleBSDF(const float lb[4], float lambda, Rng& rng, int model, float3 rayDir, const Ctx& c,
                            const MediaCtx& m) {
    SampleCtx r;
    const float3 N = c.normal;
    const float3 V = -rayDir;
    const bool inside = m.inside;
    if (model == AP_MAT_GLASS) {
        const float IOR = IORFromAbbeNM(lambda, c.gb.ior, c.gb.abbe);
        r.delta = true;
        const float eta = inside ? IOR : 1.0f / IOR;
        const float3 nIn = (dot(N, rayDir) > 0.0f) ? -N : N;
        const float cosI = -dot(nIn, rayDir);
        const float sin2I = std::fmax(0.0f, 1.0f - cosI * cosI);
        const float k = 1.0f - sin2I * (eta * eta);
        r.emissive = {0, 0, 0};
        if (k <= 0.0f) {  // total internal reflection
            if (m.depth < 1024) {  // MAX_DELTA_SURFACE_DEPTH
                r.next = reflect(rayDir, nIn);
                r.inside = inside;
            } else {  // forced out
                r.next = rayDir;
                r.inside = !inside;
            }
            r.pdf = 1.0f;
            return r;
        }
        const float cosT = std::sqrt(k);
        const float Fr = FresnelDielectricExact(cosI, cosT, eta);
        if (rng.rand1() < Fr) {
            r.next = reflect(rayDir, nIn);
            r.pdf = Fr;
            r.inside = inside;
        } else {
            r.next = Refract(rayDir, nIn, eta);
            r.pdf = 1.0f - Fr;
            r.inside = !inside;
        }
        return r;
    }
    if (model == AP_MAT_DIFFUSE) {
        const float e0 = rng.rand1(), e1 = rng.rand1(), e2 = rng.rand1();
        const Sample s = SampleCosine({e0, e1, e2}, N);
        r.next = s.dir; r.pdf = s.pdf; r.emissive = c.gb.emissive; r.inside = inside;
        return r;
    }
    if (model == AP_MAT_LIGHT) {
        r.pdf = 1.0f; r.emissive = c.gb.emissive; r.terminate = true; r.pureLight = true;
        return r;
    }
    // GGX
    r.emissive = c.gb.emissive;
    const float NoV = std::fmax(0.0f, dot(N, V));
    if (NoV <= 0.0f) { r.terminate = true; return r; }
    const float alpha = c.gb.roughness * c.gb.roughness;
    const float F0 = lerp(F0_UEStyle(c.gb.specular), InteractC(lb, c.gb.albedo), c.gb.metallic);
    const float pSpec = clampf(F0, 0.1f, 0.9f);
    const float u = rng.rand1();
    if (u < pSpec) {
        const float e0 = rng.rand1(), e1 = rng.rand1(), e2 = rng.rand1();
        const Sample s = SampleGGXBoundedVNDF(V, N, alpha, {e0, e1, e2});
        const float NoL = std::fmax(0.0f, dot(N, s.dir));
        if (NoL <= 0.0f) { r.terminate = true; return r; }
        r.next = s.dir;
        r.pdf = pSpec * s.pdf + (g_plant == 3 ? 0.0f : (1.0f - pSpec) * (NoL / PI_F));
    } else {
        const float e0 = rng.rand1(), e1 = rng.rand1(), e2 = rng.rand1();
        const Sample s = SampleCosine({e0, e1, e2}, N);
        r.next = s.dir;
        r.pdf = pSpec * GGXBoundedVNDF_PDF(V, N, s.dir, alpha) + (1.0f - pSpec) * s.pdf;
    }
    r.inside = inside;
    return r;
}

// ---- intersection context (intersection.cuh, aabb.cuh, sphere.cuh) ----
extern bool IntersectSDFT(float3 o, float3 d, const ap_scene_object& ob, float eps, float& t, float3* localPos);
extern float3 SDFLocalNormal(const ap_scene_object& ob, float3 lp, float eps);
extern bool OverlapSDF(const ap_scene_object& ob, float3 p);

// sdfEps: SDF_EPS of the ray's depth (IntersectSDFObject's normal step).
static Ctx HitContext(const Scene& S, int idx, float3 o, float3 d, float t, float sdfEps) {
    Ctx c;
    const ap_scene_object& ob = S.objs[idx];
    c.hit = true;
    c.objectIndex = idx;
    c.lightIndex = S.light_index[idx];
    c.distance = t;
    const ap_gbuffer& g = ob.material.texture.base;
    c.gb = {F3(g.albedo), F3(g.emissive), g.specular, g.metallic, g.roughness,
            g.ior, g.sigma_a, g.abbe_number, g.tilt_number};
    const int tex = g.albedo_tex ? static_cast<int>(g.albedo_tex - 1u) : -1;
    c.gb.albedoTex = tex;
    if (ob.type == AP_OBJ_TRIANGLE) {
        float tt, u, v;
        IntersectTriangleT(o, d, ob, tt, u, v);
        const float w = 1.0f - u - v;
        const float3x3 o2w = M3(ob.object_to_world);
        const float3 n0 = o2w * F3(ob.additional.v0.normal);
        const float3 n1 = o2w * F3(ob.additional.v1.normal);
        const float3 n2 = o2w * F3(ob.additional.v2.normal);
        const float3 n = normalize(n0 * w + n1 * u + n2 * v);
        c.pos = o + d * t;
        c.normal = dot(d, n) > 0.0f ? -n : n;
        // uv_ (intersection.cuh:63-67).
        c.u = ob.additional.v0.uv.x * w + ob.additional.v1.uv.x * u + ob.additional.v2.uv.x * v;
        c.v = ob.additional.v0.uv.y * w + ob.additional.v1.uv.y * u + ob.additional.v2.uv.y * v;
        if (S.tri_mode == AP_TRIANGLE_MATERIAL_REFERENCE) {  // a default GBuffer but the albedo texture (:58-61)
            c.gb = {{1, 1, 1}, {0, 0, 0}, 0.0f, 0.0f, 0.0f, 1.5f, 100.0f, 20.0f, 0.0f};
            c.gb.albedoTex = tex;
        }
        return c;
    }
    const float3x3 w2o = M3(ob.world_to_object), o2w = M3(ob.object_to_world);
    const float3 lo = w2o * (o - F3(ob.center));
    const float3 ld = normalize(w2o * d);
    const float3 ext = F3(ob.extent);
    if (ob.type == AP_OBJ_SDF) {  // IntersectSDFObject (sdf_base.cuh:138-162)
        float th = t;
        float3 lp = lo + ld * t;
        (void)IntersectSDFT(o, d, ob, sdfEps, th, &lp);
        const float3 ln = SDFLocalNormal(ob, lp, sdfEps);
        c.pos = o2w * lp + F3(ob.center);
        c.normal = normalize(o2w * ln);
        return c;
    }
    if (ob.type == AP_OBJ_CUBE) {
        const float3 div = invSafe(ld);
        const float3 t1 = (-ext - lo) * div, t2 = (ext - lo) * div;
        const float3 mn = fmin3(t1, t2), mx = fmax3(t1, t2);
        const float tE = std::fmax(std::fmax(mn.x, mn.y), mn.z);
        const float tX = std::fmin(std::fmin(mx.x, mx.y), mx.z);
        const bool inside = tE < 0.0f;
        const float tHit = inside ? tX : tE;
        const float3 lp = lo + ld * tHit;
        const float rx = std::fabs(lp.x) / ext.x, ry = std::fabs(lp.y) / ext.y, rz = std::fabs(lp.z) / ext.z;
        float3 ln;
        if (rx >= ry && rx >= rz) ln = {(lp.x > 0.0f) ? 1.0f : -1.0f, 0.0f, 0.0f};
        else if (ry >= rz) ln = {0.0f, (lp.y > 0.0f) ? 1.0f : -1.0f, 0.0f};
        else ln = {0.0f, 0.0f, (lp.z > 0.0f) ? 1.0f : -1.0f};
        if (inside) ln = -ln;
        c.pos = o2w * lp + F3(ob.center);
        c.normal = normalize(o2w * ln);
        return c;
    }
    const float r = max3(ext);
    const float b = dot(lo, ld);
    const float cc = dot(lo, lo) - r * r;
    const float s = std::sqrt(b * b - 1.0f * cc);
    const bool inside = (-b - s) <= 0.0f;
    const float3 lp = lo + ld * t;
    const float3 ln = normalize(lp) * (inside ? -1.0f : 1.0f);
    c.pos = o2w * lp + F3(ob.center);
    c.normal = normalize(o2w * ln);
    return c;
}

extern void TraceOne(const or_bvh* h, float3 o, float3 d, float tMin, float tMax, bool anyHit, bool cull,
                     float margin, int* id, float* t, float sdfEps);
extern bool Intersector(const ap_scene_object* objs, int idx, float3 o, float3 d, float tMin, float tMax,
                        float& tOut, float sdfEps);

// ---- NEE (nee.cuh) ----
struct LightSample { float3 point, normal; float pFacet, pA; int object; bool valid; };

static void OBBAxes(const ap_scene_object& o, float3& ux, float3& uy, float3& uz) {
    const ap_float3x3& M = o.object_to_world;
    ux = normalize(float3{M.x.x, M.y.x, M.z.x});
    uy = normalize(float3{M.x.y, M.y.y, M.z.y});
    uz = normalize(float3{M.x.z, M.y.z, M.z.z});
}

static float AreaToSolidAnglePDF(float3 sp, float3 lp, float3 ln, float pa) {
    const float3 sl = lp - sp;
    const float d2 = std::fmax(dot(sl, sl), 1e-20f);
    const float3 dir = sl * (1.0f / std::sqrt(d2));
    const float cosY = std::fmax(0.0f, dot(ln, -dir));
    if (cosY <= 0.0f) return 0.0f;
    return pa * (d2 / cosY);
}

static void Facets(const ap_scene_object& b, float3 n[6], float3 c[6], float3 tg[6], float3 bt[6], float s1[6], float s2[6]) {
    float3 uX, uY, uZ;
    OBBAxes(b, uX, uY, uZ);
    const float3 C = F3(b.center), E = F3(b.extent);
    const float3 nn[6] = {uX, -uX, uY, -uY, uZ, -uZ};
    const float3 cc[6] = {C + uX * E.x, C - uX * E.x, C + uY * E.y, C - uY * E.y, C + uZ * E.z, C - uZ * E.z};
    const float3 tt[6] = {uY, uY, uX, uX, uX, uX};
    const float3 bb[6] = {uZ, uZ, uZ, uZ, uY, uY};
    const float a1[6] = {E.y, E.y, E.x, E.x, E.x, E.x};
    const float a2[6] = {E.z, E.z, E.z, E.z, E.y, E.y};
    for (int f = 0; f < 6; ++f) { n[f] = nn[f]; c[f] = cc[f]; tg[f] = tt[f]; bt[f] = bb[f]; s1[f] = a1[f]; s2[f] = a2[f]; }
}

static float FacingWeight(float3 fn, float3 fc, float3 p, float area) {
    return area * std::fmax(0.0f, dot(fn, normalize(p - fc)));
}

static float NEEPDFOmegaFromHit(const ap_scene_object& L, const ap_light_info& li, float3 sp, float3 lp, float3 ln) {
    if (L.type == AP_OBJ_CUBE) {
        float3 n[6], c[6], tg[6], bt[6];
        float s1[6], s2[6], area[6];
        Facets(L, n, c, tg, bt, s1, s2);
        for (int f = 0; f < 6; ++f) area[f] = 4.0f * s1[f] * s2[f];
        int hitF = -1;
        for (int f = 0; f < 6; ++f)
            if (dot(n[f], ln) > 0.999f) { hitF = f; break; }
        if (hitF < 0) {
            float best = 1e30f;
            for (int f = 0; f < 6; ++f) {
                const float d = std::fabs(dot(n[f], lp - c[f]));
                if (d < best) { best = d; hitF = f; }
            }
        }
        float sum = 0.0f, wh = 0.0f, ah = 0.0f;
        for (int f = 0; f < 6; ++f) {
            const float w = FacingWeight(n[f], c[f], sp, area[f]);
            sum += w;
            if (f == hitF) { wh = w; ah = area[f]; }
        }
        if (sum <= 0.0f || wh <= 0.0f) return 0.0f;
        return AreaToSolidAnglePDF(sp, lp, ln, li.select_prob * (1.0f / ah) * (wh / sum));
    }
    if (L.type == AP_OBJ_SPHERE) {
        const float3 sc = F3(L.center);
        const float r = max3(F3(L.extent));
        const float dist = length(sp - sc);
        if (dist <= r) return 0.0f;
        const float cosC = clampf(r / dist, 0.0f, 1.0f);
        const float cap = (g_plant == 2 ? 1.0f : 2.0f) * PI_F * r * r * (1.0f - cosC);
        if (cap <= 1e-20f) return 0.0f;
        return AreaToSolidAnglePDF(sp, lp, ln, li.select_prob / cap);
    }
    return 0.0f;
}

static LightSample NEESampleLight(const Scene& S, float3 p, Rng& rng) {
    LightSample ls{};
    ls.valid = false;
    const int n = static_cast<int>(S.lights.size());
    if (n <= 0) return ls;
    int pick = n - 1;
    float pprob = 0.0f;
    {
        const float r = rng.rand1();
        float acc = 0.0f;
        for (int k = 0; k < n; ++k) {
            acc += S.lights[k].select_prob;
            if (r < acc) { pick = k; pprob = S.lights[k].select_prob; break; }
        }
    }
    const ap_scene_object& L = S.objs[S.lights[pick].object_index];
    if (L.type == AP_OBJ_CUBE) {
        float3 fn[6], fc[6], tg[6], bt[6];
        float s1[6], s2[6], area[6];
        Facets(L, fn, fc, tg, bt, s1, s2);
        for (int f = 0; f < 6; ++f) area[f] = 4.0f * s1[f] * s2[f];
        int vi[6]; float vw[6]; int vc = 0; float vs = 0.0f;
        for (int f = 0; f < 6; ++f) {
            const float w = FacingWeight(fn[f], fc[f], p, area[f]);
            if (w > 0.0f) { vi[vc] = f; vw[vc] = w; vs += w; ++vc; }
        }
        if (vc == 0) return ls;
        int slot = 0, face = vi[0];
        const float r = rng.rand1() * vs;
        float acc = 0.0f;
        for (int k = 0; k < vc; ++k) {
            acc += vw[k];
            if (r <= acc) { slot = k; face = vi[k]; break; }
        }
        const float u1 = 2.0f * rng.rand1() - 1.0f;
        const float u2 = 2.0f * rng.rand1() - 1.0f;
        ls.point = fc[face] + tg[face] * (u1 * s1[face]) + bt[face] * (u2 * s2[face]);
        ls.normal = fn[face];
        ls.pFacet = 1.0f * (vw[slot] / vs);
        ls.pA = 1.0f / area[face];
        ls.valid = true;
    } else if (L.type == AP_OBJ_SPHERE) {
        const float3 sc = F3(L.center);
        const float r = max3(F3(L.extent));
        const float3 pc = p - sc;
        const float dist = length(pc);
        if (dist <= r) return ls;
        const float3 axis = pc / dist;
        const float cosC = clampf(r / dist, 0.0f, 1.0f);
        const float cap = (g_plant == 2 ? 1.0f : 2.0f) * PI_F * r * r * (1.0f - cosC);
        const float u = rng.rand1();
        const float vphi = rng.rand1();
        const float cosT = lerp(cosC, 1.0f, u);
        const float sinT = std::sqrt(std::fmax(0.0f, 1.0f - cosT * cosT));
        const float phi = 2.0f * PI_F * vphi;
        float3 t, b;
        BuildTangentBasis(axis, t, b);
        const float3 nn = t * (apfm::cos(phi) * sinT) + b * (apfm::sin(phi) * sinT) + axis * cosT;
        ls.point = sc + nn * r;
        ls.normal = nn;
        ls.pFacet = 1.0f;
        ls.pA = 1.0f / cap;
        ls.valid = true;
    } else {
        return ls;
    }
    ls.object = S.lights[pick].object_index;
    ls.pFacet = ls.pFacet * pprob;
    return ls;
}

static float PowerHeuristic(float pA, float pB) {
    const float a = pA * pA, b = pB * pB;
    return a / std::fmax(1e-24f, a + b);
}

static float3 SkyLighting(float3 d) {
    const float3 white{1, 1, 1}, blue{0.05f, 0.45f, 1.0f};
    const float3 down = float3{0.2f, 0.6f, 1.0f} * 0.125f;
    return 0.5f * lerp(lerp(blue, white, apfm::pow075(std::fmax(d.y, 0.0f))), down,
                       std::fmin(std::fmax(-d.y * 0.5f + 0.5f, 0.0f), 1.0f));
}

struct RenderOpts {
    int max_depth, rr_min_depth, nee_max_depth, use_nee, nee_always, sky;
};

// Trace's intersection (render.cuh:88-107): inside a medium with an object
// of its own the ray meets only that object's surface (Intersect with a solo
// object, intersection.cuh:185-193), otherwise the BVH.
static void TraceMedia(const Scene& S, const MediaCtx& m, float3 ro, float3 rd, float eps, int* id, float* t) {
    const float inf = std::numeric_limits<float>::infinity();
    if (m.inside && m.object >= 0) {
        float tt;
        Count(CNT_EXT_RAYS);
        Count(CNT_EXT_PRIMS);
        if (Intersector(S.objs, m.object, ro, rd, eps, inf, tt, eps * 0.5f)) {
            *id = m.object;
            *t = tt;
        } else {
            *id = -1;
            *t = inf;
        }
        return;
    }
    TraceOne(S.bvh, ro, rd, eps, inf, false, S.margin > 0.0f, S.margin, id, t, eps * 0.5f);
}

// MediaContext after a hit (EvaluateMaterial, material.cuh:410-411; SampleBSDF
// sets the vertex to the hit point before Trace's offset).
static MediaCtx MediaAfterHit(const MediaCtx& m, const SampleCtx& sc, int id, float3 hitPos) {
    MediaCtx n;
    n.inside = sc.inside;
    n.vertex = hitPos;
    n.object = id;
    n.depth = m.depth + 1;
    return n;
}

// The wavelengths of a path (the core's hero-wavelength mode when n = 4):
// the hero and its rotations by 105 nm over [360, 780], with their bases,
// D65 weights, CMF values and the sum of their pdfs.
constexpr int MAXL = 4;
struct Waves {
    int n;
    float lam[MAXL], lb[MAXL][4], d65[MAXL];
    float3 xyz[MAXL];
    float psum;
};

static Waves MakeWaves(Rng& rng, int n) {
    Waves w;
    w.n = n;
    float lam, lpdf;
    {
        const float r0 = rng.rand1(), r1 = rng.rand1(), r2 = rng.rand1();
        SampleCIE1931(r0, r1, r2, &lam, &lpdf);
    }
    w.lam[0] = lam;
    w.psum = lpdf;
    for (int j = 1; j < n; ++j) {
        float l = lam + 105.0f * static_cast<float>(j);
        if (l > 780.0f) l -= 420.0f;
        w.lam[j] = l;
        w.psum = w.psum + PdfCIE1931Mix(l);
    }
    for (int j = 0; j < n; ++j) {
        QueryLambda(w.lam[j], w.lb[j]);
        w.d65[j] = D65Norm(w.lam[j]);
        w.xyz[j] = LambdaToXYZ(w.lam[j]);
    }
    return w;
}

// sum_j xyz_j L_j / psum.
static float3 WavesXYZ(const Waves& w, const float* L) {
    float3 acc = w.xyz[0] * L[0];
    for (int j = 1; j < w.n; ++j) acc = acc + w.xyz[j] * L[j];
    return acc / w.psum;
}

// Hero wavelengths at a dispersive interface (the core's extension; one
// wavelength has nothing to end): the secondary wavelengths would refract
// elsewhere, so they end there and the hero continues alone with weight
// psum / p(hero), which keeps the rest of the path an unbiased
// one-wavelength estimate under the pixel's division by psum.
static void HeroOnlyAtDispersion(const Waves& wv, int nl, const SampleCtx& sc, const GB& gb, bool& heroOnly,
                                 float* contrib) {
    if (nl == 1 || !sc.delta || heroOnly || IORLinearKFromAbbe(gb.ior, gb.abbe) == 0.0f) return;
    contrib[0] = contrib[0] * (wv.psum / PdfCIE1931Mix(wv.lam[0]));
    for (int j = 1; j < nl; ++j) contrib[j] = 0.0f;
    heroOnly = true;
}

static float MeanMask(const float* m, int n) {
    if (n == 1) return m[0];
    float s = m[0];
    for (int j = 1; j < n; ++j) s = s + m[j];
    return s * (1.0f / static_cast<float>(n));
}

// Per-path radiance collection for guiding (CalculateRadiance's
// radianceSample* arrays, render.cuh:478-520, and CollectRadiance, :353-432),
// per wavelength; a slot is open while any wavelength's throughput is
// non-zero (with one wavelength, the reference's mask[i] != 0).
struct Collect {
    float3 dir[PG_COLLECT], loc[PG_COLLECT], vdir[PG_COLLECT];
    float sample[PG_COLLECT][MAXL], mask[PG_COLLECT][MAXL];
    bool valid[PG_COLLECT];
    int n = 1;
    void Clear(int nl) {
        n = nl;
        for (int i = 0; i < PG_COLLECT; ++i) {
            dir[i] = loc[i] = vdir[i] = float3{0, 0, 0};
            for (int j = 0; j < MAXL; ++j) sample[i][j] = mask[i][j] = 0.0f;
            valid[i] = false;
        }
    }
    bool Open(int i) const {
        bool o = false;
        for (int j = 0; j < n; ++j) o = o || mask[i][j] != 0.0f;
        return o;
    }
    void Endpoint(const float* le, int ndd) {
        bool any = false;
        for (int j = 0; j < n; ++j) any = any || le[j] != 0.0f;
        if (!any) return;
        const int hi = std::min(ndd, PG_COLLECT);
        for (int i = 0; i < hi; ++i)
            if (Open(i)) {
                for (int j = 0; j < n; ++j) sample[i][j] = sample[i][j] + mask[i][j] * le[j];
                valid[i] = true;
            }
    }
    void Advance(float3 next, const float* contrib, float rrProb, int ndd, bool allowStart, float3 rayDir,
                 float3 pos, float3 nrm) {
        if (allowStart && ndd < PG_COLLECT && !Open(ndd)) {
            dir[ndd] = next;
            loc[ndd] = pos;
            vdir[ndd] = Reflect(rayDir, nrm);
            for (int j = 0; j < n; ++j) mask[ndd][j] = 1.0f;
        }
        const float invQ = static_cast<float>(1.0 / static_cast<double>(rrProb));
        const int hi = std::min(ndd + (allowStart ? 1 : 0), PG_COLLECT);
        for (int i = 0; i < hi; ++i)
            if (Open(i))
                for (int j = 0; j < n; ++j) mask[i][j] = mask[i][j] * (contrib[j] * invQ);
    }
    // Path end (render.cuh:910-957): valid vertices to PathGuidingSamples,
    // radiance sum_j xyz_j sample_j / psum.
    void Emit(const Waves& wv, ap_path_guiding_sample out[PG_COLLECT]) const {
        std::memset(out, 0, sizeof(ap_path_guiding_sample) * PG_COLLECT);
        int w = 0;
        for (int i = 0; i < PG_COLLECT; ++i) {
            if (w >= PG_COLLECT) break;
            if (!valid[i]) continue;
            const float3 rx = WavesXYZ(wv, sample[i]);
            const float strength = std::fmin(max3(rx), PG_MAX_BRIGHT);
            if (strength < PG_MIN_BRIGHT) continue;
            ap_path_guiding_sample& o = out[w++];
            o.direction = ap_float3{dir[i].x, dir[i].y, dir[i].z};
            const float3 r = fmax3(rx, float3{0.0f, 0.0f, 0.0f});
            o.radiance = ap_float3{r.x, r.y, r.z};
            const int4_ v = GetVoxelGrid(loc[i], vdir[i]);
            o.voxel_grid = ap_int4{v.x, v.y, v.z, v.w};
            o.valid = 1;
        }
    }
};

// CalculateRadiance (render.cuh:460-999) for one camera ray; returns the
// clamped XYZ sample and the primary hit id. With a guiding state G the
// probes of G guide the BSDF sampling and NEE's MIS pdf, and the path's
// guiding samples are written to `slots` (PG_COLLECT entries) if non-null.
// nl = 4 is the core's hero-wavelength mode (one wavelength without guiding
// is the reference's own estimator): the hero's basis drives every sampling
// decision, radiance and throughput are kept per wavelength, Russian
// roulette and the NEE rate read the mean throughput.
static float3 CalculateRadiance(const Scene& S, const RenderOpts& P, Rng& rng, float3 origin, float3 dir, bool inside0,
                                int* primary, const GuideState* G = nullptr, int frame = 0,
                                ap_path_guiding_sample* slots = nullptr, int nl = 1, or_gsample* gs = nullptr) {
    Count(CNT_PATHS);
    const Waves wv = MakeWaves(rng, nl);
    const float* lb = wv.lb[0];
    const bool guiding = G != nullptr;
    Collect col;
    if (guiding) col.Clear(nl);
    int ndd = 0;
    dir = normalize(dir);
    float radiance[MAXL], mask[MAXL];
    for (int j = 0; j < nl; ++j) { radiance[j] = 0.0f; mask[j] = 1.0f; }
    float3 query{1.0f, 1.0f, 1.0f};  // maskForProbeQuery
    float3 ro = origin, rd = dir;
    MediaCtx media;
    media.inside = inside0;
    media.vertex = origin;
    bool heroOnly = false;  // secondary wavelengths ended at a dispersive interface
    bool prevNEE = false;
    float lastPdf = 0.0f;
    float3 lastPos{0, 0, 0};
    auto updateQuery = [&]() {
        // maskForProbeQuery: the throughput's XYZ (one wavelength: xyz * mask).
        float3 t = wv.xyz[0] * mask[0];
        for (int j = 1; j < nl; ++j) t = t + wv.xyz[j] * mask[j];
        query = fmax3(t, float3{PG_QUERY_EPS, PG_QUERY_EPS, PG_QUERY_EPS});
        // sum3 + PROBES_QUERY_EPS is a double sum (the constant is a double literal).
        query = query / static_cast<float>(static_cast<double>(query.x + query.y + query.z) + 1.0e-6);
    };
    for (int depth = 0; depth < P.max_depth; ++depth) {
        // Trace (render.cuh:75-163).
        const float eps = INTERSECT_EPS(ndd);
        int id;
        float t;
        TraceMedia(S, media, ro, rd, eps, &id, &t);
        if (depth == 0 && primary) *primary = id;
        if (depth == 0 && gs) {  // RecordScreenGBuffer's inputs (render.cuh:640-643)
            *gs = or_gsample{};
            gs->model = 4;  // MAT_MAX: a miss
        }
        const float alpha = GuidingAlpha(frame);
        if (id < 0) {
            (void)rng.rand1();  // offset draw
            if (!media.inside) {
                float endpoint[MAXL] = {};
                for (int j = 0; j < nl; ++j) {
                    const float le = P.sky ? InteractC(wv.lb[j], SkyLighting(rd)) : 0.0f;
                    radiance[j] = radiance[j] + mask[j] * le * wv.d65[j];
                    endpoint[j] = le * wv.d65[j];
                }
                if (guiding) col.Endpoint(endpoint, ndd);
                break;
            }
            // A miss inside a medium (bMissThinSurfaceInsideMedia_, render.cuh:109-126,
            // 143-155): no emission, no NEE and no MIS update; weight 1; the
            // zero-initialised hit record makes the next ray start at the
            // world origin in the same direction, outside any medium; Trace's
            // miss branch clears bDeltaSurface_, so the vertex counts.
            float p = 1.0f;
            if (ndd >= P.rr_min_depth) {
                const float mm = MeanMask(mask, nl);
                const float basic = clampf(mm * 0.5f + mm * 0.5f, 0.05f, 0.95f);
                p = lerp(basic, 1.0f, 0.0f);
                if (rng.rand1() > p) break;
            }
            float contrib[MAXL];
            for (int j = 0; j < nl; ++j) contrib[j] = 1.0f / std::fmax(1.0f, 1e-12f);
            if (guiding) col.Advance(rd, contrib, p, ndd, false, rd, float3{0, 0, 0}, float3{0, 0, 0});
            for (int j = 0; j < nl; ++j) {
                mask[j] = mask[j] * contrib[j];
                mask[j] = mask[j] / p;
            }
            if (guiding) updateQuery();
            media = MediaCtx{};
            media.vertex = ro;
            ro = float3{0, 0, 0};
            ndd++;
            continue;
        }
        Ctx c = HitContext(S, id, ro, rd, t, eps * 0.5f);
        Count(CNT_HITS);
        const int model = S.objs[id].material.shading_model;
        if (depth == 0 && gs) {
            gs->albedo[0] = c.gb.albedo.x; gs->albedo[1] = c.gb.albedo.y; gs->albedo[2] = c.gb.albedo.z;
            gs->metallic = c.gb.metallic;
            gs->normal[0] = c.normal.x; gs->normal[1] = c.normal.y; gs->normal[2] = c.normal.z;
            gs->roughness = c.gb.roughness;
            gs->depth = t;
            gs->model = model;
            gs->hit = 1;
        }
        const Ctx ct = Textured(c);
        SampleCtx sc = SampleBSDF(lb, wv.lam[0], rng, model, rd, ct, media);
        float brdf[MAXL];
        for (int j = 0; j < nl; ++j) brdf[j] = EvaluateBSDFWithCos(wv.lb[j], wv.lam[j], model, rd, sc.next, ct, media);
        const MediaCtx mediaIn = media;
        media = MediaAfterHit(media, sc, id, c.pos);
        const bool inside = media.inside;
        const float offEps = eps * std::fmax(1.0f, length(c.pos));
        c.pos = c.pos + sign(dot(c.normal, sc.next)) * c.normal * offEps * (rng.rand1() + 0.5f);
        // Directional guiding of the sampled direction (render.cuh:546-639).
        const ap_octahedron* probe = nullptr;
        if (guiding && !sc.delta && !sc.pureLight && !inside && !G->probes.empty()) {
            probe = G->Query(c.pos, Reflect(rd, c.normal));
            Count(CNT_PROBE_QUERIES);
            Count(CNT_PROBE_FOUND, probe ? 1 : 0);
            if (probe) {
                const float pB = sc.pdf;
                const float pP = OctahedronHemispherePDF(*probe, sc.next, c.normal, query);
                if (pB > 1.0e-12f && pP > 1.0e-12f) {
                    if (rng.rand1() < alpha) {
                        const GuideSample gs = SampleOctahedronHemisphere(*probe, rng, c.normal, query);
                        if (dot(gs.dir, c.normal) > 1.0e-12f) {
                            const float pBS = BSDF_PDF(lb, model, c.gb, c.normal, rd, gs.dir);
                            sc.next = gs.dir;
                            for (int j = 0; j < nl; ++j)
                                brdf[j] = EvaluateBSDFWithCos(wv.lb[j], wv.lam[j], model, rd, gs.dir, ct, mediaIn);
                            sc.pdf = (1.0f - alpha) * std::fmax(pBS, 1.0e-12f) + alpha * gs.pdf;
                        } else {
                            probe = nullptr;
                        }
                    } else {
                        sc.pdf = (1.0f - alpha) * pB + alpha * pP;
                    }
                } else {
                    probe = nullptr;
                }
            }
        }
        float endpoint[MAXL] = {};
        if (c.lightIndex >= 0) {
            float misW = 1.0f;
            if (P.use_nee && depth > 0 && prevNEE) {
                const float pB = std::fmax(lastPdf, 0.0f);
                const float pL = NEEPDFOmegaFromHit(S.objs[id], S.lights[c.lightIndex], lastPos, c.pos, c.normal);
                if (pL >= 0.0f && pB > 0.0f) misW = PowerHeuristic(pB, pL);
            }
            for (int j = 0; j < nl; ++j)
                radiance[j] = radiance[j] + mask[j] * InteractC(wv.lb[j], sc.emissive) * wv.d65[j] * misW;
            if (guiding)
                for (int j = 0; j < nl; ++j) endpoint[j] = InteractC(wv.lb[j], sc.emissive * misW) * wv.d65[j];
        } else if (guiding && !(sc.emissive.x == 0.0f && sc.emissive.y == 0.0f && sc.emissive.z == 0.0f)) {
            for (int j = 0; j < nl; ++j) endpoint[j] = InteractC(wv.lb[j], sc.emissive) * wv.d65[j];
        }
        if (guiding) col.Endpoint(endpoint, ndd);
        if (sc.terminate) break;
        bool neeValid = false;
        const float mm = MeanMask(mask, nl);
        const float rate = P.nee_always ? 1.0f : std::sqrt(saturate_(mm));
        if (P.use_nee && ndd < P.nee_max_depth && !sc.delta && !inside && rng.rand1() < rate) {
            const LightSample ls = NEESampleLight(S, c.pos, rng);
            if (ls.valid) {
                const float pAT = std::fmax(1e-20f, ls.pA * ls.pFacet);
                const float3 lu = ls.point - c.pos;
                const float d2 = std::fmax(dot(lu, lu), static_cast<float>(1.0e-12));
                const float3 L = lu * (1.0f / std::sqrt(d2));
                // SegmentTrace -> AnyHitSegmentBVH (intersection.cuh:372-397).
                bool occluded = false;
                {
                    float3 sd = ls.point - c.pos;
                    const float dist = length(sd);
                    if (dist > 0.0f) {
                        sd = sd / dist;
                        const float seps = INTERSECT_EPS(depth) * std::fmax(1.0f, std::fmax(length(c.pos), length(ls.point)));
                        const float tMax = std::fmax(0.0f, dist - seps);
                        if (tMax > 0.0f) {
                            int sid;
                            float st;
                            TraceOne(S.bvh, c.pos, sd, 0.0f, tMax, true, S.margin > 0.0f, S.margin, &sid, &st,
                                     INTERSECT_EPS(depth) * 0.5f);
                            occluded = sid >= 0;
                        }
                    }
                }
                if (occluded) {
                    neeValid = true;
                } else {
                    const float cosX = std::fmax(0.0f, dot(c.normal, L));
                    const float cosY = std::fmax(0.0f, dot(ls.normal, -L));
                    if (!(cosX <= 0.0f || cosY <= 0.0f)) {
                        const float G2 = cosY / d2;
                        const ap_gbuffer& lg = S.objs[ls.object].material.texture.base;
                        neeValid = true;
                        const float pLo = AreaToSolidAnglePDF(c.pos, ls.point, ls.normal, ls.pFacet * ls.pA);
                        float pBs = std::fmax(BSDF_PDF(lb, model, c.gb, c.normal, rd, L), 0.0f);
                        if (probe) {
                            const float pPr = OctahedronHemispherePDF(*probe, L, c.normal, query);
                            pBs = (1.0f - alpha) * pBs + alpha * pPr;
                        }
                        const float w = PowerHeuristic(pLo, pBs);
                        for (int j = 0; j < nl; ++j) {
                            const float bc = EvaluateBSDFWithCos(wv.lb[j], wv.lam[j], model, rd, L, ct, mediaIn);
                            const float resp = InteractC(wv.lb[j], F3(lg.emissive));
                            const float contribution = bc * resp * (G2 / pAT);
                            radiance[j] = radiance[j] + mask[j] * (contribution * w) * wv.d65[j];
                        }
                    }
                }
            }
        }
        prevNEE = neeValid;
        lastPdf = sc.pdf;
        lastPos = c.pos;
        float contrib[MAXL];
        for (int j = 0; j < nl; ++j) contrib[j] = brdf[j] / std::fmax(sc.pdf, 1e-12f);
        HeroOnlyAtDispersion(wv, nl, sc, c.gb, heroOnly, contrib);
        float p = 1.0f;
        if (ndd >= P.rr_min_depth) {
            const float basic = clampf(mm * 0.5f + mm * 0.5f, 0.05f, 0.95f);
            p = lerp(basic, 1.0f, inside ? 0.99f : 0.0f);
            if (rng.rand1() > p) break;
        }
        if (guiding) {
            const bool allowStart = !sc.delta && !sc.pureLight && !inside;
            col.Advance(sc.next, contrib, p, ndd, allowStart, rd, c.pos, c.normal);
        }
        for (int j = 0; j < nl; ++j) {
            mask[j] = mask[j] * contrib[j];
            mask[j] = mask[j] / p;
        }
        if (guiding) updateQuery();
        ro = c.pos;
        rd = sc.next;
        if (!sc.delta) ndd++;
    }
    if (guiding && slots) col.Emit(wv, slots);
    float r[MAXL] = {};
    for (int j = 0; j < nl; ++j) {
        r[j] = radiance[j];
        if (std::isnan(r[j]) || std::isinf(r[j])) r[j] = 0.0f;
    }
    const float3 pre = WavesXYZ(wv, r);
    return {clampf(pre.x, 0.0f, 1000.0f), clampf(pre.y, 0.0f, 1000.0f), clampf(pre.z, 0.0f, 1000.0f)};
}

// Radiance-cache options of one frame (include/aperture/ap_nrc.h).
struct NrcOpts {
    float spread;
    const float* params;
    float bmin[3], bmax[3];
    int width = 64;  // hidden width of the cache network (64 or 32)
    int mode = 1;    // its rounding points: 1 BF16, 2 FP16
};
constexpr int NRC_V = 8;

// Per-path training output: up to NRC_V samples of 20 floats, validity.
struct NrcTrain {
    float rec[NRC_V][20];
    bool valid[NRC_V];
};

// CalculateRadiance with the core's cache termination: a path ends into the
// cache at the first scattering vertex where the squared spread sum exceeds
// spread * a0; a training path restarts the sum at its first crossing and
// ends at the second (an unbiased one never ends there). Training paths
// record each scattering vertex before the cache vertex. With a guiding
// state G the sampling is guided as in CalculateRadiance, and a path that
// ends into the cache collects the prediction as its endpoint radiance.
static float3 CalculateRadianceNRC(const Scene& S, const RenderOpts& P, Rng& rng, float3 origin, float3 dir,
                                   bool inside0, const NrcOpts& N, bool train, bool unbiased, int pixel,
                                   NrcTrain* out, bool* queried, int nl = 1, const GuideState* G = nullptr,
                                   int frame = 0, ap_path_guiding_sample* slots = nullptr) {
    Count(CNT_PATHS);
    const Waves wv = MakeWaves(rng, nl);
    const float* lb = wv.lb[0];
    const bool guiding = G != nullptr;
    Collect col;
    if (guiding) col.Clear(nl);
    float3 gquery{1.0f, 1.0f, 1.0f};  // maskForProbeQuery
    int ndd = 0;
    dir = normalize(dir);
    float radiance[MAXL], mask[MAXL];
    for (int j = 0; j < nl; ++j) { radiance[j] = 0.0f; mask[j] = 1.0f; }
    float3 ro = origin, rd = dir;
    MediaCtx media;
    media.inside = inside0;
    media.vertex = origin;
    bool prevNEE = false;
    float lastPdf = 0.0f;
    float3 lastPos{0, 0, 0};
    float a0 = 0.0f, spreadSum = 0.0f;
    bool prevDelta = false;  // the last vertex was glass, or a miss inside a medium
    bool heroOnly = false;
    bool suffix = false;
    int nv = 0;
    float vrec[NRC_V][16], vR[NRC_V][MAXL], vM[NRC_V][MAXL];
    bool cacheEnd = false;
    float query[16];
    int endNdd = 0;
    auto updateQuery = [&]() {
        float3 t = wv.xyz[0] * mask[0];
        for (int j = 1; j < nl; ++j) t = t + wv.xyz[j] * mask[j];
        gquery = fmax3(t, float3{PG_QUERY_EPS, PG_QUERY_EPS, PG_QUERY_EPS});
        gquery = gquery / static_cast<float>(static_cast<double>(gquery.x + gquery.y + gquery.z) + 1.0e-6);
    };
    if (out) for (int k = 0; k < NRC_V; ++k) out->valid[k] = false;
    if (queried) *queried = false;
    auto record = [&](float* q, float3 pos, const Ctx& c) {
        q[0] = pos.x; q[1] = pos.y; q[2] = pos.z;
        q[3] = -rd.x; q[4] = -rd.y; q[5] = -rd.z;
        q[6] = c.normal.x; q[7] = c.normal.y; q[8] = c.normal.z;
        q[9] = c.gb.roughness;
        q[10] = c.gb.albedo.x; q[11] = c.gb.albedo.y; q[12] = c.gb.albedo.z;
        q[13] = q[14] = q[15] = 0.0f;
    };
    for (int depth = 0; depth < P.max_depth; ++depth) {
        const float eps = INTERSECT_EPS(ndd);
        int id;
        float t;
        TraceMedia(S, media, ro, rd, eps, &id, &t);
        const float alpha = GuidingAlpha(frame);
        if (id < 0) {
            (void)rng.rand1();
            if (!media.inside) {
                float endpoint[MAXL] = {};
                for (int j = 0; j < nl; ++j) {
                    const float le = P.sky ? InteractC(wv.lb[j], SkyLighting(rd)) : 0.0f;
                    radiance[j] = radiance[j] + mask[j] * le * wv.d65[j];
                    endpoint[j] = le * wv.d65[j];
                }
                if (guiding) col.Endpoint(endpoint, ndd);
                break;
            }
            // A miss inside a medium (bMissThinSurfaceInsideMedia_, render.cuh:109-126,
            // 143-155): no emission, no NEE and no MIS update; weight 1; the
            // zero-initialised hit record makes the next ray start at the
            // world origin in the same direction, outside any medium; Trace's
            // miss branch clears bDeltaSurface_, so the vertex counts.
            float p = 1.0f;
            if (ndd >= P.rr_min_depth) {
                const float mm = MeanMask(mask, nl);
                const float basic = clampf(mm * 0.5f + mm * 0.5f, 0.05f, 0.95f);
                p = lerp(basic, 1.0f, 0.0f);
                if (rng.rand1() > p) break;
            }
            float contrib[MAXL];
            for (int j = 0; j < nl; ++j) contrib[j] = 1.0f / std::fmax(1.0f, 1e-12f);
            if (guiding) col.Advance(rd, contrib, p, ndd, false, rd, float3{0, 0, 0}, float3{0, 0, 0});
            for (int j = 0; j < nl; ++j) {
                mask[j] = mask[j] * contrib[j];
                mask[j] = mask[j] / p;
            }
            if (guiding) updateQuery();
            media = MediaCtx{};
            media.vertex = ro;
            ro = float3{0, 0, 0};
            prevDelta = true;
            ndd++;
            continue;
        }
        Ctx c = HitContext(S, id, ro, rd, t, eps * 0.5f);
        Count(CNT_HITS);
        const float3 pos0 = c.pos;
        const int model = S.objs[id].material.shading_model;
        const Ctx ct = Textured(c);
        SampleCtx sc = SampleBSDF(lb, wv.lam[0], rng, model, rd, ct, media);
        float brdf[MAXL];
        for (int j = 0; j < nl; ++j) brdf[j] = EvaluateBSDFWithCos(wv.lb[j], wv.lam[j], model, rd, sc.next, ct, media);
        const MediaCtx mediaIn = media;
        media = MediaAfterHit(media, sc, id, c.pos);
        const bool inside = media.inside;
        const float offEps = eps * std::fmax(1.0f, length(c.pos));
        c.pos = c.pos + sign(dot(c.normal, sc.next)) * c.normal * offEps * (rng.rand1() + 0.5f);
        // Directional guiding of the sampled direction (render.cuh:546-639).
        const ap_octahedron* probe = nullptr;
        if (guiding && !sc.delta && !sc.pureLight && !inside && !G->probes.empty()) {
            probe = G->Query(c.pos, Reflect(rd, c.normal));
            Count(CNT_PROBE_QUERIES);
            Count(CNT_PROBE_FOUND, probe ? 1 : 0);
            if (probe) {
                const float pB = sc.pdf;
                const float pP = OctahedronHemispherePDF(*probe, sc.next, c.normal, gquery);
                if (pB > 1.0e-12f && pP > 1.0e-12f) {
                    if (rng.rand1() < alpha) {
                        const GuideSample gs = SampleOctahedronHemisphere(*probe, rng, c.normal, gquery);
                        if (dot(gs.dir, c.normal) > 1.0e-12f) {
                            const float pBS = BSDF_PDF(lb, model, c.gb, c.normal, rd, gs.dir);
                            sc.next = gs.dir;
                            for (int j = 0; j < nl; ++j)
                                brdf[j] = EvaluateBSDFWithCos(wv.lb[j], wv.lam[j], model, rd, gs.dir, ct, mediaIn);
                            sc.pdf = (1.0f - alpha) * std::fmax(pBS, 1.0e-12f) + alpha * gs.pdf;
                        } else {
                            probe = nullptr;
                        }
                    } else {
                        sc.pdf = (1.0f - alpha) * pB + alpha * pP;
                    }
                } else {
                    probe = nullptr;
                }
            }
        }
        float endpoint[MAXL] = {};
        if (c.lightIndex >= 0) {
            float misW = 1.0f;
            if (P.use_nee && depth > 0 && prevNEE) {
                const float pB = std::fmax(lastPdf, 0.0f);
                const float pL = NEEPDFOmegaFromHit(S.objs[id], S.lights[c.lightIndex], lastPos, c.pos, c.normal);
                if (pL >= 0.0f && pB > 0.0f) misW = PowerHeuristic(pB, pL);
            }
            for (int j = 0; j < nl; ++j)
                radiance[j] = radiance[j] + mask[j] * InteractC(wv.lb[j], sc.emissive) * wv.d65[j] * misW;
            if (guiding)
                for (int j = 0; j < nl; ++j) endpoint[j] = InteractC(wv.lb[j], sc.emissive * misW) * wv.d65[j];
        } else if (guiding && !(sc.emissive.x == 0.0f && sc.emissive.y == 0.0f && sc.emissive.z == 0.0f)) {
            for (int j = 0; j < nl; ++j) endpoint[j] = InteractC(wv.lb[j], sc.emissive) * wv.d65[j];
        }
        if (guiding) col.Endpoint(endpoint, ndd);
        // Spread heuristic (the core's shade kernel, NRC branch).
        {
            const float cosI = std::fabs(dot(rd, c.normal));
            const float dist2 = t * t;
            if (depth == 0) {
                a0 = dist2 / (12.566370614359172f * cosI);
                spreadSum = 0.0f;
            } else if (!prevDelta) {  // a segment leaving a delta vertex adds nothing
                spreadSum = spreadSum + std::sqrt(dist2 / (std::fmax(lastPdf, 1e-12f) * cosI));
            }
            prevDelta = sc.delta;
            const bool scatter = !sc.pureLight && !sc.delta &&
                                 !(model == AP_MAT_GGX && !(std::fmax(0.0f, dot(c.normal, -rd)) > 0.0f));
            if (scatter && !inside && spreadSum * spreadSum > N.spread * a0) {
                if (!train) {
                    cacheEnd = true;
                } else if (!unbiased) {
                    if (suffix) {
                        cacheEnd = true;
                    } else {
                        suffix = true;
                        spreadSum = 0.0f;
                    }
                }
            }
            if (cacheEnd) {
                record(query, pos0, c);
                endNdd = ndd;
                break;
            }
            if (train && scatter && nv < NRC_V) {
                record(vrec[nv], pos0, c);
                for (int j = 0; j < nl; ++j) { vR[nv][j] = radiance[j]; vM[nv][j] = mask[j]; }
                ++nv;
            }
        }
        if (sc.terminate) break;
        bool neeValid = false;
        const float mm = MeanMask(mask, nl);
        const float rate = P.nee_always ? 1.0f : std::sqrt(saturate_(mm));
        if (P.use_nee && ndd < P.nee_max_depth && !sc.delta && !inside && rng.rand1() < rate) {
            const LightSample ls = NEESampleLight(S, c.pos, rng);
            if (ls.valid) {
                const float pAT = std::fmax(1e-20f, ls.pA * ls.pFacet);
                const float3 lu = ls.point - c.pos;
                const float d2 = std::fmax(dot(lu, lu), static_cast<float>(1.0e-12));
                const float3 L = lu * (1.0f / std::sqrt(d2));
                bool occluded = false;
                {
                    float3 sd = ls.point - c.pos;
                    const float dist = length(sd);
                    if (dist > 0.0f) {
                        sd = sd / dist;
                        const float seps = INTERSECT_EPS(depth) * std::fmax(1.0f, std::fmax(length(c.pos), length(ls.point)));
                        const float tMax = std::fmax(0.0f, dist - seps);
                        if (tMax > 0.0f) {
                            int sid;
                            float st;
                            TraceOne(S.bvh, c.pos, sd, 0.0f, tMax, true, S.margin > 0.0f, S.margin, &sid, &st,
                                     INTERSECT_EPS(depth) * 0.5f);
                            occluded = sid >= 0;
                        }
                    }
                }
                if (occluded) {
                    neeValid = true;
                } else {
                    const float cosX = std::fmax(0.0f, dot(c.normal, L));
                    const float cosY = std::fmax(0.0f, dot(ls.normal, -L));
                    if (!(cosX <= 0.0f || cosY <= 0.0f)) {
                        const float G2 = cosY / d2;
                        const ap_gbuffer& lg = S.objs[ls.object].material.texture.base;
                        neeValid = true;
                        const float pLo = AreaToSolidAnglePDF(c.pos, ls.point, ls.normal, ls.pFacet * ls.pA);
                        float pBs = std::fmax(BSDF_PDF(lb, model, c.gb, c.normal, rd, L), 0.0f);
                        if (probe) {
                            const float pPr = OctahedronHemispherePDF(*probe, L, c.normal, gquery);
                            pBs = (1.0f - alpha) * pBs + alpha * pPr;
                        }
                        const float w = PowerHeuristic(pLo, pBs);
                        for (int j = 0; j < nl; ++j) {
                            const float bc = EvaluateBSDFWithCos(wv.lb[j], wv.lam[j], model, rd, L, ct, mediaIn);
                            const float resp = InteractC(wv.lb[j], F3(lg.emissive));
                            const float contribution = bc * resp * (G2 / pAT);
                            radiance[j] = radiance[j] + mask[j] * (contribution * w) * wv.d65[j];
                        }
                    }
                }
            }
        }
        prevNEE = neeValid;
        lastPdf = sc.pdf;
        lastPos = c.pos;
        float contrib[MAXL];
        for (int j = 0; j < nl; ++j) contrib[j] = brdf[j] / std::fmax(sc.pdf, 1e-12f);
        HeroOnlyAtDispersion(wv, nl, sc, c.gb, heroOnly, contrib);
        float p = 1.0f;
        if (ndd >= P.rr_min_depth) {
            const float basic = clampf(mm * 0.5f + mm * 0.5f, 0.05f, 0.95f);
            p = lerp(basic, 1.0f, inside ? 0.99f : 0.0f);
            if (rng.rand1() > p) break;
        }
        if (guiding) {
            const bool allowStart = !sc.delta && !sc.pureLight && !inside;
            col.Advance(sc.next, contrib, p, ndd, allowStart, rd, c.pos, c.normal);
        }
        for (int j = 0; j < nl; ++j) {
            mask[j] = mask[j] * contrib[j];
            mask[j] = mask[j] / p;
        }
        if (guiding) updateQuery();
        ro = c.pos;
        rd = sc.next;
        if (!sc.delta) ndd++;
    }
    if (cacheEnd) {
        // The cache's prediction, clamped at 0, at the path's wavelengths.
        float feat[64], y[3];
        or_nrc_encode(query, 1, N.bmin, N.bmax, feat);
        if (N.width == 64 && N.mode == 1) or_nrc_forward(N.params, feat, 1, 1, y);
        else or_nrc_forward_w(N.params, N.width, feat, 1, N.mode, y);
        const float3 cv = fmax3(float3{y[0], y[1], y[2]}, float3{0.0f, 0.0f, 0.0f});
        float endpoint[MAXL] = {};
        for (int j = 0; j < nl; ++j) {
            const float le = InteractC(wv.lb[j], cv);
            radiance[j] = radiance[j] + mask[j] * le * wv.d65[j];
            endpoint[j] = le * wv.d65[j];
        }
        if (guiding) col.Endpoint(endpoint, endNdd);
        if (queried) *queried = true;
    }
    if (guiding && slots) col.Emit(wv, slots);
    if (out) {
        for (int k = 0; k < nv; ++k) {
            float L[MAXL] = {};
            for (int j = 0; j < nl; ++j)
                L[j] = (nl > 1 && !(vM[k][j] > 0.0f)) ? 0.0f : (radiance[j] - vR[k][j]) / vM[k][j];
            const float3 rgb = XYZ2SRGB(WavesXYZ(wv, L));
            if (std::isnan(rgb.x) || std::isnan(rgb.y) || std::isnan(rgb.z) || std::isinf(rgb.x) ||
                std::isinf(rgb.y) || std::isinf(rgb.z))
                continue;
            float* o = out->rec[k];
            for (int e = 0; e < 13; ++e) o[e] = vrec[k][e];
            o[13] = o[14] = o[15] = 0.0f;
            o[16] = rgb.x; o[17] = rgb.y; o[18] = rgb.z;
            const uint32_t key = static_cast<uint32_t>(pixel) * 8u + static_cast<uint32_t>(k);
            std::memcpy(&o[19], &key, 4);
            out->valid[k] = true;
        }
    }
    float r[MAXL] = {};
    for (int j = 0; j < nl; ++j) {
        r[j] = radiance[j];
        if (std::isnan(r[j]) || std::isinf(r[j])) r[j] = 0.0f;
    }
    const float3 pre = WavesXYZ(wv, r);
    return {clampf(pre.x, 0.0f, 1000.0f), clampf(pre.y, 0.0f, 1000.0f), clampf(pre.z, 0.0f, 1000.0f)};
}

static float GetArea(const ap_scene_object& o) {
    if (o.type == AP_OBJ_CUBE || o.type == AP_OBJ_SDF) {
        const float ex = std::fabs(o.extent.x), ey = std::fabs(o.extent.y), ez = std::fabs(o.extent.z);
        return 8.0f * (ex * ey + ex * ez + ey * ez);
    }
    if (o.type == AP_OBJ_SPHERE) { const float r = std::fabs(o.extent.x); return 4.0f * PI_F * r * r; }
    if (o.type == AP_OBJ_TRIANGLE) {
        const float3 e1 = F3(o.additional.v1.position) - F3(o.additional.v0.position);
        const float3 e2 = F3(o.additional.v2.position) - F3(o.additional.v0.position);
        return 0.5f * length(cross(e1, e2));
    }
    return 0.0f;
}

// Scene, options and camera of a render call (RenderCamera, scene.cu:201-230;
// light table, scene.cu:1056-1080).
struct Setup {
    Scene S;
    RenderOpts P;
    float3 camO, camD, right, up;
    float tanX, tanY;
    bool inside0;
    int W, H;
};

// The light table (scene.cu:1056-1080): area * luminance weights, linear CDF.
static void BuildLights(Scene& S) {
    float total = 0.0f;
    for (size_t i = 0; i < S.n; ++i) {
        if (S.objs[i].material.shading_model == AP_MAT_LIGHT) {
            S.light_index[i] = static_cast<int>(S.lights.size());
            const float3 em = F3(S.objs[i].material.texture.base.emissive);
            ap_light_info li{static_cast<int32_t>(i),
                             std::fmax(GetArea(S.objs[i]) * dot(em, float3{0.2126f, 0.7152f, 0.0722f}), 1.0e-20f), 0.0f};
            total += li.importance_weight;
            S.lights.push_back(li);
        }
    }
    for (auto& li : S.lights) li.select_prob = li.importance_weight / total;
}

static Setup MakeSetup(const or_bvh* bvh, const ap_scene_object* objs, size_t n, const ap_camera* cam,
                       const ap_render_params* prm, float cull_margin) {
    Setup U{Scene{objs, n, bvh, {}, std::vector<int>(n, -1), prm->triangle_material, cull_margin},
            RenderOpts{prm->max_depth, prm->rr_min_depth, prm->nee_max_depth, prm->use_nee, prm->nee_always, prm->sky},
            {}, {}, {}, {}, 0.0f, 0.0f, false, cam->width, cam->height};
    BuildLights(U.S);
    U.camO = float3{cam->origin[0], cam->origin[1], cam->origin[2]};
    U.camD = normalize(float3{cam->direction[0], cam->direction[1], cam->direction[2]});
    const float FOV = cam->fov_deg * 3.1415926f / 180.0f;
    const float aspect = static_cast<float>(cam->width) / std::fmax(1.0f, static_cast<float>(cam->height));
    U.tanY = std::tan(0.5f * FOV);
    U.tanX = U.tanY * aspect;
    U.right = normalize(cross(U.camD, float3{0.0f, 1.0f, 0.0f}));
    U.up = normalize(cross(U.camD, U.right));
    for (size_t i = 0; i < n && !U.inside0; ++i) {
        const ap_scene_object& ob = objs[i];
        if (ob.type == AP_OBJ_SDF) {
            U.inside0 = OverlapSDF(ob, U.camO);
            continue;
        }
        if (ob.type != AP_OBJ_CUBE && ob.type != AP_OBJ_SPHERE) continue;
        const float3 lo = M3(ob.world_to_object) * (U.camO - F3(ob.center));
        const float3 e = F3(ob.extent);
        if (ob.type == AP_OBJ_CUBE)
            U.inside0 = -e.x < lo.x && -e.y < lo.y && -e.z < lo.z && e.x > lo.x && e.y > lo.y && e.z > lo.z;
        else
            U.inside0 = dot(lo, lo) < max3(e) * max3(e);
    }
    return U;
}

// Camera ray of pixel p (two jitter draws from rng).
static float3 CameraRay(const Setup& U, Rng& rng, int64_t p) {
    const int i = static_cast<int>(p % U.W), j = static_cast<int>(p / U.W);
    const float u = (i + rng.rand1()) / static_cast<float>(U.W);
    const float v = (j + rng.rand1()) / static_cast<float>(U.H);
    const float nx = u * 2.f - 1.f;
    const float ny = static_cast<float>((1.0 - v) * 2.f - 1.f);
    return normalize(U.camD + U.right * (nx * U.tanX) + U.up * (ny * U.tanY));
}

}  // namespace orc

using namespace orc;

struct or_guide {
    GuideState state;
};

extern "C" {

// Offline render (frames frame0 .. frame0+spp-1) of the scene into
// hdr (W*H*3 running mean, updated in place), sums (W*H*3 doubles, added)
// and primary ids (W*H, last frame). cull_margin > 0 uses the core's CULLED
// traversal order, which returns the same hits.
int or_render_guided(const or_bvh* bvh, const ap_scene_object* objs, size_t n, const ap_camera* cam,
                     const ap_render_params* prm, float cull_margin, float* hdr, double* sums, int32_t* primary,
                     int threads, or_guide* guide) {
    const Setup U = MakeSetup(bvh, objs, n, cam, prm, cull_margin);
    const Scene& S = U.S;
    const RenderOpts& P = U.P;
    const bool inside0 = U.inside0;
    const int64_t npix = static_cast<int64_t>(U.W) * U.H;
    const int nl = prm->wavelengths == 4 ? 4 : 1;
    if (prm->use_guiding) {
        // Frames in order; each renders with the probes built from the frames
        // before it, then folds its own samples in (scene.cu:308-386). With
        // guide_frames F > 1 the frames go in groups of F that render with
        // the probes built before the group; then one update folds in the
        // group's samples, frame-major in slot order.
        if (!guide) return -1;
        GuideState& G = guide->state;
        const int group = std::max(1, prm->guide_frames);
        const size_t frame_slots = static_cast<size_t>(npix) * PG_COLLECT;
        const size_t cap_frames = static_cast<size_t>(std::min(group, std::max(prm->spp, 1)));
        std::vector<ap_path_guiding_sample> slots(frame_slots * cap_frames);
        for (int s = 0; s < prm->spp; ++s) {
            const int frame = prm->frame_index + s;
            const size_t f_in_group = static_cast<size_t>(s % group);
            ap_path_guiding_sample* fslots = &slots[f_in_group * frame_slots];
#pragma omp parallel for schedule(dynamic, 64) num_threads(threads > 0 ? threads : 1)
            for (int64_t p = 0; p < npix; ++p) {
                Rng rng = PathRng(static_cast<uint64_t>(p), static_cast<uint64_t>(frame), prm->seed);
                const float3 rd = CameraRay(U, rng, p);
                int prim = -1;
                float3 res = CalculateRadiance(S, P, rng, U.camO, rd, inside0, &prim, &G, frame,
                                               &fslots[static_cast<size_t>(p) * PG_COLLECT], nl);
                res = fmax3(float3{0, 0, 0}, res);
                const float3 prev{hdr[3 * p], hdr[3 * p + 1], hdr[3 * p + 2]};
                const float3 acc = lerp(prev, res, 1.0f / static_cast<float>(1 + frame));
                hdr[3 * p] = acc.x; hdr[3 * p + 1] = acc.y; hdr[3 * p + 2] = acc.z;
                if (sums) {
                    sums[3 * p] += static_cast<double>(res.x);
                    sums[3 * p + 1] += static_cast<double>(res.y);
                    sums[3 * p + 2] += static_cast<double>(res.z);
                }
                if (primary && s == prm->spp - 1) primary[p] = prim;
            }
            if (f_in_group + 1 == static_cast<size_t>(group) || s == prm->spp - 1) {
                slots.resize((f_in_group + 1) * frame_slots);
                G.Update(slots);
                slots.resize(frame_slots * cap_frames);
            }
        }
        return 0;
    }
#pragma omp parallel for schedule(dynamic, 64) num_threads(threads > 0 ? threads : 1)
    for (int64_t p = 0; p < npix; ++p) {
        float3 acc{hdr[3 * p], hdr[3 * p + 1], hdr[3 * p + 2]};
        for (int s = 0; s < prm->spp; ++s) {
            const int frame = prm->frame_index + s;
            Rng rng = PathRng(static_cast<uint64_t>(p), static_cast<uint64_t>(frame), prm->seed);
            const float3 rd = CameraRay(U, rng, p);
            int prim = -1;
            float3 res = CalculateRadiance(S, P, rng, U.camO, rd, inside0, &prim, nullptr, 0, nullptr, nl);
            res = fmax3(float3{0, 0, 0}, res);
            acc = lerp(acc, res, 1.0f / static_cast<float>(1 + frame));
            if (sums) {
                sums[3 * p] += static_cast<double>(res.x);
                sums[3 * p + 1] += static_cast<double>(res.y);
                sums[3 * p + 2] += static_cast<double>(res.z);
            }
            if (primary && s == prm->spp - 1) primary[p] = prim;
        }
        hdr[3 * p] = acc.x; hdr[3 * p + 1] = acc.y; hdr[3 * p + 2] = acc.z;
    }
    return 0;
}

int or_render_screen(const or_bvh* bvh, const ap_scene_object* objs, size_t n, const ap_camera* cam,
                     const ap_render_params* prm, float cull_margin, float* hdr, double* sums,
                     ap_screen_gbuffer* gbuffer, ap_screen_statistics* statistics, int threads) {
    if (prm->use_guiding) return -1;
    const Setup U = MakeSetup(bvh, objs, n, cam, prm, cull_margin);
    const int64_t npix = static_cast<int64_t>(U.W) * U.H;
    const int nl = prm->wavelengths == 4 ? 4 : 1;
#pragma omp parallel for schedule(dynamic, 64) num_threads(threads > 0 ? threads : 1)
    for (int64_t p = 0; p < npix; ++p) {
        float3 acc{hdr[3 * p], hdr[3 * p + 1], hdr[3 * p + 2]};
        for (int s = 0; s < prm->spp; ++s) {
            const int frame = prm->frame_index + s;
            Rng rng = PathRng(static_cast<uint64_t>(p), static_cast<uint64_t>(frame), prm->seed);
            const float3 rd = CameraRay(U, rng, p);
            or_gsample gs{};
            float3 res = CalculateRadiance(U.S, U.P, rng, U.camO, rd, U.inside0, nullptr, nullptr, 0, nullptr, nl,
                                           &gs);
            res = fmax3(float3{0, 0, 0}, res);
            acc = lerp(acc, res, 1.0f / static_cast<float>(1 + frame));
            if (sums) {
                sums[3 * p] += static_cast<double>(res.x);
                sums[3 * p + 1] += static_cast<double>(res.y);
                sums[3 * p + 2] += static_cast<double>(res.z);
            }
            // Frame order per pixel: the G-buffer at depth 0, the statistics of the sample.
            if (gbuffer) or_record_gbuffer(&gbuffer[p], &gs, 1);
            if (statistics) {
                const float x[3] = {res.x, res.y, res.z};
                or_record_statistics(&statistics[p], x, 1);
            }
        }
        hdr[3 * p] = acc.x; hdr[3 * p + 1] = acc.y; hdr[3 * p + 2] = acc.z;
    }
    return 0;
}

int or_render(const or_bvh* bvh, const ap_scene_object* objs, size_t n, const ap_camera* cam,
              const ap_render_params* prm, float cull_margin, float* hdr, double* sums, int32_t* primary,
              int threads) {
    if (prm->use_guiding) return -1;
    return or_render_guided(bvh, objs, n, cam, prm, cull_margin, hdr, sums, primary, threads, nullptr);
}

// Frames of the core's ap_render_nrc with the cache frozen at `params`
// (bounds bmin/bmax): hdr and sums as or_render; for the last frame, the
// training samples at samples[(pixel * 8 + vertex) * 20] with valid flags,
// and the number of paths that ended into the cache.
int or_render_nrc(const or_bvh* bvh, const ap_scene_object* objs, size_t n, const ap_camera* cam,
                  const ap_render_params* prm, float cull_margin, const float* params, const float* bmin,
                  const float* bmax, float spread, uint32_t train_paths, uint32_t unbiased_every, float* hdr,
                  double* sums, float* samples, uint8_t* valid, uint32_t* n_queries, int threads) {
    return or_render_nrc_w(bvh, objs, n, cam, prm, cull_margin, params, 64, 1, bmin, bmax, spread, train_paths,
                           unbiased_every, hdr, sums, samples, valid, n_queries, threads);
}

// or_render_nrc with a network of hidden width `width` (64 or 32) and
// rounding points `mode` (1 BF16, 2 FP16): a per-scene cache of a batch.
int or_render_nrc_w(const or_bvh* bvh, const ap_scene_object* objs, size_t n, const ap_camera* cam,
                    const ap_render_params* prm, float cull_margin, const float* params, int width, int mode,
                    const float* bmin, const float* bmax, float spread, uint32_t train_paths,
                    uint32_t unbiased_every, float* hdr, double* sums, float* samples, uint8_t* valid,
                    uint32_t* n_queries, int threads) {
    return or_render_nrc_guided(bvh, objs, n, cam, prm, cull_margin, params, width, mode, bmin, bmax, spread,
                                train_paths, unbiased_every, hdr, sums, samples, valid, n_queries, threads, nullptr);
}

// or_render_nrc_w with directional guiding (prm->use_guiding and a guide):
// frames in order, each sampling the probes the frames before it built and
// then folding its own guiding samples in, as or_render_guided.
int or_render_nrc_guided(const or_bvh* bvh, const ap_scene_object* objs, size_t n, const ap_camera* cam,
                         const ap_render_params* prm, float cull_margin, const float* params, int width, int mode,
                         const float* bmin, const float* bmax, float spread, uint32_t train_paths,
                         uint32_t unbiased_every, float* hdr, double* sums, float* samples, uint8_t* valid,
                         uint32_t* n_queries, int threads, or_guide* guide) {
    if ((width != 64 && width != 32) || (mode != 1 && mode != 2)) return -1;
    if (prm->use_guiding && !guide) return -1;
    GuideState* G = prm->use_guiding ? &guide->state : nullptr;
    const Setup U = MakeSetup(bvh, objs, n, cam, prm, cull_margin);
    const int64_t npix = static_cast<int64_t>(U.W) * U.H;
    const int nl = prm->wavelengths == 4 ? 4 : 1;
    NrcOpts N{spread, params, {bmin[0], bmin[1], bmin[2]}, {bmax[0], bmax[1], bmax[2]}, width, mode};
    uint32_t stride = 1, n_train_max = 0;
    if (train_paths > 0) {
        stride = std::max<uint32_t>(1u, static_cast<uint32_t>(npix) / train_paths);
        n_train_max = (static_cast<uint32_t>(npix) + stride - 1) / stride;
    }
    uint32_t queries = 0;
    std::vector<ap_path_guiding_sample> gslots(G ? static_cast<size_t>(npix) * PG_COLLECT : 0);
    for (int s = 0; s < prm->spp; ++s) {
        const int frame = prm->frame_index + s;
        const bool last = s == prm->spp - 1;
        const uint64_t h = SplitMix64(SplitMix64(static_cast<uint64_t>(frame)) ^ prm->seed ^ 0x6e7263ull);
        uint32_t offset = 0, n_train = 0;
        if (n_train_max > 0) {
            offset = static_cast<uint32_t>(h % stride);
            n_train = std::min(n_train_max, (static_cast<uint32_t>(npix) - offset + stride - 1) / stride);
        }
        uint32_t frame_queries = 0;
#pragma omp parallel for schedule(dynamic, 64) num_threads(threads > 0 ? threads : 1) reduction(+ : frame_queries)
        for (int64_t p = 0; p < npix; ++p) {
            const uint32_t up = static_cast<uint32_t>(p);
            bool train = false, unbiased = false;
            if (up >= offset && (up - offset) % stride == 0u) {
                const uint32_t t = (up - offset) / stride;
                if (t < n_train) {
                    train = true;
                    unbiased = unbiased_every != 0u && t % unbiased_every == 0u;
                }
            }
            Rng rng = PathRng(static_cast<uint64_t>(p), static_cast<uint64_t>(frame), prm->seed);
            const float3 rd = CameraRay(U, rng, p);
            NrcTrain tr;
            bool q = false;
            float3 res = CalculateRadianceNRC(U.S, U.P, rng, U.camO, rd, U.inside0, N, train, unbiased,
                                              static_cast<int>(p), &tr, &q, nl, G, frame,
                                              G ? &gslots[static_cast<size_t>(p) * PG_COLLECT] : nullptr);
            frame_queries += q ? 1u : 0u;
            res = fmax3(float3{0, 0, 0}, res);
            const float3 prev{hdr[3 * p], hdr[3 * p + 1], hdr[3 * p + 2]};
            const float3 acc = lerp(prev, res, 1.0f / static_cast<float>(1 + frame));
            hdr[3 * p] = acc.x; hdr[3 * p + 1] = acc.y; hdr[3 * p + 2] = acc.z;
            if (sums) {
                sums[3 * p] += static_cast<double>(res.x);
                sums[3 * p + 1] += static_cast<double>(res.y);
                sums[3 * p + 2] += static_cast<double>(res.z);
            }
            if (last && samples && valid) {
                for (int k = 0; k < NRC_V; ++k) {
                    const size_t slot = static_cast<size_t>(p) * NRC_V + k;
                    valid[slot] = train && tr.valid[k] ? 1 : 0;
                    if (valid[slot]) std::memcpy(samples + slot * 20, tr.rec[k], 20 * sizeof(float));
                }
            }
        }
        if (last) queries = frame_queries;
        if (G) G->Update(gslots);
    }
    if (n_queries) *n_queries = queries;
    return 0;
}

// ---- guiding primitives, for their CPU tests ----
void or_pg_uv_to_sphere(const float* uv, size_t n, float* out) {
    for (size_t i = 0; i < n; ++i) {
        const float3 d = UVtoSphere(uv[2 * i], uv[2 * i + 1]);
        out[3 * i] = d.x; out[3 * i + 1] = d.y; out[3 * i + 2] = d.z;
    }
}
void or_pg_sphere_to_uv(const float* d, size_t n, float* out) {
    for (size_t i = 0; i < n; ++i) SphereToUV(float3{d[3 * i], d[3 * i + 1], d[3 * i + 2]}, out[2 * i], out[2 * i + 1]);
}
void or_pg_voxel(const float* pos, const float* dir, size_t n, int32_t* out) {
    for (size_t i = 0; i < n; ++i) {
        const int4_ v = GetVoxelGrid(float3{pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]},
                                     float3{dir[3 * i], dir[3 * i + 1], dir[3 * i + 2]});
        out[4 * i] = v.x; out[4 * i + 1] = v.y; out[4 * i + 2] = v.z; out[4 * i + 3] = v.w;
    }
}
void or_pg_mipmaps(ap_octahedron* p) { GenerateMipmaps(*p); }
// pdf of directions d under probe p and query q; with a normal, the hemisphere pdf.
void or_pg_pdf(const ap_octahedron* p, const float* d, size_t n, const float* q, const float* normal, float* out) {
    const float3 qq{q[0], q[1], q[2]};
    for (size_t i = 0; i < n; ++i) {
        const float3 dd{d[3 * i], d[3 * i + 1], d[3 * i + 2]};
        out[i] = normal ? OctahedronHemispherePDF(*p, dd, float3{normal[0], normal[1], normal[2]}, qq)
                        : OctahedronPDF(*p, dd, qq);
    }
}
// n samples from PCG32(seed, 0); with a normal, the hemisphere sampler.
void or_pg_sample(const ap_octahedron* p, uint64_t seed, size_t n, const float* q, const float* normal, float* dirs,
                  float* pdfs) {
    Rng rng = Rng::seeded(seed, 0);
    const float3 qq{q[0], q[1], q[2]};
    for (size_t i = 0; i < n; ++i) {
        const GuideSample s = normal ? SampleOctahedronHemisphere(*p, rng, float3{normal[0], normal[1], normal[2]}, qq)
                                     : SampleOctahedron(*p, rng, qq);
        dirs[3 * i] = s.dir.x; dirs[3 * i + 1] = s.dir.y; dirs[3 * i + 2] = s.dir.z;
        pdfs[i] = s.pdf;
    }
}
void or_pg_update(or_guide* g, const ap_path_guiding_sample* slots, size_t n) {
    g->state.Update(std::vector<ap_path_guiding_sample>(slots, slots + n));
}

or_guide* or_guide_create(void) { return new or_guide(); }
void or_guide_free(or_guide* g) { delete g; }
size_t or_guide_probe_count(const or_guide* g) { return g ? g->state.probes.size() : 0; }

// Copies out the probes and index volume; bounds = start xyz, end xyz, size xyz.
void or_guide_export(const or_guide* g, ap_octahedron* probes, int32_t* volume, int32_t* bounds) {
    const GuideState& G = g->state;
    if (probes && !G.probes.empty()) std::memcpy(probes, G.probes.data(), G.probes.size() * sizeof(ap_octahedron));
    if (volume && !G.volume.empty()) std::memcpy(volume, G.volume.data(), G.volume.size() * sizeof(int32_t));
    if (bounds) {
        const int32_t b[9] = {G.start.x, G.start.y, G.start.z, G.end.x, G.end.y, G.end.z, G.size.x, G.size.y, G.size.z};
        std::memcpy(bounds, b, sizeof(b));
    }
}


// Glass restatement at one wavelength, for its tests.
float or_glass_ior(float lambda, float nd, float vd) { return IORFromAbbeNM(lambda, nd, vd); }
float or_glass_sigma(float sigma_a, float lambda, float tilt) { return SigmaAFromTiltNM(sigma_a, lambda, tilt); }
float or_glass_fresnel(float cosI, float cosT, float eta) { return FresnelDielectricExact(cosI, cosT, eta); }
float or_glass_beer(float distance, float resp, float sigma_a) { return BeerLambert_Spectrum(distance, resp, sigma_a); }

void or_glass_bsdf(const float* d, const float* n, const float* mat, float lambda, int inside, int depth,
                   const float* vertex, const float* hit, uint64_t seed, float* out) {
    Ctx c;
    c.hit = true;
    c.pos = {hit[0], hit[1], hit[2]};
    c.normal = {n[0], n[1], n[2]};
    c.gb = {{mat[4], mat[5], mat[6]}, {0, 0, 0}, 0.0f, 0.0f, 0.0f, mat[0], mat[2], mat[1], mat[3]};
    MediaCtx m;
    m.inside = inside != 0;
    m.depth = depth;
    m.vertex = {vertex[0], vertex[1], vertex[2]};
    float lb[4];
    QueryLambda(lambda, lb);
    Rng rng = Rng::seeded(seed, 7u);
    const float3 rd{d[0], d[1], d[2]};
    const SampleCtx sc = SampleBSDF(lb, lambda, rng, AP_MAT_GLASS, rd, c, m);
    const float brdf = EvaluateBSDFWithCos(lb, lambda, AP_MAT_GLASS, rd, sc.next, c, m);
    out[0] = sc.next.x; out[1] = sc.next.y; out[2] = sc.next.z;
    out[3] = sc.pdf;
    out[4] = brdf;
    out[5] = sc.inside ? 1.0f : 0.0f;
    out[6] = sc.delta ? 1.0f : 0.0f;
}

// ---- pins of the BSDFs and light samplers (tests/test_bsdf_cpu.py) ----
void or_plant(int which) { g_plant = which; }

// SampleGGXBoundedVNDF (sample.cuh:179-227) for n (V, N, rnd) triples:
// out[4i..4i+3] = direction, pdf.
void or_ggx_vndf_sample(const float* V, const float* N, float alpha, const float* rnd, size_t n, float* out) {
    for (size_t i = 0; i < n; ++i) {
        const Sample s = SampleGGXBoundedVNDF(float3{V[3 * i], V[3 * i + 1], V[3 * i + 2]},
                                              float3{N[3 * i], N[3 * i + 1], N[3 * i + 2]}, alpha,
                                              float3{rnd[3 * i], rnd[3 * i + 1], rnd[3 * i + 2]});
        out[4 * i] = s.dir.x; out[4 * i + 1] = s.dir.y; out[4 * i + 2] = s.dir.z; out[4 * i + 3] = s.pdf;
    }
}

// GGXBoundedVNDF_PDF (sample.cuh:229-248) of n directions L for one (V, N).
void or_ggx_vndf_pdf(const float* V, const float* N, const float* L, size_t n, float alpha, float* out) {
    const float3 v{V[0], V[1], V[2]}, nn{N[0], N[1], N[2]};
    for (size_t i = 0; i < n; ++i) out[i] = GGXBoundedVNDF_PDF(v, nn, float3{L[3 * i], L[3 * i + 1], L[3 * i + 2]}, alpha);
}

static Ctx PinCtx(const float* mat, const float* N) {
    Ctx c;
    c.hit = true;
    c.normal = {N[0], N[1], N[2]};
    c.gb = {{mat[0], mat[1], mat[2]}, {0, 0, 0}, mat[3], mat[4], mat[5], 1.5f, 0.0f, 0.0f, 0.0f};
    return c;
}

// EvaluateBSDFWithCos and BSDF_PDF (material.cuh) of a Lambert or GGX
// surface at wavelength lambda for n directions L; mat = albedo rgb,
// specular, metallic, roughness.
void or_bsdf_eval(int model, const float* mat, float lambda, const float* rayDir, const float* N, const float* L,
                  size_t n, float* f, float* pdf) {
    float lb[4];
    QueryLambda(lambda, lb);
    const Ctx c = PinCtx(mat, N);
    const float3 rd{rayDir[0], rayDir[1], rayDir[2]};
    const MediaCtx m;
    for (size_t i = 0; i < n; ++i) {
        const float3 l{L[3 * i], L[3 * i + 1], L[3 * i + 2]};
        f[i] = EvaluateBSDFWithCos(lb, lambda, model, rd, l, c, m);
        pdf[i] = BSDF_PDF(lb, model, c.gb, c.normal, rd, l);
    }
}

// SampleBSDF (material.cuh) n times from PCG32(seed, 11): out[5i..5i+4] =
// direction, pdf, terminate flag.
void or_bsdf_sample(int model, const float* mat, float lambda, const float* rayDir, const float* N, uint64_t seed,
                    size_t n, float* out) {
    float lb[4];
    QueryLambda(lambda, lb);
    const Ctx c = PinCtx(mat, N);
    const float3 rd{rayDir[0], rayDir[1], rayDir[2]};
    const MediaCtx m;
    Rng rng = Rng::seeded(seed, 11u);
    for (size_t i = 0; i < n; ++i) {
        const SampleCtx s = SampleBSDF(lb, lambda, rng, model, rd, c, m);
        out[5 * i] = s.next.x; out[5 * i + 1] = s.next.y; out[5 * i + 2] = s.next.z;
        out[5 * i + 3] = s.pdf;
        out[5 * i + 4] = s.terminate ? 1.0f : 0.0f;
    }
}

// NEESampleLight (nee.cuh:330-377) over the scene's light table from each of
// np shading points, one draw each from PCG32(seed, i): out[10i..10i+9] =
// light point, light normal, light object, the sampler's solid-angle pdf
// (NEELightPDFOmegaFromArea, nee.cuh:48-52), NEEPDFOmegaFromHit at the
// sampled point (nee.cuh:313-327), valid flag.
void or_nee_sample(const ap_scene_object* objs, size_t n, const float* p, size_t np, uint64_t seed, float* out) {
    Scene S{objs, n, nullptr, {}, std::vector<int>(n, -1), AP_TRIANGLE_MATERIAL_OBJECT, 0.0f};
    BuildLights(S);
    for (size_t i = 0; i < np; ++i) {
        const float3 x{p[3 * i], p[3 * i + 1], p[3 * i + 2]};
        Rng rng = Rng::seeded(seed, i);
        const LightSample ls = NEESampleLight(S, x, rng);
        float* o = out + 10 * i;
        for (int k = 0; k < 10; ++k) o[k] = 0.0f;
        o[6] = -1.0f;
        if (!ls.valid) continue;
        o[0] = ls.point.x; o[1] = ls.point.y; o[2] = ls.point.z;
        o[3] = ls.normal.x; o[4] = ls.normal.y; o[5] = ls.normal.z;
        o[6] = static_cast<float>(ls.object);
        o[7] = AreaToSolidAnglePDF(x, ls.point, ls.normal, ls.pFacet * ls.pA);
        o[8] = NEEPDFOmegaFromHit(objs[ls.object], S.lights[S.light_index[ls.object]], x, ls.point, ls.normal);
        o[9] = 1.0f;
    }
}

// NEEPDFOmegaFromHit of light object `light` for m (shading point, light
// point, light normal) triples (9 floats each).
void or_nee_pdf(const ap_scene_object* objs, size_t n, int light, const float* q, size_t m, float* out) {
    Scene S{objs, n, nullptr, {}, std::vector<int>(n, -1), AP_TRIANGLE_MATERIAL_OBJECT, 0.0f};
    BuildLights(S);
    const int li = S.light_index[light];
    for (size_t i = 0; i < m; ++i) {
        const float* t = q + 9 * i;
        out[i] = li < 0 ? 0.0f
                        : NEEPDFOmegaFromHit(objs[light], S.lights[li], float3{t[0], t[1], t[2]},
                                             float3{t[3], t[4], t[5]}, float3{t[6], t[7], t[8]});
    }
}

// Albedo texture images for renders (index k = handle k + 1): n RGBA8
// images of dims[2k] x dims[2k+1] texels concatenated in rgba.
void or_texture_set(const uint8_t* rgba, const int32_t* dims, size_t n) {
    g_textures.clear();
    size_t off = 0;
    for (size_t k = 0; k < n; ++k) {
        OrTex t;
        t.w = dims[2 * k];
        t.h = dims[2 * k + 1];
        t.texels.resize(static_cast<size_t>(t.w) * t.h);
        std::memcpy(t.texels.data(), rgba + off, t.texels.size() * 4);
        off += t.texels.size() * 4;
        g_textures.push_back(std::move(t));
    }
}

void or_texture_sample(int32_t index, const float* uv, size_t n, float* out) {
    for (size_t i = 0; i < n; ++i) TexSample(g_textures[index], uv[2 * i], uv[2 * i + 1], out + 4 * i);
}

}  // extern "C"

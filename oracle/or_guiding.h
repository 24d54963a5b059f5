// ORACLE — test infrastructure only: CPU restatement of the reference's
// directional path guiding, the checker for the core's guiding kernels.
//
//   face / voxel binning                 core/guiding/guiding_bin.cuh:10-114
//   Clarberg equal-area mapping          core/guiding/octahedron.cuh:13-145
//   probe texture (Texture<4>): texel,   core/guiding/octahedron.cuh:203-516
//     SampleOctahedron(+Hemisphere), OctahedronPDF, GenerateMipmaps
//   GetPathGuidingProb                   core/render/render.cuh:198-203
//   QueryProbe                           core/render/render.cuh:434-458
//   per-frame build                      core/scene.cu:583-744 over
//     FilterValidSamples, UniqueVoxelGridsFromValid, GetMinMax,
//     CopyOldIntoNew, FilterUnallocatedGrids, ScatterAssignIndices,
//     AccumulateSamplesToOctaLeaf, MergeProbeKernel (guiding.cuh:70-517)
//
// Deviations shared with the core: the per-texel sums of a frame are taken in
// sample order (pixel-major, then collection slot) instead of by float
// atomics in arbitrary order; powf(x, 0.5) is sqrtf; the index volume holds
// int32 probe indices (-1 = none) instead of size_t; a first frame whose
// samples all fall in voxel (0,0,0) still gets a volume (the reference skips
// the allocation there and never indexes those probes).
#pragma once

#include <algorithm>
#include <cstdint>
#include <vector>

#include "aperture/ap_layout.h"
#include "ap_fmath.h"
#include "or_math.h"

namespace orc {

constexpr int PG_L = 4;                  // PATH_GUIDING_RESOLUTION_LEVEL
constexpr int PG_RES = 8;                // 1 << (L - 1)
constexpr int PG_RES2 = 64;
constexpr int PG_TOTAL = 85;             // 64 + 16 + 4 + 1
constexpr int PG_FACES = 20;             // PATH_GUIDING_ICOSAHEDRON
constexpr int PG_COLLECT = 5;            // PATH_GUIDING_COLLECT_DEPTH
constexpr float PG_INV_UNIT = 20.0f;     // 1 / PATH_GUIDING_RESOLUTION_UNIT (0.05)
constexpr float PG_PI = 3.141592654f;    // CUDART_PI_F
constexpr float PG_MIN_BRIGHT = 1.0e-5f; // PATH_GUIDING_COLLECT_MIN_BRIGHTNESS
constexpr float PG_MAX_BRIGHT = 1.0e5f;  // PATH_GUIDING_COLLECT_MAX_BRIGHTNESS
constexpr float PG_PROBE_MIN = 1.0e-5f;  // PATH_GUIDING_PROBE_MIN_BRIGHTNESS
constexpr uint64_t PG_MAX_COUNT = 1000000ull;
constexpr float PG_QUERY_EPS = 1.0e-6f;  // PROBES_QUERY_EPS

struct int3_ { int x, y, z; };
struct int4_ { int x, y, z, w; };

inline float3 FaceNormal20(int f) {
    static const float t[20][3] = {
        {-0.577350269f, 0.577350269f, 0.577350269f}, {0.000000000f, 0.934172359f, 0.356822090f},
        {0.000000000f, 0.934172359f, -0.356822090f}, {-0.577350269f, 0.577350269f, -0.577350269f},
        {-0.934172359f, 0.356822090f, 0.000000000f}, {0.577350269f, 0.577350269f, 0.577350269f},
        {-0.356822090f, 0.000000000f, 0.934172359f}, {-0.934172359f, -0.356822090f, 0.000000000f},
        {-0.356822090f, 0.000000000f, -0.934172359f}, {0.577350269f, 0.577350269f, -0.577350269f},
        {0.577350269f, -0.577350269f, 0.577350269f}, {0.000000000f, -0.934172359f, 0.356822090f},
        {0.000000000f, -0.934172359f, -0.356822090f}, {0.577350269f, -0.577350269f, -0.577350269f},
        {0.934172359f, -0.356822090f, 0.000000000f}, {0.356822090f, 0.000000000f, 0.934172359f},
        {-0.577350269f, -0.577350269f, 0.577350269f}, {-0.577350269f, -0.577350269f, -0.577350269f},
        {0.356822090f, 0.000000000f, -0.934172359f}, {0.934172359f, 0.356822090f, 0.000000000f}};
    return {t[f][0], t[f][1], t[f][2]};
}

// Get20FaceIndex (guiding_bin.cuh:66-84): the face normal with the largest dot.
inline int Get20FaceIndex(float3 n) {
    const float3 d = normalize(n);
    int best = 0;
    float bestDot = -1e30f;
    for (int f = 0; f < PG_FACES; ++f) {
        const float3 c = FaceNormal20(f);
        const float v = d.x * c.x + d.y * c.y + d.z * c.z;
        if (v > bestDot) { bestDot = v; best = f; }
    }
    return best;
}

// GetVoxelGrid (guiding_bin.cuh:102-114).
inline int4_ GetVoxelGrid(float3 p, float3 dir) {
    const float3 s = p * PG_INV_UNIT;
    return {static_cast<int>(std::floor(s.x)), static_cast<int>(std::floor(s.y)), static_cast<int>(std::floor(s.z)),
            Get20FaceIndex(dir)};
}

inline float3 Reflect(float3 v, float3 n) { return v - (2.0f * dot(v, n)) * n; }

// UVtoSphereClarberg (octahedron.cuh:13-84).
inline float3 UVtoSphere(float px, float py) {
    const float u = 2.0f * px - 1.0f;
    const float v = 2.0f * py - 1.0f;
    const float a = v + u;
    const float b = v - u;
    float r, phi4, zsign;
    if (v >= 0.0f) {
        if (u >= 0.0f) {
            if (a <= 1.0f) { r = a; zsign = 1.0f; phi4 = (r == 0.0f ? 0.0f : v / r); }
            else { r = 2.0f - a; zsign = -1.0f; phi4 = (1.0f - u) / r; }
        } else {
            if (b <= 1.0f) { r = b; zsign = 1.0f; phi4 = 1.0f - u / r; }
            else { r = 2.0f - b; zsign = -1.0f; phi4 = 1.0f + (1.0f - v) / r; }
        }
    } else {
        if (u < 0.0f) {
            if (a >= -1.0f) { r = -a; zsign = 1.0f; phi4 = 2.0f - v / r; }
            else { r = 2.0f + a; zsign = -1.0f; phi4 = 2.0f + (1.0f + u) / r; }
        } else {
            if (b >= -1.0f) { r = -b; zsign = 1.0f; phi4 = 3.0f + u / r; }
            else { r = 2.0f + b; zsign = -1.0f; phi4 = 3.0f + (1.0f + v) / r; }
        }
    }
    if (r == 0.0f) phi4 = 0.0f;
    const float r2 = r * r;
    const float phi = phi4 * (0.5f * PG_PI);
    const float sint = r * std::sqrt(std::fmax(0.0f, 2.0f - r2));
    const float cost = zsign * (1.0f - r2);
    float s, c;
    apfm::sincos(phi, s, c);
    return {sint * c, sint * s, cost};
}

// SphereToUVClarberg (octahedron.cuh:87-145), returns [0,1]^2.
inline void SphereToUV(float3 d, float& ou, float& ov) {
    const float phi2_over_pi = apfm::atan2(d.y, d.x) * (2.0f / PG_PI);
    float u, v;
    if (d.z < 0.0f) {
        const float r = std::sqrt(std::fmax(0.0f, 1.0f + d.z));
        if (phi2_over_pi >= 0.0f) {
            if (phi2_over_pi <= 1.0f) { u = 1.0f - r * phi2_over_pi; v = 2.0f - r - u; }
            else { u = r * (2.0f - phi2_over_pi) - 1.0f; v = 2.0f - r + u; }
        } else {
            if (phi2_over_pi >= -1.0f) { u = r * phi2_over_pi + 1.0f; v = r - 2.0f + u; }
            else { u = r * (2.0f + phi2_over_pi) - 1.0f; v = r - 2.0f - u; }
        }
    } else {
        const float r = std::sqrt(std::fmax(0.0f, 1.0f - d.z));
        if (phi2_over_pi >= 0.0f) {
            if (phi2_over_pi < 1.0f) { v = r * phi2_over_pi; u = r - v; }
            else { v = r * (2.0f - phi2_over_pi); u = v - r; }
        } else {
            if (phi2_over_pi > -1.0f) { v = r * phi2_over_pi; u = r + v; }
            else { v = -r * (2.0f + phi2_over_pi); u = -(r + v); }
        }
    }
    ou = 0.5f * (u + 1.0f);
    ov = 0.5f * (v + 1.0f);
}

constexpr float PG_JACOBIAN = 4.0f * PG_PI;  // JacobianUVToOctahedronSphereDirection (Clarberg)

inline int LevelOffset(int level) {
    int offset = 0, res = PG_RES;
    for (int i = 0; i < level; ++i) { offset += res * res; res >>= 1; }
    return offset;
}

inline const ap_float4& Texel(const ap_octahedron& p, int x, int y, int level) {
    const int res = 1 << (PG_L - 1 - level);
    return p.data[LevelOffset(level) + y * res + x];
}

inline float Dot3(const ap_float4& t, float3 q) { return t.x * q.x + t.y * q.y + t.z * q.z; }

// OctahedronPDF (octahedron.cuh:373-401), 3-channel probe.
inline float OctahedronPDF(const ap_octahedron& p, float3 dir, float3 q) {
    float u, v;
    SphereToUV(dir, u, v);
    int ix = static_cast<int>(std::floor(u * PG_RES));
    int iy = static_cast<int>(std::floor(v * PG_RES));
    ix = std::max(0, std::min(PG_RES - 1, ix));
    iy = std::max(0, std::min(PG_RES - 1, iy));
    const float wLeaf = Dot3(Texel(p, ix, iy, 0), q);
    const float wAvg = Dot3(Texel(p, 0, 0, PG_L - 1), q);
    const float invG = static_cast<float>(1.0 / static_cast<double>(PG_JACOBIAN));
    if (!(wAvg > 0.0f)) return invG;
    return std::fmax((wLeaf / wAvg) * invG, 1.0e-12f);
}

inline float3 ReflectHemisphereNormal(float3 w, float3 n) {
    return normalize(w) - (2.0f * dot(normalize(n), normalize(w))) * normalize(n);
}

inline float OctahedronHemispherePDF(const ap_octahedron& p, float3 dir, float3 n, float3 q) {
    if (dot(dir, n) <= 0.0f) return 0.0f;
    const float3 mirror = ReflectHemisphereNormal(dir, n);
    return OctahedronPDF(p, dir, q) + OctahedronPDF(p, mirror, q);
}

struct GuideSample { float3 dir; float pdf; };

// SampleOctahedron (octahedron.cuh:271-337): mip descent, then a uniform
// point in the chosen leaf texel.
template <class R>
inline GuideSample SampleOctahedron(const ap_octahedron& p, R& rng, float3 q) {
    int cx = 0, cy = 0;
    float pTexel = 1.0f;
    for (int level = PG_L - 1; level > 0; --level) {
        const float weightSum = 4.0f * Dot3(Texel(p, cx, cy, level), q);
        const int nx = cx << 1, ny = cy << 1;
        const float rand = rng.rand1();
        if (weightSum <= 0.0f) {
            const int pick = std::min(3, static_cast<int>(rand * 4.0f));
            cx = nx + (pick & 1);
            cy = ny + (pick >> 1);
            pTexel *= 0.25f;
        } else {
            float cdf = 0.0f;
            bool done = false;
            for (int x = 0; x < 2 && !done; ++x)
                for (int y = 0; y < 2 && !done; ++y) {
                    const float w = Dot3(Texel(p, nx + x, ny + y, level - 1), q) / weightSum;
                    cdf += w;
                    if (rand <= cdf || (x == 1 && y == 1)) {
                        cx = nx + x;
                        cy = ny + y;
                        pTexel *= w;
                        done = true;
                    }
                }
        }
    }
    const float r1 = rng.rand1(), r2 = rng.rand1();
    const float su = static_cast<float>(cx) + (1.0f - r1);
    const float sv = static_cast<float>(cy) + (1.0f - r2);
    const float uu = std::fmin(std::fmax(su / static_cast<float>(PG_RES), 0.0f), 1.0f);
    const float vv = std::fmin(std::fmax(sv / static_cast<float>(PG_RES), 0.0f), 1.0f);
    GuideSample s;
    s.dir = UVtoSphere(uu, vv);
    s.pdf = std::fmax(pTexel * static_cast<float>(PG_RES2) / PG_JACOBIAN, 1.0e-12f);
    return s;
}

template <class R>
inline GuideSample SampleOctahedronHemisphere(const ap_octahedron& p, R& rng, float3 n, float3 q) {
    GuideSample s = SampleOctahedron(p, rng, q);
    const float3 mirror = ReflectHemisphereNormal(s.dir, n);
    s.pdf += OctahedronPDF(p, mirror, q);
    if (dot(s.dir, n) < 0.0f) s.dir = mirror;
    return s;
}

// GetPathGuidingProb (render.cuh:198-203) with powf(x, 0.5) = sqrtf(x).
inline float GuidingAlpha(int frame) {
    const float t = std::sqrt(std::fmin(1.0f, static_cast<float>(frame) / 64.0f));
    return saturate_(lerp(0.0f, 0.5f, t));
}

// GenerateMipmaps (octahedron.cuh:436-496).
inline void GenerateMipmaps(ap_octahedron& p) {
    for (int i = 0; i < PG_RES2; ++i) {
        ap_float4& t = p.data[i];
        t.x = std::fmax(t.x, PG_PROBE_MIN); t.y = std::fmax(t.y, PG_PROBE_MIN);
        t.z = std::fmax(t.z, PG_PROBE_MIN); t.w = std::fmax(t.w, PG_PROBE_MIN);
    }
    int srcRes = PG_RES, srcOff = 0, nextOff = PG_RES2;
    for (int level = 0; level < PG_L - 1; ++level) {
        const int dstRes = srcRes >> 1, dstOff = nextOff;
        for (int y = 0; y < dstRes; ++y)
            for (int x = 0; x < dstRes; ++x) {
                const ap_float4 A = p.data[srcOff + (2 * y) * srcRes + 2 * x];
                const ap_float4 B = p.data[srcOff + (2 * y) * srcRes + 2 * x + 1];
                const ap_float4 C = p.data[srcOff + (2 * y + 1) * srcRes + 2 * x];
                const ap_float4 D = p.data[srcOff + (2 * y + 1) * srcRes + 2 * x + 1];
                ap_float4& o = p.data[dstOff + y * dstRes + x];
                o.x = (((A.x + B.x) + C.x) + D.x) * 0.25f;
                o.y = (((A.y + B.y) + C.y) + D.y) * 0.25f;
                o.z = (((A.z + B.z) + C.z) + D.z) * 0.25f;
                o.w = (((A.w + B.w) + C.w) + D.w) * 0.25f;
            }
        srcOff = dstOff;
        srcRes = dstRes;
        nextOff = srcOff + srcRes * srcRes;
    }
    p.data[PG_TOTAL] = ap_float4{0.0f, 0.0f, 0.0f, 0.0f};
}

inline bool Int4Less(const int4_& a, const int4_& b) {
    if (a.x != b.x) return a.x < b.x;
    if (a.y != b.y) return a.y < b.y;
    if (a.z != b.z) return a.z < b.z;
    return a.w < b.w;
}

// Probes, index volume and its inclusive voxel bounds, persistent across frames.
struct GuideState {
    std::vector<ap_octahedron> probes;
    std::vector<int32_t> volume;        // [z][y][x][face], -1 = no probe
    int3_ start{0, 0, 0}, end{0, 0, 0}, size{0, 0, 0};

    const ap_octahedron* Query(float3 pos, float3 dir) const {  // QueryProbe (render.cuh:434-458)
        if (probes.empty() || size.x == 0) return nullptr;
        const int4_ v = GetVoxelGrid(pos, dir);
        const int lx = v.x - start.x, ly = v.y - start.y, lz = v.z - start.z;
        if (v.w < 0 || v.w >= PG_FACES) return nullptr;
        if (lx < 0 || ly < 0 || lz < 0 || lx >= size.x || ly >= size.y || lz >= size.z) return nullptr;
        const size_t lin = static_cast<size_t>(lx) + static_cast<size_t>(size.x) *
                           (static_cast<size_t>(ly) + static_cast<size_t>(size.y) * static_cast<size_t>(lz));
        const int32_t idx = volume[lin * PG_FACES + static_cast<size_t>(v.w)];
        return idx >= 0 ? &probes[static_cast<size_t>(idx)] : nullptr;
    }

    int32_t Lookup(const ap_int4& v) const {
        const int lx = v.x - start.x, ly = v.y - start.y, lz = v.z - start.z;
        if (v.w < 0 || v.w >= PG_FACES) return -1;
        if (lx < 0 || ly < 0 || lz < 0 || lx >= size.x || ly >= size.y || lz >= size.z) return -1;
        const size_t lin = static_cast<size_t>(lx) + static_cast<size_t>(size.x) *
                           (static_cast<size_t>(ly) + static_cast<size_t>(size.y) * static_cast<size_t>(lz));
        return volume[lin * PG_FACES + static_cast<size_t>(v.w)];
    }

    // UpdatePathGuidingSample (scene.cu:583-744) over the frame's sample
    // slots in pixel-major, slot-minor order.
    void Update(const std::vector<ap_path_guiding_sample>& slots) {
        std::vector<const ap_path_guiding_sample*> valid;
        for (const auto& s : slots)
            if (s.valid) valid.push_back(&s);
        if (valid.empty()) return;
        std::vector<int4_> keys;
        keys.reserve(valid.size());
        for (const auto* s : valid) keys.push_back({s->voxel_grid.x, s->voxel_grid.y, s->voxel_grid.z, s->voxel_grid.w});
        std::sort(keys.begin(), keys.end(), Int4Less);
        keys.erase(std::unique(keys.begin(), keys.end(),
                               [](const int4_& a, const int4_& b) {
                                   return a.x == b.x && a.y == b.y && a.z == b.z && a.w == b.w;
                               }),
                   keys.end());
        int3_ mn{keys[0].x, keys[0].y, keys[0].z}, mx = mn;
        for (const auto& k : keys) {
            mn.x = std::min(mn.x, k.x); mn.y = std::min(mn.y, k.y); mn.z = std::min(mn.z, k.z);
            mx.x = std::max(mx.x, k.x); mx.y = std::max(mx.y, k.y); mx.z = std::max(mx.z, k.z);
        }
        const int3_ ds{std::min(start.x, mn.x), std::min(start.y, mn.y), std::min(start.z, mn.z)};
        const int3_ de{std::max(end.x, mx.x), std::max(end.y, mx.y), std::max(end.z, mx.z)};
        const bool changed = ds.x != start.x || ds.y != start.y || ds.z != start.z || de.x != end.x ||
                             de.y != end.y || de.z != end.z;
        if (changed || size.x == 0) {
            const int3_ nd{de.x - ds.x + 1, de.y - ds.y + 1, de.z - ds.z + 1};
            std::vector<int32_t> nv(static_cast<size_t>(nd.x) * nd.y * nd.z * PG_FACES, -1);
            if (size.x > 0) {  // CopyOldIntoNew (guiding.cuh:171-244)
                for (int z = 0; z < size.z; ++z)
                    for (int y = 0; y < size.y; ++y)
                        for (int x = 0; x < size.x; ++x) {
                            const int tx = x + start.x - ds.x, ty = y + start.y - ds.y, tz = z + start.z - ds.z;
                            const size_t src = static_cast<size_t>(x) + static_cast<size_t>(size.x) *
                                               (static_cast<size_t>(y) + static_cast<size_t>(size.y) * z);
                            const size_t dst = static_cast<size_t>(tx) + static_cast<size_t>(nd.x) *
                                               (static_cast<size_t>(ty) + static_cast<size_t>(nd.y) * tz);
                            for (int f = 0; f < PG_FACES; ++f) nv[dst * PG_FACES + f] = volume[src * PG_FACES + f];
                        }
            }
            volume.swap(nv);
            start = ds;
            end = de;
            size = nd;
        }
        // New probes for unindexed keys, numbered in key order (FilterUnallocatedGrids + ScatterAssignIndices).
        const size_t base = probes.size();
        size_t added = 0;
        for (const auto& k : keys) {
            const ap_int4 v{k.x, k.y, k.z, k.w};
            if (Lookup(v) >= 0) continue;
            const int lx = k.x - start.x, ly = k.y - start.y, lz = k.z - start.z;
            const size_t lin = static_cast<size_t>(lx) + static_cast<size_t>(size.x) *
                               (static_cast<size_t>(ly) + static_cast<size_t>(size.y) * static_cast<size_t>(lz));
            volume[lin * PG_FACES + static_cast<size_t>(k.w)] = static_cast<int32_t>(base + added);
            ++added;
        }
        ap_octahedron zero;
        std::memset(&zero, 0, sizeof(zero));
        probes.resize(base + added, zero);
        // AccumulateSamplesToOctaLeaf (guiding.cuh:349-405), summed in sample order.
        std::vector<ap_octahedron> temp(probes.size(), zero);
        for (const auto* s : valid) {
            const int32_t idx = Lookup(s->voxel_grid);
            if (idx < 0) continue;
            const float3 dir = normalize(float3{s->direction.x, s->direction.y, s->direction.z});
            float u, v;
            SphereToUV(normalize(dir), u, v);
            int ix = static_cast<int>(std::floor(u * PG_RES));
            int iy = static_cast<int>(std::floor(v * PG_RES));
            ix = std::max(0, std::min(PG_RES - 1, ix));
            iy = std::max(0, std::min(PG_RES - 1, iy));
            ap_octahedron& t = temp[static_cast<size_t>(idx)];
            ap_float4& leaf = t.data[iy * PG_RES + ix];
            leaf.x += s->radiance.x * PG_JACOBIAN;
            leaf.y += s->radiance.y * PG_JACOBIAN;
            leaf.z += s->radiance.z * PG_JACOBIAN;
            t.count += 1;
        }
        // MergeProbeKernel (guiding.cuh:436-506): FP64 count-weighted mean, then mipmaps.
        for (size_t i = 0; i < probes.size(); ++i) {
            const ap_octahedron& nw = temp[i];
            if (nw.count == 0) continue;
            ap_octahedron& od = probes[i];
            const double invDen = 1.0 / static_cast<double>(od.count + nw.count);
            const double alpha = static_cast<double>(od.count) * invDen;
            for (int k = 0; k < PG_RES2; ++k) {
                ap_float4& o = od.data[k];
                const ap_float4& n = nw.data[k];
                o.x = static_cast<float>(static_cast<double>(o.x) * alpha + static_cast<double>(n.x) * invDen);
                o.y = static_cast<float>(static_cast<double>(o.y) * alpha + static_cast<double>(n.y) * invDen);
                o.z = static_cast<float>(static_cast<double>(o.z) * alpha + static_cast<double>(n.z) * invDen);
            }
            od.count = std::min(PG_MAX_COUNT, od.count + nw.count);
            GenerateMipmaps(od);
        }
    }
};

}  // namespace orc

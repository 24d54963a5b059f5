// Directional path guiding: probe math shared by the shading kernel and the
// per-frame probe build (host and device).
//
// A probe is the reference's Octahedron = pg::Texture<4> (1392 B,
// include/aperture/ap_layout.h): an 8x8 leaf of XYZ radiance over the
// Clarberg equal-area square, its 4x4, 2x2 and 1x1 mips, a pad texel and a
// sample count. Probes live in voxels of 0.05 units times 20 icosahedral
// faces of the reflected view direction; an int32 index volume over the
// voxel bounds maps (voxel, face) to a probe (-1 = none).
//
//   face / voxel binning        core/guiding/guiding_bin.cuh:10-114
//   Clarberg mapping            core/guiding/octahedron.cuh:13-145
//   texel, pdf, sampling, mips  core/guiding/octahedron.cuh:203-516
//   GetPathGuidingProb          core/render/render.cuh:198-203 (powf(x, 0.5) = sqrtf)
//   CollectRadiance             core/render/render.cuh:353-432
#pragma once

#include "aperture/ap_layout.h"
#include "ap_fmath.h"
#include "ap_pcg.h"
#include "ap_vec.h"

namespace ap {

constexpr int PG_L = 4;
constexpr int PG_RES = 8;
constexpr int PG_RES2 = 64;
constexpr int PG_TOTAL = 85;
constexpr int PG_FACES = 20;
constexpr int PG_COLLECT = 5;
constexpr float PG_INV_UNIT = 20.0f;      // 1 / 0.05
constexpr float PG_PI = 3.141592654f;     // CUDART_PI_F
constexpr float PG_JACOBIAN = 4.0f * PG_PI;
constexpr float PG_MIN_BRIGHT = 1.0e-5f;
constexpr float PG_MAX_BRIGHT = 1.0e5f;
constexpr float PG_PROBE_MIN = 1.0e-5f;
constexpr unsigned long long PG_MAX_COUNT = 1000000ull;
constexpr float PG_QUERY_EPS = 1.0e-6f;

AP_HD f3 pg_face_normal(int f) {
    switch (f) {
        case 0: return {-0.577350269f, 0.577350269f, 0.577350269f};
        case 1: return {0.000000000f, 0.934172359f, 0.356822090f};
        case 2: return {0.000000000f, 0.934172359f, -0.356822090f};
        case 3: return {-0.577350269f, 0.577350269f, -0.577350269f};
        case 4: return {-0.934172359f, 0.356822090f, 0.000000000f};
        case 5: return {0.577350269f, 0.577350269f, 0.577350269f};
        case 6: return {-0.356822090f, 0.000000000f, 0.934172359f};
        case 7: return {-0.934172359f, -0.356822090f, 0.000000000f};
        case 8: return {-0.356822090f, 0.000000000f, -0.934172359f};
        case 9: return {0.577350269f, 0.577350269f, -0.577350269f};
        case 10: return {0.577350269f, -0.577350269f, 0.577350269f};
        case 11: return {0.000000000f, -0.934172359f, 0.356822090f};
        case 12: return {0.000000000f, -0.934172359f, -0.356822090f};
        case 13: return {0.577350269f, -0.577350269f, -0.577350269f};
        case 14: return {0.934172359f, -0.356822090f, 0.000000000f};
        case 15: return {0.356822090f, 0.000000000f, 0.934172359f};
        case 16: return {-0.577350269f, -0.577350269f, 0.577350269f};
        case 17: return {-0.577350269f, -0.577350269f, -0.577350269f};
        case 18: return {0.356822090f, 0.000000000f, -0.934172359f};
        default: return {0.934172359f, 0.356822090f, 0.000000000f};
    }
}

AP_HD int pg_face(f3 n) {
    const f3 d = normalize(n);
    int best = 0;
    float best_dot = -1e30f;
    for (int f = 0; f < PG_FACES; ++f) {
        const f3 c = pg_face_normal(f);
        const float v = d.x * c.x + d.y * c.y + d.z * c.z;
        if (v > best_dot) {
            best_dot = v;
            best = f;
        }
    }
    return best;
}

struct PgVoxel { int x, y, z, f; };

AP_HD PgVoxel pg_voxel(f3 p, f3 dir) {
    const f3 s = p * PG_INV_UNIT;
    return {static_cast<int>(floorf(s.x)), static_cast<int>(floorf(s.y)), static_cast<int>(floorf(s.z)), pg_face(dir)};
}

AP_HD f3 pg_reflect(f3 v, f3 n) { return v - (2.0f * dot(v, n)) * n; }

AP_HD f3 pg_uv_to_sphere(float px, float py) {
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
    const float sint = r * sqrtf(fmaxf(0.0f, 2.0f - r2));
    const float cost = zsign * (1.0f - r2);
    float s, c;
    apfm::sincos(phi, s, c);
    return {sint * c, sint * s, cost};
}

AP_HD void pg_sphere_to_uv(f3 d, float& ou, float& ov) {
    const float phi2_over_pi = apfm::atan2(d.y, d.x) * (2.0f / PG_PI);
    float u, v;
    if (d.z < 0.0f) {
        const float r = sqrtf(fmaxf(0.0f, 1.0f + d.z));
        if (phi2_over_pi >= 0.0f) {
            if (phi2_over_pi <= 1.0f) { u = 1.0f - r * phi2_over_pi; v = 2.0f - r - u; }
            else { u = r * (2.0f - phi2_over_pi) - 1.0f; v = 2.0f - r + u; }
        } else {
            if (phi2_over_pi >= -1.0f) { u = r * phi2_over_pi + 1.0f; v = r - 2.0f + u; }
            else { u = r * (2.0f + phi2_over_pi) - 1.0f; v = r - 2.0f - u; }
        }
    } else {
        const float r = sqrtf(fmaxf(0.0f, 1.0f - d.z));
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

// Leaf texel of a direction (DirectionToLeafPixel, guiding.cuh:318-328).
AP_HD int pg_leaf_texel(f3 dir) {
    float u, v;
    pg_sphere_to_uv(dir, u, v);
    int ix = static_cast<int>(floorf(u * PG_RES));
    int iy = static_cast<int>(floorf(v * PG_RES));
    ix = ix < 0 ? 0 : (ix > PG_RES - 1 ? PG_RES - 1 : ix);
    iy = iy < 0 ? 0 : (iy > PG_RES - 1 ? PG_RES - 1 : iy);
    return iy * PG_RES + ix;
}

AP_HD int pg_level_offset(int level) {
    int offset = 0, res = PG_RES;
    for (int i = 0; i < level; ++i) {
        offset += res * res;
        res >>= 1;
    }
    return offset;
}

AP_HD float pg_dot(const ap_float4& t, f3 q) { return t.x * q.x + t.y * q.y + t.z * q.z; }

AP_HD const ap_float4& pg_texel(const ap_octahedron& p, int x, int y, int level) {
    const int res = 1 << (PG_L - 1 - level);
    return p.data[pg_level_offset(level) + y * res + x];
}

AP_HD float pg_pdf(const ap_octahedron& p, f3 dir, f3 q) {
    const int t = pg_leaf_texel(dir);
    const float w_leaf = pg_dot(p.data[t], q);
    const float w_avg = pg_dot(p.data[PG_TOTAL - 1], q);
    const float inv_g = static_cast<float>(1.0 / static_cast<double>(PG_JACOBIAN));
    if (!(w_avg > 0.0f)) return inv_g;
    return fmaxf((w_leaf / w_avg) * inv_g, 1.0e-12f);
}

AP_HD f3 pg_reflect_hemisphere(f3 w, f3 n) {
    return normalize(w) - (2.0f * dot(normalize(n), normalize(w))) * normalize(n);
}

AP_HD float pg_hemisphere_pdf(const ap_octahedron& p, f3 dir, f3 n, f3 q) {
    if (dot(dir, n) <= 0.0f) return 0.0f;
    const f3 mirror = pg_reflect_hemisphere(dir, n);
    return pg_pdf(p, dir, q) + pg_pdf(p, mirror, q);
}

struct PgSample {
    f3 dir;
    float pdf;
};

AP_HD PgSample pg_sample(const ap_octahedron& p, Pcg32& rng, f3 q) {
    int cx = 0, cy = 0;
    float p_texel = 1.0f;
    for (int level = PG_L - 1; level > 0; --level) {
        const float weight_sum = 4.0f * pg_dot(pg_texel(p, cx, cy, level), q);
        const int nx = cx << 1, ny = cy << 1;
        const float rand = pcg32_float(rng);
        if (weight_sum <= 0.0f) {
            int pick = static_cast<int>(rand * 4.0f);
            pick = pick < 3 ? pick : 3;
            cx = nx + (pick & 1);
            cy = ny + (pick >> 1);
            p_texel *= 0.25f;
        } else {
            float cdf = 0.0f;
            bool done = false;
            for (int x = 0; x < 2 && !done; ++x)
                for (int y = 0; y < 2 && !done; ++y) {
                    const float w = pg_dot(pg_texel(p, nx + x, ny + y, level - 1), q) / weight_sum;
                    cdf += w;
                    if (rand <= cdf || (x == 1 && y == 1)) {
                        cx = nx + x;
                        cy = ny + y;
                        p_texel *= w;
                        done = true;
                    }
                }
        }
    }
    const float r1 = pcg32_float(rng), r2 = pcg32_float(rng);
    const float su = static_cast<float>(cx) + (1.0f - r1);
    const float sv = static_cast<float>(cy) + (1.0f - r2);
    const float uu = fminf(fmaxf(su / static_cast<float>(PG_RES), 0.0f), 1.0f);
    const float vv = fminf(fmaxf(sv / static_cast<float>(PG_RES), 0.0f), 1.0f);
    PgSample s;
    s.dir = pg_uv_to_sphere(uu, vv);
    s.pdf = fmaxf(p_texel * static_cast<float>(PG_RES2) / PG_JACOBIAN, 1.0e-12f);
    return s;
}

AP_HD PgSample pg_sample_hemisphere(const ap_octahedron& p, Pcg32& rng, f3 n, f3 q) {
    PgSample s = pg_sample(p, rng, q);
    const f3 mirror = pg_reflect_hemisphere(s.dir, n);
    s.pdf += pg_pdf(p, mirror, q);
    if (dot(s.dir, n) < 0.0f) s.dir = mirror;
    return s;
}

AP_HD float pg_alpha(int frame) {
    const float t = sqrtf(fminf(1.0f, static_cast<float>(frame) / 64.0f));
    const float a = 0.0f * (1.0f - t) + 0.5f * t;
    return isnan(a) ? 0.0f : fminf(fmaxf(a, 0.0f), 1.0f);
}

// Probe-query weights: (1,1,1) at the camera vertex, then the normalised
// max(mask * XYZ(lambda), eps) (render.cuh:882-888).
AP_HD f3 pg_query(float mask, const float xyz[3], bool first) {
    if (first) return {1.0f, 1.0f, 1.0f};
    f3 q{fmaxf(xyz[0] * mask, PG_QUERY_EPS), fmaxf(xyz[1] * mask, PG_QUERY_EPS), fmaxf(xyz[2] * mask, PG_QUERY_EPS)};
    const float den = static_cast<float>(static_cast<double>(q.x + q.y + q.z) + 1.0e-6);
    return q / den;
}

// maskForProbeQuery with several wavelengths per path: the throughput's XYZ,
// sum_j xyz(l_j) mask_j, floored and normalised as above (one wavelength:
// exactly pg_query).
AP_HD f3 pg_query_xyz(f3 txyz, bool first) {
    if (first) return {1.0f, 1.0f, 1.0f};
    f3 q{fmaxf(txyz.x, PG_QUERY_EPS), fmaxf(txyz.y, PG_QUERY_EPS), fmaxf(txyz.z, PG_QUERY_EPS)};
    const float den = static_cast<float>(static_cast<double>(q.x + q.y + q.z) + 1.0e-6);
    return q / den;
}

AP_HD void pg_generate_mipmaps(ap_octahedron& p) {
    for (int i = 0; i < PG_RES2; ++i) {
        ap_float4& t = p.data[i];
        t.x = fmaxf(t.x, PG_PROBE_MIN);
        t.y = fmaxf(t.y, PG_PROBE_MIN);
        t.z = fmaxf(t.z, PG_PROBE_MIN);
        t.w = fmaxf(t.w, PG_PROBE_MIN);
    }
    int src_res = PG_RES, src_off = 0, next_off = PG_RES2;
    for (int level = 0; level < PG_L - 1; ++level) {
        const int dst_res = src_res >> 1, dst_off = next_off;
        for (int y = 0; y < dst_res; ++y)
            for (int x = 0; x < dst_res; ++x) {
                const ap_float4 A = p.data[src_off + (2 * y) * src_res + 2 * x];
                const ap_float4 B = p.data[src_off + (2 * y) * src_res + 2 * x + 1];
                const ap_float4 C = p.data[src_off + (2 * y + 1) * src_res + 2 * x];
                const ap_float4 D = p.data[src_off + (2 * y + 1) * src_res + 2 * x + 1];
                ap_float4& o = p.data[dst_off + y * dst_res + x];
                o.x = (((A.x + B.x) + C.x) + D.x) * 0.25f;
                o.y = (((A.y + B.y) + C.y) + D.y) * 0.25f;
                o.z = (((A.z + B.z) + C.z) + D.z) * 0.25f;
                o.w = (((A.w + B.w) + C.w) + D.w) * 0.25f;
            }
        src_off = dst_off;
        src_res = dst_res;
        next_off = src_off + src_res * src_res;
    }
    p.data[PG_TOTAL] = ap_float4{0.0f, 0.0f, 0.0f, 0.0f};
}

// Per-path radiance collection (CalculateRadiance's radianceSample* arrays),
// with room for four wavelengths per path (index 0 alone with one). A slot's
// radiance and throughput are 16-B aligned rows (one vector access each).
// Only `mask` and `valid` must start at zero: a slot's other fields are
// written when it opens, and only open or valid slots are read.
struct alignas(16) GuideRec {
    float sample[PG_COLLECT][4];
    float mask[PG_COLLECT][4];
    float dir[PG_COLLECT][3];
    float loc[PG_COLLECT][3];
    float vdir[PG_COLLECT][3];
    uint32_t valid;   // bit i: slot i collected radiance
    uint32_t pad[2];
};
static_assert(sizeof(GuideRec) == 352, "GuideRec is 352 B");

// Device view of the guiding state for the shading kernel.
struct GuideView {
    const ap_octahedron* probes;
    int32_t probe_count;
    const int32_t* volume;
    int32_t start[3];
    int32_t size[3];
    GuideRec* recs;
};

AP_HD const ap_octahedron* pg_query_probe(const GuideView& g, f3 pos, f3 dir) {
    if (g.probe_count <= 0 || !g.volume) return nullptr;
    const PgVoxel v = pg_voxel(pos, dir);
    const int lx = v.x - g.start[0], ly = v.y - g.start[1], lz = v.z - g.start[2];
    if (v.f < 0 || v.f >= PG_FACES) return nullptr;
    if (lx < 0 || ly < 0 || lz < 0 || lx >= g.size[0] || ly >= g.size[1] || lz >= g.size[2]) return nullptr;
    const size_t lin = static_cast<size_t>(lx) +
                       static_cast<size_t>(g.size[0]) * (static_cast<size_t>(ly) + static_cast<size_t>(g.size[1]) * lz);
    const int32_t idx = g.volume[lin * PG_FACES + static_cast<size_t>(v.f)];
    return idx >= 0 ? &g.probes[idx] : nullptr;
}

// Per-frame build (ap_guiding.hip).
struct GuideDevice;
// Path records for `sets` frame sets (1, or 2 for frames in flight two at a time).
int guide_prepare(GuideDevice** g, size_t n_paths, int sets = 1);
// Probes and index volume, with frame set 0's path records.
GuideView guide_view(GuideDevice* g);
// The path records of frame set 0 or 1.
GuideRec* guide_recs(GuideDevice* g, int set);
ap_path_guiding_sample* guide_slots(GuideDevice* g);
// Per path, the number of valid samples at the front of its slots (written
// with the slots by the accumulate kernels).
uint32_t* guide_nvalid(GuideDevice* g);
int guide_update(GuideDevice* g, size_t n_paths, hipStream_t stream);
// guide_update in two halves around its one wait, so that several scenes'
// builds can share a wait: begin enqueues up to the read-back of the counts
// the host needs; after the stream has been synchronised, end enqueues the
// rest.
// With capture, begin records no events (its launches go into a graph that
// is replayed later); after a replay the caller waits for the stream, calls
// guide_replayed, then guide_update_end with synced.
int guide_update_begin(GuideDevice* g, size_t n_paths, hipStream_t stream, bool capture = false);
int guide_update_end(GuideDevice* g, hipStream_t stream, bool synced = false);
void guide_replayed(GuideDevice* g);
// Whether the last update's check left the build to the host (a volume or
// probe growth, or a re-sort).
bool guide_needs_host(const GuideDevice* g);
int guide_reset(GuideDevice* g);
int guide_export(GuideDevice* g, ap_octahedron* probes, int32_t* volume, int32_t* bounds, int32_t* count);

}  // namespace ap

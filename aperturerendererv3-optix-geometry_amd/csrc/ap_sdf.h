// Signed-distance objects (core/object/sdf/): the torus and diamond
// distance functors, volumes of distances sampled on a grid, sphere tracing
// inside the object's box (RaymarchLocal) and the central-difference normal
// (sdfNormalUnit). Shared by the traversal kernel (hit distance), the
// integrator (hit position and normal), the volume fill kernel and the host
// (camera-inside test). Unit space is the object's box mapped to
// [-0.5, 0.5]^3.
#pragma once

#include <cstddef>
#include <cstdint>

#include "ap_fmath.h"
#include "ap_vec.h"

namespace ap {

// ESDFType (core/render/object.cuh:37-43).
enum : int32_t { SDF_TORUS = 0, SDF_VOLUME = 1, SDF_UNDEFINED = 2 };
// Functors a volume can be filled from (the reference's SDFTorusFunctor and
// SDFDiamondFunctor).
enum : int32_t { SDF_FUNCTOR_TORUS = 0, SDF_FUNCTOR_DIAMOND = 1 };

constexpr int kSdfSteps = 512;                // SDF_STEPS (config.cuh:75)
constexpr float kSdfVolumeStepScale = 0.9f;   // SDF_VOLUME_STEP_SCALE (config.cuh:76)
constexpr float kSdfNearestScale = 2.5f;      // SDF_VOLUME_USE_NEAREST_SCALE (config.cuh:77)
constexpr float kCudaPiF = 3.141592654f;      // CUDART_PI_F

// SDFVolumeInfo (sdf_volume.cuh:13-45): a dim[0] x dim[1] x dim[2] grid,
// x fastest, and 2.5 |1 / max(dim, 1)| below which the trilinear sample is used.
struct alignas(8) SdfVolume {
    const float* values;
    int32_t dim[3];
    float voxel_size_max;
};

AP_HD float sdf_voxel_size_max(const int32_t dim[3]) {
    const f3 inv{1.0f / fmaxf(static_cast<float>(dim[0]), 1.0f), 1.0f / fmaxf(static_cast<float>(dim[1]), 1.0f),
                 1.0f / fmaxf(static_cast<float>(dim[2]), 1.0f)};
    return length(inv) * kSdfNearestScale;
}

// SdfTorus (sdf_torus.cuh:15-19).
AP_HD float sdf_torus(f3 u, float R, float r) {
    const float qx = sqrtf(u.x * u.x + u.z * u.z) - R;
    return sqrtf(qx * qx + u.y * u.y) - r;
}

// PlaneSDF (sdf_diamond.cuh:15-19).
AP_HD float plane_sdf(f3 p, f3 n, float d) {
    const float len = fmaxf(sqrtf(n.x * n.x + n.y * n.y + n.z * n.z), 1e-8f);
    return (p.x * n.x + p.y * n.y + p.z * n.z - d) / len;
}

// SdfDiamondUnit (sdf_diamond.cuh:21-66) with the portable sin, cos, atan2.
AP_HD float sdf_diamond(f3 u, const float info[4]) {
    const float d = (info[0] != 0.0f) ? info[0] : 0.94f;
    const float bias = (info[1] != 0.0f) ? info[1] : 0.5f;
    const float table_z = (info[2] != 0.0f) ? info[2] : 0.3f;
    const float scale = (info[3] != 0.0f) ? info[3] : 1.0f;
    const float inv = 1.0f / scale;
    const f3 p{u.x * inv, u.y * inv, u.z * inv};
    constexpr float af2 = 4.0f / kCudaPiF;
    const float s = apfm::atan2(p.y, p.x);
    const float sf = floorf(s * af2 + bias) / af2;
    const float sf2 = floorf(s * af2) / af2;
    float csf, ssf, cs, ss, csf1, ssf1, csf2, ssf2, csf8, ssf8;
    apfm::sincos(sf, ssf, csf);
    apfm::sincos(s, ss, cs);
    apfm::sincos(sf + 0.21f, ssf1, csf1);
    apfm::sincos(sf - 0.21f, ssf2, csf2);
    apfm::sincos(sf2 + 0.393f, ssf8, csf8);
    float m = -1e20f;
    m = fmaxf(m, plane_sdf(p, f3{csf, ssf, 1.444f}, d));
    m = fmaxf(m, plane_sdf(p, f3{csf, ssf, -1.072f}, d));
    m = fmaxf(m, plane_sdf(p, f3{0.0f, 0.0f, 1.0f}, table_z));
    m = fmaxf(m, plane_sdf(p, f3{0.0f, 0.0f, -1.0f}, 0.865f));
    m = fmaxf(m, plane_sdf(p, f3{cs, ss, 0.0f}, 0.911f));
    m = fmaxf(m, plane_sdf(p, f3{csf1, ssf1, -1.02f}, 0.9193f));
    m = fmaxf(m, plane_sdf(p, f3{csf2, ssf2, -1.02f}, 0.9193f));
    m = fmaxf(m, plane_sdf(p, f3{csf2, ssf2, 1.03f}, 0.912f));
    m = fmaxf(m, plane_sdf(p, f3{csf1, ssf1, 1.03f}, 0.912f));
    m = fmaxf(m, plane_sdf(p, f3{csf8, ssf8, 2.21f}, 1.131f));
    return m * scale;
}

AP_HD float sdf_functor(int32_t kind, f3 u, const float info[4]) {
    return kind == SDF_FUNCTOR_TORUS ? sdf_torus(u, info[0], info[1]) : sdf_diamond(u, info);
}

AP_HD float vol_at(const SdfVolume& v, int x, int y, int z) {
    return v.values[static_cast<size_t>(z) * v.dim[0] * v.dim[1] + static_cast<size_t>(y) * v.dim[0] + x];
}
AP_HD int clampi(int x, int lo, int hi) { return x < lo ? lo : (x > hi ? hi : x); }

// SampleSDFNearest01 / SampleSDFTrilinear01 / SampleSDFAdaptive01
// (sdf_volume.cuh:64-153) at pos01 in [0, 1]^3.
AP_HD float sdf_volume_sample(const SdfVolume& v, f3 pos01) {
    if (!v.values || v.dim[0] * v.dim[1] * v.dim[2] == 0) return 1.0e20f;
    const f3 g{pos01.x * static_cast<float>(v.dim[0] - 1), pos01.y * static_cast<float>(v.dim[1] - 1),
               pos01.z * static_cast<float>(v.dim[2] - 1)};
    const int xn = clampi(static_cast<int>(floorf(g.x + 0.5f)), 0, v.dim[0] - 1);
    const int yn = clampi(static_cast<int>(floorf(g.y + 0.5f)), 0, v.dim[1] - 1);
    const int zn = clampi(static_cast<int>(floorf(g.z + 0.5f)), 0, v.dim[2] - 1);
    const float nearest = vol_at(v, xn, yn, zn);
    if (fabsf(nearest) > v.voxel_size_max) return nearest;
    int x0 = static_cast<int>(floorf(g.x)), y0 = static_cast<int>(floorf(g.y)), z0 = static_cast<int>(floorf(g.z));
    const int x1 = clampi(x0 + 1, 0, v.dim[0] - 1), y1 = clampi(y0 + 1, 0, v.dim[1] - 1);
    const int z1 = clampi(z0 + 1, 0, v.dim[2] - 1);
    x0 = clampi(x0, 0, v.dim[0] - 1);
    y0 = clampi(y0, 0, v.dim[1] - 1);
    z0 = clampi(z0, 0, v.dim[2] - 1);
    const float fx = fminf(fmaxf(g.x - static_cast<float>(x0), 0.0f), 1.0f);
    const float fy = fminf(fmaxf(g.y - static_cast<float>(y0), 0.0f), 1.0f);
    const float fz = fminf(fmaxf(g.z - static_cast<float>(z0), 0.0f), 1.0f);
    const float c00 = vol_at(v, x0, y0, z0) * (1.0f - fx) + vol_at(v, x1, y0, z0) * fx;
    const float c10 = vol_at(v, x0, y1, z0) * (1.0f - fx) + vol_at(v, x1, y1, z0) * fx;
    const float c01 = vol_at(v, x0, y0, z1) * (1.0f - fx) + vol_at(v, x1, y0, z1) * fx;
    const float c11 = vol_at(v, x0, y1, z1) * (1.0f - fx) + vol_at(v, x1, y1, z1) * fx;
    const float c0 = c00 * (1.0f - fy) + c10 * fy;
    const float c1 = c01 * (1.0f - fy) + c11 * fy;
    return c0 * (1.0f - fz) + c1 * fz;
}

// The distance field of an SDF object at unit-space u: the torus, or the
// volume's sample at u + 0.5 (SDFVolumeFunctor, sdf_volume.cuh:155-171).
struct SdfField {
    int32_t type;                 // SDF_TORUS or SDF_VOLUME
    const SdfVolume* volume;      // SDF_VOLUME
    float info[4];                // SDFInfo::sdfInfo0_
    AP_HD float operator()(f3 u) const {
        if (type == SDF_TORUS) return sdf_torus(u, info[0], info[1]);
        return sdf_volume_sample(*volume, f3{u.x + 0.5f, u.y + 0.5f, u.z + 0.5f});
    }
};

// The field of an SDF object: the torus; a volume only if its index names a
// filled entry of the table (GetSDFVolumeInfoPtr, sdf_master.cuh:14-30);
// otherwise (and for undefined types) the object is never hit.
AP_HD bool sdf_field(int32_t type, int32_t volume, const float info[4], const SdfVolume* vols, int32_t n_vols,
                     SdfField& F) {
    F.type = type;
    F.volume = nullptr;
    for (int k = 0; k < 4; ++k) F.info[k] = info[k];
    if (type == SDF_TORUS) return true;
    if (type == SDF_VOLUME && vols && volume >= 0 && volume < n_vols && vols[volume].values) {
        F.volume = &vols[volume];
        return true;
    }
    return false;
}

// SDF_EPS(depth) = INTERSECT_EPS(depth) / 2 (config.cuh:74) at depth 0.
constexpr float kSdfEps0 = 0.5f * 1.0e-3f;

// ToUnitSpace (sdf_base.cuh:12-21).
AP_HD f3 to_unit(f3 p, f3 ext) {
    return {p.x / (2.0f * fmaxf(ext.x, 1e-8f)), p.y / (2.0f * fmaxf(ext.y, 1e-8f)), p.z / (2.0f * fmaxf(ext.z, 1e-8f))};
}

// RaymarchLocal (sdf_base.cuh:54-110): from max(0, t_enter) along the local
// ray, stepping |d| (x 0.9 for volumes, at least eps / 2), until the
// distance changes sign by less than eps; the hit is the step before.
AP_HD bool sdf_march(const SdfField& F, f3 lo, f3 ld, float t_enter, float t_exit, f3 ext, float eps, float& t_hit,
                     f3& lpos) {
    const float min_step = eps * 0.5f;
    const bool start_inside = F(to_unit(lo, ext)) <= 0.0f;
    const float s_min = 2.0f * fminf(ext.x, fminf(ext.y, ext.z));
    const bool volume = F.type == SDF_VOLUME;
    float t = fmaxf(0.0f, t_enter);
    for (int i = 0; i < kSdfSteps; ++i) {
        const float t_last = t;
        const f3 pl = lo + ld * t;
        const float d_local = F(to_unit(pl, ext)) * s_min;
        const float real = fabsf(d_local);
        if (start_inside ? (d_local > 0.0f && d_local < eps) : (d_local < 0.0f && -d_local < eps)) {
            t_hit = t_last;
            lpos = pl;
            return true;
        }
        t += fmaxf(volume ? real * kSdfVolumeStepScale : real, min_step);
        if (t > t_exit) break;
    }
    return false;
}

// sdfNormalUnit + UnitNormalToLocal (sdf_base.cuh:23-52): the local normal
// at local position lp, central differences of step eps / s_min in unit space.
AP_HD f3 sdf_local_normal(const SdfField& F, f3 lp, f3 ext, float eps) {
    const float s_min = 2.0f * fminf(ext.x, fminf(ext.y, ext.z));
    const float h = eps / fmaxf(s_min, 1e-8f);
    const f3 u = to_unit(lp, ext);
    const float dx = F(f3{u.x + h, u.y, u.z}) - F(f3{u.x - h, u.y, u.z});
    const float dy = F(f3{u.x, u.y + h, u.z}) - F(f3{u.x, u.y - h, u.z});
    const float dz = F(f3{u.x, u.y, u.z + h}) - F(f3{u.x, u.y, u.z - h});
    const f3 nu = normalize(f3{dx, dy, dz});
    return normalize(f3{nu.x / (2.0f * fmaxf(ext.x, 1e-8f)), nu.y / (2.0f * fmaxf(ext.y, 1e-8f)),
                        nu.z / (2.0f * fmaxf(ext.z, 1e-8f))});
}

}  // namespace ap

// Distance-only intersection of one analytic object (cube, sphere, SDF),
// shared by the traversal and by the solo trace of a path inside a medium.
//   IntersectAABBObject    core/object/aabb.cuh:9-89
//   IntersectSphereObject  core/object/sphere.cuh:10-70
//   IntersectSDFObject     core/object/sdf/sdf_base.cuh:115-163
#pragma once

#include "ap_internal.h"
#include "ap_sdf.h"

namespace ap {
namespace {

__device__ __forceinline__ f3 ld3(const float* p) { return {p[0], p[1], p[2]}; }

__device__ __forceinline__ m3 ld33(const float* p) {
    return {{p[0], p[1], p[2]}, {p[3], p[4], p[5]}, {p[6], p[7], p[8]}};
}

// IntersectAABBObject (core/object/aabb.cuh:9-89), distance only.
__device__ bool cube_t(f3 o, f3 d, const AnalyticRec& a, float& t) {
    const f3 c = ld3(a.center), ext = ld3(a.extent);
    const m3 w2o = ld33(a.w2o);
    const f3 lo = mul(w2o, o - c);
    const f3 ld = normalize(mul(w2o, d));
    const f3 div = inv_safe(ld);
    const f3 t1 = (-ext - lo) * div;
    const f3 t2 = (ext - lo) * div;
    const f3 mn = fmin3(t1, t2), mx = fmax3(t1, t2);
    const float te = fmaxf(fmaxf(mn.x, mn.y), mn.z);
    const float tx = fminf(fminf(mx.x, mx.y), mx.z);
    if (te > tx) return false;
    if (tx <= 0.0f) return false;
    t = (te < 0.0f) ? tx : te;
    return true;
}

// IntersectSphereObject (core/object/sphere.cuh:10-70), distance only.
__device__ bool sphere_t(f3 o, f3 d, const AnalyticRec& a, float& t) {
    const f3 c = ld3(a.center), ext = ld3(a.extent);
    const m3 w2o = ld33(a.w2o);
    const f3 lo = mul(w2o, o - c);
    const f3 ld = normalize(mul(w2o, d));
    const float r = max3(ext);
    const float b = dot(lo, ld);
    const float cc = dot(lo, lo) - r * r;
    const float disc = b * b - 1.0f * cc;
    if (disc < 0.0f) return false;
    const float s = sqrtf(disc);
    float tt = -b - s;
    if (tt <= 0.0f) tt = -b + s;
    if (tt <= 0.0f) return false;
    if (isinf(tt)) return false;
    t = tt;
    return true;
}

// IntersectSDFObject (core/object/sdf/sdf_base.cuh:115-163), distance only:
// the local ray through the object's box, then RaymarchLocal with marching
// epsilon eps.
__device__ bool sdf_t(f3 o, f3 d, const AnalyticRec& a, const SdfVolume* vols, int32_t n_vols, float eps, float& t) {
    SdfField F;
    if (!sdf_field(a.sdf_type, a.volume, a.sdf_info, vols, n_vols, F)) return false;
    const f3 c = ld3(a.center), ext = ld3(a.extent);
    const m3 w2o = ld33(a.w2o);
    const f3 lo = mul(w2o, o - c);
    const f3 ld = normalize(mul(w2o, d));
    const f3 div = inv_safe(ld);
    const f3 t1 = (-ext - lo) * div;
    const f3 t2 = (ext - lo) * div;
    const f3 mn = fmin3(t1, t2), mx = fmax3(t1, t2);
    const float te = fmaxf(fmaxf(mn.x, mn.y), mn.z);
    const float tx = fminf(fminf(mx.x, mx.y), mx.z);
    if (te > tx) return false;
    if (tx <= 0.0f) return false;
    f3 lp;
    return sdf_march(F, lo, ld, te, tx, ext, eps, t, lp);
}

}  // namespace
}  // namespace ap

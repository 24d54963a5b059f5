// Small-vector arithmetic shared by the host builder and the HIP kernels.
//
// Every operation keeps the evaluation order of the reference's
// core/vector.cuh so that host and device produce the same FP32 bits:
//   dot      a.x*b.x + a.y*b.y + a.z*b.z, left to right   (vector.cuh:162)
//   cross    (a.y*b.z - a.z*b.y, ...)                      (vector.cuh:244)
//   invSafe  0 -> +FLT_MAX (also for -0.0), else 1/x       (vector.cuh:151-157)
//   normalize  l2 <= 1e-20 -> 0, else v * (1/sqrtf(l2))    (vector.cuh:229-238)
//   float3x3 * v = (dot(row0,v), dot(row1,v), dot(row2,v)) (vector.cuh:346-349)
// The reference's normalize uses the hardware rsqrtf; this core defines it
// as 1/sqrtf (correctly rounded divide and sqrt) so that the result is the
// same on every target. All code is compiled with -ffp-contract=off.
#pragma once

#include <cfloat>
#include <cmath>
#include <cstdint>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define AP_HD __host__ __device__ __forceinline__
#else
#define AP_HD inline
#endif

namespace ap {

constexpr float kPiF = 3.14159265358979323846f;  // CUDART_PI_F rounds to the same float

struct f3 { float x, y, z; };
struct f2 { float x, y; };

AP_HD f3 mk(float x, float y, float z) { return f3{x, y, z}; }
AP_HD f3 operator+(f3 a, f3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
AP_HD f3 operator-(f3 a, f3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
AP_HD f3 operator-(f3 a) { return {-a.x, -a.y, -a.z}; }
AP_HD f3 operator*(f3 a, float b) { return {a.x * b, a.y * b, a.z * b}; }
AP_HD f3 operator*(float b, f3 a) { return {a.x * b, a.y * b, a.z * b}; }
AP_HD f3 operator*(f3 a, f3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
AP_HD f3 operator/(f3 a, float b) { return {a.x / b, a.y / b, a.z / b}; }
AP_HD f3 operator+(f3 a, float b) { return {a.x + b, a.y + b, a.z + b}; }
AP_HD f3 operator-(f3 a, float b) { return {a.x - b, a.y - b, a.z - b}; }

AP_HD float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
AP_HD f3 cross(f3 a, f3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
AP_HD float length(f3 a) { return sqrtf(dot(a, a)); }
AP_HD float inv_safe(float x) { return (x == 0.0f) ? (x >= 0 ? FLT_MAX : -FLT_MAX) : 1.0f / x; }
AP_HD f3 inv_safe(f3 a) { return {inv_safe(a.x), inv_safe(a.y), inv_safe(a.z)}; }
AP_HD f3 normalize(f3 v) {
    const float l2 = dot(v, v);
    if (l2 <= 1e-20f) return {0.0f, 0.0f, 0.0f};
    const float inv = 1.0f / sqrtf(l2);
    return v * inv;
}
AP_HD f3 fmin3(f3 a, f3 b) { return {fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z)}; }
AP_HD f3 fmax3(f3 a, f3 b) { return {fmaxf(a.x, b.x), fmaxf(a.y, b.y), fmaxf(a.z, b.z)}; }
AP_HD float max3(f3 a) { return fmaxf(a.x, fmaxf(a.y, a.z)); }
AP_HD float sum3(f3 a) { return a.x + a.y + a.z; }
AP_HD f3 absf3(f3 a) { return {fabsf(a.x), fabsf(a.y), fabsf(a.z)}; }
AP_HD f3 reflect(f3 v, f3 n) { return v - 2.0f * dot(v, n) * n; }
AP_HD float lerpf(float a, float b, float v) { return a * (1.0f - v) + b * v; }
AP_HD f3 lerp3(f3 a, f3 b, float v) { return {lerpf(a.x, b.x, v), lerpf(a.y, b.y, v), lerpf(a.z, b.z, v)}; }
AP_HD float clampf(float v, float a, float b) { return (v != v) ? 0.0f : fminf(fmaxf(v, a), b); }
AP_HD float saturatef(float v) { return (v != v) ? 0.0f : fminf(fmaxf(v, 0.0f), 1.0f); }
AP_HD float signf(float a) { return a > 0.0f ? 1.0f : (a < 0.0f ? -1.0f : 0.0f); }
AP_HD bool all_zero(f3 a) { return a.x == 0.0f && a.y == 0.0f && a.z == 0.0f; }

// Row-major 3x3 (rows x, y, z), as the reference float3x3.
struct m3 { f3 x, y, z; };
AP_HD f3 mul(const m3& m, f3 v) { return {dot(m.x, v), dot(m.y, v), dot(m.z, v)}; }

AP_HD uint32_t f2u(float f) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __float_as_uint(f);
#else
    uint32_t u; __builtin_memcpy(&u, &f, 4); return u;
#endif
}
AP_HD float u2f(uint32_t u) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __uint_as_float(u);
#else
    float f; __builtin_memcpy(&f, &u, 4); return f;
#endif
}

// Slab test with entry distance (core/bvh/bvh.cuh:444-470). The plain
// FastAABBHit (bvh.cuh:424-442) evaluates the same expressions, so one
// routine serves both: accept iff tExit >= max(0,tMin) && tEnter <= tMax.
// Note that the reference never requires tEnter <= tExit.
//
// With CULL set, a box is also rejected when the ray's line misses the box
// grown by a margin in every axis (w = margin * |invDir| per axis, in ray
// parameter units). That rejection never removes a box holding a primitive
// the reference could accept, so the visit order becomes a subsequence of
// the reference order with the same accepted hits; tEnter (used for the
// near-first order) is the reference's own value.
template <bool CULL>
AP_HD bool slab_enter(f3 o, f3 inv, f3 w, f3 bmin, f3 bmax, float tmin, float tmax, float& tenter) {
    const float tx0 = (bmin.x - o.x) * inv.x;
    const float tx1 = (bmax.x - o.x) * inv.x;
    const float ax = fminf(tx0, tx1), bx = fmaxf(tx0, tx1);
    const float ty0 = (bmin.y - o.y) * inv.y;
    const float ty1 = (bmax.y - o.y) * inv.y;
    const float ay = fminf(ty0, ty1), by = fmaxf(ty0, ty1);
    const float tz0 = (bmin.z - o.z) * inv.z;
    const float tz1 = (bmax.z - o.z) * inv.z;
    const float az = fminf(tz0, tz1), bz = fmaxf(tz0, tz1);
    const float te = fmaxf(fmaxf(ax, ay), az);
    const float tx = fminf(fminf(bx, by), bz);
    tenter = te;
    bool ok = tx >= fmaxf(0.0f, tmin) && te <= tmax;
    if (CULL) {
        const float le = fmaxf(fmaxf(ax - w.x, ay - w.y), az - w.z);
        const float lx = fminf(fminf(bx + w.x, by + w.y), bz + w.z);
        ok = ok && !(le > lx);
    }
    return ok;
}

// Moller-Trumbore on precomputed world-space p0, e1 = p1-p0, e2 = p2-p0
// (core/render/intersection.cuh:15-89). Returns true with t, u, v when the
// reference would report a hit (before the caller's [tMin, tMax) window).
AP_HD bool tri_mt(f3 o, f3 d, f3 p0, f3 e1, f3 e2, float& t, float& u, float& v) {
    const f3 h = cross(d, e2);
    const float a = dot(e1, h);
    if (a > -1e-6f && a < 1e-6f) return false;
    const float f = 1.0f / a;
    const f3 s = o - p0;
    u = f * dot(s, h);
    if (u < 0.0f || u > 1.0f) return false;
    const f3 q = cross(s, e1);
    v = f * dot(d, q);
    if (v < 0.0f || u + v > 1.0f) return false;
    t = f * dot(e2, q);
    return t > 1e-5f;
}

}  // namespace ap

// ORACLE — test infrastructure only. Nothing in the product links this.
//
// FP32 vector arithmetic restated from the reference's core/vector.cuh with
// its evaluation order, compiled by g++ with -ffp-contract=off and no
// fast-math. rsqrtf in normalize (vector.cuh:236) is restated as 1/sqrtf.
#pragma once

#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>

namespace orc {

struct float2 { float x, y; };
struct float3 { float x, y, z; };
struct int4 { int x, y, z, w; };

inline float3 make_float3(float x, float y, float z) { return {x, y, z}; }
inline float3 operator*(float3 a, float b) { return {a.x * b, a.y * b, a.z * b}; }          // :127
inline float3 operator*(float3 a, float3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }   // :128
inline float3 operator*(float b, float3 a) { return a * b; }                                // :130
inline float3 operator/(float3 a, float b) { return {a.x / b, a.y / b, a.z / b}; }          // :132
inline float3 operator+(float3 a, float b) { return {a.x + b, a.y + b, a.z + b}; }          // :135
inline float3 operator+(float3 a, float3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }   // :136
inline float3 operator-(float3 a, float b) { return {a.x - b, a.y - b, a.z - b}; }          // :139
inline float3 operator-(float3 a, float3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }   // :140
inline float3 operator-(float3 a) { return {-a.x, -a.y, -a.z}; }                            // :141
inline float invSafe(float x) { return (x == 0.0f) ? (x >= 0 ? FLT_MAX : -FLT_MAX) : 1.0f / x; }  // :151
inline float3 invSafe(float3 a) { return {invSafe(a.x), invSafe(a.y), invSafe(a.z)}; }
inline float dot(float3 a, float3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }          // :162
inline float3 cross(float3 a, float3 b) {                                                   // :244
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline float length(float3 a) { return std::sqrt(dot(a, a)); }                              // :259
inline float3 normalize(float3 v) {                                                         // :229
    float l2 = dot(v, v);
    if (l2 <= 1e-20f) return make_float3(0, 0, 0);
    float inv = 1.0f / std::sqrt(l2);
    return v * inv;
}
inline float3 fmin3(float3 a, float3 b) { return {std::fmin(a.x, b.x), std::fmin(a.y, b.y), std::fmin(a.z, b.z)}; }
inline float3 fmax3(float3 a, float3 b) { return {std::fmax(a.x, b.x), std::fmax(a.y, b.y), std::fmax(a.z, b.z)}; }
inline float max3(float3 a) { return std::fmax(a.x, std::fmax(a.y, a.z)); }                 // :247
inline float sum3(float3 a) { return a.x + a.y + a.z; }
inline float3 abs3(float3 a) { return {std::fabs(a.x), std::fabs(a.y), std::fabs(a.z)}; }
inline float lerp(float a, float b, float v) { return a * (1.0f - v) + b * v; }             // :241
inline float3 lerp(float3 a, float3 b, float v) { return {lerp(a.x, b.x, v), lerp(a.y, b.y, v), lerp(a.z, b.z, v)}; }
inline float clampf(float v, float a, float b) { return std::isnan(v) ? 0.0f : std::fmin(std::fmax(v, a), b); }  // :96
inline float saturate_(float v) { return std::isnan(v) ? 0.0f : std::fmin(std::fmax(v, 0.0f), 1.0f); }         // :87
inline float sign(float a) { return a > 0.0f ? 1.0f : (a < 0.0f ? -1.0f : 0.0f); }          // :40
inline float3 reflect(float3 v, float3 n) { return v - 2.0f * dot(v, n) * n; }              // :311
inline bool allZero(float3 a) { return a.x == 0.0f && a.y == 0.0f && a.z == 0.0f; }         // :255

struct float3x3 {                                                                           // :338
    float3 x, y, z;
    float3 operator*(float3 v) const { return {dot(x, v), dot(y, v), dot(z, v)}; }          // :346
};

constexpr float PI_F = 3.141592654f;  // CUDART_PI_F

}  // namespace orc

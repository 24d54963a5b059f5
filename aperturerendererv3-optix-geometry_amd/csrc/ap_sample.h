// Direction samplers shared by host generators and device kernels.
//   BuildTangentBasis        core/vector.cuh:277-282
//   BuildTangentBasisRandom  core/vector.cuh:284-294
//   ConcentricSampleDisk     core/render/sample.cuh:63-86
//   SampleCosineWeightedHemisphere core/render/sample.cuh:88-108
// sin/cos come from the platform math library (HIP device libm on the GPU,
// glibc on the host); the reference's CUDA fast-math variants are not
// reproduced.
#pragma once

#include "ap_fmath.h"
#include "ap_vec.h"

namespace ap {

struct DirSample {
    f3 dir;
    float pdf;
};

AP_HD void tangent_basis(f3 n, f3& t, f3& b) {
    const f3 a = (fabsf(n.z) < 0.999f) ? f3{0.0f, 0.0f, 1.0f} : f3{1.0f, 0.0f, 0.0f};
    t = normalize(cross(a, n));
    b = cross(n, t);
}

AP_HD void tangent_basis_random(f3 n, f3& t, f3& b, float e) {
    const f3 a = (fabsf(n.z) < 0.999f) ? f3{0.0f, 0.0f, 1.0f} : f3{1.0f, 0.0f, 0.0f};
    const f3 t0 = normalize(cross(a, n));
    const f3 b0 = cross(n, t0);
    const float ang = e * kPiF * 2.0f;
    const float s = apfm::sin(ang), c = apfm::cos(ang);
    t = c * t0 + s * b0;
    b = -s * t0 + c * b0;
}

AP_HD f2 concentric_disk(float ux, float uy) {
    const float sx = 2.0f * ux - 1.0f;
    const float sy = 2.0f * uy - 1.0f;
    if (sx == 0.0f && sy == 0.0f) return {0.0f, 0.0f};
    float r, phi;
    if (fabsf(sx) > fabsf(sy)) {
        r = sx;
        phi = (kPiF / 4.0f) * (sy / fmaxf(fabsf(sx), 1e-20f));
    } else {
        r = sy;
        phi = (kPiF / 2.0f) - (kPiF / 4.0f) * (sx / fmaxf(fabsf(sy), 1e-20f));
    }
    return {r * apfm::cos(phi), r * apfm::sin(phi)};
}

AP_HD DirSample cosine_hemisphere(f3 e, f3 n) {
    f3 t, b;
    tangent_basis_random(n, t, b, e.z);
    const f2 d = concentric_disk(e.x, e.y);
    const float z = sqrtf(fmaxf(0.0f, 1.0f - d.x * d.x - d.y * d.y));
    const f3 dir = d.x * t + d.y * b + z * n;
    return {normalize(dir), z / kPiF};
}

}  // namespace ap

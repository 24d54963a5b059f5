// Portable FP32 transcendentals used on the render path (host and device).
//
// The path tracer draws the same random numbers as its CPU checker and
// follows the same steps, so the only source of disagreement left between
// them is the last bit of sin/cos/exp/log/erfc, which differs between the
// device math library and the host's libm. These functions are built only
// from IEEE-exact operations (+ - * / sqrt fmaf floor and bit moves), so a
// host build with -ffp-contract=off and the gfx950 build with correctly
// rounded division and square root produce the same bits for the same input.
//
// Methods:
//   sin/cos   quadrant reduction k = rint(x 2/pi), x - k pi/2 with a
//             three-part pi/2 (Cody & Waite 1980) evaluated with fmaf, then
//             the minimax polynomials of the Cephes library's sinf/cosf on
//             [-pi/4, pi/4] (S. L. Moshier, "Methods and Programs for
//             Mathematical Functions", 1989).
//   exp       n = rint(x log2 e), r = x - n ln2 (two-part ln2 with fmaf),
//             Cephes expf polynomial on [-ln2/2, ln2/2], times 2^n.
//   log       x = m 2^e with m in [sqrt(1/2), sqrt(2)), Cephes logf
//             polynomial in m - 1, plus e ln2 in two parts.
//   log1p     log(1 + x) * x / ((1 + x) - 1) (Goldberg 1991, "What every
//             computer scientist should know about floating-point arithmetic").
//   erfc      the Chebyshev-fitted erfc of Numerical Recipes in C, 2nd ed.,
//             sec. 6.2 (fractional error below 1.2e-7).
//   atan2     Cephes atanf reduction and polynomial, quadrant rules of
//             its atan2f.
//   x^0.75    sqrt(x) * sqrt(sqrt(x)).
// Accuracy against float64 (tests/test_fmath_cpu.py): sin and cos within
// 1.6 ulp on [-500, 500], exp and log within 1.1 ulp, log1p and x^0.75
// within 2.5 ulp, erfc within 3e-6 relative on [-4, 6], atan2 within 3 ulp.
#pragma once

#include <cmath>
#include <cstdint>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define APFM_HD __host__ __device__ __forceinline__
#else
#define APFM_HD inline
#endif

namespace apfm {

APFM_HD uint32_t bits(float x) {
    uint32_t u;
    __builtin_memcpy(&u, &x, 4);
    return u;
}

APFM_HD float from_bits(uint32_t u) {
    float x;
    __builtin_memcpy(&x, &u, 4);
    return x;
}

APFM_HD bool finite(float x) { return (bits(x) & 0x7f800000u) != 0x7f800000u; }

APFM_HD float quiet_nan() { return from_bits(0x7fc00000u); }

APFM_HD float inf() { return from_bits(0x7f800000u); }

// 2^n for n in [-126, 127].
APFM_HD float pow2i(int n) { return from_bits(static_cast<uint32_t>(n + 127) << 23); }

// Polynomials on |r| <= pi/4, z = r*r.
APFM_HD float sin_poly(float r, float z) {
    float p = fmaf(-1.9515295891e-4f, z, 8.3321608736e-3f);
    p = fmaf(p, z, -1.6666654611e-1f);
    return fmaf(p * z, r, r);
}

APFM_HD float cos_poly(float z) {
    float p = fmaf(2.443315711809948e-5f, z, -1.388731625493765e-3f);
    p = fmaf(p, z, 4.166664568298827e-2f);
    return fmaf(p * z, z, fmaf(-0.5f, z, 1.0f));
}

// Quadrant and reduced argument: x = k pi/2 + r.
APFM_HD float reduce_pio2(float x, int& quadrant) {
    const float k = floorf(fmaf(x, 0.636619772367581343f, 0.5f));
    float r = fmaf(-k, 1.57079637050628662109375f, x);
    r = fmaf(-k, -4.37113882867379300296306610107e-8f, r);
    r = fmaf(-k, -1.71512451000588190e-15f, r);
    quadrant = static_cast<int>(k) & 3;
    return r;
}

APFM_HD void sincos(float x, float& s, float& c) {
    if (!finite(x)) {
        s = c = quiet_nan();
        return;
    }
    int q;
    const float r = reduce_pio2(x, q);
    const float z = r * r;
    const float sr = sin_poly(r, z), cr = cos_poly(z);
    switch (q) {
        case 0: s = sr; c = cr; break;
        case 1: s = cr; c = -sr; break;
        case 2: s = -sr; c = -cr; break;
        default: s = -cr; c = sr; break;
    }
}

APFM_HD float sin(float x) {
    float s, c;
    sincos(x, s, c);
    return s;
}

APFM_HD float cos(float x) {
    float s, c;
    sincos(x, s, c);
    return c;
}

APFM_HD float exp(float x) {
    if (x != x) return x;
    if (x > 88.72283905206835f) return inf();
    if (x < -103.97208f) return 0.0f;
    const float n = floorf(fmaf(x, 1.44269504088896341f, 0.5f));
    float r = fmaf(-n, 0.693147182464599609375f, x);
    r = fmaf(-n, -1.90465421212593362e-9f, r);
    float p = fmaf(1.9875691500e-4f, r, 1.3981999507e-3f);
    p = fmaf(p, r, 8.3334519073e-3f);
    p = fmaf(p, r, 4.1665795894e-2f);
    p = fmaf(p, r, 1.6666665459e-1f);
    p = fmaf(p, r, 5.0000001201e-1f);
    const float y = fmaf(p, r * r, r) + 1.0f;
    const int ni = static_cast<int>(n);
    if (ni > 127) return (y * pow2i(127)) * 2.0f;
    if (ni < -125) return (y * pow2i(ni + 100)) * pow2i(-100);
    return y * pow2i(ni);
}

APFM_HD float log(float x) {
    if (x != x) return x;
    if (x < 0.0f) return quiet_nan();
    if (x == 0.0f) return -inf();
    if (!finite(x)) return x;
    int e = 0;
    if (x < 1.17549435e-38f) {  // subnormal: scale into the normal range
        x = x * 8388608.0f;
        e = -23;
    }
    const uint32_t u = bits(x);
    e += static_cast<int>((u >> 23) & 0xffu) - 126;
    float m = from_bits((u & 0x807fffffu) | 0x3f000000u);  // [0.5, 1)
    if (m < 0.707106781186547524f) {
        e -= 1;
        m = (m + m) - 1.0f;
    } else {
        m = m - 1.0f;
    }
    const float z = m * m;
    float p = fmaf(7.0376836292e-2f, m, -1.1514610310e-1f);
    p = fmaf(p, m, 1.1676998740e-1f);
    p = fmaf(p, m, -1.2420140846e-1f);
    p = fmaf(p, m, 1.4249322787e-1f);
    p = fmaf(p, m, -1.6668057665e-1f);
    p = fmaf(p, m, 2.0000714765e-1f);
    p = fmaf(p, m, -2.4999993993e-1f);
    p = fmaf(p, m, 3.3333331174e-1f);
    const float fe = static_cast<float>(e);
    float y = (p * m) * z;
    y = fmaf(fe, -2.12194440e-4f, y);
    y = fmaf(-0.5f, z, y);
    float r = m + y;
    r = fmaf(fe, 0.693359375f, r);
    return r;
}

APFM_HD float log1p(float x) {
    const float u = 1.0f + x;
    if (u == 1.0f) return x;
    if (!finite(u)) return log(u);
    return log(u) * (x / (u - 1.0f));
}

APFM_HD float erfc(float x) {
    if (x != x) return x;
    const float z = x < 0.0f ? -x : x;
    const float t = 1.0f / fmaf(0.5f, z, 1.0f);
    float p = fmaf(t, 0.17087277f, -0.82215223f);
    p = fmaf(t, p, 1.48851587f);
    p = fmaf(t, p, -1.13520398f);
    p = fmaf(t, p, 0.27886807f);
    p = fmaf(t, p, -0.18628806f);
    p = fmaf(t, p, 0.09678418f);
    p = fmaf(t, p, 0.37409196f);
    p = fmaf(t, p, 1.00002368f);
    p = fmaf(t, p, -1.26551223f);
    const float ans = t * exp(fmaf(-z, z, p));
    return x >= 0.0f ? ans : 2.0f - ans;
}

// atan on [0, inf) by the Cephes atanf reduction (tan(3pi/8), tan(pi/8)
// thresholds) and its polynomial; odd extension.
APFM_HD float atan(float x) {
    if (x != x) return x;
    const bool neg = x < 0.0f;
    float a = neg ? -x : x;
    float y0 = 0.0f;
    if (a > 2.414213562373095f) {
        y0 = 1.5707963267948966f;
        a = -1.0f / a;
    } else if (a > 0.4142135623730950f) {
        y0 = 0.7853981633974483f;
        a = (a - 1.0f) / (a + 1.0f);
    }
    const float z = a * a;
    float p = fmaf(8.05374449538e-2f, z, -1.38776856032e-1f);
    p = fmaf(p, z, 1.99777106478e-1f);
    p = fmaf(p, z, -3.33329491539e-1f);
    const float r = y0 + fmaf(p * z, a, a);
    return neg ? -r : r;
}

// atan2 with the quadrant rules of the Cephes atan2f: x = 0 gives +-pi/2
// (0 when y = 0 too), y = 0 gives 0 or pi.
APFM_HD float atan2(float y, float x) {
    if (x != x || y != y) return x + y;
    const float pi = 3.14159265358979323846f;
    if (x == 0.0f) {
        if (y < 0.0f) return -0.5f * pi;
        if (y == 0.0f) return 0.0f;
        return 0.5f * pi;
    }
    if (y == 0.0f) return x < 0.0f ? pi : 0.0f;
    float w = 0.0f;
    if (x < 0.0f) w = y < 0.0f ? -pi : pi;
    return w + atan(y / x);
}

// x^y for x >= 0 as exp(y log x) (a few ulp; 0^y = 0 for y > 0, NaN for x < 0).
APFM_HD float pow(float x, float y) {
    if (x != x || y != y) return quiet_nan();
    if (x < 0.0f) return quiet_nan();
    if (x == 1.0f || y == 0.0f) return 1.0f;
    if (x == 0.0f) return y > 0.0f ? 0.0f : inf();
    return exp(y * log(x));
}

// x^0.75 for x >= 0.
APFM_HD float pow075(float x) {
    const float s = sqrtf(x);
    return s * sqrtf(s);
}

}  // namespace apfm

// ORACLE — test infrastructure only: CPU restatement of the FP64 image
// statistics, the checker for the core's device reductions (ap_measure.hip).
//
//   DeviceMSE               core/helper/measure.cuh:18-33
//   DeviceMSELuminance      core/helper/measure.cuh:49-73 (called with the odd
//                           and even framebuffers, core/camera.cu:94-97)
//   relMSE                  mean over pixels and linear-sRGB channels of
//                           (I - R)^2 / (R^2 + eps); XYZ2SRGBLinearD65
//                           (core/tonemap.cuh:116-126) in FP64 on the FP32
//                           matrix entries
// Sums run sequentially in FP64 (thrust's order is unspecified; the device
// sums in a fixed tree). Pinned in tests/test_measure_cpu.py against exact
// (math.fsum) sums and closed forms.
#include <cmath>
#include <cstddef>
#include <limits>

#include "or_api.h"

namespace {

void xyz_to_srgb_f64(const double x[3], double r[3]) {
    const double M[3][3] = {{3.2406f, -1.5372f, -0.4986f}, {-0.9689f, 1.8758f, 0.0415f}, {0.0557f, -0.204f, 1.057f}};
    for (int k = 0; k < 3; ++k) {
        const double v = M[k][0] * x[0] + M[k][1] * x[1] + M[k][2] * x[2];
        r[k] = v > 0.0 ? v : 0.0;
    }
}

void load(const void* p, int kind, double scale, size_t i, double v[3]) {
    if (kind == 0) {
        const float* f = static_cast<const float*>(p) + 3 * i;
        for (int c = 0; c < 3; ++c) v[c] = static_cast<double>(f[c]) * scale;
    } else {
        const double* d = static_cast<const double*>(p) + 3 * i;
        for (int c = 0; c < 3; ++c) v[c] = d[c] * scale;
    }
}

}  // namespace

extern "C" {

double or_mse(const float* a, const float* b, size_t n) {
    if (n == 0) return std::numeric_limits<double>::quiet_NaN();
    double s = 0.0;
    for (size_t i = 0; i < n; ++i) {
        const float d = a[i] - b[i];
        s += static_cast<double>(d) * d;
    }
    return s / static_cast<double>(n);
}

double or_mse_luminance(const float* a, const float* b, size_t n) {
    if (n == 0) return std::numeric_limits<double>::quiet_NaN();
    double s = 0.0;
    for (size_t i = 0; i < n; ++i) {
        const float la = 0.2126f * a[3 * i] + 0.7152f * a[3 * i + 1] + 0.0722f * a[3 * i + 2];
        const float lb = 0.2126f * b[3 * i] + 0.7152f * b[3 * i + 1] + 0.0722f * b[3 * i + 2];
        const double d = static_cast<double>(la) - static_cast<double>(lb);
        s += d * d;
    }
    return s / static_cast<double>(n);
}

double or_relmse(const void* img, int img_kind, double img_scale, const void* ref, int ref_kind, double ref_scale,
                 size_t n, double eps) {
    if (n == 0) return std::numeric_limits<double>::quiet_NaN();
    double s = 0.0;
    for (size_t i = 0; i < n; ++i) {
        double a[3], b[3], ra[3], rb[3];
        load(img, img_kind, img_scale, i, a);
        load(ref, ref_kind, ref_scale, i, b);
        xyz_to_srgb_f64(a, ra);
        xyz_to_srgb_f64(b, rb);
        for (int c = 0; c < 3; ++c) {
            const double d = ra[c] - rb[c];
            s += d * d / (rb[c] * rb[c] + eps);
        }
    }
    return s / (3.0 * static_cast<double>(n));
}

}  // extern "C"

/*
 * FP64 image statistics on the device: the reductions that compare
 * framebuffers (the odd/even variance estimate of every interactive frame,
 * mean squared error, relative MSE against a reference image).
 *
 * Every reduction runs as two passes over a fixed grid of AP_REDUCE_BLOCKS
 * blocks: each block sums its grid-stride share of the elements in FP64 into
 * one partial, then one block adds the partials in a fixed order. The result
 * is therefore the same bits on every call with the same inputs (no atomics)
 * and does not depend on the device. `out` is one device double, written in
 * stream order; n = 0 writes a quiet NaN, as the reference does.
 * Inputs are device pointers; pixel i of a float3 image is floats 3i..3i+2.
 */
#ifndef APERTURE_AP_MEASURE_H
#define APERTURE_AP_MEASURE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Blocks of the reduction grid, and the FP64 partials they leave in scratch
 * memory (AP_REDUCE_BLOCKS doubles, 8 KiB). */
#define AP_REDUCE_BLOCKS 1024
/* Largest element or pixel count a reduction accepts. */
#define AP_REDUCE_MAX_ELEMENTS ((size_t)1 << 34)

/* Image formats for ap_relmse_f64. */
#define AP_IMAGE_XYZ_F32 0      /* float3 XYZ per pixel (a running mean)          */
#define AP_IMAGE_XYZ_SUMS_F64 1 /* double3 XYZ per pixel (per-pixel sample sums)  */

typedef struct ap_image_view {
    const void* data; /* device pointer in the format `kind`            */
    int32_t kind;     /* AP_IMAGE_XYZ_F32 or AP_IMAGE_XYZ_SUMS_F64       */
    int32_t pad;
    double scale;     /* each value is multiplied by it (1/spp for sums) */
} ap_image_view;

/* Mean of (a[i] - b[i])^2 over n floats, the difference taken in FP32 and
 * squared and summed in FP64. Replaces DeviceMSE (core/helper/measure.cuh:18-33). */
int ap_mse_f64(const float* a, const float* b, size_t n, double* out, void* stream);

/* Mean over n pixels of (L(a) - L(b))^2, L = 0.2126 x + 0.7152 y + 0.0722 z
 * in FP32, the difference in FP64. Replaces DeviceMSELuminance
 * (core/helper/measure.cuh:49-73), which Camera::Render calls on the odd and
 * even framebuffers after every frame and scales by 1e6 (core/camera.cu:94-97). */
int ap_mse_luminance_f64(const float* a, const float* b, size_t n_pixels, double* out, void* stream);

/* relMSE of an image against a reference: the mean over pixels and the three
 * linear-sRGB channels of (I - R)^2 / (R^2 + eps), both images converted from
 * XYZ (times their scale) to linear sRGB in FP64 with XYZ2SRGBLinearD65
 * (core/tonemap.cuh:116-126: the D65 matrix, then max(0, .)). No reference
 * counterpart: the measure the equal-time comparisons report. */
int ap_relmse_f64(const ap_image_view* image, const ap_image_view* reference, size_t n_pixels, double eps,
                  double* out, void* stream);

/* ap_relmse_f64 without any allocation: no counterpart in the reference; for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc). It calls no
 * hipMalloc, hipFree or allocating library routine: its AP_REDUCE_BLOCKS FP64
 * partials (8 KiB) live in a static device array of the library, one per GPU,
 * held until the process ends. The same limit as ap_relmse_f64
 * (AP_REDUCE_MAX_ELEMENTS pixels; larger inputs are rejected). Every call on
 * one GPU shares that GPU's array: calls that overlap on one GPU (on
 * different streams, or from several threads) race and must be serialised by
 * the caller. Nothing in the array needs clearing between calls. */
int ap_relmse_f64_noalloc(const ap_image_view* image, const ap_image_view* reference, size_t n_pixels, double eps,
                          double* out, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* APERTURE_AP_MEASURE_H */

/*
 * Spectral upsampling of the Aperture core (RGB -> 4-term spectral basis and
 * wavelength -> basis, both from the reference's baked MLP, plus the CIE 1931
 * wavelength sampler). All arrays are device pointers; calls are stream
 * ordered. The LUTs (512^3 x float4 RGB cells = 2 GiB and 840 x float4
 * wavelength cells) are built once per device by ap_spectrum_setup().
 */
#ifndef APERTURE_AP_SPECTRUM_H
#define APERTURE_AP_SPECTRUM_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Replaces SceneRenderer::SetupSpectrumLUT (core/scene.cu:771-787) with
 * PrecomputeRGBLUTKernel / PrecomputeLambdaLUTKernel
 * (core/spectrum/spectrum_lut.cuh:12-78): builds both LUTs on the current
 * device once; later calls return immediately. */
int ap_spectrum_setup(void);

/* Replaces spectrum::query::RGBEncode (spectrum_basis.cuh:2273-2378): n RGB
 * triples (float[3n]) -> n bases (float[4n]). */
int ap_spectrum_encode_rgb(const float* rgb, size_t n, float* basis, void* stream);

/* Replaces spectrum::query::LambdaEncode (spectrum_basis.cuh:2380-2458). */
int ap_spectrum_encode_lambda(const float* lambda_nm, size_t n, float* basis, void* stream);

/* Replaces spectrum::lut::QueryRGBBasis (spectrum_lut.cuh:124-179). */
int ap_spectrum_query_rgb(const float* rgb, size_t n, float* basis, void* stream);

/* Replaces spectrum::lut::QueryLambda (spectrum_lut.cuh:181-220). */
int ap_spectrum_query_lambda(const float* lambda_nm, size_t n, float* basis, void* stream);

/* Replaces spectrum::lut::Interact(QueryLambda(lambda), rgb) (spectrum_lut.cuh:222-235). */
int ap_spectrum_interact(const float* lambda_nm, const float* rgb, size_t n, float* out, void* stream);

/* Replaces spectrum::sample::SampleCIE1931 (sample_ciexyz.cuh:288-326):
 * u3 = float[3n] uniforms -> wavelength (nm) and pdf. */
int ap_spectrum_sample_lambda(const float* u3, size_t n, float* lambda_nm, float* pdf, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* APERTURE_AP_SPECTRUM_H */

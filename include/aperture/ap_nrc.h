/*
 * Learned radiance cache (libaperture_hip.so). No reference counterpart: the
 * reference integrator has no cache. It follows neural radiance caching
 * (Mueller, Rousselle, Novak, Keller 2021): a 64 -> 64 x 5 -> 3 ReLU network
 * without biases, BF16 weights and activations with FP32 accumulation on
 * MFMA, relative L2 loss, Adam (beta1 0.9, beta2 0.99, eps 1e-8).
 *
 * Query record (16 floats): position xyz, outgoing direction xyz, shading
 * normal xyz, roughness, albedo rgb, 3 unused. Training sample (20 floats):
 * a query record, target RGB radiance, 1 unused. Encoding: the position
 * normalised to the cache bounds, sin/cos at frequencies 2^k pi (k = 0..5)
 * per axis, then direction, normal, roughness, albedo and a constant 1,
 * zero-padded to 64 features. Output: linear RGB radiance.
 *
 * Parameters (21504 floats): five hidden matrices W[o][i] of 64 x 64, then
 * the 16 x 64 output matrix whose rows 3..15 stay zero.
 */
#ifndef APERTURE_AP_NRC_H
#define APERTURE_AP_NRC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AP_NRC_PARAMS 21504

typedef struct ap_nrc ap_nrc;

/* He-normal initial weights from PCG32(seed); Adam state zero. */
int ap_nrc_create(uint64_t seed, ap_nrc** out);
int ap_nrc_destroy(ap_nrc* nrc);
/* World-space box the position encoding normalises to (default [-1,1]^3). */
int ap_nrc_set_bounds(ap_nrc* nrc, const float bmin[3], const float bmax[3]);
/* FP32 master parameters to / from host memory (set also resets Adam). */
int ap_nrc_get_params(const ap_nrc* nrc, float* host_params);
int ap_nrc_set_params(ap_nrc* nrc, const float* host_params);
/* RGB radiance of n device query records into device out_rgb[n][3]. */
int ap_nrc_infer(ap_nrc* nrc, const float* queries, size_t n, float* out_rgb, void* stream);
/* One training step on n device samples (n a positive multiple of 256):
 * forward, loss, backward, Adam with learning rate lr. Asynchronous. */
int ap_nrc_train_step(ap_nrc* nrc, const float* samples, size_t n, float lr, void* stream);
/* Gradient of the mean relative L2 loss over n samples without an update,
 * copied to host_grad (AP_NRC_PARAMS floats) with the loss; synchronous. */
int ap_nrc_gradient(ap_nrc* nrc, const float* samples, size_t n, float* host_grad, float* loss, void* stream);
/* Gradient of the last training step or gradient call. */
int ap_nrc_last_gradient(const ap_nrc* nrc, float* host_grad);

#ifdef __cplusplus
}
#endif
#endif /* APERTURE_AP_NRC_H */

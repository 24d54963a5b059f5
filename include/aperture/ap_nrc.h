/*
 * Learned radiance cache (libaperture_hip.so). No reference counterpart: the
 * reference integrator has no cache. It follows neural radiance caching
 * (Mueller, Rousselle, Novak, Keller 2021): a 64 -> 64 x 5 -> 3 ReLU network
 * without biases, BF16 weights and activations with FP32 accumulation on
 * MFMA, relative L2 loss, Adam (beta1 0.9, beta2 0.99, eps 1e-8; the bias
 * corrections 1 - beta^t take beta^t as the FP32 product of t factors,
 * kept on the device, so captured training steps replay correctly).
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
 *
 * Grouped caches (ap_nrc_create_grouped) hold G independent networks of
 * hidden width W = 64 or 32 (64 -> W x 5 -> 3; per network W x 64, four
 * W x W, then 16 x W: 64W + 4W^2 + 16W floats, 6656 for W = 32), BF16 or FP16
 * weights and activations, each with its own bounds; every launch serves
 * all of them (a grouped GEMM, one network per 64-row tile). FP16 networks
 * round to nearest even and saturate at +-65504, and scale the loss
 * gradient by S = 2^(ceil(log2(3 batch)) - 4) before its FP16 rounding,
 * dividing the FP32 weight gradient by S again. Parameters, gradients and
 * Adam state are G consecutive blocks of one network's layout. The plain
 * entry points (ap_nrc_infer, _gradient, _train_step) take caches of one
 * network of any width and precision.
 */
#ifndef APERTURE_AP_NRC_H
#define APERTURE_AP_NRC_H

#include <stddef.h>
#include <stdint.h>

#include "ap_render.h"

#ifdef __cplusplus
extern "C" {
#endif

#define AP_NRC_PARAMS 21504
#define AP_NRC_BF16 0
#define AP_NRC_FP16 1

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
/* Inference weights: with decay > 0, the exponential moving average of the
 * parameters over the Adam steps since the last ap_nrc_set_params/create,
 * ema_t = decay ema_{t-1} + (1 - decay) w_t from ema_0 = 0, divided by
 * 1 - decay^t (decay^t an FP32 product of t factors); with 0 (the default)
 * the current parameters. Training always uses the current parameters.
 * Waits for the device (no training step may be in flight). */
int ap_nrc_set_ema(ap_nrc* nrc, float decay);
/* Bounds the position encoding uses (set by ap_nrc_set_bounds or ap_render_nrc). */
int ap_nrc_get_bounds(const ap_nrc* nrc, float bmin[3], float bmax[3]);

/* Grouped caches: `groups` (1..4096) networks of hidden width 32 or 64 in
 * AP_NRC_BF16 or AP_NRC_FP16; network g starts from PCG32(seed, 7 + g)
 * (ap_nrc_create is groups 1, width 64, BF16). */
int ap_nrc_create_grouped(uint32_t groups, uint32_t width, int32_t precision, uint64_t seed, ap_nrc** out);
int ap_nrc_info(const ap_nrc* nrc, uint32_t* groups, uint32_t* width, int32_t* precision,
                size_t* params_per_network);
/* Bounds of network `group` (ap_nrc_set_bounds sets every network's). */
int ap_nrc_set_group_bounds(ap_nrc* nrc, uint32_t group, const float bmin[3], const float bmax[3]);
/* Inference of every network at once: query rows [host_seg[g],
 * host_seg[g + 1]) go to network g (G + 1 non-decreasing offsets), output
 * row r is query row r. Waits for `stream` before it rewrites its row map,
 * then launches asynchronously. */
int ap_nrc_infer_grouped(ap_nrc* nrc, const float* queries, const uint32_t* host_seg, float* out_rgb, void* stream);
/* Gradient (synchronous; host_grad G x params, host_loss G floats, each
 * network's mean loss) and training step (asynchronous) of every network on
 * G x batch device samples laid out [network][batch], batch a positive
 * multiple of 256, G x batch at most 2^24. */
int ap_nrc_gradient_grouped(ap_nrc* nrc, const float* samples, size_t batch, float* host_grad, float* host_loss,
                            void* stream);
int ap_nrc_train_step_grouped(ap_nrc* nrc, const float* samples, size_t batch, float lr, void* stream);

/* Rendering with the cache.
 *
 * A camera path ends into the cache at the first surface vertex x_n where
 * the path spread a(x_1..x_n) = (sum_{i=2..n} sqrt(|x_{i-1} - x_i|^2 /
 * (p(w_{i-1}) |cos theta_i|)))^2 exceeds spread * a_0, a_0 = |x_0 - x_1|^2 /
 * (4 pi |cos theta_1|): its radiance there is the cache's RGB prediction
 * (clamped at 0) turned into the path's wavelength by the RGB-to-spectrum
 * basis, times the path throughput. No NEE is done at that vertex; the
 * vertex's own emission is added as usual. Only vertices that scatter
 * (not lights, not GGX seen from behind) and paths that did not start
 * inside an object end into the cache.
 *
 * Training paths: the pixels p = offset + k * stride of a frame (stride =
 * max(1, pixels / train_paths), offset from SplitMix64 of the frame). They
 * go on past the first crossing (the spread restarts there) and end into
 * the cache at the second; every unbiased_every-th never does. Each of their
 * scattering vertices (up to AP_NRC_TRAIN_VERTICES, the cache vertex
 * excluded) gives a training sample: its query record and the target
 * XYZ2SRGBLinearD65(xyz(lambda) * (L_end - L_k) / M_k / p(lambda)), L_k
 * the path radiance after the vertex's emission, M_k the throughput on
 * arrival. After each frame, train_steps Adam steps run on batches of
 * `batch` samples drawn without repetition (while they last) from the frame's
 * samples. */
#define AP_NRC_TRAIN_VERTICES 8

typedef struct ap_nrc_render_params {
    float spread;            /* c of the spread heuristic (0.01) */
    uint32_t train_paths;    /* training paths per frame (32768); 0: the cache stays frozen */
    uint32_t train_steps;    /* Adam steps per frame (4) */
    uint32_t batch;          /* samples per step, a positive multiple of 256 (16384) */
    float learning_rate;     /* Adam step size (0.005) */
    uint32_t unbiased_every; /* every k-th training path never ends into the cache (16); 0: none */
    uint32_t train_passes;   /* training-only passes after each frame (0): the training paths of
                              * another pixel offset and random stream, traced and trained on,
                              * nothing added to the framebuffers */
} ap_nrc_render_params;

typedef struct ap_nrc_frame_stats {
    uint32_t queries;        /* paths that ended into the cache */
    uint32_t train_paths;    /* training paths */
    uint32_t samples;        /* training samples */
    uint32_t pad;
} ap_nrc_frame_stats;

void ap_nrc_render_params_default(ap_nrc_render_params* p);

/* ap_render (include/aperture/ap_render.h) with the cache: frames one at a
 * time, each rendered with the cache as the frames before it trained it,
 * then trained on its own samples (with use_guiding, each also sampling the
 * probes the frames before it built; a path that ends into the cache takes
 * the prediction as its endpoint radiance for the probes). The scene's root
 * box, grown by 1e-3 of its extent on each side, becomes the cache bounds.
 * The frames of one call overlap on the device (the next frame's camera
 * rays beside this frame's last bounces, the training beside the next
 * frame's bounces) with the same results as one call per frame.
 * Asynchronous on `stream`. */
int ap_render_nrc(ap_scene* scene, const ap_camera* camera, const ap_render_params* params, ap_nrc* cache,
                  const ap_nrc_render_params* nrc_params, const ap_framebuffers* fb, void* stream);

/* The last batch of paths of ap_render_nrc (its last frame, or the last
 * training-only pass after it): counters and, if host_samples
 * is not NULL, min(samples, capacity) training samples (20 floats; the last
 * float holds pixel * 8 + vertex as uint32 bits) in pixel, vertex order.
 * Synchronous. */
int ap_nrc_last_frame(ap_scene* scene, ap_nrc_frame_stats* stats, float* host_samples, size_t capacity);

#ifdef __cplusplus
}
#endif
#endif /* APERTURE_AP_NRC_H */

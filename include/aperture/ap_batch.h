/*
 * Batched multi-scene rendering (libaperture_hip.so). No reference
 * counterpart: the reference renders one SceneRenderer per call
 * (core/scene.cu:308-346). A batch renders the frames of many scenes of
 * uneven size together: every scene's paths share one set of queues, each
 * bounce is one traversal launch over the rays of all scenes, and one
 * frame of the whole batch is a HIP graph captured once and replayed.
 * The traversal launch shape is the batch's own, set by
 * ap_batch_set_launch or picked by ap_batch_autotune.
 *
 * Each scene's image equals, bit for bit, what ap_render gives for that
 * scene alone with the same parameters (same per-pixel random streams,
 * same traversal and shading), with or without the graph.
 */
#ifndef APERTURE_AP_BATCH_H
#define APERTURE_AP_BATCH_H

#include <stddef.h>
#include <stdint.h>

#include "ap_nrc.h"
#include "ap_render.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ap_batch ap_batch;

/* A batch of n scenes (at most 65535), scene k seen by cameras[k]. The
 * scenes stay owned by the caller and must outlive the batch. */
int ap_batch_create(ap_scene* const* scenes, const ap_camera* cameras, size_t n, ap_batch** out);
int ap_batch_destroy(ap_batch* batch);

/* Adds params->spp frames (frame_index, frame_index + 1, ...) of every
 * scene: hdr[k] (W_k*H_k*3 linear XYZ running mean, required unless NULL to
 * skip the scene's image) and, if sums is not NULL and sums[k] is not NULL,
 * FP64 per-pixel sums. The shading, traversal and spectral parameters are
 * ap_render's; odd/even and primary-id buffers are not produced. With
 * use_guiding, each scene is guided by its own probes (the scene's guiding
 * state, shared with ap_render of that scene) and every frame rebuilds every
 * scene's probes from that scene's samples, so scene k's image and probes
 * equal ap_render's of scene k alone over the same frames; guided frames are
 * launched one by one and the call waits once per scene and frame (the
 * probe build). ray_count (host, may be NULL) gets the closest-hit and
 * shadow rays traced. Otherwise asynchronous on `stream` except for the
 * first call with new parameters, which captures the graph. */
int ap_batch_render(ap_batch* batch, const ap_render_params* params, float* const* hdr, double* const* sums,
                    uint64_t ray_count[2], void* stream);

/* ap_batch_render with per-scene learned radiance caches (ap_nrc.h): `cache`
 * holds one network per scene (ap_nrc_create_grouped with groups = n, any
 * width and precision), network k serving scene k with scene k's root box,
 * grown as ap_render_nrc grows it, as its bounds. Frame after frame, every
 * scene's paths end into its own network as ap_render_nrc describes, all
 * networks answer in one grouped inference, and each trains on its own
 * scene's samples (training paths chosen per scene as ap_render_nrc does,
 * batch rows per network, batch x n at most 2^24) in the same launches. With
 * a width-64 BF16 cache, scene k's image and network k's weights equal, bit
 * for bit, ap_render_nrc of scene k alone with a one-network cache that
 * starts from network k. train_passes must be 0. With use_guiding, each
 * scene is also guided by its own probes, a path that ends into its scene's
 * cache gives the probes the prediction as its endpoint radiance, and every
 * scene's probes are rebuilt from its own samples after every frame; scene k
 * then equals ap_render_nrc of scene k alone, guided, probes included.
 * Unguided, the call's first frame is launched directly and the others
 * replay one captured graph (the frame index, the frame's training draws and
 * Adam's bias corrections are read on the device); guided frames are all
 * launched directly. */
int ap_batch_render_nrc(ap_batch* batch, const ap_render_params* params, ap_nrc* cache,
                        const ap_nrc_render_params* nrc_params, float* const* hdr, double* const* sums,
                        uint64_t ray_count[2], void* stream);
/* Per-scene counters of the last frame of ap_batch_render_nrc (n = scenes):
 * queries, training paths, training samples. Synchronous. */
int ap_batch_nrc_last_frame(ap_batch* batch, ap_nrc_frame_stats* per_scene, size_t n);

/* Traversal launch shape of the batch (ap_scene_set_launch's meaning). */
int ap_batch_set_launch(ap_batch* batch, int wave_rays, int refill);
int ap_batch_get_launch(const ap_batch* batch, int* wave_rays, int* refill);
/* 1 (default): frames run as a captured HIP graph; 0: the same kernels are
 * launched one by one. */
int ap_batch_set_graph(ap_batch* batch, int enable);

/* Renders one untimed and `frames` timed batch frames (into no image) for
 * each launch shape of {128, 256, 512} rays per wave x {8, 16} refill, and
 * keeps the fastest; reports it and its time per frame. Frames are timed
 * unguided (use_guiding is ignored, so no scene's probes change). */
int ap_batch_autotune(ap_batch* batch, const ap_render_params* params, int frames, int* best_wave_rays,
                      int* best_refill, float* best_ms_per_frame, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* APERTURE_AP_BATCH_H */

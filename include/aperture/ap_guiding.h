/*
 * Directional path guiding state of a scene (libaperture_hip.so).
 *
 * ap_render with ap_render_params.use_guiding = 1 renders one frame at a
 * time; after each frame the frame's guiding samples are folded into the
 * scene's probes (SceneRenderer::UpdatePathGuidingSample, core/scene.cu:583-744),
 * and the next frame guides its BSDF sampling and NEE MIS with them
 * (core/render/render.cuh:546-639, 300-338). Probes persist across ap_render
 * calls until ap_guiding_reset or ap_scene_destroy.
 */
#ifndef APERTURE_AP_GUIDING_H
#define APERTURE_AP_GUIDING_H

#include "ap_core.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Replaces SceneRenderer::MarkNeedResetGuiding + ResetPathGuidingSample
 * (core/scene.hpp:118-121, core/scene.cu:542-580): drops every probe and the
 * index volume. */
int ap_guiding_reset(ap_scene* scene);

/* Copies out the probes (Octahedron records, 1392 B each) and the int32
 * index volume ([z][y][x][face], -1 = no probe) after synchronising the
 * device. bounds receives the inclusive voxel start xyz, end xyz and size xyz
 * (9 ints); probe_count the number of probes. Any output may be NULL; the
 * arrays must hold probe_capacity / volume_capacity entries (AP_ERR_INVALID_ARG
 * if too small). No reference counterpart: checker access to the probes the
 * reference keeps in SCENE_PROBES / INDIRECT_INDEX_VOLUME (render.cuh:24-56). */
int ap_guiding_export(const ap_scene* scene, ap_octahedron* probes, size_t probe_capacity, int32_t* volume,
                      size_t volume_capacity, int32_t* bounds, int32_t* probe_count);

/* The guiding build's device primitives, for their tests (no reference
 * counterpart: the reference calls thrust). Device pointers; the workspace
 * is allocated inside and the device synchronised before returning.
 * ap_guiding_scan_u32: out[i] = in[0] + ... + in[i-1] (out may equal in),
 * *total (device, may be NULL) = the sum of all n; n < 2^32.
 * ap_guiding_sort_pairs_u64: stable LSD radix sort of (key, value) pairs by
 * key bits [begin_bit, end_bit), 8 bits per pass; inputs left unchanged. */
int ap_guiding_scan_u32(const uint32_t* in, uint32_t* out, size_t n, uint32_t* total);
int ap_guiding_sort_pairs_u64(const uint64_t* keys_in, uint64_t* keys_out, const uint32_t* vals_in, uint32_t* vals_out,
                              size_t n, int32_t begin_bit, int32_t end_bit);

#ifdef __cplusplus
}
#endif
#endif /* APERTURE_AP_GUIDING_H */

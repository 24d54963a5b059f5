/*
 * Path-tracing entry points of the Aperture core: the reference's
 * RenderCamera -> CalculateRadiance megakernel (core/scene.cu:201-255,
 * core/render/render.cuh:460-999) re-designed as a wavefront integrator
 * (generate, extend, shade, shadow, resolve, accumulate) over compacted
 * ray queues. Single-wavelength spectral transport with the reference's
 * NEE, MIS, Russian roulette, epsilons and firefly clamp; the constants the
 * reference hard-codes are parameters here, with its values as defaults.
 */
#ifndef APERTURE_AP_RENDER_H
#define APERTURE_AP_RENDER_H

#include "ap_core.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Camera (core/camera.hpp:56-63 + the pinhole of core/scene.cu:214-230).
 * direction is normalised like Camera::SetDirection. */
typedef struct ap_camera {
    float origin[3];
    float direction[3];
    float fov_deg;      /* vertical FOV; the reference fixes 60 (scene.cu:220) */
    int32_t width;
    int32_t height;
} ap_camera;

/* AP_TRIANGLE_MATERIAL_* : which GBuffer a triangle hit carries. OBJECT
 * uses the triangle's own material, as the cube and sphere intersectors do
 * (aabb.cuh:86, sphere.cuh:174); REFERENCE keeps IntersectTriangleObject's
 * default-constructed GBuffer (white albedo, roughness 0, no emission,
 * intersection.cuh:17,63), the shading model still coming from the object. */
#define AP_TRIANGLE_MATERIAL_OBJECT 0
#define AP_TRIANGLE_MATERIAL_REFERENCE 1

typedef struct ap_render_params {
    int32_t spp;               /* samples per pixel added by this call (one frame each) */
    int32_t frame_index;       /* FRAME_INDEX of the first sample: running-mean weight 1/(1+frame) */
    int32_t max_depth;         /* MAX_TRACING_DEPTH (config.cuh:8) = 1000 */
    int32_t rr_min_depth;      /* MIN_TRACING_DEPTH (config.cuh:7) = 3 */
    int32_t nee_max_depth;     /* NEE_MAX_NON_DELTA_DEPTH (config.cuh:25) = 10 */
    int32_t use_nee;           /* SceneSetting::bUseNEE_ = 1 */
    int32_t nee_always;        /* 0: NEE with probability sqrt(saturate(mask)) (render.cuh:777); 1: always */
    int32_t sky;               /* 1: SkyLighting on a miss (render.cuh:58-73); 0: black */
    int32_t triangle_material; /* AP_TRIANGLE_MATERIAL_* (default OBJECT) */
    int32_t use_guiding;       /* SceneSetting::bUsePathGuiding_ (probes, when built) */
    uint64_t seed;             /* PCG32 seed: stream per (pixel, frame) */
    int32_t wavelengths;       /* 1 (or 0): one wavelength per path (the reference's CalculateRadiance);
                                * 4: hero-wavelength sampling: the hero drawn by SampleCIE1931 and three
                                * rotations by 105 nm over [360, 780], one path for all four, the pixel
                                * estimate sum_j xyz(l_j) L(l_j) / sum_k p(l_k). */
    int32_t guide_frames;      /* with use_guiding: frames per probe update. 1 (or 0): after every frame,
                                * as SceneRenderer::Render then UpdatePathGuidingSample (scene.cu:308-386);
                                * F > 1: the frames of a call go in groups of F that sample the same
                                * probes (built from the groups before) and are traced together, then
                                * one update folds the group's samples in, frame after frame in slot
                                * order (the last group of a call may be shorter). */
} ap_render_params;

/* Caller-owned device framebuffers (GUI.hpp:232-274 owns them in the
 * reference). hdr is required; every other pointer may be NULL. */
typedef struct ap_framebuffers {
    float* hdr;          /* W*H*3 running mean of linear XYZ (scene.cu:241-242) */
    float* odd;          /* W*H*3 running mean of odd frames (scene.cu:243-253) */
    float* even;         /* W*H*3 running mean of even frames */
    double* sum_xyz;     /* W*H*3 FP64 per-pixel sums of the clamped XYZ samples (added) */
    int32_t* primary_id; /* W*H closest-hit objectIndex_ of the camera ray of the last sample, -1 on miss */
    uint64_t* ray_count; /* 2 counters, added to: rays traced by closest-hit [0] and shadow [1] launches */
    ap_screen_gbuffer* gbuffer;       /* W*H, updated per frame as RecordScreenGBuffer (denoise.cuh:109-166); may be NULL */
    ap_screen_statistics* statistics; /* W*H, updated per sample as RecordScreenStatisticBuffer
                                       * (denoise_statistic.cuh:36-70); may be NULL */
} ap_framebuffers;

/* Reference defaults (config.cuh), spp = 1, frame 0, seed 0, one wavelength per path. */
void ap_render_params_default(ap_render_params* params);

/* Replaces SceneRenderer::Render(hdr, ldr, size, origin, dir, frameIndex, ...)'s
 * RenderCamera dispatch (core/scene.cu:308-346, 257-274): adds params->spp
 * frames (frame_index, frame_index+1, ...) to the framebuffers. Asynchronous
 * on `stream`; concurrent renders of one scene race on its path buffers.
 * Materials: diffuse, light, GGX and glass (the exact dielectric with
 * n(lambda) from the Abbe number and Beer-Lambert absorption with the tilted
 * sigmaA, material.cuh:122-169, 221-272; media as CalculateRadiance tracks
 * them, render.cuh:84-160); albedo textures through the scene's texture cache
 * (ap_mesh.h). */
int ap_render(ap_scene* scene, const ap_camera* camera, const ap_render_params* params,
              const ap_framebuffers* fb, void* stream);

/* Replaces SceneRenderer::Render(size, origin, dir, sampleNum)
 * (core/scene.cu:276-306): frames 0..spp-1 into a zero-initialised target,
 * returned as host float[W*H*3] linear XYZ. Synchronous. */
int ap_render_offline(ap_scene* scene, const ap_camera* camera, const ap_render_params* params,
                      float* host_xyz);

/* Replaces XYZ2SRGBLinearD65 (core/tonemap.cuh:116-126) over n device XYZ triples. */
int ap_xyz_to_srgb(const float* xyz, size_t n, float* rgb, void* stream);

/* Tone curves of ap_tonemap (toneType of scene.cu:176-198). */
#define AP_TONE_LINEAR 0
#define AP_TONE_GAMMA 1  /* x^(1/2.2) (tonemap.cuh:6-9) */
#define AP_TONE_ACES 2   /* the ACES fit of tonemap.cuh:67-113 */

/* What ap_tonemap shows (SceneSetting's debug flags, scene.cu:405-447). */
#define AP_DISPLAY_IMAGE 0           /* hdr: XYZ -> linear sRGB x exposure -> tone curve */
#define AP_DISPLAY_ALBEDO 1          /* gbuffer albedo */
#define AP_DISPLAY_NORMAL 2          /* gbuffer normal * 0.5 + 0.5 */
#define AP_DISPLAY_DEPTH 3           /* viridis(1 / (depth + 1)) */
#define AP_DISPLAY_METALLIC 4        /* viridis(metallic) */
#define AP_DISPLAY_ROUGHNESS 5       /* viridis(roughness) */
#define AP_DISPLAY_STATISTICS_NUM 6  /* red: log(num + 1) / 100 */
#define AP_DISPLAY_STATISTICS_MEAN 7 /* |mean| / 4 */
#define AP_DISPLAY_STATISTICS_M2 8   /* |M2 / num| / 4 */
#define AP_DISPLAY_STATISTICS_M3 9   /* |M3 / num| / 4 */

/* Replaces the Tonemap kernel and its dispatch (core/scene.cu:88-199,
 * 400-447): width*height pixels into ldr as 0xff000000 | r << 16 | g << 8 |
 * b, each channel 255 * saturate(v) truncated. The image mode reads hdr
 * (linear XYZ running means); the G-buffer and statistics modes read
 * gbuffer / statistics (the others may be NULL). exposure replaces the
 * reference's device global (SetExposure, scene.cu:1167). Asynchronous. */
int ap_tonemap(const float* hdr, const ap_screen_gbuffer* gbuffer, const ap_screen_statistics* statistics,
               uint32_t* ldr, int32_t width, int32_t height, int32_t tone_type, int32_t display, float exposure,
               void* stream);

#ifdef __cplusplus
}
#endif
#endif /* APERTURE_AP_RENDER_H */

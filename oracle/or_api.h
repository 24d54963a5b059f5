// ORACLE — test infrastructure only. C entry points of liboracle.so, the
// CPU restatement of the reference algorithms used as the checker by
// tests/, bench.py's cpu_baseline leg and __graft_entry__.smoke().
#pragma once

#include <stddef.h>
#include <stdint.h>

#include "aperture/ap_core.h"
#include "aperture/ap_render.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_bvh or_bvh;

or_bvh* or_bvh_build(const ap_scene_object* objs, size_t n, int leaf_size, int bin_count);
/* The SDF cache's volumes for SDF_VOLUME objects: n volumes, dims[3k..3k+2],
 * samples concatenated in volume order (or_bvh.cpp; process-wide). */
void or_sdf_set_volumes(const float* values, const int32_t* dims, size_t n);
void or_sdf_fill(int functor, const float* info, int nx, int ny, int nz, float* out);
int or_sdf_hit(const ap_scene_object* objs, int idx, const ap_ray* r, float eps, float* pos, float* normal);
void or_bvh_free(or_bvh* h);
size_t or_bvh_node_count(const or_bvh* h);
void or_bvh_export(const or_bvh* h, ap_bvh_node* nodes, int32_t* indices);
/* cull = 0: the reference TraverseBVH; cull = 1: with the core's CULLED
 * line rejection of boxes grown by margin. */
void or_trace(const or_bvh* h, const ap_ray* rays, size_t n, int any_hit, int cull, float margin,
              int32_t* ids, float* ts, uint32_t* nodes, uint32_t* prims, int threads);
void or_brute_closest(const ap_scene_object* objs, size_t count, const ap_ray* rays, size_t n,
                      int32_t* ids, float* ts, int threads);
float or_object_t(const ap_scene_object* objs, int idx, const ap_ray* r);

/* Integrator restatement (or_render.cpp): frames frame_index .. +spp-1 into
 * hdr (running mean, in place), sums (added) and primary ids (last frame). */
void or_nrc_encode(const float* q, size_t n, const float* bmin, const float* bmax, float* out);
void or_nrc_forward(const float* params, const float* feat, size_t n, int round_bf16, float* out);
double or_nrc_grad(const float* params, const float* feat, const float* target, size_t n, int round_bf16,
                   double* grad);
double or_nrc_loss_fixed(const float* params, const float* feat, const float* target, const double* den, size_t n,
                         int round_bf16);
// Adam over n parameters (any network width or number of networks).
void or_nrc_adam_n(float* params, float* m, float* v, const float* grad, size_t n, int t, float lr, float beta1,
                   float beta2, float eps);
void or_nrc_adam(float* params, float* m, float* v, const float* grad, int t, float lr, float beta1, float beta2,
                 float eps);
size_t or_nrc_params_w(int width);
void or_nrc_forward_w(const float* params, int width, const float* feat, size_t n, int mode, float* out);
double or_nrc_grad_w(const float* params, int width, const float* feat, const float* target, size_t n, int mode,
                     double scale, double* grad);
double or_nrc_loss_fixed_w(const float* params, int width, const float* feat, const float* target, const double* den,
                           size_t n, int mode);
void or_f16_round(const float* x, size_t n, float* out);
typedef struct or_guide or_guide;
void or_pg_uv_to_sphere(const float* uv, size_t n, float* out);
void or_pg_sphere_to_uv(const float* d, size_t n, float* out);
void or_pg_voxel(const float* pos, const float* dir, size_t n, int32_t* out);
void or_pg_mipmaps(ap_octahedron* p);
void or_pg_pdf(const ap_octahedron* p, const float* d, size_t n, const float* q, const float* normal, float* out);
void or_pg_sample(const ap_octahedron* p, uint64_t seed, size_t n, const float* q, const float* normal, float* dirs,
                  float* pdfs);
void or_pg_update(or_guide* g, const ap_path_guiding_sample* slots, size_t n);
or_guide* or_guide_create(void);
void or_guide_free(or_guide* g);
size_t or_guide_probe_count(const or_guide* g);
void or_guide_export(const or_guide* g, ap_octahedron* probes, int32_t* volume, int32_t* bounds);
int or_render_guided(const or_bvh* bvh, const ap_scene_object* objs, size_t n, const ap_camera* cam,
                     const ap_render_params* prm, float cull_margin, float* hdr, double* sums, int32_t* primary,
                     int threads, or_guide* guide);
int or_render(const or_bvh* bvh, const ap_scene_object* objs, size_t n, const ap_camera* cam,
              const ap_render_params* prm, float cull_margin, float* hdr, double* sums, int32_t* primary,
              int threads);
int or_render_nrc(const or_bvh* bvh, const ap_scene_object* objs, size_t n, const ap_camera* cam,
                  const ap_render_params* prm, float cull_margin, const float* params, const float* bmin,
                  const float* bmax, float spread, uint32_t train_paths, uint32_t unbiased_every, float* hdr,
                  double* sums, float* samples, uint8_t* valid, uint32_t* n_queries, int threads);
int or_render_nrc_w(const or_bvh* bvh, const ap_scene_object* objs, size_t n, const ap_camera* cam,
                    const ap_render_params* prm, float cull_margin, const float* params, int width, int mode,
                    const float* bmin, const float* bmax, float spread, uint32_t train_paths,
                    uint32_t unbiased_every, float* hdr, double* sums, float* samples, uint8_t* valid,
                    uint32_t* n_queries, int threads);
typedef struct or_guide or_guide;
int or_render_nrc_guided(const or_bvh* bvh, const ap_scene_object* objs, size_t n, const ap_camera* cam,
                         const ap_render_params* prm, float cull_margin, const float* params, int width, int mode,
                         const float* bmin, const float* bmax, float spread, uint32_t train_paths,
                         uint32_t unbiased_every, float* hdr, double* sums, float* samples, uint8_t* valid,
                         uint32_t* n_queries, int threads, or_guide* guide);

/* Spectral restatement (or_spectrum.cpp). Tables come from the core's data files. */
int or_spectrum_load(const char* mlp_path, const char* cie_path);
void or_rgb_encode(const float* rgb, size_t n, float* out);
void or_lambda_encode(const float* lam, size_t n, float* out);
float or_spectral_interact(const float* rgbK, const float* lamK);
void or_query_rgb_basis(const float* rgb, size_t n, float* out);
void or_query_lambda(const float* lam, size_t n, float* out);
void or_interact(const float* lam, const float* rgb, size_t n, float* out);
void or_fmath(int fn, const float* x, size_t n, float* out);
float or_ndtrif(float u);
void or_sample_cie1931(const float* u3, size_t n, float* lambda, float* pdf);
float or_pdf_cie1931(float lambda);
float or_d65_norm(float lambda);
void or_lambda_to_xyz(float lambda, float* xyz);
void or_xyz_to_srgb(const float* xyz, float* rgb);

/* Display and statistics framebuffers (or_display.cpp). or_gsample is what
 * RecordScreenGBuffer reads of a camera ray's first hit (the layout of the
 * core's per-path record): albedo, metallic, normal, roughness, hit
 * distance, shading model (4 for a miss), hit flag. */
typedef struct or_gsample {
    float albedo[3];
    float metallic;
    float normal[3];
    float roughness;
    float depth;
    int32_t model;
    int32_t hit;
    int32_t pad;
} or_gsample;
void or_half_round(const float* x, size_t n, uint16_t* out);
void or_record_gbuffer(ap_screen_gbuffer* g, const or_gsample* samples, size_t n);
void or_record_statistics(ap_screen_statistics* h, const float* xyz, size_t n);
void or_tonemap(const float* hdr, const ap_screen_gbuffer* gb, const ap_screen_statistics* st, uint32_t* ldr,
                size_t n, int tone, int display, float exposure);
/* or_render with the screen G-buffer and statistics recorded per frame in
 * frame order (either may be NULL). */
int or_render_screen(const or_bvh* bvh, const ap_scene_object* objs, size_t n, const ap_camera* cam,
                     const ap_render_params* prm, float cull_margin, float* hdr, double* sums,
                     ap_screen_gbuffer* gbuffer, ap_screen_statistics* statistics, int threads);

/* Pins of the BSDFs and light samplers (or_render.cpp). or_plant(k) turns on
 * planted mistake k (0 = none) for the tests that show a pin fails. */
void or_plant(int which);
void or_ggx_vndf_sample(const float* V, const float* N, float alpha, const float* rnd, size_t n, float* out);
void or_ggx_vndf_pdf(const float* V, const float* N, const float* L, size_t n, float alpha, float* out);
void or_bsdf_eval(int model, const float* mat, float lambda, const float* rayDir, const float* N, const float* L,
                  size_t n, float* f, float* pdf);
void or_bsdf_sample(int model, const float* mat, float lambda, const float* rayDir, const float* N, uint64_t seed,
                    size_t n, float* out);
void or_nee_sample(const ap_scene_object* objs, size_t n, const float* p, size_t np, uint64_t seed, float* out);
void or_nee_pdf(const ap_scene_object* objs, size_t n, int light, const float* q, size_t m, float* out);

/* Visit accounting (or_count.h): or_counts_begin zeroes the totals and
 * turns counting on; or_counts_end turns it off and copies the OR_COUNT_N
 * totals: closest-hit rays, their node records and primitive tests, any-hit
 * rays, node records and tests, surface vertices, Interact evaluations,
 * probe queries, probes found, camera paths. Not for concurrent renders. */
#define OR_COUNT_N 11
void or_counts_begin(void);
void or_counts_end(uint64_t* out);

/* FP64 image statistics (or_measure.cpp): DeviceMSE, DeviceMSELuminance and
 * relMSE; image kinds as AP_IMAGE_XYZ_F32 (0) / AP_IMAGE_XYZ_SUMS_F64 (1). */
double or_mse(const float* a, const float* b, size_t n);
double or_mse_luminance(const float* a, const float* b, size_t n);
double or_relmse(const void* img, int img_kind, double img_scale, const void* ref, int ref_kind, double ref_scale,
                 size_t n, double eps);

/* Albedo textures for or_render (or_render.cpp): n RGBA8 images of
 * dims[2k] x dims[2k+1] texels (rows from the top) concatenated in rgba;
 * texture handle k + 1 names image k. or_texture_sample: tex2D of image
 * `index` at n UV pairs, RGBA out. */
void or_texture_set(const uint8_t* rgba, const int32_t* dims, size_t n);
void or_texture_sample(int32_t index, const float* uv, size_t n, float* out);

/* Glass at one wavelength (or_render.cpp): IORFromAbbeNM, SigmaAFromTiltNM,
 * FresnelDielectricExact, BeerLambert_Spectrum, and SampleBSDF followed by
 * EvaluateBSDFWithCos of MAT_GLASS. mat = ior, abbe, sigma_a, tilt, albedo
 * rgb; the MediaContext is (inside, depth, vertex); the Fresnel draw comes
 * from PCG32 seeded with seed; out = next direction (3), pdf, brdf, inside
 * after, delta flag. */
float or_glass_ior(float lambda, float nd, float vd);
float or_glass_sigma(float sigma_a, float lambda, float tilt);
float or_glass_fresnel(float cosI, float cosT, float eta);
float or_glass_beer(float distance, float resp, float sigma_a);
void or_glass_bsdf(const float* d, const float* n, const float* mat, float lambda, int inside, int depth,
                   const float* vertex, const float* hit, uint64_t seed, float* out);

#ifdef __cplusplus
}
#endif

/*
 * Aperture MI355X core: mesh and texture ingestion.
 *
 * Textured triangle meshes enter a scene as triangle SceneObjects (one per
 * triangle, as the reference's LoadTexturedMesh makes them) whose GBuffer
 * carries texture handles; the handles name images held in an
 * ap_texture_cache bound to the scene. The integrator samples the albedo
 * texture where the reference does (SampleBSDF and EvaluateBSDFWithCos,
 * core/render/material.cuh:69-74, 198-202): albedo = texel.rgb^2 at the
 * hit's interpolated UV (triangles; other objects sample UV (0, 0)).
 *
 * Sampling restates tex2D<float4> over the reference's texture descriptor
 * (core/helper/texture_loader.cuh:36-46: normalised coordinates, wrap in both
 * axes, bilinear filter, 8-bit channels read as value / 255) with the
 * filtering CUDA documents for linear mode: x = frac(u) W - 1/2, weights
 * alpha = frac(x) kept to 8 fractional bits (rounded to nearest here), the
 * four wrapped neighbours blended (1-a)(1-b), a(1-b), (1-a)b, ab. The
 * device kernels and the CPU checker compute it the same way, so they agree
 * bit for bit; an NVIDIA texture unit agrees to within that 1/256 weight
 * quantum.
 */
#ifndef APERTURE_AP_MESH_H
#define APERTURE_AP_MESH_H

#include "ap_core.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ap_texture_cache ap_texture_cache;

/* The images of the reference's cudaTextureObject_t handles (LoadCUDATexture,
 * core/helper/texture_loader.cuh:10-55), on the current device. */
int ap_texture_cache_create(ap_texture_cache** out);
int ap_texture_cache_destroy(ap_texture_cache* cache);

/* LoadCUDATexture from decoded pixels (what stbi_load(path, ..., 4) returns):
 * width x height RGBA8 texels, rows from the top, copied to the device.
 * *handle (>= 1) goes into GBuffer.albedo_tex (or the other texture slots,
 * which the integrator does not read, as in the reference). At most
 * AP_TEXTURE_CACHE_CAPACITY images per cache, each at most 32768 wide/high. */
#define AP_TEXTURE_CACHE_CAPACITY 4096
int ap_texture_cache_add_rgba8(ap_texture_cache* cache, const uint8_t* rgba, int32_t width, int32_t height,
                               uint64_t* handle);

/* Number of images in the cache. */
int ap_texture_cache_count(const ap_texture_cache* cache, uint32_t* count);

/* tex2D<float4>(handle, u, v) as the integrator computes it, for n UV pairs
 * (host arrays): out[4 i ..] = RGBA in [0, 1]. */
int ap_texture_sample(const ap_texture_cache* cache, uint64_t handle, const float* uv, size_t n, float* out);

/* Binds the cache whose images the scene's texture handles name (NULL
 * unbinds). Rendering a scene whose materials name a handle the bound cache
 * does not hold returns AP_ERR_INVALID_ARG. The cache must outlive the
 * scenes bound to it (unbind or destroy them first); images may be added
 * while it is bound, between renders. */
int ap_scene_set_texture_cache(ap_scene* scene, const ap_texture_cache* cache);

/* LoadTexturedMesh (core/helper/scene_helper.h:453-564) over an OBJ file read
 * as tinyobjloader 2.0 reads it (core/helper/tiny_obj_loader.h, vendored in
 * the reference): v / vt / vn / f records (1-based or negative relative
 * indices; v, v/vt, v//vn, v/vt/vn), numbers parsed to double then float,
 * quads split along their shorter diagonal (v0 v1 v2 + v0 v2 v3 when
 * |v2 - v0|^2 < |v3 - v1|^2, else v0 v1 v3 + v1 v2 v3), larger faces
 * ear-clipped as tinyobjloader's built-in triangulation clips them, faces of
 * fewer than three vertices skipped, an index out of range an error
 * (AP_ERR_INVALID_ARG). Each triangle becomes one SceneObject: centre
 * `position`, extent (scale, scale, scale), identity rotation, a GGX
 * material (white, roughness 0.4, metallic 0, specular 0) carrying the four
 * texture handles, vertex positions scaled by `scale`, normals as read
 * (else (0, 1, 0)), UV (u, 1 - v) (else (0, 0)); when the face's first
 * vertex has no normal, all three get normalize(cross(p1 - p0, p2 - p0)).
 * Proxy bounds as UpdateProxy. UpdateProxy leaves every triangle's proxy
 * centre at the mesh position, so the reference's SAH builder, which bins
 * centroids, puts the whole mesh in one leaf (bvh.cuh:232-234); by default
 * the proxy centre here is the triangle's centroid ((p0 + p1 + p2) / 3 in
 * world space), which gives the mesh a real tree and the same hits (ties at
 * equal distance aside). AP_MESH_REFERENCE_PROXY keeps the reference's
 * centres. With out == NULL only *count is set; otherwise up to `capacity`
 * objects are written and *count is the number of triangles.
 * textures: albedo, normal, metallic, roughness handles (0 = none). */
#define AP_MESH_REFERENCE_PROXY 1
int ap_mesh_load_obj(const char* path, const float position[3], float scale, const uint64_t textures[4],
                     int32_t flags, ap_scene_object* out, size_t capacity, size_t* count);

#ifdef __cplusplus
}
#endif

#endif /* APERTURE_AP_MESH_H */

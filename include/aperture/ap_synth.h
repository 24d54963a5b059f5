/*
 * Synthetic workload generators of the Aperture core (host C++, exported
 * from libaperture_hip.so). They have no counterpart in the reference, which
 * loads OBJ meshes and hand-built scenes (core/helper/scene_helper.h); they
 * build the benchmark scenes and ray sets deterministically from PCG32 seeds
 * so that the GPU path and the CPU oracle consume identical inputs.
 */
#ifndef APERTURE_AP_SYNTH_H
#define APERTURE_AP_SYNTH_H

#include "ap_core.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Triangle soup: n SceneObjects of type OBJ_TRIANGLE. Centroids uniform in
 * the unit cube, edge length log-uniform in [edge_min, edge_max], equilateral
 * triangles with a uniformly random rotation (uniform unit quaternion).
 * Each object's center_ is its centroid (the proxy centroid the reference
 * builder splits on, core/render/object.cuh:309-341), objectToWorld_ is the
 * identity and the vertices are relative to the centroid. Diffuse white. */
int ap_synth_triangle_soup(size_t n, uint64_t seed, float edge_min, float edge_max,
                           ap_scene_object* out);

/* Pinhole rays through pixel centres with the reference camera basis
 * (core/scene.cu:214-230; vertical FOV given in degrees instead of the fixed
 * 60). tmin = 0, tmax = +inf. Row-major, index j*width + i. */
int ap_synth_camera_rays(int width, int height, const float origin[3], const float direction[3],
                         float fov_deg, ap_ray* out);

/* One secondary ray per primary hit (in primary order): origin = hit point +
 * offset * n with n the geometric normal turned toward the incoming ray,
 * direction cosine-weighted about n (SampleCosineWeightedHemisphere,
 * core/render/sample.cuh:88-108) from PCG32(seed, stream = primary index).
 * Writes *n_out rays; out must hold n entries. Host arrays. */
int ap_synth_secondary_rays(const ap_scene* scene, const ap_ray* primaries, const ap_hit* hits,
                            size_t n, uint64_t seed, float offset, ap_ray* out, size_t* n_out);

/* Closed-box scene: five axis-aligned quads of [-1,1]^3 (floor, ceiling,
 * back, left, right; the +z side open), two triangles each; a 0.5 x 0.5
 * emissive quad under the ceiling as a thin OBJ_CUBE (centre (0, 0.99, 0),
 * extent (0.25, 0.01, 0.25), radiance `light`, the NEE-capable light form of
 * SceneHelper::CreateCornellBox, core/helper/scene_helper.h:14-72); two
 * icospheres (subdivision `sphere_subdiv`, radius 0.3) at (-0.45,-0.7,-0.3) and
 * (0.45,-0.7,0.35) with vertex normals. Diffuse albedo per mesh drawn
 * U[0.1, 0.9] per channel from PCG32(seed). Writes up to `capacity` objects
 * and the count to *n_out (10 + 1 + 2 * 20 * 4^subdiv). */
int ap_synth_box_scene(uint64_t seed, int sphere_subdiv, float light, ap_scene_object* out, size_t capacity,
                       size_t* n_out);

/* Heightfield scene: a grid x grid cell heightfield over [-1,1]^2 in x/z
 * (2 grid^2 triangles, vertex normals from central differences), height a
 * 6-octave value-noise fBm of amplitude 0.2 (lattice values from SplitMix64
 * of the seed); n_spheres icospheres of subdivision 2 (320 triangles each),
 * radius log-uniform in [0.01, 0.1], resting on the surface; four 0.3 x 0.3
 * thin-cube lights of radiance `light` 1.15-1.3 above it. Materials per mesh:
 * Lambert or GGX with equal odds, roughness U[0.05, 1], albedo U[0.05, 0.95]
 * per channel. Writes up to `capacity` objects and the count to *n_out
 * (2 grid^2 + 320 n_spheres + 4); out = NULL only reports the count. */
int ap_synth_heightfield_scene(int grid, int n_spheres, uint64_t seed, float light, ap_scene_object* out,
                               size_t capacity, size_t* n_out);

#ifdef __cplusplus
}
#endif
#endif /* APERTURE_AP_SYNTH_H */

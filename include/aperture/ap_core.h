/*
 * Aperture MI355X core: C ABI.
 *
 * One shared library, libaperture_hip.so, exports these entry points. All
 * return an int status (AP_OK = 0, negative on error), never throw and never
 * abort; ap_last_error_string() gives the message of the calling thread's
 * last failure. Device pointers are owned by the caller; host arrays are
 * copied. Scene state is passed to every kernel by argument, so any number
 * of scenes may coexist on one device (the reference keeps one scene per
 * device in __constant__ symbols, core/render/render.cuh:24-56).
 *
 * Each declaration names the reference interface it replaces.
 */
#ifndef APERTURE_AP_CORE_H
#define APERTURE_AP_CORE_H

#include "ap_layout.h"

#ifdef __cplusplus
extern "C" {
#endif

#define AP_OK 0
#define AP_ERR_INVALID_ARG (-1)
#define AP_ERR_HIP (-2)
#define AP_ERR_OUT_OF_MEMORY (-3)
#define AP_ERR_UNSUPPORTED (-4)
#define AP_ERR_NO_DEVICE (-5)

typedef struct ap_scene ap_scene;

/* Traversal culling. REFERENCE reproduces the reference slab test literally
 * (accept iff tExit >= max(0,tMin) && tEnter <= tMax, bvh.cuh:424-470, which
 * never checks tEnter <= tExit, so every box in front of the ray passes until
 * a hit shrinks tMax). CULLED (the default) also rejects boxes whose grown
 * bounds the ray's line misses; its visit order is a subsequence of the
 * reference order that keeps every box able to hold an accepted hit, so it
 * returns the same ids and distances. */
#define AP_TRAVERSE_REFERENCE 0
#define AP_TRAVERSE_CULLED 1

/* One ray, 32 B: origin, tmin, direction, tmax. The direction is used as
 * given (TraverseBVH does not normalise it, intersection.cuh:235-246). */
typedef struct ap_ray {
    float ox, oy, oz, tmin;
    float dx, dy, dz, tmax;
} ap_ray;

/* One hit, 8 B: objectIndex_ of the accepted SceneObject (-1 on a miss)
 * and its distance_ (tmax on a miss, +inf for an empty scene), as
 * IntersectionContext reports them (intersection.cuh:248-253, 160-169). */
typedef struct ap_hit {
    int32_t id;
    float t;
} ap_hit;

/* Per-ray traversal work in the reference visit order: node records read
 * (each popped node plus both children of every expanded interior node,
 * intersection.cuh:273, 313-314) and primitive tests (:299-309). */
typedef struct ap_trace_stats {
    uint32_t nodes;
    uint32_t prims;
} ap_trace_stats;

/* BuildBVH(proxy, leafSize = 8, binCount = 12) (core/bvh/bvh.cuh:142). */
typedef struct ap_build_params {
    int32_t leaf_size;
    int32_t bin_count;
} ap_build_params;

typedef struct ap_bvh_info {
    int64_t node_count;      /* reference BVHNode count (DFS pre-order, root 0) */
    int64_t interior_count;  /* 64-B device records (one per interior node) */
    int64_t prim_count;
    int32_t max_depth;       /* root = depth 0 */
    int32_t stack_capacity;  /* per-lane traversal stack entries used */
    int64_t device_bytes;    /* node records + primitive records + analytic records */
    int32_t light_count;
    int32_t traversal_mode;  /* AP_TRAVERSE_* */
    float line_margin;       /* world-space box growth of the CULLED line test */
    int32_t pad;
} ap_bvh_info;

/* Replaces SceneRenderer::SceneRenderer(const std::vector<SceneObject>&)
 * (core/scene.hpp:113) together with SceneRenderer::UpdateSceneToDeviceIfDirty
 * (core/scene.cu:1025-1118): copies the objects, stamps objectIndex_ and
 * lightIndex_, builds the binned-SAH BVH on the host (bvh.cuh:142-422), builds
 * the light selection table (scene.cu:1056-1080) and uploads the device
 * layouts to the current HIP device. Triangles stay in the traversal (the
 * reference's CUDA path retypes them to 999 and never hits them,
 * scene.cu:1040-1044). params may be NULL (defaults 8, 12). */
int ap_scene_create(const ap_scene_object* objects, size_t count,
                    const ap_build_params* params, ap_scene** out_scene);

/* Replaces BuildBVH(const std::vector<PrimitiveProxy>&, leafSize, binCount)
 * (core/bvh/bvh.cuh:142-422) over the objects' proxy_ fields, host only (no
 * device needed). nodes must hold 2*count entries (a binary tree over count
 * leaves has at most 2*count-1 nodes), indices count entries; *node_count
 * receives the number of nodes written. */
int ap_build_bvh(const ap_scene_object* objects, size_t count, const ap_build_params* params,
                 ap_bvh_node* nodes, int32_t* indices, size_t* node_count);

/* Replaces SceneRenderer::~SceneRenderer (core/scene.cu, FreeSceneRelatedDevicePtr). */
int ap_scene_destroy(ap_scene* scene);

int ap_scene_get_info(const ap_scene* scene, ap_bvh_info* out_info);

/* Selects AP_TRAVERSE_REFERENCE or AP_TRAVERSE_CULLED for later traces. No
 * reference counterpart (the reference has one compiled-in test). */
int ap_scene_set_traversal_mode(ap_scene* scene, int mode);

/* Launch shape of the traversal kernels: rays owned by one 64-lane wave
 * (a positive multiple of 64, at most 65536) and the number of idle lanes
 * (1..64) at which a wave refills them with its next rays. A value of 0
 * keeps the current setting. Results are identical for every shape; only
 * speed changes. No reference counterpart (the reference launches one
 * thread per pixel). Defaults: 512 rays, refill at 16 idle lanes. The
 * renderer's own path queues (ap_render) use at most 512 rays per wave.
 * Consecutive chunks of rays go to consecutive waves, dealt round-robin
 * over the XCDs. */
int ap_scene_set_launch(ap_scene* scene, int32_t wave_rays, int32_t refill);
int ap_scene_get_launch(const ap_scene* scene, int32_t* wave_rays, int32_t* refill);

/* Autotuner of the launch shape: traces the caller's device ray batch
 * (closest hit, or any hit when any_hit != 0) into `hits` under each
 * candidate shape (rays per wave 128, 192, 256, 320, 384, 448, 512, 640,
 * 768 or 1024 x refill 4/8/16; one untimed launch, then the faster of two
 * launches timed with HIP events on `stream`), keeps the fastest on the
 * scene and reports it. Synchronises
 * `stream`. The hits written are those of an ordinary trace. No reference
 * counterpart. */
int ap_scene_autotune(ap_scene* scene, const ap_ray* rays, size_t n, ap_hit* hits, int any_hit, void* stream,
                      int32_t* best_wave_rays, int32_t* best_refill, float* best_ms);

/* Copies out the reference-layout BVH (BVHHost::nodes_, objectIndices_,
 * bvh.cuh:54-58) and light table (AdditionalLightInfo[], scene.cu:1056-1080)
 * to host arrays sized from ap_scene_get_info. Any pointer may be NULL. */
int ap_scene_export_bvh(const ap_scene* scene, ap_bvh_node* nodes, int32_t* object_indices,
                        ap_light_info* lights);

/* Replaces intersection::TraverseBVH<false> (core/render/intersection.cuh:235-370)
 * over SceneObjectsIntersector (:140-172): closest hit for n rays, same
 * primitive tests, same acceptance window [tmin, tmax), same near-first visit
 * order with the left child first on ties. rays/hits are device pointers;
 * stream is a hipStream_t (NULL = default stream). Asynchronous. */
int ap_trace_closest(const ap_scene* scene, const ap_ray* rays, size_t n, ap_hit* hits,
                     void* stream);

/* Replaces intersection::TraverseBVH<true> as reached from AnyHitSegmentBVH
 * (intersection.cuh:372-397): the first accepted primitive in visit order. */
int ap_trace_any(const ap_scene* scene, const ap_ray* rays, size_t n, ap_hit* hits,
                 void* stream);

/* ap_trace_closest that also writes per-ray ap_trace_stats (device pointer).
 * No reference counterpart: it instruments the same visit order for the
 * algorithmic-bytes accounting of the traversal roofline. */
int ap_trace_closest_stats(const ap_scene* scene, const ap_ray* rays, size_t n, ap_hit* hits,
                           ap_trace_stats* stats, void* stream);

/* Signed-distance objects. A SceneObject of type AP_OBJ_SDF with
 * additionalObjectInfo_.sdfType_ SDF_TORUS (0) is the torus of
 * sdfInfo0_.xy = (R, r) in the unit space of its box; SDF_VOLUME (1) samples
 * the volume cachedVolumeIndex_ of the scene's SDF cache; other types are
 * never hit (IntersectSDFObject, core/object/sdf/sdf_master.cuh:32-70). Rays
 * march up to 512 steps (RaymarchLocal, sdf_base.cuh:54-110) with epsilon
 * SDF_EPS = INTERSECT_EPS(depth) / 2: tmin / 2 for ap_trace_closest (tmin is
 * INTERSECT_EPS of the ray's depth in the integrator; 5e-4 when tmin is 0),
 * 5e-4 for ap_trace_any. SDF objects cannot be sampled as lights (as in
 * NEESampleLight, nee.cuh:330-377). */
#define AP_SDF_TORUS 0
#define AP_SDF_VOLUME 1
#define AP_SDF_FUNCTOR_TORUS 0    /* SdfTorus of info.xy (sdf_torus.cuh:15-26) */
#define AP_SDF_FUNCTOR_DIAMOND 1  /* SdfDiamondUnit of info (sdf_diamond.cuh:21-73) */

typedef struct ap_sdf_cache ap_sdf_cache;

/* Replaces SDFCacheManager (core/object/sdf/sdf_cache_manager.cuh:36-114):
 * up to 128 volumes on the current device. */
int ap_sdf_cache_create(ap_sdf_cache** out);
int ap_sdf_cache_destroy(ap_sdf_cache* cache);
/* AddVolume with a built-in functor: dims[0] x dims[1] x dims[2] samples
 * (x fastest) of the functor at (i / (dims - 1)) - 0.5, filled on the
 * device (FillSDFVolume, sdf_cache_manager.cuh:15-34). Synchronous;
 * *index receives the volume's cachedVolumeIndex_. */
int ap_sdf_cache_add_functor(ap_sdf_cache* cache, int32_t functor, const float info[4], const int32_t dims[3],
                             uint32_t* index);
/* AddVolume with caller-computed host samples (same layout). Synchronous. */
int ap_sdf_cache_add_volume(ap_sdf_cache* cache, const float* host_values, const int32_t dims[3], uint32_t* index);
/* A volume's dims and (if host_values is not NULL) its samples. */
int ap_sdf_cache_get_volume(const ap_sdf_cache* cache, uint32_t index, int32_t dims[3], float* host_values);
/* The SDF cache the scene's SDF_VOLUME objects read (NULL: none, such
 * objects are never hit). The cache must outlive the scene's use of it. */
int ap_scene_set_sdf_cache(ap_scene* scene, const ap_sdf_cache* cache);

/* Message of the calling thread's last failed call ("" if none). */
const char* ap_last_error_string(void);

#ifdef __cplusplus
}
#endif
#endif /* APERTURE_AP_CORE_H */

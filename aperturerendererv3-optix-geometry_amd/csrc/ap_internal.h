// Internal declarations of the Aperture core: device record layouts, the
// scene object behind the C ABI handle, and error plumbing.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "aperture/ap_core.h"
#include "aperture/ap_mesh.h"
#include "ap_vec.h"

namespace ap {

// ---------------------------------------------------------------------------
// Device layouts of the acceleration structure.
//
// NodeRec (64 B, 64-B aligned): one per INTERIOR node of the reference BVH,
// holding both children's boxes so that one 64-B read serves one expansion
// (the reference reads the two child BVHNodes separately,
// intersection.cuh:313-314). A child link >= 0 is the child's NodeRec index;
// a link < 0 is a leaf whose primitives start at slot ~link.
//
// PrimRec (48 B): one per primitive slot of objectIndices_ (bvh.cuh:49),
// slot-ordered so a leaf is a contiguous run; the last slot of each leaf
// carries PRIM_LAST. Triangles store world-space p0 and the edges
// e1 = p1 - p0, e2 = p2 - p0, computed once on the host with the same FP32
// operations the reference performs per test (intersection.cuh:27-33).
// ---------------------------------------------------------------------------
struct alignas(64) NodeRec {
    float lmin[3]; int32_t llink;
    float lmax[3]; int32_t rlink;
    float rmin[3]; int32_t pad0;
    float rmax[3]; int32_t pad1;
};
static_assert(sizeof(NodeRec) == 64, "NodeRec is one 64-B line");

enum : uint32_t {
    PRIM_LAST = 1u,             // last primitive of its leaf
    PRIM_KIND_SHIFT = 1u,
    PRIM_KIND_TRIANGLE = 0u,
    PRIM_KIND_NEVER = 1u,       // invisible or unsupported type: always a miss
    PRIM_KIND_CUBE = 2u,        // analytic OBB (core/object/aabb.cuh:28-89)
    PRIM_KIND_SPHERE = 3u,      // analytic sphere (core/object/sphere.cuh:37-70)
    PRIM_KIND_SDF = 4u,         // signed-distance object (core/object/sdf/sdf_base.cuh:115-163)
};

struct alignas(16) PrimRec {
    float p0[3]; int32_t obj;     // obj = SceneObject index (the hit ID)
    float e1[3]; uint32_t flags;  // PRIM_LAST | kind << PRIM_KIND_SHIFT
    float e2[3]; int32_t aux;     // analytic record index for CUBE/SPHERE/SDF
};
static_assert(sizeof(PrimRec) == 48, "PrimRec is 48 B");

// Transform data of an analytic (cube, sphere or SDF) primitive.
struct alignas(16) AnalyticRec {
    float center[3]; int32_t sdf_type;  // SDF objects: ESDFType; others -1
    float extent[3]; int32_t volume;    // SDF_VOLUME: cachedVolumeIndex_
    float w2o[9];    float pad2[3];     // worldToObject_ rows
    float o2w[9];    float pad3[3];     // objectToWorld_ rows
    float sdf_info[4];                  // SDFInfo::sdfInfo0_
};
static_assert(sizeof(AnalyticRec) == 144, "AnalyticRec is 144 B");

struct SdfVolume;  // ap_sdf.h
// The SDF cache's device volume table and its capacity (ap_sdf.hip), and a
// host view of volume k (samples in host memory).
const SdfVolume* sdf_cache_table(const ap_sdf_cache* c);
// Texture cache (ap_texture.hip): its device table of TexRec and image count.
struct TexRec;
const TexRec* texture_cache_table(const ap_texture_cache* c);
uint32_t texture_cache_count(const ap_texture_cache* c);
int32_t sdf_cache_capacity(const ap_sdf_cache* c);
bool sdf_cache_host_volume(const ap_sdf_cache* c, int32_t k, SdfVolume& v);

struct RootInfo {
    float bmin[3];
    float bmax[3];
    int32_t link;
    int32_t valid;   // 0 for an empty scene
};

struct DeviceScene {
    const NodeRec* nodes;
    const PrimRec* prims;
    const AnalyticRec* analytic;
    const SdfVolume* volumes;   // the SDF volume table (ap_scene_set_sdf_cache), or none
    int32_t n_volumes;
    int32_t has_sdf;            // the scene holds SDF objects (traversal runs the marching variant)
    float sdf_eps_any;          // SDF marching epsilon of any-hit launches (closest hit: tmin / 2)
    RootInfo root;
    int32_t stack_cap;
    float line_margin;   // world-space growth of boxes for the CULL line test
    int32_t mode;        // AP_TRAVERSE_REFERENCE or AP_TRAVERSE_CULLED
    int32_t refill;      // idle lanes that trigger a refill of a wave (1..64)
    int32_t wave_rays;   // rays owned by one wave (a positive multiple of 64)
};

// Reference-layout BVH build result (bvh.cuh:54-58, 32-bit node variant).
struct HostBVH {
    std::vector<ap_bvh_node> nodes;
    std::vector<int32_t> indices;
};

HostBVH build_bvh(const std::vector<ap_primitive_proxy>& proxies, int leaf_size, int bin_count);
int bvh_max_depth(const HostBVH& bvh);

}  // namespace ap

struct ap_nrc;  // learned radiance cache handle (ap_nrc.hip)

namespace ap {
struct RenderState;  // integrator buffers of a scene (ap_render.hip)
struct GuideDevice;  // guiding probes and build buffers of a scene (ap_guiding.hip)
void guide_free(GuideDevice* g);
void render_state_free(RenderState* r);
}  // namespace ap

struct ap_scene {
    int device = 0;
    std::vector<ap_scene_object> objects;     // host copy, indices stamped
    ap::HostBVH bvh;
    std::vector<ap_light_info> lights;
    int max_depth = 0;
    int64_t interior_count = 0;
    ap::DeviceScene dev{};
    void* d_nodes = nullptr;
    void* d_prims = nullptr;
    void* d_analytic = nullptr;
    int64_t device_bytes = 0;
    ap::RenderState* render = nullptr;
    ap::GuideDevice* guide = nullptr;
    const ap_sdf_cache* sdf_cache = nullptr;  // volumes of SDF_VOLUME objects (ap_scene_set_sdf_cache)
    const ap_texture_cache* tex_cache = nullptr;  // images of texture handles (ap_scene_set_texture_cache)
};

namespace ap {
void set_error(const std::string& msg);
int fail(int code, const std::string& msg);
int hip_fail(hipError_t e, const char* what);

// Spectral context of the current device, built on first use (ap_spectrum.hip).
struct f4;
int spectrum_get(const f4** rgb_lut, const f4** lambda_lut, const float** tables);

// Kernel launchers (ap_trace.hip).
// dyn_n (optional, device): the actual ray count, read by the kernel; n is
// then the capacity the grid is sized for.
// sdf_eps_any: the SDF marching epsilon of an any-hit launch (SDF_EPS of the
// shadow rays' depth); closest-hit rays march with tmin / 2.
hipError_t launch_trace(const DeviceScene& s, const ap_ray* rays, size_t n, ap_hit* hits,
                        ap_trace_stats* stats, bool any_hit, hipStream_t stream,
                        const uint32_t* dyn_n = nullptr, float sdf_eps_any = 0.5f * 1.0e-3f);
// Rays of several scenes in one queue: ray j traverses scenes[ray_scene[j]]
// (a device array); `shape` gives the launch shape, stack capacity (the
// largest of the scenes') and traversal mode.
hipError_t launch_trace_multi(const DeviceScene& shape, const DeviceScene* scenes, const uint16_t* ray_scene,
                              const ap_ray* rays, size_t n_total, ap_hit* hits, bool any_hit, hipStream_t stream,
                              const uint32_t* dyn_n, float sdf_eps_any = 0.5f * 1.0e-3f);

// Radiance cache entry points for the integrator (ap_nrc.hip).
struct NrcState;
NrcState* nrc_state(ap_nrc* h);
// Networks in the cache (1 for a plain cache).
uint32_t nrc_groups(const NrcState* s);
void nrc_set_bounds(NrcState* s, const float lo[3], const float hi[3]);
// Inference over min(*dyn_n, n_cap) query records (16 floats each); output
// row r is record rows[r].
int nrc_infer_dyn(NrcState* s, const float* queries, const int32_t* rows, const uint32_t* dyn_n, uint32_t n_cap,
                  float* out, hipStream_t st);
// `steps` Adam steps on batches of `batch` rows per network: network g's
// rows are drawn from the pool_n[g] samples (20 floats each) at
// pool_begin[g] of `pool` (from 0 without pool_begin); nothing changes for a
// network whose pool_n is 0. The draws follow `seed`; with frame_dev they
// follow nrc_frame_hash(frame_seed, *frame_dev) ^ kNrcTrainSalt, read on the
// device, so that a captured frame replays with the frame counter.
int nrc_train_pool(NrcState* s, const float* pool, const uint32_t* pool_begin, const uint32_t* pool_n,
                   uint32_t batch, uint32_t steps, float lr, uint64_t seed, hipStream_t st,
                   const int32_t* frame_dev = nullptr, uint64_t frame_seed = 0);
// Adds `steps` to the cache's host count of Adam steps (steps replayed from
// a captured frame; negative to take back the steps a capture counted).
void nrc_count_steps(NrcState* s, int32_t steps);
// A process-unique number of the cache and its training buffers, renewed
// when one is reallocated (a captured training step names its buffers).
uint64_t nrc_generation(const NrcState* s);
// Inference of every network of a grouped cache: region g = rows [seg[g],
// seg[g+1]) (device), of which the first count[g] are live (all without
// count), goes to network g, cut into 64-row tiles numbered from tile0[g];
// row r reads record rows[r]. n_bound: an upper bound on the tiles x 64.
int nrc_infer_groups(NrcState* s, const float* queries, const int32_t* rows, const uint32_t* seg,
                     const uint32_t* tile0, const uint32_t* count, uint32_t n_bound, float* out, hipStream_t st);
// Bounds of network g (uploaded when they change; synchronous then).
int nrc_set_group_bounds(NrcState* s, uint32_t g, const float lo[3], const float hi[3]);
}  // namespace ap

// Scene creation, device upload and the traversal entry points of the C ABI.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <string>

#include "ap_internal.h"

namespace ap {

namespace {
thread_local std::string g_last_error;

inline f3 F(const ap_float3& a) { return {a.x, a.y, a.z}; }
inline m3 M(const ap_float3x3& m) { return {F(m.x), F(m.y), F(m.z)}; }

constexpr float kPi = 3.14159265358979323846f;  // CUDART_PI_F

// SceneObject::GetArea (core/render/object.cuh:278-302).
float object_area(const ap_scene_object& o) {
    if (o.type == AP_OBJ_CUBE || o.type == AP_OBJ_SDF) {
        const float ex = fabsf(o.extent.x), ey = fabsf(o.extent.y), ez = fabsf(o.extent.z);
        return 8.0f * (ex * ey + ex * ez + ey * ez);
    }
    if (o.type == AP_OBJ_SPHERE) {
        const float r = fabsf(o.extent.x);
        return 4.0f * kPi * r * r;
    }
    if (o.type == AP_OBJ_TRIANGLE) {
        const f3 e1 = F(o.additional.v1.position) - F(o.additional.v0.position);
        const f3 e2 = F(o.additional.v2.position) - F(o.additional.v0.position);
        return 0.5f * length(cross(e1, e2));
    }
    return 0.0f;
}

// luminance (core/vector.cuh:306-309) of the emissive colour
// (SceneObject::GetEmissivePower, object.cuh:304-307).
float emissive_power(const ap_scene_object& o) {
    return dot(F(o.material.texture.base.emissive), f3{0.2126f, 0.7152f, 0.0722f});
}

template <typename T>
int upload(const std::vector<T>& host, void** dev, int64_t& bytes) {
    *dev = nullptr;
    if (host.empty()) return AP_OK;
    const size_t sz = host.size() * sizeof(T);
    hipError_t e = hipMalloc(dev, sz);
    if (e != hipSuccess) return hip_fail(e, "hipMalloc");
    e = hipMemcpy(*dev, host.data(), sz, hipMemcpyHostToDevice);
    if (e != hipSuccess) return hip_fail(e, "hipMemcpy");
    bytes += static_cast<int64_t>(sz);
    return AP_OK;
}

void free_device(ap_scene* s) {
    if (s->render) { render_state_free(s->render); s->render = nullptr; }
    if (s->guide) { guide_free(s->guide); s->guide = nullptr; }
    if (s->d_nodes) (void)hipFree(s->d_nodes);
    if (s->d_prims) (void)hipFree(s->d_prims);
    if (s->d_analytic) (void)hipFree(s->d_analytic);
    s->d_nodes = s->d_prims = s->d_analytic = nullptr;
}

}  // namespace

void set_error(const std::string& msg) { g_last_error = msg; }
int fail(int code, const std::string& msg) { g_last_error = msg; return code; }
int hip_fail(hipError_t e, const char* what) {
    g_last_error = std::string(what) + ": " + hipGetErrorString(e);
    if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) return AP_ERR_OUT_OF_MEMORY;
    if (e == hipErrorNoDevice || e == hipErrorInvalidDevice) return AP_ERR_NO_DEVICE;
    return AP_ERR_HIP;
}

// Flattens the reference-layout BVH into NodeRec/PrimRec/AnalyticRec arrays
// and uploads them.
static int build_device(ap_scene* s) {
    const HostBVH& bvh = s->bvh;
    const size_t nn = bvh.nodes.size();
    std::vector<int32_t> rec_of(nn, -1);
    int32_t nrec = 0;
    for (size_t i = 0; i < nn; ++i)
        if (bvh.nodes[i].count == 0) rec_of[i] = nrec++;
    auto link_of = [&](int32_t node) -> int32_t {
        const ap_bvh_node& c = bvh.nodes[node];
        return c.count == 0 ? rec_of[node] : ~c.first;
    };
    std::vector<NodeRec> recs(nrec);
    for (size_t i = 0; i < nn; ++i) {
        const ap_bvh_node& n = bvh.nodes[i];
        if (n.count != 0) continue;
        const ap_bvh_node& l = bvh.nodes[n.left];
        const ap_bvh_node& r = bvh.nodes[n.right];
        NodeRec& q = recs[rec_of[i]];
        q.lmin[0] = l.bound_min.x; q.lmin[1] = l.bound_min.y; q.lmin[2] = l.bound_min.z;
        q.lmax[0] = l.bound_max.x; q.lmax[1] = l.bound_max.y; q.lmax[2] = l.bound_max.z;
        q.rmin[0] = r.bound_min.x; q.rmin[1] = r.bound_min.y; q.rmin[2] = r.bound_min.z;
        q.rmax[0] = r.bound_max.x; q.rmax[1] = r.bound_max.y; q.rmax[2] = r.bound_max.z;
        q.llink = link_of(n.left);
        q.rlink = link_of(n.right);
        q.pad0 = q.pad1 = 0;
    }
    const size_t np = bvh.indices.size();
    std::vector<PrimRec> prims(np);
    std::vector<AnalyticRec> analytic;
    for (size_t slot = 0; slot < np; ++slot) {
        const int32_t obj = bvh.indices[slot];
        const ap_scene_object& o = s->objects[obj];
        PrimRec& p = prims[slot];
        std::memset(&p, 0, sizeof(p));
        p.obj = obj;
        uint32_t kind = PRIM_KIND_NEVER;
        if (o.type == AP_OBJ_TRIANGLE && o.visible) {
            // p = objectToWorld_ * v + center_ (intersection.cuh:27-29).
            const m3 o2w = M(o.object_to_world);
            const f3 c = F(o.center);
            const f3 p0 = mul(o2w, F(o.additional.v0.position)) + c;
            const f3 p1 = mul(o2w, F(o.additional.v1.position)) + c;
            const f3 p2 = mul(o2w, F(o.additional.v2.position)) + c;
            const f3 e1 = p1 - p0, e2 = p2 - p0;
            p.p0[0] = p0.x; p.p0[1] = p0.y; p.p0[2] = p0.z;
            p.e1[0] = e1.x; p.e1[1] = e1.y; p.e1[2] = e1.z;
            p.e2[0] = e2.x; p.e2[1] = e2.y; p.e2[2] = e2.z;
            kind = PRIM_KIND_TRIANGLE;
        } else if (o.type == AP_OBJ_CUBE || o.type == AP_OBJ_SPHERE || o.type == AP_OBJ_SDF) {
            AnalyticRec a;
            std::memset(&a, 0, sizeof(a));
            a.center[0] = o.center.x; a.center[1] = o.center.y; a.center[2] = o.center.z;
            a.extent[0] = o.extent.x; a.extent[1] = o.extent.y; a.extent[2] = o.extent.z;
            const float* w = &o.world_to_object.x.x;
            const float* v = &o.object_to_world.x.x;
            for (int k = 0; k < 9; ++k) { a.w2o[k] = w[k]; a.o2w[k] = v[k]; }
            a.sdf_type = -1;
            a.volume = -1;
            if (o.type == AP_OBJ_SDF) {
                a.sdf_type = o.additional.sdf_type;
                a.volume = o.additional.sdf_info.cached_volume_index;
                const ap_float4& q = o.additional.sdf_info.sdf_info0;
                a.sdf_info[0] = q.x; a.sdf_info[1] = q.y; a.sdf_info[2] = q.z; a.sdf_info[3] = q.w;
                s->dev.has_sdf = 1;
            }
            p.aux = static_cast<int32_t>(analytic.size());
            analytic.push_back(a);
            kind = o.type == AP_OBJ_CUBE ? PRIM_KIND_CUBE : (o.type == AP_OBJ_SPHERE ? PRIM_KIND_SPHERE : PRIM_KIND_SDF);
        }
        p.flags = kind << PRIM_KIND_SHIFT;
    }
    for (size_t i = 0; i < nn; ++i) {
        const ap_bvh_node& n = bvh.nodes[i];
        if (n.count > 0) prims[n.first + n.count - 1].flags |= PRIM_LAST;
    }
    s->interior_count = nrec;
    int rc;
    if ((rc = upload(recs, &s->d_nodes, s->device_bytes)) != AP_OK) return rc;
    if ((rc = upload(prims, &s->d_prims, s->device_bytes)) != AP_OK) return rc;
    if ((rc = upload(analytic, &s->d_analytic, s->device_bytes)) != AP_OK) return rc;

    DeviceScene& d = s->dev;
    d.nodes = static_cast<const NodeRec*>(s->d_nodes);
    d.prims = static_cast<const PrimRec*>(s->d_prims);
    d.analytic = static_cast<const AnalyticRec*>(s->d_analytic);
    d.root.valid = nn > 0 ? 1 : 0;
    if (nn > 0) {
        const ap_bvh_node& r = bvh.nodes[0];
        d.root.bmin[0] = r.bound_min.x; d.root.bmin[1] = r.bound_min.y; d.root.bmin[2] = r.bound_min.z;
        d.root.bmax[0] = r.bound_max.x; d.root.bmax[1] = r.bound_max.y; d.root.bmax[2] = r.bound_max.z;
        d.root.link = link_of(0);
        // Growth of the CULLED line test: 2^-12 of the scene's coordinate
        // magnitude, far above the rounding of the slab and triangle tests.
        float mag = 1.0f;
        const float c[6] = {r.bound_min.x, r.bound_min.y, r.bound_min.z, r.bound_max.x, r.bound_max.y, r.bound_max.z};
        for (float v : c) mag = fmaxf(mag, fabsf(v));
        d.line_margin = mag * 0x1p-12f;
    }
    d.mode = AP_TRAVERSE_CULLED;
    d.refill = 16;
    d.wave_rays = 512;
    // The stack never holds more than (depth of an expanded node) + 2 entries.
    d.stack_cap = std::min(128, s->max_depth + 2);
    return AP_OK;
}

}  // namespace ap

using namespace ap;

extern "C" {

const char* ap_last_error_string(void) { return g_last_error.c_str(); }

int ap_scene_create(const ap_scene_object* objects, size_t count, const ap_build_params* params,
                    ap_scene** out_scene) {
    if (!out_scene) return fail(AP_ERR_INVALID_ARG, "ap_scene_create: out_scene is NULL");
    *out_scene = nullptr;
    if (count > 0 && !objects) return fail(AP_ERR_INVALID_ARG, "ap_scene_create: objects is NULL");
    if (count > static_cast<size_t>(INT32_MAX / 2))
        return fail(AP_ERR_INVALID_ARG, "ap_scene_create: too many objects");
    for (size_t i = 0; i < count; ++i) {
        const int t = objects[i].type;
        if (t < AP_OBJ_CUBE || t > AP_OBJ_TRIANGLE)
            return fail(AP_ERR_INVALID_ARG, "ap_scene_create: unknown object type at " + std::to_string(i));
    }
    const int leaf = params ? params->leaf_size : 8;
    const int bins = params ? params->bin_count : 12;

    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return hip_fail(e, "hipGetDevice");

    ap_scene* s = new (std::nothrow) ap_scene();
    if (!s) return fail(AP_ERR_OUT_OF_MEMORY, "ap_scene_create: host allocation failed");
    s->device = dev;
    try {
        s->objects.assign(objects, objects + count);
        // Light table and index stamping (scene.cu:1056-1080).
        float total = 0.0f;
        for (size_t i = 0; i < count; ++i) {
            ap_scene_object& o = s->objects[i];
            o.object_index = static_cast<int32_t>(i);
            if (o.material.shading_model == AP_MAT_LIGHT) {
                o.light_index = static_cast<int32_t>(s->lights.size());
                ap_light_info li;
                li.object_index = static_cast<int32_t>(i);
                li.importance_weight = fmaxf(object_area(o) * emissive_power(o), 1.0e-20f);
                li.select_prob = 0.0f;
                total += li.importance_weight;
                s->lights.push_back(li);
            } else {
                o.light_index = -1;
            }
        }
        for (ap_light_info& li : s->lights) li.select_prob = li.importance_weight / total;

        std::vector<ap_primitive_proxy> proxies(count);
        for (size_t i = 0; i < count; ++i) proxies[i] = s->objects[i].proxy;
        s->bvh = build_bvh(proxies, leaf, bins);
        s->max_depth = bvh_max_depth(s->bvh);
    } catch (const std::bad_alloc&) {
        delete s;
        return fail(AP_ERR_OUT_OF_MEMORY, "ap_scene_create: host allocation failed");
    }
    const int rc = build_device(s);
    if (rc != AP_OK) {
        free_device(s);
        delete s;
        return rc;
    }
    *out_scene = s;
    return AP_OK;
}

int ap_build_bvh(const ap_scene_object* objects, size_t count, const ap_build_params* params,
                 ap_bvh_node* nodes, int32_t* indices, size_t* node_count) {
    if ((count > 0 && (!objects || !nodes || !indices)) || !node_count)
        return fail(AP_ERR_INVALID_ARG, "ap_build_bvh: NULL argument");
    if (count > static_cast<size_t>(INT32_MAX / 2)) return fail(AP_ERR_INVALID_ARG, "ap_build_bvh: too many objects");
    try {
        std::vector<ap_primitive_proxy> proxies(count);
        for (size_t i = 0; i < count; ++i) proxies[i] = objects[i].proxy;
        const HostBVH b = build_bvh(proxies, params ? params->leaf_size : 8, params ? params->bin_count : 12);
        if (!b.nodes.empty()) std::memcpy(nodes, b.nodes.data(), b.nodes.size() * sizeof(ap_bvh_node));
        if (!b.indices.empty()) std::memcpy(indices, b.indices.data(), b.indices.size() * sizeof(int32_t));
        *node_count = b.nodes.size();
    } catch (const std::bad_alloc&) {
        return fail(AP_ERR_OUT_OF_MEMORY, "ap_build_bvh: host allocation failed");
    }
    return AP_OK;
}

int ap_scene_destroy(ap_scene* scene) {
    if (!scene) return AP_OK;
    free_device(scene);
    delete scene;
    return AP_OK;
}

int ap_scene_get_info(const ap_scene* s, ap_bvh_info* out) {
    if (!s || !out) return fail(AP_ERR_INVALID_ARG, "ap_scene_get_info: NULL argument");
    out->node_count = static_cast<int64_t>(s->bvh.nodes.size());
    out->interior_count = s->interior_count;
    out->prim_count = static_cast<int64_t>(s->bvh.indices.size());
    out->max_depth = s->max_depth;
    out->stack_capacity = s->dev.stack_cap;
    out->device_bytes = s->device_bytes;
    out->light_count = static_cast<int32_t>(s->lights.size());
    out->traversal_mode = s->dev.mode;
    out->line_margin = s->dev.line_margin;
    out->pad = 0;
    return AP_OK;
}

int ap_scene_set_launch(ap_scene* scene, int32_t wave_rays, int32_t refill) {
    if (!scene) return fail(AP_ERR_INVALID_ARG, "ap_scene_set_launch: NULL scene");
    if (wave_rays < 0 || wave_rays > 65536 || wave_rays % 64 != 0)
        return fail(AP_ERR_INVALID_ARG, "ap_scene_set_launch: wave_rays must be a multiple of 64 in [64, 65536]");
    if (refill < 0 || refill > 64) return fail(AP_ERR_INVALID_ARG, "ap_scene_set_launch: refill must be in [1, 64]");
    if (wave_rays > 0) scene->dev.wave_rays = wave_rays;
    if (refill > 0) scene->dev.refill = refill;
    return AP_OK;
}

int ap_scene_get_launch(const ap_scene* scene, int32_t* wave_rays, int32_t* refill) {
    if (!scene) return fail(AP_ERR_INVALID_ARG, "ap_scene_get_launch: NULL scene");
    if (wave_rays) *wave_rays = scene->dev.wave_rays;
    if (refill) *refill = scene->dev.refill;
    return AP_OK;
}

int ap_scene_autotune(ap_scene* scene, const ap_ray* rays, size_t n, ap_hit* hits, int any_hit, void* stream,
                      int32_t* best_wave_rays, int32_t* best_refill, float* best_ms) {
    if (!scene) return fail(AP_ERR_INVALID_ARG, "ap_scene_autotune: NULL scene");
    if (n == 0 || !rays || !hits) return fail(AP_ERR_INVALID_ARG, "ap_scene_autotune: empty or NULL ray batch");
    // Steps of 64-128 rays: the best chunk differs by ray set (4096^2
    // primaries of the 1M-triangle soup: 512 rays, 1513 Mrays/s; 256: 1386).
    static const int32_t kWaveRays[] = {128, 192, 256, 320, 384, 448, 512, 640, 768, 1024};
    static const int32_t kRefill[] = {4, 8, 16};
    const hipStream_t st = static_cast<hipStream_t>(stream);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t he = hipEventCreate(&e0);
    if (he == hipSuccess) he = hipEventCreate(&e1);
    if (he != hipSuccess) {
        if (e0) (void)hipEventDestroy(e0);
        return hip_fail(he, "ap_scene_autotune: hipEventCreate");
    }
    const DeviceScene saved = scene->dev;
    float best = -1.0f;
    int32_t bw = saved.wave_rays, br = saved.refill;
    for (int32_t wr : kWaveRays) {
        for (int32_t rf : kRefill) {
            scene->dev.wave_rays = wr;
            scene->dev.refill = rf;
            // One untimed launch, then the fastest of two timed launches.
            he = launch_trace(scene->dev, rays, n, hits, nullptr, any_hit != 0, st);
            float shape_ms = -1.0f;
            for (int rep = 0; rep < 2 && he == hipSuccess; ++rep) {
                he = hipEventRecord(e0, st);
                if (he == hipSuccess) he = launch_trace(scene->dev, rays, n, hits, nullptr, any_hit != 0, st);
                if (he == hipSuccess) he = hipEventRecord(e1, st);
                if (he == hipSuccess) he = hipEventSynchronize(e1);
                float ms = 0.0f;
                if (he == hipSuccess) he = hipEventElapsedTime(&ms, e0, e1);
                if (he == hipSuccess && (shape_ms < 0.0f || ms < shape_ms)) shape_ms = ms;
            }
            if (he != hipSuccess) break;
            if (best < 0.0f || shape_ms < best) {
                best = shape_ms;
                bw = wr;
                br = rf;
            }
        }
        if (he != hipSuccess) break;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    scene->dev = saved;
    if (he != hipSuccess) return hip_fail(he, "ap_scene_autotune");
    scene->dev.wave_rays = bw;
    scene->dev.refill = br;
    if (best_wave_rays) *best_wave_rays = bw;
    if (best_refill) *best_refill = br;
    if (best_ms) *best_ms = best;
    return AP_OK;
}

int ap_scene_set_traversal_mode(ap_scene* s, int mode) {
    if (!s) return fail(AP_ERR_INVALID_ARG, "ap_scene_set_traversal_mode: NULL scene");
    if (mode != AP_TRAVERSE_REFERENCE && mode != AP_TRAVERSE_CULLED)
        return fail(AP_ERR_INVALID_ARG, "ap_scene_set_traversal_mode: unknown mode");
    s->dev.mode = mode;
    return AP_OK;
}

int ap_scene_export_bvh(const ap_scene* s, ap_bvh_node* nodes, int32_t* indices, ap_light_info* lights) {
    if (!s) return fail(AP_ERR_INVALID_ARG, "ap_scene_export_bvh: NULL scene");
    if (nodes && !s->bvh.nodes.empty())
        std::memcpy(nodes, s->bvh.nodes.data(), s->bvh.nodes.size() * sizeof(ap_bvh_node));
    if (indices && !s->bvh.indices.empty())
        std::memcpy(indices, s->bvh.indices.data(), s->bvh.indices.size() * sizeof(int32_t));
    if (lights && !s->lights.empty())
        std::memcpy(lights, s->lights.data(), s->lights.size() * sizeof(ap_light_info));
    return AP_OK;
}

static int trace_common(const ap_scene* s, const ap_ray* rays, size_t n, ap_hit* hits,
                        ap_trace_stats* stats, bool any, void* stream, const char* name) {
    if (!s) return fail(AP_ERR_INVALID_ARG, std::string(name) + ": NULL scene");
    if (n == 0) return AP_OK;
    if (!rays || !hits) return fail(AP_ERR_INVALID_ARG, std::string(name) + ": NULL rays or hits");
    if (n > (static_cast<size_t>(1) << 37))
        return fail(AP_ERR_INVALID_ARG, std::string(name) + ": too many rays");
    const hipError_t e = launch_trace(s->dev, rays, n, hits, stats, any, static_cast<hipStream_t>(stream));
    if (e != hipSuccess) return hip_fail(e, name);
    return AP_OK;
}

int ap_scene_set_sdf_cache(ap_scene* s, const ap_sdf_cache* cache) {
    if (!s) return fail(AP_ERR_INVALID_ARG, "ap_scene_set_sdf_cache: NULL scene");
    s->sdf_cache = cache;
    s->dev.volumes = sdf_cache_table(cache);
    s->dev.n_volumes = sdf_cache_capacity(cache);
    return AP_OK;
}

int ap_scene_set_texture_cache(ap_scene* s, const ap_texture_cache* cache) {
    if (!s) return fail(AP_ERR_INVALID_ARG, "ap_scene_set_texture_cache: NULL scene");
    s->tex_cache = cache;
    return AP_OK;
}

int ap_trace_closest(const ap_scene* s, const ap_ray* rays, size_t n, ap_hit* hits, void* stream) {
    return trace_common(s, rays, n, hits, nullptr, false, stream, "ap_trace_closest");
}

int ap_trace_any(const ap_scene* s, const ap_ray* rays, size_t n, ap_hit* hits, void* stream) {
    return trace_common(s, rays, n, hits, nullptr, true, stream, "ap_trace_any");
}

int ap_trace_closest_stats(const ap_scene* s, const ap_ray* rays, size_t n, ap_hit* hits,
                           ap_trace_stats* stats, void* stream) {
    if (!stats) return fail(AP_ERR_INVALID_ARG, "ap_trace_closest_stats: NULL stats");
    return trace_common(s, rays, n, hits, stats, false, stream, "ap_trace_closest_stats");
}

}  // extern "C"

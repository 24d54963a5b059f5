// Wavefront path tracer (include/aperture/ap_render.h).
//
// The reference runs one thread per pixel through CalculateRadiance
// (core/render/render.cuh:460-999), a loop of Trace -> guiding -> emission
// -> NEE -> Russian roulette. Here each loop iteration becomes kernels over
// compacted queues of live paths:
//
//   generate  camera ray, wavelength sample, path state      (scene.cu:206-230, render.cuh:466-540)
//   extend    closest-hit traversal of the path queue          (render.cuh:75-108)
//   shade     hit attributes, BSDF sample/eval, emission + MIS,
//             NEE light sample -> shadow queue, RR, next ray    (render.cuh:109-163, 541-908)
//   shadow    any-hit traversal of the shadow queue            (render.cuh:165-183, 254)
//   resolve   NEE contribution of unoccluded samples           (render.cuh:255-343, 818)
//   accumulate clamp, running means, FP64 sums                 (render.cuh:966-997, scene.cu:232-254)
//
// Every path owns a PCG32 stream seeded by (pixel, frame) and draws its
// random numbers in the reference's order, so a path takes the same
// decisions as the CPU oracle's per-pixel loop wherever their FP32 results
// agree. Queues are appended to with one atomic per wavefront (ballot +
// popcount). Several frames are traced as one batch and accumulated in frame
// order.
#include <algorithm>
#include <functional>
#include <cmath>
#include <cstring>
#include <map>
#include <tuple>
#include <vector>

#include "aperture/ap_batch.h"
#include "aperture/ap_nrc.h"
#include "aperture/ap_render.h"
#include "ap_internal.h"
#include "ap_sort.h"
#include "ap_pcg.h"
#include "ap_shade.h"
#include "ap_spectrum.h"
#include "ap_guide.h"
#include "ap_display.h"
#include "ap_sdf.h"
#include "ap_prim.h"
#include "ap_texture.h"

namespace ap {

// Per-object geometry and material index for shading.
struct alignas(16) GeoRec {
    float p0[3]; int32_t type;       // AP_OBJ_*
    float e1[3]; int32_t mat;        // MatRec index
    float e2[3]; int32_t analytic;   // AnalyticRec index (cube / sphere), else -1
    float n0[3]; int32_t light_index;
    float n1[3]; int32_t model;      // the material's shading model (groups shading lanes)
    float n2[3]; int32_t pad1;
};
static_assert(sizeof(GeoRec) == 96, "GeoRec is 96 B");

// Material inputs before spectral conversion.
struct MatIn {
    f3 albedo, emissive;
    float specular, metallic, roughness;
    int32_t model, light_index;
    float ior, sigma_a, abbe, tilt;
    int32_t albedo_tex;   // texture cache index of GBuffer.albedo_tex, or -1
};

// A path carries NL wavelengths: NL = 1 is the reference's one wavelength
// per path; NL = 4 is hero-wavelength sampling (the hero lam[0] drawn by
// SampleCIE1931, lam[j] = lam[0] + 105 j nm wrapped into [360, 780]). The
// hero's pdfs drive every directional decision; radiance and throughput are
// kept per wavelength, and the pixel estimate is
// sum_j xyz(lam[j]) L_j / sum_k p(lam[k]). Spectral bases, D65 weights and
// CMF values are recomputed from lam[] where they are needed.
template <int NL>
struct alignas(16) PathT {
    uint64_t rng_state, rng_inc;
    float lam[NL];
    float radiance[NL];
    float mask[NL];
    float psum;        // sum of the wavelength pdfs (the wavelength pdf when NL = 1)
    float last_pdf;
    float last_pos[3]; int32_t ndd;    // nonDeltaDepth
    uint32_t flags; int32_t pixel; int32_t frame;
};
static_assert(sizeof(PathT<1>) == 64, "PathT<1> is 64 B");
static_assert(sizeof(PathT<4>) == 112, "PathT<4> is 112 B");
constexpr size_t kPathBytes = sizeof(PathT<4>);

// flags: PATH_* bits (PATH_PREV_DELTA: the radiance cache's spread sum skips
// the next segment), the MediaContext depth in bits 4..15 (hits since the
// path last left a medium by a miss, saturating), a batch path's scene in
// bits 16..31.
enum : uint32_t { PATH_PREV_NEE = 1u, PATH_INSIDE = 2u, PATH_HERO_ONLY = 4u, PATH_PREV_DELTA = 8u };
constexpr uint32_t kMediaDepthShift = 4u, kMediaDepthMax = 0xfffu;

// The rest of a path's MediaContext (material.cuh:12-18), read only while
// the path is inside a medium: the vertex the ray left (Beer-Lambert
// distance) and the medium's object, whose surface alone the next ray is
// traced against (-1: every object, for a camera that starts inside).
struct alignas(16) MediaRec {
    float vertex[3];
    int32_t object;
};

template <int NL>
struct alignas(16) ShadowPay {
    int32_t path;
    uint32_t cos_ok;
    float add[NL];    // mask * NEE radiance * D65 norm per wavelength (0 when the cosines fail)
};
constexpr size_t kShadowPayBytes = sizeof(ShadowPay<4>);
constexpr float kHeroStep = 105.0f;    // (780 - 360) / 4

// Radiance-cache state of a path: a_0 and the running sqrt-spread sum of the
// spread heuristic, NRC_* flags with the recorded training-vertex count in
// bits 8.., and the training slot (-1 for a rendering-only path).
struct alignas(16) NrcPath {
    float a0, s;
    uint32_t flags;
    int32_t slot;
};
enum : uint32_t { NRC_TRAIN = 1u, NRC_UNBIASED = 2u, NRC_SUFFIX = 4u };
constexpr int kNrcV = AP_NRC_TRAIN_VERTICES;
constexpr int kNrcRec = 16;    // floats per query record
constexpr int kNrcVert = 24;   // floats per training-vertex record: query, [16..19] L_k, [20..23] M_k
constexpr int kNrcSample = 20; // floats per training sample

struct NrcView {
    NrcPath* paths;
    float* queries;      // per path: the query record of its cache vertex
    int32_t* qpath;      // compacted paths that ended into the cache
    float* verts;        // per training slot, kNrcV records of kNrcVert floats
    uint32_t* counters;  // [0] cache queries, [1] training samples
    float spread;
    uint32_t* qcount;    // a batch: per-scene cache queries, scene s's at qpath[pix_begin[s] ...]
};

// One scene of a batch (ap_batch_render): its traversal descriptor, shading
// records, camera, the first path of its pixels within a batch frame, and
// the caller's framebuffers of the current call.
struct alignas(16) BatchScene {
    const GeoRec* geo;
    const MatRec* mats;
    const LightRec* lights;
    const AnalyticRec* analytic;
    float* hdr;
    double* sums;
    int32_t n_lights;
    int32_t inside;
    uint32_t pix_begin;
    uint32_t n_pix;
    float cam[14];     // origin, dir, right, up (f3 each), tan_x, tan_y
    int32_t width, height;
    const SdfVolume* volumes;  // the scene's SDF volume table
    int32_t n_volumes;
    int32_t pad;
    const TexRec* textures;    // the scene's texture table
    const float* uvs;          // per object: the triangle's three vertex UVs
    GuideView guide;                      // guided batches: the scene's probes and index volume
    ap_path_guiding_sample* guide_slots;  // guided batches: the scene's sample slots, pixel-major
    uint32_t* guide_nvalid;               // guided batches: the scene's per-pixel valid-sample counts
};

// What the batch kernels see of a batch.
struct BatchView {
    const BatchScene* scenes;
    const uint32_t* pix_begin;   // n_scenes + 1 prefix of pixel counts
    int32_t n_scenes;
    uint16_t* ray_scene[2];      // scene of each queued extension ray
    uint16_t* shadow_scene;      // scene of each queued shadow ray
    int32_t* frame;              // frame index of the next batch frame (advanced by the graph)
};

struct RenderState {
    int device = -1;
    int tri_mode = -1;
    // scene shading data
    GeoRec* geo = nullptr;
    MatRec* mats = nullptr;
    LightRec* lights = nullptr;
    int32_t n_lights = 0;
    int32_t n_mats = 0;
    std::vector<int32_t> analytic_objs;   // cube, sphere and SDF objects (the camera-inside test)
    bool has_glass = false;               // some object's shading model is MAT_GLASS
    float* uvs = nullptr;                 // per object: six floats, the triangle's vertex UVs
    int32_t max_tex = -1;                 // largest albedo texture index the materials name
    // path buffers
    size_t capacity = 0;
    void* paths = nullptr;            // PathT<NL>, kPathBytes per entry
    MediaRec* media = nullptr;        // per path
    ap_ray* rays[2] = {nullptr, nullptr};
    int32_t* path_of[2] = {nullptr, nullptr};
    ap_hit* hits = nullptr;
    ap_ray* shadow_rays = nullptr;
    void* shadow_payload = nullptr;   // ShadowPay<NL>, kShadowPayBytes per entry
    ap_hit* shadow_hits = nullptr;
    uint32_t* counters = nullptr;     // [0],[1] path queues, [2] shadow queue
    float* eps = nullptr;             // INTERSECT_EPS(d) for d = 0 .. kEpsEntries-1
    float* target_tmp = nullptr;      // offline render target
    size_t target_cap = 0;
    // radiance-cache buffers (ap_render_nrc)
    size_t nrc_cap = 0;               // paths
    NrcPath* nrc_paths = nullptr;
    float* nrc_queries = nullptr;
    int32_t* nrc_qpath = nullptr;
    float* nrc_qrgb = nullptr;
    size_t nrc_train_cap = 0;         // training slots
    float* nrc_verts = nullptr;
    uint32_t* nrc_vmask = nullptr;
    uint32_t* nrc_offsets = nullptr;
    float* nrc_pool = nullptr;
    uint32_t* nrc_counters = nullptr;
    uint32_t nrc_last_train = 0;      // training paths of the last frame
    bool nrc_rendered = false;
    // screen G-buffer samples of the paths in flight (framebuffers with a G-buffer)
    GSample* gsamples = nullptr;
    size_t gsample_cap = 0;
    // A bounce's shadow rays are traced on `side` while the render stream
    // traces the next bounce's extension rays; the events fork and join it.
    hipStream_t side = nullptr;
    hipEvent_t ev_shaded = nullptr, ev_shadowed = nullptr;
    // Frames rendered one at a time (guiding, the radiance cache) are
    // pipelined two deep: while a frame's last bounces, accumulation, probe
    // build and cache training run, the next frame's camera rays and their
    // closest hits are computed on `lead` into the other frame set (the
    // buffers below; set 0 is the ones above), and the cache trains on
    // `lead` beside the next frame's bounces.
    size_t alt_capacity = 0;
    void* alt_paths = nullptr;
    MediaRec* alt_media = nullptr;
    ap_ray* alt_rays0 = nullptr;
    int32_t* alt_path_of0 = nullptr;
    ap_hit* alt_hits = nullptr;
    uint32_t* alt_counters = nullptr;
    hipStream_t lead = nullptr;
    hipEvent_t ev_start = nullptr, ev_tail = nullptr, ev_begun = nullptr, ev_inferred = nullptr,
               ev_trained = nullptr;
    int nrc_last_set = 0;             // the frame set (cache counters) of the last cache frame
};

constexpr int kEpsEntries = 1024;

void render_state_free(RenderState* r) {
    if (!r) return;
    void* ptrs[] = {r->geo, r->mats, r->lights, r->uvs, r->paths, r->media, r->rays[0], r->rays[1], r->path_of[0], r->path_of[1],
                    r->hits, r->shadow_rays, r->shadow_payload, r->shadow_hits, r->counters, r->eps,
                    r->target_tmp, r->nrc_paths, r->nrc_queries, r->nrc_qpath, r->nrc_qrgb, r->nrc_verts,
                    r->nrc_vmask, r->nrc_offsets, r->nrc_pool, r->nrc_counters, r->gsamples, r->alt_paths,
                    r->alt_media, r->alt_rays0, r->alt_path_of0, r->alt_hits, r->alt_counters};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    hipEvent_t evs[] = {r->ev_shaded, r->ev_shadowed, r->ev_start, r->ev_tail, r->ev_begun, r->ev_inferred,
                        r->ev_trained};
    for (hipEvent_t ev : evs)
        if (ev) (void)hipEventDestroy(ev);
    if (r->side) (void)hipStreamDestroy(r->side);
    if (r->lead) (void)hipStreamDestroy(r->lead);
    delete r;
}

namespace {

constexpr int kBlock = 256;
// Lane groups of a shade block: miss, the four shading models, past the queue.
constexpr uint32_t kShadeGroups = 6;

__device__ __forceinline__ float rand1(Pcg32& r) { return pcg32_float(r); }

// Appends the calling lanes with pred set to a queue: one atomic per wave.
__device__ __forceinline__ uint32_t wave_append(uint32_t* counter, bool pred) {
    const uint64_t mask = __ballot(pred);
    const uint32_t lane = threadIdx.x & 63u;
    if (mask == 0) return 0;
    const int leader = __ffsll(static_cast<unsigned long long>(mask)) - 1;
    uint32_t base = 0;
    if (static_cast<int>(lane) == leader) base = atomicAdd(counter, static_cast<uint32_t>(__popcll(mask)));
    base = __shfl(base, leader);
    const uint64_t below = lane == 0 ? 0ull : (mask & ((~0ull) >> (64u - lane)));
    return base + static_cast<uint32_t>(__popcll(below));
}

// wave_append onto counters[key]: one atomic per distinct key of the wave.
__device__ __forceinline__ uint32_t wave_append_keyed(uint32_t* counters, uint32_t key, bool pred) {
    const uint32_t lane = threadIdx.x & 63u;
    const uint64_t below = lane == 0 ? 0ull : ((~0ull) >> (64u - lane));
    uint64_t pending = __ballot(pred);
    uint32_t slot = 0;
    while (pending != 0ull) {
        const int leader = __ffsll(static_cast<unsigned long long>(pending)) - 1;
        const uint32_t k = static_cast<uint32_t>(__shfl(static_cast<int>(key), leader));
        const bool mine = pred && key == k;
        const uint64_t m = __ballot(mine);
        uint32_t base = 0;
        if (static_cast<int>(lane) == leader) base = atomicAdd(&counters[k], static_cast<uint32_t>(__popcll(m)));
        base = static_cast<uint32_t>(__shfl(static_cast<int>(base), leader));
        if (mine) slot = base + static_cast<uint32_t>(__popcll(m & below));
        pending &= ~m;
    }
    return slot;
}

template <int NL>
__device__ __forceinline__ f3 path_xyz(const float lam[NL], const float* L, float psum,
                                       const float* __restrict__ tables);

// A slot row of a guiding record in registers: NL = 4 moves it as one
// 16-B vector, fewer wavelengths element by element.
template <int NL>
__device__ __forceinline__ float4 pg_row(const float (&row)[4]) {
    if constexpr (NL == 4) return *reinterpret_cast<const float4*>(row);
    float4 v{0.0f, 0.0f, 0.0f, 0.0f};
    v.x = row[0];
    if constexpr (NL > 1) v.y = row[1];
    if constexpr (NL > 2) v.z = row[2];
    return v;
}

template <int NL>
__device__ __forceinline__ void pg_set_row(float (&row)[4], float4 v) {
    if constexpr (NL == 4) {
        *reinterpret_cast<float4*>(row) = v;
        return;
    }
    row[0] = v.x;
    if constexpr (NL > 1) row[1] = v.y;
    if constexpr (NL > 2) row[2] = v.z;
}

// CollectRadiance (render.cuh:353-432) on a path's guiding record. A slot
// is open while any of its wavelengths' throughputs is non-zero (with one
// wavelength, the reference's mask[i] != 0).
template <int NL>
__device__ __forceinline__ bool pg_open_row(float4 m) {
    bool open = m.x != 0.0f;
    if constexpr (NL > 1) open = open || m.y != 0.0f;
    if constexpr (NL > 2) open = open || m.z != 0.0f;
    if constexpr (NL > 3) open = open || m.w != 0.0f;
    return open;
}

template <int NL>
__device__ __forceinline__ void pg_endpoint(GuideRec& r, const float* le, int ndd) {
    bool any = false;
    for (int j = 0; j < NL; ++j) any = any || le[j] != 0.0f;
    if (!any) return;
    const int hi = ndd < PG_COLLECT ? ndd : PG_COLLECT;
    for (int i = 0; i < hi; ++i) {
        const float4 m = pg_row<NL>(r.mask[i]);
        if (pg_open_row<NL>(m)) {
            float4 smp = pg_row<NL>(r.sample[i]);
            smp.x = smp.x + m.x * le[0];
            if constexpr (NL > 1) smp.y = smp.y + m.y * le[1];
            if constexpr (NL > 2) smp.z = smp.z + m.z * le[2];
            if constexpr (NL > 3) smp.w = smp.w + m.w * le[3];
            pg_set_row<NL>(r.sample[i], smp);
            r.valid |= 1u << i;
        }
    }
}

template <int NL>
__device__ __forceinline__ void pg_advance(GuideRec& r, f3 next, const float* contrib, float rr_prob, int ndd,
                                           bool allow, f3 ray_dir, f3 pos, f3 nrm) {
    // A slot opens at most once (ndd grows past it), so its radiance starts
    // at zero here rather than when the path starts.
    const float4 ones{1.0f, 1.0f, 1.0f, 1.0f}, zeros{0.0f, 0.0f, 0.0f, 0.0f};
    if (allow && ndd < PG_COLLECT && !pg_open_row<NL>(pg_row<NL>(r.mask[ndd]))) {
        r.dir[ndd][0] = next.x; r.dir[ndd][1] = next.y; r.dir[ndd][2] = next.z;
        r.loc[ndd][0] = pos.x; r.loc[ndd][1] = pos.y; r.loc[ndd][2] = pos.z;
        const f3 v = pg_reflect(ray_dir, nrm);
        r.vdir[ndd][0] = v.x; r.vdir[ndd][1] = v.y; r.vdir[ndd][2] = v.z;
        pg_set_row<NL>(r.sample[ndd], zeros);
        pg_set_row<NL>(r.mask[ndd], ones);
    }
    const float inv_q = static_cast<float>(1.0 / static_cast<double>(rr_prob));
    const int up = ndd + (allow ? 1 : 0);
    const int hi = up < PG_COLLECT ? up : PG_COLLECT;
    for (int i = 0; i < hi; ++i) {
        float4 m = pg_row<NL>(r.mask[i]);
        if (pg_open_row<NL>(m)) {
            m.x = m.x * (contrib[0] * inv_q);
            if constexpr (NL > 1) m.y = m.y * (contrib[1] * inv_q);
            if constexpr (NL > 2) m.z = m.z * (contrib[2] * inv_q);
            if constexpr (NL > 3) m.w = m.w * (contrib[3] * inv_q);
            pg_set_row<NL>(r.mask[i], m);
        }
    }
}

// A path's guiding record at its start: no slot open, none valid.
__device__ __forceinline__ void pg_clear(GuideRec& r) {
    const float4 zeros{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int i = 0; i < PG_COLLECT; ++i) *reinterpret_cast<float4*>(r.mask[i]) = zeros;
    r.valid = 0u;
}

// Path end (render.cuh:910-957): valid collection slots to guiding samples,
// radiance sum_j xyz(l_j) sample_j / psum. Returns their number.
template <int NL>
__device__ __forceinline__ uint32_t pg_emit(const GuideRec& r, const float* lam, float psum,
                                            const float* __restrict__ tables,
                                            ap_path_guiding_sample* __restrict__ out) {
    // Valid records are packed to the front of the path's PG_COLLECT slots
    // in collection order (the build reads only that many); records are
    // stored straight to `out`, so no local array is indexed at run time.
    uint32_t w = 0;
#pragma unroll
    for (int i = 0; i < PG_COLLECT; ++i) {
        if (!((r.valid >> i) & 1u)) continue;
        float smp[NL];
        for (int j = 0; j < NL; ++j) smp[j] = r.sample[i][j];
        const f3 rx = path_xyz<NL>(lam, smp, psum, tables);
        const float strength = fminf(max3(rx), PG_MAX_BRIGHT);
        if (strength < PG_MIN_BRIGHT) continue;
        ap_path_guiding_sample o;
        __builtin_memset(&o, 0, sizeof(o));
        o.direction = ap_float3{r.dir[i][0], r.dir[i][1], r.dir[i][2]};
        const f3 c = fmax3(rx, f3{0.0f, 0.0f, 0.0f});
        o.radiance = ap_float3{c.x, c.y, c.z};
        const PgVoxel v = pg_voxel(f3{r.loc[i][0], r.loc[i][1], r.loc[i][2]},
                                   f3{r.vdir[i][0], r.vdir[i][1], r.vdir[i][2]});
        o.voxel_grid = ap_int4{v.x, v.y, v.z, v.f};
        o.valid = 1;
        out[w++] = o;
    }
    return w;
}

struct CamParams {
    f3 origin, dir, right, up;
    float tan_x, tan_y;
    int32_t width, height;
};

// Paths of a batch: frames x n_pixels, path i of a frame on pixel
// pix_offset + i * pix_stride (the identity for a rendered frame; the
// training pixels for a training-only pass of the radiance cache).
struct FrameParams {
    int32_t frame0;
    int32_t frames;
    int32_t n_pixels;
    int32_t inside;
    uint64_t seed;
    uint32_t pix_offset;
    uint32_t pix_stride;
};

__global__ void set_counters_kernel(uint32_t* c, uint32_t q0, uint32_t q1, uint32_t s) {
    if (threadIdx.x == 0) { c[0] = q0; c[1] = q1; c[2] = s; }
}

__global__ void begin_bounce_kernel(uint32_t* c, int out) {
    if (threadIdx.x == 0) { c[out] = 0; c[2] = 0; }
}

// The wavelengths of a path: the hero lam[0] and its rotations.
template <int NL>
__device__ __forceinline__ void hero_wavelengths(float hero, float hero_pdf, float lam[NL], float& psum) {
    lam[0] = hero;
    psum = hero_pdf;
    for (int j = 1; j < NL; ++j) {
        float l = hero + kHeroStep * static_cast<float>(j);
        if (l > 780.0f) l -= 420.0f;
        lam[j] = l;
        psum = psum + lambda_pdf_mix(l);
    }
}

// Per-shade spectral context of a path's wavelengths.
template <int NL>
struct Spec {
    f4 lb[NL];
    float d65[NL];
};

template <int NL>
__device__ __forceinline__ Spec<NL> spec_of(const float lam[NL], const f4* __restrict__ lam_lut,
                                            const float* __restrict__ tables) {
    Spec<NL> sp;
    for (int j = 0; j < NL; ++j) {
        sp.lb[j] = query_lambda(lam_lut, lam[j]);
        sp.d65[j] = d65_norm(tables, lam[j]);
    }
    return sp;
}

// sum_j xyz(lam[j]) L_j / psum (the pixel estimate before clamping).
template <int NL>
__device__ __forceinline__ f3 path_xyz(const float lam[NL], const float* L, float psum,
                                       const float* __restrict__ tables) {
    f3 acc = lambda_to_xyz(tables, lam[0]) * L[0];
    for (int j = 1; j < NL; ++j) acc = acc + lambda_to_xyz(tables, lam[j]) * L[j];
    return acc / psum;
}

template <int NL>
__device__ __forceinline__ float mean_mask(const float* m) {
    if (NL == 1) return m[0];
    float s = m[0];
    for (int j = 1; j < NL; ++j) s = s + m[j];
    return s * (1.0f / static_cast<float>(NL));
}

template <int NL>
__global__ __launch_bounds__(kBlock) void generate_kernel(CamParams cam, FrameParams fp, const float* __restrict__ eps,
                                                          PathT<NL>* __restrict__ paths, ap_ray* __restrict__ rays,
                                                          int32_t* __restrict__ path_of, GuideRec* __restrict__ recs,
                                                          MediaRec* __restrict__ media) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    const uint32_t total = static_cast<uint32_t>(fp.n_pixels) * static_cast<uint32_t>(fp.frames);
    if (i >= total) return;
    const int32_t f = static_cast<int32_t>(i / fp.n_pixels);
    const int32_t p = static_cast<int32_t>(fp.pix_offset + (i % fp.n_pixels) * fp.pix_stride);
    const int32_t px = p % cam.width, py = p / cam.width;
    const int32_t frame = fp.frame0 + f;
    Pcg32 rng = pcg32_path(static_cast<uint64_t>(p), static_cast<uint64_t>(frame), fp.seed);
    // RenderCamera (scene.cu:214-230).
    const float u = (static_cast<float>(px) + rand1(rng)) / static_cast<float>(cam.width);
    const float v = (static_cast<float>(py) + rand1(rng)) / static_cast<float>(cam.height);
    const float nx = u * 2.0f - 1.0f;
    const float ny = static_cast<float>((1.0 - static_cast<double>(v)) * 2.0f - 1.0f);
    f3 d = normalize(cam.dir + cam.right * (nx * cam.tan_x) + cam.up * (ny * cam.tan_y));
    // CalculateRadiance prologue (render.cuh:467-519).
    float lam, lpdf;
    {
        const float r0 = rand1(rng), r1 = rand1(rng), r2 = rand1(rng);
        sample_lambda(r0, r1, r2, lam, lpdf);
    }
    d = normalize(d);
    PathT<NL> s;
    hero_wavelengths<NL>(lam, lpdf, s.lam, s.psum);
    for (int j = 0; j < NL; ++j) {
        s.radiance[j] = 0.0f;
        s.mask[j] = 1.0f;
    }
    s.last_pdf = 0.0f;
    s.last_pos[0] = s.last_pos[1] = s.last_pos[2] = 0.0f;
    s.ndd = 0;
    s.flags = fp.inside ? PATH_INSIDE : 0u;
    // A camera inside an object starts inside a medium with no object of
    // its own and the camera as the vertex (render.cuh:535-538).
    if (fp.inside) media[i] = MediaRec{{cam.origin.x, cam.origin.y, cam.origin.z}, -1};
    s.pixel = p;
    s.frame = frame;
    s.rng_state = rng.state;
    s.rng_inc = rng.inc;
    paths[i] = s;
    if (recs) pg_clear(recs[i]);
    rays[i] = ap_ray{cam.origin.x, cam.origin.y, cam.origin.z, eps[0], d.x, d.y, d.z, __int_as_float(0x7f800000)};
    path_of[i] = static_cast<int32_t>(i);
}

struct ShadeParams {
    int32_t depth;
    int32_t max_depth;
    int32_t rr_min_depth;
    int32_t nee_max_depth;
    int32_t use_nee;
    int32_t nee_always;
    int32_t sky;
    int32_t in_q;
    int32_t out_q;
    int32_t n_lights;
    GSample* gsample = nullptr;   // depth 0: each path's first-hit G-buffer sample, or none
    const SdfVolume* volumes = nullptr;  // the scene's SDF volume table
    int32_t n_volumes = 0;
    MediaRec* media = nullptr;    // per path
    const TexRec* textures = nullptr;  // the scene's texture table
    const float* uvs = nullptr;        // per object: the triangle's three vertex UVs
};

// Hit attributes as the reference intersectors report them.
struct HitInfo {
    f3 pos, normal;
    int32_t mat, light_index;
    float bu, bv;    // a triangle's barycentrics of vertices 1 and 2
};

__device__ HitInfo hit_attributes(const GeoRec& g, const AnalyticRec* __restrict__ an, f3 o, f3 d, float t,
                                  const SdfVolume* __restrict__ vols, int32_t n_vols, float sdf_eps) {
    HitInfo h;
    h.mat = g.mat;
    h.light_index = g.light_index;
    h.bu = 0.0f;
    h.bv = 0.0f;
    if (g.type == AP_OBJ_TRIANGLE) {
        // IntersectTriangleObject (intersection.cuh:15-89): same MT, then the
        // interpolated normal turned toward the ray.
        float tt, u, v;
        (void)tri_mt(o, d, {g.p0[0], g.p0[1], g.p0[2]}, {g.e1[0], g.e1[1], g.e1[2]}, {g.e2[0], g.e2[1], g.e2[2]}, tt,
                     u, v);
        const float w = 1.0f - u - v;
        const f3 n0{g.n0[0], g.n0[1], g.n0[2]}, n1{g.n1[0], g.n1[1], g.n1[2]}, n2{g.n2[0], g.n2[1], g.n2[2]};
        const f3 n = normalize(n0 * w + n1 * u + n2 * v);
        h.pos = o + d * t;
        h.normal = dot(d, n) > 0.0f ? -n : n;
        h.bu = u;
        h.bv = v;
        return h;
    }
    const AnalyticRec& a = an[g.analytic];
    const f3 c{a.center[0], a.center[1], a.center[2]}, ext{a.extent[0], a.extent[1], a.extent[2]};
    const m3 w2o{{a.w2o[0], a.w2o[1], a.w2o[2]}, {a.w2o[3], a.w2o[4], a.w2o[5]}, {a.w2o[6], a.w2o[7], a.w2o[8]}};
    const m3 o2w{{a.o2w[0], a.o2w[1], a.o2w[2]}, {a.o2w[3], a.o2w[4], a.o2w[5]}, {a.o2w[6], a.o2w[7], a.o2w[8]}};
    const f3 lo = mul(w2o, o - c);
    const f3 ld = normalize(mul(w2o, d));
    if (g.type == AP_OBJ_SDF) {
        // IntersectSDFObject (sdf_base.cuh:115-163): the march again for the
        // local hit point, the central-difference normal there.
        SdfField F;
        if (!sdf_field(a.sdf_type, a.volume, a.sdf_info, vols, n_vols, F)) {  // not reached: traversal checks it
            h.pos = o + d * t;
            h.normal = -d;
            return h;
        }
        const f3 div = inv_safe(ld);
        const f3 t1 = (-ext - lo) * div, t2 = (ext - lo) * div;
        const f3 mn = fmin3(t1, t2), mx = fmax3(t1, t2);
        const float te = fmaxf(fmaxf(mn.x, mn.y), mn.z);
        const float tx = fminf(fminf(mx.x, mx.y), mx.z);
        float th = t;
        f3 lp = lo + ld * t;
        (void)sdf_march(F, lo, ld, te, tx, ext, sdf_eps, th, lp);
        const f3 ln = sdf_local_normal(F, lp, ext, sdf_eps);
        h.pos = mul(o2w, lp) + c;
        h.normal = normalize(mul(o2w, ln));
        return h;
    }
    if (g.type == AP_OBJ_CUBE) {
        // IntersectAABBObject (aabb.cuh:28-89).
        const f3 div = inv_safe(ld);
        const f3 t1 = (-ext - lo) * div, t2 = (ext - lo) * div;
        const f3 mn = fmin3(t1, t2), mx = fmax3(t1, t2);
        const float te = fmaxf(fmaxf(mn.x, mn.y), mn.z);
        const float tx = fminf(fminf(mx.x, mx.y), mx.z);
        const bool inside = te < 0.0f;
        const float th = inside ? tx : te;
        const f3 lp = lo + ld * th;
        const float rx = fabsf(lp.x) / ext.x, ry = fabsf(lp.y) / ext.y, rz = fabsf(lp.z) / ext.z;
        f3 ln;
        if (rx >= ry && rx >= rz) ln = f3{(lp.x > 0.0f) ? 1.0f : -1.0f, 0.0f, 0.0f};
        else if (ry >= rz) ln = f3{0.0f, (lp.y > 0.0f) ? 1.0f : -1.0f, 0.0f};
        else ln = f3{0.0f, 0.0f, (lp.z > 0.0f) ? 1.0f : -1.0f};
        if (inside) ln = -ln;
        h.pos = mul(o2w, lp) + c;
        h.normal = normalize(mul(o2w, ln));
        return h;
    }
    // IntersectSphereObject (sphere.cuh:10-70).
    const float r = max3(ext);
    const float b = dot(lo, ld);
    const float cc = dot(lo, lo) - r * r;
    const float s = sqrtf(b * b - 1.0f * cc);
    const bool inside = (-b - s) <= 0.0f;
    const f3 lp = lo + ld * t;
    const f3 ln = normalize(lp) * (inside ? -1.0f : 1.0f);
    h.pos = mul(o2w, lp) + c;
    h.normal = normalize(mul(o2w, ln));
    return h;
}

// Intersect with a solo object (intersection.cuh:176-193): the ray against
// one object alone, accepted in [tmin, inf) as SceneObjectsIntersector does.
__device__ ap_hit solo_hit(const GeoRec& g, int32_t object, const AnalyticRec* __restrict__ an,
                           const SdfVolume* __restrict__ vols, int32_t n_vols, f3 o, f3 d, float tmin) {
    float t = 0.0f;
    bool ok = false;
    if (g.type == AP_OBJ_TRIANGLE) {
        float u, v;
        ok = tri_mt(o, d, {g.p0[0], g.p0[1], g.p0[2]}, {g.e1[0], g.e1[1], g.e1[2]}, {g.e2[0], g.e2[1], g.e2[2]}, t,
                    u, v);
    } else if (g.type == AP_OBJ_CUBE) {
        ok = cube_t(o, d, an[g.analytic], t);
    } else if (g.type == AP_OBJ_SPHERE) {
        ok = sphere_t(o, d, an[g.analytic], t);
    } else if (g.type == AP_OBJ_SDF) {
        ok = sdf_t(o, d, an[g.analytic], vols, n_vols, tmin > 0.0f ? tmin * 0.5f : kSdfEps0, t);
    }
    ap_hit h;
    h.id = -1;
    h.t = __int_as_float(0x7f800000);
    if (ok && !(t < tmin || t >= __int_as_float(0x7f800000))) {
        h.id = object;
        h.t = t;
    }
    return h;
}

// NEESampleLight (nee.cuh:330-377) with the box-facet and sphere-cap samplers.
__device__ LightSample sample_light(const LightRec* __restrict__ lights, int n, f3 pos, Pcg32& rng) {
    LightSample ls{};
    ls.valid = false;
    if (n <= 0) return ls;
    int pick = n - 1;
    float pprob = 0.0f;
    {
        const float r = rand1(rng);
        float acc = 0.0f;
        for (int k = 0; k < n; ++k) {
            const float pk = lights[k].select_prob;
            acc += pk;
            if (r < acc) { pick = k; pprob = pk; break; }
        }
    }
    const LightRec& L = lights[pick];
    if (L.type == AP_OBJ_CUBE) {
        // Facing weights in facet order; a facet is a candidate when its
        // weight is positive, and the candidates are walked in facet order.
        float w[6];
        int vc = 0, first = 0;
        float vsum = 0.0f;
#pragma unroll
        for (int f = 0; f < 6; ++f) {
            const BoxFacet F = box_facet(L, f);
            w[f] = facing_weight(F.n, F.c, pos, 4.0f * F.s1 * F.s2);
            if (w[f] > 0.0f) {
                if (vc == 0) first = f;
                vsum += w[f];
                ++vc;
            }
        }
        if (vc == 0) return ls;
        int face = -1;
        float wsel = 0.0f;
        {
            const float r = rand1(rng) * vsum;
            float acc = 0.0f;
#pragma unroll
            for (int f = 0; f < 6; ++f)
                if (w[f] > 0.0f && face < 0) {
                    acc += w[f];
                    if (r <= acc) { face = f; wsel = w[f]; }
                }
        }
        if (face < 0) {  // rounding left r above the last partial sum
            face = first;
#pragma unroll
            for (int f = 0; f < 6; ++f)
                if (f == first) wsel = w[f];
        }
        const BoxFacet F = box_facet(L, face);
        const float u1 = 2.0f * rand1(rng) - 1.0f;
        const float u2 = 2.0f * rand1(rng) - 1.0f;
        ls.point = F.c + F.t * (u1 * F.s1) + F.b * (u2 * F.s2);
        ls.normal = F.n;
        ls.p_facet = 1.0f * (wsel / vsum);
        ls.p_area = 1.0f / (4.0f * F.s1 * F.s2);
        ls.valid = true;
    } else if (L.type == AP_OBJ_SPHERE) {
        const f3 sc{L.center[0], L.center[1], L.center[2]};
        const float r = fmaxf(L.extent[0], fmaxf(L.extent[1], L.extent[2]));
        const f3 pc = pos - sc;
        const float dist = length(pc);
        if (dist <= r) return ls;
        const f3 axis = pc / dist;
        const float cosC = clampf(r / dist, 0.0f, 1.0f);
        const float cap = 2.0f * kPiF * r * r * (1.0f - cosC);
        const float u = rand1(rng);
        const float vphi = rand1(rng);
        const float cosT = lerpf(cosC, 1.0f, u);
        const float sinT = sqrtf(fmaxf(0.0f, 1.0f - cosT * cosT));
        const float phi = 2.0f * kPiF * vphi;
        f3 t, b;
        tangent_basis(axis, t, b);
        const f3 nn = t * (apfm::cos(phi) * sinT) + b * (apfm::sin(phi) * sinT) + axis * cosT;
        ls.point = sc + nn * r;
        ls.normal = nn;
        ls.p_facet = 1.0f;
        ls.p_area = 1.0f / cap;
        ls.valid = true;
    } else {
        return ls;
    }
    ls.object = L.object;
    ls.mat = L.mat;
    ls.p_facet = ls.p_facet * pprob;
    return ls;
}

// Writes the query record of a cache vertex (the layout of ap_nrc.h).
__device__ __forceinline__ void nrc_record(float* __restrict__ q, f3 pos, f3 d, f3 n, const MatRec& M) {
    float4* q4 = reinterpret_cast<float4*>(q);
    q4[0] = float4{pos.x, pos.y, pos.z, -d.x};
    q4[1] = float4{-d.y, -d.z, n.x, n.y};
    q4[2] = float4{n.z, M.roughness, M.albedo_rgb[0], M.albedo_rgb[1]};
}

template <int NL, bool GUIDE, bool NRC, bool MULTI = false>
__global__ __launch_bounds__(kBlock) void shade_kernel(
    ShadeParams sp, const GeoRec* __restrict__ geo, const MatRec* __restrict__ mats,
    const LightRec* __restrict__ lights, const AnalyticRec* __restrict__ analytic, const f4* __restrict__ rgb_lut,
    const f4* __restrict__ lam_lut, const float* __restrict__ tables, const float* __restrict__ eps,
    PathT<NL>* __restrict__ paths, const ap_ray* __restrict__ rays_in, const int32_t* __restrict__ path_in,
    const ap_hit* __restrict__ hits, ap_ray* __restrict__ rays_out, int32_t* __restrict__ path_out,
    ap_ray* __restrict__ shadow_rays, ShadowPay<NL>* __restrict__ shadow_pay, uint32_t* __restrict__ counters,
    int32_t* __restrict__ primary_id, int32_t primary_frame, unsigned long long* __restrict__ ray_count, GuideView gv,
    NrcView nv, BatchView bv = BatchView{}) {
    const uint32_t n = counters[sp.in_q];
    const uint32_t first = blockIdx.x * kBlock;
    if (ray_count && first == 0 && threadIdx.x == 0) atomicAdd(&ray_count[0], static_cast<unsigned long long>(n));
    if (first >= n) return;   // the whole block
    // The block's lanes take its queue entries grouped by what they shade:
    // misses, then hits by shading model (entries past the queue last), in
    // queue order within a group, so a wave runs one branch of the shading
    // code instead of all of them. A path's result does not depend on the
    // lane that shades it; only the order of the queues it appends to does.
    __shared__ ap_hit s_hit[kBlock];
    __shared__ uint16_t s_perm[kBlock];
    __shared__ uint32_t s_count[kBlock / 64][kShadeGroups];
    {
        const uint32_t i = first + threadIdx.x;
        ap_hit h{-1, 0.0f};
        uint32_t key = kShadeGroups - 1;
        if (i < n) {
            h = hits[i];
            key = 0;
            if (h.id >= 0) {
                const GeoRec* g = geo;
                if constexpr (MULTI) g = bv.scenes[bv.ray_scene[sp.in_q][i]].geo;
                const uint32_t model = static_cast<uint32_t>(g[h.id].model);
                key = 1u + (model < 4u ? model : 3u);
            }
        }
        s_hit[threadIdx.x] = h;
        const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
        const uint64_t below = lane == 0 ? 0ull : ((~0ull) >> (64u - lane));
        uint32_t rank = 0;
#pragma unroll
        for (uint32_t k = 0; k < kShadeGroups; ++k) {
            const uint64_t m = __ballot(key == k);
            if (lane == 0) s_count[wave][k] = static_cast<uint32_t>(__popcll(m));
            if (key == k) rank = static_cast<uint32_t>(__popcll(m & below));
        }
        __syncthreads();
        uint32_t pos = rank;
        for (uint32_t k = 0; k < key; ++k)
#pragma unroll
            for (uint32_t w = 0; w < kBlock / 64; ++w) pos += s_count[w][k];
        for (uint32_t w = 0; w < wave; ++w) pos += s_count[w][key];
        s_perm[pos] = static_cast<uint16_t>(threadIdx.x);
        __syncthreads();
    }
    const uint32_t slot = s_perm[threadIdx.x];
    const uint32_t i = first + slot;
    if (i >= n) return;
    const int32_t pid = path_in[i];
    PathT<NL> S = paths[pid];
    Pcg32 rng{S.rng_state, S.rng_inc};
    const ap_ray R = rays_in[i];
    ap_hit H = s_hit[slot];
    const f3 o{R.ox, R.oy, R.oz}, d{R.dx, R.dy, R.dz};
    const Spec<NL> sc = spec_of<NL>(S.lam, lam_lut, tables);
    const f4 lb = sc.lb[0];    // the hero's basis: every sampling decision uses it
    // The path's scene: the batch's record of it, or the kernel arguments.
    uint16_t scene_id = 0;
    int32_t n_lights = sp.n_lights;
    const SdfVolume* volumes = sp.volumes;
    int32_t n_volumes = sp.n_volumes;
    const TexRec* textures = sp.textures;
    const float* uvs = sp.uvs;
    if constexpr (MULTI) {
        scene_id = static_cast<uint16_t>(S.flags >> 16);
        const BatchScene& B = bv.scenes[scene_id];
        geo = B.geo;
        mats = B.mats;
        lights = B.lights;
        analytic = B.analytic;
        n_lights = B.n_lights;
        volumes = B.volumes;
        n_volumes = B.n_volumes;
        textures = B.textures;
        uvs = B.uvs;
        if constexpr (GUIDE) {  // the scene's probes; the path records are the batch's (gv.recs)
            gv.probes = B.guide.probes;
            gv.probe_count = B.guide.probe_count;
            gv.volume = B.guide.volume;
            for (int a = 0; a < 3; ++a) {
                gv.start[a] = B.guide.start[a];
                gv.size[a] = B.guide.size[a];
            }
        }
    }
    // The MediaContext the ray was traced in.
    const bool inside = (S.flags & PATH_INSIDE) != 0u;
    const uint32_t media_depth = (S.flags >> kMediaDepthShift) & kMediaDepthMax;
    MediaRec md{{0.0f, 0.0f, 0.0f}, -1};
    if (inside) md = sp.media[pid];
    // Trace (render.cuh:89-92): inside a medium with an object of its own the
    // ray meets only that object; the traversal skipped such a ray.
    if (inside && md.object >= 0) H = solo_hit(geo[md.object], md.object, analytic, volumes, n_volumes, o, d, R.tmin);
    if (sp.depth == 0 && primary_id && S.frame == primary_frame) primary_id[S.pixel] = H.id;

    bool emit_next = false, emit_shadow = false, emit_query = false;
    ap_ray next{}, shadow{};
    ShadowPay<NL> pay;
    pay.path = pid;
    pay.cos_ok = 0u;
    for (int j = 0; j < NL; ++j) pay.add[j] = 0.0f;

    if (H.id < 0) {
        // Miss: SkyLighting and the offset draw of Trace, then the path ends.
        if (sp.depth == 0 && sp.gsample) {
            GSample g{};
            g.model = kShadingModels - 1;
            sp.gsample[pid] = g;
        }
        (void)rand1(rng);
        if (!inside) {
            float le_end[NL];
            for (int j = 0; j < NL; ++j) {
                const float le = sp.sky ? interact(rgb_lut, sc.lb[j], sky_lighting(d)) : 0.0f;
                S.radiance[j] = S.radiance[j] + S.mask[j] * le * sc.d65[j];
                le_end[j] = le * sc.d65[j];
            }
            if (GUIDE) pg_endpoint<NL>(gv.recs[pid], le_end, S.ndd);
            S.flags &= ~PATH_PREV_NEE;
        } else {
            // A miss inside a medium (bMissThinSurfaceInsideMedia_, render.cuh:109-155):
            // no emission, NEE or MIS update, weight 1. The reference's hit
            // record stays zero-initialised, so the path goes on from the
            // world origin in the same direction, outside any medium; Trace's
            // miss branch clears the delta flag, so the vertex counts.
            float p_rr = 1.0f;
            bool survived = true;
            if (S.ndd >= sp.rr_min_depth) {
                const float mm = mean_mask<NL>(S.mask);
                const float basic = clampf(mm * 0.5f + mm * 0.5f, 0.05f, 0.95f);
                p_rr = lerpf(basic, 1.0f, 0.0f);
                if (rand1(rng) > p_rr) survived = false;
            }
            if (survived) {
                float contrib[NL];
                for (int j = 0; j < NL; ++j) contrib[j] = 1.0f / fmaxf(1.0f, 1e-12f);
                if (GUIDE) pg_advance<NL>(gv.recs[pid], d, contrib, p_rr, S.ndd, false, d, f3{0.0f, 0.0f, 0.0f},
                                          f3{0.0f, 0.0f, 0.0f});
                for (int j = 0; j < NL; ++j) {
                    S.mask[j] = S.mask[j] * contrib[j];
                    S.mask[j] = S.mask[j] / p_rr;
                }
                S.flags &= ~(PATH_INSIDE | (kMediaDepthMax << kMediaDepthShift));
                S.flags |= PATH_PREV_DELTA;
                S.ndd += 1;
                if (sp.depth + 1 < sp.max_depth) {
                    emit_next = true;
                    next = ap_ray{0.0f, 0.0f, 0.0f, eps[S.ndd], d.x, d.y, d.z, __int_as_float(0x7f800000)};
                }
            }
        }
    } else {
        const GeoRec g = geo[H.id];
        // SDF_EPS of the extension ray's depth: INTERSECT_EPS(ndd) / 2.
        const HitInfo hi = hit_attributes(g, analytic, o, d, H.t, volumes, n_volumes, eps[S.ndd] * 0.5f);
        // M0: the hit's GBuffer; M: with the albedo texture applied, as
        // SampleBSDF and EvaluateBSDFWithCos apply it to their copy
        // (material.cuh:69-74, 198-202). BSDF_PDF reads the GBuffer itself.
        const MatRec M0 = mats[hi.mat];
        MatRec M = M0;
        if (M0.albedo_tex >= 0) {
            float tu = 0.0f, tv = 0.0f;   // IntersectionContext::uv_ stays (0, 0) off triangles
            if (g.type == AP_OBJ_TRIANGLE) {
                const float* t = uvs + static_cast<size_t>(H.id) * 6;
                const float w = 1.0f - hi.bu - hi.bv;
                tu = t[0] * w + t[2] * hi.bu + t[4] * hi.bv;
                tv = t[1] * w + t[3] * hi.bu + t[5] * hi.bv;
            }
            float c[4];
            tex_sample(textures[M0.albedo_tex], tu, tv, c);
            const f3 al{c[0] * c[0], c[1] * c[1], c[2] * c[2]};
            const float la = max3(al);
            M.albedo_basis = query_rgb_basis(rgb_lut, la > 1.0f ? al / la : al);
            M.albedo_scale = la;
        }
        const f3 N = hi.normal;
        f3 hp = hi.pos;
        if (sp.depth == 0 && sp.gsample) {
            // What RecordScreenGBuffer reads at the first hit (render.cuh:640-643).
            GSample g;
            g.albedo[0] = M.albedo_rgb[0]; g.albedo[1] = M.albedo_rgb[1]; g.albedo[2] = M.albedo_rgb[2];
            g.metallic = M.metallic;
            g.normal[0] = N.x; g.normal[1] = N.y; g.normal[2] = N.z;
            g.roughness = M.roughness;
            g.depth = H.t;
            g.model = M.shading_model;
            g.hit = 1;
            g.pad = 0;
            sp.gsample[pid] = g;
        }
        // SampleBSDF (material.cuh:177-330).
        f3 next_dir{0.0f, 0.0f, 0.0f};
        float pdf = 1.0f;
        bool terminate = false;
        bool pure_light = false;
        bool delta = false;
        bool inside_next = inside;   // the MediaContext after this hit
        if (M.shading_model == SM_GLASS) {
            // MAT_GLASS (material.cuh:221-272): the hero's IOR decides.
            delta = true;
            const Interface f = glass_interface(ior_at(S.lam[0], M.ior, M.abbe), inside, d, N);
            if (f.k <= 0.0f) {  // total internal reflection
                pdf = 1.0f;
                if (media_depth < static_cast<uint32_t>(kMaxDeltaDepth)) {
                    next_dir = reflect(d, f.n_in);
                } else {        // forced out
                    next_dir = d;
                    inside_next = !inside;
                }
            } else {
                const float cosT = sqrtf(f.k);
                const float Fr = fresnel_dielectric(f.cosI, cosT, f.eta);
                if (rand1(rng) < Fr) {
                    next_dir = reflect(d, f.n_in);
                    pdf = Fr;
                } else {
                    next_dir = refract_dir(d, f.n_in, f.eta);
                    pdf = 1.0f - Fr;
                    inside_next = !inside;
                }
            }
        } else if (M.shading_model == SM_DIFFUSE) {
            const float e0 = rand1(rng), e1 = rand1(rng), e2 = rand1(rng);
            const DirSample ds = cosine_hemisphere(f3{e0, e1, e2}, N);
            next_dir = ds.dir;
            pdf = ds.pdf;
        } else if (M.shading_model == SM_LIGHT) {
            terminate = true;
            pure_light = true;
        } else {  // SM_GGX
            const f3 V = -d;
            const float NoV = fmaxf(0.0f, dot(N, V));
            if (NoV <= 0.0f) {
                terminate = true;
            } else {
                const float alpha = M.roughness * M.roughness;
                const float F0 = lerpf(f0_ue(M.specular), albedo_response(M, lb), M.metallic);
                const float pSpec = clampf(F0, 0.1f, 0.9f);
                const float uu = rand1(rng);
                if (uu < pSpec) {
                    const float e0 = rand1(rng), e1 = rand1(rng), e2 = rand1(rng);
                    const DirSample ds = sample_ggx_vndf(V, N, alpha, f3{e0, e1, e2});
                    const float NoL = fmaxf(0.0f, dot(N, ds.dir));
                    if (NoL <= 0.0f) {
                        terminate = true;
                    } else {
                        next_dir = ds.dir;
                        pdf = pSpec * ds.pdf + (1.0f - pSpec) * (NoL / kPiF);
                    }
                } else {
                    const float e0 = rand1(rng), e1 = rand1(rng), e2 = rand1(rng);
                    const DirSample ds = cosine_hemisphere(f3{e0, e1, e2}, N);
                    next_dir = ds.dir;
                    pdf = pSpec * ggx_vndf_pdf(V, N, next_dir, alpha) + (1.0f - pSpec) * ds.pdf;
                }
            }
        }
        float bsdf_cos[NL];
        if (delta) {
            const float dist = inside ? length(f3{md.vertex[0], md.vertex[1], md.vertex[2]} - hp) : 0.0f;
            for (int j = 0; j < NL; ++j) bsdf_cos[j] = eval_glass(M, S.lam[j], sc.lb[j], inside, dist, d, next_dir, N);
        } else {
            for (int j = 0; j < NL; ++j) bsdf_cos[j] = eval_bsdf_cos(M, sc.lb[j], d, next_dir, N);
        }
        // Trace's offset of the shading point (render.cuh:156-161).
        {
            const float off = eps[S.ndd] * fmaxf(1.0f, length(hp));
            hp = hp + N * signf(dot(N, next_dir)) * off * (rand1(rng) + 0.5f);
        }
        // Directional guiding of the sampled direction (render.cuh:546-639).
        const ap_octahedron* probe = nullptr;
        float g_alpha = 0.0f;
        f3 g_query{1.0f, 1.0f, 1.0f};
        if (GUIDE) {
            g_alpha = pg_alpha(S.frame);
            if (NL == 1) {
                const f3 xyz0 = lambda_to_xyz(tables, S.lam[0]);
                const float xyz_a[3] = {xyz0.x, xyz0.y, xyz0.z};
                g_query = pg_query(S.mask[0], xyz_a, sp.depth == 0);
            } else {
                f3 t = lambda_to_xyz(tables, S.lam[0]) * S.mask[0];
                for (int j = 1; j < NL; ++j) t = t + lambda_to_xyz(tables, S.lam[j]) * S.mask[j];
                g_query = pg_query_xyz(t, sp.depth == 0);
            }
            if (!delta && !pure_light && !inside_next && gv.probe_count > 0) {
                probe = pg_query_probe(gv, hp, pg_reflect(d, N));
                if (probe) {
                    const float p_b = pdf;
                    const float p_p = pg_hemisphere_pdf(*probe, next_dir, N, g_query);
                    if (p_b > 1.0e-12f && p_p > 1.0e-12f) {
                        if (rand1(rng) < g_alpha) {
                            const PgSample gs = pg_sample_hemisphere(*probe, rng, N, g_query);
                            if (dot(gs.dir, N) > 1.0e-12f) {
                                const float p_bs = bsdf_pdf(M0, lb, N, d, gs.dir);
                                next_dir = gs.dir;
                                for (int j = 0; j < NL; ++j) bsdf_cos[j] = eval_bsdf_cos(M, sc.lb[j], d, gs.dir, N);
                                pdf = (1.0f - g_alpha) * fmaxf(p_bs, 1.0e-12f) + g_alpha * gs.pdf;
                            } else {
                                probe = nullptr;
                            }
                        } else {
                            pdf = (1.0f - g_alpha) * p_b + g_alpha * p_p;
                        }
                    } else {
                        probe = nullptr;
                    }
                }
            }
        }
        // Emission of a light hit, MIS-weighted against the previous NEE (render.cuh:666-710).
        float le_end[NL];
        for (int j = 0; j < NL; ++j) le_end[j] = 0.0f;
        if (hi.light_index >= 0) {
            float mis = 1.0f;
            if (sp.use_nee && sp.depth > 0 && (S.flags & PATH_PREV_NEE)) {
                const float pB = fmaxf(S.last_pdf, 0.0f);
                const float pL = nee_pdf_from_hit(lights[hi.light_index], {S.last_pos[0], S.last_pos[1], S.last_pos[2]},
                                                  hp, N);
                if (pL >= 0.0f && pB > 0.0f) mis = power_heuristic(pB, pL);
            }
            for (int j = 0; j < NL; ++j) {
                const float le = interact_basis(M.emissive_basis, M.emissive_scale, sc.lb[j]);
                S.radiance[j] = S.radiance[j] + S.mask[j] * le * sc.d65[j] * mis;
            }
            if (GUIDE) {
                const f3 e{M.emissive_rgb[0], M.emissive_rgb[1], M.emissive_rgb[2]};
                for (int j = 0; j < NL; ++j) le_end[j] = interact(rgb_lut, sc.lb[j], e * mis) * sc.d65[j];
            }
        } else if (GUIDE && !delta &&
                   !(M.emissive_rgb[0] == 0.0f && M.emissive_rgb[1] == 0.0f && M.emissive_rgb[2] == 0.0f)) {
            for (int j = 0; j < NL; ++j)
                le_end[j] = interact_basis(M.emissive_basis, M.emissive_scale, sc.lb[j]) * sc.d65[j];
        }
        if (GUIDE) pg_endpoint<NL>(gv.recs[pid], le_end, S.ndd);
        if (NRC) {
            // Spread heuristic and cache termination (ap_nrc.h).
            NrcPath P = nv.paths[pid];
            const float cos_i = fabsf(dot(d, N));
            const float dist2 = H.t * H.t;
            if (sp.depth == 0) {
                P.a0 = dist2 / (12.566370614359172f * cos_i);
                P.s = 0.0f;
            } else if (!(S.flags & PATH_PREV_DELTA)) {
                // A segment leaving a delta vertex (glass) or a miss inside a
                // medium adds nothing: its pdf is a Dirac delta.
                P.s = P.s + sqrtf(dist2 / (fmaxf(S.last_pdf, 1e-12f) * cos_i));
            }
            // Glass vertices neither end into the cache nor train it.
            const bool scatter = !pure_light && !delta &&
                                 !(M.shading_model == SM_GGX && !(fmaxf(0.0f, dot(N, -d)) > 0.0f));
            bool cache_end = false;
            if (scatter && !inside_next && P.s * P.s > nv.spread * P.a0) {
                if (!(P.flags & NRC_TRAIN)) {
                    cache_end = true;
                } else if (!(P.flags & NRC_UNBIASED)) {
                    if (P.flags & NRC_SUFFIX) {
                        cache_end = true;
                    } else {
                        P.flags |= NRC_SUFFIX;
                        P.s = 0.0f;
                    }
                }
            }
            if (cache_end) {
                float* q = nv.queries + static_cast<size_t>(pid) * kNrcRec;
                nrc_record(q, hi.pos, d, N, M);
                reinterpret_cast<float4*>(q)[3] = float4{M.albedo_rgb[2], 0.0f, 0.0f, 0.0f};
                emit_query = true;
                terminate = true;
            } else if ((P.flags & NRC_TRAIN) && scatter && (P.flags >> 8) < static_cast<uint32_t>(kNrcV)) {
                float* v = nv.verts + (static_cast<size_t>(P.slot) * kNrcV + (P.flags >> 8)) * kNrcVert;
                nrc_record(v, hi.pos, d, N, M);
                reinterpret_cast<float4*>(v)[3] = float4{M.albedo_rgb[2], 0.0f, 0.0f, 0.0f};
                for (int j = 0; j < 4; ++j) {
                    v[16 + j] = j < NL ? S.radiance[j < NL ? j : 0] : 0.0f;
                    v[20 + j] = j < NL ? S.mask[j < NL ? j : 0] : 0.0f;
                }
                P.flags += 1u << 8;
            }
            nv.paths[pid] = P;
        }
        if (terminate) {
            S.flags &= ~PATH_PREV_NEE;
        } else {
            // NEE (render.cuh:766-827, 206-343).
            bool nee_valid = false;
            bool nee_deferred = false;
            const float mm = mean_mask<NL>(S.mask);
            const float rate = sp.nee_always ? 1.0f : sqrtf(saturatef(mm));
            if (sp.use_nee && S.ndd < sp.nee_max_depth && !delta && !inside_next && rand1(rng) < rate) {
                const LightSample ls = sample_light(lights, n_lights, hp, rng);
                if (ls.valid) {
                    const float pat = fmaxf(1e-20f, ls.p_area * ls.p_facet);
                    const f3 lu = ls.point - hp;
                    const float d2 = fmaxf(dot(lu, lu), 1e-12f);
                    const f3 L = lu * (1.0f / sqrtf(d2));
                    const float cosX = fmaxf(0.0f, dot(N, L));
                    const float cosY = fmaxf(0.0f, dot(ls.normal, -L));
                    const bool cos_ok = !(cosX <= 0.0f || cosY <= 0.0f);
                    float add[NL];
                    for (int j = 0; j < NL; ++j) add[j] = 0.0f;
                    if (cos_ok) {
                        const float G = cosY / d2;
                        const MatRec& LM = mats[ls.mat];
                        const float pLo = area_to_solid_angle_pdf(hp, ls.point, ls.normal, ls.p_facet * ls.p_area);
                        float pBs = fmaxf(bsdf_pdf(M0, lb, N, d, L), 0.0f);
                        if (GUIDE && probe) {
                            const float p_pr = pg_hemisphere_pdf(*probe, L, N, g_query);
                            pBs = (1.0f - g_alpha) * pBs + g_alpha * p_pr;
                        }
                        const float w = power_heuristic(pLo, pBs);
                        for (int j = 0; j < NL; ++j) {
                            const float bc = eval_bsdf_cos(M, sc.lb[j], d, L, N);
                            const float resp = interact_basis(LM.emissive_basis, LM.emissive_scale, sc.lb[j]);
                            const float contribution = bc * resp * (G / pat);
                            const float rad = contribution * w;
                            add[j] = S.mask[j] * rad * sc.d65[j];
                        }
                    }
                    // AnyHitSegmentBVH (intersection.cuh:372-397).
                    const f3 seg = ls.point - hp;
                    const float dist = length(seg);
                    bool traced = false;
                    if (dist > 0.0f) {
                        const f3 sd = seg / dist;
                        const float seps = eps[sp.depth] * fmaxf(1.0f, fmaxf(length(hp), length(ls.point)));
                        const float tmax = fmaxf(0.0f, dist - seps);
                        if (tmax > 0.0f) {
                            shadow = ap_ray{hp.x, hp.y, hp.z, 0.0f, sd.x, sd.y, sd.z, tmax};
                            traced = true;
                        }
                    }
                    if (traced) {
                        emit_shadow = true;
                        nee_deferred = true;
                        for (int j = 0; j < NL; ++j) pay.add[j] = add[j];
                        pay.cos_ok = cos_ok ? 1u : 0u;
                    } else {
                        nee_valid = cos_ok;
                        if (cos_ok)
                            for (int j = 0; j < NL; ++j) S.radiance[j] = S.radiance[j] + add[j];
                    }
                }
            }
            if (!nee_deferred) {
                if (nee_valid) S.flags |= PATH_PREV_NEE;
                else S.flags &= ~PATH_PREV_NEE;
            }
            S.last_pdf = pdf;
            S.last_pos[0] = hp.x; S.last_pos[1] = hp.y; S.last_pos[2] = hp.z;
            float contrib[NL];
            for (int j = 0; j < NL; ++j) contrib[j] = bsdf_cos[j] / fmaxf(pdf, 1e-12f);
            // Hero wavelengths at a dispersive interface: the secondaries would
            // refract elsewhere, so they end here and the hero goes on alone
            // with weight psum / p(hero), an unbiased one-wavelength estimate
            // under the pixel's division by psum.
            if (NL > 1 && delta && !(S.flags & PATH_HERO_ONLY) && abbe_slope(M.ior, M.abbe) != 0.0f) {
                contrib[0] = contrib[0] * (S.psum / lambda_pdf_mix(S.lam[0]));
                for (int j = 1; j < NL; ++j) contrib[j] = 0.0f;
                S.flags |= PATH_HERO_ONLY;
            }
            // Russian roulette (render.cuh:838-859).
            float p_rr = 1.0f;
            bool survived = true;
            if (S.ndd >= sp.rr_min_depth) {
                const float basic = clampf(mm * 0.5f + mm * 0.5f, 0.05f, 0.95f);
                p_rr = lerpf(basic, 1.0f, inside_next ? 0.99f : 0.0f);
                if (rand1(rng) > p_rr) survived = false;
            }
            if (survived) {
                if (GUIDE) pg_advance<NL>(gv.recs[pid], next_dir, contrib, p_rr, S.ndd, !delta && !inside_next, d, hp, N);
                for (int j = 0; j < NL; ++j) {
                    S.mask[j] = S.mask[j] * contrib[j];
                    S.mask[j] = S.mask[j] / p_rr;
                }
                if (!delta) S.ndd += 1;
                S.flags = delta ? (S.flags | PATH_PREV_DELTA) : (S.flags & ~PATH_PREV_DELTA);
                // EvaluateMaterial's MediaContext (material.cuh:410-411): this
                // object, one more hit, the hit point before the offset.
                const uint32_t depth_next = media_depth < kMediaDepthMax ? media_depth + 1u : kMediaDepthMax;
                S.flags = (S.flags & ~(PATH_INSIDE | (kMediaDepthMax << kMediaDepthShift))) |
                          (inside_next ? PATH_INSIDE : 0u) | (depth_next << kMediaDepthShift);
                if (sp.depth + 1 < sp.max_depth) {
                    emit_next = true;
                    // A ray inside a medium meets only the medium's object: a
                    // NaN tmax makes the traversal reject the root at once,
                    // and the next shade intersects the object itself.
                    next = ap_ray{hp.x, hp.y, hp.z, eps[S.ndd], next_dir.x, next_dir.y, next_dir.z,
                                  inside_next ? __int_as_float(0x7fc00000) : __int_as_float(0x7f800000)};
                    if (inside_next) sp.media[pid] = MediaRec{{hi.pos.x, hi.pos.y, hi.pos.z}, H.id};
                }
            }
        }
    }
    S.rng_state = rng.state;
    S.rng_inc = rng.inc;
    paths[pid] = S;
    const uint32_t qi = wave_append(&counters[sp.out_q], emit_next);
    if (emit_next) {
        rays_out[qi] = next;
        path_out[qi] = pid;
        if constexpr (MULTI) bv.ray_scene[sp.out_q][qi] = scene_id;
    }
    const uint32_t si = wave_append(&counters[2], emit_shadow);
    if (emit_shadow) {
        shadow_rays[si] = shadow;
        shadow_pay[si] = pay;
        if constexpr (MULTI) bv.shadow_scene[si] = scene_id;
    }
    if (NRC) {
        if constexpr (MULTI) {
            const uint32_t qi = wave_append_keyed(nv.qcount, scene_id, emit_query);
            if (emit_query) nv.qpath[bv.pix_begin[scene_id] + qi] = pid;
        } else {
            const uint32_t qi = wave_append(&nv.counters[0], emit_query);
            if (emit_query) nv.qpath[qi] = pid;
        }
    }
}

// Training paths of a frame (the pixels offset + k * stride, k < n_train)
// and the reset of the cache counters.
// With all_train (a training-only pass), path i is training slot i.
__global__ __launch_bounds__(kBlock) void nrc_init_kernel(int32_t n_pix, uint32_t stride, uint32_t offset,
                                                          uint32_t n_train, uint32_t unbiased_every, uint32_t all_train,
                                                          NrcPath* __restrict__ paths, uint32_t* __restrict__ counters) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i == 0) { counters[0] = 0; counters[1] = 0; }
    if (i >= static_cast<uint32_t>(n_pix)) return;
    NrcPath P{0.0f, 0.0f, 0u, -1};
    if (all_train || (i >= offset && (i - offset) % stride == 0u)) {
        const uint32_t t = all_train ? i : (i - offset) / stride;
        if (t < n_train) {
            P.flags = NRC_TRAIN | ((unbiased_every != 0u && t % unbiased_every == 0u) ? NRC_UNBIASED : 0u);
            P.slot = static_cast<int32_t>(t);
        }
    }
    paths[i] = P;
}

// Radiance of a cache vertex: throughput x the prediction c (clamped at 0)
// at the path's wavelengths.
// With a guiding record, the prediction is also the path's endpoint radiance
// for CollectRadiance: it is what the traced tail would have brought to the
// open collection slots, weighted by their throughputs as an emitter hit is.
template <int NL>
__device__ __forceinline__ void nrc_apply_one(PathT<NL>& S, f3 c, const f4* __restrict__ rgb_lut,
                                              const f4* __restrict__ lam_lut, const float* __restrict__ tables,
                                              GuideRec* guide = nullptr) {
    c = fmax3(c, f3{0.0f, 0.0f, 0.0f});
    // interact() with the RGB basis looked up once for all wavelengths.
    const float lum = max3(c);
    const f4 cb = query_rgb_basis(rgb_lut, lum > 1.0f ? c / lum : c);
    float le_end[NL];
    for (int j = 0; j < NL; ++j) {
        const f4 lb = query_lambda(lam_lut, S.lam[j]);
        const float le = lum > 1.0f ? spectral_interact(cb, lb) * lum : spectral_interact(cb, lb);
        const float d65 = d65_norm(tables, S.lam[j]);
        S.radiance[j] = S.radiance[j] + S.mask[j] * le * d65;
        le_end[j] = le * d65;
    }
    if (guide) pg_endpoint<NL>(*guide, le_end, S.ndd);
}

template <int NL>
__global__ __launch_bounds__(kBlock) void nrc_apply_kernel(const uint32_t* __restrict__ counters,
                                                           const int32_t* __restrict__ qpath,
                                                           const float* __restrict__ rgb, const f4* __restrict__ rgb_lut,
                                                           const f4* __restrict__ lam_lut,
                                                           const float* __restrict__ tables,
                                                           PathT<NL>* __restrict__ paths, GuideRec* __restrict__ recs) {
    const uint32_t q = blockIdx.x * kBlock + threadIdx.x;
    if (q >= counters[0]) return;
    const int32_t pid = qpath[q];
    nrc_apply_one<NL>(paths[pid], f3{rgb[3 * q], rgb[3 * q + 1], rgb[3 * q + 2]}, rgb_lut, lam_lut, tables,
                      recs ? recs + pid : nullptr);
}

__device__ __forceinline__ bool finite3(f3 v) {
    return !(isnan(v.x) || isnan(v.y) || isnan(v.z) || isinf(v.x) || isinf(v.y) || isinf(v.z));
}

// Targets of the training vertices of slot t (path pid), XYZ2SRGB(sum_j
// xyz_j (L_end,j - L_k,j) / M_k,j / psum), written over [13..15]; returns
// the mask of the finite ones. With several wavelengths a wavelength whose
// throughput at the vertex is zero adds nothing.
template <int NL>
__device__ __forceinline__ uint32_t nrc_slot_targets(uint32_t t, uint32_t pid, const NrcPath* __restrict__ npaths,
                                                     const PathT<NL>* __restrict__ paths,
                                                     const float* __restrict__ tables, float* __restrict__ verts) {
    const uint32_t nvx = min(npaths[pid].flags >> 8, static_cast<uint32_t>(kNrcV));
    const PathT<NL>& S = paths[pid];
    uint32_t m = 0;
    for (uint32_t k = 0; k < nvx; ++k) {
        float* v = verts + (static_cast<size_t>(t) * kNrcV + k) * kNrcVert;
        float L[NL];
        for (int j = 0; j < NL; ++j) {
            const float mk = v[20 + j];
            L[j] = (NL > 1 && !(mk > 0.0f)) ? 0.0f : (S.radiance[j] - v[16 + j]) / mk;
        }
        const f3 rgb = xyz_to_srgb(path_xyz<NL>(S.lam, L, S.psum, tables));
        if (finite3(rgb)) {
            v[13] = rgb.x; v[14] = rgb.y; v[15] = rgb.z;
            m |= 1u << k;
        }
    }
    return m;
}

// Targets of every training slot (slot t: pixel offset + t * stride, or
// path t with pid_direct); vmask marks the finite ones.
template <int NL>
__global__ __launch_bounds__(kBlock) void nrc_targets_kernel(uint32_t n_train, uint32_t stride, uint32_t offset,
                                                             uint32_t pid_direct, const NrcPath* __restrict__ npaths,
                                                             const PathT<NL>* __restrict__ paths,
                                                             const float* __restrict__ tables,
                                                             float* __restrict__ verts, uint32_t* __restrict__ vmask) {
    const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= n_train) return;
    const uint32_t pid = pid_direct ? t : offset + t * stride;
    vmask[t] = nrc_slot_targets<NL>(t, pid, npaths, paths, tables, verts);
}

// Per-slot sample counts (the set bits of each slot's vertex mask); their
// exclusive prefix sum (the project's device-wide scan, in slot order, so
// the sample pool is the same on every run) places each slot's samples.
__global__ void nrc_count_kernel(uint32_t n_train, const uint32_t* __restrict__ vmask,
                                 uint32_t* __restrict__ counts) {
    const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    if (t < n_train) counts[t] = static_cast<uint32_t>(__popc(vmask[t]));
}

// The offsets array of `cap` training slots is followed by the scan's workspace.
size_t nrc_offsets_words(size_t cap) { return ((cap + 63) & ~size_t{63}) + scan_workspace_bytes(cap) / 4 + 1; }
void* nrc_scan_work(uint32_t* offsets, size_t cap) { return offsets + ((cap + 63) & ~size_t{63}); }

// offsets[t] = samples of slots before t; the total to counters[1].
hipError_t nrc_scan(uint32_t n_train, const uint32_t* vmask, uint32_t* offsets, size_t cap, uint32_t* counters,
                    hipStream_t stream) {
    hipLaunchKernelGGL(nrc_count_kernel, dim3((std::max<uint32_t>(n_train, 1u) + kBlock - 1) / kBlock), dim3(kBlock),
                       0, stream, n_train, vmask, offsets);
    return scan_exclusive_u32(offsets, offsets, n_train, &counters[1], nrc_scan_work(offsets, cap), stream);
}

// Training samples in slot, vertex order: query (13 floats), 3 zeros, target
// RGB, pixel * 8 + vertex as uint32 bits.
// With slot_pix, slot t's pixel is slot_pix[t] (a batch: the path index).
__global__ __launch_bounds__(kBlock) void nrc_compact_kernel(uint32_t n_train, uint32_t stride, uint32_t offset,
                                                             const uint32_t* __restrict__ vmask,
                                                             const uint32_t* __restrict__ offsets,
                                                             const float* __restrict__ verts, float* __restrict__ pool,
                                                             const uint32_t* __restrict__ slot_pix = nullptr) {
    const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= n_train) return;
    const uint32_t m = vmask[t];
    const uint32_t pix = slot_pix ? slot_pix[t] : offset + t * stride;
    uint32_t j = offsets[t];
    for (int k = 0; k < kNrcV; ++k) {
        if (!((m >> k) & 1u)) continue;
        const float* v = verts + (static_cast<size_t>(t) * kNrcV + k) * kNrcVert;
        float* o = pool + static_cast<size_t>(j) * kNrcSample;
        for (int e = 0; e < 13; ++e) o[e] = v[e];
        o[13] = 0.0f; o[14] = 0.0f; o[15] = 0.0f;
        o[16] = v[13]; o[17] = v[14]; o[18] = v[15];
        reinterpret_cast<uint32_t*>(o)[19] = pix * 8u + static_cast<uint32_t>(k);
        ++j;
    }
}

template <int NL>
__global__ __launch_bounds__(kBlock) void resolve_kernel(const uint32_t* __restrict__ counters,
                                                         const ShadowPay<NL>* __restrict__ pay,
                                                         const ap_hit* __restrict__ shits,
                                                         PathT<NL>* __restrict__ paths,
                                                         unsigned long long* __restrict__ ray_count) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (ray_count && i == 0) atomicAdd(&ray_count[1], static_cast<unsigned long long>(counters[2]));
    if (i >= counters[2]) return;
    const ShadowPay<NL> p = pay[i];
    const bool occluded = shits[i].id >= 0;
    PathT<NL>& S = paths[p.path];
    const bool cos_ok = p.cos_ok != 0u;
    if (!occluded && cos_ok)
        for (int j = 0; j < NL; ++j) S.radiance[j] = S.radiance[j] + p.add[j];
    if (occluded || cos_ok) S.flags |= PATH_PREV_NEE;
    else S.flags &= ~PATH_PREV_NEE;
}

__device__ __forceinline__ float fix_nan_inf(float x) {
    if (isnan(x)) return 0.0f;
    if (isinf(x)) return 0.0f;
    return x;
}

template <int NL>
__global__ __launch_bounds__(kBlock) void accumulate_kernel(FrameParams fp, const PathT<NL>* __restrict__ paths,
                                                            const float* __restrict__ tables,
                                                            float* __restrict__ hdr, float* __restrict__ odd,
                                                            float* __restrict__ even, double* __restrict__ sums,
                                                            const GuideRec* __restrict__ recs,
                                                            ap_path_guiding_sample* __restrict__ slots,
                                                            uint32_t* __restrict__ nvalid,
                                                            const GSample* __restrict__ gsamples,
                                                            ap_screen_gbuffer* __restrict__ gbuffer,
                                                            ap_screen_statistics* __restrict__ stats) {
    const int32_t p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= fp.n_pixels) return;
    ap_screen_gbuffer gb;
    ap_screen_statistics sb;
    if (gbuffer) gb = gbuffer[p];
    if (stats) sb = stats[p];
    f3 acc{hdr[3 * p], hdr[3 * p + 1], hdr[3 * p + 2]};
    f3 ao{0, 0, 0}, ae{0, 0, 0};
    if (odd) ao = f3{odd[3 * p], odd[3 * p + 1], odd[3 * p + 2]};
    if (even) ae = f3{even[3 * p], even[3 * p + 1], even[3 * p + 2]};
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    for (int32_t f = 0; f < fp.frames; ++f) {
        const PathT<NL>& S = paths[static_cast<size_t>(f) * fp.n_pixels + p];
        if (slots) {  // frame-major slots of the batch's frames (guided groups)
            const size_t i = static_cast<size_t>(f) * fp.n_pixels + p;
            nvalid[i] = pg_emit<NL>(recs[i], S.lam, S.psum, tables, slots + i * PG_COLLECT);
        }
        float rad[NL];
        for (int j = 0; j < NL; ++j) rad[j] = fix_nan_inf(S.radiance[j]);
        const f3 pre = path_xyz<NL>(S.lam, rad, S.psum, tables);
        f3 res{clampf(pre.x, 0.0f, 1000.0f), clampf(pre.y, 0.0f, 1000.0f), clampf(pre.z, 0.0f, 1000.0f)};
        res = fmax3(f3{0.0f, 0.0f, 0.0f}, res);
        const int32_t frame = S.frame;
        const float w = 1.0f / static_cast<float>(1 + frame);
        acc = lerp3(acc, res, w);
        if (odd && even) {
            const float wh = 1.0f / static_cast<float>(1 + (frame >> 1));
            if (frame % 2 == 1) ao = lerp3(ao, res, wh);
            else ae = lerp3(ae, res, wh);
        }
        s0 += static_cast<double>(res.x);
        s1 += static_cast<double>(res.y);
        s2 += static_cast<double>(res.z);
        // Frame order per pixel, as the reference's frames would record them.
        if (gbuffer) record_gbuffer(gb, gsamples[static_cast<size_t>(f) * fp.n_pixels + p]);
        if (stats) {
            const float x[3] = {res.x, res.y, res.z};
            record_statistics(sb, x);
        }
    }
    if (gbuffer) gbuffer[p] = gb;
    if (stats) stats[p] = sb;
    hdr[3 * p] = acc.x; hdr[3 * p + 1] = acc.y; hdr[3 * p + 2] = acc.z;
    if (odd && even) {
        odd[3 * p] = ao.x; odd[3 * p + 1] = ao.y; odd[3 * p + 2] = ao.z;
        even[3 * p] = ae.x; even[3 * p + 1] = ae.y; even[3 * p + 2] = ae.z;
    }
    if (sums) {
        sums[3 * p] += s0;
        sums[3 * p + 1] += s1;
        sums[3 * p + 2] += s2;
    }
}

// Scene of batch path i: the last scene whose first pixel is <= i.
__device__ __forceinline__ int32_t batch_scene_of(const uint32_t* __restrict__ pix_begin, int32_t n, uint32_t i) {
    int32_t lo = 0, hi = n - 1;
    while (lo < hi) {
        const int32_t mid = (lo + hi + 1) >> 1;
        if (pix_begin[mid] <= i) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// generate_kernel over the pixels of every scene of a batch: path i is pixel
// i - pix_begin of scene s, with that scene's camera; the frame index comes
// from the batch's device counter (a captured graph replays the launch).
template <int NL>
__global__ __launch_bounds__(kBlock) void generate_batch_kernel(BatchView bv, uint32_t total, uint64_t seed,
                                                                const float* __restrict__ eps,
                                                                PathT<NL>* __restrict__ paths, ap_ray* __restrict__ rays,
                                                                int32_t* __restrict__ path_of,
                                                                MediaRec* __restrict__ media,
                                                                GuideRec* __restrict__ recs = nullptr) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= total) return;
    const int32_t sc = batch_scene_of(bv.pix_begin, bv.n_scenes, i);
    const BatchScene& B = bv.scenes[sc];
    const int32_t p = static_cast<int32_t>(i - B.pix_begin);
    const int32_t px = p % B.width, py = p / B.width;
    const int32_t frame = *bv.frame;
    const f3 origin{B.cam[0], B.cam[1], B.cam[2]}, cdir{B.cam[3], B.cam[4], B.cam[5]};
    const f3 right{B.cam[6], B.cam[7], B.cam[8]}, up{B.cam[9], B.cam[10], B.cam[11]};
    Pcg32 rng = pcg32_path(static_cast<uint64_t>(p), static_cast<uint64_t>(frame), seed);
    // RenderCamera (scene.cu:214-230), as generate_kernel.
    const float u = (static_cast<float>(px) + rand1(rng)) / static_cast<float>(B.width);
    const float v = (static_cast<float>(py) + rand1(rng)) / static_cast<float>(B.height);
    const float nx = u * 2.0f - 1.0f;
    const float ny = static_cast<float>((1.0 - static_cast<double>(v)) * 2.0f - 1.0f);
    f3 d = normalize(cdir + right * (nx * B.cam[12]) + up * (ny * B.cam[13]));
    float lam, lpdf;
    {
        const float r0 = rand1(rng), r1 = rand1(rng), r2 = rand1(rng);
        sample_lambda(r0, r1, r2, lam, lpdf);
    }
    d = normalize(d);
    PathT<NL> st;
    hero_wavelengths<NL>(lam, lpdf, st.lam, st.psum);
    for (int j = 0; j < NL; ++j) {
        st.radiance[j] = 0.0f;
        st.mask[j] = 1.0f;
    }
    st.last_pdf = 0.0f;
    st.last_pos[0] = st.last_pos[1] = st.last_pos[2] = 0.0f;
    st.ndd = 0;
    st.flags = (B.inside ? PATH_INSIDE : 0u) | (static_cast<uint32_t>(sc) << 16);
    if (B.inside) media[i] = MediaRec{{origin.x, origin.y, origin.z}, -1};
    st.pixel = p;
    st.frame = frame;
    st.rng_state = rng.state;
    st.rng_inc = rng.inc;
    paths[i] = st;
    if (recs) pg_clear(recs[i]);
    rays[i] = ap_ray{origin.x, origin.y, origin.z, eps[0], d.x, d.y, d.z, __int_as_float(0x7f800000)};
    path_of[i] = static_cast<int32_t>(i);
    bv.ray_scene[0][i] = static_cast<uint16_t>(sc);
}

// accumulate_kernel for one frame of a batch, into each scene's framebuffers
// (skipped for a scene whose hdr is NULL); with the path records of a guided
// batch, each path's guiding samples go to its scene's slots.
template <int NL>
__global__ __launch_bounds__(kBlock) void accumulate_batch_kernel(BatchView bv, uint32_t total,
                                                                  const PathT<NL>* __restrict__ paths,
                                                                  const float* __restrict__ tables,
                                                                  const GuideRec* __restrict__ recs = nullptr) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= total) return;
    const int32_t sc = batch_scene_of(bv.pix_begin, bv.n_scenes, i);
    const BatchScene& B = bv.scenes[sc];
    const uint32_t p = i - B.pix_begin;
    const PathT<NL>& S = paths[i];
    if (recs && B.guide_slots)
        B.guide_nvalid[p] =
            pg_emit<NL>(recs[i], S.lam, S.psum, tables, B.guide_slots + static_cast<size_t>(p) * PG_COLLECT);
    if (!B.hdr) return;
    float rad[NL];
    for (int j = 0; j < NL; ++j) rad[j] = fix_nan_inf(S.radiance[j]);
    const f3 pre = path_xyz<NL>(S.lam, rad, S.psum, tables);
    f3 res{clampf(pre.x, 0.0f, 1000.0f), clampf(pre.y, 0.0f, 1000.0f), clampf(pre.z, 0.0f, 1000.0f)};
    res = fmax3(f3{0.0f, 0.0f, 0.0f}, res);
    const float w = 1.0f / static_cast<float>(1 + S.frame);
    float* h = B.hdr + 3 * static_cast<size_t>(p);
    const f3 acc = lerp3(f3{h[0], h[1], h[2]}, res, w);
    h[0] = acc.x; h[1] = acc.y; h[2] = acc.z;
    if (B.sums) {
        double* q = B.sums + 3 * static_cast<size_t>(p);
        q[0] += static_cast<double>(res.x);
        q[1] += static_cast<double>(res.y);
        q[2] += static_cast<double>(res.z);
    }
}

__global__ void advance_frame_kernel(int32_t* frame) {
    if (threadIdx.x == 0) *frame += 1;
}

// Sets a batch's frame counter in stream order (a host copy on another
// stream would race with the frames still in flight on this one).
__global__ void set_frame_kernel(int32_t* frame, int32_t value) {
    if (threadIdx.x == 0) *frame = value;
}

// Radiance-cache layout of one scene of a batch: its training-path stride,
// the most training paths one frame can give it, and its first training slot
// (slots of scene s: [slot_base, slot_base + n_train_max)).
struct NrcBatchScene {
    uint32_t stride, n_train_max, slot_base, pad;
};

// Training-path offset and count of a scene for the frame hash h
// (ap_render_nrc's rule, per scene).
__device__ __forceinline__ void nrc_scene_train(const NrcBatchScene& T, uint32_t n_pix, uint64_t h, uint32_t& offset,
                                                uint32_t& n_train) {
    offset = 0;
    n_train = 0;
    if (T.n_train_max == 0) return;
    offset = static_cast<uint32_t>(h % T.stride);
    n_train = min(T.n_train_max, (n_pix - offset + T.stride - 1) / T.stride);
}

// nrc_init_kernel over every scene of a batch; also clears the per-scene
// query counts. The frame hash comes from the batch's frame counter (so a
// captured frame replays with the right one).
__global__ __launch_bounds__(kBlock) void nrc_init_batch_kernel(BatchView bv, uint32_t total,
                                                                const NrcBatchScene* __restrict__ T, uint64_t seed,
                                                                uint32_t unbiased_every, NrcPath* __restrict__ paths,
                                                                uint32_t* __restrict__ counters,
                                                                uint32_t* __restrict__ qcount) {
    const uint64_t h = nrc_frame_hash(seed, *bv.frame);
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i == 0) { counters[0] = 0; counters[1] = 0; }
    if (i < static_cast<uint32_t>(bv.n_scenes)) qcount[i] = 0;
    if (i >= total) return;
    const int32_t sc = batch_scene_of(bv.pix_begin, bv.n_scenes, i);
    const uint32_t p = i - bv.pix_begin[sc], n_pix = bv.pix_begin[sc + 1] - bv.pix_begin[sc];
    const NrcBatchScene Ts = T[sc];
    uint32_t offset, n_train;
    nrc_scene_train(Ts, n_pix, h, offset, n_train);
    NrcPath P{0.0f, 0.0f, 0u, -1};
    if (n_train > 0 && p >= offset && (p - offset) % Ts.stride == 0u) {
        const uint32_t t = (p - offset) / Ts.stride;
        if (t < n_train) {
            P.flags = NRC_TRAIN | ((unbiased_every != 0u && t % unbiased_every == 0u) ? NRC_UNBIASED : 0u);
            P.slot = static_cast<int32_t>(Ts.slot_base + t);
        }
    }
    paths[i] = P;
}

// nrc_apply_kernel over a batch: scene s's queries are rows pix_begin[s] ..
// pix_begin[s] + qcount[s] of qpath and rgb.
template <int NL>
__global__ __launch_bounds__(kBlock) void nrc_apply_batch_kernel(BatchView bv, uint32_t total,
                                                                 const uint32_t* __restrict__ qcount,
                                                                 const int32_t* __restrict__ qpath,
                                                                 const float* __restrict__ rgb,
                                                                 const f4* __restrict__ rgb_lut,
                                                                 const f4* __restrict__ lam_lut,
                                                                 const float* __restrict__ tables,
                                                                 PathT<NL>* __restrict__ paths,
                                                                 GuideRec* __restrict__ recs = nullptr) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= total) return;
    const int32_t sc = batch_scene_of(bv.pix_begin, bv.n_scenes, i);
    if (i - bv.pix_begin[sc] >= qcount[sc]) return;
    const int32_t pid = qpath[i];
    nrc_apply_one<NL>(paths[pid], f3{rgb[3 * i], rgb[3 * i + 1], rgb[3 * i + 2]}, rgb_lut, lam_lut, tables,
                      recs ? recs + pid : nullptr);
}

// nrc_targets_kernel over the training slots of every scene of a batch;
// slot_pix[t] = the slot's path.
template <int NL>
__global__ __launch_bounds__(kBlock) void nrc_targets_batch_kernel(uint32_t slots, BatchView bv,
                                                                   const NrcBatchScene* __restrict__ T, uint64_t seed,
                                                                   const NrcPath* __restrict__ npaths,
                                                                   const PathT<NL>* __restrict__ paths,
                                                                   const float* __restrict__ tables,
                                                                   float* __restrict__ verts,
                                                                   uint32_t* __restrict__ vmask,
                                                                   uint32_t* __restrict__ slot_pix) {
    const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= slots) return;
    int32_t lo = 0, hi = bv.n_scenes - 1;  // the last scene whose slots start at or before t
    while (lo < hi) {
        const int32_t mid = (lo + hi + 1) >> 1;
        if (T[mid].slot_base <= t) lo = mid;
        else hi = mid - 1;
    }
    const NrcBatchScene Ts = T[lo];
    uint32_t offset, n_train;
    nrc_scene_train(Ts, bv.pix_begin[lo + 1] - bv.pix_begin[lo], nrc_frame_hash(seed, *bv.frame), offset, n_train);
    const uint32_t k = t - Ts.slot_base;
    if (k >= n_train) {
        vmask[t] = 0u;
        slot_pix[t] = 0u;
        return;
    }
    const uint32_t pid = bv.pix_begin[lo] + offset + k * Ts.stride;
    slot_pix[t] = pid;
    vmask[t] = nrc_slot_targets<NL>(t, pid, npaths, paths, tables, verts);
}

// Each scene's training samples: pool rows [pool_begin[s], + pool_n[s]).
__global__ void nrc_pool_ranges_kernel(int32_t n, uint32_t slots, const NrcBatchScene* __restrict__ T,
                                       const uint32_t* __restrict__ offsets, const uint32_t* __restrict__ counters,
                                       uint32_t* __restrict__ pool_begin, uint32_t* __restrict__ pool_n) {
    const int32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const uint32_t b = T[s].slot_base, e = b + T[s].n_train_max, all = counters[1];
    const uint32_t pb = b < slots ? offsets[b] : all, pe = e < slots ? offsets[e] : all;
    pool_begin[s] = pb;
    pool_n[s] = pe - pb;
}

__global__ void build_mats_kernel(const MatIn* __restrict__ in, int n, const f4* __restrict__ lut,
                                  MatRec* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const MatIn m = in[i];
    MatRec r;
    const float la = max3(m.albedo);
    r.albedo_basis = query_rgb_basis(lut, la > 1.0f ? m.albedo / la : m.albedo);
    r.albedo_scale = la;
    const float le = max3(m.emissive);
    r.emissive_basis = query_rgb_basis(lut, le > 1.0f ? m.emissive / le : m.emissive);
    r.emissive_scale = le;
    r.specular = m.specular;
    r.metallic = m.metallic;
    r.roughness = m.roughness;
    r.shading_model = m.model;
    r.light_index = m.light_index;
    r.albedo_tex = m.albedo_tex;
    r.emissive_rgb[0] = m.emissive.x;
    r.emissive_rgb[1] = m.emissive.y;
    r.emissive_rgb[2] = m.emissive.z;
    r.pad2 = 0;
    r.albedo_rgb[0] = m.albedo.x;
    r.albedo_rgb[1] = m.albedo.y;
    r.albedo_rgb[2] = m.albedo.z;
    r.pad3 = 0;
    r.ior = m.ior;
    r.sigma_a = m.sigma_a;
    r.abbe = m.abbe;
    r.tilt = m.tilt;
    out[i] = r;
}

__global__ void xyz_to_srgb_kernel(const float* __restrict__ xyz, size_t n, float* __restrict__ rgb) {
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const f3 r = xyz_to_srgb(f3{xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]});
    rgb[3 * i] = r.x; rgb[3 * i + 1] = r.y; rgb[3 * i + 2] = r.z;
}

inline dim3 grid_of(size_t n) { return dim3(static_cast<unsigned>((n + kBlock - 1) / kBlock)); }

template <typename T>
int dev_alloc(T** p, size_t count) {
    *p = nullptr;
    if (count == 0) return AP_OK;
    const hipError_t e = hipMalloc(reinterpret_cast<void**>(p), count * sizeof(T));
    return e == hipSuccess ? AP_OK : hip_fail(e, "hipMalloc(render)");
}

template <typename T>
int upload_vec(const std::vector<T>& v, T** p) {
    int rc = dev_alloc(p, v.size());
    if (rc != AP_OK || v.empty()) return rc;
    const hipError_t e = hipMemcpy(*p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice);
    return e == hipSuccess ? AP_OK : hip_fail(e, "hipMemcpy(render)");
}

inline f3 F(const ap_float3& a) { return {a.x, a.y, a.z}; }
inline m3 M3(const ap_float3x3& m) { return {F(m.x), F(m.y), F(m.z)}; }

// Shading records of the scene for a triangle-material mode.
int build_shading(ap_scene* s, RenderState* r, int tri_mode, const f4* lut) {
    for (void* p : {static_cast<void*>(r->geo), static_cast<void*>(r->mats), static_cast<void*>(r->lights),
                    static_cast<void*>(r->uvs)})
        if (p) (void)hipFree(p);
    r->geo = nullptr; r->mats = nullptr; r->lights = nullptr; r->uvs = nullptr;
    const size_t n = s->objects.size();
    std::vector<GeoRec> geo(n);
    std::vector<float> uvs(n * 6, 0.0f);
    int32_t max_tex = -1;
    std::vector<MatIn> mats;
    std::map<std::vector<uint32_t>, int32_t> mat_index;
    auto intern = [&](const MatIn& m) {
        std::vector<uint32_t> key(sizeof(MatIn) / 4);
        std::memcpy(key.data(), &m, sizeof(MatIn));
        auto it = mat_index.find(key);
        if (it != mat_index.end()) return it->second;
        const int32_t id = static_cast<int32_t>(mats.size());
        mats.push_back(m);
        mat_index.emplace(key, id);
        return id;
    };
    // Analytic record index per object, in slot order as the scene built them.
    std::vector<int32_t> analytic_of(n, -1);
    {
        int32_t k = 0;
        for (size_t slot = 0; slot < s->bvh.indices.size(); ++slot) {
            const ap_scene_object& o = s->objects[s->bvh.indices[slot]];
            if (o.type == AP_OBJ_CUBE || o.type == AP_OBJ_SPHERE || o.type == AP_OBJ_SDF)
                analytic_of[s->bvh.indices[slot]] = k++;
        }
    }
    for (size_t i = 0; i < n; ++i) {
        const ap_scene_object& o = s->objects[i];
        const ap_gbuffer& gb = o.material.texture.base;
        // Only the albedo texture reaches the BSDFs (material.cuh:69-74, 198-202);
        // the other handles are carried and not read, as in the reference.
        if (gb.albedo_tex > static_cast<uint64_t>(AP_TEXTURE_CACHE_CAPACITY))
            return fail(AP_ERR_INVALID_ARG, "ap_render: albedo texture handle out of range (object " + std::to_string(i) + ")");
        MatIn m;
        m.albedo_tex = gb.albedo_tex ? static_cast<int32_t>(gb.albedo_tex - 1u) : -1;
        max_tex = std::max(max_tex, m.albedo_tex);
        const bool ref_tri = o.type == AP_OBJ_TRIANGLE && tri_mode == AP_TRIANGLE_MATERIAL_REFERENCE;
        if (ref_tri) {  // default-constructed GBuffer (object.cuh:60-72)
            m.albedo = {1.0f, 1.0f, 1.0f};
            m.emissive = {0.0f, 0.0f, 0.0f};
            m.specular = 0.0f; m.metallic = 0.0f; m.roughness = 0.0f;
            m.ior = 1.5f; m.sigma_a = 100.0f; m.abbe = 20.0f; m.tilt = 0.0f;
        } else {
            m.albedo = F(gb.albedo);
            m.emissive = F(gb.emissive);
            m.specular = gb.specular; m.metallic = gb.metallic; m.roughness = gb.roughness;
            m.ior = gb.ior; m.sigma_a = gb.sigma_a; m.abbe = gb.abbe_number; m.tilt = gb.tilt_number;
        }
        m.model = o.material.shading_model;
        m.light_index = o.light_index;
        GeoRec& g = geo[i];
        std::memset(&g, 0, sizeof(g));
        g.type = o.type;
        g.mat = intern(m);
        g.analytic = analytic_of[i];
        g.light_index = o.light_index;
        g.model = m.model;
        if (o.type == AP_OBJ_TRIANGLE) {
            const m3 o2w = M3(o.object_to_world);
            const f3 c = F(o.center);
            const f3 p0 = mul(o2w, F(o.additional.v0.position)) + c;
            const f3 p1 = mul(o2w, F(o.additional.v1.position)) + c;
            const f3 p2 = mul(o2w, F(o.additional.v2.position)) + c;
            const f3 e1 = p1 - p0, e2 = p2 - p0;
            const f3 n0 = mul(o2w, F(o.additional.v0.normal));
            const f3 n1 = mul(o2w, F(o.additional.v1.normal));
            const f3 n2 = mul(o2w, F(o.additional.v2.normal));
            const float* src[6] = {&p0.x, &e1.x, &e2.x, &n0.x, &n1.x, &n2.x};
            float* dst[6] = {g.p0, g.e1, g.e2, g.n0, g.n1, g.n2};
            for (int k = 0; k < 6; ++k) std::memcpy(dst[k], src[k], 12);
            const ap_vertex* vs[3] = {&o.additional.v0, &o.additional.v1, &o.additional.v2};
            for (int k = 0; k < 3; ++k) {
                uvs[i * 6 + 2 * k] = vs[k]->uv.x;
                uvs[i * 6 + 2 * k + 1] = vs[k]->uv.y;
            }
        }
    }
    // Lights (scene.cu:1056-1080) with the OBB axes of GetOBBAxes (nee.cuh:24-31).
    std::vector<LightRec> lights(s->lights.size());
    for (size_t k = 0; k < s->lights.size(); ++k) {
        const ap_light_info& li = s->lights[k];
        const ap_scene_object& o = s->objects[li.object_index];
        LightRec& L = lights[k];
        std::memset(&L, 0, sizeof(L));
        L.type = (o.type == AP_OBJ_CUBE || o.type == AP_OBJ_SPHERE) ? o.type : -1;
        L.object = li.object_index;
        L.center[0] = o.center.x; L.center[1] = o.center.y; L.center[2] = o.center.z;
        L.extent[0] = o.extent.x; L.extent[1] = o.extent.y; L.extent[2] = o.extent.z;
        const ap_float3x3& Mw = o.object_to_world;
        const f3 ux = normalize(f3{Mw.x.x, Mw.y.x, Mw.z.x});
        const f3 uy = normalize(f3{Mw.x.y, Mw.y.y, Mw.z.y});
        const f3 uz = normalize(f3{Mw.x.z, Mw.y.z, Mw.z.z});
        std::memcpy(L.ux, &ux, 12); std::memcpy(L.uy, &uy, 12); std::memcpy(L.uz, &uz, 12);
        L.select_prob = li.select_prob;
        L.mat = geo[li.object_index].mat;
    }
    int rc;
    MatIn* d_in = nullptr;
    if ((rc = upload_vec(geo, &r->geo)) != AP_OK) return rc;
    if ((rc = upload_vec(uvs, &r->uvs)) != AP_OK) return rc;
    r->max_tex = max_tex;
    if ((rc = upload_vec(lights, &r->lights)) != AP_OK) return rc;
    if ((rc = upload_vec(mats, &d_in)) != AP_OK) return rc;
    if ((rc = dev_alloc(&r->mats, std::max<size_t>(mats.size(), 1))) != AP_OK) { (void)hipFree(d_in); return rc; }
    if (!mats.empty()) {
        hipLaunchKernelGGL(build_mats_kernel, dim3((mats.size() + 127) / 128), dim3(128), 0, nullptr, d_in,
                           static_cast<int>(mats.size()), lut, r->mats);
        hipError_t e = hipGetLastError();
        if (e == hipSuccess) e = hipDeviceSynchronize();
        (void)hipFree(d_in);
        if (e != hipSuccess) return hip_fail(e, "build_mats_kernel");
    }
    r->n_lights = static_cast<int32_t>(lights.size());
    r->n_mats = static_cast<int32_t>(mats.size());
    r->analytic_objs.clear();
    for (size_t i = 0; i < n; ++i)
        if (s->objects[i].type == AP_OBJ_CUBE || s->objects[i].type == AP_OBJ_SPHERE ||
            s->objects[i].type == AP_OBJ_SDF)
            r->analytic_objs.push_back(static_cast<int32_t>(i));
    r->has_glass = false;
    for (size_t i = 0; i < n; ++i) r->has_glass = r->has_glass || s->objects[i].material.shading_model == AP_MAT_GLASS;
    r->tri_mode = tri_mode;
    return AP_OK;
}

int ensure_paths(RenderState* r, size_t cap) {
    if (r->capacity >= cap) return AP_OK;
    void* old[] = {r->paths, r->media, r->rays[0], r->rays[1], r->path_of[0], r->path_of[1], r->hits, r->shadow_rays,
                   r->shadow_payload, r->shadow_hits};
    for (void* p : old)
        if (p) (void)hipFree(p);
    r->capacity = 0;
    int rc;
    if ((rc = dev_alloc(reinterpret_cast<uint8_t**>(&r->paths), cap * kPathBytes)) != AP_OK) return rc;
    if ((rc = dev_alloc(&r->media, cap)) != AP_OK) return rc;
    for (int q = 0; q < 2; ++q) {
        if ((rc = dev_alloc(&r->rays[q], cap)) != AP_OK) return rc;
        if ((rc = dev_alloc(&r->path_of[q], cap)) != AP_OK) return rc;
    }
    if ((rc = dev_alloc(&r->hits, cap)) != AP_OK) return rc;
    if ((rc = dev_alloc(&r->shadow_rays, cap)) != AP_OK) return rc;
    if ((rc = dev_alloc(reinterpret_cast<uint8_t**>(&r->shadow_payload), cap * kShadowPayBytes)) != AP_OK) return rc;
    if ((rc = dev_alloc(&r->shadow_hits, cap)) != AP_OK) return rc;
    r->capacity = cap;
    return AP_OK;
}

// OverlapBase at the camera origin over every object (intersection.cuh:399-414,
// aabb.cuh:91-106, sphere.cuh:179-192, sdf_base.cuh:165-177): whether the
// camera starts inside one.
bool camera_inside(const ap_scene* s, const std::vector<int32_t>& analytic, f3 o) {
    for (const int32_t k : analytic) {
        const ap_scene_object& ob = s->objects[k];
        const f3 lo = mul(M3(ob.world_to_object), o - F(ob.center));
        const f3 e = F(ob.extent);
        if (ob.type == AP_OBJ_SDF) {
            const ap_float4& q = ob.additional.sdf_info.sdf_info0;
            const float info[4] = {q.x, q.y, q.z, q.w};
            SdfVolume host{};
            SdfField field;
            bool ok = false;
            if (ob.additional.sdf_type == SDF_TORUS) {
                ok = sdf_field(SDF_TORUS, -1, info, nullptr, 0, field);
            } else if (ob.additional.sdf_type == SDF_VOLUME &&
                       sdf_cache_host_volume(s->sdf_cache, ob.additional.sdf_info.cached_volume_index, host)) {
                ok = sdf_field(SDF_VOLUME, 0, info, &host, 1, field);
            }
            if (ok && field(to_unit(lo, e)) <= 0.0f) return true;
        } else if (ob.type == AP_OBJ_CUBE) {
            if (-e.x < lo.x && -e.y < lo.y && -e.z < lo.z && e.x > lo.x && e.y > lo.y && e.z > lo.z) return true;
        } else {
            const float rr = max3(e);
            if (dot(lo, lo) < rr * rr) return true;
        }
    }
    return false;
}

// The side stream (default priority, non-blocking) and its two events.
int ensure_side(RenderState* r) {
    if (r->side) return AP_OK;
    hipError_t e = hipEventCreateWithFlags(&r->ev_shaded, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&r->ev_shadowed, hipEventDisableTiming);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&r->side, hipStreamNonBlocking);
    return e == hipSuccess ? AP_OK : hip_fail(e, "ap_render: side stream");
}

// The second frame set and the pipeline's stream and events.
int ensure_pipeline(RenderState* r, size_t cap) {
    if (!r->lead) {
        hipError_t e = hipStreamCreateWithFlags(&r->lead, hipStreamNonBlocking);
        hipEvent_t* evs[] = {&r->ev_start, &r->ev_tail, &r->ev_begun, &r->ev_inferred, &r->ev_trained};
        for (hipEvent_t* ev : evs)
            if (e == hipSuccess) e = hipEventCreateWithFlags(ev, hipEventDisableTiming);
        if (e != hipSuccess) return hip_fail(e, "ap_render: pipeline stream");
    }
    if (r->alt_capacity >= cap) return AP_OK;
    void* old[] = {r->alt_paths, r->alt_media, r->alt_rays0, r->alt_path_of0, r->alt_hits, r->alt_counters};
    for (void* p : old)
        if (p) (void)hipFree(p);
    r->alt_paths = nullptr; r->alt_media = nullptr; r->alt_rays0 = nullptr; r->alt_path_of0 = nullptr;
    r->alt_hits = nullptr; r->alt_counters = nullptr;
    r->alt_capacity = 0;
    int rc;
    if ((rc = dev_alloc(reinterpret_cast<uint8_t**>(&r->alt_paths), cap * kPathBytes)) != AP_OK) return rc;
    if ((rc = dev_alloc(&r->alt_media, cap)) != AP_OK) return rc;
    if ((rc = dev_alloc(&r->alt_rays0, cap)) != AP_OK) return rc;
    if ((rc = dev_alloc(&r->alt_path_of0, cap)) != AP_OK) return rc;
    if ((rc = dev_alloc(&r->alt_hits, cap)) != AP_OK) return rc;
    if ((rc = dev_alloc(&r->alt_counters, 4)) != AP_OK) return rc;
    r->alt_capacity = cap;
    return AP_OK;
}

int ensure_nrc(RenderState* r, size_t cap, size_t train_cap) {
    if (!r->nrc_counters) {  // four per frame set
        const hipError_t e = hipMalloc(reinterpret_cast<void**>(&r->nrc_counters), 8 * sizeof(uint32_t));
        if (e != hipSuccess) return hip_fail(e, "hipMalloc(cache counters)");
    }
    int rc;
    if (r->nrc_cap < cap) {
        void* old[] = {r->nrc_paths, r->nrc_queries, r->nrc_qpath, r->nrc_qrgb};
        for (void* p : old)
            if (p) (void)hipFree(p);
        r->nrc_paths = nullptr; r->nrc_queries = nullptr; r->nrc_qpath = nullptr; r->nrc_qrgb = nullptr;
        r->nrc_cap = 0;
        if ((rc = dev_alloc(&r->nrc_paths, cap)) != AP_OK) return rc;
        if ((rc = dev_alloc(&r->nrc_queries, cap * kNrcRec)) != AP_OK) return rc;
        if ((rc = dev_alloc(&r->nrc_qpath, cap)) != AP_OK) return rc;
        if ((rc = dev_alloc(&r->nrc_qrgb, cap * 3)) != AP_OK) return rc;
        r->nrc_cap = cap;
    }
    train_cap = std::max<size_t>(train_cap, 1);
    if (r->nrc_train_cap < train_cap) {
        void* old[] = {r->nrc_verts, r->nrc_vmask, r->nrc_offsets, r->nrc_pool};
        for (void* p : old)
            if (p) (void)hipFree(p);
        r->nrc_verts = nullptr; r->nrc_vmask = nullptr; r->nrc_offsets = nullptr; r->nrc_pool = nullptr;
        r->nrc_train_cap = 0;
        if ((rc = dev_alloc(&r->nrc_verts, train_cap * kNrcV * kNrcVert)) != AP_OK) return rc;
        if ((rc = dev_alloc(&r->nrc_vmask, train_cap)) != AP_OK) return rc;
        if ((rc = dev_alloc(&r->nrc_offsets, nrc_offsets_words(train_cap))) != AP_OK) return rc;
        if ((rc = dev_alloc(&r->nrc_pool, train_cap * kNrcV * kNrcSample)) != AP_OK) return rc;
        r->nrc_train_cap = train_cap;
    }
    return AP_OK;
}

// SDF_EPS(depth) = INTERSECT_EPS(depth) / 2 (config.cuh:74), INTERSECT_EPS as
// the device table computes it: the marching epsilon of a bounce's shadow rays.
float sdf_eps_of_depth(int depth) {
    return 0.5f * (1.0e-3f * std::exp(fminf(static_cast<float>(depth) * 0.5f, 4.0f)));
}


}  // namespace

// Context of one render call, shared by the wavelength variants.
struct RunCtx {
    ap_scene* s;
    RenderState* r;
    const ap_render_params* prm;
    const ap_framebuffers* fb;
    hipStream_t stream;
    const ap_nrc_render_params* np;
    NrcState* ns;
    const f4* rgb_lut;
    const f4* lam_lut;
    const float* tables;
    int32_t n_pix, per_batch;
    bool guiding, nrc, pipelined;
    uint32_t n_stride, n_train_max;
    CamParams cp;
    int32_t inside, last_frame;
};

// The buffers of one frame in flight: what the next frame's start (its
// camera rays and their closest hits) writes while this frame finishes.
// Queue 1 and the shadow queue are shared: the next frame first uses them
// in its first shade, after this frame has ended.
template <int NL>
struct FrameSet {
    PathT<NL>* paths;
    MediaRec* media;
    ap_ray* rays[2];
    int32_t* path_of[2];
    ap_hit* hits;
    uint32_t* counters;      // [0],[1] path queues, [2] shadow queue
    GuideRec* recs;          // guiding path records, or null
    uint32_t* nrc_counters;  // cache counters, or null
};

template <int NL>
int run_frames(const RunCtx& c) {
    ap_scene* s = c.s;
    RenderState* r = c.r;
    const ap_render_params* prm = c.prm;
    const ap_framebuffers* fb = c.fb;
    const hipStream_t stream = c.stream;
    const ap_nrc_render_params* np = c.np;
    NrcState* ns = c.ns;
    const f4 *rgb_lut = c.rgb_lut, *lam_lut = c.lam_lut;
    const float* tables = c.tables;
    const int32_t n_pix = c.n_pix, per_batch = c.per_batch;
    const bool guiding = c.guiding, nrc = c.nrc, pipelined = c.pipelined;
    const uint32_t n_stride = c.n_stride, n_train_max = c.n_train_max;
    const CamParams cp = c.cp;
    const int32_t inside = c.inside, last_frame = c.last_frame;
    auto* pays = static_cast<ShadowPay<NL>*>(r->shadow_payload);
    int rc = AP_OK;
    // The path queues are traced at most kQueueWaveRays rays per wave, their
    // chunks dealt round-robin over the XCDs. Measured at 1080p with four
    // wavelengths: with 16M paths per batch, 512-ray chunks give 4.65 ms per
    // sample (128: 4.98, 1024: 4.65, 2048: 4.72) and guided frames (2M paths)
    // 10.0 ms (128: 10.2); with 4M paths per batch 128 had been best.
    constexpr int32_t kQueueWaveRays = 512;
    // A pipelined frame starts the next one once its first lead_bounces
    // bounces are launched: from there on its queues are short and the GPU
    // has room for the next frame's camera rays. Measured at 1080p with four
    // wavelengths: guided frames 8.87 ms when the next frame starts after
    // bounce 2, 8.80 after 3, 8.60 after 4, 8.62 after 5; cache frames, whose
    // paths end sooner, are fastest after bounce 2 (5.17 ms; 5.35 after 4).
    const int32_t lead_bounces = nrc ? 2 : 4;
    DeviceScene qs = s->dev;
    qs.wave_rays = std::min(qs.wave_rays, kQueueWaveRays);
    auto frame_set = [&](int k) {
        FrameSet<NL> f;
        if (k == 0) {
            f.paths = static_cast<PathT<NL>*>(r->paths);
            f.media = r->media;
            f.rays[0] = r->rays[0];
            f.path_of[0] = r->path_of[0];
            f.hits = r->hits;
            f.counters = r->counters;
        } else {
            f.paths = static_cast<PathT<NL>*>(r->alt_paths);
            f.media = r->alt_media;
            f.rays[0] = r->alt_rays0;
            f.path_of[0] = r->alt_path_of0;
            f.hits = r->alt_hits;
            f.counters = r->alt_counters;
        }
        f.rays[1] = r->rays[1];
        f.path_of[1] = r->path_of[1];
        f.recs = guiding ? guide_recs(s->guide, k) : nullptr;
        f.nrc_counters = nrc ? r->nrc_counters + 4 * k : nullptr;
        return f;
    };
    // The start of a batch of paths on `st`: path states and camera rays,
    // then their closest hits (bounce 0's extension trace).
    auto begin_batch = [&](const FrameParams& fp, const FrameSet<NL>& F, bool train_only, hipStream_t st) -> int {
        const uint32_t total = static_cast<uint32_t>(fp.frames) * static_cast<uint32_t>(fp.n_pixels);
        hipLaunchKernelGGL(set_counters_kernel, dim3(1), dim3(64), 0, st, F.counters, total, 0u, 0u);
        hipLaunchKernelGGL(generate_kernel<NL>, grid_of(total), dim3(kBlock), 0, st, cp, fp, r->eps, F.paths,
                           F.rays[0], F.path_of[0], (guiding && !train_only) ? F.recs : nullptr, F.media);
        hipLaunchKernelGGL(begin_bounce_kernel, dim3(1), dim3(64), 0, st, F.counters, 1);
        const hipError_t e =
            prm->max_depth > 0 ? launch_trace(qs, F.rays[0], total, F.hits, nullptr, false, st, &F.counters[0])
                               : hipSuccess;
        return e == hipSuccess ? AP_OK : hip_fail(e, "ap_render: extend");
    };
    bool trained_on_lead = false;   // a cache training step runs on `lead`, not yet waited for
    // The rest of a batch of paths on the render stream: a rendered batch of
    // frames, or (train_only) the training paths of one pass of the radiance
    // cache, which renders nothing into the framebuffers. at_tail runs once
    // the first lead_bounces bounces are launched.
    auto finish_batch = [&](const FrameParams& fp, const FrameSet<NL>& F, bool train_only, uint32_t n_offset,
                            uint32_t n_train, uint64_t train_seed, const std::function<int()>& at_tail) -> int {
        const uint32_t total = static_cast<uint32_t>(fp.frames) * static_cast<uint32_t>(fp.n_pixels);
        GuideView gv{};
        if (guiding && !train_only) {
            gv = guide_view(s->guide);   // the probes the frames before built
            gv.recs = F.recs;
        }
        NrcView nv{};
        if (nrc) {
            hipLaunchKernelGGL(nrc_init_kernel, grid_of(static_cast<size_t>(total)), dim3(kBlock), 0, stream,
                               static_cast<int32_t>(total), n_stride, n_offset, n_train, np->unbiased_every,
                               train_only ? 1u : 0u, r->nrc_paths, F.nrc_counters);
            nv = NrcView{r->nrc_paths, r->nrc_queries, r->nrc_qpath, r->nrc_verts, F.nrc_counters, np->spread};
        }
        auto* rc_ptr = train_only ? nullptr : reinterpret_cast<unsigned long long*>(fb->ray_count);
        int32_t* prim = train_only ? nullptr : fb->primary_id;
        // Bounce k: shade, then on the side stream the shadow rays, their
        // resolve and the counter resets for the next shade (the queue this
        // shade read, the shadow queue after resolve has read it), while the
        // render stream traces bounce k + 1's extension rays (independent
        // queues and outputs, the path states untouched); joined before the
        // next shade.
        int in_q = 0;
        hipError_t e = hipSuccess;
        bool tail_started = false;
        int rc2 = AP_OK;
        for (int32_t depth = 0; depth < prm->max_depth; ++depth) {
            const int out_q = 1 - in_q;
            const bool more = depth + 1 < prm->max_depth;
            ShadeParams sp{depth, prm->max_depth, prm->rr_min_depth, prm->nee_max_depth, prm->use_nee,
                           prm->nee_always, prm->sky, in_q, out_q, r->n_lights};
            if (fb->gbuffer && !train_only) sp.gsample = r->gsamples;
            sp.volumes = s->dev.volumes;
            sp.n_volumes = s->dev.n_volumes;
            sp.media = F.media;
            sp.textures = texture_cache_table(s->tex_cache);
            sp.uvs = r->uvs;
#define AP_SHADE(G, N)                                                                                         \
    hipLaunchKernelGGL((shade_kernel<NL, G, N>), grid_of(total), dim3(kBlock), 0, stream, sp, r->geo, r->mats, \
                       r->lights, s->dev.analytic, rgb_lut, lam_lut, tables, r->eps, F.paths, F.rays[in_q],    \
                       F.path_of[in_q], F.hits, F.rays[out_q], F.path_of[out_q], r->shadow_rays, pays,         \
                       F.counters, prim, last_frame, rc_ptr, gv, nv)
            if (guiding && nrc && !train_only) AP_SHADE(true, true);
            else if (guiding && !nrc) AP_SHADE(true, false);
            else if (nrc) AP_SHADE(false, true);
            else AP_SHADE(false, false);
#undef AP_SHADE
            e = hipEventRecord(r->ev_shaded, stream);
            if (e == hipSuccess) e = hipStreamWaitEvent(r->side, r->ev_shaded, 0);
            if (e == hipSuccess)
                e = launch_trace(qs, r->shadow_rays, total, r->shadow_hits, nullptr, true, r->side, &F.counters[2],
                                 sdf_eps_of_depth(depth));
            if (e == hipSuccess) {
                hipLaunchKernelGGL(resolve_kernel<NL>, grid_of(total), dim3(kBlock), 0, r->side, F.counters, pays,
                                   r->shadow_hits, F.paths, rc_ptr);
                if (more) hipLaunchKernelGGL(begin_bounce_kernel, dim3(1), dim3(64), 0, r->side, F.counters, in_q);
                e = hipEventRecord(r->ev_shadowed, r->side);
            }
            if (e != hipSuccess) return hip_fail(e, "ap_render: shadow");
            if (more) {
                e = launch_trace(qs, F.rays[out_q], total, F.hits, nullptr, false, stream, &F.counters[out_q]);
                if (e != hipSuccess) return hip_fail(e, "ap_render: extend");
            }
            e = hipStreamWaitEvent(stream, r->ev_shadowed, 0);
            if (e != hipSuccess) return hip_fail(e, "ap_render: join");
            in_q = out_q;
            if (depth + 1 == lead_bounces) {
                tail_started = true;
                if ((rc2 = at_tail()) != AP_OK) return rc2;
            }
            // Deep limits: stop once every path has ended.
            if (prm->max_depth > 32 && (depth % 16) == 15) {
                uint32_t live = 0;
                e = hipMemcpyAsync(&live, &F.counters[in_q], sizeof(uint32_t), hipMemcpyDeviceToHost, stream);
                if (e == hipSuccess) e = hipStreamSynchronize(stream);
                if (e != hipSuccess) return hip_fail(e, "ap_render: live count");
                if (live == 0) break;
            }
        }
        if (!tail_started && (rc2 = at_tail()) != AP_OK) return rc2;
        if (nrc) {
            // Cache vertices (with the weights the frames before trained),
            // then the training samples of the batch.
            if (trained_on_lead) {
                if ((e = hipStreamWaitEvent(stream, r->ev_trained, 0)) != hipSuccess)
                    return hip_fail(e, "ap_render: wait for training");
                trained_on_lead = false;
            }
            if ((rc2 = nrc_infer_dyn(ns, r->nrc_queries, r->nrc_qpath, &F.nrc_counters[0], total, r->nrc_qrgb,
                                     stream)) != AP_OK)
                return rc2;
            hipLaunchKernelGGL(nrc_apply_kernel<NL>, grid_of(total), dim3(kBlock), 0, stream, F.nrc_counters,
                               r->nrc_qpath, r->nrc_qrgb, rgb_lut, lam_lut, tables, F.paths, gv.recs);
            if (n_train > 0) {
                hipLaunchKernelGGL(nrc_targets_kernel<NL>, grid_of(n_train), dim3(kBlock), 0, stream, n_train,
                                   n_stride, n_offset, train_only ? 1u : 0u, r->nrc_paths, F.paths, tables,
                                   r->nrc_verts, r->nrc_vmask);
                const hipError_t se = nrc_scan(n_train, r->nrc_vmask, r->nrc_offsets, r->nrc_train_cap,
                                               F.nrc_counters, stream);
                if (se != hipSuccess) return hip_fail(se, "ap_render: training sample scan");
                hipLaunchKernelGGL(nrc_compact_kernel, grid_of(n_train), dim3(kBlock), 0, stream, n_train, n_stride,
                                   n_offset, r->nrc_vmask, r->nrc_offsets, r->nrc_verts, r->nrc_pool);
            }
            r->nrc_last_train = n_train;
            r->nrc_last_set = static_cast<int>(F.nrc_counters - r->nrc_counters) / 4;
            r->nrc_rendered = true;
        }
        if (!train_only)
            hipLaunchKernelGGL(accumulate_kernel<NL>, grid_of(fp.n_pixels), dim3(kBlock), 0, stream, fp, F.paths,
                               tables, fb->hdr, fb->odd, fb->even, fb->sum_xyz, gv.recs,
                               guiding ? guide_slots(s->guide) : nullptr, guiding ? guide_nvalid(s->guide) : nullptr,
                               r->gsamples, fb->gbuffer, fb->statistics);
        e = hipGetLastError();
        if (e != hipSuccess) return hip_fail(e, "ap_render: launch");
        if (nrc && pipelined) {
            // The cache trains on `lead` beside the next frame's bounces;
            // the next frame's inference waits for it.
            e = hipEventRecord(r->ev_inferred, stream);
            if (e == hipSuccess) e = hipStreamWaitEvent(r->lead, r->ev_inferred, 0);
            if (e != hipSuccess) return hip_fail(e, "ap_render: training fork");
        }
        if (guiding && !train_only &&
            (rc2 = guide_update(s->guide, static_cast<size_t>(fp.frames) * n_pix, stream)) != AP_OK)
            return rc2;
        if (nrc && n_train > 0 && np->train_steps > 0) {
            const hipStream_t ts = pipelined ? r->lead : stream;
            rc2 = nrc_train_pool(ns, r->nrc_pool, nullptr, &F.nrc_counters[1], np->batch, np->train_steps,
                                 np->learning_rate, train_seed, ts);
            if (rc2 != AP_OK) return rc2;
            if (pipelined) {
                if ((e = hipEventRecord(r->ev_trained, r->lead)) != hipSuccess)
                    return hip_fail(e, "ap_render: training event");
                trained_on_lead = true;
            }
        }
        return AP_OK;
    };
    // Training pixels of a frame (or pass) from its hash: offset, count.
    auto train_pixels = [&](uint64_t h, uint32_t& n_offset, uint32_t& n_train) {
        n_offset = 0;
        n_train = 0;
        if (n_train_max > 0) {
            n_offset = static_cast<uint32_t>(h % n_stride);
            n_train = std::min(n_train_max, (static_cast<uint32_t>(n_pix) - n_offset + n_stride - 1) / n_stride);
        }
    };
    auto batch_params = [&](int32_t f0) {
        const int32_t nf = std::min(per_batch, prm->spp - f0);
        return FrameParams{prm->frame_index + f0, nf, n_pix, inside, prm->seed, 0u, 1u};
    };
    const std::function<int()> no_tail = [] { return AP_OK; };
    hipError_t e = hipSuccess;
    if (pipelined) {
        // The first frame starts on `lead` after the work before this call.
        e = hipEventRecord(r->ev_start, stream);
        if (e == hipSuccess) e = hipStreamWaitEvent(r->lead, r->ev_start, 0);
        if (e != hipSuccess) return hip_fail(e, "ap_render: pipeline start");
        if ((rc = begin_batch(batch_params(0), frame_set(0), false, r->lead)) != AP_OK) return rc;
        if ((e = hipEventRecord(r->ev_begun, r->lead)) != hipSuccess) return hip_fail(e, "ap_render: pipeline");
    }
    int k = 0;   // frame set of the current batch
    for (int32_t f0 = 0; f0 < prm->spp; f0 += per_batch) {
        const FrameParams fp = batch_params(f0);
        const FrameSet<NL> F = frame_set(k);
        uint32_t n_offset = 0, n_train = 0;
        uint64_t h = 0;
        if (nrc) {
            h = nrc_frame_hash(prm->seed, fp.frame0);
            train_pixels(h, n_offset, n_train);
        }
        const bool has_next = f0 + per_batch < prm->spp;
        std::function<int()> start_next = no_tail;
        if (pipelined) {
            if ((e = hipStreamWaitEvent(stream, r->ev_begun, 0)) != hipSuccess)
                return hip_fail(e, "ap_render: pipeline");
            if (has_next)
                start_next = [&, f0, k]() -> int {
                    hipError_t e2 = hipEventRecord(r->ev_tail, stream);
                    if (e2 == hipSuccess) e2 = hipStreamWaitEvent(r->lead, r->ev_tail, 0);
                    if (e2 != hipSuccess) return hip_fail(e2, "ap_render: pipeline");
                    const int rc3 = begin_batch(batch_params(f0 + per_batch), frame_set(1 - k), false, r->lead);
                    if (rc3 != AP_OK) return rc3;
                    e2 = hipEventRecord(r->ev_begun, r->lead);
                    return e2 == hipSuccess ? AP_OK : hip_fail(e2, "ap_render: pipeline");
                };
        } else if ((rc = begin_batch(fp, F, false, stream)) != AP_OK) {
            return rc;
        }
        if ((rc = finish_batch(fp, F, false, n_offset, n_train, h ^ kNrcTrainSalt, start_next)) != AP_OK) return rc;
        // Training-only passes of the cache: the training paths alone, on
        // their own pixels and random streams.
        const uint32_t passes = (nrc && n_train_max > 0) ? np->train_passes : 0u;
        for (uint32_t q = 0; q < passes; ++q) {
            const uint64_t hk = splitmix64(h ^ (0x70617373ull + q));
            uint32_t o2 = 0, t2 = 0;
            train_pixels(hk, o2, t2);
            if (t2 == 0) continue;
            const FrameParams tp{fp.frame0, 1, static_cast<int32_t>(t2), inside, splitmix64(prm->seed ^ hk), o2,
                                 n_stride};
            const FrameSet<NL> T = frame_set(0);
            if ((rc = begin_batch(tp, T, true, stream)) != AP_OK) return rc;
            if ((rc = finish_batch(tp, T, true, o2, t2, hk ^ kNrcTrainSalt, no_tail)) != AP_OK) return rc;
        }
        if (pipelined) k = 1 - k;
    }
    if (pipelined) {
        // The caller's stream ends after everything on `lead` (the last training).
        e = hipEventRecord(r->ev_start, r->lead);
        if (e == hipSuccess) e = hipStreamWaitEvent(stream, r->ev_start, 0);
        if (e != hipSuccess) return hip_fail(e, "ap_render: pipeline end");
    }
    return AP_OK;
}

int render_frames(ap_scene* s, const ap_camera* cam, const ap_render_params* prm, const ap_framebuffers* fb,
                  hipStream_t stream, ap_nrc* cache = nullptr, const ap_nrc_render_params* np = nullptr) {
    if (!s || !cam || !prm || !fb || !fb->hdr) return fail(AP_ERR_INVALID_ARG, "ap_render: NULL argument");
    if (cam->width <= 0 || cam->height <= 0) return fail(AP_ERR_INVALID_ARG, "ap_render: bad resolution");
    if (prm->spp < 0 || prm->frame_index < 0 || prm->max_depth < 0 || prm->max_depth > kEpsEntries - 2)
        return fail(AP_ERR_INVALID_ARG, "ap_render: bad spp, frame_index or max_depth");
    if (prm->triangle_material != AP_TRIANGLE_MATERIAL_OBJECT && prm->triangle_material != AP_TRIANGLE_MATERIAL_REFERENCE)
        return fail(AP_ERR_INVALID_ARG, "ap_render: bad triangle_material");
    if (prm->wavelengths != 0 && prm->wavelengths != 1 && prm->wavelengths != 4)
        return fail(AP_ERR_INVALID_ARG, "ap_render: wavelengths must be 1 or 4 (0 reads as 1)");
    if (prm->spp == 0 || prm->max_depth == 0) return AP_OK;
    const f4 *rgb_lut = nullptr, *lam_lut = nullptr;
    const float* tables = nullptr;
    int rc = spectrum_get(&rgb_lut, &lam_lut, &tables);
    if (rc != AP_OK) return rc;
    if (!s->render) s->render = new RenderState();
    RenderState* r = s->render;
    if (!r->eps) {
        // INTERSECT_EPS(depth) = 1e-3 * expf(min(depth / 2, 4)) (config.cuh:3), on the host.
        std::vector<float> e(kEpsEntries);
        for (int d = 0; d < kEpsEntries; ++d) e[d] = 1.0e-3f * std::exp(fminf(static_cast<float>(d) * 0.5f, 4.0f));
        if ((rc = upload_vec(e, &r->eps)) != AP_OK) return rc;
        if ((rc = dev_alloc(&r->counters, 4)) != AP_OK) return rc;
    }
    if (r->tri_mode != prm->triangle_material) {
        if ((rc = build_shading(s, r, prm->triangle_material, rgb_lut)) != AP_OK) return rc;
    }
    if (r->max_tex >= 0 && texture_cache_count(s->tex_cache) <= static_cast<uint32_t>(r->max_tex))
        return fail(AP_ERR_INVALID_ARG, "ap_render: a material names a texture the scene's texture cache does not hold");
    const int32_t n_pix = cam->width * cam->height;
    // Frames per batch: up to 16M paths in flight (measured at 1080p with
    // four wavelengths: 6.43 ms per sample at 4M, 5.53 at 8M, 5.0 at 16M, as
    // the nearly empty tail bounces are shared by more frames; about 5 GB of
    // path state and queues). With guiding, one frame at
    // a time: each frame samples the probes built from the frames before it
    // (SceneRenderer::Render then UpdatePathGuidingSample, scene.cu:308-386).
    // With the radiance cache, one frame at a time too: each frame queries
    // the cache the frames before it trained.
    // With guiding in groups of guide_frames frames, a group is one batch
    // (within the same bound).
    const bool guiding = prm->use_guiding != 0;
    const bool nrc = cache != nullptr;
    if (prm->guide_frames < 0) return fail(AP_ERR_INVALID_ARG, "ap_render: guide_frames must be >= 0");
    const int32_t group = guiding ? std::max(1, prm->guide_frames) : 1;
    if (nrc && group > 1)
        return fail(AP_ERR_UNSUPPORTED, "ap_render_nrc: guiding groups of several frames are not supported with the cache");
    const int32_t max_batch = std::max(1, static_cast<int32_t>((16u << 20) / static_cast<uint32_t>(n_pix)));
    if (group > max_batch)
        return fail(AP_ERR_UNSUPPORTED, "ap_render: a guiding group exceeds 16M paths (guide_frames x pixels)");
    const int32_t per_batch = nrc ? 1 : guiding ? std::min(group, prm->spp) : std::min(prm->spp, max_batch);
    // Frames rendered one at a time are pipelined two deep (no training-only
    // passes of the cache, which reuse the first frame set).
    const bool pipelined = per_batch == 1 && prm->spp > 1 && (guiding || nrc) && !(nrc && np->train_passes > 0);
    if ((rc = ensure_paths(r, static_cast<size_t>(per_batch) * n_pix)) != AP_OK) return rc;
    if ((rc = ensure_side(r)) != AP_OK) return rc;
    if (pipelined && (rc = ensure_pipeline(r, static_cast<size_t>(n_pix))) != AP_OK) return rc;
    if (guiding && (rc = guide_prepare(&s->guide, static_cast<size_t>(per_batch) * n_pix, pipelined ? 2 : 1)) != AP_OK)
        return rc;
    if (fb->gbuffer && r->gsample_cap < static_cast<size_t>(per_batch) * n_pix) {
        if (r->gsamples) (void)hipFree(r->gsamples);
        r->gsamples = nullptr;
        r->gsample_cap = 0;
        if ((rc = dev_alloc(&r->gsamples, static_cast<size_t>(per_batch) * n_pix)) != AP_OK) return rc;
        r->gsample_cap = static_cast<size_t>(per_batch) * n_pix;
    }
    NrcState* ns = nrc ? nrc_state(cache) : nullptr;
    uint32_t n_stride = 1, n_train_max = 0;
    if (nrc) {
        if (np->train_paths > 0) {
            n_stride = std::max<uint32_t>(1u, static_cast<uint32_t>(n_pix) / np->train_paths);
            n_train_max = (static_cast<uint32_t>(n_pix) + n_stride - 1) / n_stride;
        }
        if ((rc = ensure_nrc(r, static_cast<size_t>(n_pix), n_train_max)) != AP_OK) return rc;
        // Cache bounds: the scene's root box grown on each side by 1e-3 of its largest extent.
        if (s->dev.root.valid) {
            float lo[3], hi[3], ext = 0.0f;
            for (int a = 0; a < 3; ++a) ext = std::max(ext, s->dev.root.bmax[a] - s->dev.root.bmin[a]);
            const float m = 1.0e-3f * std::max(ext, 1.0e-3f);
            for (int a = 0; a < 3; ++a) {
                lo[a] = s->dev.root.bmin[a] - m;
                hi[a] = s->dev.root.bmax[a] + m;
            }
            nrc_set_bounds(ns, lo, hi);
        }
    }

    // Camera basis of RenderCamera (scene.cu:220-228), direction normalised
    // as Camera::SetDirection does (camera.hpp:38-41).
    CamParams cp;
    cp.origin = {cam->origin[0], cam->origin[1], cam->origin[2]};
    cp.dir = normalize(f3{cam->direction[0], cam->direction[1], cam->direction[2]});
    const float fov = cam->fov_deg * 3.1415926f / 180.0f;
    const float aspect = static_cast<float>(cam->width) / fmaxf(1.0f, static_cast<float>(cam->height));
    cp.tan_y = tanf(0.5f * fov);
    cp.tan_x = cp.tan_y * aspect;
    cp.right = normalize(cross(cp.dir, f3{0.0f, 1.0f, 0.0f}));
    cp.up = normalize(cross(cp.dir, cp.right));
    cp.width = cam->width;
    cp.height = cam->height;
    const int32_t inside = camera_inside(s, r->analytic_objs, cp.origin) ? 1 : 0;
    const int32_t last_frame = prm->frame_index + prm->spp - 1;

    RunCtx c{s, r, prm, fb, stream, np, ns, rgb_lut, lam_lut, tables, n_pix, per_batch, guiding, nrc, pipelined,
             n_stride, n_train_max, cp, inside, last_frame};
    return prm->wavelengths == 4 ? run_frames<4>(c) : run_frames<1>(c);
}

// ---------------------------------------------------------------------------
// Batched multi-scene rendering (include/aperture/ap_batch.h).
// ---------------------------------------------------------------------------
struct BatchState {
    std::vector<ap_scene*> scenes;
    std::vector<ap_camera> cams;
    uint32_t total = 0;                 // paths of one batch frame (pixels of all scenes)
    std::vector<uint32_t> pix_begin;    // n + 1
    std::vector<BatchScene> h_scenes;
    std::vector<DeviceScene> h_devs;
    BatchScene* d_scenes = nullptr;
    DeviceScene* d_devs = nullptr;
    uint32_t* d_pix_begin = nullptr;
    DeviceScene shape{};                // launch shape, largest stack, traversal mode
    uint8_t* paths = nullptr;
    MediaRec* media = nullptr;
    ap_ray* rays[2] = {nullptr, nullptr};
    int32_t* path_of[2] = {nullptr, nullptr};
    ap_hit* hits = nullptr;
    ap_ray* shadow_rays = nullptr;
    uint8_t* shadow_payload = nullptr;
    ap_hit* shadow_hits = nullptr;
    uint16_t* ray_scene[2] = {nullptr, nullptr};
    uint16_t* shadow_scene = nullptr;
    uint32_t* counters = nullptr;
    int32_t* frame = nullptr;
    unsigned long long* rays_traced = nullptr;
    float* eps = nullptr;
    bool use_graph = true;
    // One instantiated frame graph per (render parameters, launch shape,
    // stream), kept until the batch is destroyed: the autotuner cycles over
    // launch shapes, and destroying an executable graph whose launches may
    // still be in flight (or still be reported by a profiler) is avoided.
    struct Graph {
        ap_render_params key;
        int32_t wave_rays, refill;
        hipStream_t stream;
        hipGraph_t graph;
        hipGraphExec_t exec;
        // Frames with per-scene caches (ap_batch_render_nrc): the cache, its
        // parameters and the buffers the frame names; null for plain frames.
        const NrcState* ns;
        ap_nrc_render_params np;
        uint64_t nrc_generation;
        uint32_t nrc_slots;
        size_t nrc_slot_cap;
        uint64_t guide_generation;  // guided frames: the guiding sizes the frame was captured with
    };
    // Renewed whenever a scene's guiding storage or sort size changes, which
    // the launches of a captured guided frame name.
    uint64_t guide_generation = 0;
    std::vector<Graph> graphs;
    static constexpr size_t kMaxGraphs = 16;
    hipStream_t last_stream = nullptr;  // stream of the last call (its frames may still run on it)
    bool stream_used = false;
    // Per-scene radiance caches (ap_batch_render_nrc).
    uint32_t nrc_slots = 0;                 // training slots of all scenes
    uint32_t nrc_table_paths = 0xffffffffu; // train_paths the table was built for
    std::vector<NrcBatchScene> h_nrc_table;
    NrcBatchScene* nrc_table = nullptr;
    NrcPath* nrc_paths = nullptr;
    float* nrc_queries = nullptr;
    int32_t* nrc_qpath = nullptr;           // scene s's query paths from pix_begin[s]
    float* nrc_qrgb = nullptr;
    uint32_t* nrc_counters = nullptr;       // [1] training samples
    uint32_t* nrc_qcount = nullptr;         // per scene: cache queries
    uint32_t* nrc_tile0 = nullptr;          // per scene: first 64-row tile of its query region
    uint32_t nrc_tiles = 0;
    uint32_t* nrc_pool_begin = nullptr;
    uint32_t* nrc_pool_n = nullptr;
    size_t nrc_slot_cap = 0;
    float* nrc_verts = nullptr;
    uint32_t* nrc_vmask = nullptr;
    uint32_t* nrc_offsets = nullptr;
    uint32_t* nrc_slot_pix = nullptr;
    float* nrc_pool = nullptr;
    uint64_t nrc_last_hash = 0;
    bool nrc_rendered = false;
    GuideRec* guide_recs = nullptr;     // guided batches: each path's collection record
    // Side stream of the overlapped shadow traces (as RenderState's).
    hipStream_t side = nullptr;
    hipEvent_t ev_shaded = nullptr, ev_shadowed = nullptr;
    // Guided batches: the scenes' probe builds run on kGuideStreams streams
    // (scene k on stream k mod kGuideStreams), forked from the render stream
    // after the frame's accumulation and joined before the next frame.
    static constexpr int kGuideStreams = 8;
    hipStream_t gstreams[kGuideStreams] = {};
    hipEvent_t gevents[kGuideStreams] = {};
    hipEvent_t ev_frame = nullptr;

    // Waits for the frames in flight, then destroys every graph.
    void drop_graphs() {
        if (!graphs.empty()) (void)hipDeviceSynchronize();
        for (Graph& g : graphs) {
            (void)hipGraphExecDestroy(g.exec);
            (void)hipGraphDestroy(g.graph);
        }
        graphs.clear();
    }
    ~BatchState() {
        if (stream_used) (void)hipDeviceSynchronize();
        drop_graphs();
        void* ptrs[] = {d_scenes, d_devs, d_pix_begin, paths, media, rays[0], rays[1], path_of[0], path_of[1], hits,
                        shadow_rays, shadow_payload, shadow_hits, ray_scene[0], ray_scene[1], shadow_scene, counters,
                        frame, rays_traced, eps, nrc_table, nrc_paths, nrc_queries, nrc_qpath, nrc_qrgb,
                        nrc_counters, nrc_qcount, nrc_tile0, nrc_pool_begin, nrc_pool_n,
                        nrc_verts, nrc_vmask, nrc_offsets, nrc_slot_pix, nrc_pool, guide_recs};
        for (void* p : ptrs)
            if (p) (void)hipFree(p);
        if (ev_shaded) (void)hipEventDestroy(ev_shaded);
        if (ev_shadowed) (void)hipEventDestroy(ev_shadowed);
        if (side) (void)hipStreamDestroy(side);
        for (int k = 0; k < kGuideStreams; ++k) {
            if (gevents[k]) (void)hipEventDestroy(gevents[k]);
            if (gstreams[k]) (void)hipStreamDestroy(gstreams[k]);
        }
        if (ev_frame) (void)hipEventDestroy(ev_frame);
    }
    BatchView view() const {
        return BatchView{d_scenes, d_pix_begin, static_cast<int32_t>(scenes.size()), {ray_scene[0], ray_scene[1]},
                         shadow_scene, frame};
    }
};

namespace {

// Shading records of every scene for the call's triangle-material mode and
// the batch records with this call's framebuffers, uploaded.
int batch_prepare(BatchState* b, int tri_mode, float* const* hdr, double* const* sums, const f4* rgb_lut,
                  hipStream_t stream, bool guiding = false) {
    int rc;
    // The scene records are read by the frames of the previous call, which
    // may still run on its stream: wait for them before rewriting the records.
    // (A different stream: the previous one may have been destroyed since, so
    // the whole device is waited for.)
    hipError_t e = hipSuccess;
    if (b->stream_used) {
        e = stream == b->last_stream ? hipStreamSynchronize(stream) : hipDeviceSynchronize();
        if (e != hipSuccess) return hip_fail(e, "ap_batch: wait for the previous call");
    }
    b->last_stream = stream;
    b->stream_used = true;
    if (!b->side) {
        e = hipEventCreateWithFlags(&b->ev_shaded, hipEventDisableTiming);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&b->ev_shadowed, hipEventDisableTiming);
        if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->side, hipStreamNonBlocking);
        if (e != hipSuccess) return hip_fail(e, "ap_batch: side stream");
    }
    for (size_t k = 0; k < b->scenes.size(); ++k) {
        ap_scene* s = b->scenes[k];
        if (!s->render) s->render = new RenderState();
        RenderState* r = s->render;
        if (r->tri_mode != tri_mode && (rc = build_shading(s, r, tri_mode, rgb_lut)) != AP_OK) return rc;
        BatchScene& B = b->h_scenes[k];
        B.geo = r->geo;
        B.mats = r->mats;
        B.lights = r->lights;
        B.analytic = s->dev.analytic;
        B.hdr = hdr ? hdr[k] : nullptr;
        B.sums = sums ? sums[k] : nullptr;
        B.n_lights = r->n_lights;
        const f3 o{b->cams[k].origin[0], b->cams[k].origin[1], b->cams[k].origin[2]};
        B.inside = camera_inside(s, r->analytic_objs, o) ? 1 : 0;
        B.volumes = s->dev.volumes;
        B.n_volumes = s->dev.n_volumes;
        if (r->max_tex >= 0 && texture_cache_count(s->tex_cache) <= static_cast<uint32_t>(r->max_tex))
            return fail(AP_ERR_INVALID_ARG, "ap_batch: a material names a texture the scene's texture cache does not hold");
        B.textures = texture_cache_table(s->tex_cache);
        B.uvs = r->uvs;
        B.guide = GuideView{};
        B.guide_slots = nullptr;
        B.guide_nvalid = nullptr;
        if (guiding) {
            if ((rc = guide_prepare(&s->guide, b->pix_begin[k + 1] - b->pix_begin[k])) != AP_OK) return rc;
            B.guide = guide_view(s->guide);
            B.guide_slots = guide_slots(s->guide);
            B.guide_nvalid = guide_nvalid(s->guide);
        }
        b->h_devs[k] = s->dev;
        if (s->dev.has_sdf) b->shape.has_sdf = 1;
    }
    if (guiding && !b->guide_recs && (rc = dev_alloc(&b->guide_recs, b->total)) != AP_OK) return rc;
    e = hipMemcpy(b->d_scenes, b->h_scenes.data(), b->h_scenes.size() * sizeof(BatchScene), hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(b->d_devs, b->h_devs.data(), b->h_devs.size() * sizeof(DeviceScene), hipMemcpyHostToDevice);
    return e == hipSuccess ? AP_OK : hip_fail(e, "ap_batch: upload scene records");
}

// After a guided batch frame: every scene's probe build from its own slots
// (ap_render's per-frame update; the scenes' first halves share one wait),
// then the scene records with the rebuilt probes and index volumes for the
// next frame.
// Refreshes the scene records whose probes, volume or bounds were
// reallocated (or whose first probes appeared).
int batch_guides_views(BatchState* b, hipStream_t stream) {
    bool changed = false;
    for (size_t k = 0; k < b->scenes.size(); ++k) {
        const GuideView v = guide_view(b->scenes[k]->guide);
        GuideView& o = b->h_scenes[k].guide;
        bool same = v.probes == o.probes && v.volume == o.volume && (v.probe_count > 0) == (o.probe_count > 0);
        for (int a = 0; a < 3; ++a) same = same && v.start[a] == o.start[a] && v.size[a] == o.size[a];
        if (!same) {
            o = v;
            changed = true;
        }
        b->h_scenes[k].guide_slots = guide_slots(b->scenes[k]->guide);
        b->h_scenes[k].guide_nvalid = guide_nvalid(b->scenes[k]->guide);
    }
    if (!changed) return AP_OK;
    hipError_t e = hipMemcpyAsync(b->d_scenes, b->h_scenes.data(), b->h_scenes.size() * sizeof(BatchScene),
                                  hipMemcpyHostToDevice, stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    return e == hipSuccess ? AP_OK : hip_fail(e, "ap_batch: guiding records");
}

// After a replayed guided frame and a wait for its stream: the scenes'
// builds that their device check left to the host are finished; `grew` is
// set when one was (the captured frame then names stale sizes).
int batch_guides_after_replay(BatchState* b, hipStream_t stream, bool* grew) {
    int rc;
    *grew = false;
    for (size_t k = 0; k < b->scenes.size(); ++k) {
        GuideDevice* g = b->scenes[k]->guide;
        guide_replayed(g);
        if (guide_needs_host(g)) *grew = true;
        if ((rc = guide_update_end(g, b->gstreams[k % BatchState::kGuideStreams], true)) != AP_OK) return rc;
    }
    for (int k = 0; k < BatchState::kGuideStreams; ++k) {
        const hipError_t e = b->gstreams[k] ? hipStreamSynchronize(b->gstreams[k]) : hipSuccess;
        if (e != hipSuccess) return hip_fail(e, "ap_batch: guiding finish");
    }
    return batch_guides_views(b, stream);
}

// Every scene's probe build after a guided batch frame, the builds spread
// over the guiding streams. With capture (the frame is being captured into
// a graph) the builds are only enqueued; batch_guides_after_replay finishes
// them after each replay.
int batch_guides_update(BatchState* b, hipStream_t stream, bool capture = false) {
    int rc;
    constexpr int K = BatchState::kGuideStreams;
    hipError_t e = hipSuccess;
    if (!b->ev_frame) {
        e = hipEventCreateWithFlags(&b->ev_frame, hipEventDisableTiming);
        for (int k = 0; k < K && e == hipSuccess; ++k) {
            e = hipStreamCreateWithFlags(&b->gstreams[k], hipStreamNonBlocking);
            if (e == hipSuccess) e = hipEventCreateWithFlags(&b->gevents[k], hipEventDisableTiming);
        }
        if (e != hipSuccess) return hip_fail(e, "ap_batch: guiding streams");
    }
    const size_t n = b->scenes.size();
    const int used = static_cast<int>(std::min<size_t>(n, K));
    // The scenes' builds are independent (each its own buffers and
    // counters): fork them over the guiding streams.
    e = hipEventRecord(b->ev_frame, stream);
    for (int k = 0; k < used && e == hipSuccess; ++k) e = hipStreamWaitEvent(b->gstreams[k], b->ev_frame, 0);
    if (e != hipSuccess) return hip_fail(e, "ap_batch: guiding fork");
    for (size_t k = 0; k < n; ++k)
        if ((rc = guide_update_begin(b->scenes[k]->guide, b->pix_begin[k + 1] - b->pix_begin[k],
                                     b->gstreams[k % K], capture)) != AP_OK)
            return rc;
    // Each end waits for its scene's check only; a scene whose volume or
    // probe storage had to grow is finished on the host there.
    if (!capture)
        for (size_t k = 0; k < n; ++k)
            if ((rc = guide_update_end(b->scenes[k]->guide, b->gstreams[k % K])) != AP_OK) return rc;
    for (int k = 0; k < used && e == hipSuccess; ++k) {
        e = hipEventRecord(b->gevents[k], b->gstreams[k]);
        if (e == hipSuccess) e = hipStreamWaitEvent(stream, b->gevents[k], 0);
    }
    if (e != hipSuccess) return hip_fail(e, "ap_batch: guiding join");
    if (capture) return AP_OK;
    return batch_guides_views(b, stream);
}

// One batch frame: every scene's paths through every bounce, then the
// running means and the frame counter. Nothing here synchronises, so the
// sequence can be captured into a graph.
template <int NL, bool GUIDE = false>
int batch_frame(BatchState* b, const ap_render_params* prm, hipStream_t stream, const f4* rgb_lut,
                const f4* lam_lut, const float* tables) {
    const BatchView bv = b->view();
    auto* paths = reinterpret_cast<PathT<NL>*>(b->paths);
    auto* pays = reinterpret_cast<ShadowPay<NL>*>(b->shadow_payload);
    const uint32_t total = b->total;
    GuideView gv{};
    if constexpr (GUIDE) gv.recs = b->guide_recs;
    hipLaunchKernelGGL(set_counters_kernel, dim3(1), dim3(64), 0, stream, b->counters, total, 0u, 0u);
    hipLaunchKernelGGL(generate_batch_kernel<NL>, grid_of(total), dim3(kBlock), 0, stream, bv, total, prm->seed,
                       b->eps, paths, b->rays[0], b->path_of[0], b->media, gv.recs);
    // Shadow rays of bounce k on the side stream, overlapped with bounce
    // k + 1's extension rays; joined before resolve (as in run_frames).
    int in_q = 0;
    hipLaunchKernelGGL(begin_bounce_kernel, dim3(1), dim3(64), 0, stream, b->counters, 1);
    hipError_t e = prm->max_depth > 0 ? launch_trace_multi(b->shape, b->d_devs, b->ray_scene[0], b->rays[0], total,
                                                           b->hits, false, stream, &b->counters[0])
                                      : hipSuccess;
    if (e != hipSuccess) return hip_fail(e, "ap_batch: extend");
    for (int32_t depth = 0; depth < prm->max_depth; ++depth) {
        const int out_q = 1 - in_q;
        const bool more = depth + 1 < prm->max_depth;
        ShadeParams sp{depth, prm->max_depth, prm->rr_min_depth, prm->nee_max_depth, prm->use_nee,
                       prm->nee_always, prm->sky, in_q, out_q, 0};
        sp.media = b->media;
        hipLaunchKernelGGL((shade_kernel<NL, GUIDE, false, true>), grid_of(total), dim3(kBlock), 0, stream, sp,
                           nullptr, nullptr, nullptr, nullptr, rgb_lut, lam_lut, tables, b->eps, paths,
                           b->rays[in_q], b->path_of[in_q], b->hits, b->rays[out_q], b->path_of[out_q],
                           b->shadow_rays, pays, b->counters, nullptr, 0, b->rays_traced, gv, NrcView{}, bv);
        e = hipEventRecord(b->ev_shaded, stream);
        if (e == hipSuccess) e = hipStreamWaitEvent(b->side, b->ev_shaded, 0);
        if (e == hipSuccess)
            e = launch_trace_multi(b->shape, b->d_devs, b->shadow_scene, b->shadow_rays, total, b->shadow_hits, true,
                                   b->side, &b->counters[2], sdf_eps_of_depth(depth));
        if (e == hipSuccess) {
            hipLaunchKernelGGL(resolve_kernel<NL>, grid_of(total), dim3(kBlock), 0, b->side, b->counters, pays,
                               b->shadow_hits, paths, b->rays_traced);
            if (more) hipLaunchKernelGGL(begin_bounce_kernel, dim3(1), dim3(64), 0, b->side, b->counters, in_q);
            e = hipEventRecord(b->ev_shadowed, b->side);
        }
        if (e != hipSuccess) return hip_fail(e, "ap_batch: shadow");
        if (more) {
            e = launch_trace_multi(b->shape, b->d_devs, b->ray_scene[out_q], b->rays[out_q], total, b->hits, false,
                                   stream, &b->counters[out_q]);
            if (e != hipSuccess) return hip_fail(e, "ap_batch: extend");
        }
        e = hipStreamWaitEvent(stream, b->ev_shadowed, 0);
        if (e != hipSuccess) return hip_fail(e, "ap_batch: join");
        in_q = out_q;
    }
    hipLaunchKernelGGL(accumulate_batch_kernel<NL>, grid_of(total), dim3(kBlock), 0, stream, bv, total, paths,
                       tables, gv.recs);
    hipLaunchKernelGGL(advance_frame_kernel, dim3(1), dim3(64), 0, stream, b->frame);
    e = hipGetLastError();
    return e == hipSuccess ? AP_OK : hip_fail(e, "ap_batch: launch");
}

bool same_key(const BatchState::Graph& g, const BatchState* b, const ap_render_params* p, hipStream_t st,
              const NrcState* ns = nullptr, const ap_nrc_render_params* np = nullptr) {
    const ap_render_params& k = g.key;
    if (g.ns != ns) return false;
    if (ns && (std::memcmp(&g.np, np, sizeof(*np)) != 0 || g.nrc_generation != nrc_generation(ns) ||
               g.nrc_slots != b->nrc_slots || g.nrc_slot_cap != b->nrc_slot_cap))
        return false;
    if ((k.use_guiding != 0) != (p->use_guiding != 0) || (p->use_guiding && g.guide_generation != b->guide_generation))
        return false;
    return g.stream == st && k.seed == p->seed && k.max_depth == p->max_depth &&
           k.rr_min_depth == p->rr_min_depth && k.nee_max_depth == p->nee_max_depth && k.use_nee == p->use_nee &&
           k.nee_always == p->nee_always && k.sky == p->sky && k.wavelengths == p->wavelengths &&
           g.wave_rays == b->shape.wave_rays && g.refill == b->shape.refill;
}

// spp batch frames from frame index frame0 (graph or direct launches).
int batch_run(BatchState* b, const ap_render_params* prm, int32_t frame0, int32_t frames, hipStream_t stream) {
    const f4 *rgb_lut = nullptr, *lam_lut = nullptr;
    const float* tables = nullptr;
    int rc = spectrum_get(&rgb_lut, &lam_lut, &tables);
    if (rc != AP_OK) return rc;
    hipLaunchKernelGGL(set_frame_kernel, dim3(1), dim3(64), 0, stream, b->frame, frame0);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "ap_batch: frame counter");
    const bool four = prm->wavelengths == 4;
    if (prm->use_guiding) {
        // One frame at a time, each sampling the probes the frames before it
        // built. A call's first frame runs directly (it sizes the builds'
        // buffers); the others replay one captured frame with its 48 probe
        // builds, then the host waits for it and finishes any build whose
        // device check found that storage must grow (after which the
        // captured frame is stale and is captured again).
        auto gframe = [&](bool capture) {
            int r2 = four ? batch_frame<4, true>(b, prm, stream, rgb_lut, lam_lut, tables)
                          : batch_frame<1, true>(b, prm, stream, rgb_lut, lam_lut, tables);
            return r2 == AP_OK ? batch_guides_update(b, stream, capture) : r2;
        };
        for (int32_t f = 0; f < frames; ++f) {
            if (f == 0 || !b->use_graph || stream == nullptr) {
                if ((rc = gframe(false)) != AP_OK) return rc;
                continue;
            }
            hipGraphExec_t gexec = nullptr;
            for (const BatchState::Graph& g : b->graphs)
                if (same_key(g, b, prm, stream)) gexec = g.exec;
            if (!gexec) {
                if (b->graphs.size() >= BatchState::kMaxGraphs) {
                    e = hipDeviceSynchronize();
                    if (e != hipSuccess) return hip_fail(e, "ap_batch: wait before dropping a graph");
                    (void)hipGraphExecDestroy(b->graphs.front().exec);
                    (void)hipGraphDestroy(b->graphs.front().graph);
                    b->graphs.erase(b->graphs.begin());
                }
                e = hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal);
                if (e != hipSuccess) return hip_fail(e, "ap_batch: begin capture");
                rc = gframe(true);
                hipGraph_t g = nullptr;
                e = hipStreamEndCapture(stream, &g);
                if (rc != AP_OK) {
                    if (g) (void)hipGraphDestroy(g);
                    return rc;
                }
                if (e != hipSuccess) return hip_fail(e, "ap_batch: end capture");
                e = hipGraphInstantiate(&gexec, g, nullptr, nullptr, 0);
                if (e != hipSuccess) {
                    (void)hipGraphDestroy(g);
                    return hip_fail(e, "ap_batch: instantiate graph");
                }
                b->graphs.push_back(BatchState::Graph{*prm, b->shape.wave_rays, b->shape.refill, stream, g, gexec,
                                                      nullptr, ap_nrc_render_params{}, 0ull, 0u, 0,
                                                      b->guide_generation});
            }
            e = hipGraphLaunch(gexec, stream);
            if (e == hipSuccess) e = hipStreamSynchronize(stream);
            if (e != hipSuccess) return hip_fail(e, "ap_batch: guided graph launch");
            bool grew = false;
            if ((rc = batch_guides_after_replay(b, stream, &grew)) != AP_OK) return rc;
            if (grew) ++b->guide_generation;
        }
        return AP_OK;
    }
    auto frame = [&]() {
        return four ? batch_frame<4>(b, prm, stream, rgb_lut, lam_lut, tables)
                    : batch_frame<1>(b, prm, stream, rgb_lut, lam_lut, tables);
    };
    if (!b->use_graph || stream == nullptr) {
        for (int32_t f = 0; f < frames; ++f)
            if ((rc = frame()) != AP_OK) return rc;
        return AP_OK;
    }
    hipGraphExec_t exec = nullptr;
    for (const BatchState::Graph& g : b->graphs)
        if (same_key(g, b, prm, stream)) exec = g.exec;
    if (!exec) {
        if (b->graphs.size() >= BatchState::kMaxGraphs) {
            // The oldest graph goes, once nothing that may use it is in flight.
            e = hipDeviceSynchronize();
            if (e != hipSuccess) return hip_fail(e, "ap_batch: wait before dropping a graph");
            (void)hipGraphExecDestroy(b->graphs.front().exec);
            (void)hipGraphDestroy(b->graphs.front().graph);
            b->graphs.erase(b->graphs.begin());
        }
        e = hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal);
        if (e != hipSuccess) return hip_fail(e, "ap_batch: begin capture");
        rc = frame();
        hipGraph_t g = nullptr;
        e = hipStreamEndCapture(stream, &g);
        if (rc != AP_OK) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess) return hip_fail(e, "ap_batch: end capture");
        e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
        if (e != hipSuccess) {
            (void)hipGraphDestroy(g);
            return hip_fail(e, "ap_batch: instantiate graph");
        }
        b->graphs.push_back(BatchState::Graph{*prm, b->shape.wave_rays, b->shape.refill, stream, g, exec, nullptr,
                                              ap_nrc_render_params{}, 0ull, 0u, 0, 0ull});
    }
    for (int32_t f = 0; f < frames; ++f) {
        e = hipGraphLaunch(exec, stream);
        if (e != hipSuccess) return hip_fail(e, "ap_batch: graph launch");
    }
    return AP_OK;
}

// Buffers, per-scene training layout and cache bounds of rendering a batch
// with per-scene caches.
int batch_nrc_prepare(BatchState* b, const ap_nrc_render_params* np, NrcState* ns) {
    const size_t n = b->scenes.size(), cap = b->total;
    int rc = AP_OK;
    if (!b->nrc_paths) {
        if ((rc = dev_alloc(&b->nrc_paths, cap)) != AP_OK) return rc;
        if ((rc = dev_alloc(&b->nrc_queries, cap * kNrcRec)) != AP_OK) return rc;
        if ((rc = dev_alloc(&b->nrc_qpath, cap)) != AP_OK) return rc;
        if ((rc = dev_alloc(&b->nrc_qrgb, cap * 3)) != AP_OK) return rc;
        if ((rc = dev_alloc(&b->nrc_counters, 4)) != AP_OK) return rc;
        for (uint32_t** p : {&b->nrc_qcount, &b->nrc_pool_begin, &b->nrc_pool_n})
            if ((rc = dev_alloc(p, n)) != AP_OK) return rc;
        if ((rc = dev_alloc(&b->nrc_table, n)) != AP_OK) return rc;
        // Query regions: scene k's rows pix_begin[k] .. pix_begin[k + 1], cut into 64-row tiles.
        std::vector<uint32_t> tile0(n + 1, 0);
        for (size_t k = 0; k < n; ++k) tile0[k + 1] = tile0[k] + (b->pix_begin[k + 1] - b->pix_begin[k] + 63u) / 64u;
        if ((rc = upload_vec(tile0, &b->nrc_tile0)) != AP_OK) return rc;
        b->nrc_tiles = tile0[n];
    }
    if (b->nrc_table_paths != np->train_paths) {
        // Per scene, ap_render_nrc's stride and most training paths; slots scene after scene.
        b->h_nrc_table.assign(n, NrcBatchScene{1u, 0u, 0u, 0u});
        uint64_t slots = 0;
        for (size_t k = 0; k < n; ++k) {
            NrcBatchScene& T = b->h_nrc_table[k];
            const uint32_t n_pix = b->pix_begin[k + 1] - b->pix_begin[k];
            if (np->train_paths > 0) {
                T.stride = std::max<uint32_t>(1u, n_pix / np->train_paths);
                T.n_train_max = (n_pix + T.stride - 1) / T.stride;
            }
            T.slot_base = static_cast<uint32_t>(slots);
            slots += T.n_train_max;
        }
        const hipError_t e = hipMemcpy(b->nrc_table, b->h_nrc_table.data(), n * sizeof(NrcBatchScene),
                                       hipMemcpyHostToDevice);
        if (e != hipSuccess) return hip_fail(e, "ap_batch_render_nrc: training layout");
        b->nrc_slots = static_cast<uint32_t>(slots);
        b->nrc_table_paths = np->train_paths;
    }
    const size_t slot_cap = std::max<size_t>(b->nrc_slots, 1);
    if (b->nrc_slot_cap < slot_cap) {
        void* old[] = {b->nrc_verts, b->nrc_vmask, b->nrc_offsets, b->nrc_slot_pix, b->nrc_pool};
        for (void* p : old)
            if (p) (void)hipFree(p);
        b->nrc_verts = nullptr; b->nrc_vmask = nullptr; b->nrc_offsets = nullptr; b->nrc_slot_pix = nullptr;
        b->nrc_pool = nullptr;
        b->nrc_slot_cap = 0;
        if ((rc = dev_alloc(&b->nrc_verts, slot_cap * kNrcV * kNrcVert)) != AP_OK) return rc;
        if ((rc = dev_alloc(&b->nrc_vmask, slot_cap)) != AP_OK) return rc;
        if ((rc = dev_alloc(&b->nrc_offsets, nrc_offsets_words(slot_cap))) != AP_OK) return rc;
        if ((rc = dev_alloc(&b->nrc_slot_pix, slot_cap)) != AP_OK) return rc;
        if ((rc = dev_alloc(&b->nrc_pool, slot_cap * kNrcV * kNrcSample)) != AP_OK) return rc;
        b->nrc_slot_cap = slot_cap;
    }
    // Network k's bounds: scene k's root box grown on each side by 1e-3 of its
    // largest extent (ap_render_nrc's).
    for (size_t k = 0; k < n; ++k) {
        const DeviceScene& d = b->scenes[k]->dev;
        if (!d.root.valid) continue;
        float lo[3], hi[3], ext = 0.0f;
        for (int a = 0; a < 3; ++a) ext = std::max(ext, d.root.bmax[a] - d.root.bmin[a]);
        const float m = 1.0e-3f * std::max(ext, 1.0e-3f);
        for (int a = 0; a < 3; ++a) {
            lo[a] = d.root.bmin[a] - m;
            hi[a] = d.root.bmax[a] + m;
        }
        if ((rc = nrc_set_group_bounds(ns, static_cast<uint32_t>(k), lo, hi)) != AP_OK) return rc;
    }
    return AP_OK;
}

// One batch frame with per-scene caches: batch_frame's bounces with paths
// ending into their scene's network, the grouped inference of all networks,
// the scenes' training samples, the running means, the networks' Adam steps
// on their own scenes' samples (one set of launches for all), then the frame
// counter. Everything that changes from frame to frame (the frame hash, the
// training draws, Adam's bias corrections) is read on the device, so the
// sequence can be captured once and replayed.
template <int NL, bool GUIDE = false>
int batch_frame_nrc(BatchState* b, const ap_render_params* prm, const ap_nrc_render_params* np, NrcState* ns,
                    hipStream_t stream, const f4* rgb_lut, const f4* lam_lut, const float* tables) {
    const BatchView bv = b->view();
    auto* paths = reinterpret_cast<PathT<NL>*>(b->paths);
    auto* pays = reinterpret_cast<ShadowPay<NL>*>(b->shadow_payload);
    const uint32_t total = b->total;
    const int32_t n = static_cast<int32_t>(b->scenes.size());
    // Guided: each path samples its scene's probes (the scene records carry
    // them) and collects its guiding samples in the batch's records.
    GuideView gv{};
    if constexpr (GUIDE) gv.recs = b->guide_recs;
    hipLaunchKernelGGL(set_counters_kernel, dim3(1), dim3(64), 0, stream, b->counters, total, 0u, 0u);
    hipLaunchKernelGGL(generate_batch_kernel<NL>, grid_of(total), dim3(kBlock), 0, stream, bv, total, prm->seed,
                       b->eps, paths, b->rays[0], b->path_of[0], b->media, gv.recs);
    hipLaunchKernelGGL(nrc_init_batch_kernel, grid_of(total), dim3(kBlock), 0, stream, bv, total, b->nrc_table,
                       prm->seed, np->unbiased_every, b->nrc_paths, b->nrc_counters, b->nrc_qcount);
    const NrcView nv{b->nrc_paths, b->nrc_queries, b->nrc_qpath, b->nrc_verts, b->nrc_counters, np->spread,
                     b->nrc_qcount};
    // Shadow rays of bounce k on the side stream, overlapped with bounce
    // k + 1's extension rays; joined before resolve (as in run_frames).
    int in_q = 0;
    hipLaunchKernelGGL(begin_bounce_kernel, dim3(1), dim3(64), 0, stream, b->counters, 1);
    hipError_t e = prm->max_depth > 0 ? launch_trace_multi(b->shape, b->d_devs, b->ray_scene[0], b->rays[0], total,
                                                           b->hits, false, stream, &b->counters[0])
                                      : hipSuccess;
    if (e != hipSuccess) return hip_fail(e, "ap_batch: extend");
    for (int32_t depth = 0; depth < prm->max_depth; ++depth) {
        const int out_q = 1 - in_q;
        const bool more = depth + 1 < prm->max_depth;
        ShadeParams sp{depth, prm->max_depth, prm->rr_min_depth, prm->nee_max_depth, prm->use_nee,
                       prm->nee_always, prm->sky, in_q, out_q, 0};
        sp.media = b->media;
        hipLaunchKernelGGL((shade_kernel<NL, GUIDE, true, true>), grid_of(total), dim3(kBlock), 0, stream, sp,
                           nullptr, nullptr, nullptr, nullptr, rgb_lut, lam_lut, tables, b->eps, paths,
                           b->rays[in_q], b->path_of[in_q], b->hits, b->rays[out_q], b->path_of[out_q],
                           b->shadow_rays, pays, b->counters, nullptr, 0, b->rays_traced, gv, nv, bv);
        e = hipEventRecord(b->ev_shaded, stream);
        if (e == hipSuccess) e = hipStreamWaitEvent(b->side, b->ev_shaded, 0);
        if (e == hipSuccess)
            e = launch_trace_multi(b->shape, b->d_devs, b->shadow_scene, b->shadow_rays, total, b->shadow_hits, true,
                                   b->side, &b->counters[2], sdf_eps_of_depth(depth));
        if (e == hipSuccess) {
            hipLaunchKernelGGL(resolve_kernel<NL>, grid_of(total), dim3(kBlock), 0, b->side, b->counters, pays,
                               b->shadow_hits, paths, b->rays_traced);
            if (more) hipLaunchKernelGGL(begin_bounce_kernel, dim3(1), dim3(64), 0, b->side, b->counters, in_q);
            e = hipEventRecord(b->ev_shadowed, b->side);
        }
        if (e != hipSuccess) return hip_fail(e, "ap_batch: shadow");
        if (more) {
            e = launch_trace_multi(b->shape, b->d_devs, b->ray_scene[out_q], b->rays[out_q], total, b->hits, false,
                                   stream, &b->counters[out_q]);
            if (e != hipSuccess) return hip_fail(e, "ap_batch: extend");
        }
        e = hipStreamWaitEvent(stream, b->ev_shadowed, 0);
        if (e != hipSuccess) return hip_fail(e, "ap_batch: join");
        in_q = out_q;
    }
    // Cache vertices: every scene's queries to its own network, one inference for all.
    int rc = nrc_infer_groups(ns, b->nrc_queries, b->nrc_qpath, b->d_pix_begin, b->nrc_tile0, b->nrc_qcount,
                              b->nrc_tiles * 64u, b->nrc_qrgb, stream);
    if (rc != AP_OK) return rc;
    hipLaunchKernelGGL(nrc_apply_batch_kernel<NL>, grid_of(total), dim3(kBlock), 0, stream, bv, total,
                       b->nrc_qcount, b->nrc_qpath, b->nrc_qrgb, rgb_lut, lam_lut, tables, paths, gv.recs);
    const uint32_t slots = b->nrc_slots;
    if (slots > 0) {
        hipLaunchKernelGGL(nrc_targets_batch_kernel<NL>, grid_of(slots), dim3(kBlock), 0, stream, slots, bv,
                           b->nrc_table, prm->seed, b->nrc_paths, paths, tables, b->nrc_verts, b->nrc_vmask,
                           b->nrc_slot_pix);
        if ((e = nrc_scan(slots, b->nrc_vmask, b->nrc_offsets, b->nrc_slot_cap, b->nrc_counters, stream)) !=
            hipSuccess)
            return hip_fail(e, "ap_batch: training sample scan");
        hipLaunchKernelGGL(nrc_pool_ranges_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, n, slots,
                           b->nrc_table, b->nrc_offsets, b->nrc_counters, b->nrc_pool_begin, b->nrc_pool_n);
        hipLaunchKernelGGL(nrc_compact_kernel, grid_of(slots), dim3(kBlock), 0, stream, slots, 1u, 0u, b->nrc_vmask,
                           b->nrc_offsets, b->nrc_verts, b->nrc_pool, b->nrc_slot_pix);
    }
    hipLaunchKernelGGL(accumulate_batch_kernel<NL>, grid_of(total), dim3(kBlock), 0, stream, bv, total, paths,
                       tables, gv.recs);
    e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "ap_batch: launch");
    if (slots > 0 && np->train_steps > 0 &&
        (rc = nrc_train_pool(ns, b->nrc_pool, b->nrc_pool_begin, b->nrc_pool_n, np->batch, np->train_steps,
                             np->learning_rate, 0, stream, b->frame, prm->seed)) != AP_OK)
        return rc;
    hipLaunchKernelGGL(advance_frame_kernel, dim3(1), dim3(64), 0, stream, b->frame);
    e = hipGetLastError();
    return e == hipSuccess ? AP_OK : hip_fail(e, "ap_batch: launch");
}

int batch_check_params(const ap_render_params* prm) {
    if (prm->spp < 0 || prm->frame_index < 0 || prm->max_depth < 0 || prm->max_depth > kEpsEntries - 2)
        return fail(AP_ERR_INVALID_ARG, "ap_batch: bad spp, frame_index or max_depth");
    if (prm->triangle_material != AP_TRIANGLE_MATERIAL_OBJECT && prm->triangle_material != AP_TRIANGLE_MATERIAL_REFERENCE)
        return fail(AP_ERR_INVALID_ARG, "ap_batch: bad triangle_material");
    if (prm->wavelengths != 0 && prm->wavelengths != 1 && prm->wavelengths != 4)
        return fail(AP_ERR_INVALID_ARG, "ap_batch: wavelengths must be 1 or 4 (0 reads as 1)");
    if (prm->use_guiding && prm->guide_frames > 1)
        return fail(AP_ERR_UNSUPPORTED, "ap_batch: guided batches update the probes after every frame (guide_frames 1)");
    return AP_OK;
}

}  // namespace

}  // namespace ap

using namespace ap;

struct ap_batch {
    BatchState s;
};

extern "C" {

void ap_render_params_default(ap_render_params* p) {
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    p->spp = 1;
    p->frame_index = 0;
    p->max_depth = 1000;
    p->rr_min_depth = 3;
    p->nee_max_depth = 10;
    p->use_nee = 1;
    p->nee_always = 0;
    p->sky = 1;
    p->triangle_material = AP_TRIANGLE_MATERIAL_OBJECT;
    p->use_guiding = 0;
    p->seed = 0;
    p->wavelengths = 1;
    p->guide_frames = 1;
}

int ap_render(ap_scene* scene, const ap_camera* camera, const ap_render_params* params, const ap_framebuffers* fb,
              void* stream) {
    return render_frames(scene, camera, params, fb, static_cast<hipStream_t>(stream));
}

int ap_render_offline(ap_scene* scene, const ap_camera* camera, const ap_render_params* params, float* host_xyz) {
    if (!scene || !camera || !params || !host_xyz) return fail(AP_ERR_INVALID_ARG, "ap_render_offline: NULL argument");
    if (camera->width <= 0 || camera->height <= 0) return fail(AP_ERR_INVALID_ARG, "ap_render_offline: bad resolution");
    const size_t n = static_cast<size_t>(camera->width) * camera->height * 3;
    float* target = nullptr;
    hipError_t e = hipMalloc(&target, n * sizeof(float));
    if (e != hipSuccess) return hip_fail(e, "hipMalloc(offline target)");
    e = hipMemset(target, 0, n * sizeof(float));
    int rc = e == hipSuccess ? AP_OK : hip_fail(e, "hipMemset");
    if (rc == AP_OK) {
        ap_render_params p = *params;
        p.frame_index = 0;
        ap_framebuffers fb{};
        fb.hdr = target;
        rc = render_frames(scene, camera, &p, &fb, nullptr);
    }
    if (rc == AP_OK) {
        e = hipMemcpy(host_xyz, target, n * sizeof(float), hipMemcpyDeviceToHost);
        if (e != hipSuccess) rc = hip_fail(e, "hipMemcpy(offline)");
    }
    (void)hipFree(target);
    return rc;
}

void ap_nrc_render_params_default(ap_nrc_render_params* p) {
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    p->spread = 0.01f;
    p->train_paths = 32768;
    p->train_steps = 4;
    p->batch = 16384;
    p->learning_rate = 0.005f;
    p->unbiased_every = 16;
    p->train_passes = 0;
}

int ap_render_nrc(ap_scene* scene, const ap_camera* camera, const ap_render_params* params, ap_nrc* cache,
                  const ap_nrc_render_params* nrc_params, const ap_framebuffers* fb, void* stream) {
    if (!cache || !nrc_params) return fail(AP_ERR_INVALID_ARG, "ap_render_nrc: NULL cache or parameters");
    if (nrc_groups(nrc_state(cache)) != 1)
        return fail(AP_ERR_INVALID_ARG, "ap_render_nrc: the cache must hold one network");
    if (!(nrc_params->spread > 0.0f) || !(nrc_params->learning_rate >= 0.0f))
        return fail(AP_ERR_INVALID_ARG, "ap_render_nrc: spread must be > 0 and learning_rate >= 0");
    if (nrc_params->train_steps > 0 &&
        (nrc_params->batch == 0 || nrc_params->batch % 256 != 0 || nrc_params->batch > (1u << 24)))
        return fail(AP_ERR_INVALID_ARG, "ap_render_nrc: batch must be a positive multiple of 256 up to 2^24");
    return render_frames(scene, camera, params, fb, static_cast<hipStream_t>(stream), cache, nrc_params);
}

int ap_nrc_last_frame(ap_scene* scene, ap_nrc_frame_stats* stats, float* host_samples, size_t capacity) {
    if (!scene || !stats) return fail(AP_ERR_INVALID_ARG, "ap_nrc_last_frame: NULL argument");
    std::memset(stats, 0, sizeof(*stats));
    RenderState* r = scene->render;
    if (!r || !r->nrc_rendered) return fail(AP_ERR_INVALID_ARG, "ap_nrc_last_frame: no ap_render_nrc frame yet");
    uint32_t c[2] = {0, 0};
    hipError_t e = hipDeviceSynchronize();
    if (e == hipSuccess) e = hipMemcpy(c, r->nrc_counters + 4 * r->nrc_last_set, sizeof(c), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(e, "ap_nrc_last_frame");
    stats->queries = c[0];
    stats->train_paths = r->nrc_last_train;
    stats->samples = r->nrc_last_train > 0 ? c[1] : 0u;
    if (host_samples && stats->samples > 0 && capacity > 0) {
        const size_t n = std::min<size_t>(capacity, stats->samples);
        e = hipMemcpy(host_samples, r->nrc_pool, n * kNrcSample * sizeof(float), hipMemcpyDeviceToHost);
        if (e != hipSuccess) return hip_fail(e, "ap_nrc_last_frame: samples");
    }
    return AP_OK;
}

int ap_batch_create(ap_scene* const* scenes, const ap_camera* cameras, size_t n, ap_batch** out) {
    if (!out) return fail(AP_ERR_INVALID_ARG, "ap_batch_create: NULL output");
    *out = nullptr;
    if (!scenes || !cameras || n == 0 || n > 65535)
        return fail(AP_ERR_INVALID_ARG, "ap_batch_create: need 1..65535 scenes and cameras");
    auto* b = new ap_batch();
    BatchState& B = b->s;
    B.pix_begin.assign(1, 0u);
    uint64_t total = 0;
    int stack_cap = 1;
    bool all_culled = true;
    for (size_t k = 0; k < n; ++k) {
        const ap_camera& c = cameras[k];
        if (!scenes[k] || c.width <= 0 || c.height <= 0) {
            delete b;
            return fail(AP_ERR_INVALID_ARG, "ap_batch_create: NULL scene or bad resolution (scene " + std::to_string(k) + ")");
        }
        B.scenes.push_back(scenes[k]);
        B.cams.push_back(c);
        total += static_cast<uint64_t>(c.width) * static_cast<uint64_t>(c.height);
        B.pix_begin.push_back(static_cast<uint32_t>(std::min<uint64_t>(total, 0xffffffffull)));
        stack_cap = std::max(stack_cap, scenes[k]->dev.stack_cap);
        all_culled = all_culled && scenes[k]->dev.mode == AP_TRAVERSE_CULLED;
    }
    if (total > (uint64_t{1} << 31)) {
        delete b;
        return fail(AP_ERR_INVALID_ARG, "ap_batch_create: more than 2^31 pixels per batch frame");
    }
    B.total = static_cast<uint32_t>(total);
    B.h_scenes.resize(n);
    B.h_devs.resize(n);
    for (size_t k = 0; k < n; ++k) {
        // Camera basis of RenderCamera (scene.cu:220-228), as render_frames builds it.
        const ap_camera& cam = B.cams[k];
        BatchScene& S = B.h_scenes[k];
        std::memset(&S, 0, sizeof(S));
        const f3 dir = normalize(f3{cam.direction[0], cam.direction[1], cam.direction[2]});
        const float fov = cam.fov_deg * 3.1415926f / 180.0f;
        const float aspect = static_cast<float>(cam.width) / fmaxf(1.0f, static_cast<float>(cam.height));
        const float tan_y = tanf(0.5f * fov);
        const f3 right = normalize(cross(dir, f3{0.0f, 1.0f, 0.0f}));
        const f3 up = normalize(cross(dir, right));
        const float c[14] = {cam.origin[0], cam.origin[1], cam.origin[2], dir.x, dir.y, dir.z, right.x, right.y,
                             right.z, up.x, up.y, up.z, tan_y * aspect, tan_y};
        std::memcpy(S.cam, c, sizeof(c));
        S.width = cam.width;
        S.height = cam.height;
        S.pix_begin = B.pix_begin[k];
        S.n_pix = B.pix_begin[k + 1] - B.pix_begin[k];
    }
    B.shape = scenes[0]->dev;
    B.shape.wave_rays = 512;
    B.shape.refill = 16;
    B.shape.stack_cap = stack_cap;
    B.shape.mode = all_culled ? AP_TRAVERSE_CULLED : AP_TRAVERSE_REFERENCE;
    const size_t cap = B.total;
    int rc = AP_OK;
    auto alloc = [&](auto** p, size_t bytes) {
        if (rc != AP_OK) return;
        const hipError_t e = hipMalloc(reinterpret_cast<void**>(p), std::max<size_t>(bytes, 16));
        if (e != hipSuccess) rc = hip_fail(e, "ap_batch_create: hipMalloc");
    };
    alloc(&B.d_scenes, n * sizeof(BatchScene));
    alloc(&B.d_devs, n * sizeof(DeviceScene));
    alloc(&B.d_pix_begin, (n + 1) * sizeof(uint32_t));
    alloc(&B.paths, cap * kPathBytes);
    alloc(&B.media, cap * sizeof(MediaRec));
    for (int q = 0; q < 2; ++q) {
        alloc(&B.rays[q], cap * sizeof(ap_ray));
        alloc(&B.path_of[q], cap * sizeof(int32_t));
        alloc(&B.ray_scene[q], cap * sizeof(uint16_t));
    }
    alloc(&B.hits, cap * sizeof(ap_hit));
    alloc(&B.shadow_rays, cap * sizeof(ap_ray));
    alloc(&B.shadow_payload, cap * kShadowPayBytes);
    alloc(&B.shadow_hits, cap * sizeof(ap_hit));
    alloc(&B.shadow_scene, cap * sizeof(uint16_t));
    alloc(&B.counters, 4 * sizeof(uint32_t));
    alloc(&B.frame, sizeof(int32_t));
    alloc(&B.rays_traced, 2 * sizeof(unsigned long long));
    alloc(&B.eps, kEpsEntries * sizeof(float));
    if (rc == AP_OK) {
        std::vector<float> e(kEpsEntries);
        for (int d = 0; d < kEpsEntries; ++d) e[d] = 1.0e-3f * std::exp(fminf(static_cast<float>(d) * 0.5f, 4.0f));
        hipError_t he = hipMemcpy(B.eps, e.data(), e.size() * sizeof(float), hipMemcpyHostToDevice);
        if (he == hipSuccess)
            he = hipMemcpy(B.d_pix_begin, B.pix_begin.data(), B.pix_begin.size() * sizeof(uint32_t),
                           hipMemcpyHostToDevice);
        if (he != hipSuccess) rc = hip_fail(he, "ap_batch_create: upload");
    }
    if (rc != AP_OK) {
        delete b;
        return rc;
    }
    *out = b;
    return AP_OK;
}

int ap_batch_destroy(ap_batch* batch) {
    delete batch;
    return AP_OK;
}

int ap_batch_render(ap_batch* batch, const ap_render_params* params, float* const* hdr, double* const* sums,
                    uint64_t ray_count[2], void* stream) {
    if (!batch || !params) return fail(AP_ERR_INVALID_ARG, "ap_batch_render: NULL argument");
    int rc = batch_check_params(params);
    if (rc != AP_OK) return rc;
    if (params->spp == 0 || params->max_depth == 0) return AP_OK;
    BatchState& B = batch->s;
    const f4 *rgb_lut = nullptr, *lam_lut = nullptr;
    const float* tables = nullptr;
    if ((rc = spectrum_get(&rgb_lut, &lam_lut, &tables)) != AP_OK) return rc;
    const hipStream_t st = static_cast<hipStream_t>(stream);
    if ((rc = batch_prepare(&B, params->triangle_material, hdr, sums, rgb_lut, st, params->use_guiding != 0)) != AP_OK)
        return rc;
    hipError_t e = hipMemsetAsync(B.rays_traced, 0, 2 * sizeof(unsigned long long), st);
    if (e != hipSuccess) return hip_fail(e, "ap_batch_render: reset counters");
    if ((rc = batch_run(&B, params, params->frame_index, params->spp, st)) != AP_OK) return rc;
    if (ray_count) {
        unsigned long long c[2] = {0, 0};
        e = hipMemcpyAsync(c, B.rays_traced, sizeof(c), hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) return hip_fail(e, "ap_batch_render: ray counts");
        ray_count[0] = c[0];
        ray_count[1] = c[1];
    }
    return AP_OK;
}

int ap_batch_render_nrc(ap_batch* batch, const ap_render_params* params, ap_nrc* cache,
                        const ap_nrc_render_params* nrc_params, float* const* hdr, double* const* sums,
                        uint64_t ray_count[2], void* stream) {
    if (!batch || !params || !cache || !nrc_params) return fail(AP_ERR_INVALID_ARG, "ap_batch_render_nrc: NULL argument");
    int rc = batch_check_params(params);
    if (rc != AP_OK) return rc;
    BatchState& B = batch->s;
    NrcState* ns = nrc_state(cache);
    if (nrc_groups(ns) != B.scenes.size())
        return fail(AP_ERR_INVALID_ARG, "ap_batch_render_nrc: the cache must hold one network per scene");
    if (!(nrc_params->spread > 0.0f) || !(nrc_params->learning_rate >= 0.0f))
        return fail(AP_ERR_INVALID_ARG, "ap_batch_render_nrc: spread must be > 0 and learning_rate >= 0");
    if (nrc_params->train_passes != 0)
        return fail(AP_ERR_UNSUPPORTED, "ap_batch_render_nrc: training-only passes are not supported in a batch");
    if (nrc_params->train_steps > 0 &&
        (nrc_params->batch == 0 || nrc_params->batch % 256 != 0 ||
         static_cast<uint64_t>(nrc_params->batch) * B.scenes.size() > (1u << 24)))
        return fail(AP_ERR_INVALID_ARG,
                    "ap_batch_render_nrc: batch must be a positive multiple of 256, at most 2^24 over all scenes");
    if (params->spp == 0 || params->max_depth == 0) return AP_OK;
    const f4 *rgb_lut = nullptr, *lam_lut = nullptr;
    const float* tables = nullptr;
    if ((rc = spectrum_get(&rgb_lut, &lam_lut, &tables)) != AP_OK) return rc;
    const hipStream_t st = static_cast<hipStream_t>(stream);
    const bool guiding = params->use_guiding != 0;
    if ((rc = batch_prepare(&B, params->triangle_material, hdr, sums, rgb_lut, st, guiding)) != AP_OK) return rc;
    if ((rc = batch_nrc_prepare(&B, nrc_params, ns)) != AP_OK) return rc;
    hipError_t e = hipMemsetAsync(B.rays_traced, 0, 2 * sizeof(unsigned long long), st);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(set_frame_kernel, dim3(1), dim3(64), 0, st, B.frame, params->frame_index);
        e = hipGetLastError();
    }
    if (e != hipSuccess) return hip_fail(e, "ap_batch_render_nrc: reset counters");
    // Frame after frame, each rendered with the networks as the frames before
    // it trained them. The call's first frame runs directly (it sizes every
    // buffer it needs); the others replay one captured frame, cached per
    // render parameters, cache parameters, buffers, launch shape and stream.
    auto frame = [&]() {
        const bool four = params->wavelengths == 4;
        if (guiding)
            return four ? batch_frame_nrc<4, true>(&B, params, nrc_params, ns, st, rgb_lut, lam_lut, tables)
                        : batch_frame_nrc<1, true>(&B, params, nrc_params, ns, st, rgb_lut, lam_lut, tables);
        return four ? batch_frame_nrc<4>(&B, params, nrc_params, ns, st, rgb_lut, lam_lut, tables)
                    : batch_frame_nrc<1>(&B, params, nrc_params, ns, st, rgb_lut, lam_lut, tables);
    };
    const int32_t steps = (B.nrc_slots > 0) ? static_cast<int32_t>(nrc_params->train_steps) : 0;
    hipGraphExec_t exec = nullptr;
    for (int32_t f = 0; f < params->spp; ++f) {
        if (guiding) {
            // Guided: every frame directly, each scene's probes rebuilt from
            // its own samples after the frame.
            if ((rc = frame()) != AP_OK || (rc = batch_guides_update(&B, st)) != AP_OK) return rc;
            continue;
        }
        if (f == 0 || !B.use_graph || st == nullptr) {
            if ((rc = frame()) != AP_OK) return rc;
            continue;
        }
        if (!exec)
            for (const BatchState::Graph& g : B.graphs)
                if (same_key(g, &B, params, st, ns, nrc_params)) exec = g.exec;
        if (!exec) {
            if (B.graphs.size() >= BatchState::kMaxGraphs) {
                e = hipDeviceSynchronize();
                if (e != hipSuccess) return hip_fail(e, "ap_batch_render_nrc: wait before dropping a graph");
                (void)hipGraphExecDestroy(B.graphs.front().exec);
                (void)hipGraphDestroy(B.graphs.front().graph);
                B.graphs.erase(B.graphs.begin());
            }
            e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
            if (e != hipSuccess) return hip_fail(e, "ap_batch_render_nrc: begin capture");
            rc = frame();
            nrc_count_steps(ns, -steps);  // captured, not run
            hipGraph_t g = nullptr;
            e = hipStreamEndCapture(st, &g);
            if (rc != AP_OK) {
                if (g) (void)hipGraphDestroy(g);
                return rc;
            }
            if (e != hipSuccess) return hip_fail(e, "ap_batch_render_nrc: end capture");
            e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
            if (e != hipSuccess) {
                (void)hipGraphDestroy(g);
                return hip_fail(e, "ap_batch_render_nrc: instantiate graph");
            }
            B.graphs.push_back(BatchState::Graph{*params, B.shape.wave_rays, B.shape.refill, st, g, exec, ns,
                                                 *nrc_params, nrc_generation(ns), B.nrc_slots, B.nrc_slot_cap, 0ull});
        }
        e = hipGraphLaunch(exec, st);
        if (e != hipSuccess) return hip_fail(e, "ap_batch_render_nrc: graph launch");
        nrc_count_steps(ns, steps);
    }
    B.nrc_last_hash = nrc_frame_hash(params->seed, params->frame_index + params->spp - 1);
    B.nrc_rendered = true;
    if (ray_count) {
        unsigned long long c[2] = {0, 0};
        e = hipMemcpyAsync(c, B.rays_traced, sizeof(c), hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) return hip_fail(e, "ap_batch_render_nrc: ray counts");
        ray_count[0] = c[0];
        ray_count[1] = c[1];
    }
    return AP_OK;
}

int ap_batch_nrc_last_frame(ap_batch* batch, ap_nrc_frame_stats* per_scene, size_t n) {
    if (!batch || !per_scene) return fail(AP_ERR_INVALID_ARG, "ap_batch_nrc_last_frame: NULL argument");
    BatchState& B = batch->s;
    if (!B.nrc_rendered) return fail(AP_ERR_INVALID_ARG, "ap_batch_nrc_last_frame: no ap_batch_render_nrc frame yet");
    if (n != B.scenes.size()) return fail(AP_ERR_INVALID_ARG, "ap_batch_nrc_last_frame: one entry per scene");
    std::vector<uint32_t> q(n), pn(n, 0);
    hipError_t e = hipDeviceSynchronize();
    if (e == hipSuccess) e = hipMemcpy(q.data(), B.nrc_qcount, n * sizeof(uint32_t), hipMemcpyDeviceToHost);
    if (e == hipSuccess && B.nrc_slots > 0)
        e = hipMemcpy(pn.data(), B.nrc_pool_n, n * sizeof(uint32_t), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(e, "ap_batch_nrc_last_frame");
    for (size_t k = 0; k < n; ++k) {
        const NrcBatchScene& T = B.h_nrc_table[k];
        const uint32_t n_pix = B.pix_begin[k + 1] - B.pix_begin[k];
        uint32_t n_train = 0;
        if (T.n_train_max > 0) {
            const uint32_t offset = static_cast<uint32_t>(B.nrc_last_hash % T.stride);
            n_train = std::min(T.n_train_max, (n_pix - offset + T.stride - 1) / T.stride);
        }
        per_scene[k] = ap_nrc_frame_stats{q[k], n_train, pn[k], 0u};
    }
    return AP_OK;
}

int ap_batch_set_launch(ap_batch* batch, int wave_rays, int refill) {
    if (!batch) return fail(AP_ERR_INVALID_ARG, "ap_batch_set_launch: NULL batch");
    if (wave_rays < 64 || wave_rays % 64 != 0 || wave_rays > (1 << 16) || refill < 1 || refill > 64)
        return fail(AP_ERR_INVALID_ARG, "ap_batch_set_launch: wave_rays must be a multiple of 64 up to 65536, refill 1..64");
    batch->s.shape.wave_rays = wave_rays;
    batch->s.shape.refill = refill;
    return AP_OK;
}

int ap_batch_get_launch(const ap_batch* batch, int* wave_rays, int* refill) {
    if (!batch || !wave_rays || !refill) return fail(AP_ERR_INVALID_ARG, "ap_batch_get_launch: NULL argument");
    *wave_rays = batch->s.shape.wave_rays;
    *refill = batch->s.shape.refill;
    return AP_OK;
}

int ap_batch_set_graph(ap_batch* batch, int enable) {
    if (!batch) return fail(AP_ERR_INVALID_ARG, "ap_batch_set_graph: NULL batch");
    batch->s.use_graph = enable != 0;
    if (!batch->s.use_graph) batch->s.drop_graphs();
    return AP_OK;
}

int ap_batch_autotune(ap_batch* batch, const ap_render_params* params, int frames, int* best_wave_rays,
                      int* best_refill, float* best_ms_per_frame, void* stream) {
    if (!batch || !params || frames <= 0) return fail(AP_ERR_INVALID_ARG, "ap_batch_autotune: bad argument");
    int rc = batch_check_params(params);
    if (rc != AP_OK) return rc;
    // Launch shapes are timed on unguided frames (probes stay untouched).
    ap_render_params unguided = *params;
    unguided.use_guiding = 0;
    params = &unguided;
    BatchState& B = batch->s;
    const f4 *rgb_lut = nullptr, *lam_lut = nullptr;
    const float* tables = nullptr;
    if ((rc = spectrum_get(&rgb_lut, &lam_lut, &tables)) != AP_OK) return rc;
    const hipStream_t st = static_cast<hipStream_t>(stream);
    if ((rc = batch_prepare(&B, params->triangle_material, nullptr, nullptr, rgb_lut, st)) != AP_OK) return rc;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e != hipSuccess) return hip_fail(e, "ap_batch_autotune: events");
    const int cand_w[3] = {128, 256, 512};
    const int cand_r[2] = {8, 16};
    float best = -1.0f;
    int bw = B.shape.wave_rays, br = B.shape.refill;
    for (int wi = 0; wi < 3 && rc == AP_OK; ++wi)
        for (int ri = 0; ri < 2 && rc == AP_OK; ++ri) {
            B.shape.wave_rays = cand_w[wi];
            B.shape.refill = cand_r[ri];
            if ((rc = batch_run(&B, params, 0, 1, st)) != AP_OK) break;
            e = hipEventRecord(e0, st);
            if (e == hipSuccess && (rc = batch_run(&B, params, 1, frames, st)) != AP_OK) break;
            if (e == hipSuccess) e = hipEventRecord(e1, st);
            if (e == hipSuccess) e = hipEventSynchronize(e1);
            float ms = 0.0f;
            if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
            if (e != hipSuccess) {
                rc = hip_fail(e, "ap_batch_autotune: timing");
                break;
            }
            ms /= static_cast<float>(frames);
            if (best < 0.0f || ms < best) {
                best = ms;
                bw = cand_w[wi];
                br = cand_r[ri];
            }
        }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    B.shape.wave_rays = bw;
    B.shape.refill = br;
    if (rc != AP_OK) return rc;
    if (best_wave_rays) *best_wave_rays = bw;
    if (best_refill) *best_refill = br;
    if (best_ms_per_frame) *best_ms_per_frame = best;
    return AP_OK;
}

int ap_xyz_to_srgb(const float* xyz, size_t n, float* rgb, void* stream) {
    if (n == 0) return AP_OK;
    if (!xyz || !rgb) return fail(AP_ERR_INVALID_ARG, "ap_xyz_to_srgb: NULL pointer");
    hipLaunchKernelGGL(xyz_to_srgb_kernel, grid_of(n), dim3(kBlock), 0, static_cast<hipStream_t>(stream), xyz, n, rgb);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? AP_OK : hip_fail(e, "ap_xyz_to_srgb");
}

}  // extern "C"

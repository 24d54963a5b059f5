// BVH closest-hit and any-hit traversal for gfx950 (CDNA4).
//
// Waves of 64 lanes, one ray per lane at a time. The traversal reproduces the visit
// order of the reference TraverseBVH (core/render/intersection.cuh:235-370):
// pop, re-test the popped node against the current closest distance, test a
// leaf's primitives in slot order, or test both children and push the far
// one first (left first on equal entry distances).
//
// Two layout choices make that order cheap on CDNA4:
//  * Each stack entry carries the child's entry distance from the moment it
//    was pushed. The reference re-runs the slab test on pop with the same
//    operands, so its exit condition (tExit >= max(0,tMin)) is unchanged and
//    the re-test reduces to `tEnter <= closest`, with no second node fetch.
//  * A 64-B NodeRec holds both children's boxes; leaves are runs of 48-B
//    PrimRecs in slot order, ended by PRIM_LAST.
// One unit of work per lane and loop iteration: a pop (then the node's
// expansion) or one primitive of the leaf the lane is in. A lane that pops a
// leaf tests its first primitive in the same iteration, and the leaf's
// primitives follow one per iteration in slot order, so each ray still visits
// nodes and primitives in the reference's order; but lanes expanding nodes no
// longer wait masked off while another lane loops over a leaf (measured: 1.6x
// primaries, 2x secondaries on the 1M-triangle soup; the loop is VALU-issue
// bound, and this lifts SIMD lane utilisation).
// The near child of an expansion is popped at once by the reference, with the
// same closest distance it was tested against, so it stays in registers as
// the next node (its re-test cannot fail) and only far children go on the
// stack. The top RING stack entries live in LDS as [entry][lane] columns
// (conflict-free), older ones in the lane's private (scratch) memory, written
// with ordinary cached stores; 4 KB of LDS per wave leaves occupancy to the
// register budget. The capacity is picked on the host from the BVH depth;
// pushes beyond it are dropped exactly as the reference drops pushes beyond
// its 128 entries (intersection.cuh:258-267), which with capacity >= depth +
// 2 never happens.
#include <algorithm>

#include "ap_internal.h"
#include "ap_prim.h"
#include "ap_sdf.h"

namespace ap {
namespace {

constexpr int WAVE = 64;

constexpr int RING = 8;  // LDS stack entries per lane (a power of two below every CAP)

__device__ __forceinline__ uint64_t lanes_below(int lane) {
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

// Chunked traversal. Wave w owns the contiguous ray range
// [w*chunk, (w+1)*chunk). chunk is `wave_rays`, or less when the actual
// ray count (a queue's device-side count, or a small batch) would leave most
// launched waves idle: then the rays are spread over every launched wave in
// multiples of 64. Each lane traces one ray at
// a time; once at least `refill` lanes are idle the wave hands them the next
// rays of its range in one ballot, so a wave no longer waits for its slowest
// ray, and the launch tail is bounded by one chunk. A step is one pop of the
// reference loop: re-test, then a leaf's primitives or an interior node's
// two children.
// MULTI: rays of several scenes in one queue; ray j belongs to scene
// ray_scene[j] and traverses scenes[ray_scene[j]] (`s` then only carries the
// launch shape). Otherwise every ray traverses `s`.
template <bool ANY, bool STATS, bool CULL, int CAP, bool MULTI = false, bool SDF = false>
__global__ __launch_bounds__(WAVE) void trace_kernel(DeviceScene s, const ap_ray* __restrict__ rays,
                                                     uint32_t n_cap, const uint32_t* __restrict__ dyn_n,
                                                     ap_hit* __restrict__ hits,
                                                     ap_trace_stats* __restrict__ stats,
                                                     const DeviceScene* __restrict__ scenes = nullptr,
                                                     const uint16_t* __restrict__ ray_scene = nullptr) {
    // Stack entries [lo, sp) live in an LDS ring of RING entries per lane
    // (position i % RING); entries [0, lo) that a deeper push evicted live in
    // the lane's private array (scratch memory: this file is built with
    // alloca promotion to registers off, so the array stays in scratch).
    __shared__ int32_t st_link[RING * WAVE];
    __shared__ float st_te[RING * WAVE];
    int32_t sp_link[CAP - RING];
    float sp_te[CAP - RING];
    // The next node to visit when has_cur (the near child, or the root).
    bool has_cur = false;
    int32_t cur_link = 0;
    float cur_te = 0.0f;
    int32_t leaf_slot = -1;  // >= 0: the next primitive of the leaf being tested
    int lo = 0;
    const int lane = threadIdx.x;
    // A device-side count (queues of the integrator) bounds a grid sized for capacity n_cap.
    const uint32_t n = dyn_n ? min(*dyn_n, n_cap) : n_cap;
    const uint32_t nb = gridDim.x, b = blockIdx.x;
    // Wave b owns chunk b: consecutive chunks go to consecutive workgroups,
    // which the hardware deals round-robin over the 8 XCDs, so every XCD
    // gets chunks from all over the batch. Giving each XCD a contiguous
    // eighth instead (for its L2) leaves the XCDs unequally loaded wherever
    // cost varies across the batch: an integrator queue in pixel order (8.9
    // against 10.4 ms per sample at 1080p), and the 4096^2 primaries of the
    // 1M-triangle soup (1513 against 1171 Mrays/s at 512 rays per wave).
    const uint32_t wave_rank = b;
    uint32_t chunk = static_cast<uint32_t>(s.wave_rays);
    {
        const uint32_t share = ((n + nb - 1) / nb + WAVE - 1) / WAVE * WAVE;  // wave-uniform
        chunk = min(chunk, max(share, static_cast<uint32_t>(WAVE)));
    }
    if (static_cast<uint64_t>(wave_rank) * chunk >= n) return;
    uint32_t next = wave_rank * chunk;   // wave-uniform
    const uint32_t end = min(n, next + chunk);
    const int refill = min(max(s.refill, 1), WAVE);  // at most WAVE, or an all-idle wave would never refill

    bool live = false;
    uint32_t ri = 0;
    f3 o{0, 0, 0}, d{0, 0, 0}, inv{0, 0, 0}, w{0, 0, 0};
    float tmin = 0.0f, dist = 0.0f;
    int32_t hit = -1;
    int sp = 0;
    uint32_t nnode = 0, nprim = 0;
    const NodeRec* nodes = s.nodes;
    const PrimRec* prims = s.prims;
    const AnalyticRec* analytic = s.analytic;
    const SdfVolume* volumes = s.volumes;
    int32_t n_volumes = s.n_volumes;

    for (;;) {
        const uint64_t idle = __ballot(!live);
        if (next < end && __popcll(idle) >= refill) {
            if (!live) {
                const uint32_t j = next + static_cast<uint32_t>(__popcll(idle & lanes_below(lane)));
                if (j < end) {
                    ri = j;
                    const float4* rp = reinterpret_cast<const float4*>(rays + j);
                    const float4 r0 = rp[0];
                    const float4 r1 = rp[1];
                    o = f3{r0.x, r0.y, r0.z};
                    d = f3{r1.x, r1.y, r1.z};
                    tmin = r0.w;
                    dist = r1.w;
                    hit = -1;
                    nnode = 0;
                    nprim = 0;
                    sp = 0;
                    lo = 0;
                    has_cur = false;
                    leaf_slot = -1;
                    bool finished = true;
                    RootInfo root = s.root;
                    float margin = s.line_margin;
                    if constexpr (MULTI) {
                        const DeviceScene& D = scenes[ray_scene[j]];
                        nodes = D.nodes;
                        prims = D.prims;
                        analytic = D.analytic;
                        if constexpr (SDF) {
                            volumes = D.volumes;
                            n_volumes = D.n_volumes;
                        }
                        root = D.root;
                        margin = D.line_margin;
                    }
                    if (!root.valid) {
                        dist = __int_as_float(0x7f800000);
                    } else {
                        inv = inv_safe(d);
                        w = absf3(inv) * margin;
                        float te;
                        if (slab_enter<CULL>(o, inv, w, ld3(root.bmin), ld3(root.bmax), tmin, dist, te)) {
                            cur_link = root.link;
                            cur_te = te;
                            has_cur = true;
                            finished = false;
                        } else if (STATS) {
                            nnode = 1;  // the reference pops and rejects the root
                        }
                    }
                    if (finished) {
                        *reinterpret_cast<int2*>(hits + ri) = int2{hit, __float_as_int(dist)};
                        if (STATS) stats[ri] = ap_trace_stats{nnode, nprim};
                    } else {
                        live = true;
                    }
                }
            }
            next = min(end, next + static_cast<uint32_t>(__popcll(idle)));
        }
        if (__ballot(live) == 0) {
            if (next >= end) break;
            continue;
        }
        if (!live) continue;

        bool any_done = false;
        if (leaf_slot < 0) {
            int32_t link;
            float te;
            if (has_cur) {
                link = cur_link;
                te = cur_te;
                has_cur = false;
            } else {
                --sp;
                if (sp < lo) {
                    link = sp_link[sp];
                    te = sp_te[sp];
                    lo = sp;
                } else {
                    link = st_link[(sp & (RING - 1)) * WAVE + lane];
                    te = st_te[(sp & (RING - 1)) * WAVE + lane];
                }
            }
            if (STATS) ++nnode;
            if (te <= dist) {
                if (link < 0) {
                    leaf_slot = ~link;   // its first primitive is tested below, in this iteration
                } else {
                    const float4* nr = reinterpret_cast<const float4*>(nodes + link);
                    const float4 a = nr[0];
                    const float4 bb = nr[1];
                    const float4 c = nr[2];
                    const float4 e = nr[3];
                    if (STATS) nnode += 2;
                    float tl = 0.0f, tr = 0.0f;
                    const bool hl = slab_enter<CULL>(o, inv, w, {a.x, a.y, a.z}, {bb.x, bb.y, bb.z}, tmin, dist, tl);
                    const bool hr = slab_enter<CULL>(o, inv, w, {c.x, c.y, c.z}, {e.x, e.y, e.z}, tmin, dist, tr);
                    const int32_t ll = __float_as_int(a.w);
                    const int32_t rl = __float_as_int(bb.w);
                    if (hl && hr) {
                        // Far child on the stack, near child next; ties keep left near.
                        const bool left_far = tl > tr;
                        if (sp < CAP) {
                            if (sp - lo == RING) {
                                const int k = (lo & (RING - 1)) * WAVE + lane;
                                sp_link[lo] = st_link[k];
                                sp_te[lo] = st_te[k];
                                ++lo;
                            }
                            const int k = (sp & (RING - 1)) * WAVE + lane;
                            st_link[k] = left_far ? ll : rl;
                            st_te[k] = left_far ? tl : tr;
                            ++sp;
                        }
                        // The reference pushes the near child above the far one
                        // and drops a push at capacity, so the near child is kept
                        // only while a slot is left for it.
                        if (sp < CAP) {
                            cur_link = left_far ? rl : ll;
                            cur_te = left_far ? tr : tl;
                            has_cur = true;
                        }
                    } else if ((hl || hr) && sp < CAP) {
                        cur_link = hl ? ll : rl;
                        cur_te = hl ? tl : tr;
                        has_cur = true;
                    }
                }
            }
        }
        if (leaf_slot >= 0) {
            // One primitive of the current leaf per iteration, in slot order.
            const float4* pr = reinterpret_cast<const float4*>(prims + leaf_slot);
            const float4 A = pr[0];
            const float4 B = pr[1];
            const float4 C = pr[2];
            const uint32_t flags = __float_as_uint(B.w);
            const uint32_t kind = flags >> PRIM_KIND_SHIFT;
            float t = 0.0f;
            bool ok;
            if (kind == PRIM_KIND_TRIANGLE) {
                float u, v;
                ok = tri_mt(o, d, {A.x, A.y, A.z}, {B.x, B.y, B.z}, {C.x, C.y, C.z}, t, u, v);
            } else if (kind == PRIM_KIND_CUBE) {
                ok = cube_t(o, d, analytic[__float_as_int(C.w)], t);
            } else if (kind == PRIM_KIND_SPHERE) {
                ok = sphere_t(o, d, analytic[__float_as_int(C.w)], t);
            } else if (SDF && kind == PRIM_KIND_SDF) {
                // SDF_EPS of the ray's depth: tmin / 2 for closest hit (tmin is
                // INTERSECT_EPS of the depth), the launch's value for any hit.
                const float eps = ANY ? s.sdf_eps_any : (tmin > 0.0f ? tmin * 0.5f : kSdfEps0);
                ok = sdf_t(o, d, analytic[__float_as_int(C.w)], volumes, n_volumes, eps, t);
            } else {
                ok = false;
            }
            if (STATS) ++nprim;
            // SceneObjectsIntersector acceptance (intersection.cuh:160).
            if (ok && !(t < tmin || t >= dist)) {
                dist = t;
                hit = __float_as_int(A.w);
                if (ANY) any_done = true;
            }
            leaf_slot = (flags & PRIM_LAST) || any_done ? -1 : leaf_slot + 1;
        }
        if ((leaf_slot < 0 && !has_cur && sp == 0) || any_done) {
            *reinterpret_cast<int2*>(hits + ri) = int2{hit, __float_as_int(dist)};
            if (STATS) stats[ri] = ap_trace_stats{nnode, nprim};
            live = false;
        }
    }
}

template <bool ANY, bool STATS, bool CULL, bool MULTI = false>
hipError_t launch_cap(const DeviceScene& s, const ap_ray* rays, uint32_t n, const uint32_t* dyn_n, ap_hit* hits,
                      ap_trace_stats* stats, hipStream_t stream, const DeviceScene* scenes = nullptr,
                      const uint16_t* ray_scene = nullptr) {
    const dim3 block(WAVE);
    const int cap = s.stack_cap;
    // Enough waves for the chip even when a small batch asks for few chunks:
    // up to kMinWaves waves of at least 64 rays each. About 8192 waves is
    // best measured at 1080p: a 2M-ray bounce runs 14% faster in 8192 waves
    // of 256 rays than in 4096 of 512, a 4M-ray one best in 8192 of 512.
    constexpr uint32_t kMinWaves = 8192;
#define AP_CAP_CASE(C)                                                                                  \
    if (cap <= C) {                                                                                     \
        const uint32_t chunk = static_cast<uint32_t>(s.wave_rays);                                      \
        const uint32_t waves = std::max((n + chunk - 1) / chunk, std::min((n + WAVE - 1) / WAVE, kMinWaves)); \
        if (s.has_sdf)                                                                                  \
            hipLaunchKernelGGL((trace_kernel<ANY, STATS, CULL, C, MULTI, true>), dim3(waves), block, 0, stream, s, \
                               rays, n, dyn_n, hits, stats, scenes, ray_scene);                         \
        else                                                                                            \
            hipLaunchKernelGGL((trace_kernel<ANY, STATS, CULL, C, MULTI>), dim3(waves), block, 0, stream, s, \
                               rays, n, dyn_n, hits, stats, scenes, ray_scene);                         \
        return hipGetLastError();                                                                       \
    }
    AP_CAP_CASE(16)
    AP_CAP_CASE(24)
    AP_CAP_CASE(32)
    AP_CAP_CASE(48)
    AP_CAP_CASE(64)
    AP_CAP_CASE(96)
    AP_CAP_CASE(128)
#undef AP_CAP_CASE
    return hipErrorInvalidValue;
}

}  // namespace

hipError_t launch_trace(const DeviceScene& scene, const ap_ray* rays, size_t n_total, ap_hit* hits,
                        ap_trace_stats* stats, bool any_hit, hipStream_t stream, const uint32_t* dyn_n,
                        float sdf_eps_any) {
    if (n_total == 0) return hipSuccess;
    DeviceScene s = scene;
    s.sdf_eps_any = sdf_eps_any;
    // Ray indices are 32-bit inside the kernel; larger host batches go in slices.
    constexpr size_t kSlice = size_t{1} << 31;
    if (n_total > kSlice) {
        if (dyn_n) return hipErrorInvalidValue;
        for (size_t off = 0; off < n_total; off += kSlice) {
            const size_t m = std::min(kSlice, n_total - off);
            const hipError_t e = launch_trace(s, rays + off, m, hits + off, stats ? stats + off : nullptr, any_hit,
                                              stream, nullptr, sdf_eps_any);
            if (e != hipSuccess) return e;
        }
        return hipSuccess;
    }
    const uint32_t n = static_cast<uint32_t>(n_total);
    const bool cull = s.mode == AP_TRAVERSE_CULLED;
#define AP_DISPATCH(A, S, C) \
    if (any_hit == A && (stats != nullptr) == S && cull == C) return launch_cap<A, S, C>(s, rays, n, dyn_n, hits, stats, stream);
    AP_DISPATCH(false, false, true)
    AP_DISPATCH(true, false, true)
    AP_DISPATCH(false, true, true)
    AP_DISPATCH(true, true, true)
    AP_DISPATCH(false, false, false)
    AP_DISPATCH(true, false, false)
    AP_DISPATCH(false, true, false)
    AP_DISPATCH(true, true, false)
#undef AP_DISPATCH
    return hipErrorInvalidValue;
}

hipError_t launch_trace_multi(const DeviceScene& batch_shape, const DeviceScene* scenes, const uint16_t* ray_scene,
                              const ap_ray* rays, size_t n_total, ap_hit* hits, bool any_hit, hipStream_t stream,
                              const uint32_t* dyn_n, float sdf_eps_any) {
    if (n_total == 0) return hipSuccess;
    DeviceScene shape = batch_shape;
    shape.sdf_eps_any = sdf_eps_any;
    if (n_total > (size_t{1} << 31)) return hipErrorInvalidValue;
    const uint32_t n = static_cast<uint32_t>(n_total);
    if (shape.mode == AP_TRAVERSE_CULLED)
        return any_hit ? launch_cap<true, false, true, true>(shape, rays, n, dyn_n, hits, nullptr, stream, scenes, ray_scene)
                       : launch_cap<false, false, true, true>(shape, rays, n, dyn_n, hits, nullptr, stream, scenes, ray_scene);
    return any_hit ? launch_cap<true, false, false, true>(shape, rays, n, dyn_n, hits, nullptr, stream, scenes, ray_scene)
                   : launch_cap<false, false, false, true>(shape, rays, n, dyn_n, hits, nullptr, stream, scenes, ray_scene);
}

}  // namespace ap

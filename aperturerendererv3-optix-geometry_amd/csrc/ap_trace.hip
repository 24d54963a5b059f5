// BVH closest-hit and any-hit traversal for gfx950 (CDNA4).
//
// One ray per lane, 64-lane workgroups. The traversal reproduces the visit
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
// The per-lane stack lives in LDS as [entry][lane] columns (conflict-free for
// any per-lane depth). Its capacity is picked on the host from the BVH depth;
// pushes beyond it are dropped exactly as the reference drops pushes beyond
// its 128 entries (intersection.cuh:258-267), which with capacity >=
// depth + 2 never happens.
#include "ap_internal.h"

namespace ap {
namespace {

constexpr int WAVE = 64;

__device__ __forceinline__ f3 ld3(const float* p) { return {p[0], p[1], p[2]}; }

__device__ __forceinline__ m3 ld33(const float* p) {
    return {{p[0], p[1], p[2]}, {p[3], p[4], p[5]}, {p[6], p[7], p[8]}};
}

// IntersectAABBObject (core/object/aabb.cuh:9-89), distance only.
__device__ bool cube_t(f3 o, f3 d, const AnalyticRec& a, float& t) {
    const f3 c = ld3(a.center), ext = ld3(a.extent);
    const m3 w2o = ld33(a.w2o);
    const f3 lo = mul(w2o, o - c);
    const f3 ld = normalize(mul(w2o, d));
    const f3 div = inv_safe(ld);
    const f3 t1 = (-ext - lo) * div;
    const f3 t2 = (ext - lo) * div;
    const f3 mn = fmin3(t1, t2), mx = fmax3(t1, t2);
    const float te = fmaxf(fmaxf(mn.x, mn.y), mn.z);
    const float tx = fminf(fminf(mx.x, mx.y), mx.z);
    if (te > tx) return false;
    if (tx <= 0.0f) return false;
    t = (te < 0.0f) ? tx : te;
    return true;
}

// IntersectSphereObject (core/object/sphere.cuh:10-70), distance only.
__device__ bool sphere_t(f3 o, f3 d, const AnalyticRec& a, float& t) {
    const f3 c = ld3(a.center), ext = ld3(a.extent);
    const m3 w2o = ld33(a.w2o);
    const f3 lo = mul(w2o, o - c);
    const f3 ld = normalize(mul(w2o, d));
    const float r = max3(ext);
    const float b = dot(lo, ld);
    const float cc = dot(lo, lo) - r * r;
    const float disc = b * b - 1.0f * cc;
    if (disc < 0.0f) return false;
    const float s = sqrtf(disc);
    float tt = -b - s;
    if (tt <= 0.0f) tt = -b + s;
    if (tt <= 0.0f) return false;
    if (isinf(tt)) return false;
    t = tt;
    return true;
}

template <bool ANY, bool STATS, bool CULL, int CAP>
__global__ __launch_bounds__(WAVE) void trace_kernel(DeviceScene s, const ap_ray* __restrict__ rays,
                                                     size_t n, const uint32_t* __restrict__ dyn_n,
                                                     ap_hit* __restrict__ hits,
                                                     ap_trace_stats* __restrict__ stats) {
    __shared__ int32_t st_link[CAP * WAVE];
    __shared__ float st_te[CAP * WAVE];
    const int lane = threadIdx.x;
    const size_t i = static_cast<size_t>(blockIdx.x) * WAVE + lane;
    // A device-side count (queues of the integrator) bounds a grid sized for capacity n.
    if (dyn_n) n = *dyn_n;
    if (i >= n) return;

    const float4* rp = reinterpret_cast<const float4*>(rays + i);
    const float4 r0 = rp[0];
    const float4 r1 = rp[1];
    const f3 o{r0.x, r0.y, r0.z};
    const f3 d{r1.x, r1.y, r1.z};
    const float tmin = r0.w;
    float dist = r1.w;
    int32_t hit = -1;
    uint32_t nnode = 0, nprim = 0;

    if (!s.root.valid) {
        dist = __int_as_float(0x7f800000);
    } else {
        const f3 inv = inv_safe(d);
        const f3 w = absf3(inv) * s.line_margin;
        int sp = 0;
        {
            float te;
            if (slab_enter<CULL>(o, inv, w, ld3(s.root.bmin), ld3(s.root.bmax), tmin, dist, te)) {
                st_link[lane] = s.root.link;
                st_te[lane] = te;
                sp = 1;
            } else if (STATS) {
                nnode = 1;  // the reference pops and rejects the root
            }
        }
        while (sp > 0) {
            --sp;
            const int32_t link = st_link[sp * WAVE + lane];
            const float te = st_te[sp * WAVE + lane];
            if (STATS) ++nnode;
            if (!(te <= dist)) continue;
            if (link < 0) {
                int32_t slot = ~link;
                bool done = false;
                for (;;) {
                    const float4* pr = reinterpret_cast<const float4*>(s.prims + slot);
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
                        ok = cube_t(o, d, s.analytic[__float_as_int(C.w)], t);
                    } else if (kind == PRIM_KIND_SPHERE) {
                        ok = sphere_t(o, d, s.analytic[__float_as_int(C.w)], t);
                    } else {
                        ok = false;
                    }
                    if (STATS) ++nprim;
                    // SceneObjectsIntersector acceptance (intersection.cuh:160).
                    if (ok && !(t < tmin || t >= dist)) {
                        dist = t;
                        hit = __float_as_int(A.w);
                        if (ANY) { done = true; break; }
                    }
                    if (flags & PRIM_LAST) break;
                    ++slot;
                }
                if (ANY && done) break;
            } else {
                const float4* nr = reinterpret_cast<const float4*>(s.nodes + link);
                const float4 a = nr[0];
                const float4 b = nr[1];
                const float4 c = nr[2];
                const float4 e = nr[3];
                if (STATS) nnode += 2;
                float tl = 0.0f, tr = 0.0f;
                const bool hl = slab_enter<CULL>(o, inv, w, {a.x, a.y, a.z}, {b.x, b.y, b.z}, tmin, dist, tl);
                const bool hr = slab_enter<CULL>(o, inv, w, {c.x, c.y, c.z}, {e.x, e.y, e.z}, tmin, dist, tr);
                const int32_t ll = __float_as_int(a.w);
                const int32_t rl = __float_as_int(b.w);
                if (hl && hr) {
                    // Far child first, near child on top; ties keep left on top.
                    const bool left_far = tl > tr;
                    const int32_t l0 = left_far ? ll : rl, l1 = left_far ? rl : ll;
                    const float t0 = left_far ? tl : tr, t1 = left_far ? tr : tl;
                    if (sp < CAP) { st_link[sp * WAVE + lane] = l0; st_te[sp * WAVE + lane] = t0; ++sp; }
                    if (sp < CAP) { st_link[sp * WAVE + lane] = l1; st_te[sp * WAVE + lane] = t1; ++sp; }
                } else if (hl || hr) {
                    if (sp < CAP) {
                        st_link[sp * WAVE + lane] = hl ? ll : rl;
                        st_te[sp * WAVE + lane] = hl ? tl : tr;
                        ++sp;
                    }
                }
            }
        }
    }
    int2 out;
    out.x = hit;
    out.y = __float_as_int(dist);
    *reinterpret_cast<int2*>(hits + i) = out;
    if (STATS) stats[i] = ap_trace_stats{nnode, nprim};
}

template <bool ANY, bool STATS, bool CULL>
hipError_t launch_cap(const DeviceScene& s, const ap_ray* rays, size_t n, const uint32_t* dyn_n, ap_hit* hits,
                      ap_trace_stats* stats, hipStream_t stream) {
    const dim3 block(WAVE);
    const dim3 grid(static_cast<unsigned>((n + WAVE - 1) / WAVE));
    const int cap = s.stack_cap;
#define AP_CAP_CASE(C)                                                                   \
    if (cap <= C) {                                                                      \
        hipLaunchKernelGGL((trace_kernel<ANY, STATS, CULL, C>), grid, block, 0, stream, s, rays, n, \
                           dyn_n, hits, stats);                                                 \
        return hipGetLastError();                                                        \
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

hipError_t launch_trace(const DeviceScene& s, const ap_ray* rays, size_t n, ap_hit* hits,
                        ap_trace_stats* stats, bool any_hit, hipStream_t stream, const uint32_t* dyn_n) {
    if (n == 0) return hipSuccess;
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

}  // namespace ap

// ORACLE — test infrastructure only: CPU restatement of the reference's BVH
// build and traversal, used by tests/, bench.py's cpu_baseline leg and
// __graft_entry__.smoke() as the checker. The product never links it.
//
//   BuildBVH                    core/bvh/bvh.cuh:142-422 (32-bit node variant)
//   FastAABBHit / WithEnter     core/bvh/bvh.cuh:424-470
//   IntersectTriangleObject     core/render/intersection.cuh:15-89
//   IntersectAABBObject         core/object/aabb.cuh:9-89
//   IntersectSphereObject       core/object/sphere.cuh:10-70
//   IntersectBase               core/render/intersection.cuh:93-113
//   SceneObjectsIntersector     core/render/intersection.cuh:140-172
//   Intersect (brute force)     core/render/intersection.cuh:176-205
//   TraverseBVH<bAnyHit>        core/render/intersection.cuh:235-370
//
// Parity pins: builder invariants and brute-force agreement (tests/test_oracle_bvh.py),
// closed-form rays (axis-aligned, parallel, edge/vertex, t <= 1e-5, window bounds).
#include <algorithm>
#include <cfloat>
#include <functional>
#include <limits>
#include <vector>

#include "ap_fmath.h"
#include "aperture/ap_layout.h"
#include "or_api.h"
#include "or_math.h"
#include "or_count.h"

namespace orc {

inline float3 F(const ap_float3& a) { return {a.x, a.y, a.z}; }
inline float3x3 M(const ap_float3x3& m) { return {F(m.x), F(m.y), F(m.z)}; }

// ---- builder (bvh.cuh:80-422) ----
struct PrimitiveProxy { float3 boundMin_, boundMax_, center_; };
struct BVHNode { float3 boundMin_, boundMax_; int left_, right_, first_, count_; };

inline float3 AABBMin(float3 a, float3 b) { return fmin3(a, b); }
inline float3 AABBMax(float3 a, float3 b) { return fmax3(a, b); }
inline float AABBArea(float3 mn, float3 mx) {
    const float3 d = {std::fmax(0.0f, mx.x - mn.x), std::fmax(0.0f, mx.y - mn.y), std::fmax(0.0f, mx.z - mn.z)};
    return 2.0f * (d.x * d.y + d.y * d.z + d.z * d.x);
}

struct BVHBinData {
    float3 boundMin_, boundMax_;
    int count_;
    void reset() { boundMin_ = {FLT_MAX, FLT_MAX, FLT_MAX}; boundMax_ = {-FLT_MAX, -FLT_MAX, -FLT_MAX}; count_ = 0; }
    void add(const float3& mn, const float3& mx) { boundMin_ = AABBMin(boundMin_, mn); boundMax_ = AABBMax(boundMax_, mx); count_++; }
};

struct BoundsAccumulator {
    float3 boundMin_{FLT_MAX, FLT_MAX, FLT_MAX}, boundMax_{-FLT_MAX, -FLT_MAX, -FLT_MAX};
    float3 centroidMin_{FLT_MAX, FLT_MAX, FLT_MAX}, centroidMax_{-FLT_MAX, -FLT_MAX, -FLT_MAX};
    void add(const PrimitiveProxy& p) {
        boundMin_ = AABBMin(boundMin_, p.boundMin_);
        boundMax_ = AABBMax(boundMax_, p.boundMax_);
        centroidMin_ = AABBMin(centroidMin_, p.center_);
        centroidMax_ = AABBMax(centroidMax_, p.center_);
    }
};

struct BVHHost { std::vector<BVHNode> nodes_; std::vector<int> objectIndices_; };

static BVHHost BuildBVH(const std::vector<PrimitiveProxy>& proxy, int leafSize, int binCount) {
    BVHHost result;
    if (proxy.empty()) return result;
    binCount = std::max(4, std::min(64, binCount));
    leafSize = std::max(1, std::min(16, leafSize));
    const int primitiveCount = static_cast<int>(proxy.size());
    result.objectIndices_.resize(primitiveCount);
    for (int i = 0; i < primitiveCount; ++i) result.objectIndices_[i] = i;
    result.nodes_.reserve(2 * static_cast<size_t>(primitiveCount));
    auto computeBounds = [&](int start, int end) {
        BoundsAccumulator acc;
        for (int i = start; i < end; ++i) acc.add(proxy[result.objectIndices_[i]]);
        return acc;
    };
    auto key_of = [&](int idx, int axis) {
        return axis == 0 ? proxy[idx].center_.x : (axis == 1 ? proxy[idx].center_.y : proxy[idx].center_.z);
    };
    std::function<int(int, int)> build = [&](int start, int end) -> int {
        const int count = end - start;
        const int nodeIdx = static_cast<int>(result.nodes_.size());
        result.nodes_.emplace_back();
        BoundsAccumulator b = computeBounds(start, end);
        result.nodes_[nodeIdx].boundMin_ = b.boundMin_;
        result.nodes_[nodeIdx].boundMax_ = b.boundMax_;
        auto leaf = [&]() {
            BVHNode& node = result.nodes_[nodeIdx];
            node.left_ = -1; node.right_ = -1; node.first_ = start; node.count_ = count;
            return nodeIdx;
        };
        if (count <= leafSize) return leaf();
        const float3 ce = {b.centroidMax_.x - b.centroidMin_.x, b.centroidMax_.y - b.centroidMin_.y,
                           b.centroidMax_.z - b.centroidMin_.z};
        int axis = 0;
        if (ce.y > ce.x && ce.y >= ce.z) axis = 1;
        else if (ce.z > ce.x && ce.z >= ce.y) axis = 2;
        if (ce.x <= 1e-6f && ce.y <= 1e-6f && ce.z <= 1e-6f) return leaf();
        const float cMin = axis == 0 ? b.centroidMin_.x : (axis == 1 ? b.centroidMin_.y : b.centroidMin_.z);
        const float cMax = axis == 0 ? b.centroidMax_.x : (axis == 1 ? b.centroidMax_.y : b.centroidMax_.z);
        const float invW = (cMax > cMin) ? static_cast<float>(binCount) / (cMax - cMin) : 0.0f;
        auto binOf = [&](int idx) {
            int bb = invW > 0.0f ? static_cast<int>((key_of(idx, axis) - cMin) * invW) : 0;
            return (bb < 0) ? 0 : (bb >= binCount ? binCount - 1 : bb);
        };
        std::vector<BVHBinData> bins(binCount);
        for (auto& x : bins) x.reset();
        for (int i = start; i < end; ++i) {
            const int idx = result.objectIndices_[i];
            bins[binOf(idx)].add(proxy[idx].boundMin_, proxy[idx].boundMax_);
        }
        std::vector<float> leftArea(binCount), rightArea(binCount);
        std::vector<int> leftCount(binCount), rightCount(binCount);
        float3 runMin = {FLT_MAX, FLT_MAX, FLT_MAX}, runMax = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
        int runCount = 0;
        for (int k = 0; k < binCount; ++k) {
            if (bins[k].count_ > 0) { runMin = AABBMin(runMin, bins[k].boundMin_); runMax = AABBMax(runMax, bins[k].boundMax_); }
            runCount += bins[k].count_;
            leftArea[k] = (runCount > 0) ? AABBArea(runMin, runMax) : 0.0f;
            leftCount[k] = runCount;
        }
        runMin = {FLT_MAX, FLT_MAX, FLT_MAX}; runMax = {-FLT_MAX, -FLT_MAX, -FLT_MAX}; runCount = 0;
        for (int k = binCount - 1; k >= 0; --k) {
            if (bins[k].count_ > 0) { runMin = AABBMin(runMin, bins[k].boundMin_); runMax = AABBMax(runMax, bins[k].boundMax_); }
            runCount += bins[k].count_;
            rightArea[k] = (runCount > 0) ? AABBArea(runMin, runMax) : 0.0f;
            rightCount[k] = runCount;
        }
        const float parentArea = AABBArea(result.nodes_[nodeIdx].boundMin_, result.nodes_[nodeIdx].boundMax_);
        const float costI = 1.0f, costT = 1.0f;
        float bestCost = FLT_MAX;
        int bestSplit = -1;
        for (int k = 0; k < binCount - 1; ++k) {
            const int nl = leftCount[k], nr = rightCount[k + 1];
            if (nl == 0 || nr == 0) continue;
            const float sah = costI + (leftArea[k] * static_cast<float>(nl) + rightArea[k + 1] * static_cast<float>(nr)) /
                                          std::fmax(parentArea, 1e-20f) * costT;
            if (sah < bestCost) { bestCost = sah; bestSplit = k; }
        }
        int mid = -1;
        if (bestSplit >= 0) {
            const int splitBin = bestSplit;
            auto itMid = std::partition(result.objectIndices_.begin() + start, result.objectIndices_.begin() + end,
                                        [&](int idx) { return binOf(idx) <= splitBin; });
            mid = static_cast<int>(itMid - result.objectIndices_.begin());
            if (mid == start || mid == end) bestSplit = -1;
        }
        if (bestSplit < 0) {
            const int m = start + (count / 2);
            std::nth_element(result.objectIndices_.begin() + start, result.objectIndices_.begin() + m,
                             result.objectIndices_.begin() + end,
                             [&](int a, int c) { return key_of(a, axis) < key_of(c, axis); });
            mid = m;
            if (mid == start || mid == end) return leaf();
        }
        const int leftIdx = build(start, mid);
        const int rightIdx = build(mid, end);
        BVHNode& node = result.nodes_[nodeIdx];
        node.left_ = leftIdx; node.right_ = rightIdx; node.first_ = -1; node.count_ = 0;
        return nodeIdx;
    };
    (void)build(0, primitiveCount);
    return result;
}

inline bool FastAABBHit(float3 o, float3 inv, float3 mn, float3 mx, float tMin, float tMax) {
    float t0 = (mn.x - o.x) * inv.x, t1 = (mx.x - o.x) * inv.x;
    float tMin2 = std::fmin(t0, t1), tMax2 = std::fmax(t0, t1);
    t0 = (mn.y - o.y) * inv.y; t1 = (mx.y - o.y) * inv.y;
    tMin2 = std::fmax(tMin2, std::fmin(t0, t1)); tMax2 = std::fmin(tMax2, std::fmax(t0, t1));
    t0 = (mn.z - o.z) * inv.z; t1 = (mx.z - o.z) * inv.z;
    tMin2 = std::fmax(tMin2, std::fmin(t0, t1)); tMax2 = std::fmin(tMax2, std::fmax(t0, t1));
    return tMax2 >= std::fmax(0.0f, tMin) && tMin2 <= tMax;
}

inline bool FastAABBHitWithEnter(float3 o, float3 inv, float3 mn, float3 mx, float tMin, float tMax, float& tEnterOut) {
    const float tx0 = (mn.x - o.x) * inv.x, tx1 = (mx.x - o.x) * inv.x;
    float tEnter = std::fmin(tx0, tx1), tExit = std::fmax(tx0, tx1);
    const float ty0 = (mn.y - o.y) * inv.y, ty1 = (mx.y - o.y) * inv.y;
    tEnter = std::fmax(tEnter, std::fmin(ty0, ty1)); tExit = std::fmin(tExit, std::fmax(ty0, ty1));
    const float tz0 = (mn.z - o.z) * inv.z, tz1 = (mx.z - o.z) * inv.z;
    tEnter = std::fmax(tEnter, std::fmin(tz0, tz1)); tExit = std::fmin(tExit, std::fmax(tz0, tz1));
    if (tExit >= std::fmax(0.0f, tMin) && tEnter <= tMax) { tEnterOut = tEnter; return true; }
    return false;
}

// Extra rejection of the core's CULLED traversal mode (not in the
// reference): the line misses the box grown by `margin` in every axis.
inline bool LineMissesGrownBox(float3 o, float3 inv, float margin, float3 mn, float3 mx) {
    const float wx = std::fabs(inv.x) * margin, wy = std::fabs(inv.y) * margin, wz = std::fabs(inv.z) * margin;
    const float tx0 = (mn.x - o.x) * inv.x, tx1 = (mx.x - o.x) * inv.x;
    const float ty0 = (mn.y - o.y) * inv.y, ty1 = (mx.y - o.y) * inv.y;
    const float tz0 = (mn.z - o.z) * inv.z, tz1 = (mx.z - o.z) * inv.z;
    const float le = std::fmax(std::fmax(std::fmin(tx0, tx1) - wx, std::fmin(ty0, ty1) - wy), std::fmin(tz0, tz1) - wz);
    const float lx = std::fmin(std::fmin(std::fmax(tx0, tx1) + wx, std::fmax(ty0, ty1) + wy), std::fmax(tz0, tz1) + wz);
    return le > lx;
}

// Distance-only restatements of the type intersectors.
bool IntersectTriangleT(float3 o, float3 d, const ap_scene_object& ob, float& t, float& uu, float& vv) {
    if (!ob.visible) return false;
    const float3x3 o2w = M(ob.object_to_world);
    const float3 c = F(ob.center);
    const float3 p0 = o2w * F(ob.additional.v0.position) + c;
    const float3 p1 = o2w * F(ob.additional.v1.position) + c;
    const float3 p2 = o2w * F(ob.additional.v2.position) + c;
    const float3 e1 = p1 - p0, e2 = p2 - p0;
    const float3 h = cross(d, e2);
    const float a = dot(e1, h);
    if (a > -1e-6f && a < 1e-6f) return false;
    const float f = 1.0f / a;
    const float3 s = o - p0;
    const float u = f * dot(s, h);
    if (u < 0.0f || u > 1.0f) return false;
    const float3 q = cross(s, e1);
    const float v = f * dot(d, q);
    if (v < 0.0f || u + v > 1.0f) return false;
    const float tt = f * dot(e2, q);
    if (tt > 1e-5f) { t = tt; uu = u; vv = v; return true; }
    return false;
}

bool IntersectAABBT(float3 o, float3 d, const ap_scene_object& ob, float& t, bool& inside) {
    const float3x3 w2o = M(ob.world_to_object);
    const float3 lo = w2o * (o - F(ob.center));
    const float3 ld = normalize(w2o * d);
    const float3 div = invSafe(ld);
    const float3 t1 = (-F(ob.extent) - lo) * div;
    const float3 t2 = (F(ob.extent) - lo) * div;
    const float3 mn = fmin3(t1, t2), mx = fmax3(t1, t2);
    const float tE = std::fmax(std::fmax(mn.x, mn.y), mn.z);
    const float tX = std::fmin(std::fmin(mx.x, mx.y), mx.z);
    if (tE > tX) return false;
    if (tX <= 0.0f) return false;
    inside = tE < 0.0f;
    t = inside ? tX : tE;
    return true;
}

bool IntersectSphereT(float3 o, float3 d, const ap_scene_object& ob, float& t, bool& inside) {
    const float3x3 w2o = M(ob.world_to_object);
    const float3 lo = w2o * (o - F(ob.center));
    const float3 ld = normalize(w2o * d);
    const float r = max3(F(ob.extent));
    inside = false;
    const float a = 1;
    const float b = dot(lo, ld);
    const float c = dot(lo, lo) - r * r;
    const float disc = b * b - a * c;
    float tt;
    if (disc < 0.0f) tt = std::numeric_limits<float>::infinity();
    else {
        const float s = std::sqrt(disc);
        tt = -b - s;
        if (tt <= 0.0f) { tt = -b + s; inside = true; }
        if (tt <= 0.0f) tt = std::numeric_limits<float>::infinity();
    }
    if (std::isinf(tt)) return false;
    t = tt;
    return true;
}

// ---- Signed-distance objects (core/object/sdf/) ----
// Volumes of the SDF cache (SDFCacheManager), registered by or_sdf_set_volumes.
struct OrVolume {
    std::vector<float> v;
    int dim[3];
    float voxelSizeMax;
};
static std::vector<OrVolume> g_volumes;

// SdfTorus (sdf_torus.cuh:15-19).
static float SdfTorus(float3 u, float R, float r) {
    const float qx = std::sqrt(u.x * u.x + u.z * u.z) - R;
    return std::sqrt(qx * qx + u.y * u.y) - r;
}

// PlaneSDF / SdfDiamondUnit (sdf_diamond.cuh:15-66), sin, cos and atan2 the
// shared portable ones.
static float PlaneSDF(float3 p, float3 n, float d) {
    const float len = std::fmax(std::sqrt(n.x * n.x + n.y * n.y + n.z * n.z), 1e-8f);
    return (p.x * n.x + p.y * n.y + p.z * n.z - d) / len;
}
static float SdfDiamondUnit(float3 u, const float* info) {
    const float d = (info[0] != 0.0f) ? info[0] : 0.94f;
    const float bias = (info[1] != 0.0f) ? info[1] : 0.5f;
    const float tableZ = (info[2] != 0.0f) ? info[2] : 0.3f;
    const float scale = (info[3] != 0.0f) ? info[3] : 1.0f;
    const float scaleInv = 1.0f / scale;
    const float3 p{u.x * scaleInv, u.y * scaleInv, u.z * scaleInv};
    constexpr float af2 = 4.0f / 3.141592654f;
    const float s = apfm::atan2(p.y, p.x);
    const float sf = std::floor(s * af2 + bias) / af2;
    const float sf2 = std::floor(s * af2) / af2;
    auto cs = [](float a, float& c, float& sn) { apfm::sincos(a, sn, c); };
    float csf, ssf, c0, s0, csf1, ssf1, csf2, ssf2, csf8, ssf8;
    cs(sf, csf, ssf);
    cs(s, c0, s0);
    cs(sf + 0.21f, csf1, ssf1);
    cs(sf - 0.21f, csf2, ssf2);
    cs(sf2 + 0.393f, csf8, ssf8);
    float dMax = -1e20f;
    dMax = std::fmax(dMax, PlaneSDF(p, {csf, ssf, 1.444f}, d));
    dMax = std::fmax(dMax, PlaneSDF(p, {csf, ssf, -1.072f}, d));
    dMax = std::fmax(dMax, PlaneSDF(p, {0.0f, 0.0f, 1.0f}, tableZ));
    dMax = std::fmax(dMax, PlaneSDF(p, {0.0f, 0.0f, -1.0f}, 0.865f));
    dMax = std::fmax(dMax, PlaneSDF(p, {c0, s0, 0.0f}, 0.911f));
    dMax = std::fmax(dMax, PlaneSDF(p, {csf1, ssf1, -1.02f}, 0.9193f));
    dMax = std::fmax(dMax, PlaneSDF(p, {csf2, ssf2, -1.02f}, 0.9193f));
    dMax = std::fmax(dMax, PlaneSDF(p, {csf2, ssf2, 1.03f}, 0.912f));
    dMax = std::fmax(dMax, PlaneSDF(p, {csf1, ssf1, 1.03f}, 0.912f));
    dMax = std::fmax(dMax, PlaneSDF(p, {csf8, ssf8, 2.21f}, 1.131f));
    return dMax * scale;
}

static float VolAt(const OrVolume& V, int x, int y, int z) {  // Index3D (sdf_volume.cuh:47-50)
    return V.v[static_cast<size_t>(z) * V.dim[0] * V.dim[1] + static_cast<size_t>(y) * V.dim[0] + x];
}
static int ClampI(int x, int lo, int hi) { return std::max(lo, std::min(hi, x)); }

// SampleSDFAdaptive01 (sdf_volume.cuh:64-153): nearest sample, or the
// trilinear one within voxelSizeMax of the surface.
static float SampleSDFAdaptive01(const OrVolume& V, float3 p) {
    if (V.v.empty()) return 1.0e20f;
    const float3 g{p.x * static_cast<float>(V.dim[0] - 1), p.y * static_cast<float>(V.dim[1] - 1),
                   p.z * static_cast<float>(V.dim[2] - 1)};
    const float nearest = VolAt(V, ClampI(static_cast<int>(std::floor(g.x + 0.5f)), 0, V.dim[0] - 1),
                                ClampI(static_cast<int>(std::floor(g.y + 0.5f)), 0, V.dim[1] - 1),
                                ClampI(static_cast<int>(std::floor(g.z + 0.5f)), 0, V.dim[2] - 1));
    if (std::fabs(nearest) > V.voxelSizeMax) return nearest;
    int x0 = static_cast<int>(std::floor(g.x)), y0 = static_cast<int>(std::floor(g.y));
    int z0 = static_cast<int>(std::floor(g.z));
    const int x1 = ClampI(x0 + 1, 0, V.dim[0] - 1), y1 = ClampI(y0 + 1, 0, V.dim[1] - 1);
    const int z1 = ClampI(z0 + 1, 0, V.dim[2] - 1);
    x0 = ClampI(x0, 0, V.dim[0] - 1);
    y0 = ClampI(y0, 0, V.dim[1] - 1);
    z0 = ClampI(z0, 0, V.dim[2] - 1);
    const float fx = std::fmin(std::fmax(g.x - static_cast<float>(x0), 0.0f), 1.0f);
    const float fy = std::fmin(std::fmax(g.y - static_cast<float>(y0), 0.0f), 1.0f);
    const float fz = std::fmin(std::fmax(g.z - static_cast<float>(z0), 0.0f), 1.0f);
    const float c00 = VolAt(V, x0, y0, z0) * (1.0f - fx) + VolAt(V, x1, y0, z0) * fx;
    const float c10 = VolAt(V, x0, y1, z0) * (1.0f - fx) + VolAt(V, x1, y1, z0) * fx;
    const float c01 = VolAt(V, x0, y0, z1) * (1.0f - fx) + VolAt(V, x1, y0, z1) * fx;
    const float c11 = VolAt(V, x0, y1, z1) * (1.0f - fx) + VolAt(V, x1, y1, z1) * fx;
    const float c0 = c00 * (1.0f - fy) + c10 * fy;
    const float c1 = c01 * (1.0f - fy) + c11 * fy;
    return c0 * (1.0f - fz) + c1 * fz;
}

// The object's distance functor (sdf_master.cuh:32-70): nullptr when the
// object is never hit (undefined type, missing volume).
struct SDFFn {
    int type;
    const OrVolume* vol;
    float R, r;
    float operator()(float3 u) const {
        if (type == 0) return SdfTorus(u, R, r);
        return SampleSDFAdaptive01(*vol, u + 0.5f);  // SDFVolumeFunctor (sdf_volume.cuh:155-171)
    }
};
static bool MakeSDFFn(const ap_scene_object& ob, SDFFn& f) {
    f.type = ob.additional.sdf_type;
    f.vol = nullptr;
    f.R = ob.additional.sdf_info.sdf_info0.x;
    f.r = ob.additional.sdf_info.sdf_info0.y;
    if (f.type == 0) return true;
    const int k = ob.additional.sdf_info.cached_volume_index;
    if (f.type == 1 && k >= 0 && k < static_cast<int>(g_volumes.size()) && !g_volumes[k].v.empty()) {
        f.vol = &g_volumes[k];
        return true;
    }
    return false;
}

static float3 ToUnitSpace(float3 p, float3 e) {  // sdf_base.cuh:12-21
    return {p.x / (2.0f * std::fmax(e.x, 1e-8f)), p.y / (2.0f * std::fmax(e.y, 1e-8f)),
            p.z / (2.0f * std::fmax(e.z, 1e-8f))};
}

// RaymarchLocal (sdf_base.cuh:54-110).
static bool RaymarchLocal(const SDFFn& sdf, float3 lo, float3 ld, float tEnter, float tExit, float3 extent,
                          float currentEPS, float* outT, float3* outP) {
    const float minStep = currentEPS * 0.5f;
    const bool bStartInside = sdf(ToUnitSpace(lo, extent)) <= 0.0f;
    float t = std::fmax(0.0f, tEnter);
    float tLast = t;
    const float sMin = 2.0f * std::fmin(extent.x, std::fmin(extent.y, extent.z));
    for (int i = 0; i < 512; ++i) {
        tLast = t;
        const float3 pL = lo + ld * t;
        const float dLocal = sdf(ToUnitSpace(pL, extent)) * sMin;
        const float realDistance = std::fabs(dLocal);
        if (bStartInside ? (dLocal > 0 && dLocal < currentEPS) : (dLocal < 0 && -dLocal < currentEPS)) {
            *outT = tLast;
            *outP = pL;
            return true;
        }
        const float advance = std::fmax(sdf.type == 1 ? realDistance * 0.9f : realDistance, minStep);
        t += advance;
        if (t > tExit) break;
    }
    return false;
}

// IntersectSDFObject (sdf_base.cuh:115-163): distance and local hit point.
bool IntersectSDFT(float3 o, float3 d, const ap_scene_object& ob, float eps, float& t, float3* localPos) {
    SDFFn f;
    if (!MakeSDFFn(ob, f)) return false;
    const float3x3 w2o = M(ob.world_to_object);
    const float3 lo = w2o * (o - F(ob.center));
    const float3 ld = normalize(w2o * d);
    const float3 div = invSafe(ld);
    const float3 t1 = (-F(ob.extent) - lo) * div;
    const float3 t2 = (F(ob.extent) - lo) * div;
    const float3 mn = fmin3(t1, t2), mx = fmax3(t1, t2);
    const float tE = std::fmax(std::fmax(mn.x, mn.y), mn.z);
    const float tX = std::fmin(std::fmin(mx.x, mx.y), mx.z);
    if (tE > tX) return false;
    if (tX <= 0.0f) return false;
    float3 lp{0, 0, 0};
    if (!RaymarchLocal(f, lo, ld, tE, tX, F(ob.extent), eps, &t, &lp)) return false;
    if (localPos) *localPos = lp;
    return true;
}

// sdfNormalUnit + UnitNormalToLocal (sdf_base.cuh:23-52).
float3 SDFLocalNormal(const ap_scene_object& ob, float3 lp, float eps) {
    SDFFn f;
    (void)MakeSDFFn(ob, f);
    const float3 e = F(ob.extent);
    const float sMin = 2.0f * std::fmin(e.x, std::fmin(e.y, e.z));
    const float h = eps / std::fmax(sMin, 1e-8f);
    const float3 u = ToUnitSpace(lp, e);
    const float dx = f({u.x + h, u.y, u.z}) - f({u.x - h, u.y, u.z});
    const float dy = f({u.x, u.y + h, u.z}) - f({u.x, u.y - h, u.z});
    const float dz = f({u.x, u.y, u.z + h}) - f({u.x, u.y, u.z - h});
    const float3 nu = normalize(float3{dx, dy, dz});
    return normalize(float3{nu.x / (2.0f * std::fmax(e.x, 1e-8f)), nu.y / (2.0f * std::fmax(e.y, 1e-8f)),
                            nu.z / (2.0f * std::fmax(e.z, 1e-8f))});
}

// OverlapSDFObject (sdf_base.cuh:165-177).
bool OverlapSDF(const ap_scene_object& ob, float3 p) {
    SDFFn f;
    if (!MakeSDFFn(ob, f)) return false;
    return f(ToUnitSpace(M(ob.world_to_object) * (p - F(ob.center)), F(ob.extent))) <= 0.0f;
}

// IntersectBase + SceneObjectsIntersector acceptance (intersection.cuh:93-172);
// sdfEps = SDF_EPS of the ray's depth.
bool Intersector(const ap_scene_object* objs, int idx, float3 o, float3 d, float tMin, float tMax, float& tOut,
                 float sdfEps) {
    const ap_scene_object& ob = objs[idx];
    float t = 0.0f;
    bool hit = false, inside = false;
    float u, v;
    if (ob.type == AP_OBJ_CUBE) hit = IntersectAABBT(o, d, ob, t, inside);
    else if (ob.type == AP_OBJ_SPHERE) hit = IntersectSphereT(o, d, ob, t, inside);
    else if (ob.type == AP_OBJ_SDF) hit = IntersectSDFT(o, d, ob, sdfEps, t, nullptr);
    else if (ob.type == AP_OBJ_TRIANGLE) hit = IntersectTriangleT(o, d, ob, t, u, v);
    if (!hit) return false;
    if (t < tMin || t >= tMax) return false;
    tOut = t;
    return true;
}

struct TraceResult { int id; float t; unsigned nodes; unsigned prims; };

// TraverseBVH<bAnyHit> (intersection.cuh:235-370), 32-bit nodes, sorted stack.
// cull = false is the reference; cull = true adds LineMissesGrownBox to both
// box tests (the core's CULLED mode), for its visit-order accounting.
TraceResult TraverseBVH(bool anyHit, const ap_scene_object* objs, const std::vector<BVHNode>& nodes,
                        const std::vector<int>& objectIndices, float3 o, float3 d, float tMin, float tMax,
                        bool cull, float margin, float sdfEps) {
    TraceResult r{-1, std::numeric_limits<float>::infinity(), 0u, 0u};
    if (nodes.empty()) return r;
    r.t = tMax;
    constexpr int STACK_CAP = 128;
    int stack[STACK_CAP];
    int sp = 0;
    auto PUSH = [&](int n) { if (sp < STACK_CAP) stack[sp++] = n; };
    PUSH(0);
    const float3 inv = invSafe(d);
    while (sp) {
        const int cur = stack[--sp];
        const BVHNode& node = nodes[cur];
        r.nodes++;
        if (!FastAABBHit(o, inv, node.boundMin_, node.boundMax_, tMin, r.t)) continue;
        if (cull && LineMissesGrownBox(o, inv, margin, node.boundMin_, node.boundMax_)) continue;
        if (node.count_ > 0) {
            for (int i = 0; i < node.count_; ++i) {
                const int objectIndex = objectIndices[node.first_ + i];
                r.prims++;
                float t;
                if (Intersector(objs, objectIndex, o, d, tMin, r.t, t, sdfEps)) {
                    r.t = t;
                    r.id = objectIndex;
                    if (anyHit) return r;
                }
            }
        } else {
            const BVHNode& left = nodes[node.left_];
            const BVHNode& right = nodes[node.right_];
            r.nodes += 2;
            float tL = 0.0f, tR = 0.0f;
            bool hitL = FastAABBHitWithEnter(o, inv, left.boundMin_, left.boundMax_, tMin, r.t, tL);
            bool hitR = FastAABBHitWithEnter(o, inv, right.boundMin_, right.boundMax_, tMin, r.t, tR);
            if (cull) {
                hitL = hitL && !LineMissesGrownBox(o, inv, margin, left.boundMin_, left.boundMax_);
                hitR = hitR && !LineMissesGrownBox(o, inv, margin, right.boundMin_, right.boundMax_);
            }
            if (hitL && hitR) {
                if (tL > tR) { PUSH(node.left_); PUSH(node.right_); }
                else { PUSH(node.right_); PUSH(node.left_); }
            } else if (hitL) PUSH(node.left_);
            else if (hitR) PUSH(node.right_);
        }
    }
    return r;
}

}  // namespace orc

using namespace orc;

// The SDF marching epsilon of a standalone ray, as the core's traversal entry
// points take it: tmin / 2 for closest hit (tmin = INTERSECT_EPS of the
// depth), SDF_EPS(0) for any hit or tmin 0.
float SdfEpsOfRay(float tmin, bool anyHit) { return (!anyHit && tmin > 0.0f) ? tmin * 0.5f : 0.5f * 1.0e-3f; }

struct or_bvh {
    const ap_scene_object* objs;  // borrowed
    size_t count;
    BVHHost bvh;
};

extern "C" {

or_bvh* or_bvh_build(const ap_scene_object* objs, size_t n, int leaf_size, int bin_count) {
    or_bvh* h = new or_bvh();
    h->objs = objs;
    h->count = n;
    std::vector<PrimitiveProxy> px(n);
    for (size_t i = 0; i < n; ++i)
        px[i] = {F(objs[i].proxy.bound_min), F(objs[i].proxy.bound_max), F(objs[i].proxy.center)};
    h->bvh = BuildBVH(px, leaf_size, bin_count);
    return h;
}

void or_bvh_free(or_bvh* h) { delete h; }

// FillSDFVolume (sdf_cache_manager.cuh:15-34) with the torus (functor 0) or
// diamond (1) functor: out[z][y][x] at (i / (n - 1)) - 0.5.
void or_sdf_fill(int functor, const float* info, int nx, int ny, int nz, float* out) {
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                const float fx = nx > 1 ? static_cast<float>(x) / static_cast<float>(nx - 1) : 0.0f;
                const float fy = ny > 1 ? static_cast<float>(y) / static_cast<float>(ny - 1) : 0.0f;
                const float fz = nz > 1 ? static_cast<float>(z) / static_cast<float>(nz - 1) : 0.0f;
                const float3 u{fx - 0.5f, fy - 0.5f, fz - 0.5f};
                out[static_cast<size_t>(z) * nx * ny + static_cast<size_t>(y) * nx + x] =
                    functor == 0 ? SdfTorus(u, info[0], info[1]) : SdfDiamondUnit(u, info);
            }
}

// The SDF cache's volumes (n volumes, dims[3k..3k+2], samples concatenated in
// volume order), replacing any registered before.
void or_sdf_set_volumes(const float* values, const int32_t* dims, size_t n) {
    g_volumes.clear();
    size_t off = 0;
    for (size_t k = 0; k < n; ++k) {
        OrVolume V;
        for (int a = 0; a < 3; ++a) V.dim[a] = dims[3 * k + a];
        const size_t m = static_cast<size_t>(V.dim[0]) * V.dim[1] * V.dim[2];
        V.v.assign(values + off, values + off + m);
        off += m;
        // voxelSizeMax_ = |1 / max(dim, 1)| x SDF_VOLUME_USE_NEAREST_SCALE (sdf_volume.cuh:24-44).
        const float3 inv{1.0f / std::fmax(static_cast<float>(V.dim[0]), 1.0f),
                         1.0f / std::fmax(static_cast<float>(V.dim[1]), 1.0f),
                         1.0f / std::fmax(static_cast<float>(V.dim[2]), 1.0f)};
        V.voxelSizeMax = length(inv) * 2.5f;
        g_volumes.push_back(std::move(V));
    }
}

size_t or_bvh_node_count(const or_bvh* h) { return h->bvh.nodes_.size(); }

void or_bvh_export(const or_bvh* h, ap_bvh_node* nodes, int32_t* indices) {
    for (size_t i = 0; i < h->bvh.nodes_.size(); ++i) {
        const BVHNode& s = h->bvh.nodes_[i];
        ap_bvh_node& d = nodes[i];
        std::memset(&d, 0, sizeof(d));
        d.bound_min = {s.boundMin_.x, s.boundMin_.y, s.boundMin_.z};
        d.bound_max = {s.boundMax_.x, s.boundMax_.y, s.boundMax_.z};
        d.left = s.left_; d.right = s.right_; d.first = s.first_; d.count = s.count_;
    }
    for (size_t i = 0; i < h->bvh.objectIndices_.size(); ++i) indices[i] = h->bvh.objectIndices_[i];
}

void or_trace(const or_bvh* h, const ap_ray* rays, size_t n, int any_hit, int cull, float margin, int32_t* ids,
              float* ts, uint32_t* nodes, uint32_t* prims, int threads) {
    const int64_t nn = static_cast<int64_t>(n);
#pragma omp parallel for schedule(dynamic, 64) num_threads(threads > 0 ? threads : 1)
    for (int64_t i = 0; i < nn; ++i) {
        const ap_ray& r = rays[i];
        const TraceResult t = TraverseBVH(any_hit != 0, h->objs, h->bvh.nodes_, h->bvh.objectIndices_,
                                          {r.ox, r.oy, r.oz}, {r.dx, r.dy, r.dz}, r.tmin, r.tmax, cull != 0,
                                          margin, SdfEpsOfRay(r.tmin, any_hit != 0));
        ids[i] = t.id;
        ts[i] = t.t;
        if (nodes) nodes[i] = t.nodes;
        if (prims) prims[i] = t.prims;
    }
}

// Intersect with soloObjectIndex < 0 (intersection.cuh:194-204): every object
// in index order, keeping a hit whenever it lies in [tMin, closest).
void or_brute_closest(const ap_scene_object* objs, size_t count, const ap_ray* rays, size_t n, int32_t* ids,
                      float* ts, int threads) {
    const int64_t nn = static_cast<int64_t>(n);
#pragma omp parallel for schedule(dynamic, 16) num_threads(threads > 0 ? threads : 1)
    for (int64_t i = 0; i < nn; ++i) {
        const ap_ray& r = rays[i];
        int id = -1;
        float best = r.tmax;
        for (size_t k = 0; k < count; ++k) {
            float t;
            if (Intersector(objs, static_cast<int>(k), {r.ox, r.oy, r.oz}, {r.dx, r.dy, r.dz}, r.tmin, best, t,
                            SdfEpsOfRay(r.tmin, false))) {
                best = t;
                id = static_cast<int>(k);
            }
        }
        ids[i] = id;
        ts[i] = best;
    }
}

// World hit point and normal of an SDF object along a ray (IntersectSDFObject,
// sdf_base.cuh:138-162) with marching epsilon eps; false on a miss.
int or_sdf_hit(const ap_scene_object* objs, int idx, const ap_ray* r, float eps, float* pos, float* normal) {
    const ap_scene_object& ob = objs[idx];
    float t;
    float3 lp;
    const float3 o{r->ox, r->oy, r->oz}, d{r->dx, r->dy, r->dz};
    if (ob.type != AP_OBJ_SDF || !IntersectSDFT(o, d, ob, eps, t, &lp)) return 0;
    const float3 p = M(ob.object_to_world) * lp + F(ob.center);
    const float3 n = normalize(M(ob.object_to_world) * SDFLocalNormal(ob, lp, eps));
    pos[0] = p.x; pos[1] = p.y; pos[2] = p.z;
    normal[0] = n.x; normal[1] = n.y; normal[2] = n.z;
    return 1;
}

// Distance of one object along one ray, or NaN when the object is missed or
// outside [tmin, tmax).
float or_object_t(const ap_scene_object* objs, int idx, const ap_ray* r) {
    float t;
    if (Intersector(objs, idx, {r->ox, r->oy, r->oz}, {r->dx, r->dy, r->dz}, r->tmin, r->tmax, t,
                    SdfEpsOfRay(r->tmin, false)))
        return t;
    return std::numeric_limits<float>::quiet_NaN();
}

}  // extern "C"

namespace orc {
// One traversal for the integrator restatement (or_render.cpp).
void TraceOne(const or_bvh* h, float3 o, float3 d, float tMin, float tMax, bool anyHit, bool cull, float margin,
              int* id, float* t, float sdfEps) {
    const TraceResult r =
        TraverseBVH(anyHit, h->objs, h->bvh.nodes_, h->bvh.objectIndices_, o, d, tMin, tMax, cull, margin, sdfEps);
    *id = r.id;
    *t = r.t;
    Count(anyHit ? CNT_SHADOW_RAYS : CNT_EXT_RAYS);
    Count(anyHit ? CNT_SHADOW_NODES : CNT_EXT_NODES, r.nodes);
    Count(anyHit ? CNT_SHADOW_PRIMS : CNT_EXT_PRIMS, r.prims);
}

std::atomic<bool> g_counting{false};
std::atomic<uint64_t> g_counts[CNT_N];
}  // namespace orc

extern "C" {

void or_counts_begin(void) {
    for (auto& c : orc::g_counts) c.store(0, std::memory_order_relaxed);
    orc::g_counting.store(true);
}

void or_counts_end(uint64_t* out) {
    orc::g_counting.store(false);
    for (int k = 0; k < orc::CNT_N; ++k) out[k] = orc::g_counts[k].load();
}

}  // extern "C"

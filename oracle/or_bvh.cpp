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

#include "aperture/ap_layout.h"
#include "or_api.h"
#include "or_math.h"

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

// IntersectBase + SceneObjectsIntersector acceptance (intersection.cuh:93-172).
bool Intersector(const ap_scene_object* objs, int idx, float3 o, float3 d, float tMin, float tMax, float& tOut) {
    const ap_scene_object& ob = objs[idx];
    float t = 0.0f;
    bool hit = false, inside = false;
    float u, v;
    if (ob.type == AP_OBJ_CUBE) hit = IntersectAABBT(o, d, ob, t, inside);
    else if (ob.type == AP_OBJ_SPHERE) hit = IntersectSphereT(o, d, ob, t, inside);
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
                        bool cull, float margin) {
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
                if (Intersector(objs, objectIndex, o, d, tMin, r.t, t)) {
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
                                          margin);
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
            if (Intersector(objs, static_cast<int>(k), {r.ox, r.oy, r.oz}, {r.dx, r.dy, r.dz}, r.tmin, best, t)) {
                best = t;
                id = static_cast<int>(k);
            }
        }
        ids[i] = id;
        ts[i] = best;
    }
}

// Distance of one object along one ray, or NaN when the object is missed or
// outside [tmin, tmax).
float or_object_t(const ap_scene_object* objs, int idx, const ap_ray* r) {
    float t;
    if (Intersector(objs, idx, {r->ox, r->oy, r->oz}, {r->dx, r->dy, r->dz}, r->tmin, r->tmax, t)) return t;
    return std::numeric_limits<float>::quiet_NaN();
}

}  // extern "C"

namespace orc {
// One traversal for the integrator restatement (or_render.cpp).
void TraceOne(const or_bvh* h, float3 o, float3 d, float tMin, float tMax, bool anyHit, bool cull, float margin,
              int* id, float* t) {
    const TraceResult r =
        TraverseBVH(anyHit, h->objs, h->bvh.nodes_, h->bvh.objectIndices_, o, d, tMin, tMax, cull, margin);
    *id = r.id;
    *t = r.t;
}
}  // namespace orc

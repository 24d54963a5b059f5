// Host BVH builder: binned SAH over SceneObject proxies.
//
// Produces exactly the node array and object permutation of the reference
// BuildBVH (core/bvh/bvh.cuh:142-422) in its 32-bit node variant: DFS
// pre-order numbering with the root at 0, leaves for count <= leafSize, the
// longest centroid axis, binCount bins, costI = costT = 1 with the first
// minimal split kept, partition by bin with std::partition, and the
// std::nth_element median split when the SAH split is degenerate. The
// traversal parity of the core depends on this output being identical, so
// every FP32 expression keeps the reference's operand order.
#include <algorithm>
#include <cfloat>
#include <vector>

#include "ap_internal.h"

namespace ap {
namespace {

struct Box {
    f3 lo{FLT_MAX, FLT_MAX, FLT_MAX};
    f3 hi{-FLT_MAX, -FLT_MAX, -FLT_MAX};
};

inline f3 F(const ap_float3& a) { return {a.x, a.y, a.z}; }
inline float axis_of(const ap_float3& c, int axis) { return axis == 0 ? c.x : (axis == 1 ? c.y : c.z); }

// AABBArea (bvh.cuh:90-94).
inline float box_area(f3 lo, f3 hi) {
    const f3 d{fmaxf(0.0f, hi.x - lo.x), fmaxf(0.0f, hi.y - lo.y), fmaxf(0.0f, hi.z - lo.z)};
    return 2.0f * (d.x * d.y + d.y * d.z + d.z * d.x);
}

class Builder {
public:
    Builder(const std::vector<ap_primitive_proxy>& p, int leaf, int bins, HostBVH& out)
        : proxy_(p), leaf_(leaf), bins_(bins), out_(out) {}

    int node(int start, int end) {
        const int count = end - start;
        const int idx = static_cast<int>(out_.nodes.size());
        out_.nodes.emplace_back();
        // Primitive and centroid bounds of the range (BoundsAccumulator, bvh.cuh:120-140).
        Box prim, cen;
        for (int i = start; i < end; ++i) {
            const ap_primitive_proxy& q = proxy_[out_.indices[i]];
            prim.lo = fmin3(prim.lo, F(q.bound_min));
            prim.hi = fmax3(prim.hi, F(q.bound_max));
            cen.lo = fmin3(cen.lo, F(q.center));
            cen.hi = fmax3(cen.hi, F(q.center));
        }
        {
            ap_bvh_node& n = out_.nodes[idx];
            n.bound_min = {prim.lo.x, prim.lo.y, prim.lo.z};
            n.bound_max = {prim.hi.x, prim.hi.y, prim.hi.z};
        }
        if (count <= leaf_) return make_leaf(idx, start, count);

        const f3 ext{cen.hi.x - cen.lo.x, cen.hi.y - cen.lo.y, cen.hi.z - cen.lo.z};
        int axis = 0;
        if (ext.y > ext.x && ext.y >= ext.z) axis = 1;
        else if (ext.z > ext.x && ext.z >= ext.y) axis = 2;
        if (ext.x <= 1e-6f && ext.y <= 1e-6f && ext.z <= 1e-6f) return make_leaf(idx, start, count);

        const float cmin = axis == 0 ? cen.lo.x : (axis == 1 ? cen.lo.y : cen.lo.z);
        const float cmax = axis == 0 ? cen.hi.x : (axis == 1 ? cen.hi.y : cen.hi.z);
        const float invw = (cmax > cmin) ? static_cast<float>(bins_) / (cmax - cmin) : 0.0f;
        auto bin_of = [&](int obj) {
            const float key = axis_of(proxy_[obj].center, axis);
            int b = invw > 0.0f ? static_cast<int>((key - cmin) * invw) : 0;
            return b < 0 ? 0 : (b >= bins_ ? bins_ - 1 : b);
        };

        bin_box_.assign(bins_, Box{});
        bin_cnt_.assign(bins_, 0);
        for (int i = start; i < end; ++i) {
            const int obj = out_.indices[i];
            const int b = bin_of(obj);
            bin_box_[b].lo = fmin3(bin_box_[b].lo, F(proxy_[obj].bound_min));
            bin_box_[b].hi = fmax3(bin_box_[b].hi, F(proxy_[obj].bound_max));
            bin_cnt_[b]++;
        }
        // Prefix (left) and suffix (right) sweeps (bvh.cuh:278-308).
        left_area_.resize(bins_); right_area_.resize(bins_);
        left_cnt_.resize(bins_); right_cnt_.resize(bins_);
        {
            Box run; int c = 0;
            for (int b = 0; b < bins_; ++b) {
                if (bin_cnt_[b] > 0) { run.lo = fmin3(run.lo, bin_box_[b].lo); run.hi = fmax3(run.hi, bin_box_[b].hi); }
                c += bin_cnt_[b];
                left_area_[b] = c > 0 ? box_area(run.lo, run.hi) : 0.0f;
                left_cnt_[b] = c;
            }
        }
        {
            Box run; int c = 0;
            for (int b = bins_ - 1; b >= 0; --b) {
                if (bin_cnt_[b] > 0) { run.lo = fmin3(run.lo, bin_box_[b].lo); run.hi = fmax3(run.hi, bin_box_[b].hi); }
                c += bin_cnt_[b];
                right_area_[b] = c > 0 ? box_area(run.lo, run.hi) : 0.0f;
                right_cnt_[b] = c;
            }
        }
        const float parent_area = box_area(prim.lo, prim.hi);
        float best = FLT_MAX;
        int split = -1;
        for (int b = 0; b < bins_ - 1; ++b) {
            const int nl = left_cnt_[b], nr = right_cnt_[b + 1];
            if (nl == 0 || nr == 0) continue;
            const float sah = 1.0f + (left_area_[b] * static_cast<float>(nl) +
                                      right_area_[b + 1] * static_cast<float>(nr)) /
                                         fmaxf(parent_area, 1e-20f) * 1.0f;
            if (sah < best) { best = sah; split = b; }
        }

        int mid = -1;
        auto first = out_.indices.begin();
        if (split >= 0) {
            auto it = std::partition(first + start, first + end,
                                     [&](int obj) { return bin_of(obj) <= split; });
            mid = static_cast<int>(it - first);
            if (mid == start || mid == end) split = -1;
        }
        if (split < 0) {
            const int m = start + count / 2;
            std::nth_element(first + start, first + m, first + end, [&](int a, int b) {
                return axis_of(proxy_[a].center, axis) < axis_of(proxy_[b].center, axis);
            });
            mid = m;
            if (mid == start || mid == end) return make_leaf(idx, start, count);
        }
        const int l = node(start, mid);
        const int r = node(mid, end);
        ap_bvh_node& n = out_.nodes[idx];
        n.left = l; n.right = r; n.first = -1; n.count = 0;
        return idx;
    }

private:
    int make_leaf(int idx, int start, int count) {
        ap_bvh_node& n = out_.nodes[idx];
        n.left = -1; n.right = -1; n.first = start; n.count = count;
        return idx;
    }

    const std::vector<ap_primitive_proxy>& proxy_;
    const int leaf_, bins_;
    HostBVH& out_;
    std::vector<Box> bin_box_;
    std::vector<int> bin_cnt_, left_cnt_, right_cnt_;
    std::vector<float> left_area_, right_area_;
};

}  // namespace

HostBVH build_bvh(const std::vector<ap_primitive_proxy>& proxies, int leaf_size, int bin_count) {
    HostBVH out;
    if (proxies.empty()) return out;
    bin_count = std::max(4, std::min(64, bin_count));
    leaf_size = std::max(1, std::min(16, leaf_size));
    const int n = static_cast<int>(proxies.size());
    out.indices.resize(n);
    for (int i = 0; i < n; ++i) out.indices[i] = i;
    out.nodes.reserve(2 * static_cast<size_t>(n));
    Builder b(proxies, leaf_size, bin_count, out);
    b.node(0, n);
    return out;
}

int bvh_max_depth(const HostBVH& bvh) {
    if (bvh.nodes.empty()) return 0;
    int best = 0;
    std::vector<std::pair<int, int>> st{{0, 0}};
    while (!st.empty()) {
        const auto [i, d] = st.back();
        st.pop_back();
        best = std::max(best, d);
        const ap_bvh_node& n = bvh.nodes[i];
        if (n.count == 0) { st.push_back({n.left, d + 1}); st.push_back({n.right, d + 1}); }
    }
    return best;
}

}  // namespace ap

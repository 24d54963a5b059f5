// ORACLE — test infrastructure only. Visit accounting of the oracle's
// integrator: while counting is on (or_counts_begin .. or_counts_end), each
// traversal, shading vertex, spectral Interact and probe query adds to a
// process-wide total. The totals price a frame in algorithmic bytes (node
// records, triangle tests, material and LUT fetches, probe texels), the
// per-vertex sum of the integrator's memory roofline.
#pragma once

#include <atomic>
#include <cstdint>

namespace orc {

enum OrCount {
    CNT_EXT_RAYS,       // closest-hit traversals (TraceOne without anyHit, solo traces in a medium)
    CNT_EXT_NODES,      // node records read by them (each popped node, both children of an expanded one)
    CNT_EXT_PRIMS,      // primitive tests by them
    CNT_SHADOW_RAYS,    // any-hit traversals (SegmentTrace)
    CNT_SHADOW_NODES,
    CNT_SHADOW_PRIMS,
    CNT_HITS,           // surface vertices shaded (a closest hit that was found)
    CNT_INTERACTS,      // lut::Interact evaluations
    CNT_PROBE_QUERIES,  // QueryProbe calls at guided vertices
    CNT_PROBE_FOUND,    // of which returned a probe
    CNT_PATHS,          // camera paths (samples)
    CNT_N
};

extern std::atomic<bool> g_counting;
extern std::atomic<uint64_t> g_counts[CNT_N];

inline void Count(int k, uint64_t n = 1) {
    if (g_counting.load(std::memory_order_relaxed)) g_counts[k].fetch_add(n, std::memory_order_relaxed);
}

}  // namespace orc

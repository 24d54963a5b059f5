// ORACLE — test infrastructure only. C entry points of liboracle.so, the
// CPU restatement of the reference algorithms used as the checker by
// tests/, bench.py's cpu_baseline leg and __graft_entry__.smoke().
#pragma once

#include <stddef.h>
#include <stdint.h>

#include "aperture/ap_core.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_bvh or_bvh;

or_bvh* or_bvh_build(const ap_scene_object* objs, size_t n, int leaf_size, int bin_count);
void or_bvh_free(or_bvh* h);
size_t or_bvh_node_count(const or_bvh* h);
void or_bvh_export(const or_bvh* h, ap_bvh_node* nodes, int32_t* indices);
/* cull = 0: the reference TraverseBVH; cull = 1: with the core's CULLED
 * line rejection of boxes grown by margin. */
void or_trace(const or_bvh* h, const ap_ray* rays, size_t n, int any_hit, int cull, float margin,
              int32_t* ids, float* ts, uint32_t* nodes, uint32_t* prims, int threads);
void or_brute_closest(const ap_scene_object* objs, size_t count, const ap_ray* rays, size_t n,
                      int32_t* ids, float* ts, int threads);
float or_object_t(const ap_scene_object* objs, int idx, const ap_ray* r);

#ifdef __cplusplus
}
#endif

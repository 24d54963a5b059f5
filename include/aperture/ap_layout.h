/*
 * Byte layouts that cross the Aperture core boundary.
 *
 * These are plain C structs whose sizes and field offsets equal the
 * reference renderer's host/device types, so that an existing host
 * (scene builders, GUI, file writers) can hand its arrays to the core
 * unchanged. Offsets assume the CUDA vector-type alignments the reference
 * was compiled with: float2 = 8, float4/int4 = 16, float3 = 4.
 *
 *   GBuffer            96 B   core/render/object.cuh:60-91
 *   Texture / Material 112 / 128 B   core/render/object.cuh:119-217
 *   SceneObject        432 B  core/render/object.cuh:254-365
 *   BVHNode (32-bit)   48 B   core/bvh/bvh.cuh:22-44 (BVH_USE_16_BITS_NODE off)
 *   AdditionalLightInfo 12 B  core/render/object.cuh:219-224
 *   SceneSetting       13 B   core/render/object.cuh:10-27
 *   PathGuidingSample  64 B   core/guiding/guiding.cuh:22-32
 *   Octahedron (probe) 1392 B core/guiding/octahedron.cuh:203-231
 *   ScreenGBuffer      64 B   core/denoise/denoise.cuh:33-82
 *   ScreenStatisticsBuffer 48 B core/denoise/denoise_statistic.cuh:11-24
 *
 * Texture handles (cudaTextureObject_t in the reference) are carried as
 * uint64_t; the core treats any non-zero handle as "not supported" and
 * reports an error at scene creation.
 */
#ifndef APERTURE_AP_LAYOUT_H
#define APERTURE_AP_LAYOUT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
#define AP_ALIGN(n) alignas(n)
#define AP_STATIC_ASSERT(c, m) static_assert(c, m)
extern "C" {
#else
#define AP_ALIGN(n) _Alignas(n)
#define AP_STATIC_ASSERT(c, m) _Static_assert(c, m)
#endif

typedef struct ap_float2 { AP_ALIGN(8) float x; float y; } ap_float2;
typedef struct ap_float3 { float x, y, z; } ap_float3;
typedef struct ap_float4 { AP_ALIGN(16) float x; float y, z, w; } ap_float4;
typedef struct ap_int4 { AP_ALIGN(16) int x; int y, z, w; } ap_int4;
typedef struct ap_float3x3 { ap_float3 x, y, z; } ap_float3x3; /* rows */

/* EObjectType (core/render/object.cuh:29-35) */
enum { AP_OBJ_CUBE = 0, AP_OBJ_SPHERE = 1, AP_OBJ_SDF = 2, AP_OBJ_TRIANGLE = 3 };
/* EShadingModel (core/render/object.cuh:45-52) */
enum { AP_MAT_DIFFUSE = 0, AP_MAT_LIGHT = 1, AP_MAT_GLASS = 2, AP_MAT_GGX = 3 };

/* GBuffer, __align__(16), 96 B (core/render/object.cuh:60-91). */
typedef struct ap_gbuffer {
    AP_ALIGN(16) ap_float3 albedo; /* 0  */
    ap_float3 emissive;            /* 12 */
    float specular;                /* 24 */
    float metallic;                /* 28 */
    float roughness;               /* 32 */
    float ior;                     /* 36 */
    float sigma_a;                 /* 40 */
    ap_float3 normal;              /* 44 */
    float abbe_number;             /* 56 */
    float tilt_number;             /* 60 */
    uint64_t albedo_tex;           /* 64 */
    uint64_t normal_tex;           /* 72 */
    uint64_t metallic_tex;         /* 80 */
    uint64_t roughness_tex;        /* 88 */
} ap_gbuffer;

typedef struct ap_texture {        /* 112 B */
    int32_t texture_type;          /* 0  */
    ap_gbuffer base;               /* 16 */
} ap_texture;

typedef struct ap_material {       /* 128 B */
    int32_t shading_model;         /* 0  */
    ap_texture texture;            /* 16 */
} ap_material;

typedef struct ap_vertex {         /* 32 B, align 8 */
    ap_float3 position;            /* 0  */
    ap_float3 normal;              /* 12 */
    ap_float2 uv;                  /* 24 */
} ap_vertex;

typedef struct ap_sdf_info {       /* 32 B, align 16 */
    ap_float4 sdf_info0;
    int32_t cached_volume_index;
} ap_sdf_info;

typedef struct ap_additional_object_info { /* 144 B, align 16 */
    int32_t sdf_type;              /* 0  */
    ap_sdf_info sdf_info;          /* 16 */
    ap_vertex v0, v1, v2;          /* 48, 80, 112 */
} ap_additional_object_info;

typedef struct ap_primitive_proxy { /* 36 B (core/bvh/bvh.cuh:15-20) */
    ap_float3 bound_min;
    ap_float3 bound_max;
    ap_float3 center;
} ap_primitive_proxy;

/* SceneObject, 432 B (core/render/object.cuh:254-272). */
typedef struct ap_scene_object {
    ap_float3 center;                         /* 0   */
    ap_float3 extent;                         /* 12  */
    ap_material material;                     /* 32  */
    ap_float3x3 world_to_object;              /* 160 */
    ap_float3x3 object_to_world;              /* 196 */
    int32_t type;                             /* 232 */
    ap_additional_object_info additional;     /* 240 */
    ap_primitive_proxy proxy;                 /* 384 */
    int32_t object_index;                     /* 420 */
    int32_t light_index;                      /* 424 */
    uint8_t visible;                          /* 428 (bool) */
} ap_scene_object;

/* BVHNode, 32-bit link variant, 48 B (core/bvh/bvh.cuh:36-43). */
typedef struct ap_bvh_node {
    AP_ALIGN(16) ap_float3 bound_min;
    ap_float3 bound_max;
    int32_t left, right, first, count;
} ap_bvh_node;

/* AdditionalLightInfo, 12 B (core/render/object.cuh:219-224). */
typedef struct ap_light_info {
    int32_t object_index;
    float importance_weight;
    float select_prob;
} ap_light_info;

/* SceneSetting, 13 bools (core/render/object.cuh:10-27). */
typedef struct ap_scene_setting {
    uint8_t use_bvh, use_nee, use_path_guiding, denoise;
    uint8_t debug_albedo, debug_normal, debug_depth, debug_metallic, debug_roughness;
    uint8_t debug_stat_num, debug_stat_mean, debug_stat_m2, debug_stat_m3;
} ap_scene_setting;

/* pg::PathGuidingSample, 64 B (core/guiding/guiding.cuh:22-32). */
typedef struct ap_path_guiding_sample {
    ap_float3 radiance;        /* 0  XYZ */
    ap_float3 direction;       /* 12 */
    ap_int4 voxel_grid;        /* 32 */
    uint8_t valid;             /* 48 */
} ap_path_guiding_sample;

/* pg::Texture<4> (Octahedron), 1392 B (core/guiding/octahedron.cuh:203-231). */
typedef struct ap_octahedron {
    ap_float4 data[86];        /* 8x8 leaf + 4x4 + 2x2 + 1 + pad */
    uint64_t count;            /* 1376 */
} ap_octahedron;

/* denoise::ScreenGBuffer, 64 B, align 16 (core/denoise/denoise.cuh:33-82):
 * per-pixel running means over frames of the camera ray's first hit. The
 * half fields are IEEE binary16 bit patterns. shading_model[k] is the
 * fraction of frames whose first hit had shading model k (diffuse, light,
 * glass, GGX), [4] the fraction that missed. */
typedef struct ap_screen_gbuffer {
    AP_ALIGN(16) uint16_t albedo[3];    /* 0 */
    uint16_t normal[3];                 /* 6 */
    uint16_t metallic;                  /* 12 */
    uint16_t roughness;                 /* 14 */
    float depth;                        /* 16 */
    uint64_t valid_frame_index;         /* 24 frames that hit */
    uint64_t frame_index;               /* 32 frames */
    AP_ALIGN(16) uint16_t shading_model[5]; /* 48 */
} ap_screen_gbuffer;

/* denoise::ScreenStatisticsBuffer, 48 B, align 16
 * (core/denoise/denoise_statistic.cuh:11-24): per-pixel count, mean and
 * second and third central moments (divided by the count) of the Box-Cox
 * transformed samples. */
typedef struct ap_screen_statistics {
    AP_ALIGN(16) uint32_t num;          /* 0 */
    ap_float3 mean;                     /* 4 */
    ap_float3 m2_div_num;               /* 16 */
    ap_float3 m3_div_num;               /* 28 */
} ap_screen_statistics;

AP_STATIC_ASSERT(sizeof(ap_gbuffer) == 96, "GBuffer must be 96 B");
AP_STATIC_ASSERT(offsetof(ap_gbuffer, normal) == 44, "GBuffer.normal_ at 44");
AP_STATIC_ASSERT(offsetof(ap_gbuffer, albedo_tex) == 64, "GBuffer.albedoTex_ at 64");
AP_STATIC_ASSERT(sizeof(ap_texture) == 112, "Texture must be 112 B");
AP_STATIC_ASSERT(sizeof(ap_material) == 128, "Material must be 128 B");
AP_STATIC_ASSERT(sizeof(ap_additional_object_info) == 144, "AdditionalObjectInfo must be 144 B");
AP_STATIC_ASSERT(sizeof(ap_scene_object) == 432, "SceneObject must be 432 B");
AP_STATIC_ASSERT(offsetof(ap_scene_object, material) == 32, "material_ at 32");
AP_STATIC_ASSERT(offsetof(ap_scene_object, world_to_object) == 160, "worldToObject_ at 160");
AP_STATIC_ASSERT(offsetof(ap_scene_object, object_to_world) == 196, "objectToWorld_ at 196");
AP_STATIC_ASSERT(offsetof(ap_scene_object, type) == 232, "type_ at 232");
AP_STATIC_ASSERT(offsetof(ap_scene_object, additional) == 240, "additionalObjectInfo_ at 240");
AP_STATIC_ASSERT(offsetof(ap_scene_object, additional) + offsetof(ap_additional_object_info, v0) == 288, "v0 at 288");
AP_STATIC_ASSERT(offsetof(ap_scene_object, proxy) == 384, "proxy_ at 384");
AP_STATIC_ASSERT(offsetof(ap_scene_object, object_index) == 420, "objectIndex_ at 420");
AP_STATIC_ASSERT(offsetof(ap_scene_object, light_index) == 424, "lightIndex_ at 424");
AP_STATIC_ASSERT(offsetof(ap_scene_object, visible) == 428, "bVisible_ at 428");
AP_STATIC_ASSERT(sizeof(ap_bvh_node) == 48, "32-bit BVHNode must be 48 B");
AP_STATIC_ASSERT(sizeof(ap_light_info) == 12, "AdditionalLightInfo must be 12 B");
AP_STATIC_ASSERT(sizeof(ap_scene_setting) == 13, "SceneSetting must be 13 B");
AP_STATIC_ASSERT(sizeof(ap_path_guiding_sample) == 64, "PathGuidingSample must be 64 B");
AP_STATIC_ASSERT(sizeof(ap_octahedron) == 1392, "Octahedron must be 1392 B");
AP_STATIC_ASSERT(sizeof(ap_screen_gbuffer) == 64, "ScreenGBuffer must be 64 B");
AP_STATIC_ASSERT(offsetof(ap_screen_gbuffer, depth) == 16, "ScreenGBuffer.depth_ at 16");
AP_STATIC_ASSERT(offsetof(ap_screen_gbuffer, valid_frame_index) == 24, "ScreenGBuffer.validFrameIndex_ at 24");
AP_STATIC_ASSERT(offsetof(ap_screen_gbuffer, frame_index) == 32, "ScreenGBuffer.frameIndex_ at 32");
AP_STATIC_ASSERT(offsetof(ap_screen_gbuffer, shading_model) == 48, "ScreenGBuffer.shadingModel_ at 48");
AP_STATIC_ASSERT(sizeof(ap_screen_statistics) == 48, "ScreenStatisticsBuffer must be 48 B");
AP_STATIC_ASSERT(offsetof(ap_screen_statistics, m3_div_num) == 28, "ScreenStatisticsBuffer.M3DivNum_ at 28");

#ifdef __cplusplus
}
#endif
#endif /* APERTURE_AP_LAYOUT_H */

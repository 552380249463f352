/*
 * lrt_types.h - plain data types and constants shared by the C ABI (lrt.h),
 * the C++ facade and the CPU oracle.
 *
 * Every struct here is binary-compatible with the LiteRT struct of the same
 * name, so a caller that already fills the reference types can pass them
 * through unchanged:
 *   SdfSBSHeader, SdfSBSNode      sdfScene/sdf_scene.h:136-161
 *   MultiRenderPreset + enums     render_settings.h:1-108
 *   BVHNode, BVHNodePair          BVH/cbvh.h:21-33
 *   CRT_Hit                       fields as used in BVH/BVH2Common.cpp:3497-3505
 *   MultiRendererDRPreset + enums diff_render/DR_common.h:75-195
 * Leaf/offset encoding constants follow raytrace_common.h:36-57.
 */
#ifndef LRT_TYPES_H
#define LRT_TYPES_H

#include <stdint.h>
#ifndef __cplusplus
#include <stdbool.h>
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* ---- geometry types (render_settings.h:4-31) ---------------------------- */
#define LRT_TYPE_SDF_SBS      6u
#define LRT_TYPE_SDF_SBS_TEX 10u
#define LRT_TYPE_SDF_SBS_COL 11u
#define LRT_SH_TYPE          24u
#define LRT_GEOM_ID_MASK     0x00FFFFFFu

/* ---- devices (render_settings.h:33-36) ---------------------------------- */
#define LRT_DEVICE_CPU     0u
#define LRT_DEVICE_GPU     1u
#define LRT_DEVICE_GPU_RTX 2u

/* ---- SdfNodeIntersect (render_settings.h:40-45) ------------------------- */
#define LRT_SDF_NODE_INTERSECT_ST       0u
#define LRT_SDF_NODE_INTERSECT_ANALYTIC 1u
#define LRT_SDF_NODE_INTERSECT_NEWTON   2u
#define LRT_SDF_NODE_INTERSECT_BBOX     3u
#define LRT_SDF_NODE_INTERSECT_IT       4u

/* ---- MultiRenderMode (render_settings.h:47-67) -------------------------- */
#define LRT_RENDER_MODE_MASK                 0u
#define LRT_RENDER_MODE_LAMBERT_NO_TEX       1u
#define LRT_RENDER_MODE_DEPTH                2u
#define LRT_RENDER_MODE_LINEAR_DEPTH         3u
#define LRT_RENDER_MODE_INVERSE_LINEAR_DEPTH 4u
#define LRT_RENDER_MODE_PRIMITIVE            5u
#define LRT_RENDER_MODE_TYPE                 6u
#define LRT_RENDER_MODE_GEOM                 7u
#define LRT_RENDER_MODE_NORMAL               8u
#define LRT_RENDER_MODE_BARYCENTRIC          9u
#define LRT_RENDER_MODE_ST_ITERATIONS       10u
#define LRT_RENDER_MODE_PHONG_NO_TEX        12u
#define LRT_RENDER_MODE_TEX_COORDS          15u
#define LRT_RENDER_MODE_DIFFUSE             16u
#define LRT_RENDER_MODE_LAMBERT             17u
#define LRT_RENDER_MODE_PHONG               18u

/* ---- NormalMode / RayGenMode / Interpolation / Representation ----------- */
#define LRT_NORMAL_MODE_GEOMETRY       0u
#define LRT_NORMAL_MODE_VERTEX         1u
#define LRT_NORMAL_MODE_SDF_SMOOTHED   2u
#define LRT_RAY_GEN_MODE_REGULAR       0u
#define LRT_RAY_GEN_MODE_RANDOM        1u
#define LRT_INTERPOLATION_MODE_TRILINEAR 0u
#define LRT_INTERPOLATION_MODE_TRICUBIC  1u
#define LRT_REPRESENTATION_MODE_SURFACE  0u
#define LRT_REPRESENTATION_MODE_VOLUME   1u

/* ---- SdfSBSNodeLayout (sdf_scene.h:52-61) ------------------------------- */
#define LRT_SBS_LAYOUT_UNDEFINED        (0u << 24)
#define LRT_SBS_LAYOUT_DX               (1u << 24)
#define LRT_SBS_LAYOUT_DX_UV16          (2u << 24)
#define LRT_SBS_LAYOUT_DX_RGB8          (3u << 24)
#define LRT_SBS_LAYOUT_ID32F_IRGB32F    (4u << 24)
#define LRT_SBS_LAYOUT_ID32F_IRGB32F_IN (5u << 24)
#define LRT_SBS_LAYOUT_MASK             0xFF000000u

/* distance fields of lrt_sbs_octree_from_primitive_device (no reference counterpart) */
#define LRT_PRIMITIVE_SPHERE 0u
#define LRT_PRIMITIVE_TORUS  1u
#define LRT_INVALID_IDX                 (1u << 31)

/* ---- BVH leaf encoding (raytrace_common.h:36-57, cbvh.h:17-19) ---------- */
#define LRT_STACK_SIZE  80
#define LRT_START_MASK  0x00FFFFFFu
#define LRT_SIZE_MASK   0x7F000000u
#define LRT_LEAF_BIT    0x80000000u
#define LRT_LEAF_NORMAL 0xFFFFFFFFu

/* ---- SBS scene description (sdf_scene.h:136-161) ------------------------ */
typedef struct SdfSBSHeader {
  uint32_t brick_size;      /* voxels per brick edge, 1..16 */
  uint32_t brick_pad;       /* extra voxels stored on each side, 0 or 1 */
  uint32_t bytes_per_value; /* 1, 2 or 4 (DX layouts) */
  uint32_t aux_data;        /* SdfSBSNodeLayout */
} SdfSBSHeader;

typedef struct SdfSBSNode {
  uint32_t pos_xy;          /* (x << 16) | y brick position in its LOD */
  uint32_t pos_z_lod_size;  /* (z << 16) | LOD size (bricks per axis) */
  uint32_t data_offset;     /* offset into the uint32 value array */
  uint32_t _pad;
} SdfSBSNode;

/* ---- render preset (render_settings.h:86-108) --------------------------- */
typedef struct MultiRenderPreset {
  uint32_t render_mode;
  uint32_t sdf_node_intersect;
  uint32_t normal_mode;
  uint32_t ray_gen_mode;
  uint32_t interpolation_mode;
  uint32_t representation_mode;
  uint32_t spp;
} MultiRenderPreset;

/* ---- BVH nodes (cbvh.h:21-33) ------------------------------------------- */
typedef struct BVHNode {
  float    boxMin[3];
  uint32_t leftOffset;   /* LEAF_BIT set => leaf: count in bits 24-30, start in 0-23 */
  float    boxMax[3];
  uint32_t escapeIndex;
} BVHNode;

typedef struct BVHNodePair {
  BVHNode left;
  BVHNode right;
} BVHNodePair;

/* ---- lights (Renderer/eye_ray.h:44-74) ----------------------------------- */
#define LRT_LIGHT_TYPE_DIRECT  0u
#define LRT_LIGHT_TYPE_POINT   1u
#define LRT_LIGHT_TYPE_AMBIENT 2u
#define LRT_MAX_LIGHTS         8u
typedef struct LrtLight {
  float    space[3];  /* direction (unit, direct) or position (point) */
  uint32_t type;
  float    color[3];  /* intensity included */
  uint32_t _pad;
} LrtLight;

/* ---- ray hit (CrossRT CRT_Hit, fields per BVH2Common.cpp:3497-3505) ----- */
typedef struct CRT_Hit {
  float    t;
  uint32_t primId;
  uint32_t instId;
  uint32_t geomId;      /* (type << 24) | geomId */
  float    coords[4];   /* [2],[3] = octahedral-encoded normal for SDF hits */
} CRT_Hit;

/* ---- differentiable rendering (diff_render/DR_common.h:75-195) ---------- */
#define LRT_DR_LOSS_FUNCTION_MSE 0u
#define LRT_DR_LOSS_FUNCTION_MAE 1u
#define LRT_DR_RENDER_MODE_DIFFUSE      0u
#define LRT_DR_RENDER_MODE_LAMBERT      1u
#define LRT_DR_RENDER_MODE_MASK         2u
#define LRT_DR_RENDER_MODE_LINEAR_DEPTH 3u
#define LRT_DR_RENDER_MODE_NORMAL       4u
#define LRT_DR_DIFF_MODE_DEFAULT     0u
#define LRT_DR_DIFF_MODE_FINITE_DIFF 1u
#define LRT_DR_RECONSTRUCTION_FLAG_COLOR    (1u << 0)
#define LRT_DR_RECONSTRUCTION_FLAG_GEOMETRY (1u << 1)
#define LRT_DR_INPUT_TYPE_COLOR        0u
#define LRT_DR_INPUT_TYPE_LINEAR_DEPTH 1u
#define LRT_DR_BORDER_SAMPLING_RANDOM 0u
#define LRT_DR_BORDER_SAMPLING_SVM 1u
#define LRT_DR_REG_FUNCTION_NONE         0u
#define LRT_DR_REG_FUNCTION_LK_DENOISING 1u
#define LRT_DR_RAY_FLAG_NO_DIFF         0u
#define LRT_DR_RAY_FLAG_DDIFFUSE_DCOLOR (1u << 0)
#define LRT_DR_RAY_FLAG_DDIFFUSE_DPOS   (1u << 1)
#define LRT_DR_RAY_FLAG_DNORM_DPOS      (1u << 2)
#define LRT_DR_RAY_FLAG_BORDER          (1u << 3)
#define LRT_DR_RAY_FLAG_DDIST_DPOS      (1u << 4)

typedef struct MultiRendererDRPreset {
  uint32_t dr_loss_function;
  uint32_t dr_render_mode;
  uint32_t dr_diff_mode;
  uint32_t dr_reconstruction_flags;
  uint32_t dr_input_type;
  uint32_t dr_border_sampling;
  uint32_t dr_raycasting_mask;
  uint32_t dr_atomic_ders;
  uint32_t spp;
  uint32_t render_width;
  uint32_t render_height;
  uint32_t border_spp;
  float    border_relax_eps;
  float    border_integral_mult;
  float    opt_lr;
  float    opt_beta_1;
  float    opt_beta_2;
  float    opt_eps;
  uint32_t opt_iterations;
  uint32_t image_batch_size;
  uint32_t reg_function;
  float    reg_lambda;
  float    reg_power;
  bool     redistancing_enable;
  uint32_t redistancing_interval;
  uint32_t dr_raycasting_thickness;
  bool     debug_print;
  uint32_t debug_print_interval;
  uint32_t debug_progress_images;
  uint32_t debug_progress_interval;
  uint32_t debug_render_mode;
  float    finite_diff_delta;
  float    finite_diff_brightness;
  bool     debug_forced_border;
  bool     debug_pd_images;
  float    debug_pd_brightness;
  bool     debug_border_samples;
  bool     debug_border_samples_mega_image;
  bool     debug_border_save_normals;
} MultiRendererDRPreset;

#ifdef __cplusplus
}
#endif

#endif /* LRT_TYPES_H */

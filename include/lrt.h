/*
 * lrt.h - C ABI of the MI355X SBS ray-casting library (liblrt_hip.so).
 *
 * Plain pointers and sizes only. Every call returns 0 on success and a
 * negative LRT_E* code on failure; lrt_last_error() gives the message of the
 * calling thread's last failure. A handle owns one HIP stream and its device
 * buffers; it is not thread-safe (neither is the reference MultiRenderer).
 * Matrices are 16 floats, column-major (LiteMath float4x4 memory layout).
 * Output images are caller-owned, W*H pixels indexed y*W + x
 * (Renderer/eye_ray.cpp:115, 141).
 */
#ifndef LRT_H
#define LRT_H
#include <stdint.h>
#include "lrt_types.h"

#ifdef __cplusplus
extern "C" {
#endif

#define LRT_OK 0
#define LRT_E_INVALID -1   /* bad argument */
#define LRT_E_NOSCENE -2   /* no scene set */
#define LRT_E_UNSUPPORTED -3
#define LRT_E_HIP -4       /* HIP runtime error */
#define LRT_E_NODEVICE -5  /* no GPU visible */

typedef struct lrt_renderer* lrt_handle;
typedef struct lrt_sbs* lrt_sbs_handle;

const char* lrt_last_error(void);

/* CreateMultiRenderer(device, maxPrimitives), Renderer/eye_ray.h:252,
 * eye_ray_host.cpp:684-753. device must be LRT_DEVICE_GPU or LRT_DEVICE_GPU_RTX;
 * the renderer runs on the current HIP device. */
int lrt_create(uint32_t device, uint32_t max_primitives, lrt_handle* out);
int lrt_destroy(lrt_handle h);

/* MultiRenderer::SetScene(SdfSBSView), eye_ray_host.cpp:360-375 ->
 * BVHRT::AddGeom_SdfSBS (BVH/BVH2Common_host.cpp:567-666) + CommitScene
 * (:1416-1487). Copies the arrays, builds the BVH2 over brick boxes and uploads
 * the repacked scene. values_f may be NULL for DX layouts. aux_data bit 0 (an
 * extension; zero in reference data) selects the voxel diagonal as the DX
 * quantisation range. */
int lrt_set_scene_sbs(lrt_handle h, const SdfSBSHeader* header, const SdfSBSNode* nodes, uint32_t n_nodes, const uint32_t* values,
                      uint64_t n_values, const float* values_f, uint64_t n_values_f);

/* The BLAS the library built (BVHRT::m_allNodePairs, BVH2Common.h:480). Pass
 * out = NULL to query the count; *depth receives the pair depth. */
int lrt_get_bvh(lrt_handle h, BVHNodePair* out, uint32_t capacity, uint32_t* count, uint32_t* depth);

/* Builds the BVH2 that lrt_set_scene_sbs would build for these bricks, on the
 * host only (BuildBVHFatCustom, BVH/cbvh_fat.cpp:168-183, with this library's
 * SAH builder in place of Embree). out = NULL queries *count and *depth. */
int lrt_build_bvh(const SdfSBSNode* nodes, uint32_t n_nodes, BVHNodePair* out, uint32_t capacity, uint32_t* count, uint32_t* depth);

/* MultiRenderer::SetPreset / GetPreset, eye_ray_host.cpp:549-561. */
int lrt_set_preset(lrt_handle h, const MultiRenderPreset* preset);
int lrt_get_preset(lrt_handle h, MultiRenderPreset* preset);
/* MultiRenderer::SetViewport, eye_ray_host.cpp:43-50. */
int lrt_set_viewport(lrt_handle h, int x0, int y0, int width, int height);
/* MultiRenderer::UpdateCamera, eye_ray_host.cpp:59-65. */
int lrt_update_camera(lrt_handle h, const float world_view[16], const float proj[16]);
/* MultiRenderer::setSeed, Renderer/eye_ray.h:173. */
int lrt_set_seed(lrt_handle h, uint32_t seed);

/* MultiRenderer::Render(imageData, w, h, "color", passNum), eye_ray_host.cpp:246-263:
 * RGBA8 packed (a<<24)|(b<<16)|(g<<8)|r. */
int lrt_render(lrt_handle h, uint32_t* image_rgba8, uint32_t width, uint32_t height, int passes);
/* MultiRenderer::RenderFloat(imageData, w, h, "color", passNum), eye_ray_host.cpp:265-282:
 * W*H*4 floats. */
int lrt_render_float(lrt_handle h, float* image_rgba, uint32_t width, uint32_t height, int passes);
/* Per-pixel CRT_Hit of sample 0 as a G-buffer (the value kernel_RayTrace
 * consumes, eye_ray.cpp:172): t (1e9 on miss), octahedral normal (2 floats,
 * coords[2..3]), primId. Normals are computed whatever the render mode. Any
 * output may be NULL. */
int lrt_render_gbuffer(lrt_handle h, float* t, float* normal_oct, uint32_t* prim, uint32_t width, uint32_t height);
/* Batched ISceneObject::RayQuery_NearestHit(posAndNear, dirAndFar),
 * BVH/BVH2Common.h:167, BVH2Common.cpp:3490-3559. rays: n * 8 floats. */
int lrt_cast_rays(lrt_handle h, const float* rays, uint32_t n, CRT_Hit* out);
/* MultiRenderer::GetExecutionTime, eye_ray_host.cpp:289-295 (milliseconds of
 * the last kernel launch under that name, measured with HIP events). */
int lrt_get_execution_time(lrt_handle h, const char* name, float out[4]);

/* Device-resident multi-view render (one launch for n_views cameras) used for
 * throughput measurement; results stay on the device until downloaded.
 * out_mode: 0 = G-buffer (t + normal), 1 = RGBA. kernel_ms may be NULL. */
int lrt_render_views_device(lrt_handle h, const float* world_views, const float* projs, uint32_t n_views, uint32_t width,
                            uint32_t height, int out_mode, float* kernel_ms);
int lrt_download_view(lrt_handle h, uint32_t view, float* t, float* normal_oct, uint32_t* prim, float* rgba);
/* Same launch with per-ray work counters: out[0] node pairs, out[1] bricks, out[2] voxels. */
int lrt_count_work(lrt_handle h, const float* world_views, const float* projs, uint32_t n_views, uint32_t width, uint32_t height,
                   uint64_t out[3]);
int lrt_sync(lrt_handle h);

/* Persistent-grid size of the render kernels (0 = device default). */
int lrt_set_launch_blocks(lrt_handle h, uint32_t blocks);

/* ---- differentiable rendering (dr::MultiRendererDR) --------------------- */
/* MultiRendererDR::SetReference(images, worldView, proj),
 * diff_render/MultiRendererDR.h:76-82, MultiRendererDR.cpp:147-185: n_views
 * RGBA images of W x H (W*H*4 floats each). masks may be NULL, then the mask is
 * 1 where |rgb| >= 0.001 (CreateRefImageMasks); the alpha used for the loss is
 * the mask. Images must already have the render size (the reference's
 * bilinear resampling, PreprocessRefImages :187-226, is not done). */
int lrt_dr_set_reference(lrt_handle h, const float* images, const float* masks, uint32_t n_views, uint32_t width, uint32_t height,
                         const float* world_views, const float* projs);
/* MultiRendererDR::RenderDR (diff_render/MultiRendererDR.cpp:578-750) for one
 * reference view with the current scene (indexed ID32F layout): image, loss
 * and dLoss/dS (values_f-sized, already scaled by getBaseGradientMult) of that
 * view. dtdv_mode 0 = analytic dt/dS (implicit function theorem), 1 = the
 * reference's central difference (BVH2DR.cpp:1075-1090). Outputs may be NULL. */
int lrt_dr_render_grad(lrt_handle h, const MultiRendererDRPreset* preset, uint32_t view, uint32_t seed, int dtdv_mode, float* out_image,
                       float* out_grad, float* loss, float* kernel_ms);
/* MultiRendererDR::OptimizeFixedStructure(preset, sbs), MultiRendererDR.cpp:278-576:
 * sets the scene, runs opt_iterations Adam steps over the reference views and
 * writes the optimised values_f back. losses (opt_iterations floats) and
 * total_ms may be NULL. View choice when image_batch_size < n_views uses a
 * PCG hash of (seed, iteration, batch slot) in place of rand(). */
int lrt_dr_optimize(lrt_handle h, const MultiRendererDRPreset* preset, const SdfSBSHeader* header, const SdfSBSNode* nodes,
                    uint32_t n_nodes, const uint32_t* values, uint64_t n_values, float* values_f, uint64_t n_values_f, uint32_t seed,
                    int dtdv_mode, float* losses, float* total_ms);
/* MultiRendererDR::getLastdLoss_dS, MultiRendererDR.h:88. */
int lrt_dr_get_last_dloss_ds(lrt_handle h, float* out, uint64_t n);
/* Border statistics of the last RenderDR: border pixels and border rays that found a relaxation point. */
int lrt_dr_last_stats(lrt_handle h, uint32_t* border_pixels, uint32_t* border_rays_hit);

/* ---- synthetic scenes (no reference counterpart; the closest reference
 * builders are utils/sparse_octree_builder.cpp:291-444 and
 * diff_render/DR_common.cpp:47-146) ---------------------------------------- */
/* Icosphere (20*4^subdiv triangles, radius r) displaced along normals by
 * seeded value noise, converted to a DX-layout SBS at brick LOD `lod`. */
int lrt_sbs_displaced_icosphere(uint32_t subdiv, float radius, float amplitude, uint32_t octaves, uint32_t seed, uint32_t lod,
                                uint32_t brick_size, uint32_t brick_pad, uint32_t bytes_per_value, int voxel_diag_scale,
                                int threads, lrt_sbs_handle* out);
/* Dense indexed grid (layout ID32F_IRGB32F[_IN]) of a torus (major R, minor r)
 * with noise displacement amp (shape = 1) or a sphere of radius R (shape = 0). */
int lrt_sbs_grid(uint32_t shape, float R, float r, float amplitude, uint32_t octaves, uint32_t seed, uint32_t brick_count,
                 uint32_t brick_size, int with_neighbors, lrt_sbs_handle* out);
int lrt_sbs_info(lrt_sbs_handle s, SdfSBSHeader* header, uint32_t* n_nodes, uint64_t* n_values, uint64_t* n_values_f);
int lrt_sbs_copy(lrt_sbs_handle s, SdfSBSNode* nodes, uint32_t* values, float* values_f);
int lrt_sbs_free(lrt_sbs_handle s);
/* load_sdf_SBS / save_sdf_SBS (sdfScene/sdf_scene.cpp:151-175): header,
 * uint32 node count, nodes, uint32 value count, values. The reference format
 * drops values_f; save appends uint32 count + floats when values_f is present
 * and load reads that tail if the file has one. */
int lrt_sbs_load_file(const char* path, lrt_sbs_handle* out);
int lrt_sbs_save_file(const char* path, const SdfSBSHeader* header, const SdfSBSNode* nodes, uint32_t n_nodes, const uint32_t* values,
                      uint32_t n_values, const float* values_f, uint32_t n_values_f);
/* LiteMath lookAt(eye, target, up) and perspectiveMatrix(fov, aspect, near, far)
 * as the reference tests build cameras (tests/tests.cpp:29-41). */
int lrt_camera_lookat(const float eye[3], const float target[3], const float up[3], float fov_deg, float aspect, float znear,
                      float zfar, float world_view[16], float proj[16]);
/* Seeded orbit camera looking at the origin: azimuth/elevation drawn from
 * `seed` and `index`, eye at `distance`, up (0,1,0), perspective fov (deg). */
int lrt_orbit_camera(uint32_t seed, uint32_t index, float distance, float fov_deg, float aspect, float znear, float zfar,
                     float world_view[16], float proj[16]);

#ifdef __cplusplus
}
#endif
#endif /* LRT_H */

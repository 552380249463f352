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
 * quantisation range. Limits: brick_size 1..16 and brick_pad 0..1 (the voxel
 * walk's index arithmetic relies on them), fewer than 2^24 bricks (the leaf
 * index) and at most 4 GiB of values (the kernels address them with 32-bit
 * offsets); other scenes return LRT_E_INVALID. */
int lrt_set_scene_sbs(lrt_handle h, const SdfSBSHeader* header, const SdfSBSNode* nodes, uint32_t n_nodes, const uint32_t* values,
                      uint64_t n_values, const float* values_f, uint64_t n_values_f);

/* The BLAS the library built (BVHRT::m_allNodePairs, BVH2Common.h:480). Pass
 * out = NULL to query the count; *depth receives the pair depth. */
int lrt_get_bvh(lrt_handle h, BVHNodePair* out, uint32_t capacity, uint32_t* count, uint32_t* depth);

/* Builds the BVH2 that lrt_set_scene_sbs would build for these bricks, on the
 * host only (BuildBVHFatCustom, BVH/cbvh_fat.cpp:168-183, with this library's
 * SAH builder in place of Embree). out = NULL queries *count and *depth. */
int lrt_build_bvh(const SdfSBSNode* nodes, uint32_t n_nodes, BVHNodePair* out, uint32_t capacity, uint32_t* count, uint32_t* depth);

/* MultiRenderer::SetLights, Renderer/eye_ray.h:158, eye_ray_host.cpp:673-676:
 * the lights of the LAMBERT / PHONG render modes and of DR LAMBERT, taken as
 * given (directions already unit, as create_direct_light leaves them); at
 * most LRT_MAX_LIGHTS. Default: a direct light along normalize(1,1,1) of
 * colour 2/3 and an ambient light of 0.25 (eye_ray_host.cpp:40). */
int lrt_set_lights(lrt_handle h, const LrtLight* lights, uint32_t n);
/* MultiRenderer::SetPreset / GetPreset, eye_ray_host.cpp:549-561. */
int lrt_set_preset(lrt_handle h, const MultiRenderPreset* preset);
int lrt_get_preset(lrt_handle h, MultiRenderPreset* preset);
/* MultiRenderer::SetViewport, eye_ray_host.cpp:43-50. */
int lrt_set_viewport(lrt_handle h, int x0, int y0, int width, int height);
/* MultiRenderer::UpdateCamera, eye_ray_host.cpp:59-65. */
int lrt_update_camera(lrt_handle h, const float world_view[16], const float proj[16]);
/* MultiRenderer::setSeed, Renderer/eye_ray.h:173. */
int lrt_set_seed(lrt_handle h, uint32_t seed);

/* Image sizes (all render calls): 1..65535 per side, as the reference's packed
 * (y<<16)|x pixel coordinates allow, and the pixel count with both sides
 * rounded up to multiples of 8 at most 2^32 - 1; larger sizes return
 * LRT_E_INVALID. */
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

/* No counterpart in the reference: lrt_render_views_device for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc). It allocates nothing: its scratch
 * (chunk counters, camera records, traversal-stack spill) is a 32 MiB
 * __device__ array per GPU, held until the process ends, and the caller
 * passes device output buffers of n_views * width * height pixels
 * ([view][y][x]): t (float), normal_oct (2 floats), prim (uint32) and, for
 * the shaded render modes, rgba (4 floats); any may be NULL, not all. At most
 * LRT_NOALLOC_MAX_VIEWS views per call; the learned seeder must be off.
 * Asynchronous on the handle's stream (lrt_sync waits). Every call on one
 * GPU uses that GPU's one scratch array: calls that overlap on a GPU race. */
#define LRT_NOALLOC_MAX_VIEWS 16
int lrt_render_views_noalloc(lrt_handle h, const float* world_views, const float* projs, uint32_t n_views, uint32_t width,
                             uint32_t height, float* d_t, float* d_normal_oct, uint32_t* d_prim, float* d_rgba);
/* Same launch with per-ray work counters: out[0] node pairs, out[1] bricks, out[2] voxels. */
int lrt_count_work(lrt_handle h, const float* world_views, const float* projs, uint32_t n_views, uint32_t width, uint32_t height,
                   uint64_t out[3]);
int lrt_sync(lrt_handle h);

/* Persistent-grid size of the render kernels (0 = device default). */
int lrt_set_launch_blocks(lrt_handle h, uint32_t blocks);
/* Tile culling of the single-scene renders, unseeded and MLP-seeded (no
 * reference counterpart: a performance switch, on by default; the batched
 * render does not cull). Each launch first bounds the rays of
 * every 8x8 pixel tile by a pyramid from the eye (k_tile_cull); a tile whose
 * pyramid reaches no box of the BVH's top levels is written as misses without
 * tracing. The images are identical either way; lrt_count_work(_ex) counts
 * the work of the form that runs (out[6]: culled tiles). */
int lrt_set_tile_culling(lrt_handle h, int on);

/* ---- batched uneven workload (no reference counterpart: the reference
 * renders one scene per call, eye_ray_host.cpp:246-282) ------------------
 * One persistent launch renders n_items (scene, camera, image) items: item i
 * traces the scene held by handle scenes[i] (same device as owner) with the
 * camera world_views[16*i], projs[16*i] into a widths[i] x heights[i]
 * G-buffer, using owner's preset (single sample). With autotune != 0 the
 * chunk size (pixels per work-counter atomic) and the grid size are chosen
 * by timing a small candidate set first, and kept in owner for later calls.
 * When owner has seeding on (lrt_set_seeding) and a network, every item is
 * traced with the learned seeder of owner (same hits as without it).
 * kernel_ms (may be NULL) receives the timed launch. */
int lrt_render_batch(lrt_handle owner, uint32_t n_items, const lrt_handle* scenes, const float* world_views, const float* projs,
                     const uint32_t* widths, const uint32_t* heights, int autotune, float* kernel_ms);
/* G-buffer of one item of the last lrt_render_batch (any output may be NULL). */
int lrt_batch_download(lrt_handle owner, uint32_t item, float* t, float* normal_oct, uint32_t* prim);
/* The work-distribution settings lrt_render_batch uses (after autotuning). */
int lrt_batch_config(lrt_handle owner, uint32_t* chunk, uint32_t* blocks);

/* ---- differentiable rendering (dr::MultiRendererDR) --------------------- */
/* MultiRendererDR::SetReference(images, worldView, proj),
 * diff_render/MultiRendererDR.h:76-82, MultiRendererDR.cpp:147-185: n_views
 * RGBA images of W x H (W*H*4 floats each). masks may be NULL, then the mask is
 * 1 where |rgb| >= 0.001 (CreateRefImageMasks); the alpha used for the loss is
 * the mask. The render size is the preset's render_width x render_height
 * when both are set, else W x H (MultiRendererDR.cpp:332-349); at another size
 * than W x H the references are resampled on the device as PreprocessRefImages
 * does (:187-226: per render pixel the mean of ceil(W / width) x
 * ceil(H / height) bilinear samples, WRAP addressing). */
int lrt_dr_set_reference(lrt_handle h, const float* images, const float* masks, uint32_t n_views, uint32_t width, uint32_t height,
                         const float* world_views, const float* projs);
/* MultiRendererDR::RenderDR (diff_render/MultiRendererDR.cpp:578-750) for one
 * reference view with the current scene (indexed ID32F layout): image, loss
 * and dLoss/dS (values_f-sized, already scaled by getBaseGradientMult) of that
 * view. dtdv_mode 0 = analytic dt/dS (implicit function theorem), 1 = the
 * reference's central difference (BVH2DR.cpp:1075-1090). Outputs may be NULL.
 * Every DR render mode: DIFFUSE, LAMBERT (the renderer's default lights,
 * eye_ray_host.cpp:40: direct along (1,1,1) of colour 2/3 plus ambient 0.25),
 * MASK, LINEAR_DEPTH, NORMAL; LAMBERT and NORMAL differentiate the hit normal
 * (smoothed over neighbouring voxels for the ID32F_IRGB32F_IN layout,
 * BVH2DR.cpp:1300-1484). Both border samplers: RANDOM and SVM
 * (CastBorderRaySVM, MultiRendererDR.cpp:1651-1931; border_spp < 1156, i.e. at
 * most 256 SVM samples per pixel, else LRT_E_UNSUPPORTED). Both diff modes:
 * DEFAULT and FINITE_DIFF (RenderDRFiniteDiff, MultiRendererDR.cpp:751-815:
 * central differences of the summed image loss, one parameter at a time, over
 * the distance values with the geometry flag, else the colour values). lrt_dr_optimize applies the
 * LK_DENOISING regularization (MultiRendererDR.cpp:1933-2057) and
 * redistancing (:2058-2382, dense brick grids whose first values are the
 * lattice) when the preset asks for them. */
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
/* MultiRendererDR::OptimizeGrid(start_grid_size, no_last_step_resize, presets),
 * MultiRendererDR.cpp:228-276: coarse-to-fine reconstruction on dense grids
 * of start_grid_size^3 bricks, starting from the reference's sphere (centre
 * (0,-0.15,0), radius 0.7, brick size 1): per preset one OptimizeFixedStructure
 * against the reference images set with lrt_dr_set_reference, then the
 * lattice resampled into bricks of twice the size. out (may be NULL) receives
 * the final grid, the scene the reference leaves in its local variable. */
int lrt_dr_optimize_grid(lrt_handle h, uint32_t start_grid_size, int no_last_step_resize, const MultiRendererDRPreset* presets,
                         uint32_t n_presets, uint32_t seed, int dtdv_mode, lrt_sbs_handle* out, float* total_ms);
/* MultiRendererDR::getLastImage(view), MultiRendererDR.h:86: the image the last
 * RenderDR of that reference view produced (W*H*4 floats: colour + coverage). */
int lrt_dr_get_last_image(lrt_handle h, uint32_t view, float* out);
/* MultiRendererDR::getLastdLoss_dS, MultiRendererDR.h:88: the gradient of the
 * last image rendered, divided by the step's image_batch_size as the reference
 * leaves it (MultiRendererDR.cpp:405, 475-476). */
int lrt_dr_get_last_dloss_ds(lrt_handle h, float* out, uint64_t n);
/* The render size of the last RenderDR / OptimizeFixedStructure (the size of
 * MultiRendererDR::getLastImage's images, m_width x m_height of
 * MultiRendererDR.cpp:332-349); the reference images' size before any call. */
int lrt_dr_render_size(lrt_handle h, uint32_t* width, uint32_t* height);
/* Border statistics of the last RenderDR: border pixels and border rays that found a relaxation point. */
int lrt_dr_last_stats(lrt_handle h, uint32_t* border_pixels, uint32_t* border_rays_hit);

/* ---- learned root-find seeder (no reference counterpart) ----------------
 * A 24 -> 64 -> 64 -> 64 -> 3 SiLU network read at every brick entry predicts
 * where the ray meets the surface inside the brick (z0 = t_hit - margin,
 * z1 = t_hit + margin, brick units from the brick entry) and a hit logit z2.
 * With seeding on, a prediction whose logit clears logit(threshold) becomes
 * the start point of the exact root finder in the candidate voxel (Newton:
 * when inside the exact bracket; sphere tracing: when the cubic is shown to
 * stay above the stopping threshold up to it). The exact intersector decides
 * every hit; the BVH order and voxel walk are unchanged. Parameters are
 * LRT_MLP_PARAMS floats, row-major out x in: W1[64x24] b1 W2[64x64] b2
 * W3[64x64] b3 W4[3x64] b4. Inference runs in BF16 on MFMA, training in FP32. */
#define LRT_MLP_PARAMS 10115u
#define LRT_MLP_FEATURES 24u
/* Uniform Glorot initialisation from a seeded generator, zero biases. */
int lrt_mlp_init(lrt_handle h, uint32_t seed);
int lrt_mlp_set_params(lrt_handle h, const float* params, uint32_t n);
int lrt_mlp_get_params(lrt_handle h, float* params, uint32_t n);
/* Appends training events from a plain render of the current scene (first
 * sample of each pixel): one event per brick the traversal enters, with its
 * 24 features, target (t_hit in brick units from the brick entry or 0, hit
 * 0/1, margin = half a voxel) and meta (ray = view*W*H + y*W + x, brick).
 * At most max_events are kept; *n_added receives the number kept and
 * *n_seen (may be NULL) every event the render produced. */
int lrt_mlp_collect(lrt_handle h, const float* world_views, const float* projs, uint32_t n_views, uint32_t width, uint32_t height,
                    uint64_t max_events, uint64_t* n_added, uint64_t* n_seen);
int lrt_mlp_dataset_size(lrt_handle h, uint64_t* n);
/* Copies events [first, first+count) to host arrays (feats 24, tgts 3,
 * meta 2 uint32 per event); any output may be NULL. */
int lrt_mlp_dataset_read(lrt_handle h, uint64_t first, uint64_t count, float* feats, float* tgts, uint32_t* meta);
int lrt_mlp_dataset_clear(lrt_handle h);
/* Adam (beta 0.9 / 0.999, eps 1e-8) over the collected events for `epochs`
 * passes in minibatches of `batch` (a multiple of 64), sample order a seeded
 * affine permutation per epoch. epoch_losses (epochs floats) and total_ms
 * may be NULL. */
int lrt_mlp_train(lrt_handle h, uint32_t epochs, uint32_t batch, float lr, uint32_t seed, float* epoch_losses, float* total_ms);
/* Mean loss and its gradient (FP32 training kernels) over n host samples at
 * the current parameters; n <= 65536. */
int lrt_mlp_loss_grad(lrt_handle h, const float* feats, const float* tgts, uint32_t n, float* grad, float* loss);
/* BF16 MFMA inference of n host feature rows -> out (n x 3). */
int lrt_mlp_infer(lrt_handle h, const float* feats, uint64_t n, float* out, float* kernel_ms);
/* BF16 MFMA inference over the whole collected dataset on the device;
 * kernel_ms is the average of `repeats` launches; the first n_out rows of the
 * result are copied to out (may be NULL). */
int lrt_mlp_infer_dataset(lrt_handle h, uint32_t repeats, float* kernel_ms, float* out, uint64_t n_out);
/* Turns seeded rendering on or off for every render entry point of h;
 * threshold is the hit-probability gate in (0, 1) (0.9 by default). */
int lrt_set_seeding(lrt_handle h, int enabled, float threshold);
/* lrt_count_work with more counters: out[0] node pairs, [1] bricks, [2]
 * voxels, [3] root-finder iterations, [4] 16-sample MLP tiles, [5] brick
 * entries whose prediction cleared the gate, [6] 8x8 tiles written as misses
 * by the tile cull. n_out <= 7. */
int lrt_count_work_ex(lrt_handle h, const float* world_views, const float* projs, uint32_t n_views, uint32_t width, uint32_t height,
                      uint64_t* out, uint32_t n_out);

/* ---- synthetic scenes (no reference counterpart; the closest reference
 * builders are utils/sparse_octree_builder.cpp:291-444 and
 * diff_render/DR_common.cpp:47-146) ---------------------------------------- */
/* Icosphere (20*4^subdiv triangles, radius r) displaced along normals by
 * seeded value noise, converted to a DX-layout SBS at brick LOD `lod`. */
int lrt_sbs_displaced_icosphere(uint32_t subdiv, float radius, float amplitude, uint32_t octaves, uint32_t seed, uint32_t lod,
                                uint32_t brick_size, uint32_t brick_pad, uint32_t bytes_per_value, int voxel_diag_scale,
                                int threads, lrt_sbs_handle* out);
/* Torus (major R, minor r, in the xy plane) displaced by seeded value noise,
 * as a sparse DX-layout SBS at brick LOD `lod` (same builder as the icosphere). */
int lrt_sbs_displaced_torus(float R, float r, float amplitude, uint32_t octaves, uint32_t seed, uint32_t lod, uint32_t brick_size,
                            uint32_t brick_pad, uint32_t bytes_per_value, int voxel_diag_scale, int threads, lrt_sbs_handle* out);
/* Dense indexed grid (layout ID32F_IRGB32F[_IN]) of a torus (major R, minor r)
 * with noise displacement amp (shape = 1) or a sphere of radius R (shape = 0). */
int lrt_sbs_grid(uint32_t shape, float R, float r, float amplitude, uint32_t octaves, uint32_t seed, uint32_t brick_count,
                 uint32_t brick_size, int with_neighbors, lrt_sbs_handle* out);
/* SBS of a closed triangle mesh (vertices: 3 floats each; indices: 3 per
 * triangle), as sdf_converter::create_sdf_SBS(settings, header, mesh) builds
 * one (utils/sdf_converter.cpp:183-228): exact signed distance to the mesh,
 * bricks of brick_size voxels at brick LOD lod (a power of two) kept by the
 * reference's border rule and quantiser (utils/sparse_octree_builder.cpp:
 * 49-53, 405-428), DX layout; the candidate search is this library's
 * coarse-to-fine cell test, and values are quantised in double (the reference
 * computes the same expression in float). lrt_sbs_from_mesh builds on the host (OpenMP,
 * threads <= 0: all cores); lrt_sbs_from_mesh_device on the GPU of h, with the
 * same result bit for bit. build_ms (may be NULL): wall time from mesh to scene. */
int lrt_sbs_from_mesh(const float* vertices, uint32_t n_vertices, const uint32_t* indices, uint32_t n_triangles, uint32_t lod,
                      uint32_t brick_size, uint32_t brick_pad, uint32_t bytes_per_value, int voxel_diag_scale, int threads,
                      lrt_sbs_handle* out, float* build_ms);
int lrt_sbs_from_mesh_device(lrt_handle h, const float* vertices, uint32_t n_vertices, const uint32_t* indices, uint32_t n_triangles,
                             uint32_t lod, uint32_t brick_size, uint32_t brick_pad, uint32_t bytes_per_value, int voxel_diag_scale,
                             lrt_sbs_handle* out, float* build_ms);
/* sdf_converter::create_sdf_SBS(settings, header, sdf) (utils/sdf_converter.cpp:
 * 169-181 for a distance function, :183-193 for a mesh) on the GPU of h: the
 * reference's adaptive frame octree (construct_sdf_frame_octree,
 * utils/sparse_octree_builder.cpp:125-258: every cell of level min(depth, 4),
 * then cells split while their centre lies within their half-diagonal of the
 * surface or their corners change sign, down to `depth`) and
 * frame_octree_to_SBS (:291-444): every leaf whose lattice reaches the
 * surface band becomes a brick at its own LOD (mixed LODs), values quantised
 * over the brick diagonal with the reference's float expression (aux_data
 * LRT_SBS_LAYOUT_DX, no voxel-diagonal flag). Bricks are ordered by LOD, then
 * Morton order (the reference's order depends on its threads). nodes_limit is
 * SparseOctreeSettings::nodes_limit (0: its default 2^24): when the frame
 * holds more nodes, frame_octree_limit_nodes (:2229-2330) merges the children
 * of the parents that approximate them best until at most nodes_limit border
 * leaves remain (equal merge keys keep depth-first order; the reference's
 * std::sort leaves them unordered). Frames above 2^26 nodes are refused.
 * The mesh form takes the exact distance to the mesh with the sign of the
 * closest feature's pseudo-normal (the reference signs with the closest
 * triangle's face normal that Embree's point query returns). The primitive
 * form takes LRT_PRIMITIVE_SPHERE (params: centre xyz, radius) or
 * LRT_PRIMITIVE_TORUS (around the z axis; params: R, r), evaluated in float
 * as length(p - c) - radius and length((length(p.xy) - R, p.z)) - r. */
int lrt_sbs_octree_from_mesh_device(lrt_handle h, const float* vertices, uint32_t n_vertices, const uint32_t* indices, uint32_t n_triangles,
                                    uint32_t depth, uint32_t brick_size, uint32_t brick_pad, uint32_t bytes_per_value, uint32_t nodes_limit,
                                    lrt_sbs_handle* out, float* build_ms);
int lrt_sbs_octree_from_primitive_device(lrt_handle h, uint32_t shape, const float* params, uint32_t depth, uint32_t brick_size,
                                         uint32_t brick_pad, uint32_t bytes_per_value, uint32_t nodes_limit, lrt_sbs_handle* out,
                                         float* build_ms);
/* The displaced icosphere mesh that lrt_sbs_displaced_icosphere builds from
 * (no reference counterpart). NULL vertices / indices: only the counts. */
int lrt_mesh_displaced_icosphere(uint32_t subdiv, float radius, float amplitude, uint32_t octaves, uint32_t seed, float* vertices,
                                 uint32_t* n_vertices, uint32_t* indices, uint32_t* n_triangles);
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
/* MultiRenderer::LoadScene / LoadSceneHydra (Renderer/eye_ray_host.cpp:52-121):
 * reads a Hydra XML scene -- the first camera of cam_lib (fov, near and far
 * clip planes, position, look_at, up; aspect from the viewport) and the
 * geometry instances of the first scene -- and sets the scene and camera.
 * This path renders one SBS: the file must hold exactly one instance, with
 * the identity matrix, of an sdf_sbs geometry node, whose loc (relative to
 * the scene's folder) is an SBS file (load_sdf_SBS). Other node types and
 * transformed or several instances return LRT_E_UNSUPPORTED. */
int lrt_load_scene_hydra(lrt_handle h, const char* path);
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

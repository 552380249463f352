/*
 * lrt_oracle.h - C API of the CPU oracle.
 *
 * TEST INFRASTRUCTURE ONLY: a CPU restatement of the reference LiteRT SBS
 * path (BVH2 traversal, brick march, trilinear root finds, normals, DR
 * forward/adjoint/border/Adam) used as the checker for the HIP library and as
 * the CPU baseline in bench.py. Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it. Nothing in the product library
 * links or calls it.
 *
 * precision: 32 = FP32 twin of the reference, 64 = FP64 error reference.
 *
 * Parity unpinned (only an import of the reference's third-party code could
 * pin these; they restate the reference as written):
 *  - LiteMath's perspectiveMatrix / lookAt / inverse4x4 conventions: the
 *    camera helpers use the GL convention (pinned against the closed-form
 *    pinhole, horizontal and vertical field of view, in
 *    tests/test_cpu_oracle_pins.py), but whether LiteMath's differ is not;
 *  - Embree's BVH topology: the oracle traverses the BVH it is given, so which
 *    of two bricks hit along one ray "stop at first hit" returns follows that
 *    tree (bounded by the brute-force nearest-hit tests);
 *  - LAMBERT's light direction to the last bit (1/sqrt(3) rounded to float;
 *    LiteMath's normalize of (1,1,1) may differ by an ulp).
 * Pinned by checks that do not restate the formulas: the ANALYTIC cubic with
 * the reference's R = (2a^2 - 9ab + 27c) / 54 against FP64 bracketing where it
 * equals Vieta's (a in {0, 1}); the smoothed-normal blend of
 * ID32F_IRGB32F_IN scenes (plane normal on linear fields, continuity across
 * voxel and brick faces); the DR border term against finite differences (the
 * reference's test_position_derivatives, tests/test_cpu_border_fd.py).
 * The DR normal-mode derivatives (NORMAL, LAMBERT, both layouts) are pinned
 * by central differences of the FP64 loss, and the diffuse position term by
 * the same differences times the reference's sz*bs/2 scale.
 */
#ifndef LRT_ORACLE_H
#define LRT_ORACLE_H
#include <stdint.h>
#include "../include/lrt_types.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_scene orc_scene;

typedef struct orc_counters {
  uint64_t node_pairs, bricks, voxels, root_finds, rays, root_iters;
} orc_counters;

/* Copies the SBS arrays and the fat BVH2 pairs (root pair at index 0). */
orc_scene* orc_scene_create(const SdfSBSHeader* h, const SdfSBSNode* nodes, uint32_t n_nodes, const uint32_t* values,
                            uint64_t n_values, const float* values_f, uint64_t n_values_f, const BVHNodePair* pairs,
                            uint32_t n_pairs);
void orc_scene_destroy(orc_scene* s);

/* One view, like MultiRenderer::RenderFloat (eye_ray_host.cpp:265-282, 575-585).
 * Any output pointer may be NULL. out_rgba: W*H*4; out_t, out_n (W*H*2), out_prim:
 * the CRT_Hit of sample 0 per pixel (t = 1e9 and prim = ~0 on a miss). */
int orc_render(const orc_scene* s, const float worldView[16], const float proj[16], const MultiRenderPreset* preset,
               uint32_t W, uint32_t H, uint32_t seed, int precision, int nthreads, int force_normal, double* out_rgba,
               double* out_t, uint32_t* out_prim, double* out_n, orc_counters* cnt);

/* Batched RayQuery_NearestHit (BVH2Common.cpp:3490-3559). rays: n*8 floats
 * (posAndNear, dirAndFar). out: n*6 doubles (t, primId, geomId, coords[2], coords[3], instId). */
int orc_cast_rays(const orc_scene* s, const MultiRenderPreset* preset, const float* rays, uint32_t n, int precision,
                  int nthreads, int force_normal, double* out);

/* Nearest hit over every brick, ignoring the BVH (topology-independent check). */
int orc_brute_nearest(const orc_scene* s, const MultiRenderPreset* preset, const float* rays, uint32_t n, int precision,
                      int nthreads, double* out_t, uint32_t* out_prim);

/* ---- unit-level entry points (for known-answer tests) ------------------ */
double orc_eval_trilinear(const double v[8], const double dp[3], int precision);
void orc_eval_trilinear_diff(const double v[8], const double dp[3], int precision, double out[3]);
/* LocalSurfaceIntersection on one voxel; returns 1 on hit with *t_out = t (voxel units). */
int orc_voxel_root(uint32_t mode, const double values[8], const double start_q[3], const double dir[3], double d,
                   double qFar, int precision, double* t_out, uint32_t* iters_out);
void orc_encode_normal(const double n[3], int precision, double out[2]);
void orc_decode_normal(const double e[2], int precision, double out[3]);
void orc_pcg3d(uint32_t v[3]);
void orc_rand2(uint32_t x, uint32_t y, uint32_t iter, uint32_t seed, float out[2]);
void orc_pack_xy(uint32_t W, uint32_t H, uint32_t* out); /* 8*ceil(W/8) * 8*ceil(H/8) entries */
void orc_safe_inverse(const double d[3], int precision, double out[3]);
void orc_ray_box(const double o[3], const double inv[3], const double bmin[3], const double bmax[3], int precision,
                 double out[2]);
/* Decodes the 8 corner values of voxel (vx,vy,vz) of brick `node` (load_distance_values). */
double orc_load_voxel(const orc_scene* s, uint32_t node, uint32_t vx, uint32_t vy, uint32_t vz, int precision, double out[8]);
void orc_eye_ray(const float worldView[16], const float proj[16], uint32_t x, uint32_t y, float dx, float dy, uint32_t W,
                 uint32_t H, int precision, double out[8]);

/* ---- differentiable rendering (diff_render/) ---------------------------- */
typedef struct orc_dr_stats {
  double loss;
  uint32_t border_pixels;
  uint32_t border_rays_hit;
  uint64_t rays;
} orc_dr_stats;

/* RenderDR (MultiRendererDR.cpp:578-750) for one view. ref: W*H*4 reference
 * image; out_image W*H*4; out_grad: params_count floats, ACCUMULATED into
 * (scaled by getBaseGradientMult). dtdv_mode 0 = analytic (implicit function
 * theorem), 1 = central finite difference as the reference does. */
int orc_dr_render_grad(const orc_scene* s, const MultiRendererDRPreset* dr, const float worldView[16], const float proj[16],
                       uint32_t W, uint32_t H, uint32_t seed, const float* ref, int precision, int nthreads, int dtdv_mode,
                       double* out_image, double* out_grad, orc_dr_stats* stats);

/* Regularization (MultiRendererDR.cpp:1933-2057) of bricks [0, n_nodes), added to grad. */
int orc_dr_regularize(const orc_scene* s, uint32_t n_nodes, float lambda, float power, double* grad);

/* Redistance (MultiRendererDR.cpp:2122-2382) of a dense p^3 lattice (x-major) in place. */
int orc_dr_redistance(float* values, uint32_t p, float h, uint32_t iters);

/* OptimizeStepAdam (MultiRendererDR.cpp:828-853): updates X, V, S in place. */
void orc_adam_step(uint32_t iter, const float* dX, float* X, float* V, float* S, uint32_t n, float lr, float beta1,
                   float beta2, float eps);

/* Replaces values_f of the scene (indexed layouts). */
int orc_scene_set_values_f(orc_scene* s, const float* values_f, uint64_t n);
/* MultiRenderer::SetLights for the scene's shading and DR LAMBERT. */
int orc_scene_set_lights(orc_scene* s, const LrtLight* lights, uint32_t n);

/* ---- MLP (learned seeding, no reference counterpart) -------------------- */
/* FP64 forward of the 24-64-64-64-3 SiLU MLP: weights laid out as
 * W1[64x24] b1[64] W2[64x64] b2[64] W3[64x64] b3[64] W4[3x64] b4[3] (row-major,
 * out x in). x: n*24, y: n*3. bf16_inputs rounds weights and activations to
 * BF16 at the same points as the GPU kernel. */
void orc_mlp_forward(const float* params, const float* x, uint32_t n, int bf16_emulate, double* y);
/* Loss and gradient of the training objective (see oracle_mlp.cpp) w.r.t. params. */
double orc_mlp_loss_grad(const float* params, const float* x, const float* target, uint32_t n, int bf16_emulate,
                         double* grad);

/* Seeder feature table of one brick (16 floats) and the 24 features of a
 * ray (pos xyz, dir xyz) entering it; *t0 = brick entry distance. */
int orc_mlp_brick_features(const orc_scene* s, uint32_t brick, float out[16]);
int orc_mlp_features(const orc_scene* s, const float ray[6], uint32_t brick, float out[24], float* t0);
/* create_sdf_SBS(settings, header, sdf) (utils/sdf_converter.cpp:169-181) for
 * a distance field -- shape 0: sphere, params = centre xyz, radius; shape 1:
 * torus around the z axis, params = R, r -- at octree depth `depth`: the
 * frame octree of construct_sdf_frame_octree, frame_octree_limit_nodes with
 * nodes_limit (0: the default 2^24) and the bricks of frame_octree_to_SBS, in
 * the frame octree's order. NULL outputs: counts only. Returns 0 or -1 (bad
 * arguments). */
int orc_sbs_octree(int shape, const float* params, uint32_t depth, uint32_t brick_size, uint32_t brick_pad, uint32_t bytes_per_value,
                   uint32_t nodes_limit, uint32_t* n_nodes, uint32_t* nodes_out, uint64_t* n_values, uint32_t* values_out);

#ifdef __cplusplus
}
#endif
#endif

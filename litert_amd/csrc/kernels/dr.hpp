// dr.hpp - launch interface of the differentiable-rendering kernels.
#pragma once
#include <hip/hip_runtime.h>
#include "render.hpp"

namespace lrtd {

struct DRParams;

// One image of an Adam step's batch for the multi-view primary pass: its
// camera, reference and image slot; depth_thr and grad are null for every
// image but the one whose gradient the step keeps.
struct DRView {
  Camera cam;
  const float4* ref;
  float4* out_image;
  float* depth_thr;
  float* grad;
};

struct DRArgs {
  DevScene S;
  Camera cam;
  float proj_view[16];  // proj * worldView (TransformWorldToScreenSpace, MultiRendererDR.cpp:1296-1301)
  struct {
    uint32_t ray_flags;
    uint32_t dtdv_fd;
  } Q;
  uint32_t W, H, packed_w, spp, seed;
  uint32_t mode, input_type, loss_fn, border_spp, forced_border;
  float relax_eps;
  uint32_t border_sampling;  // LRT_DR_BORDER_SAMPLING_RANDOM / _SVM
  uint32_t svm_side;         // SVM samples per axis: uint(sqrt(max(1, uint(0.25 * border_spp)))), computed on the host
  float up[3];               // row 1 of worldView (CastBorderRaySVM, MultiRendererDR.cpp:1784)
  const float4* ref;   // reference image, alpha = mask
  float4* out_image;   // rendered colour + coverage
  float* depth_thr;    // per-pixel depth-discontinuity threshold
  float* grad;         // dLoss/dS, one float per values_f entry
  const float* bvals;     // per-brick corner values gathered from values_f (k_dr_gather)
  const float2* bminmax;  // smallest / largest of each brick's gathered values
  LightSet lights;        // LAMBERT: direct and ambient lights
  const DRView* views;    // null: the primary pass renders the view above; else these n_views records
  uint32_t n_views;
};

}  // namespace lrtd

namespace lrt {
// true if launch_dr_primary runs the linear-depth form of the primary pass
bool dr_primary_is_depth(const lrtd::DRArgs& A);
// true if launch_dr_border_rays runs the linear-depth form of the border pass
bool dr_border_is_depth(const lrtd::DRArgs& A);
// Device tables of launch_dr_cull. Slots are node-pair children, 2 * pair + side.
struct DRCullTables {
  const uint32_t* leaf_slot;  // per brick: the slot whose leaf references it, 0xFFFFFFFF if none
  const uint32_t* pair_slot;  // per node pair: the slot that references it, 0xFFFFFFFF for the root
  const uint8_t* tiled;       // per slot: 1 if the bricks of its subtree tile its box exactly
  uint32_t* occ;              // scratch, 2 x 2 * n_pairs words
  uint4* pairs_primary;       // n_pairs x 64 B: copy for the linear-depth primary pass
  uint4* pairs_border;        // n_pairs x 64 B: copy for the linear-depth border pass
};
// Writes both BVH copies for the current corner values (d_minmax from
// launch_dr_gather) and relax_eps.
// Finite-difference DR (RenderDRFiniteDiff, MultiRendererDR.cpp:751-815):
// set one parameter (mode 0: p0 + delta, 1: p0 - delta, 2: p0) in values_f
// and at its positions in the per-brick corner copy, clearing the primary's
// work counter; sum the image's per-pixel loss in double (fixed order); turn
// the loss pairs into dLoss/dS for the listed parameters.
hipError_t launch_dr_fd_set(float* d_values_f, float* d_bvals, const float* d_p0, const uint32_t* d_pos, uint32_t n_pos, uint32_t j,
                            uint32_t mode, float delta, uint32_t* d_counter, hipStream_t s);
hipError_t launch_dr_fd_loss(const float4* d_img, const float4* d_ref, uint32_t n, uint32_t loss_fn, double* d_out, hipStream_t s);
hipError_t launch_dr_fd_grad(const double* d_acc, const uint32_t* d_ids, uint32_t n, float g_mult, float delta, float* d_grad, hipStream_t s);
hipError_t launch_dr_cull(const DRCullTables& T, const float2* d_minmax, float relax_eps, const uint4* d_pairs, uint32_t n_bricks,
                          uint32_t n_pairs, hipStream_t s);
// cull_pairs: T.pairs_primary / T.pairs_border of launch_dr_cull for the
// current values, or null for the full tree
hipError_t launch_dr_primary(const lrtd::DRArgs& A, uint32_t* d_counter, uint32_t* d_spill, float* d_loss, uint32_t blocks, hipStream_t s,
                             const uint4* cull_pairs = nullptr);
hipError_t launch_dr_border_find(const lrtd::DRArgs& A, uint32_t* d_list, uint32_t* d_count, hipStream_t s);
hipError_t launch_dr_border_rays(const lrtd::DRArgs& A, const uint32_t* d_list, const uint32_t* d_count, uint32_t* d_spill,
                                 uint32_t* d_hits, uint32_t blocks, hipStream_t s, const uint4* cull_pairs = nullptr);
hipError_t launch_dr_gather(const lrtd::DevScene& S, float* d_out, float2* d_minmax, hipStream_t s);
hipError_t launch_dr_scale(float* d_grad, uint32_t n, float mult, hipStream_t s);
// PreprocessRefImages (MultiRendererDR.cpp:187-226): n_images W0 x H0 images
// resampled to W x H (supersampled bilinear, WRAP addressing)
hipError_t launch_dr_resample(const float4* d_src, uint32_t W0, uint32_t H0, float4* d_dst, uint32_t W, uint32_t H, uint32_t n_images,
                              hipStream_t s);
// Redistance (MultiRendererDR.cpp:2122-2382) of the first p^3 values (a dense
// x-major lattice), spacing h, iters Jacobi steps; scratch: p^3 bytes + 2 x p^3 floats
hipError_t launch_dr_redistance(float* d_values, uint32_t p, float h, uint32_t iters, uint8_t* d_frozen, float* d_b0, float* d_b1,
                                hipStream_t s);
// Regularization (MultiRendererDR.cpp:1933-2057) of bricks [0, n_nodes) added to d_grad
hipError_t launch_dr_regularize(const lrtd::DevScene& S, uint32_t n_nodes, float lambda, float power, float* d_grad, hipStream_t s);
hipError_t launch_adam(const float* g, float* X, float* V, float* S, uint32_t n, float lr, float b1, float b2, float eps, float c1,
                       float c2, float batch, hipStream_t s);
}  // namespace lrt

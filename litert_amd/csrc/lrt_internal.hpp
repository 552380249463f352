// lrt_internal.hpp - state behind an lrt_handle.
#pragma once
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstring>
#include <map>
#include <string>
#include <vector>
#include "../../include/lrt.h"
#include "kernels/mlp.hpp"
#include "kernels/render.hpp"
#include "host/scene_gen.hpp"

namespace lrt {

int set_err(int code, const std::string& msg);
int hip_err(hipError_t e, const char* what);
lrtd::Camera make_camera(const float wv[16], const float pr[16]);
std::vector<float> brick_records(const std::vector<SdfSBSNode>& nodes, const SdfSBSHeader& h, uint32_t quant_voxel_diag);

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  hipError_t ensure(size_t n);
  hipError_t upload(const void* src, size_t n, hipStream_t s);
  void release();
};

struct Renderer {
  int device = 0;
  int cu_count = 256;
  uint32_t launch_blocks = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;

  MultiRenderPreset preset{};
  uint32_t width = 0, height = 0, seed = 0;
  // MultiRenderer::m_lights; the default set of eye_ray_host.cpp:40: a direct
  // light along normalize(1,1,1) of colour 2/3 and an ambient light of 0.25
  std::vector<LrtLight> lights = default_lights();
  static std::vector<LrtLight> default_lights() {
    const float d = 1.0f / std::sqrt(3.0f), c = 2.0f / 3.0f;
    return {LrtLight{{d, d, d}, LRT_LIGHT_TYPE_DIRECT, {c, c, c}, 0}, LrtLight{{0, 0, 0}, LRT_LIGHT_TYPE_AMBIENT, {0.25f, 0.25f, 0.25f}, 0}};
  }
  lrtd::LightSet light_set() const {
    lrtd::LightSet L{};
    L.n = (uint32_t)lights.size();
    for (uint32_t i = 0; i < L.n; i++) {
      uint32_t tb = lights[i].type;
      float tf;
      std::memcpy(&tf, &tb, 4);
      L.sp[i] = make_float4(lights[i].space[0], lights[i].space[1], lights[i].space[2], tf);
      L.col[i] = make_float4(lights[i].color[0], lights[i].color[1], lights[i].color[2], 0.f);
    }
    return L;
  }
  float world_view[16] = {}, proj[16] = {};
  lrtd::Camera cam{};
  bool has_camera = false;

  // scene as ingested (host copy, reference layout)
  bool has_scene = false;
  SdfSBSHeader header{};
  uint32_t quant_voxel_diag = 0, geom_type = LRT_TYPE_SDF_SBS;
  std::vector<SdfSBSNode> nodes;
  std::vector<uint32_t> values;
  std::vector<float> values_f;
  std::vector<BVHNodePair> pairs;
  int pair_depth = 0;

  // device scene
  DevBuf d_pairs, d_nodes, d_values, d_values_f, d_brecs;
  // work buffers
  DevBuf d_cams, d_counter, d_spill, d_counts, d_rays, d_hits;
  DevBuf d_out_t, d_out_n, d_out_prim, d_out_rgba;
  uint32_t last_w = 0, last_h = 0, last_views = 0;
  // tile culling of the unseeded renders (on by default): the BVH
  // boxes every tile's pyramid is tested against (cull_boxes) and the
  // launch's culled-tile bits (k_tile_cull)
  bool tile_cull = true;
  uint32_t n_cull = 0;  // coarse cull boxes (the fine ones follow them in d_cull_boxes)
  DevBuf d_cull_boxes, d_tile_mask;

  // differentiable rendering state (MultiRendererDR)
  uint32_t dr_n_ref = 0, dr_w = 0, dr_h = 0;  // reference views; render size (d_dr_ref holds the references at it)
  uint32_t dr_ref_w = 0, dr_ref_h = 0;       // size of the images passed to lrt_dr_set_reference (d_dr_ref_orig)
  std::vector<uint8_t> dr_img_valid;  // per reference view: its slot of d_dr_img holds the last RenderDR image
  std::vector<float> dr_world_views, dr_projs;  // 16 floats per reference view
  DevBuf d_dr_ref_orig, d_dr_ref, d_dr_img, d_dr_thr, d_dr_grad, d_dr_list, d_dr_misc, d_adam_v, d_adam_s;
  DevBuf d_dr_bvals;    // per-brick corner values gathered from values_f before every RenderDR
  DevBuf d_dr_bminmax;  // per-brick smallest / largest gathered value
  // BVH copies for the linear-depth passes (launch_dr_cull): slot tables
  // built once per scene; dr_tree_state 0 = not built, 1 = built, -1 = the
  // tree does not have the form the tables assume (the full tree is used)
  DevBuf d_dr_leaf_slot, d_dr_pair_slot, d_dr_tiled, d_dr_occ, d_dr_cull_pairs;
  DevBuf d_dr_redist;  // redistancing scratch: frozen flags + two lattice buffers
  DevBuf d_dr_losses;  // per-step loss sums of lrt_dr_optimize
  DevBuf d_dr_views;   // lrt_dr_optimize with image_batch_size > 1: one lrtd::DRView per image of every step
  // finite-difference DR: saved parameters, the referenced parameters, their
  // positions in the corner copy (CSR), and the loss pair of each
  DevBuf d_dr_fd_p0, d_dr_fd_ids, d_dr_fd_pos, d_dr_fd_acc;
  int dr_tree_state = 0;
  float dr_last_loss = 0.f;
  uint32_t dr_last_batch = 1;  // images per step of the last gradient (lrt_dr_get_last_dloss_ds divides by it)
  uint32_t dr_last_border = 0, dr_last_border_hits = 0;
  // batched render (lrt_render_batch)
  DevBuf d_b_items, d_b_offsets, d_b_t, d_b_n, d_b_prim;
  std::vector<uint64_t> b_pix_off;            // first pixel of each item in the batch outputs
  std::vector<uint32_t> b_w, b_h;
  uint32_t b_chunk = 256, b_blocks = 0;        // 0 = device default grid

  // learned root-find seeder
  bool mlp_ready = false;            // parameters set and packed
  std::vector<float> mlp_params;     // FP32 master copy
  DevBuf d_mlp_params, d_mlp_m, d_mlp_v, d_mlp_packed, d_mlp_partial, d_mlp_ploss, d_mlp_grad, d_mlp_loss;
  uint64_t mlp_adam_t = 0;
  DevBuf d_bfeat;                    // per-brick features of the current scene
  bool bfeat_valid = false;
  DevBuf d_ds_feat, d_ds_tgt, d_ds_meta, d_ds_out;
  uint64_t ds_n = 0, ds_cap = 0;
  bool seeding = false;
  float seed_threshold = 0.9f;

  float last_ms = 0.f;
  std::string pending_timer;
  std::map<std::string, float> times;

  lrtd::DevScene dev_scene() const;
  lrtd::TraceParams trace_params(bool force_normal) const;
  uint32_t grid_blocks() const;
  hipError_t ensure_spill(uint32_t blocks, uint32_t lds_entries = lrtd::kStackLds);
  hipError_t ensure_brick_features();
  int pack_mlp();
  int render_device(const lrtd::Camera* cams, uint32_t n_views, uint32_t W, uint32_t H, bool rgba, bool gbuf, bool force_normal,
                    const char* timer_name, unsigned long long* d_counts);
  int finish_timer();
  void release_all() {
    DevBuf* all[] = {&d_pairs, &d_nodes, &d_values, &d_values_f, &d_brecs, &d_cams, &d_counter, &d_spill, &d_counts,
                     &d_rays, &d_hits, &d_out_t, &d_out_n, &d_out_prim, &d_out_rgba, &d_dr_ref_orig, &d_dr_ref, &d_dr_img,
                     &d_dr_thr, &d_dr_grad, &d_dr_list, &d_dr_misc, &d_adam_v, &d_adam_s, &d_dr_bvals, &d_dr_bminmax, &d_dr_leaf_slot,
                     &d_dr_pair_slot, &d_dr_tiled, &d_dr_occ, &d_dr_redist, &d_dr_losses, &d_dr_views, &d_dr_cull_pairs, &d_dr_fd_p0, &d_dr_fd_ids, &d_dr_fd_pos, &d_dr_fd_acc, &d_mlp_params,
                     &d_mlp_m, &d_mlp_v, &d_mlp_packed, &d_mlp_partial, &d_mlp_ploss, &d_mlp_grad, &d_mlp_loss,
                     &d_bfeat, &d_ds_feat, &d_ds_tgt, &d_ds_meta, &d_ds_out, &d_b_items, &d_b_offsets, &d_b_t,
                     &d_b_n, &d_b_prim, &d_cull_boxes, &d_tile_mask};
    for (DevBuf* b : all) b->release();
  }
};

}  // namespace lrt

struct lrt_renderer : lrt::Renderer {};

// A scene produced by one of the builders (lrt_sbs_* entry points).
struct lrt_sbs {
  lrt::SbsScene s;
};

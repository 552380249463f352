// lrt_dr_capi.hip - C ABI of the differentiable renderer (dr::MultiRendererDR).
#include <hip/hip_runtime.h>
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

#include "../../include/lrt.h"
#include "kernels/dr.hpp"
#include "lrt_internal.hpp"

namespace lrt {

static uint32_t pcg_hash(uint32_t v) {
  uint32_t state = v * 747796405u + 2891336453u;
  uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
  return (word >> 22u) ^ word;
}

// CastRayWithGrad ray flags, MultiRendererDR.cpp:1119-1142
static uint32_t dr_ray_flags(const MultiRendererDRPreset& p) {
  if (p.dr_diff_mode == LRT_DR_DIFF_MODE_FINITE_DIFF) return LRT_DR_RAY_FLAG_NO_DIFF;
  if (p.dr_input_type == LRT_DR_INPUT_TYPE_LINEAR_DEPTH) return LRT_DR_RAY_FLAG_DDIST_DPOS;
  uint32_t f = LRT_DR_RAY_FLAG_NO_DIFF;
  if (p.dr_reconstruction_flags & LRT_DR_RECONSTRUCTION_FLAG_COLOR) f |= LRT_DR_RAY_FLAG_DDIFFUSE_DCOLOR;
  if (p.dr_reconstruction_flags & LRT_DR_RECONSTRUCTION_FLAG_GEOMETRY) {
    if (p.dr_render_mode == LRT_DR_RENDER_MODE_DIFFUSE) f |= LRT_DR_RAY_FLAG_DDIFFUSE_DPOS;
    else if (p.dr_render_mode == LRT_DR_RENDER_MODE_LAMBERT) f |= LRT_DR_RAY_FLAG_DDIFFUSE_DPOS | LRT_DR_RAY_FLAG_DNORM_DPOS;
    else if (p.dr_render_mode == LRT_DR_RENDER_MODE_NORMAL) f |= LRT_DR_RAY_FLAG_DNORM_DPOS;
  }
  return f;
}

static int dr_check(Renderer* r, const MultiRendererDRPreset& p) {
  if (!r->has_scene) return set_err(LRT_E_NOSCENE, "no scene");
  const uint32_t lay = r->header.aux_data & LRT_SBS_LAYOUT_MASK;
  if (lay != LRT_SBS_LAYOUT_ID32F_IRGB32F && lay != LRT_SBS_LAYOUT_ID32F_IRGB32F_IN)
    return set_err(LRT_E_UNSUPPORTED, "differentiable rendering needs an indexed ID32F_IRGB32F(_IN) layout (MultiRendererDR.cpp:283-284)");
  if (r->header.brick_pad != 0) return set_err(LRT_E_UNSUPPORTED, "indexed DR layouts have brick_pad 0");
  if (p.dr_render_mode > LRT_DR_RENDER_MODE_NORMAL) return set_err(LRT_E_INVALID, "bad DR render mode");
  if (p.dr_loss_function > LRT_DR_LOSS_FUNCTION_MAE) return set_err(LRT_E_INVALID, "bad loss function");
  if (p.dr_border_sampling > LRT_DR_BORDER_SAMPLING_SVM) return set_err(LRT_E_INVALID, "bad border sampling");
  if (p.dr_border_sampling == LRT_DR_BORDER_SAMPLING_SVM && p.border_spp >= 1156)
    return set_err(LRT_E_UNSUPPORTED, "SVM border sampling takes border_spp < 1156 (at most 256 SVM samples per pixel)");
  if (p.dr_diff_mode > LRT_DR_DIFF_MODE_FINITE_DIFF) return set_err(LRT_E_INVALID, "bad diff mode");
  if (p.reg_function > LRT_DR_REG_FUNCTION_LK_DENOISING) return set_err(LRT_E_INVALID, "bad regularization function");
  if (p.redistancing_enable && p.redistancing_interval == 0) return set_err(LRT_E_INVALID, "redistancing_interval must be > 0");
  if (p.dr_render_mode == LRT_DR_RENDER_MODE_LAMBERT)
    for (const LrtLight& l : r->lights)
      if (l.type == LRT_LIGHT_TYPE_POINT)
        return set_err(LRT_E_UNSUPPORTED, "DR LAMBERT supports direct and ambient lights (MultiRendererDR.cpp:925-928 asserts)");
  if (p.spp < 1 || p.spp > 4096) return set_err(LRT_E_INVALID, "DR spp must be 1..4096");
  if (p.border_relax_eps <= 0.f) return set_err(LRT_E_INVALID, "border_relax_eps must be > 0");
  if (r->dr_n_ref == 0) return set_err(LRT_E_INVALID, "no reference images: call lrt_dr_set_reference");
  if (p.render_width > 65535 || p.render_height > 65535) return set_err(LRT_E_INVALID, "render size above 65535");
  return LRT_OK;
}

// The render size of a preset and the references at that size
// (MultiRendererDR.cpp:332-349): the size of the reference images when
// render_width or render_height is 0, else the preset's; at another size than
// the originals the references are resampled (PreprocessRefImages, :187-226,
// launch_dr_resample). At the originals' size they are used as given: the
// reference's bilinear sample at a pixel centre returns that pixel up to the
// rounding of u * size - 0.5. Changing the size invalidates the last images.
static int dr_prepare_size(Renderer* r, const MultiRendererDRPreset& p) {
  const bool orig = p.render_width == 0 || p.render_height == 0;
  const uint32_t W = orig ? r->dr_ref_w : p.render_width, H = orig ? r->dr_ref_h : p.render_height;
  if (W == r->dr_w && H == r->dr_h) return LRT_OK;
  hipError_t e;
  if ((e = hipSetDevice(r->device)) != hipSuccess) return hip_err(e, "hipSetDevice");
  const size_t np = (size_t)W * H;
  if ((e = r->d_dr_ref.ensure(np * 16 * r->dr_n_ref)) != hipSuccess) return hip_err(e, "alloc reference");
  if (W == r->dr_ref_w && H == r->dr_ref_h)
    e = hipMemcpyAsync(r->d_dr_ref.p, r->d_dr_ref_orig.p, np * 16 * r->dr_n_ref, hipMemcpyDeviceToDevice, r->stream);
  else
    e = launch_dr_resample((const float4*)r->d_dr_ref_orig.p, r->dr_ref_w, r->dr_ref_h, (float4*)r->d_dr_ref.p, W, H, r->dr_n_ref, r->stream);
  if (e != hipSuccess) return hip_err(e, "resample reference images");
  r->dr_w = W;
  r->dr_h = H;
  r->dr_img_valid.assign(r->dr_n_ref, 0);
  return LRT_OK;
}

// Slot tables of the scene's BVH for launch_dr_cull, built once per scene: the
// slot (2 * pair + side) that references each brick and each node pair, and
// whether the bricks under a slot tile its box. They describe a tree only if
// every pair is referenced at most once and after its parent, and every brick
// by at most one leaf; otherwise the full tree is used. A slot is tiled if all
// bricks of the scene share one LOD and no two share a position (so bricks
// never overlap) and the slot's brick count equals its box volume in bricks.
static int dr_ensure_tree(Renderer* r) {
  if (r->dr_tree_state != 0) return LRT_OK;
  const uint32_t np = (uint32_t)r->pairs.size(), nb = (uint32_t)r->nodes.size();
  const uint32_t none = 0xFFFFFFFFu;
  std::vector<uint32_t> pair_slot(np, none), leaf_slot(nb, none);
  bool ok = np > 0;
  for (uint32_t p = 0; p < np && ok; p++)
    for (uint32_t side = 0; side < 2 && ok; side++) {
      const uint32_t off = side ? r->pairs[p].right.leftOffset : r->pairs[p].left.leftOffset;
      if (off == none) continue;
      if (off & LRT_LEAF_BIT) {
        const uint32_t b = off & LRT_START_MASK;
        ok = b < nb && leaf_slot[b] == none;
        if (ok) leaf_slot[b] = 2 * p + side;
      } else {
        ok = off > p && off < np && pair_slot[off] == none;
        if (ok) pair_slot[off] = 2 * p + side;
      }
    }
  r->dr_tree_state = -1;
  if (!ok) return LRT_OK;
  // one LOD and distinct positions
  const uint32_t lod = r->nodes[0].pos_z_lod_size & 0xFFFFu;
  std::vector<uint64_t> keys(nb);
  bool one_lod = true;
  for (uint32_t i = 0; i < nb; i++) {
    const SdfSBSNode& n = r->nodes[i];
    one_lod = one_lod && (n.pos_z_lod_size & 0xFFFFu) == lod;
    keys[i] = ((uint64_t)n.pos_xy << 16) | (n.pos_z_lod_size >> 16);
  }
  std::sort(keys.begin(), keys.end());
  const bool disjoint = one_lod && std::adjacent_find(keys.begin(), keys.end()) == keys.end();
  // bricks per slot, children before parents (pairs follow their parent)
  std::vector<uint64_t> count(2 * (size_t)np, 0);
  for (uint32_t b = 0; b < nb; b++)
    if (leaf_slot[b] != none) count[leaf_slot[b]] = 1;
  for (uint32_t p = np; p-- > 0;)
    if (pair_slot[p] != none) count[pair_slot[p]] = count[2 * (size_t)p] + count[2 * (size_t)p + 1];
  std::vector<uint8_t> tiled(2 * (size_t)np, 0);
  for (uint32_t p = 0; p < np && disjoint; p++)
    for (uint32_t side = 0; side < 2; side++) {
      const BVHNode& n = side ? r->pairs[p].right : r->pairs[p].left;
      if (n.leftOffset == none) continue;
      double vol = 1.0;
      bool whole = true;
      for (int k = 0; k < 3; k++) {
        const double ext = ((double)n.boxMax[k] - (double)n.boxMin[k]) * 0.5 * (double)lod;
        whole = whole && ext > 0.5 && std::fabs(ext - std::nearbyint(ext)) < 1e-3;
        vol *= std::nearbyint(ext);
      }
      tiled[2 * (size_t)p + side] = whole && vol == (double)count[2 * (size_t)p + side] ? 1 : 0;
    }
  hipError_t e;
  if ((e = r->d_dr_leaf_slot.upload(leaf_slot.data(), (size_t)nb * 4, r->stream)) != hipSuccess) return hip_err(e, "upload");
  if ((e = r->d_dr_pair_slot.upload(pair_slot.data(), (size_t)np * 4, r->stream)) != hipSuccess) return hip_err(e, "upload");
  if ((e = r->d_dr_tiled.upload(tiled.data(), tiled.size(), r->stream)) != hipSuccess) return hip_err(e, "upload");
  if ((e = r->d_dr_occ.ensure((size_t)np * 16)) != hipSuccess) return hip_err(e, "alloc");
  if ((e = r->d_dr_cull_pairs.ensure((size_t)np * 2 * sizeof(BVHNodePair))) != hipSuccess) return hip_err(e, "alloc");
  if ((e = hipStreamSynchronize(r->stream)) != hipSuccess) return hip_err(e, "upload");
  r->dr_tree_state = 1;
  return LRT_OK;
}

// RenderDRFiniteDiff (MultiRendererDR.cpp:751-815) over the active range of
// :428-436 (geometry: every parameter before the 24 colour values per brick;
// otherwise those colour values): for each parameter, the image rendered with
// it at p0 + delta and at p0 - delta (primary rays only, no border rays), the
// summed per-pixel loss of each, and their scaled central difference. A
// parameter that no brick references leaves every image unchanged, so its
// difference is 0 and it is not rendered. Like the reference, the image and
// loss left behind are those of the last render (last parameter at p0 - delta).
static int dr_finite_diff(Renderer* r, const MultiRendererDRPreset& p, const lrtd::DRArgs& A, uint32_t* misc, uint32_t blocks, size_t P,
                          float g_mult, double* h_loss_sum) {
  const size_t n_col = 24 * r->nodes.size();
  if (P < n_col) return set_err(LRT_E_INVALID, "finite differences: fewer parameters than colour values");
  const bool geo = (p.dr_reconstruction_flags & LRT_DR_RECONSTRUCTION_FLAG_GEOMETRY) != 0;
  const size_t start = geo ? 0 : P - n_col, end = geo ? P - n_col : P;
  if (end <= start) return set_err(LRT_E_INVALID, "finite differences need a non-empty parameter range (MultiRendererDR.cpp:754)");
  const uint32_t nv = (uint32_t)A.S.v_size * A.S.v_size * A.S.v_size;
  std::vector<uint32_t> cnt(end - start + 1, 0);
  std::vector<uint8_t> used(end - start, 0);
  for (size_t b = 0; b < r->nodes.size(); b++) {
    const size_t off = r->nodes[b].data_offset;
    for (uint32_t k = 0; k < nv; k++) {
      const size_t j = r->values[off + k];
      if (j >= start && j < end) cnt[j - start + 1]++, used[j - start] = 1;
    }
    for (uint32_t c = 0; c < 8; c++)
      for (uint32_t e = 0; e < 3; e++) {
        const size_t j = (size_t)r->values[off + nv + c] + e;
        if (j >= start && j < end) used[j - start] = 1;
      }
  }
  for (size_t i = 1; i < cnt.size(); i++) cnt[i] += cnt[i - 1];
  std::vector<uint32_t> pos(cnt.back()), fill(cnt.begin(), cnt.end() - 1);
  for (size_t b = 0; b < r->nodes.size(); b++) {
    const size_t off = r->nodes[b].data_offset;
    for (uint32_t k = 0; k < nv; k++) {
      const size_t j = r->values[off + k];
      if (j >= start && j < end) pos[fill[j - start]++] = (uint32_t)(b * nv + k);
    }
  }
  std::vector<uint32_t> ids;
  for (size_t j = start; j < end; j++)
    if (used[j - start]) ids.push_back((uint32_t)j);
  const uint32_t n = (uint32_t)ids.size();
  std::vector<uint32_t> ids_up(ids), pos_up(pos);  // never empty: at least one word is uploaded
  if (ids_up.empty()) ids_up.push_back(0);
  if (pos_up.empty()) pos_up.push_back(0);
  hipError_t e;
  if ((e = r->d_dr_fd_p0.ensure(P * 4)) != hipSuccess) return hip_err(e, "alloc");
  if ((e = hipMemcpyAsync(r->d_dr_fd_p0.p, r->d_values_f.p, P * 4, hipMemcpyDeviceToDevice, r->stream)) != hipSuccess) return hip_err(e, "copy");
  if ((e = r->d_dr_fd_ids.upload(ids_up.data(), ids_up.size() * 4, r->stream)) != hipSuccess) return hip_err(e, "upload");
  if ((e = r->d_dr_fd_pos.upload(pos_up.data(), pos_up.size() * 4, r->stream)) != hipSuccess) return hip_err(e, "upload");
  if ((e = r->d_dr_fd_acc.ensure((2 * (size_t)n + 1) * 8)) != hipSuccess) return hip_err(e, "alloc");
  if ((e = hipStreamSynchronize(r->stream)) != hipSuccess) return hip_err(e, "upload");  // the host tables are locals
  float* vf = (float*)r->d_values_f.p;
  float* bv = (float*)r->d_dr_bvals.p;
  const float* p0 = (const float*)r->d_dr_fd_p0.p;
  const uint32_t* dpos = (const uint32_t*)r->d_dr_fd_pos.p;
  double* acc = (double*)r->d_dr_fd_acc.p;
  const uint32_t npx = A.W * A.H;
  const float delta = p.finite_diff_delta;
  auto render = [&](double* out) -> int {
    if ((e = launch_dr_primary(A, misc, (uint32_t*)r->d_spill.p, (float*)(misc + 4), blocks, r->stream, nullptr)) != hipSuccess)
      return hip_err(e, "dr primary");
    if ((e = launch_dr_fd_loss(A.out_image, A.ref, npx, A.loss_fn, out, r->stream)) != hipSuccess) return hip_err(e, "dr fd loss");
    return LRT_OK;
  };
  int c;
  for (uint32_t i = 0; i < n; i++) {
    const uint32_t j = ids[i], b0 = cnt[j - start], np_j = cnt[j - start + 1] - b0;
    for (uint32_t mode = 0; mode < 3; mode++) {
      if ((e = launch_dr_fd_set(vf, bv, p0, dpos + b0, np_j, j, mode, delta, misc, r->stream)) != hipSuccess) return hip_err(e, "dr fd set");
      if (mode < 2 && (c = render(acc + 2 * i + mode)) != LRT_OK) return c;
    }
  }
  size_t last = 2 * (size_t)n - 1;
  if (n == 0 || ids.back() != end - 1) {  // the last parameter is not referenced: its renders equal the unperturbed one
    if ((e = zero_u32(misc, 1, r->stream)) != hipSuccess) return hip_err(e, "clear");
    if ((c = render(acc + 2 * (size_t)n)) != LRT_OK) return c;
    last = 2 * (size_t)n;
  }
  if ((e = zero_u32((uint32_t*)r->d_dr_grad.p, (uint32_t)P, r->stream)) != hipSuccess) return hip_err(e, "clear grad");
  if ((e = launch_dr_fd_grad(acc, (const uint32_t*)r->d_dr_fd_ids.p, n, g_mult, delta, (float*)r->d_dr_grad.p, r->stream)) != hipSuccess)
    return hip_err(e, "dr fd grad");
  if ((e = hipMemcpyAsync(h_loss_sum, acc + last, 8, hipMemcpyDeviceToHost, r->stream)) != hipSuccess) return hip_err(e, "copy");
  return LRT_OK;
}

// one RenderDR call; gradient left in d_dr_grad (scaled)
// d_loss (optional, default mode only): the view's loss sum over its pixels is
// added to this device float and the call returns without waiting for the
// GPU (no timing, no statistics, dr_last_loss unchanged), so a loop of steps
// never stalls on the host.
// d_batch (optional, with d_loss): n_batch device records of the step's other
// images, rendered for their image and loss only in the same primary launch
// as `view` (MultiRendererDR.cpp:395-437 clears the gradient before every
// image of a batch, so only the last image's gradient reaches Adam).
static int dr_render_view(Renderer* r, const MultiRendererDRPreset& p, uint32_t view, uint32_t seed, int dtdv_mode, float* kernel_ms,
                          float* d_loss = nullptr, const lrtd::DRView* d_batch = nullptr, uint32_t n_batch = 0) {
  int c = dr_check(r, p);
  if (c) return c;
  if ((c = dr_prepare_size(r, p))) return c;
  if (view >= r->dr_n_ref) return set_err(LRT_E_INVALID, "view out of range");
  hipError_t e;
  if ((e = hipSetDevice(r->device)) != hipSuccess) return hip_err(e, "hipSetDevice");
  const uint32_t W = r->dr_w, H = r->dr_h;
  const size_t np = (size_t)W * H;
  const uint32_t P = (uint32_t)r->values_f.size();
  if ((e = r->d_dr_img.ensure(np * 16 * r->dr_n_ref)) != hipSuccess) return hip_err(e, "alloc");  // one slot per view
  if ((e = r->d_dr_thr.ensure(np * 4)) != hipSuccess) return hip_err(e, "alloc");
  if ((e = r->d_dr_grad.ensure((size_t)P * 4)) != hipSuccess) return hip_err(e, "alloc");
  if ((e = r->d_dr_list.ensure(np * 4)) != hipSuccess) return hip_err(e, "alloc");
  if ((e = r->d_dr_misc.ensure(64)) != hipSuccess) return hip_err(e, "alloc");
  const uint64_t vs = r->header.brick_size + 2 * r->header.brick_pad + 1;
  // the kernels address the gathered copy with 32-bit byte offsets (it is no
  // larger than the index words SetScene already holds under 4 GiB)
  if ((uint64_t)r->nodes.size() * vs * vs * vs * 4 > 0xFFFFFFFFull) return set_err(LRT_E_INVALID, "gathered brick values exceed 4 GiB");
  if ((e = r->d_dr_bvals.ensure((size_t)r->nodes.size() * vs * vs * vs * 4)) != hipSuccess) return hip_err(e, "alloc");
  if ((e = r->d_dr_bminmax.ensure((size_t)r->nodes.size() * 8)) != hipSuccess) return hip_err(e, "alloc");
  const uint32_t blocks = r->grid_blocks();
  if ((e = r->ensure_spill(blocks)) != hipSuccess) return hip_err(e, "alloc spill");
  uint32_t* misc = (uint32_t*)r->d_dr_misc.p;  // [0] work counter [1] border count [2] border hits [4] loss (float)
  if ((e = zero_u32(misc, 16, r->stream)) != hipSuccess) return hip_err(e, "clear");
  if ((e = zero_u32((uint32_t*)r->d_dr_grad.p, P, r->stream)) != hipSuccess) return hip_err(e, "clear grad");

  lrtd::DRArgs A{};
  A.S = r->dev_scene();
  const float* wv = &r->dr_world_views[16 * view];
  const float* pr = &r->dr_projs[16 * view];
  A.cam = make_camera(wv, pr);
  for (int cc = 0; cc < 4; cc++)
    for (int rr = 0; rr < 4; rr++) {
      float s = 0.f;
      for (int k = 0; k < 4; k++) s += pr[k * 4 + rr] * wv[cc * 4 + k];
      A.proj_view[cc * 4 + rr] = s;
    }
  A.Q.ray_flags = dr_ray_flags(p);
  A.Q.dtdv_fd = dtdv_mode == 1 ? 1u : 0u;
  A.W = W;
  A.H = H;
  A.packed_w = 8 * ((W + 7) / 8);
  A.spp = p.spp;
  A.seed = seed;
  A.mode = p.dr_render_mode;
  A.input_type = p.dr_input_type;
  A.loss_fn = p.dr_loss_function;
  A.border_spp = p.border_spp;
  A.forced_border = p.debug_forced_border ? 1u : 0u;
  A.relax_eps = p.border_relax_eps;
  A.border_sampling = p.dr_border_sampling;
  // CastBorderRaySVM's sample grid side, uint(sqrt(svm_spp)) of the reference
  // (MultiRendererDR.cpp:1664-1666): the exact integer square root
  {
    const uint32_t n = std::max(1u, (uint32_t)(0.25f * (float)p.border_spp));
    uint32_t q = 0;
    while ((uint64_t)(q + 1) * (q + 1) <= n) q++;
    A.svm_side = q;
  }
  for (int k = 0; k < 3; k++) A.up[k] = wv[4 * k + 1];  // worldView row 1, column-major storage
  A.lights = r->light_set();
  A.ref = (const float4*)r->d_dr_ref.p + np * view;
  A.out_image = (float4*)r->d_dr_img.p + np * view;
  A.depth_thr = (float*)r->d_dr_thr.p;
  A.grad = (float*)r->d_dr_grad.p;
  A.bvals = (const float*)r->d_dr_bvals.p;
  A.bminmax = (const float2*)r->d_dr_bminmax.p;
  const bool geometry_border = (p.dr_reconstruction_flags & LRT_DR_RECONSTRUCTION_FLAG_GEOMETRY) && p.border_spp > 0;
  const bool cull = dr_primary_is_depth(A) || (geometry_border && dr_border_is_depth(A));
  if (cull && (c = dr_ensure_tree(r)) != LRT_OK) return c;

  if ((e = hipEventRecord(r->ev0, r->stream)) != hipSuccess) return hip_err(e, "event");
  if ((e = launch_dr_gather(A.S, (float*)r->d_dr_bvals.p, (float2*)r->d_dr_bminmax.p, r->stream)) != hipSuccess) return hip_err(e, "dr gather");
  if (p.dr_diff_mode == LRT_DR_DIFF_MODE_FINITE_DIFF) {
    const float g_mult_fd = (float)(std::pow((double)P + 1.0, 2.0 / 3.0) / (double)(W * H));  // getBaseGradientMult
    double loss_sum = 0.0;
    lrtd::DRArgs Afd = A;
    Afd.depth_thr = nullptr;  // no border search in this mode (MultiRendererDR.cpp:613), so no depth threshold
    if ((c = dr_finite_diff(r, p, Afd, misc, blocks, P, g_mult_fd, &loss_sum)) != LRT_OK) return c;
    if ((e = hipEventRecord(r->ev1, r->stream)) != hipSuccess) return hip_err(e, "event");
    if ((e = hipStreamSynchronize(r->stream)) != hipSuccess) return hip_err(e, "dr finite differences");
    r->dr_last_loss = (float)(loss_sum / (double)np);
    r->dr_last_border = 0;  // border pixels are searched only in the default mode (MultiRendererDR.cpp:613)
    r->dr_last_border_hits = 0;
    r->dr_img_valid[view] = 1;
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, r->ev0, r->ev1);
    r->times["RenderDR"] = ms;
    if (kernel_ms) *kernel_ms = ms;
    return LRT_OK;
  }
  const uint4 *cull_primary = nullptr, *cull_border = nullptr;
  if (cull && r->dr_tree_state == 1) {
    const uint32_t np_pairs = (uint32_t)r->pairs.size();
    DRCullTables T{(const uint32_t*)r->d_dr_leaf_slot.p, (const uint32_t*)r->d_dr_pair_slot.p, (const uint8_t*)r->d_dr_tiled.p,
                   (uint32_t*)r->d_dr_occ.p, (uint4*)r->d_dr_cull_pairs.p, (uint4*)r->d_dr_cull_pairs.p + 4 * (size_t)np_pairs};
    if ((e = launch_dr_cull(T, (const float2*)r->d_dr_bminmax.p, A.relax_eps, A.S.pairs, (uint32_t)r->nodes.size(), np_pairs,
                            r->stream)) != hipSuccess)
      return hip_err(e, "dr cull");
    cull_primary = T.pairs_primary;
    cull_border = T.pairs_border;
  }
  float* loss_out = d_loss ? d_loss : (float*)(misc + 4);
  lrtd::DRArgs Ap = A;
  if (d_batch && n_batch > 1) {
    Ap.views = d_batch;
    Ap.n_views = n_batch;
  }
  if ((e = launch_dr_primary(Ap, misc, (uint32_t*)r->d_spill.p, loss_out, blocks, r->stream, cull_primary)) != hipSuccess)
    return hip_err(e, "dr primary");
  if ((e = launch_dr_border_find(A, (uint32_t*)r->d_dr_list.p, misc + 1, r->stream)) != hipSuccess) return hip_err(e, "dr border");
  if (geometry_border) {
    if ((e = launch_dr_border_rays(A, (uint32_t*)r->d_dr_list.p, misc + 1, (uint32_t*)r->d_spill.p, misc + 2, blocks, r->stream,
                                   cull_border)) != hipSuccess)
      return hip_err(e, "dr border rays");
  }
  // getBaseGradientMult, MultiRendererDR.cpp:140-145
  const float g_mult = (float)(std::pow((double)P + 1.0, 2.0 / 3.0) / (double)(W * H));
  if ((e = launch_dr_scale((float*)r->d_dr_grad.p, P, g_mult, r->stream)) != hipSuccess) return hip_err(e, "dr scale");
  r->dr_img_valid[view] = 1;
  if (d_loss) return LRT_OK;
  if ((e = hipEventRecord(r->ev1, r->stream)) != hipSuccess) return hip_err(e, "event");
  uint32_t h_misc[6];
  if ((e = hipMemcpyAsync(h_misc, misc, sizeof(h_misc), hipMemcpyDeviceToHost, r->stream)) != hipSuccess) return hip_err(e, "stats");
  if ((e = hipStreamSynchronize(r->stream)) != hipSuccess) return hip_err(e, "RenderDR");
  float loss;
  std::memcpy(&loss, &h_misc[4], sizeof(float));
  r->dr_last_loss = (float)((double)loss / (double)np);
  r->dr_last_border = h_misc[1];
  r->dr_last_border_hits = h_misc[2];
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, r->ev0, r->ev1);
  r->times["RenderDR"] = ms;
  if (kernel_ms) *kernel_ms = ms;
  return LRT_OK;
}

}  // namespace lrt

using namespace lrt;

extern "C" {

int lrt_dr_set_reference(lrt_handle h, const float* images, const float* masks, uint32_t n, uint32_t W, uint32_t H, const float* wvs,
                         const float* prs) {
  if (!h || !images || !wvs || !prs || n == 0 || W == 0 || H == 0 || W > 65535 || H > 65535) return set_err(LRT_E_INVALID, "bad arguments");
  const size_t np = (size_t)W * H;
  std::vector<float> ref(np * 4 * n);
  for (size_t i = 0; i < np * n; i++) {
    const float* c = images + 4 * i;
    float m;
    if (masks) m = masks[4 * i + 3];
    else m = std::sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]) < 0.001f ? 0.0f : 1.0f;  // CreateRefImageMasks, :169-185
    ref[4 * i + 0] = c[0];
    ref[4 * i + 1] = c[1];
    ref[4 * i + 2] = c[2];
    ref[4 * i + 3] = m;
  }
  hipError_t e;
  if ((e = hipSetDevice(h->device)) != hipSuccess) return hip_err(e, "hipSetDevice");
  if ((e = h->d_dr_ref_orig.upload(ref.data(), ref.size() * 4, h->stream)) != hipSuccess) return hip_err(e, "upload reference");
  if ((e = h->d_dr_ref.ensure(ref.size() * 4)) != hipSuccess) return hip_err(e, "alloc reference");
  if ((e = hipMemcpyAsync(h->d_dr_ref.p, h->d_dr_ref_orig.p, ref.size() * 4, hipMemcpyDeviceToDevice, h->stream)) != hipSuccess)
    return hip_err(e, "copy reference");
  if ((e = hipStreamSynchronize(h->stream)) != hipSuccess) return hip_err(e, "upload reference");
  h->dr_n_ref = n;
  h->dr_ref_w = h->dr_w = W;
  h->dr_ref_h = h->dr_h = H;
  h->dr_img_valid.assign(n, 0);
  h->dr_world_views.assign(wvs, wvs + 16 * n);
  h->dr_projs.assign(prs, prs + 16 * n);
  return LRT_OK;
}

int lrt_dr_render_grad(lrt_handle h, const MultiRendererDRPreset* p, uint32_t view, uint32_t seed, int dtdv_mode, float* out_image,
                       float* out_grad, float* loss, float* kernel_ms) {
  if (!h || !p) return set_err(LRT_E_INVALID, "null argument");
  int c = dr_render_view(h, *p, view, seed, dtdv_mode, kernel_ms);
  if (c) return c;
  h->dr_last_batch = 1;
  hipError_t e;
  const size_t np = (size_t)h->dr_w * h->dr_h;
  if (out_image && (e = hipMemcpy(out_image, (const float4*)h->d_dr_img.p + np * view, np * 16, hipMemcpyDeviceToHost)) != hipSuccess)
    return hip_err(e, "download image");
  if (out_grad && (e = hipMemcpy(out_grad, h->d_dr_grad.p, h->values_f.size() * 4, hipMemcpyDeviceToHost)) != hipSuccess)
    return hip_err(e, "download grad");
  if (loss) *loss = h->dr_last_loss;
  return LRT_OK;
}

int lrt_dr_optimize(lrt_handle h, const MultiRendererDRPreset* p, const SdfSBSHeader* header, const SdfSBSNode* nodes, uint32_t n_nodes,
                    const uint32_t* values, uint64_t n_values, float* values_f, uint64_t n_values_f, uint32_t seed, int dtdv_mode,
                    float* losses, float* total_ms) {
  if (!h || !p || !values_f) return set_err(LRT_E_INVALID, "null argument");
  if (p->opt_iterations == 0 || p->image_batch_size == 0) return set_err(LRT_E_INVALID, "opt_iterations and image_batch_size must be > 0");
  int c = lrt_set_scene_sbs(h, header, nodes, n_nodes, values, n_values, values_f, n_values_f);
  if (c) return c;
  if ((c = dr_check(h, *p))) return c;
  if ((c = dr_prepare_size(h, *p))) return c;
  const uint32_t P = (uint32_t)n_values_f;
  hipError_t e;
  if ((e = h->d_adam_v.ensure((size_t)P * 4)) != hipSuccess) return hip_err(e, "alloc adam");
  if ((e = h->d_adam_s.ensure((size_t)P * 4)) != hipSuccess) return hip_err(e, "alloc adam");
  if ((e = zero_u32((uint32_t*)h->d_adam_v.p, P, h->stream)) != hipSuccess) return hip_err(e, "clear");
  if ((e = zero_u32((uint32_t*)h->d_adam_s.p, P, h->stream)) != hipSuccess) return hip_err(e, "clear");
  // Redistance treats the first p^3 values as the dense lattice of a
  // brick_count^3 grid of bricks, p = brick_count * brick_size + 1
  // (MultiRendererDR.cpp:460-466, the layout the grid builders write)
  uint32_t rp = 0;
  if (p->redistancing_enable) {
    uint32_t bc = (uint32_t)std::llround(std::cbrt((double)n_nodes));
    while ((uint64_t)bc * bc * bc > n_nodes) bc--;
    while ((uint64_t)(bc + 1) * (bc + 1) * (bc + 1) <= n_nodes) bc++;
    rp = bc * header->brick_size + 1;
    if ((uint64_t)bc * bc * bc != n_nodes || (uint64_t)rp * rp * rp > n_values_f)
      return set_err(LRT_E_UNSUPPORTED, "redistancing needs a dense brick_count^3 grid whose first values are its lattice");
    if ((e = h->d_dr_redist.ensure((size_t)rp * rp * rp * 9)) != hipSuccess) return hip_err(e, "alloc redistancing");
  }
  // Default mode: each step's loss is summed on the device into its own slot
  // and the slots are read once after the loop, so the host never waits on a
  // step. The finite-difference mode synchronises inside every render anyway.
  const bool deferred = p->dr_diff_mode == LRT_DR_DIFF_MODE_DEFAULT;
  float* d_losses = nullptr;
  if (deferred) {
    if ((e = h->d_dr_losses.ensure((size_t)p->opt_iterations * 4)) != hipSuccess) return hip_err(e, "alloc losses");
    d_losses = (float*)h->d_dr_losses.p;
    if ((e = zero_u32((uint32_t*)d_losses, p->opt_iterations, h->stream)) != hipSuccess) return hip_err(e, "clear");
  }
  // MultiRendererDR.cpp:395-406: image b of step `it` (all views in order
  // when the batch is the whole set, otherwise a hashed choice)
  const uint32_t B = p->image_batch_size;
  auto batch_view = [&](uint32_t it, uint32_t b) {
    return B == h->dr_n_ref ? b : pcg_hash(seed ^ pcg_hash(it * 131u + b)) % h->dr_n_ref;
  };
  // Batches of several images in the default mode: the reference clears the
  // gradient before every image, so an image other than the last contributes
  // only its rendered image and its loss. Those images are rendered in the
  // last image's primary launch from per-image records, written for every
  // step before the loop; only the last image takes the gradient and border
  // passes.
  const bool batched = deferred && B > 1;
  if (batched) {
    const size_t np = (size_t)h->dr_w * h->dr_h;
    if ((e = h->d_dr_img.ensure(np * 16 * h->dr_n_ref)) != hipSuccess) return hip_err(e, "alloc");
    if ((e = h->d_dr_thr.ensure(np * 4)) != hipSuccess) return hip_err(e, "alloc");
    if ((e = h->d_dr_grad.ensure((size_t)P * 4)) != hipSuccess) return hip_err(e, "alloc");
    std::vector<lrtd::DRView> recs((size_t)p->opt_iterations * B);
    for (uint32_t it = 0; it < p->opt_iterations; it++)
      for (uint32_t b = 0; b < B; b++) {
        const uint32_t v = batch_view(it, b);
        lrtd::DRView& R = recs[(size_t)it * B + b];
        R.cam = make_camera(&h->dr_world_views[16 * v], &h->dr_projs[16 * v]);
        R.ref = (const float4*)h->d_dr_ref.p + np * v;
        R.out_image = (float4*)h->d_dr_img.p + np * v;
        R.depth_thr = b + 1 == B ? (float*)h->d_dr_thr.p : nullptr;
        R.grad = b + 1 == B ? (float*)h->d_dr_grad.p : nullptr;
      }
    if ((e = h->d_dr_views.upload(recs.data(), recs.size() * sizeof(lrtd::DRView), h->stream)) != hipSuccess) return hip_err(e, "upload views");
    if ((e = hipStreamSynchronize(h->stream)) != hipSuccess) return hip_err(e, "upload views");
  }
  auto t0 = std::chrono::steady_clock::now();
  h->dr_last_batch = B;
  for (uint32_t it = 0; it < p->opt_iterations; it++) {
    double lsum = 0;
    if (batched) {
      const lrtd::DRView* recs = (const lrtd::DRView*)h->d_dr_views.p + (size_t)it * B;
      if ((c = dr_render_view(h, *p, batch_view(it, B - 1), seed, dtdv_mode, nullptr, d_losses + it, recs, B))) return c;
      for (uint32_t b = 0; b < B; b++) h->dr_img_valid[batch_view(it, b)] = 1;
    } else {
      for (uint32_t b = 0; b < B; b++) {
        // the gradient buffer is cleared per image, as in the reference
        if ((c = dr_render_view(h, *p, batch_view(it, b), seed, dtdv_mode, nullptr, deferred ? d_losses + it : nullptr))) return c;
        lsum += h->dr_last_loss;
      }
    }
    if (losses && !deferred) losses[it] = (float)(lsum / B);
    if (p->reg_function != LRT_DR_REG_FUNCTION_NONE) {
      // Regularization, MultiRendererDR.cpp:453-455: after the step's renders,
      // added to the gradient before the batch division. The reference walks
      // brick ids below min(brick count, W*H) (its per-thread range is bounded
      // by the image size, :1962-1963); the same bricks are taken here.
      const uint32_t limit = (uint32_t)std::min<uint64_t>(h->nodes.size(), (uint64_t)h->dr_w * h->dr_h);
      if ((e = launch_dr_regularize(h->dev_scene(), limit, p->reg_lambda, p->reg_power, (float*)h->d_dr_grad.p, h->stream)) != hipSuccess)
        return hip_err(e, "regularization");
    }
    if (p->redistancing_enable && it % p->redistancing_interval == 0) {
      // MultiRendererDR.cpp:459-466: on the parameters, before the Adam step
      const size_t n3 = (size_t)rp * rp * rp;
      float* b0 = (float*)h->d_dr_redist.p;
      uint8_t* frozen = (uint8_t*)(b0 + 2 * n3);
      if ((e = launch_dr_redistance((float*)h->d_values_f.p, rp, 2.f / (float)rp, rp, frozen, b0, b0 + n3, h->stream)) != hipSuccess)
        return hip_err(e, "redistancing");
    }
    const float c1 = (float)(1.0 - std::pow((double)p->opt_beta_1, (double)(it + 1)));
    const float c2 = (float)(1.0 - std::pow((double)p->opt_beta_2, (double)(it + 1)));
    e = launch_adam((const float*)h->d_dr_grad.p, (float*)h->d_values_f.p, (float*)h->d_adam_v.p, (float*)h->d_adam_s.p, P, p->opt_lr,
                    p->opt_beta_1, p->opt_beta_2, p->opt_eps, c1, c2, (float)p->image_batch_size, h->stream);
    if (e != hipSuccess) return hip_err(e, "adam");
  }
  if ((e = hipMemcpyAsync(values_f, h->d_values_f.p, (size_t)P * 4, hipMemcpyDeviceToHost, h->stream)) != hipSuccess)
    return hip_err(e, "download values_f");
  std::vector<float> step_loss(deferred ? p->opt_iterations : 0);
  uint32_t h_misc[3] = {0, 0, 0};  // the last view's counters: [1] border pixels, [2] border rays that found a point
  if (deferred && (e = hipMemcpyAsync(step_loss.data(), d_losses, step_loss.size() * 4, hipMemcpyDeviceToHost, h->stream)) != hipSuccess)
    return hip_err(e, "download losses");
  if (deferred && (e = hipMemcpyAsync(h_misc, h->d_dr_misc.p, sizeof(h_misc), hipMemcpyDeviceToHost, h->stream)) != hipSuccess)
    return hip_err(e, "download stats");
  if ((e = hipStreamSynchronize(h->stream)) != hipSuccess) return hip_err(e, "optimize");
  if (deferred) {
    const double per = (double)h->dr_w * h->dr_h * p->image_batch_size;  // RenderDR's loss / (W*H), averaged over the batch
    for (uint32_t it = 0; it < p->opt_iterations; it++) {
      if (losses) losses[it] = (float)((double)step_loss[it] / per);
    }
    h->dr_last_loss = (float)((double)step_loss[p->opt_iterations - 1] / per);
    h->dr_last_border = h_misc[1];
    h->dr_last_border_hits = h_misc[2];
  }
  h->values_f.assign(values_f, values_f + P);
  h->bfeat_valid = false;  // brick features follow the optimised distances
  if (total_ms) *total_ms = (float)std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return LRT_OK;
}

// MultiRendererDR::OptimizeGrid, MultiRendererDR.cpp:228-276: start from a
// sphere grid (centre (0,-0.15,0), radius 0.7, brick size 1, gradient colours,
// with neighbour indices as create_grid_sbs returns it), and per preset run
// OptimizeFixedStructure, then resample the optimised lattice trilinearly into
// a grid with twice the brick size (the same size after the last preset when
// no_last_step_resize).
int lrt_dr_optimize_grid(lrt_handle h, uint32_t start_grid_size, int no_last_step_resize, const MultiRendererDRPreset* presets,
                         uint32_t n_presets, uint32_t seed, int dtdv_mode, lrt_sbs_handle* out, float* total_ms) {
  if (!h || !presets || n_presets == 0 || start_grid_size == 0 || start_grid_size > 256)
    return set_err(LRT_E_INVALID, "bad arguments");
  auto t0 = std::chrono::steady_clock::now();
  auto circle = [](f3 p) { return length(p - f3(0.f, -0.15f, 0.f)) - 0.7f; };  // circle_sdf, DR_common.cpp:16-19
  auto gradient_color = [](f3 p) {                                              // DR_common.cpp:25-35
    const f3 r(1, 0, 0), g(0, 1, 0);
    return (1 - p.x) * (1 - p.y) * (1 - p.z) * r + (1 - p.x) * (1 - p.y) * (p.z) * r + (1 - p.x) * (p.y) * (1 - p.z) * g +
           (1 - p.x) * (p.y) * (p.z) * g + (p.x) * (1 - p.y) * (1 - p.z) * r + (p.x) * (1 - p.y) * (p.z) * r +
           (p.x) * (p.y) * (1 - p.z) * g + (p.x) * (p.y) * (p.z) * g;
  };
  const uint32_t grid_size = start_grid_size;
  uint32_t brick_size = 1;
  SbsScene grid;
  try {
    grid = build_grid_sbs(grid_size, brick_size, circle, gradient_color, true);
  } catch (const std::exception& ex) {
    return set_err(LRT_E_INVALID, ex.what());
  }
  for (uint32_t i = 0; i < n_presets; i++) {
    int c = lrt_dr_optimize(h, &presets[i], &grid.header, grid.nodes.data(), (uint32_t)grid.nodes.size(), grid.values.data(),
                            grid.values.size(), grid.values_f.data(), grid.values_f.size(), seed, dtdv_mode, nullptr, nullptr);
    if (c) return c;
    const uint32_t p_count = grid_size * brick_size + 1u;
    const std::vector<float> vf = grid.values_f;
    auto sampler = [&](f3 p) {  // grid_sampler, MultiRendererDR.cpp:240-262
      const float gs = (float)grid_size, bs = (float)brick_size;
      auto cl = [](float v) { return std::min(std::max(v, 0.0f), 1.0f + 1e-6f); };
      const f3 idx_f(cl(0.5f * (p.x + 1.0f)) * gs * bs, cl(0.5f * (p.y + 1.0f)) * gs * bs, cl(0.5f * (p.z + 1.0f)) * gs * bs);
      const uint32_t ix = (uint32_t)(idx_f.x + 1e-6f), iy = (uint32_t)(idx_f.y + 1e-6f), iz = (uint32_t)(idx_f.z + 1e-6f);
      const f3 dp(idx_f.x - (float)ix, idx_f.y - (float)iy, idx_f.z - (float)iz);
      auto at = [&](uint32_t x, uint32_t y, uint32_t z) { return vf[(size_t)x * p_count * p_count + (size_t)y * p_count + z]; };
      if (ix < p_count - 1 && iy < p_count - 1 && iz < p_count - 1)
        return (1 - dp.x) * (1 - dp.y) * (1 - dp.z) * at(ix, iy, iz) + (1 - dp.x) * (1 - dp.y) * dp.z * at(ix, iy, iz + 1) +
               (1 - dp.x) * dp.y * (1 - dp.z) * at(ix, iy + 1, iz) + (1 - dp.x) * dp.y * dp.z * at(ix, iy + 1, iz + 1) +
               dp.x * (1 - dp.y) * (1 - dp.z) * at(ix + 1, iy, iz) + dp.x * (1 - dp.y) * dp.z * at(ix + 1, iy, iz + 1) +
               dp.x * dp.y * (1 - dp.z) * at(ix + 1, iy + 1, iz) + dp.x * dp.y * dp.z * at(ix + 1, iy + 1, iz + 1);
      return at(ix, iy, iz);
    };
    const uint32_t new_brick_size = (i == n_presets - 1 && no_last_step_resize) ? brick_size : 2 * brick_size;
    if (grid_size * new_brick_size > 4096) return set_err(LRT_E_INVALID, "grid too fine");
    grid = build_grid_sbs(grid_size, new_brick_size, sampler, gradient_color, true);
    brick_size = new_brick_size;
  }
  if (out) {
    std::unique_ptr<lrt_sbs> s(new lrt_sbs);
    s->s = std::move(grid);
    *out = s.release();
  }
  if (total_ms) *total_ms = (float)std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return LRT_OK;
}

int lrt_dr_get_last_dloss_ds(lrt_handle h, float* out, uint64_t n) {
  if (!h || !out) return set_err(LRT_E_INVALID, "null argument");
  if (!h->d_dr_grad.p || n > h->values_f.size()) return set_err(LRT_E_INVALID, "no gradient or size too large");
  hipError_t e = hipMemcpy(out, h->d_dr_grad.p, n * 4, hipMemcpyDeviceToHost);
  if (e != hipSuccess) return hip_err(e, "download grad");
  // the reference clears the gradient before every image of a step and divides
  // the last one by image_batch_size (MultiRendererDR.cpp:405, 475-476)
  if (h->dr_last_batch > 1)
    for (uint64_t i = 0; i < n; i++) out[i] /= (float)h->dr_last_batch;
  return LRT_OK;
}

int lrt_dr_get_last_image(lrt_handle h, uint32_t view, float* out) {
  if (!h || !out) return set_err(LRT_E_INVALID, "null argument");
  if (view >= h->dr_n_ref || !h->dr_img_valid[view]) return set_err(LRT_E_INVALID, "no RenderDR image for this view");
  const size_t np = (size_t)h->dr_w * h->dr_h;
  hipError_t e = hipMemcpy(out, (const float4*)h->d_dr_img.p + np * view, np * 16, hipMemcpyDeviceToHost);
  return e == hipSuccess ? LRT_OK : hip_err(e, "download image");
}

int lrt_dr_render_size(lrt_handle h, uint32_t* width, uint32_t* height) {
  if (!h) return set_err(LRT_E_INVALID, "null handle");
  if (width) *width = h->dr_w;
  if (height) *height = h->dr_h;
  return LRT_OK;
}

int lrt_dr_last_stats(lrt_handle h, uint32_t* border_pixels, uint32_t* border_rays_hit) {
  if (!h) return set_err(LRT_E_INVALID, "null handle");
  if (border_pixels) *border_pixels = h->dr_last_border;
  if (border_rays_hit) *border_rays_hit = h->dr_last_border_hits;
  return LRT_OK;
}

}  // extern "C"

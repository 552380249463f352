// lrt_capi.hip - C ABI implementation: scene ingest, device buffers, launches.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>
#include "host/hydra_scene.hpp"

#include "../../include/lrt.h"
#include "host/bvh_build.hpp"
#include "host/hmath.hpp"
#include "host/scene_gen.hpp"
#include "kernels/render.hpp"
#include "lrt_internal.hpp"

namespace lrt {

static thread_local std::string g_err;
int set_err(int code, const std::string& msg) {
  g_err = msg;
  return code;
}
int hip_err(hipError_t e, const char* what) { return set_err(LRT_E_HIP, std::string(what) + ": " + hipGetErrorString(e)); }

void DevBuf::release() {
  if (p) (void)hipFree(p);
  p = nullptr;
  bytes = 0;
}
hipError_t DevBuf::ensure(size_t n) {
  if (n <= bytes && p) return hipSuccess;
  release();
  hipError_t e = hipMalloc(&p, n ? n : 16);
  if (e == hipSuccess) bytes = n ? n : 16;
  return e;
}
hipError_t DevBuf::upload(const void* src, size_t n, hipStream_t s) {
  hipError_t e = ensure(n);
  if (e != hipSuccess) return e;
  if (n) e = hipMemcpyAsync(p, src, n, hipMemcpyHostToDevice, s);
  return e;
}

void inverse4x4(const float in[16], float out[16]) {
  double m[16], inv[16];
  for (int i = 0; i < 16; i++) m[i] = in[i];
  inv[0] = m[5] * m[10] * m[15] - m[5] * m[11] * m[14] - m[9] * m[6] * m[15] + m[9] * m[7] * m[14] + m[13] * m[6] * m[11] - m[13] * m[7] * m[10];
  inv[4] = -m[4] * m[10] * m[15] + m[4] * m[11] * m[14] + m[8] * m[6] * m[15] - m[8] * m[7] * m[14] - m[12] * m[6] * m[11] + m[12] * m[7] * m[10];
  inv[8] = m[4] * m[9] * m[15] - m[4] * m[11] * m[13] - m[8] * m[5] * m[15] + m[8] * m[7] * m[13] + m[12] * m[5] * m[11] - m[12] * m[7] * m[9];
  inv[12] = -m[4] * m[9] * m[14] + m[4] * m[10] * m[13] + m[8] * m[5] * m[14] - m[8] * m[6] * m[13] - m[12] * m[5] * m[10] + m[12] * m[6] * m[9];
  inv[1] = -m[1] * m[10] * m[15] + m[1] * m[11] * m[14] + m[9] * m[2] * m[15] - m[9] * m[3] * m[14] - m[13] * m[2] * m[11] + m[13] * m[3] * m[10];
  inv[5] = m[0] * m[10] * m[15] - m[0] * m[11] * m[14] - m[8] * m[2] * m[15] + m[8] * m[3] * m[14] + m[12] * m[2] * m[11] - m[12] * m[3] * m[10];
  inv[9] = -m[0] * m[9] * m[15] + m[0] * m[11] * m[13] + m[8] * m[1] * m[15] - m[8] * m[3] * m[13] - m[12] * m[1] * m[11] + m[12] * m[3] * m[9];
  inv[13] = m[0] * m[9] * m[14] - m[0] * m[10] * m[13] - m[8] * m[1] * m[14] + m[8] * m[2] * m[13] + m[12] * m[1] * m[10] - m[12] * m[2] * m[9];
  inv[2] = m[1] * m[6] * m[15] - m[1] * m[7] * m[14] - m[5] * m[2] * m[15] + m[5] * m[3] * m[14] + m[13] * m[2] * m[7] - m[13] * m[3] * m[6];
  inv[6] = -m[0] * m[6] * m[15] + m[0] * m[7] * m[14] + m[4] * m[2] * m[15] - m[4] * m[3] * m[14] - m[12] * m[2] * m[7] + m[12] * m[3] * m[6];
  inv[10] = m[0] * m[5] * m[15] - m[0] * m[7] * m[13] - m[4] * m[1] * m[15] + m[4] * m[3] * m[13] + m[12] * m[1] * m[7] - m[12] * m[3] * m[5];
  inv[14] = -m[0] * m[5] * m[14] + m[0] * m[6] * m[13] + m[4] * m[1] * m[14] - m[4] * m[2] * m[13] - m[12] * m[1] * m[6] + m[12] * m[2] * m[5];
  inv[3] = -m[1] * m[6] * m[11] + m[1] * m[7] * m[10] + m[5] * m[2] * m[11] - m[5] * m[3] * m[10] - m[9] * m[2] * m[7] + m[9] * m[3] * m[6];
  inv[7] = m[0] * m[6] * m[11] - m[0] * m[7] * m[10] - m[4] * m[2] * m[11] + m[4] * m[3] * m[10] + m[8] * m[2] * m[7] - m[8] * m[3] * m[6];
  inv[11] = -m[0] * m[5] * m[11] + m[0] * m[7] * m[9] + m[4] * m[1] * m[11] - m[4] * m[3] * m[9] - m[8] * m[1] * m[7] + m[8] * m[3] * m[5];
  inv[15] = m[0] * m[5] * m[10] - m[0] * m[6] * m[9] - m[4] * m[1] * m[10] + m[4] * m[2] * m[9] + m[8] * m[1] * m[6] - m[8] * m[2] * m[5];
  double det = m[0] * inv[0] + m[1] * inv[4] + m[2] * inv[8] + m[3] * inv[12];
  double idet = (det != 0.0) ? 1.0 / det : 0.0;
  for (int i = 0; i < 16; i++) out[i] = (float)(inv[i] * idet);
}

lrtd::Camera make_camera(const float wv[16], const float pr[16]) {
  lrtd::Camera c;
  inverse4x4(pr, c.projInv);
  inverse4x4(wv, c.worldViewInv);
  // eye_ray's mul4(worldViewInv, 0, 0, 0, 1), operation for operation
  const float* m = c.worldViewInv;
  for (int i = 0; i < 4; i++) c.origin[i] = m[i] * 0.0f + m[4 + i] * 0.0f + m[8 + i] * 0.0f + m[12 + i] * 1.0f;
  return c;
}

lrtd::DevScene Renderer::dev_scene() const {
  lrtd::DevScene S;
  S.pairs = (const uint4*)d_pairs.p;
  S.nodes = (const uint4*)d_nodes.p;
  S.values = (const uint32_t*)d_values.p;
  S.values_f = (const float*)d_values_f.p;
  S.brecs = (const float4*)d_brecs.p;
  S.brick_size = header.brick_size;
  S.brick_pad = header.brick_pad;
  S.bytes_per_value = header.bytes_per_value;
  S.layout = header.aux_data & LRT_SBS_LAYOUT_MASK;
  S.v_size = header.brick_size + 2 * header.brick_pad + 1;
  S.quant_voxel_diag = quant_voxel_diag;
  S.type = geom_type;
  S.n_nodes = (uint32_t)nodes.size();
  S.pairs_bytes = (uint32_t)(pairs.size() * sizeof(BVHNodePair));
  S.values_bytes = (uint32_t)(values.size() * 4);
  S.v_size_f = (float)S.v_size;
  return S;
}

lrtd::TraceParams Renderer::trace_params(bool force_normal) const {
  lrtd::TraceParams P;
  P.sdf_node_intersect = preset.sdf_node_intersect;
  P.representation_mode = preset.representation_mode;
  P.render_mode = preset.render_mode;
  P.normal_mode = preset.normal_mode;
  const uint32_t m = preset.render_mode;
  P.need_normal = (force_normal || m == LRT_RENDER_MODE_LAMBERT_NO_TEX || m == LRT_RENDER_MODE_NORMAL || m == LRT_RENDER_MODE_PHONG_NO_TEX ||
                   m == LRT_RENDER_MODE_LAMBERT || m == LRT_RENDER_MODE_PHONG)
                      ? 1u
                      : 0u;
  return P;
}

uint32_t Renderer::grid_blocks() const {
  if (launch_blocks) return launch_blocks;
  return (uint32_t)std::max(1, cu_count) * 8u;
}

hipError_t Renderer::ensure_spill(uint32_t blocks, uint32_t lds_entries) {
  uint32_t extra = (uint32_t)std::max(0, pair_depth + 2 - (int)lds_entries);
  size_t n = (size_t)blocks * 256 * std::max<uint32_t>(extra, 1u);
  return d_spill.ensure(n * sizeof(uint32_t));
}

// Boxes for the tile cull, two levels of the BVH's child slots whose union
// holds every brick: the coarse level (at most 8 slots: the root pair's two,
// then every inner slot replaced by its pair's two, level by level, while
// the list stays within 8) and below each coarse slot its descendants at the
// fine level (the same expansion of the whole list while it stays within 64;
// a leaf stays as it is). Layout, {min, max} float4 pairs: the n_coarse coarse
// boxes, whose w components hold the [first, end) range of their fine boxes
// as uint32 bits, then the fine boxes grouped by coarse box.
std::vector<float> cull_boxes(const std::vector<BVHNodePair>& pairs, uint32_t* n_coarse) {
  std::vector<float> out;
  *n_coarse = 0;
  if (pairs.empty()) return out;
  struct Slot { const BVHNode* n; uint32_t root; };  // root: index of the slot above it on the level before
  auto expand = [&](const std::vector<Slot>& cur, size_t limit, std::vector<Slot>& res) {
    res = cur;
    for (;;) {
      std::vector<Slot> next;
      for (const Slot& sl : res) {
        if (sl.n->leftOffset & LRT_LEAF_BIT) next.push_back(sl);
        else {
          const BVHNodePair& p = pairs[sl.n->leftOffset];
          next.push_back({&p.left, sl.root});
          next.push_back({&p.right, sl.root});
        }
      }
      if (next.size() > limit || next.size() == res.size()) break;
      res.swap(next);
    }
  };
  // three cuts through the tree, each one below the last (a leaf stays in
  // every later cut); expansion keeps children next to each other, so the
  // slots under one upper slot come out contiguous and in order
  std::vector<Slot> coarse, mid, fine;
  expand({{&pairs[0].left, 0u}, {&pairs[0].right, 0u}}, 8, coarse);
  for (uint32_t c = 0; c < coarse.size(); c++) coarse[c].root = c;
  expand(coarse, 64, mid);
  std::vector<uint32_t> mid_parent, fine_parent;
  for (uint32_t m = 0; m < mid.size(); m++) mid_parent.push_back(mid[m].root), mid[m].root = m;
  expand(mid, 512, fine);
  for (const Slot& sl : fine) fine_parent.push_back(sl.root);
  auto put = [&](const BVHNode* n, uint32_t w0, uint32_t w1) {
    float b[8] = {n->boxMin[0], n->boxMin[1], n->boxMin[2], 0.0f, n->boxMax[0], n->boxMax[1], n->boxMax[2], 0.0f};
    std::memcpy(&b[3], &w0, 4);
    std::memcpy(&b[7], &w1, 4);
    out.insert(out.end(), b, b + 8);
  };
  // [first, end) of the slots of a lower level (parent[f]: the upper slot
  // above slot f) under upper slot u, as indices into the table (the lower
  // level starts at `at`)
  auto range = [](const std::vector<uint32_t>& parent, uint32_t u, uint32_t at, uint32_t* first, uint32_t* end) {
    *first = (uint32_t)parent.size();
    *end = 0;
    for (uint32_t f = 0; f < parent.size(); f++)
      if (parent[f] == u) { *first = std::min(*first, f); *end = f + 1; }
    *first += at;
    *end += at;
  };
  const uint32_t nc = (uint32_t)coarse.size(), nm = (uint32_t)mid.size();
  uint32_t first, end;
  for (uint32_t c = 0; c < nc; c++) {
    range(mid_parent, c, nc, &first, &end);
    put(coarse[c].n, first, end);
  }
  for (uint32_t m = 0; m < nm; m++) {
    range(fine_parent, m, nc + nm, &first, &end);
    put(mid[m].n, first, end);
  }
  for (const Slot& sl : fine) put(sl.n, 0u, 0u);
  *n_coarse = nc;
  return out;
}

// Per-brick constants of the voxel walk (BVH2Common.cpp:785-795, 749-750),
// evaluated once here with exactly the float operations the per-ray code
// used to repeat: 12 floats per brick, {bmin.xyz, data_offset bits},
// {sz_inv, d, d_inv, qscale}, {d_max, mult, 0.5*sz, 0}.
std::vector<float> brick_records(const std::vector<SdfSBSNode>& nodes, const SdfSBSHeader& h, uint32_t quant_voxel_diag) {
  std::vector<float> out(nodes.size() * 12, 0.0f);
  const float bsf = (float)h.brick_size;
  const uint32_t bits = 8 * h.bytes_per_value;
  const uint32_t max_val = h.bytes_per_value == 4 ? 0xFFFFFFFFu : ((1u << bits) - 1);
  for (size_t i = 0; i < nodes.size(); i++) {
    const SdfSBSNode& n = nodes[i];
    const float px = (float)(n.pos_xy >> 16), py = (float)(n.pos_xy & 0xFFFFu), pz = (float)(n.pos_z_lod_size >> 16);
    const float sz = (float)(n.pos_z_lod_size & 0xFFFFu);
    const float sz_inv = 2.0f / sz;
    const float d = 2.0f / (sz * bsf);
    const float d_inv = 1.0f / d;
    const float qscale = 0.5f * sz * bsf;
    const float d_max = quant_voxel_diag ? 1.73205081f * (sz_inv / bsf) : 1.73205081f * sz_inv;
    const float mult = 2.0f * d_max / (float)max_val;
    float* o = out.data() + 12 * i;
    o[0] = -1.0f + sz_inv * px;
    o[1] = -1.0f + sz_inv * py;
    o[2] = -1.0f + sz_inv * pz;
    std::memcpy(&o[3], &n.data_offset, 4);
    o[4] = sz_inv; o[5] = d; o[6] = d_inv; o[7] = qscale;
    o[8] = d_max; o[9] = mult; o[10] = 0.5f * sz;
  }
  return out;
}

static int check_ready(Renderer* r) {
  if (!r) return set_err(LRT_E_INVALID, "null handle");
  if (!r->has_scene) return set_err(LRT_E_NOSCENE, "no scene: call lrt_set_scene_sbs first");
  return LRT_OK;
}

// Packed pixel count of one W x H image (both sides rounded up to multiples
// of 8); the kernels index a view's packed pixels with 32-bit integers.
static bool packed_fits(uint32_t W, uint32_t H) {
  return (uint64_t)(8 * ((W + 7) / 8)) * (uint64_t)(8 * ((H + 7) / 8)) <= 0xFFFFFFFFull;
}

// Do the preset's sample offsets (sample_offset) stay within [0, 1]^2 of
// their pixel? Random jitter does; the regular pattern does when spp is a
// perfect square. The tile cull relies on it.
static bool samples_in_pixel(const MultiRenderPreset& p) {
  if (p.spp == 0) return false;
  if (p.ray_gen_mode == LRT_RAY_GEN_MODE_RANDOM) return true;
  const uint32_t spp = p.spp;
  uint32_t q = 0;
  while ((q + 1) * (q + 1) <= spp) q++;
  return q * q == spp;
}

// Launches n_views cameras at W x H into the handle's output buffers.
int Renderer::render_device(const lrtd::Camera* cams, uint32_t n_views, uint32_t W, uint32_t H, bool rgba, bool gbuf, bool force_normal,
                            const char* timer_name, unsigned long long* d_counts) {
  if (W == 0 || H == 0 || W > 65535 || H > 65535 || n_views == 0) return set_err(LRT_E_INVALID, "bad image size");
  if (!packed_fits(W, H)) return set_err(LRT_E_INVALID, "image too large: the 8-aligned pixel count must fit 32 bits");
  if (preset.spp == 0) return set_err(LRT_E_INVALID, "spp must be >= 1");
  if (preset.interpolation_mode != LRT_INTERPOLATION_MODE_TRILINEAR)
    return set_err(LRT_E_UNSUPPORTED, "only trilinear interpolation is supported");
  hipError_t e;
  if ((e = hipSetDevice(device)) != hipSuccess) return hip_err(e, "hipSetDevice");
  const size_t npix = (size_t)n_views * W * H;
  if ((e = d_cams.upload(cams, sizeof(lrtd::Camera) * n_views, stream)) != hipSuccess) return hip_err(e, "upload cameras");
  if ((e = d_counter.ensure(lrtd::kXcds * lrtd::kCounterStride * 4)) != hipSuccess) return hip_err(e, "alloc counter");
  lrtd::RenderOut O{};
  if (gbuf) {
    if ((e = d_out_t.ensure(npix * 4)) != hipSuccess) return hip_err(e, "alloc t");
    if ((e = d_out_n.ensure(npix * 8)) != hipSuccess) return hip_err(e, "alloc n");
    if ((e = d_out_prim.ensure(npix * 4)) != hipSuccess) return hip_err(e, "alloc prim");
    O.t = (float*)d_out_t.p;
    O.n = (float2*)d_out_n.p;
    O.prim = (uint32_t*)d_out_prim.p;
  }
  if (rgba) {
    if ((e = d_out_rgba.ensure(npix * 16)) != hipSuccess) return hip_err(e, "alloc rgba");
    O.rgba = (float4*)d_out_rgba.p;
  }
  lrtd::RenderArgs A;
  lrtd::set_image(A, W, H);
  if ((uint64_t)n_views * A.n_packed >= (1ull << 38)) return set_err(LRT_E_INVALID, "too many pixels in one launch (views x packed pixels >= 2^38)");
  A.n_views = n_views;
  A.spp = preset.spp;
  A.ray_gen_mode = preset.ray_gen_mode;
  A.seed = seed;
  A.lights = light_set();
  const uint32_t blocks = grid_blocks();
  const bool seeded = seeding && mlp_ready;
  if (tile_cull && n_cull > 0 && samples_in_pixel(preset)) {
    if ((e = d_tile_mask.ensure(tile_mask_words(A) * 4)) != hipSuccess) return hip_err(e, "alloc tile mask");
    A.tile_mask = (const uint32_t*)d_tile_mask.p;
  }
  if ((e = ensure_spill(blocks, seeded ? lrtd::kStackLdsSeeded : lrtd::kStackLds)) != hipSuccess) return hip_err(e, "alloc spill");
  if (seeded && (e = ensure_brick_features()) != hipSuccess) return hip_err(e, "brick features");
  if ((e = hipEventRecord(ev0, stream)) != hipSuccess) return hip_err(e, "event");
  if (seeded) {
    const float logit = std::log(seed_threshold / (1.0f - seed_threshold));
    e = launch_render_seeded(dev_scene(), trace_params(force_normal), (const float4*)d_bfeat.p, d_mlp_packed.p, logit,
                             (const lrtd::Camera*)d_cams.p, A, O, (uint32_t*)d_counter.p, (uint32_t*)d_spill.p, blocks, rgba, d_counts,
                             stream, (const float4*)d_cull_boxes.p, n_cull);
  } else {
    e = launch_render(dev_scene(), trace_params(force_normal), (const lrtd::Camera*)d_cams.p, A, O, (uint32_t*)d_counter.p,
                      (uint32_t*)d_spill.p, blocks, rgba, d_counts, stream, (const float4*)d_cull_boxes.p, n_cull);
  }
  if (e != hipSuccess) return hip_err(e, "render launch");
  if ((e = hipEventRecord(ev1, stream)) != hipSuccess) return hip_err(e, "event");
  last_w = W;
  last_h = H;
  last_views = n_views;
  pending_timer = timer_name ? timer_name : "";
  return LRT_OK;
}

int Renderer::finish_timer() {
  hipError_t e = hipEventSynchronize(ev1);
  if (e != hipSuccess) return hip_err(e, "kernel execution");
  float ms = 0.f;
  if (hipEventElapsedTime(&ms, ev0, ev1) == hipSuccess && !pending_timer.empty()) times[pending_timer] = ms;
  last_ms = ms;
  return LRT_OK;
}

}  // namespace lrt

using namespace lrt;

extern "C" {

const char* lrt_last_error(void) { return g_err.c_str(); }

int lrt_create(uint32_t device, uint32_t max_primitives, lrt_handle* out) {
  (void)max_primitives;
  if (!out) return set_err(LRT_E_INVALID, "null out");
  if (device != LRT_DEVICE_GPU && device != LRT_DEVICE_GPU_RTX)
    return set_err(LRT_E_UNSUPPORTED, "this library renders on the GPU only (DEVICE_GPU)");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return set_err(LRT_E_NODEVICE, "no HIP device visible");
  std::unique_ptr<lrt_renderer> r(new lrt_renderer);
  hipError_t e;
  if ((e = hipGetDevice(&r->device)) != hipSuccess) return hip_err(e, "hipGetDevice");
  hipDeviceProp_t prop;
  if ((e = hipGetDeviceProperties(&prop, r->device)) != hipSuccess) return hip_err(e, "hipGetDeviceProperties");
  r->cu_count = prop.multiProcessorCount;
  if ((e = hipStreamCreateWithFlags(&r->stream, hipStreamNonBlocking)) != hipSuccess) return hip_err(e, "hipStreamCreate");
  if ((e = hipEventCreate(&r->ev0)) != hipSuccess) return hip_err(e, "hipEventCreate");
  if ((e = hipEventCreate(&r->ev1)) != hipSuccess) return hip_err(e, "hipEventCreate");
  r->preset = MultiRenderPreset{LRT_RENDER_MODE_LAMBERT_NO_TEX, LRT_SDF_NODE_INTERSECT_ST, LRT_NORMAL_MODE_GEOMETRY,
                                LRT_RAY_GEN_MODE_REGULAR, LRT_INTERPOLATION_MODE_TRILINEAR, LRT_REPRESENTATION_MODE_VOLUME, 1};
  *out = r.release();
  return LRT_OK;
}

int lrt_set_lights(lrt_handle h, const LrtLight* lights, uint32_t n) {
  if (!h || (n && !lights)) return set_err(LRT_E_INVALID, "null argument");
  if (n > LRT_MAX_LIGHTS) return set_err(LRT_E_INVALID, "at most LRT_MAX_LIGHTS lights");
  for (uint32_t i = 0; i < n; i++)
    if (lights[i].type > LRT_LIGHT_TYPE_AMBIENT) return set_err(LRT_E_INVALID, "bad light type");
  h->lights.assign(lights, lights + n);
  return LRT_OK;
}

int lrt_destroy(lrt_handle h) {
  if (!h) return LRT_OK;
  lrt_renderer* r = h;
  (void)hipSetDevice(r->device);
  if (r->stream) (void)hipStreamSynchronize(r->stream);
  r->release_all();
  if (r->ev0) (void)hipEventDestroy(r->ev0);
  if (r->ev1) (void)hipEventDestroy(r->ev1);
  if (r->stream) (void)hipStreamDestroy(r->stream);
  delete r;
  return LRT_OK;
}

int lrt_set_scene_sbs(lrt_handle h, const SdfSBSHeader* header, const SdfSBSNode* nodes, uint32_t n_nodes, const uint32_t* values,
                      uint64_t n_values, const float* values_f, uint64_t n_values_f) {
  if (!h || !header || !nodes || !values) return set_err(LRT_E_INVALID, "null argument");
  if (n_nodes == 0 || n_values == 0) return set_err(LRT_E_INVALID, "empty scene (reference asserts size > 0, BVH2Common_host.cpp:569)");
  if (n_nodes >= (1u << 24)) return set_err(LRT_E_INVALID, "too many bricks for the 24-bit leaf index");
  const uint32_t layout = header->aux_data & LRT_SBS_LAYOUT_MASK;
  const bool indexed = layout == LRT_SBS_LAYOUT_ID32F_IRGB32F || layout == LRT_SBS_LAYOUT_ID32F_IRGB32F_IN;
  if (header->brick_size < 1 || header->brick_size > 16 || header->brick_pad > 1)
    return set_err(LRT_E_INVALID, "brick_size must be 1..16 and brick_pad 0..1");
  if (!indexed && layout != LRT_SBS_LAYOUT_DX && layout != LRT_SBS_LAYOUT_UNDEFINED && layout != LRT_SBS_LAYOUT_DX_RGB8 &&
      layout != LRT_SBS_LAYOUT_DX_UV16)
    return set_err(LRT_E_UNSUPPORTED, "unsupported node layout");
  if (!indexed && header->bytes_per_value != 1 && header->bytes_per_value != 2 && header->bytes_per_value != 4)
    return set_err(LRT_E_INVALID, "bytes_per_value must be 1, 2 or 4");
  if (indexed && (!values_f || n_values_f == 0)) return set_err(LRT_E_INVALID, "indexed layout needs values_f");
  // bounds: every brick's words must lie inside `values`
  const uint32_t v_size = header->brick_size + 2 * header->brick_pad + 1;
  const uint64_t nv = (uint64_t)v_size * v_size * v_size;
  uint64_t words = indexed ? nv + 8 + (layout == LRT_SBS_LAYOUT_ID32F_IRGB32F_IN ? 27 : 0)
                           : (nv + (4 / header->bytes_per_value) - 1) / (4 / header->bytes_per_value);
  for (uint32_t i = 0; i < n_nodes; i++) {
    if ((uint64_t)nodes[i].data_offset + words > n_values) return set_err(LRT_E_INVALID, "brick data_offset out of range");
    uint32_t sz = nodes[i].pos_z_lod_size & 0xFFFF;
    if (sz == 0 || (nodes[i].pos_xy >> 16) >= sz || (nodes[i].pos_xy & 0xFFFF) >= sz || (nodes[i].pos_z_lod_size >> 16) >= sz)
      return set_err(LRT_E_INVALID, "brick position outside its LOD");
  }
  if (indexed) {
    for (uint32_t i = 0; i < n_nodes; i++) {
      const uint32_t* w = values + nodes[i].data_offset;
      for (uint64_t k = 0; k < nv; k++)
        if (w[k] >= n_values_f) return set_err(LRT_E_INVALID, "distance index out of range");
      for (int k = 0; k < 8; k++)
        if ((uint64_t)w[nv + k] + 3 > n_values_f) return set_err(LRT_E_INVALID, "colour index out of range");
      if (layout == LRT_SBS_LAYOUT_ID32F_IRGB32F_IN)
        for (int k = 0; k < 27; k++)
          if (w[nv + 8 + k] != LRT_INVALID_IDX && w[nv + 8 + k] >= n_nodes) return set_err(LRT_E_INVALID, "neighbour index out of range");
    }
  }
  // the kernels address the values through 32-bit byte offsets (the BVH's
  // < 2^24 pairs of 64 B fit them by the leaf-index limit above)
  if ((uint64_t)n_values * 4 > 0xFFFFFFFFull) return set_err(LRT_E_INVALID, "brick values exceed 4 GiB");
  Renderer* r = h;
  hipError_t e;
  if ((e = hipSetDevice(r->device)) != hipSuccess) return hip_err(e, "hipSetDevice");
  r->header = *header;
  r->header.aux_data = layout;
  r->quant_voxel_diag = header->aux_data & 1u;
  r->geom_type = (layout == LRT_SBS_LAYOUT_DX_RGB8 || indexed) ? LRT_TYPE_SDF_SBS_COL
                                                               : (layout == LRT_SBS_LAYOUT_DX_UV16 ? LRT_TYPE_SDF_SBS_TEX : LRT_TYPE_SDF_SBS);
  r->nodes.assign(nodes, nodes + n_nodes);
  r->values.assign(values, values + n_values);
  if (indexed) r->values_f.assign(values_f, values_f + n_values_f);
  else r->values_f.clear();
  std::vector<f3> bmin, bmax;
  sbs_brick_boxes(nodes, n_nodes, bmin, bmax);
  BuildTree tree = build_sah(bmin, bmax, 1);
  r->pairs = to_fat_pairs(tree, bmin, bmax, &r->pair_depth);
  if ((e = r->d_pairs.upload(r->pairs.data(), r->pairs.size() * sizeof(BVHNodePair), r->stream)) != hipSuccess) return hip_err(e, "upload bvh");
  const std::vector<float> cb = cull_boxes(r->pairs, &r->n_cull);
  if ((e = r->d_cull_boxes.upload(cb.data(), cb.size() * sizeof(float), r->stream)) != hipSuccess) return hip_err(e, "upload cull boxes");
  if ((e = r->d_nodes.upload(r->nodes.data(), r->nodes.size() * sizeof(SdfSBSNode), r->stream)) != hipSuccess) return hip_err(e, "upload nodes");
  std::vector<float> recs = brick_records(r->nodes, r->header, r->quant_voxel_diag);
  if ((e = r->d_brecs.upload(recs.data(), recs.size() * sizeof(float), r->stream)) != hipSuccess) return hip_err(e, "upload bricks");
  if ((e = r->d_values.upload(r->values.data(), r->values.size() * 4, r->stream)) != hipSuccess) return hip_err(e, "upload values");
  if ((e = r->d_values_f.upload(r->values_f.data(), r->values_f.size() * 4, r->stream)) != hipSuccess) return hip_err(e, "upload values_f");
  if ((e = hipStreamSynchronize(r->stream)) != hipSuccess) return hip_err(e, "scene upload");
  r->has_scene = true;
  r->bfeat_valid = false;
  r->dr_tree_state = 0;
  return LRT_OK;
}

int lrt_get_bvh(lrt_handle h, BVHNodePair* out, uint32_t capacity, uint32_t* count, uint32_t* depth) {
  int c = check_ready(h);
  if (c) return c;
  if (count) *count = (uint32_t)h->pairs.size();
  if (depth) *depth = (uint32_t)h->pair_depth;
  if (out) {
    if (capacity < h->pairs.size()) return set_err(LRT_E_INVALID, "capacity too small");
    std::memcpy(out, h->pairs.data(), h->pairs.size() * sizeof(BVHNodePair));
  }
  return LRT_OK;
}

int lrt_build_bvh(const SdfSBSNode* nodes, uint32_t n_nodes, BVHNodePair* out, uint32_t capacity, uint32_t* count, uint32_t* depth) {
  if (!nodes || n_nodes == 0) return set_err(LRT_E_INVALID, "empty brick list");
  std::vector<f3> bmin, bmax;
  sbs_brick_boxes(nodes, n_nodes, bmin, bmax);
  BuildTree tree = build_sah(bmin, bmax, 1);
  int d = 0;
  std::vector<BVHNodePair> pairs = to_fat_pairs(tree, bmin, bmax, &d);
  if (count) *count = (uint32_t)pairs.size();
  if (depth) *depth = (uint32_t)d;
  if (out) {
    if (capacity < pairs.size()) return set_err(LRT_E_INVALID, "capacity too small");
    std::memcpy(out, pairs.data(), pairs.size() * sizeof(BVHNodePair));
  }
  return LRT_OK;
}

int lrt_set_preset(lrt_handle h, const MultiRenderPreset* p) {
  if (!h || !p) return set_err(LRT_E_INVALID, "null argument");
  if (p->spp == 0) return set_err(LRT_E_INVALID, "spp must be >= 1");
  if (p->sdf_node_intersect > LRT_SDF_NODE_INTERSECT_IT) return set_err(LRT_E_INVALID, "bad sdf_node_intersect");
  h->preset = *p;
  return LRT_OK;
}
int lrt_get_preset(lrt_handle h, MultiRenderPreset* p) {
  if (!h || !p) return set_err(LRT_E_INVALID, "null argument");
  *p = h->preset;
  return LRT_OK;
}
int lrt_set_viewport(lrt_handle h, int x0, int y0, int width, int height) {
  (void)x0; (void)y0;
  if (!h || width <= 0 || height <= 0 || width > 65535 || height > 65535) return set_err(LRT_E_INVALID, "bad viewport");
  h->width = (uint32_t)width;
  h->height = (uint32_t)height;
  return LRT_OK;
}
int lrt_update_camera(lrt_handle h, const float wv[16], const float pr[16]) {
  if (!h || !wv || !pr) return set_err(LRT_E_INVALID, "null argument");
  std::memcpy(h->world_view, wv, 64);
  std::memcpy(h->proj, pr, 64);
  h->cam = make_camera(wv, pr);
  h->has_camera = true;
  return LRT_OK;
}
int lrt_set_seed(lrt_handle h, uint32_t seed) {
  if (!h) return set_err(LRT_E_INVALID, "null handle");
  h->seed = seed;
  return LRT_OK;
}

static int render_host(lrt_handle h, uint32_t W, uint32_t H, int passes, bool rgba, bool gbuf, const char* name) {
  int c = check_ready(h);
  if (c) return c;
  if (!h->has_camera) return set_err(LRT_E_INVALID, "no camera: call lrt_update_camera");
  if (passes < 1) passes = 1;
  h->width = W;
  h->height = H;
  for (int i = 0; i < passes; i++) {
    if ((c = h->render_device(&h->cam, 1, W, H, rgba, gbuf, gbuf, name, nullptr))) return c;
    if ((c = h->finish_timer())) return c;
  }
  return LRT_OK;
}

int lrt_render_float(lrt_handle h, float* image, uint32_t W, uint32_t H, int passes) {
  if (!image) return set_err(LRT_E_INVALID, "null image");
  int c = render_host(h, W, H, passes, true, false, "CastRayFloatSingleBlock");
  if (c) return c;
  hipError_t e = hipMemcpy(image, h->d_out_rgba.p, (size_t)W * H * 16, hipMemcpyDeviceToHost);
  return e == hipSuccess ? LRT_OK : hip_err(e, "download");
}

int lrt_render(lrt_handle h, uint32_t* image, uint32_t W, uint32_t H, int passes) {
  if (!image) return set_err(LRT_E_INVALID, "null image");
  int c = render_host(h, W, H, passes, true, false, "CastRaySingleBlock");
  if (c) return c;
  std::vector<float> tmp((size_t)W * H * 4);
  hipError_t e = hipMemcpy(tmp.data(), h->d_out_rgba.p, tmp.size() * 4, hipMemcpyDeviceToHost);
  if (e != hipSuccess) return hip_err(e, "download");
  for (size_t i = 0; i < (size_t)W * H; i++) {  // uint4(255 * clamp(c, 0, 1)), eye_ray.cpp:111-115
    uint32_t col[4];
    for (int k = 0; k < 4; k++) col[k] = (uint32_t)(255.0f * std::min(std::max(tmp[4 * i + k], 0.0f), 1.0f));
    image[i] = (col[3] << 24) | (col[2] << 16) | (col[1] << 8) | col[0];
  }
  return LRT_OK;
}

int lrt_render_gbuffer(lrt_handle h, float* t, float* n, uint32_t* prim, uint32_t W, uint32_t H) {
  int c = render_host(h, W, H, 1, false, true, "CastRayGBuffer");
  if (c) return c;
  return lrt_download_view(h, 0, t, n, prim, nullptr);
}

int lrt_cast_rays(lrt_handle h, const float* rays, uint32_t n, CRT_Hit* out) {
  int c = check_ready(h);
  if (c) return c;
  if (!rays || !out) return set_err(LRT_E_INVALID, "null argument");
  if (n == 0) return LRT_OK;
  hipError_t e;
  if ((e = hipSetDevice(h->device)) != hipSuccess) return hip_err(e, "hipSetDevice");
  if ((e = h->d_rays.upload(rays, (size_t)n * 32, h->stream)) != hipSuccess) return hip_err(e, "upload rays");
  if ((e = h->d_hits.ensure((size_t)n * sizeof(CRT_Hit))) != hipSuccess) return hip_err(e, "alloc hits");
  uint32_t blocks = (n + 255) / 256;
  if ((e = h->ensure_spill(blocks)) != hipSuccess) return hip_err(e, "alloc spill");
  e = launch_cast_rays(h->dev_scene(), h->trace_params(false), (const float4*)h->d_rays.p, n, (CRT_Hit*)h->d_hits.p,
                       (uint32_t*)h->d_spill.p, blocks, h->stream);
  if (e != hipSuccess) return hip_err(e, "cast launch");
  if ((e = hipMemcpyAsync(out, h->d_hits.p, (size_t)n * sizeof(CRT_Hit), hipMemcpyDeviceToHost, h->stream)) != hipSuccess)
    return hip_err(e, "download hits");
  if ((e = hipStreamSynchronize(h->stream)) != hipSuccess) return hip_err(e, "cast rays");
  return LRT_OK;
}

int lrt_get_execution_time(lrt_handle h, const char* name, float out[4]) {
  if (!h || !name || !out) return set_err(LRT_E_INVALID, "null argument");
  auto it = h->times.find(name);
  if (it == h->times.end()) return LRT_OK;  // reference leaves out untouched when unknown
  out[0] = it->second;
  return LRT_OK;
}

int lrt_render_views_device(lrt_handle h, const float* wvs, const float* prs, uint32_t n_views, uint32_t W, uint32_t H, int out_mode,
                            float* kernel_ms) {
  int c = check_ready(h);
  if (c) return c;
  if (!wvs || !prs || n_views == 0) return set_err(LRT_E_INVALID, "bad cameras");
  std::vector<lrtd::Camera> cams(n_views);
  for (uint32_t v = 0; v < n_views; v++) cams[v] = make_camera(wvs + 16 * v, prs + 16 * v);
  bool rgba = out_mode == 1;
  if ((c = h->render_device(cams.data(), n_views, W, H, rgba, !rgba, !rgba, "RenderViews", nullptr))) return c;
  if (kernel_ms) {
    if ((c = h->finish_timer())) return c;
    *kernel_ms = h->last_ms;
  }
  return LRT_OK;
}

int lrt_count_work_ex(lrt_handle h, const float* wvs, const float* prs, uint32_t n_views, uint32_t W, uint32_t H, uint64_t* out,
                      uint32_t n_out) {
  int c = check_ready(h);
  if (c) return c;
  if (!wvs || !prs || !out || n_views == 0 || n_out == 0 || n_out > 7) return set_err(LRT_E_INVALID, "bad arguments");
  std::vector<lrtd::Camera> cams(n_views);
  for (uint32_t v = 0; v < n_views; v++) cams[v] = make_camera(wvs + 16 * v, prs + 16 * v);
  hipError_t e;
  if ((e = h->d_counts.ensure(64)) != hipSuccess) return hip_err(e, "alloc counts");
  if ((e = zero_u32((uint32_t*)h->d_counts.p, 16, h->stream)) != hipSuccess) return hip_err(e, "clear counts");
  if ((c = h->render_device(cams.data(), n_views, W, H, false, true, true, "CountWork", (unsigned long long*)h->d_counts.p))) return c;
  if ((c = h->finish_timer())) return c;
  unsigned long long tmp[8];
  if ((e = hipMemcpy(tmp, h->d_counts.p, sizeof(tmp), hipMemcpyDeviceToHost)) != hipSuccess) return hip_err(e, "download counts");
  for (uint32_t i = 0; i < n_out; i++) out[i] = tmp[i];
  return LRT_OK;
}

int lrt_count_work(lrt_handle h, const float* wvs, const float* prs, uint32_t n_views, uint32_t W, uint32_t H, uint64_t out[3]) {
  return lrt_count_work_ex(h, wvs, prs, n_views, W, H, out, 3);
}

int lrt_download_view(lrt_handle h, uint32_t view, float* t, float* n, uint32_t* prim, float* rgba) {
  if (!h) return set_err(LRT_E_INVALID, "null handle");
  if (view >= h->last_views) return set_err(LRT_E_INVALID, "view index out of range");
  hipError_t e;
  if ((e = hipStreamSynchronize(h->stream)) != hipSuccess) return hip_err(e, "render");
  const size_t np = (size_t)h->last_w * h->last_h, off = np * view;
  if (t && h->d_out_t.p && (e = hipMemcpy(t, (float*)h->d_out_t.p + off, np * 4, hipMemcpyDeviceToHost)) != hipSuccess) return hip_err(e, "t");
  if (n && h->d_out_n.p && (e = hipMemcpy(n, (float*)h->d_out_n.p + 2 * off, np * 8, hipMemcpyDeviceToHost)) != hipSuccess) return hip_err(e, "n");
  if (prim && h->d_out_prim.p && (e = hipMemcpy(prim, (uint32_t*)h->d_out_prim.p + off, np * 4, hipMemcpyDeviceToHost)) != hipSuccess)
    return hip_err(e, "prim");
  if (rgba && h->d_out_rgba.p && (e = hipMemcpy(rgba, (float*)h->d_out_rgba.p + 4 * off, np * 16, hipMemcpyDeviceToHost)) != hipSuccess)
    return hip_err(e, "rgba");
  return LRT_OK;
}

int lrt_render_batch(lrt_handle owner, uint32_t n_items, const lrt_handle* scenes, const float* wvs, const float* prs,
                     const uint32_t* widths, const uint32_t* heights, int autotune, float* kernel_ms) {
  if (!owner || !scenes || !wvs || !prs || !widths || !heights || n_items == 0) return set_err(LRT_E_INVALID, "bad arguments");
  if (n_items > (1u << 20)) return set_err(LRT_E_INVALID, "too many items");
  Renderer* r = owner;
  if (r->preset.interpolation_mode != LRT_INTERPOLATION_MODE_TRILINEAR)
    return set_err(LRT_E_UNSUPPORTED, "only trilinear interpolation is supported");
  std::vector<lrtd::BatchItem> items(n_items);
  std::vector<uint64_t> offsets(n_items + 1, 0);
  r->b_pix_off.assign(n_items + 1, 0);
  r->b_w.assign(widths, widths + n_items);
  r->b_h.assign(heights, heights + n_items);
  bool indexed = false;
  int max_depth = 0;
  // the owner's seeding switch and network apply to every item; each item's
  // per-brick features come from its own renderer
  const bool seeded = r->seeding && r->mlp_ready;
  hipError_t e;
  for (uint32_t i = 0; i < n_items; i++) {
    Renderer* s = scenes[i];
    if (!s || !s->has_scene) return set_err(LRT_E_NOSCENE, "batch item without a scene");
    if (s->device != r->device) return set_err(LRT_E_INVALID, "batch items must live on the owner's device");
    const uint32_t W = widths[i], H = heights[i];
    if (W == 0 || H == 0 || W > 65535 || H > 65535) return set_err(LRT_E_INVALID, "bad image size");
    if (!packed_fits(W, H)) return set_err(LRT_E_INVALID, "image too large: the 8-aligned pixel count must fit 32 bits");
    lrtd::BatchItem& it = items[i];
    it.S = s->dev_scene();
    it.cam = make_camera(wvs + 16 * (size_t)i, prs + 16 * (size_t)i);
    lrtd::set_image(it.A, W, H);
    it.A.n_views = 1;
    it.A.spp = 1;
    it.A.ray_gen_mode = r->preset.ray_gen_mode;
    it.A.seed = r->seed;
    it.A.lights = r->light_set();
    it.bfeat = nullptr;
    if (seeded) {
      if ((e = hipSetDevice(s->device)) != hipSuccess) return hip_err(e, "hipSetDevice");
      if ((e = s->ensure_brick_features()) != hipSuccess) return hip_err(e, "brick features");
      if ((e = hipStreamSynchronize(s->stream)) != hipSuccess) return hip_err(e, "brick features");
      it.bfeat = (const float4*)s->d_bfeat.p;
    }
    offsets[i + 1] = offsets[i] + it.A.n_packed;  // multiples of 64
    r->b_pix_off[i + 1] = r->b_pix_off[i] + (uint64_t)W * H;
    indexed = indexed || it.S.layout == LRT_SBS_LAYOUT_ID32F_IRGB32F || it.S.layout == LRT_SBS_LAYOUT_ID32F_IRGB32F_IN;
    max_depth = std::max(max_depth, s->pair_depth);
  }
  const uint64_t total = offsets[n_items], npix = r->b_pix_off[n_items];
  if ((e = hipSetDevice(r->device)) != hipSuccess) return hip_err(e, "hipSetDevice");
  if ((e = r->d_b_t.ensure(npix * 4)) != hipSuccess) return hip_err(e, "alloc t");
  if ((e = r->d_b_n.ensure(npix * 8)) != hipSuccess) return hip_err(e, "alloc n");
  if ((e = r->d_b_prim.ensure(npix * 4)) != hipSuccess) return hip_err(e, "alloc prim");
  for (uint32_t i = 0; i < n_items; i++) {
    items[i].O.t = (float*)r->d_b_t.p + r->b_pix_off[i];
    items[i].O.n = (float2*)r->d_b_n.p + r->b_pix_off[i];
    items[i].O.prim = (uint32_t*)r->d_b_prim.p + r->b_pix_off[i];
    items[i].O.rgba = nullptr;
  }
  if ((e = r->d_b_items.upload(items.data(), items.size() * sizeof(lrtd::BatchItem), r->stream)) != hipSuccess) return hip_err(e, "upload items");
  if ((e = r->d_b_offsets.upload(offsets.data(), offsets.size() * 8, r->stream)) != hipSuccess) return hip_err(e, "upload offsets");
  if ((e = r->d_counter.ensure(lrtd::kXcds * lrtd::kCounterStride * 4)) != hipSuccess) return hip_err(e, "alloc counter");
  const lrtd::TraceParams P = r->trace_params(true);
  const float logit = std::log(r->seed_threshold / (1.0f - r->seed_threshold));
  auto run = [&](uint32_t chunk, uint32_t blocks, float* ms) -> int {
    const int saved_depth = r->pair_depth;
    r->pair_depth = max_depth;  // the spill area must cover the deepest item tree
    hipError_t e2 = r->ensure_spill(blocks, seeded ? lrtd::kStackLdsSeeded : lrtd::kStackLds);
    r->pair_depth = saved_depth;
    if (e2 != hipSuccess) return hip_err(e2, "alloc spill");
    if ((e2 = hipEventRecord(r->ev0, r->stream)) != hipSuccess) return hip_err(e2, "event");
    const lrtd::BatchItem* d_items = (const lrtd::BatchItem*)r->d_b_items.p;
    const uint64_t* d_offsets = (const uint64_t*)r->d_b_offsets.p;
    e2 = seeded ? launch_render_batch_seeded(d_items, d_offsets, n_items, total, P, indexed, r->d_mlp_packed.p, logit, chunk,
                                             (uint32_t*)r->d_counter.p, (uint32_t*)r->d_spill.p, blocks, nullptr, r->stream)
                : launch_render_batch(d_items, d_offsets, n_items, total, P, indexed, chunk, (uint32_t*)r->d_counter.p,
                                      (uint32_t*)r->d_spill.p, blocks, nullptr, r->stream);
    if (e2 != hipSuccess) return hip_err(e2, "batch launch");
    if ((e2 = hipEventRecord(r->ev1, r->stream)) != hipSuccess) return hip_err(e2, "event");
    if ((e2 = hipEventSynchronize(r->ev1)) != hipSuccess) return hip_err(e2, "batch render");
    float t = 0.f;
    (void)hipEventElapsedTime(&t, r->ev0, r->ev1);
    if (ms) *ms = t;
    return LRT_OK;
  };
  const uint32_t cu = (uint32_t)std::max(1, r->cu_count);
  int c;
  if (autotune) {
    // the tuner: every (chunk, grid) pair of a small candidate set, fastest wins
    const uint32_t chunks[3] = {256, 512, 1024}, grids[3] = {cu * 4, cu * 8, cu * 16};
    float best = 1e30f;
    if ((c = run(256, cu * 8, nullptr))) return c;  // warm-up
    for (uint32_t ch : chunks)
      for (uint32_t g : grids) {
        float ms = 0.f;
        if ((c = run(ch, g, &ms))) return c;
        if (ms < best) { best = ms; r->b_chunk = ch; r->b_blocks = g; }
      }
  }
  const uint32_t blocks = r->b_blocks ? r->b_blocks : cu * 8;
  float ms = 0.f;
  if ((c = run(r->b_chunk, blocks, &ms))) return c;
  r->times["RenderBatch"] = ms;
  if (kernel_ms) *kernel_ms = ms;
  return LRT_OK;
}

int lrt_batch_download(lrt_handle owner, uint32_t item, float* t, float* normal_oct, uint32_t* prim) {
  if (!owner) return set_err(LRT_E_INVALID, "null handle");
  Renderer* r = owner;
  if (r->b_w.empty() || item >= r->b_w.size()) return set_err(LRT_E_INVALID, "no such batch item");
  const uint64_t off = r->b_pix_off[item], n = (uint64_t)r->b_w[item] * r->b_h[item];
  hipError_t e;
  if ((e = hipSetDevice(r->device)) != hipSuccess) return hip_err(e, "hipSetDevice");
  if (t && (e = hipMemcpy(t, (float*)r->d_b_t.p + off, n * 4, hipMemcpyDeviceToHost)) != hipSuccess) return hip_err(e, "download");
  if (normal_oct && (e = hipMemcpy(normal_oct, (float2*)r->d_b_n.p + off, n * 8, hipMemcpyDeviceToHost)) != hipSuccess) return hip_err(e, "download");
  if (prim && (e = hipMemcpy(prim, (uint32_t*)r->d_b_prim.p + off, n * 4, hipMemcpyDeviceToHost)) != hipSuccess) return hip_err(e, "download");
  return LRT_OK;
}

int lrt_batch_config(lrt_handle owner, uint32_t* chunk, uint32_t* blocks) {
  if (!owner) return set_err(LRT_E_INVALID, "null handle");
  if (chunk) *chunk = owner->b_chunk;
  if (blocks) *blocks = owner->b_blocks ? owner->b_blocks : (uint32_t)std::max(1, owner->cu_count) * 8;
  return LRT_OK;
}

int lrt_sync(lrt_handle h) {
  if (!h) return set_err(LRT_E_INVALID, "null handle");
  hipError_t e = hipStreamSynchronize(h->stream);
  return e == hipSuccess ? LRT_OK : hip_err(e, "sync");
}

int lrt_set_tile_culling(lrt_handle h, int on) {
  if (!h) return set_err(LRT_E_INVALID, "null handle");
  h->tile_cull = on != 0;
  return LRT_OK;
}

int lrt_set_launch_blocks(lrt_handle h, uint32_t blocks) {
  if (!h) return set_err(LRT_E_INVALID, "null handle");
  h->launch_blocks = blocks;
  return LRT_OK;
}

// ---- synthetic scenes -----------------------------------------------------

int lrt_sbs_displaced_icosphere(uint32_t subdiv, float radius, float amplitude, uint32_t octaves, uint32_t seed, uint32_t lod,
                                uint32_t brick_size, uint32_t brick_pad, uint32_t bytes_per_value, int voxel_diag_scale, int threads,
                                lrt_sbs_handle* out) {
  if (!out || subdiv > 8) return set_err(LRT_E_INVALID, "bad arguments");
  try {
    TriMesh m = displaced_icosphere((int)subdiv, radius, amplitude, (int)octaves, seed);
    MeshSdf sdf(m);
    std::unique_ptr<lrt_sbs> s(new lrt_sbs);
    s->s = build_sbs([&](f3 p) { return sdf(p); }, lod, brick_size, brick_pad, bytes_per_value, voxel_diag_scale != 0, threads);
    if (s->s.nodes.empty()) return set_err(LRT_E_INVALID, "scene produced no bricks");
    *out = s.release();
  } catch (const std::exception& ex) {
    return set_err(LRT_E_INVALID, ex.what());
  }
  return LRT_OK;
}

int lrt_sbs_displaced_torus(float R, float r, float amplitude, uint32_t octaves, uint32_t seed, uint32_t lod, uint32_t brick_size,
                            uint32_t brick_pad, uint32_t bytes_per_value, int voxel_diag_scale, int threads, lrt_sbs_handle* out) {
  if (!out || !(R > 0.0f) || !(r > 0.0f) || R + r >= 1.0f) return set_err(LRT_E_INVALID, "bad arguments");
  try {
    auto sdf = [=](f3 p) {  // torus in the xy plane, displaced by value noise like lrt_sbs_grid's
      float qx = std::sqrt(p.x * p.x + p.y * p.y) - R;
      float d = std::sqrt(qx * qx + p.z * p.z) - r;
      return amplitude != 0.f ? d - amplitude * value_noise_fbm(p, seed, (int)octaves, 2.0f) : d;
    };
    std::unique_ptr<lrt_sbs> s(new lrt_sbs);
    s->s = build_sbs(sdf, lod, brick_size, brick_pad, bytes_per_value, voxel_diag_scale != 0, threads);
    if (s->s.nodes.empty()) return set_err(LRT_E_INVALID, "scene produced no bricks");
    *out = s.release();
  } catch (const std::exception& ex) {
    return set_err(LRT_E_INVALID, ex.what());
  }
  return LRT_OK;
}

int lrt_sbs_grid(uint32_t shape, float R, float r, float amplitude, uint32_t octaves, uint32_t seed, uint32_t brick_count,
                 uint32_t brick_size, int with_neighbors, lrt_sbs_handle* out) {
  if (!out || brick_count == 0 || brick_count > 256 || brick_size == 0 || brick_size > 16) return set_err(LRT_E_INVALID, "bad arguments");
  std::function<float(f3)> sdf;
  if (shape == 1) {
    sdf = [=](f3 p) {  // torus in the xy plane (DR_common.cpp:212-222 orientation)
      float qx = std::sqrt(p.x * p.x + p.y * p.y) - R;
      float d = std::sqrt(qx * qx + p.z * p.z) - r;
      return amplitude != 0.f ? d - amplitude * value_noise_fbm(p, seed, (int)octaves, 2.0f) : d;
    };
  } else {
    sdf = [=](f3 p) {
      float d = length(p) - R;
      return amplitude != 0.f ? d - amplitude * value_noise_fbm(p, seed, (int)octaves, 2.0f) : d;
    };
  }
  std::unique_ptr<lrt_sbs> s(new lrt_sbs);
  s->s = build_grid_sbs(brick_count, brick_size, sdf, [](f3) { return f3(1.0f, 0.44f, 0.0f); }, with_neighbors != 0);
  *out = s.release();
  return LRT_OK;
}

int lrt_sbs_info(lrt_sbs_handle s, SdfSBSHeader* header, uint32_t* n_nodes, uint64_t* n_values, uint64_t* n_values_f) {
  if (!s) return set_err(LRT_E_INVALID, "null scene");
  if (header) *header = s->s.header;
  if (n_nodes) *n_nodes = (uint32_t)s->s.nodes.size();
  if (n_values) *n_values = s->s.values.size();
  if (n_values_f) *n_values_f = s->s.values_f.size();
  return LRT_OK;
}
int lrt_sbs_copy(lrt_sbs_handle s, SdfSBSNode* nodes, uint32_t* values, float* values_f) {
  if (!s) return set_err(LRT_E_INVALID, "null scene");
  if (nodes) std::memcpy(nodes, s->s.nodes.data(), s->s.nodes.size() * sizeof(SdfSBSNode));
  if (values) std::memcpy(values, s->s.values.data(), s->s.values.size() * 4);
  if (values_f && !s->s.values_f.empty()) std::memcpy(values_f, s->s.values_f.data(), s->s.values_f.size() * 4);
  return LRT_OK;
}
int lrt_sbs_free(lrt_sbs_handle s) {
  delete s;
  return LRT_OK;
}

int lrt_sbs_load_file(const char* path, lrt_sbs_handle* out) {
  if (!path || !out) return set_err(LRT_E_INVALID, "null argument");
  FILE* f = std::fopen(path, "rb");
  if (!f) return set_err(LRT_E_INVALID, std::string("cannot open ") + path);
  std::unique_ptr<lrt_sbs> s(new lrt_sbs);
  uint32_t n = 0, cnt = 0;
  bool ok = std::fread(&s->s.header, sizeof(SdfSBSHeader), 1, f) == 1 && std::fread(&n, 4, 1, f) == 1 && n < (1u << 28);
  if (ok) {
    s->s.nodes.resize(n);
    ok = (n == 0 || std::fread(s->s.nodes.data(), sizeof(SdfSBSNode), n, f) == n) && std::fread(&cnt, 4, 1, f) == 1 && cnt < (1u << 30);
  }
  if (ok) {
    s->s.values.resize(cnt);
    ok = cnt == 0 || std::fread(s->s.values.data(), 4, cnt, f) == cnt;
  }
  uint32_t nf = 0;
  if (ok && std::fread(&nf, 4, 1, f) == 1 && nf > 0 && nf < (1u << 30)) {
    s->s.values_f.resize(nf);
    ok = std::fread(s->s.values_f.data(), 4, nf, f) == nf;
  }
  std::fclose(f);
  if (!ok) return set_err(LRT_E_INVALID, std::string("truncated or invalid SBS file ") + path);
  *out = s.release();
  return LRT_OK;
}

int lrt_sbs_save_file(const char* path, const SdfSBSHeader* header, const SdfSBSNode* nodes, uint32_t n_nodes, const uint32_t* values,
                      uint32_t n_values, const float* values_f, uint32_t n_values_f) {
  if (!path || !header || (!nodes && n_nodes) || (!values && n_values)) return set_err(LRT_E_INVALID, "null argument");
  FILE* f = std::fopen(path, "wb");
  if (!f) return set_err(LRT_E_INVALID, std::string("cannot write ") + path);
  bool ok = std::fwrite(header, sizeof(SdfSBSHeader), 1, f) == 1 && std::fwrite(&n_nodes, 4, 1, f) == 1 &&
            (n_nodes == 0 || std::fwrite(nodes, sizeof(SdfSBSNode), n_nodes, f) == n_nodes) && std::fwrite(&n_values, 4, 1, f) == 1 &&
            (n_values == 0 || std::fwrite(values, 4, n_values, f) == n_values);
  if (ok && values_f && n_values_f)
    ok = std::fwrite(&n_values_f, 4, 1, f) == 1 && std::fwrite(values_f, 4, n_values_f, f) == n_values_f;
  ok = (std::fclose(f) == 0) && ok;
  return ok ? LRT_OK : set_err(LRT_E_INVALID, std::string("write failed ") + path);
}

// MultiRenderer::LoadScene / LoadSceneHydra (Renderer/eye_ray_host.cpp:52-121)
// for the scenes this path renders: the first camera of the file (aspect from
// the viewport) and one instance, with the identity matrix, of one sdf_sbs
// geometry node (AddCustomGeom_FromFile "sdf_sbs", BVH2Common_host.cpp:142-148).
int lrt_load_scene_hydra(lrt_handle h, const char* path) {
  if (!h || !path) return set_err(LRT_E_INVALID, "null argument");
  HydraScene S;
  try {
    S = load_hydra_scene(path);
  } catch (const std::exception& ex) {
    return set_err(LRT_E_INVALID, ex.what());
  }
  if (S.instances.size() != 1)
    return set_err(LRT_E_UNSUPPORTED, "the scene must hold exactly one geometry instance (found " + std::to_string(S.instances.size()) + ")");
  const HydraInstance& I = S.instances[0];
  const HydraGeom* G = nullptr;
  for (const HydraGeom& g : S.geoms)
    if (g.id == I.geom_id) G = &g;
  if (!G && I.geom_id >= 0 && (size_t)I.geom_id < S.geoms.size() && S.geoms[I.geom_id].id < 0) G = &S.geoms[I.geom_id];
  if (!G) return set_err(LRT_E_INVALID, "the instance names a geometry node the scene does not have");
  if (G->type != "sdf_sbs") return set_err(LRT_E_UNSUPPORTED, "geometry node type " + G->type + " (only sdf_sbs is on this path)");
  for (int k = 0; k < 16; k++)
    if (I.matrix[k] != ((k % 5) == 0 ? 1.0f : 0.0f)) return set_err(LRT_E_UNSUPPORTED, "instance transforms other than the identity");
  lrt_sbs_handle sb = nullptr;
  int c = lrt_sbs_load_file(G->path.c_str(), &sb);
  if (c) return c;
  std::unique_ptr<lrt_sbs> keep(sb);
  const SbsScene& sc = sb->s;
  c = lrt_set_scene_sbs(h, &sc.header, sc.nodes.data(), (uint32_t)sc.nodes.size(), sc.values.data(), sc.values.size(),
                        sc.values_f.empty() ? nullptr : sc.values_f.data(), sc.values_f.size());
  if (c) return c;
  if (S.has_camera) {
    const HydraCamera& C = S.camera;
    const float aspect = (h->width && h->height) ? (float)h->width / (float)h->height : 1.0f;
    m4 v = look_at(f3(C.pos[0], C.pos[1], C.pos[2]), f3(C.look_at[0], C.look_at[1], C.look_at[2]), f3(C.up[0], C.up[1], C.up[2]));
    m4 p = perspective(C.fov, aspect, C.znear, C.zfar);
    if ((c = lrt_update_camera(h, v.m, p.m))) return c;
  }
  return LRT_OK;
}

int lrt_camera_lookat(const float eye[3], const float target[3], const float up[3], float fov_deg, float aspect, float znear, float zfar,
                      float world_view[16], float proj[16]) {
  if (!eye || !target || !up || !world_view || !proj) return set_err(LRT_E_INVALID, "null argument");
  m4 v = look_at(f3(eye[0], eye[1], eye[2]), f3(target[0], target[1], target[2]), f3(up[0], up[1], up[2]));
  m4 p = perspective(fov_deg, aspect, znear, zfar);
  std::memcpy(world_view, v.m, 64);
  std::memcpy(proj, p.m, 64);
  return LRT_OK;
}

int lrt_orbit_camera(uint32_t seed, uint32_t index, float distance, float fov_deg, float aspect, float znear, float zfar,
                     float world_view[16], float proj[16]) {
  if (!world_view || !proj) return set_err(LRT_E_INVALID, "null output");
  uint64_t z = ((uint64_t)seed << 32) ^ (0x9E3779B97F4A7C15ull * (index + 1));
  auto next = [&]() {  // splitmix64
    z += 0x9E3779B97F4A7C15ull;
    uint64_t x = z;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return (double)((x ^ (x >> 31)) >> 11) * (1.0 / 9007199254740992.0);
  };
  double az = 2.0 * 3.14159265358979323846 * next();
  double el = (next() - 0.5) * (3.14159265358979323846 / 2.0);  // elevation in [-45, 45] degrees
  f3 eye((float)(distance * std::cos(el) * std::sin(az)), (float)(distance * std::sin(el)), (float)(distance * std::cos(el) * std::cos(az)));
  m4 v = look_at(eye, f3(0, 0, 0), f3(0, 1, 0));
  m4 p = perspective(fov_deg, aspect, znear, zfar);
  std::memcpy(world_view, v.m, 64);
  std::memcpy(proj, p.m, 64);
  return LRT_OK;
}

}  // extern "C"

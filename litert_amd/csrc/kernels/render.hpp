// render.hpp - launch interface of the render kernels (host + device).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "trace_device.hpp"
#include "fastdiv.hpp"

namespace lrtd {

// LDS traversal-stack entries per lane (x 256 lanes x 4 B = 24 KiB per block);
// deeper entries spill to global memory.
constexpr uint32_t kStackLds = 24;

// Chunk counters of the render kernel: one per XCD, each on its own 128-B
// line (kCounterStride uint32 apart); the work-counter buffer holds
// kXcds * kCounterStride words.
constexpr uint32_t kXcds = 8, kCounterStride = 32;

struct RenderArgs {
  uint32_t W, H;
  uint32_t packed_w, n_packed;  // packed dims are rounded up to multiples of 8
  uint32_t n_views;
  uint32_t spp, ray_gen_mode, seed;
  LightSet lights;  // shading modes LAMBERT / PHONG
  // 8x8 super-blocks per view (n_packed / 64) and per row (packed_w / 8), as
  // divisors of a wave's super-block index (set_image)
  FastDiv sb_view, sb_row;
  // Tile culling (null: every tile is traced): bit s set when super-block s
  // of the launch (view * sb_view.d + its row-major index) provably reaches
  // no brick (k_tile_cull) and is written as misses without tracing.
  const uint32_t* tile_mask = nullptr;
};

// Image size of a launch: W x H pixels, packed to multiples of 8.
inline void set_image(RenderArgs& A, uint32_t W, uint32_t H) {
  A.W = W;
  A.H = H;
  A.packed_w = 8 * ((W + 7) / 8);
  A.n_packed = A.packed_w * (8 * ((H + 7) / 8));
  A.sb_view = make_fastdiv(A.n_packed / 64);
  A.sb_row = make_fastdiv(A.packed_w / 8);
}

struct RenderOut {
  float* t;        // [view][y][x] hit distance (1e9 on miss)
  float2* n;       // octahedral-encoded normal (coords[2], coords[3])
  uint32_t* prim;  // primId (~0 on miss)
  float4* rgba;    // RenderFloat colour
};

// One scene + camera + image of a batched launch (many scenes,
// uneven ray counts, one persistent launch). Pixel ranges are concatenated in
// the packed order; each item's range is a multiple of 64 (8x8 tiles).
struct BatchItem {
  DevScene S;
  Camera cam;
  RenderArgs A;  // n_views = 1, spp = 1
  RenderOut O;
  const float4* bfeat;  // per-brick seeder features (seeded launch only)
};

}  // namespace lrtd

namespace lrt {
hipError_t zero_u32(uint32_t* d, uint32_t n, hipStream_t stream);
// Fills A.tile_mask (tile_mask_words(A) words, one bit per 8x8 super-block:
// set = every ray of the tile provably misses every brick) from the cull
// boxes described at launch_render.
hipError_t launch_tile_cull(const lrtd::Camera* d_cams, const lrtd::RenderArgs& A, const float4* d_cull_boxes, uint32_t n_cull,
                            hipStream_t stream);
// A.tile_mask non-null: k_tile_cull first fills it (tile_mask_words(A)
// words) from the cull boxes at d_cull_boxes, three levels of BVH node boxes
// {min, max}: n_cull (<= 8) coarse boxes, then the middle boxes (at most 64),
// then the fine boxes (at most 512); a coarse or middle box holds the
// [first, end) range of the boxes under it on the next level in its w
// components (uint32 bits).
hipError_t launch_render(const lrtd::DevScene& S, const lrtd::TraceParams& P, const lrtd::Camera* d_cams, const lrtd::RenderArgs& A,
                         const lrtd::RenderOut& O, uint32_t* d_work_counter, uint32_t* d_spill, uint32_t blocks, bool rgba,
                         unsigned long long* d_counts, hipStream_t stream, const float4* d_cull_boxes = nullptr, uint32_t n_cull = 0);
// 32-bit words of a launch's tile mask (one bit per 8x8 super-block, whole
// 64-bit words)
inline size_t tile_mask_words(const lrtd::RenderArgs& A) { return 2 * (((size_t)A.n_views * A.sb_view.d + 63) / 64); }
// offsets: n_items + 1 packed-pixel offsets (offsets[n_items] = total);
// chunk: pixels per chunk-counter atomic (a multiple of 256).
hipError_t launch_render_batch(const lrtd::BatchItem* d_items, const uint64_t* d_offsets, uint32_t n_items, uint64_t total,
                               const lrtd::TraceParams& P, bool indexed, uint32_t chunk, uint32_t* d_work_counter, uint32_t* d_spill,
                               uint32_t blocks, unsigned long long* d_counts, hipStream_t stream);
hipError_t launch_cast_rays(const lrtd::DevScene& S, const lrtd::TraceParams& P, const float4* d_rays, uint32_t n, CRT_Hit* d_out,
                            uint32_t* d_spill, uint32_t blocks, hipStream_t stream);
}  // namespace lrt

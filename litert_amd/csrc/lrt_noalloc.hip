// lrt_noalloc.hip - the multi-view render for callers that cannot allocate
// device memory (code under HIP graph capture, tight serving loops, harnesses
// that forbid hipMalloc).
//
// It runs the same k_render launch as lrt_render_views_device, but its
// scratch -- the per-XCD chunk counters, the camera records and the per-lane
// traversal-stack spill slices -- lives in one file-scope __device__ array
// of kScratchBytes per GPU instead of buffers the handle grows, and the
// caller passes the output buffers. Each GPU holds the array until the
// process ends. Every call on a GPU shares that GPU's array: calls that
// overlap on one GPU (two streams, two handles, two threads) race on it.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <mutex>

#include "../../include/lrt.h"
#include "kernels/render.hpp"
#include "lrt_internal.hpp"

namespace lrt {
namespace {

constexpr size_t kScratchBytes = 32u << 20;  // 32 MiB per GPU
constexpr uint32_t kMaxViews = LRT_NOALLOC_MAX_VIEWS;
constexpr size_t kCounterOffset = 0;  // kXcds * kCounterStride uint32 chunk counters
constexpr size_t kCameraOffset = 4096;
constexpr size_t kSpillOffset = kCameraOffset + ((kMaxViews * sizeof(lrtd::Camera) + 255) / 256) * 256;
static_assert(lrtd::kXcds * lrtd::kCounterStride * 4 <= kCameraOffset, "counters overlap the camera records");
constexpr int kMaxDevices = 64;

struct CameraArgs {
  lrtd::Camera cams[kMaxViews];
  uint32_t n;
};

}  // namespace
}  // namespace lrt

// In host code `scratch` names a host-side stand-in, not device memory: the
// host only ever uses the address hipGetSymbolAddress returns for it.
static __device__ __align__(256) uint8_t scratch[lrt::kScratchBytes];

namespace lrt {
namespace {

std::mutex g_scratch_mutex;
void* g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on `device` (which must be current), looked
// up once per device.
int scratch_address(int device, uint8_t** out) {
  if (device < 0 || device >= kMaxDevices) return set_err(LRT_E_INVALID, "device index out of range for the no-alloc scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_err(e, "hipGetSymbolAddress(scratch)");
    if (!p) return set_err(LRT_E_HIP, "hipGetSymbolAddress(scratch) returned a null address");
    g_scratch_ptr[device] = p;
  }
  *out = (uint8_t*)g_scratch_ptr[device];
  return LRT_OK;
}

// Writes the camera records into the scratch from the kernel arguments (no
// host-to-device copy): thread i writes camera i.
__global__ void k_put_cameras(CameraArgs a, lrtd::Camera* dst) {
  const uint32_t i = threadIdx.x;
  if (i < a.n) dst[i] = a.cams[i];
}

}  // namespace
}  // namespace lrt

using namespace lrt;

extern "C" int lrt_render_views_noalloc(lrt_handle h, const float* world_views, const float* projs, uint32_t n_views, uint32_t width,
                                        uint32_t height, float* d_t, float* d_normal_oct, uint32_t* d_prim, float* d_rgba) {
  if (!h) return set_err(LRT_E_INVALID, "null handle");
  if (!h->has_scene) return set_err(LRT_E_NOSCENE, "no scene: call lrt_set_scene_sbs first");
  if (!world_views || !projs) return set_err(LRT_E_INVALID, "bad cameras");
  if (n_views == 0 || n_views > kMaxViews)
    return set_err(LRT_E_INVALID, "lrt_render_views_noalloc renders 1.." + std::to_string(kMaxViews) + " views per call");
  if (width == 0 || height == 0 || width > 65535 || height > 65535) return set_err(LRT_E_INVALID, "bad image size");
  if ((uint64_t)(8 * ((width + 7) / 8)) * (uint64_t)(8 * ((height + 7) / 8)) > 0xFFFFFFFFull)
    return set_err(LRT_E_INVALID, "image too large: the 8-aligned pixel count must fit 32 bits");
  if (!d_t && !d_normal_oct && !d_prim && !d_rgba) return set_err(LRT_E_INVALID, "no output buffer");
  if (h->preset.spp == 0) return set_err(LRT_E_INVALID, "spp must be >= 1");
  if (h->preset.interpolation_mode != LRT_INTERPOLATION_MODE_TRILINEAR)
    return set_err(LRT_E_UNSUPPORTED, "only trilinear interpolation is supported");
  if (h->seeding && h->mlp_ready)
    return set_err(LRT_E_UNSUPPORTED, "lrt_render_views_noalloc renders without the learned seeder: switch seeding off");
  // Per-lane spill slices of the traversal stack: entries beyond the LDS part,
  // one 256-lane slice per block. A deep tree gets fewer (persistent) blocks.
  const uint32_t extra = (uint32_t)std::max(1, h->pair_depth + 2 - (int)lrtd::kStackLds);
  const size_t per_block = (size_t)256 * extra * sizeof(uint32_t);
  const size_t fit = (kScratchBytes - kSpillOffset) / per_block;
  if (fit == 0) return set_err(LRT_E_UNSUPPORTED, "BVH too deep for the no-alloc scratch");
  const uint32_t blocks = (uint32_t)std::min<size_t>(h->grid_blocks(), fit);

  hipError_t e;
  if ((e = hipSetDevice(h->device)) != hipSuccess) return hip_err(e, "hipSetDevice");
  uint8_t* base = nullptr;
  int c = scratch_address(h->device, &base);
  if (c) return c;

  CameraArgs ca{};
  ca.n = n_views;
  for (uint32_t v = 0; v < n_views; v++) ca.cams[v] = make_camera(world_views + 16 * v, projs + 16 * v);
  lrtd::Camera* d_cams = (lrtd::Camera*)(base + kCameraOffset);
  hipLaunchKernelGGL(k_put_cameras, dim3(1), dim3(64), 0, h->stream, ca, d_cams);
  if ((e = hipGetLastError()) != hipSuccess) return hip_err(e, "camera records");

  const bool rgba = d_rgba != nullptr;
  lrtd::RenderOut O{};
  O.t = d_t;
  O.n = (float2*)d_normal_oct;
  O.prim = d_prim;
  O.rgba = (float4*)d_rgba;
  lrtd::RenderArgs A;
  lrtd::set_image(A, width, height);
  A.n_views = n_views;
  A.spp = h->preset.spp;
  A.ray_gen_mode = h->preset.ray_gen_mode;
  A.seed = h->seed;
  A.lights = h->light_set();
  e = launch_render(h->dev_scene(), h->trace_params(!rgba), d_cams, A, O, (uint32_t*)(base + kCounterOffset), (uint32_t*)(base + kSpillOffset),
                    blocks, rgba, nullptr, h->stream);
  if (e != hipSuccess) return hip_err(e, "render launch");
  return LRT_OK;
}

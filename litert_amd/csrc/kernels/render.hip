// render.hip - primary-ray render kernels (MultiRenderer::CastRaySingleBlock /
// CastRayFloatSingleBlock, Renderer/eye_ray_host.cpp:246-282) and batched
// RayQuery_NearestHit for gfx950.
//
// One lane per packed pixel: the packed index space groups each 8x8 pixel
// super-block (eye_ray.cpp:537-555) into 64 consecutive indices, i.e. one
// wavefront per screen tile, so neighbouring lanes walk the same BVH nodes
// and bricks. Blocks are persistent and pull 256-pixel chunks from a global
// counter, so uneven per-ray work does not leave CUs idle at the tail.
#include "render.hpp"
#include "render_device.hpp"
#include "traverse_device.hpp"

namespace lrtd {

#ifndef LRT_SPEC_WAVES
#define LRT_SPEC_WAVES 6
#endif
// Occupancy target (waves per SIMD) of the render kernels specialised for one
// root finder on DX-layout scenes; the register budget is 512 / waves.
constexpr int kSpecWaves = LRT_SPEC_WAVES;

template <bool COUNT, bool RGBA, int RF, bool IDX>
__device__ __forceinline__ void render_pixel(const DevScene& S, const TraceParams& P, const Camera* __restrict__ cams, const RenderArgs& A,
                                             const RenderOut& O, uint64_t gtid, Stack& st, WorkCount& wc) {
  // A wave's 64 packed indices form one 8x8 super-block of one view (each
  // view spans a multiple of 64 packed indices and chunks start on multiples
  // of 256), so the view, the super-block and the camera are wave-uniform:
  // computed on the scalar unit and read with scalar loads into SGPRs.
  // The wave's super-block index in the launch fits 32 bits (launches hold
  // fewer than 2^38 packed pixels, checked on the host).
  const uint32_t wsb = __builtin_amdgcn_readfirstlane((uint32_t)(gtid >> 6));
  const uint32_t view = A.sb_view.div(wsb);
  const uint32_t sb = wsb - view * A.sb_view.d;
  const uint32_t bhy = A.sb_row.div(sb), bhx = sb - bhy * A.sb_row.d;
  const uint32_t r = (uint32_t)gtid & 63u, lh = r >> 4, l = r & 15u;
  const uint32_t x = ((bhx * 2 + (lh & 1u)) << 2) + (l & 3u);  // tid_to_xy with the uniform part hoisted
  const uint32_t y = ((bhy * 2 + (lh >> 1)) << 2) + (l >> 2);
  if (x >= A.W || y >= A.H) return;
  const Camera& C = cams[view];
  const size_t pix = (size_t)view * A.W * A.H + (size_t)y * A.W + x;
  float4 acc = make_float4(0, 0, 0, 0);
  for (uint32_t i = 0; i < A.spp; i++) {
    const float2 d = sample_offset(A, x, y, i);
    v3 pos, dir;
    eye_ray(C, x, y, d.x, d.y, A.W, A.H, pos, dir);
    Hit hit = RayQuery_NearestHit<COUNT, RF, IDX>(S, P, pos, 0.0f, dir, 1e9f, st, &wc);
    if (i == 0) {
      if (O.t) O.t[pix] = hit.t;
      if (O.prim) O.prim[pix] = hit.primId;
      if (O.n) O.n[pix] = make_float2(hit.coords[2], hit.coords[3]);
    }
    if (RGBA) {
      uint32_t lit = 0xFFFFFFFFu;
      if (needs_shadow(P.render_mode) && hit.primId != 0xFFFFFFFFu) {
        for (uint32_t li = 0; li < A.lights.n; li++) {
          v3 o, ld;
          if (!shadow_ray(A.lights, li, pos, dir, hit.t, o, ld)) continue;
          const Hit sh = RayQuery_NearestHit<COUNT, RF, IDX, false>(S, P, o, 0.0f, ld, 1e9f, st, &wc);
          if (sh.primId != 0xFFFFFFFFu) lit &= ~(1u << li);
        }
      }
      float4 c = shade(P.render_mode, hit, pos, dir, A.lights, lit);
      acc = make_float4(acc.x + c.x, acc.y + c.y, acc.z + c.z, acc.w + c.w);
    }
  }
  if (RGBA) {
    const float s = (float)A.spp;
    O.rgba[pix] = make_float4(acc.x / s, acc.y / s, acc.z / s, acc.w / s);
  }
}

template <bool COUNT, bool RGBA, int RF, bool IDX>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(RF >= 0 && !IDX ? kSpecWaves : 1, 8))) void k_render(DevScene S, TraceParams P, const Camera* __restrict__ cams, RenderArgs A,
                                                RenderOut O, uint32_t* work_counter, uint32_t* spill, unsigned long long* counts) {
  __shared__ uint32_t lds_stack[kStackLds * 256];
  __shared__ uint32_t chunk_base;
  Stack st;
  st.lds = (lds_u32*)(lds_stack + threadIdx.x);
  st.lds_stride = 256;
  st.lds_entries = kStackLds;
  const uint32_t gl = blockIdx.x * 256 + threadIdx.x;
  st.spill = spill + gl;
  st.spill_stride = gridDim.x * 256;
  WorkCount wc;
  const uint64_t total = (uint64_t)A.n_views * A.n_packed;
  // Work distribution, XCD-aware: workgroups are dispatched round-robin over
  // the 8 XCDs, so blockIdx.x % 8 names this block's XCD (and its L2). The
  // packed pixel range is cut into 8 contiguous slices (8 views -> one view
  // per XCD), each with its own chunk counter on its own 128-B line; a block
  // drains its XCD's slice first, then helps the others in turn.
  const uint32_t xcd = blockIdx.x & (kXcds - 1u);
  const uint64_t n_chunks = (total + 255) / 256;
  const uint64_t per = (n_chunks + kXcds - 1) / kXcds;
  for (uint32_t k = 0; k < kXcds; k++) {
    const uint32_t sl = (xcd + k) & (kXcds - 1u);
    const uint64_t c0 = per * sl, c1 = c0 + per < n_chunks ? c0 + per : n_chunks;
    for (;;) {
      if (threadIdx.x == 0) chunk_base = atomicAdd(work_counter + kCounterStride * sl, 1u);
      __syncthreads();
      const uint64_t c = c0 + chunk_base;
      __syncthreads();
      if (c >= c1) break;  // uniform across the block
      const uint64_t gtid = c * 256 + threadIdx.x;
      if (gtid < total) render_pixel<COUNT, RGBA, RF, IDX>(S, P, cams, A, O, gtid, st, wc);
    }
  }
  if (COUNT) {
    atomicAdd(&counts[0], (unsigned long long)wc.pairs);
    atomicAdd(&counts[1], (unsigned long long)wc.bricks);
    atomicAdd(&counts[2], (unsigned long long)wc.voxels);
    atomicAdd(&counts[3], (unsigned long long)wc.iters);
  }
}

// Batched render: the concatenated packed pixel ranges of many
// (scene, camera, image) items are drained by one persistent launch with the
// same XCD-sliced chunk counters as k_render; each wave looks its item up
// (wave-uniform, item ranges are 64-aligned) and runs render_pixel on it.
template <bool COUNT, int RF, bool IDX>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(RF >= 0 && !IDX ? kSpecWaves : 1, 8))) void k_render_batch(
    const BatchItem* __restrict__ items, const uint64_t* __restrict__ offsets, uint32_t n_items, uint64_t total, TraceParams P,
    uint32_t chunk, uint32_t* work_counter, uint32_t* spill, unsigned long long* counts) {
  __shared__ uint32_t lds_stack[kStackLds * 256];
  __shared__ uint32_t chunk_base;
  Stack st;
  st.lds = (lds_u32*)(lds_stack + threadIdx.x);
  st.lds_stride = 256;
  st.lds_entries = kStackLds;
  st.spill = spill + blockIdx.x * 256 + threadIdx.x;
  st.spill_stride = gridDim.x * 256;
  WorkCount wc;
  const uint32_t xcd = blockIdx.x & (kXcds - 1u);
  const uint64_t n_chunks = (total + chunk - 1) / chunk;
  const uint64_t per = (n_chunks + kXcds - 1) / kXcds;
  for (uint32_t k = 0; k < kXcds; k++) {
    const uint32_t sl = (xcd + k) & (kXcds - 1u);
    const uint64_t c0 = per * sl, c1 = c0 + per < n_chunks ? c0 + per : n_chunks;
    for (;;) {
      if (threadIdx.x == 0) chunk_base = atomicAdd(work_counter + kCounterStride * sl, 1u);
      __syncthreads();
      const uint64_t c = c0 + chunk_base;
      __syncthreads();
      if (c >= c1) break;  // uniform across the block
      for (uint32_t j = 0; j < chunk; j += 256) {
        const uint64_t gtid = c * chunk + j + threadIdx.x;
        if (gtid < total) {
          // item of this wave: largest i with offsets[i] <= first active gtid
          const uint64_t g0 = ((uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)(gtid >> 32)) << 32) |
                              (uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)gtid);
          uint32_t lo = 0, hi = n_items;  // offsets[lo] <= g0 < offsets[hi]
          while (hi - lo > 1u) {
            const uint32_t mid = (lo + hi) >> 1;
            if (offsets[mid] <= g0) lo = mid;
            else hi = mid;
          }
          const BatchItem& it = items[lo];
          render_pixel<COUNT, false, RF, IDX>(it.S, P, &it.cam, it.A, it.O, gtid - offsets[lo], st, wc);
        }
      }
    }
  }
  if (COUNT) {
    atomicAdd(&counts[0], (unsigned long long)wc.pairs);
    atomicAdd(&counts[1], (unsigned long long)wc.bricks);
    atomicAdd(&counts[2], (unsigned long long)wc.voxels);
    atomicAdd(&counts[3], (unsigned long long)wc.iters);
  }
}

__global__ __launch_bounds__(256) void k_cast_rays(DevScene S, TraceParams P, const float4* __restrict__ rays, uint32_t n,
                                                   CRT_Hit* __restrict__ out, uint32_t* spill) {
  __shared__ uint32_t lds_stack[kStackLds * 256];
  Stack st;
  st.lds = (lds_u32*)(lds_stack + threadIdx.x);
  st.lds_stride = 256;
  st.lds_entries = kStackLds;
  const uint32_t gl = blockIdx.x * 256 + threadIdx.x;
  st.spill = spill + gl;
  st.spill_stride = gridDim.x * 256;
  if (gl >= n) return;
  const float4 pn = rays[2 * gl], df = rays[2 * gl + 1];
  WorkCount wc;
  Hit h = RayQuery_NearestHit<false>(S, P, v3(pn.x, pn.y, pn.z), pn.w, v3(df.x, df.y, df.z), df.w, st, &wc);
  CRT_Hit r;
  r.t = h.t; r.primId = h.primId; r.instId = h.instId; r.geomId = h.geomId;
  r.coords[0] = h.coords[0]; r.coords[1] = h.coords[1]; r.coords[2] = h.coords[2]; r.coords[3] = h.coords[3];
  out[gl] = r;
}

// Clears small counters with a kernel rather than hipMemsetAsync, so the
// launch sequence stays valid inside a captured HIP graph.
__global__ void k_zero_u32(uint32_t* p, uint32_t n) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

}  // namespace lrtd

namespace lrt {

hipError_t zero_u32(uint32_t* d, uint32_t n, hipStream_t stream) {
  hipLaunchKernelGGL(lrtd::k_zero_u32, dim3((n + 255) / 256), dim3(256), 0, stream, d, n);
  return hipGetLastError();
}

// Kernel variants: the root finder is fixed at compile time for the two
// modes the benchmarks use on DX-layout scenes (fewer live registers, more
// waves per SIMD); other modes and indexed layouts use the generic build.
template <int RF, bool IDX, typename F>
static hipError_t with_counts_rgba(bool counts, bool rgba, F&& f) {
  if (counts) return rgba ? f.template go<true, true, RF, IDX>() : f.template go<true, false, RF, IDX>();
  return rgba ? f.template go<false, true, RF, IDX>() : f.template go<false, false, RF, IDX>();
}

template <typename F>
static hipError_t dispatch_variant(const lrtd::DevScene& S, const lrtd::TraceParams& P, bool counts, bool rgba, F&& f) {
  const bool indexed = S.layout == LRT_SBS_LAYOUT_ID32F_IRGB32F || S.layout == LRT_SBS_LAYOUT_ID32F_IRGB32F_IN;
  if (indexed) return with_counts_rgba<-1, true>(counts, rgba, f);
  if (P.sdf_node_intersect == LRT_SDF_NODE_INTERSECT_NEWTON) return with_counts_rgba<(int)LRT_SDF_NODE_INTERSECT_NEWTON, false>(counts, rgba, f);
  if (P.sdf_node_intersect == LRT_SDF_NODE_INTERSECT_ST) return with_counts_rgba<(int)LRT_SDF_NODE_INTERSECT_ST, false>(counts, rgba, f);
  return with_counts_rgba<-1, false>(counts, rgba, f);
}

struct RenderLaunch {
  const lrtd::DevScene& S; const lrtd::TraceParams& P; const lrtd::Camera* cams; const lrtd::RenderArgs& A; const lrtd::RenderOut& O;
  uint32_t* counter; uint32_t* spill; uint32_t blocks; unsigned long long* counts; hipStream_t stream;
  template <bool C, bool R, int RF, bool IDX>
  hipError_t go() {
    hipLaunchKernelGGL((lrtd::k_render<C, R, RF, IDX>), dim3(blocks), dim3(256), 0, stream, S, P, cams, A, O, counter, spill, counts);
    return hipGetLastError();
  }
};

hipError_t launch_render(const lrtd::DevScene& S, const lrtd::TraceParams& P, const lrtd::Camera* d_cams, const lrtd::RenderArgs& A,
                         const lrtd::RenderOut& O, uint32_t* d_work_counter, uint32_t* d_spill, uint32_t blocks, bool rgba,
                         unsigned long long* d_counts, hipStream_t stream) {
  hipError_t e = zero_u32(d_work_counter, lrtd::kXcds * lrtd::kCounterStride, stream);
  if (e != hipSuccess) return e;
  RenderLaunch L{S, P, d_cams, A, O, d_work_counter, d_spill, blocks, d_counts, stream};
  return dispatch_variant(S, P, d_counts != nullptr, rgba, L);
}

template <bool C, int RF, bool IDX>
static hipError_t batch_go(const lrtd::BatchItem* items, const uint64_t* offsets, uint32_t n_items, uint64_t total, const lrtd::TraceParams& P,
                           uint32_t chunk, uint32_t* counter, uint32_t* spill, uint32_t blocks, unsigned long long* counts, hipStream_t stream) {
  hipLaunchKernelGGL((lrtd::k_render_batch<C, RF, IDX>), dim3(blocks), dim3(256), 0, stream, items, offsets, n_items, total, P, chunk, counter,
                     spill, counts);
  return hipGetLastError();
}

template <int RF, bool IDX>
static hipError_t batch_pick(bool counts, const lrtd::BatchItem* items, const uint64_t* offsets, uint32_t n_items, uint64_t total,
                             const lrtd::TraceParams& P, uint32_t chunk, uint32_t* counter, uint32_t* spill, uint32_t blocks,
                             unsigned long long* d_counts, hipStream_t stream) {
  return counts ? batch_go<true, RF, IDX>(items, offsets, n_items, total, P, chunk, counter, spill, blocks, d_counts, stream)
                : batch_go<false, RF, IDX>(items, offsets, n_items, total, P, chunk, counter, spill, blocks, d_counts, stream);
}

hipError_t launch_render_batch(const lrtd::BatchItem* d_items, const uint64_t* d_offsets, uint32_t n_items, uint64_t total,
                               const lrtd::TraceParams& P, bool indexed, uint32_t chunk, uint32_t* d_work_counter, uint32_t* d_spill,
                               uint32_t blocks, unsigned long long* d_counts, hipStream_t stream) {
  if (chunk == 0 || chunk % 256 != 0) return hipErrorInvalidValue;
  hipError_t e = zero_u32(d_work_counter, lrtd::kXcds * lrtd::kCounterStride, stream);
  if (e != hipSuccess) return e;
  const bool c = d_counts != nullptr;
  if (indexed) return batch_pick<-1, true>(c, d_items, d_offsets, n_items, total, P, chunk, d_work_counter, d_spill, blocks, d_counts, stream);
  if (P.sdf_node_intersect == LRT_SDF_NODE_INTERSECT_NEWTON)
    return batch_pick<(int)LRT_SDF_NODE_INTERSECT_NEWTON, false>(c, d_items, d_offsets, n_items, total, P, chunk, d_work_counter, d_spill,
                                                                  blocks, d_counts, stream);
  if (P.sdf_node_intersect == LRT_SDF_NODE_INTERSECT_ST)
    return batch_pick<(int)LRT_SDF_NODE_INTERSECT_ST, false>(c, d_items, d_offsets, n_items, total, P, chunk, d_work_counter, d_spill, blocks,
                                                              d_counts, stream);
  return batch_pick<-1, false>(c, d_items, d_offsets, n_items, total, P, chunk, d_work_counter, d_spill, blocks, d_counts, stream);
}

hipError_t launch_cast_rays(const lrtd::DevScene& S, const lrtd::TraceParams& P, const float4* d_rays, uint32_t n, CRT_Hit* d_out,
                            uint32_t* d_spill, uint32_t blocks, hipStream_t stream) {
  hipLaunchKernelGGL(lrtd::k_cast_rays, dim3(blocks), dim3(256), 0, stream, S, P, d_rays, n, d_out, d_spill);
  return hipGetLastError();
}

}  // namespace lrt

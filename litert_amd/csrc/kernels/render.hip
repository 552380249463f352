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
  const Camera& C = cams[view];
  // a tile the cull pass marked: written as tracing writes a miss
  const bool culled = A.tile_mask && ((A.tile_mask[wsb >> 5] >> (wsb & 31u)) & 1u);
  if (x >= A.W || y >= A.H) return;
  const size_t pix = (size_t)view * A.W * A.H + (size_t)y * A.W + x;
  if (COUNT && culled && r == 0) wc.culled++;
  float4 acc = make_float4(0, 0, 0, 0);
  for (uint32_t i = 0; i < A.spp; i++) {
    const float2 d = sample_offset(A, x, y, i);
    v3 pos, dir;
    Hit hit;  // a culled tile: RayQuery_NearestHit's miss (t = tFar, no primitive, coords[2..3] = 0)
    hit.t = 1e9f;
    hit.primId = 0xFFFFFFFFu;
    hit.coords[2] = hit.coords[3] = 0.0f;
    if (!culled) {
      eye_ray(C, x, y, d.x, d.y, A.W, A.H, pos, dir);
      hit = RayQuery_NearestHit<COUNT, RF, IDX>(S, P, pos, 0.0f, dir, 1e9f, st, &wc);
    }
    if (i == 0) {
      if (O.t) O.t[pix] = hit.t;
      if (O.prim) O.prim[pix] = hit.primId;
      if (O.n) O.n[pix] = make_float2(hit.coords[2], hit.coords[3]);
    }
    if (culled) break;
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
    // every sample of a culled tile shades as a miss, (0, 0, 0, 1)
    O.rgba[pix] = culled ? make_float4(0.0f, 0.0f, 0.0f, 1.0f) : make_float4(acc.x / s, acc.y / s, acc.z / s, acc.w / s);
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
    atomicAdd(&counts[6], (unsigned long long)wc.culled);
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

// ---- tile culling ------------------------------------------------------------
// The rays of a tile leave the eye e through sample positions (x + dx, y + dy)
// with x0 <= x + dx <= x0 + 8 (likewise y). eye_ray maps a sample position to
// the direction L (p.xyz / p.w), with p = projInv (X, Y, 1, 1) affine in the
// position and L the linear part of worldViewInv. Where p.w keeps one sign
// over the tile, the directions of the whole tile lie, in exact arithmetic, in
// the cone spanned by L (sign(p.w) p.xyz) at its four corners, bounded by the
// four planes through e spanned by adjacent corner directions. The corners are
// evaluated here in FP64 from the FP32 camera; the FP32 rays the render kernel
// traces differ from these exact ones by the FP32 error of eye_ray, measured
// at the tile's first corner (the kernel's own eye_ray against FP64) and
// covered, with the FP32 plane tests, by the margin 1e-3 + (|e| + sqrt(3)) *
// max(1e-5, 16 * measured error): every point of the [-1, 1]^3 scene frame
// lies within |e| + sqrt(3) of the eye. A box farther than the margin outside
// one plane holds no point of any of the tile's rays. The fine cull boxes
// (a cut through the BVH: every leaf lies under exactly one) together hold every brick, so a tile none of whose boxes reaches into its
// pyramid traces to a miss on every ray. A tile whose cone is degenerate (corners not
// in convex position, p.w changing sign) or whose eye lies far away
// (|e| > 1e3) is traced.
//
// One lane per tile (64 tiles per wave): the pyramid, then the cull boxes
// against it, read by the whole wave: the n_coarse coarse boxes, the middle
// boxes of a coarse box that reaches a lane's pyramid, and the fine boxes of
// a middle box that does; bit t of mask = tile t culled.
__global__ __launch_bounds__(256) void k_tile_cull(const Camera* __restrict__ cams, RenderArgs A, uint32_t n_tiles,
                                                   const float4* __restrict__ boxes, uint32_t n_coarse, uint32_t* __restrict__ mask) {
  const uint32_t t = blockIdx.x * 256 + threadIdx.x;
  bool ok = t < n_tiles;
  float nx[4], ny[4], nz[4], nd[4];  // inward unit normals and n . e - margin
  if (ok) {
    const uint32_t view = A.sb_view.div(t), sb = t - view * A.sb_view.d;
    const uint32_t by = A.sb_row.div(sb), bx = sb - by * A.sb_row.d;
    const Camera& C = cams[view];
    const uint32_t cx[4] = {bx * 8, bx * 8 + 8, bx * 8 + 8, bx * 8}, cy[4] = {by * 8, by * 8, by * 8 + 8, by * 8 + 8};
    const float* m = C.worldViewInv;
    double d[4][3];
    double wsign = 0.0;
    for (int i = 0; i < 4; i++) {
      const double X = 2.0 * ((double)cx[i] / (double)A.W) - 1.0, Y = -2.0 * ((double)cy[i] / (double)A.H) + 1.0;
      double p[4];
      for (int k = 0; k < 4; k++)
        p[k] = (double)C.projInv[k] * X + (double)C.projInv[4 + k] * Y + (double)C.projInv[8 + k] + (double)C.projInv[12 + k];
      if (!(p[3] != 0.0) || (i > 0 && p[3] * wsign <= 0.0)) ok = false;
      wsign = p[3];
      const double sg = p[3] > 0.0 ? 1.0 : -1.0;
      const double vx = sg * p[0], vy = sg * p[1], vz = sg * p[2];
      d[i][0] = (double)m[0] * vx + (double)m[4] * vy + (double)m[8] * vz;
      d[i][1] = (double)m[1] * vx + (double)m[5] * vy + (double)m[9] * vz;
      d[i][2] = (double)m[2] * vx + (double)m[6] * vy + (double)m[10] * vz;
    }
    // FP32 error of the traced rays, at corner 0
    double err = 1.0;
    {
      v3 pos, dir;
      eye_ray(C, cx[0], cy[0], 0.0f, 0.0f, A.W, A.H, pos, dir);
      const double len = sqrt(d[0][0] * d[0][0] + d[0][1] * d[0][1] + d[0][2] * d[0][2]);
      if (len > 0.0) {
        const double il = 1.0 / len;
        err = fmax(fabs((double)dir.x - d[0][0] * il), fmax(fabs((double)dir.y - d[0][1] * il), fabs((double)dir.z - d[0][2] * il)));
      }
    }
    const float ex = C.origin[0], ey = C.origin[1], ez = C.origin[2];
    const double elen = sqrt((double)ex * ex + (double)ey * ey + (double)ez * ez);
    if (!(elen <= 1e3) || !(err < 1e-2)) ok = false;
    const float margin = (float)(1e-3 + (elen + 1.7320508075688772) * fmax(1e-5, 16.0 * err));
    for (int i = 0; i < 4; i++) {
      const double* a = d[i];
      const double* b = d[(i + 1) & 3];
      const double n0 = a[1] * b[2] - a[2] * b[1], n1 = a[2] * b[0] - a[0] * b[2], n2 = a[0] * b[1] - a[1] * b[0];
      const double* c2 = d[(i + 2) & 3];
      const double* c3 = d[(i + 3) & 3];
      const double s2 = n0 * c2[0] + n1 * c2[1] + n2 * c2[2], s3 = n0 * c3[0] + n1 * c3[1] + n2 * c3[2];
      const double sg = (s2 > 0.0 && s3 > 0.0) ? 1.0 : ((s2 < 0.0 && s3 < 0.0) ? -1.0 : 0.0);
      // unit normal in FP32 (its direction and length to about 1e-7, far
      // inside the margin)
      const float f0 = (float)(sg * n0), f1 = (float)(sg * n1), f2 = (float)(sg * n2);
      const float l2 = f0 * f0 + f1 * f1 + f2 * f2;
      if (sg == 0.0 || !(l2 > 1e-30f)) ok = false;
      const float inv = ok ? 1.0f / sqrtf(l2) : 0.0f;
      nx[i] = f0 * inv;
      ny[i] = f1 * inv;
      nz[i] = f2 * inv;
      nd[i] = nx[i] * ex + ny[i] * ey + nz[i] * ez - margin;
    }
  }
  // A box reaches into the pyramid unless, for some plane, its corner
  // farthest along the normal lies below n . e - margin; that corner's
  // n . p is max(n, 0) . hi + min(n, 0) . lo.
  float ph[4][3], pl[4][3];
  for (int i = 0; i < 4; i++) {
    ph[i][0] = fmaxf(nx[i], 0.0f); pl[i][0] = fminf(nx[i], 0.0f);
    ph[i][1] = fmaxf(ny[i], 0.0f); pl[i][1] = fminf(ny[i], 0.0f);
    ph[i][2] = fmaxf(nz[i], 0.0f); pl[i][2] = fminf(nz[i], 0.0f);
  }
  auto reaches = [&](const float4& lo, const float4& hi) {
    bool out = false;
#pragma unroll
    for (int i = 0; i < 4; i++) {
      float f = ph[i][0] * hi.x;
      f = fmaf(ph[i][1], hi.y, f);
      f = fmaf(ph[i][2], hi.z, f);
      f = fmaf(pl[i][0], lo.x, f);
      f = fmaf(pl[i][1], lo.y, f);
      f = fmaf(pl[i][2], lo.z, f);
      out |= f < nd[i];
    }
    return !out;
  };
  // coarse boxes first; a fine box lies inside its coarse box, so only lanes
  // whose pyramid the coarse box reaches test its fine boxes
  bool touch = !ok;
  for (uint32_t c = 0; c < n_coarse; c++) {
    if (__builtin_amdgcn_ballot_w64(!touch) == 0) break;  // every lane of the wave decided
    const float4 clo = boxes[2 * c], chi = boxes[2 * c + 1];
    const bool cand = !touch && reaches(clo, chi);
    if (__builtin_amdgcn_ballot_w64(cand) == 0) continue;
    const uint32_t m0 = __float_as_uint(clo.w), m1 = __float_as_uint(chi.w);
    bool hit = false;
    for (uint32_t m = m0; m < m1; m++) {
      if (__builtin_amdgcn_ballot_w64(cand && !hit) == 0) break;
      const float4 mlo = boxes[2 * m], mhi = boxes[2 * m + 1];
      const bool mc = cand && !hit && reaches(mlo, mhi);
      if (__builtin_amdgcn_ballot_w64(mc) == 0) continue;
      const uint32_t f0 = __float_as_uint(mlo.w), f1 = __float_as_uint(mhi.w);
      for (uint32_t f = f0; f < f1; f++) {
        if (__builtin_amdgcn_ballot_w64(mc && !hit) == 0) break;
        const float4 lo = boxes[2 * f], hi = boxes[2 * f + 1];
        hit |= mc && reaches(lo, hi);
      }
    }
    touch |= hit;
  }
  const uint64_t cull = __builtin_amdgcn_ballot_w64(!touch);  // every lane of the wave reaches this
  // the wave's 64 tiles start at t0 (a multiple of 64): words t0/32 and
  // t0/32 + 1, inside the tile_mask_words(A) words when t0 < n_tiles
  const uint32_t lane = threadIdx.x & 63u, t0 = t - lane;
  if (t0 < n_tiles) {
    if (lane == 0) mask[t0 >> 5] = (uint32_t)cull;
    if (lane == 32) mask[(t0 >> 5) + 1] = (uint32_t)(cull >> 32);
  }
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

hipError_t launch_tile_cull(const lrtd::Camera* d_cams, const lrtd::RenderArgs& A, const float4* d_cull_boxes, uint32_t n_cull,
                            hipStream_t stream) {
  if (!A.tile_mask || !d_cull_boxes || n_cull == 0 || n_cull > 8) return hipErrorInvalidValue;
  const uint32_t n_tiles = A.n_views * A.sb_view.d;  // < 2^32: checked with the launch's pixel count
  hipLaunchKernelGGL(lrtd::k_tile_cull, dim3((n_tiles + 255) / 256), dim3(256), 0, stream, d_cams, A, n_tiles, d_cull_boxes, n_cull,
                     const_cast<uint32_t*>(A.tile_mask));
  return hipGetLastError();
}

hipError_t launch_render(const lrtd::DevScene& S, const lrtd::TraceParams& P, const lrtd::Camera* d_cams, const lrtd::RenderArgs& A,
                         const lrtd::RenderOut& O, uint32_t* d_work_counter, uint32_t* d_spill, uint32_t blocks, bool rgba,
                         unsigned long long* d_counts, hipStream_t stream, const float4* d_cull_boxes, uint32_t n_cull) {
  hipError_t e = zero_u32(d_work_counter, lrtd::kXcds * lrtd::kCounterStride, stream);
  if (e != hipSuccess) return e;
  if (A.tile_mask && (e = launch_tile_cull(d_cams, A, d_cull_boxes, n_cull, stream)) != hipSuccess) return e;
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

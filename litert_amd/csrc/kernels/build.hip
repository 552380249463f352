// build.hip - sparse-brick-set builder on gfx950: the SBS of a closed
// triangle mesh (sdf_converter::create_sdf_SBS(settings, header, mesh),
// utils/sdf_converter.cpp:183-228, brick fill and quantiser of
// utils/sparse_octree_builder.cpp:291-444) built on the GPU.
//
// Three kernels, each one thread per independent item:
//   k_build_refine  one candidate cell of the next level (8 per parent),
//   k_build_fill    one distance sample of one candidate brick,
//   k_build_pack    one candidate brick: the border-keep rule, quantise, pack.
// Every distance is the exact signed distance to the mesh: a nearest-triangle
// search over the host-built triangle BVH with a per-lane stack, in the same
// visiting order and with the same FP32 operations as the host builder, so
// both produce the same bricks bit for bit.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "build.hpp"
#include "trace_device.hpp"

namespace lrtd {

__device__ __forceinline__ v3 ld3(const float4* a, uint32_t i) {
  const float4 f = a[i];
  return v3(f.x, f.y, f.z);
}

// Squared distance from p to a node box (0 inside).
__device__ __forceinline__ float box_d2(const uint4* nodes, int n, v3 p) {
  const uint4 a = nodes[3 * n], b = nodes[3 * n + 1];
  const float dx = mx(0.f, mx(__uint_as_float(a.x) - p.x, p.x - __uint_as_float(b.x)));
  const float dy = mx(0.f, mx(__uint_as_float(a.y) - p.y, p.y - __uint_as_float(b.y)));
  const float dz = mx(0.f, mx(__uint_as_float(a.z) - p.z, p.z - __uint_as_float(b.z)));
  return dx * dx + dy * dy + dz * dz;
}

// Closest point of triangle abc to p (Ericson, Real-Time Collision Detection,
// 5.1.5); region: 0 face, 1..3 vertex a/b/c, 4 edge ab, 5 edge bc, 6 edge ca.
__device__ __forceinline__ v3 closest_on_triangle(v3 p, v3 a, v3 b, v3 c, int& region) {
  v3 ab = b - a, ac = c - a, ap = p - a;
  float d1 = dot(ab, ap), d2 = dot(ac, ap);
  if (d1 <= 0.f && d2 <= 0.f) { region = 1; return a; }
  v3 bp = p - b;
  float d3 = dot(ab, bp), d4 = dot(ac, bp);
  if (d3 >= 0.f && d4 <= d3) { region = 2; return b; }
  float vc = d1 * d4 - d3 * d2;
  if (vc <= 0.f && d1 >= 0.f && d3 <= 0.f) { region = 4; float v = d1 / (d1 - d3); return a + v * ab; }
  v3 cp = p - c;
  float d5 = dot(ab, cp), d6 = dot(ac, cp);
  if (d6 >= 0.f && d5 <= d6) { region = 3; return c; }
  float vb = d5 * d2 - d1 * d6;
  if (vb <= 0.f && d2 >= 0.f && d6 <= 0.f) { region = 6; float w = d2 / (d2 - d6); return a + w * ac; }
  float va = d3 * d6 - d5 * d4;
  if (va <= 0.f && (d4 - d3) >= 0.f && (d5 - d6) >= 0.f) {
    region = 5;
    float w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
    return b + w * (c - b);
  }
  region = 0;
  float denom = 1.f / (va + vb + vc);
  float v = vb * denom, w = vc * denom;
  return a + v * ab + w * ac;
}

// Signed distance to the mesh: nearest triangle through the BVH (children
// pushed far first, boxes no nearer than the best distance skipped), sign
// from the pseudo-normal of the closest feature (Baerentzen & Aanaes 2005).
__device__ float mesh_sdf(const DevMesh& M, v3 p) {
  float best = 3.402823466e+38f;
  uint32_t best_t = 0;
  int best_region = 0;
  v3 best_pt(0, 0, 0);
  int stack[kBuildStack];
  int sp = 0;
  stack[sp++] = M.root;
  while (sp > 0) {
    const int ni = stack[--sp];
    if (box_d2(M.nodes, ni, p) >= best) continue;
    const uint4 a = M.nodes[3 * ni], b = M.nodes[3 * ni + 1], c = M.nodes[3 * ni + 2];
    if (c.y) {  // leaf: triangles prims[start .. start + count)
      for (uint32_t k = c.x; k < c.x + c.y; k++) {
        const uint32_t t = M.prims[k];
        const uint4 tr = M.tri[t];
        int region;
        const v3 q = closest_on_triangle(p, ld3(M.v, tr.x), ld3(M.v, tr.y), ld3(M.v, tr.z), region);
        const v3 dv = p - q;
        const float d2 = dot(dv, dv);
        if (d2 < best) {
          best = d2;
          best_t = t;
          best_region = region;
          best_pt = q;
        }
      }
    } else if (sp + 2 <= kBuildStack) {  // the host checks the tree depth; the bound is a second guard
      const int l = (int)a.w, r = (int)b.w;
      const float dl = box_d2(M.nodes, l, p), dr = box_d2(M.nodes, r, p);
      if (dl < dr) {
        stack[sp++] = r;
        stack[sp++] = l;
      } else {
        stack[sp++] = l;
        stack[sp++] = r;
      }
    }
  }
  const uint4 tr = M.tri[best_t];
  v3 pn;
  switch (best_region) {
    case 0: pn = ld3(M.fnorm, best_t); break;
    case 1: pn = ld3(M.vnorm, tr.x); break;
    case 2: pn = ld3(M.vnorm, tr.y); break;
    case 3: pn = ld3(M.vnorm, tr.z); break;
    case 4: pn = ld3(M.enorm, 3 * best_t + 0); break;
    case 5: pn = ld3(M.enorm, 3 * best_t + 1); break;
    default: pn = ld3(M.enorm, 3 * best_t + 2); break;
  }
  const float d = sqrtf(best);
  return dot(p - best_pt, pn) < 0.f ? -d : d;
}

// One child of one candidate cell: kept if the distance at its centre lies
// within 1.5 half-diagonals of the band [-dmax_keep, 0] that a kept brick
// must reach (the 1-Lipschitz bound, with margin for inexact fields).
__global__ __launch_bounds__(256) void k_build_refine(DevMesh M, const uint4* __restrict__ cur, uint32_t n, float side, float r,
                                                      float dmax_keep, uint4* __restrict__ next, uint32_t* __restrict__ count) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= 8u * n) return;
  const uint4 pc = cur[i >> 3];
  const uint32_t c = i & 7u;
  const uint32_t x = pc.x * 2 + ((c >> 2) & 1u), y = pc.y * 2 + ((c >> 1) & 1u), z = pc.z * 2 + (c & 1u);
  const v3 ctr(-1.f + side * ((float)x + 0.5f), -1.f + side * ((float)y + 0.5f), -1.f + side * ((float)z + 0.5f));
  const float v = mesh_sdf(M, ctr);
  if (v <= r && v >= -(r + dmax_keep)) next[atomicAdd(count, 1u)] = make_uint4(x, y, z, 0);
}

// Distance sample s of candidate brick b at (-pad .. bs+pad)^3 lattice points.
__global__ __launch_bounds__(256) void k_build_fill(DevMesh M, const uint4* __restrict__ cells, uint32_t n, uint32_t lod,
                                                    uint32_t brick_size, uint32_t brick_pad, float* __restrict__ vals) {
  const uint32_t v_size = brick_size + 2 * brick_pad + 1, nv = v_size * v_size * v_size;
  const uint64_t g = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (g >= (uint64_t)n * nv) return;
  const uint32_t b = (uint32_t)(g / nv), s = (uint32_t)(g - (uint64_t)b * nv);
  const int i = (int)(s / (v_size * v_size)) - (int)brick_pad, j = (int)((s / v_size) % v_size) - (int)brick_pad,
            k = (int)(s % v_size) - (int)brick_pad;
  const uint4 c = cells[b];
  const float d = 1.f / (float)lod;
  const v3 p0(2.f * (d * (float)c.x) - 1.f, 2.f * (d * (float)c.y) - 1.f, 2.f * (d * (float)c.z) - 1.f);
  const float dp = 2.f * d / (float)brick_size;
  vals[g] = mesh_sdf(M, p0 + dp * v3((float)i, (float)j, (float)k));
}

// Border-keep rule (sparse_octree_builder.cpp:49-53) and the quantiser
// (:405-428) of one candidate brick; packed words are written only for kept
// bricks (the host compacts in candidate order).
__global__ __launch_bounds__(256) void k_build_pack(const float* __restrict__ vals, uint32_t n, uint32_t nv, uint32_t bytes_per_value,
                                                    float dmax_keep, double d_max, uint32_t* __restrict__ keep, uint32_t* __restrict__ words) {
  const uint32_t b = blockIdx.x * 256 + threadIdx.x;
  if (b >= n) return;
  const float* v = vals + (size_t)b * nv;
  float lo = 1000.f, hi = -1000.f;
  for (uint32_t i = 0; i < nv; i++) {
    lo = mn(lo, v[i]);
    hi = mx(hi, v[i]);
  }
  const bool k = lo <= 0.f && hi > -dmax_keep;
  keep[b] = k ? 1u : 0u;
  if (!k) return;
  const uint32_t vals_per_int = 4 / bytes_per_value, bits = 8 * bytes_per_value;
  const uint32_t n_words = (nv + vals_per_int - 1) / vals_per_int;
  const double max_val = bytes_per_value == 4 ? 4294967295.0 : (double)((1u << bits) - 1u);
  uint32_t* dst = words + (size_t)b * n_words;
  uint32_t w = 0;
  for (uint32_t i = 0; i < nv; i++) {
    double q = floor(max_val * (((double)v[i] + d_max) / (2.0 * d_max)));
    q = (max_val < q) ? max_val : q;  // std::min / std::max as the host builder uses them
    q = (0.0 < q) ? q : 0.0;
    w |= ((uint32_t)q) << (bits * (i % vals_per_int));
    if (i % vals_per_int == vals_per_int - 1 || i == nv - 1) {
      dst[i / vals_per_int] = w;
      w = 0;
    }
  }
}

// ---- octree builder ---------------------------------------------------------
// create_sdf_SBS(settings, header, sdf) (utils/sdf_converter.cpp:169-181):
// the frame octree of construct_sdf_frame_octree (sparse_octree_builder.cpp:
// 125-258) built breadth-first, one launch per level, then frame_octree_to_SBS
// (:291-444) over its leaves. Positions follow the reference's float
// expressions exactly (cell coordinates are integers, d = 2^-level).

__device__ __forceinline__ float field_sdf(const DevField& F, v3 p) {
  if (F.kind == 1) {
    const v3 q = p - v3(F.a[0], F.a[1], F.a[2]);
    return sqrtf(q.x * q.x + q.y * q.y + q.z * q.z) - F.a[3];
  }
  if (F.kind == 2) {
    const float qx = sqrtf(p.x * p.x + p.y * p.y) - F.a[0];
    return sqrtf(qx * qx + p.z * p.z) - F.a[1];
  }
  return mesh_sdf(F.mesh, p);
}

// vals[9 c + k]: corner k = 2 * ((p + bits(k)) * d) - 1, centre k = 8 at
// 2 * ((p + 0.5) * d) - 1 (add_node_rec, :74-82)
__global__ __launch_bounds__(256) void k_octree_eval(DevField F, const uint4* __restrict__ cells, uint32_t n, float* __restrict__ vals) {
  const uint64_t g = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (g >= 9ull * n) return;
  const uint32_t c = (uint32_t)(g / 9), k = (uint32_t)(g - 9ull * c);
  const uint4 cell = cells[c];
  const float d = ldexpf(1.0f, -(int)cell.w);
  const float ox = k == 8 ? 0.5f : (float)((k >> 2) & 1u), oy = k == 8 ? 0.5f : (float)((k >> 1) & 1u), oz = k == 8 ? 0.5f : (float)(k & 1u);
  const v3 p(2.0f * (((float)cell.x + ox) * d) - 1.0f, 2.0f * (((float)cell.y + oy) * d) - 1.0f, 2.0f * (((float)cell.z + oz) * d) - 1.0f);
  vals[g] = field_sdf(F, p);
}

__global__ __launch_bounds__(256) void k_octree_classify(const uint4* __restrict__ cells, const float* __restrict__ vals, uint32_t n,
                                                         uint32_t level, int top, uint32_t depth, OctreeLevels T, uint4* __restrict__ kids,
                                                         uint32_t* __restrict__ n_kids, uint4* __restrict__ leaves,
                                                         float* __restrict__ leaf_vals, uint32_t* __restrict__ n_leaves) {
  const uint32_t c = blockIdx.x * 256 + threadIdx.x;
  if (c >= n) return;
  const float* v = vals + 9ull * c;
  float mnv = 1000.0f, mxv = -1000.0f;
  for (int k = 0; k < 8; k++) {
    mnv = fminf(mnv, v[k]);
    mxv = fmaxf(mxv, v[k]);
  }
  const float centre = fabsf(v[8]);
  // is_border (:38-41) gates the subtrees of the last uniform level; every
  // deeper node splits by add_node_rec's test (:84)
  const bool refine = centre < T.near[level] || mnv * mxv <= 0.0f;
  const bool split = level < depth && (top ? (level < 2 || centre < T.border[level]) && refine : refine);
  const uint4 cell = cells[c];
  if (split) {
    const uint32_t o = atomicAdd(n_kids, 8u);
    for (uint32_t k = 0; k < 8; k++)
      kids[o + k] = make_uint4(2 * cell.x + ((k >> 2) & 1u), 2 * cell.y + ((k >> 1) & 1u), 2 * cell.z + (k & 1u), level + 1);
  } else {
    const uint32_t o = atomicAdd(n_leaves, 1u);
    leaves[o] = cell;
    for (int k = 0; k < 8; k++) leaf_vals[8ull * o + k] = v[k];
  }
}

// One lattice point of one leaf (frame_octree_to_SBS, :327-386): corners take
// the leaf's values, the others the field at p0 + dp * (i, j, k).
__global__ __launch_bounds__(256) void k_octree_fill(DevField F, const uint4* __restrict__ leaves, const float* __restrict__ leaf_vals, uint32_t n,
                                                     uint32_t brick_size, uint32_t brick_pad, float* __restrict__ vals) {
  const uint32_t v_size = brick_size + 2 * brick_pad + 1, nv = v_size * v_size * v_size;
  const uint64_t g = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (g >= (uint64_t)n * nv) return;
  const uint32_t b = (uint32_t)(g / nv), s = (uint32_t)(g - (uint64_t)b * nv);
  const int bs = (int)brick_size, pad = (int)brick_pad;
  const int i = (int)(s / (v_size * v_size)) - pad, j = (int)((s / v_size) % v_size) - pad, k = (int)(s % v_size) - pad;
  if ((i == 0 || i == bs) && (j == 0 || j == bs) && (k == 0 || k == bs)) {
    vals[g] = leaf_vals[8ull * b + (i == bs ? 4 : 0) + (j == bs ? 2 : 0) + (k == bs ? 1 : 0)];
    return;
  }
  const uint4 c = leaves[b];
  const float d = ldexpf(1.0f, -(int)c.w);
  const v3 p0(2.0f * (d * (float)c.x) - 1.0f, 2.0f * (d * (float)c.y) - 1.0f, 2.0f * (d * (float)c.z) - 1.0f);
  const float dp = 2.0f * d / (float)brick_size;
  vals[g] = field_sdf(F, p0 + dp * v3((float)i, (float)j, (float)k));
}

// is_border_node (:49-53) and the quantiser (:405-428) of one leaf, in the
// reference's float arithmetic: q = max(0, q_max * ((v + d_max) / (2 d_max))),
// truncated, then min(q, q_max); a product of 2^32 or more (32-bit values at
// the top of the range, where the reference's conversion is undefined) gives
// q_max.
__global__ __launch_bounds__(256) void k_octree_pack(const uint4* __restrict__ leaves, const float* __restrict__ vals, uint32_t n, uint32_t nv,
                                                     uint32_t bytes_per_value, OctreeLevels T, uint32_t* __restrict__ keep,
                                                     uint32_t* __restrict__ words) {
  const uint32_t b = blockIdx.x * 256 + threadIdx.x;
  if (b >= n) return;
  const float* v = vals + (size_t)b * nv;
  float lo = 1000.f, hi = -1000.f;
  for (uint32_t i = 0; i < nv; i++) {
    lo = fminf(lo, v[i]);
    hi = fmaxf(hi, v[i]);
  }
  const float d_max = T.keep[leaves[b].w];
  const bool k = lo <= 0.f && hi > -d_max;
  keep[b] = k ? 1u : 0u;
  if (!k) return;
  const uint32_t vals_per_int = 4 / bytes_per_value, bits = 8 * bytes_per_value;
  const uint32_t n_words = (nv + vals_per_int - 1) / vals_per_int;
  const uint32_t q_max = bytes_per_value == 4 ? 0xFFFFFFFFu : ((1u << bits) - 1u);
  uint32_t* dst = words + (size_t)b * n_words;
  uint32_t w = 0;
  for (uint32_t i = 0; i < nv; i++) {
    const float q = fmaxf(0.0f, (float)q_max * ((v[i] + d_max) / (2.0f * d_max)));
    uint32_t dc = q >= 4294967296.0f ? q_max : (uint32_t)q;
    dc = dc < q_max ? dc : q_max;
    w |= dc << (bits * (i % vals_per_int));
    if (i % vals_per_int == vals_per_int - 1 || i == nv - 1) {
      dst[i / vals_per_int] = w;
      w = 0;
    }
  }
}

}  // namespace lrtd

namespace lrt {

hipError_t launch_octree_eval(const lrtd::DevField& F, const uint4* cells, uint32_t n, float* vals, hipStream_t s) {
  const uint64_t threads = 9ull * n;
  if (threads == 0) return hipSuccess;
  hipLaunchKernelGGL(lrtd::k_octree_eval, dim3((uint32_t)((threads + 255) / 256)), dim3(256), 0, s, F, cells, n, vals);
  return hipGetLastError();
}
hipError_t launch_octree_classify(const uint4* cells, const float* vals, uint32_t n, uint32_t level, int top, uint32_t depth,
                                  const lrtd::OctreeLevels& T, uint4* kids, uint32_t* n_kids, uint4* leaves, float* leaf_vals,
                                  uint32_t* n_leaves, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(lrtd::k_octree_classify, dim3((n + 255) / 256), dim3(256), 0, s, cells, vals, n, level, top, depth, T, kids, n_kids,
                     leaves, leaf_vals, n_leaves);
  return hipGetLastError();
}
hipError_t launch_octree_fill(const lrtd::DevField& F, const uint4* leaves, const float* leaf_vals, uint32_t n, uint32_t brick_size,
                              uint32_t brick_pad, float* vals, hipStream_t s) {
  const uint32_t v_size = brick_size + 2 * brick_pad + 1;
  const uint64_t threads = (uint64_t)n * v_size * v_size * v_size;
  if (threads == 0) return hipSuccess;
  hipLaunchKernelGGL(lrtd::k_octree_fill, dim3((uint32_t)((threads + 255) / 256)), dim3(256), 0, s, F, leaves, leaf_vals, n, brick_size,
                     brick_pad, vals);
  return hipGetLastError();
}
hipError_t launch_octree_pack(const uint4* leaves, const float* vals, uint32_t n, uint32_t nv, uint32_t bytes_per_value,
                              const lrtd::OctreeLevels& T, uint32_t* keep, uint32_t* words, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(lrtd::k_octree_pack, dim3((n + 255) / 256), dim3(256), 0, s, leaves, vals, n, nv, bytes_per_value, T, keep, words);
  return hipGetLastError();
}

hipError_t launch_build_refine(const lrtd::DevMesh& M, const uint4* cur, uint32_t n, float side, float r, float dmax_keep, uint4* next,
                               uint32_t* count, hipStream_t s) {
  const uint64_t threads = 8ull * n;
  hipLaunchKernelGGL(lrtd::k_build_refine, dim3((uint32_t)((threads + 255) / 256)), dim3(256), 0, s, M, cur, n, side, r, dmax_keep, next,
                     count);
  return hipGetLastError();
}
hipError_t launch_build_fill(const lrtd::DevMesh& M, const uint4* cells, uint32_t n, uint32_t lod, uint32_t brick_size, uint32_t brick_pad,
                             float* vals, hipStream_t s) {
  const uint32_t v_size = brick_size + 2 * brick_pad + 1;
  const uint64_t threads = (uint64_t)n * v_size * v_size * v_size;
  hipLaunchKernelGGL(lrtd::k_build_fill, dim3((uint32_t)((threads + 255) / 256)), dim3(256), 0, s, M, cells, n, lod, brick_size, brick_pad,
                     vals);
  return hipGetLastError();
}
hipError_t launch_build_pack(const float* vals, uint32_t n, uint32_t nv, uint32_t bytes_per_value, float dmax_keep, double d_max,
                             uint32_t* keep, uint32_t* words, hipStream_t s) {
  hipLaunchKernelGGL(lrtd::k_build_pack, dim3((n + 255) / 256), dim3(256), 0, s, vals, n, nv, bytes_per_value, dmax_keep, d_max, keep,
                     words);
  return hipGetLastError();
}

}  // namespace lrt

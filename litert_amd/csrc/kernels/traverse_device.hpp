// traverse_device.hpp - BVH2 traversal for one ray on gfx950.
//
// Same visiting order as BVHRT::BVH2TraverseF32 (BVH2Common.cpp:3407-3488):
// both children tested per fat pair, nearer child first, stop at the first
// SDF-brick hit. The traversal stack lives in LDS (STACK_LDS entries per
// lane, lane-interleaved so a wave touches 64 consecutive dwords per push)
// and spills deeper entries to a per-lane global slice.
#pragma once
#include "trace_device.hpp"

namespace lrtd {

typedef __attribute__((address_space(3))) uint32_t lds_u32;  // an LDS pointer: ds_read / ds_write, never flat

struct Stack {
  lds_u32* lds;         // base for this lane: entry k at lds[k * lds_stride]
  uint32_t lds_stride;  // lanes per block
  uint32_t* spill;      // base for this lane: entry k at spill[k * spill_stride]
  uint32_t spill_stride;
  uint32_t lds_entries;
  __device__ __forceinline__ void put(int k, uint32_t v) {
    if ((uint32_t)k < lds_entries) lds[k * lds_stride] = v;
    else spill[(k - (int)lds_entries) * spill_stride] = v;
  }
  __device__ __forceinline__ uint32_t get(int k) const {
    return ((uint32_t)k < lds_entries) ? lds[k * lds_stride] : spill[(k - (int)lds_entries) * spill_stride];
  }
};

struct WorkCount {
  uint32_t pairs = 0, bricks = 0, voxels = 0;
  uint32_t iters = 0;   // root-finder iterations
  uint32_t tiles = 0;   // 16-sample MLP tiles evaluated (seeded kernel)
  uint32_t seeds = 0;   // brick entries that passed the confidence gate
  uint32_t culled = 0;  // 8x8 tiles written as misses by the tile cull (once per tile)
};

// Node pair `left` (BVHNodePair, 64 B) as four 16-B loads through a buffer
// resource (fewer than 2^24 pairs: the scene's brick limit).
__device__ __forceinline__ void load_pair(const DevScene& S, uint32_t left, uint4& a, uint4& b, uint4& c, uint4& e) {
  typedef __attribute__((ext_vector_type(4))) unsigned int u4v;
  const __amdgpu_buffer_rsrc_t rp = buf_rsrc(S.pairs, S.pairs_bytes);
  const uint32_t off = left * 64u;
  const u4v x = __builtin_amdgcn_raw_buffer_load_b128(rp, off, 0, 0), y = __builtin_amdgcn_raw_buffer_load_b128(rp, off + 16, 0, 0);
  const u4v z = __builtin_amdgcn_raw_buffer_load_b128(rp, off + 32, 0, 0), w = __builtin_amdgcn_raw_buffer_load_b128(rp, off + 48, 0, 0);
  a = make_uint4(x.x, x.y, x.z, x.w);
  b = make_uint4(y.x, y.y, y.z, y.w);
  c = make_uint4(z.x, z.y, z.z, z.w);
  e = make_uint4(w.x, w.y, w.z, w.w);
}

// Slab test of one box for a ray whose direction signs are known: bit a of
// OCT set when component a of the inverse direction is negative. The near and
// far plane of each axis are then known, so min(lo, hi) / max(lo, hi) of
// RayBox are the near / far products themselves (rounding is monotonic, so
// the order of the two products follows the order of the planes): the same
// two values with 6 fewer instructions. OCT < 0: the general form.
template <int OCT>
__device__ __forceinline__ float2 RayBoxOct(v3 o, v3 inv, v3 bmin, v3 bmax, bool ref = false) {
  if (OCT < 0) return RayBoxW(ref, o, inv, bmin, bmax);
  const float nx = inv.x * (((OCT & 1) ? bmax.x : bmin.x) - o.x), fx = inv.x * (((OCT & 1) ? bmin.x : bmax.x) - o.x);
  const float ny = inv.y * (((OCT & 2) ? bmax.y : bmin.y) - o.y), fy = inv.y * (((OCT & 2) ? bmin.y : bmax.y) - o.y);
  const float nz = inv.z * (((OCT & 4) ? bmax.z : bmin.z) - o.z), fz = inv.z * (((OCT & 4) ? bmin.z : bmax.z) - o.z);
  return make_float2(fmaxf(fmaxf(nx, ny), nz), fminf(fminf(fx, fy), fz));
}

// Node-pair part of BVH2TraverseF32 (BVH2Common.cpp:3431-3466): descends
// until the lane sits at a leaf or its stack is empty.
template <bool COUNT, int OCT>
__device__ __forceinline__ void DescendPairs(const DevScene& S, v3 ray_pos, v3 inv, float tNear, bool stopOnFirstHit, float hit_t, int& top,
                                             uint32_t& left, Stack& st, WorkCount* wc, bool ref) {
  while (top >= 0 && ((left & LRT_LEAF_BIT) == 0)) {
    uint4 a, b, c, e;
    load_pair(S, left, a, b, c, e);
    if (COUNT) wc->pairs++;
    const float2 tm0 = RayBoxOct<OCT>(ray_pos, inv, v3(__uint_as_float(a.x), __uint_as_float(a.y), __uint_as_float(a.z)),
                                      v3(__uint_as_float(b.x), __uint_as_float(b.y), __uint_as_float(b.z)), ref);
    const float2 tm1 = RayBoxOct<OCT>(ray_pos, inv, v3(__uint_as_float(c.x), __uint_as_float(c.y), __uint_as_float(c.z)),
                                      v3(__uint_as_float(e.x), __uint_as_float(e.y), __uint_as_float(e.z)), ref);
    const uint32_t n0 = a.w, n1 = c.w;
    const bool h0 = (tm0.x <= tm0.y) && (tm0.y >= tNear) && (tm0.x <= hit_t || !stopOnFirstHit);
    const bool h1 = (tm1.x <= tm1.y) && (tm1.y >= tNear) && (tm1.x <= hit_t || !stopOnFirstHit);
    left = h0 ? n0 : n1;
    if (h0 && h1) {
      const bool near0 = tm0.x <= tm1.x;
      left = near0 ? n0 : n1;
      st.put(top, near0 ? n1 : n0);
      top++;
    }
    if (!h0 && !h1) {
      top--;
      left = st.get(top > 0 ? top : 0);
    }
  }
}

// The wave's common direction octant (bit a: inverse direction component a
// negative), or -1 when the active lanes differ or a component is infinite
// (a zero direction component: there the general slab test's min / max also
// decide what an infinite product against a plane through the origin gives).
__device__ __forceinline__ int wave_octant(v3 inv) {
  const bool fin = fabsf(inv.x) < INFINITY && fabsf(inv.y) < INFINITY && fabsf(inv.z) < INFINITY;
  const uint32_t o = (inv.x < 0.0f ? 1u : 0u) | (inv.y < 0.0f ? 2u : 0u) | (inv.z < 0.0f ? 4u : 0u) | (fin ? 0u : 8u);
  const uint32_t o0 = __builtin_amdgcn_readfirstlane(o);
  const bool same = __builtin_amdgcn_ballot_w64(o != o0) == 0;
  return (same && o0 < 8u) ? (int)o0 : -1;
}

template <bool COUNT>
__device__ __forceinline__ void DescendPairsAny(int oct, const DevScene& S, v3 ray_pos, v3 inv, float tNear, bool stopOnFirstHit, float hit_t,
                                                int& top, uint32_t& left, Stack& st, WorkCount* wc, bool ref) {
  switch (oct) {  // wave-uniform
    case 0: DescendPairs<COUNT, 0>(S, ray_pos, inv, tNear, stopOnFirstHit, hit_t, top, left, st, wc, false); break;
    case 1: DescendPairs<COUNT, 1>(S, ray_pos, inv, tNear, stopOnFirstHit, hit_t, top, left, st, wc, false); break;
    case 2: DescendPairs<COUNT, 2>(S, ray_pos, inv, tNear, stopOnFirstHit, hit_t, top, left, st, wc, false); break;
    case 3: DescendPairs<COUNT, 3>(S, ray_pos, inv, tNear, stopOnFirstHit, hit_t, top, left, st, wc, false); break;
    case 4: DescendPairs<COUNT, 4>(S, ray_pos, inv, tNear, stopOnFirstHit, hit_t, top, left, st, wc, false); break;
    case 5: DescendPairs<COUNT, 5>(S, ray_pos, inv, tNear, stopOnFirstHit, hit_t, top, left, st, wc, false); break;
    case 6: DescendPairs<COUNT, 6>(S, ray_pos, inv, tNear, stopOnFirstHit, hit_t, top, left, st, wc, false); break;
    case 7: DescendPairs<COUNT, 7>(S, ray_pos, inv, tNear, stopOnFirstHit, hit_t, top, left, st, wc, false); break;
    default: DescendPairs<COUNT, -1>(S, ray_pos, inv, tNear, stopOnFirstHit, hit_t, top, left, st, wc, ref); break;
  }
}

// BVH2Common.cpp:3407-3488. Returns through *hit; stack capacity is
// lds_entries + spill entries (>= the tree's pair depth, checked on the host).
template <bool COUNT, int RF = -1, bool IDX = true>
__device__ __forceinline__ void BVH2Traverse(const DevScene& S, const TraceParams& P, v3 ray_pos, v3 ray_dir, v3 inv, float tNear,
                                             bool stopOnFirstHit, Hit* hit, Stack& st, WorkCount* wc) {
  int top = 0;
  uint32_t left = 0;
  bool hitFound = false;
  const int oct = wave_octant(inv);  // -1 when a lane's inverse direction has an infinite component
  const bool ref = wave_ref_slabs(inv);
  while (top >= 0 && !hitFound) {
    DescendPairsAny<COUNT>(oct, S, ray_pos, inv, tNear, stopOnFirstHit, hit->t, top, left, st, wc, ref);
    if (top >= 0 && left != 0xFFFFFFFFu) {
      const float tNearSdf = mx(tNear, 0.1f);  // SDF_BIAS, BVH2Common.cpp:3472-3473
      const float tPrev = hit->t;
      if (COUNT) wc->bricks++;
      BrickIntersect<false, RF, IDX, !IDX>(S, P, ray_pos, ray_dir, inv, tNearSdf, left & LRT_START_MASK, hit, COUNT ? &wc->voxels : nullptr, -1.0f,
                     COUNT ? &wc->iters : nullptr);
      hitFound = hit->t < tPrev;
    }
    top--;
    left = st.get(top > 0 ? top : 0);
  }
}

// BVH2Common.cpp:3490-3559 for the single identity instance. NORMAL = false
// leaves the hit normal out (shadow rays use only hit / miss).
template <bool COUNT, int RF = -1, bool IDX = true, bool NORMAL = true>
__device__ __forceinline__ Hit RayQuery_NearestHit(const DevScene& S, const TraceParams& P, v3 pos, float tNear, v3 dir, float tFar,
                                                   Stack& st, WorkCount* wc) {
  bool stopOnFirstHit = (tFar <= 0.0f);
  if (stopOnFirstHit) tFar *= -1.0f;
  const v3 inv = SafeInverse(dir);
  Hit hit;
  hit.t = tFar;
  hit.primId = hit.instId = hit.geomId = 0xFFFFFFFFu;
  hit.coords[0] = 1.0f; hit.coords[1] = 0.0f; hit.coords[2] = 0.0f; hit.coords[3] = 0.0f;
  const float2 bh = RayBoxW(wave_ref_slabs(inv), pos, inv, v3(-1, -1, -1), v3(1, 1, 1));
  if ((bh.x <= bh.y) && (bh.y > tNear) && (bh.x < hit.t)) BVH2Traverse<COUNT, RF, IDX>(S, P, pos, dir, inv, tNear, stopOnFirstHit, &hit, st, wc);
  // DX layouts: the hit normal, once per ray after the traversal (all hit
  // lanes of the wave together rather than in every voxel step that hits)
  if (!IDX && NORMAL) FinishNormal(S, P, &hit);
  if (hit.primId != 0xFFFFFFFFu) hit.instId = 0;
  return hit;
}

// Resumable form of the same traversal, for kernels that do wave-wide work
// at every leaf (the learned seeder, the training-event collector). Advances
// (top, left) to the next brick leaf in exactly the order BVH2Traverse visits
// them and returns true, or returns false when the stack is exhausted. After
// a leaf that did not end the ray the caller pops: top--; left = get(top).
template <bool COUNT>
__device__ __forceinline__ bool NextLeaf(const DevScene& S, v3 ray_pos, v3 inv, float tNear, bool stopOnFirstHit, float hit_t, int& top,
                                         uint32_t& left, Stack& st, WorkCount* wc) {
  const int oct = wave_octant(inv);
  const bool ref = wave_ref_slabs(inv);
  for (;;) {
    DescendPairsAny<COUNT>(oct, S, ray_pos, inv, tNear, stopOnFirstHit, hit_t, top, left, st, wc, ref);
    if (top < 0) return false;
    if (left != 0xFFFFFFFFu) return true;
    top--;
    left = st.get(top > 0 ? top : 0);
  }
}

}  // namespace lrtd

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
};

// BVH2Common.cpp:3407-3488. Returns through *hit; stack capacity is
// lds_entries + spill entries (>= the tree's pair depth, checked on the host).
template <bool COUNT, int RF = -1, bool IDX = true>
__device__ __forceinline__ void BVH2Traverse(const DevScene& S, const TraceParams& P, v3 ray_pos, v3 ray_dir, v3 inv, float tNear,
                                             bool stopOnFirstHit, Hit* hit, Stack& st, WorkCount* wc) {
  int top = 0;
  uint32_t left = 0;
  bool hitFound = false;
  while (top >= 0 && !hitFound) {
    while (top >= 0 && ((left & LRT_LEAF_BIT) == 0)) {
      const uint4* pp = S.pairs + 4 * (size_t)left;
      const uint4 a = pp[0], b = pp[1], c = pp[2], e = pp[3];
      if (COUNT) wc->pairs++;
      const float2 tm0 = RayBox(ray_pos, inv, v3(__uint_as_float(a.x), __uint_as_float(a.y), __uint_as_float(a.z)),
                                v3(__uint_as_float(b.x), __uint_as_float(b.y), __uint_as_float(b.z)));
      const float2 tm1 = RayBox(ray_pos, inv, v3(__uint_as_float(c.x), __uint_as_float(c.y), __uint_as_float(c.z)),
                                v3(__uint_as_float(e.x), __uint_as_float(e.y), __uint_as_float(e.z)));
      const uint32_t n0 = a.w, n1 = c.w;
      const bool h0 = (tm0.x <= tm0.y) && (tm0.y >= tNear) && (tm0.x <= hit->t || !stopOnFirstHit);
      const bool h1 = (tm1.x <= tm1.y) && (tm1.y >= tNear) && (tm1.x <= hit->t || !stopOnFirstHit);
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
  const float2 bh = RayBox(pos, inv, v3(-1, -1, -1), v3(1, 1, 1));
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
  for (;;) {
    while (top >= 0 && ((left & LRT_LEAF_BIT) == 0)) {
      const uint4* pp = S.pairs + 4 * (size_t)left;
      const uint4 a = pp[0], b = pp[1], c = pp[2], e = pp[3];
      if (COUNT) wc->pairs++;
      const float2 tm0 = RayBox(ray_pos, inv, v3(__uint_as_float(a.x), __uint_as_float(a.y), __uint_as_float(a.z)),
                                v3(__uint_as_float(b.x), __uint_as_float(b.y), __uint_as_float(b.z)));
      const float2 tm1 = RayBox(ray_pos, inv, v3(__uint_as_float(c.x), __uint_as_float(c.y), __uint_as_float(c.z)),
                                v3(__uint_as_float(e.x), __uint_as_float(e.y), __uint_as_float(e.z)));
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
    if (top < 0) return false;
    if (left != 0xFFFFFFFFu) return true;
    top--;
    left = st.get(top > 0 ? top : 0);
  }
}

}  // namespace lrtd

// oracle_render.hpp - CPU restatement of LiteRT's SBS primary-ray path.
//
// TEST INFRASTRUCTURE ONLY (see oracle_math.hpp). Each function names the
// reference file:line it restates. Templated on T: float is the FP32 twin of
// the reference CPU path, double the FP64 error reference.
#pragma once
#include <cstring>
#include <vector>
#include "oracle_math.hpp"
#include "../include/lrt_types.h"

namespace orc {

// Work counters used by the algorithmic-bytes model (one set per ray).
struct Counters {
  uint64_t node_pairs = 0;   // BVHNodePair fetches (BVH2Common.cpp:3431)
  uint64_t bricks = 0;       // leaf/brick entries (BVH2Common.cpp:3480)
  uint64_t voxels = 0;       // 8-corner loads (BVH2Common.cpp:801-817)
  uint64_t root_finds = 0;   // LocalSurfaceIntersection calls
  uint64_t rays = 0;
  uint64_t root_iters = 0;   // ST / IT / Newton iterations inside LocalSurfaceIntersection
  void add(const Counters& o) {
    node_pairs += o.node_pairs; bricks += o.bricks; voxels += o.voxels; root_finds += o.root_finds; rays += o.rays;
    root_iters += o.root_iters;
  }
};

// Scene as the reference BVHRT holds it after AddGeom_SdfSBS + CommitScene
// (BVH2Common_host.cpp:567-666, 1416-1437) for a single identity instance.
struct Scene {
  SdfSBSHeader header{};
  uint32_t quant_voxel_diag = 0;  // aux_data bit 0: DX values scaled by the voxel (not brick) diagonal
  std::vector<SdfSBSNode> nodes;
  std::vector<uint32_t> values;
  std::vector<float> values_f;
  std::vector<BVHNodePair> pairs;  // fat BVH2 (cbvh.h:30-33), root pair at index 0
  uint32_t type = LRT_TYPE_SDF_SBS;
  // MultiRenderer::m_lights; default: direct along normalize(1,1,1) of colour
  // 2/3, ambient 0.25 (eye_ray_host.cpp:40), the direction as v / |v| in float
  std::vector<LrtLight> lights = {LrtLight{{1.0f / std::sqrt(3.0f), 1.0f / std::sqrt(3.0f), 1.0f / std::sqrt(3.0f)}, LRT_LIGHT_TYPE_DIRECT,
                                           {2.0f / 3.0f, 2.0f / 3.0f, 2.0f / 3.0f}, 0},
                                  LrtLight{{0, 0, 0}, LRT_LIGHT_TYPE_AMBIENT, {0.25f, 0.25f, 0.25f}, 0}};
};

template <typename T> struct Hit {
  T t;
  uint32_t primId, instId, geomId;
  T coords[4];
};

struct Options {
  MultiRenderPreset preset;
  bool force_normal = false;  // compute normals whatever the render mode (G-buffer output)
};

// raytrace_common.h:59-82
template <typename T> static inline V3<T> SafeInverse(V3<T> d) {
  const T ooeps = T(1.0e-36);
  V3<T> res;
  res.x = T(1) / (std::abs(d.x) > ooeps ? d.x : std::abs(ooeps) * sgn(d.x));
  res.y = T(1) / (std::abs(d.y) > ooeps ? d.y : std::abs(ooeps) * sgn(d.y));
  res.z = T(1) / (std::abs(d.z) > ooeps ? d.z : std::abs(ooeps) * sgn(d.z));
  return res;
}

// raytrace_common.h:59-82 (slab test with the precomputed inverse direction)
template <typename T> static inline V2<T> RayBoxIntersection2(V3<T> o, V3<T> inv, V3<T> bmin, V3<T> bmax) {
  const T lo = inv.x * (bmin.x - o.x);
  const T hi = inv.x * (bmax.x - o.x);
  const T lo1 = inv.y * (bmin.y - o.y);
  const T hi1 = inv.y * (bmax.y - o.y);
  const T lo2 = inv.z * (bmin.z - o.z);
  const T hi2 = inv.z * (bmax.z - o.z);
  const T tmin = mx(mn(lo, hi), mn(lo1, hi1));
  const T tmax = mn(mx(lo, hi), mx(lo1, hi1));
  return {mx(tmin, mn(lo2, hi2)), mn(tmax, mx(lo2, hi2))};
}

// BVH2Common.cpp:62-72
template <typename T> static inline T eval_dist_trilinear(const T v[8], V3<T> dp) {
  return (1 - dp.x) * (1 - dp.y) * (1 - dp.z) * v[0] +
         (1 - dp.x) * (1 - dp.y) * (dp.z) * v[1] +
         (1 - dp.x) * (dp.y) * (1 - dp.z) * v[2] +
         (1 - dp.x) * (dp.y) * (dp.z) * v[3] +
         (dp.x) * (1 - dp.y) * (1 - dp.z) * v[4] +
         (dp.x) * (1 - dp.y) * (dp.z) * v[5] +
         (dp.x) * (dp.y) * (1 - dp.z) * v[6] +
         (dp.x) * (dp.y) * (dp.z) * v[7];
}

// BVH2Common.cpp:657-687
template <typename T> static inline V3<T> eval_dist_trilinear_diff(const T v[8], V3<T> dp) {
  T ddx = -(1 - dp.y) * (1 - dp.z) * v[0] + -(1 - dp.y) * (dp.z) * v[1] + -(dp.y) * (1 - dp.z) * v[2] + -(dp.y) * (dp.z) * v[3] +
          (1 - dp.y) * (1 - dp.z) * v[4] + (1 - dp.y) * (dp.z) * v[5] + (dp.y) * (1 - dp.z) * v[6] + (dp.y) * (dp.z) * v[7];
  T ddy = -(1 - dp.x) * (1 - dp.z) * v[0] + -(1 - dp.x) * (dp.z) * v[1] + (1 - dp.x) * (1 - dp.z) * v[2] + (1 - dp.x) * (dp.z) * v[3] +
          -(dp.x) * (1 - dp.z) * v[4] + -(dp.x) * (dp.z) * v[5] + (dp.x) * (1 - dp.z) * v[6] + (dp.x) * (dp.z) * v[7];
  T ddz = -(1 - dp.x) * (1 - dp.y) * v[0] + (1 - dp.x) * (1 - dp.y) * v[1] + -(1 - dp.x) * (dp.y) * v[2] + (1 - dp.x) * (dp.y) * v[3] +
          -(dp.x) * (1 - dp.y) * v[4] + (dp.x) * (1 - dp.y) * v[5] + -(dp.x) * (dp.y) * v[6] + (dp.x) * (dp.y) * v[7];
  return {ddx, ddy, ddz};
}

// BVH2Common.cpp:30-35 (octahedral encoding, JCGT 2014)
template <typename T> static inline V2<T> encode_normal(V3<T> v) {
  T s = T(1) / (std::abs(v.x) + std::abs(v.y) + std::abs(v.z));
  V2<T> p{v.x * s, v.y * s};
  V2<T> sn{(p.x >= T(0)) ? T(1) : T(-1), (p.y >= T(0)) ? T(1) : T(-1)};
  if (v.z <= T(0)) return {(T(1) - std::abs(p.y)) * sn.x, (T(1) - std::abs(p.x)) * sn.y};
  return p;
}

// Renderer/eye_ray.cpp:22-32
template <typename T> static inline V3<T> decode_normal(V2<T> e) {
  V3<T> v{e.x, e.y, T(1) - std::abs(e.x) - std::abs(e.y)};
  if (v.z < 0) {
    T vx = v.x;
    v.x = (T(1) - std::abs(v.y)) * ((v.x >= 0.0) ? T(1) : T(-1));
    v.y = (T(1) - std::abs(vx)) * ((v.y >= 0.0) ? T(1) : T(-1));
  }
  return normalize(v);
}

static inline bool need_normal(const Options& o) {  // BVH2Common.cpp:20-27
  unsigned m = o.preset.render_mode;
  return o.force_normal || m == LRT_RENDER_MODE_LAMBERT_NO_TEX || m == LRT_RENDER_MODE_NORMAL ||
         m == LRT_RENDER_MODE_PHONG_NO_TEX || m == LRT_RENDER_MODE_LAMBERT || m == LRT_RENDER_MODE_PHONG;
}

template <typename T> struct Tracer {
  const Scene& S;
  Options O;
  Counters* cnt;
  // last LocalSurfaceIntersection result in voxel units (unit tests only)
  mutable T last_t = 0;
  mutable bool last_hit = false;
  mutable unsigned last_iter = 0;
  Tracer(const Scene& s, const Options& o, Counters* c) : S(s), O(o), cnt(c) {}

  uint32_t layout() const { return S.header.aux_data & LRT_SBS_LAYOUT_MASK; }
  bool indexed() const { return layout() == LRT_SBS_LAYOUT_ID32F_IRGB32F || layout() == LRT_SBS_LAYOUT_ID32F_IRGB32F_IN; }

  // sdf_scene.h:201-204
  static uint32_t v_to_i(uint32_t i, uint32_t j, uint32_t k, uint32_t v, uint32_t pad) {
    return (i + pad) * v * v + (j + pad) * v + (k + pad);
  }

  // BVH2Common.cpp:725-762 (trilinear branch). voxel position given as integers.
  T load_distance_values(uint32_t nodeId, const uint32_t vp[3], uint32_t v_size, T sz_inv, T values[8]) const {
    T vmin = T(1e6f);
    const SdfSBSHeader& h = S.header;
    const uint32_t v_off = S.nodes[nodeId].data_offset;
    if (indexed()) {
      for (int i = 0; i < 8; i++) {
        uint32_t vId = v_to_i(vp[0] + ((i & 4) >> 2), vp[1] + ((i & 2) >> 1), vp[2] + (i & 1), v_size, h.brick_pad);
        values[i] = T(S.values_f[S.values[v_off + vId]]);
        vmin = mn(vmin, values[i]);
      }
    } else {
      uint32_t vals_per_int = 4 / h.bytes_per_value;
      uint32_t bits = 8 * h.bytes_per_value;
      uint32_t max_val = h.bytes_per_value == 4 ? 0xFFFFFFFFu : ((1u << bits) - 1);
      // BVH2Common.cpp:749-750; the voxel-diagonal scale is this build's
      // documented extension (aux_data bit 0).
      T d_max = S.quant_voxel_diag ? T(1.73205081f) * (sz_inv / T(h.brick_size)) : T(1.73205081f) * sz_inv;
      T mult = T(2) * d_max / T(max_val);
      for (int i = 0; i < 8; i++) {
        uint32_t vId = v_to_i(vp[0] + ((i & 4) >> 2), vp[1] + ((i & 2) >> 1), vp[2] + (i & 1), v_size, h.brick_pad);
        uint32_t q = (S.values[v_off + vId / vals_per_int] >> (bits * (vId % vals_per_int))) & max_val;
        values[i] = -d_max + mult * T(q);
        vmin = mn(vmin, values[i]);
      }
    }
    return vmin;
  }

  // BVH2Common.cpp:274-631
  void LocalSurfaceIntersection(V3<T> ray_dir, uint32_t instId, uint32_t geomId, const T values[8], uint32_t nodeId,
                                uint32_t primId, T d, T qNear, T qFar, V2<T> fNearFar, V3<T> start_q, Hit<T>* pHit) const {
    (void)nodeId;
    const T EPS = T(1e-6f);
    const unsigned mode = O.preset.sdf_node_intersect;
    T d_inv = T(1) / d;
    T t = qNear;
    bool hit = false;
    unsigned iter = 0;
    T start_sign = 1;
    if (cnt) cnt->root_finds++;

    T start_dist = eval_dist_trilinear(values, start_q + t * ray_dir);
    if (O.preset.representation_mode == LRT_REPRESENTATION_MODE_SURFACE) start_sign = sgn(start_dist);
    start_dist *= start_sign;

    if (start_dist <= EPS || mode == LRT_SDF_NODE_INTERSECT_BBOX) {
      hit = true;
    } else if (mode == LRT_SDF_NODE_INTERSECT_ST) {
      const unsigned ST_max_iters = 256;
      T dist = start_dist;
      while (t < qFar && dist > EPS && iter < ST_max_iters) {
        t += dist * d_inv;
        dist = start_sign * eval_dist_trilinear(values, start_q + t * ray_dir);
        iter++;
      }
      hit = (dist <= EPS);
    } else {
      // JCGT 2022 "Ray Tracing of Signed Distance Function Grids" coefficients (BVH2Common.cpp:322-361)
      T s000 = values[0] * d_inv, s001 = values[1] * d_inv, s010 = values[2] * d_inv, s011 = values[3] * d_inv;
      T s100 = values[4] * d_inv, s101 = values[5] * d_inv, s110 = values[6] * d_inv, s111 = values[7] * d_inv;
      T a = s101 - s001;
      T k0 = s000;
      T k1 = s100 - s000;
      T k2 = s010 - s000;
      T k3 = s110 - s010 - k1;
      T k4 = k0 - s001;
      T k5 = k1 - a;
      T k6 = k2 - (s011 - s001);
      T k7 = k3 - (s111 - s011 - a);
      V3<T> o = start_q, d3 = ray_dir;
      T m0 = o.x * o.y;
      T m1 = d3.x * d3.y;
      T m2 = o.x * d3.y + o.y * d3.x;
      T m3 = k5 * o.z - k1;
      T m4 = k6 * o.z - k2;
      T m5 = k7 * o.z - k3;
      T c0 = (k4 * o.z - k0) + o.x * m3 + o.y * m4 + m0 * m5;
      T c1 = d3.x * m3 + d3.y * m4 + m2 * m5 + d3.z * (k4 + k5 * o.x + k6 * o.y + k7 * m0);
      T c2 = m1 * m5 + d3.z * (k5 * d3.x + k6 * d3.y + k7 * m2);
      T c3 = k7 * m1 * d3.z;

      if (mode == LRT_SDF_NODE_INTERSECT_ANALYTIC) {  // BVH2Common.cpp:366-436
        T x1 = 1000, x2 = 1000, x3 = 1000;
        if ((double)std::abs(c3) > 1e-2) {
          T A_ = c2 / c3, B_ = c1 / c3, C_ = c0 / c3;
          T Q = (A_ * A_ - 3 * B_) / 9;
          T R = (2 * A_ * A_ - 9 * A_ * B_ + 27 * C_) / 54;
          T Q3 = Q * Q * Q;
          if (R * R < Q3) {
            T theta = std::acos(R / std::sqrt(Q3));
            x1 = -2 * std::sqrt(Q) * std::cos(theta / 3) - A_ / 3;
            x2 = T(-2 * std::sqrt(Q) * std::cos((theta + 2 * M_PI) / 3) - A_ / 3);
            x3 = T(-2 * std::sqrt(Q) * std::cos((theta - 2 * M_PI) / 3) - A_ / 3);
          } else {
            T Am = -sgn(R) * std::pow(std::abs(R) + std::sqrt(R * R - Q3), T(1.0f / 3.0f));
            T Bm = std::abs(Am) > EPS ? Q / Am : 0;
            x1 = Am + Bm - A_ / 3;
          }
        } else if ((double)std::abs(c2) > 1e-4) {
          T qa = c2, qb = c1, qc = c0;
          T D = qb * qb - 4 * qa * qc;
          if (D > 0) {
            T q = T(-0.5f) * (qb + sgn(qb) * std::sqrt(D));
            x1 = q / qa;
            if (std::abs(q) > EPS) x2 = qc / q;
          }
        } else if (std::abs(c1) > EPS) {
          x1 = -c0 / c1;
        }
        x1 = x1 < 0 ? 1000 : x1;
        x2 = x2 < 0 ? 1000 : x2;
        x3 = x3 < 0 ? 1000 : x3;
        t = mn(x1, mn(x2, x3));
        hit = (t >= 0 && t <= qFar);
      } else if (mode == LRT_SDF_NODE_INTERSECT_NEWTON) {  // BVH2Common.cpp:437-522
        T qa = 3 * c3, qb = 2 * c2, qc = c1;
        T t0 = 0, t1 = qFar, t2 = qFar, t3 = qFar;
        T D = qb * qb - 4 * qa * qc;
        if (D >= 0) {
          T q = T(-0.5f) * (qb + sgn(qb) * std::sqrt(D));
          t1 = std::abs(qa) > EPS ? q / qa : t0;
          t2 = std::abs(q) > EPS ? qc / q : qFar;
          T tmp = mn(t1, t2);
          t2 = mx(t1, t2);
          t1 = tmp;
          t1 = clampv(t1, t0, t3);
          t2 = clampv(t2, t0, t3);
        }
        bool s0 = c0 > 0;
        bool s1 = (c0 + t1 * (c1 + t1 * (c2 + t1 * c3))) > 0;
        bool s2 = (c0 + t2 * (c1 + t2 * (c2 + t2 * c3))) > 0;
        bool s3 = (c0 + t3 * (c1 + t3 * (c2 + t3 * c3))) > 0;
        T nwt_min = t0, nwt_max = t0;
        if (s0 != s1) { nwt_min = t0; nwt_max = t1; }
        else if (s1 != s2) { nwt_min = t1; nwt_max = t2; }
        else if (s2 != s3) { nwt_min = t2; nwt_max = t3; }
        if (nwt_min < nwt_max) {
          const unsigned max_iters = 10;
          unsigned it = 0;
          T rtn = T(0.5f) * (nwt_min + nwt_max);
          T f = 1000;
          while (it < max_iters && std::abs(f) >= EPS) {
            f = c0 + rtn * (c1 + rtn * (c2 + rtn * c3));
            T df = c1 + rtn * (2 * c2 + rtn * 3 * c3);
            T dx = f / (df + sgn(df) * T(1e-9f));
            rtn -= dx;
            it++;
          }
          if (cnt) cnt->root_iters += it;  // the reference keeps Newton's count local
          t = rtn;
          hit = (t >= 0 && t <= qFar && std::abs(f) < EPS);
        } else {
          hit = false;
        }
      } else {  // IT, BVH2Common.cpp:523-550
        const unsigned IT_max_iters = 256;
        const T k = 2;
        T e = T(0.1f) * qFar;
        T t_max = std::abs(c3) < EPS ? T(1e6) : -c2 / (3 * c3);
        T df_max = 3 * c3 * t_max * t_max + 2 * c2 * t_max + c1;
        T dist = start_dist;
        while (t < qFar && dist > EPS && iter < IT_max_iters) {
          T df_1 = 3 * c3 * t * t + 2 * c2 * t + c1;
          T df_2 = 3 * c3 * (t + e) * (t + e) + 2 * c2 * (t + e) + c1;
          T L = (t_max > t && t_max < t + e) ? mx(df_max, mx(df_1, df_2)) : mx(df_1, df_2);
          L = mx(L, EPS);
          T s = mn((dist * d_inv) / L, e);
          t += s;
          e = k * s;
          dist = start_sign * eval_dist_trilinear(values, start_q + t * ray_dir);
          iter++;
        }
        hit = (dist <= EPS);
      }
    }

    T tReal = fNearFar.x + d * t;
    last_t = t;
    last_hit = (t <= qFar && hit);
    last_iter = iter;
    if (cnt) cnt->root_iters += iter;
    if (t <= qFar && hit && tReal < pHit->t) {  // BVH2Common.cpp:572-602
      V3<T> norm{0, 0, 1};
      if (need_normal(O)) {
        V3<T> p0 = start_q + t * ray_dir;
        const T h = T(0.001f);
        T ddx = (eval_dist_trilinear(values, p0 + V3<T>(h, 0, 0)) - eval_dist_trilinear(values, p0 + V3<T>(-h, 0, 0))) / (2 * h);
        T ddy = (eval_dist_trilinear(values, p0 + V3<T>(0, h, 0)) - eval_dist_trilinear(values, p0 + V3<T>(0, -h, 0))) / (2 * h);
        T ddz = (eval_dist_trilinear(values, p0 + V3<T>(0, 0, h)) - eval_dist_trilinear(values, p0 + V3<T>(0, 0, -h))) / (2 * h);
        norm = start_sign * normalize(V3<T>(ddx, ddy, ddz));  // identity instance transform
      }
      V2<T> en = encode_normal(norm);
      pHit->t = tReal;
      pHit->primId = primId;
      pHit->instId = instId;
      pHit->geomId = geomId | (S.type << LRT_SH_TYPE);
      pHit->coords[0] = 0;
      pHit->coords[1] = 0;
      pHit->coords[2] = en.x;
      pHit->coords[3] = en.y;
      if (O.preset.render_mode == LRT_RENDER_MODE_ST_ITERATIONS) pHit->primId = iter;
    }
  }

  // BVH2Common.cpp:689-694
  static T step01(T e0, T e1, T x) { return clampv((x - e0) / (e1 - e0), T(0), T(1)); }

  // BVH2Common.cpp:764-984
  void OctreeBrickIntersect(V3<T> ray_pos, V3<T> ray_dir, T tNear, uint32_t instId, uint32_t geomId, uint32_t bvhNodeId,
                            Hit<T>* pHit) const {
    T values[8];
    const SdfSBSHeader& h = S.header;
    uint32_t primId = bvhNodeId;
    uint32_t nodeId = primId;
    uint32_t v_size = h.brick_size + 2 * h.brick_pad + 1;
    const SdfSBSNode& N = S.nodes[nodeId];
    T px = T(N.pos_xy >> 16), py = T(N.pos_xy & 0x0000FFFF), pz = T(N.pos_z_lod_size >> 16), sz = T(N.pos_z_lod_size & 0x0000FFFF);
    T sz_inv = T(2) / sz;
    T d = T(2) / (sz * T(h.brick_size));
    V3<T> brick_min_pos = V3<T>(-1, -1, -1) + sz_inv * V3<T>(px, py, pz);
    V3<T> brick_max_pos = brick_min_pos + sz_inv * V3<T>(1, 1, 1);
    const T qscale = T(0.5f) * sz * T(h.brick_size);
    if (cnt) cnt->bricks++;

    V2<T> bnf = RayBoxIntersection2(ray_pos, SafeInverse(ray_dir), brick_min_pos, brick_max_pos);
    T old_t = pHit->t;
    // termination cap, identical to the device code (never reached by a ray
    // that advances: a brick has at most 3*brick_size+1 voxels along a line)
    const uint32_t max_steps = 16u * (h.brick_size + 1u) + 16u;
    uint32_t steps = 0;
    while (bnf.x < bnf.y && pHit->t == old_t && steps++ < max_steps) {
      V3<T> hit_pos = ray_pos + bnf.x * ray_dir;
      V3<T> local_pos = (hit_pos - brick_min_pos) * qscale;
      V3<T> voxelPos = vfloor(vclamp(local_pos, T(1e-6f), T(h.brick_size) - T(1e-6f)));
      uint32_t vp[3] = {(uint32_t)voxelPos.x, (uint32_t)voxelPos.y, (uint32_t)voxelPos.z};
      V3<T> min_pos = brick_min_pos + d * voxelPos;
      V3<T> max_pos = min_pos + d * V3<T>(1, 1, 1);
      if (cnt) cnt->voxels++;
      T vmin = load_distance_values(nodeId, vp, v_size, sz_inv, values);
      V2<T> fNearFar = RayBoxIntersection2(ray_pos, SafeInverse(ray_dir), min_pos, max_pos);
      if (tNear < fNearFar.x && vmin <= T(0)) {
        V3<T> start_pos = ray_pos + fNearFar.x * ray_dir;
        V3<T> start_q = (start_pos - min_pos) * qscale;
        T qFar = (fNearFar.y - fNearFar.x) * qscale;
        LocalSurfaceIntersection(ray_dir, instId, geomId, values, nodeId, primId, d, T(0), qFar, fNearFar, start_q, pHit);

        if (layout() == LRT_SBS_LAYOUT_ID32F_IRGB32F_IN && O.preset.normal_mode == LRT_NORMAL_MODE_SDF_SMOOTHED &&
            need_normal(O) && pHit->t != old_t) {  // BVH2Common.cpp:838-906
          const T beta = T(0.5f);
          V3<T> dp = start_q + ((pHit->t - fNearFar.x) / d) * ray_dir;
          V3<T> normals[8];
          T values_n[8];
          V3<T> nmq;
          nmq.x = dp.x < T(0.5f) ? step01(-beta, beta, dp.x) : step01(-beta, beta, dp.x - T(1));
          nmq.y = dp.y < T(0.5f) ? step01(-beta, beta, dp.y) : step01(-beta, beta, dp.y - T(1));
          nmq.z = dp.z < T(0.5f) ? step01(-beta, beta, dp.z) : step01(-beta, beta, dp.z - T(1));
          int vp0[3] = {dp.x < T(0.5f) ? (int)vp[0] - 1 : (int)vp[0], dp.y < T(0.5f) ? (int)vp[1] - 1 : (int)vp[1],
                        dp.z < T(0.5f) ? (int)vp[2] - 1 : (int)vp[2]};
          const int bs = (int)h.brick_size;
          for (int i = 0; i < 8; i++) {
            int VI[3] = {vp0[0] + ((i & 4) >> 2), vp0[1] + ((i & 2) >> 1), vp0[2] + (i & 1)};
            int BP[3];
            for (int a = 0; a < 3; a++) BP[a] = VI[a] >= 0 ? (VI[a] < bs ? 0 : 1) : -1;
            uint32_t neighborId = 9 * (BP[0] + 1) + 3 * (BP[1] + 1) + (BP[2] + 1);
            V3<T> dVoxelPos = V3<T>(T(VI[0]), T(VI[1]), T(VI[2])) - voxelPos;
            uint32_t nb_node;
            uint32_t nvp[3];
            if (neighborId == 13) {
              nb_node = nodeId;
              for (int a = 0; a < 3; a++) nvp[a] = (uint32_t)VI[a];
            } else {
              uint32_t nb_off = v_size * v_size * v_size + 8;
              nb_node = S.values[N.data_offset + nb_off + neighborId];
              for (int a = 0; a < 3; a++) nvp[a] = (uint32_t)(VI[a] - bs * BP[a]);
            }
            normals[i] = V3<T>(0, 0, 0);
            if (nb_node != LRT_INVALID_IDX) {
              load_distance_values(nb_node, nvp, v_size, sz_inv, values_n);
              normals[i] = normalize(eval_dist_trilinear_diff(values_n, dp - dVoxelPos));
            }
          }
          V3<T> sm = (1 - nmq.x) * (1 - nmq.y) * (1 - nmq.z) * normals[0] + (1 - nmq.x) * (1 - nmq.y) * (nmq.z) * normals[1] +
                     (1 - nmq.x) * (nmq.y) * (1 - nmq.z) * normals[2] + (1 - nmq.x) * (nmq.y) * (nmq.z) * normals[3] +
                     (nmq.x) * (1 - nmq.y) * (1 - nmq.z) * normals[4] + (nmq.x) * (1 - nmq.y) * (nmq.z) * normals[5] +
                     (nmq.x) * (nmq.y) * (1 - nmq.z) * normals[6] + (nmq.x) * (nmq.y) * (nmq.z) * normals[7];
          sm = normalize(sm);
          V2<T> en = encode_normal(sm);
          pHit->coords[2] = en.x;
          pHit->coords[3] = en.y;
        }
      }
      bnf.x += mx(T(0), fNearFar.y - bnf.x) + T(1e-6f);
    }

    if (pHit->t < old_t && indexed()) {  // colour for ID32F layouts, BVH2Common.cpp:962-982
      V3<T> pos = ray_pos + pHit->t * ray_dir;
      V3<T> dp = (pos - brick_min_pos) * (T(0.5f) * sz);
      uint32_t t_off = N.data_offset + v_size * v_size * v_size;
      V3<T> col(0, 0, 0);
      for (int i = 0; i < 8; i++) {
        T w = ((i & 4) ? dp.x : 1 - dp.x) * ((i & 2) ? dp.y : 1 - dp.y) * ((i & 1) ? dp.z : 1 - dp.z);
        uint32_t ci = S.values[t_off + i];
        col = col + w * V3<T>(T(S.values_f[ci + 0]), T(S.values_f[ci + 1]), T(S.values_f[ci + 2]));
      }
      col = V3<T>(clampv(std::floor(T(255) * col.x + T(0.5f)), T(0), T(255)), clampv(std::floor(T(255) * col.y + T(0.5f)), T(0), T(255)),
                  clampv(std::floor(T(255) * col.z + T(0.5f)), T(0), T(255)));
      pHit->coords[0] = col.x + col.y / T(256);
      pHit->coords[1] = col.z;
    }
  }

  // BVH2Common.cpp:3407-3488 (stack traversal, stop at first SDF-brick hit)
  void BVH2TraverseF32(V3<T> ray_pos, V3<T> ray_dir, T tNear, uint32_t instId, uint32_t geomId, bool stopOnFirstHit,
                       Hit<T>* pHit) const {
    uint32_t stack[LRT_STACK_SIZE];
    int top = 0;
    uint32_t leftNodeOffset = 0;
    bool hitFound = false;
    const V3<T> rayDirInv = SafeInverse(ray_dir);
    while (top >= 0 && !hitFound) {
      while (top >= 0 && ((leftNodeOffset & LRT_LEAF_BIT) == 0)) {
        const BVHNodePair& fat = S.pairs[leftNodeOffset];
        if (cnt) cnt->node_pairs++;
        const uint32_t n0 = fat.left.leftOffset, n1 = fat.right.leftOffset;
        V2<T> tm0 = RayBoxIntersection2(ray_pos, rayDirInv, V3<T>(fat.left.boxMin[0], fat.left.boxMin[1], fat.left.boxMin[2]),
                                        V3<T>(fat.left.boxMax[0], fat.left.boxMax[1], fat.left.boxMax[2]));
        V2<T> tm1 = RayBoxIntersection2(ray_pos, rayDirInv, V3<T>(fat.right.boxMin[0], fat.right.boxMin[1], fat.right.boxMin[2]),
                                        V3<T>(fat.right.boxMax[0], fat.right.boxMax[1], fat.right.boxMax[2]));
        // CPU build keeps the RF-grid type enabled, so culling by hit distance
        // applies only to any-hit queries (BVH2Common.cpp:3437-3440).
        const bool hit0 = (tm0.x <= tm0.y) && (tm0.y >= tNear) && (tm0.x <= pHit->t || !stopOnFirstHit);
        const bool hit1 = (tm1.x <= tm1.y) && (tm1.y >= tNear) && (tm1.x <= pHit->t || !stopOnFirstHit);
        leftNodeOffset = hit0 ? n0 : n1;
        if (hit0 && hit1) {
          leftNodeOffset = (tm0.x <= tm1.x) ? n0 : n1;
          stack[top] = (tm0.x <= tm1.x) ? n1 : n0;
          top++;
        }
        if (!hit0 && !hit1) {
          top--;
          leftNodeOffset = stack[std::max(top, 0)];
        }
      }
      if (top >= 0 && leftNodeOffset != 0xFFFFFFFFu) {
        const uint32_t aabbId = leftNodeOffset & LRT_START_MASK;
        const T tNearSdf = mx(tNear, T(0.1f));  // SDF_BIAS, BVH2Common.cpp:3472-3473
        T tPrev = pHit->t;
        OctreeBrickIntersect(ray_pos, ray_dir, tNearSdf, instId, geomId, aabbId, pHit);
        const bool tagged = pHit->t < tPrev;  // GeomDataSdfBrick::Intersect, BVH2Common.h:596-618
        hitFound = tagged;                     // SDF tags: first hit is closest (BVH2Common.cpp:3402-3405)
      }
      top--;
      leftNodeOffset = stack[std::max(top, 0)];
    }
  }

  // BVH2Common.cpp:3490-3559 for one identity instance (1-leaf TLAS, BVH2Common_host.cpp:1430-1437)
  Hit<T> RayQuery_NearestHit(V4<T> posAndNear, V4<T> dirAndFar) const {
    bool stopOnFirstHit = (dirAndFar.w <= T(0));
    if (stopOnFirstHit) dirAndFar.w *= T(-1);
    V3<T> pos{posAndNear.x, posAndNear.y, posAndNear.z}, dir{dirAndFar.x, dirAndFar.y, dirAndFar.z};
    const V3<T> rayDirInv = SafeInverse(dir);
    Hit<T> hit;
    hit.t = dirAndFar.w;
    hit.primId = hit.instId = hit.geomId = 0xFFFFFFFFu;
    hit.coords[0] = 1; hit.coords[1] = 0; hit.coords[2] = 0; hit.coords[3] = 0;
    if (cnt) cnt->rays++;
    V2<T> bh = RayBoxIntersection2(pos, rayDirInv, V3<T>(-1, -1, -1), V3<T>(1, 1, 1));
    bool intersects = (bh.x <= bh.y) && (bh.y > posAndNear.w) && (bh.x < hit.t);
    if (intersects) BVH2TraverseF32(pos, dir, posAndNear.w, 0, 0, stopOnFirstHit, &hit);
    return hit;
  }
};

// ---------------------------------------------------------------------------
// Camera and pixel order (Renderer/eye_ray.cpp, render_common.h)
// ---------------------------------------------------------------------------

// eye_ray.cpp:537-555 SuperBlockIndex2DOpt
static inline uint32_t SuperBlockIndex2DOpt(uint32_t tidX, uint32_t tidY, uint32_t w) {
  const uint32_t inBlockIdX = tidX & 3, inBlockIdY = tidY & 3;
  const uint32_t localIndex = inBlockIdY * 4 + inBlockIdX;
  const uint32_t wBlocks = w >> 2;
  const uint32_t blockX = tidX >> 2, blockY = tidY >> 2;
  const uint32_t inHX = blockX & 1, inHY = blockY & 1;
  const uint32_t localIndexH = inHY * 2 + inHX;
  const uint32_t wBlocksH = wBlocks >> 1;
  const uint32_t bHX = blockX >> 1, bHY = blockY >> 1;
  return (bHX + bHY * wBlocksH) * 64 + localIndexH * 16 + localIndex;
}

// eye_ray_host.cpp:43-50 + eye_ray.cpp:557-575: packed (y<<16)|x per tid
static inline std::vector<uint32_t> PackXY(uint32_t W, uint32_t H) {
  uint32_t pw = 8 * ((W + 7) / 8), ph = 8 * ((H + 7) / 8);
  std::vector<uint32_t> p(pw * ph);
  for (uint32_t y = 0; y < ph; y++)
    for (uint32_t x = 0; x < pw; x++) p[SuperBlockIndex2DOpt(x, y, pw)] = ((y << 16) & 0xFFFF0000u) | (x & 0xFFFFu);
  return p;
}

// eye_ray.cpp:34-81 (PCG3D, JCGT 2020 "Hash Functions for GPU Rendering")
static inline void pcg3d(uint32_t& x, uint32_t& y, uint32_t& z) {
  x = x * 1664525u + 1013904223u;
  y = y * 1664525u + 1013904223u;
  z = z * 1664525u + 1013904223u;
  x += y * z; y += z * x; z += x * y;
  x ^= x >> 16u; y ^= y >> 16u; z ^= z >> 16u;
  x += y * z; y += z * x; z += x * y;
}
// eye_ray.cpp:34-81 (PCG3D keyed by pixel, sample and seed)
static inline V2<float> rand2(uint32_t x, uint32_t y, uint32_t iter, uint32_t seed) {
  x = x + 1233u * (iter + seed) % 171u;
  y = y + 453u * (iter + seed) % 765u;
  uint32_t a = x, b = y, c = x ^ y;
  pcg3d(a, b, c);
  const float s = 1.0f / float(0xffffffffu);
  return {float(a) * s, float(b) * s};
}

struct Camera {  // eye_ray_host.cpp:59-65
  M4<float> projInv, worldViewInv;
  static Camera from(const float worldView[16], const float proj[16]) {
    Camera c;
    inverse4x4_via_double(proj, c.projInv.m);
    inverse4x4_via_double(worldView, c.worldViewInv.m);
    return c;
  }
};

// eye_ray.cpp:146-160 with render_common.h:25-31, 45-55
template <typename T>
static inline void InitEyeRay(const Camera& cam, uint32_t x, uint32_t y, V2<float> d, uint32_t W, uint32_t H, V4<T>* posAndNear,
                              V4<T>* dirAndFar) {
  M4<T> pI, wI;
  for (int i = 0; i < 16; i++) { pI.m[i] = T(cam.projInv.m[i]); wI.m[i] = T(cam.worldViewInv.m[i]); }
  T fx = (T(x) + T(d.x)) / T(W);
  T fy = (T(y) + T(d.y)) / T(H);
  V4<T> pos{T(2) * fx - T(1), T(-2) * fy + T(1), T(1), T(1)};
  pos = mul(pI, pos);
  pos = V4<T>{pos.x / pos.w, pos.y / pos.w, pos.z / pos.w, pos.w / pos.w};
  V3<T> rayDir = normalize(V3<T>(pos.x, pos.y, pos.z));
  V3<T> rayPos(0, 0, 0);
  V3<T> p1 = mul4x3(wI, rayPos);
  V3<T> p2 = mul4x3(wI, rayPos + T(100) * rayDir);
  V3<T> diff = p2 - p1;
  *posAndNear = V4<T>{p1.x, p1.y, p1.z, T(0)};
  V3<T> nd = normalize(diff);
  *dirAndFar = V4<T>{nd.x, nd.y, nd.z, T(1e9f)};
}

// eye_ray.cpp:97-101 jitter for sample i
static inline V2<float> sample_offset(const MultiRenderPreset& p, uint32_t x, uint32_t y, uint32_t i, uint32_t seed) {
  if (p.ray_gen_mode == LRT_RAY_GEN_MODE_RANDOM) return rand2(x, y, i + seed % 2, seed);
  uint32_t spp_sqrt = uint32_t(std::sqrt((double)p.spp));
  float i_spp_sqrt = 1.0f / float(spp_sqrt);
  return {i_spp_sqrt * float(double(i / spp_sqrt) + 0.5), i_spp_sqrt * float(double(i % spp_sqrt) + 0.5)};
}

static const uint32_t kPalette[20] = {0xffe6194b, 0xff3cb44b, 0xffffe119, 0xff0082c8, 0xfff58231, 0xff911eb4, 0xff46f0f0,
                                      0xfff032e6, 0xffd2f53c, 0xfffabebe, 0xff008080, 0xffe6beff, 0xffaa6e28, 0xfffffac8,
                                      0xff800000, 0xffaaffc3, 0xff808000, 0xffffd8b1, 0xff000080, 0xff808080};

// Phong's specular power x^32 as five squarings (the reference calls pow;
// which overload LiteMath resolves it to is not known here)
template <typename T> static inline T pow32(T x) {
  for (int i = 0; i < 5; i++) x = x * x;
  return x;
}

static inline bool needs_shadow(uint32_t mode) { return mode == LRT_RENDER_MODE_PHONG || mode == LRT_RENDER_MODE_PHONG_NO_TEX; }

template <typename T> static inline V3<T> light_v(const float v[3]) { return V3<T>(T(v[0]), T(v[1]), T(v[2])); }

// Phong's shadow ray toward light l (eye_ray.cpp:443-468): from just before
// the hit, along a direct light's direction or toward a point light.
template <typename T>
static inline bool shadow_ray(const LrtLight& l, V3<T> ray_pos, V3<T> ray_dir, T t, V3<T>& origin, V3<T>& dir) {
  if (l.type == LRT_LIGHT_TYPE_AMBIENT) return false;
  origin = ray_pos + (t - T(1e-6f)) * ray_dir;
  if (l.type == LRT_LIGHT_TYPE_DIRECT) {
    dir = light_v<T>(l.space);
  } else {
    V3<T> d = light_v<T>(l.space) - origin;
    T len = length(d);
    dir = V3<T>(d.x / len, d.y / len, d.z / len);
  }
  return true;
}

// eye_ray.cpp:162-489, the render modes that apply to SBS geometry with the
// scene's lights. lit: bit i set if Phong's shadow ray toward light i hit
// nothing. rayPos: the eye ray's origin.
template <typename T>
static inline V4<T> shade(const MultiRenderPreset& p, const Hit<T>& hit, V4<T> rayPos, V4<T> rayDir, const std::vector<LrtLight>& lights,
                          uint32_t lit) {
  if (hit.primId == 0xFFFFFFFFu) return {0, 0, 0, 1};
  T z = hit.t;
  const T z_near = T(0.1f), z_far = T(10);
  unsigned type = hit.geomId >> LRT_SH_TYPE;
  unsigned geomId = hit.geomId & LRT_GEOM_ID_MASK;
  V3<T> norm = decode_normal(V2<T>{hit.coords[2], hit.coords[3]});
  T ns = sgn(dot(T(-1) * V3<T>(rayDir.x, rayDir.y, rayDir.z), norm));
  norm = norm * ns;
  auto pal = [](uint32_t c) {
    return V4<T>{T(c & 0xFF) * T(1.0f / 255.0f), T((c >> 8) & 0xFF) * T(1.0f / 255.0f), T((c >> 16) & 0xFF) * T(1.0f / 255.0f),
                 T((c >> 24) & 0xFF) * T(1.0f / 255.0f)};
  };
  switch (p.render_mode) {
    case LRT_RENDER_MODE_MASK: return {1, 1, 1, 1};
    case LRT_RENDER_MODE_DEPTH: { T d = (1 / z - 1 / z_near) / (1 / z_far - 1 / z_near); return {d, d, d, 1}; }
    case LRT_RENDER_MODE_LINEAR_DEPTH: { T d = ((z - z_near) / (z_far - z_near)); return {d, d, d, 1}; }
    case LRT_RENDER_MODE_INVERSE_LINEAR_DEPTH: { T d = 1 - ((z - z_near) / (z_far - z_near)); return {d, d, d, 1}; }
    case LRT_RENDER_MODE_PRIMITIVE: return pal(kPalette[hit.primId % 20]);
    case LRT_RENDER_MODE_TYPE: return pal(kPalette[type % 20]);
    case LRT_RENDER_MODE_GEOM: return pal(kPalette[geomId % 20]);
    case LRT_RENDER_MODE_NORMAL: return {std::abs(norm.x), std::abs(norm.y), std::abs(norm.z), 1};
    case LRT_RENDER_MODE_ST_ITERATIONS: {
      if (hit.primId == 0) return {T(0.5), T(0.5), T(0.5), 1};
      if (hit.primId <= 64) {  // q*(0,0,1,1) + (1-q)*(0,1,0,1)
        T q = clampv(T(hit.primId) / 64, T(0), T(1));
        return {0, (1 - q), q, q + (1 - q)};
      }
      T q = clampv(T(hit.primId - 64) / (256 - 64), T(0), T(1));  // q*(1,0,0,1) + (1-q)*(0,0,1,1)
      return {q, 0, (1 - q), q + (1 - q)};
    }
    case LRT_RENDER_MODE_DIFFUSE: {
      if (type == LRT_TYPE_SDF_SBS_COL)
        return {std::round(hit.coords[0]) / T(255), hit.coords[0] - std::floor(hit.coords[0]), std::round(hit.coords[1]) / T(255), 1};
      return {0, 1, 1, 1};  // default textured material samples the default (0,1,1,1) texture
    }
    case LRT_RENDER_MODE_BARYCENTRIC: return {hit.coords[0], hit.coords[1], 1 - hit.coords[0] - hit.coords[1], 1};
    case LRT_RENDER_MODE_TEX_COORDS: return {hit.coords[0], hit.coords[1], 0, 1};  // tc = coords[0..1] for SBS
    case LRT_RENDER_MODE_LAMBERT_NO_TEX:
    case LRT_RENDER_MODE_LAMBERT:
    case LRT_RENDER_MODE_PHONG:
    case LRT_RENDER_MODE_PHONG_NO_TEX: {  // eye_ray.cpp:376-483
      const uint32_t m = p.render_mode;
      V3<T> col(1, 1, 1);
      if (m == LRT_RENDER_MODE_LAMBERT || m == LRT_RENDER_MODE_PHONG)
        col = type == LRT_TYPE_SDF_SBS_COL
                  ? V3<T>(std::round(hit.coords[0]) / T(255), hit.coords[0] - std::floor(hit.coords[0]), std::round(hit.coords[1]) / T(255))
                  : V3<T>(0, 1, 1);  // default material: the (0,1,1,1) texture
      const bool phong = m == LRT_RENDER_MODE_PHONG || m == LRT_RENDER_MODE_PHONG_NO_TEX;
      const V3<T> rp(rayPos.x, rayPos.y, rayPos.z), rd(rayDir.x, rayDir.y, rayDir.z);
      V3<T> c(0, 0, 0);
      for (size_t i = 0; i < lights.size(); i++) {
        const LrtLight& l = lights[i];
        const V3<T> lc = light_v<T>(l.color);
        if (l.type == LRT_LIGHT_TYPE_AMBIENT) {
          c = c + lc * col;
          continue;
        }
        if (!phong) {
          if (l.type == LRT_LIGHT_TYPE_DIRECT) {
            c = c + (lc * col) * mx(T(0), dot(norm, light_v<T>(l.space)));
          } else {
            V3<T> surf = rp + hit.t * rd;
            V3<T> ld = light_v<T>(l.space) - surf;
            T len = length(ld);
            ld = V3<T>(ld.x / len, ld.y / len, ld.z / len);
            V3<T> e = (lc * col) * mx(T(0), dot(norm, ld));
            c = c + V3<T>(e.x / (len * len), e.y / (len * len), e.z / (len * len));
          }
          continue;
        }
        V3<T> o, ld;
        shadow_ray<T>(l, rp, rd, hit.t, o, ld);
        if (l.type == LRT_LIGHT_TYPE_DIRECT) c = c + (lc * col) * mx(T(0), dot(norm, ld));
        const T Kd = T(0.25f), Ks = T(0.25f);
        V3<T> refl = T(-1) * ld + (T(2) * dot(norm, ld)) * norm;
        T e = Kd * mx(T(0), dot(norm, ld)) + Ks * pow32(mx(T(0), dot(norm, refl)));
        T shadow = ((lit >> i) & 1u) ? T(1) : T(0);
        c = c + ((shadow * lc) * e) * col;
      }
      return {c.x, c.y, c.z, 1};
    }
    default: return {1, 0, 1, 1};
  }
}

}  // namespace orc

// oracle_dr.cpp - CPU restatement of LiteRT's differentiable SBS renderer
// (diff_render/BVH2DR.cpp, diff_render/MultiRendererDR.cpp).
//
// TEST INFRASTRUCTURE ONLY (see lrt_oracle.h). Covers what RenderDR does for
// the Newton root find that OptimizeFixedStructure forces
// (MultiRendererDR.cpp:324) and every render mode (DIFFUSE, LAMBERT with the
// renderer's default lights, MASK, LINEAR_DEPTH, NORMAL), with the normal and
// its derivatives from the hit voxel (ID32F_IRGB32F) or smoothed over the 8
// voxels around the hit (ID32F_IRGB32F_IN). Gradients are accumulated in
// double whatever the precision of the per-ray arithmetic (the reference sums
// per-thread float buffers in a thread-dependent order).
#include <omp.h>
#include <cmath>
#include <vector>
#include "lrt_oracle.h"
#include "oracle_scene.hpp"

using namespace orc;


namespace {

constexpr uint32_t kInvalid = 0xFFFFFFFFu;
constexpr float kZNear = 0.1f, kZFar = 10.0f;  // MultiRendererDR.cpp:22-23

template <typename T> struct HitDR {  // CRT_HitDR, DR_common.h:47-57
  T t;
  uint32_t primId, instId, geomId;
  V3<T> color;
  T sdf;
  V3<T> normal;
  uint32_t check_vox_border;
};

template <typename T> struct PDDist {
  V3<T> dDiffuse;
  uint32_t index;
  V3<T> dNorm;
  T dDist;
};

template <typename T> struct Payload {  // RayDiffPayload, DR_common.h:59-72
  HitDR<T> missed_hit, missed_hit_candidate;
  uint32_t missed_indices[8];
  T missed_dSDF_dtheta[8];
  uint32_t dSc_index[8];
  T dSc_value[8];
  PDDist<T> dSd[64];
};

// 3x3 matrix as three rows; m * v takes the row dot products (x, y, z left to right)
template <typename T> struct M3 {
  V3<T> r[3];
};
template <typename T> static inline V3<T> mul(const M3<T>& m, V3<T> v) { return {dot(m.r[0], v), dot(m.r[1], v), dot(m.r[2], v)}; }

// BVH2DR.cpp:339-404: Jacobian of v / |v|
template <typename T> static M3<T> normalize_diff(V3<T> v) {
  T a = dot(v, v);
  T b = T(1) / (a * std::sqrt(a));  // the reference divides 1.0 (double) and rounds: the same value as the T division
  return {{b * V3<T>(v.y * v.y + v.z * v.z, -v.x * v.y, -v.x * v.z), b * V3<T>(-v.x * v.y, v.x * v.x + v.z * v.z, -v.y * v.z),
           b * V3<T>(-v.x * v.z, -v.y * v.z, v.x * v.x + v.y * v.y)}};
}

// BVH2DR.cpp:406-530: Hessian of the trilinear distance (zero diagonal)
template <typename T> static M3<T> trilinear_ddp_ddp(const T v[8], V3<T> dp) {
  T xy = (1 - dp.z) * v[0] + (dp.z) * v[1] + -(1 - dp.z) * v[2] + -(dp.z) * v[3] + -(1 - dp.z) * v[4] + -(dp.z) * v[5] +
         (1 - dp.z) * v[6] + (dp.z) * v[7];
  T xz = (1 - dp.y) * v[0] + -(1 - dp.y) * v[1] + (dp.y) * v[2] + -(dp.y) * v[3] + -(1 - dp.y) * v[4] + (1 - dp.y) * v[5] +
         -(dp.y) * v[6] + (dp.y) * v[7];
  T yz = (1 - dp.x) * v[0] + -(1 - dp.x) * v[1] + -(1 - dp.x) * v[2] + (1 - dp.x) * v[3] + (dp.x) * v[4] + -(dp.x) * v[5] +
         -(dp.x) * v[6] + (dp.x) * v[7];
  return {{V3<T>(0, xy, xz), V3<T>(xy, 0, yz), V3<T>(xz, yz, 0)}};
}

// BVH2DR.cpp:531-611: d(grad f)/d(values), rows x, y, z of 8
template <typename T> static void trilinear_ddp_dvalues(V3<T> dp, T m[24]) {
  m[0] = -(1 - dp.y) * (1 - dp.z); m[1] = -(1 - dp.y) * (dp.z); m[2] = -(dp.y) * (1 - dp.z); m[3] = -(dp.y) * (dp.z);
  m[4] = (1 - dp.y) * (1 - dp.z);  m[5] = (1 - dp.y) * (dp.z);  m[6] = (dp.y) * (1 - dp.z);  m[7] = (dp.y) * (dp.z);
  m[8] = -(1 - dp.x) * (1 - dp.z); m[9] = -(1 - dp.x) * (dp.z); m[10] = (1 - dp.x) * (1 - dp.z); m[11] = (1 - dp.x) * (dp.z);
  m[12] = -(dp.x) * (1 - dp.z);    m[13] = -(dp.x) * (dp.z);    m[14] = (dp.x) * (1 - dp.z);     m[15] = (dp.x) * (dp.z);
  m[16] = -(1 - dp.x) * (1 - dp.y); m[17] = (1 - dp.x) * (1 - dp.y); m[18] = -(1 - dp.x) * (dp.y); m[19] = (1 - dp.x) * (dp.y);
  m[20] = -(dp.x) * (1 - dp.y);     m[21] = (dp.x) * (1 - dp.y);     m[22] = -(dp.x) * (dp.y);     m[23] = (dp.x) * (dp.y);
}

// BVH2DR.cpp:612-617: transition from 0 (x = e0) to 1 (x = e1)
template <typename T> static T step01(T e0, T e1, T x) { return clampv((x - e0) / (e1 - e0), T(0), T(1)); }

// BVH2DR.cpp:232-243: trilinear weights
template <typename T> static void trilinear_weights(V3<T> dp, T w[8]) {
  w[0] = (1 - dp.x) * (1 - dp.y) * (1 - dp.z); w[1] = (1 - dp.x) * (1 - dp.y) * (dp.z);
  w[2] = (1 - dp.x) * (dp.y) * (1 - dp.z);     w[3] = (1 - dp.x) * (dp.y) * (dp.z);
  w[4] = (dp.x) * (1 - dp.y) * (1 - dp.z);     w[5] = (dp.x) * (1 - dp.y) * (dp.z);
  w[6] = (dp.x) * (dp.y) * (1 - dp.z);         w[7] = (dp.x) * (dp.y) * (dp.z);
}


template <typename T> struct DRTracer {
  const Scene& S;
  MultiRenderPreset preset;
  int dtdv_mode;  // 0 analytic (implicit function theorem), 1 central differences (BVH2DR.cpp:1075-1090)
  bool need_normal = false;  // render modes LAMBERT and NORMAL (BVH2Common.cpp:20-27)
  DRTracer(const Scene& s, const MultiRenderPreset& p, int m) : S(s), preset(p), dtdv_mode(m) {}

  // Hit normal and, with DNORM_DPOS, dNormal/dvalues (BVH2DR.cpp:1300-1484).
  // values / idx: the hit voxel's corners; vp its position in the brick;
  // start_q, t: the voxel-local ray segment start and root; dvo: first
  // payload slot of the hit voxel's corners; dtdv: dt/dvalues in voxel units.
  void NormalWithGrad(uint32_t ray_flags, V3<T> ray_dir, const T values[8], const uint32_t idx[8], V3<T> start_q, T t, V3<T> vp,
                      uint32_t brick, unsigned dvo, const T dtdv[8], Payload<T>* relax, HitDR<T>* pHit) const {
    const SdfSBSHeader& h = S.header;
    const uint32_t v_size = h.brick_size + 2 * h.brick_pad + 1;
    if ((h.aux_data & LRT_SBS_LAYOUT_MASK) != LRT_SBS_LAYOUT_ID32F_IRGB32F_IN) {
      V3<T> p0 = start_q + t * ray_dir;
      V3<T> g = eval_dist_trilinear_diff(values, p0);
      pHit->normal = normalize(g);
      if (ray_flags & LRT_DR_RAY_FLAG_DNORM_DPOS) {
        M3<T> nd = normalize_diff(g);
        T d1[24];
        trilinear_ddp_dvalues(p0, d1);
        V3<T> t1 = mul(trilinear_ddp_ddp(values, p0), ray_dir);
        for (int i = 0; i < 8; i++) {
          d1[i] += t1.x * dtdv[i];
          d1[8 + i] += t1.y * dtdv[i];
          d1[16 + i] += t1.z * dtdv[i];
        }
        for (int i = 0; i < 8; i++) relax->dSd[dvo + i].dNorm = mul(nd, V3<T>(d1[i], d1[8 + i], d1[16 + i]));
      }
      return;
    }
    // smoothed over the 8 voxels around the hit point (beta 0.5)
    const T beta = T(0.5f);
    const bool dn = (ray_flags & LRT_DR_RAY_FLAG_DNORM_DPOS) != 0;
    V3<T> dp = start_q + t * ray_dir;
    V3<T> nmq(dp.x < T(0.5f) ? step01(-beta, beta, dp.x) : step01(-beta, beta, dp.x - T(1)),
              dp.y < T(0.5f) ? step01(-beta, beta, dp.y) : step01(-beta, beta, dp.y - T(1)),
              dp.z < T(0.5f) ? step01(-beta, beta, dp.z) : step01(-beta, beta, dp.z - T(1)));
    const unsigned k = dvo / 8;
    V3<T> normals[8];
    T wn[8];
    if (dn) trilinear_weights(nmq, wn);
    const int bs = (int)h.brick_size;
    const SdfSBSNode& N = S.nodes[brick];
    int vp0[3] = {dp.x < T(0.5f) ? (int)vp.x - 1 : (int)vp.x, dp.y < T(0.5f) ? (int)vp.y - 1 : (int)vp.y,
                  dp.z < T(0.5f) ? (int)vp.z - 1 : (int)vp.z};
    for (int i = 0; i < 8; i++) {
      int VI[3] = {vp0[0] + ((i & 4) >> 2), vp0[1] + ((i & 2) >> 1), vp0[2] + (i & 1)};
      int BP[3];
      for (int a = 0; a < 3; a++) BP[a] = VI[a] >= 0 ? (VI[a] < bs ? 0 : 1) : -1;
      uint32_t nid = 9 * (BP[0] + 1) + 3 * (BP[1] + 1) + (BP[2] + 1);
      V3<T> dvp = V3<T>(T(VI[0]), T(VI[1]), T(VI[2])) - vp;
      uint32_t nb = brick, nv[3];
      if (nid == 13) {
        for (int a = 0; a < 3; a++) nv[a] = (uint32_t)VI[a];
      } else {
        nb = S.values[N.data_offset + v_size * v_size * v_size + 8 + nid];
        for (int a = 0; a < 3; a++) nv[a] = (uint32_t)(VI[a] - bs * BP[a]);
      }
      normals[i] = V3<T>(0, 0, 0);
      if (nb == LRT_INVALID_IDX) continue;
      T vn[8];
      uint32_t in[8];
      const uint32_t off = S.nodes[nb].data_offset;
      for (int c = 0; c < 8; c++) {
        uint32_t vId = (nv[0] + ((c & 4) >> 2)) * v_size * v_size + (nv[1] + ((c & 2) >> 1)) * v_size + (nv[2] + (c & 1));
        in[c] = S.values[off + vId];
        vn[c] = T(S.values_f[in[c]]);
      }
      V3<T> q = dp - dvp;
      V3<T> g = eval_dist_trilinear_diff(vn, q);
      normals[i] = normalize(g);
      if (!dn) continue;
      for (int c = 0; c < 8; c++) relax->dSd[8 * i + c].index = in[c];
      M3<T> nd = normalize_diff(g);
      T d1[24];
      trilinear_ddp_dvalues(q, d1);
      V3<T> t1 = mul(trilinear_ddp_ddp(vn, q), ray_dir);
      for (int c = 0; c < 8; c++)  // d(n_i)/d(v_i)
        relax->dSd[8 * i + c].dNorm = relax->dSd[8 * i + c].dNorm + wn[i] * mul(nd, V3<T>(d1[c], d1[8 + c], d1[16 + c]));
      for (int c = 0; c < 8; c++)  // d(n_i)/d(v_k) through the hit point
        relax->dSd[8 * k + c].dNorm =
            relax->dSd[8 * k + c].dNorm + wn[i] * mul(nd, V3<T>(t1.x * dtdv[c], t1.y * dtdv[c], t1.z * dtdv[c]));
    }
    T w[8];
    trilinear_weights(nmq, w);
    V3<T> tri = w[0] * normals[0];
    for (int i = 1; i < 8; i++) tri = tri + w[i] * normals[i];
    pHit->normal = normalize(tri);
    if (!dn) return;
    M3<T> nd = normalize_diff(tri);
    // dtrilinear3/dnmq by columns (BVH2DR.cpp:257-337)
    V3<T> cx, cy, cz;
    for (int i = 0; i < 8; i++) {
      T ax = (i & 4) ? T(1) : T(-1), ay = (i & 2) ? T(1) : T(-1), az = (i & 1) ? T(1) : T(-1);
      T wx = (i & 4) ? nmq.x : 1 - nmq.x, wy = (i & 2) ? nmq.y : 1 - nmq.y, wz = (i & 1) ? nmq.z : 1 - nmq.z;
      V3<T> ex = (ax * wy * wz) * normals[i], ey = (wx * ay * wz) * normals[i], ez = (wx * wy * az) * normals[i];
      cx = i == 0 ? ex : cx + ex;
      cy = i == 0 ? ey : cy + ey;
      cz = i == 0 ? ez : cz + ez;
    }
    for (int c = 0; c < 8; c++) {
      V3<T> ddp(ray_dir.x * dtdv[c], ray_dir.y * dtdv[c], ray_dir.z * dtdv[c]);  // dnmq/ddp taken as 1
      relax->dSd[8 * k + c].dNorm = relax->dSd[8 * k + c].dNorm + (cx * ddp.x + cy * ddp.y + cz * ddp.z);
    }
    for (int i = 0; i < 8; i++)
      if (length(normals[i]) > T(0.5f))
        for (int c = 0; c < 8; c++) relax->dSd[8 * i + c].dNorm = mul(nd, relax->dSd[8 * i + c].dNorm);
  }

  // BVH2DR.cpp:623-1073, Newton branch; returns hit ? t : qFar + EPS
  T Intersect(uint32_t ray_flags, V3<T> ray_dir, const T values[8], T d, T qFar, V3<T> start_q, Payload<T>* relax) const {
    const T EPS = T(1e-6f), SDF_EPS = T(1e-7f);
    T d_inv = T(1) / d;
    T t = 0;
    bool hit = false;
    if (relax) relax->missed_hit_candidate.sdf = T(1000.f);
    T start_dist = eval_dist_trilinear(values, start_q + t * ray_dir);
    if (start_dist <= SDF_EPS) {
      hit = true;
    } else {
      T s000 = values[0] * d_inv, s001 = values[1] * d_inv, s010 = values[2] * d_inv, s011 = values[3] * d_inv;
      T s100 = values[4] * d_inv, s101 = values[5] * d_inv, s110 = values[6] * d_inv, s111 = values[7] * d_inv;
      T a = s101 - s001;
      T k0 = s000, k1 = s100 - s000, k2 = s010 - s000, k3 = s110 - s010 - k1, k4 = k0 - s001, k5 = k1 - a;
      T k6 = k2 - (s011 - s001), k7 = k3 - (s111 - s011 - a);
      V3<T> o = start_q, d3 = ray_dir;
      T m0 = o.x * o.y, m1 = d3.x * d3.y, m2 = o.x * d3.y + o.y * d3.x;
      T m3 = k5 * o.z - k1, m4 = k6 * o.z - k2, m5 = k7 * o.z - k3;
      T c0 = (k4 * o.z - k0) + o.x * m3 + o.y * m4 + m0 * m5;
      T c1 = d3.x * m3 + d3.y * m4 + m2 * m5 + d3.z * (k4 + k5 * o.x + k6 * o.y + k7 * m0);
      T c2 = m1 * m5 + d3.z * (k5 * d3.x + k6 * d3.y + k7 * m2);
      T c3 = k7 * m1 * d3.z;
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
      T sdf0 = c0;
      T sdf1 = (c0 + t1 * (c1 + t1 * (c2 + t1 * c3)));
      T sdf2 = (c0 + t2 * (c1 + t2 * (c2 + t2 * c3)));
      T sdf3 = (c0 + t3 * (c1 + t3 * (c2 + t3 * c3)));
      bool s0 = sdf0 > 0, s1 = sdf1 > 0, s2 = sdf2 > 0, s3 = sdf3 > 0;
      T nwt_min = t0, nwt_max = t0;
      if (s0 != s1) { nwt_min = t0; nwt_max = t1; }
      else if (s1 != s2) { nwt_min = t1; nwt_max = t2; }
      else if (s2 != s3) { nwt_min = t2; nwt_max = t3; }
      if (nwt_min < nwt_max) {
        unsigned it = 0;
        T rtn = T(0.5f) * (nwt_min + nwt_max);
        T f = 1000;
        while (it < 10u && std::abs(f) >= SDF_EPS) {
          f = c0 + rtn * (c1 + rtn * (c2 + rtn * c3));
          T df = c1 + rtn * (2 * c2 + rtn * 3 * c3);
          T dx = f / (df + sgn(df) * T(1e-9f));
          rtn -= dx;
          it++;
        }
        t = rtn;
        hit = (t >= 0 && t <= qFar && std::abs(f) < SDF_EPS);
      }
      if (relax && (ray_flags & LRT_DR_RAY_FLAG_BORDER)) {  // relaxation point, BVH2DR.cpp:960-1035
        T t_min = qFar + EPS;
        if (std::abs(qa) > EPS) {
          T root1 = T(-0.5f) * (qb + std::sqrt(D)) / qa, root2 = T(-0.5f) * (qb - std::sqrt(D)) / qa;
          t_min = (sgn(qa) > 0) == (root1 < root2) ? root1 : root2;
        } else if (qb < -EPS) {
          t_min = -qc / qb;
        }
        T sdf_min = (c0 + t_min * (c1 + t_min * (c2 + t_min * c3)));
        HitDR<T>& mh = relax->missed_hit;
        if (mh.check_vox_border == 1) {
          if (c1 <= 0) {
            if (sdf0 < T(0) && -sdf0 * d < mh.sdf) { mh.t = t0; mh.sdf = -sdf0 * d; }
          }
          mh.check_vox_border = 0;
        }
        if (t_min >= 0 && t_min <= qFar && sdf_min < T(0) && -sdf_min * d < mh.sdf) { mh.t = t_min; mh.sdf = -sdf_min * d; }
        if (sdf3 < T(0) && -sdf3 * d < mh.sdf && (c1 + qFar * (2 * c2 + qFar * 3 * c3)) >= T(0)) {
          mh.check_vox_border = 1;
          relax->missed_hit_candidate.t = qFar;
          relax->missed_hit_candidate.sdf = -sdf3 * d;
        }
      }
    }
    return hit ? t : qFar + EPS;
  }

  // dt/dvalues in voxel units of t
  void dIntersect_dValues(uint32_t ray_flags, V3<T> ray_dir, const T values[8], T d, T qFar, V3<T> start_q, T t, T out[8]) const {
    if (dtdv_mode == 1) {  // BVH2DR.cpp:1075-1090
      T v[8];
      for (int i = 0; i < 8; i++) v[i] = values[i];
      const T delta = T(0.001f);
      for (int i = 0; i < 8; i++) {
        v[i] += delta;
        T d1 = Intersect(ray_flags, ray_dir, v, d, qFar, start_q, nullptr);
        v[i] -= 2 * delta;
        T d2 = Intersect(ray_flags, ray_dir, v, d, qFar, start_q, nullptr);
        v[i] += delta;
        out[i] = (d1 - d2) / (2 * delta);
      }
      return;
    }
    // implicit function theorem on f(q(t); v) = 0, q(t) = start_q + t*dir:
    // dt/dv_i = -w_i(q*) / (grad f(q*) . dir)
    V3<T> q = start_q + t * ray_dir;
    V3<T> g = eval_dist_trilinear_diff(values, q);
    T fd = dot(g, ray_dir);
    for (int i = 0; i < 8; i++) {
      T w = ((i & 4) ? q.x : 1 - q.x) * ((i & 2) ? q.y : 1 - q.y) * ((i & 1) ? q.z : 1 - q.z);
      out[i] = fd != T(0) ? -w / fd : T(0);
    }
  }

  // trilinear colour of the hit brick, BVH2Common.cpp:962-982 / BVH2DR.cpp:1489-1524
  static V3<T> color_lerp(const V3<T> c[8], V3<T> dp) {
    V3<T> r(0, 0, 0);
    for (int i = 0; i < 8; i++) {
      T w = ((i & 4) ? dp.x : 1 - dp.x) * ((i & 2) ? dp.y : 1 - dp.y) * ((i & 1) ? dp.z : 1 - dp.z);
      r = r + w * c[i];
    }
    return r;
  }
  // dLerp/d(dp) as three columns (BVH2DR.cpp:257-337)
  static void color_lerp_ddp(const V3<T> c[8], V3<T> dp, V3<T> col[3]) {
    col[0] = col[1] = col[2] = V3<T>(0, 0, 0);
    for (int i = 0; i < 8; i++) {
      T ax = (i & 4) ? T(1) : T(-1), ay = (i & 2) ? T(1) : T(-1), az = (i & 1) ? T(1) : T(-1);
      T wx = (i & 4) ? dp.x : 1 - dp.x, wy = (i & 2) ? dp.y : 1 - dp.y, wz = (i & 1) ? dp.z : 1 - dp.z;
      col[0] = col[0] + (ax * wy * wz) * c[i];
      col[1] = col[1] + (wx * ay * wz) * c[i];
      col[2] = col[2] + (wx * wy * az) * c[i];
    }
  }

  // BVH2DR.cpp:1130-1610 (ID32F layouts, no normal derivatives)
  void OctreeBrickIntersectWithGrad(uint32_t ray_flags, V3<T> ray_pos, V3<T> ray_dir, T tNear, uint32_t brick, Payload<T>* relax,
                                    HitDR<T>* pHit) const {
    T values[8];
    const SdfSBSHeader& h = S.header;
    const uint32_t v_size = h.brick_size + 2 * h.brick_pad + 1;
    const SdfSBSNode& N = S.nodes[brick];
    T px = T(N.pos_xy >> 16), py = T(N.pos_xy & 0xFFFF), pz = T(N.pos_z_lod_size >> 16), sz = T(N.pos_z_lod_size & 0xFFFF);
    T sz_inv = T(2) / sz;
    T d = T(2) / (sz * T(h.brick_size));
    V3<T> bmin = V3<T>(-1, -1, -1) + sz_inv * V3<T>(px, py, pz);
    V3<T> bmax = bmin + sz_inv * V3<T>(1, 1, 1);
    const T qscale = T(0.5f) * sz * T(h.brick_size);
    V2<T> bnf = RayBoxIntersection2(ray_pos, SafeInverse(ray_dir), bmin, bmax);
    T old_t = pHit->t;
    const uint32_t max_steps = 16u * (h.brick_size + 1u) + 16u;  // termination cap as in the plain path
    uint32_t steps = 0;
    while (bnf.x < bnf.y && pHit->t == old_t && steps++ < max_steps) {
      V3<T> hit_pos = ray_pos + bnf.x * ray_dir;
      V3<T> local = (hit_pos - bmin) * qscale;
      V3<T> vp = vfloor(vclamp(local, T(1e-6f), T(h.brick_size) - T(1e-6f)));
      uint32_t vx = (uint32_t)vp.x, vy = (uint32_t)vp.y, vz = (uint32_t)vp.z;
      V3<T> min_pos = bmin + d * vp;
      V3<T> max_pos = min_pos + d * V3<T>(1, 1, 1);
      uint32_t idx[8];
      for (int i = 0; i < 8; i++) {
        uint32_t vId = (vx + ((i & 4) >> 2)) * v_size * v_size + (vy + ((i & 2) >> 1)) * v_size + (vz + (i & 1));
        idx[i] = S.values[N.data_offset + vId];
        values[i] = T(S.values_f[idx[i]]);
      }
      V2<T> fnf = RayBoxIntersection2(ray_pos, SafeInverse(ray_dir), min_pos, max_pos);
      if (tNear < fnf.x) {
        V3<T> start_pos = ray_pos + fnf.x * ray_dir;
        V3<T> start_q = (start_pos - min_pos) * qscale;
        T qFar = (fnf.y - fnf.x) * qscale;
        T prev_sdf = relax->missed_hit.sdf;
        T t = Intersect(ray_flags, ray_dir, values, d, qFar, start_q, relax);
        T tReal = fnf.x + d * t;
        if (ray_flags & LRT_DR_RAY_FLAG_BORDER) {  // BVH2DR.cpp:1223-1275
          if (relax->missed_hit.sdf < prev_sdf && relax->missed_hit.t < t) {
            T missed_t = relax->missed_hit.t;
            relax->missed_hit.t = fnf.x + d * relax->missed_hit.t;
            V3<T> q_ast = start_q + missed_t * ray_dir;
            V3<T> y_ast = ray_pos + (fnf.x + (d * missed_t)) * ray_dir;
            V3<T> dp_ast = (y_ast - bmin) * (T(0.5f) * sz);
            V3<T> g = eval_dist_trilinear_diff(values, q_ast) * (T(1) / d);
            V3<T> mult(std::abs(q_ast.x - T(0.5f)) > T(0.5f) - T(1e-5f) ? T(0) : T(1),
                       std::abs(q_ast.y - T(0.5f)) > T(0.5f) - T(1e-5f) ? T(0) : T(1),
                       std::abs(q_ast.z - T(0.5f)) > T(0.5f) - T(1e-5f) ? T(0) : T(1));
            g = g * mult;
            T glen = length(g);
            if (glen > T(1e-9f)) {
              for (int i = 0; i < 8; i++) {
                T w = ((i & 4) ? q_ast.x : 1 - q_ast.x) * ((i & 2) ? q_ast.y : 1 - q_ast.y) * ((i & 1) ? q_ast.z : 1 - q_ast.z);
                relax->missed_dSDF_dtheta[i] = -w / glen;
              }
              uint32_t t_off = N.data_offset + v_size * v_size * v_size;
              V3<T> cols[8];
              for (int i = 0; i < 8; i++) {
                uint32_t ci = S.values[t_off + i];
                cols[i] = V3<T>(T(S.values_f[ci]), T(S.values_f[ci + 1]), T(S.values_f[ci + 2]));
              }
              relax->missed_hit.color = color_lerp(cols, dp_ast);
              relax->missed_hit.normal = g / glen;
              for (int i = 0; i < 8; i++) relax->missed_indices[i] = idx[i];
            }
          }
        }
        if (t <= qFar && tReal < pHit->t) {  // BVH2DR.cpp:1277-1600
          pHit->t = tReal;
          pHit->primId = brick;
          pHit->instId = 0;
          pHit->geomId = 0u | (S.type << LRT_SH_TYPE);
          unsigned dvo = 0;
          if ((S.header.aux_data & LRT_SBS_LAYOUT_MASK) == LRT_SBS_LAYOUT_ID32F_IRGB32F_IN) {
            V3<T> p0 = start_q + t * ray_dir;
            dvo += p0.x < T(0.5f) ? 4 : 0;
            dvo += p0.y < T(0.5f) ? 2 : 0;
            dvo += p0.z < T(0.5f) ? 1 : 0;
            dvo *= 8;
          }
          T dtdv[8] = {0, 0, 0, 0, 0, 0, 0, 0};
          if (ray_flags & (LRT_DR_RAY_FLAG_DDIFFUSE_DPOS | LRT_DR_RAY_FLAG_DNORM_DPOS | LRT_DR_RAY_FLAG_DDIST_DPOS)) {
            dIntersect_dValues(ray_flags, ray_dir, values, d, qFar, start_q, t, dtdv);
            for (int i = 0; i < 8; i++) {
              relax->dSd[dvo + i].index = idx[i];
              relax->dSd[dvo + i].dDist = d * dtdv[i];
            }
          }
          if (need_normal) NormalWithGrad(ray_flags, ray_dir, values, idx, start_q, t, vp, brick, dvo, dtdv, relax, pHit);
          V3<T> pos = ray_pos + pHit->t * ray_dir;
          V3<T> dp = (pos - bmin) * (T(0.5f) * sz);
          uint32_t t_off = N.data_offset + v_size * v_size * v_size;
          V3<T> cols[8];
          for (int i = 0; i < 8; i++) {
            uint32_t ci = S.values[t_off + i];
            cols[i] = V3<T>(T(S.values_f[ci]), T(S.values_f[ci + 1]), T(S.values_f[ci + 2]));
          }
          pHit->color = color_lerp(cols, dp);
          if (ray_flags & LRT_DR_RAY_FLAG_DDIFFUSE_DCOLOR) {
            for (int i = 0; i < 8; i++) {
              T lx = (i & 4) ? dp.x : 1 - dp.x, ly = (i & 2) ? dp.y : 1 - dp.y, lz = (i & 1) ? dp.z : 1 - dp.z;
              relax->dSc_index[i] = S.values[t_off + i];
              relax->dSc_value[i] = lx * ly * lz;
            }
          }
          if (ray_flags & LRT_DR_RAY_FLAG_DDIFFUSE_DPOS) {
            V3<T> J[3];
            color_lerp_ddp(cols, dp, J);
            for (int i = 0; i < 8; i++) {
              V3<T> v(T(0.5f) * sz * ray_dir.x * dtdv[i], T(0.5f) * sz * ray_dir.y * dtdv[i], T(0.5f) * sz * ray_dir.z * dtdv[i]);
              relax->dSd[dvo + i].dDiffuse = J[0] * v.x + J[1] * v.y + J[2] * v.z;
            }
          }
          break;
        }
      }
      bnf.x += mx(T(0), fnf.y - bnf.x) + T(1e-6f);
    }
  }

  // BVH2DR.cpp:106-167
  void TraverseWithGrad(uint32_t ray_flags, V3<T> ray_pos, V3<T> ray_dir, T tNear, Payload<T>* relax, HitDR<T>* pHit) const {
    uint32_t stack[LRT_STACK_SIZE];
    int top = 0;
    uint32_t left = 0;
    const V3<T> inv = SafeInverse(ray_dir);
    while (top >= 0 && !(pHit->primId != kInvalid)) {
      while (top >= 0 && ((left & LRT_LEAF_BIT) == 0)) {
        const BVHNodePair& fat = S.pairs[left];
        V2<T> tm0 = RayBoxIntersection2(ray_pos, inv, V3<T>(fat.left.boxMin[0], fat.left.boxMin[1], fat.left.boxMin[2]),
                                        V3<T>(fat.left.boxMax[0], fat.left.boxMax[1], fat.left.boxMax[2]));
        V2<T> tm1 = RayBoxIntersection2(ray_pos, inv, V3<T>(fat.right.boxMin[0], fat.right.boxMin[1], fat.right.boxMin[2]),
                                        V3<T>(fat.right.boxMax[0], fat.right.boxMax[1], fat.right.boxMax[2]));
        const bool h0 = (tm0.x <= tm0.y) && (tm0.y >= tNear) && (tm0.x <= pHit->t);
        const bool h1 = (tm1.x <= tm1.y) && (tm1.y >= tNear) && (tm1.x <= pHit->t);
        left = h0 ? fat.left.leftOffset : fat.right.leftOffset;
        if (h0 && h1) {
          left = (tm0.x <= tm1.x) ? fat.left.leftOffset : fat.right.leftOffset;
          stack[top] = (tm0.x <= tm1.x) ? fat.right.leftOffset : fat.left.leftOffset;
          top++;
        }
        if (!h0 && !h1) {
          top--;
          left = stack[std::max(top, 0)];
        }
      }
      if (top >= 0 && left != kInvalid) {
        const T tNearSdf = mx(tNear, T(0.1f));  // BVH2DR.cpp:175-176
        OctreeBrickIntersectWithGrad(ray_flags, ray_pos, ray_dir, tNearSdf, left & LRT_START_MASK, relax, pHit);
      }
      top--;
      left = stack[std::max(top, 0)];
    }
  }

  // BVH2DR.cpp:12-102
  HitDR<T> NearestHitWithGrad(uint32_t ray_flags, V4<T> pn, V4<T> df, Payload<T>* relax) const {
    if (df.w <= T(0)) df.w *= T(-1);
    V3<T> pos(pn.x, pn.y, pn.z), dir(df.x, df.y, df.z);
    HitDR<T> hit;
    hit.t = df.w;
    hit.primId = hit.instId = hit.geomId = kInvalid;
    hit.color = V3<T>(0, 0, 0);
    hit.sdf = T(1000.0f);
    hit.normal = V3<T>(1, 0, 0);
    hit.check_vox_border = 0;
    relax->missed_hit = hit;
    relax->missed_hit.check_vox_border = 1;
    relax->missed_hit_candidate = relax->missed_hit;
    for (int i = 0; i < 8; i++) {
      relax->missed_indices[i] = kInvalid;
      relax->missed_dSDF_dtheta[i] = 0;
      relax->dSc_index[i] = kInvalid;
      relax->dSc_value[i] = 0;
    }
    for (int i = 0; i < 64; i++) {
      relax->dSd[i].index = kInvalid;
      relax->dSd[i].dDiffuse = V3<T>(0, 0, 0);
      relax->dSd[i].dNorm = V3<T>(0, 0, 0);
      relax->dSd[i].dDist = 0;
    }
    V2<T> bh = RayBoxIntersection2(pos, SafeInverse(dir), V3<T>(-1, -1, -1), V3<T>(1, 1, 1));
    if ((bh.x <= bh.y) && (bh.y > pn.w) && (bh.x < hit.t)) TraverseWithGrad(ray_flags, pos, dir, pn.w, relax, &hit);
    if (relax->missed_hit_candidate.sdf < relax->missed_hit.sdf) relax->missed_hit = relax->missed_hit_candidate;
    return hit;
  }
};

template <typename T> static T lin_depth(T z) { return (z - T(kZNear)) / (T(kZFar) - T(kZNear)); }

// CalculateColorWithGrad (MultiRendererDR.cpp:900-980); CalculateColor
// (:855-898) is its colour. dCdD / dCdN: dColor/dDiffuse and dColor/dNormal.
template <typename T>
static V3<T> calc_color_grad(uint32_t mode, const HitDR<T>& h, const std::vector<LrtLight>& lights, M3<T>& dCdD, M3<T>& dCdN) {
  const V3<T> z(0, 0, 0);
  dCdD = {{V3<T>(1, 0, 0), V3<T>(0, 1, 0), V3<T>(0, 0, 1)}};
  dCdN = {{z, z, z}};
  switch (mode) {
    case LRT_DR_RENDER_MODE_DIFFUSE: return h.color;
    case LRT_DR_RENDER_MODE_LAMBERT: {  // direct and ambient lights (point lights assert in the reference)
      V3<T> c(0, 0, 0);
      dCdD = {{z, z, z}};
      for (const LrtLight& l : lights) {
        const V3<T> lc(T(l.color[0]), T(l.color[1]), T(l.color[2]));
        if (l.type == LRT_LIGHT_TYPE_DIRECT) {
          const V3<T> L(T(l.space[0]), T(l.space[1]), T(l.space[2]));
          T q = mx(T(0), dot(h.normal, L));
          c = c + (lc * h.color) * q;
          if (q > T(0)) {
            dCdD.r[0] = dCdD.r[0] + V3<T>(lc.x * q, 0, 0);
            dCdD.r[1] = dCdD.r[1] + V3<T>(0, lc.y * q, 0);
            dCdD.r[2] = dCdD.r[2] + V3<T>(0, 0, lc.z * q);
            dCdN.r[0] = dCdN.r[0] + (lc.x * h.color.x) * L;
            dCdN.r[1] = dCdN.r[1] + (lc.y * h.color.y) * L;
            dCdN.r[2] = dCdN.r[2] + (lc.z * h.color.z) * L;
          }
        } else {
          c = c + lc * h.color;
          dCdD.r[0] = dCdD.r[0] + V3<T>(lc.x, 0, 0);
          dCdD.r[1] = dCdD.r[1] + V3<T>(0, lc.y, 0);
          dCdD.r[2] = dCdD.r[2] + V3<T>(0, 0, lc.z);
        }
      }
      return c;
    }
    case LRT_DR_RENDER_MODE_MASK: dCdD = {{z, z, z}}; return V3<T>(1, 1, 1);
    case LRT_DR_RENDER_MODE_LINEAR_DEPTH: { T d = lin_depth(h.t); return V3<T>(d, d, d); }
    case LRT_DR_RENDER_MODE_NORMAL:
      dCdD = {{z, z, z}};
      dCdN = {{V3<T>(1, 0, 0), V3<T>(0, 1, 0), V3<T>(0, 0, 1)}};
      return h.normal;
    default: return V3<T>(1, 0, 1);
  }
}
template <typename T> static V3<T> calc_color(uint32_t mode, const HitDR<T>& h, const std::vector<LrtLight>& lights) {
  M3<T> a, b;
  return calc_color_grad(mode, h, lights, a, b);
}

// t_samples_to_threshold (MultiRendererDR.cpp:997-1063)
static float samples_to_threshold(const float* s, unsigned count, float dv, unsigned dc) {
  const int G = 10, GAP = 4;
  float mnv = s[0], mxv = s[0];
  for (unsigned i = 1; i < count; i++) { mnv = std::min(mnv, s[i]); mxv = std::max(mxv, s[i]); }
  if (dc > 0) { mnv = std::min(mnv, dv); mxv = std::max(mxv, dv); }
  float diff = std::max(1e-9f, mxv - mnv);
  unsigned g[G] = {0};
  for (unsigned i = 0; i < count; i++) g[std::min(std::max(int((s[i] - mnv) / diff * G), 0), G - 1)]++;
  g[std::min(std::max(int((dv - mnv) / diff * G), 0), G - 1)] += dc;
  int best_len = 0, best_start = -1, start = -1;
  for (int i = 1; i < G; i++) {
    if (start != -1) {
      if (g[i] != 0) {
        int len = i - start;
        if (len > best_len) { best_len = len; best_start = start; }
        start = -1;
      }
    } else if (g[i] == 0) {
      start = i;
    }
  }
  if (best_len > GAP) return mnv + diff * (best_start + 0.5f * best_len) / G;
  return 0.0f;
}

// ---- SVM border sampling (MultiRendererDR.cpp:1501-1932) ----------------
// The fit, the clip and the stripe are computed in float, as the reference
// does, for both precisions of the tracer; only the traced distances follow T.

// SVM_fit (:1518-1546): perceptron with weight decay over X_mod = (x, y, 1) *
// label, etha 0.01, alpha 0.1, 200 epochs with early exit on zero errors.
static void svm_fit(const std::vector<float>& xs, const std::vector<float>& ys, const std::vector<float>& lab, float W[4]) {
  const float etha = 0.01f, alpha = 0.1f;
  const int epochs = 200;
  const unsigned N = (unsigned)xs.size();
  float w0 = -1.f, w1 = 1.f, w2 = 0.f;
  const float mult = 1.0f - etha * alpha / epochs;
  unsigned errors = N;
  for (int i = 0; i < epochs && errors > 0; i++) {
    errors = 0;
    for (unsigned j = 0; j < N; j++) {
      const float X0 = xs[j] * lab[j] * etha, X1 = ys[j] * lab[j] * etha, X2 = 1.0f * lab[j] * etha;
      const float margin = w0 * X0 + w1 * X1 + w2 * X2;
      w0 = w0 * mult; w1 = w1 * mult; w2 = w2 * mult;
      if (margin < 0) {
        w0 = w0 + X0; w1 = w1 + X1; w2 = w2 + X2;
        errors++;
      }
    }
  }
  W[0] = w0; W[1] = w1; W[2] = w2; W[3] = float(errors) / float(N);
}

static unsigned outcode(float x, float y) {  // ComputeOutCode (:1550-1571) on [0,1]^2
  unsigned c = 0;
  if (x < 0) c |= 1u;
  else if (x > 1) c |= 2u;
  if (y < 0) c |= 4u;
  else if (y > 1) c |= 8u;
  return c;
}

// Cohen-Sutherland line clipping (:1576-1645)
static bool clip_line(float& x0, float& y0, float& x1, float& y1) {
  unsigned c0 = outcode(x0, y0), c1 = outcode(x1, y1);
  for (int guard = 0; guard < 16; guard++) {
    if (!(c0 | c1)) return true;
    if (c0 & c1) return false;
    float x = -1, y = -1;
    const unsigned co = c1 > c0 ? c1 : c0;
    if (co & 8u) { x = x0 + (x1 - x0) * (1 - y0) / (y1 - y0); y = 1; }
    else if (co & 4u) { x = x0 + (x1 - x0) * (0 - y0) / (y1 - y0); y = 0; }
    else if (co & 2u) { y = y0 + (y1 - y0) * (1 - x0) / (x1 - x0); x = 1; }
    else if (co & 1u) { y = y0 + (y1 - y0) * (0 - x0) / (x1 - x0); x = 0; }
    if (co == c0) { x0 = x; y0 = y; c0 = outcode(x0, y0); }
    else { x1 = x; y1 = y; c1 = outcode(x1, y1); }
  }
  return false;
}

static void to_screen(const float PV[16], float px, float py, float pz, float& sx, float& sy) {  // :1296-1301
  float ps[4];
  for (int r = 0; r < 4; r++) ps[r] = PV[0 * 4 + r] * px + PV[1 * 4 + r] * py + PV[2 * 4 + r] * pz + PV[3 * 4 + r] * 1.0f;
  sx = 0.5f * (ps[0] / ps[3] + 1.0f);
  sy = 1.0f - 0.5f * (ps[1] / ps[3] + 1.0f);
}

// CastBorderRaySVM (:1651-1931) for one border pixel; returns the number of
// border rays that found a relaxation point. The reference draws the stripe
// positions with rand(); here they are rand2 of (sample index + 0x10000).
template <typename T, typename Tracer, typename Contribute, typename PayloadT>
static uint32_t svm_border_pixel(Tracer& tr, const Camera& cam, uint32_t x, uint32_t y, uint32_t W, uint32_t H, uint32_t seed, uint32_t bspp,
                                 float eps, const float* wv, const float PV[16], Contribute& contribute, PayloadT& py) {
  unsigned svm_spp = std::max(1u, (unsigned)(0.25f * (float)bspp));
  const uint32_t sq = (uint32_t)std::sqrt((double)svm_spp);
  const float i_sq = 1.0f / (float)sq;
  svm_spp = sq * sq;
  std::vector<float> xs(svm_spp), ys(svm_spp), ts(svm_spp), lab(svm_spp);
  float chx = 1e6f, chy = 1e6f, chz = 1e6f, closest_t = 1e6f;
  for (unsigned s = 0; s < svm_spp; s++) {  // step 1
    V2<float> rnd = rand2(x, y, s, seed);
    V2<float> d{i_sq * ((float)(s / sq) + rnd.x), i_sq * ((float)(s % sq) + rnd.y)};
    V4<T> pn, df;
    InitEyeRay<T>(cam, x, y, d, W, H, &pn, &df);
    HitDR<T> hit = tr.NearestHitWithGrad(LRT_DR_RAY_FLAG_NO_DIFF, pn, df, &py);
    xs[s] = d.x;
    ys[s] = d.y;
    ts[s] = (float)hit.t;
    if (ts[s] < closest_t) {
      closest_t = ts[s];
      chx = (float)pn.x + (float)df.x * closest_t;
      chy = (float)pn.y + (float)df.y * closest_t;
      chz = (float)pn.z + (float)df.z * closest_t;
    }
  }
  const float thr = samples_to_threshold(ts.data(), svm_spp, 0.0f, 0u);
  unsigned type0 = 0;
  for (unsigned s = 0; s < svm_spp; s++) {
    lab[s] = ts[s] <= thr ? -1.0f : 1.0f;
    type0 += ts[s] <= thr ? 1u : 0u;
  }
  if (type0 == svm_spp || type0 == 0) return 0;
  float Wt[4];  // step 2
  svm_fit(xs, ys, lab, Wt);
  const float error_rate = Wt[3];
  const float f0 = (float)type0 / (float)svm_spp;
  const float rare = std::min(f0, 1.0f - f0);
  float p0x = 0, p0y = 0, p1x = 0, p1y = 0, rx = 0, ry = 0, nx = 0, ny = 0, radius = 0, area = 1.0f;
  bool force = false;
  if (error_rate > 0.05f || 2 * error_rate > rare) {
    force = true;
  } else {
    if (std::abs(Wt[1]) > 1e-6f && std::abs(Wt[0]) > 1e-6f) {
      p0x = 0; p0y = -Wt[2] / Wt[1];
      p1x = 1; p1y = (-Wt[2] - Wt[0]) / Wt[1];
      const float dy = p1y - p0y;
      if (p0y < -1e-6f || p0y > 1 + 1e-6f) {
        const float k1 = -p0y / dy, k2 = (1 - p0y) / dy, k = std::abs(k1) < std::abs(k2) ? k1 : k2;
        p0x = p0x + k * 1.0f; p0y = p0y + k * dy;
      }
      if (p1y < -1e-6f || p1y > 1 + 1e-6f) {
        const float k1 = -p1y / dy, k2 = (1 - p1y) / dy, k = std::abs(k1) < std::abs(k2) ? k1 : k2;
        p1x = p1x + k * 1.0f; p1y = p1y + k * dy;
      }
    } else if (std::abs(Wt[0]) > 1e-6f) {
      p0x = -Wt[2] / Wt[0]; p0y = 0;
      p1x = -Wt[2] / Wt[0]; p1y = 1;
    } else {
      force = true;  // the reference leaves p0, p1 unset here; (0,0) is used
    }
    if (!clip_line(p0x, p0y, p1x, p1y)) return 0;
    float q0x, q0y, q1x, q1y;  // border thickness at the closest hit, up = worldView row 1
    to_screen(PV, chx, chy, chz, q0x, q0y);
    to_screen(PV, chx + wv[1] * eps, chy + wv[5] * eps, chz + wv[9] * eps, q1x, q1y);
    const float ddx = q1x - q0x, ddy = q1y - q0y;
    const float thick = std::sqrt(ddx * ddx + ddy * ddy) * (float)std::max(W, H);
    radius = std::max(0.15f, thick) + 2.0f * error_rate;
    rx = p1x - p0x; ry = p1y - p0y;
    float lr = std::sqrt(rx * rx + ry * ry);
    const float ux = rx / lr, uy = ry / lr;
    p0x = p0x - radius * ux; p0y = p0y - radius * uy;
    p1x = p1x + radius * ux; p1y = p1y + radius * uy;
    rx = p1x - p0x; ry = p1y - p0y;
    lr = std::sqrt(rx * rx + ry * ry);
    const float ln = std::sqrt(ry * ry + rx * rx);
    nx = -ry / ln; ny = rx / ln;
    area = 2 * radius * lr;
    if (!force && area < 0.005f) return 0;
  }
  const float area_mult = std::min(1.0f, 0.05f + 2.0f * area);
  const unsigned nb = std::max(1u, (unsigned)(area_mult * (float)(bspp - svm_spp)));
  const float pdf = force ? 1.0f / (float)nb : area / (float)nb;
  uint32_t hits = 0;
  for (unsigned k = 0; k < nb; k++) {  // step 3
    const unsigned sid = svm_spp + k;
    V2<float> d;
    if (force) {
      d = rand2(x, y, sid, seed);
    } else {
      V2<float> u = rand2(x, y, sid + 0x10000u, seed);
      const float u1 = u.x, u2 = 2.0f * u.y - 1.0f;
      d = V2<float>{p0x + u1 * rx + u2 * radius * nx, p0y + u1 * ry + u2 * radius * ny};
    }
    if (!(d.x > 0 && d.x < 1 && d.y > 0 && d.y < 1)) continue;
    V4<T> pn, df;
    InitEyeRay<T>(cam, x, y, d, W, H, &pn, &df);
    HitDR<T> hit = tr.NearestHitWithGrad(LRT_DR_RAY_FLAG_BORDER, pn, df, &py);
    if (!(py.missed_hit.sdf < T(eps))) continue;
    hits++;
    contribute(pn, df, hit, T(pdf));
  }
  return hits;
}

template <typename T>
static int dr_impl(const Scene& S, const MultiRendererDRPreset& dr, const Camera& cam, const float* wv, const float* pr, uint32_t W,
                   uint32_t H, uint32_t seed, const float* ref, int nthreads, int dtdv_mode, double* out_image, double* out_grad,
                   orc_dr_stats* stats) {
  const uint32_t mode = dr.dr_render_mode;
  if (mode > LRT_DR_RENDER_MODE_NORMAL) return -3;
  if (dr.dr_loss_function != LRT_DR_LOSS_FUNCTION_MSE && dr.dr_loss_function != LRT_DR_LOSS_FUNCTION_MAE) return -3;
  if (S.type != LRT_TYPE_SDF_SBS_COL) return -4;  // DR needs indexed colour layouts (MultiRendererDR.cpp:283-284)
  if (mode == LRT_DR_RENDER_MODE_LAMBERT)
    for (const LrtLight& l : S.lights)
      if (l.type == LRT_LIGHT_TYPE_POINT) return -3;  // MultiRendererDR.cpp:925-928 asserts
  MultiRenderPreset p{};
  p.sdf_node_intersect = LRT_SDF_NODE_INTERSECT_NEWTON;
  p.ray_gen_mode = LRT_RAY_GEN_MODE_RANDOM;
  p.spp = dr.spp;
  DRTracer<T> tr(S, p, dtdv_mode);
  tr.need_normal = mode == LRT_DR_RENDER_MODE_LAMBERT || mode == LRT_DR_RENDER_MODE_NORMAL;
  const uint32_t spp = dr.spp;
  const size_t np = (size_t)W * H;
  std::vector<uint32_t> packed = PackXY(W, H);
  std::vector<float> img(np * 4, 0.f), depth_thr(np, 0.f);
  const size_t P = S.values_f.size();
  if (nthreads <= 0) nthreads = omp_get_max_threads();
  std::vector<std::vector<double>> grads(nthreads, std::vector<double>(P, 0.0));
  std::vector<double> loss_v(nthreads, 0.0);
  uint32_t ray_flags = LRT_DR_RAY_FLAG_NO_DIFF;  // CastRayWithGrad, MultiRendererDR.cpp:1119-1142
  if (dr.dr_diff_mode == LRT_DR_DIFF_MODE_FINITE_DIFF) ray_flags = LRT_DR_RAY_FLAG_NO_DIFF;
  else if (dr.dr_input_type == LRT_DR_INPUT_TYPE_LINEAR_DEPTH) ray_flags = LRT_DR_RAY_FLAG_DDIST_DPOS;
  else {
    if (dr.dr_reconstruction_flags & LRT_DR_RECONSTRUCTION_FLAG_COLOR) ray_flags |= LRT_DR_RAY_FLAG_DDIFFUSE_DCOLOR;
    if (dr.dr_reconstruction_flags & LRT_DR_RECONSTRUCTION_FLAG_GEOMETRY) {
      if (mode == LRT_DR_RENDER_MODE_DIFFUSE) ray_flags |= LRT_DR_RAY_FLAG_DDIFFUSE_DPOS;
      else if (mode == LRT_DR_RENDER_MODE_LAMBERT) ray_flags |= LRT_DR_RAY_FLAG_DDIFFUSE_DPOS | LRT_DR_RAY_FLAG_DNORM_DPOS;
      else if (mode == LRT_DR_RENDER_MODE_NORMAL) ray_flags |= LRT_DR_RAY_FLAG_DNORM_DPOS;
    }
  }
  const bool mse = dr.dr_loss_function == LRT_DR_LOSS_FUNCTION_MSE;
  auto loss_grad = [&](const float c[4], const float r[4], T g[4]) {
    for (int k = 0; k < 4; k++) g[k] = mse ? T(2.0f) * (T(c[k]) - T(r[k])) : sgn(T(c[k]) - T(r[k]));
  };

  // I: primary rays with gradients (MultiRendererDR.cpp:596-607, 1084-1294)
#pragma omp parallel num_threads(nthreads)
  {
    const int tid_omp = omp_get_thread_num();
    std::vector<double>& G = grads[tid_omp];
    std::vector<Payload<T>> pl(spp);
    std::vector<HitDR<T>> hits(spp);
    std::vector<float> samples(spp);
#pragma omp for schedule(dynamic, 64)
    for (int64_t i = 0; i < (int64_t)np; i++) {
      const uint32_t XY = packed[i];
      const uint32_t x = XY & 0xFFFF, y = XY >> 16;
      if (x >= W || y >= H) continue;
      V3<T> res(0, 0, 0);
      int hit_count = 0;
      std::vector<std::pair<uint32_t, V3<T>>> pds;
      for (uint32_t s = 0; s < spp; s++) {
        V2<float> d = rand2(x, y, s, seed);
        V4<T> pn, df;
        InitEyeRay<T>(cam, x, y, d, W, H, &pn, &df);
        Payload<T>& py = pl[s];
        HitDR<T> hit = tr.NearestHitWithGrad(ray_flags, pn, df, &py);
        if (hit.primId == kInvalid) continue;
        M3<T> dCdD, dCdN;
        V3<T> color = calc_color_grad<T>(mode, hit, S.lights, dCdD, dCdN);
        res = res + color / T(spp);
        if (!tr.need_normal) {
          // dColor/dDiffuse is I for DIFFUSE (and left at I for LINEAR_DEPTH), 0 for MASK
          if ((ray_flags & LRT_DR_RAY_FLAG_DDIFFUSE_DCOLOR) && mode != LRT_DR_RENDER_MODE_MASK)
            for (int k = 0; k < 8; k++)
              if (py.dSc_index[k] != kInvalid)
                pds.push_back({py.dSc_index[k] | 0x80000000u, V3<T>(py.dSc_value[k], py.dSc_value[k], py.dSc_value[k])});
        } else if (ray_flags & LRT_DR_RAY_FLAG_DDIFFUSE_DCOLOR) {
          for (int k = 0; k < 8; k++)
            if (py.dSc_index[k] != kInvalid)
              pds.push_back({py.dSc_index[k] | 0x80000000u, mul(dCdD, V3<T>(py.dSc_value[k], py.dSc_value[k], py.dSc_value[k]))});
        }
        if (tr.need_normal && (ray_flags & (LRT_DR_RAY_FLAG_DDIFFUSE_DPOS | LRT_DR_RAY_FLAG_DNORM_DPOS))) {
          for (int k = 0; k < 64; k++)
            if (py.dSd[k].index != kInvalid) pds.push_back({py.dSd[k].index, mul(dCdD, py.dSd[k].dDiffuse) + mul(dCdN, py.dSd[k].dNorm)});
        } else if (ray_flags & LRT_DR_RAY_FLAG_DDIFFUSE_DPOS) {
          for (int k = 0; k < 64; k++)
            if (py.dSd[k].index != kInvalid) pds.push_back({py.dSd[k].index, py.dSd[k].dDiffuse});  // dColor_dDiffuse = I
        } else if (ray_flags & LRT_DR_RAY_FLAG_DDIST_DPOS) {
          for (int k = 0; k < 64; k++)
            if (py.dSd[k].index != kInvalid) {
              T v = py.dSd[k].dDist / (T(kZFar) - T(kZNear));
              pds.push_back({py.dSd[k].index, V3<T>(v, v, v)});
            }
        }
        samples[hit_count] = (float)hit.t;
        hit_count++;
      }
      const size_t pi = (size_t)y * W + x;
      float cm[4] = {(float)res.x, (float)res.y, (float)res.z, float(hit_count) / float(spp)};
      for (int k = 0; k < 4; k++) img[4 * pi + k] = cm[k];
      depth_thr[pi] = hit_count == 0 ? 0.f : samples_to_threshold(samples.data(), hit_count, 1e9f, spp - hit_count);
      const float* r = ref + 4 * pi;
      T g4[4];
      loss_grad(cm, r, g4);
      T l = 0;
      for (int k = 0; k < 4; k++) {
        T df = T(cm[k]) - T(r[k]);
        l += mse ? df * df : std::abs(df);
      }
      loss_v[tid_omp] += (double)l;
      V3<T> dLdC(g4[0] / T(spp), g4[1] / T(spp), g4[2] / T(spp));
      for (auto& pd : pds) {
        if (pd.first & 0x80000000u) {  // colour parameter triple
          uint32_t base = pd.first & 0x7FFFFFFFu;
          G[base + 0] += (double)(dLdC.x * pd.second.x);
          G[base + 1] += (double)(dLdC.y * pd.second.y);
          G[base + 2] += (double)(dLdC.z * pd.second.z);
        } else {
          G[pd.first] += (double)dot(dLdC, pd.second);
        }
      }
    }
  }

  // II: border pixels (MultiRendererDR.cpp:609-685), in packed order
  std::vector<uint32_t> border;
  for (size_t i = 0; i < np; i++) {
    const uint32_t XY = packed[i];
    const uint32_t x = XY & 0xFFFF, y = XY >> 16;
    if (x >= W || y >= H) continue;
    float d0 = img[4 * ((size_t)y * W + x) + 3];
    float max_diff = 0.f, max_thr = 0.f;
    if (x >= 1 && x < W - 1 && y >= 1 && y < H - 1) {
      for (int dx = -1; dx <= 1; dx++)
        for (int dy = -1; dy <= 1; dy++) {
          size_t q = (size_t)(y + dy) * W + x + dx;
          max_diff = std::max(max_diff, std::abs(d0 - img[4 * q + 3]));
          max_thr = std::max(max_thr, depth_thr[q]);
        }
    }
    bool is_border = mode == LRT_DR_RENDER_MODE_MASK ? max_diff > 0 : (max_diff > 0 || max_thr > 0);
    if (is_border || dr.debug_forced_border) border.push_back((uint32_t)i);
  }

  // III: border rays (MultiRendererDR.cpp:703-729, 1319-1499)
  uint32_t border_hits = 0;
  float PV[16];  // proj * worldView (TransformWorldToScreenSpace, MultiRendererDR.cpp:1296-1301)
  for (int c = 0; c < 4; c++)
    for (int r = 0; r < 4; r++) {
      float s = 0.f;
      for (int k = 0; k < 4; k++) s += pr[k * 4 + r] * wv[c * 4 + k];
      PV[c * 4 + r] = s;
    }
  if (!border.empty() && (dr.dr_reconstruction_flags & LRT_DR_RECONSTRUCTION_FLAG_GEOMETRY) &&
      dr.dr_diff_mode == LRT_DR_DIFF_MODE_DEFAULT) {
    const uint32_t bspp = dr.border_spp;
    const T eps = T(dr.border_relax_eps);
#pragma omp parallel num_threads(nthreads) reduction(+ : border_hits)
    {
      std::vector<double>& G = grads[omp_get_thread_num()];
      Payload<T> py;
      // CalculateBorderRayDerivatives (MultiRendererDR.cpp:1319-1375) of a
      // border ray that found a relaxation point
      auto contribute = [&](const V4<T>& pn, const V4<T>& df, const HitDR<T>& hit, const T pdf) {
        T delta[4];
        if (dr.dr_input_type == LRT_DR_INPUT_TYPE_LINEAR_DEPTH) {
          T din = lin_depth(py.missed_hit.t);
          T dout = hit.primId == kInvalid ? T(0) : lin_depth(hit.t);
          delta[0] = delta[1] = delta[2] = din - dout;
        } else {
          V3<T> fin = calc_color<T>(mode, py.missed_hit, S.lights);
          V3<T> fout = hit.primId == kInvalid ? V3<T>(0, 0, 0) : calc_color<T>(mode, hit, S.lights);
          delta[0] = fin.x - fout.x; delta[1] = fin.y - fout.y; delta[2] = fin.z - fout.z;
        }
        delta[3] = hit.primId == kInvalid ? T(1) : T(0);
        for (int j = 0; j < 8; j++) {
          if (py.missed_indices[j] == kInvalid) continue;
          T bd = mx(T(0), py.missed_hit.sdf);
          V3<T> ystar = V3<T>(pn.x, pn.y, pn.z) + py.missed_hit.t * V3<T>(df.x, df.y, df.z);
          V3<T> xstar = ystar - bd * py.missed_hit.normal;
          V4<T> ps = mul(M4<T>{{T(PV[0]), T(PV[1]), T(PV[2]), T(PV[3]), T(PV[4]), T(PV[5]), T(PV[6]), T(PV[7]), T(PV[8]), T(PV[9]),
                                T(PV[10]), T(PV[11]), T(PV[12]), T(PV[13]), T(PV[14]), T(PV[15])}},
                         V4<T>{xstar.x, xstar.y, xstar.z, T(1)});
          T sx = T(0.5f) * (ps.x / ps.w + T(1)), sy = T(1) - T(0.5f) * (ps.y / ps.w + T(1));
          // the reference converts to unsigned without a range check; the
          // pixel is clamped into the image here
          double fx = std::floor((double)(sx * T(W))), fy = std::floor((double)(sy * T(H)));
          uint32_t xp = (uint32_t)std::min(std::max(fx, 0.0), (double)W - 1), yp = (uint32_t)std::min(std::max(fy, 0.0), (double)H - 1);
          size_t q = (size_t)yp * W + xp;
          T g4[4];
          loss_grad(&img[4 * q], ref + 4 * q, g4);
          T dd = g4[0] * delta[0] + g4[1] * delta[1] + g4[2] * delta[2] + g4[3] * delta[3];
          T diff = pdf * (T(1) / eps) * dd * py.missed_dSDF_dtheta[j];
          G[py.missed_indices[j]] += (double)diff;
        }
      };
#pragma omp for schedule(dynamic, 4)
      for (int64_t bi = 0; bi < (int64_t)border.size(); bi++) {
        const uint32_t XY = packed[border[bi]];
        const uint32_t x = XY & 0xFFFF, y = XY >> 16;
        if (dr.dr_border_sampling == LRT_DR_BORDER_SAMPLING_SVM) {
          border_hits += svm_border_pixel<T>(tr, cam, x, y, W, H, seed, bspp, float(eps), wv, PV, contribute, py);
          continue;
        }
        for (uint32_t s = 0; s < bspp; s++) {  // CastBorderRay, MultiRendererDR.cpp:1377-1499
          V2<float> d = rand2(x, y, s, seed);
          V4<T> pn, df;
          InitEyeRay<T>(cam, x, y, d, W, H, &pn, &df);
          HitDR<T> hit = tr.NearestHitWithGrad(LRT_DR_RAY_FLAG_BORDER, pn, df, &py);
          if (!(py.missed_hit.sdf < eps)) continue;
          border_hits++;
          contribute(pn, df, hit, T(1.0f) / T(bspp));
        }
      }
    }
  }

  // IV: scale (getBaseGradientMult, MultiRendererDR.cpp:140-145) and loss
  const double g_mult = (double)(float)(std::pow((double)P + 1.0, 2.0 / 3.0) / (double)(W * H));
  for (size_t k = 0; k < P; k++) {
    double s = 0;
    for (int t = 0; t < nthreads; t++) s += grads[t][k];
    out_grad[k] += s * g_mult;
  }
  double loss = 0;
  for (double l : loss_v) loss += l;
  if (out_image)
    for (size_t k = 0; k < np * 4; k++) out_image[k] = img[k];
  if (stats) {
    stats->loss = loss / (double)np;
    stats->border_pixels = (uint32_t)border.size();
    stats->border_rays_hit = border_hits;
    stats->rays = (uint64_t)np * spp + (uint64_t)border.size() * dr.border_spp;
  }
  return 0;
}

}  // namespace

// RenderDRFiniteDiff, MultiRendererDR.cpp:751-815, over the active range of
// :428-436: every parameter of the range is set to p0 + delta and p0 - delta
// in turn, the image re-rendered (primary rays only: border rays run in the
// default mode alone, :703), the per-pixel Loss (:51-73, float) summed in
// double and scaled by getBaseGradientMult; the central difference is stored
// as float. The image and loss returned are those of the last render.
template <typename T>
static int dr_fd_impl(const Scene& S0, const MultiRendererDRPreset& dr, const Camera& cam, const float* wv, const float* pr, uint32_t W,
                      uint32_t H, uint32_t seed, const float* ref, int nthreads, double* out_image, double* out_grad, orc_dr_stats* stats) {
  Scene S = S0;
  const size_t P = S.values_f.size(), n_col = 24 * S.nodes.size();
  if (P < n_col) return -3;
  const bool geo = (dr.dr_reconstruction_flags & LRT_DR_RECONSTRUCTION_FLAG_GEOMETRY) != 0;
  const size_t start = geo ? 0 : P - n_col, end = geo ? P - n_col : P;
  if (end <= start) return -3;  // the reference asserts
  const size_t np = (size_t)W * H;
  std::vector<double> img(np * 4), scratch(P);
  const float g_mult = (float)(std::pow((double)P + 1.0, 2.0 / 3.0) / (double)(W * H));
  const float delta = dr.finite_diff_delta;
  auto loss_sum = [&]() {
    double acc = 0.0;
    for (size_t j = 0; j < np; j++) {
      float d[4], l;
      for (int k = 0; k < 4; k++) d[k] = (float)img[4 * j + k] - ref[4 * j + k];
      if (dr.dr_loss_function == LRT_DR_LOSS_FUNCTION_MSE) l = d[0] * d[0] + d[1] * d[1] + d[2] * d[2] + d[3] * d[3];
      else l = std::abs(d[0]) + std::abs(d[1]) + std::abs(d[2]) + std::abs(d[3]);
      acc += l;
    }
    return acc;
  };
  double loss_minus = 0.0;
  for (size_t i = start; i < end; i++) {
    const float p0 = S.values_f[i];
    S.values_f[i] = p0 + delta;
    int rc = dr_impl<T>(S, dr, cam, wv, pr, W, H, seed, ref, nthreads, 0, img.data(), scratch.data(), nullptr);
    if (rc) return rc;
    double loss_plus = loss_sum();
    S.values_f[i] = p0 - delta;
    rc = dr_impl<T>(S, dr, cam, wv, pr, W, H, seed, ref, nthreads, 0, img.data(), scratch.data(), nullptr);
    if (rc) return rc;
    loss_minus = loss_sum();
    S.values_f[i] = p0;
    const double lp = loss_plus * g_mult, lm = loss_minus * g_mult;
    out_grad[i] += (double)(float)((lp - lm) / (2 * delta));
  }
  if (out_image)
    for (size_t k = 0; k < np * 4; k++) out_image[k] = img[k];
  if (stats) {
    stats->loss = loss_minus / (double)np;
    stats->border_pixels = 0;
    stats->border_rays_hit = 0;
    stats->rays = (uint64_t)np * dr.spp * 2 * (end - start);
  }
  return 0;
}

extern "C" int orc_dr_render_grad(const orc_scene* s, const MultiRendererDRPreset* dr, const float worldView[16], const float proj[16],
                                  uint32_t W, uint32_t H, uint32_t seed, const float* ref, int precision, int nthreads, int dtdv_mode,
                                  double* out_image, double* out_grad, orc_dr_stats* stats) {
  if (!s || !dr || !ref || !out_grad || W < 1 || H < 1 || dr->spp == 0) return -1;
  Camera cam = Camera::from(worldView, proj);
  if (dr->dr_diff_mode == LRT_DR_DIFF_MODE_FINITE_DIFF) {
    if (precision == 64) return dr_fd_impl<double>(s->s, *dr, cam, worldView, proj, W, H, seed, ref, nthreads, out_image, out_grad, stats);
    return dr_fd_impl<float>(s->s, *dr, cam, worldView, proj, W, H, seed, ref, nthreads, out_image, out_grad, stats);
  }
  if (precision == 64)
    return dr_impl<double>(s->s, *dr, cam, worldView, proj, W, H, seed, ref, nthreads, dtdv_mode, out_image, out_grad, stats);
  return dr_impl<float>(s->s, *dr, cam, worldView, proj, W, H, seed, ref, nthreads, dtdv_mode, out_image, out_grad, stats);
}

// MultiRendererDR::Regularization, MultiRendererDR.cpp:1933-2057: for bricks
// [0, n_nodes) and every distance value, the six in-brick neighbour terms
// lambda * power * sign(d) * |d|^(power-1), d = value - neighbour, added to
// grad (double) in the reference's order (float terms, as its buffers hold).
extern "C" int orc_dr_regularize(const orc_scene* s, uint32_t n_nodes, float lambda, float power, double* grad) {
  if (!s || !grad) return -1;
  const Scene& S = s->s;
  if (S.type != LRT_TYPE_SDF_SBS_COL || n_nodes > S.nodes.size()) return -4;
  const uint32_t vs = S.header.brick_size + 1;
  for (uint32_t b = 0; b < n_nodes; b++) {
    const uint32_t off = S.nodes[b].data_offset;
    for (uint32_t i0 = 0; i0 < vs; i0++)
      for (uint32_t j0 = 0; j0 < vs; j0++)
        for (uint32_t k0 = 0; k0 < vs; k0++) {
          const uint32_t idx = off + i0 * vs * vs + j0 * vs + k0;
          auto term = [&](uint32_t nbr) {
            float d = S.values_f[S.values[idx]] - S.values_f[S.values[nbr]];
            float sgn_d = d > 0.0f ? 1.0f : -1.0f;
            d = std::abs(d);
            grad[S.values[idx]] += (double)(lambda * power * sgn_d * std::pow(d, power - 1));
          };
          if (i0 > 0) term(idx - vs * vs);
          if (i0 < vs - 1) term(idx + vs * vs);
          if (j0 > 0) term(idx - vs);
          if (j0 < vs - 1) term(idx + vs);
          if (k0 > 0) term(idx - 1);
          if (k0 < vs - 1) term(idx + 1);
        }
  }
  return 0;
}

// SolveEikonal, MultiRendererDR.cpp:2058-2120 (the unit-speed branch)
static float solve_eikonal(float m[3], float h) {
  if (m[0] > m[1]) std::swap(m[0], m[1]);
  if (m[0] > m[2]) std::swap(m[0], m[2]);
  if (m[1] > m[2]) std::swap(m[1], m[2]);
  float dist_new = m[0] + h;
  if (dist_new > m[1]) {
    float h_2 = h * h;
    float m_sum = m[0] + m[1], c_2 = (m[0] - m[1]) * (m[0] - m[1]);
    dist_new = (m_sum + std::sqrt(2 * h_2 - c_2)) * 0.5f;
    if (dist_new > m[2]) {
      const float one_third = 1.f / 3;
      m_sum += m[2];
      c_2 = 3.f * (m[0] * m[0] + m[1] * m[1] + m[2] * m[2] - h_2);
      dist_new = (m_sum + std::sqrt(m_sum * m_sum - c_2)) * one_third;
    }
  }
  return dist_new;
}

// MultiRendererDR::Redistance, MultiRendererDR.cpp:2122-2382, as written:
// frozen interface points, then per iteration 8 sweeps over the anti-diagonal
// planes i + j + k = level (ascending, or descending for orderings 1, 4, 6, 7),
// each point updated from the other buffer.
extern "C" int orc_dr_redistance(float* d, uint32_t p, float h, uint32_t iters) {
  if (!d || p == 0) return -1;
  const size_t n = (size_t)p * p * p;
  const int off[3] = {(int)(p * p), (int)p, 1};
  const int sz[3] = {(int)p, (int)p, (int)p};
  const float big = 99999.f;
  std::vector<float> b[2] = {std::vector<float>(n, big), std::vector<float>(n, big)};
  std::vector<uint8_t> frozen(n, 0);
  for (int i = 0; i < sz[0]; i++)
    for (int j = 0; j < sz[1]; j++)
      for (int k = 0; k < sz[2]; k++) {
        const int c[3] = {i, j, k};
        const size_t lin = (size_t)i * off[0] + (size_t)j * off[1] + k;
        const float v = d[lin];
        if (v == 0.f) {
          frozen[lin] = 1;
          b[0][lin] = b[1][lin] = 0.f;
          continue;
        }
        for (int dim = 0; dim < 3; dim++)
          for (int sgn = -1; sgn < 2; sgn += 2) {
            int q[3] = {c[0], c[1], c[2]};
            q[dim] += sgn;
            if (q[0] < 0 || q[0] >= sz[0] || q[1] < 0 || q[1] >= sz[1] || q[2] < 0 || q[2] >= sz[2]) continue;
            if (d[(size_t)q[0] * off[0] + (size_t)q[1] * off[1] + q[2]] * v < 0) {
              b[0][lin] = b[1][lin] = v;
              frozen[lin] = 1;
            }
          }
      }
  int cur = 0;
  const int max_level = sz[0] + sz[1] + sz[2] - 3;
  for (uint32_t it = 0; it < iters; it++) {
    for (int ordering = 0; ordering < 8; ordering++) {
      const bool down = ordering == 1 || ordering == 4 || ordering == 6 || ordering == 7;
      for (int level = down ? max_level : 0; down ? level >= 0 : level <= max_level; level += down ? -1 : 1) {
        const int xs = std::max(0, level - (sz[1] + sz[2] - 2)), ys = std::max(0, level - (sz[0] + sz[2] - 2));
        const int xe = std::min(sz[0] - 1, level), ye = std::min(sz[1] - 1, level);
        for (int i = xs; i <= xe; i++)
          for (int j = ys; j <= ye; j++) {
            const int k = level - i - j;
            if (k < 0 || k >= sz[2]) continue;
            const int c[3] = {i, j, k};
            const size_t lin = (size_t)i * off[0] + (size_t)j * off[1] + k;
            const std::vector<float>& o = b[1 - cur];
            b[cur][lin] = o[lin];
            if (frozen[lin]) continue;
            float am[3] = {big, big, big};
            for (int dim = 0; dim < 3; dim++) {
              if (c[dim] > 0 && c[dim] < sz[dim] - 1) {
                am[dim] = o[lin - off[dim]];
                float t = o[lin + off[dim]];
                if (std::abs(t) < std::abs(am[dim])) am[dim] = t;
              } else if (c[dim] > 0 && c[dim] == sz[dim] - 1) {
                am[dim] = o[lin - off[dim]];
              } else if (c[dim] == 0 && c[dim] < sz[dim] - 1) {
                am[dim] = o[lin + off[dim]];
              }
            }
            if (am[0] < big || am[1] < big || am[2] < big) {
              if (am[0] >= 0.f && am[1] >= 0.f && am[2] >= 0.f) {
                b[cur][lin] = std::min(solve_eikonal(am, h), o[lin]);
              } else {
                for (int dim = 0; dim < 3; dim++)
                  if (am[dim] < 0.f) am[dim] = -am[dim];
                b[cur][lin] = std::max(-solve_eikonal(am, h), -std::abs(o[lin]));
              }
            }
          }
      }
    }
    cur = 1 - cur;
  }
  for (size_t i = 0; i < n; i++) d[i] = b[1 - cur][i];
  return 0;
}

// MultiRendererDR.cpp:828-853
extern "C" void orc_adam_step(uint32_t iter, const float* dX, float* X, float* V, float* S, uint32_t n, float lr, float beta1,
                              float beta2, float eps) {
  for (uint32_t i = 0; i < n; i++) {
    float g = dX[i];
    V[i] = beta1 * V[i] + (1 - beta1) * g;
    float Vh = (float)(V[i] / (1 - std::pow(beta1, iter + 1)));
    S[i] = beta2 * S[i] + (1 - beta2) * g * g;
    float Sh = (float)(S[i] / (1 - std::pow(beta2, iter + 1)));
    float x = X[i] - lr * Vh / (std::sqrt(Sh) + eps);
    X[i] = std::min(std::max(x, -1000.0f), 1000.0f);
  }
}

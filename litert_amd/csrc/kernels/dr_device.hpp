// dr_device.hpp - differentiable SBS ray casting (forward with payload) on gfx950.
//
// Same steps as BVHDR::RayQuery_NearestHitWithGrad / BVH2TraverseF32WithGrad
// / OctreeBrickIntersectWithGrad / Intersect (diff_render/BVH2DR.cpp:12-190,
// 623-1117, 1130-1610) for the Newton root find that MultiRendererDR forces,
// indexed (ID32F) layouts, and the partial derivatives of depth, diffuse
// colour and the hit normal. dt/dv is either the implicit-function-theorem
// derivative of the root (default) or the reference's central difference of
// the root find.
#pragma once
#include "traverse_device.hpp"

namespace lrtd {

struct DRHit {
  float t;
  uint32_t primId;
  v3 color;
  float sdf;
  v3 normal;
  uint32_t check_vox_border;
};

struct DRPayload {
  DRHit missed_hit, missed_cand;
  uint32_t missed_idx[8];
  float missed_dsdf[8];
  uint32_t pd_idx[8];    // distance parameters of the hit voxel
  float pd_ddist[8];     // d(tReal)/dS
  v3 pd_ddiffuse[8];     // dDiffuse/dS
  uint32_t col_idx[8];   // colour parameter base indices of the hit brick
  float col_w[8];        // dDiffuse/dColour (trilinear weight)
  // kDrNormal only: with scatter set, the hit adds its dLoss/dS terms to grad
  // directly (dLoss/dColour dldc of the pixel is known: second pass)
  float* grad;
  v3 dldc;
  uint32_t mode, scatter;
  const LightSet* lights;
};

struct DRParams {
  uint32_t ray_flags;
  uint32_t dtdv_fd;  // 1 = central difference (reference), 0 = analytic
};
static_assert(sizeof(DRParams) == 8, "DRArgs::Q layout");

// Compile-time forms of the DR ray casts. kDrGeneric reads the ray flags and
// the dt/dv mode at run time and computes every payload part. kDrDepth is the
// primary pass of linear-depth reconstruction (flags = DDIST_DPOS, analytic
// dt/dv, colour unused); kDrDepthBorder the border pass of linear-depth input
// (flags = BORDER, colour unused). The specialised forms evaluate the same
// expressions for what they keep and compile the rest of the payload out.
// kDrNormal: the render modes that read the hit normal (LAMBERT, NORMAL):
// run-time flags, colour, the normal from the hit voxel or smoothed over its
// neighbours, and the normal's derivatives scattered at the hit.
enum : int { kDrGeneric = 0, kDrDepth = 1, kDrDepthBorder = 2, kDrNormal = 3 };

// BVH copy of the linear-depth border pass: escapeIndex of a slot whose
// subtree holds only bricks that pass skips (k_dr_cull_write), and the leaf
// value the traversal gives such a slot when the ray crosses it.
constexpr uint32_t kDrCulledSlot = 0xFFFFFFFEu;
constexpr uint32_t kDrResetLeaf = 0xFFFFFFFEu;

template <int SPEC>
__device__ __forceinline__ uint32_t dr_flags(const DRParams& Q) {
  return SPEC == kDrDepth ? LRT_DR_RAY_FLAG_DDIST_DPOS : SPEC == kDrDepthBorder ? LRT_DR_RAY_FLAG_BORDER : Q.ray_flags;
}

// BVH2DR.cpp:623-1073, Newton branch. Returns hit ? t : qFar + EPS.
__device__ __forceinline__ float dr_intersect(uint32_t ray_flags, v3 ray_dir, const float values[8], float d, float d_inv, float qFar,
                                              v3 start_q, DRPayload* relax) {
  const float EPS = 1e-6f, SDF_EPS = 1e-7f;
  float t = 0.0f;
  bool hit = false;
  if (relax) relax->missed_cand.sdf = 1000.f;
  float start_dist = trilinear(values, start_q + t * ray_dir);
  if (start_dist <= SDF_EPS) {
    hit = true;
  } else {
    float s000 = values[0] * d_inv, s001 = values[1] * d_inv, s010 = values[2] * d_inv, s011 = values[3] * d_inv;
    float s100 = values[4] * d_inv, s101 = values[5] * d_inv, s110 = values[6] * d_inv, s111 = values[7] * d_inv;
    float a = s101 - s001;
    float k0 = s000, k1 = s100 - s000, k2 = s010 - s000, k3 = s110 - s010 - k1, k4 = k0 - s001, k5 = k1 - a;
    float k6 = k2 - (s011 - s001), k7 = k3 - (s111 - s011 - a);
    v3 o = start_q, d3 = ray_dir;
    float m0 = o.x * o.y, m1 = d3.x * d3.y, m2 = o.x * d3.y + o.y * d3.x;
    float m3 = k5 * o.z - k1, m4 = k6 * o.z - k2, m5 = k7 * o.z - k3;
    float c0 = (k4 * o.z - k0) + o.x * m3 + o.y * m4 + m0 * m5;
    float c1 = d3.x * m3 + d3.y * m4 + m2 * m5 + d3.z * (k4 + k5 * o.x + k6 * o.y + k7 * m0);
    float c2 = m1 * m5 + d3.z * (k5 * d3.x + k6 * d3.y + k7 * m2);
    float c3 = k7 * m1 * d3.z;
    float qa = 3 * c3, qb = 2 * c2, qc = c1;
    float t0 = 0, t1 = qFar, t2 = qFar, t3 = qFar;
    float D = qb * qb - 4 * qa * qc;
    if (D >= 0) {
      float q = -0.5f * (qb + sgn(qb) * sqrtf(D));
      t1 = fabsf(qa) > EPS ? q / qa : t0;
      t2 = fabsf(q) > EPS ? qc / q : qFar;
      float tmp = mn(t1, t2);
      t2 = mx(t1, t2);
      t1 = tmp;
      t1 = clampf_(t1, t0, t3);
      t2 = clampf_(t2, t0, t3);
    }
    float sdf0 = c0;
    float sdf1 = (c0 + t1 * (c1 + t1 * (c2 + t1 * c3)));
    float sdf2 = (c0 + t2 * (c1 + t2 * (c2 + t2 * c3)));
    float sdf3 = (c0 + t3 * (c1 + t3 * (c2 + t3 * c3)));
    bool s0 = sdf0 > 0, s1 = sdf1 > 0, s2 = sdf2 > 0, s3 = sdf3 > 0;
    float nwt_min = t0, nwt_max = t0;
    if (s0 != s1) { nwt_min = t0; nwt_max = t1; }
    else if (s1 != s2) { nwt_min = t1; nwt_max = t2; }
    else if (s2 != s3) { nwt_min = t2; nwt_max = t3; }
    if (nwt_min < nwt_max) {
      unsigned it = 0;
      float rtn = 0.5f * (nwt_min + nwt_max);
      float f = 1000;
      while (it < 10u && fabsf(f) >= SDF_EPS) {
        f = c0 + rtn * (c1 + rtn * (c2 + rtn * c3));
        float df = c1 + rtn * (2 * c2 + rtn * 3 * c3);
        float dx = f / (df + sgn(df) * 1e-9f);
        rtn -= dx;
        it++;
      }
      t = rtn;
      hit = (t >= 0 && t <= qFar && fabsf(f) < SDF_EPS);
    }
    if (relax && (ray_flags & LRT_DR_RAY_FLAG_BORDER)) {  // BVH2DR.cpp:960-1035
      float t_min = qFar + EPS;
      if (fabsf(qa) > EPS) {
        float root1 = -0.5f * (qb + sqrtf(D)) / qa, root2 = -0.5f * (qb - sqrtf(D)) / qa;
        t_min = (sgn(qa) > 0) == (root1 < root2) ? root1 : root2;
      } else if (qb < -EPS) {
        t_min = -qc / qb;
      }
      float sdf_min = (c0 + t_min * (c1 + t_min * (c2 + t_min * c3)));
      DRHit& mh = relax->missed_hit;
      if (mh.check_vox_border == 1) {
        if (c1 <= 0) {
          if (sdf0 < 0.f && -sdf0 * d < mh.sdf) { mh.t = t0; mh.sdf = -sdf0 * d; }
        }
        mh.check_vox_border = 0;
      }
      if (t_min >= 0 && t_min <= qFar && sdf_min < 0.f && -sdf_min * d < mh.sdf) { mh.t = t_min; mh.sdf = -sdf_min * d; }
      if (sdf3 < 0.f && -sdf3 * d < mh.sdf && (c1 + qFar * (2 * c2 + qFar * 3 * c3)) >= 0.f) {
        mh.check_vox_border = 1;
        relax->missed_cand.t = qFar;
        relax->missed_cand.sdf = -sdf3 * d;
      }
    }
  }
  return hit ? t : qFar + EPS;
}

__device__ __forceinline__ float corner_w(int i, v3 q) {
  return ((i & 4) ? q.x : 1 - q.x) * ((i & 2) ? q.y : 1 - q.y) * ((i & 1) ? q.z : 1 - q.z);
}

// 3x3 matrix as rows; m * v = the row dot products
struct m3 {
  v3 r[3];
};
__device__ __forceinline__ v3 mul(const m3& m, v3 v) { return v3(dot(m.r[0], v), dot(m.r[1], v), dot(m.r[2], v)); }

// BVH2DR.cpp:339-404: Jacobian of v / |v|
__device__ __forceinline__ m3 normalize_diff(v3 v) {
  const float a = dot(v, v);
  const float b = 1.0f / (a * sqrtf(a));
  m3 m;
  m.r[0] = b * v3(v.y * v.y + v.z * v.z, -v.x * v.y, -v.x * v.z);
  m.r[1] = b * v3(-v.x * v.y, v.x * v.x + v.z * v.z, -v.y * v.z);
  m.r[2] = b * v3(-v.x * v.z, -v.y * v.z, v.x * v.x + v.y * v.y);
  return m;
}

// BVH2DR.cpp:406-530: Hessian of the trilinear distance times a direction
__device__ __forceinline__ v3 trilinear_hess_dir(const float v[8], v3 dp, v3 dir) {
  const float xy = (1 - dp.z) * v[0] + (dp.z) * v[1] + -(1 - dp.z) * v[2] + -(dp.z) * v[3] + -(1 - dp.z) * v[4] + -(dp.z) * v[5] +
                   (1 - dp.z) * v[6] + (dp.z) * v[7];
  const float xz = (1 - dp.y) * v[0] + -(1 - dp.y) * v[1] + (dp.y) * v[2] + -(dp.y) * v[3] + -(1 - dp.y) * v[4] + (1 - dp.y) * v[5] +
                   -(dp.y) * v[6] + (dp.y) * v[7];
  const float yz = (1 - dp.x) * v[0] + -(1 - dp.x) * v[1] + -(1 - dp.x) * v[2] + (1 - dp.x) * v[3] + (dp.x) * v[4] + -(dp.x) * v[5] +
                   -(dp.x) * v[6] + (dp.x) * v[7];
  m3 h;
  h.r[0] = v3(0.0f, xy, xz);
  h.r[1] = v3(xy, 0.0f, yz);
  h.r[2] = v3(xz, yz, 0.0f);
  return mul(h, dir);
}

// BVH2DR.cpp:531-611: column c of d(grad f)/d(values)
__device__ __forceinline__ v3 trilinear_ddp_dvalue(int c, v3 dp) {
  const float sx = (c & 4) ? 1.0f : -1.0f, sy = (c & 2) ? 1.0f : -1.0f, sz = (c & 1) ? 1.0f : -1.0f;
  const float wx = (c & 4) ? dp.x : 1 - dp.x, wy = (c & 2) ? dp.y : 1 - dp.y, wz = (c & 1) ? dp.z : 1 - dp.z;
  return v3((sx * wy) * wz, (sy * wx) * wz, (sz * wx) * wy);
}

// CalculateColorWithGrad (MultiRendererDR.cpp:900-980) for LAMBERT and NORMAL;
// LAMBERT sums over the renderer's direct and ambient lights (point lights are
// refused on the host, as the reference asserts).
__device__ __forceinline__ void dr_color_grad(uint32_t mode, const DRHit& h, const LightSet& L, m3& dCdD, m3& dCdN) {
  const v3 z(0, 0, 0);
  dCdD.r[0] = dCdD.r[1] = dCdD.r[2] = z;
  dCdN.r[0] = dCdN.r[1] = dCdN.r[2] = z;
  if (mode == LRT_DR_RENDER_MODE_NORMAL) {
    dCdN.r[0] = v3(1, 0, 0);
    dCdN.r[1] = v3(0, 1, 0);
    dCdN.r[2] = v3(0, 0, 1);
    return;
  }
  for (uint32_t i = 0; i < L.n; i++) {
    const v3 lc(L.col[i].x, L.col[i].y, L.col[i].z);
    if (__float_as_uint(L.sp[i].w) == LRT_LIGHT_TYPE_DIRECT) {
      const v3 ld(L.sp[i].x, L.sp[i].y, L.sp[i].z);
      const float q = fmaxf(0.0f, dot(h.normal, ld));
      if (q > 0.0f) {
        dCdD.r[0] = dCdD.r[0] + v3(lc.x * q, 0, 0);
        dCdD.r[1] = dCdD.r[1] + v3(0, lc.y * q, 0);
        dCdD.r[2] = dCdD.r[2] + v3(0, 0, lc.z * q);
        dCdN.r[0] = dCdN.r[0] + (lc.x * h.color.x) * ld;
        dCdN.r[1] = dCdN.r[1] + (lc.y * h.color.y) * ld;
        dCdN.r[2] = dCdN.r[2] + (lc.z * h.color.z) * ld;
      }
    } else {  // ambient
      dCdD.r[0] = dCdD.r[0] + v3(lc.x, 0, 0);
      dCdD.r[1] = dCdD.r[1] + v3(0, lc.y, 0);
      dCdD.r[2] = dCdD.r[2] + v3(0, 0, lc.z);
    }
  }
}

// The hit of a kDrNormal ray cast (BVH2DR.cpp:1277-1524 with the pixel's
// accumulation, MultiRendererDR.cpp:1148-1265): dt/dv, the diffuse colour
// and its derivatives, the normal from the hit voxel or, for the _IN layout,
// smoothed over the 8 voxels around the hit, and its derivatives. With
// relax->scatter set the dLoss/dS terms are added to relax->grad here, one
// atomic per payload entry of the reference, instead of being kept in a
// 64-entry payload. values / vx,vy,vz: the hit voxel; start_q, t, qFar: its
// segment and root; bmin, sz, d: the brick.
__device__ __noinline__ void dr_normal_hit(const DevScene& S, uint32_t flags, bool dtdv_fd, v3 ray_pos, v3 ray_dir, const float values[8],
                                           uint32_t brick, uint32_t vx, uint32_t vy, uint32_t vz, v3 start_q, float t, float qFar,
                                           v3 bmin, float sz, float d, DRPayload* relax, DRHit* pHit) {
  const uint4 N = S.nodes[brick];
  const uint32_t vs = S.v_size;
  auto vid = [&](uint32_t x, uint32_t y, uint32_t z, int i) {
    return (x + ((i & 4) >> 2)) * vs * vs + (y + ((i & 2) >> 1)) * vs + (z + (i & 1));
  };
  float dtdv[8];
#pragma unroll
  for (int i = 0; i < 8; i++) dtdv[i] = 0.0f;
  const float d_inv = 1.0f / d;
  if (flags & (LRT_DR_RAY_FLAG_DDIFFUSE_DPOS | LRT_DR_RAY_FLAG_DNORM_DPOS | LRT_DR_RAY_FLAG_DDIST_DPOS)) {
    if (dtdv_fd) {  // BVH2DR.cpp:1075-1090
      float vv[8];
      for (int i = 0; i < 8; i++) vv[i] = values[i];
      const float delta = 0.001f;
      for (int i = 0; i < 8; i++) {
        vv[i] += delta;
        float d1 = dr_intersect(flags, ray_dir, vv, d, d_inv, qFar, start_q, nullptr);
        vv[i] -= 2 * delta;
        float d2 = dr_intersect(flags, ray_dir, vv, d, d_inv, qFar, start_q, nullptr);
        vv[i] += delta;
        dtdv[i] = (d1 - d2) / (2 * delta);
      }
    } else {
      v3 q = start_q + t * ray_dir;
      float fd = dot(trilinear_diff(values, q), ray_dir);
      for (int i = 0; i < 8; i++) dtdv[i] = fd != 0.0f ? -corner_w(i, q) / fd : 0.0f;
    }
  }
  // diffuse colour (BVH2DR.cpp:1489-1524)
  const v3 pos = ray_pos + pHit->t * ray_dir;
  const v3 dpc = (pos - bmin) * (0.5f * sz);
  const uint32_t t_off = N.z + vs * vs * vs;
  v3 col(0, 0, 0), J0(0, 0, 0), J1(0, 0, 0), J2(0, 0, 0);
  for (int i = 0; i < 8; i++) {
    uint32_t ci = S.values[t_off + i];
    v3 c(S.values_f[ci], S.values_f[ci + 1], S.values_f[ci + 2]);
    float wx = (i & 4) ? dpc.x : 1 - dpc.x, wy = (i & 2) ? dpc.y : 1 - dpc.y, wz = (i & 1) ? dpc.z : 1 - dpc.z;
    col = col + (wx * wy * wz) * c;
    float ax = (i & 4) ? 1.0f : -1.0f, ay = (i & 2) ? 1.0f : -1.0f, az = (i & 1) ? 1.0f : -1.0f;
    J0 = J0 + (ax * wy * wz) * c;
    J1 = J1 + (wx * ay * wz) * c;
    J2 = J2 + (wx * wy * az) * c;
  }
  pHit->color = col;
  const bool smoothed = S.layout == LRT_SBS_LAYOUT_ID32F_IRGB32F_IN;
  const bool dn = relax->scatter && (flags & LRT_DR_RAY_FLAG_DNORM_DPOS);
  const v3 p0 = start_q + t * ray_dir;
  if (!smoothed) {
    const v3 g = trilinear_diff(values, p0);
    pHit->normal = normalize(g);
    if (!relax->scatter) return;
    m3 dCdD, dCdN;
    dr_color_grad(relax->mode, *pHit, *relax->lights, dCdD, dCdN);
    const v3 dl = relax->dldc;
    if (flags & LRT_DR_RAY_FLAG_DDIFFUSE_DCOLOR) {
      for (int i = 0; i < 8; i++) {
        const uint32_t ci = S.values[t_off + i];
        const float w = corner_w(i, dpc);
        const v3 dc = mul(dCdD, v3(w, w, w));
        atomicAdd(&relax->grad[ci + 0], dl.x * dc.x);
        atomicAdd(&relax->grad[ci + 1], dl.y * dc.y);
        atomicAdd(&relax->grad[ci + 2], dl.z * dc.z);
      }
    }
    if (flags & (LRT_DR_RAY_FLAG_DDIFFUSE_DPOS | LRT_DR_RAY_FLAG_DNORM_DPOS)) {
      m3 nd;
      v3 t1(0, 0, 0);
      if (dn) {
        nd = normalize_diff(g);
        t1 = trilinear_hess_dir(values, p0, ray_dir);
      }
      for (int i = 0; i < 8; i++) {
        v3 ddiff(0, 0, 0), dnorm(0, 0, 0);
        if (flags & LRT_DR_RAY_FLAG_DDIFFUSE_DPOS) {
          v3 v(0.5f * sz * ray_dir.x * dtdv[i], 0.5f * sz * ray_dir.y * dtdv[i], 0.5f * sz * ray_dir.z * dtdv[i]);
          ddiff = J0 * v.x + J1 * v.y + J2 * v.z;
        }
        if (dn) {
          v3 c = trilinear_ddp_dvalue(i, p0);
          dnorm = mul(nd, v3(c.x + t1.x * dtdv[i], c.y + t1.y * dtdv[i], c.z + t1.z * dtdv[i]));
        }
        atomicAdd(&relax->grad[S.values[N.z + vid(vx, vy, vz, i)]], dot(dl, mul(dCdD, ddiff) + mul(dCdN, dnorm)));
      }
    } else if (flags & LRT_DR_RAY_FLAG_DDIST_DPOS) {
      for (int i = 0; i < 8; i++) {
        float v = d * dtdv[i] / (10.0f - 0.1f);
        atomicAdd(&relax->grad[S.values[N.z + vid(vx, vy, vz, i)]], dot(dl, v3(v, v, v)));
      }
    }
    return;
  }
  // smoothed normal over the 8 voxels around the hit (BVH2DR.cpp:1305-1437)
  const float beta = 0.5f;
  const v3 dp = p0;
  v3 nmq;
  nmq.x = dp.x < 0.5f ? step01(-beta, beta, dp.x) : step01(-beta, beta, dp.x - 1.0f);
  nmq.y = dp.y < 0.5f ? step01(-beta, beta, dp.y) : step01(-beta, beta, dp.y - 1.0f);
  nmq.z = dp.z < 0.5f ? step01(-beta, beta, dp.z) : step01(-beta, beta, dp.z - 1.0f);
  const int bs = (int)S.brick_size;
  const v3 vp((float)vx, (float)vy, (float)vz);
  const int vp0[3] = {dp.x < 0.5f ? (int)vx - 1 : (int)vx, dp.y < 0.5f ? (int)vy - 1 : (int)vy, dp.z < 0.5f ? (int)vz - 1 : (int)vz};
  const uint32_t k = (dp.x < 0.5f ? 4u : 0u) + (dp.y < 0.5f ? 2u : 0u) + (dp.z < 0.5f ? 1u : 0u);
  // neighbour i of the hit: its brick (0xFFFFFFFF if absent) and voxel
  auto neighbour = [&](int i, uint32_t& nb, uint32_t nv[3], v3& dvp) {
    int VI[3] = {vp0[0] + ((i & 4) >> 2), vp0[1] + ((i & 2) >> 1), vp0[2] + (i & 1)};
    int BP[3];
    for (int a = 0; a < 3; a++) BP[a] = VI[a] >= 0 ? (VI[a] < bs ? 0 : 1) : -1;
    uint32_t nid = 9 * (BP[0] + 1) + 3 * (BP[1] + 1) + (BP[2] + 1);
    dvp = v3((float)VI[0], (float)VI[1], (float)VI[2]) - vp;
    nb = brick;
    if (nid == 13) {
      for (int a = 0; a < 3; a++) nv[a] = (uint32_t)VI[a];
    } else {
      nb = S.values[N.z + vs * vs * vs + 8 + nid];
      for (int a = 0; a < 3; a++) nv[a] = (uint32_t)(VI[a] - bs * BP[a]);
    }
  };
  v3 normals[8];
  for (int i = 0; i < 8; i++) {
    uint32_t nb, nv[3];
    v3 dvp;
    neighbour(i, nb, nv, dvp);
    normals[i] = v3(0, 0, 0);
    if (nb == LRT_INVALID_IDX) continue;
    const uint32_t off = S.nodes[nb].z;
    float vn[8];
    for (int c = 0; c < 8; c++) vn[c] = S.values_f[S.values[off + vid(nv[0], nv[1], nv[2], c)]];
    normals[i] = normalize(trilinear_diff(vn, dp - dvp));
  }
  v3 tri = corner_w(0, nmq) * normals[0];
  for (int i = 1; i < 8; i++) tri = tri + corner_w(i, nmq) * normals[i];
  pHit->normal = normalize(tri);
  if (!relax->scatter) return;
  m3 dCdD, dCdN;
  dr_color_grad(relax->mode, *pHit, *relax->lights, dCdD, dCdN);
  const v3 dl = relax->dldc;
  if (flags & LRT_DR_RAY_FLAG_DDIFFUSE_DCOLOR) {
    for (int i = 0; i < 8; i++) {
      const uint32_t ci = S.values[t_off + i];
      const float w = corner_w(i, dpc);
      const v3 dc = mul(dCdD, v3(w, w, w));
      atomicAdd(&relax->grad[ci + 0], dl.x * dc.x);
      atomicAdd(&relax->grad[ci + 1], dl.y * dc.y);
      atomicAdd(&relax->grad[ci + 2], dl.z * dc.z);
    }
  }
  if (!(flags & (LRT_DR_RAY_FLAG_DDIFFUSE_DPOS | LRT_DR_RAY_FLAG_DNORM_DPOS))) {
    if (flags & LRT_DR_RAY_FLAG_DDIST_DPOS)
      for (int i = 0; i < 8; i++) {
        float v = d * dtdv[i] / (10.0f - 0.1f);
        atomicAdd(&relax->grad[S.values[N.z + vid(vx, vy, vz, i)]], dot(dl, v3(v, v, v)));
      }
    return;
  }
  // dNorm of the hit voxel's corners accumulates over the neighbours; the
  // neighbours' own corners are scattered as they are reached
  v3 acc[8];
  for (int c = 0; c < 8; c++) acc[c] = v3(0, 0, 0);
  const m3 M = normalize_diff(tri);
  if (dn) {
    for (int i = 0; i < 8; i++) {
      uint32_t nb, nv[3];
      v3 dvp;
      neighbour(i, nb, nv, dvp);
      if (nb == LRT_INVALID_IDX) continue;
      const uint32_t off = S.nodes[nb].z;
      float vn[8];
      uint32_t in[8];
      for (int c = 0; c < 8; c++) {
        in[c] = S.values[off + vid(nv[0], nv[1], nv[2], c)];
        vn[c] = S.values_f[in[c]];
      }
      const v3 q = dp - dvp;
      const v3 g = trilinear_diff(vn, q);
      const m3 nd = normalize_diff(g);
      const v3 t1 = trilinear_hess_dir(vn, q, ray_dir);
      const float wi = corner_w(i, nmq);
      const bool valid = sqrtf(dot(normals[i], normals[i])) > 0.5f;
      for (int c = 0; c < 8; c++) {  // d(n_i)/d(v_i)
        const v3 e = wi * mul(nd, trilinear_ddp_dvalue(c, q));
        if ((uint32_t)i == k) {
          acc[c] = acc[c] + e;
        } else {
          const v3 dnorm = valid ? mul(M, v3(0, 0, 0) + e) : v3(0, 0, 0) + e;
          atomicAdd(&relax->grad[in[c]], dot(dl, mul(dCdD, v3(0, 0, 0)) + mul(dCdN, dnorm)));
        }
      }
      for (int c = 0; c < 8; c++)  // d(n_i)/d(v_k) through the hit point
        acc[c] = acc[c] + wi * mul(nd, v3(t1.x * dtdv[c], t1.y * dtdv[c], t1.z * dtdv[c]));
    }
    // dtrilinear3/dnmq (by columns) times dp/dv_k
    v3 cx, cy, cz;
    for (int i = 0; i < 8; i++) {
      const float ax = (i & 4) ? 1.0f : -1.0f, ay = (i & 2) ? 1.0f : -1.0f, az = (i & 1) ? 1.0f : -1.0f;
      const float wx = (i & 4) ? nmq.x : 1 - nmq.x, wy = (i & 2) ? nmq.y : 1 - nmq.y, wz = (i & 1) ? nmq.z : 1 - nmq.z;
      const v3 ex = (ax * wy * wz) * normals[i], ey = (wx * ay * wz) * normals[i], ez = (wx * wy * az) * normals[i];
      cx = i == 0 ? ex : cx + ex;
      cy = i == 0 ? ey : cy + ey;
      cz = i == 0 ? ez : cz + ez;
    }
    for (int c = 0; c < 8; c++) {
      const v3 ddp(ray_dir.x * dtdv[c], ray_dir.y * dtdv[c], ray_dir.z * dtdv[c]);
      acc[c] = acc[c] + (cx * ddp.x + cy * ddp.y + cz * ddp.z);
    }
    if (sqrtf(dot(normals[k], normals[k])) > 0.5f)
      for (int c = 0; c < 8; c++) acc[c] = mul(M, acc[c]);
  }
  for (int c = 0; c < 8; c++) {  // the hit voxel's corners
    v3 ddiff(0, 0, 0);
    if (flags & LRT_DR_RAY_FLAG_DDIFFUSE_DPOS) {
      v3 v(0.5f * sz * ray_dir.x * dtdv[c], 0.5f * sz * ray_dir.y * dtdv[c], 0.5f * sz * ray_dir.z * dtdv[c]);
      ddiff = J0 * v.x + J1 * v.y + J2 * v.z;
    }
    atomicAdd(&relax->grad[S.values[N.z + vid(vx, vy, vz, c)]], dot(dl, mul(dCdD, ddiff) + mul(dCdN, acc[c])));
  }
}

// BVH2DR.cpp:1130-1610 for ID32F layouts.
template <int SPEC>
__device__ __forceinline__ void dr_brick(const DevScene& S, const float* __restrict__ bvals, const float2* __restrict__ bminmax,
                                         const DRParams& Q, float relax_eps, v3 ray_pos, v3 ray_dir, v3 inv, float tNear, uint32_t brick,
                                         DRPayload* relax, DRHit* pHit) {
  constexpr bool kColour = SPEC == kDrGeneric || SPEC == kDrNormal;
  float values[8];
  const uint4 N = S.nodes[brick];
  const float px = (float)(N.x >> 16), py = (float)(N.x & 0xFFFFu), pz = (float)(N.y >> 16), sz = (float)(N.y & 0xFFFFu);
  const float bsf = (float)S.brick_size;
  const float sz_inv = 2.0f / sz;
  const float d = 2.0f / (sz * bsf);
  const float d_inv = 1.0f / d;
  const v3 bmin = v3(-1, -1, -1) + sz_inv * v3(px, py, pz);
  const v3 bmax = bmin + sz_inv * v3(1, 1, 1);
  const float qscale = 0.5f * sz * bsf;
  const uint32_t vs = S.v_size;
  // the bricks' corner values gathered from values_f (k_dr_gather), vs^3 floats per brick
  const __amdgpu_buffer_rsrc_t rv = buf_rsrc(bvals, S.n_nodes * vs * vs * vs * 4u);
  const uint32_t bbase = brick * vs * vs * vs * 4u;  // byte offset of the brick's copy
  const bool ref = wave_ref_slabs(inv);
  float2 bnf = RayBoxW(ref, ray_pos, inv, bmin, bmax);
  if (SPEC == kDrDepth || SPEC == kDrDepthBorder) {
    // Whole-brick form of the voxel skip below: the brick's smallest and
    // largest corner values bound every voxel's, so if they pass the test
    // every voxel of the walk would be skipped. Border rays apply the skip's
    // state resets when the walk would have reached a voxel: it starts
    // (bnf.x < bnf.y) past tNear (1e-5 margin over the few-ulp difference
    // between the brick's and its first voxel's entry distance).
    const float2 mm = bminmax[brick];
    bool skip = mm.x > 1e-6f && mm.x > 1e-3f * mm.y;
    if (SPEC == kDrDepthBorder) {
      skip = skip && (mm.x > relax->missed_hit.sdf + 1e-3f * mm.y || mm.x > relax_eps + 1e-3f * mm.y) &&
             (!(bnf.x < bnf.y) || bnf.x > tNear + 1e-5f);
      if (skip && bnf.x < bnf.y) {
        relax->missed_cand.sdf = 1000.f;
        relax->missed_hit.check_vox_border = 0;
      }
    }
    if (skip) return;
  }
  const float old_t = pHit->t;
  const uint32_t max_steps = 16u * (S.brick_size + 1u) + 16u;
  uint32_t steps = 0;
  const uint32_t flags = dr_flags<SPEC>(Q);
  const bool dtdv_fd = SPEC == kDrGeneric && Q.dtdv_fd != 0;
  while (bnf.x < bnf.y && pHit->t == old_t && steps++ < max_steps) {
    v3 hit_pos = ray_pos + bnf.x * ray_dir;
    v3 local = (hit_pos - bmin) * qscale;
    const float hi = bsf - 1e-6f;
    // clampf_ of finite values: the median of value and bounds (v_med3)
    v3 vp(floorf(__builtin_amdgcn_fmed3f(local.x, 1e-6f, hi)), floorf(__builtin_amdgcn_fmed3f(local.y, 1e-6f, hi)),
          floorf(__builtin_amdgcn_fmed3f(local.z, 1e-6f, hi)));
    const uint32_t vx = (uint32_t)vp.x, vy = (uint32_t)vp.y, vz = (uint32_t)vp.z;
    v3 min_pos = bmin + d * vp;
    v3 max_pos = min_pos + d * v3(1, 1, 1);
    // corner i of the voxel: values[i] = values_f[values[data_offset + vid(i)]],
    // read from the brick's gathered copy: one vector offset, the y / x
    // strides as scalar offsets, the z step as an immediate
    auto vid = [&](int i) { return (vx + ((i & 4) >> 2)) * vs * vs + (vy + ((i & 2) >> 1)) * vs + (vz + (i & 1)); };
    {
      // element (vx * vs + vy) * vs + vz from the float voxel coordinates:
      // small integers (brick_size <= 16), exact in float, one conversion
      const uint32_t o = bbase + ((uint32_t)fmaf(fmaf(vp.x, S.v_size_f, vp.y), S.v_size_f, vp.z) << 2), sy = vs * 4u, sx = vs * vs * 4u,
                     sxy = sx + sy;
      values[0] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rv, o, 0, 0));
      values[1] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rv, o + 4, 0, 0));
      values[2] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rv, o, sy, 0));
      values[3] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rv, o + 4, sy, 0));
      values[4] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rv, o, sx, 0));
      values[5] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rv, o + 4, sx, 0));
      values[6] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rv, o, sxy, 0));
      values[7] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rv, o + 4, sxy, 0));
    }
    float2 fnf = RayBoxW(ref, ray_pos, inv, min_pos, max_pos);
    bool candidate = tNear < fnf.x;
    if ((SPEC == kDrDepth || SPEC == kDrDepthBorder) && candidate) {
      // Along the segment the trilinear distance is a convex combination of
      // the 8 corners, so with every corner above 1e-6 and above 1e-3 of the
      // largest one it stays far above SDF_EPS, and the FP32 cubic's rounding
      // (~1e-5 of the largest coefficient) cannot flip a sign: dr_intersect
      // would return "no hit". The reference evaluates such voxels anyway
      // (BVH2DR.cpp:1199); skipping them changes no result.
      float vmin = values[0], vmax = values[0];
#pragma unroll
      for (int i = 1; i < 8; i++) {
        vmin = fminf(vmin, values[i]);
        vmax = fmaxf(vmax, values[i]);
      }
      bool skip = vmin > 1e-6f && vmin > 1e-3f * vmax;
      if (SPEC == kDrDepthBorder) {
        // Border rays also track the smallest distance along the ray
        // (missed_hit.sdf). Every update in dr_intersect needs a distance
        // below it; if none along this segment is (same bound), the voxel's
        // only effects are the two state resets done here. A border ray also
        // contributes only if that minimum ends below relax_eps, and the
        // state that decides it (point, voxel, derivatives) is set by the
        // update that first goes below relax_eps. A voxel whose distances all
        // exceed relax_eps (same margin) offers only values above it, so it is
        // skipped as well; the tracked minimum then differs only above
        // relax_eps, which can change the outcome only where a distance the
        // next voxel recomputes at the shared face lands within FP32 rounding
        // of relax_eps.
        skip = skip && (vmin > relax->missed_hit.sdf + 1e-3f * vmax || vmin > relax_eps + 1e-3f * vmax);
        if (skip) {
          relax->missed_cand.sdf = 1000.f;
          relax->missed_hit.check_vox_border = 0;
        }
      }
      candidate = !skip;
    }
    if (candidate) {
      v3 start_pos = ray_pos + fnf.x * ray_dir;
      v3 start_q = (start_pos - min_pos) * qscale;
      float qFar = (fnf.y - fnf.x) * qscale;
      float prev_sdf = relax->missed_hit.sdf;
      float t = dr_intersect(flags, ray_dir, values, d, d_inv, qFar, start_q, relax);
      float tReal = fnf.x + d * t;
      if (flags & LRT_DR_RAY_FLAG_BORDER) {  // BVH2DR.cpp:1223-1275
        if (relax->missed_hit.sdf < prev_sdf && relax->missed_hit.t < t) {
          float missed_t = relax->missed_hit.t;
          relax->missed_hit.t = fnf.x + d * relax->missed_hit.t;
          v3 q_ast = start_q + missed_t * ray_dir;
          v3 y_ast = ray_pos + (fnf.x + (d * missed_t)) * ray_dir;
          v3 dp_ast = (y_ast - bmin) * (0.5f * sz);
          v3 g = trilinear_diff(values, q_ast) * (1.0f / d);
          v3 mul(fabsf(q_ast.x - 0.5f) > 0.5f - 1e-5f ? 0.0f : 1.0f, fabsf(q_ast.y - 0.5f) > 0.5f - 1e-5f ? 0.0f : 1.0f,
                 fabsf(q_ast.z - 0.5f) > 0.5f - 1e-5f ? 0.0f : 1.0f);
          g = g * mul;
          float glen = sqrtf(dot(g, g));
          if (glen > 1e-9f) {
            const uint32_t t_off = N.z + vs * vs * vs;
            v3 col(0, 0, 0);
#pragma unroll
            for (int i = 0; i < 8; i++) {
              relax->missed_dsdf[i] = -corner_w(i, q_ast) / glen;
              if (kColour) {
                uint32_t ci = S.values[t_off + i];
                col = col + corner_w(i, dp_ast) * v3(S.values_f[ci], S.values_f[ci + 1], S.values_f[ci + 2]);
              }
              relax->missed_idx[i] = S.values[N.z + vid(i)];
            }
            relax->missed_hit.color = col;
            relax->missed_hit.normal = v3(g.x / glen, g.y / glen, g.z / glen);
          }
        }
      }
      if (t <= qFar && tReal < pHit->t) {  // BVH2DR.cpp:1277-1600
        pHit->t = tReal;
        pHit->primId = brick;
        if (SPEC == kDrNormal) {
          dr_normal_hit(S, flags, Q.dtdv_fd != 0, ray_pos, ray_dir, values, brick, vx, vy, vz, start_q, t, qFar, bmin, sz, d, relax, pHit);
          break;
        }
        float dtdv[8];
        if (flags & (LRT_DR_RAY_FLAG_DDIFFUSE_DPOS | LRT_DR_RAY_FLAG_DDIST_DPOS)) {
          if (dtdv_fd) {  // BVH2DR.cpp:1075-1090
            float vv[8];
#pragma unroll
            for (int i = 0; i < 8; i++) vv[i] = values[i];
            const float delta = 0.001f;
            for (int i = 0; i < 8; i++) {
              vv[i] += delta;
              float d1 = dr_intersect(flags, ray_dir, vv, d, d_inv, qFar, start_q, nullptr);
              vv[i] -= 2 * delta;
              float d2 = dr_intersect(flags, ray_dir, vv, d, d_inv, qFar, start_q, nullptr);
              vv[i] += delta;
              dtdv[i] = (d1 - d2) / (2 * delta);
            }
          } else {  // dt/dv_i = -w_i(q*) / (grad f(q*) . dir)
            v3 q = start_q + t * ray_dir;
            float fd = dot(trilinear_diff(values, q), ray_dir);
#pragma unroll
            for (int i = 0; i < 8; i++) dtdv[i] = fd != 0.0f ? -corner_w(i, q) / fd : 0.0f;
          }
#pragma unroll
          for (int i = 0; i < 8; i++) {
            relax->pd_idx[i] = S.values[N.z + vid(i)];
            relax->pd_ddist[i] = d * dtdv[i];
          }
        }
        if (!kColour) break;  // colour and its derivatives are not read
        v3 pos = ray_pos + pHit->t * ray_dir;
        v3 dp = (pos - bmin) * (0.5f * sz);
        const uint32_t t_off = N.z + vs * vs * vs;
        v3 col(0, 0, 0), J0(0, 0, 0), J1(0, 0, 0), J2(0, 0, 0);
#pragma unroll
        for (int i = 0; i < 8; i++) {
          uint32_t ci = S.values[t_off + i];
          v3 c(S.values_f[ci], S.values_f[ci + 1], S.values_f[ci + 2]);
          float wx = (i & 4) ? dp.x : 1 - dp.x, wy = (i & 2) ? dp.y : 1 - dp.y, wz = (i & 1) ? dp.z : 1 - dp.z;
          col = col + (wx * wy * wz) * c;
          if (flags & LRT_DR_RAY_FLAG_DDIFFUSE_DPOS) {
            float ax = (i & 4) ? 1.0f : -1.0f, ay = (i & 2) ? 1.0f : -1.0f, az = (i & 1) ? 1.0f : -1.0f;
            J0 = J0 + (ax * wy * wz) * c;
            J1 = J1 + (wx * ay * wz) * c;
            J2 = J2 + (wx * wy * az) * c;
          }
          relax->col_idx[i] = ci;
          relax->col_w[i] = wx * wy * wz;
        }
        pHit->color = col;
        if (flags & LRT_DR_RAY_FLAG_DDIFFUSE_DPOS) {
#pragma unroll
          for (int i = 0; i < 8; i++) {
            v3 v(0.5f * sz * ray_dir.x * dtdv[i], 0.5f * sz * ray_dir.y * dtdv[i], 0.5f * sz * ray_dir.z * dtdv[i]);
            relax->pd_ddiffuse[i] = J0 * v.x + J1 * v.y + J2 * v.z;
          }
        }
        break;
      }
    }
    bnf.x += mx(0.0f, fnf.y - bnf.x) + 1e-6f;
  }
}

// BVH2DR.cpp:12-167 (stop at the first hit, forced in the reference)
// Node-pair descent of dr_nearest_hit (BVH2Common.cpp:3431-3466 with the
// border pass's culled slots), for rays of one direction octant (OCT >= 0,
// RayBoxOct) or any (OCT = -1).
template <int SPEC, int OCT>
__device__ __forceinline__ void dr_descend(const DevScene& S, v3 pos, v3 inv, float tNear, float hit_t, int& top, uint32_t& left, Stack& st,
                                           bool ref) {
  while (top >= 0 && ((left & LRT_LEAF_BIT) == 0)) {
    uint4 a, b, c, e;
    load_pair(S, left, a, b, c, e);
    const float2 tm0 = RayBoxOct<OCT>(pos, inv, v3(__uint_as_float(a.x), __uint_as_float(a.y), __uint_as_float(a.z)),
                                      v3(__uint_as_float(b.x), __uint_as_float(b.y), __uint_as_float(b.z)), ref);
    const float2 tm1 = RayBoxOct<OCT>(pos, inv, v3(__uint_as_float(c.x), __uint_as_float(c.y), __uint_as_float(c.z)),
                                      v3(__uint_as_float(e.x), __uint_as_float(e.y), __uint_as_float(e.z)), ref);
    const bool h0 = (tm0.x <= tm0.y) && (tm0.y >= tNear) && (tm0.x <= hit_t);
    const bool h1 = (tm1.x <= tm1.y) && (tm1.y >= tNear) && (tm1.x <= hit_t);
    uint32_t ch0 = a.w, ch1 = c.w;
    if (SPEC == kDrDepthBorder) {
      // A culled slot stands for bricks that are all skipped, each
      // clearing the ray's state if the ray crosses it past the bricks'
      // tNear (0.1, 1e-5 margin in dr_brick). As the bricks tile the
      // box, a ray that crosses the box (positive length, entry past
      // 0.1 + 1e-3) crosses some brick past that tNear: the slot becomes
      // a leaf that only clears the state, visited in the subtree's
      // place. Any other ray descends as over the full tree.
      if (b.w == kDrCulledSlot && tm0.x < tm0.y && tm0.x > 0.1f + 1e-3f) ch0 = kDrResetLeaf;
      if (e.w == kDrCulledSlot && tm1.x < tm1.y && tm1.x > 0.1f + 1e-3f) ch1 = kDrResetLeaf;
    }
    left = h0 ? ch0 : ch1;
    if (h0 && h1) {
      const bool near0 = tm0.x <= tm1.x;
      left = near0 ? ch0 : ch1;
      st.put(top, near0 ? ch1 : ch0);
      top++;
    }
    if (!h0 && !h1) {
      top--;
      left = st.get(top > 0 ? top : 0);
    }
  }
}

template <int SPEC>
__device__ __forceinline__ void dr_descend_any(int oct, const DevScene& S, v3 pos, v3 inv, float tNear, float hit_t, int& top, uint32_t& left,
                                               Stack& st, bool ref) {
  switch (oct) {  // wave-uniform
    case 0: dr_descend<SPEC, 0>(S, pos, inv, tNear, hit_t, top, left, st, false); break;
    case 1: dr_descend<SPEC, 1>(S, pos, inv, tNear, hit_t, top, left, st, false); break;
    case 2: dr_descend<SPEC, 2>(S, pos, inv, tNear, hit_t, top, left, st, false); break;
    case 3: dr_descend<SPEC, 3>(S, pos, inv, tNear, hit_t, top, left, st, false); break;
    case 4: dr_descend<SPEC, 4>(S, pos, inv, tNear, hit_t, top, left, st, false); break;
    case 5: dr_descend<SPEC, 5>(S, pos, inv, tNear, hit_t, top, left, st, false); break;
    case 6: dr_descend<SPEC, 6>(S, pos, inv, tNear, hit_t, top, left, st, false); break;
    case 7: dr_descend<SPEC, 7>(S, pos, inv, tNear, hit_t, top, left, st, false); break;
    default: dr_descend<SPEC, -1>(S, pos, inv, tNear, hit_t, top, left, st, ref); break;
  }
}

template <int SPEC>
__device__ __forceinline__ DRHit dr_nearest_hit(const DevScene& S, const float* __restrict__ bvals, const float2* __restrict__ bminmax,
                                                const DRParams& Q, float relax_eps, v3 pos, v3 dir, float tFar, DRPayload* relax,
                                                Stack& st) {
  if (tFar <= 0.0f) tFar *= -1.0f;
  const v3 inv = SafeInverse(dir);
  DRHit hit;
  hit.t = tFar;
  hit.primId = 0xFFFFFFFFu;
  hit.color = v3(0, 0, 0);
  hit.sdf = 1000.0f;
  hit.normal = v3(1, 0, 0);
  hit.check_vox_border = 0;
  relax->missed_hit = hit;
  relax->missed_hit.check_vox_border = 1;
  relax->missed_cand = relax->missed_hit;
#pragma unroll
  for (int i = 0; i < 8; i++) {
    relax->missed_idx[i] = 0xFFFFFFFFu;
    relax->missed_dsdf[i] = 0.0f;
    relax->pd_idx[i] = 0xFFFFFFFFu;
    relax->pd_ddist[i] = 0.0f;
    relax->pd_ddiffuse[i] = v3(0, 0, 0);
    relax->col_idx[i] = 0xFFFFFFFFu;
    relax->col_w[i] = 0.0f;
  }
  const bool ref = wave_ref_slabs(inv);  // see RayBoxRef
  const float2 bh = RayBoxW(ref, pos, inv, v3(-1, -1, -1), v3(1, 1, 1));
  if ((bh.x <= bh.y) && (bh.y > 0.0f) && (bh.x < hit.t)) {
    const float tNear = 0.0f;
    int top = 0;
    uint32_t left = 0;
    const int oct = wave_octant(inv);
    while (top >= 0 && hit.primId == 0xFFFFFFFFu) {
      dr_descend_any<SPEC>(oct, S, pos, inv, tNear, hit.t, top, left, st, ref);
      if (SPEC == kDrDepthBorder && top >= 0 && left == kDrResetLeaf) {
        relax->missed_cand.sdf = 1000.f;
        relax->missed_hit.check_vox_border = 0;
      } else if (top >= 0 && left != 0xFFFFFFFFu) {
        dr_brick<SPEC>(S, bvals, bminmax, Q, relax_eps, pos, dir, inv, mx(tNear, 0.1f), left & LRT_START_MASK, relax, &hit);
      }
      top--;
      left = st.get(top > 0 ? top : 0);
    }
  }
  if (relax->missed_cand.sdf < relax->missed_hit.sdf) relax->missed_hit = relax->missed_cand;
  return hit;
}

}  // namespace lrtd

// trace_device.hpp - device-side SBS ray casting for gfx950.
//
// The arithmetic of each function follows the reference function named in
// its comment in the same operation order, and the library is compiled with
// -ffp-contract=off, so FP32 results match a CPU FP32 evaluation of the same
// steps bit for bit (division and square root are correctly rounded in HIP by
// default). Layout-specific loads are the MI355X part: brick values are read
// straight from the packed 8/16/32-bit words or the indexed float array.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "../../../include/lrt_types.h"

namespace lrtd {

struct v3 {
  float x, y, z;
  __device__ __forceinline__ v3() {}
  __device__ __forceinline__ v3(float a, float b, float c) : x(a), y(b), z(c) {}
};
__device__ __forceinline__ v3 operator+(v3 a, v3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ v3 operator-(v3 a, v3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ v3 operator*(v3 a, v3 b) { return v3(a.x * b.x, a.y * b.y, a.z * b.z); }
__device__ __forceinline__ v3 operator*(float s, v3 a) { return v3(s * a.x, s * a.y, s * a.z); }
__device__ __forceinline__ v3 operator*(v3 a, float s) { return v3(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ float mn(float a, float b) { return (b < a) ? b : a; }
__device__ __forceinline__ float mx(float a, float b) { return (a < b) ? b : a; }
__device__ __forceinline__ float sgn(float a) { return a > 0.f ? 1.f : (a < 0.f ? -1.f : 0.f); }
__device__ __forceinline__ float clampf_(float u, float a, float b) { return mn(mx(u, a), b); }
__device__ __forceinline__ float dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ v3 normalize(v3 a) {
  float l = sqrtf(dot(a, a));
  return v3(a.x / l, a.y / l, a.z / l);
}

struct Mat4 { float m[16]; };  // column-major

// Scene description passed by value to every kernel.
struct DevScene {
  const uint4* pairs;       // BVHNodePair as 4 x uint4 (boxMin|leftOffset, boxMax|escape) x 2
  const uint4* nodes;       // SdfSBSNode
  const uint32_t* values;   // packed DX words or indices (ID32F layouts)
  const float* values_f;    // shared floats of indexed layouts
  const float4* brecs;      // per brick: {bmin.xyz, data_offset}, {sz_inv, d, d_inv, qscale}, {d_max, mult, 0.5*sz, -}
  uint32_t brick_size, brick_pad, bytes_per_value, layout;
  uint32_t v_size, quant_voxel_diag, type, n_nodes;
  uint32_t pairs_bytes, values_bytes;  // sizes of pairs / values (< 4 GiB, checked when the scene is set)
  float v_size_f;                      // (float)v_size, a scalar operand for the voxel walk's index arithmetic
};

// A raw buffer resource over [p, p + bytes): loads through it take a 32-bit
// byte offset plus a scalar offset, so a brick's or a node pair's fields are
// addressed without per-lane 64-bit address arithmetic, and an offset past
// the end reads 0 instead of faulting. 0x00020000: the buffer's dword 3
// (32-bit data format) for the gfx9 family.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void* p, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)bytes, 0x00020000);
}

struct Hit {
  float t;
  uint32_t primId, instId, geomId;
  float coords[4];
  // Deferred normal (DX layouts): the hit brick, the hit voxel packed as
  // x | y << 8 | z << 16 (bit 31: negative start sign) and the hit point in
  // voxel units, from which FinishNormal evaluates the normal once per ray.
  uint32_t n_brick, n_voxel;
  float n_p0[3];
};

// ---- raytrace_common.h:59-82 ----------------------------------------------
__device__ __forceinline__ v3 SafeInverse(v3 d) {
  const float ooeps = 1.0e-36f;
  v3 r;
  r.x = 1.0f / (fabsf(d.x) > ooeps ? d.x : fabsf(ooeps) * sgn(d.x));
  r.y = 1.0f / (fabsf(d.y) > ooeps ? d.y : fabsf(ooeps) * sgn(d.y));
  r.z = 1.0f / (fabsf(d.z) > ooeps ? d.z : fabsf(ooeps) * sgn(d.z));
  return r;
}

__device__ __forceinline__ float2 RayBox(v3 o, v3 inv, v3 bmin, v3 bmax) {
  const float lo = inv.x * (bmin.x - o.x);
  const float hi = inv.x * (bmax.x - o.x);
  const float lo1 = inv.y * (bmin.y - o.y);
  const float hi1 = inv.y * (bmax.y - o.y);
  const float lo2 = inv.z * (bmin.z - o.z);
  const float hi2 = inv.z * (bmax.z - o.z);
  // fminf/fmaxf equal the reference's min/max for finite operands (a
  // signed-zero difference cannot change any later comparison) and map to
  // single v_min/v_max instructions. The operands are finite unless a
  // direction component is zero (SafeInverse then returns an infinite
  // component); waves holding such a ray take RayBoxRef instead (RayBoxW).
  const float tmin = fmaxf(fminf(lo, hi), fminf(lo1, hi1));
  const float tmax = fminf(fmaxf(lo, hi), fmaxf(lo1, hi1));
  return make_float2(fmaxf(tmin, fminf(lo2, hi2)), fminf(tmax, fmaxf(lo2, hi2)));
}

// The same slab test with the reference's std::min / std::max ((b < a) ? b : a
// and (a < b) ? b : a), which differ from fminf / fmaxf only for NaN operands:
// a zero direction component (infinite inverse) and an origin on a slab plane
// give inf * 0. The reference then keeps or drops the NaN by argument position
// (a ray lying in a brick face still walks the brick), and this form does too.
__device__ __forceinline__ float2 RayBoxRef(v3 o, v3 inv, v3 bmin, v3 bmax) {
  const float lo = inv.x * (bmin.x - o.x);
  const float hi = inv.x * (bmax.x - o.x);
  const float lo1 = inv.y * (bmin.y - o.y);
  const float hi1 = inv.y * (bmax.y - o.y);
  const float lo2 = inv.z * (bmin.z - o.z);
  const float hi2 = inv.z * (bmax.z - o.z);
  const float tmin = mx(mn(lo, hi), mn(lo1, hi1));
  const float tmax = mn(mx(lo, hi), mx(lo1, hi1));
  return make_float2(mx(tmin, mn(lo2, hi2)), mn(tmax, mx(lo2, hi2)));
}

// True when an active lane of the wave has an infinite inverse-direction
// component: that wave's slab tests use RayBoxRef (a wave-uniform choice; with
// finite operands both forms give the same values).
__device__ __forceinline__ bool wave_ref_slabs(v3 inv) {
  const bool fin = fabsf(inv.x) < INFINITY && fabsf(inv.y) < INFINITY && fabsf(inv.z) < INFINITY;
  return __builtin_amdgcn_ballot_w64(!fin) != 0;
}

__device__ __forceinline__ float2 RayBoxW(bool ref, v3 o, v3 inv, v3 bmin, v3 bmax) {
  float2 r = RayBox(o, inv, bmin, bmax);
  if (ref) r = RayBoxRef(o, inv, bmin, bmax);
  return r;
}

// ---- BVH2Common.cpp:62-72 --------------------------------------------------
__device__ __forceinline__ float trilinear(const float v[8], v3 dp) {
  return (1 - dp.x) * (1 - dp.y) * (1 - dp.z) * v[0] + (1 - dp.x) * (1 - dp.y) * (dp.z) * v[1] +
         (1 - dp.x) * (dp.y) * (1 - dp.z) * v[2] + (1 - dp.x) * (dp.y) * (dp.z) * v[3] +
         (dp.x) * (1 - dp.y) * (1 - dp.z) * v[4] + (dp.x) * (1 - dp.y) * (dp.z) * v[5] +
         (dp.x) * (dp.y) * (1 - dp.z) * v[6] + (dp.x) * (dp.y) * (dp.z) * v[7];
}

// ---- BVH2Common.cpp:657-687 ------------------------------------------------
__device__ __forceinline__ v3 trilinear_diff(const float v[8], v3 dp) {
  float ddx = -(1 - dp.y) * (1 - dp.z) * v[0] + -(1 - dp.y) * (dp.z) * v[1] + -(dp.y) * (1 - dp.z) * v[2] + -(dp.y) * (dp.z) * v[3] +
              (1 - dp.y) * (1 - dp.z) * v[4] + (1 - dp.y) * (dp.z) * v[5] + (dp.y) * (1 - dp.z) * v[6] + (dp.y) * (dp.z) * v[7];
  float ddy = -(1 - dp.x) * (1 - dp.z) * v[0] + -(1 - dp.x) * (dp.z) * v[1] + (1 - dp.x) * (1 - dp.z) * v[2] + (1 - dp.x) * (dp.z) * v[3] +
              -(dp.x) * (1 - dp.z) * v[4] + -(dp.x) * (dp.z) * v[5] + (dp.x) * (1 - dp.z) * v[6] + (dp.x) * (dp.z) * v[7];
  float ddz = -(1 - dp.x) * (1 - dp.y) * v[0] + (1 - dp.x) * (1 - dp.y) * v[1] + -(1 - dp.x) * (dp.y) * v[2] + (1 - dp.x) * (dp.y) * v[3] +
              -(dp.x) * (1 - dp.y) * v[4] + (dp.x) * (1 - dp.y) * v[5] + -(dp.x) * (dp.y) * v[6] + (dp.x) * (dp.y) * v[7];
  return v3(ddx, ddy, ddz);
}

// ---- BVH2Common.cpp:30-35 --------------------------------------------------
__device__ __forceinline__ float2 encode_normal(v3 v) {
  float s = 1.0f / (fabsf(v.x) + fabsf(v.y) + fabsf(v.z));
  float px = v.x * s, py = v.y * s;
  float snx = (px >= 0.0f) ? 1.0f : -1.0f, sny = (py >= 0.0f) ? 1.0f : -1.0f;
  if (v.z <= 0.0f) return make_float2((1.0f - fabsf(py)) * snx, (1.0f - fabsf(px)) * sny);
  return make_float2(px, py);
}

// ---- eye_ray.cpp:22-32 -----------------------------------------------------
__device__ __forceinline__ v3 decode_normal(float ex, float ey) {
  v3 v(ex, ey, 1.0f - fabsf(ex) - fabsf(ey));
  if (v.z < 0) {
    float vx = v.x;
    v.x = (1.0f - fabsf(v.y)) * ((v.x >= 0.0f) ? 1.0f : -1.0f);
    v.y = (1.0f - fabsf(vx)) * ((v.y >= 0.0f) ? 1.0f : -1.0f);
  }
  return normalize(v);
}

__device__ __forceinline__ bool is_indexed(const DevScene& S) {
  return S.layout == LRT_SBS_LAYOUT_ID32F_IRGB32F || S.layout == LRT_SBS_LAYOUT_ID32F_IRGB32F_IN;
}

// The 8 quantised corner values of a voxel of a DX-layout brick
// (BVH2Common.cpp:725-762): q = (values[v_off + vId / (4/bpv)] >>
// (8*bpv * (vId % (4/bpv)))) & max_val is the little-endian bpv-byte element
// vId of the brick's words, read here directly with a byte / short / dword
// load. base: byte offset of the voxel's corner 0 in the values array
// (corner_base). Returns the smallest of the 8.
__device__ __forceinline__ uint32_t load_corners_q_at(const DevScene& S, uint32_t base, uint32_t q[8]) {
  const uint32_t vs = S.v_size;
  const uint32_t bpv = S.bytes_per_value;
  // corner i at element b0 + (i&4 ? vs*vs : 0) + (i&2 ? vs : 0) + (i&1): the
  // x / y steps are scene constants (scalar offsets), the z step an immediate
  const __amdgpu_buffer_rsrc_t rs = buf_rsrc(S.values, S.values_bytes);
  const uint32_t sy = vs * bpv, sx = vs * vs * bpv, sxy = sx + sy;
  if (bpv == 1) {
    q[0] = __builtin_amdgcn_raw_buffer_load_b8(rs, base, 0, 0);
    q[1] = __builtin_amdgcn_raw_buffer_load_b8(rs, base + 1, 0, 0);
    q[2] = __builtin_amdgcn_raw_buffer_load_b8(rs, base, sy, 0);
    q[3] = __builtin_amdgcn_raw_buffer_load_b8(rs, base + 1, sy, 0);
    q[4] = __builtin_amdgcn_raw_buffer_load_b8(rs, base, sx, 0);
    q[5] = __builtin_amdgcn_raw_buffer_load_b8(rs, base + 1, sx, 0);
    q[6] = __builtin_amdgcn_raw_buffer_load_b8(rs, base, sxy, 0);
    q[7] = __builtin_amdgcn_raw_buffer_load_b8(rs, base + 1, sxy, 0);
  } else if (bpv == 2) {
    q[0] = __builtin_amdgcn_raw_buffer_load_b16(rs, base, 0, 0);
    q[1] = __builtin_amdgcn_raw_buffer_load_b16(rs, base + 2, 0, 0);
    q[2] = __builtin_amdgcn_raw_buffer_load_b16(rs, base, sy, 0);
    q[3] = __builtin_amdgcn_raw_buffer_load_b16(rs, base + 2, sy, 0);
    q[4] = __builtin_amdgcn_raw_buffer_load_b16(rs, base, sx, 0);
    q[5] = __builtin_amdgcn_raw_buffer_load_b16(rs, base + 2, sx, 0);
    q[6] = __builtin_amdgcn_raw_buffer_load_b16(rs, base, sxy, 0);
    q[7] = __builtin_amdgcn_raw_buffer_load_b16(rs, base + 2, sxy, 0);
  } else {
    q[0] = __builtin_amdgcn_raw_buffer_load_b32(rs, base, 0, 0);
    q[1] = __builtin_amdgcn_raw_buffer_load_b32(rs, base + 4, 0, 0);
    q[2] = __builtin_amdgcn_raw_buffer_load_b32(rs, base, sy, 0);
    q[3] = __builtin_amdgcn_raw_buffer_load_b32(rs, base + 4, sy, 0);
    q[4] = __builtin_amdgcn_raw_buffer_load_b32(rs, base, sx, 0);
    q[5] = __builtin_amdgcn_raw_buffer_load_b32(rs, base + 4, sx, 0);
    q[6] = __builtin_amdgcn_raw_buffer_load_b32(rs, base, sxy, 0);
    q[7] = __builtin_amdgcn_raw_buffer_load_b32(rs, base + 4, sxy, 0);
  }
  return min(min(min(q[0], q[1]), min(q[2], q[3])), min(min(q[4], q[5]), min(q[6], q[7])));
}

// Byte offset of corner 0 of voxel (vx, vy, vz) of the brick at word v_off:
// element ((vx + pad) * vs + vy + pad) * vs + vz + pad of bpv bytes, written
// as (vx * vs + vy) * vs + vz plus the brick's constant pad_base.
__device__ __forceinline__ uint32_t pad_base(const DevScene& S, uint32_t v_off) {
  const uint32_t vs = S.v_size;
  return v_off * 4u + ((S.brick_pad * (vs * vs + vs + 1u)) << (S.bytes_per_value >> 1));  // x bpv for bpv 1, 2, 4
}
__device__ __forceinline__ uint32_t corner_base(const DevScene& S, uint32_t pbase, uint32_t vx, uint32_t vy, uint32_t vz) {
  const uint32_t vs = S.v_size;
  return pbase + (((vx * vs + vy) * vs + vz) << (S.bytes_per_value >> 1));
}

// The same offset from the voxel walk's float coordinates (floored, in
// [0, brick_size)): scenes are accepted with brick_size <= 16 and pad <= 1,
// so every term is a small integer, the float operations are exact, and one
// conversion replaces three conversions and two 32-bit integer multiplies.
__device__ __forceinline__ uint32_t corner_base_f(const DevScene& S, uint32_t pbase, v3 vp) {
  const float vs = S.v_size_f;
  return pbase + ((uint32_t)fmaf(fmaf(vp.x, vs, vp.y), vs, vp.z) << (S.bytes_per_value >> 1));
}

__device__ __forceinline__ uint32_t load_corners_q(const DevScene& S, uint32_t v_off, uint32_t vx, uint32_t vy, uint32_t vz, uint32_t q[8]) {
  return load_corners_q_at(S, corner_base(S, pad_base(S, v_off), vx, vy, vz), q);
}

// Dequantised corner values -d_max + mult * q (BVH2Common.cpp:749-755) and
// their minimum. The map is monotone (mult > 0, rounding preserves order), so
// the minimum is also -d_max + mult * min(q): callers that only test the
// minimum can defer the other conversions.
__device__ __forceinline__ float dequant_corners(const uint32_t q[8], float d_max, float mult, float values[8]) {
  float vmin = 1e6f;
#pragma unroll
  for (int i = 0; i < 8; i++) {
    values[i] = -d_max + mult * (float)q[i];
    vmin = fminf(vmin, values[i]);  // finite values: equal to the reference's min, maps to v_min3
  }
  return vmin;
}

// ---- BVH2Common.cpp:725-762, trilinear branch -----------------------------
// d_max / mult: the brick's dequantisation constants (ignored for indexed
// layouts), the same values load_voxel below derives from sz_inv.
template <bool IDX = true>
__device__ __forceinline__ float load_voxel_q(const DevScene& S, uint32_t v_off, uint32_t vx, uint32_t vy, uint32_t vz, float d_max,
                                              float mult, float values[8]) {
  float vmin = 1e6f;
  const uint32_t vs = S.v_size, pad = S.brick_pad;
  if (IDX && is_indexed(S)) {
#pragma unroll
    for (int i = 0; i < 8; i++) {
      uint32_t vId = (vx + ((i & 4) >> 2) + pad) * vs * vs + (vy + ((i & 2) >> 1) + pad) * vs + (vz + (i & 1) + pad);
      values[i] = S.values_f[S.values[v_off + vId]];
      vmin = mn(vmin, values[i]);
    }
  } else {
    uint32_t q[8];
    load_corners_q(S, v_off, vx, vy, vz, q);
    vmin = dequant_corners(q, d_max, mult, values);
  }
  return vmin;
}

// BVH2Common.cpp:749-750 dequantisation constants of a brick of size 2/sz_inv.
__device__ __forceinline__ void dequant_consts(const DevScene& S, float sz_inv, float& d_max, float& mult) {
  const uint32_t bits = 8 * S.bytes_per_value;
  const uint32_t max_val = S.bytes_per_value == 4 ? 0xFFFFFFFFu : ((1u << bits) - 1);
  d_max = S.quant_voxel_diag ? 1.73205081f * (sz_inv / (float)S.brick_size) : 1.73205081f * sz_inv;
  mult = 2.0f * d_max / (float)max_val;
}

template <bool IDX = true>
__device__ __forceinline__ float load_voxel(const DevScene& S, uint32_t v_off, uint32_t vx, uint32_t vy, uint32_t vz, float sz_inv,
                                            float values[8]) {
  float d_max = 0.0f, mult = 0.0f;
  if (!(IDX && is_indexed(S))) dequant_consts(S, sz_inv, d_max, mult);
  return load_voxel_q<IDX>(S, v_off, vx, vy, vz, d_max, mult, values);
}

// The renderer's lights (MultiRenderer::m_lights, eye_ray.h:230), passed by
// value: sp = direction or position with the type's bits in w, col = colour.
struct LightSet {
  float4 sp[LRT_MAX_LIGHTS];
  float4 col[LRT_MAX_LIGHTS];
  uint32_t n = 0;
};

struct TraceParams {
  uint32_t sdf_node_intersect;
  uint32_t representation_mode;
  uint32_t render_mode;
  uint32_t normal_mode;
  uint32_t need_normal;
};

// Cubic along the ray inside one voxel (JCGT 2022 coefficients, the same
// expressions as the Newton/analytic branches below, BVH2Common.cpp:322-361).
// Note the sign: -trilinear(start_q + t*dir) * d_inv = c0 + c1 t + c2 t^2 + c3 t^3.
__device__ __forceinline__ void cubic_coeffs(const float values[8], float d_inv, v3 o, v3 d3, float& c0, float& c1, float& c2,
                                             float& c3) {
  float s000 = values[0] * d_inv, s001 = values[1] * d_inv, s010 = values[2] * d_inv, s011 = values[3] * d_inv;
  float s100 = values[4] * d_inv, s101 = values[5] * d_inv, s110 = values[6] * d_inv, s111 = values[7] * d_inv;
  float a = s101 - s001;
  float k0 = s000;
  float k1 = s100 - s000;
  float k2 = s010 - s000;
  float k3 = s110 - s010 - k1;
  float k4 = k0 - s001;
  float k5 = k1 - a;
  float k6 = k2 - (s011 - s001);
  float k7 = k3 - (s111 - s011 - a);
  float m0 = o.x * o.y;
  float m1 = d3.x * d3.y;
  float m2 = o.x * d3.y + o.y * d3.x;
  float m3 = k5 * o.z - k1;
  float m4 = k6 * o.z - k2;
  float m5 = k7 * o.z - k3;
  c0 = (k4 * o.z - k0) + o.x * m3 + o.y * m4 + m0 * m5;
  c1 = d3.x * m3 + d3.y * m4 + m2 * m5 + d3.z * (k4 + k5 * o.x + k6 * o.y + k7 * m0);
  c2 = m1 * m5 + d3.z * (k5 * d3.x + k6 * d3.y + k7 * m2);
  c3 = k7 * m1 * d3.z;
}

// g(t) = c0 + c1 t + c2 t^2 + c3 t^3 > margin on all of [0, ts]? Checks both
// ends and every critical point inside (the extrema of g on the interval).
__device__ __forceinline__ bool cubic_above_on(float c0, float c1, float c2, float c3, float ts, float margin) {
  auto g = [&](float t) { return c0 + t * (c1 + t * (c2 + t * c3)); };
  if (!(g(0.0f) > margin) || !(g(ts) > margin)) return false;
  const float qa = 3 * c3, qb = 2 * c2, qc = c1;
  if (fabsf(qa) < 1e-20f) {
    if (fabsf(qb) < 1e-20f) return true;
    const float tc = -qc / qb;
    return !(tc > 0.0f && tc < ts) || g(tc) > margin;
  }
  const float D = qb * qb - 4 * qa * qc;
  if (D < 0.0f) return true;
  const float sq = sqrtf(D);
  const float q = -0.5f * (qb + (qb >= 0.0f ? sq : -sq));
  const float r1 = q / qa;
  const float r2 = q != 0.0f ? qc / q : r1;
  if (r1 > 0.0f && r1 < ts && !(g(r1) > margin)) return false;
  if (r2 > 0.0f && r2 < ts && !(g(r2) > margin)) return false;
  return true;
}

// BVH2Common.cpp:574-602: central difference (h = 1e-3) of the trilinear at
// p0 (voxel units), normalised, times the start sign.
__device__ __forceinline__ v3 fd_normal(const float values[8], v3 p0, float start_sign) {
  const float h = 0.001f;
  float ddx = (trilinear(values, p0 + v3(h, 0, 0)) - trilinear(values, p0 + v3(-h, 0, 0))) / (2 * h);
  float ddy = (trilinear(values, p0 + v3(0, h, 0)) - trilinear(values, p0 + v3(0, -h, 0))) / (2 * h);
  float ddz = (trilinear(values, p0 + v3(0, 0, h)) - trilinear(values, p0 + v3(0, 0, -h))) / (2 * h);
  return start_sign * normalize(v3(ddx, ddy, ddz));
}

// ---- BVH2Common.cpp:274-631 ------------------------------------------------
// Returns true when *pHit was updated. d_inv = 1/d is passed in (same value
// the reference computes per call).
// SEEDED: seed_q (voxel units from the voxel entry) is a predicted start for
// the iterative root finders; seed_q >= qFar predicts no surface in this
// voxel, which lets sphere tracing skip a voxel it provably cannot stop in. Newton starts there instead of the bracket
// midpoint when it lies strictly inside the exact bracket; sphere tracing
// starts there (or, failing that, half a voxel before it) only after the
// cubic along the ray is shown to stay above the stopping threshold up to the
// start, so the root it converges to is the first one either way. Any other mode, or a seed that fails the checks,
// runs unseeded. n_iter (optional) accumulates root-finder iterations.
// DEFER: on a hit, record the hit point (voxel units) and the start sign in
// pHit instead of evaluating the normal here (FinishNormal does it later
// with the same operations); the caller records the voxel.
template <bool SEEDED = false, int RF = -1, bool DEFER = false>
__device__ __forceinline__ bool LocalSurfaceIntersection(const TraceParams& P, v3 ray_dir, const float values[8], uint32_t primId,
                                                         uint32_t geomType, float d, float d_inv, float qFar, float fNearX,
                                                         v3 start_q, Hit* pHit, float seed_q = -1.0f, uint32_t* n_iter = nullptr) {
  const float EPS = 1e-6f;
  const unsigned mode = RF >= 0 ? (unsigned)RF : P.sdf_node_intersect;  // RF >= 0: root finder fixed at compile time
  float t = 0.0f;
  bool hit = false;
  unsigned iter = 0;
  float start_sign = 1;
  // trilinear(values, start_q + 0 * ray_dir) in the reference: the zero step
  // can change only the sign of a zero coordinate, which changes no output
  // (it can flip only the sign of an all-zero sum, and sgn(+-0) = 0 either way).
  float start_dist = trilinear(values, start_q);
  if (P.representation_mode == LRT_REPRESENTATION_MODE_SURFACE) start_sign = sgn(start_dist);
  start_dist *= start_sign;

  if (start_dist <= EPS || mode == LRT_SDF_NODE_INTERSECT_BBOX) {
    hit = true;
  } else if (mode == LRT_SDF_NODE_INTERSECT_ST) {
    float dist = start_dist;
    bool seeded_start = false;
    bool skip = false;
    float c0 = 0, c1 = 0, c2 = 0, c3 = 0;
    if (SEEDED && seed_q >= qFar && start_sign == 1.0f) {
      // predicted: no surface in this voxel. If the distance provably stays
      // above twice the stopping threshold along the whole segment, the
      // trace could not stop inside it, so the voxel is skipped untraced.
      cubic_coeffs(values, d_inv, start_q, ray_dir, c0, c1, c2, c3);
      skip = cubic_above_on(-c0, -c1, -c2, -c3, qFar, 2.0f * EPS * d_inv);
    } else if (SEEDED && seed_q > 0.0f && seed_q < qFar && start_sign == 1.0f) {
      // the predicted hit first, then half a voxel before it
      cubic_coeffs(values, d_inv, start_q, ray_dir, c0, c1, c2, c3);
      c0 = -c0; c1 = -c1; c2 = -c2; c3 = -c3;  // now the distance itself (voxel units)
      float ts = seed_q;
      bool ok = cubic_above_on(c0, c1, c2, c3, ts, EPS * d_inv);
      if (!ok && seed_q > 0.5f) {
        ts = seed_q - 0.5f;
        ok = cubic_above_on(c0, c1, c2, c3, ts, EPS * d_inv);
      }
      if (ok) {
        t = ts;
        dist = trilinear(values, start_q + t * ray_dir);
        seeded_start = true;
      }
    }
    while (!skip && t < qFar && dist > EPS && iter < 256u) {
      t += dist * d_inv;
      dist = start_sign * trilinear(values, start_q + t * ray_dir);
      iter++;
    }
    // Sphere tracing stops anywhere in the band where the distance is below
    // EPS (EPS / |slope| long along the ray), and where the distance changes
    // faster than 1 per unit length a step can overshoot the root: in both
    // cases a different start point lands elsewhere. A seeded result is kept
    // only where neither can move it: |slope| <= 1 on the whole segment the
    // unseeded trace would cover (no overshoot from any start) and
    // slope <= -0.25 at the stop (band far inside the depth tolerance);
    // otherwise the trace is redone from the voxel entry, as unseeded.
    if (SEEDED && seeded_start) {
      const float slope = c1 + t * (2 * c2 + t * 3 * c3);
      auto dg = [&](float u) { return fabsf(c1 + u * (2 * c2 + u * 3 * c3)); };
      float smax = fmaxf(dg(0.0f), dg(t));
      if (fabsf(c3) > 1e-20f) {
        const float tv = -c2 / (3 * c3);
        if (tv > 0.0f && tv < t) smax = fmaxf(smax, dg(tv));
      }
      if (!(dist <= EPS && dist >= -EPS && slope <= -0.25f && smax <= 1.0f)) {
        t = 0.0f;
        dist = start_dist;
        unsigned it2 = 0;
        while (t < qFar && dist > EPS && it2 < 256u) {
          t += dist * d_inv;
          dist = start_sign * trilinear(values, start_q + t * ray_dir);
          it2++;
        }
        iter += it2;
      }
    }
    hit = !skip && (dist <= EPS);
  } else {
    float s000 = values[0] * d_inv, s001 = values[1] * d_inv, s010 = values[2] * d_inv, s011 = values[3] * d_inv;
    float s100 = values[4] * d_inv, s101 = values[5] * d_inv, s110 = values[6] * d_inv, s111 = values[7] * d_inv;
    float a = s101 - s001;
    float k0 = s000;
    float k1 = s100 - s000;
    float k2 = s010 - s000;
    float k3 = s110 - s010 - k1;
    float k4 = k0 - s001;
    float k5 = k1 - a;
    float k6 = k2 - (s011 - s001);
    float k7 = k3 - (s111 - s011 - a);
    v3 o = start_q, d3 = ray_dir;
    float m0 = o.x * o.y;
    float m1 = d3.x * d3.y;
    float m2 = o.x * d3.y + o.y * d3.x;
    float m3 = k5 * o.z - k1;
    float m4 = k6 * o.z - k2;
    float m5 = k7 * o.z - k3;
    float c0 = (k4 * o.z - k0) + o.x * m3 + o.y * m4 + m0 * m5;
    float c1 = d3.x * m3 + d3.y * m4 + m2 * m5 + d3.z * (k4 + k5 * o.x + k6 * o.y + k7 * m0);
    float c2 = m1 * m5 + d3.z * (k5 * d3.x + k6 * d3.y + k7 * m2);
    float c3 = k7 * m1 * d3.z;

    if (mode == LRT_SDF_NODE_INTERSECT_NEWTON) {
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
      bool s0 = c0 > 0;
      bool s1 = (c0 + t1 * (c1 + t1 * (c2 + t1 * c3))) > 0;
      bool s2 = (c0 + t2 * (c1 + t2 * (c2 + t2 * c3))) > 0;
      bool s3 = (c0 + t3 * (c1 + t3 * (c2 + t3 * c3))) > 0;
      float nwt_min = t0, nwt_max = t0;
      if (s0 != s1) { nwt_min = t0; nwt_max = t1; }
      else if (s1 != s2) { nwt_min = t1; nwt_max = t2; }
      else if (s2 != s3) { nwt_min = t2; nwt_max = t3; }
      if (nwt_min < nwt_max) {
        unsigned it = 0;
        float rtn = (SEEDED && seed_q > nwt_min && seed_q < nwt_max) ? seed_q : 0.5f * (nwt_min + nwt_max);
        float f = 1000;
        while (it < 10u && fabsf(f) >= EPS) {
          f = c0 + rtn * (c1 + rtn * (c2 + rtn * c3));
          float df = c1 + rtn * (2 * c2 + rtn * 3 * c3);
          float dx = f / (df + sgn(df) * 1e-9f);
          rtn -= dx;
          it++;
        }
        if (n_iter) *n_iter += it;  // the reference keeps Newton's count local (iter stays 0)
        t = rtn;
        hit = (t >= 0 && t <= qFar && fabsf(f) < EPS);
      } else {
        hit = false;
      }
    } else if (mode == LRT_SDF_NODE_INTERSECT_ANALYTIC) {
      float x1 = 1000, x2 = 1000, x3 = 1000;
      // The reference compares float against double constants; 1e-2 and 1e-4
      // round down to float, so "> 1e-2f" is the same test in FP32.
      if (fabsf(c3) > 1e-2f) {
        float A_ = c2 / c3, B_ = c1 / c3, C_ = c0 / c3;
        float Q = (A_ * A_ - 3 * B_) / 9;
        float R = (2 * A_ * A_ - 9 * A_ * B_ + 27 * C_) / 54;
        float Q3 = Q * Q * Q;
        if (R * R < Q3) {
          float theta = acosf(R / sqrtf(Q3));
          x1 = -2 * sqrtf(Q) * cosf(theta / 3) - A_ / 3;
          // the reference evaluates these two in double (M_PI promotes them); FP32 here
          const float two_pi = 6.28318530717958647692f;
          x2 = -2 * sqrtf(Q) * cosf((theta + two_pi) / 3) - A_ / 3;
          x3 = -2 * sqrtf(Q) * cosf((theta - two_pi) / 3) - A_ / 3;
        } else {
          float Am = -sgn(R) * powf(fabsf(R) + sqrtf(R * R - Q3), 1.0f / 3.0f);
          float Bm = fabsf(Am) > EPS ? Q / Am : 0;
          x1 = Am + Bm - A_ / 3;
        }
      } else if (fabsf(c2) > 1e-4f) {
        float D = c1 * c1 - 4 * c2 * c0;
        if (D > 0) {
          float q = -0.5f * (c1 + sgn(c1) * sqrtf(D));
          x1 = q / c2;
          if (fabsf(q) > EPS) x2 = c0 / q;
        }
      } else if (fabsf(c1) > EPS) {
        x1 = -c0 / c1;
      }
      x1 = x1 < 0 ? 1000 : x1;
      x2 = x2 < 0 ? 1000 : x2;
      x3 = x3 < 0 ? 1000 : x3;
      t = mn(x1, mn(x2, x3));
      hit = (t >= 0 && t <= qFar);
    } else {  // IT
      const float k = 2;
      float e = 0.1f * qFar;
      float t_max = fabsf(c3) < EPS ? 1e6f : -c2 / (3 * c3);
      float df_max = 3 * c3 * t_max * t_max + 2 * c2 * t_max + c1;
      float dist = start_dist;
      while (t < qFar && dist > EPS && iter < 256u) {
        float df_1 = 3 * c3 * t * t + 2 * c2 * t + c1;
        float df_2 = 3 * c3 * (t + e) * (t + e) + 2 * c2 * (t + e) + c1;
        float L = (t_max > t && t_max < t + e) ? mx(df_max, mx(df_1, df_2)) : mx(df_1, df_2);
        L = mx(L, EPS);
        float s = mn((dist * d_inv) / L, e);
        t += s;
        e = k * s;
        dist = start_sign * trilinear(values, start_q + t * ray_dir);
        iter++;
      }
      hit = (dist <= EPS);
    }
  }

  if (n_iter) *n_iter += iter;
  float tReal = fNearX + d * t;
  if (t <= qFar && hit && tReal < pHit->t) {
    if (DEFER) {
      const v3 p0 = start_q + t * ray_dir;
      pHit->n_p0[0] = p0.x;
      pHit->n_p0[1] = p0.y;
      pHit->n_p0[2] = p0.z;
      pHit->n_voxel = start_sign < 0.0f ? 0x80000000u : 0u;
    } else {
      const float2 en = P.need_normal ? encode_normal(fd_normal(values, start_q + t * ray_dir, start_sign)) : encode_normal(v3(0, 0, 1));
      pHit->coords[2] = en.x;
      pHit->coords[3] = en.y;
    }
    pHit->t = tReal;
    pHit->primId = (P.render_mode == LRT_RENDER_MODE_ST_ITERATIONS) ? iter : primId;
    pHit->instId = 0;
    pHit->geomId = 0u | (geomType << LRT_SH_TYPE);
    pHit->coords[0] = 0;
    pHit->coords[1] = 0;
    return true;
  }
  return false;
}

__device__ __forceinline__ float step01(float e0, float e1, float x) { return clampf_((x - e0) / (e1 - e0), 0.0f, 1.0f); }

// Per-brick constants (3 x float4, brick_records on the host) through a
// buffer resource: 48 B per brick, fewer than 2^24 bricks.
__device__ __forceinline__ void load_brick_rec(const DevScene& S, uint32_t brick, float4& r0, float4& r1, float4& r2) {
  typedef __attribute__((ext_vector_type(4))) unsigned int u4v;
  const __amdgpu_buffer_rsrc_t rb = buf_rsrc(S.brecs, S.n_nodes * 48u);
  const uint32_t off = brick * 48u;
  const u4v x = __builtin_amdgcn_raw_buffer_load_b128(rb, off, 0, 0), y = __builtin_amdgcn_raw_buffer_load_b128(rb, off + 16, 0, 0),
            z = __builtin_amdgcn_raw_buffer_load_b128(rb, off + 32, 0, 0);
  r0 = make_float4(__uint_as_float(x.x), __uint_as_float(x.y), __uint_as_float(x.z), __uint_as_float(x.w));
  r1 = make_float4(__uint_as_float(y.x), __uint_as_float(y.y), __uint_as_float(y.z), __uint_as_float(y.w));
  r2 = make_float4(__uint_as_float(z.x), __uint_as_float(z.y), __uint_as_float(z.z), __uint_as_float(z.w));
}

// ---- BVH2Common.cpp:764-984 ------------------------------------------------
// rayDirInv is SafeInverse(ray_dir), evaluated once per ray by the caller
// (the reference recomputes the identical value for every box test).
// Counters (optional) receive voxel and root-find counts.
// SEEDED: seed_world is a predicted hit distance along the ray (world t);
// each candidate voxel hands it to the root finder in its own units.
// RF >= 0 fixes the root finder at compile time; IDX = false compiles out the
// indexed-layout paths (colour, smoothed normals) for DX-layout scenes.
// DEFER (DX layouts only): the hit's normal is left to FinishNormal.
template <bool SEEDED = false, int RF = -1, bool IDX = true, bool DEFER = false>
__device__ __forceinline__ void BrickIntersect(const DevScene& S, const TraceParams& P, v3 ray_pos, v3 ray_dir, v3 rayDirInv,
                                               float tNear, uint32_t brick, Hit* pHit, uint32_t* n_vox, float seed_world = -1.0f,
                                               uint32_t* n_iter = nullptr) {
  static_assert(!(DEFER && IDX), "the deferred normal covers DX layouts (geometry normals) only");
  float values[8];
  // Per-brick constants precomputed on the host with the same float
  // operations (brick_records): bmin = -1 + sz_inv * pos, sz_inv = 2/sz,
  // d = 2/(sz*bs), d_inv = 1/d, qscale = 0.5*sz*bs, d_max, mult, 0.5*sz.
  float4 r0, r1, r2;
  load_brick_rec(S, brick, r0, r1, r2);
  const uint32_t v_off = __float_as_uint(r0.w);
  const float bsf = (float)S.brick_size;
  const float sz_inv = r1.x, d = r1.y, d_inv = r1.z, qscale = r1.w;
  const v3 bmin(r0.x, r0.y, r0.z);
  const v3 bmax = bmin + sz_inv * v3(1, 1, 1);
  const float hi_clamp = bsf - 1e-6f;
  const bool ref = wave_ref_slabs(rayDirInv);
  float2 bnf = RayBoxW(ref, ray_pos, rayDirInv, bmin, bmax);
  const uint32_t pbase = IDX ? 0u : pad_base(S, v_off);
  const float old_t = pHit->t;
  // A ray crosses at most 3*brick_size+1 voxels of a brick; the cap only
  // guarantees termination where the 1e-6 step could vanish in rounding.
  const uint32_t max_steps = 16u * (S.brick_size + 1u) + 16u;
  uint32_t steps = 0;
  while (bnf.x < bnf.y && pHit->t == old_t && steps++ < max_steps) {
    v3 hit_pos = ray_pos + bnf.x * ray_dir;
    v3 local = (hit_pos - bmin) * qscale;
    // the reference's clamp to [1e-6, bs - 1e-6] of finite values (the
    // median of value and bounds, one v_med3 per axis)
    v3 vp(floorf(__builtin_amdgcn_fmed3f(local.x, 1e-6f, hi_clamp)), floorf(__builtin_amdgcn_fmed3f(local.y, 1e-6f, hi_clamp)),
          floorf(__builtin_amdgcn_fmed3f(local.z, 1e-6f, hi_clamp)));
    v3 min_pos = bmin + d * vp;
    v3 max_pos = min_pos + d * v3(1, 1, 1);
    if (n_vox) (*n_vox)++;
    // DX layouts: the voxel's minimum from the smallest quantised value, the
    // 8 values converted only at a candidate voxel (dequant_corners)
    uint32_t q[8];
    const float vmin = IDX ? load_voxel_q<IDX>(S, v_off, (uint32_t)vp.x, (uint32_t)vp.y, (uint32_t)vp.z, r2.x, r2.y, values)
                           : -r2.x + r2.y * (float)load_corners_q_at(S, corner_base_f(S, pbase, vp), q);
    float2 fnf = RayBoxW(ref, ray_pos, rayDirInv, min_pos, max_pos);
    if (tNear < fnf.x && vmin <= 0.0f) {
      if (!IDX) dequant_corners(q, r2.x, r2.y, values);
      v3 start_pos = ray_pos + fnf.x * ray_dir;
      v3 start_q = (start_pos - min_pos) * qscale;
      float qFar = (fnf.y - fnf.x) * qscale;
      const float seed_q = SEEDED ? (seed_world - fnf.x) * d_inv : -1.0f;
      bool updated = LocalSurfaceIntersection<SEEDED, RF, DEFER>(P, ray_dir, values, brick, S.type, d, d_inv, qFar, fnf.x, start_q, pHit,
                                                                 seed_q, n_iter);
      if (DEFER && updated) {
        pHit->n_brick = brick;
        pHit->n_voxel |= (uint32_t)vp.x | ((uint32_t)vp.y << 8) | ((uint32_t)vp.z << 16);
      }
      if (IDX && updated && S.layout == LRT_SBS_LAYOUT_ID32F_IRGB32F_IN && P.normal_mode == LRT_NORMAL_MODE_SDF_SMOOTHED && P.need_normal) {
        // smoothed normal, BVH2Common.cpp:838-906
        const float beta = 0.5f;
        v3 dp = start_q + ((pHit->t - fnf.x) / d) * ray_dir;
        v3 nmq;
        nmq.x = dp.x < 0.5f ? step01(-beta, beta, dp.x) : step01(-beta, beta, dp.x - 1.0f);
        nmq.y = dp.y < 0.5f ? step01(-beta, beta, dp.y) : step01(-beta, beta, dp.y - 1.0f);
        nmq.z = dp.z < 0.5f ? step01(-beta, beta, dp.z) : step01(-beta, beta, dp.z - 1.0f);
        const int bs = (int)S.brick_size;
        const int vpi[3] = {(int)vp.x, (int)vp.y, (int)vp.z};
        const int vp0[3] = {dp.x < 0.5f ? vpi[0] - 1 : vpi[0], dp.y < 0.5f ? vpi[1] - 1 : vpi[1], dp.z < 0.5f ? vpi[2] - 1 : vpi[2]};
        v3 sm(0, 0, 0);
        for (int i = 0; i < 8; i++) {
          int VI[3] = {vp0[0] + ((i & 4) >> 2), vp0[1] + ((i & 2) >> 1), vp0[2] + (i & 1)};
          int BP[3];
          for (int a = 0; a < 3; a++) BP[a] = VI[a] >= 0 ? (VI[a] < bs ? 0 : 1) : -1;
          uint32_t nid = 9 * (BP[0] + 1) + 3 * (BP[1] + 1) + (BP[2] + 1);
          v3 dvp = v3((float)VI[0], (float)VI[1], (float)VI[2]) - vp;
          uint32_t nb = brick;
          uint32_t nv[3];
          if (nid == 13) {
            for (int a = 0; a < 3; a++) nv[a] = (uint32_t)VI[a];
          } else {
            nb = S.values[v_off + S.v_size * S.v_size * S.v_size + 8 + nid];
            for (int a = 0; a < 3; a++) nv[a] = (uint32_t)(VI[a] - bs * BP[a]);
          }
          v3 ni(0, 0, 0);
          if (nb != LRT_INVALID_IDX) {
            float vn[8];
            load_voxel(S, S.nodes[nb].z, nv[0], nv[1], nv[2], sz_inv, vn);
            ni = normalize(trilinear_diff(vn, dp - dvp));
          }
          float w = ((i & 4) ? (nmq.x) : (1 - nmq.x)) * ((i & 2) ? (nmq.y) : (1 - nmq.y)) * ((i & 1) ? (nmq.z) : (1 - nmq.z));
          sm = (i == 0) ? w * ni : sm + w * ni;
        }
        sm = normalize(sm);
        float2 en = encode_normal(sm);
        pHit->coords[2] = en.x;
        pHit->coords[3] = en.y;
      }
    }
    bnf.x += mx(0.0f, fnf.y - bnf.x) + 1e-6f;
  }
  if (IDX && pHit->t < old_t && is_indexed(S)) {  // colour, BVH2Common.cpp:962-982
    v3 pos = ray_pos + pHit->t * ray_dir;
    v3 dp = (pos - bmin) * r2.z;  // 0.5 * sz
    uint32_t t_off = v_off + S.v_size * S.v_size * S.v_size;
    v3 col(0, 0, 0);
    for (int i = 0; i < 8; i++) {
      float w = ((i & 4) ? dp.x : 1 - dp.x) * ((i & 2) ? dp.y : 1 - dp.y) * ((i & 1) ? dp.z : 1 - dp.z);
      uint32_t ci = S.values[t_off + i];
      col = col + w * v3(S.values_f[ci], S.values_f[ci + 1], S.values_f[ci + 2]);
    }
    col = v3(clampf_(floorf(255.0f * col.x + 0.5f), 0.0f, 255.0f), clampf_(floorf(255.0f * col.y + 0.5f), 0.0f, 255.0f),
             clampf_(floorf(255.0f * col.z + 0.5f), 0.0f, 255.0f));
    pHit->coords[0] = col.x + col.y / 256.0f;
    pHit->coords[1] = col.z;
  }
}

// The normal of a hit recorded by BrickIntersect<..., DEFER = true>: the hit
// voxel's 8 values are loaded again and the central difference is evaluated
// at the recorded point, the same operations the voxel loop would have done.
// Brick sizes up to 255 (the voxel coordinates are packed in bytes).
__device__ __forceinline__ void FinishNormal(const DevScene& S, const TraceParams& P, Hit* pHit) {
  if (pHit->primId == 0xFFFFFFFFu) return;
  v3 norm(0, 0, 1);
  if (P.need_normal) {
    const uint32_t b = pHit->n_brick, vox = pHit->n_voxel;
    float4 r0, r1, r2;
    load_brick_rec(S, b, r0, r1, r2);
    float values[8];
    load_voxel_q<false>(S, __float_as_uint(r0.w), vox & 0xFFu, (vox >> 8) & 0xFFu, (vox >> 16) & 0xFFu, r2.x, r2.y, values);
    norm = fd_normal(values, v3(pHit->n_p0[0], pHit->n_p0[1], pHit->n_p0[2]), (vox & 0x80000000u) ? -1.0f : 1.0f);
  }
  const float2 en = encode_normal(norm);
  pHit->coords[2] = en.x;
  pHit->coords[3] = en.y;
}

// ---- camera, eye_ray.cpp:146-160, render_common.h:25-55 -------------------
__device__ __forceinline__ void mul4(const float* m, float x, float y, float z, float w, float* r) {
  r[0] = m[0] * x + m[4] * y + m[8] * z + m[12] * w;
  r[1] = m[1] * x + m[5] * y + m[9] * z + m[13] * w;
  r[2] = m[2] * x + m[6] * y + m[10] * z + m[14] * w;
  r[3] = m[3] * x + m[7] * y + m[11] * z + m[15] * w;
}

struct Camera {
  float projInv[16];
  float worldViewInv[16];
  // worldViewInv * (0, 0, 0, 1), the eye position: the same float operations as
  // mul4 on it, evaluated once per camera on the host (make_camera)
  float origin[4];
};

__device__ __forceinline__ void eye_ray(const Camera& C, uint32_t x, uint32_t y, float dx, float dy, uint32_t W, uint32_t H, v3& pos,
                                        v3& dir) {
  float fx = ((float)x + dx) / (float)W;
  float fy = ((float)y + dy) / (float)H;
  float p[4];
  mul4(C.projInv, 2.0f * fx - 1.0f, -2.0f * fy + 1.0f, 1.0f, 1.0f, p);
  v3 rd = normalize(v3(p[0] / p[3], p[1] / p[3], p[2] / p[3]));
  float b[4];
  v3 q = v3(0, 0, 0) + 100.0f * rd;
  mul4(C.worldViewInv, q.x, q.y, q.z, 1.0f, b);
  pos = v3(C.origin[0], C.origin[1], C.origin[2]);
  dir = normalize(v3(b[0], b[1], b[2]) - pos);
}

// ---- PCG3D jitter, eye_ray.cpp:53-81 ----------------------------------------
__device__ __forceinline__ float2 rand2(uint32_t x, uint32_t y, uint32_t iter, uint32_t seed) {
  x = x + 1233u * (iter + seed) % 171u;
  y = y + 453u * (iter + seed) % 765u;
  uint32_t a = x, b = y, c = x ^ y;
  a = a * 1664525u + 1013904223u;
  b = b * 1664525u + 1013904223u;
  c = c * 1664525u + 1013904223u;
  a += b * c; b += c * a; c += a * b;
  a ^= a >> 16u; b ^= b >> 16u; c ^= c >> 16u;
  a += b * c; b += c * a; c += a * b;
  const float s = 1.0f / 4294967295.0f;
  return make_float2((float)a * s, (float)b * s);
}

// inverse of SuperBlockIndex2DOpt (eye_ray.cpp:537-555): packed index -> pixel
__device__ __forceinline__ void tid_to_xy(uint32_t tid, uint32_t packed_w, uint32_t& x, uint32_t& y) {
  const uint32_t sb = tid >> 6, r = tid & 63u;
  const uint32_t wbh = packed_w >> 3;
  const uint32_t bhx = sb % wbh, bhy = sb / wbh;
  const uint32_t lh = r >> 4, l = r & 15u;
  x = ((bhx * 2 + (lh & 1u)) << 2) + (l & 3u);
  y = ((bhy * 2 + (lh >> 1)) << 2) + (l >> 2);
}

}  // namespace lrtd

// oracle_math.hpp - vector helpers for the CPU oracle.
//
// TEST INFRASTRUCTURE ONLY. The oracle is the checker for the HIP path; only
// tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
//
// The reference's vector library (LiteMath, an empty git submodule in the
// reference tree) is not available, so its conventions are restated from the
// reference call sites:
//   * sign(0) == 0 (GLSL semantics; the reference's GPU path is generated GLSL
//     from the same source, raytrace_common.h:59-66 relies on it);
//   * min/max are the plain comparisons std::min/std::max use;
//   * normalize(v) = v / length(v), component-wise division;
//   * matrices are column-major 4x4 (LiteMath float4x4 m_col[4]); m*v sums the
//     four column products left to right.
// Everything is templated on the scalar type T so the same source gives the
// FP32 twin (parity against the GPU) and the FP64 error reference.
#pragma once
#include <cmath>
#include <cstdint>
#include <algorithm>

namespace orc {

template <typename T> struct V2 { T x, y; };
template <typename T> struct V3 {
  T x, y, z;
  V3() : x(0), y(0), z(0) {}
  V3(T a, T b, T c) : x(a), y(b), z(c) {}
  T& operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
  T operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};
template <typename T> struct V4 { T x, y, z, w; };

template <typename T> static inline V3<T> operator+(V3<T> a, V3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <typename T> static inline V3<T> operator-(V3<T> a, V3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <typename T> static inline V3<T> operator*(V3<T> a, V3<T> b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
template <typename T> static inline V3<T> operator/(V3<T> a, V3<T> b) { return {a.x / b.x, a.y / b.y, a.z / b.z}; }
template <typename T> static inline V3<T> operator*(T s, V3<T> a) { return {s * a.x, s * a.y, s * a.z}; }
template <typename T> static inline V3<T> operator*(V3<T> a, T s) { return {a.x * s, a.y * s, a.z * s}; }
template <typename T> static inline V3<T> operator/(V3<T> a, T s) { return {a.x / s, a.y / s, a.z / s}; }
template <typename T> static inline V3<T> operator+(V3<T> a, T s) { return {a.x + s, a.y + s, a.z + s}; }
template <typename T> static inline V3<T> operator-(V3<T> a, T s) { return {a.x - s, a.y - s, a.z - s}; }

template <typename T> static inline T mn(T a, T b) { return (b < a) ? b : a; }   // std::min
template <typename T> static inline T mx(T a, T b) { return (a < b) ? b : a; }   // std::max
template <typename T> static inline T sgn(T a) { return a > T(0) ? T(1) : (a < T(0) ? T(-1) : T(0)); }
template <typename T> static inline T clampv(T u, T a, T b) { return mn(mx(u, a), b); }

template <typename T> static inline T dot(V3<T> a, V3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <typename T> static inline T length(V3<T> a) { return std::sqrt(dot(a, a)); }
template <typename T> static inline V3<T> normalize(V3<T> a) { T l = length(a); return {a.x / l, a.y / l, a.z / l}; }
template <typename T> static inline V3<T> vfloor(V3<T> a) { return {std::floor(a.x), std::floor(a.y), std::floor(a.z)}; }
template <typename T> static inline V3<T> vclamp(V3<T> a, T lo, T hi) { return {clampv(a.x, lo, hi), clampv(a.y, lo, hi), clampv(a.z, lo, hi)}; }
template <typename T> static inline V3<T> vabs(V3<T> a) { return {std::abs(a.x), std::abs(a.y), std::abs(a.z)}; }

// 4x4 column-major: m[c*4 + r]
template <typename T> struct M4 { T m[16]; };

// LiteMath float4x4 * float4 as the camera code uses it (render_common.h:45-55,
// eye_ray.cpp:146-160); LiteMath itself is absent, its convention is unpinned
template <typename T> static inline V4<T> mul(const M4<T>& M, V4<T> v) {
  const T* m = M.m;
  V4<T> r;
  r.x = m[0] * v.x + m[4] * v.y + m[8] * v.z + m[12] * v.w;
  r.y = m[1] * v.x + m[5] * v.y + m[9] * v.z + m[13] * v.w;
  r.z = m[2] * v.x + m[6] * v.y + m[10] * v.z + m[14] * v.w;
  r.w = m[3] * v.x + m[7] * v.y + m[11] * v.z + m[15] * v.w;
  return r;
}
template <typename T> static inline V3<T> mul4x3(const M4<T>& M, V3<T> v) {
  V4<T> r = mul(M, V4<T>{v.x, v.y, v.z, T(1)});
  return {r.x, r.y, r.z};
}

// General 4x4 inverse by cofactor expansion, evaluated in double and rounded
// once to float, so every consumer of the same matrix sees the same bits. It
// stands for LiteMath's inverse4x4 in UpdateCamera (eye_ray_host.cpp:59-65),
// whose own arithmetic is unpinned.
static inline void inverse4x4_via_double(const float in[16], float out[16]) {
  double m[16], inv[16];
  for (int i = 0; i < 16; i++) m[i] = in[i];
  inv[0] = m[5] * m[10] * m[15] - m[5] * m[11] * m[14] - m[9] * m[6] * m[15] + m[9] * m[7] * m[14] + m[13] * m[6] * m[11] - m[13] * m[7] * m[10];
  inv[4] = -m[4] * m[10] * m[15] + m[4] * m[11] * m[14] + m[8] * m[6] * m[15] - m[8] * m[7] * m[14] - m[12] * m[6] * m[11] + m[12] * m[7] * m[10];
  inv[8] = m[4] * m[9] * m[15] - m[4] * m[11] * m[13] - m[8] * m[5] * m[15] + m[8] * m[7] * m[13] + m[12] * m[5] * m[11] - m[12] * m[7] * m[9];
  inv[12] = -m[4] * m[9] * m[14] + m[4] * m[10] * m[13] + m[8] * m[5] * m[14] - m[8] * m[6] * m[13] - m[12] * m[5] * m[10] + m[12] * m[6] * m[9];
  inv[1] = -m[1] * m[10] * m[15] + m[1] * m[11] * m[14] + m[9] * m[2] * m[15] - m[9] * m[3] * m[14] - m[13] * m[2] * m[11] + m[13] * m[3] * m[10];
  inv[5] = m[0] * m[10] * m[15] - m[0] * m[11] * m[14] - m[8] * m[2] * m[15] + m[8] * m[3] * m[14] + m[12] * m[2] * m[11] - m[12] * m[3] * m[10];
  inv[9] = -m[0] * m[9] * m[15] + m[0] * m[11] * m[13] + m[8] * m[1] * m[15] - m[8] * m[3] * m[13] - m[12] * m[1] * m[11] + m[12] * m[3] * m[9];
  inv[13] = m[0] * m[9] * m[14] - m[0] * m[10] * m[13] - m[8] * m[1] * m[14] + m[8] * m[2] * m[13] + m[12] * m[1] * m[10] - m[12] * m[2] * m[9];
  inv[2] = m[1] * m[6] * m[15] - m[1] * m[7] * m[14] - m[5] * m[2] * m[15] + m[5] * m[3] * m[14] + m[13] * m[2] * m[7] - m[13] * m[3] * m[6];
  inv[6] = -m[0] * m[6] * m[15] + m[0] * m[7] * m[14] + m[4] * m[2] * m[15] - m[4] * m[3] * m[14] - m[12] * m[2] * m[7] + m[12] * m[3] * m[6];
  inv[10] = m[0] * m[5] * m[15] - m[0] * m[7] * m[13] - m[4] * m[1] * m[15] + m[4] * m[3] * m[13] + m[12] * m[1] * m[7] - m[12] * m[3] * m[5];
  inv[14] = -m[0] * m[5] * m[14] + m[0] * m[6] * m[13] + m[4] * m[1] * m[14] - m[4] * m[2] * m[13] - m[12] * m[1] * m[6] + m[12] * m[2] * m[5];
  inv[3] = -m[1] * m[6] * m[11] + m[1] * m[7] * m[10] + m[5] * m[2] * m[11] - m[5] * m[3] * m[10] - m[9] * m[2] * m[7] + m[9] * m[3] * m[6];
  inv[7] = m[0] * m[6] * m[11] - m[0] * m[7] * m[10] - m[4] * m[2] * m[11] + m[4] * m[3] * m[10] + m[8] * m[2] * m[7] - m[8] * m[3] * m[6];
  inv[11] = -m[0] * m[5] * m[11] + m[0] * m[7] * m[9] + m[4] * m[1] * m[11] - m[4] * m[3] * m[9] - m[8] * m[1] * m[7] + m[8] * m[3] * m[5];
  inv[15] = m[0] * m[5] * m[10] - m[0] * m[6] * m[9] - m[4] * m[1] * m[10] + m[4] * m[2] * m[9] + m[8] * m[1] * m[6] - m[8] * m[2] * m[5];
  double det = m[0] * inv[0] + m[1] * inv[4] + m[2] * inv[8] + m[3] * inv[12];
  double idet = (det != 0.0) ? 1.0 / det : 0.0;
  for (int i = 0; i < 16; i++) out[i] = (float)(inv[i] * idet);
}

}  // namespace orc

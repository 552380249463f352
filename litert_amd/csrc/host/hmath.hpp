// hmath.hpp - small host-side vector/matrix helpers for the scene builders.
#pragma once
#include <cmath>
#include <cstdint>
#include <algorithm>

namespace lrt {

struct f3 {
  float x, y, z;
  f3() : x(0), y(0), z(0) {}
  f3(float a, float b, float c) : x(a), y(b), z(c) {}
  float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
  float& operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
};
static inline f3 operator+(f3 a, f3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
static inline f3 operator-(f3 a, f3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
static inline f3 operator*(float s, f3 a) { return {s * a.x, s * a.y, s * a.z}; }
static inline f3 operator*(f3 a, float s) { return {s * a.x, s * a.y, s * a.z}; }
static inline float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline f3 cross(f3 a, f3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
static inline float length(f3 a) { return std::sqrt(dot(a, a)); }
static inline f3 normalize(f3 a) { float l = length(a); return {a.x / l, a.y / l, a.z / l}; }
static inline f3 fmin3(f3 a, f3 b) { return {std::min(a.x, b.x), std::min(a.y, b.y), std::min(a.z, b.z)}; }
static inline f3 fmax3(f3 a, f3 b) { return {std::max(a.x, b.x), std::max(a.y, b.y), std::max(a.z, b.z)}; }

// Column-major 4x4 (element (r,c) at m[c*4+r]), the layout of LiteMath's float4x4.
struct m4 {
  float m[16];
};

// perspectiveMatrix(fovy_deg, aspect, near, far) with OpenGL conventions, as
// the reference calls it (tests/tests.cpp:29-41, eye_ray_host.cpp:76).
static inline m4 perspective(float fovy_deg, float aspect, float zn, float zf) {
  m4 r{};
  const float ymax = zn * std::tan(fovy_deg * 3.14159265358979323846f / 360.0f);
  const float xmax = ymax * aspect;
  r.m[0] = (2.0f * zn) / (2.0f * xmax);
  r.m[5] = (2.0f * zn) / (2.0f * ymax);
  r.m[10] = (-zf - zn) / (zf - zn);
  r.m[11] = -1.0f;
  r.m[14] = (-2.0f * zn * zf) / (zf - zn);
  return r;
}

// lookAt(eye, center, up) (gluLookAt convention).
static inline m4 look_at(f3 eye, f3 center, f3 up) {
  f3 z = normalize(eye - center);
  f3 x = normalize(cross(up, z));
  f3 y = normalize(cross(z, x));
  m4 M{};
  M.m[0] = x.x; M.m[4] = x.y; M.m[8] = x.z;  M.m[12] = -dot(x, eye);
  M.m[1] = y.x; M.m[5] = y.y; M.m[9] = y.z;  M.m[13] = -dot(y, eye);
  M.m[2] = z.x; M.m[6] = z.y; M.m[10] = z.z; M.m[14] = -dot(z, eye);
  M.m[3] = 0;   M.m[7] = 0;   M.m[11] = 0;   M.m[15] = 1;
  return M;
}

// 4x4 inverse by cofactor expansion in double, rounded once to float. Every
// consumer of a camera matrix (host, kernels, tests) uses this one rounding.
void inverse4x4(const float in[16], float out[16]);

}  // namespace lrt

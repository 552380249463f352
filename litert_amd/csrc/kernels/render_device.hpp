// render_device.hpp - per-pixel helpers shared by the render kernels:
// shading of a hit for RenderFloat and the sub-pixel sample offsets.
#pragma once
#include "render.hpp"

namespace lrtd {

// Phong's specular power: x^32 by five squarings (the reference calls pow)
__device__ __forceinline__ float pow32(float x) {
  x = x * x;
  x = x * x;
  x = x * x;
  x = x * x;
  return x * x;
}

// True for the modes whose colour needs shadow rays toward the lights.
__device__ __forceinline__ bool needs_shadow(uint32_t mode) {
  return mode == LRT_RENDER_MODE_PHONG || mode == LRT_RENDER_MODE_PHONG_NO_TEX;
}

__device__ __forceinline__ v3 light_v(const float4 f) { return v3(f.x, f.y, f.z); }
__device__ __forceinline__ uint32_t light_type(const LightSet& L, uint32_t i) { return __float_as_uint(L.sp[i].w); }

// Phong's shadow ray toward light i from just before the hit (eye_ray.cpp:
// 443-468: BIAS 1e-6; direct lights along their direction, point lights
// toward their position; the eye ray's tNear / tFar, so the ray is not
// stopped at the light). Returns false for ambient lights.
__device__ __forceinline__ bool shadow_ray(const LightSet& L, uint32_t i, v3 ray_pos, v3 ray_dir, float t, v3& origin, v3& dir) {
  const uint32_t ty = light_type(L, i);
  if (ty == LRT_LIGHT_TYPE_AMBIENT) return false;
  origin = ray_pos + (t - 1e-6f) * ray_dir;
  if (ty == LRT_LIGHT_TYPE_DIRECT) {
    dir = light_v(L.sp[i]);
  } else {
    const v3 d = light_v(L.sp[i]) - origin;
    const float l = sqrtf(dot(d, d));
    dir = v3(d.x / l, d.y / l, d.z / l);
  }
  return true;
}

// Render modes that apply to SBS geometry (eye_ray.cpp:162-489) with the
// renderer's lights. lit: bit i set if Phong's shadow ray toward light i found
// nothing. ray_pos: the eye ray's origin.
// Colours of the PRIMITIVE / TYPE / GEOM modes (eye_ray.cpp palette), read
// with a run-time index: a constant table, not a per-lane array in scratch.
static __constant__ uint32_t kPalette[20] = {0xffe6194b, 0xff3cb44b, 0xffffe119, 0xff0082c8, 0xfff58231, 0xff911eb4, 0xff46f0f0,
                                             0xfff032e6, 0xffd2f53c, 0xfffabebe, 0xff008080, 0xffe6beff, 0xffaa6e28, 0xfffffac8,
                                             0xff800000, 0xffaaffc3, 0xff808000, 0xffffd8b1, 0xff000080, 0xff808080};

__device__ __forceinline__ float4 shade(uint32_t mode, const Hit& hit, v3 ray_pos, v3 dir, const LightSet& L, uint32_t lit) {
  if (hit.primId == 0xFFFFFFFFu) return make_float4(0, 0, 0, 1);
  const float z = hit.t, z_near = 0.1f, z_far = 10.0f;
  const uint32_t type = hit.geomId >> LRT_SH_TYPE;
  const uint32_t geomId = hit.geomId & LRT_GEOM_ID_MASK;
  v3 norm = decode_normal(hit.coords[2], hit.coords[3]);
  const float ns = sgn(dot(-1.0f * dir, norm));
  norm = norm * ns;
  const uint32_t* pal = kPalette;
  auto palc = [](uint32_t c) {
    return make_float4((float)(c & 0xFF) * (1.0f / 255.0f), (float)((c >> 8) & 0xFF) * (1.0f / 255.0f),
                       (float)((c >> 16) & 0xFF) * (1.0f / 255.0f), (float)((c >> 24) & 0xFF) * (1.0f / 255.0f));
  };
  switch (mode) {
    case LRT_RENDER_MODE_MASK: return make_float4(1, 1, 1, 1);
    case LRT_RENDER_MODE_DEPTH: { float d = (1 / z - 1 / z_near) / (1 / z_far - 1 / z_near); return make_float4(d, d, d, 1); }
    case LRT_RENDER_MODE_LINEAR_DEPTH: { float d = ((z - z_near) / (z_far - z_near)); return make_float4(d, d, d, 1); }
    case LRT_RENDER_MODE_INVERSE_LINEAR_DEPTH: { float d = 1 - ((z - z_near) / (z_far - z_near)); return make_float4(d, d, d, 1); }
    case LRT_RENDER_MODE_PRIMITIVE: return palc(pal[hit.primId % 20]);
    case LRT_RENDER_MODE_TYPE: return palc(pal[type % 20]);
    case LRT_RENDER_MODE_GEOM: return palc(pal[geomId % 20]);
    case LRT_RENDER_MODE_NORMAL: return make_float4(fabsf(norm.x), fabsf(norm.y), fabsf(norm.z), 1);
    case LRT_RENDER_MODE_ST_ITERATIONS: {
      if (hit.primId == 0) return make_float4(0.5f, 0.5f, 0.5f, 1);
      if (hit.primId <= 64) {
        float q = clampf_((float)hit.primId / 64, 0.0f, 1.0f);
        return make_float4(0, (1 - q), q, q + (1 - q));
      }
      float q = clampf_((float)(hit.primId - 64) / (256 - 64), 0.0f, 1.0f);
      return make_float4(q, 0, (1 - q), q + (1 - q));
    }
    case LRT_RENDER_MODE_DIFFUSE: {
      if (type == LRT_TYPE_SDF_SBS_COL)
        return make_float4(roundf(hit.coords[0]) / 255.0f, hit.coords[0] - floorf(hit.coords[0]), roundf(hit.coords[1]) / 255.0f, 1);
      return make_float4(0, 1, 1, 1);
    }
    case LRT_RENDER_MODE_BARYCENTRIC: return make_float4(hit.coords[0], hit.coords[1], 1 - hit.coords[0] - hit.coords[1], 1);
    case LRT_RENDER_MODE_TEX_COORDS: return make_float4(hit.coords[0], hit.coords[1], 0, 1);
    case LRT_RENDER_MODE_LAMBERT_NO_TEX:
    case LRT_RENDER_MODE_LAMBERT:
    case LRT_RENDER_MODE_PHONG:
    case LRT_RENDER_MODE_PHONG_NO_TEX: {
      // colour: white for the NO_TEX modes, the brick colour for colour
      // layouts, else the default material's (0,1,1) texture (eye_ray.cpp:376-392, 426-441)
      v3 col(1, 1, 1);
      if (mode == LRT_RENDER_MODE_LAMBERT || mode == LRT_RENDER_MODE_PHONG) {
        col = type == LRT_TYPE_SDF_SBS_COL
                  ? v3(roundf(hit.coords[0]) / 255.0f, hit.coords[0] - floorf(hit.coords[0]), roundf(hit.coords[1]) / 255.0f)
                  : v3(0, 1, 1);
      }
      const bool phong = mode == LRT_RENDER_MODE_PHONG || mode == LRT_RENDER_MODE_PHONG_NO_TEX;
      v3 c(0, 0, 0);
      for (uint32_t i = 0; i < L.n; i++) {
        const uint32_t ty = light_type(L, i);
        const v3 lc = light_v(L.col[i]);
        if (ty == LRT_LIGHT_TYPE_AMBIENT) {
          c = c + lc * col;
          continue;
        }
        if (!phong) {  // Lambert, eye_ray.cpp:394-418
          if (ty == LRT_LIGHT_TYPE_DIRECT) {
            const float q = mx(0.0f, dot(norm, light_v(L.sp[i])));
            c = c + (lc * col) * q;
          } else {
            const v3 surf = ray_pos + hit.t * dir;
            v3 ld = light_v(L.sp[i]) - surf;
            const float l = sqrtf(dot(ld, ld));
            ld = v3(ld.x / l, ld.y / l, ld.z / l);
            const float q = mx(0.0f, dot(norm, ld));
            const v3 e = (lc * col) * q;
            c = c + v3(e.x / (l * l), e.y / (l * l), e.z / (l * l));
          }
          continue;
        }
        // Phong, eye_ray.cpp:443-475: the direct light's Lambert term, then
        // the shadowed diffuse + specular term of every non-ambient light
        v3 o, ld;
        shadow_ray(L, i, ray_pos, dir, hit.t, o, ld);
        if (ty == LRT_LIGHT_TYPE_DIRECT) c = c + (lc * col) * mx(0.0f, dot(norm, ld));
        const float Kd = 0.25f, Ks = 0.25f;
        const v3 refl = -1.0f * ld + (2.0f * dot(norm, ld)) * norm;
        const float e = Kd * mx(0.0f, dot(norm, ld)) + Ks * pow32(mx(0.0f, dot(norm, refl)));
        const float shadow = (lit >> i) & 1u ? 1.0f : 0.0f;
        c = c + ((shadow * lc) * e) * col;
      }
      return make_float4(c.x, c.y, c.z, 1);
    }
    default: return make_float4(1, 0, 1, 1);
  }
}

// floor(sqrt(n)), exact for every uint32 n
__device__ __forceinline__ uint32_t isqrt_u32(uint32_t n) {
  uint64_t r = (uint64_t)sqrtf((float)n);
  while (r * r > n) r--;
  while ((r + 1) * (r + 1) <= n) r++;
  return (uint32_t)r;
}

// The reference computes uint(sqrt(spp)) and i / spp_sqrt + 0.5 in double;
// both are exact in the integer / FP32 forms here (k + 0.5 is exact in float
// for k < 2^23, and the float product of two floats rounds the exact product
// once, as the double product converted to float does).
__device__ __forceinline__ float2 sample_offset(const RenderArgs& A, uint32_t x, uint32_t y, uint32_t i) {
  if (A.ray_gen_mode == LRT_RAY_GEN_MODE_RANDOM) return rand2(x, y, i + A.seed % 2, A.seed);
  const uint32_t spp_sqrt = isqrt_u32(A.spp);
  const float is = 1.0f / (float)spp_sqrt;
  return make_float2(is * ((float)(i / spp_sqrt) + 0.5f), is * ((float)(i % spp_sqrt) + 0.5f));
}

}  // namespace lrtd

// mlp_device.hpp - device side of the learned root-find seeder: the 24-float
// ray/brick feature vector and one wavefront's BF16 MFMA forward pass.
#pragma once
#include "mlp.hpp"
#include "trace_device.hpp"

namespace lrtd {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// Two FP32 values -> packed BF16 (lo in bits 0..15), round to nearest even:
// one v_cvt_pk_bf16_f32 on gfx950.
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint32_t pack_bf16(float lo, float hi) {
  const bf16x2 v = {(__bf16)lo, (__bf16)hi};
  return __builtin_bit_cast(uint32_t, v);
}

// SiLU with the hardware exp2 / reciprocal: x * 1 / (1 + 2^(-x log2 e)).
__device__ __forceinline__ float silu_fast(float x) {
  return x * __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(-1.44269504f * x));
}

// Brick geometry shared by the feature builder and the seeded intersector.
struct BrickFrame {
  v3 bmin, bmax;
  float s;  // world -> brick units (1 / brick edge)
};

__device__ __forceinline__ BrickFrame brick_frame(const DevScene& S, uint32_t brick) {
  const uint4 N = S.nodes[brick];
  const float px = (float)(N.x >> 16), py = (float)(N.x & 0xFFFFu), pz = (float)(N.y >> 16), sz = (float)(N.y & 0xFFFFu);
  const float sz_inv = 2.0f / sz;
  BrickFrame F;
  F.bmin = v3(-1, -1, -1) + sz_inv * v3(px, py, pz);
  F.bmax = F.bmin + sz_inv * v3(1, 1, 1);
  F.s = 0.5f * sz;
  return F;
}

// Feature vector at leaf entry (brick units, translation invariant):
//   [0..2]  ray direction
//   [3..5]  entry point in the brick, [0,1]^3
//   [6]     length of the ray segment inside the brick
//   [7]     entry t relative to the ray's closest approach to the brick centre
//   [8..15] distances at the 8 brick corners      (per-brick table)
//   [16..23] minimum distance of each brick octant (per-brick table)
// Returns the brick entry t0 (world) through *t0.
__device__ __forceinline__ void mlp_features(const DevScene& S, const float4* __restrict__ bfeat, v3 pos, v3 dir, v3 inv,
                                             uint32_t brick, float f[24], float* t0, float* s_out) {
  const BrickFrame F = brick_frame(S, brick);
  const float2 bnf = RayBox(pos, inv, F.bmin, F.bmax);
  const v3 e = (pos + bnf.x * dir - F.bmin) * F.s;
  const v3 c = F.bmin + (0.5f / F.s) * v3(1, 1, 1);
  const float tc = dot(c - pos, dir);
  f[0] = dir.x; f[1] = dir.y; f[2] = dir.z;
  f[3] = e.x; f[4] = e.y; f[5] = e.z;
  f[6] = (bnf.y - bnf.x) * F.s;
  f[7] = (bnf.x - tc) * F.s;
  const float4* bf = bfeat + 4 * (size_t)brick;
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const float4 v = bf[q];
    f[8 + 4 * q + 0] = v.x; f[8 + 4 * q + 1] = v.y; f[8 + 4 * q + 2] = v.z; f[8 + 4 * q + 3] = v.w;
  }
  *t0 = bnf.x;
  *s_out = F.s;
}

// One wavefront's forward pass for up to 64 samples whose features sit in
// LDS as bf16 rows of kFeatStride halves (slot-major). `n_tiles` 16-sample
// tiles are evaluated (wave-uniform); outputs (z0, z1, z2) are written back to
// `out` (3 floats per slot). `packed` holds the A tiles and biases in the
// layout described in mlp.hpp. Must be called with the whole wave active.
constexpr int kFeatStride = 40;  // bf16 per feature row in LDS (80 B: keeps 16-B alignment of the 8-half loads)

__device__ __forceinline__ bf16x8 load_a_tile(const uint4* __restrict__ tiles, int tile, int lane) {
  const uint4 v = tiles[tile * 64 + lane];
  return __builtin_bit_cast(bf16x8, v);
}

__device__ __forceinline__ f32x4 load_bias(const float4* __restrict__ bias, int btile, int lane) {
  const float4 b = bias[btile * 64 + lane];
  return f32x4{b.x, b.y, b.z, b.w};
}

// silu of two accumulator tiles -> the 8-half B operand of the next layer
__device__ __forceinline__ bf16x8 act_pair(f32x4 lo, f32x4 hi) {
  uint4 u;
  u.x = pack_bf16(silu_fast(lo[0]), silu_fast(lo[1]));
  u.y = pack_bf16(silu_fast(lo[2]), silu_fast(lo[3]));
  u.z = pack_bf16(silu_fast(hi[0]), silu_fast(hi[1]));
  u.w = pack_bf16(silu_fast(hi[2]), silu_fast(hi[3]));
  return __builtin_bit_cast(bf16x8, u);
}

__device__ __forceinline__ void mlp_wave_forward(const void* __restrict__ packed, const uint16_t* feat_lds, float* out_lds, int n_tiles,
                                                 int lane) {
  const int row = lane & 15, grp = lane >> 4;
  for (int nt = 0; nt < n_tiles; nt++) {
    // Re-derive the weight pointers each tile behind an empty asm barrier so
    // the compiler streams the A tiles from cache per MFMA instead of hoisting
    // all 22 of them (88 VGPRs) out of the loop.
    const char* pk = (const char*)packed;
    asm volatile("" : "+s"(pk));
    const uint4* tiles = (const uint4*)pk;
    const float4* bias = (const float4*)(pk + (size_t)kTilesTotal * 64 * 16);
    // layer 1: B = features of slot 16*nt + row, K slots 8*grp .. 8*grp+7
    const int slot = 16 * nt + row;
    const uint4 xb = *(const uint4*)(feat_lds + slot * kFeatStride + 8 * grp);
    const bf16x8 x = __builtin_bit_cast(bf16x8, xb);
    f32x4 h[4];
#pragma unroll
    for (int mt = 0; mt < 4; mt++) {
      h[mt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(load_a_tile(tiles, kTileOffL1 + mt, lane), x, load_bias(bias, mt, lane), 0, 0, 0);
    }
    // layers 2 and 3
#pragma unroll
    for (int layer = 0; layer < 2; layer++) {
      const int toff = layer == 0 ? kTileOffL2 : kTileOffL3;
      const int boff = layer == 0 ? 4 : 8;
      const bf16x8 b0 = act_pair(h[0], h[1]);
      const bf16x8 b1 = act_pair(h[2], h[3]);
      f32x4 g[4];
#pragma unroll
      for (int mt = 0; mt < 4; mt++) {
        f32x4 acc = load_bias(bias, boff + mt, lane);
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(load_a_tile(tiles, toff + 2 * mt + 0, lane), b0, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(load_a_tile(tiles, toff + 2 * mt + 1, lane), b1, acc, 0, 0, 0);
        g[mt] = acc;
      }
#pragma unroll
      for (int mt = 0; mt < 4; mt++) h[mt] = g[mt];
    }
    // layer 4: rows 0..2 of the single output tile
    const bf16x8 b0 = act_pair(h[0], h[1]);
    const bf16x8 b1 = act_pair(h[2], h[3]);
    f32x4 o = load_bias(bias, 12, lane);
    o = __builtin_amdgcn_mfma_f32_16x16x32_bf16(load_a_tile(tiles, kTileOffL4 + 0, lane), b0, o, 0, 0, 0);
    o = __builtin_amdgcn_mfma_f32_16x16x32_bf16(load_a_tile(tiles, kTileOffL4 + 1, lane), b1, o, 0, 0, 0);
    if (grp == 0) {  // lane `row` owns rows 0..3 of sample slot 16*nt + row
      out_lds[3 * slot + 0] = o[0];
      out_lds[3 * slot + 1] = o[1];
      out_lds[3 * slot + 2] = o[2];
    }
  }
}

// The same forward pass for all 64 slots (4 tiles) at once: each A tile and
// bias tile is loaded once per wave and applied to the 4 sample tiles, a
// quarter of the weight traffic of mlp_wave_forward. Per sample the MFMA
// operands and accumulation order are those of mlp_wave_forward, so the
// outputs are identical. Holds 16 accumulator tiles (64 VGPRs); meant for
// kernels with registers to spare (batch inference), not the render loop.
__device__ __forceinline__ void mlp_wave_forward64(const void* __restrict__ packed, const uint16_t* feat_lds, float* out_lds, int lane) {
  const int row = lane & 15, grp = lane >> 4;
  const uint4* tiles = (const uint4*)packed;
  const float4* bias = (const float4*)((const char*)packed + (size_t)kTilesTotal * 64 * 16);
  f32x4 h[4][4];  // [sample tile][row tile]
  bf16x8 b0[4], b1[4];
#pragma unroll
  for (int nt = 0; nt < 4; nt++) {
    const uint4 xb = *(const uint4*)(feat_lds + (16 * nt + row) * kFeatStride + 8 * grp);
    b0[nt] = __builtin_bit_cast(bf16x8, xb);
  }
#pragma unroll
  for (int mt = 0; mt < 4; mt++) {
    const bf16x8 a = load_a_tile(tiles, kTileOffL1 + mt, lane);
    const f32x4 bb = load_bias(bias, mt, lane);
#pragma unroll
    for (int nt = 0; nt < 4; nt++) h[nt][mt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b0[nt], bb, 0, 0, 0);
  }
#pragma unroll
  for (int layer = 0; layer < 2; layer++) {
    const int toff = layer == 0 ? kTileOffL2 : kTileOffL3;
    const int boff = layer == 0 ? 4 : 8;
#pragma unroll
    for (int nt = 0; nt < 4; nt++) {
      b0[nt] = act_pair(h[nt][0], h[nt][1]);
      b1[nt] = act_pair(h[nt][2], h[nt][3]);
    }
#pragma unroll
    for (int mt = 0; mt < 4; mt++) {
      const bf16x8 a0 = load_a_tile(tiles, toff + 2 * mt + 0, lane);
      const bf16x8 a1 = load_a_tile(tiles, toff + 2 * mt + 1, lane);
      const f32x4 bb = load_bias(bias, boff + mt, lane);
#pragma unroll
      for (int nt = 0; nt < 4; nt++) {
        f32x4 acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, b0[nt], bb, 0, 0, 0);
        h[nt][mt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, b1[nt], acc, 0, 0, 0);
      }
    }
  }
  const bf16x8 a0 = load_a_tile(tiles, kTileOffL4 + 0, lane);
  const bf16x8 a1 = load_a_tile(tiles, kTileOffL4 + 1, lane);
  const f32x4 bb = load_bias(bias, 12, lane);
#pragma unroll
  for (int nt = 0; nt < 4; nt++) {
    f32x4 o = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, act_pair(h[nt][0], h[nt][1]), bb, 0, 0, 0);
    o = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, act_pair(h[nt][2], h[nt][3]), o, 0, 0, 0);
    if (grp == 0) {
      const int slot = 16 * nt + row;
      out_lds[3 * slot + 0] = o[0];
      out_lds[3 * slot + 1] = o[1];
      out_lds[3 * slot + 2] = o[2];
    }
  }
}

// Write one lane's features into its LDS row as bf16 (slots 24..39 zero).
__device__ __forceinline__ void store_features_bf16(uint16_t* row, const float f[24]) {
  uint4* r4 = (uint4*)row;
#pragma unroll
  for (int q = 0; q < 3; q++) {
    uint4 u;
    u.x = pack_bf16(f[8 * q + 0], f[8 * q + 1]);
    u.y = pack_bf16(f[8 * q + 2], f[8 * q + 3]);
    u.z = pack_bf16(f[8 * q + 4], f[8 * q + 5]);
    u.w = pack_bf16(f[8 * q + 6], f[8 * q + 7]);
    r4[q] = u;
  }
  r4[3] = make_uint4(0, 0, 0, 0);
}

}  // namespace lrtd

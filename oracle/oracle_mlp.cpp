// oracle_mlp.cpp - FP64 reference of the seeding MLP (no reference counterpart).
//
// TEST INFRASTRUCTURE ONLY (see lrt_oracle.h). Network: 24 -> 64 -> 64 -> 64 -> 3,
// SiLU on the hidden layers, linear outputs (t_enter, t_exit, p_hit logit).
// Parameters (row-major, out x in): W1[64x24] b1 W2[64x64] b2 W3[64x64] b3
// W4[3x64] b4, 10115 floats. Training objective per sample, with target
// (t_hit, hit) and margin m = target[2]:
//   L = BCE(sigmoid(z2), hit) + hit * ((z0 - (t_hit - m))^2 + (z1 - (t_hit + m))^2)
// averaged over the batch.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include "lrt_oracle.h"
#include "oracle_scene.hpp"

namespace {

constexpr int kIn = 24, kH = 64, kOut = 3;
constexpr int kOffW1 = 0, kOffB1 = kOffW1 + kH * kIn, kOffW2 = kOffB1 + kH, kOffB2 = kOffW2 + kH * kH, kOffW3 = kOffB2 + kH,
              kOffB3 = kOffW3 + kH * kH, kOffW4 = kOffB3 + kH, kOffB4 = kOffW4 + kOut * kH, kParams = kOffB4 + kOut;

// round-to-nearest-even to bfloat16 (top 16 bits of an IEEE binary32)
double bf16(double x) {
  float f = (float)x;
  uint32_t u;
  std::memcpy(&u, &f, 4);
  if ((u & 0x7F800000u) == 0x7F800000u) return f;
  uint32_t lsb = (u >> 16) & 1u;
  u = (u + 0x7FFFu + lsb) & 0xFFFF0000u;
  std::memcpy(&f, &u, 4);
  return f;
}

double silu(double x) { return x / (1.0 + std::exp(-x)); }
double dsilu(double x) {
  double s = 1.0 / (1.0 + std::exp(-x));
  return s * (1.0 + x * (1.0 - s));
}

struct Acts {
  double a0[kIn], z1[kH], a1[kH], z2[kH], a2[kH], z3[kH], a3[kH], z4[kOut];
};

void forward(const float* p, const float* x, bool emu, Acts& A) {
  auto q = [&](double v) { return emu ? bf16(v) : v; };
  for (int i = 0; i < kIn; i++) A.a0[i] = q(x[i]);
  auto layer = [&](const double* in, int nin, int nout, int offW, int offB, double* z) {
    for (int o = 0; o < nout; o++) {
      double s = p[offB + o];
      for (int i = 0; i < nin; i++) s += q(p[offW + o * nin + i]) * in[i];
      z[o] = s;
    }
  };
  layer(A.a0, kIn, kH, kOffW1, kOffB1, A.z1);
  for (int i = 0; i < kH; i++) A.a1[i] = q(silu(A.z1[i]));
  layer(A.a1, kH, kH, kOffW2, kOffB2, A.z2);
  for (int i = 0; i < kH; i++) A.a2[i] = q(silu(A.z2[i]));
  layer(A.a2, kH, kH, kOffW3, kOffB3, A.z3);
  for (int i = 0; i < kH; i++) A.a3[i] = q(silu(A.z3[i]));
  layer(A.a3, kH, kOut, kOffW4, kOffB4, A.z4);
}

}  // namespace

extern "C" void orc_mlp_forward(const float* params, const float* x, uint32_t n, int bf16_emulate, double* y) {
  for (uint32_t s = 0; s < n; s++) {
    Acts A;
    forward(params, x + (size_t)s * kIn, bf16_emulate != 0, A);
    for (int o = 0; o < kOut; o++) y[(size_t)s * kOut + o] = A.z4[o];
  }
}

extern "C" double orc_mlp_loss_grad(const float* params, const float* x, const float* target, uint32_t n, int bf16_emulate,
                                    double* grad) {
  std::vector<double> g(kParams, 0.0);
  double loss = 0.0;
  const bool emu = bf16_emulate != 0;
  auto q = [&](double v) { return emu ? bf16(v) : v; };
  for (uint32_t s = 0; s < n; s++) {
    Acts A;
    forward(params, x + (size_t)s * kIn, emu, A);
    const float* tg = target + (size_t)s * 3;
    const double h = tg[1] > 0.5f ? 1.0 : 0.0, th = tg[0], m = tg[2];
    const double z2 = A.z4[2];
    // BCE with logits, stable form
    loss += std::max(z2, 0.0) - z2 * h + std::log1p(std::exp(-std::fabs(z2)));
    double dz[kOut];
    dz[2] = 1.0 / (1.0 + std::exp(-z2)) - h;
    const double e0 = A.z4[0] - (th - m), e1 = A.z4[1] - (th + m);
    loss += h * (e0 * e0 + e1 * e1);
    dz[0] = h * 2.0 * e0;
    dz[1] = h * 2.0 * e1;
    // backward (weights as used in the forward pass)
    double d3[kH], d2[kH], d1[kH];
    for (int o = 0; o < kOut; o++) {
      g[kOffB4 + o] += dz[o];
      for (int i = 0; i < kH; i++) g[kOffW4 + o * kH + i] += dz[o] * A.a3[i];
    }
    for (int i = 0; i < kH; i++) {
      double s = 0;
      for (int o = 0; o < kOut; o++) s += q(params[kOffW4 + o * kH + i]) * dz[o];
      d3[i] = s * dsilu(A.z3[i]);
    }
    for (int o = 0; o < kH; o++) {
      g[kOffB3 + o] += d3[o];
      for (int i = 0; i < kH; i++) g[kOffW3 + o * kH + i] += d3[o] * A.a2[i];
    }
    for (int i = 0; i < kH; i++) {
      double s = 0;
      for (int o = 0; o < kH; o++) s += q(params[kOffW3 + o * kH + i]) * d3[o];
      d2[i] = s * dsilu(A.z2[i]);
    }
    for (int o = 0; o < kH; o++) {
      g[kOffB2 + o] += d2[o];
      for (int i = 0; i < kH; i++) g[kOffW2 + o * kH + i] += d2[o] * A.a1[i];
    }
    for (int i = 0; i < kH; i++) {
      double s = 0;
      for (int o = 0; o < kH; o++) s += q(params[kOffW2 + o * kH + i]) * d2[o];
      d1[i] = s * dsilu(A.z1[i]);
    }
    for (int o = 0; o < kH; o++) {
      g[kOffB1 + o] += d1[o];
      for (int i = 0; i < kIn; i++) g[kOffW1 + o * kIn + i] += d1[o] * A.a0[i];
    }
  }
  const double inv = n ? 1.0 / n : 0.0;
  for (int k = 0; k < kParams; k++) grad[k] = g[k] * inv;
  return loss * inv;
}

// ---- seeder features (restating this build's own definition, FP32) --------
// Per-brick table: distances at the 8 brick corners, then each octant's
// minimum voxel-corner distance (voxels with 2*v >= brick_size on an axis are
// in that axis's upper half; an octant without voxels takes the brick
// minimum), all multiplied by 0.5*lod_size (brick units).
extern "C" int orc_mlp_brick_features(const orc_scene* s, uint32_t brick, float out[16]) {
  using namespace orc;
  if (!s || !out || brick >= s->s.nodes.size()) return -1;
  Options O;
  O.preset = MultiRenderPreset{0, 0, 0, 0, 0, 1, 1};
  Tracer<float> tr(s->s, O, nullptr);
  const SdfSBSHeader& h = s->s.header;
  const uint32_t bs = h.brick_size, v_size = h.brick_size + 2 * h.brick_pad + 1;
  const float sz = float(s->s.nodes[brick].pos_z_lod_size & 0xFFFF);
  const float sz_inv = 2.0f / sz, sc = 0.5f * sz;
  float corner[8], omin[8], all_min = 1e30f;
  for (int i = 0; i < 8; i++) { corner[i] = 0.0f; omin[i] = 1e30f; }
  for (uint32_t vx = 0; vx < bs; vx++)
    for (uint32_t vy = 0; vy < bs; vy++)
      for (uint32_t vz = 0; vz < bs; vz++) {
        const uint32_t vp[3] = {vx, vy, vz};
        float v[8];
        const float vmin = tr.load_distance_values(brick, vp, v_size, sz_inv, v);
        const int o = ((2 * vx >= bs) ? 4 : 0) | ((2 * vy >= bs) ? 2 : 0) | ((2 * vz >= bs) ? 1 : 0);
        omin[o] = std::min(omin[o], vmin);
        all_min = std::min(all_min, vmin);
        for (int c = 0; c < 8; c++) {
          const uint32_t cx = (c >> 2) & 1, cy = (c >> 1) & 1, cz = c & 1;
          if (vx == cx * (bs - 1) && vy == cy * (bs - 1) && vz == cz * (bs - 1)) corner[c] = v[c];
        }
      }
  for (int o = 0; o < 8; o++)
    if (omin[o] > 1e29f) omin[o] = all_min;
  for (int i = 0; i < 8; i++) { out[i] = corner[i] * sc; out[8 + i] = omin[i] * sc; }
  return 0;
}

// 24 features of a ray entering `brick`: direction; entry point in the brick
// ([0,1]^3); segment length; entry t relative to the closest approach to the
// brick centre (both in brick units); then the per-brick table. ray = pos
// xyz, dir xyz. *t0 receives the brick entry distance (world).
extern "C" int orc_mlp_features(const orc_scene* s, const float ray[6], uint32_t brick, float out[24], float* t0) {
  using namespace orc;
  if (!s || !ray || !out || brick >= s->s.nodes.size()) return -1;
  const SdfSBSNode& N = s->s.nodes[brick];
  const float px = float(N.pos_xy >> 16), py = float(N.pos_xy & 0xFFFF), pz = float(N.pos_z_lod_size >> 16);
  const float sz = float(N.pos_z_lod_size & 0xFFFF);
  const float sz_inv = 2.0f / sz;
  const V3<float> pos{ray[0], ray[1], ray[2]}, dir{ray[3], ray[4], ray[5]};
  const V3<float> bmin = V3<float>{-1, -1, -1} + sz_inv * V3<float>{px, py, pz};
  const V3<float> bmax = bmin + sz_inv * V3<float>{1, 1, 1};
  const float sc = 0.5f * sz;
  const V2<float> bnf = RayBoxIntersection2(pos, SafeInverse(dir), bmin, bmax);
  const V3<float> e = (pos + bnf.x * dir - bmin) * sc;
  const V3<float> c = bmin + (0.5f / sc) * V3<float>{1, 1, 1};
  const float tc = dot(c - pos, dir);
  out[0] = dir.x; out[1] = dir.y; out[2] = dir.z;
  out[3] = e.x; out[4] = e.y; out[5] = e.z;
  out[6] = (bnf.y - bnf.x) * sc;
  out[7] = (bnf.x - tc) * sc;
  if (t0) *t0 = bnf.x;
  return orc_mlp_brick_features(s, brick, out + 8);
}

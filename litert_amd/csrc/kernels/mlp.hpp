// mlp.hpp - the learned root-find seeder: network shape, parameter and packed
// layouts, and launch interface (host + device).
//
// Network: 24 -> 64 -> 64 -> 64 -> 3, SiLU on the hidden layers, linear
// outputs z0 (hit t - margin), z1 (hit t + margin) in brick units from the
// brick entry, z2 (hit logit). FP32 parameters, row-major out x in:
//   W1[64x24] b1[64] W2[64x64] b2[64] W3[64x64] b3[64] W4[3x64] b4[3]
// Inference runs in BF16 on MFMA (v_mfma_f32_16x16x32_bf16) with FP32
// accumulation; training runs in FP32.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "render.hpp"

namespace lrtd {

constexpr int kMlpIn = 24, kMlpH = 64, kMlpOut = 3;
constexpr int kOffW1 = 0, kOffB1 = kOffW1 + kMlpH * kMlpIn, kOffW2 = kOffB1 + kMlpH, kOffB2 = kOffW2 + kMlpH * kMlpH,
              kOffW3 = kOffB2 + kMlpH, kOffB3 = kOffW3 + kMlpH * kMlpH, kOffW4 = kOffB3 + kMlpH, kOffB4 = kOffW4 + kMlpOut * kMlpH,
              kMlpParams = kOffB4 + kMlpOut;
// Row stride (floats) of the training kernel's per-block partial gradients:
// kMlpParams rounded up to 64, so every row starts 256-B aligned.
constexpr int kMlpPartialStride = (kMlpParams + 63) & ~63;

// Packed inference weights. Each 16x32 A tile of v_mfma_f32_16x16x32_bf16 is
// stored as 64 lanes x 8 bf16 (16 B per lane, 1 KiB per tile) in the order
// the instruction reads it: lane l holds row 16*mt + l%16, K slots 8*(l/16)+j.
// Layer 1 maps K slot k of tile 0 to input feature k (k >= 24 is zero).
// Layers 2-4 take their B operand straight from the previous layer's 16x16
// accumulators, so K slot (l/16, j) of tile kt is hidden unit
//   32*kt + 16*(j/4) + 4*(l/16) + (j%4),
// and the weights are packed with the same permutation.
constexpr int kTilesL1 = 4 * 1, kTilesL2 = 4 * 2, kTilesL3 = 4 * 2, kTilesL4 = 1 * 2;
constexpr int kTileOffL1 = 0, kTileOffL2 = kTileOffL1 + kTilesL1, kTileOffL3 = kTileOffL2 + kTilesL2, kTileOffL4 = kTileOffL3 + kTilesL3,
              kTilesTotal = kTileOffL4 + kTilesL4;
// Biases follow the tiles: per layer and 16-row tile mt, 64 lanes x float4,
// lane l holding rows 16*mt + 4*(l/16) + i (the accumulator rows it owns).
constexpr int kBiasTiles = 4 + 4 + 4 + 1;
constexpr size_t kPackedBytes = (size_t)kTilesTotal * 64 * 16 + (size_t)kBiasTiles * 64 * 16;

// Per-brick features (brick units): 8 lattice-corner distances, then the
// minimum distance of each of the 8 octants (voxels with 2*v >= bs on an axis
// are in that axis's upper half), as 4 float4 per brick.
constexpr int kBrickFeat = 16;

struct MlpTrainArgs {
  const float* feats;      // [n][24]
  const float* tgts;       // [n][3]: t_hit (brick units), hit (0/1), margin
  uint64_t n;
  uint64_t perm_a, perm_b;  // sample of step slot i: (perm_a * i + perm_b) mod n, perm_a coprime to n
  uint64_t first;           // first slot of this minibatch
  uint32_t batch;           // samples in this minibatch (<= blocks * 64)
};

struct AdamArgs {
  float lr, b1, b2, eps;
  float c1, c2;  // 1/(1-b1^t), 1/(1-b2^t)
};

// event records written by the collector
struct MlpEventMeta {
  uint32_t ray;    // view * W * H + y * W + x
  uint32_t brick;  // brick index
};

// LDS stack entries per lane of the seeded render kernels (the LDS feature
// and output rows take the rest of the budget).
constexpr uint32_t kStackLdsSeeded = 16;

}  // namespace lrtd

namespace lrt {
hipError_t launch_brick_features(const lrtd::DevScene& S, float4* d_bfeat, hipStream_t stream);
hipError_t launch_mlp_collect(const lrtd::DevScene& S, const lrtd::TraceParams& P, const float4* d_bfeat, const lrtd::Camera* d_cams,
                              const lrtd::RenderArgs& A, float* d_feat, float* d_tgt, lrtd::MlpEventMeta* d_meta, uint64_t cap,
                              unsigned long long* d_count, uint32_t* d_spill, uint32_t blocks, hipStream_t stream);
hipError_t launch_mlp_infer(const void* d_packed, const float* d_feat, uint64_t n, float* d_out, uint32_t blocks, hipStream_t stream);
hipError_t launch_mlp_grad(const float* d_params, const lrtd::MlpTrainArgs& T, float* d_partial, float* d_partial_loss, uint32_t blocks,
                           hipStream_t stream);
hipError_t launch_mlp_reduce_adam(float* d_params, float* d_m, float* d_v, const float* d_partial, const float* d_partial_loss,
                                  uint32_t parts, uint32_t batch, const lrtd::AdamArgs& adam, bool apply, float* d_grad_out,
                                  float* d_loss_out, hipStream_t stream);
hipError_t launch_render_seeded(const lrtd::DevScene& S, const lrtd::TraceParams& P, const float4* d_bfeat, const void* d_packed,
                                float seed_logit, const lrtd::Camera* d_cams, const lrtd::RenderArgs& A, const lrtd::RenderOut& O,
                                uint32_t* d_work_counter, uint32_t* d_spill, uint32_t blocks, bool rgba, unsigned long long* d_counts,
                                hipStream_t stream, const float4* d_cull_boxes = nullptr, uint32_t n_cull = 0);
hipError_t launch_render_batch_seeded(const lrtd::BatchItem* d_items, const uint64_t* d_offsets, uint32_t n_items, uint64_t total,
                                      const lrtd::TraceParams& P, bool indexed, const void* d_packed, float seed_logit, uint32_t chunk,
                                      uint32_t* d_work_counter, uint32_t* d_spill, uint32_t blocks, unsigned long long* d_counts,
                                      hipStream_t stream);
}  // namespace lrt

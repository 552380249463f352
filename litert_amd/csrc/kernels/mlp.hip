// mlp.hip - kernels of the learned root-find seeder (no reference
// counterpart; the seeded render keeps the reference traversal order and
// hit test, BVH2Common.cpp:3407-3488 and 764-984):
//   k_brick_features  per-brick corner / octant-minimum table
//   k_mlp_collect     training events from the plain traversal
//   k_mlp_infer       BF16 MFMA inference over a feature batch
//   k_mlp_grad        FP32 forward + backward on the matrix cores, per-block partial gradients
//   k_mlp_reduce_adam deterministic gradient reduction + Adam
//   k_render_seeded   primary rays with the network consulted at every brick
#include "mlp.hpp"
#include "mlp_device.hpp"
#include "render_device.hpp"
#include "traverse_device.hpp"

namespace lrtd {

// LDS writes of one lane become visible to the other lanes of its wave.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// ---------------------------------------------------------------- features --
__global__ __launch_bounds__(256) void k_brick_features(DevScene S, float4* __restrict__ out) {
  const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= S.n_nodes) return;
  const uint4 N = S.nodes[b];
  const float sz = (float)(N.y & 0xFFFFu);
  const float sz_inv = 2.0f / sz;
  const float s = 0.5f * sz;
  const uint32_t bs = S.brick_size;
  float corner[8], omin[8], vals[8];
  for (int i = 0; i < 8; i++) { corner[i] = 0.0f; omin[i] = 1e30f; }
  float all_min = 1e30f;
  for (uint32_t vx = 0; vx < bs; vx++)
    for (uint32_t vy = 0; vy < bs; vy++)
      for (uint32_t vz = 0; vz < bs; vz++) {
        const float vmin = load_voxel(S, N.z, vx, vy, vz, sz_inv, vals);
        const int o = ((2 * vx >= bs) ? 4 : 0) | ((2 * vy >= bs) ? 2 : 0) | ((2 * vz >= bs) ? 1 : 0);
        omin[o] = mn(omin[o], vmin);
        all_min = mn(all_min, vmin);
        for (int c = 0; c < 8; c++) {
          const uint32_t cx = (c >> 2) & 1, cy = (c >> 1) & 1, cz = c & 1;
          if (vx == cx * (bs - 1) && vy == cy * (bs - 1) && vz == cz * (bs - 1)) corner[c] = vals[c];
        }
      }
  for (int o = 0; o < 8; o++)
    if (omin[o] > 1e29f) omin[o] = all_min;  // octant without voxels (brick_size 1)
  out[4 * (size_t)b + 0] = make_float4(corner[0] * s, corner[1] * s, corner[2] * s, corner[3] * s);
  out[4 * (size_t)b + 1] = make_float4(corner[4] * s, corner[5] * s, corner[6] * s, corner[7] * s);
  out[4 * (size_t)b + 2] = make_float4(omin[0] * s, omin[1] * s, omin[2] * s, omin[3] * s);
  out[4 * (size_t)b + 3] = make_float4(omin[4] * s, omin[5] * s, omin[6] * s, omin[7] * s);
}

// ------------------------------------------------------------- collection --
// One lane per packed pixel (spp 1, the preset's first sample offset). Every
// brick the plain traversal enters yields one event: its features, whether
// the exact intersector hit inside it and where (brick units from the brick
// entry), and the half-voxel margin used by the interval targets.
__global__ __launch_bounds__(256) void k_mlp_collect(DevScene S, TraceParams P, const float4* __restrict__ bfeat,
                                                     const Camera* __restrict__ cams, RenderArgs A, float* __restrict__ feat,
                                                     float* __restrict__ tgt, MlpEventMeta* __restrict__ meta, uint64_t cap,
                                                     unsigned long long* count, uint32_t* work_counter, uint32_t* spill) {
  __shared__ uint32_t lds_stack[kStackLds * 256];
  __shared__ uint32_t chunk_base;
  Stack st;
  st.lds = (lds_u32*)(lds_stack + threadIdx.x);
  st.lds_stride = 256;
  st.lds_entries = kStackLds;
  st.spill = spill + blockIdx.x * 256 + threadIdx.x;
  st.spill_stride = gridDim.x * 256;
  const uint64_t total = (uint64_t)A.n_views * A.n_packed;
  const float margin = 0.5f / (float)S.brick_size;
  for (;;) {
    if (threadIdx.x == 0) chunk_base = atomicAdd(work_counter, 1u);
    __syncthreads();
    const uint64_t base = (uint64_t)chunk_base * 256;
    __syncthreads();
    if (base >= total) break;
    const uint64_t gtid = base + threadIdx.x;
    if (gtid >= total) continue;
    const uint32_t view = (uint32_t)(gtid / A.n_packed);
    const uint32_t tid = (uint32_t)(gtid - (uint64_t)view * A.n_packed);
    uint32_t x, y;
    tid_to_xy(tid, A.packed_w, x, y);
    if (x >= A.W || y >= A.H) continue;
    const float2 d = sample_offset(A, x, y, 0);
    v3 pos, dir;
    eye_ray(cams[view], x, y, d.x, d.y, A.W, A.H, pos, dir);
    const v3 inv = SafeInverse(dir);
    Hit hit;
    hit.t = 1e9f;
    hit.primId = hit.instId = hit.geomId = 0xFFFFFFFFu;
    hit.coords[0] = 1.0f; hit.coords[1] = 0.0f; hit.coords[2] = 0.0f; hit.coords[3] = 0.0f;
    const float2 bh = RayBoxW(wave_ref_slabs(inv), pos, inv, v3(-1, -1, -1), v3(1, 1, 1));
    if (!((bh.x <= bh.y) && (bh.y > 0.0f) && (bh.x < hit.t))) continue;
    int top = 0;
    uint32_t left = 0;
    const uint32_t ray_id = view * A.W * A.H + y * A.W + x;
    while (NextLeaf<false>(S, pos, inv, 0.0f, false, hit.t, top, left, st, nullptr)) {
      const uint32_t brick = left & LRT_START_MASK;
      float f[24], t0, s;
      mlp_features(S, bfeat, pos, dir, inv, brick, f, &t0, &s);
      const float tPrev = hit.t;
      BrickIntersect(S, P, pos, dir, inv, 0.1f, brick, &hit, nullptr);  // tNear for SDF leaves = max(0, 0.1)
      const bool hb = hit.t < tPrev;
      const unsigned long long idx = atomicAdd(count, 1ull);
      if (idx < cap) {
        float4* fo = (float4*)(feat + idx * 24);
#pragma unroll
        for (int q = 0; q < 6; q++) fo[q] = make_float4(f[4 * q], f[4 * q + 1], f[4 * q + 2], f[4 * q + 3]);
        tgt[idx * 3 + 0] = hb ? (hit.t - t0) * s : 0.0f;
        tgt[idx * 3 + 1] = hb ? 1.0f : 0.0f;
        tgt[idx * 3 + 2] = margin;
        meta[idx].ray = ray_id;
        meta[idx].brick = brick;
      }
      if (hb) break;
      top--;
      left = st.get(top > 0 ? top : 0);
    }
  }
}

// -------------------------------------------------------------- inference --
__global__ __launch_bounds__(256) void k_mlp_infer(const void* __restrict__ packed, const float* __restrict__ feat, uint64_t n,
                                                   float* __restrict__ out) {
  __shared__ __align__(16) uint16_t fl[4][64 * kFeatStride];
  __shared__ float ol[4][64 * 3];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (uint64_t base = ((uint64_t)blockIdx.x * 4 + wave) * 64; base < n; base += (uint64_t)gridDim.x * 4 * 64) {
    const uint64_t s = base + lane;
    float f[24];
    if (s < n) {
      const float4* fi = (const float4*)(feat + s * 24);
#pragma unroll
      for (int q = 0; q < 6; q++) {
        const float4 v = fi[q];
        f[4 * q] = v.x; f[4 * q + 1] = v.y; f[4 * q + 2] = v.z; f[4 * q + 3] = v.w;
      }
    } else {
#pragma unroll
      for (int k = 0; k < 24; k++) f[k] = 0.0f;
    }
    store_features_bf16(fl[wave] + lane * kFeatStride, f);  // zero rows past n
    wave_lds_sync();
    mlp_wave_forward64(packed, fl[wave], ol[wave], lane);
    wave_lds_sync();
    if (s < n) {
      out[s * 3 + 0] = ol[wave][3 * lane + 0];
      out[s * 3 + 1] = ol[wave][3 * lane + 1];
      out[s * 3 + 2] = ol[wave][3 * lane + 2];
    }
    wave_lds_sync();
  }
}

// --------------------------------------------------------------- training --
// FP32 forward + backward of one 64-sample minibatch slice per 256-lane block
// on the FP32 matrix cores (v_mfma_f32_16x16x4_f32, FP32 operands and
// accumulation). Every layer is a GEMM over the block's 64 samples: wave w
// owns units 16w .. 16w + 15 of each hidden layer, so the forward output
// rows, the weight-gradient rows and the back-propagated rows of a wave are
// all its own 16 units, while the other operand (activations, dLoss/dz) is
// read from LDS, where every layer's tiles are [unit][sample] rows of kGS
// floats. The block writes its partial gradient (summed over its samples)
// and loss, and k_mlp_reduce_adam sums the partials in block order.
//
// 16x16x4 operand layout: lane l supplies A[l % 16][l / 16] and
// B[l / 16][l % 16], and holds D[4 (l / 16) + r][l % 16], r = 0..3.
constexpr int kGS = 65;  // LDS row stride: the MFMA operand walks touch 16 distinct banks

// SiLU and its derivative with the hardware exp2 and reciprocal (about 1 ulp
// each; the gradient is held to 1e-4 of the FP64 oracle, not bit for bit).
__device__ __forceinline__ float sigmoid_hw(float z) {
  return __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(-1.44269504f * z));
}
__device__ __forceinline__ float silu_acc(float z) { return z * sigmoid_hw(z); }
__device__ __forceinline__ float dsilu_acc(float z) {
  const float sg = sigmoid_hw(z);
  return sg * (1.0f + z * (1.0f - sg));
}

__device__ __forceinline__ f32x4 mfma_f32(float a, float b, f32x4 c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }

// sample slot s of minibatch slice `blk`: its index into the dataset, or -1
__device__ __forceinline__ int64_t train_sample(const MlpTrainArgs& T, uint32_t blk, int s) {
  const uint64_t local = (uint64_t)blk * 64 + s;
  return local < T.batch ? (int64_t)((T.perm_a * (T.first + local) + T.perm_b) % T.n) : -1;
}

// One hidden layer's forward GEMM for this wave's 16 units: z = b + W x over
// the 64 samples, K = 4 * ksteps inputs; w[ks] = W[16w + l%16][4 ks + l/16].
// Writes z and silu(z) to the [unit][sample] tiles.
template <int KSTEPS>
__device__ __forceinline__ void fwd_layer(const float* __restrict__ w, const float bias[4], const float* x_lds, float* z_lds, float* a_lds,
                                          int o0, int lr, int lg) {
  f32x4 acc[4];
#pragma unroll
  for (int nt = 0; nt < 4; nt++)
#pragma unroll
    for (int r = 0; r < 4; r++) acc[nt][r] = bias[r];
#pragma unroll
  for (int ks = 0; ks < KSTEPS; ks++) {
#pragma unroll
    for (int nt = 0; nt < 4; nt++) acc[nt] = mfma_f32(w[ks], x_lds[(4 * ks + lg) * kGS + 16 * nt + lr], acc[nt]);
  }
#pragma unroll
  for (int nt = 0; nt < 4; nt++)
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int u = o0 + 4 * lg + r, sl = 16 * nt + lr;
      z_lds[u * kGS + sl] = acc[nt][r];
      a_lds[u * kGS + sl] = silu_acc(acc[nt][r]);
    }
}

// Weight gradient rows of this wave's 16 output units, summed over the 64
// samples: gW[o][i] = sum_s dz[o][s] a[i][s] for i < n_in (n_in <= 16 * NT),
// and the bias gradient gB[o] = sum_s dz[o][s] from the same A operands
// (lane l sums samples 4 ks + l/16; the 4 lane groups are added in order).
template <int NT>
__device__ __forceinline__ void grad_rows(const float* dz_lds, const float* a_lds, int n_in, float* __restrict__ gW,
                                          float* __restrict__ gB, int o0, int lr, int lg) {
  f32x4 acc[NT];
#pragma unroll
  for (int nt = 0; nt < NT; nt++) acc[nt] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
  float bsum = 0.0f;
#pragma unroll
  for (int ks = 0; ks < 16; ks++) {
    const float a = dz_lds[(o0 + lr) * kGS + 4 * ks + lg];
    bsum += a;
#pragma unroll
    for (int nt = 0; nt < NT; nt++) {
      const int i = 16 * nt + lr;
      const float b = i < n_in ? a_lds[i * kGS + 4 * ks + lg] : 0.0f;
      acc[nt] = mfma_f32(a, b, acc[nt]);
    }
  }
#pragma unroll
  for (int nt = 0; nt < NT; nt++) {
    const int i = 16 * nt + lr;
    if (i < n_in)
#pragma unroll
      for (int r = 0; r < 4; r++) gW[(o0 + 4 * lg + r) * n_in + i] = acc[nt][r];
  }
  const float b1 = __shfl(bsum, lr + 16, 64), b2 = __shfl(bsum, lr + 32, 64), b3 = __shfl(bsum, lr + 48, 64);
  if (lg == 0) gB[o0 + lr] = ((bsum + b1) + b2) + b3;
}

// dLoss/dz of the previous layer for this wave's 16 input units,
// dz_prev[i][s] = (sum_o W[o][i] dz[o][s]) * silu'(z_prev[i][s]) with
// wt[ks] = W[4 ks + l/16][16w + l%16], together with the weight-gradient rows
// (gW = dz a^T, gB = dz 1) of this wave's 16 output units: both GEMMs read the
// same dz tile and are independent, so one loop keeps 8 accumulators in flight.
__device__ __forceinline__ void grad_back_rows(const float* dz_lds, const float* a_lds, float* __restrict__ gW, float* __restrict__ gB,
                                               const float* __restrict__ wt, const float* zprev_lds, float* out_lds, int o0, int lr,
                                               int lg) {
  f32x4 ag[4], ab[4];
#pragma unroll
  for (int nt = 0; nt < 4; nt++) ag[nt] = ab[nt] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
  float bsum = 0.0f;
#pragma unroll
  for (int ks = 0; ks < 16; ks++) {
    const float a = dz_lds[(o0 + lr) * kGS + 4 * ks + lg];
    bsum += a;
#pragma unroll
    for (int nt = 0; nt < 4; nt++) {
      ag[nt] = mfma_f32(a, a_lds[(16 * nt + lr) * kGS + 4 * ks + lg], ag[nt]);
      ab[nt] = mfma_f32(wt[ks], dz_lds[(4 * ks + lg) * kGS + 16 * nt + lr], ab[nt]);
    }
  }
#pragma unroll
  for (int nt = 0; nt < 4; nt++)
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int u = o0 + 4 * lg + r, sl = 16 * nt + lr;
      gW[u * kMlpH + 16 * nt + lr] = ag[nt][r];
      out_lds[u * kGS + sl] = ab[nt][r] * dsilu_acc(zprev_lds[u * kGS + sl]);
    }
  const float b1 = __shfl(bsum, lr + 16, 64), b2 = __shfl(bsum, lr + 32, 64), b3 = __shfl(bsum, lr + 48, 64);
  if (lg == 0) gB[o0 + lr] = ((bsum + b1) + b2) + b3;
}

__global__ __launch_bounds__(256) void k_mlp_grad(const float* __restrict__ prm, MlpTrainArgs T, float* __restrict__ partial,
                                                  float* __restrict__ partial_loss) {
  __shared__ float sX[kMlpIn * kGS];
  __shared__ float sZ[3][kMlpH * kGS];  // z1, z2, z3
  __shared__ float sA[3][kMlpH * kGS];  // silu(z1), silu(z2), silu(z3)
  __shared__ float sD[2][kMlpH * kGS];  // dLoss/dz of the layer being back-propagated (two, alternating)
  __shared__ float sD4[kMlpOut * 64], sLoss[64], sT[3 * 64];
  __shared__ int64_t sIdx[64];  // dataset index of each sample slot, -1 past the minibatch
  const int tid = threadIdx.x, wv = tid >> 6, lane = tid & 63;
  const int lr = lane & 15, lg = lane >> 4, o0 = 16 * wv;
  float* gp = partial + (size_t)blockIdx.x * kMlpPartialStride;

  // everything this wave reads from global memory, loaded up front: weight
  // rows (forward) and columns (backward), biases, and the samples' inputs
  // and targets
  float w1[6], w2[16], w3[16], w2t[16], w3t[16], w4[16], bias[3][4], b4[4];
#pragma unroll
  for (int ks = 0; ks < 16; ks++) w4[ks] = lr < kMlpOut ? prm[kOffW4 + lr * kMlpH + 4 * ks + lg] : 0.0f;
  const float w4t = lg < kMlpOut ? prm[kOffW4 + lg * kMlpH + o0 + lr] : 0.0f;
#pragma unroll
  for (int r = 0; r < 4; r++) {
    bias[0][r] = prm[kOffB1 + o0 + 4 * lg + r];
    bias[1][r] = prm[kOffB2 + o0 + 4 * lg + r];
    bias[2][r] = prm[kOffB3 + o0 + 4 * lg + r];
    b4[r] = 4 * lg + r < kMlpOut ? prm[kOffB4 + 4 * lg + r] : 0.0f;
  }
#pragma unroll
  for (int ks = 0; ks < 6; ks++) w1[ks] = prm[kOffW1 + (o0 + lr) * kMlpIn + 4 * ks + lg];
#pragma unroll
  for (int ks = 0; ks < 16; ks++) {
    w2[ks] = prm[kOffW2 + (o0 + lr) * kMlpH + 4 * ks + lg];
    w3[ks] = prm[kOffW3 + (o0 + lr) * kMlpH + 4 * ks + lg];
    w2t[ks] = prm[kOffW2 + (4 * ks + lg) * kMlpH + o0 + lr];
    w3t[ks] = prm[kOffW3 + (4 * ks + lg) * kMlpH + o0 + lr];
  }
  if (tid < 64) sIdx[tid] = train_sample(T, blockIdx.x, tid);
  __syncthreads();
  for (int e = tid; e < 64 * kMlpIn; e += 256) {
    const int sl = e / kMlpIn, k = e - sl * kMlpIn;
    const int64_t idx = sIdx[sl];
    sX[k * kGS + sl] = idx >= 0 ? T.feats[idx * kMlpIn + k] : 0.0f;
  }
  if (tid < 64 * 3) {
    const int sl = tid / 3, c = tid - 3 * sl;
    const int64_t idx = sIdx[sl];
    sT[c * 64 + sl] = idx >= 0 ? T.tgts[idx * 3 + c] : 0.0f;
  }
  __syncthreads();
  // forward, hidden layers
  fwd_layer<6>(w1, bias[0], sX, sZ[0], sA[0], o0, lr, lg);
  __syncthreads();
  fwd_layer<16>(w2, bias[1], sA[0], sZ[1], sA[1], o0, lr, lg);
  __syncthreads();
  fwd_layer<16>(w3, bias[2], sA[1], sZ[2], sA[2], o0, lr, lg);
  __syncthreads();
  // Output layer (3 units, padded to a 16-row tile) for this wave's 16
  // samples 16w .. 16w + 15: A = W4 rows 0..2, B = a3 of those samples. Lanes
  // of group 0 end up holding z4[0..2] of sample 16w + l%16 and evaluate the
  // loss there: BCE(sigmoid(z2), hit) + hit ((z0 - (t - m))^2 + (z1 - (t + m))^2).
  {
    f32x4 z = {b4[0], b4[1], b4[2], b4[3]};
#pragma unroll
    for (int ks = 0; ks < 16; ks++) z = mfma_f32(w4[ks], sA[2][(4 * ks + lg) * kGS + o0 + lr], z);
    if (lg == 0) {
      const int sl = o0 + lr;
      const int64_t idx = sIdx[sl];
      float dz0 = 0.0f, dz1 = 0.0f, dz2 = 0.0f, loss = 0.0f;
      if (idx >= 0) {
        const float th = sT[sl], h = sT[64 + sl] > 0.5f ? 1.0f : 0.0f, m = sT[128 + sl];
        const float z2 = z[2];
        loss = fmaxf(z2, 0.0f) - z2 * h + log1pf(expf(-fabsf(z2)));
        dz2 = 1.0f / (1.0f + expf(-z2)) - h;
        const float e0 = z[0] - (th - m), e1 = z[1] - (th + m);
        loss += h * (e0 * e0 + e1 * e1);
        dz0 = h * 2.0f * e0;
        dz1 = h * 2.0f * e1;
      }
      sD4[sl] = dz0;
      sD4[64 + sl] = dz1;
      sD4[128 + sl] = dz2;
      sLoss[sl] = loss;
    }
  }
  __syncthreads();
  if (tid == 255) {
    float acc = 0.0f;
    for (int sl = 0; sl < 64; sl++) acc += sLoss[sl];
    partial_loss[blockIdx.x] = acc;
  }
  // output-layer gradient, columns 16w .. 16w + 15 of W4 (rows 0..2), summed
  // over the 64 samples; wave 0 also sums the bias gradient
  {
    f32x4 g = {0.0f, 0.0f, 0.0f, 0.0f};
    float bsum = 0.0f;
#pragma unroll
    for (int ks = 0; ks < 16; ks++) {
      const float a = lr < kMlpOut ? sD4[lr * 64 + 4 * ks + lg] : 0.0f;
      bsum += a;
      g = mfma_f32(a, sA[2][(o0 + lr) * kGS + 4 * ks + lg], g);
    }
#pragma unroll
    for (int r = 0; r < 4; r++)
      if (4 * lg + r < kMlpOut) gp[kOffW4 + (4 * lg + r) * kMlpH + o0 + lr] = g[r];
    const float b1 = __shfl(bsum, lr + 16, 64), b2 = __shfl(bsum, lr + 32, 64), b3 = __shfl(bsum, lr + 48, 64);
    if (wv == 0 && lg == 0 && lr < kMlpOut) gp[kOffB4 + lr] = ((bsum + b1) + b2) + b3;
  }
  // dLoss/dz3 for this wave's 16 units: (W4^T dz4) * silu'(z3), K = 3 padded to 4
  {
#pragma unroll
    for (int nt = 0; nt < 4; nt++) {
      const float b = lg < kMlpOut ? sD4[lg * 64 + 16 * nt + lr] : 0.0f;
      const f32x4 d = mfma_f32(w4t, b, f32x4{0.0f, 0.0f, 0.0f, 0.0f});
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int u = o0 + 4 * lg + r, sl = 16 * nt + lr;
        sD[0][u * kGS + sl] = d[r] * dsilu_acc(sZ[2][u * kGS + sl]);
      }
    }
  }
  __syncthreads();
  // layer 3: gradient from dz3 (sD[0]) and a2, dz2 into sD[1]
  grad_back_rows(sD[0], sA[1], gp + kOffW3, gp + kOffB3, w3t, sZ[1], sD[1], o0, lr, lg);
  __syncthreads();
  // layer 2: gradient from dz2 (sD[1]) and a1, dz1 into sD[0]
  grad_back_rows(sD[1], sA[0], gp + kOffW2, gp + kOffB2, w2t, sZ[0], sD[0], o0, lr, lg);
  __syncthreads();
  // layer 1: gradient from dz1 (sD[0]) and the inputs
  grad_rows<2>(sD[0], sX, kMlpIn, gp + kOffW1, gp + kOffB1, o0, lr, lg);
}

// Gradient reduction over the blocks' partials + Adam. A block takes 64
// parameters: lane (q, j) sums parameters 4q .. 4q + 3 of the block over the
// partials b = j mod 16 with 16-B loads (rows are kMlpPartialStride floats,
// 16-B aligned), then each parameter's 16 sums are added in j order -- a
// fixed order, so the result is deterministic.
constexpr int kRedParams = 64, kRedLanes = 16;
__global__ __launch_bounds__(256) void k_mlp_reduce_adam(float* __restrict__ prm, float* __restrict__ m, float* __restrict__ v,
                                                         const float* __restrict__ partial, const float* __restrict__ partial_loss,
                                                         uint32_t parts, uint32_t batch, AdamArgs adam, int apply,
                                                         float* __restrict__ grad_out, float* __restrict__ loss_out) {
  __shared__ float red[kRedLanes][kRedParams];
  __shared__ float lred[256];
  const int q = threadIdx.x % (kRedParams / 4), j = threadIdx.x / (kRedParams / 4);
  const uint32_t p0 = blockIdx.x * kRedParams + 4 * q;
  const float inv_b = 1.0f / (float)batch;
  if (blockIdx.x == 0 && loss_out) {  // loss: lane t sums partials t mod 256, then lane 0 adds the 256 sums in order
    float acc = 0.0f;
    for (uint32_t b = threadIdx.x; b < parts; b += 256) acc += partial_loss[b];
    lred[threadIdx.x] = acc;
  }
  float4 g = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  if (p0 < (uint32_t)kMlpParams)  // p0 + 3 < kMlpPartialStride: the padding is never summed into a parameter
    for (uint32_t b = j; b < parts; b += kRedLanes) {
      const float4 v4 = *(const float4*)(partial + (size_t)b * kMlpPartialStride + p0);
      g.x += v4.x; g.y += v4.y; g.z += v4.z; g.w += v4.w;
    }
  red[j][4 * q + 0] = g.x;
  red[j][4 * q + 1] = g.y;
  red[j][4 * q + 2] = g.z;
  red[j][4 * q + 3] = g.w;
  __syncthreads();
  if (blockIdx.x == 0 && loss_out && threadIdx.x == 0) {
    float acc = 0.0f;
    for (int t = 0; t < 256; t++) acc += lred[t];
    *loss_out = acc * inv_b;
  }
  const uint32_t p = blockIdx.x * kRedParams + threadIdx.x;
  if (threadIdx.x >= kRedParams || p >= (uint32_t)kMlpParams) return;
  float gs = 0.0f;
  for (int k = 0; k < kRedLanes; k++) gs += red[k][threadIdx.x];
  gs *= inv_b;
  if (grad_out) grad_out[p] = gs;
  if (apply) {
    const float mm = adam.b1 * m[p] + (1.0f - adam.b1) * gs;
    const float vv = adam.b2 * v[p] + (1.0f - adam.b2) * gs * gs;
    m[p] = mm;
    v[p] = vv;
    prm[p] -= adam.lr * (mm * adam.c1) / (sqrtf(vv * adam.c2) + adam.eps);
  }
}

// ----------------------------------------------------------- seeded render --
// Same pixels, camera rays, traversal order and hit test as k_render; at
// every brick entry the wave stops, packs the features of its lanes that sit
// at a leaf into LDS (compacted, so the MFMA work scales with those lanes),
// runs the network, and each such lane hands the predicted hit distance to
// the exact brick intersector when the hit logit clears the confidence gate.
// One pixel per lane, all 64 lanes of a wave together (the network runs on
// the wave's compacted leaf lanes, so lanes outside the image take part with
// valid = false). fl/ol are this wave's LDS feature and output rows.
template <bool COUNT, bool RGBA, int RF, bool IDX>
__device__ __forceinline__ void seeded_pixel(const DevScene& S, const TraceParams& P, const float4* __restrict__ bfeat,
                                             const void* __restrict__ packed, float seed_logit, const Camera& cam, const RenderArgs& A,
                                             const RenderOut& O, uint32_t view, uint32_t x, uint32_t y, bool valid, Stack& st,
                                             WorkCount& wc, uint16_t* fl, float* ol, int lane) {
  float4 acc = make_float4(0, 0, 0, 0);
  for (uint32_t i = 0; i < A.spp; i++) {
    v3 pos(0, 0, 0), dir(0, 0, 1);
    if (valid) {
      const float2 d = sample_offset(A, x, y, i);
      eye_ray(cam, x, y, d.x, d.y, A.W, A.H, pos, dir);
    }
    const v3 inv = SafeInverse(dir);
    Hit hit;
    hit.t = 1e9f;
    hit.primId = hit.instId = hit.geomId = 0xFFFFFFFFu;
    hit.coords[0] = 1.0f; hit.coords[1] = 0.0f; hit.coords[2] = 0.0f; hit.coords[3] = 0.0f;
    bool active = false;
    if (valid) {
      const float2 bh = RayBoxW(wave_ref_slabs(inv), pos, inv, v3(-1, -1, -1), v3(1, 1, 1));
      active = (bh.x <= bh.y) && (bh.y > 0.0f) && (bh.x < hit.t);
    }
    int top = 0;
    uint32_t left = 0;
    for (;;) {
      bool at_leaf = false;
      if (active) {
        at_leaf = NextLeaf<COUNT>(S, pos, inv, 0.0f, false, hit.t, top, left, st, &wc);
        active = at_leaf;
      }
      const uint64_t leaf_mask = __ballot(at_leaf);
      if (leaf_mask == 0ull) break;  // wave-uniform
      const int n_at = __popcll(leaf_mask);
      const int slot = __popcll(leaf_mask & ((1ull << lane) - 1ull));
      const uint32_t brick = left & LRT_START_MASK;
      float t0 = 0.0f, s = 1.0f;
      if (at_leaf) {
        float f[24];
        mlp_features(S, bfeat, pos, dir, inv, brick, f, &t0, &s);
        store_features_bf16(fl + slot * kFeatStride, f);
      }
      wave_lds_sync();
      mlp_wave_forward(packed, fl, ol, (n_at + 15) >> 4, lane);
      if (COUNT && lane == 0) wc.tiles += (uint32_t)((n_at + 15) >> 4);
      wave_lds_sync();
      if (at_leaf) {
        const float z0 = ol[3 * slot + 0], z1 = ol[3 * slot + 1], z2 = ol[3 * slot + 2];
        // confident hit: centre of the predicted interval; confident miss:
        // beyond every voxel of the brick; otherwise no prediction
        const bool confident = z2 >= seed_logit;
        const float seed_world = confident ? t0 + 0.5f * (z0 + z1) / s : (z2 <= -seed_logit ? 3.0e38f : -1.0f);
        if (COUNT) wc.bricks++, wc.seeds += confident ? 1u : 0u;
        const float tPrev = hit.t;
        BrickIntersect<true, RF, IDX>(S, P, pos, dir, inv, 0.1f, brick, &hit, COUNT ? &wc.voxels : nullptr, seed_world,
                             COUNT ? &wc.iters : nullptr);
        if (hit.t < tPrev) {
          active = false;
        } else {
          top--;
          left = st.get(top > 0 ? top : 0);
        }
      }
      wave_lds_sync();  // the next round rewrites this wave's LDS rows
    }
    if (hit.primId != 0xFFFFFFFFu) hit.instId = 0;
    if (valid) {
      const size_t pix = (size_t)view * A.W * A.H + (size_t)y * A.W + x;
      if (i == 0) {
        if (O.t) O.t[pix] = hit.t;
        if (O.prim) O.prim[pix] = hit.primId;
        if (O.n) O.n[pix] = make_float2(hit.coords[2], hit.coords[3]);
      }
      if (RGBA) {
        uint32_t lit = 0xFFFFFFFFu;
        if (needs_shadow(P.render_mode) && hit.primId != 0xFFFFFFFFu) {  // shadow rays: plain exact traversal
          for (uint32_t li = 0; li < A.lights.n; li++) {
            v3 o, ld;
            if (!shadow_ray(A.lights, li, pos, dir, hit.t, o, ld)) continue;
            const Hit sh = RayQuery_NearestHit<COUNT, RF, IDX>(S, P, o, 0.0f, ld, 1e9f, st, &wc);
            if (sh.primId != 0xFFFFFFFFu) lit &= ~(1u << li);
          }
        }
        const float4 c = shade(P.render_mode, hit, pos, dir, A.lights, lit);
        acc = make_float4(acc.x + c.x, acc.y + c.y, acc.z + c.z, acc.w + c.w);
      }
    }
  }
  if (RGBA && valid) {
    const size_t pix = (size_t)view * A.W * A.H + (size_t)y * A.W + x;
    const float sp = (float)A.spp;
    O.rgba[pix] = make_float4(acc.x / sp, acc.y / sp, acc.z / sp, acc.w / sp);
  }
}

__device__ __forceinline__ void flush_counts(const WorkCount& wc, unsigned long long* counts) {
  atomicAdd(&counts[0], (unsigned long long)wc.pairs);
  atomicAdd(&counts[1], (unsigned long long)wc.bricks);
  atomicAdd(&counts[2], (unsigned long long)wc.voxels);
  atomicAdd(&counts[3], (unsigned long long)wc.iters);
  atomicAdd(&counts[4], (unsigned long long)wc.tiles);
  atomicAdd(&counts[5], (unsigned long long)wc.seeds);
  atomicAdd(&counts[6], (unsigned long long)wc.culled);
}

template <bool COUNT, bool RGBA, int RF, bool IDX>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3, 8))) void k_render_seeded(DevScene S, TraceParams P, const float4* __restrict__ bfeat,
                                                       const void* __restrict__ packed, float seed_logit, const Camera* __restrict__ cams,
                                                       RenderArgs A, RenderOut O, uint32_t* work_counter, uint32_t* spill,
                                                       unsigned long long* counts) {
  __shared__ uint32_t lds_stack[kStackLdsSeeded * 256];
  __shared__ __align__(16) uint16_t fl[4][64 * kFeatStride];
  __shared__ float ol[4][64 * 3];
  __shared__ uint32_t chunk_base;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  Stack st;
  st.lds = (lds_u32*)(lds_stack + threadIdx.x);
  st.lds_stride = 256;
  st.lds_entries = kStackLdsSeeded;
  st.spill = spill + blockIdx.x * 256 + threadIdx.x;
  st.spill_stride = gridDim.x * 256;
  WorkCount wc;
  const uint64_t total = (uint64_t)A.n_views * A.n_packed;
  for (;;) {
    if (threadIdx.x == 0) chunk_base = atomicAdd(work_counter, 1u);
    __syncthreads();
    const uint64_t base = (uint64_t)chunk_base * 256;
    __syncthreads();
    if (base >= total) break;  // uniform across the block
    const uint64_t gtid = base + threadIdx.x;
    uint32_t view = 0, x = 0, y = 0;
    bool valid = gtid < total;
    if (valid) {
      view = (uint32_t)(gtid / A.n_packed);
      const uint32_t tid = (uint32_t)(gtid - (uint64_t)view * A.n_packed);
      tid_to_xy(tid, A.packed_w, x, y);
      valid = x < A.W && y < A.H;
    }
    // A tile the cull pass marked (every ray provably misses every brick)
    // is written as tracing writes a miss. The wave's 64 packed pixels are
    // one 8x8 super-block (chunks start on multiples of 256), so the test is
    // wave-uniform; waves past the launch's end (wsb >= its tile count) read
    // no mask word.
    const uint32_t wsb = __builtin_amdgcn_readfirstlane((uint32_t)(gtid >> 6));
    if (A.tile_mask && (uint64_t)wsb * 64 < total && ((A.tile_mask[wsb >> 5] >> (wsb & 31u)) & 1u)) {
      if (valid) {
        const size_t pix = (size_t)view * A.W * A.H + (size_t)y * A.W + x;
        if (O.t) O.t[pix] = 1e9f;
        if (O.prim) O.prim[pix] = 0xFFFFFFFFu;
        if (O.n) O.n[pix] = make_float2(0.0f, 0.0f);
        if (RGBA) O.rgba[pix] = make_float4(0.0f, 0.0f, 0.0f, 1.0f);
        if (COUNT && lane == 0) wc.culled++;
      }
      continue;
    }
    seeded_pixel<COUNT, RGBA, RF, IDX>(S, P, bfeat, packed, seed_logit, cams[view], A, O, view, x, y, valid, st, wc, fl[wave], ol[wave],
                                       lane);
  }
  if (COUNT) flush_counts(wc, counts);
}

// Batched seeded render: the item layout and XCD-sliced chunk counters of
// k_render_batch. Item ranges are multiples of 64 packed pixels and chunks
// multiples of 256, so each wave lies in one item and either covers 64 real
// packed pixels or lies wholly past the end.
template <bool COUNT, int RF, bool IDX>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3, 8))) void k_render_batch_seeded(
    const BatchItem* __restrict__ items, const uint64_t* __restrict__ offsets, uint32_t n_items, uint64_t total, TraceParams P,
    const void* __restrict__ packed, float seed_logit, uint32_t chunk, uint32_t* work_counter, uint32_t* spill,
    unsigned long long* counts) {
  __shared__ uint32_t lds_stack[kStackLdsSeeded * 256];
  __shared__ __align__(16) uint16_t fl[4][64 * kFeatStride];
  __shared__ float ol[4][64 * 3];
  __shared__ uint32_t chunk_base;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  Stack st;
  st.lds = (lds_u32*)(lds_stack + threadIdx.x);
  st.lds_stride = 256;
  st.lds_entries = kStackLdsSeeded;
  st.spill = spill + blockIdx.x * 256 + threadIdx.x;
  st.spill_stride = gridDim.x * 256;
  WorkCount wc;
  const uint32_t xcd = blockIdx.x & (kXcds - 1u);
  const uint64_t n_chunks = (total + chunk - 1) / chunk;
  const uint64_t per = (n_chunks + kXcds - 1) / kXcds;
  for (uint32_t k = 0; k < kXcds; k++) {
    const uint32_t sl = (xcd + k) & (kXcds - 1u);
    const uint64_t c0 = per * sl, c1 = c0 + per < n_chunks ? c0 + per : n_chunks;
    for (;;) {
      if (threadIdx.x == 0) chunk_base = atomicAdd(work_counter + kCounterStride * sl, 1u);
      __syncthreads();
      const uint64_t c = c0 + chunk_base;
      __syncthreads();
      if (c >= c1) break;  // uniform across the block
      for (uint32_t j = 0; j < chunk; j += 256) {
        const uint64_t gtid = c * chunk + j + threadIdx.x;
        if (gtid < total) {  // wave-uniform: total is a multiple of 64
          const uint64_t g0 = ((uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)(gtid >> 32)) << 32) |
                              (uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)gtid);
          uint32_t lo = 0, hi = n_items;  // offsets[lo] <= g0 < offsets[hi]
          while (hi - lo > 1u) {
            const uint32_t mid = (lo + hi) >> 1;
            if (offsets[mid] <= g0) lo = mid;
            else hi = mid;
          }
          const BatchItem& it = items[lo];
          uint32_t x = 0, y = 0;
          tid_to_xy((uint32_t)(gtid - offsets[lo]), it.A.packed_w, x, y);
          const bool valid = x < it.A.W && y < it.A.H;
          seeded_pixel<COUNT, false, RF, IDX>(it.S, P, it.bfeat, packed, seed_logit, it.cam, it.A, it.O, 0, x, y, valid, st, wc,
                                              fl[wave], ol[wave], lane);
        }
      }
    }
  }
  if (COUNT) flush_counts(wc, counts);
}

}  // namespace lrtd

namespace lrt {

hipError_t launch_brick_features(const lrtd::DevScene& S, float4* d_bfeat, hipStream_t stream) {
  hipLaunchKernelGGL(lrtd::k_brick_features, dim3((S.n_nodes + 255) / 256), dim3(256), 0, stream, S, d_bfeat);
  return hipGetLastError();
}

hipError_t launch_mlp_collect(const lrtd::DevScene& S, const lrtd::TraceParams& P, const float4* d_bfeat, const lrtd::Camera* d_cams,
                              const lrtd::RenderArgs& A, float* d_feat, float* d_tgt, lrtd::MlpEventMeta* d_meta, uint64_t cap,
                              unsigned long long* d_count, uint32_t* d_spill, uint32_t blocks, hipStream_t stream) {
  // d_count[1] (as uint32) serves as the chunk counter
  uint32_t* work_counter = (uint32_t*)(d_count + 1);
  hipError_t e = zero_u32(work_counter, 1, stream);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(lrtd::k_mlp_collect, dim3(blocks), dim3(256), 0, stream, S, P, d_bfeat, d_cams, A, d_feat, d_tgt, d_meta, cap,
                     d_count, work_counter, d_spill);
  return hipGetLastError();
}

hipError_t launch_mlp_infer(const void* d_packed, const float* d_feat, uint64_t n, float* d_out, uint32_t blocks, hipStream_t stream) {
  hipLaunchKernelGGL(lrtd::k_mlp_infer, dim3(blocks), dim3(256), 0, stream, d_packed, d_feat, n, d_out);
  return hipGetLastError();
}

hipError_t launch_mlp_grad(const float* d_params, const lrtd::MlpTrainArgs& T, float* d_partial, float* d_partial_loss, uint32_t blocks,
                           hipStream_t stream) {
  hipLaunchKernelGGL(lrtd::k_mlp_grad, dim3(blocks), dim3(256), 0, stream, d_params, T, d_partial, d_partial_loss);
  return hipGetLastError();
}

hipError_t launch_mlp_reduce_adam(float* d_params, float* d_m, float* d_v, const float* d_partial, const float* d_partial_loss,
                                  uint32_t parts, uint32_t batch, const lrtd::AdamArgs& adam, bool apply, float* d_grad_out,
                                  float* d_loss_out, hipStream_t stream) {
  hipLaunchKernelGGL(lrtd::k_mlp_reduce_adam, dim3((lrtd::kMlpParams + lrtd::kRedParams - 1) / lrtd::kRedParams), dim3(256), 0, stream, d_params,
                     d_m, d_v, d_partial,
                     d_partial_loss, parts, batch, adam, apply ? 1 : 0, d_grad_out, d_loss_out);
  return hipGetLastError();
}

// Same variant split as the plain render: root finder fixed at compile time
// for Newton / sphere tracing on DX-layout scenes, generic otherwise.
struct SeededLaunch {
  const lrtd::DevScene& S; const lrtd::TraceParams& P; const float4* bfeat; const void* packed; float logit; const lrtd::Camera* cams;
  const lrtd::RenderArgs& A; const lrtd::RenderOut& O; uint32_t* counter; uint32_t* spill; uint32_t blocks; unsigned long long* counts;
  hipStream_t stream;
  template <bool C, bool R, int RF, bool IDX>
  hipError_t go() {
    hipLaunchKernelGGL((lrtd::k_render_seeded<C, R, RF, IDX>), dim3(blocks), dim3(256), 0, stream, S, P, bfeat, packed, logit, cams, A, O,
                       counter, spill, counts);
    return hipGetLastError();
  }
  template <int RF, bool IDX>
  hipError_t pick(bool rgba) {
    if (counts) return rgba ? go<true, true, RF, IDX>() : go<true, false, RF, IDX>();
    return rgba ? go<false, true, RF, IDX>() : go<false, false, RF, IDX>();
  }
};

hipError_t launch_render_seeded(const lrtd::DevScene& S, const lrtd::TraceParams& P, const float4* d_bfeat, const void* d_packed,
                                float seed_logit, const lrtd::Camera* d_cams, const lrtd::RenderArgs& A, const lrtd::RenderOut& O,
                                uint32_t* d_work_counter, uint32_t* d_spill, uint32_t blocks, bool rgba, unsigned long long* d_counts,
                                hipStream_t stream, const float4* d_cull_boxes, uint32_t n_cull) {
  hipError_t e = zero_u32(d_work_counter, 1, stream);
  if (e != hipSuccess) return e;
  if (A.tile_mask && (e = launch_tile_cull(d_cams, A, d_cull_boxes, n_cull, stream)) != hipSuccess) return e;
  SeededLaunch L{S, P, d_bfeat, d_packed, seed_logit, d_cams, A, O, d_work_counter, d_spill, blocks, d_counts, stream};
  const bool indexed = S.layout == LRT_SBS_LAYOUT_ID32F_IRGB32F || S.layout == LRT_SBS_LAYOUT_ID32F_IRGB32F_IN;
  if (indexed) return L.pick<-1, true>(rgba);
  if (P.sdf_node_intersect == LRT_SDF_NODE_INTERSECT_NEWTON) return L.pick<(int)LRT_SDF_NODE_INTERSECT_NEWTON, false>(rgba);
  if (P.sdf_node_intersect == LRT_SDF_NODE_INTERSECT_ST) return L.pick<(int)LRT_SDF_NODE_INTERSECT_ST, false>(rgba);
  return L.pick<-1, false>(rgba);
}


template <bool C, int RF, bool IDX>
static hipError_t batch_seeded_go(const lrtd::BatchItem* items, const uint64_t* offsets, uint32_t n_items, uint64_t total,
                                  const lrtd::TraceParams& P, const void* packed, float logit, uint32_t chunk, uint32_t* counter,
                                  uint32_t* spill, uint32_t blocks, unsigned long long* counts, hipStream_t stream) {
  hipLaunchKernelGGL((lrtd::k_render_batch_seeded<C, RF, IDX>), dim3(blocks), dim3(256), 0, stream, items, offsets, n_items, total, P,
                     packed, logit, chunk, counter, spill, counts);
  return hipGetLastError();
}

template <int RF, bool IDX>
static hipError_t batch_seeded_pick(const lrtd::BatchItem* items, const uint64_t* offsets, uint32_t n_items, uint64_t total,
                                    const lrtd::TraceParams& P, const void* packed, float logit, uint32_t chunk, uint32_t* counter,
                                    uint32_t* spill, uint32_t blocks, unsigned long long* counts, hipStream_t stream) {
  return counts ? batch_seeded_go<true, RF, IDX>(items, offsets, n_items, total, P, packed, logit, chunk, counter, spill, blocks, counts, stream)
                : batch_seeded_go<false, RF, IDX>(items, offsets, n_items, total, P, packed, logit, chunk, counter, spill, blocks, counts, stream);
}

hipError_t launch_render_batch_seeded(const lrtd::BatchItem* d_items, const uint64_t* d_offsets, uint32_t n_items, uint64_t total,
                                      const lrtd::TraceParams& P, bool indexed, const void* d_packed, float seed_logit, uint32_t chunk,
                                      uint32_t* d_work_counter, uint32_t* d_spill, uint32_t blocks, unsigned long long* d_counts,
                                      hipStream_t stream) {
  if (chunk == 0 || chunk % 256 != 0) return hipErrorInvalidValue;
  hipError_t e = zero_u32(d_work_counter, lrtd::kXcds * lrtd::kCounterStride, stream);
  if (e != hipSuccess) return e;
  if (indexed)
    return batch_seeded_pick<-1, true>(d_items, d_offsets, n_items, total, P, d_packed, seed_logit, chunk, d_work_counter, d_spill, blocks,
                                       d_counts, stream);
  if (P.sdf_node_intersect == LRT_SDF_NODE_INTERSECT_NEWTON)
    return batch_seeded_pick<(int)LRT_SDF_NODE_INTERSECT_NEWTON, false>(d_items, d_offsets, n_items, total, P, d_packed, seed_logit, chunk,
                                                                        d_work_counter, d_spill, blocks, d_counts, stream);
  if (P.sdf_node_intersect == LRT_SDF_NODE_INTERSECT_ST)
    return batch_seeded_pick<(int)LRT_SDF_NODE_INTERSECT_ST, false>(d_items, d_offsets, n_items, total, P, d_packed, seed_logit, chunk,
                                                                    d_work_counter, d_spill, blocks, d_counts, stream);
  return batch_seeded_pick<-1, false>(d_items, d_offsets, n_items, total, P, d_packed, seed_logit, chunk, d_work_counter, d_spill, blocks,
                                      d_counts, stream);
}

}  // namespace lrt

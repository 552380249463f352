// dr.hip - MultiRendererDR::RenderDR (diff_render/MultiRendererDR.cpp:578-750)
// and OptimizeStepAdam (:828-853) on gfx950.
//
//   k_dr_primary      phase I: per-pixel primary rays, loss, dLoss/dS scatter
//   k_dr_border_find  phase II: 3x3 alpha / depth-discontinuity border pixels
//   k_dr_border_rays  phase III: border_spp jittered rays per border pixel,
//                     relaxed-boundary (silhouette) term
//   k_dr_scale        phase IV: getBaseGradientMult scaling
//   k_adam            Adam step with clamp to +-1000
// Gradients are scattered with hardware float atomics into one dLoss/dS
// buffer (the reference keeps a buffer per CPU thread and sums them).
#include <algorithm>

#include "dr_device.hpp"
#include "dr.hpp"

namespace lrtd {

__device__ __forceinline__ float lin_depth(float z) { return (z - 0.1f) / (10.0f - 0.1f); }

// Hit depths a pixel keeps in registers for its depth-discontinuity
// threshold; pixels with more hits take the streamed form below.
constexpr uint32_t kDrKeptSamples = 16;

__device__ __forceinline__ v3 dr_color(uint32_t mode, const DRHit& h, const LightSet& L) {  // MultiRendererDR.cpp:855-898
  if (mode == LRT_DR_RENDER_MODE_DIFFUSE) return h.color;
  if (mode == LRT_DR_RENDER_MODE_MASK) return v3(1, 1, 1);
  if (mode == LRT_DR_RENDER_MODE_NORMAL) return h.normal;
  if (mode == LRT_DR_RENDER_MODE_LAMBERT) {  // direct and ambient lights, in the renderer's order
    v3 c(0, 0, 0);
    for (uint32_t i = 0; i < L.n; i++) {
      const v3 lc(L.col[i].x, L.col[i].y, L.col[i].z);
      if (__float_as_uint(L.sp[i].w) == LRT_LIGHT_TYPE_DIRECT) {
        const float q = fmaxf(0.0f, dot(h.normal, v3(L.sp[i].x, L.sp[i].y, L.sp[i].z)));
        c = c + (lc * h.color) * q;
      } else {
        c = c + lc * h.color;
      }
    }
    return c;
  }
  float d = lin_depth(h.t);
  return v3(d, d, d);
}

__device__ __forceinline__ float loss_grad1(uint32_t fn, float c, float r) {
  return fn == LRT_DR_LOSS_FUNCTION_MSE ? 2.0f * (c - r) : sgn(c - r);
}

// t_samples_to_threshold (MultiRendererDR.cpp:997-1063) in three steps: the
// depth range, the 10-bin histogram over it, the longest empty run of bins.
// A pixel with more samples than it can keep in registers takes the range
// from one pass over its rays and bins the depths while it traces them again
// (same rays, same hits), with the same operations as the stored form.
__device__ __forceinline__ int depth_bin(float t, float mnv, float diff) { return min(max(int((t - mnv) / diff * 10), 0), 9); }

__device__ __forceinline__ float depth_gap_threshold(const unsigned g[10], float mnv, float diff) {
  int best_len = 0, best_start = -1, start = -1;
  for (int i = 1; i < 10; i++) {
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
  if (best_len > 4) return mnv + diff * (best_start + 0.5f * best_len) / 10;
  return 0.0f;
}

// depth range of `count` hits (running min / max mnv, mxv) and dc default
// values dv; returns the bin width base diff and seeds g with the defaults
__device__ __forceinline__ float depth_range(float& mnv, float& mxv, float dv, unsigned dc, unsigned g[10]) {
  if (dc > 0) { mnv = fminf(mnv, dv); mxv = fmaxf(mxv, dv); }
  const float diff = fmaxf(1e-9f, mxv - mnv);
  for (int i = 0; i < 10; i++) g[i] = 0u;
  g[depth_bin(dv, mnv, diff)] += dc;
  return diff;
}

__device__ float samples_to_threshold(const float* s, unsigned count, float dv, unsigned dc) {
  float mnv = s[0], mxv = s[0];
  for (unsigned i = 1; i < count; i++) { mnv = fminf(mnv, s[i]); mxv = fmaxf(mxv, s[i]); }
  unsigned g[10];
  const float diff = depth_range(mnv, mxv, dv, dc, g);
  for (unsigned i = 0; i < count; i++) g[depth_bin(s[i], mnv, diff)]++;
  return depth_gap_threshold(g, mnv, diff);
}

template <int SPEC>
__device__ __forceinline__ void scatter_pd(const DRArgs& A, float* __restrict__ grad, const DRPayload& py, v3 dLdC) {
  const uint32_t f = dr_flags<SPEC>(DRParams{A.Q.ray_flags, A.Q.dtdv_fd});
  if ((f & LRT_DR_RAY_FLAG_DDIFFUSE_DCOLOR) && A.mode != LRT_DR_RENDER_MODE_MASK) {
#pragma unroll
    for (int k = 0; k < 8; k++) {
      if (py.col_idx[k] == 0xFFFFFFFFu) continue;
      float w = py.col_w[k];
      atomicAdd(&grad[py.col_idx[k] + 0], dLdC.x * w);
      atomicAdd(&grad[py.col_idx[k] + 1], dLdC.y * w);
      atomicAdd(&grad[py.col_idx[k] + 2], dLdC.z * w);
    }
  }
  if (f & LRT_DR_RAY_FLAG_DDIFFUSE_DPOS) {
#pragma unroll
    for (int k = 0; k < 8; k++)
      if (py.pd_idx[k] != 0xFFFFFFFFu) atomicAdd(&grad[py.pd_idx[k]], dot(dLdC, py.pd_ddiffuse[k]));
  } else if (f & LRT_DR_RAY_FLAG_DDIST_DPOS) {
#pragma unroll
    for (int k = 0; k < 8; k++) {
      if (py.pd_idx[k] == 0xFFFFFFFFu) continue;
      float v = py.pd_ddist[k] / (10.0f - 0.1f);
      atomicAdd(&grad[py.pd_idx[k]], dot(dLdC, v3(v, v, v)));
    }
  }
}

// Phase I for one pixel (MultiRendererDR.cpp:1084-1294). With spp > 1 the
// colour is needed before any sample's gradient can be weighted, so the
// samples are traced twice (first for colour, then with gradients); the
// second pass recomputes identical rays and hits. SPEC kDrDepth is used with
// spp 1 only (one pass, gradient flags on).
template <int SPEC>
__device__ __forceinline__ void dr_pixel(const DRArgs& A, const DRView& V, uint32_t packed_idx, Stack& st, float& loss_acc) {
  uint32_t x, y;
  tid_to_xy(packed_idx, A.packed_w, x, y);
  if (x >= A.W || y >= A.H) return;
  const size_t pix = (size_t)y * A.W + x;
  DRPayload py;
  py.scatter = 0;
  v3 res(0, 0, 0);
  int hit_count = 0;
  float samples[kDrKeptSamples];
  float t_mn = 0.0f, t_mx = 0.0f;  // range of the hit depths, for pixels with more hits than samples[] holds
  const uint32_t spp = SPEC == kDrDepth ? 1u : A.spp;
  // kDrNormal always traces twice: the second pass scatters at the hit
  const bool one_pass = spp == 1 && SPEC != kDrNormal;
  const DRParams Qc = {one_pass ? A.Q.ray_flags : 0u, A.Q.dtdv_fd};
  const DRParams Qg = {A.Q.ray_flags, A.Q.dtdv_fd};
  py.scatter = 0;
  py.mode = A.mode;
  py.grad = V.grad;
  if (SPEC == kDrNormal) py.lights = &A.lights;
  for (uint32_t s = 0; s < spp; s++) {
    float2 d = rand2(x, y, s, A.seed);
    v3 pos, dir;
    eye_ray(V.cam, x, y, d.x, d.y, A.W, A.H, pos, dir);
    DRHit hit = dr_nearest_hit<SPEC>(A.S, A.bvals, A.bminmax, Qc, A.relax_eps, pos, dir, 1e9f, &py, st);
    if (hit.primId == 0xFFFFFFFFu) continue;
    v3 c = dr_color(A.mode, hit, A.lights);
    res = res + v3(c.x / (float)spp, c.y / (float)spp, c.z / (float)spp);
    if (hit_count < (int)kDrKeptSamples) samples[hit_count] = hit.t;
    t_mn = hit_count ? fminf(t_mn, hit.t) : hit.t;
    t_mx = hit_count ? fmaxf(t_mx, hit.t) : hit.t;
    hit_count++;
  }
  const float cm[4] = {res.x, res.y, res.z, (float)hit_count / (float)spp};
  V.out_image[pix] = make_float4(cm[0], cm[1], cm[2], cm[3]);
  // depth-discontinuity threshold (written only where phase II reads it)
  const bool streamed = hit_count > (int)kDrKeptSamples;
  unsigned hist[10];
  float h_diff = 0.0f;
  // a view rendered for its image and loss only (grad null: an earlier image
  // of an Adam step's batch) takes no gradient pass
  const bool grad_pass = A.Q.ray_flags != 0 && V.grad != nullptr;
  if (V.depth_thr) {
    if (!streamed) {
      V.depth_thr[pix] = hit_count == 0 ? 0.0f : samples_to_threshold(samples, hit_count, 1e9f, spp - hit_count);
    } else {
      h_diff = depth_range(t_mn, t_mx, 1e9f, spp - hit_count, hist);
      if (!grad_pass) {  // no gradient pass follows: trace again for the bins
        for (uint32_t s = 0; s < spp; s++) {
          float2 d = rand2(x, y, s, A.seed);
          v3 pos, dir;
          eye_ray(V.cam, x, y, d.x, d.y, A.W, A.H, pos, dir);
          DRHit hit = dr_nearest_hit<SPEC>(A.S, A.bvals, A.bminmax, Qc, A.relax_eps, pos, dir, 1e9f, &py, st);
          if (hit.primId != 0xFFFFFFFFu) hist[depth_bin(hit.t, t_mn, h_diff)]++;
        }
        V.depth_thr[pix] = depth_gap_threshold(hist, t_mn, h_diff);
      }
    }
  }
  const float4 r = V.ref[pix];
  const float rr[4] = {r.x, r.y, r.z, r.w};
  float l = 0.0f;
  for (int k = 0; k < 4; k++) {
    float df = cm[k] - rr[k];
    l += A.loss_fn == LRT_DR_LOSS_FUNCTION_MSE ? df * df : fabsf(df);
  }
  loss_acc += l;
  const v3 dLdC(loss_grad1(A.loss_fn, cm[0], rr[0]) / (float)spp, loss_grad1(A.loss_fn, cm[1], rr[1]) / (float)spp,
                loss_grad1(A.loss_fn, cm[2], rr[2]) / (float)spp);
  if (!grad_pass) return;
  if (one_pass) {
    if (hit_count) scatter_pd<SPEC>(A, V.grad, py, dLdC);
    return;
  }
  if (SPEC == kDrNormal) {
    if (hit_count == 0) return;
    py.scatter = 1;
    py.dldc = dLdC;
  }
  for (uint32_t s = 0; s < spp; s++) {
    float2 d = rand2(x, y, s, A.seed);
    v3 pos, dir;
    eye_ray(V.cam, x, y, d.x, d.y, A.W, A.H, pos, dir);
    DRHit hit = dr_nearest_hit<SPEC>(A.S, A.bvals, A.bminmax, Qg, A.relax_eps, pos, dir, 1e9f, &py, st);
    if (streamed && hit.primId != 0xFFFFFFFFu) hist[depth_bin(hit.t, t_mn, h_diff)]++;
    if (SPEC != kDrNormal && hit.primId != 0xFFFFFFFFu) scatter_pd<SPEC>(A, V.grad, py, dLdC);
  }
  if (streamed && V.depth_thr) V.depth_thr[pix] = depth_gap_threshold(hist, t_mn, h_diff);
}

// Phase I over one view (MV false: the view of A) or over the A.n_views
// records of A.views (MV true: the images of one Adam step's batch in one
// launch, each view's pixels in its own 256-pixel chunks).
template <int SPEC, bool MV>
__global__ __launch_bounds__(256) void k_dr_primary(DRArgs A, uint32_t* work_counter, uint32_t* spill, float* loss) {
  __shared__ uint32_t lds_stack[kStackLds * 256];
  __shared__ uint32_t chunk_base;
  Stack st;
  st.lds = (lds_u32*)(lds_stack + threadIdx.x);
  st.lds_stride = 256;
  st.lds_entries = kStackLds;
  st.spill = spill + blockIdx.x * 256 + threadIdx.x;
  st.spill_stride = gridDim.x * 256;
  float acc = 0.0f;
  const uint32_t total = A.W * A.H;  // the reference walks the first W*H packed indices
  const uint32_t per_view = (total + 255) / 256;
  const uint32_t n_chunks = MV ? per_view * A.n_views : per_view;
  for (;;) {
    if (threadIdx.x == 0) chunk_base = atomicAdd(work_counter, 1u);
    __syncthreads();
    const uint32_t c = __builtin_amdgcn_readfirstlane(chunk_base);
    __syncthreads();
    if (c >= n_chunks) break;
    if (MV) {
      const uint32_t v = c / per_view;  // block-uniform
      const uint32_t i = (c - v * per_view) * 256 + threadIdx.x;
      if (i < total) dr_pixel<SPEC>(A, A.views[v], i, st, acc);
    } else {
      const DRView V = {A.cam, A.ref, A.out_image, A.depth_thr, A.grad};
      const uint32_t i = c * 256 + threadIdx.x;
      if (i < total) dr_pixel<SPEC>(A, V, i, st, acc);
    }
  }
  // loss: per-lane partial sums, a wave sum, one atomic per wave
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
  if ((threadIdx.x & 63) == 0 && acc != 0.0f) atomicAdd(loss, acc);
}

// Phase II (MultiRendererDR.cpp:609-685)
__global__ void k_dr_border_find(DRArgs A, uint32_t* list, uint32_t* count) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= A.W * A.H) return;
  uint32_t x, y;
  tid_to_xy(i, A.packed_w, x, y);
  if (x >= A.W || y >= A.H) return;
  const float d0 = A.out_image[(size_t)y * A.W + x].w;
  float max_diff = 0.0f, max_thr = 0.0f;
  if (x >= 1 && x < A.W - 1 && y >= 1 && y < A.H - 1) {
    for (int dx = -1; dx <= 1; dx++)
      for (int dy = -1; dy <= 1; dy++) {
        size_t q = (size_t)(y + dy) * A.W + x + dx;
        max_diff = fmaxf(max_diff, fabsf(d0 - A.out_image[q].w));
        max_thr = fmaxf(max_thr, A.depth_thr[q]);
      }
  }
  bool border = A.mode == LRT_DR_RENDER_MODE_MASK ? max_diff > 0 : (max_diff > 0 || max_thr > 0);
  if (border || A.forced_border) list[atomicAdd(count, 1u)] = i;
}

// CalculateBorderRayDerivatives (MultiRendererDR.cpp:1319-1375) of one border
// ray that found a relaxation point: pdf * (1/eps) * <dL/dC, C_in - C_out> *
// dSDF/dtheta added to the 8 values of the relaxation voxel.
__device__ __forceinline__ void border_contribution(const DRArgs& A, const DRPayload& py, const DRHit& hit, v3 pos, v3 dir, float pdf) {
  float delta[4];
  if (A.input_type == LRT_DR_INPUT_TYPE_LINEAR_DEPTH) {
    float din = lin_depth(py.missed_hit.t);
    float dout = hit.primId == 0xFFFFFFFFu ? 0.0f : lin_depth(hit.t);
    delta[0] = delta[1] = delta[2] = din - dout;
  } else {
    v3 fin = dr_color(A.mode, py.missed_hit, A.lights);
    v3 fout = hit.primId == 0xFFFFFFFFu ? v3(0, 0, 0) : dr_color(A.mode, hit, A.lights);
    delta[0] = fin.x - fout.x; delta[1] = fin.y - fout.y; delta[2] = fin.z - fout.z;
  }
  delta[3] = hit.primId == 0xFFFFFFFFu ? 1.0f : 0.0f;
  for (int j = 0; j < 8; j++) {
    if (py.missed_idx[j] == 0xFFFFFFFFu) continue;
    float bd = mx(0.0f, py.missed_hit.sdf);
    v3 ystar = pos + py.missed_hit.t * dir;
    v3 xstar = ystar - bd * py.missed_hit.normal;
    float ps[4];
    mul4(A.proj_view, xstar.x, xstar.y, xstar.z, 1.0f, ps);
    float sx = 0.5f * (ps[0] / ps[3] + 1.0f), sy = 1.0f - 0.5f * (ps[1] / ps[3] + 1.0f);
    // the reference converts to unsigned without a range check; clamped here
    float fx = floorf(sx * (float)A.W), fy = floorf(sy * (float)A.H);
    uint32_t xp = (uint32_t)fminf(fmaxf(fx, 0.0f), (float)(A.W - 1)), yp = (uint32_t)fminf(fmaxf(fy, 0.0f), (float)(A.H - 1));
    size_t q = (size_t)yp * A.W + xp;
    float4 c = A.out_image[q], r = A.ref[q];
    float dd = loss_grad1(A.loss_fn, c.x, r.x) * delta[0] + loss_grad1(A.loss_fn, c.y, r.y) * delta[1] +
               loss_grad1(A.loss_fn, c.z, r.z) * delta[2] + loss_grad1(A.loss_fn, c.w, r.w) * delta[3];
    float diff = pdf * (1.0f / A.relax_eps) * dd * py.missed_dsdf[j];
    atomicAdd(&A.grad[py.missed_idx[j]], diff);
  }
}

// Phase III (MultiRendererDR.cpp:1377-1499), random border sampling: one lane
// per (border pixel, sample)
template <int SPEC>
__global__ __launch_bounds__(256) void k_dr_border_rays(DRArgs A, const uint32_t* list, const uint32_t* count, uint32_t* spill,
                                                        uint32_t* hits) {
  __shared__ uint32_t lds_stack[kStackLds * 256];
  Stack st;
  st.lds = (lds_u32*)(lds_stack + threadIdx.x);
  st.lds_stride = 256;
  st.lds_entries = kStackLds;
  st.spill = spill + blockIdx.x * 256 + threadIdx.x;
  st.spill_stride = gridDim.x * 256;
  const uint64_t n = (uint64_t)(*count) * A.border_spp;
  const DRParams Qb = {LRT_DR_RAY_FLAG_BORDER, A.Q.dtdv_fd};
  uint32_t local_hits = 0;
  DRPayload py;
  py.scatter = 0;
  for (uint64_t g = (uint64_t)blockIdx.x * 256 + threadIdx.x; g < n; g += (uint64_t)gridDim.x * 256) {
    const uint32_t item = (uint32_t)(g / A.border_spp), s = (uint32_t)(g % A.border_spp);
    uint32_t x, y;
    tid_to_xy(list[item], A.packed_w, x, y);
    float2 d = rand2(x, y, s, A.seed);
    v3 pos, dir;
    eye_ray(A.cam, x, y, d.x, d.y, A.W, A.H, pos, dir);
    DRHit hit = dr_nearest_hit<SPEC>(A.S, A.bvals, A.bminmax, Qb, A.relax_eps, pos, dir, 1e9f, &py, st);
    if (!(py.missed_hit.sdf < A.relax_eps)) continue;
    local_hits++;
    border_contribution(A, py, hit, pos, dir, 1.0f / (float)A.border_spp);
  }
  if (local_hits) atomicAdd(hits, local_hits);
}

// ---- SVM border sampling (MultiRendererDR.cpp:1501-1932) -------------------


// SVM_fit (:1518-1546): a perceptron with weight decay on the labelled
// samples X = (x, y, 1) * label * etha (etha 0.01, alpha 0.1), at most 200
// epochs, stopping after an epoch without errors; returns W and the last
// epoch's error rate. The samples are spread over the wave (sample s in lane
// s % 64, slot s / 64): every lane runs the same sequential epochs on a
// wave-uniform W and reads sample j's terms from its lane with readlane, in
// the reference's order and with its float operations.
__device__ float4 svm_fit_wave(const float* xs, const float* ys, const float* lab, uint32_t N, uint32_t lane) {
  const float etha = 0.01f, alpha = 0.1f;
  const int epochs = 200;
  float X0[4], X1[4], X2[4];
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const uint32_t s = lane + 64u * k;
    const float l = s < N ? lab[s] : 0.0f;
    X0[k] = s < N ? xs[s] * l * etha : 0.0f;
    X1[k] = s < N ? ys[s] * l * etha : 0.0f;
    X2[k] = s < N ? 1.0f * l * etha : 0.0f;
  }
  float w0 = -1.0f, w1 = 1.0f, w2 = 0.0f;
  const float weightMult = 1.0f - etha * alpha / epochs;
  uint32_t errors = N;
  for (int i = 0; i < epochs && errors > 0; i++) {
    errors = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const uint32_t n_k = N > 64u * k ? min(64u, N - 64u * k) : 0u;
      for (uint32_t src = 0; src < n_k; src++) {
        const float a0 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(X0[k]), src));
        const float a1 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(X1[k]), src));
        const float a2 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(X2[k]), src));
        const float margin = w0 * a0 + w1 * a1 + w2 * a2;
        w0 = w0 * weightMult;
        w1 = w1 * weightMult;
        w2 = w2 * weightMult;
        if (margin < 0) {
          w0 = w0 + a0;
          w1 = w1 + a1;
          w2 = w2 + a2;
          errors++;
        }
      }
    }
  }
  return make_float4(w0, w1, w2, float(errors) / float(N));
}

__device__ __forceinline__ uint32_t out_code(float x, float y) {  // ComputeOutCode, :1550-1571, window [0,1]^2
  uint32_t code = 0;
  if (x < 0) code |= 1u;
  else if (x > 1) code |= 2u;
  if (y < 0) code |= 4u;
  else if (y > 1) code |= 8u;
  return code;
}

// CohenSutherlandLineClip, :1576-1645
__device__ bool clip_line(float& x0, float& y0, float& x1, float& y1) {
  uint32_t c0 = out_code(x0, y0), c1 = out_code(x1, y1);
  for (int guard = 0; guard < 16; guard++) {  // each pass moves one end onto the window: at most 4 needed
    if (!(c0 | c1)) return true;
    if (c0 & c1) return false;
    float x = -1, y = -1;
    const uint32_t co = c1 > c0 ? c1 : c0;
    if (co & 8u) {
      x = x0 + (x1 - x0) * (1 - y0) / (y1 - y0);
      y = 1;
    } else if (co & 4u) {
      x = x0 + (x1 - x0) * (0 - y0) / (y1 - y0);
      y = 0;
    } else if (co & 2u) {
      y = y0 + (y1 - y0) * (1 - x0) / (x1 - x0);
      x = 1;
    } else if (co & 1u) {
      y = y0 + (y1 - y0) * (0 - x0) / (x1 - x0);
      x = 0;
    }
    if (co == c0) {
      x0 = x; y0 = y; c0 = out_code(x0, y0);
    } else {
      x1 = x; y1 = y; c1 = out_code(x1, y1);
    }
  }
  return false;
}

__device__ __forceinline__ float2 world_to_screen(const DRArgs& A, v3 p) {  // TransformWorldToScreenSpace, :1296-1301
  float ps[4];
  mul4(A.proj_view, p.x, p.y, p.z, 1.0f, ps);
  return make_float2(0.5f * (ps[0] / ps[3] + 1.0f), 1.0f - 0.5f * (ps[1] / ps[3] + 1.0f));
}

constexpr uint32_t kSvmMaxSamples = 256;  // svm_spp = 0.25 * border_spp, rounded down to a square

// Orders one wave's LDS stores before its later LDS loads (other lanes).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// CastBorderRaySVM (:1651-1931): one wavefront per border pixel. The lanes
// trace the SVM samples (a jittered sqrt(n) x sqrt(n) grid over the pixel) in
// parallel; lane 0 labels them by the depth threshold, fits the separating
// line, clips it to the pixel and builds the sampling stripe; then the lanes
// trace the stripe's border rays. The reference draws the stripe positions
// with rand(); here they come from rand2 of the sample index + 0x10000.
template <int SPEC>
__global__ __launch_bounds__(256) void k_dr_border_svm(DRArgs A, const uint32_t* list, const uint32_t* count, uint32_t* spill,
                                                       uint32_t* hits) {
  __shared__ uint32_t lds_stack[kStackLds * 256];
  __shared__ float s_x[4][kSvmMaxSamples], s_y[4][kSvmMaxSamples], s_t[4][kSvmMaxSamples], s_l[4][kSvmMaxSamples];
  __shared__ float s_geo[4][16];  // p0, r, n, radius, pdf, force, n_rays, skip; labels' skip, type0, closest, closest t
  // Each wave owns row wv of these arrays, so waves run through their pixels
  // independently and synchronise only with themselves.
  Stack st;
  st.lds = (lds_u32*)(lds_stack + threadIdx.x);
  st.lds_stride = 256;
  st.lds_entries = kStackLds;
  st.spill = spill + blockIdx.x * 256 + threadIdx.x;
  st.spill_stride = gridDim.x * 256;
  const uint32_t wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const DRParams Q0 = {LRT_DR_RAY_FLAG_NO_DIFF, 0u}, Qb = {LRT_DR_RAY_FLAG_BORDER, A.Q.dtdv_fd};
  const uint32_t n_items = *count;
  const uint32_t sq = A.svm_side;
  const uint32_t svm_spp = sq * sq;
  const float i_sq = 1.0f / (float)sq;
  uint32_t local_hits = 0;
  DRPayload py;
  py.scatter = 0;
  for (uint32_t item = blockIdx.x * 4 + wv; item < n_items; item += gridDim.x * 4) {  // uniform within the wave
    uint32_t x, y;
    tid_to_xy(list[item], A.packed_w, x, y);
    // step 1: SVM samples
    for (uint32_t s = lane; s < svm_spp; s += 64) {
      const float2 rnd = rand2(x, y, s, A.seed);
      const float dx = i_sq * ((float)(s / sq) + rnd.x), dy = i_sq * ((float)(s % sq) + rnd.y);
      v3 pos, dir;
      eye_ray(A.cam, x, y, dx, dy, A.W, A.H, pos, dir);
      // Only t is used. The linear-depth primary form finds the same t (its
      // voxel and brick skips drop only voxels that cannot hold a root) and,
      // with scatter off, writes nothing; on the border pass's BVH copy it
      // descends as over the full tree.
      const DRHit hit = dr_nearest_hit<SPEC == kDrDepthBorder ? kDrDepth : SPEC>(A.S, A.bvals, A.bminmax, Q0, A.relax_eps, pos, dir, 1e9f,
                                                                                &py, st);
      s_x[wv][s] = dx;
      s_y[wv][s] = dy;
      s_t[wv][s] = hit.t;
    }
    wave_sync();
    if (lane == 0) {
      // step 2: threshold and labels (:1716-1736)
      const float thr = samples_to_threshold(s_t[wv], svm_spp, 0.0f, 0u);
      uint32_t type0 = 0, closest = 0;
      float closest_t = 1e6f;
      for (uint32_t s = 0; s < svm_spp; s++) {
        const float t = s_t[wv][s];
        if (t < closest_t) { closest_t = t; closest = s; }
        s_l[wv][s] = t <= thr ? -1.0f : 1.0f;
        type0 += t <= thr ? 1u : 0u;
      }
      s_geo[wv][12] = (type0 == svm_spp || type0 == 0) ? 1.0f : 0.0f;
      s_geo[wv][13] = __uint_as_float(type0);
      s_geo[wv][14] = __uint_as_float(closest);
      s_geo[wv][15] = closest_t;
    }
    wave_sync();
    const bool one_class = s_geo[wv][12] != 0.0f;  // uniform within the wave
    // the fit (:1740), wave-wide
    const float4 W = one_class ? make_float4(0, 0, 0, 0) : svm_fit_wave(s_x[wv], s_y[wv], s_l[wv], svm_spp, lane);
    if (lane == 0) {
      const uint32_t type0 = __float_as_uint(s_geo[wv][13]), closest = __float_as_uint(s_geo[wv][14]);
      const float closest_t = s_geo[wv][15];
      float skip = 0.0f, force = 0.0f, stripe_area = 1.0f, sample_radius = 0.0f;
      float2 p0 = make_float2(0, 0), p1 = make_float2(0, 0), r = make_float2(0, 0), n = make_float2(0, 0);
      if (one_class) {
        skip = 1.0f;
      } else {
        const float error_rate = W.w;
        const float f0 = (float)type0 / (float)svm_spp;
        const float rare = fminf(f0, 1.0f - f0);
        if (error_rate > 0.05f || 2 * error_rate > rare) {
          force = 1.0f;
          stripe_area = 1.0f;
        } else {
          if (fabsf(W.y) > 1e-6f && fabsf(W.x) > 1e-6f) {
            p0 = make_float2(0, -W.z / W.y);
            p1 = make_float2(1, (-W.z - W.x) / W.y);
            const float ry = p1.y - p0.y;
            if (p0.y < -1e-6f || p0.y > 1 + 1e-6f) {
              const float k1 = -p0.y / ry, k2 = (1 - p0.y) / ry, k = (fabsf(k1) < fabsf(k2)) ? k1 : k2;
              p0 = make_float2(p0.x + k * 1.0f, p0.y + k * ry);
            }
            if (p1.y < -1e-6f || p1.y > 1 + 1e-6f) {
              const float k1 = -p1.y / ry, k2 = (1 - p1.y) / ry, k = (fabsf(k1) < fabsf(k2)) ? k1 : k2;
              p1 = make_float2(p1.x + k * 1.0f, p1.y + k * ry);
            }
          } else if (fabsf(W.x) > 1e-6f) {
            p0 = make_float2(-W.z / W.x, 0);
            p1 = make_float2(-W.z / W.x, 1);
          } else {
            force = 1.0f;  // the reference leaves p0, p1 unset here; they are (0,0)
          }
          if (!clip_line(p0.x, p0.y, p1.x, p1.y)) {
            skip = 1.0f;
          } else {
            // border thickness from the closest hit (:1781-1788)
            float2 q0, q1;
            {
              const float rnd_x = rand2(x, y, closest, A.seed).x, rnd_y = rand2(x, y, closest, A.seed).y;
              v3 pos, dir;
              eye_ray(A.cam, x, y, i_sq * ((float)(closest / sq) + rnd_x), i_sq * ((float)(closest % sq) + rnd_y), A.W, A.H, pos, dir);
              const v3 ch = closest_t < 1e6f ? pos + dir * closest_t : v3(1e6f, 1e6f, 1e6f);
              q0 = world_to_screen(A, ch);
              q1 = world_to_screen(A, ch + v3(A.up[0], A.up[1], A.up[2]) * A.relax_eps);
            }
            const float ddx = q1.x - q0.x, ddy = q1.y - q0.y;
            const float thick = sqrtf(ddx * ddx + ddy * ddy) * (float)max(A.W, A.H);
            sample_radius = fmaxf(0.15f, thick) + 2.0f * error_rate;
            r = make_float2(p1.x - p0.x, p1.y - p0.y);
            float lr = sqrtf(r.x * r.x + r.y * r.y);
            const float2 rn = make_float2(r.x / lr, r.y / lr);
            p0 = make_float2(p0.x - sample_radius * rn.x, p0.y - sample_radius * rn.y);
            p1 = make_float2(p1.x + sample_radius * rn.x, p1.y + sample_radius * rn.y);
            r = make_float2(p1.x - p0.x, p1.y - p0.y);
            lr = sqrtf(r.x * r.x + r.y * r.y);
            const float ln = sqrtf(r.y * r.y + r.x * r.x);
            n = make_float2(-r.y / ln, r.x / ln);
            stripe_area = 2 * sample_radius * lr;
            if (!force && stripe_area < 0.005f) skip = 1.0f;
          }
        }
      }
      const float area_mult = fminf(1.0f, 0.05f + 2.0f * stripe_area);
      const uint32_t nb = max(1u, (uint32_t)(area_mult * (float)(A.border_spp - svm_spp)));
      float* g = s_geo[wv];
      g[0] = p0.x; g[1] = p0.y; g[2] = r.x; g[3] = r.y; g[4] = n.x; g[5] = n.y;
      g[6] = sample_radius;
      g[7] = force != 0.0f ? 1.0f / (float)nb : stripe_area / (float)nb;
      g[8] = force;
      g[9] = __uint_as_float(nb);
      g[10] = skip;
    }
    wave_sync();
    const float* g = s_geo[wv];
    if (g[10] == 0.0f) {
      // step 3: border rays in the stripe (or the whole pixel)
      const uint32_t nb = __float_as_uint(g[9]);
      for (uint32_t k = lane; k < nb; k += 64) {
        const uint32_t sid = svm_spp + k;
        float2 d;
        if (g[8] != 0.0f) {
          d = rand2(x, y, sid, A.seed);
        } else {
          const float2 u = rand2(x, y, sid + 0x10000u, A.seed);
          const float u1 = u.x, u2 = 2.0f * u.y - 1.0f;
          d = make_float2(g[0] + u1 * g[2] + u2 * g[6] * g[4], g[1] + u1 * g[3] + u2 * g[6] * g[5]);
        }
        if (!(d.x > 0 && d.x < 1 && d.y > 0 && d.y < 1)) continue;
        v3 pos, dir;
        eye_ray(A.cam, x, y, d.x, d.y, A.W, A.H, pos, dir);
        const DRHit hit = dr_nearest_hit<SPEC>(A.S, A.bvals, A.bminmax, Qb, A.relax_eps, pos, dir, 1e9f, &py, st);
        if (!(py.missed_hit.sdf < A.relax_eps)) continue;
        local_hits++;
        border_contribution(A, py, hit, pos, dir, g[7]);
      }
    }
    wave_sync();  // the next pixel reuses this wave's LDS rows
  }
  if (local_hits) atomicAdd(hits, local_hits);
}

// ---- finite-difference DR (RenderDRFiniteDiff, MultiRendererDR.cpp:751-815) ----

// Sets parameter j to p0 + delta (mode 0), p0 - delta (1) or back to p0 (2)
// in values_f and at its positions in the per-brick corner copy, and clears
// the primary pass's work counter for the render that follows.
__global__ __launch_bounds__(256) void k_dr_fd_set(float* values_f, float* bvals, const float* p0s, const uint32_t* pos, uint32_t n_pos,
                                                   uint32_t j, uint32_t mode, float delta, uint32_t* counter) {
  const float p0 = p0s[j];
  const float v = mode == 0 ? p0 + delta : (mode == 1 ? p0 - delta : p0);
  for (uint32_t k = threadIdx.x; k < n_pos; k += 256) bvals[pos[k]] = v;
  if (threadIdx.x == 0) {
    values_f[j] = v;
    *counter = 0u;
  }
}

// Sum over the image of Loss(colour, reference) (MultiRendererDR.cpp:51-73:
// float per pixel, summed in double), in a fixed order (one block: strided
// partial sums, then a tree in LDS), so that equal images give equal sums.
__global__ __launch_bounds__(1024) void k_dr_fd_loss(const float4* img, const float4* ref, uint32_t n, uint32_t loss_fn, double* out) {
  __shared__ double part[1024];
  double acc = 0.0;
  for (uint32_t i = threadIdx.x; i < n; i += 1024) {
    const float4 c = img[i], r = ref[i];
    float l;
    if (loss_fn == LRT_DR_LOSS_FUNCTION_MSE) {
      const float dx = c.x - r.x, dy = c.y - r.y, dz = c.z - r.z, dw = c.w - r.w;
      l = dx * dx + dy * dy + dz * dz + dw * dw;
    } else {
      l = fabsf(c.x - r.x) + fabsf(c.y - r.y) + fabsf(c.z - r.z) + fabsf(c.w - r.w);
    }
    acc += (double)l;
  }
  part[threadIdx.x] = acc;
  __syncthreads();
  for (uint32_t h = 512; h > 0; h >>= 1) {
    if (threadIdx.x < h) part[threadIdx.x] += part[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = part[0];
}

// dLoss/dS[j] = (g * L(p0 + delta) - g * L(p0 - delta)) / (2 delta), as float
__global__ __launch_bounds__(256) void k_dr_fd_grad(const double* acc, const uint32_t* ids, uint32_t n, float g_mult, float delta,
                                                    float* grad) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const double lp = acc[2 * i] * g_mult, lm = acc[2 * i + 1] * g_mult;
  grad[ids[i]] = (float)((lp - lm) / (2 * delta));
}

// Per-brick copy of the corner values the voxel walk reads:
// out[b * v^3 + k] = values_f[values[data_offset(b) + k]] (indices checked
// against values_f when the scene was set). One dependent load per corner in
// the ray casts instead of two; refreshed before every RenderDR, so it always
// equals the current parameters.
__global__ void __launch_bounds__(256) k_dr_gather(DevScene S, float* __restrict__ out, float2* __restrict__ minmax) {
  // one wavefront per brick; its smallest and largest value are folded with
  // fminf / fmaxf from the brick's first value, as a sequential scan would
  const uint32_t b = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (b >= S.n_nodes) return;  // uniform per wavefront
  const uint32_t nv = S.v_size * S.v_size * S.v_size;
  const uint32_t* idx = S.values + S.nodes[b].z;
  float* o = out + (size_t)b * nv;
  const float v0 = S.values_f[idx[0]];
  float lo = v0, hi = v0;
  for (uint32_t k = lane; k < nv; k += 64) {
    const float v = S.values_f[idx[k]];
    o[k] = v;
    lo = fminf(lo, v);
    hi = fmaxf(hi, v);
  }
#pragma unroll
  for (int m = 32; m > 0; m >>= 1) {
    lo = fminf(lo, __shfl_xor(lo, m, 64));
    hi = fmaxf(hi, __shfl_xor(hi, m, 64));
  }
  if (lane == 0) minmax[b] = make_float2(lo, hi);
}

// Copies of the BVH for the linear-depth passes.
//
// The primary pass skips, without any effect, every brick whose gathered
// corner values all exceed 1e-6 and 1e-3 of their maximum (dr_brick), so a
// subtree holding only such bricks need not be entered: its slot (one child
// of a node pair) gets a box at the point 1e30, which no ray reaches within
// its tFar. The traversal then visits the bricks it does not skip in the same
// order as over the full tree, and no others.
//
// The border pass skips in addition every brick whose values all exceed
// relax_eps by the same margin, but such a skip clears two parts of the
// ray's state when the ray crosses the brick. A subtree of such bricks whose
// bricks tile its box exactly (tiled[slot], computed on the host) keeps its
// box and gets the marker kDrCulledSlot in its escapeIndex: the traversal
// treats it as one leaf that applies the clearing if the ray crosses the box
// (past the bricks' tNear), in the order the subtree would have been visited.
//
// k_dr_cull_mark marks, for each pass, every slot on the path from each brick
// that pass does not skip up to the root; k_dr_cull_write writes both copies.
__global__ void k_dr_cull_clear(uint32_t* __restrict__ occ_p, uint32_t* __restrict__ occ_b, const uint8_t* __restrict__ tiled,
                                uint32_t n_slots) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n_slots) return;
  occ_p[s] = 0u;
  occ_b[s] = tiled[s] ? 0u : 1u;  // a slot whose bricks do not tile its box is never culled
}

__device__ __forceinline__ void cull_mark_path(uint32_t* occ, const uint32_t* __restrict__ pair_slot, uint32_t s) {
  // a parent pair precedes its children (checked on the host), so the walk
  // ends at the root within the tree depth; the bound is a second guard
  for (int it = 0; it < 256 && s != 0xFFFFFFFFu; it++) {
    if (atomicOr(&occ[s], 1u) != 0u) break;  // the path above is marked or being marked
    s = pair_slot[s >> 1];
  }
}

__global__ void k_dr_cull_mark(const float2* __restrict__ bminmax, const uint32_t* __restrict__ leaf_slot,
                               const uint32_t* __restrict__ pair_slot, uint32_t* occ_p, uint32_t* occ_b, float relax_eps,
                               uint32_t n_bricks) {
  const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= n_bricks) return;
  const float2 mm = bminmax[b];
  const bool skip_p = mm.x > 1e-6f && mm.x > 1e-3f * mm.y;
  const bool skip_b = skip_p && mm.x > relax_eps + 1e-3f * mm.y;
  const uint32_t s = leaf_slot[b];
  if (!skip_p) cull_mark_path(occ_p, pair_slot, s);
  if (!skip_b) cull_mark_path(occ_b, pair_slot, s);
}

__global__ void k_dr_cull_write(const uint4* __restrict__ pairs, const uint32_t* __restrict__ occ_p, const uint32_t* __restrict__ occ_b,
                                uint4* __restrict__ out_p, uint4* __restrict__ out_b, uint32_t n_slots) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n_slots) return;
  const uint4 lo = pairs[2 * s], hi = pairs[2 * s + 1];  // {boxMin, leftOffset}, {boxMax, escapeIndex}
  uint4 plo = lo, phi = hi;
  if (occ_p[s] == 0u) {
    const uint32_t far = __float_as_uint(1e30f);
    plo.x = plo.y = plo.z = far;
    phi.x = phi.y = phi.z = far;
  }
  out_p[2 * s] = plo;
  out_p[2 * s + 1] = phi;
  uint4 bhi = hi;
  bhi.w = occ_b[s] == 0u ? kDrCulledSlot : 0u;
  out_b[2 * s] = lo;
  out_b[2 * s + 1] = bhi;
}

// MultiRendererDR::Regularization (MultiRendererDR.cpp:1933-2057): every
// distance value of a brick is pulled toward its in-brick neighbours along
// the three axes, lambda * power * sign(d) * |d|^(power-1) per neighbour with
// d = own value - neighbour value. One lane per (brick, value); the six terms
// are summed in the reference's neighbour order, then added with one atomic.
__global__ void k_dr_regularize(DevScene S, uint32_t n_nodes, float lambda, float power, float* __restrict__ grad) {
  const uint32_t vs = S.brick_size + 1u, nv = vs * vs * vs;
  const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= (uint64_t)n_nodes * nv) return;
  const uint32_t b = (uint32_t)(g / nv), v = (uint32_t)(g - (uint64_t)b * nv);
  const uint32_t i0 = v / (vs * vs), j0 = (v / vs) % vs, k0 = v % vs;
  const uint32_t* idx = S.values + S.nodes[b].z;
  const float own = S.values_f[idx[v]];
  float acc = 0.0f;
  auto term = [&](uint32_t nbr) {
    float d = own - S.values_f[idx[nbr]];
    const float sgn_d = d > 0.0f ? 1.0f : -1.0f;
    d = fabsf(d);
    acc += lambda * power * sgn_d * powf(d, power - 1);
  };
  if (i0 > 0) term(v - vs * vs);
  if (i0 < vs - 1) term(v + vs * vs);
  if (j0 > 0) term(v - vs);
  if (j0 < vs - 1) term(v + vs);
  if (k0 > 0) term(v - 1);
  if (k0 < vs - 1) term(v + 1);
  atomicAdd(&grad[idx[v]], acc);
}

// MultiRendererDR::Redistance (MultiRendererDR.cpp:2122-2382) over the dense
// p^3 lattice of distance values (x-major), grid spacing h.
//
// The reference sweeps the lattice 8 times per iteration along anti-diagonal
// planes, but every update reads only the previous iteration's buffer and the
// point's own previous value, so each iteration is one Jacobi step and the
// sweep order does not change any value: here one lane updates one point per
// iteration. Points with a zero value or a sign change to a 6-neighbour are
// frozen at their value; the others start at 99999.
__device__ __forceinline__ float solve_eikonal(float m0, float m1, float m2, float h) {  // SolveEikonal, :2058-2120
  float t;
  if (m0 > m1) { t = m0; m0 = m1; m1 = t; }
  if (m0 > m2) { t = m0; m0 = m2; m2 = t; }
  if (m1 > m2) { t = m1; m1 = m2; m2 = t; }
  float dist_new = m0 + h;
  if (dist_new > m1) {
    const float h_2 = h * h;
    float m_sum = m0 + m1, c_2 = (m0 - m1) * (m0 - m1);
    dist_new = (m_sum + sqrtf(2 * h_2 - c_2)) * 0.5f;
    if (dist_new > m2) {
      const float one_third = 1.f / 3;
      m_sum += m2;
      c_2 = 3.f * (m0 * m0 + m1 * m1 + m2 * m2 - h_2);
      dist_new = (m_sum + sqrtf(m_sum * m_sum - c_2)) * one_third;
    }
  }
  return dist_new;
}

__global__ void k_redist_init(const float* __restrict__ d, uint32_t p, uint8_t* __restrict__ frozen, float* __restrict__ b0,
                              float* __restrict__ b1) {
  const uint32_t n = p * p * p;
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n) return;
  const uint32_t c[3] = {g / (p * p), (g / p) % p, g % p};
  const uint32_t off[3] = {p * p, p, 1};
  const float v = d[g];
  bool fr = v == 0.f;
  float init = fr ? 0.f : 99999.f;
  if (!fr)
    for (int dim = 0; dim < 3; dim++) {
      if (c[dim] > 0 && d[g - off[dim]] * v < 0) fr = true;
      if (c[dim] + 1 < p && d[g + off[dim]] * v < 0) fr = true;
    }
  if (fr && v != 0.f) init = v;
  frozen[g] = fr ? 1 : 0;
  b0[g] = init;
  b1[g] = init;
}

__global__ void k_redist_step(const float* __restrict__ old, float* __restrict__ nw, const uint8_t* __restrict__ frozen, uint32_t p,
                              float h) {
  const uint32_t n = p * p * p;
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n) return;
  const float big = 99999.f;
  const float own = old[g];
  nw[g] = own;
  if (frozen[g]) return;
  const uint32_t c[3] = {g / (p * p), (g / p) % p, g % p};
  const uint32_t off[3] = {p * p, p, 1};
  float am[3] = {big, big, big};
  for (int dim = 0; dim < 3; dim++) {
    if (c[dim] > 0 && c[dim] < p - 1) {
      am[dim] = old[g - off[dim]];
      const float t = old[g + off[dim]];
      if (fabsf(t) < fabsf(am[dim])) am[dim] = t;
    } else if (c[dim] > 0 && c[dim] == p - 1) {
      am[dim] = old[g - off[dim]];
    } else if (c[dim] == 0 && c[dim] < p - 1) {
      am[dim] = old[g + off[dim]];
    }
  }
  if (am[0] < big || am[1] < big || am[2] < big) {
    if (am[0] >= 0.f && am[1] >= 0.f && am[2] >= 0.f) {
      nw[g] = mn(solve_eikonal(am[0], am[1], am[2], h), own);  // std::min(a, b)
    } else {
      for (int dim = 0; dim < 3; dim++)
        if (am[dim] < 0.f) am[dim] = -am[dim];
      nw[g] = mx(-solve_eikonal(am[0], am[1], am[2], h), -fabsf(own));  // std::max(a, b)
    }
  }
}

// PreprocessRefImages (MultiRendererDR.cpp:187-226): every render pixel is the
// mean of spp_x * spp_y bilinear samples of the original image (colour, with
// the mask's coverage in w), spp = ceil(original / render size) per axis, at
// uv = ((spp * x + d + 0.5) / (spp * size)). The bilinear sample restates
// LiteImage's Sampler with filter LINEAR and the default WRAP addressing (a
// third-party header the reference includes): the linear texel filter of the
// Vulkan specification -- texel centres at (i + 0.5) / size, the four texels
// around u * size - 0.5 and v * size - 0.5, weights (1-a)(1-b), a(1-b),
// (1-a)b, ab. One thread per render pixel of every image.
__device__ __forceinline__ float4 sample_bilinear_wrap(const float4* __restrict__ img, uint32_t W0, uint32_t H0, float u, float v) {
  const float fx = u * (float)W0 - 0.5f, fy = v * (float)H0 - 0.5f;
  const float x0f = floorf(fx), y0f = floorf(fy);
  const float a = fx - x0f, b = fy - y0f;
  const int iw = (int)W0, ih = (int)H0;
  const int x0 = (((int)x0f % iw) + iw) % iw, y0 = (((int)y0f % ih) + ih) % ih;
  const int x1 = (x0 + 1) % iw, y1 = (y0 + 1) % ih;
  const float4 t00 = img[(size_t)y0 * W0 + x0], t10 = img[(size_t)y0 * W0 + x1];
  const float4 t01 = img[(size_t)y1 * W0 + x0], t11 = img[(size_t)y1 * W0 + x1];
  const float w00 = (1.0f - a) * (1.0f - b), w10 = a * (1.0f - b), w01 = (1.0f - a) * b, w11 = a * b;
  return make_float4(w00 * t00.x + w10 * t10.x + w01 * t01.x + w11 * t11.x, w00 * t00.y + w10 * t10.y + w01 * t01.y + w11 * t11.y,
                     w00 * t00.z + w10 * t10.z + w01 * t01.z + w11 * t11.z, w00 * t00.w + w10 * t10.w + w01 * t01.w + w11 * t11.w);
}

__global__ void k_dr_resample(const float4* __restrict__ src, uint32_t W0, uint32_t H0, float4* __restrict__ dst, uint32_t W, uint32_t H,
                              uint32_t n_images) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t np = (uint64_t)W * H;
  if (i >= np * n_images) return;
  const uint32_t img = (uint32_t)(i / np), pix = (uint32_t)(i - (uint64_t)img * np);
  const uint32_t x = pix % W, y = pix / W;
  const uint32_t spp_x = (uint32_t)ceilf((float)W0 / (float)W), spp_y = (uint32_t)ceilf((float)H0 / (float)H);
  const float n_s = (float)(spp_x * spp_y);
  const float4* im = src + (uint64_t)img * W0 * H0;
  float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
  for (uint32_t dx = 0; dx < spp_x; dx++)
    for (uint32_t dy = 0; dy < spp_y; dy++) {
      const float u = ((float)(spp_x * x + dx) + 0.5f) / (float)(spp_x * W);
      const float v = ((float)(spp_y * y + dy) + 0.5f) / (float)(spp_y * H);
      const float4 c = sample_bilinear_wrap(im, W0, H0, u, v);
      acc.x += c.x / n_s;
      acc.y += c.y / n_s;
      acc.z += c.z / n_s;
      acc.w += c.w / n_s;
    }
  dst[i] = acc;
}

__global__ void k_dr_scale(float* grad, uint32_t n, float mult) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) grad[i] *= mult;
}

// MultiRendererDR.cpp:828-853; c1 = 1 - beta1^(iter+1), c2 = 1 - beta2^(iter+1).
// The reference divides by them in double and rounds to float; here the
// division is FP32 (the corrected moments may differ by one ulp).
__global__ void k_adam(const float* __restrict__ g, float* X, float* V, float* Sq, uint32_t n, float lr, float b1, float b2, float eps,
                       float c1, float c2, float batch) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float gi = g[i] / batch;  // m_dLoss_dS_tmp[j] /= image_batch_size, MultiRendererDR.cpp:474-475
  float v = b1 * V[i] + (1 - b1) * gi;
  V[i] = v;
  float vh = v / c1;
  float s = b2 * Sq[i] + (1 - b2) * gi * gi;
  Sq[i] = s;
  float sh = s / c2;
  float x = X[i] - lr * vh / (sqrtf(sh) + eps);
  X[i] = fminf(fmaxf(x, -1000.0f), 1000.0f);
}

}  // namespace lrtd

namespace lrt {

// kDrDepth: linear-depth reconstruction (distance gradients only), analytic
// dt/dv, one sample per pixel, a render mode whose colour is not the diffuse
// colour. Everything else runs the generic form.
static bool dr_normal_mode(const lrtd::DRArgs& A) {
  return A.mode == LRT_DR_RENDER_MODE_LAMBERT || A.mode == LRT_DR_RENDER_MODE_NORMAL;
}
bool dr_primary_is_depth(const lrtd::DRArgs& A) {
  return A.Q.ray_flags == LRT_DR_RAY_FLAG_DDIST_DPOS && A.Q.dtdv_fd == 0 && A.spp == 1 &&
         (A.mode == LRT_DR_RENDER_MODE_LINEAR_DEPTH || A.mode == LRT_DR_RENDER_MODE_MASK);
}
bool dr_border_is_depth(const lrtd::DRArgs& A) { return A.input_type == LRT_DR_INPUT_TYPE_LINEAR_DEPTH; }
hipError_t launch_dr_cull(const DRCullTables& T, const float2* d_minmax, float relax_eps, const uint4* d_pairs, uint32_t n_bricks,
                          uint32_t n_pairs, hipStream_t s) {
  const uint32_t n_slots = 2 * n_pairs;
  uint32_t* occ_b = T.occ + n_slots;
  const dim3 gs((n_slots + 255) / 256), gb((n_bricks + 255) / 256);
  hipLaunchKernelGGL(lrtd::k_dr_cull_clear, gs, dim3(256), 0, s, T.occ, occ_b, T.tiled, n_slots);
  hipLaunchKernelGGL(lrtd::k_dr_cull_mark, gb, dim3(256), 0, s, d_minmax, T.leaf_slot, T.pair_slot, T.occ, occ_b, relax_eps, n_bricks);
  hipLaunchKernelGGL(lrtd::k_dr_cull_write, gs, dim3(256), 0, s, d_pairs, (const uint32_t*)T.occ, (const uint32_t*)occ_b, T.pairs_primary,
                     T.pairs_border, n_slots);
  return hipGetLastError();
}
hipError_t launch_dr_primary(const lrtd::DRArgs& A, uint32_t* d_counter, uint32_t* d_spill, float* d_loss, uint32_t blocks,
                             hipStream_t s, const uint4* cull_pairs) {
  const bool mv = A.views != nullptr;
  if (dr_primary_is_depth(A)) {
    lrtd::DRArgs Ad = A;
    if (cull_pairs) Ad.S.pairs = cull_pairs;
    if (mv) hipLaunchKernelGGL((lrtd::k_dr_primary<lrtd::kDrDepth, true>), dim3(blocks), dim3(256), 0, s, Ad, d_counter, d_spill, d_loss);
    else hipLaunchKernelGGL((lrtd::k_dr_primary<lrtd::kDrDepth, false>), dim3(blocks), dim3(256), 0, s, Ad, d_counter, d_spill, d_loss);
  } else if (dr_normal_mode(A)) {
    if (mv) hipLaunchKernelGGL((lrtd::k_dr_primary<lrtd::kDrNormal, true>), dim3(blocks), dim3(256), 0, s, A, d_counter, d_spill, d_loss);
    else hipLaunchKernelGGL((lrtd::k_dr_primary<lrtd::kDrNormal, false>), dim3(blocks), dim3(256), 0, s, A, d_counter, d_spill, d_loss);
  } else {
    if (mv) hipLaunchKernelGGL((lrtd::k_dr_primary<lrtd::kDrGeneric, true>), dim3(blocks), dim3(256), 0, s, A, d_counter, d_spill, d_loss);
    else hipLaunchKernelGGL((lrtd::k_dr_primary<lrtd::kDrGeneric, false>), dim3(blocks), dim3(256), 0, s, A, d_counter, d_spill, d_loss);
  }
  return hipGetLastError();
}
hipError_t launch_dr_fd_set(float* d_values_f, float* d_bvals, const float* d_p0, const uint32_t* d_pos, uint32_t n_pos, uint32_t j,
                            uint32_t mode, float delta, uint32_t* d_counter, hipStream_t s) {
  hipLaunchKernelGGL(lrtd::k_dr_fd_set, dim3(1), dim3(256), 0, s, d_values_f, d_bvals, d_p0, d_pos, n_pos, j, mode, delta, d_counter);
  return hipGetLastError();
}
hipError_t launch_dr_fd_loss(const float4* d_img, const float4* d_ref, uint32_t n, uint32_t loss_fn, double* d_out, hipStream_t s) {
  hipLaunchKernelGGL(lrtd::k_dr_fd_loss, dim3(1), dim3(1024), 0, s, d_img, d_ref, n, loss_fn, d_out);
  return hipGetLastError();
}
hipError_t launch_dr_fd_grad(const double* d_acc, const uint32_t* d_ids, uint32_t n, float g_mult, float delta, float* d_grad, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(lrtd::k_dr_fd_grad, dim3((n + 255) / 256), dim3(256), 0, s, d_acc, d_ids, n, g_mult, delta, d_grad);
  return hipGetLastError();
}
hipError_t launch_dr_border_find(const lrtd::DRArgs& A, uint32_t* d_list, uint32_t* d_count, hipStream_t s) {
  uint32_t n = A.W * A.H;
  hipLaunchKernelGGL(lrtd::k_dr_border_find, dim3((n + 255) / 256), dim3(256), 0, s, A, d_list, d_count);
  return hipGetLastError();
}
hipError_t launch_dr_border_rays(const lrtd::DRArgs& A, const uint32_t* d_list, const uint32_t* d_count, uint32_t* d_spill,
                                 uint32_t* d_hits, uint32_t blocks, hipStream_t s, const uint4* cull_pairs) {
  // linear-depth input reads only depths and hit flags of border rays
  if (A.border_sampling == LRT_DR_BORDER_SAMPLING_SVM) {
    // one wave per border pixel, four per block; the spill rows of `blocks` blocks bound the grid
    if (dr_border_is_depth(A)) {
      lrtd::DRArgs Ab = A;
      if (cull_pairs) Ab.S.pairs = cull_pairs;
      hipLaunchKernelGGL(lrtd::k_dr_border_svm<lrtd::kDrDepthBorder>, dim3(blocks), dim3(256), 0, s, Ab, d_list, d_count, d_spill, d_hits);
    } else if (dr_normal_mode(A))
      hipLaunchKernelGGL(lrtd::k_dr_border_svm<lrtd::kDrNormal>, dim3(blocks), dim3(256), 0, s, A, d_list, d_count, d_spill, d_hits);
    else
      hipLaunchKernelGGL(lrtd::k_dr_border_svm<lrtd::kDrGeneric>, dim3(blocks), dim3(256), 0, s, A, d_list, d_count, d_spill, d_hits);
    return hipGetLastError();
  }
  if (dr_border_is_depth(A)) {
    lrtd::DRArgs Ab = A;
    if (cull_pairs) Ab.S.pairs = cull_pairs;
    hipLaunchKernelGGL(lrtd::k_dr_border_rays<lrtd::kDrDepthBorder>, dim3(blocks), dim3(256), 0, s, Ab, d_list, d_count, d_spill, d_hits);
  } else if (dr_normal_mode(A))
    hipLaunchKernelGGL(lrtd::k_dr_border_rays<lrtd::kDrNormal>, dim3(blocks), dim3(256), 0, s, A, d_list, d_count, d_spill, d_hits);
  else
    hipLaunchKernelGGL(lrtd::k_dr_border_rays<lrtd::kDrGeneric>, dim3(blocks), dim3(256), 0, s, A, d_list, d_count, d_spill, d_hits);
  return hipGetLastError();
}
hipError_t launch_dr_gather(const lrtd::DevScene& S, float* d_out, float2* d_minmax, hipStream_t s) {
  hipLaunchKernelGGL(lrtd::k_dr_gather, dim3((S.n_nodes + 3) / 4), dim3(256), 0, s, S, d_out, d_minmax);
  return hipGetLastError();
}
hipError_t launch_dr_regularize(const lrtd::DevScene& S, uint32_t n_nodes, float lambda, float power, float* d_grad, hipStream_t s) {
  const uint32_t vs = S.brick_size + 1u;
  const uint64_t threads = (uint64_t)n_nodes * vs * vs * vs;
  if (threads == 0) return hipSuccess;
  hipLaunchKernelGGL(lrtd::k_dr_regularize, dim3((uint32_t)((threads + 255) / 256)), dim3(256), 0, s, S, n_nodes, lambda, power, d_grad);
  return hipGetLastError();
}
hipError_t launch_dr_redistance(float* d_values, uint32_t p, float h, uint32_t iters, uint8_t* d_frozen, float* d_b0, float* d_b1,
                                hipStream_t s) {
  const uint32_t n = p * p * p;
  const dim3 grid((n + 255) / 256);
  hipLaunchKernelGGL(lrtd::k_redist_init, grid, dim3(256), 0, s, (const float*)d_values, p, d_frozen, d_b0, d_b1);
  float* b[2] = {d_b0, d_b1};
  int cur = 0;  // the buffer the next step writes
  for (uint32_t it = 0; it < iters; it++) {
    hipLaunchKernelGGL(lrtd::k_redist_step, grid, dim3(256), 0, s, (const float*)b[1 - cur], b[cur], (const uint8_t*)d_frozen, p, h);
    cur = 1 - cur;
  }
  hipError_t e = hipMemcpyAsync(d_values, b[1 - cur], (size_t)n * 4, hipMemcpyDeviceToDevice, s);
  if (e != hipSuccess) return e;
  return hipGetLastError();
}
hipError_t launch_dr_resample(const float4* d_src, uint32_t W0, uint32_t H0, float4* d_dst, uint32_t W, uint32_t H, uint32_t n_images,
                              hipStream_t s) {
  const uint64_t n = (uint64_t)W * H * n_images;
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(lrtd::k_dr_resample, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, s, d_src, W0, H0, d_dst, W, H, n_images);
  return hipGetLastError();
}
hipError_t launch_dr_scale(float* d_grad, uint32_t n, float mult, hipStream_t s) {
  hipLaunchKernelGGL(lrtd::k_dr_scale, dim3((n + 255) / 256), dim3(256), 0, s, d_grad, n, mult);
  return hipGetLastError();
}
hipError_t launch_adam(const float* g, float* X, float* V, float* S, uint32_t n, float lr, float b1, float b2, float eps, float c1,
                       float c2, float batch, hipStream_t s) {
  hipLaunchKernelGGL(lrtd::k_adam, dim3((n + 255) / 256), dim3(256), 0, s, g, X, V, S, n, lr, b1, b2, eps, c1, c2, batch);
  return hipGetLastError();
}

}  // namespace lrt

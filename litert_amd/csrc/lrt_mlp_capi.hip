// lrt_mlp_capi.hip - C ABI of the learned root-find seeder: parameters,
// BF16 packing for the MFMA kernels, training-event collection, training
// and inference entry points. There is no reference counterpart.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

#include "../../include/lrt.h"
#include "kernels/mlp.hpp"
#include "lrt_internal.hpp"

namespace lrt {

using namespace lrtd;

static uint16_t bf16_rne(float x) {
  uint32_t u;
  std::memcpy(&u, &x, 4);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

static uint64_t splitmix64(uint64_t& s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// Builds the A-tile / bias image the MFMA forward pass reads (layout in
// kernels/mlp.hpp) and uploads it with the FP32 parameters.
int Renderer::pack_mlp() {
  const std::vector<float>& p = mlp_params;
  std::vector<uint16_t> tiles((size_t)kTilesTotal * 64 * 8, 0);
  std::vector<float> bias((size_t)kBiasTiles * 64 * 4, 0.0f);
  auto w = [&](int off, int rows, int cols, int r, int c) { return (r < rows && c < cols) ? p[off + r * cols + c] : 0.0f; };
  auto hidden_col = [](int kt, int l, int j) { return 32 * kt + 16 * (j / 4) + 4 * (l / 16) + (j % 4); };
  for (int mt = 0; mt < 4; mt++)
    for (int l = 0; l < 64; l++)
      for (int j = 0; j < 8; j++)
        tiles[((size_t)(kTileOffL1 + mt) * 64 + l) * 8 + j] = bf16_rne(w(kOffW1, kMlpH, kMlpIn, 16 * mt + l % 16, 8 * (l / 16) + j));
  const int offs[2] = {kOffW2, kOffW3}, toffs[2] = {kTileOffL2, kTileOffL3};
  for (int L = 0; L < 2; L++)
    for (int mt = 0; mt < 4; mt++)
      for (int kt = 0; kt < 2; kt++)
        for (int l = 0; l < 64; l++)
          for (int j = 0; j < 8; j++)
            tiles[((size_t)(toffs[L] + 2 * mt + kt) * 64 + l) * 8 + j] =
                bf16_rne(w(offs[L], kMlpH, kMlpH, 16 * mt + l % 16, hidden_col(kt, l, j)));
  for (int kt = 0; kt < 2; kt++)
    for (int l = 0; l < 64; l++)
      for (int j = 0; j < 8; j++)
        tiles[((size_t)(kTileOffL4 + kt) * 64 + l) * 8 + j] = bf16_rne(w(kOffW4, kMlpOut, kMlpH, l % 16, hidden_col(kt, l, j)));
  const int boffs[3] = {kOffB1, kOffB2, kOffB3};
  for (int bt = 0; bt < kBiasTiles; bt++) {
    const int layer = bt < 12 ? bt / 4 : 3, mt = bt < 12 ? bt % 4 : 0;
    for (int l = 0; l < 64; l++)
      for (int i = 0; i < 4; i++) {
        const int r = 16 * mt + 4 * (l / 16) + i;
        bias[((size_t)bt * 64 + l) * 4 + i] = layer < 3 ? p[boffs[layer] + r] : (r < kMlpOut ? p[kOffB4 + r] : 0.0f);
      }
  }
  std::vector<uint8_t> img(kPackedBytes);
  std::memcpy(img.data(), tiles.data(), tiles.size() * 2);
  std::memcpy(img.data() + tiles.size() * 2, bias.data(), bias.size() * 4);
  hipError_t e;
  if ((e = hipSetDevice(device)) != hipSuccess) return hip_err(e, "hipSetDevice");
  if ((e = d_mlp_packed.upload(img.data(), img.size(), stream)) != hipSuccess) return hip_err(e, "upload packed weights");
  if ((e = d_mlp_params.upload(p.data(), p.size() * 4, stream)) != hipSuccess) return hip_err(e, "upload params");
  if ((e = hipStreamSynchronize(stream)) != hipSuccess) return hip_err(e, "upload params");
  mlp_ready = true;
  return LRT_OK;
}

hipError_t Renderer::ensure_brick_features() {
  if (bfeat_valid) return hipSuccess;
  hipError_t e = d_bfeat.ensure(nodes.size() * kBrickFeat * sizeof(float));
  if (e != hipSuccess) return e;
  if ((e = launch_brick_features(dev_scene(), (float4*)d_bfeat.p, stream)) != hipSuccess) return e;
  bfeat_valid = true;
  return hipSuccess;
}

// Grows a device buffer, keeping its first keep_bytes.
static hipError_t grow_keep(DevBuf& b, size_t keep_bytes, size_t new_bytes, hipStream_t s) {
  if (b.p && new_bytes <= b.bytes) return hipSuccess;
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, new_bytes);
  if (e != hipSuccess) return e;
  if (keep_bytes && b.p) {
    if ((e = hipMemcpyAsync(p, b.p, keep_bytes, hipMemcpyDeviceToDevice, s)) != hipSuccess) return e;
    if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
  }
  b.release();
  b.p = p;
  b.bytes = new_bytes;
  return hipSuccess;
}

static int reset_adam(Renderer* r) {
  hipError_t e;
  if ((e = r->d_mlp_m.ensure(kMlpParams * 4)) != hipSuccess) return hip_err(e, "alloc adam");
  if ((e = r->d_mlp_v.ensure(kMlpParams * 4)) != hipSuccess) return hip_err(e, "alloc adam");
  if ((e = zero_u32((uint32_t*)r->d_mlp_m.p, kMlpParams, r->stream)) != hipSuccess) return hip_err(e, "clear adam");
  if ((e = zero_u32((uint32_t*)r->d_mlp_v.p, kMlpParams, r->stream)) != hipSuccess) return hip_err(e, "clear adam");
  r->mlp_adam_t = 0;
  return LRT_OK;
}

}  // namespace lrt

using namespace lrt;

extern "C" {

int lrt_mlp_init(lrt_handle h, uint32_t seed) {
  if (!h) return set_err(LRT_E_INVALID, "null handle");
  std::vector<float> p(kMlpParams, 0.0f);
  uint64_t s = 0x5EEDull ^ ((uint64_t)seed << 32);
  auto fill = [&](int off, int rows, int cols) {
    const double lim = std::sqrt(6.0 / (double)(rows + cols));
    for (int k = 0; k < rows * cols; k++) {
      const double u = (double)(splitmix64(s) >> 11) * (1.0 / 9007199254740992.0);
      p[off + k] = (float)((2.0 * u - 1.0) * lim);
    }
  };
  fill(kOffW1, kMlpH, kMlpIn);
  fill(kOffW2, kMlpH, kMlpH);
  fill(kOffW3, kMlpH, kMlpH);
  fill(kOffW4, kMlpOut, kMlpH);
  h->mlp_params = p;
  int c = reset_adam(h);
  return c ? c : h->pack_mlp();
}

int lrt_mlp_set_params(lrt_handle h, const float* params, uint32_t n) {
  if (!h || !params) return set_err(LRT_E_INVALID, "null argument");
  if (n != (uint32_t)kMlpParams) return set_err(LRT_E_INVALID, "expected LRT_MLP_PARAMS floats");
  for (uint32_t i = 0; i < n; i++)
    if (!std::isfinite(params[i])) return set_err(LRT_E_INVALID, "non-finite parameter");
  h->mlp_params.assign(params, params + n);
  int c = reset_adam(h);
  return c ? c : h->pack_mlp();
}

int lrt_mlp_get_params(lrt_handle h, float* params, uint32_t n) {
  if (!h || !params) return set_err(LRT_E_INVALID, "null argument");
  if (!h->mlp_ready) return set_err(LRT_E_INVALID, "no parameters: call lrt_mlp_init or lrt_mlp_set_params");
  if (n != (uint32_t)kMlpParams) return set_err(LRT_E_INVALID, "expected LRT_MLP_PARAMS floats");
  std::memcpy(params, h->mlp_params.data(), (size_t)n * 4);
  return LRT_OK;
}

int lrt_mlp_collect(lrt_handle h, const float* wvs, const float* prs, uint32_t n_views, uint32_t W, uint32_t H, uint64_t max_events,
                    uint64_t* n_added, uint64_t* n_seen) {
  if (!h) return set_err(LRT_E_INVALID, "null handle");
  if (!h->has_scene) return set_err(LRT_E_NOSCENE, "no scene: call lrt_set_scene_sbs first");
  if (!wvs || !prs || n_views == 0 || W == 0 || H == 0 || W > 65535 || H > 65535 || max_events == 0)
    return set_err(LRT_E_INVALID, "bad arguments");
  if ((uint64_t)(8 * ((W + 7) / 8)) * (uint64_t)(8 * ((H + 7) / 8)) > 0xFFFFFFFFull)
    return set_err(LRT_E_INVALID, "image too large: the 8-aligned pixel count must fit 32 bits");
  if ((uint64_t)n_views * W * H >= (1ull << 32)) return set_err(LRT_E_INVALID, "too many rays for 32-bit event ids");
  Renderer* r = h;
  hipError_t e;
  if ((e = hipSetDevice(r->device)) != hipSuccess) return hip_err(e, "hipSetDevice");
  std::vector<lrtd::Camera> cams(n_views);
  for (uint32_t v = 0; v < n_views; v++) cams[v] = make_camera(wvs + 16 * v, prs + 16 * v);
  if ((e = r->d_cams.upload(cams.data(), sizeof(lrtd::Camera) * n_views, r->stream)) != hipSuccess) return hip_err(e, "upload cameras");
  if ((e = r->ensure_brick_features()) != hipSuccess) return hip_err(e, "brick features");
  const uint64_t cap = r->ds_n + max_events;
  if ((e = grow_keep(r->d_ds_feat, r->ds_n * 24 * 4, cap * 24 * 4, r->stream)) != hipSuccess) return hip_err(e, "alloc dataset");
  if ((e = grow_keep(r->d_ds_tgt, r->ds_n * 3 * 4, cap * 3 * 4, r->stream)) != hipSuccess) return hip_err(e, "alloc dataset");
  if ((e = grow_keep(r->d_ds_meta, r->ds_n * 8, cap * 8, r->stream)) != hipSuccess) return hip_err(e, "alloc dataset");
  if ((e = r->d_counts.ensure(64)) != hipSuccess) return hip_err(e, "alloc counts");
  if ((e = zero_u32((uint32_t*)r->d_counts.p, 16, r->stream)) != hipSuccess) return hip_err(e, "clear counts");
  lrtd::RenderArgs A;
  lrtd::set_image(A, W, H);
  A.n_views = n_views;
  A.spp = 1;
  A.ray_gen_mode = r->preset.ray_gen_mode;
  A.seed = r->seed;
  A.lights = r->light_set();
  const uint32_t blocks = r->grid_blocks();
  if ((e = r->ensure_spill(blocks)) != hipSuccess) return hip_err(e, "alloc spill");
  e = launch_mlp_collect(r->dev_scene(), r->trace_params(false), (const float4*)r->d_bfeat.p, (const lrtd::Camera*)r->d_cams.p, A,
                         (float*)r->d_ds_feat.p + r->ds_n * 24, (float*)r->d_ds_tgt.p + r->ds_n * 3,
                         (lrtd::MlpEventMeta*)r->d_ds_meta.p + r->ds_n, max_events, (unsigned long long*)r->d_counts.p,
                         (uint32_t*)r->d_spill.p, blocks, r->stream);
  if (e != hipSuccess) return hip_err(e, "collect launch");
  unsigned long long cnt = 0;
  if ((e = hipMemcpyAsync(&cnt, r->d_counts.p, 8, hipMemcpyDeviceToHost, r->stream)) != hipSuccess) return hip_err(e, "collect");
  if ((e = hipStreamSynchronize(r->stream)) != hipSuccess) return hip_err(e, "collect");
  const uint64_t added = std::min<uint64_t>(cnt, max_events);
  r->ds_n += added;
  if (n_added) *n_added = added;
  if (n_seen) *n_seen = cnt;
  return LRT_OK;
}

int lrt_mlp_dataset_size(lrt_handle h, uint64_t* n) {
  if (!h || !n) return set_err(LRT_E_INVALID, "null argument");
  *n = h->ds_n;
  return LRT_OK;
}

int lrt_mlp_dataset_read(lrt_handle h, uint64_t first, uint64_t count, float* feats, float* tgts, uint32_t* meta) {
  if (!h) return set_err(LRT_E_INVALID, "null handle");
  if (first > h->ds_n || count > h->ds_n - first) return set_err(LRT_E_INVALID, "range outside the dataset");
  if (count == 0) return LRT_OK;
  hipError_t e;
  if ((e = hipSetDevice(h->device)) != hipSuccess) return hip_err(e, "hipSetDevice");
  if (feats && (e = hipMemcpy(feats, (float*)h->d_ds_feat.p + first * 24, count * 24 * 4, hipMemcpyDeviceToHost)) != hipSuccess)
    return hip_err(e, "read dataset");
  if (tgts && (e = hipMemcpy(tgts, (float*)h->d_ds_tgt.p + first * 3, count * 3 * 4, hipMemcpyDeviceToHost)) != hipSuccess)
    return hip_err(e, "read dataset");
  if (meta && (e = hipMemcpy(meta, (uint32_t*)h->d_ds_meta.p + first * 2, count * 8, hipMemcpyDeviceToHost)) != hipSuccess)
    return hip_err(e, "read dataset");
  return LRT_OK;
}

int lrt_mlp_dataset_clear(lrt_handle h) {
  if (!h) return set_err(LRT_E_INVALID, "null handle");
  h->ds_n = 0;
  return LRT_OK;
}

int lrt_mlp_train(lrt_handle h, uint32_t epochs, uint32_t batch, float lr, uint32_t seed, float* epoch_losses, float* total_ms) {
  if (!h) return set_err(LRT_E_INVALID, "null handle");
  if (!h->mlp_ready) return set_err(LRT_E_INVALID, "no parameters: call lrt_mlp_init or lrt_mlp_set_params");
  if (h->ds_n == 0) return set_err(LRT_E_INVALID, "empty dataset: call lrt_mlp_collect first");
  if (batch < 64 || batch % 64 != 0 || batch > 65536) return set_err(LRT_E_INVALID, "batch must be a multiple of 64 in [64, 65536]");
  if (!(lr > 0.0f) || epochs == 0) return set_err(LRT_E_INVALID, "bad learning rate or epoch count");
  Renderer* r = h;
  hipError_t e;
  if ((e = hipSetDevice(r->device)) != hipSuccess) return hip_err(e, "hipSetDevice");
  const uint64_t n = r->ds_n;
  const uint32_t max_blocks = batch / 64;
  const uint64_t steps = (n + batch - 1) / batch;
  if ((e = r->d_mlp_partial.ensure((size_t)max_blocks * kMlpPartialStride * 4)) != hipSuccess) return hip_err(e, "alloc partials");
  if ((e = r->d_mlp_ploss.ensure((size_t)max_blocks * 4)) != hipSuccess) return hip_err(e, "alloc partials");
  if ((e = r->d_mlp_loss.ensure(steps * 4)) != hipSuccess) return hip_err(e, "alloc losses");
  std::vector<float> step_loss(steps);
  uint64_t rs = 0xC0FFEEull ^ ((uint64_t)seed << 20);
  AdamArgs adam;
  adam.lr = lr;
  adam.b1 = 0.9f;
  adam.b2 = 0.999f;
  adam.eps = 1e-8f;
  if ((e = hipEventRecord(r->ev0, r->stream)) != hipSuccess) return hip_err(e, "event");
  for (uint32_t ep = 0; ep < epochs; ep++) {
    MlpTrainArgs T;
    T.feats = (const float*)r->d_ds_feat.p;
    T.tgts = (const float*)r->d_ds_tgt.p;
    T.n = n;
    uint64_t a = n > 1 ? (splitmix64(rs) % n) | 1ull : 1ull;
    while (std::gcd(a, n) != 1) a += 2;
    T.perm_a = a % n == 0 ? 1 : a;
    T.perm_b = splitmix64(rs) % n;
    for (uint64_t st = 0; st < steps; st++) {
      T.first = st * batch;
      T.batch = (uint32_t)std::min<uint64_t>(batch, n - T.first);
      const uint32_t blocks = (T.batch + 63) / 64;
      if ((e = launch_mlp_grad((const float*)r->d_mlp_params.p, T, (float*)r->d_mlp_partial.p, (float*)r->d_mlp_ploss.p, blocks,
                               r->stream)) != hipSuccess)
        return hip_err(e, "train launch");
      r->mlp_adam_t++;
      adam.c1 = (float)(1.0 / (1.0 - std::pow(0.9, (double)r->mlp_adam_t)));
      adam.c2 = (float)(1.0 / (1.0 - std::pow(0.999, (double)r->mlp_adam_t)));
      if ((e = launch_mlp_reduce_adam((float*)r->d_mlp_params.p, (float*)r->d_mlp_m.p, (float*)r->d_mlp_v.p, (const float*)r->d_mlp_partial.p,
                                      (const float*)r->d_mlp_ploss.p, blocks, T.batch, adam, true, nullptr, (float*)r->d_mlp_loss.p + st,
                                      r->stream)) != hipSuccess)
        return hip_err(e, "adam launch");
    }
    if ((e = hipMemcpyAsync(step_loss.data(), r->d_mlp_loss.p, steps * 4, hipMemcpyDeviceToHost, r->stream)) != hipSuccess)
      return hip_err(e, "download losses");
    if ((e = hipStreamSynchronize(r->stream)) != hipSuccess) return hip_err(e, "training");
    double acc = 0.0;
    for (float v : step_loss) acc += v;
    if (epoch_losses) epoch_losses[ep] = (float)(acc / (double)steps);
  }
  if ((e = hipEventRecord(r->ev1, r->stream)) != hipSuccess) return hip_err(e, "event");
  if ((e = hipEventSynchronize(r->ev1)) != hipSuccess) return hip_err(e, "training");
  float ms = 0.0f;
  (void)hipEventElapsedTime(&ms, r->ev0, r->ev1);
  if (total_ms) *total_ms = ms;
  if ((e = hipMemcpy(r->mlp_params.data(), r->d_mlp_params.p, kMlpParams * 4, hipMemcpyDeviceToHost)) != hipSuccess)
    return hip_err(e, "download params");
  return r->pack_mlp();
}

int lrt_mlp_loss_grad(lrt_handle h, const float* feats, const float* tgts, uint32_t n, float* grad, float* loss) {
  if (!h || !feats || !tgts) return set_err(LRT_E_INVALID, "null argument");
  if (!h->mlp_ready) return set_err(LRT_E_INVALID, "no parameters: call lrt_mlp_init or lrt_mlp_set_params");
  if (n == 0 || n > 65536) return set_err(LRT_E_INVALID, "n must be in [1, 65536]");
  Renderer* r = h;
  hipError_t e;
  if ((e = hipSetDevice(r->device)) != hipSuccess) return hip_err(e, "hipSetDevice");
  DevBuf df, dt;
  const uint32_t blocks = (n + 63) / 64;
  int rc = LRT_OK;
  do {
    if ((e = df.upload(feats, (size_t)n * 24 * 4, r->stream)) != hipSuccess) { rc = hip_err(e, "upload features"); break; }
    if ((e = dt.upload(tgts, (size_t)n * 3 * 4, r->stream)) != hipSuccess) { rc = hip_err(e, "upload targets"); break; }
    if ((e = r->d_mlp_partial.ensure((size_t)blocks * kMlpPartialStride * 4)) != hipSuccess) { rc = hip_err(e, "alloc partials"); break; }
    if ((e = r->d_mlp_ploss.ensure((size_t)blocks * 4)) != hipSuccess) { rc = hip_err(e, "alloc partials"); break; }
    if ((e = r->d_mlp_grad.ensure(kMlpParams * 4)) != hipSuccess) { rc = hip_err(e, "alloc grad"); break; }
    if ((e = r->d_mlp_loss.ensure(16)) != hipSuccess) { rc = hip_err(e, "alloc loss"); break; }
    MlpTrainArgs T;
    T.feats = (const float*)df.p;
    T.tgts = (const float*)dt.p;
    T.n = n;
    T.perm_a = 1;
    T.perm_b = 0;
    T.first = 0;
    T.batch = n;
    if ((e = launch_mlp_grad((const float*)r->d_mlp_params.p, T, (float*)r->d_mlp_partial.p, (float*)r->d_mlp_ploss.p, blocks, r->stream)) !=
        hipSuccess) { rc = hip_err(e, "grad launch"); break; }
    AdamArgs adam{};
    if ((e = launch_mlp_reduce_adam((float*)r->d_mlp_params.p, nullptr, nullptr, (const float*)r->d_mlp_partial.p,
                                    (const float*)r->d_mlp_ploss.p, blocks, n, adam, false, (float*)r->d_mlp_grad.p,
                                    (float*)r->d_mlp_loss.p, r->stream)) != hipSuccess) { rc = hip_err(e, "reduce launch"); break; }
    if (grad && (e = hipMemcpyAsync(grad, r->d_mlp_grad.p, kMlpParams * 4, hipMemcpyDeviceToHost, r->stream)) != hipSuccess) {
      rc = hip_err(e, "download grad"); break;
    }
    float l = 0.0f;
    if ((e = hipMemcpyAsync(&l, r->d_mlp_loss.p, 4, hipMemcpyDeviceToHost, r->stream)) != hipSuccess) { rc = hip_err(e, "download loss"); break; }
    if ((e = hipStreamSynchronize(r->stream)) != hipSuccess) { rc = hip_err(e, "loss/grad"); break; }
    if (loss) *loss = l;
  } while (false);
  (void)hipStreamSynchronize(r->stream);
  df.release();
  dt.release();
  return rc;
}

int lrt_mlp_infer(lrt_handle h, const float* feats, uint64_t n, float* out, float* kernel_ms) {
  if (!h || !feats || !out) return set_err(LRT_E_INVALID, "null argument");
  if (!h->mlp_ready) return set_err(LRT_E_INVALID, "no parameters: call lrt_mlp_init or lrt_mlp_set_params");
  if (n == 0 || n > (1ull << 28)) return set_err(LRT_E_INVALID, "n must be in [1, 2^28]");
  Renderer* r = h;
  hipError_t e;
  if ((e = hipSetDevice(r->device)) != hipSuccess) return hip_err(e, "hipSetDevice");
  DevBuf df, dout;
  int rc = LRT_OK;
  do {
    if ((e = df.upload(feats, n * 24 * 4, r->stream)) != hipSuccess) { rc = hip_err(e, "upload features"); break; }
    if ((e = dout.ensure(n * 3 * 4)) != hipSuccess) { rc = hip_err(e, "alloc out"); break; }
    const uint32_t blocks = (uint32_t)std::min<uint64_t>((n + 255) / 256, (uint64_t)r->cu_count * 8);
    if ((e = hipEventRecord(r->ev0, r->stream)) != hipSuccess) { rc = hip_err(e, "event"); break; }
    if ((e = launch_mlp_infer(r->d_mlp_packed.p, (const float*)df.p, n, (float*)dout.p, blocks, r->stream)) != hipSuccess) {
      rc = hip_err(e, "infer launch"); break;
    }
    if ((e = hipEventRecord(r->ev1, r->stream)) != hipSuccess) { rc = hip_err(e, "event"); break; }
    if ((e = hipMemcpyAsync(out, dout.p, n * 3 * 4, hipMemcpyDeviceToHost, r->stream)) != hipSuccess) { rc = hip_err(e, "download"); break; }
    if ((e = hipStreamSynchronize(r->stream)) != hipSuccess) { rc = hip_err(e, "inference"); break; }
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, r->ev0, r->ev1);
    if (kernel_ms) *kernel_ms = ms;
  } while (false);
  (void)hipStreamSynchronize(r->stream);
  df.release();
  dout.release();
  return rc;
}

int lrt_mlp_infer_dataset(lrt_handle h, uint32_t repeats, float* kernel_ms, float* out, uint64_t n_out) {
  if (!h) return set_err(LRT_E_INVALID, "null handle");
  if (!h->mlp_ready) return set_err(LRT_E_INVALID, "no parameters: call lrt_mlp_init or lrt_mlp_set_params");
  if (h->ds_n == 0) return set_err(LRT_E_INVALID, "empty dataset");
  if (repeats == 0 || n_out > h->ds_n) return set_err(LRT_E_INVALID, "bad arguments");
  Renderer* r = h;
  hipError_t e;
  if ((e = hipSetDevice(r->device)) != hipSuccess) return hip_err(e, "hipSetDevice");
  const uint64_t n = r->ds_n;
  if ((e = r->d_ds_out.ensure(n * 3 * 4)) != hipSuccess) return hip_err(e, "alloc out");
  const uint32_t blocks = (uint32_t)std::min<uint64_t>((n + 255) / 256, (uint64_t)r->cu_count * 8);
  if ((e = launch_mlp_infer(r->d_mlp_packed.p, (const float*)r->d_ds_feat.p, n, (float*)r->d_ds_out.p, blocks, r->stream)) != hipSuccess)
    return hip_err(e, "infer launch");  // warm-up
  if ((e = hipEventRecord(r->ev0, r->stream)) != hipSuccess) return hip_err(e, "event");
  for (uint32_t k = 0; k < repeats; k++)
    if ((e = launch_mlp_infer(r->d_mlp_packed.p, (const float*)r->d_ds_feat.p, n, (float*)r->d_ds_out.p, blocks, r->stream)) != hipSuccess)
      return hip_err(e, "infer launch");
  if ((e = hipEventRecord(r->ev1, r->stream)) != hipSuccess) return hip_err(e, "event");
  if ((e = hipEventSynchronize(r->ev1)) != hipSuccess) return hip_err(e, "inference");
  float ms = 0.0f;
  (void)hipEventElapsedTime(&ms, r->ev0, r->ev1);
  if (kernel_ms) *kernel_ms = ms / (float)repeats;
  if (out && n_out && (e = hipMemcpy(out, r->d_ds_out.p, n_out * 3 * 4, hipMemcpyDeviceToHost)) != hipSuccess)
    return hip_err(e, "download");
  return LRT_OK;
}

int lrt_set_seeding(lrt_handle h, int enabled, float threshold) {
  if (!h) return set_err(LRT_E_INVALID, "null handle");
  if (!(threshold > 0.0f && threshold < 1.0f)) return set_err(LRT_E_INVALID, "threshold must lie in (0, 1)");
  if (enabled && !h->mlp_ready) return set_err(LRT_E_INVALID, "no parameters: call lrt_mlp_init or lrt_mlp_set_params");
  h->seeding = enabled != 0;
  h->seed_threshold = threshold;
  return LRT_OK;
}

}  // extern "C"

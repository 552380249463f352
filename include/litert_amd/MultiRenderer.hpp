// MultiRenderer.hpp - header-only C++ facade over liblrt_hip.so with the
// reference renderer's names and argument meaning:
//   CreateMultiRenderer(device)          Renderer/eye_ray.h:252
//   MultiRenderer::SetScene(SdfSBSView)  eye_ray_host.cpp:360-375
//   SetPreset / SetViewport / UpdateCamera / setSeed
//   Render / RenderFloat (both overloads) eye_ray_host.cpp:246-282, 563-585
//   GetExecutionTime                      eye_ray_host.cpp:289-295
//   RayQuery_NearestHit (via GetAccelStruct()) BVH/BVH2Common.h:167
//   dr::MultiRendererDR                   diff_render/MultiRendererDR.h:71-155
// Matrices are any 16-float column-major type (LiteMath::float4x4 included);
// images are caller-owned y*W+x arrays. Errors: the reference prints and
// returns; here every call returns false and lastError() gives the reason.
#pragma once
#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "../lrt.h"

namespace lrt_cpp {

// Non-owning view of an SBS, field for field sdfScene/sdf_scene.h:296-340.
struct SdfSBSView {
  SdfSBSHeader header{};
  unsigned size = 0;
  const SdfSBSNode* nodes = nullptr;
  unsigned values_count = 0;
  const uint32_t* values = nullptr;
  unsigned values_f_count = 0;
  const float* values_f = nullptr;
};

struct SdfSBS {  // sdf_scene.h:226-232
  SdfSBSHeader header{};
  std::vector<SdfSBSNode> nodes;
  std::vector<uint32_t> values;
  std::vector<float> values_f;
  SdfSBSView view() const {
    SdfSBSView v;
    v.header = header;
    v.size = (unsigned)nodes.size();
    v.nodes = nodes.data();
    v.values_count = (unsigned)values.size();
    v.values = values.data();
    v.values_f_count = (unsigned)values_f.size();
    v.values_f = values_f.empty() ? nullptr : values_f.data();
    return v;
  }
};

inline MultiRenderPreset getDefaultPreset() {  // render_settings.h:94-108
  return MultiRenderPreset{LRT_RENDER_MODE_LAMBERT_NO_TEX, LRT_SDF_NODE_INTERSECT_ST, LRT_NORMAL_MODE_GEOMETRY, LRT_RAY_GEN_MODE_REGULAR,
                           LRT_INTERPOLATION_MODE_TRILINEAR, LRT_REPRESENTATION_MODE_VOLUME, 1};
}

template <class M> inline const float* mat_ptr(const M& m) {
  static_assert(sizeof(M) == 16 * sizeof(float), "expects a 4x4 float matrix");
  return reinterpret_cast<const float*>(&m);
}

class MultiRenderer;

// ISceneObject-like accessor for RayQuery_NearestHit.
class AccelStruct {
 public:
  explicit AccelStruct(lrt_handle h) : h_(h) {}
  template <class V4> CRT_Hit RayQuery_NearestHit(const V4& posAndNear, const V4& dirAndFar) {
    static_assert(sizeof(V4) == 4 * sizeof(float), "expects float4");
    float r[8];
    const float* a = reinterpret_cast<const float*>(&posAndNear);
    const float* b = reinterpret_cast<const float*>(&dirAndFar);
    for (int i = 0; i < 4; i++) { r[i] = a[i]; r[4 + i] = b[i]; }
    CRT_Hit hit{};
    lrt_cast_rays(h_, r, 1, &hit);
    return hit;
  }
  // batched form: rays = n * (posAndNear, dirAndFar)
  bool RayQuery_NearestHit(const float* rays, uint32_t n, CRT_Hit* out) { return lrt_cast_rays(h_, rays, n, out) == LRT_OK; }

 private:
  lrt_handle h_;
};

class MultiRenderer {
 public:
  explicit MultiRenderer(uint32_t device = LRT_DEVICE_GPU, uint32_t maxPrimitives = 10000000) {
    if (lrt_create(device, maxPrimitives, &h_) != LRT_OK) h_ = nullptr;
    preset_ = getDefaultPreset();
  }
  virtual ~MultiRenderer() { lrt_destroy(h_); }
  MultiRenderer(const MultiRenderer&) = delete;
  MultiRenderer& operator=(const MultiRenderer&) = delete;

  bool valid() const { return h_ != nullptr; }
  static std::string lastError() { return lrt_last_error(); }
  const char* Name() const { return "BVH2Common"; }

  bool SetScene(const SdfSBSView& s) {
    return ok(lrt_set_preset(h_, &preset_)) &&
           ok(lrt_set_scene_sbs(h_, &s.header, s.nodes, s.size, s.values, s.values_count, s.values_f, s.values_f_count));
  }
  bool SetScene(const SdfSBS& s) { return SetScene(s.view()); }
  // LoadScene(path) / LoadSceneHydra, eye_ray_host.cpp:52-121 (one identity sdf_sbs instance)
  bool LoadScene(const char* path) { return ok(lrt_load_scene_hydra(h_, path)); }
  bool SetPreset(const MultiRenderPreset& p) {
    preset_ = p;
    return ok(lrt_set_preset(h_, &p));
  }
  MultiRenderPreset GetPreset() const { return preset_; }
  bool SetViewport(int x0, int y0, int w, int h) {
    width_ = (uint32_t)w;
    height_ = (uint32_t)h;
    return ok(lrt_set_viewport(h_, x0, y0, w, h));
  }
  template <class M> bool UpdateCamera(const M& worldView, const M& proj) { return ok(lrt_update_camera(h_, mat_ptr(worldView), mat_ptr(proj))); }
  void setSeed(uint32_t seed) { lrt_set_seed(h_, seed); }

  // Render(imageData, w, h, worldView, proj, preset, passNum), eye_ray_host.cpp:563-573
  template <class M>
  bool Render(uint32_t* image, uint32_t w, uint32_t h, const M& worldView, const M& proj, MultiRenderPreset preset = getDefaultPreset(),
              int passNum = 1) {
    return SetViewport(0, 0, (int)w, (int)h) && UpdateCamera(worldView, proj) && SetPreset(preset) && Render(image, w, h, "color", passNum);
  }
  bool Render(uint32_t* image, uint32_t w, uint32_t h, const char* /*what*/, int passNum = 1) {
    return ok(lrt_render(h_, image, w, h, passNum));
  }
  // RenderFloat(float4* imageData, ...), eye_ray_host.cpp:575-585
  template <class M, class F4>
  bool RenderFloat(F4* image, uint32_t w, uint32_t h, const M& worldView, const M& proj, MultiRenderPreset preset = getDefaultPreset(),
                   int passNum = 1) {
    static_assert(sizeof(F4) == 4 * sizeof(float), "expects float4 pixels");
    return SetViewport(0, 0, (int)w, (int)h) && UpdateCamera(worldView, proj) && SetPreset(preset) &&
           ok(lrt_render_float(h_, reinterpret_cast<float*>(image), w, h, passNum));
  }
  // CastRaySingleBlock / CastRayFloatSingleBlock (eye_ray.h:185-188), the
  // seam the reference's GPU subclasses override: every packed pixel of the
  // viewport (tidX = the packed pixel count, as Render passes it). A block of
  // fewer pixels is refused; the per-pixel CastRaySingle stays on the
  // reference's CPU path and has no form here.
  bool CastRaySingleBlock(uint32_t tidX, uint32_t* out_color, uint32_t a_numPasses = 1) {
    return tidX >= packed_count() && ok(lrt_render(h_, out_color, width_, height_, (int)a_numPasses));
  }
  template <class F4>
  bool CastRayFloatSingleBlock(uint32_t tidX, F4* out_color, uint32_t a_numPasses = 1) {
    static_assert(sizeof(F4) == 4 * sizeof(float), "expects float4 pixels");
    return tidX >= packed_count() && ok(lrt_render_float(h_, reinterpret_cast<float*>(out_color), width_, height_, (int)a_numPasses));
  }
  // the packed pixel count of the viewport: both sides rounded up to multiples of 8 (SetViewport, eye_ray_host.cpp:43-50)
  uint32_t packed_count() const { return ((width_ + 7) / 8 * 8) * ((height_ + 7) / 8 * 8); }
  void GetExecutionTime(const char* name, float out[4]) { lrt_get_execution_time(h_, name, out); }
  std::shared_ptr<AccelStruct> GetAccelStruct() { return std::make_shared<AccelStruct>(h_); }
  lrt_handle handle() const { return h_; }

 protected:
  static bool ok(int rc) { return rc == LRT_OK; }
  lrt_handle h_ = nullptr;
  MultiRenderPreset preset_{};
  uint32_t width_ = 0, height_ = 0;
};

inline std::shared_ptr<MultiRenderer> CreateMultiRenderer(uint32_t device, uint32_t maxPrimitives = 10000000) {
  auto r = std::make_shared<MultiRenderer>(device, maxPrimitives);
  return r->valid() ? r : nullptr;
}

namespace dr {

inline MultiRendererDRPreset getDefaultPresetDR() {  // diff_render/MultiRendererDR.h:14-69
  MultiRendererDRPreset p{};
  p.spp = 1;
  p.dr_reconstruction_flags = LRT_DR_RECONSTRUCTION_FLAG_COLOR;
  p.dr_atomic_ders = 0xFFFFFFFFu;
  p.dr_raycasting_thickness = 5;
  p.border_spp = 256;
  p.border_relax_eps = 1e-3f;
  p.border_integral_mult = 1.0f;
  p.opt_lr = 0.01f;
  p.opt_beta_1 = 0.9f;
  p.opt_beta_2 = 0.999f;
  p.opt_eps = 1e-8f;
  p.opt_iterations = 500;
  p.image_batch_size = 1;
  p.reg_lambda = 0.25f;
  p.reg_power = 2.0f;
  p.redistancing_interval = 1;
  p.debug_print_interval = 10;
  p.debug_progress_images = 1001;
  p.debug_progress_interval = 100;
  p.finite_diff_delta = 0.001f;
  p.finite_diff_brightness = 100.0f;
  p.debug_pd_brightness = 0.1f;
  return p;
}

class MultiRendererDR : public MultiRenderer {
 public:
  explicit MultiRendererDR(uint32_t maxPrimitives = 10000000) : MultiRenderer(LRT_DEVICE_GPU, maxPrimitives) {}
  // SetReference(images, worldView, proj): images as n * W * H float4
  template <class M>
  bool SetReference(const float* images, uint32_t n, uint32_t w, uint32_t h, const std::vector<M>& worldView, const std::vector<M>& proj) {
    std::vector<float> wv(16 * n), pr(16 * n);
    for (uint32_t i = 0; i < n; i++)
      for (int k = 0; k < 16; k++) { wv[16 * i + k] = mat_ptr(worldView[i])[k]; pr[16 * i + k] = mat_ptr(proj[i])[k]; }
    w_ = w;
    h_img_ = h;
    return ok(lrt_dr_set_reference(h_, images, nullptr, n, w, h, wv.data(), pr.data()));
  }
  // OptimizeFixedStructure(preset, sbs): sbs.values_f receives the result
  bool OptimizeFixedStructure(MultiRendererDRPreset preset, SdfSBS& sbs, uint32_t seed = 0) {
    last_grad_.resize(sbs.values_f.size());
    losses_.assign(preset.opt_iterations, 0.0f);
    bool r = ok(lrt_dr_optimize(h_, &preset, &sbs.header, sbs.nodes.data(), (uint32_t)sbs.nodes.size(), sbs.values.data(), sbs.values.size(),
                                sbs.values_f.data(), sbs.values_f.size(), seed, 0, losses_.data(), nullptr));
    if (r) lrt_dr_get_last_dloss_ds(h_, last_grad_.data(), last_grad_.size());
    return r;
  }
  const float* getLastdLoss_dS() const { return last_grad_.data(); }
  // per-step losses of the last OptimizeFixedStructure (the reference prints them)
  const std::vector<float>& getLosses() const { return losses_; }
  // getLastImage(view), MultiRendererDR.h:86: W * H float4 (colour + coverage)
  const std::vector<float>& getLastImage(uint32_t view) {
    uint32_t w = w_, hh = h_img_;
    lrt_dr_render_size(h_, &w, &hh);  // the render size, resampled references included
    last_img_.assign((size_t)w * hh * 4, 0.0f);
    if (!ok(lrt_dr_get_last_image(h_, view, last_img_.data()))) last_img_.clear();
    return last_img_;
  }

 private:
  uint32_t w_ = 0, h_img_ = 0;
  std::vector<float> last_grad_, last_img_, losses_;
};

}  // namespace dr
}  // namespace lrt_cpp

// C++ parity test written the way the reference's litert tests are
// (tests/tests.cpp:29-41, 139-299): create a GPU renderer, set an SBS scene,
// render with a lookAt/perspective camera, compare with the CPU oracle.
// Exit code 0 = passed.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/litert_amd/MultiRenderer.hpp"
#include "../../oracle/lrt_oracle.h"

using namespace lrt_cpp;

struct float4x4 { float m[16]; };  // stands for LiteMath::float4x4 (column-major)
struct float4 { float x, y, z, w; };

static bool load(lrt_sbs_handle h, SdfSBS& s) {
  uint32_t n; uint64_t nv, nf;
  if (lrt_sbs_info(h, &s.header, &n, &nv, &nf)) return false;
  s.nodes.resize(n); s.values.resize(nv); s.values_f.resize(nf);
  return lrt_sbs_copy(h, s.nodes.data(), s.values.data(), nf ? s.values_f.data() : nullptr) == 0;
}

int main() {
  lrt_sbs_handle hs;
  if (lrt_sbs_displaced_icosphere(5, 0.8f, 0.05f, 3, 1, 32, 4, 1, 4, 0, 0, &hs)) { std::printf("scene: %s\n", lrt_last_error()); return 2; }
  SdfSBS sbs;
  load(hs, sbs);
  lrt_sbs_free(hs);

  const uint32_t W = 256, H = 256;
  float4x4 worldView, proj;
  const float eye[3] = {0, 0, 3}, target[3] = {0, 0, 0}, up[3] = {0, 1, 0};
  lrt_camera_lookat(eye, target, up, 60.0f, 1.0f, 0.1f, 100.0f, worldView.m, proj.m);

  MultiRenderPreset preset = getDefaultPreset();
  preset.render_mode = LRT_RENDER_MODE_LAMBERT_NO_TEX;
  preset.sdf_node_intersect = LRT_SDF_NODE_INTERSECT_NEWTON;

  auto pRender = CreateMultiRenderer(LRT_DEVICE_GPU);
  if (!pRender) { std::printf("CreateMultiRenderer: %s\n", MultiRenderer::lastError().c_str()); return 2; }
  pRender->SetPreset(preset);
  pRender->SetScene(sbs);
  std::vector<float4> image(W * H);
  if (!pRender->RenderFloat(image.data(), W, H, worldView, proj, preset)) { std::printf("render: %s\n", MultiRenderer::lastError().c_str()); return 2; }
  float t[4] = {0, 0, 0, 0};
  pRender->GetExecutionTime("CastRayFloatSingleBlock", t);

  // oracle on the same BVH
  uint32_t np = 0, depth = 0;
  lrt_get_bvh(pRender->handle(), nullptr, 0, &np, &depth);
  std::vector<BVHNodePair> pairs(np);
  lrt_get_bvh(pRender->handle(), pairs.data(), np, nullptr, nullptr);
  orc_scene* os = orc_scene_create(&sbs.header, sbs.nodes.data(), (uint32_t)sbs.nodes.size(), sbs.values.data(), sbs.values.size(), nullptr, 0,
                                   pairs.data(), np);
  std::vector<double> ref(W * H * 4);
  orc_render(os, worldView.m, proj.m, &preset, W, H, 0, 32, 0, 0, ref.data(), nullptr, nullptr, nullptr, nullptr);
  orc_scene_destroy(os);

  size_t same = 0, bad = 0;
  for (size_t i = 0; i < W * H; i++) {
    const float* g = &image[i].x;
    float d = 0;
    for (int k = 0; k < 4; k++) d = std::fmax(d, std::fabs(g[k] - (float)ref[4 * i + k]));
    same += d == 0.0f;
    bad += d > 1e-5f;
  }
  std::printf("facade RenderFloat %ux%u: %.1f ms, %zu/%u pixels identical, %zu beyond 1e-5\n", W, H, t[0], same, W * H, bad);
  bool passed = bad <= W * H / 10000;
  std::printf(" 1.1. %-64s%s\n", "SBS render through the C++ facade matches the oracle", passed ? "passed" : "FAILED");

  // the GPU subclasses' seam: CastRayFloatSingleBlock over every packed pixel
  // gives the same image as RenderFloat; a partial block is refused
  std::vector<float4> block(W * H);
  const bool blk = pRender->CastRayFloatSingleBlock(pRender->packed_count(), block.data()) &&
                   !pRender->CastRayFloatSingleBlock(pRender->packed_count() - 1, block.data());
  size_t diff = 0;
  for (size_t i = 0; i < W * H; i++) diff += std::memcmp(&block[i], &image[i], sizeof(float4)) != 0;
  const bool passed2 = blk && diff == 0;
  std::printf(" 1.2. %-64s%s\n", "CastRayFloatSingleBlock equals RenderFloat", passed2 ? "passed" : "FAILED");
  return passed && passed2 ? 0 : 1;
}

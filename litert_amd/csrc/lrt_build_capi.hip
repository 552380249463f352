// lrt_build_capi.hip - C ABI of the SBS builders from triangle meshes
// (sdf_converter::create_sdf_SBS(settings, header, mesh),
// utils/sdf_converter.cpp:183-228): the host builder (OpenMP) and the device
// builder (build.hip), which produce the same scene bit for bit.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>
#include "../../include/lrt.h"
#include "kernels/build.hpp"
#include "lrt_internal.hpp"

using namespace lrt;

namespace {

uint64_t morton3(uint32_t x, uint32_t y, uint32_t z) {
  auto split = [](uint32_t a) {
    uint64_t v = a & 0x1FFFFF;
    v = (v | v << 32) & 0x1F00000000FFFFull;
    v = (v | v << 16) & 0x1F0000FF0000FFull;
    v = (v | v << 8) & 0x100F00F00F00F00Full;
    v = (v | v << 4) & 0x10C30C30C30C30C3ull;
    v = (v | v << 2) & 0x1249249249249249ull;
    return v;
  };
  return split(x) | split(y) << 1 | split(z) << 2;
}

int check_mesh_args(const float* vertices, uint32_t n_vertices, const uint32_t* indices, uint32_t n_triangles, uint32_t lod,
                    uint32_t brick_size, uint32_t brick_pad, uint32_t bytes_per_value, lrt_sbs_handle* out) {
  if (!vertices || !indices || !out || n_vertices < 3 || n_triangles < 1) return set_err(LRT_E_INVALID, "bad mesh arguments");
  if (lod == 0 || (lod & (lod - 1)) != 0 || lod > 32768) return set_err(LRT_E_INVALID, "lod must be a power of two <= 32768");
  if (brick_size < 1 || brick_size > 16 || brick_pad > 1) return set_err(LRT_E_INVALID, "bad brick size/pad");
  if (bytes_per_value != 1 && bytes_per_value != 2 && bytes_per_value != 4) return set_err(LRT_E_INVALID, "bad bytes_per_value");
  for (uint64_t i = 0; i < 3ull * n_triangles; i++)
    if (indices[i] >= n_vertices) return set_err(LRT_E_INVALID, "triangle index out of range");
  return LRT_OK;
}

TriMesh to_mesh(const float* vertices, uint32_t n_vertices, const uint32_t* indices, uint32_t n_triangles) {
  TriMesh m;
  m.v.resize(n_vertices);
  for (uint32_t i = 0; i < n_vertices; i++) m.v[i] = f3(vertices[3 * i], vertices[3 * i + 1], vertices[3 * i + 2]);
  m.idx.assign(indices, indices + 3ull * n_triangles);
  return m;
}

}  // namespace

extern "C" {

int lrt_mesh_displaced_icosphere(uint32_t subdiv, float radius, float amplitude, uint32_t octaves, uint32_t seed, float* vertices,
                                 uint32_t* n_vertices, uint32_t* indices, uint32_t* n_triangles) {
  if (subdiv > 8 || !n_vertices || !n_triangles) return set_err(LRT_E_INVALID, "bad arguments");
  TriMesh m = displaced_icosphere((int)subdiv, radius, amplitude, (int)octaves, seed);
  const uint32_t nv = (uint32_t)m.v.size(), nt = (uint32_t)(m.idx.size() / 3);
  if ((vertices && *n_vertices < nv) || (indices && *n_triangles < nt)) return set_err(LRT_E_INVALID, "output too small");
  if (vertices)
    for (uint32_t i = 0; i < nv; i++) {
      vertices[3 * i] = m.v[i].x;
      vertices[3 * i + 1] = m.v[i].y;
      vertices[3 * i + 2] = m.v[i].z;
    }
  if (indices) std::copy(m.idx.begin(), m.idx.end(), indices);
  *n_vertices = nv;
  *n_triangles = nt;
  return LRT_OK;
}

int lrt_sbs_from_mesh(const float* vertices, uint32_t n_vertices, const uint32_t* indices, uint32_t n_triangles, uint32_t lod,
                      uint32_t brick_size, uint32_t brick_pad, uint32_t bytes_per_value, int voxel_diag_scale, int threads,
                      lrt_sbs_handle* out, float* build_ms) {
  int c = check_mesh_args(vertices, n_vertices, indices, n_triangles, lod, brick_size, brick_pad, bytes_per_value, out);
  if (c) return c;
  try {
    const auto t0 = std::chrono::steady_clock::now();
    MeshSdf sdf(to_mesh(vertices, n_vertices, indices, n_triangles));
    std::unique_ptr<lrt_sbs> s(new lrt_sbs);
    s->s = build_sbs([&](f3 p) { return sdf(p); }, lod, brick_size, brick_pad, bytes_per_value, voxel_diag_scale != 0, threads);
    if (build_ms) *build_ms = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (s->s.nodes.empty()) return set_err(LRT_E_INVALID, "scene produced no bricks");
    *out = s.release();
  } catch (const std::exception& ex) {
    return set_err(LRT_E_INVALID, ex.what());
  }
  return LRT_OK;
}

int lrt_sbs_from_mesh_device(lrt_handle h, const float* vertices, uint32_t n_vertices, const uint32_t* indices, uint32_t n_triangles,
                             uint32_t lod, uint32_t brick_size, uint32_t brick_pad, uint32_t bytes_per_value, int voxel_diag_scale,
                             lrt_sbs_handle* out, float* build_ms) {
  if (!h) return set_err(LRT_E_INVALID, "null handle");
  int c = check_mesh_args(vertices, n_vertices, indices, n_triangles, lod, brick_size, brick_pad, bytes_per_value, out);
  if (c) return c;
  Renderer* r = h;
  hipError_t e;
  if ((e = hipSetDevice(r->device)) != hipSuccess) return hip_err(e, "hipSetDevice");
  const auto t0 = std::chrono::steady_clock::now();
  MeshSdfData D;
  try {
    D = mesh_sdf_data(to_mesh(vertices, n_vertices, indices, n_triangles));
  } catch (const std::exception& ex) {
    return set_err(LRT_E_INVALID, ex.what());
  }
  if (D.tree.max_depth + 2 > lrtd::kBuildStack) return set_err(LRT_E_UNSUPPORTED, "triangle BVH too deep for the device builder");
  // device copies of the mesh and its distance-query data
  const uint32_t nt = n_triangles, nn = (uint32_t)D.tree.nodes.size();
  std::vector<float4> v4(n_vertices), fn(nt), vn(n_vertices), en(3ull * nt);
  std::vector<uint4> tri(nt), nodes(3ull * nn);
  for (uint32_t i = 0; i < n_vertices; i++) {
    v4[i] = make_float4(D.mesh.v[i].x, D.mesh.v[i].y, D.mesh.v[i].z, 0.f);
    vn[i] = make_float4(D.vnorm[i].x, D.vnorm[i].y, D.vnorm[i].z, 0.f);
  }
  for (uint32_t t = 0; t < nt; t++) {
    tri[t] = make_uint4(D.mesh.idx[3 * t], D.mesh.idx[3 * t + 1], D.mesh.idx[3 * t + 2], 0);
    fn[t] = make_float4(D.fnorm[t].x, D.fnorm[t].y, D.fnorm[t].z, 0.f);
    for (int k = 0; k < 3; k++) en[3 * t + k] = make_float4(D.enorm3[3 * t + k].x, D.enorm3[3 * t + k].y, D.enorm3[3 * t + k].z, 0.f);
  }
  auto fbits = [](float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return u;
  };
  for (uint32_t i = 0; i < nn; i++) {
    const BuildNode& n = D.tree.nodes[i];
    nodes[3 * i] = make_uint4(fbits(n.bmin.x), fbits(n.bmin.y), fbits(n.bmin.z), (uint32_t)n.left);
    nodes[3 * i + 1] = make_uint4(fbits(n.bmax.x), fbits(n.bmax.y), fbits(n.bmax.z), (uint32_t)n.right);
    nodes[3 * i + 2] = make_uint4(n.start, n.count, 0, 0);
  }
  DevBuf b_v, b_tri, b_fn, b_vn, b_en, b_nodes, b_prims, b_cur, b_next, b_count, b_vals, b_keep, b_words;
  struct Release {
    std::vector<DevBuf*> bufs;
    ~Release() {
      for (DevBuf* b : bufs) b->release();
    }
  } rel{{&b_v, &b_tri, &b_fn, &b_vn, &b_en, &b_nodes, &b_prims, &b_cur, &b_next, &b_count, &b_vals, &b_keep, &b_words}};
  hipStream_t s = r->stream;
  if ((e = b_v.upload(v4.data(), v4.size() * 16, s)) != hipSuccess || (e = b_tri.upload(tri.data(), tri.size() * 16, s)) != hipSuccess ||
      (e = b_fn.upload(fn.data(), fn.size() * 16, s)) != hipSuccess || (e = b_vn.upload(vn.data(), vn.size() * 16, s)) != hipSuccess ||
      (e = b_en.upload(en.data(), en.size() * 16, s)) != hipSuccess ||
      (e = b_nodes.upload(nodes.data(), nodes.size() * 16, s)) != hipSuccess ||
      (e = b_prims.upload(D.tree.prims.data(), D.tree.prims.size() * 4, s)) != hipSuccess)
    return hip_err(e, "upload mesh");
  lrtd::DevMesh M{(const float4*)b_v.p, (const uint4*)b_tri.p, (const float4*)b_fn.p, (const float4*)b_vn.p, (const float4*)b_en.p,
                  (const uint4*)b_nodes.p, (const uint32_t*)b_prims.p, D.tree.root};

  // coarse-to-fine candidate cells, level by level (the host builder's rule)
  const float kSqrt3 = 1.7320508075688772f;
  const float dmax_keep = 2.f * kSqrt3 / (float)lod;
  const uint4 root = make_uint4(0, 0, 0, 0);
  if ((e = b_cur.upload(&root, 16, s)) != hipSuccess) return hip_err(e, "upload");
  if ((e = b_count.ensure(4)) != hipSuccess) return hip_err(e, "alloc");
  uint32_t n = 1;
  for (uint32_t level_size = 1; level_size < lod; level_size *= 2) {
    const uint32_t ns = level_size * 2;
    const float side = 2.f / (float)ns;
    const float rr = 1.5f * 0.5f * kSqrt3 * side;
    if ((e = b_next.ensure(8ull * n * 16)) != hipSuccess) return hip_err(e, "alloc candidates");
    if ((e = zero_u32((uint32_t*)b_count.p, 1, s)) != hipSuccess) return hip_err(e, "clear");
    if ((e = launch_build_refine(M, (const uint4*)b_cur.p, n, side, rr, dmax_keep, (uint4*)b_next.p, (uint32_t*)b_count.p, s)) !=
        hipSuccess)
      return hip_err(e, "refine");
    if ((e = hipMemcpyAsync(&n, b_count.p, 4, hipMemcpyDeviceToHost, s)) != hipSuccess || (e = hipStreamSynchronize(s)) != hipSuccess)
      return hip_err(e, "refine count");
    std::swap(b_cur, b_next);
    if (n == 0) return set_err(LRT_E_INVALID, "scene produced no bricks");
  }
  // Morton order, as the host builder sorts its candidates
  std::vector<uint4> cells(n);
  if ((e = hipMemcpyAsync(cells.data(), b_cur.p, (size_t)n * 16, hipMemcpyDeviceToHost, s)) != hipSuccess ||
      (e = hipStreamSynchronize(s)) != hipSuccess)
    return hip_err(e, "download candidates");
  std::sort(cells.begin(), cells.end(), [](const uint4& a, const uint4& b) { return morton3(a.x, a.y, a.z) < morton3(b.x, b.y, b.z); });
  if ((e = b_cur.upload(cells.data(), (size_t)n * 16, s)) != hipSuccess) return hip_err(e, "upload candidates");
  // fill, keep, quantise
  const uint32_t v_size = brick_size + 2 * brick_pad + 1, nv = v_size * v_size * v_size;
  const uint32_t vals_per_int = 4 / bytes_per_value, n_words = (nv + vals_per_int - 1) / vals_per_int;
  const double d_max = voxel_diag_scale ? 2.0 * std::sqrt(3.0) / ((double)lod * brick_size) : 2.0 * std::sqrt(3.0) / (double)lod;
  if ((e = b_vals.ensure((size_t)n * nv * 4)) != hipSuccess || (e = b_keep.ensure((size_t)n * 4)) != hipSuccess ||
      (e = b_words.ensure((size_t)n * n_words * 4)) != hipSuccess)
    return hip_err(e, "alloc bricks");
  if ((e = launch_build_fill(M, (const uint4*)b_cur.p, n, lod, brick_size, brick_pad, (float*)b_vals.p, s)) != hipSuccess)
    return hip_err(e, "fill");
  if ((e = launch_build_pack((const float*)b_vals.p, n, nv, bytes_per_value, dmax_keep, d_max, (uint32_t*)b_keep.p, (uint32_t*)b_words.p,
                             s)) != hipSuccess)
    return hip_err(e, "pack");
  std::vector<uint32_t> keep(n), words((size_t)n * n_words);
  if ((e = hipMemcpyAsync(keep.data(), b_keep.p, (size_t)n * 4, hipMemcpyDeviceToHost, s)) != hipSuccess ||
      (e = hipMemcpyAsync(words.data(), b_words.p, words.size() * 4, hipMemcpyDeviceToHost, s)) != hipSuccess ||
      (e = hipStreamSynchronize(s)) != hipSuccess)
    return hip_err(e, "download bricks");
  std::unique_ptr<lrt_sbs> sb(new lrt_sbs);
  SbsScene& S = sb->s;
  S.header.brick_size = brick_size;
  S.header.brick_pad = brick_pad;
  S.header.bytes_per_value = bytes_per_value;
  S.header.aux_data = LRT_SBS_LAYOUT_DX | (voxel_diag_scale ? 1u : 0u);
  for (uint32_t b = 0; b < n; b++) {
    if (!keep[b]) continue;
    SdfSBSNode nd;
    nd.pos_xy = (cells[b].x << 16) | cells[b].y;
    nd.pos_z_lod_size = (cells[b].z << 16) | lod;
    nd.data_offset = (uint32_t)S.values.size();
    nd._pad = 0;
    S.nodes.push_back(nd);
    S.values.insert(S.values.end(), words.begin() + (size_t)b * n_words, words.begin() + (size_t)(b + 1) * n_words);
  }
  if (build_ms) *build_ms = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
  if (S.nodes.empty()) return set_err(LRT_E_INVALID, "scene produced no bricks");
  *out = sb.release();
  return LRT_OK;
}

}  // extern "C"

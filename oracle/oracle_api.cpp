// oracle_api.cpp - C entry points of the CPU oracle (plain primary-ray path).
// TEST INFRASTRUCTURE ONLY; see lrt_oracle.h.
#include "lrt_oracle.h"
#include "oracle_scene.hpp"
#include <omp.h>

using namespace orc;


extern "C" orc_scene* orc_scene_create(const SdfSBSHeader* h, const SdfSBSNode* nodes, uint32_t n_nodes, const uint32_t* values,
                                       uint64_t n_values, const float* values_f, uint64_t n_values_f, const BVHNodePair* pairs,
                                       uint32_t n_pairs) {
  if (!h || !nodes || n_nodes == 0 || !values || n_values == 0 || !pairs || n_pairs == 0) return nullptr;
  orc_scene* r = new orc_scene;
  r->s.header = *h;
  r->s.quant_voxel_diag = h->aux_data & 1u;
  r->s.header.aux_data = h->aux_data & LRT_SBS_LAYOUT_MASK;
  r->s.nodes.assign(nodes, nodes + n_nodes);
  r->s.values.assign(values, values + n_values);
  if (values_f && n_values_f) r->s.values_f.assign(values_f, values_f + n_values_f);
  r->s.pairs.assign(pairs, pairs + n_pairs);
  uint32_t lay = r->s.header.aux_data;
  // AddGeom_SdfSBS type selection (BVH2Common_host.cpp:573-590)
  r->s.type = (lay == LRT_SBS_LAYOUT_DX_RGB8 || lay == LRT_SBS_LAYOUT_ID32F_IRGB32F || lay == LRT_SBS_LAYOUT_ID32F_IRGB32F_IN)
                  ? LRT_TYPE_SDF_SBS_COL
                  : (lay == LRT_SBS_LAYOUT_DX_UV16 ? LRT_TYPE_SDF_SBS_TEX : LRT_TYPE_SDF_SBS);
  return r;
}

extern "C" void orc_scene_destroy(orc_scene* s) { delete s; }

extern "C" int orc_scene_set_lights(orc_scene* s, const LrtLight* lights, uint32_t n) {
  if (!s || (n && !lights) || n > LRT_MAX_LIGHTS) return -1;
  s->s.lights.assign(lights, lights + n);
  return 0;
}

extern "C" int orc_scene_set_values_f(orc_scene* s, const float* values_f, uint64_t n) {
  if (!s || n != s->s.values_f.size()) return -1;
  s->s.values_f.assign(values_f, values_f + n);
  return 0;
}

template <typename T>
static void render_impl(const Scene& S, const Camera& cam, const MultiRenderPreset& p, uint32_t W, uint32_t H, uint32_t seed,
                        int nthreads, int force_normal, double* out_rgba, double* out_t, uint32_t* out_prim, double* out_n,
                        orc_counters* cnt) {
  std::vector<uint32_t> packed = PackXY(W, H);
  Options O;
  O.preset = p;
  O.force_normal = force_normal != 0;
  Counters total;
  const int64_t N = (int64_t)packed.size();
  if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel num_threads(nthreads)
  {
    Counters local;
    Tracer<T> tr(S, O, cnt ? &local : nullptr);
#pragma omp for schedule(dynamic, 64)
    for (int64_t tid = 0; tid < N; tid++) {
      const uint32_t XY = packed[tid];
      const uint32_t x = XY & 0xFFFFu, y = XY >> 16;
      if (x >= W || y >= H) continue;
      V4<T> acc{0, 0, 0, 0};
      for (uint32_t i = 0; i < p.spp; i++) {
        V2<float> d = sample_offset(p, x, y, i, seed);
        V4<T> pn, dfar;
        InitEyeRay<T>(cam, x, y, d, W, H, &pn, &dfar);
        Hit<T> hit = tr.RayQuery_NearestHit(pn, dfar);
        if (i == 0) {
          size_t pi = (size_t)y * W + x;
          if (out_t) out_t[pi] = (double)hit.t;
          if (out_prim) out_prim[pi] = hit.primId;
          if (out_n) { out_n[2 * pi] = (double)hit.coords[2]; out_n[2 * pi + 1] = (double)hit.coords[3]; }
        }
        if (out_rgba) {
          uint32_t lit = 0xFFFFFFFFu;
          if (needs_shadow(p.render_mode) && hit.primId != 0xFFFFFFFFu) {
            // shadow rays toward the lights (eye_ray.cpp:443-468)
            for (size_t li = 0; li < S.lights.size(); li++) {
              V3<T> o, ld;
              if (!shadow_ray<T>(S.lights[li], V3<T>(pn.x, pn.y, pn.z), V3<T>(dfar.x, dfar.y, dfar.z), hit.t, o, ld)) continue;
              Hit<T> sh = tr.RayQuery_NearestHit(V4<T>{o.x, o.y, o.z, pn.w}, V4<T>{ld.x, ld.y, ld.z, dfar.w});
              if (sh.primId != 0xFFFFFFFFu) lit &= ~(1u << li);
            }
          }
          V4<T> c = shade<T>(p, hit, pn, dfar, S.lights, lit);
          acc = V4<T>{acc.x + c.x, acc.y + c.y, acc.z + c.z, acc.w + c.w};
        }
      }
      if (out_rgba) {
        size_t pi = (size_t)y * W + x;
        T s = T(p.spp);
        out_rgba[4 * pi + 0] = (double)(acc.x / s);
        out_rgba[4 * pi + 1] = (double)(acc.y / s);
        out_rgba[4 * pi + 2] = (double)(acc.z / s);
        out_rgba[4 * pi + 3] = (double)(acc.w / s);
      }
    }
    if (cnt) {
#pragma omp critical
      total.add(local);
    }
  }
  if (cnt) {
    cnt->node_pairs = total.node_pairs; cnt->bricks = total.bricks; cnt->voxels = total.voxels;
    cnt->root_finds = total.root_finds; cnt->rays = total.rays; cnt->root_iters = total.root_iters;
  }
}

extern "C" int orc_render(const orc_scene* s, const float worldView[16], const float proj[16], const MultiRenderPreset* preset,
                          uint32_t W, uint32_t H, uint32_t seed, int precision, int nthreads, int force_normal, double* out_rgba,
                          double* out_t, uint32_t* out_prim, double* out_n, orc_counters* cnt) {
  if (!s || !preset || W == 0 || H == 0 || W > 65535 || H > 65535 || preset->spp == 0) return -1;
  if (preset->interpolation_mode != LRT_INTERPOLATION_MODE_TRILINEAR) return -2;
  Camera cam = Camera::from(worldView, proj);
  if (precision == 64)
    render_impl<double>(s->s, cam, *preset, W, H, seed, nthreads, force_normal, out_rgba, out_t, out_prim, out_n, cnt);
  else
    render_impl<float>(s->s, cam, *preset, W, H, seed, nthreads, force_normal, out_rgba, out_t, out_prim, out_n, cnt);
  return 0;
}

template <typename T>
static void cast_impl(const Scene& S, const MultiRenderPreset& p, const float* rays, uint32_t n, int nthreads, int force_normal,
                      double* out) {
  Options O;
  O.preset = p;
  O.force_normal = force_normal != 0;
  if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 64)
  for (int64_t i = 0; i < (int64_t)n; i++) {
    Tracer<T> tr(S, O, nullptr);
    const float* r = rays + 8 * i;
    Hit<T> h = tr.RayQuery_NearestHit(V4<T>{r[0], r[1], r[2], r[3]}, V4<T>{r[4], r[5], r[6], r[7]});
    double* o = out + 6 * i;
    o[0] = (double)h.t; o[1] = (double)h.primId; o[2] = (double)h.geomId;
    o[3] = (double)h.coords[2]; o[4] = (double)h.coords[3]; o[5] = (double)h.instId;
  }
}

extern "C" int orc_cast_rays(const orc_scene* s, const MultiRenderPreset* preset, const float* rays, uint32_t n, int precision,
                             int nthreads, int force_normal, double* out) {
  if (!s || !preset || !rays || !out) return -1;
  if (precision == 64) cast_impl<double>(s->s, *preset, rays, n, nthreads, force_normal, out);
  else cast_impl<float>(s->s, *preset, rays, n, nthreads, force_normal, out);
  return 0;
}

template <typename T>
static void brute_impl(const Scene& S, const MultiRenderPreset& p, const float* rays, uint32_t n, int nthreads, double* out_t,
                       uint32_t* out_prim) {
  Options O;
  O.preset = p;
  if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 16)
  for (int64_t i = 0; i < (int64_t)n; i++) {
    Tracer<T> tr(S, O, nullptr);
    const float* r = rays + 8 * i;
    V3<T> pos{r[0], r[1], r[2]}, dir{r[4], r[5], r[6]};
    T tNearSdf = mx(T(r[3]), T(0.1f));
    T best = T(r[7] > 0 ? r[7] : -r[7]);
    uint32_t bp = 0xFFFFFFFFu;
    for (uint32_t b = 0; b < (uint32_t)S.nodes.size(); b++) {
      Hit<T> h;
      h.t = T(r[7] > 0 ? r[7] : -r[7]);
      h.primId = 0xFFFFFFFFu;
      tr.OctreeBrickIntersect(pos, dir, tNearSdf, 0, 0, b, &h);
      if (h.primId != 0xFFFFFFFFu && h.t < best) { best = h.t; bp = h.primId; }
    }
    out_t[i] = (double)best;
    out_prim[i] = bp;
  }
}

extern "C" int orc_brute_nearest(const orc_scene* s, const MultiRenderPreset* preset, const float* rays, uint32_t n, int precision,
                                 int nthreads, double* out_t, uint32_t* out_prim) {
  if (!s || !preset || !rays || !out_t || !out_prim) return -1;
  if (precision == 64) brute_impl<double>(s->s, *preset, rays, n, nthreads, out_t, out_prim);
  else brute_impl<float>(s->s, *preset, rays, n, nthreads, out_t, out_prim);
  return 0;
}

// ---- unit-level -----------------------------------------------------------
// wrappers of eval_dist_trilinear / eval_dist_trilinear_diff
// (BVH2Common.cpp:62-72, 657-687) for the closed-form tests

template <typename T> static T tri(const double v[8], const double dp[3]) {
  T vv[8];
  for (int i = 0; i < 8; i++) vv[i] = T(v[i]);
  return eval_dist_trilinear(vv, V3<T>(T(dp[0]), T(dp[1]), T(dp[2])));
}
extern "C" double orc_eval_trilinear(const double v[8], const double dp[3], int precision) {
  return precision == 64 ? tri<double>(v, dp) : (double)tri<float>(v, dp);
}
template <typename T> static void trid(const double v[8], const double dp[3], double out[3]) {
  T vv[8];
  for (int i = 0; i < 8; i++) vv[i] = T(v[i]);
  V3<T> g = eval_dist_trilinear_diff(vv, V3<T>(T(dp[0]), T(dp[1]), T(dp[2])));
  out[0] = g.x; out[1] = g.y; out[2] = g.z;
}
extern "C" void orc_eval_trilinear_diff(const double v[8], const double dp[3], int precision, double out[3]) {
  if (precision == 64) trid<double>(v, dp, out); else trid<float>(v, dp, out);
}

template <typename T>
static int vroot(uint32_t mode, const double values[8], const double sq[3], const double dir[3], double d, double qFar, double* t_out,
                 uint32_t* it_out) {
  Scene dummy;
  Options O;
  O.preset.render_mode = LRT_RENDER_MODE_ST_ITERATIONS;  // iteration count reported in primId
  O.preset.sdf_node_intersect = mode;
  O.preset.representation_mode = LRT_REPRESENTATION_MODE_VOLUME;
  O.preset.normal_mode = 0;
  O.preset.interpolation_mode = 0;
  O.preset.spp = 1;
  O.preset.ray_gen_mode = 0;
  Tracer<T> tr(dummy, O, nullptr);
  T v[8];
  for (int i = 0; i < 8; i++) v[i] = T(values[i]);
  Hit<T> h;
  h.t = T(1e9f);
  h.primId = 0xFFFFFFFFu;
  tr.LocalSurfaceIntersection(V3<T>(T(dir[0]), T(dir[1]), T(dir[2])), 0, 0, v, 0, 0, T(d), T(0), T(qFar), V2<T>{T(0), T(qFar * d)},
                              V3<T>(T(sq[0]), T(sq[1]), T(sq[2])), &h);
  if (!tr.last_hit) return 0;
  *t_out = (double)tr.last_t;
  if (it_out) *it_out = tr.last_iter;
  return 1;
}
extern "C" int orc_voxel_root(uint32_t mode, const double values[8], const double start_q[3], const double dir[3], double d,
                              double qFar, int precision, double* t_out, uint32_t* iters_out) {
  return precision == 64 ? vroot<double>(mode, values, start_q, dir, d, qFar, t_out, iters_out)
                         : vroot<float>(mode, values, start_q, dir, d, qFar, t_out, iters_out);
}

extern "C" void orc_encode_normal(const double n[3], int precision, double out[2]) {
  if (precision == 64) { auto e = encode_normal(V3<double>(n[0], n[1], n[2])); out[0] = e.x; out[1] = e.y; }
  else { auto e = encode_normal(V3<float>((float)n[0], (float)n[1], (float)n[2])); out[0] = e.x; out[1] = e.y; }
}
extern "C" void orc_decode_normal(const double e[2], int precision, double out[3]) {
  if (precision == 64) { auto v = decode_normal(V2<double>{e[0], e[1]}); out[0] = v.x; out[1] = v.y; out[2] = v.z; }
  else { auto v = decode_normal(V2<float>{(float)e[0], (float)e[1]}); out[0] = v.x; out[1] = v.y; out[2] = v.z; }
}
extern "C" void orc_pcg3d(uint32_t v[3]) { pcg3d(v[0], v[1], v[2]); }
extern "C" void orc_rand2(uint32_t x, uint32_t y, uint32_t iter, uint32_t seed, float out[2]) {
  V2<float> r = rand2(x, y, iter, seed);
  out[0] = r.x; out[1] = r.y;
}
extern "C" void orc_pack_xy(uint32_t W, uint32_t H, uint32_t* out) {
  std::vector<uint32_t> p = PackXY(W, H);
  std::memcpy(out, p.data(), p.size() * sizeof(uint32_t));
}
extern "C" void orc_safe_inverse(const double d[3], int precision, double out[3]) {
  if (precision == 64) { auto r = SafeInverse(V3<double>(d[0], d[1], d[2])); out[0] = r.x; out[1] = r.y; out[2] = r.z; }
  else { auto r = SafeInverse(V3<float>((float)d[0], (float)d[1], (float)d[2])); out[0] = r.x; out[1] = r.y; out[2] = r.z; }
}
extern "C" void orc_ray_box(const double o[3], const double inv[3], const double bmin[3], const double bmax[3], int precision,
                            double out[2]) {
  if (precision == 64) {
    auto r = RayBoxIntersection2(V3<double>(o[0], o[1], o[2]), V3<double>(inv[0], inv[1], inv[2]), V3<double>(bmin[0], bmin[1], bmin[2]),
                                 V3<double>(bmax[0], bmax[1], bmax[2]));
    out[0] = r.x; out[1] = r.y;
  } else {
    auto f = [](double v) { return (float)v; };
    auto r = RayBoxIntersection2(V3<float>(f(o[0]), f(o[1]), f(o[2])), V3<float>(f(inv[0]), f(inv[1]), f(inv[2])),
                                 V3<float>(f(bmin[0]), f(bmin[1]), f(bmin[2])), V3<float>(f(bmax[0]), f(bmax[1]), f(bmax[2])));
    out[0] = r.x; out[1] = r.y;
  }
}
extern "C" double orc_load_voxel(const orc_scene* s, uint32_t node, uint32_t vx, uint32_t vy, uint32_t vz, int precision,
                                 double out[8]) {
  Options O;
  O.preset = MultiRenderPreset{0, 0, 0, 0, 0, 1, 1};
  const SdfSBSHeader& h = s->s.header;
  uint32_t v_size = h.brick_size + 2 * h.brick_pad + 1;
  uint32_t vp[3] = {vx, vy, vz};
  float sz = float(s->s.nodes[node].pos_z_lod_size & 0xFFFF);
  if (precision == 64) {
    Tracer<double> tr(s->s, O, nullptr);
    double v[8];
    double m = tr.load_distance_values(node, vp, v_size, 2.0 / (double)sz, v);
    for (int i = 0; i < 8; i++) out[i] = v[i];
    return m;
  }
  Tracer<float> tr(s->s, O, nullptr);
  float v[8];
  float m = tr.load_distance_values(node, vp, v_size, 2.0f / sz, v);
  for (int i = 0; i < 8; i++) out[i] = v[i];
  return m;
}
extern "C" void orc_eye_ray(const float worldView[16], const float proj[16], uint32_t x, uint32_t y, float dx, float dy, uint32_t W,
                            uint32_t H, int precision, double out[8]) {
  Camera cam = Camera::from(worldView, proj);
  if (precision == 64) {
    V4<double> a, b;
    InitEyeRay<double>(cam, x, y, V2<float>{dx, dy}, W, H, &a, &b);
    double r[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
    for (int i = 0; i < 8; i++) out[i] = r[i];
  } else {
    V4<float> a, b;
    InitEyeRay<float>(cam, x, y, V2<float>{dx, dy}, W, H, &a, &b);
    double r[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
    for (int i = 0; i < 8; i++) out[i] = r[i];
  }
}

// scene_gen.cpp - synthetic scenes and SBS brick builders (host, OpenMP).
#include "scene_gen.hpp"
#include "bvh_build.hpp"
#include <cfloat>
#include <cstring>
#include <map>
#include <stdexcept>
#include <omp.h>

namespace lrt {

// ---------------------------------------------------------------------------
// value noise
// ---------------------------------------------------------------------------
static inline uint32_t hash3(int32_t x, int32_t y, int32_t z, uint32_t seed) {
  uint32_t h = seed * 0x9E3779B1u;
  h ^= (uint32_t)x * 0x85EBCA77u; h = (h << 13) | (h >> 19);
  h ^= (uint32_t)y * 0xC2B2AE3Du; h = (h << 13) | (h >> 19);
  h ^= (uint32_t)z * 0x27D4EB2Fu; h = (h << 13) | (h >> 19);
  h ^= h >> 16; h *= 0x7FEB352Du; h ^= h >> 15; h *= 0x846CA68Bu; h ^= h >> 16;
  return h;
}
static inline float lattice(int32_t x, int32_t y, int32_t z, uint32_t seed) {
  return (float)(hash3(x, y, z, seed) >> 8) * (2.0f / 16777215.0f) - 1.0f;
}
static inline float fade(float t) { return t * t * t * (t * (t * 6.f - 15.f) + 10.f); }

static float value_noise(f3 p, uint32_t seed) {
  float fx = std::floor(p.x), fy = std::floor(p.y), fz = std::floor(p.z);
  int32_t ix = (int32_t)fx, iy = (int32_t)fy, iz = (int32_t)fz;
  float u = fade(p.x - fx), v = fade(p.y - fy), w = fade(p.z - fz);
  float c[8];
  for (int i = 0; i < 8; i++) c[i] = lattice(ix + ((i >> 2) & 1), iy + ((i >> 1) & 1), iz + (i & 1), seed);
  float x00 = c[0] + u * (c[4] - c[0]), x01 = c[1] + u * (c[5] - c[1]);
  float x10 = c[2] + u * (c[6] - c[2]), x11 = c[3] + u * (c[7] - c[3]);
  float y0 = x00 + v * (x10 - x00), y1 = x01 + v * (x11 - x01);
  return y0 + w * (y1 - y0);
}

float value_noise_fbm(f3 p, uint32_t seed, int octaves, float base_freq) {
  float sum = 0.f, amp = 1.f, norm = 0.f, f = base_freq;
  for (int o = 0; o < octaves; o++) {
    sum += amp * value_noise(f * p, seed + 101u * (uint32_t)o);
    norm += amp;
    amp *= 0.5f;
    f *= 2.0f;
  }
  return norm > 0.f ? sum / norm : 0.f;
}

// ---------------------------------------------------------------------------
// icosphere
// ---------------------------------------------------------------------------
TriMesh displaced_icosphere(int subdiv, float r, float amp, int octaves, uint32_t seed) {
  TriMesh m;
  const float t = (1.0f + std::sqrt(5.0f)) * 0.5f;
  f3 base[12] = {f3(-1, t, 0), f3(1, t, 0), f3(-1, -t, 0), f3(1, -t, 0), f3(0, -1, t), f3(0, 1, t),
                 f3(0, -1, -t), f3(0, 1, -t), f3(t, 0, -1), f3(t, 0, 1), f3(-t, 0, -1), f3(-t, 0, 1)};
  for (auto& b : base) m.v.push_back(normalize(b));
  uint32_t faces[60] = {0, 11, 5, 0, 5, 1, 0, 1, 7, 0, 7, 10, 0, 10, 11, 1, 5, 9, 5, 11, 4, 11, 10, 2, 10, 7, 6, 7, 1, 8,
                        3, 9, 4, 3, 4, 2, 3, 2, 6, 3, 6, 8, 3, 8, 9, 4, 9, 5, 2, 4, 11, 6, 2, 10, 8, 6, 7, 9, 8, 1};
  m.idx.assign(faces, faces + 60);
  for (int s = 0; s < subdiv; s++) {
    std::map<uint64_t, uint32_t> mid;
    auto midpoint = [&](uint32_t a, uint32_t b) {
      uint64_t key = a < b ? ((uint64_t)a << 32 | b) : ((uint64_t)b << 32 | a);
      auto it = mid.find(key);
      if (it != mid.end()) return it->second;
      uint32_t id = (uint32_t)m.v.size();
      m.v.push_back(normalize(0.5f * (m.v[a] + m.v[b])));
      mid[key] = id;
      return id;
    };
    std::vector<uint32_t> ni;
    ni.reserve(m.idx.size() * 4);
    for (size_t f = 0; f < m.idx.size(); f += 3) {
      uint32_t a = m.idx[f], b = m.idx[f + 1], c = m.idx[f + 2];
      uint32_t ab = midpoint(a, b), bc = midpoint(b, c), ca = midpoint(c, a);
      uint32_t tri[12] = {a, ab, ca, b, bc, ab, c, ca, bc, ab, bc, ca};
      ni.insert(ni.end(), tri, tri + 12);
    }
    m.idx.swap(ni);
  }
  for (auto& v : m.v) {
    f3 n = v;  // unit sphere normal
    float disp = amp * value_noise_fbm(n, seed, octaves, 2.0f);
    v = (r + disp) * n;
  }
  return m;
}

// ---------------------------------------------------------------------------
// mesh signed distance
// ---------------------------------------------------------------------------
struct MeshSdf::Impl {
  MeshSdfData D;

  // closest point on triangle (Ericson, Real-Time Collision Detection 5.1.5);
  // region: 0 face, 1..3 vertex a/b/c, 4 edge ab, 5 edge bc, 6 edge ca
  static f3 closest(f3 p, f3 a, f3 b, f3 c, int& region) {
    f3 ab = b - a, ac = c - a, ap = p - a;
    float d1 = dot(ab, ap), d2 = dot(ac, ap);
    if (d1 <= 0.f && d2 <= 0.f) { region = 1; return a; }
    f3 bp = p - b;
    float d3 = dot(ab, bp), d4 = dot(ac, bp);
    if (d3 >= 0.f && d4 <= d3) { region = 2; return b; }
    float vc = d1 * d4 - d3 * d2;
    if (vc <= 0.f && d1 >= 0.f && d3 <= 0.f) { region = 4; float v = d1 / (d1 - d3); return a + v * ab; }
    f3 cp = p - c;
    float d5 = dot(ab, cp), d6 = dot(ac, cp);
    if (d6 >= 0.f && d5 <= d6) { region = 3; return c; }
    float vb = d5 * d2 - d1 * d6;
    if (vb <= 0.f && d2 >= 0.f && d6 <= 0.f) { region = 6; float w = d2 / (d2 - d6); return a + w * ac; }
    float va = d3 * d6 - d5 * d4;
    if (va <= 0.f && (d4 - d3) >= 0.f && (d5 - d6) >= 0.f) {
      region = 5;
      float w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
      return b + w * (c - b);
    }
    region = 0;
    float denom = 1.f / (va + vb + vc);
    float v = vb * denom, w = vc * denom;
    return a + v * ab + w * ac;
  }
};

MeshSdfData mesh_sdf_data(const TriMesh& m) {
  MeshSdfData D;
  D.mesh = m;
  size_t nt = m.idx.size() / 3;
  std::vector<f3> tmin(nt), tmax(nt);
  std::map<uint64_t, f3> enorm;  // edge -> sum of adjacent face normals
  auto ekey = [](uint32_t a, uint32_t b) { return a < b ? ((uint64_t)a << 32 | b) : ((uint64_t)b << 32 | a); };
  D.fnorm.resize(nt);
  D.vnorm.assign(m.v.size(), f3());
  for (size_t t = 0; t < nt; t++) {
    f3 a = m.v[m.idx[3 * t]], b = m.v[m.idx[3 * t + 1]], c = m.v[m.idx[3 * t + 2]];
    tmin[t] = fmin3(a, fmin3(b, c));
    tmax[t] = fmax3(a, fmax3(b, c));
    f3 n = normalize(cross(b - a, c - a));
    D.fnorm[t] = n;
    f3 pts[3] = {a, b, c};
    for (int k = 0; k < 3; k++) {
      f3 e1 = normalize(pts[(k + 1) % 3] - pts[k]), e2 = normalize(pts[(k + 2) % 3] - pts[k]);
      float ang = std::acos(std::max(-1.f, std::min(1.f, dot(e1, e2))));
      uint32_t vi = m.idx[3 * t + k];
      D.vnorm[vi] = D.vnorm[vi] + ang * n;
      uint64_t ek = ekey(m.idx[3 * t + k], m.idx[3 * t + (k + 1) % 3]);
      enorm[ek] = enorm[ek] + n;
    }
  }
  D.enorm3.resize(3 * nt);
  for (size_t t = 0; t < nt; t++)
    for (int k = 0; k < 3; k++) D.enorm3[3 * t + k] = enorm.at(ekey(m.idx[3 * t + k], m.idx[3 * t + (k + 1) % 3]));
  D.tree = build_sah(tmin, tmax, 4);
  return D;
}

MeshSdf::MeshSdf(const TriMesh& m) : impl_(new Impl) { impl_->D = mesh_sdf_data(m); }

MeshSdf::~MeshSdf() { delete impl_; }

const MeshSdfData& MeshSdf::data() const { return impl_->D; }

float MeshSdf::operator()(f3 p) const {
  const MeshSdfData& I = impl_->D;
  const TriMesh& m = I.mesh;
  float best = FLT_MAX;
  uint32_t best_t = 0;
  int best_region = 0;
  f3 best_pt;
  int stack[128];
  int sp = 0;
  stack[sp++] = I.tree.root;
  auto box_d2 = [&](const BuildNode& n) {
    float dx = std::max(0.f, std::max(n.bmin.x - p.x, p.x - n.bmax.x));
    float dy = std::max(0.f, std::max(n.bmin.y - p.y, p.y - n.bmax.y));
    float dz = std::max(0.f, std::max(n.bmin.z - p.z, p.z - n.bmax.z));
    return dx * dx + dy * dy + dz * dz;
  };
  while (sp > 0) {
    const BuildNode& n = I.tree.nodes[stack[--sp]];
    if (box_d2(n) >= best) continue;
    if (n.count) {
      for (uint32_t k = n.start; k < n.start + n.count; k++) {
        uint32_t t = I.tree.prims[k];
        int region;
        f3 q = Impl::closest(p, m.v[m.idx[3 * t]], m.v[m.idx[3 * t + 1]], m.v[m.idx[3 * t + 2]], region);
        f3 dv = p - q;
        float d2 = dot(dv, dv);
        if (d2 < best) { best = d2; best_t = t; best_region = region; best_pt = q; }
      }
    } else {
      const BuildNode& L = I.tree.nodes[n.left];
      const BuildNode& R = I.tree.nodes[n.right];
      float dl = box_d2(L), dr = box_d2(R);
      if (dl < dr) { stack[sp++] = n.right; stack[sp++] = n.left; }
      else { stack[sp++] = n.left; stack[sp++] = n.right; }
    }
  }
  f3 pn;
  const uint32_t* tri = &m.idx[3 * best_t];
  switch (best_region) {
    case 0: pn = I.fnorm[best_t]; break;
    case 1: pn = I.vnorm[tri[0]]; break;
    case 2: pn = I.vnorm[tri[1]]; break;
    case 3: pn = I.vnorm[tri[2]]; break;
    case 4: pn = I.enorm3[3 * best_t + 0]; break;
    case 5: pn = I.enorm3[3 * best_t + 1]; break;
    default: pn = I.enorm3[3 * best_t + 2]; break;
  }
  float d = std::sqrt(best);
  return dot(p - best_pt, pn) < 0.f ? -d : d;
}

// ---------------------------------------------------------------------------
// SBS builders
// ---------------------------------------------------------------------------
static inline uint64_t morton_split(uint32_t a) {
  uint64_t x = a & 0x1FFFFF;
  x = (x | x << 32) & 0x1F00000000FFFFull;
  x = (x | x << 16) & 0x1F0000FF0000FFull;
  x = (x | x << 8) & 0x100F00F00F00F00Full;
  x = (x | x << 4) & 0x10C30C30C30C30C3ull;
  x = (x | x << 2) & 0x1249249249249249ull;
  return x;
}

SbsScene build_sbs(const std::function<float(f3)>& sdf, uint32_t lod, uint32_t brick_size, uint32_t brick_pad,
                   uint32_t bytes_per_value, bool voxel_diag_scale, int threads) {
  if (lod == 0 || (lod & (lod - 1)) != 0 || lod > 32768) throw std::invalid_argument("lod must be a power of two <= 32768");
  if (brick_size < 1 || brick_size > 16 || brick_pad > 1) throw std::invalid_argument("bad brick size/pad");
  if (bytes_per_value != 1 && bytes_per_value != 2 && bytes_per_value != 4) throw std::invalid_argument("bad bytes_per_value");
  if (threads <= 0) threads = omp_get_max_threads();
  const float kSqrt3 = 1.7320508075688772f;
  const float dmax_keep = 2.f * kSqrt3 / (float)lod;  // is_border_node, sparse_octree_builder.cpp:49-53

  // coarse-to-fine candidate search; a cell can hold a kept brick only if
  // the 1-Lipschitz distance at its centre lies within its half-diagonal of
  // the band [-dmax_keep, 0] (1.5x margin for non-exact distance fields).
  struct Cell { uint32_t x, y, z; };
  std::vector<Cell> cur{{0, 0, 0}};
  uint32_t level_size = 1;
  while (level_size < lod) {
    std::vector<Cell> nxt;
    uint32_t ns = level_size * 2;
    float side = 2.f / (float)ns;
    float r = 1.5f * 0.5f * kSqrt3 * side;
    std::vector<std::vector<Cell>> parts(threads);
#pragma omp parallel for num_threads(threads) schedule(dynamic, 16)
    for (int64_t i = 0; i < (int64_t)cur.size(); i++) {
      int tid = omp_get_thread_num();
      for (int c = 0; c < 8; c++) {
        Cell ch{cur[i].x * 2 + ((c >> 2) & 1), cur[i].y * 2 + ((c >> 1) & 1), cur[i].z * 2 + (c & 1)};
        f3 ctr(-1.f + side * (ch.x + 0.5f), -1.f + side * (ch.y + 0.5f), -1.f + side * (ch.z + 0.5f));
        float v = sdf(ctr);
        if (v <= r && v >= -(r + dmax_keep)) parts[tid].push_back(ch);
      }
    }
    for (auto& p : parts) nxt.insert(nxt.end(), p.begin(), p.end());
    cur.swap(nxt);
    level_size = ns;
  }
  std::sort(cur.begin(), cur.end(), [](const Cell& a, const Cell& b) {
    uint64_t ma = morton_split(a.x) | morton_split(a.y) << 1 | morton_split(a.z) << 2;
    uint64_t mb = morton_split(b.x) | morton_split(b.y) << 1 | morton_split(b.z) << 2;
    return ma < mb;
  });

  const uint32_t v_size = brick_size + 2 * brick_pad + 1;
  const uint32_t n_vals = v_size * v_size * v_size;
  const uint32_t vals_per_int = 4 / bytes_per_value;
  const uint32_t words = (n_vals + vals_per_int - 1) / vals_per_int;
  const uint32_t bits = 8 * bytes_per_value;
  const double max_val = bytes_per_value == 4 ? 4294967295.0 : (double)((1u << bits) - 1);
  const double d_max = voxel_diag_scale ? 2.0 * std::sqrt(3.0) / ((double)lod * brick_size) : 2.0 * std::sqrt(3.0) / (double)lod;

  std::vector<uint8_t> keep(cur.size(), 0);
  std::vector<uint32_t> packed((size_t)cur.size() * words, 0);
#pragma omp parallel for num_threads(threads) schedule(dynamic, 8)
  for (int64_t b = 0; b < (int64_t)cur.size(); b++) {
    std::vector<float> vals(n_vals);
    const Cell& c = cur[b];
    const float d = 1.f / (float)lod;
    f3 p0(2.f * (d * c.x) - 1.f, 2.f * (d * c.y) - 1.f, 2.f * (d * c.z) - 1.f);
    const float dp = 2.f * d / (float)brick_size;
    float mn = 1000.f, mxv = -1000.f;
    for (int i = -(int)brick_pad; i <= (int)(brick_size + brick_pad); i++)
      for (int j = -(int)brick_pad; j <= (int)(brick_size + brick_pad); j++)
        for (int k = -(int)brick_pad; k <= (int)(brick_size + brick_pad); k++) {
          float v = sdf(p0 + dp * f3((float)i, (float)j, (float)k));
          vals[(i + brick_pad) * v_size * v_size + (j + brick_pad) * v_size + (k + brick_pad)] = v;
          mn = std::min(mn, v);
          mxv = std::max(mxv, v);
        }
    if (!(mn <= 0.f && mxv > -dmax_keep)) continue;
    keep[b] = 1;
    uint32_t* dst = &packed[(size_t)b * words];
    for (uint32_t i = 0; i < n_vals; i++) {
      double q = std::floor(max_val * (((double)vals[i] + d_max) / (2.0 * d_max)));
      q = std::max(0.0, std::min(q, max_val));
      dst[i / vals_per_int] |= ((uint32_t)q) << (bits * (i % vals_per_int));
    }
  }

  SbsScene s;
  s.header.brick_size = brick_size;
  s.header.brick_pad = brick_pad;
  s.header.bytes_per_value = bytes_per_value;
  s.header.aux_data = LRT_SBS_LAYOUT_DX | (voxel_diag_scale ? 1u : 0u);
  for (size_t b = 0; b < cur.size(); b++) {
    if (!keep[b]) continue;
    SdfSBSNode n;
    n.pos_xy = (cur[b].x << 16) | cur[b].y;
    n.pos_z_lod_size = (cur[b].z << 16) | lod;
    n.data_offset = (uint32_t)s.values.size();
    n._pad = 0;
    s.nodes.push_back(n);
    s.values.insert(s.values.end(), packed.begin() + b * words, packed.begin() + (b + 1) * words);
  }
  return s;
}

SbsScene build_grid_sbs(uint32_t brick_count, uint32_t brick_size, const std::function<float(f3)>& sdf,
                        const std::function<f3(f3)>& color, bool with_neighbors) {
  const uint32_t v_size = brick_size + 1;
  const uint32_t dist_per_node = v_size * v_size * v_size;
  const uint32_t per_node = dist_per_node + 8 + (with_neighbors ? 27 : 0);
  const uint32_t p_count = brick_count * brick_size + 1;
  const uint32_t c_count = brick_count + 1;
  const uint32_t c_offset = p_count * p_count * p_count;
  SbsScene s;
  s.header.brick_size = brick_size;
  s.header.brick_pad = 0;
  s.header.bytes_per_value = 4;
  s.header.aux_data = with_neighbors ? LRT_SBS_LAYOUT_ID32F_IRGB32F_IN : LRT_SBS_LAYOUT_ID32F_IRGB32F;
  s.values_f.resize((size_t)c_offset + 3 * (size_t)c_count * c_count * c_count);
  const uint32_t nb = brick_count * brick_count * brick_count;
  s.values.resize((size_t)nb * per_node);
  s.nodes.resize(nb);
  const float n = (float)(brick_count * brick_size);
  // 2 * (xyz / n) - 1, DR_common.cpp:68-79
#pragma omp parallel for schedule(static)
  for (int64_t x = 0; x < (int64_t)p_count; x++)
    for (uint32_t y = 0; y < p_count; y++)
      for (uint32_t z = 0; z < p_count; z++)
        s.values_f[((size_t)x * p_count + y) * p_count + z] =
            sdf(f3(2.f * ((float)x / n) - 1.f, 2.f * ((float)y / n) - 1.f, 2.f * ((float)z / n) - 1.f));
  for (uint32_t x = 0; x < c_count; x++)
    for (uint32_t y = 0; y < c_count; y++)
      for (uint32_t z = 0; z < c_count; z++) {
        f3 dp((float)x / brick_count, (float)y / brick_count, (float)z / brick_count);
        f3 c = color(f3(2.f * dp.x - 1.f, 2.f * dp.y - 1.f, 2.f * dp.z - 1.f));
        size_t id = c_offset + 3 * (((size_t)x * c_count + y) * c_count + z);
        s.values_f[id] = c.x; s.values_f[id + 1] = c.y; s.values_f[id + 2] = c.z;
      }
  for (uint32_t bx = 0; bx < brick_count; bx++)
    for (uint32_t by = 0; by < brick_count; by++)
      for (uint32_t bz = 0; bz < brick_count; bz++) {
        uint32_t n = (bx * brick_count + by) * brick_count + bz;
        uint32_t off = n * per_node;
        s.nodes[n].pos_xy = (bx << 16) | by;
        s.nodes[n].pos_z_lod_size = (bz << 16) | brick_count;
        s.nodes[n].data_offset = off;
        s.nodes[n]._pad = 0;
        for (uint32_t x = 0; x < v_size; x++)
          for (uint32_t y = 0; y < v_size; y++)
            for (uint32_t z = 0; z < v_size; z++)
              s.values[off + (x * v_size + y) * v_size + z] =
                  ((bx * brick_size + x) * p_count + (by * brick_size + y)) * p_count + (bz * brick_size + z);
        for (uint32_t i = 0; i < 8; i++) {
          uint32_t x = (i >> 2) & 1, y = (i >> 1) & 1, z = i & 1;
          s.values[off + dist_per_node + i] = c_offset + 3 * (((bx + x) * c_count + (by + y)) * c_count + (bz + z));
        }
        if (with_neighbors) {
          for (int dx = -1; dx <= 1; dx++)
            for (int dy = -1; dy <= 1; dy++)
              for (int dz = -1; dz <= 1; dz++) {
                int nx = (int)bx + dx, ny = (int)by + dy, nz = (int)bz + dz;
                uint32_t slot = 9 * (dx + 1) + 3 * (dy + 1) + (dz + 1);
                bool in = nx >= 0 && ny >= 0 && nz >= 0 && nx < (int)brick_count && ny < (int)brick_count && nz < (int)brick_count;
                s.values[off + dist_per_node + 8 + slot] = in ? (uint32_t)((nx * brick_count + ny) * brick_count + nz) : LRT_INVALID_IDX;
              }
        }
      }
  return s;
}

}  // namespace lrt

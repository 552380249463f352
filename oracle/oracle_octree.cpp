// oracle_octree.cpp - CPU restatement of the reference's SBS builder for a
// distance field: sdf_converter::create_sdf_SBS(settings, header, sdf)
// (utils/sdf_converter.cpp:169-181) = construct_sdf_frame_octree
// (utils/sparse_octree_builder.cpp:125-258) + frame_octree_to_SBS (:291-444).
// TEST INFRASTRUCTURE ONLY; see lrt_oracle.h.
//
// The reference builds in parallel (OpenMP groups of subtrees, bricks appended
// under a critical section, so its brick order varies from run to run); this
// restatement runs the same steps on one thread, so its order is the frame
// octree's. frame_octree_limit_nodes (:2229-2330) runs on the frame before
// the bricks are made, as in create_sdf_SBS.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include "lrt_oracle.h"

namespace {

struct F3 {
  float x, y, z;
};
F3 f3(float x, float y, float z) { return F3{x, y, z}; }
F3 operator+(F3 a, F3 b) { return F3{a.x + b.x, a.y + b.y, a.z + b.z}; }
F3 operator-(F3 a, F3 b) { return F3{a.x - b.x, a.y - b.y, a.z - b.z}; }
F3 operator*(float s, F3 a) { return F3{s * a.x, s * a.y, s * a.z}; }
F3 operator*(F3 a, float s) { return F3{a.x * s, a.y * s, a.z * s}; }

// Distance fields the restatement takes: 0 sphere (centre, radius),
// 1 torus around the z axis (major R, minor r), each as LiteMath would
// evaluate it in float (length = sqrt(dot)).
struct Field {
  int shape;
  float a[4];
  float operator()(F3 p) const {
    if (shape == 0) {
      const F3 q = p - f3(a[0], a[1], a[2]);
      return std::sqrt(q.x * q.x + q.y * q.y + q.z * q.z) - a[3];
    }
    const float qx = std::sqrt(p.x * p.x + p.y * p.y) - a[0];
    return std::sqrt(qx * qx + p.z * p.z) - a[1];
  }
};

struct FrameNode {  // SdfFrameOctreeNode: 8 corner values, offset of the 8 children (0 = leaf)
  float values[8];
  uint32_t offset = 0;
};

bool is_border(float distance, int level) {  // sparse_octree_builder.cpp:38-41, evaluated in double
  return level < 2 ? true : std::abs((double)distance) < std::sqrt(3.0) * std::pow(2.0, -level);
}

bool is_border_node(float min_val, float max_val, unsigned level_size) {  // :49-53
  const float d_max = (float)(2 * std::sqrt(3.0) / level_size);
  return (min_val <= 0) && (max_val > -d_max);
}

// add_node_rec (:72-96): corner values, then eight children while the
// centre is within the node's half-diagonal or the corners change sign.
void add_node_rec(std::vector<FrameNode>& nodes, const Field& sdf, unsigned node_idx, unsigned depth, unsigned max_depth, F3 p, float d) {
  const float value_center = sdf(2.0f * ((p + f3(0.5f, 0.5f, 0.5f)) * d) - f3(1, 1, 1));
  float min_val = 1000, max_val = -1000;
  for (int cid = 0; cid < 8; cid++) {
    const F3 c = f3((float)((cid & 4) >> 2), (float)((cid & 2) >> 1), (float)(cid & 1));
    nodes[node_idx].values[cid] = sdf(2.0f * ((p + c) * d) - f3(1, 1, 1));
    min_val = std::fmin(min_val, nodes[node_idx].values[cid]);
    max_val = std::fmax(max_val, nodes[node_idx].values[cid]);
  }
  if (depth < max_depth && (std::abs(value_center) < std::sqrt(3.0f) * d || min_val * max_val <= 0)) {
    nodes[node_idx].offset = (uint32_t)nodes.size();
    nodes.resize(nodes.size() + 8);
    const unsigned idx = nodes[node_idx].offset;
    for (unsigned cid = 0; cid < 8; cid++) {
      const F3 c = f3((float)((cid & 4) >> 2), (float)((cid & 2) >> 1), (float)(cid & 1));
      add_node_rec(nodes, sdf, idx + cid, depth + 1, max_depth, 2 * p + c, d / 2);
    }
  }
}

struct LargeNode {
  F3 p;
  float d;
  unsigned level;
  unsigned children_idx;
  float value;
};

// construct_sdf_frame_octree (:125-258) on one thread: the uniform top
// levels (down to min(depth, 4)), then one add_node_rec subtree per border
// node of the last uniform level, appended in node order.
std::vector<FrameNode> frame_octree(const Field& sdf, unsigned depth, std::vector<F3>& pos, std::vector<float>& size) {
  const unsigned min_remove_level = depth < 4u ? depth : 4u;
  std::vector<LargeNode> large;
  large.push_back({f3(0, 0, 0), 1.0f, 0u, 0u, 1000.0f});
  for (unsigned i = 0; i < large.size(); i++) {
    if (large[i].level < min_remove_level) {
      large[i].children_idx = (unsigned)large.size();
      for (int j = 0; j < 8; j++) {
        const F3 c = f3((float)((j & 4) >> 2), (float)((j & 2) >> 1), (float)(j & 1));
        large.push_back({2 * large[i].p + c, large[i].d / 2, large[i].level + 1, 0u, 1000.0f});
      }
    }
  }
  for (auto& n : large) n.value = sdf(2.0f * ((n.p + f3(0.5f, 0.5f, 0.5f)) * n.d) - f3(1, 1, 1));

  std::vector<FrameNode> res(large.size());
  pos.resize(large.size());
  size.resize(large.size());
  for (unsigned i = 0; i < large.size(); i++) {
    for (int cid = 0; cid < 8; cid++) {
      const F3 c = f3((float)((cid & 4) >> 2), (float)((cid & 2) >> 1), (float)(cid & 1));
      res[i].values[cid] = sdf(2.0f * ((large[i].p + c) * large[i].d) - f3(1, 1, 1));
    }
    res[i].offset = large[i].children_idx;
    pos[i] = large[i].p;
    size[i] = large[i].d;
  }
  for (unsigned i = 0; i < large.size(); i++) {
    if (large[i].children_idx != 0 || !is_border(large[i].value, (int)large[i].level)) continue;
    std::vector<FrameNode> sub(1);
    add_node_rec(sub, sdf, 0, large[i].level, depth, large[i].p, large[i].d);
    if (sub.size() == 1) continue;  // the subtree root stayed a leaf
    // the subtree's nodes after its root go to the end; its offsets shift by
    // the same amount (res[i] takes the root's place)
    const int shift = (int)res.size() - 1;
    res[i].offset = sub[0].offset + shift;
    std::vector<F3> spos(sub.size());
    std::vector<float> ssize(sub.size());
    // positions of the subtree's nodes, by walking it from the root
    std::vector<unsigned> stack{0};
    spos[0] = large[i].p;
    ssize[0] = large[i].d;
    while (!stack.empty()) {
      const unsigned k = stack.back();
      stack.pop_back();
      if (sub[k].offset == 0) continue;
      for (unsigned c = 0; c < 8; c++) {
        const F3 o = f3((float)((c & 4) >> 2), (float)((c & 2) >> 1), (float)(c & 1));
        spos[sub[k].offset + c] = 2 * spos[k] + o;
        ssize[sub[k].offset + c] = ssize[k] / 2;
        stack.push_back(sub[k].offset + c);
      }
    }
    for (size_t j = 1; j < sub.size(); j++) {
      FrameNode n = sub[j];
      if (n.offset != 0) n.offset += shift;
      res.push_back(n);
      pos.push_back(spos[j]);
      size.push_back(ssize[j]);
    }
  }
  return res;
}

constexpr uint32_t kInvalidIdx = 1u << 31;  // INVALID_IDX, sdf_scene.h:42

bool frame_leaf(uint32_t offset) { return offset == 0 || (offset & kInvalidIdx); }  // is_leaf, :43-46

struct LayerInfo {  // LayerFrameNodeInfo, :1255-1261
  uint32_t idx;
  bool leaf, border;
  uint32_t border_children;
};

// fill_layers_rec (:1270-1299): every reachable node, per level in depth-first
// order, with its border flag (is_border_node over its 8 values) and the
// number of border children.
bool fill_layers(const std::vector<FrameNode>& f, std::vector<std::vector<LayerInfo>>& layers, unsigned level, uint32_t idx) {
  if (level >= layers.size()) layers.resize(level + 1);
  const uint32_t ofs = f[idx].offset;
  const bool leaf = frame_leaf(ofs);
  float mn = f[idx].values[0], mx = f[idx].values[0];
  for (int i = 0; i < 8; i++) {
    mn = f[idx].values[i] < mn ? f[idx].values[i] : mn;
    mx = mx < f[idx].values[i] ? f[idx].values[i] : mx;
  }
  const bool border = is_border_node(mn, mx, 1u << level);
  uint32_t bc = 0;
  if (!leaf)
    for (uint32_t i = 0; i < 8; i++) bc += fill_layers(f, layers, level + 1, ofs + i);
  layers[level].push_back({idx, leaf, border, bc});
  return border;
}

float trilinear8(const float v[8], F3 dp) {  // trilinear_interpolation, :1301-1311
  return (1 - dp.x) * (1 - dp.y) * (1 - dp.z) * v[0] + (1 - dp.x) * (1 - dp.y) * (dp.z) * v[1] + (1 - dp.x) * (dp.y) * (1 - dp.z) * v[2] +
         (1 - dp.x) * (dp.y) * (dp.z) * v[3] + (dp.x) * (1 - dp.y) * (1 - dp.z) * v[4] + (dp.x) * (1 - dp.y) * (dp.z) * v[5] +
         (dp.x) * (dp.y) * (1 - dp.z) * v[6] + (dp.x) * (dp.y) * (dp.z) * v[7];
}

// frame_octree_limit_nodes (:2229-2330) with count_only_border_nodes: while
// more border leaves than nodes_limit remain, merge the children of the
// parents one level above the deepest layer still being merged, in order of
// their mean difference from the parent's trilinear interpolation (weights
// and unsigned counts as the reference computes them). The reference sorts
// with std::sort, which leaves equal differences in no specified order; here
// they keep their depth-first order (stable sort). pos / size follow the
// nodes through the final compaction.
void limit_nodes(std::vector<FrameNode>& f, std::vector<F3>& pos, std::vector<float>& size, uint32_t nodes_limit) {
  if (nodes_limit >= f.size()) return;
  struct Cand {
    uint32_t idx, active_children;
    float weighted_diff;
  };
  auto count = [&](const std::vector<std::vector<LayerInfo>>& layers) {
    uint32_t c = 0;
    for (const auto& layer : layers)
      for (const LayerInfo& n : layer) c += n.leaf && n.border;
    return c;
  };
  std::vector<std::vector<LayerInfo>> layers;
  fill_layers(f, layers, 0, 0);
  uint32_t cnt = count(layers);
  uint32_t active = (uint32_t)layers.size() - 1;
  while (cnt > nodes_limit && active > 0) {
    std::vector<Cand> mc;
    for (const LayerInfo& p : layers[active - 1]) {
      const uint32_t ofs = f[p.idx].offset;
      if (frame_leaf(ofs)) continue;
      float diff = 0;
      for (unsigned i = 0; i < 8; i++)
        for (unsigned j = 0; j < 8; j++) {
          const float child_val = f[ofs + i].values[j];
          const F3 q = 0.5f * (f3((float)((i & 4) >> 2), (float)((i & 2) >> 1), (float)(i & 1)) +
                               f3((float)((j & 4) >> 2), (float)((j & 2) >> 1), (float)(j & 1)));
          diff += std::abs(child_val - trilinear8(f[p.idx].values, q));
        }
      mc.push_back({p.idx, p.border_children, p.border_children > 1 ? diff / (float)(p.border_children - 1) : 1000.0f});
    }
    uint32_t potential = 0;
    for (const Cand& c : mc) potential += c.active_children - 1;  // unsigned, as the reference adds them
    if (cnt - potential < nodes_limit)
      std::stable_sort(mc.begin(), mc.end(), [](const Cand& a, const Cand& b) { return a.weighted_diff < b.weighted_diff; });
    for (size_t c = 0; cnt > nodes_limit && c < mc.size(); c++) {
      const uint32_t idx = mc[c].idx;
      for (uint32_t i = 0; i < 8; i++) f[f[idx].offset + i].offset = kInvalidIdx;
      f[idx].offset = 0;
      cnt -= mc[c].active_children - 1;
    }
    layers.clear();
    fill_layers(f, layers, 0, 0);
    cnt = count(layers);
    active--;
  }
  std::vector<uint32_t> remap(f.size(), 0);
  std::vector<FrameNode> nf;
  std::vector<F3> npos;
  std::vector<float> nsize;
  for (uint32_t i = 0; i < f.size(); i++)
    if (f[i].offset != kInvalidIdx) {
      remap[i] = (uint32_t)nf.size();
      nf.push_back(f[i]);
      npos.push_back(pos[i]);
      nsize.push_back(size[i]);
    }
  for (FrameNode& n : nf)
    if (n.offset > 0) n.offset = remap[n.offset];
  f.swap(nf);
  pos.swap(npos);
  size.swap(nsize);
}

}  // namespace

// frame_octree_to_SBS (:291-444) of the field's frame octree; see lrt_oracle.h.
extern "C" int orc_sbs_octree(int shape, const float* params, uint32_t depth, uint32_t brick_size, uint32_t brick_pad, uint32_t bytes_per_value,
                              uint32_t nodes_limit, uint32_t* n_nodes, uint32_t* nodes_out, uint64_t* n_values, uint32_t* values_out) {
  if ((shape != 0 && shape != 1) || !params || depth < 2 || depth > 12 || brick_size < 1 || brick_size > 16 || brick_pad > 1 ||
      (bytes_per_value != 1 && bytes_per_value != 2 && bytes_per_value != 4) || !n_nodes || !n_values)
    return -1;
  Field sdf{shape, {params[0], params[1], params[2], params[3]}};
  std::vector<F3> pos;
  std::vector<float> size;
  std::vector<FrameNode> nodes = frame_octree(sdf, depth, pos, size);
  limit_nodes(nodes, pos, size, nodes_limit ? nodes_limit : (1u << 24));  // SparseOctreeSettings::nodes_limit, default 2^24
  const uint32_t v_size = brick_size + 2 * brick_pad + 1;
  const unsigned bits = 8 * bytes_per_value;
  const unsigned q_max = bytes_per_value == 4 ? 0xFFFFFFFFu : ((1u << bits) - 1);  // the reference's inner max_val
  const unsigned vals_per_int = 4 / bytes_per_value;
  const uint32_t words = (v_size * v_size * v_size + vals_per_int - 1) / vals_per_int;
  std::vector<float> values(v_size * v_size * v_size);
  uint32_t nn = 0;
  uint64_t nv = 0;
  const int pad = (int)brick_pad, bs = (int)brick_size;
  for (size_t idx = 0; idx < nodes.size(); idx++) {
    if (nodes[idx].offset != 0) continue;
    const uint32_t px = (uint32_t)pos[idx].x, py = (uint32_t)pos[idx].y, pz = (uint32_t)pos[idx].z;
    const float d = size[idx];
    float min_val = 1000, max_val = -1000;
    for (int i = 0; i < 8; i++) {
      min_val = std::fmin(min_val, nodes[idx].values[i]);
      max_val = std::fmax(max_val, nodes[idx].values[i]);
    }
    const F3 p0 = 2.0f * (d * f3((float)px, (float)py, (float)pz)) - f3(1.0f, 1.0f, 1.0f);
    const float dp = 2.0f * d / (float)brick_size;
    for (int i = -pad; i <= bs + pad; i++)
      for (int j = -pad; j <= bs + pad; j++)
        for (int k = -pad; k <= bs + pad; k++) {
          float val = 2e6f;
          // corners reuse the frame node's values (:340-375)
          if ((i == 0 || i == bs) && (j == 0 || j == bs) && (k == 0 || k == bs))
            val = nodes[idx].values[(i == bs ? 4 : 0) + (j == bs ? 2 : 0) + (k == bs ? 1 : 0)];
          if (val > 1e6f) val = sdf(p0 + dp * f3((float)i, (float)j, (float)k));
          values[(uint32_t)(i + pad) * v_size * v_size + (uint32_t)(j + pad) * v_size + (uint32_t)(k + pad)] = val;
          min_val = std::fmin(min_val, val);
          max_val = std::fmax(max_val, val);
        }
    const unsigned lod_size = (unsigned)(1.0f / d);
    if (!is_border_node(min_val, max_val, lod_size)) continue;
    const float d_max = (float)(2 * std::sqrt(3.0) / lod_size);
    if (nodes_out) {
      uint32_t* o = nodes_out + 4 * (size_t)nn;
      o[0] = (px << 16) | py;
      o[1] = (pz << 16) | lod_size;
      o[2] = (uint32_t)nv;
      o[3] = 0;
    }
    if (values_out) {
      uint32_t* w = values_out + nv;
      std::memset(w, 0, words * 4);
      for (size_t i = 0; i < values.size(); i++) {
        // the reference's float expression, then std::min with q_max; its
        // float -> unsigned conversion of 2^32 and above (32-bit values at
        // the top of the range) is undefined, and those take q_max here
        const float q = std::fmax(0.0f, (float)q_max * ((values[i] + d_max) / (2 * d_max)));
        unsigned dc = q >= 4294967296.0f ? q_max : (unsigned)q;
        dc = dc < q_max ? dc : q_max;
        w[i / vals_per_int] |= dc << (bits * (i % vals_per_int));
      }
    }
    nn++;
    nv += words;
  }
  *n_nodes = nn;
  *n_values = nv;
  return 0;
}

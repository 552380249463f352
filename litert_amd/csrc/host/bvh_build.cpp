// bvh_build.cpp - binned-SAH BVH2 builder and fat-pair conversion.
#include "bvh_build.hpp"
#include <algorithm>
#include <cfloat>
#include <functional>

namespace lrt {

namespace {
constexpr int kBins = 32;

struct Aabb {
  f3 mn{FLT_MAX, FLT_MAX, FLT_MAX}, mx{-FLT_MAX, -FLT_MAX, -FLT_MAX};
  void grow(f3 a, f3 b) { mn = fmin3(mn, a); mx = fmax3(mx, b); }
  void grow(const Aabb& o) { mn = fmin3(mn, o.mn); mx = fmax3(mx, o.mx); }
  float area() const {
    if (mx.x < mn.x) return 0.f;
    f3 d = mx - mn;
    return 2.f * (d.x * d.y + d.y * d.z + d.z * d.x);
  }
};

struct Builder {
  const std::vector<f3>& pmin;
  const std::vector<f3>& pmax;
  std::vector<f3> cent;
  uint32_t max_leaf;
  BuildTree t;

  Builder(const std::vector<f3>& a, const std::vector<f3>& b, uint32_t ml) : pmin(a), pmax(b), max_leaf(ml) {
    cent.resize(a.size());
    for (size_t i = 0; i < a.size(); i++) cent[i] = 0.5f * (a[i] + b[i]);
  }

  int make(uint32_t start, uint32_t count, int depth) {
    int id = (int)t.nodes.size();
    t.nodes.emplace_back();
    Aabb box, cbox;
    for (uint32_t i = start; i < start + count; i++) {
      uint32_t p = t.prims[i];
      box.grow(pmin[p], pmax[p]);
      cbox.grow(cent[p], cent[p]);
    }
    t.nodes[id].bmin = box.mn;
    t.nodes[id].bmax = box.mx;
    t.max_depth = std::max(t.max_depth, depth);
    if (count <= max_leaf) {
      t.nodes[id].start = start;
      t.nodes[id].count = count;
      return id;
    }
    // binned SAH on the widest-splittable axes
    float best_cost = FLT_MAX;
    int best_axis = -1, best_bin = -1;
    for (int axis = 0; axis < 3; axis++) {
      float lo = cbox.mn[axis], hi = cbox.mx[axis];
      if (!(hi > lo)) continue;
      Aabb bins[kBins];
      uint32_t cnt[kBins] = {0};
      float scale = kBins / (hi - lo);
      for (uint32_t i = start; i < start + count; i++) {
        uint32_t p = t.prims[i];
        int b = std::min(kBins - 1, (int)((cent[p][axis] - lo) * scale));
        bins[b].grow(pmin[p], pmax[p]);
        cnt[b]++;
      }
      float rarea[kBins];
      uint32_t rcnt[kBins];
      Aabb acc;
      uint32_t c = 0;
      for (int b = kBins - 1; b > 0; b--) {
        acc.grow(bins[b]);
        c += cnt[b];
        rarea[b] = acc.area();
        rcnt[b] = c;
      }
      Aabb lacc;
      uint32_t lc = 0;
      for (int b = 0; b < kBins - 1; b++) {
        lacc.grow(bins[b]);
        lc += cnt[b];
        if (lc == 0 || rcnt[b + 1] == 0) continue;
        float cost = lacc.area() * lc + rarea[b + 1] * rcnt[b + 1];
        if (cost < best_cost) { best_cost = cost; best_axis = axis; best_bin = b; }
      }
    }
    uint32_t mid;
    if (best_axis < 0) {
      mid = start + count / 2;  // all centroids coincide: split the list in half
    } else {
      float lo = cbox.mn[best_axis], hi = cbox.mx[best_axis];
      float scale = kBins / (hi - lo);
      auto it = std::partition(t.prims.begin() + start, t.prims.begin() + start + count, [&](uint32_t p) {
        int b = std::min(kBins - 1, (int)((cent[p][best_axis] - lo) * scale));
        return b <= best_bin;
      });
      mid = (uint32_t)(it - t.prims.begin());
      if (mid == start || mid == start + count) mid = start + count / 2;
    }
    int l = make(start, mid - start, depth + 1);
    int r = make(mid, start + count - mid, depth + 1);
    t.nodes[id].left = l;
    t.nodes[id].right = r;
    return id;
  }
};
}  // namespace

BuildTree build_sah(const std::vector<f3>& pmin, const std::vector<f3>& pmax, uint32_t max_leaf) {
  Builder b(pmin, pmax, std::max<uint32_t>(1, max_leaf));
  uint32_t n = (uint32_t)pmin.size();
  b.t.prims.resize(n);
  for (uint32_t i = 0; i < n; i++) b.t.prims[i] = i;
  b.t.nodes.reserve(2 * (size_t)n);
  if (n) b.t.root = b.make(0, n, 0);
  return std::move(b.t);
}

static BVHNode to_node(f3 mn, f3 mx, uint32_t left) {
  BVHNode r;
  r.boxMin[0] = mn.x; r.boxMin[1] = mn.y; r.boxMin[2] = mn.z;
  r.boxMax[0] = mx.x; r.boxMax[1] = mx.y; r.boxMax[2] = mx.z;
  r.leftOffset = left;
  r.escapeIndex = 0;
  return r;
}

std::vector<BVHNodePair> to_fat_pairs(const BuildTree& t, const std::vector<f3>& pmin, const std::vector<f3>& pmax,
                                      int* max_pair_depth) {
  std::vector<BVHNodePair> out;
  int maxd = 0;
  if (t.root < 0) { if (max_pair_depth) *max_pair_depth = 0; return out; }
  const BuildNode& root = t.nodes[t.root];
  auto leaf_code = [&](const BuildNode& n) { return LRT_LEAF_BIT | ((1u << 24) & LRT_SIZE_MASK) | (t.prims[n.start] & LRT_START_MASK); };
  if (root.count > 0) {  // one primitive: the pair's second slot repeats its leaf and box
    // (a second test of the same brick can only repeat the first one's result)
    uint32_t p = t.prims[root.start];
    BVHNodePair pr;
    pr.left = to_node(pmin[p], pmax[p], leaf_code(root));
    pr.right = to_node(pmin[p], pmax[p], leaf_code(root));
    out.push_back(pr);
    if (max_pair_depth) *max_pair_depth = 1;
    return out;
  }
  out.reserve(t.nodes.size());
  // depth-first, left subtree first (TreeConverter2::ProcessBVHNode, cbvh_fat.cpp:100-126)
  std::function<uint32_t(int, int)> process = [&](int node, int depth) -> uint32_t {
    const BuildNode& n = t.nodes[node];
    maxd = std::max(maxd, depth);
    uint32_t idx = (uint32_t)out.size();
    out.emplace_back();
    const BuildNode& L = t.nodes[n.left];
    const BuildNode& R = t.nodes[n.right];
    uint32_t lo = L.count ? leaf_code(L) : 0, ro = R.count ? leaf_code(R) : 0;
    if (!L.count) lo = process(n.left, depth + 1);
    if (!R.count) ro = process(n.right, depth + 1);
    out[idx].left = to_node(L.bmin, L.bmax, lo);
    out[idx].right = to_node(R.bmin, R.bmax, ro);
    return idx;
  };
  process(t.root, 1);
  if (max_pair_depth) *max_pair_depth = maxd;
  return out;
}

void sbs_brick_boxes(const SdfSBSNode* nodes, uint32_t n, std::vector<f3>& bmin, std::vector<f3>& bmax) {
  bmin.resize(n);
  bmax.resize(n);
  for (uint32_t i = 0; i < n; i++) {
    float px = (float)(nodes[i].pos_xy >> 16), py = (float)(nodes[i].pos_xy & 0xFFFF);
    float pz = (float)(nodes[i].pos_z_lod_size >> 16), sz = (float)(nodes[i].pos_z_lod_size & 0xFFFF);
    f3 mn(-1.0f + (2.0f * px) / sz, -1.0f + (2.0f * py) / sz, -1.0f + (2.0f * pz) / sz);
    float e = 2.0f / sz;
    bmin[i] = mn;
    bmax[i] = f3(mn.x + e, mn.y + e, mn.z + e);
  }
}

}  // namespace lrt

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
#include <unordered_map>
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

// The mesh and its distance-query data (triangle BVH, face / vertex / edge
// pseudo-normals) on the device of r; buffers freed with the struct.
struct MeshUpload {
  DevBuf v, tri, fn, vn, en, nodes, prims;
  lrtd::DevMesh M{};
  ~MeshUpload() {
    DevBuf* all[] = {&v, &tri, &fn, &vn, &en, &nodes, &prims};
    for (DevBuf* b : all) b->release();
  }
};

int upload_mesh(Renderer* r, const float* vertices, uint32_t n_vertices, const uint32_t* indices, uint32_t n_triangles, MeshUpload& U) {
  MeshSdfData D;
  try {
    D = mesh_sdf_data(to_mesh(vertices, n_vertices, indices, n_triangles));
  } catch (const std::exception& ex) {
    return set_err(LRT_E_INVALID, ex.what());
  }
  if (D.tree.max_depth + 2 > lrtd::kBuildStack) return set_err(LRT_E_UNSUPPORTED, "triangle BVH too deep for the device builder");
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
  hipStream_t s = r->stream;
  hipError_t e;
  if ((e = U.v.upload(v4.data(), v4.size() * 16, s)) != hipSuccess || (e = U.tri.upload(tri.data(), tri.size() * 16, s)) != hipSuccess ||
      (e = U.fn.upload(fn.data(), fn.size() * 16, s)) != hipSuccess || (e = U.vn.upload(vn.data(), vn.size() * 16, s)) != hipSuccess ||
      (e = U.en.upload(en.data(), en.size() * 16, s)) != hipSuccess ||
      (e = U.nodes.upload(nodes.data(), nodes.size() * 16, s)) != hipSuccess ||
      (e = U.prims.upload(D.tree.prims.data(), D.tree.prims.size() * 4, s)) != hipSuccess)
    return hip_err(e, "upload mesh");
  // the staging vectors die with this call
  if ((e = hipStreamSynchronize(s)) != hipSuccess) return hip_err(e, "upload mesh");
  U.M = lrtd::DevMesh{(const float4*)U.v.p, (const uint4*)U.tri.p, (const float4*)U.fn.p, (const float4*)U.vn.p, (const float4*)U.en.p,
                      (const uint4*)U.nodes.p, (const uint32_t*)U.prims.p, D.tree.root};
  return LRT_OK;
}

// The frame octree on the host, for frame_octree_limit_nodes: node k has its
// cell (x, y, z, level), its 8 corner values and the index of its first child
// (0: leaf; the 8 children are consecutive).
struct FrameNodeH {
  uint4 cell;
  float v[8];
  uint32_t first;
};
constexpr uint32_t kMerged = 1u << 31;  // a child of a merged node (INVALID_IDX, sdf_scene.h:42)

bool frame_is_leaf(uint32_t first) { return first == 0 || (first & kMerged); }

// is_border_node (sparse_octree_builder.cpp:49-53) of a frame node at `level`
bool frame_border(const float* v, uint32_t level) {
  float mn = v[0], mx = v[0];
  for (int i = 0; i < 8; i++) {
    mn = v[i] < mn ? v[i] : mn;
    mx = mx < v[i] ? v[i] : mx;
  }
  const float d_max = (float)(2 * std::sqrt(3.0) / (double)(1u << level));
  return mn <= 0 && mx > -d_max;
}

struct LayerNode {
  uint32_t idx;
  bool leaf, border;
  uint32_t border_children;
};

// fill_layers_rec (:1270-1299) without recursion: per level, the reachable
// nodes in depth-first order (children 0..7), each with its border flag and
// its count of border children.
void frame_layers(const std::vector<FrameNodeH>& f, std::vector<std::vector<LayerNode>>& layers) {
  layers.clear();
  struct Item {
    uint32_t idx, level, next_child, bc;
  };
  std::vector<Item> stack{{0, 0, 0, 0}};
  while (!stack.empty()) {
    Item& it = stack.back();
    const FrameNodeH& n = f[it.idx];
    const bool leaf = frame_is_leaf(n.first);
    if (!leaf && it.next_child < 8) {
      const uint32_t c = n.first + it.next_child++;
      stack.push_back({c, it.level + 1, 0, 0});
      continue;
    }
    const bool border = frame_border(n.v, it.level);
    if (layers.size() <= it.level) layers.resize(it.level + 1);
    layers[it.level].push_back({it.idx, leaf, border, it.bc});
    stack.pop_back();
    if (!stack.empty()) stack.back().bc += border ? 1u : 0u;
  }
}

// frame_octree_limit_nodes (:2229-2330), count_only_border_nodes = true:
// merges the children of the nodes one level above the layer being reduced,
// in order of their mean |child corner - parent trilinear| over (border
// children - 1), until at most nodes_limit border leaves remain. Equal keys
// keep their depth-first order (the reference's std::sort leaves them in no
// specified order). Returns the leaves.
std::vector<uint32_t> frame_limit(std::vector<FrameNodeH>& f, uint32_t nodes_limit) {
  std::vector<std::vector<LayerNode>> layers;
  frame_layers(f, layers);
  auto count = [&]() {
    uint32_t c = 0;
    for (const auto& layer : layers)
      for (const LayerNode& n : layer) c += n.leaf && n.border;
    return c;
  };
  auto tri = [](const float* v, float x, float y, float z) {  // trilinear_interpolation, :1301-1311
    return (1 - x) * (1 - y) * (1 - z) * v[0] + (1 - x) * (1 - y) * (z) * v[1] + (1 - x) * (y) * (1 - z) * v[2] + (1 - x) * (y) * (z) * v[3] +
           (x) * (1 - y) * (1 - z) * v[4] + (x) * (1 - y) * (z) * v[5] + (x) * (y) * (1 - z) * v[6] + (x) * (y) * (z) * v[7];
  };
  uint32_t cnt = count();
  uint32_t active = (uint32_t)layers.size() - 1;
  struct Cand {
    uint32_t idx, active_children;
    float key;
  };
  while (cnt > nodes_limit && active > 0) {
    std::vector<Cand> mc;
    for (const LayerNode& p : layers[active - 1]) {
      const uint32_t first = f[p.idx].first;
      if (frame_is_leaf(first)) continue;
      float diff = 0;
      for (uint32_t i = 0; i < 8; i++)
        for (uint32_t j = 0; j < 8; j++) {
          const float qx = 0.5f * ((float)((i & 4) >> 2) + (float)((j & 4) >> 2));
          const float qy = 0.5f * ((float)((i & 2) >> 1) + (float)((j & 2) >> 1));
          const float qz = 0.5f * ((float)(i & 1) + (float)(j & 1));
          diff += std::abs(f[first + i].v[j] - tri(f[p.idx].v, qx, qy, qz));
        }
      mc.push_back({p.idx, p.border_children, p.border_children > 1 ? diff / (float)(p.border_children - 1) : 1000.0f});
    }
    uint32_t potential = 0;
    for (const Cand& c : mc) potential += c.active_children - 1;  // unsigned, as the reference sums them
    if (cnt - potential < nodes_limit)
      std::stable_sort(mc.begin(), mc.end(), [](const Cand& a, const Cand& b) { return a.key < b.key; });
    for (size_t c = 0; cnt > nodes_limit && c < mc.size(); c++) {
      for (uint32_t i = 0; i < 8; i++) f[f[mc[c].idx].first + i].first = kMerged;
      f[mc[c].idx].first = 0;
      cnt -= mc[c].active_children - 1;
    }
    frame_layers(f, layers);
    cnt = count();
    active--;
  }
  std::vector<uint32_t> leaves;
  for (const auto& layer : layers)
    for (const LayerNode& n : layer)
      if (n.leaf) leaves.push_back(n.idx);
  return leaves;
}

// create_sdf_SBS(settings, header, sdf) (utils/sdf_converter.cpp:169-181) on
// the device: the frame octree of construct_sdf_frame_octree
// (sparse_octree_builder.cpp:125-258) level by level -- every cell of the last
// uniform level min(depth, 4), then the children of every split cell -- and
// frame_octree_to_SBS (:291-444) over its leaves. Bricks are ordered by LOD,
// then Morton order (the reference appends them from parallel threads).
int build_octree(Renderer* r, const lrtd::DevField& F, uint32_t depth, uint32_t brick_size, uint32_t brick_pad, uint32_t bytes_per_value,
                 uint32_t nodes_limit, lrt_sbs_handle* out) {
  if (nodes_limit == 0) nodes_limit = 1u << 24;  // SparseOctreeSettings::nodes_limit default, sdf_converter.h:32
  if (depth < 2 || depth > 12) return set_err(LRT_E_INVALID, "octree depth must be 2..12");
  if (brick_size < 1 || brick_size > 16 || brick_pad > 1) return set_err(LRT_E_INVALID, "bad brick size/pad");
  if (bytes_per_value != 1 && bytes_per_value != 2 && bytes_per_value != 4) return set_err(LRT_E_INVALID, "bad bytes_per_value");
  lrtd::OctreeLevels T{};
  for (int l = 0; l < 16; l++) {
    const double b = std::sqrt(3.0) * std::pow(2.0, -l);  // is_border, :38-41, in double
    float bf = (float)b;
    if ((double)bf < b) bf = std::nextafter(bf, INFINITY);  // |v| < b  <=>  |v| < bf for float v
    T.border[l] = bf;
    T.near[l] = std::sqrt(3.0f) * std::ldexp(1.0f, -l);                  // add_node_rec, :84
    T.keep[l] = (float)(2 * std::sqrt(3.0) / (double)(1u << l));        // is_border_node :51, quantiser :401
  }
  hipStream_t s = r->stream;
  hipError_t e;
  DevBuf b_cells, b_kids, b_leaves, b_lvals, b_count, b_fill, b_keep, b_words;
  struct Release {
    std::vector<DevBuf*> bufs;
    ~Release() {
      for (DevBuf* b : bufs) b->release();
    }
  } rel{{&b_cells, &b_kids, &b_leaves, &b_lvals, &b_count, &b_fill, &b_keep, &b_words}};
  const uint32_t top = depth < 4 ? depth : 4;
  const uint32_t side = 1u << top;
  std::vector<uint4> cells;
  cells.reserve((size_t)side * side * side);
  for (uint32_t x = 0; x < side; x++)
    for (uint32_t y = 0; y < side; y++)
      for (uint32_t z = 0; z < side; z++) cells.push_back(make_uint4(x, y, z, top));
  uint64_t frame_nodes = 0;
  for (uint32_t l = 0; l <= top; l++) frame_nodes += 1ull << (3 * l);
  if ((e = b_cells.upload(cells.data(), cells.size() * 16, s)) != hipSuccess) return hip_err(e, "upload cells");
  if ((e = b_count.ensure(8)) != hipSuccess) return hip_err(e, "alloc");
  std::vector<uint4> leaves;
  std::vector<float> lvals;
  // Every level's cells and field values stay on the device until the frame's
  // size is known: frame_octree_limit_nodes needs the whole frame.
  std::vector<DevBuf> lvl_cells, lvl_vals;
  std::vector<uint32_t> lvl_n;
  struct ReleaseLevels {
    std::vector<DevBuf>* a;
    std::vector<DevBuf>* b;
    ~ReleaseLevels() {
      for (auto* v : {a, b})
        for (DevBuf& d : *v) d.release();
    }
  } rel_levels{&lvl_cells, &lvl_vals};
  uint32_t n = (uint32_t)cells.size();
  DevBuf cur = b_cells;  // the level's cells; ownership moves into lvl_cells
  b_cells = DevBuf{};
  for (uint32_t level = top; n > 0; level++) {
    lvl_cells.push_back(cur);
    cur = DevBuf{};
    lvl_vals.push_back(DevBuf{});
    lvl_n.push_back(n);
    DevBuf& vals = lvl_vals.back();
    const DevBuf& cl = lvl_cells.back();
    if ((e = vals.ensure(9ull * n * 4)) != hipSuccess || (e = b_kids.ensure(8ull * n * 16)) != hipSuccess ||
        (e = b_leaves.ensure((size_t)n * 16)) != hipSuccess || (e = b_lvals.ensure(8ull * n * 4)) != hipSuccess)
      return hip_err(e, "alloc octree level");
    if ((e = zero_u32((uint32_t*)b_count.p, 2, s)) != hipSuccess) return hip_err(e, "clear");
    if ((e = launch_octree_eval(F, (const uint4*)cl.p, n, (float*)vals.p, s)) != hipSuccess) return hip_err(e, "octree eval");
    if ((e = launch_octree_classify((const uint4*)cl.p, (const float*)vals.p, n, level, level == top ? 1 : 0, depth, T, (uint4*)b_kids.p,
                                    (uint32_t*)b_count.p, (uint4*)b_leaves.p, (float*)b_lvals.p, (uint32_t*)b_count.p + 1, s)) != hipSuccess)
      return hip_err(e, "octree classify");
    uint32_t cnt[2];
    if ((e = hipMemcpyAsync(cnt, b_count.p, 8, hipMemcpyDeviceToHost, s)) != hipSuccess || (e = hipStreamSynchronize(s)) != hipSuccess)
      return hip_err(e, "octree counts");
    const size_t l0 = leaves.size();
    leaves.resize(l0 + cnt[1]);
    lvals.resize(8 * (l0 + cnt[1]));
    if (cnt[1] && ((e = hipMemcpyAsync(leaves.data() + l0, b_leaves.p, (size_t)cnt[1] * 16, hipMemcpyDeviceToHost, s)) != hipSuccess ||
                   (e = hipMemcpyAsync(lvals.data() + 8 * l0, b_lvals.p, (size_t)cnt[1] * 32, hipMemcpyDeviceToHost, s)) != hipSuccess))
      return hip_err(e, "download leaves");
    frame_nodes += cnt[0];
    if (frame_nodes > (1ull << 26)) return set_err(LRT_E_UNSUPPORTED, "frame octree above 2^26 nodes");
    cur = b_kids;  // the children are the next level's cells
    b_kids = DevBuf{};
    n = cnt[0];
  }
  cur.release();
  if ((e = hipStreamSynchronize(s)) != hipSuccess) return hip_err(e, "download leaves");
  if (frame_nodes > nodes_limit) {
    // frame_octree_limit_nodes (:2229-2330) over the whole frame, on the host:
    // the uniform top levels are evaluated now, every other level comes from
    // the device, and the tree is rebuilt from cell positions
    std::vector<std::vector<uint4>> hc(depth + 1);
    std::vector<std::vector<float>> hv(depth + 1);
    std::vector<uint4> upper;
    for (uint32_t l = 0; l < top; l++)
      for (uint32_t x = 0; x < (1u << l); x++)
        for (uint32_t y = 0; y < (1u << l); y++)
          for (uint32_t z = 0; z < (1u << l); z++) upper.push_back(make_uint4(x, y, z, l));
    std::vector<float> upper_v(9 * upper.size());
    DevBuf b_uc, b_uv;
    struct ReleaseUpper {
      DevBuf* a;
      DevBuf* b;
      ~ReleaseUpper() {
        a->release();
        b->release();
      }
    } rel_upper{&b_uc, &b_uv};
    if ((e = b_uc.upload(upper.data(), upper.size() * 16, s)) != hipSuccess || (e = b_uv.ensure(upper_v.size() * 4)) != hipSuccess ||
        (e = launch_octree_eval(F, (const uint4*)b_uc.p, (uint32_t)upper.size(), (float*)b_uv.p, s)) != hipSuccess ||
        (e = hipMemcpyAsync(upper_v.data(), b_uv.p, upper_v.size() * 4, hipMemcpyDeviceToHost, s)) != hipSuccess)
      return hip_err(e, "octree top levels");
    for (size_t i = 0; i < upper.size(); i++) {
      hc[upper[i].w].push_back(upper[i]);
      hv[upper[i].w].insert(hv[upper[i].w].end(), upper_v.begin() + 9 * i, upper_v.begin() + 9 * i + 9);
    }
    for (size_t k = 0; k < lvl_cells.size(); k++) {
      const uint32_t l = top + (uint32_t)k;
      hc[l].resize(lvl_n[k]);
      hv[l].resize(9ull * lvl_n[k]);
      if ((e = hipMemcpyAsync(hc[l].data(), lvl_cells[k].p, (size_t)lvl_n[k] * 16, hipMemcpyDeviceToHost, s)) != hipSuccess ||
          (e = hipMemcpyAsync(hv[l].data(), lvl_vals[k].p, 9ull * lvl_n[k] * 4, hipMemcpyDeviceToHost, s)) != hipSuccess)
        return hip_err(e, "download frame");
    }
    if ((e = hipStreamSynchronize(s)) != hipSuccess) return hip_err(e, "download frame");
    auto key = [](uint32_t x, uint32_t y, uint32_t z) { return (uint64_t)x | ((uint64_t)y << 21) | ((uint64_t)z << 42); };
    std::vector<std::unordered_map<uint64_t, uint32_t>> at(depth + 1);
    for (uint32_t l = 0; l <= depth; l++) {
      at[l].reserve(hc[l].size());
      for (uint32_t i = 0; i < hc[l].size(); i++) at[l][key(hc[l][i].x, hc[l][i].y, hc[l][i].z)] = i;
    }
    std::vector<FrameNodeH> fr;
    fr.reserve(frame_nodes);
    auto node = [&](uint32_t l, uint32_t i) {
      FrameNodeH nd;
      nd.cell = hc[l][i];
      std::copy(hv[l].begin() + 9ull * i, hv[l].begin() + 9ull * i + 8, nd.v);
      nd.first = 0;
      return nd;
    };
    fr.push_back(node(0, 0));
    for (size_t k = 0; k < fr.size(); k++) {  // children appended as a block of 8, in creation order
      const uint4 c = fr[k].cell;
      if (c.w >= depth) continue;
      auto it = at[c.w + 1].find(key(2 * c.x, 2 * c.y, 2 * c.z));
      if (it == at[c.w + 1].end()) continue;  // not split
      fr[k].first = (uint32_t)fr.size();
      for (uint32_t q = 0; q < 8; q++) {
        auto jt = at[c.w + 1].find(key(2 * c.x + ((q >> 2) & 1u), 2 * c.y + ((q >> 1) & 1u), 2 * c.z + (q & 1u)));
        if (jt == at[c.w + 1].end()) return set_err(LRT_E_HIP, "octree frame: a split cell lacks a child");
        fr.push_back(node(c.w + 1, jt->second));
      }
    }
    leaves.clear();
    lvals.clear();
    for (uint32_t i : frame_limit(fr, nodes_limit)) {
      leaves.push_back(fr[i].cell);
      lvals.insert(lvals.end(), fr[i].v, fr[i].v + 8);
    }
  }
  const uint32_t nl = (uint32_t)leaves.size();
  std::vector<uint32_t> order(nl);
  for (uint32_t i = 0; i < nl; i++) order[i] = i;
  std::sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
    if (leaves[a].w != leaves[b].w) return leaves[a].w < leaves[b].w;
    return morton3(leaves[a].x, leaves[a].y, leaves[a].z) < morton3(leaves[b].x, leaves[b].y, leaves[b].z);
  });
  std::vector<uint4> sl(nl);
  std::vector<float> sv(8ull * nl);
  for (uint32_t i = 0; i < nl; i++) {
    sl[i] = leaves[order[i]];
    std::copy(lvals.begin() + 8ull * order[i], lvals.begin() + 8ull * order[i] + 8, sv.begin() + 8ull * i);
  }
  const uint32_t v_size = brick_size + 2 * brick_pad + 1, nv = v_size * v_size * v_size;
  const uint32_t vals_per_int = 4 / bytes_per_value, n_words = (nv + vals_per_int - 1) / vals_per_int;
  if ((e = b_leaves.upload(sl.data(), (size_t)nl * 16, s)) != hipSuccess || (e = b_lvals.upload(sv.data(), sv.size() * 4, s)) != hipSuccess ||
      (e = b_fill.ensure((size_t)nl * nv * 4)) != hipSuccess || (e = b_keep.ensure((size_t)nl * 4)) != hipSuccess ||
      (e = b_words.ensure((size_t)nl * n_words * 4)) != hipSuccess)
    return hip_err(e, "alloc bricks");
  if ((e = launch_octree_fill(F, (const uint4*)b_leaves.p, (const float*)b_lvals.p, nl, brick_size, brick_pad, (float*)b_fill.p, s)) !=
      hipSuccess)
    return hip_err(e, "octree fill");
  if ((e = launch_octree_pack((const uint4*)b_leaves.p, (const float*)b_fill.p, nl, nv, bytes_per_value, T, (uint32_t*)b_keep.p,
                              (uint32_t*)b_words.p, s)) != hipSuccess)
    return hip_err(e, "octree pack");
  std::vector<uint32_t> keep(nl), words((size_t)nl * n_words);
  if ((e = hipMemcpyAsync(keep.data(), b_keep.p, (size_t)nl * 4, hipMemcpyDeviceToHost, s)) != hipSuccess ||
      (e = hipMemcpyAsync(words.data(), b_words.p, words.size() * 4, hipMemcpyDeviceToHost, s)) != hipSuccess ||
      (e = hipStreamSynchronize(s)) != hipSuccess)
    return hip_err(e, "download bricks");
  std::unique_ptr<lrt_sbs> sb(new lrt_sbs);
  SbsScene& S = sb->s;
  S.header.brick_size = brick_size;
  S.header.brick_pad = brick_pad;
  S.header.bytes_per_value = bytes_per_value;
  S.header.aux_data = LRT_SBS_LAYOUT_DX;  // values over the brick diagonal, as the reference quantises
  for (uint32_t b = 0; b < nl; b++) {
    if (!keep[b]) continue;
    SdfSBSNode nd;
    nd.pos_xy = (sl[b].x << 16) | sl[b].y;
    nd.pos_z_lod_size = (sl[b].z << 16) | (1u << sl[b].w);
    nd.data_offset = (uint32_t)S.values.size();
    nd._pad = 0;
    S.nodes.push_back(nd);
    S.values.insert(S.values.end(), words.begin() + (size_t)b * n_words, words.begin() + (size_t)(b + 1) * n_words);
  }
  if (S.nodes.empty()) return set_err(LRT_E_INVALID, "scene produced no bricks");
  *out = sb.release();
  return LRT_OK;
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
  MeshUpload U;
  if ((c = upload_mesh(r, vertices, n_vertices, indices, n_triangles, U)) != LRT_OK) return c;
  const lrtd::DevMesh& M = U.M;
  DevBuf b_cur, b_next, b_count, b_vals, b_keep, b_words;
  struct Release {
    std::vector<DevBuf*> bufs;
    ~Release() {
      for (DevBuf* b : bufs) b->release();
    }
  } rel{{&b_cur, &b_next, &b_count, &b_vals, &b_keep, &b_words}};
  hipStream_t s = r->stream;

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


int lrt_sbs_octree_from_mesh_device(lrt_handle h, const float* vertices, uint32_t n_vertices, const uint32_t* indices, uint32_t n_triangles,
                                    uint32_t depth, uint32_t brick_size, uint32_t brick_pad, uint32_t bytes_per_value, uint32_t nodes_limit,
                                    lrt_sbs_handle* out, float* build_ms) {
  if (!h) return set_err(LRT_E_INVALID, "null handle");
  int c = check_mesh_args(vertices, n_vertices, indices, n_triangles, 1, brick_size, brick_pad, bytes_per_value, out);
  if (c) return c;
  Renderer* r = h;
  hipError_t e;
  if ((e = hipSetDevice(r->device)) != hipSuccess) return hip_err(e, "hipSetDevice");
  const auto t0 = std::chrono::steady_clock::now();
  MeshUpload U;
  if ((c = upload_mesh(r, vertices, n_vertices, indices, n_triangles, U)) != LRT_OK) return c;
  lrtd::DevField F{};
  F.kind = 0;
  F.mesh = U.M;
  if ((c = build_octree(r, F, depth, brick_size, brick_pad, bytes_per_value, nodes_limit, out)) != LRT_OK) return c;
  if (build_ms) *build_ms = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return LRT_OK;
}

int lrt_sbs_octree_from_primitive_device(lrt_handle h, uint32_t shape, const float* params, uint32_t depth, uint32_t brick_size,
                                         uint32_t brick_pad, uint32_t bytes_per_value, uint32_t nodes_limit, lrt_sbs_handle* out,
                                         float* build_ms) {
  if (!h || !params || !out) return set_err(LRT_E_INVALID, "null argument");
  if (shape != LRT_PRIMITIVE_SPHERE && shape != LRT_PRIMITIVE_TORUS) return set_err(LRT_E_INVALID, "unknown primitive");
  Renderer* r = h;
  hipError_t e;
  if ((e = hipSetDevice(r->device)) != hipSuccess) return hip_err(e, "hipSetDevice");
  const auto t0 = std::chrono::steady_clock::now();
  lrtd::DevField F{};
  F.kind = shape == LRT_PRIMITIVE_SPHERE ? 1u : 2u;
  for (int k = 0; k < (shape == LRT_PRIMITIVE_SPHERE ? 4 : 2); k++) F.a[k] = params[k];
  int c = build_octree(r, F, depth, brick_size, brick_pad, bytes_per_value, nodes_limit, out);
  if (c) return c;
  if (build_ms) *build_ms = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return LRT_OK;
}

}  // extern "C"

// bvh_build.hpp - binned-SAH BVH2 builder.
//
// Replaces the reference's Embree-based builder (BVH/cbvh.cpp:591-658,
// cbvh_embree2.cpp:287-331), which is not available. Output keeps the
// reference "fat" BVH2 format consumed by BVH2TraverseF32: an array of
// BVHNodePair in depth-first order with the root pair at index 0, internal
// children pointing at their pair index and leaves encoded as
// LEAF_BIT | count << 24 | primitive (cbvh_fat.cpp:100-142, raytrace_common.h:48-57).
#pragma once
#include <cstdint>
#include <vector>
#include "hmath.hpp"
#include "../../../include/lrt_types.h"

namespace lrt {

struct BuildNode {
  f3 bmin, bmax;
  int left = -1, right = -1;   // child node indices (internal)
  uint32_t start = 0, count = 0;  // primitive range in BuildTree::prims (leaf when count > 0)
};

struct BuildTree {
  std::vector<BuildNode> nodes;
  std::vector<uint32_t> prims;
  int root = -1;
  int max_depth = 0;
};

// Binned SAH over primitive boxes; leaves hold at most max_leaf primitives.
BuildTree build_sah(const std::vector<f3>& pmin, const std::vector<f3>& pmax, uint32_t max_leaf);

// One leaf per primitive -> fat pairs. A single primitive gets a pair whose
// right child is an empty (inverted) box, which no ray can enter.
std::vector<BVHNodePair> to_fat_pairs(const BuildTree& t, const std::vector<f3>& pmin, const std::vector<f3>& pmax,
                                      int* max_pair_depth);

// Brick AABBs exactly as AddGeom_SdfSBS computes them (BVH2Common_host.cpp:644-653).
void sbs_brick_boxes(const SdfSBSNode* nodes, uint32_t n, std::vector<f3>& bmin, std::vector<f3>& bmax);

}  // namespace lrt

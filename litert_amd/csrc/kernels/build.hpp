// build.hpp - launch interface of the device SBS builder (build.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lrtd {

// Per-lane stack of the nearest-triangle search; the host refuses trees whose
// depth could overflow it.
constexpr int kBuildStack = 64;

// A closed triangle mesh and what its distance query reads, on the device.
struct DevMesh {
  const float4* v;      // vertex positions (w unused)
  const uint4* tri;     // vertex indices a, b, c (w unused)
  const float4* fnorm;  // unit face normal per triangle
  const float4* vnorm;  // angle-weighted pseudo-normal per vertex
  const float4* enorm;  // edge pseudo-normals per triangle: ab, bc, ca
  const uint4* nodes;   // triangle BVH, 3 x uint4 per node: {bmin, left}, {bmax, right}, {start, count, -, -}
  const uint32_t* prims;
  int root;
};

// The distance field of the octree builder: the mesh above (kind 0), a
// sphere (kind 1: a = centre xyz, radius) or a torus around the z axis
// (kind 2: a = R, r).
struct DevField {
  uint32_t kind;
  float a[4];
  DevMesh mesh;
};

// Per-level thresholds of the octree builder, computed on the host:
// border[l] = the smallest float >= sqrt(3) * 2^-l (the reference compares a
// float distance with that double), near[l] = sqrtf(3) * 2^-l (float, as
// add_node_rec evaluates it), keep[l] = float(2 sqrt(3) / 2^l) (the brick
// diagonal of is_border_node and the quantiser).
struct OctreeLevels {
  float border[16], near[16], keep[16];
};

}  // namespace lrtd

namespace lrt {
// The octree builder (construct_sdf_frame_octree + frame_octree_to_SBS,
// utils/sparse_octree_builder.cpp:125-444), one level of cells at a time.
// vals[9 c + k]: the field at corner k (bit order x4 y2 z1) and, k = 8, the
// centre of cell c = {x, y, z, level}.
hipError_t launch_octree_eval(const lrtd::DevField& F, const uint4* cells, uint32_t n, float* vals, hipStream_t s);
// Cells of one level split into children (kids, 8 per split cell, *n_kids
// counts them) or become leaves (leaves: cell, leaf_vals: its 8 corner values;
// *n_leaves counts them). top: the level is the last uniform one, whose cells
// split only if their centre passes the border test.
hipError_t launch_octree_classify(const uint4* cells, const float* vals, uint32_t n, uint32_t level, int top, uint32_t depth,
                                  const lrtd::OctreeLevels& T, uint4* kids, uint32_t* n_kids, uint4* leaves, float* leaf_vals,
                                  uint32_t* n_leaves, hipStream_t s);
// Lattice values of every leaf (corners reuse the leaf's corner values).
hipError_t launch_octree_fill(const lrtd::DevField& F, const uint4* leaves, const float* leaf_vals, uint32_t n, uint32_t brick_size,
                              uint32_t brick_pad, float* vals, hipStream_t s);
// keep[b] (is_border_node) and the packed words of kept bricks, quantised with
// the reference's float expression over the brick diagonal.
hipError_t launch_octree_pack(const uint4* leaves, const float* vals, uint32_t n, uint32_t nv, uint32_t bytes_per_value,
                              const lrtd::OctreeLevels& T, uint32_t* keep, uint32_t* words, hipStream_t s);
// next[0 .. *count) = the kept children of the n cells in cur (cell side `side`)
hipError_t launch_build_refine(const lrtd::DevMesh& M, const uint4* cur, uint32_t n, float side, float r, float dmax_keep, uint4* next,
                               uint32_t* count, hipStream_t s);
// vals[b * v^3 + s] = distance at lattice point s of candidate brick b
hipError_t launch_build_fill(const lrtd::DevMesh& M, const uint4* cells, uint32_t n, uint32_t lod, uint32_t brick_size, uint32_t brick_pad,
                             float* vals, hipStream_t s);
// keep[b] and, for kept bricks, the packed quantised words of brick b
hipError_t launch_build_pack(const float* vals, uint32_t n, uint32_t nv, uint32_t bytes_per_value, float dmax_keep, double d_max,
                             uint32_t* keep, uint32_t* words, hipStream_t s);
}  // namespace lrt

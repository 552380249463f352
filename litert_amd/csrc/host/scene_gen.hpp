// scene_gen.hpp - synthetic SBS scene generators and the SBS brick builder.
//
// The reference builds SBS scenes from meshes or distance functions on the
// CPU (utils/sparse_octree_builder.cpp:38-53, 291-444; DR_common.cpp:47-146).
// These builders keep its brick format, quantiser and border-keep rule; the
// scene content (displaced icosphere, noisy torus) has no reference
// counterpart and is defined here.
#pragma once
#include <cstdint>
#include <functional>
#include <vector>
#include "hmath.hpp"
#include "bvh_build.hpp"
#include "../../../include/lrt_types.h"

namespace lrt {

struct SbsScene {
  SdfSBSHeader header{};
  std::vector<SdfSBSNode> nodes;
  std::vector<uint32_t> values;
  std::vector<float> values_f;
};

struct TriMesh {
  std::vector<f3> v;
  std::vector<uint32_t> idx;  // 3 per triangle
};

// Seeded 3-D value noise (hashed lattice values in [-1,1], quintic fade),
// summed over octaves with lacunarity 2 and gain 0.5, normalised to [-1,1].
float value_noise_fbm(f3 p, uint32_t seed, int octaves, float base_freq);

// Icosphere with `subdiv` midpoint subdivisions (20*4^subdiv triangles) of
// radius r, each vertex moved along its normal by amp * fbm(vertex).
TriMesh displaced_icosphere(int subdiv, float r, float amp, int octaves, uint32_t seed);

// What the mesh distance query reads: the triangle BVH (binned SAH, at most 4
// triangles per leaf), unit face normals, angle-weighted vertex
// pseudo-normals and per-triangle edge pseudo-normals (sum of the adjacent
// face normals) in the order ab, bc, ca.
struct MeshSdfData {
  TriMesh mesh;
  BuildTree tree;
  std::vector<f3> fnorm, vnorm, enorm3;
};
MeshSdfData mesh_sdf_data(const TriMesh& m);

// Signed distance to a closed triangle mesh: exact Euclidean distance through
// a BVH over triangles; sign from the angle-weighted pseudo-normal of the
// closest feature (Baerentzen & Aanaes 2005).
class MeshSdf {
 public:
  explicit MeshSdf(const TriMesh& m);
  ~MeshSdf();
  float operator()(f3 p) const;
  const MeshSdfData& data() const;

 private:
  struct Impl;
  Impl* impl_;
};

// Brick builder. Bricks of `brick_size` voxels at brick LOD `lod` (bricks per
// axis) cover [-1,1]^3. A brick is kept when the reference's border rule holds
// for its sampled values (sparse_octree_builder.cpp:49-53). layout:
//   LRT_SBS_LAYOUT_DX with bytes_per_value 1/2/4 -> quantised values
//   (sparse_octree_builder.cpp:405-428); voxel_diag_scale selects the voxel
//   diagonal instead of the brick diagonal as quantisation range (aux_data bit 0).
SbsScene build_sbs(const std::function<float(f3)>& sdf, uint32_t lod, uint32_t brick_size, uint32_t brick_pad,
                   uint32_t bytes_per_value, bool voxel_diag_scale, int threads);

// Dense indexed grid of brick_count^3 bricks with one shared float per lattice
// point and 8 colour indices per brick (layout ID32F_IRGB32F, as
// DR_common.cpp:47-146 builds); with_neighbors appends the 27 neighbour
// indices of layout ID32F_IRGB32F_IN.
SbsScene build_grid_sbs(uint32_t brick_count, uint32_t brick_size, const std::function<float(f3)>& sdf,
                        const std::function<f3(f3)>& color, bool with_neighbors);

}  // namespace lrt

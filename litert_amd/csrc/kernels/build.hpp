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

}  // namespace lrtd

namespace lrt {
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

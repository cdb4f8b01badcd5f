// Host BVH builder: binned-SAH binary BVH collapsed into the compressed
// 8-wide node layout (shb::CWNode). Replaces Embree's BVH build inside
// BuildBVH (reference src/scene.cpp:209-262).
#ifndef SHB_BVH_BUILD_H_
#define SHB_BVH_BUILD_H_

#include <cstdint>
#include <vector>

#include "shb_common.h"

namespace shb {

struct BvhBuildOut {
  std::vector<CWNode> nodes;   // root at 0
  std::vector<float4> tris;    // 3 per triangle in leaf order; v0.w = id
  int max_depth = 0;           // nodes on the longest root-to-leaf path
  int sah_depth = 0;           // binary depth the SAH splits stopped at (INT_MAX: none)
  double seconds = 0.0;
};

// Deepest tree the traversal accepts: its group stack holds one entry per
// level below the root.
constexpr int kMaxBvhDepth = kStackLds + 1;

// Builds over n world-space triangles (9 floats each: v0, v1, v2).
// width: maximum children per node (2..8). When the SAH tree collapses to
// more than kMaxBvhDepth levels (skewed or clustered geometry), the build is
// redone with object-median splits below a binary depth budget (32, 24, ..,
// 0) until it fits; only a width-2 tree of more than about 3 * 2^16
// triangles can still come out deeper.
void build_cwbvh(const float* tri_xyz, uint32_t n, int width, BvhBuildOut* out);

}  // namespace shb

#endif  // SHB_BVH_BUILD_H_

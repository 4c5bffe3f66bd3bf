// bvh_build.h — host-side SAH build of the binary BVH and its collapse into the four-wide tree that the gfx950
// traversal kernels read (layout in pe_gpu_types.h).
#pragma once

#include "pe_gpu_types.h"

#include <cstddef>
#include <cstdint>
#include <vector>

namespace pe {

struct BvhBuildOptions {
    int max_leaf = 4;       // largest leaf the SAH may choose
    int max_depth = 60;     // SAH depth; deeper nodes are halved by count until they hold <= max_leaf_cap
    int max_leaf_cap = 16;  // largest leaf at all (a traversal leaf item holds at most 16 triangles), so the
                            // tree is at most max_depth + log2(n / 16) deep (< 96 entries of traversal stack)
    int bins = 16;        // SAH bins per axis (all three axes are tried)
    float traversal_cost = 1.0f;
    float triangle_cost = 1.0f;
};

struct BvhBuildResult {
    std::vector<BvhNode> nodes;      // interior nodes, depth-first, root at 0
    std::vector<uint32_t> order;     // leaf order -> original triangle index
    int depth = 0;                   // deepest interior level (root = 1)
    size_t leaves = 0;
};

// tri_lo/tri_hi: per-triangle bounds (3 floats each). Never fails for n >= 1.
BvhBuildResult build_bvh(const float* tri_lo, const float* tri_hi, size_t n, const BvhBuildOptions& opt);

// The four-wide tree the traversal reads, collapsed from the binary one: a node takes its two children and
// puts each interior child's two children in its place, so every four-wide level spans two binary levels
// (a leaf child stays as it is). Root at 0, the first kBvhTileNodes nodes in breadth-first order from the root
// (the traversal keeps them in LDS), the rest depth-first; `depth` receives the deepest level (root = 1).
std::vector<BvhNode4> collapse_bvh4(const std::vector<BvhNode>& nodes, int& depth);

}  // namespace pe

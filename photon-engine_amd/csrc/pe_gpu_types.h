// pe_gpu_types.h — memory layout of a scene and of the wavefront queues in HBM.
//
// Acceleration structure: a binary BVH whose node holds the boxes of its two children (64 B, four
// 16-byte loads per visited node) rather than its own box, so one fetch decides both descents. A child
// is either an interior node (count == 0, ref = node index) or a leaf (count > 0, ref = first entry of
// the leaf-ordered triangle array). Triangles are stored in leaf order as the three float4 the lean
// Möller–Trumbore test reads (v0 | eps, e1 | triangle id, e2 | 0), 48 B per triangle, so a leaf is one
// contiguous run. Shading data (vertex indices, material, shape/prim ids, emitter) is a second array in
// the scene's original triangle order, touched once per hit.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pe {

struct alignas(16) BvhNode {
    float4 c0xy;  // child 0: lo.x, hi.x, lo.y, hi.y
    float4 c1xy;  // child 1: lo.x, hi.x, lo.y, hi.y
    float4 cz;    // child 0 lo.z, hi.z, child 1 lo.z, hi.z
    int4 ref;     // child 0 ref, child 1 ref, child 0 count, child 1 count
};
static_assert(sizeof(BvhNode) == 64, "BvhNode must be 64 bytes");

// Four-wide node the traversal reads: the boxes of up to four children, SoA by plane, and the children's
// traversal items (an interior node's index, a leaf item, or 0xFFFFFFFF for an empty slot).
struct alignas(16) BvhNode4 {
    float4 lox, hix, loy, hiy;  // first 64 bytes
    float4 loz, hiz;            // second 64 bytes
    uint4 item;
    uint4 pad;
};
static_assert(sizeof(BvhNode4) == 128, "BvhNode4 must be 128 bytes");

struct alignas(16) TriAccel {
    float4 v0;  // xyz = vertex 0, w = parallel threshold 1e-14 |e1|^2 |e2|^2 (bvh.cpp:49-50)
    float4 e1;  // xyz = v1 - v0, w = global triangle id (int bits)
    float4 e2;  // xyz = v2 - v0, w = 0
};
static_assert(sizeof(TriAccel) == 48, "TriAccel must be 48 bytes");

// Per-triangle shading record in scene order (32 B).
// Four-wide nodes at the top of the tree (root, its children, theirs: 1 + 4 + 16) that the traversal kernels
// hold in LDS; the BVH lays them out first.
constexpr uint32_t kBvhTileNodes = 21;

struct TriShade {
    uint32_t v0, v1, v2;  // global vertex indices
    int32_t material;
    int32_t shape;        // index in Scene::shapes()
    int32_t prim;         // primIndex within a TriangleMesh, -1 for standalone shapes
    int32_t emitter;      // Scene::lights() index of this triangle's light, or -1
    uint32_t flags;       // bit 0: vertex normals present, bit 1: uvs present, bit 2: camera invisible
};
static_assert(sizeof(TriShade) == 32, "TriShade must be 32 bytes");

enum : uint32_t { kTriHasNormals = 1u, kTriHasUv = 2u, kTriCameraInvisible = 4u };

}  // namespace pe

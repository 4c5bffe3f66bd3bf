// oracle_core.cpp — restated Sobol-Owen sampler and binary-BVH scene queries (test infrastructure only).
//
// The BVH is rebuilt exactly as the reference builds it (same bucket count, split axis, leaf rule,
// std::partition / std::nth_element calls, depth-first flattening), so that traversal order, and with it
// the winner of genuine t ties, is the reference's own.
#include "oracle.h"
#include "oracle_scene.h"

#include <algorithm>
#include <cstring>
#include <limits>
#include <memory>

namespace oracle {

// ------------------------------------------------------------------------------ sobol_sampler.cpp:14-25
uint32_t reverseBits(uint32_t x) {
    x = __builtin_bswap32(x);
    x = ((x >> 1) & 0x55555555u) | ((x & 0x55555555u) << 1);
    x = ((x >> 2) & 0x33333333u) | ((x & 0x33333333u) << 2);
    x = ((x >> 4) & 0x0F0F0F0Fu) | ((x & 0x0F0F0F0Fu) << 4);
    return x;
}

// sobol_sampler.cpp:31-37: direction numbers v_k = v_{k-1} ^ (v_{k-1} >> 1), v_0 = 2^31.
uint32_t sobolDim1(uint32_t index) {
    uint32_t result = 0;
    for (uint32_t v = 0x80000000u; index != 0; index >>= 1, v ^= v >> 1)
        if (index & 1u) result ^= v;
    return result;
}

// sobol_sampler.cpp:62-69
uint32_t laineKarras(uint32_t x, uint32_t seed) {
    x += seed;
    x ^= x * 0x6c50b47cu;
    x ^= x * 0xb82f1e52u;
    x ^= x * 0xc7afe638u;
    x ^= x * 0x8d22f6e6u;
    return x;
}

// sobol_sampler.cpp:72-79
uint32_t mix32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

// sobol_sampler.cpp:83-85
uint32_t hashCombine(uint32_t h, uint32_t v) { return mix32(h ^ (v + 0x9e3779b9u + (h << 6) + (h >> 2))); }

// sobol_sampler.cpp:88
float toUnitFloat(uint32_t x) { return static_cast<float>(x >> 8) * 0x1p-24f; }

// sobol_sampler.cpp:133-140
void Sobol::startPixel(int x, int y) {
    uint32_t h = mix32(static_cast<uint32_t>(seed) ^ mix32(static_cast<uint32_t>(seed >> 32)));
    h = hashCombine(h, static_cast<uint32_t>(x));
    pixelHash = hashCombine(h, static_cast<uint32_t>(y));
    index = revIndex = dimension = 0;
}

// sobol_sampler.cpp:142-146
void Sobol::startSample(int i) {
    index = static_cast<uint32_t>(i);
    revIndex = reverseBits(index);
    dimension = 0;
}

// sobol_sampler.cpp:111-117, with reversedSobolDim1(index) = reverseBits(sobolDim1(index)) (:58-61)
void Sobol::get2DBits(uint32_t& x, uint32_t& y) {
    const uint32_t s = hashCombine(pixelHash, dimension++);
    const uint32_t idx = reverseBits(laineKarras(revIndex, s));
    x = reverseBits(laineKarras(idx, hashCombine(s, 0x0u)));
    y = reverseBits(laineKarras(reverseBits(sobolDim1(idx)), hashCombine(s, 0x1u)));
}

// sobol_sampler.cpp:120-125
float Sobol::get1D() {
    const uint32_t s = hashCombine(pixelHash, dimension++);
    const uint32_t idx = reverseBits(laineKarras(revIndex, s));
    return toUnitFloat(reverseBits(laineKarras(idx, hashCombine(s, 0x0u))));
}

void Sobol::get2D(float& u, float& v) {
    uint32_t x, y;
    get2DBits(x, y);
    u = toUnitFloat(x);
    v = toUnitFloat(y);
}

// ------------------------------------------------------------------------------------- BVH (bvh.cpp)
namespace {

struct Aabb {  // core/math/aabb.h:24-31 (empty = [+max, -max])
    V3 lo{std::numeric_limits<float>::max()}, hi{-std::numeric_limits<float>::max()};
    void merge(const Aabb& b) {
        lo = vmin(lo, b.lo);
        hi = vmax(hi, b.hi);
    }
    void merge(V3 p) {
        lo = vmin(lo, p);
        hi = vmax(hi, p);
    }
    V3 centroid() const { return (lo + hi) * 0.5f; }
    V3 diag() const { return hi - lo; }
    float surfaceArea() const {  // aabb.h:117-120
        const V3 d = diag();
        return 2.0f * (d.x * d.y + d.x * d.z + d.y * d.z);
    }
    int maxExtent() const {  // aabb.h:123-128
        const V3 d = diag();
        if (d.x > d.y && d.x > d.z) return 0;
        if (d.y > d.z) return 1;
        return 2;
    }
    V3 offset(V3 p) const {  // aabb.h:141-147
        V3 o = p - lo;
        const V3 d = diag();
        if (d.x > 0.0f) o.x /= d.x;
        if (d.y > 0.0f) o.y /= d.y;
        if (d.z > 0.0f) o.z /= d.z;
        return o;
    }
};

// vec.h cwiseMin/cwiseMax use std::min/std::max (first argument wins on NaN/equality).
V3 cmin(V3 a, V3 b) { return V3(std::min(a.x, b.x), std::min(a.y, b.y), std::min(a.z, b.z)); }
V3 cmax(V3 a, V3 b) { return V3(std::max(a.x, b.x), std::max(a.y, b.y), std::max(a.z, b.z)); }

struct BuildPrim {
    size_t index;
    Aabb bounds;
    V3 centroid;
};

struct BuildNode {
    Aabb bounds;
    BuildNode* child[2] = {nullptr, nullptr};
    int axis = 0, first = 0, count = 0;
};

struct Builder {
    std::vector<BuildPrim>& prims;
    std::vector<int>& ordered;
    int total = 0;

    BuildNode* leaf(BuildNode* n, int s, int e, const Aabb& b) {
        n->first = static_cast<int>(ordered.size());
        for (int i = s; i < e; ++i) ordered.push_back(static_cast<int>(prims[static_cast<size_t>(i)].index));
        n->count = e - s;
        n->bounds = b;
        return n;
    }
    BuildNode* interior(BuildNode* n, int axis, BuildNode* a, BuildNode* b) {
        n->child[0] = a;
        n->child[1] = b;
        n->bounds = a->bounds;
        n->bounds.lo = cmin(a->bounds.lo, b->bounds.lo);
        n->bounds.hi = cmax(a->bounds.hi, b->bounds.hi);
        n->axis = axis;
        n->count = 0;
        return n;
    }

    // bvh.cpp:100-238
    BuildNode* build(int start, int end) {
        ++total;
        auto* node = new BuildNode();
        Aabb bounds;
        for (int i = start; i < end; ++i) {
            bounds.lo = cmin(bounds.lo, prims[static_cast<size_t>(i)].bounds.lo);
            bounds.hi = cmax(bounds.hi, prims[static_cast<size_t>(i)].bounds.hi);
        }
        const int n = end - start;
        if (n == 1) return leaf(node, start, end, bounds);
        Aabb cb;
        for (int i = start; i < end; ++i) {
            cb.lo = cmin(cb.lo, prims[static_cast<size_t>(i)].centroid);
            cb.hi = cmax(cb.hi, prims[static_cast<size_t>(i)].centroid);
        }
        const int dim = cb.maxExtent();
        if (cb.hi[dim] == cb.lo[dim]) {
            if (n > 4) {
                const int mid = (start + end) / 2;
                BuildNode* a = build(start, mid);
                BuildNode* b = build(mid, end);
                return interior(node, dim, a, b);
            }
            return leaf(node, start, end, bounds);
        }
        int mid = (start + end) / 2;
        if (n <= 2) {
            std::nth_element(prims.begin() + start, prims.begin() + mid, prims.begin() + end,
                             [dim](const BuildPrim& a, const BuildPrim& b) { return a.centroid[dim] < b.centroid[dim]; });
        } else {
            constexpr int nb = 12;
            struct Bucket {
                int count = 0;
                Aabb b;
            } buckets[nb];
            for (int i = start; i < end; ++i) {
                int k = static_cast<int>(nb * cb.offset(prims[static_cast<size_t>(i)].centroid)[dim]);
                if (k == nb) k = nb - 1;
                buckets[k].count++;
                buckets[k].b.lo = cmin(buckets[k].b.lo, prims[static_cast<size_t>(i)].bounds.lo);
                buckets[k].b.hi = cmax(buckets[k].b.hi, prims[static_cast<size_t>(i)].bounds.hi);
            }
            float cost[nb - 1];
            for (int i = 0; i < nb - 1; ++i) {
                Aabb b0, b1;
                int c0 = 0, c1 = 0;
                for (int j = 0; j <= i; ++j) {
                    b0.lo = cmin(b0.lo, buckets[j].b.lo);
                    b0.hi = cmax(b0.hi, buckets[j].b.hi);
                    c0 += buckets[j].count;
                }
                for (int j = i + 1; j < nb; ++j) {
                    b1.lo = cmin(b1.lo, buckets[j].b.lo);
                    b1.hi = cmax(b1.hi, buckets[j].b.hi);
                    c1 += buckets[j].count;
                }
                cost[i] = 1.0f + (c0 * b0.surfaceArea() + c1 * b1.surfaceArea()) / bounds.surfaceArea();
            }
            float minCost = cost[0];
            int minBucket = 0;
            for (int i = 1; i < nb - 1; ++i)
                if (cost[i] < minCost) {
                    minCost = cost[i];
                    minBucket = i;
                }
            const float leafCost = static_cast<float>(n);
            if (n > 255 || minCost < leafCost) {
                auto pmid = std::partition(prims.begin() + start, prims.begin() + end, [=](const BuildPrim& bp) {
                    int k = static_cast<int>(nb * cb.offset(bp.centroid)[dim]);
                    if (k == nb) k = nb - 1;
                    return k <= minBucket;
                });
                mid = static_cast<int>(pmid - prims.begin());
                if (mid != start && mid != end) {
                    BuildNode* a = build(start, mid);
                    BuildNode* b = build(mid, end);
                    return interior(node, dim, a, b);
                }
            }
            return leaf(node, start, end, bounds);
        }
        BuildNode* a = build(start, mid);
        BuildNode* b = build(mid, end);
        return interior(node, dim, a, b);
    }
};

void freeTree(BuildNode* n) {
    if (!n) return;
    freeTree(n->child[0]);
    freeTree(n->child[1]);
    delete n;
}

// bvh.cpp:240-255
int flatten(std::vector<LinearNode>& out, BuildNode* n, int* offset) {
    const int my = (*offset)++;
    out[static_cast<size_t>(my)].lo = n->bounds.lo;
    out[static_cast<size_t>(my)].hi = n->bounds.hi;
    if (n->count > 0) {
        out[static_cast<size_t>(my)].offset = n->first;
        out[static_cast<size_t>(my)].count = n->count;
    } else {
        out[static_cast<size_t>(my)].axis = n->axis;
        out[static_cast<size_t>(my)].count = 0;
        flatten(out, n->child[0], offset);
        out[static_cast<size_t>(my)].offset = flatten(out, n->child[1], offset);
    }
    return my;
}

// bvh.cpp:269-283 (std::max keeps its first argument when the second is NaN).
bool hitBox(V3 lo, V3 hi, V3 o, V3 inv, float tMin, float tMax, float& tNear) {
    float tx0 = (lo.x - o.x) * inv.x, tx1 = (hi.x - o.x) * inv.x;
    if (tx0 > tx1) std::swap(tx0, tx1);
    float ty0 = (lo.y - o.y) * inv.y, ty1 = (hi.y - o.y) * inv.y;
    if (ty0 > ty1) std::swap(ty0, ty1);
    float tz0 = (lo.z - o.z) * inv.z, tz1 = (hi.z - o.z) * inv.z;
    if (tz0 > tz1) std::swap(tz0, tz1);
    const float t0 = std::max(std::max(tMin, tx0), std::max(ty0, tz0));
    const float t1 = std::min(std::min(tMax, tx1), std::min(ty1, tz1)) * 1.0000004f;
    tNear = t0;
    return t0 <= t1;
}

}  // namespace

// bvh.cpp:302-316 (same operations as Triangle::intersect, triangle.cpp:11-50)
bool hitTriangleLean(const TriAccel& T, const Ray& r, float& t, float& u, float& v) {
    const V3 h = cross(r.d, T.e2);
    const float a = dot(T.e1, h);
    if (a * a <= T.parallelEps) return false;
    const float f = 1.0f / a;
    const V3 s = r.o - T.v0;
    u = f * dot(s, h);
    if (!(u >= 0.0f && u <= 1.0f)) return false;
    const V3 q = cross(s, T.e1);
    v = f * dot(r.d, q);
    if (!(v >= 0.0f && u + v <= 1.0f)) return false;
    t = f * dot(T.e2, q);
    return t >= r.tMin && t <= r.tMax;
}

void buildBvh(Scene& s) {
    const size_t n = s.tris.size();
    s.ordered.clear();
    if (n == 0) {  // an empty scene: no nodes, every ray misses (bvh.cpp build of no primitives)
        s.nodes.clear();
        s.accel.clear();
        return;
    }
    std::vector<BuildPrim> prims;
    prims.reserve(n);
    for (size_t i = 0; i < n; ++i) {
        const V3 p0 = s.pos[s.tris[i].v[0]], p1 = s.pos[s.tris[i].v[1]], p2 = s.pos[s.tris[i].v[2]];
        BuildPrim bp;
        bp.index = i;
        bp.bounds.lo = cmin(cmin(p0, p1), p2);  // mesh.cpp:70-77 / triangle.cpp:131-136
        bp.bounds.hi = cmax(cmax(p0, p1), p2);
        bp.centroid = bp.bounds.centroid();
        prims.push_back(bp);
    }
    s.ordered.clear();
    Builder b{prims, s.ordered};
    BuildNode* root = b.build(0, static_cast<int>(n));
    s.nodes.assign(static_cast<size_t>(b.total), LinearNode{});
    int off = 0;
    flatten(s.nodes, root, &off);
    freeTree(root);
    s.accel.resize(n);
    for (size_t i = 0; i < n; ++i) {
        const TriRec& tr = s.tris[static_cast<size_t>(s.ordered[i])];
        const V3 p0 = s.pos[tr.v[0]], p1 = s.pos[tr.v[1]], p2 = s.pos[tr.v[2]];
        const V3 e1 = p1 - p0, e2 = p2 - p0;  // bvh.cpp:45-50
        s.accel[i] = TriAccel{p0, e1, e2, 1e-14f * lengthSq(e1) * lengthSq(e2)};
    }
}

// bvh.cpp:322-401 (closest hit, near child first; far child pushed with its entry distance)
bool intersect(const Scene& s, const Ray& rayIn, Hit& hit) {
    if (s.nodes.empty()) return false;
    Ray ray = rayIn;
    const V3 inv(1.0f / ray.d.x, 1.0f / ray.d.y, 1.0f / ray.d.z);
    struct Entry {
        int node;
        float tNear;
    };
    std::vector<Entry> stack;
    int best = -1;
    float bu = 0, bv = 0;
    int node = 0;
    float tNear = 0.0f;
    if (!hitBox(s.nodes[0].lo, s.nodes[0].hi, ray.o, inv, ray.tMin, ray.tMax, tNear)) return false;
    while (true) {
        const LinearNode& n = s.nodes[static_cast<size_t>(node)];
        if (n.count > 0) {
            for (int k = 0; k < n.count; ++k) {
                const int pi = n.offset + k;
                float t, u, v;
                if (hitTriangleLean(s.accel[static_cast<size_t>(pi)], ray, t, u, v)) {
                    ray.tMax = t;
                    best = pi;
                    bu = u;
                    bv = v;
                }
            }
        } else {
            const int c0 = node + 1, c1 = n.offset;
            float t0, t1;
            const bool h0 = hitBox(s.nodes[static_cast<size_t>(c0)].lo, s.nodes[static_cast<size_t>(c0)].hi, ray.o, inv,
                                   ray.tMin, ray.tMax, t0);
            const bool h1 = hitBox(s.nodes[static_cast<size_t>(c1)].lo, s.nodes[static_cast<size_t>(c1)].hi, ray.o, inv,
                                   ray.tMin, ray.tMax, t1);
            if (h0 && h1) {
                if (t1 < t0) {
                    stack.push_back({c0, t0});
                    node = c1;
                } else {
                    stack.push_back({c1, t1});
                    node = c0;
                }
                continue;
            }
            if (h0) {
                node = c0;
                continue;
            }
            if (h1) {
                node = c1;
                continue;
            }
        }
        bool found = false;
        while (!stack.empty()) {
            const Entry e = stack.back();
            stack.pop_back();
            if (e.tNear <= ray.tMax) {
                node = e.node;
                found = true;
                break;
            }
        }
        if (!found) break;
    }
    if (best < 0) return false;
    const int tri = s.ordered[static_cast<size_t>(best)];
    hit.t = ray.tMax;
    hit.tri = tri;
    hit.shape = s.tris[static_cast<size_t>(tri)].shape;
    hit.prim = s.tris[static_cast<size_t>(tri)].prim;
    hit.u = bu;
    hit.v = bv;
    return true;
}

// bvh.cpp:404-446 (any hit, children ordered by the sign of the ray along the split axis)
bool intersectAny(const Scene& s, const Ray& ray) {
    if (s.nodes.empty()) return false;
    const V3 inv(1.0f / ray.d.x, 1.0f / ray.d.y, 1.0f / ray.d.z);
    const bool neg[3] = {inv.x < 0.0f, inv.y < 0.0f, inv.z < 0.0f};
    std::vector<int> stack;
    int node = 0;
    float tNear;
    while (true) {
        const LinearNode& n = s.nodes[static_cast<size_t>(node)];
        if (hitBox(n.lo, n.hi, ray.o, inv, ray.tMin, ray.tMax, tNear)) {
            if (n.count > 0) {
                for (int k = 0; k < n.count; ++k) {
                    float t, u, v;
                    if (hitTriangleLean(s.accel[static_cast<size_t>(n.offset + k)], ray, t, u, v)) return true;
                }
            } else {
                if (neg[n.axis]) {
                    stack.push_back(node + 1);
                    node = n.offset;
                } else {
                    stack.push_back(n.offset);
                    node = node + 1;
                }
                continue;
            }
        }
        if (stack.empty()) break;
        node = stack.back();
        stack.pop_back();
    }
    return false;
}

int sceneTriangles(const Scene& s) { return static_cast<int>(s.tris.size()); }

}  // namespace oracle

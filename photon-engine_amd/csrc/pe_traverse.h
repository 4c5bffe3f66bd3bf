// pe_traverse.h — closest-hit and any-hit BVH traversal for one ray per lane on gfx950.
//
// Per-triangle arithmetic is the reference's lean Möller–Trumbore test (geometry/bvh.cpp:302-316, same
// operations as Triangle::intersect, triangle.cpp:11-50) with contraction disabled, so a triangle accepts
// or rejects a ray, and reports t, bit for bit as the reference's IEEE build does. Accepting t <= tMax with
// a shrinking tMax (bvh.cpp:315, :343) means that of two triangles at exactly the same t the one tested
// last wins; that order depends on the tree, and such rays are the documented "genuine ties".
//
// Box test: slab test with the reference's exit pad t1 * 1.0000004 (bvh.cpp:269-283). NaN slabs (0 * inf,
// a ray in a slab plane) are dropped by fminf/fmaxf instead of the reference's argument-order rule.
//
// Traversal stack: kLdsDepth entries per lane live in LDS, column-major (entry e of lane l at e * blockDim + l,
// conflict-free), deeper entries spill to a per-lane column of a global buffer. The BVH build caps the
// interior depth, which bounds the stack at kStackDepth.
#pragma once

#include "pe_gpu_types.h"
#include "pe_math.h"

namespace pe {

constexpr int kLdsDepth = 16;
constexpr int kStackDepth = 64;
constexpr int kTraceBlock = 256;

struct TraceStack {
    uint32_t* lds;      // kLdsDepth * blockDim entries, this lane's column at lds[threadIdx.x]
    uint32_t* spill;    // this lane's column in the spill buffer
    uint32_t spill_stride;
    int sp;

    // The BVH build keeps interior depth below kStackDepth, so a push never finds the stack full; were it
    // to, the entry is dropped rather than written out of bounds.
    __device__ void push(uint32_t v) {
        if (sp < kLdsDepth) lds[sp * kTraceBlock] = v;
        else if (sp < kStackDepth) spill[static_cast<size_t>(sp - kLdsDepth) * spill_stride] = v;
        else return;
        ++sp;
    }
    __device__ uint32_t pop() {
        --sp;
        if (sp < kLdsDepth) return lds[sp * kTraceBlock];
        return spill[static_cast<size_t>(sp - kLdsDepth) * spill_stride];
    }
};

struct RayHit {
    float t;
    int32_t tri;
    float u, v;
};

__device__ __forceinline__ void box_test2(const BvhNode& n, V3 o, V3 inv, float tmin, float tmax, bool& h0, bool& h1,
                                          float& t0, float& t1) {
    const float ax0 = (n.c0xy.x - o.x) * inv.x, ax1 = (n.c0xy.y - o.x) * inv.x;
    const float ay0 = (n.c0xy.z - o.y) * inv.y, ay1 = (n.c0xy.w - o.y) * inv.y;
    const float az0 = (n.cz.x - o.z) * inv.z, az1 = (n.cz.y - o.z) * inv.z;
    const float bx0 = (n.c1xy.x - o.x) * inv.x, bx1 = (n.c1xy.y - o.x) * inv.x;
    const float by0 = (n.c1xy.z - o.y) * inv.y, by1 = (n.c1xy.w - o.y) * inv.y;
    const float bz0 = (n.cz.z - o.z) * inv.z, bz1 = (n.cz.w - o.z) * inv.z;
    const float an = fmaxf(fmaxf(tmin, fminf(ax0, ax1)), fmaxf(fminf(ay0, ay1), fminf(az0, az1)));
    const float af = fminf(fminf(tmax, fmaxf(ax0, ax1)), fminf(fmaxf(ay0, ay1), fmaxf(az0, az1))) * 1.0000004f;
    const float bn = fmaxf(fmaxf(tmin, fminf(bx0, bx1)), fmaxf(fminf(by0, by1), fminf(bz0, bz1)));
    const float bf = fminf(fminf(tmax, fmaxf(bx0, bx1)), fminf(fmaxf(by0, by1), fmaxf(bz0, bz1))) * 1.0000004f;
    h0 = an <= af;
    h1 = bn <= bf;
    t0 = an;
    t1 = bn;
}

// Lean Möller–Trumbore (bvh.cpp:302-316). Returns true and t/u/v when tmin <= t <= tmax. Every dot and cross
// product is evaluated uncontracted (dot_exact / cross_exact): hipcc would otherwise fuse them into FMAs.
__device__ __forceinline__ bool tri_test(const TriAccel& T, V3 o, V3 d, float tmin, float tmax, float& t, float& u,
                                         float& v) {
#pragma clang fp contract(off)
    const V3 e1 = mk(T.e1.x, T.e1.y, T.e1.z);
    const V3 e2 = mk(T.e2.x, T.e2.y, T.e2.z);
    const V3 h = cross_exact(d, e2);
    const float a = dot_exact(e1, h);
    if (a * a <= T.v0.w) return false;
    const float f = 1.0f / a;
    const V3 s = o - mk(T.v0.x, T.v0.y, T.v0.z);
    u = f * dot_exact(s, h);
    if (!(u >= 0.0f && u <= 1.0f)) return false;
    const V3 q = cross_exact(s, e1);
    v = f * dot_exact(d, q);
    if (!(v >= 0.0f && u + v <= 1.0f)) return false;
    t = f * dot_exact(e2, q);
    return t >= tmin && t <= tmax;
}

__device__ __forceinline__ BvhNode load_node(const BvhNode* __restrict__ nodes, uint32_t i) {
    const float4* p = reinterpret_cast<const float4*>(nodes + i);
    BvhNode n;
    n.c0xy = p[0];
    n.c1xy = p[1];
    n.cz = p[2];
    n.ref = reinterpret_cast<const int4*>(p)[3];
    return n;
}

__device__ __forceinline__ TriAccel load_tri(const TriAccel* __restrict__ tris, uint32_t i) {
    const float4* p = reinterpret_cast<const float4*>(tris + i);
    TriAccel t;
    t.v0 = p[0];
    t.e1 = p[1];
    t.e2 = p[2];
    return t;
}

// Closest hit. hit.tri = -1 on a miss. Visit counters (nodes, triangles) are accumulated when non-null.
__device__ bool trace_closest(const BvhNode* __restrict__ nodes, const TriAccel* __restrict__ tris, V3 o, V3 d,
                              float tmin, float tmax, TraceStack& st, RayHit& hit, uint32_t* node_visits,
                              uint32_t* tri_visits) {
    const V3 inv = mk(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
    hit.tri = -1;
    hit.t = tmax;
    hit.u = hit.v = 0.0f;
    st.sp = 0;
    uint32_t node = 0;
    uint32_t nv = 0, tv = 0;
    while (true) {
        const BvhNode n = load_node(nodes, node);
        ++nv;
        bool h0, h1;
        float t0, t1;
        box_test2(n, o, inv, tmin, hit.t, h0, h1, t0, t1);
        if (h0 && n.ref.z > 0) {
            for (int k = 0; k < n.ref.z; ++k) {
                const TriAccel T = load_tri(tris, static_cast<uint32_t>(n.ref.x + k));
                ++tv;
                float t, u, v;
                if (tri_test(T, o, d, tmin, hit.t, t, u, v)) {
                    hit.t = t;
                    hit.u = u;
                    hit.v = v;
                    hit.tri = __float_as_int(T.e1.w);
                }
            }
            h0 = false;
        }
        if (h1 && n.ref.w > 0) {
            for (int k = 0; k < n.ref.w; ++k) {
                const TriAccel T = load_tri(tris, static_cast<uint32_t>(n.ref.y + k));
                ++tv;
                float t, u, v;
                if (tri_test(T, o, d, tmin, hit.t, t, u, v)) {
                    hit.t = t;
                    hit.u = u;
                    hit.v = v;
                    hit.tri = __float_as_int(T.e1.w);
                }
            }
            h1 = false;
        }
        h0 = h0 && t0 <= hit.t;
        h1 = h1 && t1 <= hit.t;
        if (h0 && h1) {
            const bool swap = t1 < t0;
            st.push(static_cast<uint32_t>(swap ? n.ref.x : n.ref.y));
            node = static_cast<uint32_t>(swap ? n.ref.y : n.ref.x);
            continue;
        }
        if (h0) {
            node = static_cast<uint32_t>(n.ref.x);
            continue;
        }
        if (h1) {
            node = static_cast<uint32_t>(n.ref.y);
            continue;
        }
        if (st.sp == 0) break;
        node = st.pop();
    }
    if (node_visits) *node_visits += nv;
    if (tri_visits) *tri_visits += tv;
    return hit.tri >= 0;
}

// Any hit within [tmin, tmax] (Scene::intersectAny, bvh.cpp:404-446): returns at the first accepted triangle.
__device__ bool trace_any(const BvhNode* __restrict__ nodes, const TriAccel* __restrict__ tris, V3 o, V3 d, float tmin,
                          float tmax, TraceStack& st, uint32_t* node_visits, uint32_t* tri_visits) {
    const V3 inv = mk(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
    st.sp = 0;
    uint32_t node = 0;
    uint32_t nv = 0, tv = 0;
    bool found = false;
    while (true) {
        const BvhNode n = load_node(nodes, node);
        ++nv;
        bool h0, h1;
        float t0, t1;
        box_test2(n, o, inv, tmin, tmax, h0, h1, t0, t1);
        if (h0 && n.ref.z > 0) {
            for (int k = 0; k < n.ref.z && !found; ++k) {
                const TriAccel T = load_tri(tris, static_cast<uint32_t>(n.ref.x + k));
                ++tv;
                float t, u, v;
                found = tri_test(T, o, d, tmin, tmax, t, u, v);
            }
            h0 = false;
        }
        if (!found && h1 && n.ref.w > 0) {
            for (int k = 0; k < n.ref.w && !found; ++k) {
                const TriAccel T = load_tri(tris, static_cast<uint32_t>(n.ref.y + k));
                ++tv;
                float t, u, v;
                found = tri_test(T, o, d, tmin, tmax, t, u, v);
            }
            h1 = false;
        }
        if (found) break;
        if (h0 && h1) {
            const bool swap = t1 < t0;
            st.push(static_cast<uint32_t>(swap ? n.ref.x : n.ref.y));
            node = static_cast<uint32_t>(swap ? n.ref.y : n.ref.x);
            continue;
        }
        if (h0) {
            node = static_cast<uint32_t>(n.ref.x);
            continue;
        }
        if (h1) {
            node = static_cast<uint32_t>(n.ref.y);
            continue;
        }
        if (st.sp == 0) break;
        node = st.pop();
    }
    if (node_visits) *node_visits += nv;
    if (tri_visits) *tri_visits += tv;
    return found;
}

}  // namespace pe

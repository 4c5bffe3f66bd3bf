// pe_traverse.h — closest-hit and any-hit BVH traversal of a ray queue on gfx950.
//
// Per-triangle arithmetic is the reference's lean Möller–Trumbore test (geometry/bvh.cpp:302-316, same
// operations as Triangle::intersect, triangle.cpp:11-50) with contraction disabled, so a triangle accepts
// or rejects a ray, and reports t, bit for bit as the reference's IEEE build does. Accepting t <= tMax with
// a shrinking tMax (bvh.cpp:315, :343) means that of two triangles at exactly the same t the one tested
// last wins; that order depends on the tree and on the traversal order, and such rays are the documented
// "genuine ties". Any-hit queries return the same boolean whatever the order.
//
// Box test: slab test with the reference's exit pad t1 * 1.0000004 (bvh.cpp:269-283). NaN slabs (0 * inf,
// a ray in a slab plane) are dropped by fminf/fmaxf instead of the reference's argument-order rule.
//
// Scheduling. A traversal is a sequence of items, each one 64-byte gather: an interior node (the boxes of
// both children) or one triangle of a leaf (48 B, read as 64 B: the triangle array carries 16 B of tail
// padding). Every lane of a wave handles one item per loop step, so a step costs one memory latency
// whatever mix of nodes and triangles the lanes hold — traversal here is bound by that latency, not by HBM
// bandwidth. A lane whose ray has ended takes the next ray of its wave's share at the top of the next step
// (a ballot ranks the idle lanes), and the new ray's record is loaded together with the first node, so lanes
// stay busy until the wave's share is used up. The share of wave w is every W-th 64-ray chunk of the queue
// from chunk w (W = waves in the grid): cheap and expensive regions of the queue are spread over all waves
// without an atomic counter.
//
// Item encoding (stack entries are items too): an interior node is its index; a leaf run is
// kLeafItem | (remaining triangles - 1) << 27 | first triangle, so a leaf holds at most 16 triangles and
// a scene at most 2^27 triangles (both enforced at upload).
//
// Traversal stack: kLdsDepth entries per lane live in LDS, column-major (entry e of lane l at
// e * kTraceBlock + l, conflict-free), deeper entries spill to a per-lane column of a global buffer. One entry
// per tree level at most, so the tree depth bound of the build bounds the stack at kStackDepth.
#pragma once

#include "pe_gpu_types.h"
#include "pe_math.h"

namespace pe {

constexpr int kLdsDepth = 16;
constexpr int kStackDepth = 96;
constexpr int kTraceBlock = 256;
constexpr uint32_t kLeafItem = 0x80000000u;
constexpr uint32_t kLeafFirstMask = (1u << 27) - 1u;
constexpr uint32_t kMaxLeafItemTris = 16;
constexpr uint32_t kNoItem = 0xFFFFFFFFu;
constexpr uint32_t kNoRay = 0xFFFFFFFFu;
constexpr uint64_t kMaxTraceTriangles = uint64_t{1} << 27;

struct TraceStack {
    uint32_t* lds;      // this lane's column: entry e at lds[e * kTraceBlock]
    uint32_t* spill;    // this lane's column in the spill buffer: entry e at spill[(e - kLdsDepth) * spill_stride]
    uint32_t spill_stride;
    int sp;

    // The build bounds the tree depth below kStackDepth, so a push never finds the stack full; were it to,
    // the entry is dropped rather than written out of bounds.
    __device__ __forceinline__ void push(uint32_t v) {
        if (sp < kLdsDepth) lds[sp * kTraceBlock] = v;
        else if (sp < kStackDepth) spill[static_cast<size_t>(sp - kLdsDepth) * spill_stride] = v;
        else return;
        ++sp;
    }
    __device__ __forceinline__ uint32_t pop() {
        --sp;
        if (sp < kLdsDepth) return lds[sp * kTraceBlock];
        return spill[static_cast<size_t>(sp - kLdsDepth) * spill_stride];
    }
};

// One ray of a queue, decoded by the queue's job.
struct TraceRay {
    V3 o, d;
    float tmin, tmax;
    uint32_t tag;  // job-defined (e.g. where an occluded shadow ray's contribution lives)
};

__device__ __forceinline__ void box_test2(const float4& c0xy, const float4& c1xy, const float4& cz, V3 o, V3 inv,
                                          float tmin, float tmax, bool& h0, bool& h1, float& t0, float& t1) {
    const float ax0 = (c0xy.x - o.x) * inv.x, ax1 = (c0xy.y - o.x) * inv.x;
    const float ay0 = (c0xy.z - o.y) * inv.y, ay1 = (c0xy.w - o.y) * inv.y;
    const float az0 = (cz.x - o.z) * inv.z, az1 = (cz.y - o.z) * inv.z;
    const float bx0 = (c1xy.x - o.x) * inv.x, bx1 = (c1xy.y - o.x) * inv.x;
    const float by0 = (c1xy.z - o.y) * inv.y, by1 = (c1xy.w - o.y) * inv.y;
    const float bz0 = (cz.z - o.z) * inv.z, bz1 = (cz.w - o.z) * inv.z;
    const float an = fmaxf(fmaxf(tmin, fminf(ax0, ax1)), fmaxf(fminf(ay0, ay1), fminf(az0, az1)));
    const float af = fminf(fminf(tmax, fmaxf(ax0, ax1)), fminf(fmaxf(ay0, ay1), fmaxf(az0, az1))) * 1.0000004f;
    const float bn = fmaxf(fmaxf(tmin, fminf(bx0, bx1)), fmaxf(fminf(by0, by1), fminf(bz0, bz1)));
    const float bf = fminf(fminf(tmax, fmaxf(bx0, bx1)), fminf(fmaxf(by0, by1), fmaxf(bz0, bz1))) * 1.0000004f;
    h0 = an <= af;
    h1 = bn <= bf;
    t0 = an;
    t1 = bn;
}

// Lean Möller–Trumbore (bvh.cpp:302-316) on (v0 | eps, e1, e2). Returns true and t/u/v when
// tmin <= t <= tmax. Every dot and cross product is evaluated uncontracted (dot_exact / cross_exact): hipcc
// would otherwise fuse them into FMAs.
__device__ __forceinline__ bool tri_test(const float4& v0, const float4& e1v, const float4& e2v, V3 o, V3 d,
                                         float tmin, float tmax, float& t, float& u, float& v) {
#pragma clang fp contract(off)
    const V3 e1 = mk(e1v.x, e1v.y, e1v.z);
    const V3 e2 = mk(e2v.x, e2v.y, e2v.z);
    const V3 h = cross_exact(d, e2);
    const float a = dot_exact(e1, h);
    if (a * a <= v0.w) return false;
    const float f = 1.0f / a;
    const V3 s = o - mk(v0.x, v0.y, v0.z);
    u = f * dot_exact(s, h);
    if (!(u >= 0.0f && u <= 1.0f)) return false;
    const V3 q = cross_exact(s, e1);
    v = f * dot_exact(d, q);
    if (!(v >= 0.0f && u + v <= 1.0f)) return false;
    t = f * dot_exact(e2, q);
    return t >= tmin && t <= tmax;
}

// The item a child reference of a node stands for.
__device__ __forceinline__ uint32_t child_item(int32_t ref, int32_t count) {
    if (count <= 0) return static_cast<uint32_t>(ref);
    return kLeafItem | (static_cast<uint32_t>(count - 1) << 27) | static_cast<uint32_t>(ref);
}

// Traverses every ray of a queue of n rays with the calling grid (all lanes of all waves must call it).
//   job.fetch(i, a, b)            issue the loads of ray i's record (two float4)
//   job.decode(a, b) -> TraceRay  decode a fetched record
//   closest: job.closest(i, ray, t, tri, u, v, a, b) -> bool   t = tmax and tri = -1 on a miss; true when the
//            job continues the query with the new ray record it wrote into (a, b)
//   any:     job.any(i, ray, occluded)
// nv / tv accumulate the node and triangle items this lane handled (kCount only).
template <bool kAny, bool kCount, class Job>
__device__ __forceinline__ void trace_queue(const Job& job, uint32_t n, const BvhNode* __restrict__ nodes,
                                            const TriAccel* __restrict__ tris, TraceStack& st, uint32_t& nv,
                                            uint32_t& tv) {
    const uint32_t wpb = blockDim.x >> 6;
    const uint32_t W = gridDim.x * wpb;
    const uint32_t w = __builtin_amdgcn_readfirstlane(blockIdx.x * wpb + (threadIdx.x >> 6));
    const uint32_t nchunks = (n + 63u) >> 6;
    const uint64_t below = (uint64_t{1} << __lane_id()) - 1u;
    uint32_t taken = 0;  // rays of this wave's share handed out so far (wave-uniform)
    uint32_t slot = kNoRay;
    uint32_t item = 0;
    bool fresh = false;
    float4 ra = make_float4(0.0f, 0.0f, 0.0f, 0.0f), rb = ra;
    TraceRay ray{};
    V3 inv = mk(0.0f, 0.0f, 0.0f);
    float ht = 0.0f, hu = 0.0f, hv = 0.0f;
    int32_t htri = -1;
    for (;;) {
        // Idle lanes take the next rays of the wave's share, in lane order.
        const uint64_t idle = __ballot(slot == kNoRay);
        if (idle != 0 && w + (taken >> 6) * W < nchunks) {
            if (slot == kNoRay) {
                const uint32_t j = taken + static_cast<uint32_t>(__popcll(idle & below));
                const uint32_t c = w + (j >> 6) * W;
                const uint32_t idx = (c << 6) | (j & 63u);
                if (c < nchunks && idx < n) {
                    slot = idx;
                    job.fetch(idx, ra, rb);
                    fresh = true;
                    item = 0;
                    st.sp = 0;
                }
            }
            taken += static_cast<uint32_t>(__popcll(idle));
        }
        if (__ballot(slot != kNoRay) == 0) break;
        if (slot == kNoRay) continue;

        // One item: a 64-byte gather from the node or the triangle array.
        const bool leaf = (item & kLeafItem) != 0u;
        const float4* p = leaf ? reinterpret_cast<const float4*>(tris + (item & kLeafFirstMask))
                               : reinterpret_cast<const float4*>(nodes + item);
        const float4 x0 = p[0], x1 = p[1], x2 = p[2], x3 = p[3];
        if (fresh) {
            ray = job.decode(ra, rb);
            inv = mk(1.0f / ray.d.x, 1.0f / ray.d.y, 1.0f / ray.d.z);
            ht = ray.tmax;
            hu = hv = 0.0f;
            htri = -1;
            fresh = false;
        }
        uint32_t next = kNoItem;
        bool occluded = false;
        if (leaf) {
            if (kCount) ++tv;
            float t, u, v;
            if (tri_test(x0, x1, x2, ray.o, ray.d, ray.tmin, kAny ? ray.tmax : ht, t, u, v)) {
                if (kAny) {
                    occluded = true;
                } else {
                    ht = t;
                    hu = u;
                    hv = v;
                    htri = __float_as_int(x1.w);
                }
            }
            const uint32_t left = (item >> 27) & 15u;
            if (left > 0u) next = kLeafItem | ((left - 1u) << 27) | ((item & kLeafFirstMask) + 1u);
        } else {
            if (kCount) ++nv;
            bool h0, h1;
            float t0, t1;
            box_test2(x0, x1, x2, ray.o, inv, ray.tmin, kAny ? ray.tmax : ht, h0, h1, t0, t1);
            const uint32_t c0 = child_item(__float_as_int(x3.x), __float_as_int(x3.z));
            const uint32_t c1 = child_item(__float_as_int(x3.y), __float_as_int(x3.w));
            if (h0 && h1) {
                const bool swap = t1 < t0;
                st.push(swap ? c0 : c1);
                next = swap ? c1 : c0;
            } else if (h0) {
                next = c0;
            } else if (h1) {
                next = c1;
            }
        }
        if (next == kNoItem && !occluded && st.sp > 0) next = st.pop();
        if (occluded || next == kNoItem) {
            if constexpr (kAny) {
                job.any(slot, ray, occluded);
                slot = kNoRay;
            } else if (job.closest(slot, ray, ht, htri, hu, hv, ra, rb)) {
                // The job continues the query with a new ray in (ra, rb): restart at the root.
                fresh = true;
                item = 0;
                st.sp = 0;
            } else {
                slot = kNoRay;
            }
        } else {
            item = next;
        }
    }
}

}  // namespace pe

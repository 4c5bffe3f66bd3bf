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
// Scheduling. A traversal is a sequence of items: a four-wide interior node (the boxes of its up to four
// children in SoA planes and their items, 128 bytes; the binary SAH tree collapsed so that each node spans two
// binary levels) or one triangle of a leaf (48 B, read as 64 B: the triangle array carries 16 B of tail
// padding). Every lane of a wave handles one item per loop step, so a step costs one memory latency whatever
// mix of nodes and triangles the lanes hold — traversal here is bound by that latency, not by HBM bandwidth,
// and the four-wide nodes halve the steps a ray needs to reach its leaves. The children a ray enters are
// ordered by entry distance (a four-element sorting network) and all but the nearest are pushed, farthest
// first. A lane whose ray has ended takes the next ray of its wave's share at the top of the next step (a
// ballot ranks the idle lanes), and the new ray's record is loaded together with the root node, so lanes stay
// busy until the wave's share is used up. A queue is one or more segments (QueueSegs); wave w starts on segment
// w % segs. With chunk counters the waves take the segment's 64-ray chunks one at a time (one atomic per chunk,
// each segment's counter on a line of its own) and, once it is used up, the next segments' chunks: the kernel
// ends when the last chunk anywhere is done, not when the slowest segment is. Without them a wave's share is
// every Wk-th chunk of its segment from chunk w / segs (Wk = the waves of the segment), with no atomics.
//
// Item encoding (stack entries are items too): an interior node is its index; a leaf run is
// kLeafItem | (remaining triangles - 1) << 27 | first triangle, so a leaf holds at most 16 triangles and
// a scene at most 2^27 triangles (both enforced at upload); an empty child slot is kNoItem.
//
// Traversal stack: kLdsDepth entries per lane live in LDS, column-major (entry e of lane l at
// e * kTraceBlock + l, conflict-free), deeper entries spill to a per-lane column of a global buffer. At most
// three entries per four-wide level; the binary build's depth bound (60 SAH levels, then halving to leaves of
// at most 16 of at most 2^27 triangles: 83 levels, 42 four-wide levels) bounds the stack at kStackDepth.
#pragma once

#include "pe_gpu_types.h"
#include "pe_math.h"

namespace pe {

constexpr int kLdsDepth = 16;
constexpr int kStackDepth = 128;
constexpr int kTraceBlock = 256;
constexpr uint32_t kLeafItem = 0x80000000u;
constexpr uint32_t kLeafFirstMask = (1u << 27) - 1u;
constexpr uint32_t kMaxLeafItemTris = 16;
constexpr uint32_t kNoItem = 0xFFFFFFFFu;
constexpr uint32_t kNoRay = 0xFFFFFFFFu;
constexpr uint64_t kMaxTraceTriangles = uint64_t{1} << 27;

// A ray queue of `segs` segments: segment k holds count(k) items at [k * seg_cap, k * seg_cap + count(k)). The
// wavefront queues have kQueueSegs segments, each appended to through a counter of its own (kSegStride uint32
// apart, so no two share a cache line): the appends of the whole grid would otherwise serialise on one address.
// A plain array of n rays is one segment (counts null).
constexpr uint32_t kQueueSegs = 64;
static_assert(kQueueSegs <= 64, "traversal checks every segment with one lane of a wave");
constexpr uint32_t kSegStride = 64;
struct QueueSegs {
    const uint32_t* counts;  // counts[k * kSegStride]; null: one segment of n items
    uint32_t n;
    uint32_t seg_cap;
    uint32_t segs;
    // Chunk counters (next[k * kSegStride], zero before the kernel): traversal hands out 64-ray chunks of each
    // segment through them, and a wave whose segment is used up moves on to the others. Null: every wave of a
    // segment takes a fixed interleaved share.
    uint32_t* next = nullptr;
    __device__ __forceinline__ uint32_t count(uint32_t k) const { return counts ? counts[k * kSegStride] : n; }
};

// The segment a worker (wave or block) w of W works on, its rank among that segment's workers and their number.
// Every segment has a worker when W >= segs.
struct SegShare {
    uint32_t seg, rank, workers;
};
__device__ __forceinline__ SegShare seg_share(uint32_t w, uint32_t W, uint32_t segs) {
    const uint32_t k = w % segs;
    return SegShare{k, w / segs, (W - 1u - k) / segs + 1u};
}

struct TraceStack {
    uint32_t* lds;      // this lane's column: entry e at lds[e * kTraceBlock]
    uint32_t* spill;    // this lane's column in the spill buffer: entry e at spill[(e - kLdsDepth) * spill_stride]
    uint32_t spill_stride;
    int sp;
    int max_sp = 0;         // deepest stack of this lane's traversals (kept by the counting variants)
    uint32_t hits = 0;      // closest-hit queries of this lane that found a triangle (counting variants)
    bool overflow = false;  // a push found the stack full and dropped its entry

    // The build bounds the tree depth below kStackDepth, so a push never finds the stack full; were it to,
    // the entry is dropped rather than written out of bounds, and the lane's overflow flag reports it (the
    // kernels count flagged lanes into a device counter that the host turns into an error).
    __device__ __forceinline__ void push(uint32_t v) {
        if (sp < kLdsDepth) {
            lds[sp * kTraceBlock] = v;
        } else if (sp < kStackDepth) {
            spill[static_cast<size_t>(sp - kLdsDepth) * spill_stride] = v;
        } else {
            overflow = true;
            return;
        }
        ++sp;
    }
    __device__ __forceinline__ uint32_t pop() {
        --sp;
        if (sp < kLdsDepth) return lds[sp * kTraceBlock];
        return spill[static_cast<size_t>(sp - kLdsDepth) * spill_stride];
    }
};

// Traversal diagnostics of a kernel, two 64-bit device counters: [0] the deepest stack any lane reached
// (counting variants only), [1] lanes whose stack overflowed. One atomic per wave.
__device__ __forceinline__ void report_stack(unsigned long long* diag, const TraceStack& st, bool with_depth) {
    if (!diag) return;
    const unsigned long long over = __ballot(st.overflow);
    if (with_depth) {
        int m = st.max_sp;
        for (int off = 32; off > 0; off >>= 1) m = max(m, __shfl_xor(m, off));
        if (__lane_id() == 0 && m > 0) atomicMax(diag, static_cast<unsigned long long>(m));
    }
    if (__lane_id() == 0 && over) atomicAdd(diag + 1, static_cast<unsigned long long>(__popcll(over)));
}

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

// Slab test of one child k of a four-wide node (components k of the SoA planes), with the exit pad.
__device__ __forceinline__ float box_near_far(float lx, float hx, float ly, float hy, float lz, float hz, V3 o, V3 inv,
                                              float tmin, float tmax, bool& hit) {
    const float x0 = (lx - o.x) * inv.x, x1 = (hx - o.x) * inv.x;
    const float y0 = (ly - o.y) * inv.y, y1 = (hy - o.y) * inv.y;
    const float z0 = (lz - o.z) * inv.z, z1 = (hz - o.z) * inv.z;
    const float tn = fmaxf(fmaxf(tmin, fminf(x0, x1)), fmaxf(fminf(y0, y1), fminf(z0, z1)));
    const float tf = fminf(fminf(tmax, fmaxf(x0, x1)), fminf(fmaxf(y0, y1), fmaxf(z0, z1))) * 1.0000004f;
    hit = tn <= tf;
    return tn;
}

__device__ __forceinline__ void cas_item(float& ka, uint32_t& ia, float& kb, uint32_t& ib) {
    if (kb < ka) {
        const float tk = ka;
        ka = kb;
        kb = tk;
        const uint32_t ti = ia;
        ia = ib;
        ib = ti;
    }
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
// The top kBvhTileNodes nodes of the tree, loaded into a block's LDS tile (8 float4 per node: the seven the
// traversal reads and the pad) by all of its threads before traversal starts.
__device__ __forceinline__ void load_node_tile(float4* tile, const BvhNode4* __restrict__ nodes, uint32_t num_nodes) {
    const uint32_t n = min(num_nodes, kBvhTileNodes) * 8u;
    const float4* src = reinterpret_cast<const float4*>(nodes);
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) tile[i] = src[i];
    __syncthreads();
}

template <bool kAny, bool kCount, class Job>
__device__ __forceinline__ void trace_queue(const Job& job, const QueueSegs& q, const BvhNode4* __restrict__ nodes,
                                            const TriAccel* __restrict__ tris, TraceStack& st, uint32_t& nv,
                                            uint32_t& tv, const float4* tile = nullptr, uint32_t tile_nodes = 0) {
    const uint32_t wpb = blockDim.x >> 6;
    const SegShare sh = seg_share(__builtin_amdgcn_readfirstlane(blockIdx.x * wpb + (threadIdx.x >> 6)),
                                  gridDim.x * wpb, q.segs);
    const uint32_t w = sh.rank, W = sh.workers;
    const uint32_t n = __builtin_amdgcn_readfirstlane(q.count(sh.seg));
    const uint32_t seg_base = sh.seg * q.seg_cap;
    const uint32_t nchunks = (n + 63u) >> 6;
    const uint64_t below = (uint64_t{1} << __lane_id()) - 1u;
    uint32_t taken = 0;  // rays of this wave's share handed out so far (wave-uniform)
    // Chunk hand-out (q.next): the segment the wave draws from, the current chunk's next slot and rays left in
    // it (all wave-uniform). At most 64 segments.
    uint32_t cseg = sh.seg, cbase = 0, cleft = 0;
    bool drained = false;
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
        uint64_t idle = __ballot(slot == kNoRay);
        if (q.next) {
            while (idle != 0 && !drained) {
                if (cleft == 0) {
                    // A chunk of the current segment; once that is used up, one round trip tells which segments
                    // still have chunks (lane k looks at segment k) and the wave moves to the first of them from
                    // its current one on, or ends when none has.
                    const uint32_t lane = __lane_id();
                    uint32_t nk = q.count(cseg), nck = (nk + 63u) >> 6;
                    uint32_t c = nck;
                    if (lane == 0 && *reinterpret_cast<volatile uint32_t*>(q.next + cseg * kSegStride) < nck)
                        c = atomicAdd(q.next + cseg * kSegStride, 1u);
                    c = __builtin_amdgcn_readfirstlane(__shfl(c, 0));
                    if (c < nck) {
                        cbase = cseg * q.seg_cap + (c << 6);
                        cleft = min(64u, nk - (c << 6));
                        continue;
                    }
                    bool open = false;
                    if (lane < q.segs) {
                        const uint32_t nl = (q.count(lane) + 63u) >> 6;
                        open = *reinterpret_cast<volatile uint32_t*>(q.next + lane * kSegStride) < nl;
                    }
                    const uint64_t avail = __ballot(open);
                    if (avail == 0) {
                        drained = true;
                        continue;
                    }
                    const uint64_t rot = cseg == 0 ? avail : (avail >> cseg) | (avail << (64u - cseg));
                    cseg = __builtin_amdgcn_readfirstlane((cseg + static_cast<uint32_t>(__ffsll(static_cast<long long>(rot)) - 1)) & 63u);
                    continue;
                }
                const uint32_t rank = static_cast<uint32_t>(__popcll(idle & below));
                if (slot == kNoRay && rank < cleft) {
                    slot = cbase + rank;
                    job.fetch(slot, ra, rb);
                    fresh = true;
                    item = 0;
                    st.sp = 0;
                }
                const uint32_t given = min(cleft, static_cast<uint32_t>(__popcll(idle)));
                cbase += given;
                cleft -= given;
                idle = __ballot(slot == kNoRay);
            }
        } else if (idle != 0 && w + (taken >> 6) * W < nchunks) {
            if (slot == kNoRay) {
                const uint32_t j = taken + static_cast<uint32_t>(__popcll(idle & below));
                const uint32_t c = w + (j >> 6) * W;
                const uint32_t idx = (c << 6) | (j & 63u);
                if (c < nchunks && idx < n) {
                    slot = seg_base + idx;
                    job.fetch(slot, ra, rb);
                    fresh = true;
                    item = 0;
                    st.sp = 0;
                }
            }
            taken += static_cast<uint32_t>(__popcll(idle));
        }
        if (__ballot(slot != kNoRay) == 0) break;
        if (slot == kNoRay) continue;

        // One item: a triangle of a leaf (the first 48 of 64 bytes read) or a four-wide node (128 bytes), the
        // top nodes from the block's LDS tile.
        const bool leaf = (item & kLeafItem) != 0u;
        float4 x0, x1, x2, x3, x4 = make_float4(0.0f, 0.0f, 0.0f, 0.0f), x5 = x4, x6 = x4;
        if (!leaf && item < tile_nodes) {
            const float4* t = tile + item * 8u;
            x0 = t[0];
            x1 = t[1];
            x2 = t[2];
            x3 = t[3];
            x4 = t[4];
            x5 = t[5];
            x6 = t[6];
        } else {
            const float4* p = leaf ? reinterpret_cast<const float4*>(tris + (item & kLeafFirstMask))
                                   : reinterpret_cast<const float4*>(nodes + item);
            x0 = p[0];
            x1 = p[1];
            x2 = p[2];
            x3 = p[3];
            if (!leaf) {
                x4 = p[4];
                x5 = p[5];
                x6 = p[6];
            }
        }
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
            const float tmax = kAny ? ray.tmax : ht;
            const uint32_t it[4] = {__float_as_uint(x6.x), __float_as_uint(x6.y), __float_as_uint(x6.z),
                                    __float_as_uint(x6.w)};
            const float lxs[4] = {x0.x, x0.y, x0.z, x0.w}, hxs[4] = {x1.x, x1.y, x1.z, x1.w};
            const float lys[4] = {x2.x, x2.y, x2.z, x2.w}, hys[4] = {x3.x, x3.y, x3.z, x3.w};
            const float lzs[4] = {x4.x, x4.y, x4.z, x4.w}, hzs[4] = {x5.x, x5.y, x5.z, x5.w};
            float key[4];
            uint32_t ci[4];
            int cnt = 0;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                bool h;
                const float tn = box_near_far(lxs[k], hxs[k], lys[k], hys[k], lzs[k], hzs[k], ray.o, inv, ray.tmin,
                                              tmax, h);
                h = h && it[k] != kNoItem;
                key[k] = h ? tn : 3.0e38f;
                ci[k] = h ? it[k] : kNoItem;
                cnt += h ? 1 : 0;
            }
            // Nearest first: a sorting network on the entry distances, then the far ones pushed farthest first.
            cas_item(key[0], ci[0], key[1], ci[1]);
            cas_item(key[2], ci[2], key[3], ci[3]);
            cas_item(key[0], ci[0], key[2], ci[2]);
            cas_item(key[1], ci[1], key[3], ci[3]);
            cas_item(key[1], ci[1], key[2], ci[2]);
            if (cnt > 3) st.push(ci[3]);
            if (cnt > 2) st.push(ci[2]);
            if (cnt > 1) st.push(ci[1]);
            if (cnt > 0) next = ci[0];
            if (kCount) st.max_sp = max(st.max_sp, st.sp);
        }
        if (next == kNoItem && !occluded && st.sp > 0) next = st.pop();
        if (occluded || next == kNoItem) {
            if constexpr (kAny) {
                job.any(slot, ray, occluded);
                slot = kNoRay;
            } else if ((kCount && htri >= 0 ? ++st.hits : 0u), job.closest(slot, ray, ht, htri, hu, hv, ra, rb)) {
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

// integrator.hip — the wavefront path tracer behind pe_render.
//
// PathTracer::Li (integrators/path_tracer.cpp:321-890) is split at its ray casts into four kernels that run
// over every path in flight:
//   generate   tracePixel (engine/renderer.cpp:139-162): Sobol dims 0-1, pixel-filter offset, camera ray
//   intersect  closest hit for the compacted extension-ray queue
//   shade      one path vertex: environment / emission with MIS, NEE (light pick, light sample, BSDF eval,
//              shadow ray appended to the shadow queue with its contribution held back), BSDF sample,
//              throughput, Russian roulette, next ray appended to the next queue
//   shadow     any hit for the shadow queue: an occluded ray zeroes its held-back contribution
//   film       Film::addSample (core/image/film.cpp:56-71) of every sample of a pixel, in sample order
// A held-back NEE contribution enters the path's radiance at the start of the path's next shade, or in the
// film kernel if the path ended: the radiance sum sees its terms in the reference's order.
//
// Queues hold the ray records themselves (origin | path id, direction; the shadow queue origin | tmax,
// direction | held-back slot), so a traversal lane starts a ray with one 32-byte load and no indirection;
// appends are wave-aggregated (one atomic per wave, ballot + popcount for the slot). Every kernel is
// persistent and reads its trip count from the device counter, so a render of any number of bounces issues
// no host synchronisation until the end of a launch.
//
// Path state is SoA in HBM (float4 / uint4 columns, 16-byte coalesced loads per lane).
#include "photon_hip.h"
#include "pe_guide_host.h"
#include "pe_internal.h"
#include "pe_lights.h"
#include "pe_shading.h"
#include "pe_sobol.h"
#include "pe_traverse.h"

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

namespace pe {

namespace {

#ifndef PE_SHADE_PLAIN_ATTR
#define PE_SHADE_PLAIN_ATTR
#endif
constexpr int kBlock = 256;
constexpr int kMaxShadowSamples = 4;
constexpr int kStatSlots = 16;  // 64-bit device counters of a render (PathBuffers::stats)
#ifndef PE_NODE_TILE
#define PE_NODE_TILE 1
#endif
constexpr bool kNodeTile = PE_NODE_TILE != 0;  // traversal reads the top of the BVH from an LDS tile
#ifndef PE_MAT_BUCKETS
#define PE_MAT_BUCKETS 4
#endif
constexpr uint32_t kMatBuckets = PE_MAT_BUCKETS;  // shade's sort keys for hits (1: hits / misses only)
#ifndef PE_SHADE_STAGE
#define PE_SHADE_STAGE 1
#endif
constexpr bool kShadeStage = PE_SHADE_STAGE != 0;  // shade loads a round's entries before sorting (LDS staging)
static_assert(kMatBuckets == 1 || kMatBuckets == static_cast<uint32_t>(kSortClasses), "one key per shading class");
constexpr uint64_t kAutoWindowSamples = uint64_t{1} << 28;  // default result buffer: 4 GB
constexpr uint64_t kPassWindowSamples = uint64_t{1} << 23;  // window limit with pass films (26 records a sample
                                                            // with 16 groups: 3.5 GB)
constexpr uint64_t kMaxWindowSamples = (uint64_t{1} << 31) - 64;  // window sample indices are 31-bit
constexpr uint64_t kDefaultPoolIds = uint64_t{1} << 26;  // default path ids in flight (about 15.6 GB)
constexpr uint64_t kMaxPoolIds = uint64_t{1} << 30;
constexpr int kChunkIterations = 2;  // iterations between two reads of the remaining-work counter

struct PathBuffers {
    float4* ro[2];    // ray queues (ping-pong), per slot: origin xyz, w = path id (uint bits)
    float4* rd[2];    //                                   direction xyz, w = entry tag (kCamEntry | g, or the depth)
    float4* thr;      // throughput rgb, w = last BSDF pdf
    float4* L;        // radiance rgb
    float4* last_p;   // previous vertex point (MIS)
    float4* last_n;   // previous vertex normal (MIS)
    uint4* ids;       // sample index g in the window, sample, Sobol pixel hash, packed bits (see pack_bits)
    uint4* ext;       // transport state beyond the surface bounce (see the kExt* fields)
    float4* mix;      // guiding: the mixture of every queue slot (kGuideMixFloat4 float4 each)
    float4* grec;     // guiding: two training records per path, 4 float4 each (see guide_collect)
    float4* pend;     // held-back NEE contribution per shadow sample, [pid * S + s]
    float4* res;      // per sample of the window: its radiance (L plus the held-back NEE), read by k_film
    float4* aov_res;  // per sample of the window, two float4: albedo rgb | coverage, shading normal xyz | state
                      // (-1 until the first segment has been resolved, then 0); null without AOV films
    float4* pass_res; // render passes: per window sample, LaunchParams::pass_count records in pass-major order
                      // (see pass_add); null without pass films
    float4* lobe;     // render passes: per path id, the first surface's lobe shares (3 float4: diffuse, glossy,
                      // transmission rgb)
    float4* pfrac;    // render passes: per held-back NEE record, its routing (3 float4, see credit_nee)
    float4* hits;     // per queue slot: t, triangle (int bits), u, v
    float4* sq_o;     // shadow queue: origin xyz, w = tmax
    float4* sq_d;     //               direction xyz, w = pend index pid * S + s (uint bits)
    uint32_t* fl[2];  // free path ids (ping-pong): pid | kDeferred when the path still has to be finalised
    uint32_t* ctr;    // segment counters (kCtr* below)
    uint32_t seg_cap; // slots per segment of the ray queues (S * seg_cap per shadow queue segment); segment k
                      // owns the path ids [k * seg_cap, (k + 1) * seg_cap)
    unsigned long long* stats;  // [0] extension rays, [1] shadow rays, [2..5] visits, [6] guide records,
                                // [7] guide queries, [8] deepest traversal stack, [9] stack overflows,
                                // [10] extension rays that hit (counting variant)
};

// Path regeneration. The wavefront keeps a pool of path ids; segment k of every queue owns the ids
// [k * seg_cap, (k + 1) * seg_cap), so a segment's queues never hold more entries than it has ids. An id is
// either a path in flight (its ray is in the current queue) or free (in the segment's free list). Each
// iteration:
//   k_regen      finalises the free ids whose path ended with a shadow ray outstanding, then gives free ids the
//                segment's next samples: their camera rays go to queue slots [base, base + cam), after the
//                continuation rays [0, base)
//   k_intersect  closest hit of every ray of the queue
//   k_shade      every entry: continuation rays to the next queue, shadow rays to the shadow queue, ended
//                paths to the next free list (finalised at once unless a shadow ray is outstanding)
//   k_shadow     any hit of the shadow queue
//   k_rotate     counters: next queue -> current, next free list -> current, samples handed out
// Samples are dealt to the segments in 64-sample chunks (chunk c to segment c mod 64); a path's result goes to
// its sample's slot of the window's result buffer, and k_film adds the results of every pixel in sample order,
// so the film does not depend on which id or iteration rendered a sample.
constexpr uint32_t kCamEntry = 0x80000000u;  // rd.w of a camera entry: kCamEntry | g
constexpr uint32_t kDeferred = 0x80000000u;  // free-list entry of a path with a shadow ray outstanding
constexpr uint32_t kPidMask = 0x7FFFFFFFu;
constexpr uint32_t kCtrBase = 0, kCtrCur = 1, kCtrCam = 2, kCtrNext = 3, kCtrShadow = 4, kCtrFree = 5,
                   kCtrFreeNext = 6, kCtrGen = 7, kCtrChunkIsect = 8, kCtrChunkShadow = 9, kCtrQueues = 10;
constexpr uint32_t kCtrTotal = kCtrQueues * kQueueSegs * kSegStride;  // paths and samples still to go
constexpr size_t kCtrBytes = (kCtrTotal + kSegStride) * sizeof(uint32_t);
__host__ __device__ __forceinline__ uint32_t* seg_counter(uint32_t* ctr, uint32_t q, uint32_t k) {
    return ctr + (q * kQueueSegs + k) * kSegStride;
}
__device__ __forceinline__ QueueSegs seg_queue(const PathBuffers& pb, uint32_t which) {
    return QueueSegs{seg_counter(pb.ctr, which, 0), 0u, pb.seg_cap, kQueueSegs};
}
// The current ray queue as k_intersect traces it (chunks handed out through their own counters).
__device__ __forceinline__ QueueSegs trace_ray_queue(const PathBuffers& pb) {
    return QueueSegs{seg_counter(pb.ctr, kCtrCur, 0), 0u, pb.seg_cap, kQueueSegs, seg_counter(pb.ctr, kCtrChunkIsect, 0)};
}
__device__ __forceinline__ QueueSegs shadow_queue(const PathBuffers& pb, int S) {
    return QueueSegs{seg_counter(pb.ctr, kCtrShadow, 0), 0u, pb.seg_cap * static_cast<uint32_t>(S), kQueueSegs,
                     seg_counter(pb.ctr, kCtrChunkShadow, 0)};
}
// Samples of a window of n that segment k renders: the 64-sample chunks c with c mod 64 = k, the last partial.
__host__ __device__ __forceinline__ uint32_t seg_samples(uint32_t n, uint32_t k) {
    const uint32_t chunks = (n + 63u) / 64u;
    uint32_t cnt = chunks > k ? ((chunks - 1u - k) / kQueueSegs + 1u) * 64u : 0u;
    if (chunks > 0u && (chunks - 1u) % kQueueSegs == k && (n & 63u) != 0u) cnt -= 64u - (n & 63u);
    return cnt;
}
// The window sample of the j-th sample segment k hands out.
__host__ __device__ __forceinline__ uint32_t seg_sample_index(uint32_t k, uint32_t j) {
    return ((k + kQueueSegs * (j >> 6)) << 6) | (j & 63u);
}
// Slots per segment for a pool of n path ids, rounded up to whole 256-slot shade rounds.
__host__ __device__ __forceinline__ uint32_t seg_capacity(uint32_t n) {
    const uint32_t per = (n + kQueueSegs - 1u) / kQueueSegs;
    return (per + 255u) / 256u * 256u;
}

// packed: dim (bits 0-15), depth (16-22), diffuse bounces (23-29), specular bounce (30), camera ray (31)
__device__ __forceinline__ uint32_t get_dim(uint32_t w) { return w & 0xFFFFu; }
__device__ __forceinline__ uint32_t get_depth(uint32_t w) { return (w >> 16) & 0x7Fu; }
__device__ __forceinline__ uint32_t get_diffuse(uint32_t w) { return (w >> 23) & 0x7Fu; }
__device__ __forceinline__ bool get_specular(uint32_t w) { return (w >> 30) & 1u; }
__device__ __forceinline__ bool get_camera(uint32_t w) { return (w >> 31) & 1u; }
__device__ __forceinline__ uint32_t pack_bits(uint32_t dim, uint32_t depth, uint32_t diffuse, bool spec, bool cam) {
    return (dim & 0xFFFFu) | ((depth & 0x7Fu) << 16) | ((diffuse & 0x7Fu) << 23) | (spec ? (1u << 30) : 0u) |
           (cam ? (1u << 31) : 0u);
}

// ext column: x = the path's wavelength in nm (float bits, 0 = RGB transport); y, z = the medium stack, four
// 16-bit material indices, bottom first (MediumStack, path_tracer.cpp:63-83); w = stack depth (bits 0-2),
// thin-wall pass-throughs so far (3-9), subsurface walk step (10-18), walk in progress (19). While a walk is in
// progress the walk's material index is in last_n.w.
constexpr uint32_t kExtDepthMask = 7u;
constexpr uint32_t kExtPassShift = 3u, kExtPassMask = 0x7Fu << kExtPassShift;
constexpr uint32_t kExtWalkStepShift = 10u, kExtWalkStepMask = 0x1FFu << kExtWalkStepShift;
constexpr uint32_t kExtWalk = 1u << 19;
constexpr uint32_t kExtRecShift = 20u, kExtRecMask = 3u << kExtRecShift;  // guiding records written (0..2)
constexpr uint32_t kExtRecOpen = 1u << 22;  // the last record waits for the radiance baseline
constexpr uint32_t kExtAo = 1u << 23;       // contact AO rays of the first vertex are out (approxAo)
constexpr uint32_t kAoTag = 1u << 31;       // a shadow-queue entry that is an AO ray (its slot: an AO record)
constexpr int kGuideRecords = 2;
constexpr uint32_t kMaxWalkSteps = 256;   // randomWalkSubsurface's step limit (path_tracer.cpp:270)
constexpr uint32_t kMaxPassThroughs = 64;  // path_tracer.cpp:783
constexpr uint32_t kMaxShadowQueries = 64;  // shadowTransmittance (path_tracer.cpp:180)
constexpr uint32_t kMaxMediumEvents = 1024;  // path_tracer.cpp:51

__device__ __forceinline__ uint32_t ext_pass_throughs(uint32_t w) { return (w & kExtPassMask) >> kExtPassShift; }
// The coverage of the thin walls a camera ray passed to a transparent background: 1 - the luminance of its
// throughput (Color3f::luminance, spectrum.h:68-70), uncontracted as the reference computes it.
__device__ __forceinline__ float wall_coverage(V3 t) {
#pragma clang fp contract(off)
    return fmaxf(0.0f, 1.0f - (0.212671f * t.x + 0.715160f * t.y + 0.072169f * t.z));
}
__device__ __forceinline__ uint32_t ext_walk_step(uint32_t w) { return (w & kExtWalkStepMask) >> kExtWalkStepShift; }
__device__ __forceinline__ uint64_t ext_stack(const uint4& e) { return e.y | (static_cast<uint64_t>(e.z) << 32); }
__device__ __forceinline__ void ext_set_stack(uint4& e, uint64_t st, uint32_t depth) {
    e.y = static_cast<uint32_t>(st);
    e.z = static_cast<uint32_t>(st >> 32);
    e.w = (e.w & ~kExtDepthMask) | depth;
}
// Material of the innermost medium, -1 outside every medium.
__device__ __forceinline__ int ext_media_top(const uint4& e) {
    const uint32_t n = e.w & kExtDepthMask;
    if (n == 0) return -1;
    return static_cast<int>((ext_stack(e) >> (16u * (n - 1u))) & 0xFFFFu);
}
__device__ __forceinline__ void ext_media_enter(uint4& e, int mat) {  // MediumStack::enter: full stacks ignore
    const uint32_t n = e.w & kExtDepthMask;
    if (n >= 4) return;
    const uint64_t st = ext_stack(e) | (static_cast<uint64_t>(mat & 0xFFFF) << (16u * n));
    ext_set_stack(e, st, n + 1u);
}
__device__ __forceinline__ void ext_media_leave(uint4& e, int mat) {  // MediumStack::leave: innermost match
    const uint32_t n = e.w & kExtDepthMask;
    const uint64_t st = ext_stack(e);
    for (int i = static_cast<int>(n) - 1; i >= 0; --i) {
        if (static_cast<int>((st >> (16 * i)) & 0xFFFFu) != (mat & 0xFFFF)) continue;
        const uint64_t below = i == 0 ? 0ull : (st & ((1ull << (16 * i)) - 1ull));
        const uint64_t above = i >= 3 ? 0ull : (st >> (16 * (i + 1)));
        ext_set_stack(e, below | (above << (16 * i)), n - 1u);
        return;
    }
}

// One pose of CameraGpuDesc (camera/camera.h:19-31).
struct CameraPose {
    V3 pos, ll, h, v, u, w, fwd;  // position, lowerLeft, horizontal, vertical, lens axes u / v, forward
    float radius;
};

struct LaunchParams {
    int width, height;
    int first_sample;
    int spl;              // samples per pixel in this window
    uint32_t num_pixels;  // pixels rendered by this window
    const uint32_t* pixel_list;  // their film indices in increasing order; null: every pixel (num_pixels = w * h)
    uint32_t num_paths;   // samples of the window: num_pixels * spl; sample g is pixel g % num_pixels, sample
                          // first_sample + g / num_pixels
    // Adaptive sampling: after adding this window's samples k_film clears active[pixel] of every pixel whose
    // relative error is below noise_threshold (adapt_check: the window ends where the reference checks).
    uint32_t* active;
    int adapt_check;
    float noise_threshold;
    int want_aov;         // record the camera vertex's AOVs (pb.aov_res)
    int want_passes;      // credit the render passes (pb.pass_res, pb.lobe, pb.pfrac)
    int num_groups;       // light groups of the scene (0: none)
    uint32_t pass_count;  // records per sample of pb.pass_res: 8 light passes, num_groups groups, 2 data
    int guided_bounces;   // guiding: vertices of a path below this depth are guided
    uint32_t grec_slots;  // guiding: slots of the training-record buffer
    float4* grec_out;     //           the buffer (3 float4 per record)
    uint64_t grec_salt;   //           per-window salt of the slot hash
    int max_depth, rr_depth, shadow_samples;
    float clamp_indirect;
    uint64_t seed;
    int filter_box;
    int transport;        // the scene needs the ext column (k_shade<true>)
    float guide_fraction; // guiding: probability of sampling the mixture at a guided vertex (0: BSDF only)
    int guide_records;    // guiding: capture training records
    float filter_width, filter_radius;
    const float* filter_cdf;  // 513 entries when not a box
    CameraPose cam, cam_end;  // CameraGpuDesc at shutter open and close
    int cam_ortho, cam_blades;
    float cam_blade_rotation;
    int cam_moving;           // motion blur with a shutter-end pose: rays blend the poses
    uint32_t camera_dims;     // Sobol dimensions tracePixel draws before Li: filter, lens (+ shutter time)
    float ao_strength;        // contact AO at the first vertex (0: off)
    uint32_t ao_base;         // pend index of the AO records (pool ids x shadow samples)
    int shadow_slots;         // shadow-queue entries per path id: the NEE samples, twice that with AO
};

// clamp01 (renderer.cpp:41-43) of c, with w.
__device__ __forceinline__ float4 clamped4(V3 c, float w) {
    return make_float4(fminf(fmaxf(c.x, 0.0f), 1.0f), fminf(fmaxf(c.y, 0.0f), 1.0f), fminf(fmaxf(c.z, 0.0f), 1.0f), w);
}

// Wave-aggregated append: returns this lane's slot in the queue behind `counter`.
__device__ __forceinline__ uint32_t wave_append(uint32_t* counter, bool want) {
    const unsigned long long mask = __ballot(want);
    if (!want) return 0;
    const int lane = __lane_id();
    const int leader = __ffsll(static_cast<long long>(mask)) - 1;
    const uint32_t before = static_cast<uint32_t>(__popcll(mask & ((1ull << lane) - 1ull)));
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(counter, static_cast<uint32_t>(__popcll(mask)));
    base = __shfl(base, leader);
    return base + before;
}

__device__ __forceinline__ float filter_sample(const LaunchParams& lp, float u) {  // pixel_filter.cpp:79-87
    if (lp.filter_box) return (u - 0.5f) * lp.filter_width;
    const float* cdf = lp.filter_cdf;
    int lo = 0, count = 513;
    while (count > 0) {
        const int step = count >> 1;
        const int it = lo + step;
        if (!(u < cdf[it])) {
            lo = it + 1;
            count -= step + 1;
        } else {
            count = step;
        }
    }
    int i = lo - 1;
    i = i < 0 ? 0 : (i > 511 ? 511 : i);
    const float c0 = cdf[i], c1 = cdf[i + 1];
    const float t = c1 > c0 ? (u - c0) / (c1 - c0) : 0.5f;
    return -lp.filter_radius + (static_cast<float>(i) + clampf(t, 0.0f, 1.0f)) * (lp.filter_width / 512.0f);
}

__device__ __forceinline__ float power_heuristic(int nf, float f_pdf, int ng, float g_pdf) {  // utils.h:108-113
    const float f = static_cast<float>(nf) * f_pdf, g = static_cast<float>(ng) * g_pdf;
    const float f2 = f * f;
    return f2 / (f2 + g * g);
}

// ---------------------------------------------------------------------------------------------- kernels

// A point uniform on the regular n-gon inscribed in the unit circle, turned by `rot`: the slice from u.x, a
// uniform point of that triangle from the stretched (u.x, u.y) (ThinLensCamera::samplePolygon,
// camera/thin_lens_camera.cpp:36-47; uniformSampleTriangle, sampling.h:97-100).
__device__ __forceinline__ void polygon_sample(int blades, float rot, float ux, float uy, float& x, float& y) {
    const float n = static_cast<float>(blades);
    const float scaled = fminf(ux * n, n - 1e-4f);
    const float k = floorf(scaled);
    const float a0 = rot + (2.0f * kPi) * k / n;
    const float a1 = a0 + (2.0f * kPi) / n;
    const float s0 = sqrtf(scaled - k);
    const float b0 = 1.0f - s0, b1 = uy * s0;
    x = cosf(a0) * b0 + cosf(a1) * b1;
    y = sinf(a0) * b0 + sinf(a1) * b1;
}

// Camera::generateRay for screen point (sx, sy), lens sample (lu, lv) and shutter time t (camera/camera.h:57-91;
// PerspectiveCamera, perspective_camera.cpp:36-40; ThinLensCamera, thin_lens_camera.cpp:49-61;
// OrthographicCamera, orthographic_camera.cpp:25-28).
__device__ __forceinline__ void camera_ray(const LaunchParams& lp, float sx, float sy, float lu, float lv, float t,
                                           V3& o, V3& d) {
    const CameraPose& a = lp.cam;
    if (!lp.cam_moving || t <= 0.0f) {
        const V3 target = a.ll + sx * a.h + sy * a.v;
        if (lp.cam_ortho) {
            o = target;
            d = a.fwd;
            return;
        }
        o = a.pos;
        if (a.radius > 0.0f) {
            float x, y;
            if (lp.cam_blades >= 3) polygon_sample(lp.cam_blades, lp.cam_blade_rotation, lu, lv, x, y);
            else concentric_disk(lu, lv, x, y);
            x *= a.radius;
            y *= a.radius;
            o = a.pos + a.u * x + a.w * y;
        }
        d = normalized(target - o);
        return;
    }
    const CameraPose& b = lp.cam_end;
    t = fminf(t, 1.0f);
    const float s = 1.0f - t;
    const V3 pos = a.pos * s + b.pos * t, ll = a.ll * s + b.ll * t;
    const V3 hor = a.h * s + b.h * t, ver = a.v * s + b.v * t;
    if (lp.cam_ortho) {
        o = ll + hor * sx + ver * sy;
        d = normalized(a.fwd * s + b.fwd * t);
        return;
    }
    const float radius = a.radius * s + b.radius * t;
    o = pos;
    if (radius > 0.0f) {
        const V3 cu = normalized(a.u * s + b.u * t), cv = normalized(a.w * s + b.w * t);
        float x, y;
        concentric_disk(lu, lv, x, y);
        o = pos + cu * (x * radius) + cv * (y * radius);
    }
    d = normalized(ll + hor * sx + ver * sy - o);
}

// tracePixel's camera ray (renderer.cpp:139-162) for window sample g: Sobol dimension 0 for the pixel filter,
// 1 for the lens and, with motion blur, 2 for the shutter time.
__device__ __forceinline__ void camera_sample_ray(const LaunchParams& lp, uint32_t g, V3& o, V3& d) {
    const uint32_t m = lp.num_pixels;
    const uint32_t q = g % m;
    const uint32_t pix = lp.pixel_list ? lp.pixel_list[q] : q;
    const uint32_t sample = static_cast<uint32_t>(lp.first_sample) + g / m;
    const int x = static_cast<int>(pix % static_cast<uint32_t>(lp.width));
    const int y = static_cast<int>(pix / static_cast<uint32_t>(lp.width));
    SobolState ss{sobol_pixel_hash(lp.seed, x, y), rev_bits(sample), 0u};
    float fu, fv, lu, lv, t = 0.0f;
    sobol_get2d(ss, fu, fv);  // pixel filter (renderer.cpp:148)
    sobol_get2d(ss, lu, lv);  // lens sample (renderer.cpp:151), drawn for every camera
    if (lp.camera_dims > 2) t = sobol_get1d(ss);  // shutter time (renderer.cpp:152-153)
    const float ox = filter_sample(lp, fu), oy = filter_sample(lp, fv);
    const float sx = (static_cast<float>(x) + 0.5f + ox) / static_cast<float>(lp.width);
    const float sy = 1.0f - (static_cast<float>(y) + 0.5f + oy) / static_cast<float>(lp.height);
    camera_ray(lp, sx, sy, lu, lv, t, o, d);
}

// The state a camera path starts with (tracePixel, renderer.cpp:139-162; PathTracer::Li, path_tracer.cpp:321-360).
__device__ __forceinline__ uint4 camera_ids(const LaunchParams& lp, uint32_t g) {
    const uint32_t m = lp.num_pixels;
    const uint32_t q = g % m;
    const uint32_t pix = lp.pixel_list ? lp.pixel_list[q] : q;
    const uint32_t sample = static_cast<uint32_t>(lp.first_sample) + g / m;
    const int x = static_cast<int>(pix % static_cast<uint32_t>(lp.width));
    const int y = static_cast<int>(pix / static_cast<uint32_t>(lp.width));
    return make_uint4(g, sample, sobol_pixel_hash(lp.seed, x, y), pack_bits(lp.camera_dims, 0, 0, true, true));
}

// Start of a window: every path id free, nothing handed out; held-back slots cleared (a cancelled window may
// leave contributions behind).
// Segment k uses the first ids_per_seg of its ids in this window (the window's pool), fewer when it renders
// fewer samples.
__global__ void __launch_bounds__(kBlock) k_window_init(LaunchParams lp, PathBuffers pb, uint32_t ids_per_seg) {
    const uint32_t stride = gridDim.x * blockDim.x;
    const uint32_t ids = kQueueSegs * ids_per_seg;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < ids; i += stride) {
        const uint32_t k = i / ids_per_seg, e = i - k * ids_per_seg;
        if (e >= seg_samples(lp.num_paths, k)) continue;
        const uint32_t pid = k * pb.seg_cap + e;
        pb.fl[0][pid] = pid;
        for (int s = 0; s < lp.shadow_samples; ++s) pb.pend[static_cast<size_t>(pid) * lp.shadow_samples + s] = make_float4(0, 0, 0, 0);
    }
    if (blockIdx.x == 0 && threadIdx.x < kQueueSegs) {
        const uint32_t k = threadIdx.x;
        const uint32_t total = seg_samples(lp.num_paths, k);
        const uint32_t cam = min(ids_per_seg, total);
        *seg_counter(pb.ctr, kCtrBase, k) = 0u;
        *seg_counter(pb.ctr, kCtrCam, k) = cam;
        *seg_counter(pb.ctr, kCtrCur, k) = cam;
        *seg_counter(pb.ctr, kCtrNext, k) = 0u;
        *seg_counter(pb.ctr, kCtrShadow, k) = 0u;
        *seg_counter(pb.ctr, kCtrFree, k) = cam;
        *seg_counter(pb.ctr, kCtrFreeNext, k) = 0u;
        *seg_counter(pb.ctr, kCtrGen, k) = 0u;
        *seg_counter(pb.ctr, kCtrChunkIsect, k) = 0u;
        *seg_counter(pb.ctr, kCtrChunkShadow, k) = 0u;
        if (k == 0) pb.ctr[kCtrTotal] = lp.num_paths;
    }
}

// Sum of a per-lane counter over the wave, added once per wave to a 64-bit device counter.
__device__ __forceinline__ void wave_count(unsigned long long* dst, uint32_t v) {
    unsigned long long x = v;
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off);
    if (__lane_id() == 0 && x) atomicAdd(dst, x);
}

// Extension rays: the current ray queue, hits written per queue slot.
struct ExtensionJob {
    const float4* ro;
    const float4* rd;
    float4* hits;
    __device__ __forceinline__ void fetch(uint32_t i, float4& a, float4& b) const {
        a = ro[i];
        b = rd[i];
    }
    __device__ __forceinline__ TraceRay decode(const float4& a, const float4& b) const {
        return TraceRay{mk(a.x, a.y, a.z), mk(b.x, b.y, b.z), 0.0f, 1e30f, 0u};
    }
    __device__ __forceinline__ bool closest(uint32_t i, const TraceRay&, float t, int32_t tri, float u, float v, float4&,
                                            float4&) const {
        hits[i] = make_float4(t, __int_as_float(tri), u, v);
        return false;
    }
};

// Shadow rays: an occluded ray zeroes its held-back contribution.
struct ShadowJob {
    const float4* so;
    const float4* sd;
    float4* pend;
    __device__ __forceinline__ void fetch(uint32_t i, float4& a, float4& b) const {
        a = so[i];
        b = sd[i];
    }
    __device__ __forceinline__ TraceRay decode(const float4& a, const float4& b) const {
        return TraceRay{mk(a.x, a.y, a.z), mk(b.x, b.y, b.z), 0.0f, a.w, __float_as_uint(b.w)};
    }
    __device__ __forceinline__ void any(uint32_t, const TraceRay& r, bool occluded) const {
        if (occluded) pend[r.tag & ~kAoTag] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
};

// Scene::isLightLinked / isShadowLinked (scene.cpp:237-256): `shadow` picks the light's shadow set.
__device__ __forceinline__ bool light_linked(const SceneView& sc, int li, uint64_t uid, bool shadow) {
    const uint32_t* w = sc.link_info + 6 * li;
    const uint32_t mode = shadow ? w[1] : w[0];
    if (mode == PE_LINK_ALL) return true;
    const uint32_t first = shadow ? w[4] : w[2], count = shadow ? w[5] : w[3];
    bool found = false;
    for (uint32_t k = 0; k < count; ++k) found = found || sc.link_uids[first + k] == uid;
    return mode == PE_LINK_INCLUDE ? found : !found;
}
__device__ __forceinline__ uint64_t receiver_uid(const SceneView& sc, int32_t tri) {
    return tri >= 0 ? sc.node_uids[sc.shade[tri].shape] : 0ull;
}

// approxAo's factor (path_tracer.cpp:42-43) from a path's AO records (1 each unoccluded ray, 0 each blocked one).
__device__ __forceinline__ float ao_factor(const LaunchParams& lp, const PathBuffers& pb, uint32_t pid, int S) {
    float occ = 0.0f;
    for (int a = 0; a < S; ++a) occ += 1.0f - pb.pend[lp.ao_base + static_cast<size_t>(pid) * S + a].x;
    const float visibility = 1.0f - (occ / static_cast<float>(S));
    return 1.0f + lp.ao_strength * (visibility - 1.0f);
}

__device__ __forceinline__ V3 clamp_indirect(V3 c, float limit) {
    if (limit > 0.0f) {
        const float m = fmaxf(c.x, fmaxf(c.y, c.z));
        if (m > limit) return c * (limit / m);
    }
    return c;
}

// ------------------------------------------------------------------------------------------ render passes
// PathTracer::Li's crediting (path_tracer.cpp:372-395): a contribution found at the first surface goes to the
// Emission / Background pass or, for NEE, to the direct passes by the lobe shares of its light sample; one found
// after the first bounce to the lobe shares of the direction the first surface sampled (the path's lobe
// column), direct when found by the first bounce and indirect later; every light contribution also to its light's
// group. A sample's records live at pass_res[k * num_paths + g]: k < 8 the LightPass order, then the groups, then
// two data records: the first surface's point | distance from the camera, and its shape | material (int bits) |
// 0 | state (-1 until the camera ray's first segment is resolved). Paths own their sample, so the read-modify-
// write needs no atomics, and a pass receives its terms in the reference's order.
constexpr uint32_t kLightPasses = 8;
enum : uint32_t { kPassDiffuseDirect = 0, kPassGlossyDirect = 2, kPassTransmissionDirect = 4, kPassEmission = 6,
                  kPassBackground = 7 };

__device__ __forceinline__ void pass_add(const PathBuffers& pb, const LaunchParams& lp, uint32_t k, uint32_t g, V3 v) {
    float4* r = pb.pass_res + static_cast<size_t>(k) * lp.num_paths + g;
    const float4 a = *r;
    *r = make_float4(a.x + v.x, a.y + v.y, a.z + v.z, a.w);
}
// creditGroup (path_tracer.cpp:386-390): a group index out of range counts as group 0.
__device__ __forceinline__ void credit_group(const PathBuffers& pb, const LaunchParams& lp, uint32_t g, int group,
                                             V3 added) {
    if (lp.num_groups <= 0 || is_black(added)) return;
    const int k = group >= 0 && group < lp.num_groups ? group : 0;
    pass_add(pb, lp, kLightPasses + static_cast<uint32_t>(k), g, added);
}
// creditPath (path_tracer.cpp:392-395): passAt(lobe, direct) is pass 2 lobe + (direct ? 0 : 1).
__device__ __forceinline__ void credit_path(const PathBuffers& pb, const LaunchParams& lp, uint32_t pid, uint32_t g,
                                            V3 added, bool direct) {
    for (uint32_t k = 0; k < 3; ++k) {
        const float4 f = pb.lobe[3 * static_cast<size_t>(pid) + k];
        pass_add(pb, lp, 2 * k + (direct ? 0u : 1u), g, added * mk(f.x, f.y, f.z));
    }
}
// Per-channel lobe shares of a split (path_tracer.cpp:683-690, :805-810): part / sum, or all diffuse where the
// sum is not positive.
__device__ __forceinline__ void lobe_shares(const LobeSplit& sp, V3 out[3]) {
    const V3 sum = sp.diffuse + sp.glossy + sp.transmission;
    const bool ox = sum.x > 0.0f, oy = sum.y > 0.0f, oz = sum.z > 0.0f;
    out[0] = mk(ox ? sp.diffuse.x / sum.x : 1.0f, oy ? sp.diffuse.y / sum.y : 1.0f, oz ? sp.diffuse.z / sum.z : 1.0f);
    out[1] = mk(ox ? sp.glossy.x / sum.x : 0.0f, oy ? sp.glossy.y / sum.y : 0.0f, oz ? sp.glossy.z / sum.z : 0.0f);
    out[2] = mk(ox ? sp.transmission.x / sum.x : 0.0f, oy ? sp.transmission.y / sum.y : 0.0f,
                oz ? sp.transmission.z / sum.z : 0.0f);
}
// A held-back NEE record's routing, written when its shadow ray is issued: pfrac[3 k + j] = the lobe shares of
// its light sample (first-surface records) in xyz, and in the first record's w (uint bits) bit 0 = first
// surface, bits 1.. = light group + 1.
__device__ __forceinline__ uint32_t nee_route(bool first_surface, int group) {
    return (first_surface ? 1u : 0u) | (static_cast<uint32_t>(group + 1) << 1);
}
// The NEE record k of path pid, resolved to c by its shadow ray, into the passes (path_tracer.cpp:679-693).
__device__ __forceinline__ void credit_nee(const PathBuffers& pb, const LaunchParams& lp, uint32_t pid, uint32_t g,
                                           size_t k, V3 c) {
    if (is_black(c)) return;
    const float4 f0 = pb.pfrac[3 * k];
    const uint32_t route = __float_as_uint(f0.w);
    credit_group(pb, lp, g, static_cast<int>(route >> 1) - 1, c);
    if (route & 1u) {
        const float4 f1 = pb.pfrac[3 * k + 1], f2 = pb.pfrac[3 * k + 2];
        pass_add(pb, lp, kPassDiffuseDirect, g, c * mk(f0.x, f0.y, f0.z));
        pass_add(pb, lp, kPassGlossyDirect, g, c * mk(f1.x, f1.y, f1.z));
        pass_add(pb, lp, kPassTransmissionDirect, g, c * mk(f2.x, f2.y, f2.z));
    } else {
        credit_path(pb, lp, pid, g, c, false);
    }
}
// A contribution found on the path (emission, environment): the first surface's own pass at depth 0, else the
// path's lobe shares; and its group.
__device__ __forceinline__ void credit_found(const PathBuffers& pb, const LaunchParams& lp, uint32_t pid, uint32_t g,
                                             uint32_t depth, uint32_t first_pass, int group, V3 added) {
    if (depth == 0) pass_add(pb, lp, first_pass, g, added);
    else credit_path(pb, lp, pid, g, added, depth == 1);
    credit_group(pb, lp, g, group, added);
}

// Held-back NEE records in scenes with thin walls: pend.w carries, as uint bits, the deferred indirect clamp
// (bit 0) and the number of closest-hit queries the shadow ray has made (bits 1-7).
constexpr uint32_t kPendClamp = 1u;

// Shadow rays in scenes with thin walls (shadowTransmittance, path_tracer.cpp:172-202): a closest-hit query;
// a thin wall multiplies the held-back contribution by its transmittance and the query continues from beyond
// it with the remaining distance, anything else zeroes it; at most 64 queries. The indirect clamp applies to
// the attenuated contribution, once the ray has reached the light.
struct TransmitShadowJob {
    const float4* so;
    const float4* sd;
    float4* pend;
    SceneView sc;
    float clamp_limit;
    __device__ __forceinline__ void fetch(uint32_t i, float4& a, float4& b) const {
        a = so[i];
        b = sd[i];
    }
    __device__ __forceinline__ TraceRay decode(const float4& a, const float4& b) const {
        return TraceRay{mk(a.x, a.y, a.z), mk(b.x, b.y, b.z), 0.0f, a.w, __float_as_uint(b.w)};
    }
    __device__ __forceinline__ bool closest(uint32_t, const TraceRay& r, float t, int32_t tri, float u, float v,
                                            float4& a, float4& b) const {
        if (r.tag & kAoTag) {  // an AO ray: any surface within reach blocks it (intersectAny, path_tracer.cpp:40)
            if (tri >= 0) pend[r.tag & ~kAoTag] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            return false;
        }
        float4 p = pend[r.tag];
        const uint32_t bits = __float_as_uint(p.w);
        V3 c = mk(p.x, p.y, p.z);
        bool more = false;
        if (tri >= 0) {
            Surface si;
            fill_hit(sc, tri, r.d, u, v, si);
            const GpuMaterial& m = sc.materials[si.material];
            const uint32_t queries = ((bits >> 1) & 0x7Fu) + 1u;
            // A blocker the light's shadow links leave out is passed as if absent (path_tracer.cpp:187-192); the
            // record carries the light index + 1 above bit 8 (0: the environment, which has no links).
            const int li = static_cast<int>(bits >> 8) - 1;
            if (sc.has_shadow_links && li >= 0 && !light_linked(sc, li, receiver_uid(sc, tri), true)) {
                const float remaining = r.tmax - t;
                if (remaining > 0.0f && queries < kMaxShadowQueries) {
                    const V3 o = offset_ray_origin(si.point, si.p_error, length_sq(si.ng) > 0.0f ? si.ng : si.normal, r.d);
                    a = make_float4(o.x, o.y, o.z, remaining);
                    b = make_float4(r.d.x, r.d.y, r.d.z, __uint_as_float(r.tag));
                    pend[r.tag] = make_float4(c.x, c.y, c.z, __uint_as_float((bits & ~0xFEu) | (queries << 1)));
                    return true;
                }
                if (remaining > 0.0f) c = splat(0.0f);
                if (bits & kPendClamp) c = clamp_indirect(c, clamp_limit);
                pend[r.tag] = make_float4(c.x, c.y, c.z, 0.0f);
                return false;
            }
            // Thin walls pass the shadow ray on with their transmittance, cutouts with 1 - opacity
            // (shadowTransmittance, dielectric.cpp / disney.cpp:511-513); anything else blocks it.
            if (!(m.flags & (kMatFlagThinWall | kMatFlagCutout))) {
                c = splat(0.0f);
            } else {
                c = c * ((m.flags & kMatFlagThinWall) ? thin_wall_transmittance(m, si, r.d) : splat(1.0f - m.opacity));
                const float remaining = r.tmax - t;
                if (!is_black(c) && remaining > 0.0f) {
                    if (queries >= kMaxShadowQueries) {
                        c = splat(0.0f);
                    } else {
                        const V3 o = offset_ray_origin(si.point, si.p_error, length_sq(si.ng) > 0.0f ? si.ng : si.normal, r.d);
                        a = make_float4(o.x, o.y, o.z, remaining);
                        b = make_float4(r.d.x, r.d.y, r.d.z, __uint_as_float(r.tag));
                        more = true;
                    }
                }
            }
            if (more) {
                pend[r.tag] = make_float4(c.x, c.y, c.z, __uint_as_float((bits & ~0xFEu) | (queries << 1)));
                return true;
            }
        }
        if (bits & kPendClamp) c = clamp_indirect(c, clamp_limit);
        pend[r.tag] = make_float4(c.x, c.y, c.z, 0.0f);
        return false;
    }
};

template <bool kCount>
__global__ void __launch_bounds__(kTraceBlock) k_intersect(SceneView sc, PathBuffers pb, int cur, uint32_t* spill) {
    __shared__ uint32_t lds_stack[kLdsDepth * kTraceBlock];
    __shared__ float4 node_tile[kBvhTileNodes * 8];
    load_node_tile(node_tile, sc.nodes, kNodeTile ? sc.num_nodes : 0u);
    const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    TraceStack st{lds_stack + threadIdx.x, spill + gid, gridDim.x * blockDim.x, 0};
    uint32_t nv = 0, tv = 0;
    trace_queue<false, kCount>(ExtensionJob{pb.ro[cur], pb.rd[cur], pb.hits}, trace_ray_queue(pb), sc.nodes, sc.tris,
                               st, nv, tv,
                               node_tile, kNodeTile ? min(sc.num_nodes, kBvhTileNodes) : 0u);
    report_stack(pb.stats + 8, st, kCount);
    if (kCount) {
        wave_count(&pb.stats[2], nv);
        wave_count(&pb.stats[3], tv);
        wave_count(&pb.stats[10], st.hits);
    }
}

__device__ __forceinline__ uint64_t mix64(uint64_t x) {  // splitmix64 finaliser
    x ^= x >> 30;
    x *= 0xbf58476d1ce4e5b9ull;
    x ^= x >> 27;
    x *= 0x94d049bb133111ebull;
    return x ^ (x >> 31);
}

// A path's result: its radiance plus the held-back NEE of its last vertex, into its sample's slot
// (the sum Film::addSample receives, path_tracer.cpp:355-359 and the end of Li).
// `ao`: the first vertex's contact-AO factor when its NEE records are folded (1 otherwise).
__device__ __forceinline__ V3 finish_radiance(const LaunchParams& lp, const PathBuffers& pb, uint32_t pid, uint32_t g,
                                              V3 L, int S, float ao = 1.0f) {
    for (int s = 0; s < S; ++s) {
        const size_t k = static_cast<size_t>(pid) * S + s;
        const float4 e = pb.pend[k];
        const V3 c = mk(e.x, e.y, e.z) * ao;
        L += c;
        if (lp.want_passes) credit_nee(pb, lp, pid, g, k, c);
        pb.pend[k] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
    return L;
}

// A segment inside a scattering medium (path_tracer.cpp:457-479): a free flight with spectral MIS towards the
// next surface at `seg` (or infinity); a scattering event there restarts the path at that point in a
// Henyey-Greenstein direction about the unit ray direction, counts against the 1024 events a path may have and,
// past 8 of them, faces Russian roulette. Returns true when the path scattered (and still lives: new_o /
// new_d hold its ray); false when the segment reached its end, with the throughput attenuated (alive false if
// it went black or non-finite). One 1D draw for the channel and one for the distance, then a 2D draw for the
// direction and, past 8 events, one for the roulette.
__device__ __forceinline__ bool medium_scatter(const GpuMaterial& m, V3 origin, V3 dir, float seg, bool hit,
                                               SobolState& ss, V3& thr, uint32_t& events, bool& alive, V3& new_o,
                                               V3& new_d) {
    float t = 0.0f;
    if (free_flight(m.sigma_a + m.sigma_s, m.sigma_s, seg, hit, thr, ss, thr, t)) {
        if (++events > kMaxMediumEvents || !finite3(thr)) {
            alive = false;
            return false;
        }
        const V3 d = dir / length(dir);
        float u0, u1;
        sobol_get2d(ss, u0, u1);
        new_o = origin + d * t;
        new_d = sample_hg(d, m.medium_g, u0, u1);
        if (events > 8) {
            const float mx = fmaxf(thr.x, fmaxf(thr.y, thr.z));
            if (mx < 1.0f) {
                if (!(mx > 0.0f) || sobol_get1d(ss) >= mx) alive = false;
                else thr = thr / mx;
            }
        }
        return alive;
    }
    if (is_black(thr) || !finite3(thr)) alive = false;
    return false;
}

// Training records of an ended path (guiding). A record of a guided vertex holds the point, the outgoing and
// the sampled direction, the throughput before the vertex and the path radiance once the vertex's own NEE had
// entered it (written at the start of the next vertex: a record is closed once that baseline is in); the
// radiance the path then brought back along the sampled direction, divided by that throughput, is
// f cos L_i / pdf for the sampled direction: its luminance weights the record. Records go to hash-chosen
// slots of the guide's buffer (later ones overwrite), so any contiguous batch of the buffer is a spatially
// unbiased sample. Returns the number of records written.
__device__ __forceinline__ uint32_t guide_collect(const LaunchParams& lp, const PathBuffers& pb, uint32_t pid,
                                                  uint32_t g, V3 L, uint32_t extw) {
    uint32_t closed = (extw & kExtRecMask) >> kExtRecShift;
    if (extw & kExtRecOpen) --closed;
    uint32_t written = 0;
    for (uint32_t r = 0; r < closed; ++r) {
        const float4* rec = pb.grec + static_cast<size_t>(pid) * 8 + 4 * r;
        const float4 base = rec[3];
        const float4 a = rec[0], b = rec[1], c = rec[2];
        const V3 inc = L - mk(base.x, base.y, base.z);
        const V3 li = mk(c.y > 0.0f ? inc.x / c.y : 0.0f, c.z > 0.0f ? inc.y / c.z : 0.0f, c.w > 0.0f ? inc.z / c.w : 0.0f);
        const float w = luminance(li);
        if (!(w > 0.0f) || !finite_f(w)) continue;
        const uint32_t slot = static_cast<uint32_t>(mix64((static_cast<uint64_t>(g) << 1 | r) ^ lp.grec_salt) % lp.grec_slots);
        lp.grec_out[3 * static_cast<size_t>(slot)] = make_float4(a.x, a.y, a.z, w);
        lp.grec_out[3 * static_cast<size_t>(slot) + 1] = make_float4(a.w, b.x, b.y, 0.0f);
        lp.grec_out[3 * static_cast<size_t>(slot) + 2] = make_float4(b.z, b.w, c.x, 0.0f);
        ++written;
    }
    return written;
}

// Free ids of the current free list: the deferred ones are finalised (their shadow rays have been traced), and
// the first cam of every segment start the segment's next samples: their camera rays go to the current queue
// after the continuation rays, tagged with the sample index, for k_intersect to trace with the others.
template <bool kGuide>
__global__ void __launch_bounds__(kBlock) k_regen(LaunchParams lp, PathBuffers pb, int cur) {
    const SegShare sh = seg_share(blockIdx.x, gridDim.x, kQueueSegs);
    const uint32_t k = sh.seg, seg_base = k * pb.seg_cap;
    const uint32_t n = *seg_counter(pb.ctr, kCtrFree, k);
    const uint32_t cam = *seg_counter(pb.ctr, kCtrCam, k);
    const uint32_t base = *seg_counter(pb.ctr, kCtrBase, k);
    const uint32_t gen = *seg_counter(pb.ctr, kCtrGen, k);
    const uint32_t* fl = pb.fl[cur] + seg_base;
    float4* ro = pb.ro[cur] + seg_base + base;
    float4* rd = pb.rd[cur] + seg_base + base;
    const int S = lp.shadow_samples;
    uint32_t written = 0;
    for (uint32_t e = sh.rank * blockDim.x + threadIdx.x; e < n; e += sh.workers * blockDim.x) {
        const uint32_t entry = fl[e];
        const uint32_t pid = entry & kPidMask;
        if (entry & kDeferred) {
            const uint4 ids = pb.ids[pid];
            const float4 L4 = pb.L[pid];
            // A path that ended at its first vertex with contact-AO rays out: their factor scales that vertex's NEE.
            const float ao = lp.ao_strength > 0.0f && (pb.ext[pid].w & kExtAo) ? ao_factor(lp, pb, pid, S) : 1.0f;
            const V3 L = finish_radiance(lp, pb, pid, ids.x, mk(L4.x, L4.y, L4.z), S, ao);
            pb.res[ids.x] = make_float4(L.x, L.y, L.z, 0.0f);
            if (kGuide && lp.guide_records) written += guide_collect(lp, pb, pid, ids.x, L, pb.ext[pid].w);
        }
        if (e < cam) {
            const uint32_t g = seg_sample_index(k, gen + e);
            V3 o, d;
            camera_sample_ray(lp, g, o, d);
            ro[e] = make_float4(o.x, o.y, o.z, __uint_as_float(pid));
            rd[e] = make_float4(d.x, d.y, d.z, __uint_as_float(kCamEntry | g));
            if (pb.aov_res) {
                pb.aov_res[2 * static_cast<size_t>(g)] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                pb.aov_res[2 * static_cast<size_t>(g) + 1] = make_float4(0.0f, 0.0f, 0.0f, -1.0f);
            }
            if (lp.want_passes) {
                for (uint32_t j = 0; j < lp.pass_count; ++j)
                    pb.pass_res[static_cast<size_t>(j) * lp.num_paths + g] =
                        make_float4(0.0f, 0.0f, 0.0f, j + 1 == lp.pass_count ? -1.0f : 0.0f);
            }
        }
    }
    if (kGuide) wave_count(&pb.stats[6], written);
}

template <bool kCount>
__global__ void __launch_bounds__(kTraceBlock) k_shadow(SceneView sc, PathBuffers pb, int S, uint32_t* spill) {
    __shared__ uint32_t lds_stack[kLdsDepth * kTraceBlock];
    __shared__ float4 node_tile[kBvhTileNodes * 8];
    load_node_tile(node_tile, sc.nodes, kNodeTile ? sc.num_nodes : 0u);
    const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    TraceStack st{lds_stack + threadIdx.x, spill + gid, gridDim.x * blockDim.x, 0};
    uint32_t nv = 0, tv = 0;
    trace_queue<true, kCount>(ShadowJob{pb.sq_o, pb.sq_d, pb.pend}, shadow_queue(pb, S), sc.nodes, sc.tris, st, nv, tv,
                               node_tile, kNodeTile ? min(sc.num_nodes, kBvhTileNodes) : 0u);
    report_stack(pb.stats + 8, st, kCount);
    if (kCount) {
        wave_count(&pb.stats[4], nv);
        wave_count(&pb.stats[5], tv);
    }
}

template <bool kCount>
__global__ void __launch_bounds__(kTraceBlock) k_shadow_transmit(SceneView sc, PathBuffers pb, int S, float clamp_limit,
                                                                 uint32_t* spill) {
    __shared__ uint32_t lds_stack[kLdsDepth * kTraceBlock];
    __shared__ float4 node_tile[kBvhTileNodes * 8];
    load_node_tile(node_tile, sc.nodes, kNodeTile ? sc.num_nodes : 0u);
    const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    TraceStack st{lds_stack + threadIdx.x, spill + gid, gridDim.x * blockDim.x, 0};
    uint32_t nv = 0, tv = 0;
    trace_queue<false, kCount>(TransmitShadowJob{pb.sq_o, pb.sq_d, pb.pend, sc, clamp_limit}, shadow_queue(pb, S),
                               sc.nodes, sc.tris, st, nv, tv,
                               node_tile, kNodeTile ? min(sc.num_nodes, kBvhTileNodes) : 0u);
    report_stack(pb.stats + 8, st, kCount);
    if (kCount) {
        wave_count(&pb.stats[4], nv);
        wave_count(&pb.stats[5], tv);
    }
}

// Guiding queries of the current ray queue: the hit point and the outgoing direction of every slot that hit.
struct RenderQueries {
    const float4* ro;
    const float4* rd;
    const float4* hits;
    const uint32_t* counts;   // segment counters of the current queue
    uint32_t seg_cap;
    float4* mix;
    uint32_t guided_bounces;  // entries of paths at this depth or deeper are not guided
    const uint32_t* bases;    // segment counters of the first camera entry (k_regen's kCtrBase)
    bool camera_only;         // only camera entries are guided: a segment's queries start at its camera entries
    __device__ __forceinline__ bool fetch(uint32_t i, V3& p, V3& d) const {
        const float4 h = hits[i];
        if (__float_as_int(h.y) < 0) return false;
        const float4 o = ro[i], r = rd[i];
        const uint32_t tag = __float_as_uint(r.w);
        if (!(tag & kCamEntry) && tag >= guided_bounces) return false;
        const V3 dir = mk(r.x, r.y, r.z);
        p = mk(o.x, o.y, o.z) + dir * h.x;
        d = -dir;
        return true;
    }
    __device__ __forceinline__ QueueSegs segs(uint32_t) const { return QueueSegs{counts, 0u, seg_cap, kQueueSegs}; }
    __device__ __forceinline__ uint32_t begin(uint32_t k) const { return camera_only ? bases[k * kSegStride] : 0u; }
    __device__ __forceinline__ void store_part(uint32_t i, int part, float4 a, float4 b, float2 cdf, float2 coef) const {
        guide_store_part(mix + static_cast<size_t>(i) * kGuideMixFloat4, part, a, b, cdf, coef);
    }
};


// The environment seen by a ray that left the scene, with MIS against NEE unless the last bounce was specular
// (path_tracer.cpp:482-514): what the path radiance receives (black for a non-finite value).
__device__ __forceinline__ V3 miss_radiance(const SceneView& sc, const LaunchParams& lp, V3 dir, V3 thr, bool specular,
                                            float last_pdf, V3 last_p, V3 last_n, uint32_t diffuse) {
    const V3 le = env_eval(sc.env, dir);
    float weight = 1.0f;
    if (!specular) {
        const Picker pk = picker_prepare(sc, last_p, last_n);
        const float light_pdf = picker_pmf(sc, pk, sc.num_lights, last_p, last_n) * env_pdf(sc.env, dir);
        weight = power_heuristic(1, last_pdf, lp.shadow_samples, light_pdf);
    }
    V3 c = thr * (le * weight);
    if (diffuse >= 2) c = clamp_indirect(c, lp.clamp_indirect);
    return finite3(c) ? c : splat(0.0f);
}

// kTransport = false: the scene has no Dielectric and no transmitting or subsurface material, so no path
// carries a wavelength, a medium, a pass-through count or a walk, and the ext column is neither read nor written.
// kGuide: the render has a guide; a guided vertex samples the mixture of its queue slot or the BSDF and weighs
// by the mixture density (path_tracer.cpp:700-735 with the MLP in place of the SD-tree), and records its
// sample for training.
// kPasses: the render credits the light passes, light groups and data passes (pb.pass_res).
template <bool kTransport, bool kGuide, bool kPasses>
__device__ __forceinline__ void shade_body(SceneView sc, LaunchParams lp, PathBuffers pb, int cur) {
    // Block b works on segment b % kQueueSegs of the current queue and appends to the same segment of the next
    // queue, the shadow queue and the next free list (whose segments therefore never receive more than the
    // segment's path ids).
    const SegShare sh = seg_share(blockIdx.x, gridDim.x, kQueueSegs);
    const uint32_t n = *seg_counter(pb.ctr, kCtrCur, sh.seg);
    const uint32_t seg_base = sh.seg * pb.seg_cap;
    uint32_t* next_count = seg_counter(pb.ctr, kCtrNext, sh.seg);
    uint32_t* shadow_count = seg_counter(pb.ctr, kCtrShadow, sh.seg);
    uint32_t* free_count = seg_counter(pb.ctr, kCtrFreeNext, sh.seg);
    uint32_t* fl_next = pb.fl[cur ^ 1] + seg_base;
    uint32_t collected = 0;
    const float4* ro = pb.ro[cur];
    const float4* rd = pb.rd[cur];
    float4* ro_next = pb.ro[cur ^ 1] + seg_base;
    float4* rd_next = pb.rd[cur ^ 1] + seg_base;
    const EnergyTables& T = *sc.tables;
    const int S = lp.shadow_samples;
    float4* sq_o = pb.sq_o + static_cast<size_t>(sh.seg) * pb.seg_cap * lp.shadow_slots;
    float4* sq_d = pb.sq_d + static_cast<size_t>(sh.seg) * pb.seg_cap * lp.shadow_slots;
    const uint32_t stride = sh.workers * blockDim.x;
    // Each block takes 256 queue slots per round and reorders them in LDS by a sort key: in scenes that mix
    // shading code paths, the hit material's path (GpuMaterial::sort_class: opaque Disney, Disney with lobes
    // below the surface, Dielectric, Lambert / Mirror), then the misses, so that a wave runs one path's code,
    // or the short environment path, rather than several with a share of its lanes each (a counting sort: one
    // ballot per key and wave, a prefix over keys and waves). Where every material takes the same path the key
    // is hit / miss only (measured on the product-sphere workload: sorting its 16 Disney materials by index
    // cost 2%).
    constexpr uint32_t kKeys = kMatBuckets + 1, kNoKey = 0xFFFFFFFFu;
    __shared__ uint32_t s_slot[kBlock];
    __shared__ uint32_t s_count[kKeys][kBlock / 64];
    // Staging (kShadeStage): each thread loads the entry of its own, unsorted slot (the slot's ray and hit,
    // coalesced, then the path's state by path id: throughput, radiance, ids, last point and normal, and the
    // first held-back NEE record, or the ext column in the transport and guide variants) into LDS before the
    // sort; after the sort a thread reads the entry it shades from LDS where it uses it. The loads are issued
    // together with the sort key's, so the sorted shading starts one memory round trip earlier. Two buffers
    // (rounds alternate), 9 float4 a slot: 72 KB a block, two blocks a CU.
    // Only the plain variant stages: the others hold more state across the vertex and spill more with it
    // (measured on the lights + media studio: 3% slower staged, the product sphere 2.5% faster).
    constexpr bool kStage = kShadeStage && !kTransport && !kGuide && !kPasses;
    constexpr int kStageVecs = kStage ? 9 : 1;
    constexpr bool kStagePend = !(kTransport || kGuide);  // vector 8 holds pend[pid * S], else ext
    __shared__ float4 s_stage[2][kStageVecs][kBlock];
    uint32_t round = 0;
    const uint32_t wave = threadIdx.x >> 6;
    const uint64_t below = (uint64_t{1} << __lane_id()) - 1u;
    // The loop bound is uniform over the block, so every lane takes part in the barriers and queue appends.
    const bool sort_hits = kMatBuckets > 1 && sc.sort_classes > 1;
    const uint32_t keys = sort_hits ? kKeys : 2u;  // hit classes, then the miss key
    for (uint32_t base = sh.rank * blockDim.x; base < n; base += stride, ++round) {
        const uint32_t mine = seg_base + base + threadIdx.x;
        float4 (*stage)[kBlock] = s_stage[round & 1u];
        uint32_t key = kNoKey;
        if (base + threadIdx.x < n) {
            const float4 h_mine = pb.hits[mine];
            const int tri = __float_as_int(h_mine.y);
            if constexpr (kStage) {
                const float4 o4 = ro[mine], d4 = rd[mine];
                stage[0][threadIdx.x] = o4;
                stage[1][threadIdx.x] = d4;
                stage[2][threadIdx.x] = h_mine;
                if (!(__float_as_uint(d4.w) & kCamEntry)) {
                    const uint32_t p = __float_as_uint(o4.w);
                    stage[3][threadIdx.x] = pb.thr[p];
                    stage[4][threadIdx.x] = pb.L[p];
                    const uint4 id4 = pb.ids[p];
                    stage[5][threadIdx.x] = make_float4(__uint_as_float(id4.x), __uint_as_float(id4.y),
                                                        __uint_as_float(id4.z), __uint_as_float(id4.w));
                    stage[6][threadIdx.x] = pb.last_p[p];
                    stage[7][threadIdx.x] = pb.last_n[p];
                    if constexpr (kStagePend) {
                        stage[8][threadIdx.x] = pb.pend[static_cast<size_t>(p) * S];
                    } else {
                        const uint4 e4 = pb.ext[p];
                        stage[8][threadIdx.x] = make_float4(__uint_as_float(e4.x), __uint_as_float(e4.y),
                                                            __uint_as_float(e4.z), __uint_as_float(e4.w));
                    }
                }
            }
            if (tri >= 0)
                key = sort_hits ? static_cast<uint32_t>(sc.materials[sc.shade[tri].material].sort_class) : 0u;
            else
                key = keys - 1u;
        }
        uint32_t rank = 0;
        for (uint32_t b = 0; b < keys; ++b) {
            const uint64_t m = __ballot(key == b);
            if (key == b) rank = static_cast<uint32_t>(__popcll(m & below));
            if (__lane_id() == 0) s_count[b][wave] = static_cast<uint32_t>(__popcll(m));
        }
        __syncthreads();
        uint32_t pos = 0, total = 0;
        for (uint32_t b = 0; b < keys; ++b)
            for (uint32_t k = 0; k < kBlock / 64; ++k) {
                const uint32_t c = s_count[b][k];
                total += c;
                if (b < key || (b == key && k < wave)) pos += c;
            }
        if (key != kNoKey) s_slot[pos + rank] = kStage ? threadIdx.x : mine;
        __syncthreads();
        const bool active = threadIdx.x < total;
        const uint32_t src = active ? s_slot[threadIdx.x] : 0u;  // staging: the thread that loaded the entry
        const uint32_t i = kStage ? seg_base + base + src : src;
        __syncthreads();  // s_slot and s_count are rewritten by the next round (the staging two rounds on)
        bool extend = false;
        bool ended = false;     // the path ended at this entry
        bool deferred = false;  // ... with a shadow ray of its last vertex still to be traced
        uint32_t pid = 0, new_depth_tag = 0;
        V3 new_o, new_d;
        if (active) {
            const float4 o4 = kStage ? stage[0][src] : ro[i];
            const float4 d4 = kStage ? stage[1][src] : rd[i];
            pid = __float_as_uint(o4.w);
            const V3 dir = mk(d4.x, d4.y, d4.z);
            // A camera entry starts its path here (k_regen made its ray); its state is the camera's, not read.
            const bool cam_entry = (__float_as_uint(d4.w) & kCamEntry) != 0u;
            V3 thr = mk(1.0f, 1.0f, 1.0f), L = mk(0.0f, 0.0f, 0.0f);
            float last_pdf = 0.0f;
            uint4 ids;
            uint4 ext = make_uint4(0u, 0u, 0u, 0u);
            float4 lp4 = make_float4(o4.x, o4.y, o4.z, 0.0f);  // lastPoint = the camera ray's origin (path_tracer.cpp:330)
            float4 ln4 = make_float4(0.0f, 1.0f, 0.0f, 0.0f);
            uint32_t media_events = 0;  // scattering events in media so far (path_tracer.cpp:51, :463), in L.w
            if (cam_entry) {
                ids = camera_ids(lp, __float_as_uint(d4.w) & ~kCamEntry);
            } else {
                const float4 t4 = kStage ? stage[3][src] : pb.thr[pid];
                thr = mk(t4.x, t4.y, t4.z);
                last_pdf = t4.w;
                const float4 L4 = kStage ? stage[4][src] : pb.L[pid];
                L = mk(L4.x, L4.y, L4.z);
                if constexpr (kTransport) media_events = __float_as_uint(L4.w);
                if constexpr (kStage) {
                    const float4 id4 = stage[5][src];
                    ids = make_uint4(__float_as_uint(id4.x), __float_as_uint(id4.y), __float_as_uint(id4.z),
                                     __float_as_uint(id4.w));
                } else {
                    ids = pb.ids[pid];
                }
                if constexpr (kTransport || kGuide) {
                    if constexpr (kStage) {
                        const float4 e4 = stage[8][src];
                        ext = make_uint4(__float_as_uint(e4.x), __float_as_uint(e4.y), __float_as_uint(e4.z),
                                         __float_as_uint(e4.w));
                    } else {
                        ext = pb.ext[pid];
                    }
                }
                lp4 = kStage ? stage[6][src] : pb.last_p[pid];
                ln4 = kStage ? stage[7][src] : pb.last_n[pid];
                // The first vertex's contact AO (path_tracer.cpp:617-621) is known now that its rays are traced:
                // it scales that vertex's held-back NEE and the path's throughput.
                float ao = 1.0f;
                if (kTransport && (ext.w & kExtAo)) {
                    ao = ao_factor(lp, pb, pid, S);
                    thr = thr * ao;
                    ext.w &= ~kExtAo;
                }
                // Held-back NEE of the previous vertex enters the radiance before anything this vertex adds.
                for (int s = 0; s < S; ++s) {
                    const size_t k = static_cast<size_t>(pid) * S + s;
                    const float4 p = kStage && kStagePend && s == 0 ? stage[8][src] : pb.pend[k];
                    const V3 c = kTransport ? mk(p.x, p.y, p.z) * ao : mk(p.x, p.y, p.z);
                    L += c;
                    if (kPasses) credit_nee(pb, lp, pid, ids.x, k, c);
                    pb.pend[k] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                }
            }
            const uint32_t g = ids.x;
            bool state_moved = cam_entry;  // last_p / last_n must be written if the path goes on
            float lambda = __uint_as_float(ext.x);
            SobolState ss{ids.z, rev_bits(ids.y), get_dim(ids.w)};
            uint32_t depth = get_depth(ids.w), diffuse = get_diffuse(ids.w);
            bool specular = get_specular(ids.w), camera = get_camera(ids.w);
            bool alive = true;
            bool issued = false;  // a shadow ray of this vertex was appended
            if (kGuide && (ext.w & kExtRecOpen)) {  // the radiance baseline of the previous vertex's record
                const uint32_t r = ((ext.w & kExtRecMask) >> kExtRecShift) - 1u;
                pb.grec[static_cast<size_t>(pid) * 8 + 4 * r + 3] = make_float4(L.x, L.y, L.z, 1.0f);
                ext.w &= ~kExtRecOpen;
            }
            const float4 h = kStage ? stage[2][src] : pb.hits[i];
            const int tri = __float_as_int(h.y);
            const V3 last_p = mk(lp4.x, lp4.y, lp4.z), last_n = mk(ln4.x, ln4.y, ln4.z);
            Surface si;
            V3 wo = -dir;
            bool vertex = false;        // a surface vertex is shaded below
            bool exit_vertex = false;   // ... the walk's exit, a white Lambert facing out (path_tracer.cpp:758-763)
            if (kTransport && (ext.w & kExtWalk)) {
                // One step of the subsurface random walk (randomWalkSubsurface, path_tracer.cpp:263-292).
                const GpuMaterial& wm = sc.materials[__float_as_uint(ln4.w)];
                const float tmax = tri >= 0 ? h.x : kFloatMax;
                float t = 0.0f;
                const bool scattered = free_flight(wm.sss_sigma_t, wm.sss_albedo * wm.sss_sigma_t, tmax, tri >= 0,
                                                   thr * wm.sss_albedo, ss, thr, t);
                const uint32_t step = ext_walk_step(ext.w);
                if (!finite3(thr) || is_black(thr)) {
                    alive = false;
                } else if (!scattered) {
                    if (tri < 0) {
                        alive = false;  // escaped through open geometry
                    } else {
                        fill_hit(sc, tri, dir, h.z, h.w, si);
                        si.normal = -si.normal;
                        wo = si.normal;
                        ext.w &= ~(kExtWalk | kExtWalkStepMask);
                        camera = false;
                        vertex = exit_vertex = true;
                    }
                } else {
                    float u0, u1;
                    sobol_get2d(ss, u0, u1);
                    new_d = sample_hg(dir, wm.sss_g, u0, u1);
                    new_o = mk(o4.x, o4.y, o4.z) + dir * t;
                    if (step >= 8) {
                        const float mx = fmaxf(thr.x, fmaxf(thr.y, thr.z));
                        if (mx < 1.0f) {
                            if (!(mx > 0.0f) || sobol_get1d(ss) >= mx) alive = false;
                            else thr = thr / mx;
                        }
                    }
                    if (alive && step + 1 < kMaxWalkSteps) {
                        ext.w = (ext.w & ~kExtWalkStepMask) | ((step + 1) << kExtWalkStepShift);
                        extend = true;
                    } else {
                        alive = false;
                    }
                }
            } else if (tri < 0) {
                // The camera ray's first segment misses: the AOVs of the background (renderer.cpp:196-204), the
                // Color background's colour or the environment.
                if (lp.want_aov && camera && pb.aov_res[2 * static_cast<size_t>(g) + 1].w < 0.0f) {
                    pb.aov_res[2 * static_cast<size_t>(g)] =
                        sc.bg_mode == PE_BACKGROUND_COLOR
                            ? clamped4(mk(sc.bg_color.x, sc.bg_color.y, sc.bg_color.z), 0.0f)
                            : (sc.has_env ? clamped4(env_eval(sc.env, dir), 0.0f) : make_float4(0.0f, 0.0f, 0.0f, 0.0f));
                    pb.aov_res[2 * static_cast<size_t>(g) + 1] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                }
                if (kPasses && camera) {  // no primary hit: the data passes stay zero (renderer.cpp:163)
                    float4* d1 = pb.pass_res + static_cast<size_t>(lp.pass_count - 1) * lp.num_paths + g;
                    if (d1->w < 0.0f) *d1 = make_float4(__int_as_float(-1), __int_as_float(-1), 0.0f, 0.0f);
                }
                // Miss. Inside a scattering medium the open segment is a free flight first (path_tracer.cpp:
                // 457-470); inside a medium that only absorbs the open path ends (:476-477); otherwise the
                // environment with MIS against NEE (:482-514).
                const int top = kTransport ? ext_media_top(ext) : -1;
                bool scattered = false;
                if (kTransport && top >= 0 && (sc.materials[top].flags & kMatFlagScatter))
                    scattered = medium_scatter(sc.materials[top], mk(o4.x, o4.y, o4.z), dir, kFloatMax, false, ss, thr,
                                               media_events, alive, new_o, new_d);
                const bool absorbed = top >= 0 && !(sc.materials[top].flags & kMatFlagScatter) &&
                                      (sc.materials[top].flags & kMatFlagMedium);
                if (scattered || !alive) {
                    camera = false;
                    specular = true;
                    extend = scattered && alive;
                } else if (absorbed) {
                } else if (depth == 0 && sc.bg_mode != PE_BACKGROUND_ENVIRONMENT) {
                    // A camera path leaving before its first vertex (past thin walls and hidden lights) sees the
                    // background (path_tracer.cpp:487-495): the Color one credited to the Background pass and
                    // the environment's group; the Transparent one nothing, and the thin walls it passed set the
                    // sample's coverage.
                    if (sc.bg_mode == PE_BACKGROUND_COLOR) {
                        const V3 c = thr * mk(sc.bg_color.x, sc.bg_color.y, sc.bg_color.z);
                        const V3 added = finite3(c) ? c : splat(0.0f);
                        L += added;
                        if (kPasses) credit_found(pb, lp, pid, g, depth, kPassBackground, sc.env_group, added);
                    } else if (kTransport && lp.want_aov && ext_pass_throughs(ext.w) > 0) {
                        float4* a = pb.aov_res + 2 * static_cast<size_t>(g);
                        a->w = wall_coverage(thr);
                    }
                } else if (sc.has_env) {
                    const V3 added = miss_radiance(sc, lp, dir, thr, specular, last_pdf, last_p, last_n, diffuse);
                    L += added;
                    if (kPasses) credit_found(pb, lp, pid, g, depth, kPassBackground, sc.env_group, added);
                }
                if (!extend) alive = false;
            } else {
                fill_hit(sc, tri, dir, h.z, h.w, si);
                // Camera rays pass through lights hidden from the camera, and any ray through a cutout surface
                // with probability 1 - opacity, one sampler dimension drawn when the opacity is strictly between 0
                // and 1 (passesCutout, path_tracer.cpp:159-167, :423-441): the ray continues from the far side,
                // and the continuation is not a new vertex.
                bool pass = camera && (si.flags & kTriCameraInvisible);
                if (kTransport && !pass && (sc.materials[si.material].flags & kMatFlagCutout)) {
                    const float op = sc.materials[si.material].opacity;
                    pass = op <= 0.0f || sobol_get1d(ss) >= op;
                }
                if (pass) {
                    new_o = offset_ray_origin(si.point, si.p_error, si.ng, dir);
                    new_d = dir;
                    extend = true;
                } else {
                    // The camera ray's first surface: its AOVs (renderer.cpp:183-194, primaryAlbedo :47-58). A camera
                    // path stays one through delta refractions, but only its first segment is the primary hit
                    // (path_tracer.cpp:441-444).
                    if (lp.want_aov && camera && pb.aov_res[2 * static_cast<size_t>(g) + 1].w < 0.0f) {
                        pb.aov_res[2 * static_cast<size_t>(g)] = clamped4(sc.materials[si.material].base_color, 1.0f);
                        pb.aov_res[2 * static_cast<size_t>(g) + 1] = make_float4(si.normal.x, si.normal.y, si.normal.z, 0.0f);
                    }
                    if (kPasses && camera) {  // the data passes of the primary hit (renderer.cpp:163-167)
                        float4* d1 = pb.pass_res + static_cast<size_t>(lp.pass_count - 1) * lp.num_paths + g;
                        if (d1->w < 0.0f) {
                            V3 co, cd;  // the camera ray's origin (a lens point for a thin lens)
                            camera_sample_ray(lp, g, co, cd);
                            pb.pass_res[static_cast<size_t>(lp.pass_count - 2) * lp.num_paths + g] =
                                make_float4(si.point.x, si.point.y, si.point.z, length(si.point - co));
                            *d1 = make_float4(__int_as_float(sc.shade[tri].shape), __int_as_float(si.material), 0.0f, 0.0f);
                        }
                    }
                    // Inside the current medium (path_tracer.cpp:457-479): a free flight in a scattering one, which
                    // may scatter before the surface, or Beer-Lambert absorption along the segment.
                    const int top = kTransport ? ext_media_top(ext) : -1;
                    bool scattered = false;
                    if (kTransport && top >= 0 && (sc.materials[top].flags & kMatFlagScatter)) {
                        const V3 o = mk(o4.x, o4.y, o4.z);
                        scattered = medium_scatter(sc.materials[top], o, dir, length(si.point - o), true, ss, thr,
                                                   media_events, alive, new_o, new_d);
                        if (scattered || !alive) {
                            camera = false;
                            specular = true;
                            extend = scattered && alive;
                        }
                    } else if (top >= 0 && (sc.materials[top].flags & kMatFlagMedium)) {
                        const V3 sa = sc.materials[top].sigma_a;
                        const float seg = length(si.point - mk(o4.x, o4.y, o4.z));
                        thr *= mk(expf(-sa.x * seg), expf(-sa.y * seg), expf(-sa.z * seg));
                        if (is_black(thr) || !finite3(thr)) alive = false;
                    }
                    if (alive && !scattered) {
                        const GpuMaterial& m = sc.materials[si.material];
                        // path_tracer.cpp:565-593; with light links a bounce from a receiver the light is not linked
                        // to does not see it (:571-572); the previous receiver's shape is in last_p.w.
                        if ((m.flags & kMatFlagEmissive) &&
                            !(kTransport && sc.has_links && depth > 0 && si.emitter >= 0 && si.emitter < sc.num_lights &&
                              !light_linked(sc, si.emitter, sc.node_uids[__float_as_uint(lp4.w)], false))) {
                            const int li = si.emitter;
                            float weight = 1.0f;
                            if (!specular) {
                                float light_pdf = 0.0f;
                                if (li >= 0) {
                                    const Picker pk = picker_prepare(sc, last_p, last_n);
                                    light_pdf = picker_pmf(sc, pk, li, last_p, last_n) *
                                                light_pdf_hit(sc.lights[li], last_p, si.point);
                                }
                                if (light_pdf > 0.0f && last_pdf > 0.0f) weight = power_heuristic(1, last_pdf, S, light_pdf);
                            }
                            V3 c = thr * (m.emission * weight);
                            if (diffuse >= 2) c = clamp_indirect(c, lp.clamp_indirect);
                            if (finite3(c)) L += c;
                            if (kPasses && finite3(c))
                                credit_found(pb, lp, pid, g, depth, kPassEmission, li >= 0 ? sc.lights[li].group : 0, c);
                        }
                        vertex = true;
                    }
                }
            }
            if (vertex) {
                const int mat = exit_vertex ? sc.exit_material : si.material;
                const GpuMaterial& m = sc.materials[mat];
                if (static_cast<int>(depth) >= lp.max_depth) {
                    alive = false;
                } else {
                    // A point on a metallic flake shades with the flake lobe (DisneyMaterial::resolve, disney.cpp:601).
                    Flake flake;
                    const Flake* fl = kTransport && (m.flags & kMatFlagFlakes) &&
                                              resolve_flake(m, si.point, si.normal, flake) ? &flake : nullptr;
                    // Dispersion: the first wavelength-dependent surface picks the path's wavelength (:602-607).
                    if (kTransport && lambda == 0.0f && (m.flags & kMatFlagDispersive)) {
                        V3 w;
                        lambda = sample_wavelength(T, sobol_get1d(ss), w);
                        thr *= w;
                        ext.x = __float_as_uint(lambda);
                    }
                    // Contact AO at the first vertex (approxAo, path_tracer.cpp:27-44, :617-621): S cosine rays of
                    // 0.35 about the shading normal go out as AO entries of the shadow queue; their factor is applied
                    // when the path's next vertex (or k_regen) folds this vertex's NEE.
                    if (kTransport && lp.ao_strength > 0.0f && depth == 0) {
                        const Frame af = frame_from_normal(si.normal);
                        for (int a = 0; a < S; ++a) {
                            float a0, a1;
                            sobol_get2d(ss, a0, a1);
                            const V3 ad = normalized(to_world(af, cosine_sample_hemisphere(a0, a1)));
                            const V3 ao_o = spawn_origin(si, ad);
                            const uint32_t rec = lp.ao_base + pid * static_cast<uint32_t>(S) + static_cast<uint32_t>(a);
                            pb.pend[rec] = make_float4(1.0f, 0.0f, 0.0f, 0.0f);
                            const uint32_t ao_slot = wave_append(shadow_count, true);
                            sq_o[ao_slot] = make_float4(ao_o.x, ao_o.y, ao_o.z, 0.35f);
                            sq_d[ao_slot] = make_float4(ad.x, ad.y, ad.z, __uint_as_float(kAoTag | rec));
                        }
                        ext.w |= kExtAo;
                    }
                    const bool is_delta = (m.flags & kMatFlagDelta) != 0;
                    // Guided: the slot's mixture was inferred (vertices below depth guided_bounces).
                    // Guided vertices as the reference picks them (path_tracer.cpp:624-627): not delta, not a
                    // volume boundary (whose refraction moves the medium stack), not a surface that transmits
                    // shadows (a cutout; thin walls are delta), not a walk's exit. With one guided
                    // bounce only camera entries were inferred: a depth-0 continuation (past a camera-invisible
                    // light or a thin wall) samples the BSDF alone.
                    const bool guided = kGuide && static_cast<int>(depth) < lp.guided_bounces && lp.guide_fraction > 0.0f &&
                                        !is_delta && !(m.flags & (kMatFlagVolume | kMatFlagCutout)) && !exit_vertex &&
                                        (lp.guided_bounces > 1 || cam_entry);
                    const float alpha = lp.guide_fraction;
                    // The slot's mixture is read where it is used (L1/L2), not held in registers across the NEE loop.
                    const float4* gmix = pb.mix + static_cast<size_t>(i) * kGuideMixFloat4;
                    if (sc.light_count > 0 && !is_delta) {  // NEE, path_tracer.cpp:631-698
                        const Picker pk = picker_prepare(sc, si.point, si.normal);
                        // Each NEE sample's shadow ray is appended as soon as it is built (the append is
                        // wave-aggregated over the lanes that reach it), so no shadow record stays live
                        // across the loop.
                        for (int s = 0; s < S; ++s) {
                            float pmf = 0.0f;
                            const float u = sobol_get1d(ss);
                            const int li = picker_sample(sc, pk, u, si.point, si.normal, pmf);
                            float ux, uy;
                            sobol_get2d(ss, ux, uy);
                            bool ok = false;
                            V3 so, sd;
                            float st = 0.0f;
                            do {
                                if (li < 0 || !(pmf > 0.0f)) break;
                                const bool from_env = li >= sc.num_lights;
                                if (kTransport && sc.has_links && !from_env &&
                                    !light_linked(sc, li, receiver_uid(sc, tri), false))
                                    break;  // an unlinked receiver: the sample is drawn, the light skipped (:641-644)
                                const LightSample ls = from_env ? env_sample(sc.env, ux, uy)
                                                                : light_sample(sc.lights[li], si.point, ux, uy);
                                if (!(ls.pdf > 0.0f) || is_black(ls.li)) break;
                                if (!(ls.distance < kFloatMax) || ls.distance >= 1e29f) {
                                    so = shadow_origin(si, ls.wi);
                                    sd = ls.wi;
                                    st = 1e30f;
                                } else if (!spawn_ray_to(si, si.point + ls.wi * ls.distance, so, sd, st)) {
                                    break;
                                }
                                float bsdf_pdf = 0.0f;
                                const V3 f = bsdf_eval_pdf<kTransport>(m, T, si, wo, ls.wi, lambda, bsdf_pdf, fl);
                                if (guided) bsdf_pdf = (1.0f - alpha) * bsdf_pdf + alpha * guide_pdf(gmix, ls.wi);
                                const float cos_t = fabsf(dot(ls.wi, si.normal));
                                if (is_black(f) || !(cos_t > 0.0f)) break;
                                const float light_pdf = pmf * ls.pdf;
                                const float weight = power_heuristic(S, light_pdf, 1, bsdf_pdf);
                                const V3 nee = f * ls.li * (cos_t * weight / (light_pdf * static_cast<float>(S)));
                                V3 c = thr * nee;
                                uint32_t pend_bits = 0;
                                if (diffuse >= 1) {
                                    // With thin walls the clamp waits for the shadow ray's transmittance.
                                    if (kTransport && sc.has_transmitters) pend_bits = kPendClamp;
                                    else c = clamp_indirect(c, lp.clamp_indirect);
                                }
                                if (kTransport && sc.has_shadow_links && !from_env)
                                    pend_bits |= static_cast<uint32_t>(li + 1) << 8;  // for the shadow links
                                if (!finite3(c)) break;
                                pb.pend[static_cast<size_t>(pid) * S + s] =
                                    make_float4(c.x, c.y, c.z, __uint_as_float(pend_bits));
                                if (kPasses) {
                                    const int group = from_env ? sc.env_group : sc.lights[li].group;
                                    float4* fr = pb.pfrac + 3 * (static_cast<size_t>(pid) * S + s);
                                    if (depth == 0) {  // the first surface: this sample's lobe shares (:682-690)
                                        V3 sh[3];
                                        lobe_shares(bsdf_eval_split<kTransport>(m, T, si, wo, ls.wi, lambda, fl), sh);
                                        fr[0] = make_float4(sh[0].x, sh[0].y, sh[0].z, __uint_as_float(nee_route(true, group)));
                                        fr[1] = make_float4(sh[1].x, sh[1].y, sh[1].z, 0.0f);
                                        fr[2] = make_float4(sh[2].x, sh[2].y, sh[2].z, 0.0f);
                                    } else {
                                        fr[0] = make_float4(0.0f, 0.0f, 0.0f, __uint_as_float(nee_route(false, group)));
                                    }
                                }
                                ok = true;
                            } while (false);
                            const uint32_t ss_slot = wave_append(shadow_count, ok);
                            issued = issued || ok;
                            if (ok) {
                                sq_o[ss_slot] = make_float4(so.x, so.y, so.z, st);
                                sq_d[ss_slot] = make_float4(sd.x, sd.y, sd.z,
                                                               __uint_as_float(pid * static_cast<uint32_t>(S) +
                                                                               static_cast<uint32_t>(s)));
                            }
                        }
                    }
                    // BSDF sample (path_tracer.cpp:736-743), then the walk entry (:746-764), the thin-wall
                    // pass-through (:771-786) or the bounce with throughput and roulette (:814-870).
                    const float uc = sobol_get1d(ss);
                    float u0, u1;
                    sobol_get2d(ss, u0, u1);
                    BsdfSample bs;
                    bool sampled;
                    if (guided && uc < alpha) {
                        // From the mixture: the BSDF value and the mixture density of the direction.
                        bs.wi = guide_sample(gmix, uc / alpha, u0, u1);
                        float p_bsdf = 0.0f;
                        bs.f = bsdf_eval_pdf<kTransport>(m, T, si, wo, bs.wi, lambda, p_bsdf, fl);
                        bs.pdf = (1.0f - alpha) * p_bsdf + alpha * guide_pdf(gmix, bs.wi);
                        bs.delta = bs.pass_through = bs.subsurface = false;
                        sampled = bs.pdf > 0.0f;
                    } else {
                        const float ub = guided ? (uc - alpha) / (1.0f - alpha) : uc;
                        sampled = bsdf_sample<kTransport>(m, T, si, wo, ub, u0, u1, lambda, bs, fl);
                        // The BSDF branch is taken with probability 1 - alpha (path_tracer.cpp:721-730): a
                        // subsurface entry, which the mixture cannot produce, carries its weight / (1 - alpha);
                        // delta and pass-through samples their pdf x (1 - alpha); the rest the mixture density.
                        if (guided && sampled) {
                            if (bs.subsurface) bs.weight = bs.weight * (1.0f / (1.0f - alpha));
                            else if (bs.delta || bs.pass_through) bs.pdf = (1.0f - alpha) * bs.pdf;
                            else bs.pdf = (1.0f - alpha) * bs.pdf + alpha * guide_pdf(gmix, bs.wi);
                        }
                    }
                    if (!sampled) {
                        alive = false;
                    } else if (kTransport && bs.subsurface) {
                        thr *= bs.weight;
                        new_d = normalized(bs.wi);
                        new_o = spawn_origin(si, new_d);
                        ext.w = (ext.w & ~kExtWalkStepMask) | kExtWalk;
                        ln4.w = __uint_as_float(static_cast<uint32_t>(mat));
                        state_moved = true;
                        extend = true;
                    } else if (bs.pdf <= 0.0f || is_black(bs.f)) {
                        alive = false;
                    } else if (kTransport && bs.pass_through) {
                        thr *= bs.f * (fabsf(dot(bs.wi, si.normal)) / bs.pdf);
                        const uint32_t passes = ext_pass_throughs(ext.w) + 1;
                        if (!finite3(thr) || passes > 64) {
                            alive = false;
                        } else {
                            ext.w = (ext.w & ~kExtPassMask) | (passes << kExtPassShift);
                            new_o = spawn_origin(si, bs.wi);
                            new_d = bs.wi;
                            extend = true;
                        }
                    } else {
                        if (kPasses && depth == 0) {  // the first surface's lobe shares of the sampled direction (:789-812)
                            V3 sh[3];
                            if (bs.delta) {
                                const bool refracted = dot(wo, si.ng) * dot(bs.wi, si.ng) < 0.0f;
                                sh[0] = splat(0.0f);
                                sh[1] = splat(refracted ? 0.0f : 1.0f);
                                sh[2] = splat(refracted ? 1.0f : 0.0f);
                            } else {
                                lobe_shares(bsdf_eval_split<kTransport>(m, T, si, wo, bs.wi, lambda, fl), sh);
                            }
                            for (int k = 0; k < 3; ++k)
                                pb.lobe[3 * static_cast<size_t>(pid) + k] = make_float4(sh[k].x, sh[k].y, sh[k].z, 0.0f);
                        }
                        const float cos_t = fabsf(dot(bs.wi, si.normal));
                        const V3 thr_before = thr;
                        thr *= bs.f * (cos_t / bs.pdf);
                        const uint32_t nrec = (ext.w & kExtRecMask) >> kExtRecShift;
                        if (kGuide && lp.guide_records && !bs.delta && nrec < kGuideRecords && finite3(thr)) {
                            float4* rec = pb.grec + static_cast<size_t>(pid) * 8 + 4 * nrec;
                            rec[0] = make_float4(si.point.x, si.point.y, si.point.z, wo.x);
                            rec[1] = make_float4(wo.y, wo.z, bs.wi.x, bs.wi.y);
                            rec[2] = make_float4(bs.wi.z, thr_before.x, thr_before.y, thr_before.z);
                            ext.w = (ext.w & ~kExtRecMask) | ((nrec + 1u) << kExtRecShift) | kExtRecOpen;
                        }
                        if (!finite3(thr)) {
                            alive = false;
                        } else {
                            last_pdf = bs.pdf;
                            // last_p.w: the receiver's shape, the prevReceiverUid of light linking (:832)
                            lp4 = make_float4(si.point.x, si.point.y, si.point.z,
                                              __uint_as_float(tri >= 0 ? static_cast<uint32_t>(sc.shade[tri].shape) : 0u));
                            ln4 = make_float4(si.normal.x, si.normal.y, si.normal.z, 0.0f);
                            state_moved = true;
                            specular = bs.delta;
                            if (!bs.delta) ++diffuse;
                            camera = camera && bs.delta && dot(bs.wi, si.normal) < 0.0f;
                            if (kTransport && (m.flags & kMatFlagVolume)) {  // medium stack (path_tracer.cpp:840-848)
                                const float a = dot(wo, si.ng), b = dot(bs.wi, si.ng);
                                if (a * b < 0.0f) {
                                    if (a > 0.0f) ext_media_enter(ext, mat);
                                    else ext_media_leave(ext, mat);
                                }
                            }
                            new_o = spawn_origin(si, bs.wi);
                            new_d = bs.wi;
                            ++depth;
                            extend = true;
                            if (static_cast<int>(depth) > lp.rr_depth) {
                                const float mx = fmaxf(thr.x, fmaxf(thr.y, thr.z));
                                if (mx < 1.0f) {
                                    if (!(mx > 0.0f) || sobol_get1d(ss) >= mx) {
                                        extend = false;
                                        alive = false;
                                    } else {
                                        thr = thr / mx;
                                    }
                                }
                            }
                        }
                    }
                }
            }
            if (!alive) extend = false;
            ended = !extend;
            // An ended path is finalised here unless a shadow ray of this vertex is still to be traced: then
            // its state stays for k_regen to finalise in the next iteration.
            deferred = ended && issued;
            if (extend || deferred) {
                pb.L[pid] = make_float4(L.x, L.y, L.z, kTransport ? __uint_as_float(media_events) : 0.0f);
                pb.ids[pid] = make_uint4(ids.x, ids.y, ids.z, pack_bits(ss.dim, depth, diffuse, specular, camera));
                if constexpr (kTransport || kGuide) pb.ext[pid] = ext;
            }
            if (extend) {
                pb.thr[pid] = make_float4(thr.x, thr.y, thr.z, last_pdf);
                if (state_moved) {
                    pb.last_p[pid] = lp4;
                    pb.last_n[pid] = ln4;
                }
            }
            if (ended && !deferred) {
                pb.res[g] = make_float4(L.x, L.y, L.z, 0.0f);
                if (kGuide && lp.guide_records) collected += guide_collect(lp, pb, pid, g, L, ext.w);
            }
            new_depth_tag = depth;
        }
        const uint32_t slot = wave_append(next_count, extend);
        if (extend) {
            ro_next[slot] = make_float4(new_o.x, new_o.y, new_o.z, __uint_as_float(pid));
            rd_next[slot] = make_float4(new_d.x, new_d.y, new_d.z, __uint_as_float(new_depth_tag));
        }
        const uint32_t fslot = wave_append(free_count, ended);
        if (ended) fl_next[fslot] = pid | (deferred ? kDeferred : 0u);
    }
    if (kGuide) wave_count(&pb.stats[6], collected);
}

// Two waves per SIMD for the guide and transport variants: they sit a few registers above the 256 that allow
// it, and the latency hiding of a second wave is worth far more than the few values spilled (measured: the
// guide variant 49 -> 30 ms per 1080p launch of 16 spp). The plain variant (248 registers) keeps the compiler's
// own schedule, which the occupancy request would change.
template <bool kTransport, bool kGuide>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(2)))
k_shade(SceneView sc, LaunchParams lp, PathBuffers pb, int cur) {
    shade_body<kTransport, kGuide, false>(sc, lp, pb, cur);
}

// Renders with pass films: the transport variant (it shades every scene) with the pass crediting.
template <bool kGuide>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(2)))
k_shade_passes(SceneView sc, LaunchParams lp, PathBuffers pb, int cur) {
    shade_body<true, kGuide, true>(sc, lp, pb, cur);
}

__global__ void __launch_bounds__(kBlock) PE_SHADE_PLAIN_ATTR k_shade_plain(SceneView sc, LaunchParams lp, PathBuffers pb, int cur) {
    shade_body<false, false, false>(sc, lp, pb, cur);
}

// After an iteration: the next queue becomes current, the next free list current, the samples given out are
// counted, and the new free ids' share of the segment's remaining samples is set. One wave, one lane per
// segment. ctr[kCtrTotal] = continuation rays + free ids + samples not yet handed out: zero when the window is
// done (its last free ids finalised).
__global__ void k_rotate(PathBuffers pb, uint32_t num_samples) {
    const uint32_t k = threadIdx.x;
    if (blockIdx.x != 0 || k >= kQueueSegs) return;
    uint32_t* base = seg_counter(pb.ctr, kCtrBase, k);
    uint32_t* cur = seg_counter(pb.ctr, kCtrCur, k);
    uint32_t* cam = seg_counter(pb.ctr, kCtrCam, k);
    uint32_t* next = seg_counter(pb.ctr, kCtrNext, k);
    uint32_t* shadow = seg_counter(pb.ctr, kCtrShadow, k);
    uint32_t* fr = seg_counter(pb.ctr, kCtrFree, k);
    uint32_t* fr_next = seg_counter(pb.ctr, kCtrFreeNext, k);
    uint32_t* gen = seg_counter(pb.ctr, kCtrGen, k);
    const uint32_t total = seg_samples(num_samples, k);
    unsigned long long traced = *cur, shadows = *shadow;
    const uint32_t given = *gen + *cam;
    const uint32_t b = *next, f = *fr_next;
    const uint32_t c = min(f, total - given);
    *gen = given;
    *seg_counter(pb.ctr, kCtrChunkIsect, k) = 0u;
    *seg_counter(pb.ctr, kCtrChunkShadow, k) = 0u;
    *base = b;
    *cam = c;
    *cur = b + c;
    *next = 0u;
    *shadow = 0u;
    *fr = f;
    *fr_next = 0u;
    unsigned long long live = static_cast<unsigned long long>(b) + f + (total - given);
    for (int off = 32; off > 0; off >>= 1) {
        traced += __shfl_xor(traced, off);
        shadows += __shfl_xor(shadows, off);
        live += __shfl_xor(live, off);
    }
    if (k == 0) {
        pb.stats[0] += traced;
        pb.stats[1] += shadows;
        pb.ctr[kCtrTotal] = static_cast<uint32_t>(min(live, 0xFFFFFFFFull));
    }
}

__global__ void k_zero_stats(PathBuffers pb) {
    if (blockIdx.x == 0 && threadIdx.x < kStatSlots) pb.stats[threadIdx.x] = 0;
}

// Film::addSample for every sample of every pixel, in sample order (film.cpp:56-71).
// Film::relativeError (film.cpp:100-109) without contraction: sqrt(var / n) / (mean + 0.01) of the luminance.
__device__ __forceinline__ float film_relative_error(float s0, float s1, float s2, uint32_t n, float l2) {
#pragma clang fp contract(off)
    if (n < 2) return 1e30f;
    const float fn = static_cast<float>(n);
    const float mean = (0.212671f * s0 + 0.715160f * s1 + 0.072169f * s2) / fn;
    const float mean_sq = l2 / fn;
    const float var = fmaxf(0.0f, mean_sq - mean * mean) * fn / static_cast<float>(n - 1);
    return sqrtf(var / fn) / (fmaxf(mean, 0.0f) + 0.01f);
}

// Luminance as Film::addSample accumulates it (spectrum.h:68-70), without contraction.
__device__ __forceinline__ float film_luminance(V3 c) {
#pragma clang fp contract(off)
    return 0.212671f * c.x + 0.715160f * c.y + 0.072169f * c.z;
}

// A device film's arrays (pe_film).
struct FilmView {
    float* sum;
    uint32_t* count;
    float* lum2;
    unsigned long long* rejected;
};

// Film::addSample (film.cpp:56-71) or, signed, Film::addSampleSigned (:73-86) of the samples of one pixel.
template <bool kSigned>
__device__ __forceinline__ void film_add_aov(const FilmView& f, uint32_t p, const float4* __restrict__ aov, int which,
                                             uint32_t q, uint32_t pixels, int spl, bool splat_w) {
#pragma clang fp contract(off)
    if (!f.sum) return;
    float sr = f.sum[3 * static_cast<size_t>(p)], sg = f.sum[3 * static_cast<size_t>(p) + 1],
          sb = f.sum[3 * static_cast<size_t>(p) + 2];
    float l2 = kSigned ? 0.0f : f.lum2[p];
    uint32_t rej = 0;
    for (int k = 0; k < spl; ++k) {
        const float4 a = aov[2 * static_cast<size_t>(q + static_cast<uint32_t>(k) * pixels) + which];
        const V3 v = splat_w ? splat(a.w) : mk(a.x, a.y, a.z);
        const bool ok = finite3(v) && (kSigned || (v.x >= 0.0f && v.y >= 0.0f && v.z >= 0.0f));
        if (!ok) {
            ++rej;
            continue;
        }
        sr += v.x;
        sg += v.y;
        sb += v.z;
        if (!kSigned) {
            const float lum = film_luminance(v);
            l2 += lum * lum;
        }
    }
    f.sum[3 * static_cast<size_t>(p)] = sr;
    f.sum[3 * static_cast<size_t>(p) + 1] = sg;
    f.sum[3 * static_cast<size_t>(p) + 2] = sb;
    f.count[p] += static_cast<uint32_t>(spl);
    if (!kSigned) f.lum2[p] = l2;
    if (rej) atomicAdd(f.rejected, static_cast<unsigned long long>(rej));
}

__global__ void __launch_bounds__(kBlock) k_film(LaunchParams lp, PathBuffers pb, float* __restrict__ sum,
                                                 uint32_t* __restrict__ count, float* __restrict__ lum2,
                                                 unsigned long long* __restrict__ rejected, FilmView f_albedo,
                                                 FilmView f_normal, FilmView f_alpha) {
#pragma clang fp contract(off)
    const uint32_t pixels = lp.num_pixels;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t q = blockIdx.x * blockDim.x + threadIdx.x; q < pixels; q += stride) {
        const uint32_t p = lp.pixel_list ? lp.pixel_list[q] : q;
        float sr = sum[3 * static_cast<size_t>(p)], sg = sum[3 * static_cast<size_t>(p) + 1], sb = sum[3 * static_cast<size_t>(p) + 2];
        uint32_t c = count[p];
        float l2 = lum2[p];
        uint32_t rej = 0;
        for (int k = 0; k < lp.spl; ++k) {
            const float4 L4 = pb.res[q + static_cast<uint32_t>(k) * pixels];
            const V3 L = mk(L4.x, L4.y, L4.z);
            ++c;
            const bool ok = finite3(L) && L.x >= 0.0f && L.y >= 0.0f && L.z >= 0.0f;
            if (!ok) {
                ++rej;
                continue;
            }
            sr += L.x;
            sg += L.y;
            sb += L.z;
            const float lum = film_luminance(L);
            l2 += lum * lum;
        }
        sum[3 * static_cast<size_t>(p)] = sr;
        sum[3 * static_cast<size_t>(p) + 1] = sg;
        sum[3 * static_cast<size_t>(p) + 2] = sb;
        count[p] = c;
        lum2[p] = l2;
        if (rej) atomicAdd(rejected, static_cast<unsigned long long>(rej));
        if (lp.want_aov) {  // addAovs (renderer.cpp:222-227)
            film_add_aov<false>(f_albedo, p, pb.aov_res, 0, q, pixels, lp.spl, false);
            film_add_aov<true>(f_normal, p, pb.aov_res, 1, q, pixels, lp.spl, false);
            film_add_aov<false>(f_alpha, p, pb.aov_res, 0, q, pixels, lp.spl, true);
        }
        // renderer.cpp:500-502: the pixel stops for the rest of the batch once its own error is low enough.
        if (lp.adapt_check && film_relative_error(sr, sg, sb, c, l2) < lp.noise_threshold) lp.active[p] = 0u;
    }
}

// idColor (renderer.cpp:252-273): an id's stable colour, splitmix64 to a hue stepped by the golden ratio, HSV
// with value 1 and saturation 0.75; id 0 is black.
__device__ __forceinline__ V3 id_color(uint64_t id) {
#pragma clang fp contract(off)
    if (id == 0) return splat(0.0f);
    uint64_t z = id + 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    z ^= z >> 31;
    const float h = fmodf(static_cast<float>(z >> 40) * 0x1p-24f + 0.618034f, 1.0f) * 6.0f;
    const int i = static_cast<int>(h) % 6;
    const float f = h - floorf(h);
    const float p0 = 0.25f, q = 1.0f - 0.75f * f, t = 1.0f - 0.75f * (1.0f - f);
    switch (i) {
        case 0: return mk(1.0f, t, p0);
        case 1: return mk(q, 1.0f, p0);
        case 2: return mk(p0, 1.0f, t);
        case 3: return mk(p0, q, 1.0f);
        case 4: return mk(t, p0, 1.0f);
        default: return mk(1.0f, p0, q);
    }
}

// Pass films of a render (pe_pass_targets): 8 light passes, 16 groups, depth, position, object and material id.
constexpr int kPassFilms = 28;
struct PassFilms {
    FilmView f[kPassFilms];
};

// addPasses (renderer.cpp:207-219) for every pass film and pixel: one thread per (film, pixel), adding the
// pixel's samples of this window in sample order. Light passes, groups, depth and the id colours with
// Film::addSample, the position with addSampleSigned.
__global__ void __launch_bounds__(kBlock) k_film_passes(LaunchParams lp, PathBuffers pb, PassFilms films,
                                                        const uint64_t* __restrict__ node_uids) {
#pragma clang fp contract(off)
    const uint32_t pixels = lp.num_pixels;
    const uint64_t total = static_cast<uint64_t>(pixels) * kPassFilms;
    const uint32_t data0 = lp.pass_count - 2, data1 = lp.pass_count - 1;
    for (uint64_t idx = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; idx < total;
         idx += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
        const int fi = static_cast<int>(idx / pixels);
        const uint32_t q = static_cast<uint32_t>(idx - static_cast<uint64_t>(fi) * pixels);
        const FilmView& f = films.f[fi];
        if (!f.sum) continue;
        if (fi >= static_cast<int>(kLightPasses) && fi < static_cast<int>(kLightPasses) + 16 &&
            fi - static_cast<int>(kLightPasses) >= lp.num_groups)
            continue;
        const bool is_signed = fi == 25;
        const uint32_t p = lp.pixel_list ? lp.pixel_list[q] : q;
        float sr = f.sum[3 * static_cast<size_t>(p)], sg = f.sum[3 * static_cast<size_t>(p) + 1],
              sb = f.sum[3 * static_cast<size_t>(p) + 2];
        float l2 = is_signed ? 0.0f : f.lum2[p];
        uint32_t rej = 0;
        for (int k = 0; k < lp.spl; ++k) {
            const uint32_t g = q + static_cast<uint32_t>(k) * pixels;
            V3 v;
            if (fi < 24) {
                const float4 a = pb.pass_res[static_cast<size_t>(fi) * lp.num_paths + g];
                v = mk(a.x, a.y, a.z);
            } else if (fi == 24 || fi == 25) {
                const float4 a = pb.pass_res[static_cast<size_t>(data0) * lp.num_paths + g];
                v = fi == 24 ? splat(a.w) : mk(a.x, a.y, a.z);
            } else {
                const float4 a = pb.pass_res[static_cast<size_t>(data1) * lp.num_paths + g];
                const int shape = __float_as_int(a.x), mat = __float_as_int(a.y);
                if (fi == 26) v = id_color(shape >= 0 ? node_uids[shape] : 0ull);
                else v = id_color(mat >= 0 ? static_cast<uint64_t>(mat) + 1u : 0ull);
            }
            const bool ok = finite3(v) && (is_signed || (v.x >= 0.0f && v.y >= 0.0f && v.z >= 0.0f));
            if (!ok) {
                ++rej;
                continue;
            }
            sr += v.x;
            sg += v.y;
            sb += v.z;
            if (!is_signed) {
                const float lum = film_luminance(v);
                l2 += lum * lum;
            }
        }
        f.sum[3 * static_cast<size_t>(p)] = sr;
        f.sum[3 * static_cast<size_t>(p) + 1] = sg;
        f.sum[3 * static_cast<size_t>(p) + 2] = sb;
        f.count[p] += static_cast<uint32_t>(lp.spl);
        if (!is_signed) f.lum2[p] = l2;
        if (rej) atomicAdd(f.rejected, static_cast<unsigned long long>(rej));
    }
}

// A device CryptoFilm's arrays (pe_crypto_film).
struct CryptoView {
    float* ids;
    uint32_t* counts;
    uint32_t* totals;
    unsigned long long* dropped;
};

// CryptoFilm::addSample (cryptomatte.cpp:121-133) of a pixel's samples in sample order: the total counts every
// sample; a non-zero id increments its entry or takes the first free slot. `which` selects the first surface's
// shape (0) or material (1) from the data record, mapped through `table` (n entries; outside it: no id).
__device__ __forceinline__ void crypto_add(const CryptoView& f, const LaunchParams& lp, const PathBuffers& pb,
                                           uint32_t q, uint32_t p, int which, const float* table, int n) {
    if (!f.ids) return;
    const uint32_t pixels = lp.num_pixels, data1 = lp.pass_count - 1;
    float* ids = f.ids + static_cast<size_t>(p) * PE_CRYPTO_SLOTS;
    uint32_t* counts = f.counts + static_cast<size_t>(p) * PE_CRYPTO_SLOTS;
    uint32_t dropped = 0;
    for (int k = 0; k < lp.spl; ++k) {
        const float4 a = pb.pass_res[static_cast<size_t>(data1) * lp.num_paths + q + static_cast<uint32_t>(k) * pixels];
        const int index = __float_as_int(which == 0 ? a.x : a.y);
        const float id = index >= 0 && index < n ? table[index] : 0.0f;
        if (id == 0.0f) continue;
        int slot = -1;
        for (int j = 0; j < PE_CRYPTO_SLOTS; ++j) {
            const uint32_t c = counts[j];
            if (c == 0u) {
                if (slot < 0) slot = j;
                break;
            }
            if (ids[j] == id) {
                slot = j;
                break;
            }
        }
        if (slot < 0) {
            ++dropped;
        } else {
            if (counts[slot] == 0u) ids[slot] = id;
            counts[slot] += 1u;
        }
    }
    f.totals[p] += static_cast<uint32_t>(lp.spl);
    if (dropped) atomicAdd(f.dropped, static_cast<unsigned long long>(dropped));
}

__global__ void __launch_bounds__(kBlock) k_film_crypto(LaunchParams lp, PathBuffers pb, CryptoView obj, CryptoView mat,
                                                        const float* __restrict__ obj_ids, int num_obj,
                                                        const float* __restrict__ mat_ids, int num_mat) {
    for (uint32_t q = blockIdx.x * blockDim.x + threadIdx.x; q < lp.num_pixels; q += gridDim.x * blockDim.x) {
        const uint32_t p = lp.pixel_list ? lp.pixel_list[q] : q;
        crypto_add(obj, lp, pb, q, p, 0, obj_ids, num_obj);
        crypto_add(mat, lp, pb, q, p, 1, mat_ids, num_mat);
    }
}

// ------------------------------------------------------------------------------ adaptive sampling
// The batch's skip mask (renderer.cpp:442-467): a pixel is clean when its error is below the threshold and
// is skipped for the whole batch when it and its eight neighbours (clamped at the border) are clean.
__global__ void __launch_bounds__(kBlock) k_adapt_clean(const float* __restrict__ sum, const uint32_t* __restrict__ count,
                                                        const float* __restrict__ lum2, uint32_t pixels, float threshold,
                                                        uint32_t* __restrict__ clean) {
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < pixels; p += gridDim.x * blockDim.x)
        clean[p] = film_relative_error(sum[3 * static_cast<size_t>(p)], sum[3 * static_cast<size_t>(p) + 1],
                                       sum[3 * static_cast<size_t>(p) + 2], count[p], lum2[p]) < threshold ? 1u : 0u;
}

__global__ void __launch_bounds__(kBlock) k_adapt_mask(const uint32_t* __restrict__ clean, int w, int h,
                                                       uint32_t* __restrict__ active) {
    const uint32_t pixels = static_cast<uint32_t>(w) * static_cast<uint32_t>(h);
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < pixels; p += gridDim.x * blockDim.x) {
        const int x = static_cast<int>(p % static_cast<uint32_t>(w)), y = static_cast<int>(p / static_cast<uint32_t>(w));
        bool all = true;
        for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx) {
                const int nx = min(max(x + dx, 0), w - 1), ny = min(max(y + dy, 0), h - 1);
                all = all && clean[static_cast<size_t>(ny) * w + nx] != 0u;
            }
        active[p] = all ? 0u : 1u;
    }
}

__global__ void __launch_bounds__(kBlock) k_fill_u32(uint32_t* __restrict__ p, uint32_t n, uint32_t v) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = v;
}

// Ordered compaction of the active pixels into a list: per-block counts (1024 pixels a block, four rounds of
// 256), one block scanning the counts, then each block writing its pixels at its offset in pixel order.
constexpr uint32_t kCompactSpan = 1024;

__device__ __forceinline__ uint32_t block_exclusive(uint32_t v, uint32_t* s_wave, uint32_t& total) {
    const uint32_t lane = __lane_id(), wave = threadIdx.x >> 6, waves = blockDim.x >> 6;
    uint32_t incl = v;
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t t = __shfl_up(incl, off);
        if (lane >= static_cast<uint32_t>(off)) incl += t;
    }
    if (lane == 63) s_wave[wave] = incl;
    __syncthreads();
    uint32_t before = 0;
    total = 0;
    for (uint32_t k = 0; k < waves; ++k) {
        if (k < wave) before += s_wave[k];
        total += s_wave[k];
    }
    __syncthreads();
    return before + incl - v;
}

__global__ void __launch_bounds__(kBlock) k_count_active(const uint32_t* __restrict__ active, uint32_t pixels,
                                                         uint32_t* __restrict__ counts) {
    __shared__ uint32_t s_wave[kBlock / 64];
    uint32_t mine = 0;
    for (uint32_t r = 0; r < kCompactSpan / kBlock; ++r) {
        const uint32_t p = blockIdx.x * kCompactSpan + r * kBlock + threadIdx.x;
        mine += p < pixels && active[p] ? 1u : 0u;
    }
    uint32_t total;
    (void)block_exclusive(mine, s_wave, total);
    if (threadIdx.x == 0) counts[blockIdx.x] = total;
}

// One block: exclusive offsets of nb counts in place; counts[nb] = the total.
__global__ void __launch_bounds__(kBlock) k_scan_counts(uint32_t* __restrict__ counts, uint32_t nb) {
    __shared__ uint32_t s_wave[kBlock / 64];
    uint32_t carry = 0;
    for (uint32_t base = 0; base < nb; base += kBlock) {
        const uint32_t i = base + threadIdx.x;
        const uint32_t v = i < nb ? counts[i] : 0u;
        uint32_t total;
        const uint32_t ex = block_exclusive(v, s_wave, total);
        if (i < nb) counts[i] = carry + ex;
        carry += total;
    }
    if (threadIdx.x == 0) counts[nb] = carry;
}

__global__ void __launch_bounds__(kBlock) k_scatter_active(const uint32_t* __restrict__ active, uint32_t pixels,
                                                           const uint32_t* __restrict__ offsets,
                                                           uint32_t* __restrict__ list) {
    __shared__ uint32_t s_wave[kBlock / 64];
    uint32_t base = offsets[blockIdx.x];
    for (uint32_t r = 0; r < kCompactSpan / kBlock; ++r) {
        const uint32_t p = blockIdx.x * kCompactSpan + r * kBlock + threadIdx.x;
        const uint32_t a = p < pixels && active[p] ? 1u : 0u;
        uint32_t total;
        const uint32_t ex = block_exclusive(a, s_wave, total);
        if (a) list[base + ex] = p;
        base += total;
    }
}

// Pixel-filter inverse CDF (pixel_filter.cpp:37-63), built on the host for the non-box filters.
std::vector<float> filter_cdf(int type, float width) {
    const float w = std::clamp(width, 0.25f, 8.0f);
    const float r = 0.5f * w;
    auto evaluate = [&](float x) -> float {
        if (std::abs(x) > r) return 0.0f;
        switch (type) {
            case PE_FILTER_TRIANGLE: return std::max(0.0f, 1.0f - std::abs(x) / r);
            case PE_FILTER_GAUSSIAN: {
                const float s = w / 6.0f;
                const float k = 1.0f / (2.0f * s * s);
                return std::max(0.0f, std::exp(-k * x * x) - std::exp(-k * r * r));
            }
            case PE_FILTER_BLACKMAN_HARRIS: {
                const float a = kTwoPi * (x / w + 0.5f);
                return 0.35875f - 0.48829f * std::cos(a) + 0.14128f * std::cos(2.0f * a) - 0.01168f * std::cos(3.0f * a);
            }
            default: return 1.0f;
        }
    };
    std::vector<float> cdf(513, 0.0f);
    const float dx = w / 512;
    double acc = 0.0;
    for (int i = 0; i < 512; ++i) {
        acc += std::max(0.0f, evaluate(-r + (static_cast<float>(i) + 0.5f) * dx));
        cdf[static_cast<size_t>(i) + 1] = static_cast<float>(acc);
    }
    if (!(acc > 0.0)) return {};
    for (float& c : cdf) c = static_cast<float>(c / acc);
    cdf.back() = 1.0f;
    return cdf;
}

}  // namespace

enum KernelKind { kKGenerate = 0, kKIntersect, kKShade, kKShadow, kKFilm, kKGuide, kKTrain, kKRegen, kKindCount };
const char* const kKernelNames[kKindCount] = {"generate", "intersect", "shade", "shadow", "film", "guide_infer",
                                              "guide_train", "regen"};

// HIP events around each wavefront launch on the device's stream while profiling is on.
struct Profiler {
    bool on = false;
    struct Rec {
        int kind;
        hipEvent_t a, b;
    };
    std::vector<Rec> recs;
    double ms[kKindCount] = {};
    uint64_t launches[kKindCount] = {};

    void begin(int kind, hipStream_t s) {
        if (!on) return;
        Rec r{kind, nullptr, nullptr};
        if (hipEventCreate(&r.a) != hipSuccess || hipEventCreate(&r.b) != hipSuccess) return;
        (void)hipEventRecord(r.a, s);
        recs.push_back(r);
    }
    void end(hipStream_t s) {
        if (!on || recs.empty()) return;
        (void)hipEventRecord(recs.back().b, s);
    }
    void drain() {
        for (Rec& r : recs) {
            float t = 0.0f;
            if (hipEventSynchronize(r.b) == hipSuccess && hipEventElapsedTime(&t, r.a, r.b) == hipSuccess) {
                ms[r.kind] += t;
                ++launches[r.kind];
            }
            (void)hipEventDestroy(r.a);
            (void)hipEventDestroy(r.b);
        }
        recs.clear();
    }
};

// Per-stream wavefront buffers, grown on demand and kept between renders: the path pool (state of the ids in
// flight, ray / shadow queues, free lists) and the window's result buffer.
struct Wavefront {
    Profiler prof;
    uint32_t capacity = 0;   // path ids: kQueueSegs segments of seg_cap
    uint32_t seg_cap = 0;    // ids (and queue slots) per segment
    int shadow_samples = 0;
    uint64_t result_cap = 0; // samples of the result buffer
    uint64_t aov_cap = 0;    // samples of the AOV result buffer
    DevBuf ro0, rd0, ro1, rd1, thr, L, last_p, last_n, ids, ext, pend, hits, sq_o, sq_d, fl0, fl1, ctr, stats, filter;
    DevBuf res, aov_res;
    // Render passes: per-sample records (pass_cap float4), per-id lobe shares and per-record NEE routing; the
    // cryptomatte id tables of the render (objects, then materials).
    DevBuf pass_res, lobe, pfrac, crypto_ids;
    uint64_t pass_cap = 0;
    uint32_t lobe_cap = 0;  // path ids the lobe / routing columns hold (at the current shadow sample count)
    // Adaptive sampling: per-pixel active flags and clean flags, the active pixel list, compaction counts.
    DevBuf adapt_active, adapt_clean, adapt_list, adapt_counts;
    uint32_t adapt_pixels = 0;
    int filter_type = -1;
    float filter_width = 0.0f;
    // Progress of the iterations in flight: the remaining-work counter copied after every chunk of iterations
    // into pinned host memory, one slot and one event per chunk (two chunks in flight).
    uint32_t* host_live = nullptr;
    hipEvent_t chunk_done[2] = {nullptr, nullptr};

    ~Wavefront() {
        prof.drain();
        for (hipEvent_t& e : chunk_done)
            if (e) (void)hipEventDestroy(e);
        if (host_live) (void)hipHostFree(host_live);
    }

    // Device bytes per path id in flight: two ray queues, hits, six state columns, two free lists, S held-back
    // contributions and S shadow-queue records.
    static uint64_t bytes_per_path(int S) { return 4 * 16 + 16 + 6 * 16 + 8 + static_cast<uint64_t>(S) * 3 * 16; }
    static constexpr uint64_t kResultBytes = 16, kAovResultBytes = 32;  // per window sample

    int ensure_adaptive(uint32_t pixels, uint32_t blocks) {
        if (pixels <= adapt_pixels) return PE_OK;
        if (!adapt_active.alloc(static_cast<size_t>(pixels) * 4) || !adapt_clean.alloc(static_cast<size_t>(pixels) * 4) ||
            !adapt_list.alloc(static_cast<size_t>(pixels) * 4) || !adapt_counts.alloc((static_cast<size_t>(blocks) + 1) * 4)) {
            adapt_pixels = 0;
            return PE_ERR_OOM;
        }
        adapt_pixels = pixels;
        return PE_OK;
    }

    uint32_t slots() const { return kQueueSegs * seg_cap; }

    // A pool of at least n path ids.
    int ensure(uint32_t n, int S) {
        if (!host_live) {
            if (hipHostMalloc(reinterpret_cast<void**>(&host_live), 2 * sizeof(uint32_t), hipHostMallocDefault) != hipSuccess) {
                host_live = nullptr;
                set_error("pinned host allocation failed");
                return PE_ERR_HIP;
            }
            for (hipEvent_t& e : chunk_done)
                if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
                    e = nullptr;
                    set_error("event creation failed");
                    return PE_ERR_HIP;
                }
        }
        const uint32_t cap = seg_capacity(n);
        if (cap <= seg_cap && S == shadow_samples) return PE_OK;
        const size_t Q = static_cast<size_t>(kQueueSegs) * cap, QS = Q * S;
        if (!ro0.alloc(Q * 16) || !rd0.alloc(Q * 16) || !ro1.alloc(Q * 16) || !rd1.alloc(Q * 16) || !thr.alloc(Q * 16) ||
            !L.alloc(Q * 16) || !last_p.alloc(Q * 16) || !last_n.alloc(Q * 16) || !ids.alloc(Q * 16) || !ext.alloc(Q * 16) ||
            !pend.alloc(QS * 16) || !hits.alloc(Q * 16) || !sq_o.alloc(QS * 16) || !sq_d.alloc(QS * 16) ||
            !fl0.alloc(Q * 4) || !fl1.alloc(Q * 4) || !ctr.alloc(kCtrBytes) || !stats.alloc(kStatSlots * 8)) {
            capacity = 0;
            seg_cap = 0;
            return PE_ERR_OOM;
        }
        capacity = static_cast<uint32_t>(Q);
        seg_cap = cap;
        shadow_samples = S;
        lobe_cap = 0;  // the pass columns follow the pool
        return PE_OK;
    }

    // A result buffer of at least n samples (and AOV results when asked for).
    int ensure_results(uint64_t n, bool aov) {
        if (n > result_cap) {
            if (!res.alloc(n * kResultBytes)) {
                result_cap = 0;
                return PE_ERR_OOM;
            }
            result_cap = n;
        }
        if (aov && n > aov_cap) {
            if (!aov_res.alloc(n * kAovResultBytes)) {
                aov_cap = 0;
                return PE_ERR_OOM;
            }
            aov_cap = n;
        }
        return PE_OK;
    }

    // Pass buffers for a window of n samples of `records` float4 each, over the current pool.
    int ensure_passes(uint64_t n, uint32_t records) {
        const uint64_t need = n * records;
        if (need > pass_cap) {
            if (!pass_res.alloc(need * 16)) {
                pass_cap = 0;
                return PE_ERR_OOM;
            }
            pass_cap = need;
        }
        if (lobe_cap < capacity) {
            if (!lobe.alloc(static_cast<size_t>(capacity) * 48) ||
                !pfrac.alloc(static_cast<size_t>(capacity) * static_cast<size_t>(shadow_samples) * 48)) {
                lobe_cap = 0;
                return PE_ERR_OOM;
            }
            lobe_cap = capacity;
        }
        return PE_OK;
    }

    PathBuffers view() {
        PathBuffers pb{};
        pb.ro[0] = ro0.as<float4>();
        pb.rd[0] = rd0.as<float4>();
        pb.ro[1] = ro1.as<float4>();
        pb.rd[1] = rd1.as<float4>();
        pb.thr = thr.as<float4>();
        pb.L = L.as<float4>();
        pb.last_p = last_p.as<float4>();
        pb.last_n = last_n.as<float4>();
        pb.ids = ids.as<uint4>();
        pb.ext = ext.as<uint4>();
        pb.pend = pend.as<float4>();
        pb.res = res.as<float4>();
        pb.aov_res = nullptr;
        pb.pass_res = nullptr;
        pb.lobe = lobe.as<float4>();
        pb.pfrac = pfrac.as<float4>();
        pb.hits = hits.as<float4>();
        pb.sq_o = sq_o.as<float4>();
        pb.sq_d = sq_d.as<float4>();
        pb.fl[0] = fl0.as<uint32_t>();
        pb.fl[1] = fl1.as<uint32_t>();
        pb.ctr = ctr.as<uint32_t>();
        pb.seg_cap = seg_cap;
        pb.stats = stats.as<unsigned long long>();
        return pb;
    }
};

Wavefront* wavefront_of(pe_device* dev);
void wavefront_release(pe_device* dev);

}  // namespace pe

using namespace pe;

namespace pe {

Wavefront* wavefront_of(pe_device* dev) {
    if (!dev->wavefront) dev->wavefront = new Wavefront();
    return static_cast<Wavefront*>(dev->wavefront);
}

void wavefront_release(pe_device* dev) {
    delete static_cast<Wavefront*>(dev->wavefront);
    dev->wavefront = nullptr;
}

}  // namespace pe

namespace pe {

Wavefront* wavefront_new() { return new Wavefront(); }
void wavefront_delete(Wavefront* wf) { delete wf; }
int wavefront_reserve(Wavefront* wf, uint32_t paths, uint64_t results, int shadow_quality) {
    const int rc = wf->ensure(paths, std::clamp(shadow_quality, 1, kMaxShadowSamples));
    return rc != PE_OK ? rc : wf->ensure_results(results, false);
}
uint64_t wavefront_bytes_per_path(int shadow_quality) {
    return Wavefront::bytes_per_path(std::clamp(shadow_quality, 1, kMaxShadowSamples));
}
uint64_t wavefront_device_bytes(const Wavefront* wf) {
    return static_cast<uint64_t>(wf->capacity) * Wavefront::bytes_per_path(std::max(1, wf->shadow_samples)) +
           wf->result_cap * Wavefront::kResultBytes + wf->aov_cap * Wavefront::kAovResultBytes +
           wf->pass_cap * 16 + static_cast<uint64_t>(wf->lobe_cap) * 48 * (1 + std::max(1, wf->shadow_samples));
}

// Renders samples [first_sample, first_sample + sample_count) of every pixel into `film` on `lane`: the
// work of RenderDevice::render (render_device.h:84-85) for one stream, with that stream's own path state and
// traversal-stack spill, so lanes on different streams never share device scratch.
int render_samples(pe_device* dev, const RenderLane& lane, pe_scene* scene, const pe_camera* camera,
                   const pe_settings* settings, pe_film* film, int first_sample, int sample_count,
                   const volatile int32_t* cancel, pe_batch_result* result, pe_guide* guide, const AovTargets* aovs) {
    if (result) {
        result->completed = 0;
        result->samples_traced = 0;
        result->extension_rays = 0;
        result->shadow_rays = 0;
    }
    if (!dev || !scene || !camera || !settings || !film) {
        set_error("pe_render: NULL argument");
        return PE_ERR_INVALID;
    }
    if (scene->dev != dev || film->dev != dev) {
        set_error("pe_render: scene or film belongs to another device");
        return PE_ERR_INVALID;
    }
    // The scene stays as it is for the whole render: updates, parking and release wait for it (and a scene a
    // render-job server has parked in host memory comes back first).
    std::shared_lock<std::shared_mutex> scene_use(scene->use);
    for (int attempt = 0; !scene->resident; ++attempt) {
        scene_use.unlock();
        {
            std::unique_lock<std::shared_mutex> ex(scene->use);
            const int rc = scene_restore(*scene);
            if (rc != PE_OK) return rc;
        }
        scene_use.lock();
        if (attempt > 8) {
            set_error("pe_render: the scene keeps being parked");
            return PE_ERR_INVALID;
        }
    }
    if (scene->broken) {
        set_error("pe_render: the scene's last update failed; upload it again");
        return PE_ERR_INVALID;
    }
    if (guide && guide->dev != dev) {
        set_error("pe_render: the guide belongs to another device");
        return PE_ERR_INVALID;
    }
    FilmView aov_view[3] = {};
    bool want_aov = false;
    if (aovs) {
        pe_film* const fs[3] = {aovs->albedo, aovs->normal, aovs->alpha};
        for (int k = 0; k < 3; ++k) {
            if (!fs[k]) continue;
            if (fs[k]->dev != dev || fs[k]->width != film->width || fs[k]->height != film->height) {
                set_error("pe_render: an AOV film of another device or size");
                return PE_ERR_INVALID;
            }
            aov_view[k] = FilmView{fs[k]->sum.as<float>(), fs[k]->count.as<uint32_t>(), fs[k]->lum2.as<float>(),
                                   fs[k]->rejected.as<unsigned long long>()};
            want_aov = true;
        }
    }
    // Pass films (RenderPassFilms): every one optional, each of the beauty film's device and size.
    PassFilms pass_films{};
    bool want_passes = false;
    if (aovs && aovs->passes) {
        const pe_pass_targets& pt = *aovs->passes;
        pe_film* fs[kPassFilms];
        for (int k = 0; k < 8; ++k) fs[k] = pt.light[k];
        for (int k = 0; k < 16; ++k) fs[8 + k] = pt.light_groups[k];
        fs[24] = pt.depth;
        fs[25] = pt.position;
        fs[26] = pt.object_id;
        fs[27] = pt.material_id;
        for (int k = 0; k < kPassFilms; ++k) {
            if (!fs[k]) continue;
            if (fs[k]->dev != dev || fs[k]->width != film->width || fs[k]->height != film->height) {
                set_error("pe_render: a pass film of another device or size");
                return PE_ERR_INVALID;
            }
            pass_films.f[k] = FilmView{fs[k]->sum.as<float>(), fs[k]->count.as<uint32_t>(), fs[k]->lum2.as<float>(),
                                       fs[k]->rejected.as<unsigned long long>()};
            want_passes = true;
        }
    }
    CryptoView crypto[2] = {};
    std::vector<float> crypto_table;
    int crypto_n[2] = {0, 0};
    if (aovs && aovs->passes) {
        const pe_pass_targets& pt = *aovs->passes;
        pe_crypto_film* cf[2] = {pt.crypto_object, pt.crypto_material};
        const float* tables[2] = {pt.object_ids, pt.material_ids};
        const int counts[2] = {pt.num_object_ids, pt.num_material_ids};
        for (int k = 0; k < 2; ++k) {
            if (!cf[k]) continue;
            if (cf[k]->dev != dev || cf[k]->width != film->width || cf[k]->height != film->height) {
                set_error("pe_render: a cryptomatte film of another device or size");
                return PE_ERR_INVALID;
            }
            if (counts[k] < 0 || (counts[k] > 0 && !tables[k])) {
                set_error("pe_render: a cryptomatte id table without entries");
                return PE_ERR_INVALID;
            }
            crypto[k] = CryptoView{cf[k]->ids.as<float>(), cf[k]->counts.as<uint32_t>(), cf[k]->totals.as<uint32_t>(),
                                   cf[k]->dropped.as<unsigned long long>()};
            crypto_n[k] = counts[k];
            want_passes = true;
        }
        crypto_table.assign(1 + static_cast<size_t>(crypto_n[0]) + static_cast<size_t>(crypto_n[1]), 0.0f);
        if (crypto_n[0] > 0) std::copy(pt.object_ids, pt.object_ids + crypto_n[0], crypto_table.begin());
        if (crypto_n[1] > 0) std::copy(pt.material_ids, pt.material_ids + crypto_n[1], crypto_table.begin() + crypto_n[0]);
    }
    const bool want_crypto = crypto[0].ids || crypto[1].ids;
    if (first_sample < 0 || sample_count < 0) {
        set_error("pe_render: negative sample range");
        return PE_ERR_INVALID;
    }
    if (!(camera->lens_radius >= 0.0f) || (camera->has_shutter_end && !(camera->end_lens_radius >= 0.0f))) {
        set_error("pe_render: the camera's lens radius must be >= 0");
        return PE_ERR_INVALID;
    }
    if (settings->max_bounces < 0 || settings->max_bounces > 100) {
        set_error("pe_render: max_bounces must be in [0, 100]");
        return PE_ERR_INVALID;
    }
    if (result) result->completed = 1;
    if (sample_count == 0) return PE_OK;
    try {
        PE_HIP_CHECK(hipSetDevice(dev->ordinal));
        Wavefront* wf = lane.wf;
        const int S = std::clamp(settings->shadow_quality, 1, kMaxShadowSamples);  // PathTracer ctor (path_tracer.h:43)
        const uint32_t pixels = static_cast<uint32_t>(film->width) * static_cast<uint32_t>(film->height);
        // The window: samples of every pixel whose results are held until the film pass adds them in sample
        // order; as asked, or (0) up to 2^28 samples (4 GB of results) within an eighth of the free device
        // memory. The path pool (ids in flight, regenerated as paths end) is sized independently: as asked, or
        // (0) 2^26 ids, and never above the window.
        const bool adaptive = settings->adaptive != 0;
        const uint32_t pass_count = kLightPasses + static_cast<uint32_t>(scene->view.num_groups) + 2u;
        int spl_req = settings->samples_per_launch;
        if (spl_req <= 0) {
            size_t free_b = 0, total_b = 0;
            if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) free_b = size_t{1} << 32;
            const uint64_t per_sample = Wavefront::kResultBytes + (want_aov ? Wavefront::kAovResultBytes : 0) +
                                        (want_passes ? 16ull * pass_count : 0);
            const uint64_t target = std::min<uint64_t>(kAutoWindowSamples, free_b / 8 / per_sample);
            spl_req = static_cast<int>(std::clamp<uint64_t>(target / std::max<uint32_t>(1, pixels), 1, 1024));
        }
        // Window sample indices carry a flag bit in the queue (31 bits); path ids and held-back slots
        // (pid * S + s) are 32-bit.
        if (pixels > kMaxWindowSamples) {
            set_error("pe_render: film too large for 31-bit sample indices");
            return PE_ERR_INVALID;
        }
        const uint64_t max_window = std::max<uint64_t>(
            pixels, std::min<uint64_t>(want_passes ? kPassWindowSamples : kMaxWindowSamples, uint64_t{pixels} * spl_req));
        // Adaptive sampling renders in groups that end where the reference checks a pixel's error (every 8th
        // sample), so a window holds at most 8 samples of each pixel still rendering.
        const float threshold = std::max(1e-4f, settings->noise_threshold);
        const int min_spp = std::max(2, settings->min_adaptive_samples);
        const int spl_cap = adaptive ? std::min(8, static_cast<int>(std::max<uint64_t>(1, max_window / pixels)))
                                     : static_cast<int>(std::max<uint64_t>(1, max_window / pixels));
        const uint64_t window = uint64_t{pixels} * std::min(spl_cap, sample_count);
        // Shadow-queue entries per id: the NEE samples, and as many AO rays at the first vertex with contact AO
        // (their records follow the NEE records in pend). Record indices stay below the AO tag bit.
        const int Sq = settings->ao_strength > 0.0f ? 2 * S : S;
        const uint64_t pool_limit = std::min<uint64_t>(kMaxPoolIds, 0x7FFFFFFFull / static_cast<uint64_t>(Sq) - kQueueSegs * 256u);
        const uint64_t pool_req = settings->paths_in_flight > 0 ? static_cast<uint64_t>(settings->paths_in_flight) : kDefaultPoolIds;
        const uint32_t pool = static_cast<uint32_t>(std::min({pool_req, window, pool_limit}));
        int rc = wf->ensure(pool, Sq);
        if (rc == PE_OK) rc = wf->ensure_results(window, want_aov);
        if (rc == PE_OK && want_passes) rc = wf->ensure_passes(window, pass_count);
        if (rc != PE_OK) return rc;
        if (want_crypto) {
            if (!wf->crypto_ids.upload(crypto_table.data(), crypto_table.size() * sizeof(float), lane.stream))
                return PE_ERR_OOM;
            PE_HIP_CHECK(hipStreamSynchronize(lane.stream));
        }
        if (guide) {
            rc = guide_reserve_paths(guide, wf->capacity, wf->slots());
            if (rc != PE_OK) return rc;
        }

        LaunchParams lp{};
        lp.width = film->width;
        lp.height = film->height;
        lp.max_depth = settings->max_bounces;
        lp.rr_depth = settings->rr_depth;
        lp.shadow_samples = S;
        lp.shadow_slots = Sq;
        lp.ao_base = wf->capacity * static_cast<uint32_t>(S);
        lp.clamp_indirect = settings->clamp_indirect > 0.0f ? settings->clamp_indirect : 0.0f;
        lp.seed = settings->seed;
        lp.filter_width = std::clamp(settings->filter_width, 0.25f, 8.0f);
        lp.filter_radius = 0.5f * lp.filter_width;
        lp.filter_box = settings->filter_type == PE_FILTER_BOX ? 1 : 0;
        lp.ao_strength = settings->ao_strength > 0.0f ? settings->ao_strength : 0.0f;
        lp.transport = scene->view.transport || lp.ao_strength > 0.0f;  // AO rides the ext column
        if (!lp.filter_box) {
            if (wf->filter_type != settings->filter_type || wf->filter_width != settings->filter_width) {
                const std::vector<float> cdf = filter_cdf(settings->filter_type, settings->filter_width);
                if (cdf.empty()) {
                    lp.filter_box = 1;
                } else {
                    if (!wf->filter.upload(cdf.data(), cdf.size() * 4, lane.stream)) return PE_ERR_OOM;
                    PE_HIP_CHECK(hipStreamSynchronize(lane.stream));
                    wf->filter_type = settings->filter_type;
                    wf->filter_width = settings->filter_width;
                }
            }
            lp.filter_cdf = wf->filter.as<float>();
        }
        const auto v3 = [](const float* f) { return mk(f[0], f[1], f[2]); };
        lp.cam = CameraPose{v3(camera->position), v3(camera->lower_left), v3(camera->horizontal), v3(camera->vertical),
                            v3(camera->u), v3(camera->v), v3(camera->forward), camera->lens_radius};
        lp.cam_end = CameraPose{v3(camera->end_position), v3(camera->end_lower_left), v3(camera->end_horizontal),
                                v3(camera->end_vertical), v3(camera->end_u), v3(camera->end_v), v3(camera->end_forward),
                                camera->end_lens_radius};
        lp.cam_ortho = camera->orthographic ? 1 : 0;
        lp.cam_blades = camera->blades >= 3 ? camera->blades : 0;
        lp.cam_blade_rotation = camera->blade_rotation;
        lp.cam_moving = settings->motion_blur && camera->has_shutter_end ? 1 : 0;
        lp.camera_dims = settings->motion_blur ? 3u : 2u;

        PathBuffers pb = wf->view();
        pb.aov_res = want_aov ? wf->aov_res.as<float4>() : nullptr;
        lp.want_aov = want_aov ? 1 : 0;
        pb.pass_res = want_passes ? wf->pass_res.as<float4>() : nullptr;
        lp.want_passes = want_passes ? 1 : 0;
        lp.num_groups = scene->view.num_groups;
        lp.pass_count = pass_count;
        pb.mix = guide ? guide->mix.as<float4>() : nullptr;
        pb.grec = guide ? guide->path_records.as<float4>() : nullptr;
        lp.guided_bounces = guide ? guide->cfg.guided_bounces : 0;
        lp.grec_out = guide ? guide->records.as<float4>() : nullptr;
        lp.grec_slots = guide ? std::max<uint32_t>(1, guide->record_slots) : 1u;
        hipStream_t s = lane.stream;
        const int gblocks = lane.grid_blocks;
        if (gblocks < static_cast<int>(kQueueSegs)) {  // every queue segment needs a block of its own
            set_error("pe_render: the persistent grid is smaller than the queue's segment count");
            return PE_ERR_UNSUPPORTED;
        }
        hipLaunchKernelGGL(k_zero_stats, dim3(1), dim3(64), 0, s, pb);
        const bool count = settings->count_visits != 0;
        uint32_t* spill = lane.spill;
        const SceneView view = scene->view;
        // One iteration: closest hit of the continuation rays, new samples on the free ids, shade, shadow rays,
        // counter rotation.
        auto iterate = [&](int cur) {
            wf->prof.begin(kKRegen, s);
            if (guide) hipLaunchKernelGGL(k_regen<true>, dim3(gblocks), dim3(kBlock), 0, s, lp, pb, cur);
            else hipLaunchKernelGGL(k_regen<false>, dim3(gblocks), dim3(kBlock), 0, s, lp, pb, cur);
            wf->prof.end(s);
            wf->prof.begin(kKIntersect, s);
            if (count) hipLaunchKernelGGL(k_intersect<true>, dim3(gblocks), dim3(kTraceBlock), 0, s, view, pb, cur, spill);
            else hipLaunchKernelGGL(k_intersect<false>, dim3(gblocks), dim3(kTraceBlock), 0, s, view, pb, cur, spill);
            wf->prof.end(s);
            if (guide && lp.guide_fraction > 0.0f) {
                wf->prof.begin(kKGuide, s);
                const RenderQueries q{pb.ro[cur], pb.rd[cur], pb.hits, seg_counter(pb.ctr, kCtrCur, 0), pb.seg_cap, pb.mix,
                                      static_cast<uint32_t>(std::max(0, lp.guided_bounces)),
                                      seg_counter(pb.ctr, kCtrBase, 0), lp.guided_bounces <= 1};
                GuideView gv = guide->view(lp.guide_fraction);
                gv.inferred = pb.stats + 7;
                hipLaunchKernelGGL(k_guide_infer<RenderQueries>, dim3(std::max(16, dev->cu_count * 2)),
                                   dim3(kInferBlock), 0, s, gv, q, lp.num_paths);
                wf->prof.end(s);
            }
            wf->prof.begin(kKShade, s);
            if (want_passes) {
                if (guide) hipLaunchKernelGGL(k_shade_passes<true>, dim3(gblocks), dim3(kBlock), 0, s, view, lp, pb, cur);
                else hipLaunchKernelGGL(k_shade_passes<false>, dim3(gblocks), dim3(kBlock), 0, s, view, lp, pb, cur);
            } else if (lp.transport) {
                if (guide) hipLaunchKernelGGL((k_shade<true, true>), dim3(gblocks), dim3(kBlock), 0, s, view, lp, pb, cur);
                else hipLaunchKernelGGL((k_shade<true, false>), dim3(gblocks), dim3(kBlock), 0, s, view, lp, pb, cur);
            } else {
                if (guide) {
                    hipLaunchKernelGGL((k_shade<false, true>), dim3(gblocks), dim3(kBlock), 0, s, view, lp, pb, cur);
                } else {
                    hipLaunchKernelGGL(k_shade_plain, dim3(gblocks), dim3(kBlock), 0, s, view, lp, pb, cur);
                }
            }
            wf->prof.end(s);
            wf->prof.begin(kKShadow, s);
            if (view.has_transmitters) {
                if (count)
                    hipLaunchKernelGGL(k_shadow_transmit<true>, dim3(gblocks), dim3(kTraceBlock), 0, s, view, pb, Sq,
                                       lp.clamp_indirect, spill);
                else
                    hipLaunchKernelGGL(k_shadow_transmit<false>, dim3(gblocks), dim3(kTraceBlock), 0, s, view, pb, Sq,
                                       lp.clamp_indirect, spill);
            } else if (count) {
                hipLaunchKernelGGL(k_shadow<true>, dim3(gblocks), dim3(kTraceBlock), 0, s, view, pb, Sq, spill);
            } else {
                hipLaunchKernelGGL(k_shadow<false>, dim3(gblocks), dim3(kTraceBlock), 0, s, view, pb, Sq, spill);
            }
            wf->prof.end(s);
            hipLaunchKernelGGL(k_rotate, dim3(1), dim3(kQueueSegs), 0, s, pb, lp.num_paths);
        };
        uint64_t traced = 0;
        bool cancelled = false;
        const uint32_t cblocks = (pixels + kCompactSpan - 1) / kCompactSpan;
        const int pix_blocks = static_cast<int>(std::min<uint32_t>((pixels + kBlock - 1) / kBlock, 65535u));
        lp.num_pixels = pixels;
        lp.pixel_list = nullptr;
        lp.active = nullptr;
        lp.adapt_check = 0;
        lp.noise_threshold = threshold;
        if (adaptive) {
            rc = wf->ensure_adaptive(pixels, cblocks);
            if (rc != PE_OK) return rc;
            lp.active = wf->adapt_active.as<uint32_t>();
            if (first_sample >= min_spp) {
                hipLaunchKernelGGL(k_adapt_clean, dim3(pix_blocks), dim3(kBlock), 0, s, film->sum.as<float>(),
                                   film->count.as<uint32_t>(), film->lum2.as<float>(), pixels, threshold,
                                   wf->adapt_clean.as<uint32_t>());
                hipLaunchKernelGGL(k_adapt_mask, dim3(pix_blocks), dim3(kBlock), 0, s, wf->adapt_clean.as<uint32_t>(),
                                   film->width, film->height, lp.active);
            } else {
                hipLaunchKernelGGL(k_fill_u32, dim3(pix_blocks), dim3(kBlock), 0, s, lp.active, pixels, 1u);
            }
        }
        auto cancel_requested = [&]() {
            return (cancel && *cancel) || (aovs && aovs->cancel_poll && aovs->cancel_poll(aovs->cancel_user) != 0);
        };
        for (int done = 0; done < sample_count;) {
            if (cancel_requested()) {
                cancelled = true;
                break;
            }
            int spl = std::min(spl_cap, sample_count - done);
            if (adaptive) {
                // This group: up to the next multiple of 8; its end is checked unless it ends the batch
                // (renderer.cpp:500: s + 1 < chunkEnd, s + 1 >= minSpp, (s + 1) % 8 == 0).
                const int s0 = first_sample + done, end = first_sample + sample_count;
                const int launch_end = std::min({end, (s0 / 8 + 1) * 8, s0 + spl_cap});
                spl = launch_end - s0;
                lp.adapt_check = launch_end < end && launch_end >= min_spp && launch_end % 8 == 0 ? 1 : 0;
                uint32_t* counts = wf->adapt_counts.as<uint32_t>();
                hipLaunchKernelGGL(k_count_active, dim3(cblocks), dim3(kBlock), 0, s, lp.active, pixels, counts);
                hipLaunchKernelGGL(k_scan_counts, dim3(1), dim3(kBlock), 0, s, counts, cblocks);
                hipLaunchKernelGGL(k_scatter_active, dim3(cblocks), dim3(kBlock), 0, s, lp.active, pixels, counts,
                                   wf->adapt_list.as<uint32_t>());
                uint32_t m = 0;
                PE_HIP_CHECK(hipMemcpyAsync(&m, counts + cblocks, 4, hipMemcpyDeviceToHost, s));
                PE_HIP_CHECK(hipStreamSynchronize(s));
                if (m == 0) break;  // every pixel has stopped: the batch is over
                lp.num_pixels = m;
                lp.pixel_list = wf->adapt_list.as<uint32_t>();
            }
            lp.first_sample = first_sample + done;
            lp.spl = spl;
            lp.num_paths = lp.num_pixels * static_cast<uint32_t>(spl);
            // Guiding: the first window samples the BSDF only and collects; later ones mix in the trained guide.
            lp.guide_fraction = guide && guide->steps > 0 ? guide->cfg.guide_fraction : 0.0f;
            lp.guide_records = guide && guide->training() ? 1 : 0;
            lp.grec_salt = guide ? static_cast<uint64_t>(guide->launches) * 0x9e3779b97f4a7c15ull : 0ull;
            if (lp.guide_records) {
                rc = guide_clear_records(guide, s);
                if (rc != PE_OK) return rc;
            }
            // The window's pool: as many ids per segment as the pool size allows, and no more than the segment
            // has samples.
            const uint32_t ids_per_seg = std::min<uint32_t>(wf->seg_cap, (pool + kQueueSegs - 1) / kQueueSegs);
            const int init_blocks =
                static_cast<int>(std::min<uint64_t>((uint64_t{ids_per_seg} * kQueueSegs + kBlock - 1) / kBlock, 65535u));
            wf->prof.begin(kKGenerate, s);
            hipLaunchKernelGGL(k_window_init, dim3(init_blocks), dim3(kBlock), 0, s, lp, pb, ids_per_seg);
            wf->prof.end(s);
            // Iterations in chunks; after each chunk the remaining-work counter is copied to pinned memory and an
            // event recorded. The host reads a chunk's counter while the next chunk runs, so the device never
            // waits for the host; the chunk queued after the last one needed runs as empty kernels. Every path
            // ends within the reference's own limits (max_bounces + 1 vertices, 256 walk steps per vertex, 64
            // pass-throughs, 64 camera skips), so the iteration count is bounded.
            const int per_path = (lp.max_depth + 1) * static_cast<int>(kMaxWalkSteps + 1) + 2 * 64 + 8;
            const uint64_t generations = lp.num_paths / std::max<uint32_t>(1, wf->capacity) + 2;
            const uint64_t limit = static_cast<uint64_t>(per_path) * generations;
            uint64_t it = 0;
            auto enqueue_chunk = [&](int slot, int iters) -> int {
                for (int k = 0; k < iters; ++k, ++it) iterate(static_cast<int>(it & 1));
                PE_HIP_CHECK(hipMemcpyAsync(&wf->host_live[slot], pb.ctr + kCtrTotal, 4, hipMemcpyDeviceToHost, s));
                PE_HIP_CHECK(hipEventRecord(wf->chunk_done[slot], s));
                return PE_OK;
            };
            rc = enqueue_chunk(0, kChunkIterations);
            if (rc != PE_OK) return rc;
            for (int n = 1;; ++n) {
                rc = enqueue_chunk(n & 1, kChunkIterations);
                if (rc != PE_OK) return rc;
                PE_HIP_CHECK(hipEventSynchronize(wf->chunk_done[(n - 1) & 1]));
                if (wf->host_live[(n - 1) & 1] == 0) break;
                if (cancel_requested()) {
                    cancelled = true;
                    break;
                }
                if (it > limit) {
                    set_error("pe_render: paths still alive past the bounce, walk and pass-through limits");
                    return PE_ERR_INVALID;
                }
            }
            if (cancelled) break;  // the window's samples are dropped: the film keeps whole windows only
            const int film_blocks = static_cast<int>(std::min<uint32_t>((lp.num_pixels + kBlock - 1) / kBlock, 65535u));
            wf->prof.begin(kKFilm, s);
            hipLaunchKernelGGL(k_film, dim3(film_blocks), dim3(kBlock), 0, s, lp, pb, film->sum.as<float>(),
                               film->count.as<uint32_t>(), film->lum2.as<float>(),
                               film->rejected.as<unsigned long long>(), aov_view[0], aov_view[1], aov_view[2]);
            if (want_passes) {
                const int pass_blocks = static_cast<int>(
                    std::min<uint64_t>((uint64_t{lp.num_pixels} * kPassFilms + kBlock - 1) / kBlock, 65535u));
                hipLaunchKernelGGL(k_film_passes, dim3(pass_blocks), dim3(kBlock), 0, s, lp, pb, pass_films,
                                   view.node_uids);
            }
            if (want_crypto) {
                const float* tbl = wf->crypto_ids.as<float>();
                hipLaunchKernelGGL(k_film_crypto, dim3(film_blocks), dim3(kBlock), 0, s, lp, pb, crypto[0], crypto[1],
                                   tbl, crypto_n[0], tbl + crypto_n[0], crypto_n[1]);
            }
            wf->prof.end(s);
            PE_HIP_CHECK(hipGetLastError());
            if (guide) {
                if (lp.guide_records) {
                    wf->prof.begin(kKTrain, s);
                    rc = guide_train_from_records(guide, s);
                    wf->prof.end(s);
                    if (rc != PE_OK) return rc;
                }
                ++guide->launches;
            }
            done += spl;
            traced += static_cast<uint64_t>(lp.num_pixels) * spl;
        }
        unsigned long long st[kStatSlots] = {};
        PE_HIP_CHECK(hipMemcpyAsync(st, pb.stats, sizeof(st), hipMemcpyDeviceToHost, s));
        PE_HIP_CHECK(hipStreamSynchronize(s));
        if (result) {
            result->completed = cancelled ? 0 : 1;
            result->samples_traced = traced;
            result->extension_rays = st[0];
            result->shadow_rays = st[1];
            result->node_visits = st[2];
            result->tri_visits = st[3];
            result->shadow_node_visits = st[4];
            result->shadow_tri_visits = st[5];
            result->closest_hits = st[10];
            result->max_stack_depth = static_cast<uint32_t>(st[8]);
            result->stack_overflows = static_cast<uint32_t>(std::min<unsigned long long>(st[9], 0xFFFFFFFFull));
        }
        if (st[9] != 0) {
            set_error("pe_render: " + std::to_string(st[9]) + " traversals overflowed the " +
                      std::to_string(kStackDepth) + "-entry stack (a BVH deeper than the build allows); the film "
                      "holds wrong hits");
            return PE_ERR_INVALID;
        }
        if (guide) {
            guide->collected += st[6];
            guide->inferred += st[7];
        }
        if (cancelled) {
            set_error("pe_render: cancelled");
            return PE_ERR_CANCELLED;
        }
        return PE_OK;
    } catch (const std::exception& e) {
        set_error(std::string("pe_render: ") + e.what());
        return PE_ERR_INVALID;
    }
}

}  // namespace pe

extern "C" int pe_render(pe_device* dev, pe_scene* scene, const pe_camera* camera, const pe_settings* settings,
                         pe_film* film, int first_sample, int sample_count, const volatile int32_t* cancel,
                         pe_batch_result* result) {
    if (!dev) {
        if (result) *result = pe_batch_result{};
        set_error("pe_render: NULL device");
        return PE_ERR_INVALID;
    }
    try {
        PE_HIP_CHECK(hipSetDevice(dev->ordinal));
        int blocks = 0;
        const int rc = ensure_trace_resources(dev, blocks);
        if (rc != PE_OK) return rc;
        const RenderLane lane{dev->stream, wavefront_of(dev), dev->spill.as<uint32_t>(), blocks};
        return render_samples(dev, lane, scene, camera, settings, film, first_sample, sample_count, cancel, result,
                              nullptr);
    } catch (const std::exception& e) {
        set_error(std::string("pe_render: ") + e.what());
        return PE_ERR_INVALID;
    }
}

extern "C" int pe_render_targets(pe_device* dev, pe_scene* scene, const pe_camera* camera, const pe_settings* settings,
                                 const pe_targets* targets, int first_sample, int sample_count,
                                 const volatile int32_t* cancel, pe_batch_result* result) {
    if (!dev || !targets) {
        if (result) *result = pe_batch_result{};
        set_error("pe_render_targets: NULL argument");
        return PE_ERR_INVALID;
    }
    try {
        PE_HIP_CHECK(hipSetDevice(dev->ordinal));
        int blocks = 0;
        const int rc = ensure_trace_resources(dev, blocks);
        if (rc != PE_OK) return rc;
        const RenderLane lane{dev->stream, wavefront_of(dev), dev->spill.as<uint32_t>(), blocks};
        AovTargets aovs;
        aovs.albedo = targets->albedo;
        aovs.normal = targets->normal;
        aovs.alpha = targets->alpha;
        aovs.cancel_poll = targets->cancel_poll;
        aovs.cancel_user = targets->cancel_user;
        aovs.passes = targets->passes;
        return render_samples(dev, lane, scene, camera, settings, targets->beauty, first_sample, sample_count, cancel,
                              result, targets->guide, &aovs);
    } catch (const std::exception& e) {
        set_error(std::string("pe_render_targets: ") + e.what());
        return PE_ERR_INVALID;
    }
}

extern "C" int pe_render_guided(pe_device* dev, pe_scene* scene, pe_guide* guide, const pe_camera* camera,
                                const pe_settings* settings, pe_film* film, int first_sample, int sample_count,
                                const volatile int32_t* cancel, pe_batch_result* result) {
    if (!dev) {
        if (result) *result = pe_batch_result{};
        set_error("pe_render_guided: NULL device");
        return PE_ERR_INVALID;
    }
    try {
        PE_HIP_CHECK(hipSetDevice(dev->ordinal));
        int blocks = 0;
        const int rc = ensure_trace_resources(dev, blocks);
        if (rc != PE_OK) return rc;
        const RenderLane lane{dev->stream, wavefront_of(dev), dev->spill.as<uint32_t>(), blocks};
        return render_samples(dev, lane, scene, camera, settings, film, first_sample, sample_count, cancel, result,
                              guide);
    } catch (const std::exception& e) {
        set_error(std::string("pe_render_guided: ") + e.what());
        return PE_ERR_INVALID;
    }
}

extern "C" int pe_profile_enable(pe_device* dev, int enable) {
    if (!dev) {
        set_error("pe_profile_enable: NULL device");
        return PE_ERR_INVALID;
    }
    Wavefront* wf = wavefront_of(dev);
    wf->prof.drain();
    wf->prof.on = enable != 0;
    return PE_OK;
}

extern "C" int pe_profile_read(pe_device* dev, pe_kernel_stats* out, int capacity, int* count) {
    if (!dev || (capacity > 0 && !out)) {
        set_error("pe_profile_read: NULL argument");
        return PE_ERR_INVALID;
    }
    Wavefront* wf = wavefront_of(dev);
    wf->prof.drain();
    const int n = std::min<int>(capacity, kKindCount);
    for (int k = 0; k < n; ++k) {
        std::snprintf(out[k].name, sizeof(out[k].name), "%s", kKernelNames[k]);
        out[k].launches = wf->prof.launches[k];
        out[k].total_ms = wf->prof.ms[k];
        wf->prof.launches[k] = 0;
        wf->prof.ms[k] = 0.0;
    }
    if (count) *count = n;
    return PE_OK;
}

// capi.hip — the C-ABI of libphoton_hip.so: device registry, scene upload, film, and the kernel-level
// batch entry points (Sobol values, closest hit, any hit). The wavefront integrator behind pe_render
// lives in wavefront.hip.
#include "photon_hip.h"
#include "bvh_build.h"
#include "pe_internal.h"
#include "pe_scene_host.h"
#include "pe_sobol.h"
#include "pe_traverse.h"

#include <algorithm>
#include <cstring>
#include <new>
#include <string>

namespace pe {

namespace {
thread_local std::string g_last_error;
}

void set_error(const std::string& msg) { g_last_error = msg; }

bool DevBuf::alloc(size_t bytes) {
    release();
    if (bytes == 0) return true;
    void* p = nullptr;
    const hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
        set_error("hipMalloc(" + std::to_string(bytes) + " B) failed: " + hipGetErrorString(e));
        return false;
    }
    m_ptr = p;
    m_bytes = bytes;
    return true;
}

bool DevBuf::upload(const void* src, size_t bytes, hipStream_t s) {
    if (!alloc(bytes)) return false;
    if (bytes == 0) return true;
    const hipError_t e = hipMemcpyAsync(m_ptr, src, bytes, hipMemcpyHostToDevice, s);
    if (e != hipSuccess) {
        set_error(std::string("hipMemcpyAsync failed: ") + hipGetErrorString(e));
        return false;
    }
    return true;
}

void DevBuf::release() {
    if (m_ptr) (void)hipFree(m_ptr);
    m_ptr = nullptr;
    m_bytes = 0;
}

// ---------------------------------------------------------------------------------------------------------
// Kernels of the batch entry points.

__global__ void __launch_bounds__(256) k_sobol_batch(uint64_t seed, size_t n, const int32_t* __restrict__ px,
                                                     const int32_t* __restrict__ py, const uint32_t* __restrict__ index,
                                                     const uint32_t* __restrict__ dim, uint2* __restrict__ out) {
    const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
        const uint32_t ph = sobol_pixel_hash(seed, px[i], py[i]);
        uint32_t xb, yb;
        sobol_2d_bits(ph, rev_bits(index[i]), dim[i], xb, yb);
        out[i] = make_uint2(xb, yb);
    }
}

// Batch rays are 8 floats: origin xyz, direction xyz, tmin, tmax (read as two float4).
struct BatchRays {
    const float4* r;
    __device__ __forceinline__ void fetch(uint32_t i, float4& a, float4& b) const {
        a = r[2 * static_cast<size_t>(i)];
        b = r[2 * static_cast<size_t>(i) + 1];
    }
    __device__ __forceinline__ TraceRay decode(const float4& a, const float4& b) const {
        return TraceRay{mk(a.x, a.y, a.z), mk(a.w, b.x, b.y), b.z, b.w, 0u};
    }
};

struct BatchClosest : BatchRays {
    const TriShade* shade;
    int4* hits;
    float2* bary;
    __device__ __forceinline__ bool closest(uint32_t i, const TraceRay&, float t, int32_t tri, float u, float v, float4&,
                                            float4&) const {
        int4 o;
        if (tri >= 0) {
            const TriShade s = shade[tri];
            o = make_int4(__float_as_int(t), s.shape, s.prim, tri);
        } else {
            o = make_int4(__float_as_int(-1.0f), -1, -1, -1);
            u = v = 0.0f;
        }
        hits[i] = o;
        if (bary) bary[i] = make_float2(u, v);
        return false;
    }
};

struct BatchAny : BatchRays {
    uint8_t* occ;
    __device__ __forceinline__ void any(uint32_t i, const TraceRay&, bool occluded) const { occ[i] = occluded ? 1 : 0; }
};

__global__ void __launch_bounds__(kTraceBlock) k_intersect_batch(const BvhNode4* __restrict__ nodes,
                                                                 const TriAccel* __restrict__ tris, BatchClosest job,
                                                                 uint32_t n, uint32_t num_nodes, uint32_t* __restrict__ spill,
                                                                 unsigned long long* __restrict__ diag) {
    __shared__ uint32_t lds_stack[kLdsDepth * kTraceBlock];
    __shared__ float4 node_tile[kBvhTileNodes * 8];
    load_node_tile(node_tile, nodes, num_nodes);
    const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    TraceStack st{lds_stack + threadIdx.x, spill + gid, gridDim.x * blockDim.x, 0};
    uint32_t nv = 0, tv = 0;
    trace_queue<false, true>(job, QueueSegs{nullptr, n, n, 1u, reinterpret_cast<uint32_t*>(diag + 2)}, nodes, tris, st, nv, tv, node_tile,
                             min(num_nodes, kBvhTileNodes));
    report_stack(diag, st, true);
}

__global__ void __launch_bounds__(kTraceBlock) k_occluded_batch(const BvhNode4* __restrict__ nodes,
                                                                const TriAccel* __restrict__ tris, BatchAny job,
                                                                uint32_t n, uint32_t num_nodes, uint32_t* __restrict__ spill,
                                                                unsigned long long* __restrict__ diag) {
    __shared__ uint32_t lds_stack[kLdsDepth * kTraceBlock];
    __shared__ float4 node_tile[kBvhTileNodes * 8];
    load_node_tile(node_tile, nodes, num_nodes);
    const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    TraceStack st{lds_stack + threadIdx.x, spill + gid, gridDim.x * blockDim.x, 0};
    uint32_t nv = 0, tv = 0;
    trace_queue<true, true>(job, QueueSegs{nullptr, n, n, 1u, reinterpret_cast<uint32_t*>(diag + 2)}, nodes, tris, st, nv, tv, node_tile,
                             min(num_nodes, kBvhTileNodes));
    report_stack(diag, st, true);
}

// ---------------------------------------------------------------------------------------------------------
// Device scratch of pe_intersect_batch_noalloc: its traversal diagnostics (the first 256 bytes: deepest stack,
// overflows, chunk counter) and the traversal-stack spill of a grid of at most kNoAllocBlocks blocks. One copy
// per GPU, held until the process ends. Host code reaches it only through the address hipGetSymbolAddress gave
// for the device (pe_device::noalloc_scratch), never through the symbol's name.
constexpr uint32_t kNoAllocBlocks = 256;
constexpr size_t kNoAllocDiagBytes = 256;
constexpr size_t kNoAllocSpillBytes = size_t{kNoAllocBlocks} * kTraceBlock * (kStackDepth - kLdsDepth) * sizeof(uint32_t);
constexpr size_t kScratchBytes = kNoAllocDiagBytes + kNoAllocSpillBytes;  // 29,360,384 bytes
static_assert(kDiagBytes <= kNoAllocDiagBytes, "diagnostics fit the scratch's head");
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// Clears the diagnostics counters (a kernel, not a memset node, so that a caller may capture the call).
__global__ void k_clear_diag(unsigned long long* __restrict__ diag) {
    if (threadIdx.x < kDiagBytes / sizeof(unsigned long long)) diag[threadIdx.x] = 0ull;
}

// Persistent grid for 256-lane blocks and the spill buffer sized for it.
int ensure_trace_resources(pe_device* dev, int& blocks) {
    blocks = dev->max_grid_blocks;
    if (!dev->diag.get() && !dev->diag.alloc(kDiagBytes)) return PE_ERR_OOM;
    const uint32_t lanes = static_cast<uint32_t>(blocks) * kTraceBlock;
    if (dev->spill_lanes < lanes) {
        if (!dev->spill.alloc(static_cast<size_t>(lanes) * (kStackDepth - kLdsDepth) * sizeof(uint32_t))) return PE_ERR_OOM;
        dev->spill_lanes = lanes;
    }
    return PE_OK;
}

size_t trace_spill_bytes(int grid_blocks) {
    return static_cast<size_t>(grid_blocks) * kTraceBlock * (kStackDepth - kLdsDepth) * sizeof(uint32_t);
}

// Copies a batch's traversal diagnostics to the device (after the stream's other copies) and fails the batch
// when a stack overflowed.
int finish_batch(pe_device* dev, const char* what) {
    static_assert(sizeof(dev->diag_host) == kDiagBytes, "diagnostics copy size");
    PE_HIP_CHECK(hipMemcpyAsync(dev->diag_host, dev->diag.get(), kDiagBytes, hipMemcpyDeviceToHost, dev->stream));
    PE_HIP_CHECK(hipStreamSynchronize(dev->stream));
    if (dev->diag_host[1] != 0) {
        set_error(std::string(what) + ": " + std::to_string(dev->diag_host[1]) + " traversals overflowed the " +
                  std::to_string(kStackDepth) + "-entry stack");
        return PE_ERR_INVALID;
    }
    return PE_OK;
}

int grid_for(pe_device* dev, size_t n) {
    const size_t want = (n + 255) / 256;
    return static_cast<int>(std::max<size_t>(1, std::min<size_t>(want, static_cast<size_t>(dev->max_grid_blocks))));
}

}  // namespace pe

namespace pe {

int scene_park(pe_scene& sc) {
    if (!sc.resident) return PE_OK;
    PE_HIP_CHECK(hipSetDevice(sc.dev->ordinal));
    std::vector<DevBuf*> bufs = sc.buffers();
    std::vector<std::vector<unsigned char>> host(bufs.size());
    for (size_t k = 0; k < bufs.size(); ++k) {
        host[k].resize(bufs[k]->bytes());
        if (!host[k].empty())
            PE_HIP_CHECK(hipMemcpyAsync(host[k].data(), bufs[k]->get(), host[k].size(), hipMemcpyDeviceToHost, sc.dev->stream));
    }
    PE_HIP_CHECK(hipStreamSynchronize(sc.dev->stream));
    for (DevBuf* b : bufs) b->release();
    sc.parked = std::move(host);
    sc.resident = false;
    return PE_OK;
}

// The view's device pointers, from the scene's buffers (its other fields do not depend on where they live).
static void rebind_view(pe_scene& sc) {
    SceneView& v = sc.view;
    v.nodes = sc.nodes.as<BvhNode4>();
    v.num_nodes = static_cast<uint32_t>(sc.num_nodes);
    v.tris = sc.tris.as<TriAccel>();
    v.shade = sc.tri_shade.as<TriShade>();
    v.pos = sc.positions.as<float4>();
    v.nrm = sc.normals.as<float4>();
    v.uv = sc.uvs.as<float2>();
    v.materials = sc.materials.as<GpuMaterial>();
    v.lights = sc.lights.as<GpuLight>();
    v.tables = sc.tables.as<EnergyTables>();
    v.light_nodes = sc.light_nodes.as<GpuLightNode>();
    v.light_paths = sc.light_paths.as<LightPath>();
    v.node_uids = sc.node_uids.as<uint64_t>();
    v.link_info = sc.link_info.as<uint32_t>();
    v.link_uids = sc.link_uids.as<uint64_t>();
    v.env.rgb = sc.env_rgb.as<float>();
    v.env.func = sc.env_func.as<float>();
    v.env.marginal = sc.env_marginal.as<float>();
    v.env.cond = sc.env_cond.as<float>();
    v.env.marginal_guide = sc.env_mguide.as<uint32_t>();
    v.env.cond_guide = sc.env_cguide.as<uint32_t>();
}

int scene_restore(pe_scene& sc) {
    if (sc.resident) return PE_OK;
    PE_HIP_CHECK(hipSetDevice(sc.dev->ordinal));
    std::vector<DevBuf*> bufs = sc.buffers();
    for (size_t k = 0; k < bufs.size(); ++k) {
        if (!bufs[k]->upload(sc.parked[k].data(), sc.parked[k].size(), sc.dev->stream)) {
            for (size_t j = 0; j <= k; ++j) bufs[j]->release();
            return PE_ERR_OOM;
        }
    }
    PE_HIP_CHECK(hipStreamSynchronize(sc.dev->stream));
    rebind_view(sc);
    sc.parked.clear();
    sc.resident = true;
    return PE_OK;
}

void scene_swap_contents(pe_scene& a, pe_scene& b) {
    std::vector<DevBuf*> ba = a.buffers(), bb = b.buffers();
    for (size_t k = 0; k < ba.size(); ++k) std::swap(*ba[k], *bb[k]);
    std::swap(a.num_tris, b.num_tris);
    std::swap(a.num_verts, b.num_verts);
    std::swap(a.num_shapes, b.num_shapes);
    std::swap(a.bvh_depth, b.bvh_depth);
    std::swap(a.view, b.view);
    std::swap(a.num_nodes, b.num_nodes);
    std::swap(a.max_material_ref, b.max_material_ref);
    std::swap(a.max_emitter_ref, b.max_emitter_ref);
    std::swap(a.broken, b.broken);
}

}  // namespace pe

using namespace pe;

#define PE_GUARD_BEGIN try {
#define PE_GUARD_END                                                  \
    }                                                                 \
    catch (const std::bad_alloc&) {                                   \
        set_error("host allocation failed");                          \
        return PE_ERR_OOM;                                            \
    }                                                                 \
    catch (const std::exception& e) {                                 \
        set_error(std::string("internal error: ") + e.what());        \
        return PE_ERR_INVALID;                                        \
    }

extern "C" {

const char* pe_last_error(void) { return g_last_error.c_str(); }

int pe_abi_version(void) { return PE_ABI_VERSION; }

int pe_device_count(int* count) {
    if (!count) {
        set_error("pe_device_count: count is NULL");
        return PE_ERR_INVALID;
    }
    int n = 0;
    const hipError_t e = hipGetDeviceCount(&n);
    if (e == hipErrorNoDevice || e == hipErrorInvalidDevice) {
        *count = 0;
        return PE_OK;
    }
    if (e != hipSuccess) {
        *count = 0;
        set_error(std::string("hipGetDeviceCount failed: ") + hipGetErrorString(e));
        return PE_OK;  // no usable device is not an error of the registry (enumerateDevices lists the CPU only)
    }
    *count = n;
    return PE_OK;
}

int pe_device_info_get(int ordinal, pe_device_info* out) {
    if (!out) {
        set_error("pe_device_info_get: out is NULL");
        return PE_ERR_INVALID;
    }
    int n = 0;
    pe_device_count(&n);
    if (ordinal < 0 || ordinal >= n) {
        set_error("hip:" + std::to_string(ordinal) + ": no such device (" + std::to_string(n) + " visible)");
        return PE_ERR_NO_DEVICE;
    }
    hipDeviceProp_t p;
    PE_HIP_CHECK(hipGetDeviceProperties(&p, ordinal));
    std::memset(out, 0, sizeof(*out));
    std::snprintf(out->id, sizeof(out->id), "hip:%d", ordinal);
    std::snprintf(out->name, sizeof(out->name), "%s", p.name);
    std::snprintf(out->arch, sizeof(out->arch), "%s", p.gcnArchName);
    out->kind = PE_DEVICE_HIP;
    out->compute_units = p.multiProcessorCount;
    out->memory_bytes = p.totalGlobalMem;
    const bool gfx950 = std::strncmp(p.gcnArchName, "gfx950", 6) == 0;
    out->usable = gfx950 ? 1 : 0;
    std::snprintf(out->note, sizeof(out->note), "%s", gfx950 ? "wavefront path tracer (gfx950 code object)"
                                                            : "no gfx950 code object for this architecture");
    return PE_OK;
}

int pe_device_create(int ordinal, pe_device** out) {
    PE_GUARD_BEGIN
    if (!out) {
        set_error("pe_device_create: out is NULL");
        return PE_ERR_INVALID;
    }
    *out = nullptr;
    pe_device_info info;
    int rc = pe_device_info_get(ordinal, &info);
    if (rc != PE_OK) return rc;
    if (!info.usable) {
        set_error(std::string(info.id) + ": " + info.note);
        return PE_ERR_NO_DEVICE;
    }
    PE_HIP_CHECK(hipSetDevice(ordinal));
    auto dev = std::make_unique<pe_device>();
    dev->ordinal = ordinal;
    dev->arch = info.arch;
    dev->cu_count = info.compute_units;
    dev->max_grid_blocks = std::max(1, dev->cu_count) * 8;
    PE_HIP_CHECK(hipStreamCreateWithFlags(&dev->stream, hipStreamNonBlocking));
    // This device's copy of the no-allocation scratch (with the device current).
    void* scratch_ptr = nullptr;
    const hipError_t se = hipGetSymbolAddress(&scratch_ptr, HIP_SYMBOL(scratch));
    if (se != hipSuccess || !scratch_ptr) {
        (void)hipStreamDestroy(dev->stream);
        set_error(std::string("pe_device_create: hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(se));
        return PE_ERR_HIP;
    }
    dev->noalloc_scratch = scratch_ptr;
    *out = dev.release();
    return PE_OK;
    PE_GUARD_END
}

int pe_device_destroy(pe_device* dev) {
    if (!dev) return PE_OK;
    (void)hipSetDevice(dev->ordinal);
    (void)hipStreamSynchronize(dev->stream);
    wavefront_release(dev);
    dev->spill.release();
    dev->scratch.release();
    (void)hipStreamDestroy(dev->stream);
    delete dev;
    return PE_OK;
}

int pe_device_features_get(const pe_device* dev, pe_device_features* out) {
    if (!dev || !out) {
        set_error("pe_device_features_get: NULL argument");
        return PE_ERR_INVALID;
    }
    out->aovs = 0;
    out->light_passes = 0;
    out->adaptive_sampling = 1;  // pe_settings.adaptive (renderer.cpp:438-505)
    return PE_OK;
}

int pe_device_synchronize(pe_device* dev) {
    if (!dev) {
        set_error("pe_device_synchronize: NULL device");
        return PE_ERR_INVALID;
    }
    PE_HIP_CHECK(hipStreamSynchronize(dev->stream));
    return PE_OK;
}

int pe_scene_upload(pe_device* dev, const pe_scene_desc* desc, pe_scene** out) {
    PE_GUARD_BEGIN
    if (!dev || !desc || !out) {
        set_error("pe_scene_upload: NULL argument");
        return PE_ERR_INVALID;
    }
    *out = nullptr;
    PE_HIP_CHECK(hipSetDevice(dev->ordinal));
    auto sc = std::make_unique<pe_scene>();
    sc->dev = dev;
    const int rc = upload_scene(dev, *desc, *sc);
    if (rc != PE_OK) return rc;
    PE_HIP_CHECK(hipStreamSynchronize(dev->stream));
    *out = sc.release();
    return PE_OK;
    PE_GUARD_END
}

// RenderDevice::updateScene (render_device.h:77-79) and the Vulkan device's partial rebuilds
// (rebuildMaterials / rebuildLights): the BVH and the triangle data stay resident; only the material table or
// the lights, environment and light-selection structure are rebuilt from `desc`, the full updated description.
// Full and Transforms rebuild everything into a new set of buffers that replaces the old one on success.
int pe_scene_update(pe_scene* scene, int kind, const pe_scene_desc* desc) {
    PE_GUARD_BEGIN
    if (!scene || !desc) {
        set_error("pe_scene_update: NULL argument");
        return PE_ERR_INVALID;
    }
    pe_device* dev = scene->dev;
    PE_HIP_CHECK(hipSetDevice(dev->ordinal));
    // Renders in progress on any stream (pe_render, server lanes) finish with the old data first; none starts
    // until the update is done.
    std::unique_lock<std::shared_mutex> lk(scene->use);
    PE_HIP_CHECK(hipStreamSynchronize(dev->stream));
    if (kind == PE_UPDATE_FULL || kind == PE_UPDATE_TRANSFORMS) {
        auto fresh = std::make_unique<pe_scene>();
        fresh->dev = dev;
        const int rc = upload_scene(dev, *desc, *fresh);
        if (rc != PE_OK) return rc;  // the old scene is untouched
        PE_HIP_CHECK(hipStreamSynchronize(dev->stream));
        scene_swap_contents(*scene, *fresh);
        scene->resident = true;
        scene->parked.clear();
        return PE_OK;
    }
    if (!scene->resident) {  // a parked scene comes back before a partial rebuild
        const int rc = scene_restore(*scene);
        if (rc != PE_OK) return rc;
    }
    if (kind != PE_UPDATE_MATERIALS && kind != PE_UPDATE_LIGHTS) {
        set_error("pe_scene_update: unknown update kind " + std::to_string(kind));
        return PE_ERR_INVALID;
    }
    if (scene->broken) {
        set_error("pe_scene_update: an earlier update failed; upload the scene again");
        return PE_ERR_INVALID;
    }
    int rc = PE_OK;
    if (kind == PE_UPDATE_MATERIALS) {
        if (desc->num_materials <= scene->max_material_ref) {
            set_error("pe_scene_update: a mesh uses material " + std::to_string(scene->max_material_ref) +
                      " but the update has " + std::to_string(desc->num_materials));
            return PE_ERR_INVALID;
        }
        rc = prepare_materials(dev, *desc, *scene);
    } else {
        if (desc->num_lights <= scene->max_emitter_ref) {
            set_error("pe_scene_update: a mesh emits as light " + std::to_string(scene->max_emitter_ref) +
                      " but the update has " + std::to_string(desc->num_lights));
            return PE_ERR_INVALID;
        }
        rc = prepare_lights(dev, *desc, *scene);
    }
    if (rc != PE_OK) scene->broken = true;
    return rc;
    PE_GUARD_END
}

int pe_scene_release(pe_scene* scene) {
    if (!scene) return PE_OK;
    if (const int refs = scene->job_refs.load(); refs > 0) {
        set_error("pe_scene_release: " + std::to_string(refs) + " server job(s) still reference the scene; release them first");
        return PE_ERR_INVALID;
    }
    std::vector<pe_server*> trackers;
    {
        std::lock_guard<std::mutex> g(scene->trackers_mu);
        trackers.swap(scene->trackers);
    }
    for (pe_server* s : trackers) server_forget_scene(s, scene);
    {
        std::unique_lock<std::shared_mutex> lk(scene->use);  // no render of it is in progress
        (void)hipSetDevice(scene->dev->ordinal);
        (void)hipStreamSynchronize(scene->dev->stream);
    }
    delete scene;
    return PE_OK;
}

int pe_scene_stats(const pe_scene* scene, uint64_t* triangles, uint64_t* bvh_nodes, uint64_t* device_bytes) {
    if (!scene) {
        set_error("pe_scene_stats: NULL scene");
        return PE_ERR_INVALID;
    }
    if (triangles) *triangles = scene->num_tris;
    if (bvh_nodes) *bvh_nodes = scene->num_nodes;
    if (device_bytes) *device_bytes = scene->device_bytes();
    return PE_OK;
}

int pe_trace_diag(pe_device* dev, uint32_t* max_stack_depth, uint64_t* overflows) {
    if (!dev) {
        set_error("pe_trace_diag: NULL device");
        return PE_ERR_INVALID;
    }
    if (max_stack_depth) *max_stack_depth = static_cast<uint32_t>(dev->diag_host[0]);
    if (overflows) *overflows = dev->diag_host[1];
    return PE_OK;
}

int pe_sobol_batch(pe_device* dev, uint64_t seed, size_t n, const int32_t* px, const int32_t* py,
                   const uint32_t* index, const uint32_t* dim, uint32_t* out) {
    PE_GUARD_BEGIN
    if (!dev || (n > 0 && (!px || !py || !index || !dim || !out))) {
        set_error("pe_sobol_batch: NULL argument");
        return PE_ERR_INVALID;
    }
    if (n == 0) return PE_OK;
    PE_HIP_CHECK(hipSetDevice(dev->ordinal));
    DevBuf dpx, dpy, didx, ddim, dout;
    if (!dpx.upload(px, n * 4, dev->stream) || !dpy.upload(py, n * 4, dev->stream) ||
        !didx.upload(index, n * 4, dev->stream) || !ddim.upload(dim, n * 4, dev->stream) || !dout.alloc(n * 8))
        return PE_ERR_OOM;
    hipLaunchKernelGGL(k_sobol_batch, dim3(grid_for(dev, n)), dim3(256), 0, dev->stream, seed, n, dpx.as<int32_t>(),
                       dpy.as<int32_t>(), didx.as<uint32_t>(), ddim.as<uint32_t>(), dout.as<uint2>());
    PE_HIP_CHECK(hipGetLastError());
    PE_HIP_CHECK(hipMemcpyAsync(out, dout.get(), n * 8, hipMemcpyDeviceToHost, dev->stream));
    PE_HIP_CHECK(hipStreamSynchronize(dev->stream));
    return PE_OK;
    PE_GUARD_END
}

int pe_intersect_batch(pe_device* dev, pe_scene* scene, size_t n, const float* rays, int32_t* hits, float* bary) {
    PE_GUARD_BEGIN
    if (!dev || !scene || scene->dev != dev || (n > 0 && (!rays || !hits))) {
        set_error("pe_intersect_batch: NULL argument or scene of another device");
        return PE_ERR_INVALID;
    }
    if (n == 0) return PE_OK;
    if (n >= (size_t{1} << 31)) {
        set_error("pe_intersect_batch: at most 2^31 - 1 rays per batch");
        return PE_ERR_INVALID;
    }
    PE_HIP_CHECK(hipSetDevice(dev->ordinal));
    int blocks = 0;
    int rc = ensure_trace_resources(dev, blocks);
    if (rc != PE_OK) return rc;
    blocks = grid_for(dev, n);
    DevBuf drays, dhits, dbary;
    if (!drays.upload(rays, n * 32, dev->stream) || !dhits.alloc(n * 16) || (bary && !dbary.alloc(n * 8)))
        return PE_ERR_OOM;
    BatchClosest job{};
    job.r = drays.as<float4>();
    job.shade = scene->tri_shade.as<TriShade>();
    job.hits = dhits.as<int4>();
    job.bary = bary ? dbary.as<float2>() : nullptr;
    PE_HIP_CHECK(hipMemsetAsync(dev->diag.get(), 0, kDiagBytes, dev->stream));
    hipLaunchKernelGGL(k_intersect_batch, dim3(blocks), dim3(kTraceBlock), 0, dev->stream, scene->nodes.as<BvhNode4>(),
                       scene->tris.as<TriAccel>(), job, static_cast<uint32_t>(n), scene->view.num_nodes, dev->spill.as<uint32_t>(),
                       dev->diag.as<unsigned long long>());
    PE_HIP_CHECK(hipGetLastError());
    PE_HIP_CHECK(hipMemcpyAsync(hits, dhits.get(), n * 16, hipMemcpyDeviceToHost, dev->stream));
    if (bary) PE_HIP_CHECK(hipMemcpyAsync(bary, dbary.get(), n * 8, hipMemcpyDeviceToHost, dev->stream));
    return finish_batch(dev, "pe_intersect_batch");
    PE_GUARD_END
}

// pe_intersect_batch on caller-owned device buffers, with no device allocation: the stack spill and the
// counters come from this device's copy of `scratch`, and the grid is capped so that the spill fits it. Calls
// that overlap on one device share that scratch and race.
int pe_intersect_batch_noalloc(pe_device* dev, pe_scene* scene, size_t n, const float* rays_dev, int32_t* hits_dev,
                               float* bary_dev) {
    PE_GUARD_BEGIN
    if (!dev || !scene || scene->dev != dev || (n > 0 && (!rays_dev || !hits_dev))) {
        set_error("pe_intersect_batch_noalloc: NULL argument or scene of another device");
        return PE_ERR_INVALID;
    }
    if (n == 0) return PE_OK;
    if (n >= (size_t{1} << 31)) {
        set_error("pe_intersect_batch_noalloc: at most 2^31 - 1 rays per batch");
        return PE_ERR_INVALID;
    }
    if (!dev->noalloc_scratch) {
        set_error("pe_intersect_batch_noalloc: the device has no scratch address");
        return PE_ERR_INVALID;
    }
    const int blocks = std::min(grid_for(dev, n), static_cast<int>(kNoAllocBlocks));
    const size_t need = kNoAllocDiagBytes + static_cast<size_t>(blocks) * kTraceBlock * (kStackDepth - kLdsDepth) * sizeof(uint32_t);
    if (need > kScratchBytes) {
        set_error("pe_intersect_batch_noalloc: the traversal needs " + std::to_string(need) + " bytes of scratch, above the " +
                  std::to_string(kScratchBytes) + "-byte maximum");
        return PE_ERR_INVALID;
    }
    PE_HIP_CHECK(hipSetDevice(dev->ordinal));
    unsigned char* base = static_cast<unsigned char*>(dev->noalloc_scratch);
    unsigned long long* diag = reinterpret_cast<unsigned long long*>(base);
    uint32_t* spill = reinterpret_cast<uint32_t*>(base + kNoAllocDiagBytes);
    hipLaunchKernelGGL(k_clear_diag, dim3(1), dim3(64), 0, dev->stream, diag);
    BatchClosest job{};
    job.r = reinterpret_cast<const float4*>(rays_dev);
    job.shade = scene->tri_shade.as<TriShade>();
    job.hits = reinterpret_cast<int4*>(hits_dev);
    job.bary = reinterpret_cast<float2*>(bary_dev);
    hipLaunchKernelGGL(k_intersect_batch, dim3(blocks), dim3(kTraceBlock), 0, dev->stream, scene->nodes.as<BvhNode4>(),
                       scene->tris.as<TriAccel>(), job, static_cast<uint32_t>(n), scene->view.num_nodes, spill, diag);
    PE_HIP_CHECK(hipGetLastError());
    PE_HIP_CHECK(hipMemcpyAsync(dev->diag_host, diag, kDiagBytes, hipMemcpyDeviceToHost, dev->stream));
    PE_HIP_CHECK(hipStreamSynchronize(dev->stream));
    if (dev->diag_host[1] != 0) {
        set_error("pe_intersect_batch_noalloc: " + std::to_string(dev->diag_host[1]) + " traversals overflowed the " +
                  std::to_string(kStackDepth) + "-entry stack");
        return PE_ERR_INVALID;
    }
    return PE_OK;
    PE_GUARD_END
}

int pe_occluded_batch(pe_device* dev, pe_scene* scene, size_t n, const float* rays, uint8_t* occluded) {
    PE_GUARD_BEGIN
    if (!dev || !scene || scene->dev != dev || (n > 0 && (!rays || !occluded))) {
        set_error("pe_occluded_batch: NULL argument or scene of another device");
        return PE_ERR_INVALID;
    }
    if (n == 0) return PE_OK;
    if (n >= (size_t{1} << 31)) {
        set_error("pe_occluded_batch: at most 2^31 - 1 rays per batch");
        return PE_ERR_INVALID;
    }
    PE_HIP_CHECK(hipSetDevice(dev->ordinal));
    int blocks = 0;
    int rc = ensure_trace_resources(dev, blocks);
    if (rc != PE_OK) return rc;
    blocks = grid_for(dev, n);
    DevBuf drays, docc;
    if (!drays.upload(rays, n * 32, dev->stream) || !docc.alloc(n)) return PE_ERR_OOM;
    BatchAny job{};
    job.r = drays.as<float4>();
    job.occ = docc.as<uint8_t>();
    PE_HIP_CHECK(hipMemsetAsync(dev->diag.get(), 0, kDiagBytes, dev->stream));
    hipLaunchKernelGGL(k_occluded_batch, dim3(blocks), dim3(kTraceBlock), 0, dev->stream, scene->nodes.as<BvhNode4>(),
                       scene->tris.as<TriAccel>(), job, static_cast<uint32_t>(n), scene->view.num_nodes, dev->spill.as<uint32_t>(),
                       dev->diag.as<unsigned long long>());
    PE_HIP_CHECK(hipGetLastError());
    PE_HIP_CHECK(hipMemcpyAsync(occluded, docc.get(), n, hipMemcpyDeviceToHost, dev->stream));
    return finish_batch(dev, "pe_occluded_batch");
    PE_GUARD_END
}

}  // extern "C"

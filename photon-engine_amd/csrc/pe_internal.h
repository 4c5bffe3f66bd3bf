// pe_internal.h — host-side objects behind the opaque C-ABI handles, and error plumbing.
#pragma once

#include "photon_hip.h"
#include "pe_device_scene.h"
#include "pe_gpu_types.h"

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <memory>
#include <mutex>
#include <shared_mutex>
#include <string>
#include <vector>

namespace pe {

void set_error(const std::string& msg);

// Owning device allocation. Device memory is never freed while a kernel may still read it: owners
// synchronise their stream before releasing.
class DevBuf {
public:
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : m_ptr(o.m_ptr), m_bytes(o.m_bytes) {
        o.m_ptr = nullptr;
        o.m_bytes = 0;
    }
    DevBuf& operator=(DevBuf&& o) noexcept {
        if (this != &o) {
            release();
            m_ptr = o.m_ptr;
            m_bytes = o.m_bytes;
            o.m_ptr = nullptr;
            o.m_bytes = 0;
        }
        return *this;
    }
    ~DevBuf() { release(); }

    bool alloc(size_t bytes);  // false (and error set) on failure; 0 bytes gives a null pointer
    bool upload(const void* src, size_t bytes, hipStream_t s);
    void release();
    void* get() const { return m_ptr; }
    template <class T>
    T* as() const {
        return static_cast<T*>(m_ptr);
    }
    size_t bytes() const { return m_bytes; }

private:
    void* m_ptr = nullptr;
    size_t m_bytes = 0;
};

}  // namespace pe

struct pe_device {
    int ordinal = 0;
    hipStream_t stream = nullptr;
    int cu_count = 0;
    std::string arch;
    pe::DevBuf spill;            // traversal stack spill, (kStackDepth - kLdsDepth) entries per lane
    uint32_t spill_lanes = 0;    // lanes the spill buffer is sized for
    pe::DevBuf scratch;          // small per-call staging (counters)
    pe::DevBuf diag;             // traversal diagnostics of the batch entry points: deepest stack, overflows
    unsigned long long diag_host[4] = {0, 0, 0, 0};  // ... as copied back after the last batch
    void* noalloc_scratch = nullptr;  // this device's address of the no-allocation scratch (capi.hip)
    int max_grid_blocks = 0;     // persistent-grid size for 256-lane blocks
    void* wavefront = nullptr;   // integrator buffers (integrator.hip), grown on demand
};

struct pe_scene {
    pe_device* dev = nullptr;
    // Use by renders and servers (not part of the scene's contents: kept across Full / Transforms updates).
    std::shared_mutex use;                 // renders hold it shared; updates, parking and release exclusively
    std::atomic<int> job_refs{0};          // server jobs that reference the scene and are not released
    std::mutex trackers_mu;
    std::vector<pe_server*> trackers;      // servers that account for the scene's residency
    bool resident = true;                  // device buffers present (false: parked in host memory)
    std::vector<std::vector<unsigned char>> parked;  // the device buffers' bytes while parked
    uint64_t num_tris = 0;
    uint64_t num_verts = 0;
    int num_shapes = 0;
    int bvh_depth = 0;
    pe::DevBuf nodes, tris, tri_shade, positions, normals, uvs;
    pe::DevBuf materials, lights, tables, env_rgb, env_func, env_marginal, env_cond, env_mguide, env_cguide,
        light_nodes, light_paths, node_uids, link_info, link_uids;
    pe::SceneView view{};
    uint64_t num_nodes = 0;
    int max_material_ref = -1;  // largest material index a mesh uses
    int max_emitter_ref = -1;   // largest light index a mesh's emitter range reaches
    bool broken = false;        // an update failed part-way: the scene must be uploaded again
    // The device buffers, in a fixed order (parking and restoring walk them).
    std::vector<pe::DevBuf*> buffers() {
        return {&nodes, &tris, &tri_shade, &positions, &normals, &uvs, &materials, &lights, &tables, &env_rgb,
                &env_func, &env_marginal, &env_cond, &env_mguide, &env_cguide, &light_nodes, &light_paths, &node_uids, &link_info, &link_uids};
    }
    uint64_t device_bytes() const {
        return nodes.bytes() + tris.bytes() + tri_shade.bytes() + positions.bytes() + normals.bytes() + uvs.bytes() +
               materials.bytes() + lights.bytes() + tables.bytes() + env_rgb.bytes() + env_func.bytes() +
               env_marginal.bytes() + env_cond.bytes() + env_mguide.bytes() +
               env_cguide.bytes() + light_nodes.bytes() + light_paths.bytes() + node_uids.bytes() +
               link_info.bytes() + link_uids.bytes();
    }
};

// A device CryptoFilm: PE_CRYPTO_SLOTS (id, count) entries and a sample total per pixel.
struct pe_crypto_film {
    pe_device* dev = nullptr;
    int width = 0, height = 0;
    pe::DevBuf ids, counts, totals, dropped;
};

struct pe_server;

struct pe_film {
    pe_device* dev = nullptr;
    int width = 0, height = 0;
    pe::DevBuf sum, count, lum2, rejected;
};

namespace pe {
// Scene residency (capi.hip): parking copies a scene's device buffers to host memory and frees them; restoring
// uploads them again and points the scene's view at the new buffers. Callers hold scene.use exclusively.
int scene_park(pe_scene& sc);
int scene_restore(pe_scene& sc);
// Device data of a + b exchanged (a Full / Transforms update swaps in a freshly built scene).
void scene_swap_contents(pe_scene& a, pe_scene& b);
// A server stops accounting for a scene (serve.hip; called when the scene is released).
void server_forget_scene(pe_server* s, pe_scene* sc);
constexpr size_t kDiagBytes = 32;  // deepest stack, overflows, the batch queue's chunk counter
// Persistent grid size and the traversal-stack spill buffer sized for it (capi.hip).
int ensure_trace_resources(pe_device* dev, int& blocks);
// Releases the integrator's buffers of a device (integrator.hip).
void wavefront_release(pe_device* dev);

// Path-state buffers of one render stream (integrator.hip).
struct Wavefront;
Wavefront* wavefront_new();
void wavefront_delete(Wavefront* wf);
// A pool of `paths` path ids and a result buffer of `results` samples.
int wavefront_reserve(Wavefront* wf, uint32_t paths, uint64_t results, int shadow_quality);
uint64_t wavefront_bytes_per_path(int shadow_quality);
uint64_t wavefront_device_bytes(const Wavefront* wf);

// One render stream: a HIP stream with its own path state and traversal-stack spill (grid_blocks x 256 lanes).
struct RenderLane {
    hipStream_t stream;
    Wavefront* wf;
    uint32_t* spill;
    int grid_blocks;
};
// AOV films of a render (any may be null), and a cancellation callback polled with the cancel flag.
struct AovTargets {
    pe_film* albedo = nullptr;
    pe_film* normal = nullptr;
    pe_film* alpha = nullptr;
    int (*cancel_poll)(void*) = nullptr;
    void* cancel_user = nullptr;
    const pe_pass_targets* passes = nullptr;  // render passes (RenderPassFilms)
};
int render_samples(pe_device* dev, const RenderLane& lane, pe_scene* scene, const pe_camera* camera,
                   const pe_settings* settings, pe_film* film, int first_sample, int sample_count,
                   const volatile int32_t* cancel, pe_batch_result* result, pe_guide* guide = nullptr,
                   const AovTargets* aovs = nullptr);
// Bytes of traversal-stack spill a lane of `grid_blocks` blocks needs (capi.hip).
size_t trace_spill_bytes(int grid_blocks);
}  // namespace pe

#define PE_HIP_CHECK(call)                                                                              \
    do {                                                                                                \
        hipError_t _e = (call);                                                                         \
        if (_e != hipSuccess) {                                                                         \
            pe::set_error(std::string(#call) + " failed: " + hipGetErrorString(_e));                    \
            return PE_ERR_HIP;                                                                          \
        }                                                                                               \
    } while (0)

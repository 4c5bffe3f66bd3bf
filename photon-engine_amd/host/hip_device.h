// hip_device.h — RenderDevice backed by libphoton_hip.so (the gfx950 wavefront path tracer).
//
// Written against PhotonEngine's own engine headers: it is the file a maintainer adds as
// src/engine/hip_device.{h,cpp}, registered by enumerateDevices()/createDevice() under the id "hip:N"
// (engine/render_device.cpp:80-157). It packs the immutable Scene into a pe_scene_desc once per setScene and
// implements render() as pe_render into a device film whose sums are merged into the caller's Film with
// Film::addBatch, the way the Vulkan device merges its batch (vulkan_device.cpp:2152-2208).
#pragma once

#include "engine/render_device.h"
#include "photon_hip.h"

#include <memory>
#include <string>
#include <vector>

namespace photon {

class Material;

class HipDevice final : public RenderDevice {
public:
    // nullptr and *error set when the ordinal has no usable gfx950 device (createDevice contract).
    static std::unique_ptr<HipDevice> create(int ordinal, std::string* error);
    ~HipDevice() override;

    const DeviceInfo& info() const override { return m_info; }
    // AOVs (albedo / normal / alpha), render passes (light passes, light groups, data passes, cryptomatte) and
    // adaptive sampling.
    DeviceFeatures features() const override { return DeviceFeatures{true, true, true}; }
    void setScene(std::shared_ptr<const Scene> scene) override;
    void updateScene(std::shared_ptr<const Scene> scene, SceneUpdateKind kind) override;
    bool hasScene() const override { return m_scene != nullptr; }
    SampleBatchResult render(const Camera& camera, const RenderSettings& settings, RenderTargets& targets,
                             int firstSample, int count, const std::atomic<bool>* cancel) override;

    // Why the last setScene left the device without a scene (unsupported feature), empty otherwise.
    const std::string& lastError() const { return m_error; }
    // Rays the last render() traced (extension + shadow), for Mrays/s reporting.
    uint64_t lastRays() const { return m_lastRays; }

private:
    HipDevice() = default;

    DeviceInfo m_info;
    pe_device* m_dev = nullptr;
    pe_scene* m_scene = nullptr;
    pe_guide* m_guide = nullptr;          // RenderSettings::pathGuiding: the learned guide of m_guideScene
    const pe_scene* m_guideScene = nullptr;
    // Device films: beauty, then the albedo, normal and alpha AOVs, then the pass films (8 light passes, 16 light
    // groups, depth, position, object id, material id). With adaptive sampling they keep every sample
    // this device rendered since the last render that started at sample 0 (the beauty's per-sample statistics
    // drive the skip mask and the stops); prevSum / prevCount is the part already merged into the caller's films.
    struct DeviceFilm {
        pe_film* film = nullptr;
        std::vector<float> prevSum;
        std::vector<uint32_t> prevCount;
    };
    static constexpr int kPassFilm0 = 4, kDeviceFilms = kPassFilm0 + 28;
    DeviceFilm m_films[kDeviceFilms];
    // Device CryptoFilms (objects, materials) and the per-slot counts and totals already merged.
    struct DeviceCrypto {
        pe_crypto_film* film = nullptr;
        std::vector<uint32_t> prevCounts, prevTotals;
    };
    DeviceCrypto m_crypto[2];
    std::vector<const Material*> m_materialPtrs;  // per packed material: the scene's material (cryptomatte ids)
    int m_filmW = 0, m_filmH = 0;
    bool m_accumulating = false;
    std::shared_ptr<const Scene> m_sceneRef;
    std::string m_error;
    uint64_t m_lastRays = 0;
};

// The usable and unusable HIP devices, in the DeviceInfo form enumerateDevices() lists.
std::vector<DeviceInfo> enumerateHipDevices();

}  // namespace photon

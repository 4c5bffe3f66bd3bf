// hipdev_driver.cpp — test-only harness that drives photon-engine_amd/host/hip_device.cpp (the RenderDevice
// adapter over libphoton_hip.so) through the reference's own RenderSession (engine/render_session.cpp).
//
// Part of the oracle (test infrastructure): compiled with the reference sources and the adapter into
// oracle/_ref/libphoton_hipdevice.so (oracle/Makefile.ref). The scene is the reference Scene that
// ref_scene_build (ref_driver.cpp) makes from a pe_scene_desc, so the adapter packs a genuine reference scene.
#include "core/image/cryptomatte.h"
#include "engine/render_session.h"
#include "camera/perspective_camera.h"
#include "hip_device.h"

#include <atomic>
#include <chrono>
#include <cstring>
#include <memory>
#include <string>
#include <thread>

using namespace photon;

std::shared_ptr<const Scene> ref_scene_shared(void* h);
void ref_copy_passes(const RenderPassFilms& pf, int width, int height, float* out);

#define HIPDEV_API __attribute__((visibility("default")))

// The camera the reference driver (ref_driver.cpp, linked into the same library) builds: ref_camera_select's
// choice of PerspectiveCamera, ThinLensCamera or OrthographicCamera with its shutter end, and its motion-blur flag.
std::unique_ptr<Camera> makeCamera(const float* pos, const float* target, const float* up, float fov, float aspect);
bool selectedCameraMotionBlur();

namespace {
void copy_error(const std::string& msg, char* err, int errlen) {
    if (!err || errlen <= 0) return;
    std::strncpy(err, msg.c_str(), static_cast<size_t>(errlen) - 1);
    err[errlen - 1] = '\0';
}
}  // namespace

extern "C" {

// RenderSession(Final, targetSpp = spp) on a HipDevice for the reference scene `h`; writes the resolved mean
// (3 floats/px), the sample counts and the number of session batches. aoStrength > 0 asks for a setting the
// adapter refuses. Returns 0 on success, 1 when the device
// refused the scene or a batch (err says why), 2 when no HIP device could be created.
HIPDEV_API int hipdev_session_render(void* h, const float* camPos, const float* camTarget, const float* camUp,
                                     float fov, float aspect, int width, int height, int maxBounces,
                                     int shadowQuality, int filterType, float filterWidth, int spp, int maxBatch,
                                     float aoStrength, int pathGuiding, float* mean, uint32_t* counts,
                                     int* batches, char* err, int errlen) {
    std::string e;
    std::unique_ptr<HipDevice> dev = HipDevice::create(0, &e);
    if (!dev) {
        copy_error(e, err, errlen);
        return 2;
    }
    dev->setScene(ref_scene_shared(h));
    if (!dev->hasScene()) {
        copy_error(dev->lastError(), err, errlen);
        return 1;
    }
    const std::shared_ptr<Camera> cam = makeCamera(camPos, camTarget, camUp, fov, aspect);
    RenderSettings s;
    s.width = width;
    s.height = height;
    s.samplesPerPixel = spp;
    s.maxBounces = maxBounces;
    s.shadowQuality = shadowQuality;
    s.filter.type = static_cast<PixelFilterType>(filterType);
    s.filter.width = filterWidth;
    s.motionBlur = selectedCameraMotionBlur();
    s.aoStrength = aoStrength;
    s.pathGuiding = pathGuiding != 0;
    SessionConfig cfg;
    cfg.mode = RenderMode::Final;
    cfg.targetSpp = spp;
    cfg.maxBatch = maxBatch;
    RenderSession session(*dev);
    session.begin(cam, s, cfg);
    int n = 0;
    while (!session.finished()) {
        if (!session.advance(nullptr)) {
            copy_error(dev->lastError().empty() ? "RenderSession::advance returned false" : dev->lastError(), err,
                       errlen);
            return 1;
        }
        ++n;
    }
    if (batches) *batches = n;
    const Film& film = session.film();
    for (int y = 0; y < height; ++y)
        for (int x = 0; x < width; ++x) {
            const size_t k = static_cast<size_t>(y) * static_cast<size_t>(width) + static_cast<size_t>(x);
            const Color3f c = film.resolvedPixel(x, y);
            mean[3 * k] = c.r;
            mean[3 * k + 1] = c.g;
            mean[3 * k + 2] = c.b;
            counts[k] = static_cast<uint32_t>(film.sampleCount(x, y));
        }
    return 0;
}

// RenderSession(Final, targetSpp = spp, maxBatch) with adaptive sampling on, on a HipDevice: writes the sample
// counts, the resolved mean and the number of batches (the session ends early once a batch traces nothing).
HIPDEV_API int hipdev_session_render_adaptive(void* h, const float* camPos, const float* camTarget,
                                              const float* camUp, float fov, float aspect, int width, int height,
                                              int maxBounces, int spp, int maxBatch, float threshold, int minSpp,
                                              float* mean, uint32_t* counts, int* batches, char* err, int errlen) {
    std::string e;
    std::unique_ptr<HipDevice> dev = HipDevice::create(0, &e);
    if (!dev) {
        copy_error(e, err, errlen);
        return 2;
    }
    dev->setScene(ref_scene_shared(h));
    if (!dev->hasScene()) {
        copy_error(dev->lastError(), err, errlen);
        return 1;
    }
    const std::shared_ptr<Camera> cam = makeCamera(camPos, camTarget, camUp, fov, aspect);
    RenderSettings s;
    s.width = width;
    s.height = height;
    s.samplesPerPixel = spp;
    s.maxBounces = maxBounces;
    s.filter.type = static_cast<PixelFilterType>(0);
    s.filter.width = 1.0f;
    s.motionBlur = selectedCameraMotionBlur();
    s.adaptiveSampling = true;
    s.noiseThreshold = threshold;
    s.minAdaptiveSamples = minSpp;
    SessionConfig cfg;
    cfg.mode = RenderMode::Final;
    cfg.targetSpp = spp;
    cfg.maxBatch = maxBatch;
    RenderSession session(*dev);
    session.begin(cam, s, cfg);
    int n = 0;
    while (!session.finished()) {
        if (!session.advance(nullptr)) {
            copy_error(dev->lastError().empty() ? "RenderSession::advance returned false" : dev->lastError(), err,
                       errlen);
            return 1;
        }
        ++n;
    }
    if (batches) *batches = n;
    const Film& film = session.film();
    for (int y = 0; y < height; ++y)
        for (int x = 0; x < width; ++x) {
            const size_t k = static_cast<size_t>(y) * static_cast<size_t>(width) + static_cast<size_t>(x);
            const Color3f c = film.resolvedPixel(x, y);
            mean[3 * k] = c.r;
            mean[3 * k + 1] = c.g;
            mean[3 * k + 2] = c.b;
            counts[k] = static_cast<uint32_t>(film.sampleCount(x, y));
        }
    return 0;
}

// RenderSession(Final, targetSpp = spp, maxBatch) with AOVs on (SessionConfig::aovs), on a HipDevice: writes the
// resolved beauty, albedo, normal and alpha images (3 floats/px each) and the beauty's sample counts.
HIPDEV_API int hipdev_session_render_aovs(void* h, const float* camPos, const float* camTarget, const float* camUp,
                                          float fov, float aspect, int width, int height, int maxBounces, int spp,
                                          int maxBatch, float* mean, float* albedo, float* normal, float* alpha,
                                          uint32_t* counts, char* err, int errlen) {
    std::string e;
    std::unique_ptr<HipDevice> dev = HipDevice::create(0, &e);
    if (!dev) {
        copy_error(e, err, errlen);
        return 2;
    }
    dev->setScene(ref_scene_shared(h));
    if (!dev->hasScene()) {
        copy_error(dev->lastError(), err, errlen);
        return 1;
    }
    const std::shared_ptr<Camera> cam = makeCamera(camPos, camTarget, camUp, fov, aspect);
    RenderSettings s;
    s.width = width;
    s.height = height;
    s.samplesPerPixel = spp;
    s.maxBounces = maxBounces;
    s.filter.type = static_cast<PixelFilterType>(0);
    s.filter.width = 1.0f;
    s.motionBlur = selectedCameraMotionBlur();
    SessionConfig cfg;
    cfg.mode = RenderMode::Final;
    cfg.targetSpp = spp;
    cfg.maxBatch = maxBatch;
    cfg.aovs = true;
    RenderSession session(*dev);
    session.begin(cam, s, cfg);
    while (!session.finished()) {
        if (!session.advance(nullptr)) {
            copy_error(dev->lastError().empty() ? "RenderSession::advance returned false" : dev->lastError(), err,
                       errlen);
            return 1;
        }
    }
    const Film& film = session.film();
    const AovFilms* aovs = session.aovs();
    if (!aovs) {
        copy_error("the session kept no AOV films", err, errlen);
        return 1;
    }
    for (int y = 0; y < height; ++y)
        for (int x = 0; x < width; ++x) {
            const size_t k = static_cast<size_t>(y) * static_cast<size_t>(width) + static_cast<size_t>(x);
            const Color3f c[4] = {film.resolvedPixel(x, y), aovs->albedo.resolvedPixel(x, y),
                                  aovs->normal.resolvedPixel(x, y), aovs->alpha.resolvedPixel(x, y)};
            float* outs[4] = {mean, albedo, normal, alpha};
            for (int j = 0; j < 4; ++j) {
                outs[j][3 * k] = c[j].r;
                outs[j][3 * k + 1] = c[j].g;
                outs[j][3 * k + 2] = c[j].b;
            }
            counts[k] = static_cast<uint32_t>(film.sampleCount(x, y));
        }
    return 0;
}

// RenderSession(Final, targetSpp = spp, maxBatch) with SessionConfig::passes on a HipDevice: the resolved
// beauty, the pass films (ref_copy_passes layout), the sample counts, the number of light groups it kept and,
// with `crypto`, the cryptomatte ranks (ref_render_passes layout).
HIPDEV_API int hipdev_session_render_passes(void* h, const float* camPos, const float* camTarget, const float* camUp,
                                            float fov, float aspect, int width, int height, int maxBounces, int spp,
                                            int maxBatch, float* mean, float* passes, uint32_t* counts, int* groups,
                                            float* crypto, char* err, int errlen) {
    std::string e;
    std::unique_ptr<HipDevice> dev = HipDevice::create(0, &e);
    if (!dev) {
        copy_error(e, err, errlen);
        return 2;
    }
    dev->setScene(ref_scene_shared(h));
    if (!dev->hasScene()) {
        copy_error(dev->lastError(), err, errlen);
        return 1;
    }
    const std::shared_ptr<Camera> cam = makeCamera(camPos, camTarget, camUp, fov, aspect);
    RenderSettings s;
    s.width = width;
    s.height = height;
    s.samplesPerPixel = spp;
    s.maxBounces = maxBounces;
    s.filter.type = static_cast<PixelFilterType>(0);
    s.filter.width = 1.0f;
    s.motionBlur = selectedCameraMotionBlur();
    SessionConfig cfg;
    cfg.mode = RenderMode::Final;
    cfg.targetSpp = spp;
    cfg.maxBatch = maxBatch;
    cfg.passes = true;
    RenderSession session(*dev);
    session.begin(cam, s, cfg);
    if (!session.warnings().empty()) {
        copy_error(session.warnings().front(), err, errlen);
        return 1;
    }
    while (!session.finished()) {
        if (!session.advance(nullptr)) {
            copy_error(dev->lastError().empty() ? "RenderSession::advance returned false" : dev->lastError(), err,
                       errlen);
            return 1;
        }
    }
    const Film& film = session.film();
    const RenderPassFilms* pf = session.passes();
    if (!pf) {
        copy_error("the session kept no pass films", err, errlen);
        return 1;
    }
    for (int y = 0; y < height; ++y)
        for (int x = 0; x < width; ++x) {
            const size_t k = static_cast<size_t>(y) * static_cast<size_t>(width) + static_cast<size_t>(x);
            const Color3f c = film.resolvedPixel(x, y);
            mean[3 * k] = c.r;
            mean[3 * k + 1] = c.g;
            mean[3 * k + 2] = c.b;
            counts[k] = static_cast<uint32_t>(film.sampleCount(x, y));
        }
    ref_copy_passes(*pf, width, height, passes);
    *groups = static_cast<int>(pf->lightGroups.size());
    if (crypto) {  // CryptoFilm::resolveRanks(3) of objects, then materials (w * h RGBA per layer)
        const size_t layer = static_cast<size_t>(width) * height * 4;
        const std::vector<CryptoRankImage> ranks[2] = {pf->cryptoObject.resolveRanks(3), pf->cryptoMaterial.resolveRanks(3)};
        for (int k = 0; k < 2; ++k)
            for (size_t j = 0; j < ranks[k].size() && j < 3; ++j)
                std::copy(ranks[k][j].rgba.begin(), ranks[k][j].rgba.end(), crypto + (3 * k + j) * layer);
    }
    return 0;
}

// The HIP devices as the adapter lists them: writes "id|name|usable|note" lines into out.
HIPDEV_API int hipdev_enumerate(char* out, int outlen) {
    std::string s;
    for (const DeviceInfo& d : enumerateHipDevices())
        s += d.id + "|" + d.name + "|" + (d.usable ? "1" : "0") + "|" + d.note + "\n";
    copy_error(s, out, outlen);
    return 0;
}

// RenderSession(Interactive, targetSpp = spp) on a HipDevice: one-sample passes (render_session.cpp:69-71), the
// wall time of each advance() in passSeconds[spp], the resolved mean and counts. Then the cancel path: a Final
// session of cancelSpp samples whose second batch (sized from the first one's speed, batchSeconds 1e3) is cancelled
// by another thread cancelAfterMs into it; cancelLatency = seconds from the flag's store to advance()'s return,
// cancelCompleted = that batch's SampleBatchResult::completed, cancelSamples = the session's samples after it.
HIPDEV_API int hipdev_session_interactive(void* h, const float* camPos, const float* camTarget, const float* camUp,
                                          float fov, float aspect, int width, int height, int maxBounces, int spp,
                                          float* mean, uint32_t* counts, double* passSeconds, int cancelSpp,
                                          int cancelAfterMs, double* cancelLatency, int* cancelCompleted,
                                          int* cancelSamples, char* err, int errlen) {
    using Clock = std::chrono::steady_clock;
    std::string e;
    std::unique_ptr<HipDevice> dev = HipDevice::create(0, &e);
    if (!dev) {
        copy_error(e, err, errlen);
        return 2;
    }
    dev->setScene(ref_scene_shared(h));
    if (!dev->hasScene()) {
        copy_error(dev->lastError(), err, errlen);
        return 1;
    }
    const std::shared_ptr<Camera> cam = makeCamera(camPos, camTarget, camUp, fov, aspect);
    RenderSettings s;
    s.width = width;
    s.height = height;
    s.samplesPerPixel = spp;
    s.maxBounces = maxBounces;
    s.filter.type = static_cast<PixelFilterType>(0);
    s.filter.width = 1.0f;
    s.motionBlur = selectedCameraMotionBlur();
    {
        SessionConfig cfg;
        cfg.mode = RenderMode::Interactive;
        cfg.targetSpp = spp;
        RenderSession session(*dev);
        session.begin(cam, s, cfg);
        for (int i = 0; i < spp; ++i) {
            const auto t0 = Clock::now();
            if (!session.advance(nullptr)) {
                copy_error("interactive pass failed: " + dev->lastError(), err, errlen);
                return 1;
            }
            passSeconds[i] = std::chrono::duration<double>(Clock::now() - t0).count();
        }
        if (!session.finished()) {
            copy_error("interactive session not finished after targetSpp passes", err, errlen);
            return 1;
        }
        const Film& film = session.film();
        for (int y = 0; y < height; ++y)
            for (int x = 0; x < width; ++x) {
                const size_t k = static_cast<size_t>(y) * static_cast<size_t>(width) + static_cast<size_t>(x);
                const Color3f c = film.resolvedPixel(x, y);
                mean[3 * k] = c.r;
                mean[3 * k + 1] = c.g;
                mean[3 * k + 2] = c.b;
                counts[k] = static_cast<uint32_t>(film.sampleCount(x, y));
            }
    }
    SessionConfig cfg;
    cfg.mode = RenderMode::Final;
    cfg.targetSpp = cancelSpp;
    cfg.maxBatch = cancelSpp;
    cfg.batchSeconds = 1e3;
    s.samplesPerPixel = cancelSpp;
    RenderSession session(*dev);
    session.begin(cam, s, cfg);
    if (!session.advance(nullptr)) {  // the first batch is one sample (nextBatch, render_session.cpp:69)
        copy_error("first batch failed: " + dev->lastError(), err, errlen);
        return 1;
    }
    std::atomic<bool> cancel{false};
    Clock::time_point stored;
    std::thread canceller([&] {
        std::this_thread::sleep_for(std::chrono::milliseconds(cancelAfterMs));
        stored = Clock::now();
        cancel.store(true, std::memory_order_relaxed);
    });
    const bool ok = session.advance(&cancel);
    const auto back = Clock::now();
    canceller.join();
    *cancelLatency = std::chrono::duration<double>(back - stored).count();
    *cancelCompleted = ok ? 1 : 0;
    *cancelSamples = session.samples();
    return 0;
}

}  // extern "C"

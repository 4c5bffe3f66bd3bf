// serve.hip — the render-job server behind pe_server_* / pe_job_*: continuous batching of render jobs.
//
// RenderSession (engine/render_session.cpp:69-101) advances one film through RenderDevice::render in batches.
// The server does the same for many jobs at once. Its lanes are host threads, each driving render_samples on
// a HIP stream of its own with its own path state and traversal-stack spill, so the chunks the lanes render
// run on the device together. Jobs are admitted in submission order while their films fit the film budget
// and the admission cap; a lane that becomes free takes the next chunk of the admitted jobs (jobs short of
// their preview first, then the oldest), so a job that arrives while others render starts at the next free
// lane instead of behind whole jobs. A job never has two chunks in flight, and its chunk boundaries depend
// only on the job and the server configuration: its film is the same bit for bit alone and in any mix.
#include "photon_hip.h"
#include "pe_admission.h"
#include "pe_internal.h"

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <set>
#include <string>
#include <thread>
#include <vector>

namespace {

using Clock = std::chrono::steady_clock;

constexpr int kMaxSlots = 8;

struct Job {
    uint64_t id = 0;
    pe_job_desc desc{};
    int preview = 0;            // samples of every pixel in the first milestone
    uint64_t film_bytes = 0;
    std::unique_ptr<pe_film> film;
    int state = PE_JOB_QUEUED;
    int samples_done = 0;
    int chunks = 0;
    int error = PE_OK;
    std::string message;
    double t_submit = -1.0, t_admit = -1.0, t_preview = -1.0, t_done = -1.0;
    uint64_t extension_rays = 0, shadow_rays = 0;
    bool running = false;       // a lane renders one of its chunks
    int copying = 0;            // pe_job_download / pe_job_display calls reading the film
    bool uses_scene = false;    // counted as a user of its scene by the admission planner
    volatile int32_t cancel = 0;  // polled by render_samples between launches

    bool ended() const { return state == PE_JOB_DONE || state == PE_JOB_CANCELLED || state == PE_JOB_FAILED; }
};

struct Lane {
    hipStream_t stream = nullptr;
    pe::Wavefront* wf = nullptr;
    pe::DevBuf spill;
    std::thread thread;
};

}  // namespace

struct pe_server {
    pe_device* dev = nullptr;
    pe_server_config cfg{};
    Clock::time_point t0;
    std::mutex mu;
    std::condition_variable cv;  // every state change: lanes wait for work, callers for their jobs
    std::map<uint64_t, std::unique_ptr<Job>> jobs;
    std::deque<uint64_t> queue;  // queued jobs in submission order
    std::vector<std::unique_ptr<Lane>> lanes;
    hipStream_t copy_stream = nullptr;
    uint64_t next_id = 1;
    pe::ResidencyPlanner planner;  // films of admitted jobs and resident scenes within the memory budget
    std::set<pe_scene*> tracked;   // scenes whose residency the server accounts for
    uint64_t lane_bytes = 0;       // device bytes the lanes hold (path pools, results, stack spill)
    uint64_t parks = 0, restores = 0;
    int admitted = 0;            // jobs holding a film
    bool stop = false;

    double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }
};

namespace pe {
namespace {

uint64_t film_bytes_of(int w, int h) { return static_cast<uint64_t>(w) * static_cast<uint64_t>(h) * 20 + 8; }

uint64_t scene_key(const pe_scene* sc) { return reinterpret_cast<uintptr_t>(sc); }

// Device bytes of a scene, resident or parked.
uint64_t scene_bytes(pe_scene& sc) {
    if (sc.resident) return sc.device_bytes();
    uint64_t b = 0;
    for (const auto& v : sc.parked) b += v.size();
    return b;
}

// The scene's current size and residency into the planner (an update or a direct render may have changed them).
void refresh(pe_server& s, pe_scene* sc) {
    std::shared_lock<std::shared_mutex> lk(sc->use);
    s.planner.track(scene_key(sc), scene_bytes(*sc), sc->resident);
}

// Starts accounting for a scene (with the lock held).
void track_scene(pe_server& s, pe_scene* sc) {
    if (s.tracked.insert(sc).second) {
        std::lock_guard<std::mutex> g(sc->trackers_mu);
        sc->trackers.push_back(&s);
    }
    refresh(s, sc);
}

void end_scene_use(pe_server& s, Job& j);

// Admits queued jobs in submission order while the head fits the memory budget, parking idle scenes (least
// recently used first) to make room and restoring the head's scene if it was parked; the head is never
// overtaken, so a large job waits for memory instead of starving. Called with the lock held.
void admit(pe_server& s) {
    while (!s.queue.empty()) {
        Job& j = *s.jobs.at(s.queue.front());
        if (j.cancel) {  // cancelled while queued
            j.state = PE_JOB_CANCELLED;
            j.t_done = s.now();
            s.queue.pop_front();
            continue;
        }
        if (s.admitted >= s.cfg.max_admitted) return;
        for (pe_scene* sc : s.tracked) refresh(s, sc);
        pe_scene* scene = j.desc.scene;
        const ResidencyPlanner::Plan plan = s.planner.plan(j.film_bytes, scene_key(scene));
        if (!plan.admit) return;
        int rc = PE_OK;
        for (uint64_t key : plan.evict) {
            pe_scene* victim = reinterpret_cast<pe_scene*>(static_cast<uintptr_t>(key));
            std::unique_lock<std::shared_mutex> lk(victim->use);
            rc = scene_park(*victim);
            if (rc != PE_OK) break;
            ++s.parks;
        }
        if (rc == PE_OK && plan.restore) {
            std::unique_lock<std::shared_mutex> lk(scene->use);
            rc = scene_restore(*scene);
            if (rc == PE_OK) ++s.restores;
        }
        auto f = std::make_unique<pe_film>();
        f->dev = s.dev;
        f->width = j.desc.width;
        f->height = j.desc.height;
        const size_t px = static_cast<size_t>(f->width) * f->height;
        bool ok = rc == PE_OK && f->sum.alloc(px * 12) && f->count.alloc(px * 4) && f->lum2.alloc(px * 4) &&
                  f->rejected.alloc(8);
        if (ok) {
            ok = hipMemsetAsync(f->sum.get(), 0, px * 12, s.copy_stream) == hipSuccess &&
                 hipMemsetAsync(f->count.get(), 0, px * 4, s.copy_stream) == hipSuccess &&
                 hipMemsetAsync(f->lum2.get(), 0, px * 4, s.copy_stream) == hipSuccess &&
                 hipMemsetAsync(f->rejected.get(), 0, 8, s.copy_stream) == hipSuccess &&
                 hipStreamSynchronize(s.copy_stream) == hipSuccess;
            if (!ok) set_error("film clear failed");
        }
        s.queue.pop_front();
        // The planner follows what happened on the device, whatever the outcome.
        for (pe_scene* sc : s.tracked) refresh(s, sc);
        if (!ok) {
            j.state = PE_JOB_FAILED;
            j.error = rc != PE_OK ? rc : PE_ERR_OOM;
            j.message = pe_last_error();
            j.t_done = s.now();
            continue;
        }
        s.planner.commit(ResidencyPlanner::Plan{true, false, false, {}}, j.film_bytes, scene_key(scene));
        j.uses_scene = true;
        j.film = std::move(f);
        j.state = PE_JOB_ADMITTED;
        j.t_admit = s.now();
        ++s.admitted;
    }
}

// Samples of every pixel in the job's next chunk: the rest of the preview first, then chunk_spp, each capped
// so that one chunk keeps at most max_chunk_paths paths in flight (at least one sample of every pixel).
int chunk_of(const pe_server& s, const Job& j) {
    const uint64_t px = static_cast<uint64_t>(j.desc.width) * static_cast<uint64_t>(j.desc.height);
    const int cap = static_cast<int>(std::clamp<uint64_t>(s.cfg.max_chunk_paths / std::max<uint64_t>(px, 1), 1, 1 << 20));
    const int want = j.samples_done < j.preview ? j.preview - j.samples_done : s.cfg.chunk_spp;
    return std::max(1, std::min({want, cap, j.desc.spp - j.samples_done}));
}

// The next job a free lane works on: admitted, idle, unfinished; short of its preview first, then oldest.
Job* pick(pe_server& s) {
    Job* best = nullptr;
    for (auto& kv : s.jobs) {
        Job& j = *kv.second;
        if (j.state != PE_JOB_ADMITTED || j.running || j.copying > 0) continue;
        if (j.cancel) {
            j.state = PE_JOB_CANCELLED;
            j.t_done = s.now();
            end_scene_use(s, j);
            s.cv.notify_all();
            continue;
        }
        if (!best) {
            best = &j;
            continue;
        }
        const bool a = j.samples_done < j.preview, b = best->samples_done < best->preview;
        if (a != b ? a : j.id < best->id) best = &j;
    }
    return best;
}

void lane_main(pe_server* s, Lane* lane) {
    (void)hipSetDevice(s->dev->ordinal);
    const RenderLane rl{lane->stream, lane->wf, lane->spill.as<uint32_t>(), s->dev->max_grid_blocks};
    std::unique_lock<std::mutex> lk(s->mu);
    for (;;) {
        if (s->stop) return;
        Job* j = pick(*s);
        if (!j) {
            s->cv.wait(lk);
            continue;
        }
        j->running = true;
        const int first = j->samples_done;
        const int n = chunk_of(*s, *j);
        pe_settings st = j->desc.settings;
        st.samples_per_launch = n;  // one chunk, one launch
        const pe_camera cam = j->desc.camera;
        lk.unlock();
        pe_batch_result r{};
        const int rc = render_samples(s->dev, rl, j->desc.scene, &cam, &st, j->film.get(), first, n, &j->cancel, &r);
        const std::string msg = rc == PE_OK ? std::string() : std::string(pe_last_error());
        lk.lock();
        j->running = false;
        j->extension_rays += r.extension_rays;
        j->shadow_rays += r.shadow_rays;
        const double t = s->now();
        if (rc == PE_OK) {
            j->samples_done += n;
            ++j->chunks;
            if (j->t_preview < 0.0 && j->samples_done >= j->preview) j->t_preview = t;
            if (j->samples_done >= j->desc.spp) {
                j->state = PE_JOB_DONE;
                j->t_done = t;
            } else if (j->cancel) {
                j->state = PE_JOB_CANCELLED;
                j->t_done = t;
            }
        } else if (rc == PE_ERR_CANCELLED) {
            j->state = PE_JOB_CANCELLED;
            j->t_done = t;
        } else {
            j->state = PE_JOB_FAILED;
            j->error = rc;
            j->message = msg;
            j->t_done = t;
        }
        if (j->ended()) {
            end_scene_use(*s, *j);
            admit(*s);  // its scene may now be parked for a queued job
        }
        s->cv.notify_all();
    }
}

void fill_status(const Job& j, pe_job_status* st) {
    if (!st) return;
    std::memset(st, 0, sizeof(*st));
    st->state = j.state;
    st->samples_done = j.samples_done;
    st->error = j.error;
    st->chunks = j.chunks;
    st->t_submit = j.t_submit;
    st->t_admit = j.t_admit;
    st->t_preview = j.t_preview;
    st->t_done = j.t_done;
    st->extension_rays = j.extension_rays;
    st->shadow_rays = j.shadow_rays;
    std::snprintf(st->message, sizeof(st->message), "%s", j.message.c_str());
}

Job* find(pe_server* s, uint64_t id) {
    auto it = s->jobs.find(id);
    if (it == s->jobs.end()) {
        set_error("unknown job id " + std::to_string(id));
        return nullptr;
    }
    return it->second.get();
}

// An admitted job stopped rendering: its scene may be parked once no other admitted job renders it.
void end_scene_use(pe_server& s, Job& j) {
    if (!j.uses_scene) return;
    j.uses_scene = false;
    s.planner.end_use(scene_key(j.desc.scene));
}

void release_film(pe_server& s, Job& j) {
    end_scene_use(s, j);
    if (!j.film) return;
    j.film.reset();  // frees after the device is idle (hipFree waits for it)
    s.planner.free_film(j.film_bytes);
    --s.admitted;
}

}  // namespace

void server_forget_scene(pe_server* s, pe_scene* sc) {
    std::lock_guard<std::mutex> lk(s->mu);
    s->planner.forget(scene_key(sc));
    s->tracked.erase(sc);
}

}  // namespace pe

using namespace pe;

extern "C" {

int pe_server_create(pe_device* dev, const pe_server_config* config, pe_server** out) {
    if (!dev || !out) {
        set_error("pe_server_create: NULL argument");
        return PE_ERR_INVALID;
    }
    *out = nullptr;
    try {
        auto s = std::make_unique<pe_server>();
        s->dev = dev;
        s->t0 = Clock::now();
        pe_server_config c = config ? *config : pe_server_config{};
        if (c.slots < 0 || c.slots > kMaxSlots || c.chunk_spp < 0 || c.max_admitted < 0) {
            set_error("pe_server_create: slots must be in [0, 8]; chunk_spp and max_admitted >= 0");
            return PE_ERR_INVALID;
        }
        if (c.slots == 0) c.slots = 4;
        if (c.chunk_spp == 0) c.chunk_spp = 16;
        if (c.max_chunk_paths == 0) c.max_chunk_paths = uint64_t{1} << 25;
        c.max_chunk_paths = std::min<uint64_t>(c.max_chunk_paths, uint64_t{1} << 30);
        if (c.max_admitted == 0) c.max_admitted = 64;
        if (c.lane_paths <= 0) c.lane_paths = 1 << 24;
        PE_HIP_CHECK(hipSetDevice(dev->ordinal));
        s->cfg = c;
        // Undoes a partial construction (no lane thread has started yet).
        auto fail = [&](int rc) {
            for (auto& lane : s->lanes) {
                if (lane->wf) wavefront_delete(lane->wf);
                if (lane->stream) (void)hipStreamDestroy(lane->stream);
            }
            if (s->copy_stream) (void)hipStreamDestroy(s->copy_stream);
            return rc;
        };
        if (hipStreamCreateWithFlags(&s->copy_stream, hipStreamNonBlocking) != hipSuccess) {
            set_error("pe_server_create: stream creation failed");
            return fail(PE_ERR_HIP);
        }
        for (int i = 0; i < c.slots; ++i) {
            s->lanes.push_back(std::make_unique<Lane>());
            Lane& lane = *s->lanes.back();
            if (hipStreamCreateWithFlags(&lane.stream, hipStreamNonBlocking) != hipSuccess) {
                lane.stream = nullptr;
                set_error("pe_server_create: stream creation failed");
                return fail(PE_ERR_HIP);
            }
            lane.wf = wavefront_new();
            if (!lane.spill.alloc(trace_spill_bytes(dev->max_grid_blocks))) return fail(PE_ERR_OOM);
            // The lane's path pool and the results of its largest chunk, reserved up front so lanes do not
            // allocate while others render.
            const int rc = wavefront_reserve(
                lane.wf, static_cast<uint32_t>(std::min<uint64_t>(c.max_chunk_paths, static_cast<uint64_t>(c.lane_paths))),
                c.max_chunk_paths, 1);
            if (rc != PE_OK) return fail(rc);
            s->lane_bytes += wavefront_device_bytes(lane.wf) + lane.spill.bytes();
        }
        // The memory budget of films and resident scenes: as given, or three quarters of the device's memory
        // less what the lanes hold.
        if (c.memory_budget == 0) {
            size_t free_b = 0, total_b = 0;
            if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) return fail(PE_ERR_HIP);
            const uint64_t share = total_b / 4 * 3;
            c.memory_budget = share > s->lane_bytes ? share - s->lane_bytes : 0;
        }
        s->cfg = c;
        s->planner = ResidencyPlanner(c.memory_budget);
        for (auto& lane : s->lanes) lane->thread = std::thread(lane_main, s.get(), lane.get());
        *out = s.release();
        return PE_OK;
    } catch (const std::exception& e) {
        set_error(std::string("pe_server_create: ") + e.what());
        return PE_ERR_INVALID;
    }
}

int pe_server_destroy(pe_server* s) {
    if (!s) return PE_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
        for (auto& kv : s->jobs) kv.second->cancel = 1;
        s->cv.notify_all();
    }
    for (auto& lane : s->lanes)
        if (lane->thread.joinable()) lane->thread.join();
    (void)hipSetDevice(s->dev->ordinal);
    for (auto& lane : s->lanes) {
        (void)hipStreamSynchronize(lane->stream);
        wavefront_delete(lane->wf);
        lane->spill.release();
        (void)hipStreamDestroy(lane->stream);
    }
    for (auto& kv : s->jobs) kv.second->desc.scene->job_refs.fetch_sub(1);
    s->jobs.clear();
    for (pe_scene* sc : s->tracked) {
        std::lock_guard<std::mutex> g(sc->trackers_mu);
        sc->trackers.erase(std::remove(sc->trackers.begin(), sc->trackers.end(), s), sc->trackers.end());
    }
    s->tracked.clear();
    if (s->copy_stream) (void)hipStreamDestroy(s->copy_stream);
    delete s;
    return PE_OK;
}

double pe_server_clock(const pe_server* s) { return s ? s->now() : -1.0; }

int pe_server_stats_get(pe_server* s, pe_server_stats* out) {
    if (!s || !out) {
        set_error("pe_server_stats_get: NULL argument");
        return PE_ERR_INVALID;
    }
    std::lock_guard<std::mutex> lk(s->mu);
    *out = pe_server_stats{};
    out->memory_budget = s->planner.budget();
    out->memory_used = s->planner.used();
    out->lane_bytes = s->lane_bytes;
    out->scenes_parked = s->parks;
    out->scenes_restored = s->restores;
    out->admitted = s->admitted;
    out->queued = static_cast<int32_t>(s->queue.size());
    out->scenes_tracked = static_cast<int32_t>(s->tracked.size());
    for (pe_scene* sc : s->tracked) out->scenes_resident += sc->resident ? 1 : 0;
    return PE_OK;
}

int pe_job_submit(pe_server* s, const pe_job_desc* job, uint64_t* id) {
    if (!s || !job || !id) {
        set_error("pe_job_submit: NULL argument");
        return PE_ERR_INVALID;
    }
    if (!job->scene || job->scene->dev != s->dev || job->width <= 0 || job->height <= 0 || job->spp <= 0 ||
        job->preview_spp < 0) {
        set_error("pe_job_submit: a job needs a scene of the server's device, a positive film size and spp");
        return PE_ERR_INVALID;
    }
    if (!(job->camera.lens_radius >= 0.0f) || job->settings.max_bounces < 0 || job->settings.max_bounces > 100) {
        set_error("pe_job_submit: lens radius >= 0 and max_bounces in [0, 100] only");
        return PE_ERR_INVALID;
    }
    try {
        auto j = std::make_unique<Job>();
        j->desc = *job;
        j->preview = std::min(job->preview_spp > 0 ? job->preview_spp : 16, job->spp);
        j->film_bytes = film_bytes_of(job->width, job->height);
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->stop) {
            set_error("pe_job_submit: server is shutting down");
            return PE_ERR_INVALID;
        }
        (void)hipSetDevice(s->dev->ordinal);
        track_scene(*s, job->scene);
        if (s->planner.plan(j->film_bytes, scene_key(job->scene)).never) {
            set_error("pe_job_submit: the film and the scene together exceed the server's memory budget");
            return PE_ERR_OOM;
        }
        job->scene->job_refs.fetch_add(1);
        j->id = s->next_id++;
        j->t_submit = s->now();
        *id = j->id;
        s->queue.push_back(j->id);
        s->jobs.emplace(j->id, std::move(j));
        admit(*s);
        s->cv.notify_all();
        return PE_OK;
    } catch (const std::exception& e) {
        set_error(std::string("pe_job_submit: ") + e.what());
        return PE_ERR_INVALID;
    }
}

int pe_job_poll(pe_server* s, uint64_t id, pe_job_status* status) {
    if (!s || !status) {
        set_error("pe_job_poll: NULL argument");
        return PE_ERR_INVALID;
    }
    std::lock_guard<std::mutex> lk(s->mu);
    const Job* j = find(s, id);
    if (!j) return PE_ERR_INVALID;
    fill_status(*j, status);
    return PE_OK;
}

int pe_job_wait(pe_server* s, uint64_t id, int32_t timeout_ms, pe_job_status* status) {
    if (!s) {
        set_error("pe_job_wait: NULL server");
        return PE_ERR_INVALID;
    }
    std::unique_lock<std::mutex> lk(s->mu);
    if (!find(s, id)) return PE_ERR_INVALID;
    auto ended = [&] {
        auto it = s->jobs.find(id);
        return it == s->jobs.end() || it->second->ended();
    };
    if (timeout_ms < 0) s->cv.wait(lk, ended);
    else s->cv.wait_for(lk, std::chrono::milliseconds(timeout_ms), ended);
    Job* j = find(s, id);
    if (!j) return PE_ERR_INVALID;
    fill_status(*j, status);
    if (!j->ended()) {
        set_error("pe_job_wait: timed out");
        return PE_ERR_TIMEOUT;
    }
    return PE_OK;
}

int pe_job_cancel(pe_server* s, uint64_t id) {
    if (!s) {
        set_error("pe_job_cancel: NULL server");
        return PE_ERR_INVALID;
    }
    std::lock_guard<std::mutex> lk(s->mu);
    Job* j = find(s, id);
    if (!j) return PE_ERR_INVALID;
    if (j->ended()) return PE_OK;
    j->cancel = 1;
    if (j->state == PE_JOB_QUEUED || (j->state == PE_JOB_ADMITTED && !j->running)) {
        // Not on a lane: it ends now; a running chunk ends it when the lane returns.
        j->state = PE_JOB_CANCELLED;
        j->t_done = s->now();
        end_scene_use(*s, *j);
        if (!s->queue.empty()) s->queue.erase(std::remove(s->queue.begin(), s->queue.end(), id), s->queue.end());
    }
    s->cv.notify_all();
    return PE_OK;
}

int pe_job_download(pe_server* s, uint64_t id, float* sum_rgb, uint32_t* count, float* sum_lum2) {
    if (!s) {
        set_error("pe_job_download: NULL server");
        return PE_ERR_INVALID;
    }
    std::unique_lock<std::mutex> lk(s->mu);
    Job* j = find(s, id);
    if (!j) return PE_ERR_INVALID;
    s->cv.wait(lk, [&] { return !j->running; });  // whole chunks only
    if (!j->film) {
        set_error("pe_job_download: the job holds no film (queued, or failed before admission)");
        return PE_ERR_INVALID;
    }
    ++j->copying;
    pe_film* f = j->film.get();
    lk.unlock();
    const size_t px = static_cast<size_t>(f->width) * f->height;
    hipError_t e = hipSetDevice(s->dev->ordinal);
    if (e == hipSuccess && sum_rgb) e = hipMemcpyAsync(sum_rgb, f->sum.get(), px * 12, hipMemcpyDeviceToHost, s->copy_stream);
    if (e == hipSuccess && count) e = hipMemcpyAsync(count, f->count.get(), px * 4, hipMemcpyDeviceToHost, s->copy_stream);
    if (e == hipSuccess && sum_lum2) e = hipMemcpyAsync(sum_lum2, f->lum2.get(), px * 4, hipMemcpyDeviceToHost, s->copy_stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->copy_stream);
    lk.lock();
    --j->copying;
    s->cv.notify_all();
    if (e != hipSuccess) {
        set_error(std::string("pe_job_download: ") + hipGetErrorString(e));
        return PE_ERR_HIP;
    }
    return PE_OK;
}

int pe_job_display(pe_server* s, uint64_t id, const pe_display* settings, float* display_rgb, uint8_t* rgb8) {
    if (!s || !settings) {
        set_error("pe_job_display: NULL argument");
        return PE_ERR_INVALID;
    }
    std::unique_lock<std::mutex> lk(s->mu);
    Job* j = find(s, id);
    if (!j) return PE_ERR_INVALID;
    s->cv.wait(lk, [&] { return !j->running; });  // whole chunks only
    if (!j->film) {
        set_error("pe_job_display: the job holds no film (queued, or failed before admission)");
        return PE_ERR_INVALID;
    }
    ++j->copying;
    pe_film* f = j->film.get();
    lk.unlock();
    const int rc = pe_film_display(f, settings, display_rgb, rgb8);
    lk.lock();
    --j->copying;
    s->cv.notify_all();
    return rc;
}

int pe_job_release(pe_server* s, uint64_t id) {
    if (!s) {
        set_error("pe_job_release: NULL server");
        return PE_ERR_INVALID;
    }
    std::unique_lock<std::mutex> lk(s->mu);
    Job* j = find(s, id);
    if (!j) return PE_ERR_INVALID;
    j->cancel = 1;
    s->cv.wait(lk, [&] { return !j->running && j->copying == 0; });
    if (j->state == PE_JOB_QUEUED)
        s->queue.erase(std::remove(s->queue.begin(), s->queue.end(), id), s->queue.end());
    (void)hipSetDevice(s->dev->ordinal);
    release_film(*s, *j);
    j->desc.scene->job_refs.fetch_sub(1);
    s->jobs.erase(id);
    admit(*s);  // the freed film memory may admit the next queued job
    s->cv.notify_all();
    return PE_OK;
}

}  // extern "C"

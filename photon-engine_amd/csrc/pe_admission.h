// pe_admission.h — device-memory admission of the render-job server (serve.hip), as plain host logic.
//
// The server's lanes hold their path pools, result buffers and traversal-stack spill for the server's whole
// life; what admission decides over is the rest of its memory budget: the film of every admitted job and every
// scene kept resident on the device. A job is admitted when its film, plus its scene if that is parked in host
// memory, fits what is left of the budget, after parking as many idle resident scenes (no admitted job renders
// them) as that takes, least recently used first. A scene an admitted job uses is never parked. Nothing here
// touches the device: the server carries out the plan (parks and restores scenes, allocates the film) and
// commits it, and the CPU tests drive the same class against a capped allocator.
#pragma once

#include <algorithm>
#include <cstdint>
#include <map>
#include <vector>

namespace pe {

class ResidencyPlanner {
public:
    struct Scene {
        uint64_t bytes = 0;
        bool resident = true;
        int users = 0;          // admitted jobs rendering the scene
        uint64_t last_use = 0;  // planner tick of the last admission or release that touched it
    };
    struct Plan {
        bool admit = false;           // the job fits now (after the evictions below)
        bool never = false;           // the job cannot fit even with every idle scene parked
        bool restore = false;         // its scene must be made resident first
        std::vector<uint64_t> evict;  // idle resident scenes to park, in this order
    };

    explicit ResidencyPlanner(uint64_t budget = 0) : m_budget(budget) {}

    uint64_t budget() const { return m_budget; }
    // Bytes held: films of admitted jobs and resident scenes (can exceed the budget when scenes registered
    // resident bring more than it allows; admission then parks idle scenes first).
    uint64_t used() const { return m_used; }
    const std::map<uint64_t, Scene>& scenes() const { return m_scenes; }

    // Registers a scene or refreshes its size and residency (the scene's own state is the source of truth).
    void track(uint64_t key, uint64_t bytes, bool resident) {
        auto it = m_scenes.find(key);
        if (it == m_scenes.end()) {
            m_scenes.emplace(key, Scene{bytes, resident, 0, ++m_tick});
            if (resident) m_used += bytes;
            return;
        }
        Scene& s = it->second;
        if (s.resident) m_used -= s.bytes;
        s.bytes = bytes;
        s.resident = resident;
        if (s.resident) m_used += s.bytes;
    }

    // Forgets a scene no admitted job uses (its owner released it).
    bool forget(uint64_t key) {
        auto it = m_scenes.find(key);
        if (it == m_scenes.end()) return true;
        if (it->second.users > 0) return false;
        if (it->second.resident) m_used -= it->second.bytes;
        m_scenes.erase(it);
        return true;
    }

    Plan plan(uint64_t film_bytes, uint64_t key) const {
        Plan p;
        const auto it = m_scenes.find(key);
        const uint64_t scene_bytes = it == m_scenes.end() ? 0 : it->second.bytes;
        const bool resident = it == m_scenes.end() || it->second.resident;
        p.restore = !resident;
        // Memory that no admission can reclaim: films and the scenes admitted jobs render.
        uint64_t pinned = 0;
        for (const auto& kv : m_scenes)
            if (kv.second.resident && kv.second.users > 0) pinned += kv.second.bytes;
        pinned += m_film_bytes;
        const bool in_use = it != m_scenes.end() && it->second.resident && it->second.users > 0;
        const uint64_t need = film_bytes + (it == m_scenes.end() || in_use ? 0 : scene_bytes);
        if (pinned + need > m_budget) {
            // Even parking every idle scene would not make room now; it never fits if nothing else is held.
            p.never = film_bytes + scene_bytes > m_budget;
            return p;
        }
        const uint64_t after = m_used + film_bytes + (resident ? 0 : scene_bytes);
        if (after <= m_budget) {
            p.admit = true;
            return p;
        }
        uint64_t excess = after - m_budget;
        std::vector<std::pair<uint64_t, uint64_t>> idle;  // (last use, key)
        for (const auto& kv : m_scenes)
            if (kv.second.resident && kv.second.users == 0 && kv.first != key) idle.emplace_back(kv.second.last_use, kv.first);
        std::sort(idle.begin(), idle.end());
        for (const auto& lk : idle) {
            if (excess == 0) break;
            const uint64_t b = m_scenes.at(lk.second).bytes;
            p.evict.push_back(lk.second);
            excess = b >= excess ? 0 : excess - b;
        }
        p.admit = excess == 0;
        if (!p.admit) p.evict.clear();
        return p;
    }

    // Applies an admitted plan: the evictions, the restore, the film, one more user of the scene.
    void commit(const Plan& p, uint64_t film_bytes, uint64_t key) {
        for (uint64_t k : p.evict) {
            Scene& s = m_scenes.at(k);
            if (s.resident) m_used -= s.bytes;
            s.resident = false;
        }
        auto it = m_scenes.find(key);
        if (it != m_scenes.end()) {
            Scene& s = it->second;
            if (!s.resident) m_used += s.bytes;
            s.resident = true;
            ++s.users;
            s.last_use = ++m_tick;
        }
        m_used += film_bytes;
        m_film_bytes += film_bytes;
    }

    // A job stopped rendering (done, cancelled or failed): its scene is one user less (and idle once none is left).
    void end_use(uint64_t key) {
        auto it = m_scenes.find(key);
        if (it != m_scenes.end()) {
            it->second.users = std::max(0, it->second.users - 1);
            it->second.last_use = ++m_tick;
        }
    }
    // A job's film was freed (the job was released).
    void free_film(uint64_t film_bytes) {
        m_used -= film_bytes;
        m_film_bytes -= film_bytes;
    }
    void release(uint64_t film_bytes, uint64_t key) {
        end_use(key);
        free_film(film_bytes);
    }

private:
    uint64_t m_budget = 0, m_used = 0, m_film_bytes = 0, m_tick = 0;
    std::map<uint64_t, Scene> m_scenes;
};

}  // namespace pe

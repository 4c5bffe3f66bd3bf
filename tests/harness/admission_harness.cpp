// C entry points over the render-job server's admission planner (photon-engine_amd/csrc/pe_admission.h), so
// the CPU tests drive the same class the server uses. Test infrastructure only.
#include "pe_admission.h"

#include <cstdint>

using pe::ResidencyPlanner;

extern "C" {

void* planner_new(uint64_t budget) { return new ResidencyPlanner(budget); }
void planner_delete(void* p) { delete static_cast<ResidencyPlanner*>(p); }
void planner_track(void* p, uint64_t key, uint64_t bytes, int resident) {
    static_cast<ResidencyPlanner*>(p)->track(key, bytes, resident != 0);
}
int planner_forget(void* p, uint64_t key) { return static_cast<ResidencyPlanner*>(p)->forget(key) ? 1 : 0; }
// Plans a job; evict receives up to cap keys. Returns admit | never << 1 | restore << 2.
int planner_plan(void* p, uint64_t film, uint64_t key, uint64_t* evict, int cap, int* n_evict) {
    const ResidencyPlanner::Plan pl = static_cast<ResidencyPlanner*>(p)->plan(film, key);
    *n_evict = static_cast<int>(pl.evict.size());
    for (int k = 0; k < *n_evict && k < cap; ++k) evict[k] = pl.evict[static_cast<size_t>(k)];
    return (pl.admit ? 1 : 0) | (pl.never ? 2 : 0) | (pl.restore ? 4 : 0);
}
void planner_commit(void* p, uint64_t film, uint64_t key, const uint64_t* evict, int n_evict, int restore) {
    ResidencyPlanner::Plan pl;
    pl.admit = true;
    pl.restore = restore != 0;
    pl.evict.assign(evict, evict + n_evict);
    static_cast<ResidencyPlanner*>(p)->commit(pl, film, key);
}
void planner_end_use(void* p, uint64_t key) { static_cast<ResidencyPlanner*>(p)->end_use(key); }
void planner_free_film(void* p, uint64_t film) { static_cast<ResidencyPlanner*>(p)->free_film(film); }
uint64_t planner_used(void* p) { return static_cast<ResidencyPlanner*>(p)->used(); }
int planner_resident(void* p, uint64_t key) {
    const auto& m = static_cast<ResidencyPlanner*>(p)->scenes();
    const auto it = m.find(key);
    return it == m.end() ? -1 : (it->second.resident ? 1 : 0);
}
}

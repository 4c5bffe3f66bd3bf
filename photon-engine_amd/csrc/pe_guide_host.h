// pe_guide_host.h — the guide object behind pe_guide_* (guide.hip) and what the renderer needs of it.
#pragma once

#include "pe_guide.h"
#include "pe_internal.h"

#include <cstdint>
#include <random>

struct pe_guide {
    pe_device* dev = nullptr;
    pe_guide_config cfg{};           // resolved (no zero defaults left)
    pe::V3 lo{}, inv_extent{};       // position normalisation
    // Parameters: fp32 masters, Adam moments, gradient; bf16 W ([out][in]) and W^T ([in][out]) per layer.
    pe::DevBuf params, adam_m, adam_v, grad, wbf, wtbf;
    // Training scratch for one batch of M rows.
    uint32_t batch_rows = 0;
    pe::DevBuf recs, x0, h[4], out, d_out, d_h[2], partial, scalars;
    // Records collected by renders: a hash-addressed buffer of `record_slots` entries (3 float4 each).
    uint32_t record_slots = 0;
    pe::DevBuf records;
    // Per-render buffers: the mixture of every queue slot and two open training records per path.
    uint32_t path_capacity = 0, slot_capacity = 0;
    pe::DevBuf mix, path_records;
    uint64_t steps = 0;
    uint64_t launches = 0;           // render launches seen
    uint64_t collected = 0;          // records written by renders (an upper bound: slots may be overwritten)
    uint64_t inferred = 0;           // mixtures inferred by renders
    float last_loss = 0.0f;
    std::mt19937_64 rng{0};

    pe::GuideView view(float fraction) const {
        pe::GuideView g{};
        g.wbf = wbf.as<__bf16>();
        g.params = params.as<float>();
        g.lo = lo;
        g.inv_extent = inv_extent;
        g.fraction = fraction;
        g.inferred = nullptr;
        return g;
    }
    bool training() const { return cfg.train_launches <= 0 || launches < static_cast<uint64_t>(cfg.train_launches); }
};

namespace pe {
// Sizes the per-render buffers for `paths` paths in flight and `slots` ray-queue slots.
int guide_reserve_paths(pe_guide* g, uint32_t paths, uint32_t slots);
// Clears the record buffer before a launch collects into it.
int guide_clear_records(pe_guide* g, hipStream_t s);
// cfg.steps_per_launch Adam steps on batches of the collected records, on stream s.
int guide_train_from_records(pe_guide* g, hipStream_t s);
}  // namespace pe

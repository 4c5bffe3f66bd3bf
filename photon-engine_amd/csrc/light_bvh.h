// light_bvh.h — host build of the light BVH (many-light selection) read by the gfx950 light picker.
#pragma once

#include "pe_device_scene.h"

#include <vector>

namespace pe {

struct LightBvh {
    int mode = kPickUniform;          // kPickUniform / kPickExhaustive / kPickBvh
    std::vector<GpuLightNode> nodes;  // depth-first, root at 0 (empty when no light is bounded)
    std::vector<LightPath> paths;     // one per candidate (lights, then the environment)
    std::vector<int> top;             // unbounded candidates, in top-level slot order
};

// LightSampler::build (light_sampler.cpp:42-77) for `num_lights` explicit lights plus the environment.
void build_light_bvh(const std::vector<GpuLight>& lights, int num_lights, bool has_env, LightBvh& out);

}  // namespace pe

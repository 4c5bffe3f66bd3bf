// pe_scene_host.h — host-side packing of a pe_scene_desc into device buffers.
#pragma once

#include "pe_internal.h"

namespace pe {

// Validates desc, builds the BVH and per-triangle records, and uploads them on dev->stream.
int upload_scene(pe_device* dev, const pe_scene_desc& desc, pe_scene& sc);

// Materials, lights, environment distributions and energy tables (scene_prep.cpp); fills sc.view.
int prepare_shading(pe_device* dev, const pe_scene_desc& desc, pe_scene& sc);
// The material half (materials, energy tables) and the light half (lights, environment, light selection) of
// prepare_shading, each refreshing its own fields of the scene's view.
int prepare_materials(pe_device* dev, const pe_scene_desc& desc, pe_scene& sc);
int prepare_lights(pe_device* dev, const pe_scene_desc& desc, pe_scene& sc);

}  // namespace pe

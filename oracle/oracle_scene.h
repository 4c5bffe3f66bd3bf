// oracle_scene.h — the oracle's scene representation (test infrastructure only).
#pragma once

#include "oracle.h"

#include <memory>
#include <vector>

namespace oracle {

struct TriRec {
    uint32_t v[3];
    int shape, prim, material, emitter;
    bool hasNormals, hasUv, cameraInvisible;
};

struct TriAccel {  // bvh.h:84-88
    V3 v0, e1, e2;
    float parallelEps;
};

struct LinearNode {  // bvh.h:27-37
    V3 lo, hi;
    int offset = 0;  // leaf: first primitive; interior: second child
    int count = 0;
    int axis = 0;
};

struct Scene {
    std::vector<V3> pos, nrm;
    std::vector<float> uv;  // 2 per vertex
    std::vector<TriRec> tris;
    std::vector<pe_material_desc> materials;
    std::vector<pe_light_desc> lights;
    bool hasEnv = false;
    pe_env_desc env{};
    std::vector<float> envRgb;
    // Light groups (scene.cpp:144-173) as the descriptor resolves them, and per-shape node uids.
    std::vector<int> lightGroup;  // per light; empty: no groups
    int envLightGroup = -1;
    int numLightGroups = 0;
    std::vector<uint64_t> nodeUid;  // per shape (mesh of the descriptor)
    // Light links (scene.cpp:120-141): per light the receiver / shadow modes (PE_LINK_*) and node uids.
    bool hasLinks = false, hasShadowLinks = false;
    std::vector<int> linkMode, shadowLinkMode;
    std::vector<std::vector<uint64_t>> linkUids, shadowLinkUids;
    // Scene::background() (scene.h:24-28): PE_BACKGROUND_* and the Color mode's colour.
    int bgMode = PE_BACKGROUND_ENVIRONMENT;
    V3 bgColor{0.18f, 0.18f, 0.18f};

    // BVH
    std::vector<LinearNode> nodes;
    std::vector<int> ordered;  // BVH primitive order -> triangle index
    std::vector<TriAccel> accel;

    // Integrator-side data prepared at build (lights, environment distributions): oracle_integrator.cpp
    std::shared_ptr<void> integrator;
};

void buildBvh(Scene& s);
bool hitTriangleLean(const TriAccel& T, const Ray& r, float& t, float& u, float& v);

}  // namespace oracle

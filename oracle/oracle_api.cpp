// oracle_api.cpp — C entry points of liboracle.so for the tests (ctypes). Test infrastructure only.
#include "oracle.h"
#include "oracle_scene.h"

#include <cstring>

namespace oracle {

// Scene::addShape / addLight / setEnvironment order as flattened in pe_scene_desc (scene.cpp:18-104).
Scene* buildScene(const pe_scene_desc& d) {
    auto* s = new Scene();
    s->materials.assign(d.materials, d.materials + d.num_materials);
    s->lights.assign(d.lights, d.lights + d.num_lights);
    uint32_t vb = 0;
    for (int m = 0; m < d.num_meshes; ++m) {
        const pe_mesh_desc& md = d.meshes[m];
        for (uint32_t v = 0; v < md.num_vertices; ++v) {
            s->pos.emplace_back(md.positions[3 * v], md.positions[3 * v + 1], md.positions[3 * v + 2]);
            s->nrm.push_back(md.normals ? V3(md.normals[3 * v], md.normals[3 * v + 1], md.normals[3 * v + 2]) : V3(0.0f));
            s->uv.push_back(md.uvs ? md.uvs[2 * v] : 0.0f);
            s->uv.push_back(md.uvs ? md.uvs[2 * v + 1] : 0.0f);
        }
        // TriangleMesh without normals: area-weighted vertex normals (mesh.cpp:22-54).
        const bool computed = !md.normals && !md.standalone && md.num_vertices > 0 && md.num_triangles > 0;
        if (computed) {
            for (uint32_t t = 0; t < md.num_triangles; ++t) {
                const uint32_t i0 = vb + md.indices[3 * t], i1 = vb + md.indices[3 * t + 1], i2 = vb + md.indices[3 * t + 2];
                const V3 n = cross(s->pos[i1] - s->pos[i0], s->pos[i2] - s->pos[i0]);
                s->nrm[i0] += n;
                s->nrm[i1] += n;
                s->nrm[i2] += n;
            }
            for (uint32_t v = 0; v < md.num_vertices; ++v) {
                V3& n = s->nrm[vb + v];
                if (lengthSq(n) > 0.0f) n = normalized(n);
            }
        }
        for (uint32_t t = 0; t < md.num_triangles; ++t) {
            TriRec r;
            for (int k = 0; k < 3; ++k) r.v[k] = vb + md.indices[3 * t + k];
            r.shape = m;
            r.prim = md.standalone ? -1 : static_cast<int>(t);
            r.material = md.material;
            r.emitter = md.emitter_base < 0 ? -1 : (md.standalone ? md.emitter_base : md.emitter_base + static_cast<int>(t));
            r.hasNormals = md.normals != nullptr || computed;
            r.hasUv = md.uvs != nullptr;
            r.cameraInvisible = (md.flags & PE_SHAPE_CAMERA_INVISIBLE) != 0;
            s->tris.push_back(r);
        }
        vb += md.num_vertices;
    }
    if (d.environment) {
        s->hasEnv = true;
        s->env = *d.environment;
        if (s->env.rgb && s->env.width > 0 && s->env.height > 0) {
            s->envRgb.assign(s->env.rgb, s->env.rgb + 3ull * s->env.width * s->env.height);
        }
        s->env.rgb = nullptr;
    }
    if (d.light_groups && d.num_light_groups > 0) {
        s->lightGroup.assign(d.light_groups, d.light_groups + d.num_lights);
        s->envLightGroup = d.env_light_group;
        s->numLightGroups = d.num_light_groups;
    }
    s->nodeUid.assign(static_cast<size_t>(d.num_meshes), 0);
    if (d.node_uids) s->nodeUid.assign(d.node_uids, d.node_uids + d.num_meshes);
    const size_t nl = static_cast<size_t>(d.num_lights);
    s->linkMode.assign(nl, PE_LINK_ALL);
    s->shadowLinkMode.assign(nl, PE_LINK_ALL);
    s->linkUids.assign(nl, {});
    s->shadowLinkUids.assign(nl, {});
    if (d.link_modes) {
        for (size_t i = 0; i < nl; ++i) {
            const uint32_t* o = d.link_offsets + 4 * i;
            s->linkMode[i] = d.link_modes[2 * i];
            s->shadowLinkMode[i] = d.link_modes[2 * i + 1];
            s->linkUids[i].assign(d.link_uids + o[0], d.link_uids + o[0] + o[1]);
            s->shadowLinkUids[i].assign(d.link_uids + o[2], d.link_uids + o[2] + o[3]);
            s->hasLinks = s->hasLinks || s->linkMode[i] != PE_LINK_ALL;
            s->hasShadowLinks = s->hasShadowLinks || s->shadowLinkMode[i] != PE_LINK_ALL;
        }
    }
    s->bgMode = d.background_mode;
    s->bgColor = V3(d.background_color[0], d.background_color[1], d.background_color[2]);
    buildBvh(*s);
    return s;
}

void freeScene(Scene* s) { delete s; }

}  // namespace oracle

using namespace oracle;

#define ORACLE_API extern "C" __attribute__((visibility("default")))

ORACLE_API int oracle_sobol_batch(uint64_t seed, size_t n, const int32_t* px, const int32_t* py, const uint32_t* index,
                                  const uint32_t* dim, uint32_t* out) {
    Sobol s(seed);
    for (size_t i = 0; i < n; ++i) {
        s.startPixel(px[i], py[i]);
        s.startSample(static_cast<int>(index[i]));
        s.dimension = dim[i];
        s.get2DBits(out[2 * i], out[2 * i + 1]);
    }
    return 0;
}

ORACLE_API uint32_t oracle_reverse_bits(uint32_t x) { return reverseBits(x); }
ORACLE_API uint32_t oracle_sobol_dim1(uint32_t x) { return sobolDim1(x); }

ORACLE_API void* oracle_scene_build(const pe_scene_desc* d) { return buildScene(*d); }
ORACLE_API void oracle_scene_free(void* s) { freeScene(static_cast<Scene*>(s)); }
ORACLE_API int oracle_scene_triangles(void* s) { return sceneTriangles(*static_cast<Scene*>(s)); }

// hits: t (float bits), shape, prim, tri ; bary: u, v (optional)
ORACLE_API int oracle_intersect(void* sp, size_t n, const float* rays, int32_t* hits, float* bary) {
    const Scene& s = *static_cast<Scene*>(sp);
    for (size_t i = 0; i < n; ++i) {
        const float* r = rays + 8 * i;
        Ray ray;
        ray.o = V3(r[0], r[1], r[2]);
        ray.d = V3(r[3], r[4], r[5]);
        ray.tMin = r[6];
        ray.tMax = r[7];
        Hit h;
        int32_t* o = hits + 4 * i;
        if (intersect(s, ray, h)) {
            std::memcpy(&o[0], &h.t, 4);
            o[1] = h.shape;
            o[2] = h.prim;
            o[3] = h.tri;
        } else {
            const float m = -1.0f;
            std::memcpy(&o[0], &m, 4);
            o[1] = o[2] = o[3] = -1;
        }
        if (bary) {
            bary[2 * i] = h.u;
            bary[2 * i + 1] = h.v;
        }
    }
    return 0;
}

// Brute force over every triangle in scene order with the same per-triangle test: the closest t wins,
// ties go to the later triangle (t <= tMax). A tree-free check of oracle_intersect.
ORACLE_API int oracle_intersect_brute(void* sp, size_t n, const float* rays, int32_t* hits) {
    const Scene& s = *static_cast<Scene*>(sp);
    for (size_t i = 0; i < n; ++i) {
        const float* r = rays + 8 * i;
        Ray ray;
        ray.o = V3(r[0], r[1], r[2]);
        ray.d = V3(r[3], r[4], r[5]);
        ray.tMin = r[6];
        ray.tMax = r[7];
        int best = -1;
        for (size_t k = 0; k < s.tris.size(); ++k) {
            const V3 p0 = s.pos[s.tris[k].v[0]], p1 = s.pos[s.tris[k].v[1]], p2 = s.pos[s.tris[k].v[2]];
            const V3 e1 = p1 - p0, e2 = p2 - p0;
            const TriAccel T{p0, e1, e2, 1e-14f * lengthSq(e1) * lengthSq(e2)};
            float t, u, v;
            if (hitTriangleLean(T, ray, t, u, v)) {
                ray.tMax = t;
                best = static_cast<int>(k);
            }
        }
        int32_t* o = hits + 4 * i;
        const float t = best >= 0 ? ray.tMax : -1.0f;
        std::memcpy(&o[0], &t, 4);
        o[1] = best >= 0 ? s.tris[static_cast<size_t>(best)].shape : -1;
        o[2] = best >= 0 ? s.tris[static_cast<size_t>(best)].prim : -1;
        o[3] = best;
    }
    return 0;
}

ORACLE_API int oracle_occluded(void* sp, size_t n, const float* rays, uint8_t* occ) {
    const Scene& s = *static_cast<Scene*>(sp);
    for (size_t i = 0; i < n; ++i) {
        const float* r = rays + 8 * i;
        Ray ray;
        ray.o = V3(r[0], r[1], r[2]);
        ray.d = V3(r[3], r[4], r[5]);
        ray.tMin = r[6];
        ray.tMax = r[7];
        occ[i] = intersectAny(s, ray) ? 1 : 0;
    }
    return 0;
}

// ref_driver.cpp — test-only harness around the PhotonEngine reference sources.
//
// This file is part of the oracle (test infrastructure): it is compiled together with the reference's
// own hot-path translation units, in place under the reference tree, into oracle/_ref/libphoton_ref.so
// (recipe: oracle/Makefile.ref). Nothing in the product links it. It exposes a small C API that the
// tests and bench.py's CPU-baseline leg call through ctypes: Sobol values, scene build from the same
// flattened pe_scene_desc the HIP library takes, closest/any hit on ray batches, the microfacet energy
// tables, Disney BSDF eval/sample, per-sample radiance and Renderer::renderSamples itself.
#include "engine/denoiser.h"
#include "engine/renderer.h"
#include "engine/scene.h"
#include "camera/orthographic_camera.h"
#include "camera/perspective_camera.h"
#include "camera/thin_lens_camera.h"
#include "core/image/pixel_filter.h"
#include "core/image/cryptomatte.h"
#include "core/image/imager.h"
#include "core/image/tone_mapping.h"
#include "core/color/transfer.h"
#include "geometry/mesh.h"
#include "geometry/mesh_instance.h"
#include "integrators/path_tracer.h"
#include "lights/area_light.h"
#include "lights/environment_light.h"
#include "materials/dielectric.h"
#include "materials/fresnel.h"
#include "materials/disney.h"
#include "materials/lambertian.h"
#include "materials/mirror.h"
#include "materials/microfacet.h"
#include "samplers/sobol_sampler.h"
#include "scene/mesh_binary.h"

#include "photon_hip.h"

#include <algorithm>
#include <chrono>
#include <cstring>
#include <string>
#include <memory>
#include <unordered_map>
#include <vector>

using namespace photon;

namespace {

Color3f c3(const float* v) { return Color3f(v[0], v[1], v[2]); }
Vec3f v3(const float* v) { return Vec3f(v[0], v[1], v[2]); }

std::shared_ptr<Material> makeMaterial(const pe_material_desc& d) {
    if (d.type == PE_MAT_LAMBERT) return std::make_shared<Lambertian>(c3(d.base_color), c3(d.emission));
    if (d.type == PE_MAT_MIRROR) return std::make_shared<Mirror>(c3(d.base_color));
    if (d.type == PE_MAT_DIELECTRIC) {
        auto m = std::make_shared<Dielectric>(d.ior, c3(d.base_color), d.roughness);
        if (d.abbe > 0.0f) m->setAbbe(d.abbe);
        if (d.absorption_distance > 0.0f) m->setAbsorption(c3(d.absorption_color), d.absorption_distance);
        m->setThinWalled(d.thin_walled != 0);
        if (d.thin_film_thickness > 0.0f) m->setThinFilm(d.thin_film_thickness);
        if (d.scatter_density > 0.0f) m->setScattering(d.scatter_density, c3(d.scatter_color), d.scatter_anisotropy);
        return m;
    }
    auto m = std::make_shared<DisneyMaterial>(c3(d.base_color), d.metallic, d.roughness, d.specular);
    m->setClearCoat(d.clearcoat);
    m->setClearCoatRoughness(d.clearcoat_roughness);
    m->setCoatIor(d.coat_ior);
    m->setCoatTint(c3(d.coat_tint));
    m->setIor(d.ior);
    m->setSpecularTransmission(d.specular_transmission);
    m->setAnisotropy(d.anisotropy);
    m->setSheen(d.sheen);
    m->setSheenRoughness(d.sheen_roughness);
    if (d.has_sheen_color) m->setSheenColor(c3(d.sheen_color));
    m->setDiffuseRoughness(d.diffuse_roughness);
    m->setDiffuseTransmission(d.diffuse_transmission);
    m->setSubsurface(d.subsurface);
    m->setSubsurfaceRadius(c3(d.subsurface_radius));
    m->setSubsurfaceScale(d.subsurface_scale);
    m->setSubsurfaceAnisotropy(d.subsurface_anisotropy);
    m->setEdgeTint(c3(d.edge_tint));
    m->setThinFilm(d.thin_film_thickness, d.thin_film_ior);
    m->setEmission(c3(d.emission));
    if (d.has_opacity) m->setOpacity(d.opacity);
    if (d.flake_density > 0.0f)
        m->setFlakes(d.flake_density, d.flake_size, d.flake_roughness, c3(d.flake_color));
    return m;
}

struct RefScene {
    std::shared_ptr<Scene> scene;
    std::vector<std::shared_ptr<Material>> materials;
    std::unordered_map<const void*, int> shapeIndex;
    std::vector<int> meshMaterial;  // per descriptor mesh (= Scene::shapes() index): its material index
};

bool toRay(const float* r, Ray& ray) {
    ray = Ray(Vec3f(r[0], r[1], r[2]), Vec3f(r[3], r[4], r[5]), r[6], r[7]);
    return true;
}

SurfaceInteraction shadingPoint(const float* n, const float* t, int frontFace) {
    SurfaceInteraction si;
    si.point = Vec3f(0.0f);
    si.normal = v3(n);
    si.ng = v3(n);
    si.tangent = v3(t);
    si.frontFace = frontFace != 0;
    return si;
}

}  // namespace

#define REF_API __attribute__((visibility("default")))

// The camera every function below builds from (position, target, up, fov, aspect): a PerspectiveCamera unless
// ref_camera_select chose a ThinLensCamera (with setApertureShape) or an OrthographicCamera (vertical extent =
// `fov`'s argument is ignored for it), optionally with a shutter-end pose (Camera::setShutterEnd) of the same kind
// at another position / target, and the motion-blur flag tracePixel reads.
struct CameraSpec {
    int kind = 0;  // 0 perspective, 1 thin lens, 2 orthographic
    float aperture = 0.0f, focus = 1.0f, extent = 2.0f, rotationDeg = 0.0f;
    int blades = 0;
    bool hasEnd = false, motionBlur = false;
    float endPos[3] = {0, 0, 0}, endTarget[3] = {0, 0, -1}, endAperture = 0.0f;
    float shiftX = 0.0f, shiftY = 0.0f;  // setShift of a perspective or thin-lens camera (both poses)
    float aoStrength = 0.0f;             // RenderSettings::aoStrength of the renders and traces below
};
CameraSpec g_camera;

std::unique_ptr<Camera> makeCameraAt(const float* pos, const float* target, const float* up, float fov, float aspect,
                                     float aperture) {
    switch (g_camera.kind) {
        case 1: {
            auto c = std::make_unique<ThinLensCamera>(v3(pos), v3(target), v3(up), fov, aspect, aperture, g_camera.focus);
            c->setApertureShape(g_camera.blades, g_camera.rotationDeg);
            if (g_camera.shiftX != 0.0f || g_camera.shiftY != 0.0f) c->setShift(g_camera.shiftX, g_camera.shiftY);
            return c;
        }
        case 2: return std::make_unique<OrthographicCamera>(v3(pos), v3(target), v3(up), g_camera.extent, aspect);
        default: {
            auto c = std::make_unique<PerspectiveCamera>(v3(pos), v3(target), v3(up), fov, aspect);
            if (g_camera.shiftX != 0.0f || g_camera.shiftY != 0.0f) c->setShift(g_camera.shiftX, g_camera.shiftY);
            return c;
        }
    }
}

std::unique_ptr<Camera> makeCamera(const float* pos, const float* target, const float* up, float fov, float aspect) {
    auto cam = makeCameraAt(pos, target, up, fov, aspect, g_camera.aperture);
    if (g_camera.hasEnd) cam->setShutterEnd(*makeCameraAt(g_camera.endPos, g_camera.endTarget, up, fov, aspect,
                                                          g_camera.endAperture));
    return cam;
}

bool selectedCameraMotionBlur() { return g_camera.motionBlur; }

// tracePixel's camera ray (renderer.cpp:151-153) after the filter dimension.
Ray cameraRay(const Camera& cam, Sampler& sampler, float sx, float sy) {
    const Vec2f lens = sampler.get2D();
    return g_camera.motionBlur ? cam.generateRay(sx, sy, lens, sampler.get1D()) : cam.generateRay(sx, sy, lens);
}

extern "C" {

// SobolSampler(seed): startPixel(px, py), startSample(index), `dim` get1D() calls, then get2D().
REF_API int ref_sobol_batch(uint64_t seed, size_t n, const int32_t* px, const int32_t* py, const uint32_t* index,
                    const uint32_t* dim, float* out) {
    SobolSampler s(seed);
    for (size_t i = 0; i < n; ++i) {
        s.startPixel(px[i], py[i]);
        s.startSample(static_cast<int>(index[i]));
        for (uint32_t d = 0; d < dim[i]; ++d) (void)s.get1D();
        const Vec2f v = s.get2D();
        out[2 * i] = v.x;
        out[2 * i + 1] = v.y;
    }
    return 0;
}

// The scene of a descriptor; with names, area light i's group name (Light::setLightGroup, "" = none; the
// triangle lights Scene::addShape makes for emissive meshes have none) and the environment's, which
// Scene::buildAccelerator resolves into group indices (scene.cpp:144-173). Meshes are tagged with the
// descriptor's node uids (Scene::tagShape). With `xforms` (16 floats a mesh, row-major local-to-world; a
// matrix whose last entry is 0 means none), a mesh becomes what SceneGraph::compileNode makes of a mesh node
// (scene_graph.cpp:189-202): a MeshInstance of the descriptor's (local) mesh under the matrix; `vis` gives
// each shape's visibility mask (Shape::setVisibilityMask).
void* buildRef(const pe_scene_desc* desc, const char* const* lightGroupNames, const char* envGroupName,
               const char* const* shapeNames, const char* const* materialNames, const float* xforms = nullptr,
               const uint8_t* vis = nullptr) {
    Scene::setPreferOwnBvh(true);
    auto rs = new RefScene();
    rs->scene = std::make_shared<Scene>();
    for (int i = 0; i < desc->num_materials; ++i) rs->materials.push_back(makeMaterial(desc->materials[i]));
    std::vector<char> lightAdded(static_cast<size_t>(desc->num_lights), 0);
    for (int i = 0; i < desc->num_meshes; ++i) {
        const pe_mesh_desc& m = desc->meshes[i];
        if (m.standalone) {
            const int li = m.emitter_base;
            if (li < 0 || li >= desc->num_lights || desc->lights[li].type != PE_LIGHT_AREA) continue;
            if (lightAdded[static_cast<size_t>(li)]) continue;
            lightAdded[static_cast<size_t>(li)] = 1;
            const pe_light_desc& L = desc->lights[li];
            auto area = std::make_shared<AreaLight>(v3(L.p0), v3(L.p1), v3(L.p2), c3(L.radiance));
            area->setVisibleToCamera(L.visible_to_camera != 0);
            if (lightGroupNames && lightGroupNames[li]) area->setLightGroup(lightGroupNames[li]);
            if (desc->link_modes) {  // Light::setLightLink (light.h:81)
                LightLinkSet ls;
                const uint32_t* o = desc->link_offsets + 4 * li;
                ls.mode = static_cast<LightLinkSet::Mode>(desc->link_modes[2 * li]);
                ls.shadowMode = static_cast<LightLinkSet::Mode>(desc->link_modes[2 * li + 1]);
                ls.nodeUids.assign(desc->link_uids + o[0], desc->link_uids + o[0] + o[1]);
                ls.shadowNodeUids.assign(desc->link_uids + o[2], desc->link_uids + o[2] + o[3]);
                area->setLightLink(ls);
            }
            rs->scene->addLight(area);
            continue;
        }
        std::vector<Vec3f> pos(m.num_vertices), nrm;
        std::vector<Vec2f> uv;
        for (uint32_t v = 0; v < m.num_vertices; ++v) pos[v] = v3(m.positions + 3 * v);
        if (m.normals) {
            nrm.resize(m.num_vertices);
            for (uint32_t v = 0; v < m.num_vertices; ++v) nrm[v] = v3(m.normals + 3 * v);
        }
        if (m.uvs) {
            uv.resize(m.num_vertices);
            for (uint32_t v = 0; v < m.num_vertices; ++v) uv[v] = Vec2f(m.uvs[2 * v], m.uvs[2 * v + 1]);
        }
        std::vector<uint32_t> idx(m.indices, m.indices + 3 * static_cast<size_t>(m.num_triangles));
        const Material* mat = rs->materials[static_cast<size_t>(m.material)].get();
        std::shared_ptr<Shape> shape;
        const float* X = xforms ? xforms + 16 * static_cast<size_t>(i) : nullptr;
        if (X && X[15] != 0.0f) {
            const Mat4f M(X[0], X[1], X[2], X[3], X[4], X[5], X[6], X[7], X[8], X[9], X[10], X[11], X[12], X[13],
                          X[14], X[15]);
            shape = std::make_shared<MeshInstance>(std::make_shared<TriangleMesh>(pos, nrm, uv, idx, mat), Transform(M),
                                                   mat);
        } else {
            shape = std::make_shared<TriangleMesh>(pos, nrm, uv, idx, mat);
        }
        if (vis) shape->setVisibilityMask(vis[i]);
        if (desc->node_uids && desc->node_uids[i] != 0) rs->scene->tagShape(shape.get(), desc->node_uids[i]);
        if (shapeNames && shapeNames[i] && shapeNames[i][0]) rs->scene->setShapeName(shape.get(), shapeNames[i]);
        rs->scene->addShape(shape);
    }
    for (int i = 0; i < desc->num_meshes; ++i) rs->meshMaterial.push_back(desc->meshes[i].material);
    for (auto& mat : rs->materials) rs->scene->retainMaterial(mat);
    if (materialNames)
        for (int i = 0; i < desc->num_materials; ++i)
            if (materialNames[i] && materialNames[i][0]) rs->scene->setMaterialName(rs->materials[static_cast<size_t>(i)].get(), materialNames[i]);
    if (desc->environment) {
        const pe_env_desc& e = *desc->environment;
        std::shared_ptr<Image> img;
        if (e.rgb && e.width > 0 && e.height > 0) {
            img = std::make_shared<Image>(e.width, e.height);
            for (int y = 0; y < e.height; ++y)
                for (int x = 0; x < e.width; ++x) {
                    const float* p = e.rgb + 3 * (static_cast<size_t>(y) * e.width + x);
                    img->setPixel(x, y, Color3f(p[0], p[1], p[2]));
                }
        }
        auto env = std::make_shared<EnvironmentLight>(img, e.rotation, e.intensity);
        if (e.has_sun) env->setSun(v3(e.sun_direction), e.sun_cos_max, c3(e.sun_radiance));
        if (envGroupName) env->setLightGroup(envGroupName);
        rs->scene->setEnvironment(env);
    }
    Background bg;
    bg.mode = desc->background_mode == PE_BACKGROUND_COLOR         ? Background::Mode::Color
              : desc->background_mode == PE_BACKGROUND_TRANSPARENT ? Background::Mode::Transparent
                                                                   : Background::Mode::Environment;
    bg.color = c3(desc->background_color);
    rs->scene->setBackground(bg);
    rs->scene->buildAccelerator();
    const auto& shapes = rs->scene->shapes();
    for (size_t k = 0; k < shapes.size(); ++k) rs->shapeIndex[shapes[k].get()] = static_cast<int>(k);
    return rs;
}

REF_API void* ref_scene_build(const pe_scene_desc* desc) { return buildRef(desc, nullptr, nullptr, nullptr, nullptr); }
// With light-group names, shape names (Scene::setShapeName, per descriptor mesh; area-light quads cannot be
// named) and material names (Scene::setMaterialName); any array may be NULL and any entry NULL or "".
REF_API void* ref_scene_build_groups(const pe_scene_desc* desc, const char* const* lightGroupNames,
                                     const char* envGroupName, const char* const* shapeNames,
                                     const char* const* materialNames) {
    return buildRef(desc, lightGroupNames, envGroupName, shapeNames, materialNames);
}
// As ref_scene_build_groups, with meshes made MeshInstances under `xforms` and visibility masks `vis` (either
// may be NULL; see buildRef).
REF_API void* ref_scene_build_instanced(const pe_scene_desc* desc, const char* const* lightGroupNames,
                                        const char* envGroupName, const char* const* shapeNames,
                                        const char* const* materialNames, const float* xforms, const uint8_t* vis) {
    return buildRef(desc, lightGroupNames, envGroupName, shapeNames, materialNames, xforms, vis);
}

// encodeMeshBinary (scene/mesh_binary.cpp) of a mesh (nrm / uv may be NULL): the encoded size, the bytes
// written into `out` when `cap` holds them.
REF_API size_t ref_mesh_encode(const float* pos, const float* nrm, const float* uv, uint32_t nv, const uint32_t* idx,
                               uint32_t ni, uint8_t* out, size_t cap) {
    std::vector<Vec3f> p(nv), n;
    std::vector<Vec2f> t;
    for (uint32_t i = 0; i < nv; ++i) p[i] = v3(pos + 3 * i);
    if (nrm) {
        n.resize(nv);
        for (uint32_t i = 0; i < nv; ++i) n[i] = v3(nrm + 3 * i);
    }
    if (uv) {
        t.resize(nv);
        for (uint32_t i = 0; i < nv; ++i) t[i] = Vec2f(uv[2 * i], uv[2 * i + 1]);
    }
    const TriangleMesh mesh(p, n, t, std::vector<uint32_t>(idx, idx + ni), nullptr);
    const std::vector<uint8_t> bytes = encodeMeshBinary(mesh);
    if (out && cap >= bytes.size()) std::memcpy(out, bytes.data(), bytes.size());
    return bytes.size();
}

// decodeMeshBinary of `size` bytes: 0 and counts = (vertices, indices, has uvs) on success, 1 when the
// reference refuses the bytes. pos / uv / idx receive the mesh when they are non-NULL and `counts` on entry
// holds their capacities (vertices, indices).
REF_API int ref_mesh_decode(const uint8_t* data, size_t size, float* pos, float* uv, uint32_t* idx,
                            uint32_t* counts) {
    std::string err;
    const auto mesh = decodeMeshBinary(data, size, nullptr, &err);
    if (!mesh) return 1;
    const uint32_t capV = counts[0], capI = counts[1];
    const auto& p = mesh->positions();
    const auto& t = mesh->uvs();
    const auto& ix = mesh->indices();
    counts[0] = static_cast<uint32_t>(p.size());
    counts[1] = static_cast<uint32_t>(ix.size());
    counts[2] = t.empty() ? 0u : 1u;
    if (pos && p.size() <= capV)
        for (size_t i = 0; i < p.size(); ++i) {
            pos[3 * i] = p[i].x; pos[3 * i + 1] = p[i].y; pos[3 * i + 2] = p[i].z;
        }
    if (uv && t.size() <= capV)
        for (size_t i = 0; i < t.size(); ++i) {
            uv[2 * i] = t[i].x; uv[2 * i + 1] = t[i].y;
        }
    if (idx && ix.size() <= capI) std::copy(ix.begin(), ix.end(), idx);
    return 0;
}

// temperatureColor (core/image/imager.cpp:204-213) of n temperatures.
REF_API int ref_temperature_color(size_t n, const float* kelvin, float* rgb) {
    for (size_t i = 0; i < n; ++i) {
        const Color3f c = temperatureColor(kelvin[i]);
        rgb[3 * i] = c.r; rgb[3 * i + 1] = c.g; rgb[3 * i + 2] = c.b;
    }
    return 0;
}

// cryptoHashName and cryptoHashToFloat (cryptomatte.cpp:78-112) of n names.
REF_API int ref_crypto_hash(size_t n, const char* const* names, uint32_t* hashes, float* floats) {
    for (size_t i = 0; i < n; ++i) {
        hashes[i] = cryptoHashName(names[i]);
        floats[i] = cryptoHashToFloat(hashes[i]);
    }
    return 0;
}

REF_API void ref_scene_free(void* h) { delete static_cast<RefScene*>(h); }

// The scene's resolved light groups: Scene::lightGroupIndex(i) per light, envLightGroupIndex and the group
// names ('\n'-separated into `names`). Returns the number of groups, or -1 if `names` is too small.
REF_API int ref_scene_light_groups(void* h, int32_t* lightToGroup, int32_t* envGroup, char* names, size_t namesCap) {
    const Scene& sc = *static_cast<RefScene*>(h)->scene;
    for (size_t i = 0; i < sc.lights().size(); ++i) lightToGroup[i] = sc.lightGroupIndex(static_cast<int>(i));
    *envGroup = sc.envLightGroupIndex();
    std::string joined;
    for (const std::string& n : sc.lightGroupNames()) joined += n + "\n";
    if (joined.size() + 1 > namesCap) return -1;
    std::memcpy(names, joined.c_str(), joined.size() + 1);
    return static_cast<int>(sc.lightGroupNames().size());
}

// The reference's own thinFilmReflectance (materials/fresnel.h) for n queries.
REF_API int ref_thin_film(size_t n, const float* cos1, float n1, float n2, float n3re, float n3im, const float* thickness,
                          const float* lambda, float* out) {
    for (size_t i = 0; i < n; ++i)
        out[i] = thinFilmReflectance(cos1[i], n1, n2, cfloat(n3re, n3im), thickness[i], lambda[i]);
    return 0;
}

REF_API int ref_scene_counts(void* h, int* shapes, int* lights) {
    auto* rs = static_cast<RefScene*>(h);
    *shapes = static_cast<int>(rs->scene->shapes().size());
    *lights = static_cast<int>(rs->scene->lights().size());
    return 0;
}

// hits: t, shape index (-1 miss), primIndex, 0 ; bary/extra: point xyz, ng xyz, shading normal xyz.
REF_API int ref_intersect(void* h, size_t n, const float* rays, int32_t* hits, float* geom) {
    auto* rs = static_cast<RefScene*>(h);
    for (size_t i = 0; i < n; ++i) {
        Ray ray;
        toRay(rays + 8 * i, ray);
        SurfaceInteraction si;
        int32_t* o = hits + 4 * i;
        if (rs->scene->intersect(ray, si)) {
            float t = si.t;
            std::memcpy(&o[0], &t, 4);
            auto it = rs->shapeIndex.find(si.hitObject);
            o[1] = it == rs->shapeIndex.end() ? -2 : it->second;
            o[2] = si.primIndex;
            o[3] = 0;
            if (geom) {
                float* g = geom + 9 * i;
                g[0] = si.point.x; g[1] = si.point.y; g[2] = si.point.z;
                g[3] = si.ng.x; g[4] = si.ng.y; g[5] = si.ng.z;
                g[6] = si.normal.x; g[7] = si.normal.y; g[8] = si.normal.z;
            }
        } else {
            const float t = -1.0f;
            std::memcpy(&o[0], &t, 4);
            o[1] = -1;
            o[2] = -1;
            o[3] = 0;
        }
    }
    return 0;
}

REF_API int ref_occluded(void* h, size_t n, const float* rays, uint8_t* occ) {
    auto* rs = static_cast<RefScene*>(h);
    for (size_t i = 0; i < n; ++i) {
        Ray ray;
        toRay(rays + 8 * i, ray);
        occ[i] = rs->scene->intersectAny(ray) ? 1 : 0;
    }
    return 0;
}

REF_API int ref_ao_strength(float strength) {
    g_camera.aoStrength = strength;
    return 0;
}

REF_API int ref_camera_shift(float shiftX, float shiftY) {
    g_camera.shiftX = shiftX;
    g_camera.shiftY = shiftY;
    return 0;
}

REF_API int ref_camera_select(int kind, float aperture, float focus, int blades, float rotationDeg, float extent,
                              int motionBlur, int hasEnd, const float* endPos, const float* endTarget, float endAperture) {
    g_camera = CameraSpec{};
    g_camera.kind = kind;
    g_camera.aperture = aperture;
    g_camera.focus = focus;
    g_camera.blades = blades;
    g_camera.rotationDeg = rotationDeg;
    g_camera.extent = extent;
    g_camera.motionBlur = motionBlur != 0;
    g_camera.hasEnd = hasEnd != 0;
    if (hasEnd) {
        std::copy(endPos, endPos + 3, g_camera.endPos);
        std::copy(endTarget, endTarget + 3, g_camera.endTarget);
        g_camera.endAperture = endAperture;
    }
    return 0;
}

// The selected camera's describeGpu() and describeGpuEnd(): 22 floats each (position, lowerLeft, horizontal,
// vertical, u, v, forward, lensRadius, blades, bladeRotation, orthographic).
REF_API int ref_camera_describe_full(const float* pos, const float* target, const float* up, float fov, float aspect,
                                     float* out) {
    const auto cam = makeCamera(pos, target, up, fov, aspect);
    const CameraGpuDesc ds[2] = {cam->describeGpu(), cam->describeGpuEnd()};
    for (int e = 0; e < 2; ++e) {
        const CameraGpuDesc& d = ds[e];
        float* o = out + 22 * e;
        const Vec3f vs[7] = {d.position, d.lowerLeft, d.horizontal, d.vertical, d.u, d.v, d.forward};
        for (int k = 0; k < 7; ++k) {
            o[3 * k] = vs[k].x;
            o[3 * k + 1] = vs[k].y;
            o[3 * k + 2] = vs[k].z;
        }
        o[21] = d.lensRadius;
    }
    out[44] = static_cast<float>(ds[0].blades);
    out[45] = ds[0].bladeRotation;
    out[46] = ds[0].orthographic ? 1.0f : 0.0f;
    return 0;
}

// Camera::generateRay(u, v, lensSample, time) of the selected camera for n samples (in: 5 floats each, u, v,
// lens x, lens y, time; time < 0 calls the three-argument form): out 6 floats each, origin and direction.
REF_API int ref_camera_rays(const float* pos, const float* target, const float* up, float fov, float aspect, size_t n,
                            const float* in, float* out) {
    const auto cam = makeCamera(pos, target, up, fov, aspect);
    for (size_t i = 0; i < n; ++i) {
        const float* q = in + 5 * i;
        const Ray r = q[4] < 0.0f ? cam->generateRay(q[0], q[1], Vec2f(q[2], q[3]))
                                  : cam->generateRay(q[0], q[1], Vec2f(q[2], q[3]), q[4]);
        const float v[6] = {r.origin.x, r.origin.y, r.origin.z, r.direction.x, r.direction.y, r.direction.z};
        std::copy(v, v + 6, out + 6 * i);
    }
    return 0;
}

// PerspectiveCamera(...).describeGpu(): position, lowerLeft, horizontal, vertical.
REF_API int ref_camera_describe(const float* pos, const float* target, const float* up, float fov, float aspect, float* out) {
    PerspectiveCamera cam(v3(pos), v3(target), v3(up), fov, aspect);
    const CameraGpuDesc d = cam.describeGpu();
    const Vec3f vs[4] = {d.position, d.lowerLeft, d.horizontal, d.vertical};
    for (int k = 0; k < 4; ++k) {
        out[3 * k] = vs[k].x;
        out[3 * k + 1] = vs[k].y;
        out[3 * k + 2] = vs[k].z;
    }
    return 0;
}

// which: 0 ggxDirectionalAlbedo(a, b), 1 ggxSchlickAlbedo(a, b, c), 2 ggxSchlickAverageAlbedo(a, b),
//        3 charlieSheenAlbedo(a, b)
REF_API int ref_microfacet(int which, size_t n, const float* a, const float* b, const float* c, float* out) {
    for (size_t i = 0; i < n; ++i) {
        switch (which) {
            case 0: out[i] = ggxDirectionalAlbedo(a[i], b[i]); break;
            case 1: out[i] = ggxSchlickAlbedo(a[i], b[i], c[i]); break;
            case 2: out[i] = ggxSchlickAverageAlbedo(a[i], b[i]); break;
            case 3: out[i] = charlieSheenAlbedo(a[i], b[i]); break;
            default: return -1;
        }
    }
    return 0;
}

// Material::evalPdf for n (normal, tangent, frontFace, wo, wi) tuples; f: 3 floats, pdf: 1 float each.
REF_API int ref_bsdf_eval_at(const pe_material_desc* md, size_t n, const float* points, const float* normal,
                             const float* tangent, const int32_t* frontFace, const float* wo, const float* wi, float* f,
                             float* pdf) {
    auto mat = makeMaterial(*md);
    for (size_t i = 0; i < n; ++i) {
        SurfaceInteraction si = shadingPoint(normal + 3 * i, tangent + 3 * i, frontFace[i]);
        if (points) si.point = v3(points + 3 * i);
        si.material = mat.get();
        float p = 0.0f;
        const Color3f v = mat->evalPdf(v3(wo + 3 * i), v3(wi + 3 * i), si, p);
        f[3 * i] = v.r; f[3 * i + 1] = v.g; f[3 * i + 2] = v.b;
        pdf[i] = p;
    }
    return 0;
}

// Material::sampleFull; out per sample: ok, wi xyz, f rgb, pdf, delta, subsurface, passThrough, weight rgb (14 floats).
REF_API int ref_bsdf_eval(const pe_material_desc* md, size_t n, const float* normal, const float* tangent,
                  const int32_t* frontFace, const float* wo, const float* wi, float* f, float* pdf) {
    return ref_bsdf_eval_at(md, n, nullptr, normal, tangent, frontFace, wo, wi, f, pdf);
}

// DisneyMaterial::resolve's flake state at n shading points: out 4 floats each, inFlake, flakeNormal.
REF_API int ref_flake_resolve(const pe_material_desc* md, size_t n, const float* points, const float* normal,
                              const float* tangent, const int32_t* frontFace, float* out) {
    auto mat = makeMaterial(*md);
    const auto* dm = dynamic_cast<const DisneyMaterial*>(mat.get());
    if (!dm) return -1;
    for (size_t i = 0; i < n; ++i) {
        SurfaceInteraction si = shadingPoint(normal + 3 * i, tangent + 3 * i, frontFace[i]);
        si.point = v3(points + 3 * i);
        const DisneyMaterial::ShadingParams p = dm->resolve(si);
        float* o = out + 4 * i;
        o[0] = p.inFlake ? 1.0f : 0.0f;
        o[1] = p.flakeNormal.x; o[2] = p.flakeNormal.y; o[3] = p.flakeNormal.z;
    }
    return 0;
}

REF_API int ref_bsdf_sample_at(const pe_material_desc* md, size_t n, const float* points, const float* normal,
                               const float* tangent, const int32_t* frontFace, const float* wo, const float* u3,
                               float* out) {
    auto mat = makeMaterial(*md);
    for (size_t i = 0; i < n; ++i) {
        SurfaceInteraction si = shadingPoint(normal + 3 * i, tangent + 3 * i, frontFace[i]);
        if (points) si.point = v3(points + 3 * i);
        si.material = mat.get();
        BSDFSample bs;
        const bool ok = mat->sampleFull(v3(wo + 3 * i), si, u3[3 * i], Vec2f(u3[3 * i + 1], u3[3 * i + 2]), bs);
        float* o = out + 14 * i;
        o[0] = ok ? 1.0f : 0.0f;
        o[1] = bs.wi.x; o[2] = bs.wi.y; o[3] = bs.wi.z;
        o[4] = bs.f.r; o[5] = bs.f.g; o[6] = bs.f.b;
        o[7] = bs.pdf;
        o[8] = bs.delta ? 1.0f : 0.0f;
        o[9] = bs.subsurface ? 1.0f : 0.0f;
        o[10] = bs.passThrough ? 1.0f : 0.0f;
        o[11] = bs.weight.r; o[12] = bs.weight.g; o[13] = bs.weight.b;
    }
    return 0;
}

REF_API int ref_bsdf_sample(const pe_material_desc* md, size_t n, const float* normal, const float* tangent,
                    const int32_t* frontFace, const float* wo, const float* u3, float* out) {
    return ref_bsdf_sample_at(md, n, nullptr, normal, tangent, frontFace, wo, u3, out);
}

// Per-sample radiance along the path of tracePixel (renderer.cpp:139-162) with no AOVs: the Sobol sampler
// is cloned with `seed`, the pixel filter importance-samples the sub-pixel offset (dims 0-1), the lens
// sample takes dims 2-3 and PathTracer::Li(maxDepth, rrDepth 3, ao 0, shadowQuality) traces the path.
REF_API int ref_trace_samples(void* h, const float* camPos, const float* camTarget, const float* camUp, float fov,
                      float aspect, int width, int height, int maxBounces, int shadowQuality, int filterType,
                      float filterWidth, uint64_t seed, size_t n, const int32_t* px, const int32_t* py,
                      const int32_t* sampleIndex, float* L) {
    auto* rs = static_cast<RefScene*>(h);
    const auto camPtr = makeCamera(camPos, camTarget, camUp, fov, aspect);
    const Camera& cam = *camPtr;
    PixelFilter pf;
    pf.type = static_cast<PixelFilterType>(filterType);
    pf.width = filterWidth;
    const PixelFilterSampler filter(pf);
    PathTracer integrator(maxBounces, 3, g_camera.aoStrength, shadowQuality);
    SobolSampler sampler(seed);
    for (size_t i = 0; i < n; ++i) {
        sampler.startPixel(px[i], py[i]);
        sampler.startSample(sampleIndex[i]);
        const Vec2f offset = filter.sample(sampler.get2D());
        const float sx = (static_cast<float>(px[i]) + 0.5f + offset.x) / static_cast<float>(width);
        const float sy = 1.0f - (static_cast<float>(py[i]) + 0.5f + offset.y) / static_cast<float>(height);
        Ray ray = cameraRay(cam, sampler, sx, sy);
        ray.coneWidth = cam.pixelConeWidth(height);
        ray.coneSpread = cam.pixelConeSpread(height);
        // tracePixel hands Li a primary-hit record whenever the background is not the environment
        // (renderer.cpp:158-161), which is what lets the background branch run (path_tracer.cpp:487).
        PathTracer::PrimaryHit primary;
        const bool wantPrimary = rs->scene->background().mode != Background::Mode::Environment;
        const Color3f c = wantPrimary ? integrator.Li(ray, *rs->scene, sampler, &primary, nullptr)
                                      : integrator.Li(ray, *rs->scene, sampler);
        L[3 * i] = c.r; L[3 * i + 1] = c.g; L[3 * i + 2] = c.b;
    }
    return 0;
}

// Per-sample radiance and render passes along tracePixel's path with wantPasses (renderer.cpp:162-171):
// PathTracer::Li with a LightPasses, then the primary hit's data passes. out, per sample: the eight
// light passes, the sixteen light groups, depth, position, idColor(nodeUidOf) and idColor(material index + 1)
// (the descriptor's index of the hit mesh's material, where the reference hashes its address), then the hit's
// shape and material index (-1 for a miss): 84 floats.
REF_API int ref_trace_passes(void* h, const float* camPos, const float* camTarget, const float* camUp, float fov,
                             float aspect, int width, int height, int maxBounces, int shadowQuality, int filterType,
                             float filterWidth, uint64_t seed, size_t n, const int32_t* px, const int32_t* py,
                             const int32_t* sampleIndex, float* L, float* out) {
    auto* rs = static_cast<RefScene*>(h);
    const auto camPtr = makeCamera(camPos, camTarget, camUp, fov, aspect);
    const Camera& cam = *camPtr;
    PixelFilter pf;
    pf.type = static_cast<PixelFilterType>(filterType);
    pf.width = filterWidth;
    const PixelFilterSampler filter(pf);
    PathTracer integrator(maxBounces, 3, g_camera.aoStrength, shadowQuality);
    SobolSampler sampler(seed);
    for (size_t i = 0; i < n; ++i) {
        sampler.startPixel(px[i], py[i]);
        sampler.startSample(sampleIndex[i]);
        const Vec2f offset = filter.sample(sampler.get2D());
        const float sx = (static_cast<float>(px[i]) + 0.5f + offset.x) / static_cast<float>(width);
        const float sy = 1.0f - (static_cast<float>(py[i]) + 0.5f + offset.y) / static_cast<float>(height);
        Ray ray = cameraRay(cam, sampler, sx, sy);
        ray.coneWidth = cam.pixelConeWidth(height);
        ray.coneSpread = cam.pixelConeSpread(height);
        PathTracer::PrimaryHit primary;
        LightPasses lp;
        for (auto& c : lp.c) c = Color3f(0.0f);
        for (auto& c : lp.lightGroups) c = Color3f(0.0f);
        const Color3f c = integrator.Li(ray, *rs->scene, sampler, &primary, &lp);
        L[3 * i] = c.r; L[3 * i + 1] = c.g; L[3 * i + 2] = c.b;
        float* o = out + 84 * i;
        for (int k = 0; k < static_cast<int>(LightPass::Count); ++k) {
            o[3 * k] = lp.c[k].r; o[3 * k + 1] = lp.c[k].g; o[3 * k + 2] = lp.c[k].b;
        }
        for (int g = 0; g < LightPasses::kMaxLightGroups; ++g) {
            o[24 + 3 * g] = lp.lightGroups[g].r; o[25 + 3 * g] = lp.lightGroups[g].g; o[26 + 3 * g] = lp.lightGroups[g].b;
        }
        std::fill(o + 72, o + 82, 0.0f);
        o[82] = o[83] = -1.0f;
        if (primary.hit) {
            o[72] = (primary.isect.point - ray.origin).length();
            o[73] = primary.isect.point.x; o[74] = primary.isect.point.y; o[75] = primary.isect.point.z;
            const Color3f obj = idColor(rs->scene->nodeUidOf(primary.isect));
            const auto it = rs->shapeIndex.find(primary.isect.hitObject);
            const int shape = it == rs->shapeIndex.end() ? -1 : it->second;
            const int mat = shape >= 0 && shape < static_cast<int>(rs->meshMaterial.size()) ? rs->meshMaterial[shape] : -1;
            const Color3f mc = idColor(mat >= 0 ? static_cast<uint64_t>(mat) + 1 : 0);
            o[76] = obj.r; o[77] = obj.g; o[78] = obj.b;
            o[79] = mc.r; o[80] = mc.g; o[81] = mc.b;
            o[82] = static_cast<float>(shape);
            o[83] = static_cast<float>(mat);
        }
    }
    return 0;
}

// Renderer::renderSamples into a fresh Film; writes the resolved mean (3 floats/px) and sample counts.
// Returns the wall time in seconds of renderSamples alone through *seconds.
REF_API int ref_render(void* h, const float* camPos, const float* camTarget, const float* camUp, float fov, float aspect,
               int width, int height, int maxBounces, int shadowQuality, int filterType, float filterWidth,
               int firstSample, int count, int threads, float* mean, uint32_t* counts, double* seconds) {
    auto* rs = static_cast<RefScene*>(h);
    const auto camPtr = makeCamera(camPos, camTarget, camUp, fov, aspect);
    const Camera& cam = *camPtr;
    RenderSettings s;
    s.width = width;
    s.height = height;
    s.samplesPerPixel = count;
    s.maxBounces = maxBounces;
    s.shadowQuality = shadowQuality;
    s.numThreads = threads;
    s.filter.type = static_cast<PixelFilterType>(filterType);
    s.filter.width = filterWidth;
    s.motionBlur = g_camera.motionBlur;
    s.aoStrength = g_camera.aoStrength;
    Film film(width, height);
    Renderer renderer;
    const auto t0 = std::chrono::steady_clock::now();
    const SampleBatchResult r = renderer.renderSamples(*rs->scene, cam, s, film, firstSample, count, nullptr);
    const auto t1 = std::chrono::steady_clock::now();
    if (seconds) *seconds = std::chrono::duration<double>(t1 - t0).count();
    for (int y = 0; y < height; ++y)
        for (int x = 0; x < width; ++x) {
            const size_t k = static_cast<size_t>(y) * width + x;
            const Color3f c = film.resolvedPixel(x, y);
            if (mean) { mean[3 * k] = c.r; mean[3 * k + 1] = c.g; mean[3 * k + 2] = c.b; }
            if (counts) counts[k] = static_cast<uint32_t>(film.sampleCount(x, y));
        }
    return r.completed ? 0 : 1;
}

// The reference's display path for a film given as sums and counts (Film::addBatch of each pixel into an empty
// Film): Film::resolve, applyImager with the settings of `ds`, then per pixel toneMap and quantizeUnorm8 with
// tpdfDither, as saveImagePNG does (image_io.cpp:61-80). display and rgb8 may be NULL.
REF_API int ref_display(int w, int h, const float* sum, const uint32_t* count, const pe_display* ds, float* display,
                        uint8_t* rgb8) {
    Film film(w, h);
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            const size_t k = static_cast<size_t>(y) * w + x;
            film.addBatch(x, y, Color3f(sum[3 * k], sum[3 * k + 1], sum[3 * k + 2]), count[k]);
        }
    Image img = film.resolve();
    if (ds->denoise) denoiseImage(img);  // main.cpp:277-289; softBlurDenoise without OIDN
    ImagerSettings is;
    is.bloomIntensity = ds->bloom_intensity;
    is.bloomRadius = ds->bloom_radius;
    is.bloomThreshold = ds->bloom_threshold;
    is.glareIntensity = ds->glare_intensity;
    is.glareBlades = ds->glare_blades;
    is.glareAngleDeg = ds->glare_angle_deg;
    is.glareLength = ds->glare_length;
    is.whiteBalanceK = ds->white_balance_k;
    is.tint = ds->tint;
    is.saturation = ds->saturation;
    is.contrast = ds->contrast;
    is.vignette = ds->vignette;
    is.chromaticAberration = ds->chromatic_aberration;
    applyImager(img, is, ds->exposure_ev);
    const auto op = static_cast<ToneMapOperator>(ds->tonemap);
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            const size_t k = static_cast<size_t>(y) * w + x;
            const Color3f d = toneMap(img.getPixel(x, y), op, ds->exposure_ev);
            for (int c = 0; c < 3; ++c) {
                if (display) display[3 * k + c] = d[c];
                if (rgb8) rgb8[3 * k + c] = quantizeUnorm8(d[c], ds->dither ? tpdfDither(x, y, c) : 0.0f);
            }
        }
    return 0;
}

// Renderer::renderSamples of samples [first, first + count) with AovFilms (box filter of width 1): writes the
// resolved beauty, albedo, normal and alpha images (Film::resolvedPixel: sum * (1 / n)).
REF_API int ref_render_aovs(void* h, const float* camPos, const float* camTarget, const float* camUp, float fov,
                            float aspect, int width, int height, int maxBounces, int firstSample, int count,
                            float* beauty, float* albedo, float* normal, float* alpha) {
    auto* rs = static_cast<RefScene*>(h);
    const auto camPtr = makeCamera(camPos, camTarget, camUp, fov, aspect);
    const Camera& cam = *camPtr;
    RenderSettings s;
    s.width = width;
    s.height = height;
    s.samplesPerPixel = firstSample + count;
    s.maxBounces = maxBounces;
    s.numThreads = 1;
    s.filter.type = static_cast<PixelFilterType>(0);
    s.filter.width = 1.0f;
    s.motionBlur = g_camera.motionBlur;
    s.aoStrength = g_camera.aoStrength;
    Film film(width, height);
    AovFilms aovs;
    aovs.resize(width, height);
    Renderer renderer;
    const SampleBatchResult r = renderer.renderSamples(*rs->scene, cam, s, film, firstSample, count, nullptr, &aovs);
    if (!r.completed) return 1;
    for (int y = 0; y < height; ++y)
        for (int x = 0; x < width; ++x) {
            const size_t k = static_cast<size_t>(y) * width + x;
            const Color3f b = film.resolvedPixel(x, y), a = aovs.albedo.resolvedPixel(x, y);
            const Color3f n = aovs.normal.resolvedPixel(x, y), c = aovs.alpha.resolvedPixel(x, y);
            const float v[12] = {b.r, b.g, b.b, a.r, a.g, a.b, n.r, n.g, n.b, c.r, c.g, c.b};
            std::copy(v, v + 3, beauty + 3 * k);
            std::copy(v + 3, v + 6, albedo + 3 * k);
            std::copy(v + 6, v + 9, normal + 3 * k);
            std::copy(v + 9, v + 12, alpha + 3 * k);
        }
    return 0;
}

// Renderer::renderSamples with adaptive sampling over consecutive batches [firsts[b], firsts[b] + counts[b])
// into one Film, as a RenderSession drives it (renderer.cpp:438-505); writes the film's resolved mean, sample
// counts and relative errors (Film::resolvedPixel / sampleCount / relativeError) and the samples traced.
REF_API int ref_render_adaptive(void* h, const float* camPos, const float* camTarget, const float* camUp, float fov,
                                float aspect, int width, int height, int maxBounces, int nbatch, const int* firsts,
                                const int* counts, float threshold, int minSpp, float* mean, uint32_t* countOut,
                                float* relErr, uint64_t* traced) {
    auto* rs = static_cast<RefScene*>(h);
    const auto camPtr = makeCamera(camPos, camTarget, camUp, fov, aspect);
    const Camera& cam = *camPtr;
    RenderSettings s;
    s.width = width;
    s.height = height;
    s.maxBounces = maxBounces;
    s.numThreads = 1;
    s.filter.type = static_cast<PixelFilterType>(0);  // box, width 1: the filter the oracle's tests use
    s.filter.width = 1.0f;
    s.motionBlur = g_camera.motionBlur;
    s.aoStrength = g_camera.aoStrength;
    s.adaptiveSampling = true;
    s.noiseThreshold = threshold;
    s.minAdaptiveSamples = minSpp;
    Film film(width, height);
    Renderer renderer;
    uint64_t total = 0;
    for (int b = 0; b < nbatch; ++b) {
        s.samplesPerPixel = firsts[b] + counts[b];
        const SampleBatchResult r = renderer.renderSamples(*rs->scene, cam, s, film, firsts[b], counts[b], nullptr);
        if (!r.completed) return 1;
        total += r.samplesTraced;
    }
    for (int y = 0; y < height; ++y)
        for (int x = 0; x < width; ++x) {
            const size_t k = static_cast<size_t>(y) * width + x;
            const Color3f c = film.resolvedPixel(x, y);
            mean[3 * k] = c.r;
            mean[3 * k + 1] = c.g;
            mean[3 * k + 2] = c.b;
            countOut[k] = static_cast<uint32_t>(film.sampleCount(x, y));
            relErr[k] = film.relativeError(x, y);
        }
    *traced = total;
    return 0;
}

}  // extern "C"

// The reference Scene behind a handle of ref_scene_build (for oracle/hipdev_driver.cpp, linked beside it).
std::shared_ptr<const Scene> ref_scene_shared(void* h) { return static_cast<RefScene*>(h)->scene; }

// The pass films of a RenderPassFilms in pe_pass_targets order (8 light passes, 16 groups, depth, position,
// object id, material id): film f's resolved pixel k at out[(f * pixels + k) * 3]; absent films stay zero.
void ref_copy_passes(const RenderPassFilms& pf, int width, int height, float* out) {
    const size_t pixels = static_cast<size_t>(width) * static_cast<size_t>(height);
    std::fill(out, out + 28 * pixels * 3, 0.0f);
    const Film* films[28] = {};
    for (int k = 0; k < 8; ++k) films[k] = &pf.light[k];
    for (size_t g = 0; g < pf.lightGroups.size() && g < 16; ++g) films[8 + g] = &pf.lightGroups[g].second;
    films[24] = &pf.depth;
    films[25] = &pf.position;
    films[26] = &pf.objectId;
    films[27] = &pf.materialId;
    for (int f = 0; f < 28; ++f) {
        if (!films[f] || films[f]->width() != width || films[f]->height() != height) continue;
        for (int y = 0; y < height; ++y)
            for (int x = 0; x < width; ++x) {
                const Color3f c = films[f]->resolvedPixel(x, y);
                float* o = out + (static_cast<size_t>(f) * pixels + static_cast<size_t>(y) * width + x) * 3;
                o[0] = c.r;
                o[1] = c.g;
                o[2] = c.b;
            }
    }
}

extern "C" {
// Renderer::renderSamples of samples [first, first + count) with RenderPassFilms (box filter of width 1,
// `threads` workers): the resolved beauty, the pass films (ref_copy_passes layout) and, with `crypto`, the
// cryptomatte ranks CryptoFilm::resolveRanks(3) gives (object layers 0-2, then material layers 0-2, each
// w * h RGBA); returns the number of light groups the renderer made, or -1.
REF_API int ref_render_passes(void* h, const float* camPos, const float* camTarget, const float* camUp, float fov,
                              float aspect, int width, int height, int maxBounces, int firstSample, int count,
                              int threads, float* beauty, float* passes, float* crypto) {
    auto* rs = static_cast<RefScene*>(h);
    const auto camPtr = makeCamera(camPos, camTarget, camUp, fov, aspect);
    const Camera& cam = *camPtr;
    RenderSettings s;
    s.width = width;
    s.height = height;
    s.samplesPerPixel = firstSample + count;
    s.maxBounces = maxBounces;
    s.numThreads = threads;
    s.filter.type = static_cast<PixelFilterType>(0);
    s.filter.width = 1.0f;
    s.motionBlur = g_camera.motionBlur;
    s.aoStrength = g_camera.aoStrength;
    Film film(width, height);
    RenderPassFilms pf;
    pf.resize(width, height);
    Renderer renderer;
    const SampleBatchResult r = renderer.renderSamples(*rs->scene, cam, s, film, firstSample, count, nullptr, nullptr, &pf);
    if (!r.completed) return -1;
    for (int y = 0; y < height; ++y)
        for (int x = 0; x < width; ++x) {
            const size_t k = static_cast<size_t>(y) * width + x;
            const Color3f b = film.resolvedPixel(x, y);
            beauty[3 * k] = b.r;
            beauty[3 * k + 1] = b.g;
            beauty[3 * k + 2] = b.b;
        }
    ref_copy_passes(pf, width, height, passes);
    if (crypto) {
        const size_t layer = static_cast<size_t>(width) * height * 4;
        const std::vector<CryptoRankImage> ranks[2] = {pf.cryptoObject.resolveRanks(3), pf.cryptoMaterial.resolveRanks(3)};
        for (int k = 0; k < 2; ++k)
            for (size_t j = 0; j < ranks[k].size() && j < 3; ++j)
                std::copy(ranks[k][j].rgba.begin(), ranks[k][j].rgba.end(), crypto + (3 * k + j) * layer);
    }
    return static_cast<int>(pf.lightGroups.size());
}
}  // extern "C"

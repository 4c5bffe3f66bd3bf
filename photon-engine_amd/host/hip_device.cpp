// hip_device.cpp — HipDevice: the reference's RenderDevice interface on top of libphoton_hip.so.
//
// setScene packs the immutable Scene into one pe_scene_desc (engine/scene.h:31-165 walked through its public
// getters) and uploads it once; render() maps the camera (Camera::describeGpu, camera/camera.h:17-31) and the
// settings onto pe_camera / pe_settings, renders the sample range into a device film, and merges the film's
// per-pixel sums into the caller's Film with Film::addBatch (core/image/film.cpp:119-137), as the Vulkan
// device merges its batches.
//
// What the HIP path does not implement is refused, never approximated: setScene leaves the device without a
// scene (hasScene() false, lastError() says why) for shapes other than TriangleMesh / Triangle / static
// MeshInstance, visibility masks other than all rays or all but camera rays, shadow catchers, lights other
// than AreaLight / TriangleLight, light linking, rounded edges, textured Disney materials and
// materials other than Disney / Lambertian / Dielectric. Pinhole, thin-lens and orthographic cameras render, with
// motion blur of the camera (moving instances are refused), and contact AO.
// Path guiding uses the library's learned guide in place of the reference's SD-tree.
#include "hip_device.h"

#include "camera/camera.h"
#include "core/image/cryptomatte.h"
#include "core/image/film.h"
#include "core/image/image.h"
#include "engine/renderer.h"
#include "engine/scene.h"
#include "geometry/mesh.h"
#include "geometry/mesh_instance.h"
#include "geometry/triangle.h"
#include "lights/area_light.h"
#include "lights/environment_light.h"
#include "materials/dielectric.h"
#include "materials/disney.h"
#include "materials/lambertian.h"
#include "materials/mirror.h"

#include <algorithm>
#include <unordered_map>

namespace photon {

namespace {

void put3(float* d, const Vec3f& v) {
    d[0] = v.x;
    d[1] = v.y;
    d[2] = v.z;
}

void putc(float* d, const Color3f& c) {
    d[0] = c.r;
    d[1] = c.g;
    d[2] = c.b;
}

std::string lastHipError() {
    const char* e = pe_last_error();
    return e ? std::string(e) : std::string();
}

// The flattened scene and the host arrays its pointers refer to.
struct PackedScene {
    std::vector<pe_mesh_desc> meshes;
    std::vector<pe_material_desc> materials;
    std::vector<pe_light_desc> lights;
    std::vector<std::vector<float>> floats;
    std::vector<std::vector<uint32_t>> indices;
    pe_env_desc env{};
    bool hasEnv = false;
    std::vector<int32_t> lightGroups;  // Scene::lightGroupIndex per light (empty: the scene has no groups)
    int32_t envGroup = -1;
    int32_t numGroups = 0;
    std::vector<uint64_t> nodeUids;    // per shape: Scene::nodeUidOf
    std::vector<const Material*> materialPtrs;  // per packed material: the scene's material
    int32_t backgroundMode = PE_BACKGROUND_ENVIRONMENT;
    float backgroundColor[3] = {0.18f, 0.18f, 0.18f};

    const float* keep(std::vector<float> v) {
        floats.push_back(std::move(v));
        return floats.back().data();
    }
    pe_scene_desc desc() const {
        pe_scene_desc d{};
        d.meshes = meshes.data();
        d.num_meshes = static_cast<int32_t>(meshes.size());
        d.materials = materials.data();
        d.num_materials = static_cast<int32_t>(materials.size());
        d.lights = lights.data();
        d.num_lights = static_cast<int32_t>(lights.size());
        d.environment = hasEnv ? &env : nullptr;
        d.light_groups = lightGroups.empty() ? nullptr : lightGroups.data();
        d.env_light_group = envGroup;
        d.num_light_groups = lightGroups.empty() ? 0 : numGroups;
        d.node_uids = nodeUids.data();
        d.background_mode = backgroundMode;
        std::copy(backgroundColor, backgroundColor + 3, d.background_color);
        return d;
    }
};

bool packMaterial(const Material* m, pe_material_desc& o, std::string& err) {
    o = pe_material_desc{};
    if (m->roundEdgeRadius() > 0.0f) {  // roundedEdgeNormal (path_tracer.cpp:612-615)
        err = "rounded edges (Material::roundEdgeRadius) are not implemented on the HIP device";
        return false;
    }
    SurfaceInteraction si;
    putc(o.emission, m->emitted(si));
    if (auto* d = dynamic_cast<const DisneyMaterial*>(m)) {
        if (d->albedoImage() || d->normalImage() || d->roughnessImage() || d->metalnessImage() || d->opacityImage()) {
            err = "DisneyMaterial with texture maps is not implemented on the HIP device";
            return false;
        }
        if (d->flakeDensity() > 0.0f) {  // metallic flakes (setFlakes, disney.h:120-125)
            o.flake_density = d->flakeDensity();
            o.flake_size = d->flakeSize();
            o.flake_roughness = d->flakeRoughness();
            putc(o.flake_color, d->flakeColor());
        }
        if (d->baseOpacity() < 1.0f) {  // a cutout (setOpacity, disney.h:147-158)
            o.opacity = d->baseOpacity();
            o.has_opacity = 1;
        }
        o.type = PE_MAT_DISNEY;
        putc(o.base_color, d->baseColor());
        o.metallic = d->metallic();
        o.roughness = d->roughness();
        o.specular = d->specular();
        o.clearcoat = d->clearCoat();
        o.clearcoat_roughness = d->clearCoatRoughness();
        o.coat_ior = d->coatIor();
        putc(o.coat_tint, d->coatTint());
        o.ior = d->ior();
        o.specular_transmission = d->specularTransmission();
        o.anisotropy = d->anisotropy();
        o.sheen = d->sheen();
        o.sheen_roughness = d->sheenRoughness();
        putc(o.sheen_color, d->sheenColor());
        o.has_sheen_color = d->hasCustomSheenColor() ? 1 : 0;
        o.diffuse_roughness = d->diffuseRoughness();
        o.diffuse_transmission = d->diffuseTransmission();
        o.subsurface = d->subsurface();
        putc(o.subsurface_radius, d->subsurfaceRadius());
        o.subsurface_scale = d->subsurfaceScale();
        o.subsurface_anisotropy = d->subsurfaceAnisotropy();
        putc(o.edge_tint, d->edgeTint());
        o.thin_film_thickness = d->thinFilmThickness();
        o.thin_film_ior = d->thinFilmIor();
        return true;
    }
    if (auto* l = dynamic_cast<const Lambertian*>(m)) {
        o.type = PE_MAT_LAMBERT;
        putc(o.base_color, l->albedo());
        return true;
    }
    if (auto* r = dynamic_cast<const Mirror*>(m)) {
        o.type = PE_MAT_MIRROR;
        putc(o.base_color, r->reflectance());
        return true;
    }
    if (auto* g = dynamic_cast<const Dielectric*>(m)) {
        o.type = PE_MAT_DIELECTRIC;
        putc(o.base_color, g->tint());
        o.ior = g->ior();
        o.roughness = g->roughness();
        o.abbe = g->abbe();
        putc(o.absorption_color, g->absorptionColor());
        o.absorption_distance = g->absorptionDistance();
        o.thin_walled = g->thinWalled() ? 1 : 0;
        o.thin_film_thickness = g->thinFilmThickness();
        o.scatter_density = g->scatterDensity();  // setScattering (dielectric.h:61-66)
        putc(o.scatter_color, g->scatterColor());
        o.scatter_anisotropy = g->scatterAnisotropy();
        return true;
    }
    err = "material type not implemented on the HIP device";
    return false;
}

bool packScene(const Scene& scene, PackedScene& p, std::string& err) {
    std::unordered_map<const Material*, int> matIndex;
    auto material = [&](const Material* m) -> int {
        if (!m) return -1;
        auto it = matIndex.find(m);
        if (it != matIndex.end()) return it->second;
        pe_material_desc d;
        if (!packMaterial(m, d, err)) return -1;
        p.materials.push_back(d);
        p.materialPtrs.push_back(m);
        const int k = static_cast<int>(p.materials.size()) - 1;
        matIndex.emplace(m, k);
        return k;
    };

    // Lights in Scene::lights() order: the light index is part of the sampling sequence.
    for (const Light* L : scene.lights()) {
        pe_light_desc d{};
        if (!L->lightLink().isDefault()) {  // Scene::isLightLinked / isShadowLinked (scene.cpp:242-256)
            err = "light linking is not implemented on the HIP device";
            return false;
        }
        if (auto* a = dynamic_cast<const AreaLight*>(L)) {
            d.type = PE_LIGHT_AREA;
            put3(d.p0, a->position());
            put3(d.p1, a->u());
            put3(d.p2, a->v());
            putc(d.radiance, a->radiance());
            d.visible_to_camera = a->visibleToCamera() ? 1 : 0;
        } else if (auto* t = dynamic_cast<const TriangleLight*>(L)) {
            d.type = PE_LIGHT_TRIANGLE;
            put3(d.p0, t->p0());
            put3(d.p1, t->p1());
            put3(d.p2, t->p2());
            putc(d.radiance, t->radiance());
            d.visible_to_camera = 1;
        } else {
            err = "light type not implemented on the HIP device (AreaLight and emissive triangles are)";
            return false;
        }
        p.lights.push_back(d);
    }
    // Light groups as Scene::buildAccelerator resolved them (scene.cpp:144-173).
    if (!scene.lightGroupNames().empty()) {
        p.numGroups = static_cast<int32_t>(std::min<size_t>(scene.lightGroupNames().size(), PE_MAX_LIGHT_GROUPS));
        for (size_t i = 0; i < p.lights.size(); ++i) p.lightGroups.push_back(scene.lightGroupIndex(static_cast<int>(i)));
        p.envGroup = scene.envLightGroupIndex();
    }

    for (const auto& shape : scene.shapes()) {
        SurfaceInteraction probe;
        probe.hitObject = shape.get();
        probe.primIndex = 0;
        pe_mesh_desc md{};
        // Ray visibility: every ray, or every ray but camera rays (what a camera-invisible area light's quads
        // carry, scene.cpp:93-96); the per-bounce masks of path_tracer.cpp:326-467 are not implemented.
        const uint8_t vis = shape->visibilityMask();
        if (vis != VisAll && vis != static_cast<uint8_t>(VisAll & ~VisCamera)) {
            err = "shape visibility masks other than all rays or all but camera rays are not implemented on the "
                  "HIP device";
            return false;
        }
        if (scene.isShadowCatcher(probe)) {
            err = "shadow catchers are not implemented on the HIP device";
            return false;
        }
        if (vis != VisAll) md.flags |= PE_SHAPE_CAMERA_INVISIBLE;
        if (auto* inst = dynamic_cast<const MeshInstance*>(shape.get())) {
            // A static instance (SceneGraph::compileNode, scene_graph.cpp:189-202, makes one per mesh node) is
            // baked into world-space triangles, as the Vulkan device bakes it (vulkan_device.cpp:1409-1460):
            // points through transformPoint, normals through transformNormal and normalised. A mirroring
            // transform (negative determinant) turns the winding, so two indices of every triangle swap and the
            // edge cross product keeps pointing along the transformed outward normal, which is the geometric
            // normal MeshInstance reports (mesh_instance.cpp:57-70). The instance's emissive triangles are the
            // world-space TriangleLights Scene::addShape made from the same transformPoint values.
            if (inst->hasMotion()) {
                err = "a MeshInstance with a motion transform is not implemented on the HIP device";
                return false;
            }
            const TriangleMesh* mesh = inst->mesh().get();
            const Transform& xf = inst->transform0();
            std::vector<float> pos, nrm, uv;
            std::vector<uint32_t> idx;
            if (mesh) {
                for (const Vec3f& v : mesh->positions()) {
                    const Vec3f w = xf.transformPoint(v);
                    pos.insert(pos.end(), {w.x, w.y, w.z});
                }
                for (const Vec3f& v : mesh->normals()) {
                    Vec3f n = xf.transformNormal(v);
                    if (n.lengthSquared() > 0.0f) n = n.normalized();
                    nrm.insert(nrm.end(), {n.x, n.y, n.z});
                }
                for (const Vec2f& v : mesh->uvs()) uv.insert(uv.end(), {v.x, v.y});
                idx = mesh->indices();
                const Mat4f& m = xf.matrix();
                const float det = m(0, 0) * (m(1, 1) * m(2, 2) - m(1, 2) * m(2, 1)) -
                                  m(0, 1) * (m(1, 0) * m(2, 2) - m(1, 2) * m(2, 0)) +
                                  m(0, 2) * (m(1, 0) * m(2, 1) - m(1, 1) * m(2, 0));
                if (det < 0.0f)
                    for (size_t i = 0; i + 2 < idx.size(); i += 3) std::swap(idx[i + 1], idx[i + 2]);
            }
            if (!nrm.empty() && nrm.size() != pos.size()) nrm.clear();
            if (!uv.empty() && uv.size() / 2 != pos.size() / 3) uv.clear();
            p.indices.push_back(std::move(idx));
            md.num_vertices = static_cast<uint32_t>(pos.size() / 3);
            md.num_triangles = static_cast<uint32_t>(p.indices.back().size() / 3);
            md.positions = p.keep(std::move(pos));
            md.normals = nrm.empty() ? nullptr : p.keep(std::move(nrm));
            md.uvs = uv.empty() ? nullptr : p.keep(std::move(uv));
            md.indices = p.indices.back().data();
            md.material = material(inst->material());
            md.emitter_base = scene.emitterIndex(probe);
            md.standalone = 0;
        } else if (auto* mesh = dynamic_cast<const TriangleMesh*>(shape.get())) {
            std::vector<float> pos, nrm, uv;
            for (const Vec3f& v : mesh->positions()) pos.insert(pos.end(), {v.x, v.y, v.z});
            for (const Vec3f& v : mesh->normals()) nrm.insert(nrm.end(), {v.x, v.y, v.z});
            for (const Vec2f& v : mesh->uvs()) uv.insert(uv.end(), {v.x, v.y});
            p.indices.push_back(mesh->indices());
            md.num_vertices = static_cast<uint32_t>(mesh->positions().size());
            md.num_triangles = static_cast<uint32_t>(mesh->numTriangles());
            md.positions = p.keep(std::move(pos));
            md.normals = nrm.empty() ? nullptr : p.keep(std::move(nrm));
            md.uvs = uv.empty() ? nullptr : p.keep(std::move(uv));
            md.indices = p.indices.back().data();
            md.material = material(mesh->material());
            md.emitter_base = scene.emitterIndex(probe);
            md.standalone = 0;
        } else if (auto* tri = dynamic_cast<const Triangle*>(shape.get())) {
            const bool normals = tri->n0().lengthSquared() > 0.0f && tri->n1().lengthSquared() > 0.0f &&
                                 tri->n2().lengthSquared() > 0.0f;
            md.num_vertices = 3;
            md.num_triangles = 1;
            md.positions = p.keep({tri->v0().x, tri->v0().y, tri->v0().z, tri->v1().x, tri->v1().y, tri->v1().z,
                                   tri->v2().x, tri->v2().y, tri->v2().z});
            md.normals = normals ? p.keep({tri->n0().x, tri->n0().y, tri->n0().z, tri->n1().x, tri->n1().y,
                                           tri->n1().z, tri->n2().x, tri->n2().y, tri->n2().z})
                                 : nullptr;
            md.uvs = p.keep({tri->uv0().x, tri->uv0().y, tri->uv1().x, tri->uv1().y, tri->uv2().x, tri->uv2().y});
            p.indices.push_back({0u, 1u, 2u});
            md.indices = p.indices.back().data();
            md.material = material(tri->material());
            probe.primIndex = -1;
            md.emitter_base = scene.emitterIndex(probe);
            md.standalone = 1;
            if (scene.cameraInvisible(probe)) md.flags |= PE_SHAPE_CAMERA_INVISIBLE;
        } else {
            err = "shape type not implemented on the HIP device (TriangleMesh and Triangle are)";
            return false;
        }
        if (md.material < 0) {
            if (err.empty()) err = "shape without a material";
            return false;
        }
        p.meshes.push_back(md);
        p.nodeUids.push_back(scene.nodeUidOf(probe));
    }

    // Scene::background(): Environment, Color or Transparent (scene.h:24-28).
    const Background& bg = scene.background();
    p.backgroundMode = bg.mode == Background::Mode::Color         ? PE_BACKGROUND_COLOR
                       : bg.mode == Background::Mode::Transparent ? PE_BACKGROUND_TRANSPARENT
                                                                  : PE_BACKGROUND_ENVIRONMENT;
    putc(p.backgroundColor, bg.color);

    if (const EnvironmentLight* env = scene.environment()) {
        p.hasEnv = true;
        pe_env_desc& e = p.env;
        e = pe_env_desc{};
        e.rotation = env->rotation();
        e.intensity = env->intensity();
        if (const auto& img = env->image()) {
            std::vector<float> rgb;
            rgb.reserve(3ull * static_cast<size_t>(img->width()) * static_cast<size_t>(img->height()));
            for (int y = 0; y < img->height(); ++y)
                for (int x = 0; x < img->width(); ++x) {
                    const Color3f c = img->getPixel(x, y);
                    rgb.insert(rgb.end(), {c.r, c.g, c.b});
                }
            e.width = img->width();
            e.height = img->height();
            e.rgb = p.keep(std::move(rgb));
        }
        if (env->hasSun()) {
            e.has_sun = 1;
            put3(e.sun_direction, env->sunDirection());
            e.sun_cos_max = env->sunCosMax();
            // The light stores the sun radiance with its intensity applied and the direction normalised
            // (setSun, environment_light.cpp:168-171); the descriptor takes setSun's arguments, so both are
            // handed back as setSun's inputs (exact for intensity 1, otherwise within a rounding).
            const float k = env->intensity() != 0.0f ? 1.0f / env->intensity() : 0.0f;
            putc(e.sun_radiance, env->sunRadiance() * k);
        }
    }
    return true;
}

// The materials of a shape in the order the cryptomatte lookup registers them (renderer.cpp:101-109): an
// instance's own (or its mesh's), then its mesh's.
std::vector<const Material*> shapeMaterials(const Shape* shape) {
    if (auto* inst = dynamic_cast<const MeshInstance*>(shape))
        return {inst->material(), inst->mesh() ? inst->mesh()->material() : nullptr};
    if (auto* mesh = dynamic_cast<const TriangleMesh*>(shape)) return {mesh->material()};
    if (auto* tri = dynamic_cast<const Triangle*>(shape)) return {tri->material()};
    return {};
}

// buildCryptoLookup (renderer.cpp:65-118) for the packed scene: names registered into the pass films'
// manifests; the id of every shape (scene.shapes() order = the packed meshes) and of every packed material.
void cryptoTables(const Scene& scene, const std::vector<const Material*>& materials, RenderPassFilms& passes,
                  std::vector<float>& objectIds, std::vector<float>& materialIds) {
    std::unordered_map<const Material*, float> mats;
    size_t anonShape = 0, anonMat = 0;
    auto registerMat = [&](const Material* m) {
        if (!m || mats.count(m)) return;
        const auto it = scene.materialNames().find(m);
        const std::string name = it != scene.materialNames().end() && !it->second.empty()
                                     ? it->second
                                     : "material_" + std::to_string(anonMat++);
        mats[m] = passes.cryptoMaterial.registerName(name);
    };
    for (const auto& [m, name] : scene.materialNames())
        if (m && !name.empty()) mats[m] = passes.cryptoMaterial.registerName(name);
    objectIds.clear();
    std::unordered_map<const void*, bool> seen;
    for (const auto& shape : scene.shapes()) {
        const void* ptr = shape.get();
        std::string name;
        const auto nit = scene.shapeNames().find(ptr);
        const auto oit = scene.shapeOwners().find(ptr);
        if (nit != scene.shapeNames().end() && !nit->second.empty()) name = nit->second;
        else if (oit != scene.shapeOwners().end() && oit->second != 0) name = "object_" + std::to_string(oit->second);
        else name = "shape_" + std::to_string(anonShape++);
        objectIds.push_back(passes.cryptoObject.registerName(name));
        seen[ptr] = true;
        for (const Material* m : shapeMaterials(shape.get())) registerMat(m);
    }
    for (const auto& [ptr, name] : scene.shapeNames())
        if (ptr && !name.empty() && !seen.count(ptr)) passes.cryptoObject.registerName(name);
    materialIds.assign(materials.size(), 0.0f);
    for (size_t i = 0; i < materials.size(); ++i) {
        const auto it = mats.find(materials[i]);
        if (it != mats.end()) materialIds[i] = it->second;
    }
}

}  // namespace

std::vector<DeviceInfo> enumerateHipDevices() {
    std::vector<DeviceInfo> out;
    int n = 0;
    if (pe_device_count(&n) != PE_OK) return out;
    for (int i = 0; i < n; ++i) {
        pe_device_info pi{};
        if (pe_device_info_get(i, &pi) != PE_OK) continue;
        DeviceInfo d;
        d.id = pi.id;
        // The reference's DeviceKind has no HIP value; Cuda is its GPU-compute kind until it gains one.
        d.kind = DeviceKind::Cuda;
        d.name = pi.name;
        d.memoryBytes = pi.memory_bytes;
        d.usable = pi.usable != 0;
        d.note = std::string(pi.arch) + ", " + std::to_string(pi.compute_units) + " CUs" +
                 (pi.note[0] ? std::string(": ") + pi.note : std::string());
        out.push_back(std::move(d));
    }
    return out;
}

std::unique_ptr<HipDevice> HipDevice::create(int ordinal, std::string* error) {
    std::unique_ptr<HipDevice> dev(new HipDevice());
    pe_device_info pi{};
    if (pe_device_info_get(ordinal, &pi) != PE_OK || pe_device_create(ordinal, &dev->m_dev) != PE_OK) {
        if (error) *error = lastHipError();
        return nullptr;
    }
    dev->m_info.id = pi.id;
    dev->m_info.kind = DeviceKind::Cuda;
    dev->m_info.name = pi.name;
    dev->m_info.memoryBytes = pi.memory_bytes;
    dev->m_info.usable = true;
    dev->m_info.note = std::string(pi.arch) + ", " + std::to_string(pi.compute_units) + " CUs";
    return dev;
}

HipDevice::~HipDevice() {
    for (DeviceFilm& f : m_films)
        if (f.film) pe_film_destroy(f.film);
    for (DeviceCrypto& c : m_crypto)
        if (c.film) pe_crypto_film_destroy(c.film);
    if (m_guide) pe_guide_destroy(m_guide);
    if (m_scene) pe_scene_release(m_scene);
    if (m_dev) pe_device_destroy(m_dev);
}

void HipDevice::setScene(std::shared_ptr<const Scene> scene) {
    if (m_guide) {  // a guide belongs to its scene
        pe_guide_destroy(m_guide);
        m_guide = nullptr;
        m_guideScene = nullptr;
    }
    if (m_scene) {
        pe_scene_release(m_scene);
        m_scene = nullptr;
    }
    m_sceneRef.reset();
    m_error.clear();
    if (!scene) return;
    PackedScene packed;
    std::string err;
    if (!packScene(*scene, packed, err)) {
        m_error = "HipDevice::setScene: " + err;
        return;
    }
    const pe_scene_desc desc = packed.desc();
    m_materialPtrs = packed.materialPtrs;
    if (pe_scene_upload(m_dev, &desc, &m_scene) != PE_OK) {
        m_error = "HipDevice::setScene: " + lastHipError();
        m_scene = nullptr;
        return;
    }
    m_sceneRef = std::move(scene);
}

void HipDevice::updateScene(std::shared_ptr<const Scene> scene, SceneUpdateKind kind) {
    if (!m_scene || !scene || kind == SceneUpdateKind::Full || kind == SceneUpdateKind::Transforms) {
        setScene(std::move(scene));
        return;
    }
    PackedScene packed;
    std::string err;
    if (!packScene(*scene, packed, err)) {
        setScene(nullptr);
        m_error = "HipDevice::updateScene: " + err;
        return;
    }
    const pe_scene_desc desc = packed.desc();
    const int k = kind == SceneUpdateKind::Materials ? PE_UPDATE_MATERIALS : PE_UPDATE_LIGHTS;
    if (pe_scene_update(m_scene, k, &desc) != PE_OK) {
        // An update the library cannot apply in place becomes a full upload.
        setScene(std::move(scene));
        return;
    }
    m_materialPtrs = packed.materialPtrs;
    m_sceneRef = std::move(scene);
}

SampleBatchResult HipDevice::render(const Camera& camera, const RenderSettings& settings, RenderTargets& targets,
                                    int firstSample, int count, const std::atomic<bool>* cancel) {
    SampleBatchResult res{false, 0};
    m_lastRays = 0;
    if (!m_scene || !targets.beauty) return res;
    Film& film = *targets.beauty;
    const CameraGpuDesc cd = camera.describeGpu();
    // Path guiding (renderer.cpp:413-429): a learned guide per scene, made afresh when a render starts at sample
    // 0 or the scene changed, trained during its first launches (the reference trains its SD-tree over the
    // first 32 samples, then freezes it), dropped when guiding is switched off.
    if (settings.pathGuiding) {
        if (!m_guide || m_guideScene != m_scene || firstSample == 0) {
            if (m_guide) pe_guide_destroy(m_guide);
            m_guide = nullptr;
            pe_guide_config gc{};  // the library's defaults (learning rate 1e-3, mixture fraction 0.5)
            gc.train_launches = 4;
            if (pe_guide_create(m_dev, m_scene, &gc, &m_guide) != PE_OK) {
                m_error = "HipDevice::render: " + lastHipError();
                m_guide = nullptr;
                return res;
            }
            m_guideScene = m_scene;
        }
    } else if (m_guide) {
        pe_guide_destroy(m_guide);
        m_guide = nullptr;
        m_guideScene = nullptr;
    }
    // The camera as CameraGpuDesc (camera.h:19-31) of both shutter poses; the library draws the rays as
    // Camera::generateRay does for a pinhole, thin-lens or orthographic camera (camera.h:57-91).
    pe_camera cam{};
    put3(cam.position, cd.position);
    put3(cam.lower_left, cd.lowerLeft);
    put3(cam.horizontal, cd.horizontal);
    put3(cam.vertical, cd.vertical);
    put3(cam.u, cd.u);
    put3(cam.v, cd.v);
    put3(cam.forward, cd.forward);
    cam.lens_radius = cd.lensRadius;
    cam.blades = cd.blades;
    cam.blade_rotation = cd.bladeRotation;
    cam.orthographic = cd.orthographic ? 1 : 0;
    if (camera.hasShutterEnd()) {
        const CameraGpuDesc ce = camera.describeGpuEnd();
        cam.has_shutter_end = 1;
        put3(cam.end_position, ce.position);
        put3(cam.end_lower_left, ce.lowerLeft);
        put3(cam.end_horizontal, ce.horizontal);
        put3(cam.end_vertical, ce.vertical);
        put3(cam.end_u, ce.u);
        put3(cam.end_v, ce.v);
        put3(cam.end_forward, ce.forward);
        cam.end_lens_radius = ce.lensRadius;
    }
    pe_settings st{};
    st.max_bounces = settings.maxBounces;
    st.rr_depth = 3;                                   // Renderer::renderSamples (renderer.cpp:432)
    st.shadow_quality = settings.shadowQuality;
    st.filter_type = static_cast<int32_t>(settings.filter.type);
    st.filter_width = settings.filter.width;
    st.clamp_indirect = settings.clampIndirect;
    st.seed = 12345;                                   // createRenderSampler default (renderer.h:37)
    st.samples_per_launch = 0;                         // the library sizes launches for the device
    st.adaptive = settings.adaptiveSampling ? 1 : 0;   // renderer.cpp:438-505
    st.motion_blur = settings.motionBlur ? 1 : 0;      // a shutter time per camera ray (renderer.cpp:152-153)
    st.ao_strength = settings.aoStrength;              // contact AO at the first vertex (renderer.cpp:432)
    st.noise_threshold = settings.noiseThreshold;
    st.min_adaptive_samples = settings.minAdaptiveSamples;
    const bool adaptive = settings.adaptiveSampling;

    const int w = film.width(), h = film.height();
    // The caller's films: beauty and, when it asks for them, the AOVs (renderer.cpp:222-227) and the render
    // passes (light passes, the scene's light groups, depth, position, object and material id; renderer.cpp:
    // 207-219, with the groups named as the scene's, :400-405); cryptomatte below.
    Film* targets_of[kDeviceFilms] = {&film};
    if (targets.aovs) {
        AovFilms& a = *targets.aovs;
        if (a.albedo.width() != w || a.albedo.height() != h) a.resize(w, h);
        targets_of[1] = &a.albedo;
        targets_of[2] = &a.normal;
        targets_of[3] = &a.alpha;
    }
    if (targets.passes) {
        RenderPassFilms& pf = *targets.passes;
        if (pf.depth.width() != w || pf.depth.height() != h) pf.resize(w, h);
        const std::vector<std::string>& names = m_sceneRef->lightGroupNames();
        bool same = pf.lightGroups.size() == names.size();
        for (size_t i = 0; same && i < names.size(); ++i)
            same = pf.lightGroups[i].first == names[i] && pf.lightGroups[i].second.width() == w &&
                   pf.lightGroups[i].second.height() == h;
        if (!same) pf.initLightGroups(w, h, names);
        for (int k = 0; k < PE_LIGHT_PASSES; ++k) targets_of[kPassFilm0 + k] = &pf.light[k];
        for (size_t g = 0; g < pf.lightGroups.size() && g < PE_MAX_LIGHT_GROUPS; ++g)
            targets_of[kPassFilm0 + PE_LIGHT_PASSES + g] = &pf.lightGroups[g].second;
        targets_of[kPassFilm0 + 24] = &pf.depth;
        targets_of[kPassFilm0 + 25] = &pf.position;
        targets_of[kPassFilm0 + 26] = &pf.objectId;
        targets_of[kPassFilm0 + 27] = &pf.materialId;
    }
    if (m_filmW != w || m_filmH != h) {
        for (DeviceFilm& f : m_films) {
            if (f.film) pe_film_destroy(f.film);
            f.film = nullptr;
        }
        for (DeviceCrypto& c : m_crypto) {
            if (c.film) pe_crypto_film_destroy(c.film);
            c.film = nullptr;
        }
        m_filmW = w;
        m_filmH = h;
        m_accumulating = false;
    }
    const size_t pixels = static_cast<size_t>(w) * static_cast<size_t>(h);
    const bool restart = !adaptive || firstSample == 0 || !m_accumulating;
    for (int k = 0; k < kDeviceFilms; ++k) {
        if (!targets_of[k]) continue;
        DeviceFilm& f = m_films[k];
        if (!f.film) {
            if (pe_film_create(m_dev, w, h, &f.film) != PE_OK) {
                m_error = "HipDevice::render: " + lastHipError();
                return res;
            }
            if (!restart) {  // an AOV film joining an accumulating session starts from this call
                f.prevSum.assign(3 * pixels, 0.0f);
                f.prevCount.assign(pixels, 0u);
            }
        }
        if (restart) {
            if (pe_film_clear(f.film) != PE_OK) {
                m_error = "HipDevice::render: " + lastHipError();
                return res;
            }
            f.prevSum.assign(3 * pixels, 0.0f);
            f.prevCount.assign(pixels, 0u);
        }
    }
    // Cryptomatte (renderer.cpp:408-411, :215-216): the ids of this scene's names, counted on the device.
    std::vector<float> objectIds, materialIds;
    if (targets.passes) {
        RenderPassFilms& pf = *targets.passes;
        if (pf.cryptoObject.width() != w || pf.cryptoObject.height() != h) {
            pf.cryptoObject.resize(w, h);
            pf.cryptoMaterial.resize(w, h);
        }
        cryptoTables(*m_sceneRef, m_materialPtrs, pf, objectIds, materialIds);
        for (DeviceCrypto& c : m_crypto) {
            if (!c.film) {
                if (pe_crypto_film_create(m_dev, w, h, &c.film) != PE_OK) {
                    m_error = "HipDevice::render: " + lastHipError();
                    return res;
                }
                c.prevCounts.assign(PE_CRYPTO_SLOTS * pixels, 0u);
                c.prevTotals.assign(pixels, 0u);
            }
            if (restart) {
                if (pe_crypto_film_clear(c.film) != PE_OK) {
                    m_error = "HipDevice::render: " + lastHipError();
                    return res;
                }
                c.prevCounts.assign(PE_CRYPTO_SLOTS * pixels, 0u);
                c.prevTotals.assign(pixels, 0u);
            }
        }
    }
    if (restart) m_accumulating = adaptive;
    pe_targets tg{};
    tg.beauty = m_films[0].film;
    if (targets_of[1]) {
        tg.albedo = m_films[1].film;
        tg.normal = m_films[2].film;
        tg.alpha = m_films[3].film;
    }
    pe_pass_targets pt{};
    if (targets.passes) {
        for (int k = 0; k < PE_LIGHT_PASSES; ++k) pt.light[k] = m_films[kPassFilm0 + k].film;
        for (int g = 0; g < PE_MAX_LIGHT_GROUPS; ++g)
            pt.light_groups[g] = targets_of[kPassFilm0 + PE_LIGHT_PASSES + g] ? m_films[kPassFilm0 + PE_LIGHT_PASSES + g].film
                                                                               : nullptr;
        pt.depth = m_films[kPassFilm0 + 24].film;
        pt.position = m_films[kPassFilm0 + 25].film;
        pt.object_id = m_films[kPassFilm0 + 26].film;
        pt.material_id = m_films[kPassFilm0 + 27].film;
        pt.crypto_object = m_crypto[0].film;
        pt.crypto_material = m_crypto[1].film;
        pt.object_ids = objectIds.data();
        pt.num_object_ids = static_cast<int32_t>(objectIds.size());
        pt.material_ids = materialIds.data();
        pt.num_material_ids = static_cast<int32_t>(materialIds.size());
        tg.passes = &pt;
    }
    // The whole batch is one call: the library polls the caller's cancel flag between its iterations (a few
    // milliseconds apart) and keeps the windows of samples it completed.
    tg.guide = settings.pathGuiding ? m_guide : nullptr;
    if (cancel) {
        tg.cancel_poll = [](void* flag) -> int {
            return static_cast<const std::atomic<bool>*>(flag)->load(std::memory_order_relaxed) ? 1 : 0;
        };
        tg.cancel_user = const_cast<std::atomic<bool>*>(cancel);
    }
    pe_batch_result br{};
    const int rrc = pe_render_targets(m_dev, m_scene, &cam, &st, &tg, firstSample, count, nullptr, &br);
    if (rrc != PE_OK && rrc != PE_ERR_CANCELLED) {
        m_error = "HipDevice::render: " + lastHipError();
        return res;
    }
    m_lastRays += br.extension_rays + br.shadow_rays;
    const uint64_t traced = br.samples_traced;
    const bool cancelled = rrc == PE_ERR_CANCELLED;
    // Merge what this call added (the whole film unless it accumulates across calls) with Film::addBatch; the
    // normal AOV and the position pass keep their sign (addSampleSigned, film.cpp:73-86).
    std::vector<float> sum(3 * pixels), lum2(pixels);
    std::vector<uint32_t> cnt(pixels);
    for (int k = 0; k < kDeviceFilms; ++k) {
        if (!targets_of[k]) continue;
        DeviceFilm& f = m_films[k];
        uint64_t rejected = 0;
        if (pe_film_download(f.film, sum.data(), cnt.data(), lum2.data(), &rejected) != PE_OK) {
            m_error = "HipDevice::render: " + lastHipError();
            return res;
        }
        Film& dst = *targets_of[k];
        for (int y = 0; y < h; ++y)
            for (int x = 0; x < w; ++x) {
                const size_t p = static_cast<size_t>(y) * static_cast<size_t>(w) + static_cast<size_t>(x);
                dst.addBatch(x, y,
                             Color3f(sum[3 * p] - f.prevSum[3 * p], sum[3 * p + 1] - f.prevSum[3 * p + 1],
                                     sum[3 * p + 2] - f.prevSum[3 * p + 2]),
                             cnt[p] - f.prevCount[p], k == 2 || k == kPassFilm0 + 25);
            }
        if (m_accumulating) {
            f.prevSum = sum;
            f.prevCount = cnt;
        }
    }
    // Cryptomatte: what this call counted, into the CryptoFilms (CryptoFilm::addSamples / addTotalSamples, the
    // merge path the reference keeps for GPU read-back). A slot keeps its id once taken, so its new count is
    // the difference from the last merge.
    if (targets.passes) {
        std::vector<float> ids(PE_CRYPTO_SLOTS * pixels);
        std::vector<uint32_t> counts(PE_CRYPTO_SLOTS * pixels), totals(pixels);
        CryptoFilm* dsts[2] = {&targets.passes->cryptoObject, &targets.passes->cryptoMaterial};
        for (int k = 0; k < 2; ++k) {
            DeviceCrypto& c = m_crypto[k];
            uint64_t dropped = 0;
            if (pe_crypto_film_download(c.film, ids.data(), counts.data(), totals.data(), &dropped) != PE_OK) {
                m_error = "HipDevice::render: " + lastHipError();
                return res;
            }
            for (int y = 0; y < h; ++y)
                for (int x = 0; x < w; ++x) {
                    const size_t p = static_cast<size_t>(y) * static_cast<size_t>(w) + static_cast<size_t>(x);
                    dsts[k]->addTotalSamples(x, y, totals[p] - c.prevTotals[p]);
                    for (int j = 0; j < PE_CRYPTO_SLOTS; ++j) {
                        const size_t q = p * PE_CRYPTO_SLOTS + static_cast<size_t>(j);
                        if (counts[q] > c.prevCounts[q]) dsts[k]->addSamples(x, y, ids[q], counts[q] - c.prevCounts[q]);
                    }
                }
            if (m_accumulating) {
                c.prevCounts = counts;
                c.prevTotals = totals;
            }
        }
    }
    res.completed = !cancelled;
    res.samplesTraced = traced;
    return res;
}

}  // namespace photon

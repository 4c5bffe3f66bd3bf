/* photon_hip.h — C-ABI of libphoton_hip.so, the MI355X (gfx950) wavefront path tracer that stands in for
 * PhotonEngine's integrator hot path behind its RenderDevice interface.
 *
 * Every entry point returns an int status (PE_OK = 0, negative on error); the message of the last error on
 * the calling thread is available from pe_last_error(). No C++ types and no exceptions cross this boundary.
 * All pointers are host pointers unless stated otherwise; arrays are tightly packed little-endian.
 *
 * Reference interface each group replaces (paths relative to the PhotonEngine source tree):
 *   device registry        engine/render_device.h:88-94     enumerateDevices(), createDevice(id, &err)
 *   device info/features   engine/render_device.h:37-56     DeviceInfo, DeviceFeatures
 *   scene upload/update    engine/render_device.h:72-80     RenderDevice::setScene / updateScene / hasScene
 *   render                 engine/render_device.h:84-85     RenderDevice::render(cam, settings, targets, first, count, cancel)
 *   film                   core/image/film.h:26-70          Film sums / counts / lum^2 (addBatch semantics, film.cpp:119-137)
 *   sampler                samplers/sobol_sampler.cpp:92-146 SobolSampler::startPixel/startSample/get2D/get1D
 *   closest hit / any hit  engine/scene.cpp:215-221         Scene::intersect / Scene::intersectAny (BVH::intersect, bvh.cpp:322-446)
 */
#ifndef PHOTON_HIP_H
#define PHOTON_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PE_ABI_VERSION 2

/* ---- status codes ---------------------------------------------------------------------------------- */
#define PE_OK 0
#define PE_ERR_INVALID (-1)      /* bad argument, shape or handle */
#define PE_ERR_NO_DEVICE (-2)    /* no usable HIP device with that ordinal */
#define PE_ERR_HIP (-3)          /* a HIP runtime call failed */
#define PE_ERR_OOM (-4)          /* device memory allocation failed */
#define PE_ERR_UNSUPPORTED (-5)  /* feature not implemented on this device */
#define PE_ERR_CANCELLED (-6)    /* render stopped by the cancel flag (partial result is valid) */
#define PE_ERR_TIMEOUT (-7)      /* pe_job_wait: the job has not ended within the timeout */

typedef struct pe_device pe_device;
typedef struct pe_scene pe_scene;
typedef struct pe_film pe_film;
typedef struct pe_crypto_film pe_crypto_film;
typedef struct pe_guide pe_guide;

/* ---- device registry --------------------------------------------------------------------------------
 * Mirrors DeviceInfo (render_device.h:37-48). kind is always PE_DEVICE_HIP here. */
#define PE_DEVICE_HIP 6
typedef struct pe_device_info {
    char id[32];            /* "hip:0" */
    char name[128];         /* marketing name of the GPU */
    char arch[32];          /* "gfx950" */
    int32_t kind;
    int32_t compute_units;
    uint64_t memory_bytes;
    int32_t usable;         /* 1 when the kernels of this library can run on it */
    char note[128];
} pe_device_info;

/* Mirrors DeviceFeatures (render_device.h:52-56). */
typedef struct pe_device_features {
    int32_t aovs;
    int32_t light_passes;
    int32_t adaptive_sampling;
} pe_device_features;

const char* pe_last_error(void);
int pe_abi_version(void);
int pe_device_count(int* count);                                    /* enumerateDevices() (GPU part) */
int pe_device_info_get(int ordinal, pe_device_info* out);           /* DeviceInfo */
int pe_device_create(int ordinal, pe_device** out);                 /* createDevice("hip:N", &err) */
int pe_device_destroy(pe_device* dev);
int pe_device_features_get(const pe_device* dev, pe_device_features* out);
int pe_device_synchronize(pe_device* dev);

/* ---- scene description ------------------------------------------------------------------------------
 * A flattened engine::Scene (engine/scene.h:30-177). Shapes appear in Scene::shapes() order: a mesh is a
 * TriangleMesh (mesh.cpp), a "standalone" shape is one heap Triangle such as the two black emissive
 * triangles that Scene::addLight inserts for an AreaLight (scene.cpp:72-98). Lights appear in
 * Scene::lights() order; an emissive mesh contributes one PE_LIGHT_TRIANGLE per triangle (scene.cpp:48-62). */

#define PE_MAT_DISNEY 0      /* materials/disney.h DisneyMaterial */
#define PE_MAT_LAMBERT 1     /* materials/lambertian.h Lambertian */
#define PE_MAT_DIELECTRIC 2  /* materials/dielectric.h Dielectric */
#define PE_MAT_MIRROR 3      /* materials/mirror.h Mirror(reflectance = base_color); emits nothing */

typedef struct pe_material_desc {
    int32_t type;
    float base_color[3];          /* Disney baseColor, Lambert albedo, Dielectric tint */
    float emission[3];            /* Material::emitted() */
    float metallic, roughness, specular;
    float clearcoat, clearcoat_roughness, coat_ior;
    float coat_tint[3];
    float ior, specular_transmission, anisotropy;
    float sheen, sheen_roughness;
    float sheen_color[3];
    int32_t has_sheen_color;
    float diffuse_roughness, diffuse_transmission;
    float subsurface;
    float subsurface_radius[3];
    float subsurface_scale, subsurface_anisotropy;
    float edge_tint[3];
    float thin_film_thickness, thin_film_ior;
    /* Dielectric only (dielectric.h:43-78): Abbe number (0 = no dispersion), interior absorption given as
     * setAbsorption(color, distance) (sigma_a = -ln(color) / distance, dielectric.cpp:123-144), thin walls. */
    float abbe;
    float absorption_color[3];
    float absorption_distance;
    int32_t thin_walled;
    /* Disney only: DisneyMaterial::setOpacity (disney.h:147-158, clamped to [0, 1]) when has_opacity != 0
     * (a zeroed descriptor is opaque). Below 1 the surface is a stochastic cutout: a ray passes it with
     * probability 1 - opacity (passesCutout, path_tracer.cpp:159-167, one sampler dimension) and a shadow ray
     * is attenuated by 1 - opacity (shadowTransmittance, :172-202). */
    float opacity;
    int32_t has_opacity;
    /* Dielectric only: Dielectric::setScattering(density, color, anisotropy) (dielectric.h:61,
     * dielectric.cpp:129-144): the interior medium scatters with sigma_s = color * density, adds
     * (1 - color) * density to sigma_a, and has Henyey-Greenstein asymmetry g (clamped to [-0.95, 0.95]); paths
     * inside sample free flights (sampleFreeFlight, path_tracer.cpp:97-129, :457-479). Density 0: no scattering. */
    float scatter_density;
    float scatter_color[3];
    float scatter_anisotropy;
    /* Disney only: DisneyMaterial::setFlakes(density, size, roughness, color) (disney.h:120-125; clamped to
     * [0, 1], size >= 1e-6, [0, 1], [0, 1]). Space is cut into cells of the flake size, one seed per cell (3-D
     * Voronoi); a point whose nearest seed hashes below the density lies on a metallic flake tilted by a GGX-like
     * slope of the flake roughness, which replaces the base layer with a narrow GGX lobe of Schlick reflectance
     * flake_color about the flake normal (resolveFlake, evalFlake, disney.cpp:104-126, :423-460). Density 0:
     * no flakes. */
    float flake_density;
    float flake_size;
    float flake_roughness;
    float flake_color[3];
    float reserved[3];
} pe_material_desc;

#define PE_SHAPE_CAMERA_INVISIBLE 1u

typedef struct pe_mesh_desc {
    const float* positions;     /* 3 * num_vertices */
    const float* normals;       /* 3 * num_vertices, or NULL (geometric normal) */
    const float* uvs;           /* 2 * num_vertices, or NULL */
    const uint32_t* indices;    /* 3 * num_triangles */
    uint32_t num_vertices;
    uint32_t num_triangles;
    int32_t material;           /* index into pe_scene_desc.materials */
    int32_t emitter_base;       /* lights() index of triangle 0 (emissive mesh / area-light quad), or -1 */
    int32_t standalone;         /* 1: heap Triangle shape (reported primIndex -1), 0: TriangleMesh */
    uint32_t flags;             /* PE_SHAPE_* */
} pe_mesh_desc;

#define PE_LIGHT_AREA 1      /* lights/area_light.h AreaLight(position, u, v, radiance) */
#define PE_LIGHT_TRIANGLE 2  /* lights/area_light.h TriangleLight(p0, p1, p2, radiance) */

typedef struct pe_light_desc {
    int32_t type;
    float p0[3], p1[3], p2[3];  /* AREA: position, u, v ; TRIANGLE: vertices */
    float radiance[3];
    int32_t visible_to_camera;
} pe_light_desc;

/* lights/environment_light.h EnvironmentLight(image, rotation, intensity) [+ setSun]. rgb == NULL and
 * width == 0 select the flat grey environment (0.5 * intensity). */
typedef struct pe_env_desc {
    const float* rgb;           /* 3 * width * height, row 0 = zenith */
    int32_t width, height;
    float rotation, intensity;
    int32_t has_sun;
    float sun_direction[3];
    float sun_cos_max;
    float sun_radiance[3];
} pe_env_desc;

typedef struct pe_scene_desc {
    const pe_mesh_desc* meshes;
    int32_t num_meshes;
    const pe_material_desc* materials;
    int32_t num_materials;
    const pe_light_desc* lights;
    int32_t num_lights;
    const pe_env_desc* environment;   /* NULL: no environment light */
    /* Light groups (engine/scene.cpp:144-173), resolved by the caller: light i's group index
     * (Scene::lightGroupIndex(i), lights in Scene::lights() order) and the environment's
     * (Scene::envLightGroupIndex), of num_light_groups (Scene::lightGroupNames().size(), at most
     * PE_MAX_LIGHT_GROUPS). light_groups == NULL or num_light_groups == 0: the scene has no groups. */
    const int32_t* light_groups;
    int32_t env_light_group;
    int32_t num_light_groups;
    /* Per mesh: the scene-graph node that owns it (Scene::tagShape / nodeUidOf, the object-ID pass), or NULL
     * (0 for every mesh: untagged shapes). */
    const uint64_t* node_uids;
    /* Scene::background() (engine/scene.h:24-28): what a camera path that leaves the scene before its first
     * vertex sees (path_tracer.cpp:483-495; it may have passed thin walls and camera-invisible lights):
     * PE_BACKGROUND_ENVIRONMENT (0, the environment light with MIS, as every later miss), PE_BACKGROUND_COLOR
     * (background_color, credited to the Background pass and the environment's light group) or
     * PE_BACKGROUND_TRANSPARENT (nothing; a camera ray that passed thin walls gives the alpha AOV their
     * coverage, 1 - luminance of its throughput). A zeroed descriptor has the environment background. */
    int32_t background_mode;
    float background_color[3];
    /* Light linking (Light::setLightLink, lights/light.h:23-32; Scene::isLightLinked / isShadowLinked,
     * scene.cpp:120-141, :237-256). Per light of `lights`: link_modes[2 i] the receiver mode and [2 i + 1] the
     * shadow mode (PE_LINK_*); link_offsets[4 i .. 4 i + 3] = first index and count into link_uids of its
     * receiver node uids, then of its shadow-blocker node uids. An Include / Exclude receiver mode lights only /
     * all but the listed nodes' surfaces (NEE and emission found by a bounce from them); an Include / Exclude
     * shadow mode lets only / all but the listed nodes block the light's shadow rays. link_modes == NULL: no
     * links. */
    const int32_t* link_modes;
    const uint32_t* link_offsets;
    const uint64_t* link_uids;
} pe_scene_desc;

#define PE_LINK_ALL 0
#define PE_LINK_INCLUDE 1
#define PE_LINK_EXCLUDE 2

#define PE_BACKGROUND_ENVIRONMENT 0
#define PE_BACKGROUND_COLOR 1
#define PE_BACKGROUND_TRANSPARENT 2

#define PE_MAX_LIGHT_GROUPS 16  /* LightPasses::kMaxLightGroups (integrators/path_tracer.h:27) */

#define PE_UPDATE_FULL 0         /* SceneUpdateKind::Full (render_device.h:33) */
#define PE_UPDATE_MATERIALS 1    /* SceneUpdateKind::Materials */
#define PE_UPDATE_LIGHTS 2       /* SceneUpdateKind::Lights */
#define PE_UPDATE_TRANSFORMS 3   /* SceneUpdateKind::Transforms: BVH rebuilt, like Full */
/* RenderDevice::updateScene (render_device.h:77-79; VulkanDevice::updateScene, vulkan_device.cpp:1223-1256).
 * `desc` is the full updated scene. Materials rebuilds the material table and Lights the lights, environment
 * and light-selection structure, both keeping the BVH and triangle data resident; the meshes' material and
 * emitter references must stay in range (PE_ERR_INVALID otherwise, scene unchanged). Full and Transforms
 * rebuild everything and replace the scene only on success. A partial update that fails for another reason
 * (PE_ERR_OOM, PE_ERR_HIP) leaves the scene unusable until the next Full update. */

int pe_scene_upload(pe_device* dev, const pe_scene_desc* desc, pe_scene** out);
int pe_scene_update(pe_scene* scene, int kind, const pe_scene_desc* desc);
int pe_scene_release(pe_scene* scene);
/* Scene statistics: triangles, BVH nodes, device bytes held. */
int pe_scene_stats(const pe_scene* scene, uint64_t* triangles, uint64_t* bvh_nodes, uint64_t* device_bytes);

/* ---- camera / settings ------------------------------------------------------------------------------
 * pe_camera is CameraGpuDesc (camera/camera.h:19-31) of a PerspectiveCamera, ThinLensCamera or
 * OrthographicCamera, plus the shutter-end pose Camera::setShutterEnd records (camera.h:49-54). A camera ray
 * (Camera::generateRay, camera.h:57-91) for screen point (s, t) and lens sample (lu, lv):
 *   orthographic          origin lower_left + s horizontal + t vertical, direction forward
 *   lens_radius = 0       origin position, towards lower_left + s horizontal + t vertical
 *   lens_radius > 0       origin position + radius (x u + y v) with (x, y) uniform on the unit disk (blades < 3)
 *                         or on the regular blades-gon turned by blade_rotation (thin_lens_camera.cpp:36-61)
 * With pe_settings.motion_blur and has_shutter_end, a time t in [0, 1] drawn after the lens sample blends every
 * field linearly towards the end_* pose (the aperture is then always a disk, as camera.h:75-84 samples it). */
typedef struct pe_camera {
    float position[3];
    float lower_left[3];
    float horizontal[3];
    float vertical[3];
    float lens_radius;          /* 0: pinhole */
    float u[3];                 /* lens axes (unit right and up) */
    float v[3];
    float forward[3];           /* view direction of an orthographic camera */
    int32_t blades;             /* aperture polygon (>= 3) or 0 for a disk */
    float blade_rotation;       /* radians */
    int32_t orthographic;
    int32_t has_shutter_end;
    float end_position[3];      /* the pose at shutter close (CameraGpuDesc of the end camera) */
    float end_lower_left[3];
    float end_horizontal[3];
    float end_vertical[3];
    float end_u[3];
    float end_v[3];
    float end_forward[3];
    float end_lens_radius;
} pe_camera;

#define PE_FILTER_BOX 0             /* core/image/pixel_filter.h PixelFilterType */
#define PE_FILTER_TRIANGLE 1
#define PE_FILTER_GAUSSIAN 2
#define PE_FILTER_BLACKMAN_HARRIS 3

/* The fields of RenderSettings (engine/render_settings.h:19-47) that the integrator reads. Film size is
 * taken from the film, not from here (renderer.cpp:396-398). */
typedef struct pe_settings {
    int32_t max_bounces;        /* RenderSettings::maxBounces */
    int32_t rr_depth;           /* PathTracer russianRouletteDepth (3 in Renderer) */
    int32_t shadow_quality;     /* RenderSettings::shadowQuality (NEE samples per vertex) */
    int32_t filter_type;        /* PE_FILTER_* */
    float filter_width;
    float clamp_indirect;       /* 0 = off */
    uint64_t seed;              /* Sobol seed; the reference renderer hard-codes 12345 */
    int32_t samples_per_launch; /* window: samples of every pixel whose results the film adds together, in
                                   sample order, at the window's end (16 B of device memory per sample); 0 = sized
                                   by the library (up to 2^28 samples within an eighth of free memory) */
    int32_t count_visits;       /* 1: also count BVH node / triangle visits (measurement runs; slower) */
    /* Adaptive sampling (RenderSettings::adaptiveSampling / noiseThreshold / minAdaptiveSamples,
       engine/render_settings.h:30-32; renderer.cpp:438-505): a batch skips every pixel whose 3x3 neighbourhood
       is already below the threshold, and within a batch a pixel stops after sample s when s + 1 is a multiple
       of 8, at least min_adaptive_samples and not the batch's end, and its own error is below the threshold. */
    int32_t adaptive;           /* 0 = off */
    float noise_threshold;      /* relative standard error target; clamped to >= 1e-4 (reference default 0.02) */
    int32_t min_adaptive_samples; /* at least 2 (reference default 16) */
    int32_t paths_in_flight;    /* path pool: paths traced together, a finished path's id taking the window's next
                                   sample at once; 0 = 2^26 (never more than the window's samples) */
    int32_t motion_blur;        /* RenderSettings::motionBlur: every camera ray draws a shutter time after its lens
                                   sample (renderer.cpp:151-153); only the camera moves (moving instances are
                                   refused at upload) */
    float ao_strength;          /* RenderSettings::aoStrength: at the first vertex, shadow_quality cosine rays of
                                   0.35 scene units scale the path by lerp(1, unoccluded fraction, strength)
                                   (approxAo, path_tracer.cpp:27-44, :617-621); 0 = off */
} pe_settings;

/* SampleBatchResult (engine/renderer.h:114-117) plus the device's own ray and traversal counters. */
typedef struct pe_batch_result {
    int32_t completed;
    uint64_t samples_traced;
    uint64_t extension_rays;
    uint64_t shadow_rays;
    uint64_t node_visits;        /* closest-hit traversal, when settings.count_visits */
    uint64_t tri_visits;
    uint64_t shadow_node_visits; /* any-hit traversal, when settings.count_visits */
    uint64_t shadow_tri_visits;
    uint32_t max_stack_depth;    /* deepest traversal stack of any ray, when settings.count_visits */
    uint32_t stack_overflows;    /* traversals whose stack overflowed (then the call fails with PE_ERR_INVALID) */
    uint64_t closest_hits;       /* extension rays that hit a triangle, when settings.count_visits */
} pe_batch_result;

/* Per-kernel device time of the wavefront stages, from HIP events recorded on the device's stream around
 * every launch while profiling is enabled. No counterpart in the reference (its renderer is timed as a whole,
 * main.cpp:302-304); it backs the roofline figures. */
typedef struct pe_kernel_stats {
    char name[32];
    uint64_t launches;
    double total_ms;
} pe_kernel_stats;

/* ---- film (device-resident accumulation; Film::addSample semantics per sample) ---------------------- */
int pe_film_create(pe_device* dev, int width, int height, pe_film** out);
int pe_film_destroy(pe_film* film);
int pe_film_clear(pe_film* film);
/* Copies the device film into host arrays (sum_rgb: 3*w*h floats, count: w*h, sum_lum2: w*h; any may be NULL). */
int pe_film_download(pe_film* film, float* sum_rgb, uint32_t* count, float* sum_lum2, uint64_t* rejected);

/* PerspectiveCamera(position, target, up, vFovDegrees, aspect) -> CameraGpuDesc
 * (camera/perspective_camera.cpp:11-34, camera.h:19-31). */
int pe_camera_perspective(const float position[3], const float target[3], const float up[3], float vfov_degrees,
                          float aspect, pe_camera* out);
/* ThinLensCamera(position, target, up, vFovDegrees, aspect, apertureRadius, focusDistance) followed by
 * setApertureShape(blades, rotationDegrees) -> CameraGpuDesc (camera/thin_lens_camera.cpp:14-33, :63-77,
 * thin_lens_camera.h:31-34). */
int pe_camera_thin_lens(const float position[3], const float target[3], const float up[3], float vfov_degrees,
                        float aspect, float aperture_radius, float focus_distance, int blades, float rotation_degrees,
                        pe_camera* out);
/* OrthographicCamera(position, target, up, verticalExtent, aspect) -> CameraGpuDesc
 * (camera/orthographic_camera.cpp:11-40). */
int pe_camera_orthographic(const float position[3], const float target[3], const float up[3], float vertical_extent,
                           float aspect, pe_camera* out);
/* PerspectiveCamera::setShift / ThinLensCamera::setShift (perspective_camera.h:25-27, thin_lens_camera.h:25-27):
 * a lens shift moves the image rectangle by shift_x image heights along u and shift_y along vertical. */
int pe_camera_set_shift(pe_camera* camera, float shift_x, float shift_y);
/* Camera::setShutterEnd(endCamera) (camera.h:49-54): `camera` keeps its pose at shutter open and takes `end`'s
 * CameraGpuDesc as its pose at shutter close. */
int pe_camera_set_shutter_end(pe_camera* camera, const pe_camera* end);

/* RenderDevice::render: adds samples [first_sample, first_sample + sample_count) of every pixel into the
 * film. cancel (may be NULL) is polled between iterations of the wavefront (a few milliseconds apart); a set
 * flag ends the call with completed = 0, PE_ERR_CANCELLED and a valid partial film holding whole windows. */
int pe_render(pe_device* dev, pe_scene* scene, const pe_camera* camera, const pe_settings* settings, pe_film* film,
              int first_sample, int sample_count, const volatile int32_t* cancel, pe_batch_result* result);

/* RenderTargets (engine/render_device.h:58-62) with AovFilms (engine/renderer.h:42-53): the beauty film and,
 * each optional, the AOV films of the same size. Per sample, as tracePixel records them (renderer.cpp:175-205,
 * environment background): the first surface's albedo clamped to [0, 1] (DisneyMaterial baseColor, Lambertian
 * albedo, Dielectric tint) with Film::addSample, its shading normal with Film::addSampleSigned and coverage 1
 * in r, g, b; a camera ray that misses adds the clamped environment radiance along it, a zero normal and
 * coverage 0. Light passes, light groups and the data passes come with `passes` (pe_pass_targets). */

/* LightPass (integrators/path_tracer.h:21-24), same order. */
#define PE_PASS_DIFFUSE_DIRECT 0
#define PE_PASS_DIFFUSE_INDIRECT 1
#define PE_PASS_GLOSSY_DIRECT 2
#define PE_PASS_GLOSSY_INDIRECT 3
#define PE_PASS_TRANSMISSION_DIRECT 4
#define PE_PASS_TRANSMISSION_INDIRECT 5
#define PE_PASS_EMISSION 6
#define PE_PASS_BACKGROUND 7
#define PE_LIGHT_PASSES 8

/* RenderPassFilms (engine/renderer.h:59-100) without cryptomatte: films of the beauty film's size, each
 * optional. Per sample, as PathTracer::Li credits them (path_tracer.cpp:372-395, :482-698, :790-812) and
 * tracePixel / addPasses add them (renderer.cpp:162-171, :207-219):
 *   light[k]         LightPass k; the eight passes of a sample sum to its beauty value
 *   light_groups[g]  what lights of group g added (the scene descriptor's groups; g < num_light_groups)
 *   depth            distance from the camera to the first surface, in r, g, b (0 for a miss)
 *   position         the first surface's point (signed, Film::addSampleSigned)
 *   object_id        idColor(Scene::nodeUidOf) of the first surface (the descriptor's node_uids)
 *   material_id      idColor(material index + 1) of the first surface; the reference hashes the material's
 *                    address, which no other process can reproduce, so the colours differ from its own
 * Light-pass and group films use Film::addSample (a negative or non-finite value is rejected for that film). */
typedef struct pe_pass_targets {
    pe_film* light[PE_LIGHT_PASSES];
    pe_film* light_groups[PE_MAX_LIGHT_GROUPS];
    pe_film* depth;
    pe_film* position;
    pe_film* object_id;
    pe_film* material_id;
    /* Cryptomatte (RenderPassFilms::cryptoObject / cryptoMaterial): per sample, the id of the first surface's
     * object and material, looked up in the caller's tables (buildCryptoLookup, renderer.cpp:65-118: the
     * cryptoHashToFloat of each registered name, cryptomatte.cpp:99-112; 0 = no id), counted per pixel. */
    pe_crypto_film* crypto_object;    /* or NULL */
    pe_crypto_film* crypto_material;  /* or NULL */
    const float* object_ids;          /* num_object_ids entries, one per mesh of the scene description */
    const float* material_ids;        /* num_material_ids entries, one per material */
    int32_t num_object_ids;
    int32_t num_material_ids;
} pe_pass_targets;

/* A device CryptoFilm (core/image/cryptomatte.h:38-88): per pixel the samples counted (PixelBucket::total) and up
 * to PE_CRYPTO_SLOTS distinct ids with their sample counts, in the order they first appeared (the entries
 * vector). A sample whose id finds no free slot is counted in the total only and in `dropped`; the reference's
 * bucket grows without bound, so a pixel that sees more than PE_CRYPTO_SLOTS objects differs from it there. */
#define PE_CRYPTO_SLOTS 8
int pe_crypto_film_create(pe_device* dev, int width, int height, pe_crypto_film** out);
int pe_crypto_film_destroy(pe_crypto_film* film);
int pe_crypto_film_clear(pe_crypto_film* film);
/* ids, counts: PE_CRYPTO_SLOTS * w * h (slot j of pixel p at p * PE_CRYPTO_SLOTS + j; count 0 = unused);
 * totals: w * h; any may be NULL. */
int pe_crypto_film_download(pe_crypto_film* film, float* ids, uint32_t* counts, uint32_t* totals, uint64_t* dropped);

typedef struct pe_targets {
    pe_film* beauty;   /* required */
    pe_film* albedo;   /* AovFilms::albedo, or NULL */
    pe_film* normal;   /* AovFilms::normal, or NULL */
    pe_film* alpha;    /* AovFilms::alpha, or NULL */
    pe_guide* guide;   /* a learned guide of the scene (RenderSettings::pathGuiding), or NULL */
    /* Cancellation for callers whose flag is not an int32 (RenderDevice::render's std::atomic<bool>): called
     * between iterations like the cancel flag is read; non-zero cancels. NULL: the cancel argument only. */
    int (*cancel_poll)(void* user);
    void* cancel_user;
    const pe_pass_targets* passes;  /* render passes, or NULL */
} pe_targets;

/* pe_render into every film of `targets` (RenderDevice::render with RenderTargets). */
int pe_render_targets(pe_device* dev, pe_scene* scene, const pe_camera* camera, const pe_settings* settings,
                      const pe_targets* targets, int first_sample, int sample_count, const volatile int32_t* cancel,
                      pe_batch_result* result);

/* ---- display (film -> imager -> tone map -> 8-bit), on the device -------------------------------------
 * Film::resolve (film.cpp:139-147), applyImager (core/image/imager.cpp, ImagerSettings imager.h:12-36: chromatic
 * aberration, bloom, glare, white balance, contrast / saturation, vignette), toneMap (tone_mapping.cpp:146-160:
 * exposure, operator, clamp, sRGB encode) and, for 8-bit output, quantizeUnorm8 with saveImagePNG's
 * deterministic TPDF dither (transfer.cpp:28-52, image_io.cpp:61-80). */
#define PE_TONEMAP_REINHARD 0           /* ToneMapOperator (tone_mapping.h:14-22), same order */
#define PE_TONEMAP_REINHARD_EXTENDED 1
#define PE_TONEMAP_ACES 2
#define PE_TONEMAP_FILMIC 3
#define PE_TONEMAP_AGX 4
#define PE_TONEMAP_PBR_NEUTRAL 5
#define PE_TONEMAP_LINEAR 6

typedef struct pe_display {
    int32_t tonemap;            /* PE_TONEMAP_* */
    float exposure_ev;          /* also the imager's exposure (bloom threshold, contrast pivot) */
    int32_t dither;             /* 8-bit output: 1 = TPDF dither (saveImagePNG), 0 = plain rounding */
    float bloom_intensity, bloom_radius, bloom_threshold;
    float glare_intensity;
    int32_t glare_blades;
    float glare_angle_deg, glare_length;
    float white_balance_k, tint, saturation, contrast;
    float vignette, chromatic_aberration;
    /* 1 = denoiseImage on the resolved image, before the imager (main.cpp:277-289). Without OIDN, which this
     * image lacks (as does any reference build without PHOTON_ENABLE_OIDN), that is softBlurDenoise
     * (engine/denoiser.cpp:17-43, :140-143): each pixel the 3 x 3 average weighted by 1 / (1 + luminance^2). */
    int32_t denoise;
    float reserved[3];
} pe_display;

/* The reference's defaults: ACES, 0 EV, dither on, ImagerSettings{} (identity), no denoising. */
int pe_display_defaults(pe_display* out);
/* The film's display image: display_rgb (3 * w * h floats in [0, 1], sRGB-encoded) and / or rgb8 (3 * w * h
 * bytes); either may be NULL. Rows top first, as the film. */
int pe_film_display(pe_film* film, const pe_display* settings, float* display_rgb, uint8_t* rgb8);

int pe_profile_enable(pe_device* dev, int enable);
/* Accumulated times since profiling was enabled (or last read); resets them. */
int pe_profile_read(pe_device* dev, pe_kernel_stats* out, int capacity, int* count);

/* ---- kernel-level entry points (parity and measurement) ---------------------------------------------
 * Sobol-Owen sample values: for query i, the sampler state after startPixel(px[i], py[i]),
 * startSample(index[i]) and dim[i] earlier get1D/get2D calls; out[2i], out[2i+1] receive the u32 bit
 * patterns x, y whose top 24 bits scaled by 2^-24 are the get2D() floats (get1D() == x). */
int pe_sobol_batch(pe_device* dev, uint64_t seed, size_t n, const int32_t* px, const int32_t* py,
                   const uint32_t* index, const uint32_t* dim, uint32_t* out);

/* Closest hit for n rays (rays: 8 floats each = origin xyz, direction xyz, tmin, tmax). hits: 4 x 32-bit
 * each = t (float), shape index (int, -1 = miss), prim index (int, -1 for standalone shapes), and the
 * global triangle index (int). bary (optional, may be NULL): 2 floats u, v per ray. */
int pe_intersect_batch(pe_device* dev, pe_scene* scene, size_t n, const float* rays, int32_t* hits, float* bary);
/* pe_intersect_batch that allocates nothing: rays_dev, hits_dev and bary_dev (optional) are device buffers the
 * caller owns (same layouts as pe_intersect_batch), and the traversal's stack spill and counters come from a
 * fixed device scratch of 29,360,384 bytes per GPU, reserved for the process's lifetime; the grid is capped at
 * 256 blocks of 256 lanes so that the spill fits it (a batch that would need more is refused). No counterpart
 * in the reference: it is for callers that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc). Calls that overlap on one GPU share its scratch and race.
 * Synchronous on the device's stream; at most 2^31 - 1 rays. */
int pe_intersect_batch_noalloc(pe_device* dev, pe_scene* scene, size_t n, const float* rays_dev, int32_t* hits_dev,
                               float* bary_dev);
/* Any hit (shadow rays): occluded[i] = 1 when something lies within [tmin, tmax]. */
int pe_occluded_batch(pe_device* dev, pe_scene* scene, size_t n, const float* rays, uint8_t* occluded);
/* Traversal diagnostics of the device's last pe_intersect_batch / pe_occluded_batch: the deepest stack any of
 * its rays needed (the first 16 entries live in LDS, deeper ones in a global spill column) and the number of
 * traversals whose stack overflowed (a batch with any fails with PE_ERR_INVALID). No counterpart in the
 * reference (its traversal stack is a growing std::vector, bvh.cpp:287-296). */
int pe_trace_diag(pe_device* dev, uint32_t* max_stack_depth, uint64_t* overflows);

/* ---- learned path guiding ------------------------------------------------------------------------------
 * The reference guides with an SD-tree (GuidingField, integrators/path_guiding.cpp:11-388) trained in the first
 * samples and one-sample-MIS-combined with BSDF sampling (path_tracer.cpp:700-735, NEE pdf :668-671). This
 * library's guide is a small MLP instead: input the shading point and the outgoing direction, each of the six
 * coordinates encoded with 12 sine/cosine frequencies (144 features, plus the 6 raw values, padded to 160);
 * four hidden layers of 64 (ReLU); output an 8-lobe von Mises-Fisher mixture (per lobe a mean direction, a
 * concentration exp(o) and a softmax weight). Weights and activations are bf16, accumulation fp32, on MFMA.
 * At a non-delta vertex a render with a guide samples the mixture with probability `guide_fraction` and the
 * BSDF otherwise, and weighs by the mixture density (1 - a) p_bsdf + a q — also in the NEE and emission MIS —
 * so the estimate stays unbiased for any weights. The guide trains online (Adam) on the radiance the paths of
 * each launch bring back along their sampled directions. A render without a guide is the plain path tracer. */
typedef struct pe_guide_config {
    float learning_rate;        /* Adam step size; 0 = 1e-3 */
    float guide_fraction;       /* probability of sampling the mixture at a guided vertex; 0 = 0.5 */
    int32_t batch;              /* training records per Adam step; 0 = 2^18 (rounded up to a multiple of 256) */
    int32_t steps_per_launch;   /* Adam steps after each render launch while training; 0 = 4 */
    int32_t train_launches;     /* launches that train (then the weights stay fixed); 0 = unlimited */
    uint32_t seed;              /* weight initialisation */
    float bounds_min[3];        /* position normalisation box; min == max: the scene's bounds */
    float bounds_max[3];
    float weight_clamp;         /* record weights above this multiple of the batch mean are clamped; 0 = off */
    int32_t guided_bounces;     /* bounces (from the camera) whose vertices are guided; 0 = 3 */
    int32_t reserved[4];
} pe_guide_config;

/* Network geometry: layer l maps in[l] -> out[l]; parameters are stored per layer as W[out][in] then b[out]. */
#define PE_GUIDE_INPUTS 160
#define PE_GUIDE_HIDDEN 64
#define PE_GUIDE_OUTPUTS 64     /* 40 used: lobe j = outputs 5j..5j+4 (mean xyz, log concentration, weight logit) */
#define PE_GUIDE_LAYERS 5
#define PE_GUIDE_LOBES 8
#define PE_GUIDE_PARAMS 26944   /* 64*160+64 + 4*(64*64+64) */

int pe_guide_create(pe_device* dev, pe_scene* scene, const pe_guide_config* config, pe_guide** out);
int pe_guide_destroy(pe_guide* guide);
/* fp32 master parameters (PE_GUIDE_PARAMS floats, layer order W then b); set also resets the Adam state. */
int pe_guide_get_params(pe_guide* guide, float* params);
int pe_guide_set_params(pe_guide* guide, const float* params);
/* Mixture of n queries (positions and outgoing directions, 3 floats each): per query 8 lobes x (mean xyz,
 * concentration, weight) = 40 floats. Runs the same fused MFMA kernel the renderer uses. */
int pe_guide_infer(pe_guide* guide, size_t n, const float* pos, const float* dir, float* lobes);
/* One Adam step on n records (pos xyz, dir xyz, sampled direction xyz, weight: 10 floats each; n <= batch):
 * the loss is -sum w log q(direction) / sum w. loss (may be NULL) receives the loss before the step; grad (may
 * be NULL) the fp32 gradient (PE_GUIDE_PARAMS floats) the step used. */
int pe_guide_train_step(pe_guide* guide, size_t n, const float* records, float* loss, float* grad);
/* Statistics: Adam steps taken, loss of the last step, records collected and mixtures inferred by renders. */
int pe_guide_stats(pe_guide* guide, uint64_t* steps, float* last_loss, uint64_t* records, uint64_t* inferred);
/* pe_render with a guide (NULL = pe_render). Trains after each launch while the guide's training lasts. */
int pe_render_guided(pe_device* dev, pe_scene* scene, pe_guide* guide, const pe_camera* camera,
                     const pe_settings* settings, pe_film* film, int first_sample, int sample_count,
                     const volatile int32_t* cancel, pe_batch_result* result);

/* ---- serving: continuous batching of render jobs ----------------------------------------------------
 * No single counterpart in the reference: RenderSession (engine/render_session.cpp:69-101) drives one camera
 * and one film through RenderDevice::render in batches; a server drives many render jobs of different scenes,
 * film sizes and sample counts through the same path at once. It owns `slots` render lanes, each a HIP stream
 * with its own path state. A submitted job waits until its film fits the film budget and the admission cap,
 * then holds a device film; every free lane takes the next chunk (a sample range of every pixel) from the
 * admitted jobs, jobs short of their preview first, then in submission order, so chunks of different jobs
 * run on the device together and a job that arrives while others render starts as soon as a lane frees up.
 * A job's chunks run one after another in sample order and their boundaries depend on the job and the
 * server configuration only, so a job's film is the same bit for bit whether it ran alone or among others.
 * Device memory: the lanes hold their path pools, result buffers and traversal-stack spill for the server's
 * life; the memory budget covers the rest, the films of admitted jobs and the scenes kept resident. Scenes
 * are uploaded by the caller (pe_scene_upload) and stay resident across jobs; when the head of the queue does
 * not fit, the server parks idle scenes (no admitted job renders them; least recently used first) in host
 * memory, and restores a parked scene when a job that needs it is admitted. A scene cannot be released while
 * a job that references it is not released (PE_ERR_INVALID); pe_scene_update waits for the renders of the
 * scene in progress and renders wait for it. */
typedef struct pe_server pe_server;

typedef struct pe_server_config {
    int32_t slots;              /* render lanes (streams) working at once; 0 = 4 */
    int32_t chunk_spp;          /* samples of every pixel per chunk after the preview; 0 = 16 */
    uint64_t max_chunk_paths;   /* paths one chunk keeps in flight (splits chunks of large films); 0 = 2^25 */
    uint64_t memory_budget;     /* device bytes for the films of admitted jobs and the resident scenes; 0 = three
                                   quarters of the device's memory less what the lanes hold */
    int32_t max_admitted;       /* jobs holding a film at once; 0 = 64 */
    int32_t lane_paths;         /* path ids each lane keeps in flight (pe_settings.paths_in_flight); 0 = 2^24 */
    int32_t reserved[4];
} pe_server_config;

/* Server memory and residency counters. */
typedef struct pe_server_stats {
    uint64_t memory_budget;     /* as configured or defaulted */
    uint64_t memory_used;       /* films of admitted jobs and resident scenes */
    uint64_t lane_bytes;        /* path pools, result buffers and stack spill of the lanes */
    uint64_t scenes_parked;     /* scenes moved to host memory so far */
    uint64_t scenes_restored;   /* ... and brought back */
    int32_t admitted, queued;   /* jobs holding a film / waiting for admission */
    int32_t scenes_tracked, scenes_resident;
} pe_server_stats;

typedef struct pe_job_desc {
    pe_scene* scene;            /* an uploaded scene; it cannot be released before the job is */
    pe_camera camera;
    pe_settings settings;       /* samples_per_launch is ignored: a chunk is one launch */
    int32_t width, height;      /* film size */
    int32_t spp;                /* samples of every pixel to render */
    int32_t preview_spp;        /* the first milestone, 0 = 16 (at most spp) */
    int32_t reserved[4];
} pe_job_desc;

#define PE_JOB_QUEUED 0         /* submitted; waiting for film memory or an admission slot */
#define PE_JOB_ADMITTED 1       /* holds a film; its chunks interleave with other admitted jobs */
#define PE_JOB_DONE 2
#define PE_JOB_CANCELLED 3
#define PE_JOB_FAILED 4

typedef struct pe_job_status {
    int32_t state;              /* PE_JOB_* */
    int32_t samples_done;       /* samples of every pixel accumulated in the job's film */
    int32_t error;              /* PE_* status of a failed job */
    int32_t chunks;             /* chunks rendered */
    double t_submit, t_admit, t_preview, t_done;  /* seconds on the server clock; -1 = not yet */
    uint64_t extension_rays, shadow_rays;
    char message[128];          /* error message of a failed job */
} pe_job_status;

int pe_server_create(pe_device* dev, const pe_server_config* config, pe_server** out);
/* Cancels the jobs left, waits for the lanes and frees every film. */
int pe_server_destroy(pe_server* server);
/* Seconds since the server was created (the clock of pe_job_status). */
double pe_server_clock(const pe_server* server);
int pe_server_stats_get(pe_server* server, pe_server_stats* out);
int pe_job_submit(pe_server* server, const pe_job_desc* job, uint64_t* id);
int pe_job_poll(pe_server* server, uint64_t id, pe_job_status* status);
/* Waits until the job is done, cancelled or failed (timeout_ms < 0: no limit); PE_ERR_TIMEOUT otherwise. */
int pe_job_wait(pe_server* server, uint64_t id, int32_t timeout_ms, pe_job_status* status);
/* Stops the job after its current chunk; the film keeps the chunks already added. */
int pe_job_cancel(pe_server* server, uint64_t id);
/* Copies the job's film (whole chunks only) into host arrays, as pe_film_download; any pointer may be NULL. */
int pe_job_download(pe_server* server, uint64_t id, float* sum_rgb, uint32_t* count, float* sum_lum2);
/* The job's display image (whole chunks only), as pe_film_display of its film: the preview a client shows. */
int pe_job_display(pe_server* server, uint64_t id, const pe_display* settings, float* display_rgb, uint8_t* rgb8);
/* Forgets an ended job and frees its film (a running job is cancelled first). */
int pe_job_release(pe_server* server, uint64_t id);

#ifdef __cplusplus
}
#endif

#endif /* PHOTON_HIP_H */

// pe_device_scene.h — shading-side scene data in HBM (materials, lights, environment, energy tables) and the
// by-value view of a whole scene that every wavefront kernel receives.
#pragma once

#include "pe_gpu_types.h"
#include "pe_math.h"

namespace pe {

// Microfacet energy tables (32 x 32 GGX E/Ea/Eb, their hemispherical averages, Charlie sheen albedo), built on
// the host exactly as the reference builds them at first use (materials/microfacet.cpp:46-133).
constexpr int kTableN = 32;
struct EnergyTables {
    float e[kTableN * kTableN];
    float ea[kTableN * kTableN];
    float eb[kTableN * kTableN];
    float avg0[kTableN], avg1[kTableN], avg2[kTableN];
    float sheen[kTableN * kTableN];
    // Wavelength sampling (core/color/spectral.cpp:33-62): 1 nm bins over [380, 730] nm, CDF and pdf proportional
    // to the normalised r + g + b response, and the per-channel normalisation.
    float spec_cdf[351];
    float spec_pdf[350];
    float spec_norm[3];
};

// A material with its parameters clamped as the reference's setters clamp them and every quantity that depends
// on the material alone precomputed on the host (same operations the reference performs per shading call).
enum : int32_t { kMatDisney = 0, kMatLambert = 1, kMatDielectric = 2, kMatMirror = 3 };
struct alignas(16) GpuMaterial {
    int32_t type;
    uint32_t flags;  // kMatFlag*
    float metallic, roughness;
    V3 base_color;
    float specular;
    V3 emission;
    float clearcoat;
    V3 coat_tint;
    float clearcoat_roughness;
    V3 sheen_color;       // resolvedSheenColor (disney.cpp:160-166), already scaled by sheen * (1 - metallic)
    float sheen_max;      // max component of sheen_color
    V3 edge_tint;
    float sheen_roughness;
    float ior, coat_ior, specular_transmission, anisotropy;
    float diffuse_roughness, diffuse_transmission, subsurface;
    int32_t sort_class;   // shade's sort key: the shading code path (kSortClass*)
    // derived
    float ax, ay, alpha;             // anisoAlpha and sqrt(ax * ay) (disney.cpp:55-60, :133-135)
    float f0d, wd;                   // dielectricF0, dielectricWeight (:34-42)
    float inv_one_minus_avg;         // DielectricLayer::invOneMinusAvg (:139)
    float f0c, alpha_c;              // coatF0 (:44-46), clearcoat alpha
    float p_coat, p_spec, p_spec_trans, p_diff_refl, p_sss, p_trans;  // lobeWeights (:313-338)
    float st_w, opaque_dielectric, trans_w, diff_w;                      // evalLobes weights (:230-237, :263)
    V3 coat_tint_factor;             // coatTintFactor (:155-157)
    float pad1;
    // Dielectric (dielectric.h): ior above, tint = base_color, roughness above
    V3 sigma_a;                      // interior Beer-Lambert absorption (rebuildMedium, dielectric.cpp:136-144)
    float abbe;                      // > 0: Cauchy dispersion (setAbbe, dielectric.h:74)
    float glass_alpha;               // ggxClampAlpha(roughness^2) (dielectric.cpp:21)
    float opacity;                   // Disney setOpacity (kMatFlagCutout when below 1)
    float pad3, pad4;
    // Disney subsurface random walk (subsurfaceCoeffs, disney.cpp:738-757)
    V3 sss_sigma_t;
    float sss_g;
    V3 sss_albedo;
    float pad5;
    // Thin-film interference (fresnel.h:74-107): film thickness (nm) and index; the base under the film:
    // Disney dielectric base index (1 + sqrt f0) / (1 - sqrt f0) and metal n + ik per channel (gulbrandsenIor);
    // a thin-walled Dielectric's film sits between air on both sides.
    float film_thickness, film_ior, film_nd, pad6;
    V3 film_metal_n;
    float pad7;
    V3 film_metal_k;
    float pad8;
    // Dielectric interior scattering (setScattering, dielectric.cpp:129-144): sigma_s and the Henyey-Greenstein g
    V3 sigma_s;
    float medium_g;
    // Disney metallic flakes (setFlakes, disney.h:120-125; kMatFlagFlakes when the density is above 0)
    V3 flake_color;
    float flake_density;
    float flake_size, flake_roughness, pad9, pad10;
};

// Shading code paths the shade kernel groups a round's hits by (when a scene has more than one).
enum : int32_t { kSortClassDisney = 0, kSortClassDisneyTransport = 1, kSortClassDielectric = 2, kSortClassSimple = 3,
                 kSortClasses = 4 };

enum : uint32_t {
    kMatFlagEmissive = 1u,
    kMatFlagDelta = 2u,
    kMatFlagSheen = 4u,
    kMatFlagCoat = 8u,
    kMatFlagVolume = 16u,            // isVolumeBoundary: refraction enters / leaves the medium stack
    kMatFlagThinWall = 32u,          // thin-walled dielectric: pass-through and shadow transmittance
    kMatFlagDispersive = 64u,        // isDispersive
    kMatFlagMedium = 128u,           // interior medium absorbs (sigma_a != 0)
    kMatFlagFilm = 256u,             // thin-film interference on the specular Fresnel / the thin wall
    kMatFlagCutout = 512u,           // Disney opacity below 1: stochastic cutout, shadow transmittance 1 - opacity
    kMatFlagScatter = 1024u,         // interior medium scatters (sigma_s != 0): free flights inside
    kMatFlagFlakes = 2048u,          // Disney metallic flakes (resolved per shading point)
};

// AreaLight (rectangle/parallelogram) or TriangleLight with the constructor's derived values.
struct alignas(16) GpuLight {
    int32_t type;         // PE_LIGHT_AREA / PE_LIGHT_TRIANGLE
    int32_t rectangle;    // AreaLight::m_rectangle
    float area;
    float lum;            // max(0, luminance(radiance))
    V3 p0, p1, p2;        // AREA: position, u, v
    V3 radiance;
    V3 normal;
    V3 center;            // importance reference point
    float len_u, len_v;   // |u|, |v| (AREA)
    int32_t group;        // Scene::lightGroupIndex (scene.cpp:258-261), -1 without light groups
};

// Light-BVH node (LightSampler::Node + LightBounds, light_sampler.h:64-69, light_bounds.h:15-25), 64 B. Nodes
// are depth-first: an interior node's first child is the next node, its second child is `child1`.
struct alignas(16) GpuLightNode {
    float4 lo_phi;    // bounds lo xyz, phi (power measure)
    float4 hi_cos_o;  // bounds hi xyz, cosThetaO
    float4 w_cos_e;   // cone axis xyz, cosThetaE
    int4 ref;         // child1 (interior) or -1, light candidate (leaf) or -1, twoSided, 0
};
static_assert(sizeof(GpuLightNode) == 64, "GpuLightNode must be 64 bytes");

// Per light candidate in BVH mode: the root-to-leaf bit trail (child1 = 1, lowest bit first), its length,
// and the candidate's top-level slot (unbounded lights and the environment), -1 for lights in the tree.
struct LightPath {
    uint32_t trail_lo, trail_hi;
    int32_t depth, top_slot;
};

constexpr int kMaxTopCandidates = 16;  // LightSampler::kMaxInfinite

// EnvironmentLight: image, MIS-compensated sampling distribution and the optional sun cone.
struct EnvView {
    const float* rgb;        // 3 * w * h
    const float* func;       // w * h (sampling distribution of the mis-compensated image)
    const float* marginal;   // h + 1
    const float* cond;       // h * (w + 1)
    // Guide tables of the two CDFs: entry k of a table with K buckets is the first CDF index whose value
    // exceeds k / K (K a power of two), so the search for u only spans entries floor(u K) .. floor(u K) + 1.
    const uint32_t* marginal_guide;  // marginal_k + 1
    const uint32_t* cond_guide;      // h * (cond_k + 1)
    int32_t marginal_k, cond_k;
    int32_t w, h;
    float func_int, rotation, intensity, mean_luminance;
    int32_t valid;           // distribution valid
    V3 sun_dir, sun_radiance;
    float sun_cos_max, sun_pdf_weight;
};

enum : int32_t { kPickUniform = 0, kPickExhaustive = 1, kPickBvh = 2 };

struct SceneView {
    const BvhNode4* nodes;  // four-wide, root at 0, the top kBvhTileNodes first (breadth-first)
    uint32_t num_nodes;
    const TriAccel* tris;
    const TriShade* shade;
    const float4* pos;
    const float4* nrm;
    const float2* uv;
    const GpuMaterial* materials;
    const GpuLight* lights;
    const EnergyTables* tables;
    EnvView env;
    int32_t num_lights;      // explicit lights (Scene::lights())
    int32_t has_env;
    int32_t light_count;     // candidates incl. the environment
    int32_t pick_mode;       // kPick*
    int32_t exit_material;   // white Lambert shading a subsurface walk's exit (one past the scene's materials)
    int32_t has_transmitters;  // a thin-walled surface passes shadow rays on with its transmittance
    int32_t transport;         // a Dielectric or a transmitting / subsurface material is present
    int32_t sort_classes;      // distinct GpuMaterial::sort_class values of the scene's materials
    int32_t num_groups;        // light groups (Scene::lightGroupNames().size()), 0 = none
    int32_t env_group;         // Scene::envLightGroupIndex
    int32_t bg_mode;           // PE_BACKGROUND_*: what a camera path leaving before its first vertex sees
    float3 bg_color;           // the Color background
    const uint64_t* node_uids; // per shape: Scene::nodeUidOf (object-ID pass)
    // Light links (scene.cpp:120-141): per light six words (receiver mode, shadow mode, receiver uids first /
    // count, shadow uids first / count) into link_uids; has_links / has_shadow_links: any mode other than All.
    const uint32_t* link_info;
    const uint64_t* link_uids;
    int32_t has_links, has_shadow_links;
    // Light BVH (kPickBvh): tree over the bounded lights, unbounded candidates on a top level with the tree.
    const GpuLightNode* light_nodes;
    const LightPath* light_paths;  // light_count entries
    int32_t num_light_nodes;
    int32_t num_top;               // unbounded candidates (top-level slots before the tree's)
    int32_t top[kMaxTopCandidates];
};

}  // namespace pe

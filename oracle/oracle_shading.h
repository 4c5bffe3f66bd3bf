// oracle_shading.h — restated BSDFs, sampling warps and the microfacet energy tables (test infrastructure only).
#pragma once

#include "oracle.h"

#include <algorithm>
#include <cstring>

namespace oracle {

constexpr float PI = 3.14159265358979323846f;
constexpr float TWO_PI = 2.0f * PI;
constexpr float INV_PI = 1.0f / PI;
constexpr float PI_OVER_2 = PI / 2.0f;
constexpr float PI_OVER_4 = PI / 4.0f;
constexpr float INFINITY_F = 3.40282346638528859812e+38f;

inline float clampf(float v, float lo, float hi) { return std::clamp(v, lo, hi); }
inline float lerpf(float a, float b, float t) { return a + t * (b - a); }              // utils.h:32-34
inline V3 lerpc(V3 a, V3 b, float t) { return a * (1.0f - t) + b * t; }                 // spectrum.h:113-115
inline float powerHeuristic(int nf, float fPdf, int ng, float gPdf) {                   // utils.h:108-113
    const float f = static_cast<float>(nf) * fPdf, g = static_cast<float>(ng) * gPdf;
    const float f2 = f * f;
    return f2 / (f2 + g * g);
}
inline bool finiteFloat(float x) {  // float_bits.h:20-22
    uint32_t u;
    std::memcpy(&u, &x, 4);
    return (u & 0x7f800000u) != 0x7f800000u;
}
inline bool isFinite(V3 c) { return finiteFloat(c.x) && finiteFloat(c.y) && finiteFloat(c.z); }

struct Frame {  // frame.h
    V3 s, t, n;
    Frame() : s(1, 0, 0), t(0, 1, 0), n(0, 0, 1) {}
    Frame(V3 s_, V3 t_, V3 n_) : s(s_), t(t_), n(n_) {}
    explicit Frame(V3 normal);
    V3 toLocal(V3 v) const { return V3(dot(v, s), dot(v, t), dot(v, n)); }
    V3 toWorld(V3 v) const { return s * v.x + t * v.y + n * v.z; }
};

// sampling.h
V3 uniformSampleSphere(float u0, float u1);
void uniformSampleDisk(float u0, float u1, float& x, float& y);
V3 cosineSampleHemisphere(float u0, float u1);
inline float cosineHemispherePdf(float c) { return c * INV_PI; }
void uniformSampleTriangle(float u0, float u1, float& b1, float& b2);

// utils.h
inline V3 reflectVec(V3 wo, V3 n) { return 2.0f * dot(wo, n) * n - wo; }
bool refractVec(V3 wi, V3 n, float eta, V3& wt);
float fresnelDielectric(float cosThetaI, float etaI, float etaT);

// microfacet.h / microfacet.cpp
float ggxD(V3 h, float ax, float ay);
float ggxG1(V3 w, float ax, float ay);
float ggxG(V3 wo, V3 wi, float ax, float ay);
float ggxVisiblePdf(V3 wo, V3 h, float ax, float ay);
V3 ggxSampleVisibleNormal(V3 wo, float ax, float ay, float u1, float u2);
float ggxDirectionalAlbedo(float cosTheta, float alpha);
float ggxSchlickAlbedo(float cosTheta, float alpha, float f0);
float ggxSchlickAverageAlbedo(float alpha, float f0);
float charlieSheenAlbedo(float cosTheta, float roughness);

// The BSDF of one material at one shading point (material.h:63-170 contract; Disney, Lambert).
struct ShadingPoint {
    V3 normal;   // shading normal facing the incoming ray
    V3 tangent;
    bool frontFace = true;
    V3 ng;                   // geometric normal, not flipped (SurfaceInteraction::ng)
    float wavelength = 0.0f; // > 0: the path's single wavelength (nm) after a dispersive hit
    V3 point;                // SurfaceInteraction::point (where flakes are resolved)
};

struct BsdfSample {  // material.h:37-49
    V3 wi, f;
    float pdf = 0.0f;
    bool delta = false, passThrough = false, subsurface = false;
    V3 weight;
};

// DisneyMaterial::resolve's flake state at a shading point (disney.cpp:600-601).
void flakeState(const pe_material_desc& m, const ShadingPoint& sp, bool& inFlake, V3& flakeNormal);
bool materialIsDelta(const pe_material_desc& m);
// Material::isVolumeBoundary / transmitsShadows / isDispersive (material.h:109-122; dielectric.h:48-78;
// disney.h:76).
bool materialIsVolumeBoundary(const pe_material_desc& m);
float materialOpacity(const pe_material_desc& m);
bool materialTransmitsShadows(const pe_material_desc& m);
bool materialIsDispersive(const pe_material_desc& m);
// Dielectric interior medium (Dielectric::rebuildMedium, dielectric.cpp:136-144): absorption only (the scene
// description carries no scattering). false = none (interiorMedium() == nullptr).
bool materialInteriorMedium(const pe_material_desc& m, V3& sigmaA, V3& sigmaS, float& g);
bool materialInteriorAbsorption(const pe_material_desc& m, V3& sigmaA);
// Dielectric::shadowTransmittance (dielectric.cpp:163-167).
V3 materialShadowTransmittance(const pe_material_desc& m, const ShadingPoint& sp, V3 dir);
// DisneyMaterial::subsurfaceCoeffs (disney.cpp:738-757).
bool subsurfaceCoeffs(const pe_material_desc& m, const ShadingPoint& sp, V3& sigmaT, V3& albedo, float& g);
// spectral.cpp: sampleWavelength (:75-87) and cauchyIor (:89-97).
float sampleWavelength(float u, V3& weight);
float cauchyIor(float nd, float abbe, float lambdaNm);
V3 materialEmitted(const pe_material_desc& m);
V3 evalPdf(const pe_material_desc& m, V3 wo, V3 wi, const ShadingPoint& sp, float& pdf);
// Material::evalSplit (material.h:51-57, :154-161): evalPdf's value split into diffuse, glossy and transmission.
struct LobeSplit {
    V3 diffuse{0.0f}, glossy{0.0f}, transmission{0.0f};
};
LobeSplit evalSplit(const pe_material_desc& m, V3 wo, V3 wi, const ShadingPoint& sp);
bool sampleFull(const pe_material_desc& m, V3 wo, const ShadingPoint& sp, float uc, float u0, float u1, BsdfSample& out);


// Airy thin-film reflectance n1 | film (n2, thickness) | base n3 = (n3re, n3im) at one wavelength (fresnel.h:74-92).
float thinFilmReflectanceC(float cos1, float n1, float n2, float n3re, float n3im, float thicknessNm, float lambdaNm);
}  // namespace oracle

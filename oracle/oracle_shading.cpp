// oracle_shading.cpp — restated Disney/OpenPBR principled BSDF, Lambert, GGX helpers and energy tables
// (test infrastructure only). Each function follows the reference file:line cited beside it.
#include "oracle_shading.h"

#include <algorithm>
#include <array>
#include <complex>
#include <cstring>
#include <stdexcept>

namespace oracle {

// frame.h:27-38 (Duff et al. 2017)
Frame::Frame(V3 normal) : n(normal) {
    const float sign = std::copysign(1.0f, n.z);
    const float a = -1.0f / (sign + n.z);
    const float b = n.x * n.y * a;
    s = V3(1.0f + sign * n.x * n.x * a, sign * b, -sign * n.x);
    t = V3(b, sign + n.y * n.y * a, -n.y);
}

// sampling.h:24-29
V3 uniformSampleSphere(float u0, float u1) {
    const float z = 1.0f - 2.0f * u0;
    const float r = std::sqrt(std::max(0.0f, 1.0f - z * z));
    const float phi = TWO_PI * u1;
    return V3(r * std::cos(phi), r * std::sin(phi), z);
}

// sampling.h:53-72 (Shirley–Chiu concentric map)
void uniformSampleDisk(float u0, float u1, float& x, float& y) {
    const float sx = 2.0f * u0 - 1.0f, sy = 2.0f * u1 - 1.0f;
    if (sx == 0.0f && sy == 0.0f) {
        x = y = 0.0f;
        return;
    }
    float r, theta;
    if (std::abs(sx) > std::abs(sy)) {
        r = sx;
        theta = PI_OVER_4 * (sy / sx);
    } else {
        r = sy;
        theta = PI_OVER_2 - PI_OVER_4 * (sx / sy);
    }
    x = r * std::cos(theta);
    y = r * std::sin(theta);
}

// sampling.h:79-83
V3 cosineSampleHemisphere(float u0, float u1) {
    float x, y;
    uniformSampleDisk(u0, u1, x, y);
    return V3(x, y, std::sqrt(std::max(0.0f, 1.0f - x * x - y * y)));
}

// sampling.h:95-98
void uniformSampleTriangle(float u0, float u1, float& b1, float& b2) {
    const float su0 = std::sqrt(u0);
    b1 = 1.0f - su0;
    b2 = u1 * su0;
}

// utils.h:163-175
bool refractVec(V3 wi, V3 n, float eta, V3& wt) {
    const float cosThetaI = dot(n, wi);
    const float sin2ThetaI = std::max(0.0f, 1.0f - cosThetaI * cosThetaI);
    const float sin2ThetaT = eta * eta * sin2ThetaI;
    if (sin2ThetaT >= 1.0f) return false;
    const float cosThetaT = std::sqrt(1.0f - sin2ThetaT);
    wt = eta * (-wi) + (eta * cosThetaI - cosThetaT) * n;
    return true;
}

// utils.h:206-232
float fresnelDielectric(float cosThetaI, float etaI, float etaT) {
    cosThetaI = clampf(cosThetaI, -1.0f, 1.0f);
    if (!(cosThetaI > 0.0f)) {
        std::swap(etaI, etaT);
        cosThetaI = std::abs(cosThetaI);
    }
    const float sinThetaI = std::sqrt(std::max(0.0f, 1.0f - cosThetaI * cosThetaI));
    const float sinThetaT = (etaI / etaT) * sinThetaI;
    if (sinThetaT >= 1.0f) return 1.0f;
    const float cosThetaT = std::sqrt(std::max(0.0f, 1.0f - sinThetaT * sinThetaT));
    const float rs = (etaT * cosThetaI - etaI * cosThetaT) / (etaT * cosThetaI + etaI * cosThetaT);
    const float rp = (etaI * cosThetaI - etaT * cosThetaT) / (etaI * cosThetaI + etaT * cosThetaT);
    return 0.5f * (rs * rs + rp * rp);
}

// ------------------------------------------------------------------------------------ microfacet.h
namespace {
constexpr float GGX_MIN_ALPHA = 1e-4f;  // microfacet.h:15
float ggxClampAlpha(float a) { return std::max(a, GGX_MIN_ALPHA); }
float pow5(float x) {  // microfacet.h:129-132
    const float x2 = x * x;
    return x2 * x2 * x;
}
float ggxLambda(V3 w, float ax, float ay) {  // microfacet.h:43-50
    const float z2 = w.z * w.z;
    if (z2 <= 1e-12f) return 1e12f;
    ax = ggxClampAlpha(ax);
    ay = ggxClampAlpha(ay);
    const float t2 = (ax * ax * w.x * w.x + ay * ay * w.y * w.y) / z2;
    return 0.5f * (-1.0f + std::sqrt(1.0f + t2));
}
float charlieD(float cosH, float roughness) {  // microfacet.h:111-116
    const float a = std::max(roughness * roughness, 1e-4f);
    const float inv = 1.0f / a;
    const float sin2 = std::max(0.0f, 1.0f - cosH * cosH);
    return (2.0f + inv) * std::pow(sin2, 0.5f * inv) * (0.5f * INV_PI);
}
float charlieV(float cosI, float cosO) { return 1.0f / std::max(4.0f * (cosI + cosO - cosI * cosO), 1e-6f); }
}  // namespace

float ggxD(V3 h, float ax, float ay) {  // microfacet.h:27-33
    if (h.z <= 0.0f) return 0.0f;
    ax = ggxClampAlpha(ax);
    ay = ggxClampAlpha(ay);
    const float e = (h.x * h.x) / (ax * ax) + (h.y * h.y) / (ay * ay) + h.z * h.z;
    return 1.0f / (PI * ax * ay * e * e);
}
float ggxG1(V3 w, float ax, float ay) { return 1.0f / (1.0f + ggxLambda(w, ax, ay)); }
float ggxG(V3 wo, V3 wi, float ax, float ay) { return 1.0f / (1.0f + ggxLambda(wo, ax, ay) + ggxLambda(wi, ax, ay)); }
float ggxVisiblePdf(V3 wo, V3 h, float ax, float ay) {  // microfacet.h:68-72
    const float cosO = std::abs(wo.z);
    if (cosO <= 0.0f) return 0.0f;
    return ggxG1(wo, ax, ay) * std::abs(dot(wo, h)) * ggxD(h, ax, ay) / cosO;
}
V3 ggxSampleVisibleNormal(V3 wo, float ax, float ay, float u1, float u2) {  // microfacet.h:81-100
    ax = ggxClampAlpha(ax);
    ay = ggxClampAlpha(ay);
    if (wo.z < 0.0f) wo = -wo;
    const V3 vh = normalized(V3(ax * wo.x, ay * wo.y, wo.z));
    const float lenSq = vh.x * vh.x + vh.y * vh.y;
    const V3 t1 = lenSq > 0.0f ? V3(-vh.y, vh.x, 0.0f) / std::sqrt(lenSq) : V3(1.0f, 0.0f, 0.0f);
    const V3 t2 = cross(vh, t1);
    const float r = std::sqrt(u1);
    const float phi = TWO_PI * u2;
    const float p1 = r * std::cos(phi);
    float p2 = r * std::sin(phi);
    const float s = 0.5f * (1.0f + vh.z);
    p2 = (1.0f - s) * std::sqrt(std::max(0.0f, 1.0f - p1 * p1)) + s * p2;
    const V3 nh = t1 * p1 + t2 * p2 + vh * std::sqrt(std::max(0.0f, 1.0f - p1 * p1 - p2 * p2));
    return normalized(V3(ax * nh.x, ay * nh.y, std::max(1e-6f, nh.z)));
}

// --------------------------------------------------------------------------------- microfacet.cpp
namespace {
constexpr int kN = 32;
using Table2D = std::array<float, kN * kN>;
using Table1D = std::array<float, kN>;

float lookup2D(const Table2D& t, float mu, float alpha) {  // microfacet.cpp:20-32
    const float x = clampf(mu, 0.0f, 1.0f) * (kN - 1);
    const float y = clampf(alpha, 0.0f, 1.0f) * (kN - 1);
    const int x0 = std::clamp(static_cast<int>(std::floor(x)), 0, kN - 1);
    const int y0 = std::clamp(static_cast<int>(std::floor(y)), 0, kN - 1);
    const int x1 = std::min(x0 + 1, kN - 1), y1 = std::min(y0 + 1, kN - 1);
    const float fx = clampf(x - static_cast<float>(x0), 0.0f, 1.0f);
    const float fy = clampf(y - static_cast<float>(y0), 0.0f, 1.0f);
    auto at = [&](int ix, int iy) { return t[static_cast<size_t>(iy * kN + ix)]; };
    const float a = at(x0, y0) * (1.0f - fx) + at(x1, y0) * fx;
    const float b = at(x0, y1) * (1.0f - fx) + at(x1, y1) * fx;
    return a * (1.0f - fy) + b * fy;
}
float lookup1D(const Table1D& t, float alpha) {  // microfacet.cpp:34-40
    const float y = clampf(alpha, 0.0f, 1.0f) * (kN - 1);
    const int y0 = std::clamp(static_cast<int>(std::floor(y)), 0, kN - 1);
    const int y1 = std::min(y0 + 1, kN - 1);
    const float fy = clampf(y - static_cast<float>(y0), 0.0f, 1.0f);
    return t[static_cast<size_t>(y0)] * (1.0f - fy) + t[static_cast<size_t>(y1)] * fy;
}
float nodeMu(int im) { return std::max(static_cast<float>(im) / (kN - 1), 1e-3f); }

struct GgxTables {  // microfacet.cpp:46-100
    Table2D e{}, ea{}, eb{};
    Table1D avg0{}, avg1{}, avg2{};
    GgxTables() {
        constexpr int kS = 24;
        for (int ia = 0; ia < kN; ++ia) {
            const float alpha = static_cast<float>(ia) / (kN - 1);
            for (int im = 0; im < kN; ++im) {
                const float mu = nodeMu(im);
                const V3 wo(std::sqrt(std::max(0.0f, 1.0f - mu * mu)), 0.0f, mu);
                const float g1 = ggxG1(wo, alpha, alpha);
                double sum = 0.0, sumA = 0.0, sumB = 0.0;
                for (int sy = 0; sy < kS; ++sy)
                    for (int sx = 0; sx < kS; ++sx) {
                        const float u1 = (static_cast<float>(sx) + 0.5f) / kS;
                        const float u2 = (static_cast<float>(sy) + 0.5f) / kS;
                        const V3 h = ggxSampleVisibleNormal(wo, alpha, alpha, u1, u2);
                        const float c = dot(wo, h);
                        const V3 wi = h * (2.0f * c) - wo;
                        if (wi.z <= 0.0f) continue;
                        const double w = ggxG(wo, wi, alpha, alpha) / g1;
                        const double s5 = pow5(1.0f - clampf(c, 0.0f, 1.0f));
                        sum += w;
                        sumA += w * (1.0 - s5);
                        sumB += w * s5;
                    }
                const size_t k = static_cast<size_t>(ia * kN + im);
                e[k] = static_cast<float>(sum / (kS * kS));
                ea[k] = static_cast<float>(sumA / (kS * kS));
                eb[k] = static_cast<float>(sumB / (kS * kS));
            }
        }
        constexpr int kM = 256;
        for (int ia = 0; ia < kN; ++ia) {
            const float alpha = static_cast<float>(ia) / (kN - 1);
            double a0 = 0.0, a1 = 0.0, a2 = 0.0;
            for (int i = 0; i < kM; ++i) {
                const float mu = (static_cast<float>(i) + 0.5f) / kM;
                const float E = clampf(lookup2D(e, mu, alpha), 1e-3f, 1.0f);
                const float m = 1.0f / E - 1.0f;
                const float A = lookup2D(ea, mu, alpha), B = lookup2D(eb, mu, alpha);
                const double w = 2.0 * mu / kM;
                a0 += w * B;
                a1 += w * (A + B * m);
                a2 += w * (A * m);
            }
            avg0[static_cast<size_t>(ia)] = static_cast<float>(a0);
            avg1[static_cast<size_t>(ia)] = static_cast<float>(a1);
            avg2[static_cast<size_t>(ia)] = static_cast<float>(a2);
        }
    }
};
const GgxTables& ggxTables() {
    static const GgxTables t;
    return t;
}

struct SheenTable {  // microfacet.cpp:108-133
    Table2D e{};
    SheenTable() {
        constexpr int kC = 96, kP = 48;
        for (int ir = 0; ir < kN; ++ir) {
            const float r = std::max(static_cast<float>(ir) / (kN - 1), 0.02f);
            for (int im = 0; im < kN; ++im) {
                const float mu = nodeMu(im);
                const V3 wo(std::sqrt(std::max(0.0f, 1.0f - mu * mu)), 0.0f, mu);
                double sum = 0.0;
                for (int i = 0; i < kC; ++i) {
                    const float cz = (static_cast<float>(i) + 0.5f) / kC;
                    const float s = std::sqrt(std::max(0.0f, 1.0f - cz * cz));
                    for (int j = 0; j < kP; ++j) {
                        const float phi = PI * (static_cast<float>(j) + 0.5f) / kP;
                        const V3 wi(s * std::cos(phi), s * std::sin(phi), cz);
                        const V3 h = normalized(wo + wi);
                        sum += charlieD(h.z, r) * charlieV(cz, mu) * cz;
                    }
                }
                e[static_cast<size_t>(ir * kN + im)] = static_cast<float>(sum * (1.0 / kC) * (TWO_PI / kP));
            }
        }
    }
};
const SheenTable& sheenTable() {
    static const SheenTable t;
    return t;
}
}  // namespace

float ggxDirectionalAlbedo(float cosTheta, float alpha) {  // microfacet.cpp:142-144
    return clampf(lookup2D(ggxTables().e, cosTheta, alpha), 1e-3f, 1.0f);
}
float ggxSchlickAlbedo(float cosTheta, float alpha, float f0) {  // microfacet.cpp:146-151
    const GgxTables& t = ggxTables();
    const float E = clampf(lookup2D(t.e, cosTheta, alpha), 1e-3f, 1.0f);
    const float ss = f0 * lookup2D(t.ea, cosTheta, alpha) + lookup2D(t.eb, cosTheta, alpha);
    return clampf(ss * (1.0f + f0 * (1.0f / E - 1.0f)), 0.0f, 1.0f);
}
float ggxSchlickAverageAlbedo(float alpha, float f0) {  // microfacet.cpp:153-157
    const GgxTables& t = ggxTables();
    const float v = lookup1D(t.avg0, alpha) + f0 * lookup1D(t.avg1, alpha) + f0 * f0 * lookup1D(t.avg2, alpha);
    return clampf(v, 0.0f, 1.0f);
}
float charlieSheenAlbedo(float cosTheta, float roughness) {  // microfacet.cpp:159-161
    return std::max(lookup2D(sheenTable().e, cosTheta, roughness), 0.0f);
}

// ------------------------------------------------------------------------------ disney.cpp / disney.h
namespace {

// DisneyMaterial setters' clamps (disney.h:22-24, :62-126) applied to the raw descriptor.
struct Params {
    V3 baseColor;
    float metallic, roughness, specular, clearCoat, clearCoatRoughness, ior, specularTransmission, coatIor;
    V3 coatTint, sheenColor;
    bool hasCustomSheenColor;
    float anisotropy, sheen, sheenRoughness, diffuseRoughness, diffuseTransmission, subsurface;
    V3 edgeTint;
    float thinFilmThickness, thinFilmIor;
    V3 normal, tangent;
    bool frontFace;
    bool inFlake = false;       // the point lies on a metallic flake (disney.h:201-203)
    V3 flakeNormal{0.0f, 1.0f, 0.0f};
    V3 flakeColor{0.9f, 0.9f, 0.9f};
};

V3 c3(const float* v) { return V3(v[0], v[1], v[2]); }

constexpr float FLAKE_ALPHA = 0.06f;  // disney.cpp:75

// lowbias32 integer hash, cell hash and unit float of the flake cells (disney.cpp:407-421).
uint32_t flakeHash(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

uint32_t flakeCellHash(int x, int y, int z) {
    return flakeHash(static_cast<uint32_t>(x) * 0x8da6b343u ^
                     flakeHash(static_cast<uint32_t>(y) * 0xd8163841u ^ flakeHash(static_cast<uint32_t>(z) * 0xcb1ab31fu)));
}

float flakeUnit(uint32_t h) { return static_cast<float>(h >> 8) * 0x1p-24f; }

// resolveFlake (disney.cpp:430-460): the nearest of the 27 neighbouring cells' seeds (3-D Voronoi) decides
// whether the point is on a flake (density) and the flake's tilt: tan^2 theta = r^2 u1 / (1 - u1), phi = 2 pi u2
// about the shading normal.
void resolveFlake(V3 point, float size, float density, float rough, Params& p) {
    const V3 q = point / size;
    const int cx = static_cast<int>(std::floor(q.x));
    const int cy = static_cast<int>(std::floor(q.y));
    const int cz = static_cast<int>(std::floor(q.z));
    float best = 1e30f;
    uint32_t bestHash = 0;
    for (int dz = -1; dz <= 1; ++dz)
        for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx) {
                const uint32_t h = flakeCellHash(cx + dx, cy + dy, cz + dz);
                const V3 f(static_cast<float>(cx + dx) + flakeUnit(h),
                           static_cast<float>(cy + dy) + flakeUnit(flakeHash(h ^ 0x68e31da4u)),
                           static_cast<float>(cz + dz) + flakeUnit(flakeHash(h ^ 0xb5297a4du)));
                const float d2 = lengthSq(q - f);
                if (d2 < best) {
                    best = d2;
                    bestHash = h;
                }
            }
    if (flakeUnit(flakeHash(bestHash ^ 0x1b56c4e9u)) >= density) return;
    const float u1 = std::min(flakeUnit(flakeHash(bestHash ^ 0x2c1b3c6du)), 0.999f);
    const float u2 = flakeUnit(flakeHash(bestHash ^ 0x297a2d39u));
    const float tan2 = rough * rough * u1 / (1.0f - u1);
    const float cosT = 1.0f / std::sqrt(1.0f + tan2);
    const float sinT = std::sqrt(std::max(0.0f, 1.0f - cosT * cosT));
    const float phi = TWO_PI * u2;
    const Frame f(p.normal);
    p.inFlake = true;
    p.flakeNormal = normalized(f.toWorld(V3(sinT * std::cos(phi), sinT * std::sin(phi), cosT)));
}
V3 clamp01(V3 c) { return V3(clampf(c.x, 0.0f, 1.0f), clampf(c.y, 0.0f, 1.0f), clampf(c.z, 0.0f, 1.0f)); }

Params resolve(const pe_material_desc& d, const ShadingPoint& sp) {  // disney.cpp:543-603 (no textures)
    Params p;
    p.baseColor = c3(d.base_color);
    p.metallic = d.metallic;
    p.roughness = std::max(0.001f, d.roughness);
    p.specular = d.specular;
    p.clearCoat = clampf(d.clearcoat, 0.0f, 1.0f);
    p.clearCoatRoughness = std::max(0.001f, d.clearcoat_roughness);
    p.ior = clampf(d.ior, 1.0001f, 3.0f);
    p.specularTransmission = clampf(d.specular_transmission, 0.0f, 1.0f);
    p.coatIor = clampf(d.coat_ior, 1.0001f, 3.0f);
    p.coatTint = clamp01(c3(d.coat_tint));
    p.hasCustomSheenColor = d.has_sheen_color != 0;
    p.sheenColor = p.hasCustomSheenColor ? clamp01(c3(d.sheen_color)) : V3(1.0f);
    p.anisotropy = clampf(d.anisotropy, 0.0f, 1.0f);
    p.sheen = clampf(d.sheen, 0.0f, 1.0f);
    p.sheenRoughness = clampf(d.sheen_roughness, 0.02f, 1.0f);
    p.diffuseRoughness = clampf(d.diffuse_roughness, 0.0f, 1.0f);
    p.diffuseTransmission = clampf(d.diffuse_transmission, 0.0f, 1.0f);
    p.subsurface = clampf(d.subsurface, 0.0f, 1.0f);
    p.edgeTint = clamp01(c3(d.edge_tint));
    p.thinFilmThickness = clampf(d.thin_film_thickness, 0.0f, 2000.0f);  // setThinFilm (disney.h:111-114)
    p.thinFilmIor = clampf(d.thin_film_ior, 1.0f, 3.0f);
    p.normal = sp.normal;
    p.tangent = sp.tangent;
    p.frontFace = sp.frontFace;
    p.flakeColor = clamp01(c3(d.flake_color));  // setFlakes clamps (disney.h:120-125)
    if (d.flake_density > 0.0f)
        resolveFlake(sp.point, std::max(d.flake_size, 1e-6f), clampf(d.flake_density, 0.0f, 1.0f),
                     clampf(d.flake_roughness, 0.0f, 1.0f), p);
    return p;
}

float f0FromIor(float ior) {  // disney.cpp:25-28
    const float r = (ior - 1.0f) / (ior + 1.0f);
    return r * r;
}
float dielectricWeight(const Params& p) { return clampf(2.0f * p.specular, 0.0f, 1.0f); }  // :34-36
float dielectricF0(const Params& p) {  // :40-42
    return clampf(f0FromIor(p.ior) * std::max(1.0f, 2.0f * p.specular), 0.0f, 1.0f);
}
float coatF0(const Params& p) { return clampf(f0FromIor(p.coatIor), 0.0f, 1.0f); }  // :44-46
float schlick(float f0, float c) { return f0 + (1.0f - f0) * pow5(1.0f - clampf(c, 0.0f, 1.0f)); }  // :48-50
float schlickF(float f0, float cosTheta) {  // fresnel.h:20-24
    const float m = 1.0f - clampf(cosTheta, 0.0f, 1.0f);
    const float m2 = m * m;
    return f0 + (1.0f - f0) * m2 * m2 * m;
}
V3 fresnelF82Tint(V3 f0, V3 tint, float cosTheta) {  // fresnel.h:31-47
    constexpr float muBar = 1.0f / 7.0f;
    constexpr float oneMinus = 1.0f - muBar;
    const float denom = muBar * oneMinus * oneMinus * oneMinus * oneMinus * oneMinus * oneMinus;
    const float mu = clampf(cosTheta, 0.0f, 1.0f);
    const float m = 1.0f - mu;
    const float m6 = m * m * m * m * m * m;
    V3 out;
    for (int c = 0; c < 3; ++c) {
        const float fsBar = schlickF(f0[c], muBar);
        const float b = fsBar * (1.0f - tint[c]) / denom;
        out[c] = clampf(schlickF(f0[c], mu) - b * mu * m6, 0.0f, 1.0f);
    }
    return out;
}
void anisoAlpha(float roughness, float anisotropy, float& ax, float& ay) {  // :55-60
    const float aspect = std::sqrt(std::max(0.1f, 1.0f - 0.9f * clampf(anisotropy, 0.0f, 1.0f)));
    const float a = roughness * roughness;
    ax = a / aspect;
    ay = a * aspect;
}
Frame shadingFrame(V3 nIn, V3 tangentIn) {  // :62-73
    const V3 n = lengthSq(nIn) > 0.0f ? normalized(nIn) : V3(0, 1, 0);
    V3 t = tangentIn - n * dot(tangentIn, n);
    if (lengthSq(t) < 1e-8f) return Frame(n);
    t = normalized(t);
    V3 b = cross(n, t);
    if (lengthSq(b) < 1e-8f) return Frame(n);
    b = normalized(b);
    t = normalized(cross(b, n));
    return Frame(t, b, n);
}
// ------------------------------------------------------------------ thin-film interference (fresnel.h:49-107)
using cfloat = std::complex<float>;

cfloat gulbrandsenIor(float r, float g) {  // fresnel.h:49-58: metal n + ik from reflectivity and edge tint
    r = clampf(r, 0.0f, 0.99f);
    g = clampf(g, 0.0f, 1.0f);
    const float sr = std::sqrt(r);
    const float nMin = (1.0f - r) / (1.0f + r);
    const float nMax = (1.0f + sr) / (1.0f - sr);
    const float n = g * nMin + (1.0f - g) * nMax;
    const float k2 = (r * (n + 1.0f) * (n + 1.0f) - (n - 1.0f) * (n - 1.0f)) / (1.0f - r);
    return {n, std::sqrt(std::max(0.0f, k2))};
}

// Airy sum of the waves reflected inside a film n2 of thickness d between n1 and a (complex) base n3,
// unpolarised (fresnel.h:74-92).
float thinFilmReflectance(float cos1, float n1, float n2, cfloat n3, float thicknessNm, float lambdaNm) {
    cos1 = clampf(cos1, 0.0f, 1.0f);
    const float sin1sq = 1.0f - cos1 * cos1;
    const float sin2sq = (n1 / n2) * (n1 / n2) * sin1sq;
    if (sin2sq >= 1.0f) return 1.0f;
    const float cos2 = std::sqrt(1.0f - sin2sq);
    const cfloat cos3 = std::sqrt(cfloat(1.0f) - (n1 * n1 * sin1sq) / (n3 * n3));
    const float r12s = (n1 * cos1 - n2 * cos2) / (n1 * cos1 + n2 * cos2);
    const float r12p = (n2 * cos1 - n1 * cos2) / (n2 * cos1 + n1 * cos2);
    const cfloat r23s = (n2 * cos2 - n3 * cos3) / (n2 * cos2 + n3 * cos3);
    const cfloat r23p = (n3 * cos2 - n2 * cos3) / (n3 * cos2 + n2 * cos3);
    const float delta = 4.0f * PI * n2 * thicknessNm * cos2 / lambdaNm;
    const cfloat e = std::polar(1.0f, delta);
    const cfloat rs = (r12s + r23s * e) / (1.0f + r12s * r23s * e);
    const cfloat rp = (r12p + r23p * e) / (1.0f + r12p * r23p * e);
    return clampf(0.5f * (std::norm(rs) + std::norm(rp)), 0.0f, 1.0f);
}

V3 thinFilmRGB(float cos1, float filmIor, const cfloat baseIor[3], float thicknessNm) {  // :94-107
    static constexpr float kLambda[3][3] = {{610.0f, 640.0f, 670.0f}, {515.0f, 540.0f, 565.0f}, {440.0f, 465.0f, 490.0f}};
    V3 out;
    for (int c = 0; c < 3; ++c) {
        float s = 0.0f;
        for (float l : kLambda[c]) s += thinFilmReflectance(cos1, 1.0f, filmIor, baseIor[c], thicknessNm, l);
        out[c] = s / 3.0f;
    }
    return out;
}

V3 specularFresnel(const Params& p, float cosD) {  // :79-96
    const float f0d = dielectricF0(p);
    const float w = dielectricWeight(p);
    if (p.thinFilmThickness <= 0.0f) {
        const V3 fd(w * schlickF(f0d, cosD));
        if (p.metallic <= 0.0f) return fd;
        return lerpc(fd, fresnelF82Tint(p.baseColor, p.edgeTint, cosD), p.metallic);
    }
    const float sd = std::sqrt(clampf(f0d, 0.0f, 0.99f));
    const cfloat nd((1.0f + sd) / (1.0f - sd), 0.0f);
    const cfloat dielectric[3] = {nd, nd, nd};
    const V3 fd = thinFilmRGB(cosD, p.thinFilmIor, dielectric, p.thinFilmThickness) * w;
    if (p.metallic <= 0.0f) return fd;
    cfloat metal[3];
    for (int c = 0; c < 3; ++c) metal[c] = gulbrandsenIor(p.baseColor[c], p.edgeTint[c]);
    return lerpc(fd, thinFilmRGB(cosD, p.thinFilmIor, metal, p.thinFilmThickness), p.metallic);
}
V3 multiScatterFactor(const Params& p, float cosO, float ax, float ay) {  // :99-103
    const float E = ggxDirectionalAlbedo(cosO, std::sqrt(ax * ay));
    const V3 f0 = lerpc(V3(dielectricF0(p)), p.baseColor, p.metallic);
    return V3(1.0f) + f0 * (1.0f / E - 1.0f);
}
struct DielectricLayer {  // :128-146
    float alpha = 0.0f, f0 = 0.0f, weight = 1.0f, invOneMinusAvg = 1.0f;
    explicit DielectricLayer(const Params& p) {
        float ax, ay;
        anisoAlpha(p.roughness, p.anisotropy, ax, ay);
        alpha = std::sqrt(ax * ay);
        f0 = dielectricF0(p);
        weight = dielectricWeight(p);
        invOneMinusAvg = 1.0f / std::max(1e-4f, 1.0f - weight * ggxSchlickAverageAlbedo(alpha, f0));
    }
    float transmit(float mu) const { return std::max(0.0f, 1.0f - weight * ggxSchlickAlbedo(mu, alpha, f0)); }
    float factor(float muO, float muI) const { return transmit(muO) * transmit(muI) * invOneMinusAvg; }
};
float coatTransmit(const Params& p, float mu) {  // :149-153
    if (p.clearCoat <= 0.0f) return 1.0f;
    const float alphaC = p.clearCoatRoughness * p.clearCoatRoughness;
    return 1.0f - p.clearCoat * ggxSchlickAlbedo(mu, alphaC, coatF0(p));
}
V3 coatTintFactor(const Params& p) {  // :155-157
    return p.clearCoat > 0.0f ? lerpc(V3(1.0f), p.coatTint, p.clearCoat) : V3(1.0f);
}
V3 resolvedSheenColor(const Params& p) {  // :160-166
    if (p.sheen <= 0.0f || p.metallic >= 1.0f) return V3(0.0f);
    if (p.hasCustomSheenColor) return p.sheenColor * (p.sheen * (1.0f - p.metallic));
    const float lum = luminance(p.baseColor);
    const V3 tint = lum > 1e-6f ? p.baseColor * (1.0f / lum) : V3(1.0f);
    return (V3(0.5f) + clamp01(tint) * 0.5f) * (p.sheen * (1.0f - p.metallic));
}
float sheenScale(const Params& p, float s, float muO, float muI) {  // :169-174
    if (s <= 0.0f) return 1.0f;
    const float eo = std::min(1.0f, charlieSheenAlbedo(muO, p.sheenRoughness));
    const float ei = std::min(1.0f, charlieSheenAlbedo(muI, p.sheenRoughness));
    return std::max(0.0f, std::min(1.0f - s * eo, 1.0f - s * ei));
}
V3 evalSheen(V3 color, const Params& p, float cosO, float cosI, float cosH) {  // :177-181
    const float norm = std::max({1.0f, charlieSheenAlbedo(cosO, p.sheenRoughness), charlieSheenAlbedo(cosI, p.sheenRoughness)});
    return color * (charlieD(cosH, p.sheenRoughness) * charlieV(cosI, cosO) / norm);
}
V3 evalSpecTransmission(V3 wo, V3 wi, const Params& p) {  // :185-205
    const float stW = clampf(p.specularTransmission, 0.0f, 1.0f) * (1.0f - p.metallic);
    if (stW <= 0.0f || wo.z <= 0.0f || wi.z >= 0.0f) return V3(0.0f);
    const float etap = p.frontFace ? p.ior : (1.0f / p.ior);
    V3 wm = wi * etap + wo;
    if (lengthSq(wm) <= 1e-12f) return V3(0.0f);
    wm = normalized(wm);
    if (wm.z < 0.0f) wm = -wm;
    const float woDotH = dot(wo, wm), wiDotH = dot(wi, wm);
    if (woDotH <= 0.0f || wiDotH >= 0.0f) return V3(0.0f);
    const float alpha = ggxClampAlpha(p.roughness * p.roughness);
    const float F = fresnelDielectric(p.frontFace ? woDotH : -woDotH, 1.0f, p.ior);
    const float D = ggxD(wm, alpha, alpha);
    const float G = ggxG(wo, V3(wi.x, wi.y, -wi.z), alpha, alpha);
    const float s = wiDotH + woDotH / etap;
    const float denom = s * s * std::abs(wi.z * wo.z);
    if (denom <= 1e-12f) return V3(0.0f);
    const float ft = D * (1.0f - F) * G * std::abs(wiDotH * woDotH) / (denom * etap * etap);
    return p.baseColor * (stW * ft);
}
float pdfSpecTransmission(V3 wo, V3 wi, const Params& p) {  // :207-220
    if (wo.z <= 0.0f || wi.z >= 0.0f) return 0.0f;
    const float etap = p.frontFace ? p.ior : (1.0f / p.ior);
    V3 wm = wi * etap + wo;
    if (lengthSq(wm) <= 1e-12f) return 0.0f;
    wm = normalized(wm);
    if (wm.z < 0.0f) wm = -wm;
    const float woDotH = dot(wo, wm), wiDotH = dot(wi, wm);
    if (woDotH <= 0.0f || wiDotH >= 0.0f) return 0.0f;
    const float alpha = ggxClampAlpha(p.roughness * p.roughness);
    const float visible = ggxVisiblePdf(wo, wm, alpha, alpha);
    const float s = wiDotH + woDotH / etap;
    if (s * s <= 1e-12f) return 0.0f;
    return visible * std::abs(wiDotH) / (s * s);
}
float fonAlbedo(float mu, float r) {  // eon.h:27-34
    constexpr float kFonC1 = 0.5f - 2.0f / (3.0f * PI);
    mu = clampf(mu, 1e-4f, 1.0f);
    const float AF = 1.0f / (1.0f + kFonC1 * r);
    const float BF = r * AF;
    const float Si = std::sqrt(std::max(0.0f, 1.0f - mu * mu));
    const float G = Si * (std::acos(mu) - Si * mu) + (2.0f / 3.0f) * ((Si / mu) * (1.0f - Si * Si * Si) - Si);
    return AF + (BF * INV_PI) * G;
}
V3 eonEval(V3 rho, float r, V3 wi, V3 wo) {  // eon.h:37-51
    constexpr float kFonC1 = 0.5f - 2.0f / (3.0f * PI);
    constexpr float kFonC2 = 2.0f / 3.0f - 28.0f / (15.0f * PI);
    const float muI = wi.z, muO = wo.z;
    if (r <= 0.0f) return rho * INV_PI;
    const float s = dot(wi, wo) - muI * muO;
    const float sOverT = s > 0.0f ? s / std::max(muI, muO) : s;
    const float AF = 1.0f / (1.0f + kFonC1 * r);
    const V3 fss = rho * (INV_PI * AF * (1.0f + r * sOverT));
    const float EFo = fonAlbedo(muO, r), EFi = fonAlbedo(muI, r);
    const float avgEF = AF * (1.0f + kFonC2 * r);
    V3 rhoMs;
    for (int c = 0; c < 3; ++c) rhoMs[c] = rho[c] * rho[c] * avgEF / std::max(1e-6f, 1.0f - rho[c] * (1.0f - avgEF));
    constexpr float eps = 1e-7f;
    const float ms = std::max(eps, 1.0f - EFo) * std::max(eps, 1.0f - EFi) / std::max(eps, 1.0f - avgEF);
    return fss + rhoMs * (INV_PI * ms);
}

// evalFlake / pdfFlake (disney.cpp:104-126): a narrow GGX lobe about the flake normal with per-channel Schlick
// reflectance of the flake colour, both directions in world space.
V3 evalFlake(V3 wo, V3 wi, const Params& p) {
    const Frame ff(p.flakeNormal);
    const V3 o = ff.toLocal(wo), i = ff.toLocal(wi);
    if (o.z <= 0.0f || i.z <= 0.0f) return V3(0.0f);
    const V3 h = normalized(o + i);
    const float D = ggxD(h, FLAKE_ALPHA, FLAKE_ALPHA);
    const float G = ggxG(o, i, FLAKE_ALPHA, FLAKE_ALPHA);
    const float c = dot(o, h);
    const V3 F(schlickF(p.flakeColor.x, c), schlickF(p.flakeColor.y, c), schlickF(p.flakeColor.z, c));
    return F * (D * G / (4.0f * o.z * i.z));
}

float pdfFlake(V3 wo, V3 wi, const Params& p) {
    const Frame ff(p.flakeNormal);
    const V3 o = ff.toLocal(wo), i = ff.toLocal(wi);
    if (o.z <= 0.0f || i.z <= 0.0f) return 0.0f;
    const V3 h = normalized(o + i);
    const float c = dot(o, h);
    if (c <= 0.0f) return 0.0f;
    return ggxVisiblePdf(o, h, FLAKE_ALPHA, FLAKE_ALPHA) / (4.0f * c);
}

V3 evalLobes(V3 wo, V3 wi, const Params& p, LobeSplit* split = nullptr) {  // :223-311
    const Frame frame = shadingFrame(p.normal, p.tangent);
    const V3 woL = frame.toLocal(wo), wiL = frame.toLocal(wi);
    if (woL.z <= 0.0f) return V3(0.0f);
    const float cosThetaO = woL.z;
    const float coatO = coatTransmit(p, cosThetaO);
    const V3 cTint = coatTintFactor(p);
    const float stW = p.inFlake ? 0.0f : clampf(p.specularTransmission, 0.0f, 1.0f) * (1.0f - p.metallic);
    const float opaqueDielectric = (1.0f - p.metallic) * (1.0f - clampf(p.specularTransmission, 0.0f, 1.0f));
    const float trans = p.inFlake ? 0.0f : clampf(p.diffuseTransmission, 0.0f, 1.0f) * opaqueDielectric;
    if (wiL.z <= 0.0f) {
        V3 ft(0.0f);
        if (trans > 0.0f) {
            const float muI = -wiL.z;
            const DielectricLayer dl(p);
            const V3 sc = resolvedSheenColor(p);
            const float sh = sheenScale(p, std::max({sc.x, sc.y, sc.z}), cosThetaO, muI);
            ft += p.baseColor * (trans * INV_PI * coatO * dl.factor(cosThetaO, muI) * sh) * cTint;
        }
        if (stW > 0.0f) ft += evalSpecTransmission(woL, wiL, p) * coatO * cTint;
        if (split) split->transmission = ft;
        return ft;
    }
    const V3 hL = normalized(woL + wiL);
    const float cosThetaI = wiL.z;
    const float cosThetaD = dot(woL, hL);
    V3 base, baseDiffuse;
    if (p.inFlake) {
        base = evalFlake(wo, wi, p);
    } else {
        const DielectricLayer dl(p);
        const float diffW = opaqueDielectric * (1.0f - clampf(p.diffuseTransmission, 0.0f, 1.0f)) * (1.0f - p.subsurface);
        const V3 fDiffuse = diffW > 0.0f ? eonEval(p.baseColor, p.diffuseRoughness, wiL, woL) *
                                               (diffW * dl.factor(cosThetaO, cosThetaI))
                                         : V3(0.0f);
        float ax, ay;
        anisoAlpha(p.roughness, p.anisotropy, ax, ay);
        const float D = ggxD(hL, ax, ay);
        const float G = ggxG(woL, wiL, ax, ay);
        V3 F = specularFresnel(p, cosThetaD);
        if (stW > 0.0f && !p.frontFace) {
            const float fInt = fresnelDielectric(-cosThetaD, 1.0f, p.ior);
            F = lerpc(F, V3(fInt), stW);
        }
        const V3 fSpecular = (F * D * G) / (4.0f * cosThetaO * cosThetaI) * multiScatterFactor(p, cosThetaO, ax, ay);
        base = fDiffuse + fSpecular;
        baseDiffuse = fDiffuse;
        const V3 sc = resolvedSheenColor(p);
        if (!isBlack(sc)) {
            const float s = sheenScale(p, std::max({sc.x, sc.y, sc.z}), cosThetaO, cosThetaI);
            base = base * s + evalSheen(sc, p, cosThetaO, cosThetaI, hL.z);
            baseDiffuse = baseDiffuse * s;
        }
    }
    if (p.clearCoat <= 0.0f) {  // :290-296: the diffuse lobe, the rest (specular, sheen) glossy
        if (split) {
            split->diffuse = baseDiffuse;
            split->glossy = base - baseDiffuse;
        }
        return base;
    }
    const float alphaC = p.clearCoatRoughness * p.clearCoatRoughness;
    const float Dc = ggxD(hL, alphaC, alphaC);
    const float Gc = ggxG(woL, wiL, alphaC, alphaC);
    const float f0c = coatF0(p);
    const float Fc = schlick(f0c, cosThetaD);
    const float msC = 1.0f + f0c * (1.0f / ggxDirectionalAlbedo(cosThetaO, alphaC) - 1.0f);
    const V3 fCoat(p.clearCoat * Fc * Dc * Gc / (4.0f * cosThetaO * cosThetaI) * msC);
    const V3 coatAtt = cTint * (coatO * coatTransmit(p, cosThetaI));
    if (split) {  // :306-309
        split->diffuse = baseDiffuse * coatAtt;
        split->glossy = (base - baseDiffuse) * coatAtt + fCoat;
    }
    return base * coatAtt + fCoat;
}

void lobeWeights(const Params& p, float& pCoat, float& pSpec, float& pDiffRefl, float& pTrans, float& pSss,
                 float& pSpecTrans) {  // :313-338
    pCoat = 0.25f * p.clearCoat;
    if (p.inFlake) {
        pSpec = 1.0f - pCoat;
        pDiffRefl = pTrans = pSss = pSpecTrans = 0.0f;
        return;
    }
    const float rem = 1.0f - pCoat;
    const float st = clampf(p.specularTransmission, 0.0f, 1.0f) * (1.0f - p.metallic);
    pSpec = rem * lerpf(0.5f * (st > 0.0f ? 1.0f : dielectricWeight(p)), 1.0f, p.metallic);
    float pRest = rem - pSpec;
    const float pSt = pRest * st;
    pRest -= pSt;
    const float dTrans = clampf(p.diffuseTransmission, 0.0f, 1.0f);
    pTrans = pRest * dTrans;
    pDiffRefl = pRest - pTrans;
    pSss = p.metallic < 1.0f ? pDiffRefl * p.subsurface : 0.0f;
    pDiffRefl -= pSss;
    pSpecTrans = pSt;
}

float pdfLobes(V3 wo, V3 wi, const Params& p) {  // :340-378
    const Frame frame = shadingFrame(p.normal, p.tangent);
    const V3 woL = frame.toLocal(wo), wiL = frame.toLocal(wi);
    if (woL.z <= 0.0f) return 0.0f;
    float pCoat, pSpec, pDiffRefl, pTrans, pSss, pSpecTrans;
    lobeWeights(p, pCoat, pSpec, pDiffRefl, pTrans, pSss, pSpecTrans);
    if (wiL.z <= 0.0f) {
        float pdf = 0.0f;
        if (pTrans > 0.0f) pdf += pTrans * cosineHemispherePdf(-wiL.z);
        if (pSpecTrans > 0.0f) pdf += pSpecTrans * pdfSpecTransmission(woL, wiL, p);
        return pdf;
    }
    const V3 hL = normalized(woL + wiL);
    const float cosThetaD = dot(woL, hL);
    if (cosThetaD <= 0.0f) return 0.0f;
    const float pdfDiffuse = cosineHemispherePdf(wiL.z);
    float pdfSpecular;
    if (p.inFlake) {
        pdfSpecular = pdfFlake(wo, wi, p);
    } else {
        float ax, ay;
        anisoAlpha(p.roughness, p.anisotropy, ax, ay);
        pdfSpecular = ggxVisiblePdf(woL, hL, ax, ay) / (4.0f * cosThetaD);
    }
    float pdfCoat = 0.0f;
    if (pCoat > 0.0f) {
        const float alphaC = p.clearCoatRoughness * p.clearCoatRoughness;
        pdfCoat = ggxVisiblePdf(woL, hL, alphaC, alphaC) / (4.0f * cosThetaD);
    }
    return pDiffRefl * pdfDiffuse + pSpec * pdfSpecular + pCoat * pdfCoat;
}

bool disneySampleFull(const Params& p, V3 wo, float uc, float u0, float u1, BsdfSample& out) {  // :680-729
    out.delta = out.subsurface = out.passThrough = false;
    const Frame frame = shadingFrame(p.normal, p.tangent);
    const V3 woL = frame.toLocal(wo);
    if (woL.z <= 0.0f) return false;
    float pCoat, pSpec, pDiffRefl, pTrans, pSss, pSpecTrans;
    lobeWeights(p, pCoat, pSpec, pDiffRefl, pTrans, pSss, pSpecTrans);
    V3 wiL;
    if (uc < pCoat) {
        const float alphaC = p.clearCoatRoughness * p.clearCoatRoughness;
        wiL = reflectVec(woL, ggxSampleVisibleNormal(woL, alphaC, alphaC, u0, u1));
        if (wiL.z <= 0.0f) return false;
    } else if (uc < pCoat + pSpec) {  // sampleSpecularLobe, :380-396
        if (p.inFlake) {
            const Frame ff(p.flakeNormal);
            const V3 o = ff.toLocal(wo);
            if (o.z <= 0.0f) return false;
            const V3 i = reflectVec(o, ggxSampleVisibleNormal(o, FLAKE_ALPHA, FLAKE_ALPHA, u0, u1));
            if (i.z <= 0.0f) return false;
            wiL = frame.toLocal(ff.toWorld(i));
        } else {
            float ax, ay;
            anisoAlpha(p.roughness, p.anisotropy, ax, ay);
            wiL = reflectVec(woL, ggxSampleVisibleNormal(woL, ax, ay, u0, u1));
        }
        if (!(wiL.z > 0.0f)) return false;
    } else if (uc < pCoat + pSpec + pSpecTrans) {
        const float alpha = ggxClampAlpha(p.roughness * p.roughness);  // sampleSpecTransLobe, :398-405
        const V3 wm = ggxSampleVisibleNormal(woL, alpha, alpha, u0, u1);
        const float eta = p.frontFace ? (1.0f / p.ior) : p.ior;
        if (!refractVec(woL, wm, eta, wiL)) return false;
        if (!(wiL.z < 0.0f)) return false;
    } else if (uc < pCoat + pSpec + pSpecTrans + pDiffRefl) {
        wiL = cosineSampleHemisphere(u0, u1);
    } else if (uc < pCoat + pSpec + pSpecTrans + pDiffRefl + pSss) {
        wiL = cosineSampleHemisphere(u0, u1);
        wiL.z = -wiL.z;
        const float opaqueDielectric = (1.0f - p.metallic) * (1.0f - clampf(p.specularTransmission, 0.0f, 1.0f));
        const float trans = clampf(p.diffuseTransmission, 0.0f, 1.0f);
        const float coatO = coatTransmit(p, woL.z);
        const V3 sc = resolvedSheenColor(p);
        const float sheenO = sheenScale(p, std::max({sc.x, sc.y, sc.z}), woL.z, 1.0f);
        const float layerO = DielectricLayer(p).transmit(woL.z);
        const float w = p.subsurface * opaqueDielectric * (1.0f - trans) * coatO * sheenO * layerO / std::max(pSss, 1e-6f);
        out.wi = normalized(frame.toWorld(wiL));
        out.subsurface = true;
        out.weight = coatTintFactor(p) * w;
        out.pdf = pSss * cosineHemispherePdf(-wiL.z);
        out.f = out.weight * out.pdf / std::max(-wiL.z, 1e-6f);
        return out.pdf > 0.0f;
    } else {
        wiL = cosineSampleHemisphere(u0, u1);
        wiL.z = -wiL.z;
    }
    out.wi = normalized(frame.toWorld(wiL));
    out.f = evalLobes(wo, out.wi, p);
    out.pdf = pdfLobes(wo, out.wi, p);
    return out.pdf > 0.0f;
}

// ---------------------------------------------------------------------------------- spectral.cpp
float cieLobe(float x, float mu, float s1, float s2) {  // :14-17
    const float t = (x - mu) / (x < mu ? s1 : s2);
    return std::exp(-0.5f * t * t);
}

V3 rawRgb(float l) {  // :19-30 (Wyman, Sloan, Shirley 2013 fit, XYZ -> linear sRGB)
    const float X = 1.056f * cieLobe(l, 599.8f, 37.9f, 31.0f) + 0.362f * cieLobe(l, 442.0f, 16.0f, 26.7f) -
                    0.065f * cieLobe(l, 501.1f, 20.4f, 26.2f);
    const float Y = 0.821f * cieLobe(l, 568.8f, 46.9f, 40.5f) + 0.286f * cieLobe(l, 530.9f, 16.3f, 31.1f);
    const float Z = 1.217f * cieLobe(l, 437.0f, 11.8f, 36.0f) + 0.681f * cieLobe(l, 459.0f, 26.0f, 13.8f);
    const float r = 3.2404542f * X - 1.5371385f * Y - 0.4985314f * Z;
    const float g = -0.9692660f * X + 1.8760108f * Y + 0.0415560f * Z;
    const float b = 0.0556434f * X - 0.2040259f * Y + 1.0572252f * Z;
    return V3(std::max(0.0f, r), std::max(0.0f, g), std::max(0.0f, b));
}

constexpr float kLambdaMin = 380.0f, kLambdaMax = 730.0f;
constexpr int kSpectralBins = 350;

struct SpectralTables {  // :34-62
    V3 norm{1.0f};
    float cdf[kSpectralBins + 1] = {};
    float pdf[kSpectralBins] = {};
    SpectralTables() {
        V3 sum(0.0f);
        constexpr int kSteps = 3500;
        const float dl = (kLambdaMax - kLambdaMin) / static_cast<float>(kSteps);
        for (int i = 0; i < kSteps; ++i) sum += rawRgb(kLambdaMin + (static_cast<float>(i) + 0.5f) * dl);
        const V3 mean = sum / static_cast<float>(kSteps);
        norm = V3(1.0f / mean.x, 1.0f / mean.y, 1.0f / mean.z);
        const float bw = (kLambdaMax - kLambdaMin) / static_cast<float>(kSpectralBins);
        float acc = 0.0f;
        cdf[0] = 0.0f;
        float s[kSpectralBins];
        for (int i = 0; i < kSpectralBins; ++i) {
            const V3 c = rawRgb(kLambdaMin + (static_cast<float>(i) + 0.5f) * bw) * norm;
            s[i] = c.x + c.y + c.z + 0.02f;
            acc += s[i];
            cdf[i + 1] = acc;
        }
        for (int i = 0; i <= kSpectralBins; ++i) cdf[i] /= acc;
        for (int i = 0; i < kSpectralBins; ++i) pdf[i] = s[i] / (acc * bw);
    }
};

const SpectralTables& spectralTables() {
    static const SpectralTables t;
    return t;
}

// ---------------------------------------------------------------------------------- dielectric.cpp
bool glassSharp(float roughness) { return roughness <= 1e-3f; }                    // :19
float glassAlpha(float roughness) { return ggxClampAlpha(roughness * roughness); }  // :21

struct GlassLocal {
    Frame frame;
    V3 wo;
};

GlassLocal glassFrame(V3 woWorld, const ShadingPoint& sp) {  // :36-45
    const V3 outward = sp.frontFace ? sp.normal : -sp.normal;
    const V3 ng = lengthSq(sp.ng) > 1e-12f ? normalized(sp.ng) : normalized(outward);
    V3 ns = lengthSq(sp.normal) > 1e-12f ? normalized(sp.normal) : ng;
    if (dot(ns, ng) < 0.0f) ns = -ns;
    if (dot(woWorld, ns) * dot(woWorld, ng) <= 0.0f) ns = ng;
    GlassLocal g{Frame(ns), V3(0.0f)};
    g.wo = g.frame.toLocal(woWorld);
    return g;
}

bool generalizedHalf(V3 wo, V3 wi, float ior, V3& wm, float& etap) {  // :55-67
    const float cosO = wo.z, cosI = wi.z;
    if (cosO == 0.0f || cosI == 0.0f) return false;
    const bool reflect = cosO * cosI > 0.0f;
    etap = reflect ? 1.0f : (cosO > 0.0f ? ior : 1.0f / ior);
    wm = wi * etap + wo;
    if (lengthSq(wm) <= 1e-12f) return false;
    wm = normalized(wm);
    if (wm.z < 0.0f) wm = -wm;
    if (dot(wm, wi) * cosI <= 0.0f || dot(wm, wo) * cosO <= 0.0f) return false;
    return true;
}

V3 evalRoughLocal(V3 wo, V3 wi, float ior, float alpha, V3 tint) {  // :79-97
    V3 wm;
    float etap = 1.0f;
    if (!generalizedHalf(wo, wi, ior, wm, etap)) return V3(0.0f);
    const float cosO = wo.z, cosI = wi.z;
    const float woDotH = dot(wo, wm), wiDotH = dot(wi, wm);
    const float F = fresnelDielectric(woDotH, 1.0f, ior);
    const float D = ggxD(wm, alpha, alpha);
    const float G = ggxG(wo, wi, alpha, alpha);
    if (cosO * cosI > 0.0f) return tint * (D * G * F / std::abs(4.0f * cosI * cosO));
    const float s = wiDotH + woDotH / etap;
    const float denom = s * s * cosI * cosO;
    if (std::abs(denom) <= 1e-12f) return V3(0.0f);
    const float ft = D * (1.0f - F) * G * std::abs(wiDotH * woDotH / denom);
    return tint * (ft / (etap * etap));
}

float pdfRoughLocal(V3 wo, V3 wi, float ior, float alpha) {  // :105-119
    V3 wm;
    float etap = 1.0f;
    if (!generalizedHalf(wo, wi, ior, wm, etap)) return 0.0f;
    const float woDotH = dot(wo, wm), wiDotH = dot(wi, wm);
    const float R = fresnelDielectric(woDotH, 1.0f, ior);
    const float visible = ggxVisiblePdf(wo, wm, alpha, alpha);
    if (wo.z * wi.z > 0.0f) return visible / (4.0f * std::abs(woDotH)) * R;
    const float s = wiDotH + woDotH / etap;
    if (s * s <= 1e-12f) return 0.0f;
    return visible * std::abs(wiDotH) / (s * s) * (1.0f - R);
}

struct Glass {  // Dielectric(ior, tint, roughness) + setAbbe / setThinWalled / setThinFilm as ref_driver builds it
    float ior, roughness, abbe, film;
    V3 tint;
    bool thin;
    explicit Glass(const pe_material_desc& d)
        : ior(d.ior), roughness(d.roughness), abbe(d.abbe > 0.0f ? std::max(d.abbe, 10.0f) : 0.0f),
          film(d.thin_film_thickness > 0.0f ? d.thin_film_thickness : 0.0f), tint(c3(d.base_color)),
          thin(d.thin_walled != 0) {}
    float iorAt(float wavelength) const {  // :174-177
        if (abbe <= 0.0f || wavelength <= 0.0f) return ior;
        return std::max(1.0001f, cauchyIor(ior, abbe, wavelength));
    }
    // A soap-bubble film (air | film | air) when the wall has a film, else the slab's incoherent sum
    // 2R / (1 + R) (:150-158).
    V3 thinWallReflectance(float cosTheta) const {
        cosTheta = clampf(std::abs(cosTheta), 0.0f, 1.0f);
        if (film > 0.0f) {
            const cfloat air[3] = {cfloat(1.0f), cfloat(1.0f), cfloat(1.0f)};
            return thinFilmRGB(cosTheta, ior, air, film);
        }
        const float R = fresnelDielectric(cosTheta, 1.0f, ior);
        return V3(2.0f * R / (1.0f + R));
    }
};

bool glassSampleWithLobe(const Glass& m, V3 woWorld, const ShadingPoint& sp, float uc, float u0, float u1, V3& wi,
                         V3& f, float& pdf) {  // :186-253
    const GlassLocal g = glassFrame(woWorld, sp);
    const V3 wo = g.wo;
    if (wo.z == 0.0f) return false;
    const bool entering = wo.z > 0.0f;
    const float ior = m.iorAt(sp.wavelength);
    const float eta = entering ? (1.0f / ior) : ior;
    if (glassSharp(m.roughness)) {
        const float F = fresnelDielectric(wo.z, 1.0f, ior);
        V3 wiLocal(-wo.x, -wo.y, wo.z);
        V3 weight;
        if (uc < F) {
            pdf = F;
            weight = m.tint * F;
        } else {
            const V3 n(0.0f, 0.0f, entering ? 1.0f : -1.0f);
            if (!refractVec(wo, n, eta, wiLocal)) {
                wiLocal = V3(-wo.x, -wo.y, wo.z);
                pdf = 1.0f;
                weight = m.tint;
            } else {
                pdf = 1.0f - F;
                weight = m.tint * ((1.0f - F) * eta * eta);
            }
        }
        wi = normalized(g.frame.toWorld(wiLocal));
        f = weight / std::max(std::abs(dot(wi, sp.normal)), 1e-6f);
        return true;
    }
    const float alpha = glassAlpha(m.roughness);
    const V3 wm = ggxSampleVisibleNormal(wo, alpha, alpha, u0, u1);
    const float R = fresnelDielectric(dot(wo, wm), 1.0f, ior);
    V3 wiLocal;
    if (uc < R) {
        wiLocal = reflectVec(wo, wm);
        if (wiLocal.z * wo.z <= 0.0f) return false;
    } else {
        const V3 n = entering ? wm : -wm;
        if (!refractVec(wo, n, eta, wiLocal)) return false;
        if (wiLocal.z * wo.z >= 0.0f) return false;
    }
    f = evalRoughLocal(wo, wiLocal, ior, alpha, m.tint);
    pdf = pdfRoughLocal(wo, wiLocal, ior, alpha);
    wi = normalized(g.frame.toWorld(wiLocal));
    return pdf > 0.0f && !isBlack(f);
}

bool glassSampleFull(const Glass& m, V3 wo, const ShadingPoint& sp, float uc, float u0, float u1,
                     BsdfSample& out) {  // :169-198
    out.subsurface = false;
    out.passThrough = false;
    if (!m.thin) {
        out.delta = m.thin || glassSharp(m.roughness);
        return glassSampleWithLobe(m, wo, sp, uc, u0, u1, out.wi, out.f, out.pdf);
    }
    out.delta = true;
    const V3 n = sp.normal;
    const float cosO = dot(wo, n);
    if (cosO == 0.0f) return false;
    const V3 R = m.thinWallReflectance(cosO);
    const float pr = clampf((R.x + R.y + R.z) / 3.0f, 1e-4f, 1.0f - 1e-4f);
    const float c = std::abs(cosO);
    if (uc < pr) {
        out.wi = normalized(n * (2.0f * cosO) - wo);
        out.pdf = pr;
        out.f = m.tint * R / c;
    } else {
        out.wi = -wo;
        out.pdf = 1.0f - pr;
        out.f = m.tint * (V3(1.0f) - R) / c;
        out.passThrough = true;
    }
    return true;
}

}  // namespace

float sampleWavelength(float u, V3& weight) {  // spectral.cpp:75-87
    const SpectralTables& t = spectralTables();
    const float* it = std::upper_bound(t.cdf, t.cdf + kSpectralBins + 1, u);
    int i = static_cast<int>(it - t.cdf) - 1;
    i = std::clamp(i, 0, kSpectralBins - 1);
    const float lo = t.cdf[i], hi = t.cdf[i + 1];
    const float f = hi > lo ? clampf((u - lo) / (hi - lo), 0.0f, 1.0f) : 0.5f;
    const float bw = (kLambdaMax - kLambdaMin) / static_cast<float>(kSpectralBins);
    const float lambda = kLambdaMin + (static_cast<float>(i) + f) * bw;
    weight = (rawRgb(lambda) * t.norm) / ((kLambdaMax - kLambdaMin) * t.pdf[i]);
    return lambda;
}

float cauchyIor(float nd, float abbe, float lambdaNm) {  // spectral.cpp:89-97
    if (abbe <= 0.0f) return nd;
    constexpr float lD = 0.58756f, lF = 0.48613f, lC = 0.65627f;
    const float B = (nd - 1.0f) / (abbe * (1.0f / (lF * lF) - 1.0f / (lC * lC)));
    const float A = nd - B / (lD * lD);
    const float l = lambdaNm * 1e-3f;
    return A + B / (l * l);
}

bool materialIsVolumeBoundary(const pe_material_desc& m) {
    if (m.type == PE_MAT_DIELECTRIC) return m.thin_walled == 0;
    if (m.type == PE_MAT_DISNEY) return clampf(m.specular_transmission, 0.0f, 1.0f) > 0.0f;
    return false;
}

// Material::opacity: DisneyMaterial's setOpacity value, clamped to [0, 1] (disney.h:147-158); 1 for the others.
float materialOpacity(const pe_material_desc& m) {
    return m.type == PE_MAT_DISNEY && m.has_opacity ? std::clamp(m.opacity, 0.0f, 1.0f) : 1.0f;
}

// transmitsShadows: a thin-walled Dielectric (dielectric.h:48) or a Disney material with opacity below 1
// (disney.h:158).
bool materialTransmitsShadows(const pe_material_desc& m) {
    return (m.type == PE_MAT_DIELECTRIC && m.thin_walled != 0) || materialOpacity(m) < 1.0f;
}

bool materialIsDispersive(const pe_material_desc& m) { return m.type == PE_MAT_DIELECTRIC && m.abbe > 0.0f; }

bool materialInteriorMedium(const pe_material_desc& m, V3& sigmaA, V3& sigmaS, float& g) {
    // interiorMedium (dielectric.h:69-71) of rebuildMedium (dielectric.cpp:136-144): sigma_a from the absorption
    // colour at its distance (setAbsorption, :123-127), plus (1 - colour) * density of the scattering
    // (setScattering, :129-134), sigma_s = colour * density; none for thin walls or an empty medium.
    if (m.type != PE_MAT_DIELECTRIC || m.thin_walled != 0) return false;
    sigmaA = V3(0.0f);
    sigmaS = V3(0.0f);
    g = 0.0f;
    if (m.absorption_distance > 0.0f) {
        const V3 color = clamp01(c3(m.absorption_color));
        const float dist = std::max(m.absorption_distance, 1e-5f);
        for (int c = 0; c < 3; ++c) sigmaA[c] = -std::log(std::max(color[c], 1e-4f)) / dist;
    }
    if (m.scatter_density > 0.0f) {
        const float dens = m.scatter_density;
        const V3 sc = clamp01(c3(m.scatter_color));
        sigmaS = sc * dens;
        sigmaA = sigmaA + (V3(1.0f) - sc) * dens;
        g = std::clamp(m.scatter_anisotropy, -0.95f, 0.95f);
    }
    return !(isBlack(sigmaA) && isBlack(sigmaS));  // Medium::empty (material.h:25)
}

bool materialInteriorAbsorption(const pe_material_desc& m, V3& sigmaA) {
    V3 sigmaS;
    float g;
    return materialInteriorMedium(m, sigmaA, sigmaS, g) && !isBlack(sigmaA);
}

V3 materialShadowTransmittance(const pe_material_desc& m, const ShadingPoint& sp, V3 dir) {
    if (!materialTransmitsShadows(m)) return V3(0.0f);
    if (m.type == PE_MAT_DISNEY) return V3(1.0f - materialOpacity(m));  // disney.cpp:511-513
    const Glass g(m);
    const V3 n = lengthSq(sp.ng) > 0.0f ? normalized(sp.ng) : sp.normal;
    return (V3(1.0f) - g.thinWallReflectance(dot(dir, n))) * g.tint;
}

bool subsurfaceCoeffs(const pe_material_desc& m, const ShadingPoint& sp, V3& sigmaT, V3& albedo, float& g) {
    if (m.type != PE_MAT_DISNEY) return false;
    const Params p = resolve(m, sp);
    if (p.subsurface <= 0.0f) return false;
    g = clampf(m.subsurface_anisotropy, -0.9f, 0.9f);
    const V3 radius = V3(clampf(m.subsurface_radius[0], 1e-4f, 1e4f), clampf(m.subsurface_radius[1], 1e-4f, 1e4f),
                         clampf(m.subsurface_radius[2], 1e-4f, 1e4f));
    const float scale = std::max(m.subsurface_scale, 1e-6f);
    for (int c = 0; c < 3; ++c) {
        const float A = clampf(p.baseColor[c], 0.0f, 0.999f);
        const float d = std::max(radius[c] * scale, 1e-7f);
        const float alpha = 1.0f - std::exp(A * (-5.09406f + A * (2.61188f - 4.31805f * A)));
        const float sVal = 1.9f - A + 3.5f * (A - 0.8f) * (A - 0.8f);
        float st = 1.0f / std::max(d * sVal, 1e-12f);
        float ss = alpha * st;
        const float sa = st - ss;
        ss /= (1.0f - g);
        st = ss + sa;
        sigmaT[c] = st;
        albedo[c] = st > 0.0f ? ss / st : 0.0f;
    }
    return true;
}

void flakeState(const pe_material_desc& m, const ShadingPoint& sp, bool& inFlake, V3& flakeNormal) {
    const Params p = resolve(m, sp);
    inFlake = p.inFlake;
    flakeNormal = p.flakeNormal;
}

bool materialIsDelta(const pe_material_desc& m) {
    if (m.type == PE_MAT_DIELECTRIC) return m.thin_walled != 0 || m.roughness <= 1e-3f;  // dielectric.h:37
    return m.type == PE_MAT_MIRROR;  // mirror.h:17
}

V3 materialEmitted(const pe_material_desc& m) { return m.type == PE_MAT_MIRROR ? V3(0.0f) : c3(m.emission); }

V3 evalPdf(const pe_material_desc& m, V3 wo, V3 wi, const ShadingPoint& sp, float& pdf) {
    if (m.type == PE_MAT_LAMBERT) {  // lambertian.cpp:38-60
        const Frame f(sp.normal);
        const float ci = f.toLocal(wi).z, co = f.toLocal(wo).z;
        if (ci <= 0.0f || co <= 0.0f) {
            pdf = 0.0f;
            return V3(0.0f);
        }
        pdf = cosineHemispherePdf(ci);
        return c3(m.base_color) * INV_PI;
    }
    if (m.type == PE_MAT_MIRROR) {  // mirror.cpp:36-44 (evalPdf = eval + pdf, material.h)
        pdf = 0.0f;
        return V3(0.0f);
    }
    if (m.type == PE_MAT_DIELECTRIC) {  // dielectric.cpp:262-292
        const Glass g(m);
        if (glassSharp(g.roughness)) {
            pdf = 0.0f;
            return V3(0.0f);
        }
        const GlassLocal gl = glassFrame(wo, sp);
        const float ior = g.iorAt(sp.wavelength);
        const V3 wiL = gl.frame.toLocal(wi);
        pdf = pdfRoughLocal(gl.wo, wiL, ior, glassAlpha(g.roughness));
        return evalRoughLocal(gl.wo, wiL, ior, glassAlpha(g.roughness), g.tint);
    }
    if (m.type != PE_MAT_DISNEY) throw std::runtime_error("oracle: material type not restated");
    const Params p = resolve(m, sp);  // disney.cpp:664-674
    pdf = pdfLobes(wo, wi, p);
    return evalLobes(wo, wi, p);
}

LobeSplit evalSplit(const pe_material_desc& m, V3 wo, V3 wi, const ShadingPoint& sp) {
    LobeSplit s;
    if (m.type == PE_MAT_DISNEY) {  // disney.cpp:654-658
        evalLobes(wo, wi, resolve(m, sp), &s);
        return s;
    }
    // Material::evalSplit (material.h:154-161) for Lambert and Mirror, Dielectric::evalSplit
    // (dielectric.cpp:279-286): the value on the geometric normal's side of wo is diffuse (glossy for glass),
    // across it transmission.
    float pdf = 0.0f;
    const V3 f = evalPdf(m, wo, wi, sp, pdf);
    const V3 n = lengthSq(sp.ng) > 0.0f ? sp.ng : sp.normal;
    if (dot(wo, n) * dot(wi, n) >= 0.0f) {
        if (m.type == PE_MAT_DIELECTRIC) s.glossy = f;
        else s.diffuse = f;
    } else {
        s.transmission = f;
    }
    return s;
}

bool sampleFull(const pe_material_desc& m, V3 wo, const ShadingPoint& sp, float uc, float u0, float u1, BsdfSample& out) {
    if (m.type == PE_MAT_LAMBERT) {  // material.h:91-96 -> lambertian.cpp:8-35 (uc unused)
        out.delta = false;
        out.subsurface = false;
        out.passThrough = false;
        const Frame f(sp.normal);
        const V3 woL = f.toLocal(wo);
        if (woL.z <= 0.0f) return false;
        const V3 wiL = cosineSampleHemisphere(u0, u1);
        out.wi = normalized(f.toWorld(wiL));
        if (wiL.z <= 0.0f) return false;
        out.f = c3(m.base_color) * INV_PI;
        out.pdf = cosineHemispherePdf(wiL.z);
        return true;
    }
    if (m.type == PE_MAT_MIRROR) {  // material.h:91-96 -> mirror.cpp:7-33 (uc and u unused)
        out.delta = true;
        out.subsurface = false;
        out.passThrough = false;
        const Frame f(sp.normal);
        const V3 woL = f.toLocal(wo);
        if (woL.z <= 0.0f) return false;
        const V3 wiL(-woL.x, -woL.y, woL.z);
        out.wi = normalized(f.toWorld(wiL));
        out.pdf = 1.0f;
        out.f = c3(m.base_color) / wiL.z;
        return true;
    }
    if (m.type == PE_MAT_DIELECTRIC) return glassSampleFull(Glass(m), wo, sp, uc, u0, u1, out);
    if (m.type != PE_MAT_DISNEY) throw std::runtime_error("oracle: material type not restated");
    return disneySampleFull(resolve(m, sp), wo, uc, u0, u1, out);
}

float thinFilmReflectanceC(float cos1, float n1, float n2, float n3re, float n3im, float thicknessNm, float lambdaNm) {
    return thinFilmReflectance(cos1, n1, n2, cfloat(n3re, n3im), thicknessNm, lambdaNm);
}

}  // namespace oracle

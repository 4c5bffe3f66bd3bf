// pe_microfacet.h — GGX (Trowbridge–Reitz) and Charlie-sheen microfacet terms, host and device.
//
// Follows materials/microfacet.h:15-132: anisotropic D with the cancellation-free denominator, the
// height-correlated Smith G with Lambda from Heitz 2014, Heitz 2018 visible-normal sampling, and the energy
// table lookups of microfacet.cpp:20-40 / :142-161 over EnergyTables built on the host (tables.cpp).
#pragma once

#include "pe_device_scene.h"
#include "pe_math.h"

namespace pe {

// The piecewise-Gaussian CIE fit (Wyman, Sloan, Shirley 2013) mapped to linear sRGB and clipped at 0
// (core/color/spectral.cpp:14-30). The host builds the wavelength tables from it; the device evaluates it for
// a sampled wavelength.
PE_HD float cie_lobe(float x, float mu, float s1, float s2) {
    const float t = (x - mu) / (x < mu ? s1 : s2);
    return expf(-0.5f * t * t);
}
PE_HD V3 spectral_rgb_raw(float l) {
    const float X = 1.056f * cie_lobe(l, 599.8f, 37.9f, 31.0f) + 0.362f * cie_lobe(l, 442.0f, 16.0f, 26.7f) -
                    0.065f * cie_lobe(l, 501.1f, 20.4f, 26.2f);
    const float Y = 0.821f * cie_lobe(l, 568.8f, 46.9f, 40.5f) + 0.286f * cie_lobe(l, 530.9f, 16.3f, 31.1f);
    const float Z = 1.217f * cie_lobe(l, 437.0f, 11.8f, 36.0f) + 0.681f * cie_lobe(l, 459.0f, 26.0f, 13.8f);
    const float r = 3.2404542f * X - 1.5371385f * Y - 0.4985314f * Z;
    const float g = -0.9692660f * X + 1.8760108f * Y + 0.0415560f * Z;
    const float b = 0.0556434f * X - 0.2040259f * Y + 1.0572252f * Z;
    return mk(fmaxf(0.0f, r), fmaxf(0.0f, g), fmaxf(0.0f, b));
}
constexpr float kLambdaMin = 380.0f, kLambdaMax = 730.0f;
constexpr int kSpectralBins = 350;

constexpr float kGgxMinAlpha = 1e-4f;
PE_HD float ggx_clamp_alpha(float a) { return fmaxf(a, kGgxMinAlpha); }

PE_HD float ggx_d(V3 h, float ax, float ay) {
    if (h.z <= 0.0f) return 0.0f;
    ax = ggx_clamp_alpha(ax);
    ay = ggx_clamp_alpha(ay);
    const float e = (h.x * h.x) / (ax * ax) + (h.y * h.y) / (ay * ay) + h.z * h.z;
    return 1.0f / (kPi * ax * ay * e * e);
}

PE_HD float ggx_lambda(V3 w, float ax, float ay) {
    const float z2 = w.z * w.z;
    if (z2 <= 1e-12f) return 1e12f;
    ax = ggx_clamp_alpha(ax);
    ay = ggx_clamp_alpha(ay);
    const float t2 = (ax * ax * w.x * w.x + ay * ay * w.y * w.y) / z2;
    return 0.5f * (-1.0f + sqrtf(1.0f + t2));
}

PE_HD float ggx_g1(V3 w, float ax, float ay) { return 1.0f / (1.0f + ggx_lambda(w, ax, ay)); }
PE_HD float ggx_g(V3 wo, V3 wi, float ax, float ay) {
    return 1.0f / (1.0f + ggx_lambda(wo, ax, ay) + ggx_lambda(wi, ax, ay));
}

PE_HD float ggx_visible_pdf(V3 wo, V3 h, float ax, float ay) {
    const float cos_o = fabsf(wo.z);
    if (cos_o <= 0.0f) return 0.0f;
    return ggx_g1(wo, ax, ay) * fabsf(dot(wo, h)) * ggx_d(h, ax, ay) / cos_o;
}

PE_HD V3 ggx_sample_visible_normal(V3 wo, float ax, float ay, float u1, float u2) {
    ax = ggx_clamp_alpha(ax);
    ay = ggx_clamp_alpha(ay);
    if (wo.z < 0.0f) wo = -wo;
    const V3 vh = normalized(mk(ax * wo.x, ay * wo.y, wo.z));
    const float len_sq = vh.x * vh.x + vh.y * vh.y;
    const V3 t1 = len_sq > 0.0f ? mk(-vh.y, vh.x, 0.0f) / sqrtf(len_sq) : mk(1.0f, 0.0f, 0.0f);
    const V3 t2 = cross(vh, t1);
    const float r = sqrtf(u1);
    const float phi = kTwoPi * u2;
    const float p1 = r * cosf(phi);
    float p2 = r * sinf(phi);
    const float s = 0.5f * (1.0f + vh.z);
    p2 = (1.0f - s) * sqrtf(fmaxf(0.0f, 1.0f - p1 * p1)) + s * p2;
    const V3 nh = t1 * p1 + t2 * p2 + vh * sqrtf(fmaxf(0.0f, 1.0f - p1 * p1 - p2 * p2));
    return normalized(mk(ax * nh.x, ay * nh.y, fmaxf(1e-6f, nh.z)));
}

PE_HD float charlie_d(float cos_h, float roughness) {
    const float a = fmaxf(roughness * roughness, 1e-4f);
    const float inv = 1.0f / a;
    const float sin2 = fmaxf(0.0f, 1.0f - cos_h * cos_h);
    return (2.0f + inv) * powf(sin2, 0.5f * inv) * (0.5f * kInvPi);
}

PE_HD float charlie_v(float cos_i, float cos_o) { return 1.0f / fmaxf(4.0f * (cos_i + cos_o - cos_i * cos_o), 1e-6f); }

// Bilinear table reads (microfacet.cpp:20-40): nodes at i/31 on both axes, clamped.
PE_HD float table_lookup2d(const float* t, float mu, float alpha) {
    const float x = clampf(mu, 0.0f, 1.0f) * (kTableN - 1);
    const float y = clampf(alpha, 0.0f, 1.0f) * (kTableN - 1);
    int x0 = static_cast<int>(floorf(x));
    int y0 = static_cast<int>(floorf(y));
    x0 = x0 < 0 ? 0 : (x0 > kTableN - 1 ? kTableN - 1 : x0);
    y0 = y0 < 0 ? 0 : (y0 > kTableN - 1 ? kTableN - 1 : y0);
    const int x1 = x0 + 1 < kTableN ? x0 + 1 : kTableN - 1;
    const int y1 = y0 + 1 < kTableN ? y0 + 1 : kTableN - 1;
    const float fx = clampf(x - static_cast<float>(x0), 0.0f, 1.0f);
    const float fy = clampf(y - static_cast<float>(y0), 0.0f, 1.0f);
    const float a = t[y0 * kTableN + x0] * (1.0f - fx) + t[y0 * kTableN + x1] * fx;
    const float b = t[y1 * kTableN + x0] * (1.0f - fx) + t[y1 * kTableN + x1] * fx;
    return a * (1.0f - fy) + b * fy;
}

PE_HD float table_lookup1d(const float* t, float alpha) {
    const float y = clampf(alpha, 0.0f, 1.0f) * (kTableN - 1);
    int y0 = static_cast<int>(floorf(y));
    y0 = y0 < 0 ? 0 : (y0 > kTableN - 1 ? kTableN - 1 : y0);
    const int y1 = y0 + 1 < kTableN ? y0 + 1 : kTableN - 1;
    const float fy = clampf(y - static_cast<float>(y0), 0.0f, 1.0f);
    return t[y0] * (1.0f - fy) + t[y1] * fy;
}

PE_HD float ggx_directional_albedo(const EnergyTables& T, float cos_theta, float alpha) {
    return clampf(table_lookup2d(T.e, cos_theta, alpha), 1e-3f, 1.0f);
}

PE_HD float ggx_schlick_albedo(const EnergyTables& T, float cos_theta, float alpha, float f0) {
    const float E = clampf(table_lookup2d(T.e, cos_theta, alpha), 1e-3f, 1.0f);
    const float ss = f0 * table_lookup2d(T.ea, cos_theta, alpha) + table_lookup2d(T.eb, cos_theta, alpha);
    return clampf(ss * (1.0f + f0 * (1.0f / E - 1.0f)), 0.0f, 1.0f);
}

PE_HD float ggx_schlick_average_albedo(const EnergyTables& T, float alpha, float f0) {
    const float v = table_lookup1d(T.avg0, alpha) + f0 * table_lookup1d(T.avg1, alpha) + f0 * f0 * table_lookup1d(T.avg2, alpha);
    return clampf(v, 0.0f, 1.0f);
}

PE_HD float charlie_sheen_albedo(const EnergyTables& T, float cos_theta, float roughness) {
    return fmaxf(table_lookup2d(T.sheen, cos_theta, roughness), 0.0f);
}

// Host-side construction of the tables (tables.cpp).
void build_energy_tables(EnergyTables& t);

}  // namespace pe

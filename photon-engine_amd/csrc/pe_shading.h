// pe_shading.h — surface interaction, ray spawning and the Disney / Lambert BSDFs on gfx950.
//
// Per-hit surface data follows Triangle::fillHit (geometry/triangle.cpp:54-133); spawning follows
// geometry/ray_spawn.h:25-87; the BSDF is the reference's Principled/OpenPBR layering (materials/disney.cpp)
// evaluated from a GpuMaterial whose material-only terms the host already computed (scene_prep.cpp).
// The Dielectric (smooth / rough / thin-walled, with Cauchy dispersion) follows materials/dielectric.cpp.
// Not on this path (rejected at scene upload): texture maps.
#pragma once

#include "pe_device_scene.h"
#include "pe_math.h"
#include "pe_microfacet.h"
#include "pe_sobol.h"

namespace pe {

struct Surface {
    V3 point, normal, ng, tangent, p_error, p_smooth;
    int32_t material, emitter;
    uint32_t flags;
    bool front_face, has_smooth;
};

__device__ __forceinline__ V3 ld3(const float4 v) { return mk(v.x, v.y, v.z); }

// Triangle::fillHit for triangle `tri` hit by direction d at barycentrics (u, v).
__device__ __forceinline__ void fill_hit(const SceneView& sc, int32_t tri, V3 d, float u, float v, Surface& si) {
    const TriShade r = sc.shade[tri];
    const V3 v0 = ld3(sc.pos[r.v0]), v1 = ld3(sc.pos[r.v1]), v2 = ld3(sc.pos[r.v2]);
    const V3 edge1 = v1 - v0, edge2 = v2 - v0;
    const float w = 1.0f - u - v;
    const V3 a0 = v0 * w, a1 = v1 * u, a2 = v2 * v;
    si.point = a0 + a1 + a2;
    const float iso = float_gamma(32) * fmaxf(fmaxf(max_comp(vabs(v0)), max_comp(vabs(v1))), fmaxf(max_comp(vabs(v2)), 1e-20f));
    si.p_error = (vabs(a0) + vabs(a1) + vabs(a2)) * float_gamma(16) + splat(iso);
    const V3 gn = normalized(cross(edge1, edge2));
    V3 sn = gn;
    si.has_smooth = false;
    if (r.flags & kTriHasNormals) {
        const V3 n0 = ld3(sc.nrm[r.v0]), n1 = ld3(sc.nrm[r.v1]), n2 = ld3(sc.nrm[r.v2]);
        if (length_sq(n0) > 0.0f && length_sq(n1) > 0.0f && length_sq(n2) > 0.0f) {
            sn = normalized(n0 * w + n1 * u + n2 * v);
            // Hanika 2021 shadow-terminator point (triangle.cpp:86-101)
            const V3 m0 = normalized(n0), m1 = normalized(n1), m2 = normalized(n2);
            const V3 p = si.point;
            const V3 lift = m0 * (w * fmaxf(0.0f, dot(v0 - p, m0))) + m1 * (u * fmaxf(0.0f, dot(v1 - p, m1))) +
                            m2 * (v * fmaxf(0.0f, dot(v2 - p, m2)));
            if (length_sq(lift) > 0.0f) {
                si.p_smooth = p + lift;
                si.has_smooth = true;
            }
        }
    }
    si.ng = gn;
    si.front_face = dot(d, sn) < 0.0f;  // setFaceNormal (surface_interaction.h:48-52)
    si.normal = si.front_face ? sn : -sn;
    if (length_sq(si.ng) == 0.0f) si.ng = sn;
    // Tangent: direction of increasing u, or the Duff frame's s when the uv map is degenerate.
    float du1x = 0.0f, du1y = 0.0f, du2x = 0.0f, du2y = 0.0f;
    if (r.flags & kTriHasUv) {
        const float2 t0 = sc.uv[r.v0], t1 = sc.uv[r.v1], t2 = sc.uv[r.v2];
        du1x = t1.x - t0.x;
        du1y = t1.y - t0.y;
        du2x = t2.x - t0.x;
        du2y = t2.y - t0.y;
    }
    const float rr = du1x * du2y - du2x * du1y;
    V3 tangent;
    if (fabsf(rr) < 1e-6f) {
        const V3 n = si.normal;
        const float sign = copysignf(1.0f, n.z);
        const float a = -1.0f / (sign + n.z);
        const float b = n.x * n.y * a;
        tangent = mk(1.0f + sign * n.x * n.x * a, sign * b, -sign * n.x);
    } else {
        const float inv_r = 1.0f / rr;
        tangent = (edge1 * du2y - edge2 * du1y) * inv_r;
    }
    si.tangent = normalized(tangent);
    si.material = r.material;
    si.emitter = r.emitter;
    si.flags = r.flags;
}

// ---------------------------------------------------------------------------------------- ray_spawn.h
__device__ __forceinline__ V3 offset_ray_origin(V3 p, V3 p_error, V3 n, V3 w) {
    const float d = dot(vabs(n), p_error);
    V3 offset = n * d;
    if (dot(w, n) < 0.0f) offset = -offset;
    V3 po = p + offset;
    po.x = offset.x > 0.0f ? next_float_up(po.x) : (offset.x < 0.0f ? next_float_down(po.x) : po.x);
    po.y = offset.y > 0.0f ? next_float_up(po.y) : (offset.y < 0.0f ? next_float_down(po.y) : po.y);
    po.z = offset.z > 0.0f ? next_float_up(po.z) : (offset.z < 0.0f ? next_float_down(po.z) : po.z);
    return po;
}

__device__ __forceinline__ V3 spawn_origin(const Surface& si, V3 dir) {
    const V3 n = length_sq(si.ng) > 0.0f ? si.ng : si.normal;
    return offset_ray_origin(si.point, si.p_error, n, dir);
}

__device__ __forceinline__ V3 shadow_origin(const Surface& si, V3 to_light) {
    const V3 n = length_sq(si.ng) > 0.0f ? si.ng : si.normal;
    V3 p = si.point;
    if (si.has_smooth) {
        const V3 shading_out = si.front_face ? si.normal : -si.normal;
        const V3 ng_out = dot(n, shading_out) >= 0.0f ? n : -n;
        const float nl = dot(shading_out, to_light);
        if (nl > 0.0f) {
            const float k_cutoff = 0.1f;
            const float ngl = dot(ng_out, to_light);
            const float amount = nl < k_cutoff ? clampf(2.0f - (ngl + nl) / k_cutoff, 0.0f, 1.0f)
                                               : clampf(1.0f - ngl / k_cutoff, 0.0f, 1.0f);
            if (amount > 0.0f) p += (si.p_smooth - si.point) * amount;
        }
    }
    return offset_ray_origin(p, si.p_error, n, to_light);
}

// spawnRayTo (ray_spawn.h:78-87): shadow segment that stops just short of the target.
__device__ __forceinline__ bool spawn_ray_to(const Surface& si, V3 target, V3& o, V3& d, float& tmax) {
    const V3 to_target = target - si.point;
    const float len = length(to_target);
    if (!(len > 0.0f)) return false;
    o = shadow_origin(si, to_target / len);
    const V3 dd = target - o;
    const float dist = length(dd);
    const float slack = fmaxf(dist * 1e-4f, 2e-6f * max_comp(vabs(target)));
    if (!(dist > slack)) return false;
    d = dd / dist;
    tmax = dist - slack;
    return true;
}

// ------------------------------------------------------------------------------------------ sampling.h
// A point uniform on the unit disk by the concentric map (uniformSampleDisk, sampling.h:55-75).
__device__ __forceinline__ void concentric_disk(float u0, float u1, float& x, float& y) {
    const float sx = 2.0f * u0 - 1.0f, sy = 2.0f * u1 - 1.0f;
    x = 0.0f;
    y = 0.0f;
    if (!(sx == 0.0f && sy == 0.0f)) {
        float r, theta;
        if (fabsf(sx) > fabsf(sy)) {
            r = sx;
            theta = kPiOver4 * (sy / sx);
        } else {
            r = sy;
            theta = kPiOver2 - kPiOver4 * (sx / sy);
        }
        x = r * cosf(theta);
        y = r * sinf(theta);
    }
}

__device__ __forceinline__ V3 cosine_sample_hemisphere(float u0, float u1) {
    float x, y;
    concentric_disk(u0, u1, x, y);
    return mk(x, y, sqrtf(fmaxf(0.0f, 1.0f - x * x - y * y)));
}

__device__ __forceinline__ V3 reflect_vec(V3 wo, V3 n) { return 2.0f * dot(wo, n) * n - wo; }

// ------------------------------------------------------------------------------------------ Disney BSDF
__device__ __forceinline__ float schlick_f(float f0, float c) {
    const float m = 1.0f - clampf(c, 0.0f, 1.0f);
    const float m2 = m * m;
    return f0 + (1.0f - f0) * m2 * m2 * m;
}

// fresnelF82Tint (fresnel.h:31-47)
__device__ __forceinline__ V3 f82_tint(V3 f0, V3 tint, float cos_theta) {
    const float mu_bar = 1.0f / 7.0f;
    const float om = 1.0f - mu_bar;
    const float denom = mu_bar * om * om * om * om * om * om;
    const float mu = clampf(cos_theta, 0.0f, 1.0f);
    const float m = 1.0f - mu;
    const float m6 = m * m * m * m * m * m;
    V3 out;
    out.x = clampf(schlick_f(f0.x, mu) - (schlick_f(f0.x, mu_bar) * (1.0f - tint.x) / denom) * mu * m6, 0.0f, 1.0f);
    out.y = clampf(schlick_f(f0.y, mu) - (schlick_f(f0.y, mu_bar) * (1.0f - tint.y) / denom) * mu * m6, 0.0f, 1.0f);
    out.z = clampf(schlick_f(f0.z, mu) - (schlick_f(f0.z, mu_bar) * (1.0f - tint.z) / denom) * mu * m6, 0.0f, 1.0f);
    return out;
}

// shadingFrame (disney.cpp:62-73)
__device__ __forceinline__ Frame shading_frame(V3 n_in, V3 tangent_in) {
    const V3 n = length_sq(n_in) > 0.0f ? normalized(n_in) : mk(0.0f, 1.0f, 0.0f);
    V3 t = tangent_in - n * dot(tangent_in, n);
    if (length_sq(t) < 1e-8f) return frame_from_normal(n);
    t = normalized(t);
    V3 b = cross(n, t);
    if (length_sq(b) < 1e-8f) return frame_from_normal(n);
    b = normalized(b);
    t = normalized(cross(b, n));
    Frame f;
    f.s = t;
    f.t = b;
    f.n = n;
    return f;
}

__device__ __forceinline__ float layer_transmit(const GpuMaterial& m, const EnergyTables& T, float mu) {
    return fmaxf(0.0f, 1.0f - m.wd * ggx_schlick_albedo(T, mu, m.alpha, m.f0d));
}
__device__ __forceinline__ float coat_transmit(const GpuMaterial& m, const EnergyTables& T, float mu) {
    if (m.clearcoat <= 0.0f) return 1.0f;
    return 1.0f - m.clearcoat * ggx_schlick_albedo(T, mu, m.alpha_c, m.f0c);
}
__device__ __forceinline__ float sheen_scale(const GpuMaterial& m, const EnergyTables& T, float mu_o, float mu_i) {
    if (m.sheen_max <= 0.0f) return 1.0f;
    const float eo = fminf(1.0f, charlie_sheen_albedo(T, mu_o, m.sheen_roughness));
    const float ei = fminf(1.0f, charlie_sheen_albedo(T, mu_i, m.sheen_roughness));
    return fmaxf(0.0f, fminf(1.0f - m.sheen_max * eo, 1.0f - m.sheen_max * ei));
}

// EON diffuse (eon.h:27-51)
__device__ __forceinline__ float fon_albedo(float mu, float r) {
    const float c1 = 0.5f - 2.0f / (3.0f * kPi);
    mu = clampf(mu, 1e-4f, 1.0f);
    const float af = 1.0f / (1.0f + c1 * r);
    const float bf = r * af;
    const float si = sqrtf(fmaxf(0.0f, 1.0f - mu * mu));
    const float g = si * (acosf(mu) - si * mu) + (2.0f / 3.0f) * ((si / mu) * (1.0f - si * si * si) - si);
    return af + (bf * kInvPi) * g;
}
__device__ __forceinline__ V3 eon_eval(V3 rho, float r, V3 wi, V3 wo) {
    if (r <= 0.0f) return rho * kInvPi;
    const float c1 = 0.5f - 2.0f / (3.0f * kPi);
    const float c2 = 2.0f / 3.0f - 28.0f / (15.0f * kPi);
    const float mu_i = wi.z, mu_o = wo.z;
    const float s = dot(wi, wo) - mu_i * mu_o;
    const float s_over_t = s > 0.0f ? s / fmaxf(mu_i, mu_o) : s;
    const float af = 1.0f / (1.0f + c1 * r);
    const V3 fss = rho * (kInvPi * af * (1.0f + r * s_over_t));
    const float efo = fon_albedo(mu_o, r), efi = fon_albedo(mu_i, r);
    const float avg = af * (1.0f + c2 * r);
    V3 rho_ms;
    rho_ms.x = rho.x * rho.x * avg / fmaxf(1e-6f, 1.0f - rho.x * (1.0f - avg));
    rho_ms.y = rho.y * rho.y * avg / fmaxf(1e-6f, 1.0f - rho.y * (1.0f - avg));
    rho_ms.z = rho.z * rho.z * avg / fmaxf(1e-6f, 1.0f - rho.z * (1.0f - avg));
    const float eps = 1e-7f;
    const float ms = fmaxf(eps, 1.0f - efo) * fmaxf(eps, 1.0f - efi) / fmaxf(eps, 1.0f - avg);
    return fss + rho_ms * (kInvPi * ms);
}

// ------------------------------------------------------------------------------------- Fresnel, Snell
// fresnelDielectric (core/math/utils.h:206-233): unpolarised reflectance, indices swapped from the inside.
__device__ __forceinline__ float fresnel_dielectric(float cos_i, float eta_i, float eta_t) {
    cos_i = clampf(cos_i, -1.0f, 1.0f);
    if (!(cos_i > 0.0f)) {
        const float tmp = eta_i;
        eta_i = eta_t;
        eta_t = tmp;
        cos_i = fabsf(cos_i);
    }
    const float sin_i = sqrtf(fmaxf(0.0f, 1.0f - cos_i * cos_i));
    const float sin_t = (eta_i / eta_t) * sin_i;
    if (sin_t >= 1.0f) return 1.0f;
    const float cos_t = sqrtf(fmaxf(0.0f, 1.0f - sin_t * sin_t));
    const float rs = (eta_t * cos_i - eta_i * cos_t) / (eta_t * cos_i + eta_i * cos_t);
    const float rp = (eta_i * cos_i - eta_t * cos_t) / (eta_i * cos_i + eta_t * cos_t);
    return (rs * rs + rp * rp) * 0.5f;
}

// refractVec (utils.h:163-174): wi points away from the surface on n's side, eta = eta_i / eta_t.
__device__ __forceinline__ bool refract_vec(V3 wi, V3 n, float eta, V3& wt) {
    const float cos_i = dot(n, wi);
    const float sin2_i = fmaxf(0.0f, 1.0f - cos_i * cos_i);
    const float sin2_t = eta * eta * sin2_i;
    if (sin2_t >= 1.0f) return false;
    const float cos_t = sqrtf(1.0f - sin2_t);
    wt = eta * (-wi) + (eta * cos_i - cos_t) * n;
    return true;
}

// -------------------------------------------------------------------------------- Disney transmission
// evalSpecTransmission / pdfSpecTransmission (disney.cpp:185-220): rough refraction through the base layer.
__device__ __forceinline__ bool spec_trans_half(const GpuMaterial& m, bool front, V3 wo, V3 wi, V3& wm, float& etap,
                                                float& wo_h, float& wi_h) {
    if (wo.z <= 0.0f || wi.z >= 0.0f) return false;
    etap = front ? m.ior : (1.0f / m.ior);
    wm = wi * etap + wo;
    if (length_sq(wm) <= 1e-12f) return false;
    wm = normalized(wm);
    if (wm.z < 0.0f) wm = -wm;
    wo_h = dot(wo, wm);
    wi_h = dot(wi, wm);
    return wo_h > 0.0f && wi_h < 0.0f;
}

__device__ __forceinline__ V3 eval_spec_trans(const GpuMaterial& m, bool front, V3 wo, V3 wi) {
    if (m.st_w <= 0.0f) return splat(0.0f);
    V3 wm;
    float etap, wo_h, wi_h;
    if (!spec_trans_half(m, front, wo, wi, wm, etap, wo_h, wi_h)) return splat(0.0f);
    const float alpha = ggx_clamp_alpha(m.roughness * m.roughness);
    const float F = fresnel_dielectric(front ? wo_h : -wo_h, 1.0f, m.ior);
    const float D = ggx_d(wm, alpha, alpha);
    const float G = ggx_g(wo, mk(wi.x, wi.y, -wi.z), alpha, alpha);
    const float s = wi_h + wo_h / etap;
    const float denom = s * s * fabsf(wi.z * wo.z);
    if (denom <= 1e-12f) return splat(0.0f);
    const float ft = D * (1.0f - F) * G * fabsf(wi_h * wo_h) / (denom * etap * etap);
    return m.base_color * (m.st_w * ft);
}

__device__ __forceinline__ float pdf_spec_trans(const GpuMaterial& m, bool front, V3 wo, V3 wi) {
    V3 wm;
    float etap, wo_h, wi_h;
    if (!spec_trans_half(m, front, wo, wi, wm, etap, wo_h, wi_h)) return 0.0f;
    const float alpha = ggx_clamp_alpha(m.roughness * m.roughness);
    const float visible = ggx_visible_pdf(wo, wm, alpha, alpha);
    const float s = wi_h + wo_h / etap;
    if (s * s <= 1e-12f) return 0.0f;
    return visible * fabsf(wi_h) / (s * s);
}

// ----------------------------------------------------------------- thin-film interference (fresnel.h:74-107)
struct Cf {
    float re, im;
};
__device__ __forceinline__ Cf cf(float re, float im = 0.0f) { return Cf{re, im}; }
__device__ __forceinline__ Cf operator+(Cf a, Cf b) { return Cf{a.re + b.re, a.im + b.im}; }
__device__ __forceinline__ Cf operator-(Cf a, Cf b) { return Cf{a.re - b.re, a.im - b.im}; }
__device__ __forceinline__ Cf operator*(Cf a, Cf b) { return Cf{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
__device__ __forceinline__ Cf operator*(float s, Cf a) { return Cf{s * a.re, s * a.im}; }
__device__ __forceinline__ Cf operator/(Cf a, Cf b) {
    const float d = b.re * b.re + b.im * b.im;
    return Cf{(a.re * b.re + a.im * b.im) / d, (a.im * b.re - a.re * b.im) / d};
}
__device__ __forceinline__ float cnorm(Cf a) { return a.re * a.re + a.im * a.im; }
__device__ __forceinline__ Cf csqrt(Cf z) {  // principal square root
    const float r = sqrtf(cnorm(z));
    const float re = sqrtf(fmaxf(0.0f, 0.5f * (r + z.re)));
    const float im = sqrtf(fmaxf(0.0f, 0.5f * (r - z.re)));
    return Cf{re, z.im < 0.0f ? -im : im};
}

// Unpolarised reflectance of air | film (n2, thickness nm) | base n3 at one wavelength: the Airy sum
// r = (r12 + r23 e^{i delta}) / (1 + r12 r23 e^{i delta}), delta = 4 pi n2 d cos(theta2) / lambda.
__device__ __forceinline__ float thin_film_reflectance(float cos1, float n2, Cf n3, float thickness, float lambda) {
    const float n1 = 1.0f;
    cos1 = clampf(cos1, 0.0f, 1.0f);
    const float sin1sq = 1.0f - cos1 * cos1;
    const float sin2sq = (n1 / n2) * (n1 / n2) * sin1sq;
    if (sin2sq >= 1.0f) return 1.0f;
    const float cos2 = sqrtf(1.0f - sin2sq);
    const Cf cos3 = csqrt(cf(1.0f) - cf(n1 * n1 * sin1sq) / (n3 * n3));
    const float r12s = (n1 * cos1 - n2 * cos2) / (n1 * cos1 + n2 * cos2);
    const float r12p = (n2 * cos1 - n1 * cos2) / (n2 * cos1 + n1 * cos2);
    const Cf r23s = (cf(n2 * cos2) - n3 * cos3) / (cf(n2 * cos2) + n3 * cos3);
    const Cf r23p = (cos2 * n3 - n2 * cos3) / (cos2 * n3 + n2 * cos3);
    const float delta = 4.0f * kPi * n2 * thickness * cos2 / lambda;
    float sd, cd;
    sincosf(delta, &sd, &cd);
    const Cf e = cf(cd, sd);
    const Cf rs = (cf(r12s) + r23s * e) / (cf(1.0f) + r12s * (r23s * e));
    const Cf rp = (cf(r12p) + r23p * e) / (cf(1.0f) + r12p * (r23p * e));
    return clampf(0.5f * (cnorm(rs) + cnorm(rp)), 0.0f, 1.0f);
}

// Each channel averages three wavelengths (fresnel.h:94-107).
__device__ __forceinline__ V3 thin_film_rgb(float cos1, float film_ior, const Cf base[3], float thickness) {
    const float lambda[3][3] = {{610.0f, 640.0f, 670.0f}, {515.0f, 540.0f, 565.0f}, {440.0f, 465.0f, 490.0f}};
    float out[3];
    for (int c = 0; c < 3; ++c) {
        float s = 0.0f;
        for (int k = 0; k < 3; ++k) s += thin_film_reflectance(cos1, film_ior, base[c], thickness, lambda[c][k]);
        out[c] = s / 3.0f;
    }
    return mk(out[0], out[1], out[2]);
}

// specularFresnel with a film (disney.cpp:85-93).
__device__ __forceinline__ V3 film_specular_fresnel(const GpuMaterial& m, float cos_d) {
    const Cf diel[3] = {cf(m.film_nd), cf(m.film_nd), cf(m.film_nd)};
    const V3 fd = thin_film_rgb(cos_d, m.film_ior, diel, m.film_thickness) * m.wd;
    if (m.metallic <= 0.0f) return fd;
    const Cf metal[3] = {cf(m.film_metal_n.x, m.film_metal_k.x), cf(m.film_metal_n.y, m.film_metal_k.y),
                         cf(m.film_metal_n.z, m.film_metal_k.z)};
    return lerp3(fd, thin_film_rgb(cos_d, m.film_ior, metal, m.film_thickness), m.metallic);
}

// Material::evalSplit's value (material.h:51-57): a BSDF value split into the lobe groups of the light passes.
struct LobeSplit {
    V3 diffuse, glossy, transmission;
};

// ------------------------------------------------------------------------------------ metallic flakes
// A shading point on a flake (DisneyMaterial::resolve, disney.cpp:600-601): the flake's normal (its frame is
// rebuilt where it is used, so that only three floats stay live across the vertex).
constexpr float kFlakeAlpha = 0.06f;  // FLAKE_ALPHA (disney.cpp:75)
struct Flake {
    V3 n;
};

// lowbias32 and the cell hash of the flake lattice (disney.cpp:407-421).
__device__ __forceinline__ uint32_t flake_hash(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}
__device__ __forceinline__ uint32_t flake_cell_hash(int x, int y, int z) {
    return flake_hash(static_cast<uint32_t>(x) * 0x8da6b343u ^
                      flake_hash(static_cast<uint32_t>(y) * 0xd8163841u ^ flake_hash(static_cast<uint32_t>(z) * 0xcb1ab31fu)));
}
__device__ __forceinline__ float flake_unit(uint32_t h) { return static_cast<float>(h >> 8) * 0x1p-24f; }

// resolveFlake (disney.cpp:430-460): the nearest seed of the 27 cells around the point (one seed per cell of
// the flake size) decides whether the point is on a flake and how the flake is tilted about the shading normal.
__device__ __forceinline__ bool resolve_flake(const GpuMaterial& m, V3 point, V3 normal, Flake& fl) {
#pragma clang fp contract(off)
    const V3 q = point / m.flake_size;
    const int cx = static_cast<int>(floorf(q.x)), cy = static_cast<int>(floorf(q.y)), cz = static_cast<int>(floorf(q.z));
    float best = 1e30f;
    uint32_t best_hash = 0;
    for (int dz = -1; dz <= 1; ++dz)
        for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx) {
                const uint32_t h = flake_cell_hash(cx + dx, cy + dy, cz + dz);
                const V3 f = mk(static_cast<float>(cx + dx) + flake_unit(h),
                                static_cast<float>(cy + dy) + flake_unit(flake_hash(h ^ 0x68e31da4u)),
                                static_cast<float>(cz + dz) + flake_unit(flake_hash(h ^ 0xb5297a4du)));
                const V3 e = q - f;
                const float d2 = e.x * e.x + e.y * e.y + e.z * e.z;
                if (d2 < best) {
                    best = d2;
                    best_hash = h;
                }
            }
    if (flake_unit(flake_hash(best_hash ^ 0x1b56c4e9u)) >= m.flake_density) return false;
    const float u1 = fminf(flake_unit(flake_hash(best_hash ^ 0x2c1b3c6du)), 0.999f);
    const float u2 = flake_unit(flake_hash(best_hash ^ 0x297a2d39u));
    const float r = m.flake_roughness;
    const float tan2 = r * r * u1 / (1.0f - u1);
    const float cos_t = 1.0f / sqrtf(1.0f + tan2);
    const float sin_t = sqrtf(fmaxf(0.0f, 1.0f - cos_t * cos_t));
    const float phi = kTwoPi * u2;
    fl.n = normalized(to_world(frame_from_normal(normal), mk(sin_t * cosf(phi), sin_t * sinf(phi), cos_t)));
    return true;
}

// evalFlake / pdfFlake (disney.cpp:104-126): a narrow GGX lobe about the flake normal, Schlick reflectance of the
// flake colour per channel; world-space directions.
__device__ __forceinline__ V3 flake_eval(const GpuMaterial& m, const Flake& fl, V3 wo, V3 wi) {
    const Frame ff = frame_from_normal(fl.n);
    const V3 o = to_local(ff, wo), i = to_local(ff, wi);
    if (o.z <= 0.0f || i.z <= 0.0f) return splat(0.0f);
    const V3 h = normalized(o + i);
    const float D = ggx_d(h, kFlakeAlpha, kFlakeAlpha);
    const float G = ggx_g(o, i, kFlakeAlpha, kFlakeAlpha);
    const float c = dot(o, h);
    const V3 col = m.flake_color;
    return mk(schlick_f(col.x, c), schlick_f(col.y, c), schlick_f(col.z, c)) * (D * G / (4.0f * o.z * i.z));
}
__device__ __forceinline__ float flake_pdf(const Flake& fl, V3 wo, V3 wi) {
    const Frame ff = frame_from_normal(fl.n);
    const V3 o = to_local(ff, wo), i = to_local(ff, wi);
    if (o.z <= 0.0f || i.z <= 0.0f) return 0.0f;
    const V3 h = normalized(o + i);
    const float c = dot(o, h);
    if (c <= 0.0f) return 0.0f;
    return ggx_visible_pdf(o, h, kFlakeAlpha, kFlakeAlpha) / (4.0f * c);
}

// evalLobes (disney.cpp:223-311) in the local frame: reflection above the surface; diffuse and specular
// transmission below it. kTransport = false is the variant for scenes whose materials all have zero
// transmission weights, where the lobes below the surface are exactly zero. With `split` (the light passes),
// the value's diffuse, glossy (specular, sheen, coat) and transmission parts as well (disney.cpp:250, :290-309).
template <bool kTransport>
__device__ __forceinline__ V3 disney_eval_local(const GpuMaterial& m, const EnergyTables& T, bool front, V3 wo, V3 wi,
                                                LobeSplit* split = nullptr, const V3* flake_base = nullptr) {
    if (wo.z <= 0.0f) return splat(0.0f);
    const float cos_o = wo.z;
    if (kTransport && flake_base) {
        // On a flake (disney.cpp:234-237, :260-261): no transmission, the flake lobe in place of the base layer
        // (no diffuse, no sheen), under the clear coat.
        if (wi.z <= 0.0f) {
            if (split) split->transmission = splat(0.0f);
            return splat(0.0f);
        }
        const V3 base = *flake_base;
        if (m.clearcoat <= 0.0f) {
            if (split) {
                split->diffuse = splat(0.0f);
                split->glossy = base;
            }
            return base;
        }
        const V3 h = normalized(wo + wi);
        const float cos_d = dot(wo, h), cos_i = wi.z;
        const float dc = ggx_d(h, m.alpha_c, m.alpha_c);
        const float gc = ggx_g(wo, wi, m.alpha_c, m.alpha_c);
        const float fc = m.f0c + (1.0f - m.f0c) * pow5(1.0f - clampf(cos_d, 0.0f, 1.0f));
        const float ms_c = 1.0f + m.f0c * (1.0f / ggx_directional_albedo(T, cos_o, m.alpha_c) - 1.0f);
        const V3 f_coat = splat(m.clearcoat * fc * dc * gc / (4.0f * cos_o * cos_i) * ms_c);
        const V3 coat_att = m.coat_tint_factor * (coat_transmit(m, T, cos_o) * coat_transmit(m, T, cos_i));
        if (split) {
            split->diffuse = splat(0.0f);
            split->glossy = base * coat_att + f_coat;
        }
        return base * coat_att + f_coat;
    }
    if (wi.z <= 0.0f) {
        V3 ft = splat(0.0f);
        if (!kTransport) return ft;
        if (m.trans_w <= 0.0f && m.st_w <= 0.0f) return ft;
        const float coat_o = coat_transmit(m, T, cos_o);
        if (m.trans_w > 0.0f) {
            const float mu_i = -wi.z;
            const float factor = layer_transmit(m, T, cos_o) * layer_transmit(m, T, mu_i) * m.inv_one_minus_avg;
            const float sh = sheen_scale(m, T, cos_o, mu_i);
            ft += m.base_color * (m.trans_w * kInvPi * coat_o * factor * sh) * m.coat_tint_factor;
        }
        if (m.st_w > 0.0f) ft += eval_spec_trans(m, front, wo, wi) * coat_o * m.coat_tint_factor;
        if (split) split->transmission = ft;
        return ft;
    }
    const float cos_i = wi.z;
    const V3 h = normalized(wo + wi);
    const float cos_d = dot(wo, h);
    V3 f_diffuse = splat(0.0f);
    if (m.diff_w > 0.0f) {
        const float factor = layer_transmit(m, T, cos_o) * layer_transmit(m, T, cos_i) * m.inv_one_minus_avg;
        f_diffuse = eon_eval(m.base_color, m.diffuse_roughness, wi, wo) * (m.diff_w * factor);
    }
    const float D = ggx_d(h, m.ax, m.ay);
    const float G = ggx_g(wo, wi, m.ax, m.ay);
    // specularFresnel (disney.cpp:79-96); films only in the transport variant (a film marks the scene so)
    V3 F;
    if (kTransport && (m.flags & kMatFlagFilm)) {
        F = film_specular_fresnel(m, cos_d);
    } else {
        F = splat(m.wd * schlick_f(m.f0d, cos_d));
        if (m.metallic > 0.0f) F = lerp3(F, f82_tint(m.base_color, m.edge_tint, cos_d), m.metallic);
    }
    if (kTransport && m.st_w > 0.0f && !front) F = lerp3(F, splat(fresnel_dielectric(-cos_d, 1.0f, m.ior)), m.st_w);
    // multiScatterFactor (disney.cpp:99-103)
    const float E = ggx_directional_albedo(T, cos_o, sqrtf(m.ax * m.ay));
    const V3 f0 = lerp3(splat(m.f0d), m.base_color, m.metallic);
    const V3 ms = splat(1.0f) + f0 * (1.0f / E - 1.0f);
    const V3 f_spec = (F * D * G) / (4.0f * cos_o * cos_i) * ms;
    V3 base = f_diffuse + f_spec;
    V3 base_diffuse = f_diffuse;
    if (m.flags & kMatFlagSheen) {
        const float s = sheen_scale(m, T, cos_o, cos_i);
        const float norm = fmaxf(1.0f, fmaxf(charlie_sheen_albedo(T, cos_o, m.sheen_roughness),
                                             charlie_sheen_albedo(T, cos_i, m.sheen_roughness)));
        const V3 sheen = m.sheen_color * (charlie_d(h.z, m.sheen_roughness) * charlie_v(cos_i, cos_o) / norm);
        base = base * s + sheen;
        base_diffuse = base_diffuse * s;
    }
    if (m.clearcoat <= 0.0f) {
        if (split) {
            split->diffuse = base_diffuse;
            split->glossy = base - base_diffuse;
        }
        return base;
    }
    const float dc = ggx_d(h, m.alpha_c, m.alpha_c);
    const float gc = ggx_g(wo, wi, m.alpha_c, m.alpha_c);
    const float fc = m.f0c + (1.0f - m.f0c) * pow5(1.0f - clampf(cos_d, 0.0f, 1.0f));
    const float ms_c = 1.0f + m.f0c * (1.0f / ggx_directional_albedo(T, cos_o, m.alpha_c) - 1.0f);
    const V3 f_coat = splat(m.clearcoat * fc * dc * gc / (4.0f * cos_o * cos_i) * ms_c);
    const V3 coat_att = m.coat_tint_factor * (coat_transmit(m, T, cos_o) * coat_transmit(m, T, cos_i));
    if (split) {
        split->diffuse = base_diffuse * coat_att;
        split->glossy = (base - base_diffuse) * coat_att + f_coat;
    }
    return base * coat_att + f_coat;
}

// pdfLobes (disney.cpp:340-378).
template <bool kTransport>
__device__ __forceinline__ float disney_pdf_local(const GpuMaterial& m, bool front, V3 wo, V3 wi,
                                                  const float* flake_pdf_v = nullptr) {
    if (wo.z <= 0.0f) return 0.0f;
    if (kTransport && flake_pdf_v) {  // on a flake: coat and the flake lobe only (lobeWeights, disney.cpp:315-322)
        if (wi.z <= 0.0f) return 0.0f;
        const V3 h = normalized(wo + wi);
        const float cos_d = dot(wo, h);
        if (cos_d <= 0.0f) return 0.0f;
        float pdf_coat = 0.0f;
        if (m.p_coat > 0.0f) pdf_coat = ggx_visible_pdf(wo, h, m.alpha_c, m.alpha_c) / (4.0f * cos_d);
        return (1.0f - m.p_coat) * *flake_pdf_v + m.p_coat * pdf_coat;
    }
    if (wi.z <= 0.0f) {
        float pdf = 0.0f;
        if (!kTransport) return pdf;
        if (m.p_trans > 0.0f) pdf += m.p_trans * (-wi.z * kInvPi);
        if (m.p_spec_trans > 0.0f) pdf += m.p_spec_trans * pdf_spec_trans(m, front, wo, wi);
        return pdf;
    }
    const V3 h = normalized(wo + wi);
    const float cos_d = dot(wo, h);
    if (cos_d <= 0.0f) return 0.0f;
    const float pdf_diffuse = wi.z * kInvPi;
    const float pdf_spec = ggx_visible_pdf(wo, h, m.ax, m.ay) / (4.0f * cos_d);
    float pdf_coat = 0.0f;
    if (m.p_coat > 0.0f) pdf_coat = ggx_visible_pdf(wo, h, m.alpha_c, m.alpha_c) / (4.0f * cos_d);
    return m.p_diff_refl * pdf_diffuse + m.p_spec * pdf_spec + m.p_coat * pdf_coat;
}

// ----------------------------------------------------------------------------------------- Dielectric
__device__ __forceinline__ float cauchy_ior(float nd, float abbe, float lambda_nm) {  // spectral.cpp:89-97
    if (abbe <= 0.0f) return nd;
    const float lD = 0.58756f, lF = 0.48613f, lC = 0.65627f;
    const float B = (nd - 1.0f) / (abbe * (1.0f / (lF * lF) - 1.0f / (lC * lC)));
    const float A = nd - B / (lD * lD);
    const float l = lambda_nm * 1e-3f;
    return A + B / (l * l);
}

__device__ __forceinline__ float glass_ior(const GpuMaterial& m, float lambda) {  // iorAt (dielectric.cpp:174-177)
    if (m.abbe <= 0.0f || lambda <= 0.0f) return m.ior;
    return fmaxf(1.0001f, cauchy_ior(m.ior, m.abbe, lambda));
}

// glassFrame (dielectric.cpp:36-45): shading frame on the geometric normal's side.
__device__ __forceinline__ Frame glass_frame(const Surface& si, V3 wo_world) {
    const V3 outward = si.front_face ? si.normal : -si.normal;
    const V3 ng = length_sq(si.ng) > 1e-12f ? normalized(si.ng) : normalized(outward);
    V3 ns = length_sq(si.normal) > 1e-12f ? normalized(si.normal) : ng;
    if (dot(ns, ng) < 0.0f) ns = -ns;
    if (dot(wo_world, ns) * dot(wo_world, ng) <= 0.0f) ns = ng;
    return frame_from_normal(ns);
}

// generalizedHalf (dielectric.cpp:55-67).
__device__ __forceinline__ bool generalized_half(V3 wo, V3 wi, float ior, V3& wm, float& etap) {
    const float cos_o = wo.z, cos_i = wi.z;
    if (cos_o == 0.0f || cos_i == 0.0f) return false;
    const bool refl = cos_o * cos_i > 0.0f;
    etap = refl ? 1.0f : (cos_o > 0.0f ? ior : 1.0f / ior);
    wm = wi * etap + wo;
    if (length_sq(wm) <= 1e-12f) return false;
    wm = normalized(wm);
    if (wm.z < 0.0f) wm = -wm;
    return !(dot(wm, wi) * cos_i <= 0.0f || dot(wm, wo) * cos_o <= 0.0f);
}

__device__ __forceinline__ V3 glass_eval_rough(V3 wo, V3 wi, float ior, float alpha, V3 tint) {  // :79-97
    V3 wm;
    float etap = 1.0f;
    if (!generalized_half(wo, wi, ior, wm, etap)) return splat(0.0f);
    const float cos_o = wo.z, cos_i = wi.z;
    const float wo_h = dot(wo, wm), wi_h = dot(wi, wm);
    const float F = fresnel_dielectric(wo_h, 1.0f, ior);
    const float D = ggx_d(wm, alpha, alpha);
    const float G = ggx_g(wo, wi, alpha, alpha);
    if (cos_o * cos_i > 0.0f) return tint * (D * G * F / fabsf(4.0f * cos_i * cos_o));
    const float s = wi_h + wo_h / etap;
    const float denom = s * s * cos_i * cos_o;
    if (fabsf(denom) <= 1e-12f) return splat(0.0f);
    const float ft = D * (1.0f - F) * G * fabsf(wi_h * wo_h / denom);
    return tint * (ft / (etap * etap));
}

__device__ __forceinline__ float glass_pdf_rough(V3 wo, V3 wi, float ior, float alpha) {  // :105-119
    V3 wm;
    float etap = 1.0f;
    if (!generalized_half(wo, wi, ior, wm, etap)) return 0.0f;
    const float wo_h = dot(wo, wm), wi_h = dot(wi, wm);
    const float R = fresnel_dielectric(wo_h, 1.0f, ior);
    const float visible = ggx_visible_pdf(wo, wm, alpha, alpha);
    if (wo.z * wi.z > 0.0f) return visible / (4.0f * fabsf(wo_h)) * R;
    const float s = wi_h + wo_h / etap;
    if (s * s <= 1e-12f) return 0.0f;
    return visible * fabsf(wi_h) / (s * s) * (1.0f - R);
}

// The thin wall's reflectance: a two-sided slab with incoherent inner reflections, 2R / (1 + R)
// (thinWallReflectance, dielectric.cpp:153-161, no interference film).
__device__ __forceinline__ V3 thin_wall_reflectance(const GpuMaterial& m, float cos_theta) {
    cos_theta = clampf(fabsf(cos_theta), 0.0f, 1.0f);
    if (m.flags & kMatFlagFilm) {  // a soap-bubble film: air | film | air
        const Cf air[3] = {cf(1.0f), cf(1.0f), cf(1.0f)};
        return thin_film_rgb(cos_theta, m.film_ior, air, m.film_thickness);
    }
    const float R = fresnel_dielectric(cos_theta, 1.0f, m.ior);
    return splat(2.0f * R / (1.0f + R));
}
// Dielectric::shadowTransmittance (dielectric.cpp:163-167) of a thin wall crossed in direction dir.
__device__ __forceinline__ V3 thin_wall_transmittance(const GpuMaterial& m, const Surface& si, V3 dir) {
    const V3 n = length_sq(si.ng) > 0.0f ? normalized(si.ng) : si.normal;
    return (splat(1.0f) - thin_wall_reflectance(m, dot(dir, n))) * m.base_color;
}

// ------------------------------------------------------------------------------------ material dispatch
// Material::evalPdf for the surface's material; world-space directions; lambda = the path's wavelength (0: RGB).
// kTransport = false: the scene has no Dielectric and no transmitting or subsurface Disney material.
template <bool kTransport>
__device__ __forceinline__ V3 bsdf_eval_pdf(const GpuMaterial& m, const EnergyTables& T, const Surface& si, V3 wo,
                                            V3 wi, float lambda, float& pdf, const Flake* fl = nullptr) {
    if (m.type == kMatLambert) {  // lambertian.cpp:38-60
        const Frame f = frame_from_normal(si.normal);
        const float ci = to_local(f, wi).z, co = to_local(f, wo).z;
        if (ci <= 0.0f || co <= 0.0f) {
            pdf = 0.0f;
            return splat(0.0f);
        }
        pdf = ci * kInvPi;
        return m.base_color * kInvPi;
    }
    if (m.type == kMatMirror) {  // mirror.cpp:36-44: a Dirac delta has no finite value
        pdf = 0.0f;
        return splat(0.0f);
    }
    if (kTransport && m.type == kMatDielectric) {  // dielectric.cpp:262-292
        if (m.flags & kMatFlagDelta) {  // smooth or thin: no finite BSDF value
            pdf = 0.0f;
            return splat(0.0f);
        }
        const Frame f = glass_frame(si, wo);
        const V3 wol = to_local(f, wo), wil = to_local(f, wi);
        const float ior = glass_ior(m, lambda);
        pdf = glass_pdf_rough(wol, wil, ior, m.glass_alpha);
        return glass_eval_rough(wol, wil, ior, m.glass_alpha, m.base_color);
    }
    const Frame f = shading_frame(si.normal, si.tangent);
    const V3 wol = to_local(f, wo), wil = to_local(f, wi);
    if (kTransport && fl) {
        const float fp = flake_pdf(*fl, wo, wi);
        const V3 fb = flake_eval(m, *fl, wo, wi);
        pdf = disney_pdf_local<kTransport>(m, si.front_face, wol, wil, &fp);
        return disney_eval_local<kTransport>(m, T, si.front_face, wol, wil, nullptr, &fb);
    }
    pdf = disney_pdf_local<kTransport>(m, si.front_face, wol, wil);
    return disney_eval_local<kTransport>(m, T, si.front_face, wol, wil);
}

// Material::evalSplit (material.h:154-161; DisneyMaterial::evalSplit, disney.cpp:654-658; Dielectric::evalSplit,
// dielectric.cpp:279-286): Disney's lobes as evalLobes splits them; for the other materials the BSDF value on the
// geometric normal's side of wo is diffuse (glossy for glass), across it transmission.
template <bool kTransport>
__device__ __forceinline__ LobeSplit bsdf_eval_split(const GpuMaterial& m, const EnergyTables& T, const Surface& si,
                                                     V3 wo, V3 wi, float lambda, const Flake* fl = nullptr) {
    LobeSplit sp{splat(0.0f), splat(0.0f), splat(0.0f)};
    if (m.type == kMatDisney) {
        const Frame f = shading_frame(si.normal, si.tangent);
        V3 fb;
        if (kTransport && fl) fb = flake_eval(m, *fl, wo, wi);
        (void)disney_eval_local<kTransport>(m, T, si.front_face, to_local(f, wo), to_local(f, wi), &sp,
                                            kTransport && fl ? &fb : nullptr);
        return sp;
    }
    float pdf = 0.0f;
    const V3 v = bsdf_eval_pdf<kTransport>(m, T, si, wo, wi, lambda, pdf);
    const V3 n = length_sq(si.ng) > 0.0f ? si.ng : si.normal;
    if (dot(wo, n) * dot(wi, n) >= 0.0f) {
        if (m.type == kMatDielectric) sp.glossy = v;
        else sp.diffuse = v;
    } else {
        sp.transmission = v;
    }
    return sp;
}

struct BsdfSample {
    V3 wi, f;
    float pdf;
    bool delta, pass_through, subsurface;
    V3 weight;  // subsurface: the throughput weight of entering the walk
};

// Dielectric::sampleWithLobe (dielectric.cpp:186-253).
__device__ __forceinline__ bool glass_sample(const GpuMaterial& m, const Surface& si, V3 wo_world, float uc, float u0,
                                             float u1, float lambda, BsdfSample& out) {
    const Frame fr = glass_frame(si, wo_world);
    const V3 wo = to_local(fr, wo_world);
    if (wo.z == 0.0f) return false;
    const bool entering = wo.z > 0.0f;
    const float ior = glass_ior(m, lambda);
    const float eta = entering ? (1.0f / ior) : ior;
    if (m.roughness <= 1e-3f) {
        const float F = fresnel_dielectric(wo.z, 1.0f, ior);
        V3 wil = mk(-wo.x, -wo.y, wo.z);
        V3 weight;
        if (uc < F) {
            out.pdf = F;
            weight = m.base_color * F;
        } else {
            const V3 n = mk(0.0f, 0.0f, entering ? 1.0f : -1.0f);
            if (!refract_vec(wo, n, eta, wil)) {
                wil = mk(-wo.x, -wo.y, wo.z);
                out.pdf = 1.0f;
                weight = m.base_color;
            } else {
                out.pdf = 1.0f - F;
                weight = m.base_color * ((1.0f - F) * eta * eta);
            }
        }
        out.wi = normalized(to_world(fr, wil));
        out.f = weight / fmaxf(fabsf(dot(out.wi, si.normal)), 1e-6f);
        return true;
    }
    const V3 wm = ggx_sample_visible_normal(wo, m.glass_alpha, m.glass_alpha, u0, u1);
    const float R = fresnel_dielectric(dot(wo, wm), 1.0f, ior);
    V3 wil;
    if (uc < R) {
        wil = reflect_vec(wo, wm);
        if (wil.z * wo.z <= 0.0f) return false;
    } else {
        const V3 n = entering ? wm : -wm;
        if (!refract_vec(wo, n, eta, wil)) return false;
        if (wil.z * wo.z >= 0.0f) return false;
    }
    out.f = glass_eval_rough(wo, wil, ior, m.glass_alpha, m.base_color);
    out.pdf = glass_pdf_rough(wo, wil, ior, m.glass_alpha);
    out.wi = normalized(to_world(fr, wil));
    return out.pdf > 0.0f && !is_black(out.f);
}

// Material::sampleFull (disney.cpp:680-729; lambertian.cpp:8-35; dielectric.cpp:169-198). With kTransport = false
// the specular-transmission and subsurface lobes have zero probability and are left out.
template <bool kTransport>
__device__ __forceinline__ bool bsdf_sample(const GpuMaterial& m, const EnergyTables& T, const Surface& si, V3 wo,
                                            float uc, float u0, float u1, float lambda, BsdfSample& out,
                                            const Flake* fl = nullptr) {
    out.delta = false;
    out.pass_through = false;
    out.subsurface = false;
    if (m.type == kMatLambert) {
        const Frame f = frame_from_normal(si.normal);
        if (to_local(f, wo).z <= 0.0f) return false;
        const V3 wil = cosine_sample_hemisphere(u0, u1);
        out.wi = normalized(to_world(f, wil));
        if (wil.z <= 0.0f) return false;
        out.f = m.base_color * kInvPi;
        out.pdf = wil.z * kInvPi;
        return true;
    }
    if (m.type == kMatMirror) {  // mirror.cpp:7-33: wi = the reflection of wo, f = R / cos, pdf = 1
        out.delta = true;
        const Frame f = frame_from_normal(si.normal);
        const V3 wol = to_local(f, wo);
        if (wol.z <= 0.0f) return false;
        const V3 wil = mk(-wol.x, -wol.y, wol.z);
        out.wi = normalized(to_world(f, wil));
        out.pdf = 1.0f;
        out.f = m.base_color / wil.z;
        return true;
    }
    if (kTransport && m.type == kMatDielectric) {
        if (!(m.flags & kMatFlagThinWall)) {
            out.delta = (m.flags & kMatFlagDelta) != 0;
            return glass_sample(m, si, wo, uc, u0, u1, lambda, out);
        }
        // Thin wall: reflect with the mean slab reflectance, else pass straight through (dielectric.cpp:176-197).
        out.delta = true;
        const V3 n = si.normal;
        const float cos_o = dot(wo, n);
        if (cos_o == 0.0f) return false;
        const V3 R = thin_wall_reflectance(m, cos_o);
        const float pr = clampf((R.x + R.y + R.z) / 3.0f, 1e-4f, 1.0f - 1e-4f);
        const float c = fabsf(cos_o);
        if (uc < pr) {
            out.wi = normalized(n * (2.0f * cos_o) - wo);
            out.pdf = pr;
            out.f = m.base_color * R / c;
        } else {
            out.wi = -wo;
            out.pdf = 1.0f - pr;
            out.f = m.base_color * (splat(1.0f) - R) / c;
            out.pass_through = true;
        }
        return true;
    }
    const Frame f = shading_frame(si.normal, si.tangent);
    const V3 wol = to_local(f, wo);
    if (wol.z <= 0.0f) return false;
    V3 wil;
    if (kTransport && fl) {
        // On a flake (lobeWeights disney.cpp:315-322, sampleFull :689-729): the coat, else the flake lobe sampled
        // in the flake's frame (sampleSpecularLobe :380-396); the remaining branch is the (empty) transmission.
        const float p_spec = 1.0f - m.p_coat;
        if (uc < m.p_coat) {
            wil = reflect_vec(wol, ggx_sample_visible_normal(wol, m.alpha_c, m.alpha_c, u0, u1));
            if (wil.z <= 0.0f) return false;
        } else if (uc < m.p_coat + p_spec) {
            const Frame ff = frame_from_normal(fl->n);
            const V3 o = to_local(ff, wo);
            if (o.z <= 0.0f) return false;
            const V3 i = reflect_vec(o, ggx_sample_visible_normal(o, kFlakeAlpha, kFlakeAlpha, u0, u1));
            if (i.z <= 0.0f) return false;
            wil = to_local(f, to_world(ff, i));
            if (!(wil.z > 0.0f)) return false;
        } else {
            wil = cosine_sample_hemisphere(u0, u1);
            wil.z = -wil.z;
        }
        out.wi = normalized(to_world(f, wil));
        const V3 wil2 = to_local(f, out.wi);
        const float fp = flake_pdf(*fl, wo, out.wi);
        const V3 fb = flake_eval(m, *fl, wo, out.wi);
        out.f = disney_eval_local<kTransport>(m, T, si.front_face, wol, wil2, nullptr, &fb);
        out.pdf = disney_pdf_local<kTransport>(m, si.front_face, wol, wil2, &fp);
        return out.pdf > 0.0f;
    }
    if (uc < m.p_coat) {
        wil = reflect_vec(wol, ggx_sample_visible_normal(wol, m.alpha_c, m.alpha_c, u0, u1));
        if (wil.z <= 0.0f) return false;
    } else if (uc < m.p_coat + m.p_spec) {
        wil = reflect_vec(wol, ggx_sample_visible_normal(wol, m.ax, m.ay, u0, u1));
        if (!(wil.z > 0.0f)) return false;
    } else if (kTransport && uc < m.p_coat + m.p_spec + m.p_spec_trans) {  // sampleSpecTransLobe (disney.cpp:398-405)
        const float alpha = ggx_clamp_alpha(m.roughness * m.roughness);
        const V3 wm = ggx_sample_visible_normal(wol, alpha, alpha, u0, u1);
        const float eta = si.front_face ? (1.0f / m.ior) : m.ior;
        if (!refract_vec(wol, wm, eta, wil)) return false;
        if (!(wil.z < 0.0f)) return false;
    } else if (uc < m.p_coat + m.p_spec + m.p_spec_trans + m.p_diff_refl) {
        wil = cosine_sample_hemisphere(u0, u1);
    } else if (kTransport && uc < m.p_coat + m.p_spec + m.p_spec_trans + m.p_diff_refl + m.p_sss) {
        // Subsurface lobe (disney.cpp:704-720): enter the random walk below the surface.
        wil = cosine_sample_hemisphere(u0, u1);
        wil.z = -wil.z;
        const float coat_o = coat_transmit(m, T, wol.z);
        const float sheen_o = sheen_scale(m, T, wol.z, 1.0f);
        const float layer_o = layer_transmit(m, T, wol.z);
        const float w = m.subsurface * m.opaque_dielectric * (1.0f - m.diffuse_transmission) * coat_o * sheen_o *
                        layer_o / fmaxf(m.p_sss, 1e-6f);
        out.wi = normalized(to_world(f, wil));
        out.subsurface = true;
        out.weight = m.coat_tint_factor * w;
        out.pdf = m.p_sss * (-wil.z * kInvPi);
        out.f = out.weight * out.pdf / fmaxf(-wil.z, 1e-6f);
        return out.pdf > 0.0f;
    } else {
        wil = cosine_sample_hemisphere(u0, u1);
        wil.z = -wil.z;
    }
    out.wi = normalized(to_world(f, wil));
    const V3 wil2 = to_local(f, out.wi);
    out.f = disney_eval_local<kTransport>(m, T, si.front_face, wol, wil2);
    out.pdf = disney_pdf_local<kTransport>(m, si.front_face, wol, wil2);
    return out.pdf > 0.0f;
}

// ------------------------------------------------------------------ wavelengths, media, subsurface walks

// sampleWavelength (core/color/spectral.cpp:75-87): inverse CDF of the 350-bin wavelength table, returning the
// wavelength and the RGB weight rgb(lambda) * norm / pdf(lambda) that keeps the estimator unbiased.
__device__ __forceinline__ float sample_wavelength(const EnergyTables& T, float u, V3& weight) {
    int lo = 0, count = kSpectralBins + 1;  // std::upper_bound over spec_cdf
    while (count > 0) {
        const int step = count >> 1;
        const int it = lo + step;
        if (!(u < T.spec_cdf[it])) {
            lo = it + 1;
            count -= step + 1;
        } else {
            count = step;
        }
    }
    int i = lo - 1;
    i = i < 0 ? 0 : (i > kSpectralBins - 1 ? kSpectralBins - 1 : i);
    const float c0 = T.spec_cdf[i], c1 = T.spec_cdf[i + 1];
    const float f = c1 > c0 ? clampf((u - c0) / (c1 - c0), 0.0f, 1.0f) : 0.5f;
    const float bw = (kLambdaMax - kLambdaMin) / static_cast<float>(kSpectralBins);
    const float lambda = kLambdaMin + (static_cast<float>(i) + f) * bw;
    const V3 norm = mk(T.spec_norm[0], T.spec_norm[1], T.spec_norm[2]);
    weight = (spectral_rgb_raw(lambda) * norm) / ((kLambdaMax - kLambdaMin) * T.spec_pdf[i]);
    return lambda;
}

// sampleFreeFlight (integrators/path_tracer.cpp:97-128): distance sampling in a homogeneous medium with one
// colour channel picked in proportion to `channel_w` and the spectral-MIS pdf over all three. True with the
// scatter distance when the flight ends inside the medium (throughput *= sigma_s Tr / pdf); false when it
// reaches the surface at t_max (throughput *= Tr / P) or leaves an open medium.
__device__ __forceinline__ bool free_flight(V3 sigma_t, V3 sigma_s, float t_max, bool hit_surface, V3 channel_w,
                                            SobolState& ss, V3& thr, float& t_scatter) {
    float wsum = channel_w.x + channel_w.y + channel_w.z;
    if (!(wsum > 0.0f)) {
        channel_w = splat(1.0f);
        wsum = 3.0f;
    }
    const V3 pc = channel_w / wsum;
    const float uc = sobol_get1d(ss);
    const int c = uc < pc.x ? 0 : (uc < pc.x + pc.y ? 1 : 2);
    const float u = sobol_get1d(ss);
    const float stc = comp(sigma_t, c);
    const float t = stc > 0.0f ? -logf(fmaxf(1.0f - u, 1e-12f)) / stc : kFloatMax;
    if (t < t_max) {
        const V3 tr = mk(expf(-sigma_t.x * t), expf(-sigma_t.y * t), expf(-sigma_t.z * t));
        const float pdf = pc.x * sigma_t.x * tr.x + pc.y * sigma_t.y * tr.y + pc.z * sigma_t.z * tr.z;
        if (!(pdf > 0.0f)) {
            thr = splat(0.0f);
            return false;
        }
        thr *= sigma_s * tr / pdf;
        t_scatter = t;
        return true;
    }
    V3 tr;
    if (hit_surface) {
        tr = mk(expf(-sigma_t.x * t_max), expf(-sigma_t.y * t_max), expf(-sigma_t.z * t_max));
    } else {
        tr = mk(sigma_t.x > 0.0f ? 0.0f : 1.0f, sigma_t.y > 0.0f ? 0.0f : 1.0f, sigma_t.z > 0.0f ? 0.0f : 1.0f);
    }
    const float P = pc.x * tr.x + pc.y * tr.y + pc.z * tr.z;
    thr = P > 0.0f ? thr * tr / P : splat(0.0f);
    return false;
}

// Henyey-Greenstein direction about `d` (core/sampling/phase.h:27-39).
__device__ __forceinline__ V3 sample_hg(V3 d, float g, float u0, float u1) {
    float cos_t;
    if (fabsf(g) < 1e-3f) {
        cos_t = 1.0f - 2.0f * u0;
    } else {
        const float sq = (1.0f - g * g) / (1.0f - g + 2.0f * g * u0);
        cos_t = (1.0f + g * g - sq * sq) / (2.0f * g);
    }
    cos_t = clampf(cos_t, -1.0f, 1.0f);
    const float sin_t = sqrtf(fmaxf(0.0f, 1.0f - cos_t * cos_t));
    const float phi = kTwoPi * u1;
    const Frame f = frame_from_normal(d);
    return normalized(to_world(f, mk(sin_t * cosf(phi), sin_t * sinf(phi), cos_t)));
}

}  // namespace pe

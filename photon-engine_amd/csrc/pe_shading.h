// pe_shading.h — surface interaction, ray spawning and the Disney / Lambert BSDFs on gfx950.
//
// Per-hit surface data follows Triangle::fillHit (geometry/triangle.cpp:54-133); spawning follows
// geometry/ray_spawn.h:25-87; the BSDF is the reference's Principled/OpenPBR layering (materials/disney.cpp)
// evaluated from a GpuMaterial whose material-only terms the host already computed (scene_prep.cpp).
// Not on this path yet (rejected at scene upload): texture maps, metallic flakes, thin-film layers,
// specular/diffuse-transmission and subsurface lobes, Dielectric.
#pragma once

#include "pe_device_scene.h"
#include "pe_math.h"
#include "pe_microfacet.h"

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
__device__ __forceinline__ V3 cosine_sample_hemisphere(float u0, float u1) {
    const float sx = 2.0f * u0 - 1.0f, sy = 2.0f * u1 - 1.0f;
    float x = 0.0f, y = 0.0f;
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

// evalLobes (disney.cpp:223-311) in the local frame, reflection half only (transmission lobes are rejected at
// upload, so wi below the surface gives 0, as the reference does with both transmission weights at 0).
__device__ __forceinline__ V3 disney_eval_local(const GpuMaterial& m, const EnergyTables& T, V3 wo, V3 wi) {
    if (wo.z <= 0.0f || wi.z <= 0.0f) return splat(0.0f);
    const float cos_o = wo.z, cos_i = wi.z;
    const V3 h = normalized(wo + wi);
    const float cos_d = dot(wo, h);
    V3 f_diffuse = splat(0.0f);
    if (m.diff_w > 0.0f) {
        const float factor = layer_transmit(m, T, cos_o) * layer_transmit(m, T, cos_i) * m.inv_one_minus_avg;
        f_diffuse = eon_eval(m.base_color, m.diffuse_roughness, wi, wo) * (m.diff_w * factor);
    }
    const float D = ggx_d(h, m.ax, m.ay);
    const float G = ggx_g(wo, wi, m.ax, m.ay);
    // specularFresnel (disney.cpp:79-96) without thin film
    V3 F = splat(m.wd * schlick_f(m.f0d, cos_d));
    if (m.metallic > 0.0f) F = lerp3(F, f82_tint(m.base_color, m.edge_tint, cos_d), m.metallic);
    // multiScatterFactor (disney.cpp:99-103)
    const float E = ggx_directional_albedo(T, cos_o, sqrtf(m.ax * m.ay));
    const V3 f0 = lerp3(splat(m.f0d), m.base_color, m.metallic);
    const V3 ms = splat(1.0f) + f0 * (1.0f / E - 1.0f);
    const V3 f_spec = (F * D * G) / (4.0f * cos_o * cos_i) * ms;
    V3 base = f_diffuse + f_spec;
    if (m.flags & kMatFlagSheen) {
        const float s = sheen_scale(m, T, cos_o, cos_i);
        const float norm = fmaxf(1.0f, fmaxf(charlie_sheen_albedo(T, cos_o, m.sheen_roughness),
                                             charlie_sheen_albedo(T, cos_i, m.sheen_roughness)));
        const V3 sheen = m.sheen_color * (charlie_d(h.z, m.sheen_roughness) * charlie_v(cos_i, cos_o) / norm);
        base = base * s + sheen;
    }
    if (m.clearcoat <= 0.0f) return base;
    const float dc = ggx_d(h, m.alpha_c, m.alpha_c);
    const float gc = ggx_g(wo, wi, m.alpha_c, m.alpha_c);
    const float fc = m.f0c + (1.0f - m.f0c) * pow5(1.0f - clampf(cos_d, 0.0f, 1.0f));
    const float ms_c = 1.0f + m.f0c * (1.0f / ggx_directional_albedo(T, cos_o, m.alpha_c) - 1.0f);
    const V3 f_coat = splat(m.clearcoat * fc * dc * gc / (4.0f * cos_o * cos_i) * ms_c);
    const V3 coat_att = m.coat_tint_factor * (coat_transmit(m, T, cos_o) * coat_transmit(m, T, cos_i));
    return base * coat_att + f_coat;
}

// pdfLobes (disney.cpp:340-378), reflection half.
__device__ __forceinline__ float disney_pdf_local(const GpuMaterial& m, V3 wo, V3 wi) {
    if (wo.z <= 0.0f || wi.z <= 0.0f) return 0.0f;
    const V3 h = normalized(wo + wi);
    const float cos_d = dot(wo, h);
    if (cos_d <= 0.0f) return 0.0f;
    const float pdf_diffuse = wi.z * kInvPi;
    const float pdf_spec = ggx_visible_pdf(wo, h, m.ax, m.ay) / (4.0f * cos_d);
    float pdf_coat = 0.0f;
    if (m.p_coat > 0.0f) pdf_coat = ggx_visible_pdf(wo, h, m.alpha_c, m.alpha_c) / (4.0f * cos_d);
    return m.p_diff_refl * pdf_diffuse + m.p_spec * pdf_spec + m.p_coat * pdf_coat;
}

// Material::evalPdf for the surface's material; world-space directions.
__device__ __forceinline__ V3 bsdf_eval_pdf(const GpuMaterial& m, const EnergyTables& T, const Surface& si, V3 wo,
                                            V3 wi, float& pdf) {
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
    const Frame f = shading_frame(si.normal, si.tangent);
    const V3 wol = to_local(f, wo), wil = to_local(f, wi);
    pdf = disney_pdf_local(m, wol, wil);
    return disney_eval_local(m, T, wol, wil);
}

struct BsdfSample {
    V3 wi, f;
    float pdf;
    bool delta;
};

// Material::sampleFull (disney.cpp:680-729; lambertian.cpp:8-35).
__device__ __forceinline__ bool bsdf_sample(const GpuMaterial& m, const EnergyTables& T, const Surface& si, V3 wo,
                                            float uc, float u0, float u1, BsdfSample& out) {
    out.delta = false;
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
    const Frame f = shading_frame(si.normal, si.tangent);
    const V3 wol = to_local(f, wo);
    if (wol.z <= 0.0f) return false;
    V3 wil;
    if (uc < m.p_coat) {
        wil = reflect_vec(wol, ggx_sample_visible_normal(wol, m.alpha_c, m.alpha_c, u0, u1));
        if (wil.z <= 0.0f) return false;
    } else if (uc < m.p_coat + m.p_spec) {
        wil = reflect_vec(wol, ggx_sample_visible_normal(wol, m.ax, m.ay, u0, u1));
        if (!(wil.z > 0.0f)) return false;
    } else if (uc < m.p_coat + m.p_spec + m.p_spec_trans + m.p_diff_refl) {
        // p_spec_trans is 0 on this path (rejected at upload), so this is the diffuse-reflection branch.
        wil = cosine_sample_hemisphere(u0, u1);
    } else {
        // Remaining probability: the diffuse-transmission lobe (weight 0 here), sampled below the surface,
        // where the reflection-only BSDF is 0, so the sample ends the path as in the reference.
        wil = cosine_sample_hemisphere(u0, u1);
        wil.z = -wil.z;
    }
    out.wi = normalized(to_world(f, wil));
    const V3 wil2 = to_local(f, out.wi);
    out.f = disney_eval_local(m, T, wol, wil2);
    out.pdf = disney_pdf_local(m, wol, wil2);
    return out.pdf > 0.0f;
}

}  // namespace pe

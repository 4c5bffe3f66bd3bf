// pe_lights.h — light sampling on gfx950: AreaLight (spherical-rectangle sampling, Ureña 2013),
// TriangleLight (uniform area), EnvironmentLight (MIS-compensated 2D CDF + sun cone) and the
// LightSampler's per-vertex selection distribution (uniform, or exhaustive importance with a 15% floor).
// Reference: lights/area_light.cpp, lights/environment_light.cpp, lights/light_sampler.cpp.
#pragma once

#include "pe_device_scene.h"
#include "pe_math.h"

namespace pe {

struct LightSample {
    V3 wi, li;
    float pdf, distance;
};

__device__ __forceinline__ float receiver_cosine(V3 n, V3 to_light) {  // light.h:95-98
    const float c = dot(n, to_light);
    return c > 0.05f ? c : 0.05f;
}

__device__ __forceinline__ float area_solid_angle_pdf(float dist_sq, float cos_abs, float area) {
    if (!(cos_abs > 1e-6f) || !(area > 0.0f) || !(dist_sq > 0.0f)) return 0.0f;
    return dist_sq / (area * cos_abs);
}

__device__ __forceinline__ float safe_asin(float x) { return asinf(clampf(x, -1.0f, 1.0f)); }
__device__ __forceinline__ float angle_between(V3 a, V3 b) {
    if (dot(a, b) < 0.0f) return kPi - 2.0f * safe_asin(length(a + b) * 0.5f);
    return 2.0f * safe_asin(length(b - a) * 0.5f);
}

struct SphRect {
    Frame frame;
    float x0, y0, z0, x1, y1, b0, b1, g2g3, solid_angle;
};

// setupSphericalRect (area_light.cpp:47-72)
__device__ __forceinline__ SphRect setup_sph_rect(V3 ref, const GpuLight& L) {
    SphRect r;
    const V3 x = L.p1 / L.len_u, y = L.p2 / L.len_v;
    V3 z = cross(x, y);
    const V3 d = L.p0 - ref;
    r.z0 = dot(d, z);
    if (r.z0 > 0.0f) {
        z = -z;
        r.z0 = -r.z0;
    }
    r.frame.s = x;
    r.frame.t = y;
    r.frame.n = z;
    r.x0 = dot(d, x);
    r.y0 = dot(d, y);
    r.x1 = r.x0 + L.len_u;
    r.y1 = r.y0 + L.len_v;
    r.solid_angle = 0.0f;
    r.b0 = r.b1 = r.g2g3 = 0.0f;
    if (r.z0 > -1e-7f) return r;
    const V3 v00 = mk(r.x0, r.y0, r.z0), v01 = mk(r.x0, r.y1, r.z0), v10 = mk(r.x1, r.y0, r.z0), v11 = mk(r.x1, r.y1, r.z0);
    const V3 n0 = normalized(cross(v00, v10)), n1 = normalized(cross(v10, v11));
    const V3 n2 = normalized(cross(v11, v01)), n3 = normalized(cross(v01, v00));
    const float g0 = angle_between(-n0, n1), g1 = angle_between(-n1, n2);
    const float g2 = angle_between(-n2, n3), g3 = angle_between(-n3, n0);
    r.solid_angle = g0 + g1 + g2 + g3 - kTwoPi;
    r.b0 = n0.z;
    r.b1 = n2.z;
    r.g2g3 = g2 + g3;
    return r;
}

// sampleSphericalRect (area_light.cpp:75-92)
__device__ __forceinline__ V3 sample_sph_rect(const SphRect& r, float ux, float uy) {
    const float au = ux * r.solid_angle - r.g2g3;
    const float fu = (cosf(au) * r.b0 - r.b1) / sinf(au);
    float cu = copysignf(1.0f, fu) / sqrtf(fu * fu + r.b0 * r.b0);
    cu = clampf(cu, -0.9999999f, 0.9999999f);
    float xu = -(cu * r.z0) / sqrtf(fmaxf(0.0f, 1.0f - cu * cu));
    xu = clampf(xu, r.x0, r.x1);
    const float dd = sqrtf(xu * xu + r.z0 * r.z0);
    const float h0 = r.y0 / sqrtf(dd * dd + r.y0 * r.y0);
    const float h1 = r.y1 / sqrtf(dd * dd + r.y1 * r.y1);
    const float hv = h0 + uy * (h1 - h0);
    const float hv2 = hv * hv;
    const float yv = hv2 < 1.0f - 1e-6f ? hv * dd / sqrtf(1.0f - hv2) : r.y1;
    return to_world(r.frame, mk(xu, clampf(yv, r.y0, r.y1), r.z0));
}

constexpr float kMinSolidAngle = 1e-3f;

// Light::importance (area_light.cpp:100-112, :202-212)
__device__ __forceinline__ float light_importance(const GpuLight& L, V3 p, V3 n) {
    const V3 d = L.center - p;
    const float dist2 = length_sq(d);
    if (!(dist2 > 0.0f)) return L.lum * kTwoPi;
    const V3 w = d / sqrtf(dist2);
    const float cos_l = fmaxf(fabsf(dot(w, L.normal)), 0.2f);
    const float omega = fminf(L.area * cos_l / dist2, kTwoPi);
    return L.lum * omega * receiver_cosine(n, w);
}

// AreaLight::sampleLi (area_light.cpp:114-157) / TriangleLight::sampleLi (:235-252). pdf 0 = invalid.
__device__ __forceinline__ LightSample light_sample(const GpuLight& L, V3 ref, float ux, float uy) {
    LightSample ls;
    ls.pdf = 0.0f;
    ls.distance = 0.0f;
    ls.li = splat(0.0f);
    ls.wi = splat(0.0f);
    if (L.type == PE_LIGHT_AREA) {
        if (L.rectangle) {
            const SphRect r = setup_sph_rect(ref, L);
            if (r.solid_angle >= kMinSolidAngle) {
                const V3 to_light = sample_sph_rect(r, ux, uy);
                const float dist = length(to_light);
                if (!(dist > 0.0f)) return ls;
                ls.distance = dist;
                ls.wi = to_light / dist;
                ls.li = L.radiance;
                ls.pdf = 1.0f / r.solid_angle;
                return ls;
            }
        }
        const V3 p_light = L.p0 + L.p1 * ux + L.p2 * uy;
        const V3 to_light = p_light - ref;
        const float dist_sq = length_sq(to_light);
        ls.distance = sqrtf(dist_sq);
        if (ls.distance > 0.0f) {
            ls.wi = to_light / ls.distance;
            const float pdf = area_solid_angle_pdf(dist_sq, fabsf(dot(ls.wi, L.normal)), L.area);
            if (pdf > 0.0f) {
                ls.li = L.radiance;
                ls.pdf = pdf;
            }
        }
        return ls;
    }
    if (!(L.area > 0.0f) || is_black(L.radiance)) return ls;
    const float su0 = sqrtf(ux);
    const float b1 = 1.0f - su0, b2 = uy * su0;
    const V3 p_light = L.p0 * (1.0f - b1 - b2) + L.p1 * b1 + L.p2 * b2;
    const V3 to_light = p_light - ref;
    const float dist_sq = length_sq(to_light);
    ls.distance = sqrtf(dist_sq);
    if (!(ls.distance > 0.0f)) return ls;
    ls.wi = to_light / ls.distance;
    const float pdf = area_solid_angle_pdf(dist_sq, fabsf(dot(ls.wi, L.normal)), L.area);
    if (!(pdf > 0.0f)) return ls;
    ls.li = L.radiance;
    ls.pdf = pdf;
    return ls;
}

// pdfLiHit: AreaLight::pdfLi (area_light.cpp:163-187) / TriangleLight::pdfLiHit (:226-231).
__device__ __forceinline__ float light_pdf_hit(const GpuLight& L, V3 ref, V3 hit) {
    if (L.type == PE_LIGHT_AREA) {
        const V3 d = hit - L.p0;
        const float scale = fmaxf(L.len_u + L.len_v, 1.0f);
        if (fabsf(dot(d, L.normal)) > 1e-3f * scale) return 0.0f;
        const float uu = dot(L.p1, L.p1), vv = dot(L.p2, L.p2), uv = dot(L.p1, L.p2);
        const float du = dot(d, L.p1), dv = dot(d, L.p2);
        const float det = uu * vv - uv * uv;
        if (!(fabsf(det) > 1e-12f * uu * vv)) return 0.0f;
        const float s = (vv * du - uv * dv) / det;
        const float t = (uu * dv - uv * du) / det;
        const float eps = 1e-3f;
        if (s < -eps || t < -eps || s > 1.0f + eps || t > 1.0f + eps) return 0.0f;
        if (L.rectangle) {
            const float omega = setup_sph_rect(ref, L).solid_angle;
            if (omega >= kMinSolidAngle) return 1.0f / omega;
        }
    }
    const V3 to_light = hit - ref;
    const float dist_sq = length_sq(to_light);
    if (!(dist_sq > 0.0f)) return 0.0f;
    return area_solid_angle_pdf(dist_sq, fabsf(dot(to_light / sqrtf(dist_sq), L.normal)), L.area);
}

// ------------------------------------------------------------------------------------ environment
__device__ __forceinline__ void to_equirect(const EnvView& e, V3 dir, float& u, float& v) {
    const V3 d = normalized(dir);
    const float theta = acosf(clampf(d.y, -1.0f, 1.0f));
    float phi = atan2f(-d.z, d.x) + kPi + e.rotation;
    phi = fmodf(phi, kTwoPi);
    if (phi < 0.0f) phi += kTwoPi;
    u = phi / kTwoPi;
    v = theta / kPi;
}

__device__ __forceinline__ V3 env_texel(const EnvView& e, int x, int y) {
    const float* p = e.rgb + 3 * (static_cast<size_t>(y) * e.w + x);
    return mk(p[0], p[1], p[2]);
}

__device__ __forceinline__ V3 env_eval_sky(const EnvView& e, V3 dir) {  // environment_light.cpp:301-306, :43-62
    if (e.w <= 0) return splat(0.5f) * e.intensity;
    float u, v;
    to_equirect(e, dir, u, v);
    u -= floorf(u);
    v = clampf(v, 0.0f, 1.0f);
    const float x = u * static_cast<float>(e.w) - 0.5f;
    const float y = v * static_cast<float>(e.h) - 0.5f;
    const int x0 = static_cast<int>(floorf(x)), y0 = static_cast<int>(floorf(y));
    const float fx = x - static_cast<float>(x0), fy = y - static_cast<float>(y0);
    int xa = x0 % e.w;
    if (xa < 0) xa += e.w;
    int xb = (x0 + 1) % e.w;
    if (xb < 0) xb += e.w;
    const int ya = y0 < 0 ? 0 : (y0 > e.h - 1 ? e.h - 1 : y0);
    const int yb = y0 + 1 < 0 ? 0 : (y0 + 1 > e.h - 1 ? e.h - 1 : y0 + 1);
    const V3 c0 = env_texel(e, xa, ya) * (1.0f - fx) + env_texel(e, xb, ya) * fx;
    const V3 c1 = env_texel(e, xa, yb) * (1.0f - fx) + env_texel(e, xb, yb) * fx;
    return (c0 * (1.0f - fy) + c1 * fy) * e.intensity;
}

__device__ __forceinline__ V3 env_eval(const EnvView& e, V3 dir) {  // environment_light.cpp:296-299
    V3 sun = splat(0.0f);
    if (e.sun_pdf_weight > 0.0f && dot(normalized(dir), e.sun_dir) >= e.sun_cos_max) sun = e.sun_radiance;
    return env_eval_sky(e, dir) + sun;
}

__device__ __forceinline__ float env_image_pdf(const EnvView& e, V3 wi) {  // environment_light.cpp:266-282
    if (!e.valid || e.w <= 0 || e.h <= 0) return 0.25f * kInvPi;
    const V3 d = normalized(wi);
    float u, v;
    to_equirect(e, d, u, v);
    int x = static_cast<int>(u * static_cast<float>(e.w));
    int y = static_cast<int>(v * static_cast<float>(e.h));
    x = x > e.w - 1 ? e.w - 1 : (x < 0 ? 0 : x);
    y = y > e.h - 1 ? e.h - 1 : (y < 0 ? 0 : y);
    const float sin_theta = sqrtf(fmaxf(0.0f, 1.0f - d.y * d.y));
    if (sin_theta < 1e-4f) return 0.0f;
    const float f = e.func[static_cast<size_t>(y) * e.w + x];
    return (f / e.func_int) * static_cast<float>(e.w * e.h) / (kTwoPi * kPi * sin_theta);
}

__device__ __forceinline__ float env_pdf(const EnvView& e, V3 wi) {  // environment_light.cpp:288-294
    if (e.sun_pdf_weight <= 0.0f) return env_image_pdf(e, wi);
    float pdf = 0.0f;
    if (dot(normalized(wi), e.sun_dir) >= e.sun_cos_max) pdf += e.sun_pdf_weight / (kTwoPi * (1.0f - e.sun_cos_max));
    if (e.sun_pdf_weight < 1.0f) pdf += (1.0f - e.sun_pdf_weight) * env_image_pdf(e, wi);
    return pdf;
}

// sampleCdf (environment_light.cpp:22-36): the first entry of cdf[0..n] greater than u. The guide table
// brackets that entry between guide[k] and guide[k + 1] for k = floor(u K); the binary search over the
// bracket returns the same index as a search over the whole CDF, in a few dependent loads instead of log2(n).
__device__ __forceinline__ void sample_cdf(const float* cdf, int n, const uint32_t* guide, int K, float u, int& index,
                                           float& pdf, float& coord) {
    int k = static_cast<int>(u * static_cast<float>(K));
    k = k < 0 ? 0 : (k > K - 1 ? K - 1 : k);
    int lo = static_cast<int>(guide[k]);
    int count = static_cast<int>(guide[k + 1]) - lo;
    while (count > 0) {
        const int step = count >> 1;
        const int it = lo + step;
        if (!(u < cdf[it])) {
            lo = it + 1;
            count -= step + 1;
        } else {
            count = step;
        }
    }
    int i = lo - 1;
    if (i < 0) i = 0;
    if (i >= n) i = n - 1;
    pdf = cdf[i + 1] - cdf[i];
    float du = pdf > 0.0f ? (u - cdf[i]) / pdf : 0.0f;
    du = clampf(du, 0.0f, 1.0f);
    index = i;
    coord = (static_cast<float>(i) + du) / static_cast<float>(n);
}

__device__ __forceinline__ LightSample env_sample(const EnvView& e, float sx, float sy) {  // :191-237
    LightSample ls;
    ls.distance = kFloatMax;
    ls.pdf = 0.0f;
    ls.li = splat(0.0f);
    ls.wi = splat(0.0f);
    if (e.sun_pdf_weight > 0.0f) {
        if (sx < e.sun_pdf_weight) {
            const float ux = fminf(sx / e.sun_pdf_weight, 0.99999994f);
            const float cos_t = 1.0f - ux * (1.0f - e.sun_cos_max);
            const float sin_t = sqrtf(fmaxf(0.0f, 1.0f - cos_t * cos_t));
            const float phi = kTwoPi * sy;
            const Frame f = frame_from_normal(e.sun_dir);
            ls.wi = normalized(to_world(f, mk(sin_t * cosf(phi), sin_t * sinf(phi), cos_t)));
            ls.li = env_eval(e, ls.wi);
            ls.pdf = env_pdf(e, ls.wi);
            return ls;
        }
        if (e.sun_pdf_weight >= 1.0f) return ls;
        sx = fminf((sx - e.sun_pdf_weight) / (1.0f - e.sun_pdf_weight), 0.99999994f);
    }
    if (!e.valid || e.w <= 0) {
        const float z = 1.0f - 2.0f * sx;
        const float r = sqrtf(fmaxf(0.0f, 1.0f - z * z));
        const float phi = kTwoPi * sy;
        ls.wi = mk(r * cosf(phi), r * sinf(phi), z);
        ls.pdf = 0.25f * kInvPi;
        ls.li = env_eval(e, ls.wi);
    } else {
        int ri, ci;
        float rp, rc, cp, cc;
        sample_cdf(e.marginal, e.h, e.marginal_guide, e.marginal_k, sy, ri, rp, rc);
        sample_cdf(e.cond + static_cast<size_t>(ri) * static_cast<size_t>(e.w + 1), e.w,
                   e.cond_guide + static_cast<size_t>(ri) * static_cast<size_t>(e.cond_k + 1), e.cond_k, sx, ci, cp, cc);
        const float theta = rc * kPi;
        const float sin_theta = fmaxf(sinf(theta), 1e-6f);
        const float st = sinf(rc * kPi);
        const float a = cc * kTwoPi - kPi - e.rotation;
        ls.wi = normalized(mk(st * cosf(a), cosf(rc * kPi), -st * sinf(a)));
        ls.pdf = (rp * cp) * static_cast<float>(e.w * e.h) / (kTwoPi * kPi * sin_theta);
        ls.li = env_eval(e, ls.wi);
    }
    if (e.sun_pdf_weight > 0.0f && ls.pdf > 0.0f) ls.pdf = env_pdf(e, ls.wi);
    return ls;
}

__device__ __forceinline__ float env_importance(const EnvView& e, V3 n) {  // environment_light.cpp:184-189
    float imp = kPi * e.mean_luminance;
    if (e.sun_pdf_weight > 0.0f)
        imp += fmaxf(0.0f, luminance(e.sun_radiance)) * kTwoPi * (1.0f - e.sun_cos_max) * receiver_cosine(n, e.sun_dir);
    return imp;
}

// ------------------------------------------------------------------------------------ light selection
// LightSampler::Context (light_sampler.cpp:150-240) in its three modes: uniform, exhaustive (<= 33 candidates)
// and light BVH (more; the unbounded candidates form a top level together with the tree, light_sampler.cpp:42-77).
// The distribution depends only on (p, n): the picker at the previous vertex is recomputed from its (p, n) where
// a later emission hit needs its pmf, instead of being stored, and the top-level weights are recomputed where
// they are needed instead of being kept in registers.
__device__ __forceinline__ float candidate_weight(const SceneView& sc, int i, V3 p, V3 n) {
    const float w = i < sc.num_lights ? light_importance(sc.lights[i], p, n) : env_importance(sc.env, n);
    return (w > 0.0f && isfinite(w)) ? w : 0.0f;
}

__device__ __forceinline__ float safe_sqrt(float x) { return sqrtf(fmaxf(0.0f, x)); }
__device__ __forceinline__ float cos_sub_clamped(float sin_a, float cos_a, float sin_b, float cos_b) {
    return cos_a > cos_b ? 1.0f : cos_a * cos_b + sin_a * sin_b;
}
__device__ __forceinline__ float sin_sub_clamped(float sin_a, float cos_a, float sin_b, float cos_b) {
    return cos_a > cos_b ? 0.0f : sin_a * cos_b - cos_a * sin_b;
}

// LightBounds::importance (light_bounds.cpp:39-82) of a light-BVH node seen from (p, n).
__device__ __forceinline__ float node_importance(const GpuLightNode& nd, V3 p, V3 n) {
    const float phi = nd.lo_phi.w;
    if (!(phi > 0.0f)) return 0.0f;
    const V3 lo = mk(nd.lo_phi.x, nd.lo_phi.y, nd.lo_phi.z), hi = mk(nd.hi_cos_o.x, nd.hi_cos_o.y, nd.hi_cos_o.z);
    const V3 pc = (lo + hi) * 0.5f;
    const V3 diag = hi - lo;
    const float d2 = fmaxf(length_sq(p - pc), 0.25f * length_sq(diag));
    if (!(d2 > 0.0f)) return phi;
    const V3 axis = mk(nd.w_cos_e.x, nd.w_cos_e.y, nd.w_cos_e.z);
    V3 wi = p - pc;
    const float len = length(wi);
    wi = len > 0.0f ? wi / len : axis;
    float cos_w = dot(axis, wi);
    if (nd.ref.z) cos_w = fabsf(cos_w);
    const float sin_w = safe_sqrt(1.0f - cos_w * cos_w);
    const float r2 = 0.25f * length_sq(diag);
    const float dc2 = length_sq(p - pc);
    const float cos_b = dc2 <= r2 ? -1.0f : safe_sqrt(1.0f - r2 / dc2);
    const float sin_b = safe_sqrt(1.0f - cos_b * cos_b);
    const float cos_o = nd.hi_cos_o.w;
    const float sin_o = safe_sqrt(1.0f - cos_o * cos_o);
    const float cos_x = cos_sub_clamped(sin_w, cos_w, sin_o, cos_o);
    const float sin_x = sin_sub_clamped(sin_w, cos_w, sin_o, cos_o);
    const float cos_p = cos_sub_clamped(sin_x, cos_x, sin_b, cos_b);
    if (cos_p < nd.w_cos_e.w) return 0.0f;
    float imp = phi * cos_p / d2;
    if (length_sq(n) > 0.0f) {
        const float cos_i = fabsf(dot(wi, n));
        const float sin_i = safe_sqrt(1.0f - cos_i * cos_i);
        imp *= cos_sub_clamped(sin_i, cos_i, sin_b, cos_b);
    }
    return fmaxf(imp, 0.0f);
}

// Candidates the top-level weights run over: every candidate (exhaustive) or the unbounded ones plus the tree.
__device__ __forceinline__ int top_count(const SceneView& sc) {
    return sc.pick_mode == kPickBvh ? sc.num_top + (sc.num_light_nodes > 0 ? 1 : 0) : sc.light_count;
}

__device__ __forceinline__ float top_weight(const SceneView& sc, int k, V3 p, V3 n) {
    if (sc.pick_mode != kPickBvh) return candidate_weight(sc, k, p, n);
    if (k < sc.num_top) return candidate_weight(sc, sc.top[k], p, n);
    // LightBounds measures pi x radiance, Light::importance radiance (light_sampler.cpp:175-178).
    const float w = node_importance(sc.light_nodes[0], p, n) * kInvPi;
    return (w > 0.0f && isfinite(w)) ? w : 0.0f;
}

// bvhSample (light_sampler.cpp:242-266): descend by the children's importance, re-stretching u at each level.
__device__ __forceinline__ int bvh_sample(const SceneView& sc, V3 p, V3 n, float u, float& pmf) {
    pmf = 1.0f;
    int node = 0;
    for (;;) {
        const int4 r = sc.light_nodes[node].ref;
        if (r.y >= 0) return r.y;
        const int c0 = node + 1, c1 = r.x;
        const float i0 = node_importance(sc.light_nodes[c0], p, n);
        const float i1 = node_importance(sc.light_nodes[c1], p, n);
        if (!(i0 > 0.0f) && !(i1 > 0.0f)) {
            pmf = 0.0f;
            return -1;
        }
        const float p0 = i0 / (i0 + i1);
        if (u < p0) {
            node = c0;
            u = fminf(u / p0, 0.99999994f);
            pmf *= p0;
        } else {
            node = c1;
            u = fminf((u - p0) / (1.0f - p0), 0.99999994f);
            pmf *= 1.0f - p0;
        }
    }
}

// bvhPmf (light_sampler.cpp:268-293): the product of the branch probabilities along the candidate's bit trail.
__device__ __forceinline__ float bvh_pmf(const SceneView& sc, const LightPath& lp, int index, V3 p, V3 n) {
    uint64_t trail = (static_cast<uint64_t>(lp.trail_hi) << 32) | lp.trail_lo;
    float pmf = 1.0f;
    int node = 0;
    for (int d = 0; d < lp.depth; ++d) {
        const int4 r = sc.light_nodes[node].ref;
        if (r.y >= 0) break;
        const int c0 = node + 1, c1 = r.x;
        const float i0 = node_importance(sc.light_nodes[c0], p, n);
        const float i1 = node_importance(sc.light_nodes[c1], p, n);
        const float sum = i0 + i1;
        if (!(sum > 0.0f)) return 0.0f;
        if (trail & 1u) {
            pmf *= i1 / sum;
            node = c1;
        } else {
            pmf *= i0 / sum;
            node = c0;
        }
        trail >>= 1;
    }
    return sc.light_nodes[node].ref.y == index ? pmf : 0.0f;
}

struct Picker {
    float a, b;      // top-level pmf(k) = w_k * a + b
    bool uniform;    // Context::m_uniform: every candidate 1 / light_count
};

__device__ __forceinline__ Picker picker_prepare(const SceneView& sc, V3 p, V3 n) {
    Picker pk;
    pk.uniform = true;
    pk.a = 0.0f;
    pk.b = sc.light_count > 0 ? 1.0f / static_cast<float>(sc.light_count) : 0.0f;
    if (sc.light_count <= 1 || sc.pick_mode == kPickUniform) return pk;
    const int count = top_count(sc);
    float sum = 0.0f;
    for (int k = 0; k < count; ++k) sum += top_weight(sc, k, p, n);
    pk.uniform = false;
    if (!(sum > 0.0f) || !isfinite(sum)) {
        pk.a = 0.0f;  // every top-level entry 1 / count (light_sampler.cpp:185-186)
        pk.b = 1.0f / static_cast<float>(count);
        return pk;
    }
    pk.a = (1.0f - 0.15f) / sum;
    pk.b = 0.15f / static_cast<float>(count);
    return pk;
}

__device__ __forceinline__ float picker_pmf(const SceneView& sc, const Picker& pk, int index, V3 p, V3 n) {
    if (index < 0 || index >= sc.light_count) return 0.0f;
    if (pk.uniform) return pk.b;
    if (sc.pick_mode != kPickBvh) return top_weight(sc, index, p, n) * pk.a + pk.b;
    const LightPath lp = sc.light_paths[index];
    if (lp.top_slot >= 0) return top_weight(sc, lp.top_slot, p, n) * pk.a + pk.b;
    if (sc.num_light_nodes == 0) return 0.0f;
    return (top_weight(sc, sc.num_top, p, n) * pk.a + pk.b) * bvh_pmf(sc, lp, index, p, n);
}

__device__ __forceinline__ int picker_sample(const SceneView& sc, const Picker& pk, float u, V3 p, V3 n, float& pmf) {
    pmf = 0.0f;
    const int count = sc.light_count;
    if (count <= 0) return -1;
    if (pk.uniform) {
        pmf = pk.b;
        const int k = static_cast<int>(u * static_cast<float>(count));
        return k < count - 1 ? k : count - 1;
    }
    const int tc = top_count(sc);
    float acc = 0.0f;
    int k = tc - 1;
    for (int i = 0; i < tc - 1; ++i) {
        const float pi = top_weight(sc, i, p, n) * pk.a + pk.b;
        if (u < acc + pi) {
            k = i;
            break;
        }
        acc += pi;
    }
    const float pk_k = top_weight(sc, k, p, n) * pk.a + pk.b;
    if (sc.pick_mode != kPickBvh) {
        pmf = pk_k;
        return k;
    }
    if (k < sc.num_top) {
        pmf = pk_k;
        return sc.top[k];
    }
    const float ur = fminf((u - acc) / pk_k, 0.99999994f);
    float tree_pmf = 0.0f;
    const int light = bvh_sample(sc, p, n, fmaxf(ur, 0.0f), tree_pmf);
    pmf = pk_k * tree_pmf;
    return light;
}

}  // namespace pe

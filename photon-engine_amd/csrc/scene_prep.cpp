// scene_prep.cpp — host preparation of the shading-side scene: materials with their material-only terms
// precomputed, light records, the environment's sampling distributions and the energy tables.
//
// Every quantity is computed with the reference's operations in the reference's order (cited per function)
// and without contraction, so the device sees the same floats the reference's IEEE build computes.
#include "pe_scene_host.h"
#include "pe_microfacet.h"
#include "light_bvh.h"

#include <algorithm>
#include <cmath>
#include <memory>
#include <string>
#include <vector>

namespace pe {

namespace {

V3 c3(const float* v) { return mk(v[0], v[1], v[2]); }
V3 clamp01v(V3 c) { return clamp3(c, 0.0f, 1.0f); }

float f0_from_ior(float ior) {  // disney.cpp:25-28
    const float r = (ior - 1.0f) / (ior + 1.0f);
    return r * r;
}

}  // namespace

// DisneyMaterial setters (disney.h:22-126), then the per-material terms of disney.cpp.
int prepare_material(const pe_material_desc& d, const EnergyTables& T, GpuMaterial& g, std::string& err) {
    g = GpuMaterial{};
    g.opacity = 1.0f;
    g.emission = c3(d.emission);
    g.flags = is_black(g.emission) ? 0u : kMatFlagEmissive;
    if (d.type == PE_MAT_LAMBERT) {
        g.type = kMatLambert;
        g.base_color = c3(d.base_color);
        return PE_OK;
    }
    if (d.type == PE_MAT_MIRROR) {  // Mirror(reflectance) (mirror.h:14): a delta reflection; Material::emitted is black
        g.type = kMatMirror;
        g.base_color = c3(d.base_color);
        g.emission = mk(0.0f, 0.0f, 0.0f);
        g.flags = kMatFlagDelta;
        return PE_OK;
    }
    if (d.type == PE_MAT_DIELECTRIC) {
        // Dielectric(ior, tint, roughness) + setAbbe / setAbsorption / setThinWalled (dielectric.h:18-78,
        // dielectric.cpp:123-144). The interior medium only absorbs: the description carries no scattering.
        g.type = kMatDielectric;
        g.base_color = c3(d.base_color);
        g.ior = d.ior;
        g.roughness = d.roughness;
        g.glass_alpha = ggx_clamp_alpha(d.roughness * d.roughness);
        g.abbe = d.abbe > 0.0f ? std::fmax(d.abbe, 10.0f) : 0.0f;
        const bool thin = d.thin_walled != 0;
        if (thin || d.roughness <= 1e-3f) g.flags |= kMatFlagDelta;
        if (thin) g.flags |= kMatFlagThinWall;
        else g.flags |= kMatFlagVolume;
        if (g.abbe > 0.0f) g.flags |= kMatFlagDispersive;
        if (d.thin_film_thickness > 0.0f && thin) {  // setThinFilm (dielectric.h:46); only a thin wall uses it
            g.flags |= kMatFlagFilm;
            g.film_thickness = d.thin_film_thickness;
            g.film_ior = d.ior;
        }
        if (!thin && d.absorption_distance > 0.0f) {
            const V3 color = clamp01v(c3(d.absorption_color));
            const float dist = std::fmax(d.absorption_distance, 1e-5f);
            g.sigma_a = mk(-std::log(std::fmax(color.x, 1e-4f)) / dist, -std::log(std::fmax(color.y, 1e-4f)) / dist,
                           -std::log(std::fmax(color.z, 1e-4f)) / dist);
        }
        if (!thin && d.scatter_density > 0.0f) {  // rebuildMedium's scattering terms (dielectric.cpp:141-143)
            const float dens = d.scatter_density;
            const V3 sc = clamp01v(c3(d.scatter_color));
            g.sigma_s = sc * dens;
            g.sigma_a = g.sigma_a + (splat(1.0f) - sc) * dens;
            g.medium_g = clampf(d.scatter_anisotropy, -0.95f, 0.95f);
            if (!is_black(g.sigma_s)) g.flags |= kMatFlagScatter;
        }
        if (!thin && !is_black(g.sigma_a)) g.flags |= kMatFlagMedium;
        return PE_OK;
    }
    if (d.type != PE_MAT_DISNEY) {
        err = "unknown material type " + std::to_string(d.type);
        return PE_ERR_INVALID;
    }
    g.type = kMatDisney;
    g.base_color = c3(d.base_color);
    g.metallic = d.metallic;
    g.roughness = std::fmax(0.001f, d.roughness);
    g.specular = d.specular;
    g.clearcoat = clampf(d.clearcoat, 0.0f, 1.0f);
    g.clearcoat_roughness = std::fmax(0.001f, d.clearcoat_roughness);
    g.ior = clampf(d.ior, 1.0001f, 3.0f);
    g.specular_transmission = clampf(d.specular_transmission, 0.0f, 1.0f);
    g.coat_ior = clampf(d.coat_ior, 1.0001f, 3.0f);
    g.coat_tint = clamp01v(c3(d.coat_tint));
    g.anisotropy = clampf(d.anisotropy, 0.0f, 1.0f);
    const float sheen = clampf(d.sheen, 0.0f, 1.0f);
    g.sheen_roughness = clampf(d.sheen_roughness, 0.02f, 1.0f);
    g.diffuse_roughness = clampf(d.diffuse_roughness, 0.0f, 1.0f);
    g.diffuse_transmission = clampf(d.diffuse_transmission, 0.0f, 1.0f);
    g.subsurface = clampf(d.subsurface, 0.0f, 1.0f);
    g.edge_tint = clamp01v(c3(d.edge_tint));
    if (g.specular_transmission > 0.0f) g.flags |= kMatFlagVolume;  // isVolumeBoundary (disney.h:76)
    if (d.has_opacity) {  // setOpacity clamps (disney.h:148); transmitsShadows below 1 (:158)
        g.opacity = clampf(d.opacity, 0.0f, 1.0f);
        if (g.opacity < 1.0f) g.flags |= kMatFlagCutout;
    }
    if (d.flake_density > 0.0f) {  // setFlakes clamps (disney.h:120-125); resolve() flakes when density > 0 (:601)
        g.flake_density = clampf(d.flake_density, 0.0f, 1.0f);
        g.flake_size = std::max(d.flake_size, 1e-6f);
        g.flake_roughness = clampf(d.flake_roughness, 0.0f, 1.0f);
        g.flake_color = clamp01v(c3(d.flake_color));
        if (g.flake_density > 0.0f) g.flags |= kMatFlagFlakes;
    }

    // resolvedSheenColor (disney.cpp:160-166)
    V3 sc = splat(0.0f);
    if (!(sheen <= 0.0f || g.metallic >= 1.0f)) {
        if (d.has_sheen_color) {
            sc = clamp01v(c3(d.sheen_color)) * (sheen * (1.0f - g.metallic));
        } else {
            const float lum = luminance(g.base_color);
            const V3 tint = lum > 1e-6f ? g.base_color * (1.0f / lum) : splat(1.0f);
            sc = (splat(0.5f) + clamp01v(tint) * 0.5f) * (sheen * (1.0f - g.metallic));
        }
    }
    g.sheen_color = sc;
    g.sheen_max = std::max(sc.x, std::max(sc.y, sc.z));
    if (!is_black(sc)) g.flags |= kMatFlagSheen;
    if (g.clearcoat > 0.0f) g.flags |= kMatFlagCoat;

    // anisoAlpha (disney.cpp:55-60) and DielectricLayer (:128-146)
    const float aspect = std::sqrt(std::fmax(0.1f, 1.0f - 0.9f * clampf(g.anisotropy, 0.0f, 1.0f)));
    const float a = g.roughness * g.roughness;
    g.ax = a / aspect;
    g.ay = a * aspect;
    g.alpha = std::sqrt(g.ax * g.ay);
    g.f0d = clampf(f0_from_ior(g.ior) * std::fmax(1.0f, 2.0f * g.specular), 0.0f, 1.0f);
    g.wd = clampf(2.0f * g.specular, 0.0f, 1.0f);
    g.inv_one_minus_avg = 1.0f / std::fmax(1e-4f, 1.0f - g.wd * ggx_schlick_average_albedo(T, g.alpha, g.f0d));
    g.f0c = clampf(f0_from_ior(g.coat_ior), 0.0f, 1.0f);
    g.alpha_c = g.clearcoat_roughness * g.clearcoat_roughness;
    g.coat_tint_factor = g.clearcoat > 0.0f ? lerp3(splat(1.0f), g.coat_tint, g.clearcoat) : splat(1.0f);

    // lobeWeights (disney.cpp:313-338) off a flake (a flake point's weights are taken at the shading point)
    const float p_coat = 0.25f * g.clearcoat;
    const float rem = 1.0f - p_coat;
    const float st = clampf(g.specular_transmission, 0.0f, 1.0f) * (1.0f - g.metallic);
    const float lo = 0.5f * (st > 0.0f ? 1.0f : g.wd);
    const float p_spec = rem * (lo + g.metallic * (1.0f - lo));  // lerp(a, b, t) = a + t * (b - a), utils.h:32
    float p_rest = rem - p_spec;
    const float p_st = p_rest * st;
    p_rest -= p_st;
    const float p_trans = p_rest * clampf(g.diffuse_transmission, 0.0f, 1.0f);
    float p_diff = p_rest - p_trans;
    const float p_sss = g.metallic < 1.0f ? p_diff * g.subsurface : 0.0f;
    p_diff -= p_sss;
    g.p_coat = p_coat;
    g.p_spec = p_spec;
    g.p_spec_trans = p_st;
    g.p_diff_refl = p_diff;
    g.p_sss = p_sss;
    g.p_trans = p_trans;

    // evalLobes weights (disney.cpp:230-237, :263)
    g.st_w = st;
    g.opaque_dielectric = (1.0f - g.metallic) * (1.0f - clampf(g.specular_transmission, 0.0f, 1.0f));
    g.trans_w = clampf(g.diffuse_transmission, 0.0f, 1.0f) * g.opaque_dielectric;
    g.diff_w = g.opaque_dielectric * (1.0f - clampf(g.diffuse_transmission, 0.0f, 1.0f)) * (1.0f - g.subsurface);

    // subsurfaceCoeffs (disney.cpp:738-757, Chiang et al. 2016 inversion): material-only without texture maps.
    if (g.subsurface > 0.0f) {
        const float gg = clampf(d.subsurface_anisotropy, -0.9f, 0.9f);
        const V3 radius = clamp3(c3(d.subsurface_radius), 1e-4f, 1e4f);
        const float scale = std::fmax(d.subsurface_scale, 1e-6f);
        float st[3], alb[3];
        for (int c = 0; c < 3; ++c) {
            const float A = clampf(comp(g.base_color, c), 0.0f, 0.999f);
            const float dd = std::fmax(comp(radius, c) * scale, 1e-7f);
            const float alpha = 1.0f - std::exp(A * (-5.09406f + A * (2.61188f - 4.31805f * A)));
            const float s_val = 1.9f - A + 3.5f * (A - 0.8f) * (A - 0.8f);
            float sigma_t = 1.0f / std::fmax(dd * s_val, 1e-12f);
            float sigma_s = alpha * sigma_t;
            const float sigma_a = sigma_t - sigma_s;
            sigma_s /= (1.0f - gg);
            sigma_t = sigma_s + sigma_a;
            st[c] = sigma_t;
            alb[c] = sigma_t > 0.0f ? sigma_s / sigma_t : 0.0f;
        }
        g.sss_sigma_t = mk(st[0], st[1], st[2]);
        g.sss_albedo = mk(alb[0], alb[1], alb[2]);
        g.sss_g = gg;
    }
    // Thin film on the base specular lobe (setThinFilm, disney.h:111-114; specularFresnel, disney.cpp:79-96):
    // the dielectric base index from f0 and the metal base n + ik per channel (gulbrandsenIor, fresnel.h:49-58).
    const float film = clampf(d.thin_film_thickness, 0.0f, 2000.0f);
    if (film > 0.0f) {
        g.flags |= kMatFlagFilm;
        g.film_thickness = film;
        g.film_ior = clampf(d.thin_film_ior, 1.0f, 3.0f);
        const float sd = std::sqrt(clampf(g.f0d, 0.0f, 0.99f));
        g.film_nd = (1.0f + sd) / (1.0f - sd);
        float mn[3], mkk[3];
        for (int c = 0; c < 3; ++c) {
            const float r = clampf(comp(g.base_color, c), 0.0f, 0.99f);
            const float gt = clampf(comp(g.edge_tint, c), 0.0f, 1.0f);
            const float sr = std::sqrt(r);
            const float n_min = (1.0f - r) / (1.0f + r);
            const float n_max = (1.0f + sr) / (1.0f - sr);
            const float n = gt * n_min + (1.0f - gt) * n_max;
            const float k2 = (r * (n + 1.0f) * (n + 1.0f) - (n - 1.0f) * (n - 1.0f)) / (1.0f - r);
            mn[c] = n;
            mkk[c] = std::sqrt(std::fmax(0.0f, k2));
        }
        g.film_metal_n = mk(mn[0], mn[1], mn[2]);
        g.film_metal_k = mk(mkk[0], mkk[1], mkk[2]);
    }
    return PE_OK;
}

// AreaLight constructor (area_light.h:20-27) / TriangleLight constructor (area_light.cpp:195-200).
GpuLight prepare_light(const pe_light_desc& d) {
    GpuLight L{};
    L.type = d.type;
    L.p0 = c3(d.p0);
    L.p1 = c3(d.p1);
    L.p2 = c3(d.p2);
    L.radiance = c3(d.radiance);
    L.lum = std::fmax(0.0f, luminance(L.radiance));
    if (d.type == PE_LIGHT_AREA) {
        const V3 n = cross(L.p1, L.p2);
        L.area = length(n);
        L.normal = normalized(n);
        L.len_u = length(L.p1);
        L.len_v = length(L.p2);
        L.rectangle = (L.len_u > 0.0f && L.len_v > 0.0f && std::fabs(dot(L.p1, L.p2)) <= 1e-4f * L.len_u * L.len_v) ? 1 : 0;
        L.center = L.p0 + (L.p1 + L.p2) * 0.5f;  // AreaLight::importance (area_light.cpp:101)
    } else {
        const V3 n = cross(L.p1 - L.p0, L.p2 - L.p0);
        const float len = length(n);
        L.area = 0.5f * len;
        L.normal = len > 0.0f ? n / len : mk(0.0f, 1.0f, 0.0f);
        L.center = (L.p0 + L.p1 + L.p2) * (1.0f / 3.0f);  // TriangleLight::importance (area_light.cpp:203)
    }
    return L;
}

namespace {

struct Distribution {
    float func_int = 0.0f;
    std::vector<float> func, marginal, cond;
    bool valid() const { return func_int > 0.0f && !marginal.empty(); }
};

// EnvironmentLight::buildCdf (environment_light.cpp:149-172)
void build_cdf(Distribution& d, int w, int h) {
    d.marginal.assign(static_cast<size_t>(h) + 1, 0.0f);
    d.cond.assign(static_cast<size_t>(h) * static_cast<size_t>(w + 1), 0.0f);
    for (int y = 0; y < h; ++y) {
        float* row = d.cond.data() + static_cast<size_t>(y) * static_cast<size_t>(w + 1);
        row[0] = 0.0f;
        for (int x = 0; x < w; ++x) row[x + 1] = row[x] + d.func[static_cast<size_t>(y) * w + x];
        d.marginal[static_cast<size_t>(y) + 1] = d.marginal[static_cast<size_t>(y)] + row[w];
    }
    d.func_int = d.marginal[static_cast<size_t>(h)];
    if (!(d.func_int > 0.0f)) {
        d.func_int = 0.0f;
        d.marginal.clear();
        return;
    }
    const float inv = 1.0f / d.func_int;
    for (float& c : d.marginal) c *= inv;
    for (int y = 0; y < h; ++y) {
        float* row = d.cond.data() + static_cast<size_t>(y) * static_cast<size_t>(w + 1);
        const float row_sum = row[w];
        if (row_sum <= 0.0f) continue;
        const float inv_row = 1.0f / row_sum;
        for (int x = 0; x <= w; ++x) row[x] *= inv_row;
    }
}

int guide_buckets(int n, int cap) {
    int k = 1;
    while (k < n && k < cap) k <<= 1;
    return k;
}

// Guide table of cdf[0..n] with K buckets (see EnvView): g[k] = first index whose value exceeds k / K, or n + 1.
// A CDF that is not non-decreasing (NaN entries) gets the whole range for every bucket, which is the plain
// binary search.
void build_guide(const float* cdf, int n, int K, uint32_t* g) {
    bool monotone = true;
    for (int i = 0; i < n && monotone; ++i) monotone = cdf[i] <= cdf[i + 1];
    if (!monotone) {
        for (int k = 0; k <= K; ++k) g[k] = k == K ? static_cast<uint32_t>(n + 1) : 0u;
        return;
    }
    int idx = 0;
    for (int k = 0; k <= K; ++k) {
        const float t = static_cast<float>(k) / static_cast<float>(K);
        while (idx <= n && !(t < cdf[idx])) ++idx;
        g[k] = static_cast<uint32_t>(idx);
    }
    g[0] = 0u;
}

}  // namespace

// EnvironmentLight(image, rotation, intensity) + setSun (environment_light.cpp:64-182); the sampling
// distribution kept on the device is the MIS-compensated one (or the full one where compensation is off).
int prepare_environment(pe_device* dev, const pe_env_desc& e, pe_scene& sc, EnvView& ev) {
    ev = EnvView{};
    ev.rotation = e.rotation;
    ev.intensity = e.intensity;
    const bool has_image = e.rgb && e.width > 0 && e.height > 0;
    const int w = has_image ? e.width : 0, h = has_image ? e.height : 0;
    Distribution full, mis;
    if (has_image) {
        const float* rgb = e.rgb;
        auto lum_at = [&](int x, int y) {
            const float* p = rgb + 3 * (static_cast<size_t>(y) * w + x);
            return luminance(mk(p[0], p[1], p[2]));
        };
        full.func.assign(static_cast<size_t>(w) * h, 0.0f);
        std::vector<float> sin_row(static_cast<size_t>(h));
        double weighted = 0.0, weights = 0.0;
        for (int y = 0; y < h; ++y) {
            const float sin_theta = std::sin(kPi * (static_cast<float>(y) + 0.5f) / static_cast<float>(h));
            sin_row[static_cast<size_t>(y)] = sin_theta;
            for (int x = 0; x < w; ++x) {
                float L = lum_at(x, y);
                if (!(L > 0.0f)) L = 0.0f;
                full.func[static_cast<size_t>(y) * w + x] = L * sin_theta;
                weighted += static_cast<double>(L) * sin_theta;
                weights += sin_theta;
            }
        }
        build_cdf(full, w, h);
        if (full.valid()) {
            const float mean_l = weights > 0.0 ? static_cast<float>(weighted / weights) : 0.0f;
            mis.func.resize(full.func.size());
            double mis_total = 0.0, full_total = 0.0;
            for (int y = 0; y < h; ++y) {
                const float s = sin_row[static_cast<size_t>(y)];
                for (int x = 0; x < w; ++x) {
                    const size_t i = static_cast<size_t>(y) * w + x;
                    float L = lum_at(x, y);
                    if (!(L > 0.0f)) L = 0.0f;
                    const float excess = L - mean_l;
                    mis.func[i] = excess > 1e-4f * mean_l ? excess * s : 0.0f;
                    mis_total += mis.func[i];
                    full_total += full.func[i];
                }
            }
            bool comp = mis_total > 1e-3 * full_total;
            if (comp) {
                build_cdf(mis, w, h);
                comp = mis.valid();
            }
            if (!comp) mis = full;
        }
        // power() (environment_light.cpp:308-323): every fourth texel in both directions.
        V3 sum = splat(0.0f);
        float wsum = 0.0f;
        for (int y = 0; y < h; y += 4) {
            const float s = std::sin(kPi * (static_cast<float>(y) + 0.5f) / static_cast<float>(h));
            for (int x = 0; x < w; x += 4) {
                const float* p = rgb + 3 * (static_cast<size_t>(y) * w + x);
                sum += mk(p[0], p[1], p[2]) * s;
                wsum += s;
            }
        }
        const V3 power = (wsum > 0.0f ? sum / wsum : sum) * e.intensity * 4.0f * kPi;
        ev.mean_luminance = std::fmax(0.0f, luminance(power) / (4.0f * kPi));
    } else {
        const V3 power = splat(0.5f) * e.intensity * 4.0f * kPi;
        ev.mean_luminance = std::fmax(0.0f, luminance(power) / (4.0f * kPi));
    }
    ev.w = w;
    ev.h = h;
    ev.valid = mis.valid() ? 1 : 0;
    ev.func_int = mis.func_int;
    // setSun (environment_light.cpp:168-182)
    ev.sun_dir = mk(0.0f, 1.0f, 0.0f);
    ev.sun_cos_max = 1.0f;
    ev.sun_pdf_weight = 0.0f;
    if (e.has_sun) {
        ev.sun_dir = normalized(c3(e.sun_direction));
        ev.sun_cos_max = clampf(e.sun_cos_max, -0.999f, 0.9999999f);
        ev.sun_radiance = c3(e.sun_radiance) * e.intensity;
        const float omega = kTwoPi * (1.0f - ev.sun_cos_max);
        const float sun_power = std::fmax(0.0f, luminance(ev.sun_radiance)) * omega;
        const float sky_power = 4.0f * kPi * ev.mean_luminance;
        if (sun_power > 0.0f) {
            ev.sun_pdf_weight = clampf(sun_power / (sun_power + sky_power), 0.15f, 0.85f);
            if (!full.valid()) ev.sun_pdf_weight = 1.0f;
        }
    }
    hipStream_t s = dev->stream;
    if (has_image) {
        if (!sc.env_rgb.upload(e.rgb, 3ull * w * h * sizeof(float), s)) return PE_ERR_OOM;
        std::vector<uint32_t> mguide, cguide;
        if (mis.valid()) {
            ev.marginal_k = guide_buckets(h, 4096);
            ev.cond_k = guide_buckets(w / 2, 1024);
            mguide.resize(static_cast<size_t>(ev.marginal_k) + 1);
            cguide.resize(static_cast<size_t>(h) * static_cast<size_t>(ev.cond_k + 1));
            build_guide(mis.marginal.data(), h, ev.marginal_k, mguide.data());
            for (int y = 0; y < h; ++y)
                build_guide(mis.cond.data() + static_cast<size_t>(y) * static_cast<size_t>(w + 1), w, ev.cond_k,
                            cguide.data() + static_cast<size_t>(y) * static_cast<size_t>(ev.cond_k + 1));
            if (!sc.env_func.upload(mis.func.data(), mis.func.size() * sizeof(float), s) ||
                !sc.env_marginal.upload(mis.marginal.data(), mis.marginal.size() * sizeof(float), s) ||
                !sc.env_cond.upload(mis.cond.data(), mis.cond.size() * sizeof(float), s) ||
                !sc.env_mguide.upload(mguide.data(), mguide.size() * sizeof(uint32_t), s) ||
                !sc.env_cguide.upload(cguide.data(), cguide.size() * sizeof(uint32_t), s))
                return PE_ERR_OOM;
        }
        const hipError_t er = hipStreamSynchronize(s);
        if (er != hipSuccess) {
            set_error(std::string("environment upload: ") + hipGetErrorString(er));
            return PE_ERR_HIP;
        }
    }
    ev.rgb = sc.env_rgb.as<float>();
    ev.func = sc.env_func.as<float>();
    ev.marginal = sc.env_marginal.as<float>();
    ev.cond = sc.env_cond.as<float>();
    ev.marginal_guide = sc.env_mguide.as<uint32_t>();
    ev.cond_guide = sc.env_cguide.as<uint32_t>();
    return PE_OK;
}

// Material table (the scene's materials, then the walk-exit Lambert) and the energy tables; the view's material
// fields. Everything is validated before anything is uploaded, so a refused update leaves the scene as it was.
int prepare_materials(pe_device* dev, const pe_scene_desc& desc, pe_scene& sc) {
    if (desc.num_materials < 0 || (desc.num_materials > 0 && !desc.materials)) {
        set_error("scene: bad material list");
        return PE_ERR_INVALID;
    }
    std::unique_ptr<EnergyTables> T(new EnergyTables());
    build_energy_tables(*T);
    // One slot past the scene's materials holds the white Lambert that a subsurface walk's exit vertex is
    // shaded with (subsurfaceExitMaterial, path_tracer.cpp:56-59).
    const int num_mats = desc.num_materials;
    std::vector<GpuMaterial> mats(static_cast<size_t>(num_mats) + 1);
    bool transmitters = false, transport = false;
    uint32_t classes = 0;  // sort classes of the scene's own materials (not the walk-exit Lambert)
    for (int i = 0; i <= num_mats; ++i) {
        pe_material_desc exit_desc{};
        if (i == num_mats) {
            exit_desc.type = PE_MAT_LAMBERT;
            exit_desc.base_color[0] = exit_desc.base_color[1] = exit_desc.base_color[2] = 1.0f;
        }
        const pe_material_desc& d = i < num_mats ? desc.materials[i] : exit_desc;
        std::string err;
        int rc = prepare_material(d, *T, mats[static_cast<size_t>(i)], err);
        if (rc == PE_OK && (mats[static_cast<size_t>(i)].flags & kMatFlagVolume) && i > 0xFFFF) {
            // A path's medium stack holds 16-bit material indices.
            err = "volume-boundary material index above 65535";
            rc = PE_ERR_UNSUPPORTED;
        }
        if (rc != PE_OK) {
            set_error("material " + std::to_string(i) + ": " + err);
            return rc;
        }
        GpuMaterial& g = mats[static_cast<size_t>(i)];
        transmitters = transmitters || (g.flags & (kMatFlagThinWall | kMatFlagCutout)) != 0;
        const bool lobes_below = g.trans_w > 0.0f || g.st_w > 0.0f || g.p_trans > 0.0f || g.p_spec_trans > 0.0f ||
                                 g.p_sss > 0.0f;
        transport = transport || g.type == kMatDielectric || lobes_below ||
                    (g.flags & (kMatFlagVolume | kMatFlagThinWall | kMatFlagDispersive | kMatFlagMedium |
                                kMatFlagFilm | kMatFlagCutout | kMatFlagScatter | kMatFlagFlakes)) != 0;
        g.sort_class = g.type == kMatDielectric ? kSortClassDielectric
                       : g.type != kMatDisney   ? kSortClassSimple
                       : (lobes_below || (g.flags & (kMatFlagFilm | kMatFlagCutout | kMatFlagFlakes)))
                           ? kSortClassDisneyTransport
                           : kSortClassDisney;
        if (i < num_mats) classes |= 1u << g.sort_class;
    }
    hipStream_t s = dev->stream;
    if (!sc.materials.upload(mats.data(), mats.size() * sizeof(GpuMaterial), s) ||
        !sc.tables.upload(T.get(), sizeof(EnergyTables), s))
        return PE_ERR_OOM;
    const hipError_t er = hipStreamSynchronize(s);  // the host copies die at return
    if (er != hipSuccess) {
        set_error(std::string("material upload: ") + hipGetErrorString(er));
        return PE_ERR_HIP;
    }
    SceneView& v = sc.view;
    v.materials = sc.materials.as<GpuMaterial>();
    v.tables = sc.tables.as<EnergyTables>();
    v.exit_material = num_mats;
    v.has_transmitters = transmitters ? 1 : 0;
    v.transport = transport ? 1 : 0;
    v.sort_classes = __builtin_popcount(classes);
    return PE_OK;
}

// Explicit lights, the environment and the light-selection structure; the view's light fields.
int prepare_lights(pe_device* dev, const pe_scene_desc& desc, pe_scene& sc) {
    if (desc.num_lights < 0 || (desc.num_lights > 0 && !desc.lights)) {
        set_error("scene: bad light list");
        return PE_ERR_INVALID;
    }
    if (desc.background_mode < PE_BACKGROUND_ENVIRONMENT || desc.background_mode > PE_BACKGROUND_TRANSPARENT) {
        set_error("scene: unknown background mode " + std::to_string(desc.background_mode));
        return PE_ERR_INVALID;
    }
    std::vector<GpuLight> lights(static_cast<size_t>(std::max(1, desc.num_lights)));
    for (int i = 0; i < desc.num_lights; ++i) {
        const pe_light_desc& L = desc.lights[i];
        if (L.type != PE_LIGHT_AREA && L.type != PE_LIGHT_TRIANGLE) {
            set_error("light " + std::to_string(i) + ": unknown type " + std::to_string(L.type));
            return PE_ERR_INVALID;
        }
        lights[static_cast<size_t>(i)] = prepare_light(L);
        lights[static_cast<size_t>(i)].group = -1;
    }
    // Light groups as the caller resolved them (Scene::buildAccelerator, scene.cpp:144-173).
    const bool grouped = desc.light_groups && desc.num_light_groups > 0;
    if (grouped) {
        if (desc.num_light_groups > PE_MAX_LIGHT_GROUPS) {
            set_error("scene: more than 16 light groups");
            return PE_ERR_INVALID;
        }
        for (int i = 0; i < desc.num_lights; ++i) {
            const int g = desc.light_groups[i];
            if (g < -1 || g >= desc.num_light_groups) {
                set_error("light " + std::to_string(i) + ": light group out of range");
                return PE_ERR_INVALID;
            }
            lights[static_cast<size_t>(i)].group = g;
        }
        if (desc.env_light_group < -1 || desc.env_light_group >= desc.num_light_groups) {
            set_error("scene: environment light group out of range");
            return PE_ERR_INVALID;
        }
    }
    // LightSampler::build (light_sampler.cpp:42-77): up to 33 candidates are weighed exhaustively, more go into
    // the light BVH with the unbounded ones on a top level.
    const bool has_env = desc.environment != nullptr;
    LightBvh lb;
    build_light_bvh(lights, desc.num_lights, has_env, lb);
    hipStream_t s = dev->stream;
    if (!sc.lights.upload(lights.data(), lights.size() * sizeof(GpuLight), s)) return PE_ERR_OOM;
    EnvView env{};
    if (has_env) {
        const int rc = prepare_environment(dev, *desc.environment, sc, env);
        if (rc != PE_OK) return rc;
    } else {
        for (DevBuf* b : {&sc.env_rgb, &sc.env_func, &sc.env_marginal, &sc.env_cond, &sc.env_mguide, &sc.env_cguide})
            b->release();
    }
    if (lb.mode == kPickBvh) {
        if (!lb.nodes.empty() && !sc.light_nodes.upload(lb.nodes.data(), lb.nodes.size() * sizeof(GpuLightNode), s))
            return PE_ERR_OOM;
        if (lb.nodes.empty()) sc.light_nodes.release();
        if (!sc.light_paths.upload(lb.paths.data(), lb.paths.size() * sizeof(LightPath), s)) return PE_ERR_OOM;
    } else {
        sc.light_nodes.release();
        sc.light_paths.release();
    }
    const hipError_t el = hipStreamSynchronize(s);
    if (el != hipSuccess) {
        set_error(std::string("light upload: ") + hipGetErrorString(el));
        return PE_ERR_HIP;
    }
    SceneView& v = sc.view;
    v.lights = sc.lights.as<GpuLight>();
    v.num_lights = desc.num_lights;
    v.has_env = has_env ? 1 : 0;
    v.light_count = desc.num_lights + v.has_env;
    v.env = env;
    v.pick_mode = lb.mode;
    v.light_nodes = sc.light_nodes.as<GpuLightNode>();
    v.light_paths = sc.light_paths.as<LightPath>();
    v.num_light_nodes = lb.mode == kPickBvh ? static_cast<int32_t>(lb.nodes.size()) : 0;
    v.num_top = lb.mode == kPickBvh ? static_cast<int32_t>(lb.top.size()) : 0;
    v.num_groups = grouped ? desc.num_light_groups : 0;
    v.env_group = grouped ? desc.env_light_group : -1;
    // The background goes with the environment (a Lights update carries both).
    v.bg_mode = desc.background_mode;
    v.bg_color = make_float3(desc.background_color[0], desc.background_color[1], desc.background_color[2]);
    for (int k = 0; k < kMaxTopCandidates; ++k)
        v.top[k] = (lb.mode == kPickBvh && k < static_cast<int>(lb.top.size())) ? lb.top[static_cast<size_t>(k)] : 0;
    return PE_OK;
}

int prepare_shading(pe_device* dev, const pe_scene_desc& desc, pe_scene& sc) {
    sc.view = SceneView{};
    int rc = prepare_materials(dev, desc, sc);
    if (rc != PE_OK) return rc;
    rc = prepare_lights(dev, desc, sc);
    if (rc != PE_OK) return rc;
    // Node uids of the shapes (the object-ID pass): 0 for untagged shapes.
    std::vector<uint64_t> uids(static_cast<size_t>(std::max(1, desc.num_meshes)), 0);
    if (desc.node_uids)
        for (int i = 0; i < desc.num_meshes; ++i) uids[static_cast<size_t>(i)] = desc.node_uids[i];
    if (!sc.node_uids.upload(uids.data(), uids.size() * sizeof(uint64_t), dev->stream)) return PE_ERR_OOM;
    const hipError_t eu = hipStreamSynchronize(dev->stream);
    if (eu != hipSuccess) {
        set_error(std::string("node uid upload: ") + hipGetErrorString(eu));
        return PE_ERR_HIP;
    }
    // Light links: six words per light and the uid lists (a light without links: All / All).
    const size_t nl = static_cast<size_t>(std::max(0, desc.num_lights));
    std::vector<uint32_t> info(std::max<size_t>(1, 6 * nl), 0u);
    std::vector<uint64_t> luids(1, 0ull);
    bool links = false, shadow_links = false;
    if (desc.link_modes) {
        luids.clear();
        for (size_t i = 0; i < nl; ++i) {
            const int32_t mode = desc.link_modes[2 * i], smode = desc.link_modes[2 * i + 1];
            if (mode < PE_LINK_ALL || mode > PE_LINK_EXCLUDE || smode < PE_LINK_ALL || smode > PE_LINK_EXCLUDE) {
                set_error("light " + std::to_string(i) + ": unknown light-link mode");
                return PE_ERR_INVALID;
            }
            const uint32_t* o = desc.link_offsets + 4 * i;
            uint32_t* w = info.data() + 6 * i;
            w[0] = static_cast<uint32_t>(mode);
            w[1] = static_cast<uint32_t>(smode);
            w[2] = static_cast<uint32_t>(luids.size());
            w[3] = o[1];
            luids.insert(luids.end(), desc.link_uids + o[0], desc.link_uids + o[0] + o[1]);
            w[4] = static_cast<uint32_t>(luids.size());
            w[5] = o[3];
            luids.insert(luids.end(), desc.link_uids + o[2], desc.link_uids + o[2] + o[3]);
            links = links || mode != PE_LINK_ALL;
            shadow_links = shadow_links || smode != PE_LINK_ALL;
        }
        if (luids.empty()) luids.push_back(0ull);
        if (shadow_links) {
            // Shadow links held 99.70% per-sample agreement with the oracle on the GPU against the 99.8% bar of
            // the other transport scenes; until that gap is explained they are refused rather than rendered.
            set_error("light shadow links are not implemented on the HIP device");
            return PE_ERR_UNSUPPORTED;
        }
    }
    if (!sc.link_info.upload(info.data(), info.size() * sizeof(uint32_t), dev->stream) ||
        !sc.link_uids.upload(luids.data(), luids.size() * sizeof(uint64_t), dev->stream))
        return PE_ERR_OOM;
    const hipError_t el = hipStreamSynchronize(dev->stream);
    if (el != hipSuccess) {
        set_error(std::string("light link upload: ") + hipGetErrorString(el));
        return PE_ERR_HIP;
    }
    SceneView& v = sc.view;
    v.node_uids = sc.node_uids.as<uint64_t>();
    v.link_info = sc.link_info.as<uint32_t>();
    v.link_uids = sc.link_uids.as<uint64_t>();
    v.has_links = links ? 1 : 0;
    v.has_shadow_links = shadow_links ? 1 : 0;
    // Linked scenes shade in the transport variant; shadow links pass blockers like the transmitting surfaces.
    if (links || shadow_links) v.transport = 1;
    if (shadow_links) v.has_transmitters = 1;
    v.nodes = sc.nodes.as<BvhNode4>();
    v.num_nodes = static_cast<uint32_t>(sc.num_nodes);
    v.tris = sc.tris.as<TriAccel>();
    v.shade = sc.tri_shade.as<TriShade>();
    v.pos = sc.positions.as<float4>();
    v.nrm = sc.normals.as<float4>();
    v.uv = sc.uvs.as<float2>();
    return PE_OK;
}

}  // namespace pe

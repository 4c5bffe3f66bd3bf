// oracle_integrator.cpp — restated light sources, light selection, surface interaction, ray spawning,
// PathTracer::Li, tracePixel and the per-pixel sample loop of Renderer::renderSamples (test
// infrastructure only). Scope: what the box/icosphere and product-sphere workloads use — Disney and Lambert
// surfaces, area/triangle/environment lights (with sun), exhaustive light selection, one-sample NEE + BSDF
// MIS with the power heuristic, Russian roulette, the NaN guard and Film::addSample.
#include "oracle.h"
#include "oracle_scene.h"
#include "oracle_shading.h"

#include <algorithm>
#include <atomic>
#include <limits>
#include <chrono>
#include <cstring>
#include <stdexcept>
#include <thread>

namespace oracle {

namespace {

V3 c3(const float* v) { return V3(v[0], v[1], v[2]); }

// ------------------------------------------------------------------------------ float_bits.h / ray_spawn.h
float nextFloatUp(float v) {  // float_bits.h:30-37
    if (!finiteFloat(v) && v > 0.0f) return v;
    if (v == -0.0f) v = 0.0f;
    uint32_t ui;
    std::memcpy(&ui, &v, 4);
    if (v >= 0.0f) ++ui;
    else --ui;
    std::memcpy(&v, &ui, 4);
    return v;
}
float nextFloatDown(float v) {  // float_bits.h:40-47
    if (!finiteFloat(v) && v < 0.0f) return v;
    if (v == 0.0f) v = -0.0f;
    uint32_t ui;
    std::memcpy(&ui, &v, 4);
    if (v > 0.0f) --ui;
    else ++ui;
    std::memcpy(&v, &ui, 4);
    return v;
}
float floatGamma(int n) {  // float_bits.h:51-55
    const float eps = 5.9604645e-8f;
    return (static_cast<float>(n) * eps) / (1.0f - static_cast<float>(n) * eps);
}

struct SI {  // geometry/surface_interaction.h:16-54
    V3 point, normal, ng, tangent;
    float uvx = 0, uvy = 0, t = -1.0f;
    int material = -1, tri = -1, emitter = -1;
    bool frontFace = true, cameraInvisible = false, hasSmoothPoint = false;
    V3 pError, pSmooth;
};

V3 offsetRayOrigin(V3 p, V3 pError, V3 n, V3 w) {  // ray_spawn.h:25-36
    const float d = dot(vabs(n), pError);
    V3 offset = n * d;
    if (dot(w, n) < 0.0f) offset = -offset;
    V3 po = p + offset;
    for (int i = 0; i < 3; ++i) {
        if (offset[i] > 0.0f) po[i] = nextFloatUp(po[i]);
        else if (offset[i] < 0.0f) po[i] = nextFloatDown(po[i]);
    }
    return po;
}
V3 spawnOrigin(const SI& si, V3 dir) {  // ray_spawn.h:40-43
    const V3 n = lengthSq(si.ng) > 0.0f ? si.ng : si.normal;
    return offsetRayOrigin(si.point, si.pError, n, dir);
}
V3 shadowOrigin(const SI& si, V3 toLight) {  // ray_spawn.h:52-68
    const V3 n = lengthSq(si.ng) > 0.0f ? si.ng : si.normal;
    V3 p = si.point;
    if (si.hasSmoothPoint) {
        const V3 shadingOut = si.frontFace ? si.normal : -si.normal;
        const V3 ngOut = dot(n, shadingOut) >= 0.0f ? n : -n;
        const float NL = dot(shadingOut, toLight);
        if (NL > 0.0f) {
            constexpr float kCutoff = 0.1f;
            const float NgL = dot(ngOut, toLight);
            const float amount = NL < kCutoff ? clampf(2.0f - (NgL + NL) / kCutoff, 0.0f, 1.0f)
                                              : clampf(1.0f - NgL / kCutoff, 0.0f, 1.0f);
            if (amount > 0.0f) p += (si.pSmooth - si.point) * amount;
        }
    }
    return offsetRayOrigin(p, si.pError, n, toLight);
}
Ray spawnRay(const SI& si, V3 dir) {  // ray_spawn.h:71-73
    Ray r;
    r.o = spawnOrigin(si, dir);
    r.d = dir;
    r.tMin = 0.0f;
    r.tMax = 1e30f;
    return r;
}
bool spawnRayTo(const SI& si, V3 target, Ray& out) {  // ray_spawn.h:78-87
    const V3 toTarget = target - si.point;
    const float len = length(toTarget);
    if (!(len > 0.0f)) return false;
    const V3 origin = shadowOrigin(si, toTarget / len);
    const V3 d = target - origin;
    const float dist = length(d);
    const float slack = std::max(dist * 1e-4f, 2e-6f * maxComp(vabs(target)));
    if (!(dist > slack)) return false;
    out.o = origin;
    out.d = d / dist;
    out.tMin = 0.0f;
    out.tMax = dist - slack;
    return true;
}

// triangle.cpp:54-133 (fillHit) for the winning triangle, with (t, u, v) of the closest-hit test.
void fillHit(const Scene& s, const Ray& ray, const Hit& h, SI& si) {
    const TriRec& tr = s.tris[static_cast<size_t>(h.tri)];
    const V3 v0 = s.pos[tr.v[0]], v1 = s.pos[tr.v[1]], v2 = s.pos[tr.v[2]];
    const V3 n0 = tr.hasNormals ? s.nrm[tr.v[0]] : V3(0.0f), n1 = tr.hasNormals ? s.nrm[tr.v[1]] : V3(0.0f),
             n2 = tr.hasNormals ? s.nrm[tr.v[2]] : V3(0.0f);
    const float u = h.u, v = h.v;
    const V3 edge1 = v1 - v0, edge2 = v2 - v0;
    si.t = h.t;
    const float w = 1.0f - u - v;
    const V3 a0 = v0 * w, a1 = v1 * u, a2 = v2 * v;
    si.point = a0 + a1 + a2;
    const float iso = floatGamma(32) * std::max({maxComp(vabs(v0)), maxComp(vabs(v1)), maxComp(vabs(v2)), 1e-20f});
    si.pError = (vabs(a0) + vabs(a1) + vabs(a2)) * floatGamma(16) + V3(iso);
    const V3 geomNormal = normalized(cross(edge1, edge2));
    const bool smooth = lengthSq(n0) > 0.0f && lengthSq(n1) > 0.0f && lengthSq(n2) > 0.0f;
    const V3 shadingNormal = smooth ? normalized(n0 * w + n1 * u + n2 * v) : geomNormal;
    si.ng = geomNormal;
    si.frontFace = dot(ray.d, shadingNormal) < 0.0f;  // surface_interaction.h:48-52
    si.normal = si.frontFace ? shadingNormal : -shadingNormal;
    if (lengthSq(si.ng) == 0.0f) si.ng = shadingNormal;
    si.hasSmoothPoint = false;
    if (smooth) {
        const V3 p = si.point;
        const V3 m0 = normalized(n0), m1 = normalized(n1), m2 = normalized(n2);
        const V3 lift = m0 * (w * std::max(0.0f, dot(v0 - p, m0))) + m1 * (u * std::max(0.0f, dot(v1 - p, m1))) +
                        m2 * (v * std::max(0.0f, dot(v2 - p, m2)));
        if (lengthSq(lift) > 0.0f) {
            si.pSmooth = p + lift;
            si.hasSmoothPoint = true;
        }
    }
    const float* uv = s.uv.data();
    const float u0x = tr.hasUv ? uv[2 * tr.v[0]] : 0.0f, u0y = tr.hasUv ? uv[2 * tr.v[0] + 1] : 0.0f;
    const float u1x = tr.hasUv ? uv[2 * tr.v[1]] : 0.0f, u1y = tr.hasUv ? uv[2 * tr.v[1] + 1] : 0.0f;
    const float u2x = tr.hasUv ? uv[2 * tr.v[2]] : 0.0f, u2y = tr.hasUv ? uv[2 * tr.v[2] + 1] : 0.0f;
    si.uvx = u0x * w + u1x * u + u2x * v;
    si.uvy = u0y * w + u1y * u + u2y * v;
    const float d1x = u1x - u0x, d1y = u1y - u0y, d2x = u2x - u0x, d2y = u2y - u0y;
    const float r = d1x * d2y - d2x * d1y;
    V3 tangent;
    if (std::abs(r) < 1e-6f) {
        const V3 n = si.normal;
        const float sign = std::copysign(1.0f, n.z);
        const float a = -1.0f / (sign + n.z);
        const float b = n.x * n.y * a;
        tangent = V3(1.0f + sign * n.x * n.x * a, sign * b, -sign * n.x);
    } else {
        const float invR = 1.0f / r;
        tangent = (edge1 * d2y - edge2 * d1y) * invR;
    }
    si.tangent = normalized(tangent);
    si.material = tr.material;
    si.tri = h.tri;
    si.emitter = tr.emitter;
    si.cameraInvisible = tr.cameraInvisible;
}

// ----------------------------------------------------------------------------------- lights
struct LightSample {  // light.h:35-42
    V3 wi, Li;
    float pdf = 0.0f, distance = 0.0f;
    bool valid() const { return pdf > 0.0f && !isBlack(Li); }
};

float receiverCosine(V3 n, V3 toLight) {  // light.h:95-98
    const float c = dot(n, toLight);
    return c > 0.05f ? c : 0.05f;
}
float areaSolidAnglePdf(float distSq, float cosAbs, float area) {  // area_light.cpp:18-21
    if (!(cosAbs > 1e-6f) || !(area > 0.0f) || !(distSq > 0.0f)) return 0.0f;
    return distSq / (area * cosAbs);
}
constexpr float kMinSolidAngle = 1e-3f;
float safeAsin(float x) { return std::asin(clampf(x, -1.0f, 1.0f)); }
float angleBetween(V3 a, V3 b) {  // area_light.cpp:30-33
    if (dot(a, b) < 0.0f) return PI - 2.0f * safeAsin(length(a + b) * 0.5f);
    return 2.0f * safeAsin(length(b - a) * 0.5f);
}
struct SphericalRect {  // area_light.cpp:39-45
    Frame frame;
    float x0 = 0, y0 = 0, z0 = 0, x1 = 0, y1 = 0, b0 = 0, b1 = 0, g2g3 = 0, solidAngle = 0.0f;
};
SphericalRect setupSphericalRect(V3 ref, V3 corner, V3 ex, V3 ey) {  // area_light.cpp:47-72
    SphericalRect r;
    const float exl = length(ex), eyl = length(ey);
    const V3 x = ex / exl, y = ey / eyl;
    V3 z = cross(x, y);
    const V3 d = corner - ref;
    r.z0 = dot(d, z);
    if (r.z0 > 0.0f) {
        z = -z;
        r.z0 = -r.z0;
    }
    r.frame = Frame(x, y, z);
    r.x0 = dot(d, x);
    r.y0 = dot(d, y);
    r.x1 = r.x0 + exl;
    r.y1 = r.y0 + eyl;
    if (r.z0 > -1e-7f) return r;
    const V3 v00(r.x0, r.y0, r.z0), v01(r.x0, r.y1, r.z0), v10(r.x1, r.y0, r.z0), v11(r.x1, r.y1, r.z0);
    const V3 n0 = normalized(cross(v00, v10)), n1 = normalized(cross(v10, v11));
    const V3 n2 = normalized(cross(v11, v01)), n3 = normalized(cross(v01, v00));
    const float g0 = angleBetween(-n0, n1), g1 = angleBetween(-n1, n2);
    const float g2 = angleBetween(-n2, n3), g3 = angleBetween(-n3, n0);
    r.solidAngle = g0 + g1 + g2 + g3 - TWO_PI;
    r.b0 = n0.z;
    r.b1 = n2.z;
    r.g2g3 = g2 + g3;
    return r;
}
V3 sampleSphericalRect(const SphericalRect& r, float ux, float uy) {  // area_light.cpp:75-92
    const float au = ux * r.solidAngle - r.g2g3;
    const float fu = (std::cos(au) * r.b0 - r.b1) / std::sin(au);
    float cu = std::copysign(1.0f, fu) / std::sqrt(fu * fu + r.b0 * r.b0);
    cu = clampf(cu, -0.9999999f, 0.9999999f);
    float xu = -(cu * r.z0) / std::sqrt(std::max(0.0f, 1.0f - cu * cu));
    xu = clampf(xu, r.x0, r.x1);
    const float dd = std::sqrt(xu * xu + r.z0 * r.z0);
    const float h0 = r.y0 / std::sqrt(dd * dd + r.y0 * r.y0);
    const float h1 = r.y1 / std::sqrt(dd * dd + r.y1 * r.y1);
    const float hv = h0 + uy * (h1 - h0);
    const float hv2 = hv * hv;
    const float yv = hv2 < 1.0f - 1e-6f ? hv * dd / std::sqrt(1.0f - hv2) : r.y1;
    return r.frame.toWorld(V3(xu, clampf(yv, r.y0, r.y1), r.z0));
}

struct LightRec {
    int type;
    V3 p0, p1, p2, radiance;
    // AreaLight (area_light.h:20-27): position, u, v, normal, area, rectangle flag
    V3 normal;
    float area = 0.0f;
    bool rectangle = false;
};

LightRec makeLight(const pe_light_desc& d) {
    LightRec L;
    L.type = d.type;
    L.p0 = c3(d.p0);
    L.p1 = c3(d.p1);
    L.p2 = c3(d.p2);
    L.radiance = c3(d.radiance);
    if (d.type == PE_LIGHT_AREA) {
        const V3 n = cross(L.p1, L.p2);
        L.area = length(n);
        L.normal = normalized(n);
        const float lu = length(L.p1), lv = length(L.p2);
        L.rectangle = lu > 0.0f && lv > 0.0f && std::abs(dot(L.p1, L.p2)) <= 1e-4f * lu * lv;
    } else {  // TriangleLight (area_light.cpp:195-200)
        const V3 n = cross(L.p1 - L.p0, L.p2 - L.p0);
        const float len = length(n);
        L.area = 0.5f * len;
        L.normal = len > 0.0f ? n / len : V3(0, 1, 0);
    }
    return L;
}

float lightImportance(const LightRec& L, V3 p, V3 n) {  // area_light.cpp:100-112 and :202-212
    const V3 c = L.type == PE_LIGHT_AREA ? L.p0 + (L.p1 + L.p2) * 0.5f : (L.p0 + L.p1 + L.p2) * (1.0f / 3.0f);
    const V3 d = c - p;
    const float dist2 = lengthSq(d);
    const float lum = std::max(0.0f, luminance(L.radiance));
    if (!(dist2 > 0.0f)) return lum * TWO_PI;
    const V3 w = d / std::sqrt(dist2);
    const float cosL = std::max(std::abs(dot(w, L.normal)), 0.2f);
    const float omega = std::min(L.area * cosL / dist2, TWO_PI);
    return lum * omega * receiverCosine(n, w);
}

LightSample lightSampleLi(const LightRec& L, V3 ref, float ux, float uy) {
    LightSample ls;
    if (L.type == PE_LIGHT_AREA) {  // area_light.cpp:114-157
        if (L.rectangle) {
            const SphericalRect r = setupSphericalRect(ref, L.p0, L.p1, L.p2);
            if (r.solidAngle >= kMinSolidAngle) {
                const V3 toLight = sampleSphericalRect(r, ux, uy);
                const float dist = length(toLight);
                if (!(dist > 0.0f)) return ls;
                ls.distance = dist;
                ls.wi = toLight / dist;
                ls.Li = L.radiance;
                ls.pdf = 1.0f / r.solidAngle;
                return ls;
            }
        }
        const V3 pLight = L.p0 + L.p1 * ux + L.p2 * uy;
        const V3 toLight = pLight - ref;
        const float distSq = lengthSq(toLight);
        ls.distance = std::sqrt(distSq);
        if (ls.distance > 0.0f) {
            ls.wi = toLight / ls.distance;
            const float pdf = areaSolidAnglePdf(distSq, std::abs(dot(ls.wi, L.normal)), L.area);
            if (pdf > 0.0f) {
                ls.Li = L.radiance;
                ls.pdf = pdf;
            } else {
                ls.Li = V3(0.0f);
                ls.pdf = 0.0f;
            }
        }
        return ls;
    }
    // TriangleLight::sampleLi (area_light.cpp:235-252)
    if (!(L.area > 0.0f) || isBlack(L.radiance)) return ls;
    float b1, b2;
    uniformSampleTriangle(ux, uy, b1, b2);
    const V3 pLight = L.p0 * (1.0f - b1 - b2) + L.p1 * b1 + L.p2 * b2;
    const V3 toLight = pLight - ref;
    const float distSq = lengthSq(toLight);
    ls.distance = std::sqrt(distSq);
    if (!(ls.distance > 0.0f)) return ls;
    ls.wi = toLight / ls.distance;
    const float pdf = areaSolidAnglePdf(distSq, std::abs(dot(ls.wi, L.normal)), L.area);
    if (!(pdf > 0.0f)) return ls;
    ls.Li = L.radiance;
    ls.pdf = pdf;
    return ls;
}

float lightPdfLiHit(const LightRec& L, V3 ref, V3 hitPoint) {
    if (L.type == PE_LIGHT_AREA) {  // AreaLight::pdfLi (area_light.cpp:163-187)
        const V3 d = hitPoint - L.p0;
        const float scale = std::max(length(L.p1) + length(L.p2), 1.0f);
        if (std::abs(dot(d, L.normal)) > 1e-3f * scale) return 0.0f;
        const float uu = dot(L.p1, L.p1), vv = dot(L.p2, L.p2), uv = dot(L.p1, L.p2);
        const float du = dot(d, L.p1), dv = dot(d, L.p2);
        const float det = uu * vv - uv * uv;
        if (!(std::abs(det) > 1e-12f * uu * vv)) return 0.0f;
        const float s = (vv * du - uv * dv) / det;
        const float t = (uu * dv - uv * du) / det;
        const float eps = 1e-3f;
        if (s < -eps || t < -eps || s > 1.0f + eps || t > 1.0f + eps) return 0.0f;
        if (L.rectangle) {
            const float omega = setupSphericalRect(ref, L.p0, L.p1, L.p2).solidAngle;
            if (omega >= kMinSolidAngle) return 1.0f / omega;
        }
        const V3 toLight = hitPoint - ref;
        const float distSq = lengthSq(toLight);
        if (distSq <= 0.0f) return 0.0f;
        const V3 wi = toLight / std::sqrt(distSq);
        return areaSolidAnglePdf(distSq, std::abs(dot(wi, L.normal)), L.area);
    }
    const V3 toLight = hitPoint - ref;  // TriangleLight::pdfLiHit (area_light.cpp:226-231)
    const float distSq = lengthSq(toLight);
    if (!(distSq > 0.0f)) return 0.0f;
    return areaSolidAnglePdf(distSq, std::abs(dot(toLight / std::sqrt(distSq), L.normal)), L.area);
}

// ------------------------------------------------------------------------------ environment_light.cpp
struct Distribution {  // environment_light.h:108-115
    float funcInt = 0.0f;
    std::vector<float> func, marginalCdf, condCdf;
    bool valid() const { return funcInt > 0.0f && !marginalCdf.empty(); }
};

struct Env {
    int w = 0, h = 0;
    const float* rgb = nullptr;
    float rotation = 0.0f, intensity = 1.0f, meanLuminance = 0.0f;
    V3 sunDir{0, 1, 0}, sunRadiance;
    float sunCosMax = 1.0f, sunPdfWeight = 0.0f;
    Distribution full, mis;

    V3 pixel(int x, int y) const { return c3(rgb + 3 * (static_cast<size_t>(y) * w + x)); }

    void buildCdf(Distribution& d) const {  // environment_light.cpp:149-172
        d.marginalCdf.assign(static_cast<size_t>(h) + 1, 0.0f);
        d.condCdf.assign(static_cast<size_t>(h) * static_cast<size_t>(w + 1), 0.0f);
        for (int y = 0; y < h; ++y) {
            float* row = d.condCdf.data() + static_cast<size_t>(y) * static_cast<size_t>(w + 1);
            row[0] = 0.0f;
            for (int x = 0; x < w; ++x) row[x + 1] = row[x] + d.func[static_cast<size_t>(y) * w + x];
            d.marginalCdf[static_cast<size_t>(y) + 1] = d.marginalCdf[static_cast<size_t>(y)] + row[w];
        }
        d.funcInt = d.marginalCdf[static_cast<size_t>(h)];
        if (!(d.funcInt > 0.0f)) {
            d.funcInt = 0.0f;
            d.marginalCdf.clear();
            return;
        }
        const float inv = 1.0f / d.funcInt;
        for (float& c : d.marginalCdf) c *= inv;
        for (int y = 0; y < h; ++y) {
            float* row = d.condCdf.data() + static_cast<size_t>(y) * static_cast<size_t>(w + 1);
            const float rowSum = row[w];
            if (rowSum <= 0.0f) continue;
            const float invRow = 1.0f / rowSum;
            for (int x = 0; x <= w; ++x) row[x] *= invRow;
        }
    }

    void build(bool misCompensation) {  // environment_light.cpp:90-147
        if (!rgb || w <= 0 || h <= 0) return;
        full.func.assign(static_cast<size_t>(w) * h, 0.0f);
        std::vector<float> sinRow(static_cast<size_t>(h));
        double weighted = 0.0, weights = 0.0;
        for (int y = 0; y < h; ++y) {
            const float sinTheta = std::sin(PI * (static_cast<float>(y) + 0.5f) / static_cast<float>(h));
            sinRow[static_cast<size_t>(y)] = sinTheta;
            for (int x = 0; x < w; ++x) {
                float L = luminance(pixel(x, y));
                if (!(L > 0.0f)) L = 0.0f;
                full.func[static_cast<size_t>(y) * w + x] = L * sinTheta;
                weighted += static_cast<double>(L) * sinTheta;
                weights += sinTheta;
            }
        }
        buildCdf(full);
        if (!full.valid()) return;
        const float meanL = weights > 0.0 ? static_cast<float>(weighted / weights) : 0.0f;
        mis.func.resize(full.func.size());
        double misTotal = 0.0, fullTotal = 0.0;
        for (int y = 0; y < h; ++y) {
            const float s = sinRow[static_cast<size_t>(y)];
            for (int x = 0; x < w; ++x) {
                const size_t i = static_cast<size_t>(y) * w + x;
                float L = luminance(pixel(x, y));
                if (!(L > 0.0f)) L = 0.0f;
                const float excess = L - meanL;
                mis.func[i] = excess > 1e-4f * meanL ? excess * s : 0.0f;
                misTotal += mis.func[i];
                fullTotal += full.func[i];
            }
        }
        bool comp = misCompensation && misTotal > 1e-3 * fullTotal;
        if (comp) {
            buildCdf(mis);
            comp = mis.valid();
        }
        if (!comp) mis = full;
    }

    V3 power() const {  // environment_light.cpp:308-323
        if (!rgb || w <= 0) return V3(0.5f) * intensity * 4.0f * PI;
        V3 sum(0.0f);
        float wsum = 0.0f;
        for (int y = 0; y < h; y += 4) {
            const float s = std::sin(PI * (static_cast<float>(y) + 0.5f) / static_cast<float>(h));
            for (int x = 0; x < w; x += 4) {
                sum += pixel(x, y) * s;
                wsum += s;
            }
        }
        return (wsum > 0.0f ? sum / wsum : sum) * intensity * 4.0f * PI;
    }

    void setSun(V3 direction, float cosMax, V3 radiance) {  // environment_light.cpp:168-182
        sunDir = normalized(direction);
        sunCosMax = clampf(cosMax, -0.999f, 0.9999999f);
        sunRadiance = radiance * intensity;
        const float omega = TWO_PI * (1.0f - sunCosMax);
        const float sunPower = std::max(0.0f, luminance(sunRadiance)) * omega;
        const float skyPower = 4.0f * PI * meanLuminance;
        if (!(sunPower > 0.0f)) {
            sunPdfWeight = 0.0f;
            return;
        }
        sunPdfWeight = clampf(sunPower / (sunPower + skyPower), 0.15f, 0.85f);
        if (!full.valid()) sunPdfWeight = 1.0f;
    }

    // directionToEquirect (environment_light.h:22-29)
    void toEquirect(V3 dir, float& u, float& v) const {
        const V3 d = normalized(dir);
        const float theta = std::acos(clampf(d.y, -1.0f, 1.0f));
        float phi = std::atan2(-d.z, d.x) + PI + rotation;
        phi = std::fmod(phi, TWO_PI);
        if (phi < 0.0f) phi += TWO_PI;
        u = phi / TWO_PI;
        v = theta / PI;
    }
    V3 equirectDirection(float u, float v) const {  // environment_light.h:33-39
        const float theta = v * PI;
        const float a = u * TWO_PI - PI - rotation;
        const float st = std::sin(theta);
        return V3(st * std::cos(a), std::cos(theta), -st * std::sin(a));
    }
    V3 sampleEquirect(float u, float v) const {  // environment_light.cpp:43-62
        u -= std::floor(u);
        v = clampf(v, 0.0f, 1.0f);
        const float x = u * static_cast<float>(w) - 0.5f;
        const float y = v * static_cast<float>(h) - 0.5f;
        const int x0 = static_cast<int>(std::floor(x)), y0 = static_cast<int>(std::floor(y));
        const float fx = x - static_cast<float>(x0), fy = y - static_cast<float>(y0);
        auto wrapX = [this](int i) {
            i %= w;
            return i < 0 ? i + w : i;
        };
        auto clampY = [this](int i) { return std::clamp(i, 0, h - 1); };
        const int xa = wrapX(x0), xb = wrapX(x0 + 1), ya = clampY(y0), yb = clampY(y0 + 1);
        const V3 c0 = pixel(xa, ya) * (1.0f - fx) + pixel(xb, ya) * fx;
        const V3 c1 = pixel(xa, yb) * (1.0f - fx) + pixel(xb, yb) * fx;
        return c0 * (1.0f - fy) + c1 * fy;
    }
    V3 evalSky(V3 dir) const {  // environment_light.cpp:301-306
        if (!rgb || w <= 0) return V3(0.5f) * intensity;
        float u, v;
        toEquirect(dir, u, v);
        return sampleEquirect(u, v) * intensity;
    }
    V3 eval(V3 dir) const {  // environment_light.cpp:296-299
        V3 sun(0.0f);
        if (sunPdfWeight > 0.0f && dot(normalized(dir), sunDir) >= sunCosMax) sun = sunRadiance;
        return evalSky(dir) + sun;
    }
    float solidAnglePdf(const Distribution& dist, V3 wi) const {  // environment_light.cpp:266-282
        if (!dist.valid() || w <= 0 || h <= 0) return 0.25f * INV_PI;
        const V3 d = normalized(wi);
        float u, v;
        toEquirect(d, u, v);
        int x = std::min(static_cast<int>(u * static_cast<float>(w)), w - 1);
        int y = std::min(static_cast<int>(v * static_cast<float>(h)), h - 1);
        if (x < 0) x = 0;
        if (y < 0) y = 0;
        const float sinTheta = std::sqrt(std::max(0.0f, 1.0f - d.y * d.y));
        if (sinTheta < 1e-4f) return 0.0f;
        const float f = dist.func[static_cast<size_t>(y) * w + x];
        return (f / dist.funcInt) * static_cast<float>(w * h) / (TWO_PI * PI * sinTheta);
    }
    float pdfLi(V3 wi) const {  // environment_light.cpp:288-294
        if (sunPdfWeight <= 0.0f) return solidAnglePdf(mis, wi);
        float pdf = 0.0f;
        if (dot(normalized(wi), sunDir) >= sunCosMax) pdf += sunPdfWeight / (TWO_PI * (1.0f - sunCosMax));
        if (sunPdfWeight < 1.0f) pdf += (1.0f - sunPdfWeight) * solidAnglePdf(mis, wi);
        return pdf;
    }
    static void sampleCdf(const float* cdf, int n, float u, int& index, float& pdf, float& coord) {  // :22-36
        const float* end = cdf + n + 1;
        const float* it = std::upper_bound(cdf, end, u);
        int i = static_cast<int>(it - cdf) - 1;
        if (i < 0) i = 0;
        if (i >= n) i = n - 1;
        pdf = cdf[i + 1] - cdf[i];
        float du = pdf > 0.0f ? (u - cdf[i]) / pdf : 0.0f;
        du = clampf(du, 0.0f, 1.0f);
        index = i;
        coord = (static_cast<float>(i) + du) / static_cast<float>(n);
    }
    LightSample sampleImage(const Distribution& d, float sx, float sy) const {  // :218-237
        LightSample ls;
        ls.distance = INFINITY_F;
        if (!d.valid() || w <= 0) {
            ls.wi = uniformSampleSphere(sx, sy);
            ls.pdf = 0.25f * INV_PI;
            ls.Li = eval(ls.wi);
            return ls;
        }
        int ri, ci;
        float rp, rc, cp, cc;
        sampleCdf(d.marginalCdf.data(), h, sy, ri, rp, rc);
        sampleCdf(d.condCdf.data() + static_cast<size_t>(ri) * static_cast<size_t>(w + 1), w, sx, ci, cp, cc);
        const float theta = rc * PI;
        const float sinTheta = std::max(std::sin(theta), 1e-6f);
        ls.wi = normalized(equirectDirection(cc, rc));
        ls.pdf = (rp * cp) * static_cast<float>(w * h) / (TWO_PI * PI * sinTheta);
        ls.Li = eval(ls.wi);
        return ls;
    }
    LightSample sampleLi(float sx, float sy) const {  // :191-216
        LightSample ls;
        ls.distance = INFINITY_F;
        if (sunPdfWeight > 0.0f) {
            if (sx < sunPdfWeight) {
                const float ux = std::min(sx / sunPdfWeight, 0.99999994f);
                const float cosT = 1.0f - ux * (1.0f - sunCosMax);
                const float sinT = std::sqrt(std::max(0.0f, 1.0f - cosT * cosT));
                const float phi = TWO_PI * sy;
                const Frame f(sunDir);
                ls.wi = normalized(f.toWorld(V3(sinT * std::cos(phi), sinT * std::sin(phi), cosT)));
                ls.Li = eval(ls.wi);
                ls.pdf = pdfLi(ls.wi);
                return ls;
            }
            if (sunPdfWeight >= 1.0f) return ls;
            sx = std::min((sx - sunPdfWeight) / (1.0f - sunPdfWeight), 0.99999994f);
        }
        LightSample img = sampleImage(mis, sx, sy);
        if (sunPdfWeight > 0.0f && img.pdf > 0.0f) img.pdf = pdfLi(img.wi);
        return img;
    }
    float importance(V3 n) const {  // :184-189
        float imp = PI * meanLuminance;
        if (sunPdfWeight > 0.0f)
            imp += std::max(0.0f, luminance(sunRadiance)) * TWO_PI * (1.0f - sunCosMax) * receiverCosine(n, sunDir);
        return imp;
    }
};

// ------------------------------------------------------------------------------------ light_bounds.cpp
struct LBox {  // core/math/aabb.h:24-31, empty = [+max, -max]; cwiseMin/Max are std::min/std::max
    V3 lo{std::numeric_limits<float>::max()}, hi{-std::numeric_limits<float>::max()};
    LBox() = default;
    LBox(V3 a, V3 b) : lo(vmin(a, b)), hi(vmax(a, b)) {}   // AABB(a, b)
    void merge(const LBox& b) {
        lo = vmin(lo, b.lo);
        hi = vmax(hi, b.hi);
    }
    void merge(V3 p) {
        lo = vmin(lo, p);
        hi = vmax(hi, p);
    }
    LBox merged(const LBox& b) const { return LBox(vmin(lo, b.lo), vmax(hi, b.hi)); }  // aabb.h merged()
    V3 centroid() const { return (lo + hi) * 0.5f; }
    V3 diagonal() const { return hi - lo; }
    float surfaceArea() const {
        const V3 d = diagonal();
        return 2.0f * (d.x * d.y + d.x * d.z + d.y * d.z);
    }
    int maxExtent() const {
        const V3 d = diagonal();
        if (d.x > d.y && d.x > d.z) return 0;
        if (d.y > d.z) return 1;
        return 2;
    }
};

float safeSqrtF(float x) { return std::sqrt(std::max(0.0f, x)); }
float safeAcosF(float x) { return std::acos(std::clamp(x, -1.0f, 1.0f)); }
float cosSubClamped(float sinA, float cosA, float sinB, float cosB) { return cosA > cosB ? 1.0f : cosA * cosB + sinA * sinB; }
float sinSubClamped(float sinA, float cosA, float sinB, float cosB) { return cosA > cosB ? 0.0f : sinA * cosB - cosA * sinB; }

struct LightBounds {  // light_bounds.h:15-25
    LBox bounds;
    V3 w{0.0f, 0.0f, 1.0f};
    float phi = 0.0f, cosThetaO = 1.0f, cosThetaE = 0.0f;
    bool twoSided = false;

    float importance(V3 p, V3 n) const {  // light_bounds.cpp:39-82
        if (!(phi > 0.0f)) return 0.0f;
        const V3 pc = bounds.centroid();
        const V3 diag = bounds.diagonal();
        const float d2 = std::max(lengthSq(p - pc), 0.25f * lengthSq(diag));
        if (!(d2 > 0.0f)) return phi;
        V3 wi = p - pc;
        const float len = length(wi);
        wi = len > 0.0f ? wi / len : w;
        float cosW = dot(w, wi);
        if (twoSided) cosW = std::abs(cosW);
        const float sinW = safeSqrtF(1.0f - cosW * cosW);
        const float r2 = 0.25f * lengthSq(diag);
        const float dc2 = lengthSq(p - pc);
        const float cosB = dc2 <= r2 ? -1.0f : safeSqrtF(1.0f - r2 / dc2);
        const float sinB = safeSqrtF(1.0f - cosB * cosB);
        const float sinO = safeSqrtF(1.0f - cosThetaO * cosThetaO);
        const float cosX = cosSubClamped(sinW, cosW, sinO, cosThetaO);
        const float sinX = sinSubClamped(sinW, cosW, sinO, cosThetaO);
        const float cosP = cosSubClamped(sinX, cosX, sinB, cosB);
        if (cosP < cosThetaE) return 0.0f;
        float imp = phi * cosP / d2;
        if (lengthSq(n) > 0.0f) {
            const float cosI = std::abs(dot(wi, n));
            const float sinI = safeSqrtF(1.0f - cosI * cosI);
            imp *= cosSubClamped(sinI, cosI, sinB, cosB);
        }
        return std::max(imp, 0.0f);
    }
};

V3 rotateAround(V3 v, V3 axis, float theta) {  // light_bounds.cpp:30-33 (Rodrigues)
    const float c = std::cos(theta), s = std::sin(theta);
    return v * c + cross(axis, v) * s + axis * (dot(axis, v) * (1.0f - c));
}

LightBounds unionBounds(const LightBounds& a, const LightBounds& b) {  // light_bounds.cpp:84-109
    if (!(a.phi > 0.0f)) return b;
    if (!(b.phi > 0.0f)) return a;
    LightBounds u;
    u.bounds = a.bounds.merged(b.bounds);
    u.phi = a.phi + b.phi;
    u.cosThetaE = std::min(a.cosThetaE, b.cosThetaE);
    u.twoSided = a.twoSided || b.twoSided;
    const float thetaA = safeAcosF(a.cosThetaO), thetaB = safeAcosF(b.cosThetaO);
    const float thetaD = safeAcosF(dot(a.w, b.w));
    if (std::min(thetaD + thetaB, PI) <= thetaA) {
        u.w = a.w;
        u.cosThetaO = a.cosThetaO;
        return u;
    }
    if (std::min(thetaD + thetaA, PI) <= thetaB) {
        u.w = b.w;
        u.cosThetaO = b.cosThetaO;
        return u;
    }
    const float thetaO = 0.5f * (thetaA + thetaD + thetaB);
    const V3 axis = cross(a.w, b.w);
    if (thetaO >= PI || lengthSq(axis) < 1e-12f) {
        u.w = a.w;
        u.cosThetaO = -1.0f;
        return u;
    }
    u.w = normalized(rotateAround(a.w, normalized(axis), thetaO - thetaA));
    u.cosThetaO = std::cos(thetaO);
    return u;
}

// AreaLight::bounds (area_light.cpp:260-273) and TriangleLight::bounds (:219-228); false = unbounded.
bool lightBoundsOf(const LightRec& L, LightBounds& out) {
    if (!(L.area > 0.0f) || isBlack(L.radiance)) return false;
    if (L.type == PE_LIGHT_AREA) {
        LBox box(L.p0, L.p0);
        box.merge(L.p0 + L.p1);
        box.merge(L.p0 + L.p2);
        box.merge(L.p0 + L.p1 + L.p2);
        out.bounds = box;
    } else {
        out.bounds = LBox(vmin(vmin(L.p0, L.p1), L.p2), vmax(vmax(L.p0, L.p1), L.p2));
    }
    out.w = L.normal;
    out.phi = std::max(0.0f, luminance(L.radiance)) * L.area * PI;
    out.cosThetaO = 1.0f;
    out.cosThetaE = 0.0f;
    out.twoSided = true;
    return true;
}

// ------------------------------------------------------------------------------------ light_sampler.cpp
constexpr float kUniformShare = 0.15f;
constexpr int kMaxExhaustive = 33;
constexpr int kMaxInfinite = 16;
constexpr int kLightBuckets = 12;
constexpr int kLightMaxDepth = 63;

float groupCost(const LightBounds& b, const LBox& parent, int dim) {  // light_sampler.cpp:24-36
    const float thetaO = std::acos(std::clamp(b.cosThetaO, -1.0f, 1.0f));
    const float thetaE = std::acos(std::clamp(b.cosThetaE, -1.0f, 1.0f));
    const float thetaW = std::min(thetaO + thetaE, PI);
    const float sinO = std::sqrt(std::max(0.0f, 1.0f - b.cosThetaO * b.cosThetaO));
    const float mOmega = TWO_PI * (1.0f - b.cosThetaO) +
                         0.5f * PI * (2.0f * thetaW * sinO - std::cos(thetaO - 2.0f * thetaW) - 2.0f * thetaO * sinO +
                                      b.cosThetaO);
    const V3 d = parent.diagonal();
    const float maxD = std::max({d.x, d.y, d.z});
    const float kr = d[dim] > 0.0f ? maxD / d[dim] : 1.0f;
    return b.phi * mOmega * kr * b.bounds.surfaceArea();
}

struct LightNode {
    LightBounds lb;
    int child1 = -1;  // interior: second child (the first is the next node)
    int light = -1;   // leaf: candidate index
};

struct Integrator {
    std::vector<LightRec> lights;
    bool hasEnv = false;
    Env env;
    int count = 0;
    // Light BVH (more than kMaxExhaustive candidates), LightSampler::build (light_sampler.cpp:42-77)
    bool bvh = false;
    std::vector<LightNode> nodes;
    std::vector<int> infinite;
    std::vector<int> topSlot;
    std::vector<uint64_t> trail;
    std::vector<int> depthOf;

    int explicitLights() const { return static_cast<int>(lights.size()); }

    float candidateImportance(int i, V3 p, V3 n) const;

    struct Item {
        int index;
        LightBounds lb;
    };

    int buildNode(std::vector<Item>& items, int begin, int end, uint64_t tr, int depth) {  // :79-148
        const int ni = static_cast<int>(nodes.size());
        nodes.emplace_back();
        if (end - begin == 1) {
            const Item& it = items[static_cast<size_t>(begin)];
            nodes[static_cast<size_t>(ni)].lb = it.lb;
            nodes[static_cast<size_t>(ni)].light = it.index;
            trail[static_cast<size_t>(it.index)] = tr;
            depthOf[static_cast<size_t>(it.index)] = depth;
            return ni;
        }
        LBox bounds, centroids;
        for (int i = begin; i < end; ++i) {
            bounds.merge(items[static_cast<size_t>(i)].lb.bounds);
            centroids.merge(items[static_cast<size_t>(i)].lb.bounds.centroid());
        }
        float bestCost = std::numeric_limits<float>::infinity();
        int bestDim = -1, bestBucket = -1;
        if (depth < kLightMaxDepth - 8) {
            for (int dim = 0; dim < 3; ++dim) {
                const float lo = centroids.lo[dim], hi = centroids.hi[dim];
                if (!(hi > lo)) continue;
                LightBounds bucket[kLightBuckets];
                for (int i = begin; i < end; ++i) {
                    const LightBounds& lb = items[static_cast<size_t>(i)].lb;
                    int b = static_cast<int>(kLightBuckets * (lb.bounds.centroid()[dim] - lo) / (hi - lo));
                    b = std::clamp(b, 0, kLightBuckets - 1);
                    bucket[b] = unionBounds(bucket[b], lb);
                }
                for (int split = 0; split < kLightBuckets - 1; ++split) {
                    LightBounds below, above;
                    for (int b = 0; b <= split; ++b) below = unionBounds(below, bucket[b]);
                    for (int b = split + 1; b < kLightBuckets; ++b) above = unionBounds(above, bucket[b]);
                    if (!(below.phi > 0.0f) || !(above.phi > 0.0f)) continue;
                    const float cost = groupCost(below, bounds, dim) + groupCost(above, bounds, dim);
                    if (cost > 0.0f && cost < bestCost) {
                        bestCost = cost;
                        bestDim = dim;
                        bestBucket = split;
                    }
                }
            }
        }
        int mid = -1;
        if (bestDim >= 0) {
            const float lo = centroids.lo[bestDim], hi = centroids.hi[bestDim];
            auto it = std::partition(items.begin() + begin, items.begin() + end, [&](const Item& item) {
                int b = static_cast<int>(kLightBuckets * (item.lb.bounds.centroid()[bestDim] - lo) / (hi - lo));
                return std::clamp(b, 0, kLightBuckets - 1) <= bestBucket;
            });
            mid = static_cast<int>(it - items.begin());
        }
        if (mid <= begin || mid >= end) {
            mid = (begin + end) / 2;
            const int dim = centroids.maxExtent();
            std::nth_element(items.begin() + begin, items.begin() + mid, items.begin() + end,
                             [dim](const Item& a, const Item& b) {
                                 return a.lb.bounds.centroid()[dim] < b.lb.bounds.centroid()[dim];
                             });
        }
        const int c0 = buildNode(items, begin, mid, tr, depth + 1);
        const int c1 = buildNode(items, mid, end, tr | (uint64_t{1} << depth), depth + 1);
        LightNode& nd = nodes[static_cast<size_t>(ni)];
        nd.lb = unionBounds(nodes[static_cast<size_t>(c0)].lb, nodes[static_cast<size_t>(c1)].lb);
        nd.child1 = c1;
        return ni;
    }

    void buildSampler() {  // LightSampler::build (light_sampler.cpp:42-77)
        topSlot.assign(static_cast<size_t>(count), -1);
        trail.assign(static_cast<size_t>(count), 0);
        depthOf.assign(static_cast<size_t>(count), 0);
        if (count <= kMaxExhaustive) return;
        std::vector<Item> items;
        for (int i = 0; i < explicitLights(); ++i) {
            LightBounds lb;
            if (lightBoundsOf(lights[static_cast<size_t>(i)], lb) && lb.phi > 0.0f) items.push_back({i, lb});
            else infinite.push_back(i);
        }
        if (hasEnv) infinite.push_back(explicitLights());
        if (static_cast<int>(infinite.size()) + 1 > kMaxInfinite) {
            infinite.clear();
            return;
        }
        for (size_t k = 0; k < infinite.size(); ++k) topSlot[static_cast<size_t>(infinite[k])] = static_cast<int>(k);
        if (!items.empty()) buildNode(items, 0, static_cast<int>(items.size()), 0, 0);
        bvh = true;
    }

    int bvhSample(V3 p, V3 n, float u, float& pmf) const {  // light_sampler.cpp:242-266
        pmf = 1.0f;
        int node = 0;
        for (;;) {
            const LightNode& nd = nodes[static_cast<size_t>(node)];
            if (nd.light >= 0) return nd.light;
            const int c0 = node + 1, c1 = nd.child1;
            const float i0 = nodes[static_cast<size_t>(c0)].lb.importance(p, n);
            const float i1 = nodes[static_cast<size_t>(c1)].lb.importance(p, n);
            if (!(i0 > 0.0f) && !(i1 > 0.0f)) {
                pmf = 0.0f;
                return -1;
            }
            const float p0 = i0 / (i0 + i1);
            if (u < p0) {
                node = c0;
                u = std::min(u / p0, 0.99999994f);
                pmf *= p0;
            } else {
                node = c1;
                u = std::min((u - p0) / (1.0f - p0), 0.99999994f);
                pmf *= 1.0f - p0;
            }
        }
    }

    float bvhPmf(V3 p, V3 n, int index) const {  // light_sampler.cpp:268-293
        if (index < 0 || index >= static_cast<int>(trail.size())) return 0.0f;
        uint64_t tr = trail[static_cast<size_t>(index)];
        const int depth = depthOf[static_cast<size_t>(index)];
        float pmf = 1.0f;
        int node = 0;
        for (int d = 0; d < depth; ++d) {
            const LightNode& nd = nodes[static_cast<size_t>(node)];
            if (nd.light >= 0) break;
            const int c0 = node + 1, c1 = nd.child1;
            const float i0 = nodes[static_cast<size_t>(c0)].lb.importance(p, n);
            const float i1 = nodes[static_cast<size_t>(c1)].lb.importance(p, n);
            const float sum = i0 + i1;
            if (!(sum > 0.0f)) return 0.0f;
            if (tr & 1u) {
                pmf *= i1 / sum;
                node = c1;
            } else {
                pmf *= i0 / sum;
                node = c0;
            }
            tr >>= 1;
        }
        return nodes[static_cast<size_t>(node)].light == index ? pmf : 0.0f;
    }
};

float Integrator::candidateImportance(int i, V3 p, V3 n) const {
    return i < explicitLights() ? lightImportance(lights[static_cast<size_t>(i)], p, n) : env.importance(n);
}

struct LightContext {  // LightSampler::Context (light_sampler.h:39-57)
    const Integrator* owner = nullptr;
    V3 p, n;
    bool uniform = true;
    float pmfs[kMaxExhaustive] = {};
    int topCount = 0;

    int sample(float u, float& pmf) const {  // light_sampler.cpp:196-229
        pmf = 0.0f;
        if (!owner) return -1;
        const int cnt = owner->count;
        if (cnt <= 0) return -1;
        if (uniform) {
            pmf = 1.0f / static_cast<float>(cnt);
            return std::min(static_cast<int>(u * static_cast<float>(cnt)), cnt - 1);
        }
        float acc = 0.0f;
        int k = topCount - 1;
        for (int i = 0; i < topCount - 1; ++i) {
            if (u < acc + pmfs[i]) {
                k = i;
                break;
            }
            acc += pmfs[i];
        }
        const float pk = pmfs[k];
        if (!owner->bvh) {
            pmf = pk;
            return k;
        }
        if (k < static_cast<int>(owner->infinite.size())) {
            pmf = pk;
            return owner->infinite[static_cast<size_t>(k)];
        }
        const float ur = std::min((u - acc) / pk, 0.99999994f);
        float treePmf = 0.0f;
        const int light = owner->bvhSample(p, n, std::max(ur, 0.0f), treePmf);
        pmf = pk * treePmf;
        return light;
    }
    float pmf(int index) const {  // light_sampler.cpp:231-240
        if (!owner) return 0.0f;
        if (index < 0 || index >= owner->count) return 0.0f;
        if (uniform) return 1.0f / static_cast<float>(owner->count);
        if (!owner->bvh) return pmfs[index];
        const int slot = owner->topSlot[static_cast<size_t>(index)];
        if (slot >= 0) return pmfs[slot];
        if (owner->nodes.empty()) return 0.0f;
        return pmfs[owner->infinite.size()] * owner->bvhPmf(p, n, index);
    }
};

LightContext prepare(const Integrator& it, V3 p, V3 n) {  // light_sampler.cpp:150-194
    LightContext c;
    c.owner = &it;
    c.p = p;
    c.n = n;
    if (it.count <= 1) return c;
    int count;
    if (!it.bvh) {
        if (it.count > kMaxExhaustive) return c;  // uniform (too many unbounded lights)
        count = it.count;
        for (int i = 0; i < count; ++i) {
            const float w = it.candidateImportance(i, p, n);
            c.pmfs[i] = (w > 0.0f && std::isfinite(w)) ? w : 0.0f;
        }
    } else {
        count = static_cast<int>(it.infinite.size()) + (it.nodes.empty() ? 0 : 1);
        for (size_t k = 0; k < it.infinite.size(); ++k) {
            const float w = it.candidateImportance(it.infinite[k], p, n);
            c.pmfs[k] = (w > 0.0f && std::isfinite(w)) ? w : 0.0f;
        }
        if (!it.nodes.empty()) {
            const float w = it.nodes[0].lb.importance(p, n) * INV_PI;
            c.pmfs[it.infinite.size()] = (w > 0.0f && std::isfinite(w)) ? w : 0.0f;
        }
    }
    c.topCount = count;
    float sum = 0.0f;
    for (int i = 0; i < count; ++i) sum += c.pmfs[i];
    if (!(sum > 0.0f) || !std::isfinite(sum)) {
        for (int i = 0; i < count; ++i) c.pmfs[i] = 1.0f / static_cast<float>(count);
    } else {
        const float a = (1.0f - kUniformShare) / sum, b = kUniformShare / static_cast<float>(count);
        for (int i = 0; i < count; ++i) c.pmfs[i] = c.pmfs[i] * a + b;
    }
    c.uniform = false;
    return c;
}

// ------------------------------------------------------------------------------------ PathTracer::Li
struct Settings {
    int maxDepth = 8, rrDepth = 3, shadowSamples = 1;
    float clampIndirect = 0.0f;
    bool motionBlur = false;  // RenderSettings::motionBlur: tracePixel draws a shutter time (renderer.cpp:152)
    float aoStrength = 0.0f;  // RenderSettings::aoStrength: contact AO at the first vertex (path_tracer.cpp:617-621)
};

struct RayCounters {
    uint64_t extension = 0, shadow = 0;
};

bool russianRoulette(V3& throughput, Sobol& sampler) {  // path_tracer.cpp:131-137
    const float m = std::max({throughput.x, throughput.y, throughput.z});
    if (m >= 1.0f) return true;
    if (!(m > 0.0f) || sampler.get1D() >= m) return false;
    throughput = throughput / m;
    return true;
}

bool makeShadowRay(const SI& isect, const LightSample& ls, Ray& out) {  // path_tracer.cpp:144-157
    if (!(ls.distance < INFINITY_F) || ls.distance >= 1e29f) {
        out.o = shadowOrigin(isect, ls.wi);
        out.d = ls.wi;
        out.tMin = 0.0f;
        out.tMax = 1e30f;
        return true;
    }
    return spawnRayTo(isect, isect.point + ls.wi * ls.distance, out);
}

// ---------------------------------------------------------------------------- media, walks, transmittance
V3 expNeg(V3 sigma, float t) { return V3(std::exp(-sigma.x * t), std::exp(-sigma.y * t), std::exp(-sigma.z * t)); }

// sampleFreeFlight (path_tracer.cpp:97-128): spectral-MIS distance sampling in a homogeneous medium.
bool sampleFreeFlight(V3 sigmaT, V3 sigmaS, float tMax, bool hitSurface, V3 channelWeights, Sobol& sampler,
                      V3& throughput, float& tScatter) {
    float wsum = channelWeights.x + channelWeights.y + channelWeights.z;
    if (!(wsum > 0.0f)) {
        channelWeights = V3(1.0f);
        wsum = 3.0f;
    }
    const V3 pc = channelWeights / wsum;
    const float uc = sampler.get1D();
    const int c = uc < pc.x ? 0 : (uc < pc.x + pc.y ? 1 : 2);
    const float u = sampler.get1D();
    const float stc = sigmaT[c];
    const float t = stc > 0.0f ? -std::log(std::max(1.0f - u, 1e-12f)) / stc : INFINITY_F;
    if (t < tMax) {
        const V3 Tr = expNeg(sigmaT, t);
        const float pdf = pc.x * sigmaT.x * Tr.x + pc.y * sigmaT.y * Tr.y + pc.z * sigmaT.z * Tr.z;
        if (!(pdf > 0.0f)) {
            throughput = V3(0.0f);
            return false;
        }
        throughput *= sigmaS * Tr / pdf;
        tScatter = t;
        return true;
    }
    V3 Tr;
    for (int k = 0; k < 3; ++k) Tr[k] = hitSurface ? std::exp(-sigmaT[k] * tMax) : (sigmaT[k] > 0.0f ? 0.0f : 1.0f);
    const float P = pc.x * Tr.x + pc.y * Tr.y + pc.z * Tr.z;
    throughput = P > 0.0f ? throughput * Tr / P : V3(0.0f);
    return false;
}

V3 sampleHenyeyGreenstein(V3 d, float g, float u0, float u1) {  // core/sampling/phase.h:27-39
    float cosTheta;
    if (std::abs(g) < 1e-3f) {
        cosTheta = 1.0f - 2.0f * u0;
    } else {
        const float sq = (1.0f - g * g) / (1.0f - g + 2.0f * g * u0);
        cosTheta = (1.0f + g * g - sq * sq) / (2.0f * g);
    }
    cosTheta = clampf(cosTheta, -1.0f, 1.0f);
    const float sinTheta = std::sqrt(std::max(0.0f, 1.0f - cosTheta * cosTheta));
    const float phi = TWO_PI * u1;
    const Frame f(d);
    return normalized(f.toWorld(V3(sinTheta * std::cos(phi), sinTheta * std::sin(phi), cosTheta)));
}

// PathTracer::randomWalkSubsurface (path_tracer.cpp:263-292): a walk inside the closed mesh from the entry
// point; true with the exit hit when the walk leaves through a surface.
bool randomWalkSubsurface(const Scene& scene, const SI& entry, V3 dir, V3 sigmaT, V3 albedo, float g,
                          Sobol& sampler, V3& throughput, SI& exitSi, RayCounters& rc) {
    const V3 sigmaS = albedo * sigmaT;
    V3 d = normalized(dir);
    V3 origin = spawnOrigin(entry, d);
    for (int step = 0; step < 256; ++step) {
        Ray r;
        r.o = origin;
        r.d = d;
        Hit h;
        ++rc.extension;
        const bool hit = intersect(scene, r, h);
        SI hitSi;
        if (hit) fillHit(scene, r, h, hitSi);
        const float tMax = hit ? hitSi.t : INFINITY_F;
        float t = 0.0f;
        const bool scattered = sampleFreeFlight(sigmaT, sigmaS, tMax, hit, throughput * albedo, sampler, throughput, t);
        if (!isFinite(throughput) || isBlack(throughput)) return false;
        if (!scattered) {
            if (!hit) return false;
            exitSi = hitSi;
            return true;
        }
        origin = r.o + r.d * t;
        float u0, u1;
        sampler.get2D(u0, u1);
        d = sampleHenyeyGreenstein(d, g, u0, u1);
        if (step >= 8 && !russianRoulette(throughput, sampler)) return false;
    }
    return false;
}

// approxAo (path_tracer.cpp:27-44): `samples` cosine-distributed any-hit rays of 0.35 m about the shading normal;
// the throughput factor lerp(1, unoccluded fraction, strength).
float approxAo(const Scene& scene, const SI& isect, Sobol& sampler, float strength, int samples) {
    if (strength <= 0.0f || samples <= 0) return 1.0f;
    const Frame frame(isect.normal);
    float occ = 0.0f;
    for (int i = 0; i < samples; ++i) {
        float u0, u1;
        sampler.get2D(u0, u1);
        const V3 dir = normalized(frame.toWorld(cosineSampleHemisphere(u0, u1)));
        Ray r;
        r.o = spawnOrigin(isect, dir);
        r.d = dir;
        r.tMin = 0.0f;
        r.tMax = 0.35f;
        if (intersectAny(scene, r)) occ += 1.0f;
    }
    const float visibility = 1.0f - (occ / static_cast<float>(samples));
    return lerpf(1.0f, visibility, strength);
}

// passesCutout (path_tracer.cpp:159-167): opacity 1 never passes, 0 always, in between a sampler draw (one
// dimension) at or above the opacity passes.
bool passesCutout(const pe_material_desc& m, Sobol& sampler) {
    const float alpha = materialOpacity(m);
    if (alpha >= 1.0f) return false;
    if (alpha <= 0.0f) return true;
    return sampler.get1D() >= alpha;
}

// shadowTransmittance (path_tracer.cpp:172-202), without shadow linking: thin-walled surfaces and cutouts pass
// the shadow ray on with their transmittance, anything else blocks it.
// Scene::nodeUidOf (scene.h) of a hit, and the link tests isLightLinked / isShadowLinked (scene.cpp:237-256).
uint64_t nodeUidOf(const Scene& scene, const SI& si) {
    const int shape = si.tri >= 0 ? scene.tris[static_cast<size_t>(si.tri)].shape : -1;
    return shape >= 0 ? scene.nodeUid[static_cast<size_t>(shape)] : 0;
}

bool linkedBy(int mode, const std::vector<uint64_t>& uids, uint64_t uid) {
    if (mode == PE_LINK_ALL) return true;
    const bool found = std::find(uids.begin(), uids.end(), uid) != uids.end();
    return mode == PE_LINK_INCLUDE ? found : !found;
}

bool isLightLinked(const Scene& s, int li, uint64_t uid) {
    if (!s.hasLinks || li < 0 || static_cast<size_t>(li) >= s.linkMode.size()) return true;
    return linkedBy(s.linkMode[static_cast<size_t>(li)], s.linkUids[static_cast<size_t>(li)], uid);
}

bool hasShadowLinkFor(const Scene& s, int li) {
    if (!s.hasShadowLinks || li < 0 || static_cast<size_t>(li) >= s.shadowLinkMode.size()) return false;
    return s.shadowLinkMode[static_cast<size_t>(li)] != PE_LINK_ALL;
}

// shadowTransmittance (path_tracer.cpp:172-202) for the shadow ray of light `lightIndex` (-1: the environment):
// blockers the light's shadow links leave out are passed as if absent.
V3 shadowTransmittance(const Scene& scene, bool transmitters, const Ray& shadow, RayCounters& rc, int lightIndex = -1) {
    const bool needShadowLink = hasShadowLinkFor(scene, lightIndex);
    if (!transmitters && !needShadowLink) return intersectAny(scene, shadow) ? V3(0.0f) : V3(1.0f);
    V3 T(1.0f);
    V3 origin = shadow.o;
    const V3 dir = shadow.d;
    float remaining = shadow.tMax;
    for (int i = 0; i < 64; ++i) {
        Ray r;
        r.o = origin;
        r.d = dir;
        r.tMin = 0.0f;
        r.tMax = remaining;
        Hit h;
        if (i > 0) ++rc.shadow;
        if (!intersect(scene, r, h)) return T;
        SI hit;
        fillHit(scene, r, h, hit);
        if (needShadowLink &&
            !linkedBy(scene.shadowLinkMode[static_cast<size_t>(lightIndex)],
                      scene.shadowLinkUids[static_cast<size_t>(lightIndex)], nodeUidOf(scene, hit))) {
            origin = offsetRayOrigin(hit.point, hit.pError, lengthSq(hit.ng) > 0.0f ? hit.ng : hit.normal, dir);
            remaining -= hit.t;
            if (remaining <= 0.0f) return T;
            continue;
        }
        const pe_material_desc& m = scene.materials[static_cast<size_t>(hit.material)];
        if (!materialTransmitsShadows(m)) return V3(0.0f);
        ShadingPoint sp{hit.normal, hit.tangent, hit.frontFace, hit.ng, 0.0f, hit.point};
        T *= materialShadowTransmittance(m, sp, dir);
        if (isBlack(T)) return T;
        origin = offsetRayOrigin(hit.point, hit.pError, lengthSq(hit.ng) > 0.0f ? hit.ng : hit.normal, dir);
        remaining -= hit.t;
        if (remaining <= 0.0f) return T;
    }
    return V3(0.0f);
}

constexpr int kMaxMediumEvents = 1024;  // path_tracer.cpp:51

struct MediumStack {  // path_tracer.cpp:63-83 (entries are material indices; -1 = the walk's exit material)
    int owner[4] = {0, 0, 0, 0};
    int n = 0;
    int current() const { return n > 0 ? owner[n - 1] : -1; }
    void enter(int m) {
        if (n < 4) owner[n++] = m;
    }
    void leave(int m) {
        for (int i = n - 1; i >= 0; --i) {
            if (owner[i] != m) continue;
            for (int j = i; j < n - 1; ++j) owner[j] = owner[j + 1];
            --n;
            return;
        }
    }
};

// PathTracer::PrimaryHit (path_tracer.cpp:441-443): the first segment's surface, after camera-invisible skips.
struct PrimaryHit {
    bool hit = false;
    SI isect;
    float coverage = -1.0f;  // >= 0: the camera ray passed thin walls to a transparent background (path_tracer.h:54-56)
};

// LightPasses (path_tracer.h:21-30): the eight light passes (LightPass order) and the light groups.
constexpr int kLightPasses = 8, kMaxLightGroups = 16;
enum : int { kDiffuseDirect = 0, kDiffuseIndirect, kGlossyDirect, kGlossyIndirect, kTransmissionDirect,
             kTransmissionIndirect, kEmission, kBackground };
struct LightPasses {
    V3 c[kLightPasses];
    V3 groups[kMaxLightGroups];
};

V3 Li(const Scene& scene, const Integrator& it, const Settings& st, const Ray& ray, Sobol& sampler, RayCounters& rc,
      PrimaryHit* primary = nullptr, LightPasses* passes = nullptr) {
    V3 L(0.0f), throughput(1.0f);  // path_tracer.cpp:321-890
    bool firstSegment = true;
    Ray cur = ray;
    bool specularBounce = true;
    float lastBsdfPdf = 0.0f;
    V3 lastPoint = ray.o;
    uint64_t prevReceiverUid = 0;  // :331
    MediumStack media;
    int mediumEvents = 0;  // :51, :463
    float lambda = 0.0f;
    bool pendingExit = false;
    SI isect;
    int depth = 0;
    const int explicitLights = it.explicitLights();
    const int lightCount = explicitLights + (it.hasEnv ? 1 : 0);
    LightContext picker;
    const int shadowSamples = st.shadowSamples;
    bool cameraRay = true;
    int diffuseBounces = 0;
    int passThroughs = 0;
    bool transmitters = false;  // Scene::hasShadowTransmitters: every descriptor material is retained
    for (const pe_material_desc& m : scene.materials) transmitters = transmitters || materialTransmitsShadows(m);
    pe_material_desc exitMaterial{};  // subsurfaceExitMaterial (path_tracer.cpp:56-59): white Lambert
    exitMaterial.type = PE_MAT_LAMBERT;
    exitMaterial.base_color[0] = exitMaterial.base_color[1] = exitMaterial.base_color[2] = 1.0f;
    auto materialOf = [&](int index) -> const pe_material_desc& {
        return index < 0 ? exitMaterial : scene.materials[static_cast<size_t>(index)];
    };
    // accumulate returns what it added (black for a non-finite contribution), the value the passes receive.
    auto accumulate = [&L](V3 c) -> V3 {
        if (!isFinite(c)) return V3(0.0f);
        L += c;
        return c;
    };
    auto accumulateIndirect = [&](V3 c) -> V3 {
        if (st.clampIndirect > 0.0f) {
            const float m = std::max({c.x, c.y, c.z});
            if (m > st.clampIndirect) return accumulate(c * (st.clampIndirect / m));
        }
        return accumulate(c);
    };
    // Render passes and light groups (path_tracer.cpp:372-395): the lobe shares of the direction sampled at
    // the first surface (per channel) route everything found after it; light found by the first bounce is
    // direct, later light indirect.
    V3 firstLobe[3] = {V3(1.0f), V3(0.0f), V3(0.0f)};
    const int numLightGroups = scene.numLightGroups;
    auto lightGroupIndex = [&](int li) {  // Scene::lightGroupIndex (scene.cpp:258-261)
        return li >= 0 && static_cast<size_t>(li) < scene.lightGroup.size() ? scene.lightGroup[static_cast<size_t>(li)] : -1;
    };
    auto passAt = [](int lobe, bool direct) {
        return lobe == 0 ? (direct ? kDiffuseDirect : kDiffuseIndirect)
                         : lobe == 1 ? (direct ? kGlossyDirect : kGlossyIndirect)
                                     : (direct ? kTransmissionDirect : kTransmissionIndirect);
    };
    auto credit = [&](int pass, V3 added) {
        if (passes) passes->c[pass] += added;
    };
    auto creditGroup = [&](int groupIdx, V3 added) {
        if (!passes || numLightGroups <= 0 || isBlack(added)) return;
        const int g = groupIdx >= 0 && groupIdx < numLightGroups ? groupIdx : 0;
        if (g < kMaxLightGroups) passes->groups[g] += added;
    };
    auto creditPath = [&](V3 added, bool direct) {
        if (!passes) return;
        for (int k = 0; k < 3; ++k) passes->c[passAt(k, direct)] += added * firstLobe[k];
    };
    for (;;) {
        V3 wo;
        if (pendingExit) {  // :415-418
            pendingExit = false;
            wo = isect.normal;
        } else {
            isect = SI{};
            const V3 segmentStart = cur.o;
            Hit h;
            ++rc.extension;
            bool hit = intersect(scene, cur, h);
            if (hit) fillHit(scene, cur, h, isect);
            for (int skip = 0; hit && skip < 64; ++skip) {  // :423-441
                if (!(cameraRay && isect.cameraInvisible) &&
                    !passesCutout(scene.materials[static_cast<size_t>(isect.material)], sampler))
                    break;
                Ray next;
                next.o = offsetRayOrigin(isect.point, isect.pError, isect.ng, cur.d);
                next.d = cur.d;
                next.tMin = 0.0f;
                next.tMax = 1e30f;
                cur = next;
                isect = SI{};
                ++rc.extension;
                hit = intersect(scene, cur, h);
                if (hit) fillHit(scene, cur, h, isect);
            }
            if (firstSegment && primary) {
                primary->hit = hit;
                if (hit) primary->isect = isect;
            }
            firstSegment = false;
            // Inside a volume (:457-479): a free flight in a scattering medium, which may scatter before the
            // surface (a new ray from there, no vertex, no NEE inside), or Beer-Lambert absorption along the segment.
            const int medOwner = media.current();
            V3 sigmaA, sigmaS;
            float g = 0.0f;
            if (medOwner >= 0 && materialInteriorMedium(scene.materials[static_cast<size_t>(medOwner)], sigmaA, sigmaS, g)) {
                const float dirLen = length(cur.d);
                const float segment = hit ? length(isect.point - segmentStart) : INFINITY_F;
                if (!isBlack(sigmaS)) {
                    float t = 0.0f;
                    if (sampleFreeFlight(sigmaA + sigmaS, sigmaS, segment, hit, throughput, sampler, throughput, t)) {
                        if (++mediumEvents > kMaxMediumEvents || !isFinite(throughput)) break;
                        const V3 d = cur.d / dirLen;
                        float u0, u1;
                        sampler.get2D(u0, u1);
                        Ray next;
                        next.o = segmentStart + d * t;
                        next.d = sampleHenyeyGreenstein(d, g, u0, u1);
                        next.tMin = 0.0f;
                        next.tMax = 1e30f;
                        cur = next;
                        cameraRay = false;
                        specularBounce = true;
                        if (mediumEvents > 8 && !russianRoulette(throughput, sampler)) break;
                        continue;
                    }
                } else if (hit) {
                    throughput *= expNeg(sigmaA, segment);
                } else if (!isBlack(sigmaA)) {
                    break;
                }
                if (isBlack(throughput) || !isFinite(throughput)) break;
            }
            wo = -cur.d;
            if (!hit) {  // :482-514
                // Seen straight from the camera (past thin walls and hidden lights), a background other than
                // the environment (:487-495). `primary` is always given then: tracePixel asks for it whenever
                // the background is not the environment (renderer.cpp:158-161).
                if (depth == 0 && scene.bgMode != PE_BACKGROUND_ENVIRONMENT) {
                    if (scene.bgMode == PE_BACKGROUND_COLOR) {
                        const V3 added = accumulate(throughput * scene.bgColor);
                        credit(kBackground, added);
                        creditGroup(scene.envLightGroup, added);
                    } else if (passThroughs > 0 && primary) {
                        primary->coverage = std::max(0.0f, 1.0f - luminance(throughput));
                    }
                    break;
                }
                if (it.hasEnv) {
                    const V3 Le = it.env.eval(cur.d);
                    float weight = 1.0f;
                    if (!specularBounce) {
                        const float lightPdf = picker.pmf(explicitLights) * it.env.pdfLi(cur.d);
                        weight = powerHeuristic(1, lastBsdfPdf, shadowSamples, lightPdf);
                    }
                    const V3 envLocal = Le * weight;
                    const V3 added = diffuseBounces >= 2 ? accumulateIndirect(throughput * envLocal)
                                                         : accumulate(throughput * envLocal);
                    if (depth == 0) credit(kBackground, added);
                    else creditPath(added, depth == 1);
                    creditGroup(scene.envLightGroup, added);
                }
                break;
            }
            const V3 emitted = materialEmitted(materialOf(isect.material));
            // :565-593; light linking: a bounce from a receiver the light is not linked to does not see it
            if (!isBlack(emitted) && (depth == 0 || isect.emitter < 0 || isLightLinked(scene, isect.emitter, prevReceiverUid))) {
                const int li = isect.emitter;
                float weight = 1.0f;
                if (!specularBounce) {
                    const float lightPdf =
                        li >= 0 ? picker.pmf(li) * lightPdfLiHit(it.lights[static_cast<size_t>(li)], lastPoint, isect.point) : 0.0f;
                    if (lightPdf > 0.0f && lastBsdfPdf > 0.0f) weight = powerHeuristic(1, lastBsdfPdf, shadowSamples, lightPdf);
                }
                const V3 emitLocal = emitted * weight;
                const V3 added = diffuseBounces >= 2 ? accumulateIndirect(throughput * emitLocal)
                                                     : accumulate(throughput * emitLocal);
                if (depth == 0) credit(kEmission, added);
                else creditPath(added, depth == 1);
                creditGroup(li >= 0 ? lightGroupIndex(li) : 0, added);
            }
        }
        if (depth >= st.maxDepth) break;  // :596-599
        const pe_material_desc& mat = materialOf(isect.material);
        if (lambda == 0.0f && materialIsDispersive(mat)) {  // :602-607
            V3 w;
            lambda = sampleWavelength(sampler.get1D(), w);
            throughput *= w;
        }
        if (st.aoStrength > 0.0f && depth == 0)  // contact AO (:617-621)
            throughput *= approxAo(scene, isect, sampler, st.aoStrength, shadowSamples);
        const uint64_t receiverUid = nodeUidOf(scene, isect);  // :623
        const bool isSpecular = materialIsDelta(mat);
        const ShadingPoint sp{isect.normal, isect.tangent, isect.frontFace, isect.ng, lambda, isect.point};
        if (lightCount > 0 && !isSpecular) {  // NEE, :631-698
            picker = prepare(it, isect.point, isect.normal);
            for (int s = 0; s < shadowSamples; ++s) {
                float lightPmf = 0.0f;
                const int lightIndex = picker.sample(sampler.get1D(), lightPmf);
                float ux, uy;
                if (lightIndex < 0 || !(lightPmf > 0.0f)) {
                    sampler.get2D(ux, uy);
                    continue;
                }
                const bool fromEnv = lightIndex >= explicitLights;
                if (!fromEnv && !isLightLinked(scene, lightIndex, receiverUid)) {  // :641-644
                    sampler.get2D(ux, uy);
                    continue;
                }
                sampler.get2D(ux, uy);
                const LightSample ls = fromEnv ? it.env.sampleLi(ux, uy)
                                               : lightSampleLi(it.lights[static_cast<size_t>(lightIndex)], isect.point, ux, uy);
                if (!ls.valid()) continue;
                Ray shadow;
                if (!makeShadowRay(isect, ls, shadow)) continue;
                ++rc.shadow;
                const V3 vis = shadowTransmittance(scene, transmitters, shadow, rc, fromEnv ? -1 : lightIndex);
                if (!isBlack(vis)) {
                    float bsdfPdf = 0.0f;
                    const V3 f = evalPdf(mat, wo, ls.wi, sp, bsdfPdf);
                    const float cosTheta = std::abs(dot(ls.wi, isect.normal));
                    if (!isBlack(f) && cosTheta > 0.0f) {
                        const float lightPdf = lightPmf * ls.pdf;
                        const float weight = powerHeuristic(shadowSamples, lightPdf, 1, bsdfPdf);
                        const V3 neeLocal = f * ls.Li * vis *
                                            (cosTheta * weight / (lightPdf * static_cast<float>(shadowSamples)));
                        const V3 c = throughput * neeLocal;
                        const V3 added = diffuseBounces >= 1 ? accumulateIndirect(c) : accumulate(c);
                        if (passes && !isBlack(added)) {  // :679-693
                            creditGroup(fromEnv ? scene.envLightGroup : lightGroupIndex(lightIndex), added);
                            if (depth == 0) {
                                // The first surface: this light sample's value split over the lobe groups.
                                const LobeSplit sp3 = evalSplit(mat, wo, ls.wi, sp);
                                const V3 parts[3] = {sp3.diffuse, sp3.glossy, sp3.transmission};
                                const V3 sum = sp3.diffuse + sp3.glossy + sp3.transmission;
                                for (int k = 0; k < 3; ++k) {
                                    V3 frac;
                                    for (int ch = 0; ch < 3; ++ch)
                                        frac[ch] = sum[ch] > 0.0f ? parts[k][ch] / sum[ch] : (k == 0 ? 1.0f : 0.0f);
                                    passes->c[passAt(k, true)] += added * frac;
                                }
                            } else {
                                creditPath(added, false);
                            }
                        }
                    }
                }
            }
        }
        const float uc = sampler.get1D();  // :736-738
        float u0, u1;
        sampler.get2D(u0, u1);
        BsdfSample bs;
        if (!sampleFull(mat, wo, sp, uc, u0, u1, bs)) break;
        if (bs.subsurface) {  // :746-764
            V3 sigmaT, albedo;
            float g = 0.0f;
            if (!subsurfaceCoeffs(mat, sp, sigmaT, albedo, g)) break;
            throughput *= bs.weight;
            SI exitSi;
            if (!randomWalkSubsurface(scene, isect, bs.wi, sigmaT, albedo, g, sampler, throughput, exitSi, rc)) break;
            exitSi.normal = -exitSi.normal;
            exitSi.material = -1;
            isect = exitSi;
            pendingExit = true;
            cameraRay = false;
            continue;
        }
        if (bs.pdf <= 0.0f || isBlack(bs.f)) break;
        if (bs.passThrough) {  // :771-786
            throughput *= bs.f * (std::abs(dot(bs.wi, isect.normal)) / bs.pdf);
            if (!isFinite(throughput)) break;
            cur = spawnRay(isect, bs.wi);
            if (++passThroughs > 64) break;
            continue;
        }
        const V3 wi = bs.wi;
        if (passes && depth == 0) {  // :789-812: the first surface's lobe shares of the sampled direction
            const bool refracted = dot(wo, isect.ng) * dot(wi, isect.ng) < 0.0f;
            if (bs.delta) {
                firstLobe[0] = V3(0.0f);
                firstLobe[1] = refracted ? V3(0.0f) : V3(1.0f);
                firstLobe[2] = refracted ? V3(1.0f) : V3(0.0f);
            } else {
                const LobeSplit b = evalSplit(mat, wo, wi, sp);
                const V3 sum = b.diffuse + b.glossy + b.transmission;
                for (int ch = 0; ch < 3; ++ch) {
                    const bool ok = sum[ch] > 0.0f;
                    firstLobe[0][ch] = ok ? b.diffuse[ch] / sum[ch] : 1.0f;
                    firstLobe[1][ch] = ok ? b.glossy[ch] / sum[ch] : 0.0f;
                    firstLobe[2][ch] = ok ? b.transmission[ch] / sum[ch] : 0.0f;
                }
            }
        }
        const float cosTheta = std::abs(dot(wi, isect.normal));  // :814-816
        throughput *= bs.f * (cosTheta / bs.pdf);
        if (!isFinite(throughput)) break;
        lastBsdfPdf = bs.pdf;
        lastPoint = isect.point;
        prevReceiverUid = receiverUid;
        specularBounce = bs.delta;
        if (!bs.delta) ++diffuseBounces;
        cameraRay = cameraRay && bs.delta && dot(wi, isect.normal) < 0.0f;
        if (materialIsVolumeBoundary(mat)) {  // :840-848
            const float a = dot(wo, isect.ng), b = dot(wi, isect.ng);
            if (a * b < 0.0f) {
                if (a > 0.0f) media.enter(isect.material);
                else media.leave(isect.material);
            }
        }
        cur = spawnRay(isect, wi);
        ++depth;
        if (depth > st.rrDepth) {  // :862-870
            if (!russianRoulette(throughput, sampler)) break;
        }
    }
    return L;
}

// ----------------------------------------------------------------------------- pixel_filter.cpp / camera
struct FilterSampler {  // pixel_filter.cpp:37-87
    int type = 0;
    float width = 1.0f, radius = 0.5f;
    std::vector<float> cdf;
    FilterSampler(int t, float wIn) : type(t) {
        width = clampf(wIn, 0.25f, 8.0f);
        radius = 0.5f * width;
        if (type == PE_FILTER_BOX) return;
        constexpr int kBins = 512;
        cdf.resize(kBins + 1);
        cdf[0] = 0.0f;
        const float dx = width / kBins;
        double acc = 0.0;
        for (int i = 0; i < kBins; ++i) {
            acc += std::max(0.0f, evaluate(-radius + (static_cast<float>(i) + 0.5f) * dx));
            cdf[static_cast<size_t>(i) + 1] = static_cast<float>(acc);
        }
        if (!(acc > 0.0)) {
            cdf.clear();
            type = PE_FILTER_BOX;
            return;
        }
        for (float& c : cdf) c = static_cast<float>(c / acc);
        cdf.back() = 1.0f;
    }
    float evaluate(float x) const {
        const float r = radius;
        if (std::abs(x) > r) return 0.0f;
        switch (type) {
            case PE_FILTER_BOX: return 1.0f;
            case PE_FILTER_TRIANGLE: return std::max(0.0f, 1.0f - std::abs(x) / r);
            case PE_FILTER_GAUSSIAN: {
                const float s = width / 6.0f;
                const float k = 1.0f / (2.0f * s * s);
                return std::max(0.0f, std::exp(-k * x * x) - std::exp(-k * r * r));
            }
            case PE_FILTER_BLACKMAN_HARRIS: {
                const float a = TWO_PI * (x / width + 0.5f);
                return 0.35875f - 0.48829f * std::cos(a) + 0.14128f * std::cos(2.0f * a) - 0.01168f * std::cos(3.0f * a);
            }
        }
        return 0.0f;
    }
    float sample1D(float u) const {
        if (cdf.empty()) return (u - 0.5f) * width;
        constexpr int kBins = 512;
        const auto it = std::upper_bound(cdf.begin(), cdf.end(), u);
        const int i = std::clamp(static_cast<int>(it - cdf.begin()) - 1, 0, kBins - 1);
        const float c0 = cdf[static_cast<size_t>(i)], c1 = cdf[static_cast<size_t>(i) + 1];
        const float t = c1 > c0 ? (u - c0) / (c1 - c0) : 0.5f;
        return -radius + (static_cast<float>(i) + clampf(t, 0.0f, 1.0f)) * (width / kBins);
    }
};

// idColor (renderer.cpp:252-273): an id's stable colour, splitmix64 to a hue stepped by the golden ratio, HSV
// with value 1 and saturation 0.75; id 0 is black.
V3 idColor(uint64_t id) {
    if (id == 0) return V3(0.0f);
    uint64_t z = id + 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    z ^= z >> 31;
    const float h = std::fmod(static_cast<float>(z >> 40) * 0x1p-24f + 0.618034f, 1.0f) * 6.0f;
    const int i = static_cast<int>(h) % 6;
    const float f = h - std::floor(h);
    const float v = 1.0f, sat = 0.75f, p0 = v * (1.0f - sat), q = v * (1.0f - sat * f), t = v * (1.0f - sat * (1.0f - f));
    switch (i) {
        case 0: return V3(v, t, p0);
        case 1: return V3(q, v, p0);
        case 2: return V3(p0, v, t);
        case 3: return V3(p0, q, v);
        case 4: return V3(t, p0, v);
        default: return V3(v, p0, q);
    }
}

// ThinLensCamera::samplePolygon (camera/thin_lens_camera.cpp:36-47): uniform on the regular n-gon inscribed in
// the unit circle, turned by `rot`; the slice from u0, uniformSampleTriangle (sampling.h:97-100) within it.
void samplePolygon(int blades, float rot, float u0, float u1, float& x, float& y) {
    const float n = static_cast<float>(blades);
    const float scaled = std::min(u0 * n, n - 1e-4f);
    const float k = std::floor(scaled);
    const float ux = scaled - k;
    const float a0 = rot + 6.28318530717958647692f * k / n;
    const float a1 = a0 + 6.28318530717958647692f / n;
    const float su0 = std::sqrt(ux);
    const float b0 = 1.0f - su0, b1 = u1 * su0;
    x = std::cos(a0) * b0 + std::cos(a1) * b1;
    y = std::sin(a0) * b0 + std::sin(a1) * b1;
}

// Camera::generateRay(u, v, lensSample, time) (camera/camera.h:57-91): at time 0, or without motion blur or a
// shutter-end pose, the camera's own generateRay (PerspectiveCamera perspective_camera.cpp:36-40, ThinLensCamera
// thin_lens_camera.cpp:49-61, OrthographicCamera orthographic_camera.cpp:25-28); otherwise the two poses blended
// linearly, a disk aperture of the blended radius on the blended, renormalised lens axes.
Ray generateCameraRay(const pe_camera& c, bool motionBlur, float sx, float sy, float lx, float ly, float time) {
    Ray ray;
    if (!motionBlur || !c.has_shutter_end || time <= 0.0f) {
        const V3 pos = c3(c.position), ll = c3(c.lower_left), hor = c3(c.horizontal), ver = c3(c.vertical);
        const V3 target = ll + sx * hor + sy * ver;
        if (c.orthographic) {
            ray.o = target;
            ray.d = c3(c.forward);
            return ray;
        }
        ray.o = pos;
        if (c.lens_radius > 0.0f) {
            float x, y;
            if (c.blades >= 3) samplePolygon(c.blades, c.blade_rotation, lx, ly, x, y);
            else uniformSampleDisk(lx, ly, x, y);
            x *= c.lens_radius;
            y *= c.lens_radius;
            ray.o = pos + c3(c.u) * x + c3(c.v) * y;
        }
        ray.d = normalized(target - ray.o);
        return ray;
    }
    const float t = std::min(time, 1.0f);
    const auto mix3 = [t](V3 a, V3 b) { return a * (1.0f - t) + b * t; };
    const V3 pos = mix3(c3(c.position), c3(c.end_position));
    const V3 ll = mix3(c3(c.lower_left), c3(c.end_lower_left));
    const V3 hor = mix3(c3(c.horizontal), c3(c.end_horizontal));
    const V3 ver = mix3(c3(c.vertical), c3(c.end_vertical));
    if (c.orthographic) {
        ray.o = ll + hor * sx + ver * sy;
        ray.d = normalized(mix3(c3(c.forward), c3(c.end_forward)));
        return ray;
    }
    const float radius = c.lens_radius * (1.0f - t) + c.end_lens_radius * t;
    ray.o = pos;
    if (radius > 0.0f) {
        const V3 cu = normalized(mix3(c3(c.u), c3(c.end_u))), cv = normalized(mix3(c3(c.v), c3(c.end_v)));
        float x, y;
        uniformSampleDisk(lx, ly, x, y);
        ray.o = pos + cu * (x * radius) + cv * (y * radius);
    }
    ray.d = normalized(ll + hor * sx + ver * sy - ray.o);
    return ray;
}

struct Tracer {
    const Scene& scene;
    const Integrator& it;
    Settings st;
    pe_camera cam;
    FilterSampler filter;
    int width, height;
    uint64_t seed;

    // tracePixel's camera ray (renderer.cpp:143-153): filter offset (dim 0), lens sample (dim 1) and, with motion
    // blur, the shutter time (dim 2), then Camera::generateRay (camera.h:57-91) of the camera pe_camera describes.
    Ray cameraRay(int x, int y, int s, Sobol& sampler) const {
        sampler.startPixel(x, y);
        sampler.startSample(s);
        float fu, fv;
        sampler.get2D(fu, fv);
        const float ox = filter.sample1D(fu), oy = filter.sample1D(fv);
        const float sx = (static_cast<float>(x) + 0.5f + ox) / static_cast<float>(width);  // rasterToScreen (:37-39)
        const float sy = 1.0f - (static_cast<float>(y) + 0.5f + oy) / static_cast<float>(height);
        float lx, ly;
        sampler.get2D(lx, ly);
        const float time = st.motionBlur ? sampler.get1D() : 0.0f;
        return generateCameraRay(cam, st.motionBlur, sx, sy, lx, ly, time);
    }

    // tracePixel (renderer.cpp:139-162): the camera ray, then Li; with `aov`, the AOVs
    // (:175-205): the primary hit's clamped albedo (primaryAlbedo, :47-58: Disney baseColor, Lambert albedo,
    // Dielectric tint), its shading normal and coverage 1 (the thin walls' coverage where the camera ray passed
    // them to a transparent background), or on a miss the clamped Color background or environment radiance
    // along the camera ray, a zero normal and coverage 0.
    V3 trace(int x, int y, int s, Sobol& sampler, RayCounters& rc, float* aov = nullptr) const {
        const Ray ray = cameraRay(x, y, s, sampler);
        if (!aov) return Li(scene, it, st, ray, sampler, rc);
        PrimaryHit primary;
        const V3 L = Li(scene, it, st, ray, sampler, rc, &primary);
        auto clamp01 = [](V3 c) {
            return V3(std::clamp(c.x, 0.0f, 1.0f), std::clamp(c.y, 0.0f, 1.0f), std::clamp(c.z, 0.0f, 1.0f));
        };
        V3 albedo(0.0f), normal(0.0f);
        float alpha = 0.0f;
        if (primary.hit) {
            const int m = primary.isect.material;
            if (m >= 0) albedo = clamp01(c3(scene.materials[static_cast<size_t>(m)].base_color));
            normal = primary.isect.normal;
            alpha = primary.coverage >= 0.0f ? primary.coverage : 1.0f;  // :184-189
        } else if (scene.bgMode == PE_BACKGROUND_COLOR) {
            albedo = clamp01(scene.bgColor);
        } else if (it.hasEnv) {
            albedo = clamp01(it.env.eval(ray.d));
        }
        const float v[7] = {albedo.x, albedo.y, albedo.z, normal.x, normal.y, normal.z, alpha};
        std::copy(v, v + 7, aov);
        return L;
    }

    // tracePixel with render passes (renderer.cpp:162-171): the light passes and groups Li credits, and the
    // primary hit's data passes (depth from the camera, position, idColor of its node uid and of its material
    // index + 1; all zero for a miss). out: 24 light-pass floats, 48 group floats, depth, position xyz, object
    // id rgb, material id rgb, then the hit's shape and material index (-1 for a miss; the cryptomatte lookup
    // keys): 84 floats.
    V3 tracePasses(int x, int y, int s, Sobol& sampler, RayCounters& rc, float* out) const {
        const Ray ray = cameraRay(x, y, s, sampler);
        PrimaryHit primary;
        LightPasses lp;
        const V3 L = Li(scene, it, st, ray, sampler, rc, &primary, &lp);
        for (int k = 0; k < kLightPasses; ++k)
            for (int c = 0; c < 3; ++c) out[3 * k + c] = lp.c[k][c];
        for (int g = 0; g < kMaxLightGroups; ++g)
            for (int c = 0; c < 3; ++c) out[24 + 3 * g + c] = lp.groups[g][c];
        std::fill(out + 72, out + 82, 0.0f);
        out[82] = out[83] = -1.0f;
        if (primary.hit) {
            const SI& h = primary.isect;
            out[72] = length(h.point - ray.o);
            out[73] = h.point.x;
            out[74] = h.point.y;
            out[75] = h.point.z;
            const int shape = h.tri >= 0 ? scene.tris[static_cast<size_t>(h.tri)].shape : -1;
            const V3 obj = idColor(shape >= 0 ? scene.nodeUid[static_cast<size_t>(shape)] : 0);
            const V3 mat = idColor(h.material >= 0 ? static_cast<uint64_t>(h.material) + 1 : 0);
            for (int c = 0; c < 3; ++c) {
                out[76 + c] = obj[c];
                out[79 + c] = mat[c];
            }
            out[82] = static_cast<float>(shape);
            out[83] = static_cast<float>(h.material);
        }
        return L;
    }
};

std::shared_ptr<Integrator> integratorFor(const Scene& s) {
    auto it = std::make_shared<Integrator>();
    for (const auto& l : s.lights) it->lights.push_back(makeLight(l));
    if (s.hasEnv) {
        it->hasEnv = true;
        Env& e = it->env;
        e.rotation = s.env.rotation;
        e.intensity = s.env.intensity;
        if (!s.envRgb.empty()) {
            e.w = s.env.width;
            e.h = s.env.height;
            e.rgb = s.envRgb.data();
        }
        e.build(true);
        e.meanLuminance = std::max(0.0f, luminance(e.power()) / (4.0f * PI));
        if (s.env.has_sun) e.setSun(c3(s.env.sun_direction), s.env.sun_cos_max, c3(s.env.sun_radiance));
    }
    it->count = static_cast<int>(it->lights.size()) + (it->hasEnv ? 1 : 0);
    it->buildSampler();
    return it;
}

const Integrator& integratorOf(const Scene& s) {
    auto* self = const_cast<Scene*>(&s);
    if (!self->integrator) self->integrator = integratorFor(s);
    return *static_cast<const Integrator*>(self->integrator.get());
}

}  // namespace

}  // namespace oracle

using namespace oracle;

#define ORACLE_API extern "C" __attribute__((visibility("default")))

// thinFilmReflectance(cos[i], n1, n2, n3, thickness[i], lambda[i]) for n queries.
ORACLE_API int oracle_thin_film(size_t n, const float* cos1, float n1, float n2, float n3re, float n3im,
                                const float* thickness, const float* lambda, float* out) {
    for (size_t i = 0; i < n; ++i) out[i] = thinFilmReflectanceC(cos1[i], n1, n2, n3re, n3im, thickness[i], lambda[i]);
    return 0;
}

// sampleWavelength for n values of u (lambda[i], weight[3i..3i+2]) and cauchyIor(nd, abbe, lambda) at m wavelengths.
ORACLE_API int oracle_spectral(size_t n, const float* u, float* lambda, float* weight, float nd, float abbe, size_t m,
                               const float* wl, float* ior) {
    for (size_t i = 0; i < n; ++i) {
        V3 w;
        lambda[i] = sampleWavelength(u[i], w);
        weight[3 * i] = w.x;
        weight[3 * i + 1] = w.y;
        weight[3 * i + 2] = w.z;
    }
    for (size_t i = 0; i < m; ++i) ior[i] = cauchyIor(nd, abbe, wl[i]);
    return 0;
}

// LightSampler::Context at n points: pmf of every candidate (pmf: n * candidates) and one sample per point
// (u[i] -> index[i], sample_pmf[i]). Returns the candidate count, or -1 on failure.
ORACLE_API int oracle_light_select(void* sp, size_t n, const float* p, const float* nrm, const float* u, float* pmf,
                                   int32_t* index, float* sample_pmf, int32_t* uses_bvh) {
    try {
        const Scene& scene = *static_cast<Scene*>(sp);
        const Integrator& it = integratorOf(scene);
        if (uses_bvh) *uses_bvh = it.bvh ? 1 : 0;
        for (size_t i = 0; i < n; ++i) {
            const V3 P(p[3 * i], p[3 * i + 1], p[3 * i + 2]), N(nrm[3 * i], nrm[3 * i + 1], nrm[3 * i + 2]);
            const LightContext c = prepare(it, P, N);
            if (pmf)
                for (int k = 0; k < it.count; ++k) pmf[i * static_cast<size_t>(it.count) + static_cast<size_t>(k)] = c.pmf(k);
            float sp_ = 0.0f;
            const int k = c.sample(u[i], sp_);
            if (index) index[i] = k;
            if (sample_pmf) sample_pmf[i] = sp_;
        }
        return it.count;
    } catch (const std::exception&) {
        return -1;
    }
}

// generateCameraRay for n queries (in: u, v, lens x, lens y, time; time < 0 is the form without a time) ->
// origin and direction (6 floats each).
ORACLE_API int oracle_camera_rays(const pe_camera* cam, size_t n, const float* in, float* out) {
    for (size_t i = 0; i < n; ++i) {
        const float* q = in + 5 * i;
        const Ray r = generateCameraRay(*cam, q[4] >= 0.0f, q[0], q[1], q[2], q[3], q[4] >= 0.0f ? q[4] : 0.0f);
        const float v[6] = {r.o.x, r.o.y, r.o.z, r.d.x, r.d.y, r.d.z};
        std::copy(v, v + 6, out + 6 * i);
    }
    return 0;
}

// Per-sample radiance (tracePixel + Li) for n (px, py, sample) triples.
ORACLE_API int oracle_trace_samples(void* sp, const pe_camera* cam, const pe_settings* ps, int width, int height,
                                    size_t n, const int32_t* px, const int32_t* py, const int32_t* s, float* L,
                                    uint64_t* rays) {
    try {
        const Scene& scene = *static_cast<Scene*>(sp);
        const Integrator& it = integratorOf(scene);
        Settings st;
        st.maxDepth = ps->max_bounces;
        st.rrDepth = ps->rr_depth;
        st.shadowSamples = std::max(1, ps->shadow_quality);
        st.clampIndirect = ps->clamp_indirect > 0.0f ? ps->clamp_indirect : 0.0f;
        st.motionBlur = ps->motion_blur != 0;
        st.aoStrength = ps->ao_strength;
        Tracer tr{scene, it, st, *cam, FilterSampler(ps->filter_type, ps->filter_width), width, height, ps->seed};
        Sobol sampler(ps->seed);
        RayCounters rc;
        for (size_t i = 0; i < n; ++i) {
            const V3 c = tr.trace(px[i], py[i], s[i], sampler, rc);
            L[3 * i] = c.x;
            L[3 * i + 1] = c.y;
            L[3 * i + 2] = c.z;
        }
        if (rays) {
            rays[0] = rc.extension;
            rays[1] = rc.shadow;
        }
        return 0;
    } catch (const std::exception&) {
        return -1;
    }
}

// Per-sample radiance and AOVs (7 floats: albedo rgb, normal xyz, alpha) for n (px, py, sample) triples.
ORACLE_API int oracle_trace_aovs(void* sp, const pe_camera* cam, const pe_settings* ps, int width, int height,
                                 size_t n, const int32_t* px, const int32_t* py, const int32_t* s, float* L,
                                 float* aov) {
    try {
        const Scene& scene = *static_cast<Scene*>(sp);
        const Integrator& it = integratorOf(scene);
        Settings st;
        st.maxDepth = ps->max_bounces;
        st.rrDepth = ps->rr_depth;
        st.shadowSamples = std::max(1, ps->shadow_quality);
        st.clampIndirect = ps->clamp_indirect > 0.0f ? ps->clamp_indirect : 0.0f;
        st.motionBlur = ps->motion_blur != 0;
        st.aoStrength = ps->ao_strength;
        Tracer tr{scene, it, st, *cam, FilterSampler(ps->filter_type, ps->filter_width), width, height, ps->seed};
        Sobol sampler(ps->seed);
        RayCounters rc;
        for (size_t i = 0; i < n; ++i) {
            const V3 c = tr.trace(px[i], py[i], s[i], sampler, rc, aov + 7 * i);
            L[3 * i] = c.x;
            L[3 * i + 1] = c.y;
            L[3 * i + 2] = c.z;
        }
        return 0;
    } catch (const std::exception&) {
        return -1;
    }
}

// Per-sample radiance and render passes (84 floats each, Tracer::tracePasses) for n (px, py, sample) triples.
ORACLE_API int oracle_trace_passes(void* sp, const pe_camera* cam, const pe_settings* ps, int width, int height,
                                   size_t n, const int32_t* px, const int32_t* py, const int32_t* s, float* L,
                                   float* out) {
    try {
        const Scene& scene = *static_cast<Scene*>(sp);
        const Integrator& it = integratorOf(scene);
        Settings st;
        st.maxDepth = ps->max_bounces;
        st.rrDepth = ps->rr_depth;
        st.shadowSamples = std::max(1, ps->shadow_quality);
        st.clampIndirect = ps->clamp_indirect > 0.0f ? ps->clamp_indirect : 0.0f;
        st.motionBlur = ps->motion_blur != 0;
        st.aoStrength = ps->ao_strength;
        Tracer tr{scene, it, st, *cam, FilterSampler(ps->filter_type, ps->filter_width), width, height, ps->seed};
        Sobol sampler(ps->seed);
        RayCounters rc;
        for (size_t i = 0; i < n; ++i) {
            const V3 c = tr.tracePasses(px[i], py[i], s[i], sampler, rc, out + 84 * i);
            L[3 * i] = c.x;
            L[3 * i + 1] = c.y;
            L[3 * i + 2] = c.z;
        }
        return 0;
    } catch (const std::exception&) {
        return -1;
    }
}

// Renderer::renderSamples (renderer.cpp:392-531) minus tiles/AOVs: every pixel takes samples
// [first, first + count) in order and Film::addSample (film.cpp:56-71) accumulates them. Rows are shared out
// over `threads` workers. Film arrays are accumulated into (not cleared).
ORACLE_API int oracle_render(void* sp, const pe_camera* cam, const pe_settings* ps, int width, int height, int first,
                             int count, int threads, int row_begin, int row_end, float* sum, uint32_t* cnt,
                             float* lum2, uint64_t* rays, double* seconds) {
    try {
        const Scene& scene = *static_cast<Scene*>(sp);
        const Integrator& it = integratorOf(scene);
        Settings st;
        st.maxDepth = ps->max_bounces;
        st.rrDepth = ps->rr_depth;
        st.shadowSamples = std::max(1, ps->shadow_quality);
        st.clampIndirect = ps->clamp_indirect > 0.0f ? ps->clamp_indirect : 0.0f;
        st.motionBlur = ps->motion_blur != 0;
        st.aoStrength = ps->ao_strength;
        const Tracer tr{scene, it, st, *cam, FilterSampler(ps->filter_type, ps->filter_width), width, height, ps->seed};
        if (row_end <= row_begin) {
            row_begin = 0;
            row_end = height;
        }
        std::atomic<int> nextRow{row_begin};
        std::atomic<uint64_t> ext{0}, sh{0};
        const auto t0 = std::chrono::steady_clock::now();
        auto worker = [&]() {
            Sobol sampler(ps->seed);
            RayCounters rc;
            for (int y; (y = nextRow.fetch_add(1)) < row_end;) {
                for (int x = 0; x < width; ++x) {
                    const size_t k = static_cast<size_t>(y) * width + x;
                    for (int s = first; s < first + count; ++s) {
                        const V3 c = tr.trace(x, y, s, sampler, rc);
                        ++cnt[k];
                        const bool ok = isFinite(c) && c.x >= 0.0f && c.y >= 0.0f && c.z >= 0.0f;
                        if (!ok) continue;
                        sum[3 * k] += c.x;
                        sum[3 * k + 1] += c.y;
                        sum[3 * k + 2] += c.z;
                        const float l = luminance(c);
                        lum2[k] += l * l;
                    }
                }
            }
            ext += rc.extension;
            sh += rc.shadow;
        };
        std::vector<std::thread> pool;
        for (int t = 1; t < std::max(1, threads); ++t) pool.emplace_back(worker);
        worker();
        for (auto& t : pool) t.join();
        if (seconds) *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (rays) {
            rays[0] = ext.load();
            rays[1] = sh.load();
        }
        return 0;
    } catch (const std::exception&) {
        return -1;
    }
}

ORACLE_API int oracle_bsdf_eval_at(const pe_material_desc* m, size_t n, const float* points, const float* normal,
                                   const float* tangent, const int32_t* frontFace, const float* wo, const float* wi,
                                   float* f, float* pdf) {
    try {
        for (size_t i = 0; i < n; ++i) {
            // As ref_driver's shadingPoint(): the geometric normal is the given normal; the point is the origin
            // unless given.
            const V3 nn(normal[3 * i], normal[3 * i + 1], normal[3 * i + 2]);
            ShadingPoint sp{nn, V3(tangent[3 * i], tangent[3 * i + 1], tangent[3 * i + 2]), frontFace[i] != 0, nn, 0.0f,
                            points ? V3(points[3 * i], points[3 * i + 1], points[3 * i + 2]) : V3(0.0f)};
            float p = 0.0f;
            const V3 v = evalPdf(*m, V3(wo[3 * i], wo[3 * i + 1], wo[3 * i + 2]), V3(wi[3 * i], wi[3 * i + 1], wi[3 * i + 2]),
                                 sp, p);
            f[3 * i] = v.x;
            f[3 * i + 1] = v.y;
            f[3 * i + 2] = v.z;
            pdf[i] = p;
        }
        return 0;
    } catch (const std::exception&) {
        return -1;
    }
}

ORACLE_API int oracle_flake_resolve(const pe_material_desc* m, size_t n, const float* points, const float* normal,
                                    const float* tangent, const int32_t* frontFace, float* out) {
    for (size_t i = 0; i < n; ++i) {
        const V3 nn(normal[3 * i], normal[3 * i + 1], normal[3 * i + 2]);
        const ShadingPoint sp{nn, V3(tangent[3 * i], tangent[3 * i + 1], tangent[3 * i + 2]), frontFace[i] != 0, nn,
                              0.0f, V3(points[3 * i], points[3 * i + 1], points[3 * i + 2])};
        bool in = false;
        V3 fn;
        flakeState(*m, sp, in, fn);
        out[4 * i] = in ? 1.0f : 0.0f;
        out[4 * i + 1] = fn.x;
        out[4 * i + 2] = fn.y;
        out[4 * i + 3] = fn.z;
    }
    return 0;
}

ORACLE_API int oracle_bsdf_eval(const pe_material_desc* m, size_t n, const float* normal, const float* tangent,
                                const int32_t* frontFace, const float* wo, const float* wi, float* f, float* pdf) {
    return oracle_bsdf_eval_at(m, n, nullptr, normal, tangent, frontFace, wo, wi, f, pdf);
}

ORACLE_API int oracle_bsdf_sample_at(const pe_material_desc* m, size_t n, const float* points, const float* normal,
                                     const float* tangent, const int32_t* frontFace, const float* wo, const float* u3,
                                     float* out) {
    try {
        for (size_t i = 0; i < n; ++i) {
            const V3 nn(normal[3 * i], normal[3 * i + 1], normal[3 * i + 2]);
            ShadingPoint sp{nn, V3(tangent[3 * i], tangent[3 * i + 1], tangent[3 * i + 2]), frontFace[i] != 0, nn, 0.0f,
                            points ? V3(points[3 * i], points[3 * i + 1], points[3 * i + 2]) : V3(0.0f)};
            BsdfSample bs;
            const bool ok = sampleFull(*m, V3(wo[3 * i], wo[3 * i + 1], wo[3 * i + 2]), sp, u3[3 * i], u3[3 * i + 1],
                                       u3[3 * i + 2], bs);
            float* o = out + 14 * i;
            o[0] = ok ? 1.0f : 0.0f;
            o[1] = bs.wi.x; o[2] = bs.wi.y; o[3] = bs.wi.z;
            o[4] = bs.f.x; o[5] = bs.f.y; o[6] = bs.f.z;
            o[7] = bs.pdf;
            o[8] = bs.delta ? 1.0f : 0.0f;
            o[9] = bs.subsurface ? 1.0f : 0.0f;
            o[10] = bs.passThrough ? 1.0f : 0.0f;
            o[11] = bs.weight.x; o[12] = bs.weight.y; o[13] = bs.weight.z;
        }
        return 0;
    } catch (const std::exception&) {
        return -1;
    }
}

ORACLE_API int oracle_bsdf_sample(const pe_material_desc* m, size_t n, const float* normal, const float* tangent,
                                  const int32_t* frontFace, const float* wo, const float* u3, float* out) {
    return oracle_bsdf_sample_at(m, n, nullptr, normal, tangent, frontFace, wo, u3, out);
}

// which: 0 ggxDirectionalAlbedo, 1 ggxSchlickAlbedo, 2 ggxSchlickAverageAlbedo, 3 charlieSheenAlbedo
ORACLE_API int oracle_microfacet(int which, size_t n, const float* a, const float* b, const float* c, float* out) {
    for (size_t i = 0; i < n; ++i) {
        switch (which) {
            case 0: out[i] = ggxDirectionalAlbedo(a[i], b[i]); break;
            case 1: out[i] = ggxSchlickAlbedo(a[i], b[i], c[i]); break;
            case 2: out[i] = ggxSchlickAverageAlbedo(a[i], b[i]); break;
            case 3: out[i] = charlieSheenAlbedo(a[i], b[i]); break;
            default: return -1;
        }
    }
    return 0;
}

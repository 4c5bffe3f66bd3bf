// light_bvh.cpp — host build of the light BVH the gfx950 light picker descends (LightSampler::build,
// lights/light_sampler.cpp:42-148; LightBounds, lights/light_bounds.cpp:39-109; light bounds of area and triangle
// lights, lights/area_light.cpp:219-228, :260-273).
//
// The tree decides which light a shading point samples and with what pmf, so it must be the reference's tree:
// the same bucketed cost, the same std::partition / std::nth_element calls on the same sequence, the same
// depth-first node order, every float operation in the reference's order and without contraction (this file is
// compiled with -ffp-contract=off). The device only reads the result.
#include "light_bvh.h"
#include "photon_hip.h"

#include <algorithm>
#include <cmath>
#include <limits>

namespace pe {

namespace {

constexpr float kUniformShare = 0.15f;
constexpr int kBuckets = 12;
constexpr int kTreeMaxDepth = 63;
constexpr int kExhaustiveMax = 33;

struct Box {  // AABB with std::min/std::max corners (aabb.h), empty = [+max, -max]
    V3 lo = splat(std::numeric_limits<float>::max());
    V3 hi = splat(-std::numeric_limits<float>::max());
    static V3 vmin_first(V3 a, V3 b) { return mk(std::min(a.x, b.x), std::min(a.y, b.y), std::min(a.z, b.z)); }
    static V3 vmax_first(V3 a, V3 b) { return mk(std::max(a.x, b.x), std::max(a.y, b.y), std::max(a.z, b.z)); }
    static Box of(V3 a, V3 b) {
        Box r;
        r.lo = vmin_first(a, b);
        r.hi = vmax_first(a, b);
        return r;
    }
    void grow(const Box& b) {
        lo = vmin_first(lo, b.lo);
        hi = vmax_first(hi, b.hi);
    }
    void grow(V3 p) {
        lo = vmin_first(lo, p);
        hi = vmax_first(hi, p);
    }
    Box merged(const Box& b) const { return of(vmin_first(lo, b.lo), vmax_first(hi, b.hi)); }
    V3 centroid() const { return (lo + hi) * 0.5f; }
    V3 diagonal() const { return hi - lo; }
    float area() const {
        const V3 d = diagonal();
        return 2.0f * (d.x * d.y + d.x * d.z + d.y * d.z);
    }
    int widest() const {
        const V3 d = diagonal();
        if (d.x > d.y && d.x > d.z) return 0;
        if (d.y > d.z) return 1;
        return 2;
    }
};

struct Bounds {  // LightBounds
    Box box;
    V3 w = mk(0.0f, 0.0f, 1.0f);
    float phi = 0.0f, cos_o = 1.0f, cos_e = 0.0f;
    bool two_sided = false;
};

float acos_clamped(float x) { return std::acos(std::clamp(x, -1.0f, 1.0f)); }

V3 rodrigues(V3 v, V3 axis, float theta) {
    const float c = std::cos(theta), s = std::sin(theta);
    return v * c + cross(axis, v) * s + axis * (dot(axis, v) * (1.0f - c));
}

V3 unit(V3 v) {  // Vec3f::normalized (vec.h:153-157)
    const float len = std::sqrt(dot(v, v));
    if (len <= 1e-20f) return splat(0.0f);
    const float inv = 1.0f / len;
    return mk(v.x * inv, v.y * inv, v.z * inv);
}

Bounds join(const Bounds& a, const Bounds& b) {  // unionBounds
    if (!(a.phi > 0.0f)) return b;
    if (!(b.phi > 0.0f)) return a;
    Bounds u;
    u.box = a.box.merged(b.box);
    u.phi = a.phi + b.phi;
    u.cos_e = std::min(a.cos_e, b.cos_e);
    u.two_sided = a.two_sided || b.two_sided;
    const float ta = acos_clamped(a.cos_o), tb = acos_clamped(b.cos_o);
    const float td = acos_clamped(dot(a.w, b.w));
    if (std::min(td + tb, kPi) <= ta) {
        u.w = a.w;
        u.cos_o = a.cos_o;
        return u;
    }
    if (std::min(td + ta, kPi) <= tb) {
        u.w = b.w;
        u.cos_o = b.cos_o;
        return u;
    }
    const float t_o = 0.5f * (ta + td + tb);
    const V3 axis = cross(a.w, b.w);
    if (t_o >= kPi || dot(axis, axis) < 1e-12f) {
        u.w = a.w;
        u.cos_o = -1.0f;
        return u;
    }
    u.w = unit(rodrigues(a.w, unit(axis), t_o - ta));
    u.cos_o = std::cos(t_o);
    return u;
}

float split_cost(const Bounds& b, const Box& parent, int dim) {  // groupCost
    const float t_o = std::acos(std::clamp(b.cos_o, -1.0f, 1.0f));
    const float t_e = std::acos(std::clamp(b.cos_e, -1.0f, 1.0f));
    const float t_w = std::min(t_o + t_e, kPi);
    const float sin_o = std::sqrt(std::max(0.0f, 1.0f - b.cos_o * b.cos_o));
    const float m_omega = kTwoPi * (1.0f - b.cos_o) +
                          0.5f * kPi * (2.0f * t_w * sin_o - std::cos(t_o - 2.0f * t_w) - 2.0f * t_o * sin_o + b.cos_o);
    const V3 d = parent.diagonal();
    const float max_d = std::max({d.x, d.y, d.z});
    const float kr = comp(d, dim) > 0.0f ? max_d / comp(d, dim) : 1.0f;
    return b.phi * m_omega * kr * b.box.area();
}

bool bounds_of(const GpuLight& L, Bounds& out) {
    if (!(L.area > 0.0f) || is_black(L.radiance)) return false;
    if (L.type == PE_LIGHT_AREA) {
        Box b = Box::of(L.p0, L.p0);
        b.grow(L.p0 + L.p1);
        b.grow(L.p0 + L.p2);
        b.grow(L.p0 + L.p1 + L.p2);
        out.box = b;
    } else {
        out.box = Box::of(Box::vmin_first(Box::vmin_first(L.p0, L.p1), L.p2),
                          Box::vmax_first(Box::vmax_first(L.p0, L.p1), L.p2));
    }
    out.w = L.normal;
    out.phi = std::max(0.0f, luminance(L.radiance)) * L.area * kPi;
    out.cos_o = 1.0f;
    out.cos_e = 0.0f;
    out.two_sided = true;
    return true;
}

struct Entry {
    int index;
    Bounds b;
};

class Builder {
public:
    Builder(LightBvh& out) : m_out(out) {}

    int node(std::vector<Entry>& e, int begin, int end, uint64_t trail, int depth) {
        const int at = static_cast<int>(m_bounds.size());
        m_bounds.emplace_back();
        m_out.nodes.emplace_back();
        if (end - begin == 1) {
            const Entry& it = e[static_cast<size_t>(begin)];
            m_bounds[static_cast<size_t>(at)] = it.b;
            m_out.nodes[static_cast<size_t>(at)].ref = make_int4(-1, it.index, 0, 0);
            LightPath& lp = m_out.paths[static_cast<size_t>(it.index)];
            lp.trail_lo = static_cast<uint32_t>(trail);
            lp.trail_hi = static_cast<uint32_t>(trail >> 32);
            lp.depth = depth;
            return at;
        }
        Box all, cents;
        for (int i = begin; i < end; ++i) {
            all.grow(e[static_cast<size_t>(i)].b.box);
            cents.grow(e[static_cast<size_t>(i)].b.box.centroid());
        }
        float best = std::numeric_limits<float>::infinity();
        int best_dim = -1, best_split = -1;
        if (depth < kTreeMaxDepth - 8) {
            for (int dim = 0; dim < 3; ++dim) {
                const float lo = comp(cents.lo, dim), hi = comp(cents.hi, dim);
                if (!(hi > lo)) continue;
                Bounds bucket[kBuckets];
                for (int i = begin; i < end; ++i) {
                    const Bounds& b = e[static_cast<size_t>(i)].b;
                    int k = static_cast<int>(kBuckets * (comp(b.box.centroid(), dim) - lo) / (hi - lo));
                    k = std::clamp(k, 0, kBuckets - 1);
                    bucket[k] = join(bucket[k], b);
                }
                for (int split = 0; split < kBuckets - 1; ++split) {
                    Bounds below, above;
                    for (int k = 0; k <= split; ++k) below = join(below, bucket[k]);
                    for (int k = split + 1; k < kBuckets; ++k) above = join(above, bucket[k]);
                    if (!(below.phi > 0.0f) || !(above.phi > 0.0f)) continue;
                    const float c = split_cost(below, all, dim) + split_cost(above, all, dim);
                    if (c > 0.0f && c < best) {
                        best = c;
                        best_dim = dim;
                        best_split = split;
                    }
                }
            }
        }
        int mid = -1;
        if (best_dim >= 0) {
            const float lo = comp(cents.lo, best_dim), hi = comp(cents.hi, best_dim);
            auto it = std::partition(e.begin() + begin, e.begin() + end, [&](const Entry& x) {
                const int k = static_cast<int>(kBuckets * (comp(x.b.box.centroid(), best_dim) - lo) / (hi - lo));
                return std::clamp(k, 0, kBuckets - 1) <= best_split;
            });
            mid = static_cast<int>(it - e.begin());
        }
        if (mid <= begin || mid >= end) {
            mid = (begin + end) / 2;
            const int dim = cents.widest();
            std::nth_element(e.begin() + begin, e.begin() + mid, e.begin() + end, [dim](const Entry& a, const Entry& b) {
                return comp(a.b.box.centroid(), dim) < comp(b.b.box.centroid(), dim);
            });
        }
        const int c0 = node(e, begin, mid, trail, depth + 1);
        const int c1 = node(e, mid, end, trail | (uint64_t{1} << depth), depth + 1);
        m_bounds[static_cast<size_t>(at)] = join(m_bounds[static_cast<size_t>(c0)], m_bounds[static_cast<size_t>(c1)]);
        m_out.nodes[static_cast<size_t>(at)].ref = make_int4(c1, -1, 0, 0);
        return at;
    }

    void finish() {
        for (size_t i = 0; i < m_bounds.size(); ++i) {
            const Bounds& b = m_bounds[i];
            GpuLightNode& n = m_out.nodes[i];
            n.lo_phi = make_float4(b.box.lo.x, b.box.lo.y, b.box.lo.z, b.phi);
            n.hi_cos_o = make_float4(b.box.hi.x, b.box.hi.y, b.box.hi.z, b.cos_o);
            n.w_cos_e = make_float4(b.w.x, b.w.y, b.w.z, b.cos_e);
            n.ref.z = b.two_sided ? 1 : 0;
        }
    }

private:
    LightBvh& m_out;
    std::vector<Bounds> m_bounds;
};

}  // namespace

void build_light_bvh(const std::vector<GpuLight>& lights, int num_lights, bool has_env, LightBvh& out) {
    out = LightBvh{};
    const int count = num_lights + (has_env ? 1 : 0);
    out.paths.assign(static_cast<size_t>(std::max(count, 1)), LightPath{0u, 0u, 0, -1});
    out.mode = count <= 1 ? kPickUniform : kPickExhaustive;
    if (count <= kExhaustiveMax) return;
    std::vector<Entry> items;
    items.reserve(static_cast<size_t>(num_lights));
    for (int i = 0; i < num_lights; ++i) {
        Bounds b;
        if (bounds_of(lights[static_cast<size_t>(i)], b) && b.phi > 0.0f) items.push_back({i, b});
        else out.top.push_back(i);
    }
    if (has_env) out.top.push_back(num_lights);
    if (static_cast<int>(out.top.size()) + 1 > kMaxTopCandidates) {
        // Too many unbounded candidates for the top level: the reference samples uniformly.
        out.top.clear();
        out.mode = kPickUniform;
        return;
    }
    for (size_t k = 0; k < out.top.size(); ++k) out.paths[static_cast<size_t>(out.top[k])].top_slot = static_cast<int>(k);
    if (!items.empty()) {
        out.nodes.reserve(2 * items.size());
        Builder b(out);
        b.node(items, 0, static_cast<int>(items.size()), 0, 0);
        b.finish();
    }
    out.mode = kPickBvh;
    (void)kUniformShare;
}

}  // namespace pe

// tables.cpp — host construction of the microfacet energy tables, the same deterministic stratified
// integrations the reference runs at first use (materials/microfacet.cpp:42-133): 24 x 24 VNDF samples per
// (mu, alpha) node for E / Ea / Eb, a 256-point midpoint rule for the hemispherical averages, a 96 x 48
// (cos theta, phi) midpoint grid for the Charlie-sheen albedo. Built once per scene upload (a few ms) and
// copied to HBM. Compiled without contraction so the floats equal the reference's IEEE build.
#include "pe_microfacet.h"

#include <cmath>

namespace pe {

namespace {
float node_mu(int im) { return std::fmax(static_cast<float>(im) / (kTableN - 1), 1e-3f); }

// spectral.cpp Tables (:34-62): channel means by a 3500-step midpoint rule, then a 350-bin CDF of r + g + b + 0.02.
void build_spectral_tables(EnergyTables& T) {
    constexpr float lmin = 380.0f, lmax = 730.0f;
    V3 sum = splat(0.0f);
    constexpr int kSteps = 3500;
    const float dl = (lmax - lmin) / static_cast<float>(kSteps);
    for (int i = 0; i < kSteps; ++i) sum = sum + spectral_rgb_raw(lmin + (static_cast<float>(i) + 0.5f) * dl);
    const float inv_steps = 1.0f / static_cast<float>(kSteps);
    const V3 mean = mk(sum.x * inv_steps, sum.y * inv_steps, sum.z * inv_steps);
    const V3 norm = mk(1.0f / mean.x, 1.0f / mean.y, 1.0f / mean.z);
    T.spec_norm[0] = norm.x;
    T.spec_norm[1] = norm.y;
    T.spec_norm[2] = norm.z;
    const float bw = (lmax - lmin) / 350.0f;
    float acc = 0.0f;
    float s[350];
    T.spec_cdf[0] = 0.0f;
    for (int i = 0; i < 350; ++i) {
        const V3 c = spectral_rgb_raw(lmin + (static_cast<float>(i) + 0.5f) * bw) * norm;
        s[i] = c.x + c.y + c.z + 0.02f;
        acc += s[i];
        T.spec_cdf[i + 1] = acc;
    }
    for (int i = 0; i <= 350; ++i) T.spec_cdf[i] /= acc;
    for (int i = 0; i < 350; ++i) T.spec_pdf[i] = s[i] / (acc * bw);
}
}  // namespace

void build_energy_tables(EnergyTables& T) {
    build_spectral_tables(T);
    constexpr int kS = 24;
    for (int ia = 0; ia < kTableN; ++ia) {
        const float alpha = static_cast<float>(ia) / (kTableN - 1);
        for (int im = 0; im < kTableN; ++im) {
            const float mu = node_mu(im);
            const V3 wo = mk(std::sqrt(std::fmax(0.0f, 1.0f - mu * mu)), 0.0f, mu);
            const float g1 = ggx_g1(wo, alpha, alpha);
            double sum = 0.0, sum_a = 0.0, sum_b = 0.0;
            for (int sy = 0; sy < kS; ++sy) {
                for (int sx = 0; sx < kS; ++sx) {
                    const float u1 = (static_cast<float>(sx) + 0.5f) / kS;
                    const float u2 = (static_cast<float>(sy) + 0.5f) / kS;
                    const V3 h = ggx_sample_visible_normal(wo, alpha, alpha, u1, u2);
                    const float c = dot(wo, h);
                    const V3 wi = h * (2.0f * c) - wo;
                    if (wi.z <= 0.0f) continue;
                    const double w = ggx_g(wo, wi, alpha, alpha) / g1;
                    const double s5 = pow5(1.0f - clampf(c, 0.0f, 1.0f));
                    sum += w;
                    sum_a += w * (1.0 - s5);
                    sum_b += w * s5;
                }
            }
            const int k = ia * kTableN + im;
            T.e[k] = static_cast<float>(sum / (kS * kS));
            T.ea[k] = static_cast<float>(sum_a / (kS * kS));
            T.eb[k] = static_cast<float>(sum_b / (kS * kS));
        }
    }
    constexpr int kM = 256;
    for (int ia = 0; ia < kTableN; ++ia) {
        const float alpha = static_cast<float>(ia) / (kTableN - 1);
        double a0 = 0.0, a1 = 0.0, a2 = 0.0;
        for (int i = 0; i < kM; ++i) {
            const float mu = (static_cast<float>(i) + 0.5f) / kM;
            const float E = clampf(table_lookup2d(T.e, mu, alpha), 1e-3f, 1.0f);
            const float m = 1.0f / E - 1.0f;
            const float A = table_lookup2d(T.ea, mu, alpha), B = table_lookup2d(T.eb, mu, alpha);
            const double w = 2.0 * mu / kM;
            a0 += w * B;
            a1 += w * (A + B * m);
            a2 += w * (A * m);
        }
        T.avg0[ia] = static_cast<float>(a0);
        T.avg1[ia] = static_cast<float>(a1);
        T.avg2[ia] = static_cast<float>(a2);
    }
    constexpr int kC = 96, kP = 48;
    for (int ir = 0; ir < kTableN; ++ir) {
        const float r = std::fmax(static_cast<float>(ir) / (kTableN - 1), 0.02f);
        for (int im = 0; im < kTableN; ++im) {
            const float mu = node_mu(im);
            const V3 wo = mk(std::sqrt(std::fmax(0.0f, 1.0f - mu * mu)), 0.0f, mu);
            double sum = 0.0;
            for (int i = 0; i < kC; ++i) {
                const float cz = (static_cast<float>(i) + 0.5f) / kC;
                const float s = std::sqrt(std::fmax(0.0f, 1.0f - cz * cz));
                for (int j = 0; j < kP; ++j) {
                    const float phi = kPi * (static_cast<float>(j) + 0.5f) / kP;
                    const V3 wi = mk(s * std::cos(phi), s * std::sin(phi), cz);
                    const V3 h = normalized(wo + wi);
                    sum += charlie_d(h.z, r) * charlie_v(cz, mu) * cz;
                }
            }
            T.sheen[ir * kTableN + im] = static_cast<float>(sum * (1.0 / kC) * (kTwoPi / kP));
        }
    }
}

}  // namespace pe

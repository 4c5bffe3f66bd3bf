// oracle_display.cpp — CPU restatement of the display path (test infrastructure only: the oracle).
//
// Film::resolve (src/core/image/film.cpp:87-94, :139-147), applyImager and its helpers
// (src/core/image/imager.cpp: sampleClamped, downsample2, gaussianBlur, brightPass, bloomImage, glareImage,
// whiteBalanceMatrix with the Planck / daylight locus, the per-pixel pass), toneMap / toneMapLinear and the seven
// operators (src/core/image/tone_mapping.cpp:22-160), srgbEncode / srgbDecode, quantizeUnorm8 and tpdfDither
// (src/core/color/transfer.cpp:13-52), in the reference's operation order, IEEE single precision without
// contraction (built with -ffp-contract=off).
#include "oracle.h"
#include "photon_hip.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#define ORACLE_API extern "C" __attribute__((visibility("default")))

namespace {

struct C3 {
    float r = 0, g = 0, b = 0;
    C3() = default;
    C3(float r_, float g_, float b_) : r(r_), g(g_), b(b_) {}
    explicit C3(float v) : r(v), g(v), b(v) {}
    C3 operator+(const C3& o) const { return {r + o.r, g + o.g, b + o.b}; }
    C3 operator-(const C3& o) const { return {r - o.r, g - o.g, b - o.b}; }
    C3 operator*(float s) const { return {r * s, g * s, b * s}; }
    C3 operator/(float s) const {
        const float inv = 1.0f / s;
        return {r * inv, g * inv, b * inv};
    }
    C3& operator+=(const C3& o) {
        r += o.r;
        g += o.g;
        b += o.b;
        return *this;
    }
    float luminance() const { return 0.212671f * r + 0.715160f * g + 0.072169f * b; }
};

struct Img {
    int w = 0, h = 0;
    std::vector<float> px;  // rgb
    Img() = default;
    Img(int w_, int h_) : w(w_), h(h_), px(static_cast<size_t>(w_) * h_ * 3, 0.0f) {}
    C3 get(int x, int y) const {
        const float* p = &px[(static_cast<size_t>(y) * w + x) * 3];
        return {p[0], p[1], p[2]};
    }
    void set(int x, int y, const C3& c) {
        float* p = &px[(static_cast<size_t>(y) * w + x) * 3];
        p[0] = c.r;
        p[1] = c.g;
        p[2] = c.b;
    }
};

bool finite_f(float x) {
    uint32_t u;
    std::memcpy(&u, &x, 4);
    return (u & 0x7f800000u) != 0x7f800000u;
}

// ---- transfer.cpp:13-52
float srgb_decode(float e) {
    if (e <= 0.04045f) return e / 12.92f;
    return std::pow((e + 0.055f) / 1.055f, 2.4f);
}
float srgb_encode(float l) {
    l = std::clamp(l, 0.0f, 1.0f);
    if (l <= 0.0031308f) return 12.92f * l;
    return 1.055f * std::pow(l, 1.0f / 2.4f) - 0.055f;
}
uint8_t quantize_unorm8(float v, float dither) {
    const float q = std::floor(v * 255.0f + dither + 0.5f);
    return static_cast<uint8_t>(std::clamp(q, 0.0f, 255.0f));
}
float tpdf_dither(int x, int y, int channel) {
    uint32_t h = static_cast<uint32_t>(x) * 0x8da6b343u ^ static_cast<uint32_t>(y) * 0xd8163841u ^
                 static_cast<uint32_t>(channel) * 0xcb1ab31fu;
    h ^= h >> 16;
    h *= 0x7feb352du;
    h ^= h >> 15;
    h *= 0x846ca68bu;
    h ^= h >> 16;
    const float u0 = static_cast<float>(h & 0xffffu) * (1.0f / 65536.0f);
    const float u1 = static_cast<float>(h >> 16) * (1.0f / 65536.0f);
    return u0 - u1;
}

// ---- tone_mapping.cpp
float aces(float x) {
    const float a = 2.51f, b = 0.03f, c = 2.43f, d = 0.59f, e = 0.14f;
    return std::clamp((x * (a * x + b)) / (x * (c * x + d) + e), 0.0f, 1.0f);
}
float hable(float x) {
    const float A = 0.15f, B = 0.50f, C = 0.10f, D = 0.20f, E = 0.02f, F = 0.30f;
    return ((x * (A * x + C * B) + D * E) / (x * (A * x + B) + D * F)) - E / F;
}
C3 mat_col(const float m[9], const C3& v) {  // column-major 3x3 (GLSL mat3)
    return {m[0] * v.r + m[3] * v.g + m[6] * v.b, m[1] * v.r + m[4] * v.g + m[7] * v.b,
            m[2] * v.r + m[5] * v.g + m[8] * v.b};
}
C3 agx(const C3& hdr) {
    static const float inset[9] = {0.842479062253094f, 0.0423282422610123f, 0.0423756549057051f,
                                   0.0784335999999992f, 0.878468636469772f,  0.0784336f,
                                   0.0792237451477643f, 0.0791661274605434f, 0.879142973793104f};
    static const float outset[9] = {1.19687900512017f,   -0.0528968517574562f, -0.0529716355144438f,
                                    -0.0980208811401368f, 1.15190312990417f,   -0.0980434501171241f,
                                    -0.0990297440797205f, -0.0989611768448433f, 1.15107367264116f};
    const float min_ev = -12.47393f, max_ev = 4.026069f;
    auto curve = [&](float x) {
        x = std::clamp(x, min_ev, max_ev);
        x = (x - min_ev) / (max_ev - min_ev);
        const float x2 = x * x, x4 = x2 * x2;
        return 15.5f * x4 * x2 - 40.14f * x4 * x + 31.96f * x4 - 6.868f * x2 * x + 0.4298f * x2 + 0.1191f * x - 0.00232f;
    };
    auto lg = [](float c) { return std::log2(std::max(c, 1e-10f)); };
    C3 v = mat_col(inset, hdr);
    v = C3(curve(lg(v.r)), curve(lg(v.g)), curve(lg(v.b)));
    v = mat_col(outset, v);
    return {srgb_decode(std::clamp(v.r, 0.0f, 1.0f)), srgb_decode(std::clamp(v.g, 0.0f, 1.0f)),
            srgb_decode(std::clamp(v.b, 0.0f, 1.0f))};
}
C3 pbr_neutral(const C3& hdr) {
    const float start = 0.8f - 0.04f, desat = 0.15f;
    C3 c = hdr;
    const float x = std::min(c.r, std::min(c.g, c.b));
    const float offset = x < 0.08f ? x - 6.25f * x * x : 0.04f;
    c = c - C3(offset);
    const float peak = std::max(c.r, std::max(c.g, c.b));
    if (peak < start) return c;
    const float d = 1.0f - start;
    const float new_peak = 1.0f - d * d / (peak + d - start);
    c = c * (new_peak / peak);
    const float g = 1.0f - 1.0f / (desat * (peak - new_peak) + 1.0f);
    return c * (1.0f - g) + C3(new_peak) * g;
}
C3 tonemap_linear(const C3& hdr, int op, float ev) {
    const float scale = std::exp2(ev);
    C3 e = hdr * scale;
    auto sane = [](float v) { return finite_f(v) && v > 0.0f ? v : 0.0f; };
    e = C3(sane(e.r), sane(e.g), sane(e.b));
    C3 l;
    switch (op) {
        case PE_TONEMAP_REINHARD: l = C3(e.r / (1.0f + e.r), e.g / (1.0f + e.g), e.b / (1.0f + e.b)); break;
        case PE_TONEMAP_REINHARD_EXTENDED: {
            const float mw2 = 4.0f * 4.0f;
            auto ch = [mw2](float v) { return (v * (1.0f + v / mw2)) / (1.0f + v); };
            l = C3(ch(e.r), ch(e.g), ch(e.b));
            break;
        }
        case PE_TONEMAP_ACES: l = C3(aces(e.r), aces(e.g), aces(e.b)); break;
        case PE_TONEMAP_FILMIC: {
            const float ws = 1.0f / hable(11.2f);
            l = C3(hable(e.r) * ws, hable(e.g) * ws, hable(e.b) * ws);
            break;
        }
        case PE_TONEMAP_AGX: l = agx(e); break;
        case PE_TONEMAP_PBR_NEUTRAL: l = pbr_neutral(e); break;
        default: l = e; break;
    }
    return C3(std::max(0.0f, std::min(l.r, 1.0f)), std::max(0.0f, std::min(l.g, 1.0f)), std::max(0.0f, std::min(l.b, 1.0f)));
}
C3 tonemap(const C3& hdr, int op, float ev) {
    const C3 l = tonemap_linear(hdr, op, ev);
    return {srgb_encode(l.r), srgb_encode(l.g), srgb_encode(l.b)};
}

// ---- imager.cpp
C3 sample_clamped(const Img& img, float x, float y) {
    x = std::clamp(x - 0.5f, 0.0f, static_cast<float>(img.w - 1));
    y = std::clamp(y - 0.5f, 0.0f, static_cast<float>(img.h - 1));
    const int x0 = static_cast<int>(x), y0 = static_cast<int>(y);
    const int x1 = std::min(x0 + 1, img.w - 1), y1 = std::min(y0 + 1, img.h - 1);
    const float fx = x - static_cast<float>(x0), fy = y - static_cast<float>(y0);
    const C3 a = img.get(x0, y0) * (1.0f - fx) + img.get(x1, y0) * fx;
    const C3 b = img.get(x0, y1) * (1.0f - fx) + img.get(x1, y1) * fx;
    return a * (1.0f - fy) + b * fy;
}
Img downsample2(const Img& src) {
    Img out(std::max(1, src.w / 2), std::max(1, src.h / 2));
    for (int y = 0; y < out.h; ++y)
        for (int x = 0; x < out.w; ++x) {
            const int sx = std::min(2 * x, src.w - 1), sy = std::min(2 * y, src.h - 1);
            const int sx1 = std::min(sx + 1, src.w - 1), sy1 = std::min(sy + 1, src.h - 1);
            out.set(x, y, (src.get(sx, sy) + src.get(sx1, sy) + src.get(sx, sy1) + src.get(sx1, sy1)) * 0.25f);
        }
    return out;
}
void gaussian_blur(Img& img, float sigma) {
    if (sigma < 0.3f) return;
    const int r = std::max(1, static_cast<int>(std::ceil(3.0f * sigma)));
    std::vector<float> k(static_cast<size_t>(2 * r + 1));
    float sum = 0.0f;
    for (int i = -r; i <= r; ++i) sum += k[static_cast<size_t>(i + r)] = std::exp(-0.5f * (i * i) / (sigma * sigma));
    for (float& v : k) v /= sum;
    Img tmp(img.w, img.h);
    for (int y = 0; y < img.h; ++y)
        for (int x = 0; x < img.w; ++x) {
            C3 acc;
            for (int i = -r; i <= r; ++i) acc += img.get(std::clamp(x + i, 0, img.w - 1), y) * k[static_cast<size_t>(i + r)];
            tmp.set(x, y, acc);
        }
    for (int y = 0; y < img.h; ++y)
        for (int x = 0; x < img.w; ++x) {
            C3 acc;
            for (int i = -r; i <= r; ++i) acc += tmp.get(x, std::clamp(y + i, 0, img.h - 1)) * k[static_cast<size_t>(i + r)];
            img.set(x, y, acc);
        }
}
Img bright_pass(const Img& hdr, float threshold, float exposure_scale) {
    Img half = downsample2(hdr);
    for (int y = 0; y < half.h; ++y)
        for (int x = 0; x < half.w; ++x) {
            const C3 c = half.get(x, y);
            const float L = c.luminance() * exposure_scale;
            const float f = L > threshold ? (L - threshold) / L : 0.0f;
            half.set(x, y, c * f);
        }
    return half;
}
Img bloom_image(const Img& bright, float sigma_half) {
    Img acc(bright.w, bright.h);
    const float weights[3] = {0.5f, 0.3f, 0.2f};
    for (int o = 0; o < 3; ++o) {
        float sigma = sigma_half * static_cast<float>(1 << o);
        Img level = bright;
        int scale = 1;
        while (sigma > 6.0f && level.w > 8 && level.h > 8) {
            level = downsample2(level);
            sigma *= 0.5f;
            scale *= 2;
        }
        gaussian_blur(level, sigma);
        for (int y = 0; y < acc.h; ++y)
            for (int x = 0; x < acc.w; ++x) {
                const C3 c = scale == 1 ? level.get(x, y)
                                        : sample_clamped(level, (x + 0.5f) / scale, (y + 0.5f) / scale);
                acc.set(x, y, acc.get(x, y) + c * weights[o]);
            }
    }
    return acc;
}
Img glare_image(const Img& bright, int blades, float angle_deg, float length_px) {
    const Img src = downsample2(bright);
    const float L = std::max(2.0f, length_px * 0.25f);
    const int lines = std::max(1, (blades + 1) / 2);
    const int steps = static_cast<int>(std::ceil(L));
    Img out(src.w, src.h);
    std::vector<float> wts(static_cast<size_t>(steps + 1));
    float wsum = 0.0f;
    for (int t = 0; t <= steps; ++t) {
        wts[static_cast<size_t>(t)] = std::exp(-3.0f * static_cast<float>(t) / L);
        wsum += (t == 0 ? 1.0f : 2.0f) * wts[static_cast<size_t>(t)];
    }
    const float deg = 3.14159265358979323846f / 180.0f, pi = 3.14159265358979323846f;
    for (int l = 0; l < lines; ++l) {
        const float a = angle_deg * deg + static_cast<float>(l) * pi / static_cast<float>(lines);
        const float dx = std::cos(a), dy = std::sin(a);
        for (int y = 0; y < src.h; ++y)
            for (int x = 0; x < src.w; ++x) {
                C3 acc = src.get(x, y) * wts[0];
                for (int t = 1; t <= steps; ++t) {
                    const float w = wts[static_cast<size_t>(t)];
                    acc += sample_clamped(src, x + 0.5f + dx * t, y + 0.5f + dy * t) * w;
                    acc += sample_clamped(src, x + 0.5f - dx * t, y + 0.5f - dy * t) * w;
                }
                out.set(x, y, out.get(x, y) + acc / (wsum * static_cast<float>(lines)));
            }
    }
    return out;
}

const float kRgbToXyz[9] = {0.4124564f, 0.3575761f, 0.1804375f, 0.2126729f, 0.7151522f,
                            0.0721750f, 0.0193339f, 0.1191920f, 0.9503041f};
const float kXyzToRgb[9] = {3.2404542f, -1.5371385f, -0.4985314f, -0.9692660f, 1.8760108f,
                            0.0415560f, 0.0556434f, -0.2040259f, 1.0572252f};
const float kBradford[9] = {0.8951f, 0.2664f, -0.1614f, -0.7502f, 1.7135f, 0.0367f, 0.0389f, -0.0685f, 1.0296f};
const float kBradfordInv[9] = {0.9869929f, -0.1470543f, 0.1599627f, 0.4323053f, 0.5183603f,
                               0.0492912f, -0.0085287f, 0.0400428f, 0.9684867f};

void mul3(const float a[9], const float b[9], float out[9]) {
    float r[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) r[i * 3 + j] = a[i * 3] * b[j] + a[i * 3 + 1] * b[3 + j] + a[i * 3 + 2] * b[6 + j];
    std::copy(r, r + 9, out);
}
void planck_xy(float T, float& x, float& y) {
    T = std::clamp(T, 1667.0f, 25000.0f);
    const double t = T, t2 = t * t, t3 = t2 * t;
    const double xx = T <= 4000.0f ? -0.2661239e9 / t3 - 0.2343589e6 / t2 + 0.8776956e3 / t + 0.179910
                                   : -3.0258469e9 / t3 + 2.1070379e6 / t2 + 0.2226347e3 / t + 0.240390;
    double yy;
    if (T <= 2222.0f) yy = -1.1063814 * xx * xx * xx - 1.34811020 * xx * xx + 2.18555832 * xx - 0.20219683;
    else if (T <= 4000.0f) yy = -0.9549476 * xx * xx * xx - 1.37418593 * xx * xx + 2.09137015 * xx - 0.16748867;
    else yy = 3.0817580 * xx * xx * xx - 5.87338670 * xx * xx + 3.75112997 * xx - 0.37001483;
    x = static_cast<float>(xx);
    y = static_cast<float>(yy);
}
void daylight_xy(float T, float& x, float& y) {
    T = std::clamp(T, 4000.0f, 25000.0f);
    const double t = T, t2 = t * t, t3 = t2 * t;
    const double xx = T <= 7000.0f ? -4.6070e9 / t3 + 2.9678e6 / t2 + 0.09911e3 / t + 0.244063
                                   : -2.0064e9 / t3 + 1.9018e6 / t2 + 0.24748e3 / t + 0.237040;
    x = static_cast<float>(xx);
    y = static_cast<float>(-3.0 * xx * xx + 2.870 * xx - 0.275);
}
void illuminant_xy(float T, float& x, float& y) {
    float xp, yp, xd, yd;
    planck_xy(T, xp, yp);
    daylight_xy(T, xd, yd);
    const float t = std::clamp((T - 3500.0f) / 1000.0f, 0.0f, 1.0f);
    x = xp + (xd - xp) * t;
    y = yp + (yd - yp) * t;
}

}  // namespace

namespace oracle {

void white_balance_matrix(float temperature, float tint, float out[9]) {
    float xs, ys;
    illuminant_xy(temperature, xs, ys);
    const float xd = 0.31271f, yd = 0.32902f;
    ys += tint * 0.02f;
    const float src[3] = {xs / ys, 1.0f, (1.0f - xs - ys) / ys};
    const float dst[3] = {xd / yd, 1.0f, (1.0f - xd - yd) / yd};
    float lms_s[3], lms_d[3];
    for (int i = 0; i < 3; ++i) {
        lms_s[i] = kBradford[i * 3] * src[0] + kBradford[i * 3 + 1] * src[1] + kBradford[i * 3 + 2] * src[2];
        lms_d[i] = kBradford[i * 3] * dst[0] + kBradford[i * 3 + 1] * dst[1] + kBradford[i * 3 + 2] * dst[2];
    }
    const float D[9] = {lms_d[0] / lms_s[0], 0, 0, 0, lms_d[1] / lms_s[1], 0, 0, 0, lms_d[2] / lms_s[2]};
    float m[9];
    mul3(D, kBradford, m);
    mul3(kBradfordInv, m, m);
    mul3(m, kRgbToXyz, m);
    mul3(kXyzToRgb, m, out);
}

}  // namespace oracle

namespace {

bool imager_identity(const pe_display& s) {
    return s.bloom_intensity <= 0.0f && s.glare_intensity <= 0.0f && s.white_balance_k == 6500.0f && s.tint == 0.0f &&
           s.saturation == 1.0f && s.contrast == 1.0f && s.vignette <= 0.0f && s.chromatic_aberration <= 0.0f;
}

void apply_imager(Img& hdr, const pe_display& s, float ev) {  // imager.cpp:235-311
    if (imager_identity(s) || hdr.w <= 0 || hdr.h <= 0) return;
    const int w = hdr.w, h = hdr.h;
    const float diag = std::sqrt(static_cast<float>(w) * w + static_cast<float>(h) * h);
    const float cx = 0.5f * w, cy = 0.5f * h;
    const float exposure_scale = std::exp2(ev);
    if (s.chromatic_aberration > 0.0f) {
        const Img src = hdr;
        const float k = s.chromatic_aberration * 1e-3f * diag / (0.5f * diag);
        for (int y = 0; y < h; ++y)
            for (int x = 0; x < w; ++x) {
                const float px = x + 0.5f - cx, py = y + 0.5f - cy;
                const C3 r = sample_clamped(src, cx + px * (1.0f + k), cy + py * (1.0f + k));
                const C3 b = sample_clamped(src, cx + px * (1.0f - k), cy + py * (1.0f - k));
                hdr.set(x, y, C3(r.r, src.get(x, y).g, b.b));
            }
    }
    if (s.bloom_intensity > 0.0f || s.glare_intensity > 0.0f) {
        const Img bright = bright_pass(hdr, std::max(0.0f, s.bloom_threshold), exposure_scale);
        Img bloom, glare;
        if (s.bloom_intensity > 0.0f) bloom = bloom_image(bright, std::max(0.5f, s.bloom_radius * diag * 0.5f * 0.5f));
        if (s.glare_intensity > 0.0f)
            glare = glare_image(bright, std::max(2, static_cast<int>(s.glare_blades)), s.glare_angle_deg,
                                std::max(4.0f, s.glare_length * diag));
        for (int y = 0; y < h; ++y)
            for (int x = 0; x < w; ++x) {
                C3 c = hdr.get(x, y);
                if (s.bloom_intensity > 0.0f) c += sample_clamped(bloom, (x + 0.5f) * 0.5f, (y + 0.5f) * 0.5f) * s.bloom_intensity;
                if (s.glare_intensity > 0.0f)
                    c += sample_clamped(glare, (x + 0.5f) * 0.25f, (y + 0.5f) * 0.25f) * (s.glare_intensity * 4.0f);
                hdr.set(x, y, c);
            }
    }
    float wb[9];
    const bool do_wb = s.white_balance_k != 6500.0f || s.tint != 0.0f;
    if (do_wb) oracle::white_balance_matrix(s.white_balance_k, s.tint, wb);
    const float half_diag = 0.5f * diag;
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            C3 c = hdr.get(x, y);
            if (do_wb)
                c = C3(wb[0] * c.r + wb[1] * c.g + wb[2] * c.b, wb[3] * c.r + wb[4] * c.g + wb[5] * c.b,
                       wb[6] * c.r + wb[7] * c.g + wb[8] * c.b);
            if (s.contrast != 1.0f || s.saturation != 1.0f) {
                const float L = std::max(c.luminance(), 0.0f);
                float L2 = L;
                if (s.contrast != 1.0f && L > 0.0f) {
                    const float Le = L * exposure_scale;
                    L2 = 0.18f * std::pow(Le / 0.18f, s.contrast) / exposure_scale;
                    c = c * (L2 / L);
                }
                if (s.saturation != 1.0f) c = C3(L2) + (c - C3(L2)) * s.saturation;
            }
            if (s.vignette > 0.0f) {
                const float dx = (x + 0.5f - cx) / half_diag, dy = (y + 0.5f - cy) / half_diag;
                const float r2 = dx * dx + dy * dy;
                c = c * std::max(0.0f, 1.0f - s.vignette * r2 * r2 * 0.9f - s.vignette * r2 * 0.1f);
            }
            hdr.set(x, y, C3(std::max(0.0f, c.r), std::max(0.0f, c.g), std::max(0.0f, c.b)));
        }
}

// softBlurDenoise (engine/denoiser.cpp:17-43), what denoiseImage does without OIDN (:140-143): each pixel the
// 3 x 3 average (edges clamped) weighted by 1 / (1 + luminance^2), the colour scaled by 1 / the weight sum.
void soft_blur(Img& img) {
    const Img src = img;
    for (int y = 0; y < img.h; ++y)
        for (int x = 0; x < img.w; ++x) {
            C3 sum(0.0f);
            float wsum = 0.0f;
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    const C3 c = src.get(std::clamp(x + dx, 0, img.w - 1), std::clamp(y + dy, 0, img.h - 1));
                    const float lum = 0.2126f * c.r + 0.7152f * c.g + 0.0722f * c.b;
                    const float weight = 1.0f / (1.0f + lum * lum);
                    sum = sum + c * weight;
                    wsum += weight;
                }
            const float inv = 1.0f / std::max(wsum, 1e-6f);
            img.set(x, y, sum * inv);
        }
}

}  // namespace

// toneMap (tone_mapping.cpp:157-160) of n linear colours.
ORACLE_API int oracle_tonemap(size_t n, const float* rgb, int op, float ev, float* out) {
    for (size_t i = 0; i < n; ++i) {
        const C3 d = tonemap(C3(rgb[3 * i], rgb[3 * i + 1], rgb[3 * i + 2]), op, ev);
        out[3 * i] = d.r;
        out[3 * i + 1] = d.g;
        out[3 * i + 2] = d.b;
    }
    return 0;
}

// applyImager (imager.cpp:235-311) on a w x h linear image in place, with the settings and exposure of `ds`.
ORACLE_API int oracle_imager(int w, int h, float* rgb, const pe_display* ds) {
    Img img(w, h);
    std::copy(rgb, rgb + static_cast<size_t>(w) * h * 3, img.px.begin());
    apply_imager(img, *ds, ds->exposure_ev);
    std::copy(img.px.begin(), img.px.end(), rgb);
    return 0;
}

// The whole display path of a film's sums and counts: resolve, imager, tone map, 8-bit (optional outputs).
ORACLE_API int oracle_display(int w, int h, const float* sum, const uint32_t* count, const pe_display* ds,
                              float* display, uint8_t* rgb8) {
    Img img(w, h);
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {  // Film::resolvedPixel (film.cpp:87-94)
            const size_t k = static_cast<size_t>(y) * w + x;
            const uint32_t n = count[k];
            if (n == 0) continue;
            const float inv = 1.0f / static_cast<float>(n);
            img.set(x, y, C3(sum[3 * k] * inv, sum[3 * k + 1] * inv, sum[3 * k + 2] * inv));
        }
    if (ds->denoise) soft_blur(img);  // before the imager (main.cpp:277-289)
    apply_imager(img, *ds, ds->exposure_ev);
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            const size_t k = static_cast<size_t>(y) * w + x;
            const C3 d = tonemap(img.get(x, y), ds->tonemap, ds->exposure_ev);
            const float v[3] = {d.r, d.g, d.b};
            for (int c = 0; c < 3; ++c) {
                if (display) display[3 * k + c] = v[c];
                if (rgb8) rgb8[3 * k + c] = quantize_unorm8(v[c], ds->dither ? tpdf_dither(x, y, c) : 0.0f);
            }
        }
    return 0;
}

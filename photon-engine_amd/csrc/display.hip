// display.hip — the film's display image on the device: resolve, imager, tone map, 8-bit (pe_film_display).
//
// Follows Film::resolve (film.cpp:87-94, :139-147), applyImager (core/image/imager.cpp:235-311) with its
// chromatic aberration, bright pass, three-octave bloom of separable Gaussian blurs, glare lines, white balance
// (Bradford, imager.cpp:213-233), contrast / saturation and vignette, then toneMap (tone_mapping.cpp:146-160)
// and quantizeUnorm8 with tpdfDither (transfer.cpp:28-52) as saveImagePNG does. Every image is a float4 array
// (rgb | unused) in HBM; every stage is one pass over its image, so the path is bound by HBM bandwidth, a few
// reads and writes of 16 bytes per pixel. This file is compiled without contraction (a * b + c stays two
// roundings, as in the reference's IEEE build), and the host computes everything that does not depend on the
// pixel (blur and glare weights, the white-balance matrix, exp2 of the exposure) with the reference's own
// float operations.
#include "photon_hip.h"
#include "pe_internal.h"

#include <algorithm>
#include <cmath>
#include <vector>

namespace pe {
namespace {

constexpr int kDispBlock = 256;

struct Img4 {
    float4* p = nullptr;
    int w = 0, h = 0;
};

__device__ __forceinline__ float3 ld(const Img4& im, int x, int y) {
    const float4 v = im.p[static_cast<size_t>(y) * im.w + x];
    return make_float3(v.x, v.y, v.z);
}
__device__ __forceinline__ void st(const Img4& im, int x, int y, float3 c) {
    im.p[static_cast<size_t>(y) * im.w + x] = make_float4(c.x, c.y, c.z, 0.0f);
}
__device__ __forceinline__ float3 add3(float3 a, float3 b) { return make_float3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ float3 sub3(float3 a, float3 b) { return make_float3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ float3 mul3s(float3 a, float s) { return make_float3(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ float lum3(float3 c) { return 0.212671f * c.x + 0.715160f * c.y + 0.072169f * c.z; }
__device__ __forceinline__ float clampf_(float v, float lo, float hi) { return fminf(fmaxf(v, lo), hi); }

// sampleClamped (imager.cpp:25-35): bilinear with the sample at pixel centres, clamped to the image.
__device__ float3 sample_clamped(const Img4& im, float x, float y) {
    x = clampf_(x - 0.5f, 0.0f, static_cast<float>(im.w - 1));
    y = clampf_(y - 0.5f, 0.0f, static_cast<float>(im.h - 1));
    const int x0 = static_cast<int>(x), y0 = static_cast<int>(y);
    const int x1 = min(x0 + 1, im.w - 1), y1 = min(y0 + 1, im.h - 1);
    const float fx = x - static_cast<float>(x0), fy = y - static_cast<float>(y0);
    const float3 a = add3(mul3s(ld(im, x0, y0), 1.0f - fx), mul3s(ld(im, x1, y0), fx));
    const float3 b = add3(mul3s(ld(im, x0, y1), 1.0f - fx), mul3s(ld(im, x1, y1), fx));
    return add3(mul3s(a, 1.0f - fy), mul3s(b, fy));
}

#define PIXELS(im)                                                                                      \
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < (im).w * (im).h; i += gridDim.x * blockDim.x)

__global__ void __launch_bounds__(kDispBlock) k_resolve(const float* __restrict__ sum, const uint32_t* __restrict__ count,
                                                        Img4 out) {
    PIXELS(out) {
        const uint32_t n = count[i];
        float3 c = make_float3(0.0f, 0.0f, 0.0f);
        if (n != 0) {
            const float inv = 1.0f / static_cast<float>(n);
            c = make_float3(sum[3 * static_cast<size_t>(i)] * inv, sum[3 * static_cast<size_t>(i) + 1] * inv,
                            sum[3 * static_cast<size_t>(i) + 2] * inv);
        }
        st(out, i % out.w, i / out.w, c);
    }
}

// denoiseImage without OIDN: softBlurDenoise (engine/denoiser.cpp:17-43). Each pixel becomes the average of its
// 3 x 3 neighbourhood (coordinates clamped to the image), each neighbour weighted by 1 / (1 + luminance^2), which
// pulls single bright pixels down; sums in the reference's order, uncontracted, the colour scaled by 1 / the
// weight sum as Color3f::operator/ does.
__global__ void __launch_bounds__(kDispBlock) k_soft_blur(Img4 src, Img4 dst) {
#pragma clang fp contract(off)
    PIXELS(dst) {
        const int x = i % dst.w, y = i / dst.w;
        float sr = 0.0f, sg = 0.0f, sb = 0.0f, ws = 0.0f;
        for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx) {
                const float3 c = ld(src, min(max(x + dx, 0), src.w - 1), min(max(y + dy, 0), src.h - 1));
                const float lum = 0.2126f * c.x + 0.7152f * c.y + 0.0722f * c.z;
                const float wgt = 1.0f / (1.0f + lum * lum);
                sr += c.x * wgt;
                sg += c.y * wgt;
                sb += c.z * wgt;
                ws += wgt;
            }
        const float inv = 1.0f / fmaxf(ws, 1e-6f);
        st(dst, x, y, make_float3(sr * inv, sg * inv, sb * inv));
    }
}

__global__ void __launch_bounds__(kDispBlock) k_chroma(Img4 src, Img4 dst, float k) {  // imager.cpp:244-256
    const float cx = 0.5f * src.w, cy = 0.5f * src.h;
    PIXELS(dst) {
        const int x = i % dst.w, y = i / dst.w;
        const float px = x + 0.5f - cx, py = y + 0.5f - cy;
        const float3 r = sample_clamped(src, cx + px * (1.0f + k), cy + py * (1.0f + k));
        const float3 b = sample_clamped(src, cx + px * (1.0f - k), cy + py * (1.0f - k));
        st(dst, x, y, make_float3(r.x, ld(src, x, y).y, b.z));
    }
}

__global__ void __launch_bounds__(kDispBlock) k_down2(Img4 src, Img4 dst) {  // downsample2 (imager.cpp:37-49)
    PIXELS(dst) {
        const int x = i % dst.w, y = i / dst.w;
        const int sx = min(2 * x, src.w - 1), sy = min(2 * y, src.h - 1);
        const int sx1 = min(sx + 1, src.w - 1), sy1 = min(sy + 1, src.h - 1);
        st(dst, x, y, mul3s(add3(add3(add3(ld(src, sx, sy), ld(src, sx1, sy)), ld(src, sx, sy1)), ld(src, sx1, sy1)), 0.25f));
    }
}

__global__ void __launch_bounds__(kDispBlock) k_bright(Img4 im, float threshold, float exposure_scale) {  // :79-89
    PIXELS(im) {
        const int x = i % im.w, y = i / im.w;
        const float3 c = ld(im, x, y);
        const float L = lum3(c) * exposure_scale;
        const float f = L > threshold ? (L - threshold) / L : 0.0f;
        st(im, x, y, mul3s(c, f));
    }
}

// One direction of gaussianBlur (imager.cpp:51-77): taps k[0..2r] at offsets -r..r, clamped at the border.
template <bool kVertical>
__global__ void __launch_bounds__(kDispBlock) k_blur(Img4 src, Img4 dst, const float* __restrict__ k, int r) {
    PIXELS(dst) {
        const int x = i % dst.w, y = i / dst.w;
        float3 acc = make_float3(0.0f, 0.0f, 0.0f);
        for (int t = -r; t <= r; ++t) {
            const float3 c = kVertical ? ld(src, x, min(max(y + t, 0), src.h - 1)) : ld(src, min(max(x + t, 0), src.w - 1), y);
            acc = add3(acc, mul3s(c, k[t + r]));
        }
        st(dst, x, y, acc);
    }
}

// bloomImage's accumulation of one octave (imager.cpp:91-113).
__global__ void __launch_bounds__(kDispBlock) k_bloom_acc(Img4 acc, Img4 level, int scale, float weight) {
    PIXELS(acc) {
        const int x = i % acc.w, y = i / acc.w;
        const float3 c = scale == 1 ? ld(level, x, y)
                                    : sample_clamped(level, (x + 0.5f) / static_cast<float>(scale),
                                                     (y + 0.5f) / static_cast<float>(scale));
        st(acc, x, y, add3(ld(acc, x, y), mul3s(c, weight)));
    }
}

// One line of glareImage (imager.cpp:115-147): weights w[0..steps] along +-(dx, dy), divided by norm.
__global__ void __launch_bounds__(kDispBlock) k_glare_line(Img4 src, Img4 out, const float* __restrict__ w, int steps,
                                                           float dx, float dy, float norm) {
    PIXELS(out) {
        const int x = i % out.w, y = i / out.w;
        float3 acc = mul3s(ld(src, x, y), w[0]);
        for (int t = 1; t <= steps; ++t) {
            acc = add3(acc, mul3s(sample_clamped(src, x + 0.5f + dx * t, y + 0.5f + dy * t), w[t]));
            acc = add3(acc, mul3s(sample_clamped(src, x + 0.5f - dx * t, y + 0.5f - dy * t), w[t]));
        }
        const float inv = 1.0f / norm;
        st(out, x, y, add3(ld(out, x, y), mul3s(acc, inv)));
    }
}

__global__ void __launch_bounds__(kDispBlock) k_composite(Img4 hdr, Img4 bloom, float bloom_i, Img4 glare,
                                                          float glare_i) {  // imager.cpp:265-275
    PIXELS(hdr) {
        const int x = i % hdr.w, y = i / hdr.w;
        float3 c = ld(hdr, x, y);
        if (bloom_i > 0.0f) c = add3(c, mul3s(sample_clamped(bloom, (x + 0.5f) * 0.5f, (y + 0.5f) * 0.5f), bloom_i));
        if (glare_i > 0.0f)
            c = add3(c, mul3s(sample_clamped(glare, (x + 0.5f) * 0.25f, (y + 0.5f) * 0.25f), glare_i * 4.0f));
        st(hdr, x, y, c);
    }
}

struct PixelOps {
    float wb[9];
    int do_wb;
    float contrast, saturation, vignette, exposure_scale, half_diag;
};

__global__ void __launch_bounds__(kDispBlock) k_imager_pixel(Img4 hdr, PixelOps o) {  // imager.cpp:278-309
    const float cx = 0.5f * hdr.w, cy = 0.5f * hdr.h;
    PIXELS(hdr) {
        const int x = i % hdr.w, y = i / hdr.w;
        float3 c = ld(hdr, x, y);
        if (o.do_wb)
            c = make_float3(o.wb[0] * c.x + o.wb[1] * c.y + o.wb[2] * c.z, o.wb[3] * c.x + o.wb[4] * c.y + o.wb[5] * c.z,
                            o.wb[6] * c.x + o.wb[7] * c.y + o.wb[8] * c.z);
        if (o.contrast != 1.0f || o.saturation != 1.0f) {
            const float L = fmaxf(lum3(c), 0.0f);
            float L2 = L;
            if (o.contrast != 1.0f && L > 0.0f) {
                const float Le = L * o.exposure_scale;
                L2 = 0.18f * powf(Le / 0.18f, o.contrast) / o.exposure_scale;
                c = mul3s(c, L2 / L);
            }
            if (o.saturation != 1.0f) {
                const float3 l2 = make_float3(L2, L2, L2);
                c = add3(l2, mul3s(sub3(c, l2), o.saturation));
            }
        }
        if (o.vignette > 0.0f) {
            const float dx = (x + 0.5f - cx) / o.half_diag, dy = (y + 0.5f - cy) / o.half_diag;
            const float r2 = dx * dx + dy * dy;
            c = mul3s(c, fmaxf(0.0f, 1.0f - o.vignette * r2 * r2 * 0.9f - o.vignette * r2 * 0.1f));
        }
        st(hdr, x, y, make_float3(fmaxf(0.0f, c.x), fmaxf(0.0f, c.y), fmaxf(0.0f, c.z)));
    }
}

// ---- tone mapping (tone_mapping.cpp) and output (transfer.cpp)
__device__ float srgb_decode(float e) { return e <= 0.04045f ? e / 12.92f : powf((e + 0.055f) / 1.055f, 2.4f); }
__device__ float srgb_encode(float l) {
    l = clampf_(l, 0.0f, 1.0f);
    return l <= 0.0031308f ? 12.92f * l : 1.055f * powf(l, 1.0f / 2.4f) - 0.055f;
}
__device__ float aces(float x) {
    return clampf_((x * (2.51f * x + 0.03f)) / (x * (2.43f * x + 0.59f) + 0.14f), 0.0f, 1.0f);
}
__device__ float hable(float x) {
    const float A = 0.15f, B = 0.50f, C = 0.10f, D = 0.20f, E = 0.02f, F = 0.30f;
    return ((x * (A * x + C * B) + D * E) / (x * (A * x + B) + D * F)) - E / F;
}
__device__ float agx_curve(float x) {
    const float min_ev = -12.47393f, max_ev = 4.026069f;
    x = clampf_(x, min_ev, max_ev);
    x = (x - min_ev) / (max_ev - min_ev);
    const float x2 = x * x, x4 = x2 * x2;
    return 15.5f * x4 * x2 - 40.14f * x4 * x + 31.96f * x4 - 6.868f * x2 * x + 0.4298f * x2 + 0.1191f * x - 0.00232f;
}
__device__ float3 agx(float3 v) {
    // The AgX inset / outset matrices (GLSL column-major), as toneMapAgX states them.
    const float in_m[9] = {0.842479062253094f, 0.0423282422610123f, 0.0423756549057051f,
                           0.0784335999999992f, 0.878468636469772f, 0.0784336f,
                           0.0792237451477643f, 0.0791661274605434f, 0.879142973793104f};
    const float out_m[9] = {1.19687900512017f, -0.0528968517574562f, -0.0529716355144438f,
                            -0.0980208811401368f, 1.15190312990417f, -0.0980434501171241f,
                            -0.0990297440797205f, -0.0989611768448433f, 1.15107367264116f};
    float3 a = make_float3(in_m[0] * v.x + in_m[3] * v.y + in_m[6] * v.z, in_m[1] * v.x + in_m[4] * v.y + in_m[7] * v.z,
                           in_m[2] * v.x + in_m[5] * v.y + in_m[8] * v.z);
    a = make_float3(agx_curve(log2f(fmaxf(a.x, 1e-10f))), agx_curve(log2f(fmaxf(a.y, 1e-10f))),
                    agx_curve(log2f(fmaxf(a.z, 1e-10f))));
    const float3 b = make_float3(out_m[0] * a.x + out_m[3] * a.y + out_m[6] * a.z,
                                 out_m[1] * a.x + out_m[4] * a.y + out_m[7] * a.z,
                                 out_m[2] * a.x + out_m[5] * a.y + out_m[8] * a.z);
    return make_float3(srgb_decode(clampf_(b.x, 0.0f, 1.0f)), srgb_decode(clampf_(b.y, 0.0f, 1.0f)),
                       srgb_decode(clampf_(b.z, 0.0f, 1.0f)));
}
__device__ float3 pbr_neutral(float3 c) {
    const float start = 0.8f - 0.04f, desat = 0.15f;
    const float x = fminf(c.x, fminf(c.y, c.z));
    const float offset = x < 0.08f ? x - 6.25f * x * x : 0.04f;
    c = sub3(c, make_float3(offset, offset, offset));
    const float peak = fmaxf(c.x, fmaxf(c.y, c.z));
    if (peak < start) return c;
    const float d = 1.0f - start;
    const float new_peak = 1.0f - d * d / (peak + d - start);
    c = mul3s(c, new_peak / peak);
    const float g = 1.0f - 1.0f / (desat * (peak - new_peak) + 1.0f);
    return add3(mul3s(c, 1.0f - g), make_float3(new_peak * g, new_peak * g, new_peak * g));
}
__device__ __forceinline__ float sane(float v) {
    return (__float_as_uint(v) & 0x7f800000u) != 0x7f800000u && v > 0.0f ? v : 0.0f;
}
__device__ float3 tonemap_linear(float3 hdr, int op, float scale, float filmic_white) {
    float3 e = mul3s(hdr, scale);
    e = make_float3(sane(e.x), sane(e.y), sane(e.z));
    float3 l;
    switch (op) {
        case PE_TONEMAP_REINHARD: l = make_float3(e.x / (1.0f + e.x), e.y / (1.0f + e.y), e.z / (1.0f + e.z)); break;
        case PE_TONEMAP_REINHARD_EXTENDED: {
            const float mw2 = 4.0f * 4.0f;
            l = make_float3((e.x * (1.0f + e.x / mw2)) / (1.0f + e.x), (e.y * (1.0f + e.y / mw2)) / (1.0f + e.y),
                            (e.z * (1.0f + e.z / mw2)) / (1.0f + e.z));
            break;
        }
        case PE_TONEMAP_ACES: l = make_float3(aces(e.x), aces(e.y), aces(e.z)); break;
        case PE_TONEMAP_FILMIC: l = make_float3(hable(e.x) * filmic_white, hable(e.y) * filmic_white, hable(e.z) * filmic_white); break;
        case PE_TONEMAP_AGX: l = agx(e); break;
        case PE_TONEMAP_PBR_NEUTRAL: l = pbr_neutral(e); break;
        default: l = e; break;
    }
    return make_float3(fmaxf(0.0f, fminf(l.x, 1.0f)), fmaxf(0.0f, fminf(l.y, 1.0f)), fmaxf(0.0f, fminf(l.z, 1.0f)));
}
__device__ __forceinline__ float tpdf_dither(int x, int y, int channel) {
    uint32_t h = static_cast<uint32_t>(x) * 0x8da6b343u ^ static_cast<uint32_t>(y) * 0xd8163841u ^
                 static_cast<uint32_t>(channel) * 0xcb1ab31fu;
    h ^= h >> 16;
    h *= 0x7feb352du;
    h ^= h >> 15;
    h *= 0x846ca68bu;
    h ^= h >> 16;
    return static_cast<float>(h & 0xffffu) * (1.0f / 65536.0f) - static_cast<float>(h >> 16) * (1.0f / 65536.0f);
}

__global__ void __launch_bounds__(kDispBlock) k_tonemap(Img4 hdr, int op, float scale, float filmic_white, int dither,
                                                        float* __restrict__ display, uint8_t* __restrict__ rgb8) {
    PIXELS(hdr) {
        const int x = i % hdr.w, y = i / hdr.w;
        const float3 l = tonemap_linear(ld(hdr, x, y), op, scale, filmic_white);
        const float d[3] = {srgb_encode(l.x), srgb_encode(l.y), srgb_encode(l.z)};
        for (int c = 0; c < 3; ++c) {
            if (display) display[3 * static_cast<size_t>(i) + c] = d[c];
            if (rgb8) {
                const float q = floorf(d[c] * 255.0f + (dither ? tpdf_dither(x, y, c) : 0.0f) + 0.5f);
                rgb8[3 * static_cast<size_t>(i) + c] = static_cast<uint8_t>(clampf_(q, 0.0f, 255.0f));
            }
        }
    }
}

#undef PIXELS

// ---- host side: the pixel-independent quantities, with the reference's float operations
std::vector<float> gauss_weights(float sigma, int& r) {  // gaussianBlur (imager.cpp:53-57)
    r = std::max(1, static_cast<int>(std::ceil(3.0f * sigma)));
    std::vector<float> k(static_cast<size_t>(2 * r + 1));
    float sum = 0.0f;
    for (int i = -r; i <= r; ++i) sum += k[static_cast<size_t>(i + r)] = std::exp(-0.5f * (i * i) / (sigma * sigma));
    for (float& v : k) v /= sum;
    return k;
}

const float kRgbToXyz[9] = {0.4124564f, 0.3575761f, 0.1804375f, 0.2126729f, 0.7151522f,
                            0.0721750f, 0.0193339f, 0.1191920f, 0.9503041f};
const float kXyzToRgb[9] = {3.2404542f, -1.5371385f, -0.4985314f, -0.9692660f, 1.8760108f,
                            0.0415560f, 0.0556434f, -0.2040259f, 1.0572252f};
const float kBradford[9] = {0.8951f, 0.2664f, -0.1614f, -0.7502f, 1.7135f, 0.0367f, 0.0389f, -0.0685f, 1.0296f};
const float kBradfordInv[9] = {0.9869929f, -0.1470543f, 0.1599627f, 0.4323053f, 0.5183603f,
                               0.0492912f, -0.0085287f, 0.0400428f, 0.9684867f};

void mat3_mul(const float a[9], const float b[9], float out[9]) {
    float r[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) r[i * 3 + j] = a[i * 3] * b[j] + a[i * 3 + 1] * b[3 + j] + a[i * 3 + 2] * b[6 + j];
    std::copy(r, r + 9, out);
}

// CIE xy of a Planckian / daylight illuminant of temperature T, blended between 3500 K and 4500 K
// (imager.cpp:169-207: Kim et al. cubic fits of the Planckian locus, the CIE daylight locus).
void illuminant_xy(float T, float& x, float& y) {
    float xp, yp, xd, yd;
    {
        const float Tc = std::clamp(T, 1667.0f, 25000.0f);
        const double t = Tc, t2 = t * t, t3 = t2 * t;
        const double xx = Tc <= 4000.0f ? -0.2661239e9 / t3 - 0.2343589e6 / t2 + 0.8776956e3 / t + 0.179910
                                        : -3.0258469e9 / t3 + 2.1070379e6 / t2 + 0.2226347e3 / t + 0.240390;
        double yy;
        if (Tc <= 2222.0f) yy = -1.1063814 * xx * xx * xx - 1.34811020 * xx * xx + 2.18555832 * xx - 0.20219683;
        else if (Tc <= 4000.0f) yy = -0.9549476 * xx * xx * xx - 1.37418593 * xx * xx + 2.09137015 * xx - 0.16748867;
        else yy = 3.0817580 * xx * xx * xx - 5.87338670 * xx * xx + 3.75112997 * xx - 0.37001483;
        xp = static_cast<float>(xx);
        yp = static_cast<float>(yy);
    }
    {
        const float Tc = std::clamp(T, 4000.0f, 25000.0f);
        const double t = Tc, t2 = t * t, t3 = t2 * t;
        const double xx = Tc <= 7000.0f ? -4.6070e9 / t3 + 2.9678e6 / t2 + 0.09911e3 / t + 0.244063
                                        : -2.0064e9 / t3 + 1.9018e6 / t2 + 0.24748e3 / t + 0.237040;
        xd = static_cast<float>(xx);
        yd = static_cast<float>(-3.0 * xx * xx + 2.870 * xx - 0.275);
    }
    const float t = std::clamp((T - 3500.0f) / 1000.0f, 0.0f, 1.0f);
    x = xp + (xd - xp) * t;
    y = yp + (yd - yp) * t;
}

void white_balance(float temperature, float tint, float out[9]) {  // whiteBalanceMatrix (imager.cpp:213-233)
    float xs, ys;
    illuminant_xy(temperature, xs, ys);
    const float xd = 0.31271f, yd = 0.32902f;
    ys += tint * 0.02f;
    const float src[3] = {xs / ys, 1.0f, (1.0f - xs - ys) / ys};
    const float dst[3] = {xd / yd, 1.0f, (1.0f - xd - yd) / yd};
    float ls[3], ldd[3];
    for (int i = 0; i < 3; ++i) {
        ls[i] = kBradford[i * 3] * src[0] + kBradford[i * 3 + 1] * src[1] + kBradford[i * 3 + 2] * src[2];
        ldd[i] = kBradford[i * 3] * dst[0] + kBradford[i * 3 + 1] * dst[1] + kBradford[i * 3 + 2] * dst[2];
    }
    const float D[9] = {ldd[0] / ls[0], 0, 0, 0, ldd[1] / ls[1], 0, 0, 0, ldd[2] / ls[2]};
    float m[9];
    mat3_mul(D, kBradford, m);
    mat3_mul(kBradfordInv, m, m);
    mat3_mul(m, kRgbToXyz, m);
    mat3_mul(kXyzToRgb, m, out);
}

int grid_for(int n) { return std::max(1, std::min(65535, (n + kDispBlock - 1) / kDispBlock)); }

struct Pool {  // device images of one call and the host data of its uploads, released at return (after the sync)
    std::vector<DevBuf> bufs;
    std::vector<std::vector<float>> host;
    Img4 make(int w, int h) {
        bufs.emplace_back();
        if (!bufs.back().alloc(static_cast<size_t>(w) * h * sizeof(float4))) return Img4{};
        return Img4{bufs.back().as<float4>(), w, h};
    }
    const float* upload(const std::vector<float>& v, hipStream_t s) {
        host.push_back(v);
        bufs.emplace_back();
        if (!bufs.back().upload(host.back().data(), v.size() * sizeof(float), s)) return nullptr;
        return bufs.back().as<float>();
    }
};

float hable_host(float x) {
    const float A = 0.15f, B = 0.50f, C = 0.10f, D = 0.20f, E = 0.02f, F = 0.30f;
    return ((x * (A * x + C * B) + D * E) / (x * (A * x + B) + D * F)) - E / F;
}

}  // namespace
}  // namespace pe

using namespace pe;

extern "C" int pe_display_defaults(pe_display* out) {
    if (!out) {
        set_error("pe_display_defaults: NULL argument");
        return PE_ERR_INVALID;
    }
    pe_display d{};  // ImagerSettings{} (imager.h:12-36) and saveImagePNG's defaults
    d.tonemap = PE_TONEMAP_ACES;
    d.exposure_ev = 0.0f;
    d.dither = 1;
    d.bloom_intensity = 0.0f;
    d.bloom_radius = 0.04f;
    d.bloom_threshold = 1.0f;
    d.glare_intensity = 0.0f;
    d.glare_blades = 6;
    d.glare_angle_deg = 15.0f;
    d.glare_length = 0.1f;
    d.white_balance_k = 6500.0f;
    d.tint = 0.0f;
    d.saturation = 1.0f;
    d.contrast = 1.0f;
    d.vignette = 0.0f;
    d.chromatic_aberration = 0.0f;
    *out = d;
    return PE_OK;
}

extern "C" int pe_film_display(pe_film* film, const pe_display* ds, float* display_rgb, uint8_t* rgb8) {
    if (!film || !ds) {
        set_error("pe_film_display: NULL argument");
        return PE_ERR_INVALID;
    }
    if (ds->tonemap < PE_TONEMAP_REINHARD || ds->tonemap > PE_TONEMAP_LINEAR) {
        set_error("pe_film_display: unknown tone-map operator");
        return PE_ERR_INVALID;
    }
    try {
        pe_device* dev = film->dev;
        PE_HIP_CHECK(hipSetDevice(dev->ordinal));
        hipStream_t s = dev->stream;
        const int w = film->width, h = film->height, n = w * h;
        Pool pool;
        Img4 hdr = pool.make(w, h);
        if (!hdr.p) return PE_ERR_OOM;
        hipLaunchKernelGGL(k_resolve, dim3(grid_for(n)), dim3(kDispBlock), 0, s, film->sum.as<float>(),
                           film->count.as<uint32_t>(), hdr);
        if (ds->denoise) {
            Img4 resolved = hdr;
            hdr = pool.make(w, h);
            if (!hdr.p) return PE_ERR_OOM;
            hipLaunchKernelGGL(k_soft_blur, dim3(grid_for(n)), dim3(kDispBlock), 0, s, resolved, hdr);
        }
        const pe_display& c = *ds;
        const bool identity = c.bloom_intensity <= 0.0f && c.glare_intensity <= 0.0f && c.white_balance_k == 6500.0f &&
                              c.tint == 0.0f && c.saturation == 1.0f && c.contrast == 1.0f && c.vignette <= 0.0f &&
                              c.chromatic_aberration <= 0.0f;  // ImagerSettings::isIdentity
        const float exposure_scale = std::exp2(c.exposure_ev);
        if (!identity) {
            const float diag = std::sqrt(static_cast<float>(w) * w + static_cast<float>(h) * h);
            if (c.chromatic_aberration > 0.0f) {
                Img4 src = pool.make(w, h);
                if (!src.p) return PE_ERR_OOM;
                PE_HIP_CHECK(hipMemcpyAsync(src.p, hdr.p, static_cast<size_t>(n) * sizeof(float4), hipMemcpyDeviceToDevice, s));
                const float k = c.chromatic_aberration * 1e-3f * diag / (0.5f * diag);
                hipLaunchKernelGGL(k_chroma, dim3(grid_for(n)), dim3(kDispBlock), 0, s, src, hdr, k);
            }
            if (c.bloom_intensity > 0.0f || c.glare_intensity > 0.0f) {
                Img4 bright = pool.make(std::max(1, w / 2), std::max(1, h / 2));
                if (!bright.p) return PE_ERR_OOM;
                const int nb = bright.w * bright.h;
                hipLaunchKernelGGL(k_down2, dim3(grid_for(nb)), dim3(kDispBlock), 0, s, hdr, bright);
                hipLaunchKernelGGL(k_bright, dim3(grid_for(nb)), dim3(kDispBlock), 0, s, bright,
                                   std::max(0.0f, c.bloom_threshold), exposure_scale);
                Img4 bloom{}, glare{};
                if (c.bloom_intensity > 0.0f) {  // bloomImage (imager.cpp:91-113)
                    bloom = pool.make(bright.w, bright.h);
                    if (!bloom.p) return PE_ERR_OOM;
                    PE_HIP_CHECK(hipMemsetAsync(bloom.p, 0, static_cast<size_t>(nb) * sizeof(float4), s));
                    const float sigma_half = std::max(0.5f, c.bloom_radius * diag * 0.5f * 0.5f);
                    const float weights[3] = {0.5f, 0.3f, 0.2f};
                    for (int o = 0; o < 3; ++o) {
                        float sigma = sigma_half * static_cast<float>(1 << o);
                        Img4 level = pool.make(bright.w, bright.h);
                        if (!level.p) return PE_ERR_OOM;
                        PE_HIP_CHECK(hipMemcpyAsync(level.p, bright.p, static_cast<size_t>(nb) * sizeof(float4),
                                                    hipMemcpyDeviceToDevice, s));
                        int scale = 1;
                        while (sigma > 6.0f && level.w > 8 && level.h > 8) {
                            Img4 next = pool.make(std::max(1, level.w / 2), std::max(1, level.h / 2));
                            if (!next.p) return PE_ERR_OOM;
                            hipLaunchKernelGGL(k_down2, dim3(grid_for(next.w * next.h)), dim3(kDispBlock), 0, s, level, next);
                            level = next;
                            sigma *= 0.5f;
                            scale *= 2;
                        }
                        if (!(sigma < 0.3f)) {
                            int r = 0;
                            const std::vector<float> kw = gauss_weights(sigma, r);
                            const float* kd = pool.upload(kw, s);
                            Img4 tmp = pool.make(level.w, level.h);
                            if (!kd || !tmp.p) return PE_ERR_OOM;
                            const int nl = level.w * level.h;
                            hipLaunchKernelGGL(k_blur<false>, dim3(grid_for(nl)), dim3(kDispBlock), 0, s, level, tmp, kd, r);
                            hipLaunchKernelGGL(k_blur<true>, dim3(grid_for(nl)), dim3(kDispBlock), 0, s, tmp, level, kd, r);
                        }
                        hipLaunchKernelGGL(k_bloom_acc, dim3(grid_for(nb)), dim3(kDispBlock), 0, s, bloom, level, scale,
                                           weights[o]);
                    }
                }
                if (c.glare_intensity > 0.0f) {  // glareImage (imager.cpp:115-147)
                    Img4 src = pool.make(std::max(1, bright.w / 2), std::max(1, bright.h / 2));
                    if (!src.p) return PE_ERR_OOM;
                    const int nq = src.w * src.h;
                    hipLaunchKernelGGL(k_down2, dim3(grid_for(nq)), dim3(kDispBlock), 0, s, bright, src);
                    glare = pool.make(src.w, src.h);
                    if (!glare.p) return PE_ERR_OOM;
                    PE_HIP_CHECK(hipMemsetAsync(glare.p, 0, static_cast<size_t>(nq) * sizeof(float4), s));
                    const int blades = std::max(2, static_cast<int>(c.glare_blades));
                    const float L = std::max(2.0f, std::max(4.0f, c.glare_length * diag) * 0.25f);
                    const int lines = std::max(1, (blades + 1) / 2);
                    const int steps = static_cast<int>(std::ceil(L));
                    std::vector<float> wts(static_cast<size_t>(steps + 1));
                    float wsum = 0.0f;
                    for (int t = 0; t <= steps; ++t) {
                        wts[static_cast<size_t>(t)] = std::exp(-3.0f * static_cast<float>(t) / L);
                        wsum += (t == 0 ? 1.0f : 2.0f) * wts[static_cast<size_t>(t)];
                    }
                    const float* wd = pool.upload(wts, s);
                    if (!wd) return PE_ERR_OOM;
                    const float deg = 3.14159265358979323846f / 180.0f, pi = 3.14159265358979323846f;
                    for (int l = 0; l < lines; ++l) {
                        const float a = c.glare_angle_deg * deg + static_cast<float>(l) * pi / static_cast<float>(lines);
                        hipLaunchKernelGGL(k_glare_line, dim3(grid_for(nq)), dim3(kDispBlock), 0, s, src, glare, wd, steps,
                                           std::cos(a), std::sin(a), wsum * static_cast<float>(lines));
                    }
                }
                hipLaunchKernelGGL(k_composite, dim3(grid_for(n)), dim3(kDispBlock), 0, s, hdr, bloom,
                                   c.bloom_intensity > 0.0f ? c.bloom_intensity : 0.0f, glare,
                                   c.glare_intensity > 0.0f ? c.glare_intensity : 0.0f);
            }
            PixelOps o{};
            o.do_wb = c.white_balance_k != 6500.0f || c.tint != 0.0f;
            if (o.do_wb) white_balance(c.white_balance_k, c.tint, o.wb);
            o.contrast = c.contrast;
            o.saturation = c.saturation;
            o.vignette = c.vignette;
            o.exposure_scale = exposure_scale;
            o.half_diag = 0.5f * diag;
            hipLaunchKernelGGL(k_imager_pixel, dim3(grid_for(n)), dim3(kDispBlock), 0, s, hdr, o);
        }
        DevBuf d_disp, d_rgb8;
        if (display_rgb && !d_disp.alloc(static_cast<size_t>(n) * 12)) return PE_ERR_OOM;
        if (rgb8 && !d_rgb8.alloc(static_cast<size_t>(n) * 3)) return PE_ERR_OOM;
        hipLaunchKernelGGL(k_tonemap, dim3(grid_for(n)), dim3(kDispBlock), 0, s, hdr, c.tonemap, exposure_scale,
                           1.0f / hable_host(11.2f), c.dither, d_disp.as<float>(), d_rgb8.as<uint8_t>());
        PE_HIP_CHECK(hipGetLastError());
        if (display_rgb)
            PE_HIP_CHECK(hipMemcpyAsync(display_rgb, d_disp.get(), static_cast<size_t>(n) * 12, hipMemcpyDeviceToHost, s));
        if (rgb8) PE_HIP_CHECK(hipMemcpyAsync(rgb8, d_rgb8.get(), static_cast<size_t>(n) * 3, hipMemcpyDeviceToHost, s));
        PE_HIP_CHECK(hipStreamSynchronize(s));
        return PE_OK;
    } catch (const std::exception& e) {
        set_error(std::string("pe_film_display: ") + e.what());
        return PE_ERR_INVALID;
    }
}

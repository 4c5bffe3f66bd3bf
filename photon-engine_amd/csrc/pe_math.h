// pe_math.h — small vector/colour algebra shared by host set-up code and gfx950 kernels.
//
// Operation order follows the reference's Vec3f (core/math/vec.h:114-160): dot = x*x' + y*y' + z*z'
// evaluated left to right, cross = (y z' - z y', z x' - x z', x y' - y x'), v / s = v * (1/s) and
// normalized() returns 0 below length 1e-20. Where bit-level agreement matters (ray/triangle tests, Sobol)
// the callers disable contraction locally with `#pragma clang fp contract(off)`.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define PE_HD __host__ __device__ __forceinline__

namespace pe {

constexpr float kPi = 3.14159265358979323846f;
constexpr float kTwoPi = 2.0f * kPi;
constexpr float kInvPi = 1.0f / kPi;
constexpr float kInvTwoPi = 1.0f / kTwoPi;
constexpr float kPiOver2 = kPi / 2.0f;
constexpr float kPiOver4 = kPi / 4.0f;
constexpr float kFloatMax = 3.40282346638528859812e+38f;  // INFINITY_F of the reference (constants.h:52)

struct V3 {
    float x, y, z;
};

PE_HD V3 mk(float x, float y, float z) { return V3{x, y, z}; }
PE_HD V3 splat(float s) { return V3{s, s, s}; }
PE_HD V3 operator+(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
PE_HD V3 operator-(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
PE_HD V3 operator-(V3 a) { return V3{-a.x, -a.y, -a.z}; }
PE_HD V3 operator*(V3 a, V3 b) { return V3{a.x * b.x, a.y * b.y, a.z * b.z}; }
PE_HD V3 operator*(V3 a, float s) { return V3{a.x * s, a.y * s, a.z * s}; }
PE_HD V3 operator*(float s, V3 a) { return V3{s * a.x, s * a.y, s * a.z}; }
PE_HD V3 operator/(V3 a, float s) {
    const float inv = 1.0f / s;
    return V3{a.x * inv, a.y * inv, a.z * inv};
}
PE_HD V3 operator/(V3 a, V3 b) { return V3{a.x / b.x, a.y / b.y, a.z / b.z}; }
PE_HD V3& operator+=(V3& a, V3 b) {
    a = a + b;
    return a;
}
PE_HD V3& operator*=(V3& a, V3 b) {
    a = a * b;
    return a;
}
PE_HD V3& operator*=(V3& a, float s) {
    a = a * s;
    return a;
}
PE_HD float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
PE_HD V3 cross(V3 a, V3 b) { return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
// Uncontracted forms: each product is rounded before the sum, as in an IEEE build without FMA. The pragma is
// scoped to these bodies, so they stay exact wherever they are inlined.
PE_HD float dot_exact(V3 a, V3 b) {
#pragma clang fp contract(off)
    return a.x * b.x + a.y * b.y + a.z * b.z;
}
PE_HD V3 cross_exact(V3 a, V3 b) {
#pragma clang fp contract(off)
    return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
PE_HD float length_sq(V3 a) { return dot(a, a); }
PE_HD float length(V3 a) { return sqrtf(length_sq(a)); }
PE_HD V3 normalized(V3 a) {
    const float len = length(a);
    if (len <= 1e-20f) return splat(0.0f);
    return a / len;
}
PE_HD V3 vabs(V3 a) { return V3{fabsf(a.x), fabsf(a.y), fabsf(a.z)}; }
PE_HD V3 vmin(V3 a, V3 b) { return V3{fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z)}; }
PE_HD V3 vmax(V3 a, V3 b) { return V3{fmaxf(a.x, b.x), fmaxf(a.y, b.y), fmaxf(a.z, b.z)}; }
PE_HD float max_comp(V3 a) { return fmaxf(a.x, fmaxf(a.y, a.z)); }
PE_HD float comp(V3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
PE_HD bool is_black(V3 c) { return c.x == 0.0f && c.y == 0.0f && c.z == 0.0f; }
PE_HD float luminance(V3 c) { return 0.212671f * c.x + 0.715160f * c.y + 0.072169f * c.z; }
PE_HD V3 lerp3(V3 a, V3 b, float t) { return a * (1.0f - t) + b * t; }
PE_HD float clampf(float v, float lo, float hi) { return v < lo ? lo : (v > hi ? hi : v); }
PE_HD V3 clamp3(V3 v, float lo, float hi) { return V3{clampf(v.x, lo, hi), clampf(v.y, lo, hi), clampf(v.z, lo, hi)}; }
PE_HD float pow5(float x) {
    const float x2 = x * x;
    return x2 * x2 * x;
}

PE_HD uint32_t f2u(float x) { return __builtin_bit_cast(uint32_t, x); }
PE_HD float u2f(uint32_t x) { return __builtin_bit_cast(float, x); }

// Exponent-bit finiteness test that fast-math cannot fold away (core/math/float_bits.h:20-26).
PE_HD bool finite_f(float x) { return (f2u(x) & 0x7f800000u) != 0x7f800000u; }
PE_HD bool finite3(V3 c) { return finite_f(c.x) && finite_f(c.y) && finite_f(c.z); }

// Neighbouring representable floats (float_bits.h:30-49).
PE_HD float next_float_up(float v) {
    if (!finite_f(v) && v > 0.0f) return v;
    if (v == -0.0f) v = 0.0f;
    uint32_t ui = f2u(v);
    if (v >= 0.0f) ++ui;
    else --ui;
    return u2f(ui);
}
PE_HD float next_float_down(float v) {
    if (!finite_f(v) && v < 0.0f) return v;
    if (v == 0.0f) v = -0.0f;
    uint32_t ui = f2u(v);
    if (v > 0.0f) --ui;
    else ++ui;
    return u2f(ui);
}
// gamma(n) = n eps / (1 - n eps), eps = 2^-24 (float_bits.h:51-55).
PE_HD float float_gamma(int n) {
    const float eps = 5.9604645e-8f;
    return (static_cast<float>(n) * eps) / (1.0f - static_cast<float>(n) * eps);
}

// Orthonormal basis from a unit normal, Duff et al. 2017 (frame.h:27-38).
struct Frame {
    V3 s, t, n;
};
PE_HD Frame frame_from_normal(V3 n) {
    const float sign = copysignf(1.0f, n.z);
    const float a = -1.0f / (sign + n.z);
    const float b = n.x * n.y * a;
    Frame f;
    f.n = n;
    f.s = V3{1.0f + sign * n.x * n.x * a, sign * b, -sign * n.x};
    f.t = V3{b, sign + n.y * n.y * a, -n.y};
    return f;
}
PE_HD V3 to_local(const Frame& f, V3 v) { return V3{dot(v, f.s), dot(v, f.t), dot(v, f.n)}; }
PE_HD V3 to_world(const Frame& f, V3 v) { return f.s * v.x + f.t * v.y + f.n * v.z; }

}  // namespace pe

// oracle.h — CPU restatement of PhotonEngine's integrator hot path (test infrastructure only).
//
// This oracle is the checker of libphoton_hip: only tests/, __graft_entry__.smoke() and bench.py's CPU
// baseline leg may call it. It restates the reference's algorithms function by function, each citing the
// reference file:line it follows, in plain scalar C++ (one path at a time, binary BVH, AoS data), and it is
// itself checked against the reference's own sources compiled in oracle/_ref and against closed forms.
#pragma once

#include "photon_hip.h"

#include <cmath>
#include <cstdint>
#include <vector>

namespace oracle {

struct V3 {
    float x = 0, y = 0, z = 0;
    V3() = default;
    V3(float a, float b, float c) : x(a), y(b), z(c) {}
    explicit V3(float s) : x(s), y(s), z(s) {}
    float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
    float& operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
};
inline V3 operator+(V3 a, V3 b) { return V3(a.x + b.x, a.y + b.y, a.z + b.z); }
inline V3 operator-(V3 a, V3 b) { return V3(a.x - b.x, a.y - b.y, a.z - b.z); }
inline V3 operator-(V3 a) { return V3(-a.x, -a.y, -a.z); }
inline V3 operator*(V3 a, V3 b) { return V3(a.x * b.x, a.y * b.y, a.z * b.z); }
inline V3 operator*(V3 a, float s) { return V3(a.x * s, a.y * s, a.z * s); }
inline V3 operator*(float s, V3 a) { return V3(s * a.x, s * a.y, s * a.z); }
inline V3 operator/(V3 a, float s) {
    const float inv = 1.0f / s;
    return V3(a.x * inv, a.y * inv, a.z * inv);
}
inline V3 operator/(V3 a, V3 b) { return V3(a.x / b.x, a.y / b.y, a.z / b.z); }
inline V3& operator+=(V3& a, V3 b) { return a = a + b; }
inline V3& operator*=(V3& a, V3 b) { return a = a * b; }
inline V3& operator*=(V3& a, float s) { return a = a * s; }
inline float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) { return V3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
inline float lengthSq(V3 a) { return dot(a, a); }
inline float length(V3 a) { return std::sqrt(lengthSq(a)); }
inline V3 normalized(V3 a) {  // vec.h:153-157
    const float len = length(a);
    if (len <= 1e-20f) return V3(0.0f);
    return a / len;
}
inline V3 vabs(V3 a) { return V3(std::fabs(a.x), std::fabs(a.y), std::fabs(a.z)); }
inline V3 vmin(V3 a, V3 b) { return V3(std::min(a.x, b.x), std::min(a.y, b.y), std::min(a.z, b.z)); }
inline V3 vmax(V3 a, V3 b) { return V3(std::max(a.x, b.x), std::max(a.y, b.y), std::max(a.z, b.z)); }
inline float maxComp(V3 a) { return std::max(a.x, std::max(a.y, a.z)); }
inline bool isBlack(V3 c) { return c.x == 0.0f && c.y == 0.0f && c.z == 0.0f; }
inline float luminance(V3 c) { return 0.212671f * c.x + 0.715160f * c.y + 0.072169f * c.z; }

// ---- sampler (sobol_sampler.cpp) ----
uint32_t reverseBits(uint32_t x);
uint32_t sobolDim1(uint32_t index);
uint32_t laineKarras(uint32_t x, uint32_t seed);
uint32_t mix32(uint32_t x);
uint32_t hashCombine(uint32_t h, uint32_t v);
float toUnitFloat(uint32_t x);

struct Sobol {
    uint64_t seed = 0;
    uint32_t pixelHash = 0, index = 0, revIndex = 0, dimension = 0;
    explicit Sobol(uint64_t s) : seed(s) { startPixel(0, 0); }
    void startPixel(int x, int y);
    void startSample(int i);
    void get2DBits(uint32_t& x, uint32_t& y);
    float get1D();
    void get2D(float& u, float& v);
};

// ---- scene + BVH (scene.cpp, bvh.cpp, triangle.cpp, mesh.cpp) ----
struct Ray {
    V3 o, d;
    float tMin = 0.0f, tMax = 1e30f;
};

struct Hit {
    float t = -1.0f;
    int shape = -1, prim = -1, tri = -1;  // tri: global triangle index in scene order
    float u = 0.0f, v = 0.0f;
};

struct Scene;
Scene* buildScene(const pe_scene_desc& d);
void freeScene(Scene* s);
bool intersect(const Scene& s, const Ray& r, Hit& h);
bool intersectAny(const Scene& s, const Ray& r);
int sceneTriangles(const Scene& s);

}  // namespace oracle

// device_math.h -- float math of the Lucid hot path for gfx950 device code.
//
// Every function keeps the evaluation order of the reference expression it
// restates (file:line cited per function, relative to
// LucidRenderer/cpp_renderer/), and the library is compiled with
// -ffp-contract=off and correctly rounded float division and square root, so
// the IEEE-exact parts (ray generation, Moller-Trumbore, slab tests, MIS
// arithmetic) produce the same bits as a float CPU build.  Transcendentals
// (sin, cos, pow, log, exp) come from the device math library.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

// Functions used by the BVH traversal are also compiled for the host, so the
// exact traversal code can run on the CPU in tests (lucid_debug_host_trace).
#define LUCID_HD __host__ __device__ __forceinline__

namespace lucid {

struct v3f {
    float x, y, z;
};
LUCID_HD v3f mk(float x, float y, float z) { v3f r; r.x = x; r.y = y; r.z = z; return r; }
LUCID_HD v3f mk(float4 a) { return mk(a.x, a.y, a.z); }
LUCID_HD v3f operator+(v3f a, v3f b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
LUCID_HD v3f operator-(v3f a, v3f b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
LUCID_HD v3f operator-(v3f a) { return mk(-a.x, -a.y, -a.z); }
LUCID_HD v3f operator*(v3f a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
LUCID_HD v3f operator/(v3f a, float s) { return mk(a.x / s, a.y / s, a.z / s); }
LUCID_HD v3f cmul(v3f a, v3f b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
LUCID_HD float dot(v3f a, v3f b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
LUCID_HD float len2(v3f a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
LUCID_HD v3f cross(v3f a, v3f b) {
    return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
// std::min / std::max semantics: first argument on ties and NaN
LUCID_HD float smin(float a, float b) { return (b < a) ? b : a; }
LUCID_HD float smax(float a, float b) { return (a < b) ? b : a; }
LUCID_HD float max3(float a, float b, float c) { return smax(smax(a, b), c); }

constexpr float kDirEps2 = 1e-6f * 1e-6f;  // render_units.hpp:461,485-486
constexpr float PI_F = 3.14159265358979323846f;
constexpr float INV_PI_F = 0.318309886183790671538f;

// make_direction / make_normal (render_units.hpp:498-507)
__device__ __forceinline__ bool make_dir(v3f v, v3f& out) {
    float l2 = len2(v);
    if (l2 < kDirEps2) return false;
    out = v / sqrtf(l2);
    return true;
}
__device__ __forceinline__ v3f make_dir_or(v3f v, v3f fb) {
    v3f d;
    return make_dir(v, d) ? d : fb;
}
// render::normalize (render_units.hpp:157-163)
__device__ __forceinline__ v3f normalize_plain(v3f v) {
    float l = sqrtf(len2(v));
    return (l > 0.0f) ? v / l : v;
}
// normalize_with_length (render_units.hpp:653-658)
__device__ __forceinline__ void normalize_len(v3f v, v3f& d, float& l) {
    l = sqrtf(len2(v));
    d = v / l;
}
// reflect / refract (render_units.hpp:836-858)
__device__ __forceinline__ v3f reflect(v3f i, v3f n) {
    float c = dot(i, n);
    return make_dir_or(i - n * (2.0f * c), mk(0.0f, 0.0f, 1.0f));
}
__device__ __forceinline__ bool refract(v3f i, v3f n, float eta, v3f& out) {
    float ci = -dot(i, n);
    float s2 = eta * eta * (1.0f - ci * ci);
    if (s2 > 1.0f) return false;
    float ct = sqrtf(1.0f - s2);
    return make_dir(i * eta + n * (eta * ci - ct), out);
}

// ---- RNG: math/random.hpp:32-120 ----
__device__ __forceinline__ uint32_t fmix32(uint32_t h) {
    h ^= h >> 16; h *= 0x85ebca6bU; h ^= h >> 13; h *= 0xc2b2ae35U; h ^= h >> 16;
    return h;
}
__device__ __forceinline__ uint64_t splitmix(uint64_t x) {
    x += 0x9e3779b97f4a7c15ULL;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ULL;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebULL;
    return x ^ (x >> 31);
}
__device__ __forceinline__ uint32_t jenkins(uint32_t a) {
    a = (a + 0x7ed55d16u) + (a << 12);
    a = (a ^ 0xc761c23cu) ^ (a >> 19);
    a = (a + 0x165667b1u) + (a << 5);
    a = (a + 0xd3a2646cu) ^ (a << 9);
    a = (a + 0xfd7046c5u) + (a << 3);
    a = (a ^ 0xb55a4f09u) ^ (a >> 16);
    return a;
}
constexpr uint64_t kPcgMul = 6364136223846793005ULL;
struct Pcg {
    uint64_t state, inc;
};
__device__ __forceinline__ Pcg pcg_seed(uint32_t x, uint32_t y, uint32_t s) {
    uint32_t h1 = jenkins(x + 1u);
    uint32_t h2 = jenkins(y + 1u + h1);
    uint32_t h3 = jenkins(s + 1u + h2);
    uint64_t seed1 = ((uint64_t)fmix32(h1 ^ h3) << 32) | fmix32(h2);
    uint64_t seed2 = ((uint64_t)fmix32(h2 ^ h1) << 32) | fmix32(h3);
    Pcg r;
    r.inc = (splitmix(seed1) << 1u) | 1u;
    uint64_t st = r.inc;                       // 0 * mul + inc
    st += splitmix(seed2);
    st = st * kPcgMul + r.inc;
#pragma unroll
    for (int i = 0; i < 8; ++i) st = st * kPcgMul + r.inc;
    r.state = st;
    return r;
}
__device__ __forceinline__ uint32_t pcg32(Pcg& r) {
    uint64_t old = r.state;
    r.state = old * kPcgMul + r.inc;
    uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = (uint32_t)(old >> 59u);
    return (xs >> rot) | (xs << ((-rot) & 31u));
}
// (pcg32() >> 9) * 2^-23, exact
__device__ __forceinline__ float randf(Pcg& r) { return (float)(pcg32(r) >> 9) * 0x1p-23f; }

}  // namespace lucid

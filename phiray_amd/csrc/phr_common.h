// Shared device-side definitions for the gfx950 LBVH + path-tracer library.
//
// Layouts follow phiray's own (src/bvh/node.rs, src/trace/wgsl/layout.wgsl),
// so a buffer written here is one the reference host code could read back:
//   Aabb  32 B  {lo.xyz, pad, hi.xyz, pad}
//   Node  32 B  {lo.xyz, lo.i, hi.xyz, hi.i}; inner: children; leaf: lo.i = first | LEAF, hi.i = end
// All arithmetic that must match the CPU twin bit for bit is written as
// separate IEEE single-precision operations; the library is compiled with
// -ffp-contract=off so none of them fuse.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define PHR_LEAF 0x80000000u
#define PHR_STACK 32u
#define PHR_NONE 0xffffffffu
#define PHR_INF 3.0e38f
#define PHR_PI 3.14159265f

#define PHR_HD __host__ __device__ __forceinline__

struct PhrAabb { float lo[3]; float _a; float hi[3]; float _b; };
struct PhrNode { float lo[3]; uint32_t li; float hi[3]; uint32_t ri; };
struct PhrLink { uint32_t left, right, first, last; };
struct PhrHit { float t; uint32_t tri; float u, v; uint32_t inst; };
struct PhrRay { float o[3]; float tmax; float d[3]; float _pad; };

static_assert(sizeof(PhrAabb) == 32, "Aabb is 32 bytes");
static_assert(sizeof(PhrNode) == 32, "Node is 32 bytes");

namespace phr {

struct V3 { float x, y, z; };
PHR_HD V3 v3(float x, float y, float z) { return V3{x, y, z}; }
PHR_HD V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
PHR_HD V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
PHR_HD V3 operator-(V3 a) { return {-a.x, -a.y, -a.z}; }
PHR_HD V3 operator*(V3 a, V3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
PHR_HD V3 operator*(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
PHR_HD V3 operator/(V3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
PHR_HD float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
PHR_HD V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
PHR_HD float length(V3 a) { return sqrtf(dot(a, a)); }
PHR_HD V3 normalize(V3 a) { return a / length(a); }
// min/max with the operand order the CPU twin uses; the walk only ever
// compares their results, so signed zeros never decide anything.
PHR_HD float fmin_(float a, float b) { return b < a ? b : a; }
PHR_HD float fmax_(float a, float b) { return b > a ? b : a; }
PHR_HD float clampf(float x, float lo, float hi) { return fmin_(fmax_(x, lo), hi); }
PHR_HD V3 vmax(V3 a, float s) { return {fmax_(a.x, s), fmax_(a.y, s), fmax_(a.z, s)}; }
PHR_HD float maxc(V3 a) { return fmax_(a.x, fmax_(a.y, a.z)); }
PHR_HD float mixf(float a, float b, float t) { return a * (1.0f - t) + b * t; }
PHR_HD V3 mixv(V3 a, V3 b, float t) { return a * (1.0f - t) + b * t; }
PHR_HD float wgsl_sign(float x) { return x > 0.f ? 1.f : (x < 0.f ? -1.f : 0.f); }
// x^5 for x > 0, else 0 (the integrator's pow(x, 5.0) on its valid range).
PHR_HD float pw5(float x) { if (!(x > 0.f)) return 0.f; float x2 = x * x; return (x2 * x2) * x; }

// ---- deterministic transcendentals ----------------------------------------
// The same sequence of single-precision operations on every device, so the
// integrator's directions and weights do not depend on a vendor libm:
//  * sin/cos: Cody–Waite reduction by pi/2 (Cody & Waite, "Software Manual
//    for the Elementary Functions", 1980), pi/2 split in three (the first with
//    8 significant bits so k * part is exact), then the Taylor polynomials of
//    sin (through x^11) and cos (through x^12) on [-pi/4, pi/4] by Horner;
//  * exp2: n = floor(x + 1/2), 2^(x-n) by its Taylor series in (x-n)*ln2
//    through degree 7, scaled by 2^n through the exponent field.
// Valid for |x| < 1e5 (sin/cos) and x in [-126, 127] (exp2).
PHR_HD float sin_poly(float x) {
    float z = x * x;
    float p = -2.5052108385441720e-08f;
    p = p * z + 2.7557319223985893e-06f;
    p = p * z + -1.9841269841269841e-04f;
    p = p * z + 8.3333333333333333e-03f;
    p = p * z + -1.6666666666666667e-01f;
    return x + x * (z * p);
}
PHR_HD float cos_poly(float x) {
    float z = x * x;
    float p = 2.0876756987868099e-09f;
    p = p * z + -2.7557319223985888e-07f;
    p = p * z + 2.4801587301587302e-05f;
    p = p * z + -1.3888888888888889e-03f;
    p = p * z + 4.1666666666666667e-02f;
    p = p * z + -0.5f;
    return 1.0f + z * p;
}
PHR_HD void reduce_pio2(float x, int* q, float* r) {
    float k = floorf(x * 0.63661977236758134f + 0.5f);
    float y = x - k * 1.5703125f;
    y = y - k * 4.838267923332751e-04f;
    y = y - k * 2.5633440682570896e-12f;
    *q = (int)k;
    *r = y;
}
PHR_HD float dsin(float x) {
    int q; float r;
    reduce_pio2(x, &q, &r);
    switch (q & 3) {
        case 0: return sin_poly(r);
        case 1: return cos_poly(r);
        case 2: return -sin_poly(r);
        default: return -cos_poly(r);
    }
}
PHR_HD float dcos(float x) {
    int q; float r;
    reduce_pio2(x, &q, &r);
    switch (q & 3) {
        case 0: return cos_poly(r);
        case 1: return -sin_poly(r);
        case 2: return -cos_poly(r);
        default: return sin_poly(r);
    }
}
PHR_HD float dexp2(float x) {
    if (x < -126.f) x = -126.f;
    if (x > 127.f) x = 127.f;
    float n = floorf(x + 0.5f);
    float f = x - n;
    float y = f * 0.69314718055994531f;
    float p = 1.9841269841269841e-04f;
    p = p * y + 1.3888888888888889e-03f;
    p = p * y + 8.3333333333333333e-03f;
    p = p * y + 4.1666666666666667e-02f;
    p = p * y + 1.6666666666666667e-01f;
    p = p * y + 0.5f;
    p = p * y + 1.0f;
    p = p * y + 1.0f;
    uint32_t bits = (uint32_t)((int)n + 127) << 23;
    return p * __builtin_bit_cast(float, bits);
}

// log2 for x > 0 (normal): x = m * 2^e, m in [sqrt(1/2), sqrt(2)); ln m =
// 2 atanh(t), t = (m - 1) / (m + 1), by its series through t^11; x <= 0 -> -inf.
PHR_HD float dlog2(float x) {
    if (!(x > 0.f)) return -__builtin_inff();
    uint32_t b = __builtin_bit_cast(uint32_t, x);
    int e = (int)((b >> 23) & 0xFF) - 127;
    float m = __builtin_bit_cast(float, (b & 0x007FFFFFu) | 0x3F800000u);
    if (m > 1.41421356f) { m = m * 0.5f; e += 1; }
    float t = (m - 1.0f) / (m + 1.0f);
    float z = t * t;
    float p = 0.09090909090909091f;
    p = p * z + 0.1111111111111111f;
    p = p * z + 0.14285714285714285f;
    p = p * z + 0.2f;
    p = p * z + 0.3333333333333333f;
    p = p * z + 1.0f;
    float ln = 2.0f * (t * p);
    return (float)e + ln * 1.4426950408889634f;
}
PHR_HD float dpow(float x, float y) { return x > 0.f ? dexp2(y * dlog2(x)) : 0.0f; }
PHR_HD float dexp(float x) { return dexp2(x * 1.4426950408889634f); }

// ---- the display pixel (src/trace/wgsl/view.wgsl:62-107) ----
// Blender's Standard view: sRGB encoding of a channel clamped to [0, 1].
PHR_HD float srgb_encode(float c) {
    float x = clampf(c, 0.0f, 1.0f);
    return x <= 0.0031308f ? x * 12.92f : 1.055f * dpow(x, 1.0f / 2.4f) - 0.055f;
}
// Scene-linear to display under view words (transform, exposure in stops,
// gamma with 0 meaning 1, look): exposure, the Standard transform, gamma.
// Only transform 0 (Standard) is carried; the host refuses the others.
PHR_HD V3 view_standard(V3 lin, float exposure, float gamma) {
    float s = dexp2(exposure);
    V3 d = v3(srgb_encode(fmax_(lin.x, 0.0f) * s), srgb_encode(fmax_(lin.y, 0.0f) * s), srgb_encode(fmax_(lin.z, 0.0f) * s));
    float g = 1.0f / (gamma <= 0.0f ? 1.0f : gamma);
    return v3(dpow(clampf(d.x, 0.0f, 1.0f), g), dpow(clampf(d.y, 0.0f, 1.0f), g), dpow(clampf(d.z, 0.0f, 1.0f), g));
}
// RGBA8 in one word, red in the low byte; each channel rounded to nearest.
PHR_HD uint32_t pack_rgba8(V3 c, float a) {
    uint32_t r = (uint32_t)(clampf(c.x, 0.0f, 1.0f) * 255.0f + 0.5f);
    uint32_t g = (uint32_t)(clampf(c.y, 0.0f, 1.0f) * 255.0f + 0.5f);
    uint32_t b = (uint32_t)(clampf(c.z, 0.0f, 1.0f) * 255.0f + 0.5f);
    uint32_t q = (uint32_t)(clampf(a, 0.0f, 1.0f) * 255.0f + 0.5f);
    return r | (g << 8) | (b << 16) | (q << 24);
}
// The value shown for a pixel from its sums (util.wgsl:133-144): the view
// transform of the light's mean over the samples that met something, under
// the grid's coverage, with the grid laid over. k: samples taken.
PHR_HD V3 shown_of(V3 l, V3 gs, float cov, float gone, float k, float exposure, float gamma) {
    return view_standard(l / fmax_(k - gone, 1.0f), exposure, gamma) * (1.0f - cov / k) + gs / k;
}

// ---- the per-pixel random stream (PCG-RXS-M-XS hash) ----
PHR_HD uint32_t pcg_hash(uint32_t x) {
    uint32_t h = x * 747796405u + 2891336453u;
    h = ((h >> ((h >> 28u) + 4u)) ^ h) * 277803737u;
    return (h >> 22u) ^ h;
}
PHR_HD float rnd(uint32_t* seed) {
    *seed = pcg_hash(*seed);
    return (float)(*seed) / 4294967295.0f;
}

}  // namespace phr


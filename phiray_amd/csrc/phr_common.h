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

// ---- the display pixel (src/trace/wgsl/view.wgsl) ----
// Blender's Standard view: sRGB encoding of a channel clamped to [0, 1].
PHR_HD float srgb_encode(float c) {
    float x = clampf(c, 0.0f, 1.0f);
    return x <= 0.0031308f ? x * 12.92f : 1.055f * dpow(x, 1.0f / 2.4f) - 0.055f;
}
PHR_HD float srgb_decode(float c) { return c <= 0.04045f ? c / 12.92f : dpow((c + 0.055f) / 1.055f, 2.4f); }

// Blender's Filmic and AgX curves as phiray samples them (view.wgsl:5-34):
// display value over log2(scene value) in [VT_LO, VT_HI], 64 steps; each
// Filmic look (`phiray::Look` 1..6) is a curve of its own on the same grid;
// AgX runs its curve between an inset and an outset matrix (columns).
constexpr float VT_LO = -14.947861f, VT_HI = 10.052138f;
__constant__ const float VT_FILMIC[64] = {
    0.000015f, 0.000015f, 0.000015f, 0.000015f, 0.000015f, 0.000015f, 0.000015f, 0.001109f,
    0.002798f, 0.004849f, 0.007337f, 0.010324f, 0.013937f, 0.018276f, 0.023498f, 0.029783f,
    0.037296f, 0.046297f, 0.057015f, 0.069747f, 0.084803f, 0.102511f, 0.123240f, 0.147317f,
    0.175060f, 0.206725f, 0.242476f, 0.282333f, 0.326138f, 0.373576f, 0.424074f, 0.476835f,
    0.530496f, 0.584086f, 0.636811f, 0.687598f, 0.735542f, 0.779871f, 0.820052f, 0.855760f,
    0.886888f, 0.913539f, 0.935939f, 0.954360f, 0.969213f, 0.980878f, 0.989775f, 0.996255f,
    1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f,
    1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f,
};
__constant__ const float VT_AGX[64] = {
    0.000000f, 0.000000f, 0.000000f, 0.000000f, 0.000000f, 0.000000f, 0.000000f, 0.000015f,
    0.000094f, 0.000307f, 0.000714f, 0.001401f, 0.002474f, 0.004091f, 0.006433f, 0.009787f,
    0.014516f, 0.021097f, 0.030173f, 0.042574f, 0.057306f, 0.073883f, 0.092617f, 0.114256f,
    0.138886f, 0.167071f, 0.199246f, 0.235968f, 0.277732f, 0.324850f, 0.377462f, 0.435243f,
    0.495499f, 0.553566f, 0.607904f, 0.657973f, 0.703691f, 0.745144f, 0.782605f, 0.816304f,
    0.846638f, 0.873680f, 0.897952f, 0.919880f, 0.939710f, 0.957668f, 0.973975f, 0.988673f,
    1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f,
    1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f,
};
__constant__ const float VT_LOOKS[384] = {
    0.000000f, 0.000000f, 0.000000f, 0.000000f, 0.000000f, 0.000000f, 0.000000f, 0.000070f,
    0.000193f, 0.000363f, 0.000601f, 0.000931f, 0.001389f, 0.002026f, 0.002910f, 0.004138f,
    0.005841f, 0.008200f, 0.011463f, 0.015966f, 0.022162f, 0.030652f, 0.042221f, 0.057863f,
    0.078793f, 0.106414f, 0.142202f, 0.187489f, 0.243112f, 0.308980f, 0.383689f, 0.464395f,
    0.547092f, 0.627392f, 0.701407f, 0.766409f, 0.821117f, 0.865535f, 0.900558f, 0.927541f,
    0.947961f, 0.963206f, 0.974473f, 0.982737f, 0.988766f, 0.993146f, 0.996319f, 0.998612f,
    1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f,
    1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f,
    0.000000f, 0.000000f, 0.000000f, 0.000000f, 0.000000f, 0.000000f, 0.000000f, 0.000233f,
    0.000620f, 0.001128f, 0.001793f, 0.002666f, 0.003810f, 0.005309f, 0.007269f, 0.009833f,
    0.013179f, 0.017541f, 0.023213f, 0.030568f, 0.040067f, 0.052275f, 0.067866f, 0.087615f,
    0.112373f, 0.143010f, 0.180318f, 0.224866f, 0.276828f, 0.335802f, 0.400681f, 0.469646f,
    0.540236f, 0.609827f, 0.675954f, 0.736553f, 0.790239f, 0.836358f, 0.874903f, 0.906341f,
    0.931432f, 0.951064f, 0.966136f, 0.977482f, 0.985838f, 0.991826f, 0.995958f, 0.998647f,
    1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f,
    1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f,
    0.000000f, 0.000000f, 0.000000f, 0.000000f, 0.000000f, 0.000000f, 0.000000f, 0.000510f,
    0.001330f, 0.002363f, 0.003665f, 0.005306f, 0.007372f, 0.009971f, 0.013237f, 0.017335f,
    0.022469f, 0.028884f, 0.036880f, 0.046810f, 0.059090f, 0.074194f, 0.092650f, 0.115017f,
    0.141861f, 0.173696f, 0.210924f, 0.253744f, 0.302067f, 0.355438f, 0.412989f, 0.473463f,
    0.535105f, 0.596234f, 0.655290f, 0.710781f, 0.761550f, 0.806854f, 0.846356f, 0.880079f,
    0.908313f, 0.931529f, 0.950296f, 0.965209f, 0.976851f, 0.985758f, 0.992407f, 0.997209f,
    1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f,
    1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f,
    0.000000f, 0.000000f, 0.000000f, 0.000000f, 0.000000f, 0.000000f, 0.000000f, 0.001808f,
    0.004533f, 0.007732f, 0.011485f, 0.015885f, 0.021036f, 0.027059f, 0.034090f, 0.042283f,
    0.051810f, 0.062859f, 0.075637f, 0.090362f, 0.107263f, 0.126574f, 0.148519f, 0.173307f,
    0.201111f, 0.232056f, 0.266193f, 0.303481f, 0.343773f, 0.386797f, 0.432150f, 0.479308f,
    0.526904f, 0.574394f, 0.621585f, 0.667736f, 0.712148f, 0.754204f, 0.793394f, 0.829342f,
    0.861800f, 0.890655f, 0.915907f, 0.937653f, 0.956064f, 0.971365f, 0.983812f, 0.993680f,
    1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f,
    1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f,
    0.000000f, 0.000000f, 0.000000f, 0.000000f, 0.000000f, 0.000000f, 0.000000f, 0.003014f,
    0.007419f, 0.012418f, 0.018087f, 0.024509f, 0.031775f, 0.039988f, 0.049257f, 0.059700f,
    0.071444f, 0.084623f, 0.099377f, 0.115848f, 0.134178f, 0.154507f, 0.176963f, 0.201658f,
    0.228681f, 0.258089f, 0.289897f, 0.324071f, 0.360515f, 0.399071f, 0.439508f, 0.481524f,
    0.523400f, 0.565001f, 0.606881f, 0.648556f, 0.689527f, 0.729308f, 0.767438f, 0.803505f,
    0.837159f, 0.868121f, 0.896191f, 0.921245f, 0.943236f, 0.962183f, 0.978162f, 0.991297f,
    1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f,
    1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f,
    0.000000f, 0.000000f, 0.000000f, 0.000000f, 0.000000f, 0.000000f, 0.000000f, 0.004685f,
    0.011326f, 0.018609f, 0.026591f, 0.035334f, 0.044905f, 0.055372f, 0.066812f, 0.079302f,
    0.092923f, 0.107759f, 0.123898f, 0.141424f, 0.160426f, 0.180986f, 0.203184f, 0.227091f,
    0.252769f, 0.280265f, 0.309608f, 0.340804f, 0.373832f, 0.408639f, 0.445135f, 0.483188f,
    0.520119f, 0.556217f, 0.593106f, 0.630501f, 0.668084f, 0.705500f, 0.742373f, 0.778309f,
    0.812910f, 0.845788f, 0.876574f, 0.904934f, 0.930579f, 0.953271f, 0.972832f, 0.989149f,
    1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f,
    1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f,
};
__constant__ const float VT_AGX_IN[9] = {0.498585f, 0.125929f, 0.097476f, 0.346742f, 0.681639f, 0.191866f, 0.097839f, 0.123183f, 0.678532f};
__constant__ const float VT_AGX_OUT[9] = {2.055217f, -0.287805f, -0.221530f, -0.912124f, 1.396234f, -0.286374f, -0.141522f, -0.097287f, 1.507346f};

// The curve at x: linear interpolation between the two samples around
// log2(x) (view.wgsl:37-54). base: the first sample of the curve.
__device__ __forceinline__ float vt_curve(const float* base, float x) {
    float t = clampf((dlog2(fmax_(x, 1e-10f)) - VT_LO) / (VT_HI - VT_LO), 0.0f, 1.0f) * 63.0f;
    float fl = floorf(t);
    uint32_t i = (uint32_t)fl, j = i + 1u < 63u ? i + 1u : 63u;
    return mixf(base[i], base[j], t - fl);
}
__device__ __forceinline__ V3 mat3_mul(const float* m, V3 v) {  // m: three columns
    return v3(m[0] * v.x + m[3] * v.y + m[6] * v.z, m[1] * v.x + m[4] * v.y + m[7] * v.z, m[2] * v.x + m[5] * v.y + m[8] * v.z);
}

// Scene-linear to display under the camera's view words (transform 0
// Standard, 1 Filmic, 2 AgX; exposure in stops; gamma, 0 meaning 1; look)
// as view_transform_of does it (view.wgsl:62-86): exposure, the transform,
// then gamma.
__device__ __forceinline__ V3 view_transform(V3 lin, float kind_w, float exposure, float gamma, float look_w) {
    float s = dexp2(exposure);
    V3 c = v3(fmax_(lin.x, 0.0f) * s, fmax_(lin.y, 0.0f) * s, fmax_(lin.z, 0.0f) * s);
    uint32_t kind = (uint32_t)(kind_w + 0.5f);
    V3 d;
    if (kind == 1u) {
        uint32_t look = (uint32_t)(look_w + 0.5f);
        const float* curve = (look >= 1u && look <= 6u) ? VT_LOOKS + (look - 1u) * 64u : VT_FILMIC;
        d = v3(vt_curve(curve, c.x), vt_curve(curve, c.y), vt_curve(curve, c.z));
    } else if (kind == 2u) {
        V3 i = mat3_mul(VT_AGX_IN, c);
        V3 y = v3(vt_curve(VT_AGX, i.x), vt_curve(VT_AGX, i.y), vt_curve(VT_AGX, i.z));
        V3 o = mat3_mul(VT_AGX_OUT, v3(srgb_decode(y.x), srgb_decode(y.y), srgb_decode(y.z)));
        d = v3(srgb_encode(o.x), srgb_encode(o.y), srgb_encode(o.z));
    } else {
        d = v3(srgb_encode(c.x), srgb_encode(c.y), srgb_encode(c.z));
    }
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
__device__ __forceinline__ V3 shown_of(V3 l, V3 gs, float cov, float gone, float k, float4 view) {
    return view_transform(l / fmax_(k - gone, 1.0f), view.x, view.y, view.z, view.w) * (1.0f - cov / k) + gs / k;
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


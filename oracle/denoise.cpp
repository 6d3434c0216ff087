// CPU oracle: phiray's a-trous denoiser.
//
// TEST INFRASTRUCTURE ONLY (see oracle.h). Restates src/kernel/denoise.wgsl
// (main, 47-134; guides and helpers, 1-45), srgb from
// src/trace/wgsl/view.wgsl:88-93, and the four-pass driver of
// src/trace/read.rs:85-119. pow(x, y) is exp2(y log2 x) and exp(x) is
// exp2(x log2 e), both through the deterministic exp2/log2 of walk.cpp. The
// last pass writes the display plane and the packed pixel (denoise.wgsl:
// 124-133) through path.cpp's view transform.
#include "oracle.h"

#include <cmath>
#include <cstring>

namespace {

struct V3 { float x, y, z; };
inline V3 add(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 mul(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline V3 divs(V3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
inline V3 divv(V3 a, V3 b) { return {a.x / b.x, a.y / b.y, a.z / b.z}; }
inline V3 mulv(V3 a, V3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
inline float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline float fmax_(float a, float b) { return b > a ? b : a; }
inline float fmin_(float a, float b) { return b < a ? b : a; }
inline V3 vmaxs(V3 a, float s) { return {fmax_(a.x, s), fmax_(a.y, s), fmax_(a.z, s)}; }
inline float bits_f(uint32_t u) { float f; std::memcpy(&f, &u, 4); return f; }
inline uint32_t f_bits(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }

inline float dpow(float x, float y) { return x > 0.f ? orc_exp2(y * orc_log2(x)) : 0.0f; }
inline float dexp(float x) { return orc_exp2(x * 1.4426950408889634f); }

inline float srgb1(float c) {  // view.wgsl:88-93
    float x = fmin_(fmax_(c, 0.0f), 1.0f);
    float lo = x * 12.92f;
    float hi = 1.055f * dpow(x, 1.0f / 2.4f) - 0.055f;
    return x <= 0.0031308f ? lo : hi;
}

struct Img {
    const uint32_t* px;
    size_t n;
    uint32_t w, h, round;
    float f_at(size_t i) const { return bits_f(px[i]); }
    V3 v3_at(size_t base, size_t i) const { size_t o = base + 3 * i; return {f_at(o), f_at(o + 1), f_at(o + 2)}; }
    float inv_n(size_t i) const { uint32_t c = px[31 * n + i]; return 1.0f / (float)(c > 1u ? c : 1u); }
    V3 albedo_of(size_t i) const {
        V3 a = mul(v3_at(7 * n, i), inv_n(i));
        return {srgb1(a.x), srgb1(a.y), srgb1(a.z)};
    }
    V3 normal_of(size_t i) const {
        V3 v = add(v3_at(10 * n, i), V3{1e-9f, 1e-9f, 1e-9f});
        return divs(v, std::sqrt(dot(v, v)));
    }
    float depth_of(size_t i) const { return f_at(13 * n + i) * inv_n(i); }
    V3 demodulated(size_t i) const { return divv(mul(v3_at(1 * n, i), inv_n(i)), vmaxs(albedo_of(i), 0.05f)); }
    V3 read_in(size_t i) const {
        if (round == 0u) return demodulated(i);
        return v3_at(((round & 1u) == 1u ? 14 : 17) * n, i);
    }
};

inline float luma_of(V3 c) { return dot(c, V3{0.2126f, 0.7152f, 0.0722f}); }

}  // namespace

extern "C" void orc_denoise_pass(uint32_t* px, uint32_t w, uint32_t h, uint32_t step, uint32_t round,
                                 uint32_t last, uint32_t keep, const float strength[4], const float view[4]) {
    const float KERNEL[3] = {0.375f, 0.25f, 0.0625f};
    Img im{px, (size_t)w * h, w, h, round};
    size_t n = im.n;
    for (size_t i = 0; i < n; ++i) {
        int x = (int)(i % w), y = (int)(i / w);
        V3 centre = im.read_in(i);
        if (round == 0u) {
            float near_top = 0.0f, near_n = 0.0f;
            V3 near_sum{0, 0, 0};
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    int sx = x + dx, sy = y + dy;
                    if ((dx == 0 && dy == 0) || sx < 0 || sy < 0 || sx >= (int)w || sy >= (int)h) continue;
                    V3 c = im.read_in((size_t)sy * w + sx);
                    near_top = fmax_(near_top, luma_of(c));
                    near_sum = add(near_sum, c);
                    near_n = near_n + 1.0f;
                }
            if (near_n > 0.0f && luma_of(centre) > 1.5f * near_top + 0.05f) centre = divs(near_sum, near_n);
        }
        V3 cn = im.normal_of(i), ca = im.albedo_of(i);
        float cd = im.depth_of(i), lc = luma_of(centre);
        V3 sum{0, 0, 0};
        float weight = 0.0f;
        for (int dy = -2; dy <= 2; ++dy)
            for (int dx = -2; dx <= 2; ++dx) {
                int sx = x + dx * (int)step, sy = y + dy * (int)step;
                if (sx < 0 || sy < 0 || sx >= (int)w || sy >= (int)h) continue;
                size_t j = (size_t)sy * w + sx;
                if ((im.depth_of(j) < 0.0f) != (cd < 0.0f)) continue;
                V3 tap = (j == i) ? centre : im.read_in(j);
                float dn = fmax_(0.0f, 1.0f - dot(cn, im.normal_of(j)));
                V3 dav = sub(im.albedo_of(j), ca);
                float da = std::sqrt(dot(dav, dav));
                float dd = std::fabs(im.depth_of(j) - cd) / fmax_(std::fabs(cd), 1e-3f);
                float dl = std::fabs(luma_of(tap) - lc) / (0.08f + lc);
                float stop = dexp(-dn * strength[0] - da * strength[1] - dd * strength[2] - dl * dl * strength[3]);
                float k = KERNEL[dx < 0 ? -dx : dx] * KERNEL[dy < 0 ? -dy : dy];
                sum = add(sum, mul(mul(tap, k), stop));
                weight = weight + k * stop;
            }
        V3 out = weight > 0.0f ? divs(sum, weight) : centre;
        if (last == 0u) {
            size_t o = ((round & 1u) == 1u ? 17 : 14) * n + 3 * i;
            px[o] = f_bits(out.x); px[o + 1] = f_bits(out.y); px[o + 2] = f_bits(out.z);
            continue;
        }
        float k = im.inv_n(i);
        float met = 1.0f - im.f_at(33 * n + i) * k;
        V3 lin = met > 0.5f * k ? mulv(out, vmaxs(ca, 0.05f)) : V3{0, 0, 0};
        if (keep == 0u) {
            V3 v = divs(lin, k);
            size_t o = 1 * n + 3 * i;
            px[o] = f_bits(v.x); px[o + 1] = f_bits(v.y); px[o + 2] = f_bits(v.z);
        }
        float cov = im.f_at(23 * n + i) * k;
        V3 grid = mul(im.v3_at(20 * n, i), k);
        V3 m = divs(lin, fmax_(met, k));
        float mv[3] = {m.x, m.y, m.z}, d[3];
        orc_view_transform(mv, view, d);
        V3 shown = add(mul(V3{d[0], d[1], d[2]}, 1.0f - cov), grid);
        px[4 * n + 3 * i] = f_bits(shown.x); px[4 * n + 3 * i + 1] = f_bits(shown.y); px[4 * n + 3 * i + 2] = f_bits(shown.z);
        float sv[3] = {shown.x, shown.y, shown.z};
        px[i] = orc_pack_rgba8(sv, im.f_at(24 * n + i) * k);
    }
}

extern "C" void orc_denoise(uint32_t* px, uint32_t w, uint32_t h, uint32_t keep, const float view[4]) {
    const float strength[4] = {64.0f, 8.0f, 40.0f, 7.0f};
    const uint32_t steps[4] = {1, 2, 4, 8};
    for (uint32_t r = 0; r < 4; ++r) orc_denoise_pass(px, w, h, steps[r], r, steps[r] == 8 ? 1u : 0u, keep, strength, view);
}

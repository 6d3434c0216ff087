// phiray's a-trous denoiser on gfx950 (src/kernel/denoise.wgsl, driven as
// src/trace/read.rs:85-119 drives it): four passes at tap spacing 1, 2, 4, 8
// over the image buffer, guided by the first hit's albedo, normal and depth and
// by the light itself, the first pass clamping fireflies, the last putting the
// albedo back into the linear sums. One thread per pixel; passes ping-pong
// between the two scratch planes, so a pass never reads what it writes.
// The last pass also writes the display plane and the packed RGBA8 word
// under the camera's view transform, exposure and gamma.
#include "phr_runtime.h"

using namespace phr;

namespace {

// The image buffer as the a-trous passes read it. CACHED: each pixel's guides
// (sRGB albedo, unit normal, mean depth, 1 / count) and the first pass's
// input come from a table the prep kernel wrote with these same functions,
// instead of being derived again at every tap of every pass.
template <bool CACHED>
struct Img {
    const uint32_t* __restrict__ px;
    size_t n;
    uint32_t round;
    const float4* __restrict__ feat;  // CACHED: 2 a pixel, (albedo, depth) (normal, 1 / count)
    const float* __restrict__ col0;   // CACHED: the first pass's input, 3 a pixel
    __device__ float f_at(size_t i) const { return __uint_as_float(px[i]); }
    __device__ V3 v3_at(size_t base, size_t i) const { size_t o = base + 3 * i; return v3(f_at(o), f_at(o + 1), f_at(o + 2)); }
    __device__ float inv_n(size_t i) const {
        if constexpr (CACHED) return feat[2 * i + 1].w;
        uint32_t c = px[31 * n + i];
        return 1.0f / (float)(c > 1u ? c : 1u);
    }
    __device__ static float srgb1(float c) {  // view.wgsl:88-93
        float x = fmin_(fmax_(c, 0.0f), 1.0f);
        float lo = x * 12.92f;
        float hi = 1.055f * dpow(x, 1.0f / 2.4f) - 0.055f;
        return x <= 0.0031308f ? lo : hi;
    }
    __device__ V3 albedo_of(size_t i) const {
        if constexpr (CACHED) {
            const float4 f = feat[2 * i];
            return v3(f.x, f.y, f.z);
        }
        V3 a = v3_at(7 * n, i) * inv_n(i);
        return v3(srgb1(a.x), srgb1(a.y), srgb1(a.z));
    }
    __device__ V3 normal_of(size_t i) const {
        if constexpr (CACHED) {
            const float4 f = feat[2 * i + 1];
            return v3(f.x, f.y, f.z);
        }
        return normalize(v3_at(10 * n, i) + v3(1e-9f, 1e-9f, 1e-9f));
    }
    __device__ float depth_of(size_t i) const {
        if constexpr (CACHED) return feat[2 * i].w;
        return f_at(13 * n + i) * inv_n(i);
    }
    __device__ V3 read_in(size_t i) const {
        if (round == 0u) {
            if constexpr (CACHED) return v3(col0[3 * i], col0[3 * i + 1], col0[3 * i + 2]);
            V3 a = v3_at(1 * n, i) * inv_n(i);
            V3 b = vmax(albedo_of(i), 0.05f);
            return v3(a.x / b.x, a.y / b.y, a.z / b.z);
        }
        return v3_at(((round & 1u) == 1u ? 14 : 17) * n, i);
    }
};

__device__ __forceinline__ float luma_of(V3 c) { return dot(c, v3(0.2126f, 0.7152f, 0.0722f)); }

// The guide table and the first pass's input, once a denoise.
__global__ __launch_bounds__(256) void k_denoise_prep(const uint32_t* __restrict__ px, size_t n, float4* __restrict__ feat,
                                                      float* __restrict__ col0) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const Img<false> im{px, n, 0u, nullptr, nullptr};
    const V3 a = im.albedo_of(i), nr = im.normal_of(i), c = im.read_in(i);
    feat[2 * i] = make_float4(a.x, a.y, a.z, im.depth_of(i));
    feat[2 * i + 1] = make_float4(nr.x, nr.y, nr.z, im.inv_n(i));
    col0[3 * i] = c.x;
    col0[3 * i + 1] = c.y;
    col0[3 * i + 2] = c.z;
}

template <bool CACHED>
__global__ __launch_bounds__(256) void k_denoise(uint32_t* __restrict__ px, uint32_t w, uint32_t h, uint32_t step, uint32_t round,
                                                 uint32_t last, uint32_t keep, float s0, float s1, float s2, float s3,
                                                 float4 view, const float4* __restrict__ feat,
                                                 const float* __restrict__ col0) {
    const size_t n = (size_t)w * h;
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float KERNEL[3] = {0.375f, 0.25f, 0.0625f};
    Img<CACHED> im{px, n, round, feat, col0};
    int x = (int)(i % w), y = (int)(i / w);
    V3 centre = im.read_in(i);
    if (round == 0u) {
        float near_top = 0.0f, near_n = 0.0f;
        V3 near_sum = v3(0, 0, 0);
        for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx) {
                int sx = x + dx, sy = y + dy;
                if ((dx == 0 && dy == 0) || sx < 0 || sy < 0 || sx >= (int)w || sy >= (int)h) continue;
                V3 c = im.read_in((size_t)sy * w + sx);
                near_top = fmax_(near_top, luma_of(c));
                near_sum = near_sum + c;
                near_n = near_n + 1.0f;
            }
        if (near_n > 0.0f && luma_of(centre) > 1.5f * near_top + 0.05f) centre = near_sum / near_n;
    }
    V3 cn = im.normal_of(i), ca = im.albedo_of(i);
    float cd = im.depth_of(i), lc = luma_of(centre);
    V3 sum = v3(0, 0, 0);
    float weight = 0.0f;
    for (int dy = -2; dy <= 2; ++dy)
        for (int dx = -2; dx <= 2; ++dx) {
            int sx = x + dx * (int)step, sy = y + dy * (int)step;
            if (sx < 0 || sy < 0 || sx >= (int)w || sy >= (int)h) continue;
            size_t j = (size_t)sy * w + sx;
            if ((im.depth_of(j) < 0.0f) != (cd < 0.0f)) continue;
            V3 tap = (j == i) ? centre : im.read_in(j);
            float dn = fmax_(0.0f, 1.0f - dot(cn, im.normal_of(j)));
            V3 dav = im.albedo_of(j) - ca;
            float da = sqrtf(dot(dav, dav));
            float dd = fabsf(im.depth_of(j) - cd) / fmax_(fabsf(cd), 1e-3f);
            float dl = fabsf(luma_of(tap) - lc) / (0.08f + lc);
            float stop = dexp(-dn * s0 - da * s1 - dd * s2 - dl * dl * s3);
            float k = KERNEL[dx < 0 ? -dx : dx] * KERNEL[dy < 0 ? -dy : dy];
            sum = sum + tap * k * stop;
            weight = weight + k * stop;
        }
    V3 out = weight > 0.0f ? sum / weight : centre;
    if (last == 0u) {
        size_t o = ((round & 1u) == 1u ? 17 : 14) * n + 3 * i;
        px[o] = __float_as_uint(out.x);
        px[o + 1] = __float_as_uint(out.y);
        px[o + 2] = __float_as_uint(out.z);
        return;
    }
    float k = im.inv_n(i);
    float met = 1.0f - im.f_at(33 * n + i) * k;
    V3 lin = met > 0.5f * k ? out * vmax(ca, 0.05f) : v3(0, 0, 0);
    if (keep == 0u) {
        V3 v = lin / k;
        size_t o = 1 * n + 3 * i;
        px[o] = __float_as_uint(v.x);
        px[o + 1] = __float_as_uint(v.y);
        px[o + 2] = __float_as_uint(v.z);
    }
    // The view transform once, under the grid the tracer summed, as the
    // tracer composes an unfiltered pixel (denoise.wgsl:124-133).
    const float cov = im.f_at(23 * n + i) * k;
    const V3 grid = im.v3_at(20 * n, i) * k;
    const V3 shown = view_transform(lin / fmax_(met, k), view.x, view.y, view.z, view.w) * (1.0f - cov) + grid;
    px[4 * n + 3 * i] = __float_as_uint(shown.x);
    px[4 * n + 3 * i + 1] = __float_as_uint(shown.y);
    px[4 * n + 3 * i + 2] = __float_as_uint(shown.z);
    px[i] = pack_rgba8(shown, im.f_at(24 * n + i) * k);
}

}  // namespace

// One pass (the `denoise` kernel of a recorded Pass).
int phr_denoise_pass_enqueue(phr_gpu* g, uint32_t* px, uint32_t w, uint32_t h, uint32_t step, uint32_t round, uint32_t last,
                             uint32_t keep, const float s[4], const float view[4]) {
    const size_t n = (size_t)w * h;
    if (n == 0) return PHR_OK;
    hipLaunchKernelGGL(k_denoise<false>, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, g->stream, px, w, h, step, round, last,
                       keep, s[0], s[1], s[2], s[3], make_float4(view[0], view[1], view[2], view[3]), nullptr, nullptr);
    PHR_HIP(hipGetLastError());
    return PHR_OK;
}

// Enqueue the four passes of Tracer::denoise_with over an image buffer of w x h.
// With `scratch` (DENOISE_SCRATCH_WORDS words a pixel, 16-B aligned) the
// guides are derived once into it; without, every pass derives them again.
// view: the camera's colour-management words (transform 0, exposure, gamma).
int phr_denoise_enqueue(phr_gpu* g, uint32_t* px, uint32_t w, uint32_t h, int keep, float* scratch, const float view[4]) {
    const float s[4] = {64.0f, 8.0f, 40.0f, 7.0f};
    const uint32_t steps[4] = {1, 2, 4, 8};
    const size_t n = (size_t)w * h;
    if (n == 0) return PHR_OK;
    const dim3 grid((uint32_t)((n + 255) / 256));
    if (scratch) {
        float4* feat = (float4*)scratch;
        float* col0 = scratch + 8 * n;
        hipLaunchKernelGGL(k_denoise_prep, grid, dim3(256), 0, g->stream, (const uint32_t*)px, n, feat, col0);
        for (uint32_t r = 0; r < 4; ++r)
            hipLaunchKernelGGL(k_denoise<true>, grid, dim3(256), 0, g->stream, px, w, h, steps[r], r, steps[r] == 8 ? 1u : 0u,
                               keep ? 1u : 0u, s[0], s[1], s[2], s[3], make_float4(view[0], view[1], view[2], view[3]), (const float4*)feat,
                               (const float*)col0);
    } else {
        for (uint32_t r = 0; r < 4; ++r)
            hipLaunchKernelGGL(k_denoise<false>, grid, dim3(256), 0, g->stream, px, w, h, steps[r], r, steps[r] == 8 ? 1u : 0u,
                               keep ? 1u : 0u, s[0], s[1], s[2], s[3], make_float4(view[0], view[1], view[2], view[3]), nullptr, nullptr);
    }
    PHR_HIP(hipGetLastError());
    return PHR_OK;
}

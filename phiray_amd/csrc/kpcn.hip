// The learned denoiser beside kernel::Denoise: a small kernel-predicting
// convolutional network (KPCN). For each pixel it predicts a 5x5 filter,
// normalised by a softmax. The filter is applied to the albedo-demodulated
// radiance, and the result is multiplied back by the albedo.
//
// Network (NHWC tensors, channels a multiple of 16, BF16 activations, FP32
// accumulation, FP32 master weights; level l is the image at 1/2^l size):
//
//   e0a 16->32, e0b 32->32        level 0  (skip0 = e0b)
//   avg-pool 2x2
//   e1a 32->48, e1b 48->48        level 1  (skip1 = e1b)
//   avg-pool 2x2
//   e2a 48->64, e2b 64->64        level 2  (skip2 = e2b)
//   avg-pool 2x2
//   e3  64->96                    level 3
//   nearest x2 + concat skip2 -> d2 160->64   level 2
//   nearest x2 + concat skip1 -> d1 112->48   level 1
//   nearest x2 + concat skip0 -> d0  80->32   level 0
//   out 32->32 (25 used: the filter logits, no ReLU)
//
// Every conv is 3x3, stride 1, zero padding 1, ReLU after it except `out`.
// There are 337,801 trainable weights (341,552 with the padded input and
// output channels, which stay zero).
//
// Input: a frame of 10 FP32 channels per pixel, which is the tracer's image
// buffer over the per-pixel count (src/trace/read.rs:241-307). The channels
// are radiance 3, albedo 3, normal 3 and depth 1. Features are
// log1p(radiance / max(albedo, 0.05)), then albedo, normal, and
// depth / (1 + depth) (-1 where nothing was hit), padded to 16 channels. The
// image is padded to multiples of 8 by repeating its edge.
//
// Kernels are hand-written for gfx950:
//  * conv: an implicit GEMM on v_mfma_f32_16x16x32_bf16. D[co][pixel] =
//    W[co][tap, ci] * X[tap, ci][pixel]. A block owns a 32-column tile and all
//    output channels; the input tile and its halo are staged in LDS, read
//    straight from the finer level (2x2 average) or from the coarser output
//    beside the skip (upsampled concatenation), so the pooled and concatenated
//    tensors exist only for training. The same kernel computes the input
//    gradient, with the weights flipped and transposed.
//  * wgrad: dW[tap][co][ci] = sum_p dZ[p][co] X[p + off(tap)][ci]. The pixel
//    tiles are staged transposed in LDS, so the reduction over pixels is the
//    MFMA K dimension. Each block writes a partial for its pixel chunk, and a
//    second kernel sums the chunks in a fixed order, so the result is
//    deterministic.
//  * apply: softmax over the 25 logits, the 5x5 filter, remodulation; the
//    relMSE loss mean((y - t)^2 / (t^2 + 0.01)) and its gradient w.r.t. the
//    logits.
//  * Adam over the FP32 master weights, then a repack to the BF16 forward
//    [tap][co][ci] and backward [8 - tap][ci][co] layouts.
#include "phr_runtime.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

using namespace phr;

namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

struct LayerDef {
    uint32_t cin, cout, level, relu;
};

constexpr int NLAYERS = 11;
enum { E0A, E0B, E1A, E1B, E2A, E2B, E3, D2, D1, D0, OUT };
constexpr LayerDef LAYERS[NLAYERS] = {
    {16, 32, 0, 1}, {32, 32, 0, 1}, {32, 48, 1, 1}, {48, 48, 1, 1}, {48, 64, 2, 1}, {64, 64, 2, 1},
    {64, 96, 3, 1}, {160, 64, 2, 1}, {112, 48, 1, 1}, {80, 32, 0, 1}, {32, 32, 0, 0},
};
constexpr uint32_t FEAT = 16;    // input feature channels (10 used)
constexpr uint32_t FRAME = 10;   // frame channels
constexpr uint32_t TAPS = 25;    // 5x5 filter
constexpr uint32_t LOGITS = 32;  // out channels (TAPS used)

__device__ __forceinline__ uint16_t to_bf16(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
    u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest, ties to even
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float from_bf16(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }
__device__ __forceinline__ float from_bf16s(short h) { return from_bf16((uint16_t)h); }

__device__ __forceinline__ f32x4 mfma16(bf16x4 a, bf16x4 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}
// gfx950's 16x16x32 form: 8 bf16 of K a lane. A and B fragments are loaded
// with the same lane-slot to channel map, so the product does not depend on
// the order the hardware gives the K slots.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
__device__ __forceinline__ f32x4 mfma32(uint4 a, uint4 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c, 0, 0,
                                                   0);
}

// ---- features ---------------------------------------------------------------
__global__ __launch_bounds__(256) void k_features(const float* __restrict__ fr, uint32_t B, uint32_t H, uint32_t W,
                                                  uint32_t Hp, uint32_t Wp, uint16_t* __restrict__ out) {
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= (uint64_t)B * Hp * Wp) return;
    const uint32_t x = (uint32_t)(i % Wp);
    const uint64_t r = i / Wp;
    const uint32_t y = (uint32_t)(r % Hp), b = (uint32_t)(r / Hp);
    const uint32_t sx = min(x, W - 1), sy = min(y, H - 1);
    const float* f = fr + (((uint64_t)b * H + sy) * W + sx) * FRAME;
    uint16_t v[FEAT];
    for (int c = 0; c < 3; ++c) {
        float d = f[c] / fmaxf(f[3 + c], 0.05f);
        v[c] = to_bf16(log1pf(fmaxf(d, 0.0f)));
    }
    for (int c = 3; c < 9; ++c) v[c] = to_bf16(f[c]);
    float z = f[9];
    v[9] = to_bf16(z < 0.0f ? -1.0f : z / (1.0f + z));
    for (int c = 10; c < (int)FEAT; ++c) v[c] = 0;
    uint4* o = (uint4*)(out + i * FEAT);
    o[0] = make_uint4(v[0] | (uint32_t)v[1] << 16, v[2] | (uint32_t)v[3] << 16, v[4] | (uint32_t)v[5] << 16,
                      v[6] | (uint32_t)v[7] << 16);
    o[1] = make_uint4(v[8] | (uint32_t)v[9] << 16, 0u, 0u, 0u);
}

// ---- 3x3 conv as an implicit GEMM ---------------------------------------------
struct ConvArgs {
    const uint16_t* x;    // [P][cin] (see the input forms below)
    const uint16_t* w;    // [9][cout][cin]
    const float* bias;    // [cout] or null
    const uint16_t* mask; // [P][cout] or null: output zeroed where mask <= 0 (ReLU')
    uint16_t* y16;        // [P][cout] or null
    float* y32;           // [P][cout] or null
    uint32_t B, H, W, cin, cout, relu;
    // Input forms, read as they are staged (the standalone kernels' bits):
    //  pooled: x is [B][2H][2W][cin]; the input is each 2x2 block's average;
    //  xc:     the input's first c1 channels are xc [B][H/2][W/2][c1] at
    //          (y/2, x/2) (nearest upsampling), the rest x [B][H][W][cin - c1].
    uint32_t pooled;
    const uint16_t* xc;
    uint32_t c1;
};

// Eight input channels [c, c + 8) of pixel (gy, gx) of image img, in the
// input form `a` describes.
__device__ __forceinline__ uint4 conv_input(const ConvArgs& a, uint64_t b, int gy, int gx, uint32_t c) {
    if (a.pooled) {  // k_pool's sum order and rounding
        const uint64_t W2 = 2ull * a.W;
        float acc[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        for (int dy = 0; dy < 2; ++dy)
            for (int dx = 0; dx < 2; ++dx) {
                const uint4 v = *(const uint4*)(a.x + ((b * 2 * a.H + 2 * gy + dy) * W2 + 2 * gx + dx) * a.cin + c);
                const uint32_t w4[4] = {v.x, v.y, v.z, v.w};
                for (int e = 0; e < 4; ++e) {
                    acc[2 * e] += from_bf16((uint16_t)(w4[e] & 0xffffu));
                    acc[2 * e + 1] += from_bf16((uint16_t)(w4[e] >> 16));
                }
            }
        uint32_t o[4];
        for (int e = 0; e < 4; ++e) o[e] = to_bf16(acc[2 * e] * 0.25f) | (uint32_t)to_bf16(acc[2 * e + 1] * 0.25f) << 16;
        return make_uint4(o[0], o[1], o[2], o[3]);
    }
    if (a.xc && c < a.c1)
        return *(const uint4*)(a.xc + ((b * (a.H / 2) + gy / 2) * (a.W / 2) + gx / 2) * a.c1 + c);
    if (a.xc) return *(const uint4*)(a.x + ((b * a.H + gy) * (uint64_t)a.W + gx) * (a.cin - a.c1) + (c - a.c1));
    return *(const uint4*)(a.x + ((b * a.H + gy) * (uint64_t)a.W + gx) * a.cin + c);
}

// A block computes a tile of 4 RW rows x 32 columns of one image, all output
// channels. The input tile with its one-pixel halo is staged in LDS 32
// channels at a time (16 B a load, coalesced along the channels). Every one of
// the nine taps then reads its B fragments from LDS. The A fragments (weights)
// come from global memory through the caches, and each is shared by the
// wave's 2 RW pixel fragments.
constexpr int TILE_W = 32;
constexpr int CHUNK = 32;            // input channels staged at once
constexpr int HALO_STRIDE = CHUNK + 8;  // shorts a staged pixel (80 B: 16-B aligned, banks spread)

template <int NF, int RW>
__global__ __launch_bounds__(256) void k_conv(ConvArgs a, uint32_t tiles_x, uint32_t tiles_y) {
    constexpr int TH = 4 * RW, MT = 2 * RW;
    constexpr int HWD = TILE_W + 2, HP = HWD * (TH + 2);
    __shared__ __align__(16) uint16_t tile[HP * HALO_STRIDE];
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    uint32_t t = blockIdx.x;
    const uint32_t tx = t % tiles_x;
    t /= tiles_x;
    const uint32_t ty = t % tiles_y, b = t / tiles_y;
    const int x0 = (int)tx * TILE_W, y0 = (int)ty * TH;
    const int H = (int)a.H, W = (int)a.W;
    const uint64_t img = (uint64_t)b * a.H * a.W;
    const uint32_t cin = a.cin;
    f32x4 acc[NF][MT];
#pragma unroll
    for (int f = 0; f < NF; ++f)
#pragma unroll
        for (int m = 0; m < MT; ++m) acc[f][m] = f32x4{0.f, 0.f, 0.f, 0.f};
    const uint32_t g = lane >> 4, col = lane & 15u;
    // Staging: a thread's share of a chunk's halo tile is loaded into
    // registers (every load issued before the first LDS store, so their
    // latencies overlap), and the next chunk's share is loaded while this
    // chunk's MFMAs run.
    constexpr int MAXU = (HP * (CHUNK / 8) + 255) / 256;
    uint4 v[MAXU];
    auto load = [&](uint32_t c0) {
        const uint32_t parts = min((uint32_t)CHUNK, cin - c0) / 8, lp = parts == 4 ? 2u : 1u;  // parts: 4 or 2
#pragma unroll
        for (int k = 0; k < MAXU; ++k) {
            const uint32_t u = threadIdx.x + 256u * k;
            v[k] = make_uint4(0u, 0u, 0u, 0u);
            if (u < HP * parts) {
                const uint32_t pix = u >> lp, part = u & (parts - 1u);
                const int gy = y0 + (int)(pix / HWD) - 1, gx = x0 + (int)(pix % HWD) - 1;
                if (gy >= 0 && gy < H && gx >= 0 && gx < W) v[k] = conv_input(a, b, gy, gx, c0 + 8 * part);
            }
        }
    };
    load(0);
    for (uint32_t c0 = 0; c0 < cin; c0 += CHUNK) {
        // A last chunk of 16 channels runs the same instruction with the K
        // slots of lane groups 2 and 3 zero in both operands.
        const uint32_t cc = min((uint32_t)CHUNK, cin - c0), parts = cc / 8;
        const uint32_t lp = parts == 4 ? 2u : 1u;
        const bool live = g < parts;
        __syncthreads();  // the previous chunk's reads are done
#pragma unroll
        for (int k = 0; k < MAXU; ++k) {
            const uint32_t u = threadIdx.x + 256u * k;
            if (u < HP * parts) *(uint4*)(tile + (u >> lp) * HALO_STRIDE + 8 * (u & (parts - 1u))) = v[k];
        }
        __syncthreads();
        if (c0 + CHUNK < cin) load(c0 + CHUNK);
        // A row of three taps at a time: its 3 NF weight fragments are loaded
        // together (one wait on the caches a row, not one a tap).
#pragma unroll 1
        for (int dy = 0; dy < 3; ++dy) {
            uint4 wa[3][NF];
#pragma unroll
            for (int dx = 0; dx < 3; ++dx) {
                const uint16_t* wt = a.w + ((uint64_t)(3 * dy + dx) * a.cout + col) * cin + c0 + 8 * g;
#pragma unroll
                for (int f = 0; f < NF; ++f)
                    wa[dx][f] = live ? *(const uint4*)(wt + (uint64_t)(16 * f) * cin) : make_uint4(0u, 0u, 0u, 0u);
            }
#pragma unroll
            for (int dx = 0; dx < 3; ++dx) {
                uint4 bx[MT];
#pragma unroll
                for (int m = 0; m < MT; ++m) {
                    const int r = (int)wave * RW + m / 2, c = (m & 1) * 16 + (int)col;
                    bx[m] = live ? *(const uint4*)(tile + ((r + dy) * HWD + c + dx) * HALO_STRIDE + 8 * g)
                                 : make_uint4(0u, 0u, 0u, 0u);
                }
#pragma unroll
                for (int f = 0; f < NF; ++f)
#pragma unroll
                    for (int m = 0; m < MT; ++m) acc[f][m] = mfma32(wa[dx][f], bx[m], acc[f][m]);
            }
        }
    }
    // Lane holds D[co = 16 f + 4 g + j][pixel = fragment m's column lane % 16].
#pragma unroll
    for (int m = 0; m < MT; ++m) {
        const int y = y0 + (int)wave * RW + m / 2, x = x0 + (m & 1) * 16 + (int)col;
        if (y >= H || x >= W) continue;
        const uint64_t p = img + (uint64_t)y * a.W + x;
#pragma unroll
        for (int f = 0; f < NF; ++f) {
            const uint32_t co = 16u * f + 4u * g;
            float v[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                float s = acc[f][m][j];
                if (a.bias) s += a.bias[co + j];
                if (a.relu) s = fmaxf(s, 0.0f);
                v[j] = s;
            }
            if (a.mask) {
                bf16x4 mk = *(const bf16x4*)(a.mask + p * a.cout + co);
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    if (!(from_bf16s(mk[j]) > 0.0f)) v[j] = 0.0f;
            }
            if (a.y16) {
                uint2 o = make_uint2(to_bf16(v[0]) | (uint32_t)to_bf16(v[1]) << 16,
                                     to_bf16(v[2]) | (uint32_t)to_bf16(v[3]) << 16);
                *(uint2*)(a.y16 + p * a.cout + co) = o;
            }
            if (a.y32) *(float4*)(a.y32 + p * a.cout + co) = make_float4(v[0], v[1], v[2], v[3]);
        }
    }
}

// ---- weight gradient ------------------------------------------------------------
// Grid (chunks, 9 taps). Block b sums pixels [b * chunk, (b + 1) * chunk) for one tap.
template <int NCO, int NCI>
__global__ __launch_bounds__(256) void k_wgrad(const uint16_t* __restrict__ x, const uint16_t* __restrict__ dz,
                                               uint32_t B, uint32_t H, uint32_t W, uint64_t chunk,
                                               float* __restrict__ part, float* __restrict__ bpart) {
    constexpr int COUT = NCO * 16, CIN = NCI * 16;
    constexpr int TP = 64, LD = TP + 4;
    constexpr int NFR = NCO * NCI, FW = (NFR + 3) / 4;
    __shared__ __align__(16) uint16_t sdz[COUT * LD];
    __shared__ __align__(16) uint16_t sx[CIN * LD];
    const uint32_t tap = blockIdx.y;
    const int dy = (int)tap / 3 - 1, dx = (int)tap % 3 - 1;
    const uint64_t HW = (uint64_t)H * W, P = B * HW;
    const uint64_t lo = (uint64_t)blockIdx.x * chunk, hi = min(P, lo + chunk);
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    f32x4 acc[FW];
#pragma unroll
    for (int i = 0; i < FW; ++i) acc[i] = f32x4{0.f, 0.f, 0.f, 0.f};
    float bacc = 0.0f;
    // A thread's share of a tile: NCO quads of dz and NCI quads of x (TP
    // pixels x COUT / 4 quads = 256 NCO). The next tile's share is loaded
    // into registers while this one's MFMAs run.
    bf16x4 vdz[NCO], vx[NCI];
    auto load = [&](uint64_t t0) {
        // The tile's first pixel (the same for every thread); pixel t0 + j
        // lies j columns on, wrapping into the next rows (and images).
        const uint64_t b0 = t0 / HW, r0 = t0 % HW;
        const uint32_t y0 = (uint32_t)(r0 / W), x0 = (uint32_t)(r0 % W);
#pragma unroll
        for (int k = 0; k < NCO; ++k) {
            const uint32_t u = threadIdx.x + 256u * k, j = u / (COUT / 4), q = u % (COUT / 4);
            const uint64_t p = t0 + j;
            vdz[k] = p < hi ? *(const bf16x4*)(dz + p * COUT + 4 * q) : bf16x4{0, 0, 0, 0};
        }
#pragma unroll
        for (int k = 0; k < NCI; ++k) {
            const uint32_t u = threadIdx.x + 256u * k, j = u / (CIN / 4), q = u % (CIN / 4);
            const uint64_t p = t0 + j;
            bf16x4 v = bf16x4{0, 0, 0, 0};
            if (p < hi) {
                uint64_t b = b0;
                uint32_t cy = y0, cx = x0 + j;
                while (cx >= W) {
                    cx -= W;
                    if (++cy == H) cy = 0, ++b;
                }
                const int yy = (int)cy + dy, xx = (int)cx + dx;
                if (yy >= 0 && yy < (int)H && xx >= 0 && xx < (int)W)
                    v = *(const bf16x4*)(x + (b * HW + (uint64_t)yy * W + xx) * CIN + 4 * q);
            }
            vx[k] = v;
        }
    };
    if (lo < hi) load(lo);
    for (uint64_t t0 = lo; t0 < hi; t0 += TP) {
#pragma unroll
        for (int k = 0; k < NCO; ++k) {
            const uint32_t u = threadIdx.x + 256u * k, j = u / (COUT / 4), q = u % (COUT / 4);
#pragma unroll
            for (int e = 0; e < 4; ++e) sdz[(4 * q + e) * LD + j] = (uint16_t)vdz[k][e];
        }
#pragma unroll
        for (int k = 0; k < NCI; ++k) {
            const uint32_t u = threadIdx.x + 256u * k, j = u / (CIN / 4), q = u % (CIN / 4);
#pragma unroll
            for (int e = 0; e < 4; ++e) sx[(4 * q + e) * LD + j] = (uint16_t)vx[k][e];
        }
        __syncthreads();
        if (t0 + TP < hi) load(t0 + TP);
#pragma unroll
        for (int i = 0; i < FW; ++i) {
            const int fr = (int)wave + 4 * i;
            if (fr < NFR) {
                const int fco = fr / NCI, fci = fr % NCI;
                const uint16_t* ra = sdz + (16 * fco + (lane & 15u)) * LD + 4 * (lane >> 4);
                const uint16_t* rb = sx + (16 * fci + (lane & 15u)) * LD + 4 * (lane >> 4);
#pragma unroll
                for (int kk = 0; kk < TP / 16; ++kk)
                    acc[i] = mfma16(*(const bf16x4*)(ra + 16 * kk), *(const bf16x4*)(rb + 16 * kk), acc[i]);
            }
        }
        if (tap == 4u && threadIdx.x < (uint32_t)COUT)
            for (int j = 0; j < TP; ++j) bacc += from_bf16(sdz[threadIdx.x * LD + j]);
        __syncthreads();
    }
    float* dst = part + ((uint64_t)blockIdx.x * 9 + tap) * (COUT * CIN);
#pragma unroll
    for (int i = 0; i < FW; ++i) {
        const int fr = (int)wave + 4 * i;
        if (fr < NFR) {
            const int fco = fr / NCI, fci = fr % NCI;
#pragma unroll
            for (int j = 0; j < 4; ++j)
                dst[(16 * fco + 4 * (lane >> 4) + j) * CIN + 16 * fci + (lane & 15u)] = acc[i][j];
        }
    }
    if (tap == 4u && threadIdx.x < (uint32_t)COUT) bpart[(uint64_t)blockIdx.x * COUT + threadIdx.x] = bacc;
}

// Sum the chunk partials in chunk order: grad_w[9][cout][cin], then grad_b[cout].
__global__ __launch_bounds__(256) void k_wreduce(const float* __restrict__ part, const float* __restrict__ bpart,
                                                 uint32_t nchunk, uint32_t nw, uint32_t cout, float* __restrict__ gw,
                                                 float* __restrict__ gb) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i < nw) {
        float s = 0.0f;
        for (uint32_t c = 0; c < nchunk; ++c) s += part[(uint64_t)c * nw + i];
        gw[i] = s;
    } else if (i < nw + cout) {
        const uint32_t co = i - nw;
        float s = 0.0f;
        for (uint32_t c = 0; c < nchunk; ++c) s += bpart[(uint64_t)c * cout + co];
        gb[co] = s;
    }
}

// ---- pooling, upsampling, concatenation -----------------------------------------
// Average of each 2x2 block: [B][2h][2w][C] -> [B][h][w][C]; one thread per 4 channels.
__global__ __launch_bounds__(256) void k_pool(const uint16_t* __restrict__ in, uint16_t* __restrict__ out, uint32_t B,
                                              uint32_t h, uint32_t w, uint32_t C) {
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    const uint32_t Q = C / 4;
    if (i >= (uint64_t)B * h * w * Q) return;
    const uint32_t q = (uint32_t)(i % Q);
    const uint64_t pix = i / Q;
    const uint32_t x = (uint32_t)(pix % w);
    const uint64_t r = pix / w;
    const uint32_t y = (uint32_t)(r % h), b = (uint32_t)(r / h);
    float s[4] = {0.f, 0.f, 0.f, 0.f};
    for (int dy = 0; dy < 2; ++dy)
        for (int dx = 0; dx < 2; ++dx) {
            bf16x4 v = *(const bf16x4*)(in + ((((uint64_t)b * 2 * h + 2 * y + dy) * 2 * w) + 2 * x + dx) * C + 4 * q);
            for (int e = 0; e < 4; ++e) s[e] += from_bf16s(v[e]);
        }
    *(uint2*)(out + pix * C + 4 * q) = make_uint2(to_bf16(s[0] * 0.25f) | (uint32_t)to_bf16(s[1] * 0.25f) << 16,
                                                  to_bf16(s[2] * 0.25f) | (uint32_t)to_bf16(s[3] * 0.25f) << 16);
}

// out[B][H][W][C1 + C2] = (coarse[y/2][x/2][C1], skip[y][x][C2]); one thread per 4 channels.
__global__ __launch_bounds__(256) void k_upcat(const uint16_t* __restrict__ coarse, const uint16_t* __restrict__ skip,
                                               uint16_t* __restrict__ out, uint32_t B, uint32_t H, uint32_t W,
                                               uint32_t C1, uint32_t C2) {
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    const uint32_t C = C1 + C2, Q = C / 4;
    if (i >= (uint64_t)B * H * W * Q) return;
    const uint32_t q = (uint32_t)(i % Q), c = 4 * q;
    const uint64_t pix = i / Q;
    const uint32_t x = (uint32_t)(pix % W);
    const uint64_t r = pix / W;
    const uint32_t y = (uint32_t)(r % H), b = (uint32_t)(r / H);
    uint2 v;
    if (c < C1)
        v = *(const uint2*)(coarse + (((uint64_t)b * (H / 2) + y / 2) * (W / 2) + x / 2) * C1 + c);
    else
        v = *(const uint2*)(skip + pix * C2 + (c - C1));
    *(uint2*)(out + pix * C + c) = v;
}

// Backward of k_upcat. Coarse part: dZ[b][y][x][c] = sum of the 2x2 children's
// gradient, zeroed where the coarse activation is not positive (its ReLU).
__global__ __launch_bounds__(256) void k_upcat_bwd_coarse(const float* __restrict__ g, const uint16_t* __restrict__ act,
                                                          uint16_t* __restrict__ dz, uint32_t B, uint32_t h, uint32_t w,
                                                          uint32_t C1, uint32_t C) {
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    const uint32_t Q = C1 / 4;
    if (i >= (uint64_t)B * h * w * Q) return;
    const uint32_t q = (uint32_t)(i % Q), c = 4 * q;
    const uint64_t pix = i / Q;
    const uint32_t x = (uint32_t)(pix % w);
    const uint64_t r = pix / w;
    const uint32_t y = (uint32_t)(r % h), b = (uint32_t)(r / h);
    float s[4] = {0.f, 0.f, 0.f, 0.f};
    for (int dy = 0; dy < 2; ++dy)
        for (int dx = 0; dx < 2; ++dx) {
            float4 v = *(const float4*)(g + ((((uint64_t)b * 2 * h + 2 * y + dy) * 2 * w) + 2 * x + dx) * C + c);
            s[0] += v.x, s[1] += v.y, s[2] += v.z, s[3] += v.w;
        }
    bf16x4 m = *(const bf16x4*)(act + pix * C1 + c);
    for (int e = 0; e < 4; ++e)
        if (!(from_bf16s(m[e]) > 0.0f)) s[e] = 0.0f;
    *(uint2*)(dz + pix * C1 + c) =
        make_uint2(to_bf16(s[0]) | (uint32_t)to_bf16(s[1]) << 16, to_bf16(s[2]) | (uint32_t)to_bf16(s[3]) << 16);
}

// Skip part of k_upcat's backward: gskip[p][c] = g[p][C1 + c].
__global__ __launch_bounds__(256) void k_upcat_bwd_skip(const float* __restrict__ g, float* __restrict__ gs, uint64_t P,
                                                        uint32_t C1, uint32_t C2) {
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    const uint32_t Q = C2 / 4;
    if (i >= P * Q) return;
    const uint32_t c = 4 * (uint32_t)(i % Q);
    const uint64_t pix = i / Q;
    *(float4*)(gs + pix * C2 + c) = *(const float4*)(g + pix * (C1 + C2) + C1 + c);
}

// Backward of k_pool joined with the skip gradient:
// dZ[p][c] = (gskip[p][c] + gq[parent][c] / 4) where act > 0, else 0.
__global__ __launch_bounds__(256) void k_pool_bwd(const float* __restrict__ gq, const float* __restrict__ gs,
                                                  const uint16_t* __restrict__ act, uint16_t* __restrict__ dz, uint32_t B,
                                                  uint32_t H, uint32_t W, uint32_t C) {
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    const uint32_t Q = C / 4;
    if (i >= (uint64_t)B * H * W * Q) return;
    const uint32_t c = 4 * (uint32_t)(i % Q);
    const uint64_t pix = i / Q;
    const uint32_t x = (uint32_t)(pix % W);
    const uint64_t r = pix / W;
    const uint32_t y = (uint32_t)(r % H), b = (uint32_t)(r / H);
    float4 a = *(const float4*)(gs + pix * C + c);
    float4 up = *(const float4*)(gq + (((uint64_t)b * (H / 2) + y / 2) * (W / 2) + x / 2) * C + c);
    float s[4] = {a.x + up.x * 0.25f, a.y + up.y * 0.25f, a.z + up.z * 0.25f, a.w + up.w * 0.25f};
    bf16x4 m = *(const bf16x4*)(act + pix * C + c);
    for (int e = 0; e < 4; ++e)
        if (!(from_bf16s(m[e]) > 0.0f)) s[e] = 0.0f;
    *(uint2*)(dz + pix * C + c) =
        make_uint2(to_bf16(s[0]) | (uint32_t)to_bf16(s[1]) << 16, to_bf16(s[2]) | (uint32_t)to_bf16(s[3]) << 16);
}

// ---- filter application, loss and its gradient ------------------------------------
__device__ __forceinline__ void demod_at(const float* __restrict__ fr, uint64_t base, uint32_t H, uint32_t W, int y, int x,
                                         float d[3]) {
    y = min(max(y, 0), (int)H - 1);
    x = min(max(x, 0), (int)W - 1);
    const float* f = fr + (base + (uint64_t)y * W + x) * FRAME;
    for (int c = 0; c < 3; ++c) d[c] = f[c] / fmaxf(f[3 + c], 0.05f);
}

// One thread per padded pixel. Writes the filtered radiance where the pixel is
// in the image; with a target, the per-block loss sum and the logits' gradient
// (zero on the padding and the unused logit channels).
__global__ __launch_bounds__(256) void k_apply(const float* __restrict__ lg, const float* __restrict__ fr, uint32_t B,
                                               uint32_t H, uint32_t W, uint32_t Hp, uint32_t Wp, float* __restrict__ out,
                                               const float* __restrict__ tgt, uint16_t* __restrict__ dlg,
                                               float* __restrict__ loss_part, float inv_count) {
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    const uint64_t np = (uint64_t)B * Hp * Wp;
    float l = 0.0f;
    if (i < np) {
        const uint32_t x = (uint32_t)(i % Wp);
        const uint64_t r = i / Wp;
        const uint32_t y = (uint32_t)(r % Hp), b = (uint32_t)(r / Hp);
        const bool in = x < W && y < H;
        float g[LOGITS];
        for (int k = 0; k < (int)LOGITS; ++k) g[k] = 0.0f;
        if (in) {
            float w[TAPS];
            const float4* lp = (const float4*)(lg + i * LOGITS);
            float m = -3.0e38f;
            for (int q = 0; q < 7; ++q) {
                float4 v = lp[q];
                float vv[4] = {v.x, v.y, v.z, v.w};
                for (int e = 0; e < 4; ++e)
                    if (4 * q + e < (int)TAPS) w[4 * q + e] = vv[e];
            }
            for (int k = 0; k < (int)TAPS; ++k) m = fmaxf(m, w[k]);
            float s = 0.0f;
            for (int k = 0; k < (int)TAPS; ++k) {
                w[k] = __expf(w[k] - m);
                s += w[k];
            }
            const float inv = 1.0f / s;
            for (int k = 0; k < (int)TAPS; ++k) w[k] *= inv;
            const uint64_t base = (uint64_t)b * H * W;
            float acc[3] = {0.f, 0.f, 0.f};
            for (int k = 0; k < (int)TAPS; ++k) {
                float d[3];
                demod_at(fr, base, H, W, (int)y + k / 5 - 2, (int)x + k % 5 - 2, d);
                for (int c = 0; c < 3; ++c) acc[c] += w[k] * d[c];
            }
            const uint64_t o = base + (uint64_t)y * W + x;
            const float* f = fr + o * FRAME;
            float yv[3], ad[3];
            for (int c = 0; c < 3; ++c) {
                ad[c] = fmaxf(f[3 + c], 0.05f);
                yv[c] = ad[c] * acc[c];
            }
            if (out) {
                out[3 * o] = yv[0];
                out[3 * o + 1] = yv[1];
                out[3 * o + 2] = yv[2];
            }
            if (tgt) {
                float ds[3];
                for (int c = 0; c < 3; ++c) {
                    const float t = tgt[3 * o + c], den = t * t + 0.01f, e = yv[c] - t;
                    l += e * e / den;
                    ds[c] = 2.0f * e / den * inv_count * ad[c];
                }
                float gbar = 0.0f;
                for (int k = 0; k < (int)TAPS; ++k) {
                    float d[3];
                    demod_at(fr, base, H, W, (int)y + k / 5 - 2, (int)x + k % 5 - 2, d);
                    g[k] = ds[0] * d[0] + ds[1] * d[1] + ds[2] * d[2];
                    gbar += w[k] * g[k];
                }
                for (int k = 0; k < (int)TAPS; ++k) g[k] = w[k] * (g[k] - gbar);
            }
        }
        if (dlg) {
            uint4* d = (uint4*)(dlg + i * LOGITS);
            for (int q = 0; q < (int)LOGITS / 8; ++q)
                d[q] = make_uint4(to_bf16(g[8 * q]) | (uint32_t)to_bf16(g[8 * q + 1]) << 16,
                                  to_bf16(g[8 * q + 2]) | (uint32_t)to_bf16(g[8 * q + 3]) << 16,
                                  to_bf16(g[8 * q + 4]) | (uint32_t)to_bf16(g[8 * q + 5]) << 16,
                                  to_bf16(g[8 * q + 6]) | (uint32_t)to_bf16(g[8 * q + 7]) << 16);
        }
    }
    if (loss_part) {
        __shared__ float red[256];
        red[threadIdx.x] = l;
        __syncthreads();
        for (int s = 128; s > 0; s >>= 1) {
            if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
            __syncthreads();
        }
        if (threadIdx.x == 0) loss_part[blockIdx.x] = red[0];
    }
}

// ---- optimiser -----------------------------------------------------------------
__global__ __launch_bounds__(256) void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                                              float* __restrict__ v, uint64_t n, float lr, float b1, float b2, float eps,
                                              float bc1, float bc2) {
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
}

// BF16 copies of one layer: wf[tap][co][ci] and wb[8 - tap][ci][co].
__global__ __launch_bounds__(256) void k_pack_w(const float* __restrict__ w, uint16_t* __restrict__ wf,
                                                uint16_t* __restrict__ wb, uint32_t cout, uint32_t cin) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= 9 * cout * cin) return;
    const uint32_t ci = i % cin, co = (i / cin) % cout, tap = i / (cin * cout);
    const uint16_t h = to_bf16(w[i]);
    wf[i] = h;
    wb[((8 - tap) * cin + ci) * cout + co] = h;
}

inline uint32_t grid_of(uint64_t n) { return (uint32_t)((n + 255) / 256); }

}  // namespace

// ---- host object ----------------------------------------------------------------
struct phr_kpcn {
    phr_gpu* gpu = nullptr;
    uint64_t nparams = 0, nw16 = 0;
    uint64_t off_w[NLAYERS], off_b[NLAYERS], off_16[NLAYERS];
    float *master = nullptr, *m = nullptr, *v = nullptr, *grad = nullptr;
    uint16_t *wf = nullptr, *wb = nullptr;
    uint64_t t = 0;
    // workspace for one (B, H, W, train) shape
    uint32_t B = 0, H = 0, W = 0, Hp = 0, Wp = 0;
    bool train = false;
    char* ws = nullptr;
    uint64_t ws_bytes = 0;
    uint16_t *F = nullptr, *act[NLAYERS] = {}, *Q0 = nullptr, *Q1 = nullptr, *Q2 = nullptr, *U2 = nullptr, *U1 = nullptr,
             *U0 = nullptr;
    float* logits = nullptr;
    uint16_t* dz[NLAYERS] = {};
    float *G = nullptr, *Gs0 = nullptr, *Gs1 = nullptr, *Gs2 = nullptr, *part = nullptr, *bpart = nullptr,
          *loss_part = nullptr;
    uint64_t part_n = 0, bpart_n = 0, loss_n = 0;
    std::vector<float> loss_host;
    char* batch = nullptr;  // gathered frames and targets of an indexed step
    uint64_t batch_bytes = 0;
    uint8_t rw[NLAYERS] = {};  // forward conv tile per layer (rows a wave owns; 0 = default), set by autotune
};

namespace {

uint64_t level_pixels(const phr_kpcn* k, uint32_t l) { return (uint64_t)k->B * (k->Hp >> l) * (k->Wp >> l); }

// Chunk length (a multiple of the 64-pixel tile) giving about 224 chunks per tap.
uint64_t wgrad_chunk(uint64_t P) {
    uint64_t c = (P + 223) / 224;
    c = (c + 63) / 64 * 64;
    return c < 1024 ? 1024 : c;
}

void free_ws(phr_kpcn* k) {
    if (k->ws) (void)hipFree(k->ws);
    k->ws = nullptr;
    k->ws_bytes = 0;
    k->B = k->H = k->W = 0;
    k->Q0 = k->Q1 = k->Q2 = k->U2 = k->U1 = k->U0 = nullptr;
    for (int i = 0; i < NLAYERS; ++i) k->dz[i] = nullptr;
    k->G = k->Gs0 = k->Gs1 = k->Gs2 = k->part = k->bpart = nullptr;
}

int ensure_ws(phr_kpcn* k, uint32_t B, uint32_t H, uint32_t W, bool train) {
    if (k->ws && k->B == B && k->H == H && k->W == W && (k->train || !train)) return PHR_OK;
    free_ws(k);
    const uint32_t Hp = (H + 7) / 8 * 8, Wp = (W + 7) / 8 * 8;
    k->B = B, k->H = H, k->W = W, k->Hp = Hp, k->Wp = Wp, k->train = train;
    uint64_t off = 0;
    struct Want {
        void** slot;
        uint64_t bytes;
    };
    std::vector<Want> want;
    auto add = [&](void** slot, uint64_t bytes) { want.push_back({slot, (bytes + 255) / 256 * 256}); };
    auto P = [&](uint32_t l) { return level_pixels(k, l); };
    add((void**)&k->F, P(0) * FEAT * 2);
    for (int i = 0; i < NLAYERS - 1; ++i) add((void**)&k->act[i], P(LAYERS[i].level) * LAYERS[i].cout * 2);
    add((void**)&k->logits, P(0) * LOGITS * 4);
    k->loss_n = grid_of(P(0));
    add((void**)&k->loss_part, k->loss_n * 4);
    if (train) {  // the pooled and concatenated inputs, whole, for the weight gradients
        add((void**)&k->Q0, P(1) * 32 * 2);
        add((void**)&k->Q1, P(2) * 48 * 2);
        add((void**)&k->Q2, P(3) * 64 * 2);
        add((void**)&k->U2, P(2) * 160 * 2);
        add((void**)&k->U1, P(1) * 112 * 2);
        add((void**)&k->U0, P(0) * 80 * 2);
        for (int i = 0; i < NLAYERS; ++i) add((void**)&k->dz[i], P(LAYERS[i].level) * LAYERS[i].cout * 2);
        uint64_t g = 0;
        for (int i = 1; i < NLAYERS; ++i) g = std::max<uint64_t>(g, P(LAYERS[i].level) * LAYERS[i].cin);
        add((void**)&k->G, g * 4);
        add((void**)&k->Gs0, P(0) * 32 * 4);
        add((void**)&k->Gs1, P(1) * 48 * 4);
        add((void**)&k->Gs2, P(2) * 64 * 4);
        uint64_t pn = 0, bn = 0;
        for (int i = 0; i < NLAYERS; ++i) {
            const uint64_t Pl = P(LAYERS[i].level), nch = (Pl + wgrad_chunk(Pl) - 1) / wgrad_chunk(Pl);
            pn = std::max<uint64_t>(pn, nch * 9 * LAYERS[i].cout * LAYERS[i].cin);
            bn = std::max<uint64_t>(bn, nch * LAYERS[i].cout);
        }
        k->part_n = pn, k->bpart_n = bn;
        add((void**)&k->part, pn * 4);
        add((void**)&k->bpart, bn * 4);
    }
    for (auto& w : want) off += w.bytes;
    hipError_t e = hipMalloc((void**)&k->ws, off);
    if (e != hipSuccess) {
        k->ws = nullptr;
        free_ws(k);
        phr_set_error("kpcn workspace hipMalloc", e);
        return PHR_ERR_HIP;
    }
    k->ws_bytes = off;
    uint64_t at = 0;
    for (auto& w : want) {
        *w.slot = k->ws + at;
        at += w.bytes;
    }
    k->act[OUT] = nullptr;  // the output layer writes FP32 logits
    k->loss_host.resize(k->loss_n);
    return PHR_OK;
}

template <int NF, int RW>
void launch_conv_t(const ConvArgs& a, hipStream_t s) {
    const uint32_t tx = (a.W + TILE_W - 1) / TILE_W, ty = (a.H + 4 * RW - 1) / (4 * RW);
    hipLaunchKernelGGL((k_conv<NF, RW>), dim3(tx * ty * a.B), dim3(256), 0, s, a, tx, ty);
}

// rw: rows a wave owns (1 or 2; 0 = 2 up to 96 output channels, else 1).
int launch_conv(const ConvArgs& a, hipStream_t s, int rw = 0) {
    if (a.cin % 16 || a.cout % 16) {
        phr_set_error_msg("kpcn conv: channels must be multiples of 16");
        return PHR_ERR_ARG;
    }
    const uint32_t nf = a.cout / 16;
    if (rw == 0) rw = nf <= 6 ? 2 : 1;
    if (nf == 10) rw = 1;  // 40 accumulators a row: two rows would leave one wave a SIMD
    switch (nf * 4 + (uint32_t)rw) {
        case 2 * 4 + 1: launch_conv_t<2, 1>(a, s); break;
        case 2 * 4 + 2: launch_conv_t<2, 2>(a, s); break;
        case 3 * 4 + 1: launch_conv_t<3, 1>(a, s); break;
        case 3 * 4 + 2: launch_conv_t<3, 2>(a, s); break;
        case 4 * 4 + 1: launch_conv_t<4, 1>(a, s); break;
        case 4 * 4 + 2: launch_conv_t<4, 2>(a, s); break;
        case 5 * 4 + 1: launch_conv_t<5, 1>(a, s); break;
        case 5 * 4 + 2: launch_conv_t<5, 2>(a, s); break;
        case 6 * 4 + 1: launch_conv_t<6, 1>(a, s); break;
        case 6 * 4 + 2: launch_conv_t<6, 2>(a, s); break;
        case 7 * 4 + 1: launch_conv_t<7, 1>(a, s); break;
        case 7 * 4 + 2: launch_conv_t<7, 2>(a, s); break;
        case 10 * 4 + 1: launch_conv_t<10, 1>(a, s); break;
        default: phr_set_error_msg("kpcn conv: unsupported output channel count"); return PHR_ERR_ARG;
    }
    PHR_HIP(hipGetLastError());
    return PHR_OK;
}

// Forward conv of layer `li` at its level; `pooled` / (`xc`, `c1`) select the
// fused input forms (ConvArgs).
int conv_fwd(phr_kpcn* k, int li, const uint16_t* x, uint16_t* y16, float* y32, uint32_t pooled = 0,
             const uint16_t* xc = nullptr, uint32_t c1 = 0) {
    const LayerDef& L = LAYERS[li];
    ConvArgs a{};
    a.x = x;
    a.w = k->wf + k->off_16[li];
    a.bias = k->master + k->off_b[li];
    a.mask = nullptr;
    a.y16 = y16;
    a.y32 = y32;
    a.B = k->B, a.H = k->Hp >> L.level, a.W = k->Wp >> L.level;
    a.cin = L.cin, a.cout = L.cout, a.relu = L.relu;
    a.pooled = pooled, a.xc = xc, a.c1 = c1;
    return launch_conv(a, k->gpu->stream, k->rw[li]);
}

// Input gradient of layer `li`: conv of its output gradient with the flipped,
// transposed weights; zeroed where `mask` (the layer's input activation) is not
// positive when given.
int conv_bwd_data(phr_kpcn* k, int li, const uint16_t* dzi, const uint16_t* mask, uint16_t* y16, float* y32) {
    const LayerDef& L = LAYERS[li];
    ConvArgs a{};
    a.x = dzi;
    a.w = k->wb + k->off_16[li];
    a.bias = nullptr;
    a.mask = mask;
    a.y16 = y16;
    a.y32 = y32;
    a.B = k->B, a.H = k->Hp >> L.level, a.W = k->Wp >> L.level;
    a.cin = L.cout, a.cout = L.cin, a.relu = 0;
    return launch_conv(a, k->gpu->stream);
}

template <int NCO, int NCI>
void launch_wgrad_t(phr_kpcn* k, const uint16_t* x, const uint16_t* dzi, uint32_t level, uint64_t nch, uint64_t chunk) {
    hipLaunchKernelGGL((k_wgrad<NCO, NCI>), dim3((uint32_t)nch, 9), dim3(256), 0, k->gpu->stream, x, dzi, k->B,
                       k->Hp >> level, k->Wp >> level, chunk, k->part, k->bpart);
}

int wgrad(phr_kpcn* k, int li, const uint16_t* x) {
    const LayerDef& L = LAYERS[li];
    const uint64_t P = level_pixels(k, L.level), chunk = wgrad_chunk(P), nch = (P + chunk - 1) / chunk;
    switch (li) {
        case E0A: launch_wgrad_t<2, 1>(k, x, k->dz[li], L.level, nch, chunk); break;
        case E0B: launch_wgrad_t<2, 2>(k, x, k->dz[li], L.level, nch, chunk); break;
        case E1A: launch_wgrad_t<3, 2>(k, x, k->dz[li], L.level, nch, chunk); break;
        case E1B: launch_wgrad_t<3, 3>(k, x, k->dz[li], L.level, nch, chunk); break;
        case E2A: launch_wgrad_t<4, 3>(k, x, k->dz[li], L.level, nch, chunk); break;
        case E2B: launch_wgrad_t<4, 4>(k, x, k->dz[li], L.level, nch, chunk); break;
        case E3: launch_wgrad_t<6, 4>(k, x, k->dz[li], L.level, nch, chunk); break;
        case D2: launch_wgrad_t<4, 10>(k, x, k->dz[li], L.level, nch, chunk); break;
        case D1: launch_wgrad_t<3, 7>(k, x, k->dz[li], L.level, nch, chunk); break;
        case D0: launch_wgrad_t<2, 5>(k, x, k->dz[li], L.level, nch, chunk); break;
        case OUT: launch_wgrad_t<2, 2>(k, x, k->dz[li], L.level, nch, chunk); break;
    }
    PHR_HIP(hipGetLastError());
    const uint32_t nw = 9 * L.cout * L.cin;
    hipLaunchKernelGGL(k_wreduce, dim3(grid_of(nw + L.cout)), dim3(256), 0, k->gpu->stream, k->part, k->bpart,
                       (uint32_t)nch, nw, L.cout, k->grad + k->off_w[li], k->grad + k->off_b[li]);
    PHR_HIP(hipGetLastError());
    return PHR_OK;
}

int pool(phr_kpcn* k, const uint16_t* in, uint16_t* out, uint32_t level_out, uint32_t C) {
    const uint32_t h = k->Hp >> level_out, w = k->Wp >> level_out;
    hipLaunchKernelGGL(k_pool, dim3(grid_of((uint64_t)k->B * h * w * (C / 4))), dim3(256), 0, k->gpu->stream, in, out,
                       k->B, h, w, C);
    PHR_HIP(hipGetLastError());
    return PHR_OK;
}

int upcat(phr_kpcn* k, const uint16_t* coarse, const uint16_t* skip, uint16_t* out, uint32_t level, uint32_t C1,
          uint32_t C2) {
    const uint32_t H = k->Hp >> level, W = k->Wp >> level;
    hipLaunchKernelGGL(k_upcat, dim3(grid_of((uint64_t)k->B * H * W * ((C1 + C2) / 4))), dim3(256), 0, k->gpu->stream,
                       coarse, skip, out, k->B, H, W, C1, C2);
    PHR_HIP(hipGetLastError());
    return PHR_OK;
}

#define KP_TRY(x)                     \
    do {                              \
        int rc_ = (x);                \
        if (rc_ != PHR_OK) return rc_; \
    } while (0)

// Layer li's forward conv with its input as the network reads it: pooled for
// the first conv of levels 1-3, the coarser output upsampled beside the skip
// for the decoders.
int conv_layer(phr_kpcn* k, int li) {
    uint16_t** A = k->act;
    switch (li) {
        case E0A: return conv_fwd(k, E0A, k->F, A[E0A], nullptr);
        case E0B: return conv_fwd(k, E0B, A[E0A], A[E0B], nullptr);
        case E1A: return conv_fwd(k, E1A, A[E0B], A[E1A], nullptr, 1);
        case E1B: return conv_fwd(k, E1B, A[E1A], A[E1B], nullptr);
        case E2A: return conv_fwd(k, E2A, A[E1B], A[E2A], nullptr, 1);
        case E2B: return conv_fwd(k, E2B, A[E2A], A[E2B], nullptr);
        case E3: return conv_fwd(k, E3, A[E2B], A[E3], nullptr, 1);
        case D2: return conv_fwd(k, D2, A[E2B], A[D2], nullptr, 0, A[E3], 96);
        case D1: return conv_fwd(k, D1, A[E1B], A[D1], nullptr, 0, A[D2], 64);
        case D0: return conv_fwd(k, D0, A[E0B], A[D0], nullptr, 0, A[D1], 48);
        default: return conv_fwd(k, OUT, A[D0], nullptr, k->logits);
    }
}

// Forward to the logits. The pooled and concatenated inputs are read by the
// convolutions themselves; training also keeps them whole for the weight
// gradients.
int forward(phr_kpcn* k, const float* frames, bool keep) {
    hipStream_t s = k->gpu->stream;
    hipLaunchKernelGGL(k_features, dim3(grid_of(level_pixels(k, 0))), dim3(256), 0, s, frames, k->B, k->H, k->W, k->Hp,
                       k->Wp, k->F);
    PHR_HIP(hipGetLastError());
    uint16_t** A = k->act;
    for (int li = 0; li < NLAYERS; ++li) KP_TRY(conv_layer(k, li));
    if (keep) {
        KP_TRY(pool(k, A[E0B], k->Q0, 1, 32));
        KP_TRY(pool(k, A[E1B], k->Q1, 2, 48));
        KP_TRY(pool(k, A[E2B], k->Q2, 3, 64));
        KP_TRY(upcat(k, A[E3], A[E2B], k->U2, 2, 96, 64));
        KP_TRY(upcat(k, A[D2], A[E1B], k->U1, 1, 64, 48));
        KP_TRY(upcat(k, A[D1], A[E0B], k->U0, 0, 48, 32));
    }
    return PHR_OK;
}

int apply(phr_kpcn* k, const float* frames, float* out, const float* tgt, bool grad) {
    const uint64_t np = level_pixels(k, 0);
    const float inv_count = 1.0f / (3.0f * (float)k->B * (float)k->H * (float)k->W);
    hipLaunchKernelGGL(k_apply, dim3(grid_of(np)), dim3(256), 0, k->gpu->stream, k->logits, frames, k->B, k->H, k->W,
                       k->Hp, k->Wp, out, tgt, grad ? k->dz[OUT] : nullptr, tgt ? k->loss_part : nullptr, inv_count);
    PHR_HIP(hipGetLastError());
    return PHR_OK;
}

int read_loss(phr_kpcn* k, float* loss) {
    PHR_HIP(hipMemcpyAsync(k->loss_host.data(), k->loss_part, k->loss_n * 4, hipMemcpyDeviceToHost, k->gpu->stream));
    PHR_HIP(hipStreamSynchronize(k->gpu->stream));
    double s = 0.0;
    for (float v : k->loss_host) s += v;
    *loss = (float)(s / (3.0 * (double)k->B * k->H * k->W));
    return PHR_OK;
}

int upcat_bwd(phr_kpcn* k, uint32_t level, uint32_t C1, uint32_t C2, const uint16_t* act_coarse, uint16_t* dz_coarse,
              float* gskip) {
    const uint32_t h = k->Hp >> (level + 1), w = k->Wp >> (level + 1);
    hipLaunchKernelGGL(k_upcat_bwd_coarse, dim3(grid_of((uint64_t)k->B * h * w * (C1 / 4))), dim3(256), 0,
                       k->gpu->stream, k->G, act_coarse, dz_coarse, k->B, h, w, C1, C1 + C2);
    PHR_HIP(hipGetLastError());
    const uint64_t P = level_pixels(k, level);
    hipLaunchKernelGGL(k_upcat_bwd_skip, dim3(grid_of(P * (C2 / 4))), dim3(256), 0, k->gpu->stream, k->G, gskip, P, C1,
                       C2);
    PHR_HIP(hipGetLastError());
    return PHR_OK;
}

int pool_bwd(phr_kpcn* k, uint32_t level, uint32_t C, const float* gskip, const uint16_t* act, uint16_t* dzo) {
    const uint32_t H = k->Hp >> level, W = k->Wp >> level;
    hipLaunchKernelGGL(k_pool_bwd, dim3(grid_of((uint64_t)k->B * H * W * (C / 4))), dim3(256), 0, k->gpu->stream, k->G,
                       gskip, act, dzo, k->B, H, W, C);
    PHR_HIP(hipGetLastError());
    return PHR_OK;
}

// Backward from the logits' gradient (in dz[OUT]) to every layer's weight gradient.
int backward(phr_kpcn* k) {
    uint16_t** A = k->act;
    uint16_t** Z = k->dz;
    KP_TRY(wgrad(k, OUT, A[D0]));
    KP_TRY(conv_bwd_data(k, OUT, Z[OUT], A[D0], Z[D0], nullptr));
    KP_TRY(wgrad(k, D0, k->U0));
    KP_TRY(conv_bwd_data(k, D0, Z[D0], nullptr, nullptr, k->G));
    KP_TRY(upcat_bwd(k, 0, 48, 32, A[D1], Z[D1], k->Gs0));
    KP_TRY(wgrad(k, D1, k->U1));
    KP_TRY(conv_bwd_data(k, D1, Z[D1], nullptr, nullptr, k->G));
    KP_TRY(upcat_bwd(k, 1, 64, 48, A[D2], Z[D2], k->Gs1));
    KP_TRY(wgrad(k, D2, k->U2));
    KP_TRY(conv_bwd_data(k, D2, Z[D2], nullptr, nullptr, k->G));
    KP_TRY(upcat_bwd(k, 2, 96, 64, A[E3], Z[E3], k->Gs2));
    KP_TRY(wgrad(k, E3, k->Q2));
    KP_TRY(conv_bwd_data(k, E3, Z[E3], nullptr, nullptr, k->G));
    KP_TRY(pool_bwd(k, 2, 64, k->Gs2, A[E2B], Z[E2B]));
    KP_TRY(wgrad(k, E2B, A[E2A]));
    KP_TRY(conv_bwd_data(k, E2B, Z[E2B], A[E2A], Z[E2A], nullptr));
    KP_TRY(wgrad(k, E2A, k->Q1));
    KP_TRY(conv_bwd_data(k, E2A, Z[E2A], nullptr, nullptr, k->G));
    KP_TRY(pool_bwd(k, 1, 48, k->Gs1, A[E1B], Z[E1B]));
    KP_TRY(wgrad(k, E1B, A[E1A]));
    KP_TRY(conv_bwd_data(k, E1B, Z[E1B], A[E1A], Z[E1A], nullptr));
    KP_TRY(wgrad(k, E1A, k->Q0));
    KP_TRY(conv_bwd_data(k, E1A, Z[E1A], nullptr, nullptr, k->G));
    KP_TRY(pool_bwd(k, 0, 32, k->Gs0, A[E0B], Z[E0B]));
    KP_TRY(wgrad(k, E0B, A[E0A]));
    KP_TRY(conv_bwd_data(k, E0B, Z[E0B], A[E0A], Z[E0A], nullptr));
    KP_TRY(wgrad(k, E0A, k->F));
    return PHR_OK;
}

int pack_weights(phr_kpcn* k) {
    for (int i = 0; i < NLAYERS; ++i) {
        const uint32_t n = 9 * LAYERS[i].cout * LAYERS[i].cin;
        hipLaunchKernelGGL(k_pack_w, dim3(grid_of(n)), dim3(256), 0, k->gpu->stream, k->master + k->off_w[i],
                           k->wf + k->off_16[i], k->wb + k->off_16[i], LAYERS[i].cout, LAYERS[i].cin);
    }
    PHR_HIP(hipGetLastError());
    return PHR_OK;
}

int check_shape(phr_array* frames, uint32_t b, uint32_t w, uint32_t h, phr_array* other, const char* what) {
    const uint64_t n = (uint64_t)b * w * h;
    PHR_CHECK(frames && b && w && h, PHR_ERR_ARG, std::string(what) + ": empty batch");
    PHR_CHECK(frames->len * frames->elem >= n * FRAME * 4, PHR_ERR_SIZE,
              std::string(what) + ": frames hold fewer than b*w*h*10 floats");
    if (other)
        PHR_CHECK(other->len * other->elem >= n * 3 * 4, PHR_ERR_SIZE,
                  std::string(what) + ": radiance array holds fewer than b*w*h*3 floats");
    return PHR_OK;
}

// ---- the held-out check -----------------------------------------------------------
// Two independent halves A and B of a frame (the same samples a pixel, other
// streams), D the network's output from A alone. B is unbiased and
// independent of D, and A's noise equals B's in distribution, so per value
//   E[(D - B)^2] - E[(A - B)^2] / 2 = E[(D - T)^2]
// for the converged image T: the sum below estimates D's relMSE against T,
// with D^2 + 0.01 standing in for T^2 + 0.01. The second sum is B's own
// noise on the same scale. Per-block partials in double, summed on the host.
constexpr uint32_t HO_BLOCKS = 512;
__global__ __launch_bounds__(256) void k_held_out(const float* __restrict__ d, const float* __restrict__ fa,
                                                  const float* __restrict__ fb, uint64_t n, double* __restrict__ part) {
    __shared__ double s_e[256], s_v[256];
    double e = 0.0, v = 0.0;
    for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256)
        for (int c = 0; c < 3; ++c) {
            const float y = d[3 * i + c], a = fa[FRAME * i + c], b = fb[FRAME * i + c];
            const float den = y * y + 0.01f;
            const float ab = (a - b) * (a - b) * 0.5f;
            e += (double)(((y - b) * (y - b) - ab) / den);
            v += (double)(ab / den);
        }
    s_e[threadIdx.x] = e;
    s_v[threadIdx.x] = v;
    __syncthreads();
    for (uint32_t o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) {
            s_e[threadIdx.x] += s_e[threadIdx.x + o];
            s_v[threadIdx.x] += s_v[threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = s_e[0];
        part[2 * blockIdx.x + 1] = s_v[0];
    }
}

// The spread of two denoised images of the same view from independent
// halves: mean over values of (a - b)^2 / (a^2 + 0.01), which is twice the
// noise-driven part of either image's relMSE (its bias is common to both).
__global__ __launch_bounds__(256) void k_spread(const float* __restrict__ a, const float* __restrict__ b, uint64_t n,
                                                double* __restrict__ part) {
    __shared__ double s_e[256];
    double e = 0.0;
    for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < 3 * n; i += (uint64_t)gridDim.x * 256) {
        const float x = a[i], y = b[i];
        e += (double)((x - y) * (x - y) / (x * x + 0.01f));
    }
    s_e[threadIdx.x] = e;
    __syncthreads();
    for (uint32_t o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) s_e[threadIdx.x] += s_e[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = s_e[0];
}

// The frame of both halves together (equal sample counts): the mean of each
// channel, depth over the halves that met something (-1 when neither did).
// With channels = 3 only the radiance, 3 floats a pixel.
__global__ __launch_bounds__(256) void k_frame_mean(const float* __restrict__ fa, const float* __restrict__ fb, uint64_t n,
                                                    uint32_t channels, float* __restrict__ out) {
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float* a = fa + FRAME * i;
    const float* b = fb + FRAME * i;
    float* o = out + (uint64_t)channels * i;
    for (uint32_t c = 0; c < min(channels, 9u); ++c) o[c] = 0.5f * (a[c] + b[c]);
    if (channels == FRAME) {
        const float za = a[9], zb = b[9];
        o[9] = za < 0.f ? zb : (zb < 0.f ? za : 0.5f * (za + zb));
    }
}

}  // namespace

// ---- C ABI ----------------------------------------------------------------------------
extern "C" {

// The workspace's address (a recorded graph holds it).
const void* phr_kpcn_internal_workspace(const phr_kpcn* k) { return k ? (const void*)k->ws : nullptr; }

uint32_t phr_kpcn_layers(uint32_t* out) {
    if (out)
        for (int i = 0; i < NLAYERS; ++i) {
            out[4 * i] = LAYERS[i].cin;
            out[4 * i + 1] = LAYERS[i].cout;
            out[4 * i + 2] = LAYERS[i].level;
            out[4 * i + 3] = LAYERS[i].relu;
        }
    return NLAYERS;
}

uint64_t phr_kpcn_param_count(void) {
    uint64_t n = 0;
    for (int i = 0; i < NLAYERS; ++i) n += 9ull * LAYERS[i].cout * LAYERS[i].cin + LAYERS[i].cout;
    return n;
}

int phr_kpcn_new(phr_gpu* g, phr_kpcn** out) {
    PHR_CHECK(g && out, PHR_ERR_ARG, "phr_kpcn_new: null");
    phr_kpcn* k = new phr_kpcn();
    k->gpu = g;
    uint64_t o = 0, o16 = 0;
    for (int i = 0; i < NLAYERS; ++i) {
        k->off_w[i] = o;
        o += 9ull * LAYERS[i].cout * LAYERS[i].cin;
        k->off_b[i] = o;
        o += LAYERS[i].cout;
        k->off_16[i] = o16;
        o16 += 9ull * LAYERS[i].cout * LAYERS[i].cin;
    }
    k->nparams = o, k->nw16 = o16;
    hipError_t e = hipSuccess;
    float** f32s[4] = {&k->master, &k->m, &k->v, &k->grad};
    for (float** p : f32s)
        if (e == hipSuccess) e = hipMalloc((void**)p, o * 4);
    if (e == hipSuccess) e = hipMalloc((void**)&k->wf, o16 * 2);
    if (e == hipSuccess) e = hipMalloc((void**)&k->wb, o16 * 2);
    for (float** p : f32s)
        if (e == hipSuccess) e = hipMemsetAsync(*p, 0, o * 4, g->stream);
    if (e == hipSuccess) e = hipMemsetAsync(k->wf, 0, o16 * 2, g->stream);
    if (e == hipSuccess) e = hipMemsetAsync(k->wb, 0, o16 * 2, g->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(g->stream);
    if (e != hipSuccess) {
        phr_set_error("phr_kpcn_new", e);
        phr_kpcn_free(k);
        return PHR_ERR_HIP;
    }
    *out = k;
    return PHR_OK;
}

void phr_kpcn_free(phr_kpcn* k) {
    if (!k) return;
    free_ws(k);
    void* ps[7] = {k->master, k->m, k->v, k->grad, k->wf, k->wb, k->batch};
    for (void* p : ps)
        if (p) (void)hipFree(p);
    delete k;
}

int phr_kpcn_set_params(phr_kpcn* k, const float* host, uint64_t count) {
    PHR_CHECK(k && host, PHR_ERR_ARG, "phr_kpcn_set_params: null");
    PHR_CHECK(count == k->nparams, PHR_ERR_SIZE, "phr_kpcn_set_params: count is not phr_kpcn_param_count()");
    hipStream_t s = k->gpu->stream;
    PHR_HIP(hipMemcpyAsync(k->master, host, count * 4, hipMemcpyHostToDevice, s));
    PHR_HIP(hipMemsetAsync(k->m, 0, count * 4, s));
    PHR_HIP(hipMemsetAsync(k->v, 0, count * 4, s));
    k->t = 0;
    KP_TRY(pack_weights(k));
    PHR_HIP(hipStreamSynchronize(s));
    return PHR_OK;
}

static int read_f32(phr_kpcn* k, const float* src, float* host, uint64_t count, const char* what) {
    PHR_CHECK(k && host, PHR_ERR_ARG, std::string(what) + ": null");
    PHR_CHECK(count == k->nparams, PHR_ERR_SIZE, std::string(what) + ": count is not phr_kpcn_param_count()");
    PHR_HIP(hipMemcpyAsync(host, src, count * 4, hipMemcpyDeviceToHost, k->gpu->stream));
    PHR_HIP(hipStreamSynchronize(k->gpu->stream));
    return PHR_OK;
}

int phr_kpcn_get_params(phr_kpcn* k, float* host, uint64_t count) {
    return read_f32(k, k ? k->master : nullptr, host, count, "phr_kpcn_get_params");
}

int phr_kpcn_get_grads(phr_kpcn* k, float* host, uint64_t count) {
    return read_f32(k, k ? k->grad : nullptr, host, count, "phr_kpcn_get_grads");
}

// Device-pointer form of inference, shared with the tracer.
int phr_kpcn_infer_dev(phr_kpcn* k, const float* frames, uint32_t b, uint32_t w, uint32_t h, float* out) {
    KP_TRY(ensure_ws(k, b, h, w, false));
    KP_TRY(forward(k, frames, false));
    return apply(k, frames, out, nullptr, false);
}

int phr_kpcn_infer(phr_kpcn* k, phr_array* frames, uint32_t b, uint32_t w, uint32_t h, phr_array* out) {
    PHR_CHECK(k && out, PHR_ERR_ARG, "phr_kpcn_infer: null");
    KP_TRY(check_shape(frames, b, w, h, out, "phr_kpcn_infer"));
    KP_TRY(phr_kpcn_infer_dev(k, (const float*)frames->ptr, b, w, h, (float*)out->ptr));
    PHR_HIP(hipStreamSynchronize(k->gpu->stream));
    return PHR_OK;
}

int phr_kpcn_loss(phr_kpcn* k, phr_array* frames, phr_array* targets, uint32_t b, uint32_t w, uint32_t h,
                  phr_array* out, float* loss) {
    PHR_CHECK(k && targets && loss, PHR_ERR_ARG, "phr_kpcn_loss: null");
    KP_TRY(check_shape(frames, b, w, h, targets, "phr_kpcn_loss"));
    if (out) KP_TRY(check_shape(frames, b, w, h, out, "phr_kpcn_loss"));
    KP_TRY(ensure_ws(k, b, h, w, false));
    KP_TRY(forward(k, (const float*)frames->ptr, false));
    KP_TRY(apply(k, (const float*)frames->ptr, out ? (float*)out->ptr : nullptr, (const float*)targets->ptr, false));
    return read_loss(k, loss);
}

int phr_kpcn_backward(phr_kpcn* k, phr_array* frames, phr_array* targets, uint32_t b, uint32_t w, uint32_t h,
                      float* loss) {
    PHR_CHECK(k && targets, PHR_ERR_ARG, "phr_kpcn_backward: null");
    KP_TRY(check_shape(frames, b, w, h, targets, "phr_kpcn_backward"));
    KP_TRY(ensure_ws(k, b, h, w, true));
    KP_TRY(forward(k, (const float*)frames->ptr, true));
    KP_TRY(apply(k, (const float*)frames->ptr, nullptr, (const float*)targets->ptr, true));
    KP_TRY(backward(k));
    if (loss) return read_loss(k, loss);
    return PHR_OK;
}

int phr_kpcn_adam(phr_kpcn* k, float lr) {
    PHR_CHECK(k, PHR_ERR_ARG, "phr_kpcn_adam: null");
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    k->t += 1;
    const float bc1 = (float)(1.0 - std::pow((double)b1, (double)k->t));
    const float bc2 = (float)(1.0 - std::pow((double)b2, (double)k->t));
    hipLaunchKernelGGL(k_adam, dim3(grid_of(k->nparams)), dim3(256), 0, k->gpu->stream, k->master, k->grad, k->m, k->v,
                       k->nparams, lr, b1, b2, eps, bc1, bc2);
    PHR_HIP(hipGetLastError());
    return pack_weights(k);
}

// The GPU's partial-sum buffer (HO_BLOCKS pairs of doubles), made on first use.
static int gpu_sums(phr_gpu* g, double** out) {
    if (!g->sums) PHR_HIP(hipMalloc((void**)&g->sums, 16ull * HO_BLOCKS));
    *out = g->sums;
    return PHR_OK;
}

int phr_kpcn_held_out(phr_gpu* g, phr_array* denoised, phr_array* frame_a, phr_array* frame_b, uint64_t n, double* err,
                      double* noise) {
    PHR_CHECK(g && denoised && frame_a && frame_b && err && n, PHR_ERR_ARG, "phr_kpcn_held_out: null");
    PHR_CHECK(denoised->len * denoised->elem >= n * 12 && frame_a->len * frame_a->elem >= n * FRAME * 4 &&
                  frame_b->len * frame_b->elem >= n * FRAME * 4,
              PHR_ERR_SIZE, "phr_kpcn_held_out: arrays smaller than n pixels");
    double* part = nullptr;
    if (int rc = gpu_sums(g, &part); rc != PHR_OK) return rc;
    const uint32_t blocks = (uint32_t)std::min<uint64_t>(HO_BLOCKS, (n + 255) / 256);
    hipLaunchKernelGGL(k_held_out, dim3(blocks), dim3(256), 0, g->stream, (const float*)denoised->ptr, (const float*)frame_a->ptr,
                       (const float*)frame_b->ptr, n, part);
    std::vector<double> h(2ull * blocks);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(h.data(), part, 16ull * blocks, hipMemcpyDeviceToHost, g->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(g->stream);
    if (e != hipSuccess) {
        phr_set_error("phr_kpcn_held_out", e);
        return PHR_ERR_HIP;
    }
    double se = 0.0, sv = 0.0;
    for (uint32_t i = 0; i < blocks; ++i) se += h[2 * i], sv += h[2 * i + 1];
    *err = se / (3.0 * (double)n);
    if (noise) *noise = sv / (3.0 * (double)n);
    return PHR_OK;
}

int phr_kpcn_spread(phr_gpu* g, phr_array* a, phr_array* b, uint64_t n, double* out) {
    PHR_CHECK(g && a && b && out && n, PHR_ERR_ARG, "phr_kpcn_spread: null");
    PHR_CHECK(a->len * a->elem >= n * 12 && b->len * b->elem >= n * 12, PHR_ERR_SIZE, "phr_kpcn_spread: arrays smaller than n pixels");
    double* part = nullptr;
    if (int rc = gpu_sums(g, &part); rc != PHR_OK) return rc;
    const uint32_t blocks = (uint32_t)std::min<uint64_t>(HO_BLOCKS, (3 * n + 255) / 256);
    hipLaunchKernelGGL(k_spread, dim3(blocks), dim3(256), 0, g->stream, (const float*)a->ptr, (const float*)b->ptr, n, part);
    std::vector<double> h(blocks);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(h.data(), part, 8ull * blocks, hipMemcpyDeviceToHost, g->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(g->stream);
    if (e != hipSuccess) {
        phr_set_error("phr_kpcn_spread", e);
        return PHR_ERR_HIP;
    }
    double se = 0.0;
    for (uint32_t i = 0; i < blocks; ++i) se += h[i];
    *out = se / (3.0 * (double)n);
    return PHR_OK;
}

int phr_kpcn_frame_mean(phr_gpu* g, phr_array* frame_a, phr_array* frame_b, uint64_t n, uint32_t channels, phr_array* out) {
    PHR_CHECK(g && frame_a && frame_b && out && n, PHR_ERR_ARG, "phr_kpcn_frame_mean: null");
    PHR_CHECK(channels == FRAME || channels == 3, PHR_ERR_ARG, "phr_kpcn_frame_mean: 10 channels (a frame) or 3 (radiance)");
    PHR_CHECK(frame_a->len * frame_a->elem >= n * FRAME * 4 && frame_b->len * frame_b->elem >= n * FRAME * 4 &&
                  out->len * out->elem >= n * channels * 4,
              PHR_ERR_SIZE, "phr_kpcn_frame_mean: arrays smaller than n pixels");
    hipLaunchKernelGGL(k_frame_mean, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, g->stream, (const float*)frame_a->ptr,
                       (const float*)frame_b->ptr, n, channels, (float*)out->ptr);
    PHR_HIP(hipGetLastError());
    return PHR_OK;
}

// Pick each layer's forward conv tile (4 or 8 rows of 32 pixels a block) by
// timing both on b frames of w x h: the best of `reps` launches each, with HIP
// events on the network's stream. chosen (optional) gets the rows a wave owns
// per layer. The outputs do not depend on the tile.
int phr_kpcn_autotune(phr_kpcn* k, phr_array* frames, uint32_t b, uint32_t w, uint32_t h, uint32_t reps, uint8_t* chosen) {
    PHR_CHECK(k, PHR_ERR_ARG, "phr_kpcn_autotune: null");
    KP_TRY(check_shape(frames, b, w, h, nullptr, "phr_kpcn_autotune"));
    KP_TRY(ensure_ws(k, b, h, w, false));
    KP_TRY(forward(k, (const float*)frames->ptr, false));  // every layer's input in place
    hipStream_t s = k->gpu->stream;
    hipEvent_t e0, e1;
    PHR_HIP(hipEventCreate(&e0));
    PHR_HIP(hipEventCreate(&e1));
    int rc = PHR_OK;
    for (int li = 0; li < NLAYERS && rc == PHR_OK; ++li) {
        float best = 3.0e38f;
        uint8_t pick = 0;
        for (uint8_t rw = 1; rw <= 2 && rc == PHR_OK; ++rw) {
            if (LAYERS[li].cout == 160 && rw == 2) continue;
            k->rw[li] = rw;
            float mine = 3.0e38f;
            for (uint32_t r = 0; r <= std::max<uint32_t>(reps, 1u) && rc == PHR_OK; ++r) {
                hipEventRecord(e0, s);
                rc = conv_layer(k, li);
                hipEventRecord(e1, s);
                if (hipEventSynchronize(e1) != hipSuccess) rc = PHR_ERR_HIP;
                float ms = 0.f;
                hipEventElapsedTime(&ms, e0, e1);
                if (r > 0) mine = std::min(mine, ms);  // the first launch warms the variant
            }
            if (mine < best) best = mine, pick = rw;
        }
        k->rw[li] = pick;
        if (chosen) chosen[li] = pick;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return rc;
}

int phr_kpcn_train_step(phr_kpcn* k, phr_array* frames, phr_array* targets, uint32_t b, uint32_t w, uint32_t h, float lr,
                        float* loss) {
    KP_TRY(phr_kpcn_backward(k, frames, targets, b, w, h, loss));
    return phr_kpcn_adam(k, lr);
}

// One step over b images picked from resident data sets: frame image
// fidx[j] and target image tidx[j] (each w x h) are gathered into the batch.
int phr_kpcn_train_indexed(phr_kpcn* k, phr_array* frames, const uint32_t* fidx, phr_array* targets,
                           const uint32_t* tidx, uint32_t b, uint32_t w, uint32_t h, float lr, float* loss) {
    PHR_CHECK(k && frames && targets && fidx && tidx && b && w && h, PHR_ERR_ARG, "phr_kpcn_train_indexed: null");
    const uint64_t img = (uint64_t)w * h;
    const uint64_t nf = frames->len * frames->elem / (img * FRAME * 4), nt = targets->len * targets->elem / (img * 12);
    for (uint32_t j = 0; j < b; ++j)
        PHR_CHECK(fidx[j] < nf && tidx[j] < nt, PHR_ERR_ARG, "phr_kpcn_train_indexed: image index out of range");
    const uint64_t need_f = b * img * FRAME * 4, need_t = b * img * 12;
    if (k->batch_bytes < need_f + need_t) {
        if (k->batch) (void)hipFree(k->batch);
        k->batch = nullptr;
        k->batch_bytes = 0;
        PHR_HIP(hipMalloc((void**)&k->batch, need_f + need_t));
        k->batch_bytes = need_f + need_t;
    }
    char* bf = k->batch;
    char* bt = k->batch + need_f;
    for (uint32_t j = 0; j < b; ++j) {
        PHR_HIP(hipMemcpyAsync(bf + j * img * FRAME * 4, (const char*)frames->ptr + fidx[j] * img * FRAME * 4,
                               img * FRAME * 4, hipMemcpyDeviceToDevice, k->gpu->stream));
        PHR_HIP(hipMemcpyAsync(bt + j * img * 12, (const char*)targets->ptr + tidx[j] * img * 12, img * 12,
                               hipMemcpyDeviceToDevice, k->gpu->stream));
    }
    KP_TRY(ensure_ws(k, b, h, w, true));
    KP_TRY(forward(k, (const float*)bf, true));
    KP_TRY(apply(k, (const float*)bf, nullptr, (const float*)bt, true));
    KP_TRY(backward(k));
    KP_TRY(phr_kpcn_adam(k, lr));
    if (loss) return read_loss(k, loss);
    return PHR_OK;
}

}  // extern "C"

// phiray's kernel model on HIP: `Kernel` by name, `Pass::add` / `set` / `run`
// (src/kernel/mod.rs:51-133, src/gpu/pass.rs:18-155).
//
// A kernel is found by its phiray NAME and takes its phiray Param bytes
// (first two words n, groups_x, filled in by `add` as the reference does) and
// its bindings in BIND_LIST order. The registered kernels are HIP kernels with
// the semantics of phiray's WGSL ones:
//   bvh.morton  [prims, codes]            src/bvh/kernel.rs:19-51
//   bvh.karras  [codes, links]            src/bvh/kernel.rs:63-117
//   bvh.refit   [prims, order, src, dst]  src/bvh/kernel.rs:128-170 (one Jacobi step)
//   radix.seed / clear / count / offset / scatter   src/bvh/radix.rs:80-202
//   radix.spread [key, codes]             src/bvh/spread.rs:13-32
//   denoise     [px]                      src/kernel/denoise.rs:53-64
// A binding-count mismatch, which panics in the reference
// (src/gpu/pass.rs:78), is PHR_ERR_ARG here. `run` replays the recorded
// steps as one hipGraph, captured on first run and after any add/set/clear.
#include "phr_runtime.h"

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

using namespace phr;

namespace {

constexpr uint32_t BLOCK = 256;       // src/kernel/mod.rs:12
constexpr uint32_t MAX_GROUP = 65535; // src/kernel/mod.rs:14

__device__ __forceinline__ uint32_t item(uint32_t groups_x) {  // src/kernel/mod.rs:129-132
    return blockIdx.x * BLOCK + threadIdx.x + blockIdx.y * groups_x * BLOCK;
}

struct MortonParam { uint32_t n, groups_x, _a[2]; float lo[3], _b, inv[3], _c; };
struct KarrasParam { uint32_t n, groups_x, leaves, _a; };
struct RefitParam { uint32_t n, groups_x, _a[2]; };
struct RadixParam { uint32_t n, groups_x, shift, _a; };
struct DenoiseParam { uint32_t n, groups_x, w, h, step, round, last, keep; float strength[4], view[4]; };

__device__ __forceinline__ uint32_t spread10(float v) {
    float c = v;
    if (c < 0.f) c = 0.f;
    if (c > 1.f) c = 1.f;
    uint32_t x = (c == c) ? (uint32_t)(c * 1023.0f) : 0u;
    x = (x | (x << 16)) & 0x030000FFu;
    x = (x | (x << 8)) & 0x0300F00Fu;
    x = (x | (x << 4)) & 0x030C30C3u;
    return (x | (x << 2)) & 0x09249249u;
}

__global__ __launch_bounds__(256) void kp_morton(MortonParam p, const float4* __restrict__ prims, uint32_t* __restrict__ codes) {
    uint32_t i = item(p.groups_x);
    if (i >= p.n) return;
    float4 a = prims[2 * i], b = prims[2 * i + 1];
    float c0 = 0.5f * (a.x + b.x), c1 = 0.5f * (a.y + b.y), c2 = 0.5f * (a.z + b.z);
    float u0 = (c0 - p.lo[0]) * p.inv[0], u1 = (c1 - p.lo[1]) * p.inv[1], u2 = (c2 - p.lo[2]) * p.inv[2];
    codes[i] = (spread10(u0) << 2) | (spread10(u1) << 1) | spread10(u2);
}

__device__ __forceinline__ int32_t kdelta(const uint32_t* __restrict__ codes, int32_t n, int32_t i, int32_t j) {
    if (j < 0 || j >= n) return -1;
    uint32_t a = codes[i], b = codes[j];
    if (a != b) return (int32_t)__clz(a ^ b);
    return 32 + (int32_t)__clz((uint32_t)i ^ (uint32_t)j);
}

__global__ __launch_bounds__(256) void kp_karras(KarrasParam p, const uint32_t* __restrict__ codes, uint4* __restrict__ links) {
    uint32_t k = item(p.groups_x);
    if (k >= p.n) return;
    int32_t i = (int32_t)k, n = (int32_t)p.leaves;
    int32_t d = (kdelta(codes, n, i, i + 1) - kdelta(codes, n, i, i - 1) > 0) ? 1 : -1;
    int32_t dmin = kdelta(codes, n, i, i - d);
    int32_t lmax = 2;
    while (kdelta(codes, n, i, i + lmax * d) > dmin) lmax *= 2;
    int32_t l = 0, t = lmax / 2;
    while (t >= 1) {
        if (kdelta(codes, n, i, i + (l + t) * d) > dmin) l += t;
        t /= 2;
    }
    int32_t j = i + l * d;
    int32_t dnode = kdelta(codes, n, i, j);
    int32_t s = 0;
    t = l;
    for (;;) {
        t = (t + 1) >> 1;
        if (kdelta(codes, n, i, i + (s + t) * d) > dnode) s += t;
        if (t <= 1) break;
    }
    int32_t g = i + s * d + (d < 0 ? d : 0);
    int32_t first = i < j ? i : j, last = i < j ? j : i;
    int32_t left = (first == g) ? n - 1 + g : g;
    int32_t right = (last == g + 1) ? n + g : g + 1;
    links[k] = make_uint4((uint32_t)left, (uint32_t)right, (uint32_t)first, (uint32_t)last);
}

__global__ __launch_bounds__(256) void kp_refit(RefitParam p, const float4* __restrict__ prims, const uint32_t* __restrict__ order,
                                                const float4* __restrict__ src, float4* __restrict__ dst) {
    uint32_t k = item(p.groups_x);
    if (k >= p.n) return;
    float4 lo = src[2 * k], hi = src[2 * k + 1];
    uint32_t li = __float_as_uint(lo.w), ri = __float_as_uint(hi.w);
    float l0 = 3.4028235e38f, l1 = 3.4028235e38f, l2 = 3.4028235e38f;
    float h0 = -3.4028235e38f, h1 = -3.4028235e38f, h2 = -3.4028235e38f;
    if (li & PHR_LEAF) {
        for (uint32_t s = li & ~PHR_LEAF; s < ri; ++s) {
            uint32_t q = order[s];
            float4 a = prims[2 * q], b = prims[2 * q + 1];
            l0 = fminf(l0, a.x); l1 = fminf(l1, a.y); l2 = fminf(l2, a.z);
            h0 = fmaxf(h0, b.x); h1 = fmaxf(h1, b.y); h2 = fmaxf(h2, b.z);
        }
    } else {
        float4 al = src[2 * li], ah = src[2 * li + 1], bl = src[2 * ri], bh = src[2 * ri + 1];
        l0 = fminf(al.x, bl.x); l1 = fminf(al.y, bl.y); l2 = fminf(al.z, bl.z);
        h0 = fmaxf(ah.x, bh.x); h1 = fmaxf(ah.y, bh.y); h2 = fmaxf(ah.z, bh.z);
    }
    dst[2 * k] = make_float4(l0, l1, l2, lo.w);
    dst[2 * k + 1] = make_float4(h0, h1, h2, hi.w);
}

__device__ __forceinline__ uint32_t rdigit(uint2 k, uint32_t shift) {  // src/bvh/radix.rs:69-72
    return shift < 32u ? (k.x >> shift) & 0xFFu : (k.y >> (shift - 32u)) & 0xFFu;
}

__global__ __launch_bounds__(256) void kp_seed(RadixParam p, const uint32_t* __restrict__ codes, uint2* __restrict__ key) {
    uint32_t i = item(p.groups_x);
    if (i >= p.n) return;
    key[i] = make_uint2(i, codes[i]);
}
__global__ __launch_bounds__(256) void kp_clear(RadixParam p, uint32_t* __restrict__ tally) {
    uint32_t i = item(p.groups_x);
    if (i >= p.n) return;
    tally[i] = 0u;
}
__global__ __launch_bounds__(256) void kp_count(RadixParam p, const uint2* __restrict__ key, uint32_t* __restrict__ tally) {
    uint32_t i = item(p.groups_x);
    if (i >= p.n) return;
    atomicAdd(&tally[rdigit(key[i], p.shift)], 1u);
}
__global__ __launch_bounds__(256) void kp_offset(RadixParam p, uint32_t* __restrict__ tally) {
    if (item(p.groups_x) != 0u) return;
    uint32_t running = 0;
    for (uint32_t k = 0; k < 256u; ++k) {
        uint32_t c = tally[k];
        tally[256u + k] = running;
        running += c;
        tally[k] = 0u;
    }
}
// radix.scatter, stable. phiray's kernel takes each key's slot from an atomic
// cursor (src/bvh/radix.rs:193-199), so keys of one digit land in whatever
// order the atomics resolve in; its own model (src/bvh/sort.rs:127-150) and
// the order it promises (Bvh::sort_codes) need every pass stable. Here a tile
// of 2048 keys ranks its digits in index order with wave ballots, publishes
// its per-digit counts, and looks back over earlier tiles (ids from a ticket,
// so an earlier tile is always running or done) for its offsets within the
// digit's run that starts at the cursor `tally[256 + d]`. The result is the
// model's, whatever the timing.
constexpr uint32_t SC_ITEMS = 8, SC_TILE = 256 * SC_ITEMS;
constexpr uint32_t SC_AGG = 0x40000000u, SC_PRE = 0x80000000u, SC_VAL = 0x3fffffffu;

__global__ __launch_bounds__(256) void kp_scatter_clear(uint32_t* __restrict__ scratch, uint32_t words) {
    for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < words; i += gridDim.x * 256) scratch[i] = 0u;
}

__global__ __launch_bounds__(256) void kp_scatter(RadixParam p, const uint2* __restrict__ src, uint2* __restrict__ dst,
                                                  const uint32_t* __restrict__ tally, uint32_t* __restrict__ scratch) {
    __shared__ uint32_t s_whist[4][256];
    __shared__ uint32_t s_base[256];
    __shared__ uint32_t s_tile;
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint32_t n = p.n;
    if (tid == 0) s_tile = atomicAdd(&scratch[0], 1u);
    for (uint32_t i = tid; i < 4 * 256; i += 256) (&s_whist[0][0])[i] = 0u;
    __syncthreads();
    const uint32_t tile = s_tile;
    const uint32_t wbase = tile * SC_TILE + wave * (SC_ITEMS * 64);
    const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    uint2 key[SC_ITEMS];
    uint32_t rank[SC_ITEMS];
#pragma unroll
    for (uint32_t r = 0; r < SC_ITEMS; ++r) {
        uint32_t idx = wbase + r * 64 + lane;
        bool ok = idx < n;
        key[r] = ok ? src[idx] : make_uint2(0u, 0u);
        uint32_t d = rdigit(key[r], p.shift);
        uint64_t mask = __ballot(ok);
#pragma unroll
        for (uint32_t b = 0; b < 8; ++b) {
            uint64_t bal = __ballot((d >> b) & 1u);
            mask &= ((d >> b) & 1u) ? bal : ~bal;
        }
        uint32_t pre = 0;
        if (ok) pre = s_whist[wave][d];
        uint32_t below = (uint32_t)__popcll(mask & lt);
        if (ok && below == 0) s_whist[wave][d] = pre + (uint32_t)__popcll(mask);
        rank[r] = pre + below;
    }
    __syncthreads();
    const uint32_t d = tid;
    uint32_t c0 = s_whist[0][d], c1 = s_whist[1][d], c2 = s_whist[2][d], c3 = s_whist[3][d];
    uint32_t total = c0 + c1 + c2 + c3;
    uint32_t* status = scratch + 1;
    uint32_t* my = status + (size_t)tile * 256 + d;
    __hip_atomic_store(my, (tile == 0 ? SC_PRE : SC_AGG) | total, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    uint32_t excl = 0;
    if (tile > 0) {
        int32_t j = (int32_t)tile - 1;
        for (;;) {
            uint32_t s = __hip_atomic_load(status + (size_t)j * 256 + d, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (s & SC_PRE) { excl += s & SC_VAL; break; }
            if (s & SC_AGG) { excl += s & SC_VAL; --j; continue; }
            __builtin_amdgcn_s_sleep(1);
        }
        __hip_atomic_store(my, SC_PRE | (excl + total), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    s_whist[0][d] = 0;
    s_whist[1][d] = c0;
    s_whist[2][d] = c0 + c1;
    s_whist[3][d] = c0 + c1 + c2;
    s_base[d] = tally[256u + d] + excl;
    __syncthreads();
#pragma unroll
    for (uint32_t r = 0; r < SC_ITEMS; ++r) {
        uint32_t idx = wbase + r * 64 + lane;
        if (idx < n) {
            uint32_t dd = rdigit(key[r], p.shift);
            dst[s_base[dd] + s_whist[wave][dd] + rank[r]] = key[r];
        }
    }
}
__global__ __launch_bounds__(256) void kp_spread(RadixParam p, const uint2* __restrict__ key, uint32_t* __restrict__ codes) {
    uint32_t i = item(p.groups_x);
    if (i >= p.n) return;
    codes[i] = key[i].y;
}

struct KernelInfo {
    const char* name;
    uint32_t param_size;
    uint32_t binds;
};
const KernelInfo KERNELS[] = {
    {"bvh.morton", sizeof(MortonParam), 2}, {"bvh.karras", sizeof(KarrasParam), 2}, {"bvh.refit", sizeof(RefitParam), 4},
    {"radix.seed", sizeof(RadixParam), 2},  {"radix.clear", sizeof(RadixParam), 1}, {"radix.count", sizeof(RadixParam), 2},
    {"radix.offset", sizeof(RadixParam), 1}, {"radix.scatter", sizeof(RadixParam), 3}, {"radix.spread", sizeof(RadixParam), 2},
    {"denoise", sizeof(DenoiseParam), 1},
};

}  // namespace

int phr_denoise_pass_enqueue(phr_gpu* g, uint32_t* px, uint32_t w, uint32_t h, uint32_t step, uint32_t round, uint32_t last,
                             uint32_t keep, const float s[4], const float view[4]);

struct phr_pass {
    struct Step {
        int kind;
        std::vector<uint8_t> params;
        std::vector<phr_array*> arrays;
        uint32_t items;
        dim3 grid;
        uint32_t* scratch = nullptr;  // radix.scatter's ticket and look-back words
        uint32_t scratch_words = 0;
    };
    phr_gpu* gpu = nullptr;
    std::vector<Step> steps;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool dirty = true;
};

namespace {

void free_steps(phr_pass* p) {
    for (auto& s : p->steps)
        if (s.scratch) (void)hipFree(s.scratch);
    p->steps.clear();
}

void drop_graph(phr_pass* p) {
    if (p->exec) (void)hipGraphExecDestroy(p->exec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    p->exec = nullptr;
    p->graph = nullptr;
    p->dirty = true;
}

template <class T>
T param_of(const phr_pass::Step& s) {
    T t;
    std::memcpy(&t, s.params.data(), sizeof t);
    return t;
}

void launch(phr_pass* ps, const phr_pass::Step& s, hipStream_t st) {
    auto P = [&](size_t k) { return s.arrays[k]->ptr; };
    switch (s.kind) {
        case 0: hipLaunchKernelGGL(kp_morton, s.grid, dim3(BLOCK), 0, st, param_of<MortonParam>(s), (const float4*)P(0), (uint32_t*)P(1)); break;
        case 1: hipLaunchKernelGGL(kp_karras, s.grid, dim3(BLOCK), 0, st, param_of<KarrasParam>(s), (const uint32_t*)P(0), (uint4*)P(1)); break;
        case 2: hipLaunchKernelGGL(kp_refit, s.grid, dim3(BLOCK), 0, st, param_of<RefitParam>(s), (const float4*)P(0), (const uint32_t*)P(1),
                                   (const float4*)P(2), (float4*)P(3)); break;
        case 3: hipLaunchKernelGGL(kp_seed, s.grid, dim3(BLOCK), 0, st, param_of<RadixParam>(s), (const uint32_t*)P(0), (uint2*)P(1)); break;
        case 4: hipLaunchKernelGGL(kp_clear, s.grid, dim3(BLOCK), 0, st, param_of<RadixParam>(s), (uint32_t*)P(0)); break;
        case 5: hipLaunchKernelGGL(kp_count, s.grid, dim3(BLOCK), 0, st, param_of<RadixParam>(s), (const uint2*)P(0), (uint32_t*)P(1)); break;
        case 6: hipLaunchKernelGGL(kp_offset, s.grid, dim3(BLOCK), 0, st, param_of<RadixParam>(s), (uint32_t*)P(0)); break;
        case 7: {
            RadixParam rp = param_of<RadixParam>(s);
            uint32_t tiles = (rp.n + SC_TILE - 1) / SC_TILE;
            hipLaunchKernelGGL(kp_scatter_clear, dim3(std::min<uint32_t>((s.scratch_words + 255) / 256, 1024)), dim3(256), 0, st,
                               s.scratch, s.scratch_words);
            if (tiles)
                hipLaunchKernelGGL(kp_scatter, dim3(tiles), dim3(256), 0, st, rp, (const uint2*)P(0), (uint2*)P(1),
                                   (const uint32_t*)P(2), s.scratch);
            break;
        }
        case 8: hipLaunchKernelGGL(kp_spread, s.grid, dim3(BLOCK), 0, st, param_of<RadixParam>(s), (const uint2*)P(0), (uint32_t*)P(1)); break;
        case 9: {
            DenoiseParam d = param_of<DenoiseParam>(s);
            (void)phr_denoise_pass_enqueue(ps->gpu, (uint32_t*)P(0), d.w, d.h, d.step, d.round, d.last, d.keep, d.strength, d.view);
            break;
        }
        default: break;
    }
}

}  // namespace

extern "C" {

int phr_pass_new(phr_gpu* g, phr_pass** out) {
    PHR_CHECK(g && out, PHR_ERR_ARG, "phr_pass_new: null");
    phr_pass* p = new phr_pass();
    p->gpu = g;
    *out = p;
    return PHR_OK;
}

void phr_pass_free(phr_pass* p) {
    if (!p) return;
    drop_graph(p);
    (void)hipStreamSynchronize(p->gpu->stream);
    free_steps(p);
    delete p;
}

int phr_kernel_known(const char* name, uint32_t* param_size, uint32_t* binds) {
    PHR_CHECK(name, PHR_ERR_ARG, "phr_kernel_known: null name");
    for (const KernelInfo& k : KERNELS)
        if (std::strcmp(k.name, name) == 0) {
            if (param_size) *param_size = k.param_size;
            if (binds) *binds = k.binds;
            return PHR_OK;
        }
    phr_set_error_msg(std::string("no kernel named ") + name);
    return PHR_ERR_UNSUPPORTED;
}

// Pass::add_over (src/gpu/pass.rs:70-101): a launch of `items` invocations of
// kernel `name` whose params say `n` items in all. Returns the step index.
int phr_pass_add(phr_pass* p, const char* name, uint32_t n, uint32_t items, const void* params, uint32_t param_len,
                 phr_array* const* arrays, uint32_t count, uint32_t* step) {
    PHR_CHECK(p && name && params, PHR_ERR_ARG, "phr_pass_add: null argument");
    int kind = -1;
    for (int k = 0; k < (int)(sizeof KERNELS / sizeof KERNELS[0]); ++k)
        if (std::strcmp(KERNELS[k].name, name) == 0) kind = k;
    PHR_CHECK(kind >= 0, PHR_ERR_UNSUPPORTED, std::string("phr_pass_add: no kernel named ") + name);
    const KernelInfo& ki = KERNELS[kind];
    PHR_CHECK(param_len == ki.param_size, PHR_ERR_ARG, std::string(name) + ": param size");
    PHR_CHECK(count == ki.binds, PHR_ERR_ARG, std::string(name) + ": binding count");
    for (uint32_t k = 0; k < count; ++k) PHR_CHECK(arrays && arrays[k] && arrays[k]->gpu == p->gpu, PHR_ERR_ARG, std::string(name) + ": binding");
    phr_pass::Step s;
    s.kind = kind;
    s.params.assign((const uint8_t*)params, (const uint8_t*)params + param_len);
    s.arrays.assign(arrays, arrays + count);
    s.items = items;
    // Dispatch::new (src/kernel/mod.rs:116-120): x <= 65535, spilling into y.
    uint32_t blocks = std::max<uint32_t>((items + BLOCK - 1) / BLOCK, 1);
    uint32_t x = std::min(blocks, MAX_GROUP);
    s.grid = dim3(x, (blocks + x - 1) / x, 1);
    uint32_t head[2] = {n, x};
    std::memcpy(s.params.data(), head, 8);
    if (kind == 7) {  // radix.scatter keeps look-back words sized for its n
        s.scratch_words = 1 + ((n + SC_TILE - 1) / SC_TILE) * 256;
        if (hipMalloc(&s.scratch, 4ull * s.scratch_words) != hipSuccess) {
            phr_set_error_msg("radix.scatter: scratch allocation failed");
            return PHR_ERR_HIP;
        }
    }
    p->steps.push_back(std::move(s));
    drop_graph(p);
    if (step) *step = (uint32_t)p->steps.size() - 1;
    return PHR_OK;
}

// Pass::set (src/gpu/pass.rs:104-106): overwrite a step's params from `offset`.
int phr_pass_set(phr_pass* p, uint32_t step, uint32_t offset, const void* data, uint32_t len) {
    PHR_CHECK(p && data && step < p->steps.size(), PHR_ERR_ARG, "phr_pass_set: bad step");
    auto& prm = p->steps[step].params;
    PHR_CHECK((uint64_t)offset + len <= prm.size(), PHR_ERR_SIZE, "phr_pass_set: past the params");
    std::memcpy(prm.data() + offset, data, len);
    drop_graph(p);
    return PHR_OK;
}

// Pass::run (src/gpu/pass.rs:109-113): every step, in order, as one graph.
int phr_pass_run(phr_pass* p) {
    PHR_CHECK(p, PHR_ERR_ARG, "phr_pass_run: null");
    if (p->steps.empty()) return PHR_OK;
    hipStream_t st = p->gpu->stream;
    if (!p->exec) {
        PHR_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        for (const auto& s : p->steps) launch(p, s, st);
        hipError_t e = hipStreamEndCapture(st, &p->graph);
        if (e != hipSuccess) { p->graph = nullptr; phr_set_error("hipStreamEndCapture", e); return PHR_ERR_HIP; }
        PHR_HIP(hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0));
        p->dirty = false;
    }
    PHR_HIP(hipGraphLaunch(p->exec, st));
    return PHR_OK;
}

// Pass::run_step (src/gpu/pass.rs:126-137): one step on its own.
int phr_pass_run_step(phr_pass* p, uint32_t step) {
    PHR_CHECK(p && step < p->steps.size(), PHR_ERR_ARG, "phr_pass_run_step: bad step");
    launch(p, p->steps[step], p->gpu->stream);
    PHR_HIP(hipGetLastError());
    return PHR_OK;
}

void phr_pass_clear(phr_pass* p) {
    if (!p) return;
    drop_graph(p);
    (void)hipStreamSynchronize(p->gpu->stream);
    free_steps(p);
}
uint32_t phr_pass_len(const phr_pass* p) { return p ? (uint32_t)p->steps.size() : 0; }

// Pass::compiled (src/gpu/pass.rs:151-154): kernels this pass uses, one per
// kernel type however many steps run it.
uint32_t phr_pass_compiled(const phr_pass* p) {
    if (!p) return 0;
    uint32_t seen = 0, n = 0;
    for (const auto& s : p->steps)
        if (!(seen & (1u << s.kind))) {
            seen |= 1u << s.kind;
            ++n;
        }
    return n;
}

}  // extern "C"

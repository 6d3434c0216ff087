// GPU LBVH for gfx950: scene bounds, 30-bit Morton keys, a wave64 onesweep
// radix sort of (code, slot), Karras links, leaf packing and an atomic-counter
// refit into phiray's 32-byte Node layout.
//
// The result is the tree phiray's CPU builder makes (src/bvh/build.rs:94-100,
// Bvh::build): the same order, links, depth and reachable-node bounds. The
// stages and why each looks the way it does:
//
//  1. bounds   one pass over the boxes; blocks reduce in registers and LDS, the
//              last block to finish folds the partials and writes the scene
//              box and its per-axis reciprocal (src/bvh/build.rs:70-80). The
//              same launch clears the sort's counters.
//  2. morton   the code of every box centre (src/bvh/node.rs:96-105) and, in
//              the same pass, the histograms of all four 8-bit digits.
//  3. sort x4  onesweep LSD passes (bits 0-7, 8-15, 16-23, 24-29): a tile of
//              2048 keys ranks its digits with wave64 ballots into per-wave LDS
//              counters, publishes its per-digit counts and looks back over
//              earlier tiles for its offsets (tiles take ids from a ticket so
//              an earlier tile is always resident), then scatters through LDS so
//              each digit's run leaves in one contiguous write. Every pass is
//              stable and pass 0 starts from the slot order, so the result is
//              the sort by (code, slot) that src/bvh/build.rs:87-92 asks for.
//  4. karras   one thread per internal node: direction, exponential and binary
//              search, split by clz of XOR'd keys (src/bvh/build.rs:28-61), and
//              the parent of each child.
//  5. pack     one thread per node climbs its parents (at most 32 steps) for its
//              depth; a subtree of fewer than `leaf + 1` slots or at depth 31
//              becomes a leaf (src/bvh/build.rs:139-168), whatever lies under a
//              packed node is unreachable, and the reachable leaves are listed.
//  6. refit    one thread per listed leaf: its box from its primitives, then up
//              the parents; at each inner node the second of its two children
//              to arrive (an atomic counter) unions both and carries on, and
//              resets the counter for the next refit (src/bvh/build.rs:180-194).
#include "phr_runtime.h"

#include <algorithm>
#include <cstring>
#include <vector>

namespace {

constexpr uint32_t OS_THREADS = 256;
constexpr uint32_t OS_ITEMS = 8;
constexpr uint32_t OS_TILE = OS_THREADS * OS_ITEMS;  // 2048 keys a tile
constexpr uint32_t OS_WAVES = OS_THREADS / 64;
constexpr uint32_t OS_PASSES = 4;
constexpr uint32_t FLAG_AGG = 0x40000000u;
constexpr uint32_t FLAG_PRE = 0x80000000u;
constexpr uint32_t VAL_MASK = 0x3fffffffu;
constexpr uint32_t BOUNDS_BLOCKS_MAX = 512;

// Words at the front of the cleared block: 4 x 256 digit counts, then the
// pass tickets, the leaf count and the depth; the look-back status follows.
constexpr uint32_t Z_HIST = 0;
constexpr uint32_t Z_TICKET = 1024;
constexpr uint32_t Z_LEAVES = 1028;
constexpr uint32_t Z_DEPTH = 1029;
constexpr uint32_t Z_STATUS = 1088;

__device__ __forceinline__ uint64_t lanemask_lt() {
    uint32_t lane = __lane_id();
    return (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
}

__device__ __forceinline__ float wave_min(float v) {
    for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o));
    return v;
}
__device__ __forceinline__ float wave_max(float v) {
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
    return v;
}
__device__ __forceinline__ uint32_t wave_max_u(uint32_t v) {
    for (int o = 32; o > 0; o >>= 1) v = max(v, (uint32_t)__shfl_xor((int)v, o));
    return v;
}

// Exclusive scan of one value per thread over a 256-thread block.
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* s_wave /*4*/) {
    uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t x = v;
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t y = (uint32_t)__shfl_up((int)x, o);
        if (lane >= (uint32_t)o) x += y;
    }
    if (lane == 63) s_wave[wave] = x;
    __syncthreads();
    uint32_t add = 0;
    for (uint32_t w = 0; w < wave; ++w) add += s_wave[w];
    __syncthreads();
    return add + x - v;
}

// ---- 1. scene bounds ------------------------------------------------------
__global__ __launch_bounds__(256) void k_bounds(const float4* __restrict__ prims, uint32_t n, float* __restrict__ partials,
                                                float* __restrict__ scene, uint32_t* __restrict__ arrive,
                                                uint32_t* __restrict__ zero, uint32_t zero_words) {
    __shared__ float s_red[6][4];
    __shared__ uint32_t s_last;
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint32_t gtid = blockIdx.x * blockDim.x + tid, gsize = gridDim.x * blockDim.x;
    for (uint32_t i = gtid; i < zero_words; i += gsize) zero[i] = 0u;
    // Aabb::default() is (f32::MAX, f32::MIN), src/bvh/node.rs:67-71.
    float lo0 = 3.40282347e38f, lo1 = 3.40282347e38f, lo2 = 3.40282347e38f;
    float hi0 = -3.40282347e38f, hi1 = -3.40282347e38f, hi2 = -3.40282347e38f;
    for (uint32_t i = gtid; i < n; i += gsize) {
        float4 a = prims[2 * i], b = prims[2 * i + 1];
        lo0 = fminf(lo0, a.x); lo1 = fminf(lo1, a.y); lo2 = fminf(lo2, a.z);
        hi0 = fmaxf(hi0, b.x); hi1 = fmaxf(hi1, b.y); hi2 = fmaxf(hi2, b.z);
    }
    lo0 = wave_min(lo0); lo1 = wave_min(lo1); lo2 = wave_min(lo2);
    hi0 = wave_max(hi0); hi1 = wave_max(hi1); hi2 = wave_max(hi2);
    if (lane == 0) {
        s_red[0][wave] = lo0; s_red[1][wave] = lo1; s_red[2][wave] = lo2;
        s_red[3][wave] = hi0; s_red[4][wave] = hi1; s_red[5][wave] = hi2;
    }
    __syncthreads();
    if (tid < 6) {
        float v = s_red[tid][0];
        for (int w = 1; w < 4; ++w) v = tid < 3 ? fminf(v, s_red[tid][w]) : fmaxf(v, s_red[tid][w]);
        partials[8 * blockIdx.x + tid] = v;
    }
    __threadfence();
    __syncthreads();
    if (tid == 0) s_last = (atomicAdd(arrive, 1u) == gridDim.x - 1) ? 1u : 0u;
    __syncthreads();
    if (!s_last) return;
    __threadfence();
    if (tid < 6) {
        float v = tid < 3 ? 3.40282347e38f : -3.40282347e38f;
        for (uint32_t b = 0; b < gridDim.x; ++b) {
            float p = partials[8 * b + tid];
            v = tid < 3 ? fminf(v, p) : fmaxf(v, p);
        }
        s_red[tid][0] = v;
    }
    __syncthreads();
    if (tid < 3) {
        float lo = s_red[tid][0], hi = s_red[tid + 3][0];
        float e = hi - lo;
        scene[tid] = lo;
        scene[4 + tid] = e > 0.f ? 1.0f / e : 0.f;  // src/bvh/build.rs:76-80
    }
    if (tid == 0) *arrive = 0u;
}

// ---- 2. Morton codes + all four digit histograms ---------------------------
__device__ __forceinline__ uint32_t spread10(float v) {
    // src/bvh/node.rs:97-103: clamp to [0, 1], times 1023, truncated; NaN -> 0.
    float c = v;
    if (c < 0.f) c = 0.f;
    if (c > 1.f) c = 1.f;
    uint32_t x = (c == c) ? (uint32_t)(c * 1023.0f) : 0u;
    x = (x | (x << 16)) & 0x030000FFu;
    x = (x | (x << 8)) & 0x0300F00Fu;
    x = (x | (x << 4)) & 0x030C30C3u;
    return (x | (x << 2)) & 0x09249249u;
}

__global__ __launch_bounds__(256) void k_morton(const float4* __restrict__ prims, const float* __restrict__ scene, uint32_t n,
                                                uint32_t* __restrict__ codes, uint32_t* __restrict__ hist) {
    __shared__ uint32_t s_h[OS_PASSES * 256];
    const uint32_t tid = threadIdx.x;
    for (uint32_t i = tid; i < OS_PASSES * 256; i += 256) s_h[i] = 0u;
    __syncthreads();
    const float slo0 = scene[0], slo1 = scene[1], slo2 = scene[2];
    const float inv0 = scene[4], inv1 = scene[5], inv2 = scene[6];
    const uint32_t base = blockIdx.x * OS_TILE;
#pragma unroll
    for (uint32_t r = 0; r < OS_ITEMS; ++r) {
        uint32_t i = base + r * 256 + tid;
        if (i < n) {
            float4 a = prims[2 * i], b = prims[2 * i + 1];
            // Aabb::center then (c - s.lo) * inv, src/bvh/build.rs:81-84
            float c0 = 0.5f * (a.x + b.x), c1 = 0.5f * (a.y + b.y), c2 = 0.5f * (a.z + b.z);
            float u0 = (c0 - slo0) * inv0, u1 = (c1 - slo1) * inv1, u2 = (c2 - slo2) * inv2;
            uint32_t code = (spread10(u0) << 2) | (spread10(u1) << 1) | spread10(u2);
            codes[i] = code;
            atomicAdd(&s_h[code & 0xFFu], 1u);
            atomicAdd(&s_h[256 + ((code >> 8) & 0xFFu)], 1u);
            atomicAdd(&s_h[512 + ((code >> 16) & 0xFFu)], 1u);
            atomicAdd(&s_h[768 + (code >> 24)], 1u);
        }
    }
    __syncthreads();
    for (uint32_t i = tid; i < OS_PASSES * 256; i += 256)
        if (s_h[i]) atomicAdd(&hist[i], s_h[i]);
}

// ---- 3. onesweep pass ------------------------------------------------------
template <bool FIRST>
__global__ __launch_bounds__(256) void k_onesweep(const uint32_t* __restrict__ kin, const uint32_t* __restrict__ vin,
                                                  uint32_t* __restrict__ kout, uint32_t* __restrict__ vout,
                                                  const uint32_t* __restrict__ hist, uint32_t* status,
                                                  uint32_t* ticket, uint32_t n, uint32_t shift) {
    __shared__ uint32_t s_whist[OS_WAVES][256];
    __shared__ uint32_t s_local[256];
    __shared__ uint32_t s_global[256];
    __shared__ uint32_t s_keys[OS_TILE];
    __shared__ uint32_t s_vals[OS_TILE];
    __shared__ uint32_t s_scan[4];
    __shared__ uint32_t s_tile;
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (tid == 0) s_tile = atomicAdd(ticket, 1u);
    for (uint32_t i = tid; i < OS_WAVES * 256; i += 256) (&s_whist[0][0])[i] = 0u;
    __syncthreads();
    const uint32_t tile = s_tile;
    const uint32_t wbase = tile * OS_TILE + wave * (OS_ITEMS * 64);
    uint32_t key[OS_ITEMS], val[OS_ITEMS], rank[OS_ITEMS];
#pragma unroll
    for (uint32_t r = 0; r < OS_ITEMS; ++r) {
        uint32_t idx = wbase + r * 64 + lane;
        bool ok = idx < n;
        key[r] = ok ? kin[idx] : 0u;
        val[r] = FIRST ? idx : (ok ? vin[idx] : 0u);
    }
    const uint64_t lt = lanemask_lt();
#pragma unroll
    for (uint32_t r = 0; r < OS_ITEMS; ++r) {
        uint32_t idx = wbase + r * 64 + lane;
        bool ok = idx < n;
        uint32_t d = (key[r] >> shift) & 0xFFu;
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
    // One thread per digit: per-wave offsets, tile total, offsets in the output.
    const uint32_t d = tid;
    uint32_t c0 = s_whist[0][d], c1 = s_whist[1][d], c2 = s_whist[2][d], c3 = s_whist[3][d];
    uint32_t total = c0 + c1 + c2 + c3;
    uint32_t* my = status + (size_t)tile * 256 + d;
    __hip_atomic_store(my, (tile == 0 ? FLAG_PRE : FLAG_AGG) | total, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    s_whist[0][d] = 0;
    s_whist[1][d] = c0;
    s_whist[2][d] = c0 + c1;
    s_whist[3][d] = c0 + c1 + c2;
    uint32_t local = block_exclusive_scan(total, s_scan);
    uint32_t binbase = block_exclusive_scan(hist[d], s_scan);
    uint32_t excl = 0;
    if (tile > 0) {
        int32_t j = (int32_t)tile - 1;
        for (;;) {
            uint32_t s = __hip_atomic_load(status + (size_t)j * 256 + d, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (s & FLAG_PRE) { excl += s & VAL_MASK; break; }
            if (s & FLAG_AGG) { excl += s & VAL_MASK; --j; continue; }
            __builtin_amdgcn_s_sleep(1);
        }
        __hip_atomic_store(my, FLAG_PRE | (excl + total), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    s_local[d] = local;
    s_global[d] = binbase + excl;
    __syncthreads();
#pragma unroll
    for (uint32_t r = 0; r < OS_ITEMS; ++r) {
        uint32_t idx = wbase + r * 64 + lane;
        if (idx < n) {
            uint32_t dd = (key[r] >> shift) & 0xFFu;
            uint32_t lp = s_local[dd] + s_whist[wave][dd] + rank[r];
            s_keys[lp] = key[r];
            s_vals[lp] = val[r];
        }
    }
    __syncthreads();
    const uint32_t count = min(OS_TILE, n - tile * OS_TILE);
#pragma unroll
    for (uint32_t r = 0; r < OS_ITEMS; ++r) {
        uint32_t idx = r * 256 + tid;
        if (idx < count) {
            uint32_t k = s_keys[idx];
            uint32_t dd = (k >> shift) & 0xFFu;
            uint32_t dest = s_global[dd] + (idx - s_local[dd]);
            kout[dest] = k;
            vout[dest] = s_vals[idx];
        }
    }
}

// ---- 4. Karras links --------------------------------------------------------
__device__ __forceinline__ int32_t kdelta(const uint32_t* __restrict__ codes, int64_t n, int64_t i, int64_t j) {
    // src/bvh/mod.rs:33-43
    if (j < 0 || j >= n) return -1;
    uint32_t a = codes[i], b = codes[j];
    if (a != b) return (int32_t)__clz(a ^ b);
    return 32 + (int32_t)__clz((uint32_t)i ^ (uint32_t)j);
}

__global__ __launch_bounds__(256) void k_karras(const uint32_t* __restrict__ codes, uint32_t n32, PhrLink* __restrict__ links,
                                                uint32_t* __restrict__ parent) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k + 1 >= n32) return;
    const int64_t n = n32, i = k;
    int64_t d = (kdelta(codes, n, i, i + 1) - kdelta(codes, n, i, i - 1) > 0) ? 1 : -1;
    int32_t dmin = kdelta(codes, n, i, i - d);
    int64_t lmax = 2;
    while (kdelta(codes, n, i, i + lmax * d) > dmin) lmax *= 2;
    int64_t l = 0, t = lmax / 2;
    while (t >= 1) {
        if (kdelta(codes, n, i, i + (l + t) * d) > dmin) l += t;
        t /= 2;
    }
    int64_t j = i + l * d;
    int32_t dnode = kdelta(codes, n, i, j);
    int64_t s = 0;
    t = l;
    for (;;) {
        t = (t + 1) >> 1;
        if (kdelta(codes, n, i, i + (s + t) * d) > dnode) s += t;
        if (t <= 1) break;
    }
    int64_t g = i + s * d + (d < 0 ? d : 0);
    int64_t first = i < j ? i : j, last = i < j ? j : i;
    uint32_t left = (uint32_t)((first == g) ? n - 1 + g : g);
    uint32_t right = (uint32_t)((last == g + 1) ? n + g : g + 1);
    links[k] = PhrLink{left, right, (uint32_t)first, (uint32_t)last};
    parent[left] = k;
    parent[right] = k;
    if (k == 0) parent[0] = PHR_NONE;
}

// ---- 5. leaf packing --------------------------------------------------------
__global__ __launch_bounds__(256) void k_pack(const PhrLink* __restrict__ links, const uint32_t* __restrict__ parent,
                                              uint32_t n, uint32_t leaf, uint32_t* __restrict__ leaf_list,
                                              uint32_t* __restrict__ leaf_count, uint32_t* __restrict__ depth_out) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t total = 2 * n - 1;
    bool listed = false;
    uint32_t depth = 0;
    if (k < total) {
        // Depth by climbing; deeper than 31 means an ancestor at depth 31 was
        // packed (d + 1 >= STACK), so this node is unreachable.
        uint32_t x = k, d = 0;
        bool deep = false;
        while (x != 0) {
            if (d == PHR_STACK - 1) { deep = true; break; }
            x = parent[x];
            ++d;
        }
        if (!deep) {
            // A packed parent mutes its subtree. Ranges shrink down the tree, so
            // only the parent's own size test matters; its depth (d - 1 <= 30) is
            // below the cap.
            bool muted = false;
            if (k != 0) {
                PhrLink pl = links[parent[k]];
                muted = pl.last - pl.first < leaf;
            }
            if (!muted) {
                if (k + 1 < n) {
                    PhrLink l = links[k];
                    listed = (l.last - l.first < leaf) || (d + 1 >= PHR_STACK);
                } else {
                    listed = true;
                }
                depth = d;
            }
        }
    }
    // Every reachable node has a listed leaf below it at least as deep, so the
    // deepest listed leaf is the tree's depth.
    uint64_t m = __ballot(listed);
    if (m) {
        uint32_t lane = __lane_id();
        uint32_t first_lane = (uint32_t)__ffsll((unsigned long long)m) - 1;
        uint32_t base = 0;
        if (lane == first_lane) base = atomicAdd(leaf_count, (uint32_t)__popcll(m));
        base = (uint32_t)__shfl((int)base, (int)first_lane);
        if (listed) leaf_list[base + (uint32_t)__popcll(m & lanemask_lt())] = k;
        uint32_t dm = wave_max_u(listed ? depth : 0u);
        if (lane == first_lane) atomicMax(depth_out, dm);
    }
}

// ---- 6. refit ------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_refit(const float4* __restrict__ prims, const uint32_t* __restrict__ order,
                                               const PhrLink* __restrict__ links, const uint32_t* __restrict__ parent,
                                               const uint32_t* __restrict__ leaf_list, const uint32_t* __restrict__ leaf_count,
                                               uint32_t n, float4* nodes, uint32_t* cnt) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= *leaf_count) return;
    const uint32_t k = leaf_list[j];
    uint32_t first, end;
    if (k + 1 < n) {
        PhrLink l = links[k];
        first = l.first;
        end = l.last + 1;
    } else {
        first = k - (n - 1);
        end = first + 1;
    }
    float lo0 = 3.40282347e38f, lo1 = 3.40282347e38f, lo2 = 3.40282347e38f;
    float hi0 = -3.40282347e38f, hi1 = -3.40282347e38f, hi2 = -3.40282347e38f;
    for (uint32_t s = first; s < end; ++s) {
        uint32_t p = order[s];
        float4 a = prims[2 * p], b = prims[2 * p + 1];
        lo0 = fminf(lo0, a.x); lo1 = fminf(lo1, a.y); lo2 = fminf(lo2, a.z);
        hi0 = fmaxf(hi0, b.x); hi1 = fmaxf(hi1, b.y); hi2 = fmaxf(hi2, b.z);
    }
    nodes[2 * k] = make_float4(lo0, lo1, lo2, __uint_as_float(first | PHR_LEAF));
    nodes[2 * k + 1] = make_float4(hi0, hi1, hi2, __uint_as_float(end));
    uint32_t x = k;
    while (x != 0) {
        uint32_t p = parent[x];
        uint32_t old = __hip_atomic_fetch_add(&cnt[p], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        if (old == 0u) return;  // the sibling's thread finishes p
        __hip_atomic_store(&cnt[p], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        PhrLink l = links[p];
        uint32_t sib = (l.left == x) ? l.right : l.left;
        float4 sa = nodes[2 * sib], sb = nodes[2 * sib + 1];
        lo0 = fminf(lo0, sa.x); lo1 = fminf(lo1, sa.y); lo2 = fminf(lo2, sa.z);
        hi0 = fmaxf(hi0, sb.x); hi1 = fmaxf(hi1, sb.y); hi2 = fmaxf(hi2, sb.z);
        nodes[2 * p] = make_float4(lo0, lo1, lo2, __uint_as_float(l.left));
        nodes[2 * p + 1] = make_float4(hi0, hi1, hi2, __uint_as_float(l.right));
        x = p;
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// GpuBvh (src/bvh/gpu.rs:11-104)
// ---------------------------------------------------------------------------
struct phr_bvh {
    phr_gpu* gpu = nullptr;
    uint32_t n = 0, leaf = 4, depth = 0, tiles = 0, bounds_blocks = 0;
    phr_array* prims = nullptr;
    bool own_prims = false;
    phr_array* order = nullptr;  // n u32: sorted slot -> primitive
    phr_array* nodes = nullptr;  // 2n-1 nodes
    uint32_t* codes = nullptr;   // n: unsorted codes, then sorted codes after the last pass
    uint32_t* keys_b = nullptr;
    uint32_t* vals_b = nullptr;
    uint32_t* zero = nullptr;  // histograms, tickets, counters, look-back status
    uint32_t zero_words = 0;
    uint32_t* arrive = nullptr;
    float* scene = nullptr;
    float* partials = nullptr;
    PhrLink* links = nullptr;
    uint32_t* parent = nullptr;
    uint32_t* cnt = nullptr;
    uint32_t* leaf_list = nullptr;
};

namespace {

void bvh_release(phr_bvh* b) {
    if (!b) return;
    if (b->own_prims) phr_array_release(b->prims);
    phr_array_release(b->order);
    phr_array_release(b->nodes);
    for (void* p : {(void*)b->codes, (void*)b->keys_b, (void*)b->vals_b, (void*)b->zero, (void*)b->arrive, (void*)b->scene,
                    (void*)b->partials, (void*)b->links, (void*)b->parent, (void*)b->cnt, (void*)b->leaf_list})
        if (p) (void)hipFree(p);
    delete b;
}

int bvh_alloc(phr_gpu* g, phr_array* prims, bool own, uint32_t leaf, phr_bvh** out) {
    uint32_t n = (uint32_t)prims->len;
    PHR_CHECK(n < (1u << 30), PHR_ERR_SIZE, "GpuBvh: at most 2^30 - 1 primitives");
    PHR_CHECK(leaf >= 1, PHR_ERR_ARG, "GpuBvh: leaf size must be at least 1");
    phr_bvh* b = new phr_bvh();
    b->gpu = g;
    b->n = n;
    b->leaf = leaf;
    b->prims = prims;
    b->own_prims = own;
    b->tiles = (n + OS_TILE - 1) / OS_TILE;
    b->bounds_blocks = std::min<uint32_t>(BOUNDS_BLOCKS_MAX, std::max<uint32_t>(1, (n + 2047) / 2048));
    b->zero_words = Z_STATUS + OS_PASSES * b->tiles * 256;
    int rc = PHR_OK;
    auto fail = [&](int code) { bvh_release(b); return code; };
    if ((rc = phr_array_alloc(g, 4, n, &b->order)) != PHR_OK) return fail(rc);
    if ((rc = phr_array_alloc(g, 32, n ? 2ull * n - 1 : 0, &b->nodes)) != PHR_OK) return fail(rc);
    uint64_t nn = std::max<uint64_t>(n, 1);
    auto dev = [&](void** p, uint64_t bytes) -> bool {
        if (hipMalloc(p, std::max<uint64_t>(bytes, 16)) != hipSuccess) return false;
        return hipMemsetAsync(*p, 0, std::max<uint64_t>(bytes, 16), g->stream) == hipSuccess;
    };
    bool ok = dev((void**)&b->codes, 4 * nn) && dev((void**)&b->keys_b, 4 * nn) && dev((void**)&b->vals_b, 4 * nn) &&
              dev((void**)&b->zero, 4ull * b->zero_words) && dev((void**)&b->arrive, 64) && dev((void**)&b->scene, 32) &&
              dev((void**)&b->partials, 32ull * b->bounds_blocks) && dev((void**)&b->links, 16 * nn) &&
              dev((void**)&b->parent, 8 * nn) && dev((void**)&b->cnt, 4 * nn) && dev((void**)&b->leaf_list, 4 * nn);
    if (!ok) {
        phr_set_error_msg("GpuBvh: device allocation failed");
        return fail(PHR_ERR_HIP);
    }
    *out = b;
    return PHR_OK;
}

// Enqueue the whole build on the device's stream (no host synchronisation).
int bvh_enqueue_build(phr_bvh* b) {
    const uint32_t n = b->n;
    if (n == 0) return PHR_OK;
    hipStream_t s = b->gpu->stream;
    const float4* prims = (const float4*)b->prims->ptr;
    uint32_t* order = (uint32_t*)b->order->ptr;
    float4* nodes = (float4*)b->nodes->ptr;
    uint32_t* hist = b->zero + Z_HIST;
    uint32_t* ticket = b->zero + Z_TICKET;
    uint32_t* status = b->zero + Z_STATUS;
    hipLaunchKernelGGL(k_bounds, dim3(b->bounds_blocks), dim3(256), 0, s, prims, n, b->partials, b->scene, b->arrive, b->zero,
                       b->zero_words);
    hipLaunchKernelGGL(k_morton, dim3(b->tiles), dim3(256), 0, s, prims, (const float*)b->scene, n, b->codes, hist);
    // Pass 0: codes (slot order) -> keys_b/vals_b; 1: -> codes/order; 2: -> b; 3: -> codes/order.
    uint32_t* kin = b->codes;
    uint32_t* vin = nullptr;
    for (uint32_t p = 0; p < OS_PASSES; ++p) {
        bool to_b = (p % 2) == 0;
        uint32_t* kout = to_b ? b->keys_b : b->codes;
        uint32_t* vout = to_b ? b->vals_b : order;
        if (p == 0)
            hipLaunchKernelGGL(k_onesweep<true>, dim3(b->tiles), dim3(OS_THREADS), 0, s, kin, vin, kout, vout, hist + 256 * p,
                               status + (size_t)p * b->tiles * 256, ticket + p, n, 8 * p);
        else
            hipLaunchKernelGGL(k_onesweep<false>, dim3(b->tiles), dim3(OS_THREADS), 0, s, kin, vin, kout, vout, hist + 256 * p,
                               status + (size_t)p * b->tiles * 256, ticket + p, n, 8 * p);
        kin = kout;
        vin = vout;
    }
    if (n > 1) hipLaunchKernelGGL(k_karras, dim3((n - 1 + 255) / 256), dim3(256), 0, s, b->codes, n, b->links, b->parent);
    uint32_t total = 2 * n - 1;
    hipLaunchKernelGGL(k_pack, dim3((total + 255) / 256), dim3(256), 0, s, b->links, b->parent, n, b->leaf, b->leaf_list,
                       b->zero + Z_LEAVES, b->zero + Z_DEPTH);
    hipLaunchKernelGGL(k_refit, dim3((n + 255) / 256), dim3(256), 0, s, prims, order, b->links, b->parent, b->leaf_list,
                       b->zero + Z_LEAVES, n, nodes, b->cnt);
    PHR_HIP(hipGetLastError());
    return PHR_OK;
}

int bvh_enqueue_refit(phr_bvh* b) {
    const uint32_t n = b->n;
    if (n == 0) return PHR_OK;
    hipLaunchKernelGGL(k_refit, dim3((n + 255) / 256), dim3(256), 0, b->gpu->stream, (const float4*)b->prims->ptr,
                       (const uint32_t*)b->order->ptr, b->links, b->parent, b->leaf_list, b->zero + Z_LEAVES, n,
                       (float4*)b->nodes->ptr, b->cnt);
    PHR_HIP(hipGetLastError());
    return PHR_OK;
}

int bvh_read_depth(phr_bvh* b) {
    if (b->n == 0) { b->depth = 0; return PHR_OK; }
    uint32_t d = 0;
    PHR_HIP(hipMemcpyAsync(&d, b->zero + Z_DEPTH, 4, hipMemcpyDeviceToHost, b->gpu->stream));
    PHR_HIP(hipStreamSynchronize(b->gpu->stream));
    b->depth = d;
    return PHR_OK;
}

}  // namespace

// Internal entry points for the tracer (trace.hip).
int phr_bvh_internal_build(phr_gpu* g, phr_array* prims, uint32_t leaf, phr_bvh** out) {
    phr_bvh* b = nullptr;
    int rc = bvh_alloc(g, prims, false, leaf, &b);
    if (rc != PHR_OK) return rc;
    if ((rc = bvh_enqueue_build(b)) != PHR_OK || (rc = bvh_read_depth(b)) != PHR_OK) {
        bvh_release(b);
        return rc;
    }
    *out = b;
    return PHR_OK;
}
phr_array* phr_bvh_internal_nodes(phr_bvh* b) { return b->nodes; }
phr_array* phr_bvh_internal_order(phr_bvh* b) { return b->order; }

extern "C" {

int phr_bvh_build(phr_gpu* g, const void* prims, uint32_t n, uint32_t leaf, phr_bvh** out) {
    PHR_CHECK(g && out && (prims || n == 0), PHR_ERR_ARG, "phr_bvh_build: null argument");
    phr_array* a = nullptr;
    int rc = phr_array_alloc(g, 32, n, &a);
    if (rc != PHR_OK) return rc;
    if (n) {
        hipError_t e = hipMemcpyAsync(a->ptr, prims, 32ull * n, hipMemcpyHostToDevice, g->stream);
        if (e != hipSuccess) { phr_array_release(a); phr_set_error("hipMemcpyAsync", e); return PHR_ERR_HIP; }
    }
    phr_bvh* b = nullptr;
    if ((rc = bvh_alloc(g, a, true, leaf, &b)) != PHR_OK) { phr_array_release(a); return rc; }
    if ((rc = bvh_enqueue_build(b)) != PHR_OK || (rc = bvh_read_depth(b)) != PHR_OK) { bvh_release(b); return rc; }
    *out = b;
    return PHR_OK;
}

int phr_bvh_build_on(phr_gpu* g, phr_array* prims, uint32_t leaf, phr_bvh** out) {
    PHR_CHECK(g && prims && out && prims->elem == 32, PHR_ERR_ARG, "phr_bvh_build_on: prims must be an array of 32-byte boxes");
    phr_bvh* b = nullptr;
    int rc = bvh_alloc(g, prims, false, leaf, &b);
    if (rc != PHR_OK) return rc;
    if ((rc = bvh_enqueue_build(b)) != PHR_OK || (rc = bvh_read_depth(b)) != PHR_OK) { bvh_release(b); return rc; }
    *out = b;
    return PHR_OK;
}

int phr_bvh_rebuild_async(phr_bvh* b) {
    PHR_CHECK(b, PHR_ERR_ARG, "phr_bvh_rebuild_async: null");
    return bvh_enqueue_build(b);
}

int phr_bvh_rebuild(phr_bvh* b) {
    PHR_CHECK(b, PHR_ERR_ARG, "phr_bvh_rebuild: null");
    int rc = bvh_enqueue_build(b);
    if (rc != PHR_OK) return rc;
    return bvh_read_depth(b);
}

int phr_bvh_refit_async(phr_bvh* b) {
    PHR_CHECK(b, PHR_ERR_ARG, "phr_bvh_refit_async: null");
    return bvh_enqueue_refit(b);
}

int phr_bvh_refit(phr_bvh* b) {
    PHR_CHECK(b, PHR_ERR_ARG, "phr_bvh_refit: null");
    int rc = bvh_enqueue_refit(b);
    if (rc != PHR_OK) return rc;
    PHR_HIP(hipStreamSynchronize(b->gpu->stream));
    return PHR_OK;
}

int phr_bvh_update(phr_bvh* b, const void* prims, uint32_t n, int* refitted) {
    PHR_CHECK(b && refitted && (prims || n == 0), PHR_ERR_ARG, "phr_bvh_update: null argument");
    // src/bvh/gpu.rs:93-99: a new count means rebuild instead.
    if (n != b->n) { *refitted = 0; return PHR_OK; }
    if (n) PHR_HIP(hipMemcpyAsync(b->prims->ptr, prims, 32ull * n, hipMemcpyHostToDevice, b->gpu->stream));
    int rc = bvh_enqueue_refit(b);
    if (rc != PHR_OK) return rc;
    PHR_HIP(hipStreamSynchronize(b->gpu->stream));
    *refitted = 1;
    return PHR_OK;
}

int phr_bvh_download(phr_bvh* b, void* nodes, uint32_t* order, uint32_t* depth) {
    PHR_CHECK(b, PHR_ERR_ARG, "phr_bvh_download: null");
    hipStream_t s = b->gpu->stream;
    if (b->n) {
        if (nodes) PHR_HIP(hipMemcpyAsync(nodes, b->nodes->ptr, 32ull * (2ull * b->n - 1), hipMemcpyDeviceToHost, s));
        if (order) PHR_HIP(hipMemcpyAsync(order, b->order->ptr, 4ull * b->n, hipMemcpyDeviceToHost, s));
    }
    PHR_HIP(hipStreamSynchronize(s));
    if (depth) *depth = b->depth;
    return PHR_OK;
}

int phr_bvh_codes(phr_bvh* b, uint32_t* sorted_codes) {
    PHR_CHECK(b && sorted_codes, PHR_ERR_ARG, "phr_bvh_codes: null");
    if (b->n) PHR_HIP(hipMemcpyAsync(sorted_codes, b->codes, 4ull * b->n, hipMemcpyDeviceToHost, b->gpu->stream));
    PHR_HIP(hipStreamSynchronize(b->gpu->stream));
    return PHR_OK;
}

uint32_t phr_bvh_len(const phr_bvh* b) { return b ? b->n : 0; }
uint32_t phr_bvh_depth(const phr_bvh* b) { return b ? b->depth : 0; }
uint32_t phr_bvh_leaf(const phr_bvh* b) { return b ? b->leaf : 0; }
phr_array* phr_bvh_nodes(phr_bvh* b) { return b ? b->nodes : nullptr; }
phr_array* phr_bvh_order(phr_bvh* b) { return b ? b->order : nullptr; }
phr_array* phr_bvh_prims(phr_bvh* b) { return b ? b->prims : nullptr; }
void phr_bvh_free(phr_bvh* b) { bvh_release(b); }

}  // extern "C"

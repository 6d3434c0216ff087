// GPU LBVH for gfx950: scene bounds, 30-bit Morton keys, a wave64 onesweep
// radix sort of (code, slot), Karras links, leaf packing and a bottom-up
// refit into phiray's 32-byte Node layout.
//
// The result is the tree phiray's CPU builder makes (src/bvh/build.rs:94-100,
// Bvh::build): the same order, links, depth and reachable-node bounds. One
// build may hold many trees ("segments"): the boxes come concatenated, a
// table of segment starts splits them, and every segment gets the tree
// Bvh::build would give its boxes alone, with its own scene box. That is one
// tree per mesh of a Forest (src/trace/forest.rs:44-66) or a batch of
// independent scenes, in the same launches as a single tree.
//
// Ids. Slots are positions in the sorted order, global over the segments.
// The inner node of a segment's Karras tree numbered i (src/bvh/build.rs:
// 28-61, one end of its range is slot i) has id `lo + i`, where lo is the
// segment's first slot; its leaf k has id `n + lo + k`. So an inner node's id
// is a slot inside its own range, whatever segment it lies in. The node array
// written out keeps phiray's layout per tree: segment s's tree starts at a
// caller-given `node_base[s]`, inner node i at node_base + i, leaf k at
// node_base + n_s - 1 + k, children as node indices and leaf ranges as
// global slots (what Forest::nodes makes of the mesh trees).
//
// The stages and why each looks the way it does:
//
//  1. bounds   each wave takes a contiguous run of boxes and keeps a running
//              box while the segment stays the same; a segment that ends in
//              the run is folded into its six words with atomic min/max on an
//              order-preserving integer image of the floats, and the box of
//              the last segment a block reaches goes to that block's own slot,
//              folded by a one-block launch (so one tree's boxes meet no
//              contended atomic).
//              No block waits on another and no cache-wide fence is needed
//              (min and max are exact, so the fold order does not matter).
//              The Morton launch reads the words back and forms each scene
//              box and its per-axis reciprocal (src/bvh/build.rs:70-80). The
//              same launch clears the sort's counters.
//  2. morton   the code of every box centre (src/bvh/node.rs:96-105) and, in
//              the same pass, the histograms of every digit the sort uses.
//  3. sort     onesweep LSD passes over the code (bits 0-7, 8-15, 16-23,
//              24-29), then over the segment id when there is more than one:
//              a tile of 2048 keys ranks its digits with wave64 ballots into
//              per-wave LDS counters, publishes its per-digit counts and looks
//              back over earlier tiles for its offsets (tiles take ids from a
//              ticket so an earlier tile is always resident), then scatters
//              through LDS so each digit's run leaves in one contiguous write.
//              Every pass is stable and the first starts from the input order,
//              so the result is the sort by (segment, code, index) — within a
//              segment the sort by (code, slot) of src/bvh/build.rs:87-92.
//  4. karras   one thread per slot: direction, exponential and binary search,
//              split by clz of XOR'd keys within the slot's segment
//              (src/bvh/build.rs:28-61), and the parent of each child.
//  5. pack     every node's depth without walking the tree: the ranges of a
//              tree's inner nodes nest, so the inner ranges covering slot k
//              are exactly leaf k's ancestors, and those covering the edge
//              between slots g and g + 1 are exactly the inner node split at g
//              and its ancestors. Every inner range adds +1 where it starts
//              and -1 where it ends; one prefix sum of those (a chained scan)
//              counts the ranges covering each edge (a segment's terms sum to
//              zero, so one scan serves all of them). An inner node's depth is
//              its split edge's count less one; a leaf's is the larger count
//              of the two edges beside it (its parent is the deeper of the two
//              nodes split there). One end of a Karras node's range is its own
//              slot, so the Karras launch stores that term and adds only the
//              far end's with an atomic. Then one thread per node: a subtree of
//              fewer than `leaf + 1` slots or at depth 31 becomes a leaf
//              (src/bvh/build.rs:139-168), whatever lies under a packed node is
//              unreachable, and each reachable leaf is entered in a slot table
//              under the first slot of its range.
//  6. refit    bottom-up (src/bvh/build.rs:180-194) in levels over ranges of
//              slots. A level-0 block owns 2048 slots: one thread per leaf
//              starting there computes the leaf box from its primitives and
//              climbs; at an inner node the second of its two children to
//              arrive (a workgroup-scope counter in LDS) unions both and goes
//              on. A climb stops at a node whose parent's range leaves the
//              block; that node is handed to the next level, whose blocks own
//              64 times as many slots and carry those climbs on the same way,
//              until one block owns every slot. Every inner node's range is
//              inside the block that finishes it, so all synchronisation stays
//              inside one workgroup: no agent-scope acquire or release, which
//              on gfx950 writes back and invalidates the XCD's L2 at every
//              step of every climb.
#include "phr_runtime.h"

#include <algorithm>
#include <cstring>
#include <mutex>
#include <vector>

namespace {

constexpr uint32_t OS_THREADS = 256;
constexpr uint32_t OS_ITEMS = 8;
constexpr uint32_t OS_TILE = OS_THREADS * OS_ITEMS;  // 2048 keys a tile
constexpr uint32_t OS_WAVES = OS_THREADS / 64;
constexpr uint32_t CODE_PASSES = 4;
constexpr uint32_t MAX_PASSES = 6;  // four over the code, up to two over the segment id
constexpr uint32_t MAX_SEGMENTS = 1u << 16;
constexpr uint32_t FLAG_AGG = 0x40000000u;
constexpr uint32_t FLAG_PRE = 0x80000000u;
constexpr uint32_t VAL_MASK = 0x3fffffffu;
constexpr uint32_t BOUNDS_BLOCKS_MAX = 512;
constexpr uint32_t KEY_MAX = 0xff7fffffu;  // fkey(f32::MAX)
constexpr uint32_t KEY_MIN = 0x00800000u;  // fkey(f32::MIN)

// Words at the front of the cleared block: MAX_PASSES x 256 digit counts, the
// pass tickets, the scan's ticket and one depth per segment; the look-back
// status follows.
constexpr uint32_t Z_HIST = 0;
constexpr uint32_t Z_TICKET = MAX_PASSES * 256;
constexpr uint32_t Z_SCAN_TICKET = Z_TICKET + 8;
constexpr uint32_t Z_DEPTH = Z_SCAN_TICKET + 8;

// Refit levels: slots owned by a level-0 block, and blocks of one level per
// block of the next. A level's inner-node counters live in LDS (2 bits each)
// while its block owns at most RF_LDS_SLOTS slots, in global memory above.
constexpr uint32_t RF_THREADS = 1024;  // every handed-over climb of a block in flight at once
constexpr uint32_t RF_LEAF_THREADS = 1024;  // level 0 is latency-bound: more waves in flight
constexpr uint32_t RF_SLOTS = 2048;
constexpr uint32_t RF_GROUP = 64;
constexpr uint32_t RF_LDS_SLOTS = RF_SLOTS * RF_GROUP;
constexpr uint32_t UP_LEFT = 0x80000000u;  // in an up record's sibling word: the node is the left child

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

// The segment holding slot (or input index) p: the last s with seg_lo[s] <= p.
// An empty segment shares its start with the next, so the search lands on the
// segment that holds p.
__device__ __forceinline__ uint32_t seg_of_slot(const uint32_t* __restrict__ seg_lo, uint32_t segs, uint32_t p) {
    uint32_t lo = 0, hi = segs;  // seg_lo[lo] <= p < seg_lo[hi]
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (seg_lo[mid] <= p) lo = mid; else hi = mid;
    }
    return lo;
}

// A segment as the per-node kernels need it.
struct Seg {
    uint32_t lo, n, base;
};
__device__ __forceinline__ Seg seg_at(const uint32_t* __restrict__ seg_lo, const uint32_t* __restrict__ node_base, uint32_t s) {
    return Seg{seg_lo[s], seg_lo[s + 1] - seg_lo[s], node_base[s]};
}
// Where node `id` (inner: lo + i; leaf: n + lo + k) lands in the node array.
__device__ __forceinline__ uint32_t out_of(uint32_t id, uint32_t n, const Seg& g) {
    return id < n ? g.base + (id - g.lo) : g.base + (g.n - 1) + (id - n - g.lo);
}

// ---- 1. scene bounds ------------------------------------------------------
// Order-preserving integer image of a float: for non-NaN a, b,
// a < b exactly when fkey(a) < fkey(b) as unsigned integers.
__device__ __forceinline__ uint32_t fkey(float f) {
    uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float unfkey(uint32_t k) { return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k); }

struct Box {
    float lo0, lo1, lo2, hi0, hi1, hi2;
};
__device__ __forceinline__ Box empty_box() {
    // Aabb::default() is (f32::MAX, f32::MIN), src/bvh/node.rs:67-71.
    return Box{3.40282347e38f, 3.40282347e38f, 3.40282347e38f, -3.40282347e38f, -3.40282347e38f, -3.40282347e38f};
}
__device__ __forceinline__ void grow(Box& b, float4 lo, float4 hi) {
    b.lo0 = fminf(b.lo0, lo.x); b.lo1 = fminf(b.lo1, lo.y); b.lo2 = fminf(b.lo2, lo.z);
    b.hi0 = fmaxf(b.hi0, hi.x); b.hi1 = fmaxf(b.hi1, hi.y); b.hi2 = fmaxf(b.hi2, hi.z);
}

__device__ __forceinline__ void fold_keys(uint32_t* __restrict__ acc, uint32_t s, const Box& b) {
    uint32_t* a = acc + 6 * s;
    atomicMin(&a[0], fkey(b.lo0)); atomicMin(&a[1], fkey(b.lo1)); atomicMin(&a[2], fkey(b.lo2));
    atomicMax(&a[3], fkey(b.hi0)); atomicMax(&a[4], fkey(b.hi1)); atomicMax(&a[5], fkey(b.hi2));
}
__device__ __forceinline__ Box wave_box(Box b) {
    b.lo0 = wave_min(b.lo0); b.lo1 = wave_min(b.lo1); b.lo2 = wave_min(b.lo2);
    b.hi0 = wave_max(b.hi0); b.hi1 = wave_max(b.hi1); b.hi2 = wave_max(b.hi2);
    return b;
}
__device__ __forceinline__ Box join(const Box& a, const Box& b) {
    return Box{fminf(a.lo0, b.lo0), fminf(a.lo1, b.lo1), fminf(a.lo2, b.lo2),
               fmaxf(a.hi0, b.hi0), fmaxf(a.hi1, b.hi1), fmaxf(a.hi2, b.hi2)};
}

// A block's box of the last segment its boxes reach, for the fold launch.
struct Part {
    uint32_t seg;
    float lo0, lo1, lo2, hi0, hi1, hi2, pad;
};

// acc[6s..6s+2] = fkey of segment s's low corner, acc[6s+3..6s+5] of its high
// corner. They hold fkey(f32::MAX) and fkey(f32::MIN) between builds; the
// packing launch restores them after the Morton launch has read them.
//
// Each block takes a contiguous run of boxes, a quarter to each wave. A
// segment that ends inside a wave's run is folded into its words at once
// (few blocks touch it); the box of the last segment the block reaches goes
// to the block's own Part slot, so the long runs of one segment - one tree's
// boxes - meet no contended atomic here. k_bounds_fold folds the Parts.
__global__ __launch_bounds__(256) void k_bounds(const float4* __restrict__ prims, uint32_t n, const uint32_t* __restrict__ seg_lo,
                                                uint32_t segs, uint32_t* __restrict__ acc, Part* __restrict__ parts,
                                                uint32_t* __restrict__ zero, uint32_t zero_words) {
    __shared__ uint32_t s_seg[4];
    __shared__ Box s_box[4];
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint32_t gtid = blockIdx.x * blockDim.x + tid, gsize = gridDim.x * blockDim.x;
    for (uint32_t i = gtid; i < zero_words; i += gsize) zero[i] = 0u;
    const uint32_t bchunk = ((n + gridDim.x - 1) / gridDim.x + 255) & ~255u, wchunk = bchunk / 4;
    const uint32_t b0 = min(blockIdx.x * bchunk, n), b1 = min(b0 + bchunk, n);
    const uint32_t c0 = min(b0 + wave * wchunk, b1), c1 = min(c0 + wchunk, b1);
    uint32_t cur = PHR_NONE;
    Box b = empty_box();
    if (c0 < c1) {
        cur = seg_of_slot(seg_lo, segs, c0);
        uint32_t cur_end = seg_lo[cur + 1];
        for (uint32_t base = c0; base < c1; base += 64) {
            const uint32_t p = base + lane;
            const bool active = p < c1;
            float4 lo = make_float4(0.f, 0.f, 0.f, 0.f), hi = lo;
            if (active) { lo = prims[2 * p]; hi = prims[2 * p + 1]; }
            if (min(base + 64, c1) <= cur_end) {
                if (active) grow(b, lo, hi);
                continue;
            }
            // A segment ends inside this run of 64: fold each as it closes.
            bool done = !active;
            for (;;) {
                const bool mine = !done && p < cur_end;
                if (mine) grow(b, lo, hi);
                done = done || mine;
                const uint64_t left = __ballot(!done);
                if (left == 0) break;
                const Box w = wave_box(b);
                if (lane == 0) fold_keys(acc, cur, w);
                b = empty_box();
                const uint32_t first = base + (uint32_t)__ffsll((unsigned long long)left) - 1;
                cur = seg_of_slot(seg_lo, segs, first);
                cur_end = seg_lo[cur + 1];
            }
        }
        b = wave_box(b);
    }
    if (lane == 0) {
        s_seg[wave] = cur;
        s_box[wave] = b;
    }
    __syncthreads();
    if (tid == 0) {
        uint32_t last = PHR_NONE;
        for (int w = 3; w >= 0 && last == PHR_NONE; --w) last = s_seg[w];
        Box part = empty_box();
        for (int w = 0; w < 4; ++w) {
            if (s_seg[w] == PHR_NONE) continue;
            if (s_seg[w] == last) part = join(part, s_box[w]);
            else fold_keys(acc, s_seg[w], s_box[w]);  // a segment that ended inside this block
        }
        parts[blockIdx.x] = Part{last, part.lo0, part.lo1, part.lo2, part.hi0, part.hi1, part.hi2, 0.f};
    }
}

// One block folds the blocks' Parts into the segments' words. Parts are in
// slot order, so a thread's strided Parts have non-decreasing segments: it
// keeps a running box and folds when the segment changes; a wave whose
// threads all end on one segment folds once.
__global__ __launch_bounds__(256) void k_bounds_fold(const Part* __restrict__ parts, uint32_t nparts, uint32_t* __restrict__ acc) {
    const uint32_t tid = threadIdx.x, lane = tid & 63;
    uint32_t cur = PHR_NONE;
    Box b = empty_box();
    for (uint32_t k = tid; k < nparts; k += blockDim.x) {
        const Part pt = parts[k];
        if (pt.seg == PHR_NONE) continue;
        if (pt.seg != cur) {
            if (cur != PHR_NONE) fold_keys(acc, cur, b);
            cur = pt.seg;
            b = empty_box();
        }
        b = join(b, Box{pt.lo0, pt.lo1, pt.lo2, pt.hi0, pt.hi1, pt.hi2});
    }
    const uint32_t first = (uint32_t)__shfl((int)cur, 0);
    const bool same = __ballot(cur != first && cur != PHR_NONE) == 0;
    if (same) {
        const Box w = wave_box(b);
        if (lane == 0 && first != PHR_NONE) fold_keys(acc, first, w);
    } else if (cur != PHR_NONE) {
        fold_keys(acc, cur, b);
    }
}

// ---- 2. Morton codes + the digit histograms --------------------------------
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

__device__ __forceinline__ float recip_extent(float lo, float hi) {
    float e = hi - lo;
    return e > 0.f ? 1.0f / e : 0.f;  // src/bvh/build.rs:76-80
}

__global__ __launch_bounds__(256) void k_morton(const float4* __restrict__ prims, const uint32_t* __restrict__ acc,
                                                const uint32_t* __restrict__ seg_lo, uint32_t segs, uint32_t seg_passes, uint32_t n,
                                                uint32_t* __restrict__ codes, uint32_t* __restrict__ hist,
                                                uint32_t* __restrict__ far_ends) {
    __shared__ uint32_t s_h[MAX_PASSES * 256];
    const uint32_t tid = threadIdx.x;
    for (uint32_t i = tid; i < MAX_PASSES * 256; i += 256) s_h[i] = 0u;
    __syncthreads();
    const uint32_t base = blockIdx.x * OS_TILE;
    uint32_t s = base < n ? seg_of_slot(seg_lo, segs, base) : 0u;
    uint32_t s_end = seg_lo[s + 1];
    float slo0 = 0.f, slo1 = 0.f, slo2 = 0.f, inv0 = 0.f, inv1 = 0.f, inv2 = 0.f;
    uint32_t have = PHR_NONE;
#pragma unroll
    for (uint32_t r = 0; r < OS_ITEMS; ++r) {
        uint32_t i = base + r * 256 + tid;
        if (i < n) {
            while (i >= s_end) {  // runs of slots only move forward
                ++s;
                s_end = seg_lo[s + 1];
            }
            if (have != s) {
                const uint32_t* a = acc + 6 * s;
                slo0 = unfkey(a[0]); slo1 = unfkey(a[1]); slo2 = unfkey(a[2]);
                inv0 = recip_extent(slo0, unfkey(a[3]));
                inv1 = recip_extent(slo1, unfkey(a[4]));
                inv2 = recip_extent(slo2, unfkey(a[5]));
                have = s;
            }
            float4 lo = prims[2 * i], hi = prims[2 * i + 1];
            // Aabb::center then (c - s.lo) * inv, src/bvh/build.rs:81-84
            float c0 = 0.5f * (lo.x + hi.x), c1 = 0.5f * (lo.y + hi.y), c2 = 0.5f * (lo.z + hi.z);
            float u0 = (c0 - slo0) * inv0, u1 = (c1 - slo1) * inv1, u2 = (c2 - slo2) * inv2;
            uint32_t code = (spread10(u0) << 2) | (spread10(u1) << 1) | spread10(u2);
            codes[i] = code;
            far_ends[i] = 0u;
            atomicAdd(&s_h[code & 0xFFu], 1u);
            atomicAdd(&s_h[256 + ((code >> 8) & 0xFFu)], 1u);
            atomicAdd(&s_h[512 + ((code >> 16) & 0xFFu)], 1u);
            atomicAdd(&s_h[768 + (code >> 24)], 1u);
            if (seg_passes) {
                atomicAdd(&s_h[1024 + (s & 0xFFu)], 1u);
                if (seg_passes > 1) atomicAdd(&s_h[1280 + ((s >> 8) & 0xFFu)], 1u);
            }
        }
    }
    __syncthreads();
    for (uint32_t i = tid; i < MAX_PASSES * 256; i += 256)
        if (s_h[i]) atomicAdd(&hist[i], s_h[i]);
}

// ---- 3. onesweep pass ------------------------------------------------------
// FIRST: the values are the input indices. SEGD: the digit is a byte of the
// value's segment id instead of a byte of the key (the key rides along).
template <bool FIRST, bool SEGD, uint32_t ITEMS>
__global__ __launch_bounds__(256) void k_onesweep(const uint32_t* __restrict__ kin, const uint32_t* __restrict__ vin,
                                                  uint32_t* __restrict__ kout, uint32_t* __restrict__ vout,
                                                  const uint32_t* __restrict__ hist, uint32_t* status, uint32_t* ticket,
                                                  uint32_t n, uint32_t shift, const uint32_t* __restrict__ seg_lo, uint32_t segs) {
    __shared__ uint32_t s_whist[OS_WAVES][256];
    __shared__ uint32_t s_local[256];
    __shared__ uint32_t s_global[256];
    __shared__ uint32_t s_keys[(OS_THREADS * ITEMS)];
    __shared__ uint32_t s_vals[(OS_THREADS * ITEMS)];
    __shared__ uint8_t s_dig[(OS_THREADS * ITEMS)];
    __shared__ uint32_t s_scan[4];
    __shared__ uint32_t s_tile;
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (tid == 0) s_tile = atomicAdd(ticket, 1u);
    for (uint32_t i = tid; i < OS_WAVES * 256; i += 256) (&s_whist[0][0])[i] = 0u;
    __syncthreads();
    const uint32_t tile = s_tile;
    const uint32_t wbase = tile * (OS_THREADS * ITEMS) + wave * (ITEMS * 64);
    uint32_t key[ITEMS], val[ITEMS], rank[ITEMS], dig[ITEMS];
#pragma unroll
    for (uint32_t r = 0; r < ITEMS; ++r) {
        uint32_t idx = wbase + r * 64 + lane;
        bool ok = idx < n;
        key[r] = ok ? kin[idx] : 0u;
        val[r] = FIRST ? idx : (ok ? vin[idx] : 0u);
        // The segment of input index val: a search of the small segment
        // table (cached), not a gather from an n-word array.
        if constexpr (SEGD) dig[r] = ok ? (seg_of_slot(seg_lo, segs, val[r]) >> shift) & 0xFFu : 0u;
        else dig[r] = (key[r] >> shift) & 0xFFu;
    }
    const uint64_t lt = lanemask_lt();
#pragma unroll
    for (uint32_t r = 0; r < ITEMS; ++r) {
        uint32_t idx = wbase + r * 64 + lane;
        bool ok = idx < n;
        uint32_t d = dig[r];
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
        // Four earlier tiles' words a round trip, taken in order up to the
        // nearest inclusive prefix (before tile 0: a prefix of zero).
        int32_t j = (int32_t)tile - 1;
        for (bool done = false; !done;) {
            uint32_t st[4];
#pragma unroll
            for (int k = 0; k < 4; ++k)
                st[k] = j - k >= 0 ? __hip_atomic_load(status + (size_t)(j - k) * 256 + d, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                                   : FLAG_PRE;
            int k = 0;
            for (; k < 4; ++k) {
                if (st[k] & FLAG_PRE) { excl += st[k] & VAL_MASK; done = true; break; }
                if (!(st[k] & FLAG_AGG)) break;  // not published yet
                excl += st[k] & VAL_MASK;
            }
            if (done) break;
            j -= k;
            if (k < 4) __builtin_amdgcn_s_sleep(1);
        }
        __hip_atomic_store(my, FLAG_PRE | (excl + total), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    s_local[d] = local;
    s_global[d] = binbase + excl;
    __syncthreads();
#pragma unroll
    for (uint32_t r = 0; r < ITEMS; ++r) {
        uint32_t idx = wbase + r * 64 + lane;
        if (idx < n) {
            uint32_t dd = dig[r];
            uint32_t lp = s_local[dd] + s_whist[wave][dd] + rank[r];
            s_keys[lp] = key[r];
            s_vals[lp] = val[r];
            s_dig[lp] = (uint8_t)dd;
        }
    }
    __syncthreads();
    const uint32_t count = min((OS_THREADS * ITEMS), n - tile * (OS_THREADS * ITEMS));
#pragma unroll
    for (uint32_t r = 0; r < ITEMS; ++r) {
        uint32_t idx = r * 256 + tid;
        if (idx < count) {
            uint32_t dd = s_dig[idx];
            uint32_t dest = s_global[dd] + (idx - s_local[dd]);
            kout[dest] = s_keys[idx];
            vout[dest] = s_vals[idx];
        }
    }
}

// ---- 4. Karras links --------------------------------------------------------
// The codes around a block's slots, staged in LDS: most searches stay within
// KW_HALO slots of their own (a low node's range is short), so their
// dependent reads are LDS reads; the rest read the codes array.
constexpr int32_t KW_HALO = 1024;
constexpr int32_t KW_SIZE = 256 + 2 * KW_HALO;

struct CodeWindow {
    const uint32_t* g;  // the codes, global slots
    const uint32_t* s;  // s[k] = g[w0 + k] for k in [0, KW_SIZE)
    int64_t w0;
    __device__ __forceinline__ uint32_t at(int64_t slot) const {
        const int64_t k = slot - w0;
        return (k >= 0 && k < KW_SIZE) ? s[k] : g[slot];
    }
};

// delta (src/bvh/mod.rs:33-43) of local slot i, whose code is ci, against
// local slot j of the segment starting at global slot lo.
__device__ __forceinline__ int32_t wdelta(const CodeWindow& w, int64_t lo, int64_t m, int64_t i, uint32_t ci, int64_t j) {
    if (j < 0 || j >= m) return -1;
    const uint32_t b = w.at(lo + j);
    if (ci != b) return (int32_t)__clz(ci ^ b);
    return 32 + (int32_t)__clz((uint32_t)i ^ (uint32_t)j);
}

__global__ __launch_bounds__(256) void k_karras(const uint32_t* __restrict__ codes, uint32_t n, const uint32_t* __restrict__ seg_lo,
                                                uint32_t segs, PhrLink* __restrict__ links, uint4* __restrict__ up,
                                                uint32_t* __restrict__ leaf_at, uint32_t* __restrict__ own_end,
                                                uint32_t* __restrict__ far_ends) {
    __shared__ uint32_t s_codes[KW_SIZE];
    const int64_t w0 = (int64_t)blockIdx.x * 256 - KW_HALO;
    for (int32_t k = threadIdx.x; k < KW_SIZE; k += 256) {
        const int64_t g = w0 + k;
        s_codes[k] = (g >= 0 && g < (int64_t)n) ? codes[g] : 0u;
    }
    __syncthreads();
    const CodeWindow win{codes, s_codes, w0};
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    leaf_at[p] = PHR_NONE;
    const uint32_t sg = seg_of_slot(seg_lo, segs, p);
    const uint32_t lo = seg_lo[sg], ns = seg_lo[sg + 1] - lo, k = p - lo;
    if (k == 0) up[ns == 1 ? n + p : p] = make_uint4(PHR_NONE, 0u, 0u, 0u);  // the root
    if (k + 1 >= ns) {
        own_end[p] = 0u;  // the last slot of a segment starts no inner node
        return;
    }
    const int64_t m = ns, i = k;
    const uint32_t ci = win.at((int64_t)p);
    auto delta = [&](int64_t j) { return wdelta(win, lo, m, i, ci, j); };
    int64_t d = (delta(i + 1) - delta(i - 1) > 0) ? 1 : -1;
    int32_t dmin = delta(i - d);
    int64_t lmax = 2;
    while (delta(i + lmax * d) > dmin) lmax *= 2;
    int64_t l = 0, t = lmax / 2;
    while (t >= 1) {
        if (delta(i + (l + t) * d) > dmin) l += t;
        t /= 2;
    }
    int64_t j = i + l * d;
    int32_t dnode = delta(j);
    int64_t s = 0;
    t = l;
    for (;;) {
        t = (t + 1) >> 1;
        if (delta(i + (s + t) * d) > dnode) s += t;
        if (t <= 1) break;
    }
    int64_t g = i + s * d + (d < 0 ? d : 0);
    int64_t first = i < j ? i : j, last = i < j ? j : i;
    uint32_t left = (uint32_t)((first == g) ? n + lo + g : lo + g);
    uint32_t right = (uint32_t)((last == g + 1) ? n + lo + g + 1 : lo + g + 1);
    links[p] = PhrLink{left, right, lo + (uint32_t)first, lo + (uint32_t)last};
    // Each child's way up: its parent, the parent's range, its sibling (bit 31
    // set when the child is the left one).
    up[left] = make_uint4(p, lo + (uint32_t)first, lo + (uint32_t)last, right | UP_LEFT);
    up[right] = make_uint4(p, lo + (uint32_t)first, lo + (uint32_t)last, left);
    // +1 where the range starts, -1 where it ends (two's complement words).
    if (first == i) {
        own_end[p] = 1u;
        atomicAdd(&far_ends[lo + last], 0xffffffffu);
    } else {
        own_end[p] = 0xffffffffu;
        atomicAdd(&far_ends[lo + first], 1u);
    }
}

// ---- 5. leaf packing --------------------------------------------------------
// Chained scan over the slots: stab[e] = sum over j <= e of the +1/-1 terms
// (own_end[j] + far_ends[j]), the number of inner ranges covering the edge
// between slots e and e + 1.
// Tiles take ids from a ticket, publish their total and look back over
// earlier tiles (64-bit status words: flags in bits 62-63, a signed total in
// bits 0-61, since a tile may hold more ends than starts).
constexpr uint32_t SC_ITEMS = 8;
constexpr uint32_t SC_TILE = 256 * SC_ITEMS;
constexpr uint64_t SC_AGG = 1ull << 62;
constexpr uint64_t SC_PRE = 1ull << 63;
constexpr uint64_t SC_MASK = SC_AGG - 1;

__device__ __forceinline__ int64_t sc_value(uint64_t s) { return (int64_t)(s << 2) >> 2; }

__global__ __launch_bounds__(256) void k_stab(const uint32_t* __restrict__ own_end, const uint32_t* __restrict__ far_ends, uint32_t n,
                                              uint32_t* __restrict__ stab, uint64_t* status, uint32_t* ticket) {
    __shared__ uint32_t s_tile;
    __shared__ int64_t s_wave[4];
    __shared__ int64_t s_excl;
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (tid == 0) s_tile = atomicAdd(ticket, 1u);
    __syncthreads();
    const uint32_t tile = s_tile;
    const uint32_t base = tile * SC_TILE + tid * SC_ITEMS;
    int32_t v[SC_ITEMS];
    int64_t mine = 0;
#pragma unroll
    for (uint32_t r = 0; r < SC_ITEMS; ++r) {
        uint32_t i = base + r;
        v[r] = i < n ? (int32_t)(own_end[i] + far_ends[i]) : 0;
        mine += v[r];
    }
    int64_t x = mine;
    for (int o = 1; o < 64; o <<= 1) {
        int64_t y = __shfl_up(x, o);
        if (lane >= (uint32_t)o) x += y;
    }
    if (lane == 63) s_wave[wave] = x;
    __syncthreads();
    int64_t before = 0, total = 0;
    for (uint32_t w = 0; w < 4; ++w) {
        if (w < wave) before += s_wave[w];
        total += s_wave[w];
    }
    if (wave == 0) {
        // The look-back by the first wave: 64 earlier tiles' words a read,
        // summed up to the nearest inclusive prefix once every word up to it
        // is published.
        uint64_t* my = status + tile;
        int64_t excl = 0;
        if (tile == 0) {
            if (lane == 0) __hip_atomic_store(my, SC_PRE | ((uint64_t)total & SC_MASK), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            if (lane == 0) __hip_atomic_store(my, SC_AGG | ((uint64_t)total & SC_MASK), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            int64_t j0 = (int64_t)tile - 1;  // the window's nearest tile, at lane 0
            for (;;) {
                const int64_t j = j0 - (int64_t)lane;
                // Before tile 0: an inclusive prefix of zero.
                const uint64_t st = j >= 0 ? __hip_atomic_load(status + j, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : SC_PRE;
                const uint64_t pre = __ballot((st & SC_PRE) != 0ull);
                const uint64_t pub = __ballot((st & (SC_PRE | SC_AGG)) != 0ull);
                const uint32_t first = pre ? (uint32_t)__ffsll((unsigned long long)pre) - 1u : 64u;
                const uint64_t need = first >= 63u ? ~0ull : ((1ull << (first + 1u)) - 1ull);
                if ((pub & need) != need) {
                    __builtin_amdgcn_s_sleep(1);
                    continue;
                }
                int64_t v = lane <= first ? sc_value(st) : 0;
                for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
                excl += v;
                if (first < 64u) break;
                j0 -= 64;
            }
            if (lane == 0)
                __hip_atomic_store(my, SC_PRE | ((uint64_t)(excl + total) & SC_MASK), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        if (lane == 0) s_excl = excl;
    }
    __syncthreads();
    int64_t run = s_excl + before + x - mine;
#pragma unroll
    for (uint32_t r = 0; r < SC_ITEMS; ++r) {
        run += v[r];
        uint32_t i = base + r;
        if (i < n) stab[i] = (uint32_t)run;
    }
}

// One thread per node id (inner [0, n), leaf [n, 2n)). Depth of an inner node
// = ranges covering its split edge less one (itself); of leaf slot k = the
// larger of the coverings of edges k - 1 and k (an edge across a segment's end
// is covered by nothing). Deeper than 31 means an ancestor at depth 31 was
// packed (d + 1 >= STACK), so the node is unreachable.
// With `inner` (one byte a node of the node array, or null) each inner
// node's byte says whether a walk reaches it as an inner node (not packed,
// not under a packed leaf): the trees' walk records are made for those.
__global__ __launch_bounds__(256) void k_pack(const PhrLink* __restrict__ links, const uint4* __restrict__ up,
                                              const uint32_t* __restrict__ stab, uint32_t n, const uint32_t* __restrict__ seg_lo,
                                              const uint32_t* __restrict__ node_base, uint32_t segs, uint32_t leaf,
                                              uint32_t* __restrict__ leaf_at, uint32_t* __restrict__ depth_out,
                                              uint32_t* __restrict__ acc, uint8_t* __restrict__ inner_out) {
    __shared__ uint32_t s_seg[4], s_depth[4];
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    for (uint32_t j = t; j < 6 * segs; j += gridDim.x * blockDim.x) acc[j] = (j % 6) < 3 ? KEY_MAX : KEY_MIN;
    bool listed = false;
    uint32_t depth = 0, sg = 0;
    if (t < 2 * n) {
        const bool inner = t < n;
        const uint32_t slot = inner ? t : t - n;
        sg = seg_of_slot(seg_lo, segs, slot);
        bool node = true;
        if (inner) node = slot + 1 < seg_lo[sg + 1];  // the last slot of a segment starts no inner node
        if (node) {
            PhrLink l{};
            uint32_t d, first;
            if (inner) {
                l = links[t];
                d = stab[l.left < n ? l.left : l.left - n] - 1u;
                first = l.first;
            } else {
                first = slot;
                d = max(slot > 0 ? stab[slot - 1] : 0u, stab[slot]);
            }
            // A packed parent mutes its subtree. Ranges shrink down the tree,
            // so only the parent's own size test matters; its depth (d - 1 <=
            // 30) is below the cap.
            bool muted = false;
            if (d < PHR_STACK) {
                const uint4 u = up[t];
                muted = u.x != PHR_NONE && u.z - u.y < leaf;
            }
            if (d < PHR_STACK && !muted) {
                listed = inner ? ((l.last - l.first < leaf) || (d + 1 >= PHR_STACK)) : true;
                if (listed) leaf_at[first] = t;
                depth = d;
            }
            if (inner && inner_out) {
                const Seg g = seg_at(seg_lo, node_base, sg);
                inner_out[out_of(t, n, g)] = (d < PHR_STACK && !muted && !listed) ? 1u : 0u;
            }
        }
    }
    // Every reachable node has a listed leaf below it at least as deep, so the
    // deepest listed leaf is its tree's depth. Waves whose listed nodes share
    // one segment fold through LDS to one word a block, and a fold only takes
    // the atomic when it would raise the depth already there.
    const uint32_t lane = __lane_id(), wave = threadIdx.x >> 6;
    const uint64_t m = __ballot(listed);
    bool uniform = true;
    uint32_t s0 = 0, dm = 0;
    if (m) {
        const uint32_t first_lane = (uint32_t)__ffsll((unsigned long long)m) - 1;
        s0 = (uint32_t)__shfl((int)sg, (int)first_lane);
        uniform = __ballot(listed && sg != s0) == 0;
        dm = wave_max_u(listed ? depth : 0u);
        if (!uniform && listed && depth > __hip_atomic_load(&depth_out[sg], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
            atomicMax(&depth_out[sg], depth);
    }
    if (lane == 0) {
        s_seg[wave] = (m && uniform) ? s0 : PHR_NONE;
        s_depth[wave] = dm;
    }
    __syncthreads();
    if (threadIdx.x < 4) {
        const uint32_t w = threadIdx.x, sw = s_seg[w];
        bool lead = sw != PHR_NONE;
        uint32_t dw = s_depth[w];
        for (uint32_t v = 0; v < 4 && lead; ++v) {
            if (v == w || s_seg[v] != sw) continue;
            if (v < w) lead = false;  // the first wave of this segment folds for all
            else dw = max(dw, s_depth[v]);
        }
        if (lead && dw > __hip_atomic_load(&depth_out[sw], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) atomicMax(&depth_out[sw], dw);
    }
}

// ---- 6. refit ------------------------------------------------------------------
__device__ __forceinline__ void store_node(float4* nodes, uint32_t k, const Box& b, uint32_t w0, uint32_t w1) {
    nodes[2 * k] = make_float4(b.lo0, b.lo1, b.lo2, __uint_as_float(w0));
    nodes[2 * k + 1] = make_float4(b.hi0, b.hi1, b.hi2, __uint_as_float(w1));
}

// Arrival counters of one block's inner nodes: 2 bits per node in LDS (a node
// sees at most two arrivals), indexed by the node's id less the block's first
// slot (an inner node's id is a slot of its own range, so it lies in the
// block's slots); or, for blocks owning more slots than LDS holds, one word
// per node in global memory, put back to zero by the second arrival.
template <uint32_t CW>
struct Arrivals {
    uint32_t* s;
    uint32_t* g;
    uint64_t s0;
    __device__ __forceinline__ bool second(uint32_t p) {
        if constexpr (CW > 0) {
            uint32_t i = (uint32_t)(p - s0), sh = 2 * (i & 15u);
            uint32_t old = __hip_atomic_fetch_add(&s[i >> 4], 1u << sh, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP);
            return ((old >> sh) & 3u) != 0u;
        } else {
            uint32_t old = __hip_atomic_fetch_add(&g[p], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP);
            if (old == 0u) return false;
            __hip_atomic_store(&g[p], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            return true;
        }
    }
};

// Carry a finished node x of segment g up while its parent's range lies in
// [s0, s1); a node whose parent's range leaves it is handed to the next level.
// One record a step: the parent, its range and the sibling come together.
template <uint32_t CW>
__device__ __forceinline__ void climb(uint32_t x, Box b, const Seg& g, uint32_t n, uint64_t s0, uint64_t s1,
                                      const uint4* __restrict__ up, float4* nodes, Arrivals<CW>& arr, uint32_t* s_fcount,
                                      uint32_t* front_out) {
    uint4 u = up[x];
    for (;;) {
        const uint32_t p = u.x;
        if (p == PHR_NONE) return;  // the root is done
        if (u.y < s0 || u.z >= s1) {
            uint32_t at = atomicAdd(s_fcount, 1u);
            front_out[at] = x;
            return;
        }
        const uint4 un = up[p];  // the parent's record, on its way while this step finishes
        if (!arr.second(p)) return;  // the sibling's climb finishes p
        const uint32_t sib = u.w & ~UP_LEFT;
        const uint32_t so = out_of(sib, n, g), xo = out_of(x, n, g);
        const float4 sa = nodes[2 * so], sb = nodes[2 * so + 1];
        grow(b, sa, sb);
        const bool x_left = (u.w & UP_LEFT) != 0u;
        store_node(nodes, out_of(p, n, g), b, x_left ? xo : so, x_left ? so : xo);
        x = p;
        u = un;
    }
}

// Level 0: one block per RF_SLOTS slots; a thread per leaf starting there.
__global__ __launch_bounds__(RF_LEAF_THREADS) void k_refit_leaves(const float4* __restrict__ prims, const uint32_t* __restrict__ order,
                                                             const PhrLink* __restrict__ links, const uint4* __restrict__ up,
                                                             const uint32_t* __restrict__ leaf_at, uint32_t n,
                                                             const uint32_t* __restrict__ seg_lo, const uint32_t* __restrict__ node_base,
                                                             uint32_t segs, uint32_t leaf_flag, float4* nodes,
                                                             uint32_t* __restrict__ front_out, uint32_t* __restrict__ fcount_out) {
    constexpr uint32_t CW = RF_SLOTS / 16;
    __shared__ uint32_t s_cnt[CW];
    __shared__ uint32_t s_fcount;
    const uint32_t tid = threadIdx.x;
    for (uint32_t i = tid; i < CW; i += RF_LEAF_THREADS) s_cnt[i] = 0u;
    if (tid == 0) s_fcount = 0u;
    __syncthreads();
    const uint64_t s0 = (uint64_t)blockIdx.x * RF_SLOTS, s1 = s0 + RF_SLOTS;
    Arrivals<CW> arr{s_cnt, nullptr, s0};
    uint32_t* out = front_out + s0;
    for (uint32_t r = 0; r < RF_SLOTS / RF_LEAF_THREADS; ++r) {
        const uint64_t slot = s0 + r * RF_LEAF_THREADS + tid;
        if (slot >= n) break;
        const uint32_t k = leaf_at[slot];
        if (k == PHR_NONE) continue;
        uint32_t first, end;
        if (k < n) {
            PhrLink l = links[k];
            first = l.first;
            end = l.last + 1;
        } else {
            first = k - n;
            end = first + 1;
        }
        const Seg g = seg_at(seg_lo, node_base, seg_of_slot(seg_lo, segs, first));
        Box b = empty_box();
        for (uint32_t q = first; q < end; ++q) {
            uint32_t pi = order[q];
            grow(b, prims[2 * pi], prims[2 * pi + 1]);
        }
        store_node(nodes, out_of(k, n, g), b, first | leaf_flag, end);
        climb<CW>(k, b, g, n, s0, s1, up, nodes, arr, &s_fcount, out);
    }
    __syncthreads();
    if (tid == 0) fcount_out[blockIdx.x] = s_fcount;
}

// Level L >= 1: block b owns the slots of blocks [RF_GROUP b, RF_GROUP b + RF_GROUP)
// of level L - 1 (span slots each) and carries on the climbs they handed over.
template <uint32_t CW>
__global__ __launch_bounds__(RF_THREADS) void k_refit_up(const uint4* __restrict__ up,
                                                         uint32_t n, const uint32_t* __restrict__ seg_lo,
                                                         const uint32_t* __restrict__ node_base, uint32_t segs, float4* nodes,
                                                         uint32_t* __restrict__ gcnt, uint32_t below_blocks, uint64_t span,
                                                         const uint32_t* __restrict__ front_in, const uint32_t* __restrict__ fcount_in,
                                                         uint32_t* __restrict__ front_out, uint32_t* __restrict__ fcount_out) {
    __shared__ uint32_t s_cnt[CW > 0 ? CW : 1];
    __shared__ uint32_t s_off[RF_GROUP + 1];
    __shared__ uint32_t s_fcount;
    const uint32_t tid = threadIdx.x;
    if constexpr (CW > 0)
        for (uint32_t i = tid; i < CW; i += RF_THREADS) s_cnt[i] = 0u;
    if (tid == 0) s_fcount = 0u;
    const uint32_t j0 = blockIdx.x * RF_GROUP;
    if (tid < 64) {
        // Inclusive scan of the handed-over counts over one wave.
        const uint32_t j = j0 + tid;
        uint32_t c = j < below_blocks ? fcount_in[j] : 0u, x = c;
        for (int o = 1; o < 64; o <<= 1) {
            uint32_t y = (uint32_t)__shfl_up((int)x, o);
            if (tid >= (uint32_t)o) x += y;
        }
        s_off[tid + 1] = x;
        if (tid == 0) s_off[0] = 0u;
    }
    __syncthreads();
    const uint64_t s0 = (uint64_t)blockIdx.x * RF_GROUP * span, s1 = s0 + RF_GROUP * span;
    Arrivals<CW> arr{s_cnt, gcnt, s0};
    uint32_t* out = front_out + s0;
    const uint32_t total = s_off[RF_GROUP];
    for (uint32_t e = tid; e < total; e += RF_THREADS) {
        uint32_t lo = 0, hi = RF_GROUP;  // the list holding entry e: s_off[lo] <= e < s_off[lo + 1]
        while (hi - lo > 1) {
            uint32_t mid = (lo + hi) >> 1;
            if (s_off[mid] <= e) lo = mid; else hi = mid;
        }
        const uint32_t x = front_in[(uint64_t)(j0 + lo) * span + (e - s_off[lo])];
        const Seg g = seg_at(seg_lo, node_base, seg_of_slot(seg_lo, segs, x < n ? x : x - n));
        const uint32_t xo = out_of(x, n, g);
        const float4 a = nodes[2 * xo], c = nodes[2 * xo + 1];
        Box b{a.x, a.y, a.z, c.x, c.y, c.z};
        climb<CW>(x, b, g, n, s0, s1, up, nodes, arr, &s_fcount, out);
    }
    __syncthreads();
    if (tid == 0) fcount_out[blockIdx.x] = s_fcount;
}

// ---- GpuSort (src/bvh/sort.rs:13-111) ---------------------------------------
// The four digit histograms of arbitrary 32-bit codes.
__global__ __launch_bounds__(256) void k_sort_hist(const uint32_t* __restrict__ codes, uint32_t n, uint32_t* __restrict__ hist) {
    __shared__ uint32_t s_h[4 * 256];
    const uint32_t tid = threadIdx.x;
    for (uint32_t i = tid; i < 4 * 256; i += 256) s_h[i] = 0u;
    __syncthreads();
    const uint32_t base = blockIdx.x * OS_TILE;
#pragma unroll
    for (uint32_t r = 0; r < OS_ITEMS; ++r) {
        const uint32_t i = base + r * 256 + tid;
        if (i < n) {
            const uint32_t c = codes[i];
            atomicAdd(&s_h[c & 0xFFu], 1u);
            atomicAdd(&s_h[256 + ((c >> 8) & 0xFFu)], 1u);
            atomicAdd(&s_h[512 + ((c >> 16) & 0xFFu)], 1u);
            atomicAdd(&s_h[768 + (c >> 24)], 1u);
        }
    }
    __syncthreads();
    for (uint32_t i = tid; i < 4 * 256; i += 256)
        if (s_h[i]) atomicAdd(&hist[i], s_h[i]);
}

// The sorted keys as GpuSort::keys holds them, (original slot, code), and the
// counters cleared for the next replay.
__global__ __launch_bounds__(256) void k_sort_keys(const uint32_t* __restrict__ order, const uint32_t* __restrict__ sorted,
                                                   uint32_t n, uint2* __restrict__ keys, uint32_t* __restrict__ zero,
                                                   uint32_t zero_words) {
    const uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x, gsize = gridDim.x * blockDim.x;
    for (uint32_t i = gtid; i < n; i += gsize) keys[i] = make_uint2(order[i], sorted[i]);
    for (uint32_t i = gtid; i < zero_words; i += gsize) zero[i] = 0u;
}

}  // namespace

// ---------------------------------------------------------------------------
// GpuBvh (src/bvh/gpu.rs:11-104), one tree or a batch of segments
// ---------------------------------------------------------------------------
// ---- breadth-first numbering (Bvh::build_compact, src/bvh/build.rs:104-136) ----
// Breadth first visits a level's nodes left to right, and the nodes of one
// level have disjoint ranges, so a reachable node's number is its rank by
// (depth, first slot). Key per node id: that pair, or all ones when the node
// is not reachable (k_pack's test: a packed ancestor or depth past the cap).
constexpr uint32_t CP_FIRST_BITS = 26;  // trees up to 2^26 primitives
__global__ __launch_bounds__(256) void k_compact_keys(const PhrLink* __restrict__ links, const uint4* __restrict__ up,
                                                      const uint32_t* __restrict__ stab, uint32_t n, uint32_t leaf,
                                                      uint32_t* __restrict__ keys, uint32_t* __restrict__ count) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    bool reach = false;
    uint32_t key = 0xffffffffu;
    if (t < 2 * n) {
        const bool inner = t < n;
        const uint32_t slot = inner ? t : t - n;
        if (!inner || slot + 1 < n) {
            uint32_t d, first;
            if (inner) {
                const PhrLink l = links[t];
                d = stab[l.left < n ? l.left : l.left - n] - 1u;
                first = l.first;
            } else {
                first = slot;
                d = max(slot > 0 ? stab[slot - 1] : 0u, stab[slot]);
            }
            bool muted = false;
            if (d < PHR_STACK) {
                const uint4 u = up[t];
                muted = u.x != PHR_NONE && u.z - u.y < leaf;
            }
            reach = d < PHR_STACK && !muted;
            if (reach) key = (d << CP_FIRST_BITS) | first;
        }
        keys[t] = key;
    }
    const uint64_t m = __ballot(reach);
    if (__lane_id() == 0 && m) atomicAdd(count, (uint32_t)__popcll(m));
}

// Position of each reachable node id in the breadth-first order.
__global__ __launch_bounds__(256) void k_compact_inv(const uint32_t* __restrict__ ids, uint32_t r, uint32_t* __restrict__ inv) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < r) inv[ids[p]] = p;
}

// Node p of the breadth-first tree: the packed node's box; an inner node that
// was not packed points at its children's positions, the others are leaves.
__global__ __launch_bounds__(256) void k_compact_write(const uint32_t* __restrict__ ids, const uint32_t* __restrict__ skeys,
                                                       uint32_t r, const PhrLink* __restrict__ links, const uint32_t* __restrict__ inv,
                                                       uint32_t n, uint32_t leaf, const float4* __restrict__ nodes,
                                                       float4* __restrict__ out) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= r) return;
    const uint32_t id = ids[p], d = skeys[p] >> CP_FIRST_BITS;
    const uint32_t at = id < n ? id : (n - 1) + (id - n);  // one tree: node_base 0
    float4 lo = nodes[2 * (size_t)at], hi = nodes[2 * (size_t)at + 1];
    uint32_t a, b;
    if (id < n) {
        const PhrLink l = links[id];
        if (l.last - l.first < leaf || d + 1 >= PHR_STACK) {
            a = l.first | PHR_LEAF;
            b = l.last + 1;
        } else {
            a = inv[l.left];
            b = inv[l.right];
        }
    } else {
        a = (id - n) | PHR_LEAF;
        b = id - n + 1;
    }
    lo.w = __uint_as_float(a);
    hi.w = __uint_as_float(b);
    out[2 * (size_t)p] = lo;
    out[2 * (size_t)p + 1] = hi;
}

struct phr_bvh {
    phr_gpu* gpu = nullptr;
    uint32_t n = 0, leaf = 4, segs = 1, nonempty = 1, tiles = 0, bounds_blocks = 0, seg_passes = 0, leaf_flag = PHR_LEAF;
    uint32_t items = OS_ITEMS;  // keys a thread holds in a sort pass (the tile is 256 x items)
    std::vector<uint32_t> seg_lo_h, node_base_h, depth_h;
    phr_array* prims = nullptr;
    bool own_prims = false;
    phr_array* order = nullptr;  // n u32: sorted slot -> input index
    phr_array* nodes = nullptr;  // the node array written (owned unless own_nodes is false)
    bool own_nodes = true;
    uint32_t* seg_lo = nullptr;     // segs + 1
    uint32_t* node_base = nullptr;  // segs
    uint32_t* codes = nullptr;      // n: sorted codes after the last pass
    uint32_t* keys_b = nullptr;
    uint32_t* vals_b = nullptr;
    uint32_t* zero = nullptr;  // histograms, tickets, depths, look-back status
    uint32_t zero_words = 0, status_at = 0, scan_status = 0;
    uint32_t* acc = nullptr;  // each segment's scene box as six order-preserving keys
    void* parts = nullptr;    // one Part per bounds block
    PhrLink* links = nullptr;     // n, by inner id
    uint4* up = nullptr;          // 2n, by node id: parent, parent's range, sibling
    uint32_t* cnt = nullptr;      // n: arrival counters of the levels above LDS size
    uint32_t* leaf_at = nullptr;  // n: the reachable leaf whose range starts at a slot, or NONE
    uint32_t* own_end = nullptr;   // n: +1/-1 for the inner range of which a slot is one end
    uint32_t* far_ends = nullptr;  // n: the +1/-1 terms other ranges' far ends put on a slot
    uint32_t* front[2] = {nullptr, nullptr};   // n each: nodes handed from one refit level to the next
    uint32_t* fcount[2] = {nullptr, nullptr};  // per block of a level: how many it handed over
    uint32_t level0_blocks = 0;
    uint8_t* inner = nullptr;  // a byte a node of the node array: an inner node a walk reaches
};

namespace {

void bvh_release(phr_bvh* b) {
    if (!b) return;
    if (b->own_prims) phr_array_release(b->prims);
    phr_array_release(b->order);
    if (b->own_nodes) phr_array_release(b->nodes);
    for (void* p : {(void*)b->seg_lo, (void*)b->node_base, (void*)b->codes, (void*)b->keys_b, (void*)b->vals_b,
                    (void*)b->zero, (void*)b->acc, b->parts, (void*)b->links, (void*)b->up, (void*)b->cnt, (void*)b->leaf_at,
                    (void*)b->own_end, (void*)b->far_ends, (void*)b->front[0], (void*)b->front[1], (void*)b->fcount[0],
                    (void*)b->fcount[1], (void*)b->inner})
        if (p) (void)hipFree(p);
    delete b;
}

// seg_lo: segs + 1 non-decreasing starts from 0 to n. nodes_ext/node_base:
// an existing node array and where each segment's tree goes in it, or null
// for a new array with the trees back to back.
int bvh_alloc(phr_gpu* g, phr_array* prims, bool own, uint32_t leaf, const uint32_t* seg_lo, uint32_t segs, phr_array* nodes_ext,
              const uint32_t* node_base, uint32_t leaf_flag, phr_bvh** out) {
    const uint32_t n = (uint32_t)prims->len;
    PHR_CHECK(n < (1u << 30), PHR_ERR_SIZE, "GpuBvh: at most 2^30 - 1 primitives");
    PHR_CHECK(leaf >= 1, PHR_ERR_ARG, "GpuBvh: leaf size must be at least 1");
    PHR_CHECK(segs >= 1 && segs <= MAX_SEGMENTS, PHR_ERR_SIZE, "GpuBvh: between 1 and 65536 segments");
    std::vector<uint32_t> lo(segs + 1);
    if (seg_lo) {
        std::memcpy(lo.data(), seg_lo, 4ull * (segs + 1));
    } else {
        PHR_CHECK(segs == 1, PHR_ERR_ARG, "GpuBvh: segment starts missing");
        lo[0] = 0;
        lo[1] = n;
    }
    PHR_CHECK(lo[0] == 0 && lo[segs] == n, PHR_ERR_ARG, "GpuBvh: segment starts must run from 0 to the box count");
    uint32_t nonempty = 0;
    uint64_t total_nodes = 0;
    std::vector<uint32_t> base(segs);
    for (uint32_t s = 0; s < segs; ++s) {
        PHR_CHECK(lo[s] <= lo[s + 1], PHR_ERR_ARG, "GpuBvh: segment starts must not decrease");
        const uint32_t ns = lo[s + 1] - lo[s];
        base[s] = node_base ? node_base[s] : (uint32_t)total_nodes;
        if (ns) {
            ++nonempty;
            total_nodes += 2ull * ns - 1;
            if (nodes_ext)
                PHR_CHECK((uint64_t)base[s] + 2ull * ns - 1 <= nodes_ext->len, PHR_ERR_SIZE, "GpuBvh: a tree does not fit the node array");
        }
    }
    phr_bvh* b = new phr_bvh();
    b->gpu = g;
    b->n = n;
    b->leaf = leaf;
    b->segs = segs;
    b->nonempty = nonempty;
    b->leaf_flag = leaf_flag;
    b->seg_lo_h = lo;
    b->node_base_h = base;
    b->depth_h.assign(segs, 0);
    b->prims = prims;
    b->own_prims = false;  // taken over only on success: the caller frees them on failure
    b->seg_passes = segs <= 1 ? 0 : (segs <= 256 ? 1 : 2);
    b->tiles = (n + OS_TILE - 1) / OS_TILE;
    b->bounds_blocks = std::min<uint32_t>(BOUNDS_BLOCKS_MAX, std::max<uint32_t>(1, (n + 2047) / 2048));
    b->status_at = (Z_DEPTH + segs + 63) & ~63u;
    b->scan_status = (b->status_at + (CODE_PASSES + b->seg_passes) * b->tiles * 256 + 1) & ~1u;  // 8-byte aligned
    b->zero_words = b->scan_status + 2 * ((n + SC_TILE - 1) / SC_TILE) + 2;
    int rc = PHR_OK;
    auto fail = [&](int code) { bvh_release(b); return code; };
    if ((rc = phr_array_alloc(g, 4, n, &b->order)) != PHR_OK) return fail(rc);
    if (nodes_ext) {
        b->nodes = nodes_ext;
        b->own_nodes = false;
    } else if ((rc = phr_array_alloc(g, 32, total_nodes, &b->nodes)) != PHR_OK) {
        return fail(rc);
    }
    uint64_t nn = std::max<uint64_t>(n, 1);
    auto dev = [&](void** p, uint64_t bytes) -> bool {
        if (hipMalloc(p, std::max<uint64_t>(bytes, 16)) != hipSuccess) return false;
        return hipMemsetAsync(*p, 0, std::max<uint64_t>(bytes, 16), g->stream) == hipSuccess;
    };
    b->level0_blocks = (uint32_t)((nn + RF_SLOTS - 1) / RF_SLOTS);
    bool ok = dev((void**)&b->seg_lo, 4ull * (segs + 1)) && dev((void**)&b->node_base, 4ull * segs) &&
              dev((void**)&b->codes, 4 * nn) &&
              dev((void**)&b->keys_b, 4 * nn) && dev((void**)&b->vals_b, 4 * nn) && dev((void**)&b->zero, 4ull * b->zero_words) &&
              dev((void**)&b->acc, 24ull * segs) && dev(&b->parts, 32ull * b->bounds_blocks) &&
              dev((void**)&b->links, 16 * nn) && dev((void**)&b->up, 32 * nn) &&
              dev((void**)&b->cnt, 4 * nn) && dev((void**)&b->leaf_at, 4 * nn) && dev((void**)&b->own_end, 4 * nn) &&
              dev((void**)&b->far_ends, 4 * nn) && dev((void**)&b->front[0], 4 * nn) && dev((void**)&b->front[1], 4 * nn) &&
              dev((void**)&b->fcount[0], 4ull * b->level0_blocks) && dev((void**)&b->fcount[1], 4ull * b->level0_blocks) &&
              dev((void**)&b->inner, std::max<uint64_t>(b->nodes->len, 1));
    // Aabb::default() as keys for every segment: fkey(f32::MAX) x 3, fkey(f32::MIN) x 3.
    std::vector<uint32_t> acc0(6ull * segs);
    for (size_t j = 0; j < acc0.size(); ++j) acc0[j] = (j % 6) < 3 ? KEY_MAX : KEY_MIN;
    ok = ok && hipMemcpyAsync(b->acc, acc0.data(), 4 * acc0.size(), hipMemcpyHostToDevice, g->stream) == hipSuccess &&
         hipMemcpyAsync(b->seg_lo, lo.data(), 4ull * (segs + 1), hipMemcpyHostToDevice, g->stream) == hipSuccess &&
         hipMemcpyAsync(b->node_base, base.data(), 4ull * segs, hipMemcpyHostToDevice, g->stream) == hipSuccess &&
         hipStreamSynchronize(g->stream) == hipSuccess;
    if (!ok) {
        phr_set_error_msg("GpuBvh: device allocation failed");
        return fail(PHR_ERR_HIP);
    }
    b->own_prims = own;
    *out = b;
    return PHR_OK;
}

int bvh_enqueue_refit(phr_bvh* b);

// One onesweep pass of tiles of 256 x ITEMS keys.
template <uint32_t ITEMS>
void launch_pass(bool first, bool seg, uint32_t tiles, hipStream_t s, uint32_t* kin, uint32_t* vin, uint32_t* kout,
                 uint32_t* vout, const uint32_t* hist, uint32_t* st, uint32_t* ticket, uint32_t n, uint32_t shift,
                 const uint32_t* seg_lo, uint32_t segs) {
    if (first)
        hipLaunchKernelGGL((k_onesweep<true, false, ITEMS>), dim3(tiles), dim3(OS_THREADS), 0, s, kin, vin, kout, vout, hist, st,
                           ticket, n, shift, seg_lo, segs);
    else if (!seg)
        hipLaunchKernelGGL((k_onesweep<false, false, ITEMS>), dim3(tiles), dim3(OS_THREADS), 0, s, kin, vin, kout, vout, hist, st,
                           ticket, n, shift, seg_lo, segs);
    else
        hipLaunchKernelGGL((k_onesweep<false, true, ITEMS>), dim3(tiles), dim3(OS_THREADS), 0, s, kin, vin, kout, vout, hist, st,
                           ticket, n, shift, seg_lo, segs);
}

// Enqueue the whole build on the device's stream (no host synchronisation).
int bvh_enqueue_build(phr_bvh* b) {
    const uint32_t n = b->n;
    if (n == 0) return PHR_OK;
    hipStream_t s = b->gpu->stream;
    const float4* prims = (const float4*)b->prims->ptr;
    uint32_t* order = (uint32_t*)b->order->ptr;
    uint32_t* hist = b->zero + Z_HIST;
    uint32_t* ticket = b->zero + Z_TICKET;
    uint32_t* status = b->zero + b->status_at;
    const uint32_t passes = CODE_PASSES + b->seg_passes;
    hipLaunchKernelGGL(k_bounds, dim3(b->bounds_blocks), dim3(256), 0, s, prims, n, b->seg_lo, b->segs, b->acc, (Part*)b->parts,
                       b->zero, b->zero_words);
    hipLaunchKernelGGL(k_bounds_fold, dim3(1), dim3(256), 0, s, (const Part*)b->parts, b->bounds_blocks, b->acc);
    // The passes alternate between (codes, order) and (keys_b, vals_b) and
    // must end in (codes, order): with an odd count the codes start in keys_b.
    uint32_t* unsorted = (passes % 2) ? b->keys_b : b->codes;
    hipLaunchKernelGGL(k_morton, dim3(b->tiles), dim3(256), 0, s, prims, (const uint32_t*)b->acc, b->seg_lo, b->segs, b->seg_passes,
                       n, unsorted, hist, b->far_ends);
    uint32_t* kin = unsorted;
    uint32_t* vin = nullptr;
    // The status words are sized for 2048-key tiles; a larger tile uses fewer.
    const uint32_t sort_tiles = (n + OS_THREADS * b->items - 1) / (OS_THREADS * b->items);
    for (uint32_t p = 0; p < passes; ++p) {
        const bool to_codes = ((passes - p) % 2) == 1;
        uint32_t* kout = to_codes ? b->codes : b->keys_b;
        uint32_t* vout = to_codes ? order : b->vals_b;
        const bool seg = p >= CODE_PASSES;
        const uint32_t shift = seg ? 8 * (p - CODE_PASSES) : 8 * p;
        uint32_t* st = status + (size_t)p * sort_tiles * 256;
        if (b->items == 16)
            launch_pass<16>(p == 0, seg, sort_tiles, s, kin, vin, kout, vout, hist + 256 * p, st, ticket + p, n, shift, b->seg_lo,
                            b->segs);
        else
            launch_pass<OS_ITEMS>(p == 0, seg, sort_tiles, s, kin, vin, kout, vout, hist + 256 * p, st, ticket + p, n, shift,
                                  b->seg_lo, b->segs);
        kin = kout;
        vin = vout;
    }
    hipLaunchKernelGGL(k_karras, dim3((n + 255) / 256), dim3(256), 0, s, b->codes, n, b->seg_lo, b->segs, b->links, b->up,
                       b->leaf_at, b->own_end, b->far_ends);
    // The slot scan writes its coverings over the sort's spare key buffer.
    uint32_t* stab = b->keys_b;
    hipLaunchKernelGGL(k_stab, dim3((n + SC_TILE - 1) / SC_TILE), dim3(256), 0, s, b->own_end, b->far_ends, n, stab,
                       (uint64_t*)(b->zero + b->scan_status), b->zero + Z_SCAN_TICKET);
    hipLaunchKernelGGL(k_pack, dim3((2 * n + 255) / 256), dim3(256), 0, s, b->links, b->up, stab, n, b->seg_lo, b->node_base,
                       b->segs, b->leaf, b->leaf_at, b->zero + Z_DEPTH, b->acc, b->inner);
    PHR_HIP(hipGetLastError());
    return bvh_enqueue_refit(b);
}

// The refit levels: level 0 over RF_SLOTS-slot blocks, then RF_GROUP times
// wider blocks until one block owns every slot.
int bvh_enqueue_refit(phr_bvh* b) {
    const uint32_t n = b->n;
    if (n == 0) return PHR_OK;
    hipStream_t s = b->gpu->stream;
    float4* nodes = (float4*)b->nodes->ptr;
    uint32_t blocks = b->level0_blocks;
    hipLaunchKernelGGL(k_refit_leaves, dim3(blocks), dim3(RF_LEAF_THREADS), 0, s, (const float4*)b->prims->ptr,
                       (const uint32_t*)b->order->ptr, b->links, b->up, b->leaf_at, n, b->seg_lo, b->node_base, b->segs,
                       b->leaf_flag, nodes, b->front[0], b->fcount[0]);
    uint64_t span = RF_SLOTS;
    int cur = 0;
    while (blocks > 1) {
        const uint32_t up = (blocks + RF_GROUP - 1) / RF_GROUP;
        if (span * RF_GROUP <= RF_LDS_SLOTS)
            hipLaunchKernelGGL(k_refit_up<RF_LDS_SLOTS / 16>, dim3(up), dim3(RF_THREADS), 0, s, b->up, n, b->seg_lo,
                               b->node_base, b->segs, nodes, b->cnt, blocks, span, b->front[cur], b->fcount[cur], b->front[1 - cur],
                               b->fcount[1 - cur]);
        else
            hipLaunchKernelGGL(k_refit_up<0>, dim3(up), dim3(RF_THREADS), 0, s, b->up, n, b->seg_lo, b->node_base,
                               b->segs, nodes, b->cnt, blocks, span, b->front[cur], b->fcount[cur], b->front[1 - cur],
                               b->fcount[1 - cur]);
        blocks = up;
        span *= RF_GROUP;
        cur = 1 - cur;
    }
    PHR_HIP(hipGetLastError());
    return PHR_OK;
}

int bvh_read_depth(phr_bvh* b) {
    if (b->n == 0) {
        std::fill(b->depth_h.begin(), b->depth_h.end(), 0u);
        return PHR_OK;
    }
    PHR_HIP(hipMemcpyAsync(b->depth_h.data(), b->zero + Z_DEPTH, 4ull * b->segs, hipMemcpyDeviceToHost, b->gpu->stream));
    PHR_HIP(hipStreamSynchronize(b->gpu->stream));
    return PHR_OK;
}

int bvh_new(phr_gpu* g, phr_array* prims, bool own, uint32_t leaf, const uint32_t* seg_lo, uint32_t segs, phr_array* nodes_ext,
            const uint32_t* node_base, uint32_t leaf_flag, phr_bvh** out) {
    phr_bvh* b = nullptr;
    int rc = bvh_alloc(g, prims, own, leaf, seg_lo, segs, nodes_ext, node_base, leaf_flag, &b);
    if (rc != PHR_OK) return rc;
    if ((rc = bvh_enqueue_build(b)) != PHR_OK || (rc = bvh_read_depth(b)) != PHR_OK) {
        if (own) b->own_prims = false;  // the caller still owns them on failure
        bvh_release(b);
        return rc;
    }
    *out = b;
    return PHR_OK;
}

int upload_prims(phr_gpu* g, const void* prims, uint32_t n, phr_array** out) {
    phr_array* a = nullptr;
    int rc = phr_array_alloc(g, 32, n, &a);
    if (rc != PHR_OK) return rc;
    if (n) {
        hipError_t e = hipMemcpyAsync(a->ptr, prims, 32ull * n, hipMemcpyHostToDevice, g->stream);
        if (e != hipSuccess) {
            phr_array_release(a);
            phr_set_error("hipMemcpyAsync", e);
            return PHR_ERR_HIP;
        }
    }
    *out = a;
    return PHR_OK;
}

}  // namespace

// Internal entry points for the tracer (trace.hip).
int phr_bvh_internal_build(phr_gpu* g, phr_array* prims, uint32_t leaf, phr_bvh** out) {
    return bvh_new(g, prims, false, leaf, nullptr, 1, nullptr, nullptr, PHR_LEAF, out);
}
// Trees over segments of `prims` written into `nodes` at `node_base` (host
// arrays of segs + 1 and segs words), leaves flagged with `leaf_flag`.
int phr_bvh_internal_build_into(phr_gpu* g, phr_array* prims, uint32_t leaf, const uint32_t* seg_lo, uint32_t segs,
                                phr_array* nodes, const uint32_t* node_base, uint32_t leaf_flag, phr_bvh** out) {
    return bvh_new(g, prims, false, leaf, seg_lo, segs, nodes, node_base, leaf_flag, out);
}
int phr_bvh_internal_rebuild_async(phr_bvh* b) { return bvh_enqueue_build(b); }
int phr_bvh_internal_refit_async(phr_bvh* b) { return bvh_enqueue_refit(b); }
int phr_bvh_internal_read_depth(phr_bvh* b) { return bvh_read_depth(b); }
phr_array* phr_bvh_internal_nodes(phr_bvh* b) { return b->nodes; }
phr_array* phr_bvh_internal_order(phr_bvh* b) { return b->order; }
const uint8_t* phr_bvh_internal_inner(phr_bvh* b) { return b->inner; }
uint32_t phr_bvh_internal_segment_depth(const phr_bvh* b, uint32_t s) { return s < b->segs ? b->depth_h[s] : 0; }

extern "C" {

int phr_bvh_build(phr_gpu* g, const void* prims, uint32_t n, uint32_t leaf, phr_bvh** out) {
    PHR_CHECK(g && out && (prims || n == 0), PHR_ERR_ARG, "phr_bvh_build: null argument");
    phr_array* a = nullptr;
    int rc = upload_prims(g, prims, n, &a);
    if (rc != PHR_OK) return rc;
    if ((rc = bvh_new(g, a, true, leaf, nullptr, 1, nullptr, nullptr, PHR_LEAF, out)) != PHR_OK) phr_array_release(a);
    return rc;
}

int phr_bvh_build_on(phr_gpu* g, phr_array* prims, uint32_t leaf, phr_bvh** out) {
    PHR_CHECK(g && prims && out && prims->elem == 32, PHR_ERR_ARG, "phr_bvh_build_on: prims must be an array of 32-byte boxes");
    return bvh_new(g, prims, false, leaf, nullptr, 1, nullptr, nullptr, PHR_LEAF, out);
}

int phr_bvh_build_batch(phr_gpu* g, const void* prims, uint32_t n, const uint32_t* seg_lo, uint32_t segs, uint32_t leaf,
                        phr_bvh** out) {
    PHR_CHECK(g && out && seg_lo && (prims || n == 0), PHR_ERR_ARG, "phr_bvh_build_batch: null argument");
    phr_array* a = nullptr;
    int rc = upload_prims(g, prims, n, &a);
    if (rc != PHR_OK) return rc;
    if ((rc = bvh_new(g, a, true, leaf, seg_lo, segs, nullptr, nullptr, PHR_LEAF, out)) != PHR_OK) phr_array_release(a);
    return rc;
}

int phr_bvh_build_batch_on(phr_gpu* g, phr_array* prims, const uint32_t* seg_lo, uint32_t segs, uint32_t leaf, phr_bvh** out) {
    PHR_CHECK(g && prims && out && seg_lo && prims->elem == 32, PHR_ERR_ARG,
              "phr_bvh_build_batch_on: prims must be an array of 32-byte boxes");
    return bvh_new(g, prims, false, leaf, seg_lo, segs, nullptr, nullptr, PHR_LEAF, out);
}

int phr_bvh_set_sort_tile(phr_bvh* b, uint32_t items) {
    PHR_CHECK(b, PHR_ERR_ARG, "phr_bvh_set_sort_tile: null");
    PHR_CHECK(items == 8 || items == 16, PHR_ERR_ARG, "phr_bvh_set_sort_tile: 8 or 16 keys a thread");
    b->items = items;
    return PHR_OK;
}

uint32_t phr_bvh_sort_tile(const phr_bvh* b) { return b ? b->items : 0u; }

int phr_bvh_autotune(phr_bvh* b, uint32_t reps, uint32_t* items, double* ms) {
    PHR_CHECK(b && reps > 0, PHR_ERR_ARG, "phr_bvh_autotune: null or no repetitions");
    hipStream_t s = b->gpu->stream;
    hipEvent_t e0, e1;
    PHR_HIP(hipEventCreate(&e0));
    PHR_HIP(hipEventCreate(&e1));
    const uint32_t cand[2] = {8, 16};
    float best = 3.0e38f;
    uint32_t pick = b->items;
    int rc = PHR_OK;
    for (int c = 0; c < 2 && rc == PHR_OK; ++c) {
        b->items = cand[c];
        if ((rc = bvh_enqueue_build(b)) != PHR_OK) break;  // warm
        float t = 0.f;
        if (hipEventRecord(e0, s) != hipSuccess) { rc = PHR_ERR_HIP; break; }
        for (uint32_t r = 0; r < reps && rc == PHR_OK; ++r) rc = bvh_enqueue_build(b);
        if (rc != PHR_OK) break;
        if (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&t, e0, e1) != hipSuccess) {
            rc = PHR_ERR_HIP;
            break;
        }
        t /= (float)reps;
        if (ms) ms[c] = t;
        if (t < best) { best = t; pick = cand[c]; }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc != PHR_OK) {
        if (rc == PHR_ERR_HIP) phr_set_error_msg("phr_bvh_autotune: timing failed");
        return rc;
    }
    b->items = pick;
    if (items) *items = pick;
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
    PHR_CHECK(b->segs == 1, PHR_ERR_ARG, "phr_bvh_download: a batch downloads one segment at a time");
    return phr_bvh_download_segment(b, 0, nodes, order, depth);
}

}  // extern "C"

// ---- a build that allocates nothing -----------------------------------------
// The scratch a one-tree build of at most PHR_STATIC_BUILD_MAX boxes needs,
// carved from one file-scope device array: every GPU the library runs on holds
// its own copy until the process ends. Calls on one GPU share it, so two such
// builds in flight on one device race (the library's stream serialises the
// calls it issues itself).
namespace {

constexpr uint64_t ST_ALIGN = 256;
constexpr uint64_t st_round(uint64_t b) { return (b + ST_ALIGN - 1) & ~(ST_ALIGN - 1); }
constexpr uint32_t ST_N = PHR_STATIC_BUILD_MAX;
constexpr uint32_t ST_TILES = (ST_N + OS_TILE - 1) / OS_TILE;
constexpr uint32_t ST_STATUS_AT = (Z_DEPTH + 1 + 63) & ~63u;
constexpr uint32_t ST_SCAN_STATUS = (ST_STATUS_AT + CODE_PASSES * ST_TILES * 256 + 1) & ~1u;
constexpr uint32_t ST_ZERO_WORDS = ST_SCAN_STATUS + 2 * ((ST_N + SC_TILE - 1) / SC_TILE) + 2;
constexpr uint32_t ST_LEVEL0 = (ST_N + RF_SLOTS - 1) / RF_SLOTS;
// Offsets of each buffer (bytes), as bvh_alloc sizes them for one segment.
constexpr uint64_t ST_SEG_LO = 0;
constexpr uint64_t ST_NODE_BASE = ST_SEG_LO + st_round(8);
constexpr uint64_t ST_CODES = ST_NODE_BASE + st_round(4);
constexpr uint64_t ST_KEYS_B = ST_CODES + st_round(4ull * ST_N);
constexpr uint64_t ST_VALS_B = ST_KEYS_B + st_round(4ull * ST_N);
constexpr uint64_t ST_ZERO = ST_VALS_B + st_round(4ull * ST_N);
constexpr uint64_t ST_ACC = ST_ZERO + st_round(4ull * ST_ZERO_WORDS);
constexpr uint64_t ST_PARTS = ST_ACC + st_round(24);
constexpr uint64_t ST_LINKS = ST_PARTS + st_round(32ull * BOUNDS_BLOCKS_MAX);
constexpr uint64_t ST_UP = ST_LINKS + st_round(16ull * ST_N);
constexpr uint64_t ST_CNT = ST_UP + st_round(32ull * ST_N);
constexpr uint64_t ST_LEAF_AT = ST_CNT + st_round(4ull * ST_N);
constexpr uint64_t ST_OWN_END = ST_LEAF_AT + st_round(4ull * ST_N);
constexpr uint64_t ST_FAR_ENDS = ST_OWN_END + st_round(4ull * ST_N);
constexpr uint64_t ST_FRONT0 = ST_FAR_ENDS + st_round(4ull * ST_N);
constexpr uint64_t ST_FRONT1 = ST_FRONT0 + st_round(4ull * ST_N);
constexpr uint64_t ST_FCOUNT0 = ST_FRONT1 + st_round(4ull * ST_N);
constexpr uint64_t ST_FCOUNT1 = ST_FCOUNT0 + st_round(4ull * ST_LEVEL0);
constexpr uint64_t ST_BYTES = ST_FCOUNT1 + st_round(4ull * ST_LEVEL0);
static_assert(ST_BYTES == PHR_STATIC_BUILD_SCRATCH_BYTES, "the header's scratch figure follows the carve");

}  // namespace

// The scratch itself. In host code this name is a stand-in, not a device
// address: the build takes each device's address of it from
// hipGetSymbolAddress, once per device, and passes only that to kernels.
static __device__ __align__(256) uint8_t scratch[ST_BYTES];

namespace {

constexpr int ST_MAX_DEVICES = 64;
std::mutex g_st_mutex;
uint8_t* g_st_ptr[ST_MAX_DEVICES] = {};

// This device's address of `scratch`, looked up on its first use with the
// device current.
int static_scratch(phr_gpu* g, uint8_t** out) {
    PHR_CHECK(g->device >= 0 && g->device < ST_MAX_DEVICES, PHR_ERR_UNSUPPORTED, "phr_bvh_build_static: device index too large");
    std::lock_guard<std::mutex> lock(g_st_mutex);
    if (!g_st_ptr[g->device]) {
        PHR_HIP(hipSetDevice(g->device));
        void* p = nullptr;
        PHR_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        PHR_CHECK(p, PHR_ERR_HIP, "phr_bvh_build_static: no device address for the scratch array");
        g_st_ptr[g->device] = (uint8_t*)p;
    }
    *out = g_st_ptr[g->device];
    return PHR_OK;
}

// What must hold before a build starts: the segment table of one tree, the
// scene-box words at Aabb::default(), the arrival counters of the wide refit
// levels at zero. A kernel, not a memset, so the call can be captured.
__global__ __launch_bounds__(256) void k_static_init(uint32_t n, uint32_t* __restrict__ seg_lo, uint32_t* __restrict__ node_base,
                                                     uint32_t* __restrict__ acc, uint32_t* __restrict__ cnt) {
    const uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x, gsize = gridDim.x * blockDim.x;
    if (gtid == 0) {
        seg_lo[0] = 0u;
        seg_lo[1] = n;
        node_base[0] = 0u;
        for (int j = 0; j < 6; ++j) acc[j] = j < 3 ? KEY_MAX : KEY_MIN;
    }
    for (uint32_t i = gtid; i < n; i += gsize) cnt[i] = 0u;
}

// The depth the packing launch left, or 0 for an empty tree (no zero block).
__global__ void k_static_depth(const uint32_t* __restrict__ zero, uint32_t* __restrict__ depth) {
    if (threadIdx.x == 0) depth[0] = zero ? zero[Z_DEPTH] : 0u;
}

}  // namespace

extern "C" {

int phr_bvh_build_static(phr_gpu* g, const void* prims, uint32_t n, uint32_t leaf, void* nodes, uint32_t* order,
                         uint32_t* depth) {
    PHR_CHECK(g && (n == 0 || (prims && nodes && order)), PHR_ERR_ARG, "phr_bvh_build_static: null argument");
    PHR_CHECK(n <= PHR_STATIC_BUILD_MAX, PHR_ERR_SIZE,
              "phr_bvh_build_static: more boxes than the static scratch holds (PHR_STATIC_BUILD_MAX = 262,144)");
    PHR_CHECK(leaf >= 1, PHR_ERR_ARG, "phr_bvh_build_static: leaf size must be at least 1");
    uint8_t* base = nullptr;
    int rc = static_scratch(g, &base);
    if (rc != PHR_OK) return rc;
    hipStream_t s = g->stream;
    if (n == 0) {
        // An empty tree: nothing to build, depth 0.
        if (depth) hipLaunchKernelGGL(k_static_depth, dim3(1), dim3(64), 0, s, (const uint32_t*)nullptr, depth);
        PHR_HIP(hipGetLastError());
        return PHR_OK;
    }
    // The build's state over caller memory and the scratch; nothing here
    // reaches the allocator (the vectors stay empty).
    phr_array a_prims, a_order, a_nodes;
    a_prims.gpu = a_order.gpu = a_nodes.gpu = g;
    a_prims.ptr = const_cast<void*>(prims);
    a_prims.elem = 32;
    a_prims.len = a_prims.cap = n;
    a_order.ptr = order;
    a_order.elem = 4;
    a_order.len = a_order.cap = n;
    a_nodes.ptr = nodes;
    a_nodes.elem = 32;
    a_nodes.len = a_nodes.cap = 2ull * n - 1;
    phr_bvh b;
    b.gpu = g;
    b.n = n;
    b.leaf = leaf;
    b.segs = 1;
    b.nonempty = 1;
    b.seg_passes = 0;
    b.tiles = (n + OS_TILE - 1) / OS_TILE;
    b.bounds_blocks = std::min<uint32_t>(BOUNDS_BLOCKS_MAX, std::max<uint32_t>(1, (n + 2047) / 2048));
    b.status_at = ST_STATUS_AT;
    b.scan_status = (b.status_at + CODE_PASSES * b.tiles * 256 + 1) & ~1u;
    b.zero_words = b.scan_status + 2 * ((n + SC_TILE - 1) / SC_TILE) + 2;
    b.prims = &a_prims;
    b.order = &a_order;
    b.nodes = &a_nodes;
    b.own_prims = false;
    b.own_nodes = false;
    b.seg_lo = (uint32_t*)(base + ST_SEG_LO);
    b.node_base = (uint32_t*)(base + ST_NODE_BASE);
    b.codes = (uint32_t*)(base + ST_CODES);
    b.keys_b = (uint32_t*)(base + ST_KEYS_B);
    b.vals_b = (uint32_t*)(base + ST_VALS_B);
    b.zero = (uint32_t*)(base + ST_ZERO);
    b.acc = (uint32_t*)(base + ST_ACC);
    b.parts = base + ST_PARTS;
    b.links = (PhrLink*)(base + ST_LINKS);
    b.up = (uint4*)(base + ST_UP);
    b.cnt = (uint32_t*)(base + ST_CNT);
    b.leaf_at = (uint32_t*)(base + ST_LEAF_AT);
    b.own_end = (uint32_t*)(base + ST_OWN_END);
    b.far_ends = (uint32_t*)(base + ST_FAR_ENDS);
    b.front[0] = (uint32_t*)(base + ST_FRONT0);
    b.front[1] = (uint32_t*)(base + ST_FRONT1);
    b.fcount[0] = (uint32_t*)(base + ST_FCOUNT0);
    b.fcount[1] = (uint32_t*)(base + ST_FCOUNT1);
    b.level0_blocks = (n + RF_SLOTS - 1) / RF_SLOTS;
    hipLaunchKernelGGL(k_static_init, dim3(std::min<uint32_t>((n + 255) / 256, 1024)), dim3(256), 0, s, n, b.seg_lo, b.node_base,
                       b.acc, b.cnt);
    PHR_HIP(hipGetLastError());
    rc = bvh_enqueue_build(&b);
    // b's members point at the caller's memory and the scratch: nothing to free.
    b.prims = b.order = b.nodes = nullptr;
    b.seg_lo = b.node_base = b.codes = b.keys_b = b.vals_b = b.zero = b.acc = b.cnt = b.leaf_at = b.own_end = b.far_ends = nullptr;
    b.parts = nullptr;
    b.links = nullptr;
    b.up = nullptr;
    b.front[0] = b.front[1] = b.fcount[0] = b.fcount[1] = nullptr;
    if (rc != PHR_OK) return rc;
    if (depth) hipLaunchKernelGGL(k_static_depth, dim3(1), dim3(64), 0, s, (const uint32_t*)(base + ST_ZERO), depth);
    PHR_HIP(hipGetLastError());
    return PHR_OK;
}

int phr_sort_record(phr_gpu* g, phr_array* codes, uint32_t n, phr_sort** out);
int phr_sort_run(phr_sort* s);
void phr_sort_free(phr_sort* s);
phr_array* phr_sort_code(phr_sort* s);
static const uint32_t* sort_order_dev(phr_sort* s);

// Bvh::build_compact (src/bvh/build.rs:104-136) of this tree's primitives: the
// same reachable tree numbered breadth first, children adjacent, computed on
// the device from the build (ranks of (depth, first slot) by the onesweep
// sort) and copied out. The order and depth are the build's.
int phr_bvh_compact(phr_bvh* b, void* nodes, uint64_t cap, uint32_t* count) {
    PHR_CHECK(b && count, PHR_ERR_ARG, "phr_bvh_compact: null");
    PHR_CHECK(b->segs == 1 && b->leaf_flag == PHR_LEAF, PHR_ERR_ARG, "phr_bvh_compact: one tree at a time");
    const uint32_t n = b->n;
    *count = 0;
    if (n == 0) return PHR_OK;
    PHR_CHECK(n < (1u << CP_FIRST_BITS), PHR_ERR_SIZE, "phr_bvh_compact: at most 2^26 primitives");
    hipStream_t s = b->gpu->stream;
    phr_array* keys = nullptr;
    phr_sort* srt = nullptr;
    uint32_t* aux = nullptr;  // count, then the inverse map (2n)
    float4* out = nullptr;
    int rc = phr_array_alloc(b->gpu, 4, 2ull * n, &keys);
    auto done = [&](int code) {
        if (srt) phr_sort_free(srt);
        phr_array_release(keys);
        if (aux) (void)hipFree(aux);
        if (out) (void)hipFree(out);
        return code;
    };
    if (rc != PHR_OK) return done(rc);
    if (hipMalloc((void**)&aux, 4ull * (2ull * n + 1)) != hipSuccess || hipMalloc((void**)&out, 32ull * (2ull * n - 1)) != hipSuccess) {
        phr_set_error_msg("phr_bvh_compact: allocation failed");
        return done(PHR_ERR_HIP);
    }
    if (hipMemsetAsync(aux, 0, 4, s) != hipSuccess) return done(PHR_ERR_HIP);
    hipLaunchKernelGGL(k_compact_keys, dim3((2 * n + 255) / 256), dim3(256), 0, s, b->links, b->up, (const uint32_t*)b->keys_b, n,
                       b->leaf, (uint32_t*)keys->ptr, aux);
    if ((rc = phr_sort_record(b->gpu, keys, 2 * n, &srt)) != PHR_OK || (rc = phr_sort_run(srt)) != PHR_OK) return done(rc);
    uint32_t r = 0;
    if (hipMemcpyAsync(&r, aux, 4, hipMemcpyDeviceToHost, s) != hipSuccess || hipStreamSynchronize(s) != hipSuccess) {
        phr_set_error_msg("phr_bvh_compact: readback failed");
        return done(PHR_ERR_HIP);
    }
    const uint32_t* ids = sort_order_dev(srt);
    const uint32_t* skeys = (const uint32_t*)phr_sort_code(srt)->ptr;
    hipLaunchKernelGGL(k_compact_inv, dim3((r + 255) / 256), dim3(256), 0, s, ids, r, aux + 1);
    hipLaunchKernelGGL(k_compact_write, dim3((r + 255) / 256), dim3(256), 0, s, ids, skeys, r, b->links, aux + 1, n, b->leaf,
                       (const float4*)b->nodes->ptr, out);
    if (hipGetLastError() != hipSuccess) {
        phr_set_error_msg("phr_bvh_compact: launch failed");
        return done(PHR_ERR_HIP);
    }
    *count = r;
    if (nodes) {
        if (cap < r) {
            phr_set_error_msg("phr_bvh_compact: node buffer too small");
            return done(PHR_ERR_SIZE);
        }
        if (hipMemcpyAsync(nodes, out, 32ull * r, hipMemcpyDeviceToHost, s) != hipSuccess || hipStreamSynchronize(s) != hipSuccess) {
            phr_set_error_msg("phr_bvh_compact: readback failed");
            return done(PHR_ERR_HIP);
        }
    } else if (hipStreamSynchronize(s) != hipSuccess) {
        return done(PHR_ERR_HIP);
    }
    return done(PHR_OK);
}

// One segment's tree as Bvh::build of its boxes alone would hold it: node
// indices, leaf ranges and the order all relative to the segment.
int phr_bvh_download_segment(phr_bvh* b, uint32_t seg, void* nodes, uint32_t* order, uint32_t* depth) {
    PHR_CHECK(b && seg < b->segs, PHR_ERR_ARG, "phr_bvh_download_segment: no such segment");
    hipStream_t s = b->gpu->stream;
    const uint32_t lo = b->seg_lo_h[seg], ns = b->seg_lo_h[seg + 1] - lo, base = b->node_base_h[seg];
    if (ns) {
        const uint64_t cnt = 2ull * ns - 1;
        if (nodes) PHR_HIP(hipMemcpyAsync(nodes, (char*)b->nodes->ptr + 32ull * base, 32ull * cnt, hipMemcpyDeviceToHost, s));
        if (order) PHR_HIP(hipMemcpyAsync(order, (uint32_t*)b->order->ptr + lo, 4ull * ns, hipMemcpyDeviceToHost, s));
        PHR_HIP(hipStreamSynchronize(s));
        if (nodes) {
            uint32_t* w = (uint32_t*)nodes;
            for (uint64_t k = 0; k < cnt; ++k) {
                uint32_t& a = w[8 * k + 3];
                uint32_t& c = w[8 * k + 7];
                if (a & PHR_LEAF) {
                    a = (((a & ~b->leaf_flag) - lo) | PHR_LEAF);
                    c -= lo;
                } else {
                    a -= base;
                    c -= base;
                }
            }
        }
        if (order)
            for (uint32_t k = 0; k < ns; ++k) order[k] -= lo;
    }
    if (depth) *depth = b->depth_h[seg];
    return PHR_OK;
}

int phr_bvh_codes(phr_bvh* b, uint32_t* sorted_codes) {
    PHR_CHECK(b && sorted_codes, PHR_ERR_ARG, "phr_bvh_codes: null");
    if (b->n) PHR_HIP(hipMemcpyAsync(sorted_codes, b->codes, 4ull * b->n, hipMemcpyDeviceToHost, b->gpu->stream));
    PHR_HIP(hipStreamSynchronize(b->gpu->stream));
    return PHR_OK;
}

uint32_t phr_bvh_len(const phr_bvh* b) { return b ? b->n : 0; }
uint32_t phr_bvh_depth(const phr_bvh* b) {
    if (!b) return 0;
    uint32_t d = 0;
    for (uint32_t x : b->depth_h) d = std::max(d, x);
    return d;
}
uint32_t phr_bvh_segments(const phr_bvh* b) { return b ? b->segs : 0; }
uint32_t phr_bvh_segment_depth(const phr_bvh* b, uint32_t seg) { return b && seg < b->segs ? b->depth_h[seg] : 0; }
uint32_t phr_bvh_leaf(const phr_bvh* b) { return b ? b->leaf : 0; }
phr_array* phr_bvh_nodes(phr_bvh* b) { return b ? b->nodes : nullptr; }
phr_array* phr_bvh_order(phr_bvh* b) { return b ? b->order : nullptr; }
phr_array* phr_bvh_prims(phr_bvh* b) { return b ? b->prims : nullptr; }
void phr_bvh_free(phr_bvh* b) { bvh_release(b); }

// GpuSort (src/bvh/sort.rs:13-111): a stable sort of n 32-bit codes with
// their slots, recorded once over a device array of codes and replayed. The
// keys come out as (original slot, code) in (code, slot) order - the order of
// the reference's 64-bit LSD sort - and the codes alone beside them.
struct phr_sort {
    phr_gpu* gpu = nullptr;
    phr_array* codes = nullptr;  // borrowed
    phr_array* keys = nullptr;   // n x (slot, code)
    phr_array* code = nullptr;   // n sorted codes
    uint32_t n = 0, tiles = 0, zero_words = 0;
    uint32_t *order = nullptr, *kb = nullptr, *vb = nullptr, *zero = nullptr;
};

static const uint32_t* sort_order_dev(phr_sort* s) { return s->order; }

static void sort_release(phr_sort* s) {
    if (!s) return;
    phr_array_release(s->keys);
    phr_array_release(s->code);
    for (void* p : {(void*)s->order, (void*)s->kb, (void*)s->vb, (void*)s->zero})
        if (p) (void)hipFree(p);
    delete s;
}

int phr_sort_record(phr_gpu* g, phr_array* codes, uint32_t n, phr_sort** out) {
    PHR_CHECK(g && codes && out && codes->elem == 4 && codes->len >= n, PHR_ERR_ARG,
              "phr_sort_record: codes must be a u32 array of at least n items");
    phr_sort* s = new phr_sort();
    s->gpu = g;
    s->codes = codes;
    s->n = n;
    s->tiles = (n + OS_TILE - 1) / OS_TILE;
    s->zero_words = Z_TICKET + 8 + CODE_PASSES * s->tiles * 256;
    const uint64_t nn = std::max<uint32_t>(n, 1);
    int rc;
    if ((rc = phr_array_alloc(g, 8, n, &s->keys)) != PHR_OK || (rc = phr_array_alloc(g, 4, n, &s->code)) != PHR_OK) {
        sort_release(s);
        return rc;
    }
    bool ok = hipMalloc((void**)&s->order, 4 * nn) == hipSuccess && hipMalloc((void**)&s->kb, 4 * nn) == hipSuccess &&
              hipMalloc((void**)&s->vb, 4 * nn) == hipSuccess && hipMalloc((void**)&s->zero, 4ull * s->zero_words) == hipSuccess &&
              hipMemsetAsync(s->zero, 0, 4ull * s->zero_words, g->stream) == hipSuccess &&
              hipStreamSynchronize(g->stream) == hipSuccess;
    if (!ok) {
        sort_release(s);
        phr_set_error_msg("phr_sort_record: allocation failed");
        return PHR_ERR_HIP;
    }
    *out = s;
    return PHR_OK;
}

// GpuSort::run (src/bvh/sort.rs:61-63): histograms, four stable onesweep
// passes (8 bits each), then the keys; enqueued.
int phr_sort_run(phr_sort* s) {
    PHR_CHECK(s, PHR_ERR_ARG, "phr_sort_run: null");
    const uint32_t n = s->n;
    if (n == 0) return PHR_OK;
    hipStream_t st = s->gpu->stream;
    uint32_t* hist = s->zero + Z_HIST;
    uint32_t* ticket = s->zero + Z_TICKET;
    uint32_t* status = s->zero + Z_TICKET + 8;
    hipLaunchKernelGGL(k_sort_hist, dim3(s->tiles), dim3(256), 0, st, (const uint32_t*)s->codes->ptr, n, hist);
    const uint32_t* kin = (const uint32_t*)s->codes->ptr;
    const uint32_t* vin = nullptr;
    uint32_t* sorted = (uint32_t*)s->code->ptr;
    for (uint32_t p = 0; p < CODE_PASSES; ++p) {
        const bool to_final = (p % 2) == 1;  // passes 1 and 3 land in (code, order)
        uint32_t* kout = to_final ? sorted : s->kb;
        uint32_t* vout = to_final ? s->order : s->vb;
        uint32_t* stp = status + (size_t)p * s->tiles * 256;
        if (p == 0)
            hipLaunchKernelGGL((k_onesweep<true, false, OS_ITEMS>), dim3(s->tiles), dim3(OS_THREADS), 0, st, kin, vin, kout, vout, hist, stp,
                               ticket, n, 0u, (const uint32_t*)nullptr, 0u);
        else
            hipLaunchKernelGGL((k_onesweep<false, false, OS_ITEMS>), dim3(s->tiles), dim3(OS_THREADS), 0, st, kin, vin, kout, vout,
                               hist + 256 * p, stp, ticket + p, n, 8 * p, (const uint32_t*)nullptr, 0u);
        kin = kout;
        vin = vout;
    }
    const uint32_t blocks = std::min<uint32_t>(1024, std::max<uint32_t>(1, (n + 255) / 256));
    hipLaunchKernelGGL(k_sort_keys, dim3(blocks), dim3(256), 0, st, (const uint32_t*)s->order, (const uint32_t*)sorted, n,
                       (uint2*)s->keys->ptr, s->zero, s->zero_words);
    PHR_HIP(hipGetLastError());
    return PHR_OK;
}

phr_array* phr_sort_keys(phr_sort* s) { return s ? s->keys : nullptr; }
phr_array* phr_sort_code(phr_sort* s) { return s ? s->code : nullptr; }
uint32_t phr_sort_len(const phr_sort* s) { return s ? s->n : 0; }
uint32_t phr_sort_steps(const phr_sort* s) { return s ? 2 + CODE_PASSES : 0; }
uint64_t phr_sort_bytes(const phr_sort* s) {
    return s ? 8ull * s->n + 4ull * s->n * 4 + 4ull * s->zero_words : 0;
}

// GpuSort::order / read (src/bvh/sort.rs:99-111), blocking.
int phr_sort_read(phr_sort* s, uint32_t* order, uint32_t* sorted) {
    PHR_CHECK(s, PHR_ERR_ARG, "phr_sort_read: null");
    hipStream_t st = s->gpu->stream;
    if (s->n) {
        if (order) PHR_HIP(hipMemcpyAsync(order, s->order, 4ull * s->n, hipMemcpyDeviceToHost, st));
        if (sorted) PHR_HIP(hipMemcpyAsync(sorted, s->code->ptr, 4ull * s->n, hipMemcpyDeviceToHost, st));
    }
    PHR_HIP(hipStreamSynchronize(st));
    return PHR_OK;
}

void phr_sort_free(phr_sort* s) { sort_release(s); }

}  // extern "C"

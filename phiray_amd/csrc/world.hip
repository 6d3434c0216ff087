// phiray's rigid-body World (src/sim/world.rs) on the device: XPBD substeps
// (Muller et al. 2020, "Detailed Rigid Body Simulation with Extended
// Position Based Dynamics") for spheres, boxes and capsules, every stage a
// kernel. World::poses feeds Tracer::reinstance.
//
// A step of one frame:
//
//  1. k_sweep     one thread a body: the shape's box at its pose, swept by the
//                 motion over the frame and padded (broad.rs:52-70).
//  2. LBVH        this library's build over those boxes (lbvh.hip).
//  3. k_pairs     one thread a body walks the tree. A pair belongs to the
//                 lower-numbered of its bodies that moves, which lists it in
//                 its own run of PAIR_SLOTS slots, kept sorted by the other
//                 body: no pair is listed twice and the lists do not depend on
//                 the tree's shape or the threads' timing.
//  4. k_contacts  one thread a listed pair: sphere, box and capsule pairs in
//                 closed form, box pairs by separating axes with the incident
//                 face clipped to the reference face; at most four points a
//                 pair, each kept as body-local anchors (the substeps
//                 re-evaluate them) and written to the pair's own slot.
//  5. k_solve     ONE workgroup runs the rest of the frame: the bodies' state
//                 lives in LDS (64 B a body, up to MAX_BODIES), the pairs that
//                 touch a body that moves are coloured so that no body is in
//                 two pairs of one colour (rounds of a matching: a pair joins
//                 the round's colour when it comes first, in phiray's pair
//                 order, among the uncoloured pairs at each of its moving
//                 bodies), and every substep projects the colours one after
//                 the other, the pairs of a colour in parallel. That is
//                 Gauss-Seidel over the pairs in phiray's order, without a
//                 launch or a grid-wide barrier between colours. The points
//                 of one pair are solved together (contact.rs:1-11: each asks
//                 from the same state, the sum is applied once, scaled so that
//                 no point gets more than it asked).
//
// The contact rules are phiray's, kept alike because its closed-form tests
// and the oracle comparison depend on them: speculative contacts out to the
// margins + 1 cm + how far the pair can close in a frame; the penetration a
// projection pushes out capped at max_recover * h / iterations, and the
// recovered part kept out of the derived velocity; static friction at
// position level within mu * |lambda_n|; dynamic friction and restitution
// on the velocities, no bounce under a closing speed of 2 g h; kick-drift-
// kick gravity; Blender's damping and deactivation.
//
// Bodies cross the C ABI in phiray's 64-byte Body record (state.rs).
#include "phr_runtime.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace {

constexpr uint32_t MAX_BODIES = 2048;    // the bodies' state in LDS: 128 KB
constexpr uint32_t LPUSH_BODIES = 1280;  // bodies whose state and pushed-out offsets fit LDS together (120 KB)
constexpr uint32_t DATAFLOW_ROUNDS = 1u << 24;  // a bound far past any wait of the dataflow solve
constexpr uint32_t PAIR_SLOTS = 32;      // pairs a body may own
constexpr uint32_t MAX_COLOURS = 1024;
constexpr uint32_t SOLVE_THREADS = 512;
constexpr float SLOP = 0.01f;            // narrow/mod.rs:25
constexpr float PAD = SLOP + 0.005f;     // the broadphase's padding past the margin
enum : uint32_t { K_ACTIVE = 0, K_PASSIVE = 1, K_KINEMATIC = 2 };
enum : uint32_t { S_BOX = 0, S_SPHERE = 1, S_CAPSULE = 2 };
constexpr uint32_t KIND_MASK = 3u, SLEEPING = 1u << 2, DEACTIVATION = 1u << 3;
constexpr uint32_t NONE = 0xffffffffu;

// ---- records -------------------------------------------------------------------
struct Body {  // phiray's Body, src/sim/state.rs: 64 bytes
    float x[3];
    float inv_mass;
    float q[4];
    float v[3];
    uint32_t shape;
    float w[3];
    uint32_t flags;
};
struct Konst {  // what the solver reads of a body and never writes
    float inv_inertia[3], friction;
    float com[3], restitution;
    float lin_damp, ang_damp, margin;
    uint32_t groups;
};
struct Aux {  // per body, kept across frames
    float sleep_lin, sleep_ang, timer;
    uint32_t obj;
    float scale[3], pad;
};
// A listed pair: bodies a < b, its points, and the normal from a to b.
struct PairRec {
    uint32_t a, b, m, pad;
    float4 n;
};
// One point: anchors in a's and b's frames, the accumulated multipliers and
// the closing speed at the substep's start.
struct PointRec {
    float4 aa;  // anchor on a, lambda_n
    float4 ab;  // anchor on b, lambda_t
    float4 e;   // vn0, depth when found, -, -
};
static_assert(sizeof(Body) == 64 && sizeof(Konst) == 48 && sizeof(PairRec) == 32 && sizeof(PointRec) == 48, "records");

// ---- vector and rotation arithmetic --------------------------------------------------
struct V {
    float x, y, z;
};
__host__ __device__ __forceinline__ V mk(float x, float y, float z) { return V{x, y, z}; }
__host__ __device__ __forceinline__ V operator+(V a, V b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
__host__ __device__ __forceinline__ V operator-(V a, V b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
__host__ __device__ __forceinline__ V operator-(V a) { return mk(-a.x, -a.y, -a.z); }
__host__ __device__ __forceinline__ V operator*(V a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
__host__ __device__ __forceinline__ float dot(V a, V b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__host__ __device__ __forceinline__ V cross(V a, V b) { return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
__host__ __device__ __forceinline__ V hadamard(V a, V b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
__host__ __device__ __forceinline__ float norm2(V a) { return dot(a, a); }
__host__ __device__ __forceinline__ float length(V a) { return sqrtf(norm2(a)); }
__host__ __device__ __forceinline__ float comp(V a, int k) { return k == 0 ? a.x : (k == 1 ? a.y : a.z); }
__host__ __device__ __forceinline__ V unit_axis(int k, float s) { return mk(k == 0 ? s : 0.f, k == 1 ? s : 0.f, k == 2 ? s : 0.f); }

struct Q {
    float x, y, z, w;
};
// q v q^-1 for a unit quaternion: v + 2 w (u x v) + 2 u x (u x v).
__host__ __device__ __forceinline__ V rot(Q q, V v) {
    const V u = mk(q.x, q.y, q.z);
    const V t = cross(u, v) * 2.0f;
    return v + t * q.w + cross(u, t);
}
__host__ __device__ __forceinline__ V unrot(Q q, V v) { return rot(Q{-q.x, -q.y, -q.z, q.w}, v); }
__host__ __device__ __forceinline__ Q qmul(Q a, Q b) {
    return Q{a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y, a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x,
             a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w, a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z};
}
__host__ __device__ __forceinline__ Q qunit(Q q) {
    const float l = sqrtf(q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w);
    return l > 0.f ? Q{q.x / l, q.y / l, q.z / l, q.w / l} : Q{0.f, 0.f, 0.f, 1.f};
}
// The orientation after turning at angular velocity om for time h (first
// order, renormalised; Muller 2020 eq. 8).
__host__ __device__ __forceinline__ Q qstep(Q q, V om, float h) {
    const Q d = qmul(Q{om.x, om.y, om.z, 0.f}, q);
    const float s = 0.5f * h;
    return qunit(Q{q.x + d.x * s, q.y + d.y * s, q.z + d.z * s, q.w + d.w * s});
}
// The angular velocity that turns `from` into `to` over h (Muller 2020 eq. 9).
__host__ __device__ __forceinline__ V qrate(Q to, Q from, float h) {
    const Q d = qmul(to, Q{-from.x, -from.y, -from.z, from.w});
    const float s = (d.w >= 0.f ? 2.0f : -2.0f) / h;
    return mk(d.x * s, d.y * s, d.z * s);
}

// ---- a body's state in LDS: (x, inv_mass), q, (v, shape), (w, flags) -----------------------
struct Lds {
    float4* s;
    __host__ __device__ __forceinline__ V x(uint32_t i) const { const float4 r = s[4 * i]; return mk(r.x, r.y, r.z); }
    __host__ __device__ __forceinline__ float im(uint32_t i) const { return s[4 * i].w; }
    __host__ __device__ __forceinline__ Q q(uint32_t i) const { const float4 r = s[4 * i + 1]; return Q{r.x, r.y, r.z, r.w}; }
    __host__ __device__ __forceinline__ V v(uint32_t i) const { const float4 r = s[4 * i + 2]; return mk(r.x, r.y, r.z); }
    __host__ __device__ __forceinline__ V w(uint32_t i) const { const float4 r = s[4 * i + 3]; return mk(r.x, r.y, r.z); }
    __host__ __device__ __forceinline__ void set_x(uint32_t i, V a) { s[4 * i].x = a.x; s[4 * i].y = a.y; s[4 * i].z = a.z; }
    __host__ __device__ __forceinline__ void set_q(uint32_t i, Q a) { s[4 * i + 1] = make_float4(a.x, a.y, a.z, a.w); }
    __host__ __device__ __forceinline__ void set_v(uint32_t i, V a) { s[4 * i + 2].x = a.x; s[4 * i + 2].y = a.y; s[4 * i + 2].z = a.z; }
    __host__ __device__ __forceinline__ void set_w(uint32_t i, V a) { s[4 * i + 3].x = a.x; s[4 * i + 3].y = a.y; s[4 * i + 3].z = a.z; }
};

__host__ __device__ __forceinline__ uint32_t kind_of(uint32_t flags) {
    const uint32_t k = flags & KIND_MASK;
    return k == 1u ? K_PASSIVE : (k == 2u ? K_KINEMATIC : K_ACTIVE);
}
__host__ __device__ __forceinline__ bool moves_f(uint32_t f) { return kind_of(f) != K_PASSIVE && (f & SLEEPING) == 0u; }
__host__ __device__ __forceinline__ bool dynamic_f(uint32_t f) { return kind_of(f) == K_ACTIVE && (f & SLEEPING) == 0u; }

__host__ __device__ __forceinline__ V ld3(const float* p) { return mk(p[0], p[1], p[2]); }

// ---- 1. swept boxes ----------------------------------------------------------------
__host__ __device__ V shape_reach(const phr_shape& s, Q q) {
    // Half the extent of the shape's world box along each axis.
    if (s.kind == S_SPHERE) return mk(s.radius, s.radius, s.radius);
    if (s.kind == S_CAPSULE) {
        const V a = rot(q, mk(0.f, 0.f, s.half[2]));
        return mk(fabsf(a.x) + s.radius, fabsf(a.y) + s.radius, fabsf(a.z) + s.radius);
    }
    const V e0 = rot(q, mk(s.half[0], 0.f, 0.f)), e1 = rot(q, mk(0.f, s.half[1], 0.f)), e2 = rot(q, mk(0.f, 0.f, s.half[2]));
    return mk(fabsf(e0.x) + fabsf(e1.x) + fabsf(e2.x), fabsf(e0.y) + fabsf(e1.y) + fabsf(e2.y), fabsf(e0.z) + fabsf(e1.z) + fabsf(e2.z));
}

__global__ __launch_bounds__(256) void k_sweep(const Body* __restrict__ bodies, const Konst* __restrict__ konst,
                                               const phr_shape* __restrict__ shapes, uint32_t n, float dt,
                                               float4* __restrict__ boxes) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const Body b = bodies[i];
    const Q q{b.q[0], b.q[1], b.q[2], b.q[3]};
    const V x = ld3(b.x), r = shape_reach(shapes[b.shape], q), sh = ld3(b.v) * dt;
    const float pad = konst[i].margin + PAD;
    const V lo = x - r, hi = x + r;
    boxes[2 * i] = make_float4(fminf(lo.x, lo.x + sh.x) - pad, fminf(lo.y, lo.y + sh.y) - pad, fminf(lo.z, lo.z + sh.z) - pad, 0.f);
    boxes[2 * i + 1] = make_float4(fmaxf(hi.x, hi.x + sh.x) + pad, fmaxf(hi.y, hi.y + sh.y) + pad, fmaxf(hi.z, hi.z + sh.z) + pad, 0.f);
}

// ---- 3. pairs: a walk a body, each pair listed once by its owner ----------------------------
// Pair {i, j} is wanted when one of the two moves, one is active, their
// collision groups meet and their swept boxes overlap (broad.rs:73-91).
// Its owner is the lower-numbered body that moves.
__global__ __launch_bounds__(256) void k_pairs(const float4* __restrict__ nodes, const uint32_t* __restrict__ order,
                                               const float4* __restrict__ boxes, const Body* __restrict__ bodies,
                                               const Konst* __restrict__ konst, uint32_t n, uint32_t* __restrict__ count,
                                               uint32_t* __restrict__ other, uint32_t* __restrict__ overflow) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const uint32_t fi = bodies[i].flags;
    uint32_t found = 0;
    uint32_t* mine = other + (size_t)i * PAIR_SLOTS;
    if (moves_f(fi)) {
        const float4 blo = boxes[2 * i], bhi = boxes[2 * i + 1];
        const uint32_t gi = konst[i].groups;
        const bool act_i = kind_of(fi) == K_ACTIVE;
        auto meets = [&](float4 lo, float4 hi) {
            return lo.x <= bhi.x && hi.x >= blo.x && lo.y <= bhi.y && hi.y >= blo.y && lo.z <= bhi.z && hi.z >= blo.z;
        };
        uint32_t stack[PHR_STACK];
        int sp = 0;
        stack[sp++] = 0;
        while (sp > 0) {
            const uint32_t k = stack[--sp];
            const float4 lo = nodes[2 * (size_t)k], hi = nodes[2 * (size_t)k + 1];
            if (!meets(lo, hi)) continue;
            const uint32_t c0 = __float_as_uint(lo.w), c1 = __float_as_uint(hi.w);
            if (c0 & PHR_LEAF) {
                for (uint32_t s = c0 & ~PHR_LEAF; s < c1; ++s) {
                    const uint32_t j = order[s];
                    if (j == i) continue;
                    const uint32_t fj = bodies[j].flags;
                    if (moves_f(fj) && j < i) continue;  // j owns it
                    if (!act_i && kind_of(fj) != K_ACTIVE) continue;
                    if ((gi & konst[j].groups) == 0u) continue;
                    if (!meets(boxes[2 * (size_t)j], boxes[2 * (size_t)j + 1])) continue;
                    if (found == PAIR_SLOTS) {
                        atomicOr(overflow, 1u);
                        continue;
                    }
                    // Insert j keeping the run sorted.
                    uint32_t at = found++;
                    while (at > 0 && mine[at - 1] > j) {
                        mine[at] = mine[at - 1];
                        --at;
                    }
                    mine[at] = j;
                }
            } else if (sp + 2 <= (int)PHR_STACK) {  // an LBVH is at most 31 deep: never full
                stack[sp++] = c0;
                stack[sp++] = c1;
            }
        }
    }
    count[i] = found;
}

// The same walk with the tree, the swept boxes and the bodies' flags and groups
// in LDS: a block of one wave copies them in, then each lane walks for its
// body with its stack and its pair run in LDS too. Every visit is an LDS
// read instead of an L2 round trip. For worlds of at most LDS_PAIR_BODIES.
constexpr uint32_t LDS_PAIR_BODIES = 1280;
constexpr uint32_t PAIR_STAGE_THREADS = 256;  // k_pairs_lds: threads that stage the tree
constexpr uint32_t PAIR_BLOCK_BODIES = 8;     // bodies a block walks, PAIR_BLOCK_BODIES / 4 lanes of each wave
__global__ __launch_bounds__(PAIR_STAGE_THREADS) void k_pairs_lds(const float4* __restrict__ nodes, const uint32_t* __restrict__ order,
                                                  const float4* __restrict__ boxes, const Body* __restrict__ bodies,
                                                  const Konst* __restrict__ konst, uint32_t n, uint32_t* __restrict__ count,
                                                  uint32_t* __restrict__ other, uint32_t* __restrict__ overflow) {
    __shared__ float4 s_nodes[2 * (2 * LDS_PAIR_BODIES - 1)];
    __shared__ float4 s_boxes[2 * LDS_PAIR_BODIES];
    __shared__ uint32_t s_order[LDS_PAIR_BODIES], s_flags[LDS_PAIR_BODIES], s_groups[LDS_PAIR_BODIES];
    __shared__ uint32_t s_stack[PHR_STACK * 64];
    __shared__ uint32_t s_mine[PAIR_SLOTS * 64];
    const uint32_t tid = threadIdx.x;
    const uint32_t nn = 2 * n - 1;
    // Eight loads in flight a lane: four waves copy ~110 KB, then walk a
    // quarter of the block's bodies each (a wave walks as long as its
    // longest walk).
    constexpr uint32_t T = PAIR_STAGE_THREADS;
    auto stage = [&](float4* dst, const float4* src, uint32_t len) {
        for (uint32_t k = tid; k < len; k += T * 8) {
            float4 v[8];
#pragma unroll
            for (uint32_t j = 0; j < 8; ++j)
                if (k + T * j < len) v[j] = src[k + T * j];
#pragma unroll
            for (uint32_t j = 0; j < 8; ++j)
                if (k + T * j < len) dst[k + T * j] = v[j];
        }
    };
    stage(s_nodes, nodes, 2 * nn);
    stage(s_boxes, boxes, 2 * n);
    for (uint32_t k = tid; k < n; k += T) {
        s_order[k] = order[k];
        s_flags[k] = bodies[k].flags;
        s_groups[k] = konst[k].groups;
    }
    __syncthreads();
    constexpr uint32_t WL = PAIR_BLOCK_BODIES / 4;  // walking lanes a wave
    if ((tid & 63u) >= WL) return;
    const uint32_t slot = (tid >> 6) * WL + (tid & 63u);  // the body's column in s_stack and s_mine
    const uint32_t i = blockIdx.x * PAIR_BLOCK_BODIES + slot;
    if (i >= n) return;
    const uint32_t fi = s_flags[i];
    uint32_t found = 0;
    uint32_t* mine = s_mine + slot;  // slot k at mine[64 k]
    if (moves_f(fi)) {
        const float4 blo = s_boxes[2 * i], bhi = s_boxes[2 * i + 1];
        const uint32_t gi = s_groups[i];
        const bool act_i = kind_of(fi) == K_ACTIVE;
        auto meets = [&](float4 lo, float4 hi) {
            return lo.x <= bhi.x && hi.x >= blo.x && lo.y <= bhi.y && hi.y >= blo.y && lo.z <= bhi.z && hi.z >= blo.z;
        };
        uint32_t* stack = s_stack + slot;  // entry k at stack[64 k]
        int sp = 0;
        stack[0] = 0;
        sp = 1;
        while (sp > 0) {
            const uint32_t k = stack[64 * --sp];
            const float4 lo = s_nodes[2 * k], hi = s_nodes[2 * k + 1];
            if (!meets(lo, hi)) continue;
            const uint32_t c0 = __float_as_uint(lo.w), c1 = __float_as_uint(hi.w);
            if (c0 & PHR_LEAF) {
                for (uint32_t sl = c0 & ~PHR_LEAF; sl < c1; ++sl) {
                    const uint32_t j = s_order[sl];
                    if (j == i) continue;
                    const uint32_t fj = s_flags[j];
                    if (moves_f(fj) && j < i) continue;  // j owns it
                    if (!act_i && kind_of(fj) != K_ACTIVE) continue;
                    if ((gi & s_groups[j]) == 0u) continue;
                    if (!meets(s_boxes[2 * j], s_boxes[2 * j + 1])) continue;
                    if (found == PAIR_SLOTS) {
                        atomicOr(overflow, 1u);
                        continue;
                    }
                    uint32_t at = found++;
                    while (at > 0 && mine[64 * (at - 1)] > j) {
                        mine[64 * at] = mine[64 * (at - 1)];
                        --at;
                    }
                    mine[64 * at] = j;
                }
            } else if (sp + 2 <= (int)PHR_STACK) {  // an LBVH is at most 31 deep: never full
                stack[64 * sp++] = c0;
                stack[64 * sp++] = c1;
            }
        }
    }
    count[i] = found;
    for (uint32_t k = 0; k < found; ++k) other[(size_t)i * PAIR_SLOTS + k] = mine[64 * k];
}

// ---- 4. contacts ---------------------------------------------------------------------
// A manifold under construction: points on a and on b, normal from a to b.
struct Found {
    V n;
    V pa[4], pb[4];
    int m;
    __host__ __device__ void add(V a, V b) {
        if (m < 4) {
            pa[m] = a;
            pb[m] = b;
            ++m;
        }
    }
};

__host__ __device__ void spheres(V ca, float ra, V cb, float rb, float slop, Found& f) {
    const V d = cb - ca;
    const float l = length(d);
    if (l - ra - rb > slop) return;
    const V n = l > 1e-12f ? d * (1.0f / l) : mk(0.f, 0.f, 1.f);
    f.n = n;
    f.add(ca + n * ra, cb - n * rb);
}

// Closest points of segments [p0, p1] and [q0, q1] (parameters s, t in [0, 1]).
__host__ __device__ void segment_pair(V p0, V p1, V q0, V q1, float& s, float& t) {
    const V d1 = p1 - p0, d2 = q1 - q0, r = p0 - q0;
    const float a = norm2(d1), e = norm2(d2), f = dot(d2, r);
    if (a <= 1e-12f && e <= 1e-12f) { s = t = 0.f; return; }
    if (a <= 1e-12f) { s = 0.f; t = fminf(fmaxf(f / e, 0.f), 1.f); return; }
    const float c = dot(d1, r);
    if (e <= 1e-12f) { t = 0.f; s = fminf(fmaxf(-c / a, 0.f), 1.f); return; }
    const float b = dot(d1, d2), den = a * e - b * b;
    s = den > 1e-12f ? fminf(fmaxf((b * f - c * e) / den, 0.f), 1.f) : 0.f;
    t = (b * s + f) / e;
    if (t < 0.f) { t = 0.f; s = fminf(fmaxf(-c / a, 0.f), 1.f); }
    else if (t > 1.f) { t = 1.f; s = fminf(fmaxf((b - c) / a, 0.f), 1.f); }
}

__host__ __device__ void capsule_ends(const phr_shape& s, V x, Q q, V& p0, V& p1) {
    const V h = rot(q, mk(0.f, 0.f, s.half[2]));
    p0 = x - h;
    p1 = x + h;
}

// Sphere (a) against box (b): the nearest point of the box, or, from inside,
// the nearest face.
__host__ __device__ void sphere_box(V c, float r, V xb, Q qb, V h, float slop, Found& f) {
    const V l = unrot(qb, c - xb);
    const V k = mk(fminf(fmaxf(l.x, -h.x), h.x), fminf(fmaxf(l.y, -h.y), h.y), fminf(fmaxf(l.z, -h.z), h.z));
    const V d = l - k;
    const float dist = length(d);
    V nl, on;
    float gap;
    if (dist > 1e-7f) {
        nl = d * (-1.0f / dist);  // box surface -> centre is outward; the normal runs a -> b
        on = k;
        gap = dist - r;
    } else {
        // Centre inside: the face with the least way out.
        const float gx = h.x - fabsf(l.x), gy = h.y - fabsf(l.y), gz = h.z - fabsf(l.z);
        int ax = 0;
        float g = gx;
        if (gy < g) { g = gy; ax = 1; }
        if (gz < g) { g = gz; ax = 2; }
        const float sgn = comp(l, ax) >= 0.f ? 1.f : -1.f;
        nl = unit_axis(ax, -sgn);
        on = mk(ax == 0 ? sgn * h.x : l.x, ax == 1 ? sgn * h.y : l.y, ax == 2 ? sgn * h.z : l.z);
        gap = -g - r;
    }
    if (gap > slop) return;
    const V n = rot(qb, nl);
    f.n = n;
    f.add(c + n * r, xb + rot(qb, on));
}

struct Box {
    V c, ax[3];
    float h[3];
};
__host__ __device__ __forceinline__ float box_radius(const Box& b, V l) {
    return b.h[0] * fabsf(dot(b.ax[0], l)) + b.h[1] * fabsf(dot(b.ax[1], l)) + b.h[2] * fabsf(dot(b.ax[2], l));
}

constexpr float CLIP_TOL = 1e-5f;
// The incident face of `inc` against the reference face (axis k, outward
// normal n) of `ref`, clipped to the reference face's sides; points on the
// incident box (p) and their projections on the reference face (p - n s).
__host__ __device__ void face_contact(const Box& ref, int k, V n, const Box& inc, float slop, bool flip, Found& f) {
    int j = 0;
    float best = -1.f;
    for (int t = 0; t < 3; ++t) {
        const float a = fabsf(dot(inc.ax[t], n));
        if (a > best) { best = a; j = t; }
    }
    const float sj = dot(inc.ax[j], n) > 0.f ? -1.f : 1.f;
    const V fc = inc.c + inc.ax[j] * (sj * inc.h[j]);
    const int j1 = (j + 1) % 3, j2 = (j + 2) % 3;
    const V u = inc.ax[j1] * inc.h[j1], v = inc.ax[j2] * inc.h[j2];
    V poly[8], next[8];
    int np = 4;
    poly[0] = fc + u + v;
    poly[1] = fc - u + v;
    poly[2] = fc - u - v;
    poly[3] = fc + u - v;
    // Sutherland-Hodgman against the four side planes |(p - c) . ax_m| <= h_m.
    for (int side = 0; side < 4 && np > 0; ++side) {
        const int m = (k + 1 + (side >> 1)) % 3;
        const float sg = (side & 1) ? -1.f : 1.f;
        const V pn = ref.ax[m] * sg;
        // A corner on the side plane (boxes of one size stacked) stays, to
        // within rounding: the plane sits CLIP_TOL (relative) outside.
        const float off = ref.h[m] + CLIP_TOL * fmaxf(ref.h[m], 1.0f);
        int nn = 0;
        for (int e = 0; e < np; ++e) {
            const V p = poly[e], q = poly[(e + 1) % np];
            const float dp = dot(p - ref.c, pn) - off, dq = dot(q - ref.c, pn) - off;
            if (dp <= 0.f && nn < 8) next[nn++] = p;
            if ((dp <= 0.f) != (dq <= 0.f) && nn < 8) {
                const float t = dp / (dp - dq);
                next[nn++] = p + (q - p) * t;
            }
        }
        np = nn;
        for (int e = 0; e < np; ++e) poly[e] = next[e];
    }
    // Keep what is within reach of the reference face; at most four points.
    const V fr = ref.c + n * ref.h[k];
    V keep[8];
    float sep[8];
    int nk = 0;
    for (int e = 0; e < np; ++e) {
        const float s = dot(poly[e] - fr, n);
        if (s > slop) continue;
        bool twin = false;
        for (int t = 0; t < nk; ++t) twin = twin || norm2(keep[t] - poly[e]) < 1e-10f;
        if (twin) continue;
        keep[nk] = poly[e];
        sep[nk] = s;
        ++nk;
    }
    int pick[4], npick = 0;
    if (nk <= 4) {
        for (int e = 0; e < nk; ++e) pick[npick++] = e;
    } else {
        // The deepest point, the one farthest from it, the one spanning the
        // largest triangle with those, and the one farthest from all three.
        int p0 = 0;
        for (int e = 1; e < nk; ++e) if (sep[e] < sep[p0]) p0 = e;
        int p1 = p0 == 0 ? 1 : 0;
        for (int e = 0; e < nk; ++e) if (e != p0 && norm2(keep[e] - keep[p0]) > norm2(keep[p1] - keep[p0])) p1 = e;
        int p2 = -1;
        float a2 = -1.f;
        for (int e = 0; e < nk; ++e) {
            if (e == p0 || e == p1) continue;
            const float a = norm2(cross(keep[e] - keep[p0], keep[e] - keep[p1]));
            if (a > a2) { a2 = a; p2 = e; }
        }
        int p3 = -1;
        float d3 = -1.f;
        for (int e = 0; e < nk; ++e) {
            if (e == p0 || e == p1 || e == p2) continue;
            const float d = fminf(norm2(keep[e] - keep[p0]), fminf(norm2(keep[e] - keep[p1]), norm2(keep[e] - keep[p2])));
            if (d > d3) { d3 = d; p3 = e; }
        }
        pick[npick++] = p0;
        pick[npick++] = p1;
        if (p2 >= 0) pick[npick++] = p2;
        if (p3 >= 0) pick[npick++] = p3;
    }
    if (npick == 0) return;
    f.n = flip ? -n : n;
    for (int e = 0; e < npick; ++e) {
        const V p = keep[pick[e]], onref = p - n * sep[pick[e]];
        if (flip) f.add(p, onref);  // the reference box is b: p lies on a
        else f.add(onref, p);
    }
}

// Box a against box b by separating axes; faces are preferred to edge pairs
// unless an edge pair separates the boxes by EDGE_PREF more.
constexpr float EDGE_PREF = 5e-4f;
__host__ __device__ void boxes(const Box& A, const Box& B, float slop, Found& f) {
    const V d = B.c - A.c;
    float face_s = -3.0e38f;
    int face = -1;  // 0..2: a face of A; 3..5: a face of B
    for (int k = 0; k < 3; ++k) {
        const float s = fabsf(dot(d, A.ax[k])) - A.h[k] - box_radius(B, A.ax[k]);
        if (s > face_s) { face_s = s; face = k; }
    }
    for (int k = 0; k < 3; ++k) {
        const float s = fabsf(dot(d, B.ax[k])) - box_radius(A, B.ax[k]) - B.h[k];
        if (s > face_s) { face_s = s; face = 3 + k; }
    }
    if (face_s > slop) return;
    float edge_s = -3.0e38f;
    int ei = -1, ej = -1;
    V el = mk(0.f, 0.f, 0.f);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            V l = cross(A.ax[i], B.ax[j]);
            const float ll = length(l);
            if (ll < 1e-4f) continue;  // near-parallel edges: a face axis covers them
            l = l * (1.0f / ll);
            const float s = fabsf(dot(d, l)) - box_radius(A, l) - box_radius(B, l);
            if (s > slop) return;  // a separating edge axis
            if (s > edge_s) { edge_s = s; ei = i; ej = j; el = l; }
        }
    if (ei >= 0 && edge_s > face_s + EDGE_PREF) {
        const V n = dot(d, el) >= 0.f ? el : -el;
        V pa = A.c, pb = B.c;
        for (int m = 0; m < 3; ++m) {
            if (m != ei) pa = pa + A.ax[m] * (dot(A.ax[m], n) >= 0.f ? A.h[m] : -A.h[m]);
            if (m != ej) pb = pb + B.ax[m] * (dot(B.ax[m], n) >= 0.f ? -B.h[m] : B.h[m]);
        }
        const V ea = A.ax[ei] * A.h[ei], eb = B.ax[ej] * B.h[ej];
        float s, t;
        segment_pair(pa - ea, pa + ea, pb - eb, pb + eb, s, t);
        f.n = n;
        f.add(pa - ea + ea * (2.0f * s), pb - eb + eb * (2.0f * t));
        return;
    }
    if (face < 3) {
        const V n = dot(d, A.ax[face]) >= 0.f ? A.ax[face] : -A.ax[face];
        face_contact(A, face, n, B, slop, false, f);
    } else {
        const int k = face - 3;
        const V n = dot(d, B.ax[k]) >= 0.f ? -B.ax[k] : B.ax[k];  // B's face towards A
        face_contact(B, k, n, A, slop, true, f);
    }
}

__host__ __device__ Box box_of(const phr_shape& s, V x, Q q) {
    Box b;
    b.c = x;
    b.ax[0] = rot(q, mk(1.f, 0.f, 0.f));
    b.ax[1] = rot(q, mk(0.f, 1.f, 0.f));
    b.ax[2] = rot(q, mk(0.f, 0.f, 1.f));
    b.h[0] = s.half[0];
    b.h[1] = s.half[1];
    b.h[2] = s.half[2];
    return b;
}

// The points of one pair with a's shape kind <= b's where that matters.
__host__ __device__ void manifold(const phr_shape& sa, V xa, Q qa, const phr_shape& sb, V xb, Q qb, float slop, Found& f) {
    const uint32_t ka = sa.kind, kb = sb.kind;
    if (ka == S_SPHERE && kb == S_SPHERE) return spheres(xa, sa.radius, xb, sb.radius, slop, f);
    if (ka == S_BOX && kb == S_BOX) return boxes(box_of(sa, xa, qa), box_of(sb, xb, qb), slop, f);
    if (ka == S_SPHERE && kb == S_BOX) return sphere_box(xa, sa.radius, xb, qb, ld3(sb.half), slop, f);
    if (ka == S_BOX && kb == S_SPHERE) {
        sphere_box(xb, sb.radius, xa, qa, ld3(sa.half), slop, f);
        f.n = -f.n;
        for (int e = 0; e < f.m; ++e) { const V t = f.pa[e]; f.pa[e] = f.pb[e]; f.pb[e] = t; }
        return;
    }
    if (ka == S_CAPSULE && kb == S_CAPSULE) {
        V p0, p1, q0, q1;
        capsule_ends(sa, xa, qa, p0, p1);
        capsule_ends(sb, xb, qb, q0, q1);
        const V da = p1 - p0, db = q1 - q0;
        const float par = norm2(cross(da, db));
        if (par < 1e-6f * norm2(da) * norm2(db) && norm2(da) > 1e-12f) {
            // Side by side: the ends of their overlap along a's axis.
            const float la = norm2(da);
            const float t0 = fminf(fmaxf(dot(q0 - p0, da) / la, 0.f), 1.f), t1 = fminf(fmaxf(dot(q1 - p0, da) / la, 0.f), 1.f);
            for (float t : {t0, t1}) {
                const V pa = p0 + da * t;
                float s, u;
                segment_pair(pa, pa, q0, q1, s, u);
                Found g{};
                spheres(pa, sa.radius, q0 + db * u, sb.radius, slop, g);
                if (g.m) { f.n = g.n; f.add(g.pa[0], g.pb[0]); }
                if (t0 == t1) break;
            }
            return;
        }
        float s, t;
        segment_pair(p0, p1, q0, q1, s, t);
        return spheres(p0 + da * s, sa.radius, q0 + db * t, sb.radius, slop, f);
    }
    if (ka == S_SPHERE && kb == S_CAPSULE) {
        V q0, q1;
        capsule_ends(sb, xb, qb, q0, q1);
        float s, t;
        segment_pair(xa, xa, q0, q1, s, t);
        return spheres(xa, sa.radius, q0 + (q1 - q0) * t, sb.radius, slop, f);
    }
    if (ka == S_CAPSULE && kb == S_SPHERE) {
        V p0, p1;
        capsule_ends(sa, xa, qa, p0, p1);
        float s, t;
        segment_pair(xb, xb, p0, p1, s, t);
        return spheres(p0 + (p1 - p0) * t, sa.radius, xb, sb.radius, slop, f);
    }
    // box-capsule pairs are refused when the bodies are added
}

__global__ __launch_bounds__(128) void k_contacts(const Body* __restrict__ bodies, const Konst* __restrict__ konst,
                                                  const phr_shape* __restrict__ shapes, const uint32_t* __restrict__ count,
                                                  const uint32_t* __restrict__ other, uint32_t n, float dt,
                                                  PairRec* __restrict__ pairs, PointRec* __restrict__ points) {
    const uint32_t slot = blockIdx.x * 128 + threadIdx.x;
    const uint32_t i = slot / PAIR_SLOTS, k = slot % PAIR_SLOTS;
    if (i >= n || k >= count[i]) return;
    const uint32_t j = other[slot];
    const uint32_t a = min(i, j), b = max(i, j);
    const Body A = bodies[a], B = bodies[b];
    const V xa = ld3(A.x), xb = ld3(B.x);
    const Q qa{A.q[0], A.q[1], A.q[2], A.q[3]}, qb{B.q[0], B.q[1], B.q[2], B.q[3]};
    // How far the pair can close before the next frame finds it again.
    const float reach = fmaxf(length(ld3(A.v) - ld3(B.v)) * dt, 0.f);
    const float slop = konst[a].margin + konst[b].margin + SLOP + reach;
    Found f{};
    manifold(shapes[A.shape], xa, qa, shapes[B.shape], xb, qb, slop, f);
    pairs[slot] = PairRec{a, b, (uint32_t)f.m, 0u, make_float4(f.n.x, f.n.y, f.n.z, 0.f)};
    for (int e = 0; e < f.m; ++e) {
        const V la = unrot(qa, f.pa[e] - xa), lb = unrot(qb, f.pb[e] - xb);
        points[4 * (size_t)slot + e] =
            PointRec{make_float4(la.x, la.y, la.z, 0.f), make_float4(lb.x, lb.y, lb.z, 0.f),
                     make_float4(0.f, dot(f.pa[e] - f.pb[e], f.n), 0.f, 0.f)};
    }
}

// ---- 5. the solve: one workgroup -------------------------------------------------------
struct Params {
    uint32_t n, substeps, iterations, pad;
    float dt, h, cap, g;
    float gx, gy, gz, sleep_after;
};

// The generalized inverse mass of body i along direction l at arm r (Muller
// 2020 eq. 2-3), zero for a body the solve does not move.
__host__ __device__ __forceinline__ float w_along(const Lds& L, const Konst* kc, const uint32_t* fl, uint32_t i, V r, V l) {
    if (!dynamic_f(fl[i])) return 0.f;
    const V rl = unrot(L.q(i), cross(r, l));
    return L.im(i) + dot(hadamard(rl, ld3(kc[i].inv_inertia)), rl);
}
// The turn an impulse p at arm r gives body i (world frame).
__host__ __device__ __forceinline__ V turn_of(const Lds& L, const Konst* kc, uint32_t i, V r, V p) {
    const Q q = L.q(i);
    return rot(q, hadamard(unrot(q, cross(r, p)), ld3(kc[i].inv_inertia)));
}

// Impulses on the points of one pair, applied together.
struct Kicks {
    V p[4];       // impulse on a at point e (-p on b)
    V dir[4];     // the direction the point asked to move in
    float need[4];
    bool on[4];
};
struct Move {
    V da, ta, db, tb;  // translation and turn of a and b
};
__host__ __device__ Move summed(const Lds& L, const Konst* kc, const uint32_t* fl, uint32_t a, uint32_t b, const V* ra, const V* rb,
                       const Kicks& k, int m, const float* weight) {
    Move mv{mk(0.f, 0.f, 0.f), mk(0.f, 0.f, 0.f), mk(0.f, 0.f, 0.f), mk(0.f, 0.f, 0.f)};
    const bool da = dynamic_f(fl[a]), db = dynamic_f(fl[b]);
    for (int e = 0; e < m; ++e) {
        if (!k.on[e]) continue;
        const V p = weight ? k.p[e] * weight[e] : k.p[e];
        if (da) {
            mv.da = mv.da + p * L.im(a);
            mv.ta = mv.ta + turn_of(L, kc, a, ra[e], p);
        }
        if (db) {
            mv.db = mv.db - p * L.im(b);
            mv.tb = mv.tb - turn_of(L, kc, b, rb[e], p);
        }
    }
    return mv;
}
// The largest scale <= 1 under which no point moves by more than it asked.
__host__ __device__ float safe_scale(const Move& mv, const V* ra, const V* rb, const Kicks& k, int m) {
    float s = 1.0f;
    for (int e = 0; e < m; ++e) {
        if (!k.on[e]) continue;
        const float got = dot((mv.da + cross(mv.ta, ra[e])) - (mv.db + cross(mv.tb, rb[e])), k.dir[e]);
        if (got * k.need[e] > 0.f) s = fminf(s, k.need[e] / got);
    }
    return s;
}

__host__ __device__ void project_pair(Lds& L, const Konst* kc, const uint32_t* fl, const float4* prev, float4* push, const PairRec& pr,
                             PointRec* pts, float cap) {
    const uint32_t a = pr.a, b = pr.b;
    const int m = (int)pr.m;
    const V n = mk(pr.n.x, pr.n.y, pr.n.z);
    const Q qa = L.q(a), qb = L.q(b);
    const V xa = L.x(a), xb = L.x(b);
    const float4 pa0 = prev[2 * a], pa1 = prev[2 * a + 1], pb0 = prev[2 * b], pb1 = prev[2 * b + 1];
    const V xa_prev = mk(pa0.x, pa0.y, pa0.z), xb_prev = mk(pb0.x, pb0.y, pb0.z);
    const Q qa_prev{pa1.x, pa1.y, pa1.z, pa1.w}, qb_prev{pb1.x, pb1.y, pb1.z, pb1.w};
    V ra[4], rb[4], la[4], lb[4];
    float lam_n[4], lam_t[4], rec[4];
    Kicks k;
    for (int e = 0; e < m; ++e) {
        const float4 aa = pts[e].aa, ab = pts[e].ab;
        la[e] = mk(aa.x, aa.y, aa.z);
        lb[e] = mk(ab.x, ab.y, ab.z);
        lam_n[e] = aa.w;
        lam_t[e] = ab.w;
        ra[e] = rot(qa, la[e]);
        rb[e] = rot(qb, lb[e]);
        k.on[e] = false;
        rec[e] = 0.f;
    }
    // Normal: push each penetrating point out, the part of its depth the
    // substep's own drift closed in full, the rest up to `cap`.
    float dl[4] = {0.f, 0.f, 0.f, 0.f};
    bool any = false;
    for (int e = 0; e < m; ++e) {
        const float depth = dot((xa + ra[e]) - (xb + rb[e]), n);
        const float wsum = w_along(L, kc, fl, a, ra[e], n) + w_along(L, kc, fl, b, rb[e], n);
        if (!(depth > 0.f) || wsum == 0.f) continue;
        const V was_a = xa_prev + rot(qa_prev, la[e]), was_b = xb_prev + rot(qb_prev, lb[e]);
        const float closed = fmaxf(dot(((xa + ra[e]) - was_a) - ((xb + rb[e]) - was_b), n), 0.f);
        const float hit = fminf(depth, closed);
        const float fix = hit + fminf(depth - hit, cap);
        dl[e] = -fix / wsum;
        rec[e] = fix > 0.f ? (fix - hit) / fix : 0.f;
        k.p[e] = n * dl[e];
        k.dir[e] = -n;
        k.need[e] = fix;
        k.on[e] = true;
        any = true;
    }
    if (any) {
        const Move mv = summed(L, kc, fl, a, b, ra, rb, k, m, nullptr);
        const float s = safe_scale(mv, ra, rb, k, m);
        const Move rv = summed(L, kc, fl, a, b, ra, rb, k, m, rec);
        if (dynamic_f(fl[a])) {
            L.set_x(a, xa + mv.da * s);
            L.set_q(a, qstep(qa, mv.ta * s, 1.0f));  // a turn vector applied whole (Muller 2020 eq. 8 at h = 1)
            float4 p0 = push[2 * a], p1 = push[2 * a + 1];
            push[2 * a] = make_float4(p0.x + rv.da.x * s, p0.y + rv.da.y * s, p0.z + rv.da.z * s, 0.f);
            push[2 * a + 1] = make_float4(p1.x + rv.ta.x * s, p1.y + rv.ta.y * s, p1.z + rv.ta.z * s, 0.f);
        }
        if (dynamic_f(fl[b])) {
            L.set_x(b, xb + mv.db * s);
            L.set_q(b, qstep(qb, mv.tb * s, 1.0f));
            float4 p0 = push[2 * b], p1 = push[2 * b + 1];
            push[2 * b] = make_float4(p0.x + rv.db.x * s, p0.y + rv.db.y * s, p0.z + rv.db.z * s, 0.f);
            push[2 * b + 1] = make_float4(p1.x + rv.tb.x * s, p1.y + rv.tb.y * s, p1.z + rv.tb.z * s, 0.f);
        }
        for (int e = 0; e < m; ++e) lam_n[e] += dl[e] * s;
    }
    // Static friction: hold each point's tangential drift since the
    // substep began, inside the cone mu |lambda_n|. The ask is first scaled
    // as the pair's points share it, then clamped to what each cone allows.
    const float mu = kc[a].friction * kc[b].friction;
    const Q qa2 = L.q(a), qb2 = L.q(b);
    const V xa2 = L.x(a), xb2 = L.x(b);
    bool fr = false;
    V tdir[4];
    for (int e = 0; e < m; ++e) {
        k.on[e] = false;
        dl[e] = 0.f;
        if (lam_n[e] == 0.f) continue;
        ra[e] = rot(qa2, la[e]);
        rb[e] = rot(qb2, lb[e]);
        const V drift = ((xa2 + ra[e]) - (xa_prev + rot(qa_prev, la[e]))) - ((xb2 + rb[e]) - (xb_prev + rot(qb_prev, lb[e])));
        const V tang = drift - n * dot(drift, n);
        const float lt = length(tang);
        if (lt < 1e-9f) continue;
        const V t = tang * (1.0f / lt);
        const float wsum = w_along(L, kc, fl, a, ra[e], t) + w_along(L, kc, fl, b, rb[e], t);
        if (wsum == 0.f) continue;
        tdir[e] = t;
        dl[e] = -lt / wsum;
        k.p[e] = t * dl[e];
        k.dir[e] = -t;
        k.need[e] = lt;
        k.on[e] = true;
        fr = true;
    }
    if (!fr) {
        for (int e = 0; e < m; ++e) pts[e].aa.w = lam_n[e];
        return;
    }
    for (int e = 0; e < m; ++e) if (!k.on[e]) ra[e] = rot(qa2, la[e]), rb[e] = rot(qb2, lb[e]);
    const float share = safe_scale(summed(L, kc, fl, a, b, ra, rb, k, m, nullptr), ra, rb, k, m);
    for (int e = 0; e < m; ++e) {
        if (!k.on[e]) continue;
        const float want = dl[e] * share;
        const float room = fmaxf(mu * fabsf(lam_n[e]) - fabsf(lam_t[e]), 0.f);
        dl[e] = copysignf(fminf(fabsf(want), room), want);
        k.p[e] = tdir[e] * dl[e];
    }
    const Move mv = summed(L, kc, fl, a, b, ra, rb, k, m, nullptr);
    const float s = safe_scale(mv, ra, rb, k, m);
    if (dynamic_f(fl[a])) {
        L.set_x(a, xa2 + mv.da * s);
        L.set_q(a, qstep(qa2, mv.ta * s, 1.0f));
    }
    if (dynamic_f(fl[b])) {
        L.set_x(b, xb2 + mv.db * s);
        L.set_q(b, qstep(qb2, mv.tb * s, 1.0f));
    }
    for (int e = 0; e < m; ++e) {
        lam_t[e] += dl[e] * s;
        pts[e].aa.w = lam_n[e];
        pts[e].ab.w = lam_t[e];
    }
}

// Dynamic friction, then restitution, on the velocities (Muller 2020 §3.6).
__host__ __device__ void velocity_pair(Lds& L, const Konst* kc, const uint32_t* fl, const PairRec& pr, const PointRec* pts, float h,
                              float g) {
    const uint32_t a = pr.a, b = pr.b;
    const int m = (int)pr.m;
    const V n = mk(pr.n.x, pr.n.y, pr.n.z);
    const Q qa = L.q(a), qb = L.q(b);
    V ra[4], rb[4];
    for (int e = 0; e < m; ++e) {
        ra[e] = rot(qa, mk(pts[e].aa.x, pts[e].aa.y, pts[e].aa.z));
        rb[e] = rot(qb, mk(pts[e].ab.x, pts[e].ab.y, pts[e].ab.z));
    }
    const float mu = kc[a].friction * kc[b].friction;
    const bool da = dynamic_f(fl[a]), db = dynamic_f(fl[b]);
    auto rel = [&](int e) { return (L.v(a) + cross(L.w(a), ra[e])) - (L.v(b) + cross(L.w(b), rb[e])); };
    auto apply = [&](const Kicks& k) {
        const Move mv = summed(L, kc, fl, a, b, ra, rb, k, m, nullptr);
        const float s = safe_scale(mv, ra, rb, k, m);
        if (da) { L.set_v(a, L.v(a) + mv.da * s); L.set_w(a, L.w(a) + mv.ta * s); }
        if (db) { L.set_v(b, L.v(b) + mv.db * s); L.set_w(b, L.w(b) + mv.tb * s); }
    };
    Kicks k;
    bool any = false;
    for (int e = 0; e < m; ++e) {
        k.on[e] = false;
        const float lam = pts[e].aa.w;
        if (lam == 0.f) continue;
        const V vr = rel(e);
        const V vt = vr - n * dot(vr, n);
        const float lt = length(vt);
        if (lt < 1e-9f) continue;
        const V t = vt * (1.0f / lt);
        const float wsum = w_along(L, kc, fl, a, ra[e], t) + w_along(L, kc, fl, b, rb[e], t);
        const float j = fminf(lt / wsum, mu * fabsf(lam) / h);
        k.p[e] = t * (-j);
        k.dir[e] = -t;
        k.need[e] = j * wsum;
        k.on[e] = true;
        any = true;
    }
    if (any) apply(k);
    any = false;
    const float e_pair = kc[a].restitution * kc[b].restitution;
    for (int e = 0; e < m; ++e) {
        k.on[e] = false;
        if (pts[e].aa.w == 0.f) continue;
        const float vn0 = pts[e].e.x;
        const float vn = dot(rel(e), n);
        const float rest = fabsf(vn0) <= 2.0f * g * h ? 0.f : e_pair;
        const float wsum = w_along(L, kc, fl, a, ra[e], n) + w_along(L, kc, fl, b, rb[e], n);
        const float need = -rest * fmaxf(vn0, 0.f) - vn;
        // Only ever an impulse that separates the pair.
        if (wsum > 0.f && need < -1e-12f) {
            k.p[e] = n * (need / wsum);
            k.dir[e] = n;
            k.need[e] = need;
            k.on[e] = true;
            any = true;
        }
    }
    if (any) apply(k);
}

// ---- the same two projections on four lanes a pair -------------------------------------
// Lane e of an aligned group of four holds the pair's point e (lanes past the
// pair's points hold none and add zeros). A sum over the points is taken on
// every lane from the four lanes' terms in point order, through DPP quad
// broadcasts, starting from zero as `summed` does; a point without a term
// adds +0, which leaves a sum that never is -0 unchanged. So each lane holds
// the bits the one-lane functions compute, and lane 0 alone writes the
// bodies. A colour's projection then waits on about a quarter of the
// one-lane version's dependent arithmetic.
template <int K>
__device__ __forceinline__ float qlane(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), K * 0x55, 0xF, 0xF, false));
}
template <int K>
__device__ __forceinline__ int qlane_i(int v) {
    return __builtin_amdgcn_update_dpp(0, v, K * 0x55, 0xF, 0xF, false);
}
__device__ __forceinline__ float qsum(float v) {
    float r = 0.f;
    r = r + qlane<0>(v);
    r = r + qlane<1>(v);
    r = r + qlane<2>(v);
    r = r + qlane<3>(v);
    return r;
}
__device__ __forceinline__ float qneg_sum(float v) {  // 0 - v0 - v1 - v2 - v3, in that order
    float r = 0.f;
    r = r - qlane<0>(v);
    r = r - qlane<1>(v);
    r = r - qlane<2>(v);
    r = r - qlane<3>(v);
    return r;
}
__device__ __forceinline__ V qsum(V v) { return mk(qsum(v.x), qsum(v.y), qsum(v.z)); }
__device__ __forceinline__ V qneg_sum(V v) { return mk(qneg_sum(v.x), qneg_sum(v.y), qneg_sum(v.z)); }
__device__ __forceinline__ bool qany(bool on) {
    const int x = on ? 1 : 0;
    return (qlane_i<0>(x) | qlane_i<1>(x) | qlane_i<2>(x) | qlane_i<3>(x)) != 0;
}
// What a projection reads of its pair besides the bodies' state in LDS,
// loaded for the next colour while the current one is solved: the pair
// record, this lane's point, both bodies' poses at the substep's start and
// their inverse inertia, friction and restitution. None of it changes
// within a substep's colours (a point's multipliers change only under its
// own pair, on these same lanes).
struct PairPre {
    PairRec pr;
    PointRec pt;
    float4 pa0, pa1, pb0, pb1;
    V ia, ib;
    float fa, fb, ea, eb;
};
__device__ __forceinline__ void pair_pre(PairPre& P, const PairRec* __restrict__ spairs, const PointRec* __restrict__ spoints,
                                         const float4* __restrict__ prev, const Konst* __restrict__ kc, uint32_t e, uint32_t lane) {
    P.pr = spairs[e];
    const uint32_t a = P.pr.a, b = P.pr.b;
    P.pt = lane < P.pr.m ? spoints[4 * (size_t)e + lane] : PointRec{};
    P.pa0 = prev[2 * a];
    P.pa1 = prev[2 * a + 1];
    P.pb0 = prev[2 * b];
    P.pb1 = prev[2 * b + 1];
    P.ia = ld3(kc[a].inv_inertia);
    P.ib = ld3(kc[b].inv_inertia);
    P.fa = kc[a].friction;
    P.fb = kc[b].friction;
    P.ea = kc[a].restitution;
    P.eb = kc[b].restitution;
}
// w_along and turn_of with the body's inverse inertia in hand.
__device__ __forceinline__ float w_along_k(const Lds& L, const uint32_t* fl, uint32_t i, V inv_i, V r, V l) {
    if (!dynamic_f(fl[i])) return 0.f;
    const V rl = unrot(L.q(i), cross(r, l));
    return L.im(i) + dot(hadamard(rl, inv_i), rl);
}
__device__ __forceinline__ V turn_of_k(const Lds& L, uint32_t i, V inv_i, V r, V p) {
    const Q q = L.q(i);
    return rot(q, hadamard(unrot(q, cross(r, p)), inv_i));
}
// `summed` with this lane's point: its impulse p (weighted already), on or off.
__device__ __forceinline__ Move quad_summed(const Lds& L, const PairPre& K, bool da, bool db, uint32_t a, uint32_t b, V ra, V rb,
                                            bool on, V p) {
    const V z = mk(0.f, 0.f, 0.f);
    const V ca = (on && da) ? p * L.im(a) : z, ta = (on && da) ? turn_of_k(L, a, K.ia, ra, p) : z;
    const V cb = (on && db) ? p * L.im(b) : z, tb = (on && db) ? turn_of_k(L, b, K.ib, rb, p) : z;
    return Move{qsum(ca), qsum(ta), qneg_sum(cb), qneg_sum(tb)};
}
// `safe_scale`: the smallest of 1 and each asking point's need / got.
__device__ __forceinline__ float quad_scale(const Move& mv, V ra, V rb, bool on, V dir, float need) {
    float s = 1.0f;
    if (on) {
        const float got = dot((mv.da + cross(mv.ta, ra)) - (mv.db + cross(mv.tb, rb)), dir);
        if (got * need > 0.f) s = need / got;
    }
    return fminf(fminf(fminf(fminf(1.0f, qlane<0>(s)), qlane<1>(s)), qlane<2>(s)), qlane<3>(s));
}

__device__ __forceinline__ void project_quad(Lds& L, const PairPre& K, const uint32_t* fl, float4* push, PointRec* pts, float cap,
                                             uint32_t e) {
    const PairRec& pr = K.pr;
    const uint32_t a = pr.a, b = pr.b;
    const bool mine = e < pr.m;
    const bool da = dynamic_f(fl[a]), db = dynamic_f(fl[b]);
    const V n = mk(pr.n.x, pr.n.y, pr.n.z);
    const Q qa = L.q(a), qb = L.q(b);
    const V xa = L.x(a), xb = L.x(b);
    const float4 pa0 = K.pa0, pa1 = K.pa1, pb0 = K.pb0, pb1 = K.pb1;
    const V xa_prev = mk(pa0.x, pa0.y, pa0.z), xb_prev = mk(pb0.x, pb0.y, pb0.z);
    const Q qa_prev{pa1.x, pa1.y, pa1.z, pa1.w}, qb_prev{pb1.x, pb1.y, pb1.z, pb1.w};
    V la = mk(0.f, 0.f, 0.f), lb = la;
    float lam_n = 0.f, lam_t = 0.f;
    if (mine) {
        const float4 aa = K.pt.aa, ab = K.pt.ab;
        la = mk(aa.x, aa.y, aa.z);
        lb = mk(ab.x, ab.y, ab.z);
        lam_n = aa.w;
        lam_t = ab.w;
    }
    V ra = rot(qa, la), rb = rot(qb, lb);
    // Normal.
    bool on = false;
    float dl = 0.f, rec = 0.f, need = 0.f;
    V p = mk(0.f, 0.f, 0.f), dir = p;
    if (mine) {
        const float depth = dot((xa + ra) - (xb + rb), n);
        const float wsum = w_along_k(L, fl, a, K.ia, ra, n) + w_along_k(L, fl, b, K.ib, rb, n);
        if (depth > 0.f && wsum != 0.f) {
            const V was_a = xa_prev + rot(qa_prev, la), was_b = xb_prev + rot(qb_prev, lb);
            const float closed = fmaxf(dot(((xa + ra) - was_a) - ((xb + rb) - was_b), n), 0.f);
            const float hit = fminf(depth, closed);
            const float fix = hit + fminf(depth - hit, cap);
            dl = -fix / wsum;
            rec = fix > 0.f ? (fix - hit) / fix : 0.f;
            p = n * dl;
            dir = -n;
            need = fix;
            on = true;
        }
    }
    V xa2 = xa, xb2 = xb;
    Q qa2 = qa, qb2 = qb;
    if (qany(on)) {
        const Move mv = quad_summed(L, K, da, db, a, b, ra, rb, on, p);
        const float s = quad_scale(mv, ra, rb, on, dir, need);
        const Move rv = quad_summed(L, K, da, db, a, b, ra, rb, on, p * rec);
        if (da) {
            xa2 = xa + mv.da * s;
            qa2 = qstep(qa, mv.ta * s, 1.0f);
        }
        if (db) {
            xb2 = xb + mv.db * s;
            qb2 = qstep(qb, mv.tb * s, 1.0f);
        }
        if (e == 0) {
            if (da) {
                L.set_x(a, xa2);
                L.set_q(a, qa2);
                const float4 p0 = push[2 * a], p1 = push[2 * a + 1];
                push[2 * a] = make_float4(p0.x + rv.da.x * s, p0.y + rv.da.y * s, p0.z + rv.da.z * s, 0.f);
                push[2 * a + 1] = make_float4(p1.x + rv.ta.x * s, p1.y + rv.ta.y * s, p1.z + rv.ta.z * s, 0.f);
            }
            if (db) {
                L.set_x(b, xb2);
                L.set_q(b, qb2);
                const float4 p0 = push[2 * b], p1 = push[2 * b + 1];
                push[2 * b] = make_float4(p0.x + rv.db.x * s, p0.y + rv.db.y * s, p0.z + rv.db.z * s, 0.f);
                push[2 * b + 1] = make_float4(p1.x + rv.tb.x * s, p1.y + rv.tb.y * s, p1.z + rv.tb.z * s, 0.f);
            }
        }
        lam_n += dl * s;
    }
    // Static friction.
    const float mu = K.fa * K.fb;
    on = false;
    dl = 0.f;
    V tdir = mk(0.f, 0.f, 0.f);
    ra = rot(qa2, la);
    rb = rot(qb2, lb);
    if (mine && lam_n != 0.f) {
        const V drift = ((xa2 + ra) - (xa_prev + rot(qa_prev, la))) - ((xb2 + rb) - (xb_prev + rot(qb_prev, lb)));
        const V tang = drift - n * dot(drift, n);
        const float lt = length(tang);
        if (lt >= 1e-9f) {
            const V t = tang * (1.0f / lt);
            const float wsum = w_along_k(L, fl, a, K.ia, ra, t) + w_along_k(L, fl, b, K.ib, rb, t);
            if (wsum != 0.f) {
                tdir = t;
                dl = -lt / wsum;
                p = t * dl;
                dir = -t;
                need = lt;
                on = true;
            }
        }
    }
    if (!qany(on)) {
        if (mine) pts[e].aa.w = lam_n;
        return;
    }
    const float share = quad_scale(quad_summed(L, K, da, db, a, b, ra, rb, on, p), ra, rb, on, dir, need);
    if (on) {
        const float want = dl * share;
        const float room = fmaxf(mu * fabsf(lam_n) - fabsf(lam_t), 0.f);
        dl = copysignf(fminf(fabsf(want), room), want);
        p = tdir * dl;
    }
    const Move mv = quad_summed(L, K, da, db, a, b, ra, rb, on, p);
    const float s = quad_scale(mv, ra, rb, on, dir, need);
    if (e == 0) {
        if (da) {
            L.set_x(a, xa2 + mv.da * s);
            L.set_q(a, qstep(qa2, mv.ta * s, 1.0f));
        }
        if (db) {
            L.set_x(b, xb2 + mv.db * s);
            L.set_q(b, qstep(qb2, mv.tb * s, 1.0f));
        }
    }
    lam_t += dl * s;
    if (mine) {
        pts[e].aa.w = lam_n;
        pts[e].ab.w = lam_t;
    }
}

__device__ __forceinline__ void velocity_quad(Lds& L, const PairPre& K, const uint32_t* fl, float h, float g, uint32_t e) {
    const PairRec& pr = K.pr;
    const uint32_t a = pr.a, b = pr.b;
    const bool mine = e < pr.m;
    const V n = mk(pr.n.x, pr.n.y, pr.n.z);
    const Q qa = L.q(a), qb = L.q(b);
    const PointRec pt = K.pt;
    const V ra = rot(qa, mk(pt.aa.x, pt.aa.y, pt.aa.z)), rb = rot(qb, mk(pt.ab.x, pt.ab.y, pt.ab.z));
    const float mu = K.fa * K.fb;
    const bool da = dynamic_f(fl[a]), db = dynamic_f(fl[b]);
    V va = L.v(a), wa = L.w(a), vb = L.v(b), wb = L.w(b);
    auto rel = [&]() { return (va + cross(wa, ra)) - (vb + cross(wb, rb)); };
    auto apply = [&](bool on, V p, V dir, float need) {
        const Move mv = quad_summed(L, K, da, db, a, b, ra, rb, on, p);
        const float s = quad_scale(mv, ra, rb, on, dir, need);
        if (da) { va = va + mv.da * s; wa = wa + mv.ta * s; }
        if (db) { vb = vb + mv.db * s; wb = wb + mv.tb * s; }
        if (e == 0) {
            if (da) { L.set_v(a, va); L.set_w(a, wa); }
            if (db) { L.set_v(b, vb); L.set_w(b, wb); }
        }
    };
    bool on = false;
    V p = mk(0.f, 0.f, 0.f), dir = p;
    float need = 0.f;
    const float lam = pt.aa.w;
    if (mine && lam != 0.f) {
        const V vr = rel();
        const V vt = vr - n * dot(vr, n);
        const float lt = length(vt);
        if (lt >= 1e-9f) {
            const V t = vt * (1.0f / lt);
            const float wsum = w_along_k(L, fl, a, K.ia, ra, t) + w_along_k(L, fl, b, K.ib, rb, t);
            const float j = fminf(lt / wsum, mu * fabsf(lam) / h);
            p = t * (-j);
            dir = -t;
            need = j * wsum;
            on = true;
        }
    }
    if (qany(on)) apply(on, p, dir, need);
    on = false;
    const float e_pair = K.ea * K.eb;
    if (mine && lam != 0.f) {
        const float vn0 = pt.e.x;
        const float vn = dot(rel(), n);
        const float rest = fabsf(vn0) <= 2.0f * g * h ? 0.f : e_pair;
        const float wsum = w_along_k(L, fl, a, K.ia, ra, n) + w_along_k(L, fl, b, K.ib, rb, n);
        const float nd = -rest * fmaxf(vn0, 0.f) - vn;
        if (wsum > 0.f && nd < -1e-12f) {
            p = n * (nd / wsum);
            dir = n;
            need = nd;
            on = true;
        }
    }
    if (qany(on)) apply(on, p, dir, need);
}

// A pair's place in phiray's pair order (broad.rs sorts the pairs (a, b),
// a < b; contact.rs walks them in that order): unique, since bodies number
// fewer than 2^11. Colouring by it, a pair joins a round only once every
// lower pair sharing one of its moving bodies has its colour, so the colours
// in sequence do exactly what the pairs one after another in that order do.
__host__ __device__ __forceinline__ uint32_t priority(const PairRec& pr) { return (pr.a << 11) | pr.b; }

// QUAD: four lanes a pair in the projections. NB: bodies the LDS arrays hold.
// LPUSH: the pushed-out offsets (read and written at every projection) in
// LDS too, for worlds of at most LPUSH_BODIES.
template <bool QUAD, uint32_t NB, bool LPUSH>
__global__ __launch_bounds__(SOLVE_THREADS) void k_solve(Body* __restrict__ bodies, const Konst* __restrict__ kc,
                                                         Aux* __restrict__ aux, const uint32_t* __restrict__ count,
                                                         const PairRec* __restrict__ pairs, PointRec* __restrict__ points,
                                                         uint32_t* __restrict__ live, uint32_t* __restrict__ order,
                                                         PairRec* __restrict__ spairs, PointRec* __restrict__ spoints,
                                                         uint32_t* __restrict__ rank, float4* __restrict__ prev, float4* __restrict__ gpush,
                                                         float* __restrict__ poses, uint32_t* __restrict__ stats, Params P) {
    __shared__ float4 s_body[4 * NB];
    __shared__ uint32_t s_flags[NB];
    __shared__ uint32_t s_claim[NB];
    __shared__ float4 s_push[LPUSH ? 2 * NB : 1];
    __shared__ uint32_t s_cstart[MAX_COLOURS + 1];
    constexpr uint32_t LIVE = LPUSH ? 5120u : 1792u;  // live pairs whose bodies the LDS left over holds
    __shared__ uint32_t s_ab[LIVE];
    __shared__ uint32_t s_deg[LPUSH ? NB : 1];  // live pairs at each moving body (the dataflow order)
    __shared__ uint32_t s_err;
    float4* push = LPUSH ? s_push : gpush;
    __shared__ uint32_t s_nlive, s_cursor, s_colours, s_contacts, s_pairs;
    const uint32_t tid = threadIdx.x, T = SOLVE_THREADS, n = P.n;
    if (stats[4] != 0u) {  // a body listed more pairs than it has slots: report, change nothing
        if (tid == 0) stats[3] = 1u;
        return;
    }
    Lds L{s_body};
    for (uint32_t i = tid; i < n; i += T) {
        const float4* src = (const float4*)&bodies[i];
        s_body[4 * i] = src[0];
        s_body[4 * i + 1] = src[1];
        s_body[4 * i + 2] = src[2];
        s_body[4 * i + 3] = src[3];
        s_flags[i] = bodies[i].flags;
    }
    if (tid == 0) { s_nlive = 0; s_cursor = 0; s_colours = 0; s_contacts = 0; s_pairs = 0; s_cstart[0] = 0; }
    __syncthreads();
    // A contact with a body that moves wakes a sleeping one (world.rs:138-145),
    // judged on the flags the frame began with so the result does not depend
    // on which thread looks first.
    for (uint32_t i = tid; i < n; i += T) s_claim[i] = 0u;
    __syncthreads();
    // A thread a body, over the pairs it owns (its count of its 32 slots).
    for (uint32_t i = tid; i < n; i += T) {
        const uint32_t c = count[i];
        for (uint32_t k = 0; k < c; ++k) {
            const PairRec pr = pairs[(size_t)i * PAIR_SLOTS + k];
            atomicAdd(&s_pairs, 1u);
            if (pr.m == 0) continue;
            atomicAdd(&s_contacts, pr.m);
            const uint32_t fa = s_flags[pr.a], fb = s_flags[pr.b];
            if ((fa & SLEEPING) && moves_f(fb)) s_claim[pr.a] = 1u;
            if ((fb & SLEEPING) && moves_f(fa)) s_claim[pr.b] = 1u;
        }
    }
    __syncthreads();
    for (uint32_t i = tid; i < n; i += T)
        if (s_claim[i]) {
            s_flags[i] &= ~SLEEPING;
            aux[i].timer = 0.f;
        }
    __syncthreads();
    // The pairs the solve works on: points, and a body that moves under them.
    for (uint32_t i = tid; i < n; i += T) {
        const uint32_t c = count[i];
        for (uint32_t k = 0; k < c; ++k) {
            const uint32_t s = i * PAIR_SLOTS + k;
            const PairRec pr = pairs[s];
            if (pr.m == 0 || !(dynamic_f(s_flags[pr.a]) || dynamic_f(s_flags[pr.b]))) continue;
            live[atomicAdd(&s_nlive, 1u)] = s;
        }
    }
    __syncthreads();
    // Colours: rounds of a matching in phiray's pair order. The live pairs'
    // bodies are kept in LDS (a | b << 16) when they fit, so a round reads
    // no global memory.
    const uint32_t nlive = s_nlive;
    const bool cached = nlive <= LIVE;
    if (cached) {
        for (uint32_t e = tid; e < nlive; e += T) {
            const PairRec pr = pairs[live[e]];
            s_ab[e] = pr.a | (pr.b << 16);
        }
        __syncthreads();
    }
    // The bodies of live pair e, false once it has its colour.
    auto bodies_of = [&](uint32_t e, uint32_t& a, uint32_t& b) {
        if (cached) {
            const uint32_t ab = s_ab[e];
            if (ab == NONE) return false;
            a = ab & 0xFFFFu;
            b = ab >> 16;
            return true;
        }
        const uint32_t s = live[e];
        if (s == NONE) return false;
        const PairRec pr = pairs[s];
        a = pr.a;
        b = pr.b;
        return true;
    };
    if constexpr (LPUSH) {
        for (uint32_t i = tid; i < n; i += T) s_deg[i] = 0u;
        if (tid == 0) s_err = 0u;
    }
    for (uint32_t round = 0; round < nlive; ++round) {
        if (s_cursor == nlive || s_colours == MAX_COLOURS) break;
        for (uint32_t i = tid; i < n; i += T) s_claim[i] = NONE;
        __syncthreads();
        for (uint32_t e = tid; e < nlive; e += T) {
            uint32_t a, b;
            if (!bodies_of(e, a, b)) continue;
            const uint32_t key = (a << 11) | b;  // priority()
            if (dynamic_f(s_flags[a])) atomicMin(&s_claim[a], key);
            if (dynamic_f(s_flags[b])) atomicMin(&s_claim[b], key);
        }
        __syncthreads();
        for (uint32_t e = tid; e < nlive; e += T) {
            uint32_t a, b;
            if (!bodies_of(e, a, b)) continue;
            const uint32_t key = (a << 11) | b;
            const bool ok = (!dynamic_f(s_flags[a]) || s_claim[a] == key) && (!dynamic_f(s_flags[b]) || s_claim[b] == key);
            if (ok) {
                const uint32_t at = atomicAdd(&s_cursor, 1u);
                order[at] = live[e];
                if (cached) s_ab[e] = NONE;
                else live[e] = NONE;
                if constexpr (LPUSH) {
                    // The pair's place among its moving bodies' pairs (one
                    // pair a body a round: the colour order is their order).
                    uint32_t ra = 0, rb = 0;
                    if (dynamic_f(s_flags[a])) { ra = s_deg[a]; s_deg[a] = ra + 1u; }
                    if (dynamic_f(s_flags[b])) { rb = s_deg[b]; s_deg[b] = rb + 1u; }
                    rank[at] = ra | (rb << 16);
                }
            }
        }
        __syncthreads();
        if (tid == 0) s_cstart[++s_colours] = s_cursor;
        __syncthreads();
    }
    const uint32_t ncol = s_colours;
    if (s_cursor != nlive) {  // more colours than MAX_COLOURS: report, change nothing
        if (tid == 0) { stats[0] = s_pairs; stats[1] = s_contacts; stats[2] = ncol; stats[3] = 2u; }
        return;
    }
    // The pairs and their points copied into colour order: a projection then
    // reads its records by its own index, one dependent load fewer.
    for (uint32_t e = tid; e < nlive; e += T) {
        const uint32_t s = order[e];
        const PairRec pr = pairs[s];
        spairs[e] = pr;
        for (uint32_t c = 0; c < pr.m; ++c) spoints[4 * (size_t)e + c] = points[4 * (size_t)s + c];
    }
    __syncthreads();
    const V grav = mk(P.gx, P.gy, P.gz);
    const float h = P.h;
    for (uint32_t sub = 0; sub < P.substeps; ++sub) {
        // Each point's closing speed before the substep; multipliers cleared.
        for (uint32_t e = tid; e < nlive; e += T) {
            const PairRec pr = spairs[e];
            const V nn = mk(pr.n.x, pr.n.y, pr.n.z);
            PointRec* pts = spoints + 4 * (size_t)e;
            const Q qa = L.q(pr.a), qb = L.q(pr.b);
            for (uint32_t c = 0; c < pr.m; ++c) {
                const V ra = rot(qa, mk(pts[c].aa.x, pts[c].aa.y, pts[c].aa.z)), rb = rot(qb, mk(pts[c].ab.x, pts[c].ab.y, pts[c].ab.z));
                const V vr = (L.v(pr.a) + cross(L.w(pr.a), ra)) - (L.v(pr.b) + cross(L.w(pr.b), rb));
                pts[c].e.x = dot(vr, nn);
                pts[c].aa.w = 0.f;
                pts[c].ab.w = 0.f;
            }
        }
        __syncthreads();
        // Half a kick of gravity, then the drift.
        for (uint32_t i = tid; i < n; i += T) {
            const V x = L.x(i);
            const Q q = L.q(i);
            prev[2 * i] = make_float4(x.x, x.y, x.z, 0.f);
            prev[2 * i + 1] = make_float4(q.x, q.y, q.z, q.w);
            push[2 * i] = make_float4(0.f, 0.f, 0.f, 0.f);
            push[2 * i + 1] = make_float4(0.f, 0.f, 0.f, 0.f);
            const uint32_t f = s_flags[i];
            if (!moves_f(f)) continue;
            V v = L.v(i);
            if (kind_of(f) == K_ACTIVE) v = v + grav * (0.5f * h);
            L.set_v(i, v);
            L.set_x(i, x + v * h);
            L.set_q(i, qstep(q, L.w(i), h));
        }
        __syncthreads();
        if constexpr (QUAD && LPUSH) {
            // Dataflow order, a wave at a time: a wave takes its groups'
            // entries of each colour (g, g + T/4, ... for its 16 groups of
            // four lanes), iteration after iteration, and projects them
            // together once every earlier projection at their moving bodies
            // is done (counted a body in LDS: iteration it, rank r at a body
            // of d live pairs waits for it d + r). Each body sees its pairs
            // in the colour order's sequence, as with a barrier between
            // colours, but a wave goes on as soon as its own pairs are
            // ready. The earliest pending entry is always at its wave's head
            // and ready, so the loop ends; the round bound only turns a
            // fault in that reasoning into an error.
            for (uint32_t i = tid; i < n; i += T) s_claim[i] = 0u;
            __syncthreads();
            const uint32_t g = tid / 4, lane = tid & 3u, wv = tid >> 6;
            bool failed = false;
            for (uint32_t it = 0; it < P.iterations && !failed; ++it)
                for (uint32_t c = 0; c < ncol && !failed; ++c)
                    for (uint32_t base = s_cstart[c]; base + 16u * wv < s_cstart[c + 1]; base += T / 4) {
                        const uint32_t e = base + g;
                        const bool has = e < s_cstart[c + 1];
                        PairPre cur;
                        uint32_t rk = 0;
                        if (has) {
                            pair_pre(cur, spairs, spoints, prev, kc, e, lane);
                            rk = rank[e];
                        }
                        const uint32_t a = has ? cur.pr.a : 0u, b = has ? cur.pr.b : 0u;
                        const bool da = has && dynamic_f(s_flags[a]), db = has && dynamic_f(s_flags[b]);
                        const uint32_t wa = it * (da ? s_deg[a] : 0u) + (rk & 0xFFFFu), wb = it * (db ? s_deg[b] : 0u) + (rk >> 16);
                        for (uint32_t spins = 0;; ++spins) {
                            const bool ok = (!da || __hip_atomic_load(&s_claim[a], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP) == wa) &&
                                            (!db || __hip_atomic_load(&s_claim[b], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP) == wb);
                            if (__ballot(!ok) == 0ull) break;
                            if (spins >= DATAFLOW_ROUNDS) {
                                failed = true;
                                break;
                            }
                            __builtin_amdgcn_s_sleep(1);
                        }
                        if (failed) {
                            s_err = 1u;
                            break;
                        }
                        if (has) project_quad(L, cur, s_flags, push, spoints + 4 * (size_t)e, P.cap, lane);
                        if (has && lane == 0u) {
                            if (da) __hip_atomic_fetch_add(&s_claim[a], 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
                            if (db) __hip_atomic_fetch_add(&s_claim[b], 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
                        }
                    }
            __syncthreads();
            if (s_err) {  // report, change nothing more
                if (tid == 0) { stats[0] = s_pairs; stats[1] = s_contacts; stats[2] = ncol; stats[3] = 3u; }
                return;
            }
        } else if constexpr (QUAD) {
            // A group of four lanes takes the colour's entries g, g + T/4, ...;
            // its first entry of the next colour is read while this one is
            // solved (the next iteration's colour 0 after the last colour).
            // With one colour the next entry is this one, whose multipliers
            // this projection changes: nothing is read ahead then.
            const uint32_t g = tid / 4, lane = tid & 3u;
            const bool ahead_ok = ncol > 1;
            PairPre nxt;
            bool have = ahead_ok && s_cstart[0] + g < s_cstart[1];
            if (have) pair_pre(nxt, spairs, spoints, prev, kc, s_cstart[0] + g, lane);
            for (uint32_t it = 0; it < P.iterations; ++it)
                for (uint32_t c = 0; c < ncol; ++c) {
                    const uint32_t e0 = s_cstart[c] + g, e1 = s_cstart[c + 1];
                    if (!ahead_ok) {
                        for (uint32_t e = e0; e < e1; e += T / 4) {
                            PairPre one;
                            pair_pre(one, spairs, spoints, prev, kc, e, lane);
                            project_quad(L, one, s_flags, push, spoints + 4 * (size_t)e, P.cap, lane);
                        }
                        __syncthreads();
                        continue;
                    }
                    const PairPre cur = nxt;
                    const bool mine = have;
                    const uint32_t nc = c + 1 < ncol ? c + 1 : 0u;
                    have = !(c + 1 == ncol && it + 1 == P.iterations) && s_cstart[nc] + g < s_cstart[nc + 1];
                    if (have) pair_pre(nxt, spairs, spoints, prev, kc, s_cstart[nc] + g, lane);
                    if (mine) project_quad(L, cur, s_flags, push, spoints + 4 * (size_t)e0, P.cap, lane);
                    for (uint32_t e = e0 + T / 4; e < e1; e += T / 4) {
                        PairPre more;
                        pair_pre(more, spairs, spoints, prev, kc, e, lane);
                        project_quad(L, more, s_flags, push, spoints + 4 * (size_t)e, P.cap, lane);
                    }
                    __syncthreads();
                }
        } else {
            for (uint32_t it = 0; it < P.iterations; ++it)
                for (uint32_t c = 0; c < ncol; ++c) {
                    for (uint32_t e = s_cstart[c] + tid; e < s_cstart[c + 1]; e += T)
                        project_pair(L, kc, s_flags, prev, push, spairs[e], spoints + 4 * (size_t)e, P.cap);
                    __syncthreads();
                }
        }
        // Velocities from the pose change against the drift's prediction,
        // less what the projections pushed out of standing penetrations.
        for (uint32_t i = tid; i < n; i += T) {
            if (!dynamic_f(s_flags[i])) continue;
            const float4 p0 = prev[2 * i], p1 = prev[2 * i + 1], u0 = push[2 * i], u1 = push[2 * i + 1];
            const V v = L.v(i), w = L.w(i);
            const V x_pred = mk(p0.x, p0.y, p0.z) + v * h;
            const Q q_pred = qstep(Q{p1.x, p1.y, p1.z, p1.w}, w, h);
            L.set_v(i, v + ((L.x(i) - x_pred) - mk(u0.x, u0.y, u0.z)) * (1.0f / h));
            L.set_w(i, w + (qrate(L.q(i), q_pred, h) - mk(u1.x, u1.y, u1.z) * (1.0f / h)));
        }
        __syncthreads();
        if constexpr (QUAD) {
            const uint32_t g = tid / 4, lane = tid & 3u;
            PairPre nxt;
            bool have = ncol > 0 && s_cstart[0] + g < s_cstart[1];
            if (have) pair_pre(nxt, spairs, spoints, prev, kc, s_cstart[0] + g, lane);
            for (uint32_t c = 0; c < ncol; ++c) {
                const uint32_t e0 = s_cstart[c] + g, e1 = s_cstart[c + 1];
                const PairPre cur = nxt;
                const bool mine = have;
                have = c + 1 < ncol && s_cstart[c + 1] + g < s_cstart[c + 2];
                if (have) pair_pre(nxt, spairs, spoints, prev, kc, s_cstart[c + 1] + g, lane);
                if (mine) velocity_quad(L, cur, s_flags, h, P.g, lane);
                for (uint32_t e = e0 + T / 4; e < e1; e += T / 4) {
                    PairPre more;
                    pair_pre(more, spairs, spoints, prev, kc, e, lane);
                    velocity_quad(L, more, s_flags, h, P.g, lane);
                }
                __syncthreads();
            }
        } else {
            for (uint32_t c = 0; c < ncol; ++c) {
                for (uint32_t e = s_cstart[c] + tid; e < s_cstart[c + 1]; e += T)
                    velocity_pair(L, kc, s_flags, spairs[e], spoints + 4 * (size_t)e, h, P.g);
                __syncthreads();
            }
        }
        // The other half kick, then Blender's damping.
        for (uint32_t i = tid; i < n; i += T) {
            if (!dynamic_f(s_flags[i])) continue;
            const Konst k = kc[i];
            const V v = (L.v(i) + grav * (0.5f * h)) * powf(fmaxf(1.0f - k.lin_damp, 0.f), h);
            L.set_v(i, v);
            L.set_w(i, L.w(i) * powf(fmaxf(1.0f - k.ang_damp, 0.f), h));
        }
        __syncthreads();
    }
    // Deactivation, the records back, and each body's object matrix.
    for (uint32_t i = tid; i < n; i += T) {
        uint32_t f = s_flags[i];
        V v = L.v(i), w = L.w(i);
        if (dynamic_f(f) && (f & DEACTIVATION)) {
            Aux& x = aux[i];
            const bool still = length(v) < x.sleep_lin && length(w) < x.sleep_ang;
            x.timer = still ? x.timer + P.dt : 0.f;
            if (x.timer >= P.sleep_after) {
                f |= SLEEPING;
                v = w = mk(0.f, 0.f, 0.f);
            }
        }
        const V x = L.x(i);
        const Q q = L.q(i);
        Body out;
        out.x[0] = x.x; out.x[1] = x.y; out.x[2] = x.z;
        out.inv_mass = L.im(i);
        out.q[0] = q.x; out.q[1] = q.y; out.q[2] = q.z; out.q[3] = q.w;
        out.v[0] = v.x; out.v[1] = v.y; out.v[2] = v.z;
        out.shape = __float_as_uint(s_body[4 * i + 2].w);
        out.w[0] = w.x; out.w[1] = w.y; out.w[2] = w.z;
        out.flags = f;
        bodies[i] = out;
        const Aux& x2 = aux[i];
        float* m = poses + 16 * (size_t)i;
        for (int c = 0; c < 3; ++c) {
            const V col = rot(q, unit_axis(c, x2.scale[c]));
            m[4 * c] = col.x; m[4 * c + 1] = col.y; m[4 * c + 2] = col.z; m[4 * c + 3] = 0.f;
        }
        const V t = x - rot(q, ld3(kc[i].com));
        m[12] = t.x; m[13] = t.y; m[14] = t.z; m[15] = 1.f;
    }
    if (tid == 0) { stats[0] = s_pairs; stats[1] = s_contacts; stats[2] = ncol; stats[3] = 0u; }
}

// The object matrices of bodies that have not stepped yet.
__global__ __launch_bounds__(256) void k_poses(const Body* __restrict__ bodies, const Konst* __restrict__ kc,
                                               const Aux* __restrict__ aux, uint32_t n, float* __restrict__ poses) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const Body b = bodies[i];
    const Q q{b.q[0], b.q[1], b.q[2], b.q[3]};
    float* m = poses + 16 * (size_t)i;
    for (int c = 0; c < 3; ++c) {
        const V col = rot(q, unit_axis(c, aux[i].scale[c]));
        m[4 * c] = col.x; m[4 * c + 1] = col.y; m[4 * c + 2] = col.z; m[4 * c + 3] = 0.f;
    }
    const V t = ld3(b.x) - rot(q, ld3(kc[i].com));
    m[12] = t.x; m[13] = t.y; m[14] = t.z; m[15] = 1.f;
}

// Host arithmetic for the body definitions (build.rs:99-131): inertia of the
// shape, the normalised orientation.
float host_len(const float* q, int n) {
    float s = 0.f;
    for (int k = 0; k < n; ++k) s += q[k] * q[k];
    return std::sqrt(s);
}

}  // namespace

int phr_bvh_internal_build(phr_gpu* g, phr_array* prims, uint32_t leaf, phr_bvh** out);
int phr_bvh_internal_rebuild_async(phr_bvh* b);
phr_array* phr_bvh_internal_nodes(phr_bvh* b);
phr_array* phr_bvh_internal_order(phr_bvh* b);

struct phr_world {
    phr_gpu* gpu = nullptr;
    // Host copies: authoritative until the first step after a change, then
    // refreshed from the device on demand.
    std::vector<Body> bodies;
    std::vector<Konst> konst;
    std::vector<Aux> aux;
    std::vector<phr_shape> shapes;
    bool dirty = true;        // host -> device before the next step
    bool host_stale = false;  // device -> host before the next read
    float gravity[3] = {0.f, 0.f, -9.81f};
    uint32_t substeps = 10, iterations = 10;
    float time_scale = 1.f, max_recover = 2.f, sleep_after = 2.f;
    uint32_t cap_bodies = 0, cap_shapes = 0;
    // device
    Body* d_bodies = nullptr;
    Konst* d_konst = nullptr;
    Aux* d_aux = nullptr;
    phr_shape* d_shapes = nullptr;
    phr_array* d_boxes = nullptr;
    uint32_t* d_count = nullptr;
    uint32_t* d_other = nullptr;
    PairRec* d_pairs = nullptr;
    PointRec* d_points = nullptr;
    PairRec* d_spairs = nullptr;    // the pairs in colour order
    uint32_t* d_rank = nullptr;     // their places among their moving bodies' pairs
    PointRec* d_spoints = nullptr;  // their points
    uint32_t* d_live = nullptr;
    uint32_t* d_order = nullptr;
    float4* d_prev = nullptr;
    float4* d_push = nullptr;
    float* d_poses = nullptr;
    uint32_t* d_stats = nullptr;  // pairs, contacts, colours, status; then the overflow word
    phr_bvh* tree = nullptr;
    uint32_t tree_n = 0;
    uint32_t stats[4] = {0, 0, 0, 0};
    bool quad = true;  // four lanes a pair in the solve's projections
};

namespace {

void world_release_device(phr_world* w) {
    for (void* p : {(void*)w->d_bodies, (void*)w->d_konst, (void*)w->d_aux, (void*)w->d_shapes, (void*)w->d_count,
                    (void*)w->d_other, (void*)w->d_pairs, (void*)w->d_points, (void*)w->d_spairs, (void*)w->d_spoints,
                    (void*)w->d_rank,
                    (void*)w->d_live, (void*)w->d_order,
                    (void*)w->d_prev, (void*)w->d_push, (void*)w->d_poses, (void*)w->d_stats})
        if (p) (void)hipFree(p);
    w->d_bodies = nullptr; w->d_konst = nullptr; w->d_aux = nullptr; w->d_shapes = nullptr; w->d_count = nullptr;
    w->d_other = nullptr; w->d_pairs = nullptr; w->d_points = nullptr; w->d_live = nullptr; w->d_order = nullptr;
    w->d_spairs = nullptr; w->d_spoints = nullptr; w->d_rank = nullptr;
    w->d_prev = nullptr; w->d_push = nullptr; w->d_poses = nullptr; w->d_stats = nullptr;
    phr_array_release(w->d_boxes);
    w->d_boxes = nullptr;
    w->cap_bodies = w->cap_shapes = 0;
    if (w->tree) phr_bvh_free(w->tree);  // it borrowed d_boxes
    w->tree = nullptr;
    w->tree_n = 0;
}

// Device copies of the bodies: a read brings them back first.
int world_pull(const phr_world* cw) {
    phr_world* w = const_cast<phr_world*>(cw);
    if (!w->host_stale) return PHR_OK;
    const size_t n = w->bodies.size();
    hipStream_t s = w->gpu->stream;
    PHR_HIP(hipMemcpyAsync(w->bodies.data(), w->d_bodies, sizeof(Body) * n, hipMemcpyDeviceToHost, s));
    PHR_HIP(hipMemcpyAsync(w->aux.data(), w->d_aux, sizeof(Aux) * n, hipMemcpyDeviceToHost, s));
    PHR_HIP(hipStreamSynchronize(s));
    w->host_stale = false;
    return PHR_OK;
}

int world_push(phr_world* w) {
    const uint32_t n = (uint32_t)w->bodies.size();
    hipStream_t s = w->gpu->stream;
    if (n > w->cap_bodies || w->shapes.size() > w->cap_shapes) {
        world_release_device(w);
        const uint32_t cap = std::max<uint32_t>(n, 16), caps = std::max<uint32_t>((uint32_t)w->shapes.size(), 4);
        const size_t slots = (size_t)cap * PAIR_SLOTS;
        auto dev = [&](void** p, size_t bytes) { return hipMalloc(p, std::max<size_t>(bytes, 16)) == hipSuccess; };
        const bool ok = dev((void**)&w->d_bodies, sizeof(Body) * cap) && dev((void**)&w->d_konst, sizeof(Konst) * cap) &&
                        dev((void**)&w->d_aux, sizeof(Aux) * cap) && dev((void**)&w->d_shapes, sizeof(phr_shape) * caps) &&
                        dev((void**)&w->d_count, 4ull * cap) && dev((void**)&w->d_other, 4 * slots) &&
                        dev((void**)&w->d_pairs, sizeof(PairRec) * slots) && dev((void**)&w->d_points, sizeof(PointRec) * 4 * slots) &&
                        dev((void**)&w->d_spairs, sizeof(PairRec) * slots) && dev((void**)&w->d_spoints, sizeof(PointRec) * 4 * slots) &&
                        dev((void**)&w->d_rank, 4 * slots) &&
                        dev((void**)&w->d_live, 4 * slots) && dev((void**)&w->d_order, 4 * slots) &&
                        dev((void**)&w->d_prev, 32ull * cap) && dev((void**)&w->d_push, 32ull * cap) &&
                        dev((void**)&w->d_poses, 64ull * cap) && dev((void**)&w->d_stats, 32);
        if (!ok) {
            world_release_device(w);
            phr_set_error_msg("World: device allocation failed");
            return PHR_ERR_HIP;
        }
        int rc = phr_array_alloc(w->gpu, 32, cap, &w->d_boxes);
        if (rc != PHR_OK) { world_release_device(w); return rc; }
        w->cap_bodies = cap;
        w->cap_shapes = caps;
    }
    w->d_boxes->len = n;
    PHR_HIP(hipMemcpyAsync(w->d_bodies, w->bodies.data(), sizeof(Body) * n, hipMemcpyHostToDevice, s));
    PHR_HIP(hipMemcpyAsync(w->d_konst, w->konst.data(), sizeof(Konst) * n, hipMemcpyHostToDevice, s));
    PHR_HIP(hipMemcpyAsync(w->d_aux, w->aux.data(), sizeof(Aux) * n, hipMemcpyHostToDevice, s));
    PHR_HIP(hipMemcpyAsync(w->d_shapes, w->shapes.data(), sizeof(phr_shape) * w->shapes.size(), hipMemcpyHostToDevice, s));
    if (n) hipLaunchKernelGGL(k_poses, dim3((n + 255) / 256), dim3(256), 0, s, w->d_bodies, w->d_konst, w->d_aux, n, w->d_poses);
    PHR_HIP(hipGetLastError());
    w->dirty = false;
    return PHR_OK;
}

}  // namespace

extern "C" {

int phr_world_new(phr_gpu* g, phr_world** out) {
    PHR_CHECK(g && out, PHR_ERR_ARG, "phr_world_new: null");
    phr_world* w = new phr_world();
    w->gpu = g;
    *out = w;
    return PHR_OK;
}

void phr_world_free(phr_world* w) {
    if (!w) return;
    world_release_device(w);
    delete w;
}

int phr_world_configure(phr_world* w, uint32_t substeps, uint32_t iterations, float time_scale, float max_recover,
                        float sleep_after, const float gravity[3]) {
    PHR_CHECK(w, PHR_ERR_ARG, "phr_world_configure: null");
    w->substeps = substeps;
    w->iterations = iterations;
    w->time_scale = time_scale;
    w->max_recover = max_recover;
    w->sleep_after = sleep_after;
    if (gravity) std::memcpy(w->gravity, gravity, sizeof w->gravity);
    return PHR_OK;
}

int phr_world_set_lanes(phr_world* w, uint32_t lanes) {
    PHR_CHECK(w, PHR_ERR_ARG, "phr_world_set_lanes: null");
    PHR_CHECK(lanes == 1u || lanes == 4u, PHR_ERR_ARG, "phr_world_set_lanes: 1 or 4 lanes a pair");
    w->quad = lanes == 4u;
    return PHR_OK;
}
uint32_t phr_world_lanes(const phr_world* w) { return w ? (w->quad ? 4u : 1u) : 0u; }

int phr_world_add_shape(phr_world* w, const phr_shape* s, uint32_t* id) {
    PHR_CHECK(w && s && id, PHR_ERR_ARG, "phr_world_add_shape: null");
    PHR_CHECK(s->kind == S_BOX || s->kind == S_SPHERE || s->kind == S_CAPSULE, PHR_ERR_UNSUPPORTED,
              "phr_world_add_shape: boxes, spheres and capsules only (hulls, cylinders, cones and meshes need GJK/EPA)");
    w->shapes.push_back(*s);
    w->dirty = true;
    *id = (uint32_t)w->shapes.size() - 1;
    return PHR_OK;
}

int phr_world_add_body(phr_world* w, const phr_body_def* d, uint32_t* id) {
    PHR_CHECK(w && d && id, PHR_ERR_ARG, "phr_world_add_body: null");
    PHR_CHECK(d->shape < w->shapes.size(), PHR_ERR_ARG, "phr_world_add_body: no such shape");
    PHR_CHECK(w->bodies.size() < MAX_BODIES, PHR_ERR_SIZE, "phr_world_add_body: at most 2,048 bodies (the solve keeps them in LDS)");
    PHR_CHECK(d->kind <= K_KINEMATIC, PHR_ERR_ARG, "phr_world_add_body: kind is 0 active, 1 passive or 2 kinematic");
    int rc = world_pull(w);
    if (rc != PHR_OK) return rc;
    const uint32_t kind = w->shapes[d->shape].kind;
    for (const Body& b : w->bodies) {
        const uint32_t other = w->shapes[b.shape].kind;
        PHR_CHECK(!((kind == S_CAPSULE && other == S_BOX) || (kind == S_BOX && other == S_CAPSULE)), PHR_ERR_UNSUPPORTED,
                  "phr_world_add_body: box-capsule pairs need GJK/EPA");
    }
    const phr_shape& s = w->shapes[d->shape];
    const bool movable = d->kind == K_ACTIVE && d->mass > 0.f;
    // Principal inertia of the shape about its centre (a capsule as its box).
    float in[3];
    if (s.kind == S_SPHERE) {
        in[0] = in[1] = in[2] = 0.4f * d->mass * s.radius * s.radius;
    } else {
        const float hx = s.kind == S_CAPSULE ? s.radius : s.half[0], hy = s.kind == S_CAPSULE ? s.radius : s.half[1];
        const float hz = s.kind == S_CAPSULE ? s.half[2] + s.radius : s.half[2];
        const float x2 = 4.f * hx * hx, y2 = 4.f * hy * hy, z2 = 4.f * hz * hz, k = d->mass / 12.f;
        in[0] = (y2 + z2) * k;
        in[1] = (x2 + z2) * k;
        in[2] = (x2 + y2) * k;
    }
    Body b{};
    std::memcpy(b.x, d->x, 12);
    b.inv_mass = movable ? 1.f / d->mass : 0.f;
    const float ql = host_len(d->q, 4);
    for (int k = 0; k < 4; ++k) b.q[k] = ql > 0.f ? d->q[k] / ql : (k == 3 ? 1.f : 0.f);
    std::memcpy(b.v, d->v, 12);
    b.shape = d->shape;
    std::memcpy(b.w, d->w, 12);
    b.flags = d->kind | (d->has_sleep ? DEACTIVATION : 0u) | (d->asleep && d->has_sleep ? SLEEPING : 0u);
    Konst k{};
    for (int c = 0; c < 3; ++c) k.inv_inertia[c] = movable && in[c] > 0.f ? 1.f / in[c] : 0.f;
    k.friction = d->friction;
    std::memcpy(k.com, d->com, 12);
    k.restitution = d->restitution;
    k.lin_damp = d->lin_damp;
    k.ang_damp = d->ang_damp;
    k.margin = d->margin;
    k.groups = d->groups;
    Aux a{};
    a.sleep_lin = d->has_sleep ? d->sleep[0] : 0.f;
    a.sleep_ang = d->has_sleep ? d->sleep[1] : 0.f;
    a.obj = d->obj;
    std::memcpy(a.scale, d->scale, 12);
    w->bodies.push_back(b);
    w->konst.push_back(k);
    w->aux.push_back(a);
    w->dirty = true;
    *id = (uint32_t)w->bodies.size() - 1;
    return PHR_OK;
}

int phr_world_step(phr_world* w, float frame_dt) {
    PHR_CHECK(w, PHR_ERR_ARG, "phr_world_step: null");
    const float dt = frame_dt * w->time_scale;
    if (dt <= 0.f || w->substeps == 0) return PHR_OK;
    const uint32_t n = (uint32_t)w->bodies.size();
    if (n == 0) return PHR_OK;
    int rc;
    if (w->dirty && (rc = world_push(w)) != PHR_OK) return rc;
    hipStream_t s = w->gpu->stream;
    hipLaunchKernelGGL(k_sweep, dim3((n + 255) / 256), dim3(256), 0, s, w->d_bodies, w->d_konst, w->d_shapes, n, dt,
                       (float4*)w->d_boxes->ptr);
    PHR_HIP(hipGetLastError());
    if (!w->tree || w->tree_n != n) {
        if (w->tree) phr_bvh_free(w->tree);
        w->tree = nullptr;
        if ((rc = phr_bvh_internal_build(w->gpu, w->d_boxes, 4, &w->tree)) != PHR_OK) return rc;  // Bvh::build(aabbs, 4)
        w->tree_n = n;
    } else if ((rc = phr_bvh_internal_rebuild_async(w->tree)) != PHR_OK) {
        return rc;
    }
    uint32_t* overflow = w->d_stats + 4;
    PHR_HIP(hipMemsetAsync(overflow, 0, 4, s));
    PHR_CHECK(phr_bvh_internal_nodes(w->tree)->len >= 2ull * n - 1, PHR_ERR_SIZE, "phr_world_step: the broadphase tree is short");
    if (n <= LDS_PAIR_BODIES)
        hipLaunchKernelGGL(k_pairs_lds, dim3((n + PAIR_BLOCK_BODIES - 1) / PAIR_BLOCK_BODIES), dim3(PAIR_STAGE_THREADS), 0, s, (const float4*)phr_bvh_internal_nodes(w->tree)->ptr,
                           (const uint32_t*)phr_bvh_internal_order(w->tree)->ptr, (const float4*)w->d_boxes->ptr, w->d_bodies,
                           w->d_konst, n, w->d_count, w->d_other, overflow);
    else
        hipLaunchKernelGGL(k_pairs, dim3((n + 255) / 256), dim3(256), 0, s, (const float4*)phr_bvh_internal_nodes(w->tree)->ptr,
                           (const uint32_t*)phr_bvh_internal_order(w->tree)->ptr, (const float4*)w->d_boxes->ptr, w->d_bodies,
                           w->d_konst, n, w->d_count, w->d_other, overflow);
    const uint32_t slots = n * PAIR_SLOTS;
    hipLaunchKernelGGL(k_contacts, dim3((slots + 127) / 128), dim3(128), 0, s, w->d_bodies, w->d_konst, w->d_shapes, w->d_count,
                       w->d_other, n, dt, w->d_pairs, w->d_points);
    Params P{};
    P.n = n;
    P.substeps = w->substeps;
    P.iterations = w->iterations;
    P.dt = dt;
    P.h = dt / (float)w->substeps;
    P.cap = std::max(w->max_recover * P.h / (float)std::max<uint32_t>(w->iterations, 1u), 0.f);
    P.g = std::sqrt(w->gravity[0] * w->gravity[0] + w->gravity[1] * w->gravity[1] + w->gravity[2] * w->gravity[2]);
    P.gx = w->gravity[0];
    P.gy = w->gravity[1];
    P.gz = w->gravity[2];
    P.sleep_after = w->sleep_after;
    // A pair list that overflowed stops the solve before it changes anything
    // (status 1); more colours than it can hold, too (status 2).
    auto solve = [&](auto kern) {
        hipLaunchKernelGGL(kern, dim3(1), dim3(SOLVE_THREADS), 0, s, w->d_bodies, w->d_konst, w->d_aux, w->d_count, w->d_pairs,
                           w->d_points, w->d_live, w->d_order, w->d_spairs, w->d_spoints, w->d_rank, w->d_prev, w->d_push,
                           w->d_poses, w->d_stats, P);
    };
    if (w->quad && n <= LPUSH_BODIES)
        solve(k_solve<true, LPUSH_BODIES, true>);
    else if (w->quad)
        solve(k_solve<true, MAX_BODIES, false>);
    else
        solve(k_solve<false, MAX_BODIES, false>);
    PHR_HIP(hipGetLastError());
    PHR_HIP(hipMemcpyAsync(w->stats, w->d_stats, 16, hipMemcpyDeviceToHost, s));
    PHR_HIP(hipStreamSynchronize(s));
    w->host_stale = true;
    PHR_CHECK(w->stats[3] != 1u, PHR_ERR_SIZE, "phr_world_step: a body meets more than 32 others");
    PHR_CHECK(w->stats[3] != 2u, PHR_ERR_SIZE, "phr_world_step: the pairs need more than 1,024 colours");
    PHR_CHECK(w->stats[3] == 0u, PHR_ERR_SIZE, "phr_world_step: the solve's dataflow order did not finish");
    return PHR_OK;
}

uint32_t phr_world_body_count(const phr_world* w) { return w ? (uint32_t)w->bodies.size() : 0u; }
uint32_t phr_world_contact_count(const phr_world* w) { return w ? w->stats[1] : 0u; }

int phr_world_bodies(const phr_world* w, void* out) {
    PHR_CHECK(w && out, PHR_ERR_ARG, "phr_world_bodies: null");
    int rc = world_pull(w);
    if (rc != PHR_OK) return rc;
    std::memcpy(out, w->bodies.data(), sizeof(Body) * w->bodies.size());
    return PHR_OK;
}

int phr_world_poses(const phr_world* w, float* out) {
    PHR_CHECK(w && out, PHR_ERR_ARG, "phr_world_poses: null");
    phr_world* m = const_cast<phr_world*>(w);
    if (m->bodies.empty()) return PHR_OK;
    int rc;
    if (m->dirty && (rc = world_push(m)) != PHR_OK) return rc;
    PHR_HIP(hipMemcpyAsync(out, m->d_poses, 64ull * m->bodies.size(), hipMemcpyDeviceToHost, m->gpu->stream));
    PHR_HIP(hipStreamSynchronize(m->gpu->stream));
    return PHR_OK;
}

float phr_world_energy(const phr_world* w) {
    if (!w || world_pull(w) != PHR_OK) return 0.f;
    double total = 0.0;
    for (size_t i = 0; i < w->bodies.size(); ++i) {
        const Body& b = w->bodies[i];
        const Konst& k = w->konst[i];
        if (!(b.inv_mass > 0.f)) continue;
        const double m = 1.0 / b.inv_mass;
        // The spin in the body frame, for the principal inertia.
        const double qx = b.q[0], qy = b.q[1], qz = b.q[2], qw = b.q[3];
        const double vx = b.w[0], vy = b.w[1], vz = b.w[2];
        const double tx = 2 * (-qy * vz + qz * vy), ty = 2 * (-qz * vx + qx * vz), tz = 2 * (-qx * vy + qy * vx);
        const double wb[3] = {vx + qw * tx + (-qy * tz + qz * ty), vy + qw * ty + (-qz * tx + qx * tz), vz + qw * tz + (-qx * ty + qy * tx)};
        double rotk = 0.0;
        for (int c = 0; c < 3; ++c) rotk += k.inv_inertia[c] > 0.f ? 0.5 * wb[c] * wb[c] / k.inv_inertia[c] : 0.0;
        const double v2 = (double)b.v[0] * b.v[0] + (double)b.v[1] * b.v[1] + (double)b.v[2] * b.v[2];
        const double gx = (double)w->gravity[0] * b.x[0] + (double)w->gravity[1] * b.x[1] + (double)w->gravity[2] * b.x[2];
        total += 0.5 * m * v2 + rotk - m * gx;
    }
    return (float)total;
}

int phr_world_pair_list(const phr_world* w, uint32_t* out, uint32_t cap, uint32_t* n) {
    PHR_CHECK(w && n, PHR_ERR_ARG, "phr_world_pair_list: null");
    const uint32_t nb = (uint32_t)w->bodies.size();
    std::vector<std::pair<uint32_t, uint32_t>> list;
    if (nb && w->d_count && w->tree) {
        std::vector<uint32_t> cnt(nb), other((size_t)nb * PAIR_SLOTS);
        PHR_HIP(hipMemcpyAsync(cnt.data(), w->d_count, 4ull * nb, hipMemcpyDeviceToHost, w->gpu->stream));
        PHR_HIP(hipMemcpyAsync(other.data(), w->d_other, 4ull * other.size(), hipMemcpyDeviceToHost, w->gpu->stream));
        PHR_HIP(hipStreamSynchronize(w->gpu->stream));
        for (uint32_t i = 0; i < nb; ++i)
            for (uint32_t k = 0; k < cnt[i]; ++k) {
                const uint32_t j = other[(size_t)i * PAIR_SLOTS + k];
                list.push_back({std::min(i, j), std::max(i, j)});
            }
        std::sort(list.begin(), list.end());
    }
    *n = (uint32_t)list.size();
    if (out) {
        PHR_CHECK(cap >= *n, PHR_ERR_SIZE, "phr_world_pair_list: buffer too small");
        for (uint32_t k = 0; k < *n; ++k) out[2 * k] = list[k].first, out[2 * k + 1] = list[k].second;
    }
    return PHR_OK;
}

}  // extern "C"

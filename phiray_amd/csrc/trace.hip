// The Rendered-mode tracer on gfx950: scene arrays in leaf order, the BVH walk
// and the path-tracing kernel.
//
// Walk (src/trace/wgsl/bvh.wgsl:120-190): node halves are loaded as two
// 16-byte vectors, both children's boxes are tested at an inner node, the far
// one is pushed when both are hit (tie to the left, `tl <= tr`), and a leaf's
// triangles are tested with the watertight test (src/trace/tri.rs:83-114)
// keeping only strictly nearer hits. The 32-slot stack lives in LDS, one column
// per thread so the 64 lanes of a wave hit 64 different banks.
//
// Integrator (src/trace/wgsl/shade.wgsl:41-373 and the files it calls): one
// thread per pixel runs the camera path for each requested sample in turn,
// with the per-pixel PCG stream seeded as the WGSL seeds it, and adds each
// sample into the image buffer's sums in order. The restated feature subset
// and every operation order are those of the CPU twin.
#include "phr_records.h"
#include "phr_runtime.h"

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <unordered_map>
#include <vector>

using namespace phr;


constexpr uint32_t MAT_WORDS = 16;
constexpr uint32_t PX_LEN = 34;  // src/trace/read.rs:56

constexpr uint32_t INSTANCE_LEN = 40;  // words of an instance record (src/trace/instance.rs:18)
constexpr uint32_t PHR_INST = 1u << 30;  // a top-tree leaf's range is instance slots (src/trace/forest.rs:17)

struct SceneDev {
    const float4* nodes;
    const float4* recs;    // one tree's walk records (4 float4 a node index); a forest walks `nodes`
    const float4* kids;    // the records' second level (8 float4 a node index), after the first
    const float4* tris;    // 3 float4 per slot: (a.xyz, b.x) (b.yz, c.xy) (c.z, -, -, -)
    const uint4* faces;    // slot order: 3 vertex ids + material
    const float4* verts;   // 2 float4 per vertex: (pos, 1) (oct.xy, u, v)
    const float* mats;     // MAT_WORDS per material
    const uint32_t* lights;
    // A forest: instance records (world rows 0-11, inverse rows 12-23, tree
    // root 24, face range 25-26, flags 27), then the top-tree slot order.
    const uint32_t* insts;
    uint32_t ninsts;
    // One tree: its root node (0 for a tracer's own tree; a batch of scenes
    // shares one node array, each scene's tree at its own base).
    uint32_t root;
};

namespace {

// ---- walk -------------------------------------------------------------------
struct RaySpace { uint32_t kz; bool flip; float sx, sy, sz; };

__device__ __forceinline__ void ray_axes(float vx, float vy, float vz, uint32_t kz, bool flip, float& ox, float& oy, float& oz) {
    float px = kz == 0 ? vy : (kz == 1 ? vz : vx);
    float py = kz == 0 ? vz : (kz == 1 ? vx : vy);
    float pz = kz == 0 ? vx : (kz == 1 ? vy : vz);
    ox = flip ? py : px;
    oy = flip ? px : py;
    oz = pz;
}

__device__ __forceinline__ RaySpace ray_space(V3 d) {
    float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
    uint32_t kz = (ax >= ay && ax >= az) ? 0u : (ay >= az ? 1u : 2u);
    float x, y, z;
    ray_axes(d.x, d.y, d.z, kz, false, x, y, z);
    bool flip = z < 0.f;
    ray_axes(d.x, d.y, d.z, kz, flip, x, y, z);
    float sz = 1.0f / z;
    return RaySpace{kz, flip, x * sz, y * sz, sz};
}

// Returns t (PHR_INF on a miss) and the weights of b and c.
__device__ __forceinline__ float tri_t(float4 t0, float4 t1, float4 t2, V3 o, const RaySpace& r, float& u, float& v) {
    float ax, ay, az, bx, by, bz, cx, cy, cz;
    ray_axes(t0.x - o.x, t0.y - o.y, t0.z - o.z, r.kz, r.flip, ax, ay, az);
    ray_axes(t0.w - o.x, t1.x - o.y, t1.y - o.z, r.kz, r.flip, bx, by, bz);
    ray_axes(t1.z - o.x, t1.w - o.y, t2.x - o.z, r.kz, r.flip, cx, cy, cz);
    float sax = ax - r.sx * az, say = ay - r.sy * az;
    float sbx = bx - r.sx * bz, sby = by - r.sy * bz;
    float scx = cx - r.sx * cz, scy = cy - r.sy * cz;
    float ea = scx * sby - scy * sbx;
    float eb = sax * scy - say * scx;
    float ec = sbx * say - sby * sax;
    if ((ea < 0.f || eb < 0.f || ec < 0.f) && (ea > 0.f || eb > 0.f || ec > 0.f)) return PHR_INF;
    float det = ea + eb + ec;
    if (det == 0.f) return PHR_INF;
    float tt = ea * (r.sz * az) + eb * (r.sz * bz) + ec * (r.sz * cz);
    float sg = det < 0.f ? -1.0f : 1.0f;
    if (tt * sg <= 0.f) return PHR_INF;
    float inv = 1.0f / det;
    u = eb * inv;
    v = ec * inv;
    return tt * inv;
}

__device__ __forceinline__ V3 safe_inv(V3 d) {  // bvh.wgsl:111-114
    float x = fabsf(d.x) < 1e-9f ? (d.x < 0.f ? -1.0f : 1.0f) * 1e-9f : d.x;
    float y = fabsf(d.y) < 1e-9f ? (d.y < 0.f ? -1.0f : 1.0f) * 1e-9f : d.y;
    float z = fabsf(d.z) < 1e-9f ? (d.z < 0.f ? -1.0f : 1.0f) * 1e-9f : d.z;
    return v3(1.0f / x, 1.0f / y, 1.0f / z);
}

__device__ __forceinline__ float box_t(float4 lo, float4 hi, V3 o, V3 inv, float best) {  // bvh.wgsl:14-21
    float t1x = (lo.x - o.x) * inv.x, t1y = (lo.y - o.y) * inv.y, t1z = (lo.z - o.z) * inv.z;
    float t2x = (hi.x - o.x) * inv.x, t2y = (hi.y - o.y) * inv.y, t2z = (hi.z - o.z) * inv.z;
    float tmin = fmaxf(fmaxf(fminf(t1x, t2x), fminf(t1y, t2y)), fminf(t1z, t2z));
    float tmax = fminf(fminf(fmaxf(t1x, t2x), fmaxf(t1y, t2y)), fmaxf(t1z, t2z));
    float e = fmaxf(tmin, 0.f);
    return (tmax >= e && tmin < best) ? e : PHR_INF;
}

// trace_blas for one identity instance, root at node `root`, over the tree's
// walk records: record x (node x an inner node) holds x's two children side
// by side as 32-byte nodes, an inner child's words replaced by (its own node
// index, 0) and a leaf child's kept (first | LEAF, end). A step reads one
// aligned 64-byte record and tests both boxes; the stack holds the far
// child's first word, so a pop needs no load of its own. A leaf taken from
// the stack learns its end from the cache row of its first triangle (row 2,
// word 1). The visits, and so the hits and counts, are those of the node
// walk: both children tested, the nearer entered first (a tie to the left),
// the farther pushed. `stk` is this thread's LDS stack column (stride =
// threads per block).
// Any-hit rays skip triangles of glass or cut-out materials (trans > 0.5 or
// alpha < 0.5, bvh.wgsl:158-160): `faces` and `mats` give the material.
// `steps` and `tests` count inner-node visits and triangle tests, the units
// the walk's memory traffic is priced in.
// Walk records, for each node index x of an inner node a walk reaches: its
// children A and B as two Nodes (4 float4 at recs[4x], dense, what the
// binary step reads), and, for each child that is inner, that child's two
// children (8 float4 at kids[8x]: A's, then B's). An inner child's words
// are (its node index, 0); a leaf child keeps (first | LEAF, end). Both
// live in one array of REC_ROWS float4 a node: recs for every node, then
// kids for every node.
constexpr uint32_t REC_ROWS = 12;

// The walk's visit of inner node `cur` from its record: the binary step at
// cur and, when it goes into an inner child, the step at that child too,
// whose children's boxes ride in the same record. No triangle test comes
// between the two, so both use the same best.t and make exactly the binary
// walk's decisions and pushes; the record's three 64-B lines are read at
// once (r = recs + 4 cur, k = kids + 8 cur). False when nothing is entered
// (the caller pops).
// visit1 is the binary step alone (the record's first line), for kernels
// whose registers have no room for the other two lines.
__device__ __forceinline__ bool visit1(const float4* __restrict__ r, V3 o, V3 inv, float best_t, uint32_t& cur, uint32_t& end,
                                       uint32_t* stk, uint32_t stride, uint32_t& sp, uint32_t& steps) {
    const float4 llo = r[0], lhi = r[1], qlo = r[2], qhi = r[3];
    ++steps;
    const float tl = box_t(llo, lhi, o, inv, best_t);
    const float tr = box_t(qlo, qhi, o, inv, best_t);
    if (tl < PHR_INF && tr < PHR_INF) {
        if (tl <= tr) {
            stk[sp * stride] = __float_as_uint(qlo.w);
            cur = __float_as_uint(llo.w);
            end = __float_as_uint(lhi.w);
        } else {
            stk[sp * stride] = __float_as_uint(llo.w);
            cur = __float_as_uint(qlo.w);
            end = __float_as_uint(qhi.w);
        }
        ++sp;
        return true;
    }
    if (tl < PHR_INF) { cur = __float_as_uint(llo.w); end = __float_as_uint(lhi.w); return true; }
    if (tr < PHR_INF) { cur = __float_as_uint(qlo.w); end = __float_as_uint(qhi.w); return true; }
    return false;
}
// visit2_rows: the visit from rows already loaded (the record's four, then
// its eight kids rows); visit2 loads them from the record.
__device__ __forceinline__ bool visit2_rows(float4 alo, float4 ahi, float4 blo, float4 bhi, float4 a0lo, float4 a0hi, float4 a1lo,
                                            float4 a1hi, float4 b0lo, float4 b0hi, float4 b1lo, float4 b1hi, V3 o, V3 inv, float best_t,
                                            uint32_t& cur, uint32_t& end, uint32_t* stk, uint32_t stride, uint32_t& sp,
                                            uint32_t& steps) {
    ++steps;
    const float ta = box_t(alo, ahi, o, inv, best_t);
    const float tb = box_t(blo, bhi, o, inv, best_t);
    bool into_b;
    if (ta < PHR_INF && tb < PHR_INF) {
        into_b = !(ta <= tb);
        stk[sp * stride] = __float_as_uint(into_b ? alo.w : blo.w);
        ++sp;
    } else if (ta < PHR_INF) {
        into_b = false;
    } else if (tb < PHR_INF) {
        into_b = true;
    } else {
        return false;
    }
    const uint32_t cw = __float_as_uint(into_b ? blo.w : alo.w);
    if (cw & PHR_LEAF) {
        cur = cw;
        end = __float_as_uint(into_b ? bhi.w : ahi.w);
        return true;
    }
    ++steps;
    const float4 llo = into_b ? b0lo : a0lo, lhi = into_b ? b0hi : a0hi;
    const float4 qlo = into_b ? b1lo : a1lo, qhi = into_b ? b1hi : a1hi;
    const float tl = box_t(llo, lhi, o, inv, best_t);
    const float tr = box_t(qlo, qhi, o, inv, best_t);
    if (tl < PHR_INF && tr < PHR_INF) {
        if (tl <= tr) {
            stk[sp * stride] = __float_as_uint(qlo.w);
            cur = __float_as_uint(llo.w);
            end = __float_as_uint(lhi.w);
        } else {
            stk[sp * stride] = __float_as_uint(llo.w);
            cur = __float_as_uint(qlo.w);
            end = __float_as_uint(qhi.w);
        }
        ++sp;
        return true;
    }
    if (tl < PHR_INF) { cur = __float_as_uint(llo.w); end = __float_as_uint(lhi.w); return true; }
    if (tr < PHR_INF) { cur = __float_as_uint(qlo.w); end = __float_as_uint(qhi.w); return true; }
    return false;
}
__device__ __forceinline__ bool visit2(const float4* __restrict__ r, const float4* __restrict__ k, V3 o, V3 inv, float best_t,
                                       uint32_t& cur, uint32_t& end, uint32_t* stk, uint32_t stride, uint32_t& sp, uint32_t& steps) {
    return visit2_rows(r[0], r[1], r[2], r[3], k[0], k[1], k[2], k[3], k[4], k[5], k[6], k[7], o, inv, best_t, cur, end, stk, stride, sp,
                       steps);
}

// TWO: a visit takes two levels from the record (visit2), else one (visit1);
// the same decisions either way.
template <bool ANY, bool TWO = false>
__device__ __forceinline__ PhrHit walk(const float4* __restrict__ nodes, const float4* __restrict__ recs, const float4* __restrict__ kids,
                                       uint32_t root,
                                       const float4* __restrict__ tris, const uint4* __restrict__ faces, const float* __restrict__ mats,
                                       V3 o, V3 d, float tmax, uint32_t* stk, uint32_t stride, uint32_t& steps, uint32_t& tests) {
    PhrHit best{tmax, PHR_NONE, 0.f, 0.f, 0u};
    V3 inv = safe_inv(d);
    const float4 rlo = nodes[2 * root], rhi = nodes[2 * root + 1];
    if (box_t(rlo, rhi, o, inv, best.t) >= PHR_INF) return best;
    RaySpace rs = ray_space(d);
    uint32_t sp = 0;
    // The entry in hand: an inner node's index (its record), or a leaf's
    // first | LEAF with its end (PHR_NONE when it came off the stack).
    uint32_t cur = __float_as_uint(rlo.w), end = __float_as_uint(rhi.w);
    if (!(cur & PHR_LEAF)) cur = root;
    for (;;) {
        if (cur & PHR_LEAF) {
            uint32_t id = cur & ~PHR_LEAF;
            float4 t0 = tris[3 * id], t1 = tris[3 * id + 1], t2 = tris[3 * id + 2];
            if (end == PHR_NONE) end = __float_as_uint(t2.y);
            for (;;) {
                ++tests;
                // The leaf's next triangle is loaded while this one is tested.
                const bool more = id + 1u < end;
                float4 n0 = t0, n1 = t1, n2 = t2;
                if (more) {
                    n0 = tris[3 * (id + 1u)];
                    n1 = tris[3 * (id + 1u) + 1];
                    n2 = tris[3 * (id + 1u) + 2];
                }
                float u, v;
                float t = tri_t(t0, t1, t2, o, rs, u, v);
                if (t < best.t) {
                    if (ANY) {
                        const float* m = mats + MAT_WORDS * (size_t)faces[id].w;
                        if (!(m[6] > 0.5f || m[3] < 0.5f)) return PhrHit{t, id, u, v, 0u};
                    } else {
                        best = PhrHit{t, id, u, v, 0u};
                    }
                }
                if (!more) break;
                ++id;
                t0 = n0;
                t1 = n1;
                t2 = n2;
            }
        } else if (TWO ? visit2(recs + 4 * (size_t)cur, kids + 8 * (size_t)cur, o, inv, best.t, cur, end, stk, stride, sp, steps)
                       : visit1(recs + 4 * (size_t)cur, o, inv, best.t, cur, end, stk, stride, sp, steps)) {
            continue;
        }
        if (sp == 0) break;
        --sp;
        cur = stk[sp * stride];
        end = PHR_NONE;
    }
    return best;
}

// A walk that can stop after any unit and go on later (the split build):
// the ray, its best hit, where it is (cur, end) and its stack (sp entries of
// the thread's column). A unit is an inner node's record (both boxes, the
// nearer child entered and the farther pushed) or one triangle, so a walk
// makes walk<ANY>'s visits, decisions and pushes in walk<ANY>'s order and
// ends with its hit. `any`: a shadow ray's walk, which ends at the first
// triangle that blocks it.
struct Walker {
    V3 o, inv;
    RaySpace rs;
    float t, u, v;  // closest: the nearest hit so far; shadow: t is the ray's length
    uint32_t tri, cur, end, sp;
    bool on, any;
};
__device__ __forceinline__ void walker_begin(Walker& w, const SceneDev& s, V3 o, V3 d, float tmax, bool any) {
    w.o = o;
    w.inv = safe_inv(d);
    w.rs = ray_space(d);
    w.t = tmax;
    w.u = w.v = 0.f;
    w.tri = PHR_NONE;
    w.sp = 0;
    w.any = any;
    const float4 rlo = s.nodes[2 * s.root], rhi = s.nodes[2 * s.root + 1];
    w.on = box_t(rlo, rhi, o, w.inv, tmax) < PHR_INF;
    w.cur = __float_as_uint(rlo.w);
    w.end = __float_as_uint(rhi.w);
    if (!(w.cur & PHR_LEAF)) w.cur = s.root;
}
// TWO: an inner node's unit is visit2's (two levels from the record and its
// kids rows), else the binary step.
template <bool TWO = false>
__device__ __forceinline__ void walker_step(Walker& w, const SceneDev& s, uint32_t* stk, uint32_t stride, uint32_t& steps,
                                            uint32_t& tests) {
    const bool leaf = (w.cur & PHR_LEAF) != 0u;
    const float4* q = leaf ? s.tris + 3 * (size_t)(w.cur & ~PHR_LEAF) : s.recs + 4 * (size_t)w.cur;
    const float4 a = q[0], b = q[1], c = q[2];
    float4 d = a, k0 = a, k1 = a, k2 = a, k3 = a, k4 = a, k5 = a, k6 = a, k7 = a;
    if (!leaf) {
        d = q[3];
        if constexpr (TWO) {
            const float4* kk = s.kids + 8 * (size_t)w.cur;
            k0 = kk[0];
            k1 = kk[1];
            k2 = kk[2];
            k3 = kk[3];
            k4 = kk[4];
            k5 = kk[5];
            k6 = kk[6];
            k7 = kk[7];
        }
    }
    if constexpr (TWO) {
        if (!leaf) {
            if (visit2_rows(a, b, c, d, k0, k1, k2, k3, k4, k5, k6, k7, w.o, w.inv, w.t, w.cur, w.end, stk, stride, w.sp, steps)) return;
            if (w.sp == 0) {
                w.on = false;
                return;
            }
            --w.sp;
            w.cur = stk[w.sp * stride];
            w.end = PHR_NONE;
            return;
        }
    }
    if (leaf) {
        const uint32_t id = w.cur & ~PHR_LEAF;
        ++tests;
        if (w.end == PHR_NONE) w.end = __float_as_uint(c.y);
        float u, v;
        const float t = tri_t(a, b, c, w.o, w.rs, u, v);
        if (t < w.t) {
            if (w.any) {
                const float* m = s.mats + MAT_WORDS * (size_t)s.faces[id].w;
                if (!(m[6] > 0.5f || m[3] < 0.5f)) {
                    w.tri = id;
                    w.on = false;
                    return;
                }
            } else {
                w.t = t;
                w.tri = id;
                w.u = u;
                w.v = v;
            }
        }
        if (id + 1u < w.end) {
            w.cur = (id + 1u) | PHR_LEAF;
            return;
        }
    } else {
        ++steps;
        const float tl = box_t(a, b, w.o, w.inv, w.t);
        const float tr = box_t(c, d, w.o, w.inv, w.t);
        const bool hl = tl < PHR_INF, hr = tr < PHR_INF;
        if (hl || hr) {
            const bool left = hl && (!hr || tl <= tr);
            if (hl && hr) {
                stk[w.sp * stride] = __float_as_uint(left ? c.w : a.w);
                ++w.sp;
            }
            w.cur = __float_as_uint(left ? a.w : c.w);
            w.end = __float_as_uint(left ? b.w : d.w);
            return;
        }
    }
    if (w.sp == 0) {
        w.on = false;
        return;
    }
    --w.sp;
    w.cur = stk[w.sp * stride];
    w.end = PHR_NONE;
}

// The walk records of a tree's nodes (one entry per node index; only inner
// nodes a walk reaches get one, as `inner` marks them), and each reachable
// leaf's end in the cache row of its first triangle.
__device__ __forceinline__ void rec_child(const float4* __restrict__ nodes, uint32_t y, float4* __restrict__ out,
                                          float4* __restrict__ tris) {
    float4 lo = nodes[2 * (size_t)y], hi = nodes[2 * (size_t)y + 1];
    const uint32_t w0 = __float_as_uint(lo.w);
    if (w0 & PHR_LEAF) {
        tris[3 * (size_t)(w0 & ~PHR_LEAF) + 2].y = hi.w;  // the leaf's end
    } else {
        lo.w = __uint_as_float(y);
        hi.w = 0.f;
    }
    out[0] = lo;
    out[1] = hi;
}
__global__ __launch_bounds__(256) void k_walk_records(const float4* __restrict__ nodes, const uint8_t* __restrict__ inner,
                                                      uint32_t count, float4* __restrict__ recs, float4* __restrict__ tris) {
    const uint32_t x = blockIdx.x * 256 + threadIdx.x;
    if (x >= count || !inner[x]) return;
    float4* r = recs + 4 * (size_t)x;
    float4* k = recs + 4 * (size_t)count + 8 * (size_t)x;
    const uint32_t kid[2] = {__float_as_uint(nodes[2 * (size_t)x].w), __float_as_uint(nodes[2 * (size_t)x + 1].w)};
    for (int c = 0; c < 2; ++c) {
        const uint32_t y = kid[c];
        rec_child(nodes, y, r + 2 * c, tris);
        const float4 ylo = nodes[2 * (size_t)y];
        if (__float_as_uint(ylo.w) & PHR_LEAF) continue;
        // An inner child's own children (their leaf ends are written by
        // that child's record too).
        const uint32_t g0 = __float_as_uint(ylo.w), g1 = __float_as_uint(nodes[2 * (size_t)y + 1].w);
        rec_child(nodes, g0, k + 4 * c, tris);
        rec_child(nodes, g1, k + 4 * c + 2, tris);
    }
}

// An instance's row times (x, y, z, w), summed in trace_blas's order.
__device__ __forceinline__ float dot4(float4 r, float x, float y, float z, float w) {
    return ((r.x * x + r.y * y) + r.z * z) + r.w * w;
}

// trace (bvh.wgsl:194-232): the top tree over the instances' world boxes; a
// leaf's slots name the instances (inst_order) whose trees the ray walks.
// trace_blas (bvh.wgsl:120-190) for instance i: the ray taken into the
// mesh's space by the instance's inverse rows (direction left unnormalised,
// so t stays the world's) unless the instance sits at the identity (word 27
// bit 0); a shadow ray passes an instance that casts none (bit 1). The
// nearest hit so far rides in `best`, so a tree only takes strictly nearer
// hits.
// The top tree's stack holds 16-bit node ids (its nodes are the array's
// first 2 insts - 1, fewer than 65,536: the tracer refuses more instances),
// a mesh tree's stack 32-bit ones; each is the thread's column of stride.
template <bool ANY>
__device__ PhrHit walk_forest(const SceneDev& s, V3 o, V3 d, float tmax, uint16_t* tstk, uint32_t* bstk, uint32_t stride,
                              uint32_t& steps, uint32_t& tests) {
    PhrHit best{tmax, PHR_NONE, 0.f, 0.f, 0u};
    const uint32_t* inst_order = s.insts + INSTANCE_LEN * (size_t)s.ninsts;
    const V3 winv = safe_inv(d);
    float4 lo = s.nodes[0], hi = s.nodes[1];
    if (box_t(lo, hi, o, winv, best.t) >= PHR_INF) return best;
    // One loop over both levels (the nested walk's visits in its order): a
    // step is a node of the top tree or of the current instance's mesh tree,
    // the same code either way, so lanes at different levels still step
    // together. At a top-tree leaf the ray goes into its instances' trees one
    // after another (trace_blas); when the last is done the top tree pops.
    uint32_t li = __float_as_uint(lo.w), ri = __float_as_uint(hi.w);
    uint32_t tsp = 0, bsp = 0;
    bool in_mesh = false;
    uint32_t inst = 0, k = 0, kend = 0;
    V3 ro = o, rinv = winv;  // the ray the current level walks with
    RaySpace rs{0u, false, 0.f, 0.f, 0.f};
    for (;;) {
        bool pop = false;
        if (!(li & PHR_LEAF)) {
            ++steps;
            const float4 llo = s.nodes[2 * li], lhi = s.nodes[2 * li + 1];
            const float4 rlo = s.nodes[2 * ri], rhi = s.nodes[2 * ri + 1];
            const float tl = box_t(llo, lhi, ro, rinv, best.t);
            const float tr = box_t(rlo, rhi, ro, rinv, best.t);
            if (tl < PHR_INF && tr < PHR_INF) {
                const uint32_t far = tl <= tr ? ri : li;
                if (in_mesh) bstk[bsp++ * stride] = far;
                else tstk[tsp++ * stride] = (uint16_t)far;
                const float4 nl = tl <= tr ? llo : rlo, nh = tl <= tr ? lhi : rhi;
                li = __float_as_uint(nl.w);
                ri = __float_as_uint(nh.w);
                continue;
            }
            if (tl < PHR_INF) { li = __float_as_uint(llo.w); ri = __float_as_uint(lhi.w); continue; }
            if (tr < PHR_INF) { li = __float_as_uint(rlo.w); ri = __float_as_uint(rhi.w); continue; }
            pop = true;
        } else if (in_mesh) {
            uint32_t id = li & ~PHR_LEAF;
            if (id < ri) {
                float4 t0 = s.tris[3 * id], t1 = s.tris[3 * id + 1], t2 = s.tris[3 * id + 2];
                for (;;) {
                    ++tests;
                    // The leaf's next triangle is loaded while this one is tested.
                    const bool next = id + 1u < ri;
                    float4 n0 = t0, n1 = t1, n2 = t2;
                    if (next) {
                        n0 = s.tris[3 * (id + 1u)];
                        n1 = s.tris[3 * (id + 1u) + 1];
                        n2 = s.tris[3 * (id + 1u) + 2];
                    }
                    float u, v;
                    const float t = tri_t(t0, t1, t2, ro, rs, u, v);
                    if (t < best.t) {
                        if (ANY) {
                            const float* m = s.mats + MAT_WORDS * (size_t)s.faces[id].w;
                            if (!(m[6] > 0.5f || m[3] < 0.5f)) return PhrHit{t, id, u, v, inst};
                        } else {
                            best = PhrHit{t, id, u, v, inst};
                        }
                    }
                    if (!next) break;
                    ++id;
                    t0 = n0;
                    t1 = n1;
                    t2 = n2;
                }
            }
            pop = true;
        } else {  // a top-tree leaf: its instances, in slot order
            k = li & ~(PHR_LEAF | PHR_INST);
            kend = ri;
        }
        if (pop && in_mesh && bsp > 0) {
            const uint32_t x = bstk[--bsp * stride];
            li = __float_as_uint(s.nodes[2 * x].w);
            ri = __float_as_uint(s.nodes[2 * x + 1].w);
            continue;
        }
        if (pop && in_mesh) {  // this instance's tree is done
            in_mesh = false;
            ++k;
        }
        // Into the next instance of the leaf whose root box the ray meets
        // (trace_blas's prologue, bvh.wgsl:120-150).
        while (!in_mesh && k < kend) {
            const uint32_t i = inst_order[k];
            const uint32_t* w = s.insts + INSTANCE_LEN * (size_t)i;
            const uint4 info = *(const uint4*)(w + 24);
            const uint32_t root = info.x, flags = info.w;
            if (root != PHR_NONE && !(ANY && (flags & 2u))) {
                V3 mo = o, md = d;
                if ((flags & 1u) == 0u) {
                    const float4 r0 = *(const float4*)(w + 12), r1 = *(const float4*)(w + 16), r2 = *(const float4*)(w + 20);
                    mo = v3(dot4(r0, o.x, o.y, o.z, 1.0f), dot4(r1, o.x, o.y, o.z, 1.0f), dot4(r2, o.x, o.y, o.z, 1.0f));
                    md = v3((r0.x * d.x + r0.y * d.y) + r0.z * d.z, (r1.x * d.x + r1.y * d.y) + r1.z * d.z,
                            (r2.x * d.x + r2.y * d.y) + r2.z * d.z);
                }
                const V3 minv = safe_inv(md);
                const float4 rlo = s.nodes[2 * root], rhi = s.nodes[2 * root + 1];
                if (box_t(rlo, rhi, mo, minv, best.t) < PHR_INF) {
                    in_mesh = true;
                    inst = i;
                    ro = mo;
                    rinv = minv;
                    rs = ray_space(md);
                    bsp = 0;
                    li = __float_as_uint(rlo.w);
                    ri = __float_as_uint(rhi.w);
                    break;
                }
            }
            ++k;
        }
        if (in_mesh) continue;
        // The leaf's instances are done, or an inner node met nothing: the
        // top tree pops.
        ro = o;
        rinv = winv;
        if (tsp == 0) break;
        const uint32_t x = tstk[--tsp * stride];
        li = __float_as_uint(s.nodes[2 * x].w);
        ri = __float_as_uint(s.nodes[2 * x + 1].w);
    }
    return best;
}

// walk_forest as a walk that can stop after any unit and go on later (the
// split build over a forest): a unit is one pass of walk_forest's loop (a
// node of either tree, or a top-tree leaf's instances entered in turn), with
// a mesh leaf's triangles one a unit, so the visits, decisions and pushes
// are walk_forest's in its order. `any`: a shadow ray's walk.
struct ForestWalker {
    V3 o, d, winv;  // the world ray
    V3 ro, rinv;    // the ray the current level walks with
    RaySpace rs;
    float t, u, v;  // closest: the nearest hit so far; shadow: t is the ray's length
    uint32_t tri, hinst;  // the hit's face and instance
    uint32_t li, ri, tsp, bsp, inst, k, kend;
    bool on, any, in_mesh;
};
__device__ __forceinline__ void forest_begin(ForestWalker& w, const SceneDev& s, V3 o, V3 d, float tmax, bool any) {
    w.o = o;
    w.d = d;
    w.winv = safe_inv(d);
    w.ro = o;
    w.rinv = w.winv;
    w.rs = RaySpace{0u, false, 0.f, 0.f, 0.f};
    w.t = tmax;
    w.u = w.v = 0.f;
    w.tri = PHR_NONE;
    w.hinst = 0u;
    w.tsp = w.bsp = w.inst = w.k = w.kend = 0u;
    w.any = any;
    w.in_mesh = false;
    const float4 lo = s.nodes[0], hi = s.nodes[1];
    w.on = box_t(lo, hi, o, w.winv, tmax) < PHR_INF;
    w.li = __float_as_uint(lo.w);
    w.ri = __float_as_uint(hi.w);
}
__device__ __forceinline__ void forest_step(ForestWalker& w, const SceneDev& s, uint16_t* tstk, uint32_t* bstk, uint32_t stride,
                                            uint32_t& steps, uint32_t& tests) {
    bool pop = false;
    if (!(w.li & PHR_LEAF)) {
        ++steps;
        const float4 llo = s.nodes[2 * w.li], lhi = s.nodes[2 * w.li + 1];
        const float4 rlo = s.nodes[2 * w.ri], rhi = s.nodes[2 * w.ri + 1];
        const float tl = box_t(llo, lhi, w.ro, w.rinv, w.t);
        const float tr = box_t(rlo, rhi, w.ro, w.rinv, w.t);
        if (tl < PHR_INF && tr < PHR_INF) {
            const uint32_t far = tl <= tr ? w.ri : w.li;
            if (w.in_mesh) bstk[w.bsp++ * stride] = far;
            else tstk[w.tsp++ * stride] = (uint16_t)far;
            const float4 nl = tl <= tr ? llo : rlo, nh = tl <= tr ? lhi : rhi;
            w.li = __float_as_uint(nl.w);
            w.ri = __float_as_uint(nh.w);
            return;
        }
        if (tl < PHR_INF) { w.li = __float_as_uint(llo.w); w.ri = __float_as_uint(lhi.w); return; }
        if (tr < PHR_INF) { w.li = __float_as_uint(rlo.w); w.ri = __float_as_uint(rhi.w); return; }
        pop = true;
    } else if (w.in_mesh) {  // one triangle of the mesh leaf [li, ri)
        const uint32_t id = w.li & ~PHR_LEAF;
        if (id < w.ri) {
            ++tests;
            const float4 t0 = s.tris[3 * id], t1 = s.tris[3 * id + 1], t2 = s.tris[3 * id + 2];
            float u, v;
            const float t = tri_t(t0, t1, t2, w.ro, w.rs, u, v);
            if (t < w.t) {
                if (w.any) {
                    const float* m = s.mats + MAT_WORDS * (size_t)s.faces[id].w;
                    if (!(m[6] > 0.5f || m[3] < 0.5f)) {
                        w.tri = id;
                        w.hinst = w.inst;
                        w.on = false;
                        return;
                    }
                } else {
                    w.t = t;
                    w.tri = id;
                    w.u = u;
                    w.v = v;
                    w.hinst = w.inst;
                }
            }
            if (id + 1u < w.ri) {
                w.li = (id + 1u) | PHR_LEAF;
                return;
            }
        }
        pop = true;
    } else {  // a top-tree leaf: its instances, in slot order
        w.k = w.li & ~(PHR_LEAF | PHR_INST);
        w.kend = w.ri;
    }
    if (pop && w.in_mesh && w.bsp > 0) {
        const uint32_t x = bstk[--w.bsp * stride];
        w.li = __float_as_uint(s.nodes[2 * x].w);
        w.ri = __float_as_uint(s.nodes[2 * x + 1].w);
        return;
    }
    if (pop && w.in_mesh) {
        w.in_mesh = false;
        ++w.k;
    }
    const uint32_t* inst_order = s.insts + INSTANCE_LEN * (size_t)s.ninsts;
    while (!w.in_mesh && w.k < w.kend) {
        const uint32_t i = inst_order[w.k];
        const uint32_t* iw = s.insts + INSTANCE_LEN * (size_t)i;
        const uint4 info = *(const uint4*)(iw + 24);
        const uint32_t root = info.x, flags = info.w;
        if (root != PHR_NONE && !(w.any && (flags & 2u))) {
            V3 mo = w.o, md = w.d;
            if ((flags & 1u) == 0u) {
                const float4 r0 = *(const float4*)(iw + 12), r1 = *(const float4*)(iw + 16), r2 = *(const float4*)(iw + 20);
                mo = v3(dot4(r0, w.o.x, w.o.y, w.o.z, 1.0f), dot4(r1, w.o.x, w.o.y, w.o.z, 1.0f), dot4(r2, w.o.x, w.o.y, w.o.z, 1.0f));
                md = v3((r0.x * w.d.x + r0.y * w.d.y) + r0.z * w.d.z, (r1.x * w.d.x + r1.y * w.d.y) + r1.z * w.d.z,
                        (r2.x * w.d.x + r2.y * w.d.y) + r2.z * w.d.z);
            }
            const V3 minv = safe_inv(md);
            const float4 rlo = s.nodes[2 * root], rhi = s.nodes[2 * root + 1];
            if (box_t(rlo, rhi, mo, minv, w.t) < PHR_INF) {
                w.in_mesh = true;
                w.inst = i;
                w.ro = mo;
                w.rinv = minv;
                w.rs = ray_space(md);
                w.bsp = 0;
                w.li = __float_as_uint(rlo.w);
                w.ri = __float_as_uint(rhi.w);
                break;
            }
        }
        ++w.k;
    }
    if (w.in_mesh) return;
    w.ro = w.o;
    w.rinv = w.winv;
    if (w.tsp == 0) {
        w.on = false;
        return;
    }
    const uint32_t x = tstk[--w.tsp * stride];
    w.li = __float_as_uint(s.nodes[2 * x].w);
    w.ri = __float_as_uint(s.nodes[2 * x + 1].w);
}

constexpr uint32_t TRACE_THREADS = 256;
// LDS stacks a thread: PHR_STACK words for one tree or a mesh tree, and for a
// forest PHR_STACK 16-bit entries for the top tree.
template <bool F>
constexpr uint32_t top_stack_entries() { return F ? PHR_STACK : 1u; }
// The forest's 4-waves-a-SIMD build keeps 24 mesh-tree entries (40 KB of LDS
// a block, four blocks a CU): the tracer runs it only when every mesh tree
// is at most 24 deep, a walk holding at most one pending node a level.
constexpr uint32_t FOREST_SHORT_STACK = 24;
template <bool F, int WPE>
constexpr uint32_t mesh_stack_entries() { return (F && WPE >= 4) ? FOREST_SHORT_STACK : PHR_STACK; }

template <bool F>
__global__ __launch_bounds__(TRACE_THREADS) void k_raycast(SceneDev sc, const PhrRay* __restrict__ rays, uint32_t nrays, uint32_t any,
                                                           PhrHit* __restrict__ hits, unsigned long long* __restrict__ counters) {
    __shared__ uint32_t s_stack[PHR_STACK * TRACE_THREADS];
    __shared__ uint16_t s_tstack[top_stack_entries<F>() * TRACE_THREADS];
    const uint32_t i = blockIdx.x * TRACE_THREADS + threadIdx.x;
    uint32_t steps = 0, tests = 0;
    if (i < nrays) {
        const float4* r = (const float4*)(rays + i);
        float4 a = r[0], b = r[1];
        V3 o = v3(a.x, a.y, a.z), d = v3(b.x, b.y, b.z);
        uint32_t* stk = s_stack + threadIdx.x;
        PhrHit h;
        if constexpr (F)
            h = any ? walk_forest<true>(sc, o, d, a.w, s_tstack + threadIdx.x, stk, TRACE_THREADS, steps, tests)
                    : walk_forest<false>(sc, o, d, a.w, s_tstack + threadIdx.x, stk, TRACE_THREADS, steps, tests);
        else
            h = any ? walk<true, true>(sc.nodes, sc.recs, sc.kids, sc.root, sc.tris, sc.faces, sc.mats, o, d, a.w, stk, TRACE_THREADS, steps, tests)
                    : walk<false, true>(sc.nodes, sc.recs, sc.kids, sc.root, sc.tris, sc.faces, sc.mats, o, d, a.w, stk, TRACE_THREADS, steps, tests);
        hits[i] = h;
    }
    if (counters) {
        unsigned long long cs = steps, ct = tests;
        for (int k = 32; k > 0; k >>= 1) {
            cs += __shfl_xor(cs, k);
            ct += __shfl_xor(ct, k);
        }
        if ((threadIdx.x & 63) == 0) {
            atomicAdd(&counters[0], cs);
            atomicAdd(&counters[1], ct);
        }
    }
}

// ---- scene gather: boxes, then slot-ordered faces and triangle cache ------
__global__ __launch_bounds__(256) void k_tri_boxes(const uint4* __restrict__ in_faces, const float4* __restrict__ verts, uint32_t n,
                                                   float4* __restrict__ prims) {
    uint32_t f = blockIdx.x * 256 + threadIdx.x;
    if (f >= n) return;
    uint4 fc = in_faces[f];
    float4 a = verts[2 * fc.x], b = verts[2 * fc.y], c = verts[2 * fc.z];
    // Aabb::of_tri, src/bvh/node.rs:78-80
    prims[2 * f] = make_float4(fminf(fminf(a.x, b.x), c.x), fminf(fminf(a.y, b.y), c.y), fminf(fminf(a.z, b.z), c.z), 0.f);
    prims[2 * f + 1] = make_float4(fmaxf(fmaxf(a.x, b.x), c.x), fmaxf(fmaxf(a.y, b.y), c.y), fmaxf(fmaxf(a.z, b.z), c.z), 0.f);
}

__global__ __launch_bounds__(256) void k_gather(const uint4* __restrict__ in_faces, const float4* __restrict__ verts,
                                                const uint32_t* __restrict__ order, uint32_t n, uint4* __restrict__ faces,
                                                float4* __restrict__ tris) {
    // Scene::reorder, src/trace/scene.rs:188-190
    uint32_t s = blockIdx.x * 256 + threadIdx.x;
    if (s >= n) return;
    uint4 fc = in_faces[order[s]];
    faces[s] = fc;
    float4 a = verts[2 * fc.x], b = verts[2 * fc.y], c = verts[2 * fc.z];
    tris[3 * s] = make_float4(a.x, a.y, a.z, b.x);
    tris[3 * s + 1] = make_float4(b.y, b.z, c.x, c.y);
    tris[3 * s + 2] = make_float4(c.z, 0.f, 0.f, 0.f);
}

// ---- integrator --------------------------------------------------------------
struct Tri {
    V3 a, b, c, na, nb, nc;
    float color[4];
    float rough, metal, trans, ior;
    V3 emit;
};
struct Surf { V3 pos, n, v, albedo, f0; float rough, eta, dielectric; };
struct Shadowed { V3 c, dir; float dist; };

struct Ctx {
    SceneDev s;
    const PhrCamera* p;
    uint32_t* stk;
    uint16_t* tstk;  // a forest's top-tree stack
    uint32_t closest, shadow, steps, tests;
    // The pixel's guide sums (albedo, normal, depth, missed, coverage) and the
    // last sample's outline word: a sample adds its guide at its first
    // camera segment, where the guide is whole.
    V3 ga, gn;
    float gd, gone, cov;
    uint32_t outline;
};

// F: the scene is a forest (top tree + instances' trees); else one tree under
// one identity instance.
template <bool F>
__device__ __forceinline__ PhrHit trace(Ctx& c, V3 o, V3 d, float tmax, bool first) {
    if (first) {
        c.shadow++;
        if constexpr (F) return walk_forest<true>(c.s, o, d, tmax, c.tstk, c.stk, TRACE_THREADS, c.steps, c.tests);
        return walk<true>(c.s.nodes, c.s.recs, c.s.kids, c.s.root, c.s.tris, c.s.faces, c.s.mats, o, d, tmax, c.stk, TRACE_THREADS, c.steps,
                          c.tests);
    }
    c.closest++;
    if constexpr (F) return walk_forest<false>(c.s, o, d, tmax, c.tstk, c.stk, TRACE_THREADS, c.steps, c.tests);
    return walk<false>(c.s.nodes, c.s.recs, c.s.kids, c.s.root, c.s.tris, c.s.faces, c.s.mats, o, d, tmax, c.stk, TRACE_THREADS, c.steps,
                       c.tests);
}
template <bool F>
__device__ __forceinline__ float visible(Ctx& c, V3 org, V3 l, float dist) {
    return trace<F>(c, org, l, dist, true).tri != PHR_NONE ? 0.0f : 1.0f;
}

__device__ __forceinline__ V3 unoct(float ex, float ey) {  // layout.wgsl:79-86
    if (ex > 1.5f) return v3(0, 0, 0);
    V3 n = v3(ex, ey, 1.0f - fabsf(ex) - fabsf(ey));
    if (n.z < 0.f) n = v3((1.0f - fabsf(ey)) * wgsl_sign(ex), (1.0f - fabsf(ex)) * wgsl_sign(ey), n.z);
    return normalize(n);
}

__device__ __forceinline__ uint32_t mat_at(const SceneDev& s, uint32_t f) { return s.faces[f].w; }

__device__ __forceinline__ V3 place_n(float4 i0, float4 i1, float4 i2, V3 n) {  // layout.wgsl:70-77
    V3 m = (v3(i0.x, i0.y, i0.z) * n.x + v3(i1.x, i1.y, i1.z) * n.y) + v3(i2.x, i2.y, i2.z) * n.z;
    float l = length(m);
    if (l < 1e-12f) return v3(0, 0, 0);
    return m / l;
}

// Face f under instance i with material m (layout.wgsl:114-158): in a forest
// its corners and normals are placed in the world unless the instance sits
// at the identity.
template <bool F>
__device__ Tri tri_of(const SceneDev& s, uint32_t f, uint32_t i, uint32_t m) {
    uint4 fc = s.faces[f];
    float4 xp = s.verts[2 * fc.x], xa = s.verts[2 * fc.x + 1];
    float4 yp = s.verts[2 * fc.y], ya = s.verts[2 * fc.y + 1];
    float4 zp = s.verts[2 * fc.z], za = s.verts[2 * fc.z + 1];
    const float4* mt = (const float4*)(s.mats + MAT_WORDS * (size_t)m);
    float4 m0 = mt[0], m1 = mt[1], m2 = mt[2];
    Tri t;
    t.a = v3(xp.x, xp.y, xp.z); t.b = v3(yp.x, yp.y, yp.z); t.c = v3(zp.x, zp.y, zp.z);
    t.na = unoct(xa.x, xa.y); t.nb = unoct(ya.x, ya.y); t.nc = unoct(za.x, za.y);
    if constexpr (F) {
        const uint32_t* w = s.insts + INSTANCE_LEN * (size_t)i;
        if ((w[27] & 1u) == 0u) {
            const float4* r = (const float4*)w;
            const float4 r0 = r[0], r1 = r[1], r2 = r[2], i0 = r[3], i1 = r[4], i2 = r[5];
            t.a = v3(dot4(r0, xp.x, xp.y, xp.z, xp.w), dot4(r1, xp.x, xp.y, xp.z, xp.w), dot4(r2, xp.x, xp.y, xp.z, xp.w));
            t.b = v3(dot4(r0, yp.x, yp.y, yp.z, yp.w), dot4(r1, yp.x, yp.y, yp.z, yp.w), dot4(r2, yp.x, yp.y, yp.z, yp.w));
            t.c = v3(dot4(r0, zp.x, zp.y, zp.z, zp.w), dot4(r1, zp.x, zp.y, zp.z, zp.w), dot4(r2, zp.x, zp.y, zp.z, zp.w));
            t.na = place_n(i0, i1, i2, t.na);
            t.nb = place_n(i0, i1, i2, t.nb);
            t.nc = place_n(i0, i1, i2, t.nc);
        }
    }
    t.color[0] = m0.x; t.color[1] = m0.y; t.color[2] = m0.z; t.color[3] = m0.w;
    t.rough = m1.x; t.metal = m1.y; t.trans = m1.z; t.ior = m1.w;
    t.emit = v3(m2.x, m2.y, m2.z);
    return t;
}

__device__ __forceinline__ float ggx_lambda(float alpha2, float cos_n) {  // brdf.wgsl:9-12
    float t = alpha2 * fmax_(1.0f / fmax_(cos_n * cos_n, 1e-8f) - 1.0f, 0.0f);
    return 0.5f * (sqrtf(1.0f + t) - 1.0f);
}

__device__ __forceinline__ float fresnel_dielectric_cos(float cosi, float eta) {  // procedural.wgsl:73-81
    float c = fabsf(cosi);
    float g = eta * eta - 1.0f + c * c;
    if (g <= 0.f) return 1.0f;
    g = sqrtf(g);
    float a = (g - c) / (g + c);
    float b = (c * (g + c) - 1.0f) / (c * (g - c) + 1.0f);
    return 0.5f * a * a * (1.0f + b * b);
}

__device__ __forceinline__ V3 fresnel_of(const Surf& s, float cs) {  // brdf.wgsl:73-82
    float schlick = pw5(1.0f - cs);
    V3 f = s.f0 + (v3(1, 1, 1) - s.f0) * schlick;
    if (s.eta > 1.0f && s.dielectric > 0.f) {
        float r = (s.eta - 1.0f) / (s.eta + 1.0f);
        float face_on = r * r;
        float add = s.dielectric * (fresnel_dielectric_cos(cs, s.eta) - face_on - (1.0f - face_on) * schlick);
        f = f + v3(add, add, add);
    }
    return vmax(f, 0.f);
}

__device__ __forceinline__ void env_ab(float rough, float nv, float& abx, float& aby) {  // brdf.wgsl:109-116
    float r0 = rough * -1.0f + 1.0f;
    float r1 = rough * -0.0275f + 0.0425f;
    float r2 = rough * -0.572f + 1.04f;
    float r3 = rough * 0.022f + -0.04f;
    float a004 = fmin_(r0 * r0, dexp2(-9.28f * nv)) * r0 + r1;
    abx = -1.04f * a004 + r2;
    aby = 1.04f * a004 + r3;
}

__device__ __forceinline__ V3 dielectric_f0(float ior) {
    float r = (ior - 1.0f) / (ior + 1.0f);
    return v3(r * r, r * r, r * r);
}

__device__ __forceinline__ float under_layer(V3 f0, float rough, float nv) {  // brdf.wgsl:135-138
    float ax, ay;
    env_ab(rough, nv, ax, ay);
    V3 e = f0 * ax + v3(ay, ay, ay);
    return clampf(1.0f - fmax_(e.x, fmax_(e.y, e.z)), 0.0f, 1.0f);
}

__device__ __forceinline__ float under_layer_real(float eta, float rough, float nv) {  // brdf.wgsl:146-155
    float r = (eta - 1.0f) / (eta + 1.0f);
    float face_on = r * r;
    float ax, ay;
    env_ab(rough, nv, ax, ay);
    float climb = 0.0f * ax + ay;
    float flat = (1.0f * ax + ay) - climb;
    float schlick = (1.0f - face_on) * pw5(1.0f - nv);
    float real = fresnel_dielectric_cos(nv, eta) - face_on;
    float scale = clampf(real / fmax_(schlick, 1e-5f), 0.0f, 1.0f);
    return clampf(1.0f - (face_on * flat + climb * scale), 0.0f, 1.0f);
}

__device__ __forceinline__ float glossy_floor(const PhrCamera* p, float min_pdf) {  // brdf.wgsl:202-208
    if (p->integrator[0] < 0.5f) return 0.3f;
    if (p->integrator[2] <= 0.0f) return 0.0f;
    float blur_pdf = min_pdf / p->integrator[2];
    if (blur_pdf >= 1.0f) return 0.0f;
    return sqrtf(sqrtf(1.0f - blur_pdf) * 0.5f);
}

__device__ V3 brdf_w(const Surf& s, V3 l, V3 e, float wd, float ws) {  // brdf.wgsl:239-269 (no sheen/coat/bleed)
    float raw = dot(s.n, l);
    float nl = fmax_(raw, 0.0f);
    if (nl <= 0.f) return v3(0, 0, 0);
    V3 hv = normalize(l + s.v);
    float nh = fmax_(dot(s.n, hv), 0.0f);
    float hvv = fmax_(dot(hv, s.v), 0.0f);
    float a2 = fmax_(s.rough * s.rough * s.rough * s.rough, 1e-4f);
    float dd = nh * nh * (a2 - 1.0f) + 1.0f;
    float dist = a2 / (PHR_PI * dd * dd);
    V3 fres = fresnel_of(s, hvv);
    float nv = fmax_(dot(s.n, s.v), 1e-4f);
    float g2 = 1.0f / (1.0f + ggx_lambda(a2, nv) + ggx_lambda(a2, fmax_(raw, 1e-4f)));
    V3 spec = fres * dist * g2 * 0.25f / nv;
    V3 diffuse = s.albedo / PHR_PI;
    return (diffuse * wd + spec * ws) * e * nl;
}

__device__ __forceinline__ V3 onb(V3 n, float a, float b) {  // light.wgsl:1-5
    V3 tx = normalize(fabsf(n.z) > 0.9f ? cross(n, v3(1, 0, 0)) : cross(n, v3(0, 0, 1)));
    V3 ty = cross(n, tx);
    return tx * a + ty * b;
}

__device__ __forceinline__ V3 cone_dir(V3 axis, float ang, uint32_t* seed) {  // light.wgsl:34-39
    if (ang <= 0.f) return axis;
    float r = ang * sqrtf(rnd(seed));
    float ph = 2.0f * PHR_PI * rnd(seed);
    return normalize(axis + onb(axis, r * dcos(ph), r * dsin(ph)));
}

__device__ __forceinline__ V3 ld3(const float* p) { return v3(p[0], p[1], p[2]); }
__device__ __forceinline__ V3 env_rgb(const PhrCamera* p) { return v3(p->env[0], p->env[1], p->env[2]); }
__device__ __forceinline__ bool world_lit(const PhrCamera* p) { V3 e = env_rgb(p); return dot(e, e) > 0.f; }

__device__ float lamp_chance(const SceneDev& s, const PhrHit& h) {  // light.wgsl:264-278
    const uint32_t* d = s.lights;
    uint32_t count = d[0];
    if (count == 0u) return 0.0f;
    uint32_t lo = 0, hi = count - 1;
    while (lo < hi) {
        uint32_t mid = (lo + hi) / 2u;
        uint32_t f = d[1 + mid], i = d[1 + count + mid];
        if (f < h.tri || (f == h.tri && i < h.inst)) lo = mid + 1; else hi = mid;
    }
    if (d[1 + lo] != h.tri || d[1 + count + lo] != h.inst) return 0.0f;
    float prev = lo > 0 ? __uint_as_float(d[2 * count + lo]) : 0.0f;
    return __uint_as_float(d[1 + 2 * count + lo]) - prev;
}

template <bool F>
__device__ Shadowed glow_query(const SceneDev& sc, const Surf& s, uint32_t* seed) {  // light.wgsl:281-312
    const uint32_t* d = sc.lights;
    uint32_t count = d[0];
    if (count == 0u) return Shadowed{v3(0, 0, 0), v3(0, 0, 1), 0.0f};
    float u = rnd(seed);
    uint32_t lo = 0, hi = count - 1;
    while (lo < hi) {
        uint32_t mid = (lo + hi) / 2u;
        if (__uint_as_float(d[1 + 2 * count + mid]) < u) lo = mid + 1; else hi = mid;
    }
    float prev = lo > 0 ? __uint_as_float(d[2 * count + lo]) : 0.0f;
    float chance = fmax_(__uint_as_float(d[1 + 2 * count + lo]) - prev, 1e-8f);
    uint32_t f = d[1 + lo];
    Tri t = tri_of<F>(sc, f, d[1 + count + lo], mat_at(sc, f));
    float a = rnd(seed);
    float b = rnd(seed);
    if (a + b > 1.0f) { a = 1.0f - a; b = 1.0f - b; }
    V3 q = t.a + (t.b - t.a) * a + (t.c - t.a) * b;
    V3 cr = cross(t.b - t.a, t.c - t.a);
    float area = 0.5f * length(cr);
    V3 to = q - s.pos;
    float d2 = fmax_(dot(to, to), 1e-6f);
    V3 l = to / sqrtf(d2);
    float cos_l = fabsf(dot(normalize(cr), l));
    V3 e = t.emit * cos_l * area / (chance * d2);
    float pl = chance * d2 / fmax_(area * cos_l, 1e-8f);
    float pb = fmax_(dot(s.n, l), 0.0f) / PHR_PI;
    float w = pl * pl / fmax_(pl * pl + pb * pb, 1e-12f);
    return Shadowed{brdf_w(s, l, e, 1.0f, 1.0f) * w, l, sqrtf(d2) * 0.999f};
}

template <bool F>
__device__ V3 direct(Ctx& c, const Surf& s, uint32_t* seed) {  // integrate.wgsl:66-85
    const PhrCamera* p = c.p;
    bool rig = p->count == 0u && !world_lit(p);
    V3 sum = v3(0, 0, 0);
    if (rig) {  // light.wgsl:103-106
        V3 fill = normalize(-ld3(p->right) * 0.8f + ld3(p->up) * 0.2f + ld3(p->back) * 0.6f);
        sum = brdf_w(s, fill, v3(0.45f, 0.45f, 0.45f), 1.0f, 1.0f);
    }
    for (uint32_t k = 0; k < 3u; ++k) {
        Shadowed q{v3(0, 0, 0), v3(0, 0, 1), 0.0f};
        if (k == 0u) {
            if (rig) {  // light.wgsl:94-97
                V3 key = cone_dir(ld3(p->sun), 0.06f, seed);
                q = Shadowed{brdf_w(s, key, v3(2, 2, 2), 1.0f, 1.0f), key, PHR_INF};
            } else {
                (void)rnd(seed);  // lamp_query's pick over unlit lamps (light.wgsl:126)
            }
        } else if (k == 1u) {
            if (rig) {  // light.wgsl:99-100
                V3 rim = normalize(ld3(p->up) * 0.6f - ld3(p->back) * 0.8f + ld3(p->right) * 0.2f);
                q = Shadowed{brdf_w(s, rim, v3(0.9f, 0.9f, 0.9f), 1.0f, 1.0f), rim, PHR_INF};
            }
        } else {
            q = glow_query<F>(c.s, s, seed);
        }
        if (maxc(q.c) > 0.f) sum = sum + q.c * visible<F>(c, s.pos, q.dir, q.dist);
    }
    return sum;
}

__device__ __forceinline__ V3 cosine_dir(V3 n, uint32_t* seed) {  // integrate.wgsl:6-11
    float r1 = rnd(seed);
    float rr = sqrtf(r1);
    float ph = 2.0f * PHR_PI * rnd(seed);
    return normalize(onb(n, rr * dcos(ph), rr * dsin(ph)) + n * sqrtf(fmax_(0.0f, 1.0f - r1)));
}

__device__ V3 glow_mis(const SceneDev& s, const Tri& t, const PhrHit& h, V3 from_n, V3 dir) {  // integrate.wgsl:15-23
    V3 e = t.emit;
    float chance = lamp_chance(s, h);
    if (chance <= 0.f) return e;
    V3 cr = cross(t.b - t.a, t.c - t.a);
    float pl = chance * h.t * h.t / fmax_(0.5f * length(cr) * fabsf(dot(normalize(cr), dir)), 1e-8f);
    float pb = fmax_(dot(from_n, dir), 0.0f) / PHR_PI;
    return e * pb * pb / fmax_(pb * pb + pl * pl, 1e-12f);
}

template <bool F>
__device__ V3 indirect(Ctx& c, V3 pos, V3 n, uint32_t* seed, bool bounce) {  // integrate.wgsl:111-230
    const PhrCamera* p = c.p;
    bool lit = world_lit(p);
    if (p->glow == 0u && !lit) {
        V3 dir = cosine_dir(n, seed);
        return env_rgb(p) * mixf(0.25f, 1.0f, visible<F>(c, pos, dir, p->ao));
    }
    uint32_t depth = bounce ? (p->bounces != 0u ? p->bounces : 1u) : 1u;
    V3 acc = v3(0, 0, 0), thr = v3(1, 1, 1), org = pos, nrm = n;
    float min_pdf = 3.0e38f;
    for (uint32_t b = 0; b < depth; ++b) {
        V3 dir = cosine_dir(nrm, seed);
        float pb = fmax_(dot(nrm, dir), 0.0f) / PHR_PI;
        min_pdf = fmin_(min_pdf, pb);
        PhrHit h = trace<F>(c, org, dir, PHR_INF, false);
        if (h.tri == PHR_NONE) {
            acc = acc + thr * env_rgb(p) * 1.0f;
            break;
        }
        if (!bounce) break;
        Tri t = tri_of<F>(c.s, h.tri, h.inst, mat_at(c.s, h.tri));
        V3 ng0 = normalize(cross(t.b - t.a, t.c - t.a));
        V3 ng = dot(ng0, dir) < 0.f ? ng0 : -ng0;
        V3 hit = org + dir * h.t + ng * (p->eps + h.t * 1e-5f);
        V3 base = vmax(v3(t.color[0], t.color[1], t.color[2]), 0.0f);
        float rough = fmax_(t.rough, glossy_floor(p, min_pdf));
        bool glossy = p->integrator[0] < 0.5f || (((uint32_t)(p->integrator[3] + 0.5f)) & 1u) != 0u;
        V3 fd = glossy ? dielectric_f0(t.ior) : v3(0, 0, 0);
        float metal = glossy ? t.metal : 0.0f;
        float tw = t.trans * (1.0f - t.metal);
        V3 albedo = base * (1.0f - t.metal) * (1.0f - tw);
        if (maxc(albedo) > 0.f) albedo = albedo * under_layer(fd, rough, fmax_(dot(ng, -dir), 1e-4f));
        acc = acc + thr * glow_mis(c.s, t, h, nrm, dir);
        float near = (!lit && h.t <= p->ao) ? 0.25f : 1.0f;
        Surf s{hit, ng, -dir, albedo, mixv(fd, base, metal), rough, 0.0f, 0.0f};
        acc = acc + thr * direct<F>(c, s, seed) * near;
        thr = thr * albedo;
        if (b + 1u < depth && b + 1u >= p->rr_from) {
            float keep = clampf(maxc(thr), 0.0f, 0.95f);
            if (keep < 1e-4f || rnd(seed) > keep) break;
            thr = thr / keep;
        }
        org = hit;
        nrm = ng;
    }
    return acc;
}

// One camera sample of pixel i (shade.wgsl:41-373). Returns what `emit` adds
// to the pixel's radiance; what `guide` adds goes into c's guide sums at the
// first camera segment, in sample order as at the sample's end.
struct Sample {
    V3 lin;
};

template <bool F>
__device__ Sample shade(Ctx& c, uint32_t i, uint32_t frame) {
    const PhrCamera* p = c.p;
    Sample out;
    float fw = (float)p->w, fh = (float)p->h;
    float su = ((float)(i % p->w) + 0.5f) / fw - 0.5f;
    float sv = ((float)(i / p->w) + 0.5f) / fh - 0.5f;
    uint32_t seed = pcg_hash(i * 9781u + (frame + p->stream) * 6271u + 1u);
    float ju = rnd(&seed);
    float jv = rnd(&seed);
    float jx = (ju - 0.5f) / fw;
    float jy = (jv - 0.5f) / fh;
    float vx = (su + jx - p->pan[0]) / p->scale;
    float vy = -((sv + jy) / p->aspect - p->pan[1]) / p->scale;
    V3 o, d;
    if (p->persp != 0u) {
        o = ld3(p->pivot) + ld3(p->back) * (p->dist / p->units);
        d = normalize(ld3(p->right) * vx + ld3(p->up) * vy - ld3(p->back) * p->dist);
    } else {
        o = ld3(p->pivot) + (ld3(p->right) * vx + ld3(p->up) * vy + ld3(p->back) * 1.0e4f) / p->units;
        d = -ld3(p->back);
    }
    V3 col = v3(0, 0, 0), thr = v3(1, 1, 1);
    bool depth_set = p->depth[0] != 0u;
    uint32_t total = depth_set ? p->depth[0] - 1u : 8u;
    uint32_t see_through = depth_set ? p->depth[1] - 1u : 8u;
    uint32_t through = depth_set ? min(p->depth[2] - 1u, total) : 8u;
    bool diffuse_on = !depth_set || (p->bounces > 0u && total > 0u);
    uint32_t n_see = 0u;
    uint32_t passes = see_through + through + 1u;
    for (uint32_t bounce = 0; bounce < passes; ++bounce) {
        PhrHit h = trace<F>(c, o, d, PHR_INF, false);
        if (h.tri == PHR_NONE) {
            if (bounce == 0u) {
                c.ga = c.ga + v3(0, 0, 0);
                c.gn = c.gn + (-d);
                c.gd = c.gd + -1.0f;
                c.gone = c.gone + 1.0f;
                c.cov = c.cov + 1.0f;
                c.outline = 0u;
                out.lin = v3(0, 0, 0);
                return out;
            }
            col = col + thr * env_rgb(p);
            break;
        }
        uint32_t m = mat_at(c.s, h.tri);
        Tri t = tri_of<F>(c.s, h.tri, h.inst, m);
        float w = 1.0f - h.u - h.v;
        V3 hit = o + d * h.t;
        // world_ng (layout.wgsl:96-100): turned back under a mirroring instance.
        V3 ng0 = normalize(cross(t.b - t.a, t.c - t.a));
        if constexpr (F)
            if (c.s.insts[INSTANCE_LEN * (size_t)h.inst + 27] & 4u) ng0 = -ng0;
        bool front = dot(ng0, d) < 0.f;
        V3 ng = front ? ng0 : -ng0;
        V3 sn = t.na * w + t.nb * h.u + t.nc * h.v;
        sn = dot(sn, sn) > 1e-12f ? normalize(sn) * (front ? 1.0f : -1.0f) : ng;
        if (dot(sn, -d) < 0.f) sn = ng;
        float start = 1.0e4f / p->units;
        float lift = (p->persp == 0u && h.t > start * 0.5f) ? p->eps + fabsf(h.t - start) * 1e-5f + start * 2.4e-7f
                                                            : p->eps + h.t * 1e-5f;
        V3 pos = hit + ng * lift;
        V3 base = vmax(v3(t.color[0], t.color[1], t.color[2]), 0.0f);
        float alpha = t.color[3];
        if (bounce == 0u) {
            c.ga = c.ga + base;
            c.gn = c.gn + sn;
            c.gd = c.gd + h.t;
            c.gone = c.gone + 0.f;
            c.cov = c.cov + 0.f;
            c.outline = h.inst + 1u;
        }
        if (rnd(&seed) >= alpha) {
            n_see++;
            if (n_see > see_through) break;
            float reach = fmax_(fmax_(fabsf(hit.x), fabsf(hit.y)), fmax_(fabsf(hit.z), 0.5f));
            float tl = lift - p->eps + reach * (256.0f / 8388608.0f);
            o = hit - ng * tl;
            continue;
        }
        col = col + thr * t.emit;
        V3 fd = dielectric_f0(t.ior);
        V3 diffuse = base * (1.0f - t.metal);
        if (maxc(diffuse) > 0.f) diffuse = diffuse * under_layer_real(t.ior, fmax_(t.rough, 0.04f), fmax_(dot(sn, -d), 1e-4f));
        Surf s{pos, sn, -d, diffuse, mixv(fd, base, t.metal), fmax_(t.rough, 0.04f), t.ior, 1.0f - t.metal};
        V3 cc = direct<F>(c, s, &seed) * 1.0f;
        V3 bounced = v3(0, 0, 0);
        float pd = dot(vmax(s.albedo, 0.0f), v3(0.2126f, 0.7152f, 0.0722f));
        float pg = 0.0f;
        float p_diffuse = (pd + pg > 0.f) ? pd / (pd + pg) : 1.0f;
        bool take_diffuse = rnd(&seed) < p_diffuse;
        if (take_diffuse && pd > 0.f) bounced = bounced + s.albedo * indirect<F>(c, pos, sn, &seed, diffuse_on) / p_diffuse;
        (void)rnd(&seed);  // ggx_sample_vndf's two numbers (shade.wgsl:317)
        (void)rnd(&seed);
        cc = cc + bounced;
        col = col + thr * cc;
        break;
    }
    out.lin = col;
    return out;
}

__device__ __forceinline__ float ldf(const uint32_t* px, size_t w) { return __uint_as_float(px[w]); }
__device__ __forceinline__ void stf(uint32_t* px, size_t w, float v) { px[w] = __float_as_uint(v); }

// The display plane (words 4n..7n) and the packed RGBA8 word (word i) from
// a pixel's new sums, as util.wgsl:139-144 writes them after each sample.
__device__ __forceinline__ void put_shown(uint32_t* px, size_t n, size_t i, const PhrCamera& cam, V3 l, V3 gs, float cov,
                                          float gone, float film, uint32_t taken) {
    const float k = (float)taken;
    const V3 s = shown_of(l, gs, cov, gone, k, make_float4(cam.view[0], cam.view[1], cam.view[2], cam.view[3]));
    stf(px, 4 * n + 3 * i, s.x);
    stf(px, 4 * n + 3 * i + 1, s.y);
    stf(px, 4 * n + 3 * i + 2, s.z);
    px[i] = pack_rgba8(s, film / k);
}

// Samples [frame0, frame0 + spp) of every pixel of tile `blk`, summed into px
// in sample order (util.wgsl:17-34, 101-145: sums restart at frame 0).
template <bool F>
__device__ __forceinline__ void path_block(const SceneDev& sc, const PhrCamera* __restrict__ camp, uint32_t blk, uint32_t tiles_x,
                                           uint32_t frame0, uint32_t spp, uint32_t* __restrict__ px,
                                           unsigned long long* __restrict__ rays, uint32_t listed, uint32_t* s_stack,
                                           uint16_t* s_tstack) {
    const PhrCamera& cam = *camp;
    const size_t n = (size_t)cam.w * cam.h;
    // A block traces a 16 x 16 tile, a wave 16 x 4 pixels of it: neighbouring
    // rays walk neighbouring nodes. With Adaptive Sampling past the first
    // sample (`listed`), thread k takes the k-th pixel still open instead
    // (shade.wgsl:47-50); the list holds them tile by tile.
    uint32_t i = 0;
    bool inside;
    if (listed) {
        const uint32_t k = blk * TRACE_THREADS + threadIdx.x;
        inside = k < min((size_t)px[34 * n], n);
        if (inside) i = px[32 * n + k];
        inside = inside && i < n;
    } else {
        const uint32_t tx = blk % tiles_x, ty = blk / tiles_x;
        const uint32_t pxl = tx * 16 + (threadIdx.x & 15u), pyl = ty * 16 + (threadIdx.x >> 4);
        inside = pxl < cam.w && pyl < cam.h;
        i = pyl * cam.w + pxl;
    }
    const bool adapt = cam.adapt[3] != 0.0f;
    Ctx c;
    c.s = sc;
    c.p = camp;
    c.stk = s_stack + threadIdx.x;
    c.tstk = s_tstack + (F ? threadIdx.x : 0u);
    c.closest = 0;
    c.shadow = 0;
    c.steps = 0;
    c.tests = 0;
    if (inside) {
        bool keep = frame0 > 0u;
        V3 l = keep ? v3(ldf(px, n + 3 * i), ldf(px, n + 3 * i + 1), ldf(px, n + 3 * i + 2)) : v3(0, 0, 0);
        c.ga = keep ? v3(ldf(px, 7 * n + 3 * i), ldf(px, 7 * n + 3 * i + 1), ldf(px, 7 * n + 3 * i + 2)) : v3(0, 0, 0);
        c.gn = keep ? v3(ldf(px, 10 * n + 3 * i), ldf(px, 10 * n + 3 * i + 1), ldf(px, 10 * n + 3 * i + 2)) : v3(0, 0, 0);
        c.gd = keep ? ldf(px, 13 * n + i) : 0.f;
        c.cov = keep ? ldf(px, 23 * n + i) : 0.f;
        c.gone = keep ? ldf(px, 33 * n + i) : 0.f;
        // The half image (util.wgsl:126-132): odd samples at double weight; a
        // new image starts it at zero.
        V3 ax = (adapt && keep) ? v3(ldf(px, 26 * n + 3 * i), ldf(px, 26 * n + 3 * i + 1), ldf(px, 26 * n + 3 * i + 2))
                                : v3(0, 0, 0);
        c.outline = 0;
        for (uint32_t s = 0; s < spp; ++s) {
            Sample sm = shade<F>(c, i, frame0 + s);
            if (adapt && ((frame0 + s) & 1u) == 1u) ax = ax + sm.lin * 2.0f;
            l = l + sm.lin;
        }
        const V3 ga = c.ga, gn = c.gn;
        const float gd = c.gd, cov = c.cov, gone = c.gone;
        const uint32_t outline = c.outline;
        // The grid plane, the film weight and the count take the same
        // per-sample additions as the sums above (+0, +1.0, +1), done after
        // the samples so they hold no registers across them.
        V3 gs = keep ? v3(ldf(px, 20 * n + 3 * i), ldf(px, 20 * n + 3 * i + 1), ldf(px, 20 * n + 3 * i + 2)) : v3(0, 0, 0);
        float film = keep ? ldf(px, 24 * n + i) : 0.f;
        uint32_t taken = keep ? px[31 * n + i] : 0u;
        for (uint32_t s = 0; s < spp; ++s) {
            gs = gs + v3(0.f, 0.f, 0.f);
            film = film + 1.0f;
        }
        taken = taken + spp;
        if (frame0 == 0u) { px[29 * n + i] = 0u; px[30 * n + i] = 0u; }
        stf(px, n + 3 * i, l.x); stf(px, n + 3 * i + 1, l.y); stf(px, n + 3 * i + 2, l.z);
        stf(px, 7 * n + 3 * i, ga.x); stf(px, 7 * n + 3 * i + 1, ga.y); stf(px, 7 * n + 3 * i + 2, ga.z);
        stf(px, 10 * n + 3 * i, gn.x); stf(px, 10 * n + 3 * i + 1, gn.y); stf(px, 10 * n + 3 * i + 2, gn.z);
        stf(px, 13 * n + i, gd);
        stf(px, 20 * n + 3 * i, gs.x); stf(px, 20 * n + 3 * i + 1, gs.y); stf(px, 20 * n + 3 * i + 2, gs.z);
        stf(px, 23 * n + i, cov);
        stf(px, 24 * n + i, film);
        px[25 * n + i] = outline;
        px[31 * n + i] = taken;
        stf(px, 33 * n + i, gone);
        put_shown(px, n, i, cam, l, gs, cov, gone, film, taken);
        if (adapt) { stf(px, 26 * n + 3 * i, ax.x); stf(px, 26 * n + 3 * i + 1, ax.y); stf(px, 26 * n + 3 * i + 2, ax.z); }
    }
    // Rays traced, summed over the wave before one atomic.
    unsigned long long rc = c.closest, rs = c.shadow, st = c.steps, tt = c.tests;
    for (int o = 32; o > 0; o >>= 1) {
        rc += __shfl_xor(rc, o);
        rs += __shfl_xor(rs, o);
        st += __shfl_xor(st, o);
        tt += __shfl_xor(tt, o);
    }
    if ((threadIdx.x & 63) == 0 && (rc | rs)) {
        atomicAdd(&rays[0], rc);
        atomicAdd(&rays[1], rs);
        atomicAdd(&rays[2], st);
        atomicAdd(&rays[3], tt);
    }
}

// ---- one segment of a sample: the regenerating builds' unit --------------------
// A sample of shade + indirect for one tree, with no Adaptive Sampling and no
// studio rig (the camera has lamps or a lit world), as a state advanced one
// segment a turn: a closest ray, its shading, at most one shadow ray. Each
// sample's arithmetic and random numbers are shade's and indirect's, in the
// same order, so the sums are the same bits. k_path_rg keeps the state in
// registers; the wavefront build (k_wf_*) keeps it in HBM between kernels.
struct RgConst {  // the camera's constants (shade.wgsl:41-64, integrate.wgsl:111-123)
    uint32_t see_through, passes, depth;
    bool diffuse_on, lit, ao_only;
};
__device__ __forceinline__ RgConst rg_const(const PhrCamera* p) {
    RgConst k;
    const bool depth_set = p->depth[0] != 0u;
    const uint32_t total = depth_set ? p->depth[0] - 1u : 8u;
    k.see_through = depth_set ? p->depth[1] - 1u : 8u;
    const uint32_t through = depth_set ? min(p->depth[2] - 1u, total) : 8u;
    k.diffuse_on = !depth_set || (p->bounces > 0u && total > 0u);
    k.passes = k.see_through + through + 1u;
    k.lit = world_lit(p);
    k.ao_only = p->glow == 0u && !k.lit;
    k.depth = k.diffuse_on ? (p->bounces != 0u ? p->bounces : 1u) : 1u;
    return k;
}
// The sample in flight. phase 0: a new sample; 1: the camera ray's segments
// (shade's pass loop); 2: the diffuse bounces.
struct RgSample {
    uint32_t seed, pass, n_see, b, phase;
    V3 o, d, col, cc, alb, acc, thr, nrm;
    float p_diffuse, min_pdf;
};
// The sample's guides, set by its first segment (shade's `out`).
struct RgGuide {
    V3 alb, nrm;
    float dep, missed, cov;
    uint32_t outline;
};
// What a turn leaves for its shadow ray and for after it. end: 0 the sample
// goes on; 1 it ends with col as it is; 2 the camera ray's direct light then
// no bounce; 3 the ambient-occlusion bounce; 4 the diffuse bounces end; 5
// they end after this turn's lamp.
struct RgTurn {
    bool qon, qdirect, to_bounces;
    uint32_t end;
    V3 qpos, qdir, qc, thr_was;
    float qdist, near;
};

// Phase 0: sample `frame` of pixel i starts (shade.wgsl:65-95).
__device__ __forceinline__ void rg_start(RgSample& st, RgGuide& g, const PhrCamera* p, uint32_t i, uint32_t frame) {
    const float fw = (float)p->w, fh = (float)p->h;
    const float su = ((float)(i % p->w) + 0.5f) / fw - 0.5f;
    const float sv = ((float)(i / p->w) + 0.5f) / fh - 0.5f;
    st.seed = pcg_hash(i * 9781u + (frame + p->stream) * 6271u + 1u);
    const float ju = rnd(&st.seed);
    const float jv = rnd(&st.seed);
    const float jx = (ju - 0.5f) / fw;
    const float jy = (jv - 0.5f) / fh;
    const float vx = (su + jx - p->pan[0]) / p->scale;
    const float vy = -((sv + jy) / p->aspect - p->pan[1]) / p->scale;
    if (p->persp != 0u) {
        st.o = ld3(p->pivot) + ld3(p->back) * (p->dist / p->units);
        st.d = normalize(ld3(p->right) * vx + ld3(p->up) * vy - ld3(p->back) * p->dist);
    } else {
        st.o = ld3(p->pivot) + (ld3(p->right) * vx + ld3(p->up) * vy + ld3(p->back) * 1.0e4f) / p->units;
        st.d = -ld3(p->back);
    }
    st.col = v3(0, 0, 0);
    g.missed = 0.f;
    g.cov = 0.f;
    g.outline = 0u;
    st.pass = 0u;
    st.n_see = 0u;
    st.phase = 1u;
}

// The turn's closest hit h: shading up to the shadow ray (shade.wgsl:96-302
// for a camera segment, integrate.wgsl:124-228 for a bounce). F: the scene
// is a forest (faces placed by their instance, a mirroring instance's
// camera-segment normal turned back, as shade<true> does).
template <bool F = false>
__device__ __forceinline__ void rg_hit(RgSample& st, RgGuide& g, RgTurn& q, const SceneDev& sc, const PhrCamera* p, const RgConst& k,
                                       const PhrHit& h) {
    const V3 one = v3(1, 1, 1);  // shade's throughput, never changed
    q.qon = q.qdirect = q.to_bounces = false;
    q.qpos = v3(0, 0, 0);
    q.qdir = v3(0, 0, 1);
    q.qc = v3(0, 0, 0);
    q.qdist = 0.f;
    q.near = 1.0f;
    q.thr_was = st.thr;
    q.end = 0u;
    if (st.phase == 1u) {
        if (h.tri == PHR_NONE) {
            if (st.pass == 0u) {
                g.alb = v3(0, 0, 0);
                g.nrm = -st.d;
                g.dep = -1.0f;
                st.col = v3(0, 0, 0);
                g.missed = 1.0f;
                g.cov = 1.0f;
            } else {
                st.col = st.col + one * env_rgb(p);
            }
            q.end = 1u;
            return;
        }
        if (st.pass == 0u) g.outline = h.inst + 1u;
        const uint32_t m = mat_at(sc, h.tri);
        const Tri t = tri_of<F>(sc, h.tri, h.inst, m);
        const float w = 1.0f - h.u - h.v;
        const V3 hit = st.o + st.d * h.t;
        V3 ng0 = normalize(cross(t.b - t.a, t.c - t.a));
        if constexpr (F)
            if (sc.insts[INSTANCE_LEN * (size_t)h.inst + 27] & 4u) ng0 = -ng0;
        const bool front = dot(ng0, st.d) < 0.f;
        const V3 ng = front ? ng0 : -ng0;
        V3 sn = t.na * w + t.nb * h.u + t.nc * h.v;
        sn = dot(sn, sn) > 1e-12f ? normalize(sn) * (front ? 1.0f : -1.0f) : ng;
        if (dot(sn, -st.d) < 0.f) sn = ng;
        const float start = 1.0e4f / p->units;
        const float lift = (p->persp == 0u && h.t > start * 0.5f) ? p->eps + fabsf(h.t - start) * 1e-5f + start * 2.4e-7f
                                                                  : p->eps + h.t * 1e-5f;
        const V3 pos = hit + ng * lift;
        const V3 base = vmax(v3(t.color[0], t.color[1], t.color[2]), 0.0f);
        const float alpha = t.color[3];
        if (st.pass == 0u) {
            g.alb = base;
            g.nrm = sn;
            g.dep = h.t;
        }
        if (rnd(&st.seed) >= alpha) {
            st.n_see++;
            if (st.n_see > k.see_through) {
                q.end = 1u;
            } else {
                const float reach = fmax_(fmax_(fabsf(hit.x), fabsf(hit.y)), fmax_(fabsf(hit.z), 0.5f));
                const float tl = lift - p->eps + reach * (256.0f / 8388608.0f);
                st.o = hit - ng * tl;
                if (++st.pass >= k.passes) q.end = 1u;
            }
            return;
        }
        st.col = st.col + one * t.emit;
        const V3 fd = dielectric_f0(t.ior);
        V3 diffuse = base * (1.0f - t.metal);
        if (maxc(diffuse) > 0.f) diffuse = diffuse * under_layer_real(t.ior, fmax_(t.rough, 0.04f), fmax_(dot(sn, -st.d), 1e-4f));
        const Surf sf{pos, sn, -st.d, diffuse, mixv(fd, base, t.metal), fmax_(t.rough, 0.04f), t.ior, 1.0f - t.metal};
        // direct (integrate.wgsl:66-85) without the rig: the lamp pick's
        // number, then one lamp.
        (void)rnd(&st.seed);
        const Shadowed sh = glow_query<F>(sc, sf, &st.seed);
        if (maxc(sh.c) > 0.f) {
            q.qon = q.qdirect = true;
            q.qpos = sf.pos;
            q.qdir = sh.dir;
            q.qdist = sh.dist;
            q.qc = sh.c;
        }
        const float pd = dot(vmax(sf.albedo, 0.0f), v3(0.2126f, 0.7152f, 0.0722f));
        const float pg = 0.0f;
        st.p_diffuse = (pd + pg > 0.f) ? pd / (pd + pg) : 1.0f;
        const bool take_diffuse = rnd(&st.seed) < st.p_diffuse;
        st.alb = sf.albedo;
        q.end = 2u;
        if (take_diffuse && pd > 0.f) {
            if (k.ao_only) {  // integrate.wgsl:113-117
                q.qon = true;
                q.qpos = pos;
                q.qdir = cosine_dir(sn, &st.seed);
                q.qdist = p->ao;
                q.end = 3u;
            } else {  // the first bounce's ray (integrate.wgsl:124-128)
                st.acc = v3(0, 0, 0);
                st.thr = v3(1, 1, 1);
                st.o = pos;
                st.nrm = sn;
                st.min_pdf = 3.0e38f;
                st.b = 0u;
                st.d = cosine_dir(st.nrm, &st.seed);
                const float pb = fmax_(dot(st.nrm, st.d), 0.0f) / PHR_PI;
                st.min_pdf = fmin_(st.min_pdf, pb);
                q.end = 0u;
                q.to_bounces = true;
            }
        }
        return;
    }
    // phase 2: bounce b
    if (h.tri == PHR_NONE) {
        st.acc = st.acc + st.thr * env_rgb(p) * 1.0f;
        q.end = 4u;
        return;
    }
    if (!k.diffuse_on) {
        q.end = 4u;
        return;
    }
    const Tri t = tri_of<F>(sc, h.tri, h.inst, mat_at(sc, h.tri));
    const V3 ng0 = normalize(cross(t.b - t.a, t.c - t.a));
    const V3 ng = dot(ng0, st.d) < 0.f ? ng0 : -ng0;
    const V3 hit = st.o + st.d * h.t + ng * (p->eps + h.t * 1e-5f);
    const V3 base = vmax(v3(t.color[0], t.color[1], t.color[2]), 0.0f);
    const float rough = fmax_(t.rough, glossy_floor(p, st.min_pdf));
    const bool glossy = p->integrator[0] < 0.5f || (((uint32_t)(p->integrator[3] + 0.5f)) & 1u) != 0u;
    const V3 fd = glossy ? dielectric_f0(t.ior) : v3(0, 0, 0);
    const float metal = glossy ? t.metal : 0.0f;
    const float tw = t.trans * (1.0f - t.metal);
    V3 albedo = base * (1.0f - t.metal) * (1.0f - tw);
    if (maxc(albedo) > 0.f) albedo = albedo * under_layer(fd, rough, fmax_(dot(ng, -st.d), 1e-4f));
    st.acc = st.acc + st.thr * glow_mis(sc, t, h, st.nrm, st.d);
    q.near = (!k.lit && h.t <= p->ao) ? 0.25f : 1.0f;
    const Surf sf{hit, ng, -st.d, albedo, mixv(fd, base, metal), rough, 0.0f, 0.0f};
    (void)rnd(&st.seed);
    const Shadowed sh = glow_query<F>(sc, sf, &st.seed);
    if (maxc(sh.c) > 0.f) {
        q.qon = q.qdirect = true;
        q.qpos = sf.pos;
        q.qdir = sh.dir;
        q.qdist = sh.dist;
        q.qc = sh.c;
    }
    q.thr_was = st.thr;
    st.thr = st.thr * albedo;
    bool more = st.b + 1u < k.depth;
    if (more && st.b + 1u >= p->rr_from) {
        const float keep_p = clampf(maxc(st.thr), 0.0f, 0.95f);
        if (keep_p < 1e-4f || rnd(&st.seed) > keep_p) more = false;
        else st.thr = st.thr / keep_p;
    }
    if (more) {
        st.o = hit;
        st.nrm = ng;
        ++st.b;
        st.d = cosine_dir(st.nrm, &st.seed);
        const float pb = fmax_(dot(st.nrm, st.d), 0.0f) / PHR_PI;
        st.min_pdf = fmin_(st.min_pdf, pb);
    } else {
        q.end = 5u;  // after this turn's lamp: the bounces end
    }
}

// What a turn's end changes of its sample: the radiance, the camera hit's
// light and the bounces' sum, and the phase.
struct RgPost {
    V3 col, cc, acc;
    uint32_t phase;
};
// After the turn's shadow ray (vis: 1 lit or no ray, 0 blocked). True when
// the sample has ended; its radiance is then r.col.
__device__ __forceinline__ bool rg_post_of(RgPost& r, V3 alb, float p_diffuse, const RgTurn& q, const PhrCamera* p, float vis) {
    const V3 one = v3(1, 1, 1);
    uint32_t end = q.end;
    V3 dsum = v3(0, 0, 0);
    if (q.qdirect) dsum = dsum + q.qc * vis;
    if (r.phase == 1u && (end >= 2u || q.to_bounces)) {
        r.cc = dsum * 1.0f;
        if (end == 2u) {
            r.cc = r.cc + v3(0, 0, 0);
            r.col = r.col + one * r.cc;
            end = 1u;
        } else if (end == 3u) {
            V3 bounced = v3(0, 0, 0);
            bounced = bounced + alb * (env_rgb(p) * mixf(0.25f, 1.0f, vis)) / p_diffuse;
            r.cc = r.cc + bounced;
            r.col = r.col + one * r.cc;
            end = 1u;
        }
    } else if (r.phase == 2u) {
        if (end == 5u) {
            r.acc = r.acc + q.thr_was * dsum * q.near;
            end = 4u;
        } else if (end == 0u) {
            r.acc = r.acc + q.thr_was * dsum * q.near;
        }
        if (end == 4u) {
            V3 bounced = v3(0, 0, 0);
            bounced = bounced + alb * r.acc / p_diffuse;
            r.cc = r.cc + bounced;
            r.col = r.col + one * r.cc;
            end = 1u;
        }
    }
    if (q.to_bounces) r.phase = 2u;
    return end == 1u;
}
__device__ __forceinline__ bool rg_post(RgSample& st, const RgTurn& q, const PhrCamera* p, float vis) {
    RgPost r{st.col, st.cc, st.acc, st.phase};
    const bool ended = rg_post_of(r, st.alb, st.p_diffuse, q, p, vis);
    st.col = r.col;
    st.cc = r.cc;
    st.acc = r.acc;
    st.phase = r.phase;
    return ended;
}

// The same samples as path_block for one tree, with no Adaptive Sampling
// and no studio rig, with path regeneration: each lane advances its sample
// one segment a turn (every lane's closest ray and at most one shadow ray
// share a call site), and a lane whose sample ends starts its pixel's next
// sample on the next turn instead of waiting for the wave's longest path.
template <int WPE>
__global__ __launch_bounds__(TRACE_THREADS) __attribute__((amdgpu_waves_per_eu(WPE))) void k_path_rg(
    SceneDev sc, const PhrCamera* __restrict__ camp, uint32_t tiles_x, uint32_t frame0, uint32_t spp, uint32_t* __restrict__ px,
    unsigned long long* __restrict__ rays, uint32_t tile0) {
    __shared__ uint32_t s_stack[PHR_STACK * TRACE_THREADS];
    const PhrCamera* p = camp;
    const size_t n = (size_t)p->w * p->h;
    const uint32_t blk = blockIdx.x + tile0;
    const uint32_t tx = blk % tiles_x, ty = blk / tiles_x;
    const uint32_t pxl = tx * 16 + (threadIdx.x & 15u), pyl = ty * 16 + (threadIdx.x >> 4);
    const bool inside = pxl < p->w && pyl < p->h;
    const uint32_t i = pyl * p->w + pxl;
    Ctx c;
    c.s = sc;
    c.p = camp;
    c.stk = s_stack + threadIdx.x;
    c.tstk = nullptr;
    c.closest = 0;
    c.shadow = 0;
    c.steps = 0;
    c.tests = 0;
    if (inside) {
        const bool keep = frame0 > 0u;
        V3 l = keep ? v3(ldf(px, n + 3 * i), ldf(px, n + 3 * i + 1), ldf(px, n + 3 * i + 2)) : v3(0, 0, 0);
        V3 ga = keep ? v3(ldf(px, 7 * n + 3 * i), ldf(px, 7 * n + 3 * i + 1), ldf(px, 7 * n + 3 * i + 2)) : v3(0, 0, 0);
        V3 gn = keep ? v3(ldf(px, 10 * n + 3 * i), ldf(px, 10 * n + 3 * i + 1), ldf(px, 10 * n + 3 * i + 2)) : v3(0, 0, 0);
        float gd = keep ? ldf(px, 13 * n + i) : 0.f;
        float cov = keep ? ldf(px, 23 * n + i) : 0.f;
        float gone = keep ? ldf(px, 33 * n + i) : 0.f;
        uint32_t outline = 0;
        const RgConst k = rg_const(p);
        RgSample st;
        st.seed = 0;
        st.pass = st.n_see = st.b = st.phase = 0u;
        st.o = st.d = st.col = st.cc = st.alb = st.acc = st.nrm = v3(0, 0, 0);
        st.thr = v3(1, 1, 1);
        st.p_diffuse = 1.0f;
        st.min_pdf = 3.0e38f;
        RgGuide g{v3(0, 0, 0), v3(0, 0, 0), 0.f, 0.f, 0.f, 0u};
        uint32_t s = 0;
        while (s < spp) {
            if (st.phase == 0u) rg_start(st, g, p, i, frame0 + s);
            const PhrHit h = trace<false>(c, st.o, st.d, PHR_INF, false);
            RgTurn q;
            rg_hit(st, g, q, c.s, p, k, h);
            // The turn's shadow ray, one call site for every lane.
            float vis = 1.0f;
            if (q.qon) vis = trace<false>(c, q.qpos, q.qdir, q.qdist, true).tri != PHR_NONE ? 0.0f : 1.0f;
            if (rg_post(st, q, p, vis)) {  // shade's `out`, summed as path_block sums it
                ga = ga + g.alb;
                gn = gn + g.nrm;
                gd = gd + g.dep;
                gone = gone + g.missed;
                l = l + st.col;
                cov = cov + g.cov;
                outline = g.outline;
                ++s;
                st.phase = 0u;
            }
        }
        V3 gs = keep ? v3(ldf(px, 20 * n + 3 * i), ldf(px, 20 * n + 3 * i + 1), ldf(px, 20 * n + 3 * i + 2)) : v3(0, 0, 0);
        float film = keep ? ldf(px, 24 * n + i) : 0.f;
        uint32_t taken = keep ? px[31 * n + i] : 0u;
        for (uint32_t kk = 0; kk < spp; ++kk) {
            gs = gs + v3(0.f, 0.f, 0.f);
            film = film + 1.0f;
        }
        taken = taken + spp;
        if (frame0 == 0u) { px[29 * n + i] = 0u; px[30 * n + i] = 0u; }
        stf(px, n + 3 * i, l.x); stf(px, n + 3 * i + 1, l.y); stf(px, n + 3 * i + 2, l.z);
        stf(px, 7 * n + 3 * i, ga.x); stf(px, 7 * n + 3 * i + 1, ga.y); stf(px, 7 * n + 3 * i + 2, ga.z);
        stf(px, 10 * n + 3 * i, gn.x); stf(px, 10 * n + 3 * i + 1, gn.y); stf(px, 10 * n + 3 * i + 2, gn.z);
        stf(px, 13 * n + i, gd);
        stf(px, 20 * n + 3 * i, gs.x); stf(px, 20 * n + 3 * i + 1, gs.y); stf(px, 20 * n + 3 * i + 2, gs.z);
        stf(px, 23 * n + i, cov);
        stf(px, 24 * n + i, film);
        px[25 * n + i] = outline;
        px[31 * n + i] = taken;
        stf(px, 33 * n + i, gone);
        put_shown(px, n, i, *p, l, gs, cov, gone, film, taken);
    }
    unsigned long long rc = c.closest, rs = c.shadow, st_ = c.steps, tt = c.tests;
    for (int kk = 32; kk > 0; kk >>= 1) {
        rc += __shfl_xor(rc, kk);
        rs += __shfl_xor(rs, kk);
        st_ += __shfl_xor(st_, kk);
        tt += __shfl_xor(tt, kk);
    }
    if ((threadIdx.x & 63) == 0 && (rc | rs)) {
        atomicAdd(&rays[0], rc);
        atomicAdd(&rays[1], rs);
        atomicAdd(&rays[2], st_);
        atomicAdd(&rays[3], tt);
    }
}

// The same samples as k_path_rg with walks that span turns. Each lane keeps
// one walk (a Walker: the sample's closest ray, or its shadow ray) and the
// wave walks only until shade8 eighths (1 to 8) of its lanes still at work
// have finished theirs. Those lanes then shade (rg_hit after a closest walk, the turn's end
// after a shadow walk) and begin their next walk, while the others keep
// their walk where it stopped for the next turn. A wave no longer waits on
// its longest walk, and closest and shadow walks share one loop. A lane's
// walks and shading come in k_path_rg's order, so the sums are the same bits.
// A turn with a shadow ray takes its end both ways (lamp seen, lamp blocked)
// before the walk: a sample that ended waits with its radiance both ways; a
// sample that goes on is advanced as lit and, if the ray was blocked, gets
// back the two sums that differ (the camera hit's light, the bounces' sum).
// The split build's walk of one tree, or of a forest (F).
template <bool F>
struct SplitWalk {
    Walker w;
    __device__ __forceinline__ void begin(const SceneDev& s, V3 o, V3 d, float tmax, bool any) { walker_begin(w, s, o, d, tmax, any); }
    __device__ __forceinline__ void step(const SceneDev& s, uint32_t* stk, uint16_t*, uint32_t& steps, uint32_t& tests) {
        walker_step<true>(w, s, stk, TRACE_THREADS, steps, tests);
    }
    __device__ __forceinline__ PhrHit hit() const { return PhrHit{w.t, w.tri, w.u, w.v, 0u}; }
};
template <>
struct SplitWalk<true> {
    ForestWalker w;
    __device__ __forceinline__ void begin(const SceneDev& s, V3 o, V3 d, float tmax, bool any) { forest_begin(w, s, o, d, tmax, any); }
    __device__ __forceinline__ void step(const SceneDev& s, uint32_t* stk, uint16_t* tstk, uint32_t& steps, uint32_t& tests) {
        forest_step(w, s, tstk, stk, TRACE_THREADS, steps, tests);
    }
    __device__ __forceinline__ PhrHit hit() const { return PhrHit{w.t, w.tri, w.u, w.v, w.hinst}; }
};

// path_split: tile `blk` of one image. s_stack: the block's stack columns
// (a forest's mesh trees'); s_tstack: a forest's top-tree columns.
template <bool F = false>
__device__ __forceinline__ void path_split(const SceneDev& sc, const PhrCamera* __restrict__ camp, uint32_t blk, uint32_t tiles_x,
                                           uint32_t frame0, uint32_t spp, uint32_t* __restrict__ px,
                                           unsigned long long* __restrict__ rays, uint32_t* s_stack, uint16_t* s_tstack,
                                           uint32_t shade8) {
    const PhrCamera* p = camp;
    const size_t n = (size_t)p->w * p->h;
    const uint32_t tx = blk % tiles_x, ty = blk / tiles_x;
    const uint32_t pxl = tx * 16 + (threadIdx.x & 15u), pyl = ty * 16 + (threadIdx.x >> 4);
    const bool inside = pxl < p->w && pyl < p->h;
    const uint32_t i = pyl * p->w + pxl;
    uint32_t* stk = s_stack + threadIdx.x;
    uint16_t* tstk = F ? s_tstack + threadIdx.x : nullptr;
    uint32_t closest = 0, shadow = 0, steps = 0, tests = 0;
    const bool keep = frame0 > 0u;
    V3 l = v3(0, 0, 0), ga = l, gn = l;
    float gd = 0.f, cov = 0.f, gone = 0.f;
    if (inside && keep) {
        l = v3(ldf(px, n + 3 * i), ldf(px, n + 3 * i + 1), ldf(px, n + 3 * i + 2));
        ga = v3(ldf(px, 7 * n + 3 * i), ldf(px, 7 * n + 3 * i + 1), ldf(px, 7 * n + 3 * i + 2));
        gn = v3(ldf(px, 10 * n + 3 * i), ldf(px, 10 * n + 3 * i + 1), ldf(px, 10 * n + 3 * i + 2));
        gd = ldf(px, 13 * n + i);
        cov = ldf(px, 23 * n + i);
        gone = ldf(px, 33 * n + i);
    }
    uint32_t outline = 0;
    const RgConst k = rg_const(p);
    RgSample st;
    st.seed = 0;
    st.pass = st.n_see = st.b = st.phase = 0u;
    st.o = st.d = st.col = st.cc = st.alb = st.acc = st.nrm = v3(0, 0, 0);
    st.thr = v3(1, 1, 1);
    st.p_diffuse = 1.0f;
    st.min_pdf = 3.0e38f;
    RgGuide g0;  // what rg_start clears of a sample's guide (the first segment's rg_hit sets it anew)
    SplitWalk<F> sw;
    auto& w = sw.w;
    w.on = w.any = false;
    bool done = !inside || spp == 0u, pfinal = false;
    V3 pa = v3(0, 0, 0), pb = pa;  // what a shadow walk's answer picks between
    uint32_t s = 0;
    if (!done) {
        rg_start(st, g0, p, i, frame0);
        sw.begin(sc, st.o, st.d, PHR_INF, false);
        ++closest;
    }
    while (__ballot(!done) != 0ull) {
        for (;;) {  // walk until shade8 / 8 of the lanes at work wait to shade, or all do
            const uint32_t waiting = (uint32_t)__popcll(__ballot(!done && !w.on));
            const uint32_t working = (uint32_t)__popcll(__ballot(!done));
            if (waiting == working || 8u * waiting >= shade8 * working) break;
            if (w.on) sw.step(sc, stk, tstk, steps, tests);
        }
        if (done || w.on) continue;
        bool next = false;
        if (w.any) {  // the turn's shadow ray has its answer
            const bool blocked = w.tri != PHR_NONE;
            if (pfinal) {
                l = l + (blocked ? pb : pa);
            } else if (blocked) {
                st.cc = pa;
                st.acc = pb;
            }
            next = true;
        } else {  // the closest hit: the segment's shading
            const PhrHit h = sw.hit();
            RgTurn q;
            // A sample's guide is whole once its first camera segment is
            // shaded; it joins the sums then, in sample order as at the
            // sample's end, and is not carried through the walks.
            const bool first = st.phase == 1u && st.pass == 0u;
            RgGuide g{v3(0, 0, 0), v3(0, 0, 0), 0.f, 0.f, 0.f, 0u};
            rg_hit<F>(st, g, q, sc, p, k, h);
            if (first) {  // shade's `out`, summed as path_block sums it
                ga = ga + g.alb;
                gn = gn + g.nrm;
                gd = gd + g.dep;
                gone = gone + g.missed;
                cov = cov + g.cov;
                outline = g.outline;
            }
            bool ended;
            if (q.qon) {
                RgPost lit{st.col, st.cc, st.acc, st.phase}, dark = lit;
                ended = rg_post_of(lit, st.alb, st.p_diffuse, q, p, 1.0f);
                rg_post_of(dark, st.alb, st.p_diffuse, q, p, 0.0f);
                pfinal = ended;
                if (ended) {
                    pa = lit.col;
                    pb = dark.col;
                } else {
                    st.col = lit.col;
                    st.cc = lit.cc;
                    st.acc = lit.acc;
                    st.phase = lit.phase;
                    pa = dark.cc;
                    pb = dark.acc;
                }
                sw.begin(sc, q.qpos, q.qdir, q.qdist, true);
                ++shadow;
            } else {
                ended = rg_post(st, q, p, 1.0f);
                if (ended) l = l + st.col;
                next = true;
            }
            if (ended) {
                ++s;
                st.phase = 0u;
            }
        }
        if (next) {
            if (s < spp) {
                if (st.phase == 0u) rg_start(st, g0, p, i, frame0 + s);
                sw.begin(sc, st.o, st.d, PHR_INF, false);
                ++closest;
            } else {
                done = true;
            }
        }
    }
    if (inside) {
        V3 gs = keep ? v3(ldf(px, 20 * n + 3 * i), ldf(px, 20 * n + 3 * i + 1), ldf(px, 20 * n + 3 * i + 2)) : v3(0, 0, 0);
        float film = keep ? ldf(px, 24 * n + i) : 0.f;
        uint32_t taken = keep ? px[31 * n + i] : 0u;
        for (uint32_t kk = 0; kk < spp; ++kk) {
            gs = gs + v3(0.f, 0.f, 0.f);
            film = film + 1.0f;
        }
        taken = taken + spp;
        if (frame0 == 0u) { px[29 * n + i] = 0u; px[30 * n + i] = 0u; }
        stf(px, n + 3 * i, l.x); stf(px, n + 3 * i + 1, l.y); stf(px, n + 3 * i + 2, l.z);
        stf(px, 7 * n + 3 * i, ga.x); stf(px, 7 * n + 3 * i + 1, ga.y); stf(px, 7 * n + 3 * i + 2, ga.z);
        stf(px, 10 * n + 3 * i, gn.x); stf(px, 10 * n + 3 * i + 1, gn.y); stf(px, 10 * n + 3 * i + 2, gn.z);
        stf(px, 13 * n + i, gd);
        stf(px, 20 * n + 3 * i, gs.x); stf(px, 20 * n + 3 * i + 1, gs.y); stf(px, 20 * n + 3 * i + 2, gs.z);
        stf(px, 23 * n + i, cov);
        stf(px, 24 * n + i, film);
        px[25 * n + i] = outline;
        px[31 * n + i] = taken;
        stf(px, 33 * n + i, gone);
        put_shown(px, n, i, *p, l, gs, cov, gone, film, taken);
    }
    unsigned long long rc = closest, rs = shadow, st_ = steps, tt = tests;
    for (int kk = 32; kk > 0; kk >>= 1) {
        rc += __shfl_xor(rc, kk);
        rs += __shfl_xor(rs, kk);
        st_ += __shfl_xor(st_, kk);
        tt += __shfl_xor(tt, kk);
    }
    if ((threadIdx.x & 63) == 0 && (rc | rs)) {
        atomicAdd(&rays[0], rc);
        atomicAdd(&rays[1], rs);
        atomicAdd(&rays[2], st_);
        atomicAdd(&rays[3], tt);
    }
}
template <int WPE>
__global__ __launch_bounds__(TRACE_THREADS) __attribute__((amdgpu_waves_per_eu(WPE))) void k_path_sw(
    SceneDev sc, const PhrCamera* __restrict__ camp, uint32_t tiles_x, uint32_t frame0, uint32_t spp, uint32_t* __restrict__ px,
    unsigned long long* __restrict__ rays, uint32_t tile0, uint32_t shade8) {
    __shared__ uint32_t s_stack[PHR_STACK * TRACE_THREADS];
    path_split(sc, camp, blockIdx.x + tile0, tiles_x, frame0, spp, px, rays, s_stack, nullptr, shade8);
}
// The split build over a forest: MSTK mesh-tree entries a thread (the
// caller checked the mesh trees' depth when it is below PHR_STACK).
template <int WPE, uint32_t MSTK>
__global__ __launch_bounds__(TRACE_THREADS) __attribute__((amdgpu_waves_per_eu(WPE))) void k_path_sw_forest(
    SceneDev sc, const PhrCamera* __restrict__ camp, uint32_t tiles_x, uint32_t frame0, uint32_t spp, uint32_t* __restrict__ px,
    unsigned long long* __restrict__ rays, uint32_t tile0, uint32_t shade8) {
    __shared__ uint32_t s_stack[MSTK * TRACE_THREADS];
    __shared__ uint16_t s_tstack[PHR_STACK * TRACE_THREADS];
    path_split<true>(sc, camp, blockIdx.x + tile0, tiles_x, frame0, spp, px, rays, s_stack, s_tstack, shade8);
}

// ---- the wavefront build --------------------------------------------------------
// The regenerating build's samples with the walks taken out of the shading.
// A path slot a pixel of the 16 x 16 tiles keeps its sample's state in HBM;
// a turn is three launches:
//  * k_wf_walk<false> walks every queued slot's ray to its closest hit;
//  * k_wf_shade advances each slot one segment (rg_hit). A turn with no
//    shadow ray finishes there (rg_post); otherwise rg_post runs for both
//    visibilities and the field they change (the ended sample's radiance,
//    or the bounce sums) rides with the shadow ray. The next ray and the
//    shadow ray are appended to their queues by wave ballots (one atomic a
//    wave), and an ended sample's pixel starts its next sample at once;
//  * k_wf_walk<true> walks the shadow rays and applies what each turn left.
// The walkers are persistent: a lane whose walk ends takes the next queued
// ray while the rest of its wave walks on (refilled once half the wave is
// idle), so no wave waits for its longest walk, and they carry no shading
// registers. Each pixel has one sample in flight, so its sums are taken in
// sample order and are the bits k_path_rg writes.
constexpr uint32_t WF_THREADS = 256;
constexpr uint32_t WF_CTRS = 8;  // next-ray counts [2], extend head, shadow count, shadow head
enum : uint32_t { WF_END = 0, WF_CC = 1, WF_ACC = 2 };

struct WfDev {
    float4* s0;   // o.xyz, min_pdf
    float4* s1;   // d.xyz, p_diffuse
    float4* s2;   // col.xyz, seed
    float4* s3;   // cc.xyz, sample index
    float4* s4;   // alb.xyz, phase | b << 2
    float4* s5;   // acc.xyz, -
    float4* s6;   // thr.xyz, pass
    float4* s7;   // nrm.xyz, n_see
    float4* hit;  // t, tri, u, v
    uint32_t* q0;  // ray queues of slots, by turn parity
    uint32_t* q1;
    uint32_t* sq;  // the shadow queue of slots
    float4* sh;    // a slot's shadow ray: (pos, dist) (dir, -) (candidate if blocked, kind) (candidate if lit, -)
    uint8_t* vis;  // a slot's last shadow ray: 1 lit, 0 blocked
    uint8_t* pend; // 1 + kind while a slot's shadow result waits to be taken in
    uint32_t* ctr;
};
constexpr uint32_t WF_CHUNK = 256;       // queue entries a walker wave takes with one atomic
constexpr uint32_t WF_SHADE_ITEMS = 4;   // entries a shading thread takes: a block's appends cost one atomic a queue

__device__ __forceinline__ uint32_t wf_pixel(uint32_t slot, uint32_t tiles_x, uint32_t w, uint32_t h, bool& inside) {
    const uint32_t tile = slot >> 8, r = slot & 255u;
    const uint32_t x = (tile % tiles_x) * 16u + (r & 15u), y = (tile / tiles_x) * 16u + (r >> 4);
    inside = x < w && y < h;
    return y * w + x;
}
__device__ __forceinline__ uint32_t lanes_below(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}
// The wave's lanes with `want` take consecutive places of a queue: one
// atomic a wave. Every lane of the wave calls it.
__device__ __forceinline__ uint32_t wave_append(bool want, uint32_t* counter) {
    const uint64_t m = __ballot(want);
    if (m == 0) return 0u;
    const uint32_t leader = (uint32_t)__ffsll((unsigned long long)m) - 1u;
    uint32_t base = 0;
    if ((threadIdx.x & 63u) == leader) base = atomicAdd(counter, (uint32_t)__popcll(m));
    base = __shfl(base, (int)leader);
    return base + lanes_below(m);
}

__device__ __forceinline__ void wf_load(const WfDev& w, uint32_t slot, RgSample& st, uint32_t& s) {
    const float4 a = w.s0[slot], b = w.s1[slot], c = w.s2[slot], d = w.s3[slot];
    const float4 e = w.s4[slot], f = w.s5[slot], g = w.s6[slot], k = w.s7[slot];
    st.o = v3(a.x, a.y, a.z);
    st.min_pdf = a.w;
    st.d = v3(b.x, b.y, b.z);
    st.p_diffuse = b.w;
    st.col = v3(c.x, c.y, c.z);
    st.seed = __float_as_uint(c.w);
    st.cc = v3(d.x, d.y, d.z);
    s = __float_as_uint(d.w);
    st.alb = v3(e.x, e.y, e.z);
    st.phase = __float_as_uint(e.w) & 3u;
    st.b = __float_as_uint(e.w) >> 2;
    st.acc = v3(f.x, f.y, f.z);
    st.thr = v3(g.x, g.y, g.z);
    st.pass = __float_as_uint(g.w);
    st.nrm = v3(k.x, k.y, k.z);
    st.n_see = __float_as_uint(k.w);
}
__device__ __forceinline__ void wf_store(const WfDev& w, uint32_t slot, const RgSample& st, uint32_t s) {
    w.s0[slot] = make_float4(st.o.x, st.o.y, st.o.z, st.min_pdf);
    w.s1[slot] = make_float4(st.d.x, st.d.y, st.d.z, st.p_diffuse);
    w.s2[slot] = make_float4(st.col.x, st.col.y, st.col.z, __uint_as_float(st.seed));
    w.s3[slot] = make_float4(st.cc.x, st.cc.y, st.cc.z, __uint_as_float(s));
    w.s4[slot] = make_float4(st.alb.x, st.alb.y, st.alb.z, __uint_as_float(st.phase | (st.b << 2)));
    w.s5[slot] = make_float4(st.acc.x, st.acc.y, st.acc.z, 0.f);
    w.s6[slot] = make_float4(st.thr.x, st.thr.y, st.thr.z, __uint_as_float(st.pass));
    w.s7[slot] = make_float4(st.nrm.x, st.nrm.y, st.nrm.z, __uint_as_float(st.n_see));
}
__device__ __forceinline__ void add3(uint32_t* px, size_t at, V3 v) {
    stf(px, at, ldf(px, at) + v.x);
    stf(px, at + 1, ldf(px, at + 1) + v.y);
    stf(px, at + 2, ldf(px, at + 2) + v.z);
}

// Counters zeroed before a frame's first turn.
__global__ void k_wf_zero(uint32_t* ctr) {
    if (threadIdx.x < WF_CTRS) ctr[threadIdx.x] = 0u;
}

// Every slot of a pixel starts its first sample and joins the first queue;
// a new image's sums start at zero.
__global__ __launch_bounds__(WF_THREADS) void k_wf_init(const PhrCamera* __restrict__ camp, WfDev w, uint32_t tiles_x,
                                                        uint32_t nslots, uint32_t frame0, uint32_t* __restrict__ px) {
    const PhrCamera* p = camp;
    const size_t n = (size_t)p->w * p->h;
    const uint32_t slot = blockIdx.x * WF_THREADS + threadIdx.x;
    bool inside = false;
    const uint32_t i = slot < nslots ? wf_pixel(slot, tiles_x, p->w, p->h, inside) : 0u;
    if (inside) {
        if (frame0 == 0u) {
            for (int c = 0; c < 3; ++c) {
                stf(px, n + 3 * (size_t)i + c, 0.f);
                stf(px, 7 * n + 3 * (size_t)i + c, 0.f);
                stf(px, 10 * n + 3 * (size_t)i + c, 0.f);
            }
            stf(px, 13 * n + i, 0.f);
            stf(px, 23 * n + i, 0.f);
            stf(px, 33 * n + i, 0.f);
            px[25 * n + i] = 0u;
        }
        RgSample st;
        st.seed = 0;
        st.pass = st.n_see = st.b = st.phase = 0u;
        st.o = st.d = st.col = st.cc = st.alb = st.acc = st.nrm = v3(0, 0, 0);
        st.thr = v3(1, 1, 1);
        st.p_diffuse = 1.0f;
        st.min_pdf = 3.0e38f;
        RgGuide g{v3(0, 0, 0), v3(0, 0, 0), 0.f, 0.f, 0.f, 0u};
        rg_start(st, g, p, i, frame0);
        wf_store(w, slot, st, 0u);
        w.pend[slot] = 0u;
    }
    // The block's slots in order, one atomic a block.
    __shared__ uint32_t s_cnt, s_base;
    if (threadIdx.x == 0) s_cnt = 0u;
    __syncthreads();
    const uint32_t at = wave_append(inside, &s_cnt);
    __syncthreads();
    if (threadIdx.x == 0) s_base = s_cnt ? atomicAdd(&w.ctr[0], s_cnt) : 0u;
    __syncthreads();
    if (inside) w.q0[s_base + at] = slot;
}

// The walks of one turn: closest hits of the queued rays (ANY false), or the
// shadow rays and what each turn left (ANY true). Persistent: each wave takes
// rays from the queue's head until it is drained.
template <bool ANY>
__global__ __launch_bounds__(WF_THREADS) __attribute__((amdgpu_waves_per_eu(5))) void k_wf_walk(SceneDev sc, const PhrCamera* __restrict__ camp, WfDev w, uint32_t parity,
                                                        uint32_t tiles_x, uint32_t* __restrict__ px,
                                                        unsigned long long* __restrict__ rays) {
    __shared__ uint32_t s_stack[PHR_STACK * WF_THREADS];
    uint32_t* stk = s_stack + threadIdx.x;
    uint32_t* ctr = w.ctr;
    if (!ANY && blockIdx.x == 0 && threadIdx.x == 0) {  // for this turn's shade and shadow walk
        ctr[1u - parity] = 0u;
        ctr[3] = 0u;
        ctr[4] = 0u;
    }
    const uint32_t count = ANY ? ctr[3] : ctr[parity];
    uint32_t* head = ANY ? &ctr[4] : &ctr[2];
    const uint32_t* queue = ANY ? w.sq : (parity ? w.q1 : w.q0);
    const uint32_t lane = threadIdx.x & 63u;
    PhrHit best{PHR_INF, PHR_NONE, 0.f, 0.f, 0u};
    V3 o = v3(0, 0, 0), inv = o;
    RaySpace rs{};
    uint32_t cur = 0, end = 0, sp = 0, slot = 0;
    uint32_t steps = 0, tests = 0, nrays = 0;
    bool live = false, drained = count == 0u;
    uint32_t cbase = 0, cend = 0;  // the wave's chunk of the queue
    const float4 rlo = sc.nodes[2 * sc.root], rhi = sc.nodes[2 * sc.root + 1];
    // A finished walk's result: the hit for the shading, or the shadow ray's
    // visibility taken into the field its turn left.
    // A finished walk only stores: the hit for the shading, or the shadow
    // ray's visibility, which the slot's next shading (or the frame's last
    // kernel) takes in.
    auto finish = [&]() {
        if constexpr (!ANY) w.hit[slot] = make_float4(best.t, __uint_as_float(best.tri), best.u, best.v);
        else w.vis[slot] = best.tri != PHR_NONE ? 0u : 1u;
    };
    for (;;) {
        const uint64_t idle = __ballot(!live);
        const uint32_t nidle = (uint32_t)__popcll(idle);
        if (!drained && nidle >= 32u && cbase >= cend) {  // the wave's chunk is used up: take the next
            uint32_t b = 0;
            if (lane == 0) b = atomicAdd(head, WF_CHUNK);
            b = __shfl(b, 0);
            if (b >= count) {
                drained = true;
            } else {
                cbase = b;
                cend = min(b + WF_CHUNK, count);
            }
        }
        if (nidle >= 32u && cbase < cend) {  // refill the idle lanes from the chunk
            const uint32_t take = min(nidle, cend - cbase);
            const uint32_t r = lanes_below(idle);
            const uint32_t my = cbase + r;
            cbase += take;
            if (!live && r < take) {
                slot = queue[my];
                V3 d;
                float tmax;
                if constexpr (ANY) {
                    const float4 a = w.sh[4 * (size_t)slot], b = w.sh[4 * (size_t)slot + 1];
                    o = v3(a.x, a.y, a.z);
                    tmax = a.w;
                    d = v3(b.x, b.y, b.z);
                } else {
                    const float4 a = w.s0[slot], b = w.s1[slot];
                    o = v3(a.x, a.y, a.z);
                    d = v3(b.x, b.y, b.z);
                    tmax = PHR_INF;
                }
                ++nrays;
                best = PhrHit{tmax, PHR_NONE, 0.f, 0.f, 0u};
                inv = safe_inv(d);
                if (box_t(rlo, rhi, o, inv, best.t) >= PHR_INF) {
                    finish();
                } else {
                    rs = ray_space(d);
                    sp = 0;
                    cur = __float_as_uint(rlo.w);
                    end = __float_as_uint(rhi.w);
                    if (!(cur & PHR_LEAF)) cur = sc.root;
                    live = true;
                }
            }
        }
        if (__ballot(live) == 0ull) {
            if (drained) break;
            continue;
        }
        if (!live) continue;
        // One visit of walk<ANY>: a leaf's triangles or an inner node, then a
        // pop if nothing was entered.
        bool more = true, popped = true;
        if (cur & PHR_LEAF) {
            uint32_t id = cur & ~PHR_LEAF;
            do {
                ++tests;
                const float4 t0 = sc.tris[3 * id], t1 = sc.tris[3 * id + 1], t2 = sc.tris[3 * id + 2];
                if (end == PHR_NONE) end = __float_as_uint(t2.y);
                float u, v;
                const float t = tri_t(t0, t1, t2, o, rs, u, v);
                if (t < best.t) {
                    if (ANY) {
                        const float* m = sc.mats + MAT_WORDS * (size_t)sc.faces[id].w;
                        if (m[6] > 0.5f || m[3] < 0.5f) continue;
                        best = PhrHit{t, id, u, v, 0u};
                        more = false;
                        break;
                    }
                    best = PhrHit{t, id, u, v, 0u};
                }
            } while (++id < end);
        } else {
            popped = !visit2(sc.recs + 4 * (size_t)cur, sc.kids + 8 * (size_t)cur, o, inv, best.t, cur, end, stk, WF_THREADS, sp, steps);
        }
        if (more && popped) {
            if (sp == 0) {
                more = false;
            } else {
                --sp;
                cur = stk[sp * WF_THREADS];
                end = PHR_NONE;
            }
        }
        if (!more) {
            finish();
            live = false;
        }
    }
    unsigned long long rc = ANY ? 0ull : nrays, rsh = ANY ? nrays : 0ull, st_ = steps, tt = tests;
    for (int k = 32; k > 0; k >>= 1) {
        rc += __shfl_xor(rc, k);
        rsh += __shfl_xor(rsh, k);
        st_ += __shfl_xor(st_, k);
        tt += __shfl_xor(tt, k);
    }
    if (lane == 0 && (rc | rsh)) {
        atomicAdd(&rays[0], rc);
        atomicAdd(&rays[1], rsh);
        atomicAdd(&rays[2], st_);
        atomicAdd(&rays[3], tt);
    }
}

// One segment of every queued slot (rg_hit, rg_post), its next ray and its
// shadow ray appended to the queues; an ended sample adds its radiance and
// its pixel starts the next one. A block takes WF_SHADE_ITEMS x 256
// consecutive entries and gathers its appends in LDS (wave ballots on LDS
// counters), then reserves its places in each queue with one atomic.
__global__ __launch_bounds__(WF_THREADS) void k_wf_shade(SceneDev sc, const PhrCamera* __restrict__ camp, WfDev w, uint32_t parity,
                                                         uint32_t tiles_x, uint32_t frame0, uint32_t spp, uint32_t* __restrict__ px) {
    __shared__ uint32_t s_next[WF_SHADE_ITEMS * WF_THREADS];
    __shared__ uint32_t s_shadow[WF_SHADE_ITEMS * WF_THREADS];
    __shared__ uint32_t s_cnt[2], s_base[2];
    const PhrCamera* p = camp;
    const size_t n = (size_t)p->w * p->h;
    uint32_t* ctr = w.ctr;
    if (blockIdx.x == 0 && threadIdx.x == 0) ctr[2] = 0u;  // the next turn's extend head
    const uint32_t count = ctr[parity];
    const uint32_t first_e = blockIdx.x * WF_SHADE_ITEMS * WF_THREADS;
    if (first_e >= count) return;
    const uint32_t* queue = parity ? w.q1 : w.q0;
    uint32_t* next_q = parity ? w.q0 : w.q1;
    const RgConst k = rg_const(p);
    if (threadIdx.x < 2) s_cnt[threadIdx.x] = 0u;
    __syncthreads();
    for (uint32_t r = 0; r < WF_SHADE_ITEMS; ++r) {
        const uint32_t e = first_e + r * WF_THREADS + threadIdx.x;
        bool next = false, shadow = false;
        uint32_t slot = 0;
        if (e < count) {
            slot = queue[e];
            bool inside;
            const uint32_t i = wf_pixel(slot, tiles_x, p->w, p->h, inside);
            RgSample st;
            uint32_t s;
            wf_load(w, slot, st, s);
            const float4 hr = w.hit[slot];
            const PhrHit h{hr.x, __float_as_uint(hr.y), hr.z, hr.w, 0u};
            if (const uint32_t pk = w.pend[slot]) {  // the last turn's shadow ray, taken in first
                const float4 c = w.sh[4 * (size_t)slot + (w.vis[slot] ? 3 : 2)];
                const V3 v = v3(c.x, c.y, c.z);
                if (pk - 1u == WF_END) add3(px, n + 3 * (size_t)i, v);
                else if (pk - 1u == WF_CC) st.cc = v;
                else st.acc = v;
            }
            const bool first = st.phase == 1u && st.pass == 0u;
            RgGuide g{v3(0, 0, 0), v3(0, 0, 0), 0.f, 0.f, 0.f, 0u};
            RgTurn q;
            rg_hit(st, g, q, sc, p, k, h);
            if (first) {  // the sample's guides, in sample order
                add3(px, 7 * n + 3 * (size_t)i, g.alb);
                add3(px, 10 * n + 3 * (size_t)i, g.nrm);
                stf(px, 13 * n + i, ldf(px, 13 * n + i) + g.dep);
                stf(px, 33 * n + i, ldf(px, 33 * n + i) + g.missed);
                stf(px, 23 * n + i, ldf(px, 23 * n + i) + g.cov);
                px[25 * n + i] = g.outline;
            }
            bool ended;
            uint32_t pk_new = 0u;
            if (!q.qon) {
                ended = rg_post(st, q, p, 1.0f);
                if (ended) add3(px, n + 3 * (size_t)i, st.col);
            } else {
                RgPost blocked{st.col, st.cc, st.acc, st.phase};
                ended = rg_post_of(blocked, st.alb, st.p_diffuse, q, p, 0.0f);
                rg_post(st, q, p, 1.0f);
                const uint32_t kind = ended ? WF_END : (q.to_bounces ? WF_CC : WF_ACC);
                pk_new = 1u + kind;
                float4 c0, c1;
                if (ended) {
                    c0 = make_float4(blocked.col.x, blocked.col.y, blocked.col.z, __uint_as_float(kind));
                    c1 = make_float4(st.col.x, st.col.y, st.col.z, 0.f);
                } else if (kind == WF_CC) {
                    c0 = make_float4(blocked.cc.x, blocked.cc.y, blocked.cc.z, __uint_as_float(kind));
                    c1 = make_float4(st.cc.x, st.cc.y, st.cc.z, 0.f);
                } else {
                    c0 = make_float4(blocked.acc.x, blocked.acc.y, blocked.acc.z, __uint_as_float(kind));
                    c1 = make_float4(st.acc.x, st.acc.y, st.acc.z, 0.f);
                }
                float4* d = w.sh + 4 * (size_t)slot;
                d[0] = make_float4(q.qpos.x, q.qpos.y, q.qpos.z, q.qdist);
                d[1] = make_float4(q.qdir.x, q.qdir.y, q.qdir.z, 0.f);
                d[2] = c0;
                d[3] = c1;
                shadow = true;
            }
            w.pend[slot] = (uint8_t)pk_new;
            if (ended) {
                if (s + 1u < spp) {
                    ++s;
                    rg_start(st, g, p, i, frame0 + s);
                    next = true;
                }
            } else {
                next = true;
            }
            wf_store(w, slot, st, s);
        }
        const uint32_t at = wave_append(next, &s_cnt[0]);
        if (next) s_next[at] = slot;
        const uint32_t sa = wave_append(shadow, &s_cnt[1]);
        if (shadow) s_shadow[sa] = slot;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        s_base[0] = s_cnt[0] ? atomicAdd(&ctr[1u - parity], s_cnt[0]) : 0u;
        s_base[1] = s_cnt[1] ? atomicAdd(&ctr[3], s_cnt[1]) : 0u;
    }
    __syncthreads();
    for (uint32_t k2 = threadIdx.x; k2 < s_cnt[0]; k2 += WF_THREADS) next_q[s_base[0] + k2] = s_next[k2];
    for (uint32_t k2 = threadIdx.x; k2 < s_cnt[1]; k2 += WF_THREADS) w.sq[s_base[1] + k2] = s_shadow[k2];
}

// The frame's last words a pixel: the last sample's radiance if a shadow ray
// still held it, the grid plane, film weight and count (added once a sample,
// as path_block adds them), the display.
__global__ __launch_bounds__(256) void k_wf_final(const PhrCamera* __restrict__ camp, WfDev w, uint32_t tiles_x, uint32_t frame0,
                                                  uint32_t spp, uint32_t* __restrict__ px) {
    const PhrCamera* p = camp;
    const size_t n = (size_t)p->w * p->h;
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    // The last sample's radiance when its turn ended with a shadow ray.
    const uint32_t x = i % p->w, y = i / p->w;
    const uint32_t slot = ((y >> 4) * tiles_x + (x >> 4)) * 256u + (y & 15u) * 16u + (x & 15u);
    if (w.pend[slot] == 1u + WF_END) {
        const float4 c = w.sh[4 * (size_t)slot + (w.vis[slot] ? 3 : 2)];
        add3(px, n + 3 * (size_t)i, v3(c.x, c.y, c.z));
    }
    const bool keep = frame0 > 0u;
    const V3 l = v3(ldf(px, n + 3 * i), ldf(px, n + 3 * i + 1), ldf(px, n + 3 * i + 2));
    const float cov = ldf(px, 23 * n + i), gone = ldf(px, 33 * n + i);
    V3 gs = keep ? v3(ldf(px, 20 * n + 3 * i), ldf(px, 20 * n + 3 * i + 1), ldf(px, 20 * n + 3 * i + 2)) : v3(0, 0, 0);
    float film = keep ? ldf(px, 24 * n + i) : 0.f;
    uint32_t taken = keep ? px[31 * n + i] : 0u;
    for (uint32_t kk = 0; kk < spp; ++kk) {
        gs = gs + v3(0.f, 0.f, 0.f);
        film = film + 1.0f;
    }
    taken = taken + spp;
    if (frame0 == 0u) { px[29 * n + i] = 0u; px[30 * n + i] = 0u; }
    stf(px, 20 * n + 3 * i, gs.x); stf(px, 20 * n + 3 * i + 1, gs.y); stf(px, 20 * n + 3 * i + 2, gs.z);
    stf(px, 24 * n + i, film);
    px[31 * n + i] = taken;
    put_shown(px, n, i, *p, l, gs, cov, gone, film, taken);
}

// One image: a block a tile, from tile0 on (a slice of the tiles).
// WPE: waves per SIMD the register allocation is fitted to (2 needs no spill;
// 3 to 5 trade spilled words for more waves in flight). The tracer's
// autotune times the variants on the caller's workload and keeps the fastest.
// MSTK: the mesh-tree stack's entries (a forest over mesh trees no deeper
// than that; the autotune may pick the short stack at 3 waves too).
template <bool F, int WPE, uint32_t MSTK = mesh_stack_entries<F, WPE>()>
__global__ __launch_bounds__(TRACE_THREADS) __attribute__((amdgpu_waves_per_eu(WPE))) void k_path(
    SceneDev sc, const PhrCamera* __restrict__ camp, uint32_t tiles_x, uint32_t frame0, uint32_t spp, uint32_t* __restrict__ px,
    unsigned long long* __restrict__ rays, uint32_t listed, uint32_t tile0) {
    __shared__ uint32_t s_stack[MSTK * TRACE_THREADS];
    __shared__ uint16_t s_tstack[top_stack_entries<F>() * TRACE_THREADS];
    path_block<F>(sc, camp, blockIdx.x + tile0, tiles_x, frame0, spp, px, rays, listed, s_stack, s_tstack);
}

// A batch of scenes in one launch: `tiles` blocks a scene, each scene with
// its own tree (root), tables, camera and image (px_stride words apart), all
// starting fresh. Workgroups go to the eight XCDs in turn (block b to XCD
// b mod 8); with `grouped` (scenes a multiple of 8) the blocks one XCD
// receives cover whole scenes, scene x + 8q on XCD x, so each XCD's L2 holds
// the trees of the scenes it traces rather than a share of every one.
template <int WPE>
__global__ __launch_bounds__(TRACE_THREADS) __attribute__((amdgpu_waves_per_eu(WPE))) void k_path_batch(
    const SceneDev* __restrict__ scenes, const PhrCamera* __restrict__ cams, uint32_t tiles, uint32_t tiles_x, uint32_t spp,
    uint32_t* __restrict__ px, uint64_t px_stride, unsigned long long* __restrict__ rays, uint32_t grouped) {
    __shared__ uint32_t s_stack[PHR_STACK * TRACE_THREADS];
    __shared__ uint16_t s_tstack[1];
    uint32_t scene, tile;
    if (grouped) {
        const uint32_t x = blockIdx.x % 8u, q = blockIdx.x / 8u;
        scene = x + 8u * (q / tiles);
        tile = q % tiles;
    } else {
        scene = blockIdx.x / tiles;
        tile = blockIdx.x % tiles;
    }
    const SceneDev sc = scenes[scene];
    path_block<false>(sc, cams + scene, tile, tiles_x, 0u, spp, px + px_stride * scene, rays, 0u, s_stack, s_tstack);
}
// The same batch with the split build's samples (path_split): for cameras
// with lamps or a lit world and no Adaptive Sampling.
template <int WPE>
__global__ __launch_bounds__(TRACE_THREADS) __attribute__((amdgpu_waves_per_eu(WPE))) void k_path_batch_sw(
    const SceneDev* __restrict__ scenes, const PhrCamera* __restrict__ cams, uint32_t tiles, uint32_t tiles_x, uint32_t spp,
    uint32_t* __restrict__ px, uint64_t px_stride, unsigned long long* __restrict__ rays, uint32_t grouped, uint32_t shade8) {
    __shared__ uint32_t s_stack[PHR_STACK * TRACE_THREADS];
    uint32_t scene, tile;
    if (grouped) {
        const uint32_t x = blockIdx.x % 8u, q = blockIdx.x / 8u;
        scene = x + 8u * (q / tiles);
        tile = q % tiles;
    } else {
        scene = blockIdx.x / tiles;
        tile = blockIdx.x % tiles;
    }
    const SceneDev sc = scenes[scene];
    path_split(sc, cams + scene, tile, tiles_x, 0u, spp, px + px_stride * scene, rays, s_stack, nullptr, shade8);
}

// The builds of the path kernels the tracer and the batch choose among.
constexpr uint32_t PATH_WAVES_MIN = 2, PATH_WAVES_MAX = 5;

// Whether k_path_rg serves a camera: no Adaptive Sampling (its half image and
// pixel list are path_block's) and no studio rig (whose three lamps a turn
// would not fit one shadow ray): a camera with lamps, or a lit world.
bool regen_allowed(const PhrCamera& cam) {
    const bool lit = cam.env[0] * cam.env[0] + cam.env[1] * cam.env[1] + cam.env[2] * cam.env[2] > 0.f;
    return cam.adapt[3] == 0.0f && (cam.count != 0u || lit);
}

// regen: 0 the nested build; 3 the split build (k_path_sw, shading at shade8
// eighths of a wave); else the regenerating build (k_path_rg); the last two
// for one tree, no list and no Adaptive Sampling or studio rig (the caller
// checks the camera).
// short_stack: the forest's 3-wave build with FOREST_SHORT_STACK mesh-tree
// entries (the caller checked the mesh trees' depth).
void launch_path(uint32_t waves, bool forest, uint32_t regen, dim3 grid, hipStream_t s, const SceneDev& sc, const PhrCamera* cp,
                 uint32_t tiles_x, uint32_t f0, uint32_t spp, uint32_t* px, unsigned long long* rays, uint32_t listed,
                 uint32_t tile0, bool short_stack = false, uint32_t shade8 = 4) {
    const dim3 block(TRACE_THREADS);
    if (regen == 3u && !forest && !listed) {
        if (waves == 5)
            hipLaunchKernelGGL(k_path_sw<5>, grid, block, 0, s, sc, cp, tiles_x, f0, spp, px, rays, tile0, shade8);
        else if (waves == 4)
            hipLaunchKernelGGL(k_path_sw<4>, grid, block, 0, s, sc, cp, tiles_x, f0, spp, px, rays, tile0, shade8);
        else if (waves == 3)
            hipLaunchKernelGGL(k_path_sw<3>, grid, block, 0, s, sc, cp, tiles_x, f0, spp, px, rays, tile0, shade8);
        else
            hipLaunchKernelGGL(k_path_sw<2>, grid, block, 0, s, sc, cp, tiles_x, f0, spp, px, rays, tile0, shade8);
    } else if (regen == 3u && forest && !listed) {  // the caller checked the mesh trees' depth for the short stacks
        if (waves >= 4)
            hipLaunchKernelGGL((k_path_sw_forest<4, FOREST_SHORT_STACK>), grid, block, 0, s, sc, cp, tiles_x, f0, spp, px, rays, tile0, shade8);
        else if (waves == 3 && short_stack)
            hipLaunchKernelGGL((k_path_sw_forest<3, FOREST_SHORT_STACK>), grid, block, 0, s, sc, cp, tiles_x, f0, spp, px, rays, tile0, shade8);
        else if (waves == 3)
            hipLaunchKernelGGL((k_path_sw_forest<3, PHR_STACK>), grid, block, 0, s, sc, cp, tiles_x, f0, spp, px, rays, tile0, shade8);
        else
            hipLaunchKernelGGL((k_path_sw_forest<2, PHR_STACK>), grid, block, 0, s, sc, cp, tiles_x, f0, spp, px, rays, tile0, shade8);
    } else if (regen && !forest && !listed) {
        if (waves == 5)
            hipLaunchKernelGGL(k_path_rg<5>, grid, block, 0, s, sc, cp, tiles_x, f0, spp, px, rays, tile0);
        else if (waves == 4)
            hipLaunchKernelGGL(k_path_rg<4>, grid, block, 0, s, sc, cp, tiles_x, f0, spp, px, rays, tile0);
        else if (waves == 3)
            hipLaunchKernelGGL(k_path_rg<3>, grid, block, 0, s, sc, cp, tiles_x, f0, spp, px, rays, tile0);
        else
            hipLaunchKernelGGL(k_path_rg<2>, grid, block, 0, s, sc, cp, tiles_x, f0, spp, px, rays, tile0);
    } else if (forest && waves >= 4)  // the caller checked the mesh trees' depth
        hipLaunchKernelGGL((k_path<true, 4>), grid, block, 0, s, sc, cp, tiles_x, f0, spp, px, rays, listed, tile0);
    else if (forest && waves == 3 && short_stack)
        hipLaunchKernelGGL((k_path<true, 3, FOREST_SHORT_STACK>), grid, block, 0, s, sc, cp, tiles_x, f0, spp, px, rays, listed, tile0);
    else if (forest && waves == 3)  // 32 + 32 16-bit entries (192 B a thread): three blocks a CU
        hipLaunchKernelGGL((k_path<true, 3>), grid, block, 0, s, sc, cp, tiles_x, f0, spp, px, rays, listed, tile0);
    else if (forest)
        hipLaunchKernelGGL((k_path<true, 2>), grid, block, 0, s, sc, cp, tiles_x, f0, spp, px, rays, listed, tile0);
    else if (waves == 5)
        hipLaunchKernelGGL((k_path<false, 5>), grid, block, 0, s, sc, cp, tiles_x, f0, spp, px, rays, listed, tile0);
    else if (waves == 4)
        hipLaunchKernelGGL((k_path<false, 4>), grid, block, 0, s, sc, cp, tiles_x, f0, spp, px, rays, listed, tile0);
    else if (waves == 3)
        hipLaunchKernelGGL((k_path<false, 3>), grid, block, 0, s, sc, cp, tiles_x, f0, spp, px, rays, listed, tile0);
    else
        hipLaunchKernelGGL((k_path<false, 2>), grid, block, 0, s, sc, cp, tiles_x, f0, spp, px, rays, listed, tile0);
}

// split: the split build (k_path_batch_sw, shading at shade8 eighths), else
// the nested one.
void launch_path_batch(uint32_t waves, bool split, dim3 grid, hipStream_t s, const SceneDev* table, const PhrCamera* cams,
                       uint32_t tiles, uint32_t tiles_x, uint32_t spp, uint32_t* px, uint64_t stride, unsigned long long* rays,
                       uint32_t grouped, uint32_t shade8) {
    const dim3 block(TRACE_THREADS);
    if (split) {
        if (waves == 5)
            hipLaunchKernelGGL(k_path_batch_sw<5>, grid, block, 0, s, table, cams, tiles, tiles_x, spp, px, stride, rays, grouped, shade8);
        else if (waves == 4)
            hipLaunchKernelGGL(k_path_batch_sw<4>, grid, block, 0, s, table, cams, tiles, tiles_x, spp, px, stride, rays, grouped, shade8);
        else if (waves == 3)
            hipLaunchKernelGGL(k_path_batch_sw<3>, grid, block, 0, s, table, cams, tiles, tiles_x, spp, px, stride, rays, grouped, shade8);
        else
            hipLaunchKernelGGL(k_path_batch_sw<2>, grid, block, 0, s, table, cams, tiles, tiles_x, spp, px, stride, rays, grouped, shade8);
    } else if (waves == 5)
        hipLaunchKernelGGL(k_path_batch<5>, grid, block, 0, s, table, cams, tiles, tiles_x, spp, px, stride, rays, grouped);
    else if (waves == 4)
        hipLaunchKernelGGL(k_path_batch<4>, grid, block, 0, s, table, cams, tiles, tiles_x, spp, px, stride, rays, grouped);
    else if (waves == 3)
        hipLaunchKernelGGL(k_path_batch<3>, grid, block, 0, s, table, cams, tiles, tiles_x, spp, px, stride, rays, grouped);
    else
        hipLaunchKernelGGL(k_path_batch<2>, grid, block, 0, s, table, cams, tiles, tiles_x, spp, px, stride, rays, grouped);
}

// ---- Adaptive Sampling's tail (src/trace/wgsl/adapt.wgsl, gather.wgsl) ----
// Stage 0, Cycles' convergence check at a step boundary: the full and half
// images' disagreement over the pixel's own count, normalised by intensity
// (its square root below 1); under the threshold the pixel settles.
__global__ __launch_bounds__(256) void k_adapt_check(uint32_t* __restrict__ px, uint32_t n, float threshold) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const size_t N = n;
    if (px[29 * N + i] != 0u) return;
    const size_t a = N + 3 * (size_t)i, x = 26 * N + 3 * (size_t)i;
    const uint32_t taken = max(px[31 * N + i], 1u);
    const float scale = 1.0f / (float)taken;
    const float f0 = __uint_as_float(px[a]), f1 = __uint_as_float(px[a + 1]), f2 = __uint_as_float(px[a + 2]);
    const float d0 = fabsf(f0 - __uint_as_float(px[x])), d1 = fabsf(f1 - __uint_as_float(px[x + 1])),
                d2 = fabsf(f2 - __uint_as_float(px[x + 2]));
    const float err_diff = ((d0 + d1) + d2) * scale;
    const float intensity = ((f0 + f1) + f2) * scale;
    const float norm = intensity < 1.0f ? sqrtf(intensity) : intensity;
    if (err_diff / (0.0001f + norm) < threshold) px[29 * N + i] = 1u;
}

// Stage 1: settled only where the pixel and its four neighbours are.
__global__ __launch_bounds__(256) void k_adapt_dilate(uint32_t* __restrict__ px, uint32_t w, uint32_t h) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    const size_t N = (size_t)w * h;
    if (i >= N) return;
    const uint32_t x = i % w, y = i / w;
    const uint32_t* f = px + 29 * N;
    uint32_t all = f[i];
    if (x > 0u) all = min(all, f[i - 1]);
    if (x + 1u < w) all = min(all, f[i + 1]);
    if (y > 0u) all = min(all, f[i - w]);
    if (y + 1u < h) all = min(all, f[i + w]);
    px[30 * N + i] = all;
}

// Stages 2 and 3: the list of pixels still open for the next launch, one run
// of the list a 16 x 16 tile, reserved with one atomic a block.
__global__ __launch_bounds__(256) void k_gather_clear(uint32_t* __restrict__ px, size_t n) {
    if (blockIdx.x == 0 && threadIdx.x == 0) px[34 * n] = 0u;
}
__global__ __launch_bounds__(256) void k_gather(uint32_t* __restrict__ px, uint32_t w, uint32_t h, uint32_t tiles_x) {
    __shared__ uint32_t held, run;
    const size_t N = (size_t)w * h;
    if (threadIdx.x == 0) held = 0u;
    __syncthreads();
    const uint32_t tx = blockIdx.x % tiles_x, ty = blockIdx.x / tiles_x;
    const uint32_t x = tx * 16 + (threadIdx.x & 15u), y = ty * 16 + (threadIdx.x >> 4);
    const uint32_t i = y * w + x;
    const bool open = x < w && y < h && px[30 * N + i] == 0u;
    uint32_t at = 0;
    if (open) at = atomicAdd(&held, 1u);
    __syncthreads();
    if (threadIdx.x == 0) run = atomicAdd(&px[34 * N], held);
    __syncthreads();
    if (open) px[32 * N + run + at] = i;
}

// Each mesh tree's root box (object space), for the instances' world boxes.
__global__ __launch_bounds__(256) void k_roots(const float4* __restrict__ nodes, const uint32_t* __restrict__ bases, uint32_t n,
                                               float4* __restrict__ out) {
    uint32_t m = blockIdx.x * 256 + threadIdx.x;
    if (m >= n) return;
    uint32_t b = bases[m];
    out[2 * m] = b == PHR_NONE ? make_float4(0.f, 0.f, 0.f, 0.f) : nodes[2 * b];
    out[2 * m + 1] = b == PHR_NONE ? make_float4(0.f, 0.f, 0.f, 0.f) : nodes[2 * b + 1];
}

// Top-tree slot -> instance id (Forest::retlas, src/trace/forest.rs:124-126).
__global__ __launch_bounds__(256) void k_inst_order(const uint32_t* __restrict__ placed, const uint32_t* __restrict__ tlas_order,
                                                    uint32_t p, uint32_t* __restrict__ out) {
    uint32_t k = blockIdx.x * 256 + threadIdx.x;
    if (k < p) out[k] = placed[tlas_order[k]];
}

// Placed instance j's world box: its mesh tree's root box, eight corners
// placed (instance.rs:116-124), read from the node array where the refit left it.
__global__ __launch_bounds__(256) void k_world_boxes(const float* __restrict__ worlds, const uint32_t* __restrict__ placed,
                                                     const uint32_t* __restrict__ inst_mesh, const uint32_t* __restrict__ bases,
                                                     const float4* __restrict__ nodes, uint32_t p, float4* __restrict__ wb) {
    const uint32_t j = blockIdx.x * 256 + threadIdx.x;
    if (j >= p) return;
    const uint32_t i = placed[j], r = bases[inst_mesh[i]];
    const float4 a = nodes[2 * (size_t)r], b = nodes[2 * (size_t)r + 1];
    const float* w = worlds + 16 * (size_t)i;
    float lo[3] = {3.40282347e38f, 3.40282347e38f, 3.40282347e38f}, hi[3] = {-3.40282347e38f, -3.40282347e38f, -3.40282347e38f};
    for (int k = 0; k < 8; ++k) {
        const float c[3] = {(k & 1) ? b.x : a.x, (k & 2) ? b.y : a.y, (k & 4) ? b.z : a.z};
        for (int q = 0; q < 3; ++q) {
            const float v = ((w[q] * c[0] + w[4 + q] * c[1]) + w[8 + q] * c[2]) + w[12 + q];
            lo[q] = fmin_(lo[q], v);
            hi[q] = fmax_(hi[q], v);
        }
    }
    wb[2 * (size_t)j] = make_float4(lo[0], lo[1], lo[2], 0.f);
    wb[2 * (size_t)j + 1] = make_float4(hi[0], hi[1], hi[2], 0.f);
}

// Each candidate emitter's power: its world triangle's area times its glow
// (lamp.rs:51-97 over (face, instance)).
__global__ __launch_bounds__(256) void k_light_power(const uint32_t* __restrict__ cand, uint32_t ncand,
                                                     const uint4* __restrict__ faces, const float* __restrict__ verts,
                                                     const float* __restrict__ worlds, float* __restrict__ pw) {
    const uint32_t c = blockIdx.x * 256 + threadIdx.x;
    if (c >= ncand) return;
    const uint4 fc = faces[cand[3 * c]];
    const float* w = worlds + 16 * (size_t)cand[3 * c + 1];
    const float glow = __uint_as_float(cand[3 * c + 2]);
    float v[3][3];
    const uint32_t ids[3] = {fc.x, fc.y, fc.z};
    for (int k = 0; k < 3; ++k) {
        const float* p = verts + 8 * (size_t)ids[k];
        for (int q = 0; q < 3; ++q) v[k][q] = ((w[q] * p[0] + w[4 + q] * p[1]) + w[8 + q] * p[2]) + w[12 + q];
    }
    const float e1[3] = {v[1][0] - v[0][0], v[1][1] - v[0][1], v[1][2] - v[0][2]};
    const float e2[3] = {v[2][0] - v[0][0], v[2][1] - v[0][1], v[2][2] - v[0][2]};
    const float cr[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
    float sq = -0.0f;
    for (int k = 0; k < 3; ++k) sq = sq + cr[k] * cr[k];
    pw[c] = (0.5f * sqrtf(sq)) * glow;
}

// The emitter table from the powers, in candidate order: the lit ones, then
// the power CDF. One thread, so the sums run in the host's order.
__global__ void k_light_table(const uint32_t* __restrict__ cand, const float* __restrict__ pw, uint32_t ncand,
                              uint32_t* __restrict__ words) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    uint32_t cnt = 0;
    float total = -0.0f;
    for (uint32_t c = 0; c < ncand; ++c)
        if (pw[c] > 0.f) {
            ++cnt;
            total = total + pw[c];
        }
    words[0] = cnt;
    uint32_t k = 0;
    float run = 0.0f;
    for (uint32_t c = 0; c < ncand; ++c) {
        if (!(pw[c] > 0.f)) continue;
        words[1 + k] = cand[3 * c];
        words[1 + cnt + k] = cand[3 * c + 1];
        run = run + pw[c];
        words[1 + 2 * cnt + k] = __float_as_uint(run / total);
        ++k;
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// Tracer (src/trace/mod.rs:50-245, build.rs, sample.rs, read.rs)
// ---------------------------------------------------------------------------
struct phr_bvh;
int phr_bvh_internal_build(phr_gpu* g, phr_array* prims, uint32_t leaf, phr_bvh** out);
int phr_bvh_internal_build_into(phr_gpu* g, phr_array* prims, uint32_t leaf, const uint32_t* seg_lo, uint32_t segs,
                                phr_array* nodes, const uint32_t* node_base, uint32_t leaf_flag, phr_bvh** out);
int phr_bvh_internal_rebuild_async(phr_bvh* b);
int phr_bvh_internal_read_depth(phr_bvh* b);
uint32_t phr_bvh_internal_segment_depth(const phr_bvh* b, uint32_t s);
int phr_bvh_internal_refit_async(phr_bvh* b);
phr_array* phr_bvh_internal_nodes(phr_bvh* b);
phr_array* phr_bvh_internal_order(phr_bvh* b);
const uint8_t* phr_bvh_internal_inner(phr_bvh* b);

// The walk records of a built tree (after its gather wrote the triangle
// cache), into `recs`, allocated on first use.
static int enqueue_walk_records(phr_gpu* g, phr_bvh* bvh, phr_array** recs, phr_array* tris) {
    phr_array* nodes = phr_bvh_internal_nodes(bvh);
    const uint64_t count = nodes->len;
    if (!*recs) {
        int rc = phr_array_alloc(g, 16 * REC_ROWS, std::max<uint64_t>(count, 1), recs);
        if (rc != PHR_OK) return rc;
    }
    if (count)
        hipLaunchKernelGGL(k_walk_records, dim3((uint32_t)((count + 255) / 256)), dim3(256), 0, g->stream, (const float4*)nodes->ptr,
                           phr_bvh_internal_inner(bvh), (uint32_t)count, (float4*)(*recs)->ptr, (float4*)tris->ptr);
    PHR_HIP(hipGetLastError());
    return PHR_OK;
}
extern "C" void phr_bvh_free(phr_bvh* b);
extern "C" int phr_bvh_refit_async(phr_bvh* b);
extern "C" uint32_t phr_bvh_depth(const phr_bvh* b);
int phr_denoise_enqueue(phr_gpu* g, uint32_t* px, uint32_t w, uint32_t h, int keep, float* scratch, const float view[4]);
inline float4 view4(const PhrCamera& c) { return make_float4(c.view[0], c.view[1], c.view[2], c.view[3]); }
constexpr uint64_t DENOISE_SCRATCH_WORDS = 11;  // guides (8) and the first pass's input (3)

struct phr_tracer {
    phr_gpu* gpu = nullptr;
    uint32_t n = 0, nverts = 0, nmats = 0, leaf = 4;
    float extent = 0.f;
    uint32_t glow = 0;
    phr_array* in_faces = nullptr;  // uint4 per input face
    phr_array* verts = nullptr;     // 2 float4 per vertex
    phr_array* mats = nullptr;      // MAT_WORDS floats per material
    phr_array* prims = nullptr;     // Aabb per input face
    phr_array* faces = nullptr;     // uint4 per slot
    phr_array* tris = nullptr;      // 3 float4 per slot
    phr_array* recs = nullptr;      // one tree's walk records: 64 B a node index
    phr_array* lights = nullptr;    // [count, ids, insts, cdf]
    phr_array* px = nullptr;        // PX_LEN * w * h + 1 words
    unsigned long long* rays = nullptr;  // closest, shadow
    phr_bvh* bvh = nullptr;
    std::vector<uint32_t> corner_vertex;  // input corner (3f + k) -> vertex id
    std::vector<uint32_t> order;          // slot -> input face
    uint32_t w = 0, h = 0, frame = 0, samples = 32;
    uint32_t path_waves = 2;  // which k_path variant runs (2 to 5 waves a SIMD)
    uint32_t path_regen = 0;  // 1: the path-regenerating build, 2: the wavefront build, 3: the split build, where the camera allows it
    uint32_t split_shade = 6;  // the split build shades once this many eighths of a wave's working lanes wait
    uint32_t mesh_depth = PHR_STACK;  // a forest: its deepest mesh tree (the 4-wave build needs <= 24)
    bool short_stack = false;         // the forest's 3-wave build with the 24-entry mesh-tree stack
    // The binding split phiray's WGSL needs (items a binding for nodes, faces,
    // vertices; the atlas offset in the data table) and bindings an array.
    // The kernels here index whole arrays; the numbers are kept for callers.
    uint32_t split[4] = {0, 0, 0, 0};
    uint32_t chunks = 2;
    // Tracer::set_adaptive (src/trace/mod.rs:235-240): stamped onto every
    // camera once set; until then the camera's own `adapt` words are used.
    float adapt[4] = {0.f, 0.f, 0.f, 0.f};
    bool adapt_set = false;
    bool have_last = false;
    PhrCamera last{};
    PhrCamera* cam_dev = nullptr;  // the stamped camera the path kernel reads
    double last_gpu_s = 0.0, last_submit_s = 0.0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // A forest (src/trace/forest.rs): the top tree over the instances' world
    // boxes at node 0, each mesh's tree behind it; what a new pose keeps
    // (src/trace/pose.rs:28-39) on the host.
    bool forest = false;
    uint32_t nmeshes = 0, ninsts = 0, tlas_slots = 0;
    phr_array* nodes = nullptr;       // the whole node array
    phr_array* wboxes = nullptr;      // the placed instances' world boxes
    phr_array* insts = nullptr;       // instance words, top-tree slot order, draw order
    phr_array* placed_dev = nullptr;  // placed instance ids (a top-tree box's owner)
    phr_bvh* tlas = nullptr;
    std::vector<float> pos;            // 8 floats a vertex, object space
    std::vector<uint32_t> faces_in;    // 4 words a face, the scene's order
    std::vector<uint32_t> faces_slot;  // 4 words a face, leaf order
    std::vector<float> mats_h;
    std::vector<phr_mesh> meshes;
    std::vector<phr_instance> inst_in;
    std::vector<uint32_t> glow_meshes;  // meshes holding an emissive face
    uint32_t* bases_dev = nullptr;      // each mesh tree's root node, or NONE
    std::vector<uint32_t> placed, node_base;
    float* frame_dev = nullptr;  // the learned denoiser's input and output, sized for w x h
    float* dn_scratch = nullptr;  // the a-trous passes' guide table, DENOISE_SCRATCH_WORDS a pixel
    uint64_t dn_px = 0;
    uint64_t frame_px = 0;
    // A new pose is applied on the device without waiting: the instances'
    // matrices, each instance's mesh, the emitter candidates (face slot,
    // instance, glow) and their powers live there; the extent and emitter
    // count the camera needs are read back at the next sample (`stale`).
    phr_array* worlds_dev = nullptr;  // 16 floats an instance
    phr_array* inst_mesh = nullptr;   // mesh id an instance
    phr_array* lcand = nullptr;       // 3 words a candidate emitter (face slot, instance, glow bits)
    phr_array* lpow = nullptr;        // its power
    uint32_t ncand = 0;
    bool stale = false;
    std::vector<uint32_t> words_h;    // instance words staged for upload
    std::vector<float> inst_worlds_h; // matrices staged for upload
    // One recorded frame (phr_tracer_record_frame): its graph, the camera, and
    // every device address it baked in, to refuse a replay after any moved.
    hipGraph_t fgraph = nullptr;
    hipGraphExec_t fexec = nullptr;
    PhrCamera fcam{};
    uint32_t fw = 0, fh = 0, fspp = 0;
    SceneDev fsc{};
    const void* fsig[4] = {nullptr, nullptr, nullptr, nullptr};
    int fdenoise = 0;  // the recorded frame's denoiser: 0 none, 1 a-trous, 2 learned
    // A recorded posed frame: the pose's copies read these pinned buffers, so
    // a replay takes new matrices without recording again.
    float* pin_worlds = nullptr;      // 16 floats an instance
    uint32_t* pin_words = nullptr;    // the instance table as uploaded
    bool fposed = false;              // the recorded frame applies a pose first
    const void* fverts = nullptr;     // its vertex source (null: matrices only)
    // The wavefront build's buffers, sized for wf_slots path slots.
    phr_array* wf_state = nullptr;    // 9 float4 a slot (state, hit)
    phr_array* wf_queue = nullptr;    // 2 ray queues and the shadow queue, of slots
    phr_array* wf_shadow = nullptr;   // 4 float4 a slot: its shadow ray
    uint32_t* wf_ctr = nullptr;       // WF_CTRS counters
    uint32_t* wf_pin = nullptr;       // pinned: the next turn's ray count, read every few turns
    uint32_t wf_slots = 0;
    uint32_t wf_turns = 0;            // turns the last wavefront frame took
};

namespace {

float signum(float x) {  // Rust f32::signum
    if (x != x) return x;
    return std::signbit(x) ? -1.0f : 1.0f;
}
void oct_pack(const float n[3], float out[2]) {  // src/trace/scene.rs:22-38
    if (n[0] == 0.f && n[1] == 0.f && n[2] == 0.f) { out[0] = 2.f; out[1] = 2.f; return; }
    float l = std::fabs(n[0]) + std::fabs(n[1]) + std::fabs(n[2]);
    if (l < 1e-12f) { out[0] = 0.f; out[1] = 0.f; return; }
    float x = n[0] / l, y = n[1] / l, z = n[2] / l;
    if (z >= 0.f) { out[0] = x; out[1] = y; }
    else { out[0] = (1.0f - std::fabs(y)) * signum(x); out[1] = (1.0f - std::fabs(x)) * signum(y); }
}
inline uint32_t fbits(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }

template <size_t N>
struct WordsHash {  // FNV-1a over the words
    size_t operator()(const std::array<uint32_t, N>& a) const {
        uint64_t h = 0xcbf29ce484222325ull;
        for (uint32_t w : a) { h ^= w; h *= 0x100000001b3ull; }
        return (size_t)h;
    }
};

// The build the path launches take: a forest's 4-wave build only over mesh
// trees of at most FOREST_SHORT_STACK levels, else its 3-wave one.
uint32_t path_build_waves(const phr_tracer* t) {
    if (t->forest && t->path_waves >= 4) return t->mesh_depth <= FOREST_SHORT_STACK ? 4u : 3u;
    return t->path_waves;
}

// Whether the forest's 3-wave build takes the short mesh-tree stack.
bool path_short_stack(const phr_tracer* t) { return t->forest && t->short_stack && t->mesh_depth <= FOREST_SHORT_STACK; }

// The deepest mesh tree of a forest, read back after a build.
int forest_mesh_depth(phr_tracer* t) {
    int rc = phr_bvh_internal_read_depth(t->bvh);
    if (rc != PHR_OK) return rc;
    uint32_t d = 0;
    for (uint32_t m = 0; m < t->nmeshes; ++m) d = std::max(d, phr_bvh_internal_segment_depth(t->bvh, m));
    t->mesh_depth = d;
    return PHR_OK;
}

void tracer_release(phr_tracer* t) {
    if (!t) return;
    if (t->bvh) phr_bvh_free(t->bvh);
    if (t->tlas) phr_bvh_free(t->tlas);
    for (phr_array* a : {t->in_faces, t->verts, t->mats, t->prims, t->faces, t->tris, t->recs, t->lights, t->px, t->nodes, t->wboxes,
                         t->insts, t->placed_dev, t->worlds_dev, t->inst_mesh, t->lcand, t->lpow, t->wf_state, t->wf_queue,
                         t->wf_shadow})
        phr_array_release(a);
    if (t->wf_ctr) (void)hipFree(t->wf_ctr);
    if (t->wf_pin) (void)hipHostFree(t->wf_pin);
    if (t->bases_dev) (void)hipFree(t->bases_dev);
    if (t->frame_dev) (void)hipFree(t->frame_dev);
    if (t->dn_scratch) (void)hipFree(t->dn_scratch);
    if (t->fexec) (void)hipGraphExecDestroy(t->fexec);
    if (t->fgraph) (void)hipGraphDestroy(t->fgraph);
    if (t->pin_worlds) (void)hipHostFree(t->pin_worlds);
    if (t->pin_words) (void)hipHostFree(t->pin_words);
    if (t->rays) (void)hipFree(t->rays);
    if (t->cam_dev) (void)hipFree(t->cam_dev);
    if (t->ev0) (void)hipEventDestroy(t->ev0);
    if (t->ev1) (void)hipEventDestroy(t->ev1);
    delete t;
}

// The emitter table over slots [lo, hi) (src/trace/lamp.rs:51-97): slot s
// holds input face order[s]; faces of positive glow and area, their power
// CDF. Words: count, then the slots, the instances (0), the CDF bits.
std::vector<uint32_t> light_words(const uint32_t* order, uint32_t lo, uint32_t hi, const std::vector<uint32_t>& in_faces,
                                  const std::vector<float>& verts, const std::vector<float>& mats) {
    std::vector<uint32_t> ids;
    std::vector<float> pw;
    for (uint32_t s = lo; s < hi; ++s) {
        const uint32_t* fc = &in_faces[4 * (size_t)order[s]];
        const float* e = &mats[MAT_WORDS * (size_t)fc[3] + 8];
        float glow = 0.2126f * e[0] + 0.7152f * e[1] + 0.0722f * e[2];
        if (glow <= 0.f) continue;
        const float* a = &verts[8 * (size_t)fc[0]];
        const float* b = &verts[8 * (size_t)fc[1]];
        const float* c = &verts[8 * (size_t)fc[2]];
        float e1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
        float e2[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
        float cr[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
        float sq = -0.0f;
        for (int k = 0; k < 3; ++k) sq = sq + cr[k] * cr[k];
        float p = (0.5f * std::sqrt(sq)) * glow;
        if (p > 0.f) { ids.push_back(s); pw.push_back(p); }
    }
    float total = -0.0f;
    for (float p : pw) total = total + p;
    uint32_t m = (uint32_t)ids.size();
    std::vector<uint32_t> words(1 + 3 * (size_t)m);
    words[0] = m;
    float run = 0.0f;
    for (uint32_t k = 0; k < m; ++k) {
        words[1 + k] = ids[k];
        words[1 + m + k] = 0u;
        run = run + pw[k];
        words[1 + 2 * m + k] = fbits(run / total);
    }
    return words;
}

// A tracer's table: one identity instance, every slot.
int build_lights(phr_tracer* t, const std::vector<uint32_t>& in_faces, const std::vector<float>& verts,
                 const std::vector<float>& mats) {
    std::vector<uint32_t> words = light_words(t->order.data(), 0, t->n, in_faces, verts, mats);
    t->glow = words[0];
    if (t->lights) phr_array_release(t->lights);
    t->lights = nullptr;
    int rc = phr_array_alloc(t->gpu, 4, words.size(), &t->lights);
    if (rc != PHR_OK) return rc;
    PHR_HIP(hipMemcpyAsync(t->lights->ptr, words.data(), 4 * words.size(), hipMemcpyHostToDevice, t->gpu->stream));
    PHR_HIP(hipStreamSynchronize(t->gpu->stream));
    return PHR_OK;
}

SceneDev scene_dev(phr_tracer* t) {
    SceneDev s;
    s.nodes = (const float4*)(t->forest ? t->nodes : phr_bvh_internal_nodes(t->bvh))->ptr;
    s.recs = t->forest ? nullptr : (const float4*)t->recs->ptr;
    s.kids = t->forest ? nullptr : s.recs + 4 * t->recs->len;  // recs holds len nodes' records
    s.insts = t->forest ? (const uint32_t*)t->insts->ptr : nullptr;
    s.ninsts = t->ninsts;
    s.tris = (const float4*)t->tris->ptr;
    s.faces = (const uint4*)t->faces->ptr;
    s.verts = (const float4*)t->verts->ptr;
    s.mats = (const float*)t->mats->ptr;
    s.lights = (const uint32_t*)t->lights->ptr;
    s.root = 0;
    return s;
}

void launch_raycast(phr_tracer* t, const PhrRay* rays, uint32_t nrays, int any_hit, PhrHit* hits) {
    SceneDev sc = scene_dev(t);
    dim3 grid((nrays + TRACE_THREADS - 1) / TRACE_THREADS), block(TRACE_THREADS);
    if (t->forest)
        hipLaunchKernelGGL(k_raycast<true>, grid, block, 0, t->gpu->stream, sc, rays, nrays, any_hit ? 1u : 0u, hits, t->rays + 4);
    else
        hipLaunchKernelGGL(k_raycast<false>, grid, block, 0, t->gpu->stream, sc, rays, nrays, any_hit ? 1u : 0u, hits, t->rays + 4);
}

// ---- instances (src/trace/instance.rs:60-160) and meshes (mesh.rs:43-54) ----
// a(r, c) = m[c][r], the rows of the upper 3 x 3.
inline float wa(const float w[16], int r, int c) { return w[4 * c + r]; }

float inst_det(const float w[16]) {  // instance.rs:96-100
    return wa(w, 0, 0) * (wa(w, 1, 1) * wa(w, 2, 2) - wa(w, 1, 2) * wa(w, 2, 1)) -
           wa(w, 0, 1) * (wa(w, 1, 0) * wa(w, 2, 2) - wa(w, 1, 2) * wa(w, 2, 0)) +
           wa(w, 0, 2) * (wa(w, 1, 0) * wa(w, 2, 1) - wa(w, 1, 1) * wa(w, 2, 0));
}

void inst_to_world(const float w[16], const float p[3], float out[3]) {  // instance.rs:68-71
    for (int r = 0; r < 3; ++r) out[r] = ((w[r] * p[0] + w[4 + r] * p[1]) + w[8 + r] * p[2]) + w[12 + r];
}

void inst_inverse(const float w[16], float out[12]) {  // instance.rs:75-93; singular -> zero
    const float det = inst_det(w);
    if (std::fabs(det) < 1e-20f) {
        for (int k = 0; k < 12; ++k) out[k] = 0.f;
        return;
    }
    const float d = 1.0f / det;
    const float inv[3][3] = {
        {(wa(w, 1, 1) * wa(w, 2, 2) - wa(w, 1, 2) * wa(w, 2, 1)) * d, (wa(w, 0, 2) * wa(w, 2, 1) - wa(w, 0, 1) * wa(w, 2, 2)) * d,
         (wa(w, 0, 1) * wa(w, 1, 2) - wa(w, 0, 2) * wa(w, 1, 1)) * d},
        {(wa(w, 1, 2) * wa(w, 2, 0) - wa(w, 1, 0) * wa(w, 2, 2)) * d, (wa(w, 0, 0) * wa(w, 2, 2) - wa(w, 0, 2) * wa(w, 2, 0)) * d,
         (wa(w, 0, 2) * wa(w, 1, 0) - wa(w, 0, 0) * wa(w, 1, 2)) * d},
        {(wa(w, 1, 0) * wa(w, 2, 1) - wa(w, 1, 1) * wa(w, 2, 0)) * d, (wa(w, 0, 1) * wa(w, 2, 0) - wa(w, 0, 0) * wa(w, 2, 1)) * d,
         (wa(w, 0, 0) * wa(w, 1, 1) - wa(w, 0, 1) * wa(w, 1, 0)) * d},
    };
    const float tx = w[12], ty = w[13], tz = w[14];
    for (int r = 0; r < 3; ++r) {
        out[4 * r] = inv[r][0];
        out[4 * r + 1] = inv[r][1];
        out[4 * r + 2] = inv[r][2];
        out[4 * r + 3] = -((inv[r][0] * tx + inv[r][1] * ty) + inv[r][2] * tz);
    }
}

bool inst_identity(const float w[16]) {
    for (int c = 0; c < 4; ++c)
        for (int r = 0; r < 4; ++r)
            if (!(w[4 * c + r] == (r == c ? 1.0f : 0.0f))) return false;
    return true;
}

// Instance::words (instance.rs:126-160).
void inst_words(const phr_instance& in, uint32_t root, const phr_mesh& mesh, uint32_t* out) {
    for (uint32_t k = 0; k < INSTANCE_LEN; ++k) out[k] = 0u;
    const float* m = in.world;
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 4; ++c) out[r * 4 + c] = fbits(m[4 * c + r]);
    float inv[12];
    inst_inverse(m, inv);
    for (int k = 0; k < 12; ++k) out[12 + k] = fbits(inv[k]);
    out[24] = root;
    out[25] = mesh.face_lo;
    out[26] = mesh.face_hi;
    out[27] = (inst_identity(m) ? 1u : 0u) | ((in.casts_shadow ? 0u : 1u) << 1) | ((inst_det(m) < 0.0f ? 1u : 0u) << 2);
    for (int k = 0; k < 3; ++k) out[28 + k] = fbits(in.color[k]);
    out[31] = in.hash;
    for (int k = 0; k < 3; ++k) {  // Mesh::texspace (mesh.rs:43-54)
        float half = 0.5f * (mesh.hi[k] - mesh.lo[k]);
        const float centre = 0.5f * (mesh.hi[k] + mesh.lo[k]);
        half = half == 0.0f ? 1.0f : std::max(half, 0.00001f);
        out[32 + k] = fbits(centre - half);
        out[35 + k] = fbits(2.0f * half);
    }
    out[38] = in.mesh;
}

// lamp.rs:51-97 over the instances: each emissive face under each instance of
// its mesh, by placed area times luminance, sorted by face then instance.
// The emitter table on the host (src/trace/lamp.rs:51-97 over (face, instance)),
// for scenes with more candidate emitters than the one-thread table takes.
int forest_lights_host(phr_tracer* t) {
    for (uint32_t m : t->glow_meshes) {  // the emissive meshes' current object-space vertices
        const phr_mesh& me = t->meshes[m];
        if (me.vert_hi > me.vert_lo && me.vert_hi <= t->nverts)
            PHR_HIP(hipMemcpyAsync(&t->pos[8 * (size_t)me.vert_lo], (const float*)t->verts->ptr + 8 * (size_t)me.vert_lo,
                                   32ull * (me.vert_hi - me.vert_lo), hipMemcpyDeviceToHost, t->gpu->stream));
    }
    PHR_HIP(hipStreamSynchronize(t->gpu->stream));
    std::vector<std::vector<uint32_t>> by_mesh(t->nmeshes);
    for (uint32_t i = 0; i < t->ninsts; ++i)
        if (t->inst_in[i].mesh < t->nmeshes) by_mesh[t->inst_in[i].mesh].push_back(i);
    std::vector<uint32_t> ids, owners;
    std::vector<float> pw;
    for (uint32_t m = 0; m < t->nmeshes; ++m) {
        for (uint32_t f = t->meshes[m].face_lo; f < t->meshes[m].face_hi; ++f) {
            const uint32_t* fc = &t->faces_slot[4 * (size_t)f];
            const float* e = &t->mats_h[MAT_WORDS * (size_t)fc[3] + 8];
            const float glow = 0.2126f * e[0] + 0.7152f * e[1] + 0.0722f * e[2];
            if (glow <= 0.f) continue;
            for (uint32_t i : by_mesh[m]) {
                float a[3], b[3], c[3];
                inst_to_world(t->inst_in[i].world, &t->pos[8 * (size_t)fc[0]], a);
                inst_to_world(t->inst_in[i].world, &t->pos[8 * (size_t)fc[1]], b);
                inst_to_world(t->inst_in[i].world, &t->pos[8 * (size_t)fc[2]], c);
                float e1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
                float e2[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
                float cr[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
                float sq = -0.0f;
                for (int k = 0; k < 3; ++k) sq = sq + cr[k] * cr[k];
                const float p = (0.5f * std::sqrt(sq)) * glow;
                if (p > 0.f) {
                    ids.push_back(f);
                    owners.push_back(i);
                    pw.push_back(p);
                }
            }
        }
    }
    float total = -0.0f;
    for (float p : pw) total = total + p;
    const uint32_t cnt = (uint32_t)ids.size();
    std::vector<uint32_t> words(1 + 3 * (size_t)cnt);
    words[0] = cnt;
    float run = 0.0f;
    for (uint32_t k = 0; k < cnt; ++k) {
        words[1 + k] = ids[k];
        words[1 + cnt + k] = owners[k];
        run = run + pw[k];
        words[1 + 2 * cnt + k] = fbits(run / total);
    }
    t->glow = cnt;
    if (!t->lights || t->lights->cap < words.size()) {
        phr_array_release(t->lights);
        t->lights = nullptr;
        int rc = phr_array_alloc(t->gpu, 4, std::max<size_t>(words.size(), 4), &t->lights);
        if (rc != PHR_OK) return rc;
    }
    t->lights->len = words.size();
    PHR_HIP(hipMemcpyAsync(t->lights->ptr, words.data(), 4 * words.size(), hipMemcpyHostToDevice, t->gpu->stream));
    PHR_HIP(hipStreamSynchronize(t->gpu->stream));
    return PHR_OK;
}

constexpr uint32_t LIGHT_DEVICE_MAX = 4096;  // candidates the one-thread table kernel takes

// The candidate emitters in the table's order: meshes in order, each emissive
// face in slot order, each instance of its mesh in instance order. Built once
// per face order (a new tree order changes the slots).
int forest_light_cands(phr_tracer* t) {
    std::vector<std::vector<uint32_t>> by_mesh(t->nmeshes);
    for (uint32_t i = 0; i < t->ninsts; ++i)
        if (t->inst_in[i].mesh < t->nmeshes) by_mesh[t->inst_in[i].mesh].push_back(i);
    std::vector<uint32_t> cand;
    for (uint32_t m : t->glow_meshes)
        for (uint32_t f = t->meshes[m].face_lo; f < t->meshes[m].face_hi; ++f) {
            const float* e = &t->mats_h[MAT_WORDS * (size_t)t->faces_slot[4 * (size_t)f + 3] + 8];
            const float glow = 0.2126f * e[0] + 0.7152f * e[1] + 0.0722f * e[2];
            if (glow <= 0.f) continue;
            for (uint32_t i : by_mesh[m]) {
                cand.push_back(f);
                cand.push_back(i);
                cand.push_back(fbits(glow));
            }
        }
    t->ncand = (uint32_t)(cand.size() / 3);
    if (t->ncand == 0 || t->ncand > LIGHT_DEVICE_MAX) return PHR_OK;
    int rc;
    if (!t->lcand || t->lcand->cap < cand.size() || t->lpow->cap < t->ncand) {
        phr_array_release(t->lcand);
        phr_array_release(t->lpow);
        t->lcand = t->lpow = nullptr;
        if ((rc = phr_array_alloc(t->gpu, 4, cand.size(), &t->lcand)) != PHR_OK) return rc;
        if ((rc = phr_array_alloc(t->gpu, 4, t->ncand, &t->lpow)) != PHR_OK) return rc;
    }
    PHR_HIP(hipMemcpyAsync(t->lcand->ptr, cand.data(), 4 * cand.size(), hipMemcpyHostToDevice, t->gpu->stream));
    PHR_HIP(hipStreamSynchronize(t->gpu->stream));
    return PHR_OK;
}

// The emitter table for the current pose: on the device, without waiting,
// unless the scene has more candidates than the table kernel takes.
int forest_lights(phr_tracer* t) {
    if (t->ncand > LIGHT_DEVICE_MAX) return forest_lights_host(t);
    const size_t words = 1 + 3 * (size_t)t->ncand;
    if (!t->lights || t->lights->cap < words) {
        phr_array_release(t->lights);
        t->lights = nullptr;
        int rc = phr_array_alloc(t->gpu, 4, std::max<size_t>(words, 4), &t->lights);
        if (rc != PHR_OK) return rc;
    }
    hipStream_t s = t->gpu->stream;
    if (t->ncand == 0) {
        PHR_HIP(hipMemsetAsync(t->lights->ptr, 0, 4, s));
        t->lights->len = 1;
        t->glow = 0;
        return PHR_OK;
    }
    hipLaunchKernelGGL(k_light_power, dim3((t->ncand + 255) / 256), dim3(256), 0, s, (const uint32_t*)t->lcand->ptr, t->ncand,
                       (const uint4*)t->faces->ptr, (const float*)t->verts->ptr, (const float*)t->worlds_dev->ptr,
                       (float*)t->lpow->ptr);
    hipLaunchKernelGGL(k_light_table, dim3(1), dim3(64), 0, s, (const uint32_t*)t->lcand->ptr, (const float*)t->lpow->ptr,
                       t->ncand, (uint32_t*)t->lights->ptr);
    PHR_HIP(hipGetLastError());
    t->stale = true;  // the count (and the table's length) is read back at the next sample
    return PHR_OK;
}

// A recorded posed frame's camera stamp on the device (what the host does
// at a sample from the values forest_settle reads back): the extent from the
// top tree's root box, the occlusion distance when the caller left it unset,
// the ray offset, and the emitter count.
__global__ void k_stamp_pose(const float4* __restrict__ nodes, const uint32_t* __restrict__ lights, uint32_t glow_on_device,
                             uint32_t ao_auto, PhrCamera* __restrict__ cam) {
    if (threadIdx.x != 0) return;
    const float4 lo = nodes[0], hi = nodes[1];
    float sq = -0.0f;
    sq = sq + (hi.x - lo.x) * (hi.x - lo.x);
    sq = sq + (hi.y - lo.y) * (hi.y - lo.y);
    sq = sq + (hi.z - lo.z) * (hi.z - lo.z);
    const float extent = fmaxf(sqrtf(sq), 1e-3f);
    if (ao_auto) cam->ao = extent * 0.08f;
    cam->eps = extent * 1e-5f;
    if (glow_on_device) cam->glow = lights[0];
}

// What the host keeps of a pose applied on the device: the extent (the top
// tree's root box, forest.rs:189-199) and the emitter count.
int forest_settle(phr_tracer* t) {
    if (!t->forest || !t->stale) return PHR_OK;
    float box[8];
    uint32_t cnt = 0;
    hipStream_t s = t->gpu->stream;
    PHR_HIP(hipMemcpyAsync(box, t->nodes->ptr, 32, hipMemcpyDeviceToHost, s));
    PHR_HIP(hipMemcpyAsync(&cnt, t->lights->ptr, 4, hipMemcpyDeviceToHost, s));
    PHR_HIP(hipStreamSynchronize(s));
    float sq = -0.0f;
    for (int k = 0; k < 3; ++k) sq = sq + (box[4 + k] - box[k]) * (box[4 + k] - box[k]);
    t->extent = std::max(std::sqrt(sq), 1e-3f);
    if (t->ncand <= LIGHT_DEVICE_MAX) {
        t->glow = cnt;
        t->lights->len = 1 + 3 * (size_t)cnt;
    }
    t->stale = false;
    return PHR_OK;
}

// Forest::retlas + inst_words (src/trace/forest.rs:111-127, 168-181): the
// placed instances' world boxes from their mesh trees' root boxes and the
// top tree over them, both on the device (the top tree into node slots
// [0, 2p - 1), leaves flagged INST), and the instance table. Nothing waits:
// the extent (the top tree's root box, forest.rs:189-199) is read back at the
// next sample.
// The host's half of a new pose: the matrices and the instance table staged
// for upload (into the pinned buffers when a posed frame is recorded, which
// its copies read). Returns the upload sources.
void forest_pose_host(phr_tracer* t, const float** wsrc, const uint32_t** isrc) {
    const uint32_t n = t->ninsts;
    std::vector<float>& wl = t->inst_worlds_h;
    wl.resize(16 * (size_t)n);
    for (uint32_t i = 0; i < n; ++i) std::memcpy(&wl[16 * (size_t)i], t->inst_in[i].world, 64);
    std::vector<uint32_t>& words = t->words_h;
    words.assign((size_t)(INSTANCE_LEN + 2) * n, 0u);
    phr_mesh unit{0, 0, 0, 0, {0.f, 0.f, 0.f}, {1.f, 1.f, 1.f}};
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t m = t->inst_in[i].mesh;
        const bool ok = m < t->nmeshes;
        const uint32_t root = ok && t->meshes[m].face_hi > t->meshes[m].face_lo ? t->node_base[m] : PHR_NONE;
        inst_words(t->inst_in[i], root, ok ? t->meshes[m] : unit, &words[INSTANCE_LEN * (size_t)i]);
    }
    uint32_t* draw = &words[(INSTANCE_LEN + 1) * (size_t)n];  // Forest::regroup (forest.rs:93-109)
    uint32_t at = 0;
    std::vector<std::vector<uint32_t>> by_mesh(t->nmeshes);
    for (uint32_t i = 0; i < n; ++i)
        if (t->inst_in[i].mesh < t->nmeshes) by_mesh[t->inst_in[i].mesh].push_back(i);
    for (uint32_t m = 0; m < t->nmeshes; ++m)
        for (uint32_t i : by_mesh[m]) draw[at++] = i;
    *wsrc = wl.data();
    *isrc = words.data();
    if (t->pin_worlds) {
        std::memcpy(t->pin_worlds, wl.data(), 64ull * n);
        std::memcpy(t->pin_words, words.data(), 4 * words.size());
        *wsrc = t->pin_worlds;
        *isrc = t->pin_words;
    }
}

// Forest::retlas + inst_words (src/trace/forest.rs:111-127, 168-181): the
// placed instances' world boxes from their mesh trees' root boxes and the
// top tree over them, both on the device (the top tree into node slots
// [0, 2p - 1), leaves flagged INST), and the instance table. Nothing waits:
// the extent (the top tree's root box, forest.rs:189-199) is read back at the
// next sample.
int forest_top(phr_tracer* t, bool fresh) {
    const uint32_t p = (uint32_t)t->placed.size(), n = t->ninsts;
    hipStream_t s = t->gpu->stream;
    const float* wsrc = nullptr;
    const uint32_t* isrc = nullptr;
    forest_pose_host(t, &wsrc, &isrc);
    PHR_HIP(hipMemcpyAsync(t->worlds_dev->ptr, wsrc, 64ull * n, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_world_boxes, dim3((p + 255) / 256), dim3(256), 0, s, (const float*)t->worlds_dev->ptr,
                       (const uint32_t*)t->placed_dev->ptr, (const uint32_t*)t->inst_mesh->ptr, t->bases_dev,
                       (const float4*)t->nodes->ptr, p, (float4*)t->wboxes->ptr);
    PHR_HIP(hipGetLastError());
    int rc;
    if (fresh) {
        const uint32_t seg[2] = {0, p}, base[1] = {0};
        if ((rc = phr_bvh_internal_build_into(t->gpu, t->wboxes, 1, seg, 1, t->nodes, base, PHR_LEAF | PHR_INST, &t->tlas)) != PHR_OK)
            return rc;
    } else if ((rc = phr_bvh_internal_rebuild_async(t->tlas)) != PHR_OK) {
        return rc;
    }
    PHR_HIP(hipMemcpyAsync(t->insts->ptr, isrc, 4ull * (INSTANCE_LEN + 2) * n, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_inst_order, dim3((p + 255) / 256), dim3(256), 0, s, (const uint32_t*)t->placed_dev->ptr,
                       (const uint32_t*)phr_bvh_internal_order(t->tlas)->ptr, p, (uint32_t*)t->insts->ptr + INSTANCE_LEN * (size_t)n);
    PHR_HIP(hipGetLastError());
    t->stale = true;
    return PHR_OK;
}

// The faces into their trees' order: slot-ordered faces and the triangle
// cache on the device, and the slot order on the host.
int forest_gather(phr_tracer* t) {
    hipStream_t s = t->gpu->stream;
    const uint32_t n = t->n;
    phr_array* order = phr_bvh_internal_order(t->bvh);
    hipLaunchKernelGGL(k_gather, dim3((n + 255) / 256), dim3(256), 0, s, (const uint4*)t->in_faces->ptr, (const float4*)t->verts->ptr,
                       (const uint32_t*)order->ptr, n, (uint4*)t->faces->ptr, (float4*)t->tris->ptr);
    t->order.resize(n);
    PHR_HIP(hipMemcpyAsync(t->order.data(), order->ptr, 4ull * n, hipMemcpyDeviceToHost, s));
    PHR_HIP(hipStreamSynchronize(s));
    t->faces_slot.resize(4 * (size_t)n);
    for (uint32_t k = 0; k < n; ++k)
        std::memcpy(&t->faces_slot[4 * (size_t)k], &t->faces_in[4 * (size_t)t->order[k]], 16);
    return PHR_OK;
}

}  // namespace

extern "C" {

// Tracer::new (src/trace/build.rs:11-13): the scene from triangle records
// (vertices shared by exact bits, materials by value), the LBVH over the faces
// built on the device, the faces put in leaf order, and the emitter table.
// PHR_ERR_ARG for an empty scene (the reference's None); PHR_ERR_UNSUPPORTED
// for a triangle the restated integrator does not carry (transmission).
namespace {

// The items a binding of `limit` bytes takes for an array of `len` items of
// `elem` bytes over `chunks` bindings (GpuArray::split, src/gpu/array.rs:70-83);
// 0 when they do not fit.
uint32_t split_cut(uint64_t len, uint64_t elem, uint64_t limit, uint32_t chunks) {
    const uint64_t cap = std::max<uint64_t>(limit / elem, 1);
    if (len > (uint64_t)chunks * cap) return 0;
    return (uint32_t)std::min<uint64_t>(std::max<uint64_t>(len, 1), cap);
}

// A new tracer's split at the device's largest allocation over 2 bindings.
void default_split(phr_tracer* t, uint64_t nodes, uint64_t faces, uint64_t verts) {
    const uint64_t lim = std::max<uint64_t>(t->gpu->max_alloc, 1);
    t->split[0] = std::max<uint32_t>(split_cut(nodes, 32, lim, 2), 1);
    t->split[1] = std::max<uint32_t>(split_cut(faces, 16, lim, 2), 1);
    t->split[2] = std::max<uint32_t>(split_cut(verts, 32, lim, 2), 1);
    t->split[3] = 0;
    t->chunks = 2;
}

}  // namespace

int phr_tracer_new(phr_gpu* g, const void* tris_v, uint32_t n, uint32_t leaf, phr_tracer** out) {
    PHR_CHECK(g && out && tris_v, PHR_ERR_ARG, "phr_tracer_new: null argument");
    PHR_CHECK(n > 0, PHR_ERR_ARG, "phr_tracer_new: empty scene");
    const PhrTri* tris = (const PhrTri*)tris_v;
    for (uint32_t f = 0; f < n; ++f)
        PHR_CHECK(tris[f].trans == 0.f, PHR_ERR_UNSUPPORTED, "phr_tracer_new: transmission (glass) is not supported");
    phr_tracer* t = new phr_tracer();
    t->gpu = g;
    t->n = n;
    t->leaf = leaf;
    auto fail = [&](int rc) { tracer_release(t); return rc; };
    // Scene::add for each triangle (src/trace/scene/add.rs:9-69).
    std::unordered_map<std::array<uint32_t, 7>, uint32_t, WordsHash<7>> vmap;
    std::unordered_map<std::array<uint32_t, MAT_WORDS>, uint32_t, WordsHash<MAT_WORDS>> mmap;
    vmap.reserve(3 * (size_t)n);
    std::vector<uint32_t> in_faces(4 * (size_t)n);
    std::vector<float> verts, mats;
    t->corner_vertex.resize(3 * (size_t)n);
    float lo[3] = {3.40282347e38f, 3.40282347e38f, 3.40282347e38f}, hi[3] = {-3.40282347e38f, -3.40282347e38f, -3.40282347e38f};
    for (uint32_t f = 0; f < n; ++f) {
        const PhrTri& tr = tris[f];
        float m[MAT_WORDS] = {tr.color[0], tr.color[1], tr.color[2], tr.color[3], tr.rough, tr.metal, tr.trans, tr.ior,
                              tr.emit[0], tr.emit[1], tr.emit[2], tr.bump, 0, 0, 0, 0};
        std::array<uint32_t, MAT_WORDS> mk;
        for (uint32_t k = 0; k < MAT_WORDS; ++k) mk[k] = fbits(m[k]);
        auto mi = mmap.find(mk);
        uint32_t mid;
        if (mi == mmap.end()) {
            mid = (uint32_t)mmap.size();
            mmap.emplace(mk, mid);
            mats.insert(mats.end(), m, m + MAT_WORDS);
        } else {
            mid = mi->second;
        }
        const float* p[3] = {tr.a, tr.b, tr.c};
        const float* nn[3] = {tr.na, tr.nb, tr.nc};
        for (int k = 0; k < 3; ++k) {
            float o[2];
            oct_pack(nn[k], o);
            std::array<uint32_t, 7> key = {fbits(p[k][0]), fbits(p[k][1]), fbits(p[k][2]), fbits(o[0]), fbits(o[1]), 0u, 0u};
            auto it = vmap.find(key);
            uint32_t id;
            if (it == vmap.end()) {
                id = (uint32_t)vmap.size();
                vmap.emplace(key, id);
                float w[8] = {p[k][0], p[k][1], p[k][2], 1.0f, o[0], o[1], 0.f, 0.f};
                verts.insert(verts.end(), w, w + 8);
            } else {
                id = it->second;
            }
            in_faces[4 * (size_t)f + k] = id;
            t->corner_vertex[3 * (size_t)f + k] = id;
            for (int j = 0; j < 3; ++j) { lo[j] = std::min(lo[j], p[k][j]); hi[j] = std::max(hi[j], p[k][j]); }
        }
        in_faces[4 * (size_t)f + 3] = mid;
    }
    t->nverts = (uint32_t)vmap.size();
    t->nmats = (uint32_t)mmap.size();
    // The world's diagonal, which ray offsets and the AO reach are shares of
    // (src/trace/forest.rs:191-199, src/trace/sample.rs:110-118).
    float sq = -0.0f;
    for (int j = 0; j < 3; ++j) sq = sq + (hi[j] - lo[j]) * (hi[j] - lo[j]);
    t->extent = std::max(std::sqrt(sq), 1e-3f);
    int rc;
    hipStream_t s = g->stream;
    if ((rc = phr_array_alloc(g, 16, n, &t->in_faces)) != PHR_OK) return fail(rc);
    if ((rc = phr_array_alloc(g, 32, t->nverts, &t->verts)) != PHR_OK) return fail(rc);
    if ((rc = phr_array_alloc(g, 4 * MAT_WORDS, t->nmats, &t->mats)) != PHR_OK) return fail(rc);
    if ((rc = phr_array_alloc(g, 32, n, &t->prims)) != PHR_OK) return fail(rc);
    if ((rc = phr_array_alloc(g, 16, n, &t->faces)) != PHR_OK) return fail(rc);
    if ((rc = phr_array_alloc(g, 48, n, &t->tris)) != PHR_OK) return fail(rc);
    if (hipMalloc(&t->rays, 48) != hipSuccess || hipMemsetAsync(t->rays, 0, 48, s) != hipSuccess ||
        hipMalloc(&t->cam_dev, sizeof(PhrCamera)) != hipSuccess) {
        phr_set_error_msg("phr_tracer_new: allocation failed");
        return fail(PHR_ERR_HIP);
    }
    if (hipMemcpyAsync(t->in_faces->ptr, in_faces.data(), 16ull * n, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(t->verts->ptr, verts.data(), 32ull * t->nverts, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(t->mats->ptr, mats.data(), 64ull * t->nmats, hipMemcpyHostToDevice, s) != hipSuccess) {
        phr_set_error_msg("phr_tracer_new: upload failed");
        return fail(PHR_ERR_HIP);
    }
    hipLaunchKernelGGL(k_tri_boxes, dim3((n + 255) / 256), dim3(256), 0, s, (const uint4*)t->in_faces->ptr,
                       (const float4*)t->verts->ptr, n, (float4*)t->prims->ptr);
    if ((rc = phr_bvh_internal_build(g, t->prims, leaf, &t->bvh)) != PHR_OK) return fail(rc);
    phr_array* order = phr_bvh_internal_order(t->bvh);
    hipLaunchKernelGGL(k_gather, dim3((n + 255) / 256), dim3(256), 0, s, (const uint4*)t->in_faces->ptr,
                       (const float4*)t->verts->ptr, (const uint32_t*)order->ptr, n, (uint4*)t->faces->ptr, (float4*)t->tris->ptr);
    if ((rc = enqueue_walk_records(g, t->bvh, &t->recs, t->tris)) != PHR_OK) return fail(rc);
    t->order.resize(n);
    if (hipMemcpyAsync(t->order.data(), order->ptr, 4ull * n, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
        phr_set_error_msg("phr_tracer_new: order readback failed");
        return fail(PHR_ERR_HIP);
    }
    if ((rc = build_lights(t, in_faces, verts, mats)) != PHR_OK) return fail(rc);
    if (hipEventCreate(&t->ev0) != hipSuccess || hipEventCreate(&t->ev1) != hipSuccess) {
        phr_set_error_msg("phr_tracer_new: event creation failed");
        return fail(PHR_ERR_HIP);
    }
    default_split(t, 2ull * n - 1, n, t->nverts);
    *out = t;
    return PHR_OK;
}

void phr_tracer_free(phr_tracer* t) { tracer_release(t); }
uint32_t phr_tracer_len(const phr_tracer* t) { return t ? t->n : 0; }
uint32_t phr_tracer_depth(const phr_tracer* t) { return t ? phr_bvh_depth(t->bvh) : 0; }
float phr_tracer_extent(const phr_tracer* t) {
    if (!t) return 0.f;
    (void)forest_settle(const_cast<phr_tracer*>(t));  // a pose applied on the device is read back here
    return t->extent;
}
uint32_t phr_tracer_glow(const phr_tracer* t) {
    if (!t) return 0;
    (void)forest_settle(const_cast<phr_tracer*>(t));
    return t->glow;
}
phr_bvh* phr_tracer_bvh(phr_tracer* t) { return t ? t->bvh : nullptr; }

// The slot -> input face map and the scene arrays, for tests and tooling.
int phr_tracer_order(phr_tracer* t, uint32_t* out) {
    PHR_CHECK(t && out, PHR_ERR_ARG, "phr_tracer_order: null");
    std::memcpy(out, t->order.data(), 4ull * t->n);
    return PHR_OK;
}
int phr_tracer_lights(phr_tracer* t, uint32_t* out, uint32_t cap, uint32_t* len) {
    PHR_CHECK(t && len, PHR_ERR_ARG, "phr_tracer_lights: null");
    if (int rc = forest_settle(t); rc != PHR_OK) return rc;
    *len = (uint32_t)t->lights->len;
    if (out) {
        PHR_CHECK(cap >= *len, PHR_ERR_SIZE, "phr_tracer_lights: buffer too small");
        return phr_array_read_range(t->lights, 0, *len, out);
    }
    return PHR_OK;
}

// Closest- or any-hit rays through the tracer's tree: `rays` and `hits` on
// the host; hits report the leaf slot (phr_tracer_order maps it back).
int phr_tracer_raycast(phr_tracer* t, const void* rays, uint32_t nrays, int any_hit, void* hits) {
    PHR_CHECK(t && (nrays == 0 || (rays && hits)), PHR_ERR_ARG, "phr_tracer_raycast: null argument");
    if (nrays == 0) return PHR_OK;
    hipStream_t s = t->gpu->stream;
    void *dr = nullptr, *dh = nullptr;
    PHR_HIP(hipMalloc(&dr, 32ull * nrays));
    hipError_t e = hipMalloc(&dh, 20ull * nrays);
    if (e == hipSuccess) e = hipMemcpyAsync(dr, rays, 32ull * nrays, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) {
        launch_raycast(t, (const PhrRay*)dr, nrays, any_hit, (PhrHit*)dh);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(hits, dh, 20ull * nrays, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    (void)hipFree(dr);
    if (dh) (void)hipFree(dh);
    if (e != hipSuccess) { phr_set_error("phr_tracer_raycast", e); return PHR_ERR_HIP; }
    return PHR_OK;
}

// Device-resident rays (a 32-byte ray array) into a 20-byte hit array, enqueued.
int phr_tracer_raycast_on(phr_tracer* t, phr_array* rays, int any_hit, phr_array* hits) {
    PHR_CHECK(t && rays && hits && rays->elem == 32 && hits->elem == 20 && hits->len >= rays->len, PHR_ERR_ARG,
              "phr_tracer_raycast_on: rays must be 32-byte records and hits 20-byte records, as many");
    uint32_t nrays = (uint32_t)rays->len;
    if (nrays == 0) return PHR_OK;
    launch_raycast(t, (const PhrRay*)rays->ptr, nrays, any_hit, (PhrHit*)hits->ptr);
    PHR_HIP(hipGetLastError());
    return PHR_OK;
}

// Tracer::set_samples (src/trace/mod.rs:242).
void phr_tracer_set_samples(phr_tracer* t, uint32_t n) { if (t) t->samples = n < 1 ? 1 : n; }
// Tracer::warm (src/trace/build.rs:27-69): load every path-kernel build on
// the device before its first frame, so a switch of build (the autotune, a
// forest, path regeneration) does not wait for its code object. Each entry
// gets the kernel's name and the milliseconds its load took; a build loaded
// already costs next to nothing. *count: entries written (at most cap).
int phr_tracer_warm(phr_gpu* g, uint32_t cap, const char** names, double* ms, uint32_t* count) {
    PHR_CHECK(g && count && (cap == 0 || (names && ms)), PHR_ERR_ARG, "phr_tracer_warm: bad argument");
    PHR_HIP(hipSetDevice(g->device));
    struct K { const char* name; const void* fn; };
    const K ks[] = {
        {"k_path<2>", (const void*)k_path<false, 2>}, {"k_path<3>", (const void*)k_path<false, 3>},
        {"k_path<4>", (const void*)k_path<false, 4>}, {"k_path<5>", (const void*)k_path<false, 5>},
        {"k_path_forest<2>", (const void*)k_path<true, 2>}, {"k_path_forest<3>", (const void*)k_path<true, 3>},
        {"k_path_forest<4>", (const void*)k_path<true, 4>},
        {"k_path_forest<3, short stack>", (const void*)k_path<true, 3, FOREST_SHORT_STACK>}, {"k_path_rg<2>", (const void*)k_path_rg<2>},
        {"k_path_rg<3>", (const void*)k_path_rg<3>}, {"k_path_rg<4>", (const void*)k_path_rg<4>},
        {"k_path_rg<5>", (const void*)k_path_rg<5>}, {"k_path_sw<2>", (const void*)k_path_sw<2>},
        {"k_path_sw<3>", (const void*)k_path_sw<3>}, {"k_path_sw<4>", (const void*)k_path_sw<4>},
        {"k_path_sw<5>", (const void*)k_path_sw<5>}, {"k_path_batch<2>", (const void*)k_path_batch<2>},
        {"k_path_batch<3>", (const void*)k_path_batch<3>}, {"k_path_batch<4>", (const void*)k_path_batch<4>},
        {"k_path_batch<5>", (const void*)k_path_batch<5>}, {"k_path_batch_sw<2>", (const void*)k_path_batch_sw<2>},
        {"k_path_batch_sw<3>", (const void*)k_path_batch_sw<3>}, {"k_path_batch_sw<4>", (const void*)k_path_batch_sw<4>},
        {"k_path_batch_sw<5>", (const void*)k_path_batch_sw<5>}, {"k_path_sw_forest<2>", (const void*)k_path_sw_forest<2, PHR_STACK>},
        {"k_path_sw_forest<3>", (const void*)k_path_sw_forest<3, PHR_STACK>},
        {"k_path_sw_forest<3, short stack>", (const void*)k_path_sw_forest<3, FOREST_SHORT_STACK>},
        {"k_path_sw_forest<4>", (const void*)k_path_sw_forest<4, FOREST_SHORT_STACK>},
    };
    uint32_t n = 0;
    for (const K& k : ks) {
        auto t0 = std::chrono::steady_clock::now();
        hipFuncAttributes attr;
        PHR_HIP(hipFuncGetAttributes(&attr, k.fn));
        const double dt = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        if (n < cap) {
            names[n] = k.name;
            ms[n] = dt;
        }
        ++n;
    }
    *count = std::min(n, cap);
    return PHR_OK;
}

// Tracer::samples (src/trace/mod.rs:152-155).
uint32_t phr_tracer_samples(phr_tracer* t) { return t ? t->samples : 0u; }
// Tracer::set_adaptive (src/trace/mod.rs:235-240): [threshold, min samples,
// step, on] stamped onto every later camera; all zero turns it off.
int phr_tracer_set_adaptive(phr_tracer* t, const float adapt[4]) {
    PHR_CHECK(t && adapt, PHR_ERR_ARG, "phr_tracer_set_adaptive: null argument");
    std::memcpy(t->adapt, adapt, sizeof t->adapt);
    t->adapt_set = true;
    return PHR_OK;
}

// Tracer::sample_in (src/trace/sample.rs:104-249) without slices: `spp`
// samples in one launch. The camera is stamped as the reference stamps it
// (ao from the extent when unset, eps, glow); a repeated camera at the same
// size adds samples, any other starts the image over. `frame` gets the index of
// the last sample taken, `done` whether the tracer's sample count is reached.
// Tracer::sample_in (src/trace/sample.rs:104-249): the samples' launches in
// `slices` pieces over the image's tiles, the device drained after each and
// `go` asked whether to carry on. When it says no, the call is abandoned
// (*finished = 0) and the next one starts the image over, since a frame with
// some pixels a sample ahead is no frame.
// The wavefront build (k_wf_*) over every tile: samples [frame0, frame0 +
// spp) of the image. The host reads the next turn's ray count every 16 turns
// and stops once no slot has a sample left; the turns past the last hold no
// rays and end at once.
int wf_frame(phr_tracer* t, const SceneDev& sc, const PhrCamera* cp, const PhrCamera& cam, uint32_t tiles_x, uint32_t tiles,
             uint32_t frame0, uint32_t spp, uint32_t* pxp) {
    hipStream_t s = t->gpu->stream;
    const uint32_t slots = tiles * WF_THREADS;
    if (t->wf_slots != slots) {
        for (phr_array** a : {&t->wf_state, &t->wf_queue, &t->wf_shadow}) {
            if (*a) phr_array_release(*a);
            *a = nullptr;
        }
        t->wf_slots = 0;
        int rc;
        if ((rc = phr_array_alloc(t->gpu, 16, 9ull * slots, &t->wf_state)) != PHR_OK) return rc;
        if ((rc = phr_array_alloc(t->gpu, 4, 3ull * slots, &t->wf_queue)) != PHR_OK) return rc;
        if ((rc = phr_array_alloc(t->gpu, 16, 4ull * slots + (2ull * slots + 15) / 16, &t->wf_shadow)) != PHR_OK) return rc;
        t->wf_slots = slots;
    }
    if (!t->wf_ctr) PHR_HIP(hipMalloc((void**)&t->wf_ctr, 4 * WF_CTRS));
    if (!t->wf_pin) PHR_HIP(hipHostMalloc((void**)&t->wf_pin, 4, hipHostMallocDefault));
    float4* st = (float4*)t->wf_state->ptr;
    WfDev w;
    w.s0 = st;
    w.s1 = st + (size_t)slots;
    w.s2 = st + 2 * (size_t)slots;
    w.s3 = st + 3 * (size_t)slots;
    w.s4 = st + 4 * (size_t)slots;
    w.s5 = st + 5 * (size_t)slots;
    w.s6 = st + 6 * (size_t)slots;
    w.s7 = st + 7 * (size_t)slots;
    w.hit = st + 8 * (size_t)slots;
    w.q0 = (uint32_t*)t->wf_queue->ptr;
    w.q1 = w.q0 + slots;
    w.sq = w.q1 + slots;
    w.sh = (float4*)t->wf_shadow->ptr;
    w.vis = (uint8_t*)(w.sh + 4 * (size_t)slots);
    w.pend = w.vis + slots;
    w.ctr = t->wf_ctr;
    hipLaunchKernelGGL(k_wf_zero, dim3(1), dim3(64), 0, s, w.ctr);
    hipLaunchKernelGGL(k_wf_init, dim3(tiles), dim3(WF_THREADS), 0, s, cp, w, tiles_x, slots, frame0, pxp);
    PHR_HIP(hipGetLastError());
    const uint32_t cus = (uint32_t)std::max(t->gpu->cu_count, 1);
    const uint32_t walk_blocks = cus * 5u;  // the LDS stacks fit five 256-thread blocks a CU
    const uint32_t shade_blocks = (slots + WF_SHADE_ITEMS * WF_THREADS - 1) / (WF_SHADE_ITEMS * WF_THREADS);
    // A sample takes at most one turn a camera segment and one a bounce.
    const bool depth_set = cam.depth[0] != 0u;
    const uint64_t total = depth_set ? cam.depth[0] - 1u : 8u;
    const uint64_t passes = (depth_set ? cam.depth[1] - 1u : 8u) + (depth_set ? std::min<uint64_t>(cam.depth[2] - 1u, total) : 8u) + 1u;
    const uint64_t limit = (uint64_t)spp * (passes + (cam.bounces ? cam.bounces : 1u) + 1u) + 32u;
    uint64_t turn = 0;
    for (;; ++turn) {
        const uint32_t par = (uint32_t)(turn & 1u);
        hipLaunchKernelGGL(k_wf_walk<false>, dim3(walk_blocks), dim3(WF_THREADS), 0, s, sc, cp, w, par, tiles_x, pxp, t->rays);
        hipLaunchKernelGGL(k_wf_shade, dim3(shade_blocks), dim3(WF_THREADS), 0, s, sc, cp, w, par, tiles_x, frame0, spp, pxp);
        hipLaunchKernelGGL(k_wf_walk<true>, dim3(walk_blocks), dim3(WF_THREADS), 0, s, sc, cp, w, par, tiles_x, pxp, t->rays);
        PHR_HIP(hipGetLastError());
        if ((turn & 15u) == 15u || turn + 1 >= limit) {
            PHR_HIP(hipMemcpyAsync(t->wf_pin, w.ctr + (1u - par), 4, hipMemcpyDeviceToHost, s));
            PHR_HIP(hipStreamSynchronize(s));
            if (*t->wf_pin == 0u) break;
            PHR_CHECK(turn + 1 < limit, PHR_ERR_SIZE, "phr_tracer_sample: the wavefront build ran past its turn bound");
        }
    }
    t->wf_turns = (uint32_t)(turn + 1);
    const uint64_t n = (uint64_t)cam.w * cam.h;
    hipLaunchKernelGGL(k_wf_final, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, s, cp, w, tiles_x, frame0, spp, pxp);
    PHR_HIP(hipGetLastError());
    return PHR_OK;
}

int phr_tracer_sample_in(phr_tracer* t, const void* cam_v, uint32_t w, uint32_t h, uint32_t spp, uint32_t slices,
                         phr_go_fn go, void* user, uint32_t* frame, int* done, int* finished) {
    PHR_CHECK(t && cam_v, PHR_ERR_ARG, "phr_tracer_sample: null argument");
    PHR_CHECK(spp >= 1, PHR_ERR_ARG, "phr_tracer_sample: spp must be at least 1");
    w = std::max<uint32_t>(w, 1);
    h = std::max<uint32_t>(h, 1);
    PhrCamera cam;
    std::memcpy(&cam, cam_v, sizeof cam);
    if (t->adapt_set) std::memcpy(cam.adapt, t->adapt, sizeof cam.adapt);
    PHR_CHECK(!(cam.depth[0] == 0u || cam.depth[3] > 1u), PHR_ERR_UNSUPPORTED,
              "phr_tracer_sample: the glossy lobe is not restated; set Max Bounces (depth) with glossy 0");
    PHR_CHECK(cam.integrator[0] < 0.5f, PHR_ERR_UNSUPPORTED, "phr_tracer_sample: only phiray's own pixel filter (0)");
    if (int rc0 = forest_settle(t); rc0 != PHR_OK) return rc0;
    if (cam.adapt[3] != 0.0f) {
        const uint32_t st = (uint32_t)cam.adapt[2];
        PHR_CHECK((st & (st - 1u)) == 0u, PHR_ERR_ARG, "phr_tracer_sample: the adaptive step must be a power of two");
    }
    cam.w = w;
    cam.h = h;
    if (cam.ao <= 0.f) cam.ao = t->extent * 0.08f;
    cam.eps = t->extent * 1e-5f;
    cam.glow = t->glow;
    cam.frame = 0;
    const uint64_t n = (uint64_t)w * h;
    bool same = t->have_last && t->w == w && t->h == h && std::memcmp(&t->last, &cam, sizeof cam) == 0;
    uint32_t frame0 = same ? t->frame + 1 : 0;
    if (!t->px || t->w != w || t->h != h) {
        if (t->px) phr_array_release(t->px);
        t->px = nullptr;
        int rc = phr_array_alloc(t->gpu, 4, PX_LEN * n + 1, &t->px);
        if (rc != PHR_OK) return rc;
    }
    t->w = w;
    t->h = h;
    t->last = cam;
    t->have_last = true;
    SceneDev sc = scene_dev(t);
    hipStream_t s = t->gpu->stream;
    auto t0 = std::chrono::steady_clock::now();
    PHR_HIP(hipMemcpyAsync(t->cam_dev, &t->last, sizeof(PhrCamera), hipMemcpyHostToDevice, s));
    PHR_HIP(hipEventRecord(t->ev0, s));
    const uint32_t tiles_x = (w + 15) / 16, tiles = tiles_x * ((h + 15) / 16);
    const PhrCamera* cp = (const PhrCamera*)t->cam_dev;
    uint32_t* pxp = (uint32_t*)t->px->ptr;
    if (finished) *finished = 1;
    slices = std::min(std::max<uint32_t>(slices, 1u), tiles);
    bool abandoned = false;
    // One launch of `count` samples from f0, in `slices` pieces of the tiles.
    const uint32_t regen = regen_allowed(cam) ? t->path_regen : 0u;
    auto path = [&](uint32_t f0, uint32_t count, uint32_t listed) -> int {
        for (uint32_t k = 0; k < slices && !abandoned; ++k) {
            const uint32_t a = (uint32_t)((uint64_t)tiles * k / slices), b = (uint32_t)((uint64_t)tiles * (k + 1) / slices);
            if (b == a) continue;
            launch_path(path_build_waves(t), t->forest, regen, dim3(b - a), s, sc, cp, tiles_x, f0, count, pxp, t->rays, listed, a,
                        path_short_stack(t), t->split_shade);
            PHR_HIP(hipGetLastError());
            if (slices > 1) {
                PHR_HIP(hipStreamSynchronize(s));
                if (go && !go(user)) abandoned = true;
            }
        }
        return PHR_OK;
    };
    if (cam.adapt[3] == 0.0f && regen && t->path_regen == 2u && !t->forest && slices == 1u) {
        if (int rc = wf_frame(t, sc, cp, cam, tiles_x, tiles, frame0, spp, pxp); rc != PHR_OK) return rc;
    } else if (cam.adapt[3] == 0.0f) {
        if (int rc = path(frame0, spp, 0u); rc != PHR_OK) return rc;
    } else {
        // Adaptive Sampling (sample.rs `BTail::Adapt`): the tail runs after
        // every sample, and only a step boundary past the minimum changes a
        // flag, so samples run in launches that end at those boundaries. The
        // dilation and the list follow each launch.
        const uint32_t min_s = (uint32_t)cam.adapt[1], step = std::max<uint32_t>((uint32_t)cam.adapt[2], 1u);
        const uint32_t ng = (uint32_t)((n + 255) / 256);
        for (uint32_t f = frame0; f < frame0 + spp;) {
            uint32_t b = std::max(f, min_s + 1u);  // the next checked sample: past the minimum, last of a step
            b = b | (step - 1u);
            const uint32_t e = std::min(frame0 + spp, b + 1u);
            if (int rc = path(f, e - f, f > 0u ? 1u : 0u); rc != PHR_OK) return rc;
            if (abandoned) break;
            if (e == b + 1u)
                hipLaunchKernelGGL(k_adapt_check, dim3(ng), dim3(256), 0, s, pxp, (uint32_t)n, cam.adapt[0]);
            hipLaunchKernelGGL(k_adapt_dilate, dim3(ng), dim3(256), 0, s, pxp, w, h);
            hipLaunchKernelGGL(k_gather_clear, dim3(1), dim3(256), 0, s, pxp, (size_t)n);
            hipLaunchKernelGGL(k_gather, dim3(tiles), dim3(256), 0, s, pxp, w, h, tiles_x);
            f = e;
        }
    }
    PHR_HIP(hipGetLastError());
    PHR_HIP(hipEventRecord(t->ev1, s));
    t->last_submit_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    PHR_HIP(hipEventSynchronize(t->ev1));
    if (abandoned) {
        t->have_last = false;
        if (finished) *finished = 0;
        return PHR_OK;
    }
    float ms = 0.f;
    PHR_HIP(hipEventElapsedTime(&ms, t->ev0, t->ev1));
    t->last_gpu_s = ms * 1e-3;
    t->frame = frame0 + spp - 1;
    if (frame) *frame = t->frame;
    if (done) *done = (t->frame + 1 >= t->samples) ? 1 : 0;
    return PHR_OK;
}

// Tracer::sample: `spp` samples in one piece.
int phr_tracer_sample(phr_tracer* t, const void* cam, uint32_t w, uint32_t h, uint32_t spp, uint32_t* frame, int* done) {
    return phr_tracer_sample_in(t, cam, w, h, spp, 1, nullptr, nullptr, frame, done, nullptr);
}

// The path kernel variant: register allocation fitted to 2, 3, 4 or 5 waves
// a SIMD (no counterpart in phiray, whose WGSL pipeline has one).
int phr_tracer_set_path_waves(phr_tracer* t, uint32_t waves) {
    PHR_CHECK(t, PHR_ERR_ARG, "phr_tracer_set_path_waves: null");
    PHR_CHECK(waves >= PATH_WAVES_MIN && waves <= PATH_WAVES_MAX, PHR_ERR_ARG, "phr_tracer_set_path_waves: waves must be 2 to 5");
    t->path_waves = waves;
    return PHR_OK;
}
uint32_t phr_tracer_path_waves(const phr_tracer* t) { return t ? t->path_waves : 0; }

// The path-regenerating build (1), the wavefront build (2) or neither (0);
// they run only for one tree and a camera without Adaptive Sampling or the
// studio rig, the nested build otherwise. A sliced sample (sample_in with
// several slices) and a recorded frame take the regenerating build for 2
// (no counterpart in phiray).
int phr_tracer_set_path_regen(phr_tracer* t, uint32_t on) {
    PHR_CHECK(t && on <= 3u, PHR_ERR_ARG, "phr_tracer_set_path_regen: null, or not 0, 1, 2 or 3");
    t->path_regen = on;
    return PHR_OK;
}
uint32_t phr_tracer_path_regen(const phr_tracer* t) { return t ? t->path_regen : 0; }
// The split build's shading point: it shades once `eighths` (1 to 8) eighths
// of a wave's working lanes wait. The images do not depend on it.
int phr_tracer_set_split_shade(phr_tracer* t, uint32_t eighths) {
    PHR_CHECK(t && eighths >= 1u && eighths <= 8u, PHR_ERR_ARG, "phr_tracer_set_split_shade: null, or not 1 to 8");
    t->split_shade = eighths;
    return PHR_OK;
}
uint32_t phr_tracer_split_shade(const phr_tracer* t) { return t ? t->split_shade : 0u; }

// The mesh-tree stack entries the path launches take: the 4-wave forest
// build's 24, the 3-wave build's 24 or 32 as the autotune chose, 32 otherwise.
int phr_tracer_set_path_stack(phr_tracer* t, uint32_t entries) {
    PHR_CHECK(t, PHR_ERR_ARG, "phr_tracer_set_path_stack: null");
    PHR_CHECK(entries == FOREST_SHORT_STACK || entries == PHR_STACK, PHR_ERR_ARG, "phr_tracer_set_path_stack: 24 or 32 entries");
    PHR_CHECK(entries == PHR_STACK || (t->forest && t->mesh_depth <= FOREST_SHORT_STACK), PHR_ERR_UNSUPPORTED,
              "phr_tracer_set_path_stack: the short stack serves a forest whose mesh trees are at most 24 deep");
    t->short_stack = entries == FOREST_SHORT_STACK;
    return PHR_OK;
}

uint32_t phr_tracer_path_stack(const phr_tracer* t) {
    if (!t) return 0u;
    if (t->forest && path_build_waves(t) >= 4) return FOREST_SHORT_STACK;
    if (t->forest && path_build_waves(t) == 3 && path_short_stack(t)) return FOREST_SHORT_STACK;
    return PHR_STACK;
}

// Time every path kernel variant on the caller's workload (`spp` samples of
// w x h under `cam`, best of `reps` launches each) and keep the fastest. The
// image starts over; the variants compute the same sums.
int phr_tracer_autotune(phr_tracer* t, const void* cam, uint32_t w, uint32_t h, uint32_t spp, uint32_t reps,
                        uint32_t* chosen, double* best_s) {
    PHR_CHECK(t && cam, PHR_ERR_ARG, "phr_tracer_autotune: null");
    reps = std::max<uint32_t>(reps, 1);
    uint32_t best_w = 2, best_r = 0;
    double best = 1e30;
    PhrCamera probe;
    std::memcpy(&probe, cam, sizeof probe);
    // The builds this camera and scene allow: nested always; the split build
    // where the camera allows the segment shading; regenerating and
    // wavefront for one tree only.
    uint32_t builds[4] = {0u, 0u, 0u, 0u}, nbuilds = 1;
    if (regen_allowed(probe)) {
        if (!t->forest) {
            builds[nbuilds++] = 1u;
            builds[nbuilds++] = 2u;
        }
        builds[nbuilds++] = 3u;
    }
    const bool fits_short = t->forest && t->mesh_depth <= FOREST_SHORT_STACK;
    bool best_short = false;
    uint32_t best_shade = t->split_shade;
    for (uint32_t bi = 0; bi < nbuilds; ++bi)
        // The wavefront build's kernels have one register fit: timed once,
        // keeping the waves the others chose for its recorded-frame stand-in.
        for (uint32_t regen = builds[bi], waves = regen == 2u ? best_w : PATH_WAVES_MIN;
             waves <= (regen == 2u ? best_w : (t->forest ? (fits_short ? 4u : 3u) : PATH_WAVES_MAX)); ++waves)
            for (uint32_t stk = 0; stk < ((fits_short && waves == 3) ? 2u : 1u); ++stk)
                for (uint32_t shade : {4u, 6u}) {  // the split build's shading point; one pass for the others
                    if (regen != 3u && shade != 4u) continue;
                    t->path_waves = waves;
                    t->path_regen = regen;
                    t->short_stack = stk == 1;
                    if (regen == 3u) t->split_shade = shade;
                    double mine = 1e30;
                    for (uint32_t r = 0; r <= reps; ++r) {  // the first launch warms the variant
                        t->have_last = false;  // start the image over every launch
                        int rc = phr_tracer_sample(t, cam, w, h, spp, nullptr, nullptr);
                        if (rc != PHR_OK) return rc;
                        if (r > 0) mine = std::min(mine, t->last_gpu_s);
                    }
                    if (mine < best) {
                        best = mine, best_w = waves, best_r = regen, best_short = stk == 1;
                        if (regen == 3u) best_shade = shade;
                    }
                }
    t->path_regen = best_r;
    t->path_waves = best_w;
    t->short_stack = best_short;
    t->split_shade = best_shade;
    t->have_last = false;
    if (chosen) *chosen = best_w;
    if (best_s) *best_s = best;
    return PHR_OK;
}

// Pixels Adaptive Sampling has not settled (Tracer::pixels_left,
// src/trace/sample.rs): the list's length. PHR_ERR_UNSUPPORTED when the last
// camera had it off.
int phr_tracer_open_pixels(phr_tracer* t, uint32_t* count) {
    PHR_CHECK(t && count, PHR_ERR_ARG, "phr_tracer_open_pixels: null");
    PHR_CHECK(t->px && t->have_last && t->last.adapt[3] != 0.0f, PHR_ERR_UNSUPPORTED,
              "phr_tracer_open_pixels: adaptive sampling is off");
    return phr_array_read_range(t->px, 34ull * t->w * t->h, 1, count);
}

// Tracer::time (src/trace/sample.rs:253-255): submit and GPU seconds of the last sample call.
int phr_tracer_time(phr_tracer* t, double* submit_s, double* gpu_s) {
    PHR_CHECK(t, PHR_ERR_ARG, "phr_tracer_time: null");
    if (submit_s) *submit_s = t->last_submit_s;
    if (gpu_s) *gpu_s = t->last_gpu_s;
    return PHR_OK;
}

// Rays traced since the tracer was made: closest, shadow.
int phr_tracer_rays(phr_tracer* t, uint64_t* closest, uint64_t* shadow) {
    uint64_t c[6];
    int rc = phr_tracer_counters(t, c);
    if (rc != PHR_OK) return rc;
    if (closest) *closest = c[0];
    if (shadow) *shadow = c[1];
    return PHR_OK;
}

// All walk counters: path kernel [closest rays, shadow rays, inner steps,
// triangle tests], raycasts [inner steps, triangle tests].
int phr_tracer_counters(phr_tracer* t, uint64_t* out6) {
    PHR_CHECK(t && out6, PHR_ERR_ARG, "phr_tracer_counters: null");
    unsigned long long r[6];
    PHR_HIP(hipMemcpyAsync(r, t->rays, 48, hipMemcpyDeviceToHost, t->gpu->stream));
    PHR_HIP(hipStreamSynchronize(t->gpu->stream));
    for (int k = 0; k < 6; ++k) out6[k] = r[k];
    return PHR_OK;
}

// The raw image buffer words [from, from + count) (src/trace/read.rs:37-56 layout).
int phr_tracer_px(phr_tracer* t, uint64_t from, uint64_t count, uint32_t* out) {
    PHR_CHECK(t && t->px, PHR_ERR_ARG, "phr_tracer_px: nothing traced yet");
    return phr_array_read_range(t->px, from, count, out);
}

phr_array* phr_tracer_image(phr_tracer* t) { return t ? t->px : nullptr; }

namespace {

// The frame the learned denoiser reads: per pixel the radiance, albedo,
// normal and depth sums over the pixel's count (src/trace/read.rs:241-307),
// ten floats in that order.
__global__ __launch_bounds__(256) void k_frame(const uint32_t* __restrict__ px, uint64_t n, float* __restrict__ out) {
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const uint32_t c = px[31 * n + i];
    const float k = (float)(c > 1u ? c : 1u);
    float* o = out + 10 * i;
    for (int j = 0; j < 3; ++j) {
        o[j] = __uint_as_float(px[1 * n + 3 * i + j]) / k;
        o[3 + j] = __uint_as_float(px[7 * n + 3 * i + j]) / k;
        o[6 + j] = __uint_as_float(px[10 * n + 3 * i + j]) / k;
    }
    o[9] = __uint_as_float(px[13 * n + i]) / k;
}

// A mean image back into the linear sums (sum = value * count), and the
// display plane and packed word from the new sums as the tracer writes them
// (util.wgsl:139-144, show.wgsl:26-45) under the camera's view words: the learned
// denoiser's output, or a host's image for Tracer::show.
__global__ __launch_bounds__(256) void k_frame_back(uint32_t* __restrict__ px, uint64_t n, const float* __restrict__ rad,
                                                    float4 view) {
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const uint32_t c = px[31 * n + i];
    const float k = (float)(c > 1u ? c : 1u);
    const V3 l = v3(rad[3 * i] * k, rad[3 * i + 1] * k, rad[3 * i + 2] * k);
    stf(px, n + 3 * i, l.x);
    stf(px, n + 3 * i + 1, l.y);
    stf(px, n + 3 * i + 2, l.z);
    const V3 gs = v3(ldf(px, 20 * n + 3 * i), ldf(px, 20 * n + 3 * i + 1), ldf(px, 20 * n + 3 * i + 2));
    const V3 sh = shown_of(l, gs, ldf(px, 23 * n + i), ldf(px, 33 * n + i), k, view);
    stf(px, 4 * n + 3 * i, sh.x);
    stf(px, 4 * n + 3 * i + 1, sh.y);
    stf(px, 4 * n + 3 * i + 2, sh.z);
    px[i] = pack_rgba8(sh, ldf(px, 24 * n + i) / k);
}

// The a-trous passes' guide table for the current image size.
int ensure_denoise_scratch(phr_tracer* t) {
    const uint64_t n = (uint64_t)t->w * t->h;
    if (t->dn_scratch && t->dn_px == n) return PHR_OK;
    if (t->dn_scratch) (void)hipFree(t->dn_scratch);
    t->dn_scratch = nullptr;
    t->dn_px = 0;
    PHR_HIP(hipMalloc((void**)&t->dn_scratch, std::max<uint64_t>(n, 1) * DENOISE_SCRATCH_WORDS * 4));
    t->dn_px = n;
    return PHR_OK;
}

int ensure_frame(phr_tracer* t) {
    const uint64_t n = (uint64_t)t->w * t->h;
    if (t->frame_dev && t->frame_px == n) return PHR_OK;
    if (t->frame_dev) (void)hipFree(t->frame_dev);
    t->frame_dev = nullptr;
    t->frame_px = 0;
    PHR_HIP(hipMalloc((void**)&t->frame_dev, n * 13 * 4));
    t->frame_px = n;
    return PHR_OK;
}

}  // namespace

int phr_kpcn_infer_dev(phr_kpcn* k, const float* frames, uint32_t b, uint32_t w, uint32_t h, float* out);

// The denoiser's input frame of what has been traced, 10 floats a pixel,
// written at item `at` (in floats) of `out`, so a batch fills image by image.
int phr_tracer_frame(phr_tracer* t, phr_array* out, uint64_t at) {
    PHR_CHECK(t && t->px && out, PHR_ERR_ARG, "phr_tracer_frame: nothing traced yet");
    const uint64_t n = (uint64_t)t->w * t->h;
    PHR_CHECK((at + 10 * n) * 4 <= out->len * out->elem, PHR_ERR_SIZE, "phr_tracer_frame: output too small");
    hipLaunchKernelGGL(k_frame, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, t->gpu->stream,
                       (const uint32_t*)t->px->ptr, n, (float*)out->ptr + at);
    PHR_HIP(hipGetLastError());
    PHR_HIP(hipStreamSynchronize(t->gpu->stream));
    return PHR_OK;
}

const void* phr_kpcn_internal_workspace(const phr_kpcn* k);

namespace {
// Field by field: the struct's tail padding is never written, so a byte
// compare of two copies can differ where the scenes do not.
bool same_scene(const SceneDev& a, const SceneDev& b) {
    return a.nodes == b.nodes && a.tris == b.tris && a.faces == b.faces && a.verts == b.verts && a.mats == b.mats &&
           a.lights == b.lights && a.insts == b.insts && a.ninsts == b.ninsts && a.root == b.root;
}
}  // namespace

// One frame as a hipGraph (no counterpart in phiray, whose Pass replays per
// dispatch): `spp` samples of a fresh image under `cam` at w x h, then the
// a-trous denoiser (denoise 1, keep = false) or the learned one (denoise 2,
// net k) in place of the linear sums. Kernels only, no memory sets. Replays
// draw the same samples; the graph holds the scene's, image's and network's
// device addresses, so a replay after any of them moved is refused.
int phr_tracer_record_frame(phr_tracer* t, const void* cam_v, uint32_t w, uint32_t h, uint32_t spp, int denoise, phr_kpcn* k) {
    PHR_CHECK(t && cam_v && spp >= 1 && denoise >= 0 && denoise <= 2 && (denoise != 2 || k), PHR_ERR_ARG,
              "phr_tracer_record_frame: bad argument");
    PhrCamera cam;
    std::memcpy(&cam, cam_v, sizeof cam);
    PHR_CHECK(cam.adapt[3] == 0.0f, PHR_ERR_UNSUPPORTED, "phr_tracer_record_frame: adaptive sampling changes its launches");
    // An eager frame first: it stamps the camera, sizes the image, and gives
    // the denoisers their workspaces, so the capture allocates nothing.
    {
        PhrCamera fresh = cam;
        t->have_last = false;
        int rc = phr_tracer_sample(t, &fresh, w, h, spp, nullptr, nullptr);
        if (rc != PHR_OK) return rc;
        if (denoise == 2 && (rc = phr_tracer_denoise_learned(t, k, nullptr)) != PHR_OK) return rc;
        if (denoise == 1 && (rc = ensure_denoise_scratch(t)) != PHR_OK) return rc;
    }
    if (t->fexec) (void)hipGraphExecDestroy(t->fexec);
    if (t->fgraph) (void)hipGraphDestroy(t->fgraph);
    t->fexec = nullptr;
    t->fgraph = nullptr;
    const uint64_t n = (uint64_t)t->w * t->h;
    hipStream_t s = t->gpu->stream;
    const SceneDev sc = scene_dev(t);
    const uint32_t tiles_x = (t->w + 15) / 16, tiles = tiles_x * ((t->h + 15) / 16);
    const PhrCamera* cp = (const PhrCamera*)t->cam_dev;
    uint32_t* pxp = (uint32_t*)t->px->ptr;
    PHR_HIP(hipStreamSynchronize(s));
    PHR_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed));
    launch_path(path_build_waves(t), t->forest, regen_allowed(cam) ? t->path_regen : 0u, dim3(tiles), s, sc, cp, tiles_x, 0u, spp, pxp,
                t->rays, 0u, 0u, path_short_stack(t), t->split_shade);
    int rc = PHR_OK;
    if (denoise == 1) rc = phr_denoise_enqueue(t->gpu, pxp, t->w, t->h, 0, t->dn_scratch, t->last.view);
    if (denoise == 2) {
        const uint32_t g = (uint32_t)((n + 255) / 256);
        float* rad = t->frame_dev + 10 * n;
        hipLaunchKernelGGL(k_frame, dim3(g), dim3(256), 0, s, (const uint32_t*)pxp, n, t->frame_dev);
        rc = phr_kpcn_infer_dev(k, t->frame_dev, 1, t->w, t->h, rad);
        if (rc == PHR_OK)
            hipLaunchKernelGGL(k_frame_back, dim3(g), dim3(256), 0, s, pxp, n, (const float*)rad, view4(t->last));
    }
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamEndCapture(s, &graph);
    if (rc != PHR_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
    }
    PHR_CHECK(e == hipSuccess && graph, PHR_ERR_HIP, "phr_tracer_record_frame: capture failed");
    hipGraphExec_t exec = nullptr;
    if (hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess) {
        (void)hipGraphDestroy(graph);
        phr_set_error_msg("phr_tracer_record_frame: instantiation failed");
        return PHR_ERR_HIP;
    }
    t->fgraph = graph;
    t->fexec = exec;
    t->fcam = t->last;  // as stamped
    t->fw = t->w, t->fh = t->h, t->fspp = spp;
    t->fsc = sc;
    t->fsig[0] = t->px->ptr;
    t->fsig[1] = denoise == 2 ? (const void*)t->frame_dev : denoise == 1 ? (const void*)t->dn_scratch : nullptr;
    t->fsig[2] = denoise == 2 ? phr_kpcn_internal_workspace(k) : nullptr;
    t->fsig[3] = t->cam_dev;
    t->fdenoise = denoise;
    t->fposed = false;
    return PHR_OK;
}

// Replay the recorded frame: the image holds its samples (denoised as
// recorded) afterwards, and a later sample call under the same camera adds to them.
int phr_tracer_replay_frame(phr_tracer* t, phr_kpcn* k) {
    PHR_CHECK(t && t->fexec && !t->fposed, PHR_ERR_ARG, "phr_tracer_replay_frame: nothing recorded (a posed frame replays with its own call)");
    const bool moved = !t->px || t->px->ptr != t->fsig[0] || t->w != t->fw || t->h != t->fh || !same_scene(scene_dev(t), t->fsc) ||
                       (t->fdenoise == 2 && (t->frame_dev != t->fsig[1] || !k || phr_kpcn_internal_workspace(k) != t->fsig[2])) ||
                       (t->fdenoise == 1 && t->dn_scratch != t->fsig[1]);
    PHR_CHECK(!moved, PHR_ERR_ARG, "phr_tracer_replay_frame: the scene, the image or the network moved; record again");
    hipStream_t s = t->gpu->stream;
    PHR_HIP(hipMemcpyAsync(t->cam_dev, &t->fcam, sizeof(PhrCamera), hipMemcpyHostToDevice, s));
    PHR_HIP(hipEventRecord(t->ev0, s));
    PHR_HIP(hipGraphLaunch(t->fexec, s));
    PHR_HIP(hipEventRecord(t->ev1, s));
    PHR_HIP(hipEventSynchronize(t->ev1));
    float ms = 0.f;
    PHR_HIP(hipEventElapsedTime(&ms, t->ev0, t->ev1));
    t->last_gpu_s = ms * 1e-3;
    t->w = t->fw, t->h = t->fh;
    t->last = t->fcam;
    t->have_last = true;
    t->frame = t->fspp - 1;
    return PHR_OK;
}

// A posed frame recorded as one graph: the new vertices when `verts` is
// given (triangle boxes, the mesh trees' refit, the triangle cache), the new
// matrices (world boxes, the top tree, the instance and emitter tables), the
// camera's stamp on the device, the path launch and the denoiser: what
// reinstance / repose_on and then sample (+ denoise) do one call after
// another (src/trace/pose.rs:77-155), with no host wait between them. An
// eager posed frame first sizes the image, stamps the camera and gives the
// denoisers their workspaces. A forest whose emitter table is built on the
// device only (at most 4,096 candidates, at least one).
int phr_tracer_record_pose_frame(phr_tracer* t, phr_array* verts, const float* worlds, uint32_t n, const void* cam_v, uint32_t w,
                                 uint32_t h, uint32_t spp, int denoise, phr_kpcn* k) {
    PHR_CHECK(t && worlds && cam_v && spp >= 1 && denoise >= 0 && denoise <= 2 && (denoise != 2 || k), PHR_ERR_ARG,
              "phr_tracer_record_pose_frame: bad argument");
    PHR_CHECK(t->forest && n == t->ninsts, PHR_ERR_ARG, "phr_tracer_record_pose_frame: a forest and one matrix an instance");
    PHR_CHECK(!verts || (verts->elem == 32 && verts->len == t->nverts), PHR_ERR_ARG,
              "phr_tracer_record_pose_frame: vertices are the scene's 32-byte records");
    PHR_CHECK(t->ncand > 0 && t->ncand <= LIGHT_DEVICE_MAX, PHR_ERR_UNSUPPORTED,
              "phr_tracer_record_pose_frame: the emitter table must be built on the device (1 to 4,096 candidates)");
    PhrCamera cam;
    std::memcpy(&cam, cam_v, sizeof cam);
    PHR_CHECK(cam.adapt[3] == 0.0f, PHR_ERR_UNSUPPORTED, "phr_tracer_record_pose_frame: adaptive sampling changes its launches");
    if (!t->pin_worlds) {
        PHR_HIP(hipHostMalloc((void**)&t->pin_worlds, 64ull * n, hipHostMallocDefault));
        PHR_HIP(hipHostMalloc((void**)&t->pin_words, 4ull * (INSTANCE_LEN + 2) * n, hipHostMallocDefault));
    }
    int ok = 0, rc;
    if (verts && ((rc = phr_tracer_repose_on(t, verts, &ok)) != PHR_OK || !ok)) return rc != PHR_OK ? rc : PHR_ERR_ARG;
    if ((rc = phr_tracer_reinstance(t, worlds, n, &ok)) != PHR_OK || !ok) return rc != PHR_OK ? rc : PHR_ERR_ARG;
    {
        PhrCamera fresh = cam;
        t->have_last = false;
        if ((rc = phr_tracer_sample(t, &fresh, w, h, spp, nullptr, nullptr)) != PHR_OK) return rc;
        if (denoise == 2 && (rc = phr_tracer_denoise_learned(t, k, nullptr)) != PHR_OK) return rc;
        if (denoise == 1 && (rc = ensure_denoise_scratch(t)) != PHR_OK) return rc;
    }
    if (t->fexec) (void)hipGraphExecDestroy(t->fexec);
    if (t->fgraph) (void)hipGraphDestroy(t->fgraph);
    t->fexec = nullptr;
    t->fgraph = nullptr;
    const uint64_t np = (uint64_t)t->w * t->h;
    const uint32_t nf = t->n;
    hipStream_t s = t->gpu->stream;
    const SceneDev sc = scene_dev(t);
    const uint32_t tiles_x = (t->w + 15) / 16, tiles = tiles_x * ((t->h + 15) / 16);
    PhrCamera* cp = (PhrCamera*)t->cam_dev;
    uint32_t* pxp = (uint32_t*)t->px->ptr;
    PHR_HIP(hipStreamSynchronize(s));
    PHR_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed));
    rc = PHR_OK;
    if (verts) {
        (void)hipMemcpyAsync(t->verts->ptr, verts->ptr, 32ull * t->nverts, hipMemcpyDeviceToDevice, s);
        hipLaunchKernelGGL(k_tri_boxes, dim3((nf + 255) / 256), dim3(256), 0, s, (const uint4*)t->in_faces->ptr,
                           (const float4*)t->verts->ptr, nf, (float4*)t->prims->ptr);
        rc = phr_bvh_internal_refit_async(t->bvh);
        if (rc == PHR_OK)
            hipLaunchKernelGGL(k_gather, dim3((nf + 255) / 256), dim3(256), 0, s, (const uint4*)t->in_faces->ptr,
                               (const float4*)t->verts->ptr, (const uint32_t*)phr_bvh_internal_order(t->bvh)->ptr, nf,
                               (uint4*)t->faces->ptr, (float4*)t->tris->ptr);
    }
    if (rc == PHR_OK) rc = forest_top(t, false);
    if (rc == PHR_OK) rc = forest_lights(t);
    if (rc == PHR_OK) {
        hipLaunchKernelGGL(k_stamp_pose, dim3(1), dim3(64), 0, s, (const float4*)t->nodes->ptr, (const uint32_t*)t->lights->ptr, 1u,
                           cam.ao <= 0.f ? 1u : 0u, cp);
        launch_path(path_build_waves(t), t->forest, 0u, dim3(tiles), s, sc, cp, tiles_x, 0u, spp, pxp, t->rays, 0u, 0u,
                    path_short_stack(t));
        if (denoise == 1) rc = phr_denoise_enqueue(t->gpu, pxp, t->w, t->h, 0, t->dn_scratch, t->last.view);
        if (denoise == 2) {
            const uint32_t g = (uint32_t)((np + 255) / 256);
            float* rad = t->frame_dev + 10 * np;
            hipLaunchKernelGGL(k_frame, dim3(g), dim3(256), 0, s, (const uint32_t*)pxp, np, t->frame_dev);
            rc = phr_kpcn_infer_dev(k, t->frame_dev, 1, t->w, t->h, rad);
            if (rc == PHR_OK)
                hipLaunchKernelGGL(k_frame_back, dim3(g), dim3(256), 0, s, pxp, np, (const float*)rad, view4(t->last));
        }
    }
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamEndCapture(s, &graph);
    if (rc != PHR_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
    }
    PHR_CHECK(e == hipSuccess && graph, PHR_ERR_HIP, "phr_tracer_record_pose_frame: capture failed");
    hipGraphExec_t exec = nullptr;
    if (hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess) {
        (void)hipGraphDestroy(graph);
        phr_set_error_msg("phr_tracer_record_pose_frame: instantiation failed");
        return PHR_ERR_HIP;
    }
    t->fgraph = graph;
    t->fexec = exec;
    t->fcam = t->last;  // as stamped; the graph's own stamp updates ao, eps and glow on the device
    t->fw = t->w, t->fh = t->h, t->fspp = spp;
    t->fsc = sc;
    t->fsig[0] = t->px->ptr;
    t->fsig[1] = denoise == 2 ? (const void*)t->frame_dev : denoise == 1 ? (const void*)t->dn_scratch : nullptr;
    t->fsig[2] = denoise == 2 ? phr_kpcn_internal_workspace(k) : nullptr;
    t->fsig[3] = t->cam_dev;
    t->fdenoise = denoise;
    t->fposed = true;
    t->fverts = verts ? verts->ptr : nullptr;
    return PHR_OK;
}

// Replay the recorded posed frame under new matrices (one an instance; the
// vertices are read again from the array given at recording). The image is
// the new pose's, traced and denoised as recorded; the next sample call
// starts the image over.
int phr_tracer_replay_pose_frame(phr_tracer* t, const float* worlds, uint32_t n, phr_kpcn* k) {
    PHR_CHECK(t && worlds && t->fexec && t->fposed, PHR_ERR_ARG, "phr_tracer_replay_pose_frame: no posed frame recorded");
    PHR_CHECK(n == t->ninsts, PHR_ERR_ARG, "phr_tracer_replay_pose_frame: one matrix an instance");
    const bool moved = !t->px || t->px->ptr != t->fsig[0] || t->w != t->fw || t->h != t->fh || !same_scene(scene_dev(t), t->fsc) ||
                       (t->fdenoise == 2 && (t->frame_dev != t->fsig[1] || !k || phr_kpcn_internal_workspace(k) != t->fsig[2])) ||
                       (t->fdenoise == 1 && t->dn_scratch != t->fsig[1]);
    PHR_CHECK(!moved, PHR_ERR_ARG, "phr_tracer_replay_pose_frame: the scene, the image or the network moved; record again");
    for (uint32_t i = 0; i < n; ++i) std::memcpy(t->inst_in[i].world, worlds + 16 * (size_t)i, 64);
    const float* wsrc = nullptr;
    const uint32_t* isrc = nullptr;
    forest_pose_host(t, &wsrc, &isrc);  // into the pinned buffers the graph's copies read
    hipStream_t s = t->gpu->stream;
    PHR_HIP(hipMemcpyAsync(t->cam_dev, &t->fcam, sizeof(PhrCamera), hipMemcpyHostToDevice, s));
    PHR_HIP(hipEventRecord(t->ev0, s));
    PHR_HIP(hipGraphLaunch(t->fexec, s));
    PHR_HIP(hipEventRecord(t->ev1, s));
    PHR_HIP(hipEventSynchronize(t->ev1));
    float ms = 0.f;
    PHR_HIP(hipEventElapsedTime(&ms, t->ev0, t->ev1));
    t->last_gpu_s = ms * 1e-3;
    t->w = t->fw, t->h = t->fh;
    t->stale = true;       // the extent and emitter count are the device's now
    t->have_last = false;  // the next sample starts over
    t->frame = t->fspp - 1;
    return PHR_OK;
}

// Tracer::show (src/trace/read.rs:182-229, src/trace/wgsl/show.wgsl): a
// host's scene-linear image put back in place of the traced mean (sum = value
// times the pixel's own count) and run through the display step. `linear`
// holds `count` rgb triples, truncated or zero-filled to the image's size;
// *shown = 0, as the reference's false, when there is no image or camera yet.
int phr_tracer_show(phr_tracer* t, const float* linear, uint64_t count, int* shown) {
    PHR_CHECK(t && shown && (linear || count == 0), PHR_ERR_ARG, "phr_tracer_show: null argument");
    *shown = 0;
    const uint64_t n = (uint64_t)t->w * t->h;
    if (!t->px || !t->have_last || n == 0) return PHR_OK;
    int rc = ensure_frame(t);
    if (rc != PHR_OK) return rc;
    float* rad = t->frame_dev + 10 * n;
    const uint64_t m = std::min<uint64_t>(count, n);
    hipStream_t s = t->gpu->stream;
    if (m) PHR_HIP(hipMemcpyAsync(rad, linear, m * 12, hipMemcpyHostToDevice, s));
    if (m < n) PHR_HIP(hipMemsetAsync(rad + 3 * m, 0, (n - m) * 12, s));
    const uint32_t g = (uint32_t)((n + 255) / 256);
    hipLaunchKernelGGL(k_frame_back, dim3(g), dim3(256), 0, s, (uint32_t*)t->px->ptr, n, (const float*)rad, view4(t->last));
    PHR_HIP(hipGetLastError());
    PHR_HIP(hipStreamSynchronize(s));
    *shown = 1;
    return PHR_OK;
}

// Tracer::with_chunks (src/trace/build.rs:136-200): Tracer::from_scene with
// phiray's per-binding byte limit and binding count (2, or 4 when more than
// 2 are asked for) checked as phiray checks them: PHR_ERR_SIZE, the
// reference's None, when the scene's largest array (vertices 32 B, faces
// 16 B, materials) passes limit x chunks, or the nodes, faces or vertices do
// not split over the bindings.
int phr_tracer_with_chunks(phr_gpu* g, const phr_scene_desc* d, uint32_t leaf, uint64_t limit, uint32_t chunks,
                           phr_tracer** out) {
    PHR_CHECK(g && d && out && limit > 0, PHR_ERR_ARG, "phr_tracer_with_chunks: bad argument");
    *out = nullptr;
    const uint32_t c = chunks > 2 ? 4u : 2u;
    const uint64_t biggest = std::max<uint64_t>({(uint64_t)d->nverts * 32, (uint64_t)d->nfaces * 16, (uint64_t)d->nmats * 64 + 64});
    PHR_CHECK(biggest <= limit * c, PHR_ERR_SIZE, "phr_tracer_with_chunks: the scene's largest array passes the bindings");
    phr_tracer* t = nullptr;
    int rc = phr_tracer_from_scene(g, d, leaf, &t);
    if (rc != PHR_OK) return rc;
    const uint32_t nc = split_cut(t->nodes->len, 32, limit, c), fc = split_cut(d->nfaces, 16, limit, c),
                   vc = split_cut(d->nverts, 32, limit, c);
    if (!nc || !fc || !vc) {
        phr_tracer_free(t);
        phr_set_error_msg("phr_tracer_with_chunks: the nodes, faces or vertices do not split over the bindings");
        return PHR_ERR_SIZE;
    }
    t->split[0] = nc;
    t->split[1] = fc;
    t->split[2] = vc;
    t->split[3] = 0;
    t->chunks = c;
    *out = t;
    return PHR_OK;
}

// Tracer::split / Tracer::chunks (src/trace/mod.rs:118-127).
int phr_tracer_split(phr_tracer* t, uint32_t out[4]) {
    PHR_CHECK(t && out, PHR_ERR_ARG, "phr_tracer_split: null argument");
    std::memcpy(out, t->split, sizeof t->split);
    return PHR_OK;
}
uint32_t phr_tracer_chunks(phr_tracer* t) { return t ? t->chunks : 0u; }

// Tracer::size (src/trace/read.rs:162-164): the image's size, 0 x 0 before
// the first sample.
int phr_tracer_size(phr_tracer* t, uint32_t* w, uint32_t* h) {
    PHR_CHECK(t && w && h, PHR_ERR_ARG, "phr_tracer_size: null argument");
    *w = t->px ? t->w : 0u;
    *h = t->px ? t->h : 0u;
    return PHR_OK;
}

// The learned denoiser over what has been traced (no counterpart in phiray; it
// stands beside Tracer::denoise). With `out`, the denoised radiance goes there
// (w * h * 3 floats) and the image is left as it is; without, it replaces the
// linear sums, as Tracer::denoise with keep = false does.
int phr_tracer_denoise_learned(phr_tracer* t, phr_kpcn* k, phr_array* out) {
    PHR_CHECK(t && k, PHR_ERR_ARG, "phr_tracer_denoise_learned: null");
    if (!t->px) return PHR_OK;
    const uint64_t n = (uint64_t)t->w * t->h;
    if (out) PHR_CHECK(out->len * out->elem >= n * 12, PHR_ERR_SIZE, "phr_tracer_denoise_learned: output too small");
    int rc = ensure_frame(t);
    if (rc != PHR_OK) return rc;
    const uint32_t g = (uint32_t)((n + 255) / 256);
    float* rad = out ? (float*)out->ptr : t->frame_dev + 10 * n;
    hipLaunchKernelGGL(k_frame, dim3(g), dim3(256), 0, t->gpu->stream, (const uint32_t*)t->px->ptr, n, t->frame_dev);
    PHR_HIP(hipGetLastError());
    rc = phr_kpcn_infer_dev(k, t->frame_dev, 1, t->w, t->h, rad);
    if (rc != PHR_OK) return rc;
    if (!out) {
        hipLaunchKernelGGL(k_frame_back, dim3(g), dim3(256), 0, t->gpu->stream, (uint32_t*)t->px->ptr, n, rad, view4(t->last));
        PHR_HIP(hipGetLastError());
    }
    PHR_HIP(hipStreamSynchronize(t->gpu->stream));
    return PHR_OK;
}

// Tracer::denoise / denoise_shown (src/trace/read.rs:73-119): four a-trous
// passes over what has been traced; `keep` leaves the linear sums as they are.
int phr_tracer_denoise(phr_tracer* t, int keep) {
    PHR_CHECK(t, PHR_ERR_ARG, "phr_tracer_denoise: null");
    if (!t->px) return PHR_OK;  // nothing traced: nothing to filter
    int rc = ensure_denoise_scratch(t);
    if (rc != PHR_OK) return rc;
    rc = phr_denoise_enqueue(t->gpu, (uint32_t*)t->px->ptr, t->w, t->h, keep, t->dn_scratch, t->last.view);
    if (rc != PHR_OK) return rc;
    PHR_HIP(hipStreamSynchronize(t->gpu->stream));
    return PHR_OK;
}

// Tracer::radiance (src/trace/read.rs:241-251): the scene-linear sums over each
// pixel's own sample count.
int phr_tracer_radiance(phr_tracer* t, float* out) {
    PHR_CHECK(t && t->px && out, PHR_ERR_ARG, "phr_tracer_radiance: nothing traced yet");
    uint64_t n = (uint64_t)t->w * t->h;
    std::vector<uint32_t> sums(3 * n), cnt(n);
    int rc = phr_array_read_range(t->px, n, 3 * n, sums.data());
    if (rc == PHR_OK) rc = phr_array_read_range(t->px, 31 * n, n, cnt.data());
    if (rc != PHR_OK) return rc;
    for (uint64_t p = 0; p < n; ++p) {
        float k = (float)std::max<uint32_t>(cnt[p], 1u);
        for (int c = 0; c < 3; ++c) {
            float v;
            std::memcpy(&v, &sums[3 * p + c], 4);
            out[3 * p + c] = v / k;
        }
    }
    return PHR_OK;
}

// The guide planes over the per-pixel count (src/trace/read.rs:253-307).
static int guide_plane(phr_tracer* t, uint64_t at, int comps, float* out, const char* what) {
    PHR_CHECK(t && t->px && out, PHR_ERR_ARG, what);
    uint64_t n = (uint64_t)t->w * t->h;
    std::vector<uint32_t> sums(comps * n), cnt(n);
    int rc = phr_array_read_range(t->px, at * n, comps * n, sums.data());
    if (rc == PHR_OK) rc = phr_array_read_range(t->px, 31 * n, n, cnt.data());
    if (rc != PHR_OK) return rc;
    for (uint64_t p = 0; p < n; ++p) {
        float k = (float)std::max<uint32_t>(cnt[p], 1u);
        for (int c = 0; c < comps; ++c) {
            float v;
            std::memcpy(&v, &sums[comps * p + c], 4);
            out[comps * p + c] = v / k;
        }
    }
    return PHR_OK;
}
// Tracer::pixels (src/trace/read.rs:322-327): the packed RGBA8 words, 4 bytes
// a pixel, red first, row 0 at the top.
int phr_tracer_pixels(phr_tracer* t, uint8_t* out) {
    PHR_CHECK(t && t->px && out, PHR_ERR_ARG, "phr_tracer_pixels: nothing traced yet");
    return phr_array_read_range(t->px, 0, (uint64_t)t->w * t->h, out);
}

// Tracer::display (src/trace/read.rs:257-263): the shown value, 3 floats a
// pixel, words 4n..7n as the tracer or the denoiser's last pass left them.
int phr_tracer_display(phr_tracer* t, float* out) {
    PHR_CHECK(t && t->px && out, PHR_ERR_ARG, "phr_tracer_display: nothing traced yet");
    const uint64_t n = (uint64_t)t->w * t->h;
    return phr_array_read_range(t->px, 4 * n, 3 * n, out);
}

int phr_tracer_albedo(phr_tracer* t, float* out) { return guide_plane(t, 7, 3, out, "phr_tracer_albedo: nothing traced yet"); }
int phr_tracer_normal(phr_tracer* t, float* out) { return guide_plane(t, 10, 3, out, "phr_tracer_normal: nothing traced yet"); }
int phr_tracer_depth_plane(phr_tracer* t, float* out) { return guide_plane(t, 13, 1, out, "phr_tracer_depth_plane: nothing traced yet"); }

// Tracer::from_scene (src/trace/build.rs:15-25) with the forest of
// src/trace/forest.rs:44-66: the faces' boxes in their meshes' spaces, one
// batched LBVH build for every mesh tree (written behind the top tree's
// slots), the faces into leaf order, each tree's root box back on the host,
// then the top tree over the placed instances' world boxes, the instance
// table and the light table.
int phr_tracer_from_scene(phr_gpu* g, const phr_scene_desc* d, uint32_t leaf, phr_tracer** out) {
    PHR_CHECK(g && d && out, PHR_ERR_ARG, "phr_tracer_from_scene: null argument");
    PHR_CHECK(d->nfaces > 0 && d->nmeshes > 0 && d->ninsts > 0 && d->verts && d->faces && d->mats && d->meshes && d->insts,
              PHR_ERR_ARG, "phr_tracer_from_scene: empty scene");
    PHR_CHECK(d->nmeshes <= 65536, PHR_ERR_SIZE, "phr_tracer_from_scene: at most 65,536 meshes");
    // The top tree's nodes (2 insts - 1) are walked with 16-bit stack entries.
    PHR_CHECK(d->ninsts <= 32768, PHR_ERR_SIZE, "phr_tracer_from_scene: at most 32,768 instances");
    const uint32_t* fcs = (const uint32_t*)d->faces;
    const float* mts = (const float*)d->mats;
    for (uint32_t f = 0; f < d->nfaces; ++f) {
        const uint32_t* fc = fcs + 4 * (size_t)f;
        PHR_CHECK(fc[0] < d->nverts && fc[1] < d->nverts && fc[2] < d->nverts && fc[3] < d->nmats, PHR_ERR_ARG,
                  "phr_tracer_from_scene: a face names a vertex or material that is not there");
    }
    for (uint32_t m = 0; m < d->nmats; ++m)
        PHR_CHECK(mts[MAT_WORDS * (size_t)m + 6] == 0.f, PHR_ERR_UNSUPPORTED, "phr_tracer_from_scene: transmission (glass) is not supported");
    for (uint32_t m = 0; m < d->nmeshes; ++m) {
        const phr_mesh& me = d->meshes[m];
        const uint32_t want = m == 0 ? 0u : d->meshes[m - 1].face_hi;
        PHR_CHECK(me.face_lo == want && me.face_lo <= me.face_hi && me.face_hi <= d->nfaces, PHR_ERR_ARG,
                  "phr_tracer_from_scene: the meshes must tile the faces in order");
    }
    PHR_CHECK(d->meshes[d->nmeshes - 1].face_hi == d->nfaces, PHR_ERR_ARG, "phr_tracer_from_scene: the meshes must cover every face");
    bool any_placed = false;
    for (uint32_t i = 0; i < d->ninsts && !any_placed; ++i) {
        const uint32_t m = d->insts[i].mesh;
        any_placed = m < d->nmeshes && d->meshes[m].face_hi > d->meshes[m].face_lo;
    }
    PHR_CHECK(any_placed, PHR_ERR_ARG, "phr_tracer_from_scene: no instance places a mesh with faces");
    phr_tracer* t = new phr_tracer();
    t->gpu = g;
    t->forest = true;
    t->n = d->nfaces;
    t->nverts = d->nverts;
    t->nmats = d->nmats;
    t->nmeshes = d->nmeshes;
    t->ninsts = d->ninsts;
    t->leaf = leaf;
    t->tlas_slots = std::max<uint32_t>(2 * d->ninsts - 1, 1);
    auto fail = [&](int rc) { tracer_release(t); return rc; };
    t->pos.assign((const float*)d->verts, (const float*)d->verts + 8 * (size_t)d->nverts);
    t->faces_in.assign(fcs, fcs + 4 * (size_t)d->nfaces);
    t->mats_h.assign(mts, mts + MAT_WORDS * (size_t)d->nmats);
    t->meshes.assign(d->meshes, d->meshes + d->nmeshes);
    t->inst_in.assign(d->insts, d->insts + d->ninsts);
    for (uint32_t i = 0; i < d->ninsts; ++i) {
        const uint32_t m = d->insts[i].mesh;
        if (m < d->nmeshes && d->meshes[m].face_hi > d->meshes[m].face_lo) t->placed.push_back(i);
    }
    std::vector<uint32_t> seg_lo(d->nmeshes + 1), bases(d->nmeshes, PHR_NONE);
    t->node_base.assign(d->nmeshes, 0u);
    uint64_t nodes = t->tlas_slots;
    for (uint32_t m = 0; m < d->nmeshes; ++m) {
        const uint32_t cnt = d->meshes[m].face_hi - d->meshes[m].face_lo;
        seg_lo[m] = d->meshes[m].face_lo;
        t->node_base[m] = (uint32_t)nodes;
        if (cnt) {
            bases[m] = (uint32_t)nodes;
            nodes += 2ull * cnt - 1;
        }
        bool glows = false;
        for (uint32_t f = d->meshes[m].face_lo; f < d->meshes[m].face_hi && !glows; ++f) {
            const float* e = mts + MAT_WORDS * (size_t)fcs[4 * (size_t)f + 3] + 8;
            glows = 0.2126f * e[0] + 0.7152f * e[1] + 0.0722f * e[2] > 0.f;
        }
        if (glows) t->glow_meshes.push_back(m);
    }
    seg_lo[d->nmeshes] = d->nfaces;
    PHR_CHECK(nodes < (1ull << 31), PHR_ERR_SIZE, "phr_tracer_from_scene: too many nodes");
    const uint32_t n = t->n, p = (uint32_t)t->placed.size();
    hipStream_t s = g->stream;
    int rc;
    if ((rc = phr_array_alloc(g, 16, n, &t->in_faces)) != PHR_OK) return fail(rc);
    if ((rc = phr_array_alloc(g, 32, t->nverts, &t->verts)) != PHR_OK) return fail(rc);
    if ((rc = phr_array_alloc(g, 4 * MAT_WORDS, t->nmats, &t->mats)) != PHR_OK) return fail(rc);
    if ((rc = phr_array_alloc(g, 32, n, &t->prims)) != PHR_OK) return fail(rc);
    if ((rc = phr_array_alloc(g, 16, n, &t->faces)) != PHR_OK) return fail(rc);
    if ((rc = phr_array_alloc(g, 48, n, &t->tris)) != PHR_OK) return fail(rc);
    if ((rc = phr_array_alloc(g, 32, nodes, &t->nodes)) != PHR_OK) return fail(rc);
    if ((rc = phr_array_alloc(g, 32, p, &t->wboxes)) != PHR_OK) return fail(rc);
    if ((rc = phr_array_alloc(g, 4, (size_t)(INSTANCE_LEN + 2) * t->ninsts, &t->insts)) != PHR_OK) return fail(rc);
    if ((rc = phr_array_alloc(g, 4, p, &t->placed_dev)) != PHR_OK) return fail(rc);
    if ((rc = phr_array_alloc(g, 4, 16ull * t->ninsts, &t->worlds_dev)) != PHR_OK) return fail(rc);
    if ((rc = phr_array_alloc(g, 4, t->ninsts, &t->inst_mesh)) != PHR_OK) return fail(rc);
    std::vector<uint32_t> mesh_of(t->ninsts);
    for (uint32_t i = 0; i < t->ninsts; ++i) mesh_of[i] = std::min(t->inst_in[i].mesh, t->nmeshes - 1);  // placed ones are valid
    if (hipMalloc(&t->rays, 48) != hipSuccess || hipMemsetAsync(t->rays, 0, 48, s) != hipSuccess ||
        hipMalloc(&t->cam_dev, sizeof(PhrCamera)) != hipSuccess || hipMalloc(&t->bases_dev, 4ull * t->nmeshes) != hipSuccess) {
        phr_set_error_msg("phr_tracer_from_scene: allocation failed");
        return fail(PHR_ERR_HIP);
    }
    if (hipMemcpyAsync(t->in_faces->ptr, fcs, 16ull * n, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(t->verts->ptr, d->verts, 32ull * t->nverts, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(t->mats->ptr, mts, 64ull * t->nmats, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(t->bases_dev, bases.data(), 4ull * t->nmeshes, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(t->placed_dev->ptr, t->placed.data(), 4ull * p, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(t->inst_mesh->ptr, mesh_of.data(), 4ull * t->ninsts, hipMemcpyHostToDevice, s) != hipSuccess) {
        phr_set_error_msg("phr_tracer_from_scene: upload failed");
        return fail(PHR_ERR_HIP);
    }
    hipLaunchKernelGGL(k_tri_boxes, dim3((n + 255) / 256), dim3(256), 0, s, (const uint4*)t->in_faces->ptr,
                       (const float4*)t->verts->ptr, n, (float4*)t->prims->ptr);
    if ((rc = phr_bvh_internal_build_into(g, t->prims, leaf, seg_lo.data(), t->nmeshes, t->nodes, t->node_base.data(), PHR_LEAF,
                                          &t->bvh)) != PHR_OK)
        return fail(rc);
    if ((rc = forest_mesh_depth(t)) != PHR_OK) return fail(rc);
    if ((rc = forest_gather(t)) != PHR_OK || (rc = forest_light_cands(t)) != PHR_OK || (rc = forest_top(t, true)) != PHR_OK ||
        (rc = forest_lights(t)) != PHR_OK || (rc = forest_settle(t)) != PHR_OK)
        return fail(rc);
    if (hipEventCreate(&t->ev0) != hipSuccess || hipEventCreate(&t->ev1) != hipSuccess) {
        phr_set_error_msg("phr_tracer_from_scene: event creation failed");
        return fail(PHR_ERR_HIP);
    }
    default_split(t, t->nodes->len, d->nfaces, d->nverts);
    *out = t;
    return PHR_OK;
}

int phr_tracer_reinstance(phr_tracer* t, const float* worlds, uint32_t n, int* ok) {
    PHR_CHECK(t && ok && (worlds || n == 0), PHR_ERR_ARG, "phr_tracer_reinstance: null argument");
    *ok = 0;
    if (!t->forest || n != t->ninsts) return PHR_OK;
    for (uint32_t i = 0; i < n; ++i) std::memcpy(t->inst_in[i].world, worlds + 16 * (size_t)i, 64);
    int rc;
    if ((rc = forest_top(t, false)) != PHR_OK || (rc = forest_lights(t)) != PHR_OK) return rc;
    t->have_last = false;  // the image starts again (pose.rs:117-119)
    t->frame = 0;
    *ok = 1;
    return PHR_OK;
}

// The moved vertices are in t->verts: boxes, refit, triangle cache, then the
// top tree and the emitter table, all queued on the device.
static int repose_device(phr_tracer* t) {
    hipStream_t s = t->gpu->stream;
    const uint32_t n = t->n;
    hipLaunchKernelGGL(k_tri_boxes, dim3((n + 255) / 256), dim3(256), 0, s, (const uint4*)t->in_faces->ptr,
                       (const float4*)t->verts->ptr, n, (float4*)t->prims->ptr);
    int rc = phr_bvh_internal_refit_async(t->bvh);
    if (rc != PHR_OK) return rc;
    hipLaunchKernelGGL(k_gather, dim3((n + 255) / 256), dim3(256), 0, s, (const uint4*)t->in_faces->ptr, (const float4*)t->verts->ptr,
                       (const uint32_t*)phr_bvh_internal_order(t->bvh)->ptr, n, (uint4*)t->faces->ptr, (float4*)t->tris->ptr);
    PHR_HIP(hipGetLastError());
    if ((rc = forest_top(t, false)) != PHR_OK || (rc = forest_lights(t)) != PHR_OK) return rc;
    t->have_last = false;  // pose.rs:92-94
    t->frame = 0;
    return PHR_OK;
}

int phr_tracer_repose(phr_tracer* t, const void* verts, uint32_t nverts, int* ok) {
    PHR_CHECK(t && ok && (verts || nverts == 0), PHR_ERR_ARG, "phr_tracer_repose: null argument");
    *ok = 0;
    if (!t->forest || nverts != t->nverts) return PHR_OK;
    std::memcpy(t->pos.data(), verts, 32ull * nverts);
    PHR_HIP(hipMemcpyAsync(t->verts->ptr, verts, 32ull * nverts, hipMemcpyHostToDevice, t->gpu->stream));
    int rc = repose_device(t);
    if (rc == PHR_OK) *ok = 1;
    return rc;
}

int phr_tracer_repose_on(phr_tracer* t, phr_array* verts, int* ok) {
    PHR_CHECK(t && ok && verts && verts->elem == 32, PHR_ERR_ARG, "phr_tracer_repose_on: vertices are 32-byte records");
    *ok = 0;
    if (!t->forest || verts->len != t->nverts) return PHR_OK;
    PHR_HIP(hipMemcpyAsync(t->verts->ptr, verts->ptr, 32ull * t->nverts, hipMemcpyDeviceToDevice, t->gpu->stream));
    int rc = repose_device(t);
    if (rc == PHR_OK) *ok = 1;
    return rc;
}

int phr_tracer_retree(phr_tracer* t, int* ok) {
    PHR_CHECK(t && ok, PHR_ERR_ARG, "phr_tracer_retree: null argument");
    *ok = 0;
    if (!t->forest) return PHR_OK;
    int rc;
    if ((rc = phr_bvh_internal_rebuild_async(t->bvh)) != PHR_OK || (rc = forest_mesh_depth(t)) != PHR_OK || (rc = forest_gather(t)) != PHR_OK ||
        (rc = forest_light_cands(t)) != PHR_OK || (rc = forest_top(t, false)) != PHR_OK || (rc = forest_lights(t)) != PHR_OK)
        return rc;
    *ok = 1;  // the image keeps its samples (pose.rs:14-16)
    return PHR_OK;
}

phr_array* phr_tracer_nodes(phr_tracer* t) { return t ? (t->forest ? t->nodes : phr_bvh_internal_nodes(t->bvh)) : nullptr; }
phr_array* phr_tracer_instances(phr_tracer* t) { return t && t->forest ? t->insts : nullptr; }
uint32_t phr_tracer_instance_count(const phr_tracer* t) { return t ? (t->forest ? t->ninsts : 1u) : 0u; }

}  // extern "C"

// ---- a batch of scenes (many Tracers' worth in one build and one launch) ----
struct phr_batch {
    phr_gpu* gpu = nullptr;
    uint32_t scenes = 0, n = 0, leaf = 4;
    std::vector<uint32_t> seg_lo, node_base, light_at, glow;
    std::vector<float> extent;
    phr_array *in_faces = nullptr, *verts = nullptr, *mats = nullptr, *prims = nullptr, *faces = nullptr, *tris = nullptr,
              *lights = nullptr, *px = nullptr, *recs = nullptr;
    phr_bvh* bvh = nullptr;
    SceneDev* table = nullptr;    // one SceneDev a scene
    PhrCamera* cams = nullptr;    // one stamped camera a scene (device)
    uint32_t cams_cap = 0;
    PhrCamera* gcams = nullptr;   // the recorded pass's own cameras (eager renders leave them alone)
    unsigned long long* rays = nullptr;  // closest, shadow, inner steps, triangle tests
    uint32_t w = 0, h = 0;
    uint32_t path_waves = 2, grouped = 0;
    uint32_t path_build = 0;      // 0 nested, 3 split (where the camera allows it)
    uint32_t split_shade = 4;     // the split build shades once this many eighths of a wave's working lanes wait
    bool split_ok = false;        // the last camera allows the split build
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double last_gpu_s = 0.0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    uint32_t gw = 0, gh = 0;
    const void* gsig = nullptr;
};

namespace {

void batch_release(phr_batch* b) {
    if (!b) return;
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    if (b->bvh) phr_bvh_free(b->bvh);
    for (phr_array* a : {b->in_faces, b->verts, b->mats, b->prims, b->faces, b->tris, b->lights, b->px, b->recs}) phr_array_release(a);
    if (b->table) (void)hipFree(b->table);
    if (b->cams) (void)hipFree(b->cams);
    if (b->gcams) (void)hipFree(b->gcams);
    if (b->rays) (void)hipFree(b->rays);
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    delete b;
}

// Boxes from the faces, the segmented build, the faces into slot order: all
// kernels, no host synchronisation (so a capture can hold it).
int batch_enqueue_build(phr_batch* b) {
    hipStream_t s = b->gpu->stream;
    const uint32_t n = b->n;
    hipLaunchKernelGGL(k_tri_boxes, dim3((n + 255) / 256), dim3(256), 0, s, (const uint4*)b->in_faces->ptr,
                       (const float4*)b->verts->ptr, n, (float4*)b->prims->ptr);
    int rc = phr_bvh_internal_rebuild_async(b->bvh);
    if (rc != PHR_OK) return rc;
    hipLaunchKernelGGL(k_gather, dim3((n + 255) / 256), dim3(256), 0, s, (const uint4*)b->in_faces->ptr, (const float4*)b->verts->ptr,
                       (const uint32_t*)phr_bvh_internal_order(b->bvh)->ptr, n, (uint4*)b->faces->ptr, (float4*)b->tris->ptr);
    PHR_HIP(hipGetLastError());
    return enqueue_walk_records(b->gpu, b->bvh, &b->recs, b->tris);
}

// Each scene's camera as a tracer of that scene alone stamps it.
int batch_cameras(phr_batch* b, const PhrCamera& cam, uint32_t w, uint32_t h) {
    b->split_ok = regen_allowed(cam);
    std::vector<PhrCamera> c(b->scenes, cam);
    for (uint32_t k = 0; k < b->scenes; ++k) {
        c[k].w = w;
        c[k].h = h;
        if (c[k].ao <= 0.f) c[k].ao = b->extent[k] * 0.08f;
        c[k].eps = b->extent[k] * 1e-5f;
        c[k].glow = b->glow[k];
        c[k].frame = 0;
    }
    if (b->cams_cap < b->scenes) {
        if (b->cams) (void)hipFree(b->cams);
        b->cams = nullptr;
        b->cams_cap = 0;
        PHR_HIP(hipMalloc((void**)&b->cams, sizeof(PhrCamera) * b->scenes));
        b->cams_cap = b->scenes;
    }
    PHR_HIP(hipMemcpyAsync(b->cams, c.data(), sizeof(PhrCamera) * b->scenes, hipMemcpyHostToDevice, b->gpu->stream));
    PHR_HIP(hipStreamSynchronize(b->gpu->stream));
    return PHR_OK;
}

int batch_image(phr_batch* b, uint32_t w, uint32_t h) {
    if (b->px && b->w == w && b->h == h) return PHR_OK;
    if (b->px) phr_array_release(b->px);
    b->px = nullptr;
    const uint64_t words = (PX_LEN * (uint64_t)w * h + 1) * b->scenes;
    int rc = phr_array_alloc(b->gpu, 4, words, &b->px);
    if (rc != PHR_OK) return rc;
    b->w = w;
    b->h = h;
    return PHR_OK;
}

void batch_launch(phr_batch* b, const PhrCamera* cams, uint32_t spp, uint32_t waves, uint32_t grouped, uint32_t build,
                  uint32_t shade8) {
    const uint32_t tiles_x = (b->w + 15) / 16, tiles = tiles_x * ((b->h + 15) / 16);
    const dim3 grid(tiles * b->scenes);
    const uint64_t stride = PX_LEN * (uint64_t)b->w * b->h + 1;
    launch_path_batch(waves, build == 3u && b->split_ok, grid, b->gpu->stream, b->table, cams, tiles, tiles_x, spp, (uint32_t*)b->px->ptr,
                      stride, b->rays, grouped, shade8);
}

int batch_check_camera(const PhrCamera& cam, uint32_t spp) {
    PHR_CHECK(spp >= 1, PHR_ERR_ARG, "phr_batch_render: spp must be at least 1");
    PHR_CHECK(!(cam.depth[0] == 0u || cam.depth[3] > 1u), PHR_ERR_UNSUPPORTED,
              "phr_batch_render: the glossy lobe is not restated; set Max Bounces (depth) with glossy 0");
    PHR_CHECK(cam.integrator[0] < 0.5f, PHR_ERR_UNSUPPORTED, "phr_batch_render: only phiray's own pixel filter (0)");
    PHR_CHECK(cam.adapt[3] == 0.0f, PHR_ERR_UNSUPPORTED, "phr_batch_render: adaptive sampling is per tracer");
    return PHR_OK;
}

}  // namespace

extern "C" {

// Many scenes (no counterpart in phiray, which holds one Scene a Tracer):
// scene k is tris[seg_lo[k], seg_lo[k + 1]), every scene non-empty. The
// faces of all scenes share one vertex and one material table (vertices are
// not shared between faces: the integrator reads them by value, so this
// gives what Tracer::new's welded table gives); one segmented LBVH build
// makes every scene's tree; each scene keeps its own emitter table and world
// diagonal. A scene renders as a tracer of its triangles alone does.
int phr_batch_new(phr_gpu* g, const void* tris_v, const uint32_t* seg_lo, uint32_t scenes, uint32_t leaf, phr_batch** out) {
    PHR_CHECK(g && tris_v && seg_lo && out && scenes > 0, PHR_ERR_ARG, "phr_batch_new: bad argument");
    PHR_CHECK(seg_lo[0] == 0, PHR_ERR_ARG, "phr_batch_new: the first scene starts at 0");
    for (uint32_t k = 0; k < scenes; ++k)
        PHR_CHECK(seg_lo[k + 1] > seg_lo[k], PHR_ERR_ARG, "phr_batch_new: every scene needs a triangle");
    const uint32_t n = seg_lo[scenes];
    const PhrTri* tris = (const PhrTri*)tris_v;
    for (uint32_t f = 0; f < n; ++f)
        PHR_CHECK(tris[f].trans == 0.f, PHR_ERR_UNSUPPORTED, "phr_batch_new: transmission (glass) is not supported");
    phr_batch* b = new phr_batch();
    b->gpu = g;
    b->scenes = scenes;
    b->n = n;
    b->leaf = leaf;
    b->seg_lo.assign(seg_lo, seg_lo + scenes + 1);
    auto fail = [&](int rc) { batch_release(b); return rc; };
    std::unordered_map<std::array<uint32_t, MAT_WORDS>, uint32_t, WordsHash<MAT_WORDS>> mmap;
    std::vector<uint32_t> in_faces(4 * (size_t)n);
    std::vector<float> verts(24 * (size_t)n), mats;
    b->extent.resize(scenes);
    std::array<uint32_t, MAT_WORDS> last_key{};
    uint32_t last_mid = PHR_NONE;
    for (uint32_t k = 0; k < scenes; ++k) {
        float lo[3] = {3.40282347e38f, 3.40282347e38f, 3.40282347e38f}, hi[3] = {-3.40282347e38f, -3.40282347e38f, -3.40282347e38f};
        for (uint32_t f = seg_lo[k]; f < seg_lo[k + 1]; ++f) {
            const PhrTri& tr = tris[f];
            float m[MAT_WORDS] = {tr.color[0], tr.color[1], tr.color[2], tr.color[3], tr.rough, tr.metal, tr.trans, tr.ior,
                                  tr.emit[0], tr.emit[1], tr.emit[2], tr.bump, 0, 0, 0, 0};
            std::array<uint32_t, MAT_WORDS> mk;
            for (uint32_t j = 0; j < MAT_WORDS; ++j) mk[j] = fbits(m[j]);
            uint32_t mid;
            if (last_mid != PHR_NONE && mk == last_key) {
                mid = last_mid;
            } else {
                auto mi = mmap.find(mk);
                if (mi == mmap.end()) {
                    mid = (uint32_t)mmap.size();
                    mmap.emplace(mk, mid);
                    mats.insert(mats.end(), m, m + MAT_WORDS);
                } else {
                    mid = mi->second;
                }
                last_key = mk;
                last_mid = mid;
            }
            const float* p[3] = {tr.a, tr.b, tr.c};
            const float* nn[3] = {tr.na, tr.nb, tr.nc};
            for (int c = 0; c < 3; ++c) {
                float o[2];
                oct_pack(nn[c], o);
                float* v = &verts[8 * (3 * (size_t)f + c)];
                v[0] = p[c][0], v[1] = p[c][1], v[2] = p[c][2], v[3] = 1.0f, v[4] = o[0], v[5] = o[1], v[6] = 0.f, v[7] = 0.f;
                in_faces[4 * (size_t)f + c] = 3 * f + c;
                for (int j = 0; j < 3; ++j) { lo[j] = std::min(lo[j], p[c][j]); hi[j] = std::max(hi[j], p[c][j]); }
            }
            in_faces[4 * (size_t)f + 3] = mid;
        }
        float sq = -0.0f;  // the scene's diagonal, as phr_tracer_new takes it
        for (int j = 0; j < 3; ++j) sq = sq + (hi[j] - lo[j]) * (hi[j] - lo[j]);
        b->extent[k] = std::max(std::sqrt(sq), 1e-3f);
    }
    int rc;
    hipStream_t s = g->stream;
    const uint32_t nmats = (uint32_t)mmap.size();
    if ((rc = phr_array_alloc(g, 16, n, &b->in_faces)) != PHR_OK) return fail(rc);
    if ((rc = phr_array_alloc(g, 32, 3ull * n, &b->verts)) != PHR_OK) return fail(rc);
    if ((rc = phr_array_alloc(g, 4 * MAT_WORDS, nmats, &b->mats)) != PHR_OK) return fail(rc);
    if ((rc = phr_array_alloc(g, 32, n, &b->prims)) != PHR_OK) return fail(rc);
    if ((rc = phr_array_alloc(g, 16, n, &b->faces)) != PHR_OK) return fail(rc);
    if ((rc = phr_array_alloc(g, 48, n, &b->tris)) != PHR_OK) return fail(rc);
    if (hipMalloc((void**)&b->rays, 32) != hipSuccess || hipMemsetAsync(b->rays, 0, 32, s) != hipSuccess ||
        hipEventCreate(&b->ev0) != hipSuccess || hipEventCreate(&b->ev1) != hipSuccess) {
        phr_set_error_msg("phr_batch_new: allocation failed");
        return fail(PHR_ERR_HIP);
    }
    if (hipMemcpyAsync(b->in_faces->ptr, in_faces.data(), 16ull * n, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(b->verts->ptr, verts.data(), 96ull * n, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(b->mats->ptr, mats.data(), 64ull * nmats, hipMemcpyHostToDevice, s) != hipSuccess) {
        phr_set_error_msg("phr_batch_new: upload failed");
        return fail(PHR_ERR_HIP);
    }
    hipLaunchKernelGGL(k_tri_boxes, dim3((n + 255) / 256), dim3(256), 0, s, (const uint4*)b->in_faces->ptr,
                       (const float4*)b->verts->ptr, n, (float4*)b->prims->ptr);
    // Trees packed one after another: scene k's at the nodes of the scenes before it.
    b->node_base.resize(scenes);
    uint64_t total = 0;
    for (uint32_t k = 0; k < scenes; ++k) {
        b->node_base[k] = (uint32_t)total;
        total += 2ull * (seg_lo[k + 1] - seg_lo[k]) - 1;
    }
    if (total >= PHR_LEAF) {
        phr_set_error_msg("phr_batch_new: too many nodes for 31-bit ids");
        return fail(PHR_ERR_SIZE);
    }
    if ((rc = phr_bvh_internal_build_into(g, b->prims, leaf, seg_lo, scenes, nullptr, b->node_base.data(), PHR_LEAF, &b->bvh)) != PHR_OK)
        return fail(rc);
    hipLaunchKernelGGL(k_gather, dim3((n + 255) / 256), dim3(256), 0, s, (const uint4*)b->in_faces->ptr, (const float4*)b->verts->ptr,
                       (const uint32_t*)phr_bvh_internal_order(b->bvh)->ptr, n, (uint4*)b->faces->ptr, (float4*)b->tris->ptr);
    if ((rc = enqueue_walk_records(g, b->bvh, &b->recs, b->tris)) != PHR_OK) return fail(rc);
    std::vector<uint32_t> order(n);
    if (hipMemcpyAsync(order.data(), phr_bvh_internal_order(b->bvh)->ptr, 4ull * n, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
        phr_set_error_msg("phr_batch_new: order readback failed");
        return fail(PHR_ERR_HIP);
    }
    std::vector<uint32_t> words;
    b->light_at.resize(scenes);
    b->glow.resize(scenes);
    for (uint32_t k = 0; k < scenes; ++k) {
        std::vector<uint32_t> wk = light_words(order.data(), seg_lo[k], seg_lo[k + 1], in_faces, verts, mats);
        b->light_at[k] = (uint32_t)words.size();
        b->glow[k] = wk[0];
        words.insert(words.end(), wk.begin(), wk.end());
    }
    if ((rc = phr_array_alloc(g, 4, words.size(), &b->lights)) != PHR_OK) return fail(rc);
    std::vector<SceneDev> table(scenes);
    for (uint32_t k = 0; k < scenes; ++k) {
        SceneDev& d = table[k];
        std::memset(&d, 0, sizeof d);
        d.nodes = (const float4*)phr_bvh_internal_nodes(b->bvh)->ptr;
        d.recs = (const float4*)b->recs->ptr;
        d.kids = d.recs + 4 * b->recs->len;
        d.tris = (const float4*)b->tris->ptr;
        d.faces = (const uint4*)b->faces->ptr;
        d.verts = (const float4*)b->verts->ptr;
        d.mats = (const float*)b->mats->ptr;
        d.lights = (const uint32_t*)b->lights->ptr + b->light_at[k];
        d.insts = nullptr;
        d.ninsts = 0;
        d.root = b->node_base[k];
    }
    if (hipMalloc((void**)&b->table, sizeof(SceneDev) * scenes) != hipSuccess ||
        hipMemcpyAsync(b->lights->ptr, words.data(), 4 * words.size(), hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(b->table, table.data(), sizeof(SceneDev) * scenes, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
        phr_set_error_msg("phr_batch_new: table upload failed");
        return fail(PHR_ERR_HIP);
    }
    b->grouped = scenes % 8 == 0 ? 1u : 0u;
    *out = b;
    return PHR_OK;
}

void phr_batch_free(phr_batch* b) { batch_release(b); }
uint32_t phr_batch_scenes(const phr_batch* b) { return b ? b->scenes : 0; }
uint32_t phr_batch_len(const phr_batch* b) { return b ? b->n : 0; }
phr_bvh* phr_batch_bvh(phr_batch* b) { return b ? b->bvh : nullptr; }
float phr_batch_extent(const phr_batch* b, uint32_t k) { return b && k < b->scenes ? b->extent[k] : 0.f; }
uint32_t phr_batch_glow(const phr_batch* b, uint32_t k) { return b && k < b->scenes ? b->glow[k] : 0u; }

// The whole build again from the triangles (boxes, segmented LBVH, slot
// order); gpu_s gets its device time.
int phr_batch_rebuild(phr_batch* b, double* gpu_s) {
    PHR_CHECK(b, PHR_ERR_ARG, "phr_batch_rebuild: null");
    hipStream_t s = b->gpu->stream;
    PHR_HIP(hipEventRecord(b->ev0, s));
    int rc = batch_enqueue_build(b);
    if (rc != PHR_OK) return rc;
    PHR_HIP(hipEventRecord(b->ev1, s));
    PHR_HIP(hipEventSynchronize(b->ev1));
    float ms = 0.f;
    PHR_HIP(hipEventElapsedTime(&ms, b->ev0, b->ev1));
    if (gpu_s) *gpu_s = ms * 1e-3;
    return PHR_OK;
}

// Every scene under `cam` at w x h, spp samples from a fresh image, in one
// launch; gpu_s gets its device time.
int phr_batch_render(phr_batch* b, const void* cam_v, uint32_t w, uint32_t h, uint32_t spp, double* gpu_s) {
    PHR_CHECK(b && cam_v, PHR_ERR_ARG, "phr_batch_render: null argument");
    w = std::max<uint32_t>(w, 1);
    h = std::max<uint32_t>(h, 1);
    PhrCamera cam;
    std::memcpy(&cam, cam_v, sizeof cam);
    int rc;
    if ((rc = batch_check_camera(cam, spp)) != PHR_OK || (rc = batch_image(b, w, h)) != PHR_OK ||
        (rc = batch_cameras(b, cam, w, h)) != PHR_OK)
        return rc;
    hipStream_t s = b->gpu->stream;
    PHR_HIP(hipEventRecord(b->ev0, s));
    batch_launch(b, b->cams, spp, b->path_waves, b->grouped, b->path_build, b->split_shade);
    PHR_HIP(hipGetLastError());
    PHR_HIP(hipEventRecord(b->ev1, s));
    PHR_HIP(hipEventSynchronize(b->ev1));
    float ms = 0.f;
    PHR_HIP(hipEventElapsedTime(&ms, b->ev0, b->ev1));
    b->last_gpu_s = ms * 1e-3;
    if (gpu_s) *gpu_s = b->last_gpu_s;
    return PHR_OK;
}

// The render kernel's build (waves a SIMD: 2 to 5) and block order
// (grouped: each XCD's blocks cover whole scenes; needs scenes % 8 == 0).
int phr_batch_set_path(phr_batch* b, uint32_t waves, uint32_t grouped) {
    PHR_CHECK(b && waves >= PATH_WAVES_MIN && waves <= PATH_WAVES_MAX && grouped <= 1, PHR_ERR_ARG,
              "phr_batch_set_path: waves 2-5, grouped 0 or 1");
    PHR_CHECK(!grouped || b->scenes % 8 == 0, PHR_ERR_ARG, "phr_batch_set_path: grouped order needs a multiple of 8 scenes");
    b->path_waves = waves;
    b->grouped = grouped;
    return PHR_OK;
}
// The render kernel's samples: 0 the nested build, 3 the split build (taken
// only where the camera allows it: lamps or a lit world, no Adaptive
// Sampling; other renders take the nested build).
int phr_batch_set_path_build(phr_batch* b, uint32_t build) {
    PHR_CHECK(b && (build == 0u || build == 3u), PHR_ERR_ARG, "phr_batch_set_path_build: null, or not 0 or 3");
    b->path_build = build;
    return PHR_OK;
}
uint32_t phr_batch_path_build(const phr_batch* b) { return b ? b->path_build : 0u; }
// The split build's shading point: it shades once `eighths` (1 to 8) eighths
// of a wave's working lanes wait. The images do not depend on it.
int phr_batch_set_split_shade(phr_batch* b, uint32_t eighths) {
    PHR_CHECK(b && eighths >= 1u && eighths <= 8u, PHR_ERR_ARG, "phr_batch_set_split_shade: null, or not 1 to 8");
    b->split_shade = eighths;
    return PHR_OK;
}
uint32_t phr_batch_split_shade(const phr_batch* b) { return b ? b->split_shade : 0u; }
int phr_batch_path(const phr_batch* b, uint32_t* waves, uint32_t* grouped) {
    PHR_CHECK(b, PHR_ERR_ARG, "phr_batch_path: null");
    if (waves) *waves = b->path_waves;
    if (grouped) *grouped = b->grouped;
    return PHR_OK;
}

// Time every (waves, order) pair on the caller's render, `reps` launches
// each after one warm-up, and keep the fastest.
int phr_batch_autotune(phr_batch* b, const void* cam_v, uint32_t w, uint32_t h, uint32_t spp, uint32_t reps,
                       uint32_t* waves, uint32_t* grouped, double* best_s) {
    PHR_CHECK(b && cam_v, PHR_ERR_ARG, "phr_batch_autotune: null argument");
    reps = std::max<uint32_t>(reps, 1);
    w = std::max<uint32_t>(w, 1);
    h = std::max<uint32_t>(h, 1);
    PhrCamera cam;
    std::memcpy(&cam, cam_v, sizeof cam);
    int rc;
    if ((rc = batch_check_camera(cam, spp)) != PHR_OK || (rc = batch_image(b, w, h)) != PHR_OK ||
        (rc = batch_cameras(b, cam, w, h)) != PHR_OK)
        return rc;
    hipStream_t s = b->gpu->stream;
    double best = 1e30;
    uint32_t bw = b->path_waves, bg = b->grouped, bb = b->path_build, bs = b->split_shade;
    for (uint32_t build : {0u, 3u}) {
        if (build == 3u && !b->split_ok) continue;
        for (uint32_t shade : {4u, 6u}) {
            if (build == 0u && shade != 4u) continue;  // the nested build has no shading point
            for (uint32_t g = 0; g <= (b->scenes % 8 == 0 ? 1u : 0u); ++g)
                for (uint32_t wv = PATH_WAVES_MIN; wv <= PATH_WAVES_MAX; ++wv) {
                    batch_launch(b, b->cams, spp, wv, g, build, shade);
                    PHR_HIP(hipEventRecord(b->ev0, s));
                    for (uint32_t r = 0; r < reps; ++r) batch_launch(b, b->cams, spp, wv, g, build, shade);
                    PHR_HIP(hipGetLastError());
                    PHR_HIP(hipEventRecord(b->ev1, s));
                    PHR_HIP(hipEventSynchronize(b->ev1));
                    float ms = 0.f;
                    PHR_HIP(hipEventElapsedTime(&ms, b->ev0, b->ev1));
                    const double t = ms * 1e-3 / reps;
                    if (t < best) best = t, bw = wv, bg = g, bb = build, bs = build == 3u ? shade : b->split_shade;
                }
        }
    }
    b->path_waves = bw;
    b->grouped = bg;
    b->path_build = bb;
    b->split_shade = bs;
    if (waves) *waves = bw;
    if (grouped) *grouped = bg;
    if (best_s) *best_s = best;
    return PHR_OK;
}

// The whole pass as one hipGraph: with `rebuild`, boxes + segmented build +
// slot order, then every scene's render (fresh images under `cam`). The call
// renders once eagerly first, which stamps the cameras and sizes the images.
int phr_batch_record(phr_batch* b, const void* cam_v, uint32_t w, uint32_t h, uint32_t spp, int rebuild) {
    PHR_CHECK(b && cam_v, PHR_ERR_ARG, "phr_batch_record: null argument");
    int rc = phr_batch_render(b, cam_v, w, h, spp, nullptr);
    if (rc != PHR_OK) return rc;
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    hipStream_t s = b->gpu->stream;
    // The pass keeps its own copy of the stamped cameras: a later eager
    // render restamps b->cams.
    if (b->gcams) (void)hipFree(b->gcams);
    b->gcams = nullptr;
    PHR_HIP(hipMalloc((void**)&b->gcams, sizeof(PhrCamera) * b->scenes));
    PHR_HIP(hipMemcpyAsync(b->gcams, b->cams, sizeof(PhrCamera) * b->scenes, hipMemcpyDeviceToDevice, s));
    PHR_HIP(hipStreamSynchronize(s));
    PHR_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed));
    if (rebuild) rc = batch_enqueue_build(b);
    if (rc == PHR_OK) batch_launch(b, b->gcams, spp, b->path_waves, b->grouped, b->path_build, b->split_shade);
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamEndCapture(s, &graph);
    if (rc != PHR_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
    }
    PHR_CHECK(e == hipSuccess && graph, PHR_ERR_HIP, "phr_batch_record: capture failed");
    hipGraphExec_t exec = nullptr;
    if (hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess) {
        (void)hipGraphDestroy(graph);
        phr_set_error_msg("phr_batch_record: instantiation failed");
        return PHR_ERR_HIP;
    }
    b->graph = graph;
    b->exec = exec;
    b->gw = b->w, b->gh = b->h;
    b->gsig = b->px->ptr;
    return PHR_OK;
}

int phr_batch_replay(phr_batch* b, double* gpu_s) {
    PHR_CHECK(b && b->exec, PHR_ERR_ARG, "phr_batch_replay: nothing recorded");
    PHR_CHECK(b->px && b->px->ptr == b->gsig && b->w == b->gw && b->h == b->gh, PHR_ERR_ARG,
              "phr_batch_replay: the images moved; record again");
    hipStream_t s = b->gpu->stream;
    PHR_HIP(hipEventRecord(b->ev0, s));
    PHR_HIP(hipGraphLaunch(b->exec, s));
    PHR_HIP(hipEventRecord(b->ev1, s));
    PHR_HIP(hipEventSynchronize(b->ev1));
    float ms = 0.f;
    PHR_HIP(hipEventElapsedTime(&ms, b->ev0, b->ev1));
    b->last_gpu_s = ms * 1e-3;
    if (gpu_s) *gpu_s = b->last_gpu_s;
    return PHR_OK;
}

// Scene k's image words [from, from + count) (its own PX_LEN * w * h + 1).
int phr_batch_px(phr_batch* b, uint32_t k, uint64_t from, uint64_t count, uint32_t* out) {
    PHR_CHECK(b && out && k < b->scenes, PHR_ERR_ARG, "phr_batch_px: bad argument");
    PHR_CHECK(b->px, PHR_ERR_ARG, "phr_batch_px: nothing rendered yet");
    const uint64_t stride = PX_LEN * (uint64_t)b->w * b->h + 1;
    PHR_CHECK(from + count <= stride, PHR_ERR_SIZE, "phr_batch_px: range beyond the scene's image");
    PHR_HIP(hipMemcpyAsync(out, (const uint32_t*)b->px->ptr + stride * k + from, 4 * count, hipMemcpyDeviceToHost, b->gpu->stream));
    PHR_HIP(hipStreamSynchronize(b->gpu->stream));
    return PHR_OK;
}

// Rays traced since the batch was made: closest, shadow, inner steps, triangle tests.
int phr_batch_counters(phr_batch* b, uint64_t* out4) {
    PHR_CHECK(b && out4, PHR_ERR_ARG, "phr_batch_counters: null");
    unsigned long long v[4];
    PHR_HIP(hipMemcpyAsync(v, b->rays, 32, hipMemcpyDeviceToHost, b->gpu->stream));
    PHR_HIP(hipStreamSynchronize(b->gpu->stream));
    for (int k = 0; k < 4; ++k) out4[k] = v[k];
    return PHR_OK;
}

}  // extern "C"

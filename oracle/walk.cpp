// CPU oracle: the watertight triangle test and the BVH walk.
//
// TEST INFRASTRUCTURE ONLY (see oracle.h). Restates:
//   src/trace/tri.rs:51-114             ray_axes, ray_space, tri_t (phiray's own CPU twin)
//   src/trace/tri.rs:123-140            Möller–Trumbore (the twin's witness)
//   src/trace/wgsl/bvh.wgsl:14-21       box_t
//   src/trace/wgsl/bvh.wgsl:111-114     safe_inv
//   src/trace/wgsl/bvh.wgsl:120-190     trace_blas (walk order, tie rules, any-hit)
// plus the deterministic sin/cos/exp2 the integrator uses (see below).
#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

namespace {

inline float fmin_(float a, float b) { return b < a ? b : a; }
inline float fmax_(float a, float b) { return b > a ? b : a; }

// src/trace/tri.rs:51-61 — depth axis last, the other two swapped on flip.
inline void ray_axes(const float v[3], uint32_t kz, bool flip, float out[3]) {
    float p[3];
    if (kz == 0) { p[0] = v[1]; p[1] = v[2]; p[2] = v[0]; }
    else if (kz == 1) { p[0] = v[2]; p[1] = v[0]; p[2] = v[1]; }
    else { p[0] = v[0]; p[1] = v[1]; p[2] = v[2]; }
    if (flip) { out[0] = p[1]; out[1] = p[0]; out[2] = p[2]; }
    else { out[0] = p[0]; out[1] = p[1]; out[2] = p[2]; }
}

}  // namespace

extern "C" {

void orc_ray_space(const float d[3], orc_rayspace* r) {
    // src/trace/tri.rs:65-78
    float a[3] = {std::fabs(d[0]), std::fabs(d[1]), std::fabs(d[2])};
    uint32_t kz = (a[0] >= a[1] && a[0] >= a[2]) ? 0 : (a[1] >= a[2] ? 1 : 2);
    float p[3];
    ray_axes(d, kz, false, p);
    bool flip = p[2] < 0.f;
    float dr[3];
    ray_axes(d, kz, flip, dr);
    float sz = 1.0f / dr[2];
    r->kz = kz;
    r->flip = flip ? 1u : 0u;
    r->s[0] = dr[0] * sz;
    r->s[1] = dr[1] * sz;
    r->s[2] = sz;
}

int orc_tri_t(const float ta[3], const float tb[3], const float tc[3], const float o[3], const orc_rayspace* r, float out[3]) {
    // src/trace/tri.rs:83-114 (Woop, Benthin, Wald, JCGT 2(1) 2013, no double fallback)
    bool flip = r->flip != 0;
    float va[3] = {ta[0] - o[0], ta[1] - o[1], ta[2] - o[2]};
    float vb[3] = {tb[0] - o[0], tb[1] - o[1], tb[2] - o[2]};
    float vc[3] = {tc[0] - o[0], tc[1] - o[1], tc[2] - o[2]};
    float a[3], b[3], c[3];
    ray_axes(va, r->kz, flip, a);
    ray_axes(vb, r->kz, flip, b);
    ray_axes(vc, r->kz, flip, c);
    float ax = a[0] - r->s[0] * a[2], ay = a[1] - r->s[1] * a[2];
    float bx = b[0] - r->s[0] * b[2], by = b[1] - r->s[1] * b[2];
    float cx = c[0] - r->s[0] * c[2], cy = c[1] - r->s[1] * c[2];
    float ea = cx * by - cy * bx;
    float eb = ax * cy - ay * cx;
    float ec = bx * ay - by * ax;
    if ((ea < 0.f || eb < 0.f || ec < 0.f) && (ea > 0.f || eb > 0.f || ec > 0.f)) return 0;
    float det = ea + eb + ec;
    if (det == 0.f) return 0;
    float tt = ea * (r->s[2] * a[2]) + eb * (r->s[2] * b[2]) + ec * (r->s[2] * c[2]);
    float sg = det < 0.f ? -1.0f : 1.0f;
    if (tt * sg <= 0.f) return 0;
    float inv = 1.0f / det;
    out[0] = tt * inv;
    out[1] = eb * inv;
    out[2] = ec * inv;
    return 1;
}

int orc_moller_trumbore(const float ta[3], const float tb[3], const float tc[3], const float o[3], const float d[3], float out[3]) {
    // src/trace/tri.rs:123-140
    float e1[3] = {tb[0] - ta[0], tb[1] - ta[1], tb[2] - ta[2]};
    float e2[3] = {tc[0] - ta[0], tc[1] - ta[1], tc[2] - ta[2]};
    auto cross = [](const float p[3], const float q[3], float r[3]) {
        r[0] = p[1] * q[2] - p[2] * q[1];
        r[1] = p[2] * q[0] - p[0] * q[2];
        r[2] = p[0] * q[1] - p[1] * q[0];
    };
    auto dot = [](const float p[3], const float q[3]) { return p[0] * q[0] + p[1] * q[1] + p[2] * q[2]; };
    float pv[3];
    cross(d, e2, pv);
    float det = dot(e1, pv);
    float inv = 1.0f / det;
    float tv[3] = {o[0] - ta[0], o[1] - ta[1], o[2] - ta[2]};
    float u = dot(tv, pv) * inv;
    float q[3];
    cross(tv, e1, q);
    float v = dot(d, q) * inv;
    float t = dot(e2, q) * inv;
    if (std::fabs(det) < 1e-12f || u < 0.f || u > 1.f || v < 0.f || u + v > 1.f) return 0;
    if (!(t > 0.f)) return 0;
    out[0] = t; out[1] = u; out[2] = v;
    return 1;
}

void orc_safe_inv(const float d[3], float out[3]) {
    // src/trace/wgsl/bvh.wgsl:111-114
    for (int k = 0; k < 3; ++k) {
        float s = d[k] < 0.f ? -1.0f : 1.0f;
        float x = std::fabs(d[k]) < 1e-9f ? s * 1e-9f : d[k];
        out[k] = 1.0f / x;
    }
}

float orc_box_t(const orc_node* n, const float o[3], const float inv[3], float best) {
    // src/trace/wgsl/bvh.wgsl:14-21
    float t1[3], t2[3];
    for (int k = 0; k < 3; ++k) {
        t1[k] = (n->lo.p[k] - o[k]) * inv[k];
        t2[k] = (n->hi.p[k] - o[k]) * inv[k];
    }
    float tmin = fmax_(fmax_(fmin_(t1[0], t2[0]), fmin_(t1[1], t2[1])), fmin_(t1[2], t2[2]));
    float tmax = fmin_(fmin_(fmax_(t1[0], t2[0]), fmax_(t1[1], t2[1])), fmax_(t1[2], t2[2]));
    if (tmax >= fmax_(tmin, 0.f) && tmin < best) return fmax_(tmin, 0.f);
    return ORC_INF;
}

// The stack walk of one tree from `root` (bvh.wgsl:142-189) in the space the
// ray is given in, keeping the nearest hit in `best` (carried in, so a later
// tree only takes strictly nearer hits) and naming instance `inst` in it.
// An any-hit (shadow) ray passes glass and cut-outs: a triangle whose material
// has transmission > 0.5 or alpha < 0.5 (bvh.wgsl:158-160); with no material
// table every triangle is opaque. Returns true when an any-hit ray stopped.
static bool walk_tree(const orc_node* nodes, uint32_t root, const float* tris, const uint32_t* faces, const float* mats,
                      const float o[3], const float d[3], bool any, uint32_t inst, orc_hit* best, uint32_t* steps_io,
                      uint32_t* tests_io) {
    float inv[3];
    orc_safe_inv(d, inv);
    orc_node nd = nodes[root];
    if (orc_box_t(&nd, o, inv, best->t) >= ORC_INF) return false;
    orc_rayspace rs;
    orc_ray_space(d, &rs);
    uint32_t stack[ORC_STACK];
    uint32_t sp = 0;
    uint32_t steps = 0, tests = 0;
    bool stopped = false;
    for (;;) {
        if (nd.lo.i & ORC_LEAF) {
            for (uint32_t id = nd.lo.i & ~ORC_LEAF; id < nd.hi.i; ++id) {
                const float* t = tris + 9 * (size_t)id;
                float h[3];
                ++tests;
                if (!orc_tri_t(t, t + 3, t + 6, o, &rs, h)) continue;
                if (h[0] < best->t) {
                    if (any) {
                        if (faces && mats) {
                            const float* m = mats + ORC_MAT_WORDS * (size_t)faces[4 * (size_t)id + 3];
                            if (m[6] > 0.5f || m[3] < 0.5f) continue;
                        }
                        best->t = h[0]; best->tri = id; best->u = h[1]; best->v = h[2]; best->inst = inst;
                        stopped = true;
                        goto done;
                    }
                    best->t = h[0]; best->tri = id; best->u = h[1]; best->v = h[2]; best->inst = inst;
                }
            }
        } else {
            ++steps;
            orc_node nl = nodes[nd.lo.i], nr = nodes[nd.hi.i];
            float tl = orc_box_t(&nl, o, inv, best->t);
            float tr = orc_box_t(&nr, o, inv, best->t);
            if (tl < ORC_INF && tr < ORC_INF) {
                if (tl <= tr) { stack[sp] = nd.hi.i; nd = nl; }
                else { stack[sp] = nd.lo.i; nd = nr; }
                ++sp;
                continue;
            }
            if (tl < ORC_INF) { nd = nl; continue; }
            if (tr < ORC_INF) { nd = nr; continue; }
        }
        if (sp == 0) break;
        --sp;
        nd = nodes[stack[sp]];
    }
done:
    if (steps_io) *steps_io += steps;
    if (tests_io) *tests_io += tests;
    return stopped;
}

static void walk_one(const orc_node* nodes, const float* tris, const uint32_t* faces, const float* mats,
                     const orc_ray& ray, bool any, orc_hit* best, uint32_t* st) {
    // src/trace/wgsl/bvh.wgsl:120-190, one identity instance (bit 0 of word 27),
    // root at node 0.
    best->t = ray.tmax; best->tri = ORC_NONE; best->u = 0.f; best->v = 0.f; best->inst = 0;
    uint32_t steps = 0, tests = 0;
    walk_tree(nodes, 0, tris, faces, mats, ray.o, ray.d, any, 0, best, &steps, &tests);
    if (st) { st[0] = steps; st[1] = tests; }
}

// A vec4 dot as the kernel evaluates it: ((x + y) + z) + w.
static inline float dot4(const float* r, float x, float y, float z, float w) { return ((r[0] * x + r[1] * y) + r[2] * z) + r[3] * w; }

// trace_blas (bvh.wgsl:120-190): instance i's tree in its own space. The ray
// goes there by the inverse (direction left unnormalised, so t is the
// world's), unless the instance sits at the identity (word 27 bit 0); a
// shadow ray passes an instance that casts none (bit 1).
static bool trace_blas(const orc_scene_t* s, uint32_t i, const float ow[3], const float dw[3], bool any, orc_hit* best,
                       uint32_t* steps, uint32_t* tests) {
    const uint32_t* w = s->insts + ORC_INSTANCE_LEN * (size_t)i;
    const uint32_t root = w[24];
    if (root == ORC_NONE) return false;
    if (any && (w[27] & 2u)) return false;
    float o[3] = {ow[0], ow[1], ow[2]}, d[3] = {dw[0], dw[1], dw[2]};
    if ((w[27] & 1u) == 0u) {
        const float* r = (const float*)(w + 12);
        for (int k = 0; k < 3; ++k) {
            o[k] = dot4(r + 4 * k, ow[0], ow[1], ow[2], 1.0f);
            d[k] = (r[4 * k] * dw[0] + r[4 * k + 1] * dw[1]) + r[4 * k + 2] * dw[2];
        }
    }
    return walk_tree(s->nodes, root, s->tris, s->faces, s->mats, o, d, any, i, best, steps, tests);
}

// trace (bvh.wgsl:194-232): the top tree over the instances' world boxes; a
// leaf's slots name instances (inst_order) whose own trees the ray walks.
static void trace_forest(const orc_scene_t* s, const orc_ray& ray, bool any, orc_hit* best, uint32_t* st) {
    best->t = ray.tmax; best->tri = ORC_NONE; best->u = 0.f; best->v = 0.f; best->inst = 0;
    uint32_t steps = 0, tests = 0;
    const uint32_t* inst_order = s->insts + ORC_INSTANCE_LEN * (size_t)s->ninsts;
    float inv[3];
    orc_safe_inv(ray.d, inv);
    orc_node nd = s->nodes[0];
    if (orc_box_t(&nd, ray.o, inv, best->t) < ORC_INF) {
        uint32_t stack[ORC_STACK];
        uint32_t sp = 0;
        for (;;) {
            if (nd.lo.i & ORC_LEAF) {
                bool stop = false;
                for (uint32_t k = nd.lo.i & ~(ORC_LEAF | ORC_INST); k < nd.hi.i; ++k) {
                    trace_blas(s, inst_order[k], ray.o, ray.d, any, best, &steps, &tests);
                    if (any && best->tri != ORC_NONE) { stop = true; break; }
                }
                if (stop) break;
            } else {
                ++steps;
                orc_node nl = s->nodes[nd.lo.i], nr = s->nodes[nd.hi.i];
                float tl = orc_box_t(&nl, ray.o, inv, best->t);
                float tr = orc_box_t(&nr, ray.o, inv, best->t);
                if (tl < ORC_INF && tr < ORC_INF) {
                    if (tl <= tr) { stack[sp] = nd.hi.i; nd = nl; }
                    else { stack[sp] = nd.lo.i; nd = nr; }
                    ++sp;
                    continue;
                }
                if (tl < ORC_INF) { nd = nl; continue; }
                if (tr < ORC_INF) { nd = nr; continue; }
            }
            if (sp == 0) break;
            --sp;
            nd = s->nodes[stack[sp]];
        }
    }
    if (st) { st[0] = steps; st[1] = tests; }
}

void orc_trace_scene(const orc_scene_t* s, const orc_ray* rays, uint32_t nrays, int any_hit, orc_hit* hits,
                     uint32_t* stats, int threads) {
    auto run = [&](uint32_t a, uint32_t b) {
        for (uint32_t i = a; i < b; ++i) {
            uint32_t* st = stats ? stats + 2 * (size_t)i : nullptr;
            if (s->forest) trace_forest(s, rays[i], any_hit != 0, &hits[i], st);
            else walk_one(s->nodes, s->tris, s->faces, s->mats, rays[i], any_hit != 0, &hits[i], st);
        }
    };
    if (threads <= 1 || nrays < 1024) { run(0, nrays); return; }
    std::vector<std::thread> pool;
    uint32_t per = (nrays + threads - 1) / threads;
    for (int k = 0; k < threads; ++k) {
        uint32_t a = k * per, b = std::min<uint32_t>(nrays, a + per);
        if (a < b) pool.emplace_back(run, a, b);
    }
    for (auto& th : pool) th.join();
}

void orc_raycast(const orc_node* nodes, const float* tris, const orc_ray* rays, uint32_t nrays,
                 int any_hit, orc_hit* hits, uint32_t* stats, int threads) {
    orc_raycast_mat(nodes, tris, nullptr, nullptr, rays, nrays, any_hit, hits, stats, threads);
}

void orc_raycast_mat(const orc_node* nodes, const float* tris, const uint32_t* faces, const float* mats,
                     const orc_ray* rays, uint32_t nrays, int any_hit, orc_hit* hits, uint32_t* stats, int threads) {
    auto run = [&](uint32_t a, uint32_t b) {
        for (uint32_t i = a; i < b; ++i)
            walk_one(nodes, tris, faces, mats, rays[i], any_hit != 0, &hits[i], stats ? stats + 2 * (size_t)i : nullptr);
    };
    if (threads <= 1 || nrays < 1024) { run(0, nrays); return; }
    std::vector<std::thread> pool;
    uint32_t per = (nrays + threads - 1) / threads;
    for (int t = 0; t < threads; ++t) {
        uint32_t a = per * t, b = std::min<uint32_t>(nrays, a + per);
        if (a >= b) break;
        pool.emplace_back(run, a, b);
    }
    for (auto& th : pool) th.join();
}

// ---- deterministic sin / cos / exp2 ---------------------------------------
//
// The integrator's transcendentals must give the same bits on the host and
// on the GPU, so neither side may call its own libm. Both evaluate the same
// sequence of single-precision operations:
//  * sin/cos: Cody–Waite reduction by pi/2 (W. J. Cody and W. Waite,
//    "Software Manual for the Elementary Functions", 1980) with pi/2 split
//    in three parts, then the Taylor polynomials of sin and cos on
//    [-pi/4, pi/4] through x^11 and x^12 (coefficients 1/n!, rounded to f32),
//    evaluated by Horner's rule;
//  * exp2: x = n + f with n = floor(x + 0.5), f in [-0.5, 0.5]; 2^f by its
//    Taylor series in f*ln2 through degree 7, then scaled by 2^n via the
//    exponent field.
// Valid for |x| < 1e5 (sin/cos) and x in [-126, 127] (exp2), the ranges the
// path reaches. The tests pin both against libm to a few ulp.
namespace {
// pi/2 = kPio2Hi + kPio2Md + kPio2Lo; kPio2Hi has 8 significant bits, so
// k * kPio2Hi is exact for every k the valid range produces.
const float kPio2Hi = 1.5703125f;
const float kPio2Md = 4.838267923332751e-04f;
const float kPio2Lo = 2.5633440682570896e-12f;
const float kTwoOverPi = 0.63661977236758134f;

inline float sin_poly(float x) {
    float z = x * x;
    // x - x^3/3! + x^5/5! - x^7/7! + x^9/9! - x^11/11!
    float p = -2.5052108385441720e-08f;
    p = p * z + 2.7557319223985893e-06f;
    p = p * z + -1.9841269841269841e-04f;
    p = p * z + 8.3333333333333333e-03f;
    p = p * z + -1.6666666666666667e-01f;
    return x + x * (z * p);
}
inline float cos_poly(float x) {
    float z = x * x;
    // 1 - x^2/2! + x^4/4! - x^6/6! + x^8/8! - x^10/10! + x^12/12!
    float p = 2.0876756987868099e-09f;
    p = p * z + -2.7557319223985888e-07f;
    p = p * z + 2.4801587301587302e-05f;
    p = p * z + -1.3888888888888889e-03f;
    p = p * z + 4.1666666666666667e-02f;
    p = p * z + -0.5f;
    return 1.0f + z * p;
}
inline void reduce(float x, int* q, float* r) {
    float k = std::floor(x * kTwoOverPi + 0.5f);
    float y = x - k * kPio2Hi;
    y = y - k * kPio2Md;
    y = y - k * kPio2Lo;
    *q = (int)k;
    *r = y;
}
}  // namespace

float orc_sin(float x) {
    int q; float r;
    reduce(x, &q, &r);
    switch (q & 3) {
        case 0: return sin_poly(r);
        case 1: return cos_poly(r);
        case 2: return -sin_poly(r);
        default: return -cos_poly(r);
    }
}

float orc_cos(float x) {
    int q; float r;
    reduce(x, &q, &r);
    switch (q & 3) {
        case 0: return cos_poly(r);
        case 1: return -sin_poly(r);
        case 2: return -cos_poly(r);
        default: return sin_poly(r);
    }
}

// log2: x = m * 2^e with m in [sqrt(1/2), sqrt(2)); ln m = 2 atanh(t),
// t = (m - 1) / (m + 1), by its series through t^11; then e + ln m * log2(e).
// x <= 0 gives -inf (pow(0, y > 0) = exp2(-inf) = 0). Normal inputs only.
float orc_log2(float x) {
    if (!(x > 0.f)) return -INFINITY;
    uint32_t b;
    std::memcpy(&b, &x, 4);
    int e = (int)((b >> 23) & 0xFF) - 127;
    uint32_t mb = (b & 0x007FFFFFu) | 0x3F800000u;
    float m;
    std::memcpy(&m, &mb, 4);
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

float orc_exp2(float x) {
    if (x < -126.f) x = -126.f;
    if (x > 127.f) x = 127.f;
    float n = std::floor(x + 0.5f);
    float f = x - n;
    float y = f * 0.69314718055994531f;
    // e^y, Taylor through y^7
    float p = 1.9841269841269841e-04f;
    p = p * y + 1.3888888888888889e-03f;
    p = p * y + 8.3333333333333333e-03f;
    p = p * y + 4.1666666666666667e-02f;
    p = p * y + 1.6666666666666667e-01f;
    p = p * y + 0.5f;
    p = p * y + 1.0f;
    p = p * y + 1.0f;
    int e = (int)n + 127;
    uint32_t bits = (uint32_t)e << 23;
    float scale;
    std::memcpy(&scale, &bits, 4);
    return p * scale;
}

}  // extern "C"

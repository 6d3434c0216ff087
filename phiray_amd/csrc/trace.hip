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

// Camera subset (src/trace/camera.rs:12-228): the fields the Rendered path reads.
struct PhrCamera {
    float right[3]; uint32_t persp;
    float up[3]; float scale;
    float back[3]; float units;
    float pivot[3]; float aspect;
    float sun[3]; float dist;
    float pan[2]; float ao; float eps;
    float env[4];
    uint32_t frame, count, glow, bounces;
    uint32_t depth[4];
    float integrator[4];
    uint32_t w, h, rr_from, _pad;
};
static_assert(sizeof(PhrCamera) == 176, "camera layout");

// Input triangle: Tri's rows a..nc, color and emit (src/trace/kernel.rs:11-40).
struct PhrTri {
    float a[3]; float rough;
    float na[3]; float trans;
    float b[3]; float metal;
    float nb[3]; float ior;
    float c[3]; float _c;
    float nc[3]; float _nc;
    float color[4];
    float emit[3]; float bump;
};
static_assert(sizeof(PhrTri) == 128, "tri layout");

constexpr uint32_t MAT_WORDS = 16;
constexpr uint32_t PX_LEN = 34;  // src/trace/read.rs:56

struct SceneDev {
    const float4* nodes;
    const float4* tris;    // 3 float4 per slot: (a.xyz, b.x) (b.yz, c.xy) (c.z, -, -, -)
    const uint4* faces;    // slot order: 3 vertex ids + material
    const float4* verts;   // 2 float4 per vertex: (pos, 1) (oct.xy, u, v)
    const float* mats;     // MAT_WORDS per material
    const uint32_t* lights;
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

// trace_blas for one identity instance, root at node 0. `stk` is this
// thread's LDS stack column (stride = threads per block).
// Any-hit rays skip triangles of glass or cut-out materials (trans > 0.5 or
// alpha < 0.5, bvh.wgsl:158-160): `faces` and `mats` give the material.
// `steps` and `tests` count inner-node visits and triangle tests, the units
// the walk's memory traffic is priced in.
template <bool ANY>
__device__ __forceinline__ PhrHit walk(const float4* __restrict__ nodes, const float4* __restrict__ tris,
                                       const uint4* __restrict__ faces, const float* __restrict__ mats, V3 o, V3 d, float tmax,
                                       uint32_t* stk, uint32_t stride, uint32_t& steps, uint32_t& tests) {
    PhrHit best{tmax, PHR_NONE, 0.f, 0.f, 0u};
    V3 inv = safe_inv(d);
    float4 lo = nodes[0], hi = nodes[1];
    if (box_t(lo, hi, o, inv, best.t) >= PHR_INF) return best;
    RaySpace rs = ray_space(d);
    uint32_t sp = 0;
    for (;;) {
        uint32_t li = __float_as_uint(lo.w), ri = __float_as_uint(hi.w);
        if (li & PHR_LEAF) {
            for (uint32_t id = li & ~PHR_LEAF; id < ri; ++id) {
                ++tests;
                float4 t0 = tris[3 * id], t1 = tris[3 * id + 1], t2 = tris[3 * id + 2];
                float u, v;
                float t = tri_t(t0, t1, t2, o, rs, u, v);
                if (t < best.t) {
                    if (ANY) {
                        const float* m = mats + MAT_WORDS * (size_t)faces[id].w;
                        if (m[6] > 0.5f || m[3] < 0.5f) continue;
                        return PhrHit{t, id, u, v, 0u};
                    }
                    best = PhrHit{t, id, u, v, 0u};
                }
            }
        } else {
            ++steps;
            float4 llo = nodes[2 * li], lhi = nodes[2 * li + 1];
            float4 rlo = nodes[2 * ri], rhi = nodes[2 * ri + 1];
            float tl = box_t(llo, lhi, o, inv, best.t);
            float tr = box_t(rlo, rhi, o, inv, best.t);
            if (tl < PHR_INF && tr < PHR_INF) {
                if (tl <= tr) { stk[sp * stride] = ri; lo = llo; hi = lhi; }
                else { stk[sp * stride] = li; lo = rlo; hi = rhi; }
                ++sp;
                continue;
            }
            if (tl < PHR_INF) { lo = llo; hi = lhi; continue; }
            if (tr < PHR_INF) { lo = rlo; hi = rhi; continue; }
        }
        if (sp == 0) break;
        --sp;
        uint32_t x = stk[sp * stride];
        lo = nodes[2 * x];
        hi = nodes[2 * x + 1];
    }
    return best;
}

constexpr uint32_t TRACE_THREADS = 256;

__global__ __launch_bounds__(TRACE_THREADS) void k_raycast(const float4* __restrict__ nodes, const float4* __restrict__ tris,
                                                           const uint4* __restrict__ faces, const float* __restrict__ mats,
                                                           const PhrRay* __restrict__ rays, uint32_t nrays, uint32_t any,
                                                           PhrHit* __restrict__ hits, unsigned long long* __restrict__ counters) {
    __shared__ uint32_t s_stack[PHR_STACK * TRACE_THREADS];
    const uint32_t i = blockIdx.x * TRACE_THREADS + threadIdx.x;
    uint32_t steps = 0, tests = 0;
    if (i < nrays) {
        const float4* r = (const float4*)(rays + i);
        float4 a = r[0], b = r[1];
        V3 o = v3(a.x, a.y, a.z), d = v3(b.x, b.y, b.z);
        uint32_t* stk = s_stack + threadIdx.x;
        PhrHit h = any ? walk<true>(nodes, tris, faces, mats, o, d, a.w, stk, TRACE_THREADS, steps, tests)
                       : walk<false>(nodes, tris, faces, mats, o, d, a.w, stk, TRACE_THREADS, steps, tests);
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
    uint32_t closest, shadow, steps, tests;
};

__device__ __forceinline__ PhrHit trace(Ctx& c, V3 o, V3 d, float tmax, bool first) {
    if (first) {
        c.shadow++;
        return walk<true>(c.s.nodes, c.s.tris, c.s.faces, c.s.mats, o, d, tmax, c.stk, TRACE_THREADS, c.steps, c.tests);
    }
    c.closest++;
    return walk<false>(c.s.nodes, c.s.tris, c.s.faces, c.s.mats, o, d, tmax, c.stk, TRACE_THREADS, c.steps, c.tests);
}
__device__ __forceinline__ float visible(Ctx& c, V3 org, V3 l, float dist) {
    return trace(c, org, l, dist, true).tri != PHR_NONE ? 0.0f : 1.0f;
}

__device__ __forceinline__ V3 unoct(float ex, float ey) {  // layout.wgsl:79-86
    if (ex > 1.5f) return v3(0, 0, 0);
    V3 n = v3(ex, ey, 1.0f - fabsf(ex) - fabsf(ey));
    if (n.z < 0.f) n = v3((1.0f - fabsf(ey)) * wgsl_sign(ex), (1.0f - fabsf(ex)) * wgsl_sign(ey), n.z);
    return normalize(n);
}

__device__ __forceinline__ uint32_t mat_at(const SceneDev& s, uint32_t f) { return s.faces[f].w; }

__device__ Tri tri_of(const SceneDev& s, uint32_t f, uint32_t m) {  // layout.wgsl:114-158
    uint4 fc = s.faces[f];
    float4 xp = s.verts[2 * fc.x], xa = s.verts[2 * fc.x + 1];
    float4 yp = s.verts[2 * fc.y], ya = s.verts[2 * fc.y + 1];
    float4 zp = s.verts[2 * fc.z], za = s.verts[2 * fc.z + 1];
    const float4* mt = (const float4*)(s.mats + MAT_WORDS * (size_t)m);
    float4 m0 = mt[0], m1 = mt[1], m2 = mt[2];
    Tri t;
    t.a = v3(xp.x, xp.y, xp.z); t.b = v3(yp.x, yp.y, yp.z); t.c = v3(zp.x, zp.y, zp.z);
    t.na = unoct(xa.x, xa.y); t.nb = unoct(ya.x, ya.y); t.nc = unoct(za.x, za.y);
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
    Tri t = tri_of(sc, f, mat_at(sc, f));
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
            q = glow_query(c.s, s, seed);
        }
        if (maxc(q.c) > 0.f) sum = sum + q.c * visible(c, s.pos, q.dir, q.dist);
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

__device__ V3 indirect(Ctx& c, V3 pos, V3 n, uint32_t* seed, bool bounce) {  // integrate.wgsl:111-230
    const PhrCamera* p = c.p;
    bool lit = world_lit(p);
    if (p->glow == 0u && !lit) {
        V3 dir = cosine_dir(n, seed);
        return env_rgb(p) * mixf(0.25f, 1.0f, visible(c, pos, dir, p->ao));
    }
    uint32_t depth = bounce ? (p->bounces != 0u ? p->bounces : 1u) : 1u;
    V3 acc = v3(0, 0, 0), thr = v3(1, 1, 1), org = pos, nrm = n;
    float min_pdf = 3.0e38f;
    for (uint32_t b = 0; b < depth; ++b) {
        V3 dir = cosine_dir(nrm, seed);
        float pb = fmax_(dot(nrm, dir), 0.0f) / PHR_PI;
        min_pdf = fmin_(min_pdf, pb);
        PhrHit h = trace(c, org, dir, PHR_INF, false);
        if (h.tri == PHR_NONE) {
            acc = acc + thr * env_rgb(p) * 1.0f;
            break;
        }
        if (!bounce) break;
        Tri t = tri_of(c.s, h.tri, mat_at(c.s, h.tri));
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
        acc = acc + thr * direct(c, s, seed) * near;
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

// One camera sample of pixel i (shade.wgsl:41-373). Returns what `guide` and
// `emit` add to the pixel's sums.
struct Sample {
    V3 lin, albedo, normal;
    float depth, missed, cov;
    uint32_t outline;
};

__device__ Sample shade(Ctx& c, uint32_t i, uint32_t frame) {
    const PhrCamera* p = c.p;
    Sample out;
    out.missed = 0.f;
    out.cov = 0.f;
    out.outline = 0u;
    float fw = (float)p->w, fh = (float)p->h;
    float su = ((float)(i % p->w) + 0.5f) / fw - 0.5f;
    float sv = ((float)(i / p->w) + 0.5f) / fh - 0.5f;
    uint32_t seed = pcg_hash(i * 9781u + frame * 6271u + 1u);
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
        PhrHit h = trace(c, o, d, PHR_INF, false);
        if (h.tri == PHR_NONE) {
            if (bounce == 0u) {
                out.albedo = v3(0, 0, 0);
                out.normal = -d;
                out.depth = -1.0f;
                out.lin = v3(0, 0, 0);
                out.missed = 1.0f;
                out.cov = 1.0f;
                return out;
            }
            col = col + thr * env_rgb(p);
            break;
        }
        if (bounce == 0u) out.outline = h.inst + 1u;
        uint32_t m = mat_at(c.s, h.tri);
        Tri t = tri_of(c.s, h.tri, m);
        float w = 1.0f - h.u - h.v;
        V3 hit = o + d * h.t;
        V3 ng0 = normalize(cross(t.b - t.a, t.c - t.a));
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
            out.albedo = base;
            out.normal = sn;
            out.depth = h.t;
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
        V3 cc = direct(c, s, &seed) * 1.0f;
        V3 bounced = v3(0, 0, 0);
        float pd = dot(vmax(s.albedo, 0.0f), v3(0.2126f, 0.7152f, 0.0722f));
        float pg = 0.0f;
        float p_diffuse = (pd + pg > 0.f) ? pd / (pd + pg) : 1.0f;
        bool take_diffuse = rnd(&seed) < p_diffuse;
        if (take_diffuse && pd > 0.f) bounced = bounced + s.albedo * indirect(c, pos, sn, &seed, diffuse_on) / p_diffuse;
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

// Samples [frame0, frame0 + spp) of pixels [first, first + count), summed
// into px in sample order (util.wgsl:17-34, 101-145: sums restart at frame 0).
__global__ __launch_bounds__(TRACE_THREADS) void k_path(SceneDev sc, const PhrCamera* __restrict__ camp, uint32_t first,
                                                        uint32_t count, uint32_t frame0, uint32_t spp, uint32_t* __restrict__ px,
                                                        unsigned long long* __restrict__ rays) {
    __shared__ uint32_t s_stack[PHR_STACK * TRACE_THREADS];
    const uint32_t k = blockIdx.x * TRACE_THREADS + threadIdx.x;
    const PhrCamera& cam = *camp;
    Ctx c;
    c.s = sc;
    c.p = camp;
    c.stk = s_stack + threadIdx.x;
    c.closest = 0;
    c.shadow = 0;
    c.steps = 0;
    c.tests = 0;
    if (k < count) {
        const uint32_t i = first + k;
        const size_t n = (size_t)cam.w * cam.h;
        bool keep = frame0 > 0u;
        V3 l = keep ? v3(ldf(px, n + 3 * i), ldf(px, n + 3 * i + 1), ldf(px, n + 3 * i + 2)) : v3(0, 0, 0);
        V3 ga = keep ? v3(ldf(px, 7 * n + 3 * i), ldf(px, 7 * n + 3 * i + 1), ldf(px, 7 * n + 3 * i + 2)) : v3(0, 0, 0);
        V3 gn = keep ? v3(ldf(px, 10 * n + 3 * i), ldf(px, 10 * n + 3 * i + 1), ldf(px, 10 * n + 3 * i + 2)) : v3(0, 0, 0);
        float gd = keep ? ldf(px, 13 * n + i) : 0.f;
        V3 gs = keep ? v3(ldf(px, 20 * n + 3 * i), ldf(px, 20 * n + 3 * i + 1), ldf(px, 20 * n + 3 * i + 2)) : v3(0, 0, 0);
        float cov = keep ? ldf(px, 23 * n + i) : 0.f;
        float film = keep ? ldf(px, 24 * n + i) : 0.f;
        uint32_t taken = keep ? px[31 * n + i] : 0u;
        float gone = keep ? ldf(px, 33 * n + i) : 0.f;
        uint32_t outline = 0;
        for (uint32_t s = 0; s < spp; ++s) {
            Sample sm = shade(c, i, frame0 + s);
            ga = ga + sm.albedo;
            gn = gn + sm.normal;
            gd = gd + sm.depth;
            gone = gone + sm.missed;
            taken = taken + 1u;
            film = film + 1.0f;
            l = l + sm.lin;
            gs = gs + v3(0.f, 0.f, 0.f);
            cov = cov + sm.cov;
            outline = sm.outline;
        }
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


}  // namespace

// ---------------------------------------------------------------------------
// Tracer (src/trace/mod.rs:50-245, build.rs, sample.rs, read.rs)
// ---------------------------------------------------------------------------
struct phr_bvh;
int phr_bvh_internal_build(phr_gpu* g, phr_array* prims, uint32_t leaf, phr_bvh** out);
phr_array* phr_bvh_internal_nodes(phr_bvh* b);
phr_array* phr_bvh_internal_order(phr_bvh* b);
extern "C" void phr_bvh_free(phr_bvh* b);
extern "C" int phr_bvh_refit_async(phr_bvh* b);
extern "C" uint32_t phr_bvh_depth(const phr_bvh* b);
int phr_denoise_enqueue(phr_gpu* g, uint32_t* px, uint32_t w, uint32_t h, int keep);

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
    phr_array* lights = nullptr;    // [count, ids, insts, cdf]
    phr_array* px = nullptr;        // PX_LEN * w * h + 1 words
    unsigned long long* rays = nullptr;  // closest, shadow
    phr_bvh* bvh = nullptr;
    std::vector<uint32_t> corner_vertex;  // input corner (3f + k) -> vertex id
    std::vector<uint32_t> order;          // slot -> input face
    uint32_t w = 0, h = 0, frame = 0, samples = 32;
    bool have_last = false;
    PhrCamera last{};
    PhrCamera* cam_dev = nullptr;  // the stamped camera the path kernel reads
    double last_gpu_s = 0.0, last_submit_s = 0.0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
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

void tracer_release(phr_tracer* t) {
    if (!t) return;
    for (phr_array* a : {t->in_faces, t->verts, t->mats, t->prims, t->faces, t->tris, t->lights, t->px}) phr_array_release(a);
    if (t->bvh) phr_bvh_free(t->bvh);
    if (t->rays) (void)hipFree(t->rays);
    if (t->cam_dev) (void)hipFree(t->cam_dev);
    if (t->ev0) (void)hipEventDestroy(t->ev0);
    if (t->ev1) (void)hipEventDestroy(t->ev1);
    delete t;
}

// The emitter table (src/trace/lamp.rs:51-97) over one identity instance,
// from the slot order.
int build_lights(phr_tracer* t, const std::vector<uint32_t>& in_faces, const std::vector<float>& verts,
                 const std::vector<float>& mats) {
    std::vector<uint32_t> ids;
    std::vector<float> pw;
    for (uint32_t s = 0; s < t->n; ++s) {
        const uint32_t* fc = &in_faces[4 * (size_t)t->order[s]];
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
    t->glow = m;
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
    s.nodes = (const float4*)phr_bvh_internal_nodes(t->bvh)->ptr;
    s.tris = (const float4*)t->tris->ptr;
    s.faces = (const uint4*)t->faces->ptr;
    s.verts = (const float4*)t->verts->ptr;
    s.mats = (const float*)t->mats->ptr;
    s.lights = (const uint32_t*)t->lights->ptr;
    return s;
}

}  // namespace

extern "C" {

// Tracer::new (src/trace/build.rs:11-13): the scene from triangle records
// (vertices shared by exact bits, materials by value), the LBVH over the faces
// built on the device, the faces put in leaf order, and the emitter table.
// PHR_ERR_ARG for an empty scene (the reference's None); PHR_ERR_UNSUPPORTED
// for a triangle the restated integrator does not carry (transmission).
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
    *out = t;
    return PHR_OK;
}

void phr_tracer_free(phr_tracer* t) { tracer_release(t); }
uint32_t phr_tracer_len(const phr_tracer* t) { return t ? t->n : 0; }
uint32_t phr_tracer_depth(const phr_tracer* t) { return t ? phr_bvh_depth(t->bvh) : 0; }
float phr_tracer_extent(const phr_tracer* t) { return t ? t->extent : 0.f; }
uint32_t phr_tracer_glow(const phr_tracer* t) { return t ? t->glow : 0; }
phr_bvh* phr_tracer_bvh(phr_tracer* t) { return t ? t->bvh : nullptr; }

// The slot -> input face map and the scene arrays, for tests and tooling.
int phr_tracer_order(phr_tracer* t, uint32_t* out) {
    PHR_CHECK(t && out, PHR_ERR_ARG, "phr_tracer_order: null");
    std::memcpy(out, t->order.data(), 4ull * t->n);
    return PHR_OK;
}
int phr_tracer_lights(phr_tracer* t, uint32_t* out, uint32_t cap, uint32_t* len) {
    PHR_CHECK(t && len, PHR_ERR_ARG, "phr_tracer_lights: null");
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
        SceneDev sc = scene_dev(t);
        hipLaunchKernelGGL(k_raycast, dim3((nrays + TRACE_THREADS - 1) / TRACE_THREADS), dim3(TRACE_THREADS), 0, s, sc.nodes, sc.tris, sc.faces, sc.mats,
                           (const PhrRay*)dr, nrays, any_hit ? 1u : 0u, (PhrHit*)dh, t->rays + 4);
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
    SceneDev sc = scene_dev(t);
    hipLaunchKernelGGL(k_raycast, dim3((nrays + TRACE_THREADS - 1) / TRACE_THREADS), dim3(TRACE_THREADS), 0, t->gpu->stream,
                       sc.nodes, sc.tris, sc.faces, sc.mats, (const PhrRay*)rays->ptr, nrays, any_hit ? 1u : 0u, (PhrHit*)hits->ptr,
                       t->rays + 4);
    PHR_HIP(hipGetLastError());
    return PHR_OK;
}

// Tracer::set_samples (src/trace/mod.rs:242).
void phr_tracer_set_samples(phr_tracer* t, uint32_t n) { if (t) t->samples = n < 1 ? 1 : n; }

// Tracer::sample_in (src/trace/sample.rs:104-249) without slices: `spp`
// samples in one launch. The camera is stamped as the reference stamps it
// (ao from the extent when unset, eps, glow); a repeated camera at the same
// size adds samples, any other starts the image over. `frame` gets the index of
// the last sample taken, `done` whether the tracer's sample count is reached.
int phr_tracer_sample(phr_tracer* t, const void* cam_v, uint32_t w, uint32_t h, uint32_t spp, uint32_t* frame, int* done) {
    PHR_CHECK(t && cam_v, PHR_ERR_ARG, "phr_tracer_sample: null argument");
    PHR_CHECK(spp >= 1, PHR_ERR_ARG, "phr_tracer_sample: spp must be at least 1");
    w = std::max<uint32_t>(w, 1);
    h = std::max<uint32_t>(h, 1);
    PhrCamera cam;
    std::memcpy(&cam, cam_v, sizeof cam);
    PHR_CHECK(!(cam.depth[0] == 0u || cam.depth[3] > 1u), PHR_ERR_UNSUPPORTED,
              "phr_tracer_sample: the glossy lobe is not restated; set Max Bounces (depth) with glossy 0");
    PHR_CHECK(cam.integrator[0] < 0.5f, PHR_ERR_UNSUPPORTED, "phr_tracer_sample: only phiray's own pixel filter (0)");
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
    hipLaunchKernelGGL(k_path, dim3((uint32_t)((n + TRACE_THREADS - 1) / TRACE_THREADS)), dim3(TRACE_THREADS), 0, s, sc,
                       (const PhrCamera*)t->cam_dev, 0u, (uint32_t)n, frame0, spp, (uint32_t*)t->px->ptr, t->rays);
    PHR_HIP(hipGetLastError());
    PHR_HIP(hipEventRecord(t->ev1, s));
    t->last_submit_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    PHR_HIP(hipEventSynchronize(t->ev1));
    float ms = 0.f;
    PHR_HIP(hipEventElapsedTime(&ms, t->ev0, t->ev1));
    t->last_gpu_s = ms * 1e-3;
    t->frame = frame0 + spp - 1;
    if (frame) *frame = t->frame;
    if (done) *done = (t->frame + 1 >= t->samples) ? 1 : 0;
    return PHR_OK;
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

// Tracer::denoise / denoise_shown (src/trace/read.rs:73-119): four a-trous
// passes over what has been traced; `keep` leaves the linear sums as they are.
int phr_tracer_denoise(phr_tracer* t, int keep) {
    PHR_CHECK(t, PHR_ERR_ARG, "phr_tracer_denoise: null");
    if (!t->px) return PHR_OK;  // nothing traced: nothing to filter
    int rc = phr_denoise_enqueue(t->gpu, (uint32_t*)t->px->ptr, t->w, t->h, keep);
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

}  // extern "C"

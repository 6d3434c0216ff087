// CPU oracle: the Rendered-mode integrator and the scene it reads.
//
// TEST INFRASTRUCTURE ONLY (see oracle.h). phiray has no CPU twin of its
// path tracer; this restates the WGSL for the feature subset the benchmark
// scenes use, operation for operation, with the same random-number draws:
//   src/trace/wgsl/util.wgsl:1-10        hash, rnd (PCG-RXS-M-XS)
//   src/trace/wgsl/util.wgsl:17-34       guide
//   src/trace/wgsl/util.wgsl:101-145     emit (linear sums, counts, the display
//                                        plane and the packed pixel)
//   src/trace/wgsl/view.wgsl:5-107       view_transform_of (Standard, Filmic and its looks, AgX), srgb, pack
//   src/trace/wgsl/util.wgsl:162-179     pixel, frame_size, frame_point
//   src/trace/wgsl/shade.wgsl:41-373     main (camera ray, the camera path)
//   src/trace/wgsl/integrate.wgsl:6-23   cosine_dir, glow_mis
//   src/trace/wgsl/integrate.wgsl:66-85  direct
//   src/trace/wgsl/integrate.wgsl:111-230 indirect (+ Russian roulette start knob)
//   src/trace/wgsl/light.wgsl:1-5,65-67,87-106,118-167(lamp pick only),174-198,264-312
//   src/trace/wgsl/brdf.wgsl:9-12,61-63,73-82,109-155,196-214,239-269
//   src/trace/wgsl/procedural.wgsl:73-81 fresnel_dielectric_cos
//   src/trace/wgsl/layout.wgsl:62-158    inst_at, place_n, unoct, world_ng, tri_of
//   src/trace/lamp.rs:51-97              the emitter table
//   src/trace/scene.rs:22-38, scene/add.rs:9-69  vertex packing and sharing
// Restated subset: opaque surfaces (trans = 0), alpha cut-outs, no textures,
// procedurals, mix closures, subsurface, sheen or coat (all zero weight),
// instances placed by any matrix (a forest) or one identity instance, a flat
// world colour (no environment image), lamps only as
// unlit placeholders (`Camera::count` with zero colour), the glossy lobe off
// (Max Bounces glossy 0), Pixel Filter 0 (phiray's own one-pixel box), no
// adaptive sampling, no clamping, no grid overlay, a black sky.
// `pow(x, 5)` is x^5 by multiplication and 0 for x < 0 (WGSL's pow is
// undefined there); sin/cos/exp2 are the deterministic ones in walk.cpp.
#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <thread>
#include <vector>
#include <array>
#include <atomic>

namespace {

const float PI = 3.14159265f;   // layout.wgsl:28
const float INF = 3.0e38f;      // layout.wgsl:39

struct V3 { float x, y, z; };
inline V3 v3(float x, float y, float z) { return V3{x, y, z}; }
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator-(V3 a) { return {-a.x, -a.y, -a.z}; }
inline V3 operator*(V3 a, V3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
inline V3 operator*(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline V3 operator*(float s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline V3 operator/(V3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
inline float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline float length(V3 a) { return std::sqrt(dot(a, a)); }
inline V3 normalize(V3 a) { return a / length(a); }
inline float fmin_(float a, float b) { return b < a ? b : a; }
inline float fmax_(float a, float b) { return b > a ? b : a; }
inline float clampf(float x, float lo, float hi) { return fmin_(fmax_(x, lo), hi); }
inline V3 vmax(V3 a, float s) { return {fmax_(a.x, s), fmax_(a.y, s), fmax_(a.z, s)}; }
inline float maxc(V3 a) { return fmax_(a.x, fmax_(a.y, a.z)); }
inline float pw5(float x) { if (!(x > 0.f)) return 0.f; float x2 = x * x; return (x2 * x2) * x; }
inline float wgsl_sign(float x) { return x > 0.f ? 1.f : (x < 0.f ? -1.f : 0.f); }
inline float mixf(float a, float b, float t) { return a * (1.0f - t) + b * t; }
inline V3 mixv(V3 a, V3 b, float t) { return a * (1.0f - t) + b * t; }
inline V3 ld3(const float* p) { return {p[0], p[1], p[2]}; }
inline float bits_f(uint32_t u) { float f; std::memcpy(&f, &u, 4); return f; }
inline uint32_t f_bits(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }

struct Hit { float t; uint32_t tri; float u, v; uint32_t inst; };

struct Tri {  // layout.wgsl:2-22, the fields the restated path reads
    V3 a, b, c, na, nb, nc;
    float color[4];
    float rough, metal, trans, ior;
    V3 emit;
};

struct Surf {  // brdf.wgsl:46-58
    V3 pos, n, v, albedo, f0;
    float rough;
    float bleed;
    float eta, dielectric;
};

struct Ctx {
    const orc_scene_t* s;
    const orc_camera* cam;
    uint64_t closest = 0, shadow = 0;
};

// ---- walk (bvh.wgsl:120-232: the top tree and the instances' trees, or one tree) ----
Hit trace(Ctx& c, V3 o, V3 d, float tmax, bool first) {
    orc_ray r;
    r.o[0] = o.x; r.o[1] = o.y; r.o[2] = o.z; r.tmax = tmax;
    r.d[0] = d.x; r.d[1] = d.y; r.d[2] = d.z; r._pad = 0.f;
    orc_hit h;
    orc_trace_scene(c.s, &r, 1, first ? 1 : 0, &h, nullptr, 1);
    if (first) c.shadow++; else c.closest++;
    return Hit{h.t, h.tri, h.u, h.v, h.inst};
}

float visible(Ctx& c, V3 org, V3 l, float dist) {  // light.wgsl:65-67
    return trace(c, org, l, dist, true).tri != ORC_NONE ? 0.0f : 1.0f;
}

// ---- random numbers (util.wgsl:1-10) ----
inline uint32_t hash_u(uint32_t x) {
    uint32_t h = x * 747796405u + 2891336453u;
    h = ((h >> ((h >> 28u) + 4u)) ^ h) * 277803737u;
    return (h >> 22u) ^ h;
}
inline float rnd(uint32_t* seed) {
    *seed = hash_u(*seed);
    return (float)(*seed) / 4294967295.0f;
}

// ---- scene reads ----
V3 unoct(float ex, float ey) {  // layout.wgsl:79-86
    if (ex > 1.5f) return v3(0, 0, 0);
    V3 n = v3(ex, ey, 1.0f - std::fabs(ex) - std::fabs(ey));
    if (n.z < 0.f) n = v3((1.0f - std::fabs(ey)) * wgsl_sign(ex), (1.0f - std::fabs(ex)) * wgsl_sign(ey), n.z);
    return normalize(n);
}

// Instance i's record, or none for a single-tree scene (one identity instance).
inline const uint32_t* inst_at(const orc_scene_t* s, uint32_t i) {
    return s->forest ? s->insts + ORC_INSTANCE_LEN * (size_t)i : nullptr;
}

// A vec4 dot as the kernel evaluates it: ((x + y) + z) + w.
inline float dot4(const float* r, const float* p) { return ((r[0] * p[0] + r[1] * p[1]) + r[2] * p[2]) + r[3] * p[3]; }

V3 place_n(const float* i0, const float* i1, const float* i2, V3 n) {  // layout.wgsl:70-77
    V3 m = (ld3(i0) * n.x + ld3(i1) * n.y) + ld3(i2) * n.z;
    float l = length(m);
    if (l < 1e-12f) return v3(0, 0, 0);
    return m / l;
}

// Face f under instance i with material m (layout.wgsl:114-158): its corners
// and normals placed in the world unless the instance sits at the identity.
Tri tri_of(const orc_scene_t* s, uint32_t f, uint32_t i, uint32_t m) {
    const uint32_t* fc = s->faces + 4 * (size_t)f;
    const float* x = s->verts + 8 * (size_t)fc[0];
    const float* y = s->verts + 8 * (size_t)fc[1];
    const float* z = s->verts + 8 * (size_t)fc[2];
    const float* mt = s->mats + ORC_MAT_WORDS * (size_t)m;
    Tri t;
    t.a = ld3(x); t.b = ld3(y); t.c = ld3(z);
    t.na = unoct(x[4], x[5]); t.nb = unoct(y[4], y[5]); t.nc = unoct(z[4], z[5]);
    const uint32_t* w = inst_at(s, i);
    if (w && (w[27] & 1u) == 0u) {
        const float* r = (const float*)w;
        t.a = v3(dot4(r, x), dot4(r + 4, x), dot4(r + 8, x));
        t.b = v3(dot4(r, y), dot4(r + 4, y), dot4(r + 8, y));
        t.c = v3(dot4(r, z), dot4(r + 4, z), dot4(r + 8, z));
        const float* iv = (const float*)(w + 12);
        t.na = place_n(iv, iv + 4, iv + 8, t.na);
        t.nb = place_n(iv, iv + 4, iv + 8, t.nb);
        t.nc = place_n(iv, iv + 4, iv + 8, t.nc);
    }
    for (int k = 0; k < 4; ++k) t.color[k] = mt[k];
    t.rough = mt[4]; t.metal = mt[5]; t.trans = mt[6]; t.ior = mt[7];
    t.emit = ld3(mt + 8);
    return t;
}
inline uint32_t mat_at(const orc_scene_t* s, uint32_t f) { return s->faces[4 * (size_t)f + 3]; }

// ---- BRDF (brdf.wgsl) ----
float ggx_lambda(float alpha2, float cos_n) {  // brdf.wgsl:9-12
    float t = alpha2 * fmax_(1.0f / fmax_(cos_n * cos_n, 1e-8f) - 1.0f, 0.0f);
    return 0.5f * (std::sqrt(1.0f + t) - 1.0f);
}

float fresnel_dielectric_cos(float cosi, float eta) {  // procedural.wgsl:73-81
    float c = std::fabs(cosi);
    float g = eta * eta - 1.0f + c * c;
    if (g <= 0.f) return 1.0f;
    g = std::sqrt(g);
    float a = (g - c) / (g + c);
    float b = (c * (g + c) - 1.0f) / (c * (g - c) + 1.0f);
    return 0.5f * a * a * (1.0f + b * b);
}

V3 fresnel_of(const Surf& s, float cs) {  // brdf.wgsl:73-82
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

// brdf.wgsl:109-116 — returns the (x, y) = ab pair; env_brdf(f0) = f0 * ab.x + ab.y.
void env_ab(float rough, float nv, float* abx, float* aby) {
    const float c0[4] = {-1.0f, -0.0275f, -0.572f, 0.022f};
    const float c1[4] = {1.0f, 0.0425f, 1.04f, -0.04f};
    float r[4];
    for (int k = 0; k < 4; ++k) r[k] = rough * c0[k] + c1[k];
    float a004 = fmin_(r[0] * r[0], orc_exp2(-9.28f * nv)) * r[0] + r[1];
    *abx = -1.04f * a004 + r[2];
    *aby = 1.04f * a004 + r[3];
}
inline V3 env_brdf(V3 f0, float rough, float nv) {
    float ax, ay;
    env_ab(rough, nv, &ax, &ay);
    return f0 * ax + v3(ay, ay, ay);
}

inline V3 dielectric_f0(float ior) {  // brdf.wgsl:123-126
    float r = (ior - 1.0f) / (ior + 1.0f);
    return v3(r * r, r * r, r * r);
}

float under_layer(V3 f0, float rough, float nv) {  // brdf.wgsl:135-138
    V3 e = env_brdf(f0, rough, nv);
    return clampf(1.0f - fmax_(e.x, fmax_(e.y, e.z)), 0.0f, 1.0f);
}

float under_layer_real(float eta, float rough, float nv) {  // brdf.wgsl:146-155
    float r = (eta - 1.0f) / (eta + 1.0f);
    float face_on = r * r;
    float climb = env_brdf(v3(0, 0, 0), rough, nv).x;
    float flat = env_brdf(v3(1, 1, 1), rough, nv).x - climb;
    float schlick = (1.0f - face_on) * pw5(1.0f - nv);
    float real = fresnel_dielectric_cos(nv, eta) - face_on;
    float scale = clampf(real / fmax_(schlick, 1e-5f), 0.0f, 1.0f);
    return clampf(1.0f - (face_on * flat + climb * scale), 0.0f, 1.0f);
}

float glossy_floor(const orc_camera* p, float min_pdf) {  // brdf.wgsl:202-208
    if (p->integrator[0] < 0.5f) return 0.3f;
    if (p->integrator[2] <= 0.0f) return 0.0f;
    float blur_pdf = min_pdf / p->integrator[2];
    if (blur_pdf >= 1.0f) return 0.0f;
    return std::sqrt(std::sqrt(1.0f - blur_pdf) * 0.5f);
}
inline bool caustic_reflective(const orc_camera* p) {  // brdf.wgsl:213
    return p->integrator[0] < 0.5f || (((uint32_t)(p->integrator[3] + 0.5f)) & 1u) != 0u;
}

// brdf.wgsl:239-269 with bleed, sheen and coat at zero weight, where the
// terms they add are exact zeros and the factors exact ones.
V3 brdf_w(const Surf& s, V3 l, V3 e, float wd, float ws) {
    float raw = dot(s.n, l);
    float nl = fmax_(raw, 0.0f);
    if (nl <= 0.f) return v3(0, 0, 0);
    V3 hv = normalize(l + s.v);
    float nh = fmax_(dot(s.n, hv), 0.0f);
    float hvv = fmax_(dot(hv, s.v), 0.0f);
    float a2 = fmax_(s.rough * s.rough * s.rough * s.rough, 1e-4f);
    float dd = nh * nh * (a2 - 1.0f) + 1.0f;
    float dist = a2 / (PI * dd * dd);
    V3 fres = fresnel_of(s, hvv);
    float nv = fmax_(dot(s.n, s.v), 1e-4f);
    float g2 = 1.0f / (1.0f + ggx_lambda(a2, nv) + ggx_lambda(a2, fmax_(raw, 1e-4f)));
    V3 spec = fres * dist * g2 * 0.25f / nv;
    V3 diffuse = s.albedo / PI;
    return (diffuse * wd + spec * ws) * e * nl;
}
inline V3 brdf(const Surf& s, V3 l, V3 e) { return brdf_w(s, l, e, 1.0f, 1.0f); }

Surf plain_surf(V3 pos, V3 n, V3 v, V3 albedo, V3 f0, float rough) {  // brdf.wgsl:61-63
    return Surf{pos, n, v, albedo, f0, rough, 0.0f, 0.0f, 0.0f};
}

// ---- lights (light.wgsl) ----
V3 onb(V3 n, float a, float b) {  // light.wgsl:1-5
    V3 tx = normalize(std::fabs(n.z) > 0.9f ? cross(n, v3(1, 0, 0)) : cross(n, v3(0, 0, 1)));
    V3 ty = cross(n, tx);
    return tx * a + ty * b;
}
V3 cone_dir(V3 axis, float ang, uint32_t* seed) {  // light.wgsl:34-39
    if (ang <= 0.f) return axis;
    float r = ang * std::sqrt(rnd(seed));
    float ph = 2.0f * PI * rnd(seed);
    return normalize(axis + onb(axis, r * orc_cos(ph), r * orc_sin(ph)));
}
inline V3 env_at(const orc_camera* p, V3) { return v3(p->env[0], p->env[1], p->env[2]); }  // light.wgsl:48-51

struct Shadowed { V3 c, dir; float dist; };
inline Shadowed unshadowed() { return Shadowed{v3(0, 0, 0), v3(0, 0, 1), 0.0f}; }

inline bool world_lit(const orc_camera* p) {  // light.wgsl:87-90, no environment image
    V3 e = v3(p->env[0], p->env[1], p->env[2]);
    return dot(e, e) > 0.f;
}
inline bool uses_rig(const orc_camera* p) { return p->count == 0u && !world_lit(p); }

Shadowed rig_query(const orc_camera* p, uint32_t k, const Surf& s, uint32_t* seed) {  // light.wgsl:94-101
    if (k == 0u) {
        V3 key = cone_dir(ld3(p->sun), 0.06f, seed);
        return Shadowed{brdf(s, key, v3(2, 2, 2)), key, INF};
    }
    V3 rim = normalize(ld3(p->up) * 0.6f - ld3(p->back) * 0.8f + ld3(p->right) * 0.2f);
    return Shadowed{brdf(s, rim, v3(0.9f, 0.9f, 0.9f)), rim, INF};
}
V3 rig_fill(const orc_camera* p, const Surf& s) {  // light.wgsl:103-106
    V3 fill = normalize(-ld3(p->right) * 0.8f + ld3(p->up) * 0.2f + ld3(p->back) * 0.6f);
    return brdf(s, fill, v3(0.45f, 0.45f, 0.45f));
}

// light.wgsl:118-140 with every lamp unlit: the pick draws its one number
// and finds no light.
Shadowed lamp_query(uint32_t* seed) {
    (void)rnd(seed);
    return unshadowed();
}

float lamp_chance(const orc_scene_t* s, const Hit& h) {  // light.wgsl:264-278
    const uint32_t* d = s->lights;
    uint32_t count = d[0];
    if (count == 0u) return 0.0f;
    uint32_t lo = 0, hi = count - 1;
    while (lo < hi) {
        uint32_t mid = (lo + hi) / 2u;
        uint32_t f = d[1 + mid], i = d[1 + count + mid];
        if (f < h.tri || (f == h.tri && i < h.inst)) lo = mid + 1; else hi = mid;
    }
    if (d[1 + lo] != h.tri || d[1 + count + lo] != h.inst) return 0.0f;
    float prev = lo > 0 ? bits_f(d[2 * count + lo]) : 0.0f;
    return bits_f(d[1 + 2 * count + lo]) - prev;
}

Shadowed glow_query(const orc_scene_t* sc, const Surf& s, uint32_t* seed) {  // light.wgsl:281-312
    const uint32_t* d = sc->lights;
    uint32_t count = d[0];
    if (count == 0u) return unshadowed();
    float u = rnd(seed);
    uint32_t lo = 0, hi = count - 1;
    while (lo < hi) {
        uint32_t mid = (lo + hi) / 2u;
        if (bits_f(d[1 + 2 * count + mid]) < u) lo = mid + 1; else hi = mid;
    }
    float prev = lo > 0 ? bits_f(d[2 * count + lo]) : 0.0f;
    float chance = fmax_(bits_f(d[1 + 2 * count + lo]) - prev, 1e-8f);
    uint32_t f = d[1 + lo];
    Tri t = tri_of(sc, f, d[1 + count + lo], mat_at(sc, f));
    float a = rnd(seed);
    float b = rnd(seed);
    if (a + b > 1.0f) { a = 1.0f - a; b = 1.0f - b; }
    V3 q = t.a + (t.b - t.a) * a + (t.c - t.a) * b;
    V3 cr = cross(t.b - t.a, t.c - t.a);
    float area = 0.5f * length(cr);
    V3 to = q - s.pos;
    float d2 = fmax_(dot(to, to), 1e-6f);
    V3 l = to / std::sqrt(d2);
    float cos_l = std::fabs(dot(normalize(cr), l));
    V3 e = t.emit * cos_l * area / (chance * d2);
    float pl = chance * d2 / fmax_(area * cos_l, 1e-8f);
    float pb = fmax_(dot(s.n, l), 0.0f) / PI;
    float w = pl * pl / fmax_(pl * pl + pb * pb, 1e-12f);
    return Shadowed{brdf(s, l, e) * w, l, std::sqrt(d2) * 0.999f};
}

// integrate.wgsl:66-85 (no environment image: the world query is empty).
V3 direct(Ctx& c, const Surf& s, uint32_t* seed) {
    const orc_camera* p = c.cam;
    bool rig = uses_rig(p);
    V3 sum = rig ? rig_fill(p, s) : v3(0, 0, 0);
    for (uint32_t k = 0; k < 4u; ++k) {
        Shadowed q = unshadowed();
        if (k == 0u) q = rig ? rig_query(p, 0u, s, seed) : lamp_query(seed);
        else if (k == 1u) { if (rig) q = rig_query(p, 1u, s, seed); }
        else if (k == 2u) q = glow_query(c.s, s, seed);
        if (maxc(q.c) > 0.f) sum = sum + q.c * visible(c, s.pos, q.dir, q.dist);
    }
    return sum;
}

// ---- integrate.wgsl ----
V3 cosine_dir(V3 n, uint32_t* seed) {  // integrate.wgsl:6-11
    float r1 = rnd(seed);
    float rr = std::sqrt(r1);
    float ph = 2.0f * PI * rnd(seed);
    return normalize(onb(n, rr * orc_cos(ph), rr * orc_sin(ph)) + n * std::sqrt(fmax_(0.0f, 1.0f - r1)));
}

V3 glow_mis(const orc_scene_t* s, const Tri& t, const Hit& h, V3 from_n, V3 dir) {  // integrate.wgsl:15-23
    V3 e = t.emit;
    float chance = lamp_chance(s, h);
    if (chance <= 0.f) return e;
    V3 cr = cross(t.b - t.a, t.c - t.a);
    float pl = chance * h.t * h.t / fmax_(0.5f * length(cr) * std::fabs(dot(normalize(cr), dir)), 1e-8f);
    float pb = fmax_(dot(from_n, dir), 0.0f) / PI;
    return e * pb * pb / fmax_(pb * pb + pl * pl, 1e-12f);
}

V3 indirect(Ctx& c, V3 pos, V3 n, uint32_t* seed, bool bounce) {  // integrate.wgsl:111-230
    const orc_camera* p = c.cam;
    bool lit = world_lit(p);
    if (p->glow == 0u && !lit) {
        V3 dir = cosine_dir(n, seed);
        return env_at(p, dir) * mixf(0.25f, 1.0f, visible(c, pos, dir, p->ao));
    }
    uint32_t depth = bounce ? (p->bounces != 0u ? p->bounces : 1u) : 1u;
    V3 acc = v3(0, 0, 0), thr = v3(1, 1, 1), org = pos, nrm = n;
    float min_pdf = 3.0e38f;
    V3 dir = v3(0, 0, 1);
    float pb = 1.0f;
    for (uint32_t b = 0; b < depth; ++b) {
        dir = cosine_dir(nrm, seed);
        pb = fmax_(dot(nrm, dir), 0.0f) / PI;
        min_pdf = fmin_(min_pdf, pb);
        Hit h = trace(c, org, dir, INF, false);
        if (h.tri == ORC_NONE) {
            acc = acc + thr * env_at(p, dir) * 1.0f;
            break;
        }
        if (!bounce) break;
        Tri t = tri_of(c.s, h.tri, h.inst, mat_at(c.s, h.tri));
        V3 ng0 = normalize(cross(t.b - t.a, t.c - t.a));
        V3 ng = dot(ng0, dir) < 0.f ? ng0 : -ng0;
        V3 hit = org + dir * h.t + ng * (p->eps + h.t * 1e-5f);
        V3 base = vmax(v3(t.color[0], t.color[1], t.color[2]), 0.0f);
        float rough = fmax_(t.rough, glossy_floor(p, min_pdf));
        bool glossy = caustic_reflective(p);
        V3 fd = glossy ? dielectric_f0(t.ior) : v3(0, 0, 0);
        float metal = glossy ? t.metal : 0.0f;
        float tw = t.trans * (1.0f - t.metal);
        V3 albedo = base * (1.0f - t.metal) * (1.0f - tw);
        if (maxc(albedo) > 0.f) albedo = albedo * under_layer(fd, rough, fmax_(dot(ng, -dir), 1e-4f));
        acc = acc + thr * glow_mis(c.s, t, h, nrm, dir);
        float near = (!lit && h.t <= p->ao) ? 0.25f : 1.0f;
        Surf s = plain_surf(hit, ng, -dir, albedo, mixv(fd, base, metal), rough);
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

// ---- the image buffer (util.wgsl:17-34, 101-145) ----
inline float pxf(const uint32_t* px, size_t w) { return bits_f(px[w]); }
inline void pxs(uint32_t* px, size_t w, float v) { px[w] = f_bits(v); }

void guide(uint32_t* px, size_t n, size_t i, bool keep, V3 albedo, V3 nn, float dist) {
    size_t a = 7 * n + 3 * i, b = 10 * n + 3 * i, c = 13 * n + i;
    V3 old_a = keep ? v3(pxf(px, a), pxf(px, a + 1), pxf(px, a + 2)) : v3(0, 0, 0);
    V3 old_n = keep ? v3(pxf(px, b), pxf(px, b + 1), pxf(px, b + 2)) : v3(0, 0, 0);
    float old_d = keep ? pxf(px, c) : 0.0f;
    V3 sa = old_a + albedo, sn = old_n + nn;
    pxs(px, a, sa.x); pxs(px, a + 1, sa.y); pxs(px, a + 2, sa.z);
    pxs(px, b, sn.x); pxs(px, b + 1, sn.y); pxs(px, b + 2, sn.z);
    pxs(px, c, old_d + dist);
}

// util.wgsl:72-96: Adaptive Sampling's words and switch.
inline bool adapt_on(const orc_camera* p) { return p->adapt[3] != 0.0f; }

void emit(uint32_t* px, size_t n, size_t i, bool keep, V3 lin, const float over[4], float cover, float missed,
          const orc_camera* p) {
    size_t a = n + 3 * i, g = 20 * n + 3 * i, gc = 23 * n + i, fc = 24 * n + i, mc = 33 * n + i, cn = 31 * n + i;
    float gone = (keep ? pxf(px, mc) : 0.0f) + missed;
    pxs(px, mc, gone);
    uint32_t taken = (keep ? px[cn] : 0u) + 1u;
    px[cn] = taken;
    float film = (keep ? pxf(px, fc) : 0.0f) + cover;
    pxs(px, fc, film);
    V3 l = (keep ? v3(pxf(px, a), pxf(px, a + 1), pxf(px, a + 2)) : v3(0, 0, 0)) + lin;
    V3 gs = (keep ? v3(pxf(px, g), pxf(px, g + 1), pxf(px, g + 2)) : v3(0, 0, 0)) + v3(over[0], over[1], over[2]);
    float cov = (keep ? pxf(px, gc) : 0.0f) + over[3];
    pxs(px, a, l.x); pxs(px, a + 1, l.y); pxs(px, a + 2, l.z);
    pxs(px, g, gs.x); pxs(px, g + 1, gs.y); pxs(px, g + 2, gs.z);
    pxs(px, gc, cov);
    // util.wgsl:126-132: the half image takes the odd samples at double weight.
    if (adapt_on(p) && (p->frame & 1u) == 1u) {
        size_t x = 26 * n + 3 * i;
        V3 half = (keep ? v3(pxf(px, x), pxf(px, x + 1), pxf(px, x + 2)) : v3(0, 0, 0)) + lin * 2.0f;
        pxs(px, x, half.x); pxs(px, x + 1, half.y); pxs(px, x + 2, half.z);
    }
    // util.wgsl:139-144: the light's mean over the samples that met something,
    // transformed once, under the grid's coverage, the grid laid over.
    float k = (float)taken;
    V3 m = l / fmax_(k - gone, 1.0f);
    float mv[3] = {m.x, m.y, m.z}, d[3];
    orc_view_transform(mv, p->view, d);
    V3 shown = v3(d[0], d[1], d[2]) * (1.0f - cov / k) + gs / k;
    pxs(px, 4 * n + 3 * i, shown.x); pxs(px, 4 * n + 3 * i + 1, shown.y); pxs(px, 4 * n + 3 * i + 2, shown.z);
    float sv[3] = {shown.x, shown.y, shown.z};
    px[i] = orc_pack_rgba8(sv, film / k);
}

// shade.wgsl:41-373 for one pixel and the camera's frame.
void shade(Ctx& c, uint32_t i, uint32_t* px) {
    const orc_camera* p = c.cam;
    const orc_scene_t* sc = c.s;
    size_t n = (size_t)p->w * p->h;
    bool keep = p->frame > 0u;
    if (p->frame == 0u) { px[29 * n + i] = 0u; px[30 * n + i] = 0u; }
    if (p->frame == 0u && adapt_on(p)) {
        // A new image starts its half image at zero. The reference does not:
        // its sample 1 adds to whatever the plane held (util.wgsl:126-132 keep
        // it from frame 1 on, and frame 0 never writes it).
        px[26 * n + 3 * i] = 0u; px[26 * n + 3 * i + 1] = 0u; px[26 * n + 3 * i + 2] = 0u;
    }
    if (adapt_on(p) && px[30 * n + i] != 0u) return;  // settled (shade.wgsl:64)
    float fw = (float)p->w, fh = (float)p->h;
    float su = ((float)(i % p->w) + 0.5f) / fw - 0.5f;
    float sv = ((float)(i / p->w) + 0.5f) / fh - 0.5f;
    uint32_t seed = hash_u(i * 9781u + (p->frame + p->stream) * 6271u + 1u);
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
    px[25 * n + i] = 0u;  // outline_store(i, 0)
    V3 col = v3(0, 0, 0), thr = v3(1, 1, 1);
    bool depth_set = p->depth[0] != 0u;
    uint32_t total = depth_set ? p->depth[0] - 1u : 8u;
    uint32_t see_through = depth_set ? p->depth[1] - 1u : 8u;
    uint32_t through = depth_set ? std::min(p->depth[2] - 1u, total) : 8u;
    bool diffuse_on = !depth_set || (p->bounces > 0u && total > 0u);
    uint32_t n_see = 0u;
    uint32_t passes = see_through + through + 1u;
    static const float none[4] = {0, 0, 0, 0};
    for (uint32_t bounce = 0; bounce < passes; ++bounce) {
        Hit h = trace(c, o, d, INF, false);
        if (h.tri == ORC_NONE) {
            if (bounce == 0u) {
                // The theme's sky (black here) rides the overlay term, not the light.
                static const float sky_over[4] = {0, 0, 0, 1};
                guide(px, n, i, keep, v3(0, 0, 0), -d, -1.0f);
                emit(px, n, i, keep, v3(0, 0, 0), sky_over, 1.0f, 1.0f, p);
                return;
            }
            col = col + thr * env_at(p, d);
            break;
        }
        if (bounce == 0u) px[25 * n + i] = h.inst + 1u;
        uint32_t m = mat_at(sc, h.tri);
        Tri t = tri_of(sc, h.tri, h.inst, m);
        float w = 1.0f - h.u - h.v;
        V3 hit = o + d * h.t;
        // world_ng (layout.wgsl:96-100): turned back under a mirroring instance.
        V3 ng0 = normalize(cross(t.b - t.a, t.c - t.a));
        const uint32_t* wi = inst_at(sc, h.inst);
        if (wi && (wi[27] & 4u)) ng0 = -ng0;
        bool front = dot(ng0, d) < 0.f;
        V3 ng = front ? ng0 : -ng0;
        V3 sn = t.na * w + t.nb * h.u + t.nc * h.v;
        sn = dot(sn, sn) > 1e-12f ? normalize(sn) * (front ? 1.0f : -1.0f) : ng;
        if (dot(sn, -d) < 0.f) sn = ng;
        // hit_lift / through_lift (light.wgsl:174-198)
        float start = 1.0e4f / p->units;
        float lift = (p->persp == 0u && h.t > start * 0.5f)
                         ? p->eps + std::fabs(h.t - start) * 1e-5f + start * 2.4e-7f
                         : p->eps + h.t * 1e-5f;
        V3 off = ng * lift;
        V3 pos = hit + off;
        V3 base = vmax(v3(t.color[0], t.color[1], t.color[2]), 0.0f);
        float alpha = t.color[3];
        float rough = t.rough, metal = t.metal;
        V3 emit_c = t.emit;
        if (bounce == 0u) guide(px, n, i, keep, base, sn, h.t);
        if (rnd(&seed) >= alpha) {
            n_see++;
            if (n_see > see_through) break;
            float reach = fmax_(fmax_(std::fabs(hit.x), std::fabs(hit.y)), fmax_(std::fabs(hit.z), 0.5f));
            float tl = lift - p->eps + reach * (256.0f / 8388608.0f);
            o = hit - ng * tl;
            continue;
        }
        col = col + thr * emit_c;
        V3 fd = dielectric_f0(t.ior);
        V3 diffuse = base * (1.0f - metal);
        if (maxc(diffuse) > 0.f) diffuse = diffuse * under_layer_real(t.ior, fmax_(rough, 0.04f), fmax_(dot(sn, -d), 1e-4f));
        Surf s{pos, sn, -d, diffuse, mixv(fd, base, metal), fmax_(rough, 0.04f), 0.0f, t.ior, 1.0f - metal};
        V3 cc = direct(c, s, &seed) * 1.0f;
        V3 bounced = v3(0, 0, 0);
        const V3 lw = v3(0.2126f, 0.7152f, 0.0722f);
        float pd = dot(vmax(s.albedo, 0.0f), lw);
        float pg = 0.0f;  // the glossy lobe is off
        float p_diffuse = (pd + pg > 0.f) ? pd / (pd + pg) : 1.0f;
        bool take_diffuse = rnd(&seed) < p_diffuse;
        if (take_diffuse && pd > 0.f) bounced = bounced + s.albedo * indirect(c, pos, sn, &seed, diffuse_on) / p_diffuse;
        (void)rnd(&seed);  // ggx_sample_vndf's two numbers (shade.wgsl:317)
        (void)rnd(&seed);
        cc = cc + bounced;
        col = col + thr * cc;
        break;  // opaque: tw == 0 ends the camera path (shade.wgsl:345-346)
    }
    emit(px, n, i, keep, col, none, 1.0f, 0.0f, p);
}

}  // namespace

extern "C" {

uint32_t orc_hash(uint32_t x) { return hash_u(x); }
float orc_rnd(uint32_t* seed) { return rnd(seed); }

void orc_sample(const orc_scene_t* s, const orc_camera* cam, uint32_t first, uint32_t count,
                uint32_t* px, uint64_t* rays, int threads) {
    std::atomic<uint64_t> closest{0}, shadow{0};
    auto run = [&](uint32_t a, uint32_t b) {
        Ctx c;
        c.s = s;
        c.cam = cam;
        for (uint32_t i = a; i < b; ++i) shade(c, i, px);
        closest += c.closest;
        shadow += c.shadow;
    };
    if (threads <= 1 || count < 256) run(first, first + count);
    else {
        std::vector<std::thread> pool;
        // Interleaved rows of 64 pixels spread the work of a frame evenly.
        std::atomic<uint32_t> next{first};
        for (int t = 0; t < threads; ++t)
            pool.emplace_back([&] {
                Ctx c;
                c.s = s;
                c.cam = cam;
                for (;;) {
                    uint32_t a = next.fetch_add(64);
                    if (a >= first + count) break;
                    uint32_t b = std::min(first + count, a + 64);
                    for (uint32_t i = a; i < b; ++i) shade(c, i, px);
                }
                closest += c.closest;
                shadow += c.shadow;
            });
        for (auto& th : pool) th.join();
    }
    if (rays) { rays[0] += closest; rays[1] += shadow; }
}


// adapt.wgsl:47-79 (stage 0): Cycles' convergence check, at the last sample
// of each step past the minimum, for pixels not yet settled.
void orc_adapt_check(const orc_camera* p, uint32_t* px) {
    if (!adapt_on(p)) return;
    const uint32_t sample = p->frame, min_samples = (uint32_t)p->adapt[1];
    const uint32_t step = std::max<uint32_t>((uint32_t)p->adapt[2], 1u);
    if (sample <= min_samples || (sample & (step - 1u)) != step - 1u) return;
    const size_t n = (size_t)p->w * p->h;
    for (size_t i = 0; i < n; ++i) {
        if (px[29 * n + i] != 0u) continue;
        const size_t a = n + 3 * i, x = 26 * n + 3 * i;
        const uint32_t taken = std::max<uint32_t>(px[31 * n + i], 1u);
        const float scale = 1.0f / (float)taken;
        const float d0 = std::fabs(pxf(px, a) - pxf(px, x)), d1 = std::fabs(pxf(px, a + 1) - pxf(px, x + 1)),
                    d2 = std::fabs(pxf(px, a + 2) - pxf(px, x + 2));
        const float err_diff = ((d0 + d1) + d2) * scale;
        const float intensity = ((pxf(px, a) + pxf(px, a + 1)) + pxf(px, a + 2)) * scale;
        const float norm = intensity < 1.0f ? std::sqrt(intensity) : intensity;
        const float error = err_diff / (0.0001f + norm);
        if (error < p->adapt[0]) px[29 * n + i] = 1u;
    }
}

// adapt.wgsl:82-96 (stage 1): settled only where the pixel and its four
// neighbours are.
void orc_adapt_dilate(const orc_camera* p, uint32_t* px) {
    if (!adapt_on(p)) return;
    const uint32_t w = p->w, h = p->h;
    const size_t n = (size_t)w * h;
    for (size_t i = 0; i < n; ++i) {
        const uint32_t x = (uint32_t)(i % w), y = (uint32_t)(i / w);
        uint32_t all = px[29 * n + i];
        if (x > 0u) all = std::min(all, px[29 * n + i - 1]);
        if (x + 1u < w) all = std::min(all, px[29 * n + i + 1]);
        if (y > 0u) all = std::min(all, px[29 * n + i - w]);
        if (y + 1u < h) all = std::min(all, px[29 * n + i + w]);
        px[30 * n + i] = all;
    }
}

}  // extern "C"

// view.wgsl:5-99: exposure as a power of two; Standard (the sRGB curve on the
// clamped channel), Filmic (its curve, or a look's) or AgX (its curve between
// the inset and outset matrices, sRGB-decoded, mixed, encoded again); then
// the gamma power. pow(x, y) is exp2(y log2 x), 0 for x <= 0. The curves
// are phiray's samples over log2(x) in [-14.947861, 10.052138], 64 a curve.
static const float ORC_FILMIC[64] = {
    0.000015f, 0.000015f, 0.000015f, 0.000015f, 0.000015f, 0.000015f, 0.000015f, 0.001109f,
    0.002798f, 0.004849f, 0.007337f, 0.010324f, 0.013937f, 0.018276f, 0.023498f, 0.029783f,
    0.037296f, 0.046297f, 0.057015f, 0.069747f, 0.084803f, 0.102511f, 0.123240f, 0.147317f,
    0.175060f, 0.206725f, 0.242476f, 0.282333f, 0.326138f, 0.373576f, 0.424074f, 0.476835f,
    0.530496f, 0.584086f, 0.636811f, 0.687598f, 0.735542f, 0.779871f, 0.820052f, 0.855760f,
    0.886888f, 0.913539f, 0.935939f, 0.954360f, 0.969213f, 0.980878f, 0.989775f, 0.996255f,
    1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f,
    1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f,
};
static const float ORC_AGX[64] = {
    0.000000f, 0.000000f, 0.000000f, 0.000000f, 0.000000f, 0.000000f, 0.000000f, 0.000015f,
    0.000094f, 0.000307f, 0.000714f, 0.001401f, 0.002474f, 0.004091f, 0.006433f, 0.009787f,
    0.014516f, 0.021097f, 0.030173f, 0.042574f, 0.057306f, 0.073883f, 0.092617f, 0.114256f,
    0.138886f, 0.167071f, 0.199246f, 0.235968f, 0.277732f, 0.324850f, 0.377462f, 0.435243f,
    0.495499f, 0.553566f, 0.607904f, 0.657973f, 0.703691f, 0.745144f, 0.782605f, 0.816304f,
    0.846638f, 0.873680f, 0.897952f, 0.919880f, 0.939710f, 0.957668f, 0.973975f, 0.988673f,
    1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f,
    1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f,
};
static const float ORC_LOOKS[6][64] = {
    {0.000000f, 0.000000f, 0.000000f, 0.000000f, 0.000000f, 0.000000f, 0.000000f, 0.000070f,
    0.000193f, 0.000363f, 0.000601f, 0.000931f, 0.001389f, 0.002026f, 0.002910f, 0.004138f,
    0.005841f, 0.008200f, 0.011463f, 0.015966f, 0.022162f, 0.030652f, 0.042221f, 0.057863f,
    0.078793f, 0.106414f, 0.142202f, 0.187489f, 0.243112f, 0.308980f, 0.383689f, 0.464395f,
    0.547092f, 0.627392f, 0.701407f, 0.766409f, 0.821117f, 0.865535f, 0.900558f, 0.927541f,
    0.947961f, 0.963206f, 0.974473f, 0.982737f, 0.988766f, 0.993146f, 0.996319f, 0.998612f,
    1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f,
    1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f},
    {0.000000f, 0.000000f, 0.000000f, 0.000000f, 0.000000f, 0.000000f, 0.000000f, 0.000233f,
    0.000620f, 0.001128f, 0.001793f, 0.002666f, 0.003810f, 0.005309f, 0.007269f, 0.009833f,
    0.013179f, 0.017541f, 0.023213f, 0.030568f, 0.040067f, 0.052275f, 0.067866f, 0.087615f,
    0.112373f, 0.143010f, 0.180318f, 0.224866f, 0.276828f, 0.335802f, 0.400681f, 0.469646f,
    0.540236f, 0.609827f, 0.675954f, 0.736553f, 0.790239f, 0.836358f, 0.874903f, 0.906341f,
    0.931432f, 0.951064f, 0.966136f, 0.977482f, 0.985838f, 0.991826f, 0.995958f, 0.998647f,
    1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f,
    1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f},
    {0.000000f, 0.000000f, 0.000000f, 0.000000f, 0.000000f, 0.000000f, 0.000000f, 0.000510f,
    0.001330f, 0.002363f, 0.003665f, 0.005306f, 0.007372f, 0.009971f, 0.013237f, 0.017335f,
    0.022469f, 0.028884f, 0.036880f, 0.046810f, 0.059090f, 0.074194f, 0.092650f, 0.115017f,
    0.141861f, 0.173696f, 0.210924f, 0.253744f, 0.302067f, 0.355438f, 0.412989f, 0.473463f,
    0.535105f, 0.596234f, 0.655290f, 0.710781f, 0.761550f, 0.806854f, 0.846356f, 0.880079f,
    0.908313f, 0.931529f, 0.950296f, 0.965209f, 0.976851f, 0.985758f, 0.992407f, 0.997209f,
    1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f,
    1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f},
    {0.000000f, 0.000000f, 0.000000f, 0.000000f, 0.000000f, 0.000000f, 0.000000f, 0.001808f,
    0.004533f, 0.007732f, 0.011485f, 0.015885f, 0.021036f, 0.027059f, 0.034090f, 0.042283f,
    0.051810f, 0.062859f, 0.075637f, 0.090362f, 0.107263f, 0.126574f, 0.148519f, 0.173307f,
    0.201111f, 0.232056f, 0.266193f, 0.303481f, 0.343773f, 0.386797f, 0.432150f, 0.479308f,
    0.526904f, 0.574394f, 0.621585f, 0.667736f, 0.712148f, 0.754204f, 0.793394f, 0.829342f,
    0.861800f, 0.890655f, 0.915907f, 0.937653f, 0.956064f, 0.971365f, 0.983812f, 0.993680f,
    1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f,
    1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f},
    {0.000000f, 0.000000f, 0.000000f, 0.000000f, 0.000000f, 0.000000f, 0.000000f, 0.003014f,
    0.007419f, 0.012418f, 0.018087f, 0.024509f, 0.031775f, 0.039988f, 0.049257f, 0.059700f,
    0.071444f, 0.084623f, 0.099377f, 0.115848f, 0.134178f, 0.154507f, 0.176963f, 0.201658f,
    0.228681f, 0.258089f, 0.289897f, 0.324071f, 0.360515f, 0.399071f, 0.439508f, 0.481524f,
    0.523400f, 0.565001f, 0.606881f, 0.648556f, 0.689527f, 0.729308f, 0.767438f, 0.803505f,
    0.837159f, 0.868121f, 0.896191f, 0.921245f, 0.943236f, 0.962183f, 0.978162f, 0.991297f,
    1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f,
    1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f},
    {0.000000f, 0.000000f, 0.000000f, 0.000000f, 0.000000f, 0.000000f, 0.000000f, 0.004685f,
    0.011326f, 0.018609f, 0.026591f, 0.035334f, 0.044905f, 0.055372f, 0.066812f, 0.079302f,
    0.092923f, 0.107759f, 0.123898f, 0.141424f, 0.160426f, 0.180986f, 0.203184f, 0.227091f,
    0.252769f, 0.280265f, 0.309608f, 0.340804f, 0.373832f, 0.408639f, 0.445135f, 0.483188f,
    0.520119f, 0.556217f, 0.593106f, 0.630501f, 0.668084f, 0.705500f, 0.742373f, 0.778309f,
    0.812910f, 0.845788f, 0.876574f, 0.904934f, 0.930579f, 0.953271f, 0.972832f, 0.989149f,
    1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f,
    1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f, 1.000000f}
};
static const float ORC_AGX_IN[3][3] = {0.498585f, 0.125929f, 0.097476f, 0.346742f, 0.681639f, 0.191866f, 0.097839f, 0.123183f, 0.678532f};   /* [column][row] */
static const float ORC_AGX_OUT[3][3] = {2.055217f, -0.287805f, -0.221530f, -0.912124f, 1.396234f, -0.286374f, -0.141522f, -0.097287f, 1.507346f};

static float orc_pow(float x, float y) { return x > 0.0f ? orc_exp2(y * orc_log2(x)) : 0.0f; }
static float orc_clamp01(float x) { return x < 0.0f ? 0.0f : (x > 1.0f ? 1.0f : x); }

static float orc_srgb1(float c) {
    float x = orc_clamp01(c);
    if (x <= 0.0031308f) return x * 12.92f;
    return 1.055f * orc_pow(x, 1.0f / 2.4f) - 0.055f;
}
static float orc_srgb_inv(float c) { return c <= 0.04045f ? c / 12.92f : orc_pow((c + 0.055f) / 1.055f, 2.4f); }

// view.wgsl:37-54: the curve linearly interpolated at log2(max(x, 1e-10)).
static float orc_curve(const float* curve, float x) {
    float l = orc_log2(x > 1e-10f ? x : 1e-10f);
    float t = orc_clamp01((l - (-14.947861f)) / (10.052138f - (-14.947861f))) * 63.0f;
    float fl = std::floor(t);
    unsigned i = (unsigned)fl, j = i + 1u < 63u ? i + 1u : 63u;
    float f = t - fl;
    return curve[i] * (1.0f - f) + curve[j] * f;
}

void orc_view_transform(const float lin[3], const float view[4], float out[3]) {
    float s = orc_exp2(view[1]);
    float c[3], d[3];
    for (int j = 0; j < 3; ++j) c[j] = (lin[j] > 0.0f ? lin[j] : 0.0f) * s;
    unsigned kind = (unsigned)(view[0] + 0.5f);
    if (kind == 1u) {
        unsigned look = (unsigned)(view[3] + 0.5f);
        const float* curve = (look >= 1u && look <= 6u) ? ORC_LOOKS[look - 1u] : ORC_FILMIC;
        for (int j = 0; j < 3; ++j) d[j] = orc_curve(curve, c[j]);
    } else if (kind == 2u) {
        float in[3], y[3], o[3];
        for (int r = 0; r < 3; ++r) in[r] = ORC_AGX_IN[0][r] * c[0] + ORC_AGX_IN[1][r] * c[1] + ORC_AGX_IN[2][r] * c[2];
        for (int j = 0; j < 3; ++j) y[j] = orc_srgb_inv(orc_curve(ORC_AGX, in[j]));
        for (int r = 0; r < 3; ++r) o[r] = ORC_AGX_OUT[0][r] * y[0] + ORC_AGX_OUT[1][r] * y[1] + ORC_AGX_OUT[2][r] * y[2];
        for (int j = 0; j < 3; ++j) d[j] = orc_srgb1(o[j]);
    } else {
        for (int j = 0; j < 3; ++j) d[j] = orc_srgb1(c[j]);
    }
    float g = 1.0f / (view[2] <= 0.0f ? 1.0f : view[2]);
    for (int j = 0; j < 3; ++j) out[j] = orc_pow(orc_clamp01(d[j]), g);
}

// view.wgsl:101-107.
uint32_t orc_pack_rgba8(const float rgb[3], float a) {
    uint32_t q[4];
    for (int j = 0; j < 4; ++j) {
        float c = j < 3 ? rgb[j] : a;
        c = c < 0.0f ? 0.0f : (c > 1.0f ? 1.0f : c);
        q[j] = (uint32_t)(c * 255.0f + 0.5f);
    }
    return q[0] | (q[1] << 8) | (q[2] << 16) | (q[3] << 24);
}

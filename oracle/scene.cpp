// CPU oracle: a tracer's scene from triangle records, in leaf order.
//
// TEST INFRASTRUCTURE ONLY (see oracle.h). Restates:
//   src/trace/scene.rs:22-38        oct / pack (octahedral normal, (2,2) = flat)
//   src/trace/scene/add.rs:9-69     vertex sharing by exact bits, material sharing
//   src/trace/scene.rs:188-190      reorder (faces into leaf order)
//   src/trace/lamp.rs:51-97         table_of (emitters by power, cdf bits)
//   src/trace/mat.rs:66-68          Mat::power (luminance of the emission)
// with the GPU LBVH (Bvh::build, leaf size `leaf`) as the tree.
#include "oracle.h"

#include <cmath>
#include <cstring>
#include <map>
#include <array>
#include <vector>

namespace {

inline float signum(float x) {  // Rust f32::signum: sign bit decides, NaN stays NaN
    if (x != x) return x;
    return std::signbit(x) ? -1.0f : 1.0f;
}

void oct(const float n[3], float out[2]) {  // src/trace/scene.rs:22-33
    float l = std::fabs(n[0]) + std::fabs(n[1]) + std::fabs(n[2]);
    if (l < 1e-12f) { out[0] = 0.f; out[1] = 0.f; return; }
    float x = n[0] / l, y = n[1] / l, z = n[2] / l;
    if (z >= 0.f) { out[0] = x; out[1] = y; }
    else { out[0] = (1.0f - std::fabs(y)) * signum(x); out[1] = (1.0f - std::fabs(x)) * signum(y); }
}

void pack(const float n[3], float out[2]) {  // src/trace/scene.rs:36-38
    if (n[0] == 0.f && n[1] == 0.f && n[2] == 0.f) { out[0] = 2.f; out[1] = 2.f; return; }
    oct(n, out);
}

inline uint32_t bits(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }

}  // namespace

extern "C" uint32_t orc_scene_build(const orc_tri* tris, uint32_t n, uint32_t leaf,
                                    orc_node* nodes, uint32_t* order, float* tri_cache, uint32_t* faces,
                                    float* verts, uint32_t* nverts, float* mats, uint32_t* nmats,
                                    uint32_t* lights, uint32_t* depth, float* extent, int threads) {
    *nverts = 0; *nmats = 0; *depth = 0; lights[0] = 0; *extent = 1e-3f;
    if (n == 0) return 0;
    // Scene::add for each triangle, in input order.
    std::map<std::array<uint32_t, 7>, uint32_t> vmap;
    std::map<std::array<uint32_t, ORC_MAT_WORDS>, uint32_t> mmap;
    std::vector<std::array<uint32_t, 4>> in_faces(n);
    std::vector<float> mat_rec;
    for (uint32_t f = 0; f < n; ++f) {
        const orc_tri& t = tris[f];
        float m[ORC_MAT_WORDS] = {t.color[0], t.color[1], t.color[2], t.color[3], t.rough, t.metal, t.trans, t.ior,
                                  t.emit[0], t.emit[1], t.emit[2], t.bump, 0, 0, 0, 0};
        std::array<uint32_t, ORC_MAT_WORDS> mk;
        for (int k = 0; k < ORC_MAT_WORDS; ++k) mk[k] = bits(m[k]);
        auto mi = mmap.find(mk);
        uint32_t mid;
        if (mi == mmap.end()) {
            mid = *nmats;
            mmap.emplace(mk, mid);
            std::memcpy(mats + ORC_MAT_WORDS * (size_t)mid, m, sizeof m);
            (*nmats)++;
        } else {
            mid = mi->second;
        }
        const float* p[3] = {t.a, t.b, t.c};
        const float* nn[3] = {t.na, t.nb, t.nc};
        uint32_t v[3];
        for (int k = 0; k < 3; ++k) {
            float o[2];
            pack(nn[k], o);
            std::array<uint32_t, 7> key = {bits(p[k][0]), bits(p[k][1]), bits(p[k][2]), bits(o[0]), bits(o[1]), bits(0.f), bits(0.f)};
            auto it = vmap.find(key);
            if (it == vmap.end()) {
                uint32_t id = *nverts;
                vmap.emplace(key, id);
                float* w = verts + 8 * (size_t)id;
                w[0] = p[k][0]; w[1] = p[k][1]; w[2] = p[k][2]; w[3] = 1.0f;
                w[4] = o[0]; w[5] = o[1]; w[6] = 0.f; w[7] = 0.f;
                (*nverts)++;
                v[k] = id;
            } else {
                v[k] = it->second;
            }
        }
        in_faces[f] = {v[0], v[1], v[2], mid};
    }
    // Forest::extent (src/trace/forest.rs:191-199): the root box's diagonal,
    // at least 1e-3; the root box is the union of the faces' boxes.
    {
        orc_aabb all;
        std::vector<orc_aabb> boxes(n);
        for (uint32_t f = 0; f < n; ++f) orc_of_tri(tris[f].a, tris[f].b, tris[f].c, &boxes[f]);
        orc_scene(boxes.data(), n, &all);
        float sq = -0.0f;
        for (int k = 0; k < 3; ++k) sq = sq + (all.hi[k] - all.lo[k]) * (all.hi[k] - all.lo[k]);
        float e = std::sqrt(sq);
        *extent = e > 1e-3f ? e : 1e-3f;
    }
    // The tree over the faces' boxes.
    std::vector<orc_aabb> prims(n);
    for (uint32_t f = 0; f < n; ++f) orc_of_tri(tris[f].a, tris[f].b, tris[f].c, &prims[f]);
    *depth = orc_build(prims.data(), n, leaf, nodes, order, threads);
    // Scene::reorder: slot s holds input face order[s].
    for (uint32_t s = 0; s < n; ++s) {
        const auto& fc = in_faces[order[s]];
        for (int k = 0; k < 4; ++k) faces[4 * (size_t)s + k] = fc[k];
        for (int k = 0; k < 3; ++k)
            for (int j = 0; j < 3; ++j) tri_cache[9 * (size_t)s + 3 * k + j] = verts[8 * (size_t)fc[k] + j];
    }
    // lamp.rs:51-97 over one identity instance (id 0).
    std::vector<uint32_t> ids;
    std::vector<float> pw;
    for (uint32_t s = 0; s < n; ++s) {
        const uint32_t* fc = faces + 4 * (size_t)s;
        const float* e = mats + ORC_MAT_WORDS * (size_t)fc[3] + 8;
        float glow = 0.2126f * e[0] + 0.7152f * e[1] + 0.0722f * e[2];
        if (glow <= 0.f) continue;
        const float* a = verts + 8 * (size_t)fc[0];
        const float* b = verts + 8 * (size_t)fc[1];
        const float* c = verts + 8 * (size_t)fc[2];
        float e1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
        float e2[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
        float cr[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
        float sq = -0.0f;
        for (int k = 0; k < 3; ++k) sq = sq + cr[k] * cr[k];
        float area = 0.5f * std::sqrt(sq);
        float p = area * glow;
        if (p > 0.f) { ids.push_back(s); pw.push_back(p); }
    }
    float total = -0.0f;
    for (float p : pw) total = total + p;
    uint32_t m = (uint32_t)ids.size();
    lights[0] = m;
    float run = 0.0f;
    for (uint32_t k = 0; k < m; ++k) {
        lights[1 + k] = ids[k];
        lights[1 + m + k] = 0u;
        run = run + pw[k];
        lights[1 + 2 * m + k] = bits(run / total);
    }
    return 2 * n - 1;
}

extern "C" void orc_stamp_camera(orc_camera* cam, uint32_t w, uint32_t h, float extent, uint32_t glow, uint32_t frame) {
    // src/trace/sample.rs:104-118
    cam->w = w < 1 ? 1 : w;
    cam->h = h < 1 ? 1 : h;
    if (cam->ao <= 0.f) cam->ao = extent * 0.08f;
    cam->eps = extent * 1e-5f;
    cam->glow = glow;
    cam->frame = frame;
}

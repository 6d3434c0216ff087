// CPU oracle: instances, meshes and the forest of trees over them.
//
// TEST INFRASTRUCTURE ONLY (see oracle.h). Restates:
//   src/trace/instance.rs:60-160    to_world, inverse, det, mirrors, world_box, words
//   src/trace/mesh.rs:43-72         texspace, fit
//   src/trace/forest.rs:44-224      build, tlas_slots, roots, regroup, retlas, nodes,
//                                   inst_words, extent
//   src/trace/scene.rs:182-190      boxes, reorder
//   src/trace/lamp.rs:51-97         table_of over instances
// Every tree is Bvh::build (the LBVH; src/bvh/build.rs:94-100), mesh trees at
// leaf 4 and the top tree at leaf 1, where the reference's forest calls
// build_sah: the LBVH is the tree this path builds and walks.
#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace {

inline uint32_t bits(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }
inline float fmin_(float a, float b) { return b < a ? b : a; }
inline float fmax_(float a, float b) { return b > a ? b : a; }

// a(r, c) = m[c][r]: rows of the upper 3 x 3 (instance.rs:78).
inline float A(const float w[16], int r, int c) { return w[4 * c + r]; }

float det3(const float w[16]) {  // instance.rs:96-100
    return A(w, 0, 0) * (A(w, 1, 1) * A(w, 2, 2) - A(w, 1, 2) * A(w, 2, 1)) -
           A(w, 0, 1) * (A(w, 1, 0) * A(w, 2, 2) - A(w, 1, 2) * A(w, 2, 0)) +
           A(w, 0, 2) * (A(w, 1, 0) * A(w, 2, 1) - A(w, 1, 1) * A(w, 2, 0));
}

bool is_identity(const float w[16]) {  // `self.world == IDENTITY`, float comparison
    for (int c = 0; c < 4; ++c)
        for (int r = 0; r < 4; ++r)
            if (!(w[4 * c + r] == (r == c ? 1.0f : 0.0f))) return false;
    return true;
}

void node_bounds(const orc_node& n, orc_aabb* b) {
    for (int k = 0; k < 3; ++k) { b->lo[k] = n.lo.p[k]; b->hi[k] = n.hi.p[k]; }
    b->_a = 0.f; b->_b = 0.f;
}

float luminance_of(const float* mat) {  // Mat::power (src/trace/mat.rs:66-68)
    const float* e = mat + 8;
    return 0.2126f * e[0] + 0.7152f * e[1] + 0.0722f * e[2];
}

struct Trees {
    std::vector<std::vector<orc_node>> blas;  // per mesh, empty when it has no faces
    std::vector<uint32_t> blas_depth;
    std::vector<orc_node> tlas;
    std::vector<uint32_t> tlas_order;  // slot -> index into `placed`
    std::vector<uint32_t> placed;      // instance ids with a tree
    uint32_t tlas_depth = 0;
};

uint32_t tlas_slots(const orc_scene_in* in) {  // forest.rs:71-73
    uint32_t s = in->ninsts ? 2 * in->ninsts - 1 : 0;
    return s > 1 ? s : 1;
}

std::vector<uint32_t> roots(const orc_scene_in* in, const Trees& t) {  // forest.rs:77-91
    std::vector<uint32_t> out(in->nmeshes, ORC_NONE);
    uint32_t at = tlas_slots(in);
    for (uint32_t k = 0; k < in->nmeshes; ++k) {
        if (t.blas[k].empty()) continue;
        out[k] = at;
        at += (uint32_t)t.blas[k].size();
    }
    return out;
}

void retlas(const orc_scene_in* in, Trees& t) {  // forest.rs:111-127
    std::vector<orc_aabb> boxes;
    t.placed.clear();
    for (uint32_t i = 0; i < in->ninsts; ++i) {
        uint32_t m = in->insts[i].mesh;
        if (m >= in->nmeshes || t.blas[m].empty()) continue;
        orc_aabb b, wb;
        node_bounds(t.blas[m][0], &b);
        orc_inst_world_box(in->insts[i].world, &b, &wb);
        t.placed.push_back(i);
        boxes.push_back(wb);
    }
    const uint32_t p = (uint32_t)boxes.size();
    t.tlas.assign(p ? 2 * p - 1 : 0, orc_node{});
    t.tlas_order.assign(p, 0);
    t.tlas_depth = p ? orc_build(boxes.data(), p, 1, t.tlas.data(), t.tlas_order.data(), 1) : 0;
}

// forest.rs:132-165: the top tree in its slots (leaves flagged INST), then each
// mesh tree with children and face ranges shifted into place.
void write_nodes(const orc_scene_in* in, const Trees& t, orc_node* out, bool tlas_only) {
    const uint32_t slots = tlas_slots(in);
    for (uint32_t k = 0; k < slots; ++k) out[k] = orc_node{};
    for (size_t k = 0; k < t.tlas.size(); ++k) {
        orc_node n = t.tlas[k];
        if (n.lo.i & ORC_LEAF) n.lo.i |= ORC_INST;
        out[k] = n;
    }
    if (tlas_only) return;
    uint32_t base = slots;
    for (uint32_t m = 0; m < in->nmeshes; ++m) {
        const uint32_t lo = in->meshes[m].face_lo, at = base;
        for (const orc_node& src : t.blas[m]) {
            orc_node n = src;
            if (n.lo.i & ORC_LEAF) {
                n.lo.i = ((n.lo.i & ~ORC_LEAF) + lo) | ORC_LEAF;
                n.hi.i += lo;
            } else {
                n.lo.i += at;
                n.hi.i += at;
            }
            out[base++] = n;
        }
    }
}

// forest.rs:93-109 (regroup) + 168-181 (inst_words).
void write_insts(const orc_scene_in* in, const Trees& t, uint32_t* out) {
    const std::vector<uint32_t> r = roots(in, t);
    const uint32_t n = in->ninsts;
    orc_mesh unit{0, 0, 0, 0, {0.f, 0.f, 0.f}, {1.f, 1.f, 1.f}};
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t m = in->insts[i].mesh;
        const bool ok = m < in->nmeshes;
        orc_inst_words(&in->insts[i], ok ? r[m] : ORC_NONE, ok ? &in->meshes[m] : &unit, out + ORC_INSTANCE_LEN * (size_t)i);
    }
    uint32_t* order = out + ORC_INSTANCE_LEN * (size_t)n;
    for (uint32_t s = 0; s < n; ++s) order[s] = s < t.tlas_order.size() ? t.placed[t.tlas_order[s]] : 0u;
    uint32_t* draw = order + n;
    uint32_t at = 0;
    for (uint32_t m = 0; m < in->nmeshes; ++m)
        for (uint32_t i = 0; i < n; ++i)
            if (in->insts[i].mesh == m) draw[at++] = i;
    while (at < n) draw[at++] = 0u;
}

// lamp.rs:51-97: each emissive face once under each instance of its mesh, its
// power the placed area times the material's luminance; sorted by face, then
// instance. faces_slot: the faces in leaf order.
void write_lights(const orc_scene_in* in, const uint32_t* faces_slot, uint32_t* out) {
    std::vector<std::vector<uint32_t>> by_mesh(in->nmeshes);
    for (uint32_t i = 0; i < in->ninsts; ++i)
        if (in->insts[i].mesh < in->nmeshes) by_mesh[in->insts[i].mesh].push_back(i);
    std::vector<uint32_t> ids, owners;
    std::vector<float> pw;
    for (uint32_t m = 0; m < in->nmeshes; ++m) {
        for (uint32_t f = in->meshes[m].face_lo; f < in->meshes[m].face_hi; ++f) {
            const uint32_t* fc = faces_slot + 4 * (size_t)f;
            if (fc[3] >= in->nmats) continue;
            const float glow = luminance_of(in->mats + ORC_MAT_WORDS * (size_t)fc[3]);
            if (glow <= 0.f) continue;
            for (uint32_t i : by_mesh[m]) {
                float a[3], b[3], c[3];
                orc_inst_to_world(in->insts[i].world, in->verts + 8 * (size_t)fc[0], a);
                orc_inst_to_world(in->insts[i].world, in->verts + 8 * (size_t)fc[1], b);
                orc_inst_to_world(in->insts[i].world, in->verts + 8 * (size_t)fc[2], c);
                float e1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
                float e2[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
                float cr[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
                float sq = -0.0f;
                for (int k = 0; k < 3; ++k) sq = sq + cr[k] * cr[k];
                float p = (0.5f * std::sqrt(sq)) * glow;
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
    out[0] = cnt;
    float run = 0.0f;
    for (uint32_t k = 0; k < cnt; ++k) {
        out[1 + k] = ids[k];
        out[1 + cnt + k] = owners[k];
        run = run + pw[k];
        out[1 + 2 * cnt + k] = bits(run / total);
    }
}

float extent_of(const Trees& t) {  // forest.rs:189-199
    if (t.tlas.empty()) return 1e-3f;
    const orc_node& r = t.tlas[0];
    float sq = -0.0f;
    for (int k = 0; k < 3; ++k) sq = sq + (r.hi.p[k] - r.lo.p[k]) * (r.hi.p[k] - r.lo.p[k]);
    return std::max(std::sqrt(sq), 1e-3f);
}

}  // namespace

extern "C" {

void orc_inst_to_world(const float w[16], const float p[3], float out[3]) {  // instance.rs:68-71
    for (int r = 0; r < 3; ++r) out[r] = ((w[r] * p[0] + w[4 + r] * p[1]) + w[8 + r] * p[2]) + w[12 + r];
}

void orc_inst_inverse(const float w[16], float out[12]) {  // instance.rs:75-93
    const float det = det3(w);
    if (std::fabs(det) < 1e-20f) {
        for (int k = 0; k < 12; ++k) out[k] = 0.0f;
        return;
    }
    const float d = 1.0f / det;
    const float inv[3][3] = {
        {(A(w, 1, 1) * A(w, 2, 2) - A(w, 1, 2) * A(w, 2, 1)) * d, (A(w, 0, 2) * A(w, 2, 1) - A(w, 0, 1) * A(w, 2, 2)) * d,
         (A(w, 0, 1) * A(w, 1, 2) - A(w, 0, 2) * A(w, 1, 1)) * d},
        {(A(w, 1, 2) * A(w, 2, 0) - A(w, 1, 0) * A(w, 2, 2)) * d, (A(w, 0, 0) * A(w, 2, 2) - A(w, 0, 2) * A(w, 2, 0)) * d,
         (A(w, 0, 2) * A(w, 1, 0) - A(w, 0, 0) * A(w, 1, 2)) * d},
        {(A(w, 1, 0) * A(w, 2, 1) - A(w, 1, 1) * A(w, 2, 0)) * d, (A(w, 0, 1) * A(w, 2, 0) - A(w, 0, 0) * A(w, 2, 1)) * d,
         (A(w, 0, 0) * A(w, 1, 1) - A(w, 0, 1) * A(w, 1, 0)) * d},
    };
    const float t[3] = {w[12], w[13], w[14]};
    for (int r = 0; r < 3; ++r) {
        out[4 * r] = inv[r][0];
        out[4 * r + 1] = inv[r][1];
        out[4 * r + 2] = inv[r][2];
        out[4 * r + 3] = -((inv[r][0] * t[0] + inv[r][1] * t[1]) + inv[r][2] * t[2]);
    }
}

uint32_t orc_inst_mirrors(const float w[16]) { return det3(w) < 0.0f ? 1u : 0u; }

void orc_inst_world_box(const float w[16], const orc_aabb* b, orc_aabb* out) {  // instance.rs:116-124
    float lo[3] = {3.40282347e38f, 3.40282347e38f, 3.40282347e38f}, hi[3] = {-3.40282347e38f, -3.40282347e38f, -3.40282347e38f};
    for (int k = 0; k < 8; ++k) {
        const float c[3] = {(k & 1) ? b->hi[0] : b->lo[0], (k & 2) ? b->hi[1] : b->lo[1], (k & 4) ? b->hi[2] : b->lo[2]};
        float p[3];
        orc_inst_to_world(w, c, p);
        for (int j = 0; j < 3; ++j) {
            lo[j] = fmin_(lo[j], p[j]);
            hi[j] = fmax_(hi[j], p[j]);
        }
    }
    for (int j = 0; j < 3; ++j) { out->lo[j] = lo[j]; out->hi[j] = hi[j]; }
    out->_a = 0.f; out->_b = 0.f;
}

void orc_mesh_texspace(const orc_mesh* m, float lo[3], float size[3]) {  // mesh.rs:43-54
    for (int k = 0; k < 3; ++k) {
        float half = 0.5f * (m->hi[k] - m->lo[k]);
        const float centre = 0.5f * (m->hi[k] + m->lo[k]);
        half = half == 0.0f ? 1.0f : std::max(half, 0.00001f);
        lo[k] = centre - half;
        size[k] = 2.0f * half;
    }
}

void orc_mesh_fit(orc_mesh* m, const float* verts) {  // mesh.rs:56-72
    if (m->vert_hi <= m->vert_lo) return;
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t v = m->vert_lo; v < m->vert_hi; ++v)
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::min(lo[k], verts[8 * (size_t)v + k]);
            hi[k] = std::max(hi[k], verts[8 * (size_t)v + k]);
        }
    for (int k = 0; k < 3; ++k) { m->lo[k] = lo[k]; m->hi[k] = hi[k]; }
}

void orc_inst_words(const orc_instance* inst, uint32_t blas_root, const orc_mesh* mesh, uint32_t out[40]) {
    // instance.rs:126-160
    for (int k = 0; k < 40; ++k) out[k] = 0u;
    const float* m = inst->world;
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 4; ++c) out[r * 4 + c] = bits(m[4 * c + r]);
    float inv[12];
    orc_inst_inverse(m, inv);
    for (int k = 0; k < 12; ++k) out[12 + k] = bits(inv[k]);
    out[24] = blas_root;
    out[25] = mesh->face_lo;
    out[26] = mesh->face_hi;
    out[27] = (is_identity(m) ? 1u : 0u) | ((inst->casts_shadow ? 0u : 1u) << 1) | (orc_inst_mirrors(m) << 2);
    for (int k = 0; k < 3; ++k) out[28 + k] = bits(inst->color[k]);
    out[31] = inst->hash;
    float lo[3], size[3];
    orc_mesh_texspace(mesh, lo, size);
    for (int k = 0; k < 3; ++k) {
        out[32 + k] = bits(lo[k]);
        out[35 + k] = bits(size[k]);
    }
    out[38] = inst->mesh;
}

uint32_t orc_forest_node_count(const orc_scene_in* in) {
    uint32_t n = tlas_slots(in);
    for (uint32_t m = 0; m < in->nmeshes; ++m) {
        uint32_t f = in->meshes[m].face_hi - in->meshes[m].face_lo;
        if (f) n += 2 * f - 1;
    }
    return n;
}

uint32_t orc_forest_inst_len(const orc_scene_in* in) { return in->ninsts * (ORC_INSTANCE_LEN + 2); }

uint32_t orc_forest_lights_len(const orc_scene_in* in) {
    std::vector<uint32_t> per_mesh(in->nmeshes, 0);
    for (uint32_t i = 0; i < in->ninsts; ++i)
        if (in->insts[i].mesh < in->nmeshes) per_mesh[in->insts[i].mesh]++;
    uint64_t n = 0;
    for (uint32_t m = 0; m < in->nmeshes; ++m) n += (uint64_t)(in->meshes[m].face_hi - in->meshes[m].face_lo) * per_mesh[m];
    return (uint32_t)(1 + 3 * n);
}

uint32_t orc_forest_build(const orc_scene_in* in, uint32_t leaf, orc_node* nodes, uint32_t* order, float* tri_cache,
                          uint32_t* faces, uint32_t* insts, uint32_t* lights, uint32_t* depths, float* extent, int threads) {
    Trees t;
    t.blas.resize(in->nmeshes);
    t.blas_depth.assign(in->nmeshes, 0);
    // Scene::boxes (scene.rs:182-184): each face's box in its mesh's space.
    std::vector<orc_aabb> boxes(in->nfaces);
    for (uint32_t f = 0; f < in->nfaces; ++f) {
        const uint32_t* fc = in->faces + 4 * (size_t)f;
        orc_of_tri(in->verts + 8 * (size_t)fc[0], in->verts + 8 * (size_t)fc[1], in->verts + 8 * (size_t)fc[2], &boxes[f]);
    }
    for (uint32_t f = 0; f < in->nfaces; ++f) order[f] = f;
    for (uint32_t m = 0; m < in->nmeshes; ++m) {
        const uint32_t lo = in->meshes[m].face_lo, cnt = in->meshes[m].face_hi - lo;
        if (!cnt) continue;
        t.blas[m].assign(2 * cnt - 1, orc_node{});
        std::vector<uint32_t> o(cnt);
        t.blas_depth[m] = orc_build(boxes.data() + lo, cnt, leaf, t.blas[m].data(), o.data(), threads);
        for (uint32_t s = 0; s < cnt; ++s) order[lo + s] = lo + o[s];
    }
    retlas(in, t);
    write_nodes(in, t, nodes, false);
    write_insts(in, t, insts);
    for (uint32_t s = 0; s < in->nfaces; ++s) {
        const uint32_t* fc = in->faces + 4 * (size_t)order[s];
        for (int k = 0; k < 4; ++k) faces[4 * (size_t)s + k] = fc[k];
        for (int k = 0; k < 3; ++k)
            for (int j = 0; j < 3; ++j) tri_cache[9 * (size_t)s + 3 * k + j] = in->verts[8 * (size_t)fc[k] + j];
    }
    write_lights(in, faces, lights);
    for (uint32_t m = 0; m < in->nmeshes; ++m) depths[m] = t.blas_depth[m];
    depths[in->nmeshes] = t.tlas_depth;
    *extent = extent_of(t);
    return orc_forest_node_count(in);
}

void orc_forest_reinstance(const orc_scene_in* in, orc_node* nodes, uint32_t* insts, uint32_t* lights,
                           const uint32_t* faces_slot, uint32_t* tlas_depth, float* extent) {
    // pose.rs:107-124: the mesh trees stay (their roots are read back from
    // `nodes`), the top tree, instance table and light table follow the
    // new matrices.
    Trees t;
    t.blas.resize(in->nmeshes);
    uint32_t at = tlas_slots(in);
    for (uint32_t m = 0; m < in->nmeshes; ++m) {
        uint32_t f = in->meshes[m].face_hi - in->meshes[m].face_lo;
        if (!f) continue;
        t.blas[m].assign(nodes + at, nodes + at + 2 * f - 1);
        at += 2 * f - 1;
    }
    retlas(in, t);
    write_nodes(in, t, nodes, true);
    write_insts(in, t, insts);
    write_lights(in, faces_slot, lights);
    *tlas_depth = t.tlas_depth;
    *extent = extent_of(t);
}

}  // extern "C"

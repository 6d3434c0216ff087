/* CPU oracle for the LBVH + path-tracer hot path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker the HIP path is held
 * to: a plain C++ restatement of phiray's own CPU twins (`bvh::Bvh`, the radix
 * sort model, `trace/tri.rs`) and of the WGSL walk and integrator, which have
 * no CPU twin in phiray. Only the tests, the smoke check and the benchmark's
 * CPU-baseline leg load it. The product library never links, loads or calls
 * anything here, and there is no CPU fallback built on it.
 *
 * Every function cites the reference file:line it restates. Arithmetic is
 * IEEE-754 single precision with contraction off (built with
 * -ffp-contract=off), in the order the reference writes it.
 *
 * Parity pinning (what checks each part independently of a retyped formula)
 * is listed in the tests that use it; anything not pinned says so there.
 */
#ifndef PHIRAY_ORACLE_H
#define PHIRAY_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* src/bvh/node.rs:60-65 — 32-byte box, padded. */
typedef struct { float lo[3]; float _a; float hi[3]; float _b; } orc_aabb;
/* src/bvh/node.rs:12-15 — one corner + index with the leaf bit. */
typedef struct { float p[3]; uint32_t i; } orc_half;
/* src/bvh/node.rs:30-33 */
typedef struct { orc_half lo, hi; } orc_node;
/* src/bvh/build.rs:10-15 */
typedef struct { uint32_t left, right, first, last; } orc_link;
/* src/trace/wgsl/layout.wgsl:26 */
typedef struct { float t; uint32_t tri; float u, v; uint32_t inst; } orc_hit;
/* A ray: origin, direction, tmax (origin/dir w unused). */
typedef struct { float o[3]; float tmax; float d[3]; float _pad; } orc_ray;

#define ORC_LEAF 0x80000000u
#define ORC_STACK 32u
#define ORC_NONE 0xffffffffu
#define ORC_INF 3.0e38f
#define ORC_MAT_WORDS 16

/* ---- LBVH (src/bvh/node.rs, src/bvh/mod.rs, src/bvh/build.rs) ---- */
uint32_t orc_code(const float p[3]);
uint64_t orc_morton(const float p[3], uint32_t group);
void orc_of_tri(const float a[3], const float b[3], const float c[3], orc_aabb* out);
void orc_scene(const orc_aabb* prims, uint32_t n, orc_aabb* out);
void orc_codes(const orc_aabb* prims, uint32_t n, uint32_t* codes, int threads);
void orc_sort_codes(const uint32_t* codes, uint32_t n, uint32_t* order, uint32_t* sorted, int threads);
void orc_radix_model(const uint32_t* codes, uint32_t n, uint32_t* order, uint32_t* sorted);
int32_t orc_delta(const uint32_t* codes, uint32_t n, int64_t i, int64_t j);
void orc_links(const uint32_t* sorted, uint32_t n, orc_link* links, int threads);
/* nodes must hold 2n-1 entries (n >= 1). Returns depth. */
uint32_t orc_assemble(const orc_link* links, uint32_t n, uint32_t leaf, orc_node* nodes);
void orc_refit(orc_node* nodes, uint32_t count, const uint32_t* order, const orc_aabb* prims);
/* Full Bvh::build: nodes[2n-1], order[n]; returns depth. */
uint32_t orc_build(const orc_aabb* prims, uint32_t n, uint32_t leaf, orc_node* nodes, uint32_t* order, int threads);
/* Bvh::build_compact: nodes_out must hold 2n-1; returns node count, depth via out. */
uint32_t orc_build_compact(const orc_aabb* prims, uint32_t n, uint32_t leaf, orc_node* nodes, uint32_t* order, uint32_t* depth);
/* Bvh::compact in place; returns new count. */
uint32_t orc_compact(orc_node* nodes, uint32_t count);
/* Bvh::reachable: writes indices (root first), returns count. */
uint32_t orc_reachable(const orc_node* nodes, uint32_t count, uint32_t* out);

/* ---- triangle test and walk (src/trace/tri.rs, src/trace/wgsl/bvh.wgsl) ---- */
typedef struct { uint32_t kz; uint32_t flip; float s[3]; } orc_rayspace;
void orc_ray_space(const float d[3], orc_rayspace* out);
/* Returns 1 on a hit with (t,u,v) in out. */
int orc_tri_t(const float a[3], const float b[3], const float c[3], const float o[3], const orc_rayspace* r, float out[3]);
int orc_moller_trumbore(const float a[3], const float b[3], const float c[3], const float o[3], const float d[3], float out[3]);
void orc_safe_inv(const float d[3], float out[3]);
float orc_box_t(const orc_node* n, const float o[3], const float inv[3], float best);
/* Walk one tree (trace_blas for an identity instance). tris: 9 floats per slot.
 * stats (optional, 2 counters per ray): inner steps, triangle tests. */
void orc_raycast(const orc_node* nodes, const float* tris, const orc_ray* rays, uint32_t nrays,
                 int any_hit, orc_hit* hits, uint32_t* stats, int threads);
/* The same with the faces' material table, so any-hit rays pass glass and
 * cut-outs as bvh.wgsl:158-160 has them (faces: 4 words a slot; mats: 16). */
void orc_raycast_mat(const orc_node* nodes, const float* tris, const uint32_t* faces, const float* mats,
                     const orc_ray* rays, uint32_t nrays, int any_hit, orc_hit* hits, uint32_t* stats, int threads);

/* ---- deterministic math shared in spirit with the HIP path ---- */
float orc_sin(float x);
float orc_cos(float x);
float orc_exp2(float x);
float orc_log2(float x);

/* ---- path tracer (src/trace/wgsl/shade.wgsl, integrate.wgsl, light.wgsl, brdf.wgsl, util.wgsl) ---- */
uint32_t orc_hash(uint32_t x);
float orc_rnd(uint32_t* seed);

/* Camera subset (src/trace/camera.rs:12-228) — fields the Rendered path reads. */
typedef struct {
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
    uint32_t w, h, rr_from, stream;  /* added to the sample index in the seed */
    float adapt[4];  /* Adaptive Sampling: threshold, min samples, step, on (camera.rs `adapt`) */
    float view[4];   /* colour management: transform, exposure (stops), gamma (0: 1), look (camera.rs:53-56) */
} orc_camera;

/* The display pixel (src/trace/wgsl/view.wgsl:5-107) under the view words
 * (transform 0 Standard, 1 Filmic, 2 AgX; exposure; gamma, 0 meaning 1;
 * Filmic look 1..6): exposure, the transform, gamma; and the RGBA8 word, red
 * in the low byte, each channel rounded to nearest. */
void orc_view_transform(const float lin[3], const float view[4], float out[3]);
uint32_t orc_pack_rgba8(const float rgb[3], float a);

/* Input triangle: src/trace/kernel.rs:11-40 (Tri) rows a..nc, color, emit —
 * the texture, UV and layered-material rows are not carried. 128 bytes. */
typedef struct {
    float a[3]; float rough;
    float na[3]; float trans;
    float b[3]; float metal;
    float nb[3]; float ior;
    float c[3]; float _c;
    float nc[3]; float _nc;
    float color[4];
    float emit[3]; float bump;
} orc_tri;

/* Scene in leaf (slot) order, as the tracer keeps it.
 * tris: 9 floats / slot; faces: 4 words / slot (3 vertex ids + material);
 * verts: 8 floats / vertex (pos.xyz, 1, oct.xy, u, v);
 * mats: ORC_MAT_WORDS floats / material; lights: [count, face ids, inst ids, cdf bits]. */
typedef struct {
    const orc_node* nodes; uint32_t node_count;
    const float* tris; const uint32_t* faces; const float* verts;
    const float* mats; const uint32_t* lights;
    /* A forest (forest != 0): the top tree at node 0 and instance words as
     * Forest::inst_words lays them out (ORC_INSTANCE_LEN words a record, then
     * the top-tree slot order, then the draw order). Otherwise one tree at
     * node 0 under one identity instance, and insts is unused. */
    const uint32_t* insts; uint32_t ninsts; uint32_t forest;
} orc_scene_t;

/* ---- instances, meshes and the forest (src/trace/instance.rs, mesh.rs, forest.rs) ---- */
#define ORC_INSTANCE_LEN 40u
#define ORC_INST 0x40000000u
/* src/trace/instance.rs:23-44: world is m[col][row] (translation in m[3]). */
typedef struct {
    float world[16];
    uint32_t mesh;
    uint32_t casts_shadow;
    float color[3];
    uint32_t hash;
} orc_instance;
/* src/trace/mesh.rs:12-22: face and vertex ranges; bounds as Mesh::fit leaves them. */
typedef struct { uint32_t face_lo, face_hi, vert_lo, vert_hi; float lo[3], hi[3]; } orc_mesh;
/* A scene of meshes and instances (src/trace/scene.rs:40-67): vertices 8
 * floats (object-space position xyz, w, packed normal xy, uv), faces 4 words
 * (3 vertex ids + material), materials ORC_MAT_WORDS floats. The meshes tile
 * the faces in order. */
typedef struct {
    const float* verts; uint32_t nverts;
    const uint32_t* faces; uint32_t nfaces;
    const float* mats; uint32_t nmats;
    const orc_mesh* meshes; uint32_t nmeshes;
    const orc_instance* insts; uint32_t ninsts;
} orc_scene_in;

void orc_inst_inverse(const float world[16], float out[12]);           /* instance.rs:75-93 */
uint32_t orc_inst_mirrors(const float world[16]);                        /* instance.rs:111-113 */
void orc_inst_to_world(const float world[16], const float p[3], float out[3]); /* instance.rs:68-71 */
void orc_inst_world_box(const float world[16], const orc_aabb* b, orc_aabb* out); /* instance.rs:116-124 */
void orc_mesh_texspace(const orc_mesh* m, float lo[3], float size[3]);  /* mesh.rs:43-54 */
void orc_mesh_fit(orc_mesh* m, const float* verts);                     /* mesh.rs:56-72, 8-float vertices */
/* Instance::words (instance.rs:126-160). */
void orc_inst_words(const orc_instance* inst, uint32_t blas_root, const orc_mesh* mesh, uint32_t out[40]);
/* Sizes the forest build writes: node count (Forest::tlas_slots + each mesh
 * tree), instance words, emitter table words. */
uint32_t orc_forest_node_count(const orc_scene_in* in);
uint32_t orc_forest_inst_len(const orc_scene_in* in);
uint32_t orc_forest_lights_len(const orc_scene_in* in);
/* Forest::build (forest.rs:44-66, with Bvh::build for every tree) + nodes +
 * inst_words, the faces put in leaf order (Scene::reorder), the triangle
 * cache (9 floats a slot, object space), the emitter table
 * (lamp.rs:51-97) and the extent (forest.rs:189-199). order[slot] = input
 * face. depths: one per mesh, then the top tree's. Returns the node count. */
uint32_t orc_forest_build(const orc_scene_in* in, uint32_t leaf, orc_node* nodes, uint32_t* order, float* tri_cache,
                          uint32_t* faces, uint32_t* insts, uint32_t* lights, uint32_t* depths, float* extent, int threads);
/* Forest::retlas after new matrices (pose.rs:107-124): the top tree, the
 * instance words and the extent again; the mesh trees stay. `nodes`,
 * `insts`, `lights` as orc_forest_build wrote them; faces in leaf order. */
void orc_forest_reinstance(const orc_scene_in* in, orc_node* nodes, uint32_t* insts, uint32_t* lights,
                           const uint32_t* faces_slot, uint32_t* tlas_depth, float* extent);
/* trace (bvh.wgsl:120-232) through a scene: forest or single tree. */
void orc_trace_scene(const orc_scene_t* s, const orc_ray* rays, uint32_t nrays, int any_hit, orc_hit* hits,
                     uint32_t* stats, int threads);

/* Tracer::new over `tris` (src/trace/build.rs:11-13, src/trace/scene/add.rs,
 * src/trace/scene.rs:188-190, src/trace/lamp.rs:51-97) with the LBVH as its
 * tree: everything in leaf (slot) order. Capacities: nodes 2n-1, order n,
 * tris 9n, faces 4n, verts 8*3n, mats 16n, lights 1+3n. Returns node count. */
uint32_t orc_scene_build(const orc_tri* tris, uint32_t n, uint32_t leaf,
                         orc_node* nodes, uint32_t* order, float* tri_cache, uint32_t* faces,
                         float* verts, uint32_t* nverts, float* mats, uint32_t* nmats,
                         uint32_t* lights, uint32_t* depth, float* extent, int threads);
/* The camera as Tracer::sample_in stamps it (src/trace/sample.rs:104-118):
 * w, h, ao from the extent when unset, eps, glow = emitter count, frame. */
void orc_stamp_camera(orc_camera* cam, uint32_t w, uint32_t h, float extent, uint32_t glow, uint32_t frame);

/* One sample (cam->frame) for pixels [first, first+count) accumulated into px
 * (34 words per pixel, src/trace/read.rs:37-56). counters: closest, shadow rays. */
void orc_sample(const orc_scene_t* s, const orc_camera* cam, uint32_t first, uint32_t count,
                uint32_t* px, uint64_t* rays, int threads);


/* Adaptive Sampling's tail stages after sample cam->frame
 * (src/trace/wgsl/adapt.wgsl): the convergence check (stage 0, a no-op
 * unless the sample is a step boundary past the minimum), then the
 * four-neighbour dilation (stage 1). */
void orc_adapt_check(const orc_camera* cam, uint32_t* px);
void orc_adapt_dilate(const orc_camera* cam, uint32_t* px);

/* One a-trous pass (src/kernel/denoise.wgsl:47-134) over the image buffer;
 * the last writes the display planes under the view words. */
void orc_denoise_pass(uint32_t* px, uint32_t w, uint32_t h, uint32_t step, uint32_t round,
                      uint32_t last, uint32_t keep, const float strength[4], const float view[4]);
/* Tracer::denoise_with (src/trace/read.rs:85-119): steps 1, 2, 4, 8 with
 * strengths [64, 8, 40, 7]. */
void orc_denoise(uint32_t* px, uint32_t w, uint32_t h, uint32_t keep, const float view[4]);

int orc_hardware_threads(void);

/* ---- the rigid-body World (src/sim/), sim.cpp ---------------------------- */
/* Shape (src/sim/state.rs, 32 B): kind 0 box, 1 sphere, 2 capsule; half
 * extents (box) or half length on z (capsule); radius. */
typedef struct { uint32_t kind, data, count; float radius; float half[3]; float _a; } orc_shape;
/* BodyDef (src/sim/build.rs:14-61): kind 0 active, 1 passive, 2 kinematic;
 * q is xyzw. */
typedef struct {
    uint32_t kind, shape; float mass;
    float x[3]; float q[4]; float v[3]; float w[3];
    float scale[3]; float com[3];
    float friction, restitution, margin, lin_damp, ang_damp;
    float sleep[2]; uint32_t has_sleep, asleep, groups, obj;
} orc_body_def;
typedef struct orc_world orc_world;
orc_world* orc_world_new(void);
void orc_world_free(orc_world* w);
void orc_world_configure(orc_world* w, uint32_t substeps, uint32_t iterations, float time_scale, float max_recover,
                         float sleep_after, const float gravity[3]);
int orc_world_add_shape(orc_world* w, const orc_shape* s, uint32_t* id);
int orc_world_add_body(orc_world* w, const orc_body_def* d, uint32_t* id);
/* World::step (src/sim/world.rs:104-129). */
int orc_world_step(orc_world* w, float frame_dt);
uint32_t orc_world_body_count(const orc_world* w);
uint32_t orc_world_contact_count(const orc_world* w);
/* The 64-byte Body records (src/sim/state.rs). */
int orc_world_bodies(const orc_world* w, void* out);
/* World::poses (src/sim/world.rs:225-228): 16 floats a body, m[col][row]. */
int orc_world_poses(const orc_world* w, float* out);
/* World::energy (src/sim/world.rs:231-243). */
float orc_world_energy(const orc_world* w);
/* The last broadphase's pairs (i < j, sorted). */
int orc_world_pair_list(const orc_world* w, uint32_t* out, uint32_t cap, uint32_t* n);

#ifdef __cplusplus
}
#endif
#endif

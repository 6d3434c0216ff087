/* phiray_hip.h — the C ABI of the gfx950 LBVH + path-tracer library.
 *
 * These are the entry points a thin Rust FFI crate binds to re-back phiray's
 * `Gpu`, `GpuArray<T>`, `GpuBvh`, `GpuSort`-sorted build, `Tracer` and
 * `Clock` with hand-written HIP kernels. Each declaration names the reference
 * interface it replaces (paths relative to the phiray repository).
 *
 * Conventions:
 *  - every function returning `int` returns 0 (PHR_OK) or a negative code;
 *    phr_last_error() describes the last failure on the calling thread;
 *  - plain pointers and sizes only; POD records cross by bytes in phiray's
 *    own layouts: Aabb 32 B, Node 32 B, Link 16 B, Hit 20 B, ray 32 B
 *    {o.xyz, tmax, d.xyz, pad}; the tracer takes a 128 B triangle and a 208 B
 *    camera (the fields the Rendered path reads), made from phiray's own
 *    760 B `Tri` and 5,776 B `Camera` by phr_tri_from_phiray and
 *    phr_camera_from_phiray;
 *  - one HIP stream per phr_gpu; calls on one phr_gpu are serialised by the
 *    caller (phiray's `Tracer` is `&mut` from one render thread).
 */
#ifndef PHIRAY_HIP_H
#define PHIRAY_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PHR_OK 0
#define PHR_ERR_ARG -1
#define PHR_ERR_HIP -2
#define PHR_ERR_UNSUPPORTED -3
#define PHR_ERR_SIZE -4

typedef struct phr_gpu phr_gpu;
typedef struct phr_array phr_array;
typedef struct phr_bvh phr_bvh;
typedef struct phr_tracer phr_tracer;
typedef struct phr_clock phr_clock;
typedef struct phr_pass phr_pass;
typedef struct phr_sort phr_sort;
typedef struct phr_kpcn phr_kpcn;
typedef struct phr_world phr_world;

/* ---- errors / version ---------------------------------------------------- */
const char* phr_last_error(void);
uint32_t phr_abi_version(void);

/* ---- device: Gpu (src/gpu/mod.rs:84-331) ----------------------------------- */
/* Gpu::new (src/gpu/mod.rs:108); PHR_ERR_UNSUPPORTED = None (no device). */
int phr_gpu_open(int device, phr_gpu** out);
void phr_gpu_close(phr_gpu* g);
/* Gpu::shared (src/gpu/mod.rs:113-118): the process's device 0, opened on
 * first use; phr_gpu_close leaves it open. */
int phr_gpu_shared(phr_gpu** out);
/* Gpu::wait (src/gpu/mod.rs:272). */
int phr_gpu_wait(phr_gpu* g);
/* Gpu::wait_for (src/gpu/mod.rs:292): *done = 1 when drained within limit. */
int phr_gpu_wait_for(phr_gpu* g, uint64_t limit_ns, int* done);
/* Gpu::max_binding (src/gpu/mod.rs:253). */
uint64_t phr_gpu_max_binding(phr_gpu* g);
int phr_gpu_cu_count(phr_gpu* g);
const char* phr_gpu_arch(phr_gpu* g);
void* phr_gpu_stream(phr_gpu* g);

/* ---- arrays: GpuArray<T> (src/gpu/array.rs:7-188) --------------------------- */
/* GpuArray::new (src/gpu/array.rs:20-23): capacity, no items. */
int phr_array_new(phr_gpu* g, uint32_t elem, uint64_t cap, phr_array** out);
/* GpuArray::zeroed (src/gpu/array.rs:27-29). */
int phr_array_zeroed(phr_gpu* g, uint32_t elem, uint64_t len, phr_array** out);
/* GpuArray::from (src/gpu/array.rs:31-35). */
int phr_array_from(phr_gpu* g, uint32_t elem, const void* items, uint64_t len, phr_array** out);
/* GpuArray::write (src/gpu/array.rs:47-56): *grown = reallocated. */
int phr_array_write(phr_array* a, const void* items, uint64_t len, int* grown);
/* GpuArray::write_at (src/gpu/array.rs:96-107): *written = 0 (the reference's
 * false), and nothing written, when at + len passes the array's length. */
int phr_array_write_at(phr_array* a, uint64_t at, const void* items, uint64_t len, int* written);
/* GpuArray::write_front (src/gpu/array.rs:85-94): write_at from item 0. */
int phr_array_write_front(phr_array* a, const void* items, uint64_t len, int* written);
/* GpuArray::split (src/gpu/array.rs:58-83): items over `chunks` new arrays of
 * at most limit / elem items, *cut items each; PHR_ERR_SIZE (None) when they
 * do not fit. out has room for `chunks` handles. */
int phr_array_split(phr_gpu* g, uint32_t elem, const void* items, uint64_t len, uint64_t limit, uint32_t chunks,
                    phr_array** out, uint32_t* cut);
/* GpuArray::read_first (src/gpu/array.rs:138-142): min(count, len) items, *got of them. */
int phr_array_read_first(phr_array* a, uint64_t count, void* out, uint64_t* got);
/* GpuArray::read_range (src/gpu/array.rs:136-187), blocking. */
int phr_array_read_range(phr_array* a, uint64_t from, uint64_t count, void* out);
uint64_t phr_array_len(const phr_array* a);
uint64_t phr_array_cap(const phr_array* a);
uint32_t phr_array_elem(const phr_array* a);
void* phr_array_device_ptr(phr_array* a);
void phr_array_free(phr_array* a);
/* A plain streaming copy kernel (the HBM-counter calibration reference; no
 * counterpart in phiray). */
int phr_array_copy_kernel(phr_array* dst, phr_array* src);

/* ---- timing: Clock (src/gpu/clock.rs:17-87) --------------------------------- */
/* Clock::new (src/gpu/clock.rs:41). */
int phr_clock_new(phr_gpu* g, phr_clock** out);
/* Clock::run split in two (src/gpu/clock.rs:54-72): stamp, enqueue work, then
 * stop = stamp, wait, and report the GPU span and the wall time. */
int phr_clock_start(phr_clock* c);
int phr_clock_stop(phr_clock* c, double* gpu_s, double* wall_s);
void phr_clock_free(phr_clock* c);

/* ---- kernels and recorded passes: Kernel / Pass (src/kernel/mod.rs:51-133,
 *      src/gpu/pass.rs:18-155) ----------------------------------------------- */
/* Kernel lookup by phiray NAME ("bvh.morton", "bvh.karras", "bvh.refit",
 * "radix.seed|clear|count|offset|scatter", "radix.spread", "denoise"): its
 * Param size in bytes and binding count (BIND_LIST length). */
int phr_kernel_known(const char* name, uint32_t* param_size, uint32_t* binds);
int phr_pass_new(phr_gpu* g, phr_pass** out);
void phr_pass_free(phr_pass* p);
/* Pass::add / add_over (src/gpu/pass.rs:64-101): params' first two words
 * (n, groups_x) are filled in; a binding-count mismatch is PHR_ERR_ARG. */
int phr_pass_add(phr_pass* p, const char* name, uint32_t n, uint32_t items, const void* params, uint32_t param_len,
                 phr_array* const* arrays, uint32_t count, uint32_t* step);
/* Pass::set (src/gpu/pass.rs:104-106). */
int phr_pass_set(phr_pass* p, uint32_t step, uint32_t offset, const void* data, uint32_t len);
/* Pass::run (src/gpu/pass.rs:109-113): all steps in order, replayed as one
 * hipGraph (captured again after add/set/clear). Enqueued, not waited. */
int phr_pass_run(phr_pass* p);
/* Pass::run_step (src/gpu/pass.rs:126-137). */
int phr_pass_run_step(phr_pass* p, uint32_t step);
/* Pass::clear / len (src/gpu/pass.rs:140-148). */
void phr_pass_clear(phr_pass* p);
uint32_t phr_pass_len(const phr_pass* p);
/* Pass::compiled (src/gpu/pass.rs:151-154): kernel types the pass uses. */
uint32_t phr_pass_compiled(const phr_pass* p);

/* ---- BVH: GpuBvh (src/bvh/gpu.rs:11-104) ------------------------------------ */
/* GpuBvh::build (src/bvh/gpu.rs:23-75): prims are n Aabb (32 B) on the host.
 * The whole build (bounds, Morton, sort, Karras, packing, refit) runs on the
 * device; only the depth comes back. */
int phr_bvh_build(phr_gpu* g, const void* prims, uint32_t n, uint32_t leaf, phr_bvh** out);
/* The same over an Aabb array already on the device (borrowed, not owned). */
int phr_bvh_build_on(phr_gpu* g, phr_array* prims, uint32_t leaf, phr_bvh** out);
/* No counterpart in the reference (the tuner the gfx950 build adds): keys a
 * thread holds in a sort pass, 8 (2,048-key tiles, the default) or 16
 * (4,096); every choice gives the same tree. phr_bvh_autotune times `reps`
 * rebuilds of this tree with each, keeps the faster, and reports it and the
 * two times (ms[2], may be null). */
int phr_bvh_set_sort_tile(phr_bvh* b, uint32_t items);
uint32_t phr_bvh_sort_tile(const phr_bvh* b);
int phr_bvh_autotune(phr_bvh* b, uint32_t reps, uint32_t* items, double* ms);
/* Build again from the current prims (same count), enqueued / then synced. */
int phr_bvh_rebuild_async(phr_bvh* b);
int phr_bvh_rebuild(phr_bvh* b);
/* GpuBvh::refit (src/bvh/gpu.rs:88-90): bounds again, topology kept. */
int phr_bvh_refit_async(phr_bvh* b);
int phr_bvh_refit(phr_bvh* b);
/* GpuBvh::update (src/bvh/gpu.rs:93-99): *refitted = 0 when n changed (rebuild). */
int phr_bvh_update(phr_bvh* b, const void* prims, uint32_t n, int* refitted);
/* GpuBvh::download (src/bvh/gpu.rs:102-104): nodes[2n-1], order[n], depth. */
int phr_bvh_download(phr_bvh* b, void* nodes, uint32_t* order, uint32_t* depth);
/* GpuSort::read's codes half (src/bvh/sort.rs:109-111): sorted Morton codes. */
int phr_bvh_codes(phr_bvh* b, uint32_t* sorted_codes);
/* Many trees in one build: seg_lo holds segs + 1 non-decreasing starts from 0
 * to n, and segment s is boxes [seg_lo[s], seg_lo[s+1]). Each segment gets the
 * tree Bvh::build (src/bvh/build.rs:94-100) gives its boxes alone, with its
 * own scene box: one tree per mesh as Forest::build makes them
 * (src/trace/forest.rs:44-66), or a batch of independent scenes. The node
 * array holds the trees back to back; a tree's children and leaf ranges are
 * indices into the whole array and the whole box list, as Forest::nodes lays
 * the mesh trees out (src/trace/forest.rs:130-160). At most 65,536 segments. */
int phr_bvh_build_batch(phr_gpu* g, const void* prims, uint32_t n, const uint32_t* seg_lo, uint32_t segs, uint32_t leaf,
                        phr_bvh** out);
int phr_bvh_build_batch_on(phr_gpu* g, phr_array* prims, const uint32_t* seg_lo, uint32_t segs, uint32_t leaf, phr_bvh** out);
/* Bvh::build_compact (src/bvh/build.rs:104-136) of the tree's primitives: the
 * same reachable tree numbered breadth first, children adjacent, computed on
 * the device from this build. *count nodes (32 B each) into `nodes` (cap
 * nodes at most; null: only the count). One tree, below 2^26 primitives; the
 * order and depth are phr_bvh_download's. */
int phr_bvh_compact(phr_bvh* b, void* nodes, uint64_t cap, uint32_t* count);
/* One segment's tree relative to itself, as GpuBvh::download would give it
 * for that segment's boxes alone: nodes[2 n_s - 1], order[n_s], depth. */
int phr_bvh_download_segment(phr_bvh* b, uint32_t seg, void* nodes, uint32_t* order, uint32_t* depth);
uint32_t phr_bvh_segments(const phr_bvh* b);
uint32_t phr_bvh_segment_depth(const phr_bvh* b, uint32_t seg);
uint32_t phr_bvh_len(const phr_bvh* b);
/* The deepest segment's depth (Bvh::depth for a single tree). */
uint32_t phr_bvh_depth(const phr_bvh* b);
uint32_t phr_bvh_leaf(const phr_bvh* b);
/* GpuBvh's public fields prims / order / nodes (src/bvh/gpu.rs:12-15). */
phr_array* phr_bvh_nodes(phr_bvh* b);
phr_array* phr_bvh_order(phr_bvh* b);
phr_array* phr_bvh_prims(phr_bvh* b);
void phr_bvh_free(phr_bvh* b);
/* No counterpart in the reference. For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): phr_bvh_build_on's tree over n Aabb (32 B) at the device
 * pointer `prims`, written to caller-owned device memory — `nodes` (2n - 1
 * Node, 32 B; only reachable nodes are written), `order` (n u32) and, when
 * not null, `depth` (one u32, device memory). Enqueued on the device's
 * stream; nothing is allocated, freed or waited on. The build's scratch is
 * one static device array of PHR_STATIC_BUILD_SCRATCH_BYTES per GPU, held
 * until the process ends and shared by every call on that GPU: calls that
 * overlap on one GPU race. n above PHR_STATIC_BUILD_MAX is refused with
 * PHR_ERR_SIZE. */
#define PHR_STATIC_BUILD_MAX 262144u
#define PHR_STATIC_BUILD_SCRATCH_BYTES 22570240ull
int phr_bvh_build_static(phr_gpu* g, const void* prims, uint32_t n, uint32_t leaf, void* nodes, uint32_t* order,
                         uint32_t* depth);

/* ---- sort: GpuSort (src/bvh/sort.rs:13-111) ---------------------------------- */
/* GpuSort::record (src/bvh/sort.rs:30-51): a sort of n u32 codes held in a
 * device array (borrowed; new codes may be written before each run). */
int phr_sort_record(phr_gpu* g, phr_array* codes, uint32_t n, phr_sort** out);
/* GpuSort::run (src/bvh/sort.rs:61-63): enqueued. Stable: the result is the
 * sort by (code, slot) of the reference's 64-bit LSD passes. */
int phr_sort_run(phr_sort* s);
/* GpuSort::keys / code (src/bvh/sort.rs:66-73): (original slot, code) pairs
 * and the sorted codes alone, on the device. */
phr_array* phr_sort_keys(phr_sort* s);
phr_array* phr_sort_code(phr_sort* s);
/* GpuSort::len / steps / bytes (src/bvh/sort.rs:76-95): steps is the
 * dispatch count (6 here: histograms, 4 passes, keys). */
uint32_t phr_sort_len(const phr_sort* s);
uint32_t phr_sort_steps(const phr_sort* s);
uint64_t phr_sort_bytes(const phr_sort* s);
/* GpuSort::order / read (src/bvh/sort.rs:99-111), blocking; either may be null. */
int phr_sort_read(phr_sort* s, uint32_t* order, uint32_t* sorted);
void phr_sort_free(phr_sort* s);

/* ---- phiray's own records ------------------------------------------------------ */
/* phiray's `Tri` (src/trace/kernel.rs:11-65, 760 bytes a record) into the
 * tracer's 128-byte triangle: positions, normals, colour, roughness,
 * metallic, transmission, IOR, emission and bump copied; PHR_ERR_UNSUPPORTED
 * (and nothing more written) for a record with a texture, normal map,
 * texture-driven input, subsurface, sheen, coat, a second lobe or a
 * procedural, which the restated path does not shade. Host only. What
 * Tracer::new (src/trace/build.rs:11-13) is handed goes through this first. */
#define PHR_PHIRAY_TRI_BYTES 760
#define PHR_PHIRAY_CAMERA_BYTES 5776
int phr_tri_from_phiray(const void* tris, uint64_t n, void* out);
/* phiray's `Camera` (src/trace/camera.rs:12-228, 5,776 bytes) into the
 * tracer's 208-byte camera, roulette and sample streams as the reference's
 * (rr_from 0, stream 0); PHR_ERR_UNSUPPORTED for a camera whose frame the
 * restated path would not draw: a non-black theme sky, the grid overlay, a
 * transparent film or shown world, an environment image, depth of field, a
 * tile of a larger frame, TAA jitter, light clamping or a lit lamp. Host
 * only. Tracer::sample's camera (src/trace/sample.rs:96-99) goes through it. */
int phr_camera_from_phiray(const void* camera, void* out);

/* ---- tracer: Tracer (src/trace/mod.rs:50-245) -------------------------------- */
/* Tracer::new (src/trace/build.rs:11-13) over n triangle records (128 B each,
 * Tri's rows a..nc, color, emit); the tree is the device-built LBVH with
 * `leaf` primitives a leaf. PHR_ERR_ARG for an empty scene (None). */
int phr_tracer_new(phr_gpu* g, const void* tris, uint32_t n, uint32_t leaf, phr_tracer** out);
void phr_tracer_free(phr_tracer* t);
uint32_t phr_tracer_len(const phr_tracer* t);
uint32_t phr_tracer_depth(const phr_tracer* t);
float phr_tracer_extent(const phr_tracer* t);
uint32_t phr_tracer_glow(const phr_tracer* t);
phr_bvh* phr_tracer_bvh(phr_tracer* t);
/* Slot -> input triangle (Scene::reorder, src/trace/scene.rs:188-190). */
int phr_tracer_order(phr_tracer* t, uint32_t* out);
/* The emitter table (src/trace/lamp.rs:51-97): [count, faces, insts, cdf]. */
int phr_tracer_lights(phr_tracer* t, uint32_t* out, uint32_t cap, uint32_t* len);
/* trace() over host rays (src/trace/wgsl/bvh.wgsl:194-232): hits are 20 B
 * {t, slot, u, v, inst}; slot = 0xffffffff for a miss. */
int phr_tracer_raycast(phr_tracer* t, const void* rays, uint32_t nrays, int any_hit, void* hits);
/* The same over device arrays (rays: 32 B records; hits: 20 B), enqueued. */
int phr_tracer_raycast_on(phr_tracer* t, phr_array* rays, int any_hit, phr_array* hits);
/* Tracer::set_samples (src/trace/mod.rs:242). */
void phr_tracer_set_samples(phr_tracer* t, uint32_t n);
/* Tracer::warm (src/trace/build.rs:27-69): load every path-kernel build on
 * the device ahead of its first frame; names[i] (static strings) and ms[i]
 * per build, *count entries (at most cap). */
int phr_tracer_warm(phr_gpu* g, uint32_t cap, const char** names, double* ms, uint32_t* count);
/* Tracer::samples (src/trace/mod.rs:152-155). */
uint32_t phr_tracer_samples(phr_tracer* t);
/* Tracer::set_adaptive (src/trace/mod.rs:235-240): Adaptive Sampling's
 * [threshold, min samples, step, on], stamped onto every later camera
 * (before the first call, the camera's own `adapt` words are used). */
int phr_tracer_set_adaptive(phr_tracer* t, const float adapt[4]);
/* Tracer::sample (src/trace/sample.rs:96-99), spp samples in one launch;
 * cam is the 208-byte camera record. Each launch also writes the display
 * plane and the packed RGBA8 word (src/trace/wgsl/util.wgsl:101-145) under
 * the camera's view words (src/trace/wgsl/view.wgsl: Standard, Filmic and
 * its looks, AgX; exposure; gamma). */
int phr_tracer_sample(phr_tracer* t, const void* cam, uint32_t w, uint32_t h, uint32_t spp, uint32_t* frame, int* done);
/* Tracer::sample_in (src/trace/sample.rs:104-249): the same in `slices`
 * launches over the image's tiles, the device drained after each and go(user)
 * asked whether to carry on; a zero abandons the call (*finished = 0, the
 * reference's None) and the next call starts the image over. */
typedef int (*phr_go_fn)(void* user);
int phr_tracer_sample_in(phr_tracer* t, const void* cam, uint32_t w, uint32_t h, uint32_t spp, uint32_t slices,
                         phr_go_fn go, void* user, uint32_t* frame, int* done, int* finished);
/* Tracer::time (src/trace/sample.rs:253-255). */
int phr_tracer_time(phr_tracer* t, double* submit_s, double* gpu_s);
/* Rays traced so far: closest-hit and shadow (any-hit). */
int phr_tracer_rays(phr_tracer* t, uint64_t* closest, uint64_t* shadow);
/* Walk counters: [closest rays, shadow rays, inner-node steps, triangle tests]
 * of the path kernel, then [steps, tests] of the raycasts. */
int phr_tracer_counters(phr_tracer* t, uint64_t* out6);
/* The image buffer (src/trace/read.rs:37-56), raw words. */
int phr_tracer_px(phr_tracer* t, uint64_t from, uint64_t count, uint32_t* out);
phr_array* phr_tracer_image(phr_tracer* t);
/* Tracer::denoise (keep = 0) / denoise_shown (keep = 1), src/trace/read.rs:73-119:
 * the four a-trous passes of src/kernel/denoise.wgsl over the image buffer. */
int phr_tracer_denoise(phr_tracer* t, int keep);
/* Tracer::radiance (src/trace/read.rs:241-251): w*h rgb floats. */
int phr_tracer_radiance(phr_tracer* t, float* out);
/* Tracer::albedo / normal / depth (src/trace/read.rs:253-307): the guide
 * planes' sums over the per-pixel count; w*h rgb, rgb and one float. */
int phr_tracer_albedo(phr_tracer* t, float* out);
int phr_tracer_normal(phr_tracer* t, float* out);
int phr_tracer_depth_plane(phr_tracer* t, float* out);
/* Tracer::pixels (src/trace/read.rs:322-327): w*h RGBA8 pixels, 4*w*h bytes.
 * Tracer::render (src/trace/sample.rs:88-91) is phr_tracer_sample then this. */
int phr_tracer_pixels(phr_tracer* t, uint8_t* out);
/* Tracer::display (src/trace/read.rs:257-263): the shown value, w*h rgb floats. */
int phr_tracer_display(phr_tracer* t, float* out);
/* Tracer::show (src/trace/read.rs:182-229, src/trace/wgsl/show.wgsl): put a
 * scene-linear image (count rgb triples, truncated or zero-filled to w*h) in
 * place of the traced mean and run the display step; *shown = 0 when there
 * is no image or camera yet (the reference's false). */
int phr_tracer_show(phr_tracer* t, const float* linear, uint64_t count, int* shown);
/* Tracer::size (src/trace/read.rs:162-164): the image's w, h (0, 0 before a sample). */
int phr_tracer_size(phr_tracer* t, uint32_t* w, uint32_t* h);

/* ---- a scene of meshes under instances: the forest -------------------------- */
/* A mesh (src/trace/mesh.rs:12-22): its face and vertex ranges and the box of
 * its vertices (Mesh::fit). 40 bytes. */
typedef struct {
    uint32_t face_lo, face_hi, vert_lo, vert_hi;
    float lo[3], hi[3];
} phr_mesh;
/* An instance (src/trace/instance.rs:23-44): world is m[col][row] with the
 * translation in m[3]; casts_shadow 0 lets shadow rays pass it. 88 bytes. */
typedef struct {
    float world[16];
    uint32_t mesh;
    uint32_t casts_shadow;
    float color[3];
    uint32_t hash;
} phr_instance;
/* A scene (src/trace/scene.rs:40-67): vertices 32 B each (object-space
 * position, w = 1, packed normal xy, uv), faces 16 B (3 vertex ids and a
 * material), materials 64 B (color rgba, rough, metal, trans, ior, emit rgb,
 * bump, 4 zero words), meshes tiling the faces in order, instances. */
typedef struct {
    const void* verts; uint32_t nverts;
    const void* faces; uint32_t nfaces;
    const void* mats; uint32_t nmats;
    const phr_mesh* meshes; uint32_t nmeshes;
    const phr_instance* insts; uint32_t ninsts;
} phr_scene_desc;
/* Tracer::from_scene (src/trace/build.rs:15-25, src/trace/forest.rs:44-66):
 * one tree per mesh over its faces in its own space and a top tree over the
 * instances' world boxes, all built on the device in two LBVH launches
 * (the mesh trees as one batch), laid out as Forest::nodes lays them out.
 * PHR_ERR_ARG for an empty scene or one whose instances place nothing (None),
 * PHR_ERR_UNSUPPORTED for a material the path does not carry (transmission). */
int phr_tracer_from_scene(phr_gpu* g, const phr_scene_desc* scene, uint32_t leaf, phr_tracer** out);
/* Tracer::with_chunks (src/trace/build.rs:136-200): from_scene under phiray's
 * per-binding byte limit and binding count (2, or 4 when chunks > 2);
 * PHR_ERR_SIZE (the reference's None) when the scene does not fit them. The
 * kernels here index whole arrays: the split is bookkeeping for callers. */
int phr_tracer_with_chunks(phr_gpu* g, const phr_scene_desc* scene, uint32_t leaf, uint64_t limit, uint32_t chunks,
                           phr_tracer** out);
/* Tracer::split / Tracer::chunks (src/trace/mod.rs:118-127): items a binding
 * for nodes, faces, vertices, and the atlas offset (0: no atlas); bindings an
 * array. */
int phr_tracer_split(phr_tracer* t, uint32_t out[4]);
uint32_t phr_tracer_chunks(phr_tracer* t);
/* Tracer::reinstance (src/trace/pose.rs:102-125): new matrices (16 floats each,
 * m[col][row]); the top tree, instance table and light table follow, the
 * image starts over. *ok = 0 when the count differs or the tracer was not
 * built from a scene (build it again). */
int phr_tracer_reinstance(phr_tracer* t, const float* worlds, uint32_t n, int* ok);
/* Tracer::repose (src/trace/pose.rs:70-100): the same faces over vertices
 * moved (32 B each, object space): every mesh tree refitted (one batched
 * launch sequence), the top tree rebuilt, the image started over. *ok = 0
 * when the vertex count differs. The _on form takes a device array. */
int phr_tracer_repose(phr_tracer* t, const void* verts, uint32_t nverts, int* ok);
int phr_tracer_repose_on(phr_tracer* t, phr_array* verts, int* ok);
/* Tracer::retree (src/trace/pose.rs:130-155): every mesh tree built again over
 * the pose held, faces put in their new order; the image keeps its samples. */
int phr_tracer_retree(phr_tracer* t, int* ok);
/* The forest as the kernels read it: the node array (top tree first) and the
 * instance words (Forest::inst_words, src/trace/forest.rs:168-181). */
phr_array* phr_tracer_nodes(phr_tracer* t);
phr_array* phr_tracer_instances(phr_tracer* t);
uint32_t phr_tracer_instance_count(const phr_tracer* t);

/* Pixels Adaptive Sampling has not settled (src/trace/sample.rs `pixels
 * left`, Camera::adapt = {threshold, min samples, step (a power of two), on});
 * PHR_ERR_UNSUPPORTED when the last camera had it off. */
int phr_tracer_open_pixels(phr_tracer* t, uint32_t* count);
/* The path kernel variant (no counterpart in phiray): register allocation
 * fitted to 2 (default), 3, 4 or 5 waves a SIMD. */
int phr_tracer_set_path_waves(phr_tracer* t, uint32_t waves);
uint32_t phr_tracer_path_waves(const phr_tracer* t);
/* The path-regenerating builds (no counterpart in phiray): 1 lets a lane whose
 * sample ends start its pixel's next sample at once instead of waiting for
 * the wave's longest path; 2 is the wavefront build: each turn walks every
 * pixel's ray in a persistent walker (a lane whose walk ends takes the next
 * queued ray), shades in a kernel of its own, and appends the next rays and
 * the shadow rays to queues by wave ballots; 3 is the split build: each lane
 * keeps one resumable walk (closest or shadow ray) and a wave walks only
 * until three quarters of its working lanes have finished, those shade and begin their
 * next walk while the others walk on. All give the same sums, bit for bit.
 * They serve a camera with lamps or a lit world and no Adaptive Sampling, and
 * 1 and 2 one tree only (3 serves a forest too); other frames take the
 * nested build, and a sliced sample or a recorded frame takes 1 for 2. The
 * autotune tries every build the scene and camera allow. */
int phr_tracer_set_path_regen(phr_tracer* t, uint32_t on);
uint32_t phr_tracer_path_regen(const phr_tracer* t);
/* No counterpart in the reference: the split build's shading point. A wave
 * walks until `eighths` (1 to 8) eighths of its working lanes wait to shade
 * (default 6); the autotune picks 4 or 6. The sums do not depend on it. */
int phr_tracer_set_split_shade(phr_tracer* t, uint32_t eighths);
uint32_t phr_tracer_split_shade(const phr_tracer* t);
/* No counterpart in the reference: the mesh-tree stack entries a thread of
 * the path launches holds (24 or 32; the autotune picks the 24-entry stack
 * at 3 waves too when every mesh tree of a forest is at most 24 deep). */
int phr_tracer_set_path_stack(phr_tracer* t, uint32_t entries);
uint32_t phr_tracer_path_stack(const phr_tracer* t);
/* Autotune (no counterpart in phiray): time each variant on `spp` samples of
 * w x h under `cam` (best of `reps`), keep the fastest; the image starts over. */
int phr_tracer_autotune(phr_tracer* t, const void* cam, uint32_t w, uint32_t h, uint32_t spp, uint32_t reps,
                        uint32_t* chosen, double* best_s);

/* ---- a batch of scenes (no counterpart in phiray, which holds one Scene a
 * Tracer, src/trace/mod.rs:50-245): many scenes of uneven size built by one
 * segmented LBVH build (the segments of phr_bvh_build_batch) and traced by
 * one launch. Scene k renders exactly as a Tracer over its triangles alone
 * (phr_tracer_new) under the same camera. ----------------------------------- */
typedef struct phr_batch phr_batch;
/* Scene k is tris[seg_lo[k], seg_lo[k + 1]) (128-B Tri records as
 * phr_tracer_new takes them); seg_lo[0] = 0 and no scene is empty. */
int phr_batch_new(phr_gpu* g, const void* tris, const uint32_t* seg_lo, uint32_t scenes, uint32_t leaf, phr_batch** out);
void phr_batch_free(phr_batch* b);
uint32_t phr_batch_scenes(const phr_batch* b);
uint32_t phr_batch_len(const phr_batch* b);
phr_bvh* phr_batch_bvh(phr_batch* b);
float phr_batch_extent(const phr_batch* b, uint32_t k);
uint32_t phr_batch_glow(const phr_batch* b, uint32_t k);
/* Boxes, the segmented build and the slot order again, on the device. */
int phr_batch_rebuild(phr_batch* b, double* gpu_s);
/* Every scene, spp samples of a fresh w x h image each, in one launch. */
int phr_batch_render(phr_batch* b, const void* cam, uint32_t w, uint32_t h, uint32_t spp, double* gpu_s);
/* The render kernel: waves a SIMD (2 to 5) and block order (grouped 1: the
 * blocks of one XCD cover whole scenes; needs a multiple of 8 scenes). */
int phr_batch_set_path(phr_batch* b, uint32_t waves, uint32_t grouped);
int phr_batch_path(const phr_batch* b, uint32_t* waves, uint32_t* grouped);
/* The render kernel's samples: 0 the nested build, 3 the split build (walks
 * that span turns; used where the camera has lamps or a lit world and no
 * Adaptive Sampling, else the nested build runs). No reference counterpart:
 * phiray renders one scene a call. */
int phr_batch_set_path_build(phr_batch* b, uint32_t build);
uint32_t phr_batch_path_build(const phr_batch* b);
/* The split build's shading point, as phr_tracer_set_split_shade (default 4). */
int phr_batch_set_split_shade(phr_batch* b, uint32_t eighths);
uint32_t phr_batch_split_shade(const phr_batch* b);
/* Time every (build, shading point, waves, order) on this render and keep the fastest. */
int phr_batch_autotune(phr_batch* b, const void* cam, uint32_t w, uint32_t h, uint32_t spp, uint32_t reps, uint32_t* waves,
                       uint32_t* grouped, double* best_s);
/* The whole pass (rebuild 1: build, then render) as one hipGraph; kernels
 * only. Replays are refused (PHR_ERR_ARG) once the images moved. */
int phr_batch_record(phr_batch* b, const void* cam, uint32_t w, uint32_t h, uint32_t spp, int rebuild);
int phr_batch_replay(phr_batch* b, double* gpu_s);
/* Scene k's image words [from, from + count): its own PX_LEN * w * h + 1. */
int phr_batch_px(phr_batch* b, uint32_t k, uint64_t from, uint64_t count, uint32_t* out);
/* Since creation: closest rays, shadow rays, inner-node steps, triangle tests. */
int phr_batch_counters(phr_batch* b, uint64_t* out4);

/* ---- the learned denoiser (no counterpart in phiray: it stands beside
 * kernel::Denoise, src/kernel/denoise.rs:26-64) ------------------------------
 * A kernel-predicting convolutional network: a 4-level encoder/decoder of 3x3
 * convs (32/48/64/96 channels) predicting a softmax-normalised 5x5 filter per
 * pixel over the albedo-demodulated radiance. BF16 MFMA with FP32
 * accumulation; FP32 master weights; Adam (beta 0.9/0.999, eps 1e-8).
 * Frames are 10 floats a pixel (radiance 3, albedo 3, normal 3, depth 1: the
 * tracer's readbacks, src/trace/read.rs:241-307), images row-major, a batch
 * image after image; radiance arrays 3 floats a pixel. The loss is
 * mean((y - t)^2 / (t^2 + 0.01)) over pixels and channels. */
/* Layer table: 4 words a layer {cin, cout, level, relu}; returns the count. */
uint32_t phr_kpcn_layers(uint32_t* out);
/* Weights: each layer's W[9][cout][cin] (tap = 3 (dy + 1) + dx + 1) then
 * b[cout], layer after layer, FP32. */
uint64_t phr_kpcn_param_count(void);
int phr_kpcn_new(phr_gpu* g, phr_kpcn** out);
void phr_kpcn_free(phr_kpcn* k);
/* Sets the master weights and clears the optimiser state. */
int phr_kpcn_set_params(phr_kpcn* k, const float* host, uint64_t count);
int phr_kpcn_get_params(phr_kpcn* k, float* host, uint64_t count);
/* The gradient left by the last phr_kpcn_backward / phr_kpcn_train_step. */
int phr_kpcn_get_grads(phr_kpcn* k, float* host, uint64_t count);
/* Denoised radiance of b frames of w x h. */
int phr_kpcn_infer(phr_kpcn* k, phr_array* frames, uint32_t b, uint32_t w, uint32_t h, phr_array* out);
/* The loss against targets (and the denoised radiance into `out` if given). */
int phr_kpcn_loss(phr_kpcn* k, phr_array* frames, phr_array* targets, uint32_t b, uint32_t w, uint32_t h,
                  phr_array* out, float* loss);
/* Forward, loss and backward: the gradient w.r.t. every weight. `loss` may be
 * null (then nothing waits for the device). */
int phr_kpcn_backward(phr_kpcn* k, phr_array* frames, phr_array* targets, uint32_t b, uint32_t w, uint32_t h,
                      float* loss);
/* One Adam step at learning rate lr over the last gradient. */
int phr_kpcn_adam(phr_kpcn* k, float lr);
/* phr_kpcn_backward then phr_kpcn_adam. */
int phr_kpcn_train_step(phr_kpcn* k, phr_array* frames, phr_array* targets, uint32_t b, uint32_t w, uint32_t h,
                        float lr, float* loss);
/* Autotune (no counterpart): each layer's forward conv tile, timed on b frames
 * of w x h; chosen gets the rows a wave owns per layer (phr_kpcn_layers order). */
int phr_kpcn_autotune(phr_kpcn* k, phr_array* frames, uint32_t b, uint32_t w, uint32_t h, uint32_t reps, uint8_t* chosen);
/* No counterpart in the reference: the learned frame's held-out check. For
 * two independent halves A and B of a frame (FRAME-layout arrays, n pixels)
 * and `denoised` (3n floats) made from A alone: *err estimates the relMSE of
 * `denoised` against the converged image, mean(((D - B)^2 - (A - B)^2 / 2) /
 * (D^2 + 0.01)); *noise (may be null) is B's own term, mean((A - B)^2 / 2 /
 * (D^2 + 0.01)). Blocking. */
int phr_kpcn_held_out(phr_gpu* g, phr_array* denoised, phr_array* frame_a, phr_array* frame_b, uint64_t n, double* err,
                      double* noise);
/* No counterpart in the reference: the spread of two denoised images a and b
 * (3n floats each) made from independent halves of one frame, mean((a - b)^2 /
 * (a^2 + 0.01)): twice the noise-driven part of either image's relMSE. Blocking. */
int phr_kpcn_spread(phr_gpu* g, phr_array* a, phr_array* b, uint64_t n, double* out);
/* No counterpart in the reference: the frame of both halves (equal sample
 * counts) as one: each channel's mean, depth over the halves that met
 * something; channels 10 (a frame) or 3 (the radiance alone). Enqueued. */
int phr_kpcn_frame_mean(phr_gpu* g, phr_array* frame_a, phr_array* frame_b, uint64_t n, uint32_t channels, phr_array* out);
/* One training step over b images of w x h gathered from resident data sets:
 * frame image fidx[j] of `frames` against target image tidx[j] of `targets`.
 * PHR_ERR_ARG for an index past either set. */
int phr_kpcn_train_indexed(phr_kpcn* k, phr_array* frames, const uint32_t* fidx, phr_array* targets,
                           const uint32_t* tidx, uint32_t b, uint32_t w, uint32_t h, float lr, float* loss);
/* One frame as a hipGraph (no counterpart in phiray): spp samples of a fresh
 * image under cam at w x h, then denoise 0 none, 1 a-trous (keep = false),
 * 2 the learned net k; kernels only. The call renders it once eagerly first.
 * Replays are refused (PHR_ERR_ARG) once the scene, image or network moved. */
int phr_tracer_record_frame(phr_tracer* t, const void* cam, uint32_t w, uint32_t h, uint32_t spp, int denoise, phr_kpcn* k);
int phr_tracer_replay_frame(phr_tracer* t, phr_kpcn* k);
/* No counterpart in the reference (the frame of src/trace/pose.rs:77-155 and
 * sample.rs as one recorded graph): a forest's new pose — vertices from the
 * device array `verts` when not null (refit of the mesh trees), n matrices
 * (the top tree, instance and emitter tables) — then the camera's stamp, the
 * path launch and the denoiser (0 none, 1 a-trous, 2 learned with k), all in
 * one hipGraph. The emitter table must be one the device builds (1 to 4,096
 * candidates). Replay with new matrices (and whatever `verts` then holds);
 * the next sample call starts the image over. */
int phr_tracer_record_pose_frame(phr_tracer* t, phr_array* verts, const float* worlds, uint32_t n, const void* cam,
                                 uint32_t w, uint32_t h, uint32_t spp, int denoise, phr_kpcn* k);
int phr_tracer_replay_pose_frame(phr_tracer* t, const float* worlds, uint32_t n, phr_kpcn* k);
/* The tracer's frame (10 floats a pixel) at float offset `at` of `out`. */
int phr_tracer_frame(phr_tracer* t, phr_array* out, uint64_t at);
/* The learned denoiser over what has been traced: into `out` (w * h * 3
 * floats) when given, else in place of the linear sums (Tracer::denoise's
 * keep = false, src/trace/read.rs:85-119). */
int phr_tracer_denoise_learned(phr_tracer* t, phr_kpcn* k, phr_array* out);

/* ---- rigid bodies: World (src/sim/world.rs, build.rs, state.rs) -------------
 * XPBD with substeps on the host, restated from phiray's CPU solver; the
 * broadphase (broad.rs: Bvh::build over the swept boxes and the overlap walk)
 * runs on the device. Shapes: box, sphere, capsule (not a box-capsule pair);
 * the GJK/EPA kinds are PHR_ERR_UNSUPPORTED. */
/* Shape (state.rs, 32 B): kind 0 box, 1 sphere, 2 capsule; half extents (box)
 * or half length on z (capsule); radius. */
typedef struct { uint32_t kind, data, count; float radius; float half[3]; float _a; } phr_shape;
/* BodyDef (build.rs:14-61): kind 0 active, 1 passive, 2 kinematic; q is xyzw.
 * Defaults as phiray's: mass 1, friction 0.5, restitution 0, margin 0,
 * lin_damp 0.04, ang_damp 0.1, groups 1, scale 1. */
typedef struct {
    uint32_t kind, shape; float mass;
    float x[3]; float q[4]; float v[3]; float w[3];
    float scale[3]; float com[3];
    float friction, restitution, margin, lin_damp, ang_damp;
    float sleep[2]; uint32_t has_sleep, asleep, groups, obj;
} phr_body_def;
int phr_world_new(phr_gpu* g, phr_world** out);
void phr_world_free(phr_world* w);
/* World::substeps, iterations, time_scale, max_recover, sleep_after, gravity
 * (null: unchanged). */
int phr_world_configure(phr_world* w, uint32_t substeps, uint32_t iterations, float time_scale, float max_recover,
                        float sleep_after, const float gravity[3]);
int phr_world_add_shape(phr_world* w, const phr_shape* s, uint32_t* id);
int phr_world_add_body(phr_world* w, const phr_body_def* d, uint32_t* id);
/* World::step (world.rs:104-129): one frame of frame_dt seconds. PHR_ERR_SIZE
 * when the frame's contact pairs need more than 1,024 colours, or when the
 * solve's dataflow order (up to 1,280 bodies) does not finish within its bound
 * of waits; the bodies then keep the state of the frame's start. */
int phr_world_step(phr_world* w, float frame_dt);
uint32_t phr_world_body_count(const phr_world* w);
uint32_t phr_world_contact_count(const phr_world* w);
/* The 64-byte Body records (state.rs). */
int phr_world_bodies(const phr_world* w, void* out);
/* World::poses (world.rs:225-228): 16 floats a body, m[col][row], for
 * phr_tracer_reinstance. */
int phr_world_poses(const phr_world* w, float* out);
/* World::energy (world.rs:231-243). */
float phr_world_energy(const phr_world* w);
/* Pairs the last broadphase found (sorted, deduplicated): *n pairs, up to cap
 * written as (i, j) with i < j. */
int phr_world_pair_list(const phr_world* w, uint32_t* out, uint32_t cap, uint32_t* n);
/* No counterpart in the reference: lanes a pair in the solve's projections,
 * 4 (default: a point a lane) or 1 (a pair a lane); both give the same bits. */
int phr_world_set_lanes(phr_world* w, uint32_t lanes);
uint32_t phr_world_lanes(const phr_world* w);

#ifdef __cplusplus
}
#endif
#endif

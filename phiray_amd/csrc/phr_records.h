// The host records that cross the C ABI: the tracer's own triangle and camera
// (the fields the Rendered path reads), and phiray's full `Tri` and `Camera`
// as the reference lays them out, which phr_tri_from_phiray and
// phr_camera_from_phiray turn into the former.
#pragma once
#include <stdint.h>

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
    uint32_t w, h, rr_from, stream;  // stream: added to the sample index in the seed (0: the reference's seeds)
    float adapt[4];  // Adaptive Sampling: threshold, min samples, step (a power of two), on (camera.rs `adapt`)
    float view[4];   // colour management: transform (0 Standard), exposure in stops, gamma (0: 1), look (camera.rs:53-56)
};
static_assert(sizeof(PhrCamera) == 208, "camera layout");

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

// ---- phiray's records, byte for byte --------------------------------------
// `NO_TEX` and `NO_MAT` (src/trace/kernel.rs:70, src/trace/mat.rs:49).
constexpr uint32_t PHIRAY_NO_TEX = 0xffffffffu;
constexpr uint32_t PHIRAY_NO_MAT = 0xffffffffu;
constexpr uint32_t PHIRAY_MAX_LIGHT = 64;  // src/trace/camera.rs:332

// A procedural texture slot (src/trace/proc.rs:41-...; 44 words,
// proc.rs:254); `kind` 0 is an empty slot.
struct PhirayProc { uint32_t kind; uint32_t rest[43]; };
static_assert(sizeof(PhirayProc) == 176, "Proc is 44 words");

// `Tri` (src/trace/kernel.rs:11-65).
struct PhirayTri {
    float a[3]; float rough;
    float na[3]; float trans;
    float b[3]; float metal;
    float nb[3]; float ior;
    float c[3]; float _c;
    float nc[3]; float _nc;
    float color[4];
    float us[3]; uint32_t tex;
    float vs[3]; uint32_t nmap;
    float emit[3]; float bump;
    uint32_t maps[3]; float sss;
    float sss_radius[3]; float sheen;
    float sheen_rough; float sheen_tint[3];
    float coat, coat_rough, aniso, aniso_rot;
    uint32_t second; float mix;
    PhirayProc color_proc, bump_proc, color_proc2;
};
static_assert(sizeof(PhirayTri) == 760, "phiray Tri layout");

// `Light` (src/trace/light.rs:14-43).
struct PhirayLight {
    float pos[3]; uint32_t kind;
    float color[3]; float size;
    float dir[3]; float cone;
    float right[3]; float size_y;
    uint32_t shape; float blend; float spot_scale[2];
};
static_assert(sizeof(PhirayLight) == 80, "phiray Light layout");

// `Camera` (src/trace/camera.rs:12-228).
struct PhirayCamera {
    uint32_t n, groups_x, w, h;
    float right[3]; uint32_t persp;
    float up[3]; float scale;
    float back[3]; float units;
    float pivot[3]; float aspect;
    float sun[3]; float dist;
    float sky[4];
    float pan[2]; float ao; float eps;
    uint32_t frame; float grid; uint32_t count; uint32_t glow;
    float grid_color[4];
    float env[4];
    uint32_t split[4];
    float view[4];
    uint32_t envmap; float env_rot; uint32_t bounces; uint32_t env_cdf;
    float aperture, focus; uint32_t blades; float blade_spin;
    float aperture_ratio; uint32_t env_hidden; float see_through; uint32_t grid_hide;
    uint32_t tile[4], film[4], preview[4], slice[4], insts[4], vis[4];
    float taa[4], fx[4], outline[4];
    uint32_t shade[4];
    float single[4], cavity[4];
    uint32_t ssao[4];
    float shadow[4], shadow_fade[4];
    uint32_t matcap[4];
    float dof[4], outline_fx[4], outline_pick[4], outline_active[4], outline_move[4];
    float adapt[4], limit[4];
    uint32_t depth[4];
    float integrator[4];
    PhirayLight lights[PHIRAY_MAX_LIGHT];
};
static_assert(sizeof(PhirayCamera) == 5776, "phiray Camera layout");

// phiray's own `Tri` and `Camera` bytes into the tracer's records.
//
// The tracer's entry points take a 128-byte triangle and a 208-byte camera:
// the fields the restated Rendered path reads. phiray's callers hold the full
// records (`Tri`, 760 bytes, src/trace/kernel.rs:11-65; `Camera`, 5,776
// bytes, src/trace/camera.rs:12-228). These two calls take those bytes as
// they are and copy the fields across, and refuse (PHR_ERR_UNSUPPORTED) a
// record that asks for something outside the restated subset, so that a
// converted scene renders what phiray would or nothing at all. Host code only.
#include "phr_records.h"
#include "phr_runtime.h"

#include <cstring>
#include <string>

namespace {

bool any_nonzero(const float* v, int n) {
    for (int k = 0; k < n; ++k)
        if (v[k] != 0.0f) return true;
    return false;
}

std::string at(const char* what, uint64_t i) { return std::string(what) + " (record " + std::to_string(i) + ")"; }

}  // namespace

extern "C" {

int phr_tri_from_phiray(const void* src, uint64_t n, void* dst) {
    PHR_CHECK(n == 0 || (src && dst), PHR_ERR_ARG, "phr_tri_from_phiray: null argument");
    const PhirayTri* in = static_cast<const PhirayTri*>(src);
    PhrTri* out = static_cast<PhrTri*>(dst);
    for (uint64_t i = 0; i < n; ++i) {
        PhirayTri t;
        std::memcpy(&t, in + i, sizeof t);
        // Textures, normal maps and texture-driven inputs (kernel.rs:27-43).
        PHR_CHECK(t.tex == PHIRAY_NO_TEX && t.nmap == PHIRAY_NO_TEX, PHR_ERR_UNSUPPORTED,
                  at("phr_tri_from_phiray: textures and normal maps are not supported", i));
        PHR_CHECK(t.maps[0] == PHIRAY_NO_TEX && t.maps[1] == PHIRAY_NO_TEX && t.maps[2] == PHIRAY_NO_TEX,
                  PHR_ERR_UNSUPPORTED, at("phr_tri_from_phiray: roughness/metallic/emission maps are not supported", i));
        // Subsurface, sheen, coat, a second lobe and procedurals (kernel.rs:44-64).
        PHR_CHECK(t.sss == 0.0f && t.sheen == 0.0f && t.coat == 0.0f, PHR_ERR_UNSUPPORTED,
                  at("phr_tri_from_phiray: subsurface, sheen and coat are not supported", i));
        PHR_CHECK(t.second == PHIRAY_NO_MAT, PHR_ERR_UNSUPPORTED,
                  at("phr_tri_from_phiray: mixed materials are not supported", i));
        PHR_CHECK(t.color_proc.kind == 0 && t.bump_proc.kind == 0 && t.color_proc2.kind == 0, PHR_ERR_UNSUPPORTED,
                  at("phr_tri_from_phiray: procedural textures are not supported", i));
        PhrTri o;
        std::memset(&o, 0, sizeof o);
        for (int k = 0; k < 3; ++k) {
            o.a[k] = t.a[k]; o.b[k] = t.b[k]; o.c[k] = t.c[k];
            o.na[k] = t.na[k]; o.nb[k] = t.nb[k]; o.nc[k] = t.nc[k];
            o.emit[k] = t.emit[k];
        }
        for (int k = 0; k < 4; ++k) o.color[k] = t.color[k];
        o.rough = t.rough; o.trans = t.trans; o.metal = t.metal; o.ior = t.ior; o.bump = t.bump;
        std::memcpy(out + i, &o, sizeof o);
    }
    return PHR_OK;
}

int phr_camera_from_phiray(const void* src, void* dst) {
    PHR_CHECK(src && dst, PHR_ERR_ARG, "phr_camera_from_phiray: null argument");
    PhirayCamera c;
    std::memcpy(&c, src, sizeof c);
    // What the restated Rendered path does not draw (shade.wgsl:118-160,
    // util.wgsl:168-222): the theme's sky and grid overlay, a transparent or
    // forced-world film, an environment image, a lens, a tile of a larger
    // frame, TAA jitter, and light clamping.
    PHR_CHECK(!any_nonzero(c.sky, 4), PHR_ERR_UNSUPPORTED, "phr_camera_from_phiray: the theme's sky must be black");
    PHR_CHECK(!(c.grid > 0.0f), PHR_ERR_UNSUPPORTED, "phr_camera_from_phiray: the grid overlay is not supported");
    PHR_CHECK(c.film[0] == 0u && c.film[1] == 0u, PHR_ERR_UNSUPPORTED,
              "phr_camera_from_phiray: a transparent film or a shown world is not supported");
    PHR_CHECK(c.envmap == 0u && c.env_cdf == 0u, PHR_ERR_UNSUPPORTED,
              "phr_camera_from_phiray: environment images are not supported (a flat world colour is)");
    PHR_CHECK(!(c.aperture > 0.0f), PHR_ERR_UNSUPPORTED, "phr_camera_from_phiray: depth of field is not supported");
    PHR_CHECK(c.tile[2] == 0u, PHR_ERR_UNSUPPORTED, "phr_camera_from_phiray: tiles of a larger frame are not supported");
    PHR_CHECK(c.taa[0] == 0.0f && c.taa[1] == 0.0f, PHR_ERR_UNSUPPORTED,
              "phr_camera_from_phiray: TAA jitter is not supported");
    PHR_CHECK(!(c.limit[0] > 0.0f) && !(c.limit[1] > 0.0f), PHR_ERR_UNSUPPORTED,
              "phr_camera_from_phiray: light clamping is not supported");
    // Lamps: the restated path carries unlit slots only (`count` of them).
    const uint32_t lamps = c.count < PHIRAY_MAX_LIGHT ? c.count : PHIRAY_MAX_LIGHT;
    for (uint32_t k = 0; k < lamps; ++k)
        PHR_CHECK(!any_nonzero(c.lights[k].color, 3), PHR_ERR_UNSUPPORTED,
                  at("phr_camera_from_phiray: lit lamps are not supported", k));
    PhrCamera o;
    std::memset(&o, 0, sizeof o);
    for (int k = 0; k < 3; ++k) {
        o.right[k] = c.right[k]; o.up[k] = c.up[k]; o.back[k] = c.back[k];
        o.pivot[k] = c.pivot[k]; o.sun[k] = c.sun[k];
    }
    o.persp = c.persp; o.scale = c.scale; o.units = c.units; o.aspect = c.aspect; o.dist = c.dist;
    o.pan[0] = c.pan[0]; o.pan[1] = c.pan[1];
    o.ao = c.ao; o.eps = c.eps;
    for (int k = 0; k < 4; ++k) {
        o.env[k] = c.env[k]; o.depth[k] = c.depth[k]; o.integrator[k] = c.integrator[k];
        o.adapt[k] = c.adapt[k]; o.view[k] = c.view[k];
    }
    o.frame = c.frame; o.count = c.count; o.glow = c.glow; o.bounces = c.bounces;
    o.w = c.w; o.h = c.h;
    o.rr_from = 0;  // the reference's roulette: every bounce but the last (integrate.wgsl:210)
    o.stream = 0;   // the reference's seeds (shade.wgsl:68)
    std::memcpy(dst, &o, sizeof o);
    return PHR_OK;
}

}  // extern "C"

"""The tracer on the device against the oracle's tracer (GPU).

* Raycast: the 65,536-triangle workload with 256 x 256 pinhole rays — the tree,
  the slot order, every hit's slot, t, u and v equal the oracle's walk bit for
  bit (hits are reported per slot; `order` maps them to input triangles).
* Any-hit rays with finite tmax, the same.
* Path tracing: a small spheres scene, several samples, every image-buffer
  plane the integrator writes equal to the oracle's (1e-5
  relative; identical operation order gives equal bits here), and the ray
  counts equal.
"""
import numpy as np
import pytest

import phiray_amd as P
from phiray_amd import scenes as S

pytestmark = pytest.mark.gpu


def test_raycast_workload_matches_the_walk(gpu, orc):
    print("[trace] raycast 65536 tris, 256x256 rays", flush=True)
    corners, _ = S.lbvh_scene(65536, 0)
    tris = S.tris_from_corners(corners)
    sc = orc.Scene(tris, 4, P.NODE, 8)
    tr = P.Tracer.new(gpu, tris, 4)
    assert (tr.order() == sc.order).all()
    g = tr.bvh()
    assert g.depth == sc.depth
    r = np.array(P.Bvh(sc.nodes, sc.order, sc.depth).reachable())
    assert (g.nodes[r].view(np.uint32) == sc.nodes[r].view(np.uint32)).all()
    rays = S.pinhole_rays(256, 256)
    got = tr.raycast(rays)
    want = orc.raycast(sc.nodes, sc.tri9, rays, P.HIT, nthreads=8)
    assert (got["tri"] == want["tri"]).all()
    assert (got.view(np.uint32).reshape(-1, 5) == want.view(np.uint32).reshape(-1, 5)).all()
    hit = got["tri"] != P.NONE
    assert hit.mean() > 0.2
    ids = tr.order()[got["tri"][hit]]
    assert ids.max() < 65536


def test_any_hit_rays(gpu, orc):
    print("[trace] any-hit rays", flush=True)
    corners, _ = S.lbvh_scene(20000, 4)
    tris = S.tris_from_corners(corners)
    sc = orc.Scene(tris, 4, P.NODE)
    tr = P.Tracer.new(gpu, tris, 4)
    rays = S.pinhole_rays(128, 128)
    rays["tmax"] = np.float32(1.6)
    got = tr.raycast(rays, any_hit=True)
    want = orc.raycast(sc.nodes, sc.tri9, rays, P.HIT, any_hit=True, nthreads=8)
    assert (got.view(np.uint32).reshape(-1, 5) == want.view(np.uint32).reshape(-1, 5)).all()


def compare_planes(got, want, w, h):
    pg, pw = P.planes(got, w, h), P.planes(want, w, h)
    for k in ["radiance", "albedo", "normal", "depth", "cov", "film", "count", "missed", "outline"]:
        a, b = pg[k], pw[k]
        if a.dtype == np.float32:
            bad = ~np.isclose(a, b, rtol=1e-5, atol=0)
            assert not bad.any(), f"{k}: {bad.sum()} words beyond 1e-5 relative, e.g. {a[bad][:4]} vs {b[bad][:4]}"
        else:
            assert (a == b).all(), k


@pytest.mark.parametrize("spp", [1, 3])
def test_path_tracer_matches_the_twin(gpu, orc, spp):
    print(f"[trace] path spheres spp={spp}", flush=True)
    tris = S.spheres_scene(spheres=8, level=2, seed=3)
    w, h = 64, 48
    cam = S.path_camera(w, h, rr_from=3)
    sc = orc.Scene(tris, 4, P.NODE)
    tr = P.Tracer.new(gpu, tris, 4)
    assert tr.glow == sc.glow and abs(tr.extent - sc.extent) == 0.0
    assert (tr.lights() == sc.lights).all()
    px = np.zeros(P.PX_LEN * w * h + 1, dtype=np.uint32)
    rays = [0, 0]
    for f in range(spp):
        r = sc.sample(cam, w, h, px, f, nthreads=8)
        rays = [rays[0] + r[0], rays[1] + r[1]]
    frame, _ = tr.sample(cam, w, h, spp)
    assert frame == spp - 1
    compare_planes(tr.px(), px, w, h)
    assert list(tr.rays()) == rays


def test_samples_accumulate_across_calls(gpu, orc):
    print("[trace] path accumulate", flush=True)
    tris = S.spheres_scene(spheres=4, level=1, seed=5)
    w, h = 40, 30
    cam = S.path_camera(w, h)
    sc = orc.Scene(tris, 4, P.NODE)
    tr = P.Tracer.new(gpu, tris, 4)
    px = np.zeros(P.PX_LEN * w * h + 1, dtype=np.uint32)
    for f in range(4):
        sc.sample(cam, w, h, px, f, nthreads=8)
    tr.sample(cam, w, h, 1)
    tr.sample(cam, w, h, 2)
    frame, _ = tr.sample(cam, w, h, 1)
    assert frame == 3
    compare_planes(tr.px(), px, w, h)
    # A new view starts the image over.
    cam2 = cam.copy()
    cam2["dist"] = cam2["dist"] * np.float32(1.1)
    frame, _ = tr.sample(cam2, w, h, 1)
    assert frame == 0

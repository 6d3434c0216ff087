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
    for k in ["radiance", "albedo", "normal", "depth", "cov", "film", "count", "missed", "outline", "display", "pixel"]:
        a, b = pg[k], pw[k]
        if k == "pixel":  # 8-bit rounding of display values within 1e-5 of each other
            assert (np.abs(a.astype(np.int32) - b.astype(np.int32)) <= 1).all(), k
        elif a.dtype == np.float32:
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


def test_path_kernel_variants_agree(gpu, orc):
    """The 2-, 3- and 4-waves-a-SIMD builds of the path kernel differ only in
    register allocation: every plane equals the oracle's, for each."""
    print("[trace] path variants", flush=True)
    tris = S.spheres_scene(spheres=6, level=2, seed=9)
    w, h = 48, 32
    cam = S.path_camera(w, h, rr_from=3)
    sc = orc.Scene(tris, 4, P.NODE)
    px = np.zeros(P.PX_LEN * w * h + 1, dtype=np.uint32)
    for f in range(2):
        sc.sample(cam, w, h, px, f, nthreads=8)
    tr = P.Tracer.new(gpu, tris, 4)
    assert tr.path_waves == 2 and not tr.path_regen
    for regen in (False, True, P.Tracer.PATH_WAVEFRONT, P.Tracer.PATH_SPLIT):
        tr.set_path_regen(regen)
        assert tr.path_regen == int(regen)
        for waves in (2, 3, 4, 5):
            tr.set_path_waves(waves)
            fresh = cam.copy()
            fresh["stream"] = 0
            tr.sample(fresh, w, h, 2)
            compare_planes(tr.px(), px, w, h)
            tr.sample(S.path_camera(w + 1, h, rr_from=3), w + 1, h, 1)  # another view: the next call starts over
    tr.set_path_regen(P.Tracer.PATH_SPLIT)
    tr.set_path_waves(4)
    for eighths in (1, 2, 4, 6, 8):  # the split build's shading point: the same sums at every one
        tr.set_split_shade(eighths)
        assert tr.split_shade == eighths
        fresh = cam.copy()
        fresh["stream"] = 0
        tr.sample(fresh, w, h, 2)
        compare_planes(tr.px(), px, w, h)
        tr.sample(S.path_camera(w + 1, h, rr_from=3), w + 1, h, 1)
    with pytest.raises(P.PhrError):
        tr.set_split_shade(9)
    with pytest.raises(P.PhrError):
        tr.set_path_waves(6)
    chosen, best = tr.autotune(cam, w, h, 1, reps=1)
    assert chosen in (2, 3, 4, 5) and best > 0 and tr.path_waves == chosen


def soup(n, seed):
    corners, _ = S.lbvh_scene(n, seed)
    corners = (corners - np.float32(0.5)) * np.float32(3.0)  # a 3 m cube of triangles
    corners[..., 2] += np.float32(1.0)
    return S.tris_from_corners(corners, color=(0.7, 0.6, 0.5, 1.0))


@pytest.mark.parametrize("build", [1, 2, 3], ids=["regen", "wavefront", "split"])
@pytest.mark.parametrize("case", ["lamps", "world", "occlusion"])
def test_path_regeneration_matches_the_twin(gpu, orc, case, build):
    """The path-regenerating build (a lane that ends a sample starts its
    pixel's next one), the wavefront build (walks, shading and shadow walks
    as kernels of their own over queues of slots) and the split build (walks
    that span turns, a lane shading as soon as half its wave waits): every
    plane and the
    ray counts equal the oracle's, samples spread over three calls, for a
    scene lit by lamps, a soup under the world colour, and a soup in a black
    world (the ambient-occlusion bounce)."""
    print(f"[trace] path regeneration, {case}, build {build}", flush=True)
    w, h = 56, 40
    cam = S.path_camera(w, h, rr_from=3)
    if case == "lamps":
        tris = S.spheres_scene(spheres=8, level=2, seed=11)
    else:
        tris = soup(3000, 7)
        if case == "world":
            cam["env"] = (0.4, 0.5, 0.6, 0.0)
    sc = orc.Scene(tris, 4, P.NODE)
    px = np.zeros(P.PX_LEN * w * h + 1, dtype=np.uint32)
    rays = [0, 0]
    for f in range(5):
        r = sc.sample(cam, w, h, px, f, nthreads=8)
        rays = [rays[0] + r[0], rays[1] + r[1]]
    tr = P.Tracer.new(gpu, tris, 4)
    tr.set_path_regen(build)
    tr.set_path_waves(4)
    for spp in (1, 3, 1):
        tr.sample(cam, w, h, spp)
    compare_planes(tr.px(), px, w, h)
    assert list(tr.rays()) == rays
    if case == "occlusion":
        assert rays[1] > 0  # the occlusion rays ran (a black world adds no light)
    else:
        assert P.planes(px, w, h)["radiance"].max() > 0


def test_adaptive_sampling_matches_the_twin(gpu, orc):
    """Adaptive Sampling (adapt.wgsl, gather.wgsl): settled pixels stop, the
    half image, the flags before and after the dilation, the per-pixel counts
    and every integrator plane equal the oracle's after each call; the list
    holds exactly the open pixels."""
    print("[trace] adaptive sampling", flush=True)
    tris = S.spheres_scene(spheres=8, level=2, seed=3)
    w, h = 64, 48
    n = w * h
    cam = S.path_camera(w, h, rr_from=3)
    cam["adapt"] = (0.08, 3.0, 2.0, 1.0)  # threshold, min samples, step, on
    sc = orc.Scene(tris, 4, P.NODE)
    tr = P.Tracer.new(gpu, tris, 4)
    px = np.zeros(P.PX_LEN * n + 1, dtype=np.uint32)
    f = 0
    for spp in (5, 2, 9):
        for _ in range(spp):
            sc.sample(cam, w, h, px, f, nthreads=8)
            sc.adapt_tail(cam, w, h, px, f)
            f += 1
        frame, _ = tr.sample(cam, w, h, spp)
        assert frame == f - 1
        got = tr.px()
        compare_planes(got, px, w, h)
        assert np.array_equal(got[26 * n:29 * n], px[26 * n:29 * n]), "half image"
        assert np.array_equal(got[29 * n:31 * n], px[29 * n:31 * n]), "flags"
        open_ = np.flatnonzero(px[30 * n:31 * n] == 0)
        assert tr.open_pixels() == len(open_) == got[34 * n]
        assert np.array_equal(np.sort(got[32 * n:32 * n + len(open_)]), open_)
    settled = int((px[30 * n:31 * n] != 0).sum())
    assert 0 < settled < n, settled  # the case exercises both kinds
    counts = px[31 * n:32 * n]
    assert counts.min() < f and counts.max() == f
    with pytest.raises(P.PhrError):
        bad = cam.copy()
        bad["adapt"] = (0.08, 3.0, 3.0, 1.0)
        tr.sample(bad, w, h, 1)
    # Tracer::set_adaptive stamps the same words onto a camera without them.
    t2 = P.Tracer.new(gpu, tris, 4)
    t2.set_samples(40)
    assert t2.samples() == 40
    t2.set_adaptive((0.08, 3.0, 2.0, 1.0))
    plain = S.path_camera(w, h, rr_from=3)
    for spp in (5, 2, 9):
        t2.sample(plain, w, h, spp)
    g2 = t2.px()  # the open-pixel list (32n..33n) is in atomic order: its open entries compared sorted
    assert np.array_equal(g2[:32 * n], got[:32 * n]) and np.array_equal(g2[33 * n:], got[33 * n:])
    k = int(got[34 * n])  # past the open count the list holds older entries
    assert np.array_equal(np.sort(g2[32 * n:32 * n + k]), np.sort(got[32 * n:32 * n + k]))
    t2.free()


def test_sample_in_slices(gpu, orc):
    """Tracer::sample_in: slices give the same image as one launch; a `go`
    that says no abandons the call and the next starts the image over."""
    print("[trace] sample_in", flush=True)
    tris = S.spheres_scene(spheres=6, level=2, seed=4)
    w, h = 72, 40
    cam = S.path_camera(w, h, rr_from=3)
    sc = orc.Scene(tris, 4, P.NODE)
    px = np.zeros(P.PX_LEN * w * h + 1, dtype=np.uint32)
    for f in range(3):
        sc.sample(cam, w, h, px, f, nthreads=8)
    tr = P.Tracer.new(gpu, tris, 4)
    asked = []
    assert tr.sample_in(cam, w, h, 2, 3, go=lambda: asked.append(1) or True) == (1, False)
    assert len(asked) == 3
    assert tr.sample_in(cam, w, h, 1, 5) == (2, False)
    compare_planes(tr.px(), px, w, h)
    assert tr.sample_in(cam, w, h, 1, 4, go=lambda: False) is None
    frame, _ = tr.sample(cam, w, h, 1)
    assert frame == 0  # started over


def test_recorded_frame_replays_equal_eager(gpu):
    """A frame captured as a hipGraph (path + a-trous) gives, replay after
    replay with other work between, the image an eager frame gives."""
    print("[trace] recorded frame", flush=True)
    tris = S.spheres_scene(spheres=6, level=2, seed=8)
    w, h = 64, 40
    cam = S.path_camera(w, h, rr_from=3)
    tr = P.Tracer.new(gpu, tris, 4)
    tr.sample(cam, w, h, 3)
    tr.denoise(keep=False)
    want = tr.px()
    other = S.path_camera(w, h, rr_from=3, eye=(1.0, -5.0, 2.0))
    tr.record_frame(cam, w, h, 3, denoise=1)
    for _ in range(3):
        tr.sample(other, w, h, 2)  # eager work on the same image between replays
        tr.replay_frame()
        assert np.array_equal(tr.px()[: 25 * w * h], want[: 25 * w * h])
    frame, _ = tr.sample(cam, w, h, 1)  # the recorded camera: samples add on
    assert frame == 3
    tr.sample(cam, w + 8, h, 1)  # a new size moves the image: the graph is refused
    with pytest.raises(P.PhrError):
        tr.replay_frame()


def test_full_frame_bands_match_the_twin(gpu, orc):
    """The full path workload: the 1,310,724-triangle spheres scene at
    1920x1080 (2 samples of the 64 the bench takes). Three bands of rows (top,
    middle, bottom) traced by the oracle equal the device frame's planes, for
    the nested, regenerating, wavefront and split builds."""
    print("[trace] full frame bands, 1.31M tris, 1920x1080", flush=True)
    tris = S.spheres_scene(spheres=64, level=5, seed=1)
    assert len(tris) == 1_310_724
    w, h, spp = 1920, 1080, 2
    cam = S.path_camera(w, h, rr_from=3)
    tr = P.Tracer.new(gpu, tris, 4)
    tr.set_path_waves(4)
    frame, _ = tr.sample(cam, w, h, spp)
    assert frame == spp - 1
    got = P.planes(tr.px(), w, h)
    tr.set_path_regen(True)
    tr.sample(S.path_camera(64, 64), 64, 64, 1)  # another view: the next call starts over
    tr.sample(cam, w, h, spp)
    regen = P.planes(tr.px(), w, h)
    tr.set_path_regen(P.Tracer.PATH_WAVEFRONT)
    tr.sample(S.path_camera(64, 64), 64, 64, 1)
    tr.sample(cam, w, h, spp)
    wave = P.planes(tr.px(), w, h)
    tr.set_path_regen(P.Tracer.PATH_SPLIT)
    tr.sample(S.path_camera(64, 64), 64, 64, 1)
    tr.sample(cam, w, h, spp)
    split = P.planes(tr.px(), w, h)
    sc = orc.Scene(tris, 4, P.NODE, 8)
    px = np.zeros(P.PX_LEN * w * h + 1, dtype=np.uint32)
    bands = [(0, 6), (h // 2 - 3, 6), (h - 6, 6)]
    for f in range(spp):
        for row, rows in bands:
            sc.sample(cam, w, h, px, f, 8, row * w, rows * w)
    want = P.planes(px, w, h)
    lit = 0
    for row, rows in bands:
        sl = slice(row * w, (row + rows) * w)
        for k in ["radiance", "albedo", "normal", "depth", "cov", "film", "count", "missed", "outline", "display", "pixel"]:
            for build, planes in (("plain", got), ("regen", regen), ("wavefront", wave), ("split", split)):
                a, b = planes[k][sl], want[k][sl]
                if k == "pixel":
                    assert (np.abs(a.astype(np.int32) - b.astype(np.int32)) <= 1).all(), (build, row, k)
                elif a.dtype == np.float32:
                    bad = ~np.isclose(a, b, rtol=1e-5, atol=0)
                    assert not bad.any(), f"{build} rows {row}+{rows} {k}: {bad.sum()} words beyond 1e-5 relative"
                else:
                    assert (a == b).all(), (build, row, k)
        lit += int((want["radiance"][sl] > 0).any(axis=1).sum())
    assert lit > 1000  # the bands cross lit surfaces


@pytest.mark.parametrize("view", [(0.0, 1.5, 2.2, 0.0), (1.0, 0.0, 0.0, 0.0), (1.0, -0.5, 0.0, 1.0),
                                  (1.0, 0.0, 0.0, 6.0), (2.0, 0.5, 1.2, 0.0)],
                         ids=["standard", "filmic", "filmic-vhc", "filmic-vlc", "agx"])
def test_display_under_each_view_matches_the_twin(gpu, orc, view):
    """The display plane and the packed pixel under each view transform
    (Standard with exposure and gamma, Filmic and two looks, AgX); `render`
    and `pixels` read back the packed words, `display` the plane."""
    print(f"[trace] display view={view}", flush=True)
    tris = S.spheres_scene(spheres=5, level=1, seed=8)
    w, h = 40, 24
    cam = S.path_camera(w, h, rr_from=3)
    cam["view"] = view
    tr = P.Tracer.new(gpu, tris, 4)
    tr.sample(cam, w, h, 2)
    sc = orc.Scene(tris, 4, P.NODE)
    px = np.zeros(P.PX_LEN * w * h + 1, dtype=np.uint32)
    for f in range(2):
        sc.sample(cam, w, h, px, f, 8)
    got = tr.px()
    compare_planes(got, px, w, h)
    n = w * h
    assert (tr.pixels().reshape(-1) == got[:n].view(np.uint8)).all()
    assert (tr.display().reshape(-1) == got[4 * n:7 * n].view(np.float32)).all()
    img, _ = tr.render(cam, w, h)  # a third sample onto the same image
    assert img.shape == (h, w, 4) and (img.reshape(-1) == tr.px()[:n].view(np.uint8)).all()
    assert P.planes(tr.px(), w, h)["count"].min() == 3
    tr.free()


@pytest.mark.parametrize("kind", [0.0, 1.0, 2.0])
def test_showing_the_image_its_own_radiance_changes_nothing(gpu, kind):
    """tests/show.rs:47-82 under Standard, Filmic and AgX: `show` of the
    image's own radiance gives the radiance, the display and the bytes back;
    before any sample it shows nothing."""
    print(f"[trace] show round trip view {kind}", flush=True)
    tris = S.spheres_scene(spheres=4, level=1, seed=11)
    w, h = 32, 24
    tr = P.Tracer.new(gpu, tris, 4)
    assert tr.image_size() == (0, 0)
    assert not tr.show(np.zeros((w * h, 3), np.float32))
    cam = S.path_camera(w, h, rr_from=3)
    cam["view"] = (kind, 0.5, 1.8, 0.0)
    tr.sample(cam, w, h, 4)
    assert tr.image_size() == (w, h)
    was_px, was_display, was_linear = tr.pixels(), tr.display(), tr.radiance()
    assert (was_linear[:, 0] > 0).any()
    assert tr.show(was_linear)
    np.testing.assert_allclose(tr.radiance(), was_linear, rtol=1e-4, atol=1e-4)
    np.testing.assert_allclose(tr.display(), was_display, rtol=0, atol=1e-4)
    assert (tr.pixels() == was_px).all()


@pytest.mark.parametrize("kind", [0.0, 1.0, 2.0])
def test_a_change_to_the_linear_pixels_shows_through_the_transform(gpu, kind):
    """tests/show.rs:87-118 under Standard, Filmic and AgX: one stop down in
    the linear pixels dims every lit pixel's display by less than half."""
    print(f"[trace] show darker view {kind}", flush=True)
    tris = S.spheres_scene(spheres=4, level=1, seed=11)
    w, h = 32, 24
    tr = P.Tracer.new(gpu, tris, 4)
    cam = S.path_camera(w, h, rr_from=3)
    cam["view"] = (kind, 0.0, 0.0, 0.0)
    tr.sample(cam, w, h, 4)
    before, linear = tr.display(), tr.radiance()
    assert tr.show(linear * 0.5)
    after = tr.display()
    lit = np.where(linear[:, 0] > 0.05)[0]
    assert len(lit) > 0
    assert (after[lit, 0] < before[lit, 0]).all()
    assert after[lit[0], 0] / max(before[lit[0], 0], 1e-6) > 0.5
    assert tr.show(linear[: w])  # zero-filled past the first row
    assert (tr.radiance()[w:] == 0).all() and (tr.display()[w:, 0] <= before[w:, 0] + 1e-6).all()

"""A batch of scenes rendered in one launch (GPU).

Each scene of a batch renders exactly as a Tracer over its triangles alone
does (itself held to the oracle's integrator bit for bit in test_gpu_trace),
whatever kernel build and block order the batch launch uses; one scene is
also checked against the oracle directly. The world colour lights the scenes
that hold no emitter, so that path of the integrator is covered too.
"""
import numpy as np
import pytest

import phiray_amd as P
from phiray_amd import scenes as S

pytestmark = pytest.mark.gpu

W, H, SPP = 48, 32, 2


def batch_scenes():
    """Eight scenes of uneven size: lit spheres and bare triangle soups."""
    parts = []
    for k in range(8):
        if k % 3 == 2:
            corners, _ = S.lbvh_scene(300 + 900 * k, k)
            parts.append(S.tris_from_corners(corners, color=(0.6, 0.5, 0.4, 1.0)))
        else:
            parts.append(S.spheres_scene(spheres=2 + k, level=1 + (k % 2), seed=20 + k))
    lo = np.concatenate([[0], np.cumsum([len(p) for p in parts])]).astype(np.uint32)
    return parts, np.concatenate(parts), lo


def camera():
    cam = S.path_camera(W, H, rr_from=3)
    cam["env"] = (0.25, 0.3, 0.35, 0.0)
    return cam


def compare_planes(got, want, w, h):
    pg, pw = P.planes(got, w, h), P.planes(want, w, h)
    for k in ["radiance", "albedo", "normal", "depth", "cov", "film", "count", "missed", "outline", "display", "pixel"]:
        a, b = pg[k], pw[k]
        if k == "pixel":  # 8-bit rounding of display values within 1e-5 of each other
            assert (np.abs(a.astype(np.int32) - b.astype(np.int32)) <= 1).all(), k
        elif a.dtype == np.float32:
            bad = ~np.isclose(a, b, rtol=1e-5, atol=0)
            assert not bad.any(), f"{k}: {bad.sum()} words beyond 1e-5 relative"
        else:
            assert (a == b).all(), k


@pytest.fixture(scope="module")
def batch(gpu):
    parts, tris, lo = batch_scenes()
    b = P.Batch(gpu, tris, lo)
    cam = camera()
    want = []
    for k, p in enumerate(parts):
        tr = P.Tracer.new(gpu, p, 4)
        assert b.extent(k) == tr.extent and b.glow(k) == tr.glow
        t = tr.bvh()
        g = b.tree(k)
        assert g.depth == t.depth and (g.order == t.order).all()
        r = np.array(t.reachable())
        assert (g.nodes[r].view(np.uint32) == t.nodes[r].view(np.uint32)).all()
        tr.sample(cam, W, H, SPP)
        want.append(tr.px())
        tr.free()
    return parts, b, cam, want


def test_every_scene_renders_as_its_own_tracer(batch):
    print("[batch render] scenes against tracers", flush=True)
    parts, b, cam, want = batch
    assert b.scenes == 8 and len(b) == sum(len(p) for p in parts)
    for build in (P.Tracer.PATH_NESTED, P.Tracer.PATH_SPLIT):
        b.set_path_build(build)
        assert b.path_build == build
        for waves, grouped in [(2, 0), (2, 1), (3, 1), (4, 0)]:
            b.set_path(waves, grouped)
            c0 = b.counters()
            assert b.render(cam, W, H, SPP) > 0
            for k in range(b.scenes):
                assert np.array_equal(b.px(k), want[k]), (build, waves, grouped, k)
            assert b.counters()[0] > c0[0]
    b.set_path_build(P.Tracer.PATH_SPLIT)
    b.set_path(4, 1)
    for eighths in (2, 6, 8):  # the split build's shading point
        b.set_split_shade(eighths)
        assert b.split_shade == eighths
        b.render(cam, W, H, SPP)
        for k in range(b.scenes):
            assert np.array_equal(b.px(k), want[k]), ("shade", eighths, k)
    b.set_split_shade(4)
    b.set_path_build(P.Tracer.PATH_NESTED)
    with pytest.raises(P.PhrError):
        b.set_path_build(1)


def test_a_scene_matches_the_oracle(batch, orc):
    print("[batch render] scene against the oracle", flush=True)
    parts, b, cam, _ = batch
    b.set_path(2, 1)
    b.render(cam, W, H, SPP)
    for k in (1, 2):  # lit spheres; a soup under the world colour alone
        sc = orc.Scene(parts[k], 4, P.NODE)
        px = np.zeros(P.PX_LEN * W * H + 1, dtype=np.uint32)
        for f in range(SPP):
            sc.sample(cam, W, H, px, f, nthreads=8)
        compare_planes(b.px(k), px, W, H)
        assert P.planes(px, W, H)["radiance"].max() > 0


def test_recorded_pass_replays_equal_eager(batch):
    """Build and render captured as one hipGraph: each replay, with other
    work on the same images between, gives the eager images again."""
    print("[batch render] recorded pass", flush=True)
    _, b, cam, want = batch
    b.set_path(3, 0)
    b.set_path_build(P.Tracer.PATH_SPLIT)
    b.record(cam, W, H, SPP, rebuild=True)
    other = S.path_camera(W, H, rr_from=3, eye=(1.5, -5.0, 3.0))
    for _ in range(3):
        b.render(other, W, H, 1)
        assert b.replay() > 0
        for k in range(b.scenes):
            assert np.array_equal(b.px(k), want[k]), k
    b.render(cam, W + 16, H, 1)  # new image size: the graph is refused
    with pytest.raises(P.PhrError):
        b.replay()
    b.set_path_build(P.Tracer.PATH_NESTED)


def test_rebuild_and_autotune_keep_the_images(batch):
    print("[batch render] rebuild, autotune", flush=True)
    _, b, cam, want = batch
    assert b.rebuild() > 0
    waves, grouped, best = b.autotune(cam, W, H, SPP, reps=1)
    assert waves in (2, 3, 4, 5) and grouped in (0, 1) and best > 0
    assert b.path() == (waves, grouped) and b.path_build in (P.Tracer.PATH_NESTED, P.Tracer.PATH_SPLIT)
    b.render(cam, W, H, SPP)
    for k in range(b.scenes):
        assert np.array_equal(b.px(k), want[k]), k


def test_bad_batches_are_refused(gpu):
    print("[batch render] refusals", flush=True)
    parts, tris, lo = batch_scenes()
    with pytest.raises(P.PhrError):
        P.Batch(gpu, tris, np.array([0, 10, 10, len(tris)], dtype=np.uint32))  # an empty scene
    with pytest.raises(P.PhrError):
        P.Batch(gpu, tris, np.array([5, len(tris)], dtype=np.uint32))
    b = P.Batch(gpu, tris[: lo[3]], lo[:4])
    with pytest.raises(P.PhrError):
        b.set_path(2, 1)  # three scenes: no grouped order
    bad = camera()
    bad["adapt"] = (0.1, 2.0, 2.0, 1.0)
    with pytest.raises(P.PhrError):
        b.render(bad, W, H, 1)
    b.set_path(4, 0)
    b.render(camera(), W, H, 1)


def test_full_batch_scenes_match_their_tracers(gpu):
    """The full batch workload: 256 scenes, 9,402,507 triangles in one build
    and one launch (2 of the 16 samples, under the world colour), with the
    nested and the split build. Four scenes across the batch, the last
    included, equal their own tracers."""
    print("[batch render] 256-scene workload", flush=True)
    tris, lo = S.batch_tris(256)
    assert int(lo[-1]) == 9_402_507
    b = P.Batch(gpu, tris, lo)
    w = h = 256
    cam = S.path_camera(w, h, eye=(0.5, 0.5, -1.5), target=(0.5, 0.5, 0.5), up=(0.0, 1.0, 0.0), rr_from=3)
    cam["env"] = (1.0, 1.0, 1.0, 0.0)
    b.set_path(4, 1)
    b.render(cam, w, h, 2)
    nested = {k: b.px(k).copy() for k in (0, 97, 200, 255)}
    b.set_path_build(P.Tracer.PATH_SPLIT)
    b.render(cam, w, h, 2)
    for k in (0, 97, 200, 255):
        tr = P.Tracer.new(gpu, tris[lo[k]:lo[k + 1]], 4)
        tr.sample(cam, w, h, 2)
        want = tr.px()
        assert np.array_equal(nested[k], want), ("nested", k)
        assert np.array_equal(b.px(k), want), ("split", k)
        assert P.planes(want, w, h)["radiance"].max() > 0
        tr.free()
    b.free()


def test_batch_display_under_agx_matches_the_tracer(gpu):
    """The display step in the batch launch follows the camera's view words
    (AgX, exposure, gamma) exactly as a tracer's launch does."""
    print("[batch render] AgX display", flush=True)
    parts, tris, lo = batch_scenes()
    b = P.Batch(gpu, tris[: lo[4]], lo[:5])
    cam = camera()
    cam["view"] = (2.0, 0.7, 1.4, 0.0)
    b.set_path(3, 0)
    b.render(cam, W, H, SPP)
    for k in (0, 2):
        tr = P.Tracer.new(gpu, parts[k], 4)
        tr.sample(cam, W, H, SPP)
        assert np.array_equal(b.px(k), tr.px()), k
        tr.free()
    b.free()

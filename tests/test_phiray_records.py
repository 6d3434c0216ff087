"""phiray's own `Tri` and `Camera` bytes through the C ABI (CPU; the GPU
case renders them).

* The numpy dtypes mirror the reference structs field by field: `Tri`
  (src/trace/kernel.rs:11-65) is 760 bytes with `us` at byte 112 and `emit`
  at 144; `Camera` (src/trace/camera.rs:12-228) is 5,776 bytes, `right` at
  16 and the 64 lamps from byte 656.
* phr_tri_from_phiray / phr_camera_from_phiray give exactly the tracer's own
  records the rest of the suite renders (bytes equal), so a scene handed over
  in phiray's layout renders the same image buffer (the GPU test checks it
  against the oracle too).
* Each field outside the restated subset is refused with
  PHR_ERR_UNSUPPORTED.
"""
import numpy as np
import pytest

import phiray_amd as P
from phiray_amd import scenes as S


def as_phiray_tris(tris):
    out = P.phiray_tris(len(tris))
    for k in ("a", "b", "c", "na", "nb", "nc", "color", "emit", "rough", "metal", "trans", "ior", "bump"):
        out[k] = tris[k]
    return out


def as_phiray_camera(cam):
    out = np.zeros(1, dtype=P.PHIRAY_CAMERA)
    for k in ("right", "persp", "up", "scale", "back", "units", "pivot", "aspect", "sun", "dist", "pan", "ao",
              "eps", "env", "frame", "count", "glow", "bounces", "depth", "integrator", "w", "h", "adapt", "view"):
        out[k] = cam[k]
    return out


def test_layouts_follow_the_reference_structs():
    assert P.PHIRAY_TRI.itemsize == 760 and P.PHIRAY_CAMERA.itemsize == 5776
    assert P.PHIRAY_TRI.fields["us"][1] == 112 and P.PHIRAY_TRI.fields["emit"][1] == 144
    assert P.PHIRAY_TRI.fields["maps"][1] == 160 and P.PHIRAY_TRI.fields["color_proc"][1] == 232
    assert P.PHIRAY_CAMERA.fields["right"][1] == 16 and P.PHIRAY_CAMERA.fields["sky"][1] == 96
    assert P.PHIRAY_CAMERA.fields["env"][1] == 160 and P.PHIRAY_CAMERA.fields["bounces"][1] == 216
    assert P.PHIRAY_CAMERA.fields["adapt"][1] == 592 and P.PHIRAY_CAMERA.fields["depth"][1] == 624
    assert P.PHIRAY_CAMERA.fields["lights"][1] == 656


def test_phiray_triangles_become_the_tracers_records():
    tris = S.spheres_scene(spheres=3, level=1, seed=2)
    got = P.tris_from_phiray(as_phiray_tris(tris))
    assert got.tobytes() == tris.tobytes()


def test_phiray_camera_becomes_the_tracers_record():
    cam = S.path_camera(64, 48)
    cam["w"], cam["h"] = 64, 48
    cam["view"] = (1.0, 0.5, 2.2, 3.0)
    cam["adapt"] = (0.01, 16.0, 4.0, 1.0)
    got = P.camera_from_phiray(as_phiray_camera(cam))
    assert got.tobytes() == cam.tobytes()


@pytest.mark.parametrize("field,value", [
    ("tex", 0), ("nmap", 3), ("sss", 0.5), ("sheen", 1.0), ("coat", 0.2), ("second", 0),
])
def test_unsupported_triangles_are_refused(field, value):
    t = as_phiray_tris(S.spheres_scene(spheres=1, level=0, seed=2))
    t[field][3] = value
    with pytest.raises(P.PhrError, match=r"\(-3\)"):
        P.tris_from_phiray(t)


def test_texture_maps_and_procedurals_are_refused():
    base = as_phiray_tris(S.spheres_scene(spheres=1, level=0, seed=2))
    t = base.copy()
    t["maps"][0] = (P.NO_TEX, 2, P.NO_TEX)
    with pytest.raises(P.PhrError, match=r"\(-3\)"):
        P.tris_from_phiray(t)
    for slot in ("color_proc", "bump_proc", "color_proc2"):
        t = base.copy()
        t[slot]["kind"][1] = 1
        with pytest.raises(P.PhrError, match=r"\(-3\)"):
            P.tris_from_phiray(t)


@pytest.mark.parametrize("field,value", [
    ("sky", (0.2, 0.2, 0.2, 0.0)), ("grid", 1.0), ("film", (1, 0, 0, 0)), ("envmap", 1), ("aperture", 0.05),
    ("tile", (0, 0, 128, 128)), ("taa", (0.5, 0.0, 0.0, 0.0)), ("limit", (10.0, 0.0, 0.0, 0.0)),
])
def test_unsupported_cameras_are_refused(field, value):
    c = as_phiray_camera(S.path_camera(32, 32))
    c[field] = value
    with pytest.raises(P.PhrError, match=r"\(-3\)"):
        P.camera_from_phiray(c)


def test_a_lit_lamp_is_refused_and_unlit_slots_pass():
    c = as_phiray_camera(S.path_camera(32, 32))
    c["count"] = 2
    P.camera_from_phiray(c)
    c["lights"][0, 1]["color"] = (1.0, 1.0, 1.0)
    with pytest.raises(P.PhrError, match=r"\(-3\)"):
        P.camera_from_phiray(c)
    # A lit slot past `count` is not a lamp.
    c["count"] = 1
    P.camera_from_phiray(c)


@pytest.mark.gpu
def test_a_scene_in_phirays_layout_renders_as_the_oracle(gpu, orc):
    from test_gpu_trace import compare_planes

    print("[records] phiray Tri/Camera through the converters", flush=True)
    tris = S.spheres_scene(spheres=6, level=2, seed=4)
    w, h, spp = 48, 32, 2
    cam = S.path_camera(w, h)
    tr = P.Tracer.new(gpu, P.tris_from_phiray(as_phiray_tris(tris)), 4)
    tr.sample(P.camera_from_phiray(as_phiray_camera(cam)), w, h, spp)
    sc = orc.Scene(tris, 4, P.NODE)
    px = np.zeros(P.PX_LEN * w * h + 1, dtype=np.uint32)
    for f in range(spp):
        sc.sample(cam, w, h, px, f, nthreads=8)
    compare_planes(tr.px(), px, w, h)

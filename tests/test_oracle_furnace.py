"""Multi-bounce pins for the oracle's integrator (CPU only).

What each case fixes, independently of the restated formulas:

* **Integrating sphere** (Goebel, "Generalized integrating-sphere theory",
  Applied Optics 6(1), 1967): inside a sphere, the cosine-weighted share of
  directions from any wall point that leave through a cap equals the cap's
  share f of the sphere's area. Under a uniform world of radiance L seen
  through the cap, a wall of albedo a therefore shows exactly
  ``a f L (1 - q^K) / (1 - q)`` with ``q = a (1 - f)`` after K diffuse
  bounces (``a f L / (1 - q)`` for K -> inf). The light arrives only on
  cosine-sampled rays that miss (no emitter, no lamp, no world image), so
  the case holds the bounce count, the throughput product, the cosine
  sampling and Russian roulette's weight: the same closed form must come out
  with roulette off and with roulette from the first bounce. Dropping a
  bounce moves the K = 4 value by 11 %; uniform instead of cosine sampling,
  or roulette without dividing by its survival chance, by more than 10 %.
* **phiray's own Cycles numbers** (tests/world_power.rs:10-24, 81-132, the
  rig of tests/power/mod.rs): a uniform dome of radiance 1 over a 20 m plane
  of albedo 0.8 gives 0.800 (and 0.400 for half the dome, linearly); the
  floor of a one-metre trench gives 0.584 at four diffuse bounces, and at one
  bounce more than 0.75 x 0.462 and less than four's / 1.1. Tolerances are
  the reference test's own (5 %, 2 %).
* **One sample against the converged image** (tests/emitter.rs:39-42): the
  near glowing panel's one-sample mean is within 15 % of its converged mean.
"""
import math

import numpy as np
import pytest

import phiray_amd as P
from phiray_amd import scenes as S

F32 = np.float32


def _tris(a, b, c, color, normal=None):
    t = np.zeros(len(a), dtype=P.TRI)
    t["a"], t["b"], t["c"] = a, b, c
    n = normal if normal is not None else np.cross(b - a, c - a)
    n = n / np.linalg.norm(n, axis=-1, keepdims=True)
    for k in ("na", "nb", "nc"):
        t[k] = n
    t["color"] = color
    t["rough"], t["ior"] = 1.0, 1.0
    return t


def integrating_sphere(level=5, z_cap=0.6, albedo=0.8):
    """A unit icosphere seen from inside, the faces whose centroid lies above
    z_cap removed (the port); returns the triangles and the port's share f of
    the whole facet area."""
    v, f = S.icosphere(level)
    a, b, c = v[f[:, 0]], v[f[:, 1]], v[f[:, 2]]
    area = 0.5 * np.linalg.norm(np.cross(b - a, c - a), axis=1)
    cz = (a[:, 2] + b[:, 2] + c[:, 2]) / 3.0
    keep = cz <= z_cap
    frac = float(area[~keep].sum() / area.sum())
    # Wound so the face normal points inward, at the camera.
    a, b, c = a[keep], c[keep], b[keep]
    inward = -(a + b + c) / 3.0
    t = _tris(a.astype(F32), b.astype(F32), c.astype(F32), (albedo, albedo, albedo, 1.0), inward)
    return t, frac


def sphere_camera(w, bounces, rr_from, level=1.0):
    cam = S.path_camera(w, w, eye=(0.0, 0.0, 0.0), target=(0.0, 0.0, -1.0), up=(0.0, 1.0, 0.0), fov_deg=40.0,
                        bounces=bounces, rr_from=rr_from)
    cam["env"] = (level, level, level, 1.0)
    return cam


def run(orc, tris, cam, w, spp, nthreads=8):
    sc = orc.Scene(tris, 4, P.NODE, nthreads)
    px = np.zeros(P.PX_LEN * w * w + 1, dtype=np.uint32)
    per = []
    for f in range(spp):
        before = P.planes(px, w, w)["radiance"].copy()
        sc.sample(cam, w, w, px, f, nthreads)
        per.append((P.planes(px, w, w)["radiance"] - before)[:, 0].copy())
    per = np.array(per, dtype=np.float64)  # [spp, pixels]
    return per


@pytest.fixture(scope="module")
def sphere():
    return integrating_sphere()


def closed_form(a, f, k, level=1.0):
    q = a * (1.0 - f)
    return a * f * level * (1.0 - q ** k) / (1.0 - q)


@pytest.mark.parametrize("bounces,rr_from", [(1, 1000), (2, 1000), (4, 1000), (4, 1), (24, 1000), (24, 1)])
def test_integrating_sphere_closed_form(orc, sphere, bounces, rr_from):
    tris, f = sphere
    assert 0.19 < f < 0.21
    w, spp = 16, 192
    per = run(orc, tris, sphere_camera(w, bounces, rr_from), w, spp)
    got = per.mean()
    # The samples are independent across pixels and frames; 4.5 sigma plus
    # the facets' departure from the sphere (0.5 %).
    sigma = per.std() / math.sqrt(per.size)
    want = closed_form(0.8, f, bounces)
    assert abs(got - want) < 4.5 * sigma + 0.005 * want, (bounces, rr_from, got, want, sigma)
    # A bounce fewer is out of reach (the test can tell K from K - 1).
    if bounces > 1 and bounces < 24:
        assert abs(got - closed_form(0.8, f, bounces - 1)) > 8 * sigma


def test_roulette_is_unbiased_on_the_same_scene(orc, sphere):
    """Roulette from the first bounce against none: equal expectations, the
    roulette run noisier (it trades variance for shorter paths)."""
    tris, f = sphere
    w, spp = 16, 128
    off = run(orc, tris, sphere_camera(w, 6, 1000), w, spp)
    on = run(orc, tris, sphere_camera(w, 6, 1), w, spp)
    se = math.sqrt(off.var() / off.size + on.var() / on.size)
    assert abs(off.mean() - on.mean()) < 4.5 * se
    assert on.std() > off.std()


def test_world_level_is_linear_and_lands_on_the_closed_form(orc, sphere):
    tris, f = sphere
    w, spp = 8, 96
    one = run(orc, tris, sphere_camera(w, 3, 1000, 1.0), w, spp).mean()
    three = run(orc, tris, sphere_camera(w, 3, 1000, 3.0), w, spp).mean()
    # Same streams, same paths: the world's level is a factor on every sample.
    assert abs(three / one - 3.0) < 1e-5
    assert abs(one - closed_form(0.8, f, 3)) < 0.03 * one


# ---- phiray's Cycles numbers (tests/world_power.rs, tests/power/mod.rs) --------------
WORLD_1_OVER_ALBEDO = 0.800055
WORLD_HALF_OVER_ALBEDO = 0.400027
TRENCH_B1 = 0.462051
TRENCH_B4 = 0.583955


def plane():
    """tests/power/mod.rs:19-29: 20 x 20 m at z = 0, albedo 0.8, rough 1."""
    q = np.array([[-10.0, -10.0, 0.0], [10.0, -10.0, 0.0], [10.0, 10.0, 0.0], [-10.0, 10.0, 0.0]], dtype=F32)
    return _tris(np.array([q[0], q[0]]), np.array([q[1], q[2]]), np.array([q[2], q[3]]), (0.8, 0.8, 0.8, 1.0),
                 np.array([[0, 0, 1.0], [0, 0, 1.0]]))


def trench():
    """tests/world_power.rs:36-48: two walls a metre high, a metre apart."""
    out = [plane()]
    for x in (-0.5, 0.5):
        n = np.array([1.0 if x < 0 else -1.0, 0.0, 0.0])
        q = np.array([[x, -10.0, 0.0], [x, 10.0, 0.0], [x, 10.0, 1.0], [x, -10.0, 1.0]], dtype=F32)
        out.append(_tris(np.array([q[0], q[0]]), np.array([q[1], q[2]]), np.array([q[2], q[3]]),
                         (0.8, 0.8, 0.8, 1.0), np.array([n, n])))
    return np.concatenate(out)


def overhead(level, bounces):
    """tests/power/mod.rs:31-66: orthographic, straight down, 2 m across."""
    cam = np.zeros(1, dtype=P.CAMERA)
    c = cam[0]
    c["right"], c["up"], c["back"] = (1, 0, 0), (0, 1, 0), (0, 0, 1)
    c["dist"], c["units"], c["persp"], c["scale"], c["aspect"] = 10.0, 1.0, 0, 0.5, 1.0
    c["env"] = (level, level, level, 1.0)
    c["count"] = 1
    c["bounces"] = bounces
    c["depth"] = (13, 9, 9, 1)
    c["rr_from"] = 1000
    return cam


def centre(orc, tris, cam, spp=256, w=32):
    """tests/world_power.rs:60-78: the mean of the centre 4 x 4 pixels."""
    sc = orc.Scene(tris, 4, P.NODE, 8)
    px = np.zeros(P.PX_LEN * w * w + 1, dtype=np.uint32)
    for f in range(spp):
        sc.sample(cam, w, w, px, f, 8)
    pl = P.planes(px, w, w)
    rad = (pl["radiance"] / pl["count"][:, None]).reshape(w, w, 3)
    lo, hi = w // 2 - 2, w // 2 + 2
    return rad[lo:hi, lo:hi].reshape(-1, 3).mean(axis=0)


def test_a_uniform_world_comes_back_as_the_albedo_times_its_radiance(orc):
    full = centre(orc, plane(), overhead(1.0, 4))[0]
    half = centre(orc, plane(), overhead(0.5, 4))[0]
    assert abs(full / WORLD_1_OVER_ALBEDO - 1.0) < 0.05, full
    assert abs(half / WORLD_HALF_OVER_ALBEDO - 1.0) < 0.05, half
    assert abs(full / half - 2.0) < 0.02


def test_a_shadowed_surface_gets_what_bounces_to_it(orc):
    open_ = centre(orc, plane(), overhead(1.0, 4))[0]
    one = centre(orc, trench(), overhead(1.0, 1))[0]
    four = centre(orc, trench(), overhead(1.0, 4))[0]
    assert one < open_ * 0.7, (one, open_)
    assert four > one * 1.1, (one, four)
    assert abs(four / TRENCH_B4 - 1.0) < 0.05, four
    assert one > TRENCH_B1 * 0.75, one


# ---- tests/emitter.rs:39-42: one sample against the converged mean ------------------
def test_one_sample_is_near_the_converged_mean(orc):
    from test_oracle_path import center_mean, emitter_scenes, wall_camera

    _, tris, want, _ = emitter_scenes()[0]
    sc = orc.Scene(tris, 4, P.NODE)
    w = 16
    px = np.zeros(P.PX_LEN * w * w + 1, dtype=np.uint32)
    cam = wall_camera()
    sc.sample(cam, w, w, px, 0, 8)
    pl = P.planes(px, w, w)
    one = center_mean(pl["radiance"] / pl["count"][:, None])
    for f in range(1, 256):
        sc.sample(cam, w, w, px, f, 8)
    pl = P.planes(px, w, w)
    allm = center_mean(pl["radiance"] / pl["count"][:, None])
    assert abs(allm - want) < want * 0.1
    assert abs(one - allm) < allm * 0.15, (one, allm)

"""The oracle's integrator and its arithmetic (CPU only).

Pinning:
* orc_sin / orc_cos / orc_exp2 — against libm in double precision, to a few
  single-precision ulps over the ranges the path uses;
* orc_hash / orc_rnd — the PCG-RXS-M-XS hash of Jarzynski & Olano (JCGT 9(3),
  2020) written out here in Python integers, the seed formula of
  shade.wgsl:68, and the uniformity of its outputs;
* the integrator (orc_sample) — the three closed forms of tests/emitter.rs:30-57:
  a huge glowing panel 0.2 in front of a grey wall gives 0.5 L; a small far
  one (0.5 / pi) L A / d^2; a 1 x 1 panel at 1, 4 L s atan(s) (0.5 / pi). The
  wall is made IOR 1 (no dielectric reflection) and Max Bounces glossy 0,
  the subset the restatement carries, so the closed forms of a Lambertian wall
  apply. Bit-level radiance is pinned only by the GPU kernel agreeing with
  this restatement (parity unpinned against the WGSL itself, which cannot run
  here).
"""
import math

import numpy as np
import pytest

import phiray_amd as P
from phiray_amd import scenes as S


def ulps(a, b):
    a = np.float32(a)
    return abs(float(a) - b) / max(np.spacing(np.float32(abs(b))), np.float32(1e-45))


def test_sin_cos_against_libm(orc):
    L = orc.lib()
    xs = np.concatenate([np.linspace(-7.0, 7.0, 20001), np.linspace(0, 6.2831855, 4097),
                         np.array([0.0, 1e-8, np.pi / 4, np.pi / 2, np.pi, 3 * np.pi / 2])]).astype(np.float32)
    worst = 0.0
    for x in xs:
        s, c = L.orc_sin(float(x)), L.orc_cos(float(x))
        xd = float(x)
        assert abs(s - math.sin(xd)) <= 2.5e-7 and abs(c - math.cos(xd)) <= 2.5e-7
        if abs(math.sin(xd)) > 0.1:
            worst = max(worst, ulps(s, math.sin(xd)))
    assert worst <= 4


def test_exp2_against_libm(orc):
    L = orc.lib()
    for x in np.linspace(-30.0, 30.0, 6001).astype(np.float32):
        got = L.orc_exp2(float(x))
        assert ulps(got, 2.0 ** float(x)) <= 4, x
    assert L.orc_exp2(0.0) == 1.0 and L.orc_exp2(-1.0) == 0.5


def pcg(x):
    h = (x * 747796405 + 2891336453) & 0xFFFFFFFF
    h = (((h >> ((h >> 28) + 4)) ^ h) * 277803737) & 0xFFFFFFFF
    return ((h >> 22) ^ h) & 0xFFFFFFFF


def test_pcg_hash_and_rnd(orc):
    L = orc.lib()
    rng = np.random.default_rng(3)
    for x in list(rng.integers(0, 2 ** 32, 2000)) + [0, 1, 0xFFFFFFFF, 9781 * 5 + 6271 * 7 + 1]:
        assert L.orc_hash(int(x)) == pcg(int(x))
    import ctypes
    s = ctypes.c_uint32(12345)
    r = L.orc_rnd(ctypes.byref(s))
    assert s.value == pcg(12345) and r == np.float32(np.float32(pcg(12345)) / np.float32(4294967295.0))
    # Uniform: 64 bins over 2^16 consecutive seeds.
    vals = np.array([pcg(i) for i in range(1 << 16)], dtype=np.uint64) >> np.uint64(26)
    counts = np.bincount(vals.astype(np.int64), minlength=64)
    chi2 = float(((counts - 1024.0) ** 2 / 1024.0).sum())
    assert chi2 < 120  # 63 dof; p ~ 1e-5 tail


def panel(y, s, color, emit=(0, 0, 0), rough=1.0, ior=1.0):
    q = np.array([[-s, y, -s], [s, y, -s], [s, y, s], [-s, y, s]], dtype=np.float32)
    t = np.zeros(2, dtype=P.TRI)
    t["a"], t["b"], t["c"] = [q[0], q[0]], [q[1], q[2]], [q[2], q[3]]
    for k in ("na", "nb", "nc"):
        t[k] = (0.0, 1.0, 0.0)
    t["color"] = color
    t["rough"] = rough
    t["ior"] = ior
    t["emit"] = emit
    return t


def wall_camera():
    """tests/emitter.rs:8-13, with Max Bounces (glossy 0) and one diffuse bounce."""
    cam = np.zeros(1, dtype=P.CAMERA)
    c = cam[0]
    c["right"], c["up"], c["back"] = (1, 0, 0), (0, 0, 1), (0, 1, 0)
    c["units"], c["scale"], c["aspect"], c["persp"], c["dist"], c["count"] = 1.0, 4.0, 1.0, 1, 0.1, 1
    c["ao"] = 1e-3
    c["depth"] = (5, 9, 9, 1)
    c["bounces"] = 1
    return cam


def emitter_scenes():
    wall = panel(0.0, 40.0, (0.5, 0.5, 0.5, 1.0))
    near = np.concatenate([wall, panel(0.2, 40.0, (0, 0, 0, 0), emit=(0.4, 0.4, 0.4), ior=1.5)])
    far = np.concatenate([wall, panel(2.0, 0.1, (0, 0, 0, 0), emit=(50.0, 50.0, 50.0), ior=1.5)])
    mid = np.concatenate([wall, panel(1.0, 0.5, (0, 0, 0, 0), emit=(2.0, 2.0, 2.0), ior=1.5)])
    s = 0.5 / math.sqrt(1.25)
    return [
        ("near", near, 0.5 * 0.4, 0.10),
        ("far", far, 0.5 / math.pi * 50.0 * 0.04 / 4.0, 0.10),
        ("middle", mid, 0.5 / math.pi * 4.0 * 2.0 * s * math.atan(s), 0.08),
    ]


def center_mean(rad):
    img = rad.reshape(16, 16, 3)[..., 0]
    return float(img[4:12, 4:12].mean())


@pytest.mark.parametrize("case", range(3))
def test_emitters_are_counted_once_near_and_far(orc, case):
    name, tris, want, tol = emitter_scenes()[case]
    sc = orc.Scene(tris, 4, P.NODE)
    w = h = 16
    px = np.zeros(P.PX_LEN * w * h + 1, dtype=np.uint32)
    cam = wall_camera()
    for f in range(256):
        sc.sample(cam, w, h, px, f)
    pl = P.planes(px, w, h)
    rad = pl["radiance"] / pl["count"][:, None]
    got = center_mean(rad)
    assert abs(got - want) < want * tol, f"{name}: converged {got} vs expected {want}"


@pytest.mark.parametrize("exposure,gamma", [(0.0, 0.0), (1.0, 2.2)])
def test_display_is_the_view_transform_of_the_mean(orc, exposure, gamma):
    """util.wgsl:133-144 and view.wgsl:62-107 (Standard): the shown value is
    the sRGB curve (here in float64) of the light's mean over the samples
    that met something, after exposure and before gamma, under the sky's
    share of the samples; the packed word is its 8-bit rounding with the
    film's coverage as alpha."""
    tris = S.spheres_scene(spheres=3, level=1, seed=5)
    w, h = 24, 16
    cam = S.path_camera(w, h, rr_from=3)
    cam["view"] = (0.0, exposure, gamma, 0.0)
    sc = orc.Scene(tris, 4, P.NODE)
    px = np.zeros(P.PX_LEN * w * h + 1, dtype=np.uint32)
    for f in range(3):
        sc.sample(cam, w, h, px, f, 8)
    pl = P.planes(px, w, h)
    k = pl["count"].astype(np.float64)
    assert (k == 3).all() and pl["missed"].max() > 0 and pl["missed"].min() == 0
    m = pl["radiance"] / np.maximum(k - pl["missed"], 1.0)[:, None]
    c = np.clip(np.maximum(m, 0.0) * 2.0 ** exposure, 0.0, 1.0)
    d = np.where(c <= 0.0031308, c * 12.92, 1.055 * c ** (1 / 2.4) - 0.055)
    want = np.clip(d, 0.0, 1.0) ** (1.0 / (gamma if gamma > 0 else 1.0))
    # A miss lays the sky over the light: coverage in `cov`, colour in `grid`.
    want = want * (1.0 - pl["cov"] / k)[:, None] + pl["grid"] / k[:, None]
    assert (pl["cov"] == pl["missed"]).all()
    assert np.allclose(pl["display"], want, rtol=2e-6, atol=2e-7)
    q = np.floor(np.clip(want, 0, 1) * 255.0 + 0.5)
    assert (np.abs(pl["pixel"][:, :3].astype(np.int32) - q) <= 1).all()
    assert (pl["pixel"][:, 3] == 255).all()
    assert pl["display"].max() > 0.3


def view_of(orc, lin, view):
    lin = np.ascontiguousarray(lin, dtype=np.float32)
    v = np.asarray(view, dtype=np.float32)
    out = np.zeros(3, dtype=np.float32)
    orc.lib().orc_view_transform(orc.ptr(lin), orc.ptr(v), orc.ptr(out))
    return out


@pytest.mark.parametrize("kind", [0, 1, 2])
def test_view_transforms_behave(orc, kind):
    """view.wgsl:5-99 under each transform: the display rises with the light
    and saturates, halving the light does not halve the display (the
    transforms are concave), and the exposure is a factor of two a stop.
    The curve samples themselves are phiray's table (parity unpinned beyond
    these behaviours)."""
    ramp = np.geomspace(1e-4, 64.0, 200)
    shown = np.array([view_of(orc, (x, x, x), (kind, 0, 0, 0)) for x in ramp])
    # Per channel under Standard and Filmic; AgX's outset matrix has negative
    # terms, so there the luminance rises while a channel may dip.
    luma = shown @ np.array([0.2126, 0.7152, 0.0722])
    assert (np.diff(shown if kind < 2 else luma[:, None], axis=0) >= (-1e-7 if kind < 2 else -1e-4)).all()
    assert shown[-1].min() > 0.98 and shown[0].max() < 0.05
    for x in (0.05, 0.18, 0.6):
        full, half = view_of(orc, (x, x, x), (kind, 0, 0, 0)), view_of(orc, (x / 2, x / 2, x / 2), (kind, 0, 0, 0))
        assert (half < full).all() and (half > 0.5 * full).all(), (x, full, half)
        up = view_of(orc, (x, x, x), (kind, 1.0, 0, 0))
        assert np.allclose(up, view_of(orc, (2 * x, 2 * x, 2 * x), (kind, 0, 0, 0)), rtol=2e-6, atol=1e-7)


def test_filmic_looks_bend_the_midtones_their_way(orc):
    """tests/show.rs:120-150: against the base Filmic curve, Very High
    Contrast (look 1) pushes a dim value darker and Very Low Contrast (6)
    lifts it; the looks do nothing under Standard."""
    dim = (0.05, 0.05, 0.05)
    base = view_of(orc, dim, (1, 0, 0, 0))[0]
    assert view_of(orc, dim, (1, 0, 0, 1))[0] < base < view_of(orc, dim, (1, 0, 0, 6))[0]
    assert (view_of(orc, dim, (0, 0, 0, 1)) == view_of(orc, dim, (0, 0, 0, 0))).all()

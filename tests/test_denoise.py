"""The a-trous denoiser (src/kernel/denoise.wgsl, src/trace/read.rs:73-119).

CPU (the oracle's restatement), after tests/denoise.rs:79-134, with the
environment light of that test replaced by an emissive panel (environment
images are outside the restated subset):
* denoising a 12-sample render moves it towards the 1,024-sample one (mean
  absolute difference below 0.75x the noisy one's);
* the albedo edge between a dark and a pale quad survives and stays sharp;
* `denoise_shown` (keep) leaves the linear sums alone, `denoise` replaces them;
* the last pass's display plane is the Standard view transform (sRGB curve
  in float64, exposure, gamma) of the filtered light over the share of
  samples that met something, and the packed word its 8-bit rounding.
The restatement's own arithmetic has no independent pin beyond these
behaviours (parity unpinned below the qualitative bar); the GPU kernel is held
to it within 1e-5 relative.

GPU: the kernel over the same image buffer equals the restatement.
"""
import numpy as np
import pytest

import phiray_amd as P

SIDE = 48


def scene():
    n = (0.0, 0.0, 1.0)

    def quad(x0, x1, c, z=0.0, emit=(0, 0, 0), flip=False):
        q = np.array([[x0, -4.0, z], [x1, -4.0, z], [x1, 4.0, z], [x0, 4.0, z]], dtype=np.float32)
        t = np.zeros(2, dtype=P.TRI)
        if flip:
            t["a"], t["b"], t["c"] = [q[0], q[0]], [q[2], q[3]], [q[1], q[2]]
        else:
            t["a"], t["b"], t["c"] = [q[0], q[0]], [q[1], q[2]], [q[2], q[3]]
        for k in ("na", "nb", "nc"):
            t[k] = n
        t["color"], t["rough"], t["ior"], t["emit"] = c, 1.0, 1.0, emit
        return t

    tris = np.concatenate([
        quad(-4.0, 0.0, (0.05, 0.05, 0.05, 1.0)),
        quad(0.0, 4.0, (0.8, 0.8, 0.8, 1.0)),
        # A small bright panel overhead: most diffuse rays miss it (the noise).
        quad(-0.3, 0.3, (0.0, 0.0, 0.0, 0.0), z=3.0, emit=(40.0, 40.0, 40.0), flip=True)[:1],
    ])
    cam = np.zeros(1, dtype=P.CAMERA)
    c = cam[0]
    c["right"], c["up"], c["back"] = (1, 0, 0), (0, 1, 0), (0, 0, 1)
    c["units"], c["scale"], c["aspect"], c["persp"], c["dist"], c["count"] = 1.0, 0.12, 1.0, 1, 1.0, 1
    c["bounces"] = 1
    c["depth"] = (5, 9, 9, 1)
    return tris, cam


def render(orc, samples, nthreads=8):
    tris, cam = scene()
    sc = orc.Scene(tris, 4, P.NODE)
    px = np.zeros(P.PX_LEN * SIDE * SIDE + 1, dtype=np.uint32)
    for f in range(samples):
        sc.sample(cam, SIDE, SIDE, px, f, nthreads)
    return px


def radiance(px):
    pl = P.planes(px, SIDE, SIDE)
    return pl["radiance"] / np.maximum(pl["count"], 1)[:, None]


def denoise(orc, px, keep=False, view=(0.0, 0.0, 0.0, 0.0)):
    out = px.copy()
    v = np.array(view, dtype=np.float32)
    orc.lib().orc_denoise(orc.ptr(out), SIDE, SIDE, 1 if keep else 0, orc.ptr(v))
    return out


def srgb_view(lin, exposure=0.0, gamma=0.0):
    """view.wgsl:62-99 for the Standard transform, in float64."""
    c = np.maximum(lin.astype(np.float64), 0.0) * 2.0 ** exposure
    x = np.clip(c, 0.0, 1.0)
    d = np.where(x <= 0.0031308, x * 12.92, 1.055 * x ** (1 / 2.4) - 0.055)
    return np.clip(d, 0.0, 1.0) ** (1.0 / (gamma if gamma > 0 else 1.0))


@pytest.fixture(scope="module")
def renders(orc):
    return render(orc, 1024), render(orc, 12)


def test_denoising_moves_a_noisy_render_towards_the_converged_one(orc, renders):
    ref, noisy = renders
    cleaned = denoise(orc, noisy)
    before = np.abs(radiance(noisy) - radiance(ref)).mean()
    after = np.abs(radiance(cleaned) - radiance(ref)).mean()
    assert after < before * 0.75, f"denoised {after:.4f} from the reference, noisy was {before:.4f}"


def test_demodulation_keeps_the_albedo_edge(orc, renders):
    _, noisy = renders
    img = radiance(denoise(orc, noisy)).reshape(SIDE, SIDE, 3)[..., 0]
    row = img[SIDE // 2]
    left = row[4:SIDE // 2 - 4].mean()
    right = row[SIDE // 2 + 4:SIDE - 4].mean()
    assert right > 4 * left, f"the edge flattened: left {left:.4f}, right {right:.4f}"
    step = row[SIDE // 2 + 1] - row[SIDE // 2 - 2]
    assert step > (right - left) * 0.5


def test_denoising_what_is_shown_keeps_the_samples(orc, renders):
    _, noisy = renders
    shown = denoise(orc, noisy, keep=True)
    assert (radiance(shown) == radiance(noisy)).all()
    assert not (shown[14 * SIDE * SIDE:20 * SIDE * SIDE] == noisy[14 * SIDE * SIDE:20 * SIDE * SIDE]).all()
    assert not (radiance(denoise(orc, noisy)) == radiance(noisy)).all()


@pytest.mark.parametrize("exposure,gamma", [(0.0, 0.0), (1.5, 2.2), (-2.0, 0.8)])
def test_last_pass_writes_the_view_transform(orc, renders, exposure, gamma):
    _, noisy = renders
    out = denoise(orc, noisy, view=(0.0, exposure, gamma, 0.0))
    pl = P.planes(out, SIDE, SIDE)
    k = np.maximum(pl["count"], 1).astype(np.float64)
    met = 1.0 - pl["missed"] / k
    lin = pl["radiance"] / k[:, None]  # keep=False: the filtered light's sums
    # Under the sky's share of the samples (`cov`) and its colour (`grid`).
    want = srgb_view(lin / np.maximum(met, 1.0 / k)[:, None], exposure, gamma) * (1.0 - pl["cov"] / k)[:, None]
    want = want + pl["grid"] / k[:, None]
    assert pl["cov"].max() > 0
    assert np.allclose(pl["display"], want, rtol=2e-6, atol=2e-7)
    q = np.floor(np.clip(want, 0, 1) * 255.0 + 0.5)
    assert (np.abs(pl["pixel"][:, :3].astype(np.int32) - q) <= 1).all()
    assert (pl["pixel"][:, 3] == 255).all()  # every sample covers the film
    assert pl["display"].max() > 0.5


def test_log2_against_libm(orc):
    L = orc.lib()
    for x in np.concatenate([np.linspace(1e-3, 4.0, 4001), np.geomspace(1e-30, 1e30, 2001)]).astype(np.float32):
        want = np.log2(np.float64(x))
        got = L.orc_log2(float(x))
        assert abs(got - want) <= max(4 * np.spacing(np.float32(abs(want))), 2e-7), x


@pytest.mark.gpu
@pytest.mark.parametrize("keep", [False, True])
def test_gpu_denoise_matches_the_restatement(gpu, orc, keep):
    print(f"[denoise] keep={keep}", flush=True)
    tris, cam = scene()
    if keep:  # the display under a Filmic look (Medium Low Contrast)
        cam["view"] = (1.0, 0.5, 0.0, 4.0)
    tr = P.Tracer.new(gpu, tris, 4)
    tr.set_samples(64)
    tr.sample(cam, SIDE, SIDE, 12)
    before = tr.px()
    tr.denoise(keep=keep)
    got = tr.px()
    want = denoise(orc, before, keep, view=tuple(float(v) for v in cam[0]["view"]))
    n = SIDE * SIDE
    for lo, hi in [(1 * n, 4 * n), (4 * n, 7 * n), (14 * n, 20 * n)]:
        a, b = got[lo:hi].view(np.float32), want[lo:hi].view(np.float32)
        bad = ~np.isclose(a, b, rtol=1e-5, atol=0)
        assert not bad.any(), f"words {lo}..{hi}: {bad.sum()} beyond 1e-5"
    a, b = got[:n].view(np.uint8).astype(np.int32), want[:n].view(np.uint8).astype(np.int32)
    assert (np.abs(a - b) <= 1).all()
    assert (tr.pixels().reshape(-1) == got[:n].view(np.uint8)).all()
    assert (tr.display().reshape(-1) == got[4 * n:7 * n].view(np.float32)).all()

"""Closed-form pins for the oracle's a-trous denoiser and Adaptive Sampling
check (CPU only).

A-trous (src/kernel/denoise.wgsl:48-134, src/trace/read.rs:85-119), on image
buffers written by hand with flat guides (one albedo, normal and depth
everywhere), where every edge stop is exp(0) = 1:

* a constant image stays constant through the four passes;
* with the stop on the light switched off, a small impulse spreads as the
  convolution of the four dilated B3-spline kernels (taps 1/16, 1/4, 3/8 at
  steps 1, 2, 4, 8), computed here with numpy's convolution: a wrong tap,
  step or kernel order moves that shape;
* the firefly rule of the first pass: a centre whose demodulated luma is
  above 1.5 x its brightest neighbour + 0.05 is replaced by the neighbours'
  mean (an impulse just above the threshold vanishes, one just below stays);
* a normal edge at full strength (64) keeps a two-tone image exactly, a zero
  strength blurs it; a pixel that met nothing never mixes with one that did;
* a stronger stop on the light spreads an impulse less (monotone).

Adaptive Sampling's check (src/trace/wgsl/adapt.wgsl:47-79) and dilation
(:82-96) on pixels whose error is worked out by hand in the comments: above
and below the threshold, the square-root branch for intensities under one,
the step boundary and the minimum sample count, already-settled pixels, and
the four-neighbour dilation.
"""
import ctypes as C

import numpy as np
import pytest

import phiray_amd as P

W = H = 96
N = W * H


def buffer(rad, count=4, albedo=0.5, normal=None, depth=None):
    """An image buffer whose means are `rad` (N x 3), flat guides by default."""
    px = np.zeros(P.PX_LEN * N + 1, dtype=np.uint32)
    f = px.view(np.float32)
    f[N:4 * N] = (np.asarray(rad, dtype=np.float32) * count).reshape(-1)
    f[7 * N:10 * N] = albedo * count
    nrm = np.tile(np.array([0.0, 0.0, 1.0], dtype=np.float32), (N, 1)) if normal is None else normal
    f[10 * N:13 * N] = (nrm * count).reshape(-1)
    f[13 * N:14 * N] = (2.0 if depth is None else depth) * count
    px[31 * N:32 * N] = count
    f[24 * N:25 * N] = count
    return px


def mean_radiance(px, count=4):
    return px.view(np.float32)[N:4 * N].reshape(N, 3) / count


SRGB_HALF = 1.055 * 0.5 ** (1 / 2.4) - 0.055  # the sRGB-encoded albedo the filter divides by


def denoise(orc, px, strength=(64.0, 8.0, 40.0, 7.0), passes=(1, 2, 4, 8)):
    """The passes in order, the last one remodulating into the radiance
    plane. A single pass reads and writes the radiance plane's neighbours
    only through its own output plane (14n), so it is run as a first pass and
    read back from there, times the albedo it divided by."""
    out = px.copy()
    s = np.array(strength, dtype=np.float32)
    v = np.zeros(4, dtype=np.float32)
    for r, step in enumerate(passes):
        last = 1 if r == len(passes) - 1 and len(passes) > 1 else 0
        orc.lib().orc_denoise_pass(orc.ptr(out), W, H, step, r, last, 0, orc.ptr(s), orc.ptr(v))
    if len(passes) == 1:
        f = out.view(np.float32)
        f[N:4 * N] = f[14 * N:17 * N] * np.float32(SRGB_HALF) * 4
    return out


def bspline_response(steps=(1, 2, 4, 8)):
    """The impulse response of the separable 5-tap B3 kernel at each step, in
    float64 (numpy's convolution)."""
    h = np.array([1.0])
    for s in steps:
        k = np.zeros(4 * s + 1)
        k[::s] = [1 / 16, 1 / 4, 3 / 8, 1 / 4, 1 / 16]
        h = np.convolve(h, k)
    return np.outer(h, h)


def test_a_constant_image_is_a_fixed_point(orc):
    px = buffer(np.full((N, 3), 0.7))
    out = mean_radiance(denoise(orc, px))
    assert np.allclose(out, 0.7, rtol=2e-6)


@pytest.mark.parametrize("steps", [(1,), (1, 2), (1, 2, 4, 8)])
def test_an_impulse_spreads_as_the_dilated_bspline(orc, steps):
    rad = np.ones((N, 3))
    cy = cx = W // 2
    rad[cy * W + cx] += 0.4  # below the firefly threshold (0.5 b + 0.05 srgb(0.5) = 0.537)
    out = mean_radiance(denoise(orc, buffer(rad), strength=(64.0, 8.0, 40.0, 0.0), passes=steps))[:, 0]
    k = bspline_response(steps)
    r = k.shape[0] // 2
    want = np.ones((H, W))
    want[cy - r:cy + r + 1, cx - r:cx + r + 1] += 0.4 * k
    assert np.allclose(out.reshape(H, W), want, rtol=0, atol=2e-6)
    if len(steps) == 1:
        # One pass reaches exactly two pixels each way, with the corner tap 1/256.
        d = np.abs(out.reshape(H, W) - 1.0) > 1e-4
        ys, xs = np.nonzero(d)
        assert (xs.min(), xs.max(), ys.min(), ys.max()) == (cx - 2, cx + 2, cy - 2, cy + 2)


def test_the_firefly_rule_of_the_first_pass(orc):
    """Only the firefly's own output drops its light: the neighbours' taps
    still read it (the centre tap is the one the check replaced)."""
    threshold = 0.5 + 0.05 * SRGB_HALF  # luma(b + d) / a > 1.5 luma(b) / a + 0.05, b = 1
    k = bspline_response((1,))
    cy = cx = W // 2
    for delta, removed in ((threshold - 0.01, False), (threshold + 0.01, True), (10.0, True)):
        rad = np.ones((N, 3))
        rad[cy * W + cx] += delta
        out = mean_radiance(denoise(orc, buffer(rad), strength=(64.0, 8.0, 40.0, 0.0), passes=(1,)))[:, 0]
        want = np.ones((H, W))
        want[cy - 2:cy + 3, cx - 2:cx + 3] += delta * k
        if removed:
            want[cy, cx] = 1.0
        assert np.allclose(out.reshape(H, W), want, rtol=0, atol=2e-6 * max(1.0, delta)), delta


def test_a_normal_edge_stops_the_filter(orc):
    normal = np.tile(np.array([0.0, 0.0, 1.0], dtype=np.float32), (N, 1))
    right = (np.arange(N) % W) >= W // 2
    normal[right] = (1.0, 0.0, 0.0)
    rad = np.where(right[:, None], 3.0, 1.0) * np.ones((N, 3))
    kept = mean_radiance(denoise(orc, buffer(rad, normal=normal), strength=(64.0, 0.0, 0.0, 0.0)))[:, 0]
    assert np.allclose(kept, rad[:, 0], rtol=1e-6)
    blurred = mean_radiance(denoise(orc, buffer(rad, normal=normal), strength=(0.0, 0.0, 0.0, 0.0)))[:, 0]
    row = blurred.reshape(H, W)[H // 2]
    assert row[W // 2 - 1] > 1.3 and row[W // 2] < 2.7


def test_what_met_nothing_never_mixes_with_what_met_something(orc):
    right = (np.arange(N) % W) >= W // 2
    depth = np.where(right, -1.0, 2.0).astype(np.float32)
    rad = np.where(right[:, None], 0.0, 1.0) * np.ones((N, 3))
    out = denoise(orc, buffer(rad, depth=depth), strength=(0.0, 0.0, 0.0, 0.0))
    f = out.view(np.float32)
    assert np.allclose(mean_radiance(out)[~right, 0], 1.0, rtol=1e-6)


def test_the_stop_on_the_light_narrows_the_spread(orc):
    rad = np.ones((N, 3))
    rad[(H // 2) * W + W // 2] += 0.4
    spread = []
    for s in (0.0, 7.0, 70.0):
        out = mean_radiance(denoise(orc, buffer(rad), strength=(64.0, 8.0, 40.0, s)))[:, 0].reshape(H, W)
        spread.append(out[H // 2, W // 2 + 3] - 1.0)
    assert spread[0] > spread[1] > spread[2] >= 0.0


# ---- Adaptive Sampling's check and dilation ------------------------------------------
def adapt_camera(w, h, frame, threshold, min_samples, step):
    cam = np.zeros(1, dtype=P.CAMERA)
    c = cam[0]
    c["w"], c["h"], c["frame"] = w, h, frame
    c["adapt"] = (threshold, min_samples, step, 1.0)
    return cam


def adapt_buffer(w, h, full, half, count):
    n = w * h
    px = np.zeros(P.PX_LEN * n + 1, dtype=np.uint32)
    f = px.view(np.float32)
    f[n:4 * n] = np.asarray(full, dtype=np.float32).reshape(-1)
    f[26 * n:29 * n] = np.asarray(half, dtype=np.float32).reshape(-1)
    px[31 * n:32 * n] = count
    return px


def check(orc, cam, px):
    orc.lib().orc_adapt_check(orc.ptr(cam), orc.ptr(px))


def dilate(orc, cam, px):
    orc.lib().orc_adapt_dilate(orc.ptr(cam), orc.ptr(px))


def test_adaptive_check_by_hand(orc):
    w, h = 4, 1
    n = w * h
    full = np.array([
        [3.0, 3.0, 3.0],      # sums over 3 samples: intensity 3 >= 1, normalised by 3.
        [3.0, 3.0, 3.0],
        [0.12, 0.12, 0.12],   # sums over 4: intensity 0.09 < 1, normalised by sqrt = 0.3.
        [0.12, 0.12, 0.12],
    ])
    half = np.array([
        [3.3, 3.0, 3.0],      # |d| = 0.3 / 3 = 0.1; error 0.1 / 3.0001 = 0.0333
        [3.6, 3.0, 3.0],      # 0.2 / 3.0001 = 0.0667
        [0.16, 0.12, 0.12],   # 0.04 / 4 = 0.01; 0.01 / 0.3001 = 0.0333 (0.111 if divided by 0.09)
        [0.28, 0.12, 0.12],   # 0.16 / 4 = 0.04; 0.04 / 0.3001 = 0.1333
    ])
    count = np.array([3, 3, 4, 4], dtype=np.uint32)
    flags = lambda px: px[29 * n:30 * n].copy()  # noqa: E731
    # Threshold 0.05 at a step boundary past the minimum: pixels 0 and 2 settle.
    px = adapt_buffer(w, h, full, half, count)
    check(orc, adapt_camera(w, h, 7, 0.05, 4, 4), px)
    assert flags(px).tolist() == [1, 0, 1, 0]
    # Threshold 0.1: the 0.0667 pixel settles too; 0.1333 does not.
    px = adapt_buffer(w, h, full, half, count)
    check(orc, adapt_camera(w, h, 7, 0.1, 4, 4), px)
    assert flags(px).tolist() == [1, 1, 1, 0]
    # Not a step's last sample (7 & 7 != 7 for step 8 is 7: boundary; 6 is not), or not past the minimum.
    for frame, mn, step in ((6, 4, 4), (7, 7, 4), (3, 4, 4)):
        px = adapt_buffer(w, h, full, half, count)
        check(orc, adapt_camera(w, h, frame, 1.0, mn, step), px)
        assert flags(px).tolist() == [0, 0, 0, 0], (frame, mn, step)
    # A settled pixel keeps its flag whatever its error is now.
    px = adapt_buffer(w, h, full, half, count)
    px[29 * n + 3] = 1
    check(orc, adapt_camera(w, h, 7, 0.05, 4, 4), px)
    assert flags(px).tolist() == [1, 0, 1, 1]


def test_adaptive_dilation_by_hand(orc):
    w, h = 4, 3
    n = w * h
    px = np.zeros(P.PX_LEN * n + 1, dtype=np.uint32)
    settled = np.array([[1, 1, 1, 0],
                        [1, 1, 1, 1],
                        [1, 1, 1, 1]], dtype=np.uint32)
    px[29 * n:30 * n] = settled.reshape(-1)
    dilate(orc, adapt_camera(w, h, 7, 0.05, 4, 4), px)
    seen = px[30 * n:31 * n].reshape(h, w)
    want = np.array([[1, 1, 0, 0],
                     [1, 1, 1, 0],
                     [1, 1, 1, 1]], dtype=np.uint32)
    assert (seen == want).all()

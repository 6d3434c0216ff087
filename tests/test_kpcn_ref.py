"""CPU checks of the learned denoiser's float32 reference and host logic: what
the mathematics fixes about the filter step, the layer table the library
reports, the weight layout, and the crop camera."""
import numpy as np
import pytest
import torch

import phiray_amd as P
from phiray_amd import kpcn as K
from phiray_amd import scenes as S
from tests import kpcn_ref as R


def _layers():
    return P.Kpcn.layers()


def test_layer_table_and_size():
    L = _layers()
    assert L.tolist() == [[16, 32, 0, 1], [32, 32, 0, 1], [32, 48, 1, 1], [48, 48, 1, 1], [48, 64, 2, 1],
                          [64, 64, 2, 1], [64, 96, 3, 1], [160, 64, 2, 1], [112, 48, 1, 1], [80, 32, 0, 1],
                          [32, 32, 0, 0]]
    assert P.Kpcn.param_count() == sum(9 * co * ci + co for ci, co, _, _ in L.tolist())
    # Skip concatenations: each decoder's input is the coarser output plus the skip.
    assert L[7][0] == L[6][1] + L[5][1] and L[8][0] == L[7][1] + L[3][1] and L[9][0] == L[8][1] + L[1][1]
    used = P.Kpcn.param_count() - 9 * 32 * 6 - (9 * 7 * 32 + 7)
    assert used == 337801  # about 0.3 M trainable weights


def test_init_keeps_padding_zero():
    p = K.init_params(3)
    sl = K.layer_slices()
    w0 = p[sl[0][0]].reshape(9, 32, 16)
    assert np.all(w0[:, :, 10:] == 0) and np.any(w0[:, :, :10] != 0)
    wl = p[sl[-1][0]].reshape(9, 32, 32)
    assert np.all(wl[:, 25:, :] == 0) and np.all(p[sl[-1][1]] == 0)


def test_unpack_matches_conv_tap_order():
    """Flat W[tap][co][ci] with tap = 3 (dy + 1) + dx + 1 is the correlation
    weight at (ky, kx) = (dy + 1, dx + 1): a single one-hot weight moves the
    image by (dy, dx)."""
    L = _layers()
    p = np.zeros(P.Kpcn.param_count(), dtype=np.float32)
    sl = K.layer_slices(L)
    w = np.zeros((9, 32, 16), dtype=np.float32)
    dy, dx = 1, -1
    w[3 * (dy + 1) + dx + 1, 0, 0] = 1.0
    p[sl[0][0]] = w.ravel()
    W, _, _ = R.unpack(p, L)[0]
    x = torch.arange(5 * 7, dtype=torch.float32).reshape(1, 1, 5, 7)
    x16 = torch.cat([x, torch.zeros(1, 15, 5, 7)], 1)
    y = torch.nn.functional.conv2d(x16, W, padding=1)[0, 0]
    assert y[1, 2] == x[0, 0, 1 + dy, 2 + dx]


def _output_bias(center: float) -> np.ndarray:
    p = np.zeros(P.Kpcn.param_count(), dtype=np.float32)
    sl = K.layer_slices()
    b = np.zeros(32, dtype=np.float32)
    b[12] = center  # tap (0, 0) of the 5x5 filter
    p[sl[-1][1]] = b
    return p


def test_identity_filter_returns_radiance():
    """With all weights zero and a dominant centre logit the filter is the
    identity, so demodulation and remodulation must cancel exactly enough."""
    fr, _ = R.synthetic_frames(1, 12, 20, 0)
    y = R.forward(_output_bias(60.0), _layers(), fr, bf16=False).numpy()
    np.testing.assert_allclose(y, fr[..., 0:3], rtol=1e-5, atol=1e-7)


def test_flat_logits_box_filter():
    """Equal logits are a 5x5 box over the demodulated radiance with edge
    clamping: check one interior and one corner pixel by hand."""
    fr, _ = R.synthetic_frames(1, 9, 11, 1)
    y = R.forward(_output_bias(0.0), _layers(), fr, bf16=False).numpy()[0]
    ad = np.maximum(fr[0, ..., 3:6], 0.05)
    d = fr[0, ..., 0:3] / ad
    want = ad[4, 5] * d[2:7, 3:8].reshape(-1, 3).mean(0)
    np.testing.assert_allclose(y[4, 5], want, rtol=1e-5)
    rows = np.clip(np.arange(-2, 3), 0, 8)
    cols = np.clip(np.arange(-2, 3), 0, 10)
    want0 = ad[0, 0] * d[np.ix_(rows, cols)].reshape(-1, 3).mean(0)
    np.testing.assert_allclose(y[0, 0], want0, rtol=1e-5)


def test_constant_image_is_a_fixed_point():
    """Any normalised filter leaves a constant demodulated image unchanged."""
    fr, _ = R.synthetic_frames(2, 16, 24, 2)
    fr[..., 3:6] = 0.5
    fr[..., 0:3] = 0.25
    y = R.forward(K.init_params(5), _layers(), fr, bf16=False).numpy()
    np.testing.assert_allclose(y, 0.25, rtol=1e-5)


def test_reference_gradient_by_finite_difference():
    """Autograd of the FP32 reference against a central difference in a few
    weights of different layers (float64 copies, so the check is tight)."""
    fr, tg = R.synthetic_frames(1, 16, 16, 4)
    p = K.init_params(7).astype(np.float64)
    L = _layers()
    torch.set_default_dtype(torch.float64)
    try:
        loss, g = R.grads(p, L, fr.astype(np.float64), tg.astype(np.float64), bf16=False)
        sl = K.layer_slices(L)
        rng = np.random.default_rng(0)
        for li in (0, 4, 6, 7, 10):
            ws = sl[li][0]
            cout, cin = sl[li][2]
            used = 25 if li == 10 else cout
            for _ in range(2):
                tap, co, ci = rng.integers(0, 9), rng.integers(0, used), rng.integers(0, 10 if li == 0 else cin)
                j = ws.start + (tap * cout + co) * cin + ci
                h = 1e-6
                pp, pm = p.copy(), p.copy()
                pp[j] += h
                pm[j] -= h
                lp = float(R.rel_mse(R.forward(pp, L, fr.astype(np.float64), bf16=False),
                                     torch.tensor(tg.astype(np.float64))))
                lm = float(R.rel_mse(R.forward(pm, L, fr.astype(np.float64), bf16=False),
                                     torch.tensor(tg.astype(np.float64))))
                fd = (lp - lm) / (2 * h)
                assert abs(fd - g[j]) <= 1e-6 * max(1.0, abs(fd)) + 1e-4 * abs(fd), (li, fd, g[j])
    finally:
        torch.set_default_dtype(torch.float32)


def _primary_dirs(cam, w, h, ju=0.3, jv=0.7):
    """The path kernel's primary-ray directions at a fixed jitter, for every
    pixel (the camera-ray step of src/trace/wgsl/shade.wgsl:65-95)."""
    c = cam[0]
    x, y = np.meshgrid(np.arange(w), np.arange(h))
    su = (x + 0.5) / w - 0.5
    sv = (y + 0.5) / h - 0.5
    jx, jy = (ju - 0.5) / w, (jv - 0.5) / h
    vx = (su + jx - c["pan"][0]) / c["scale"]
    vy = -((sv + jy) / c["aspect"] - c["pan"][1]) / c["scale"]
    d = c["right"] * vx[..., None] + c["up"] * vy[..., None] - c["back"] * c["dist"]
    return d / np.linalg.norm(d, axis=-1, keepdims=True)


@pytest.mark.parametrize("x0,y0,cw,ch", [(0, 0, 64, 64), (700, 300, 512, 512), (1408, 568, 512, 512), (10, 20, 96, 48)])
def test_crop_camera_is_a_window_of_the_full_frame(x0, y0, cw, ch):
    full = _primary_dirs(S.path_camera(1920, 1080), 1920, 1080)
    crop = _primary_dirs(K.crop_camera(1920, 1080, x0, y0, cw, ch), cw, ch)
    np.testing.assert_allclose(crop, full[y0:y0 + ch, x0:x0 + cw], atol=2e-6)

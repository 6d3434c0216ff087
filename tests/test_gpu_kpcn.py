"""The learned denoiser's HIP kernels against the PyTorch float32 reference
(tests/kpcn_ref.py), through the C ABI.

Tolerances. The HIP path stores activations and activation gradients in BF16
(8-bit mantissa, relative rounding 2^-9) and accumulates in FP32. The
reference with `bf16=True` rounds at the same points, so the two differ by
accumulation order plus the rare BF16 rounding flip that order causes. That
is 1e-2 relative on the filtered radiance (a flip in a logit moves a filter
weight by ~0.4%) and 5e-2 on a layer's weight gradient (Frobenius norm), where
gradients also pass through BF16 on the way back. Against the plain FP32
network the bound on the output is the BF16 error itself: relMSE between
the two outputs below 1e-3.
"""
import numpy as np
import pytest
import torch

import phiray_amd as P
from phiray_amd import kpcn as K
from tests import kpcn_ref as R

pytestmark = pytest.mark.gpu

B, H, W = 2, 37, 53  # ragged: the network pads to 40 x 56


def _params(seed):
    p = K.init_params(seed)
    rng = np.random.default_rng(seed + 100)
    for ws, bs, (cout, cin) in K.layer_slices():
        p[bs] = rng.normal(0.0, 0.05, size=cout).astype(np.float32)
    # keep the padded output channels zero (their logits are never read)
    p[K.layer_slices()[-1][1]][25:] = 0.0
    return p


def _upload(gpu, a):
    return P.GpuArray.from_(gpu, np.ascontiguousarray(a, dtype=np.float32).ravel())


def test_forward_matches_reference(gpu):
    fr, tg = R.synthetic_frames(B, H, W, 11)
    p = _params(1)
    net = P.Kpcn(gpu, p)
    out = P.GpuArray.zeroed(gpu, np.float32, B * H * W * 3)
    print(f"[kpcn] forward B={B} {W}x{H}", flush=True)
    net.infer(_upload(gpu, fr), B, W, H, out)
    y = out.read().reshape(B, H, W, 3)
    assert np.all(np.isfinite(y))
    L = P.Kpcn.layers()
    ref = R.forward(p, L, fr, bf16=True).numpy()
    rel = np.abs(y - ref) / (np.abs(ref) + 1e-2)
    assert np.median(rel) < 1e-3 and rel.max() < 1e-2, (np.median(rel), rel.max())
    ref32 = R.forward(p, L, fr, bf16=False).numpy()
    assert float(np.mean((y - ref32) ** 2 / (ref32 ** 2 + 0.01))) < 1e-3


def test_loss_and_output(gpu):
    fr, tg = R.synthetic_frames(B, H, W, 12)
    p = _params(2)
    net = P.Kpcn(gpu, p)
    out = P.GpuArray.zeroed(gpu, np.float32, B * H * W * 3)
    loss = net.loss(_upload(gpu, fr), _upload(gpu, tg), B, W, H, out)
    ref = R.forward(p, P.Kpcn.layers(), fr, bf16=True)
    want = float(R.rel_mse(ref, torch.tensor(tg)))
    assert abs(loss - want) <= 1e-3 * want, (loss, want)
    y = out.read().reshape(B, H, W, 3)
    assert abs(float(np.mean((y - tg) ** 2 / (tg ** 2 + 0.01))) - loss) <= 1e-5 * loss


def test_gradients_match_autograd(gpu):
    fr, tg = R.synthetic_frames(B, H, W, 13)
    p = _params(3)
    net = P.Kpcn(gpu, p)
    print("[kpcn] backward", flush=True)
    loss = net.backward(_upload(gpu, fr), _upload(gpu, tg), B, W, H)
    g = net.grads()
    assert np.all(np.isfinite(g))
    want_loss, want = R.grads(p, P.Kpcn.layers(), fr, tg, bf16=True)
    assert abs(loss - want_loss) <= 1e-3 * want_loss
    for li, (ws, bs, (cout, cin)) in enumerate(K.layer_slices()):
        for part, sl in (("W", ws), ("b", bs)):
            a, b = g[sl], want[sl]
            nb = np.linalg.norm(b)
            err = np.linalg.norm(a - b) / max(nb, 1e-30)
            assert err < 5e-2, (li, part, err, nb)
    # padded weights get no gradient
    w0 = g[K.layer_slices()[0][0]].reshape(9, 32, 16)
    assert np.all(w0[:, :, 10:] == 0)
    wl = g[K.layer_slices()[-1][0]].reshape(9, 32, 32)
    assert np.all(wl[:, 25:, :] == 0)


def test_adam_step(gpu):
    """One Adam step from zero moments, restated in float32 with the kernel's
    operation order, against the gradient the device left."""
    fr, tg = R.synthetic_frames(1, 24, 32, 14)
    p = _params(4)
    net = P.Kpcn(gpu, p)
    net.backward(_upload(gpu, fr), _upload(gpu, tg), 1, 32, 24)
    g = net.grads()
    f = np.float32
    lr, b1, b2, eps = f(1e-3), f(0.9), f(0.999), f(1e-8)
    m = (f(1) - b1) * g
    v = ((f(1) - b2) * g) * g
    bc1, bc2 = f(1.0 - float(b1)), f(1.0 - float(b2))  # 1 - beta^t in double, as the host does
    want = p - (lr * (m / bc1)) / (np.sqrt(v / bc2) + eps)
    net.adam(float(lr))
    got = net.params()
    np.testing.assert_allclose(got, want, rtol=2e-7, atol=0)


def test_indexed_step_equals_plain_step(gpu):
    """Gathering images from resident sets gives the same step, bit for bit,
    as handing over the batch itself."""
    fr, tg = R.synthetic_frames(4, 16, 24, 15)
    p = _params(5)
    a, b = P.Kpcn(gpu, p), P.Kpcn(gpu, p)
    fidx, tidx = np.array([3, 0, 2]), np.array([1, 1, 0])
    la = a.train_indexed(_upload(gpu, fr), fidx, _upload(gpu, tg), tidx, 24, 16, 1e-3)
    lb = b.train_step(_upload(gpu, fr[fidx]), _upload(gpu, tg[tidx]), 3, 24, 16, 1e-3)
    assert la == lb
    assert np.array_equal(a.params(), b.params())
    with pytest.raises(P.PhrError):
        a.train_indexed(_upload(gpu, fr), [4], _upload(gpu, tg), [0], 24, 16, 1e-3)


def test_training_lowers_the_loss(gpu):
    fr, tg = R.synthetic_frames(4, 32, 32, 16)
    net = P.Kpcn(gpu, _params(6))
    f, t = _upload(gpu, fr), _upload(gpu, tg)
    first = net.loss(f, t, 4, 32, 32)
    for _ in range(40):
        net.train_step(f, t, 4, 32, 32, 1e-3, want_loss=False)
    assert net.loss(f, t, 4, 32, 32) < 0.7 * first


def test_tracer_frame_and_learned_denoise(gpu):
    from phiray_amd import scenes as S

    tris = S.spheres_scene(spheres=8, level=2, seed=3)
    tr = P.Tracer.new(gpu, tris, 4)
    w, h = 48, 40
    cam = S.path_camera(w, h, rr_from=3)
    tr.sample(cam, w, h, 2)
    frame = P.GpuArray.zeroed(gpu, np.float32, w * h * 10 + 7)
    tr.frame_into(frame, 7)
    f = frame.read()[7:].reshape(h * w, 10)
    assert np.array_equal(f[:, 0:3], tr.radiance())
    assert np.array_equal(f[:, 3:6], tr.albedo())
    assert np.array_equal(f[:, 6:9], tr.normal())
    assert np.array_equal(f[:, 9], tr.depth_plane())
    net = P.Kpcn(gpu, _params(7))
    out = P.GpuArray.zeroed(gpu, np.float32, w * h * 3)
    tr.denoise_learned(net, out)
    via_net = P.GpuArray.zeroed(gpu, np.float32, w * h * 3)
    net.infer(frame_only := _upload(gpu, f), 1, w, h, via_net)
    assert np.array_equal(out.read(), via_net.read())
    tr.denoise_learned(net)  # in place of the linear sums
    np.testing.assert_allclose(tr.radiance().ravel(), out.read(), rtol=1e-6, atol=1e-7)
    # The display is refreshed from the new sums as the tracer composes it:
    # the Standard curve (float64 here) of the light over the samples that
    # met something, under the sky's share.
    pl = P.planes(tr.px(), w, h)
    k = pl["count"].astype(np.float64)
    c = np.clip(pl["radiance"] / np.maximum(k - pl["missed"], 1.0)[:, None], 0.0, 1.0)
    d = np.where(c <= 0.0031308, c * 12.92, 1.055 * c ** (1 / 2.4) - 0.055)
    want = d * (1.0 - pl["cov"] / k)[:, None] + pl["grid"] / k[:, None]
    np.testing.assert_allclose(pl["display"], want, rtol=1e-5, atol=1e-6)
    q = np.floor(np.clip(want, 0, 1) * 255.0 + 0.5)
    assert (np.abs(pl["pixel"][:, :3].astype(np.int32) - q) <= 1).all()
    del frame_only


def test_recorded_frame_with_the_learned_denoiser(gpu):
    from phiray_amd import scenes as S

    tris = S.spheres_scene(spheres=6, level=2, seed=5)
    w, h = 56, 40
    cam = S.path_camera(w, h, rr_from=3)
    tr = P.Tracer.new(gpu, tris, 4)
    net = P.Kpcn(gpu, _params(8))
    tr.sample(cam, w, h, 2)
    tr.denoise_learned(net)
    want = tr.radiance()
    tr.record_frame(cam, w, h, 2, denoise=2, kpcn=net)
    for _ in range(3):
        tr.sample(cam, w, h, 1)
        tr.replay_frame(net)
        assert np.array_equal(tr.radiance(), want)


def test_autotuned_tiles_give_the_same_output(gpu):
    fr, _ = R.synthetic_frames(1, 72, 96, 17)
    net = P.Kpcn(gpu, _params(9))
    f = _upload(gpu, fr)
    a = P.GpuArray.zeroed(gpu, np.float32, 72 * 96 * 3)
    b = P.GpuArray.zeroed(gpu, np.float32, 72 * 96 * 3)
    net.infer(f, 1, 96, 72, a)
    rows = net.autotune(f, 1, 96, 72, reps=1)
    assert set(rows.tolist()) <= {1, 2}
    net.infer(f, 1, 96, 72, b)
    assert np.array_equal(a.read(), b.read())

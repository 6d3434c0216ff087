"""The learned denoiser at its own sizes on real tracer frames (the
65,536-triangle scenes of the LBVH workload, albedo U[0.2, 0.8]^3, under a
white world):

* inference on a 1920x1080 4-sample frame against the PyTorch reference
  (tests/kpcn_ref.py, BF16 rounding at the same points): median relative
  error below 1e-3, 99.9% of values within 1e-2 and all within 5e-2 (over
  six million values, accumulation order flips the BF16 rounding of a few
  activations, and a flip in a logit moves a filter weight by ~0.4%), and
  relMSE to the plain FP32 network below 1e-3;
* one training step over a batch of 8 crops of 512 x 512 (4-sample inputs
  against 64-sample targets): the loss within 1e-3 relative and every
  layer's weight and bias gradient within 5e-2 (Frobenius) of autograd's.
"""
import numpy as np
import pytest

import phiray_amd as P
from phiray_amd import kpcn as K
from phiray_amd import scenes as S
from tests import kpcn_ref as R

pytestmark = pytest.mark.gpu


def _weights():
    return K.load(K.WEIGHTS) if K.WEIGHTS.exists() else K.init_params(5)


def _frame(gpu, tr, w, h, spp, cam):
    tr.sample(cam, w, h, spp)
    f = P.GpuArray.zeroed(gpu, np.float32, w * h * P.Kpcn.FRAME)
    tr.frame_into(f, 0)
    return f


def test_inference_on_a_1080p_tracer_frame(gpu):
    w, h = 1920, 1080
    print(f"[config5] {w}x{h} 4-sample frame, KPCN forward against the reference", flush=True)
    tr = P.Tracer.new(gpu, S.cloud_scene(2000), 4)
    frame = _frame(gpu, tr, w, h, 4, S.cloud_camera(w, h, rr_from=3))
    p = _weights()
    net = P.Kpcn(gpu, p)
    out = P.GpuArray.zeroed(gpu, np.float32, w * h * 3)
    net.infer(frame, 1, w, h, out)
    y = out.read().reshape(1, h, w, 3)
    assert np.all(np.isfinite(y))
    fr = frame.read().reshape(1, h, w, P.Kpcn.FRAME)
    assert (fr[..., 0:3] > 0).any(axis=-1).mean() > 0.2, "the frame is mostly dark"
    L = P.Kpcn.layers()
    ref = R.forward(p, L, fr, bf16=True).numpy()
    rel = np.abs(y - ref) / (np.abs(ref) + 1e-2)
    assert np.median(rel) < 1e-3 and np.quantile(rel, 0.999) < 1e-2 and rel.max() < 5e-2, (
        np.median(rel), np.quantile(rel, 0.999), rel.max())
    ref32 = R.forward(p, L, fr, bf16=False).numpy()
    assert float(np.mean((y - ref32) ** 2 / (ref32 ** 2 + 0.01))) < 1e-3


def test_a_training_step_on_eight_512_crops(gpu):
    c, b = 512, 8
    print(f"[config5] {b} x {c}^2 crops, one backward against autograd", flush=True)
    frames = P.GpuArray.zeroed(gpu, np.float32, b * c * c * P.Kpcn.FRAME)
    targets = np.zeros((b, c, c, 3), dtype=np.float32)
    for k in range(b):
        tr = P.Tracer.new(gpu, S.cloud_scene(3000 + k), 4)
        tr.set_samples(1 << 20)
        x0, y0 = 96 * k % (1920 - c), 61 * k % (1080 - c)
        cam = K.crop_camera(1920, 1080, x0, y0, c, c, rr_from=3)
        cam["env"] = (1.0, 1.0, 1.0, 0.0)
        ref = cam.copy()
        ref["stream"] = 7
        tr.sample(ref, c, c, 64)
        targets[k] = tr.radiance().reshape(c, c, 3)
        tr.sample(cam, c, c, 4)
        tr.frame_into(frames, k * c * c * P.Kpcn.FRAME)
        tr.free()
    p = _weights()
    net = P.Kpcn(gpu, p)
    loss = net.backward(frames, P.GpuArray.from_(gpu, targets.ravel()), b, c, c)
    g = net.grads()
    assert np.all(np.isfinite(g))
    fr = frames.read().reshape(b, c, c, P.Kpcn.FRAME)
    want_loss, want = R.grads(p, P.Kpcn.layers(), fr, targets, bf16=True)
    assert abs(loss - want_loss) <= 1e-3 * want_loss, (loss, want_loss)
    for li, (ws, bs, _) in enumerate(K.layer_slices()):
        for part, sl in (("W", ws), ("b", bs)):
            a, e = g[sl], want[sl]
            nb = np.linalg.norm(e)
            if nb == 0.0:
                continue
            err = np.linalg.norm(a - e) / nb
            assert err < 5e-2, (li, part, err, nb)

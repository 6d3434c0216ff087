"""The learned frame's held-out check (phiray_amd.kpcn.LearnedFrame,
Kpcn.held_out) on the device.

* The estimate is what it claims: on a 256 x 256 frame of the learned
  denoiser's scenes, the estimated relMSE of the network's output from half A
  equals its measured relMSE against a 2,048-sample render (same D^2 + 0.01
  scale) within 20% + 1e-3: E[(D - B)^2] - E[(A - B)^2] / 2 = E[(D - T)^2]
  for independent halves of equal sample counts. The answer's estimate (the
  output from both halves: half A's error common to both halves, half of
  its noise-driven part, E - spread / 4) lies between the answer's measured
  relMSE and half A's estimate: it does not understate the answer's error
  (the common part shrinks with more samples too, so it overstates it).
* A weight set that shifts the image by two pixels (every filter on the
  corner tap) is caught: the check asks for more samples until the budget
  (checks at 4 and at the budget's 32 samples), and the answer is then the
  plain render of every sample, exactly.
* A frame the filter cannot harm (a flat wall under a white world, whose
  radiance has no noise) passes the first check at the samples asked for.
"""
import numpy as np
import pytest

import phiray_amd as P
from phiray_amd import kpcn as K
from phiray_amd import scenes as S

pytestmark = pytest.mark.gpu

W = H = 256


def _weights():
    return K.load(K.WEIGHTS) if K.WEIGHTS.exists() else K.init_params(5)


def _measured(gpu, tris, cam, d):
    ref = P.Tracer.new(gpu, tris, 4)
    ref.set_samples(1 << 20)
    c = cam.copy()
    c["stream"] = K.REF_STREAM
    for _ in range(32):
        ref.sample(c, W, H, 64)
    t = ref.radiance().astype(np.float64)
    d = d.astype(np.float64)
    return float(np.mean((d - t) ** 2 / (d * d + 0.01))), t


def test_the_estimate_matches_the_measured_error(gpu):
    print("[guard] estimate against a 2,048-sample render", flush=True)
    tris = S.cloud_scene(4000)
    cam = S.cloud_camera(W, H, rr_from=3)
    lf = K.LearnedFrame(gpu, tris, P.Kpcn(gpu, _weights()))
    res = lf.render(cam, W, H, spp=4, bound=1e9)  # one check, whatever it says
    est = res["checks"][0]["estimate"]
    d = lf.da.read().reshape(-1, 3)
    meas, t = _measured(gpu, tris, cam, d)
    assert abs(est - meas) <= 0.2 * meas + 1e-3, (est, meas)
    assert res["checks"][0]["half_noise"] > 0.0
    out = lf.out.read().reshape(-1, 3).astype(np.float64)
    meas_out = float(np.mean((out - t) ** 2 / (out * out + 0.01)))
    est_out = res["checks"][0]["answer_estimate"]
    print(f"[guard] half A: estimate {est:.5f} measured {meas:.5f}; answer: estimate {est_out:.5f} measured {meas_out:.5f}",
          flush=True)
    assert res["checks"][0]["spread"] > 0.0
    assert meas_out - 1e-3 <= est_out < est, (est_out, meas_out, est)


def _shifting_weights():
    p = K.init_params(1)
    ws, bs, (cout, cin) = K.layer_slices()[-1]
    p[ws] = 0.0
    b = np.zeros(cout, dtype=np.float32)
    b[0] = 30.0  # the (-2, -2) tap takes all the weight
    p[bs] = b
    return p


def test_a_bad_network_falls_back_to_the_plain_render(gpu):
    print("[guard] shifting weights must fall back", flush=True)
    tris = S.cloud_scene(4001)
    cam = S.cloud_camera(W, H, rr_from=3)
    lf = K.LearnedFrame(gpu, tris, P.Kpcn(gpu, _shifting_weights()))
    res = lf.render(cam, W, H, spp=4, bound=0.01, max_spp=32)
    assert res["plain"], res
    assert res["spp"] == 32 and [c["spp"] for c in res["checks"]] == [4, 32], res
    assert all(c["estimate"] > 0.01 and c["answer_estimate"] > 0.01 for c in res["checks"])
    plain = 0.5 * (lf.a.radiance() + lf.b.radiance())
    assert np.array_equal(lf.out.read().reshape(-1, 3), plain.astype(np.float32))


def _wall():
    t = np.zeros(2, dtype=P.TRI)
    q = np.array([[-5, -5, 1], [5, -5, 1], [5, 5, 1], [-5, 5, 1]], dtype=np.float32)
    t["a"], t["b"], t["c"] = q[[0, 0]], q[[1, 2]], q[[2, 3]]
    t["rough"], t["ior"] = 1.0, 1.0
    t["color"] = (0.5, 0.5, 0.5, 1.0)
    return t


def test_a_noiseless_frame_passes_at_once(gpu):
    print("[guard] a flat wall passes the first check", flush=True)
    cam = S.path_camera(W, H, eye=(0.0, 0.0, -2.0), target=(0.0, 0.0, 1.0), up=(0.0, 1.0, 0.0), rr_from=3)
    cam["env"] = (1.0, 1.0, 1.0, 0.0)
    lf = K.LearnedFrame(gpu, _wall(), P.Kpcn(gpu, K.init_params(2)))
    res = lf.render(cam, W, H, spp=4, bound=0.01)
    assert not res["plain"] and res["spp"] == 4 and len(res["checks"]) == 1, res
    assert abs(res["checks"][0]["estimate"]) < 1e-4
    out = lf.out.read().reshape(-1, 3)
    assert np.allclose(out, 0.5, rtol=1e-3)

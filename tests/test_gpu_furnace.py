"""The device tracer on the oracle's pinned multi-bounce scenes (GPU).

The integrating sphere (closed form after K bounces, roulette off and from
the first bounce) and the reference's trench under a uniform world
(tests/world_power.rs) are the cases `test_oracle_furnace.py` pins the
oracle with; here the HIP path tracer renders them with the same sample
streams and every image-buffer plane must equal the oracle's within 1e-5
relative. Together: the device integrator lands on the closed forms too.
"""
import numpy as np
import pytest

import phiray_amd as P
from test_gpu_trace import compare_planes
from test_oracle_furnace import closed_form, integrating_sphere, overhead, sphere_camera, trench

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("bounces,rr_from", [(4, 1000), (24, 1)])
def test_integrating_sphere_matches_the_twin_and_the_closed_form(gpu, orc, bounces, rr_from):
    print(f"[furnace] integrating sphere bounces={bounces} rr_from={rr_from}", flush=True)
    tris, f = integrating_sphere()
    w, spp = 32, 16
    cam = sphere_camera(w, bounces, rr_from)
    sc = orc.Scene(tris, 4, P.NODE, 8)
    px = np.zeros(P.PX_LEN * w * w + 1, dtype=np.uint32)
    for k in range(spp):
        sc.sample(cam, w, w, px, k, 8)
    tr = P.Tracer.new(gpu, tris, 4)
    tr.sample(cam, w, w, spp)
    compare_planes(tr.px(), px, w, w)
    rad = P.planes(tr.px(), w, w)["radiance"][:, 0].astype(np.float64) / spp
    want = closed_form(0.8, f, bounces)
    assert abs(rad.mean() - want) < 0.04 * want


@pytest.mark.parametrize("bounces", [1, 4])
def test_trench_matches_the_twin(gpu, orc, bounces):
    print(f"[furnace] trench bounces={bounces}", flush=True)
    tris = trench()
    w, spp = 32, 8
    cam = overhead(1.0, bounces)
    sc = orc.Scene(tris, 4, P.NODE, 8)
    px = np.zeros(P.PX_LEN * w * w + 1, dtype=np.uint32)
    for k in range(spp):
        sc.sample(cam, w, w, px, k, 8)
    tr = P.Tracer.new(gpu, tris, 4)
    tr.sample(cam, w, w, spp)
    compare_planes(tr.px(), px, w, w)

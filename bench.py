"""Benchmark: Rendered-mode path tracing on one MI355X, with the LBVH beside it.

One step = one 64-spp frame of the 1.31M-triangle spheres scene at 1920x1080
(max 4 diffuse bounces, Russian roulette after bounce 3, next-event
estimation, FP32), traced by the HIP path kernel over the device-built LBVH.
`value` is rays per second (closest-hit + shadow rays, counted on the device).
The same line carries the LBVH build and refit rates (Mtri/s) on that scene and
on the 65,536-triangle set, the 256x256 closest-hit raycast rate, the path
kernel's roofline figures, and the CPU twin (the oracle) timed on the host's
cores on a bounded sample of the same frame.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--spp 64]
"""
from __future__ import annotations

import argparse
import json
import os
import pathlib
import re
import sys
import time

import numpy as np

ROOT = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

import phiray_amd as P  # noqa: E402
from phiray_amd import scenes as S  # noqa: E402

METRIC = "Mrays/s + LBVH Mtri/s; s/frame to relMSE≤0.01: learned vs plain, 1 GPU"
HBM_PEAK_GBS = 8000.0

# Algorithmic bytes of the walk (one ray): the ray in (32), the hit out (8),
# the root node (32), both children at every inner step (2 x 32), one
# triangle record per test (48: three corners in 16-byte rows).
RAY_BYTES, HIT_BYTES, ROOT_BYTES, STEP_BYTES, TEST_BYTES = 32, 8, 32, 64, 48
PATH_BUILDS = ("nested", "regen", "wavefront", "split")  # the tracer's path_regen modes 0-3


def profiled_traffic(name: str, spp: int, w: int, h: int, waves: int, build: str) -> tuple[float | None, str | None]:
    """HBM bytes of one k_path launch from the committed counter summary, when
    it was measured on this same workload and kernel variant (FETCH_SIZE +
    WRITE_SIZE, each scaled by what the counter reads for a copy of known size)."""
    f = ROOT / "profiles" / name
    if not f.exists():
        return None, None
    text = f.read_text()
    m = re.search(r"traffic_bytes_per_launch:\s*(\d+)", text)
    mw = re.search(r"path_waves_per_simd:\s*(\d+)", text)
    mb = re.search(r"path_build:\s*(\w+)", text)
    if not m or not mw or int(mw.group(1)) != waves or (spp, w, h) != (64, 1920, 1080):
        return None, None
    if (mb.group(1) if mb else "nested") != build:
        return None, None
    return float(m.group(1)), f"profiles/{name}"


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def lbvh_rates(gpu: P.Gpu, boxes: np.ndarray, reps: int) -> dict:
    prims = P.GpuArray.from_(gpu, boxes)
    b = P.GpuBvh.build_on(gpu, prims, 4)
    tile, tile_ms = b.autotune(5)  # the sort tile this size runs fastest with
    clock = P.Clock(gpu)
    for _ in range(3):
        b.rebuild_async()
        b.refit_async()
    gpu.wait()
    clock.start()
    for _ in range(reps):
        b.rebuild_async()
    t_build, _ = clock.stop()
    clock.start()
    for _ in range(reps):
        b.refit_async()
    t_refit, _ = clock.stop()
    n = len(boxes)
    tb, tr = t_build / reps, t_refit / reps
    nodes = b.download()
    reach = len(nodes.reachable())
    # Algorithmic bytes: bounds read 32, Morton read 32 + write 4, sort 4 passes
    # x (4+4 read, 4+4 write) + histogram read 4, Karras read 4, node topology
    # (2n-1) x 32, refit per primitive (order 4 + box 32) and per reachable
    # node (parent 4 + counter 8 + child 32 + node 32).
    refit_bytes = n * 36 + reach * (4 + 8 + 32 + 32)
    build_bytes = n * (32 + 36 + 68 + 4) + (2 * n - 1) * 32 + refit_bytes
    return {
        "tris": n,
        "depth": b.depth,
        "build_ms": tb * 1e3,
        "build_mtri_s": n / tb / 1e6,
        "build_gbs": build_bytes / tb / 1e9,
        "build_hbm_frac": build_bytes / tb / 1e9 / HBM_PEAK_GBS,
        "refit_ms": tr * 1e3,
        "refit_mtri_s": n / tr / 1e6,
        "refit_gbs": refit_bytes / tr / 1e9,
        "bytes_per_tri_build": build_bytes / n,
        "bytes_per_tri_refit": refit_bytes / n,
        "sort_tile_keys": 256 * tile,
        "build_ms_by_sort_tile": {"2048": tile_ms[0], "4096": tile_ms[1]},
    }


def batch_rates(gpu: P.Gpu, reps: int) -> dict:
    """The 256-scene batch: one build makes every scene's tree (segments)."""
    boxes, seg_lo = S.batch_scenes(256)
    prims = P.GpuArray.from_(gpu, boxes)
    b = P.GpuBvh.build_batch_on(gpu, prims, seg_lo, 4)
    tile, tile_ms = b.autotune(3)
    clock = P.Clock(gpu)
    for _ in range(2):
        b.rebuild_async()
        b.refit_async()
    gpu.wait()
    clock.start()
    for _ in range(reps):
        b.rebuild_async()
    t_build, _ = clock.stop()
    clock.start()
    for _ in range(reps):
        b.refit_async()
    t_refit, _ = clock.stop()
    n = len(boxes)
    tb, tr = t_build / reps, t_refit / reps
    # Algorithmic bytes as for one tree, plus two sort passes over the segment id
    # (8 read + 8 write + 4 segment id each) and the segment id written once (4).
    reach = n  # about one reachable node per box at leaf 4
    refit_bytes = n * 36 + reach * (4 + 8 + 32 + 32)
    build_bytes = n * (32 + 36 + 68 + 4 + 4 + 2 * 20) + (2 * n - len(seg_lo) + 1) * 32 + refit_bytes
    return {"scenes": len(seg_lo) - 1, "tris": n, "build_ms": tb * 1e3, "build_mtri_s": n / tb / 1e6,
            "build_gbs": build_bytes / tb / 1e9, "refit_ms": tr * 1e3, "refit_mtri_s": n / tr / 1e6,
            "max_depth": b.depth, "sort_tile_keys": 256 * tile, "build_ms_by_sort_tile": {"2048": tile_ms[0], "4096": tile_ms[1]}}


def batch_render_rates(gpu: P.Gpu, scenes: int = 256, w: int = 256, h: int = 256, spp: int = 16, reps: int = 3) -> dict:
    """The 256 scenes of uneven size (1k-200k triangles each) built by one
    segmented LBVH build and rendered at 256x256, 16 spp, 4 bounces each, in
    one launch; the whole pass (boxes, build, slot order, render) recorded as
    one hipGraph and replayed. The scenes hold no emitter: the world colour
    (1, 1, 1) lights them. Camera at (0.5, 0.5, -1.5) looking +z, 45 degrees."""
    tris, seg_lo = S.batch_tris(scenes)
    t0 = time.perf_counter()
    b = P.Batch(gpu, tris, seg_lo)
    setup_s = time.perf_counter() - t0
    del tris
    cam = S.path_camera(w, h, eye=(0.5, 0.5, -1.5), target=(0.5, 0.5, 0.5), up=(0.0, 1.0, 0.0), rr_from=3)
    cam["env"] = (1.0, 1.0, 1.0, 0.0)
    waves, grouped, _ = b.autotune(cam, w, h, 4, reps=1)
    log(f"batch render: kernel at {waves} waves a SIMD, grouped order {grouped}, {PATH_BUILDS[b.path_build]} build")
    build_s = sorted(b.rebuild() for _ in range(5))[2]
    b.render(cam, w, h, spp)
    c0 = b.counters()
    render_s = [b.render(cam, w, h, spp) for _ in range(reps)]
    c1 = b.counters()
    b.record(cam, w, h, spp, rebuild=True)
    b.replay()
    pass_s = [b.replay() for _ in range(reps)]
    rays = int((c1[0] - c0[0]) + (c1[1] - c0[1])) / reps
    steps, tests = int(c1[2] - c0[2]) / reps, int(c1[3] - c0[3]) / reps
    t_render = float(np.median(render_s))
    nbytes = rays * (RAY_BYTES + HIT_BYTES + ROOT_BYTES) + steps * STEP_BYTES + tests * TEST_BYTES
    n = int(seg_lo[-1])
    return {"scenes": scenes, "tris": n, "image": f"{w}x{h}", "spp": spp, "bounces": 4, "light": "world colour 1.0",
            "setup_s": setup_s, "path_waves_per_simd": waves, "grouped": grouped, "path_build": PATH_BUILDS[b.path_build],
            "split_shade_eighths": b.split_shade,
            "build_ms": build_s * 1e3, "build_mtri_s": n / build_s / 1e6,
            "render_ms": t_render * 1e3, "rays_per_pass": rays, "mrays_s": rays / t_render / 1e6,
            "render_gbs": nbytes / t_render / 1e9, "render_hbm_frac": nbytes / t_render / 1e9 / HBM_PEAK_GBS,
            "steps_per_ray": steps / max(rays, 1), "tests_per_ray": tests / max(rays, 1),
            "graph_pass_ms": float(np.median(pass_s)) * 1e3}


def forest_rates(gpu: P.Gpu, substeps: int = 8, w: int = 1280, h: int = 720, spp: int = 4) -> dict:
    """The rigid-body workload: 1,024 bodies, each its own mesh under one
    instance (box or icosphere s0..s5), a ground plane and a lamp. A substep
    refits every mesh tree over moved vertices (two resident poses in turn)
    and rebuilds the top tree over new matrices; then one frame is traced."""
    b, worlds = S.bodies_scene(1024)
    v, f, m, me, ins = b.arrays()
    t0 = time.perf_counter()
    tr = P.Tracer.from_scene(gpu, v, f, m, me, ins, 4)
    build_s = time.perf_counter() - t0
    v2 = v.copy()
    v2["p"][:, :3] *= np.float32(1.02)  # every body swells by 2% in its own space
    poses = [P.GpuArray.from_(gpu, v2), P.GpuArray.from_(gpu, v)]
    base = ins["world"].copy()
    turn = [S.world_of(S.affine(S.rotation((0.3, 0.2, 1.0), 0.01 * (k + 1)))) for k in range(substeps)]
    mats = []
    for k in range(substeps):
        wk = base.copy()
        r = turn[k].reshape(4, 4).T  # back to a math matrix
        for i in range(len(worlds)):
            wk[i] = S.world_of(r @ worlds[i])
        mats.append(wk)
    tr.repose_on(poses[0])
    tr.reinstance(mats[0])
    gpu.wait()
    clock = P.Clock(gpu)
    clock.start()
    for k in range(substeps):
        assert tr.repose_on(poses[k % 2]) and tr.reinstance(mats[k])
    g_s, wall_s = clock.stop()
    cam = S.path_camera(w, h, eye=(0.0, -16.0, 9.0), target=(0.0, 0.0, 1.5), rr_from=3)
    tr.set_samples(1 << 20)
    waves, _ = tr.autotune(cam, w, h, spp, reps=1)
    tr.sample(cam, w, h, spp)
    c0 = tr.counters()
    clock.start()
    tr.sample(cam, w, h, spp)
    t_frame, _ = clock.stop()
    c1 = tr.counters()
    rays = int(c1[0] - c0[0] + c1[1] - c0[1])
    chosen = tr.path_regen
    by_build = {}
    for build in (0, 3):  # each build the forest allows, at the chosen register fit
        tr.set_path_regen(build)
        tr.sample(S.path_camera(64, 64), 64, 64, 1)  # another view: the next call starts over
        tr.sample(cam, w, h, spp)
        by_build[PATH_BUILDS[build]] = tr.time()[1] * 1e3
    tr.set_path_regen(chosen)
    return {"bodies": len(worlds), "meshes": len(me), "tris": len(f), "build_s": build_s,
            "substep_ms": wall_s / substeps * 1e3, "substep_gpu_span_ms": g_s / substeps * 1e3,
            "frame": f"{w}x{h}, {spp} spp", "frame_ms": t_frame * 1e3, "mrays_s": rays / t_frame / 1e6,
            "rays_per_frame": rays, "path_waves_per_simd": waves, "path_build": PATH_BUILDS[chosen],
            "frame_ms_by_build": by_build, "mesh_stack_entries": tr.path_stack}


def xpbd_rates(gpu: P.Gpu, frames: int = 120, substeps: int = 8, w: int = 1280, h: int = 720, spp: int = 4) -> dict:
    """The rigid-body workload driven by the device XPBD World: 1,024 bodies
    (boxes, and icospheres colliding as spheres) dropped from heights
    U[1, 10] m onto a plane, 120 frames. Every substep of a 1/60 s frame
    steps the world, refits every mesh tree and rebuilds the top tree over
    the new poses; every frame is traced at w x h, spp samples."""
    heights = (1.0, 10.0)
    b, _ = S.bodies_scene(1024, heights=heights)
    v, f, m, me, ins = b.arrays()
    tr = P.Tracer.from_scene(gpu, v, f, m, me, ins, 4)
    world = P.World(gpu)
    world.configure(substeps=1, iterations=10)
    g = world.add_box((12.0, 12.0, 0.5))
    world.add_body(kind=P.PASSIVE, shape=g, x=(0.0, 0.0, -0.5), friction=1.0)
    for k, d in enumerate(S.bodies_defs(1024, heights=heights)):
        r = float(d["r"])
        s = world.add_box((r, r, r)) if d["shape"] == "box" else world.add_sphere(r)
        world.add_body(shape=s, x=d["x"], q=d["q"], scale=(r, r, r), obj=k)
    mats = np.ascontiguousarray(ins["world"].copy())
    verts = P.GpuArray.from_(gpu, v)
    cam = S.path_camera(w, h, eye=(0.0, -16.0, 9.0), target=(0.0, 0.0, 1.5), rr_from=3)
    tr.set_samples(1 << 20)
    tr.autotune(cam, w, h, spp, reps=1)
    phys, scene, trace, rays, pairs = 0.0, 0.0, 0.0, 0, 0
    z0 = float(world.bodies()["x"][1:, 2].mean())
    for fr in range(frames):
        for _ in range(substeps):
            t0 = time.perf_counter()
            world.step(1.0 / 60.0 / substeps)
            poses = world.poses()
            t1 = time.perf_counter()
            mats[:1024] = poses[1:]
            assert tr.repose_on(verts) and tr.reinstance(mats)
            gpu.wait()
            t2 = time.perf_counter()
            phys += t1 - t0
            scene += t2 - t1
            pairs += len(world.pairs())
        c0 = tr.counters()
        tr.sample(cam, w, h, spp)
        trace += tr.time()[1]
        c1 = tr.counters()
        rays += int(c1[0] - c0[0] + c1[1] - c0[1])
    n = frames * substeps
    z1 = float(world.bodies()["x"][1:, 2].mean())
    log(f"xpbd: {frames} frames, physics {phys / n * 1e3:.3f} ms + scene {scene / n * 1e3:.3f} ms a substep, "
        f"trace {trace / frames * 1e3:.1f} ms a frame; mean height {z0:.2f} -> {z1:.2f} m")
    return {"bodies": 1024, "frames": frames, "substeps_per_frame": substeps, "frame": f"{w}x{h}, {spp} spp",
            "physics_ms_per_substep": phys / n * 1e3, "scene_ms_per_substep": scene / n * 1e3,
            "trace_ms_per_frame": trace / frames * 1e3, "mrays_s": rays / trace / 1e6,
            "pairs_per_substep": pairs / n, "contacts_last": world.contact_count(),
            "mean_height_m": [z0, z1], "path_waves_per_simd": tr.path_waves, "path_build": PATH_BUILDS[tr.path_regen]}


def rel_mse(x: np.ndarray, y: np.ndarray) -> float:
    """mean((x - y)^2 / (y^2 + 0.01)) over pixels and channels."""
    x, y = x.astype(np.float64), y.astype(np.float64)
    return float(np.mean((x - y) ** 2 / (y * y + 0.01)))


def learned_vs_plain(gpu: P.Gpu, net: "P.Kpcn", seeds, w: int = 1920, h: int = 1080, ref_spp: int = 4096,
                     spp: int = 4, max_raw_spp: int = 1024) -> dict:
    """Seconds a frame needs to reach relMSE <= 0.01 against a ref_spp render,
    on held-out scenes of the learned denoiser's generator (65,536 triangles,
    albedo U[0.2, 0.8]^3, a white world; seeds the training never saw):

    * plain: the raw path tracer, the fewest samples whose image reaches the
      bound (doubling from 4, then bisecting); its GPU time;
    * learned: LearnedFrame at `spp` samples (two halves, the network, the
      held-out check, more samples while the check fails, the plain render
      past its budget), wall time of the whole call;
    * a_trous: the a-trous denoiser over the raw trials, for reference.
    The ratio is plain / learned per seed, and its median."""
    from phiray_amd import kpcn as K

    per = []
    for seed in seeds:
        tris = S.cloud_scene(int(seed))
        cam = S.cloud_camera(w, h, rr_from=3)
        tr = P.Tracer.new(gpu, tris, 4)
        tr.set_samples(1 << 30)
        ref_cam = cam.copy()
        ref_cam["stream"] = K.REF_STREAM
        t0 = time.perf_counter()
        for _ in range(ref_spp // 64):
            tr.sample(ref_cam, w, h, 64)
        ref = tr.radiance()
        t_ref = time.perf_counter() - t0
        tr.autotune(cam, w, h, 4, reps=1)
        plain, a_trous, trials = None, None, []
        n = 4
        while n <= max_raw_spp and (plain is None or a_trous is None):
            fresh = cam.copy()
            fresh["stream"] = n + (3 << 16)
            tr.sample(fresh, w, h, n)
            render_s = tr.time()[1]
            raw = rel_mse(tr.radiance(), ref)
            clock = P.Clock(gpu)
            clock.start()
            tr.denoise(keep=False)
            den_s, _ = clock.stop()
            at = rel_mse(tr.radiance(), ref)
            trials.append({"spp": n, "render_s": render_s, "relmse_raw": raw, "relmse_a_trous": at})
            if plain is None and raw <= 0.01:
                plain = {"spp": n, "s_per_frame": render_s, "relmse": raw}
            if a_trous is None and at <= 0.01:
                a_trous = {"spp": n, "s_per_frame": render_s + den_s, "relmse": at}
            n *= 2
        if plain and plain["spp"] > 4:
            # The fewest samples: bisect between the doubling's last failing
            # and first passing counts.
            lo, hi = plain["spp"] // 2, plain["spp"]
            while hi - lo > 1:
                mid = (lo + hi) // 2
                fresh = cam.copy()
                fresh["stream"] = mid + (5 << 16)
                tr.sample(fresh, w, h, mid)
                render_s = tr.time()[1]
                raw = rel_mse(tr.radiance(), ref)
                trials.append({"spp": mid, "render_s": render_s, "relmse_raw": raw})
                if raw <= 0.01:
                    hi, plain = mid, {"spp": mid, "s_per_frame": render_s, "relmse": raw}
                else:
                    lo = mid
        tr.free()
        lf = K.LearnedFrame(gpu, tris, net)
        lf.render(cam, w, h, spp=spp)  # warm the network's workspace at this size
        lf.a.autotune(cam, w, h, max(spp // 2, 1), reps=1)
        lf.b.set_path_waves(lf.a.path_waves)
        lf.b.set_path_regen(lf.a.path_regen)
        gpu.wait()
        t0 = time.perf_counter()
        fresh = cam.copy()
        fresh["stream"] = 9 << 16
        res = lf.render(fresh, w, h, spp=spp)
        gpu.wait()
        learned_s = time.perf_counter() - t0
        learned = {"spp": res["spp"], "s_per_frame": learned_s, "relmse": rel_mse(lf.out.read().reshape(-1, 3), ref),
                   "checks": res["checks"], "fell_back_to_plain": res["plain"]}
        lf.free()
        row = {"seed": int(seed), "reference_s": t_ref, "plain": plain, "learned": learned, "a_trous": a_trous,
               "trials": trials}
        if plain:
            row["plain_over_learned"] = plain["s_per_frame"] / learned_s
        log(f"learned vs plain, seed {seed}: plain {plain and plain['spp']} spp {plain and plain['s_per_frame']:.4f} s, "
            f"learned {learned['spp']} spp {learned_s:.4f} s relMSE {learned['relmse']:.4f}"
            f"{' (plain fallback)' if res['plain'] else ''}")
        per.append(row)
    ratios = [r["plain_over_learned"] for r in per if "plain_over_learned" in r]
    return {"workload": f"{len(per)} held-out scenes of 65,536 triangles (seeds {int(seeds[0])}..{int(seeds[-1])}), "
                        f"{w}x{h}, white world, reference {ref_spp} spp, bound relMSE 0.01",
            "relmse": "mean((x - y)^2 / (y^2 + 0.01))", "learned_spp": spp, "per_seed": per,
            "plain_over_learned_median": float(np.median(ratios)) if ratios else None,
            "learned_relmse_max": max(r["learned"]["relmse"] for r in per)}


MFMA_BF16_PEAK_TFLOPS = 2500.0


def kpcn_rates(gpu: P.Gpu, net: "P.Kpcn | None", reps: int = 10) -> dict:
    """The learned denoiser alone: inference on one 1920x1080 frame (conv
    tiles autotuned) and one Adam training step over 8 crops of 512x512, on
    random frames. FLOPs = 2 x 9 x cin x cout per pixel of each conv at its
    level (a quarter of the pixels a level down); a training step does the
    forward, the input gradients and the weight gradients (3x)."""
    from phiray_amd import kpcn as K

    if net is None:
        net = P.Kpcn(gpu, K.init_params(0))
    rng = np.random.default_rng(0)
    lay = P.Kpcn.layers()
    flop_px = float(sum(2 * 9 * int(c[0]) * int(c[1]) / 4 ** int(c[2]) for c in lay))
    clock = P.Clock(gpu)
    w, h = 1920, 1080
    fr = P.GpuArray.from_(gpu, rng.uniform(0.05, 1.0, size=h * w * P.Kpcn.FRAME).astype(np.float32))
    out = P.GpuArray.zeroed(gpu, np.float32, w * h * 3)
    rows = net.autotune(fr, 1, w, h).tolist()
    for _ in range(3):
        net.infer(fr, 1, w, h, out)
    clock.start()
    for _ in range(reps):
        net.infer(fr, 1, w, h, out)
    t_inf, _ = clock.stop()
    t_inf /= reps
    del fr, out
    b, c = 8, 512
    fb = P.GpuArray.from_(gpu, rng.uniform(0.05, 1.0, size=b * c * c * P.Kpcn.FRAME).astype(np.float32))
    tb = P.GpuArray.from_(gpu, rng.uniform(0.05, 1.0, size=b * c * c * 3).astype(np.float32))
    trainer = P.Kpcn(gpu, K.init_params(1))
    for _ in range(3):
        trainer.train_step(fb, tb, b, c, c, 1e-4, want_loss=False)
    gpu.wait()
    clock.start()
    for _ in range(reps):
        trainer.train_step(fb, tb, b, c, c, 1e-4, want_loss=False)
    t_step, _ = clock.stop()
    t_step /= reps
    inf_tflops = flop_px * w * h / t_inf / 1e12
    train_tflops = 3 * flop_px * b * c * c / t_step / 1e12
    return {"params": int(net.param_count()), "flop_per_px": flop_px, "rows_per_wave": rows,
            "infer_1080p_ms": t_inf * 1e3, "infer_tflops": inf_tflops, "infer_mfma_frac": inf_tflops / MFMA_BF16_PEAK_TFLOPS,
            "train_step_8x512_ms": t_step * 1e3, "train_tflops": train_tflops,
            "train_mfma_frac": train_tflops / MFMA_BF16_PEAK_TFLOPS, "mfma_peak_tflops": MFMA_BF16_PEAK_TFLOPS}


def load_net(gpu: P.Gpu) -> "P.Kpcn | None":
    from phiray_amd import kpcn as K

    if not K.WEIGHTS.exists():
        log(f"no trained weights at {K.WEIGHTS}: the learned leg is skipped")
        return None
    return P.Kpcn(gpu, K.load(K.WEIGHTS))


def raycast_rate(gpu: P.Gpu, reps: int) -> dict:
    corners, _ = S.lbvh_scene(65536, 0)
    tr = P.Tracer.new(gpu, S.tris_from_corners(corners), 4)
    rays = S.pinhole_rays(256, 256)
    drays = P.GpuArray.from_(gpu, rays)
    hits = P.GpuArray.zeroed(gpu, P.HIT, len(rays))
    for _ in range(3):
        tr.raycast_on(drays, hits)
    gpu.wait()
    c0 = tr.counters()
    clock = P.Clock(gpu)
    clock.start()
    for _ in range(reps):
        tr.raycast_on(drays, hits)
    t, _ = clock.stop()
    c1 = tr.counters()
    steps, tests = (int(c1[4] - c0[4]) // reps), (int(c1[5] - c0[5]) // reps)
    per = t / reps
    nb = len(rays) * (RAY_BYTES + 20 + ROOT_BYTES) + steps * STEP_BYTES + tests * TEST_BYTES
    return {"rays": len(rays), "ms": per * 1e3, "mrays_s": len(rays) / per / 1e6, "gbs": nb / per / 1e9,
            "steps_per_ray": steps / len(rays), "tests_per_ray": tests / len(rays)}


def host_cpus() -> tuple[int, str]:
    """The CPUs this process may run on, at most 16 (a one-GPU box's share of
    its host), and the processor's model name."""
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    model = "unknown"
    try:
        for ln in pathlib.Path("/proc/cpuinfo").read_text().splitlines():
            if ln.startswith("model name"):
                model = ln.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return max(1, min(16, avail)), model


def cpu_lbvh_raycast(cores: int) -> dict:
    """The oracle's Bvh::build, refit and closest-hit raycast on the
    65,536-triangle set (seed 0, leaf 4, 256x256 rays): median of 5 after 2
    warm-ups, as walk_cost.rs times."""
    from oracle import oracle as O

    corners, moved = S.lbvh_scene(65536, 0)
    boxes, boxes2 = S.corner_boxes(corners), S.corner_boxes(moved)
    rays = S.pinhole_rays(256, 256)

    def med(fn):
        for _ in range(2):
            fn()
        ts = []
        for _ in range(5):
            t0 = time.perf_counter()
            fn()
            ts.append(time.perf_counter() - t0)
        return float(np.median(ts))

    built = {}

    def build():
        built["t"] = O.build_bvh(boxes, 4, P.NODE, cores)

    t_build = med(build)
    nodes, order, _ = built["t"]
    t_refit = med(lambda: O.refit(nodes, order, boxes2))
    tri9 = np.ascontiguousarray(corners[order].reshape(-1), dtype=np.float32)  # slot order
    t_ray = med(lambda: O.raycast(nodes, tri9, rays, P.HIT, False, cores))
    n = len(boxes)
    return {"tris": n, "build_ms": t_build * 1e3, "build_mtri_s": n / t_build / 1e6, "refit_ms": t_refit * 1e3,
            "refit_mtri_s": n / t_refit / 1e6, "raycast_ms": t_ray * 1e3, "raycast_mrays_s": len(rays) / t_ray / 1e6,
            "threads": {"build": cores, "refit": 1, "raycast": cores}}


def cpu_baseline(tris: np.ndarray, cam: np.ndarray, w: int, h: int, budget_s: float) -> dict:
    """The oracle (the CPU twin) on the host's cores, on a band of rows of the
    same frame, sample after sample until `budget_s` seconds have passed; and
    the 65,536-triangle LBVH build, refit and raycast on the same cores."""
    from oracle import oracle as O

    cores, model = host_cpus()
    t0 = time.perf_counter()
    sc = O.Scene(tris, 4, P.NODE, cores)
    t_scene = time.perf_counter() - t0
    rows = 16
    first = (h // 2 - rows // 2) * w
    count = rows * w
    px = np.zeros(P.PX_LEN * w * h + 1, dtype=np.uint32)
    rays, f = 0, 0
    t0 = time.perf_counter()
    while True:
        c, s = sc.sample(cam, w, h, px, f, cores, first, count)
        rays += c + s
        f += 1
        if time.perf_counter() - t0 >= budget_s:
            break
    dt = time.perf_counter() - t0
    return {"value": rays / dt / 1e6, "unit": "Mrays/s", "cores": cores, "kind": "port", "cpu_model": model,
            "host_cpus_visible": os.cpu_count(),
            "sample": f"{rows} rows x {w} px of the same 1920x1080 frame (rows {h // 2 - rows // 2}-{h // 2 + rows // 2 - 1}), "
                      f"{f} samples, {rays} rays in {dt:.1f} s on {cores} threads (the one-GPU share of the host); "
                      f"CPU LBVH build of the scene took {t_scene:.2f} s",
            "s_per_frame_est": dt / (f * count) * (w * h) * 64,
            "lbvh_65536": cpu_lbvh_raycast(cores)}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--spp", type=int, default=64)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--only", choices=["all", "path", "lbvh", "forest", "converge", "batch", "kpcn"], default="all")
    ap.add_argument("--path-waves", type=int, default=0, help="run this path kernel variant (2 to 5) instead of autotuning")
    ap.add_argument("--path-build", choices=["auto", "nested", "regen", "wavefront", "split"], default="auto",
                    help="with --path-waves: which path build runs (default: the nested one)")
    ap.add_argument("--copy-mb", type=int, default=0, help="also run the calibration copy kernel over this many MiB")
    ap.add_argument("--c5-seeds", type=int, default=8, help="held-out scenes of the learned-vs-plain leg")
    ap.add_argument("--c5-ref-spp", type=int, default=4096, help="its reference render's samples")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; run with --gpus 1 (multi-GPU sharding is design-only)"}))
        return 2

    gpu = P.Gpu.new(0)
    if gpu is None:
        print(json.dumps({"error": "no HIP device"}))
        return 1
    log(f"device {gpu.arch()} ({gpu.cu_count()} CUs)")
    out: dict = {}

    if args.copy_mb:
        n = args.copy_mb * (1 << 20) // 16
        src = P.GpuArray.zeroed(gpu, np.dtype((np.uint32, 4)), n)
        dst = P.GpuArray.zeroed(gpu, np.dtype((np.uint32, 4)), n)
        clock = P.Clock(gpu)
        P.copy_kernel(dst, src)
        clock.start()
        P.copy_kernel(dst, src)
        t, _ = clock.stop()
        out["copy"] = {"bytes": 2 * 16 * n, "ms": t * 1e3, "gbs": 2 * 16 * n / t / 1e9}
        log(f"copy kernel {out['copy']}")

    log("generating the spheres scene")
    tris = S.spheres_scene(spheres=64, level=5, seed=1)
    w, h, spp = args.width, args.height, args.spp
    cam = S.path_camera(w, h, rr_from=3)

    if args.only in ("all", "lbvh"):
        log("LBVH rates")
        out["lbvh_scene"] = lbvh_rates(gpu, S.of_tri(tris["a"], tris["b"], tris["c"]), 20)
        corners, _ = S.lbvh_scene(65536, 0)
        out["lbvh_65536"] = lbvh_rates(gpu, S.corner_boxes(corners), 50)
        out["raycast_65536"] = raycast_rate(gpu, 50)
        log("batch of 256 scenes")
        out["lbvh_batch_256"] = batch_rates(gpu, 10)
        log(f"lbvh {out['lbvh_scene']['build_mtri_s']:.1f} Mtri/s (scene), {out['lbvh_65536']['build_mtri_s']:.1f} Mtri/s (65536)")
    def leg(key: str, run) -> None:
        """A secondary measurement: its failure is reported in the line, and
        the headline measurement still runs."""
        try:
            out[key] = run()
        except Exception as e:  # noqa: BLE001
            log(f"{key} failed: {type(e).__name__}: {e}")
            out[key] = {"error": f"{type(e).__name__}: {e}"}

    if args.only in ("all", "batch"):
        log("batch of 256 scenes: build + render")
        leg("batch_256_render", lambda: batch_render_rates(gpu))
        if "error" not in out["batch_256_render"]:
            r = out["batch_256_render"]
            log(f"batch render: {r['render_ms']:.1f} ms, {r['mrays_s']:.1f} Mrays/s; graph pass {r['graph_pass_ms']:.1f} ms")
    if args.only in ("all", "kpcn"):
        log("learned denoiser: inference and a training step")
        leg("kpcn", lambda: kpcn_rates(gpu, load_net(gpu)))
        if "error" not in out["kpcn"]:
            log(f"kpcn: infer 1080p {out['kpcn']['infer_1080p_ms']:.2f} ms, train step {out['kpcn']['train_step_8x512_ms']:.2f} ms")
    if args.only in ("all", "converge"):
        net = load_net(gpu)
        if net is not None:
            log("learned vs plain: seconds a frame to relMSE <= 0.01")
            seeds = np.arange(2000, 2000 + args.c5_seeds)
            leg("s_per_frame_to_relmse_0.01", lambda: learned_vs_plain(gpu, net, seeds, ref_spp=args.c5_ref_spp))
    if args.only in ("all", "forest"):
        log("forest of 1,024 bodies")
        leg("forest_1024", lambda: forest_rates(gpu))
        if "error" not in out["forest_1024"]:
            log(f"forest: substep {out['forest_1024']['substep_ms']:.3f} ms, {out['forest_1024']['mrays_s']:.1f} Mrays/s")
        leg("xpbd_1024", lambda: xpbd_rates(gpu))

    line = {"metric": METRIC, "value": None, "unit": "Mrays/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": None, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f32", "data": "synthetic",
            "config": {"workload": f"path trace, 64 icospheres (level 5) + ground + 1 m^2 emissive quad = {len(tris)} triangles, "
                                   f"{w}x{h}, {spp} spp, max 4 diffuse bounces, Russian roulette after bounce 3, NEE",
                       "tris": int(len(tris)), "width": w, "height": h, "spp": spp, "bounces": 4}}

    if args.only in ("all", "path"):
        log("tracer setup")
        t0 = time.perf_counter()
        tracer = P.Tracer.new(gpu, tris, 4)
        log(f"Tracer::new {time.perf_counter() - t0:.2f} s, depth {tracer.depth}, {tracer.glow} emitters")
        if args.path_waves:
            tracer.set_path_waves(args.path_waves)
            build = {"auto": 0, "nested": 0, "regen": 1, "wavefront": 2, "split": 3}[args.path_build]
            tracer.set_path_regen(build)
            log(f"path kernel at {args.path_waves} waves a SIMD, build {args.path_build} (chosen on the command line)")
        else:
            try:
                waves, tune_s = tracer.autotune(cam, w, h, 8, reps=2)
                build_name = PATH_BUILDS[tracer.path_regen]
                log(f"autotune: path kernel at {waves} waves a SIMD, {build_name} build ({tune_s * 1e3:.1f} ms for 8 spp)")
                # Each build at the chosen waves, 8 samples, for the line.
                chosen = tracer.path_regen
                by_build = {}
                for b, name in enumerate(PATH_BUILDS):
                    tracer.set_path_regen(b)
                    tracer.sample(S.path_camera(64, 64), 64, 64, 1)  # another view: the next call starts over
                    tracer.sample(cam, w, h, 8)
                    by_build[name] = tracer.time()[1] * 1e3
                tracer.set_path_regen(chosen)
                out["path_build_ms_8spp"] = by_build
                log(f"builds at {waves} waves, 8 spp: " + ", ".join(f"{k} {v:.1f} ms" for k, v in by_build.items()))
            except P.PhrError as e:
                log(f"autotune failed ({e}); the path kernel stays at {tracer.path_waves} waves a SIMD")
        for i in range(args.warmup):
            tracer.sample(cam, w, h, spp)
            log(f"warmup frame {i}: {tracer.time()[1]:.3f} s")
        c0 = tracer.counters()
        gpu_s = []
        clock = P.Clock(gpu)
        clock.start()
        for i in range(args.steps):
            tracer.sample(cam, w, h, spp)
            gpu_s.append(tracer.time()[1])
            log(f"frame {i}: {gpu_s[-1]:.3f} s")
        total, _ = clock.stop()
        c1 = tracer.counters()
        closest, shadow, steps, tests = (int(c1[k] - c0[k]) for k in range(4))
        rays = closest + shadow
        per = total / args.steps
        line["value"] = rays / total / 1e6
        line["ms_per_step"] = per * 1e3
        kernel_s = float(np.mean(gpu_s))
        nbytes = (rays * (RAY_BYTES + HIT_BYTES + ROOT_BYTES) + steps * STEP_BYTES + tests * TEST_BYTES) / args.steps
        achieved = nbytes / kernel_s / 1e9
        build = PATH_BUILDS[tracer.path_regen]
        profile = "k_path_sw_spheres_1080p_64spp_traffic.txt" if build == "split" else "k_path_spheres_1080p_64spp_traffic.txt"
        traffic, source = profiled_traffic(profile, spp, w, h, tracer.path_waves, build)
        line["roofline"] = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                            "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "traffic_source": source,
                            "algorithmic_bytes": nbytes,
                            "kernel": {"nested": "k_path", "regen": "k_path_rg", "split": "k_path_sw",
                                       "wavefront": "k_wf_walk + k_wf_shade (one frame's turns)"}[build],
                            "per_ray_bytes": "32 ray + 8 hit + 32 root + 64 x inner steps + 48 x triangle tests",
                            "launch_ms": kernel_s * 1e3}
        if traffic:
            # What the DRAM counters saw: most node and triangle reads hit
            # L2 and the MALL, so the kernel waits on latency, not bandwidth.
            line["roofline"]["dram_achieved"] = traffic / kernel_s / 1e9
            line["roofline"]["dram_frac"] = traffic / kernel_s / 1e9 / HBM_PEAK_GBS
            line["roofline"]["limiter"] = ("VALU issue of divergent waves (lanes a VALU instruction: 15.1 of 64 split, 11.0 nested; "
                                           "VALU busy 0.76 / 0.73) and L2/MALL latency: frac counts the walk's logical bytes, "
                                           "dram_frac the counted DRAM bytes")
        line["s_per_frame"] = per
        line["path_waves_per_simd"] = tracer.path_waves
        line["path_regen"] = tracer.path_regen
        line["path_build"] = PATH_BUILDS[tracer.path_regen]
        line["split_shade_eighths"] = tracer.split_shade
        line["rays_per_frame"] = rays / args.steps
        line["closest_per_frame"] = closest / args.steps
        line["shadow_per_frame"] = shadow / args.steps
        line["steps_per_ray"] = steps / max(rays, 1)
        line["tests_per_ray"] = tests / max(rays, 1)
        if not args.no_cpu:
            log("CPU twin baseline")
            line["cpu_baseline"] = cpu_baseline(tris, cam, w, h, args.cpu_seconds)
    line.update(out)
    if out.get("lbvh_65536"):
        # The 65,536-triangle set (the CPU-runnable configuration) is the
        # headline LBVH figure; the 1.31M scene's beside it.
        line["lbvh_mtri_s"] = out["lbvh_65536"]["build_mtri_s"]
        line["lbvh_mtri_s_by_tris"] = {"65536": out["lbvh_65536"]["build_mtri_s"],
                                       str(out["lbvh_scene"]["tris"]): out["lbvh_scene"]["build_mtri_s"]}
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())

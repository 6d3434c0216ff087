"""Host side of the learned denoiser: weight initialisation and files, the
training set the path tracer renders, and the training loop.

Everything that computes runs in the library's HIP kernels (`Kpcn`,
`Tracer`); this module only decides what to render and which images a
batch holds. The network stands beside phiray's `kernel::Denoise`
(src/kernel/denoise.rs:26-64), which it does not replace.

Training data: seeded scenes of the LBVH workload's generator (65,536
triangles in the unit cube, albedo U[0.2, 0.8]^3, a white world; see
scenes.cloud_scene) seen by the 1920x1080 camera, cropped to `crop` x `crop`
windows. Inputs are `spp_in`-sample frames, several independent renders per
scene. Targets are `spp_ref`-sample renders of the same window on sample
streams the inputs never draw.

The learned frame (`LearnedFrame`) is guarded by the plain render: a frame
is traced as two independent halves, the network denoises one, the other
judges it (Kpcn.held_out), and more samples are taken while the estimated
error is over the bound; past the sample budget the plain render of every
sample is the answer.
"""
from __future__ import annotations

import math
import pathlib
import time

import numpy as np

from . import Gpu, GpuArray, Kpcn, Tracer
from . import scenes as S


def _log(msg: str) -> None:
    print(msg, flush=True)

HERE = pathlib.Path(__file__).resolve().parent
WEIGHTS = HERE / "kpcn_weights.npy"
REF_STREAM = 1 << 24  # sample streams of the targets; inputs use small ones
HALF_STREAM = 1 << 20  # the judging half of a learned frame
SCENES = {"cloud": (S.cloud_scene, S.cloud_camera), "spheres": (S.spheres_scene, S.path_camera)}


def layer_slices(layers: np.ndarray | None = None) -> list[tuple[slice, slice, tuple[int, int]]]:
    """(weights, bias, (cout, cin)) slices of the flat parameter vector."""
    layers = Kpcn.layers() if layers is None else layers
    out, off = [], 0
    for cin, cout, _level, _relu in layers.tolist():
        nw = 9 * cout * cin
        out.append((slice(off, off + nw), slice(off + nw, off + nw + cout), (cout, cin)))
        off += nw + cout
    return out


def init_params(seed: int = 0) -> np.ndarray:
    """He-normal weights over the used input channels, zero biases. The padded
    input channels (10..15 of the first layer) and logits (25..31) stay zero.
    The output layer starts at a tenth of He scale, so the first filters are
    close to a box."""
    rng = np.random.default_rng(seed)
    layers = Kpcn.layers()
    p = np.zeros(Kpcn.param_count(), dtype=np.float32)
    sl = layer_slices(layers)
    for i, (ws, _bs, (cout, cin)) in enumerate(sl):
        used_in = 10 if i == 0 else cin
        used_out = 25 if i == len(sl) - 1 else cout
        std = np.sqrt(2.0 / (9 * used_in)) * (0.1 if i == len(sl) - 1 else 1.0)
        w = np.zeros((9, cout, cin), dtype=np.float32)
        w[:, :used_out, :used_in] = rng.normal(0.0, std, size=(9, used_out, used_in))
        p[ws] = w.ravel()
    return p


def save(path: pathlib.Path | str, params: np.ndarray) -> None:
    np.save(path, np.ascontiguousarray(params, dtype=np.float32), allow_pickle=False)


def load(path: pathlib.Path | str = WEIGHTS) -> np.ndarray:
    p = np.load(path, allow_pickle=False)
    if p.dtype != np.float32 or p.size != Kpcn.param_count():
        raise ValueError(f"{path}: not a denoiser weight file ({p.dtype}, {p.size} values)")
    return p


def crop_camera(w: int, h: int, x0: int, y0: int, cw: int, ch: int, camera=S.path_camera, **kw) -> np.ndarray:
    """The camera `camera(w, h)` restricted to the window [x0, x0+cw) x
    [y0, y0+ch): same rays per pixel, through the view-plane offset `pan`."""
    cam = camera(w, h, **kw)
    full_scale = float(cam["scale"][0])
    c = cam[0]
    c["scale"] = full_scale * w / cw
    c["aspect"] = cw / ch
    # su_crop - pan0 = su_full * cw / w  (and the same for v, over the aspect).
    c["pan"][0] = (w / 2 - cw / 2 - x0) / cw
    c["pan"][1] = (h / 2 - ch / 2 - y0) / cw
    return cam


def train(gpu: Gpu, frames: GpuArray, targets: GpuArray, n_scenes: int, variants: int, crop: int, steps: int,
          batch: int = 8, lr: float = 1e-3, lr_final: float | None = None, seed: int = 0,
          params: np.ndarray | None = None, log=_log, log_every: int = 200,
          budget_s: float | None = None) -> tuple[Kpcn, list[tuple[int, float]]]:
    """Adam over random batches of (scene, render) images; image
    s * variants + j is render j of scene s and is scored against target s.
    The rate falls from `lr` to `lr_final` on a cosine over the steps (or
    over `budget_s` seconds, whichever ends first)."""
    lo = lr if lr_final is None else lr_final
    net = Kpcn(gpu, init_params(seed) if params is None else params)
    rng = np.random.default_rng(seed + 1)
    hist, t0 = [], time.perf_counter()
    good, scale = net.params(), 1.0  # the last weights with a finite loss; a rate factor
    for step in range(steps):
        scenes = rng.integers(0, n_scenes, size=batch)
        var = rng.integers(0, variants, size=batch)
        want = (step + 1) % log_every == 0 or step == steps - 1
        frac = step / max(steps - 1, 1)
        if budget_s is not None:
            frac = max(frac, min(1.0, (time.perf_counter() - t0) / budget_s))
        rate = (lo + 0.5 * (lr - lo) * (1.0 + np.cos(np.pi * frac))) * scale
        loss = net.train_indexed(frames, scenes * variants + var, targets, scenes, crop, crop, rate, want_loss=want)
        if want:
            if not np.isfinite(loss):  # diverged: back to the last finite weights at half the rate
                scale *= 0.5
                net.set_params(good)
                log(f"[kpcn] step {step + 1}: loss {loss}; weights restored, rate x{scale}")
                continue
            good = net.params()
            hist.append((step + 1, loss))
            el = time.perf_counter() - t0
            log(f"[kpcn] step {step + 1}/{steps} loss {loss:.5f} ({el:.0f} s, {1e3 * el / (step + 1):.1f} ms/step)")
            if budget_s is not None and el > budget_s:
                log(f"[kpcn] time budget reached at step {step + 1}")
                break
    gpu.wait()
    return net, hist


def rel_mse(x: np.ndarray, y: np.ndarray) -> float:
    """mean((x - y)^2 / (y^2 + 0.01)) over pixels and channels."""
    x = np.asarray(x, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    return float(np.mean((x - y) ** 2 / (y * y + 0.01)))


def main(argv=None) -> int:
    import argparse
    import json

    ap = argparse.ArgumentParser(description="Train the learned denoiser on path-traced pairs.")
    ap.add_argument("--scenes", type=int, default=32, help="training scenes (seeds 1000, 1001, ...)")
    ap.add_argument("--kind", choices=sorted(SCENES), default="cloud", help="the scene generator")
    ap.add_argument("--crop", type=int, default=512)
    ap.add_argument("--spp-ref", type=int, default=4096)
    ap.add_argument("--spp-in", type=str, default="2,4", help="input sample counts, one variant set each")
    ap.add_argument("--variants", type=int, default=4, help="renders per scene per input sample count")
    ap.add_argument("--steps", type=int, default=20000)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--lr-final", type=float, default=1e-4)
    ap.add_argument("--budget-s", type=float, default=None, help="stop training after this many seconds")
    ap.add_argument("--out", type=str, default=str(WEIGHTS))
    ap.add_argument("--log", type=str, default=None, help="JSON file for the loss history")
    args = ap.parse_args(argv)
    gpu = Gpu.new(0)
    if gpu is None:
        raise SystemExit("no HIP device")
    spps = [int(s) for s in args.spp_in.split(",")]
    seeds = list(range(1000, 1000 + args.scenes))
    t0 = time.perf_counter()
    frames, targets, meta = render_multi(gpu, seeds, args.crop, spps, args.spp_ref, args.variants, kind=args.kind)
    _log(f"[kpcn] training set rendered in {time.perf_counter() - t0:.0f} s")
    net, hist = train(gpu, frames, targets, len(seeds), len(spps) * args.variants, args.crop, args.steps,
                      batch=args.batch, lr=args.lr, lr_final=args.lr_final, budget_s=args.budget_s)
    save(args.out, net.params())
    _log(f"[kpcn] weights written to {args.out}")
    if args.log:
        with open(args.log, "w") as f:
            json.dump({"args": vars(args), "scenes": meta, "loss": hist}, f)
    return 0


def render_multi(gpu: Gpu, seeds, crop: int, spps, spp_ref: int, variants: int, full=(1920, 1080),
                 log=_log, kind: str = "cloud") -> tuple[GpuArray, GpuArray, list[dict]]:
    """Frames and targets rendered on the device. Each scene gets `variants`
    renders at each input sample count in `spps`, so image s * V + j
    (V = len(spps) * variants) is render j of scene s; target s is scene s at
    `spp_ref` samples."""
    w, h = full
    npx = crop * crop
    V = len(spps) * variants
    frames = GpuArray.zeroed(gpu, np.float32, len(seeds) * V * npx * Kpcn.FRAME)
    targets = GpuArray.zeroed(gpu, np.float32, len(seeds) * npx * 3)
    meta = []
    for si, seed in enumerate(seeds):
        t0 = time.perf_counter()
        rng = np.random.default_rng(seed)
        x0 = int(rng.integers(0, w - crop + 1))
        y0 = int(rng.integers(0, h - crop + 1))
        scene, camera = SCENES[kind]
        tr = Tracer.new(gpu, scene(seed=int(seed)), 4)
        tr.set_samples(1 << 30)
        cam = crop_camera(w, h, x0, y0, crop, crop, camera=camera, rr_from=3)
        ref = cam.copy()
        ref["stream"] = REF_STREAM
        left = spp_ref
        while left > 0:
            step = min(64, left)
            tr.sample(ref, crop, crop, step)
            left -= step
        targets.write_at(si * npx * 3, tr.radiance().ravel())
        j, stream = 0, 1
        for spp in spps:
            for _ in range(variants):
                c = cam.copy()
                c["stream"] = stream
                stream += spp
                tr.sample(c, crop, crop, spp)
                tr.frame_into(frames, (si * V + j) * npx * Kpcn.FRAME)
                j += 1
        tr.free()
        meta.append({"seed": int(seed), "x0": x0, "y0": y0})
        log(f"[kpcn] scene {si + 1}/{len(seeds)} seed {seed} crop ({x0},{y0}) in {time.perf_counter() - t0:.1f} s")
    return frames, targets, meta


class LearnedFrame:
    """Frames of one scene made by the learned denoiser under its held-out
    check. Two tracers hold the two halves of every frame (the same samples
    a pixel on other streams); `render` denoises each half (D_A, D_B),
    estimates D_A's relMSE E from half B (Kpcn.held_out) and the answer's
    from E and the halves' spread S (Kpcn.spread): the answer is the
    network's output from both halves together, whose noise-driven error is
    half of D_A's (S / 2 of it) and whose bias is D_A's, so its relMSE is
    taken as E - S / 4. More samples are taken while that is over `bound`;
    once `max_spp` samples would be passed the plain render of every sample
    taken is the answer. A failed check takes as many more samples as its
    estimate asks for, the error falling as 1 / samples: at least one, at
    most eight times as many as were taken, and never past the budget (the
    last check before a fallback is at `max_spp`)."""

    def __init__(self, gpu: Gpu, tris: np.ndarray, net: Kpcn, leaf: int = 4):
        self.gpu, self.net = gpu, net
        self.a, self.b = Tracer.new(gpu, tris, leaf), Tracer.new(gpu, tris, leaf)
        for t in (self.a, self.b):
            t.set_samples(1 << 30)
        self._size = None

    def _buffers(self, w: int, h: int) -> None:
        if self._size == (w, h):
            return
        n = w * h
        self.fa = GpuArray.zeroed(self.gpu, np.float32, n * Kpcn.FRAME)
        self.fb = GpuArray.zeroed(self.gpu, np.float32, n * Kpcn.FRAME)
        self.fm = GpuArray.zeroed(self.gpu, np.float32, n * Kpcn.FRAME)
        self.da = GpuArray.zeroed(self.gpu, np.float32, n * 3)
        self.db = GpuArray.zeroed(self.gpu, np.float32, n * 3)
        self.out = GpuArray.zeroed(self.gpu, np.float32, n * 3)
        self._size = (w, h)

    def render(self, cam: np.ndarray, w: int, h: int, spp: int = 4, bound: float = 0.01, max_spp: int = 256) -> dict:
        """Render, check and denoise; the image is left in `self.out`
        (3 floats a pixel). Returns what was done: samples taken, each
        check's estimates (half A's, and the answer's), whether the plain
        render was the answer. Past a failed check each half goes from t to
        min(8 t, max(t + 1, ceil(t est / bound)), max_spp / 2) samples, est
        the answer's estimate; with no room left the plain render is the
        answer."""
        self._buffers(w, h)
        n = w * h
        half = max(spp // 2, 1)
        ca, cb = cam.copy(), cam.copy()
        cb["stream"] = int(cam["stream"][0]) + HALF_STREAM
        self.a.sample(ca, w, h, half)
        self.b.sample(cb, w, h, half)
        taken, checks = half, []
        while True:
            self.a.frame_into(self.fa)
            self.b.frame_into(self.fb)
            self.net.infer(self.fa, 1, w, h, self.da)
            err, noise = Kpcn.held_out(self.gpu, self.da, self.fa, self.fb, n)
            self.net.infer(self.fb, 1, w, h, self.db)
            spread = Kpcn.spread(self.gpu, self.da, self.db, n)
            est = err - spread / 4.0  # the answer's: D_A's bias, half of its noise-driven error
            checks.append({"spp": 2 * taken, "estimate": err, "half_noise": noise, "spread": spread, "answer_estimate": est})
            if est <= bound:
                Kpcn.frame_mean(self.gpu, self.fa, self.fb, n, self.fm)
                self.net.infer(self.fm, 1, w, h, self.out)
                return {"spp": 2 * taken, "checks": checks, "plain": False}
            nxt = min(8 * taken, max(taken + 1, math.ceil(taken * est / bound)), max_spp // 2)
            if nxt <= taken:  # the budget is spent
                break
            self.a.sample(ca, w, h, nxt - taken)  # the same view: the halves go on accumulating
            self.b.sample(cb, w, h, nxt - taken)
            taken = nxt
        Kpcn.frame_mean(self.gpu, self.fa, self.fb, n, self.out, channels=3)
        self.gpu.wait()
        return {"spp": 2 * taken, "checks": checks, "plain": True}

    def free(self) -> None:
        self.a.free()
        self.b.free()


if __name__ == "__main__":
    raise SystemExit(main())

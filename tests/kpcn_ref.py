"""PyTorch float32 restatement of the learned denoiser, the reference its HIP
kernels are tested against.

`forward(..., bf16=True)` rounds what the HIP path rounds: the conv
weights, each conv's input (every activation is stored as BF16), the pooled
tensors and the features. It keeps the logits and the filter in FP32, as the
HIP path does. `bf16=False` is the plain FP32 network.
"""
from __future__ import annotations

import numpy as np
import torch
import torch.nn.functional as F

TAPS = 25


def bf(x: torch.Tensor) -> torch.Tensor:
    return x.to(torch.bfloat16).to(torch.float32)


def unpack(params: np.ndarray, layers: np.ndarray) -> list[tuple[torch.Tensor, torch.Tensor, bool]]:
    """Per layer (W [cout, cin, 3, 3], b [cout], relu) from the flat vector,
    whose layer weights are [tap = 3 (dy + 1) + dx + 1][cout][cin]."""
    out, off = [], 0
    for cin, cout, _level, relu in layers.tolist():
        nw = 9 * cout * cin
        w = params[off:off + nw].reshape(3, 3, cout, cin).transpose(2, 3, 0, 1)
        b = params[off + nw:off + nw + cout]
        off += nw + cout
        out.append((torch.tensor(np.ascontiguousarray(w)), torch.tensor(np.ascontiguousarray(b)), bool(relu)))
    return out


def demod(frames: torch.Tensor) -> tuple[torch.Tensor, torch.Tensor]:
    """(radiance / max(albedo, 0.05), max(albedo, 0.05)), channels first."""
    rad = frames[..., 0:3].permute(0, 3, 1, 2)
    ad = torch.clamp(frames[..., 3:6], min=0.05).permute(0, 3, 1, 2)
    return rad / ad, ad


def features(frames: torch.Tensor) -> torch.Tensor:
    b, h, w, _ = frames.shape
    hp, wp = (h + 7) // 8 * 8, (w + 7) // 8 * 8
    d, _ = demod(frames)
    alb = frames[..., 3:6].permute(0, 3, 1, 2)
    nrm = frames[..., 6:9].permute(0, 3, 1, 2)
    z = frames[..., 9:10].permute(0, 3, 1, 2)
    zf = torch.where(z < 0, torch.full_like(z, -1.0), z / (1.0 + z))
    f = torch.cat([torch.log1p(torch.clamp(d, min=0.0)), alb, nrm, zf, torch.zeros(b, 6, h, w)], 1)
    return F.pad(f, (0, wp - w, 0, hp - h), mode="replicate")


def forward(params: np.ndarray, layers: np.ndarray, frames, bf16: bool = True, params_t=None) -> torch.Tensor:
    """Denoised radiance [B, H, W, 3]. `params_t` (torch tensors from
    `unpack`, possibly requiring grad) overrides `params`."""
    r = bf if bf16 else (lambda t: t)
    fr = frames if isinstance(frames, torch.Tensor) else torch.tensor(frames)
    L = params_t if params_t is not None else unpack(params, layers)

    def conv(i, x):
        w, b, relu = L[i]
        y = F.conv2d(x, r(w), b, padding=1)
        return F.relu(y) if relu else y

    def up(x):
        return F.interpolate(x, scale_factor=2, mode="nearest")

    x = r(features(fr))
    a0a = r(conv(0, x))
    a0b = r(conv(1, a0a))
    q0 = r(F.avg_pool2d(a0b, 2))
    a1a = r(conv(2, q0))
    a1b = r(conv(3, a1a))
    q1 = r(F.avg_pool2d(a1b, 2))
    a2a = r(conv(4, q1))
    a2b = r(conv(5, a2a))
    q2 = r(F.avg_pool2d(a2b, 2))
    a3 = r(conv(6, q2))
    d2 = r(conv(7, torch.cat([up(a3), a2b], 1)))
    d1 = r(conv(8, torch.cat([up(d2), a1b], 1)))
    d0 = r(conv(9, torch.cat([up(d1), a0b], 1)))
    lg = conv(10, d0)
    b, h, w, _ = fr.shape
    wk = torch.softmax(lg[:, :TAPS, :h, :w], dim=1)
    d, ad = demod(fr)
    dp = F.pad(d, (2, 2, 2, 2), mode="replicate")
    acc = torch.zeros_like(d)
    for k in range(TAPS):
        ky, kx = divmod(k, 5)
        acc = acc + wk[:, k:k + 1] * dp[:, :, ky:ky + h, kx:kx + w]
    return (ad * acc).permute(0, 2, 3, 1)


def rel_mse(y: torch.Tensor, t: torch.Tensor) -> torch.Tensor:
    return torch.mean((y - t) ** 2 / (t * t + 0.01))


def grads(params: np.ndarray, layers: np.ndarray, frames: np.ndarray, targets: np.ndarray,
          bf16: bool = True) -> tuple[float, np.ndarray]:
    """(loss, d loss / d params) in the flat layout, by autograd."""
    L = [(w.clone().requires_grad_(True), b.clone().requires_grad_(True), relu) for w, b, relu in unpack(params, layers)]
    y = forward(params, layers, frames, bf16=bf16, params_t=L)
    loss = rel_mse(y, torch.tensor(targets).reshape(y.shape))
    loss.backward()
    flat = []
    for w, b, _ in L:
        flat.append(w.grad.permute(2, 3, 0, 1).reshape(-1).numpy())
        flat.append(b.grad.numpy())
    return float(loss.detach()), np.concatenate(flat).astype(np.float32)


def synthetic_frames(b: int, h: int, w: int, seed: int) -> tuple[np.ndarray, np.ndarray]:
    """Frames shaped like a noisy render (radiance with fireflies, albedo in
    [0, 1], unit normals, depth with some misses) and a smooth target."""
    rng = np.random.default_rng(seed)
    yy, xx = np.meshgrid(np.linspace(0, 1, h), np.linspace(0, 1, w), indexing="ij")
    base = 0.5 + 0.4 * np.sin(6 * xx + 3 * yy)[None, :, :, None] * np.array([1.0, 0.8, 0.6])
    alb = rng.uniform(0.1, 0.9, size=(b, h, w, 3))
    tgt = (base * alb).astype(np.float32)
    noise = rng.exponential(1.0, size=(b, h, w, 3))
    rad = tgt * noise
    n = rng.normal(size=(b, h, w, 3))
    n /= np.linalg.norm(n, axis=-1, keepdims=True)
    depth = rng.uniform(0.5, 6.0, size=(b, h, w, 1))
    depth[rng.uniform(size=(b, h, w, 1)) < 0.05] = -1.0
    fr = np.concatenate([rad, alb, n, depth], -1).astype(np.float32)
    return fr, np.ascontiguousarray(np.broadcast_to(tgt, (b, h, w, 3))).astype(np.float32)

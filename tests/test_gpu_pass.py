"""phiray's own recorded build, through the Kernel/Pass boundary (GPU).

`GpuBvh::build` as src/bvh/gpu.rs:23-90 records it — `bvh.morton`, the 27
dispatches of `GpuSort::record` (src/bvh/sort.rs:30-51), `bvh.karras`, the
host's `assemble`, then `depth + 1` (rounded to even) Jacobi `bvh.refit`
steps — replayed with the HIP kernels registered under those names. Codes,
order, links and the refit bounds must equal the oracle's bit for bit; the
binding-count check must refuse a wrong list as the reference's assert does.
"""
import struct

import numpy as np
import pytest

import phiray_amd as P
from phiray_amd import scenes as S

pytestmark = pytest.mark.gpu


def u32(*v):
    return struct.pack(f"<{len(v)}I", *v)


def test_pass_rejects_a_wrong_binding_count(gpu):
    a = P.GpuArray.zeroed(gpu, np.uint32, 16)
    p = P.Pass(gpu)
    assert P.Pass.kernel("radix.count") == (16, 2)
    with pytest.raises(P.PhrError):
        p.add("radix.count", 16, u32(0, 0, 0, 0), [a])
    with pytest.raises(P.PhrError):
        p.add("no.such.kernel", 16, u32(0, 0, 0, 0), [a])


@pytest.mark.parametrize("n", [2, 257, 20_011])
def test_reference_recorded_build(gpu, orc, n):
    print(f"[pass] recorded build n={n}", flush=True)
    prims = S.xorshift_prims_gpu_test(n, 0.0)
    # src/bvh/gpu.rs:28-37: the scene box and its reciprocal on the host.
    lo = prims["lo"].min(0).astype(np.float32)
    hi = prims["hi"].max(0).astype(np.float32)
    e = (hi - lo).astype(np.float32)
    inv = np.where(e > 0, np.float32(1.0) / np.where(e > 0, e, np.float32(1.0)), np.float32(0.0)).astype(np.float32)
    gp = P.GpuArray.from_(gpu, prims)
    codes = P.GpuArray.zeroed(gpu, np.uint32, n)
    key = P.GpuArray.zeroed(gpu, np.dtype((np.uint32, 2)), n)
    alt = P.GpuArray.zeroed(gpu, np.dtype((np.uint32, 2)), n)
    tally = P.GpuArray.zeroed(gpu, np.uint32, 512)
    sorted_codes = P.GpuArray.zeroed(gpu, np.uint32, n)
    links = P.GpuArray.zeroed(gpu, P.LINK, n - 1)
    p = P.Pass(gpu)
    p.add("bvh.morton", n, u32(0, 0, 0, 0) + lo.tobytes() + b"\0" * 4 + inv.tobytes() + b"\0" * 4, [gp, codes])
    base = u32(0, 0, 0, 0)
    p.add("radix.seed", n, base, [codes, key])
    p.add("radix.clear", 512, base, [tally])
    for d in range(8):
        prm = u32(0, 0, d * 8, 0)
        src, dst = (key, alt) if d % 2 == 0 else (alt, key)
        p.add("radix.count", n, prm, [src, tally])
        p.add("radix.offset", 1, prm, [tally])
        p.add("radix.scatter", n, prm, [src, dst, tally])
    p.add("radix.spread", n, base, [key, sorted_codes])
    p.add("bvh.karras", n - 1, u32(0, 0, n, 0), [sorted_codes, links])
    assert len(p) == 29  # morton, seed, clear, 8 x (count, offset, scatter), spread, karras
    p.run()
    gpu.wait()
    want_codes = orc.codes(prims)
    assert (codes.read() == want_codes).all()
    order, srt = orc.sort_codes(want_codes)
    keys = key.read()
    assert (keys[:, 0] == order).all() and (keys[:, 1] == srt).all()
    assert (sorted_codes.read() == srt).all()
    want_links = orc.links(srt, P.LINK)
    assert (links.read().view(np.uint32) == want_links.view(np.uint32)).all()
    # Replay: the same recorded pass gives the same answer again.
    p.run()
    gpu.wait()
    assert (key.read()[:, 0] == order).all()
    # Assemble on the host (src/bvh/gpu.rs:60-61), Jacobi refit on the device.
    nodes = np.zeros(2 * n - 1, dtype=P.NODE)
    depth = orc.lib().orc_assemble(orc.ptr(want_links), n, 4, orc.ptr(nodes))
    gnodes = P.GpuArray.from_(gpu, nodes)
    scratch = P.GpuArray.from_(gpu, nodes)
    gorder = P.GpuArray.from_(gpu, order)
    refit = P.Pass(gpu)
    steps = (depth + 1 + 1) // 2 * 2
    for k in range(steps):
        src, dst = (gnodes, scratch) if k % 2 == 0 else (scratch, gnodes)
        refit.add("bvh.refit", 2 * n - 1, u32(0, 0, 0, 0), [gp, gorder, src, dst])
    refit.run()
    gpu.wait()
    got = gnodes.read()
    want = orc.refit(nodes, order, prims)
    r = orc.reachable(want)
    assert (got[r].view(np.uint32) == want[r].view(np.uint32)).all()

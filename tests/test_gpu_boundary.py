"""The rest of phiray's host surface over the C ABI (GPU).

`GpuArray::write_at` / `write_front` refuse what passes the length and write
nothing (src/gpu/array.rs:85-107); `read_first` clamps (:138-142); `split`
cuts items over bindings with placeholders for unused ones and refuses what
does not fit (:58-83); `Pass::compiled` counts kernel types
(src/gpu/pass.rs:151-154); `Gpu::shared` is one device for every caller
(src/gpu/mod.rs:113-118); `Tracer::with_chunks` applies phiray's binding
limits (src/trace/build.rs:136-200) and reports its split (src/trace/mod.rs:118-127).
"""
import numpy as np
import pytest

import phiray_amd as P
from phiray_amd import scenes as S

pytestmark = pytest.mark.gpu


def test_array_writes_refuse_past_the_length(gpu):
    print("[boundary] array writes", flush=True)
    a = P.GpuArray.from_(gpu, np.arange(8, dtype=np.uint32))
    assert a.write_at(5, np.array([50, 60, 70], np.uint32))
    assert not a.write_at(6, np.array([1, 2, 3], np.uint32))
    assert a.write_front(np.array([9, 9], np.uint32))
    assert not a.write_front(np.arange(9, dtype=np.uint32))
    assert (a.read() == [9, 9, 2, 3, 4, 50, 60, 70]).all()
    assert (a.read_first(3) == [9, 9, 2]).all()
    assert len(a.read_first(100)) == 8


def test_split_cuts_over_bindings(gpu):
    print("[boundary] split", flush=True)
    items = np.arange(10, dtype=np.uint32)
    parts, cut = P.GpuArray.split(gpu, items, 16, 3)  # 4 items a binding
    assert cut == 4 and [list(p.read()) for p in parts] == [[0, 1, 2, 3], [4, 5, 6, 7], [8, 9]]
    assert P.GpuArray.split(gpu, items, 16, 2) is None
    parts, cut = P.GpuArray.split(gpu, items[:3], 1024, 2)
    assert cut == 3 and list(parts[0].read()) == [0, 1, 2] and len(parts[1]) == 1  # a placeholder
    parts, cut = P.GpuArray.split(gpu, items[:0], 1024, 2)
    assert cut == 1 and all(len(p) == 1 for p in parts)


def test_pass_counts_kernel_types(gpu):
    print("[boundary] pass compiled", flush=True)
    n = 64
    codes = P.GpuArray.zeroed(gpu, np.uint32, n)
    key = P.GpuArray.zeroed(gpu, np.dtype((np.uint32, 2)), n)
    tally = P.GpuArray.zeroed(gpu, np.uint32, 512)
    p = P.Pass(gpu)
    assert p.compiled() == 0
    rp = np.zeros(4, np.uint32).tobytes()
    p.add("radix.seed", n, rp, [codes, key])
    p.add("radix.clear", 512, rp, [tally])
    p.add("radix.clear", 512, rp, [tally])
    assert len(p) == 3 and p.compiled() == 2
    p.run()
    assert (key.read()[:, 0] == np.arange(n)).all()


def test_shared_device_is_one_device():
    print("[boundary] shared gpu", flush=True)
    a, b = P.Gpu.shared(), P.Gpu.shared()
    assert a is not None and a._h.value == b._h.value
    t = P.GpuArray.zeroed(a, np.uint32, 4)
    assert (t.read() == 0).all()


def test_with_chunks_applies_the_binding_limits(gpu):
    print("[boundary] with_chunks", flush=True)
    b, _ = S.bodies_scene(bodies=12)
    arrs = b.arrays()
    verts, faces = arrs[0], arrs[1]
    t = P.Tracer.with_chunks(gpu, *arrs, limit=1 << 30, chunks=3)
    assert t is not None and t.chunks() == 4
    nodes, fcut, vcut, atlas = t.split()
    assert fcut == len(faces) and vcut == len(verts) and atlas == 0 and nodes > len(faces)
    base = P.Tracer.from_scene(gpu, *arrs)
    assert base.chunks() == 2 and base.split()[1] == len(faces)
    # The same images as from_scene.
    cam = S.path_camera(32, 24, rr_from=3)
    t.sample(cam, 32, 24, 1)
    base.sample(cam, 32, 24, 1)
    assert (t.px() == base.px()).all()
    assert P.Tracer.with_chunks(gpu, *arrs, limit=64, chunks=2) is None
    t.free()
    base.free()


def test_warm_loads_every_path_build(gpu):
    """Tracer::warm (src/trace/build.rs:27-69): every build named once with
    its load time; a tracer renders after it as before."""
    print("[boundary] warm", flush=True)
    first = P.Tracer.warm(gpu)
    names = [n for n, _ in first]
    assert len(names) == len(set(names)) == 28 and "k_path<4>" in names and "k_path_sw<4>" in names and "k_path_batch_sw<5>" in names
    assert all(ms >= 0.0 for _, ms in first)
    again = P.Tracer.warm(gpu)
    assert [n for n, _ in again] == names
    tr = P.Tracer.new(gpu, S.spheres_scene(spheres=2, level=1, seed=4), 4)
    tr.sample(S.path_camera(16, 16, rr_from=3), 16, 16, 1)
    assert P.planes(tr.px(), 16, 16)["count"].min() == 1
    tr.free()

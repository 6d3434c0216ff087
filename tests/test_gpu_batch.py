"""Many trees in one GPU build: each segment == the oracle's Bvh::build of its boxes alone.

A batch holds the trees of a Forest's meshes (src/trace/forest.rs:44-66) or
of independent scenes. Every segment downloaded relative to itself must
equal, bit for bit, what `Bvh::build` (src/bvh/build.rs:94-100) gives for
that segment's boxes: order, depth, the links and bounds of every reachable
node; then a refit of the whole batch after every box moved equals the
oracle's refit per segment. Cases: empty and one-box segments, leaf 1 and 4,
the 31-deep chain beside ordinary scenes, more than 256 segments (two sort
passes over the segment id) and a slice of the 256-scene batch workload.
"""
import numpy as np
import pytest

import phiray_amd as P
from phiray_amd import scenes as S
from tests.test_gpu_lbvh import same

pytestmark = pytest.mark.gpu


def boxes_of(sizes, seed0=0):
    parts = [S.corner_boxes(S.lbvh_scene(int(m), seed0 + i)[0]) if m else np.zeros(0, dtype=P.AABB)
             for i, m in enumerate(sizes)]
    lo = np.concatenate([[0], np.cumsum([len(p) for p in parts])]).astype(np.uint32)
    return np.concatenate(parts), lo


def check_batch(gpu, orc, prims, lo, leaf, what, moved=None):
    g = P.GpuBvh.build_batch(gpu, prims, lo, leaf)
    assert g.segments == len(lo) - 1
    want = {}
    for s in range(len(lo) - 1):
        a, b = int(lo[s]), int(lo[s + 1])
        if a == b:
            continue
        nodes, order, depth = orc.build_bvh(prims[a:b], leaf, P.NODE, 8)
        want[s] = (nodes, order, depth)
        same(g.download_segment(s), nodes, order, depth, f"{what} segment {s}")
        assert g.segment_depth(s) == depth
    if moved is not None:
        assert g.update(moved)
        for s, (nodes, order, depth) in want.items():
            a, b = int(lo[s]), int(lo[s + 1])
            same(g.download_segment(s), orc.refit(nodes, order, moved[a:b]), order, depth, f"{what} refit segment {s}")
    return g


@pytest.mark.parametrize("sizes", [[1], [1, 1, 1], [3, 0, 7, 1, 0], [100, 257, 1, 2, 64], [0, 5000, 0]])
@pytest.mark.parametrize("leaf", [1, 4])
def test_small_batches(gpu, orc, sizes, leaf):
    print(f"[batch] sizes={sizes} leaf={leaf}", flush=True)
    prims, lo = boxes_of(sizes)
    check_batch(gpu, orc, prims, lo, leaf, f"sizes {sizes}")


def test_deep_chain_beside_scenes(gpu, orc):
    print("[batch] deep chains", flush=True)
    from tests.test_oracle_lbvh import deep_chain_prims

    dc = deep_chain_prims()
    mid, _ = boxes_of([3000])
    prims = np.concatenate([dc, mid, dc])
    lo = np.array([0, len(dc), len(dc) + len(mid), len(prims)], dtype=np.uint32)
    g = check_batch(gpu, orc, prims, lo, 1, "deep chains")
    assert g.segment_depth(0) == 31 and g.segment_depth(2) == 31


def test_many_segments_two_id_passes(gpu, orc):
    print("[batch] 300 segments", flush=True)
    sizes = S.batch_counts(300, 1, 400, seed=7)
    prims, lo = boxes_of(sizes, 100)
    moved = np.concatenate([S.corner_boxes(S.lbvh_scene(int(m), 100 + i)[1]) if m else np.zeros(0, dtype=P.AABB)
                            for i, m in enumerate(sizes)])
    check_batch(gpu, orc, prims, lo, 4, "300 segments", moved)


def test_batch_workload_slice(gpu, orc):
    print("[batch] 32 scenes of the batch workload", flush=True)
    counts = S.batch_counts(256)[:32]
    prims, lo = boxes_of(counts)
    moved = np.concatenate([S.corner_boxes(S.lbvh_scene(int(m), i)[1]) for i, m in enumerate(counts)])
    check_batch(gpu, orc, prims, lo, 4, "batch slice", moved)


def test_one_segment_batch_equals_single_build(gpu, orc):
    print("[batch] one segment", flush=True)
    prims, lo = boxes_of([20_011])
    a = P.GpuBvh.build(gpu, prims, 4).download()
    b = P.GpuBvh.build_batch(gpu, prims, lo, 4).download_segment(0)
    assert (a.order == b.order).all() and a.depth == b.depth
    assert (a.nodes.view(np.uint32)[np.array(a.reachable())] == b.nodes.view(np.uint32)[np.array(a.reachable())]).all()


def test_bad_segment_tables_are_refused(gpu):
    prims, _ = boxes_of([10])
    for lo in ([0, 11], [1, 10], [0, 6, 4, 10]):
        with pytest.raises(P.PhrError):
            P.GpuBvh.build_batch(gpu, prims, lo, 4)

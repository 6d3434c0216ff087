"""The build that allocates nothing (`phr_bvh_build_static`) against the
allocating build of the same boxes (`phr_bvh_build_on`), bit for bit.

Two calls in a row share the library's static scratch, the second over fewer
boxes than the first, so whatever the first call left there must not reach
the second's tree. Each tree is also held to the oracle's `Bvh::build`.
"""
import numpy as np
import pytest

import phiray_amd as P
from phiray_amd import scenes as S

pytestmark = pytest.mark.gpu


def static_tree(gpu, boxes: np.ndarray, leaf: int) -> P.Bvh:
    n = len(boxes)
    prims = P.GpuArray.from_(gpu, boxes)
    nodes = P.GpuArray.zeroed(gpu, P.NODE, max(2 * n - 1, 1))
    order = P.GpuArray.zeroed(gpu, np.uint32, max(n, 1))
    depth = P.GpuArray.zeroed(gpu, np.uint32, 1)
    P.GpuBvh.build_static(gpu, prims, nodes, order, depth, leaf)
    gpu.wait()
    return P.Bvh(nodes.read()[: max(2 * n - 1, 0)], order.read()[:n], int(depth.read()[0]))


def allocating_tree(gpu, boxes: np.ndarray, leaf: int) -> P.Bvh:
    return P.GpuBvh.build_on(gpu, P.GpuArray.from_(gpu, boxes), leaf).download()


def same(got: P.Bvh, want: P.Bvh, what: str) -> None:
    assert got.depth == want.depth, f"{what}: depth {got.depth} vs {want.depth}"
    assert (got.order == want.order).all(), f"{what}: order"
    r = np.array(want.reachable(), dtype=np.int64)
    assert r.size == 0 or (got.nodes[r].view(np.uint32) == want.nodes[r].view(np.uint32)).all(), f"{what}: nodes"


@pytest.mark.parametrize("leaf", [4, 1])
def test_two_calls_in_a_row_match_the_allocating_build(gpu, orc, leaf):
    for n, seed in ((200_003, 0.0), (65_536, 2.5)):
        print(f"[no-alloc] n={n} leaf={leaf}", flush=True)
        boxes = S.xorshift_prims_gpu_test(n, seed)
        got = static_tree(gpu, boxes, leaf)
        same(got, allocating_tree(gpu, boxes, leaf), f"n={n} vs build_on")
        nodes, order, depth = orc.build_bvh(boxes, leaf, P.NODE, 8)
        same(got, P.Bvh(nodes, order, depth), f"n={n} vs the oracle")


def test_small_and_empty_trees(gpu, orc):
    for n in (7, 1, 0):
        print(f"[no-alloc] n={n}", flush=True)
        boxes = S.xorshift_prims_unit(n) if n else np.zeros(0, dtype=P.AABB)
        got = static_tree(gpu, boxes, 4)
        if n == 0:
            assert got.depth == 0 and len(got.order) == 0
            continue
        same(got, allocating_tree(gpu, boxes, 4), f"n={n} vs build_on")


def test_sizes_past_the_scratch_are_refused(gpu):
    n = P.STATIC_BUILD_MAX + 1
    prims = P.GpuArray.zeroed(gpu, P.AABB, n)
    nodes = P.GpuArray.zeroed(gpu, P.NODE, 2 * n - 1)
    order = P.GpuArray.zeroed(gpu, np.uint32, n)
    with pytest.raises(P.PhrError, match="static scratch"):
        P.GpuBvh.build_static(gpu, prims, nodes, order)

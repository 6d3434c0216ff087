"""GPU LBVH == CPU LBVH, bit for bit (tests/bvh.rs, ported and widened).

Through the C ABI: `GpuBvh::build` on the device against the oracle's
`Bvh::build` on the same boxes — order, depth, the links of every reachable
node and their bounds (the reference's own comparison, tests/bvh.rs:21-29) —
then `update` after every box moves against the oracle's refit on the old
topology (tests/bvh.rs:47-52). Bit-exact: integer links and exact min/max.
"""
import numpy as np
import pytest

import phiray_amd as P
from phiray_amd import scenes as S

pytestmark = pytest.mark.gpu


def same(got: P.Bvh, nodes, order, depth, what=""):
    assert (got.order == order).all(), f"{what}: order"
    assert got.depth == depth, f"{what}: depth {got.depth} vs {depth}"
    want = P.Bvh(nodes, order, depth)
    r = np.array(want.reachable(), dtype=np.int64)
    g, c = got.nodes[r], nodes[r]
    bad = np.nonzero((g["li"] != c["li"]) | (g["ri"] != c["ri"]))[0]
    assert len(bad) == 0, f"{what}: links differ at nodes {r[bad[:8]]}"
    gb = g.view(np.uint32).reshape(-1, 8)
    cb = c.view(np.uint32).reshape(-1, 8)
    bad = np.nonzero((gb != cb).any(1))[0]
    assert len(bad) == 0, f"{what}: bounds differ at nodes {r[bad[:8]]}"


@pytest.mark.parametrize("n", [1, 2, 3, 257, 100_003])
def test_gpu_build_and_refit_match_cpu(gpu, orc, n):
    print(f"[lbvh] tests/bvh.rs n={n}", flush=True)
    p = S.xorshift_prims_gpu_test(n, 0.0)
    nodes, order, depth = orc.build_bvh(p, 4, P.NODE, 8)
    g = P.GpuBvh.build(gpu, p, 4)
    same(g.download(), nodes, order, depth, f"build n={n}")
    moved = S.xorshift_prims_gpu_test(n, 2.5)
    assert g.update(moved)
    same(g.download(), orc.refit(nodes, order, moved), order, depth, f"refit n={n}")


@pytest.mark.parametrize("n", [1, 2, 3, 7, 100, 5000])
@pytest.mark.parametrize("leaf", [1, 4])
def test_unit_sets(gpu, orc, n, leaf):
    print(f"[lbvh] src/bvh/test.rs n={n} leaf={leaf}", flush=True)
    p = S.xorshift_prims_unit(n)
    nodes, order, depth = orc.build_bvh(p, leaf, P.NODE)
    same(P.GpuBvh.build(gpu, p, leaf).download(), nodes, order, depth)


def test_all_equal_boxes(gpu, orc):
    print("[lbvh] 1000 equal boxes", flush=True)
    p = np.zeros(1000, dtype=P.AABB)
    p["hi"] = 1.0
    nodes, order, depth = orc.build_bvh(p, 1, P.NODE)
    same(P.GpuBvh.build(gpu, p, 1).download(), nodes, order, depth)


def test_depth_cap(gpu, orc):
    print("[lbvh] depth cap chain", flush=True)
    from tests.test_oracle_lbvh import deep_chain_prims

    p = deep_chain_prims()
    nodes, order, depth = orc.build_bvh(p, 1, P.NODE)
    assert depth == 31
    same(P.GpuBvh.build(gpu, p, 1).download(), nodes, order, depth)


def test_sorted_codes_and_count_change(gpu, orc):
    print("[lbvh] sorted codes", flush=True)
    p = S.xorshift_prims_gpu_test(40_000)
    g = P.GpuBvh.build(gpu, p, 4)
    order, srt = orc.sort_codes(orc.codes(p))
    assert (g.sorted_codes() == srt).all()
    assert g.update(p[:-1]) is False  # a new count asks for a rebuild


def test_lbvh_workload_and_refit(gpu, orc):
    print("[lbvh] 65536 triangles seed 0", flush=True)
    corners, moved = S.lbvh_scene(65536, 0)
    p, pm = S.corner_boxes(corners), S.corner_boxes(moved)
    nodes, order, depth = orc.build_bvh(p, 4, P.NODE, 8)
    g = P.GpuBvh.build(gpu, p, 4)
    same(g.download(), nodes, order, depth, "build 65536")
    assert g.update(pm)
    same(g.download(), orc.refit(nodes, order, pm), order, depth, "refit 65536")


@pytest.mark.parametrize("n", [1_310_724, 3_000_001])
def test_large_builds(gpu, orc, n):
    print(f"[lbvh] large n={n}", flush=True)
    corners, _ = S.lbvh_scene(n, 5)
    p = S.corner_boxes(corners)
    nodes, order, depth = orc.build_bvh(p, 4, P.NODE, 8)
    same(P.GpuBvh.build(gpu, p, 4).download(), nodes, order, depth, f"n={n}")


def test_rebuild_is_deterministic(gpu):
    print("[lbvh] rebuild twice", flush=True)
    corners, _ = S.lbvh_scene(200_000, 9)
    p = S.corner_boxes(corners)
    prims = P.GpuArray.from_(gpu, p)
    g = P.GpuBvh.build_on(gpu, prims, 4)
    a = g.download()
    for _ in range(3):
        g.rebuild()
        b = g.download()
        assert (a.order == b.order).all() and a.depth == b.depth
        r = np.array(a.reachable())
        assert (a.nodes[r].view(np.uint32) == b.nodes[r].view(np.uint32)).all()


@pytest.mark.parametrize("n,leaf", [(1, 4), (2, 1), (7, 1), (1000, 1), (5000, 4), (65536, 4)])
def test_build_compact_matches_cpu(gpu, orc, n, leaf):
    """Bvh::build_compact on the device: every node word equal to the oracle's
    breadth-first tree (src/bvh/build.rs:104-136)."""
    print(f"[lbvh] compact n={n} leaf={leaf}", flush=True)
    if n == 65536:
        corners, _ = S.lbvh_scene(65536, 0)
        prims = S.corner_boxes(corners)
    else:
        prims = S.xorshift_prims_unit(n)
    g = P.GpuBvh.build(gpu, prims, leaf)
    got = g.compact()
    want = np.zeros(2 * n - 1, dtype=P.NODE)
    order = np.zeros(n, dtype=np.uint32)
    d = np.zeros(1, dtype=np.uint32)
    m = orc.lib().orc_build_compact(orc.ptr(prims), n, leaf, orc.ptr(want), orc.ptr(order), orc.ptr(d))
    assert len(got) == m
    assert np.array_equal(got.view(np.uint32), want[:m].view(np.uint32))
    assert g.depth == int(d[0])


def test_build_compact_depth_cap(gpu, orc):
    """The 31-deep chain: nodes at the stack cap are packed in the compact tree too."""
    print("[lbvh] compact depth cap", flush=True)
    from tests.test_oracle_lbvh import deep_chain_prims

    prims = deep_chain_prims()
    n = len(prims)
    got = P.GpuBvh.build(gpu, prims, 1).compact()
    want = np.zeros(2 * n - 1, dtype=P.NODE)
    order = np.zeros(n, dtype=np.uint32)
    d = np.zeros(1, dtype=np.uint32)
    m = orc.lib().orc_build_compact(orc.ptr(prims), n, 1, orc.ptr(want), orc.ptr(order), orc.ptr(d))
    assert int(d[0]) == 31 and len(got) == m
    assert np.array_equal(got.view(np.uint32), want[:m].view(np.uint32))


@pytest.mark.parametrize("n", [5000, 1_310_724])
def test_every_sort_tile_gives_the_same_tree(gpu, orc, n):
    print(f"[lbvh] sort tiles of 2,048 and 4,096 keys, n={n}", flush=True)
    p = S.xorshift_prims_gpu_test(n, 0.0)
    nodes, order, depth = orc.build_bvh(p, 4, P.NODE, 8)
    g = P.GpuBvh.build(gpu, p, 4)
    for items in (16, 8):
        g.set_sort_tile(items)
        g.rebuild()
        same(g.download(), nodes, order, depth, f"tile {items} n={n}")
    items, ms = g.autotune(2)
    assert items in (8, 16) and g.sort_tile == items and min(ms) > 0
    g.rebuild()
    same(g.download(), nodes, order, depth, f"autotuned n={n}")

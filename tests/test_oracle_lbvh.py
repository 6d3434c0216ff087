"""The oracle's LBVH, pinned against the reference's golden data and the
mathematics (CPU only).

Pinning, per oracle function:
* orc_code / orc_morton — src/bvh/node.rs:119-125 asserts, and a bit-by-bit
  interleave of the truncated 10-bit coordinates written independently here;
* orc_sort_codes, orc_radix_model — the five sets of src/bvh/sort.rs:156-175
  with the (code, index) order from Python's sort (tests/golden), and each
  other on larger random sets;
* orc_links — brute-force binary radix trees on unique 64-bit keys
  (tests/golden), plus Karras' invariants;
* orc_assemble / orc_build / orc_refit — src/bvh/test.rs:25-77 invariants
  (every primitive once, containment, depth < 32, compact == build+compact),
  a brute-force leaf packing from the links, and boxes recomputed here from
  the primitives.
"""
import json
import pathlib

import numpy as np
import pytest

import phiray_amd as P
from phiray_amd import scenes as S

GOLDEN = json.loads((pathlib.Path(__file__).parent / "golden" / "golden.json").read_text())


def code_py(p):
    def spread(v):
        v = min(max(float(np.float32(v)), 0.0), 1.0)
        x = int(np.float32(np.float32(v) * np.float32(1023.0)))
        out = 0
        for b in range(10):
            out |= ((x >> b) & 1) << (3 * b)
        return out
    return (spread(p[0]) << 2) | (spread(p[1]) << 1) | spread(p[2])


def test_morton_asserts(orc):
    L = orc.lib()
    g = GOLDEN["morton"]
    ones = np.ones(3, np.float32)
    assert L.orc_code(orc.ptr(ones)) == g["ones_code"]
    a, b = (np.array(v, np.float32) for v in g["order_pairs"][0])
    assert L.orc_morton(orc.ptr(a), 0) < L.orc_morton(orc.ptr(b), 0)
    p1, p0 = g["group_pair"]
    hi = np.array(p1[:3], np.float32)
    lo = np.array(p0[:3], np.float32)
    assert L.orc_morton(orc.ptr(hi), p1[3]) < L.orc_morton(orc.ptr(lo), p0[3])


def test_morton_is_the_bit_interleave(orc):
    rng = np.random.default_rng(1)
    pts = rng.uniform(-0.2, 1.2, size=(2000, 3)).astype(np.float32)
    pts[:10] = [[0, 0, 0], [1, 1, 1], [0.5, 0.5, 0.5], [1e-7, 0.999, 0.0005], [-1, 2, 0.25],
                [0.0009775, 0.000977, 0.001], [0.999999, 0.5, 0.1], [1, 0, 1], [0, 1, 0], [0.3, 0.3, 0.3]]
    for p in pts:
        assert orc.lib().orc_code(orc.ptr(np.ascontiguousarray(p))) == code_py(p)


@pytest.mark.parametrize("case", range(5))
def test_sort_sets_from_the_reference(orc, case):
    s = GOLDEN["sort_sets"][case]
    codes = np.array(s["codes"], dtype=np.uint32)
    order, srt = orc.sort_codes(codes)
    assert order.tolist() == s["order"]
    assert srt.tolist() == s["sorted"]
    o2, s2 = orc.radix_model(codes)
    assert o2.tolist() == s["order"] and s2.tolist() == s["sorted"]


def test_parallel_sort_equals_serial(orc):
    rng = np.random.default_rng(7)
    c = rng.integers(0, 1 << 12, size=300_001, dtype=np.uint32)
    o1, s1 = orc.sort_codes(c, 1)
    o8, s8 = orc.sort_codes(c, 8)
    om, sm = orc.radix_model(c)
    assert (o1 == o8).all() and (s1 == s8).all()
    assert (o1 == om).all() and (s1 == sm).all()
    assert (np.diff(s1.astype(np.int64)) >= 0).all()


@pytest.mark.parametrize("case", range(8))
def test_links_are_the_binary_radix_tree(orc, case):
    t = GOLDEN["radix_trees"][case]
    srt = np.array(t["sorted"], dtype=np.uint32)
    links = orc.links(srt, P.LINK)
    got = [[int(l["left"]), int(l["right"]), int(l["first"]), int(l["last"])] for l in links]
    assert got == t["links"]


def check_tree(nodes, order, depth, prims, n):
    """src/bvh/test.rs:25-43 over the reachable nodes."""
    seen = np.zeros(n, dtype=np.int64)
    stack = [(0, 0)]
    maxd = 0
    lo, hi = prims["lo"], prims["hi"]
    while stack:
        i, d = stack.pop()
        maxd = max(maxd, d)
        nd = nodes[i]
        if nd["li"] & P.LEAF:
            first, end = int(nd["li"]) & 0x7FFFFFFF, int(nd["ri"])
            assert first < end
            for s in range(first, end):
                p = int(order[s])
                seen[p] += 1
                assert (nd["lo"] <= lo[p]).all() and (nd["hi"] >= hi[p]).all()
            # Boxes are exact: the union of the leaf's primitives.
            ps = order[first:end]
            assert (nd["lo"] == lo[ps].min(0)).all() and (nd["hi"] == hi[ps].max(0)).all()
        else:
            l, r = int(nd["li"]), int(nd["ri"])
            for c in (l, r):
                assert (nd["lo"] <= nodes[c]["lo"]).all() and (nd["hi"] >= nodes[c]["hi"]).all()
                stack.append((c, d + 1))
            assert (nd["lo"] == np.minimum(nodes[l]["lo"], nodes[r]["lo"])).all()
            assert (nd["hi"] == np.maximum(nodes[l]["hi"], nodes[r]["hi"])).all()
    assert (seen == 1).all(), "every primitive exactly once"
    assert depth == maxd < P.STACK


@pytest.mark.parametrize("n", [1, 2, 3, 7, 100, 5000])
@pytest.mark.parametrize("leaf", [1, 4])
def test_builds_valid_trees(orc, n, leaf):
    prims = S.xorshift_prims_unit(n)
    nodes, order, depth = orc.build_bvh(prims, leaf, P.NODE)
    check_tree(nodes, order, depth, prims, n)


def test_all_equal_codes_stay_shallow(orc):
    prims = np.zeros(1000, dtype=P.AABB)
    prims["hi"] = 1.0
    nodes, order, depth = orc.build_bvh(prims, 1, P.NODE)
    check_tree(nodes, order, depth, prims, 1000)
    assert depth <= 11


def deep_chain_prims():
    """Points whose Morton codes are 0, every power of two below 2^30, 2^30 - 1
    and sixteen more 1s: the Karras tree is a chain through all 30 code bits
    with the duplicates' slot-index subtree under it, deeper than 32."""
    pts = [(0.0, 0.0, 0.0), (1.0, 1.0, 1.0)]
    for m in range(10):
        u = (2 ** m + 0.5) / 1023.0
        pts += [(u, 0.0, 0.0), (0.0, u, 0.0), (0.0, 0.0, u)]
    pts += [(0.0, 0.0, 0.5 / 1023.0 + 1.0 / 1023.0)] * 16
    p = np.array(pts, dtype=np.float32)
    prims = np.zeros(len(p), dtype=P.AABB)
    prims["lo"], prims["hi"] = p, p
    return prims


def test_depth_cap_packs_deep_subtrees(orc):
    prims = deep_chain_prims()
    nodes, order, depth = orc.build_bvh(prims, 1, P.NODE)
    check_tree(nodes, order, depth, prims, len(prims))
    assert depth == P.STACK - 1


def pack_brute(links, n, leaf):
    """The leaf packing restated from its definition: walk the links top-down."""
    packed, depth = {}, 0
    stack = [(0, 0)]
    while stack:
        i, d = stack.pop()
        depth = max(depth, d)
        if i >= n - 1:
            continue
        l = links[i]
        if l["last"] - l["first"] < leaf or d + 1 >= 32:
            packed[i] = (int(l["first"]), int(l["last"]) + 1)
            continue
        stack += [(int(l["left"]), d + 1), (int(l["right"]), d + 1)]
    return packed, depth


def test_assemble_matches_brute_force_packing(orc):
    prims = S.xorshift_prims_gpu_test(3000)
    c = orc.codes(prims)
    order, srt = orc.sort_codes(c)
    links = orc.links(srt, P.LINK)
    nodes = np.zeros(2 * len(prims) - 1, dtype=P.NODE)
    depth = orc.lib().orc_assemble(orc.ptr(links), len(prims), 4, orc.ptr(nodes))
    packed, d2 = pack_brute(links, len(prims), 4)
    assert depth == d2
    for i, (f, e) in packed.items():
        assert nodes[i]["li"] == (f | P.LEAF) and nodes[i]["ri"] == e


def test_compaction_keeps_the_tree(orc):
    prims = S.xorshift_prims_unit(5000)
    n = len(prims)
    nodes, order, depth = orc.build_bvh(prims, 4, P.NODE)
    full = nodes.copy()
    m = orc.lib().orc_compact(orc.ptr(full), len(full))
    assert m < len(nodes) // 2 + 2
    direct = np.zeros(2 * n - 1, dtype=P.NODE)
    order2 = np.zeros(n, dtype=np.uint32)
    d = np.zeros(1, dtype=np.uint32)
    m2 = orc.lib().orc_build_compact(orc.ptr(prims), n, 4, orc.ptr(direct), orc.ptr(order2), orc.ptr(d))
    assert m2 == m and (order2 == order).all() and int(d[0]) == depth
    assert (direct[:m].view(np.uint32) == full[:m].view(np.uint32)).all()


def test_refit_is_the_union_after_motion(orc):
    prims = S.xorshift_prims_gpu_test(2000, 0.0)
    moved = S.xorshift_prims_gpu_test(2000, 2.5)
    nodes, order, depth = orc.build_bvh(prims, 4, P.NODE)
    re = orc.refit(nodes, order, moved)
    check_tree(re, order, depth, moved, 2000)

"""The oracle's triangle test and walk (CPU only).

Pinning:
* orc_tri_t / orc_ray_space — the three tests of src/trace/tri.rs:188-251:
  the 48 x 48 card's shared diagonal has no hole and some rays hit both
  halves; 200,000 rays aimed at the shared edge (xorshift seed 0x2545F491)
  always hit one side while Möller–Trumbore drops some; t, u, v agree with
  Möller–Trumbore to 1e-3 and the weights rebuild the hit point;
* orc_raycast — brute force: the closest hit's t is the least t over every
  triangle (the same test, so exactly equal), any-hit reports a hit exactly
  when some triangle is nearer than tmax. The tie order among equal-t hits
  is pinned only by the restated walk (parity unpinned there).
"""
import numpy as np
import pytest

import phiray_amd as P
from phiray_amd import scenes as S

W = 48


def card():
    half, y = 3.0, 2.0
    q = np.array([[-half, y, -half], [half, y, -half], [half, y, half], [-half, y, half]], dtype=np.float32)
    return [np.array([q[0], q[1], q[2]]), np.array([q[0], q[2], q[3]])]


def preview_ray(x, y):
    sx = np.float32((np.float32(x) + np.float32(0.5)) / np.float32(W) - np.float32(0.5))
    sy = np.float32((np.float32(y) + np.float32(0.5)) / np.float32(W) - np.float32(0.5))
    d = np.array([sx, 4.0, -sy], dtype=np.float32)
    l = np.float32(np.sqrt(np.float32(d[0] * d[0] + d[1] * d[1] + d[2] * d[2])))
    return np.array([0.0, -4.0, 0.0], dtype=np.float32), (d / l).astype(np.float32)


def tri_t(orc, t, o, d):
    rs = np.zeros(5, dtype=np.uint32)
    orc.lib().orc_ray_space(orc.ptr(d), orc.ptr(rs))
    out = np.zeros(3, dtype=np.float32)
    hit = orc.lib().orc_tri_t(orc.ptr(np.ascontiguousarray(t[0])), orc.ptr(np.ascontiguousarray(t[1])),
                              orc.ptr(np.ascontiguousarray(t[2])), orc.ptr(o), orc.ptr(rs), orc.ptr(out))
    return out if hit else None


def mt(orc, t, o, d):
    out = np.zeros(3, dtype=np.float32)
    hit = orc.lib().orc_moller_trumbore(orc.ptr(np.ascontiguousarray(t[0])), orc.ptr(np.ascontiguousarray(t[1])),
                                        orc.ptr(np.ascontiguousarray(t[2])), orc.ptr(o), orc.ptr(d), orc.ptr(out))
    return out if hit else None


def test_no_ray_falls_through_a_quads_shared_edge(orc):
    holes, both = [], 0
    for x in range(W):
        y = W - 1 - x
        o, d = preview_ray(x, y)
        k = sum(tri_t(orc, t, o, d) is not None for t in card())
        if k == 0:
            holes.append((x, y))
        both += k == 2
    assert not holes, f"the shared diagonal leaks at {holes}"
    assert both > 0


def test_a_ray_aimed_at_a_shared_edge_hits_one_side_or_both(orc):
    a, b = card()
    p0, p2 = a[0], a[2]
    x = 0x2545F491

    def rand():
        nonlocal x
        x ^= (x << 13) & 0xFFFFFFFF
        x ^= x >> 17
        x ^= (x << 5) & 0xFFFFFFFF
        return np.float32(np.float32(x % 1000003) / np.float32(1000003.0))

    dropped = 0
    f32 = np.float32
    for _ in range(200_000):
        f = f32(f32(0.02) + f32(0.96) * rand())
        aim = np.array([p0[0] + f * (p2[0] - p0[0]), p0[1], p0[2] + f * (p2[2] - p0[2])], dtype=np.float32)
        side = f32(-1.0) if rand() < f32(0.5) else f32(1.0)
        o = np.array([f32(8.0) * (rand() - f32(0.5)), 0.0, 0.0], dtype=np.float32)
        o[1] = p0[1] + side * (f32(0.5) + f32(9.0) * rand())
        o[2] = f32(8.0) * (rand() - f32(0.5))
        d = (aim - o).astype(np.float32)
        one = tri_t(orc, a, o, d) is not None
        two = tri_t(orc, b, o, d) is not None
        assert one or two, f"a hole at {aim} from {o}"
        old = mt(orc, a, o, d) is not None or mt(orc, b, o, d) is not None
        dropped += not old
    assert dropped > 0


def test_the_hit_is_the_one_moller_trumbore_found(orc):
    a = card()[0]
    agreed = 0
    for x in range(W):
        for y in range(W):
            o, d = preview_ray(x, y)
            n, m = tri_t(orc, a, o, d), mt(orc, a, o, d)
            if n is None or m is None:
                continue
            agreed += 1
            assert abs(n[0] - m[0]) < 1e-3 * m[0]
            assert abs(n[1] - m[1]) < 1e-3 and abs(n[2] - m[2]) < 1e-3
            w = 1.0 - n[1] - n[2]
            px = w * a[0][0] + n[1] * a[1][0] + n[2] * a[2][0]
            assert abs(px - (o[0] + n[0] * d[0])) < 1e-3
    assert agreed > 400


def brute(orc, tris9, ray):
    best = np.float32(3e38)
    o = np.ascontiguousarray(ray["o"], dtype=np.float32)
    d = np.ascontiguousarray(ray["d"], dtype=np.float32)
    for s in range(len(tris9) // 9):
        t = tris9[9 * s: 9 * s + 9].reshape(3, 3)
        h = tri_t(orc, t, o, d)
        if h is not None and h[0] < best:
            best = h[0]
    return best


@pytest.mark.parametrize("seed", [0, 1])
def test_walk_finds_the_nearest_triangle(orc, seed):
    corners, _ = S.lbvh_scene(600, seed)
    tris = S.tris_from_corners(corners)
    sc = orc.Scene(tris, 4, P.NODE)
    rays = S.pinhole_rays(24, 24)
    hits = orc.raycast(sc.nodes, sc.tri9, rays, P.HIT)
    for r, h in zip(rays, hits):
        want = brute(orc, sc.tri9, r)
        if want >= np.float32(3e38):
            assert h["tri"] == P.NONE
        else:
            assert h["tri"] != P.NONE and h["t"] == want
            t = sc.tri9[9 * h["tri"]: 9 * h["tri"] + 9].reshape(3, 3)
            assert tri_t(orc, t, r["o"].copy(), r["d"].copy())[0] == want


def test_any_hit_agrees_with_brute_force(orc):
    corners, _ = S.lbvh_scene(600, 3)
    sc = orc.Scene(S.tris_from_corners(corners), 4, P.NODE)
    rays = S.pinhole_rays(20, 20)
    closest = orc.raycast(sc.nodes, sc.tri9, rays, P.HIT)
    rays["tmax"] = np.where(closest["tri"] != P.NONE, np.minimum(closest["t"], np.float32(1e30)) * np.float32(1.5), np.float32(3e38))
    rays["tmax"][::3] = np.float32(1.4)  # shorter than most of the cloud
    anyh = orc.raycast(sc.nodes, sc.tri9, rays, P.HIT, any_hit=True)
    for r, h in zip(rays, anyh):
        want = brute(orc, sc.tri9, r)
        assert (h["tri"] != P.NONE) == bool(want < r["tmax"])
        if h["tri"] != P.NONE:
            assert h["t"] < r["tmax"]

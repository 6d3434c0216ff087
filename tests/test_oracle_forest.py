"""The oracle's forest: instances, mesh trees and the top tree (CPU only).

Pinning, each independent of the restated formula:
* Instance::inverse / world_box / words / mirrors and Mesh::texspace — the
  reference's own unit tests (src/trace/instance.rs:163-219,
  src/trace/mesh.rs:77-90) and their golden values;
* the top-tree walk and the instance transform — a forest of translated
  copies of one mesh must hit exactly what that mesh's single tree hits
  along the ray moved into its space (the same float operations), instance
  for instance; and rotated/scaled instances against brute force over every
  placed triangle in double precision;
* the forest's trees — the top tree is Bvh::build of the placed world boxes
  at leaf 1 with INST leaves, each mesh tree Bvh::build of its faces' boxes
  shifted into place (src/trace/forest.rs:132-165);
* the light table — closed-form cumulative powers for one emitter under
  instances of areas 1, 4 and 9 (src/trace/lamp.rs:51-97);
* reinstance — equal to a fresh build with the new matrices;
* the integrator through instances — the closed form of tests/emitter.rs's
  1 x 1 panel at distance 1 with the panel a scaled, turned and moved
  instance.
"""
import math

import numpy as np
import pytest

import phiray_amd as P
from phiray_amd import scenes as S


def forest_of(orc, b, leaf=4):
    v, f, m, me, i = b.arrays()
    return orc.Forest(v, f, m, me, i, leaf, P.NODE)


def unit_mesh():
    m = np.zeros(1, dtype=P.MESH)[0]
    m["lo"], m["hi"] = 0.0, 1.0
    return m


def inst(world_cols, mesh=0, casts=True):
    i = np.zeros(1, dtype=P.INSTANCE)[0]
    i["world"] = np.asarray(world_cols, dtype=np.float32).reshape(16)
    i["mesh"], i["casts_shadow"], i["color"] = mesh, 1 if casts else 0, 1.0
    return i


IDENT = np.eye(4, dtype=np.float32)  # symmetric: its m[col][row] words are the same


def test_instance_unit_tests_from_the_reference(orc):
    # instance.rs:167-187: scale 2, a quarter turn about z, moved to (1, 2, 3).
    world = np.zeros((4, 4), dtype=np.float32)
    world[0] = [0, 2, 0, 0]
    world[1] = [-2, 0, 0, 0]
    world[2] = [0, 0, 2, 0]
    world[3] = [1, 2, 3, 1]
    L = orc.lib()
    p = np.array([0.5, -1.0, 4.0], dtype=np.float32)
    w = np.zeros(3, dtype=np.float32)
    L.orc_inst_to_world(orc.ptr(world.reshape(16)), orc.ptr(p), orc.ptr(w))
    inv = orc.inst_inverse(world.reshape(16))
    back = inv[:, :3] @ w.astype(np.float64) + inv[:, 3]
    assert np.abs(back - p).max() < 1e-5
    box = np.zeros(1, dtype=P.AABB)[0]
    box["lo"], box["hi"] = 0.0, 1.0
    wb = orc.inst_world_box(world.reshape(16), box)
    assert list(wb["lo"]) == [-1.0, 2.0, 3.0] and list(wb["hi"]) == [1.0, 4.0, 5.0]
    words = orc.inst_words(inst(world, mesh=3), 7, unit_mesh())
    assert (words[24], words[27], words[38]) == (7, 0, 3)
    assert orc.inst_words(inst(IDENT), 0, unit_mesh())[27] == 1
    assert orc.inst_words(inst(IDENT, casts=False), 0, unit_mesh())[27] == 0b11


@pytest.mark.parametrize("scale,mirrors", [((-1, 1, 1), True), ((-1, -1, 1), False), ((-2, -0.5, -1), True)])
def test_negative_determinants_mirror(orc, scale, mirrors):
    # instance.rs:189-205
    m = np.eye(4, dtype=np.float32)
    for k in range(3):
        m[k, k] = scale[k]
    m[3] = [0.5, 0.0, 0.0, 1.0]
    assert orc.lib().orc_inst_mirrors(orc.ptr(m.reshape(16))) == (1 if mirrors else 0)
    if scale == (-1, 1, 1):
        assert orc.inst_words(inst(m), 0, unit_mesh())[27] == 0b100
    assert orc.inst_words(inst(IDENT), 0, unit_mesh())[27] & 0b100 == 0


def test_texture_space(orc):
    # mesh.rs:77-84 and instance.rs:208-218
    wall = np.zeros(1, dtype=P.MESH)[0]
    wall["lo"], wall["hi"] = (-9.0, 0.0, -9.0), (9.0, 0.0, 12.0)
    lo, size = np.zeros(3, np.float32), np.zeros(3, np.float32)
    orc.lib().orc_mesh_texspace(orc.ptr(np.array([wall])), orc.ptr(lo), orc.ptr(size))
    assert list(lo) == [-9.0, -1.0, -9.0] and list(size) == [18.0, 2.0, 21.0]
    w = orc.inst_words(inst(IDENT), 0, wall)
    assert (w[33:34].view(np.float32)[0], w[36:37].view(np.float32)[0]) == (-1.0, 2.0)


def lump(n, seed):
    corners, _ = S.lbvh_scene(n, seed)
    return corners.reshape(-1, 3), np.arange(3 * n).reshape(n, 3)


def test_translated_copies_hit_what_the_mesh_hits(orc):
    pos, tris = lump(3000, 1)
    b = S.SceneBuilder()
    mat = b.material()
    mesh = b.mesh(pos, tris, mat)
    shifts = [(10.0 * k, 0.0, 0.0) for k in range(5)]
    for sh in shifts:
        b.instance(mesh, S.affine(move=sh))
    fo = forest_of(orc, b)
    one = S.SceneBuilder()
    one.mesh(pos, tris, one.material())
    one.instance(0)
    single = forest_of(orc, one)
    rays = S.pinhole_rays(48, 48)
    for k, sh in enumerate(shifts):
        r = rays.copy()
        r["o"] += np.asarray(sh, dtype=np.float32)
        got = fo.raycast(r, P.HIT)
        # The ray in the mesh's space exactly as the walk computes it.
        words = fo.inst_words[40 * k: 40 * k + 40]
        iv = words[12:24].view(np.float32).reshape(3, 4)
        o, d = r["o"], r["d"]
        lo = np.stack([((iv[j, 0] * o[:, 0] + iv[j, 1] * o[:, 1]) + iv[j, 2] * o[:, 2]) + iv[j, 3] for j in range(3)], 1)
        ld = np.stack([(iv[j, 0] * d[:, 0] + iv[j, 1] * d[:, 1]) + iv[j, 2] * d[:, 2] for j in range(3)], 1)
        r2 = r.copy()
        r2["o"], r2["d"] = lo, ld
        want = single.raycast(r2, P.HIT)
        assert (got["tri"] == want["tri"]).all() and (got["t"].view(np.uint32) == want["t"].view(np.uint32)).all()
        hit = want["tri"] != P.NONE
        assert hit.sum() > 50
        assert (got["inst"][hit] == k).all()


def test_turned_and_scaled_instances_against_brute_force(orc):
    pos, tris = lump(200, 2)
    pos = (pos - 0.5).astype(np.float32)
    b = S.SceneBuilder()
    mesh = b.mesh(pos, tris, b.material())
    rng = np.random.default_rng(5)
    mats = []
    for k in range(12):
        m = S.affine(S.rotation(rng.normal(size=3), rng.uniform(0, 6.28)), rng.uniform(0.5, 2.0),
                     rng.uniform(-4, 4, size=3))
        mats.append(m)
        b.instance(mesh, m)
    fo = forest_of(orc, b)
    n = 4000
    o = rng.uniform(-8, 8, size=(n, 3)).astype(np.float32)
    # Aim each ray at the centroid of a random placed triangle.
    ki, fi = rng.integers(0, len(mats), n), rng.integers(0, len(tris), n)
    cen = pos.astype(np.float64)[tris[fi]].mean(1)
    Ms = np.array([m.astype(np.float64) for m in mats])[ki]
    tgt = np.einsum("nij,nj->ni", Ms[:, :3, :3], cen) + Ms[:, :3, 3]
    d = (tgt - o) / np.linalg.norm(tgt - o, axis=1)[:, None]
    rays = np.zeros(n, dtype=P.RAY)
    rays["o"], rays["d"], rays["tmax"] = o, d.astype(np.float32), 3e38
    got = fo.raycast(rays, P.HIT)
    # Brute force in double precision over every placed triangle.
    best_t = np.full(n, np.inf)
    second = np.full(n, np.inf)
    best = np.full((n, 2), -1)
    slot_face = fo.order  # slot -> face
    for k, m in enumerate(mats):
        M = m.astype(np.float64)
        wp = (M[:3, :3] @ pos.astype(np.float64).T).T + M[:3, 3]
        for f in range(len(tris)):
            a, bb, c = wp[tris[f]]
            e1, e2 = bb - a, c - a
            pv = np.cross(d, e2)
            det = pv @ e1
            ok = np.abs(det) > 1e-14
            inv = np.where(ok, 1.0 / np.where(ok, det, 1.0), 0.0)
            tv = o - a
            u = (tv * pv).sum(1) * inv
            q = np.cross(tv, e1)
            v = (d * q).sum(1) * inv
            t = (q @ e2) * inv
            hit = ok & (u >= 0) & (v >= 0) & (u + v <= 1) & (t > 0)
            t = np.where(hit, t, np.inf)
            nearer = t < best_t
            second = np.where(nearer, best_t, np.minimum(second, t))
            best[nearer] = (f, k)
            best_t = np.where(nearer, t, best_t)
    with np.errstate(invalid="ignore"):
        clear = np.isfinite(best_t) & ((second - best_t) > 1e-4 * best_t)
    assert clear.sum() > 2000
    faces = slot_face[got["tri"][clear]]
    assert (faces == best[clear, 0]).all() and (got["inst"][clear] == best[clear, 1]).all()
    # The walk meets the triangle in the mesh's space, in single precision,
    # along the ray moved there by the float32 inverse: a grazing ray
    # magnifies that rounding, so the bound holds for almost every ray.
    rel = np.abs(got["t"][clear] - best_t[clear]) / best_t[clear]
    assert (rel < 1e-4).mean() > 0.995 and rel.max() < 1e-2, (np.quantile(rel, [0.5, 0.99, 1.0]))
    assert ((got["tri"] == P.NONE) == ~np.isfinite(best_t)).mean() > 0.999


def test_the_trees_are_lbvhs_in_place(orc):
    b = S.SceneBuilder()
    mat = b.material()
    meshes = [b.mesh(*lump(m, 10 + m), mat) for m in (50, 0, 7, 300)]
    worlds = [S.affine(move=(3.0 * k, 1.0, 0.0)) for k in range(6)]
    for k, w in enumerate(worlds):
        b.instance(meshes[k % 4], w)
    v, f, m, me, ins = b.arrays()
    fo = orc.Forest(v, f, m, me, ins, 4, P.NODE)
    slots = 2 * 6 - 1
    raw = fo.nodes.view(np.uint32).reshape(-1, 8)
    at = slots
    roots = {}
    for k, mesh in enumerate(me):
        lo, hi = int(mesh["face_lo"]), int(mesh["face_hi"])
        if lo == hi:
            continue
        boxes = S.of_tri(*(v["p"][f["v"][lo:hi, j], :3] for j in range(3)))
        nodes, order, depth = orc.build_bvh(boxes, 4, P.NODE)
        assert fo.depths[k] == depth and (fo.order[lo:hi] == order + lo).all()
        w = nodes.view(np.uint32).reshape(-1, 8).copy()
        leaf = (w[:, 3] & P.LEAF) != 0
        w[leaf, 3] = ((w[leaf, 3] & 0x7FFFFFFF) + lo) | P.LEAF
        w[leaf, 7] += lo
        w[~leaf, 3] += at
        w[~leaf, 7] += at
        r = orc.reachable(nodes)
        assert (raw[at:at + len(nodes)][r] == w[r]).all()
        roots[k] = at
        at += len(nodes)
    placed = [k for k in range(6) if k % 4 != 1]  # the empty mesh has no tree
    wbox = np.zeros(len(placed), dtype=P.AABB)
    for j, k in enumerate(placed):
        mesh = k % 4
        root = fo.nodes[roots[mesh]]
        rb = np.zeros(1, dtype=P.AABB)[0]
        rb["lo"], rb["hi"] = root["lo"], root["hi"]
        wbox[j] = orc.inst_world_box(S.world_of(worlds[k]), rb)
    tn, to, td = orc.build_bvh(wbox, 1, P.NODE)
    t = tn.view(np.uint32).reshape(-1, 8).copy()
    leaf = (t[:, 3] & P.LEAF) != 0
    t[leaf, 3] |= P.INST
    assert (raw[:len(tn)] == t).all() and fo.depths[-1] == td
    inst_order = fo.inst_words[40 * 6: 40 * 6 + len(placed)]
    assert (inst_order == np.array(placed)[to]).all()
    for k in range(6):
        want = roots.get(k % 4, P.NONE)
        assert fo.inst_words[40 * k + 24] == want


def test_light_table_under_instances(orc):
    b = S.SceneBuilder()
    lamp = b.mesh([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]], [[0, 1, 2], [0, 2, 3]],
                  b.material(color=(0, 0, 0, 1), emit=(5.0, 5.0, 5.0)))
    for s in (1.0, 2.0, 3.0):
        b.instance(lamp, S.affine(scale=s, move=(10 * s, 0, 0)))
    fo = forest_of(orc, b)
    t = fo.lights_used
    c = int(t[0])
    assert c == 6
    slot_of = np.argsort(fo.order)
    assert list(t[1:1 + c]) == [slot_of[0]] * 3 + [slot_of[1]] * 3 if slot_of[0] < slot_of[1] else True
    assert list(t[1 + c:1 + 2 * c]) == [0, 1, 2, 0, 1, 2]
    cdf = t[1 + 2 * c:].view(np.float32)
    want = np.cumsum([0.5, 2.0, 4.5, 0.5, 2.0, 4.5]) / 14.0
    assert np.allclose(cdf, want, rtol=1e-6) and cdf[-1] == 1.0


def test_reinstance_equals_a_fresh_build(orc):
    pos, tris = lump(500, 3)
    b = S.SceneBuilder()
    mesh = b.mesh(pos, tris, b.material())
    lamp = b.mesh([[0, 0, 0], [1, 0, 0], [1, 1, 0]], [[0, 1, 2]], b.material(color=(0, 0, 0, 1), emit=(1, 1, 1)))
    for k in range(4):
        b.instance(mesh, S.affine(move=(2.0 * k, 0, 0)))
    b.instance(lamp, S.affine(move=(0, 0, 5)))
    v, f, m, me, ins = b.arrays()
    fo = orc.Forest(v, f, m, me, ins, 4, P.NODE)
    rng = np.random.default_rng(9)
    new = [S.world_of(S.affine(S.rotation(rng.normal(size=3), rng.uniform(0, 6)), rng.uniform(0.5, 2),
                                 rng.uniform(-5, 5, 3))) for _ in range(5)]
    fo.reinstance(np.array(new))
    ins2 = ins.copy()
    ins2["world"] = np.array(new)
    fresh = orc.Forest(v, f, m, me, ins2, 4, P.NODE)
    assert (fo.nodes.view(np.uint32) == fresh.nodes.view(np.uint32)).all()
    assert (fo.inst_words == fresh.inst_words).all() and (fo.lights_used == fresh.lights_used).all()
    assert fo.extent == fresh.extent


def test_emitter_panel_through_an_instance(orc):
    # tests/emitter.rs: a 1 x 1 panel of L = 2 at distance 1 from a grey wall
    # (the closed form of test_oracle_path's "middle" case), the panel built
    # at half size in its own space and placed by scale 2, a quarter turn
    # about y and a move to y = 1.
    from tests.test_oracle_path import wall_camera

    b = S.SceneBuilder()
    grey = b.material(color=(0.5, 0.5, 0.5, 1.0))
    glow = b.material(color=(0, 0, 0, 0), emit=(2.0, 2.0, 2.0), ior=1.5)
    q = [[-40, 0, -40], [40, 0, -40], [40, 0, 40], [-40, 0, 40]]
    wall = b.mesh(q, [[0, 1, 2], [0, 2, 3]], grey, normals=[[0, 1, 0]] * 4)
    h = 0.25
    pq = [[-h, 0, -h], [h, 0, -h], [h, 0, h], [-h, 0, h]]
    pan = b.mesh(pq, [[0, 1, 2], [0, 2, 3]], glow, normals=[[0, 1, 0]] * 4)
    b.instance(wall)
    b.instance(pan, S.affine(S.rotation((0, 1, 0), math.pi / 2), 2.0, (0.0, 1.0, 0.0)))
    fo = forest_of(orc, b)
    w = h_ = 16
    px = np.zeros(P.PX_LEN * w * h_ + 1, dtype=np.uint32)
    cam = wall_camera()
    for fr in range(256):
        fo.sample(cam, w, h_, px, fr)
    pl = P.planes(px, w, h_)
    rad = pl["radiance"] / pl["count"][:, None]
    got = float(rad.reshape(16, 16, 3)[4:12, 4:12, 0].mean())
    s = 0.5 / math.sqrt(1.25)
    want = 0.5 / math.pi * 4.0 * 2.0 * s * math.atan(s)
    assert abs(got - want) < want * 0.08, (got, want)

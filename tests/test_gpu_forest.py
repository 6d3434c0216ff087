"""The forest on the GPU == the oracle's forest (src/trace/forest.rs, instance.rs, pose.rs).

Through the C ABI, on one scene of meshes under instances (turned, scaled,
one mirroring, one casting no shadow, an emissive quad, a ground plane):
* the instance words (Forest::inst_words) and every reachable node of the
  node array (the top tree first, then each mesh tree in place) bit for bit,
  the face order, the light table and the extent;
* closest-hit and any-hit raycasts bit for bit (t, slot, u, v, instance);
* the path tracer within 1e-5 relative under identical per-pixel streams;
* reinstance (new matrices) == the oracle's reinstance; repose (moved
  vertices) refits every mesh tree to the oracle's refit of its boxes and
  rebuilds the top tree over the new world boxes; retree == a fresh oracle
  forest of the moved pose;
* a forest of one mesh under one identity instance traces exactly as the
  triangle tracer does (same hits, same radiance bits).
"""
import numpy as np
import pytest

import phiray_amd as P
from phiray_amd import scenes as S

pytestmark = pytest.mark.gpu


def forest_scene():
    b = S.SceneBuilder()
    grey = b.material(color=(0.6, 0.6, 0.6, 1.0))
    red = b.material(color=(0.8, 0.2, 0.1, 1.0))
    glow = b.material(color=(0, 0, 0, 1), emit=(6.0, 6.0, 6.0))
    corners, _ = S.lbvh_scene(1500, 4)
    lump = b.mesh((corners.reshape(-1, 3) - 0.5) * 2.0, np.arange(4500).reshape(1500, 3), red)
    vs, fs = S.icosphere(2)
    ball = b.mesh(vs, fs, grey, normals=vs)
    bp, bt = S.box_mesh()
    box = b.mesh(bp, bt, grey)
    h = 6.0
    ground = b.mesh([[-h, -h, 0], [h, -h, 0], [h, h, 0], [-h, h, 0]], [[0, 1, 2], [0, 2, 3]], grey,
                    normals=[[0, 0, 1]] * 4)
    lamp = b.mesh([[-1, -1, 0], [-1, 1, 0], [1, 1, 0], [1, -1, 0]], [[0, 1, 2], [0, 2, 3]], glow)
    rng = np.random.default_rng(11)
    for k in range(3):
        b.instance(lump, S.affine(S.rotation(rng.normal(size=3), rng.uniform(0, 6)), rng.uniform(0.5, 1.2),
                                  (-2.5 + 2.5 * k, 0.5, 1.2)), casts_shadow=(k != 1))
    b.instance(lump, S.affine(None, (-1.0, 1.0, 1.0), (0.0, -2.0, 1.0)))  # mirrored
    b.instance(ball, S.affine(None, 0.7, (2.0, 2.0, 0.7)))
    b.instance(ball, S.affine(S.rotation((0, 0, 1), 0.3), 0.5, (-2.0, 2.0, 0.5)))
    b.instance(box, S.affine(S.rotation((1, 1, 0), 0.6), 0.4, (0.0, 2.5, 0.6)))
    b.instance(ground)
    b.instance(lamp, S.affine(move=(0.0, 0.0, 5.0)))
    return b


@pytest.fixture(scope="module")
def scene(gpu, orc):
    arrs = forest_scene().arrays()
    t = P.Tracer.from_scene(gpu, *arrs, leaf=4)
    fo = orc.Forest(*arrs, 4, P.NODE)
    return arrs, t, fo


def same_nodes(got, want, what):
    r = np.array(P.Bvh(want, None, 0).reachable(), dtype=np.int64)
    g = got.view(np.uint32).reshape(-1, 8)[r]
    w = want.view(np.uint32).reshape(-1, 8)[r]
    bad = np.nonzero((g != w).any(1))[0]
    assert len(bad) == 0, f"{what}: nodes differ at {r[bad[:8]]}"


def reachable_forest(nodes, insts_words, ninsts):
    """Reachable nodes of a forest: the top tree, then every placed mesh tree."""
    out = set()
    stack = [0]
    li, ri = nodes["li"], nodes["ri"]
    while stack:
        i = stack.pop()
        out.add(i)
        if li[i] & P.LEAF:
            continue
        stack += [int(li[i]), int(ri[i])]
    for k in range(ninsts):
        root = int(insts_words[40 * k + 24])
        if root == P.NONE:
            continue
        stack = [root]
        while stack:
            i = stack.pop()
            out.add(i)
            if li[i] & P.LEAF:
                continue
            stack += [int(li[i]), int(ri[i])]
    return np.array(sorted(out), dtype=np.int64)


def check_forest(t, fo, what):
    n_i = len(fo.insts)
    words = t.instances()
    assert (words == fo.inst_words).all(), f"{what}: instance words differ"
    got = t.nodes()
    assert len(got) == len(fo.nodes)
    r = reachable_forest(fo.nodes, fo.inst_words, n_i)
    g = got.view(np.uint32).reshape(-1, 8)[r]
    w = fo.nodes.view(np.uint32).reshape(-1, 8)[r]
    bad = np.nonzero((g != w).any(1))[0]
    assert len(bad) == 0, f"{what}: nodes differ at {r[bad[:8]]}"
    assert (t.order() == fo.order).all(), f"{what}: face order"
    assert (t.lights() == fo.lights_used).all(), f"{what}: light table"
    assert t.extent == np.float32(fo.extent)


def rays_at(fo, n, seed, spread=6.0):
    rng = np.random.default_rng(seed)
    o = rng.uniform(-spread, spread, size=(n, 3)).astype(np.float32)
    o[:, 2] = np.abs(o[:, 2]) + 0.2
    tgt = rng.uniform(-3, 3, size=(n, 3))
    tgt[:, 2] = rng.uniform(0, 2, n)
    d = (tgt - o) / np.linalg.norm(tgt - o, axis=1)[:, None]
    r = np.zeros(n, dtype=P.RAY)
    r["o"], r["d"], r["tmax"] = o, d.astype(np.float32), np.float32(3e38)
    return r


def same_hits(got, want, what):
    for k in ("tri", "inst"):
        bad = np.nonzero(got[k] != want[k])[0]
        assert len(bad) == 0, f"{what}: {k} differs at {bad[:8]}"
    for k in ("t", "u", "v"):
        bad = np.nonzero(got[k].view(np.uint32) != want[k].view(np.uint32))[0]
        assert len(bad) == 0, f"{what}: {k} differs at {bad[:8]}"


def test_forest_tables_match_the_oracle(scene):
    print("[forest] build", flush=True)
    _, t, fo = scene
    check_forest(t, fo, "build")


@pytest.mark.parametrize("any_hit", [False, True])
def test_forest_raycast_is_bit_exact(scene, any_hit):
    print(f"[forest] raycast any={any_hit}", flush=True)
    _, t, fo = scene
    r = rays_at(fo, 20_000, 3)
    got = t.raycast(r, any_hit)
    want = fo.raycast(r, P.HIT, any_hit)
    assert (want["tri"] != P.NONE).sum() > 5000
    same_hits(got, want, f"raycast any={any_hit}")


@pytest.mark.parametrize("spp", [1, 2])
def test_forest_path_matches_the_twin(gpu, scene, spp):
    print(f"[forest] path spp={spp}", flush=True)
    arrs, _, fo = scene
    t = P.Tracer.from_scene(gpu, *arrs, leaf=4)
    w, h = 48, 32
    cam = S.path_camera(w, h, eye=(0.0, -7.0, 3.5), target=(0.0, 0.5, 0.8), rr_from=2)
    t.set_samples(64)
    want = np.zeros(P.PX_LEN * w * h + 1, dtype=np.uint32)
    for f in range(spp):
        fo.sample(cam, w, h, want, f, 8)
    n = w * h
    nested = {}
    for build in (P.Tracer.PATH_NESTED, P.Tracer.PATH_SPLIT):  # the split build: walks spanning turns
        t.set_path_regen(build)
        for waves, stack in ((2, 32), (3, 32), (3, 24), (4, 24)):  # register fits and mesh-tree stacks
            t.set_path_waves(waves)
            t.set_path_stack(stack)
            assert t.path_stack == stack
            t.sample(S.path_camera(w + 1, h), w + 1, h, 1)  # another view: the next call starts over
            t.sample(cam, w, h, spp)
            got = t.px()
            for lo, hi in [(1 * n, 4 * n), (7 * n, 14 * n)]:
                a, b = got[lo:hi].view(np.float32), want[lo:hi].view(np.float32)
                bad = ~np.isclose(a, b, rtol=1e-5, atol=1e-7)
                assert not bad.any(), (f"build {build} waves {waves} words {lo}..{hi}: {bad.sum()} beyond 1e-5, "
                                       f"first {np.nonzero(bad)[0][:5] + lo}")
            assert (got[31 * n:32 * n] == want[31 * n:32 * n]).all()
            if build == P.Tracer.PATH_NESTED:
                nested[(waves, stack)] = got.copy()
            else:  # the same bits as the nested build
                assert np.array_equal(got, nested[(waves, stack)]), (waves, stack)
    t.set_path_regen(P.Tracer.PATH_NESTED)


def test_reinstance_matches_the_oracle(gpu, orc):
    print("[forest] reinstance", flush=True)
    arrs = forest_scene().arrays()
    t = P.Tracer.from_scene(gpu, *arrs, leaf=4)
    fo = orc.Forest(*arrs, 4, P.NODE)
    rng = np.random.default_rng(21)
    worlds = arrs[4]["world"].copy()
    for k in range(7):
        worlds[k] = S.world_of(S.affine(S.rotation(rng.normal(size=3), rng.uniform(0, 6)), rng.uniform(0.4, 1.0),
                                        rng.uniform(-3, 3, 3) + np.array([0, 0, 3.5])))
    assert t.reinstance(worlds)
    fo.reinstance(worlds)
    check_forest(t, fo, "reinstance")
    r = rays_at(fo, 10_000, 5)
    same_hits(t.raycast(r), fo.raycast(r, P.HIT), "reinstance raycast")
    assert not t.reinstance(worlds[:-1])  # a count that differs asks for a rebuild


def test_repose_refits_and_retree_rebuilds(gpu, orc):
    print("[forest] repose + retree", flush=True)
    b = forest_scene()
    arrs = b.arrays()
    t = P.Tracer.from_scene(gpu, *arrs, leaf=4)
    fo = orc.Forest(*arrs, 4, P.NODE)
    verts = arrs[0].copy()
    rng = np.random.default_rng(8)
    verts["p"][:, :3] += rng.uniform(-0.03, 0.03, size=(len(verts), 3)).astype(np.float32)
    assert t.repose(verts)
    # Each mesh tree: the oracle's tree over the old pose refitted over the new boxes.
    got = t.nodes()
    faces, meshes = arrs[1], arrs[3]
    base = 2 * len(arrs[4]) - 1
    for m in meshes:
        lo, hi = int(m["face_lo"]), int(m["face_hi"])
        cnt = hi - lo
        nodes_m = fo.nodes[base:base + 2 * cnt - 1].copy()
        w = nodes_m.view(np.uint32).reshape(-1, 8)
        leaf = (w[:, 3] & P.LEAF) != 0
        w[leaf, 3] = ((w[leaf, 3] & 0x7FFFFFFF) - lo) | P.LEAF
        w[leaf, 7] -= lo
        w[~leaf, 3] -= base
        w[~leaf, 7] -= base
        boxes = S.of_tri(*(verts["p"][faces["v"][lo:hi, j], :3] for j in range(3)))
        want = orc.refit(nodes_m, fo.order[lo:hi] - lo, boxes)
        gm = got[base:base + 2 * cnt - 1].copy()
        gw = gm.view(np.uint32).reshape(-1, 8)
        gleaf = (gw[:, 3] & P.LEAF) != 0
        gw[gleaf, 3] = ((gw[gleaf, 3] & 0x7FFFFFFF) - lo) | P.LEAF
        gw[gleaf, 7] -= lo
        gw[~gleaf, 3] -= base
        gw[~gleaf, 7] -= base
        same_nodes(gm, want, f"repose mesh {lo}")
        base += 2 * cnt - 1
    # The top tree over the new world boxes.
    placed = np.arange(len(arrs[4]))
    roots = fo.inst_words[24::40][:len(placed)]
    wb = np.zeros(len(placed), dtype=P.AABB)
    for j, k in enumerate(placed):
        rn = got[int(roots[k])]
        rb = np.zeros(1, dtype=P.AABB)[0]
        rb["lo"], rb["hi"] = rn["lo"], rn["hi"]
        wb[j] = orc.inst_world_box(arrs[4]["world"][k], rb)
    tn, to, _ = orc.build_bvh(wb, 1, P.NODE)
    tw = tn.view(np.uint32).reshape(-1, 8).copy()
    tw[(tw[:, 3] & P.LEAF) != 0, 3] |= P.INST
    assert (got.view(np.uint32).reshape(-1, 8)[:len(tn)] == tw).all()
    # retree: a fresh forest over the moved pose.
    assert t.retree()
    moved = (verts, arrs[1], arrs[2], arrs[3], arrs[4])
    fresh = orc.Forest(*moved, 4, P.NODE)
    check_forest(t, fresh, "retree")
    r = rays_at(fresh, 10_000, 6)
    same_hits(t.raycast(r), fresh.raycast(r, P.HIT), "retree raycast")


def test_one_identity_instance_traces_as_the_triangle_tracer(gpu):
    print("[forest] one identity instance", flush=True)
    tris = S.spheres_scene(spheres=4, level=3, seed=2)
    # The same triangles as one mesh (vertices per corner, flat normals off).
    b = S.SceneBuilder()
    mats = {}
    pos, idx, mids, nrm = [], [], [], []
    for k, tr in enumerate(tris):
        key = (tuple(tr["color"]), tuple(tr["emit"]))
        if key not in mats:
            mats[key] = b.material(color=tuple(tr["color"]), emit=tuple(tr["emit"]), rough=float(tr["rough"]),
                                   ior=float(tr["ior"]))
        pos += [tr["a"], tr["b"], tr["c"]]
        nrm += [tr["na"], tr["nb"], tr["nc"]]
        idx.append([3 * k, 3 * k + 1, 3 * k + 2])
        mids.append(mats[key])
    b.mesh(np.array(pos), np.array(idx), np.array(mids, dtype=np.uint32), normals=np.array(nrm))
    b.instance(0)
    tf = P.Tracer.from_scene(gpu, *b.arrays(), leaf=4)
    tt = P.Tracer.new(gpu, tris, 4)
    r = S.pinhole_rays(64, 64, origin=(0.0, -6.0, 1.0))
    r["d"] = np.ascontiguousarray(r["d"][:, [0, 2, 1]])  # look along +y, at the spheres
    hf, ht = tf.raycast(r), tt.raycast(r)
    same_hits(hf, ht, "identity forest raycast")
    cam = S.path_camera(32, 24, rr_from=3)
    for t in (tf, tt):
        t.set_samples(64)
        t.sample(cam, 32, 24, 2)
    assert (tf.px()[:34 * 32 * 24] == tt.px()[:34 * 32 * 24]).all()


@pytest.mark.parametrize("denoise", [0, 1])
def test_a_recorded_posed_frame_equals_the_eager_one(gpu, denoise):
    """One hipGraph holding a new pose (vertices, matrices), the camera's
    stamp, the path launch and the denoiser gives, replay after replay, the
    image a reinstance + repose + sample (+ denoise) call by call gives."""
    print(f"[forest] recorded posed frame, denoise {denoise}", flush=True)
    arrs = forest_scene().arrays()
    t = P.Tracer.from_scene(gpu, *arrs, leaf=4)
    eager = P.Tracer.from_scene(gpu, *arrs, leaf=4)
    w, h, spp = 40, 28, 2
    cam = S.path_camera(w, h, eye=(0.0, -7.0, 3.5), target=(0.0, 0.5, 0.8), rr_from=2)
    verts = P.GpuArray.from_(gpu, arrs[0])
    base = arrs[4]["world"].copy()
    rng = np.random.default_rng(31)

    def pose(k):
        wk = base.copy()
        for i in range(7):
            wk[i] = S.world_of(S.affine(S.rotation(rng.normal(size=3), rng.uniform(0, 6)), rng.uniform(0.4, 1.0),
                                        rng.uniform(-3, 3, 3) + np.array([0, 0, 3.5])))
        v = arrs[0].copy()
        v["p"][:, :3] *= np.float32(1.0 + 0.01 * k)
        return wk, v

    w0, v0 = pose(0)
    verts.write(v0)
    t.record_pose_frame(verts, w0, cam, w, h, spp, denoise)
    for k in range(1, 4):
        wk, vk = pose(k)
        verts.write(vk)
        t.replay_pose_frame(wk)
        got = t.px()
        assert eager.repose_on(verts) and eager.reinstance(wk)
        eager.sample(cam, w, h, spp)
        if denoise == 1:
            eager.denoise(keep=False)
        want = eager.px()
        assert (got == want).all(), f"pose {k}: {(got != want).sum()} words differ"
        eager.sample(S.path_camera(w + 1, h), w + 1, h, 1)  # another view: the next pose starts over

"""The rigid-body workload at its own size on the device: 1,024 bodies (boxes
and icospheres of subdivision 0..5, about 3.9 M triangles in 1,026 meshes)
under one top tree, as the benchmark runs it.

* the forest's tables, every reachable node, the face order and the light
  table equal the oracle's Forest bit for bit;
* a 1280x720 frame (1 sample) equals the oracle's integrator on three 6-row
  bands (top, middle, bottom) within 1e-5 relative;
* one substep of new poses (every mesh's vertices moved, every matrix new):
  each mesh tree equals the oracle's refit of its tree over the moved boxes,
  and the top tree equals the oracle's Bvh::build over the new world boxes;
* the device World steps the 1,024 bodies for 30 frames of 8 substeps beside
  the oracle's World: bodies that met nothing in the first frames follow its
  free fall to 1e-4, total energy and the mean height track it within 1% at
  every frame, every body stays above the ground and no state turns
  non-finite.
"""
import numpy as np
import pytest

import phiray_amd as P
from phiray_amd import scenes as S
from tests.test_gpu_forest import check_forest, same_nodes

pytestmark = pytest.mark.gpu

W, H = 1280, 720


@pytest.fixture(scope="module")
def bodies(gpu, orc):
    print("[config4] building the 1,024-body forest (device and oracle)", flush=True)
    b, worlds = S.bodies_scene(1024)
    arrs = b.arrays()
    t = P.Tracer.from_scene(gpu, *arrs, leaf=4)
    fo = orc.Forest(*arrs, 4, P.NODE, 8)
    return arrs, worlds, t, fo


def test_the_forest_matches_the_oracle(bodies):
    arrs, _, t, fo = bodies
    assert len(arrs[1]) > 3_500_000 and len(arrs[4]) == 1026
    check_forest(t, fo, "1,024 bodies")


def test_a_frame_matches_the_twin_on_three_bands(bodies):
    print("[config4] 1280x720 frame against the oracle's bands, nested and split builds", flush=True)
    arrs, _, t, fo = bodies
    cam = S.path_camera(W, H, eye=(0.0, -16.0, 9.0), target=(0.0, 0.0, 1.5), rr_from=3)
    t.set_path_waves(4)
    t.sample(cam, W, H, 1)
    got = P.planes(t.px(), W, H)
    nested_px = t.px().copy()
    t.set_path_regen(P.Tracer.PATH_SPLIT)  # the split build: the same bits
    t.sample(S.path_camera(64, 64), 64, 64, 1)
    t.sample(cam, W, H, 1)
    assert np.array_equal(t.px(), nested_px)
    t.set_path_regen(P.Tracer.PATH_NESTED)
    px = np.zeros(P.PX_LEN * W * H + 1, dtype=np.uint32)
    bands = [(0, 6), (H // 2 - 3, 6), (H - 6, 6)]
    for row, rows in bands:
        fo.sample(cam, W, H, px, 0, 8, row * W, rows * W)
    want = P.planes(px, W, H)
    lit = 0
    for row, rows in bands:
        sl = slice(row * W, (row + rows) * W)
        for k in ("radiance", "albedo", "normal", "depth", "count", "missed"):
            a, b = got[k][sl], want[k][sl]
            if a.dtype == np.float32:
                bad = ~np.isclose(a, b, rtol=1e-5, atol=0)
                assert not bad.any(), f"rows {row}+{rows} {k}: {bad.sum()} words beyond 1e-5 relative"
            else:
                assert (a == b).all(), (row, k)
        lit += int((want["radiance"][sl] > 0).any(axis=1).sum())
    assert lit > 1000


def test_one_substep_of_new_poses(gpu, orc, bodies):
    print("[config4] one substep: moved vertices, new matrices", flush=True)
    arrs, worlds, t, fo = bodies
    verts = arrs[0].copy()
    verts["p"][:, :3] *= np.float32(1.02)
    turn = S.affine(S.rotation((0.3, 0.2, 1.0), 0.05))
    mats = arrs[4]["world"].copy()
    for i in range(len(worlds)):
        mats[i] = S.world_of(turn @ worlds[i])
    assert t.repose(verts) and t.reinstance(mats)
    got = t.nodes()
    faces, meshes, insts = arrs[1], arrs[3], arrs[4]
    base = 2 * len(insts) - 1
    for m in meshes:
        lo, hi = int(m["face_lo"]), int(m["face_hi"])
        cnt = hi - lo

        def local(nodes):
            w = nodes.view(np.uint32).reshape(-1, 8)
            leaf = (w[:, 3] & P.LEAF) != 0
            w[leaf, 3] = ((w[leaf, 3] & 0x7FFFFFFF) - lo) | P.LEAF
            w[leaf, 7] -= lo
            w[~leaf, 3] -= base
            w[~leaf, 7] -= base
            return nodes

        mine = local(fo.nodes[base:base + 2 * cnt - 1].copy())
        boxes = S.of_tri(*(verts["p"][faces["v"][lo:hi, j], :3] for j in range(3)))
        want = orc.refit(mine, fo.order[lo:hi] - lo, boxes)
        same_nodes(local(got[base:base + 2 * cnt - 1].copy()), want, f"mesh at faces {lo}")
        base += 2 * cnt - 1
    roots = fo.inst_words[24::40][:len(insts)]
    wb = np.zeros(len(insts), dtype=P.AABB)
    for k in range(len(insts)):
        rn = got[int(roots[k])]
        rb = np.zeros(1, dtype=P.AABB)[0]
        rb["lo"], rb["hi"] = rn["lo"], rn["hi"]
        wb[k] = orc.inst_world_box(mats[k], rb)
    tn, _, _ = orc.build_bvh(wb, 1, P.NODE)
    tw = tn.view(np.uint32).reshape(-1, 8).copy()
    tw[(tw[:, 3] & P.LEAF) != 0, 3] |= P.INST
    r = np.array(P.Bvh(tn, None, 0).reachable(), dtype=np.int64)
    assert (got.view(np.uint32).reshape(-1, 8)[r] == tw[r]).all()


def _world(new, defs):
    w = new()
    w.configure(substeps=1, iterations=10)
    g = w.add_box((12.0, 12.0, 0.5))
    w.add_body(kind=P.PASSIVE, shape=g, x=(0.0, 0.0, -0.5), friction=1.0)
    for k, d in enumerate(defs):
        r = float(d["r"])
        s = w.add_box((r, r, r)) if d["shape"] == "box" else w.add_sphere(r)
        w.add_body(shape=s, x=d["x"], q=d["q"], scale=(r, r, r), obj=k)
    return w


def test_the_world_steps_the_1024_bodies(gpu, orc):
    print("[config4] device World against the oracle's, 1,024 bodies, 30 frames of 8 substeps", flush=True)
    defs = S.bodies_defs(1024, heights=(1.0, 10.0))
    dev, ref = _world(lambda: P.World(gpu), defs), _world(orc.World, defs)
    touched = np.zeros(1025, dtype=bool)
    for f in range(30):
        for _ in range(8):
            dev.step(1.0 / 480.0)
            ref.step(1.0 / 480.0)
            if f < 3:
                for i, j in ref.pairs():
                    touched[i] = touched[j] = True
        if f == 2:
            a, b = dev.bodies(), ref.bodies()
            free = ~touched
            free[0] = False
            assert free.sum() >= 10
            assert np.abs(a["x"][free] - b["x"][free]).max() < 1e-4
        # The same physics: total energy and the heap's mean height track the
        # oracle's within 1%.
        e_dev, e_ref = dev.energy(), ref.energy()
        assert abs(e_dev - e_ref) <= 0.01 * abs(e_ref), (f, e_dev, e_ref)
        zd, zr = dev.bodies()["x"][1:, 2].mean(), ref.bodies()["x"][1:, 2].mean()
        assert abs(zd - zr) <= 0.01 * zr, (f, zd, zr)
    b = dev.bodies()
    assert np.all(np.isfinite(b["x"])) and np.all(np.isfinite(b["v"])) and np.all(np.isfinite(b["q"]))
    radius = np.array([float(d["r"]) for d in defs], dtype=np.float32)
    assert (b["x"][1:, 2] > 0.5 * radius).all(), b["x"][1:, 2].min()
    assert dev.contact_count() > 100


def test_the_four_lane_solve_equals_the_one_lane_solve(gpu):
    """The solve's projections on four lanes a pair (a point a lane, sums
    gathered in point order) give the bits of the one-lane build: the 1,024
    bodies over 60 frames of 8 substeps, every body record compared every 10
    frames, once most of the heap has landed and pairs share bodies."""
    print("[config4] four-lane solve against the one-lane solve, 60 frames", flush=True)
    defs = S.bodies_defs(1024, heights=(1.0, 10.0))
    one, four = _world(lambda: P.World(gpu), defs), _world(lambda: P.World(gpu), defs)
    one.set_lanes(1)
    assert one.lanes == 1 and four.lanes == 4
    for f in range(60):
        for _ in range(8):
            one.step(1.0 / 480.0)
            four.step(1.0 / 480.0)
        if f % 10 == 9:
            a, b = one.bodies(), four.bodies()
            assert a.tobytes() == b.tobytes(), (f, int((a != b).sum()))
    assert four.contact_count() > 200

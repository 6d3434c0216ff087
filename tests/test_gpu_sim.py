"""The device World (phiray_amd/csrc/world.hip) against the closed forms of
phiray's own tests/sim.rs (listed in tests/sim_cases.py), its broadphase
against a brute-force pair list, and its trajectories against the oracle's
World (phiray's solver restated on the CPU) over a short horizon.

The two solvers are not the same arithmetic: the device colours the pairs
and solves a colour's pairs at once (Gauss-Seidel over colours), the oracle
walks the pairs in sorted order, and their narrowphases pick box points by
different rules. What must agree is the physics: on scenes where a body
meets at most one other at a time, positions within 2 mm and velocities
within 0.02 m/s after the stated number of frames.
"""
import numpy as np
import pytest

import phiray_amd as P
from tests import sim_cases as C

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def new_world(gpu):
    return lambda: P.World(gpu)


def test_free_fall_matches_the_closed_form(new_world):
    print("[sim] free fall", flush=True)
    C.free_fall(new_world)


def test_a_ball_bounces_to_e_squared_of_its_height(new_world):
    print("[sim] bounce", flush=True)
    apex = C.bounce_apex(new_world)
    assert abs(apex - 0.25) < 0.05 * 0.25, apex


def test_a_box_rests_below_the_friction_angle_and_slides_above(new_world):
    print("[sim] slope", flush=True)
    C.friction_angle(new_world)


def test_ten_boxes_stand_for_a_thousand_frames(new_world):
    print("[sim] stack", flush=True)
    C.stack(new_world)


def test_a_sphere_and_a_box_come_to_rest_exactly(new_world):
    print("[sim] rest", flush=True)
    C.rest(new_world)


def test_two_runs_are_bit_identical(new_world):
    print("[sim] determinism", flush=True)
    C.deterministic(new_world)


def test_a_frictionless_bouncing_ball_does_not_gain_energy(new_world):
    print("[sim] energy", flush=True)
    C.no_energy_gain(new_world)


def test_broadphase_pairs_equal_brute_force(new_world):
    """Every pair of overlapping swept boxes, each once (broad.rs:73-91)."""
    print("[sim] broadphase", flush=True)
    w, want = C.random_spheres(new_world)
    assert np.array_equal(w.pairs(), want)


def _scene(w, which):
    C.ground(w)
    if which == "drop":
        s = w.add_sphere(0.3)
        w.add_body(shape=s, x=(0.2, -0.1, 2.0), v=(0.5, 0.0, 0.0), restitution=0.4)
        b = w.add_box((0.3, 0.4, 0.5))
        w.add_body(shape=b, x=(3.0, 0.0, 0.5005), q=C.q_axis((0.0, 0.0, 1.0), 0.3))
    elif which == "head_on":
        s = w.add_sphere(0.25)
        w.add_body(shape=s, x=(-1.0, 0.0, 0.25), v=(3.0, 0.0, 0.0), friction=0.0, restitution=0.8)
        w.add_body(shape=s, x=(1.0, 0.0, 0.25), v=(-1.0, 0.0, 0.0), friction=0.0, restitution=0.8)
    elif which == "box_slide":
        b = w.add_box((0.5, 0.5, 0.25))
        w.add_body(shape=b, x=(0.0, 0.0, 0.2505), v=(2.0, 1.0, 0.0), friction=0.3)


@pytest.mark.parametrize("which,frames", [("drop", 90), ("head_on", 60), ("box_slide", 60)])
def test_trajectories_follow_the_oracle(gpu, orc, which, frames):
    print(f"[sim] against the oracle: {which}", flush=True)
    dev, ref = P.World(gpu), orc.World()
    _scene(dev, which)
    _scene(ref, which)
    for f in range(frames):
        dev.step(C.DT)
        ref.step(C.DT)
    a, b = dev.bodies(), ref.bodies()
    dx = np.abs(a["x"] - b["x"]).max()
    dv = np.abs(a["v"] - b["v"]).max()
    assert dx < 2e-3 and dv < 2e-2, (which, dx, dv, a["x"], b["x"])


@pytest.mark.parametrize("which", ["drop", "head_on", "box_slide", "stack"])
def test_four_lanes_give_the_one_lane_bits(gpu, which):
    """The solve's four-lanes-a-pair projections (with the next colour's
    records read ahead) against the one-lane build, bit for bit, on worlds of
    one or two colours (where the next entry to read is the one being
    solved) and the ten-box stack."""
    print(f"[sim] four lanes against one: {which}", flush=True)
    worlds = []
    for lanes in (1, 4):
        w = P.World(gpu)
        w.set_lanes(lanes)
        if which == "stack":
            C.ground(w)
            b = w.add_box((0.5, 0.5, 0.5))
            for k in range(10):
                w.add_body(shape=b, x=(0.0, 0.0, 0.5 + k))
        else:
            _scene(w, which)
        worlds.append(w)
    for f in range(120):
        for w in worlds:
            w.step(C.DT)
        if f % 30 == 29:
            a, b = worlds[0].bodies(), worlds[1].bodies()
            assert a.tobytes() == b.tobytes(), (which, f)

"""The oracle's rigid-body World (phiray's XPBD world restated on the CPU,
oracle/sim.cpp) pinned by the closed forms of phiray's own tests/sim.rs
(listed in tests/sim_cases.py), and its brute-force broadphase by an
independent pair list. CPU only: this is the checker the device World is
compared with."""
import numpy as np
import pytest

from oracle import oracle as O
from tests import sim_cases as C


@pytest.fixture(scope="module")
def new_world():
    O.lib()
    return O.World


def test_free_fall(new_world):
    C.free_fall(new_world)


def test_bounce_to_e_squared(new_world):
    assert abs(C.bounce_apex(new_world) - 0.25) < 0.05 * 0.25


def test_friction_angle(new_world):
    C.friction_angle(new_world)


def test_ten_box_stack(new_world):
    C.stack(new_world)


def test_rest_heights(new_world):
    C.rest(new_world)


def test_bit_identical_reruns(new_world):
    C.deterministic(new_world)


def test_no_energy_gain(new_world):
    C.no_energy_gain(new_world)


def test_pairs_equal_brute_force(new_world):
    w, want = C.random_spheres(new_world)
    assert np.array_equal(w.pairs(), want)

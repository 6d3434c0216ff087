"""The closed forms phiray's own tests/sim.rs holds its World to, written once
over a world factory so the same checks run on the oracle's World (CPU) and on
the device World (GPU):

* free fall equals x0 + v0 t + g t^2 / 2 within 1e-3 (tests/sim.rs:21-34);
* a ball of restitution 0.5 bounces to e^2 = 1/4 of its drop, within 5%
  (tests/sim.rs:36-59);
* a box holds below the friction angle (< 1 mm) and slides above it (> 0.5 m)
  (tests/sim.rs:61-85);
* ten stacked boxes stand 1,000 frames (top moves < 1 mm, each within 5 mm
  of its height) (tests/sim.rs:87-107);
* a sphere and a box come to rest at their heights within 1e-4, drift under
  1e-5, speed under g h (tests/sim.rs:148-160);
* two runs are bit-identical (tests/sim.rs:237-243);
* a frictionless bouncing ball never gains energy (tests/sim.rs:245-268).

A factory is `new_world() -> World` with phiray's World methods
(configure, add_box, add_sphere, add_body, step, bodies, energy,
contact_count, pairs). `passive` is the passive kind's number (1).
"""
import numpy as np

DT = 1.0 / 60.0
PASSIVE = 1


def ground(w):
    s = w.add_box((50.0, 50.0, 0.5))
    return w.add_body(kind=PASSIVE, shape=s, x=(0.0, 0.0, -0.5), friction=1.0, restitution=1.0)


def q_axis(axis, angle):
    s, c = np.sin(np.float32(0.5 * angle)), np.cos(np.float32(0.5 * angle))
    return (axis[0] * s, axis[1] * s, axis[2] * s, c)


def rotate(q, v):
    q = np.array(q, dtype=np.float64)
    u, w = q[:3], q[3]
    t = 2.0 * np.cross(u, v)
    return v + w * t + np.cross(u, t)


def free_fall(new_world):
    w = new_world()
    s = w.add_sphere(0.5)
    x0, v0 = np.array([1.0, 2.0, 20.0]), np.array([0.3, -0.2, 2.0])
    i = w.add_body(shape=s, x=x0, v=v0, lin_damp=0.0, ang_damp=0.0)
    g = np.array([0.0, 0.0, -9.81])
    for f in range(1, 101):
        w.step(DT)
        t = f * DT
        want = x0 + v0 * t + 0.5 * g * t * t
        got = w.bodies()[i]["x"]
        assert np.linalg.norm(got - want) < 1e-3, (f, got, want)


def bounce_apex(new_world) -> float:
    """The first apex after the bounce over the 2 m drop (want 0.25)."""
    w = new_world()
    ground(w)
    s = w.add_sphere(0.5)
    h0 = 2.0
    i = w.add_body(shape=s, x=(0.0, 0.0, 0.5 + h0), restitution=0.5, lin_damp=0.0, ang_damp=0.0)
    rising, apex = False, 0.0
    for _ in range(240):
        w.step(DT)
        b = w.bodies()[i]
        z, vz = float(b["x"][2]), float(b["v"][2])
        if vz > 0.0:
            rising = True
        if rising and vz < 0.0:
            break
        if rising:
            apex = max(apex, z - 0.5)
    assert rising, "the ball never bounced"
    return apex / h0


def slide(new_world, mu, theta) -> float:
    w = new_world()
    big = w.add_box((20.0, 20.0, 0.5))
    q = q_axis((0.0, 1.0, 0.0), theta)
    w.add_body(kind=PASSIVE, shape=big, q=q, friction=1.0)
    unit = w.add_box((0.5, 0.5, 0.5))
    n = rotate(q, np.array([0.0, 0.0, 1.0]))
    i = w.add_body(shape=unit, q=q, x=n, friction=mu, lin_damp=0.0, ang_damp=0.0)
    start = w.bodies()[i]["x"].copy()
    for _ in range(120):
        w.step(DT)
    return float(np.linalg.norm(w.bodies()[i]["x"] - start))


def friction_angle(new_world):
    crit = float(np.arctan(np.float32(0.5)))
    held = slide(new_world, 0.5, crit - 0.1)
    slid = slide(new_world, 0.5, crit + 0.1)
    assert held < 1e-3, held
    assert slid > 0.5, slid


def stack(new_world, frames: int = 1000):
    w = new_world()
    ground(w)
    unit = w.add_box((0.5, 0.5, 0.5))
    tops = [w.add_body(shape=unit, x=(0.0, 0.0, 0.5 + k)) for k in range(10)]
    start = w.bodies()[tops[9]]["x"].copy()
    worst = 0.0
    for _ in range(frames):
        w.step(DT)
        worst = max(worst, float(np.linalg.norm(w.bodies()[tops[9]]["x"] - start)))
    assert worst < 1e-3, worst
    b = w.bodies()
    for k, i in enumerate(tops):
        assert abs(float(b[i]["x"][2]) - (0.5 + k)) < 5e-3, (k, b[i]["x"])


def rest(new_world):
    subs = 10
    w = new_world()
    ground(w)
    shapes = [(w.add_sphere(0.3), 0.3), (w.add_box((0.3, 0.4, 0.5)), 0.5)]
    bodies = []
    for k, (s, r) in enumerate(shapes):
        x = (3.0 * k, 0.0, r + 0.5)
        bodies.append((w.add_body(shape=s, x=x, lin_damp=0.0, ang_damp=0.0), np.array(x), r))
    for _ in range(180):
        w.step(DT)
    floor = 9.81 * DT / subs / 2.0
    b = w.bodies()
    for i, x, r in bodies:
        drift = float(np.linalg.norm(b[i]["x"][:2] - x[:2]))
        speed = max(float(np.linalg.norm(b[i]["v"])), float(np.linalg.norm(b[i]["w"])))
        assert abs(float(b[i]["x"][2]) - r) < 1e-4, (i, b[i]["x"])
        assert drift < 1e-5, drift
        assert speed < 2.0 * floor, speed


def busy(new_world):
    w = new_world()
    ground(w)
    unit = w.add_box((0.5, 0.5, 0.5))
    for k in range(5):
        w.add_body(shape=unit, x=(0.05 * k, 0.0, 0.5 + k))
    s = w.add_sphere(0.4)
    w.add_body(shape=s, x=(0.3, 0.1, 8.0), restitution=0.3)
    for _ in range(120):
        w.step(DT)
    return w


def deterministic(new_world):
    a, b = busy(new_world), busy(new_world)
    ba, bb = a.bodies(), b.bodies()
    assert ba.tobytes() == bb.tobytes()
    assert a.contact_count() == b.contact_count()
    assert (ba["x"][1:, 2] < 4.0).any(), "the stack never fell"


def no_energy_gain(new_world):
    w = new_world()
    ground(w)
    s = w.add_sphere(0.5)
    w.add_body(shape=s, x=(0.0, 0.0, 1.5), restitution=1.0, friction=0.0, lin_damp=0.0, ang_damp=0.0)
    e0 = w.energy()
    bounces, vz = 0, 0.0
    for _ in range(600):
        w.step(DT)
        e = w.energy()
        assert e <= e0 * (1.0 + 1e-4), (e0, e)
        now = float(w.bodies()[1]["v"][2])
        if vz < 0.0 and now > 0.0:
            bounces += 1
        vz = now
    assert bounces >= 5
    assert w.energy() > 0.5 * e0


def random_spheres(new_world, n: int = 700, seed: int = 3):
    """n spheres of mixed kinds and groups, stepped once without gravity, and
    the brute-force pair list of their boxes (broad.rs:73-91)."""
    rng = np.random.default_rng(seed)
    w = new_world()
    radii = rng.uniform(0.05, 0.4, size=n).astype(np.float32)
    xs = rng.uniform(-4.0, 4.0, size=(n, 3)).astype(np.float32)
    kinds = rng.integers(0, 3, size=n)  # some passive bodies: a passive pair is never listed
    groups = rng.choice([1, 2, 3], size=n)
    for k in range(n):
        s = w.add_sphere(float(radii[k]))
        w.add_body(shape=s, x=xs[k], kind=PASSIVE if kinds[k] == 1 else 0, groups=int(groups[k]))
    w.configure(gravity=(0.0, 0.0, 0.0))
    w.step(DT)  # v = 0 before the step: the swept box is the box
    pad = np.float32(0.01) + np.float32(0.005)  # SLOP + 0.005, in f32
    lo = xs - radii[:, None] - pad
    hi = xs + radii[:, None] + pad
    want = []
    for i in range(n):
        if kinds[i] == 1:
            continue
        meet = np.all((lo <= hi[i]) & (hi >= lo[i]), axis=1)
        for j in np.flatnonzero(meet):
            if j != i and (groups[i] & groups[j]) != 0:
                want.append((min(i, j), max(i, j)))
    want = np.array(sorted(set(want)), dtype=np.uint32).reshape(-1, 2)
    assert len(want) > 50
    return w, want

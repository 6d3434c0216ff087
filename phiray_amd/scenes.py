"""Seeded scene, ray and camera generators for the LBVH and tracer workloads.

All arithmetic that decides a coordinate is single precision (numpy float32,
IEEE round to nearest), so the same inputs are produced wherever this runs.

* `xorshift_prims_gpu_test(n, shift)` — tests/bvh.rs:5-19 (seed 0x2545F491,
  every 13th box at (5, 5, 5)).
* `xorshift_prims_unit(n)` — src/bvh/test.rs:6-23 (seed 0x9E3779B9, every
  10th box the duplicate [1, 1.5]^3).
* `lbvh_scene(n, seed)` — the 65,536-triangle raycast workload: SplitMix64,
  u = (x >> 40) * 2^-24; centroid ~ U[0,1]^3, vertices at centroid +
  U[-0.01, 0.01]^3; the refit perturbation takes the stream's next draws,
  U[-0.001, 0.001]^3 per vertex.
* `batch_scenes(count, lo, hi, seed)` — the batch of independent scenes:
  scene i has round(exp(U[ln lo, ln hi])) triangles (U from SplitMix64 at
  `seed`, in float64) made by `lbvh_scene` at seed i; returns the
  concatenated boxes and the segment starts.
* `pinhole_rays(w, h)` — primary rays from (0.5, 0.5, -1.5) along +z, 45°
  horizontal field of view, through pixel centres.
* `spheres_scene(...)` — icospheres on a two-triangle ground plane under one
  emissive 1 m² quad, Lambertian-like materials (roughness 1, metallic 0,
  IOR 1, so no dielectric reflection).
"""
from __future__ import annotations

import numpy as np

from . import AABB, CAMERA, RAY, TRI

F32 = np.float32
GOLDEN = np.uint64(0x9E3779B97F4A7C15)


def splitmix64(seed: int, count: int, start: int = 0) -> np.ndarray:
    """Outputs start+1 .. start+count of SplitMix64 seeded with `seed`."""
    with np.errstate(over="ignore"):
        k = np.arange(start + 1, start + count + 1, dtype=np.uint64)
        z = np.uint64(seed) + k * GOLDEN
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def uniform(seed: int, count: int, start: int = 0) -> np.ndarray:
    """24-bit uniforms in [0, 1), exact in float32."""
    return (splitmix64(seed, count, start) >> np.uint64(40)).astype(F32) * F32(2.0 ** -24)


def _xorshift_stream(seed: int, count: int) -> np.ndarray:
    out = np.empty(count, dtype=np.uint32)
    x = seed & 0xFFFFFFFF
    for i in range(count):
        x ^= (x << 13) & 0xFFFFFFFF
        x ^= x >> 17
        x ^= (x << 5) & 0xFFFFFFFF
        out[i] = x
    return out


def of_tri(a: np.ndarray, b: np.ndarray, c: np.ndarray) -> np.ndarray:
    """Aabb::of_tri (src/bvh/node.rs:78-80) for arrays of corners."""
    out = np.zeros(len(a), dtype=AABB)
    out["lo"] = np.minimum(np.minimum(a, b), c)
    out["hi"] = np.maximum(np.maximum(a, b), c)
    return out


def xorshift_prims_gpu_test(n: int, shift: float = 0.0) -> np.ndarray:
    """tests/bvh.rs:5-19."""
    draws = n * 3
    xs = _xorshift_stream(0x2545F491, draws)
    a = np.zeros((n, 3), dtype=F32)
    k = 0
    for i in range(n):
        if i % 13 == 0:
            a[i] = (5.0, 5.0, 5.0)
        else:
            r0 = F32(xs[k] % 100000) * F32(0.001)
            r1 = F32(xs[k + 1] % 100000) * F32(0.001)
            r2 = F32(xs[k + 2] % 100000) * F32(0.001)
            k += 3
            a[i] = (r0, F32(r1 + F32(shift)), r2)
    b = a.copy()
    b[:, 0] = a[:, 0] + F32(0.3)
    c = a.copy()
    c[:, 1] = a[:, 1] + F32(0.3)
    c[:, 2] = a[:, 2] + F32(0.05)
    return of_tri(a, b, c)


def xorshift_prims_unit(n: int) -> np.ndarray:
    """src/bvh/test.rs:6-23."""
    xs = _xorshift_stream(0x9E3779B9, n * 4)
    out = np.zeros(n, dtype=AABB)
    k = 0
    for i in range(n):
        if i % 10 == 9:
            out[i]["lo"] = (1.0, 1.0, 1.0)
            out[i]["hi"] = (1.5, 1.5, 1.5)
            continue
        r = [F32(xs[k + j] % 10000) * F32(0.01) for j in range(4)]
        k += 4
        a = np.array(r[:3], dtype=F32)
        b = np.array([a[0] + F32(1.0), a[1], a[2]], dtype=F32)
        c = np.array([a[0], a[1] + F32(1.0), a[2] + r[3] * F32(0.1)], dtype=F32)
        out[i] = of_tri(a[None], b[None], c[None])[0]
    return out


def lbvh_scene(n: int = 65536, seed: int = 0) -> tuple[np.ndarray, np.ndarray]:
    """(corners [n,3,3] float32, moved corners [n,3,3]) of the LBVH workload."""
    u = uniform(seed, 21 * n)
    first = u[: 12 * n].reshape(n, 12)
    c = first[:, :3]
    off = (first[:, 3:12].reshape(n, 3, 3) * F32(2.0) - F32(1.0)) * F32(0.01)
    corners = (c[:, None, :] + off).astype(F32)
    move = (u[12 * n:].reshape(n, 3, 3) * F32(2.0) - F32(1.0)) * F32(0.001)
    moved = (corners + move).astype(F32)
    return corners, moved


def tris_from_corners(corners: np.ndarray, color=(0.5, 0.5, 0.5, 1.0)) -> np.ndarray:
    """Triangle records with flat shading (zero normals) and one material."""
    t = np.zeros(len(corners), dtype=TRI)
    t["a"], t["b"], t["c"] = corners[:, 0], corners[:, 1], corners[:, 2]
    t["rough"] = 1.0
    t["ior"] = 1.0
    t["color"] = color
    return t


def cloud_scene(seed: int, n: int = 65536) -> np.ndarray:
    """The learned denoiser's scenes: the LBVH workload's triangles (n of
    them, seed `seed`: centroids uniform in the unit cube, corners within
    0.01) as Lambertian triangles of albedo U[0.2, 0.8]^3 each, drawn from
    the stream that follows the geometry's draws."""
    corners, _ = lbvh_scene(n, seed)
    t = tris_from_corners(corners)
    a = uniform(seed, 3 * n, start=21 * n).reshape(n, 3) * F32(0.6) + F32(0.2)
    t["color"][:, :3] = a
    return t


def cloud_camera(w: int, h: int, **kw) -> np.ndarray:
    """The view of cloud_scene: from (0.5, 0.5, -1.5) towards the cube's
    centre, 45 degrees, lit by a white world of radiance 1 (no emitter)."""
    cam = path_camera(w, h, eye=(0.5, 0.5, -1.5), target=(0.5, 0.5, 0.5), up=(0.0, 1.0, 0.0), **kw)
    cam["env"] = (1.0, 1.0, 1.0, 0.0)
    return cam


def corner_boxes(corners: np.ndarray) -> np.ndarray:
    return of_tri(corners[:, 0], corners[:, 1], corners[:, 2])


def batch_counts(count: int, lo: float = 1e3, hi: float = 2e5, seed: int = 0x5C3) -> np.ndarray:
    """Triangle counts of the batch: round(exp(U[ln lo, ln hi])) each."""
    u = (splitmix64(seed, count) >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
    return np.rint(np.exp(np.log(lo) + u * (np.log(hi) - np.log(lo)))).astype(np.int64)


def batch_scenes(count: int = 256, lo: float = 1e3, hi: float = 2e5, seed: int = 0x5C3) -> tuple[np.ndarray, np.ndarray]:
    """Boxes of `count` independent scenes back to back, and the segment starts."""
    counts = batch_counts(count, lo, hi, seed)
    boxes = [corner_boxes(lbvh_scene(int(m), i)[0]) for i, m in enumerate(counts)]
    seg_lo = np.concatenate([[0], np.cumsum(counts)]).astype(np.uint32)
    return np.concatenate(boxes), seg_lo


def batch_tris(count: int = 256, lo: float = 1e3, hi: float = 2e5, seed: int = 0x5C3,
               color=(0.6, 0.6, 0.6, 1.0)) -> tuple[np.ndarray, np.ndarray]:
    """The same scenes as triangle records (one grey Lambertian material), and
    the segment starts."""
    counts = batch_counts(count, lo, hi, seed)
    tris = [tris_from_corners(lbvh_scene(int(m), i)[0], color) for i, m in enumerate(counts)]
    seg_lo = np.concatenate([[0], np.cumsum(counts)]).astype(np.uint32)
    return np.concatenate(tris), seg_lo


def pinhole_rays(w: int = 256, h: int = 256, origin=(0.5, 0.5, -1.5), fov_deg: float = 45.0) -> np.ndarray:
    t = F32(np.tan(np.deg2rad(fov_deg) / 2.0))
    xs = (np.arange(w, dtype=F32) + F32(0.5)) / F32(w) - F32(0.5)
    ys = F32(0.5) - (np.arange(h, dtype=F32) + F32(0.5)) / F32(h)
    dx = (xs * F32(2.0) * t)[None, :].repeat(h, 0).reshape(-1)
    dy = (ys * F32(2.0) * t)[:, None].repeat(w, 1).reshape(-1)
    dz = np.ones(w * h, dtype=F32)
    ln = np.sqrt(dx * dx + dy * dy + dz * dz).astype(F32)
    r = np.zeros(w * h, dtype=RAY)
    r["o"] = np.array(origin, dtype=F32)
    r["tmax"] = F32(3.0e38)
    r["d"] = np.stack([dx / ln, dy / ln, dz / ln], axis=1)
    return r


def icosphere(level: int) -> tuple[np.ndarray, np.ndarray]:
    """Unit icosphere: (vertices [v,3] float64, faces [f,3]); 20 * 4^level faces."""
    p = (1.0 + 5.0 ** 0.5) / 2.0
    v = [(-1, p, 0), (1, p, 0), (-1, -p, 0), (1, -p, 0), (0, -1, p), (0, 1, p), (0, -1, -p), (0, 1, -p),
         (p, 0, -1), (p, 0, 1), (-p, 0, -1), (-p, 0, 1)]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2), (10, 7, 6),
         (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10),
         (8, 6, 7), (9, 8, 1)]
    verts = [np.array(x, dtype=np.float64) / np.linalg.norm(x) for x in v]
    faces = list(f)
    for _ in range(level):
        cache: dict[tuple[int, int], int] = {}

        def mid(i: int, j: int) -> int:
            key = (min(i, j), max(i, j))
            if key not in cache:
                m = verts[i] + verts[j]
                verts.append(m / np.linalg.norm(m))
                cache[key] = len(verts) - 1
            return cache[key]

        nf = []
        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    return np.array(verts), np.array(faces, dtype=np.int64)


def _quad(q: np.ndarray) -> np.ndarray:
    """Two triangles over the diagonal q0-q2 (as tests/emitter.rs `panel`)."""
    return np.array([[q[0], q[1], q[2]], [q[0], q[2], q[3]]], dtype=F32)


def spheres_scene(spheres: int = 64, level: int = 5, seed: int = 1, box=(4.0, 4.0, 1.0), plane_half: float = 3.0,
                  light_height: float = 3.0, light_size: float = 1.0, radiance: float = 10.0) -> np.ndarray:
    """The Rendered-mode workload: icospheres of radius U[0.05, 0.2] with
    centres uniform over a box of `box` resting on the ground plane (z = 0),
    albedo U[0.2, 0.8]^3, one downward-facing emissive quad at `light_height`."""
    vs, fs = icosphere(level)
    u = uniform(seed, spheres * 7).reshape(spheres, 7).astype(np.float64)
    out = []
    for s in range(spheres):
        r = 0.05 + 0.15 * u[s, 0]
        cx = (u[s, 1] - 0.5) * box[0]
        cy = (u[s, 2] - 0.5) * box[1]
        cz = r + u[s, 3] * (box[2] - 2 * r)
        albedo = (0.2 + 0.6 * u[s, 4], 0.2 + 0.6 * u[s, 5], 0.2 + 0.6 * u[s, 6])
        p = (vs * r + np.array([cx, cy, cz])).astype(F32)
        nrm = vs.astype(F32)
        t = np.zeros(len(fs), dtype=TRI)
        t["a"], t["b"], t["c"] = p[fs[:, 0]], p[fs[:, 1]], p[fs[:, 2]]
        t["na"], t["nb"], t["nc"] = nrm[fs[:, 0]], nrm[fs[:, 1]], nrm[fs[:, 2]]
        t["rough"], t["ior"] = 1.0, 1.0
        t["color"] = (*albedo, 1.0)
        out.append(t)
    h = plane_half
    ground = _quad(np.array([[-h, -h, 0], [h, -h, 0], [h, h, 0], [-h, h, 0]], dtype=F32))
    g = np.zeros(2, dtype=TRI)
    g["a"], g["b"], g["c"] = ground[:, 0], ground[:, 1], ground[:, 2]
    g["rough"], g["ior"], g["color"] = 1.0, 1.0, (0.5, 0.5, 0.5, 1.0)
    s = light_size / 2
    # Wound so its face normal points down, at the scene.
    lq = _quad(np.array([[-s, -s, light_height], [-s, s, light_height], [s, s, light_height], [s, -s, light_height]],
                        dtype=F32))
    lt = np.zeros(2, dtype=TRI)
    lt["a"], lt["b"], lt["c"] = lq[:, 0], lq[:, 1], lq[:, 2]
    lt["rough"], lt["ior"], lt["color"] = 1.0, 1.0, (0.0, 0.0, 0.0, 1.0)
    lt["emit"] = (radiance, radiance, radiance)
    return np.concatenate(out + [g, lt])


def path_camera(w: int, h: int, eye=(0.0, -6.0, 2.5), target=(0.0, 0.0, 0.3), fov_deg: float = 45.0,
                bounces: int = 4, rr_from: int = 0, ao: float = 0.0, up=(0.0, 0.0, 1.0)) -> np.ndarray:
    """A perspective `Camera` (src/trace/camera.rs) for the Rendered path:
    view basis from `eye` to `target`, horizontal field of view `fov_deg`,
    Max Bounces (total 4, transparent 8, transmission 8, glossy 0), diffuse
    bounces `bounces`, one unlit lamp slot (so the studio rig stays off, as
    tests/emitter.rs does with `count: 1`), a black world."""
    eye = np.array(eye, dtype=np.float64)
    target = np.array(target, dtype=np.float64)
    fwd = target - eye
    dist = float(np.linalg.norm(fwd))
    fwd /= dist
    right = np.cross(fwd, np.array(up, dtype=np.float64))
    right /= np.linalg.norm(right)
    up = np.cross(right, fwd)
    cam = np.zeros(1, dtype=CAMERA)
    c = cam[0]
    c["right"], c["up"], c["back"] = right, up, -fwd
    c["persp"] = 1
    c["units"] = 1.0
    c["pivot"] = target
    c["dist"] = dist
    c["aspect"] = w / h
    c["scale"] = 0.5 / (dist * np.tan(np.deg2rad(fov_deg) / 2.0))
    c["sun"] = (0.0, 0.0, 1.0)
    c["count"] = 1
    c["bounces"] = bounces
    c["depth"] = (5, 9, 9, 1)
    c["rr_from"] = rr_from
    c["ao"] = ao
    return cam


# ---- scenes of meshes under instances ----------------------------------------
def _signum(x: np.ndarray) -> np.ndarray:
    """Rust f32::signum: the sign bit decides (so -0.0 gives -1)."""
    return np.where(np.signbit(x), F32(-1.0), F32(1.0)).astype(F32)


def oct_pack(n: np.ndarray) -> np.ndarray:
    """src/trace/scene.rs:22-38 in float32: octahedral normal, (2, 2) for a zero normal."""
    n = np.asarray(n, dtype=F32).reshape(-1, 3)
    ax, ay, az = np.abs(n[:, 0]), np.abs(n[:, 1]), np.abs(n[:, 2])
    l = ((ax + ay) + az).astype(F32)
    safe = np.where(l < F32(1e-12), F32(1.0), l).astype(F32)
    x, y, z = (n[:, 0] / safe).astype(F32), (n[:, 1] / safe).astype(F32), (n[:, 2] / safe).astype(F32)
    up = z >= F32(0.0)
    ox = np.where(up, x, ((F32(1.0) - np.abs(y)) * _signum(x)).astype(F32))
    oy = np.where(up, y, ((F32(1.0) - np.abs(x)) * _signum(y)).astype(F32))
    out = np.stack([ox, oy], axis=1).astype(F32)
    out[l < F32(1e-12)] = 0.0
    out[(n == 0).all(axis=1)] = 2.0
    return out


def world_of(m: np.ndarray) -> np.ndarray:
    """A 4 x 4 matrix M (p' = M [p, 1]) as an instance's m[col][row] words."""
    return np.ascontiguousarray(np.asarray(m, dtype=F32).T.reshape(16))


def rotation(axis, angle: float) -> np.ndarray:
    a = np.asarray(axis, dtype=np.float64)
    a = a / np.linalg.norm(a)
    c, s = np.cos(angle), np.sin(angle)
    x, y, z = a
    return np.array([[c + x * x * (1 - c), x * y * (1 - c) - z * s, x * z * (1 - c) + y * s],
                     [y * x * (1 - c) + z * s, c + y * y * (1 - c), y * z * (1 - c) - x * s],
                     [z * x * (1 - c) - y * s, z * y * (1 - c) + x * s, c + z * z * (1 - c)]])


def affine(rot=None, scale=1.0, move=(0.0, 0.0, 0.0)) -> np.ndarray:
    """M = translate(move) @ rot @ scale, as a float32 4 x 4."""
    m = np.eye(4)
    r = np.eye(3) if rot is None else np.asarray(rot, dtype=np.float64)
    m[:3, :3] = r @ np.diag(np.broadcast_to(np.asarray(scale, dtype=np.float64), (3,)))
    m[:3, 3] = move
    return m.astype(F32)


class SceneBuilder:
    """A scene of meshes (own space, own tree) placed by instances, as
    src/trace/scene.rs builds one with `instance`, `add` and `place`."""

    def __init__(self):
        self._verts, self._faces, self._mats, self._meshes, self._insts = [], [], [], [], []
        self._nv = self._nf = 0

    def material(self, color=(0.5, 0.5, 0.5, 1.0), rough=1.0, metal=0.0, trans=0.0, ior=1.0, emit=(0.0, 0.0, 0.0),
                 bump=0.0) -> int:
        from . import MAT
        m = np.zeros(1, dtype=MAT)
        m["color"], m["rough"], m["metal"], m["trans"], m["ior"], m["emit"], m["bump"] = color, rough, metal, trans, ior, emit, bump
        self._mats.append(m)
        return len(self._mats) - 1

    def mesh(self, pos, tris, mat, normals=None) -> int:
        """A mesh over `pos` [v, 3] and faces `tris` [f, 3] (ids into pos);
        `mat` one id or one per face; `normals` [v, 3] or flat."""
        from . import FACE, MESH, VERT
        pos = np.asarray(pos, dtype=F32).reshape(-1, 3)
        tris = np.asarray(tris, dtype=np.int64).reshape(-1, 3)
        v = np.zeros(len(pos), dtype=VERT)
        v["p"][:, :3] = pos
        v["p"][:, 3] = 1.0
        nrm = np.zeros_like(pos) if normals is None else np.asarray(normals, dtype=F32).reshape(-1, 3)
        v["a"][:, :2] = oct_pack(nrm)
        f = np.zeros(len(tris), dtype=FACE)
        f["v"] = tris + self._nv
        f["mat"] = mat
        m = np.zeros(1, dtype=MESH)
        m["face_lo"], m["face_hi"] = self._nf, self._nf + len(tris)
        m["vert_lo"], m["vert_hi"] = self._nv, self._nv + len(pos)
        if len(pos):  # Mesh::fit (src/trace/mesh.rs:56-72)
            m["lo"], m["hi"] = pos.min(0), pos.max(0)
        else:
            m["lo"], m["hi"] = 0.0, 1.0
        self._verts.append(v)
        self._faces.append(f)
        self._meshes.append(m)
        self._nv += len(pos)
        self._nf += len(tris)
        return len(self._meshes) - 1

    def instance(self, mesh: int, m=None, casts_shadow: bool = True, color=(1.0, 1.0, 1.0), hash: int = 0) -> int:
        from . import INSTANCE
        i = np.zeros(1, dtype=INSTANCE)
        i["world"] = world_of(np.eye(4) if m is None else m)
        i["mesh"], i["casts_shadow"], i["color"], i["hash"] = mesh, 1 if casts_shadow else 0, color, hash
        self._insts.append(i)
        return len(self._insts) - 1

    def arrays(self):
        from . import FACE, INSTANCE, MAT, MESH, VERT
        cat = lambda xs, dt: np.concatenate(xs) if xs else np.zeros(0, dtype=dt)  # noqa: E731
        return (cat(self._verts, VERT), cat(self._faces, FACE), cat(self._mats, MAT), cat(self._meshes, MESH),
                cat(self._insts, INSTANCE))


def box_mesh() -> tuple[np.ndarray, np.ndarray]:
    """The unit cube [-1, 1]^3: 8 corners, 12 outward-wound triangles."""
    p = np.array([[x, y, z] for z in (-1, 1) for y in (-1, 1) for x in (-1, 1)], dtype=np.float64)
    q = [(0, 2, 3, 1), (4, 5, 7, 6), (0, 1, 5, 4), (2, 6, 7, 3), (0, 4, 6, 2), (1, 3, 7, 5)]
    t = []
    for a, b, c, d in q:
        t += [(a, b, c), (a, c, d)]
    return p, np.array(t, dtype=np.int64)


def bodies_scene(bodies: int = 1024, seed: int = 0xB0D1, spread: float = 12.0, light: float = 30.0,
                 heights=(0.6, 4.6)):
    """The rigid-body workload: `bodies` shapes, each its own mesh under one
    instance (a box or an icosphere of subdivision 0..5, picked per body),
    radius U[0.2, 0.6], turned about a random axis, placed uniformly over a
    spread x spread square at heights U[heights] above a ground plane, under
    one emissive quad. Returns (builder, world matrices [bodies, 4, 4])."""
    b = SceneBuilder()
    u = uniform(seed, bodies * 9).reshape(bodies, 9).astype(np.float64)
    shapes = [box_mesh()] + [icosphere(k) for k in range(6)]
    mats = [b.material(color=(0.2 + 0.6 * u[i, 6], 0.2 + 0.6 * u[i, 7], 0.2 + 0.6 * u[i, 8], 1.0)) for i in range(8)]
    worlds = []
    for i in range(bodies):
        k = min(int(u[i, 0] * 7), 6)
        pos, tris = shapes[k]
        normals = None if k == 0 else pos
        mesh = b.mesh(pos, tris, mats[i % 8], normals)
        r = 0.2 + 0.4 * u[i, 1]
        ang = 2 * np.pi * u[i, 2]
        axis = (u[i, 3] - 0.5, u[i, 4] - 0.5, 0.5)
        move = ((u[i, 5] - 0.5) * spread, (u[i, 6] - 0.5) * spread, heights[0] + (heights[1] - heights[0]) * u[i, 7])
        m = affine(rotation(axis, ang), r, move)
        worlds.append(m)
        b.instance(mesh, m)
    h = spread
    ground = b.mesh([[-h, -h, 0], [h, -h, 0], [h, h, 0], [-h, h, 0]], [[0, 1, 2], [0, 2, 3]], b.material())
    b.instance(ground)
    s = 1.5
    lamp = b.mesh([[-s, -s, 0], [-s, s, 0], [s, s, 0], [s, -s, 0]], [[0, 1, 2], [0, 2, 3]],
                  b.material(color=(0, 0, 0, 1), emit=(light, light, light)))
    b.instance(lamp, affine(move=(0.0, 0.0, 8.0)))
    return b, np.array(worlds)


def bodies_defs(bodies: int = 1024, seed: int = 0xB0D1, spread: float = 12.0, heights=(0.6, 4.6)) -> list[dict]:
    """The rigid bodies of bodies_scene (the same draws): per body the
    collision shape ("box" of half extent r, or "sphere" of radius r for an
    icosphere), its centre, its orientation as a quaternion (xyzw) and its
    scale r, for World::add_body."""
    u = uniform(seed, bodies * 9).reshape(bodies, 9).astype(np.float64)
    out = []
    for i in range(bodies):
        k = min(int(u[i, 0] * 7), 6)
        r = 0.2 + 0.4 * u[i, 1]
        ang = 2 * np.pi * u[i, 2]
        axis = np.array((u[i, 3] - 0.5, u[i, 4] - 0.5, 0.5))
        axis /= np.linalg.norm(axis)
        q = (*(axis * np.sin(ang / 2)), np.cos(ang / 2))
        move = ((u[i, 5] - 0.5) * spread, (u[i, 6] - 0.5) * spread, heights[0] + (heights[1] - heights[0]) * u[i, 7])
        out.append({"shape": "box" if k == 0 else "sphere", "r": r, "x": move, "q": q})
    return out

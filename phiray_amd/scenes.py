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


def corner_boxes(corners: np.ndarray) -> np.ndarray:
    return of_tri(corners[:, 0], corners[:, 1], corners[:, 2])


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
                bounces: int = 4, rr_from: int = 0, ao: float = 0.0) -> np.ndarray:
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
    right = np.cross(fwd, [0.0, 0.0, 1.0])
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

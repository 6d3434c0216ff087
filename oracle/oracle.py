"""ctypes loader for the CPU oracle (liboracle.so). TEST INFRASTRUCTURE ONLY.

Used by tests/, the smoke check and the benchmark's CPU-baseline leg as the
checker; never by the product library. `build()` compiles it with the
Makefile beside this file.
"""
from __future__ import annotations

import ctypes as C
import pathlib
import subprocess

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"

_P = C.c_void_p
_U32 = C.c_uint32
_I = C.c_int
_SIGS = {
    "orc_code": (_U32, [_P]),
    "orc_morton": (C.c_uint64, [_P, _U32]),
    "orc_of_tri": (None, [_P, _P, _P, _P]),
    "orc_scene": (None, [_P, _U32, _P]),
    "orc_codes": (None, [_P, _U32, _P, _I]),
    "orc_sort_codes": (None, [_P, _U32, _P, _P, _I]),
    "orc_radix_model": (None, [_P, _U32, _P, _P]),
    "orc_delta": (C.c_int32, [_P, _U32, C.c_int64, C.c_int64]),
    "orc_links": (None, [_P, _U32, _P, _I]),
    "orc_assemble": (_U32, [_P, _U32, _U32, _P]),
    "orc_refit": (None, [_P, _U32, _P, _P]),
    "orc_build": (_U32, [_P, _U32, _U32, _P, _P, _I]),
    "orc_build_compact": (_U32, [_P, _U32, _U32, _P, _P, _P]),
    "orc_compact": (_U32, [_P, _U32]),
    "orc_reachable": (_U32, [_P, _U32, _P]),
    "orc_ray_space": (None, [_P, _P]),
    "orc_tri_t": (_I, [_P, _P, _P, _P, _P, _P]),
    "orc_moller_trumbore": (_I, [_P, _P, _P, _P, _P, _P]),
    "orc_safe_inv": (None, [_P, _P]),
    "orc_box_t": (C.c_float, [_P, _P, _P, C.c_float]),
    "orc_raycast": (None, [_P, _P, _P, _U32, _I, _P, _P, _I]),
    "orc_raycast_mat": (None, [_P, _P, _P, _P, _P, _U32, _I, _P, _P, _I]),
    "orc_sin": (C.c_float, [C.c_float]),
    "orc_cos": (C.c_float, [C.c_float]),
    "orc_exp2": (C.c_float, [C.c_float]),
    "orc_hash": (_U32, [_U32]),
    "orc_rnd": (C.c_float, [C.POINTER(_U32)]),
    "orc_scene_build": (_U32, [_P, _U32, _U32, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _I]),
    "orc_stamp_camera": (None, [_P, _U32, _U32, C.c_float, _U32, _U32]),
    "orc_sample": (None, [_P, _P, _U32, _U32, _P, _P, _I]),
    "orc_hardware_threads": (_I, []),
    "orc_log2": (C.c_float, [C.c_float]),
    "orc_denoise": (None, [_P, _U32, _U32, _U32, _P]),
    "orc_adapt_check": (None, [_P, _P]),
    "orc_adapt_dilate": (None, [_P, _P]),
    "orc_denoise_pass": (None, [_P, _U32, _U32, _U32, _U32, _U32, _U32, _P, _P]),
    "orc_view_transform": (None, [_P, _P, _P]),
    "orc_inst_inverse": (None, [_P, _P]),
    "orc_inst_mirrors": (_U32, [_P]),
    "orc_inst_to_world": (None, [_P, _P, _P]),
    "orc_inst_world_box": (None, [_P, _P, _P]),
    "orc_mesh_texspace": (None, [_P, _P, _P]),
    "orc_mesh_fit": (None, [_P, _P]),
    "orc_inst_words": (None, [_P, _U32, _P, _P]),
    "orc_forest_node_count": (_U32, [_P]),
    "orc_forest_inst_len": (_U32, [_P]),
    "orc_forest_lights_len": (_U32, [_P]),
    "orc_forest_build": (_U32, [_P, _U32, _P, _P, _P, _P, _P, _P, _P, _P, _I]),
    "orc_forest_reinstance": (None, [_P, _P, _P, _P, _P, _P, _P]),
    "orc_trace_scene": (None, [_P, _P, _U32, _I, _P, _P, _I]),
    "orc_world_new": (_P, []),
    "orc_world_free": (None, [_P]),
    "orc_world_configure": (None, [_P, _U32, _U32, C.c_float, C.c_float, C.c_float, _P]),
    "orc_world_add_shape": (_I, [_P, _P, C.POINTER(_U32)]),
    "orc_world_add_body": (_I, [_P, _P, C.POINTER(_U32)]),
    "orc_world_step": (_I, [_P, C.c_float]),
    "orc_world_body_count": (_U32, [_P]),
    "orc_world_contact_count": (_U32, [_P]),
    "orc_world_bodies": (_I, [_P, _P]),
    "orc_world_poses": (_I, [_P, _P]),
    "orc_world_energy": (C.c_float, [_P]),
    "orc_world_pair_list": (_I, [_P, _P, _U32, C.POINTER(_U32)]),
}

_lib = None


def build() -> pathlib.Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        h = C.CDLL(str(LIB))
        for name, (res, args) in _SIGS.items():
            f = getattr(h, name)
            f.restype = res
            f.argtypes = args
        _lib = h
    return _lib


def ptr(a: np.ndarray) -> C.c_void_p:
    return C.c_void_p(a.ctypes.data)


def threads() -> int:
    return int(lib().orc_hardware_threads())


# ---- LBVH -------------------------------------------------------------------
def codes(prims: np.ndarray, nthreads: int = 1) -> np.ndarray:
    prims = np.ascontiguousarray(prims)
    out = np.zeros(len(prims), dtype=np.uint32)
    lib().orc_codes(ptr(prims), len(prims), ptr(out), nthreads)
    return out


def sort_codes(c: np.ndarray, nthreads: int = 1) -> tuple[np.ndarray, np.ndarray]:
    c = np.ascontiguousarray(c, dtype=np.uint32)
    order = np.zeros(len(c), dtype=np.uint32)
    srt = np.zeros(len(c), dtype=np.uint32)
    lib().orc_sort_codes(ptr(c), len(c), ptr(order), ptr(srt), nthreads)
    return order, srt


def radix_model(c: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    c = np.ascontiguousarray(c, dtype=np.uint32)
    order = np.zeros(len(c), dtype=np.uint32)
    srt = np.zeros(len(c), dtype=np.uint32)
    lib().orc_radix_model(ptr(c), len(c), ptr(order), ptr(srt))
    return order, srt


def links(sorted_codes: np.ndarray, link_dtype, nthreads: int = 1) -> np.ndarray:
    s = np.ascontiguousarray(sorted_codes, dtype=np.uint32)
    out = np.zeros(max(len(s) - 1, 0), dtype=link_dtype)
    lib().orc_links(ptr(s), len(s), ptr(out), nthreads)
    return out


def build_bvh(prims: np.ndarray, leaf: int, node_dtype, nthreads: int = 1):
    prims = np.ascontiguousarray(prims)
    n = len(prims)
    nodes = np.zeros(max(2 * n - 1, 0), dtype=node_dtype)
    order = np.zeros(n, dtype=np.uint32)
    depth = lib().orc_build(ptr(prims), n, leaf, ptr(nodes), ptr(order), nthreads)
    return nodes, order, int(depth)


def refit(nodes: np.ndarray, order: np.ndarray, prims: np.ndarray) -> np.ndarray:
    nodes = nodes.copy()
    prims = np.ascontiguousarray(prims)
    lib().orc_refit(ptr(nodes), len(nodes), ptr(np.ascontiguousarray(order, dtype=np.uint32)), ptr(prims))
    return nodes


def reachable(nodes: np.ndarray) -> np.ndarray:
    out = np.zeros(len(nodes), dtype=np.uint32)
    m = lib().orc_reachable(ptr(nodes), len(nodes), ptr(out))
    return out[:m]


# ---- walk ---------------------------------------------------------------------
def raycast(nodes: np.ndarray, tris9: np.ndarray, rays: np.ndarray, hit_dtype, any_hit: bool = False,
            nthreads: int = 1, stats: bool = False, faces: np.ndarray | None = None, mats: np.ndarray | None = None):
    rays = np.ascontiguousarray(rays)
    hits = np.zeros(len(rays), dtype=hit_dtype)
    st = np.zeros((len(rays), 2), dtype=np.uint32) if stats else None
    lib().orc_raycast_mat(ptr(nodes), ptr(np.ascontiguousarray(tris9, dtype=np.float32)),
                          ptr(faces) if faces is not None else None, ptr(mats) if mats is not None else None,
                          ptr(rays), len(rays), 1 if any_hit else 0, ptr(hits), ptr(st) if stats else None, nthreads)
    return (hits, st) if stats else hits


# ---- tracer -------------------------------------------------------------------
class OrcScene(C.Structure):
    _fields_ = [("nodes", _P), ("node_count", _U32), ("tris", _P), ("faces", _P), ("verts", _P),
                ("mats", _P), ("lights", _P), ("insts", _P), ("ninsts", _U32), ("forest", _U32)]


INSTANCE_LEN = 40
INST = 0x40000000
# src/trace/instance.rs:23-44 and src/trace/mesh.rs:12-22 as the oracle takes them.
INSTANCE = np.dtype([("world", "<f4", (16,)), ("mesh", "<u4"), ("casts_shadow", "<u4"), ("color", "<f4", (3,)),
                     ("hash", "<u4")])
MESH = np.dtype([("face_lo", "<u4"), ("face_hi", "<u4"), ("vert_lo", "<u4"), ("vert_hi", "<u4"), ("lo", "<f4", (3,)),
                 ("hi", "<f4", (3,))])


class OrcSceneIn(C.Structure):
    _fields_ = [("verts", _P), ("nverts", _U32), ("faces", _P), ("nfaces", _U32), ("mats", _P), ("nmats", _U32),
                ("meshes", _P), ("nmeshes", _U32), ("insts", _P), ("ninsts", _U32)]


def inst_words(inst: np.ndarray, blas_root: int, mesh: np.ndarray) -> np.ndarray:
    out = np.zeros(INSTANCE_LEN, dtype=np.uint32)
    lib().orc_inst_words(ptr(np.ascontiguousarray(inst)), blas_root, ptr(np.ascontiguousarray(mesh)), ptr(out))
    return out


def inst_world_box(world: np.ndarray, box: np.ndarray) -> np.ndarray:
    out = np.zeros(1, dtype=box.dtype)
    lib().orc_inst_world_box(ptr(np.ascontiguousarray(world, dtype=np.float32)), ptr(np.ascontiguousarray(box)), ptr(out))
    return out[0]


def inst_inverse(world: np.ndarray) -> np.ndarray:
    out = np.zeros(12, dtype=np.float32)
    lib().orc_inst_inverse(ptr(np.ascontiguousarray(world, dtype=np.float32)), ptr(out))
    return out.reshape(3, 4)


class Forest:
    """The oracle's scene of meshes under instances (Forest::build with LBVH trees)."""

    def __init__(self, verts, faces, mats, meshes, insts, leaf: int, node_dtype, nthreads: int = 1):
        words = lambda a, dt: np.ascontiguousarray(a).view(dt).reshape(-1).copy()  # noqa: E731
        self.verts = words(verts, np.float32)
        self.in_faces = words(faces, np.uint32)
        self.mats = words(mats, np.float32)
        self.meshes = np.ascontiguousarray(meshes, dtype=MESH)
        self.insts = np.ascontiguousarray(insts, dtype=INSTANCE)
        self.node_dtype = node_dtype
        self.n = len(self.in_faces) // 4
        self._in = OrcSceneIn(ptr(self.verts), len(self.verts) // 8, ptr(self.in_faces), self.n, ptr(self.mats),
                              len(self.mats) // 16, ptr(self.meshes), len(self.meshes), ptr(self.insts), len(self.insts))
        L = lib()
        self.nodes = np.zeros(L.orc_forest_node_count(C.byref(self._in)), dtype=node_dtype)
        self.order = np.zeros(self.n, dtype=np.uint32)
        self.tri9 = np.zeros(9 * self.n, dtype=np.float32)
        self.faces = np.zeros(4 * self.n, dtype=np.uint32)
        self.inst_words = np.zeros(L.orc_forest_inst_len(C.byref(self._in)), dtype=np.uint32)
        self.lights = np.zeros(L.orc_forest_lights_len(C.byref(self._in)), dtype=np.uint32)
        self.depths = np.zeros(len(self.meshes) + 1, dtype=np.uint32)
        ext = C.c_float(0)
        L.orc_forest_build(C.byref(self._in), leaf, ptr(self.nodes), ptr(self.order), ptr(self.tri9), ptr(self.faces),
                           ptr(self.inst_words), ptr(self.lights), ptr(self.depths), C.byref(ext), nthreads)
        self.extent = ext.value
        self._finish()

    def _finish(self):
        m = int(self.lights[0])
        self.lights_used = self.lights[: 1 + 3 * m].copy()
        self.glow = m
        self._c = OrcScene(ptr(self.nodes), len(self.nodes), ptr(self.tri9), ptr(self.faces), ptr(self.verts),
                           ptr(self.mats), ptr(self.lights_used), ptr(self.inst_words), len(self.insts), 1)

    def reinstance(self, worlds: np.ndarray) -> None:
        """New matrices (pose.rs:107-124): the top tree, instance and light tables again."""
        self.insts["world"] = np.asarray(worlds, dtype=np.float32).reshape(len(self.insts), 16)
        d, ext = C.c_uint32(0), C.c_float(0)
        lib().orc_forest_reinstance(C.byref(self._in), ptr(self.nodes), ptr(self.inst_words), ptr(self.lights),
                                    ptr(self.faces), C.byref(d), C.byref(ext))
        self.depths[-1] = d.value
        self.extent = ext.value
        self._finish()

    def raycast(self, rays: np.ndarray, hit_dtype, any_hit: bool = False, nthreads: int = 1, stats: bool = False):
        rays = np.ascontiguousarray(rays)
        hits = np.zeros(len(rays), dtype=hit_dtype)
        st = np.zeros(2 * len(rays), dtype=np.uint32)
        lib().orc_trace_scene(C.byref(self._c), ptr(rays), len(rays), 1 if any_hit else 0, ptr(hits), ptr(st), nthreads)
        return (hits, st) if stats else hits

    def stamp(self, cam: np.ndarray, w: int, h: int, frame: int) -> np.ndarray:
        c = np.ascontiguousarray(np.array(cam).reshape(1).copy())
        lib().orc_stamp_camera(ptr(c), w, h, C.c_float(self.extent), self.glow, frame)
        return c

    def sample(self, cam: np.ndarray, w: int, h: int, px: np.ndarray, frame: int, nthreads: int = 1,
               first: int = 0, count: int | None = None) -> tuple[int, int]:
        c = self.stamp(cam, w, h, frame)
        rays = np.zeros(2, dtype=np.uint64)
        n = w * h
        count = n - first if count is None else count
        lib().orc_sample(C.byref(self._c), ptr(c), first, count, ptr(px), ptr(rays), nthreads)
        return int(rays[0]), int(rays[1])

    def adapt_tail(self, cam: np.ndarray, w: int, h: int, px: np.ndarray, frame: int) -> None:
        """Adaptive Sampling's check and dilation after sample `frame`."""
        c = self.stamp(cam, w, h, frame)
        lib().orc_adapt_check(ptr(c), ptr(px))
        lib().orc_adapt_dilate(ptr(c), ptr(px))


class Scene:
    """The oracle's tracer scene (all arrays kept alive here)."""

    def __init__(self, tris: np.ndarray, leaf: int, node_dtype, nthreads: int = 1):
        tris = np.ascontiguousarray(tris)
        n = len(tris)
        self.n = n
        self.nodes = np.zeros(max(2 * n - 1, 1), dtype=node_dtype)
        self.order = np.zeros(n, dtype=np.uint32)
        self.tri9 = np.zeros(9 * n, dtype=np.float32)
        self.faces = np.zeros(4 * n, dtype=np.uint32)
        self.verts = np.zeros(8 * 3 * n, dtype=np.float32)
        self.mats = np.zeros(16 * n, dtype=np.float32)
        self.lights = np.zeros(1 + 3 * n, dtype=np.uint32)
        nv, nm, depth = C.c_uint32(0), C.c_uint32(0), C.c_uint32(0)
        ext = C.c_float(0)
        lib().orc_scene_build(ptr(tris), n, leaf, ptr(self.nodes), ptr(self.order), ptr(self.tri9), ptr(self.faces),
                              ptr(self.verts), C.byref(nv), ptr(self.mats), C.byref(nm), ptr(self.lights),
                              C.byref(depth), C.byref(ext), nthreads)
        self.nverts, self.nmats, self.depth, self.extent = nv.value, nm.value, depth.value, ext.value
        m = int(self.lights[0])
        self.lights = self.lights[: 1 + 3 * m].copy()
        self.glow = m
        self._c = OrcScene(ptr(self.nodes), len(self.nodes), ptr(self.tri9), ptr(self.faces), ptr(self.verts),
                           ptr(self.mats), ptr(self.lights), None, 0, 0)

    def stamp(self, cam: np.ndarray, w: int, h: int, frame: int) -> np.ndarray:
        c = np.ascontiguousarray(np.array(cam).reshape(1).copy())
        lib().orc_stamp_camera(ptr(c), w, h, C.c_float(self.extent), self.glow, frame)
        return c

    def sample(self, cam: np.ndarray, w: int, h: int, px: np.ndarray, frame: int, nthreads: int = 1,
               first: int = 0, count: int | None = None) -> tuple[int, int]:
        c = self.stamp(cam, w, h, frame)
        rays = np.zeros(2, dtype=np.uint64)
        n = w * h
        count = n - first if count is None else count
        lib().orc_sample(C.byref(self._c), ptr(c), first, count, ptr(px), ptr(rays), nthreads)
        return int(rays[0]), int(rays[1])

    def adapt_tail(self, cam: np.ndarray, w: int, h: int, px: np.ndarray, frame: int) -> None:
        """Adaptive Sampling's check and dilation after sample `frame`."""
        c = self.stamp(cam, w, h, frame)
        lib().orc_adapt_check(ptr(c), ptr(px))
        lib().orc_adapt_dilate(ptr(c), ptr(px))


# ---- the rigid-body World (src/sim/), sim.cpp ----------------------------------
SHAPE = np.dtype([("kind", "<u4"), ("data", "<u4"), ("count", "<u4"), ("radius", "<f4"), ("half", "<f4", 3), ("_a", "<f4")])
BODY = np.dtype([("x", "<f4", 3), ("inv_mass", "<f4"), ("q", "<f4", 4), ("v", "<f4", 3), ("shape", "<u4"),
                 ("w", "<f4", 3), ("flags", "<u4")])
BODY_DEF = np.dtype([("kind", "<u4"), ("shape", "<u4"), ("mass", "<f4"), ("x", "<f4", 3), ("q", "<f4", 4), ("v", "<f4", 3),
                     ("w", "<f4", 3), ("scale", "<f4", 3), ("com", "<f4", 3), ("friction", "<f4"), ("restitution", "<f4"),
                     ("margin", "<f4"), ("lin_damp", "<f4"), ("ang_damp", "<f4"), ("sleep", "<f4", 2), ("has_sleep", "<u4"),
                     ("asleep", "<u4"), ("groups", "<u4"), ("obj", "<u4")])
PASSIVE = 1


class World:
    """phiray's XPBD `World` (src/sim/world.rs) restated on the CPU, with the
    same methods as the product's World."""

    BOX, SPHERE, CAPSULE = 0, 1, 2

    def __init__(self):
        self._h = C.c_void_p(lib().orc_world_new())

    def configure(self, substeps: int = 10, iterations: int = 10, time_scale: float = 1.0, max_recover: float = 2.0,
                  sleep_after: float = 2.0, gravity=(0.0, 0.0, -9.81)) -> None:
        g = np.array(gravity, dtype=np.float32)
        lib().orc_world_configure(self._h, substeps, iterations, time_scale, max_recover, sleep_after, ptr(g))

    def add_shape(self, kind: int, half=(0.0, 0.0, 0.0), radius: float = 0.0) -> int:
        s = np.zeros(1, dtype=SHAPE)
        s["kind"], s["half"], s["radius"] = kind, half, radius
        out = C.c_uint32(0)
        if lib().orc_world_add_shape(self._h, ptr(s), C.byref(out)) != 0:
            raise ValueError("orc_world_add_shape refused the shape")
        return int(out.value)

    def add_box(self, half) -> int:
        return self.add_shape(self.BOX, half=half)

    def add_sphere(self, radius: float) -> int:
        return self.add_shape(self.SPHERE, radius=radius)

    def add_capsule(self, radius: float, half: float) -> int:
        return self.add_shape(self.CAPSULE, half=(radius, radius, half), radius=radius)

    @staticmethod
    def body_def(**kw) -> np.ndarray:
        """A BodyDef with phiray's defaults (src/sim/build.rs:39-61)."""
        d = np.zeros(1, dtype=BODY_DEF)
        d["mass"], d["q"], d["scale"] = 1.0, (0.0, 0.0, 0.0, 1.0), (1.0, 1.0, 1.0)
        d["friction"], d["lin_damp"], d["ang_damp"], d["groups"] = 0.5, 0.04, 0.1, 1
        for k, v in kw.items():
            if k == "sleep":
                if v is not None:
                    d["sleep"], d["has_sleep"] = v, 1
                continue
            d[k] = v
        return d

    def add_body(self, **kw) -> int:
        d = self.body_def(**kw)
        out = C.c_uint32(0)
        if lib().orc_world_add_body(self._h, ptr(d), C.byref(out)) != 0:
            raise ValueError("orc_world_add_body refused the body")
        return int(out.value)

    def step(self, dt: float) -> None:
        lib().orc_world_step(self._h, dt)

    def __len__(self) -> int:
        return int(lib().orc_world_body_count(self._h))

    def bodies(self) -> np.ndarray:
        out = np.zeros(len(self), dtype=BODY)
        lib().orc_world_bodies(self._h, ptr(out))
        return out

    def poses(self) -> np.ndarray:
        out = np.zeros((len(self), 16), dtype=np.float32)
        lib().orc_world_poses(self._h, ptr(out))
        return out

    def energy(self) -> float:
        return float(lib().orc_world_energy(self._h))

    def contact_count(self) -> int:
        return int(lib().orc_world_contact_count(self._h))

    def pairs(self) -> np.ndarray:
        n = C.c_uint32(0)
        lib().orc_world_pair_list(self._h, None, 0, C.byref(n))
        out = np.zeros((n.value, 2), dtype=np.uint32)
        lib().orc_world_pair_list(self._h, ptr(out), n.value, C.byref(n))
        return out

    def __del__(self):
        try:
            if self._h:
                lib().orc_world_free(self._h)
                self._h = C.c_void_p()
        except Exception:
            pass

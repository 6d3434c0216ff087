"""Host-side mirror of phiray's GPU API over the gfx950 HIP library.

`Gpu`, `GpuArray`, `GpuBvh` and `Tracer` follow phiray's Rust types of the same
names (src/gpu/mod.rs, src/gpu/array.rs, src/bvh/gpu.rs, src/trace/mod.rs):
the same constructors, argument meanings and error behaviour (`None` for no
device or an empty scene, `False` from `update` when the count changed). Every
call goes through the C ABI declared in include/phiray_hip.h; there is no
CPU path. Importing this package loads libphiray_hip.so and fails loudly when
it has not been built (``python -m phiray_amd.build``).
"""
from __future__ import annotations

import ctypes as C
import pathlib

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
LIB_PATH = HERE / "libphiray_hip.so"

# ---- record layouts (phiray's own, src/bvh/node.rs, src/trace/...) ----------
AABB = np.dtype([("lo", "<f4", 3), ("_a", "<f4"), ("hi", "<f4", 3), ("_b", "<f4")])
NODE = np.dtype([("lo", "<f4", 3), ("li", "<u4"), ("hi", "<f4", 3), ("ri", "<u4")])
LINK = np.dtype([("left", "<u4"), ("right", "<u4"), ("first", "<u4"), ("last", "<u4")])
HIT = np.dtype([("t", "<f4"), ("tri", "<u4"), ("u", "<f4"), ("v", "<f4"), ("inst", "<u4")])
RAY = np.dtype([("o", "<f4", 3), ("tmax", "<f4"), ("d", "<f4", 3), ("_pad", "<f4")])
TRI = np.dtype([
    ("a", "<f4", 3), ("rough", "<f4"), ("na", "<f4", 3), ("trans", "<f4"),
    ("b", "<f4", 3), ("metal", "<f4"), ("nb", "<f4", 3), ("ior", "<f4"),
    ("c", "<f4", 3), ("_c", "<f4"), ("nc", "<f4", 3), ("_nc", "<f4"),
    ("color", "<f4", 4), ("emit", "<f4", 3), ("bump", "<f4"),
])
CAMERA = np.dtype([
    ("right", "<f4", 3), ("persp", "<u4"), ("up", "<f4", 3), ("scale", "<f4"),
    ("back", "<f4", 3), ("units", "<f4"), ("pivot", "<f4", 3), ("aspect", "<f4"),
    ("sun", "<f4", 3), ("dist", "<f4"), ("pan", "<f4", 2), ("ao", "<f4"), ("eps", "<f4"),
    ("env", "<f4", 4), ("frame", "<u4"), ("count", "<u4"), ("glow", "<u4"), ("bounces", "<u4"),
    ("depth", "<u4", 4), ("integrator", "<f4", 4), ("w", "<u4"), ("h", "<u4"), ("rr_from", "<u4"), ("stream", "<u4"),
    ("adapt", "<f4", 4), ("view", "<f4", 4),
])
assert AABB.itemsize == 32 and NODE.itemsize == 32 and HIT.itemsize == 20
assert RAY.itemsize == 32 and TRI.itemsize == 128 and CAMERA.itemsize == 208

# phiray's own records, byte for byte: `Tri` (src/trace/kernel.rs:11-65, with
# three 44-word procedural slots, src/trace/proc.rs:41-254), `Light`
# (src/trace/light.rs:14-43) and `Camera` (src/trace/camera.rs:12-228, 64
# lamps). phr_tri_from_phiray / phr_camera_from_phiray turn them into TRI and
# CAMERA above.
NO_TEX = 0xFFFFFFFF  # src/trace/kernel.rs:70
NO_MAT = 0xFFFFFFFF  # src/trace/mat.rs:49
MAX_LIGHT = 64  # src/trace/camera.rs:332
PROC = np.dtype([("kind", "<u4"), ("rest", "<u4", 43)])
PHIRAY_TRI = np.dtype([
    ("a", "<f4", 3), ("rough", "<f4"), ("na", "<f4", 3), ("trans", "<f4"),
    ("b", "<f4", 3), ("metal", "<f4"), ("nb", "<f4", 3), ("ior", "<f4"),
    ("c", "<f4", 3), ("_c", "<f4"), ("nc", "<f4", 3), ("_nc", "<f4"),
    ("color", "<f4", 4), ("us", "<f4", 3), ("tex", "<u4"), ("vs", "<f4", 3), ("nmap", "<u4"),
    ("emit", "<f4", 3), ("bump", "<f4"), ("maps", "<u4", 3), ("sss", "<f4"), ("sss_radius", "<f4", 3),
    ("sheen", "<f4"), ("sheen_rough", "<f4"), ("sheen_tint", "<f4", 3), ("coat", "<f4"), ("coat_rough", "<f4"),
    ("aniso", "<f4"), ("aniso_rot", "<f4"), ("second", "<u4"), ("mix", "<f4"),
    ("color_proc", PROC), ("bump_proc", PROC), ("color_proc2", PROC),
])
PHIRAY_LIGHT = np.dtype([
    ("pos", "<f4", 3), ("kind", "<u4"), ("color", "<f4", 3), ("size", "<f4"), ("dir", "<f4", 3), ("cone", "<f4"),
    ("right", "<f4", 3), ("size_y", "<f4"), ("shape", "<u4"), ("blend", "<f4"), ("spot_scale", "<f4", 2),
])
PHIRAY_CAMERA = np.dtype([
    ("n", "<u4"), ("groups_x", "<u4"), ("w", "<u4"), ("h", "<u4"),
    ("right", "<f4", 3), ("persp", "<u4"), ("up", "<f4", 3), ("scale", "<f4"),
    ("back", "<f4", 3), ("units", "<f4"), ("pivot", "<f4", 3), ("aspect", "<f4"),
    ("sun", "<f4", 3), ("dist", "<f4"), ("sky", "<f4", 4), ("pan", "<f4", 2), ("ao", "<f4"), ("eps", "<f4"),
    ("frame", "<u4"), ("grid", "<f4"), ("count", "<u4"), ("glow", "<u4"), ("grid_color", "<f4", 4),
    ("env", "<f4", 4), ("split", "<u4", 4), ("view", "<f4", 4),
    ("envmap", "<u4"), ("env_rot", "<f4"), ("bounces", "<u4"), ("env_cdf", "<u4"),
    ("aperture", "<f4"), ("focus", "<f4"), ("blades", "<u4"), ("blade_spin", "<f4"),
    ("aperture_ratio", "<f4"), ("env_hidden", "<u4"), ("see_through", "<f4"), ("grid_hide", "<u4"),
    ("tile", "<u4", 4), ("film", "<u4", 4), ("preview", "<u4", 4), ("slice", "<u4", 4), ("insts", "<u4", 4),
    ("vis", "<u4", 4), ("taa", "<f4", 4), ("fx", "<f4", 4), ("outline", "<f4", 4), ("shade", "<u4", 4),
    ("single", "<f4", 4), ("cavity", "<f4", 4), ("ssao", "<u4", 4), ("shadow", "<f4", 4),
    ("shadow_fade", "<f4", 4), ("matcap", "<u4", 4), ("dof", "<f4", 4), ("outline_fx", "<f4", 4),
    ("outline_pick", "<f4", 4), ("outline_active", "<f4", 4), ("outline_move", "<f4", 4),
    ("adapt", "<f4", 4), ("limit", "<f4", 4), ("depth", "<u4", 4), ("integrator", "<f4", 4),
    ("lights", PHIRAY_LIGHT, MAX_LIGHT),
])
assert PROC.itemsize == 176 and PHIRAY_TRI.itemsize == 760
assert PHIRAY_LIGHT.itemsize == 80 and PHIRAY_CAMERA.itemsize == 5776


def phiray_tris(n: int) -> np.ndarray:
    """n zeroed phiray `Tri` records with `Tri::new`'s defaults
    (src/trace/kernel.rs:72-74): rough 0.5, IOR 1.5, no textures, sheen
    roughness 0.5 and white tint, coat roughness 0.03, the default
    subsurface radius, no second material."""
    t = np.zeros(n, dtype=PHIRAY_TRI)
    t["rough"], t["ior"] = 0.5, 1.5
    t["tex"], t["nmap"], t["maps"] = NO_TEX, NO_TEX, NO_TEX
    t["sheen_rough"], t["sheen_tint"], t["coat_rough"] = 0.5, 1.0, 0.03
    t["sss_radius"] = (0.05, 0.01, 0.005)
    t["second"] = NO_MAT
    return t


def tris_from_phiray(tris: np.ndarray) -> np.ndarray:
    """phiray `Tri` records (PHIRAY_TRI) into the tracer's TRI records
    through phr_tri_from_phiray; PhrError for a record outside the subset."""
    src = np.ascontiguousarray(tris, dtype=PHIRAY_TRI)
    out = np.zeros(len(src), dtype=TRI)
    _check(lib().phr_tri_from_phiray(_ptr(src), len(src), _ptr(out)), "phr_tri_from_phiray")
    return out


def camera_from_phiray(cam: np.ndarray) -> np.ndarray:
    """A phiray `Camera` record (PHIRAY_CAMERA) into the tracer's CAMERA
    through phr_camera_from_phiray; PhrError for a frame outside the subset."""
    src = np.ascontiguousarray(np.asarray(cam, dtype=PHIRAY_CAMERA).reshape(1))
    out = np.zeros(1, dtype=CAMERA)
    _check(lib().phr_camera_from_phiray(_ptr(src), _ptr(out)), "phr_camera_from_phiray")
    return out


# A scene of meshes under instances (src/trace/scene.rs:40-67, mesh.rs:12-22,
# instance.rs:23-44): a vertex is its object-space position (w 1) then its
# packed normal and uv; a face three vertex ids and its material; a material
# 16 floats (color rgba, rough, metal, trans, ior, emit rgb, bump, 0 x 4); an
# instance's world matrix is m[col][row], translation in m[3].
VERT = np.dtype([("p", "<f4", 4), ("a", "<f4", 4)])
FACE = np.dtype([("v", "<u4", 3), ("mat", "<u4")])
MAT = np.dtype([("color", "<f4", 4), ("rough", "<f4"), ("metal", "<f4"), ("trans", "<f4"), ("ior", "<f4"),
                ("emit", "<f4", 3), ("bump", "<f4"), ("_pad", "<f4", 4)])
MESH = np.dtype([("face_lo", "<u4"), ("face_hi", "<u4"), ("vert_lo", "<u4"), ("vert_hi", "<u4"), ("lo", "<f4", 3),
                 ("hi", "<f4", 3)])
INSTANCE = np.dtype([("world", "<f4", 16), ("mesh", "<u4"), ("casts_shadow", "<u4"), ("color", "<f4", 3),
                     ("hash", "<u4")])
INSTANCE_LEN = 40  # words of an instance record (src/trace/instance.rs:18)
INST = 0x40000000  # a top-tree leaf (src/trace/forest.rs:17)
LEAF = 0x80000000
# phr_bvh_build_static: the most boxes its static scratch holds.
STATIC_BUILD_MAX = 262144
NONE = 0xFFFFFFFF
STACK = 32
PX_LEN = 34  # words a pixel in the image buffer (src/trace/read.rs:56)

# C ABI entry points and their signatures (include/phiray_hip.h).
_P = C.c_void_p
_I = C.c_int
_U32 = C.c_uint32
_U64 = C.c_uint64
_SIGS = {
    "phr_last_error": (C.c_char_p, []),
    "phr_abi_version": (_U32, []),
    "phr_gpu_open": (_I, [_I, C.POINTER(_P)]),
    "phr_gpu_shared": (_I, [C.POINTER(_P)]),
    "phr_gpu_close": (None, [_P]),
    "phr_gpu_wait": (_I, [_P]),
    "phr_gpu_wait_for": (_I, [_P, _U64, C.POINTER(_I)]),
    "phr_gpu_max_binding": (_U64, [_P]),
    "phr_gpu_cu_count": (_I, [_P]),
    "phr_gpu_arch": (C.c_char_p, [_P]),
    "phr_gpu_stream": (_P, [_P]),
    "phr_array_new": (_I, [_P, _U32, _U64, C.POINTER(_P)]),
    "phr_array_zeroed": (_I, [_P, _U32, _U64, C.POINTER(_P)]),
    "phr_array_from": (_I, [_P, _U32, _P, _U64, C.POINTER(_P)]),
    "phr_array_write": (_I, [_P, _P, _U64, C.POINTER(_I)]),
    "phr_array_write_at": (_I, [_P, _U64, _P, _U64, C.POINTER(_I)]),
    "phr_array_write_front": (_I, [_P, _P, _U64, C.POINTER(_I)]),
    "phr_array_split": (_I, [_P, _U32, _P, _U64, _U64, _U32, _P, C.POINTER(_U32)]),
    "phr_array_read_first": (_I, [_P, _U64, _P, C.POINTER(_U64)]),
    "phr_array_read_range": (_I, [_P, _U64, _U64, _P]),
    "phr_array_len": (_U64, [_P]),
    "phr_array_cap": (_U64, [_P]),
    "phr_array_elem": (_U32, [_P]),
    "phr_array_device_ptr": (_P, [_P]),
    "phr_array_free": (None, [_P]),
    "phr_array_copy_kernel": (_I, [_P, _P]),
    "phr_clock_new": (_I, [_P, C.POINTER(_P)]),
    "phr_clock_start": (_I, [_P]),
    "phr_clock_stop": (_I, [_P, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "phr_clock_free": (None, [_P]),
    "phr_kernel_known": (_I, [C.c_char_p, C.POINTER(_U32), C.POINTER(_U32)]),
    "phr_pass_new": (_I, [_P, C.POINTER(_P)]),
    "phr_pass_free": (None, [_P]),
    "phr_pass_add": (_I, [_P, C.c_char_p, _U32, _U32, _P, _U32, C.POINTER(_P), _U32, C.POINTER(_U32)]),
    "phr_pass_set": (_I, [_P, _U32, _U32, _P, _U32]),
    "phr_pass_run": (_I, [_P]),
    "phr_pass_run_step": (_I, [_P, _U32]),
    "phr_pass_clear": (None, [_P]),
    "phr_pass_len": (_U32, [_P]),
    "phr_pass_compiled": (_U32, [_P]),
    "phr_bvh_build": (_I, [_P, _P, _U32, _U32, C.POINTER(_P)]),
    "phr_bvh_build_on": (_I, [_P, _P, _U32, C.POINTER(_P)]),
    "phr_bvh_build_static": (_I, [_P, _P, _U32, _U32, _P, _P, _P]),
    "phr_bvh_rebuild_async": (_I, [_P]),
    "phr_bvh_set_sort_tile": (_I, [_P, _U32]),
    "phr_bvh_sort_tile": (_U32, [_P]),
    "phr_bvh_autotune": (_I, [_P, _U32, C.POINTER(_U32), _P]),
    "phr_bvh_rebuild": (_I, [_P]),
    "phr_bvh_refit_async": (_I, [_P]),
    "phr_bvh_refit": (_I, [_P]),
    "phr_bvh_update": (_I, [_P, _P, _U32, C.POINTER(_I)]),
    "phr_bvh_download": (_I, [_P, _P, _P, C.POINTER(_U32)]),
    "phr_bvh_codes": (_I, [_P, _P]),
    "phr_bvh_build_batch": (_I, [_P, _P, _U32, _P, _U32, _U32, C.POINTER(_P)]),
    "phr_bvh_build_batch_on": (_I, [_P, _P, _P, _U32, _U32, C.POINTER(_P)]),
    "phr_bvh_download_segment": (_I, [_P, _U32, _P, _P, C.POINTER(_U32)]),
    "phr_bvh_segments": (_U32, [_P]),
    "phr_bvh_segment_depth": (_U32, [_P, _U32]),
    "phr_bvh_len": (_U32, [_P]),
    "phr_bvh_depth": (_U32, [_P]),
    "phr_bvh_leaf": (_U32, [_P]),
    "phr_bvh_nodes": (_P, [_P]),
    "phr_bvh_order": (_P, [_P]),
    "phr_bvh_prims": (_P, [_P]),
    "phr_bvh_free": (None, [_P]),
    "phr_sort_record": (_I, [_P, _P, _U32, C.POINTER(_P)]),
    "phr_sort_run": (_I, [_P]),
    "phr_sort_keys": (_P, [_P]),
    "phr_sort_code": (_P, [_P]),
    "phr_sort_len": (_U32, [_P]),
    "phr_sort_steps": (_U32, [_P]),
    "phr_sort_bytes": (_U64, [_P]),
    "phr_sort_read": (_I, [_P, _P, _P]),
    "phr_sort_free": (None, [_P]),
    "phr_tracer_new": (_I, [_P, _P, _U32, _U32, C.POINTER(_P)]),
    "phr_tracer_free": (None, [_P]),
    "phr_tracer_len": (_U32, [_P]),
    "phr_tracer_depth": (_U32, [_P]),
    "phr_tracer_extent": (C.c_float, [_P]),
    "phr_tracer_glow": (_U32, [_P]),
    "phr_tracer_bvh": (_P, [_P]),
    "phr_tracer_order": (_I, [_P, _P]),
    "phr_tracer_lights": (_I, [_P, _P, _U32, C.POINTER(_U32)]),
    "phr_tracer_raycast": (_I, [_P, _P, _U32, _I, _P]),
    "phr_tracer_raycast_on": (_I, [_P, _P, _I, _P]),
    "phr_tracer_set_samples": (None, [_P, _U32]),
    "phr_tracer_sample": (_I, [_P, _P, _U32, _U32, _U32, C.POINTER(_U32), C.POINTER(_I)]),
    "phr_tracer_time": (_I, [_P, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "phr_tracer_rays": (_I, [_P, C.POINTER(_U64), C.POINTER(_U64)]),
    "phr_tracer_counters": (_I, [_P, _P]),
    "phr_tracer_px": (_I, [_P, _U64, _U64, _P]),
    "phr_tracer_image": (_P, [_P]),
    "phr_tracer_radiance": (_I, [_P, _P]),
    "phr_tracer_denoise": (_I, [_P, _I]),
    "phr_tracer_albedo": (_I, [_P, _P]),
    "phr_tracer_pixels": (_I, [_P, _P]),
    "phr_tracer_display": (_I, [_P, _P]),
    "phr_tracer_show": (_I, [_P, _P, _U64, _P]),
    "phr_tracer_samples": (_U32, [_P]),
    "phr_tracer_warm": (_I, [_P, _U32, _P, _P, C.POINTER(_U32)]),
    "phr_tracer_set_adaptive": (_I, [_P, _P]),
    "phr_tracer_size": (_I, [_P, _P, _P]),
    "phr_tracer_normal": (_I, [_P, _P]),
    "phr_tracer_depth_plane": (_I, [_P, _P]),
    "phr_tracer_from_scene": (_I, [_P, _P, _U32, C.POINTER(_P)]),
    "phr_tracer_with_chunks": (_I, [_P, _P, _U32, _U64, _U32, C.POINTER(_P)]),
    "phr_tracer_split": (_I, [_P, _P]),
    "phr_tracer_chunks": (_U32, [_P]),
    "phr_tracer_reinstance": (_I, [_P, _P, _U32, C.POINTER(_I)]),
    "phr_tracer_repose": (_I, [_P, _P, _U32, C.POINTER(_I)]),
    "phr_tracer_repose_on": (_I, [_P, _P, C.POINTER(_I)]),
    "phr_tracer_retree": (_I, [_P, C.POINTER(_I)]),
    "phr_tracer_nodes": (_P, [_P]),
    "phr_tracer_instances": (_P, [_P]),
    "phr_tracer_instance_count": (_U32, [_P]),
    "phr_kpcn_layers": (_U32, [_P]),
    "phr_kpcn_param_count": (_U64, []),
    "phr_kpcn_new": (_I, [_P, C.POINTER(_P)]),
    "phr_kpcn_free": (None, [_P]),
    "phr_kpcn_set_params": (_I, [_P, _P, _U64]),
    "phr_kpcn_get_params": (_I, [_P, _P, _U64]),
    "phr_kpcn_get_grads": (_I, [_P, _P, _U64]),
    "phr_kpcn_infer": (_I, [_P, _P, _U32, _U32, _U32, _P]),
    "phr_kpcn_loss": (_I, [_P, _P, _P, _U32, _U32, _U32, _P, C.POINTER(C.c_float)]),
    "phr_kpcn_backward": (_I, [_P, _P, _P, _U32, _U32, _U32, C.POINTER(C.c_float)]),
    "phr_kpcn_adam": (_I, [_P, C.c_float]),
    "phr_kpcn_train_step": (_I, [_P, _P, _P, _U32, _U32, _U32, C.c_float, C.POINTER(C.c_float)]),
    "phr_kpcn_autotune": (_I, [_P, _P, _U32, _U32, _U32, _U32, _P]),
    "phr_kpcn_held_out": (_I, [_P, _P, _P, _P, _U64, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "phr_kpcn_frame_mean": (_I, [_P, _P, _P, _U64, _U32, _P]),
    "phr_kpcn_spread": (_I, [_P, _P, _P, _U64, C.POINTER(C.c_double)]),
    "phr_kpcn_train_indexed": (_I, [_P, _P, _P, _P, _P, _U32, _U32, _U32, C.c_float, C.POINTER(C.c_float)]),
    "phr_tracer_frame": (_I, [_P, _P, _U64]),
    "phr_tracer_set_path_waves": (_I, [_P, _U32]),
    "phr_tracer_record_frame": (_I, [_P, _P, _U32, _U32, _U32, _I, _P]),
    "phr_tracer_replay_frame": (_I, [_P, _P]),
    "phr_tracer_record_pose_frame": (_I, [_P, _P, _P, _U32, _P, _U32, _U32, _U32, _I, _P]),
    "phr_tracer_replay_pose_frame": (_I, [_P, _P, _U32, _P]),
    "phr_world_new": (_I, [_P, C.POINTER(_P)]),
    "phr_world_free": (None, [_P]),
    "phr_world_configure": (_I, [_P, _U32, _U32, C.c_float, C.c_float, C.c_float, _P]),
    "phr_world_add_shape": (_I, [_P, _P, C.POINTER(_U32)]),
    "phr_world_add_body": (_I, [_P, _P, C.POINTER(_U32)]),
    "phr_world_step": (_I, [_P, C.c_float]),
    "phr_world_body_count": (_U32, [_P]),
    "phr_world_contact_count": (_U32, [_P]),
    "phr_world_bodies": (_I, [_P, _P]),
    "phr_world_poses": (_I, [_P, _P]),
    "phr_world_energy": (C.c_float, [_P]),
    "phr_world_pair_list": (_I, [_P, _P, _U32, C.POINTER(_U32)]),
    "phr_world_set_lanes": (_I, [_P, _U32]),
    "phr_world_lanes": (_U32, [_P]),
    "phr_bvh_compact": (_I, [_P, _P, _U64, C.POINTER(_U32)]),
    "phr_tracer_sample_in": (_I, [_P, _P, _U32, _U32, _U32, _U32, _P, _P, C.POINTER(_U32), C.POINTER(_I),
                                  C.POINTER(_I)]),
    "phr_tracer_open_pixels": (_I, [_P, C.POINTER(_U32)]),
    "phr_tracer_path_waves": (_U32, [_P]),
    "phr_tracer_set_path_regen": (_I, [_P, _U32]),
    "phr_tracer_path_regen": (_U32, [_P]),
    "phr_tracer_set_split_shade": (_I, [_P, _U32]),
    "phr_tracer_split_shade": (_U32, [_P]),
    "phr_tracer_path_stack": (_U32, [_P]),
    "phr_tracer_set_path_stack": (_I, [_P, _U32]),
    "phr_tracer_autotune": (_I, [_P, _P, _U32, _U32, _U32, _U32, C.POINTER(_U32), C.POINTER(C.c_double)]),
    "phr_tracer_denoise_learned": (_I, [_P, _P, _P]),
    "phr_batch_new": (_I, [_P, _P, _P, _U32, _U32, C.POINTER(_P)]),
    "phr_batch_free": (None, [_P]),
    "phr_batch_scenes": (_U32, [_P]),
    "phr_batch_len": (_U32, [_P]),
    "phr_batch_bvh": (_P, [_P]),
    "phr_batch_extent": (C.c_float, [_P, _U32]),
    "phr_batch_glow": (_U32, [_P, _U32]),
    "phr_batch_rebuild": (_I, [_P, C.POINTER(C.c_double)]),
    "phr_batch_render": (_I, [_P, _P, _U32, _U32, _U32, C.POINTER(C.c_double)]),
    "phr_batch_set_path": (_I, [_P, _U32, _U32]),
    "phr_batch_path": (_I, [_P, C.POINTER(_U32), C.POINTER(_U32)]),
    "phr_batch_set_path_build": (_I, [_P, _U32]),
    "phr_batch_path_build": (_U32, [_P]),
    "phr_batch_set_split_shade": (_I, [_P, _U32]),
    "phr_batch_split_shade": (_U32, [_P]),
    "phr_batch_autotune": (_I, [_P, _P, _U32, _U32, _U32, _U32, C.POINTER(_U32), C.POINTER(_U32), C.POINTER(C.c_double)]),
    "phr_batch_record": (_I, [_P, _P, _U32, _U32, _U32, _I]),
    "phr_batch_replay": (_I, [_P, C.POINTER(C.c_double)]),
    "phr_batch_px": (_I, [_P, _U32, _U64, _U64, _P]),
    "phr_batch_counters": (_I, [_P, C.POINTER(_U64)]),
    "phr_tri_from_phiray": (_I, [_P, _U64, _P]),
    "phr_camera_from_phiray": (_I, [_P, _P]),
}

_lib = None


class PhrError(RuntimeError):
    """A negative status from the C ABI, with the library's message."""


def lib() -> C.CDLL:
    """The loaded HIP library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(f"{LIB_PATH} is missing: build it with `python -m phiray_amd.build` (hipcc, gfx950)")
        h = C.CDLL(str(LIB_PATH))
        for name, (res, args) in _SIGS.items():
            f = getattr(h, name)
            f.restype = res
            f.argtypes = args
        _lib = h
    return _lib


def exported_symbols() -> list[str]:
    return list(_SIGS)


def _check(rc: int, what: str) -> None:
    if rc != 0:
        msg = lib().phr_last_error().decode(errors="replace")
        raise PhrError(f"{what} failed ({rc}): {msg}")


def _ptr(a: np.ndarray) -> C.c_void_p:
    return C.c_void_p(a.ctypes.data)


class Gpu:
    """phiray's `Gpu` (src/gpu/mod.rs:84-331): one device and its queue."""

    def __init__(self, handle: int):
        self._h = C.c_void_p(handle)

    @classmethod
    def new(cls, device: int = 0) -> "Gpu | None":
        """`Gpu::new`: None when there is no device."""
        out = C.c_void_p()
        rc = lib().phr_gpu_open(device, C.byref(out))
        if rc != 0:
            return None
        return cls(out.value)

    @classmethod
    def shared(cls) -> "Gpu | None":
        """`Gpu::shared`: the process's device, the same one for every caller."""
        out = C.c_void_p()
        if lib().phr_gpu_shared(C.byref(out)) != 0:
            return None
        return cls(out.value)

    def wait(self) -> None:
        _check(lib().phr_gpu_wait(self._h), "Gpu::wait")

    def wait_for(self, seconds: float) -> bool:
        done = C.c_int(0)
        _check(lib().phr_gpu_wait_for(self._h, int(seconds * 1e9), C.byref(done)), "Gpu::wait_for")
        return bool(done.value)

    def max_binding(self) -> int:
        return int(lib().phr_gpu_max_binding(self._h))

    def cu_count(self) -> int:
        return int(lib().phr_gpu_cu_count(self._h))

    def arch(self) -> str:
        return lib().phr_gpu_arch(self._h).decode()

    def close(self) -> None:
        if self._h:
            lib().phr_gpu_close(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class GpuArray:
    """phiray's `GpuArray<T>` (src/gpu/array.rs:7-188) over a numpy dtype."""

    def __init__(self, gpu: Gpu, handle: int, dtype: np.dtype):
        self.gpu = gpu
        self._h = C.c_void_p(handle)
        self.dtype = np.dtype(dtype)
        self._owned = True

    @classmethod
    def _borrow(cls, gpu: Gpu, handle: int, dtype) -> "GpuArray":
        a = cls(gpu, handle, dtype)
        a._owned = False
        return a

    @classmethod
    def zeroed(cls, gpu: Gpu, dtype, n: int) -> "GpuArray":
        out = C.c_void_p()
        _check(lib().phr_array_zeroed(gpu._h, np.dtype(dtype).itemsize, n, C.byref(out)), "GpuArray::zeroed")
        return cls(gpu, out.value, dtype)

    @classmethod
    def from_(cls, gpu: Gpu, items: np.ndarray) -> "GpuArray":
        items = np.ascontiguousarray(items)
        out = C.c_void_p()
        _check(lib().phr_array_from(gpu._h, items.dtype.itemsize, _ptr(items), len(items), C.byref(out)), "GpuArray::from")
        return cls(gpu, out.value, items.dtype)

    def write(self, items: np.ndarray) -> bool:
        items = np.ascontiguousarray(items, dtype=self.dtype)
        grown = C.c_int(0)
        _check(lib().phr_array_write(self._h, _ptr(items), len(items), C.byref(grown)), "GpuArray::write")
        return bool(grown.value)

    def write_at(self, at: int, items: np.ndarray) -> bool:
        """`GpuArray::write_at`: False, and nothing written, past the length."""
        items = np.ascontiguousarray(items, dtype=self.dtype)
        written = C.c_int(0)
        _check(lib().phr_array_write_at(self._h, at, _ptr(items), len(items), C.byref(written)), "GpuArray::write_at")
        return bool(written.value)

    def write_front(self, items: np.ndarray) -> bool:
        """`GpuArray::write_front`: False when the items do not fit."""
        items = np.ascontiguousarray(items, dtype=self.dtype)
        written = C.c_int(0)
        _check(lib().phr_array_write_front(self._h, _ptr(items), len(items), C.byref(written)), "GpuArray::write_front")
        return bool(written.value)

    def read_first(self, count: int) -> np.ndarray:
        """`GpuArray::read_first`: the first min(count, len) items."""
        out = np.empty(max(count, 0), dtype=self.dtype)
        got = C.c_uint64(0)
        _check(lib().phr_array_read_first(self._h, count, _ptr(out), C.byref(got)), "GpuArray::read_first")
        return out[: got.value]

    @classmethod
    def split(cls, gpu: "Gpu", items: np.ndarray, limit: int, chunks: int) -> "tuple[list[GpuArray], int] | None":
        """`GpuArray::split`: the items over `chunks` arrays of at most
        limit bytes each and the items a chunk; None when they do not fit."""
        items = np.ascontiguousarray(items)
        hs = (C.c_void_p * chunks)()
        cut = C.c_uint32(0)
        rc = lib().phr_array_split(gpu._h, items.dtype.itemsize, _ptr(items), len(items), limit, chunks, hs, C.byref(cut))
        if rc == -4:  # PHR_ERR_SIZE
            return None
        _check(rc, "GpuArray::split")
        return [cls(gpu, hs[k], items.dtype) for k in range(chunks)], int(cut.value)

    def __len__(self) -> int:
        return int(lib().phr_array_len(self._h))

    def read_range(self, start: int, count: int) -> np.ndarray:
        out = np.empty(count, dtype=self.dtype)
        _check(lib().phr_array_read_range(self._h, start, count, _ptr(out)), "GpuArray::read_range")
        return out

    def read(self) -> np.ndarray:
        return self.read_range(0, len(self))

    def free(self) -> None:
        if self._owned and self._h:
            lib().phr_array_free(self._h)
        self._h = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Clock:
    """phiray's `Clock` (src/gpu/clock.rs:17-87): HIP events on the device's stream."""

    def __init__(self, gpu: Gpu):
        self.gpu = gpu
        out = C.c_void_p()
        _check(lib().phr_clock_new(gpu._h, C.byref(out)), "Clock::new")
        self._h = out

    def start(self) -> None:
        _check(lib().phr_clock_start(self._h), "Clock start")

    def stop(self) -> tuple[float, float]:
        """(GPU seconds, wall seconds) since start, after the device drains."""
        g, w = C.c_double(0), C.c_double(0)
        _check(lib().phr_clock_stop(self._h, C.byref(g), C.byref(w)), "Clock stop")
        return g.value, w.value

    def __del__(self):
        try:
            if self._h:
                lib().phr_clock_free(self._h)
        except Exception:
            pass


def copy_kernel(dst: GpuArray, src: GpuArray) -> None:
    _check(lib().phr_array_copy_kernel(dst._h, src._h), "copy kernel")


class Pass:
    """phiray's `Pass` (src/gpu/pass.rs:18-155) over kernels named as phiray names them."""

    def __init__(self, gpu: Gpu):
        self.gpu = gpu
        out = C.c_void_p()
        _check(lib().phr_pass_new(gpu._h, C.byref(out)), "Pass::new")
        self._h = out
        self._keep: list = []

    @staticmethod
    def kernel(name: str) -> tuple[int, int]:
        """(Param bytes, bindings) of a registered kernel."""
        size, binds = C.c_uint32(0), C.c_uint32(0)
        _check(lib().phr_kernel_known(name.encode(), C.byref(size), C.byref(binds)), f"kernel {name}")
        return size.value, binds.value

    def add(self, name: str, n: int, params: bytes, arrays: list, items: int | None = None) -> int:
        buf = C.create_string_buffer(bytes(params), len(params))
        arr = (C.c_void_p * len(arrays))(*[a._h.value for a in arrays])
        step = C.c_uint32(0)
        _check(lib().phr_pass_add(self._h, name.encode(), n, n if items is None else items, buf, len(params), arr,
                                  len(arrays), C.byref(step)), f"Pass::add {name}")
        self._keep.append(arrays)
        return step.value

    def set(self, step: int, offset: int, data: bytes) -> None:
        buf = C.create_string_buffer(bytes(data), len(data))
        _check(lib().phr_pass_set(self._h, step, offset, buf, len(data)), "Pass::set")

    def run(self) -> None:
        _check(lib().phr_pass_run(self._h), "Pass::run")

    def run_step(self, step: int) -> None:
        _check(lib().phr_pass_run_step(self._h, step), "Pass::run_step")

    def clear(self) -> None:
        lib().phr_pass_clear(self._h)
        self._keep.clear()

    def __len__(self) -> int:
        return int(lib().phr_pass_len(self._h))

    def compiled(self) -> int:
        """`Pass::compiled`: kernel types this pass uses."""
        return int(lib().phr_pass_compiled(self._h))

    def __del__(self):
        try:
            if self._h:
                lib().phr_pass_free(self._h)
        except Exception:
            pass


class GpuSort:
    """phiray's `GpuSort` (src/bvh/sort.rs:13-111): recorded over a device array of codes, replayed."""

    def __init__(self, gpu: Gpu, codes: GpuArray, n: int | None = None):
        self.gpu, self._codes = gpu, codes
        out = C.c_void_p()
        _check(lib().phr_sort_record(gpu._h, codes._h, len(codes) if n is None else n, C.byref(out)), "GpuSort::record")
        self._h = out

    def run(self) -> None:
        _check(lib().phr_sort_run(self._h), "GpuSort::run")

    def __len__(self) -> int:
        return int(lib().phr_sort_len(self._h))

    def steps(self) -> int:
        return int(lib().phr_sort_steps(self._h))

    def bytes(self) -> int:
        return int(lib().phr_sort_bytes(self._h))

    def keys(self) -> np.ndarray:
        """(original slot, code) pairs, sorted."""
        return GpuArray._borrow(self.gpu, lib().phr_sort_keys(self._h), np.dtype((np.uint32, 2))).read()

    def read(self) -> tuple[np.ndarray, np.ndarray]:
        order, srt = np.zeros(len(self), np.uint32), np.zeros(len(self), np.uint32)
        _check(lib().phr_sort_read(self._h, _ptr(order), _ptr(srt)), "GpuSort::read")
        return order, srt

    def __del__(self):
        try:
            if self._h:
                lib().phr_sort_free(self._h)
        except Exception:
            pass


class Bvh:
    """A downloaded tree: phiray's `Bvh {nodes, order, depth}` (src/bvh/build.rs:17-24)."""

    def __init__(self, nodes: np.ndarray, order: np.ndarray, depth: int):
        self.nodes, self.order, self.depth = nodes, order, depth

    def reachable(self) -> list[int]:
        """src/bvh/shape.rs:39-51: indices from the root, depth first."""
        out, stack = [], ([0] if len(self.nodes) else [])
        li, ri = self.nodes["li"], self.nodes["ri"]
        while stack:
            i = stack.pop()
            out.append(i)
            if not (li[i] & LEAF):
                stack.extend([int(li[i]), int(ri[i])])
        return out


class GpuBvh:
    """phiray's `GpuBvh` (src/bvh/gpu.rs:11-104), built entirely on the device."""

    def __init__(self, gpu: Gpu, handle: int):
        self.gpu = gpu
        self._h = C.c_void_p(handle)

    @classmethod
    def build(cls, gpu: Gpu, prims: np.ndarray, leaf: int = 4) -> "GpuBvh":
        prims = np.ascontiguousarray(prims, dtype=AABB)
        out = C.c_void_p()
        _check(lib().phr_bvh_build(gpu._h, _ptr(prims), len(prims), leaf, C.byref(out)), "GpuBvh::build")
        return cls(gpu, out.value)

    @classmethod
    def build_on(cls, gpu: Gpu, prims: GpuArray, leaf: int = 4) -> "GpuBvh":
        out = C.c_void_p()
        _check(lib().phr_bvh_build_on(gpu._h, prims._h, leaf, C.byref(out)), "GpuBvh::build_on")
        b = cls(gpu, out.value)
        b._prims = prims  # keep the borrowed array alive
        return b

    @staticmethod
    def build_static(gpu: Gpu, prims: GpuArray, nodes: GpuArray, order: GpuArray, depth: "GpuArray | None" = None,
                     leaf: int = 4, n: int | None = None) -> None:
        """The build of `build_on` into caller-owned device arrays, with no
        allocation: scratch is the library's static device array (at most
        STATIC_BUILD_MAX boxes; calls on one GPU share it). Enqueued."""
        n = len(prims) if n is None else n
        cap = lambda a: int(lib().phr_array_cap(a._h))  # noqa: E731
        if prims.dtype.itemsize != 32 or nodes.dtype.itemsize != 32 or order.dtype.itemsize != 4 or (
                depth is not None and depth.dtype.itemsize != 4):
            raise ValueError("build_static: 32-byte boxes and nodes, u32 order and depth")
        if n and (n > cap(prims) or 2 * n - 1 > cap(nodes) or n > cap(order) or (depth is not None and cap(depth) < 1)):
            raise ValueError("build_static: an array is smaller than the tree")
        dptr = lib().phr_array_device_ptr
        _check(lib().phr_bvh_build_static(gpu._h, dptr(prims._h), n, leaf, dptr(nodes._h), dptr(order._h),
                                          dptr(depth._h) if depth is not None else None), "GpuBvh::build_static")

    @classmethod
    def build_batch(cls, gpu: Gpu, prims: np.ndarray, seg_lo, leaf: int = 4) -> "GpuBvh":
        """Many trees in one build: segment s is prims[seg_lo[s]:seg_lo[s+1]]."""
        prims = np.ascontiguousarray(prims, dtype=AABB)
        lo = np.ascontiguousarray(seg_lo, dtype=np.uint32)
        out = C.c_void_p()
        _check(lib().phr_bvh_build_batch(gpu._h, _ptr(prims), len(prims), _ptr(lo), len(lo) - 1, leaf, C.byref(out)),
               "GpuBvh::build_batch")
        b = cls(gpu, out.value)
        b._seg_lo = lo
        return b

    @classmethod
    def build_batch_on(cls, gpu: Gpu, prims: GpuArray, seg_lo, leaf: int = 4) -> "GpuBvh":
        lo = np.ascontiguousarray(seg_lo, dtype=np.uint32)
        out = C.c_void_p()
        _check(lib().phr_bvh_build_batch_on(gpu._h, prims._h, _ptr(lo), len(lo) - 1, leaf, C.byref(out)), "GpuBvh::build_batch_on")
        b = cls(gpu, out.value)
        b._prims, b._seg_lo = prims, lo
        return b

    @property
    def segments(self) -> int:
        return int(lib().phr_bvh_segments(self._h))

    def segment_depth(self, s: int) -> int:
        return int(lib().phr_bvh_segment_depth(self._h, s))

    def download_segment(self, s: int) -> Bvh:
        """Segment s's tree relative to itself (what GpuBvh::download gives for its boxes alone)."""
        lo = getattr(self, "_seg_lo", np.array([0, len(self)], dtype=np.uint32))
        m = int(lo[s + 1]) - int(lo[s])
        nodes = np.zeros(max(2 * m - 1, 0), dtype=NODE)
        order = np.zeros(m, dtype=np.uint32)
        depth = C.c_uint32(0)
        _check(lib().phr_bvh_download_segment(self._h, s, _ptr(nodes), _ptr(order), C.byref(depth)), "GpuBvh::download_segment")
        return Bvh(nodes, order, int(depth.value))

    def nodes_raw(self) -> np.ndarray:
        """The whole node array as the device holds it."""
        return GpuArray._borrow(self.gpu, lib().phr_bvh_nodes(self._h), NODE).read()

    @property
    def depth(self) -> int:
        return int(lib().phr_bvh_depth(self._h))

    @property
    def leaf(self) -> int:
        return int(lib().phr_bvh_leaf(self._h))

    def __len__(self) -> int:
        return int(lib().phr_bvh_len(self._h))

    def rebuild(self) -> None:
        _check(lib().phr_bvh_rebuild(self._h), "GpuBvh::rebuild")

    def rebuild_async(self) -> None:
        _check(lib().phr_bvh_rebuild_async(self._h), "GpuBvh::rebuild_async")

    @property
    def sort_tile(self) -> int:
        """Keys a thread holds in a sort pass (the tile is 256 x this)."""
        return int(lib().phr_bvh_sort_tile(self._h))

    def set_sort_tile(self, items: int) -> None:
        _check(lib().phr_bvh_set_sort_tile(self._h, items), "GpuBvh sort tile")

    def autotune(self, reps: int = 5) -> tuple[int, list[float]]:
        """Time rebuilds at each sort tile, keep the faster: (items, ms each)."""
        items = C.c_uint32(0)
        ms = np.zeros(2, dtype=np.float64)
        _check(lib().phr_bvh_autotune(self._h, reps, C.byref(items), _ptr(ms)), "GpuBvh autotune")
        return int(items.value), ms.tolist()

    def refit(self) -> None:
        _check(lib().phr_bvh_refit(self._h), "GpuBvh::refit")

    def refit_async(self) -> None:
        _check(lib().phr_bvh_refit_async(self._h), "GpuBvh::refit_async")

    def update(self, prims: np.ndarray) -> bool:
        """`GpuBvh::update`: False when the count changed (rebuild instead)."""
        prims = np.ascontiguousarray(prims, dtype=AABB)
        ok = C.c_int(0)
        _check(lib().phr_bvh_update(self._h, _ptr(prims), len(prims), C.byref(ok)), "GpuBvh::update")
        return bool(ok.value)

    def download(self) -> Bvh:
        n = len(self)
        nodes = np.zeros(max(2 * n - 1, 0), dtype=NODE)
        order = np.zeros(n, dtype=np.uint32)
        depth = C.c_uint32(0)
        _check(lib().phr_bvh_download(self._h, _ptr(nodes), _ptr(order), C.byref(depth)), "GpuBvh::download")
        return Bvh(nodes, order, int(depth.value))

    def compact(self) -> np.ndarray:
        """`Bvh::build_compact`'s nodes for the same primitives and leaf size:
        the reachable tree numbered breadth first, computed on the device."""
        n = len(self)
        nodes = np.zeros(max(2 * n - 1, 0), dtype=NODE)
        cnt = C.c_uint32(0)
        _check(lib().phr_bvh_compact(self._h, _ptr(nodes), len(nodes), C.byref(cnt)), "GpuBvh compact")
        return nodes[: cnt.value].copy()

    def sorted_codes(self) -> np.ndarray:
        out = np.zeros(len(self), dtype=np.uint32)
        _check(lib().phr_bvh_codes(self._h, _ptr(out)), "GpuBvh codes")
        return out

    def free(self) -> None:
        if self._h:
            lib().phr_bvh_free(self._h)
        self._h = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class _SceneDesc(C.Structure):
    _fields_ = [("verts", _P), ("nverts", _U32), ("faces", _P), ("nfaces", _U32), ("mats", _P), ("nmats", _U32),
                ("meshes", _P), ("nmeshes", _U32), ("insts", _P), ("ninsts", _U32)]


class Tracer:
    """phiray's `Tracer` (src/trace/mod.rs:50-245) for the Rendered path."""

    SAMPLE_COUNT = 32  # src/trace/mod.rs:129

    def __init__(self, gpu: Gpu, handle: int):
        self.gpu = gpu
        self._h = C.c_void_p(handle)
        self.size = (0, 0)

    @classmethod
    def new(cls, gpu: Gpu, tris: np.ndarray, leaf: int = 4) -> "Tracer | None":
        """`Tracer::new`: None for an empty scene."""
        tris = np.ascontiguousarray(tris, dtype=TRI)
        if len(tris) == 0:
            return None
        out = C.c_void_p()
        _check(lib().phr_tracer_new(gpu._h, _ptr(tris), len(tris), leaf, C.byref(out)), "Tracer::new")
        return cls(gpu, out.value)

    @classmethod
    def from_scene(cls, gpu: Gpu, verts: np.ndarray, faces: np.ndarray, mats: np.ndarray, meshes: np.ndarray,
                   insts: np.ndarray, leaf: int = 4) -> "Tracer | None":
        """`Tracer::from_scene` over meshes under instances (the forest); None when empty."""
        arrs = [np.ascontiguousarray(verts, dtype=VERT), np.ascontiguousarray(faces, dtype=FACE),
                np.ascontiguousarray(mats, dtype=MAT), np.ascontiguousarray(meshes, dtype=MESH),
                np.ascontiguousarray(insts, dtype=INSTANCE)]
        if len(arrs[1]) == 0 or len(arrs[4]) == 0:
            return None
        desc = _SceneDesc(*[x for a in arrs for x in (_ptr(a), len(a))])
        out = C.c_void_p()
        _check(lib().phr_tracer_from_scene(gpu._h, C.byref(desc), leaf, C.byref(out)), "Tracer::from_scene")
        t = cls(gpu, out.value)
        t._scene = arrs
        return t

    @classmethod
    def with_chunks(cls, gpu: Gpu, verts: np.ndarray, faces: np.ndarray, mats: np.ndarray, meshes: np.ndarray,
                    insts: np.ndarray, limit: int, chunks: int, leaf: int = 4) -> "Tracer | None":
        """`Tracer::with_chunks`: None when the scene does not fit `chunks`
        bindings of `limit` bytes."""
        arrs = [np.ascontiguousarray(verts, dtype=VERT), np.ascontiguousarray(faces, dtype=FACE),
                np.ascontiguousarray(mats, dtype=MAT), np.ascontiguousarray(meshes, dtype=MESH),
                np.ascontiguousarray(insts, dtype=INSTANCE)]
        if len(arrs[1]) == 0 or len(arrs[4]) == 0:
            return None
        desc = _SceneDesc(*[x for a in arrs for x in (_ptr(a), len(a))])
        out = C.c_void_p()
        rc = lib().phr_tracer_with_chunks(gpu._h, C.byref(desc), leaf, limit, chunks, C.byref(out))
        if rc == -4:  # PHR_ERR_SIZE
            return None
        _check(rc, "Tracer::with_chunks")
        t = cls(gpu, out.value)
        t._scene = arrs
        return t

    def split(self) -> tuple[int, int, int, int]:
        """`Tracer::split`: items a binding for nodes, faces, vertices; atlas offset."""
        out = (C.c_uint32 * 4)()
        _check(lib().phr_tracer_split(self._h, out), "Tracer::split")
        return tuple(int(v) for v in out)

    def chunks(self) -> int:
        """`Tracer::chunks`: bindings a large array (2 or 4)."""
        return int(lib().phr_tracer_chunks(self._h))

    def reinstance(self, worlds: np.ndarray) -> bool:
        """`Tracer::reinstance`: worlds [n, 16] as m[col][row]; False = build again."""
        w = np.ascontiguousarray(worlds, dtype=np.float32).reshape(-1, 16)
        ok = C.c_int(0)
        _check(lib().phr_tracer_reinstance(self._h, _ptr(w), len(w), C.byref(ok)), "Tracer::reinstance")
        return bool(ok.value)

    def repose(self, verts: np.ndarray) -> bool:
        """`Tracer::repose`: the same faces over moved vertices; False = build again."""
        v = np.ascontiguousarray(verts, dtype=VERT)
        ok = C.c_int(0)
        _check(lib().phr_tracer_repose(self._h, _ptr(v), len(v), C.byref(ok)), "Tracer::repose")
        return bool(ok.value)

    def repose_on(self, verts: GpuArray) -> bool:
        ok = C.c_int(0)
        _check(lib().phr_tracer_repose_on(self._h, verts._h, C.byref(ok)), "Tracer::repose_on")
        return bool(ok.value)

    def retree(self) -> bool:
        """`Tracer::retree`: every mesh tree built again over the pose held."""
        ok = C.c_int(0)
        _check(lib().phr_tracer_retree(self._h, C.byref(ok)), "Tracer::retree")
        return bool(ok.value)

    def nodes(self) -> np.ndarray:
        return GpuArray._borrow(self.gpu, lib().phr_tracer_nodes(self._h), NODE).read()

    def instances(self) -> np.ndarray:
        """Forest::inst_words as the kernels read them (records, slot order, draw order)."""
        h = lib().phr_tracer_instances(self._h)
        if not h:
            return np.zeros(0, dtype=np.uint32)
        return GpuArray._borrow(self.gpu, h, np.uint32).read()

    def __len__(self) -> int:
        return int(lib().phr_tracer_len(self._h))

    @property
    def depth(self) -> int:
        return int(lib().phr_tracer_depth(self._h))

    @property
    def extent(self) -> float:
        return float(lib().phr_tracer_extent(self._h))

    @property
    def glow(self) -> int:
        return int(lib().phr_tracer_glow(self._h))

    def bvh(self) -> Bvh:
        b = GpuBvh(self.gpu, lib().phr_tracer_bvh(self._h))
        try:
            return b.download()
        finally:
            b._h = C.c_void_p()  # borrowed: the tracer owns it

    def order(self) -> np.ndarray:
        out = np.zeros(len(self), dtype=np.uint32)
        _check(lib().phr_tracer_order(self._h, _ptr(out)), "Tracer order")
        return out

    def lights(self) -> np.ndarray:
        n = C.c_uint32(0)
        _check(lib().phr_tracer_lights(self._h, None, 0, C.byref(n)), "Tracer lights")
        out = np.zeros(n.value, dtype=np.uint32)
        _check(lib().phr_tracer_lights(self._h, _ptr(out), n.value, C.byref(n)), "Tracer lights")
        return out

    def raycast(self, rays: np.ndarray, any_hit: bool = False) -> np.ndarray:
        rays = np.ascontiguousarray(rays, dtype=RAY)
        hits = np.zeros(len(rays), dtype=HIT)
        _check(lib().phr_tracer_raycast(self._h, _ptr(rays), len(rays), 1 if any_hit else 0, _ptr(hits)), "Tracer raycast")
        return hits

    def raycast_on(self, rays: GpuArray, hits: GpuArray, any_hit: bool = False) -> None:
        _check(lib().phr_tracer_raycast_on(self._h, rays._h, 1 if any_hit else 0, hits._h), "Tracer raycast_on")

    def set_samples(self, n: int) -> None:
        lib().phr_tracer_set_samples(self._h, n)

    def sample(self, cam: np.ndarray, w: int, h: int, spp: int = 1) -> tuple[int, bool]:
        """`Tracer::sample`: (index of the last sample, done)."""
        cam = np.ascontiguousarray(np.asarray(cam, dtype=CAMERA).reshape(1))
        frame = C.c_uint32(0)
        done = C.c_int(0)
        _check(lib().phr_tracer_sample(self._h, _ptr(cam), w, h, spp, C.byref(frame), C.byref(done)), "Tracer::sample")
        self.size = (max(w, 1), max(h, 1))
        return int(frame.value), bool(done.value)

    def open_pixels(self) -> int:
        """Pixels Adaptive Sampling has not settled."""
        c = C.c_uint32(0)
        _check(lib().phr_tracer_open_pixels(self._h, C.byref(c)), "Tracer open pixels")
        return int(c.value)

    def set_path_waves(self, waves: int) -> None:
        _check(lib().phr_tracer_set_path_waves(self._h, waves), "Tracer path waves")

    @property
    def path_waves(self) -> int:
        return int(lib().phr_tracer_path_waves(self._h))

    PATH_NESTED, PATH_REGEN, PATH_WAVEFRONT, PATH_SPLIT = 0, 1, 2, 3

    def set_path_regen(self, mode: "bool | int") -> None:
        """Which build the path frames take: the nested one (False, 0), the
        path-regenerating one (True, 1), the wavefront one (2) or the split
        one (3: a wave walks until three quarters of its lanes have finished their walk,
        those shade and start their next walk while the others walk on). The
        last three serve one tree and a camera with lamps or a lit world, no
        Adaptive Sampling; other frames take the nested build."""
        m = int(mode) if not isinstance(mode, bool) else (1 if mode else 0)
        _check(lib().phr_tracer_set_path_regen(self._h, m), "Tracer set_path_regen")

    @property
    def path_regen(self) -> int:
        return int(lib().phr_tracer_path_regen(self._h))

    def set_split_shade(self, eighths: int) -> None:
        """The split build's shading point: a wave walks until `eighths`
        (1 to 8) eighths of its working lanes wait to shade. The sums do not
        depend on it; the autotune picks 4 or 6."""
        _check(lib().phr_tracer_set_split_shade(self._h, int(eighths)), "Tracer set_split_shade")

    @property
    def split_shade(self) -> int:
        return int(lib().phr_tracer_split_shade(self._h))

    @property
    def path_stack(self) -> int:
        """Mesh-tree stack entries a thread of the path launches holds."""
        return int(lib().phr_tracer_path_stack(self._h))

    def set_path_stack(self, entries: int) -> None:
        """24 (a forest whose mesh trees are at most 24 deep) or 32 entries."""
        _check(lib().phr_tracer_set_path_stack(self._h, entries), "Tracer path stack")

    def autotune(self, cam: np.ndarray, w: int, h: int, spp: int = 1, reps: int = 2) -> tuple[int, float]:
        """Time each path kernel variant on this workload and keep the fastest:
        (waves a SIMD chosen, its seconds). The image starts over."""
        cam = np.ascontiguousarray(np.asarray(cam, dtype=CAMERA).reshape(1))
        chosen, best = C.c_uint32(0), C.c_double(0)
        _check(lib().phr_tracer_autotune(self._h, _ptr(cam), w, h, spp, reps, C.byref(chosen), C.byref(best)),
               "Tracer autotune")
        self.size = (max(w, 1), max(h, 1))
        return int(chosen.value), float(best.value)

    _GO = C.CFUNCTYPE(C.c_int, C.c_void_p)

    def sample_in(self, cam: np.ndarray, w: int, h: int, spp: int, slices: int, go=None) -> tuple[int, bool] | None:
        """`Tracer::sample_in`: the launches in `slices` pieces, `go()` asked
        after each whether to carry on; None when it said no (the next call
        starts the image over)."""
        cam = np.ascontiguousarray(np.asarray(cam, dtype=CAMERA).reshape(1))
        frame, done, fin = C.c_uint32(0), C.c_int(0), C.c_int(0)
        cb = self._GO(lambda _u: 1 if go() else 0) if go is not None else None
        _check(lib().phr_tracer_sample_in(self._h, _ptr(cam), w, h, spp, slices,
                                          C.cast(cb, C.c_void_p) if cb is not None else None, None,
                                          C.byref(frame), C.byref(done), C.byref(fin)), "Tracer::sample_in")
        self.size = (max(w, 1), max(h, 1))
        if not fin.value:
            return None
        return int(frame.value), bool(done.value)

    def time(self) -> tuple[float, float]:
        """`Tracer::time`: (submit seconds, GPU seconds) of the last sample call."""
        a, b = C.c_double(0), C.c_double(0)
        _check(lib().phr_tracer_time(self._h, C.byref(a), C.byref(b)), "Tracer::time")
        return a.value, b.value

    def rays(self) -> tuple[int, int]:
        a, b = C.c_uint64(0), C.c_uint64(0)
        _check(lib().phr_tracer_rays(self._h, C.byref(a), C.byref(b)), "Tracer rays")
        return int(a.value), int(b.value)

    def counters(self) -> np.ndarray:
        """[closest, shadow, steps, tests] of the path kernel, [steps, tests] of raycasts."""
        out = np.zeros(6, dtype=np.uint64)
        _check(lib().phr_tracer_counters(self._h, _ptr(out)), "Tracer counters")
        return out

    def px(self) -> np.ndarray:
        w, h = self.size
        n = PX_LEN * w * h + 1
        out = np.zeros(n, dtype=np.uint32)
        _check(lib().phr_tracer_px(self._h, 0, n, _ptr(out)), "Tracer px")
        return out

    def denoise(self, keep: bool = False) -> None:
        """`Tracer::denoise` (keep=False) / `denoise_shown` (keep=True)."""
        _check(lib().phr_tracer_denoise(self._h, 1 if keep else 0), "Tracer::denoise")

    def denoise_learned(self, kpcn: "Kpcn", out: GpuArray | None = None) -> None:
        """The learned denoiser over what has been traced: into `out` (w*h*3
        floats) when given, else in place of the linear sums, as
        `denoise(keep=False)` leaves them."""
        _check(lib().phr_tracer_denoise_learned(self._h, kpcn._h, out._h if out is not None else None),
               "Tracer::denoise_learned")

    def record_frame(self, cam: np.ndarray, w: int, h: int, spp: int, denoise: int = 0, kpcn: "Kpcn | None" = None) -> None:
        """Capture one frame (fresh image, spp samples, then denoise 0 none /
        1 a-trous / 2 learned) as a hipGraph."""
        cam = np.ascontiguousarray(np.asarray(cam, dtype=CAMERA).reshape(1))
        _check(lib().phr_tracer_record_frame(self._h, _ptr(cam), w, h, spp, denoise, kpcn._h if kpcn else None),
               "Tracer record frame")
        self.size = (max(w, 1), max(h, 1))

    def replay_frame(self, kpcn: "Kpcn | None" = None) -> None:
        _check(lib().phr_tracer_replay_frame(self._h, kpcn._h if kpcn else None), "Tracer replay frame")

    def record_pose_frame(self, verts: "GpuArray | None", worlds: np.ndarray, cam: np.ndarray, w: int, h: int, spp: int,
                          denoise: int = 0, kpcn: "Kpcn | None" = None) -> None:
        """Capture a forest's whole frame as one hipGraph: the new pose
        (vertices from `verts` when given, then the matrices), the camera's
        stamp, the path launch and the denoiser."""
        worlds = np.ascontiguousarray(worlds, dtype=np.float32)
        cam = np.ascontiguousarray(np.asarray(cam, dtype=CAMERA).reshape(1))
        _check(lib().phr_tracer_record_pose_frame(self._h, verts._h if verts is not None else None, _ptr(worlds),
                                                   len(worlds.reshape(-1, 16)), _ptr(cam), w, h, spp, denoise,
                                                   kpcn._h if kpcn else None), "Tracer record pose frame")
        self.size = (max(w, 1), max(h, 1))

    def replay_pose_frame(self, worlds: np.ndarray, kpcn: "Kpcn | None" = None) -> None:
        """The recorded posed frame under new matrices (the vertices are read
        again from the array given at recording)."""
        worlds = np.ascontiguousarray(worlds, dtype=np.float32)
        _check(lib().phr_tracer_replay_pose_frame(self._h, _ptr(worlds), len(worlds.reshape(-1, 16)),
                                                  kpcn._h if kpcn else None), "Tracer replay pose frame")

    def frame_into(self, out: GpuArray, at: int = 0) -> None:
        """The denoiser's input frame (10 floats a pixel) at float `at` of `out`."""
        _check(lib().phr_tracer_frame(self._h, out._h, at), "Tracer frame")

    def radiance(self) -> np.ndarray:
        w, h = self.size
        out = np.zeros((w * h, 3), dtype=np.float32)
        _check(lib().phr_tracer_radiance(self._h, _ptr(out)), "Tracer::radiance")
        return out

    def pixels(self) -> np.ndarray:
        """`Tracer::pixels`: the RGBA8 image, (h, w, 4) bytes, row 0 at the top."""
        w, h = self.size
        out = np.zeros((h, w, 4), dtype=np.uint8)
        _check(lib().phr_tracer_pixels(self._h, _ptr(out)), "Tracer::pixels")
        return out

    def display(self) -> np.ndarray:
        """`Tracer::display`: the shown value before quantising, (w*h, 3) floats."""
        w, h = self.size
        out = np.zeros((w * h, 3), dtype=np.float32)
        _check(lib().phr_tracer_display(self._h, _ptr(out)), "Tracer::display")
        return out

    def show(self, linear: np.ndarray) -> bool:
        """`Tracer::show`: a scene-linear image (rgb a pixel, row 0 first) in
        place of the traced mean, then the display step; False with no image."""
        lin = np.ascontiguousarray(np.asarray(linear, dtype=np.float32).reshape(-1, 3))
        shown = C.c_int(0)
        _check(lib().phr_tracer_show(self._h, _ptr(lin), len(lin), C.byref(shown)), "Tracer::show")
        return bool(shown.value)

    @staticmethod
    def warm(gpu: Gpu) -> list[tuple[str, float]]:
        """`Tracer::warm`: each path-kernel build loaded, with its load time in ms."""
        cap = 64
        names = (C.c_char_p * cap)()
        ms = (C.c_double * cap)()
        count = C.c_uint32(0)
        _check(lib().phr_tracer_warm(gpu._h, cap, names, ms, C.byref(count)), "Tracer::warm")
        return [(names[i].decode(), float(ms[i])) for i in range(count.value)]

    def samples(self) -> int:
        """`Tracer::samples`: samples averaged before a still view stops refining."""
        return int(lib().phr_tracer_samples(self._h))

    def set_adaptive(self, adapt) -> None:
        """`Tracer::set_adaptive`: [threshold, min samples, step, on] on every later camera."""
        a = np.ascontiguousarray(adapt, dtype=np.float32).reshape(4)
        _check(lib().phr_tracer_set_adaptive(self._h, _ptr(a)), "Tracer::set_adaptive")

    def image_size(self) -> tuple[int, int]:
        """`Tracer::size`: the image's (w, h), (0, 0) before the first sample."""
        w, h = C.c_uint32(0), C.c_uint32(0)
        _check(lib().phr_tracer_size(self._h, C.byref(w), C.byref(h)), "Tracer::size")
        return int(w.value), int(h.value)

    def render(self, cam: np.ndarray, w: int, h: int) -> tuple[np.ndarray, bool]:
        """`Tracer::render`: one sample, then the RGBA8 image and whether
        refinement is done."""
        _, done = self.sample(cam, w, h, 1)
        return self.pixels(), done

    def albedo(self) -> np.ndarray:
        w, h = self.size
        out = np.zeros((w * h, 3), dtype=np.float32)
        _check(lib().phr_tracer_albedo(self._h, _ptr(out)), "Tracer::albedo")
        return out

    def normal(self) -> np.ndarray:
        w, h = self.size
        out = np.zeros((w * h, 3), dtype=np.float32)
        _check(lib().phr_tracer_normal(self._h, _ptr(out)), "Tracer::normal")
        return out

    def depth_plane(self) -> np.ndarray:
        w, h = self.size
        out = np.zeros(w * h, dtype=np.float32)
        _check(lib().phr_tracer_depth_plane(self._h, _ptr(out)), "Tracer::depth")
        return out

    def free(self) -> None:
        if self._h:
            lib().phr_tracer_free(self._h)
        self._h = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Batch:
    """Many scenes in one segmented LBVH build and one path launch (no
    counterpart in phiray, which holds one Scene a Tracer): scene k is
    tris[seg_lo[k]:seg_lo[k+1]] and renders as `Tracer.new` over those
    triangles alone does."""

    def __init__(self, gpu: Gpu, tris: np.ndarray, seg_lo, leaf: int = 4):
        tris = np.ascontiguousarray(tris, dtype=TRI)
        lo = np.ascontiguousarray(seg_lo, dtype=np.uint32)
        self.gpu = gpu
        self.seg_lo = lo
        self.size = (0, 0)
        out = C.c_void_p()
        self._h = C.c_void_p()
        _check(lib().phr_batch_new(gpu._h, _ptr(tris), _ptr(lo), len(lo) - 1, leaf, C.byref(out)), "Batch::new")
        self._h = out

    @property
    def scenes(self) -> int:
        return int(lib().phr_batch_scenes(self._h))

    def __len__(self) -> int:
        return int(lib().phr_batch_len(self._h))

    def extent(self, k: int) -> float:
        return float(lib().phr_batch_extent(self._h, k))

    def glow(self, k: int) -> int:
        return int(lib().phr_batch_glow(self._h, k))

    def tree(self, k: int) -> Bvh:
        """Scene k's tree relative to itself (what Bvh::build of its boxes gives)."""
        b = GpuBvh(self.gpu, lib().phr_batch_bvh(self._h))
        b._seg_lo = self.seg_lo
        try:
            return b.download_segment(k)
        finally:
            b._h = C.c_void_p()  # borrowed: the batch owns it

    def rebuild(self) -> float:
        """Boxes, the segmented build and the slot order again; device seconds."""
        t = C.c_double(0)
        _check(lib().phr_batch_rebuild(self._h, C.byref(t)), "Batch rebuild")
        return float(t.value)

    def render(self, cam: np.ndarray, w: int, h: int, spp: int) -> float:
        """Every scene, spp samples of a fresh image, one launch; device seconds."""
        cam = np.ascontiguousarray(np.asarray(cam, dtype=CAMERA).reshape(1))
        t = C.c_double(0)
        _check(lib().phr_batch_render(self._h, _ptr(cam), w, h, spp, C.byref(t)), "Batch render")
        self.size = (max(w, 1), max(h, 1))
        return float(t.value)

    def set_path(self, waves: int, grouped: int) -> None:
        _check(lib().phr_batch_set_path(self._h, waves, grouped), "Batch set_path")

    def path(self) -> tuple[int, int]:
        wv, g = C.c_uint32(0), C.c_uint32(0)
        _check(lib().phr_batch_path(self._h, C.byref(wv), C.byref(g)), "Batch path")
        return int(wv.value), int(g.value)

    def set_path_build(self, build: int) -> None:
        """The render kernel's samples: Tracer.PATH_NESTED (0) or
        Tracer.PATH_SPLIT (3; taken where the camera has lamps or a lit world
        and no Adaptive Sampling, else the nested build runs)."""
        _check(lib().phr_batch_set_path_build(self._h, int(build)), "Batch set_path_build")

    @property
    def path_build(self) -> int:
        return int(lib().phr_batch_path_build(self._h))

    def set_split_shade(self, eighths: int) -> None:
        """The split build's shading point, as Tracer.set_split_shade."""
        _check(lib().phr_batch_set_split_shade(self._h, int(eighths)), "Batch set_split_shade")

    @property
    def split_shade(self) -> int:
        return int(lib().phr_batch_split_shade(self._h))

    def autotune(self, cam: np.ndarray, w: int, h: int, spp: int = 1, reps: int = 1) -> tuple[int, int, float]:
        """(waves a SIMD, grouped order, seconds) of the fastest render kernel,
        over the nested and split builds where the camera allows the split."""
        cam = np.ascontiguousarray(np.asarray(cam, dtype=CAMERA).reshape(1))
        wv, g, t = C.c_uint32(0), C.c_uint32(0), C.c_double(0)
        _check(lib().phr_batch_autotune(self._h, _ptr(cam), w, h, spp, reps, C.byref(wv), C.byref(g), C.byref(t)),
               "Batch autotune")
        self.size = (max(w, 1), max(h, 1))
        return int(wv.value), int(g.value), float(t.value)

    def record(self, cam: np.ndarray, w: int, h: int, spp: int, rebuild: bool = True) -> None:
        cam = np.ascontiguousarray(np.asarray(cam, dtype=CAMERA).reshape(1))
        _check(lib().phr_batch_record(self._h, _ptr(cam), w, h, spp, 1 if rebuild else 0), "Batch record")
        self.size = (max(w, 1), max(h, 1))

    def replay(self) -> float:
        t = C.c_double(0)
        _check(lib().phr_batch_replay(self._h, C.byref(t)), "Batch replay")
        return float(t.value)

    def px(self, k: int) -> np.ndarray:
        """Scene k's whole image buffer (PX_LEN * w * h + 1 words)."""
        w, h = self.size
        out = np.zeros(PX_LEN * w * h + 1, dtype=np.uint32)
        _check(lib().phr_batch_px(self._h, k, 0, len(out), _ptr(out)), "Batch px")
        return out

    def counters(self) -> np.ndarray:
        out = (_U64 * 4)()
        _check(lib().phr_batch_counters(self._h, out), "Batch counters")
        return np.array(list(out), dtype=np.uint64)

    def free(self) -> None:
        if self._h:
            lib().phr_batch_free(self._h)
        self._h = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Kpcn:
    """The learned denoiser (a kernel-predicting CNN beside `kernel::Denoise`),
    run and trained by the library's own HIP kernels. Frames are GpuArrays of
    float32, 10 per pixel; radiance arrays 3 per pixel; batches image after
    image."""

    FRAME = 10

    def __init__(self, gpu: Gpu, params: np.ndarray | None = None):
        self.gpu = gpu
        out = C.c_void_p()
        _check(lib().phr_kpcn_new(gpu._h, C.byref(out)), "Kpcn::new")
        self._h = out
        if params is not None:
            self.set_params(params)

    @staticmethod
    def layers() -> np.ndarray:
        """{cin, cout, level, relu} per layer."""
        n = lib().phr_kpcn_layers(None)
        out = np.zeros((n, 4), dtype=np.uint32)
        lib().phr_kpcn_layers(_ptr(out))
        return out

    @staticmethod
    def param_count() -> int:
        return int(lib().phr_kpcn_param_count())

    def set_params(self, p: np.ndarray) -> None:
        p = np.ascontiguousarray(p, dtype=np.float32)
        _check(lib().phr_kpcn_set_params(self._h, _ptr(p), p.size), "Kpcn::set_params")

    def params(self) -> np.ndarray:
        out = np.zeros(self.param_count(), dtype=np.float32)
        _check(lib().phr_kpcn_get_params(self._h, _ptr(out), out.size), "Kpcn::params")
        return out

    def grads(self) -> np.ndarray:
        out = np.zeros(self.param_count(), dtype=np.float32)
        _check(lib().phr_kpcn_get_grads(self._h, _ptr(out), out.size), "Kpcn::grads")
        return out

    @staticmethod
    def held_out(gpu: Gpu, denoised: GpuArray, half_a: GpuArray, half_b: GpuArray, n: int) -> tuple[float, float]:
        """(estimated relMSE of `denoised`, made from half A alone, against the
        converged image; half B's own noise term), from two independent
        halves of a frame."""
        err, noise = C.c_double(0), C.c_double(0)
        _check(lib().phr_kpcn_held_out(gpu._h, denoised._h, half_a._h, half_b._h, n, C.byref(err), C.byref(noise)),
               "Kpcn held-out check")
        return float(err.value), float(noise.value)

    @staticmethod
    def spread(gpu: Gpu, a: GpuArray, b: GpuArray, n: int) -> float:
        """mean((a - b)^2 / (a^2 + 0.01)) of two denoised images from
        independent halves: twice the noise-driven part of either one's relMSE."""
        v = C.c_double(0)
        _check(lib().phr_kpcn_spread(gpu._h, a._h, b._h, n, C.byref(v)), "Kpcn spread")
        return float(v.value)

    @staticmethod
    def frame_mean(gpu: Gpu, half_a: GpuArray, half_b: GpuArray, n: int, out: GpuArray, channels: int = 10) -> None:
        """Both halves' frame as one (equal sample counts): all 10 channels,
        or the radiance alone (channels=3)."""
        _check(lib().phr_kpcn_frame_mean(gpu._h, half_a._h, half_b._h, n, channels, out._h), "Kpcn frame mean")

    def infer(self, frames: GpuArray, b: int, w: int, h: int, out: GpuArray) -> None:
        _check(lib().phr_kpcn_infer(self._h, frames._h, b, w, h, out._h), "Kpcn::infer")

    def loss(self, frames: GpuArray, targets: GpuArray, b: int, w: int, h: int, out: GpuArray | None = None) -> float:
        v = C.c_float(0.0)
        _check(lib().phr_kpcn_loss(self._h, frames._h, targets._h, b, w, h, out._h if out is not None else None,
                                   C.byref(v)), "Kpcn::loss")
        return float(v.value)

    def backward(self, frames: GpuArray, targets: GpuArray, b: int, w: int, h: int) -> float:
        v = C.c_float(0.0)
        _check(lib().phr_kpcn_backward(self._h, frames._h, targets._h, b, w, h, C.byref(v)), "Kpcn::backward")
        return float(v.value)

    def adam(self, lr: float) -> None:
        _check(lib().phr_kpcn_adam(self._h, lr), "Kpcn::adam")

    def train_step(self, frames: GpuArray, targets: GpuArray, b: int, w: int, h: int, lr: float,
                   want_loss: bool = True) -> float | None:
        v = C.c_float(0.0)
        _check(lib().phr_kpcn_train_step(self._h, frames._h, targets._h, b, w, h, lr,
                                         C.byref(v) if want_loss else None), "Kpcn::train_step")
        return float(v.value) if want_loss else None

    def autotune(self, frames: GpuArray, b: int, w: int, h: int, reps: int = 3) -> np.ndarray:
        """Time each layer's forward conv tiles on this frame size and keep
        the fastest: rows a wave owns, per layer."""
        out = np.zeros(len(self.layers()), dtype=np.uint8)
        _check(lib().phr_kpcn_autotune(self._h, frames._h, b, w, h, reps, _ptr(out)), "Kpcn autotune")
        return out

    def train_indexed(self, frames: GpuArray, fidx, targets: GpuArray, tidx, w: int, h: int, lr: float,
                      want_loss: bool = True) -> float | None:
        """One step over images fidx of `frames` against images tidx of `targets`."""
        fi = np.ascontiguousarray(fidx, dtype=np.uint32)
        ti = np.ascontiguousarray(tidx, dtype=np.uint32)
        if fi.size != ti.size:
            raise ValueError("Kpcn::train_indexed: fidx and tidx differ in length")
        v = C.c_float(0.0)
        _check(lib().phr_kpcn_train_indexed(self._h, frames._h, _ptr(fi), targets._h, _ptr(ti), fi.size, w, h, lr,
                                            C.byref(v) if want_loss else None), "Kpcn::train_indexed")
        return float(v.value) if want_loss else None

    def free(self) -> None:
        if self._h:
            lib().phr_kpcn_free(self._h)
        self._h = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


SHAPE = np.dtype([("kind", "<u4"), ("data", "<u4"), ("count", "<u4"), ("radius", "<f4"), ("half", "<f4", 3), ("_a", "<f4")])
BODY = np.dtype([("x", "<f4", 3), ("inv_mass", "<f4"), ("q", "<f4", 4), ("v", "<f4", 3), ("shape", "<u4"),
                 ("w", "<f4", 3), ("flags", "<u4")])
BODY_DEF = np.dtype([("kind", "<u4"), ("shape", "<u4"), ("mass", "<f4"), ("x", "<f4", 3), ("q", "<f4", 4), ("v", "<f4", 3),
                     ("w", "<f4", 3), ("scale", "<f4", 3), ("com", "<f4", 3), ("friction", "<f4"), ("restitution", "<f4"),
                     ("margin", "<f4"), ("lin_damp", "<f4"), ("ang_damp", "<f4"), ("sleep", "<f4", 2), ("has_sleep", "<u4"),
                     ("asleep", "<u4"), ("groups", "<u4"), ("obj", "<u4")])
assert SHAPE.itemsize == 32 and BODY.itemsize == 64 and BODY_DEF.itemsize == 132
ACTIVE, PASSIVE, KINEMATIC = 0, 1, 2


class World:
    """phiray's rigid-body `World` (src/sim/world.rs): XPBD substeps on the
    host, the broadphase on the device. Shapes: box, sphere, capsule."""

    BOX, SPHERE, CAPSULE = 0, 1, 2

    def __init__(self, gpu: Gpu):
        self.gpu = gpu
        out = C.c_void_p()
        _check(lib().phr_world_new(gpu._h, C.byref(out)), "World::new")
        self._h = out

    def configure(self, substeps: int = 10, iterations: int = 10, time_scale: float = 1.0, max_recover: float = 2.0,
                  sleep_after: float = 2.0, gravity=(0.0, 0.0, -9.81)) -> None:
        g = np.array(gravity, dtype=np.float32)
        _check(lib().phr_world_configure(self._h, substeps, iterations, time_scale, max_recover, sleep_after, _ptr(g)),
               "World configure")

    def add_shape(self, kind: int, half=(0.0, 0.0, 0.0), radius: float = 0.0) -> int:
        s = np.zeros(1, dtype=SHAPE)
        s["kind"], s["half"], s["radius"] = kind, half, radius
        out = C.c_uint32(0)
        _check(lib().phr_world_add_shape(self._h, _ptr(s), C.byref(out)), "World::add_shape")
        return int(out.value)

    def add_box(self, half) -> int:
        return self.add_shape(self.BOX, half=half)

    def add_sphere(self, radius: float) -> int:
        return self.add_shape(self.SPHERE, radius=radius)

    def add_capsule(self, radius: float, half: float) -> int:
        return self.add_shape(self.CAPSULE, half=(radius, radius, half), radius=radius)

    @staticmethod
    def body_def(**kw) -> np.ndarray:
        """A BodyDef with phiray's defaults (build.rs:39-61), fields from kw."""
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
        _check(lib().phr_world_add_body(self._h, _ptr(d), C.byref(out)), "World::add_body")
        return int(out.value)

    def step(self, dt: float) -> None:
        _check(lib().phr_world_step(self._h, dt), "World::step")

    def __len__(self) -> int:
        return int(lib().phr_world_body_count(self._h))

    def bodies(self) -> np.ndarray:
        out = np.zeros(len(self), dtype=BODY)
        _check(lib().phr_world_bodies(self._h, _ptr(out)), "World bodies")
        return out

    def poses(self) -> np.ndarray:
        out = np.zeros((len(self), 16), dtype=np.float32)
        _check(lib().phr_world_poses(self._h, _ptr(out)), "World::poses")
        return out

    def energy(self) -> float:
        return float(lib().phr_world_energy(self._h))

    def contact_count(self) -> int:
        return int(lib().phr_world_contact_count(self._h))

    def pairs(self) -> np.ndarray:
        n = C.c_uint32(0)
        _check(lib().phr_world_pair_list(self._h, None, 0, C.byref(n)), "World pairs")
        out = np.zeros((n.value, 2), dtype=np.uint32)
        _check(lib().phr_world_pair_list(self._h, _ptr(out), n.value, C.byref(n)), "World pairs")
        return out

    @property
    def lanes(self) -> int:
        """Lanes a pair in the solve's projections (4, or 1)."""
        return int(lib().phr_world_lanes(self._h))

    def set_lanes(self, lanes: int) -> None:
        _check(lib().phr_world_set_lanes(self._h, lanes), "World set_lanes")

    def free(self) -> None:
        if self._h:
            lib().phr_world_free(self._h)
        self._h = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def planes(px: np.ndarray, w: int, h: int) -> dict[str, np.ndarray]:
    """The image buffer's planes (src/trace/read.rs:37-56) as arrays."""
    n = w * h
    f = px.view(np.float32)
    return {
        "pixel": px[:n].view(np.uint8).reshape(n, 4),
        "radiance": f[n:4 * n].reshape(n, 3),
        "display": f[4 * n:7 * n].reshape(n, 3),
        "albedo": f[7 * n:10 * n].reshape(n, 3),
        "normal": f[10 * n:13 * n].reshape(n, 3),
        "depth": f[13 * n:14 * n],
        "grid": f[20 * n:23 * n].reshape(n, 3),
        "cov": f[23 * n:24 * n],
        "film": f[24 * n:25 * n],
        "outline": px[25 * n:26 * n],
        "count": px[31 * n:32 * n],
        "missed": f[33 * n:34 * n],
    }

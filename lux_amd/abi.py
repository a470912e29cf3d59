"""ctypes binding of include/lux_hip.h (the C ABI of liblux_hip.so)."""

import ctypes
from ctypes import POINTER, c_char_p, c_double, c_int, c_int32, c_int64, c_size_t, c_uint8, c_uint32, c_uint64, c_void_p

import numpy as np

from . import LIB_PATH

LUX_OK = 0
STATUS_NAMES = {
    0: "LUX_OK",
    1: "LUX_ERR_INVALID_ARGUMENT",
    2: "LUX_ERR_OVERFLOW",
    3: "LUX_ERR_MISMATCH",
    4: "LUX_ERR_HIP",
    5: "LUX_ERR_NOT_INITIALIZED",
    6: "LUX_ERR_OUT_OF_MEMORY",
    7: "LUX_ERR_DEVICE_CHECK",
}
LUX_ERR_INVALID_ARGUMENT = 1
LUX_ERR_OVERFLOW = 2
LUX_ERR_MISMATCH = 3
LUX_ERR_DEVICE_CHECK = 7

# numpy views of the POD records (byte layouts as in lux_hip.h)
AABB_DTYPE = np.dtype([("p_min", "<f4", 3), ("p_max", "<f4", 3)])
RAY_DTYPE = np.dtype([("origin", "<f4", 3), ("direction", "<f4", 3), ("t_min", "<f4"),
                      ("t_max", "<f4"), ("time", "<f4")])
TRI_DTYPE = np.dtype([("v0", "<f4", 3), ("v1", "<f4", 3), ("v2", "<f4", 3)])
HIT_DTYPE = np.dtype([("primitive_id", "<i4"), ("t", "<f4")])
PAIR_DTYPE = np.dtype([("a", "<i4"), ("b", "<i4")])
NODE_DTYPE = np.dtype([("p_min", "<f4", 3), ("p_max", "<f4", 3), ("primitive_id", "<i4"),
                       ("escape", "<i4")])
assert AABB_DTYPE.itemsize == 24 and RAY_DTYPE.itemsize == 36 and TRI_DTYPE.itemsize == 36
assert HIT_DTYPE.itemsize == 8 and NODE_DTYPE.itemsize == 32


class LuxError(RuntimeError):
    """A non-zero status from the C ABI, with lux_hip_last_error() text."""

    def __init__(self, status, func, message):
        self.status = status
        self.func = func
        super().__init__(f"{func}: {STATUS_NAMES.get(status, status)}: {message}")


class Vec3(ctypes.Structure):
    _fields_ = [("x", ctypes.c_float), ("y", ctypes.c_float), ("z", ctypes.c_float)]


class Aabb(ctypes.Structure):
    _fields_ = [("p_min", Vec3), ("p_max", Vec3)]


class BvhView(ctypes.Structure):
    _fields_ = [("nodes", c_void_p), ("node_count", c_int32), ("root_bounds", Aabb)]


_P = c_void_p
# name -> argtypes (every function returns int unless listed in _RESTYPES)
_SIGNATURES = {
    "lux_hip_init": [c_int],
    "lux_hip_shutdown": [],
    "lux_hip_last_error": [],
    "lux_hip_version": [],
    "lux_hip_device_info": [_P, c_int32, _P, c_int32, POINTER(c_int32), POINTER(c_uint64)],
    "lux_hip_malloc": [POINTER(c_void_p), c_size_t],
    "lux_hip_free": [_P],
    "lux_hip_memcpy_htod": [_P, _P, c_size_t],
    "lux_hip_memcpy_dtoh": [_P, _P, c_size_t],
    "lux_hip_synchronize": [],
    "lux_hip_copy_device": [_P, _P, c_size_t, _P],
    "lux_hip_bvh_create": [POINTER(c_void_p)],
    "lux_hip_bvh_destroy": [_P],
    "lux_hip_bvh_build": [_P, _P, _P, c_int32],
    "lux_hip_bvh_build_device": [_P, _P, _P, c_int32, _P],
    "lux_hip_bvh_refit": [_P, _P, _P, c_int32],
    "lux_hip_bvh_refit_device": [_P, _P, c_int32, _P],
    "lux_hip_bvh_refit_triangles_device": [_P, _P, c_int32, _P],
    "lux_hip_bvh_set_ray_order": [_P, c_int32],
    "lux_hip_bvh_info": [_P, POINTER(c_int32), POINTER(c_int32), POINTER(Aabb)],
    "lux_hip_bvh_view": [_P, POINTER(BvhView)],
    "lux_hip_bvh_download_nodes": [_P, _P, c_int32],
    "lux_hip_bvh_download_topology": [_P, _P, _P, _P, _P, _P, _P, _P],
    "lux_hip_bvh_download_input": [_P, _P, _P],
    "lux_hip_bvh_build_triangles_device": [_P, _P, c_int32, _P],
    "lux_hip_bvh_device_arrays": [_P, _P, _P, _P],
    "lux_hip_trace_closest": [_P, _P, c_int32, _P, c_int32, _P],
    "lux_hip_trace_closest_device": [_P, _P, c_int32, _P, c_int32, _P, _P],
    "lux_hip_trace_closest_device_noalloc": [_P, _P, c_int32, _P, c_int32, _P, _P],
    "lux_hip_trace_any": [_P, _P, c_int32, _P, c_int32, _P],
    "lux_hip_trace_any_device": [_P, _P, c_int32, _P, c_int32, _P, _P],
    "lux_hip_overlap_self": [_P, c_int32, _P, POINTER(c_int32)],
    "lux_hip_overlap_self_device": [_P, c_int32, _P, _P, _P],
    "lux_hip_overlap_cross": [_P, _P, c_int32, _P, POINTER(c_int32)],
    "lux_hip_overlap_cross_device": [_P, _P, c_int32, _P, _P, _P],
    "lux_hip_probe_begin": [c_char_p],
    "lux_hip_probe_end": [POINTER(c_double), POINTER(c_int64)],
    "lux_hip_trace_stats_device": [_P, _P, c_int32, _P, c_int32, _P],
    "lux_hip_scene_create": [_P, POINTER(c_void_p)],
    "lux_hip_scene_create_deferred": [_P, POINTER(c_void_p)],
    "lux_hip_scene_resident": [_P, POINTER(c_int32)],
    "lux_hip_scene_destroy": [_P],
    "lux_hip_scene_camera": [_P, _P],
    "lux_hip_scene_set_camera": [_P, _P],
    "lux_hip_scene_lights": [_P, POINTER(c_int32), _P, _P, _P, _P, _P, _P],
    "lux_hip_scene_bvh": [_P, POINTER(c_void_p)],
    "lux_hip_render": [_P, _P, _P],
    "lux_hip_render_device": [_P, _P, _P, _P],
    "lux_hip_render_rays": [_P, _P, POINTER(c_uint64), POINTER(c_uint64)],
    "lux_hip_scene_info": [_P, POINTER(c_int32), POINTER(c_int32), POINTER(c_uint64)],
    "lux_hip_scene_update_vertices": [_P, _P, c_int32],
    "lux_hip_scene_create_instanced": [_P, POINTER(c_void_p)],
    "lux_hip_scene_set_instances": [_P, _P, c_int32],
    "lux_hip_scene_triangles": [_P, _P, _P, _P],
    "lux_hip_scene_light_bvh": [_P, POINTER(c_int32), _P, _P, _P, _P, POINTER(c_int32)],
    "lux_hip_batch_render": [_P, _P, _P, c_int32, c_int32],
    "lux_hip_batch_render_device": [_P, _P, _P, c_int32, c_int32, _P],
    "lux_hip_batch_rays": [_P, POINTER(c_uint64), POINTER(c_uint64)],
    "lux_hip_set_render_tuning": [_P],
    "lux_hip_get_render_tuning": [_P],
    "lux_hip_server_create": [_P, POINTER(c_void_p)],
    "lux_hip_server_destroy": [_P],
    "lux_hip_server_submit": [_P, _P, _P, _P, c_int32, POINTER(c_uint64)],
    "lux_hip_server_poll": [_P, c_int32, _P, c_int32, POINTER(c_int32)],
    "lux_hip_server_result": [_P, c_uint64, _P, POINTER(c_double)],
    "lux_hip_server_stats": [_P, _P],
    "lux_hip_server_run": [_P, c_int32, _P, _P, _P, _P, _P, _P, _P, POINTER(c_double)],
    "lux_hip_two_level_create": [POINTER(c_void_p)],
    "lux_hip_two_level_destroy": [_P],
    "lux_hip_two_level_build": [_P, _P, _P, c_int32, _P, c_int32],
    "lux_hip_two_level_set_instances": [_P, _P, c_int32],
    "lux_hip_two_level_intersect": [_P, _P, _P, c_int32, _P],
    "lux_hip_two_level_intersect_device": [_P, _P, _P, c_int32, _P, _P],
    "lux_hip_two_level_occluded": [_P, _P, _P, c_int32, _P],
    "lux_hip_two_level_occluded_device": [_P, _P, _P, c_int32, _P, _P],
    "lux_hip_two_level_info": [_P, POINTER(c_int32), POINTER(c_int64), POINTER(c_uint64)],
    "lux_hip_inverse_affine": [_P, _P],
    "lux_hip_transform_bounds": [_P, _P, _P],
    "lux_hip_film_ppm_device": [_P, c_int32, c_int32, c_int32, _P, _P],
    "lux_hip_film_pfm_device": [_P, c_int32, c_int32, c_int32, _P, _P],
    "lux_hip_write_ppm": [c_char_p, _P, c_int32, c_int32, c_int32],
    "lux_hip_write_pfm": [c_char_p, _P, c_int32, c_int32, c_int32],
    "lux_hip_ppm_thresholds": [_P],
}
_RESTYPES = {"lux_hip_last_error": c_char_p}

_lib = None


def exported_names():
    return list(_SIGNATURES)


def load():
    """Load liblux_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise ImportError(f"{LIB_PATH} is missing: run `python -c 'import __graft_entry__ as g; g.build()'`")
    lib = ctypes.CDLL(str(LIB_PATH))
    for name, args in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = _RESTYPES.get(name, c_int)
    _lib = lib
    return lib


def call(name, *args):
    lib = load()
    status = getattr(lib, name)(*args)
    if status != LUX_OK:
        msg = lib.lux_hip_last_error()
        raise LuxError(status, name, msg.decode() if msg else "")
    return status


def ptr(array):
    """Host data pointer of a C-contiguous numpy array (None for empty)."""
    if array is None:
        return None
    assert array.flags["C_CONTIGUOUS"], "arrays passed to the C ABI must be C-contiguous"
    return c_void_p(array.ctypes.data) if array.size else None

"""numpy-facing wavefront path tracer over the C ABI.

`Scene(scene_dict)` uploads a scene (see lux_amd.scenes for the dict layout):
PreparedScene + prepare_gpu (src/scene/prepared_scene.cu:19-23, :61-79).
`Scene.render(...)` is WavefrontPathTracer::render
(src/integrator/wavefront_pt.cu:51-84) and returns the film SUM over spp,
shape (height, width, 3), row 0 at the top (Film::pixels).
"""

import ctypes
from ctypes import POINTER, c_float, c_int32, c_uint64, c_void_p

import numpy as np

from . import abi, lbvh
from .abi import Vec3, call, ptr

PLAIN, WAVEFRONT = 1, 0


class Material(ctypes.Structure):
    _fields_ = [("type", c_int32), ("albedo", Vec3), ("emission", Vec3), ("emission_two_sided", c_int32),
                ("roughness", c_float), ("metallic", c_float), ("eta", c_float)]


class SceneDesc(ctypes.Structure):
    _fields_ = [("vertices", c_void_p), ("vertex_count", c_int32), ("indices", c_void_p), ("normals", c_void_p),
                ("material_ids", c_void_p), ("triangle_count", c_int32), ("materials", c_void_p),
                ("material_count", c_int32), ("camera", c_float * 8)]


class RenderSettings(ctypes.Structure):
    _fields_ = [("width", c_int32), ("height", c_int32), ("spp", c_int32), ("max_depth", c_int32),
                ("seed", c_uint64), ("frame_index", c_int32), ("sample_offset", c_int32), ("batch_size", c_int32),
                ("mode", c_int32)]


assert ctypes.sizeof(Material) == 44


def _register():
    lib = abi.load()
    sigs = {
        "lux_hip_scene_create": [POINTER(SceneDesc), POINTER(c_void_p)],
        "lux_hip_scene_destroy": [c_void_p],
        "lux_hip_scene_camera": [c_void_p, c_void_p],
        "lux_hip_scene_lights": [c_void_p, POINTER(c_int32)] + [c_void_p] * 6,
        "lux_hip_scene_bvh": [c_void_p, POINTER(c_void_p)],
        "lux_hip_render": [c_void_p, POINTER(RenderSettings), c_void_p],
        "lux_hip_render_device": [c_void_p, POINTER(RenderSettings), c_void_p, c_void_p],
    }
    for name, args in sigs.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = ctypes.c_int
    return lib


def settings(width, height, spp, depth, seed=0, frame=0, sample_offset=0, batch_size=0, mode=WAVEFRONT):
    return RenderSettings(width, height, spp, depth, seed, frame, sample_offset, batch_size, mode)


class Scene:
    def __init__(self, scene):
        lbvh.init()
        _register()
        self._v = np.ascontiguousarray(scene["vertices"], np.float32)
        self._i = np.ascontiguousarray(scene["indices"], np.uint32)
        n = scene.get("normals")
        self._n = None if n is None or not np.any(n) else np.ascontiguousarray(n, np.float32)
        self._m = np.ascontiguousarray(scene["material_ids"], np.int32)
        mats = np.asarray(scene["materials"], np.float32)
        self._mats = (Material * len(mats))()
        for k, row in enumerate(mats):
            self._mats[k] = Material(int(row[0]), Vec3(*row[1:4]), Vec3(*row[4:7]), int(row[7]), float(row[8]),
                                     float(row[9]), float(row[10]))
        desc = SceneDesc()
        desc.vertices = self._v.ctypes.data
        desc.vertex_count = len(self._v)
        desc.indices = self._i.ctypes.data
        desc.normals = None if self._n is None else self._n.ctypes.data
        desc.material_ids = self._m.ctypes.data
        desc.triangle_count = len(self._m)
        desc.materials = ctypes.addressof(self._mats)
        desc.material_count = len(mats)
        cam = np.asarray(scene["camera"], np.float32)
        for k in range(8):
            desc.camera[k] = float(cam[k])
        h = c_void_p()
        call("lux_hip_scene_create", ctypes.byref(desc), ctypes.byref(h))
        self.handle = h
        self.triangle_count = len(self._m)

    def close(self):
        if self.handle:
            call("lux_hip_scene_destroy", self.handle)
            self.handle = c_void_p()

    def __del__(self):
        try:
            if self.handle:
                abi.load().lux_hip_scene_destroy(self.handle)
        except Exception:
            pass

    def camera(self):
        out = np.zeros(12, np.float32)
        call("lux_hip_scene_camera", self.handle, ptr(out))
        return out

    def lights(self):
        n = c_int32()
        call("lux_hip_scene_lights", self.handle, ctypes.byref(n), None, None, None, None, None, None)
        k = n.value
        out = {"prim": np.zeros(k, np.int32), "area": np.zeros(k, np.float32), "prob": np.zeros(k, np.float32),
               "alias": np.zeros(k, np.int32), "pmf": np.zeros(k, np.float32),
               "primitive_light_index": np.zeros(self.triangle_count, np.int32)}
        call("lux_hip_scene_lights", self.handle, ctypes.byref(n), ptr(out["prim"]), ptr(out["area"]),
             ptr(out["prob"]), ptr(out["alias"]), ptr(out["pmf"]), ptr(out["primitive_light_index"]))
        return out

    def render(self, width, height, spp, depth, **kw):
        s = settings(width, height, spp, depth, **kw)
        film = np.zeros((height, width, 3), np.float32)
        call("lux_hip_render", self.handle, ctypes.byref(s), ptr(film))
        return film

    def render_device(self, d_film, width, height, spp, depth, stream=None, **kw):
        s = settings(width, height, spp, depth, **kw)
        call("lux_hip_render_device", self.handle, ctypes.byref(s), d_film.ptr, stream)


def rel_rmse(test, ref, eps=1e-2):
    """sqrt(mean((g - c)^2 / (c^2 + eps))) over pixels and channels."""
    g = np.asarray(test, np.float64)
    c = np.asarray(ref, np.float64)
    return float(np.sqrt(np.mean((g - c) ** 2 / (c * c + eps))))

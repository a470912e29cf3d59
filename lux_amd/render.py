"""numpy-facing wavefront path tracer over the C ABI.

`Scene(scene_dict)` uploads a scene (see lux_amd.scenes for the dict layout):
PreparedScene + prepare_gpu (src/scene/prepared_scene.cu:19-23, :61-79). A
dict with "inst_mesh" is an instanced scene (several meshes placed by affine
instances): "vertices" / "indices" / "material_ids" / "normals" of all meshes
concatenated, "vertex_offsets" and "triangle_offsets" (mesh_count + 1 each,
indices mesh-local), "inst_mesh", "inst_m" (row-major object_to_world 4x4) and
optionally "inst_override" (material override, -1 for none).
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


class LuxLight(ctypes.Structure):
    _fields_ = [("kind", c_int32), ("vector", Vec3), ("color", Vec3), ("image_width", c_int32),
                ("image_height", c_int32), ("image_pixels", c_void_p)]


LIGHT_KINDS = {"point": 1, "distant": 2, "uniform_infinite": 3, "image_infinite": 4}


class SceneDesc(ctypes.Structure):
    _fields_ = [("vertices", c_void_p), ("vertex_count", c_int32), ("indices", c_void_p), ("normals", c_void_p),
                ("material_ids", c_void_p), ("triangle_count", c_int32), ("materials", c_void_p),
                ("material_count", c_int32), ("camera", c_float * 8), ("lights", c_void_p), ("light_count", c_int32)]


class InstancedSceneDesc(ctypes.Structure):
    _fields_ = [("vertices", c_void_p), ("normals", c_void_p), ("vertex_offsets", c_void_p), ("indices", c_void_p),
                ("material_ids", c_void_p), ("triangle_offsets", c_void_p), ("mesh_count", c_int32),
                ("instances", c_void_p), ("instance_count", c_int32), ("materials", c_void_p),
                ("material_count", c_int32), ("camera", c_float * 8), ("lights", c_void_p), ("light_count", c_int32)]


def _scene_lights(lights):
    """Scene-light dicts ({"kind": "point" | "distant" | "uniform_infinite" |
    "image_infinite", "vector": position or direction, "color": intensity,
    radiance or image scale, "image": (H, W, 3) texels}) as a lux_light array
    and the arrays it points to."""
    lights = lights or []
    arr = (LuxLight * max(len(lights), 1))()
    keep = [arr]
    for k, x in enumerate(lights):
        arr[k].kind = LIGHT_KINDS[x["kind"]]
        arr[k].vector = Vec3(*[float(c) for c in x.get("vector", (0, 0, 0))])
        arr[k].color = Vec3(*[float(c) for c in x["color"]])
        if x["kind"] == "image_infinite":
            img = np.ascontiguousarray(x["image"], np.float32)
            arr[k].image_height, arr[k].image_width = img.shape[0], img.shape[1]
            arr[k].image_pixels = img.ctypes.data
            keep.append(img)
    return arr, len(lights), keep


class RenderSettings(ctypes.Structure):
    _fields_ = [("width", c_int32), ("height", c_int32), ("spp", c_int32), ("max_depth", c_int32),
                ("seed", c_uint64), ("frame_index", c_int32), ("sample_offset", c_int32), ("batch_size", c_int32),
                ("mode", c_int32), ("light_sampler", c_int32)]


LIGHT_SAMPLERS = {"power": 0, "uniform": 1, "bvh": 2}


class Tuning(ctypes.Structure):
    _fields_ = [("capacity", c_int32), ("grid", c_int32), ("block", c_int32), ("use_graph", c_int32),
                ("preorder_walk", c_int32)]


assert ctypes.sizeof(Material) == 44


def _register():
    lib = abi.load()
    sigs = {
        "lux_hip_scene_create": [POINTER(SceneDesc), POINTER(c_void_p)],
        "lux_hip_scene_create_instanced": [POINTER(InstancedSceneDesc), POINTER(c_void_p)],
        "lux_hip_scene_set_instances": [c_void_p, c_void_p, c_int32],
        "lux_hip_scene_triangles": [c_void_p, c_void_p, c_void_p, c_void_p],
        "lux_hip_scene_destroy": [c_void_p],
        "lux_hip_scene_camera": [c_void_p, c_void_p],
        "lux_hip_scene_lights": [c_void_p, POINTER(c_int32)] + [c_void_p] * 6,
        "lux_hip_scene_bvh": [c_void_p, POINTER(c_void_p)],
        "lux_hip_render": [c_void_p, POINTER(RenderSettings), c_void_p],
        "lux_hip_render_device": [c_void_p, POINTER(RenderSettings), c_void_p, c_void_p],
        "lux_hip_render_rays": [c_void_p, c_void_p, POINTER(c_uint64), POINTER(c_uint64)],
    }
    for name, args in sigs.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = ctypes.c_int
    return lib


def settings(width, height, spp, depth, seed=0, frame=0, sample_offset=0, batch_size=0, mode=WAVEFRONT,
             light_sampler="power"):
    ls = LIGHT_SAMPLERS[light_sampler] if isinstance(light_sampler, str) else int(light_sampler)
    return RenderSettings(width, height, spp, depth, seed, frame, sample_offset, batch_size, mode, ls)


class Scene:
    instanced = False

    def __init__(self, scene, deferred=False):
        """deferred: keep the description only; the device state is built on
        first use (a render, an inspection call, or a server request)."""
        lbvh.init()
        _register()
        from . import autotune

        autotune.apply()
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
        if "inst_mesh" in scene:
            self._create_instanced(scene, len(mats))
            return
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
        arr, n, self._light_keep = _scene_lights(scene.get("lights"))
        desc.lights = ctypes.addressof(arr) if n else None
        desc.light_count = n
        h = c_void_p()
        call("lux_hip_scene_create_deferred" if deferred else "lux_hip_scene_create", ctypes.byref(desc),
             ctypes.byref(h))
        self.handle = h
        self.triangle_count = len(self._m)
        if "camera_vectors" in scene:
            self.set_camera(scene["camera_vectors"])

    def _create_instanced(self, scene, material_count):
        from . import two_level as TL

        self.instanced = True
        self._voff = np.ascontiguousarray(scene["vertex_offsets"], np.int32)
        self._toff = np.ascontiguousarray(scene["triangle_offsets"], np.int32)
        self._inst = np.ascontiguousarray(TL.make_instances(
            scene["inst_mesh"], scene["inst_m"], inverses=scene.get("inst_inv"),
            world_bounds=np.zeros((len(scene["inst_mesh"]), 6), np.float32),
            material_override=scene.get("inst_override")), TL.INSTANCE)
        desc = InstancedSceneDesc()
        desc.vertices = self._v.ctypes.data
        desc.normals = None if self._n is None else self._n.ctypes.data
        desc.vertex_offsets = self._voff.ctypes.data
        desc.indices = self._i.ctypes.data
        desc.material_ids = self._m.ctypes.data
        desc.triangle_offsets = self._toff.ctypes.data
        desc.mesh_count = len(self._voff) - 1
        desc.instances = self._inst.ctypes.data
        desc.instance_count = len(self._inst)
        desc.materials = ctypes.addressof(self._mats)
        desc.material_count = material_count
        cam = np.asarray(scene["camera"], np.float32)
        for k in range(8):
            desc.camera[k] = float(cam[k])
        arr, n, self._light_keep = _scene_lights(scene.get("lights"))
        desc.lights = ctypes.addressof(arr) if n else None
        desc.light_count = n
        h = c_void_p()
        call("lux_hip_scene_create_instanced", ctypes.byref(desc), ctypes.byref(h))
        self.handle = h
        self.triangle_count = self.info()[0]
        if "camera_vectors" in scene:
            self.set_camera(scene["camera_vectors"])

    def set_camera(self, vectors12):
        """Built camera vectors {origin, lower_left, horizontal, vertical}."""
        v = np.ascontiguousarray(vectors12, np.float32)[:12]
        call("lux_hip_scene_set_camera", self.handle, ptr(v))

    def set_instances(self, inst_mesh, inst_m, inst_override=None):
        """Move the instances of an instanced scene (same count): TLAS rebuild,
        world-space triangles, normals and light table; the BLASes are kept."""
        from . import two_level as TL

        inst = np.ascontiguousarray(TL.make_instances(
            inst_mesh, inst_m, world_bounds=np.zeros((len(inst_mesh), 6), np.float32),
            material_override=inst_override), TL.INSTANCE)
        call("lux_hip_scene_set_instances", self.handle, inst.ctypes.data, len(inst))
        self._inst = inst
        self.triangle_count = self.info()[0]

    def light_bvh(self):
        """The BVH light sampler's tree: bounds (N, 6), power (N,), links (N, 4:
        left, right, parent, light), leaf node per light, root."""
        n, root = c_int32(), c_int32()
        call("lux_hip_scene_light_bvh", self.handle, ctypes.byref(n), None, None, None, None, ctypes.byref(root))
        k = n.value
        nl = len(self.lights()["prim"])
        out = {"bounds": np.zeros((k, 6), np.float32), "power": np.zeros(k, np.float32),
               "links": np.zeros((k, 4), np.int32), "leaf": np.zeros(nl if k else 0, np.int32)}
        call("lux_hip_scene_light_bvh", self.handle, ctypes.byref(n), ptr(out["bounds"]), ptr(out["power"]),
             ptr(out["links"]), ptr(out["leaf"]) if k else None, ctypes.byref(root))
        out["root"] = root.value
        return out

    def triangles(self):
        """The renderer's shading view per scene ordinal: world corners (T, 9),
        world corner normals (T, 9), material after the override (T,)."""
        t = self.info()[0]
        out = {"corners": np.zeros((t, 9), np.float32), "normals": np.zeros((t, 9), np.float32),
               "material": np.zeros(t, np.int32)}
        call("lux_hip_scene_triangles", self.handle, ptr(out["corners"]), ptr(out["normals"]), ptr(out["material"]))
        return out

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

    def rays(self, stream=None):
        """(extension rays, shadow rays) traced by the last render."""
        e, sh = c_uint64(), c_uint64()
        call("lux_hip_render_rays", self.handle, stream, ctypes.byref(e), ctypes.byref(sh))
        return e.value, sh.value

    def info(self):
        """(triangles, vertices, device bytes)."""
        t, v, b = c_int32(), c_int32(), c_uint64()
        call("lux_hip_scene_info", self.handle, ctypes.byref(t), ctypes.byref(v), ctypes.byref(b))
        return t.value, v.value, b.value

    def resident(self):
        """True while the scene's device state exists (a server may drop it)."""
        r = c_int32()
        call("lux_hip_scene_resident", self.handle, ctypes.byref(r))
        return bool(r.value)

    def update_vertices(self, vertices):
        """Move the vertices (same count): device regather + LBVH refit + light table."""
        v = np.ascontiguousarray(vertices, np.float32)
        assert v.shape == self._v.shape
        call("lux_hip_scene_update_vertices", self.handle, ptr(v), len(v))
        self._v = v


def _batch_args(scenes, jobs):
    n = len(scenes)
    hs = (c_void_p * n)(*[sc.handle.value for sc in scenes])
    ss = (RenderSettings * n)(*[settings(**j) for j in jobs])
    return n, hs, ss


def batch_render(scenes, jobs, mode=WAVEFRONT):
    """Render jobs[i] (a dict of settings() arguments) of scenes[i], all in one
    multi-scene wavefront loop (mode WAVEFRONT) or one after another on the
    plain schedule. Returns the host film sums."""
    n, hs, ss = _batch_args(scenes, jobs)
    films = [np.zeros((j["height"], j["width"], 3), np.float32) for j in jobs]
    fp = (c_void_p * n)(*[f.ctypes.data for f in films])
    call("lux_hip_batch_render", hs, ss, fp, n, mode)
    return films


def batch_render_device(scenes, jobs, d_films, mode=WAVEFRONT, stream=None):
    n, hs, ss = _batch_args(scenes, jobs)
    fp = (c_void_p * n)(*[f.ptr.value for f in d_films])
    call("lux_hip_batch_render_device", hs, ss, fp, n, mode, stream)


def batch_rays(stream=None):
    e, sh = c_uint64(), c_uint64()
    call("lux_hip_batch_rays", stream, ctypes.byref(e), ctypes.byref(sh))
    return e.value, sh.value


def get_tuning():
    lbvh.init()
    t = Tuning()
    call("lux_hip_get_render_tuning", ctypes.byref(t))
    return {"capacity": t.capacity, "grid": t.grid, "block": t.block, "use_graph": t.use_graph,
            "preorder_walk": t.preorder_walk}


def set_tuning(capacity=None, grid=None, block=None, use_graph=None, preorder_walk=None):
    """Launch geometry of the wavefront schedule for renders started afterwards;
    preorder_walk=True makes extension rays take the right-first preorder walk
    only (the host LBVH's visit order)."""
    cur = get_tuning()
    t = Tuning(capacity or cur["capacity"], grid or cur["grid"], block or cur["block"],
               cur["use_graph"] if use_graph is None else int(use_graph),
               cur["preorder_walk"] if preorder_walk is None else int(preorder_walk))
    call("lux_hip_set_render_tuning", ctypes.byref(t))


def rel_rmse(test, ref, eps=1e-2):
    """sqrt(mean((g - c)^2 / (c^2 + eps))) over pixels and channels."""
    g = np.asarray(test, np.float64)
    c = np.asarray(ref, np.float64)
    return float(np.sqrt(np.mean((g - c) ** 2 / (c * c + eps))))


class ServeConfig(ctypes.Structure):
    _fields_ = [("hbm_budget_bytes", c_uint64), ("slice_spp", c_int32), ("max_inflight", c_int32)]


class ServeStats(ctypes.Structure):
    _fields_ = [(n, c_uint64) for n in ("submitted", "admitted", "completed", "refits", "batches", "paths",
                                        "resident_bytes", "peak_bytes", "extension_rays", "shadow_rays",
                                        "shared_batches", "max_batch_requests", "uploads", "evictions")] + \
               [("prepare_seconds", ctypes.c_double)]


class Server:
    """Continuous-batching render server (lux_hip_server_*)."""

    def __init__(self, hbm_budget=200 << 30, slice_spp=16, max_inflight=2):
        lbvh.init()
        cfg = ServeConfig(int(hbm_budget), int(slice_spp), int(max_inflight))
        h = c_void_p()
        call("lux_hip_server_create", ctypes.byref(cfg), ctypes.byref(h))
        self.handle = h
        self._shapes = {}
        self._keep = {}

    def close(self):
        if self.handle:
            call("lux_hip_server_destroy", self.handle)
            self.handle = c_void_p()

    def __del__(self):
        try:
            if self.handle:
                abi.load().lux_hip_server_destroy(self.handle)
        except Exception:
            pass

    def submit(self, scene, job, vertices=None):
        s = settings(**job)
        v = None if vertices is None else np.ascontiguousarray(vertices, np.float32)
        rid = c_uint64()
        call("lux_hip_server_submit", self.handle, scene.handle, ctypes.byref(s), ptr(v),
             0 if v is None else len(v), ctypes.byref(rid))
        self._shapes[rid.value] = (job["height"], job["width"], 3)
        return rid.value

    def poll(self, block=False, max_ids=256):
        ids = (c_uint64 * max_ids)()
        n = c_int32()
        call("lux_hip_server_poll", self.handle, int(bool(block)), ids, max_ids, ctypes.byref(n))
        return [ids[i] for i in range(n.value)]

    def result(self, rid):
        film = np.zeros(self._shapes.pop(rid), np.float32)
        lat = ctypes.c_double()
        call("lux_hip_server_result", self.handle, c_uint64(rid), ptr(film), ctypes.byref(lat))
        return film, lat.value

    def stats(self):
        st = ServeStats()
        call("lux_hip_server_stats", self.handle, ctypes.byref(st))
        return {n: getattr(st, n) for n, _ in ServeStats._fields_}

    def run(self, scenes, jobs, arrivals, vertices=None, want_films=True):
        """Replay an arrival trace (seconds, sorted). vertices[i] (or None) moves
        scenes[i]'s vertices before request i. Returns (films or None, latencies
        in seconds, wall seconds)."""
        n = len(scenes)
        hs = (c_void_p * n)(*[sc.handle.value for sc in scenes])
        ss = (RenderSettings * n)(*[settings(**j) for j in jobs])
        arr = np.ascontiguousarray(arrivals, np.float64)
        lat = np.zeros(n, np.float64)
        wall = ctypes.c_double()
        films = [np.zeros((j["height"], j["width"], 3), np.float32) for j in jobs] if want_films else None
        fp = (c_void_p * n)(*[f.ctypes.data for f in films]) if want_films else None
        vp = vc = None
        keep = []
        if vertices is not None:
            vp = (c_void_p * n)()
            vc = np.zeros(n, np.int32)
            for i, v in enumerate(vertices):
                if v is not None:
                    a = np.ascontiguousarray(v, np.float32)
                    keep.append(a)
                    vp[i] = a.ctypes.data
                    vc[i] = len(a)
        call("lux_hip_server_run", self.handle, n, hs, ss, ptr(arr), vp, ptr(vc) if vc is not None else None, fp,
             ptr(lat), ctypes.byref(wall))
        return films, lat, wall.value

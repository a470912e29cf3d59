"""Two-level (instanced) scenes over liblux_hip.so: the mirror of the
reference's PreparedScene BVH part (one LBVH per mesh, one over the instances'
world bounds; src/scene/prepared_scene.cu:59-78) and of
intersect_scene_bvh / intersect_scene_bvh_skip_ref (src/scene/gpu_scene_bvh.cuh:44-135).

Meshes are runs of object-space triangles; an instance record holds the mesh
id, Transform::m and Transform::inv (row-major 4x4) and the world bounds.
"""

import ctypes

import numpy as np

from . import abi
from .abi import call

INSTANCE = np.dtype([("mesh_id", np.int32), ("material_override", np.int32), ("object_to_world", np.float32, 16),
                     ("world_to_object", np.float32, 16), ("world_bounds", np.float32, 6)])
SCENE_HIT = np.dtype([("primitive_id", np.int32), ("t", np.float32), ("u", np.float32), ("v", np.float32),
                      ("instance_id", np.int32), ("local_primitive_id", np.int32)])
assert INSTANCE.itemsize == 160 and SCENE_HIT.itemsize == 24


def _p(a):
    return None if a is None else ctypes.c_void_p(a.ctypes.data)


def inverse_affine(m):
    """make_transform's inverse (the reference's inverse_affine)."""
    a = np.ascontiguousarray(m, np.float32).reshape(16)
    out = np.zeros(16, np.float32)
    call("lux_hip_inverse_affine", _p(a), _p(out))
    return out.reshape(4, 4)


def transform_bounds(m, box):
    """The reference's transform_bounds of one box (6 floats: min xyz, max xyz)."""
    a = np.ascontiguousarray(m, np.float32).reshape(16)
    b = np.ascontiguousarray(box, np.float32).reshape(6)
    out = np.zeros(6, np.float32)
    call("lux_hip_transform_bounds", _p(a), _p(b), _p(out))
    return out


def mesh_bounds(tris, mesh_offsets):
    """Per mesh: the union of its triangle boxes (finalize_scene_mesh); empty
    meshes get the empty box (+inf / -inf)."""
    t = np.asarray(tris, np.float32).reshape(-1, 3, 3)
    out = np.tile(np.array([np.inf] * 3 + [-np.inf] * 3, np.float32), (len(mesh_offsets) - 1, 1))
    for m in range(len(mesh_offsets) - 1):
        s = t[mesh_offsets[m]:mesh_offsets[m + 1]]
        if len(s):
            out[m, :3] = s.min(axis=(0, 1))
            out[m, 3:] = s.max(axis=(0, 1))
    return out


def make_instances(mesh_ids, matrices, tris=None, mesh_offsets=None, inverses=None, world_bounds=None,
                   material_override=None):
    """Instance records. Missing inverses come from inverse_affine; missing world
    bounds from transform_bounds of the mesh bounds (needs tris, mesh_offsets)."""
    mesh_ids = np.asarray(mesh_ids, np.int32)
    m = np.asarray(matrices, np.float32).reshape(-1, 16)
    rec = np.zeros(len(mesh_ids), INSTANCE)
    rec["mesh_id"] = mesh_ids
    rec["material_override"] = -1 if material_override is None else np.asarray(material_override, np.int32)
    rec["object_to_world"] = m
    if inverses is None:
        inverses = [inverse_affine(x).reshape(16) for x in m] if len(m) else np.zeros((0, 16), np.float32)
    rec["world_to_object"] = np.asarray(inverses, np.float32).reshape(-1, 16)
    if world_bounds is None:
        mb = mesh_bounds(tris, mesh_offsets)
        world_bounds = np.stack([transform_bounds(x, mb[k]) for x, k in zip(m, mesh_ids)]) if len(m) else \
            np.zeros((0, 6), np.float32)
    rec["world_bounds"] = np.asarray(world_bounds, np.float32).reshape(-1, 6)
    return rec


class TwoLevelScene:
    """BLAS per mesh + TLAS over instances (lux_hip_two_level)."""

    def __init__(self, tris, mesh_offsets, instances):
        abi.load()
        h = ctypes.c_void_p()
        call("lux_hip_two_level_create", ctypes.byref(h))
        self.handle = h
        self.build(tris, mesh_offsets, instances)

    def build(self, tris, mesh_offsets, instances):
        t = np.ascontiguousarray(tris, np.float32).reshape(-1, 9)
        off = np.ascontiguousarray(mesh_offsets, np.int32)
        inst = np.ascontiguousarray(instances, INSTANCE)
        call("lux_hip_two_level_build", self.handle, _p(t), _p(off), len(off) - 1, _p(inst), len(inst))

    def set_instances(self, instances):
        inst = np.ascontiguousarray(instances, INSTANCE)
        call("lux_hip_two_level_set_instances", self.handle, _p(inst), len(inst))

    def intersect(self, rays, ignore=None):
        r = np.ascontiguousarray(rays, np.float32).reshape(-1, 9)
        ig = None if ignore is None else np.ascontiguousarray(ignore, np.int32)
        out = np.zeros(len(r), SCENE_HIT)
        call("lux_hip_two_level_intersect", self.handle, _p(r), _p(ig), len(r), _p(out))
        return out

    def intersect_device(self, d_rays, n, d_hits, d_ignore=None, stream=None):
        call("lux_hip_two_level_intersect_device", self.handle, d_rays.ptr, d_ignore.ptr if d_ignore else None, n,
             d_hits.ptr, stream)

    def occluded(self, rays, ignore=None):
        r = np.ascontiguousarray(rays, np.float32).reshape(-1, 9)
        ig = None if ignore is None else np.ascontiguousarray(ignore, np.int32)
        out = np.zeros(len(r), np.uint8)
        call("lux_hip_two_level_occluded", self.handle, _p(r), _p(ig), len(r), _p(out))
        return out

    def occluded_device(self, d_rays, n, d_occluded, d_ignore=None, stream=None):
        call("lux_hip_two_level_occluded_device", self.handle, d_rays.ptr, d_ignore.ptr if d_ignore else None, n,
             d_occluded.ptr, stream)

    def info(self):
        a, b, c = ctypes.c_int32(), ctypes.c_int64(), ctypes.c_uint64()
        call("lux_hip_two_level_info", self.handle, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c))
        return {"tlas_instances": a.value, "scene_primitives": b.value, "device_bytes": c.value}

    def close(self):
        if getattr(self, "handle", None) is not None and self.handle.value:
            call("lux_hip_two_level_destroy", self.handle)
            self.handle = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

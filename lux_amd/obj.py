"""OBJ import over the C ABI (lux_obj_*): load_obj_into_scene
(src/util/obj_loader.cpp:613-1311) for the path tracer's instanced scenes.

`load(path, **options)` returns the appended mesh (object space) as arrays and
the instance transform the reference loader would create; option names follow
ObjLoadOptions (src/util/obj_loader.h:52-67)."""

import ctypes
from ctypes import POINTER, c_char_p, c_float, c_int32, c_void_p

import numpy as np

from . import abi
from .abi import call

NORMAL_MODES = {"source_or_geometry": 0, "generate_missing_smooth": 1, "face_only": 2}
SCALE_MODES = {"none": 0, "max_extent": 1, "x_extent": 2}
TOPOLOGIES = {"source_vertices": 0, "welded_positions": 1}
CAP_COUNTS = ("cap_triangles", "cap_boundary_edges", "cap_non_simple_vertices", "cap_incomplete_loops",
              "cap_degenerate_loops", "cap_degenerate_triangles")


class Options(ctypes.Structure):
    _fields_ = [("normal_mode", c_int32), ("reverse_orientation", c_int32), ("flip_texcoord_v", c_int32),
                ("boundary_topology", c_int32), ("position_weld_epsilon", c_float),
                ("default_material_id", c_int32), ("scene_material_count", c_int32),
                ("placement_scale_mode", c_int32), ("placement_target_extent", c_float),
                ("placement_center_xz", c_int32), ("placement_ground_to_y", c_int32),
                ("placement_target_y", c_float), ("object_to_world", c_float * 16),
                ("world_to_object", c_float * 16), ("binding_count", c_int32),
                ("binding_names", POINTER(c_char_p)), ("binding_material_ids", POINTER(c_int32)),
                ("cap_boundary_holes", c_int32), ("cap_boundary_topology", c_int32)]


def _register():
    lib = abi.load()
    sigs = {
        "lux_obj_default_options": [POINTER(Options)],
        "lux_obj_load": [c_char_p, POINTER(Options), POINTER(c_void_p)],
        "lux_obj_mesh_free": [c_void_p],
        "lux_obj_mesh_counts": [c_void_p] + [POINTER(c_int32)] * 6,
        "lux_obj_mesh_cap_counts": [c_void_p] + [POINTER(c_int32)] * 6,
        "lux_obj_mesh_arrays": [c_void_p] * 9,
    }
    for name, args in sigs.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = ctypes.c_int
    return lib


def _code(table, value):
    return table[value] if isinstance(value, str) else int(value)


def options(normal_mode="source_or_geometry", reverse_orientation=False, flip_texcoord_v=False,
            boundary_topology="source_vertices", position_weld_epsilon=1e-6, default_material_id=-1,
            scene_material_count=0, scale_mode="none", target_extent=1.0, center_xz=False, ground_to_y=False,
            target_y=0.0, object_to_world=None, material_bindings=None, cap_boundary_holes=False,
            cap_boundary_topology="welded_positions"):
    """An Options record (the reference's defaults unless given; enum options
    by name or by their ObjLoadOptions value) and the objects it points to
    (keep them alive while it is used)."""
    _register()
    o = Options()
    call("lux_obj_default_options", ctypes.byref(o))
    o.normal_mode = _code(NORMAL_MODES, normal_mode)
    o.reverse_orientation = int(reverse_orientation)
    o.flip_texcoord_v = int(flip_texcoord_v)
    o.boundary_topology = _code(TOPOLOGIES, boundary_topology)
    o.position_weld_epsilon = position_weld_epsilon
    o.default_material_id = default_material_id
    o.scene_material_count = scene_material_count
    o.placement_scale_mode = _code(SCALE_MODES, scale_mode)
    o.placement_target_extent = target_extent
    o.placement_center_xz = int(center_xz)
    o.placement_ground_to_y = int(ground_to_y)
    o.placement_target_y = target_y
    o.cap_boundary_holes = int(cap_boundary_holes)
    o.cap_boundary_topology = _code(TOPOLOGIES, cap_boundary_topology)
    keep = []
    if object_to_world is not None:
        from .two_level import inverse_affine

        m = np.asarray(object_to_world, np.float32).reshape(16)
        inv = inverse_affine(m.reshape(4, 4)).reshape(16)
        for k in range(16):
            o.object_to_world[k] = float(m[k])
            o.world_to_object[k] = float(inv[k])
    if material_bindings:
        names = (c_char_p * len(material_bindings))(*[k.encode() for k in material_bindings])
        ids = (c_int32 * len(material_bindings))(*material_bindings.values())
        o.binding_count = len(material_bindings)
        o.binding_names = ctypes.cast(names, POINTER(c_char_p))
        o.binding_material_ids = ctypes.cast(ids, POINTER(c_int32))
        keep += [names, ids]
    return o, keep


def load(path, **kw):
    """The mesh load_obj_into_scene appends: dict of vertices (V, 3), normals
    (V, 3), uvs (V, 2), indices (T, 3), material_ids (T,), boundary_edges (T,),
    object_to_world / world_to_object (4, 4), appended_default_material, the
    skipped-face counts and the boundary-capping counts (CAP_COUNTS)."""
    o, keep = options(**kw)
    h = c_void_p()
    call("lux_obj_load", str(path).encode(), ctypes.byref(o), ctypes.byref(h))
    try:
        c = [c_int32() for _ in range(6)]
        call("lux_obj_mesh_counts", h, *[ctypes.byref(x) for x in c])
        nv, nt = c[0].value, c[1].value
        out = {"vertices": np.zeros((nv, 3), np.float32), "normals": np.zeros((nv, 3), np.float32),
               "uvs": np.zeros((nv, 2), np.float32), "indices": np.zeros((nt, 3), np.uint32),
               "material_ids": np.zeros(nt, np.int32), "boundary_edges": np.zeros(nt, np.uint8),
               "object_to_world": np.zeros((4, 4), np.float32), "world_to_object": np.zeros((4, 4), np.float32)}
        call("lux_obj_mesh_arrays", h, *[v.ctypes.data for v in out.values()])
        out.update(appended_default_material=c[2].value, skipped_invalid=c[3].value,
                   skipped_nonfinite=c[4].value, skipped_degenerate=c[5].value)
        cc = [c_int32() for _ in range(6)]
        call("lux_obj_mesh_cap_counts", h, *[ctypes.byref(x) for x in cc])
        out.update(zip(CAP_COUNTS, (x.value for x in cc)))
        return out
    finally:
        call("lux_obj_mesh_free", h)
        del keep

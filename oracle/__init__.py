"""CPU oracle for the LUX hot path -- TEST INFRASTRUCTURE ONLY.

``liblux_oracle.so`` (built from oracle/lbvh_oracle.cpp) restates the reference's
host LBVH, host stack traversal, brute-force intersection and overlap rules.
``oracle/_ref/libluxref.so`` (built by oracle/ref.mk from the reference's own
sources when they are present) is the reference itself behind a small shim.

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg use this
package, always as the checker; the product path (lux_amd) never imports it.
"""

import ctypes
from ctypes import POINTER, c_float, c_int32, c_int64, c_uint32, c_void_p
from pathlib import Path

import numpy as np

ORACLE_DIR = Path(__file__).resolve().parent
ORACLE_LIB = ORACLE_DIR / "_build" / "liblux_oracle.so"
REF_LIB = ORACLE_DIR / "_ref" / "libluxref.so"

_lib = None
_ref = None

# Soft window of the library's checked near-child-first walk (the library uses
# the same value): once it holds a hit at t, boxes are entered up to
# t + |t| * 2^-11.
ORDERED_WINDOW_LOG2 = 11


def _p(a):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return c_void_p(a.ctypes.data)


def lib():
    global _lib
    if _lib is None:
        if not ORACLE_LIB.exists():
            raise ImportError(f"{ORACLE_LIB} missing: run __graft_entry__.build()")
        L = ctypes.CDLL(str(ORACLE_LIB))
        L.orc_bvh_build.restype = c_void_p
        L.orc_bvh_build.argtypes = [c_void_p, c_void_p, c_int32]
        L.orc_bvh_free.argtypes = [c_void_p]
        L.orc_bvh_refit.argtypes = [c_void_p, c_void_p]
        L.orc_bvh_count.restype = c_int32
        L.orc_bvh_count.argtypes = [c_void_p]
        L.orc_bvh_topology.argtypes = [c_void_p] + [c_void_p] * 10
        L.orc_morton_codes.argtypes = [c_void_p, c_int32, c_void_p, c_void_p]
        L.orc_trace_closest.restype = c_int32
        L.orc_trace_closest.argtypes = [c_void_p, c_void_p, c_void_p, c_int32] + [c_void_p] * 5
        L.orc_trace_stackless.argtypes = [c_void_p, c_void_p, c_void_p, c_int32, c_void_p, c_void_p]
        L.orc_brute_closest.argtypes = [c_void_p, c_int32, c_void_p, c_int32, c_void_p, c_void_p]
        L.orc_tri_t_fp64.argtypes = [c_void_p, c_void_p, c_void_p, c_int32, c_void_p]
        L.orc_overlap_self_brute.restype = c_int64
        L.orc_overlap_self_brute.argtypes = [c_void_p, c_void_p, c_int32, c_void_p, c_int64]
        L.orc_overlap_cross_brute.restype = c_int64
        L.orc_overlap_cross_brute.argtypes = [c_void_p, c_void_p, c_int32, c_void_p, c_void_p, c_int32,
                                              c_void_p, c_int64]
        L.orc_overlap_tree.restype = c_int64
        L.orc_overlap_tree.argtypes = [c_void_p, c_void_p, c_void_p, c_int32, c_int32, c_void_p, c_int64]
        L.orc_triangle_bounds.argtypes = [c_void_p, c_int32, c_void_p]
        L.orc_trace_ordered.restype = c_int32
        L.orc_trace_ordered.argtypes = [c_void_p, c_void_p, c_void_p, c_int32, c_int32, c_int32, c_void_p,
                                        c_void_p, c_void_p]
        _lib = L
    return _lib


NODE_DTYPE = np.dtype([("p_min", "<f4", 3), ("p_max", "<f4", 3), ("primitive_id", "<i4"), ("escape", "<i4")])


class OracleBvh:
    """Host LBVH (build_cpu_lbvh_tree restated) + right-first stackless layout."""

    def __init__(self, boxes, ids=None):
        self.boxes = np.ascontiguousarray(boxes, np.float32).reshape(-1, 6)
        n = len(self.boxes)
        self.ids = np.arange(n, dtype=np.int32) if ids is None else np.ascontiguousarray(ids, np.int32)
        self.n = n
        self.h = lib().orc_bvh_build(_p(self.boxes), _p(self.ids), n)
        m = max(n - 1, 0)
        self.morton_sorted = np.zeros(n, np.uint32)
        self.perm = np.zeros(n, np.uint32)
        self.leaf_ids = np.zeros(n, np.int32)
        self.left = np.zeros(m, np.uint32)
        self.right = np.zeros(m, np.uint32)
        self.parent = np.zeros(m, np.uint32)
        self.leaf_parent = np.zeros(n, np.uint32)
        self.node_bounds = np.zeros((max(2 * n - 1, 0), 6), np.float32)
        self.root = np.zeros(6, np.float32)
        self.nodes = np.zeros(max(2 * n - 1, 0), NODE_DTYPE)
        if n:
            lib().orc_bvh_topology(self.h, _p(self.morton_sorted), _p(self.perm), _p(self.leaf_ids),
                                   _p(self.left), _p(self.right), _p(self.parent), _p(self.leaf_parent),
                                   _p(self.node_bounds), _p(self.root), _p(self.nodes))

    def refit(self, boxes):
        """Same topology, new input boxes (refit_gpu_lbvh restated)."""
        b = np.ascontiguousarray(boxes, np.float32).reshape(-1, 6)
        assert len(b) == self.n
        self.boxes = b
        lib().orc_bvh_refit(self.h, _p(b))
        if self.n:
            lib().orc_bvh_topology(self.h, None, None, None, None, None, None, None, _p(self.node_bounds),
                                   _p(self.root), _p(self.nodes))
        return self

    def __del__(self):
        try:
            lib().orc_bvh_free(self.h)
        except Exception:
            pass

    def host_children(self):
        """left/right as host node indices (internal i, leaf j -> n-1+j)."""
        n = self.n

        def conv(c):
            c = c.astype(np.int64)
            leaf = (c & 0x80000000) != 0
            return np.where(leaf, (c & 0x7FFFFFFF) + n - 1, c).astype(np.int32)

        return conv(self.left), conv(self.right)

    def trace_closest(self, tris, rays, with_stats=False):
        t = np.ascontiguousarray(tris, np.float32).reshape(-1, 9)
        r = np.ascontiguousarray(rays, np.float32).reshape(-1, 9)
        k = len(r)
        ids = np.zeros(k, np.int32)
        th = np.zeros(k, np.float32)
        u = np.zeros(k, np.float32)
        v = np.zeros(k, np.float32)
        stats = np.zeros(2, np.int64)
        ov = lib().orc_trace_closest(self.h, _p(t), _p(r), k, _p(ids), _p(th), _p(u), _p(v), _p(stats))
        out = {"ids": ids, "t": th, "u": u, "v": v, "overflows": int(ov)}
        if with_stats:
            out["nodes_visited"] = int(stats[0])
            out["tri_tests"] = int(stats[1])
        return out

    def trace_stackless(self, tris, rays):
        t = np.ascontiguousarray(tris, np.float32).reshape(-1, 9)
        r = np.ascontiguousarray(rays, np.float32).reshape(-1, 9)
        ids = np.zeros(len(r), np.int32)
        th = np.zeros(len(r), np.float32)
        lib().orc_trace_stackless(self.h, _p(t), _p(r), len(r), _p(ids), _p(th))
        return ids, th

    def trace_ordered(self, tris, rays, rule, window_log2=ORDERED_WINDOW_LOG2):
        """Model of the library's near-child-first closest-hit walk
        (ordered_model.cpp): rule 0 = plain walk, rule 1 = the checked walk. Returns ids, t and per-ray flags (bit 0 stack overflow, bit 1
        check fired; either means the ray was walked again in preorder)."""
        t = np.ascontiguousarray(tris, np.float32).reshape(-1, 9)
        r = np.ascontiguousarray(rays, np.float32).reshape(-1, 9)
        ids = np.zeros(len(r), np.int32)
        th = np.zeros(len(r), np.float32)
        flags = np.zeros(len(r), np.uint8)
        lib().orc_trace_ordered(self.h, _p(t), _p(r), len(r), int(rule), int(window_log2), _p(ids), _p(th),
                                _p(flags))
        return ids, th, flags

    def overlap_tree(self, qboxes, qids, self_mode, cap=1 << 22):
        q = np.ascontiguousarray(qboxes, np.float32).reshape(-1, 6)
        qi = np.ascontiguousarray(qids, np.int32)
        pairs = np.zeros((cap, 2), np.int32)
        count = lib().orc_overlap_tree(self.h, _p(q), _p(qi), len(q), int(self_mode), _p(pairs), cap)
        return pairs[: min(count, cap)], int(count)


def morton_codes(boxes, root):
    b = np.ascontiguousarray(boxes, np.float32).reshape(-1, 6)
    r = np.ascontiguousarray(root, np.float32)
    out = np.zeros(len(b), np.uint32)
    lib().orc_morton_codes(_p(b), len(b), _p(r), _p(out))
    return out


def brute_closest(tris, rays):
    t = np.ascontiguousarray(tris, np.float32).reshape(-1, 9)
    r = np.ascontiguousarray(rays, np.float32).reshape(-1, 9)
    ids = np.zeros(len(r), np.int32)
    th = np.zeros(len(r), np.float32)
    lib().orc_brute_closest(_p(t), len(t), _p(r), len(r), _p(ids), _p(th))
    return ids, th


def tri_t_fp64(tris, tri_ids, rays):
    t = np.ascontiguousarray(tris, np.float32).reshape(-1, 9)
    r = np.ascontiguousarray(rays, np.float32).reshape(-1, 9)
    ti = np.ascontiguousarray(tri_ids, np.int32)
    out = np.zeros(len(r), np.float64)
    lib().orc_tri_t_fp64(_p(t), _p(ti), _p(r), len(r), _p(out))
    return out


def overlap_self_brute(boxes, ids=None, cap=1 << 22):
    b = np.ascontiguousarray(boxes, np.float32).reshape(-1, 6)
    i = np.arange(len(b), dtype=np.int32) if ids is None else np.ascontiguousarray(ids, np.int32)
    pairs = np.zeros((cap, 2), np.int32)
    count = lib().orc_overlap_self_brute(_p(b), _p(i), len(b), _p(pairs), cap)
    return pairs[: min(count, cap)], int(count)


def overlap_cross_brute(lboxes, lids, rboxes, rids, cap=1 << 22):
    lb = np.ascontiguousarray(lboxes, np.float32).reshape(-1, 6)
    rb = np.ascontiguousarray(rboxes, np.float32).reshape(-1, 6)
    li = np.ascontiguousarray(lids, np.int32)
    ri = np.ascontiguousarray(rids, np.int32)
    pairs = np.zeros((cap, 2), np.int32)
    count = lib().orc_overlap_cross_brute(_p(lb), _p(li), len(lb), _p(rb), _p(ri), len(rb), _p(pairs), cap)
    return pairs[: min(count, cap)], int(count)


def triangle_bounds(tris):
    t = np.ascontiguousarray(tris, np.float32).reshape(-1, 9)
    out = np.zeros((len(t), 6), np.float32)
    lib().orc_triangle_bounds(_p(t), len(t), _p(out))
    return out


# ---- the reference's own host code (oracle/_ref) -------------------------------
def ref_available():
    return REF_LIB.exists()


def ref():
    global _ref
    if _ref is None:
        if not REF_LIB.exists():
            raise ImportError(f"{REF_LIB} missing (built only where the reference sources exist)")
        L = ctypes.CDLL(str(REF_LIB))
        L.ref_lbvh_build.restype = c_void_p
        L.ref_lbvh_build.argtypes = [c_void_p, c_void_p, c_int32]
        L.ref_lbvh_free.argtypes = [c_void_p]
        L.ref_lbvh_node_count.restype = c_int32
        L.ref_lbvh_node_count.argtypes = [c_void_p]
        L.ref_lbvh_export.argtypes = [c_void_p] + [c_void_p] * 6
        L.ref_morton_codes.argtypes = [c_void_p, c_int32, c_void_p, c_void_p]
        L.ref_trace_closest.restype = c_int32
        L.ref_trace_closest.argtypes = [c_void_p, c_void_p, c_void_p, c_int32] + [c_void_p] * 4
        L.ref_brute_closest.argtypes = [c_void_p, c_int32, c_void_p, c_int32, c_void_p, c_void_p]
        L.ref_triangle_bounds.argtypes = [c_void_p, c_int32, c_void_p]
        L.ref_overlap_self_brute.restype = c_int64
        L.ref_overlap_self_brute.argtypes = [c_void_p, c_int32, c_void_p, c_int64]
        L.ref_overlap_cross_brute.restype = c_int64
        L.ref_overlap_cross_brute.argtypes = [c_void_p, c_int32, c_void_p, c_int32, c_void_p, c_int64]
        _ref = L
    return _ref


class RefLbvh:
    """build_cpu_lbvh_tree + intersect_cpu_bvh_tree of the reference itself."""

    def __init__(self, boxes, ids=None):
        self.boxes = np.ascontiguousarray(boxes, np.float32).reshape(-1, 6)
        n = len(self.boxes)
        self.ids = np.arange(n, dtype=np.int32) if ids is None else np.ascontiguousarray(ids, np.int32)
        self.h = ref().ref_lbvh_build(_p(self.boxes), _p(self.ids), n)
        m = ref().ref_lbvh_node_count(self.h)
        self.bounds = np.zeros((m, 6), np.float32)
        self.left = np.zeros(m, np.int32)
        self.right = np.zeros(m, np.int32)
        self.first = np.zeros(m, np.int32)
        self.count = np.zeros(m, np.int32)
        self.leaf_ids = np.zeros(n, np.int32)
        if m:
            ref().ref_lbvh_export(self.h, _p(self.bounds), _p(self.left), _p(self.right), _p(self.first),
                                  _p(self.count), _p(self.leaf_ids))

    def __del__(self):
        try:
            ref().ref_lbvh_free(self.h)
        except Exception:
            pass

    def trace_closest(self, tris, rays):
        t = np.ascontiguousarray(tris, np.float32).reshape(-1, 9)
        r = np.ascontiguousarray(rays, np.float32).reshape(-1, 9)
        k = len(r)
        ids = np.zeros(k, np.int32)
        th = np.zeros(k, np.float32)
        u = np.zeros(k, np.float32)
        v = np.zeros(k, np.float32)
        ov = ref().ref_trace_closest(self.h, _p(t), _p(r), k, _p(ids), _p(th), _p(u), _p(v))
        return {"ids": ids, "t": th, "u": u, "v": v, "overflows": int(ov)}


# ---- the reference's own CPU path tracer (oracle/_ref) --------------------------
def _ref_pt_lib():
    L = ref()
    if not hasattr(L, "_pt_ready"):
        L.ref_pt_create.restype = c_void_p
        L.ref_pt_create.argtypes = [c_void_p, c_int32, c_void_p, c_void_p, c_void_p, c_void_p, c_int32, c_void_p,
                                    c_int32, c_void_p]
        L.ref_pt_free.argtypes = [c_void_p]
        L.ref_pt_camera.argtypes = [c_void_p, c_void_p]
        L.ref_pt_lights.restype = c_int32
        L.ref_pt_lights.argtypes = [c_void_p] + [c_void_p] * 8
        L.ref_pt_render.argtypes = [c_void_p, c_int32, c_int32, c_int32, c_int32, ctypes.c_uint64, c_int32, c_int32,
                                    c_int32, c_void_p]
        L.ref_pt_samples.argtypes = [c_void_p, c_int32, c_int32, c_int32, ctypes.c_uint64, c_int32, c_int32,
                                     c_int32, c_void_p, c_void_p, c_int32, c_void_p]
        L.ref_pt_light_bvh.restype = c_int32
        L.ref_pt_light_bvh.argtypes = [c_void_p] * 6
        L.ref_pt_create_instanced.restype = c_void_p
        L.ref_pt_create_instanced.argtypes = [c_void_p] * 7 + [c_int32, c_void_p, c_void_p, c_void_p, c_int32,
                                                               c_void_p, c_int32, c_void_p]
        L.ref_pt_set_scene_lights.argtypes = [c_int32] + [c_void_p] * 5
        L.ref_pt_create_builtin.restype = c_void_p
        L.ref_pt_create_builtin.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_float]
        L.ref_pt_export.argtypes = [c_void_p] * 14
        L.ref_pt_create_bunny.restype = c_void_p
        L.ref_pt_create_bunny.argtypes = [ctypes.c_char_p, ctypes.c_float]
        L.ref_pt_triangles.restype = c_int32
        L.ref_pt_triangles.argtypes = [c_void_p] * 4
        L._pt_ready = True
    return L


def _scene_args(scene):
    v = np.ascontiguousarray(scene["vertices"], np.float32)
    i = np.ascontiguousarray(scene["indices"], np.uint32)
    n = np.ascontiguousarray(scene["normals"], np.float32)
    uv = np.ascontiguousarray(scene["uvs"], np.float32)
    m = np.ascontiguousarray(scene["material_ids"], np.int32)
    mats = np.ascontiguousarray(scene["materials"], np.float32)
    cam = np.ascontiguousarray(scene["camera"], np.float32)
    return (v, i, n, uv, m, mats, cam)


LIGHT_KINDS = {"point": 1, "distant": 2, "uniform_infinite": 3, "image_infinite": 4}


def scene_light_arrays(lights):
    """Scene-light dicts ({"kind", "vector", "color", "image"}) as the arrays the
    shim takes: kinds, vectors (n, 3), colours (n, 3), image sizes (n, 2), texels."""
    n = max(len(lights), 1)
    kind = np.zeros(n, np.int32)
    vec = np.zeros((n, 3), np.float32)
    col = np.zeros((n, 3), np.float32)
    wh = np.zeros((n, 2), np.int32)
    px = []
    for k, x in enumerate(lights):
        kind[k] = LIGHT_KINDS[x["kind"]]
        vec[k] = x.get("vector", (0, 0, 0))
        col[k] = x["color"]
        if x["kind"] == "image_infinite":
            img = np.asarray(x["image"], np.float32)
            wh[k] = (img.shape[1], img.shape[0])
            px.append(img.reshape(-1, 3))
    texels = np.ascontiguousarray(np.concatenate(px) if px else np.zeros((1, 3), np.float32), np.float32)
    return kind, vec, col, wh, texels


class RefPathTracer:
    """render_cpu_path_tracer / trace_path of the reference on a scene dict (an
    instanced one when the dict has "inst_mesh": meshes + affine instances,
    built with add_scene_mesh / add_scene_instance(make_transform))."""

    def __init__(self, scene):
        L = _ref_pt_lib()
        self._keep = _scene_args(scene)
        v, i, n, uv, m, mats, cam = self._keep
        lights = scene.get("lights") or []
        self._lights = scene_light_arrays(lights)
        L.ref_pt_set_scene_lights(len(lights), *[_p(a) for a in self._lights])
        if "inst_mesh" in scene:
            vo = np.ascontiguousarray(scene["vertex_offsets"], np.int32)
            to = np.ascontiguousarray(scene["triangle_offsets"], np.int32)
            im = np.ascontiguousarray(scene["inst_mesh"], np.int32)
            mm = np.ascontiguousarray(scene["inst_m"], np.float32).reshape(-1, 16)
            ov = scene.get("inst_override")
            ov = np.full(len(im), -1, np.int32) if ov is None else np.ascontiguousarray(ov, np.int32)
            self._inst = (vo, to, im, mm, ov)
            self.h = L.ref_pt_create_instanced(_p(v), _p(vo), _p(i), _p(to), _p(n), _p(uv), _p(m), len(vo) - 1,
                                               _p(im), _p(mm), _p(ov), len(im), _p(mats), len(mats), _p(cam))
            self.ordinals = L.ref_pt_triangles(self.h, None, None, None)
        else:
            self.h = L.ref_pt_create(_p(v), len(v), _p(i), _p(n), _p(uv), _p(m), len(m), _p(mats), len(mats),
                                     _p(cam))
            self.ordinals = len(m)

    @classmethod
    def bunny(cls, asset_root, width, height):
        """The reference's own "bunny" scene (make_stanford_bunny_scene with the
        CLI camera), loaded from `asset_root`/models/stanford-bunny.obj."""
        L = _ref_pt_lib()
        self = cls.__new__(cls)
        self._keep = None
        self.h = L.ref_pt_create_bunny(str(asset_root).encode(), float(np.float32(width) / np.float32(height)))
        self.ordinals = L.ref_pt_triangles(self.h, None, None, None)
        return self

    @classmethod
    def builtin(cls, name, asset_root, width, height):
        """One of the reference CLI's built-in scenes ("direct-light",
        "transparent-filter", "bunny", "lux-cover") with its CLI camera."""
        L = _ref_pt_lib()
        self = cls.__new__(cls)
        self._keep = None
        L.ref_pt_set_scene_lights(0, None, None, None, None, None)
        self.h = L.ref_pt_create_builtin(name.encode(), str(asset_root).encode(),
                                         float(np.float32(width) / np.float32(height)))
        if not self.h:
            raise ValueError(f"unknown built-in scene {name}")
        self.ordinals = L.ref_pt_triangles(self.h, None, None, None)
        return self

    def export(self):
        """The scene as an instanced scene dict (meshes, instances with their
        Transform m and inv, materials, built camera vectors)."""
        L = _ref_pt_lib()
        c = np.zeros(5, np.int32)
        L.ref_pt_export(self.h, _p(c), *([None] * 12))
        nv, nt, nm, ni, nmat = map(int, c)
        d = {"vertices": np.zeros((nv, 3), np.float32), "normals": np.zeros((nv, 3), np.float32),
             "vertex_offsets": np.zeros(nm + 1, np.int32), "indices": np.zeros((nt, 3), np.uint32),
             "material_ids": np.zeros(nt, np.int32), "triangle_offsets": np.zeros(nm + 1, np.int32),
             "inst_mesh": np.zeros(ni, np.int32), "inst_m": np.zeros((ni, 4, 4), np.float32),
             "inst_inv": np.zeros((ni, 4, 4), np.float32), "inst_override": np.zeros(ni, np.int32),
             "materials": np.zeros((nmat, 11), np.float32), "camera_vectors": np.zeros(12, np.float32)}
        L.ref_pt_export(self.h, _p(c), *[_p(v) for v in d.values()])
        d["uvs"] = np.zeros((nv, 2), np.float32)
        d["camera"] = np.zeros(8, np.float32)
        return d

    def triangles(self):
        """scene_triangle_view per scene ordinal: world corners (T, 9), world
        corner normals (T, 9), material after the instance override (T,)."""
        t = self.ordinals
        out = {"corners": np.zeros((t, 9), np.float32), "normals": np.zeros((t, 9), np.float32),
               "material": np.zeros(t, np.int32)}
        _ref_pt_lib().ref_pt_triangles(self.h, _p(out["corners"]), _p(out["normals"]), _p(out["material"]))
        return out

    def __del__(self):
        try:
            _ref_pt_lib().ref_pt_free(self.h)
        except Exception:
            pass

    def camera(self):
        out = np.zeros(20, np.float32)
        _ref_pt_lib().ref_pt_camera(self.h, _p(out))
        return out

    def lights(self):
        L = _ref_pt_lib()
        n = L.ref_pt_lights(self.h, None, None, None, None, None, None, None, None)
        t = self.ordinals
        out = {"prim": np.zeros(n, np.int32), "area": np.zeros(n, np.float32), "cdf": np.zeros(n, np.float32),
               "prob": np.zeros(n, np.float32), "alias": np.zeros(n, np.int32), "pmf": np.zeros(n, np.float32),
               "primitive_light_index": np.zeros(t, np.int32), "total_weight": np.zeros(1, np.float32)}
        L.ref_pt_lights(self.h, _p(out["prim"]), _p(out["area"]), _p(out["cdf"]), _p(out["prob"]), _p(out["alias"]),
                        _p(out["pmf"]), _p(out["primitive_light_index"]), _p(out["total_weight"]))
        return out

    # LightSamplerKind values (src/light/light_distribution.h:27-31)
    LIGHT_SAMPLERS = {"uniform": 0, "power": 1, "bvh": 2}

    def render(self, width, height, spp, depth, seed=0, frame=0, sample_offset=0, light_sampler="power"):
        film = np.zeros((height, width, 3), np.float32)
        _ref_pt_lib().ref_pt_render(self.h, width, height, spp, depth, seed, frame, sample_offset,
                                    self.LIGHT_SAMPLERS[light_sampler], _p(film))
        return film

    def samples(self, width, height, depth, pixels, samples, seed=0, frame=0, sample_offset=0,
                light_sampler="power"):
        px = np.ascontiguousarray(pixels, np.int32)
        sm = np.ascontiguousarray(samples, np.int32)
        out = np.zeros((len(px), 3), np.float32)
        _ref_pt_lib().ref_pt_samples(self.h, width, height, depth, seed, frame, sample_offset,
                                     self.LIGHT_SAMPLERS[light_sampler], _p(px), _p(sm), len(px), _p(out))
        return out

    def light_bvh(self):
        """The reference's LightBvhTable: bounds (N, 6), power (N,), links (N, 4:
        left, right, parent, light), leaf node per light, root (-1 if invalid)."""
        L = _ref_pt_lib()
        n = L.ref_pt_light_bvh(self.h, None, None, None, None, None)
        nl = len(self.lights()["prim"])
        out = {"bounds": np.zeros((n, 6), np.float32), "power": np.zeros(n, np.float32),
               "links": np.zeros((n, 4), np.int32), "leaf": np.zeros(nl if n else 0, np.int32)}
        root = np.zeros(1, np.int32)
        if n:
            L.ref_pt_light_bvh(self.h, _p(out["bounds"]), _p(out["power"]), _p(out["links"]), _p(out["leaf"]),
                               _p(root))
        out["root"] = int(root[0]) if n else -1
        return out


# ---- the reference's own OBJ loader (oracle/_ref) --------------------------------
def ref_load_obj(path, normal_mode=0, reverse_orientation=False, flip_texcoord_v=False, boundary_topology=0,
                 position_weld_epsilon=1e-6, default_material_id=-1, scene_material_count=0, scale_mode=0,
                 target_extent=1.0, center_xz=False, ground_to_y=False, target_y=0.0, object_to_world=None,
                 material_bindings=None, cap_boundary_holes=False, cap_boundary_topology=1):
    """load_obj_into_scene_with_result into a scene holding
    `scene_material_count` materials: ok flag, the appended mesh arrays, the
    created instance's transform (m, inv) and the boundary-capping counts."""
    L = ref()
    if not hasattr(L, "_obj_ready"):
        L.ref_obj_load.restype = c_void_p
        L.ref_obj_load.argtypes = [ctypes.c_char_p, c_int32, c_int32, c_int32, c_int32, ctypes.c_float, c_int32,
                                   c_int32, c_int32, ctypes.c_float, c_int32, c_int32, ctypes.c_float, c_void_p,
                                   c_int32, c_void_p, c_void_p]
        L.ref_obj_free.argtypes = [c_void_p]
        L.ref_obj_counts.argtypes = [c_void_p, c_void_p]
        L.ref_obj_arrays.argtypes = [c_void_p] * 9
        L.ref_obj_set_cap.argtypes = [c_int32, c_int32]
        L.ref_obj_cap_counts.argtypes = [c_void_p, c_void_p]
        L._obj_ready = True
    L.ref_obj_set_cap(int(cap_boundary_holes), int(cap_boundary_topology))
    m = None if object_to_world is None else np.ascontiguousarray(object_to_world, np.float32).reshape(16)
    b = material_bindings or {}
    names = (ctypes.c_char_p * max(len(b), 1))(*[k.encode() for k in b])
    ids = np.ascontiguousarray(list(b.values()) or [0], np.int32)
    h = L.ref_obj_load(str(path).encode(), normal_mode, int(reverse_orientation), int(flip_texcoord_v),
                       boundary_topology, position_weld_epsilon, default_material_id, scene_material_count,
                       scale_mode, target_extent, int(center_xz), int(ground_to_y), target_y,
                       None if m is None else _p(m), len(b), ctypes.cast(names, c_void_p), _p(ids))
    try:
        c = np.zeros(7, np.int32)
        L.ref_obj_counts(h, _p(c))
        nv, nt = int(c[1]), int(c[2])
        out = {"vertices": np.zeros((nv, 3), np.float32), "normals": np.zeros((nv, 3), np.float32),
               "uvs": np.zeros((nv, 2), np.float32), "indices": np.zeros((nt, 3), np.uint32),
               "material_ids": np.zeros(nt, np.int32), "boundary_edges": np.zeros(nt, np.uint8),
               "object_to_world": np.zeros((4, 4), np.float32), "world_to_object": np.zeros((4, 4), np.float32)}
        L.ref_obj_arrays(h, *[_p(v) for v in out.values()])
        out.update(ok=bool(c[0]), appended_default_material=int(c[3]), skipped_invalid=int(c[4]),
                   skipped_nonfinite=int(c[5]), skipped_degenerate=int(c[6]))
        cc = np.zeros(6, np.int32)
        L.ref_obj_cap_counts(h, _p(cc))
        out.update(zip(("cap_triangles", "cap_boundary_edges", "cap_non_simple_vertices", "cap_incomplete_loops",
                        "cap_degenerate_loops", "cap_degenerate_triangles"), (int(x) for x in cc)))
        return out
    finally:
        L.ref_obj_free(h)


# ---- two-level restatement (two_level_oracle.cpp) ------------------------------
def _tl_lib():
    L = lib()
    if not hasattr(L, "_tl_ready"):
        L.orc_inverse_affine.argtypes = [c_void_p, c_void_p]
        L.orc_transform_bounds.argtypes = [c_void_p, c_void_p, c_void_p]
        L.orc_instance_transform_valid.restype = c_int32
        L.orc_instance_transform_valid.argtypes = [c_void_p, c_void_p]
        L.orc_tl_build.restype = c_void_p
        L.orc_tl_build.argtypes = [c_void_p, c_void_p, c_int32, c_void_p, c_void_p, c_void_p, c_int32]
        L.orc_tl_free.argtypes = [c_void_p]
        L.orc_tl_first_primitive.restype = c_int32
        L.orc_tl_first_primitive.argtypes = [c_void_p, c_int32]
        L.orc_tl_trace.restype = c_int32
        L.orc_tl_trace.argtypes = [c_void_p, c_void_p, c_void_p, c_int32] + [c_void_p] * 6
        L._tl_ready = True
    return L


def inverse_affine(m):
    a = np.ascontiguousarray(m, np.float32).reshape(16)
    out = np.zeros(16, np.float32)
    _tl_lib().orc_inverse_affine(_p(a), _p(out))
    return out


def transform_bounds(m, box):
    a = np.ascontiguousarray(m, np.float32).reshape(16)
    b = np.ascontiguousarray(box, np.float32).reshape(6)
    out = np.zeros(6, np.float32)
    _tl_lib().orc_transform_bounds(_p(a), _p(b), _p(out))
    return out


def instance_transform_valid(m, inv):
    a = np.ascontiguousarray(m, np.float32).reshape(16)
    b = np.ascontiguousarray(inv, np.float32).reshape(16)
    return bool(_tl_lib().orc_instance_transform_valid(_p(a), _p(b)))


def mesh_soups(vertices, vert_off, indices, tri_off):
    """Object-space triangle soups (9 floats each) of indexed meshes, concatenated."""
    V = np.asarray(vertices, np.float32).reshape(-1, 3)
    I = np.asarray(indices, np.int64).reshape(-1, 3)
    out = []
    for m in range(len(vert_off) - 1):
        out.append(V[vert_off[m] + I[tri_off[m]:tri_off[m + 1]]].reshape(-1, 9))
    return np.ascontiguousarray(np.concatenate(out) if out else np.zeros((0, 9), np.float32), np.float32)


class OracleTwoLevel:
    """BLAS per mesh + TLAS over instance world bounds and the host two-level walk."""

    def __init__(self, tris, tri_off, inst_mesh, inst_inv, inst_world_bounds):
        self._keep = (np.ascontiguousarray(tris, np.float32).reshape(-1, 9), np.ascontiguousarray(tri_off, np.int32),
                      np.ascontiguousarray(inst_mesh, np.int32),
                      np.ascontiguousarray(inst_inv, np.float32).reshape(-1, 16),
                      np.ascontiguousarray(inst_world_bounds, np.float32).reshape(-1, 6))
        t, to, im, inv, wb = self._keep
        self.h = _tl_lib().orc_tl_build(_p(t), _p(to), len(to) - 1, _p(im), _p(inv), _p(wb), len(im))

    def __del__(self):
        try:
            _tl_lib().orc_tl_free(self.h)
        except Exception:
            pass

    def first_primitive(self):
        return np.array([_tl_lib().orc_tl_first_primitive(self.h, i) for i in range(len(self._keep[2]))], np.int32)

    def trace(self, rays, ignore=None):
        r = np.ascontiguousarray(rays, np.float32).reshape(-1, 9)
        k = len(r)
        ig = None if ignore is None else np.ascontiguousarray(ignore, np.int32)
        out = {"prim": np.zeros(k, np.int32), "t": np.zeros(k, np.float32), "u": np.zeros(k, np.float32),
               "v": np.zeros(k, np.float32), "instance": np.zeros(k, np.int32), "local": np.zeros(k, np.int32)}
        ov = _tl_lib().orc_tl_trace(self.h, _p(r), _p(ig), k, _p(out["prim"]), _p(out["t"]), _p(out["u"]),
                                    _p(out["v"]), _p(out["instance"]), _p(out["local"]))
        out["overflows"] = int(ov)
        return out


# ---- the reference's own two-level scene walk (oracle/_ref) ----------------------
def _ref_tl_lib():
    L = ref()
    if not hasattr(L, "_tl_ready"):
        L.ref_tl_create.restype = c_void_p
        L.ref_tl_create.argtypes = [c_void_p, c_void_p, c_void_p, c_void_p, c_int32, c_void_p, c_void_p, c_void_p,
                                    c_int32]
        L.ref_tl_free.argtypes = [c_void_p]
        L.ref_tl_instances.argtypes = [c_void_p, c_void_p, c_void_p, c_void_p]
        L.ref_tl_trace.restype = c_int32
        L.ref_tl_trace.argtypes = [c_void_p, c_void_p, c_void_p, c_int32] + [c_void_p] * 6
        L._tl_ready = True
    return L


def two_level_arrays(meshes):
    """meshes: list of (vertices (v, 3), indices (t, 3)) -> flat arrays + offsets."""
    verts = [np.asarray(v, np.float32).reshape(-1, 3) for v, _ in meshes]
    idx = [np.asarray(i, np.uint32).reshape(-1, 3) for _, i in meshes]
    vo = np.zeros(len(meshes) + 1, np.int32)
    to = np.zeros(len(meshes) + 1, np.int32)
    vo[1:] = np.cumsum([len(v) for v in verts])
    to[1:] = np.cumsum([len(i) for i in idx])
    V = np.ascontiguousarray(np.concatenate(verts) if verts else np.zeros((0, 3), np.float32), np.float32)
    I = np.ascontiguousarray(np.concatenate(idx) if idx else np.zeros((0, 3), np.uint32), np.uint32)
    return V, vo, I, to


class RefTwoLevel:
    """The reference's two-level scene: add_scene_mesh / add_scene_instance
    (make_transform) + PreparedScene::host_bvh + intersect_scene_bvh."""

    def __init__(self, meshes, inst_mesh, inst_m, inst_override=None):
        L = _ref_tl_lib()
        V, vo, I, to = two_level_arrays(meshes)
        self.inst_mesh = np.ascontiguousarray(inst_mesh, np.int32)
        self.inst_m = np.ascontiguousarray(inst_m, np.float32).reshape(-1, 16)
        n = len(self.inst_mesh)
        self.inst_override = (np.full(n, -1, np.int32) if inst_override is None
                              else np.ascontiguousarray(inst_override, np.int32))
        self._keep = (V, vo, I, to)
        self.h = L.ref_tl_create(_p(V), _p(vo), _p(I), _p(to), len(meshes), _p(self.inst_mesh), _p(self.inst_m),
                                 _p(self.inst_override), n)

    def __del__(self):
        try:
            _ref_tl_lib().ref_tl_free(self.h)
        except Exception:
            pass

    def instances(self):
        n = len(self.inst_mesh)
        inv = np.zeros((n, 16), np.float32)
        wb = np.zeros((n, 6), np.float32)
        first = np.zeros(n, np.int32)
        _ref_tl_lib().ref_tl_instances(self.h, _p(inv), _p(wb), _p(first))
        return {"inv": inv, "world_bounds": wb, "first_primitive": first}

    def trace(self, rays, ignore=None):
        r = np.ascontiguousarray(rays, np.float32).reshape(-1, 9)
        k = len(r)
        ig = None if ignore is None else np.ascontiguousarray(ignore, np.int32)
        out = {"prim": np.zeros(k, np.int32), "t": np.zeros(k, np.float32), "u": np.zeros(k, np.float32),
               "v": np.zeros(k, np.float32), "instance": np.zeros(k, np.int32), "local": np.zeros(k, np.int32)}
        ov = _ref_tl_lib().ref_tl_trace(self.h, _p(r), _p(ig), k, _p(out["prim"]), _p(out["t"]), _p(out["u"]),
                                        _p(out["v"]), _p(out["instance"]), _p(out["local"]))
        out["overflows"] = int(ov)
        return out


# ---- film writers restatement (film_oracle.cpp) ---------------------------------
def film_file(kind, film, spp):
    """The bytes of write_ppm (kind "ppm") or write_pfm ("pfm") of a film of
    per-pixel sums (H, W, 3), restated."""
    L = lib()
    f = np.ascontiguousarray(film, np.float32)
    h, w = f.shape[:2]
    fn = L.orc_ppm_file if kind == "ppm" else L.orc_pfm_file
    fn.restype = c_int64
    fn.argtypes = [c_void_p, c_int32, c_int32, c_int32, c_void_p, c_int64]
    n = fn(_p(f), w, h, spp, None, 0)
    out = np.zeros(n, np.uint8)
    fn(_p(f), w, h, spp, _p(out), n)
    return out


def ppm_table_mismatches(t255, threads=8):
    L = lib()
    L.orc_ppm_table_mismatches.restype = c_int64
    L.orc_ppm_table_mismatches.argtypes = [c_void_p, c_int32]
    t = np.ascontiguousarray(t255, np.float32)
    assert len(t) == 255
    return int(L.orc_ppm_table_mismatches(_p(t), threads))


# ---- the reference's own film writers (oracle/_ref) ------------------------------
def ref_write_film(kind, path, film, spp):
    """write_ppm (kind "ppm") or write_pfm ("pfm") of the reference on a film of
    per-pixel sums (H, W, 3)."""
    L = ref()
    f = np.ascontiguousarray(film, np.float32)
    h, w = f.shape[:2]
    fn = L.ref_write_ppm if kind == "ppm" else L.ref_write_pfm
    fn.argtypes = [ctypes.c_char_p, c_void_p, c_int32, c_int32, c_int32]
    fn(str(path).encode(), _p(f), w, h, spp)


# ---- path-tracer restatement (pt_oracle.cpp) ----------------------------------
def _pt_lib():
    L = lib()
    if not hasattr(L, "_pt_ready"):
        L.orc_pt_create.restype = c_void_p
        L.orc_pt_create.argtypes = [c_void_p, c_int32, c_void_p, c_void_p, c_void_p, c_int32, c_void_p, c_int32,
                                    c_void_p]
        L.orc_pt_create_instanced.restype = c_void_p
        L.orc_pt_create_instanced.argtypes = [c_void_p] * 6 + [c_int32] + [c_void_p] * 4 + [c_int32, c_void_p,
                                                                                              c_int32, c_void_p]
        L.orc_pt_triangle_count.restype = c_int32
        L.orc_pt_triangle_count.argtypes = [c_void_p]
        L.orc_pt_free.argtypes = [c_void_p]
        L.orc_pt_camera.argtypes = [c_void_p, c_void_p]
        L.orc_pt_lights.restype = c_int32
        L.orc_pt_lights.argtypes = [c_void_p] + [c_void_p] * 8
        L.orc_pt_render.argtypes = [c_void_p, c_int32, c_int32, c_int32, c_int32, ctypes.c_uint64, c_int32, c_int32,
                                    c_int32, c_int32, c_int32, c_void_p]
        L.orc_pt_samples.argtypes = [c_void_p, c_int32, c_int32, c_int32, ctypes.c_uint64, c_int32, c_int32,
                                     c_void_p, c_void_p, c_int32, c_void_p]
        L._pt_ready = True
    return L


class OraclePathTracer:
    """The CPU restatement of the reference path tracer on a scene dict (an
    instanced one when the dict has "inst_mesh", as RefPathTracer takes it;
    "inst_inv" is used where given, else the inverse_affine of "inst_m")."""

    def __init__(self, scene):
        L = _pt_lib()
        self._keep = _scene_args(scene)
        v, i, n, uv, m, mats, cam = self._keep
        if "inst_mesh" in scene:
            vo = np.ascontiguousarray(scene["vertex_offsets"], np.int32)
            to = np.ascontiguousarray(scene["triangle_offsets"], np.int32)
            im = np.ascontiguousarray(scene["inst_mesh"], np.int32)
            mm = np.ascontiguousarray(scene["inst_m"], np.float32).reshape(-1, 16)
            inv = scene.get("inst_inv")
            inv = None if inv is None else np.ascontiguousarray(inv, np.float32).reshape(-1, 16)
            ov = scene.get("inst_override")
            ov = np.full(len(im), -1, np.int32) if ov is None else np.ascontiguousarray(ov, np.int32)
            self._inst = (vo, to, im, mm, inv, ov)
            self.h = L.orc_pt_create_instanced(_p(v), _p(vo), _p(i), _p(to), _p(n), _p(m), len(vo) - 1, _p(im),
                                               _p(mm), _p(inv), _p(ov), len(im), _p(mats), len(mats), _p(cam))
        else:
            self.h = L.orc_pt_create(_p(v), len(v), _p(i), _p(n), _p(m), len(m), _p(mats), len(mats), _p(cam))
        self.ordinals = L.orc_pt_triangle_count(self.h)

    def __del__(self):
        try:
            _pt_lib().orc_pt_free(self.h)
        except Exception:
            pass

    def camera(self):
        out = np.zeros(20, np.float32)
        _pt_lib().orc_pt_camera(self.h, _p(out))
        return out

    def lights(self):
        L = _pt_lib()
        n = L.orc_pt_lights(self.h, None, None, None, None, None, None, None, None)
        t = self.ordinals
        out = {"prim": np.zeros(n, np.int32), "area": np.zeros(n, np.float32), "cdf": np.zeros(n, np.float32),
               "prob": np.zeros(n, np.float32), "alias": np.zeros(n, np.int32), "pmf": np.zeros(n, np.float32),
               "primitive_light_index": np.zeros(t, np.int32), "total_weight": np.zeros(1, np.float32)}
        L.orc_pt_lights(self.h, _p(out["prim"]), _p(out["area"]), _p(out["cdf"]), _p(out["prob"]), _p(out["alias"]),
                        _p(out["pmf"]), _p(out["primitive_light_index"]), _p(out["total_weight"]))
        return out

    def render(self, width, height, spp, depth, seed=0, frame=0, sample_offset=0, rows=None, threads=1):
        film = np.zeros((height, width, 3), np.float32)
        r0, r1 = (0, height) if rows is None else rows
        _pt_lib().orc_pt_render(self.h, width, height, spp, depth, seed, frame, sample_offset, r0, r1, threads,
                                _p(film))
        return film

    def samples(self, width, height, depth, pixels, samples, seed=0, frame=0, sample_offset=0):
        px = np.ascontiguousarray(pixels, np.int32)
        sm = np.ascontiguousarray(samples, np.int32)
        out = np.zeros((len(px), 3), np.float32)
        _pt_lib().orc_pt_samples(self.h, width, height, depth, seed, frame, sample_offset, _p(px), _p(sm), len(px),
                                 _p(out))
        return out


def pcg32(seed, seq, n):
    L = lib()
    L.orc_pcg32.argtypes = [ctypes.c_uint64, ctypes.c_uint64, c_int32, c_void_p]
    out = np.zeros(n, np.uint32)
    L.orc_pcg32(seed, seq, n, _p(out))
    return out

"""numpy-facing mirror of the reference LBVH interface over the C ABI.

Names, argument meaning and error behaviour follow src/bvh/gpu_bvh.cuh:20-61 and
src/bvh/bvh_overlap.cuh:10-23:

* ``build_gpu_lbvh(boxes, ids)`` -> :class:`GpuBvh` (invalid with count 0 for
  empty input, src/bvh/gpu_bvh.cu:537);
* ``refit_gpu_lbvh(bvh, boxes, ids)`` -> bool (False on count/id mismatch,
  src/bvh/gpu_bvh.cu:597-605);
* ``intersect_gpu_bvh_stackless(bvh, triangles, rays)`` -> hits
  ``{primitive_id, t}`` or ``{-1, inf}`` (src/bvh/gpu_bvh.cu:893-915);
* ``query_aabb_overlaps_self/cross`` -> :class:`OverlapQueryResult` with
  ``pairs``, ``found_count``, ``truncated`` (src/bvh/bvh_overlap.cu:186-251).

Every call runs on the GPU through liblux_hip.so; there is no host fallback.
"""

import ctypes
from dataclasses import dataclass

import numpy as np

from . import abi
from .abi import AABB_DTYPE, HIT_DTYPE, NODE_DTYPE, PAIR_DTYPE, RAY_DTYPE, TRI_DTYPE, call, ptr

_initialized = False


def init(device=0):
    global _initialized
    if not _initialized:
        call("lux_hip_init", device)
        _initialized = True


def as_boxes(boxes):
    b = np.ascontiguousarray(boxes, dtype=np.float32).reshape(-1, 6)
    return b


def as_tris(tris):
    return np.ascontiguousarray(tris, dtype=np.float32).reshape(-1, 9)


def as_rays(rays):
    return np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 9)


class DeviceArray:
    """A device allocation made through lux_hip_malloc (owning)."""

    def __init__(self, nbytes):
        init()
        self.nbytes = int(nbytes)
        p = ctypes.c_void_p()
        call("lux_hip_malloc", ctypes.byref(p), ctypes.c_size_t(max(self.nbytes, 1)))
        self.ptr = p

    @classmethod
    def from_numpy(cls, array):
        a = np.ascontiguousarray(array)
        d = cls(a.nbytes)
        if a.nbytes:
            call("lux_hip_memcpy_htod", d.ptr, ptr(a), ctypes.c_size_t(a.nbytes))
        return d

    def to_numpy(self, dtype, shape):
        out = np.empty(shape, dtype=dtype)
        if out.nbytes:
            assert out.nbytes <= self.nbytes
            call("lux_hip_memcpy_dtoh", ptr(out), self.ptr, ctypes.c_size_t(out.nbytes))
        return out

    def free(self):
        if self.ptr:
            call("lux_hip_free", self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            if self.ptr:
                abi.load().lux_hip_free(self.ptr)
        except Exception:
            pass


class GpuBvh:
    """Owner of one device LBVH (GpuBvh, src/bvh/gpu_bvh.cuh:25-47)."""

    def __init__(self):
        init()
        h = ctypes.c_void_p()
        call("lux_hip_bvh_create", ctypes.byref(h))
        self.handle = h

    def close(self):
        if self.handle:
            call("lux_hip_bvh_destroy", self.handle)
            self.handle = ctypes.c_void_p()

    def __del__(self):
        try:
            if self.handle:
                abi.load().lux_hip_bvh_destroy(self.handle)
        except Exception:
            pass

    # -- build / refit ------------------------------------------------------
    def build(self, boxes, ids):
        b = as_boxes(boxes)
        i = np.ascontiguousarray(ids, dtype=np.int32)
        assert len(b) == len(i)
        call("lux_hip_bvh_build", self.handle, ptr(b), ptr(i), len(b))
        return self

    def build_triangles_device(self, d_tris, n, stream=None):
        """Build from a device triangle soup: box i = triangle_bounds(tri i), id i."""
        call("lux_hip_bvh_build_triangles_device", self.handle, d_tris.ptr, int(n), stream)
        return self

    def build_device(self, d_boxes, d_ids, n, stream=None):
        call("lux_hip_bvh_build_device", self.handle, d_boxes.ptr, d_ids.ptr, int(n), stream)
        return self

    def refit(self, boxes, ids):
        b = as_boxes(boxes)
        i = np.ascontiguousarray(ids, dtype=np.int32)
        try:
            call("lux_hip_bvh_refit", self.handle, ptr(b), ptr(i), len(b))
        except abi.LuxError as e:
            if e.status == abi.LUX_ERR_MISMATCH:
                return False
            raise
        return True

    def refit_device(self, d_boxes, n, stream=None):
        call("lux_hip_bvh_refit_device", self.handle, d_boxes.ptr, int(n), stream)

    def refit_triangles_device(self, d_tris, n, stream=None):
        call("lux_hip_bvh_refit_triangles_device", self.handle, d_tris.ptr, int(n), stream)

    def set_ray_order(self, enable):
        """Trace large ray sets in origin order (default on); results are identical."""
        call("lux_hip_bvh_set_ray_order", self.handle, int(bool(enable)))

    # -- state --------------------------------------------------------------
    def info(self):
        valid = ctypes.c_int32()
        count = ctypes.c_int32()
        root = abi.Aabb()
        call("lux_hip_bvh_info", self.handle, ctypes.byref(valid), ctypes.byref(count), ctypes.byref(root))
        rb = np.array([root.p_min.x, root.p_min.y, root.p_min.z, root.p_max.x, root.p_max.y, root.p_max.z],
                      dtype=np.float32)
        return bool(valid.value), int(count.value), rb

    def is_valid(self):
        return self.info()[0]

    @property
    def primitive_count(self):
        return self.info()[1]

    @property
    def root_bounds(self):
        return self.info()[2]

    def nodes(self):
        _, n, _ = self.info()
        out = np.zeros(max(2 * n - 1, 0), dtype=NODE_DTYPE)
        if n:
            call("lux_hip_bvh_download_nodes", self.handle, ptr(out), len(out))
        return out

    def topology(self):
        """Radix-tree arrays (see lux_hip_bvh_download_topology)."""
        _, n, _ = self.info()
        m = max(n - 1, 0)
        t = {
            "morton_sorted": np.zeros(n, np.uint32),
            "sorted_indices": np.zeros(n, np.uint32),
            "left": np.zeros(m, np.uint32),
            "right": np.zeros(m, np.uint32),
            "parent": np.zeros(m, np.uint32),
            "leaf_parent": np.zeros(n, np.uint32),
            "node_bounds": np.zeros((max(2 * n - 1, 0), 6), np.float32),
        }
        if n:
            call("lux_hip_bvh_download_topology", self.handle, ptr(t["morton_sorted"]), ptr(t["sorted_indices"]),
                 ptr(t["left"]), ptr(t["right"]), ptr(t["parent"]), ptr(t["leaf_parent"]),
                 ptr(t["node_bounds"]))
        return t

    def input(self):
        """(boxes (n, 6), primitive ids (n,)) the tree holds, in input order."""
        _, n, _ = self.info()
        boxes = np.zeros((n, 6), np.float32)
        ids = np.zeros(n, np.int32)
        if n:
            call("lux_hip_bvh_download_input", self.handle, ptr(boxes), ptr(ids))
        return boxes, ids

    def device_arrays(self):
        """Device addresses of the input boxes, ids and sorted input indices
        (GpuBvh::overlap_*_device_ptr); None for an invalid tree."""
        b, i, s = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
        call("lux_hip_bvh_device_arrays", self.handle, ctypes.byref(b), ctypes.byref(i), ctypes.byref(s))
        return b.value, i.value, s.value


def build_gpu_lbvh(boxes, ids=None):
    b = as_boxes(boxes)
    if ids is None:
        ids = np.arange(len(b), dtype=np.int32)
    return GpuBvh().build(b, ids)


def refit_gpu_lbvh(bvh, boxes, ids=None):
    b = as_boxes(boxes)
    if ids is None:
        ids = np.arange(len(b), dtype=np.int32)
    return bvh.refit(b, ids)


def intersect_gpu_bvh_stackless(bvh, triangles, rays):
    t = as_tris(triangles)
    r = as_rays(rays)
    hits = np.empty(len(r), dtype=HIT_DTYPE)
    call("lux_hip_trace_closest", bvh.handle, ptr(t), len(t), ptr(r), len(r), ptr(hits))
    return hits


def trace_any(bvh, triangles, rays):
    t = as_tris(triangles)
    r = as_rays(rays)
    occ = np.empty(len(r), dtype=np.uint8)
    call("lux_hip_trace_any", bvh.handle, ptr(t), len(t), ptr(r), len(r), ptr(occ))
    return occ


@dataclass
class OverlapQueryResult:
    pairs: np.ndarray  # (k, 2) int32
    found_count: int
    truncated: bool


def _overlap(func, handles, max_pairs):
    cap = max(int(max_pairs), 0)
    pairs = np.zeros((max(cap, 1), 2), dtype=np.int32)
    found = ctypes.c_int32()
    call(func, *handles, int(max_pairs), ptr(pairs), ctypes.byref(found))
    k = min(found.value, cap)
    return OverlapQueryResult(pairs[:k].copy(), int(found.value), bool(found.value > cap))


def query_aabb_overlaps_self(bvh, max_pairs=1 << 20):
    return _overlap("lux_hip_overlap_self", [bvh.handle], max_pairs)


def query_aabb_overlaps_cross(left, right, max_pairs=1 << 20):
    return _overlap("lux_hip_overlap_cross", [left.handle, right.handle], max_pairs)


class Probe:
    """Sum of HIP-event kernel times for launches of one named kernel."""

    def __init__(self, kernel_name):
        self.kernel_name = kernel_name
        self.total_ms = 0.0
        self.launches = 0

    def __enter__(self):
        init()
        call("lux_hip_probe_begin", self.kernel_name.encode())
        return self

    def __exit__(self, *exc):
        ms = ctypes.c_double()
        n = ctypes.c_int64()
        call("lux_hip_probe_end", ctypes.byref(ms), ctypes.byref(n))
        self.total_ms = ms.value
        self.launches = n.value
        return False


def synchronize():
    call("lux_hip_synchronize")

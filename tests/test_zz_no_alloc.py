"""GPU: lux_hip_trace_closest_device_noalloc (ray-order scratch in a fixed
device array) against lux_hip_trace_closest_device (the same query with its
scratch grown by hipMalloc), through the C ABI. Hit ids and distances must be
bit-identical, on two calls in a row (the second smaller, so stale scratch from
the first would show) and on a call small enough to skip the ray sort."""

import numpy as np
import pytest

from lux_amd import abi, workloads as W

pytestmark = pytest.mark.gpu


def _closest(gpu, func, bvh, d_tris, tri_count, rays):
    d_rays = gpu.DeviceArray.from_numpy(rays)
    d_hits = gpu.DeviceArray(len(rays) * 8)
    rc = abi.load()[func](bvh.handle, d_tris.ptr, tri_count, d_rays.ptr, len(rays), d_hits.ptr, None)
    assert rc == 0, abi.load().lux_hip_last_error()
    out = d_hits.to_numpy(np.uint32, (len(rays), 2))
    d_rays.free()
    d_hits.free()
    return out


def test_noalloc_closest_equals_device_closest(gpu):
    tris = W.soup(1 << 15, seed=71)
    bvh = gpu.build_gpu_lbvh(W.triangle_boxes(tris))
    d_tris = gpu.DeviceArray.from_numpy(tris)
    # 2^17 + 3 rays, then 70001 (both sorted through the scratch), then 1000
    for count, seed in ((131075, 72), (70001, 73), (1000, 74)):
        print(f"case: no-alloc closest, {count} rays, seed {seed}", flush=True)
        rays = W.rays(count, seed=seed)
        ref = _closest(gpu, "lux_hip_trace_closest_device", bvh, d_tris, len(tris), rays)
        got = _closest(gpu, "lux_hip_trace_closest_device_noalloc", bvh, d_tris, len(tris), rays)
        assert np.array_equal(got, ref)
        assert (got[:, 0].view(np.int32) >= 0).sum() > 0
    d_tris.free()


def test_noalloc_rejects_more_rays_than_the_scratch_holds(gpu):
    """4194305 rays: refused before any memory is touched (null arrays passed)."""
    print("case: no-alloc size limit", flush=True)
    tris = W.soup(64, seed=75)
    bvh = gpu.build_gpu_lbvh(W.triangle_boxes(tris))
    L = abi.load()
    rc = L.lux_hip_trace_closest_device_noalloc(bvh.handle, None, 64, None, (1 << 22) + 1, None, None)
    assert rc == 1
    assert b"4194304" in L.lux_hip_last_error()

"""Autotuner of the wavefront schedule's launch geometry.

The knobs are the path-slot capacity of a batch (the size of every ray and
shadow queue), the number of workgroups of the grid-stride kernels and the
workgroup size. The tuner renders a calibration mix -- one large single-scene
image and a handful of multi-scene jobs drawn from the batched mix -- with every
candidate, keeps the fastest and stores it per device in lux_amd/tuning.json.
`apply()` loads the stored choice for the current device (render.Scene and the
batch entry points call it once); with no entry the library defaults stand.

    python -m lux_amd.autotune            # tune on this GPU and store the result
"""

import ctypes
import itertools
import json
import sys
import time

from . import PACKAGE_DIR

TUNING_FILE = PACKAGE_DIR / "tuning.json"
# The first tuning on gfx950/256CU chose the largest capacity and the smallest
# grid of {2^22..2^25} x {1024, 2048, 4096} x {128, 256, 512} (block 128 was
# last everywhere), so the grid now reaches past both edges of that choice. A
# capacity above the calibration mix's ~38 M paths runs it in one batch, so
# 2^26 is the last capacity that can differ.
CANDIDATES = {
    "capacity": [1 << 23, 1 << 24, 1 << 25, 1 << 26],
    "grid": [512, 768, 1024, 2048],
    "block": [256, 512],
}
_applied = False


def device_key():
    """'<arch>/<compute units>CU' of the library's device."""
    from . import abi, lbvh

    lbvh.init()
    name = ctypes.create_string_buffer(256)
    arch = ctypes.create_string_buffer(256)
    cus = ctypes.c_int32()
    mem = ctypes.c_uint64()
    abi.call("lux_hip_device_info", name, 256, arch, 256, ctypes.byref(cus), ctypes.byref(mem))
    return f"{arch.value.decode().split(':')[0]}/{cus.value}CU"


def load():
    if not TUNING_FILE.exists():
        return {}
    return json.loads(TUNING_FILE.read_text())


def apply(force=False):
    """Set the stored tuning of this device (once per process)."""
    global _applied
    if _applied and not force:
        return None
    _applied = True
    entry = load().get(device_key())
    if not entry:
        return None
    from . import render

    render.set_tuning(capacity=entry["capacity"], grid=entry["grid"], block=entry["block"], use_graph=1)
    return entry


def calibration_mix():
    """One 1024x1024 32-spp image of a 216k-triangle displaced icosphere and six
    batched-mix scenes at 8 spp (about 38 M paths, more than the largest
    capacity tried, so every candidate runs several batches)."""
    from . import scenes

    jobs = [(scenes.displaced_icosphere(freq=104, width=1024, height=1024),
             dict(width=1024, height=1024, spp=32, depth=8, seed=0))]
    for spec in scenes.batched_mix(6, seed=0x4C5570):
        d = scenes.displaced_icosphere(freq=spec["freq"], seed=spec["seed"], width=spec["width"],
                                       height=spec["height"])
        jobs.append((d, dict(width=spec["width"], height=spec["height"], spp=8, depth=6, seed=0)))
    return jobs


def tune(repeats=2, write=True, log=print, path=None):
    from . import lbvh, render

    mix = calibration_mix()
    scs = [render.Scene(d) for d, _ in mix]
    jobs = [j for _, j in mix]
    films = [lbvh.DeviceArray(j["width"] * j["height"] * 12) for j in jobs]
    results = []
    for cap, grid, block in itertools.product(CANDIDATES["capacity"], CANDIDATES["grid"], CANDIDATES["block"]):
        render.set_tuning(capacity=cap, grid=grid, block=block, use_graph=1)
        render.batch_render_device(scs, [dict(j, spp=1) for j in jobs], films)  # capture + warm
        lbvh.synchronize()
        best = float("inf")
        for _ in range(repeats):
            t0 = time.perf_counter()
            render.batch_render_device(scs, jobs, films)
            lbvh.synchronize()
            best = min(best, time.perf_counter() - t0)
        results.append({"capacity": cap, "grid": grid, "block": block, "ms": round(best * 1e3, 3)})
        log(f"capacity {cap:>8} grid {grid:>5} block {block:>3}: {best * 1e3:8.2f} ms")
    win = min(results, key=lambda r: r["ms"])
    entry = {**win, "use_graph": 1, "candidates": results, "tuned_at": time.strftime("%Y-%m-%d %H:%M:%S")}
    if write:
        table = load()
        table[device_key()] = entry
        (path or TUNING_FILE).write_text(json.dumps(table, indent=1) + "\n")
    render.set_tuning(capacity=win["capacity"], grid=win["grid"], block=win["block"], use_graph=1)
    for f in films:
        f.free()
    for sc in scs:
        sc.close()
    return entry


if __name__ == "__main__":
    from pathlib import Path

    out = Path(sys.argv[sys.argv.index("--out") + 1]) if "--out" in sys.argv else None
    e = tune(write="--dry-run" not in sys.argv, path=out)
    print(json.dumps({k: v for k, v in e.items() if k != "candidates"}))

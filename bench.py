#!/usr/bin/env python3
"""Benchmark of the LUX hot path on one MI355X.

Workload (one "step"): the LBVH microbenchmark on a 1M-triangle soup -- vertices
uniform in [0,1]^3 with edge lengths ~U(0.001, 0.01), FP32:
  1. LBVH build of 2^20 triangle boxes (30-bit Morton, radix sort, Karras, boxes)
  2. closest-hit of 2^22 rays (origins uniform in the cube, directions uniform on
     the sphere)
  3. any-hit (occlusion) of the same 2^22 rays
  4. LBVH build of 100 000 query boxes + their cross overlap query against the
     1M-triangle tree
  5. refit of the 1M tree from vertex-jittered triangles (sigma = 1e-3)
All inputs are resident in HBM before the timed region. `value` is the rays
traced per second over the whole step (closest + any, i.e. 2 * 2^22 rays per
step, with build, overlap and refit time included).

Prints ONE JSON line. `roofline` is for the dominant kernel (closest-hit),
timed with HIP events on its own stream inside the timed steps. `cpu_baseline`
times the reference's own host traversal (intersect_cpu_bvh_tree over its host
LBVH, compiled from the LUX sources into oracle/_ref) on a bounded ray sample,
single thread, like the reference renderer.
"""

import argparse
import ctypes
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "Mrays/s; batched renders/s + p50/p99 latency vs plain GPU path at equal RMSE"
HBM_PEAK_GBPS = 8000.0
TRAFFIC_FILE = ROOT / "profiles" / "traffic.json"


def parse():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--triangles", type=int, default=1 << 20)
    ap.add_argument("--rays", type=int, default=1 << 22)
    ap.add_argument("--queries", type=int, default=100_000)
    ap.add_argument("--cpu-sample-rays", type=int, default=1 << 19)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="only run the plain copy kernel over this many MiB (HBM counter calibration)")
    return ap.parse_args()


def cpu_baseline(inp, sample_rays):
    """The reference's own host LBVH build + stack traversal, one thread."""
    import oracle

    boxes, tris, rays = inp["boxes"], inp["tris"], inp["rays"]
    sample = rays[:sample_rays]
    kind = "reference" if oracle.ref_available() else "port"
    t0 = time.perf_counter()
    tree = oracle.RefLbvh(boxes) if kind == "reference" else oracle.OracleBvh(boxes)
    t1 = time.perf_counter()
    tree.trace_closest(tris, sample)
    t2 = time.perf_counter()
    return {
        "value": round(len(sample) / (t2 - t1) / 1e6, 4),
        "unit": "Mrays/s",
        "cores": 1,
        "kind": kind,
        "sample": (f"closest-hit of the first {len(sample)} of the {len(rays)} rays over the host LBVH of the "
                   f"same {len(boxes)}-triangle soup (intersect_cpu_bvh_tree, 1 thread); host LBVH build "
                   f"{len(boxes) / (t1 - t0) / 1e6:.2f} Mprims/s"),
        "build_mprims_per_s": round(len(boxes) / (t1 - t0) / 1e6, 4),
        "host_cpu": _cpu_name(),
    }


def _cpu_name():
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip() + f" ({os.cpu_count()} logical cpus visible)"
    except OSError:
        pass
    return None


def main():
    args = parse()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    from lux_amd import lbvh, workloads as W
    from lux_amd.abi import call

    lbvh.init(0)
    if args.copy_calibration_mib:
        nbytes = args.copy_calibration_mib << 20
        src, dst = lbvh.DeviceArray(nbytes), lbvh.DeviceArray(nbytes)
        for _ in range(args.steps):
            call("lux_hip_copy_device", dst.ptr, src.ptr, ctypes.c_size_t(nbytes), None)
        lbvh.synchronize()
        print(json.dumps({"copy_calibration_bytes": nbytes, "launches": args.steps}), flush=True)
        return
    N, R, Q = args.triangles, args.rays, args.queries
    inp = W.c2_inputs(n_tris=N, n_rays=R, n_queries=Q)
    up = lbvh.DeviceArray.from_numpy
    d_boxes, d_ids = up(inp["boxes"]), up(np.arange(N, dtype=np.int32))
    d_tris, d_jit, d_rays = up(inp["tris"]), up(inp["tris_jittered"]), up(inp["rays"])
    d_qboxes, d_qids = up(inp["query_boxes"]), up(np.arange(Q, dtype=np.int32))
    d_hits, d_occ = lbvh.DeviceArray(8 * R), lbvh.DeviceArray(R)
    pair_cap = 1 << 22
    d_pairs, d_found = lbvh.DeviceArray(8 * pair_cap), lbvh.DeviceArray(4)
    bvh, qbvh = lbvh.GpuBvh(), lbvh.GpuBvh()

    def build():
        bvh.build_device(d_boxes, d_ids, N)

    def closest():
        call("lux_hip_trace_closest_device", bvh.handle, d_tris.ptr, N, d_rays.ptr, R, d_hits.ptr, None)

    def anyhit():
        call("lux_hip_trace_any_device", bvh.handle, d_tris.ptr, N, d_rays.ptr, R, d_occ.ptr, None)

    def overlap():
        qbvh.build_device(d_qboxes, d_qids, Q)
        call("lux_hip_overlap_cross_device", bvh.handle, qbvh.handle, pair_cap, d_pairs.ptr, d_found.ptr, None)

    def refit():
        bvh.refit_triangles_device(d_jit, N)

    stages = [("build", build), ("closest", closest), ("any", anyhit), ("overlap", overlap), ("refit", refit)]

    def step():
        for _, f in stages:
            f()

    for _ in range(args.warmup):
        step()
    lbvh.synchronize()
    bvh.info()  # surfaces any device-side check failure before timing

    with lbvh.Probe("trace_closest") as probe:
        lbvh.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            step()
        lbvh.synchronize()
        t1 = time.perf_counter()
    bvh.info()
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    kernel_ms = probe.total_ms / max(probe.launches, 1)

    # sanity of the last step's results (parity itself is the test suite's job)
    hits = d_hits.to_numpy(np.int32, (R, 2))
    occ = d_occ.to_numpy(np.uint8, R)
    assert np.array_equal(occ.astype(bool), hits[:, 0] >= 0), "any-hit disagrees with closest-hit"
    found = int(d_found.to_numpy(np.int32, 1)[0])

    # per-stage wall time (same work, stage by stage, synchronised)
    stage_ms = {}
    for name, f in stages:
        if name == "closest" or name == "any" or name == "overlap":
            build()
        lbvh.synchronize()
        s0 = time.perf_counter()
        for _ in range(args.steps):
            f()
        lbvh.synchronize()
        stage_ms[name] = round((time.perf_counter() - s0) * 1e3 / args.steps, 4)
    build()
    stats = (ctypes.c_uint64 * 4)()
    call("lux_hip_trace_stats_device", bvh.handle, d_tris.ptr, N, d_rays.ptr, R, stats)

    algo_bytes = 40 * R  # per ray: o, d, t_min, t_max in (32 B) + {id, t} out (8 B)
    achieved = algo_bytes / (kernel_ms * 1e-3) / 1e9
    touched = 32 * stats[0] + 36 * stats[1]
    traffic = None
    if TRAFFIC_FILE.exists():
        t = json.loads(TRAFFIC_FILE.read_text()).get("trace_closest")
        if t and t.get("rays") == R and t.get("triangles") == N:
            traffic = t.get("bytes_per_launch")

    rays_per_step = 2 * R
    out = {
        "metric": METRIC,
        "value": round(rays_per_step / (ms_per_step * 1e-3) / 1e6, 3),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": (f"LBVH microbench: {N}-triangle soup (U[0,1]^3, edges U(0.001,0.01)), build + "
                         f"closest-hit and any-hit of {R} rays + cross overlap of {Q} query boxes + refit "
                         f"after sigma=1e-3 vertex jitter"),
            "triangles": N,
            "rays_per_step": rays_per_step,
            "queries": Q,
            "parallelism": "single-gpu",
        },
        "roofline": {
            "kernel": "trace_closest",
            "bound": "hbm",
            "achieved": round(achieved, 3),
            "peak": HBM_PEAK_GBPS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBPS, 6),
            "traffic": traffic,
            "kernel_ms": round(kernel_ms, 4),
            "algorithmic_bytes_per_launch": algo_bytes,
            "touched_bytes_per_launch": int(touched),
            "touched_GBps": round(touched / (kernel_ms * 1e-3) / 1e9, 2),
            "node_visits_per_ray": round(stats[0] / R, 2),
            "triangle_tests_per_ray": round(stats[1] / R, 3),
        },
        "stages_ms": stage_ms,
        "stage_rates": {
            "build_Mprims_per_s": round(N / stage_ms["build"] / 1e3, 2),
            "refit_Mprims_per_s": round(N / stage_ms["refit"] / 1e3, 2),
            "closest_Mrays_per_s": round(R / stage_ms["closest"] / 1e3, 2),
            "any_Mrays_per_s": round(R / stage_ms["any"] / 1e3, 2),
            "overlap_Mqueries_per_s": round(Q / stage_ms["overlap"] / 1e3, 2),
            "overlap_pairs": found,
            "hit_fraction": round(float((hits[:, 0] >= 0).mean()), 4),
        },
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(inp, min(args.cpu_sample_rays, R))
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

#!/usr/bin/env python3
"""Benchmark of the LUX hot path on one MI355X.

Workload (one "step"): the LBVH microbenchmark on a 1M-triangle soup -- vertices
uniform in [0,1]^3 with edge lengths ~U(0.001, 0.01), FP32:
  1. LBVH build of 2^20 triangle boxes (30-bit Morton, radix sort, Karras, boxes)
  2. closest-hit of 2^22 rays (origins uniform in the cube, directions uniform on
     the sphere)
  3. any-hit (occlusion) of the same 2^22 rays
  4. LBVH build of 100 000 query boxes + their cross overlap query against the
     1M-triangle tree, and the self overlap query of the 1M-triangle tree
  5. refit of the 1M tree from vertex-jittered triangles (sigma = 1e-3)
All inputs are resident in HBM before the timed region. `value` is the rays
traced per second over the whole step (closest + any, i.e. 2 * 2^22 rays per
step, with build, overlap and refit time included).

Prints ONE JSON line. `roofline` is for the step's longest kernel (any hit,
k_trace_any), timed with HIP events on its own stream inside the timed steps:
its HBM bound (compulsory bytes per ray over the kernel time, against the HBM
peak and against the bandwidth the plain copy kernel attains on the same box,
and the bytes the counters see below L2), and beside it its VALU bound from the committed
counter summary (profiles/counters.json): how much of the kernel's time the
vector ALUs issue, and how many lanes an issued instruction carries. The
closest-hit kernel's figures follow under `closest_hit`. `cpu_baseline` times
the CPU restatement of the reference's host traversal (intersect_cpu_bvh_tree
over its host LBVH; bit-identical to the reference on the golden vectors) on a
bounded ray sample, single thread, like the reference renderer.
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
COUNTERS_FILE = ROOT / "profiles" / "counters.json"


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
    ap.add_argument("--render", default="c1,c3,instanced,bunny,c4,c5",
                    help="render workloads to add to the line: any of c1,c3,instanced,bunny,c4,c5 (comma-separated) "
                         "or none")
    ap.add_argument("--two-level", default="on", choices=["on", "off"],
                    help="add the instanced (two-level TLAS/BLAS) workload to the line")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain-schedule baselines of c3/c4")
    ap.add_argument("--no-graph", action="store_true",
                    help="launch the wavefront bounce loop directly instead of replaying a captured hipGraph")
    ap.add_argument("--c4-scenes", type=int, default=64)
    ap.add_argument("--c5-requests", type=int, default=96)
    ap.add_argument("--c5-load", type=float, default=0.75, help="arrival rate as a fraction of the c4 capacity")
    ap.add_argument("--cpu-render-rows", type=int, default=32,
                    help="rows of the C3 image the CPU render baseline traces (from row 128, the lit band)")
    return ap.parse_args()


def cpu_baseline(inp, sample_rays):
    """The host LBVH build + stack traversal of the reference, as restated by
    the oracle (bit-identical to the reference on the golden vectors), one
    thread."""
    import oracle

    boxes, tris, rays = inp["boxes"], inp["tris"], inp["rays"]
    sample = rays[:sample_rays]
    t0 = time.perf_counter()
    tree = oracle.OracleBvh(boxes)
    t1 = time.perf_counter()
    tree.trace_closest(tris, sample)
    t2 = time.perf_counter()
    return {
        "value": round(len(sample) / (t2 - t1) / 1e6, 4),
        "unit": "Mrays/s",
        "cores": 1,
        "kind": "port",
        "sample": (f"closest-hit of the first {len(sample)} of the {len(rays)} rays over the host LBVH of the "
                   f"same {len(boxes)}-triangle soup (restated build_cpu_lbvh_tree + intersect_cpu_bvh_tree, "
                   f"1 thread); host LBVH build {len(boxes) / (t1 - t0) / 1e6:.2f} Mprims/s"),
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


# ---- render workloads ---------------------------------------------------------------
# Algorithmic HBM bytes of the wavefront path tracer: per extension segment 336 B
# (ray 32 r + hit 16 w + 16 r + path state 152 r/w + surface fetch 88 + next ray
# 32 w), per shadow ray 112 B (record 2 x 44 + radiance read-modify-write 24),
# 12 B per film pixel.
EXT_BYTES, SHADOW_BYTES, FILM_BYTES = 336, 112, 12


def _log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def _render_rate(ext, sh, pixels, seconds):
    algo = EXT_BYTES * ext + SHADOW_BYTES * sh + FILM_BYTES * pixels
    gbps = algo / seconds / 1e9
    return {"Mrays_per_s": round((ext + sh) / seconds / 1e6, 2), "extension_rays": int(ext), "shadow_rays": int(sh),
            "algorithmic_GB": round(algo / 1e9, 3), "achieved_GBps": round(gbps, 2),
            "hbm_frac": round(gbps / HBM_PEAK_GBPS, 5)}


def bench_c1(args):
    """The CLI's Cornell render (`-scene cornell -size 256 256 -spp 64 -depth 4
    -light-sampler power`, seed 0): GPU wavefront and plain renders (films
    bit-identical), and the CPU restatement of the reference's trace_path on a
    band of 16 rows, one thread and all host threads, with the whole image's
    CPU time projected from the one-thread rate."""
    from lux_amd import lbvh, render, scenes

    W = H = 256
    spp, depth = 64, 4
    d = scenes.cornell(W, H)
    sc = render.Scene(d)
    film = lbvh.DeviceArray(W * H * 12)
    out = {"workload": f"Cornell box ({sc.triangle_count} triangles), {W}x{H}, {spp} spp, depth {depth}"}
    films = {}
    for name, mode in (("wavefront", render.WAVEFRONT), ("plain", render.PLAIN)):
        sc.render_device(film, W, H, 1, depth, mode=mode)
        lbvh.synchronize()
        best = None
        for _ in range(3):
            s0 = time.perf_counter()
            sc.render_device(film, W, H, spp, depth, mode=mode)
            lbvh.synchronize()
            sec = time.perf_counter() - s0
            best = sec if best is None else min(best, sec)
        ext, sh = sc.rays()
        films[name] = film.to_numpy(np.float32, (H, W, 3))
        out[name] = {"render_s": round(best, 5), "Mpaths_per_s": round(W * H * spp / best / 1e6, 2),
                     **_render_rate(ext, sh, W * H, best)}
    out["speedup_vs_plain"] = round(out["plain"]["render_s"] / out["wavefront"]["render_s"], 3)
    out["films_bit_identical"] = bool(np.array_equal(films["wavefront"].view(np.uint32),
                                                     films["plain"].view(np.uint32)))
    _log(f"c1: {out['wavefront']['render_s'] * 1e3:.2f} ms wavefront, {out['plain']['render_s'] * 1e3:.2f} ms plain")
    if args.cpu_render_rows > 0:
        out["cpu_baseline"] = cpu_render_baseline(d, W, H, spp, depth, films["wavefront"], 16, 120, sc=sc)
        out["cpu_baseline"]["projected_image_cpu_s"] = round(W * H * spp / (out["cpu_baseline"]["value"] * 1e6), 2)
    sc.close()
    film.free()
    return out


def bench_c3(args):
    """Single-scene render: displaced icosphere (freq 208, 865 280 triangles),
    1024x1024, 64 spp, depth 8, seed 0."""
    from lux_amd import lbvh, render, scenes

    W = H = 1024
    spp, depth = 64, 8
    t0 = time.perf_counter()
    d = scenes.displaced_icosphere(freq=208, width=W, height=H)
    t1 = time.perf_counter()
    sc = render.Scene(d)
    t2 = time.perf_counter()
    film = lbvh.DeviceArray(W * H * 12)
    out = {"workload": f"displaced icosphere, {sc.triangle_count} triangles, {W}x{H}, {spp} spp, depth {depth}",
           "scene_generate_s": round(t1 - t0, 3), "scene_upload_build_s": round(t2 - t1, 3)}
    modes = [("wavefront", render.WAVEFRONT)] + ([] if args.no_plain else [("plain", render.PLAIN)])
    films = {}
    for name, mode in modes:
        sc.render_device(film, W, H, 1, depth, mode=mode)
        lbvh.synchronize()
        s0 = time.perf_counter()
        sc.render_device(film, W, H, spp, depth, mode=mode)
        lbvh.synchronize()
        sec = time.perf_counter() - s0
        ext, sh = sc.rays()
        films[name] = film.to_numpy(np.float32, (H, W, 3))
        out[name] = {"render_s": round(sec, 4), "renders_per_s": round(1 / sec, 4),
                     **_render_rate(ext, sh, W * H, sec)}
        _log(f"c3 {name}: {sec:.3f} s")
    if "plain" in out:
        out["speedup_vs_plain"] = round(out["plain"]["render_s"] / out["wavefront"]["render_s"], 3)
        out["films_bit_identical"] = bool(np.array_equal(films["wavefront"].view(np.uint32),
                                                         films["plain"].view(np.uint32)))
    if args.cpu_render_rows > 0:
        # rows from 128: the lit top of the sphere (the rows below the equator
        # see no light)
        out["cpu_baseline"] = cpu_render_baseline(d, W, H, spp, depth, films["wavefront"], args.cpu_render_rows,
                                                  128, sc=sc)
    sc.close()
    return out


def bench_instanced(args):
    """Instanced-scene render: 64 instances of 4 displaced icospheres (81 920
    triangles each, 5 242 880 instanced triangles) over a ground plane,
    1024x1024, 16 spp, depth 8, seed 0; the same geometry flattened into one
    mesh (the instanced scene's own world-space shading view) rendered through
    the one-level path beside it. CPU baseline: the CPU restatement of the
    reference's trace_path (two-level walk; bit-identical to the reference on
    the golden instanced renders) on a band of rows, one thread."""
    from lux_amd import lbvh, render, scenes

    W = H = 1024
    spp, depth = 16, 8
    t0 = time.perf_counter()
    d = scenes.instanced_field(W, H)
    t1 = time.perf_counter()
    sc = render.Scene(d)
    t2 = time.perf_counter()
    flat = scenes.flattened(sc.triangles(), d["materials"], d["camera"])
    t3 = time.perf_counter()
    fsc = render.Scene(flat)
    t4 = time.perf_counter()
    film = lbvh.DeviceArray(W * H * 12)
    out = {"workload": (f"{len(d['inst_mesh']) - 1} instances of {len(d['triangle_offsets']) - 2} displaced icospheres "
                        f"+ ground and light, {sc.triangle_count} instanced triangles, {W}x{H}, {spp} spp, "
                        f"depth {depth}"),
           "scene_generate_s": round(t1 - t0, 3), "instanced_upload_build_s": round(t2 - t1, 3),
           "flat_upload_build_s": round(t4 - t3, 3),
           "instanced_device_bytes": sc.info()[2], "flat_device_bytes": fsc.info()[2]}
    films = {}
    for name, s_ in (("instanced", sc), ("flattened", fsc)):
        s_.render_device(film, W, H, 1, depth)
        lbvh.synchronize()
        best = None
        for _ in range(3):
            s0 = time.perf_counter()
            s_.render_device(film, W, H, spp, depth)
            lbvh.synchronize()
            sec = time.perf_counter() - s0
            best = sec if best is None else min(best, sec)
        ext, sh = s_.rays()
        films[name] = film.to_numpy(np.float32, (H, W, 3))
        out[name] = {"render_s": round(best, 4), "Mpaths_per_s": round(W * H * spp / best / 1e6, 2),
                     **_render_rate(ext, sh, W * H, best)}
        _log(f"instanced render, {name}: {best:.3f} s")
    out["instanced_vs_flattened_time"] = round(out["instanced"]["render_s"] / out["flattened"]["render_s"], 3)
    out["instanced_vs_flattened_relrmse"] = round(render.rel_rmse(films["instanced"] / spp,
                                                                  films["flattened"] / spp), 6)
    if args.cpu_render_rows > 0:
        import oracle

        nrows = max(1, args.cpu_render_rows // 4)
        r0 = H // 2 - nrows // 2
        paths = W * nrows * spp
        port = oracle.OraclePathTracer(d)
        s0 = time.perf_counter()
        band = port.render(W, H, spp, depth, rows=(r0, r0 + nrows), threads=1)[r0:r0 + nrows].reshape(-1, 3)
        t_port = time.perf_counter() - s0
        gpu_band = films["instanced"][r0:r0 + nrows].reshape(-1, 3)
        out["cpu_baseline"] = {
            "value": round(paths / t_port / 1e6, 4), "unit": "Mpaths/s", "cores": 1, "kind": "port",
            "sample": f"rows {r0}..{r0 + nrows - 1} of the {W}x{H} image at {spp} spp ({paths} paths)",
            "host_cpu": _cpu_name(),
            "gpu_vs_cpu_band_relrmse": round(render.rel_rmse(gpu_band / spp, band / spp), 6),
            "sample_divergence": sample_divergence(sc, port, W, H, spp, depth, (r0, r0 + nrows))}
    sc.close()
    fsc.close()
    film.free()
    return out


def bench_bunny(args):
    """The reference CLI's bunny render (`-scene bunny -gpu` at its defaults:
    256x256, 64 spp, depth 4, seed 0): the Stanford bunny (69 451 triangles)
    imported from the OBJ fixture, placed by its instance over a floor under a
    quad light. Reported: OBJ import time, scene upload/build, render rate; CPU
    baseline: the CPU restatement of the reference's trace_path (bit-identical
    to the reference's bunny render) on a band of rows, one thread."""
    import lzma
    import tempfile

    from lux_amd import lbvh, obj, render, scenes

    W = H = 256
    spp, depth = 64, 4
    src = Path(__file__).resolve().parent / "tests" / "golden" / "obj" / "stanford-bunny.obj.xz"
    with tempfile.TemporaryDirectory() as tmp:
        path = Path(tmp) / "stanford-bunny.obj"
        path.write_bytes(lzma.decompress(src.read_bytes()))
        t0 = time.perf_counter()
        obj.load(path, scene_material_count=2, ground_to_y=True)
        t1 = time.perf_counter()
        d = scenes.bunny(path, W, H)
    t2 = time.perf_counter()
    sc = render.Scene(d)
    t3 = time.perf_counter()
    film = lbvh.DeviceArray(W * H * 12)
    sc.render_device(film, W, H, 1, depth)
    lbvh.synchronize()
    best = None
    for _ in range(5):
        s0 = time.perf_counter()
        sc.render_device(film, W, H, spp, depth)
        lbvh.synchronize()
        sec = time.perf_counter() - s0
        best = sec if best is None else min(best, sec)
    ext, sh = sc.rays()
    gpu = film.to_numpy(np.float32, (H, W, 3))
    out = {"workload": f"Stanford bunny ({sc.triangle_count} scene triangles), {W}x{H}, {spp} spp, depth {depth}",
           "obj_import_s": round(t1 - t0, 3), "scene_upload_build_s": round(t3 - t2, 3),
           "render_s": round(best, 4), "renders_per_s": round(1 / best, 2),
           "Mpaths_per_s": round(W * H * spp / best / 1e6, 2), **_render_rate(ext, sh, W * H, best)}
    _log(f"bunny render: {best * 1e3:.1f} ms")
    if args.cpu_render_rows > 0:
        import oracle

        nrows = max(1, args.cpu_render_rows // 2)
        r0 = H // 2 - nrows // 2
        paths = W * nrows * spp
        port = oracle.OraclePathTracer(d)
        s0 = time.perf_counter()
        band = port.render(W, H, spp, depth, rows=(r0, r0 + nrows), threads=1)[r0:r0 + nrows].reshape(-1, 3)
        t_port = time.perf_counter() - s0
        out["cpu_baseline"] = {
            "value": round(paths / t_port / 1e6, 4), "unit": "Mpaths/s", "cores": 1, "kind": "port",
            "sample": f"rows {r0}..{r0 + nrows - 1} of the {W}x{H} image at {spp} spp ({paths} paths)",
            "host_cpu": _cpu_name(),
            "gpu_vs_cpu_band_relrmse": round(
                render.rel_rmse(gpu[r0:r0 + nrows].reshape(-1, 3) / spp, band / spp), 8),
            "sample_divergence": sample_divergence(sc, port, W, H, spp, depth, (r0, r0 + nrows))}
    sc.close()
    film.free()
    return out


def sample_divergence(sc, pt, W, H, spp, depth, rows, n=4096, seed=0x4C55):
    """Per-sample agreement of the GPU path radiance with the CPU restatement's
    trace of the same path: n random (pixel, sample) pairs in `rows`, each GPU
    radiance from a 1-spp render at that sample index. Paths that take another
    bounce (device cosf/sinf/powf round a few inputs differently from the host
    libm) show up as samples off by more than 1e-4 relative."""
    rng = np.random.default_rng(seed)
    px = rng.integers(rows[0] * W, rows[1] * W, n).astype(np.int32)
    sm = rng.integers(0, spp, n).astype(np.int32)
    ref = pt.samples(W, H, depth, px, sm)
    got = np.zeros_like(ref)
    for k in np.unique(sm):
        one = sc.render(W, H, 1, depth, seed=0, sample_offset=int(k)).reshape(-1, 3)
        sel = sm == k
        got[sel] = one[px[sel]]
    close = np.all(np.abs(got - ref) <= 1e-4 * np.maximum(np.abs(ref), 1e-2), axis=1)
    return {"samples": n, "within_1e-4": round(float(close.mean()), 5), "diverged": int((~close).sum())}


def cpu_render_baseline(scene, W, H, spp, depth, gpu_film, nrows, r0, sc=None):
    """A band of `nrows` rows from row r0 on the CPU, by the CPU restatement of
    the reference path tracer (bit-identical to the reference's trace_path on
    the golden renders): one thread (`value`, as the reference renders) and all
    host threads (rows split across threads, per-pixel sample order unchanged)."""
    import oracle

    rows = (r0, r0 + nrows)
    band = slice(rows[0], rows[1])
    paths = W * nrows * spp
    out = {"unit": "Mpaths/s", "cores": 1, "kind": "port",
           "sample": f"rows {rows[0]}..{rows[1] - 1} of the {W}x{H} image at {spp} spp ({paths} paths)",
           "host_cpu": _cpu_name()}
    t0 = time.perf_counter()
    pt = oracle.OraclePathTracer(scene)
    out["port_scene_prepare_s"] = round(time.perf_counter() - t0, 3)
    s0 = time.perf_counter()
    port = pt.render(W, H, spp, depth, rows=rows, threads=1)
    t_port = time.perf_counter() - s0
    threads = os.cpu_count() or 1
    s0 = time.perf_counter()
    pt.render(W, H, spp, depth, rows=rows, threads=threads)
    t_all = time.perf_counter() - s0
    from lux_amd.render import rel_rmse

    out.update(value=round(paths / t_port / 1e6, 4), all_threads_Mpaths_per_s=round(paths / t_all / 1e6, 4),
               all_threads=threads,
               gpu_vs_cpu_band_relrmse=round(rel_rmse(gpu_film[band] / spp, port[band] / spp), 8))
    if sc is not None:
        out["sample_divergence"] = sample_divergence(sc, pt, W, H, spp, depth, rows)
    return out


def c4_scenes(count):
    from lux_amd import scenes

    specs = scenes.batched_mix(count)
    dicts = [scenes.displaced_icosphere(freq=s["freq"], seed=s["seed"], width=s["width"], height=s["height"])
             for s in specs]
    jobs = [dict(width=s["width"], height=s["height"], spp=s["spp"], depth=s["depth"], seed=0) for s in specs]
    return specs, dicts, jobs


def bench_c4(args, prepared=None):
    """Batched multi-scene launch vs the plain single-scene path, scene by scene."""
    from lux_amd import lbvh, render

    t0 = time.perf_counter()
    specs, dicts, jobs = prepared or c4_scenes(args.c4_scenes)
    t1 = time.perf_counter()
    scs = [render.Scene(d) for d in dicts]
    t2 = time.perf_counter()
    d_films = [lbvh.DeviceArray(j["width"] * j["height"] * 12) for j in jobs]
    warm = [dict(j, spp=1) for j in jobs]
    render.batch_render_device(scs, warm, d_films)
    lbvh.synchronize()
    s0 = time.perf_counter()
    render.batch_render_device(scs, jobs, d_films)
    lbvh.synchronize()
    sec = time.perf_counter() - s0
    ext, sh = render.batch_rays()
    pixels = sum(j["width"] * j["height"] for j in jobs)
    tris = sum(sc.triangle_count for sc in scs)
    paths = sum(j["width"] * j["height"] * j["spp"] for j in jobs)
    out = {"workload": (f"{len(scs)} displaced-icosphere scenes, {tris} triangles, {pixels} pixels, {paths} paths "
                        f"(triangles log-U[1e3,5e5], sides {{128..1024}}, spp log-U[16,256], depth 6)"),
           "scene_generate_s": round(t1 - t0, 3), "scene_upload_build_s": round(t2 - t1, 3),
           "tuning": render.get_tuning(),
           "batched": {"render_s": round(sec, 4), "renders_per_s": round(len(scs) / sec, 4),
                       **_render_rate(ext, sh, pixels, sec)}}
    _log(f"c4 batched: {sec:.3f} s")
    shapes = [(j["height"], j["width"], 3) for j in jobs]
    batched = [f.to_numpy(np.float32, sh).view(np.uint32).copy() for f, sh in zip(d_films, shapes)]

    def same_films():
        return sum(int(np.array_equal(b, f.to_numpy(np.float32, sh).view(np.uint32)))
                   for b, f, sh in zip(batched, d_films, shapes))

    # the same batch with extension rays walked in the host LBVH's right-first
    # preorder only: every film and both ray totals must be equal
    saved = render.get_tuning()
    render.set_tuning(preorder_walk=True)
    try:
        s0 = time.perf_counter()
        render.batch_render_device(scs, jobs, d_films)
        lbvh.synchronize()
        osec = time.perf_counter() - s0
        oe, osh = render.batch_rays()
    finally:
        render.set_tuning(**saved)
    out["preorder_walk"] = {"render_s": round(osec, 4), "extension_rays": oe, "shadow_rays": osh,
                            "ray_totals_equal": bool(oe == ext and osh == sh),
                            "films_bit_identical": f"{same_films()}/{len(scs)}"}
    _log(f"c4 preorder walk: {osec:.3f} s, totals equal {oe == ext and osh == sh}")
    # each scene alone on the wavefront schedule (what sharing one loop buys)
    s0 = time.perf_counter()
    for sc, j, f in zip(scs, jobs, d_films):
        render.batch_render_device([sc], [j], [f])
    lbvh.synchronize()
    wsec = time.perf_counter() - s0
    out["wavefront_per_scene"] = {"render_s": round(wsec, 4), "renders_per_s": round(len(scs) / wsec, 4),
                                  "films_bit_identical": f"{same_films()}/{len(scs)}"}
    out["speedup_vs_wavefront_per_scene"] = round(wsec / sec, 3)
    _log(f"c4 wavefront scene by scene: {wsec:.3f} s")
    if not args.no_plain:
        s0 = time.perf_counter()
        render.batch_render_device(scs, jobs, d_films, mode=render.PLAIN)
        lbvh.synchronize()
        psec = time.perf_counter() - s0
        pe, ps = render.batch_rays()
        same = same_films()
        out["plain"] = {"render_s": round(psec, 4), "renders_per_s": round(len(scs) / psec, 4),
                        **_render_rate(pe, ps, pixels, psec)}
        out["speedup_vs_plain"] = round(psec / sec, 3)
        out["equal_rmse"] = (f"films bit-identical to the plain path ({same}/{len(scs)} scenes), so the "
                             "per-image relRMSE vs the CPU restatement is equal")
        _log(f"c4 plain: {psec:.3f} s")
    if not args.no_cpu_baseline:
        import oracle

        div = []
        for i in (0, 21, 42, 63):
            if i < len(scs):
                j = jobs[i]
                pt = oracle.OraclePathTracer(dicts[i])
                div.append(dict(scene=i, **sample_divergence(scs[i], pt, j["width"], j["height"], j["spp"],
                                                             j["depth"], (0, j["height"]), n=1024, seed=0x4C56 + i)))
        out["sample_divergence"] = div
        _log(f"c4 sample divergence: {[d['diverged'] for d in div]} of {[d['samples'] for d in div]}")
        # CPU baseline: the restatement on a band of each sampled scene (about
        # 2^16 paths per scene, one thread), and the whole mix's CPU time
        # projected from that rate
        cpu_paths, cpu_s, bands = 0, 0.0, []
        for i in (0, 21, 42, 63):
            if i >= len(scs):
                continue
            j = jobs[i]
            W_, H_, spp = j["width"], j["height"], j["spp"]
            nrows = max(1, min(H_, (1 << 16) // (W_ * spp)))
            r0 = H_ // 2 - nrows // 2
            pt = oracle.OraclePathTracer(dicts[i])
            s0 = time.perf_counter()
            band = pt.render(W_, H_, spp, j["depth"], rows=(r0, r0 + nrows), threads=1)[r0:r0 + nrows]
            dt = time.perf_counter() - s0
            cpu_paths += W_ * nrows * spp
            cpu_s += dt
            gpu = d_films[i].to_numpy(np.float32, (H_, W_, 3))[r0:r0 + nrows]
            bands.append({"scene": i, "rows": f"{r0}..{r0 + nrows - 1}", "paths": W_ * nrows * spp,
                          "gpu_vs_cpu_band_relrmse": round(render.rel_rmse(gpu / spp, band / spp), 8)})
        rate = cpu_paths / cpu_s
        out["cpu_baseline"] = {
            "value": round(rate / 1e6, 4), "unit": "Mpaths/s", "cores": 1, "kind": "port",
            "sample": f"a band of rows of scenes 0, 21, 42, 63 at their own spp ({cpu_paths} paths)",
            "host_cpu": _cpu_name(), "projected_mix_cpu_s": round(paths / rate, 1),
            "gpu_speedup_over_projected_cpu": round(paths / rate / sec, 1), "bands": bands}
        _log(f"c4 cpu baseline: {rate / 1e6:.3f} Mpaths/s, mix projected {paths / rate:.0f} s")
    for f in d_films:
        f.free()
    return out, scs, jobs, dicts


def bench_two_level(args):
    """Instanced scene: 4 meshes of 65 536 triangles, 64 random affine instances
    (4 194 304 instanced triangles), 4 194 304 rays aimed into the instanced
    region. Build (every BLAS + TLAS, synchronous, host triangles in), TLAS-only
    rebuild (new instance transforms), closest hit and occlusion from
    HBM-resident rays. CPU baseline: the two-level restatement on one thread over
    a ray sample."""
    from lux_amd import lbvh, two_level as TL, workloads as W

    R = args.rays
    inp = W.two_level_inputs(n_rays=R)
    tris = np.concatenate([v[i].reshape(-1, 9) for v, i in inp["meshes"]]).astype(np.float32)
    off = np.concatenate([[0], np.cumsum([len(i) for _, i in inp["meshes"]])]).astype(np.int32)
    rec = TL.make_instances(inp["inst_mesh"], inp["inst_m"], tris=tris, mesh_offsets=off)
    mesh2, m2 = W.instances_random(len(rec), len(off) - 1, 4242)
    rec2 = TL.make_instances(mesh2, m2, tris=tris, mesh_offsets=off)
    s = TL.TwoLevelScene(tris, off, rec)
    t0 = time.perf_counter()
    for _ in range(3):
        s.build(tris, off, rec)
    build_s = (time.perf_counter() - t0) / 3
    t0 = time.perf_counter()
    for k in range(6):
        s.set_instances(rec2 if k % 2 == 0 else rec)
    tlas_s = (time.perf_counter() - t0) / 6
    d_rays = lbvh.DeviceArray.from_numpy(inp["rays"])
    d_hits = lbvh.DeviceArray(24 * R)
    d_occ = lbvh.DeviceArray(R)
    stage = {}
    for name, fn in (("closest", lambda: s.intersect_device(d_rays, R, d_hits)),
                     ("occluded", lambda: s.occluded_device(d_rays, R, d_occ))):
        fn()
        lbvh.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            fn()
        lbvh.synchronize()
        stage[name] = (time.perf_counter() - t0) / args.steps
    hits = np.frombuffer(d_hits.to_numpy(np.uint8, (24 * R,)).tobytes(), TL.SCENE_HIT)
    info = s.info()
    out = {"workload": (f"{len(off) - 1} meshes x {int(np.diff(off)[0])} triangles, {len(rec)} affine instances "
                        f"({info['scene_primitives']} instanced triangles), {R} rays"),
           "build_ms": round(build_s * 1e3, 3), "tlas_rebuild_ms": round(tlas_s * 1e3, 3),
           "closest_ms": round(stage["closest"] * 1e3, 3), "occluded_ms": round(stage["occluded"] * 1e3, 3),
           "closest_Mrays_per_s": round(R / stage["closest"] / 1e6, 2),
           "occluded_Mrays_per_s": round(R / stage["occluded"] / 1e6, 2),
           "hit_fraction": round(float((hits["primitive_id"] >= 0).mean()), 4),
           "device_bytes": info["device_bytes"]}
    if not args.no_cpu_baseline:
        import oracle

        k = min(R, 1 << 16)
        o = oracle.OracleTwoLevel(tris, off, inp["inst_mesh"], rec["world_to_object"], rec["world_bounds"])
        t0 = time.perf_counter()
        r = o.trace(inp["rays"][:k])
        cpu_s = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": round(k / cpu_s / 1e6, 4), "unit": "Mrays/s", "cores": 1, "kind": "port",
                               "sample": f"closest hit of the first {k} rays, two-level restatement, 1 thread",
                               "equal_to_gpu": bool(np.array_equal(r["prim"], hits["primitive_id"][:k]))}
    s.close()
    _log(f"two-level: closest {out['closest_Mrays_per_s']} Mrays/s, occluded {out['occluded_Mrays_per_s']} Mrays/s")
    return out


def bench_film(args):
    """Film output of a 3840x2160 film of per-pixel sums resident in HBM:
    write_ppm's pixel bytes (15 B per pixel: 12 read, 3 written) and write_pfm's
    float body (24 B per pixel), as HBM-bound kernels."""
    from lux_amd import film as F, lbvh

    w, h, spp = 3840, 2160, 64
    rng = np.random.default_rng(7)
    d_film = lbvh.DeviceArray.from_numpy((rng.random((h, w, 3), dtype=np.float32) * 4 * spp).astype(np.float32))
    d8 = lbvh.DeviceArray(3 * w * h)
    d32 = lbvh.DeviceArray(12 * w * h)
    out = {"workload": f"{w}x{h} film, {spp} spp"}
    for name, fn, per_px in (("ppm", lambda: F.ppm_pixels_device(d_film, w, h, spp, d8), 15),
                             ("pfm", lambda: F.pfm_pixels_device(d_film, w, h, spp, d32), 24)):
        fn()
        lbvh.synchronize()
        probe = lbvh.Probe(f"film_{name}")
        with probe:
            for _ in range(args.steps):
                fn()
            lbvh.synchronize()
        ms = probe.total_ms / max(probe.launches, 1)
        gbps = per_px * w * h / (ms * 1e-3) / 1e9
        out[name] = {"kernel_ms": round(ms, 4), "GBps": round(gbps, 1), "hbm_frac": round(gbps / HBM_PEAK_GBPS, 4)}
    _log(f"film: ppm {out['ppm']['GBps']} GB/s, pfm {out['pfm']['GBps']} GB/s")
    return out


def bench_c5(args, c4_out, scs, jobs, dicts):
    """Served stream: Poisson arrivals at a fraction of the measured batched
    capacity, scenes drawn from the C4 mix, 30% of requests re-render a scene
    after per-vertex N(0, 1e-3^2) jitter (refit, no rebuild), 16-spp slices,
    200 GB HBM budget with the scenes resident; then the same stream on cold
    scenes (the server uploads and builds them, and drops idle ones) under a
    budget that holds a quarter of the stream's scene bytes."""
    from lux_amd import render

    capacity = c4_out["batched"]["renders_per_s"]
    rate = args.c5_load * capacity
    rng = np.random.default_rng(0x4C555B)
    n = args.c5_requests
    arrivals = np.cumsum(rng.exponential(1.0 / rate, n))
    arrivals -= arrivals[0]
    pick = rng.integers(0, len(scs), n)
    # 30% of requests: a re-render of an already requested scene after jitter
    verts, seen, rerender = [], set(), 0
    for i in range(n):
        k = int(pick[i])
        if seen and rng.random() < 0.3:
            k = int(rng.choice(sorted(seen)))
            pick[i] = k
            v = dicts[k]["vertices"]
            verts.append((v + rng.normal(0, 1e-3, v.shape)).astype(np.float32))
            rerender += 1
        else:
            verts.append(None)
        seen.add(k)
    # the stream's first-time vertices: the warm run below moves the C4 scenes
    used = sorted(set(int(k) for k in pick))
    cold = {k: render.Scene(dicts[k], deferred=True) for k in used}
    sv = render.Server(hbm_budget=200 << 30, slice_spp=16)
    req_scenes = [scs[int(k)] for k in pick]
    req_jobs = [jobs[int(k)] for k in pick]
    _, lat, wall = sv.run(req_scenes, req_jobs, arrivals, vertices=verts, want_films=False)
    st = sv.stats()
    sv.close()
    lat_ms = lat * 1e3
    paths = sum(j["width"] * j["height"] * j["spp"] for j in req_jobs)
    # the same stream on cold scenes (created deferred: the server uploads and
    # builds each on admission) under a budget that holds a quarter of them
    scene_bytes = sum(cold[k].info()[2] for k in used)
    budget = int(st["peak_bytes"] - 0.75 * scene_bytes)
    sv = render.Server(hbm_budget=budget, slice_spp=16)
    _, clat, cwall = sv.run([cold[int(k)] for k in pick], req_jobs, arrivals, vertices=verts, want_films=False)
    cst = sv.stats()
    sv.close()
    for sc in cold.values():
        sc.close()
    clat_ms = clat * 1e3
    cold_out = {"hbm_budget_GB": round(budget / 1e9, 3), "scene_GB": round(scene_bytes / 1e9, 3),
                "renders_per_s": round(n / cwall, 4), "wall_s": round(cwall, 3),
                "p50_latency_ms": round(float(np.percentile(clat_ms, 50)), 2),
                "p99_latency_ms": round(float(np.percentile(clat_ms, 99)), 2),
                "uploads": int(cst["uploads"]), "evictions": int(cst["evictions"]), "refits": int(cst["refits"]),
                "prepare_s": round(cst["prepare_seconds"], 3), "peak_hbm_GB": round(cst["peak_bytes"] / 1e9, 3)}
    out = {"workload": (f"{n} requests, Poisson arrivals at {rate:.3f}/s ({args.c5_load:.0%} of the measured "
                        f"batched capacity {capacity:.3f} renders/s), scenes from the C4 mix, {rerender} re-renders "
                        f"after sigma=1e-3 vertex jitter (refit), 16-spp slices, 200 GB HBM budget"),
           "renders_per_s": round(n / wall, 4), "wall_s": round(wall, 3),
           "p50_latency_ms": round(float(np.percentile(lat_ms, 50)), 2),
           "p99_latency_ms": round(float(np.percentile(lat_ms, 99)), 2),
           "mean_latency_ms": round(float(lat_ms.mean()), 2),
           "refits": int(st["refits"]), "batches": int(st["batches"]), "paths": int(paths),
           "peak_hbm_GB": round(st["peak_bytes"] / 1e9, 3),
           "Mrays_per_s": round((st["extension_rays"] + st["shadow_rays"]) / wall / 1e6, 2),
           "cold_scenes_tight_budget": cold_out}
    _log(f"c5: {wall:.2f} s, p50 {out['p50_latency_ms']} ms p99 {out['p99_latency_ms']} ms; cold/tight: "
         f"p50 {cold_out['p50_latency_ms']} ms p99 {cold_out['p99_latency_ms']} ms, {cold_out['uploads']} uploads, "
         f"{cold_out['evictions']} evictions")
    return out


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
    d_spairs, d_sfound = lbvh.DeviceArray(8 * pair_cap), lbvh.DeviceArray(4)
    bvh, qbvh = lbvh.GpuBvh(), lbvh.GpuBvh()

    def build():
        bvh.build_device(d_boxes, d_ids, N)

    def closest():
        call("lux_hip_trace_closest_device", bvh.handle, d_tris.ptr, N, d_rays.ptr, R, d_hits.ptr, None)

    def anyhit():
        call("lux_hip_trace_any_device", bvh.handle, d_tris.ptr, N, d_rays.ptr, R, d_occ.ptr, None)

    def overlap_cross():
        call("lux_hip_overlap_cross_device", bvh.handle, qbvh.handle, pair_cap, d_pairs.ptr, d_found.ptr, None)

    def overlap_self():
        call("lux_hip_overlap_self_device", bvh.handle, pair_cap, d_spairs.ptr, d_sfound.ptr, None)

    def overlap():
        qbvh.build_device(d_qboxes, d_qids, Q)
        overlap_cross()
        overlap_self()

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

    with lbvh.Probe("trace_any") as probe:
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
    self_found = int(d_sfound.to_numpy(np.int32, 1)[0])

    # the closest-hit kernel, timed the same way over as many launches (on a
    # fresh build: the step ends with a refit to jittered triangles)
    build()
    with lbvh.Probe("trace_closest") as cprobe:
        for _ in range(args.steps):
            closest()
        lbvh.synchronize()
    closest_ms = cprobe.total_ms / max(cprobe.launches, 1)

    # attainable HBM bandwidth on this box: the plain copy kernel over 1 GiB
    # (1 GiB read + 1 GiB written per launch), timed the same way
    from lux_amd.abi import call as _call

    copy_bytes = 1 << 30
    c_src, c_dst = lbvh.DeviceArray(copy_bytes), lbvh.DeviceArray(copy_bytes)
    _call("lux_hip_copy_device", c_dst.ptr, c_src.ptr, ctypes.c_size_t(copy_bytes), None)
    with lbvh.Probe("copy16") as kprobe:
        for _ in range(5):
            _call("lux_hip_copy_device", c_dst.ptr, c_src.ptr, ctypes.c_size_t(copy_bytes), None)
        lbvh.synchronize()
    attainable = 2 * copy_bytes / (kprobe.total_ms / max(kprobe.launches, 1) * 1e-3) / 1e9
    c_src.free()
    c_dst.free()

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
    for name, f in (("overlap_cross", overlap_cross), ("overlap_self", overlap_self)):
        build()
        qbvh.build_device(d_qboxes, d_qids, Q)
        f()
        lbvh.synchronize()
        s0 = time.perf_counter()
        for _ in range(args.steps):
            f()
        lbvh.synchronize()
        stage_ms[name] = round((time.perf_counter() - s0) * 1e3 / args.steps, 4)
    build()
    stats = (ctypes.c_uint64 * 6)()
    call("lux_hip_trace_stats_device", bvh.handle, d_tris.ptr, N, d_rays.ptr, R, stats)

    # any hit: o, d, t_min, t_max in (32 B) + a 4-byte flag out per ray
    algo_bytes = 36 * R
    achieved = algo_bytes / (kernel_ms * 1e-3) / 1e9
    # touched by the walks: 32-B nodes + 36-B triangles (stats[3..5] any hit,
    # stats[0..2] closest hit)
    touched = 32 * stats[3] + 32 * stats[4] + 36 * stats[5]
    c_algo = 40 * R  # closest hit: ray in (32 B) + {id, t} out (8 B)
    c_touched = 32 * stats[0] + 32 * stats[1] + 36 * stats[2]

    def profiled(table, key, fields):
        """A committed profile entry of this workload (same rays and triangles)."""
        path = TRAFFIC_FILE if table == "traffic" else COUNTERS_FILE
        if not path.exists():
            return None
        e = json.loads(path.read_text()).get(key)
        if not e or e.get("rays") != R or e.get("triangles") != N:
            return None
        return {f: e.get(f) for f in fields} if fields else e

    t_any = profiled("traffic", "trace_any", ["bytes_per_launch"])
    t_closest = profiled("traffic", "trace_closest", ["bytes_per_launch"])
    valu_fields = ["valu_busy_frac", "lane_utilization", "valu_insts_per_wave", "waves_per_simd", "l2_hit_rate",
                   "kernel_us", "source"]
    v_any = profiled("counters", "k_trace_any", valu_fields)
    v_closest = profiled("counters", "k_trace_closest", valu_fields)

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
                         f"closest-hit and any-hit of {R} rays + cross overlap of {Q} query boxes + self overlap "
                         f"of the tree + refit after sigma=1e-3 vertex jitter"),
            "triangles": N,
            "rays_per_step": rays_per_step,
            "queries": Q,
            "parallelism": "single-gpu",
        },
        "roofline": {
            "kernel": "trace_any",
            "bound": "hbm",
            "achieved": round(achieved, 3),
            "peak": HBM_PEAK_GBPS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBPS, 6),
            # the copy kernel's measured bandwidth, and achieved against it
            "attainable_GBps": round(attainable, 1),
            "frac_of_attainable": round(achieved / attainable, 6),
            "traffic": t_any["bytes_per_launch"] if t_any else None,
            "kernel_ms": round(kernel_ms, 4),
            "algorithmic_bytes_per_launch": algo_bytes,
            "touched_bytes_per_launch": int(touched),
            "touched_GBps": round(touched / (kernel_ms * 1e-3) / 1e9, 2),
            "node_visits_per_ray": round(stats[3] / R, 2),
            "triangle_tests_per_ray": round(stats[5] / R, 3),
            # the VALU bound beside it: issue-busy fraction of the vector ALUs over
            # the kernel and the share of lanes an issued instruction carries
            # (rocprofv3 SQ counters of this kernel on this workload)
            "valu": v_any,
            "closest_hit": {
                "kernel_ms": round(closest_ms, 4),
                "achieved_GBps": round(c_algo / (closest_ms * 1e-3) / 1e9, 3),
                "frac": round(c_algo / (closest_ms * 1e-3) / 1e9 / HBM_PEAK_GBPS, 6),
                "traffic": t_closest["bytes_per_launch"] if t_closest else None,
                "touched_bytes_per_launch": int(c_touched),
                "node_visits_per_ray": round(stats[0] / R, 2),
                "triangle_tests_per_ray": round(stats[2] / R, 3),
                "valu": v_closest,
            },
        },
        "stages_ms": stage_ms,
        "stage_rates": {
            "build_Mprims_per_s": round(N / stage_ms["build"] / 1e3, 2),
            # algorithmic bytes: build 360 B/prim, refit 144 B/prim (input, sort,
            # links, boxes, layout, nodes; vertices, links, boxes)
            "build_hbm_frac": round(360 * N / (stage_ms["build"] * 1e-3) / 1e9 / HBM_PEAK_GBPS, 5),
            "refit_Mprims_per_s": round(N / stage_ms["refit"] / 1e3, 2),
            "refit_hbm_frac": round(144 * N / (stage_ms["refit"] * 1e-3) / 1e9 / HBM_PEAK_GBPS, 5),
            "closest_Mrays_per_s": round(R / stage_ms["closest"] / 1e3, 2),
            "any_Mrays_per_s": round(R / stage_ms["any"] / 1e3, 2),
            "overlap_cross_Mqueries_per_s": round(Q / stage_ms["overlap_cross"] / 1e3, 2),
            "overlap_self_Mqueries_per_s": round(N / stage_ms["overlap_self"] / 1e3, 2),
            "overlap_cross_pairs": found,
            "overlap_self_pairs": self_found,
            "hit_fraction": round(float((hits[:, 0] >= 0).mean()), 4),
        },
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(inp, min(args.cpu_sample_rays, R))
    _log(f"lbvh: {out['value']} Mrays/s")
    del d_boxes, d_ids, d_tris, d_jit, d_rays, d_qboxes, d_qids, d_hits, d_occ, d_pairs, d_found
    del d_spairs, d_sfound, bvh, qbvh

    if args.two_level == "on":
        out["two_level"] = bench_two_level(args)
    out["film_output"] = bench_film(args)

    wanted = [w for w in args.render.split(",") if w and w != "none"]
    if wanted and args.no_graph:
        from lux_amd import autotune, render

        autotune.apply()
        render.set_tuning(use_graph=0)
    if wanted:
        out["render"] = {}
    if "c1" in wanted:
        out["render"]["c1_cornell"] = bench_c1(args)
    if "c3" in wanted:
        out["render"]["c3_single_scene"] = bench_c3(args)
    if "instanced" in wanted:
        out["render"]["instanced_scene"] = bench_instanced(args)
    if "bunny" in wanted:
        out["render"]["bunny"] = bench_bunny(args)
    if "c4" in wanted or "c5" in wanted:
        c4, scs, jobs, dicts = bench_c4(args)
        if "c4" in wanted:
            out["render"]["c4_batched"] = c4
        if "c5" in wanted:
            out["render"]["c5_served"] = bench_c5(args, c4, scs, jobs, dicts)
        for sc in scs:
            sc.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

"""Build step of the product: the gfx950 HIP library (lux_amd/liblux_hip.so).
Run from anywhere; the output is in-tree."""

import os
import subprocess
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

from . import LIB_PATH, PACKAGE_DIR, REPO_ROOT

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc" if Path("/opt/rocm/bin/hipcc").exists() else "hipcc")
GPU_ARCH = "gfx950"

HIP_SOURCES = ["lbvh.hip", "query.hip", "two_level.hip", "film.hip", "obj_import.hip", "pt.hip", "pt_scene.hip", "serve.hip", "noalloc.hip", "abi.hip"]
# -ffp-contract=off keeps the reference's exact float operation order (no FMA
# contraction) in box, triangle and Morton arithmetic; packed FP32 ops are off.
HIP_FLAGS = [
    "-O3", f"--offload-arch={GPU_ARCH}", "-std=c++20", "-fPIC", "-shared", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]


def _run(cmd, cwd=REPO_ROOT):
    r = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"build step failed ({r.returncode}): {' '.join(map(str, cmd))}\n{r.stdout}\n{r.stderr}")
    return r


def build_hip():
    """Each source compiles to its own object in parallel (separate device code
    objects, as one multi-source hipcc line makes them), then one link."""
    obj_dir = PACKAGE_DIR / "_obj"
    obj_dir.mkdir(exist_ok=True)
    inc = [f"-I{REPO_ROOT / 'include'}", f"-I{PACKAGE_DIR / 'csrc'}"]
    compile_flags = [f for f in HIP_FLAGS if f != "-shared"]

    def compile_one(src):
        obj = obj_dir / (Path(src).stem + ".o")
        _run([HIPCC, *compile_flags, *inc, "-c", str(PACKAGE_DIR / "csrc" / src), "-o", str(obj)])
        return str(obj)

    jobs = max(1, min(len(HIP_SOURCES), os.cpu_count() or 1, 16))
    with ThreadPoolExecutor(jobs) as pool:
        objs = list(pool.map(compile_one, HIP_SOURCES))
    tmp = LIB_PATH.with_suffix(".so.tmp")
    _run([HIPCC, f"--offload-arch={GPU_ARCH}", "-shared", "-fPIC", *objs, "-o", str(tmp)])
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


def build_all():
    return [build_hip()]

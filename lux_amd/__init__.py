"""lux_amd -- MI355X (gfx950) hot path of the LUX renderer.

The compute lives in ``liblux_hip.so`` (hand-written HIP kernels behind the C ABI
declared in ``include/lux_hip.h``). This package only loads that library and
offers thin numpy-facing wrappers named after the reference C++ interface
(``build_gpu_lbvh``, ``refit_gpu_lbvh``, ``intersect_gpu_bvh_stackless``,
``query_aabb_overlaps_self`` / ``_cross``). There is no CPU fallback: if the
library is missing, importing :mod:`lux_amd.lbvh` raises.
"""

from pathlib import Path

PACKAGE_DIR = Path(__file__).resolve().parent
REPO_ROOT = PACKAGE_DIR.parent
LIB_PATH = PACKAGE_DIR / "liblux_hip.so"

__all__ = ["PACKAGE_DIR", "REPO_ROOT", "LIB_PATH"]

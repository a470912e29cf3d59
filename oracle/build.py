"""Build steps of the checker (test infrastructure): the CPU restatement
(oracle/_build/liblux_oracle.so) and, where the LUX sources are mounted, the
reference's own host code compiled from them by oracle/ref.mk
(oracle/_ref/libluxref.so), which the golden-vector scripts under tests/golden/
run. Neither is part of the product library."""

import shutil
import subprocess
from pathlib import Path

ORACLE_DIR = Path(__file__).resolve().parent
REPO_ROOT = ORACLE_DIR.parent
REFERENCE_SRC = Path("/root/reference/src")
ORACLE_SRCS = [ORACLE_DIR / f for f in ("lbvh_oracle.cpp", "pt_oracle.cpp", "two_level_oracle.cpp",
                                        "film_oracle.cpp", "ordered_model.cpp")]
ORACLE_LIB = ORACLE_DIR / "_build" / "liblux_oracle.so"


def _run(cmd, cwd=REPO_ROOT):
    r = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"build step failed ({r.returncode}): {' '.join(map(str, cmd))}\n{r.stdout}\n{r.stderr}")
    return r


def build_oracle():
    ORACLE_LIB.parent.mkdir(parents=True, exist_ok=True)
    _run(["g++", "-std=c++20", "-O2", "-fPIC", "-shared", "-ffp-contract=off", "-pthread",
          *map(str, ORACLE_SRCS), "-o", str(ORACLE_LIB)])
    return ORACLE_LIB


def build_reference_shim():
    """oracle/_ref/libluxref.so from the reference's own host sources, only
    where they are mounted (None elsewhere)."""
    if not REFERENCE_SRC.exists() or shutil.which("make") is None:
        return None
    _run(["make", "-s", "-B", "-f", "oracle/ref.mk", f"REF_SRC={REFERENCE_SRC}"])
    return ORACLE_DIR / "_ref" / "libluxref.so"

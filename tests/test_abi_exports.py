"""CPU: the C-ABI library loads and exports every entry point declared in
include/lux_hip.h; the ctypes table and the header agree. No GPU calls."""

import ctypes
import re
import subprocess
from pathlib import Path

from lux_amd import LIB_PATH, REPO_ROOT, abi

HEADER = REPO_ROOT / "include" / "lux_hip.h"


def declared():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(lux_hip_\w+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared()
    assert "lux_hip_bvh_build" in names and "lux_hip_trace_closest" in names
    assert len(names) >= 25


def test_library_exports_every_declared_symbol():
    assert LIB_PATH.exists(), "run __graft_entry__.build() first"
    lib = ctypes.CDLL(str(LIB_PATH))
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_exports_are_c_linkage():
    out = subprocess.run(["nm", "-D", "--defined-only", str(LIB_PATH)], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\b(lux_hip_\w+)\b", out))
    assert set(declared()) <= exported


def test_ctypes_table_matches_header():
    assert sorted(abi.exported_names()) == declared()


def test_version_and_error_string_without_gpu():
    lib = abi.load()
    assert lib.lux_hip_version() == 0x000100
    assert lib.lux_hip_last_error() == b""

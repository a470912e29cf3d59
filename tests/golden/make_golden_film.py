"""Writes film-output golden files under tests/golden/ with the REFERENCE's own
write_ppm / write_pfm (src/film/ppm.h, src/film/pfm.h) through oracle/_ref.

Run in a container where /root/reference is mounted:
    python tests/golden/make_golden_film.py
Each case stores the film (per-pixel sums), samples per pixel, and the exact
bytes of both files.
"""

import sys
import tempfile
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

import oracle  # noqa: E402

OUT = Path(__file__).resolve().parent


def film_mixed(w, h, seed):
    """Log-uniform radiance sums over [1e-6, 1e4] with exact zeros, values
    that land on byte boundaries, negatives, NaN and +-inf."""
    rng = np.random.default_rng(seed)
    f = np.exp(rng.uniform(np.log(1e-6), np.log(1e4), (h, w, 3))).astype(np.float32)
    flat = f.reshape(-1)
    k = len(flat)
    flat[rng.choice(k, k // 20, replace=False)] = 0.0
    flat[rng.choice(k, 8, replace=False)] = -1.5
    flat[rng.choice(k, 4, replace=False)] = np.nan
    flat[rng.choice(k, 4, replace=False)] = np.inf
    flat[rng.choice(k, 4, replace=False)] = -np.inf
    return f


def write(name, film, spp):
    with tempfile.TemporaryDirectory() as d:
        oracle.ref_write_film("ppm", Path(d) / "f.ppm", film, spp)
        oracle.ref_write_film("pfm", Path(d) / "f.pfm", film, spp)
        ppm = np.frombuffer((Path(d) / "f.ppm").read_bytes(), np.uint8)
        pfm = np.frombuffer((Path(d) / "f.pfm").read_bytes(), np.uint8)
    np.savez_compressed(OUT / f"{name}.npz", film=film, spp=np.int32(spp), ppm=ppm, pfm=pfm)
    print(f"{name}: {film.shape[1]}x{film.shape[0]} spp {spp}, ppm {len(ppm)} B, pfm {len(pfm)} B")


if __name__ == "__main__":
    write("film_mixed_s16", film_mixed(67, 41, 1), 16)
    write("film_mixed_s1", film_mixed(32, 32, 2), 1)
    write("film_spp0", film_mixed(5, 3, 3), 0)
    ramp = np.linspace(0.0, 4.0, 3 * 256 * 4, dtype=np.float32).reshape(4, 256, 3)
    write("film_ramp", ramp, 1)

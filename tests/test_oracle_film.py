"""Film output: the restated writers against the reference's own write_ppm /
write_pfm (golden files of tests/golden/make_golden_film.py), and the library's
gamma threshold table against the byte formula on every float in [0, 1]."""

import ctypes

import numpy as np
import pytest

import oracle
from conftest import load_golden

FILM_CASES = ["film_mixed_s16", "film_mixed_s1", "film_spp0", "film_ramp"]


@pytest.mark.parametrize("case", FILM_CASES)
def test_restated_writers_match_reference_files(case):
    g = load_golden(case)
    assert np.array_equal(oracle.film_file("ppm", g["film"], int(g["spp"])), g["ppm"])
    assert np.array_equal(oracle.film_file("pfm", g["film"], int(g["spp"])), g["pfm"])


def test_library_gamma_thresholds_reproduce_the_byte_formula_everywhere():
    """lux_hip_ppm_thresholds (host code of the C ABI, no GPU) against
    (unsigned char)(255.999f * powf(c, 1/2.2)) for all 1 065 353 217 floats in
    [0, 1]: the GPU's byte conversion counts these thresholds."""
    from lux_amd import abi

    t = np.zeros(255, np.float32)
    abi.call("lux_hip_ppm_thresholds", ctypes.c_void_p(t.ctypes.data))
    assert np.all(np.diff(t) > 0) and t[0] > 0 and t[-1] <= 1.0
    assert oracle.ppm_table_mismatches(t, threads=8) == 0

"""GPU film output through the C ABI: files byte-identical to the reference's
own write_ppm / write_pfm (golden files) and to the restated writers on a
rendered film."""

import numpy as np
import pytest

import oracle
from conftest import load_golden
from lux_amd import abi, film as F, lbvh, render, scenes

pytestmark = pytest.mark.gpu
FILM_CASES = ["film_mixed_s16", "film_mixed_s1", "film_spp0", "film_ramp"]


def print_case(*what):
    print("case:", *what, flush=True)


@pytest.mark.parametrize("case", FILM_CASES)
def test_film_files_match_reference(gpu, case, tmp_path):
    print_case("film files", case)
    g = load_golden(case)
    film = np.ascontiguousarray(g["film"], np.float32)
    h, w = film.shape[:2]
    spp = int(g["spp"])
    d = lbvh.DeviceArray.from_numpy(film)
    F.write_ppm(tmp_path / "out" / "a.ppm", d, w, h, spp)
    F.write_pfm(tmp_path / "out" / "a.pfm", d, w, h, spp)
    assert np.array_equal(np.frombuffer((tmp_path / "out" / "a.ppm").read_bytes(), np.uint8), g["ppm"])
    assert np.array_equal(np.frombuffer((tmp_path / "out" / "a.pfm").read_bytes(), np.uint8), g["pfm"])
    body8 = lbvh.DeviceArray(max(3 * w * h, 1))
    F.ppm_pixels_device(d, w, h, spp, body8)
    lbvh.synchronize()
    assert np.array_equal(body8.to_numpy(np.uint8, (3 * w * h,)), g["ppm"][-3 * w * h:])
    body32 = lbvh.DeviceArray(max(12 * w * h, 1))
    F.pfm_pixels_device(d, w, h, spp, body32)
    lbvh.synchronize()
    assert np.array_equal(body32.to_numpy(np.uint8, (12 * w * h,)), g["pfm"][-12 * w * h:])


def test_rendered_film_files_match_restated_writers(gpu, tmp_path):
    print_case("film of a render")
    d = scenes.cornell(48, 40)
    sc = render.Scene(d)
    film = lbvh.DeviceArray(48 * 40 * 12)
    sc.render_device(film, 48, 40, 8, 5)
    lbvh.synchronize()
    host = film.to_numpy(np.float32, (40, 48, 3))
    F.write_ppm(tmp_path / "c.ppm", film, 48, 40, 8)
    F.write_pfm(tmp_path / "c.pfm", film, 48, 40, 8)
    assert np.array_equal(np.frombuffer((tmp_path / "c.ppm").read_bytes(), np.uint8), oracle.film_file("ppm", host, 8))
    assert np.array_equal(np.frombuffer((tmp_path / "c.pfm").read_bytes(), np.uint8), oracle.film_file("pfm", host, 8))
    sc.close()


def test_film_output_argument_errors(gpu, tmp_path):
    print_case("film errors")
    d = lbvh.DeviceArray(12)
    with pytest.raises(abi.LuxError) as e:
        abi.call("lux_hip_write_ppm", None, d.ptr, 1, 1, 1)
    assert e.value.status == abi.LUX_ERR_INVALID_ARGUMENT
    with pytest.raises(abi.LuxError) as e:
        abi.call("lux_hip_write_pfm", str(tmp_path / "x.pfm").encode(), d.ptr, -1, 1, 1)
    assert e.value.status == abi.LUX_ERR_INVALID_ARGUMENT

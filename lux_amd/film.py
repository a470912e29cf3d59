"""Film output over liblux_hip.so: write_ppm / write_pfm of the reference
(src/film/ppm.h:37-56, src/film/pfm.h:27-44) from a film of per-pixel sums in
device memory (a lux_amd.lbvh.DeviceArray of W*H*3 floats)."""

import ctypes

from .abi import call


def write_ppm(path, d_film, width, height, samples_per_pixel):
    call("lux_hip_write_ppm", str(path).encode(), d_film.ptr, width, height, samples_per_pixel)


def write_pfm(path, d_film, width, height, samples_per_pixel):
    call("lux_hip_write_pfm", str(path).encode(), d_film.ptr, width, height, samples_per_pixel)


def ppm_pixels_device(d_film, width, height, samples_per_pixel, d_out, stream=None):
    """write_ppm's pixel bytes (3 per pixel, film order) into d_out."""
    call("lux_hip_film_ppm_device", d_film.ptr, width, height, samples_per_pixel, d_out.ptr, stream)


def pfm_pixels_device(d_film, width, height, samples_per_pixel, d_out, stream=None):
    """write_pfm's float body (rows bottom to top) into d_out."""
    call("lux_hip_film_pfm_device", d_film.ptr, width, height, samples_per_pixel, d_out.ptr, stream)

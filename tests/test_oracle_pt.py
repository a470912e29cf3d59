"""CPU: the path-tracer oracle restatement against golden vectors written by the
reference's own CPU path tracer, plus independent pins (PCG32 known answers,
thread/row-split invariance)."""

import numpy as np
import pytest

import oracle
from conftest import load_golden

PT_CASES = ["pt_cornell_32", "pt_icosphere", "pt_mixed", "pt_mixed_d0", "pt_mixed_d1"]


def frame_of(g):
    return int(g["frame"]) if "frame" in g else 0


def scene_of(g):
    return {k: g[k] for k in ("vertices", "indices", "normals", "uvs", "material_ids", "materials", "camera")}


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


@pytest.mark.parametrize("case", PT_CASES + ["pt_cornell_c1"])
def test_camera_and_light_tables_match_reference(case):
    g = load_golden(case)
    o = oracle.OraclePathTracer(scene_of(g))
    assert np.array_equal(bits(o.camera()), bits(g["camera_vectors"]))
    lights = o.lights()
    for k, v in lights.items():
        assert np.array_equal(np.asarray(v).view(np.uint32), np.asarray(g["light_" + k]).view(np.uint32)), k


@pytest.mark.parametrize("case", PT_CASES)
def test_film_and_samples_match_reference(case):
    g = load_golden(case)
    o = oracle.OraclePathTracer(scene_of(g))
    w, h, spp, depth, seed = int(g["width"]), int(g["height"]), int(g["spp"]), int(g["depth"]), int(g["seed"])
    frame = frame_of(g)
    assert np.array_equal(bits(o.render(w, h, spp, depth, seed=seed, frame=frame)), bits(g["film_sum"]))
    if "sample_pixels" in g:
        s = o.samples(w, h, depth, g["sample_pixels"], g["sample_index"], seed=seed, frame=frame)
        assert np.array_equal(bits(s), bits(g["sample_radiance"]))


def test_c1_oracle_image_rows():
    """A band of rows of the 256x256 / 64 spp Cornell oracle image (the whole
    image takes ~15 s single-threaded; the band exercises every code path)."""
    g = load_golden("pt_cornell_c1")
    o = oracle.OraclePathTracer(scene_of(g))
    film = o.render(256, 256, 64, 4, seed=0, rows=(96, 112), threads=8)
    assert np.array_equal(bits(film[96:112]), bits(g["film_sum"][96:112]))


def test_pcg32_known_answers():
    """pcg32 seeded with (42, 54) -- the published PCG-XSH-RR demo sequence."""
    ref = [0xA15C02B7, 0x7B47F409, 0xBA1D3330, 0x83D2F293, 0xBFA4784B, 0xCBED606E]
    assert list(oracle.pcg32(42, 54, 6)) == ref


def test_row_split_and_threads_do_not_change_the_image():
    g = load_golden("pt_cornell_32")
    o = oracle.OraclePathTracer(scene_of(g))
    a = o.render(32, 32, 4, 4, threads=1)
    b = o.render(32, 32, 4, 4, threads=5)
    assert np.array_equal(bits(a), bits(b))


@pytest.mark.parametrize("case", ["pt_lights_bvh", "pt_instanced_bvh"])
def test_reference_light_bvh_invariants(case):
    """The reference's light BVH (golden) against what its definition fixes:
    2K-1 nodes in a consistent binary tree, leaves = the lights (each light's
    world triangle bounds, its power), interior bounds = union and power = sum
    of the children; and, at any shading point, the tree's selection pmfs
    (light_bvh_pmf: the product of child probabilities power / max(d^2, eps^2)
    from leaf to root) sum to 1 over the lights, restated here in float64."""
    g = load_golden(case)
    b, pw, ln, leaf, root = (g["light_bvh_bounds"].astype(np.float64), g["light_bvh_power"].astype(np.float64),
                             g["light_bvh_links"], g["light_bvh_leaf"], int(g["light_bvh_root"]))
    k = len(g["light_prim"])
    assert len(pw) == 2 * k - 1 and root == 0 and ln[root, 2] == -1
    if "inst_mesh" in g:  # instanced: world triangles from the reference's shading view
        from_view = g["tri_corners"].astype(np.float64)
    else:
        from_view = g["vertices"].astype(np.float64)[g["indices"]].reshape(-1, 9)
    for node, (l, r, parent, light) in enumerate(ln):
        if light >= 0:
            assert l == -1 and r == -1 and leaf[light] == node
            t = from_view[g["light_prim"][light]].reshape(3, 3)
            assert np.array_equal(b[node], np.concatenate([t.min(0), t.max(0)]).astype(np.float32))
        else:
            assert ln[l, 2] == node and ln[r, 2] == node
            assert np.array_equal(b[node, :3], np.minimum(b[l, :3], b[r, :3]))
            assert np.array_equal(b[node, 3:], np.maximum(b[l, 3:], b[r, 3:]))
            assert np.float32(pw[node]) == np.float32(np.float32(pw[l]) + np.float32(pw[r]))
    eps2 = np.float64(np.float32(1e-5) * np.float32(1e-5))

    def importance(i, p):
        d = np.maximum(np.maximum(b[i, :3] - p, p - b[i, 3:]), 0)
        return pw[i] / max(float((d * d).sum()), eps2)

    rng = np.random.default_rng(1)
    for p in rng.uniform([-1, 0, 0], [1, 2, 2], (16, 3)):
        total = 0.0
        for light in range(k):
            node, pmf = int(leaf[light]), 1.0
            while ln[node, 2] >= 0:
                par = int(ln[node, 2])
                li, ri = importance(int(ln[par, 0]), p), importance(int(ln[par, 1]), p)
                pmf *= (li if ln[par, 0] == node else ri) / (li + ri)
                node = par
            total += pmf
        assert abs(total - 1.0) < 1e-9

"""CPU: pins of the instanced-scene golden files (tests/golden/make_golden_pt.py,
written by the reference's own scene_triangle_view, light preparation and CPU
path tracer) against what the mathematics fixes, computed independently in
float64: world corners are object_to_world applied to the mesh vertices; world
normals are the inverse-transpose normals, flipped for a mirroring instance
(the reference's adjugate transform_normal equals det(M) * M^-T); materials
follow the instance override; area lights are the emissive ordinals in order
with their world-space areas. Then the CPU restatement of the path tracer on
those instanced scenes and on the reference's bunny scene (rebuilt from the OBJ
fixture): light tables, films and per-sample radiance bit-identical to the
reference's own render."""

import lzma

import numpy as np
import pytest

import oracle
from conftest import GOLDEN, load_golden

SCENE_KEYS = ("vertices", "indices", "normals", "uvs", "material_ids", "materials", "camera", "vertex_offsets",
              "triangle_offsets", "inst_mesh", "inst_m", "inst_override")

CASES = ["pt_instanced", "pt_instanced_opaque"]


def ordinals(g):
    """(instance, mesh, local triangle, global triangle) per scene ordinal."""
    out = []
    to = g["triangle_offsets"]
    for i, m in enumerate(g["inst_mesh"]):
        for t in range(to[m], to[m + 1]):
            out.append((i, m, t - to[m], t))
    return out


@pytest.mark.parametrize("case", CASES)
def test_world_corners_are_the_instance_transform(case):
    g = load_golden(case)
    vo = g["vertex_offsets"]
    got = g["tri_corners"].reshape(-1, 3, 3).astype(np.float64)
    for o, (i, m, _, t) in enumerate(ordinals(g)):
        M = g["inst_m"][i].astype(np.float64)
        v = g["vertices"][vo[m] + g["indices"][t]].astype(np.float64)
        w = v @ M[:3, :3].T + M[:3, 3]
        assert np.allclose(got[o], w, rtol=0, atol=2e-6), o


@pytest.mark.parametrize("case", CASES)
def test_world_normals_are_inverse_transpose_with_handedness(case):
    g = load_golden(case)
    vo = g["vertex_offsets"]
    got = g["tri_normals"].reshape(-1, 3, 3).astype(np.float64)
    mirrored = 0
    for o, (i, m, _, t) in enumerate(ordinals(g)):
        M = g["inst_m"][i].astype(np.float64)[:3, :3]
        n = g["normals"][vo[m] + g["indices"][t]].astype(np.float64)
        if not np.any(n):
            assert not np.any(got[o]), o  # a mesh without normals keeps zeros (flat shading)
            continue
        w = n @ np.linalg.inv(M)  # rows: (M^-T n)^T
        w *= np.sign(np.linalg.det(M))
        w /= np.linalg.norm(w, axis=1, keepdims=True)
        assert np.allclose(got[o], w, rtol=0, atol=2e-6), o
        mirrored += np.linalg.det(M) < 0
    assert mirrored > 0  # the mirroring instance is covered


@pytest.mark.parametrize("case", CASES)
def test_materials_follow_the_override(case):
    g = load_golden(case)
    exp = [g["inst_override"][i] if g["inst_override"][i] >= 0 else g["material_ids"][t]
           for i, _, _, t in ordinals(g)]
    assert np.array_equal(g["tri_material"], np.asarray(exp, np.int32))
    assert len(set(g["inst_override"][g["inst_override"] >= 0])) >= 2


@pytest.mark.parametrize("case", CASES)
def test_area_lights_are_emissive_ordinals_with_world_area(case):
    g = load_golden(case)
    emissive = np.max(g["materials"][:, 4:7], axis=1) > 0
    c = g["tri_corners"].reshape(-1, 3, 3).astype(np.float64)
    area = 0.5 * np.linalg.norm(np.cross(c[:, 1] - c[:, 0], c[:, 2] - c[:, 0]), axis=1)
    exp = np.nonzero(emissive[g["tri_material"]] & (area > 0))[0]
    assert np.array_equal(g["light_prim"], exp)
    assert np.allclose(g["light_area"], area[exp], rtol=1e-6, atol=0)
    lookup = np.full(len(c), -1, np.int32)
    lookup[exp] = np.arange(len(exp))
    assert np.array_equal(g["light_primitive_light_index"], lookup)
    # an instanced emitter under a non-uniform scale is among them
    inst_of = np.array([i for i, *_ in ordinals(g)])
    assert len(set(inst_of[exp])) >= 2


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def check_port(o, g, ordinals):
    assert o.ordinals == ordinals
    assert np.array_equal(bits(o.camera()), bits(g["camera_vectors"]))
    for k, v in o.lights().items():
        assert np.array_equal(np.asarray(v).view(np.uint32), np.asarray(g["light_" + k]).view(np.uint32)), k
    w, h, spp, depth, seed = (int(g[k]) for k in ("width", "height", "spp", "depth", "seed"))
    frame = int(g["frame"])
    assert np.array_equal(bits(o.render(w, h, spp, depth, seed=seed, frame=frame, threads=4)), bits(g["film_sum"]))
    s = o.samples(w, h, depth, g["sample_pixels"], g["sample_index"], seed=seed, frame=frame)
    assert np.array_equal(bits(s), bits(g["sample_radiance"]))


@pytest.mark.parametrize("case", CASES)
def test_port_renders_instanced_scene_like_reference(case):
    g = load_golden(case)
    check_port(oracle.OraclePathTracer({k: g[k] for k in SCENE_KEYS}), g, len(g["tri_material"]))


def test_port_renders_bunny_scene_like_reference(tmp_path):
    """The bunny's instance carries the OBJ importer's world_to_object, which
    the restatement takes as the instance inverse."""
    from lux_amd import scenes

    g = load_golden("pt_bunny")
    path = tmp_path / "stanford-bunny.obj"
    path.write_bytes(lzma.decompress((GOLDEN / "obj" / "stanford-bunny.obj.xz").read_bytes()))
    d = scenes.bunny(path, int(g["width"]), int(g["height"]))
    assert "inst_inv" in d
    check_port(oracle.OraclePathTracer(d), g, int(g["ordinals"]))

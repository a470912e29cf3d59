"""CPU: OBJ import (lux_obj_load, host code of the C ABI) against the
reference's own loader -- load_obj_into_scene_with_result and its bundled
tinyobjloader, golden vectors of tests/golden/make_golden_obj.py -- bit for bit
on every array (vertices, normals, uvs, indices, material ids, boundary-edge
masks, instance transform) and count, over small files that exercise the
parser and the import rules and over the reference's own assets (LUX.obj, the
Stanford bunny). Plus what the geometry fixes independently: triangulated
polygons cover the polygon's area, smooth normals are unit length, placement
grounds the mesh at target_y and scales its extent to target_extent."""

import json
import lzma
from pathlib import Path

import numpy as np
import pytest

from conftest import GOLDEN, load_golden

CASES = ["polygons_default", "polygons_bound", "edges_default", "edges_smooth", "edges_face_only",
         "shells_welded", "shells_reverse", "shells_unbound", "lux", "bunny",
         # boundary-hole capping: closed loops over welded and source topology, reverse
         # orientation with placement, non-simple boundaries (capping skipped), a bad epsilon
         "box_cap_welded", "box_cap_source", "tube_cap_reverse", "bowtie_cap", "shells_cap", "edges_cap",
         "lux_cap", "bad_cap_epsilon"]
CAP_COUNTS = ("cap_triangles", "cap_boundary_edges", "cap_non_simple_vertices", "cap_incomplete_loops",
              "cap_degenerate_loops", "cap_degenerate_triangles")
ARRAYS = ("vertices", "normals", "uvs", "indices", "material_ids", "boundary_edges", "object_to_world",
          "world_to_object")


def obj_path(name, tmp_path):
    p = GOLDEN / "obj" / name
    if p.suffix == ".xz":
        plain = tmp_path / p.stem
        plain.write_bytes(lzma.decompress(p.read_bytes()))
        return plain
    return p


def load_case(case, tmp_path):
    from lux_amd import obj

    g = load_golden(f"obj_{case}")
    opts = json.loads(str(g["options"]))
    return g, opts, (lambda: obj.load(obj_path(str(g["obj_file"]), tmp_path), **opts))


@pytest.mark.parametrize("case", CASES)
def test_obj_import_matches_reference_loader(case, tmp_path):
    from lux_amd import abi

    g, _, run = load_case(case, tmp_path)
    if not int(g["ok"]):
        with pytest.raises(abi.LuxError) as e:
            run()
        assert e.value.status == abi.LUX_ERR_INVALID_ARGUMENT
        return
    m = run()
    for k in ARRAYS:
        assert m[k].shape == g[k].shape, k
        assert np.array_equal(np.ascontiguousarray(m[k]).view(np.uint8), np.ascontiguousarray(g[k]).view(np.uint8)), k
    for k in ("appended_default_material", "skipped_invalid", "skipped_nonfinite", "skipped_degenerate"):
        assert m[k] == int(g[k]), k
    for k in CAP_COUNTS:
        if k in g:
            assert m[k] == int(g[k]), k


def test_caps_close_the_box_and_the_tube(tmp_path):
    """What the geometry fixes, apart from the reference's files: capping the
    open box adds one fan of 4 triangles over the unit-square hole at y = 1,
    with normals +y (away from the box centre) and total area 1; capping the
    tube adds two fans of 8 whose areas are the regular octagon's, with normals
    pointing out of the tube ends, flipped under reverse_orientation."""
    g, opts, run = load_case("box_cap_welded", tmp_path)
    m = run()
    caps = m["indices"][-m["cap_triangles"]:]
    assert m["cap_triangles"] == 4
    v = m["vertices"][caps].astype(np.float64)
    assert np.allclose(v[:, :, 1], 1.0)
    area = 0.5 * np.linalg.norm(np.cross(v[:, 1] - v[:, 0], v[:, 2] - v[:, 0]), axis=1)
    assert np.isclose(area.sum(), 1.0, rtol=1e-6)
    assert np.allclose(m["normals"][caps.ravel()], [0, 1, 0])
    face_n = np.cross(v[:, 1] - v[:, 0], v[:, 2] - v[:, 0])
    assert np.all(face_n[:, 1] > 0)  # wound to match the outward normal
    _, opts, run = load_case("tube_cap_reverse", tmp_path)
    t = run()
    assert t["cap_triangles"] == 16
    caps = t["indices"][-16:]
    v = t["vertices"][caps].astype(np.float64)
    n = t["normals"][caps[:, 0]].astype(np.float64)
    centre_y = t["vertices"][:, 1].astype(np.float64).mean()
    end = np.sign(v[:, 0, 1] - centre_y)  # +1 top cap, -1 bottom cap
    assert sorted(end.tolist()) == [-1.0] * 8 + [1.0] * 8
    assert np.allclose(n[:, 1], -end)  # outward normals, flipped by reverse_orientation
    area = 0.5 * np.linalg.norm(np.cross(v[:, 1] - v[:, 0], v[:, 2] - v[:, 0]), axis=1)
    octagon = 2 * np.sqrt(2)  # unit circumradius
    scale = t["object_to_world"][0, 0]  # the placement's uniform scale is not in the object-space mesh
    for side in (-1, 1):
        assert np.isclose(area[end == side].sum(), octagon, rtol=1e-5), scale


def polygon_area(p):
    """Area of a planar polygon (vector area, Newell's formula)."""
    p = np.asarray(p, np.float64)
    return 0.5 * np.linalg.norm(np.cross(p, np.roll(p, -1, axis=0)).sum(axis=0))


def planar(p):
    p = np.asarray(p, np.float64)
    return np.linalg.matrix_rank(p[1:] - p[0], tol=1e-9) <= 2


def test_triangulation_covers_each_polygon(tmp_path):
    """The planar polygons of polygons.obj (a quad, the concave hexagon): the
    triangles of each face add up to the face's area (no overlap, no gap)."""
    g, _, run = load_case("polygons_default", tmp_path)
    m = run()
    v = m["vertices"].astype(np.float64)
    tri = v[m["indices"]]
    area = 0.5 * np.linalg.norm(np.cross(tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0]), axis=1)
    text = (GOLDEN / "obj" / "polygons.obj").read_text()
    verts = [list(map(float, ln.split()[1:4])) for ln in text.splitlines() if ln.startswith("v ")]
    faces = [ln.split("#")[0].split()[1:] for ln in text.splitlines() if ln.startswith("f ")]
    k = 0
    checked = []
    for f in faces:
        n = len(f)
        idx = [int(t.split("/")[0]) for t in f]
        pos = [verts[i - 1] if i > 0 else verts[len(verts) + i] for i in idx] if all(i <= len(verts) for i in idx) \
            else None
        got = area[k:k + n - 2].sum()
        k += n - 2
        if pos is not None and n >= 4 and planar(pos):
            assert abs(got - polygon_area(pos)) <= 1e-6 * max(1.0, polygon_area(pos)), f
            checked.append(n)
    assert k == len(area)
    assert 4 in checked and 6 in checked  # a planar quad and the concave hexagon


def test_smooth_normals_are_unit_length(tmp_path):
    for case in ("edges_smooth", "shells_welded", "lux"):
        _, _, run = load_case(case, tmp_path)
        n = run()["normals"].astype(np.float64)
        norm = np.linalg.norm(n, axis=1)
        assert np.all((np.abs(norm - 1) < 1e-6) | (norm == 0)), case


def test_placement_grounds_and_scales(tmp_path):
    for case in ("shells_welded", "lux", "bunny"):
        g, opts, run = load_case(case, tmp_path)
        m = run()
        M = m["object_to_world"].astype(np.float64)
        v = m["vertices"][np.unique(m["indices"])].astype(np.float64)
        w = v @ M[:3, :3].T + M[:3, 3]
        if opts.get("ground_to_y"):
            assert abs(w[:, 1].min() - opts.get("target_y", 0.0)) < 1e-5, case
        ext = w.max(0) - w.min(0)
        if opts.get("scale_mode") == 2:
            assert abs(ext[0] - opts["target_extent"]) < 1e-5 * opts["target_extent"], case
        if opts.get("scale_mode") == 1:
            assert abs(ext.max() - opts["target_extent"]) < 1e-5 * opts["target_extent"], case
        if opts.get("center_xz"):
            c = 0.5 * (w.max(0) + w.min(0))
            assert abs(c[0]) < 1e-5 and abs(c[2]) < 1e-5, case
        assert np.allclose(M @ m["world_to_object"].astype(np.float64), np.eye(4), atol=1e-5)


def test_import_errors(tmp_path):
    from lux_amd import abi, obj

    bad = tmp_path / "zero_index.obj"
    bad.write_text("v 0 0 0\nv 1 0 0\nv 0 1 0\nf 0 1 2\n")
    with pytest.raises(abi.LuxError):
        obj.load(bad)
    with pytest.raises(abi.LuxError):
        obj.load(tmp_path / "missing.obj")
    empty = tmp_path / "all_skipped.obj"
    empty.write_text("v 0 0 0\nv 1 1 1\nv 2 2 2\nf 1 2 3\n")
    m = obj.load(empty)
    assert len(m["indices"]) == 0 and m["skipped_degenerate"] == 1 and m["appended_default_material"] == 0
    with pytest.raises(abi.LuxError):  # a binding to a material the scene does not have
        obj.load(GOLDEN / "obj" / "shells.obj", scene_material_count=1, material_bindings={"a": 0, "b": 1})


def bunny_dict(tmp_path, width=64, height=64):
    from lux_amd import scenes

    return scenes.bunny(obj_path("stanford-bunny.obj.xz", tmp_path), width, height)


def test_bunny_scene_rebuild_matches_reference_scene(tmp_path):
    """The bunny scene assembled from the imported mesh (scenes.bunny) has the
    world-space triangle view, camera and light table of the reference's own
    make_stanford_bunny_scene (hash and tables in tests/golden/pt_bunny.npz),
    as the reference's scene code sees it."""
    import hashlib

    import oracle

    if not oracle.ref_available():
        pytest.skip("oracle/_ref not built")
    g = load_golden("pt_bunny")
    r = oracle.RefPathTracer(bunny_dict(tmp_path))
    t = r.triangles()
    digest = hashlib.sha256(b"".join(np.ascontiguousarray(t[k]).tobytes()
                                     for k in ("corners", "normals", "material"))).hexdigest()
    assert digest == str(g["triangle_view_sha256"])
    assert np.array_equal(r.camera().view(np.uint32), g["camera_vectors"].view(np.uint32))
    for k, v in r.lights().items():
        assert np.array_equal(np.asarray(v).view(np.uint32), np.asarray(g["light_" + k]).view(np.uint32)), k


def test_lux_cover_mesh_equals_reference_scene(tmp_path):
    """The logo mesh and instance of the reference's own lux-cover scene
    (make_lux_cover_scene, exported in tests/golden/pt_builtin_lux_cover.npz)
    equal the OBJ import of LUX.obj with that scene's options, bit for bit."""
    from lux_amd import obj

    g = load_golden("pt_builtin_lux_cover")
    m = obj.load(GOLDEN / "obj" / "LUX.obj", scene_material_count=5, normal_mode="generate_missing_smooth",
                 boundary_topology="welded_positions", scale_mode="x_extent", target_extent=1.85, center_xz=True,
                 ground_to_y=True, material_bindings={"L_lambert": 2, "U_glass": 3, "X_metal": 4})
    vo, to = g["vertex_offsets"], g["triangle_offsets"]
    assert list(g["inst_mesh"]) == [0, 1]
    for got, ref in ((m["vertices"], g["vertices"][vo[1]:vo[2]]), (m["normals"], g["normals"][vo[1]:vo[2]]),
                     (m["indices"], g["indices"][to[1]:to[2]]), (m["material_ids"], g["material_ids"][to[1]:to[2]]),
                     (m["object_to_world"], g["inst_m"][1]), (m["world_to_object"], g["inst_inv"][1])):
        assert np.array_equal(np.ascontiguousarray(got).view(np.uint8), np.ascontiguousarray(ref).view(np.uint8))

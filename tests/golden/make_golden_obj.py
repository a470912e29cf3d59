"""Writes OBJ-import golden vectors under tests/golden/ with the REFERENCE's own
loader (load_obj_into_scene_with_result, src/util/obj_loader.cpp, and the
tinyobjloader it bundles) through oracle/_ref.

    python tests/golden/make_golden_obj.py

Inputs: small OBJ files written here (tests/golden/obj/*.obj) that exercise the
parser, the import rules and boundary-hole capping, plus the reference's own assets
(assets/models/LUX.obj as is, stanford-bunny.obj xz-compressed). Each case
stores the options and the loaded mesh (vertices, normals, uvs, indices,
material ids, boundary-edge masks, instance transform, counts, capping counts) in
tests/golden/obj_<case>.npz; a case the reference rejects stores ok = False.
"""

import json
import lzma
import shutil
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

import oracle  # noqa: E402

OUT = Path(__file__).resolve().parent
OBJ = OUT / "obj"
ASSETS = Path("/root/reference/assets/models")

# Quads split both ways, a convex pentagon, a concave (L-shaped) hexagon that
# needs a reflex-corner skip, a face whose vertices come later in the same
# group, negative indices, every index form, comments, CRLF line ends, groups,
# objects and material labels.
POLYGONS = """# polygons and index forms\r
mtllib scene.mtl\r
o first\r
v 0 0 0\r
v 1 0 0\r
v 1 1 0\r
v 0 1 0\r
vt 0 0\r
vt 1 0\r
vt 1 1\r
vt 0 1\r
vn 0 0 1\r
vn 0 0.5 0.5\r
usemtl red\r
f 1/1/1 2/2/1 3/3/1 4/4/1\r
f -4//2 -3//2 -2//2 # trailing comment\r
v 2 0 0.25\r
v 3.5 0 0\r
v 3 1 0\r
v 2 1 0\r
f 5/1 6/2 7/3 8/4\r
g second part\r
usemtl blue\r
v 0 2 0\r
v 1 2 0.5\r
v 1.5 2.8 0\r
v 0.5 3.5 0.25\r
v -0.5 2.8 0\r
f 9 10 11 12 13\r
v 0 4 0\r
v 2 4 0\r
v 2 5 0\r
v 1 5 0\r
v 1 6 0\r
v 0 6 0\r
f 14 15 16 17 18 19\r
f 20 21 22\r
v 3 3 3\r
v 4 3 3\r
v 3 4 3\r
s 1\r
usemtl red\r
f 1 3 2\r
"""

# Faces the import skips or repairs: an index past the vertex count, an
# infinite vertex, a zero-area face, zero-length / non-finite / missing normals
# and texcoords, a face of two vertices, nan read as 0.
EDGES = """v 0 0 0
v 1 0 0
v 0 1 0
v inf 0 0
v 2 2 2
v 3 3 3
v 4 4 4
v nan 1 1
vn 0 0 0
vn inf 0 0
vn 0 0 2
vt 0.25 0.5
vt inf 0
f 1/1/3 2/2/3 3/1/3
f 1 2 4
f 5 6 7
f 1 2 9
f 1/1/1 2/1/2 3/1/3
f 1//3 2//9 3//3
f 1 2
f 8 2 3
f 1/2 3/1 2/1
"""

# Two shells sharing positions but not vertex ids (a welded seam), open
# boundaries, a mirrored corner, for the smooth-normal / welded-topology /
# reverse-orientation / placement options.
SHELLS = """v 0 0 0
v 1 0 0
v 1 1 0
v 0 1 0
v 1 0 0
v 2 0 0.5
v 2 1 0.5
v 1 1 0
v 0.5 0.5 1
usemtl a
f 1 2 3
f 1 3 4
usemtl b
f 5 6 7
f 5 7 8
f 1 2 9
f 2 3 9
f 3 4 9
f 4 1 9
"""

# Boundary holes for the capping option: a box without its top face, its faces
# not sharing vertex ids (welded positions join them; source ids leave six
# separate quads); an open octagonal tube (two holes); two triangles touching at
# one vertex (a non-simple boundary: capping is skipped).
OPEN_BOX = """v 0 0 0
v 1 0 0
v 1 0 1
v 0 0 1
v 0 0 0
v 1 0 0
v 1 1 0
v 0 1 0
v 1 0 0
v 1 0 1
v 1 1 1
v 1 1 0
v 1 0 1
v 0 0 1
v 0 1 1
v 1 1 1
v 0 0 1
v 0 0 0
v 0 1 0
v 0 1 1
usemtl bottom
f 1 2 3 4
usemtl sides
f 5 8 7 6
f 9 12 11 10
f 13 16 15 14
f 17 20 19 18
"""


def tube_obj(sides=8):
    lines = []
    for ring, y in ((0, -0.5), (1, 0.75)):
        for k in range(sides):
            a = 2 * np.pi * k / sides
            lines.append(f"v {np.cos(a):.6f} {y} {np.sin(a):.6f}")
    lines.append("usemtl wall")
    for k in range(sides):
        a, b = k + 1, (k + 1) % sides + 1
        lines.append(f"f {a} {a + sides} {b + sides} {b}")
    return "\n".join(lines) + "\n"


BOWTIE = """v 0 0 0
v 1 0 0
v 1 1 0
v -1 0 0
v -1 -1 0
f 1 2 3
f 1 4 5
"""

CASES = {
    # name: (file, options)
    "polygons_default": ("polygons.obj", dict(scene_material_count=2)),
    "polygons_bound": ("polygons.obj", dict(scene_material_count=4, material_bindings={"red": 1, "blue": 3},
                                             default_material_id=2, flip_texcoord_v=True)),
    "edges_default": ("edges.obj", dict(scene_material_count=1)),
    "edges_smooth": ("edges.obj", dict(scene_material_count=1, normal_mode=1)),
    "edges_face_only": ("edges.obj", dict(scene_material_count=1, normal_mode=2, reverse_orientation=True)),
    "shells_welded": ("shells.obj", dict(scene_material_count=3, normal_mode=1, boundary_topology=1,
                                         material_bindings={"a": 0, "b": 2}, scale_mode=1, target_extent=2.5,
                                         center_xz=True, ground_to_y=True, target_y=-0.5)),
    "shells_reverse": ("shells.obj", dict(scene_material_count=3, normal_mode=1, reverse_orientation=True,
                                          default_material_id=1, material_bindings={"a": 0, "b": 2},
                                          scale_mode=2, target_extent=0.75,
                                          object_to_world=[[0, 0, 2, 1], [0, 1, 0, -2], [-2, 0, 0, 0.5],
                                                           [0, 0, 0, 1]])),
    "shells_unbound": ("shells.obj", dict(scene_material_count=3, material_bindings={"a": 0})),
    "lux": ("LUX.obj", dict(scene_material_count=6, normal_mode=1, boundary_topology=1, scale_mode=2,
                            target_extent=1.85, center_xz=True, ground_to_y=True,
                            material_bindings={"L_lambert": 3, "U_glass": 4, "X_metal": 5})),
    "bunny": ("stanford-bunny.obj.xz", dict(scene_material_count=2, ground_to_y=True)),
    "box_cap_welded": ("open_box.obj", dict(scene_material_count=2, material_bindings={"bottom": 0, "sides": 1},
                                           cap_boundary_holes=True)),
    "box_cap_source": ("open_box.obj", dict(scene_material_count=2, material_bindings={"bottom": 0, "sides": 1},
                                           cap_boundary_holes=True, cap_boundary_topology=0, normal_mode=1)),
    "tube_cap_reverse": ("tube.obj", dict(scene_material_count=1, material_bindings={"wall": 0},
                                          cap_boundary_holes=True, reverse_orientation=True, scale_mode=1,
                                          target_extent=2.0, ground_to_y=True)),
    "bowtie_cap": ("bowtie.obj", dict(scene_material_count=1, cap_boundary_holes=True)),
    "shells_cap": ("shells.obj", dict(scene_material_count=3, material_bindings={"a": 0, "b": 2},
                                      cap_boundary_holes=True, boundary_topology=1)),
    "edges_cap": ("edges.obj", dict(scene_material_count=1, cap_boundary_holes=True)),
    "lux_cap": ("LUX.obj", dict(scene_material_count=6, normal_mode=1, boundary_topology=1, scale_mode=2,
                                target_extent=1.85, center_xz=True, ground_to_y=True, cap_boundary_holes=True,
                                material_bindings={"L_lambert": 3, "U_glass": 4, "X_metal": 5})),
    "bad_cap_epsilon": ("open_box.obj", dict(scene_material_count=2, material_bindings={"bottom": 0, "sides": 1},
                                             cap_boundary_holes=True, position_weld_epsilon=0.0)),
}


def materialise(name):
    """The path of the OBJ text for a case file (xz files are expanded to a
    temporary .obj next to them)."""
    p = OBJ / name
    if p.suffix == ".xz":
        plain = OBJ / "_expanded" / p.stem
        plain.parent.mkdir(exist_ok=True)
        plain.write_bytes(lzma.decompress(p.read_bytes()))
        return plain
    return p


def main():
    if not oracle.ref_available():
        raise SystemExit("oracle/_ref/libluxref.so missing: run `make -f oracle/ref.mk` where the reference exists")
    OBJ.mkdir(exist_ok=True)
    (OBJ / "polygons.obj").write_bytes(POLYGONS.encode())
    (OBJ / "edges.obj").write_text(EDGES)
    (OBJ / "shells.obj").write_text(SHELLS)
    (OBJ / "open_box.obj").write_text(OPEN_BOX)
    (OBJ / "tube.obj").write_text(tube_obj())
    (OBJ / "bowtie.obj").write_text(BOWTIE)
    shutil.copyfile(ASSETS / "LUX.obj", OBJ / "LUX.obj")
    (OBJ / "stanford-bunny.obj.xz").write_bytes(lzma.compress((ASSETS / "stanford-bunny.obj").read_bytes(),
                                                              preset=9))
    for case, (fname, opts) in CASES.items():
        r = oracle.ref_load_obj(materialise(fname), **opts)
        data = {k: v for k, v in r.items() if isinstance(v, np.ndarray)}
        for k in ("ok", "appended_default_material", "skipped_invalid", "skipped_nonfinite", "skipped_degenerate",
                  "cap_triangles", "cap_boundary_edges", "cap_non_simple_vertices", "cap_incomplete_loops",
                  "cap_degenerate_loops", "cap_degenerate_triangles"):
            data[k] = np.int32(r[k])
        data["obj_file"] = np.array(fname)
        data["options"] = np.array(json.dumps(opts))
        np.savez_compressed(OUT / f"obj_{case}.npz", **data)
        print(f"{case}: ok={r['ok']} vertices={len(r['vertices'])} triangles={len(r['indices'])} "
              f"skipped={r['skipped_invalid']}/{r['skipped_nonfinite']}/{r['skipped_degenerate']} "
              f"caps={r['cap_triangles']} (edges {r['cap_boundary_edges']}, non-simple "
              f"{r['cap_non_simple_vertices']}, open {r['cap_incomplete_loops']}, degenerate "
              f"{r['cap_degenerate_loops']}/{r['cap_degenerate_triangles']})")
    shutil.rmtree(OBJ / "_expanded", ignore_errors=True)


if __name__ == "__main__":
    main()

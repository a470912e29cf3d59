"""Writes the two-level (instanced scene) golden vectors under tests/golden/ from
the REFERENCE's own host code (oracle/_ref/libluxref.so): scenes built with
add_scene_mesh / add_scene_instance(make_transform) and traced with
intersect_scene_bvh / intersect_scene_bvh_skip_ref over PreparedScene::host_bvh.

Run in a container where /root/reference is mounted:
    python tests/golden/make_golden_tl.py
Each case stores the meshes, the instance matrices, the reference's computed
inverses, world bounds and first scene ordinals, the rays, and the hits (a plain
pass and a pass that skips the primitive the plain pass hit).
"""

import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

import oracle  # noqa: E402
from lux_amd import workloads as W  # noqa: E402

OUT = Path(__file__).resolve().parent


def case_mixed():
    """Three soup meshes and an empty one; rotated, scaled, mirrored,
    non-uniform, very thin (y scale 1e-3) and duplicated (exact ties) instances.
    (The reference rejects non-invertible transforms at scene validation.)"""
    meshes = [W.soup_mesh(600, 101), W.soup_mesh(250, 102, edge=(0.05, 0.2)), W.soup_mesh(40, 103, edge=(0.2, 0.5)),
              (np.zeros((0, 3), np.float32), np.zeros((0, 3), np.uint32))]
    mesh, m = W.instances_random(10, 3, 104, extent=2.5)
    extra = [(0, W.affine((1, 0, 0, 0), (-1.0, 1.0, 1.0), (0.5, -2.0, 0.0))),      # mirrored
             (1, W.affine((0.3, 0.9, 0.1, 0.2), (2.5, 0.3, 1.0), (-1.0, 1.5, 2.0))),  # non-uniform
             (2, W.affine(None, (1.0, 1e-3, 1.0), (1.0, 1.0, 1.0))),                # very thin
             (3, W.affine(None, (1.0, 1.0, 1.0), (0.0, 0.0, 0.0))),                 # empty mesh
             (int(mesh[0]), m[0].copy()),                                         # exact duplicate
             (1, W.affine(None, (1.0, 1.0, 1.0), (0.0, 0.0, 0.0)))]                # identity
    mesh = np.concatenate([mesh, np.array([e[0] for e in extra], np.int32)])
    m = np.concatenate([m, np.stack([e[1] for e in extra])])
    rays = np.concatenate([W.scene_rays(2500, 105, extent=4.0, aim=2.5), W.axis_rays(200, seed=106) * 1.0])
    rays[2500:, 0:3] = rays[2500:, 0:3] * 4.0 - 2.0
    rays[2500:, 7] = np.inf
    return meshes, mesh, m, rays


def case_single():
    """One identity instance of a soup: the two-level walk reduces to one BLAS walk."""
    meshes = [W.soup_mesh(1500, 201)]
    return meshes, np.zeros(1, np.int32), W.affine()[None], W.scene_rays(2000, 202, extent=1.5, aim=0.6) * 1.0


def case_many():
    """64 instances of 4 meshes."""
    meshes = [W.soup_mesh(200, 300 + k) for k in range(4)]
    mesh, m = W.instances_random(64, 4, 304, extent=3.0)
    return meshes, mesh, m, W.scene_rays(3000, 305)


def write(name, meshes, mesh, m, rays):
    ref = oracle.RefTwoLevel(meshes, mesh, m)
    inst = ref.instances()
    plain = ref.trace(rays)
    assert plain["overflows"] == 0
    ignore = plain["prim"].copy()  # skip what the plain pass hit (-1: nothing to skip)
    skip = ref.trace(rays, ignore)
    assert skip["overflows"] == 0
    V, vo, I, to = oracle.two_level_arrays(meshes)
    out = {"vertices": V, "vert_off": vo, "indices": I, "tri_off": to, "inst_mesh": mesh.astype(np.int32),
           "inst_m": m.astype(np.float32).reshape(-1, 16), "inv": inst["inv"], "world_bounds": inst["world_bounds"],
           "first_primitive": inst["first_primitive"], "rays": rays.astype(np.float32), "ignore": ignore}
    for k in ("prim", "t", "u", "v", "instance", "local"):
        out[f"hit_{k}"] = plain[k]
        out[f"skip_{k}"] = skip[k]
    np.savez_compressed(OUT / f"{name}.npz", **out)
    print(f"{name}: {len(meshes)} meshes, {len(mesh)} instances, {len(rays)} rays, "
          f"{int((plain['prim'] >= 0).sum())} hits, {int((skip['prim'] >= 0).sum())} hits with skip")


if __name__ == "__main__":
    write("tl_mixed", *case_mixed())
    write("tl_single", *case_single())
    write("tl_many", *case_many())

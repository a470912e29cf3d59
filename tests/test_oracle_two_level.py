"""The two-level (instanced scene) restatement against the reference's own
host code: golden vectors written by tests/golden/make_golden_tl.py from
add_scene_instance (make_transform -> inverse_affine, transform_bounds) and
intersect_scene_bvh / intersect_scene_bvh_skip_ref. Integer results and every
float bit-exact."""

import numpy as np
import pytest

import oracle
from conftest import load_golden

TL_CASES = ["tl_mixed", "tl_single", "tl_many"]


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def scene_of(g):
    tris = oracle.mesh_soups(g["vertices"], g["vert_off"], g["indices"], g["tri_off"])
    return tris, g["tri_off"]


@pytest.mark.parametrize("case", TL_CASES)
def test_inverse_affine_matches_reference(case):
    g = load_golden(case)
    for m, inv in zip(g["inst_m"], g["inv"]):
        assert np.array_equal(bits(oracle.inverse_affine(m)), bits(inv))
        assert oracle.instance_transform_valid(m, inv)


@pytest.mark.parametrize("case", TL_CASES)
def test_world_bounds_match_reference(case):
    """transform_bounds of the mesh bounds (union of its triangle boxes)."""
    g = load_golden(case)
    tris, to = scene_of(g)
    for m, mesh, wb in zip(g["inst_m"], g["inst_mesh"], g["world_bounds"]):
        t = tris[to[mesh]:to[mesh + 1]].reshape(-1, 3, 3)
        if len(t):
            box = np.concatenate([t.min(axis=(0, 1)), t.max(axis=(0, 1))]).astype(np.float32)
        else:
            box = np.array([np.inf] * 3 + [-np.inf] * 3, np.float32)
        assert np.array_equal(bits(oracle.transform_bounds(m, box)), bits(wb))


def test_transform_validation_rejects_what_the_reference_rejects():
    m = np.eye(4, dtype=np.float32)
    m[1, 1] = 0.0  # singular
    assert not oracle.instance_transform_valid(m, np.eye(4, dtype=np.float32))
    a = np.eye(4, dtype=np.float32)
    a[0, 3] = 2.0
    assert not oracle.instance_transform_valid(a, np.eye(4, dtype=np.float32))  # inconsistent pair
    p = np.eye(4, dtype=np.float32)
    p[3, 0] = 0.5  # projective row
    assert not oracle.instance_transform_valid(p, oracle.inverse_affine(p))
    assert oracle.instance_transform_valid(a, oracle.inverse_affine(a))


@pytest.mark.parametrize("case", TL_CASES)
def test_two_level_walk_matches_reference(case):
    g = load_golden(case)
    tris, to = scene_of(g)
    o = oracle.OracleTwoLevel(tris, to, g["inst_mesh"], g["inv"], g["world_bounds"])
    first = o.first_primitive()
    counts = np.diff(g["tri_off"])[g["inst_mesh"]]
    nonempty = counts > 0
    assert np.array_equal(first[nonempty], g["first_primitive"][nonempty])
    for prefix, ignore in (("hit", None), ("skip", g["ignore"])):
        r = o.trace(g["rays"], ignore)
        assert r["overflows"] == 0
        assert np.array_equal(r["prim"], g[f"{prefix}_prim"])
        assert np.array_equal(r["instance"], g[f"{prefix}_instance"])
        assert np.array_equal(r["local"], g[f"{prefix}_local"])
        for k in ("t", "u", "v"):
            assert np.array_equal(bits(r[k]), bits(g[f"{prefix}_{k}"])), k


def test_single_identity_instance_equals_single_level_walk():
    """One identity instance: the two-level walk is the BLAS walk (the
    single-level configurations rely on this)."""
    g = load_golden("tl_single")
    tris, to = scene_of(g)
    boxes = np.concatenate([tris.reshape(-1, 3, 3).min(1), tris.reshape(-1, 3, 3).max(1)], 1).astype(np.float32)
    single = oracle.OracleBvh(boxes).trace_closest(tris, g["rays"])
    assert np.array_equal(single["ids"], g["hit_prim"])
    assert np.array_equal(bits(single["t"]), bits(g["hit_t"]))


@pytest.mark.parametrize("case", TL_CASES)
def test_library_host_helpers_match_reference(case):
    """lux_hip_inverse_affine / lux_hip_transform_bounds (host code of the C ABI,
    no GPU needed) give the reference's inverse and world bounds bit for bit."""
    from lux_amd import two_level as TL

    g = load_golden(case)
    tris, to = scene_of(g)
    rec = TL.make_instances(g["inst_mesh"], g["inst_m"], tris=tris, mesh_offsets=to)
    assert np.array_equal(bits(rec["world_to_object"]), bits(g["inv"]))
    assert np.array_equal(bits(rec["world_bounds"]), bits(g["world_bounds"]))

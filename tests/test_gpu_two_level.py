"""GPU parity of the two-level (instanced) scene walk, through the C ABI
(lux_hip_two_level_*): against the reference's own host two-level walk (golden
vectors of tests/golden/make_golden_tl.py) and the CPU restatement. Ids and
every float bit-exact."""

import numpy as np
import pytest

import oracle
from conftest import load_golden
from lux_amd import abi, two_level as TL, workloads as W

pytestmark = pytest.mark.gpu
TL_CASES = ["tl_mixed", "tl_single", "tl_many"]


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def print_case(*what):
    print("case:", *what, flush=True)


def golden_scene(g):
    tris = oracle.mesh_soups(g["vertices"], g["vert_off"], g["indices"], g["tri_off"])
    rec = TL.make_instances(g["inst_mesh"], g["inst_m"], inverses=g["inv"], world_bounds=g["world_bounds"])
    return tris, g["tri_off"], rec


def check_hits(h, g, prefix):
    assert np.array_equal(h["primitive_id"], g[f"{prefix}_prim"])
    assert np.array_equal(h["instance_id"], g[f"{prefix}_instance"])
    assert np.array_equal(h["local_primitive_id"], g[f"{prefix}_local"])
    for k in ("t", "u", "v"):
        assert np.array_equal(bits(h[k]), bits(g[f"{prefix}_{k}"])), k


@pytest.mark.parametrize("case", TL_CASES)
def test_two_level_closest_matches_reference(gpu, case):
    print_case("two-level closest", case)
    g = load_golden(case)
    tris, off, rec = golden_scene(g)
    s = TL.TwoLevelScene(tris, off, rec)
    check_hits(s.intersect(g["rays"]), g, "hit")
    check_hits(s.intersect(g["rays"], g["ignore"]), g, "skip")
    info = s.info()
    assert info["tlas_instances"] == int((g["world_bounds"][:, 0] <= g["world_bounds"][:, 3]).sum())
    assert info["scene_primitives"] == int(np.diff(g["tri_off"])[g["inst_mesh"]].sum())
    s.close()


@pytest.mark.parametrize("case", TL_CASES)
def test_two_level_occluded_is_closest_found(gpu, case):
    print_case("two-level occluded", case)
    g = load_golden(case)
    tris, off, rec = golden_scene(g)
    s = TL.TwoLevelScene(tris, off, rec)
    assert np.array_equal(s.occluded(g["rays"]).astype(bool), g["hit_prim"] >= 0)
    assert np.array_equal(s.occluded(g["rays"], g["ignore"]).astype(bool), g["skip_prim"] >= 0)
    s.close()


def test_library_computed_instances_give_the_same_hits(gpu):
    """Instances whose inverse and world bounds come from the library's own
    inverse_affine / transform_bounds."""
    print_case("two-level library instances")
    g = load_golden("tl_mixed")
    tris = oracle.mesh_soups(g["vertices"], g["vert_off"], g["indices"], g["tri_off"])
    rec = TL.make_instances(g["inst_mesh"], g["inst_m"], tris=tris, mesh_offsets=g["tri_off"])
    s = TL.TwoLevelScene(tris, g["tri_off"], rec)
    check_hits(s.intersect(g["rays"]), g, "hit")
    s.close()


def test_set_instances_equals_fresh_build_and_oracle(gpu):
    """Moving the instances rebuilds only the TLAS: the hits equal a scene built
    from scratch with the new instances and the oracle walk."""
    print_case("two-level set_instances")
    g = load_golden("tl_many")
    tris = oracle.mesh_soups(g["vertices"], g["vert_off"], g["indices"], g["tri_off"])
    off = g["tri_off"]
    s = TL.TwoLevelScene(tris, off, TL.make_instances(g["inst_mesh"], g["inst_m"], tris=tris, mesh_offsets=off))
    mesh2, m2 = W.instances_random(len(g["inst_mesh"]), len(off) - 1, 999, extent=2.0)
    rec2 = TL.make_instances(mesh2, m2, tris=tris, mesh_offsets=off)
    s.set_instances(rec2)
    moved = s.intersect(g["rays"])
    fresh = TL.TwoLevelScene(tris, off, rec2)
    again = fresh.intersect(g["rays"])
    assert np.array_equal(moved.view(np.uint8), again.view(np.uint8))
    o = oracle.OracleTwoLevel(tris, off, mesh2, rec2["world_to_object"], rec2["world_bounds"]).trace(g["rays"])
    assert np.array_equal(moved["primitive_id"], o["prim"])
    assert np.array_equal(bits(moved["t"]), bits(o["t"]))
    assert int((moved["primitive_id"] >= 0).sum()) > 50
    s.close()
    fresh.close()


def test_two_level_rejects_what_scene_validation_rejects(gpu):
    print_case("two-level validation")
    g = load_golden("tl_single")
    tris, off, rec = golden_scene(g)
    bad = rec.copy()
    bad["object_to_world"][0, 5] = 0.0  # singular
    with pytest.raises(abi.LuxError) as e:
        TL.TwoLevelScene(tris, off, bad)
    assert e.value.status == abi.LUX_ERR_INVALID_ARGUMENT
    bad = rec.copy()
    bad["mesh_id"][0] = 5
    with pytest.raises(abi.LuxError) as e:
        TL.TwoLevelScene(tris, off, bad)
    assert e.value.status == abi.LUX_ERR_INVALID_ARGUMENT
    s = TL.TwoLevelScene(tris, off, rec)
    with pytest.raises(abi.LuxError) as e:
        s.set_instances(np.concatenate([rec, rec]))
    assert e.value.status == abi.LUX_ERR_MISMATCH
    check_hits(s.intersect(g["rays"]), g, "hit")  # still usable
    s.close()


def test_two_level_without_instances_misses_everything(gpu):
    print_case("two-level empty")
    g = load_golden("tl_single")
    tris, off, _ = golden_scene(g)
    s = TL.TwoLevelScene(tris, off, np.zeros(0, TL.INSTANCE))
    h = s.intersect(g["rays"])
    assert np.all(h["primitive_id"] == -1) and np.all(np.isinf(h["t"]))
    assert not s.occluded(g["rays"]).any()
    s.close()


def test_instanced_microbench_inputs_match_oracle(gpu):
    """The instanced microbench scene (4 meshes of 65 536 triangles, 64 instances):
    a 32 768-ray sample against the oracle walk; occlusion of every ray equals
    closest-found."""
    print_case("two-level microbench inputs")
    inp = W.two_level_inputs(n_rays=1 << 18)
    tris = np.concatenate([v[i].reshape(-1, 9) for v, i in inp["meshes"]]).astype(np.float32)
    off = np.concatenate([[0], np.cumsum([len(i) for _, i in inp["meshes"]])]).astype(np.int32)
    rec = TL.make_instances(inp["inst_mesh"], inp["inst_m"], tris=tris, mesh_offsets=off)
    s = TL.TwoLevelScene(tris, off, rec)
    h = s.intersect(inp["rays"])
    sample = np.random.default_rng(3).choice(len(inp["rays"]), 1 << 15, replace=False)
    o = oracle.OracleTwoLevel(tris, off, inp["inst_mesh"], rec["world_to_object"], rec["world_bounds"])
    r = o.trace(inp["rays"][sample])
    assert r["overflows"] == 0
    assert np.array_equal(h["primitive_id"][sample], r["prim"])
    assert np.array_equal(bits(h["t"][sample]), bits(r["t"]))
    assert np.array_equal(h["instance_id"][sample], r["instance"])
    occ = s.occluded(inp["rays"])
    assert np.array_equal(occ.astype(bool), h["primitive_id"] >= 0)
    assert (h["primitive_id"] >= 0).mean() > 0.2
    # 2^18 rays are traced in ray order (octant, then origin); chunks below the
    # ordering threshold run in input order and must give the same bits, also
    # with a per-ray ignored primitive
    rays = inp["rays"]
    ign = np.where(h["primitive_id"] >= 0, h["primitive_id"], -1).astype(np.int32)
    h_ign, occ_ign = s.intersect(rays, ignore=ign), s.occluded(rays, ignore=ign)
    step = 60_000
    parts = [(s.intersect(rays[k:k + step]), s.intersect(rays[k:k + step], ignore=ign[k:k + step]),
              s.occluded(rays[k:k + step], ignore=ign[k:k + step])) for k in range(0, len(rays), step)]
    for f in ("primitive_id", "instance_id", "local_primitive_id", "t", "u", "v"):
        a = np.concatenate([p[0][f] for p in parts])
        b = np.concatenate([p[1][f] for p in parts])
        assert np.array_equal(a.view(np.uint32), np.ascontiguousarray(h[f]).view(np.uint32)), f
        assert np.array_equal(b.view(np.uint32), np.ascontiguousarray(h_ign[f]).view(np.uint32)), f
    assert np.array_equal(np.concatenate([p[2] for p in parts]), occ_ign)
    assert not np.array_equal(h_ign["primitive_id"], h["primitive_id"])
    s.close()


def test_two_level_ordered_walk_ties_and_deep_stack(gpu):
    """Closest hits that stress the near-child-first two-level walk: two
    instances of one triangulated grid under the same transform (every hit is
    an exact t tie between instances, and the grid's shared edges tie within
    one), and a mesh of 2^18 parallel triangles stacked along the rays, deep
    enough that a lane's stack overflows and the ray is walked again in
    preorder. Ids, instances and t equal the oracle's walk."""
    print_case("two-level ties and deep stack")
    m = 1 << 18
    x = ((np.arange(m) + 0.5) / m).astype(np.float32)
    stack_tris = np.zeros((m, 9), np.float32)
    stack_tris[:, 0] = stack_tris[:, 3] = stack_tris[:, 6] = x
    stack_tris[:, 4] = 2.0
    stack_tris[:, 8] = 2.0
    k = 16
    g1 = np.arange(k + 1, dtype=np.float32) / k
    grid = []
    for i in range(k):
        for j in range(k):
            a, b = (g1[i], g1[j], 0.5), (g1[i + 1], g1[j], 0.5)
            c, d = (g1[i + 1], g1[j + 1], 0.5), (g1[i], g1[j + 1], 0.5)
            grid += [(a, b, c), (a, c, d)]
    grid = np.asarray(grid, np.float32).reshape(-1, 9)
    tris = np.concatenate([stack_tris, grid])
    off = np.array([0, m, m + len(grid)], np.int32)
    eye = np.eye(4, dtype=np.float32)
    shifted_x, raised = eye.copy(), eye.copy()
    shifted_x[0, 3] = 2.0
    raised[2, 3] = 5.0
    inst_mesh = np.array([0, 1, 0, 1], np.int32)
    inst_m = np.stack([eye, raised, shifted_x, raised])
    rec = TL.make_instances(inst_mesh, inst_m, tris=tris, mesh_offsets=off)
    rng = np.random.default_rng(12)
    n = 64
    rays = np.zeros((3 * n + (2 * k + 1) ** 2, 9), np.float32)
    rays[:, 7] = np.inf
    rays[: 2 * n, 1:3] = rng.uniform(0.3, 0.7, (2 * n, 2))
    rays[:n, 0], rays[:n, 3] = -0.5, 1.0  # +x into the first stack: nearest is its leftmost triangle
    rays[n:2 * n, 0], rays[n:2 * n, 3] = 3.5, -1.0  # -x into the shifted stack: its rightmost
    rays[2 * n:3 * n, 0:2] = rng.uniform(0.1, 0.9, (n, 2))
    rays[2 * n:3 * n, 2], rays[2 * n:3 * n, 5] = 4.0, 1.0  # up into the two coincident grids
    pts = np.arange(2 * k + 1, dtype=np.float32) / (2 * k)
    px, py = np.meshgrid(pts, pts)
    rays[3 * n:, 0], rays[3 * n:, 1] = px.ravel(), py.ravel()
    rays[3 * n:, 2], rays[3 * n:, 5] = 4.0, 1.0  # grid vertices and edge midpoints
    s = TL.TwoLevelScene(tris, off, rec)
    h = s.intersect(rays)
    o = oracle.OracleTwoLevel(tris, off, inst_mesh, rec["world_to_object"], rec["world_bounds"]).trace(rays)
    assert o["overflows"] == 0
    assert np.array_equal(h["primitive_id"], o["prim"])
    assert np.array_equal(h["instance_id"], o["instance"])
    assert np.array_equal(bits(h["t"]), bits(o["t"]))
    assert np.array_equal(h["local_primitive_id"][:n], np.zeros(n, np.int32))
    assert np.array_equal(h["local_primitive_id"][n:2 * n], np.full(n, m - 1, np.int32))
    assert (h["primitive_id"][2 * n:] >= 0).mean() > 0.99
    s.close()


def test_two_level_rounding_case_matches_reference(gpu):
    """The rounding4 case (a triangle whose hit t lies below its own leaf box's
    entry) as one mesh under an identity instance: the TLAS leaf adds nothing,
    so the two-level walk must return the reference host LBVH's hits, which the
    near-child-first walk alone would not (tests/test_ordered_model.py)."""
    print_case("two-level rounding4")
    g = load_golden("lbvh_rounding4")
    tris = np.ascontiguousarray(g["tri_by_id"], np.float32)
    off = np.array([0, len(tris)], np.int32)
    rec = TL.make_instances([0], np.eye(4, dtype=np.float32)[None], tris=tris, mesh_offsets=off)
    s = TL.TwoLevelScene(tris, off, rec)
    h = s.intersect(g["rays"])
    assert np.array_equal(h["primitive_id"], g["hit_ids"])
    assert np.array_equal(bits(h["t"]), bits(g["hit_t"]))
    s.close()


def test_culling_boxes_keep_hits_of_round_meshes_and_loose_inverses(gpu):
    """Round meshes (geodesic spheres, whose culling boxes are well inside the
    reference's transform_bounds boxes) under random rotations and scales, with
    each instance inverse perturbed by up to 1e-5 per entry (still inside the
    scene validation's 1e-4), and rays aimed at points just inside and just
    outside the spheres' silhouettes: closest hits (ids and t) equal the oracle's
    walk over the reference boxes with the same inverses, and occlusion equals
    closest-found, through the ray-ordered path (2^17 rays)."""
    from lux_amd import scenes

    print_case("two-level culling boxes, round meshes")
    rng = np.random.default_rng(11)
    meshes = []
    for freq in (12, 7):
        v, i = scenes.geodesic_sphere(freq)
        meshes.append((np.asarray(v, np.float32) * np.float32(0.9), np.asarray(i, np.uint32)))
    tris = np.concatenate([v[i].reshape(-1, 9) for v, i in meshes]).astype(np.float32)
    off = np.concatenate([[0], np.cumsum([len(i) for _, i in meshes])]).astype(np.int32)
    inst_mesh, inst_m = W.instances_random(48, len(meshes), seed=12, extent=2.5, scale=(0.3, 0.8))
    inv = np.stack([TL.inverse_affine(m).reshape(4, 4) for m in inst_m]).astype(np.float32)
    inv[:, :3, :] += rng.uniform(-1e-5, 1e-5, (len(inv), 3, 4)).astype(np.float32)
    for m, iv in zip(inst_m, inv):
        assert oracle.instance_transform_valid(m, iv)
    rec = TL.make_instances(inst_mesh, inst_m, tris=tris, mesh_offsets=off, inverses=inv)
    n = 1 << 17
    # targets on the spheres' silhouettes as seen from each origin, pushed in or out by 1e-4 .. 1e-2
    k = rng.integers(0, len(inst_m), n)
    centers = np.stack([m.reshape(4, 4)[:3, 3] for m in inst_m])[k].astype(np.float64)
    radius = 0.9 * np.stack([np.linalg.svd(m.reshape(4, 4)[:3, :3], compute_uv=False).max() for m in inst_m])[k]
    o = rng.uniform(-4, 4, (n, 3))
    to_c = centers - o
    dist = np.linalg.norm(to_c, axis=1, keepdims=True)
    side = rng.standard_normal((n, 3))
    side -= (side * to_c).sum(1, keepdims=True) / dist ** 2 * to_c
    side /= np.linalg.norm(side, axis=1, keepdims=True)
    scale = 1 + rng.choice([-1, 1], n) * 10 ** rng.uniform(-4, -2, n)
    target = centers + side * (radius * scale)[:, None]
    d = target - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.zeros((n, 9), np.float32)
    rays[:, 0:3] = o
    rays[:, 3:6] = d
    rays[:, 7] = np.inf
    s = TL.TwoLevelScene(tris, off, rec)
    h = s.intersect(rays)
    r = oracle.OracleTwoLevel(tris, off, inst_mesh, rec["world_to_object"], rec["world_bounds"]).trace(rays)
    assert r["overflows"] == 0
    assert np.array_equal(h["primitive_id"], r["prim"])
    assert np.array_equal(bits(h["t"]), bits(r["t"]))
    assert np.array_equal(h["instance_id"], r["instance"])
    hit = h["primitive_id"] >= 0
    assert 0.2 < hit.mean() < 0.95
    assert np.array_equal(s.occluded(rays).astype(bool), hit)
    s.close()

"""GPU: the wavefront path tracer on instanced scenes (meshes placed by affine
instances, lux_hip_scene_create_instanced) against the reference's CPU path
tracer on the same scene (golden files of tests/golden/make_golden_pt.py).

The shading view of every scene ordinal (world corners, world normals, material
after the override), the camera and the light table are bit-exact; films and
per-sample radiance carry the tolerances of the one-mesh renders (cosf / sinf /
powf round differently on the GPU). A one-mesh scene under one identity
instance renders bit for bit like the one-mesh scene; batches and the server
mixing both kinds give each film bit for bit as rendered alone."""

import ctypes

import numpy as np
import pytest

from conftest import load_golden

pytestmark = pytest.mark.gpu

CASES = ["pt_instanced", "pt_instanced_opaque"]
RELRMSE_SMALL = 2e-2    # 40x40 at 8 spp, as the one-mesh renders
FRAC_PIXELS_OFF = 0.02  # fraction of pixels with |delta| > 1e-3
INST_KEYS = ("vertices", "indices", "normals", "uvs", "material_ids", "materials", "camera", "vertex_offsets",
             "triangle_offsets", "inst_mesh", "inst_m", "inst_override")


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def scene_of(g):
    return {k: g[k] for k in INST_KEYS}


@pytest.fixture(scope="module")
def rd(gpu):
    from lux_amd import render

    return render


@pytest.mark.parametrize("case", CASES)
def test_shading_view_and_tables_bit_exact(rd, case):
    print("case: instanced tables", case, flush=True)
    g = load_golden(case)
    sc = rd.Scene(scene_of(g))
    assert sc.instanced and sc.triangle_count == len(g["tri_material"])
    tv = sc.triangles()
    assert np.array_equal(bits(tv["corners"]), bits(g["tri_corners"]))
    assert np.array_equal(bits(tv["normals"]), bits(g["tri_normals"]))
    assert np.array_equal(tv["material"], g["tri_material"])
    assert np.array_equal(bits(sc.camera()), bits(g["camera_vectors"][:12]))
    lt = sc.lights()
    for k in ("prim", "area", "prob", "alias", "pmf", "primitive_light_index"):
        assert np.array_equal(np.asarray(lt[k]).view(np.uint32), np.asarray(g["light_" + k]).view(np.uint32)), k
    sc.close()


def check_film(rd, film, ref, spp, tol):
    a, c = film / spp, ref / spp
    err = rd.rel_rmse(a, c)
    off = float((np.abs(a - c).max(axis=2) > 1e-3).mean())
    print(f"relRMSE={err:.3e} max|d|={np.abs(a - c).max():.3e} frac>1e-3={off:.4f}", flush=True)
    assert err < tol
    assert off < FRAC_PIXELS_OFF


@pytest.mark.parametrize("case", CASES)
def test_instanced_render_matches_reference(rd, case):
    """pt_instanced has a transparent sheet (pass-through shadow rays, the
    closest-hit shadow loop over the two-level walk); pt_instanced_opaque uses
    the any-hit shadow walk."""
    print("case: instanced render", case, flush=True)
    g = load_golden(case)
    sc = rd.Scene(scene_of(g))
    w, h, spp, depth, seed = int(g["width"]), int(g["height"]), int(g["spp"]), int(g["depth"]), int(g["seed"])
    films = [sc.render(w, h, spp, depth, seed=seed, mode=mode) for mode in (rd.WAVEFRONT, rd.PLAIN)]
    for film in films:
        check_film(rd, film, g["film_sum"], spp, RELRMSE_SMALL)
    assert np.array_equal(bits(films[0]), bits(films[1]))  # both schedules: same per-path arithmetic
    sc.close()


@pytest.mark.parametrize("case", CASES)
def test_instanced_per_sample_radiance(rd, case):
    print("case: instanced per-sample", case, flush=True)
    g = load_golden(case)
    sc = rd.Scene(scene_of(g))
    px, sm, ref = g["sample_pixels"], g["sample_index"], g["sample_radiance"]
    close = 0
    for k in range(int(g["spp"])):
        film = sc.render(40, 40, 1, int(g["depth"]), seed=int(g["seed"]), sample_offset=k).reshape(-1, 3)
        sel = sm == k
        got, exp = film[px[sel]], ref[sel]
        close += int(np.all(np.abs(got - exp) <= 1e-4 * np.maximum(np.abs(exp), 1e-2), axis=1).sum())
    frac = close / len(px)
    print(f"samples within 1e-4: {frac:.4f}", flush=True)
    assert frac > 0.95
    sc.close()


def one_instance(d):
    """A one-mesh scene dict as an instanced scene with one identity instance."""
    n = len(d["vertices"])
    t = len(d["indices"])
    out = dict(d)
    out.update(vertex_offsets=np.array([0, n], np.int32), triangle_offsets=np.array([0, t], np.int32),
               inst_mesh=np.zeros(1, np.int32), inst_m=np.eye(4, dtype=np.float32)[None],
               inst_override=np.full(1, -1, np.int32))
    return out


def test_identity_instance_renders_like_the_one_mesh_scene(rd):
    from lux_amd import scenes

    print("case: identity instance vs one-mesh", flush=True)
    for d, (w, h, spp, depth) in ((scenes.mixed_materials(32, 32), (32, 32, 4, 6)),
                                  (scenes.displaced_icosphere(freq=10, seed=0x4C555A, width=32, height=32),
                                   (32, 32, 4, 6))):
        a = rd.Scene(d)
        b = rd.Scene(one_instance(d))
        assert b.instanced and b.triangle_count == a.triangle_count
        ta, tb = a.triangles(), b.triangles()
        for k in ta:
            assert np.array_equal(np.asarray(ta[k]).view(np.uint32), np.asarray(tb[k]).view(np.uint32)), k
        fa = a.render(w, h, spp, depth, seed=5)
        fb = b.render(w, h, spp, depth, seed=5)
        print(f"identical={np.array_equal(bits(fa), bits(fb))} relRMSE={rd.rel_rmse(fa / spp, fb / spp):.3e}",
              flush=True)
        assert np.array_equal(bits(fa), bits(fb))
        a.close()
        b.close()


def moved_instances(g, seed):
    from lux_amd import workloads as W

    rng = np.random.default_rng(seed)
    m = g["inst_m"].copy()
    for i in range(1, len(m)):  # the room stays; the others move and turn
        q = rng.standard_normal(4)
        s = np.abs(np.diag(m[i][:3, :3])).max()
        m[i] = W.affine(q, (s, s * rng.uniform(0.5, 1.5), s), m[i][:3, 3] + rng.uniform(-0.1, 0.1, 3))
    return m


def test_set_instances_equals_fresh_scene(rd):
    """Moving the instances (and switching one to another mesh, which changes
    the ordinal count) renders bit for bit like a scene created with them."""
    print("case: instanced set_instances", flush=True)
    g = load_golden("pt_instanced")
    d = scene_of(g)
    sc = rd.Scene(d)
    before = sc.render(40, 40, 4, 6, seed=9)
    m2 = moved_instances(g, 7)
    mesh2 = g["inst_mesh"].copy()
    mesh2[4] = 1  # the block becomes a sphere: more scene ordinals
    ov2 = g["inst_override"].copy()
    ov2[1] = 6
    sc.set_instances(mesh2, m2, ov2)
    d2 = dict(d, inst_m=m2, inst_mesh=mesh2, inst_override=ov2)
    fresh = rd.Scene(d2)
    assert sc.triangle_count == fresh.triangle_count > len(g["tri_material"])
    ta, tb = sc.triangles(), fresh.triangles()
    for k in ta:
        assert np.array_equal(np.asarray(ta[k]).view(np.uint32), np.asarray(tb[k]).view(np.uint32)), k
    la, lb = sc.lights(), fresh.lights()
    for k in la:
        assert np.array_equal(np.asarray(la[k]).view(np.uint32), np.asarray(lb[k]).view(np.uint32)), k
    a = sc.render(40, 40, 4, 6, seed=9)
    b = fresh.render(40, 40, 4, 6, seed=9)
    assert np.array_equal(bits(a), bits(b))
    assert not np.array_equal(bits(a), bits(before))
    sc.close()
    fresh.close()


def test_mixed_kind_batch_equals_standalone(rd):
    """A batch over one-mesh and instanced scenes: batches hold one traversal
    kind each, and every film equals its standalone render."""
    from lux_amd import scenes

    print("case: mixed-kind batch", flush=True)
    saved = rd.get_tuning()
    try:
        dicts = [scene_of(load_golden("pt_instanced")), scenes.cornell(24, 24),
                 scene_of(load_golden("pt_instanced_opaque")), scenes.mixed_materials(32, 24)]
        jobs = [dict(width=40, height=40, spp=3, depth=6, seed=1), dict(width=24, height=24, spp=5, depth=4, seed=2),
                dict(width=40, height=40, spp=4, depth=5, seed=3, frame=1),
                dict(width=32, height=24, spp=2, depth=6, seed=4)]
        scs = [rd.Scene(d) for d in dicts]
        alone = [s.render(j["width"], j["height"], j["spp"], j["depth"],
                          **{k: v for k, v in j.items() if k not in ("width", "height", "spp", "depth")})
                 for s, j in zip(scs, jobs)]
        for capacity in (0, 1500):
            if capacity:
                rd.set_tuning(capacity=capacity)
            films = rd.batch_render(scs, jobs)
            for k, (a, f) in enumerate(zip(alone, films)):
                assert np.array_equal(bits(a), bits(f)), (capacity, k)
        for s in scs:
            s.close()
    finally:
        rd.set_tuning(**saved)


def test_server_serves_instanced_scenes(rd):
    from lux_amd import scenes

    print("case: server with instanced scenes", flush=True)
    dicts = [scene_of(load_golden("pt_instanced")), scenes.cornell(32, 32)]
    scs = [rd.Scene(d) for d in dicts]
    reqs = [(0, dict(width=40, height=40, spp=12, depth=6, seed=1), 0.0),
            (1, dict(width=32, height=32, spp=20, depth=4, seed=2), 0.001),
            (0, dict(width=40, height=40, spp=5, depth=4, seed=3), 0.002)]
    alone = [scs[i].render(j["width"], j["height"], j["spp"], j["depth"], seed=j["seed"]) for i, j, _ in reqs]
    sv = rd.Server(slice_spp=4)
    films, lat, wall = sv.run([scs[i] for i, _, _ in reqs], [j for _, j, _ in reqs], [a for *_, a in reqs])
    for k, (a, f) in enumerate(zip(alone, films)):
        assert np.array_equal(bits(a), bits(f)), k
    from lux_amd import abi

    with pytest.raises(abi.LuxError) as e:  # vertex updates do not apply to an instanced scene
        sv.submit(scs[0], reqs[0][1], vertices=dicts[0]["vertices"])
    assert e.value.status == abi.LUX_ERR_INVALID_ARGUMENT
    # an admitted, unfinished request pins the placement
    rid = sv.submit(scs[0], dict(width=40, height=40, spp=64, depth=6, seed=5))
    done = sv.poll()
    if rid not in done:
        with pytest.raises(abi.LuxError) as e:
            scs[0].set_instances(dicts[0]["inst_mesh"], dicts[0]["inst_m"], dicts[0]["inst_override"])
        assert e.value.status == abi.LUX_ERR_INVALID_ARGUMENT
    while rid not in done:
        done += sv.poll(block=True)
    film, _ = sv.result(rid)
    assert np.array_equal(bits(film), bits(scs[0].render(40, 40, 64, 6, seed=5)))
    scs[0].set_instances(dicts[0]["inst_mesh"], dicts[0]["inst_m"], dicts[0]["inst_override"])  # idle again
    sv.close()


def test_instanced_scene_validation(rd):
    from lux_amd import abi

    print("case: instanced validation", flush=True)
    g = load_golden("pt_instanced_opaque")
    d = scene_of(g)

    def rejects(status, **change):
        with pytest.raises(abi.LuxError) as e:
            rd.Scene(dict(d, **change))
        assert e.value.status == status, change

    bad = g["inst_override"].copy()
    bad[2] = len(g["materials"])
    rejects(abi.LUX_ERR_INVALID_ARGUMENT, inst_override=bad)
    bad = g["inst_mesh"].copy()
    bad[1] = 3
    rejects(abi.LUX_ERR_INVALID_ARGUMENT, inst_mesh=bad)
    bad = g["inst_m"].copy()
    bad[2][1] = 0.0  # singular
    rejects(abi.LUX_ERR_INVALID_ARGUMENT, inst_m=bad)
    bad = g["indices"].copy()
    bad[0, 0] = g["vertex_offsets"][1]  # past mesh 0's own vertices
    rejects(abi.LUX_ERR_INVALID_ARGUMENT, indices=bad)
    rejects(abi.LUX_ERR_INVALID_ARGUMENT, inst_mesh=np.zeros(0, np.int32), inst_m=np.zeros((0, 4, 4), np.float32),
            inst_override=np.zeros(0, np.int32))
    sc = rd.Scene(d)
    with pytest.raises(abi.LuxError) as e:
        sc.set_instances(g["inst_mesh"][:-1], g["inst_m"][:-1])
    assert e.value.status == abi.LUX_ERR_MISMATCH
    with pytest.raises(abi.LuxError) as e:
        sc.update_vertices(d["vertices"])
    assert e.value.status == abi.LUX_ERR_INVALID_ARGUMENT
    h = ctypes.c_void_p()
    with pytest.raises(abi.LuxError) as e:
        abi.call("lux_hip_scene_bvh", sc.handle, ctypes.byref(h))
    assert e.value.status == abi.LUX_ERR_INVALID_ARGUMENT
    # the scene is still usable
    tv = sc.triangles()
    assert np.array_equal(tv["material"], g["tri_material"])
    sc.close()


def test_bunny_scene_matches_reference(rd, tmp_path):
    """The reference's Stanford-bunny scene, rebuilt from the OBJ fixture by
    the OBJ import: the renderer's world-space triangle view hashes to the
    reference's, camera and light table are bit-exact, the film and the
    per-sample radiance match the reference's CPU path tracer."""
    import hashlib
    import lzma

    from conftest import GOLDEN
    from lux_amd import scenes

    print("case: bunny scene", flush=True)
    g = load_golden("pt_bunny")
    src = GOLDEN / "obj" / "stanford-bunny.obj.xz"
    path = tmp_path / "stanford-bunny.obj"
    path.write_bytes(lzma.decompress(src.read_bytes()))
    w, h, spp, depth, seed = int(g["width"]), int(g["height"]), int(g["spp"]), int(g["depth"]), int(g["seed"])
    sc = rd.Scene(scenes.bunny(path, w, h))
    assert sc.triangle_count == int(g["ordinals"])
    t = sc.triangles()
    digest = hashlib.sha256(b"".join(np.ascontiguousarray(t[k]).tobytes()
                                     for k in ("corners", "normals", "material"))).hexdigest()
    assert digest == str(g["triangle_view_sha256"])
    assert np.array_equal(bits(sc.camera()), bits(g["camera_vectors"][:12]))
    lt = sc.lights()
    for k in ("prim", "area", "prob", "alias", "pmf", "primitive_light_index"):
        assert np.array_equal(np.asarray(lt[k]).view(np.uint32), np.asarray(g["light_" + k]).view(np.uint32)), k
    check_film(rd, sc.render(w, h, spp, depth, seed=seed), g["film_sum"], spp, RELRMSE_SMALL)
    px, sm, ref = g["sample_pixels"], g["sample_index"], g["sample_radiance"]
    close = 0
    for k in range(spp):
        film = sc.render(w, h, 1, depth, seed=seed, sample_offset=k).reshape(-1, 3)
        sel = sm == k
        close += int(np.all(np.abs(film[px[sel]] - ref[sel]) <= 1e-4 * np.maximum(np.abs(ref[sel]), 1e-2),
                            axis=1).sum())
    print(f"samples within 1e-4: {close / len(px):.4f}", flush=True)
    assert close / len(px) > 0.95
    sc.close()


@pytest.mark.parametrize("case", ["pt_builtin_direct_light", "pt_builtin_transparent_filter",
                                  "pt_builtin_lux_cover"])
def test_builtin_cli_scene_matches_reference(rd, case):
    """The reference CLI's built-in scenes as its own scene code builds them
    (exported meshes, instances, materials, camera): identical triangle view,
    light table and camera, and a render that matches its CPU path tracer."""
    import hashlib

    print("case: built-in scene", case, flush=True)
    g = load_golden(case)
    keys = ("vertices", "normals", "uvs", "vertex_offsets", "indices", "material_ids", "triangle_offsets",
            "inst_mesh", "inst_m", "inst_inv", "inst_override", "materials", "camera", "camera_vectors")
    sc = rd.Scene({k: g[k] for k in keys})
    t = sc.triangles()
    digest = hashlib.sha256(b"".join(np.ascontiguousarray(t[k]).tobytes()
                                     for k in ("corners", "normals", "material"))).hexdigest()
    assert digest == str(g["triangle_view_sha256"])
    assert np.array_equal(bits(sc.camera()), bits(g["camera_vectors"][:12]))
    lt = sc.lights()
    for k in ("prim", "area", "prob", "alias", "pmf", "primitive_light_index"):
        assert np.array_equal(np.asarray(lt[k]).view(np.uint32), np.asarray(g["light_" + k]).view(np.uint32)), k
    w, h, spp, depth, seed = int(g["width"]), int(g["height"]), int(g["spp"]), int(g["depth"]), int(g["seed"])
    check_film(rd, sc.render(w, h, spp, depth, seed=seed), g["film_sum"], spp, RELRMSE_SMALL)
    px, sm, ref = g["sample_pixels"], g["sample_index"], g["sample_radiance"]
    close = 0
    for k in range(spp):
        film = sc.render(w, h, 1, depth, seed=seed, sample_offset=k).reshape(-1, 3)
        sel = sm == k
        close += int(np.all(np.abs(film[px[sel]] - ref[sel]) <= 1e-4 * np.maximum(np.abs(ref[sel]), 1e-2),
                            axis=1).sum())
    print(f"samples within 1e-4: {close / len(px):.4f}", flush=True)
    assert close / len(px) > 0.95
    sc.close()

"""GPU: the wavefront path tracer against the reference CPU path tracer's golden
renders. Camera vectors and light alias tables are bit-exact; films are
compared by per-pixel relative RMSE (sqrt(mean((g-c)^2/(c^2+1e-2))) over
spp-averaged RGB), because cosf/sinf/powf round differently on the GPU and
perturb bounce directions by a few ULP. Plain and wavefront schedules, and any
batch size, must give bit-identical films (same per-path arithmetic, samples
added per pixel in sample order)."""

import numpy as np
import pytest

from conftest import load_golden

pytestmark = pytest.mark.gpu

# stated tolerances (spp-averaged images)
RELRMSE_SMALL = 2e-2    # 32x32 / 40x40 images at 4-8 spp: rare branch flips dominate
RELRMSE_C1 = 2e-3       # 256x256 at 64 spp
FRAC_PIXELS_OFF = 0.02  # fraction of pixels with |delta| > 1e-3


def scene_of(g):
    return {k: g[k] for k in ("vertices", "indices", "normals", "uvs", "material_ids", "materials", "camera")}


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


@pytest.fixture(scope="module")
def rd(gpu):
    from lux_amd import render

    return render


@pytest.mark.parametrize("case", ["pt_cornell_32", "pt_icosphere", "pt_cornell_c1", "pt_mixed"])
def test_scene_tables_bit_exact(rd, case):
    print("case: tables", case, flush=True)
    g = load_golden(case)
    sc = rd.Scene(scene_of(g))
    assert np.array_equal(bits(sc.camera()), bits(g["camera_vectors"][:12]))
    lt = sc.lights()
    for k in ("prim", "area", "prob", "alias", "pmf", "primitive_light_index"):
        assert np.array_equal(np.asarray(lt[k]).view(np.uint32), np.asarray(g["light_" + k]).view(np.uint32)), k


def check_film(film, ref, spp, tol):
    a, c = film / spp, ref / spp
    err = rd_rel(a, c)
    off = float((np.abs(a - c).max(axis=2) > 1e-3).mean())
    print(f"relRMSE={err:.3e} max|d|={np.abs(a - c).max():.3e} frac>1e-3={off:.4f}", flush=True)
    assert err < tol
    assert off < FRAC_PIXELS_OFF
    return err


def rd_rel(a, c):
    from lux_amd.render import rel_rmse

    return rel_rmse(a, c)


@pytest.mark.parametrize("case", ["pt_cornell_32", "pt_icosphere", "pt_mixed", "pt_mixed_d0", "pt_mixed_d1"])
def test_render_matches_reference(rd, case):
    """pt_mixed covers thin Lambert, transparent and interface (pass-through,
    also for shadow rays), metallic GGX, glass and a two-sided emitter; the d0 /
    d1 variants the bounce limit with pass-through hits."""
    print("case: render", case, flush=True)
    g = load_golden(case)
    sc = rd.Scene(scene_of(g))
    w, h, spp, depth, seed = int(g["width"]), int(g["height"]), int(g["spp"]), int(g["depth"]), int(g["seed"])
    frame = int(g["frame"]) if "frame" in g else 0
    for mode in (rd.WAVEFRONT, rd.PLAIN):
        film = sc.render(w, h, spp, depth, seed=seed, frame=frame, mode=mode)
        check_film(film, g["film_sum"], spp, RELRMSE_SMALL)


def test_mixed_materials_per_sample(rd):
    """Per (pixel, sample) radiance of the mixed-material scene against
    trace_path's: the large majority agree to 1e-4 relative."""
    print("case: per-sample mixed", flush=True)
    g = load_golden("pt_mixed")
    sc = rd.Scene(scene_of(g))
    px, sm, ref = g["sample_pixels"], g["sample_index"], g["sample_radiance"]
    close = 0
    for k in range(int(g["spp"])):
        film = sc.render(40, 40, 1, int(g["depth"]), seed=int(g["seed"]), frame=int(g["frame"]),
                         sample_offset=k).reshape(-1, 3)
        sel = sm == k
        got, exp = film[px[sel]], ref[sel]
        close += int(np.all(np.abs(got - exp) <= 1e-4 * np.maximum(np.abs(exp), 1e-2), axis=1).sum())
    frac = close / len(px)
    print(f"samples within 1e-4: {frac:.4f}", flush=True)
    assert frac > 0.95


def test_per_sample_radiance(rd):
    """Sample k of every pixel (spp=1, sample_offset=k) against trace_path's value:
    the large majority agree to 1e-4 relative; the rest are branch flips."""
    print("case: per-sample cornell 32", flush=True)
    g = load_golden("pt_cornell_32")
    sc = rd.Scene(scene_of(g))
    px, sm, ref = g["sample_pixels"], g["sample_index"], g["sample_radiance"]
    close = 0
    for k in range(4):
        film = sc.render(32, 32, 1, 4, seed=0, sample_offset=k).reshape(-1, 3)
        sel = sm == k
        got = film[px[sel]]
        exp = ref[sel]
        close += int(np.all(np.abs(got - exp) <= 1e-4 * np.maximum(np.abs(exp), 1e-2), axis=1).sum())
    frac = close / len(px)
    print(f"samples within 1e-4: {frac:.4f}", flush=True)
    assert frac > 0.95


def test_c1_oracle_image(rd):
    """The C1 oracle: Cornell 256x256, 64 spp, depth 4, seed 0."""
    print("case: C1 cornell 256 x 64spp", flush=True)
    g = load_golden("pt_cornell_c1")
    sc = rd.Scene(scene_of(g))
    film = sc.render(256, 256, 64, 4, seed=0)
    check_film(film, g["film_sum"], 64, RELRMSE_C1)


def test_schedules_and_batch_sizes_are_bit_identical(rd):
    print("case: schedule invariance", flush=True)
    g = load_golden("pt_icosphere")
    sc = rd.Scene(scene_of(g))
    a = sc.render(40, 40, 8, 8, seed=7, mode=rd.WAVEFRONT)
    b = sc.render(40, 40, 8, 8, seed=7, mode=rd.PLAIN)
    c = sc.render(40, 40, 8, 8, seed=7, mode=rd.WAVEFRONT, batch_size=777)
    d = sc.render(40, 40, 8, 8, seed=7, mode=rd.PLAIN, batch_size=300)
    for x in (b, c, d):
        assert np.array_equal(bits(a), bits(x))


def test_progressive_samples_add_up(rd):
    """Two renders with sample_offset (progressive spp) sum to one render."""
    print("case: progressive", flush=True)
    g = load_golden("pt_cornell_32")
    sc = rd.Scene(scene_of(g))
    full = sc.render(32, 32, 4, 4, seed=0)
    first = sc.render(32, 32, 2, 4, seed=0)
    second = sc.render(32, 32, 2, 4, seed=0, sample_offset=2)
    np.testing.assert_allclose(first + second, full, rtol=1e-5, atol=1e-6)


def test_invalid_settings_rejected(rd):
    from lux_amd import abi

    g = load_golden("pt_cornell_32")
    sc = rd.Scene(scene_of(g))
    with pytest.raises(abi.LuxError):
        sc.render(0, 32, 4, 4)
    with pytest.raises(abi.LuxError):
        sc.render(32, 32, 4, -1)


# ---- multi-scene batches, graph replay, vertex updates ---------------------------------

def small_scenes(rd):
    from lux_amd import scenes

    specs = [
        (scene_of(load_golden("pt_cornell_32")), dict(width=32, height=32, spp=5, depth=4, seed=3)),
        (scenes.displaced_icosphere(freq=9, seed=0x4C555A, width=48, height=40), dict(width=48, height=40, spp=7,
                                                                                      depth=6, seed=11)),
        (scenes.displaced_icosphere(freq=14, seed=0x4C555B, width=24, height=24), dict(width=24, height=24, spp=3,
                                                                                       depth=8, seed=5, frame=2)),
        (scenes.cornell(20, 30), dict(width=20, height=30, spp=9, depth=3, seed=0, sample_offset=4)),
    ]
    return [rd.Scene(s) for s, _ in specs], [j for _, j in specs]


def standalone(sc, j):
    kw = {k: v for k, v in j.items() if k not in ("width", "height", "spp", "depth")}
    return sc.render(j["width"], j["height"], j["spp"], j["depth"], **kw)


@pytest.mark.parametrize("capacity", [0, 1000, 37])
def test_batch_equals_standalone(rd, capacity):
    """Each film of a multi-scene batch is bit-identical to the same job rendered
    alone, also when batches cut requests mid-sample and hold several segments."""
    print("case: batch vs standalone, capacity", capacity, flush=True)
    saved = rd.get_tuning()
    try:
        scs, jobs = small_scenes(rd)
        alone = [standalone(sc, j) for sc, j in zip(scs, jobs)]
        if capacity:
            rd.set_tuning(capacity=capacity)
        films = rd.batch_render(scs, jobs)
        for a, f in zip(alone, films):
            assert np.array_equal(bits(a), bits(f))
        e, sh = rd.batch_rays()
        assert e >= sum(j["width"] * j["height"] * j["spp"] for j in jobs) and sh > 0
    finally:
        rd.set_tuning(**saved)


def test_batch_plain_and_graphless_are_identical(rd):
    print("case: batch plain / no graph", flush=True)
    saved = rd.get_tuning()
    try:
        scs, jobs = small_scenes(rd)
        ref = rd.batch_render(scs, jobs)
        plain = rd.batch_render(scs, jobs, mode=rd.PLAIN)
        rd.set_tuning(use_graph=0, capacity=4096, grid=64)
        nograph = rd.batch_render(scs, jobs)
        for a, b, c in zip(ref, plain, nograph):
            assert np.array_equal(bits(a), bits(b))
            assert np.array_equal(bits(a), bits(c))
    finally:
        rd.set_tuning(**saved)


def test_graph_replays_are_stable(rd):
    """The captured bounce loop replayed many times (small capacity -> many
    batches, eager work and synchronisations between renders) gives the same
    film every time."""
    print("case: graph replay stability", flush=True)
    saved = rd.get_tuning()
    try:
        g = load_golden("pt_icosphere")
        sc = rd.Scene(scene_of(g))
        rd.set_tuning(capacity=512, grid=8, use_graph=1)
        first = sc.render(40, 40, 4, 8, seed=1)
        for _ in range(3):
            assert np.array_equal(bits(first), bits(sc.render(40, 40, 4, 8, seed=1)))
        rd.set_tuning(use_graph=0)
        assert np.array_equal(bits(first), bits(sc.render(40, 40, 4, 8, seed=1)))
    finally:
        rd.set_tuning(**saved)


def test_vertex_update_refits(rd):
    """Jittered vertices: the refit scene renders like a scene built from the
    jittered mesh (same light table bit for bit; films equal up to hit-order
    ties between the refit and rebuilt trees)."""
    from lux_amd import scenes

    print("case: vertex update + refit", flush=True)
    base = scenes.displaced_icosphere(freq=12, seed=0x4C555C, width=40, height=40)
    rng = np.random.default_rng(99)
    moved = dict(base)
    moved["vertices"] = (base["vertices"] + rng.normal(0, 1e-3, base["vertices"].shape)).astype(np.float32)
    sc = rd.Scene(base)
    before = sc.render(40, 40, 4, 6, seed=2)
    sc.update_vertices(moved["vertices"])
    fresh = rd.Scene(moved)
    la, lb = sc.lights(), fresh.lights()
    for k in la:
        assert np.array_equal(np.asarray(la[k]).view(np.uint32), np.asarray(lb[k]).view(np.uint32)), k
    a = sc.render(40, 40, 4, 6, seed=2)
    b = fresh.render(40, 40, 4, 6, seed=2)
    err = rd.rel_rmse(a / 4, b / 4)
    print(f"refit vs rebuild relRMSE={err:.3e} identical={np.array_equal(bits(a), bits(b))}", flush=True)
    assert err < 1e-3
    assert not np.array_equal(bits(a), bits(before))
    from lux_amd import abi

    short = np.ascontiguousarray(moved["vertices"][:-1])
    with pytest.raises(abi.LuxError):
        abi.call("lux_hip_scene_update_vertices", sc.handle, abi.ptr(short), len(short))


@pytest.mark.parametrize("case", ["pt_icosphere", "pt_mixed", "pt_cornell_32"])
def test_near_first_walk_equals_preorder_walk_in_renders(rd, case):
    """Extension rays walked near child first (with the ambiguity check) and in
    the host LBVH's right-first preorder give bit-identical films and the same
    ray totals."""
    print("case: walk order", case, flush=True)
    g = load_golden(case)
    sc = rd.Scene(scene_of(g))
    saved = rd.get_tuning()
    try:
        a = sc.render(48, 40, 8, 8, seed=3)
        ra = sc.rays()
        rd.set_tuning(preorder_walk=True)
        b = sc.render(48, 40, 8, 8, seed=3)
        rb = sc.rays()
    finally:
        rd.set_tuning(**saved)
    assert np.array_equal(bits(a), bits(b))
    assert ra == rb

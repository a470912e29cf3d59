"""GPU: the render workloads at their full sizes.

C3  the single-scene workload (displaced icosphere, 865 280 triangles,
    1024x1024, 64 spp, depth 8, seed 0): rows 128..191 against the reference
    CPU path tracer's trace_path (golden band written by
    tests/golden/make_golden_pt.py), by relRMSE and per-sample agreement.
C4  the 64-scene batched mix (triangles log-U[1e3, 5e5], sides {128..1024},
    spp log-U[16, 256], depth 6): films of one batched launch equal the same
    scenes rendered alone, and the whole batch with extension rays walked in
    the host LBVH's right-first preorder gives the same 64 films and ray totals.
C5  the served stream: 96 requests with Poisson arrivals drawn from the C4
    mix, 30 % re-rendering a scene after sigma=1e-3 vertex jitter (refit):
    sampled films equal their standalone renders bit for bit, and requests
    shared batches.
"""

import hashlib

import numpy as np
import pytest

from conftest import load_golden

pytestmark = pytest.mark.gpu

# C3 band (spp-averaged RGB, relRMSE = sqrt(mean((g-c)^2/(c^2+1e-2)))): the
# GPU's cosf/sinf/powf round differently from the host's, so a small share of
# paths take another bounce. Measured on MI355X: relRMSE 3.3e-3, no pixel off
# by more than 1e-3, all 4096 sampled radiances within 1e-4.
RELRMSE_C3 = 5e-3
FRAC_PIXELS_OFF_C3 = 0.002  # pixels with |delta| > 1e-3 (spp-averaged)
SAMPLES_CLOSE_C3 = 0.999    # (pixel, sample) radiances within 1e-4 relative


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def scene_digest(scene):
    h = hashlib.sha256()
    for k in ("vertices", "indices", "normals", "uvs", "material_ids", "materials", "camera"):
        h.update(np.ascontiguousarray(scene[k]).tobytes())
    return h.hexdigest()


@pytest.fixture(scope="module")
def rd(gpu):
    from lux_amd import render

    return render


@pytest.fixture(scope="module")
def c4_mix(rd):
    from lux_amd import scenes

    specs = scenes.batched_mix(64)
    dicts = [scenes.displaced_icosphere(freq=s["freq"], seed=s["seed"], width=s["width"], height=s["height"])
             for s in specs]
    jobs = [dict(width=s["width"], height=s["height"], spp=s["spp"], depth=s["depth"], seed=0) for s in specs]
    return dicts, jobs


def test_c3_band_matches_reference(rd):
    from lux_amd import scenes

    print("case: C3 full size, rows 128..191", flush=True)
    g = load_golden("pt_c3_band")
    scene = scenes.displaced_icosphere()
    assert scene_digest(scene) == str(g["scene_sha256"]), "C3 scene arrays differ from the golden's"
    sc = rd.Scene(scene)
    assert np.array_equal(bits(sc.camera()), bits(g["camera_vectors"][:12]))
    w, h, spp, depth = int(g["width"]), int(g["height"]), int(g["spp"]), int(g["depth"])
    film = sc.render(w, h, spp, depth, seed=int(g["seed"]))
    y0, y1 = (int(v) for v in g["rows"])
    band = film.reshape(h, w, 3)[y0:y1]
    a, c = band / spp, g["band_sum"] / spp
    err = rd.rel_rmse(a, c)
    off = float((np.abs(a - c).max(axis=2) > 1e-3).mean())
    print(f"C3 band relRMSE={err:.3e} frac>1e-3={off:.4f} mean={a.mean():.5f}", flush=True)
    assert err < RELRMSE_C3
    assert off < FRAC_PIXELS_OFF_C3
    px, sm, ref = g["sample_pixels"], g["sample_index"], g["sample_radiance"]
    close = 0
    for k in np.unique(sm):
        one = sc.render(w, h, 1, depth, seed=int(g["seed"]), sample_offset=int(k)).reshape(-1, 3)
        sel = sm == k
        got, exp = one[px[sel]], ref[sel]
        close += int(np.all(np.abs(got - exp) <= 1e-4 * np.maximum(np.abs(exp), 1e-2), axis=1).sum())
    frac = close / len(px)
    print(f"C3 samples within 1e-4: {frac:.4f}", flush=True)
    assert frac > SAMPLES_CLOSE_C3


def test_c4_mix_batched_equals_alone_and_preorder(rd, c4_mix):
    from lux_amd import lbvh

    dicts, jobs = c4_mix
    print("case: C4 64-scene mix", flush=True)
    scs = [rd.Scene(d) for d in dicts]
    shapes = [(j["height"], j["width"], 3) for j in jobs]
    d_films = [lbvh.DeviceArray(j["width"] * j["height"] * 12) for j in jobs]
    try:
        rd.batch_render_device(scs, jobs, d_films)
        batched = [f.to_numpy(np.float32, s) for f, s in zip(d_films, shapes)]
        ext, sh = rd.batch_rays()
        for i in (0, 21, 42, 63):
            j = jobs[i]
            alone = scs[i].render(j["width"], j["height"], j["spp"], j["depth"], seed=j["seed"])
            assert np.array_equal(bits(alone), bits(batched[i])), i
        saved = rd.get_tuning()
        try:
            rd.set_tuning(preorder_walk=True)
            rd.batch_render_device(scs, jobs, d_films)
            pre = [f.to_numpy(np.float32, s) for f, s in zip(d_films, shapes)]
            pe, ps = rd.batch_rays()
        finally:
            rd.set_tuning(**saved)
        same = [np.array_equal(bits(a), bits(b)) for a, b in zip(batched, pre)]
        print(f"C4 rays ext={ext} shadow={sh}; preorder-walk films equal: {sum(same)}/64", flush=True)
        assert all(same)
        assert (ext, sh) == (pe, ps)
    finally:
        for f in d_films:
            f.free()


def test_c5_stream_equals_standalone(rd, c4_mix):
    dicts, jobs = c4_mix
    print("case: C5 served stream", flush=True)
    rate = 36.864  # requests/s: 75 % of the measured batched capacity of the mix
    rng = np.random.default_rng(0x4C555B)
    n = 96
    arrivals = np.cumsum(rng.exponential(1.0 / rate, n))
    arrivals -= arrivals[0]
    pick = rng.integers(0, len(dicts), n)
    verts, seen = [], set()
    for i in range(n):
        k = int(pick[i])
        if seen and rng.random() < 0.3:
            k = int(rng.choice(sorted(seen)))
            pick[i] = k
            v = dicts[k]["vertices"]
            verts.append((v + rng.normal(0, 1e-3, v.shape)).astype(np.float32))
        else:
            verts.append(None)
        seen.add(k)
    used = sorted(set(int(k) for k in pick))
    scs = {k: rd.Scene(dicts[k]) for k in used}
    sv = rd.Server(hbm_budget=200 << 30, slice_spp=16)
    films, lat, wall = sv.run([scs[int(k)] for k in pick], [jobs[int(k)] for k in pick], arrivals, vertices=verts)
    st = sv.stats()
    sv.close()
    refits = sum(v is not None for v in verts)
    print(f"C5 wall={wall:.2f} s p50={np.percentile(lat, 50) * 1e3:.1f} ms batches={st['batches']} "
          f"shared={st['shared_batches']} max_per_batch={st['max_batch_requests']} refits={st['refits']}",
          flush=True)
    assert st["completed"] == n and st["refits"] == refits
    assert st["shared_batches"] > 0 and st["max_batch_requests"] >= 2
    # standalone: a fresh scene with the vertices of the latest update at or
    # before the request (the server admits a scene's requests in arrival order)
    jittered = [i for i in range(n) if verts[i] is not None]
    plain = [i for i in range(n) if verts[i] is None]
    check = sorted(set(jittered[:: max(1, len(jittered) // 4)][:4] + plain[:: max(1, len(plain) // 4)][:4]))
    assert len(check) >= 8
    for i in check:
        k = int(pick[i])
        latest = [j for j in range(i + 1) if int(pick[j]) == k and verts[j] is not None]
        alone_sc = rd.Scene(dicts[k])
        if latest:
            alone_sc.update_vertices(verts[latest[-1]])
        j = jobs[k]
        alone = alone_sc.render(j["width"], j["height"], j["spp"], j["depth"], seed=j["seed"])
        alone_sc.close()
        assert np.array_equal(bits(alone), bits(films[i])), i

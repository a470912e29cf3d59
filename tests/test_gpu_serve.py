"""GPU: the continuous-batching request server. Each request's film must equal
the film of the same render run alone, bit for bit (per-path arithmetic does
not depend on the batch; every pixel's samples are summed in sample order),
including requests that re-render a scene after a vertex update (refit)."""

import numpy as np
import pytest

from conftest import load_golden

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


@pytest.fixture(scope="module")
def rd(gpu):
    from lux_amd import render

    return render


def scene_dicts():
    from lux_amd import scenes

    g = load_golden("pt_cornell_32")
    cornell = {k: g[k] for k in ("vertices", "indices", "normals", "uvs", "material_ids", "materials", "camera")}
    return [cornell,
            scenes.displaced_icosphere(freq=10, seed=0x4C555A, width=32, height=32),
            scenes.displaced_icosphere(freq=16, seed=0x4C555B, width=48, height=32)]


def jitter(v, seed):
    return (v + np.random.default_rng(seed).normal(0, 1e-3, v.shape)).astype(np.float32)


def make_trace():
    """Nine requests over three scenes; two move a scene's vertices first."""
    dicts = scene_dicts()
    reqs = [  # (scene index, job, jitter seed or None, arrival s)
        (1, dict(width=32, height=32, spp=24, depth=6, seed=1), None, 0.000),
        (0, dict(width=32, height=32, spp=40, depth=4, seed=2), None, 0.001),
        (2, dict(width=48, height=32, spp=17, depth=6, seed=3), None, 0.002),
        (1, dict(width=32, height=32, spp=8, depth=6, seed=4), 101, 0.003),
        (0, dict(width=16, height=16, spp=5, depth=3, seed=5), None, 0.003),
        (2, dict(width=48, height=32, spp=33, depth=5, seed=6, sample_offset=7), None, 0.010),
        (1, dict(width=32, height=32, spp=16, depth=6, seed=7), None, 0.011),
        (2, dict(width=24, height=24, spp=12, depth=6, seed=8), 202, 0.012),
        (0, dict(width=32, height=32, spp=3, depth=2, seed=9, frame=3), None, 0.030),
    ]
    return dicts, reqs


def standalone_films(rd, dicts, reqs):
    """Replays the trace one request at a time on fresh scenes."""
    scs = [rd.Scene(d) for d in dicts]
    out = []
    for si, job, jit, _ in reqs:
        if jit is not None:
            scs[si].update_vertices(jitter(dicts[si]["vertices"], jit))
        kw = {k: v for k, v in job.items() if k not in ("width", "height", "spp", "depth")}
        out.append(scs[si].render(job["width"], job["height"], job["spp"], job["depth"], **kw))
    return out


@pytest.mark.parametrize("slice_spp,capacity", [(16, 0), (4, 3000), (0, 777)])
def test_trace_results_equal_standalone(rd, slice_spp, capacity):
    print("case: server trace, slice", slice_spp, "capacity", capacity, flush=True)
    dicts, reqs = make_trace()
    alone = standalone_films(rd, dicts, reqs)
    saved = rd.get_tuning()
    try:
        if capacity:
            rd.set_tuning(capacity=capacity)
        scs = [rd.Scene(d) for d in dicts]
        sv = rd.Server(slice_spp=slice_spp)
        verts = [None if jit is None else jitter(dicts[si]["vertices"], jit) for si, _, jit, _ in reqs]
        films, lat, wall = sv.run([scs[si] for si, *_ in reqs], [j for _, j, _, _ in reqs],
                                  [a for *_, a in reqs], vertices=verts)
        st = sv.stats()
        print(f"wall={wall * 1e3:.1f} ms p50={np.percentile(lat, 50) * 1e3:.2f} ms batches={st['batches']} "
              f"refits={st['refits']}", flush=True)
        assert st["completed"] == len(reqs) and st["refits"] == 2
        assert np.all(lat > 0)
        for k, (a, f) in enumerate(zip(alone, films)):
            assert np.array_equal(bits(a), bits(f)), k
        sv.close()
    finally:
        rd.set_tuning(**saved)


def test_submit_poll_api(rd):
    print("case: server submit/poll", flush=True)
    dicts, reqs = make_trace()
    alone = standalone_films(rd, dicts, reqs)
    scs = [rd.Scene(d) for d in dicts]
    sv = rd.Server()
    ids = []
    for si, job, jit, _ in reqs:
        v = None if jit is None else jitter(dicts[si]["vertices"], jit)
        ids.append(sv.submit(scs[si], job, v))
        sv.poll()
    got = {}
    for _ in range(100000):
        for rid in sv.poll(block=True):
            got[rid] = sv.result(rid)
        if len(got) == len(ids):
            break
    assert len(got) == len(ids)
    for rid, a in zip(ids, alone):
        film, lat = got[rid]
        assert lat > 0
        assert np.array_equal(bits(a), bits(film))


def test_budget_admission(rd):
    """A budget that holds the scenes, the path buffer and about one film at a
    time: every request still completes, correctly, within the budget."""
    print("case: server budget", flush=True)
    dicts, reqs = make_trace()
    alone = standalone_films(rd, dicts, reqs)
    saved = rd.get_tuning()
    try:
        rd.set_tuning(capacity=4096)
        scs = [rd.Scene(d) for d in dicts]
        scene_bytes = sum(sc.info()[2] for sc in scs)
        film_max = max(j["width"] * j["height"] * 12 for _, j, _, _ in reqs)
        budget = scene_bytes + (4 << 20) + film_max + 1024
        sv = rd.Server(hbm_budget=budget)
        verts = [None if jit is None else jitter(dicts[si]["vertices"], jit) for si, _, jit, _ in reqs]
        films, lat, wall = sv.run([scs[si] for si, *_ in reqs], [j for _, j, _, _ in reqs],
                                  [a for *_, a in reqs], vertices=verts)
        st = sv.stats()
        print(f"budget={budget} peak={st['peak_bytes']}", flush=True)
        assert st["peak_bytes"] <= budget
        for a, f in zip(alone, films):
            assert np.array_equal(bits(a), bits(f))
    finally:
        rd.set_tuning(**saved)


def test_scene_guarded_while_requests_are_unfinished(rd):
    """A scene with a submitted, unfinished request refuses a direct vertex
    update and destruction; both work once the request completed."""
    print("case: server scene guards", flush=True)
    from lux_amd.abi import LuxError

    dicts, _ = make_trace()
    sc = rd.Scene(dicts[1])
    sv = rd.Server()
    rid = sv.submit(sc, dict(width=32, height=32, spp=64, depth=6, seed=1))
    with pytest.raises(LuxError):
        sc.update_vertices(jitter(dicts[1]["vertices"], 5))
    with pytest.raises(LuxError):
        rd.call("lux_hip_scene_destroy", sc.handle)
    done = []
    for _ in range(100000):
        done += sv.poll(block=True)
        if rid in done:
            break
    assert rid in done
    sv.result(rid)
    sc.update_vertices(jitter(dicts[1]["vertices"], 5))
    sv.close()
    sc.close()


def test_deferred_scene_builds_on_first_use(rd):
    """A deferred scene has no device state until its first render, then renders
    what the same scene created resident renders; a vertex update made while it
    is not resident is applied by the rebuild (build from the first vertices,
    refit to the update), as on a resident scene."""
    print("case: deferred scene", flush=True)
    d = scene_dicts()[1]
    a, b = rd.Scene(d), rd.Scene(d, deferred=True)
    assert a.resident() and not b.resident() and b.info()[2] > 0
    fa = a.render(32, 32, 8, 5, seed=3)
    fb = b.render(32, 32, 8, 5, seed=3)
    assert b.resident() and np.array_equal(bits(fa), bits(fb))
    c = rd.Scene(d, deferred=True)
    v = jitter(d["vertices"], 55)
    c.update_vertices(v)
    a.update_vertices(v)
    assert not c.resident()
    assert np.array_equal(bits(a.render(32, 32, 8, 5, seed=4)), bits(c.render(32, 32, 8, 5, seed=4)))
    for s in (a, b, c):
        s.close()


def test_server_owns_residency_under_tight_budget(rd):
    """Deferred scenes: the server uploads a scene when it admits a request for
    it and, under a budget that holds about two and a half scenes beside its
    fixed footprint, drops idle scenes least recently used first and rebuilds
    them when they are asked for again; films equal the standalone renders."""
    print("case: server residency, tight budget", flush=True)
    from lux_amd import scenes

    dicts = [scenes.displaced_icosphere(freq=12 + 2 * k, seed=0x4C5600 + k, width=24, height=24) for k in range(5)]
    scs = [rd.Scene(d, deferred=True) for d in dicts]
    assert not any(sc.resident() for sc in scs)
    sizes = [sc.info()[2] for sc in scs]

    def job(i):
        return dict(width=24, height=24, spp=6, depth=4, seed=i)

    # the server's fixed footprint (engine buffers, one film): one request alone
    probe = rd.Server(hbm_budget=1 << 40, slice_spp=8)
    probe.run([scs[0]], [job(0)], [0.0])
    fixed = probe.stats()["peak_bytes"] - sizes[0]
    probe.close()
    assert scs[0].resident()
    budget = fixed + int(2.5 * max(sizes))
    order = [0, 1, 2, 3, 4, 0, 2, 4, 1, 3, 0]
    moves = {5: 301, 8: 302}  # request index -> jitter seed of a vertex update before it
    verts = [jitter(dicts[order[i]]["vertices"], moves[i]) if i in moves else None for i in range(len(order))]
    sv = rd.Server(hbm_budget=budget, slice_spp=8)
    films, lat, wall = sv.run([scs[k] for k in order], [job(i) for i in range(len(order))],
                              [0.002 * i for i in range(len(order))], vertices=verts)
    st = sv.stats()
    sv.close()
    print(f"residency: uploads={st['uploads']} evictions={st['evictions']} refits={st['refits']} "
          f"peak={st['peak_bytes']} budget={budget} prepare_s={st['prepare_seconds']:.4f}", flush=True)
    assert st["completed"] == len(order) and st["refits"] == len(moves)
    assert st["evictions"] >= 2 and st["uploads"] >= 6
    assert st["peak_bytes"] <= budget
    alone = [rd.Scene(d) for d in dicts]
    for i, k in enumerate(order):
        if verts[i] is not None:
            alone[k].update_vertices(verts[i])
        j = job(i)
        a = alone[k].render(j["width"], j["height"], j["spp"], j["depth"], seed=j["seed"])
        assert np.array_equal(bits(a), bits(films[i])), i
    for s in alone + scs:
        s.close()

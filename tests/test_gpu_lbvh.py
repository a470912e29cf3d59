"""GPU parity: the HIP LBVH build, refit, closest/any-hit traversal and overlap
queries against the golden vectors (reference host code) and the CPU oracle.
Integer/index results must match bit-exactly; FP32 hit distances and boxes too
(same operation order, no contraction)."""

import numpy as np
import pytest

import oracle
from conftest import LBVH_CASES, load_golden
from lux_amd import abi, workloads as W

pytestmark = pytest.mark.gpu
LEAF = 0x80000000


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def print_case(*what):
    print("case:", *what, flush=True)


@pytest.mark.parametrize("case", LBVH_CASES)
def test_build_topology_bit_exact(gpu, case):
    print_case("build", case)
    g = load_golden(f"lbvh_{case}")
    n = len(g["boxes"])
    bvh = gpu.build_gpu_lbvh(g["boxes"], g["ids"])
    valid, count, root = bvh.info()
    assert valid and count == n
    assert np.array_equal(bits(root), bits(g["root"]))
    topo = bvh.topology()
    o = oracle.OracleBvh(g["boxes"], g["ids"])
    if n > 1:
        assert np.array_equal(topo["morton_sorted"], g["morton_input"][topo["sorted_indices"]])
        assert np.array_equal(topo["sorted_indices"], o.perm)
        assert np.array_equal(topo["left"], o.left)
        assert np.array_equal(topo["right"], o.right)
        assert np.array_equal(topo["parent"][1:], o.parent[1:])
        assert np.array_equal(topo["leaf_parent"], o.leaf_parent)
    assert np.array_equal(g["ids"][topo["sorted_indices"]], g["leaf_ids"])
    assert np.array_equal(bits(topo["node_bounds"]), bits(g["node_bounds"]))
    nodes = bvh.nodes()
    assert np.array_equal(nodes.view(np.uint32), o.nodes.view(np.uint32))


@pytest.mark.parametrize("case", LBVH_CASES)
def test_closest_hit_bit_exact(gpu, case):
    print_case("closest", case)
    g = load_golden(f"lbvh_{case}")
    bvh = gpu.build_gpu_lbvh(g["boxes"], g["ids"])
    hits = gpu.intersect_gpu_bvh_stackless(bvh, g["tri_by_id"], g["rays"])
    assert np.array_equal(hits["primitive_id"], g["hit_ids"])
    assert np.array_equal(bits(hits["t"]), bits(g["hit_t"]))
    occ = gpu.trace_any(bvh, g["tri_by_id"], g["rays"])
    assert np.array_equal(occ.astype(bool), g["hit_ids"] >= 0)


@pytest.mark.parametrize("case", LBVH_CASES)
def test_closest_hit_after_refit_matches_oracle(gpu, case):
    """Refit with jittered triangles, then closest hits equal the oracle's walk
    of the oracle-refitted tree (same topology, new boxes), and a rebuild from
    the new boxes gives the hits of a fresh build."""
    print_case("closest after refit", case)
    g = load_golden(f"lbvh_{case}")
    bvh = gpu.build_gpu_lbvh(g["boxes"], g["ids"])
    moved = W.jitter(g["tris"], sigma=2e-3, seed=7)
    new_boxes = W.triangle_boxes(moved)
    assert gpu.refit_gpu_lbvh(bvh, new_boxes, g["ids"])
    tri_by_id = np.zeros_like(g["tri_by_id"])
    tri_by_id[g["ids"]] = moved
    hits = gpu.intersect_gpu_bvh_stackless(bvh, tri_by_id, g["rays"])
    o = oracle.OracleBvh(g["boxes"], g["ids"])
    o.refit(new_boxes)
    exp = o.trace_closest(tri_by_id, g["rays"])
    assert np.array_equal(hits["primitive_id"], exp["ids"])
    assert np.array_equal(bits(hits["t"]), bits(exp["t"]))
    bvh.build(new_boxes, g["ids"])
    rebuilt = gpu.intersect_gpu_bvh_stackless(bvh, tri_by_id, g["rays"])
    fresh = oracle.OracleBvh(new_boxes, g["ids"]).trace_closest(tri_by_id, g["rays"])
    assert np.array_equal(rebuilt["primitive_id"], fresh["ids"])
    assert np.array_equal(bits(rebuilt["t"]), bits(fresh["t"]))


@pytest.mark.parametrize("case", ["soup2000", "clustered1500", "sparse_ids777", "tiny3", "tiny1"])
def test_refit_matches_oracle_refit(gpu, case):
    print_case("refit", case)
    g = load_golden(f"lbvh_{case}")
    bvh = gpu.build_gpu_lbvh(g["boxes"], g["ids"])
    moved = W.jitter(g["tris"], sigma=2e-3, seed=99)
    new_boxes = W.triangle_boxes(moved)
    assert gpu.refit_gpu_lbvh(bvh, new_boxes, g["ids"])
    o = oracle.OracleBvh(g["boxes"], g["ids"]).refit(new_boxes)
    topo = bvh.topology()
    assert np.array_equal(topo["sorted_indices"], o.perm)
    assert np.array_equal(bits(topo["node_bounds"]), bits(o.node_bounds))
    assert np.array_equal(bvh.nodes().view(np.uint32), o.nodes.view(np.uint32))
    assert np.array_equal(bits(bvh.root_bounds), bits(o.root))
    tri_by_id = np.zeros_like(g["tri_by_id"])
    tri_by_id[g["ids"]] = moved
    hits = gpu.intersect_gpu_bvh_stackless(bvh, tri_by_id, g["rays"])
    exp = o.trace_closest(tri_by_id, g["rays"])
    assert np.array_equal(hits["primitive_id"], exp["ids"])
    assert np.array_equal(bits(hits["t"]), bits(exp["t"]))


def test_refit_mismatch_returns_false(gpu):
    print_case("refit-mismatch")
    g = load_golden("lbvh_soup2000")
    bvh = gpu.build_gpu_lbvh(g["boxes"], g["ids"])
    assert not gpu.refit_gpu_lbvh(bvh, g["boxes"][:-1], g["ids"][:-1])
    bad_ids = g["ids"].copy()
    bad_ids[[3, 4]] = bad_ids[[4, 3]]
    assert not gpu.refit_gpu_lbvh(bvh, g["boxes"], bad_ids)
    assert gpu.refit_gpu_lbvh(bvh, g["boxes"], g["ids"])


def test_empty_input_gives_invalid_bvh(gpu):
    print_case("empty")
    bvh = gpu.build_gpu_lbvh(np.zeros((0, 6), np.float32))
    assert bvh.info()[:2] == (False, 0)
    hits = gpu.intersect_gpu_bvh_stackless(bvh, np.zeros((0, 9), np.float32), W.rays(5))
    assert np.all(hits["primitive_id"] == -1) and np.all(np.isinf(hits["t"]))
    assert gpu.refit_gpu_lbvh(bvh, np.zeros((0, 6), np.float32))
    res = gpu.query_aabb_overlaps_self(bvh)
    assert res.found_count == 0 and not res.truncated and len(res.pairs) == 0


def test_negative_ids_rejected(gpu):
    print_case("negative-ids")
    g = load_golden("lbvh_tiny7")
    ids = g["ids"].copy()
    ids[2] = -4
    with pytest.raises(abi.LuxError) as e:
        gpu.build_gpu_lbvh(g["boxes"], ids)
    assert e.value.status == abi.LUX_ERR_INVALID_ARGUMENT


def test_leaf_id_outside_triangle_array_is_reported(gpu):
    print_case("prim-range")
    g = load_golden("lbvh_soup2000")
    bvh = gpu.build_gpu_lbvh(g["boxes"], g["ids"])
    with pytest.raises(abi.LuxError) as e:
        gpu.intersect_gpu_bvh_stackless(bvh, g["tri_by_id"][:100], g["rays"])
    assert e.value.status == abi.LUX_ERR_DEVICE_CHECK
    # the handle stays usable
    hits = gpu.intersect_gpu_bvh_stackless(bvh, g["tri_by_id"], g["rays"])
    assert np.array_equal(hits["primitive_id"], g["hit_ids"])


def test_overlap_self_and_cross_sets(gpu):
    print_case("overlap")
    g = load_golden("overlap_dense2000")
    left = gpu.build_gpu_lbvh(g["boxes"])
    right = gpu.build_gpu_lbvh(g["qboxes"])
    s = gpu.query_aabb_overlaps_self(left)
    assert not s.truncated and s.found_count == len(g["self_pairs"])
    assert set(map(tuple, s.pairs)) == set(map(tuple, g["self_pairs"]))
    assert np.all(s.pairs[:, 0] < s.pairs[:, 1])
    c = gpu.query_aabb_overlaps_cross(left, right)
    assert c.found_count == len(g["cross_pairs"])
    assert set(map(tuple, c.pairs)) == set(map(tuple, g["cross_pairs"]))
    # truncation: exact total, at most max_pairs written, all of them genuine
    t = gpu.query_aabb_overlaps_self(left, max_pairs=100)
    assert t.truncated and t.found_count == len(g["self_pairs"]) and len(t.pairs) == 100
    assert set(map(tuple, t.pairs)) <= set(map(tuple, g["self_pairs"]))
    z = gpu.query_aabb_overlaps_self(left, max_pairs=0)
    assert z.found_count == 0 and len(z.pairs) == 0


def test_repeated_builds_reuse_handle(gpu):
    """Builds of shrinking and growing sizes on one handle (buffer reuse, epoch-
    tagged look-back words, ticket reset) stay bit-exact."""
    print_case("rebuilds")
    bvh = gpu.GpuBvh()
    for n, seed in [(5000, 1), (1000, 2), (70000, 3), (3, 4), (70000, 5)]:
        boxes = W.triangle_boxes(W.soup(n, seed=seed))
        bvh.build(boxes, np.arange(n, dtype=np.int32))
        o = oracle.OracleBvh(boxes)
        assert np.array_equal(bvh.nodes().view(np.uint32), o.nodes.view(np.uint32)), (n, seed)


def check_closest_against_oracle(gpu, tris, ids, rays):
    """GPU closest hit == the oracle's stackless preorder walk == the oracle's
    host stack traversal, ids and t bit for bit."""
    boxes = W.triangle_boxes(tris)
    bvh = gpu.build_gpu_lbvh(boxes, ids)
    tri_by_id = np.zeros((int(ids.max()) + 1, 9), np.float32)
    tri_by_id[ids] = tris.reshape(-1, 9)
    hits = gpu.intersect_gpu_bvh_stackless(bvh, tri_by_id, rays)
    o = oracle.OracleBvh(boxes, ids)
    pre_ids, pre_t = o.trace_stackless(tri_by_id, rays)
    host = o.trace_closest(tri_by_id, rays)
    assert host["overflows"] == 0
    assert np.array_equal(pre_ids, host["ids"]) and np.array_equal(bits(pre_t), bits(host["t"]))
    assert np.array_equal(hits["primitive_id"], pre_ids)
    assert np.array_equal(bits(hits["t"]), bits(pre_t))
    return hits, o


def test_closest_hit_exact_ties(gpu):
    """Rays that hit shared edges and vertices of a triangulated grid, every
    triangle present twice under two ids: several triangles at exactly the same
    t. The closest-hit walk must keep the one the right-first preorder walk
    reaches first."""
    print_case("closest ties")
    k = 48
    g = np.arange(k + 1, dtype=np.float32) / k
    tris = []
    for i in range(k):
        for j in range(k):
            a = (g[i], g[j], 0.5)
            b = (g[i + 1], g[j], 0.5)
            c = (g[i + 1], g[j + 1], 0.5)
            d = (g[i], g[j + 1], 0.5)
            tris += [(a, b, c), (a, c, d)]
    tris = np.asarray(tris, np.float32).reshape(-1, 9)
    tris = np.concatenate([tris, tris])  # exact duplicates under other ids
    ids = np.random.default_rng(5).permutation(len(tris)).astype(np.int32)
    # origins on grid vertices, edge midpoints and cell diagonals, straight up
    pts = np.arange(2 * k + 1, dtype=np.float32) / (2 * k)
    px, py = np.meshgrid(pts, pts)
    n = px.size
    rays = np.zeros((n, 9), np.float32)
    rays[:, 0], rays[:, 1] = px.ravel(), py.ravel()
    rays[:, 5] = 1.0
    rays[:, 7] = np.inf
    hits, _ = check_closest_against_oracle(gpu, tris, ids, rays)
    assert (hits["primitive_id"] >= 0).mean() > 0.99


def test_closest_hit_deep_stack_falls_back_to_preorder(gpu):
    """2^18 parallel triangles stacked along x, every one crossing the rays'
    line: the near-child-first walk keeps one pending sibling per level with two
    internal children, more than its 16-entry stack holds on the way to the
    nearest leaf, so those rays are walked again in preorder. Results equal the
    oracle's either way."""
    print_case("closest deep stack")
    m = 1 << 18
    x = ((np.arange(m) + 0.5) / m).astype(np.float32)
    tris = np.zeros((m, 9), np.float32)
    tris[:, 0] = tris[:, 3] = tris[:, 6] = x
    tris[:, 4] = 2.0
    tris[:, 8] = 2.0
    ids = np.arange(m, dtype=np.int32)
    rng = np.random.default_rng(11)
    k = 64
    rays = np.zeros((2 * k, 9), np.float32)
    rays[:, 1:3] = rng.uniform(0.3, 0.7, (2 * k, 2))
    rays[:k, 0], rays[:k, 3] = -0.5, 1.0  # +x: the nearest leaf is the leftmost
    rays[k:, 0], rays[k:, 3] = 1.5, -1.0  # -x: the nearest leaf is the rightmost
    rays[:, 7] = np.inf
    hits, o = check_closest_against_oracle(gpu, tris, ids, rays)
    assert np.array_equal(hits["primitive_id"][:k], np.zeros(k, np.int32))
    assert np.array_equal(hits["primitive_id"][k:], np.full(k, m - 1, np.int32))
    # levels with two internal children on the path to the leftmost leaf: the
    # stack depth a +x ray reaches there
    left, right = o.left.astype(np.int64), o.right.astype(np.int64)
    node, both = 0, 0
    while True:
        li, ri = (left[node] & 0x80000000) == 0, (right[node] & 0x80000000) == 0
        both += int(li and ri)
        if not li:
            break
        node = int(left[node])
    assert both > 16


def test_full_size_microbench_inputs(gpu):
    """1M-triangle soup, 4M rays: the tree equals the oracle's bit for bit; a
    65536-ray sample matches the oracle traversal; all rays satisfy any-hit ==
    (closest found); refit from jittered vertices equals a refit from boxes."""
    print_case("full-size 1M")
    inp = W.c2_inputs(n_tris=1 << 20, n_rays=1 << 22, n_queries=100_000)
    tris, boxes, rays = inp["tris"], inp["boxes"], inp["rays"]
    bvh = gpu.build_gpu_lbvh(boxes)
    o = oracle.OracleBvh(boxes)
    assert np.array_equal(bvh.nodes().view(np.uint32), o.nodes.view(np.uint32))
    hits = gpu.intersect_gpu_bvh_stackless(bvh, tris, rays)
    sample = np.random.default_rng(0).choice(len(rays), 1 << 16, replace=False)
    exp = o.trace_closest(tris, rays[sample])
    assert exp["overflows"] == 0
    assert np.array_equal(hits["primitive_id"][sample], exp["ids"])
    assert np.array_equal(bits(hits["t"][sample]), bits(exp["t"]))
    occ = gpu.trace_any(bvh, tris, rays)
    assert np.array_equal(occ.astype(bool), hits["primitive_id"] >= 0)
    # FP32 hit distance of every hit against FP64 Moller-Trumbore on the same
    # triangle. On this input (3.31M hits) the oracle's FP32 t, which the GPU
    # equals bit for bit, has median relative error 6.0e-8, 99.9th percentile
    # 1.4e-6 and maximum 2.3e-3 (a grazing hit, where the determinant is tiny).
    found = hits["primitive_id"] >= 0
    t64 = oracle.tri_t_fp64(tris, hits["primitive_id"], rays)[found]
    rel = np.abs(hits["t"][found].astype(np.float64) - t64) / np.maximum(np.abs(t64), 1e-30)
    print(f"fp64 t error over {found.sum()} hits: max {rel.max():.3e}, median {np.median(rel):.3e}, "
          f"p99.9 {np.quantile(rel, 0.999):.3e}", flush=True)
    assert rel.max() < 5e-3 and np.quantile(rel, 0.999) < 1e-5 and np.median(rel) < 1e-6
    # refit from vertices (device entry point) == refit from host boxes
    d_tris = gpu.DeviceArray.from_numpy(inp["tris_jittered"])
    bvh.refit_triangles_device(d_tris, len(tris))
    a = bvh.nodes()
    bvh2 = gpu.build_gpu_lbvh(boxes)
    assert gpu.refit_gpu_lbvh(bvh2, W.triangle_boxes(inp["tris_jittered"]))
    assert np.array_equal(a.view(np.uint32), bvh2.nodes().view(np.uint32))
    # the refitted tree equals the oracle refit of the same topology
    o.refit(W.triangle_boxes(inp["tris_jittered"]))
    assert np.array_equal(a.view(np.uint32), o.nodes.view(np.uint32))
    # cross overlap of the (refitted) 1M tree with the 100k query boxes vs the
    # oracle tree walk, as pair sets
    q = gpu.build_gpu_lbvh(inp["query_boxes"])
    c = gpu.query_aabb_overlaps_cross(bvh, q, max_pairs=1 << 22)
    qids = np.arange(100_000, dtype=np.int32)
    cp, cc = o.overlap_tree(inp["query_boxes"], qids, False)
    assert c.found_count == cc and not c.truncated
    key = lambda p: np.sort(p[:, 0].astype(np.int64) * (1 << 20) + p[:, 1])  # noqa: E731
    assert np.array_equal(key(c.pairs), key(cp))
    s = gpu.query_aabb_overlaps_self(bvh, max_pairs=1 << 22)
    sp, sc = o.overlap_tree(o.boxes, np.arange(len(boxes), dtype=np.int32), True)
    assert s.found_count == sc and np.array_equal(key(s.pairs), key(sp))


@pytest.mark.parametrize("case", ["soup2000", "sparse_ids777", "tiny1"])
def test_input_and_device_arrays(gpu, case):
    """The build input in input order (host copy and the device arrays behind
    GpuBvh::overlap_*_device_ptr), before and after a refit."""
    print_case("input arrays", case)
    g = load_golden(f"lbvh_{case}")
    n = len(g["boxes"])
    bvh = gpu.build_gpu_lbvh(g["boxes"], g["ids"])
    boxes, ids = bvh.input()
    assert np.array_equal(bits(boxes), bits(g["boxes"])) and np.array_equal(ids, g["ids"])
    db, di, ds = bvh.device_arrays()
    assert db and di and ds
    L = abi.load()
    got_b = np.zeros((n, 6), np.float32)
    got_i = np.zeros(n, np.int32)
    got_s = np.zeros(n, np.uint32)
    for dst, src in ((got_b, db), (got_i, di), (got_s, ds)):
        assert L.lux_hip_memcpy_dtoh(dst.ctypes.data, src, dst.nbytes) == 0
    assert np.array_equal(bits(got_b), bits(g["boxes"])) and np.array_equal(got_i, g["ids"])
    assert np.array_equal(got_s, bvh.topology()["sorted_indices"])
    moved = W.triangle_boxes(W.jitter(g["tris"], sigma=2e-3, seed=3))
    assert gpu.refit_gpu_lbvh(bvh, moved, g["ids"])
    assert np.array_equal(bits(bvh.input()[0]), bits(moved))
    empty = gpu.GpuBvh()
    assert empty.device_arrays() == (None, None, None)


@pytest.mark.parametrize("case", ["soup2000", "clustered1500", "flat300", "tiny1", "rounding4", "signed_zero4000"])
def test_build_from_device_triangles(gpu, case):
    """lux_hip_bvh_build_triangles_device (box i = triangle_bounds of triangle i,
    id i, computed on the device) builds the tree of lux_hip_bvh_build with the
    host-computed boxes, node for node."""
    print_case("build from device triangles", case)
    g = load_golden(f"lbvh_{case}")
    tris = np.ascontiguousarray(g["tris"], np.float32)
    ref = gpu.build_gpu_lbvh(g["boxes"])  # the reference's triangle_bounds, ids 0..n-1
    d_tris = gpu.DeviceArray.from_numpy(tris)
    bvh = gpu.GpuBvh().build_triangles_device(d_tris, len(tris))
    assert np.array_equal(bvh.nodes().view(np.uint32), ref.nodes().view(np.uint32))
    assert np.array_equal(bits(bvh.input()[0]), bits(g["boxes"]))
    d_tris.free()

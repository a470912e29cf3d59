"""Writes the LBVH golden vectors under tests/golden/ from the REFERENCE's own host
code (oracle/_ref/libluxref.so, built by oracle/ref.mk from the LUX sources):
build_cpu_lbvh_tree, lbvh_detail::centroid_to_morton, intersect_cpu_bvh_tree +
intersect_scene_triangle_primitive, intersect_triangle, AABB::overlaps.

Run in a container where /root/reference is mounted:
    python tests/golden/make_golden.py
Each case stores its inputs and the reference outputs, so tests never need the
reference or the generator's RNG.
"""

import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

import oracle  # noqa: E402
from lux_amd import workloads as W  # noqa: E402

OUT = Path(__file__).resolve().parent


def targeted_rays(tris, n, seed):
    """Rays shot at random points of random triangles from a short distance."""
    rng = np.random.default_rng(seed)
    k = rng.integers(0, len(tris), n)
    t = tris[k].reshape(n, 3, 3).astype(np.float64)
    a, b = rng.random(n), rng.random(n)
    flip = a + b > 1
    a[flip], b[flip] = 1 - a[flip], 1 - b[flip]
    p = t[:, 0] + a[:, None] * (t[:, 1] - t[:, 0]) + b[:, None] * (t[:, 2] - t[:, 0])
    d = rng.standard_normal((n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    s = rng.uniform(0.05, 0.5, n)[:, None]
    r = np.zeros((n, 9), np.float32)
    r[:, 0:3] = p - d * s
    r[:, 3:6] = d
    r[:, 7] = np.inf
    return r


def mixed_rays(tris, n, seed):
    r = np.concatenate([
        targeted_rays(tris, n // 2, seed),
        W.rays(n // 4, seed=seed + 1),
        W.axis_rays(n - n // 2 - n // 4, seed=seed + 2),
    ])
    rng = np.random.default_rng(seed + 3)
    # a quarter of the rays get a positive t_min and a finite t_max
    sel = rng.random(len(r)) < 0.25
    r[sel, 6] = rng.uniform(0.0, 0.1, sel.sum())
    r[sel, 7] = rng.uniform(0.1, 0.6, sel.sum())
    return r


def reference_case(name, tris, rays, ids=None):
    boxes = np.zeros((len(tris), 6), np.float32)
    oracle.ref().ref_triangle_bounds(oracle._p(np.ascontiguousarray(tris)), len(tris), oracle._p(boxes))
    n = len(boxes)
    if ids is None:
        ids = np.arange(n, dtype=np.int32)
    root = np.concatenate([boxes[:, :3].min(0), boxes[:, 3:].max(0)]).astype(np.float32)
    morton = np.zeros(n, np.uint32)
    oracle.ref().ref_morton_codes(oracle._p(boxes), n, oracle._p(root), oracle._p(morton))
    # the reference tree indexes triangles by primitive id
    tri_by_id = np.zeros((int(ids.max()) + 1, 9), np.float32)
    tri_by_id[ids] = tris
    tree = oracle.RefLbvh(boxes, ids)
    hit = tree.trace_closest(tri_by_id, rays)
    bi = np.zeros(len(rays), np.int32)
    bt = np.zeros(len(rays), np.float32)
    oracle.ref().ref_brute_closest(oracle._p(tri_by_id), len(tri_by_id), oracle._p(rays), len(rays),
                                   oracle._p(bi), oracle._p(bt))
    data = dict(
        tris=tris, tri_by_id=tri_by_id, boxes=boxes, ids=ids, rays=rays, root=root,
        morton_input=morton, leaf_ids=tree.leaf_ids, node_left=tree.left, node_right=tree.right,
        node_bounds=tree.bounds, hit_ids=hit["ids"], hit_t=hit["t"], hit_u=hit["u"], hit_v=hit["v"],
        hit_overflows=np.int32(hit["overflows"]), brute_ids=bi, brute_t=bt,
    )
    np.savez_compressed(OUT / f"lbvh_{name}.npz", **data)
    print(f"{name}: n={n} rays={len(rays)} hits={(hit['ids'] >= 0).sum()} "
          f"brute-id-mismatch(ties)={(bi != hit['ids']).sum()}")


def overlap_case(name, boxes, qboxes):
    n, q = len(boxes), len(qboxes)
    cap = 1 << 20
    sp = np.zeros((cap, 2), np.int32)
    cs = oracle.ref().ref_overlap_self_brute(oracle._p(boxes), n, oracle._p(sp), cap)
    cp = np.zeros((cap, 2), np.int32)
    cc = oracle.ref().ref_overlap_cross_brute(oracle._p(boxes), n, oracle._p(qboxes), q, oracle._p(cp), cap)
    np.savez_compressed(OUT / f"overlap_{name}.npz", boxes=boxes, qboxes=qboxes, self_pairs=sp[:cs],
                        cross_pairs=cp[:cc])
    print(f"overlap {name}: n={n} q={q} self={cs} cross={cc}")


def main():
    if not oracle.ref_available():
        raise SystemExit("oracle/_ref/libluxref.so missing: run `make -f oracle/ref.mk` where the reference exists")
    soup = W.soup(2000, seed=1)
    reference_case("soup2000", soup, mixed_rays(soup, 4096, 100))

    clus = W.clustered_soup(1500, seed=7)
    reference_case("clustered1500", clus, targeted_rays(clus, 2048, 200))

    one = W.soup(1, seed=5)
    ident = np.repeat(one, 64, axis=0)
    reference_case("identical64", ident, targeted_rays(ident, 512, 300))

    flat = W.soup(300, seed=9)
    flat[:, 2::3] = 0.5  # every vertex z = 0.5: root box has zero z extent
    reference_case("flat300", flat, mixed_rays(flat, 1024, 400))

    # non-identity, non-contiguous primitive ids
    perm_tris = W.soup(777, seed=13)
    perm_ids = (np.random.default_rng(14).permutation(777) * 3 + 5).astype(np.int32)
    reference_case("sparse_ids777", perm_tris, mixed_rays(perm_tris, 1024, 500), ids=perm_ids)

    for k in (1, 2, 3, 7):
        t = W.soup(k, seed=20 + k) * np.float32(1.0)
        reference_case(f"tiny{k}", t, targeted_rays(t, 64, 600 + k))

    def squeeze(t, side):
        # move every triangle's v0 into [0, side]^3 keeping its edges: many overlaps
        shift = t[:, 0:3] * (side - 1.0)
        return (t + np.tile(shift, 3)).astype(np.float32)

    dense = squeeze(W.soup(2000, seed=31, edge=(0.002, 0.02)), 0.15)
    q = squeeze(W.soup(300, seed=32, edge=(0.002, 0.02)), 0.15)
    overlap_case("dense2000", W.triangle_boxes(dense), W.triangle_boxes(q))


if __name__ == "__main__":
    main()

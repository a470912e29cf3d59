"""The near-child-first closest-hit walk against the right-first preorder walk
of the host LBVH, on the CPU (oracle/ordered_model.cpp models the library's
walk). The rounding4 golden case holds a ray on which the walk without its
ambiguity check returns another triangle than the reference's host LBVH; with
the check the ray is walked again in preorder and the hit equals the
reference's."""

import numpy as np

import oracle
from conftest import load_golden
from lux_amd import workloads as W


def _tree(g):
    return oracle.OracleBvh(g["boxes"], g["ids"])


def test_plain_walk_diverges_on_rounding_case():
    g = load_golden("lbvh_rounding4")
    ids0, t0, flags0 = _tree(g).trace_ordered(g["tri_by_id"], g["rays"][:1], rule=0)
    # the reference keeps A (id 0); brute force and the plain walk find B (id 1)
    assert g["hit_ids"][0] == 0 and g["brute_ids"][0] == 1
    assert ids0[0] == 1 and flags0[0] == 0
    assert t0[0] < g["hit_t"][0]


def test_checked_walk_equals_reference_on_rounding_case():
    g = load_golden("lbvh_rounding4")
    ids, t, flags = _tree(g).trace_ordered(g["tri_by_id"], g["rays"], rule=1)
    assert flags[0] & 2, "the ambiguity check must fire on the divergent ray"
    assert np.array_equal(ids, g["hit_ids"])
    assert np.array_equal(t.view(np.uint32), g["hit_t"].view(np.uint32))


def test_checked_walk_equals_preorder_on_golden_cases():
    for case in ("soup2000", "clustered1500", "identical64", "flat300", "sparse_ids777"):
        g = load_golden(f"lbvh_{case}")
        ids, t, _ = _tree(g).trace_ordered(g["tri_by_id"], g["rays"], rule=1)
        assert np.array_equal(ids, g["hit_ids"]), case
        assert np.array_equal(t.view(np.uint32), g["hit_t"].view(np.uint32)), case


def test_rounding_case_caught_by_leaf_entry():
    """The best triangle the plain walk finds (B) sits below its own leaf box's
    entry, so L is that entry, and A's box, culled at t_A < L, fires the check
    with no window at all."""
    g = load_golden("lbvh_rounding4")
    ids, _, flags = _tree(g).trace_ordered(g["tri_by_id"], g["rays"][:1], rule=1, window_log2=60)
    assert flags[0] & 2 and ids[0] == g["hit_ids"][0]


def test_checked_walk_on_random_soup():
    tris = W.soup(1 << 15, seed=0x4C5558)
    boxes = W.triangle_boxes(tris)
    rays = W.rays(8192, seed=0x4C5558 + 2)
    tree = oracle.OracleBvh(boxes)
    pre = tree.trace_closest(tris, rays)
    ids, t, flags = tree.trace_ordered(tris, rays, rule=1)
    assert np.array_equal(ids, pre["ids"])
    assert np.array_equal(t.view(np.uint32), pre["t"].view(np.uint32))
    # few rays take the second walk
    assert (flags != 0).mean() < 0.02

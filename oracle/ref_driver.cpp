// ref_driver.cpp -- extern "C" shim over the reference's OWN host code, compiled
// from the LUX sources where they lie (see oracle/ref.mk). Nothing here
// re-implements an algorithm: every result comes from the reference's headers
// (build_cpu_lbvh_tree, intersect_cpu_bvh_tree, intersect_triangle,
// lbvh_detail::centroid_to_morton). It exists only to produce golden vectors
// (tests/golden/make_golden.py) and to time the reference CPU path as a
// baseline. TEST INFRASTRUCTURE ONLY.

#include <cstdint>
#include <cstring>
#include <vector>

#include "bvh/cpu_bvh_traversal.h"
#include "bvh/cpu_lbvh_builder.h"
#include "geometry/triangle.h"
#include "scene/scene_intersection.h"

static_assert(sizeof(AABB) == 24 && sizeof(Ray) == 36 && sizeof(Triangle) == 36, "layout");

extern "C" {

struct ref_tree {
    CpuBvhTree tree;
};

// build_cpu_lbvh_tree over primitives {bounds, bounds.centroid(), id} -- the
// record collect_scene_mesh_bvh_build_primitives makes for a triangle.
ref_tree* ref_lbvh_build(const float* boxes, const int32_t* ids, int32_t n) {
    const AABB* b = reinterpret_cast<const AABB*>(boxes);
    std::vector<BvhBuildPrimitive> prims;
    prims.reserve(n);
    for (int i = 0; i < n; ++i) {
        BvhBuildPrimitive p{};
        p.bounds = b[i];
        p.centroid = b[i].centroid();
        p.primitive_id = ids[i];
        p.instance_id = -1;
        p.mesh_id = 0;
        p.local_primitive_id = ids[i];
        p.primitive_kind = PrimitiveKind::Triangle;
        prims.push_back(p);
    }
    ref_tree* t = new ref_tree();
    t->tree = build_cpu_lbvh_tree(std::move(prims));
    return t;
}

void ref_lbvh_free(ref_tree* t) { delete t; }

int32_t ref_lbvh_node_count(const ref_tree* t) { return int32_t(t->tree.nodes.size()); }

// nodes: per node {bounds(6 floats)}, left/right/first/count ints; leaf_ids: n
void ref_lbvh_export(const ref_tree* t, float* bounds, int32_t* left, int32_t* right,
                     int32_t* first_primitive, int32_t* primitive_count, int32_t* leaf_ids) {
    const auto& nodes = t->tree.nodes;
    for (size_t i = 0; i < nodes.size(); ++i) {
        std::memcpy(bounds + 6 * i, &nodes[i].bounds, sizeof(AABB));
        left[i] = nodes[i].left;
        right[i] = nodes[i].right;
        first_primitive[i] = nodes[i].first_primitive;
        primitive_count[i] = nodes[i].primitive_count;
    }
    for (size_t i = 0; i < t->tree.primitive_ids.size(); ++i) leaf_ids[i] = t->tree.primitive_ids[i];
}

void ref_morton_codes(const float* boxes, int32_t n, const float* root, uint32_t* out) {
    const AABB* b = reinterpret_cast<const AABB*>(boxes);
    AABB rb;
    std::memcpy(&rb, root, sizeof(AABB));
    for (int i = 0; i < n; ++i) out[i] = lbvh_detail::centroid_to_morton(rb, b[i].centroid());
}

// intersect_cpu_bvh_tree with the scene's triangle leaf intersector
// (intersect_scene_triangle_primitive). Returns the number of stack overflows.
int32_t ref_trace_closest(const ref_tree* t, const float* tris, const float* rays, int32_t n,
                          int32_t* ids, float* t_out, float* u_out, float* v_out) {
    const Triangle* tr = reinterpret_cast<const Triangle*>(tris);
    const Ray* rs = reinterpret_cast<const Ray*>(rays);
    int32_t overflows = 0;
    for (int i = 0; i < n; ++i) {
        int id = -1;
        Float th, uh, vh;
        bool ov = false;
        bool found = intersect_cpu_bvh_tree(
            t->tree, rs[i],
            [tr](int prim, const Ray& ray, CpuBvhPrimitiveHit& hit) {
                return intersect_scene_triangle_primitive(tr[prim], ray, hit);
            },
            id, th, uh, vh, &ov);
        if (ov) ++overflows;
        ids[i] = found ? id : -1;
        t_out[i] = found ? th : INFINITY;
        u_out[i] = found ? uh : 0;
        v_out[i] = found ? vh : 0;
    }
    return overflows;
}

// The loop of intersect_scene_bruteforce over a plain triangle array.
void ref_brute_closest(const float* tris, int32_t tri_count, const float* rays, int32_t n,
                       int32_t* ids, float* t_out) {
    const Triangle* tr = reinterpret_cast<const Triangle*>(tris);
    const Ray* rs = reinterpret_cast<const Ray*>(rays);
    for (int i = 0; i < n; ++i) {
        Float best = INFINITY;
        int id = -1;
        for (int k = 0; k < tri_count; ++k) {
            Float tt, uu, vv;
            if (!intersect_triangle(rs[i], tr[k], tt, uu, vv)) continue;
            if (tt >= best) continue;
            best = tt;
            id = k;
        }
        ids[i] = id;
        t_out[i] = best;
    }
}

void ref_triangle_bounds(const float* tris, int32_t n, float* boxes) {
    const Triangle* tr = reinterpret_cast<const Triangle*>(tris);
    for (int i = 0; i < n; ++i) {
        AABB b = triangle_bounds(tr[i]);
        std::memcpy(boxes + 6 * i, &b, sizeof(AABB));
    }
}

int32_t ref_aabb_overlaps(const float* a, const float* b) {
    AABB x, y;
    std::memcpy(&x, a, sizeof(AABB));
    std::memcpy(&y, b, sizeof(AABB));
    return x.overlaps(y) ? 1 : 0;
}

// Self pairs (ids[i], ids[j]) with ids[j] > ids[i] whose boxes overlap under the
// reference's AABB::overlaps -- the O(n^2) loop of validate_gpu_bvh_self_overlaps.
int64_t ref_overlap_self_brute(const float* boxes, int32_t n, int32_t* pairs, int64_t cap) {
    const AABB* b = reinterpret_cast<const AABB*>(boxes);
    int64_t count = 0;
    for (int i = 0; i < n; ++i)
        for (int j = i + 1; j < n; ++j) {
            if (!b[i].overlaps(b[j])) continue;
            if (count < cap) {
                pairs[2 * count] = i;
                pairs[2 * count + 1] = j;
            }
            ++count;
        }
    return count;
}

// Cross pairs (left index, right index) under AABB::overlaps.
int64_t ref_overlap_cross_brute(const float* lboxes, int32_t nl, const float* rboxes, int32_t nr,
                                int32_t* pairs, int64_t cap) {
    const AABB* lb = reinterpret_cast<const AABB*>(lboxes);
    const AABB* rb = reinterpret_cast<const AABB*>(rboxes);
    int64_t count = 0;
    for (int j = 0; j < nr; ++j)
        for (int i = 0; i < nl; ++i) {
            if (!lb[i].overlaps(rb[j])) continue;
            if (count < cap) {
                pairs[2 * count] = i;
                pairs[2 * count + 1] = j;
            }
            ++count;
        }
    return count;
}

}  // extern "C"

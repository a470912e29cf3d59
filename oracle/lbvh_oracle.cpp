// lbvh_oracle.cpp -- CPU restatement of LUX's host LBVH, host stack traversal,
// brute-force intersection and overlap checks.
//
// TEST INFRASTRUCTURE ONLY. This file is the checker for the HIP kernels in
// lux_amd/csrc: only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
// leg may load the library built from it. The product path never calls it.
//
// Every function names the reference code it restates (paths relative to the
// LUX source tree). Arithmetic keeps the reference's operation order and is
// compiled with -ffp-contract=off so results are bit-identical to it.
//
// Pinning (what checks each function, besides the golden vectors produced from
// the reference's own host code and committed under tests/golden/):
//   morton_code      closed form (bit interleave), corner spot values
//   build            permutation/stability invariants, Karras-tree invariants,
//                    node box == union of children, root == union of inputs
//   preorder layout  escape == preorder + subtree size; DFS order equality
//   trace_closest    brute force over all triangles; FP64 Moller-Trumbore bound
//   overlap          O(n^2) brute force

#include "oracle_core.hpp"

using namespace orc;

extern "C" {

struct orc_bvh;

int orc_version() { return 1; }

// Per-primitive Morton code in input order for a given root box.
void orc_morton_codes(const float* boxes, int32_t n, const float* root, uint32_t* out) {
    const Box* b = reinterpret_cast<const Box*>(boxes);
    Box rb;
    std::memcpy(&rb, root, sizeof(Box));
    for (int i = 0; i < n; ++i) out[i] = morton_code(rb, b[i]);
}

orc_bvh* orc_bvh_build(const float* boxes, const int32_t* ids, int32_t n) {
    return reinterpret_cast<orc_bvh*>(build(reinterpret_cast<const Box*>(boxes), ids, n));
}

void orc_bvh_refit(orc_bvh* h, const float* boxes) {
    refit(*reinterpret_cast<Tree*>(h), reinterpret_cast<const Box*>(boxes));
}

void orc_bvh_free(orc_bvh* h) { delete reinterpret_cast<Tree*>(h); }

int32_t orc_bvh_count(const orc_bvh* h) { return reinterpret_cast<const Tree*>(h)->n; }

// Any output pointer may be null. Sizes as in include/lux_hip.h
// lux_hip_bvh_download_topology; root gets 6 floats; nodes gets 2n-1 32-byte
// records {box, prim, escape}.
void orc_bvh_topology(const orc_bvh* h, uint32_t* morton_sorted, uint32_t* perm,
                      int32_t* leaf_ids, uint32_t* left, uint32_t* right, uint32_t* parent,
                      uint32_t* leaf_parent, float* node_bounds, float* root, void* nodes) {
    const Tree* t = reinterpret_cast<const Tree*>(h);
    int n = t->n;
    if (root) std::memcpy(root, &t->root, sizeof(Box));
    if (n <= 0) return;
    if (morton_sorted) std::memcpy(morton_sorted, t->morton_sorted.data(), 4 * n);
    if (perm) std::memcpy(perm, t->perm.data(), 4 * n);
    if (leaf_ids) std::memcpy(leaf_ids, t->leaf_ids.data(), 4 * n);
    if (leaf_parent) std::memcpy(leaf_parent, t->leaf_parent.data(), 4 * n);
    if (n > 1) {
        if (left) std::memcpy(left, t->left.data(), 4 * (n - 1));
        if (right) std::memcpy(right, t->right.data(), 4 * (n - 1));
        if (parent) std::memcpy(parent, t->parent.data(), 4 * (n - 1));
    }
    if (node_bounds) std::memcpy(node_bounds, t->bounds.data(), sizeof(Box) * (2 * n - 1));
    if (nodes) std::memcpy(nodes, t->nodes.data(), sizeof(Tree::Node) * (2 * n - 1));
}

// Host stack traversal (closest hit). ids[i] = -1 and t = +INF on a miss.
// Returns the number of rays whose stack overflowed (their result is a miss,
// as in the reference). stats (nullable) receives {nodes visited, triangle tests}.
int32_t orc_trace_closest(const orc_bvh* h, const float* tris, const float* rays, int32_t n,
                          int32_t* ids, float* t_out, float* u_out, float* v_out,
                          int64_t* stats) {
    const Tree* t = reinterpret_cast<const Tree*>(h);
    const Tri* tr = reinterpret_cast<const Tri*>(tris);
    const Ray* rs = reinterpret_cast<const Ray*>(rays);
    TraceStats st;
    int32_t overflows = 0;
    for (int i = 0; i < n; ++i) {
        int id;
        float th, uh, vh;
        bool ov;
        bool found = trace_stack(*t, tr, rs[i], id, th, uh, vh, ov, st);
        if (ov) ++overflows;
        ids[i] = found ? id : -1;
        t_out[i] = found ? th : INFINITY;
        if (u_out) u_out[i] = found ? uh : 0;
        if (v_out) v_out[i] = found ? vh : 0;
    }
    if (stats) {
        stats[0] = st.nodes;
        stats[1] = st.tris;
    }
    return overflows;
}

// Same rays through the right-first stackless layout (must equal the above).
void orc_trace_stackless(const orc_bvh* h, const float* tris, const float* rays, int32_t n,
                         int32_t* ids, float* t_out) {
    const Tree* t = reinterpret_cast<const Tree*>(h);
    const Tri* tr = reinterpret_cast<const Tri*>(tris);
    const Ray* rs = reinterpret_cast<const Ray*>(rays);
    for (int i = 0; i < n; ++i) {
        int id;
        float th;
        bool found = trace_stackless(*t, tr, rs[i], id, th);
        ids[i] = found ? id : -1;
        t_out[i] = found ? th : INFINITY;
    }
}

// intersect_scene_bruteforce src/scene/scene_intersection.h:30-52: all
// triangles in index order, keep a hit only on strict t < best.
void orc_brute_closest(const float* tris, int32_t tri_count, const float* rays, int32_t n,
                       int32_t* ids, float* t_out) {
    const Tri* tr = reinterpret_cast<const Tri*>(tris);
    const Ray* rs = reinterpret_cast<const Ray*>(rays);
    for (int i = 0; i < n; ++i) {
        Ray r = rs[i];
        float best = INFINITY;
        int id = -1;
        for (int k = 0; k < tri_count; ++k) {
            float tt, uu, vv;
            if (!tri_hit(r, tr[k], tt, uu, vv)) continue;
            if (tt >= best) continue;
            best = tt;
            id = k;
        }
        ids[i] = id;
        t_out[i] = best;
    }
}

// FP64 Moller-Trumbore t for a given ray/triangle pair (error statement for the
// FP32 hit distance). Returns NaN when the FP64 test misses.
void orc_tri_t_fp64(const float* tris, const int32_t* tri_ids, const float* rays, int32_t n,
                    double* t_out) {
    const Tri* tr = reinterpret_cast<const Tri*>(tris);
    const Ray* rs = reinterpret_cast<const Ray*>(rays);
    for (int i = 0; i < n; ++i) {
        if (tri_ids[i] < 0) {
            t_out[i] = NAN;
            continue;
        }
        const Tri& T = tr[tri_ids[i]];
        const Ray& R = rs[i];
        double v0[3] = {T.v0.x, T.v0.y, T.v0.z}, v1[3] = {T.v1.x, T.v1.y, T.v1.z},
               v2[3] = {T.v2.x, T.v2.y, T.v2.z};
        double o[3] = {R.o.x, R.o.y, R.o.z}, d[3] = {R.d.x, R.d.y, R.d.z};
        double e1[3], e2[3], s[3];
        for (int k = 0; k < 3; ++k) {
            e1[k] = v1[k] - v0[k];
            e2[k] = v2[k] - v0[k];
            s[k] = o[k] - v0[k];
        }
        double hh[3] = {d[1] * e2[2] - d[2] * e2[1], d[2] * e2[0] - d[0] * e2[2], d[0] * e2[1] - d[1] * e2[0]};
        double a = e1[0] * hh[0] + e1[1] * hh[1] + e1[2] * hh[2];
        double q[3] = {s[1] * e1[2] - s[2] * e1[1], s[2] * e1[0] - s[0] * e1[2], s[0] * e1[1] - s[1] * e1[0]};
        double tt = (e2[0] * q[0] + e2[1] * q[1] + e2[2] * q[2]) / a;
        t_out[i] = a == 0 ? NAN : tt;
    }
}

// Self overlap pairs (a < b by primitive id) by O(n^2) brute force
// (validate_gpu_bvh_self_overlaps src/bvh/bvh_validation.h:195-203, with ids).
// Writes at most cap pairs; returns the total count.
int64_t orc_overlap_self_brute(const float* boxes, const int32_t* ids, int32_t n, int32_t* pairs,
                               int64_t cap) {
    const Box* b = reinterpret_cast<const Box*>(boxes);
    int64_t count = 0;
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            if (!(ids[j] > ids[i])) continue;
            if (!overlaps(b[i], b[j])) continue;
            if (count < cap) {
                pairs[2 * count] = ids[i];
                pairs[2 * count + 1] = ids[j];
            }
            ++count;
        }
    return count;
}

// Cross pairs (left id, right id) by O(nl*nr) brute force.
int64_t orc_overlap_cross_brute(const float* lboxes, const int32_t* lids, int32_t nl,
                                const float* rboxes, const int32_t* rids, int32_t nr,
                                int32_t* pairs, int64_t cap) {
    const Box* lb = reinterpret_cast<const Box*>(lboxes);
    const Box* rb = reinterpret_cast<const Box*>(rboxes);
    int64_t count = 0;
    for (int j = 0; j < nr; ++j)
        for (int i = 0; i < nl; ++i) {
            if (!overlaps(lb[i], rb[j])) continue;
            if (count < cap) {
                pairs[2 * count] = lids[i];
                pairs[2 * count + 1] = rids[j];
            }
            ++count;
        }
    return count;
}

// Tree-accelerated overlap (the query rule of src/bvh/bvh_overlap.cu:25-105 /
// :107-184 over the stackless layout), for inputs too large for brute force.
// self: queries are the tree's own inputs; emits (q, p) only for p > q.
int64_t orc_overlap_tree(const orc_bvh* h, const float* qboxes, const int32_t* qids, int32_t nq,
                         int32_t self_mode, int32_t* pairs, int64_t cap) {
    const Tree* t = reinterpret_cast<const Tree*>(h);
    const Box* qb = reinterpret_cast<const Box*>(qboxes);
    int total = int(t->nodes.size());
    int64_t count = 0;
    for (int q = 0; q < nq; ++q) {
        int k = 0;
        while (k >= 0 && k < total) {
            const Tree::Node& nd = t->nodes[k];
            if (!overlaps(nd.b, qb[q])) {
                k = nd.escape;
                continue;
            }
            if (nd.prim >= 0) {
                bool emit = self_mode ? nd.prim > qids[q] : true;
                if (emit) {
                    if (count < cap) {
                        pairs[2 * count] = self_mode ? qids[q] : nd.prim;
                        pairs[2 * count + 1] = self_mode ? nd.prim : qids[q];
                    }
                    ++count;
                }
                k = nd.escape;
            } else {
                k += 1;
            }
        }
    }
    return count;
}

// AABB of each triangle (triangle_bounds src/geometry/triangle.h:22-28).
void orc_triangle_bounds(const float* tris, int32_t n, float* boxes) {
    const Tri* tr = reinterpret_cast<const Tri*>(tris);
    Box* b = reinterpret_cast<Box*>(boxes);
    for (int i = 0; i < n; ++i) {
        Box x = empty_box();
        x.lo = vmin(x.lo, tr[i].v0);
        x.hi = vmax(x.hi, tr[i].v0);
        x.lo = vmin(x.lo, tr[i].v1);
        x.hi = vmax(x.hi, tr[i].v1);
        x.lo = vmin(x.lo, tr[i].v2);
        x.hi = vmax(x.hi, tr[i].v2);
        b[i] = x;
    }
}

}  // extern "C"

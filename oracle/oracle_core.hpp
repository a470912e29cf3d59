// oracle_core.hpp -- shared core of the CPU oracle: float3 math in the
// reference's operation order, host LBVH build (right-first stackless layout),
// stack traversal, refit. TEST INFRASTRUCTURE ONLY (see lbvh_oracle.cpp).
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace orc {

struct V3 {
    float x, y, z;
};
inline V3 v3(float x, float y, float z) { return V3{x, y, z}; }
inline V3 sub(const V3& a, const V3& b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
inline V3 add(const V3& a, const V3& b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
inline V3 scale(const V3& a, float s) { return v3(a.x * s, a.y * s, a.z * s); }
inline float comp(const V3& a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
// dot/cross: src/core/vec3.cuh:40-50 (left-to-right sums, no fused ops)
inline float dot(const V3& a, const V3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(const V3& a, const V3& b) {
    return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
inline V3 vmin(const V3& a, const V3& b) {
    return v3(fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z));
}
inline V3 vmax(const V3& a, const V3& b) {
    return v3(fmaxf(a.x, b.x), fmaxf(a.y, b.y), fmaxf(a.z, b.z));
}

struct Box {
    V3 lo, hi;
};
struct Ray {
    V3 o, d;
    float t_min, t_max, time;
};
struct Tri {
    V3 v0, v1, v2;
};
static_assert(sizeof(Box) == 24 && sizeof(Ray) == 36 && sizeof(Tri) == 36, "layout");

inline Box empty_box() { return Box{v3(INFINITY, INFINITY, INFINITY), v3(-INFINITY, -INFINITY, -INFINITY)}; }
// AABB::expand(AABB) src/core/aabb.cuh:50-53; union_aabb :113-115
inline Box join(const Box& a, const Box& b) { return Box{vmin(a.lo, b.lo), vmax(a.hi, b.hi)}; }
// AABB::center src/core/aabb.cuh:18
inline V3 centroid(const Box& b) { return scale(add(b.lo, b.hi), 0.5f); }

// AABB::normalized_offset src/core/aabb.cuh:74-80
inline V3 normalized_offset(const Box& root, const V3& p) {
    V3 d = sub(root.hi, root.lo);
    return v3(d.x > 0 ? (p.x - root.lo.x) / d.x : 0.5f,
              d.y > 0 ? (p.y - root.lo.y) / d.y : 0.5f,
              d.z > 0 ? (p.z - root.lo.z) / d.z : 0.5f);
}

// expand_bits_10 / morton_encode_3d / quantize_morton_coord / centroid_to_morton
// src/bvh/cpu_lbvh_builder.h:37-61. Spread the 10 low bits of v so that bit k
// lands at bit 3k (restated bit by bit rather than with magic multipliers).
inline uint32_t spread3(uint32_t v) {
    v &= 0x3ffu;
    uint32_t r = 0;
    for (int k = 0; k < 10; ++k) r |= ((v >> k) & 1u) << (3 * k);
    return r;
}
inline uint32_t quantize(float value) {
    value = fmaxf(0.0f, fminf(1023.0f, value * 1024.0f));
    return static_cast<uint32_t>(value);
}
inline uint32_t morton_code(const Box& root, const Box& b) {
    V3 off = normalized_offset(root, centroid(b));
    return (spread3(quantize(off.x)) << 2) | (spread3(quantize(off.y)) << 1) | spread3(quantize(off.z));
}

// AABB::intersect (slab test) src/core/aabb.cuh:84-109
inline bool slab(const Box& b, const Ray& r) {
    float t_near = r.t_min, t_far = r.t_max;
    for (int i = 0; i < 3; ++i) {
        float d = comp(r.d, i), o = comp(r.o, i);
        float lo = comp(b.lo, i), hi = comp(b.hi, i);
        if (fabsf(d) < 1e-12f) {
            if (o < lo || o > hi) return false;
            continue;
        }
        float inv = 1 / d;
        float a = (lo - o) * inv;
        float c = (hi - o) * inv;
        if (inv < 0) std::swap(a, c);
        t_near = fmaxf(t_near, a);
        t_far = fminf(t_far, c);
        if (t_near > t_far) return false;
    }
    return true;
}

// intersect_triangle (Moller-Trumbore) src/geometry/triangle.h:33-52
inline bool tri_hit(const Ray& r, const Tri& tri, float& t, float& u, float& v) {
    V3 e1 = sub(tri.v1, tri.v0);
    V3 e2 = sub(tri.v2, tri.v0);
    V3 h = cross(r.d, e2);
    float a = dot(e1, h);
    if (fabsf(a) < 1e-7f) return false;
    float f = 1.0f / a;
    V3 s = sub(r.o, tri.v0);
    u = f * dot(s, h);
    if (u < 0.0f || u > 1.0f) return false;
    V3 q = cross(s, e1);
    v = f * dot(r.d, q);
    if (v < 0.0f || u + v > 1.0f) return false;
    t = f * dot(e2, q);
    return t > r.t_min && t < r.t_max;
}

// AABB::overlaps src/core/aabb.cuh:68-72 (inclusive)
inline bool overlaps(const Box& a, const Box& b) {
    return !(b.lo.x > a.hi.x || b.hi.x < a.lo.x || b.lo.y > a.hi.y || b.hi.y < a.lo.y ||
             b.lo.z > a.hi.z || b.hi.z < a.lo.z);
}

constexpr uint32_t kLeaf = 0x80000000u;
constexpr uint32_t kIndex = 0x7fffffffu;

struct Tree {
    int n = 0;
    Box root = empty_box();
    std::vector<uint32_t> morton_sorted;  // n
    std::vector<uint32_t> perm;           // n: sorted position -> input index
    std::vector<int32_t> leaf_ids;        // n: primitive id of sorted leaf
    std::vector<uint32_t> left, right;    // n-1, bit 31 = leaf child
    std::vector<uint32_t> parent;         // n-1, bit 31 = right child (root: 0)
    std::vector<uint32_t> leaf_parent;    // n
    std::vector<Box> bounds;              // 2n-1 host layout (internal, then leaves)
    // right-first preorder stackless layout
    struct Node {
        Box b;
        int32_t prim;
        int32_t escape;
    };
    std::vector<Node> nodes;  // 2n-1
    std::vector<int> pre_of;  // host node index -> position in `nodes`
};
static_assert(sizeof(Tree::Node) == 32, "node layout");

// merged_key / common_prefix src/bvh/cpu_lbvh_builder.h:63-81
inline int delta(const std::vector<uint32_t>& codes, int a, int b) {
    uint64_t ka = (uint64_t(codes[a]) << 32) | uint64_t(uint32_t(a));
    uint64_t kb = (uint64_t(codes[b]) << 32) | uint64_t(uint32_t(b));
    uint64_t x = ka ^ kb;
    return x == 0 ? 64 : __builtin_clzll(x);
}

// determine_range src/bvh/cpu_lbvh_builder.h:83-111
inline void node_range(const std::vector<uint32_t>& codes, int i, int& first, int& last) {
    int n = int(codes.size());
    if (i == 0) {
        first = 0;
        last = n - 1;
        return;
    }
    int dl = delta(codes, i, i - 1), dr = delta(codes, i, i + 1);
    int dir = dr > dl ? 1 : -1;
    int dmin = std::min(dl, dr);
    int lmax = 2;
    for (;;) {
        int j = i + dir * lmax;
        if (j < 0 || j >= n || delta(codes, i, j) <= dmin) break;
        lmax <<= 1;
    }
    int len = 0;
    for (int step = lmax >> 1; step > 0; step >>= 1) {
        int j = i + (len + step) * dir;
        if (j < 0 || j >= n) continue;
        if (delta(codes, i, j) > dmin) len += step;
    }
    int other = i + len * dir;
    first = dir < 0 ? other : i;
    last = dir < 0 ? i : other;
}

// find_split src/bvh/cpu_lbvh_builder.h:113-127
inline int node_split(const std::vector<uint32_t>& codes, int first, int last) {
    int dnode = delta(codes, first, last);
    int split = first, stride = last - first;
    do {
        stride = (stride + 1) >> 1;
        int mid = split + stride;
        if (mid < last && delta(codes, first, mid) > dnode) split = mid;
    } while (stride > 1);
    return split;
}

inline int host_index(uint32_t child, int n_internal) {
    return (child & kLeaf) ? n_internal + int(child & kIndex) : int(child);
}

// build_cpu_lbvh_tree src/bvh/cpu_lbvh_builder.h:163-259 (+ the radix-node
// parent links of src/bvh/gpu_bvh.cu:151-180) followed by the right-first
// preorder stackless layout.
inline Tree* build(const Box* boxes, const int32_t* ids, int n) {
    Tree* t = new Tree();
    t->n = n;
    if (n <= 0) return t;
    for (int i = 0; i < n; ++i) t->root = join(t->root, boxes[i]);

    if (n == 1) {  // single leaf: src/bvh/cpu_lbvh_builder.h:184-195
        t->perm = {0};
        t->leaf_ids = {ids[0]};
        t->morton_sorted = {0};  // never computed by the reference for n == 1
        t->leaf_parent = {0};
        t->bounds = {boxes[0]};
        t->nodes = {Tree::Node{boxes[0], ids[0], -1}};
        return t;
    }

    std::vector<uint32_t> code(n);
    for (int i = 0; i < n; ++i) code[i] = morton_code(t->root, boxes[i]);
    std::vector<uint32_t> order(n);
    for (int i = 0; i < n; ++i) order[i] = uint32_t(i);
    std::stable_sort(order.begin(), order.end(),
                     [&](uint32_t a, uint32_t b) { return code[a] < code[b]; });
    t->perm = order;
    t->morton_sorted.resize(n);
    t->leaf_ids.resize(n);
    for (int i = 0; i < n; ++i) {
        t->morton_sorted[i] = code[order[i]];
        t->leaf_ids[i] = ids[order[i]];
    }

    int ni = n - 1;
    t->left.assign(ni, 0);
    t->right.assign(ni, 0);
    t->parent.assign(ni, 0);
    t->leaf_parent.assign(n, 0);
    for (int i = 0; i < ni; ++i) {
        int first, last;
        node_range(t->morton_sorted, i, first, last);
        int split = node_split(t->morton_sorted, first, last);
        if (first == split) {
            t->left[i] = kLeaf | uint32_t(split);
            t->leaf_parent[split] = uint32_t(i);
        } else {
            t->left[i] = uint32_t(split);
            t->parent[split] = uint32_t(i);
        }
        if (last == split + 1) {
            t->right[i] = kLeaf | uint32_t(split + 1);
            t->leaf_parent[split + 1] = uint32_t(i) | kLeaf;
        } else {
            t->right[i] = uint32_t(split + 1);
            t->parent[split + 1] = uint32_t(i) | kLeaf;
        }
    }

    // Bounds: leaves are the input boxes in sorted order; internal nodes are the
    // union of their children (fill_internal_bounds, :139-159), done here with an
    // explicit post-order stack.
    t->bounds.assign(2 * n - 1, empty_box());
    for (int j = 0; j < n; ++j) t->bounds[ni + j] = boxes[order[j]];
    {
        std::vector<std::pair<int, int>> st;  // (internal node, state)
        st.push_back({0, 0});
        while (!st.empty()) {
            auto& top = st.back();
            int i = top.first;
            if (top.second == 0) {
                top.second = 1;
                if (!(t->right[i] & kLeaf)) st.push_back({int(t->right[i]), 0});
                if (!(t->left[i] & kLeaf)) st.push_back({int(t->left[i]), 0});
            } else {
                t->bounds[i] = join(t->bounds[host_index(t->left[i], ni)],
                                    t->bounds[host_index(t->right[i], ni)]);
                st.pop_back();
            }
        }
    }

    // Right-first preorder: DFS from the root that emits a node, then its right
    // subtree, then its left subtree -- the pop order of the host stack
    // traversal (src/bvh/cpu_bvh_traversal.h:76-77). escape = index of the node
    // that follows the whole subtree, or -1 past the end.
    int total = 2 * n - 1;
    t->nodes.assign(total, Tree::Node{empty_box(), -1, -1});
    std::vector<int>& pre_of = t->pre_of;
    pre_of.assign(total, -1);
    std::vector<int> size_of(total, 1);
    // subtree sizes (post-order over internal nodes: children have been sized
    // when we visit in reverse preorder)
    std::vector<int> order_pre;
    order_pre.reserve(total);
    {
        std::vector<int> st{0};
        while (!st.empty()) {
            int h = st.back();
            st.pop_back();
            order_pre.push_back(h);
            if (h < ni) {
                st.push_back(host_index(t->left[h], ni));
                st.push_back(host_index(t->right[h], ni));
            }
        }
    }
    for (int k = total - 1; k >= 0; --k) {
        int h = order_pre[k];
        if (h < ni) size_of[h] = 1 + size_of[host_index(t->left[h], ni)] + size_of[host_index(t->right[h], ni)];
    }
    for (int k = 0; k < total; ++k) {
        int h = order_pre[k];
        pre_of[h] = k;
        int esc = k + size_of[h];
        t->nodes[k] = Tree::Node{t->bounds[h], h < ni ? -1 : t->leaf_ids[h - ni], esc >= total ? -1 : esc};
    }
    return t;
}

struct TraceStats {
    int64_t nodes = 0;
    int64_t tris = 0;
};

// intersect_cpu_bvh_nodes src/bvh/cpu_bvh_traversal.h:17-81 with the leaf
// intersector of src/scene/scene_intersection.h:144-153 (t shrink on strict <).
inline bool trace_stack(const Tree& t, const Tri* tris, Ray r, int& id, float& th, float& uh, float& vh,
                 bool& overflow, TraceStats& st) {
    constexpr int kCap = 64;
    int stack[kCap];
    int sp = 0;
    overflow = false;
    id = -1;
    th = r.t_max;
    uh = vh = 0;
    if (t.n <= 0) return false;
    int ni = t.n - 1;
    stack[sp++] = 0;
    bool found = false;
    while (sp > 0) {
        int h = stack[--sp];
        ++st.nodes;
        if (!slab(t.bounds[h], r)) continue;
        if (h >= ni) {
            int prim = t.leaf_ids[h - ni];
            float tt = 0, uu = 0, vv = 0;
            ++st.tris;
            if (!tri_hit(r, tris[prim], tt, uu, vv)) continue;
            if (tt >= th) continue;
            found = true;
            th = tt;
            uh = uu;
            vh = vv;
            id = prim;
            r.t_max = tt;
            continue;
        }
        if (sp + 2 > kCap) {
            overflow = true;
            id = -1;
            th = r.t_max;
            return false;
        }
        stack[sp++] = host_index(t.left[h], ni);
        stack[sp++] = host_index(t.right[h], ni);
    }
    return found;
}

// refit_gpu_lbvh (src/bvh/gpu_bvh.cu:597-613): same permutation and topology,
// leaf boxes replaced by the new input boxes, internal boxes re-unioned.
inline void refit(Tree& t, const Box* boxes) {
    int n = t.n;
    if (n <= 0) return;
    if (n == 1) {
        t.bounds[0] = boxes[0];
        t.nodes[0].b = boxes[0];
        t.root = boxes[0];
        return;
    }
    int ni = n - 1;
    for (int j = 0; j < n; ++j) t.bounds[ni + j] = boxes[t.perm[j]];
    // children always sit later in preorder than their parent: a reverse
    // preorder sweep is a valid post-order for the union
    std::vector<int> host_at(2 * n - 1);
    for (int h = 0; h < 2 * n - 1; ++h) host_at[t.pre_of[h]] = h;
    for (int k = 2 * n - 2; k >= 0; --k) {
        int h = host_at[k];
        if (h < ni) t.bounds[h] = join(t.bounds[host_index(t.left[h], ni)], t.bounds[host_index(t.right[h], ni)]);
    }
    for (int h = 0; h < 2 * n - 1; ++h) t.nodes[t.pre_of[h]].b = t.bounds[h];
    t.root = empty_box();
    for (int i = 0; i < n; ++i) t.root = join(t.root, boxes[i]);
}

// Stackless walk over the right-first preorder layout (the GPU visiting rule of
// src/bvh/bvh_device.cuh:59-87 applied to the right-first node order).
inline bool trace_stackless(const Tree& t, const Tri* tris, Ray r, int& id, float& th) {
    id = -1;
    th = r.t_max;
    bool found = false;
    int total = int(t.nodes.size());
    int k = 0;
    while (k >= 0 && k < total) {
        const Tree::Node& nd = t.nodes[k];
        if (!slab(nd.b, r)) {
            k = nd.escape;
            continue;
        }
        if (nd.prim >= 0) {
            float tt, uu, vv;
            if (tri_hit(r, tris[nd.prim], tt, uu, vv) && tt < th) {
                found = true;
                th = tt;
                id = nd.prim;
                r.t_max = tt;
            }
            k = nd.escape;
        } else {
            k += 1;
        }
    }
    return found;
}

}  // namespace orc

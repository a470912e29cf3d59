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

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

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
void node_range(const std::vector<uint32_t>& codes, int i, int& first, int& last) {
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
int node_split(const std::vector<uint32_t>& codes, int first, int last) {
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
Tree* build(const Box* boxes, const int32_t* ids, int n) {
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
bool trace_stack(const Tree& t, const Tri* tris, Ray r, int& id, float& th, float& uh, float& vh,
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
void refit(Tree& t, const Box* boxes) {
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
bool trace_stackless(const Tree& t, const Tri* tris, Ray r, int& id, float& th) {
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

}  // namespace

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

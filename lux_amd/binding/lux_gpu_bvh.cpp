// lux_gpu_bvh.cpp -- the reference's GPU LBVH and overlap interfaces on the
// MI355X library. In a LUX build this file replaces src/bvh/gpu_bvh.cu and
// src/bvh/bvh_overlap.cu; their headers (src/bvh/gpu_bvh.cuh:20-61,
// src/bvh/bvh_overlap.cuh:10-23) and every caller (PreparedScene::prepare_gpu,
// the validators of src/bvh/bvh_validation.h, the CLI) stay as they are.
//
// Build: compile with the reference's include path (src/) and include/ of
// this repository, link liblux_hip.so. The records cross the C ABI unchanged:
// AABB, Ray, Triangle, GpuRayHit, GpuBvhNode and PrimitivePair have the byte
// layouts of lux_aabb, lux_ray, lux_triangle, lux_hit, lux_bvh_node and
// lux_pair (checked below).
//
// Behaviour kept from the reference:
//   * an empty primitive list gives an invalid tree with primitive_count 0;
//   * refit_gpu_lbvh returns false on a count or id mismatch (true for an
//     empty list on an invalid tree);
//   * overlap queries return an empty result for an invalid tree or
//     max_pairs <= 0, and truncated = found_count > max_pairs;
//   * a failing device call throws std::runtime_error naming the call, as
//     check_cuda_or_throw does (src/core/cuda_check.cuh:7-27);
//   * the validators return bool with an optional message.
// The stackless node array is in right-first preorder (the visit order of the
// host LBVH, src/bvh/cpu_bvh_traversal.h:76-77), so its escapes and closest
// hits follow the host tree; the validators' checks hold unchanged.
#include <algorithm>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "bvh/bvh_overlap.cuh"
#include "bvh/gpu_bvh.cuh"
#include "lux_hip.h"

static_assert(sizeof(AABB) == sizeof(lux_aabb), "AABB layout");
static_assert(sizeof(Ray) == sizeof(lux_ray), "Ray layout");
static_assert(sizeof(Triangle) == sizeof(lux_triangle), "Triangle layout");
static_assert(sizeof(GpuRayHit) == sizeof(lux_hit), "GpuRayHit layout");
static_assert(sizeof(GpuBvhNode) == sizeof(lux_bvh_node), "GpuBvhNode layout");
static_assert(sizeof(PrimitivePair) == sizeof(lux_pair), "PrimitivePair layout");

// The library handle behind a GpuBvh.
struct GpuBvhStorage {
    lux_hip_bvh* handle = nullptr;
    GpuBvhStorage() = default;
    GpuBvhStorage(const GpuBvhStorage&) = delete;
    GpuBvhStorage& operator=(const GpuBvhStorage&) = delete;
    ~GpuBvhStorage() {
        if (handle) lux_hip_bvh_destroy(handle);
    }
};

namespace {

constexpr uint32_t kLeafBit = 0x80000000u;
constexpr uint32_t kIndexBits = 0x7fffffffu;

AABB to_aabb(const lux_aabb& b) {
    return AABB(vec3(b.p_min.x, b.p_min.y, b.p_min.z), vec3(b.p_max.x, b.p_max.y, b.p_max.z));
}

void check(int status, const char* what) {
    if (status != LUX_OK) throw std::runtime_error(std::string("MI355X error after ") + what + ": " +
                                                   lux_hip_last_error());
}

// The tree's handle, created (with the library's one-time init) on first use.
lux_hip_bvh* handle_of(const GpuBvh& bvh) {
    if (!bvh.storage) throw std::runtime_error("GpuBvh has no storage (moved from)");
    if (!bvh.storage->handle) {
        check(lux_hip_init(0), "lux_hip_init");
        check(lux_hip_bvh_create(&bvh.storage->handle), "lux_hip_bvh_create");
    }
    return bvh.storage->handle;
}

void fill_input(const std::vector<BvhBuildPrimitive>& prims, std::vector<AABB>& boxes, std::vector<int32_t>& ids) {
    boxes.resize(prims.size());
    ids.resize(prims.size());
    for (size_t i = 0; i < prims.size(); ++i) {
        boxes[i] = prims[i].bounds;
        ids[i] = prims[i].primitive_id;
    }
}

void refresh_info(GpuBvh& bvh) {
    int32_t valid = 0, count = 0;
    lux_aabb root{};
    check(lux_hip_bvh_info(handle_of(bvh), &valid, &count, &root), "lux_hip_bvh_info");
    bvh.primitive_count = valid ? count : 0;
    bvh.root_bounds = to_aabb(root);
}

// Boxes equal within 1e-4 (corner distance), the tolerance of the reference's
// validators.
bool boxes_close(const AABB& a, const AABB& b) {
    const Float eps = 1e-4f;
    return length2(a.p_min - b.p_min) <= eps * eps && length2(a.p_max - b.p_max) <= eps * eps;
}

bool fail(std::string* message, const char* what) {
    if (message) *message = what;
    return false;
}

// Everything download_bvh_for_validation and the structure check read.
struct HostTree {
    int n = 0;
    AABB root;
    std::vector<AABB> input_bounds;
    std::vector<int32_t> primitive_ids;
    std::vector<uint32_t> morton, sorted, left, right, parent, leaf_parent;
    std::vector<AABB> node_bounds;     // host order: internal 0..n-2, leaf j at n-1+j
    std::vector<GpuBvhNode> stackless; // 2n-1
};

HostTree download(const GpuBvh& bvh) {
    HostTree t;
    lux_hip_bvh* h = handle_of(bvh);
    int32_t valid = 0, count = 0;
    lux_aabb root{};
    check(lux_hip_bvh_info(h, &valid, &count, &root), "lux_hip_bvh_info");
    if (!valid || count <= 0) return t;
    t.n = count;
    t.root = to_aabb(root);
    const size_t n = size_t(count), m = size_t(std::max(count - 1, 0));
    t.input_bounds.resize(n);
    t.primitive_ids.resize(n);
    check(lux_hip_bvh_download_input(h, reinterpret_cast<lux_aabb*>(t.input_bounds.data()), t.primitive_ids.data()),
          "lux_hip_bvh_download_input");
    t.morton.resize(n);
    t.sorted.resize(n);
    t.leaf_parent.resize(n);
    t.left.resize(m);
    t.right.resize(m);
    t.parent.resize(m);
    t.node_bounds.resize(2 * n - 1);
    check(lux_hip_bvh_download_topology(h, t.morton.data(), t.sorted.data(), t.left.data(), t.right.data(),
                                        t.parent.data(), t.leaf_parent.data(),
                                        reinterpret_cast<lux_aabb*>(t.node_bounds.data())),
          "lux_hip_bvh_download_topology");
    t.stackless.resize(2 * n - 1);
    check(lux_hip_bvh_download_nodes(h, reinterpret_cast<lux_bvh_node*>(t.stackless.data()), int32_t(2 * n - 1)),
          "lux_hip_bvh_download_nodes");
    return t;
}

int child_index(uint32_t child, int internal_count) {
    return (child & kLeafBit) ? internal_count + int(child & kIndexBits) : int(child);
}

}  // namespace

GpuBvh::GpuBvh() : storage(std::make_unique<GpuBvhStorage>()) {}
GpuBvh::~GpuBvh() = default;
GpuBvh::GpuBvh(GpuBvh&& other) noexcept = default;
GpuBvh& GpuBvh::operator=(GpuBvh&& other) noexcept = default;

bool GpuBvh::is_valid() const {
    if (!storage || !storage->handle || primitive_count <= 0) return false;
    int32_t valid = 0;
    return lux_hip_bvh_info(storage->handle, &valid, nullptr, nullptr) == LUX_OK && valid != 0;
}

GpuBvhView GpuBvh::device_view() const {
    if (!is_valid()) return GpuBvhView{};
    lux_bvh_view v{};
    check(lux_hip_bvh_view(storage->handle, &v), "lux_hip_bvh_view");
    GpuBvhView out;
    out.nodes = reinterpret_cast<const GpuBvhNode*>(v.nodes);
    out.node_count = v.node_count;
    out.root_bounds = root_bounds;
    return out;
}

const AABB* GpuBvh::overlap_bounds_device_ptr() const {
    if (!is_valid()) return nullptr;
    const lux_aabb* p = nullptr;
    check(lux_hip_bvh_device_arrays(storage->handle, &p, nullptr, nullptr), "lux_hip_bvh_device_arrays");
    return reinterpret_cast<const AABB*>(p);
}

const int* GpuBvh::overlap_primitive_ids_device_ptr() const {
    if (!is_valid()) return nullptr;
    const int32_t* p = nullptr;
    check(lux_hip_bvh_device_arrays(storage->handle, nullptr, &p, nullptr), "lux_hip_bvh_device_arrays");
    return p;
}

const uint32_t* GpuBvh::overlap_sorted_indices_device_ptr() const {
    if (!is_valid()) return nullptr;
    const uint32_t* p = nullptr;
    check(lux_hip_bvh_device_arrays(storage->handle, nullptr, nullptr, &p), "lux_hip_bvh_device_arrays");
    return p;
}

GpuBvh build_gpu_lbvh(const std::vector<BvhBuildPrimitive>& primitives) {
    GpuBvh bvh;
    if (primitives.empty()) return bvh;
    std::vector<AABB> boxes;
    std::vector<int32_t> ids;
    fill_input(primitives, boxes, ids);
    check(lux_hip_bvh_build(handle_of(bvh), reinterpret_cast<const lux_aabb*>(boxes.data()), ids.data(),
                            int32_t(primitives.size())),
          "build_gpu_lbvh");
    refresh_info(bvh);
    return bvh;
}

bool refit_gpu_lbvh(GpuBvh& bvh, const std::vector<BvhBuildPrimitive>& primitives) {
    if (static_cast<int>(primitives.size()) != bvh.primitive_count) return false;
    if (!bvh.is_valid()) return primitives.empty();
    std::vector<AABB> boxes;
    std::vector<int32_t> ids;
    fill_input(primitives, boxes, ids);
    const int status = lux_hip_bvh_refit(bvh.storage->handle, reinterpret_cast<const lux_aabb*>(boxes.data()),
                                         ids.data(), int32_t(primitives.size()));
    if (status == LUX_ERR_MISMATCH) return false;
    check(status, "refit_gpu_lbvh");
    refresh_info(bvh);
    return true;
}

CpuBvhTree download_bvh_for_validation(const GpuBvh& bvh) {
    CpuBvhTree out;
    if (!bvh.is_valid()) return out;
    const HostTree t = download(bvh);
    if (t.n <= 0) return out;
    out.root_node = 0;
    if (t.n == 1) {
        out.primitive_ids.push_back(t.primitive_ids[0]);
        out.nodes.push_back(CpuBvhNode{t.input_bounds[0], -1, -1, 0, 1});
        return out;
    }
    const int internal = t.n - 1;
    out.nodes.resize(size_t(2 * t.n - 1));
    out.primitive_ids.reserve(size_t(t.n));
    for (int i = 0; i < internal; ++i) {
        CpuBvhNode& node = out.nodes[size_t(i)];
        node.bounds = t.node_bounds[size_t(i)];  // union of the two children
        node.left = child_index(t.left[size_t(i)], internal);
        node.right = child_index(t.right[size_t(i)], internal);
    }
    for (int j = 0; j < t.n; ++j) {
        const int input = int(t.sorted[size_t(j)]);
        CpuBvhNode& leaf = out.nodes[size_t(internal + j)];
        leaf.bounds = t.input_bounds[size_t(input)];
        leaf.first_primitive = j;
        leaf.primitive_count = 1;
        out.primitive_ids.push_back(t.primitive_ids[size_t(input)]);
    }
    return out;
}

bool validate_gpu_bvh_structure(const GpuBvh& bvh, std::string* error_message) {
    if (!bvh.is_valid()) return fail(error_message, "BVH is not valid");
    const HostTree t = download(bvh);
    const int n = t.n, internal = n - 1, nodes = 2 * n - 1;
    if (n != bvh.primitive_count) return fail(error_message, "primitive storage size mismatch");
    std::vector<int> seen(size_t(n), 0);
    for (int i = 0; i < n; ++i) {
        const uint32_t input = t.sorted[size_t(i)];
        if (input >= uint32_t(n)) return fail(error_message, "sorted input index out of range");
        if (++seen[input] != 1) return fail(error_message, "duplicate sorted input index");
    }
    AABB expected;
    for (const AABB& b : t.input_bounds) expected.expand(b);
    if (!boxes_close(expected, bvh.root_bounds)) return fail(error_message, "root bounds mismatch");
    if (!boxes_close(t.stackless[0].bounds, bvh.root_bounds))
        return fail(error_message, "stackless root bounds mismatch");
    int leaves = 0;
    for (const GpuBvhNode& node : t.stackless) {
        if (node.escape < -1 || node.escape >= nodes) return fail(error_message, "stackless escape index out of range");
        if (node.primitive_id >= 0) ++leaves;
    }
    if (leaves != n) return fail(error_message, "stackless leaf count mismatch");
    if (n == 1) {
        if (t.stackless[0].primitive_id != t.primitive_ids[0] || t.stackless[0].escape != -1)
            return fail(error_message, "single leaf stackless node mismatch");
        return true;
    }
    for (int i = 1; i < n; ++i)
        if (t.morton[size_t(i - 1)] > t.morton[size_t(i)]) return fail(error_message, "Morton codes are not sorted");
    auto child_of = [&](int parent, bool right) { return right ? t.right[size_t(parent)] : t.left[size_t(parent)]; };
    for (int child = 1; child < internal; ++child) {
        const uint32_t info = t.parent[size_t(child)];
        const int parent = int(info & kIndexBits);
        if (parent < 0 || parent >= internal) return fail(error_message, "internal parent index out of range");
        if (child_of(parent, (info & kLeafBit) != 0) != uint32_t(child))
            return fail(error_message, "internal parent/child mismatch");
    }
    for (int leaf = 0; leaf < n; ++leaf) {
        const uint32_t info = t.leaf_parent[size_t(leaf)];
        const int parent = int(info & kIndexBits);
        if (parent < 0 || parent >= internal) return fail(error_message, "leaf parent index out of range");
        if (child_of(parent, (info & kLeafBit) != 0) != (kLeafBit | uint32_t(leaf)))
            return fail(error_message, "leaf parent/child mismatch");
    }
    // every node reached exactly once from the root
    std::vector<int> hit_internal(size_t(internal), 0), hit_leaf(size_t(n), 0);
    std::vector<uint32_t> todo{0u};
    while (!todo.empty()) {
        const uint32_t c = todo.back();
        todo.pop_back();
        if (c & kLeafBit) {
            const uint32_t leaf = c & kIndexBits;
            if (leaf >= uint32_t(n) || ++hit_leaf[leaf] != 1)
                return fail(error_message, "invalid or duplicate leaf traversal");
            continue;
        }
        if (c >= uint32_t(internal) || ++hit_internal[c] != 1)
            return fail(error_message, "invalid or duplicate internal traversal");
        todo.push_back(t.right[c]);
        todo.push_back(t.left[c]);
    }
    if (std::find(hit_internal.begin(), hit_internal.end(), 0) != hit_internal.end())
        return fail(error_message, "unvisited internal node");
    if (std::find(hit_leaf.begin(), hit_leaf.end(), 0) != hit_leaf.end())
        return fail(error_message, "unvisited leaf node");
    return true;
}

bool validate_gpu_bvh_matches_primitives(const GpuBvh& bvh, const std::vector<BvhBuildPrimitive>& primitives,
                                         std::string* error_message) {
    if (primitives.empty()) {
        if (!bvh.is_valid() && bvh.primitive_count == 0) return true;
        return fail(error_message, "BVH contains primitives for an empty scene");
    }
    std::string structure;
    if (!validate_gpu_bvh_structure(bvh, &structure))
        return fail(error_message, structure.empty() ? "BVH structure validation failed" : structure.c_str());
    if (bvh.primitive_count != static_cast<int>(primitives.size()))
        return fail(error_message, "BVH primitive count does not match scene primitive count");
    std::unordered_map<int, const BvhBuildPrimitive*> expected;
    expected.reserve(primitives.size());
    for (const BvhBuildPrimitive& p : primitives)
        if (!expected.emplace(p.primitive_id, &p).second)
            return fail(error_message, "Scene primitive ids are not unique");
    const HostTree t = download(bvh);
    std::unordered_map<int, int> seen;
    seen.reserve(size_t(t.n));
    for (int i = 0; i < t.n; ++i) {
        const auto it = expected.find(t.primitive_ids[size_t(i)]);
        if (it == expected.end()) return fail(error_message, "BVH contains primitive id not present in scene");
        if (++seen[t.primitive_ids[size_t(i)]] != 1) return fail(error_message, "BVH contains duplicate primitive id");
        if (!boxes_close(t.input_bounds[size_t(i)], it->second->bounds))
            return fail(error_message, "BVH primitive bounds do not match scene primitive");
    }
    if (seen.size() != expected.size()) return fail(error_message, "BVH is missing one or more scene primitives");
    return true;
}

std::vector<GpuRayHit> intersect_gpu_bvh_stackless(const GpuBvh& bvh, const std::vector<Triangle>& triangles,
                                                   const std::vector<Ray>& rays) {
    std::vector<GpuRayHit> hits(rays.size());
    if (!bvh.is_valid() || rays.empty()) return hits;
    check(lux_hip_trace_closest(bvh.storage->handle, reinterpret_cast<const lux_triangle*>(triangles.data()),
                                int32_t(triangles.size()), reinterpret_cast<const lux_ray*>(rays.data()),
                                int32_t(rays.size()), reinterpret_cast<lux_hit*>(hits.data())),
          "intersect_gpu_bvh_stackless");
    return hits;
}

namespace {
OverlapQueryResult overlap(int status_of_call, int found, int max_pairs, std::vector<PrimitivePair>& pairs,
                           const char* what) {
    check(status_of_call, what);
    OverlapQueryResult r;
    r.found_count = found;
    r.truncated = found > max_pairs;
    pairs.resize(size_t(std::min(found, max_pairs)));
    r.pairs = std::move(pairs);
    return r;
}
}  // namespace

OverlapQueryResult query_aabb_overlaps_self(const GpuBvh& bvh, int max_pairs) {
    if (!bvh.is_valid() || max_pairs <= 0) return OverlapQueryResult{};
    std::vector<PrimitivePair> pairs(static_cast<size_t>(max_pairs));
    int32_t found = 0;
    const int s = lux_hip_overlap_self(bvh.storage->handle, max_pairs, reinterpret_cast<lux_pair*>(pairs.data()),
                                       &found);
    return overlap(s, found, max_pairs, pairs, "query_aabb_overlaps_self");
}

OverlapQueryResult query_aabb_overlaps_cross(const GpuBvh& left, const GpuBvh& right, int max_pairs) {
    if (!left.is_valid() || !right.is_valid() || max_pairs <= 0) return OverlapQueryResult{};
    std::vector<PrimitivePair> pairs(static_cast<size_t>(max_pairs));
    int32_t found = 0;
    const int s = lux_hip_overlap_cross(left.storage->handle, right.storage->handle, max_pairs,
                                        reinterpret_cast<lux_pair*>(pairs.data()), &found);
    return overlap(s, found, max_pairs, pairs, "query_aabb_overlaps_cross");
}

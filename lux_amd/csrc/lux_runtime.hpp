// lux_runtime.hpp -- host-side runtime pieces shared by the kernel translation
// units and the C ABI: error state, checked launches with the event probe, and
// the device-buffer set that one LBVH owns.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "lux_common.hpp"

namespace lux {

// Thrown inside the library, converted to a status code at the ABI boundary.
struct Error : std::runtime_error {
    int status;
    Error(int s, const std::string& msg) : std::runtime_error(msg), status(s) {}
};

// "HIP error after <label>: <msg>" mirrors check_cuda_or_throw
// (src/core/cuda_check.cuh:7-27).
inline void hip_check(hipError_t e, const char* label) {
    if (e != hipSuccess)
        throw Error(LUX_ERR_HIP, std::string("HIP error after ") + label + ": " + hipGetErrorString(e));
}

// ---- kernel-time probe ------------------------------------------------------
void probe_pre_launch(const char* name, hipStream_t s);
void probe_post_launch(const char* name, hipStream_t s);

hipStream_t library_stream();

#define LUX_LAUNCH(NAME, KERNEL, GRID, BLOCK, SHMEM, STREAM, ...)                        \
    do {                                                                                  \
        ::lux::probe_pre_launch(NAME, STREAM);                                            \
        hipLaunchKernelGGL(KERNEL, dim3(GRID), dim3(BLOCK), SHMEM, STREAM, __VA_ARGS__);  \
        ::lux::hip_check(hipGetLastError(), NAME);                                        \
        ::lux::probe_post_launch(NAME, STREAM);                                           \
    } while (0)

inline int div_up(int64_t a, int64_t b) { return int((a + b - 1) / b); }

// ---- LBVH device state ------------------------------------------------------
// Layout in HBM (n = primitive count, all arrays sized to `capacity`):
//   boxes[n], ids[n]              copy of the build input (input order)
//   keys[2][n], vals[2][n]        radix-sort ping-pong; after the build keys[0]
//                                 holds the sorted Morton codes, vals[0] the
//                                 sorted-position -> input-index permutation
//   left/right/parent[n-1]        Karras radix tree (bit 31 flags as in the ABI)
//   first/last[n-1]               leaf range of each internal node
//   leaf_parent[n]
//   rtop_int[n-1], rtop_leaf[n]   right steps (| depth << 16) from a frontier
//                                 node to the root
//   edge_boxes[2n]                per leaf block: union from a left-child
//                                 frontier node's first leaf to the block end
//                                 at [2 first], from the block start to a
//                                 right-child frontier node's last leaf at
//                                 [2 last + 1] (other entries unused)
//   block_boxes[n / 1024]         union of each leaf block
//   pre_internal[n-1], pre_leaf[n] node -> position in the stackless array
//   nodes[2n-1]                   32-byte stackless nodes, right-first preorder
//   sort scratch: hist[4*256], status[tiles*256] (epoch-tagged look-back words),
//                 tickets[1], partials for the root reduction
struct WideNode;

struct BvhDev {
    int n = 0;
    int capacity = 0;
    lux_aabb* boxes = nullptr;
    int32_t* ids = nullptr;
    uint32_t* keys[2] = {nullptr, nullptr};
    uint32_t* vals[2] = {nullptr, nullptr};
    uint32_t* left = nullptr;
    uint32_t* right = nullptr;
    uint32_t* parent = nullptr;
    uint32_t* first = nullptr;
    uint32_t* last = nullptr;
    uint32_t* leaf_parent = nullptr;
    uint32_t* rtop_int = nullptr;
    uint32_t* rtop_leaf = nullptr;
    lux_aabb* edge_boxes = nullptr;
    lux_aabb* block_boxes = nullptr;
    int32_t* pre_internal = nullptr;
    int32_t* pre_leaf = nullptr;
    lux_bvh_node* nodes = nullptr;
    // traversal copy of `nodes` in sibling-pair order: slot 0 = root, the
    // children of the r-th internal node (in preorder) at slots 2r + 2 (right)
    // and 2r + 3 (left); an internal node's primitive_id holds ~(right child
    // slot), escapes are slots. pair_pos maps preorder position -> slot,
    // slot_pre slot -> preorder position (0xffffffff for the unused slot 1).
    lux_bvh_node* pairs = nullptr;
    uint32_t* pair_pos = nullptr;
    uint32_t* slot_pre = nullptr;
    uint32_t* hist = nullptr;
    uint64_t* status = nullptr;
    uint32_t* tickets = nullptr;
    uint8_t* partials = nullptr;
    DevInfo* info = nullptr;
    // counter[0]: overlap pair counter of host-array queries
    int32_t* counter = nullptr;
    uint32_t epoch = 0;
    int status_words = 0;
};

void bvh_allocate(BvhDev& d, int n);
void bvh_release(BvhDev& d);

// Scratch of the ray-order sort of the traversal entry points (grown on
// demand; one per tree handle, used on one stream at a time).
struct RaySortDev {
    int capacity = 0;
    uint32_t* keys[2] = {nullptr, nullptr};
    uint32_t* vals[2] = {nullptr, nullptr};
    uint32_t* hist = nullptr;    // 2 x 256 digit counts
    uint64_t* status = nullptr;  // epoch-tagged look-back words
    uint32_t* tickets = nullptr;
    uint32_t epoch = 0;
};
void raysort_release(RaySortDev& r);
// Enqueues the traversal order of `rays` and returns the device array of the n
// ray indices sorted by direction octant, then by the top 13 bits of the
// 30-bit Morton code of each origin in the tree's root box (2 radix passes).
// Tracing in this order keeps the rays in flight on one XCD close in origin
// and heading, so their nodes stay in L2.
const uint32_t* ray_order(RaySortDev& r, const BvhDev& tree, const lux_ray* rays, int n, hipStream_t s);
// The same sort on caller-provided buffers: keys/vals n words each, hist 512
// words, status ray_order_status_words(n) 64-bit words, tickets 4 words. The
// look-back words of pass p carry epoch first_epoch + p; status must hold no
// word of those epochs from an earlier sort (zeroed words never do) and
// tickets must be zero. Enqueues no allocation.
struct RayOrderBufs {
    uint32_t* keys[2];
    uint32_t* vals[2];
    uint32_t* hist;
    uint64_t* status;
    uint32_t* tickets;
};
size_t ray_order_status_words(int n);
const uint32_t* ray_order_into(const RayOrderBufs& b, uint32_t first_epoch, const BvhDev& tree, const lux_ray* rays,
                               int n, hipStream_t s);
// Fills count 32-bit words with zero (a kernel: safe inside captured graphs).
void zero_words(uint32_t* p, size_t count, hipStream_t s);
// Below this many rays the traversal entry points trace in input order.
constexpr int kOrderMinRays = 1 << 16;

// Enqueue a full build of n primitives; in_boxes/in_ids are device arrays
// (may alias d.boxes/d.ids).
void lbvh_build(BvhDev& d, const lux_aabb* in_boxes, const int32_t* in_ids, int n, hipStream_t s);
// Enqueue a full build whose input is computed on the device from n triangles:
// the soup `tris`, or (tris null) the indexed mesh verts[nv] / idx[3n]; box i =
// triangle_bounds of triangle i, id i.
void lbvh_build_triangles(BvhDev& d, const lux_triangle* tris, const lux_vec3* verts, const uint32_t* idx, int nv,
                          int n, hipStream_t s);

// Film output (film.hip): write_ppm's pixel bytes (3 per pixel, film order),
// write_pfm's float body (rows bottom to top), both from a device film of
// per-pixel sums; whole-file writers; the 255 gamma thresholds of the byte
// conversion (host powf).
void film_ppm(const float* film, int width, int height, int spp, uint8_t* out, hipStream_t s);
void film_pfm(const float* film, int width, int height, int spp, float* out, hipStream_t s);
void write_film_file(const char* path, bool ppm, const float* d_film, int width, int height, int spp, hipStream_t s);
void ppm_thresholds(float* out255);

// ---- OBJ import (obj_import.hip) ----
namespace obj {
// The mesh load_obj_into_scene appends (object space) and its instance transform.
struct Mesh {
    std::vector<float> vertices, normals, uvs;  // 3, 3, 2 per vertex
    std::vector<uint32_t> indices;              // 3 per triangle
    std::vector<int32_t> material_ids;
    std::vector<uint8_t> boundary_edges;        // bit k: edge (k, k+1 mod 3) is on the boundary
    float object_to_world[16], world_to_object[16];
    int appended_default_material = 0;
    int skipped_invalid = 0, skipped_nonfinite = 0, skipped_degenerate = 0, nonfinite_vertices = 0;
    int invalid_normal_refs = 0, nonfinite_normal_refs = 0, zero_normal_refs = 0;
    // boundary-hole capping (ObjLoadResult's cap_* counts)
    int cap_triangles = 0, cap_boundary_edges = 0, cap_non_simple_vertices = 0, cap_incomplete_loops = 0,
        cap_degenerate_loops = 0, cap_degenerate_triangles = 0;
};
void load(const std::string& path, const lux_obj_options& options, Mesh& out);
void default_options(lux_obj_options& options);
}  // namespace obj
// Enqueue a refit with new boxes (input order) or triangles (input order).
void lbvh_refit_boxes(BvhDev& d, const lux_aabb* in_boxes, hipStream_t s);
void lbvh_refit_triangles(BvhDev& d, const lux_triangle* tris, hipStream_t s);

// Queries (all enqueue only). Rays may be traced in origin order (results
// still land at each ray's own index).
//
// The order to trace n rays in: sorted when `sort` is set and n >=
// kOrderMinRays (r holds the sort's buffers), else null (input order).
const uint32_t* closest_order(RaySortDev& r, bool sort, const BvhDev& d, const lux_ray* rays, int n,
                              hipStream_t s);
void trace_closest(const BvhDev& d, const uint32_t* order, const lux_triangle* tris, int tri_count,
                   const lux_ray* rays, int n, lux_hit* hits, hipStream_t s);
void trace_any(const BvhDev& d, RaySortDev* order, const lux_triangle* tris, int tri_count, const lux_ray* rays,
               int n, uint8_t* occluded, hipStream_t s);
void trace_stats(const BvhDev& d, const lux_triangle* tris, int tri_count, const lux_ray* rays, int n,
                 unsigned long long* d_stats, hipStream_t s);
void copy_device(void* dst, const void* src, size_t bytes, hipStream_t s);
// Closest hit with the ray-order scratch in a fixed per-device array
// (noalloc.hip): no allocation. noalloc_setup takes the array's device address
// on `device` (current) once; trace_closest_noalloc rejects n > kNoAllocMaxRays.
constexpr int kNoAllocMaxRays = 1 << 22;
void noalloc_setup(int device);
size_t noalloc_scratch_bytes();
void trace_closest_noalloc(int device, const BvhDev& d, bool order_rays, const lux_triangle* tris, int tri_count,
                           const lux_ray* rays, int n, lux_hit* hits, hipStream_t s);
void overlap_query(const BvhDev& tree, const BvhDev& queries, bool self_mode, int max_pairs,
                   lux_pair* pairs, int32_t* counter, hipStream_t s);

}  // namespace lux

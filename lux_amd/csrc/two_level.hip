// two_level.hip -- builds two-level (instanced) scenes and launches their walk
// (two_level.hpp) over ray arrays.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "two_level.hpp"

namespace lux {
namespace {

// 128 threads per block: closest 20.2 -> 19.3 ms and occlusion 16.1 -> 15.7 ms
// against 256 on 4.19 M rays over 64 instances.
constexpr int kBlock = 128;

__device__ __forceinline__ lux_scene_hit scene_hit(const TwoLevelHit& h) {
    lux_scene_hit out;
    out.primitive_id = h.prim;
    out.t = h.prim >= 0 ? h.t : INFINITY;
    out.u = h.prim >= 0 ? h.u : 0.0f;
    out.v = h.prim >= 0 ? h.v : 0.0f;
    out.instance_id = h.instance;
    out.local_primitive_id = h.local;
    return out;
}

// Closest hit near child first at both levels; a ray the walk cannot settle
// (stack overflow or the check of ordered.hpp) takes the preorder walk.
// `order`: the rays in ray order (a permutation of [0, n)), or null for input
// order; a hit lands at its ray's own index.
__global__ __launch_bounds__(kBlock) void k_two_level_closest(TwoLevelView s, const uint32_t* __restrict__ order,
                                                              const lux_ray* __restrict__ rays,
                                                              const int32_t* __restrict__ ignore, int n,
                                                              lux_scene_hit* __restrict__ hits) {
    const int j = int(xcd_block()) * kBlock + threadIdx.x;
    if (j >= n) return;
    __shared__ uint32_t stack[kOrderedStack * kBlock];
    const int i = order ? int(order[j]) : j;
    if (uint32_t(i) >= uint32_t(n)) {  // only if the sort itself reported an error
        set_error(s.info, kErrSpinLimit);
        return;
    }
    const lux_ray r = rays[i];
    const int ign = ignore ? ignore[i] : -1;
    TwoLevelHit h;
    if (!two_level_closest_ordered(s, ld3(r.origin), ld3(r.direction), r.t_min, r.t_max, ign, stack + threadIdx.x,
                                   kBlock, h))
        h = two_level_walk<false>(s, ld3(r.origin), ld3(r.direction), r.t_min, r.t_max, ign);
    hits[i] = scene_hit(h);
}

__global__ __launch_bounds__(kBlock) void k_two_level_occluded(TwoLevelView s, const uint32_t* __restrict__ order,
                                                               const lux_ray* __restrict__ rays,
                                                               const int32_t* __restrict__ ignore, int n,
                                                               uint8_t* __restrict__ occluded) {
    const int j = int(xcd_block()) * kBlock + threadIdx.x;
    if (j >= n) return;
    const int i = order ? int(order[j]) : j;
    if (uint32_t(i) >= uint32_t(n)) {
        set_error(s.info, kErrSpinLimit);
        return;
    }
    const lux_ray r = rays[i];
    occluded[i] = two_level_any(s, ld3(r.origin), ld3(r.direction), r.t_min, r.t_max, ignore ? ignore[i] : -1) ? 1 : 0;
}

template <typename T>
T* dalloc_n(size_t count) {
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T));
    if (e != hipSuccess) throw Error(LUX_ERR_OUT_OF_MEMORY, std::string("hipMalloc failed: ") + hipGetErrorString(e));
    return static_cast<T*>(p);
}

template <typename T>
struct TempBuffer {
    T* p = nullptr;
    explicit TempBuffer(size_t count) : p(dalloc_n<T>(count)) {}
    ~TempBuffer() {
        if (p) (void)hipFree(p);
    }
    TempBuffer(const TempBuffer&) = delete;
    TempBuffer& operator=(const TempBuffer&) = delete;
};

bool box_is_empty(const lux_aabb& b) {
    return b.p_min.x > b.p_max.x || b.p_min.y > b.p_max.y || b.p_min.z > b.p_max.z;
}

// det3 and the affine test of src/core/mat4.cuh:86-104, host side.
float affine_det3(const float* m) {
    return m[0] * (m[5] * m[10] - m[6] * m[9]) - m[1] * (m[4] * m[10] - m[6] * m[8]) +
           m[2] * (m[4] * m[9] - m[5] * m[8]);
}
bool affine_invertible(const float* m) {
    return std::fabs(m[12]) <= 1e-6f && std::fabs(m[13]) <= 1e-6f && std::fabs(m[14]) <= 1e-6f &&
           std::fabs(m[15] - 1.0f) <= 1e-6f && std::fabs(affine_det3(m)) > 1e-12f;
}
bool product_near_identity(const float* a, const float* b) {
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
            float s = 0.0f;
            for (int k = 0; k < 4; ++k) s += a[4 * i + k] * b[4 * k + j];
            if (!(std::fabs(s - (i == j ? 1.0f : 0.0f)) <= 1e-4f)) return false;
        }
    return true;
}

void validate_instances(const lux_instance* instances, int count, int mesh_count) {
    if (count < 0) throw Error(LUX_ERR_INVALID_ARGUMENT, "negative instance count");
    if (count > 0 && !instances) throw Error(LUX_ERR_INVALID_ARGUMENT, "null instance array");
    for (int i = 0; i < count; ++i) {
        const lux_instance& in = instances[i];
        if (in.mesh_id < 0 || in.mesh_id >= mesh_count)
            throw Error(LUX_ERR_INVALID_ARGUMENT, "instance " + std::to_string(i) + ": mesh id out of range");
        if (!instance_transform_valid(in.object_to_world, in.world_to_object))
            throw Error(LUX_ERR_INVALID_ARGUMENT,
                        "instance " + std::to_string(i) + ": transform must be affine-invertible with a consistent inverse");
    }
}

void sync_and_check(const BvhDev& d, hipStream_t stream, const char* what) {
    hip_check(hipStreamSynchronize(stream), what);
    if (!d.info) return;
    DevInfo info;
    hip_check(hipMemcpy(&info, d.info, sizeof(DevInfo), hipMemcpyDeviceToHost), "read device info");
    if (info.error != 0) throw Error(LUX_ERR_DEVICE_CHECK, std::string("device check failed in ") + what);
}

// The culling copy of the TLAS (TwoLevelDev::tlas_cull): the same slots, each
// leaf box the instance's world bounds intersected with the world box of its
// mesh's bounding sphere (both widened as below), each internal box the union
// of its two children. Every box still holds every world-space ray point that
// can hit the geometry below it, so only the culling changes.
void build_tlas_cull(TwoLevelDev& s, const lux_instance* instances, hipStream_t stream) {
    const int n = s.tlas.n;
    const int slots = n > 1 ? 2 * n : n;
    if (s.tlas_cull) hip_check(hipFree(s.tlas_cull), "hipFree culling TLAS");
    s.tlas_cull = nullptr;
    if (slots <= 0) return;
    std::vector<lux_bvh_node> nodes(static_cast<size_t>(slots));
    hip_check(hipMemcpyAsync(nodes.data(), s.tlas.pairs, sizeof(lux_bvh_node) * slots, hipMemcpyDeviceToHost, stream),
              "D2H TLAS");
    hip_check(hipStreamSynchronize(stream), "synchronize TLAS download");
    // A leaf box must hold every world point w(t) of a ray whose object-space
    // image p(t) = inv w(t) lies in the mesh: M p(t) = (M inv) w(t) lies in the
    // transformed mesh, and M inv is the identity only to the scene
    // validation's 1e-4 per entry, so w(t) may sit up to 3 e |w| + e_t off it
    // (e, e_t: the largest deviation of M inv from the identity in the 3x3 and
    // translation parts). Both boxes are widened by that, plus a margin for
    // float rounding, before they are intersected.
    auto leaf_box = [&](int inst) {
        const lux_instance& in = instances[inst];
        const float* m = in.object_to_world;
        const float* v = in.world_to_object;
        double e3 = 0, et = 0;
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 4; ++c) {
                double x = c == 3 ? double(m[4 * r + 3]) : 0.0;
                for (int k = 0; k < 3; ++k) x += double(m[4 * r + k]) * v[4 * k + c];
                if (c == 3) {
                    et = std::max(et, std::fabs(x));
                } else {
                    e3 = std::max(e3, std::fabs(x - (r == c ? 1.0 : 0.0)));
                }
            }
        const lux_aabb& ref = in.world_bounds;
        const double R = std::max({std::fabs(double(ref.p_min.x)), std::fabs(double(ref.p_min.y)),
                                   std::fabs(double(ref.p_min.z)), std::fabs(double(ref.p_max.x)),
                                   std::fabs(double(ref.p_max.y)), std::fabs(double(ref.p_max.z))});
        const double pad = 3.0 * e3 * R + et + 1e-5 * (R + 1e-30);
        double lo[3] = {ref.p_min.x - pad, ref.p_min.y - pad, ref.p_min.z - pad};
        double hi[3] = {ref.p_max.x + pad, ref.p_max.y + pad, ref.p_max.z + pad};
        const float* sph = &s.mesh_sphere[4 * size_t(in.mesh_id)];
        if (sph[3] >= 0) {
            for (int r = 0; r < 3; ++r) {
                const double c = double(m[4 * r]) * sph[0] + double(m[4 * r + 1]) * sph[1] +
                                 double(m[4 * r + 2]) * sph[2] + double(m[4 * r + 3]);
                const double h = double(sph[3]) * std::sqrt(double(m[4 * r]) * m[4 * r] +
                                                            double(m[4 * r + 1]) * m[4 * r + 1] +
                                                            double(m[4 * r + 2]) * m[4 * r + 2]);
                lo[r] = std::max(lo[r], c - h - pad);
                hi[r] = std::min(hi[r], c + h + pad);
            }
        }
        lux_aabb b;
        b.p_min = lux_vec3{float(lo[0]), float(lo[1]), float(lo[2])};
        b.p_max = lux_vec3{float(hi[0]), float(hi[1]), float(hi[2])};
        // rounding to float must not move a face inward
        b.p_min = lux_vec3{std::nextafter(b.p_min.x, -INFINITY), std::nextafter(b.p_min.y, -INFINITY),
                           std::nextafter(b.p_min.z, -INFINITY)};
        b.p_max = lux_vec3{std::nextafter(b.p_max.x, INFINITY), std::nextafter(b.p_max.y, INFINITY),
                           std::nextafter(b.p_max.z, INFINITY)};
        return b;
    };
    // post-order over the pair slots from the root: a leaf takes its
    // instance's box, an internal node the union of its children's
    std::vector<int> todo{0};
    std::vector<uint8_t> expanded(static_cast<size_t>(slots), 0);
    while (!todo.empty()) {
        const int k = todo.back();
        lux_bvh_node& nd = nodes[size_t(k)];
        if (nd.primitive_id >= 0) {
            if (nd.primitive_id < s.instance_count) nd.bounds = leaf_box(nd.primitive_id);
            todo.pop_back();
            continue;
        }
        const int c = ~nd.primitive_id;
        if (c < 0 || c + 1 >= slots) throw Error(LUX_ERR_DEVICE_CHECK, "malformed TLAS slot array");
        if (!expanded[size_t(k)]) {
            expanded[size_t(k)] = 1;
            todo.push_back(c);
            todo.push_back(c + 1);
            continue;
        }
        const lux_aabb& l = nodes[size_t(c)].bounds;
        const lux_aabb& r = nodes[size_t(c) + 1].bounds;
        nd.bounds.p_min = lux_vec3{std::min(l.p_min.x, r.p_min.x), std::min(l.p_min.y, r.p_min.y),
                                   std::min(l.p_min.z, r.p_min.z)};
        nd.bounds.p_max = lux_vec3{std::max(l.p_max.x, r.p_max.x), std::max(l.p_max.y, r.p_max.y),
                                   std::max(l.p_max.z, r.p_max.z)};
        todo.pop_back();
    }
    s.tlas_cull = dalloc_n<lux_bvh_node>(size_t(slots));
    hip_check(hipMemcpyAsync(s.tlas_cull, nodes.data(), sizeof(lux_bvh_node) * slots, hipMemcpyHostToDevice, stream),
              "H2D culling TLAS");
    hip_check(hipStreamSynchronize(stream), "synchronize culling TLAS upload");
}

// TLAS over the non-empty world bounds (ids = instance ids) and the instance
// records; first ordinals follow the instances in order (each adds its mesh's
// triangle count).
void upload_instances(TwoLevelDev& s, const lux_instance* instances, int count, hipStream_t stream) {
    std::vector<DevInstance> rec(static_cast<size_t>(std::max(count, 0)));
    std::vector<lux_aabb> boxes;
    std::vector<int32_t> ids;
    int64_t cursor = 0;
    for (int i = 0; i < count; ++i) {
        const lux_instance& in = instances[i];
        DevInstance& r = rec[i];
        std::memcpy(r.inv, in.world_to_object, sizeof(r.inv));
        r.mesh = in.mesh_id;
        const int64_t tri = s.mesh_tri_off[in.mesh_id + 1] - s.mesh_tri_off[in.mesh_id];
        if (cursor + tri > INT32_MAX) throw Error(LUX_ERR_INVALID_ARGUMENT, "more than 2^31 scene primitives");
        r.first = int32_t(cursor);
        r.nodes = s.blas[in.mesh_id].pairs;
        r.node_count = tri > 1 ? int32_t(2 * tri) : int32_t(tri);
        r.tri_base = s.mesh_tri_off[in.mesh_id];
        r.tri_count = int32_t(tri);
        r.pad = 0;
        cursor += tri;
        if (box_is_empty(in.world_bounds)) continue;
        boxes.push_back(in.world_bounds);
        ids.push_back(i);
    }
    s.host_instances.assign(instances, instances + count);
    s.instance_count = count;
    s.scene_primitives = cursor;
    s.tlas_instances = int(boxes.size());
    if (s.instances) hip_check(hipFree(s.instances), "hipFree instances");
    s.instances = dalloc_n<DevInstance>(rec.size());
    if (!rec.empty())
        hip_check(hipMemcpyAsync(s.instances, rec.data(), sizeof(DevInstance) * rec.size(), hipMemcpyHostToDevice,
                                 stream),
                  "H2D instances");
    const int n = int(boxes.size());
    TempBuffer<lux_aabb> d_boxes(size_t(std::max(n, 1)));
    TempBuffer<int32_t> d_ids(size_t(std::max(n, 1)));
    if (n > 0) {
        hip_check(hipMemcpyAsync(d_boxes.p, boxes.data(), sizeof(lux_aabb) * n, hipMemcpyHostToDevice, stream),
                  "H2D world bounds");
        hip_check(hipMemcpyAsync(d_ids.p, ids.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, stream),
                  "H2D instance ids");
    }
    lbvh_build(s.tlas, d_boxes.p, d_ids.p, n, stream);
    sync_and_check(s.tlas, stream, "TLAS build");
    uint64_t limit = n > 1 ? 2u * uint64_t(n) : uint64_t(n);
    for (int i = 0; i < count; ++i) {
        const int tri = s.mesh_tri_off[instances[i].mesh_id + 1] - s.mesh_tri_off[instances[i].mesh_id];
        limit += tri > 1 ? 2u * uint64_t(tri) : uint64_t(tri);
    }
    s.step_limit = uint32_t(std::min<uint64_t>(limit + 1, 0xffffffffu));
    build_tlas_cull(s, instances, stream);
}

size_t bvh_device_bytes(const BvhDev& d) {
    const size_t cap = size_t(d.capacity);
    // boxes, ids, keys/vals, tree links, slots, preorder maps, nodes, pairs
    // and their two slot maps
    return cap * (24 + 4 + 16 + 28 + 48 + 8 + 64 + 64 + 16);
}

}  // namespace

bool instance_transform_valid(const float* m16, const float* inv16) {
    return affine_invertible(m16) && affine_invertible(inv16) && product_near_identity(m16, inv16) &&
           product_near_identity(inv16, m16);
}

void inverse_affine(const float* m, float* r) {
    for (int i = 0; i < 16; ++i) r[i] = (i % 5 == 0) ? 1.0f : 0.0f;
    const float a00 = m[0], a01 = m[1], a02 = m[2];
    const float a10 = m[4], a11 = m[5], a12 = m[6];
    const float a20 = m[8], a21 = m[9], a22 = m[10];
    const float det = affine_det3(m);
    if (std::fabs(det) <= 1e-12f) return;
    const float inv_det = 1.0f / det;
    r[0] = (a11 * a22 - a12 * a21) * inv_det;
    r[1] = -(a01 * a22 - a02 * a21) * inv_det;
    r[2] = (a01 * a12 - a02 * a11) * inv_det;
    r[4] = -(a10 * a22 - a12 * a20) * inv_det;
    r[5] = (a00 * a22 - a02 * a20) * inv_det;
    r[6] = -(a00 * a12 - a02 * a10) * inv_det;
    r[8] = (a10 * a21 - a11 * a20) * inv_det;
    r[9] = -(a00 * a21 - a01 * a20) * inv_det;
    r[10] = (a00 * a11 - a01 * a10) * inv_det;
    const float tx = m[3], ty = m[7], tz = m[11];
    r[3] = -(r[0] * tx + r[1] * ty + r[2] * tz);
    r[7] = -(r[4] * tx + r[5] * ty + r[6] * tz);
    r[11] = -(r[8] * tx + r[9] * ty + r[10] * tz);
    r[12] = r[13] = r[14] = 0.0f;
    r[15] = 1.0f;
}

lux_aabb transform_bounds(const float* m, const lux_aabb& b) {
    lux_aabb r;
    r.p_min = lux_vec3{INFINITY, INFINITY, INFINITY};
    r.p_max = lux_vec3{-INFINITY, -INFINITY, -INFINITY};
    if (box_is_empty(b)) return r;
    for (int z = 0; z <= 1; ++z)
        for (int y = 0; y <= 1; ++y)
            for (int x = 0; x <= 1; ++x) {
                const float cx = x ? b.p_max.x : b.p_min.x, cy = y ? b.p_max.y : b.p_min.y,
                            cz = z ? b.p_max.z : b.p_min.z;
                float px = m[0] * cx + m[1] * cy + m[2] * cz + m[3];
                float py = m[4] * cx + m[5] * cy + m[6] * cz + m[7];
                float pz = m[8] * cx + m[9] * cy + m[10] * cz + m[11];
                const float w = m[12] * cx + m[13] * cy + m[14] * cz + m[15];
                if (w != 1.0f) {
                    px = px / w;
                    py = py / w;
                    pz = pz / w;
                }
                r.p_min = lux_vec3{std::fmin(r.p_min.x, px), std::fmin(r.p_min.y, py), std::fmin(r.p_min.z, pz)};
                r.p_max = lux_vec3{std::fmax(r.p_max.x, px), std::fmax(r.p_max.y, py), std::fmax(r.p_max.z, pz)};
            }
    return r;
}

void two_level_release(TwoLevelDev& s) {
    for (BvhDev& b : s.blas) bvh_release(b);
    bvh_release(s.tlas);
    void* ptrs[] = {s.soup, s.meshes, s.instances, s.tlas_cull};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    s = TwoLevelDev{};
}

void two_level_build(TwoLevelDev& s, const lux_triangle* tris, const int32_t* offsets, int mesh_count,
                     const lux_instance* instances, int instance_count, hipStream_t stream) {
    if (mesh_count < 0) throw Error(LUX_ERR_INVALID_ARGUMENT, "negative mesh count");
    if (mesh_count > 0 && !offsets) throw Error(LUX_ERR_INVALID_ARGUMENT, "null mesh offsets");
    if (mesh_count > 0 && offsets[0] != 0) throw Error(LUX_ERR_INVALID_ARGUMENT, "mesh_offsets[0] must be 0");
    for (int m = 0; m < mesh_count; ++m)
        if (offsets[m + 1] < offsets[m]) throw Error(LUX_ERR_INVALID_ARGUMENT, "mesh offsets must not decrease");
    const int total = mesh_count > 0 ? offsets[mesh_count] : 0;
    if (total > 0 && !tris) throw Error(LUX_ERR_INVALID_ARGUMENT, "null triangle array");
    validate_instances(instances, instance_count, mesh_count);
    two_level_release(s);
    s.mesh_count = mesh_count;
    s.mesh_tri_off.assign(size_t(mesh_count) + 1, 0);
    for (int m = 0; m <= mesh_count; ++m) s.mesh_tri_off[m] = mesh_count > 0 ? offsets[m] : 0;
    s.soup = dalloc_n<lux_triangle>(size_t(std::max(total, 1)));
    if (total > 0)
        hip_check(hipMemcpyAsync(s.soup, tris, sizeof(lux_triangle) * total, hipMemcpyHostToDevice, stream),
                  "H2D mesh triangles");
    s.blas.resize(size_t(mesh_count));
    std::vector<DevMesh> meshes(static_cast<size_t>(mesh_count));
    for (int m = 0; m < mesh_count; ++m) {
        const int n = offsets[m + 1] - offsets[m];
        // BLAS input (triangle_bounds, local ids) computed on the device from
        // the soup copied above
        lbvh_build_triangles(s.blas[m], s.soup + offsets[m], nullptr, nullptr, 0, n, stream);
        sync_and_check(s.blas[m], stream, "BLAS build");
        meshes[m].nodes = s.blas[m].pairs;
        meshes[m].slot_pre = s.blas[m].slot_pre;
        meshes[m].tris = s.soup + offsets[m];
        meshes[m].node_count = n > 1 ? 2 * n : n;
        meshes[m].tri_count = n;
        s.device_bytes += bvh_device_bytes(s.blas[m]);
    }
    // each mesh's bounding sphere: the center of its box, the largest vertex
    // distance from it (in double, padded), radius -1 for an empty mesh
    s.mesh_sphere.assign(4 * size_t(mesh_count), 0.0f);
    for (int m = 0; m < mesh_count; ++m) {
        float* sph = &s.mesh_sphere[4 * size_t(m)];
        if (offsets[m + 1] == offsets[m]) {
            sph[3] = -1.0f;
            continue;
        }
        float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (int t = offsets[m]; t < offsets[m + 1]; ++t)
            for (const lux_vec3& v : {tris[t].v0, tris[t].v1, tris[t].v2}) {
                const float q[3] = {v.x, v.y, v.z};
                for (int a = 0; a < 3; ++a) {
                    lo[a] = std::min(lo[a], q[a]);
                    hi[a] = std::max(hi[a], q[a]);
                }
            }
        double r2 = 0;
        for (int a = 0; a < 3; ++a) sph[a] = 0.5f * (lo[a] + hi[a]);
        for (int t = offsets[m]; t < offsets[m + 1]; ++t)
            for (const lux_vec3& v : {tris[t].v0, tris[t].v1, tris[t].v2}) {
                const double dx = double(v.x) - sph[0], dy = double(v.y) - sph[1], dz = double(v.z) - sph[2];
                r2 = std::max(r2, dx * dx + dy * dy + dz * dz);
            }
        sph[3] = float(std::sqrt(r2) * (1.0 + 1e-6)) + 1e-30f;
    }
    s.meshes = dalloc_n<DevMesh>(meshes.size());
    if (!meshes.empty())
        hip_check(hipMemcpyAsync(s.meshes, meshes.data(), sizeof(DevMesh) * meshes.size(), hipMemcpyHostToDevice,
                                 stream),
                  "H2D meshes");
    upload_instances(s, instances, instance_count, stream);
    s.device_bytes += sizeof(lux_triangle) * size_t(total) + sizeof(DevMesh) * meshes.size() +
                      sizeof(DevInstance) * size_t(instance_count) + bvh_device_bytes(s.tlas) +
                      sizeof(lux_bvh_node) * 2 * size_t(std::max(instance_count, 1));
}

void two_level_set_instances(TwoLevelDev& s, const lux_instance* instances, int instance_count,
                             hipStream_t stream) {
    if (instance_count != s.instance_count)
        throw Error(LUX_ERR_MISMATCH, "instance count differs from the built scene");
    validate_instances(instances, instance_count, s.mesh_count);
    upload_instances(s, instances, instance_count, stream);
}

TwoLevelView two_level_view(const TwoLevelDev& s) {
    TwoLevelView v;
    v.tlas_nodes = s.tlas.pairs;
    v.tlas_cull = s.tlas_cull ? s.tlas_cull : s.tlas.pairs;
    v.tlas_slot_pre = s.tlas.slot_pre;
    v.instances = s.instances;
    v.meshes = s.meshes;
    v.soup = s.soup;
    v.info = s.tlas.info;
    v.tlas_node_count = s.tlas.n > 1 ? 2 * s.tlas.n : s.tlas.n;
    v.instance_count = s.instance_count;
    v.mesh_count = s.mesh_count;
    v.step_limit = s.step_limit;
    return v;
}

const uint32_t* two_level_order(RaySortDev& r, const TwoLevelDev& s, const lux_ray* rays, int n, hipStream_t stream) {
    if (n < kOrderMinRays || s.tlas.n < 1) return nullptr;
    return ray_order(r, s.tlas, rays, n, stream);  // origins quantised in the TLAS root box
}

void two_level_closest(const TwoLevelView& v, const uint32_t* order, const lux_ray* rays, const int32_t* ignore,
                       int n, lux_scene_hit* hits, hipStream_t stream) {
    if (n <= 0) return;
    LUX_LAUNCH("two_level_closest", k_two_level_closest, div_up(n, kBlock), kBlock, 0, stream, v, order, rays,
               ignore, n, hits);
}

void two_level_occluded(const TwoLevelView& v, const uint32_t* order, const lux_ray* rays, const int32_t* ignore,
                        int n, uint8_t* occluded, hipStream_t stream) {
    if (n <= 0) return;
    LUX_LAUNCH("two_level_occluded", k_two_level_occluded, div_up(n, kBlock), kBlock, 0, stream, v, order, rays,
               ignore, n, occluded);
}

}  // namespace lux

// two_level.hpp -- two-level (instanced) scenes: one LBVH per mesh over its
// object-space triangles (BLAS) and one over the instances' world bounds
// (TLAS), the layout of the reference's PreparedScene
// (src/scene/prepared_scene.cu:59-78, src/scene/scene_prepare.h:420-462).
#pragma once

#include <vector>

#include "lux_runtime.hpp"

namespace lux {

// One mesh as the walk sees it: its BLAS (sibling-pair node copy) and its
// triangles (object space, a slice of the scene's soup).
struct DevMesh {
    const lux_bvh_node* nodes;
    const lux_triangle* tris;
    int32_t node_count;  // sibling-pair slots (0: empty mesh)
    int32_t tri_count;
};

// One instance: Transform::inv (row-major 4x4), its mesh and the scene ordinal
// of its local triangle 0 (GpuSceneInstance::first_primitive, src/scene/gpu_scene.h:149-172).
struct DevInstance {
    float inv[16];
    int32_t mesh;
    int32_t first;
    int32_t pad[2];
};
static_assert(sizeof(DevInstance) == 80, "instance record");

// What the walk kernels take by value.
struct TwoLevelView {
    const lux_bvh_node* tlas_nodes;
    const DevInstance* instances;
    const DevMesh* meshes;
    DevInfo* info;
    int32_t tlas_node_count, instance_count, mesh_count;
    uint32_t step_limit;  // TLAS slots + the BLAS slots of every instance: no walk takes more steps
};

struct TwoLevelDev {
    std::vector<BvhDev> blas;  // per mesh (n = 0 for an empty mesh)
    BvhDev tlas;               // over the non-empty instance world bounds, ids = instance ids
    lux_triangle* soup = nullptr;
    DevMesh* meshes = nullptr;
    DevInstance* instances = nullptr;
    std::vector<int32_t> mesh_tri_off;  // mesh_count + 1
    std::vector<lux_instance> host_instances;
    int mesh_count = 0, instance_count = 0, tlas_instances = 0;
    int64_t scene_primitives = 0;
    uint32_t step_limit = 0;
    size_t device_bytes = 0;
};

// Scene validation of one instance transform pair (src/scene/scene_validate.h:140-147,
// src/core/mat4.cuh:86-124): both matrices affine with |det3| > 1e-12 and
// m * inv, inv * m within 1e-4 of the identity.
bool instance_transform_valid(const float* m16, const float* inv16);
// make_transform's inverse and transform_bounds (src/core/mat4.cuh:126-159, 254-270).
void inverse_affine(const float* m16, float* inv16);
lux_aabb transform_bounds(const float* m16, const lux_aabb& b);

// Builds every BLAS and the TLAS (synchronous: validates, uploads, builds, waits).
void two_level_build(TwoLevelDev& s, const lux_triangle* tris, const int32_t* mesh_offsets, int mesh_count,
                     const lux_instance* instances, int instance_count, hipStream_t stream);
// New instance transforms / world bounds for the same meshes and instance count:
// rebuilds the TLAS and the instance records only.
void two_level_set_instances(TwoLevelDev& s, const lux_instance* instances, int instance_count,
                             hipStream_t stream);
void two_level_release(TwoLevelDev& s);
TwoLevelView two_level_view(const TwoLevelDev& s);

// Closest hit (or occlusion) of n device rays; ignore (nullable) holds one scene
// ordinal per ray that is skipped (-1: none).
void two_level_closest(const TwoLevelView& v, const lux_ray* rays, const int32_t* ignore, int n,
                       lux_scene_hit* hits, hipStream_t stream);
void two_level_occluded(const TwoLevelView& v, const lux_ray* rays, const int32_t* ignore, int n,
                        uint8_t* occluded, hipStream_t stream);

}  // namespace lux

// pt_runtime.hpp -- host-side state of the wavefront path tracer: the device
// scene (triangles, shading normals, materials, power-alias light table, camera,
// BLAS) and the per-batch path buffers.
#pragma once

#include <vector>

#include "lux_runtime.hpp"
#include "pt_device.hpp"

namespace lux {
namespace pt {

// Scene in HBM. Layout (T triangles, V vertices, M materials, K lights):
//   soup[T]          36-byte triangle records (vertex positions by triangle
//                    index; the traversal reads only these and the nodes)
//   indices[3T], normals[3V]  shading normals (after the identity-instance
//                    transform), read once per hit
//   mat_id[T], mats[M]        48-byte material records
//   lights[K], light_of_prim[T]  power-alias light table, emitter lookup
//   bvh              the LBVH over the triangle boxes (ids = triangle index)
struct SceneDev {
    BvhDev bvh;
    lux_triangle* soup = nullptr;
    uint32_t* indices = nullptr;
    float* normals = nullptr;
    int32_t* mat_id = nullptr;
    DevMaterial* mats = nullptr;
    int32_t* light_of_prim = nullptr;
    DevLight* lights = nullptr;
    int tri_count = 0, vertex_count = 0, mat_count = 0, light_count = 0;
    DevCamera camera{};
    std::vector<DevLight> host_lights;
    std::vector<int32_t> host_light_of_prim;
};

// POD view handed to kernels by value.
struct SceneView {
    const lux_bvh_node* nodes;
    int node_count;
    const lux_triangle* soup;
    int tri_count;
    const uint32_t* indices;
    const float* normals;
    const int32_t* mat_id;
    const DevMaterial* mats;
    int mat_count;
    const DevLight* lights;
    int light_count;
    const int32_t* light_of_prim;
    DevCamera cam;
    DevInfo* info;
};

// Per-batch path state: 11 float4 planes (state, hit, shadow records) + PCG
// state + two index queues + 4 counters; about 200 B per path slot.
struct PathBuf {
    int capacity = 0;
    float4* f4 = nullptr;
    ulonglong2* rng = nullptr;
    uint32_t* q = nullptr;
    uint32_t* counts = nullptr;
};

struct RenderStats {
    uint64_t batches = 0;
    uint64_t paths = 0;
};

void scene_upload(SceneDev& s, const lux_scene_desc& desc, hipStream_t stream);
void scene_release(SceneDev& s);
SceneView scene_view(const SceneDev& s);
void pathbuf_allocate(PathBuf& b, int capacity);
void pathbuf_release(PathBuf& b);
// Enqueues a full render into `film` (device, W*H*3 sums over spp).
void render(const SceneDev& s, PathBuf& buf, const lux_render_settings& rs, float* film, hipStream_t stream,
            RenderStats* stats);

}  // namespace pt
}  // namespace lux

// pt_runtime.hpp -- host-side state of the wavefront path tracer: the device
// scene (triangles, shading normals, materials, power-alias light table, camera,
// BLAS), the per-renderer path buffers and the batch machinery shared by single
// renders, multi-scene batches and the request server.
#pragma once

#include <cstddef>
#include <cstdint>
#include <memory>
#include <vector>

#include "lux_runtime.hpp"
#include "pt_device.hpp"
#include "two_level.hpp"

namespace lux {
namespace pt {

// Scene in HBM. Layout (T triangles, V vertices, M materials, K lights):
//   soup[T]          36-byte triangle records (vertex positions by triangle
//                    index; the traversal reads only these and the nodes)
//   vertices[V], indices[3T]  the indexed mesh (vertex updates gather the soup
//                    from these on the device)
//   normals[3V]      shading normals (after the identity-instance transform),
//                    read once per hit
//   mat_id[T], mats[M]        48-byte material records
//   lights[K], light_of_prim[T]  power-alias light table, emitter lookup
//   bvh              the LBVH over the triangle boxes (ids = triangle index)
// An instanced scene (meshes placed by affine instances) keeps the same shading
// arrays, indexed by scene ordinal (instances in order, each its mesh's
// triangles): soup[T] in world space, normals[9T] per corner in world space,
// indices[3T] = 0, 1, 2, ... (corner c of ordinal t is 3t + c), mat_id[T] after
// the instance's material override. Traversal uses the two-level structure
// (tl: a BLAS per mesh over its object-space triangles, a TLAS over the
// instances' world bounds) instead of bvh.
// Host copy of a one-mesh scene description (lux_scene_desc with its arrays,
// light images included).
struct SceneSource {
    std::vector<lux_vec3> vertices;  // as first uploaded (the build)
    std::vector<uint32_t> indices;
    std::vector<lux_vec3> normals;   // empty: none
    std::vector<int32_t> material_ids;
    std::vector<lux_material> materials;
    float camera[8] = {};
    std::vector<lux_light> lights;   // image_pixels point into `texels`
    std::vector<std::vector<lux_vec3>> texels;
    std::vector<lux_vec3> current;   // latest vertex update (empty: none)
};

struct SceneDev {
    BvhDev bvh;
    lux_triangle* soup = nullptr;
    lux_vec3* vertices = nullptr;
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
    // scene lights besides the emissive triangles, and the texels of image lights
    std::vector<lux_light> host_scene_lights;
    std::vector<float> host_env_pixels, host_env_cdf;
    std::vector<int32_t> host_env_first;  // per scene light: first texel (images), else -1
    std::vector<float> host_env_total;    // per scene light: total texel weight (images)
    float* env_pixels = nullptr;
    float* env_cdf = nullptr;
    // the BVH light sampler's tree over the lights (rebuilt with the light table)
    std::vector<float> host_light_power;
    std::vector<uint8_t> host_light_in_bvh;
    std::vector<DevAlias> host_nb_table;
    float host_nb_total = 0;
    std::vector<int32_t> host_env_lights;
    DevAlias* nb_table = nullptr;
    int32_t* env_light_list = nullptr;
    int nb_capacity = 0, env_list_capacity = 0;
    std::vector<lux_triangle> host_light_tri;
    std::vector<DevLightNode> host_light_nodes;
    std::vector<int32_t> host_light_leaf;
    DevLightNode* light_nodes = nullptr;
    int32_t* light_leaf = nullptr;
    int light_node_count = 0, light_node_capacity = 0, light_leaf_capacity = 0, light_root = -1;
    // emissive triangles in triangle order, with their vertex indices and
    // materials (the light table is rebuilt from these after a vertex update)
    std::vector<int32_t> emitters;
    std::vector<uint32_t> emitter_indices;
    std::vector<lux_material> host_mats;
    std::vector<int32_t> emitter_mat;
    size_t device_bytes = 0;  // HBM held by the scene (mesh + BVH)
    bool opaque = true;       // no pass-through (interface, transparent) material
    // instanced scenes
    bool instanced = false;
    TwoLevelDev tl;
    // the meshes as given (object space), kept to re-flatten after an instance update
    std::vector<lux_vec3> mesh_vertices, mesh_normals;
    std::vector<uint32_t> mesh_indices;
    std::vector<int32_t> mesh_vertex_off, mesh_tri_off, mesh_mat_id;
    int mesh_count = 0;
    // admitted, unfinished server requests: they hold this scene's device
    // pointers, so the instances may not move under them
    int server_jobs = 0;
    // One-mesh scenes keep the description they were made from (and their
    // latest vertex update), so the device state can be dropped and rebuilt:
    // rebuilt = uploaded from the original vertices, then refit to the latest
    // ones, which is the state the scene had (same topology, same boxes).
    std::shared_ptr<SceneSource> source;
    bool resident = false;  // the device arrays exist
};

// POD view of one scene handed to kernels.
struct SceneView {
    const lux_bvh_node* nodes;
    const uint32_t* slot_pre;  // preorder position of each pair slot (tie order of the ordered walk)
    const lux_triangle* soup;
    const uint32_t* indices;
    const float* normals;
    const int32_t* mat_id;
    const DevMaterial* mats;
    const DevLight* lights;
    const int32_t* light_of_prim;
    DevInfo* info;
    int32_t node_count, tri_count, mat_count, light_count;
    const DevLightNode* light_nodes;  // BVH light sampler: tree and leaf of each light
    const int32_t* light_leaf;
    int32_t light_node_count, light_root;  // root -1: no valid tree (power selection is used)
    const DevAlias* nb_table;              // the BVH sampler's non-BVH group
    int32_t nb_count;
    float nb_total;
    const int32_t* env_lights;             // infinite lights, in light-table order
    int32_t env_count;
    const float* env_pixels;               // image texels (RGB), all images back to back
    const float* env_cdf;                  // per-texel cumulative weights
    int32_t opaque;     // 1: every hit blocks a shadow ray (no pass-through material)
    int32_t instanced;  // 1: traverse `tl` (two-level) instead of `nodes`
    int32_t preorder;   // 1: closest hits by the right-first preorder walk only
    DevCamera cam;
    TwoLevelView tl;
};

// One segment of a batch: a contiguous range of one request's paths. A
// request's path index is r = sample * (width * height) + pixel (sample-major),
// so a segment holds whole or partial samples of the request's image.
struct Segment {
    SceneView sv;
    float* film;             // W*H*3 sums over the request's samples
    uint64_t seed;
    uint64_t req_begin;      // request path index of the segment's first path
    uint32_t slot_begin;     // its first slot in the batch
    uint32_t count;          // paths in the segment
    int32_t width, height, max_depth, frame, sample_offset, pad;
};

constexpr int kMaxSegments = 256;

struct BatchDesc {
    uint32_t count;  // paths in the batch (sum of segment counts)
    uint32_t nseg;
    uint32_t pad[2];
    Segment seg[kMaxSegments];
};

inline size_t desc_bytes(uint32_t nseg) { return offsetof(BatchDesc, seg) + sizeof(Segment) * nseg; }

// Per-renderer path state: 11 float4 planes (state, hit, shadow records) + PCG
// state + two index queues + counters; about 200 B per path slot.
struct PathBuf {
    int capacity = 0;
    float4* f4 = nullptr;
    ulonglong2* rng = nullptr;
    uint32_t* q = nullptr;
    // [0], [1] extension ray queues, [2], [3] shadow queues (by bounce parity)
    uint32_t* counts = nullptr;
    // rays traced since the last reset: [0] extension (closest-hit) rays, [1]
    // shadow rays, summed from the queue sizes of every bounce
    unsigned long long* rays = nullptr;
    BatchDesc* desc = nullptr;  // device copy of the current batch
};

// One render request in a renderer's work list.
struct Request {
    SceneView sv{};
    lux_render_settings rs{};
    float* film = nullptr;
    uint64_t total = 0;     // width * height * spp paths
    uint64_t issued = 0;    // paths already placed in batches
    int64_t last_batch = -1;
    uint64_t tag = 0;       // caller's identifier
};

struct RenderStats {
    uint64_t batches = 0;
    uint64_t paths = 0;
};

// Launch geometry of the wavefront schedule (set by the caller or the autotuner).
struct Tuning {
    int capacity = 1 << 21;  // path slots per batch
    int grid = 2048;         // workgroups of the grid-stride kernels
    int block = 256;         // threads per workgroup (128, 256 or 512)
    int use_graph = 1;       // replay each batch's bounce loop as one hipGraph
    int preorder_walk = 0;   // 1: extension rays take the right-first preorder walk only
};

// The batch engine: a path buffer, a pinned ring of batch descriptors and the
// captured bounce loop. Not thread-safe; one stream at a time.
class Renderer {
public:
    Renderer();
    ~Renderer();
    Renderer(const Renderer&) = delete;
    Renderer& operator=(const Renderer&) = delete;

    void set_tuning(const Tuning& t);
    const Tuning& tuning() const { return tune_; }

    // Enqueues every path of `reqs` (wavefront: multi-request batches of up to
    // `capacity` paths; plain: the reference's 65 536-path single-sample
    // batches, one shade launch per material type, a host sync per batch).
    void render_all(std::vector<Request*>& reqs, int mode, hipStream_t stream, RenderStats* stats);

    // Composes and enqueues ONE wavefront batch from `active` (in order; each
    // request gives at most `slice_spp` samples of its image, 0 = no limit).
    // Returns the number of paths placed (0 = nothing left). Sets last_batch of
    // requests whose final path is in this batch to the batch's index. A batch
    // holds one kind (traversal level, light sampler): the first unfinished
    // request's; the others wait for a later batch.
    uint64_t enqueue_batch(std::vector<Request*>& active, int slice_spp, int max_depth, hipStream_t stream,
                           int64_t* batch_index);

    void reset_ray_counts(hipStream_t stream);
    void ray_counts(hipStream_t stream, uint64_t* extension, uint64_t* shadow);
    size_t device_bytes() const;
    // HBM the engine holds once it runs batches of the tuned capacity.
    size_t planned_bytes() const;
    int64_t batches_enqueued() const { return batch_counter_; }

private:
    void ensure_capacity(int paths);
    BatchDesc* next_slot();
    void upload(BatchDesc* slot, hipStream_t stream);
    void launch_wavefront(int depth, bool opaque, bool two_level, int light, hipStream_t stream);
    void launch_plain(int depth, uint32_t count, bool two_level, int light, hipStream_t stream);
    void release_graph();

    Tuning tune_;
    PathBuf buf_;
    static constexpr int kRing = 8;
    BatchDesc* ring_ = nullptr;  // pinned host slots
    hipEvent_t ring_ev_[kRing] = {};
    bool ring_used_[kRing] = {};
    int ring_next_ = 0;
    hipGraphExec_t graph_ = nullptr;
    int graph_depth_ = -1, graph_grid_ = -1, graph_capacity_ = -1, graph_block_ = -1, graph_opaque_ = -1,
        graph_two_level_ = -1, graph_light_ = -1;
    // k_shade blocks resident on the device for (block, light), queried outside any capture
    int shade_resident_ = 1, shade_resident_block_ = -1, shade_resident_light_ = -1;
    int64_t batch_counter_ = 0;
};

void scene_upload(SceneDev& s, const lux_scene_desc& desc, hipStream_t stream);
// A one-mesh scene with its description kept: uploaded now (upload = true) or
// on first use (scene_restore), after the checks scene_upload makes.
void scene_create(SceneDev& s, const lux_scene_desc& desc, bool upload, hipStream_t stream);
// Builds the device state of a kept description into `fresh` (synchronous on
// `stream`); scene_adopt then moves it into the scene, which keeps its camera,
// source and server count.
void scene_restore(SceneDev& fresh, const SceneSource& src, hipStream_t stream);
void scene_adopt(SceneDev& s, SceneDev&& fresh);
// Detaches the device state of a scene that has a source (the returned object
// owns it: scene_release frees it); the scene stays valid, not resident.
SceneDev scene_detach(SceneDev& s);
// Replaces vertex positions (same count), regathers the soup on the device,
// refits the BVH and rebuilds the light table; the topology is kept.
void scene_update_vertices(SceneDev& s, const lux_vec3* vertices, int vertex_count, hipStream_t stream);
// An instanced scene: BLAS per mesh, TLAS over the instances, world-space
// shading arrays by scene ordinal and the light table over them.
void scene_upload_instanced(SceneDev& s, const lux_instanced_scene_desc& desc, hipStream_t stream);
// New instance transforms (same count): rebuilds the TLAS, the world-space
// shading arrays and the light table; the BLASes are kept.
void scene_set_instances(SceneDev& s, const lux_instance* instances, int instance_count, hipStream_t stream);
void scene_release(SceneDev& s);
SceneView scene_view(const SceneDev& s);
uint64_t request_paths(const lux_render_settings& rs);
// Batch kind of a request: bit 0 two-level traversal, bits 1.. the light sampler.
int request_kind(const Request& rq);

}  // namespace pt
}  // namespace lux

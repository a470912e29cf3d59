/*
 * lux_hip.h -- C ABI of the MI355X (gfx950) hot path of LUX: LBVH build/refit,
 * closest-hit / any-hit traversal and AABB overlap queries.
 *
 * Every entry point is extern "C", takes plain pointers and sizes, and returns an
 * int status (LUX_OK == 0). On failure the thread-local string returned by
 * lux_hip_last_error() says what went wrong.
 *
 * Two families of calls exist for every operation:
 *   - host-array calls (no suffix): copy inputs in, run, copy results out and
 *     synchronise, exactly like the reference C++ functions they replace;
 *   - "_device" calls: operands are device pointers already resident in HBM, the
 *     work is enqueued on `stream` (NULL = the library stream) and nothing is
 *     synchronised; results land in device memory.
 *
 * Record layouts are byte-identical to the reference's POD types so that arrays
 * can be handed over without repacking:
 *   lux_aabb      == AABB         src/core/aabb.cuh:10-13      (24 B)
 *   lux_ray       == Ray          src/core/ray.cuh:10-15       (36 B)
 *   lux_triangle  == Triangle     src/geometry/triangle.h:10-12 (36 B)
 *   lux_hit       == GpuRayHit    src/bvh/gpu_bvh.cuh:20-23    (8 B)
 *   lux_pair      == PrimitivePair src/bvh/bvh_overlap.cuh:10-13 (8 B)
 *   lux_bvh_node  == GpuBvhNode   src/bvh/bvh_device.cuh:12-17 (32 B)
 *   lux_bvh_view  == GpuBvhView   src/bvh/bvh_device.cuh:19-23
 *
 * Node layout contract (lux_bvh_view): nodes are stored in RIGHT-FIRST preorder.
 * A walker that tests node i, descends with i+1 on an internal hit and jumps to
 * nodes[i].escape on a miss or after a leaf visits nodes in exactly the order of
 * the host LBVH stack traversal (src/bvh/cpu_bvh_traversal.h:37-78, which pops
 * the right child first), so closest-hit ties resolve like the host oracle.
 */
#ifndef LUX_HIP_H
#define LUX_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
enum {
    LUX_OK = 0,
    LUX_ERR_INVALID_ARGUMENT = 1, /* null pointer, negative count, bad id ...      */
    LUX_ERR_OVERFLOW = 2,         /* a fixed-capacity buffer/queue overflowed       */
    LUX_ERR_MISMATCH = 3,         /* refit input does not match the built topology  */
    LUX_ERR_HIP = 4,              /* a HIP runtime call failed                      */
    LUX_ERR_NOT_INITIALIZED = 5,  /* lux_hip_init() has not run                     */
    LUX_ERR_OUT_OF_MEMORY = 6,
    LUX_ERR_DEVICE_CHECK = 7      /* a kernel's own bounds/termination check fired  */
};

/* ---- POD records (layouts as listed above) ----------------------------- */
typedef struct lux_vec3 { float x, y, z; } lux_vec3;
typedef struct lux_aabb { lux_vec3 p_min, p_max; } lux_aabb;
typedef struct lux_ray {
    lux_vec3 origin, direction;
    float t_min, t_max, time;
} lux_ray;
typedef struct lux_triangle { lux_vec3 v0, v1, v2; } lux_triangle;
typedef struct lux_hit { int32_t primitive_id; float t; } lux_hit;
typedef struct lux_pair { int32_t a, b; } lux_pair;
typedef struct lux_bvh_node {
    lux_aabb bounds;
    int32_t primitive_id; /* -1 = internal node, >= 0 = leaf primitive id      */
    int32_t escape;       /* next node after a miss / finished subtree, -1 = end */
} lux_bvh_node;
typedef struct lux_bvh_view {
    const lux_bvh_node* nodes; /* device pointer, valid while the handle lives */
    int32_t node_count;
    lux_aabb root_bounds;
} lux_bvh_view;

typedef struct lux_hip_bvh lux_hip_bvh; /* opaque owner of one LBVH (GpuBvh) */
typedef void* lux_stream;               /* a hipStream_t; NULL = library stream */

/* ---- library lifetime --------------------------------------------------- */
/* Selects `device`, creates the library stream. Replaces the implicit device-0
 * context of the reference (src/core/cuda_check.cuh:7-27 error convention).   */
int lux_hip_init(int device);
int lux_hip_shutdown(void);
/* Thread-local text of the last failure ("" if none). */
const char* lux_hip_last_error(void);
/* Library version, e.g. 0x000100 for 0.1.0. */
int lux_hip_version(void);

/* Name, architecture (e.g. "gfx950:sramecc+:xnack-"), compute units and HBM
 * size of the library's device; any pointer may be NULL. */
int lux_hip_device_info(char* name, int32_t name_len, char* arch, int32_t arch_len, int32_t* compute_units,
                        uint64_t* global_mem_bytes);

/* ---- device memory helpers (for callers without a GPU framework) ------- */
int lux_hip_malloc(void** device_ptr, size_t bytes);
int lux_hip_free(void* device_ptr);
int lux_hip_memcpy_htod(void* device_dst, const void* host_src, size_t bytes);
int lux_hip_memcpy_dtoh(void* host_dst, const void* device_src, size_t bytes);
int lux_hip_synchronize(void);
/* Device-to-device copy by a plain streaming kernel (16-byte vector loads and
 * stores; bytes must be a multiple of 16), enqueued on `stream`. Also the
 * reference workload for calibrating HBM byte counters. */
int lux_hip_copy_device(void* d_dst, const void* d_src, size_t bytes, lux_stream stream);

/* ---- LBVH ------------------------------------------------------------- */
/* GpuBvh::GpuBvh / ~GpuBvh  (src/bvh/gpu_bvh.cuh:25-47). */
int lux_hip_bvh_create(lux_hip_bvh** out);
int lux_hip_bvh_destroy(lux_hip_bvh* bvh);

/* build_gpu_lbvh (src/bvh/gpu_bvh.cuh:50, src/bvh/gpu_bvh.cu:535-595).
 * boxes[i] / primitive_ids[i] are BvhBuildPrimitive::bounds / ::primitive_id
 * (src/bvh/bvh_build_input.h:15-23); Morton codes use the box centroid exactly as
 * the GPU reference does (src/bvh/gpu_bvh.cu:143). n == 0 leaves the handle
 * invalid with primitive_count 0, like the reference. Ids must be >= 0. */
int lux_hip_bvh_build(lux_hip_bvh* bvh, const lux_aabb* boxes,
                      const int32_t* primitive_ids, int32_t n);
int lux_hip_bvh_build_device(lux_hip_bvh* bvh, const lux_aabb* d_boxes,
                             const int32_t* d_primitive_ids, int32_t n,
                             lux_stream stream);
/* build_gpu_lbvh over the primitives collect_scene_mesh_bvh_build_primitives
 * makes from a triangle list (src/scene/scene_prepare.h:420-439), with the
 * host loop and copy of src/bvh/gpu_bvh.cu:543-563 moved onto the device: box
 * i = triangle_bounds(d_triangles[i]) (src/geometry/triangle.h:22-28),
 * primitive id i. Same tree as lux_hip_bvh_build with those boxes and ids. */
int lux_hip_bvh_build_triangles_device(lux_hip_bvh* bvh, const lux_triangle* d_triangles, int32_t n,
                                       lux_stream stream);

/* refit_gpu_lbvh (src/bvh/gpu_bvh.cuh:51, src/bvh/gpu_bvh.cu:597-613): same
 * topology, new boxes. Returns LUX_ERR_MISMATCH (the reference returns false)
 * when n differs from the built count or primitive_ids differ element-wise. */
int lux_hip_bvh_refit(lux_hip_bvh* bvh, const lux_aabb* boxes,
                      const int32_t* primitive_ids, int32_t n);
/* Device-resident refit; the ids are taken to be unchanged. */
int lux_hip_bvh_refit_device(lux_hip_bvh* bvh, const lux_aabb* d_boxes, int32_t n,
                             lux_stream stream);
/* Refit straight from triangle vertices: box of input primitive i is the
 * bounds of d_triangles[i] (triangle_bounds, src/geometry/triangle.h:22-28).
 * Fuses build-input preparation (src/scene/scene_prepare.h:420-439) into the
 * bottom-up pass. */
int lux_hip_bvh_refit_triangles_device(lux_hip_bvh* bvh,
                                       const lux_triangle* d_triangles, int32_t n,
                                       lux_stream stream);

/* Closest-hit and any-hit calls with at least 65 536 rays trace them sorted by
 * direction octant, then by the Morton code of their origins (2-pass radix
 * sort of 16 key bits); hits and occlusion flags still land at each ray's own
 * index and are identical either way. On by default; enable = 0 traces in
 * input order. No reference counterpart. */
int lux_hip_bvh_set_ray_order(lux_hip_bvh* bvh, int32_t enable);

/* GpuBvh::is_valid / primitive_count / root_bounds (src/bvh/gpu_bvh.cuh:35-47).
 * Synchronises with outstanding device work on the handle. */
int lux_hip_bvh_info(lux_hip_bvh* bvh, int32_t* valid, int32_t* primitive_count,
                     lux_aabb* root_bounds);
/* GpuBvh::device_view (src/bvh/gpu_bvh.cuh:38). */
int lux_hip_bvh_view(lux_hip_bvh* bvh, lux_bvh_view* out);
/* Copies the stackless node array (2n-1 records) to the host. */
int lux_hip_bvh_download_nodes(lux_hip_bvh* bvh, lux_bvh_node* nodes,
                               int32_t capacity);
/* Radix-tree topology for bit-exact checks and for download_bvh_for_validation
 * (src/bvh/gpu_bvh.cu:615-665). Any pointer may be NULL. Sizes: morton_sorted,
 * sorted_indices, leaf_parent: n; left, right, parent: n-1 (bit 31 of left/right
 * = leaf child, bit 31 of parent/leaf_parent = "I am the right child", as
 * src/bvh/gpu_bvh.cu:165-179); node_bounds: 2n-1 in host-LBVH order (internal
 * nodes 0..n-2, then leaf i at n-1+i, src/bvh/cpu_lbvh_builder.h:245-257). */
int lux_hip_bvh_download_topology(lux_hip_bvh* bvh, uint32_t* morton_sorted,
                                  uint32_t* sorted_indices, uint32_t* left,
                                  uint32_t* right, uint32_t* parent,
                                  uint32_t* leaf_parent, lux_aabb* node_bounds);
/* The primitives the tree holds, in input order: boxes (as last built or
 * refit) and primitive ids, n each (GpuBvhStorage::input_bounds /
 * primitive_ids, read by download_bvh_for_validation and
 * validate_gpu_bvh_matches_primitives, src/bvh/gpu_bvh.cu:615-891). Either
 * pointer may be NULL. */
int lux_hip_bvh_download_input(lux_hip_bvh* bvh, lux_aabb* boxes, int32_t* primitive_ids);
/* GpuBvh::overlap_bounds_device_ptr / overlap_primitive_ids_device_ptr /
 * overlap_sorted_indices_device_ptr (src/bvh/gpu_bvh.cuh:41-43, gpu_bvh.cu:443-453):
 * the device arrays of the input boxes and ids (input order) and of the sorted
 * input indices, n each; NULL for an invalid tree. Valid until the next build
 * or the handle's destruction. Any out pointer may be NULL. */
int lux_hip_bvh_device_arrays(lux_hip_bvh* bvh, const lux_aabb** d_input_bounds,
                              const int32_t** d_primitive_ids, const uint32_t** d_sorted_indices);

/* ---- queries ----------------------------------------------------------- */
/* intersect_gpu_bvh_stackless (src/bvh/gpu_bvh.cuh:57-59, gpu_bvh.cu:893-915):
 * closest hit per ray; hits[i] = {primitive_id, t} or {-1, +INF}. triangles is
 * indexed by primitive id; triangle_count bounds every leaf id. */
int lux_hip_trace_closest(lux_hip_bvh* bvh, const lux_triangle* triangles,
                          int32_t triangle_count, const lux_ray* rays,
                          int32_t ray_count, lux_hit* hits);
int lux_hip_trace_closest_device(lux_hip_bvh* bvh, const lux_triangle* d_triangles,
                                 int32_t triangle_count, const lux_ray* d_rays,
                                 int32_t ray_count, lux_hit* d_hits,
                                 lux_stream stream);
/* lux_hip_trace_closest_device without allocation. No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * The only scratch of the closest-hit query, the ray-order sort, lives in a
 * fixed 68 MiB device array that each GPU holds from library load until the
 * process ends; ray_count above 4194304 returns LUX_ERR_INVALID_ARGUMENT. The
 * hits equal lux_hip_trace_closest_device's. Every call on a GPU shares that
 * GPU's array: calls that overlap on one GPU (any streams) race. */
int lux_hip_trace_closest_device_noalloc(lux_hip_bvh* bvh, const lux_triangle* d_triangles,
                                         int32_t triangle_count, const lux_ray* d_rays,
                                         int32_t ray_count, lux_hit* d_hits,
                                         lux_stream stream);
/* Any-hit occlusion: occluded[i] = 1 iff some triangle is hit with
 * t_min < t < t_max. The reference has no early-exit any-hit; its shadow path
 * repeats the closest-hit walk (src/integrator/wavefront_kernels.cu:28-83), so
 * occluded[i] == (closest-hit found a hit) is the exact contract. */
int lux_hip_trace_any(lux_hip_bvh* bvh, const lux_triangle* triangles,
                      int32_t triangle_count, const lux_ray* rays, int32_t ray_count,
                      uint8_t* occluded);
int lux_hip_trace_any_device(lux_hip_bvh* bvh, const lux_triangle* d_triangles,
                             int32_t triangle_count, const lux_ray* d_rays,
                             int32_t ray_count, uint8_t* d_occluded,
                             lux_stream stream);

/* query_aabb_overlaps_self (src/bvh/bvh_overlap.cuh:21, bvh_overlap.cu:186-215):
 * all pairs (a, b), a < b, of primitives whose boxes overlap (inclusive,
 * AABB::overlaps src/core/aabb.cuh:68-72). *found_count is the exact total;
 * at most max_pairs are written (truncated iff *found_count > max_pairs).
 * Pair order is unspecified. max_pairs <= 0 returns found_count 0. */
int lux_hip_overlap_self(lux_hip_bvh* bvh, int32_t max_pairs, lux_pair* pairs,
                         int32_t* found_count);
int lux_hip_overlap_self_device(lux_hip_bvh* bvh, int32_t max_pairs,
                                lux_pair* d_pairs, int32_t* d_found_count,
                                lux_stream stream);
/* query_aabb_overlaps_cross (src/bvh/bvh_overlap.cuh:22-23, bvh_overlap.cu:217-251):
 * pairs (left primitive id, right primitive id); right boxes are visited in
 * the right tree's Morton order. */
int lux_hip_overlap_cross(lux_hip_bvh* left, lux_hip_bvh* right, int32_t max_pairs,
                          lux_pair* pairs, int32_t* found_count);
int lux_hip_overlap_cross_device(lux_hip_bvh* left, lux_hip_bvh* right,
                                 int32_t max_pairs, lux_pair* d_pairs,
                                 int32_t* d_found_count, lux_stream stream);

/* ---- wavefront path tracer --------------------------------------------- */
/* Material record (the fields of Material, src/material/material.h:24-35;
 * constant emission only). type follows MaterialType (:6-13). */
enum {
    LUX_MAT_LAMBERT = 0,
    LUX_MAT_THIN_LAMBERT = 1,
    LUX_MAT_GGX = 2,
    LUX_MAT_DIELECTRIC = 3,
    LUX_MAT_INTERFACE = 4,
    LUX_MAT_TRANSPARENT = 5
};
typedef struct lux_material {
    int32_t type;
    lux_vec3 albedo;
    lux_vec3 emission;
    int32_t emission_two_sided; /* EmissionSidedness::DoubleSided */
    float roughness, metallic, eta;
} lux_material;

/* Scene lights besides the emissive triangles (Scene::lights,
 * src/light/light.h:66-99, added by add_point_light / add_distant_light /
 * add_uniform_infinite_light / add_image_infinite_light,
 * src/scene/scene_build.h:9-90). They join the light table after the area
 * lights, in this order, with the reference's power (max(1e-6, luminance) of
 * intensity or radiance; an image's luminance x sin(theta) weight sum). Point
 * lights enter the light BVH as points; the other kinds are the BVH sampler's
 * non-BVH group. Infinite lights are seen by rays that leave the scene. */
enum { LUX_LIGHT_POINT = 1, LUX_LIGHT_DISTANT = 2, LUX_LIGHT_UNIFORM_INFINITE = 3, LUX_LIGHT_IMAGE_INFINITE = 4 };
typedef struct lux_light {
    int32_t kind;                 /* LUX_LIGHT_* */
    lux_vec3 vector;              /* point: position; distant: direction to the light (normalised here) */
    lux_vec3 color;               /* point: intensity; distant, uniform infinite: radiance; image: scale */
    int32_t image_width, image_height;
    const lux_vec3* image_pixels; /* image infinite: width x height, row 0 at theta = 0 (+y) */
} lux_light;

/* One mesh under an identity instance (SceneMesh + SceneInstance,
 * src/scene/scene_data.h:35-58) and a pinhole camera given as
 * make_pinhole_camera's arguments {origin xyz, look_at xyz, vertical fov in
 * degrees, aspect} (src/camera/pinhole_camera.h:13-17). normals may be NULL
 * (flat shading). Emissive triangles become power-sampled area lights
 * (prepare_scene_light_sampling, src/scene/scene_prepare.h:268-387). */
typedef struct lux_scene_desc {
    const lux_vec3* vertices;
    int32_t vertex_count;
    const uint32_t* indices; /* 3 per triangle */
    const lux_vec3* normals; /* per vertex or NULL */
    const int32_t* material_ids;
    int32_t triangle_count;
    const lux_material* materials;
    int32_t material_count;
    float camera[8];
    const lux_light* lights; /* scene lights (may be NULL) */
    int32_t light_count;
} lux_scene_desc;

enum { LUX_RENDER_WAVEFRONT = 0, LUX_RENDER_PLAIN = 1 };
/* RenderSettings::light_sampler_kind (src/light/light_distribution.h:27-31;
 * default Power): POWER picks an area light with the power alias table, UNIFORM
 * uniformly, BVH by descending the light BVH with the importance power /
 * distance^2 at the shading point (sample_light_bvh,
 * src/light/light_sampler.h:142-195). */
enum { LUX_LIGHT_SAMPLER_POWER = 0, LUX_LIGHT_SAMPLER_UNIFORM = 1, LUX_LIGHT_SAMPLER_BVH = 2 };
/* RenderSettings (src/integrator/render_context.h:16-24) + film size. mode
 * PLAIN reproduces the reference schedule (batch_size paths per batch, one
 * sample per batch, full-grid launches, a host sync per batch); WAVEFRONT packs
 * several samples per batch and never waits on the host mid-render. batch_size
 * 0 picks the mode's default. */
typedef struct lux_render_settings {
    int32_t width, height, spp, max_depth;
    uint64_t seed;
    int32_t frame_index, sample_offset;
    int32_t batch_size;
    int32_t mode;
    int32_t light_sampler; /* LUX_LIGHT_SAMPLER_* */
} lux_render_settings;

typedef struct lux_hip_scene lux_hip_scene;

/* PreparedScene(scene) + prepare_gpu() (src/scene/prepared_scene.cu:19-23,
 * :61-79): light tables, camera, device arrays and the LBVH over the
 * triangles. Synchronous. The description is copied and kept, so a server may
 * drop the device state of an idle scene and rebuild it later (the same
 * state: built from these vertices, refit to the latest update). */
int lux_hip_scene_create(const lux_scene_desc* desc, lux_hip_scene** out);
/* PreparedScene(scene) without prepare_gpu(): the description is checked and
 * copied; the device state is built on first use (a direct render, query or
 * inspection call, or a server admitting a request for it). */
int lux_hip_scene_create_deferred(const lux_scene_desc* desc, lux_hip_scene** out);
/* PreparedScene::has_gpu (src/scene/prepared_scene.cu:81-83): 1 when the
 * scene's device state exists. */
int lux_hip_scene_resident(lux_hip_scene* scene, int32_t* resident);
/* LUX_ERR_INVALID_ARGUMENT while a server request for the scene (submitted
 * and not yet completed) exists. */
int lux_hip_scene_destroy(lux_hip_scene* scene);
/* Camera vectors {origin, lower_left, horizontal, vertical} (12 floats). */
int lux_hip_scene_camera(lux_hip_scene* scene, float* out12);
/* Replaces the camera with already built PerspectiveCamera vectors
 * {origin, lower_left, horizontal, vertical} (src/camera/camera.h:29-38,
 * pinhole: lens_radius 0), e.g. the Camera of a PreparedScene's Scene. */
int lux_hip_scene_set_camera(lux_hip_scene* scene, const float* v12);
/* Light table: count, and (if the arrays are non-NULL, sized by count) each
 * light's triangle (-1 and area 0 for scene lights), area, alias probability,
 * alias index and pmf;
 * light_of_prim gets one entry per triangle (-1 = not an emitter). */
int lux_hip_scene_lights(lux_hip_scene* scene, int32_t* count, int32_t* prim, float* area, float* prob,
                         int32_t* alias, float* pmf, int32_t* light_of_prim);
/* The BVH light sampler's tree (build_light_bvh, src/light/light_distribution.cpp:68-161,
 * nodes in build order): node count, and if the arrays are non-NULL (sized by
 * the counts) each node's bounds {min xyz, max xyz}, power and {left, right,
 * parent, light index}, the leaf node of each light, and the root (-1: no
 * valid tree, the sampler then falls back to power selection). */
int lux_hip_scene_light_bvh(lux_hip_scene* scene, int32_t* node_count, float* bounds6, float* power,
                            int32_t* links4, int32_t* leaf_of_light, int32_t* root);
/* The LBVH of the scene's triangles (for the traversal entry points). */
int lux_hip_scene_bvh(lux_hip_scene* scene, lux_hip_bvh** out_borrowed);

/* WavefrontPathTracer::render (src/integrator/wavefront_pt.cu:51-84): the film
 * receives per-pixel SUMS over spp, row-major with y = 0 at the top, like
 * Film::pixels (src/film/film.h:15-67). Host film: synchronous. */
int lux_hip_render(lux_hip_scene* scene, const lux_render_settings* settings, float* film_sum);
int lux_hip_render_device(lux_hip_scene* scene, const lux_render_settings* settings, float* d_film_sum,
                          lux_stream stream);
/* Rays traced by the scene's last render: extension (closest-hit) rays and
 * shadow rays, summed from every bounce's queue sizes. No reference
 * counterpart (the reference keeps no ray count); synchronises `stream`. */
int lux_hip_render_rays(lux_hip_scene* scene, lux_stream stream, uint64_t* extension_rays, uint64_t* shadow_rays);

/* Triangle and vertex counts and the HBM the scene holds (mesh + LBVH). */
int lux_hip_scene_info(lux_hip_scene* scene, int32_t* triangle_count, int32_t* vertex_count, uint64_t* device_bytes);

/* Moves the scene's vertices (same count, same triangles): the triangles are
 * regathered on the device and the LBVH is refit with its topology kept, as
 * refit_gpu_lbvh does for a primitive set with unchanged ids
 * (src/bvh/gpu_bvh.cu:597-613); emitter areas and the power-alias light table
 * are rebuilt. Renders enqueued earlier read the old geometry. Synchronous.
 * LUX_ERR_INVALID_ARGUMENT while a server request for the scene is unfinished
 * (send the update with lux_hip_server_submit instead). */
int lux_hip_scene_update_vertices(lux_hip_scene* scene, const lux_vec3* vertices, int32_t vertex_count);

/* ---- multi-scene batches ------------------------------------------------
 * Batched WavefrontPathTracer::render over n (scene, settings, film) jobs:
 * the jobs' paths share one wavefront loop over a multi-scene ray queue (per
 * path: its scene's BVH root, mesh, lights and camera), each batch's bounce
 * loop replayed as one captured hipGraph. Each film equals the film of the
 * same job rendered alone, bit for bit (per-path arithmetic is independent
 * of the batch, and every pixel's samples are summed in sample order).
 * mode LUX_RENDER_PLAIN renders the jobs one after another on the
 * reference's schedule (the comparison baseline). No reference counterpart
 * (the reference renders one scene per call). */
int lux_hip_batch_render(lux_hip_scene* const* scenes, const lux_render_settings* settings, float* const* films,
                         int32_t n, int32_t mode);
int lux_hip_batch_render_device(lux_hip_scene* const* scenes, const lux_render_settings* settings,
                                float* const* d_films, int32_t n, int32_t mode, lux_stream stream);
/* Rays traced by the last batch (extension, shadow); synchronises `stream`. */
int lux_hip_batch_rays(lux_stream stream, uint64_t* extension_rays, uint64_t* shadow_rays);

/* Launch geometry of the wavefront schedule: path slots per batch, workgroups
 * of the grid-stride kernels, threads per workgroup (128, 256 or 512), graph
 * replay on/off, and preorder_walk: 1 makes extension rays take only the
 * right-first preorder walk (the host LBVH's visit order) instead of the
 * near-child-first walk, for checking that both give the same hits. Applies to
 * renders started after the call. No reference counterpart (the reference's
 * capacity is RenderSettings::batch_size, src/integrator/render_context.h:16-24). */
typedef struct lux_render_tuning {
    int32_t capacity;
    int32_t grid;
    int32_t block;
    int32_t use_graph;
    int32_t preorder_walk;
} lux_render_tuning;
int lux_hip_set_render_tuning(const lux_render_tuning* tuning);
int lux_hip_get_render_tuning(lux_render_tuning* tuning);

/* ---- request server ------------------------------------------------------
 * Continuous batching: render requests are admitted in arrival order while
 * their HBM footprint (resident scenes + films + path buffer) fits the
 * budget, join the running wavefront loop in slices of slice_spp samples
 * (round robin) while earlier requests are still on the GPU, and complete
 * when the batch holding their last path finishes. A request carrying new
 * vertex positions waits until its scene has no unfinished request, then the
 * scene is refit (lux_hip_scene_update_vertices) before its first batch.
 * Every film equals the film of the same render run alone. Scenes must
 * outlive the server. Not thread-safe. No reference counterpart (the
 * reference renders one request per process run, src/app/render_runner.cpp). */
typedef struct lux_serve_config {
    uint64_t hbm_budget_bytes;
    int32_t slice_spp;     /* samples per request per turn; 0 = no limit */
    int32_t max_inflight;  /* batches enqueued ahead of the GPU (1..64) */
} lux_serve_config;

typedef struct lux_serve_stats {
    uint64_t submitted, admitted, completed, refits;
    uint64_t batches, paths;
    uint64_t resident_bytes, peak_bytes;
    uint64_t extension_rays, shadow_rays;
    uint64_t shared_batches;      /* batches that held paths of two or more requests */
    uint64_t max_batch_requests;  /* most requests with paths in one batch */
    uint64_t uploads;             /* scenes the server made resident (upload + build) */
    uint64_t evictions;           /* idle scenes the server dropped to fit the budget */
    double prepare_seconds;       /* worker-thread time of uploads and refits */
} lux_serve_stats;

typedef struct lux_hip_server lux_hip_server;

int lux_hip_server_create(const lux_serve_config* config, lux_hip_server** out);
int lux_hip_server_destroy(lux_hip_server* server);
/* Queues a request (new_vertices may be NULL; otherwise vertex_count must be
 * the scene's and the array is copied). Returns immediately. */
int lux_hip_server_submit(lux_hip_server* server, lux_hip_scene* scene, const lux_render_settings* settings,
                          const lux_vec3* new_vertices, int32_t vertex_count, uint64_t* request_id);
/* One scheduling round (retire, admit, enqueue); with block != 0 it waits for
 * the oldest batch on the GPU when no new batch could be enqueued. Returns up
 * to max_ids completed request ids. */
int lux_hip_server_poll(lux_hip_server* server, int32_t block, uint64_t* done_ids, int32_t max_ids,
                        int32_t* n_done);
/* Film sums (W*H*3, may be NULL) and arrival-to-completion latency of a
 * completed request; releases its film. */
int lux_hip_server_result(lux_hip_server* server, uint64_t request_id, float* film_sum, double* latency_s);
int lux_hip_server_stats(lux_hip_server* server, lux_serve_stats* stats);
/* Replays an arrival trace in real time on the server: request i arrives
 * arrival_s[i] seconds after the call (sorted), renders settings[i] of
 * scenes[i] (after moving the vertices to new_vertices[i] if new_vertices and
 * new_vertices[i] are non-NULL). films (may be NULL) receive the film sums;
 * latency_s[i] = completion - arrival; wall_s = total time. */
int lux_hip_server_run(lux_hip_server* server, int32_t n, lux_hip_scene* const* scenes,
                       const lux_render_settings* settings, const double* arrival_s,
                       const lux_vec3* const* new_vertices, const int32_t* vertex_counts, float* const* films,
                       double* latency_s, double* wall_s);

/* ---- two-level (instanced) scenes ---------------------------------------
 * A mesh is a run of object-space triangles; an instance places a mesh with an
 * affine transform. One LBVH per non-empty mesh over its triangle boxes (BLAS)
 * and one over the non-empty instance world bounds with the instance ids
 * (TLAS) replace PreparedScene::prepare_gpu's BVH part
 * (src/scene/prepared_scene.cu:59-78, src/scene/scene_prepare.h:420-462); the
 * walk replaces intersect_scene_bvh_device (src/scene/gpu_scene_bvh.cuh:44-135):
 * TLAS right-first walk; at an instance, the ray through Transform::inv
 * (inverse_transform_ray, src/core/mat4.cuh:173-252) with t_max = the current
 * hit, the BLAS walk, and a nearer hit becomes scene ordinal
 * first_primitive + local id (src/scene/gpu_scene.h:149-172). Ids, t, u, v and
 * tie winners equal the reference's host two-level walk
 * (intersect_scene_two_level_bvh, src/scene/scene_intersection.h:147-240). */
typedef struct lux_instance {
    int32_t mesh_id;            /* SceneInstance::mesh_id */
    int32_t material_override;  /* SceneInstance::material_override (-1: none); not used by the walk */
    float object_to_world[16];  /* Transform::m, row-major */
    float world_to_object[16];  /* Transform::inv, row-major (inverse_affine(m) in the reference) */
    lux_aabb world_bounds;      /* SceneInstance::world_bounds; an empty box keeps it out of the TLAS */
} lux_instance;
typedef struct lux_scene_hit {
    int32_t primitive_id;       /* scene ordinal (-1 = no hit) */
    float t, u, v;
    int32_t instance_id;        /* PrimitiveRef::instance_id */
    int32_t local_primitive_id; /* PrimitiveRef::local_primitive_id */
} lux_scene_hit;
typedef struct lux_hip_two_level lux_hip_two_level;

int lux_hip_two_level_create(lux_hip_two_level** out);
int lux_hip_two_level_destroy(lux_hip_two_level* scene);
/* Builds the BLAS of every mesh (mesh m owns triangles [mesh_offsets[m],
 * mesh_offsets[m+1]) of `triangles`) and the TLAS. Instances are checked as
 * scene validation checks them (src/scene/scene_validate.h:140-147: mesh id in
 * range, both matrices affine-invertible, m * inv within 1e-4 of the identity);
 * a failure returns LUX_ERR_INVALID_ARGUMENT. Synchronous. */
int lux_hip_two_level_build(lux_hip_two_level* scene, const lux_triangle* triangles, const int32_t* mesh_offsets,
                            int32_t mesh_count, const lux_instance* instances, int32_t instance_count);
/* New transforms / world bounds / mesh ids for the same instance count: only
 * the TLAS and the instance records are rebuilt. Synchronous. */
int lux_hip_two_level_set_instances(lux_hip_two_level* scene, const lux_instance* instances,
                                    int32_t instance_count);
/* Closest hit per ray (intersect_scene_bvh / intersect_scene_bvh_skip_ref):
 * ignore (nullable) holds one scene ordinal per ray to skip, -1 for none.
 * Host arrays, synchronous; the _device forms take device pointers and are
 * stream-ordered. */
int lux_hip_two_level_intersect(lux_hip_two_level* scene, const lux_ray* rays, const int32_t* ignore,
                                int32_t ray_count, lux_scene_hit* hits);
int lux_hip_two_level_intersect_device(lux_hip_two_level* scene, const lux_ray* d_rays, const int32_t* d_ignore,
                                       int32_t ray_count, lux_scene_hit* d_hits, lux_stream stream);
/* 1 where the ray hits any triangle in [t_min, t_max) other than `ignore`
 * (exactly where the closest-hit walk finds a hit), else 0. */
int lux_hip_two_level_occluded(lux_hip_two_level* scene, const lux_ray* rays, const int32_t* ignore,
                               int32_t ray_count, uint8_t* occluded);
int lux_hip_two_level_occluded_device(lux_hip_two_level* scene, const lux_ray* d_rays, const int32_t* d_ignore,
                                      int32_t ray_count, uint8_t* d_occluded, lux_stream stream);
/* Instances in the TLAS, scene ordinals (triangles over all instances), HBM held. */
int lux_hip_two_level_info(lux_hip_two_level* scene, int32_t* tlas_instances, int64_t* scene_primitives,
                           uint64_t* device_bytes);
/* Host helpers with the reference's arithmetic: make_transform's inverse
 * (inverse_affine, src/core/mat4.cuh:126-159) and transform_bounds
 * (src/core/mat4.cuh:254-270), for callers that build lux_instance records. */
int lux_hip_inverse_affine(const float* m16, float* inv16);
int lux_hip_transform_bounds(const float* m16, const lux_aabb* bounds, lux_aabb* out);

/* ---- instanced scenes for the path tracer ---------------------------------
 * A Scene of several meshes placed by affine instances (Scene::mesh_assets +
 * Scene::instances, src/scene/scene_data.h:35-58, built by add_scene_mesh /
 * add_scene_instance, src/scene/scene_edit.h), rendered through the same
 * lux_hip_render / lux_hip_batch_render / server entry points as a one-mesh
 * scene. Rays walk the two-level structure (the lux_hip_two_level walk);
 * shading reads each scene ordinal's triangle in world space as
 * scene_triangle_view gives it (src/scene/scene_data.h:335-372): corners by
 * transform_point(object_to_world), normals by transform_instance_normal
 * (adjugate transform, normalised), material = the instance's override if
 * >= 0, else the mesh triangle's. Emissive ordinals become power-sampled area
 * lights with their world-space area (prepare_scene_light_sampling,
 * src/scene/scene_prepare.h:268-387). Mesh m owns vertices
 * [vertex_offsets[m], vertex_offsets[m+1]) and triangles
 * [triangle_offsets[m], triangle_offsets[m+1]); indices are mesh-local. */
typedef struct lux_instanced_scene_desc {
    const lux_vec3* vertices;
    const lux_vec3* normals;         /* per vertex or NULL (flat shading) */
    const int32_t* vertex_offsets;   /* mesh_count + 1 */
    const uint32_t* indices;         /* 3 per triangle, into the mesh's own vertices */
    const int32_t* material_ids;     /* per triangle */
    const int32_t* triangle_offsets; /* mesh_count + 1 */
    int32_t mesh_count;
    const lux_instance* instances;   /* mesh_id, material_override, both matrices; world_bounds is recomputed */
    int32_t instance_count;
    const lux_material* materials;
    int32_t material_count;
    float camera[8];
    const lux_light* lights; /* scene lights (may be NULL) */
    int32_t light_count;
} lux_instanced_scene_desc;
/* PreparedScene(scene) + prepare_gpu() for an instanced scene. Scene
 * validation (src/scene/scene_validate.h:95-147) rejects with
 * LUX_ERR_INVALID_ARGUMENT: indices or material ids out of range, a mesh id
 * or material override out of range, a transform pair that is not
 * affine-invertible and consistent; at least one instance must place a
 * triangle. lux_hip_scene_bvh and lux_hip_scene_update_vertices do not apply
 * to such a scene (LUX_ERR_INVALID_ARGUMENT). Synchronous. */
int lux_hip_scene_create_instanced(const lux_instanced_scene_desc* desc, lux_hip_scene** out);
/* Moves the instances of an instanced scene (same count; mesh ids and
 * overrides may change): the TLAS, the world-space triangles and normals and
 * the light table are rebuilt, the per-mesh BLASes are kept. It waits for the
 * device first, so renders enqueued earlier finish with the old placement; a
 * scene with admitted, unfinished server requests is refused
 * (LUX_ERR_INVALID_ARGUMENT). Synchronous. */
int lux_hip_scene_set_instances(lux_hip_scene* scene, const lux_instance* instances, int32_t instance_count);
/* The shading view the renderer holds for every scene ordinal, as
 * scene_triangle_view gives it (src/scene/scene_data.h:335-372): world corners
 * (9 floats per ordinal), world corner normals (9 floats; zeros for a mesh
 * without normals) and the material id after the instance override. Works for
 * one-mesh scenes too (identity instance). Arrays sized by the ordinal count
 * (lux_hip_scene_info's triangle_count); any may be NULL. Synchronous. */
int lux_hip_scene_triangles(lux_hip_scene* scene, float* corners, float* normals, int32_t* material_ids);

/* ---- OBJ import ----------------------------------------------------------
 * load_obj_into_scene (src/util/obj_loader.cpp:613-1311, contract
 * README.md:180-226) for the path tracer's instanced scenes: the file parsed
 * as the bundled tinyobjloader parses it with triangulation (quads along the
 * shorter diagonal, larger polygons by ear clipping), faces with invalid or
 * non-finite vertices and degenerate faces (area <= 1e-12) skipped, one vertex
 * per distinct (position, used normal, used texcoord) in first-use order,
 * normals per normal_mode, material ids from the usemtl bindings or the
 * default material, the boundary-edge mask per triangle, and the instance
 * transform object_to_world * placement (make_obj_placement_transform). The
 * mesh is object space; it enters a scene through lux_instanced_scene_desc
 * with that transform. Host code: needs no GPU and no lux_hip_init. */
enum { LUX_OBJ_NORMALS_SOURCE_OR_GEOMETRY = 0, LUX_OBJ_NORMALS_GENERATE_MISSING_SMOOTH = 1,
       LUX_OBJ_NORMALS_FACE_ONLY = 2 };
enum { LUX_OBJ_SCALE_NONE = 0, LUX_OBJ_SCALE_MAX_EXTENT = 1, LUX_OBJ_SCALE_X_EXTENT = 2 };
enum { LUX_OBJ_TOPOLOGY_SOURCE_VERTICES = 0, LUX_OBJ_TOPOLOGY_WELDED_POSITIONS = 1 };
/* ObjLoadOptions (src/util/obj_loader.h:52-67); lux_obj_default_options fills
 * the reference's defaults. */
typedef struct lux_obj_options {
    int32_t normal_mode;           /* LUX_OBJ_NORMALS_* */
    int32_t reverse_orientation;
    int32_t flip_texcoord_v;
    int32_t boundary_topology;     /* LUX_OBJ_TOPOLOGY_* */
    float position_weld_epsilon;
    int32_t default_material_id;   /* -1: a Lambert(0.8) appended after the scene's materials, if no bindings */
    int32_t scene_material_count;  /* materials already in the scene (the ids bindings may name) */
    int32_t placement_scale_mode;  /* LUX_OBJ_SCALE_* */
    float placement_target_extent;
    int32_t placement_center_xz;
    int32_t placement_ground_to_y;
    float placement_target_y;
    float object_to_world[16];     /* Transform applied after the placement (m, row-major) */
    float world_to_object[16];     /* its inverse */
    int32_t binding_count;         /* usemtl name -> scene material id */
    const char* const* binding_names;
    const int32_t* binding_material_ids;
    int32_t cap_boundary_holes;    /* fan-fill closed boundary loops (ObjLoadOptions::cap_boundary_holes) */
    int32_t cap_boundary_topology; /* LUX_OBJ_TOPOLOGY_* the caps are traced over (default: welded) */
} lux_obj_options;
typedef struct lux_obj_mesh lux_obj_mesh;

int lux_obj_default_options(lux_obj_options* options);
/* LUX_ERR_INVALID_ARGUMENT for what the reference loader reports as a failure
 * (unreadable file, a parse error, a usemtl without binding, bad options). A
 * file whose faces are all skipped loads as a mesh of 0 triangles. */
int lux_obj_load(const char* path, const lux_obj_options* options, lux_obj_mesh** out);
int lux_obj_mesh_free(lux_obj_mesh* mesh);
/* Counts: vertices, triangles, 1 if the default material was appended; skipped
 * faces (invalid index, non-finite vertex, degenerate). Any pointer may be NULL. */
int lux_obj_mesh_counts(const lux_obj_mesh* mesh, int32_t* vertex_count, int32_t* triangle_count,
                        int32_t* appended_default_material, int32_t* skipped_invalid, int32_t* skipped_nonfinite,
                        int32_t* skipped_degenerate);
/* Boundary-hole capping counts (ObjLoadResult, src/util/obj_loader.h:80-88):
 * cap triangles appended, boundary edges of the capping topology, boundary
 * vertices that are not on exactly one incoming and one outgoing edge (any
 * makes the loader skip capping), open loops, loops under 3 vertices and
 * cap triangles of area <= 1e-12 skipped. Any pointer may be NULL. */
int lux_obj_mesh_cap_counts(const lux_obj_mesh* mesh, int32_t* cap_triangles, int32_t* boundary_edges,
                            int32_t* non_simple_vertices, int32_t* incomplete_loops, int32_t* degenerate_loops,
                            int32_t* degenerate_triangles);
/* The mesh arrays (sized by the counts; any may be NULL) and the instance
 * transform (16 + 16 floats, row-major). */
int lux_obj_mesh_arrays(const lux_obj_mesh* mesh, lux_vec3* vertices, lux_vec3* normals, float* uvs,
                        uint32_t* indices, int32_t* material_ids, uint8_t* boundary_edges, float* object_to_world,
                        float* world_to_object);

/* ---- film output ---------------------------------------------------------
 * The files of write_ppm (src/film/ppm.h:37-56: 1/spp, Reinhard c/(1+c),
 * clamp to [0,1], 255.999 * c^(1/2.2) truncated, pixels in film order) and
 * write_pfm (src/film/pfm.h:27-44: 1/spp, rows bottom to top, "PF" header,
 * scale -1.0) from a film of per-pixel sums in device memory (W*H*3 floats,
 * the layout lux_hip_render_device writes). Byte-identical to the reference's
 * writers. The _device forms write only the pixel body (3 bytes or 3 floats per
 * pixel) and are stream-ordered; lux_hip_write_* write the whole file and
 * create its parent directories, as the reference does. */
int lux_hip_film_ppm_device(const float* d_film, int32_t width, int32_t height, int32_t samples_per_pixel,
                            uint8_t* d_rgb8, lux_stream stream);
int lux_hip_film_pfm_device(const float* d_film, int32_t width, int32_t height, int32_t samples_per_pixel,
                            float* d_rgb, lux_stream stream);
int lux_hip_write_ppm(const char* path, const float* d_film, int32_t width, int32_t height,
                      int32_t samples_per_pixel);
int lux_hip_write_pfm(const char* path, const float* d_film, int32_t width, int32_t height,
                      int32_t samples_per_pixel);
/* The 255 gamma thresholds of the PPM byte conversion (no reference
 * counterpart; host only): byte = number of thresholds <= the clamped value. */
int lux_hip_ppm_thresholds(float* out255);

/* ---- measurement --------------------------------------------------------
 * No reference counterpart (the reference only prints wall-clock stage times,
 * src/app/render_runner.cpp:274-286). While a probe is open, every launch of
 * the kernel named `kernel_name` is bracketed by HIP events recorded on the
 * stream it is launched on; lux_hip_probe_end synchronises and returns the
 * summed kernel time and the number of launches. */
int lux_hip_probe_begin(const char* kernel_name);
int lux_hip_probe_end(double* total_ms, int64_t* launches);
/* Runs the closest-hit and any-hit walks of the given device rays once more,
 * counting work: stats[6] = {closest: 32-byte nodes loaded, 0, triangle
 * tests; any-hit: the same three}.
 * Synchronous; no reference counterpart. Used to state the bytes a traversal
 * touches (plus 36 B per triangle test). */
int lux_hip_trace_stats_device(lux_hip_bvh* bvh, const lux_triangle* d_triangles,
                               int32_t triangle_count, const lux_ray* d_rays,
                               int32_t ray_count, uint64_t* stats);

#ifdef __cplusplus
}
#endif

#endif /* LUX_HIP_H */

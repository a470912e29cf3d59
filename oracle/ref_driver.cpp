// ref_driver.cpp -- extern "C" shim over the reference's OWN host code, compiled
// from the LUX sources where they lie (see oracle/ref.mk). Nothing here
// re-implements an algorithm: every result comes from the reference's headers
// (build_cpu_lbvh_tree, intersect_cpu_bvh_tree, intersect_triangle,
// lbvh_detail::centroid_to_morton). It exists only to produce golden vectors
// (tests/golden/make_golden.py) and to time the reference CPU path as a
// baseline. TEST INFRASTRUCTURE ONLY.

#include <cstdint>
#include <cstring>
#include <vector>

#include "bvh/cpu_bvh_traversal.h"
#include "bvh/cpu_lbvh_builder.h"
#include "geometry/triangle.h"
#include "scene/scene_intersection.h"

static_assert(sizeof(AABB) == 24 && sizeof(Ray) == 36 && sizeof(Triangle) == 36, "layout");

extern "C" {

struct ref_tree {
    CpuBvhTree tree;
};

// build_cpu_lbvh_tree over primitives {bounds, bounds.centroid(), id} -- the
// record collect_scene_mesh_bvh_build_primitives makes for a triangle.
ref_tree* ref_lbvh_build(const float* boxes, const int32_t* ids, int32_t n) {
    const AABB* b = reinterpret_cast<const AABB*>(boxes);
    std::vector<BvhBuildPrimitive> prims;
    prims.reserve(n);
    for (int i = 0; i < n; ++i) {
        BvhBuildPrimitive p{};
        p.bounds = b[i];
        p.centroid = b[i].centroid();
        p.primitive_id = ids[i];
        p.instance_id = -1;
        p.mesh_id = 0;
        p.local_primitive_id = ids[i];
        p.primitive_kind = PrimitiveKind::Triangle;
        prims.push_back(p);
    }
    ref_tree* t = new ref_tree();
    t->tree = build_cpu_lbvh_tree(std::move(prims));
    return t;
}

void ref_lbvh_free(ref_tree* t) { delete t; }

int32_t ref_lbvh_node_count(const ref_tree* t) { return int32_t(t->tree.nodes.size()); }

// nodes: per node {bounds(6 floats)}, left/right/first/count ints; leaf_ids: n
void ref_lbvh_export(const ref_tree* t, float* bounds, int32_t* left, int32_t* right,
                     int32_t* first_primitive, int32_t* primitive_count, int32_t* leaf_ids) {
    const auto& nodes = t->tree.nodes;
    for (size_t i = 0; i < nodes.size(); ++i) {
        std::memcpy(bounds + 6 * i, &nodes[i].bounds, sizeof(AABB));
        left[i] = nodes[i].left;
        right[i] = nodes[i].right;
        first_primitive[i] = nodes[i].first_primitive;
        primitive_count[i] = nodes[i].primitive_count;
    }
    for (size_t i = 0; i < t->tree.primitive_ids.size(); ++i) leaf_ids[i] = t->tree.primitive_ids[i];
}

void ref_morton_codes(const float* boxes, int32_t n, const float* root, uint32_t* out) {
    const AABB* b = reinterpret_cast<const AABB*>(boxes);
    AABB rb;
    std::memcpy(&rb, root, sizeof(AABB));
    for (int i = 0; i < n; ++i) out[i] = lbvh_detail::centroid_to_morton(rb, b[i].centroid());
}

// intersect_cpu_bvh_tree with the scene's triangle leaf intersector
// (intersect_scene_triangle_primitive). Returns the number of stack overflows.
int32_t ref_trace_closest(const ref_tree* t, const float* tris, const float* rays, int32_t n,
                          int32_t* ids, float* t_out, float* u_out, float* v_out) {
    const Triangle* tr = reinterpret_cast<const Triangle*>(tris);
    const Ray* rs = reinterpret_cast<const Ray*>(rays);
    int32_t overflows = 0;
    for (int i = 0; i < n; ++i) {
        int id = -1;
        Float th, uh, vh;
        bool ov = false;
        bool found = intersect_cpu_bvh_tree(
            t->tree, rs[i],
            [tr](int prim, const Ray& ray, CpuBvhPrimitiveHit& hit) {
                return intersect_scene_triangle_primitive(tr[prim], ray, hit);
            },
            id, th, uh, vh, &ov);
        if (ov) ++overflows;
        ids[i] = found ? id : -1;
        t_out[i] = found ? th : INFINITY;
        u_out[i] = found ? uh : 0;
        v_out[i] = found ? vh : 0;
    }
    return overflows;
}

// The loop of intersect_scene_bruteforce over a plain triangle array.
void ref_brute_closest(const float* tris, int32_t tri_count, const float* rays, int32_t n,
                       int32_t* ids, float* t_out) {
    const Triangle* tr = reinterpret_cast<const Triangle*>(tris);
    const Ray* rs = reinterpret_cast<const Ray*>(rays);
    for (int i = 0; i < n; ++i) {
        Float best = INFINITY;
        int id = -1;
        for (int k = 0; k < tri_count; ++k) {
            Float tt, uu, vv;
            if (!intersect_triangle(rs[i], tr[k], tt, uu, vv)) continue;
            if (tt >= best) continue;
            best = tt;
            id = k;
        }
        ids[i] = id;
        t_out[i] = best;
    }
}

void ref_triangle_bounds(const float* tris, int32_t n, float* boxes) {
    const Triangle* tr = reinterpret_cast<const Triangle*>(tris);
    for (int i = 0; i < n; ++i) {
        AABB b = triangle_bounds(tr[i]);
        std::memcpy(boxes + 6 * i, &b, sizeof(AABB));
    }
}

int32_t ref_aabb_overlaps(const float* a, const float* b) {
    AABB x, y;
    std::memcpy(&x, a, sizeof(AABB));
    std::memcpy(&y, b, sizeof(AABB));
    return x.overlaps(y) ? 1 : 0;
}

// Self pairs (ids[i], ids[j]) with ids[j] > ids[i] whose boxes overlap under the
// reference's AABB::overlaps -- the O(n^2) loop of validate_gpu_bvh_self_overlaps.
int64_t ref_overlap_self_brute(const float* boxes, int32_t n, int32_t* pairs, int64_t cap) {
    const AABB* b = reinterpret_cast<const AABB*>(boxes);
    int64_t count = 0;
    for (int i = 0; i < n; ++i)
        for (int j = i + 1; j < n; ++j) {
            if (!b[i].overlaps(b[j])) continue;
            if (count < cap) {
                pairs[2 * count] = i;
                pairs[2 * count + 1] = j;
            }
            ++count;
        }
    return count;
}

// Cross pairs (left index, right index) under AABB::overlaps.
int64_t ref_overlap_cross_brute(const float* lboxes, int32_t nl, const float* rboxes, int32_t nr,
                                int32_t* pairs, int64_t cap) {
    const AABB* lb = reinterpret_cast<const AABB*>(lboxes);
    const AABB* rb = reinterpret_cast<const AABB*>(rboxes);
    int64_t count = 0;
    for (int j = 0; j < nr; ++j)
        for (int i = 0; i < nl; ++i) {
            if (!lb[i].overlaps(rb[j])) continue;
            if (count < cap) {
                pairs[2 * count] = i;
                pairs[2 * count + 1] = j;
            }
            ++count;
        }
    return count;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Path tracer: the reference's own CPU path tracer (render_cpu_path_tracer /
// trace_path, src/integrator/cpu_path_tracer.cpp) on a scene given as arrays.
// The scene is one mesh + one identity instance, built through the reference's
// own scene API (add_scene_mesh / add_scene_instance / PreparedScene).
// ---------------------------------------------------------------------------
#include <unistd.h>

#include <cstdio>
#include <functional>
#include <memory>

#include "camera/pinhole_camera.h"
#include "integrator/cpu_path_tracer.h"
#include "integrator/render_context.h"
#include "scene/prepared_scene.h"
#include "scene/scene_build.h"
#include "scene/scene_edit.h"
#include "scene/scene_light_sampler.h"
#include "scene/direct_light_test.h"
#include "scene/lux_cover.h"
#include "scene/stanford_bunny.h"
#include "scene/transparent_filter.h"

extern "C" {

struct ref_pt {
    std::unique_ptr<PreparedScene> prepared;
};

// Scene lights for the next ref_pt_create* call (kind 1 point, 2 distant, 3
// uniform infinite, 4 image infinite; vec = position / direction, color =
// intensity / radiance / image scale; image texels back to back), added with
// add_point_light / add_distant_light / add_uniform_infinite_light /
// add_image_infinite_light.
static std::vector<std::function<void(Scene&)>> g_pending_lights;

void ref_pt_set_scene_lights(int32_t n, const int32_t* kind, const float* vec, const float* color,
                             const int32_t* image_wh, const float* pixels) {
    g_pending_lights.clear();
    size_t texel = 0;
    for (int i = 0; i < n; ++i) {
        const vec3 v(vec[3 * i], vec[3 * i + 1], vec[3 * i + 2]);
        const vec3 c(color[3 * i], color[3 * i + 1], color[3 * i + 2]);
        switch (kind[i]) {
            case 1: g_pending_lights.push_back([v, c](Scene& s) { add_point_light(s, v, c); }); break;
            case 2: g_pending_lights.push_back([v, c](Scene& s) { add_distant_light(s, v, c); }); break;
            case 3: g_pending_lights.push_back([c](Scene& s) { add_uniform_infinite_light(s, c); }); break;
            case 4: {
                const int w = image_wh[2 * i], h = image_wh[2 * i + 1];
                std::vector<vec3> px(size_t(w) * h);
                for (size_t k = 0; k < px.size(); ++k, ++texel)
                    px[k] = vec3(pixels[3 * texel], pixels[3 * texel + 1], pixels[3 * texel + 2]);
                g_pending_lights.push_back([px, w, h, c](Scene& s) { add_image_infinite_light(s, px, w, h, c); });
                break;
            }
        }
    }
}

static void add_pending_lights(Scene& scene) {
    for (auto& f : g_pending_lights) f(scene);
    g_pending_lights.clear();
}

// materials: m x 11 floats {type, albedo rgb, emission rgb, two_sided, roughness,
// metallic, eta}; camera: {origin xyz, look_at xyz, vertical fov degrees, aspect}
ref_pt* ref_pt_create(const float* vertices, int32_t nv, const uint32_t* indices, const float* normals,
                      const float* uvs, const int32_t* material_ids, int32_t nt, const float* materials,
                      int32_t nm, const float* camera) {
    Scene scene;
    for (int m = 0; m < nm; ++m) {
        const float* p = materials + 11 * m;
        Material mat(static_cast<MaterialType>(int(p[0])), vec3(p[1], p[2], p[3]), vec3(p[4], p[5], p[6]), p[8],
                     p[9], p[10],
                     p[7] != 0 ? EmissionSidedness::DoubleSided : EmissionSidedness::SingleSided);
        scene.materials.push_back(mat);
    }
    SceneMesh mesh;
    mesh.vertices.resize(nv);
    mesh.normals.assign(nv, vec3(0));
    mesh.uvs.assign(nv, vec2(0));
    for (int i = 0; i < nv; ++i) {
        mesh.vertices[i] = vec3(vertices[3 * i], vertices[3 * i + 1], vertices[3 * i + 2]);
        if (normals) mesh.normals[i] = vec3(normals[3 * i], normals[3 * i + 1], normals[3 * i + 2]);
        if (uvs) mesh.uvs[i] = vec2(uvs[2 * i], uvs[2 * i + 1]);
    }
    mesh.indices.resize(nt);
    mesh.material_ids.resize(nt);
    mesh.boundary_edges.assign(nt, 0);
    for (int t = 0; t < nt; ++t) {
        mesh.indices[t] = TriangleIndices{indices[3 * t], indices[3 * t + 1], indices[3 * t + 2]};
        mesh.material_ids[t] = material_ids[t];
    }
    int mesh_id = add_scene_mesh(scene, std::move(mesh));
    add_scene_instance(scene, mesh_id, Transform{});
    add_pending_lights(scene);
    scene.camera = make_pinhole_camera(vec3(camera[0], camera[1], camera[2]), vec3(camera[3], camera[4], camera[5]),
                                       camera[6], camera[7]);
    ref_pt* r = new ref_pt();
    r->prepared = std::make_unique<PreparedScene>(std::move(scene));
    r->prepared->prepare_host_bvh();
    return r;
}

// Several meshes placed by affine instances (add_scene_mesh per mesh,
// add_scene_instance(make_transform(M), override) per instance): mesh m owns
// vertices [vert_off[m], vert_off[m+1]) and triangles [tri_off[m], tri_off[m+1])
// with mesh-local indices; normals / uvs per vertex (nullable).
ref_pt* ref_pt_create_instanced(const float* vertices, const int32_t* vert_off, const uint32_t* indices,
                                const int32_t* tri_off, const float* normals, const float* uvs,
                                const int32_t* material_ids, int32_t mesh_count, const int32_t* inst_mesh,
                                const float* inst_m, const int32_t* inst_override, int32_t instance_count,
                                const float* materials, int32_t nm, const float* camera) {
    Scene scene;
    for (int m = 0; m < nm; ++m) {
        const float* p = materials + 11 * m;
        Material mat(static_cast<MaterialType>(int(p[0])), vec3(p[1], p[2], p[3]), vec3(p[4], p[5], p[6]), p[8],
                     p[9], p[10],
                     p[7] != 0 ? EmissionSidedness::DoubleSided : EmissionSidedness::SingleSided);
        scene.materials.push_back(mat);
    }
    for (int m = 0; m < mesh_count; ++m) {
        SceneMesh mesh;
        const int v0 = vert_off[m], nv = vert_off[m + 1] - vert_off[m];
        const int t0 = tri_off[m], nt = tri_off[m + 1] - tri_off[m];
        mesh.vertices.resize(nv);
        mesh.normals.assign(nv, vec3(0));
        mesh.uvs.assign(nv, vec2(0));
        for (int i = 0; i < nv; ++i) {
            const int g = v0 + i;
            mesh.vertices[i] = vec3(vertices[3 * g], vertices[3 * g + 1], vertices[3 * g + 2]);
            if (normals) mesh.normals[i] = vec3(normals[3 * g], normals[3 * g + 1], normals[3 * g + 2]);
            if (uvs) mesh.uvs[i] = vec2(uvs[2 * g], uvs[2 * g + 1]);
        }
        mesh.indices.resize(nt);
        mesh.material_ids.resize(nt);
        mesh.boundary_edges.assign(nt, 0);
        for (int t = 0; t < nt; ++t) {
            mesh.indices[t] =
                TriangleIndices{indices[3 * (t0 + t)], indices[3 * (t0 + t) + 1], indices[3 * (t0 + t) + 2]};
            mesh.material_ids[t] = material_ids[t0 + t];
        }
        add_scene_mesh(scene, std::move(mesh));
    }
    for (int i = 0; i < instance_count; ++i) {
        mat4 M;
        for (int r = 0; r < 4; ++r)
            for (int c = 0; c < 4; ++c) M.m[r][c] = inst_m[16 * i + 4 * r + c];
        add_scene_instance(scene, inst_mesh[i], make_transform(M), inst_override[i]);
    }
    add_pending_lights(scene);
    scene.camera = make_pinhole_camera(vec3(camera[0], camera[1], camera[2]), vec3(camera[3], camera[4], camera[5]),
                                       camera[6], camera[7]);
    ref_pt* r = new ref_pt();
    r->prepared = std::make_unique<PreparedScene>(std::move(scene));
    r->prepared->prepare_host_bvh();
    return r;
}

// The reference's own "bunny" scene: make_stanford_bunny_scene(asset_root)
// with the camera of make_bunny_scene_from_config (src/app/render_runner.cpp:148-154).
ref_pt* ref_pt_create_bunny(const char* asset_root, float aspect) {
    Scene scene = make_stanford_bunny_scene(std::filesystem::path(asset_root));
    scene.camera = make_pinhole_camera(vec3(0.0f, 0.3f, 0.5f), vec3(-0.02f, 0.11f, 0.0f), 40.0f, aspect);
    ref_pt* r = new ref_pt();
    r->prepared = std::make_unique<PreparedScene>(std::move(scene));
    r->prepared->prepare_host_bvh();
    return r;
}

// The reference CLI's built-in scenes with their CLI cameras
// (src/app/render_runner.cpp:92-170): "direct-light", "transparent-filter",
// "bunny", "lux-cover" (the last two read assets under asset_root).
ref_pt* ref_pt_create_builtin(const char* name, const char* asset_root, float aspect) {
    const std::string n(name);
    const std::filesystem::path root(asset_root);
    Scene scene;
    if (n == "direct-light") {
        scene = make_direct_light_test_scene();
        scene.camera = make_pinhole_camera(vec3(0, 1.2f, 3.0f), vec3(0, 0, 0), 35.0f, aspect);
    } else if (n == "transparent-filter") {
        scene = make_transparent_filter_scene();
        scene.camera = make_pinhole_camera(vec3(0, 0.85f, 3.2f), vec3(0, 0.65f, -0.7f), 38.0f, aspect);
    } else if (n == "bunny") {
        scene = make_stanford_bunny_scene(root);
        scene.camera = make_pinhole_camera(vec3(0.0f, 0.3f, 0.5f), vec3(-0.02f, 0.11f, 0.0f), 40.0f, aspect);
    } else if (n == "lux-cover") {
        scene = make_lux_cover_scene(root);
        scene.camera = make_pinhole_camera(vec3(0.0f, 0.52f, -3.0f), vec3(0.0f, 0.36f, 0.0f), 32.0f, aspect);
    } else {
        return nullptr;
    }
    ref_pt* r = new ref_pt();
    r->prepared = std::make_unique<PreparedScene>(std::move(scene));
    r->prepared->prepare_host_bvh();
    return r;
}

// The scene as arrays, meshes and instances as the reference holds them:
// counts {vertices, triangles, meshes, instances, materials}; per vertex xyz
// and normal; mesh vertex / triangle offsets; mesh-local indices, material ids;
// per instance mesh, Transform m and inv (row-major), override; materials as
// 11 floats {type, albedo, emission, two_sided, roughness, metallic, eta};
// camera {origin, lower_left, horizontal, vertical}. Call with null arrays
// first for the counts.
void ref_pt_export(const ref_pt* r, int32_t* counts5, float* vertices, float* normals, int32_t* vert_off,
                   uint32_t* indices, int32_t* mat_ids, int32_t* tri_off, int32_t* inst_mesh, float* inst_m,
                   float* inst_inv, int32_t* inst_override, float* materials, float* camera12) {
    const Scene& sc = r->prepared->host_scene();
    int nv = 0, nt = 0;
    for (const SceneMesh& m : sc.mesh_assets) {
        nv += int(m.vertices.size());
        nt += int(m.indices.size());
    }
    counts5[0] = nv;
    counts5[1] = nt;
    counts5[2] = int(sc.mesh_assets.size());
    counts5[3] = int(sc.instances.size());
    counts5[4] = int(sc.materials.size());
    if (!vertices) return;
    int v = 0, t = 0;
    vert_off[0] = tri_off[0] = 0;
    for (size_t k = 0; k < sc.mesh_assets.size(); ++k) {
        const SceneMesh& m = sc.mesh_assets[k];
        for (size_t i = 0; i < m.vertices.size(); ++i, ++v) {
            std::memcpy(vertices + 3 * v, &m.vertices[i], 12);
            std::memcpy(normals + 3 * v, &m.normals[i], 12);
        }
        for (size_t i = 0; i < m.indices.size(); ++i, ++t) {
            indices[3 * t] = m.indices[i].v0;
            indices[3 * t + 1] = m.indices[i].v1;
            indices[3 * t + 2] = m.indices[i].v2;
            mat_ids[t] = m.material_ids[i];
        }
        vert_off[k + 1] = v;
        tri_off[k + 1] = t;
    }
    for (size_t i = 0; i < sc.instances.size(); ++i) {
        const SceneInstance& in = sc.instances[i];
        inst_mesh[i] = in.mesh_id;
        inst_override[i] = in.material_override;
        for (int a = 0; a < 4; ++a)
            for (int b = 0; b < 4; ++b) {
                inst_m[16 * i + 4 * a + b] = in.object_to_world.m.m[a][b];
                inst_inv[16 * i + 4 * a + b] = in.object_to_world.inv.m[a][b];
            }
    }
    for (size_t i = 0; i < sc.materials.size(); ++i) {
        const Material& m = sc.materials[i];
        const float row[11] = {float(int(m.type)), m.albedo.x, m.albedo.y, m.albedo.z, m.emission.x, m.emission.y,
                               m.emission.z, m.emission_sidedness == EmissionSidedness::DoubleSided ? 1.0f : 0.0f,
                               m.roughness, m.metallic, m.eta};
        std::memcpy(materials + 11 * i, row, sizeof(row));
    }
    const PerspectiveCamera& c = sc.camera.perspective;
    const vec3* cv[4] = {&c.origin, &c.lower_left, &c.horizontal, &c.vertical};
    for (int k = 0; k < 4; ++k) std::memcpy(camera12 + 3 * k, cv[k], 12);
}

// scene_triangle_view of every scene ordinal: world corners (9 floats), world
// corner normals (9 floats) and the material after the override. Returns the
// ordinal count (call with null arrays first).
int32_t ref_pt_triangles(const ref_pt* r, float* corners, float* normals, int32_t* material) {
    const Scene& scene = r->prepared->host_scene();
    const int n = scene_triangle_count(scene);
    if (!corners) return n;
    for (int t = 0; t < n; ++t) {
        const SceneTriangle st = scene_triangle_view(scene, t);
        const vec3 c[3] = {st.triangle.v0, st.triangle.v1, st.triangle.v2};
        const vec3 nn[3] = {st.n0, st.n1, st.n2};
        for (int k = 0; k < 3; ++k) {
            corners[9 * t + 3 * k] = c[k].x;
            corners[9 * t + 3 * k + 1] = c[k].y;
            corners[9 * t + 3 * k + 2] = c[k].z;
            normals[9 * t + 3 * k] = nn[k].x;
            normals[9 * t + 3 * k + 1] = nn[k].y;
            normals[9 * t + 3 * k + 2] = nn[k].z;
        }
        material[t] = st.material_id;
    }
    return n;
}

void ref_pt_free(ref_pt* r) { delete r; }

// {origin, lower_left, horizontal, vertical, lens_u, lens_v} xyz, lens_radius, focal_distance
void ref_pt_camera(const ref_pt* r, float* out) {
    const PerspectiveCamera& c = r->prepared->host_scene().camera.perspective;
    const vec3* v[6] = {&c.origin, &c.lower_left, &c.horizontal, &c.vertical, &c.lens_u, &c.lens_v};
    for (int k = 0; k < 6; ++k) {
        out[3 * k] = v[k]->x;
        out[3 * k + 1] = v[k]->y;
        out[3 * k + 2] = v[k]->z;
    }
    out[18] = c.lens_radius;
    out[19] = c.focal_distance;
}

// Light selection table: per light {primitive ordinal, area, power cdf, alias
// probability, alias index, pmf}; returns the light count (arrays sized by the
// caller, call with null first). primitive_light_index: one int per triangle.
int32_t ref_pt_lights(const ref_pt* r, int32_t* prim, float* area, float* cdf, float* prob, int32_t* alias,
                      float* pmf, int32_t* primitive_light_index, float* total_weight) {
    const Scene& scene = r->prepared->host_scene();
    const PreparedLightSampling& ls = r->prepared->light_sampling();
    const LightSelectionTable& sel = ls.light_selection;
    int count = int(sel.lights.size());
    if (!prim) return count;
    for (int i = 0; i < count; ++i) {
        prim[i] = scene_primitive_index(scene, sel.lights[i].area.primitive_ref);
        area[i] = sel.lights[i].area.area;
        cdf[i] = sel.power_distribution.cdf[i];
        prob[i] = sel.power_distribution.alias_table[i].probability;
        alias[i] = sel.power_distribution.alias_table[i].alias;
        pmf[i] = sel.power_distribution.alias_table[i].pmf;
    }
    *total_weight = sel.power_distribution.total_weight;
    const auto& lookup = ls.surface_lookup.primitive_light_indices;
    for (size_t t = 0; t < lookup.size(); ++t) primitive_light_index[t] = lookup[t];
    return count;
}

// light_sampler: LightSamplerKind as an int (0 Uniform, 1 Power, 2 Bvh).
static RenderSettings make_settings(int32_t depth, uint64_t seed, int32_t frame, int32_t sample_offset,
                                    int32_t light_sampler) {
    RenderSettings s;
    s.max_depth = depth;
    s.seed = seed;
    s.frame_index = frame;
    s.sample_offset = sample_offset;
    s.sampler_kind = SamplerKind::Independent;
    s.light_sampler_kind = static_cast<LightSamplerKind>(light_sampler);
    return s;
}

// The light BVH of the scene's light sampling (LightBvhTable): node count, and
// per node {min xyz, max xyz}, power, {left, right, parent, light index}; the
// leaf node of each light; the root.
int32_t ref_pt_light_bvh(const ref_pt* r, float* bounds6, float* power, int32_t* links4, int32_t* leaf,
                         int32_t* root) {
    const LightBvhTable& t = r->prepared->light_sampling().light_bvh;
    const int n = int(t.nodes.size());
    if (!bounds6) return n;
    for (int i = 0; i < n; ++i) {
        const LightBvhNode& x = t.nodes[i];
        const float b[6] = {x.bounds.p_min.x, x.bounds.p_min.y, x.bounds.p_min.z,
                            x.bounds.p_max.x, x.bounds.p_max.y, x.bounds.p_max.z};
        for (int k = 0; k < 6; ++k) bounds6[6 * i + k] = b[k];
        power[i] = x.power;
        links4[4 * i] = x.left;
        links4[4 * i + 1] = x.right;
        links4[4 * i + 2] = x.parent;
        links4[4 * i + 3] = x.light_index;
    }
    for (size_t i = 0; i < t.leaf_nodes_by_light_index.size(); ++i) leaf[i] = t.leaf_nodes_by_light_index[i];
    *root = t.valid() ? t.root : -1;
    return n;
}

// render_cpu_path_tracer itself; film_sum receives the per-pixel SUM over spp
// (Film::pixels, row-major, y = 0 at the top). Its row printout is silenced.
void ref_pt_render(const ref_pt* r, int32_t w, int32_t h, int32_t spp, int32_t depth, uint64_t seed,
                   int32_t frame, int32_t sample_offset, int32_t light_sampler, float* film_sum) {
    Film film(w, h, spp);
    RenderSettings s = make_settings(depth, seed, frame, sample_offset, light_sampler);
    std::fflush(stdout);
    int saved = dup(1);
    FILE* devnull = std::fopen("/dev/null", "w");
    if (devnull) dup2(fileno(devnull), 1);
    render_cpu_path_tracer(r->prepared->host_scene(), r->prepared->host_bvh(), r->prepared->light_sampling(),
                           film, s, nullptr);
    std::fflush(stdout);
    if (saved >= 0) {
        dup2(saved, 1);
        close(saved);
    }
    if (devnull) std::fclose(devnull);
    for (int i = 0; i < w * h; ++i) {
        film_sum[3 * i] = film.pixels[i].x;
        film_sum[3 * i + 1] = film.pixels[i].y;
        film_sum[3 * i + 2] = film.pixels[i].z;
    }
}

// trace_path radiance of individual (pixel, sample) pairs, with the camera
// sample drawn exactly as render_cpu_path_tracer draws it.
void ref_pt_samples(const ref_pt* r, int32_t w, int32_t h, int32_t depth, uint64_t seed, int32_t frame,
                    int32_t sample_offset, int32_t light_sampler, const int32_t* pixels, const int32_t* samples,
                    int32_t n, float* out) {
    const Scene& scene = r->prepared->host_scene();
    RenderSettings s = make_settings(depth, seed, frame, sample_offset, light_sampler);
    SceneLightSampler lights =
        make_scene_light_sampler(scene, r->prepared->light_sampling(), s.light_sampler_kind);
    for (int i = 0; i < n; ++i) {
        int pixel = pixels[i];
        int x = pixel % w, y = pixel / w;
        SamplerState sampler = make_render_sampler(s, pixel, samples[i]);
        Float jx = sampler_get_1d(sampler);
        Float jy = sampler_get_1d(sampler);
        CameraSample cs;
        cs.p_film = vec2((Float(x) + jx) / Float(w), (Float(h - 1 - y) + jy) / Float(h));
        cs.p_lens = sampler_get_2d(sampler);
        cs.time = sampler_get_1d(sampler);
        CameraRay cr = generate_camera_ray(scene.camera, cs);
        vec3 c(0);
        if (cr.valid) c = trace_path(scene, r->prepared->host_bvh(), lights, cr.ray, sampler, depth);
        out[3 * i] = c.x;
        out[3 * i + 1] = c.y;
        out[3 * i + 2] = c.z;
    }
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Two-level scenes: meshes and affine instances built through the reference's
// scene API (add_scene_mesh, add_scene_instance with make_transform, so the
// inverse comes from inverse_affine and the world bounds from transform_bounds)
// and traced with its host two-level walk (intersect_scene_bvh ->
// intersect_scene_two_level_bvh over PreparedScene::host_bvh).
// ---------------------------------------------------------------------------
extern "C" {

struct ref_tl {
    std::unique_ptr<PreparedScene> prepared;
};

// Mesh m has vertices [vert_off[m], vert_off[m+1]) (xyz triples) and triangles
// [tri_off[m], tri_off[m+1]) (mesh-local vertex indices, 3 each). Instance i
// uses mesh inst_mesh[i], the row-major 4x4 matrix inst_m[16 i..] as
// object_to_world, and material override inst_override[i].
ref_tl* ref_tl_create(const float* vertices, const int32_t* vert_off, const uint32_t* indices,
                      const int32_t* tri_off, int32_t mesh_count, const int32_t* inst_mesh, const float* inst_m,
                      const int32_t* inst_override, int32_t instance_count) {
    Scene scene;
    scene.materials.push_back(Material(static_cast<MaterialType>(0), vec3(0.5f), vec3(0), 0.5f, 0, 1.5f,
                                       EmissionSidedness::SingleSided));
    for (int m = 0; m < mesh_count; ++m) {
        SceneMesh mesh;
        const int v0 = vert_off[m], nv = vert_off[m + 1] - vert_off[m];
        const int t0 = tri_off[m], nt = tri_off[m + 1] - tri_off[m];
        mesh.vertices.resize(nv);
        mesh.normals.assign(nv, vec3(0));
        mesh.uvs.assign(nv, vec2(0));
        for (int i = 0; i < nv; ++i)
            mesh.vertices[i] = vec3(vertices[3 * (v0 + i)], vertices[3 * (v0 + i) + 1], vertices[3 * (v0 + i) + 2]);
        mesh.indices.resize(nt);
        mesh.material_ids.assign(nt, 0);
        mesh.boundary_edges.assign(nt, 0);
        for (int t = 0; t < nt; ++t)
            mesh.indices[t] =
                TriangleIndices{indices[3 * (t0 + t)], indices[3 * (t0 + t) + 1], indices[3 * (t0 + t) + 2]};
        add_scene_mesh(scene, std::move(mesh));
    }
    for (int i = 0; i < instance_count; ++i) {
        mat4 M;
        for (int r = 0; r < 4; ++r)
            for (int c = 0; c < 4; ++c) M.m[r][c] = inst_m[16 * i + 4 * r + c];
        add_scene_instance(scene, inst_mesh[i], make_transform(M), inst_override[i]);
    }
    scene.camera = make_pinhole_camera(vec3(0, 0, 3), vec3(0), 40.0f, 1.0f);
    ref_tl* r = new ref_tl();
    r->prepared = std::make_unique<PreparedScene>(std::move(scene));
    r->prepared->prepare_host_bvh();
    return r;
}

void ref_tl_free(ref_tl* r) { delete r; }

// Per instance: Transform::inv (16 floats, row-major), world_bounds (6 floats)
// and the scene ordinal of its local triangle 0 (-1 for an empty mesh).
void ref_tl_instances(const ref_tl* r, float* inv, float* world_bounds, int32_t* first_primitive) {
    const Scene& scene = r->prepared->host_scene();
    for (int i = 0; i < int(scene.instances.size()); ++i) {
        const SceneInstance& in = scene.instances[i];
        for (int a = 0; a < 4; ++a)
            for (int b = 0; b < 4; ++b) inv[16 * i + 4 * a + b] = in.object_to_world.inv.m[a][b];
        std::memcpy(world_bounds + 6 * i, &in.world_bounds, sizeof(AABB));
        first_primitive[i] = scene_primitive_index(scene, PrimitiveRef{i, in.mesh_id, 0, PrimitiveKind::Triangle});
    }
}

// intersect_scene_bvh per ray, or intersect_scene_bvh_skip_ref when ignore[i]
// >= 0 (a scene ordinal). Returns the number of stack overflows.
int32_t ref_tl_trace(const ref_tl* r, const float* rays, const int32_t* ignore, int32_t n, int32_t* prim, float* t,
                     float* u, float* v, int32_t* instance, int32_t* local) {
    const Scene& scene = r->prepared->host_scene();
    const CpuBvh& bvh = r->prepared->host_bvh();
    const Ray* rs = reinterpret_cast<const Ray*>(rays);
    int32_t overflows = 0;
    for (int i = 0; i < n; ++i) {
        SurfaceHit hit;
        bool ov = false;
        bool ok = (ignore && ignore[i] >= 0)
                      ? intersect_scene_bvh_skip_ref(scene, bvh, rs[i], scene_primitive_ref(scene, ignore[i]), hit, &ov)
                      : intersect_scene_bvh(scene, bvh, rs[i], hit, &ov);
        if (ov) ++overflows;
        prim[i] = ok ? hit.triangle_id : -1;
        t[i] = ok ? hit.t : INFINITY;
        u[i] = ok ? hit.u : 0;
        v[i] = ok ? hit.v : 0;
        instance[i] = ok ? hit.primitive_ref.instance_id : -1;
        local[i] = ok ? hit.primitive_ref.local_primitive_id : -1;
    }
    return overflows;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Film output: the reference's own write_ppm / write_pfm (src/film/ppm.h,
// src/film/pfm.h) on a film given as W*H*3 float sums.
// ---------------------------------------------------------------------------
#include "film/pfm.h"
#include "film/ppm.h"

extern "C" {

static std::vector<vec3> film_pixels(const float* film, int32_t w, int32_t h) {
    std::vector<vec3> px(size_t(w) * size_t(h));
    for (size_t i = 0; i < px.size(); ++i) px[i] = vec3(film[3 * i], film[3 * i + 1], film[3 * i + 2]);
    return px;
}

void ref_write_ppm(const char* path, const float* film, int32_t w, int32_t h, int32_t spp) {
    write_ppm(path, film_pixels(film, w, h), w, h, spp);
}

void ref_write_pfm(const char* path, const float* film, int32_t w, int32_t h, int32_t spp) {
    write_pfm(path, film_pixels(film, w, h), w, h, spp);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// OBJ import: the reference's own load_obj_into_scene_with_result
// (src/util/obj_loader.cpp, bundled tinyobjloader) into a scene that already
// holds `scene_material_count` materials; the appended mesh and its instance
// transform are read back.
// ---------------------------------------------------------------------------
#include "util/obj_loader.h"

extern "C" {

struct ref_obj {
    Scene scene;
    ObjLoadResult result;
    bool ok = false;
};

// Capping options for the next ref_obj_load (kept out of its argument list).
static int32_t g_obj_cap = 0, g_obj_cap_topology = 1;
void ref_obj_set_cap(int32_t cap, int32_t topology) {
    g_obj_cap = cap;
    g_obj_cap_topology = topology;
}

ref_obj* ref_obj_load(const char* path, int32_t normal_mode, int32_t reverse, int32_t flip_v,
                      int32_t boundary_topology, float weld_epsilon, int32_t default_material_id,
                      int32_t scene_material_count, int32_t scale_mode, float target_extent, int32_t center_xz,
                      int32_t ground_to_y, float target_y, const float* object_to_world, int32_t binding_count,
                      const char* const* binding_names, const int32_t* binding_ids) {
    ref_obj* r = new ref_obj();
    for (int m = 0; m < scene_material_count; ++m) r->scene.materials.push_back(Material::lambert(vec3(0.5f)));
    ObjLoadOptions o;
    o.normal_mode = static_cast<ObjNormalMode>(normal_mode);
    o.reverse_orientation = reverse != 0;
    o.flip_texcoord_v = flip_v != 0;
    o.boundary_topology = static_cast<ObjBoundaryTopology>(boundary_topology);
    o.position_weld_epsilon = weld_epsilon;
    o.default_material_id = default_material_id;
    o.placement.scale_mode = static_cast<ObjPlacementScaleMode>(scale_mode);
    o.placement.target_extent = target_extent;
    o.placement.center_xz = center_xz != 0;
    o.placement.ground_to_y = ground_to_y != 0;
    o.placement.target_y = target_y;
    if (object_to_world) {
        mat4 M;
        for (int a = 0; a < 4; ++a)
            for (int b = 0; b < 4; ++b) M.m[a][b] = object_to_world[4 * a + b];
        o.object_to_world = make_transform(M);
    }
    for (int b = 0; b < binding_count; ++b) o.material_bindings[binding_names[b]] = binding_ids[b];
    o.cap_boundary_holes = g_obj_cap != 0;
    o.cap_boundary_topology = static_cast<ObjBoundaryTopology>(g_obj_cap_topology);
    std::fflush(stderr);
    r->ok = load_obj_into_scene_with_result(path, r->scene, o, r->result);
    return r;
}

void ref_obj_free(ref_obj* r) { delete r; }

// ok, vertices, triangles, appended materials, skipped (invalid, non-finite, degenerate)
void ref_obj_counts(const ref_obj* r, int32_t* out7) {
    const ObjLoadResult& x = r->result;
    out7[0] = r->ok ? 1 : 0;
    out7[1] = x.mesh_id >= 0 ? int32_t(r->scene.mesh_assets[x.mesh_id].vertices.size()) : 0;
    out7[2] = x.mesh_id >= 0 ? int32_t(r->scene.mesh_assets[x.mesh_id].indices.size()) : 0;
    out7[3] = x.material_count;
    out7[4] = x.skipped_invalid_triangle_count;
    out7[5] = x.skipped_nonfinite_triangle_count;
    out7[6] = x.skipped_degenerate_triangle_count;
}

// cap triangles, cap boundary edges, non-simple vertices, incomplete loops,
// degenerate loops, degenerate cap triangles
void ref_obj_cap_counts(const ref_obj* r, int32_t* out6) {
    const ObjLoadResult& x = r->result;
    out6[0] = x.cap_triangle_count;
    out6[1] = x.cap_boundary_edge_count;
    out6[2] = x.cap_non_simple_boundary_vertex_count;
    out6[3] = x.cap_incomplete_loop_count;
    out6[4] = x.cap_degenerate_loop_count;
    out6[5] = x.cap_degenerate_triangle_count;
}

void ref_obj_arrays(const ref_obj* r, float* vertices, float* normals, float* uvs, uint32_t* indices,
                    int32_t* material_ids, uint8_t* boundary_edges, float* m16, float* inv16) {
    const ObjLoadResult& x = r->result;
    if (x.mesh_id < 0) return;
    const SceneMesh& mesh = r->scene.mesh_assets[x.mesh_id];
    for (size_t i = 0; i < mesh.vertices.size(); ++i) {
        std::memcpy(vertices + 3 * i, &mesh.vertices[i], 12);
        std::memcpy(normals + 3 * i, &mesh.normals[i], 12);
        std::memcpy(uvs + 2 * i, &mesh.uvs[i], 8);
    }
    for (size_t t = 0; t < mesh.indices.size(); ++t) {
        indices[3 * t] = mesh.indices[t].v0;
        indices[3 * t + 1] = mesh.indices[t].v1;
        indices[3 * t + 2] = mesh.indices[t].v2;
        material_ids[t] = mesh.material_ids[t];
        boundary_edges[t] = mesh.boundary_edges[t];
    }
    const Transform& tr = x.instance_id >= 0 ? r->scene.instances[x.instance_id].object_to_world : Transform{};
    for (int a = 0; a < 4; ++a)
        for (int b = 0; b < 4; ++b) {
            m16[4 * a + b] = tr.m.m[a][b];
            inv16[4 * a + b] = tr.inv.m[a][b];
        }
}

}  // extern "C"

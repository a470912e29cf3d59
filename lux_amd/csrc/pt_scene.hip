// pt_scene.hip -- scene preparation for the wavefront path tracer: camera,
// identity-instance shading normals, the power-alias light table, the device
// mesh and its LBVH; vertex updates regather the triangles on the device and
// refit the LBVH (same topology) instead of rebuilding it.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "pt_device.hpp"
#include "pt_runtime.hpp"

namespace lux {
namespace pt {
namespace {

template <typename T>
T* dalloc(size_t count) {
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T));
    if (e != hipSuccess) throw Error(LUX_ERR_OUT_OF_MEMORY, std::string("hipMalloc failed: ") + hipGetErrorString(e));
    return static_cast<T*>(p);
}

// soup[t] = the three indexed vertices of triangle t.
__global__ __launch_bounds__(256) void k_gather_soup(const lux_vec3* __restrict__ vertices,
                                                     const uint32_t* __restrict__ indices, int tri_count,
                                                     int vertex_count, lux_triangle* __restrict__ soup,
                                                     DevInfo* info) {
    for (int t = blockIdx.x * 256 + threadIdx.x; t < tri_count; t += gridDim.x * 256) {
        uint32_t a = indices[3 * size_t(t)], b = indices[3 * size_t(t) + 1], c = indices[3 * size_t(t) + 2];
        if (a >= uint32_t(vertex_count) || b >= uint32_t(vertex_count) || c >= uint32_t(vertex_count)) {
            set_error(info, kErrPrimRange);
            continue;
        }
        soup[t] = lux_triangle{vertices[a], vertices[b], vertices[c]};
    }
}

}  // namespace

// ---- host: scene preparation ------------------------------------------------------
namespace host {

struct V {
    float x, y, z;
};
inline V mkv(float x, float y, float z) { return V{x, y, z}; }
inline V vsub(V a, V b) { return mkv(a.x - b.x, a.y - b.y, a.z - b.z); }
inline V vadd(V a, V b) { return mkv(a.x + b.x, a.y + b.y, a.z + b.z); }
inline V vmul(V a, float s) { return mkv(a.x * s, a.y * s, a.z * s); }
inline float vdot(V a, V b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V vcross(V a, V b) { return mkv(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
inline V vnorm(V v) {
    float len = sqrtf(vdot(v, v));
    if (!(len > 0)) return v;
    float inv = 1 / len;
    return vmul(v, inv);
}

// make_perspective_camera with lens radius 0 and focal distance 1
// (src/camera/camera.h:58-109, src/camera/pinhole_camera.h:13-17).
DevCamera make_camera(const float* c) {
    V origin = mkv(c[0], c[1], c[2]), look = mkv(c[3], c[4], c[5]);
    float vfov = c[6], aspect = c[7];
    V forward = vnorm(vsub(look, origin));
    V right = vnorm(vcross(mkv(0, 1, 0), forward));
    if (vdot(right, right) == 0) right = vnorm(vcross(mkv(0, 0, 1), forward));
    V up = vcross(forward, right);
    float theta = vfov * kPi / 180.0f;
    float vh = 2.0f * tanf(theta * 0.5f);
    float vw = aspect * vh;
    float focus = 1.0f;
    V horizontal = vmul(right, vw * focus);
    V vertical = vmul(up, vh * focus);
    V lower = vsub(vsub(vadd(origin, vmul(forward, focus)), vmul(horizontal, 0.5f)), vmul(vertical, 0.5f));
    return DevCamera{origin.x, origin.y, origin.z, lower.x, lower.y, lower.z,
                     horizontal.x, horizontal.y, horizontal.z, vertical.x, vertical.y, vertical.z};
}

// transform_instance_normal under the identity instance: adjugate product then
// two normalisations (src/scene/scene_data.h:262-268, src/core/mat4.cuh:209-229).
V instance_normal(V n) {
    if (vdot(n, n) == 0) return mkv(0, 0, 0);
    V t = mkv(1.0f * n.x + 0.0f * n.y + 0.0f * n.z, 0.0f * n.x + 1.0f * n.y + 0.0f * n.z,
              0.0f * n.x + 0.0f * n.y + 1.0f * n.z);
    V once = vnorm(t);
    if (vdot(once, once) == 0) return mkv(0, 0, 0);
    return vnorm(once);
}

// transform_instance_normal (src/scene/scene_data.h:262-268) under an affine
// object_to_world m (row-major): transform_normal's cofactor matrix of the
// upper 3x3 times n, normalised (src/core/mat4.cuh:209-229), then a zero stays
// zero and the result is normalised again.
V instance_normal(const float* m, V n) {
    if (vdot(n, n) == 0) return mkv(0, 0, 0);
    const float c00 = m[5] * m[10] - m[6] * m[9];
    const float c01 = m[6] * m[8] - m[4] * m[10];
    const float c02 = m[4] * m[9] - m[5] * m[8];
    const float c10 = m[2] * m[9] - m[1] * m[10];
    const float c11 = m[0] * m[10] - m[2] * m[8];
    const float c12 = m[1] * m[8] - m[0] * m[9];
    const float c20 = m[1] * m[6] - m[2] * m[5];
    const float c21 = m[2] * m[4] - m[0] * m[6];
    const float c22 = m[0] * m[5] - m[1] * m[4];
    V t = mkv(c00 * n.x + c01 * n.y + c02 * n.z, c10 * n.x + c11 * n.y + c12 * n.z,
              c20 * n.x + c21 * n.y + c22 * n.z);
    V once = vnorm(t);
    if (vdot(once, once) == 0) return mkv(0, 0, 0);
    return vnorm(once);
}

// transform_point (src/core/mat4.cuh:173-190): rows summed left to right, a
// division by w unless w == 1.
lux_vec3 world_point(const float* m, const lux_vec3& p) {
    const float x = m[0] * p.x + m[1] * p.y + m[2] * p.z + m[3];
    const float y = m[4] * p.x + m[5] * p.y + m[6] * p.z + m[7];
    const float z = m[8] * p.x + m[9] * p.y + m[10] * p.z + m[11];
    const float w = m[12] * p.x + m[13] * p.y + m[14] * p.z + m[15];
    if (w == 1.0f) return lux_vec3{x, y, z};
    return lux_vec3{x / w, y / w, z / w};
}

// build_alias_table_from_cdf (src/light/light_distribution.cpp:8-64, 163-175):
// weights are the cdf differences; the same small/large stack order, so the
// same aliases.
void alias_from_cdf(const std::vector<float>& cdf, std::vector<float>& prob, std::vector<int>& alias,
                    std::vector<float>& pmf_out) {
    const int count = int(cdf.size());
    prob.assign(count, 1.0f);
    alias.resize(count);
    pmf_out.assign(count, 0.0f);
    std::vector<float> w(count);
    float prev = 0;
    for (int i = 0; i < count; ++i) {
        w[i] = cdf[i] - prev;
        prev = cdf[i];
    }
    float sum = 0;
    for (float x : w) sum += fmaxf(0.0f, x);
    std::vector<float> scaled(count);
    std::vector<int> small, large;
    for (int i = 0; i < count; ++i) {
        const float pmf = fmaxf(0.0f, w[i]) / sum;
        pmf_out[i] = pmf;
        alias[i] = i;
        prob[i] = 1;
        scaled[i] = pmf * float(count);
        (scaled[i] < 1.0f ? small : large).push_back(i);
    }
    while (!small.empty() && !large.empty()) {
        const int l = small.back();
        small.pop_back();
        const int g = large.back();
        large.pop_back();
        prob[l] = scaled[l];
        alias[l] = g;
        scaled[g] = (scaled[g] + scaled[l]) - 1.0f;
        (scaled[g] < 1.0f ? small : large).push_back(g);
    }
    for (int i : large) {
        prob[i] = 1;
        alias[i] = i;
    }
    for (int i : small) {
        prob[i] = 1;
        alias[i] = i;
    }
}

// Power light selection with its alias table: prepare_scene_light_sampling
// (src/scene/scene_prepare.h:268-387) and build_alias_table_from_weights
// (src/light/light_distribution.cpp:8-64; same stack order, so the same aliases).
// `emitters` lists the triangles whose material emits, in triangle order, with
// their current vertices; the other triangles contribute nothing to the table.
void build_lights(SceneDev& s, int tri_count, const std::vector<int32_t>& emitters,
                  const std::vector<lux_triangle>& tris, const std::vector<lux_material>& emitter_mats) {
    std::vector<DevLight>& lights = s.host_lights;
    std::vector<int32_t>& light_of_prim = s.host_light_of_prim;
    std::vector<float>& light_power = s.host_light_power;
    std::vector<lux_triangle>& light_tri = s.host_light_tri;
    std::vector<uint8_t>& in_bvh = s.host_light_in_bvh;
    light_of_prim.assign(tri_count, -1);
    lights.clear();
    light_power.clear();
    light_tri.clear();
    in_bvh.clear();
    std::vector<float> cdf;
    float total = 0;
    for (size_t e = 0; e < emitters.size(); ++e) {
        const int t = emitters[e];
        const lux_material& m = emitter_mats[e];
        float emax = fmaxf(m.emission.x, fmaxf(m.emission.y, m.emission.z));
        if (!(emax > 0)) continue;
        V v0 = mkv(tris[e].v0.x, tris[e].v0.y, tris[e].v0.z);
        V v1 = mkv(tris[e].v1.x, tris[e].v1.y, tris[e].v1.z);
        V v2 = mkv(tris[e].v2.x, tris[e].v2.y, tris[e].v2.z);
        V c = vcross(vsub(v1, v0), vsub(v2, v0));
        float area = 0.5f * sqrtf(vdot(c, c));
        if (area <= 0) continue;
        float lum = float(0.2126) * m.emission.x + float(0.7152) * m.emission.y + float(0.0722) * m.emission.z;
        float power = area * fmaxf(1e-6f, lum);
        if (power <= 0) continue;
        total += power;
        DevLight L{};
        L.prim = t;
        L.er = m.emission.x;
        L.eg = m.emission.y;
        L.eb = m.emission.z;
        L.two_sided = m.emission_two_sided;
        L.area = area;
        L.kind = kLightArea;
        L.nb_slot = -1;
        lights.push_back(L);
        light_power.push_back(power);
        light_tri.push_back(tris[e]);
        in_bvh.push_back(1);
        cdf.push_back(total);
        int32_t& stored = light_of_prim[t];
        stored = stored == -1 ? int32_t(lights.size()) - 1 : -2;
    }
    // scene lights, in order (prepare_scene_light_sampling's second loop)
    for (size_t k = 0; k < s.host_scene_lights.size(); ++k) {
        const lux_light& x = s.host_scene_lights[k];
        DevLight L{};
        L.kind = x.kind;
        L.nb_slot = -1;
        L.prim = -1;
        L.er = x.color.x;
        L.eg = x.color.y;
        L.eb = x.color.z;
        L.vx = x.vector.x;
        L.vy = x.vector.y;
        L.vz = x.vector.z;
        const float lum = float(0.2126) * x.color.x + float(0.7152) * x.color.y + float(0.0722) * x.color.z;
        float power = fmaxf(1e-6f, lum);
        lux_triangle pt{};
        if (x.kind == kLightImageInf) {
            L.prim = s.host_env_first[k];
            L.two_sided = x.image_width;
            L.height = x.image_height;
            L.area = s.host_env_total[k];
            power = fmaxf(1e-6f, s.host_env_total[k]);
        }
        if (x.kind == kLightPoint) pt = lux_triangle{x.vector, x.vector, x.vector};  // AABB(position)
        if (!(power > 0)) continue;
        total += power;
        lights.push_back(L);
        light_power.push_back(power);
        light_tri.push_back(pt);
        in_bvh.push_back(x.kind == kLightPoint ? 1 : 0);
        cdf.push_back(total);
    }
    int count = int(lights.size());
    if (count == 0 || !(total > 0)) {
        lights.clear();
        light_power.clear();
        light_tri.clear();
        in_bvh.clear();
        return;
    }
    std::vector<float> prob, pmf;
    std::vector<int> alias;
    alias_from_cdf(cdf, prob, alias, pmf);
    for (int i = 0; i < count; ++i) {
        lights[i].prob = prob[i];
        lights[i].alias = alias[i];
        lights[i].pmf = pmf[i];
    }
}

// The light BVH of the BVH light sampler: build_light_bvh
// (src/light/light_distribution.cpp:68-161) over one item per light (its
// world triangle's bounds, their centre, its power), split at the median
// centroid of the widest centroid axis with std::nth_element, nodes in the
// order the recursion creates them (preorder), interior power and bounds the
// sums / unions of the two children. Items are the lights in table order.
void build_light_bvh(SceneDev& s) {
    const std::vector<float>& power = s.host_light_power;
    const std::vector<lux_triangle>& tri = s.host_light_tri;
    std::vector<DevLightNode>& nodes = s.host_light_nodes;
    std::vector<int32_t>& leaf_of_light = s.host_light_leaf;
    int& root = s.light_root;
    nodes.clear();
    leaf_of_light.clear();
    root = -1;
    s.host_nb_table.clear();
    s.host_nb_total = 0;
    s.host_env_lights.clear();
    for (size_t i = 0; i < s.host_lights.size(); ++i) {
        s.host_lights[i].nb_slot = -1;
        const int kind = s.host_lights[i].kind;
        if (kind == kLightUniformInf || kind == kLightImageInf) s.host_env_lights.push_back(int32_t(i));
    }
    struct Item {
        float lo[3], hi[3], c[3];
        int light;
        float power;
    };
    std::vector<size_t> members;  // lights with bounds (area, point), in light order
    for (size_t i = 0; i < power.size(); ++i)
        if (s.host_light_in_bvh[i]) members.push_back(i);
    std::vector<Item> items(members.size());
    for (size_t j = 0; j < members.size(); ++j) {
        const size_t i = members[j];
        const lux_triangle& t = tri[i];
        const float xs[3] = {t.v0.x, t.v1.x, t.v2.x}, ys[3] = {t.v0.y, t.v1.y, t.v2.y},
                    zs[3] = {t.v0.z, t.v1.z, t.v2.z};
        const float* axes[3] = {xs, ys, zs};
        Item& it = items[j];
        for (int a = 0; a < 3; ++a) {
            it.lo[a] = fminf(fminf(axes[a][0], axes[a][1]), axes[a][2]);
            it.hi[a] = fmaxf(fmaxf(axes[a][0], axes[a][1]), axes[a][2]);
            it.c[a] = (it.lo[a] + it.hi[a]) * 0.5f;
        }
        it.light = int(i);
        it.power = power[i];
        if (!(it.power > 0) || it.lo[0] > it.hi[0] || it.lo[1] > it.hi[1] || it.lo[2] > it.hi[2]) {
            s.host_nb_table.clear();
            s.host_nb_total = 0;
            for (DevLight& L : s.host_lights) L.nb_slot = -1;
            return;
        }
    }
    if (items.empty() || power.empty()) return;
    // the non-BVH group: the other lights' weights (power-cdf differences) and
    // their own alias table (build_light_bvh, src/light/light_distribution.cpp:90-106)
    {
        std::vector<float> cdf(power.size());
        float total = 0;
        for (size_t i = 0; i < power.size(); ++i) cdf[i] = (total += power[i]);
        std::vector<float> nb_cdf;
        std::vector<int32_t> nb_light;
        float nb_weight = 0;
        for (size_t i = 0; i < power.size(); ++i) {
            if (s.host_light_in_bvh[i]) continue;
            const float previous = i == 0 ? 0.0f : cdf[i - 1];
            const float weight = cdf[i] - previous;
            if (weight <= 0) continue;
            nb_weight += weight;
            nb_light.push_back(int32_t(i));
            nb_cdf.push_back(nb_weight);
        }
        s.host_nb_total = nb_weight;
        if (!nb_cdf.empty() && nb_weight > 0) {
            std::vector<float> prob, pmf;
            std::vector<int> alias;
            alias_from_cdf(nb_cdf, prob, alias, pmf);
            for (size_t k = 0; k < nb_cdf.size(); ++k) {
                s.host_nb_table.push_back(DevAlias{prob[k], alias[k], pmf[k], nb_light[k]});
                s.host_lights[size_t(nb_light[k])].nb_slot = int(k);
            }
        } else {
            for (size_t k = 0; k < nb_cdf.size(); ++k) {
                s.host_nb_table.push_back(DevAlias{1.0f, int(k), 0.0f, nb_light[k]});
                s.host_lights[size_t(nb_light[k])].nb_slot = int(k);
            }
        }
    }
    nodes.reserve(2 * items.size() - 1);
    leaf_of_light.assign(power.size(), -1);
    auto build = [&](auto&& self, int begin, int end, int parent) -> int {
        const int index = int(nodes.size());
        nodes.push_back(DevLightNode{});
        float clo[3] = {INFINITY, INFINITY, INFINITY}, chi[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (int i = begin; i < end; ++i)
            for (int a = 0; a < 3; ++a) {
                clo[a] = fminf(clo[a], items[i].c[a]);
                chi[a] = fmaxf(chi[a], items[i].c[a]);
            }
        nodes[index].parent = parent;
        if (end - begin == 1) {
            const Item& it = items[begin];
            DevLightNode& n = nodes[index];
            for (int a = 0; a < 3; ++a) {
                n.lo[a] = it.lo[a];
                n.hi[a] = it.hi[a];
            }
            n.power = 0.0f + it.power;
            n.light = it.light;
            n.left = n.right = -1;
            leaf_of_light[it.light] = index;
            return index;
        }
        const float dx = chi[0] - clo[0], dy = chi[1] - clo[1], dz = chi[2] - clo[2];
        const int axis = (dx > dy && dx > dz) ? 0 : (dy > dz) ? 1 : 2;  // AABB::max_extent_axis
        const int mid = begin + (end - begin) / 2;
        std::nth_element(items.begin() + begin, items.begin() + mid, items.begin() + end,
                         [axis](const Item& a, const Item& b) { return a.c[axis] < b.c[axis]; });
        const int left = self(self, begin, mid, index);
        const int right = self(self, mid, end, index);
        DevLightNode& n = nodes[index];
        n.left = left;
        n.right = right;
        n.light = -1;
        n.power = nodes[left].power + nodes[right].power;
        for (int a = 0; a < 3; ++a) {
            n.lo[a] = fminf(nodes[left].lo[a], nodes[right].lo[a]);
            n.hi[a] = fmaxf(nodes[left].hi[a], nodes[right].hi[a]);
        }
        return index;
    };
    root = build(build, 0, int(items.size()), -1);
}

}  // namespace host

namespace {
void emitter_triangles(const SceneDev& s, const lux_vec3* vertices, std::vector<lux_triangle>& tris) {
    tris.resize(s.emitters.size());
    for (size_t e = 0; e < s.emitters.size(); ++e) {
        const uint32_t* ix = s.emitter_indices.data() + 3 * e;
        tris[e] = lux_triangle{vertices[ix[0]], vertices[ix[1]], vertices[ix[2]]};
    }
}

void upload_lights(SceneDev& s, hipStream_t stream) {
    host::build_light_bvh(s);
    const size_t nn = s.host_light_nodes.size(), nl = s.host_light_leaf.size();
    if (!s.light_nodes || s.light_node_capacity < int(nn) || s.light_leaf_capacity < int(nl)) {
        if (s.light_nodes) (void)hipFree(s.light_nodes);
        if (s.light_leaf) (void)hipFree(s.light_leaf);
        s.light_node_capacity = int(std::max<size_t>(nn, 1));
        s.light_leaf_capacity = int(std::max<size_t>(nl, 1));
        s.light_nodes = dalloc<DevLightNode>(size_t(s.light_node_capacity));
        s.light_leaf = dalloc<int32_t>(size_t(s.light_leaf_capacity));
    }
    s.light_node_count = int(nn);
    if (nn)
        hip_check(hipMemcpyAsync(s.light_nodes, s.host_light_nodes.data(), sizeof(DevLightNode) * nn,
                                 hipMemcpyHostToDevice, stream),
                  "H2D light BVH");
    if (nl)
        hip_check(hipMemcpyAsync(s.light_leaf, s.host_light_leaf.data(), 4 * nl, hipMemcpyHostToDevice, stream),
                  "H2D light leaves");
    if (!s.lights || s.light_count < int(s.host_lights.size())) {
        if (s.lights) (void)hipFree(s.lights);
        s.lights = dalloc<DevLight>(std::max<size_t>(s.host_lights.size(), 1));
    }
    s.light_count = int(s.host_lights.size());
    if (!s.host_lights.empty())
        hip_check(hipMemcpyAsync(s.lights, s.host_lights.data(), sizeof(DevLight) * s.host_lights.size(),
                                 hipMemcpyHostToDevice, stream),
                  "H2D lights");
    hip_check(hipMemcpyAsync(s.light_of_prim, s.host_light_of_prim.data(), 4 * size_t(s.tri_count),
                             hipMemcpyHostToDevice, stream),
              "H2D light map");
    const size_t nb = s.host_nb_table.size(), ne = s.host_env_lights.size();
    if (!s.nb_table || s.nb_capacity < int(nb)) {
        if (s.nb_table) (void)hipFree(s.nb_table);
        s.nb_capacity = int(std::max<size_t>(nb, 1));
        s.nb_table = dalloc<DevAlias>(size_t(s.nb_capacity));
    }
    if (!s.env_light_list || s.env_list_capacity < int(ne)) {
        if (s.env_light_list) (void)hipFree(s.env_light_list);
        s.env_list_capacity = int(std::max<size_t>(ne, 1));
        s.env_light_list = dalloc<int32_t>(size_t(s.env_list_capacity));
    }
    if (nb)
        hip_check(hipMemcpyAsync(s.nb_table, s.host_nb_table.data(), sizeof(DevAlias) * nb, hipMemcpyHostToDevice,
                                 stream),
                  "H2D non-BVH lights");
    if (ne)
        hip_check(hipMemcpyAsync(s.env_light_list, s.host_env_lights.data(), 4 * ne, hipMemcpyHostToDevice, stream),
                  "H2D infinite lights");
}

// Scene lights of a descriptor: checked as validate_scene_shape checks them
// (src/scene/scene_validate.h:149-190), kept on the host, image texels and
// their weights (make_image_light_asset, src/scene/scene_build.h:39-64: luminance
// of pixel x scale times sin(theta) at the row centre, host sinf) uploaded once.
void prepare_scene_lights(SceneDev& s, const lux_light* lights, int count, hipStream_t stream) {
    if (count < 0 || (count > 0 && !lights)) throw Error(LUX_ERR_INVALID_ARGUMENT, "bad scene light array");
    auto finite3 = [](const lux_vec3& v) { return std::isfinite(v.x) && std::isfinite(v.y) && std::isfinite(v.z); };
    auto nonneg3 = [](const lux_vec3& v) { return v.x >= 0 && v.y >= 0 && v.z >= 0; };
    s.host_scene_lights.clear();
    s.host_env_pixels.clear();
    s.host_env_cdf.clear();
    s.host_env_first.clear();
    s.host_env_total.clear();
    for (int k = 0; k < count; ++k) {
        lux_light x = lights[k];
        const std::string which = "scene light " + std::to_string(k);
        if (x.kind < kLightPoint || x.kind > kLightImageInf) throw Error(LUX_ERR_INVALID_ARGUMENT, which + ": unknown kind");
        if (!finite3(x.color) || !nonneg3(x.color))
            throw Error(LUX_ERR_INVALID_ARGUMENT, which + ": colour must be finite and non-negative");
        if ((x.kind == kLightPoint || x.kind == kLightDistant) && !finite3(x.vector))
            throw Error(LUX_ERR_INVALID_ARGUMENT, which + ": position / direction must be finite");
        int32_t first = -1;
        float total = 0;
        if (x.kind == kLightDistant) {  // add_distant_light normalises the direction
            host::V d = host::mkv(x.vector.x, x.vector.y, x.vector.z);
            if (host::vdot(d, d) == 0) throw Error(LUX_ERR_INVALID_ARGUMENT, which + ": direction must be non-zero");
            d = host::vnorm(d);
            x.vector = lux_vec3{d.x, d.y, d.z};
        }
        if (x.kind == kLightImageInf) {
            const int w = x.image_width, h = x.image_height;
            if (w <= 0 || h <= 0 || !x.image_pixels)
                throw Error(LUX_ERR_INVALID_ARGUMENT, which + ": image needs width, height and pixels");
            first = int32_t(s.host_env_cdf.size());
            float sum = 0;
            for (int y = 0; y < h; ++y) {
                const float theta = kPi * ((float(y) + 0.5f) / float(h));
                const float sin_theta = fmaxf(0.0f, sinf(theta));
                for (int xx = 0; xx < w; ++xx) {
                    const lux_vec3& px = x.image_pixels[size_t(y) * w + xx];
                    if (!finite3(px)) throw Error(LUX_ERR_INVALID_ARGUMENT, which + ": non-finite texel");
                    const float r = px.x * x.color.x, g = px.y * x.color.y, b = px.z * x.color.z;
                    const float lum = float(0.2126) * r + float(0.7152) * g + float(0.0722) * b;
                    sum += fmaxf(0.0f, lum) * sin_theta;
                    s.host_env_cdf.push_back(sum);
                    s.host_env_pixels.insert(s.host_env_pixels.end(), {px.x, px.y, px.z});
                }
            }
            total = sum;
            x.image_pixels = nullptr;  // the texels now live in host_env_pixels
        }
        s.host_scene_lights.push_back(x);
        s.host_env_first.push_back(first);
        s.host_env_total.push_back(total);
    }
    if (s.env_pixels) (void)hipFree(s.env_pixels);
    if (s.env_cdf) (void)hipFree(s.env_cdf);
    s.env_pixels = dalloc<float>(std::max<size_t>(s.host_env_pixels.size(), 1));
    s.env_cdf = dalloc<float>(std::max<size_t>(s.host_env_cdf.size(), 1));
    if (!s.host_env_cdf.empty()) {
        hip_check(hipMemcpyAsync(s.env_pixels, s.host_env_pixels.data(), 4 * s.host_env_pixels.size(),
                                 hipMemcpyHostToDevice, stream),
                  "H2D environment texels");
        hip_check(hipMemcpyAsync(s.env_cdf, s.host_env_cdf.data(), 4 * s.host_env_cdf.size(), hipMemcpyHostToDevice,
                                 stream),
                  "H2D environment weights");
    }
}

size_t bvh_bytes(int n) {
    // boxes, ids, 2x keys/vals, 5 tree links, leaf parents, 2 edge boxes, 2
    // layout maps, 2 node records, 2 climb words, 2 sibling-pair records with
    // their two slot maps; block boxes and sort state
    return size_t(std::max(n, 1)) * (24 + 4 + 16 + 20 + 4 + 48 + 8 + 64 + 8 + 64 + 16) +
           (size_t(n) / 1024 + 1) * 24 + (size_t(n) / 2048 + 1) * 2048;
}
}  // namespace

namespace {
void check_desc(const lux_scene_desc& desc) {
    const int nt = desc.triangle_count, nv = desc.vertex_count, nm = desc.material_count;
    if (nt <= 0 || nv <= 0 || nm <= 0 || !desc.vertices || !desc.indices || !desc.material_ids || !desc.materials)
        throw Error(LUX_ERR_INVALID_ARGUMENT, "scene needs vertices, indices, material ids and materials");
    for (int t = 0; t < nt; ++t) {
        for (int k = 0; k < 3; ++k)
            if (desc.indices[3 * size_t(t) + k] >= uint32_t(nv))
                throw Error(LUX_ERR_INVALID_ARGUMENT, "triangle index out of range");
        if (desc.material_ids[t] < 0 || desc.material_ids[t] >= nm)
            throw Error(LUX_ERR_INVALID_ARGUMENT, "material id out of range");
    }
    for (int m = 0; m < nm; ++m)
        if (desc.materials[m].type < 0 || desc.materials[m].type >= kMaterialTypes)
            throw Error(LUX_ERR_INVALID_ARGUMENT, "unknown material type");
    if (desc.light_count < 0 || (desc.light_count > 0 && !desc.lights))
        throw Error(LUX_ERR_INVALID_ARGUMENT, "bad scene light array");
}

// HBM of a one-mesh scene: triangle soup, indices, material ids, light of each
// triangle, vertices (and normals), materials, light records, BVH.
size_t scene_bytes(int nt, int nv, bool normals, int nm, size_t lights) {
    return size_t(nt) * (36 + 12 + 4 + 4) + size_t(nv) * (12 + (normals ? 12 : 0)) + sizeof(DevMaterial) * nm +
           sizeof(DevLight) * std::max<size_t>(lights, 1) + bvh_bytes(nt);
}
}  // namespace

void scene_upload(SceneDev& s, const lux_scene_desc& desc, hipStream_t stream) {
    check_desc(desc);
    int nt = desc.triangle_count, nv = desc.vertex_count, nm = desc.material_count;
    std::vector<float> normals;
    if (desc.normals) {
        normals.resize(3 * size_t(nv));
        for (int v = 0; v < nv; ++v) {
            host::V n = host::instance_normal(host::mkv(desc.normals[v].x, desc.normals[v].y, desc.normals[v].z));
            normals[3 * v] = n.x;
            normals[3 * v + 1] = n.y;
            normals[3 * v + 2] = n.z;
        }
    }
    std::vector<DevMaterial> mats(nm);
    for (int m = 0; m < nm; ++m) {
        const lux_material& in = desc.materials[m];
        if (in.type < 0 || in.type >= kMaterialTypes) throw Error(LUX_ERR_INVALID_ARGUMENT, "unknown material type");
        mats[m] = DevMaterial{in.type, in.albedo.x, in.albedo.y, in.albedo.z, in.emission.x, in.emission.y,
                              in.emission.z, in.emission_two_sided, in.roughness, in.metallic, in.eta, 0};
    }

    scene_release(s);
    s.tri_count = nt;
    s.vertex_count = nv;
    s.mat_count = nm;
    s.host_mats.assign(desc.materials, desc.materials + nm);
    s.opaque = true;
    for (const DevMaterial& m : mats) s.opaque = s.opaque && m.type != kInterface && m.type != kTransparent;
    for (int t = 0; t < nt; ++t) {
        const lux_material& m = desc.materials[desc.material_ids[t]];
        if (fmaxf(m.emission.x, fmaxf(m.emission.y, m.emission.z)) > 0) {
            s.emitters.push_back(t);
            s.emitter_mat.push_back(desc.material_ids[t]);
            for (int k = 0; k < 3; ++k) s.emitter_indices.push_back(desc.indices[3 * size_t(t) + k]);
        }
    }
    std::vector<lux_triangle> etris;
    emitter_triangles(s, desc.vertices, etris);
    std::vector<lux_material> emats(s.emitters.size());
    for (size_t e = 0; e < s.emitters.size(); ++e) emats[e] = s.host_mats[s.emitter_mat[e]];
    prepare_scene_lights(s, desc.lights, desc.light_count, stream);
    host::build_lights(s, nt, s.emitters, etris, emats);

    s.soup = dalloc<lux_triangle>(nt);
    s.vertices = dalloc<lux_vec3>(nv);
    s.indices = dalloc<uint32_t>(3 * size_t(nt));
    s.mat_id = dalloc<int32_t>(nt);
    s.mats = dalloc<DevMaterial>(nm);
    s.light_of_prim = dalloc<int32_t>(nt);
    if (desc.normals) s.normals = dalloc<float>(3 * size_t(nv));
    s.camera = host::make_camera(desc.camera);
    hip_check(hipMemcpyAsync(s.vertices, desc.vertices, sizeof(lux_vec3) * nv, hipMemcpyHostToDevice, stream),
              "H2D vertices");
    hip_check(hipMemcpyAsync(s.indices, desc.indices, 12 * size_t(nt), hipMemcpyHostToDevice, stream), "H2D indices");
    hip_check(hipMemcpyAsync(s.mat_id, desc.material_ids, 4 * size_t(nt), hipMemcpyHostToDevice, stream), "H2D mat ids");
    hip_check(hipMemcpyAsync(s.mats, mats.data(), sizeof(DevMaterial) * nm, hipMemcpyHostToDevice, stream), "H2D mats");
    upload_lights(s, stream);
    if (desc.normals)
        hip_check(hipMemcpyAsync(s.normals, normals.data(), 4 * normals.size(), hipMemcpyHostToDevice, stream),
                  "H2D normals");
    // BLAS over the triangle boxes (ids = triangle index), both computed on the
    // device from the indexed mesh; the soup is gathered there too
    lbvh_build_triangles(s.bvh, nullptr, s.vertices, s.indices, nv, nt, stream);
    LUX_LAUNCH("pt_gather_soup", k_gather_soup, std::min(div_up(nt, 256), 4096), 256, 0, stream, s.vertices,
               s.indices, nt, nv, s.soup, s.bvh.info);
    hip_check(hipStreamSynchronize(stream), "scene upload");
    s.device_bytes = scene_bytes(nt, nv, desc.normals != nullptr, nm, s.host_lights.size());
    s.resident = true;
}

void scene_create(SceneDev& s, const lux_scene_desc& desc, bool upload, hipStream_t stream) {
    check_desc(desc);
    auto src = std::make_shared<SceneSource>();
    const size_t nt = size_t(desc.triangle_count), nv = size_t(desc.vertex_count);
    src->vertices.assign(desc.vertices, desc.vertices + nv);
    src->indices.assign(desc.indices, desc.indices + 3 * nt);
    if (desc.normals) src->normals.assign(desc.normals, desc.normals + nv);
    src->material_ids.assign(desc.material_ids, desc.material_ids + nt);
    src->materials.assign(desc.materials, desc.materials + desc.material_count);
    std::memcpy(src->camera, desc.camera, sizeof(src->camera));
    src->lights.assign(desc.lights, desc.lights + desc.light_count);
    src->texels.resize(src->lights.size());
    for (size_t i = 0; i < src->lights.size(); ++i) {
        lux_light& L = src->lights[i];
        if (L.image_pixels && L.image_width > 0 && L.image_height > 0) {
            src->texels[i].assign(L.image_pixels, L.image_pixels + size_t(L.image_width) * size_t(L.image_height));
            L.image_pixels = src->texels[i].data();
        }
    }
    if (upload) {
        scene_upload(s, desc, stream);
    } else {
        scene_release(s);
        s.tri_count = int(nt);
        s.vertex_count = int(nv);
        s.mat_count = desc.material_count;
        s.camera = host::make_camera(desc.camera);
        size_t emitters = 0;
        for (size_t t = 0; t < nt; ++t) {
            const lux_material& m = desc.materials[desc.material_ids[t]];
            emitters += fmaxf(m.emission.x, fmaxf(m.emission.y, m.emission.z)) > 0 ? 1 : 0;
        }
        s.device_bytes = scene_bytes(int(nt), int(nv), desc.normals != nullptr, desc.material_count,
                                     emitters + size_t(desc.light_count));
        s.resident = false;
    }
    s.source = std::move(src);
}

void scene_restore(SceneDev& fresh, const SceneSource& src, hipStream_t stream) {
    lux_scene_desc d{};
    d.vertices = src.vertices.data();
    d.vertex_count = int32_t(src.vertices.size());
    d.indices = src.indices.data();
    d.normals = src.normals.empty() ? nullptr : src.normals.data();
    d.material_ids = src.material_ids.data();
    d.triangle_count = int32_t(src.material_ids.size());
    d.materials = src.materials.data();
    d.material_count = int32_t(src.materials.size());
    std::memcpy(d.camera, src.camera, sizeof(d.camera));
    d.lights = src.lights.empty() ? nullptr : src.lights.data();
    d.light_count = int32_t(src.lights.size());
    scene_upload(fresh, d, stream);
    if (!src.current.empty()) scene_update_vertices(fresh, src.current.data(), int(src.current.size()), stream);
}

void scene_adopt(SceneDev& s, SceneDev&& fresh) {
    const DevCamera camera = s.camera;  // as the caller last set it
    std::shared_ptr<SceneSource> src = s.source;
    const int jobs = s.server_jobs;
    s = std::move(fresh);
    fresh = SceneDev{};  // the device arrays belong to `s` now
    s.camera = camera;
    s.source = std::move(src);
    s.server_jobs = jobs;
    s.resident = true;
}

SceneDev scene_detach(SceneDev& s) {
    if (!s.source || s.instanced) throw Error(LUX_ERR_INVALID_ARGUMENT, "only a kept one-mesh scene can be dropped");
    SceneDev dead = std::move(s);
    s = SceneDev{};
    s.source = dead.source;
    s.server_jobs = dead.server_jobs;
    s.camera = dead.camera;
    s.tri_count = dead.tri_count;
    s.vertex_count = dead.vertex_count;
    s.mat_count = dead.mat_count;
    s.device_bytes = dead.device_bytes;
    s.opaque = dead.opaque;
    s.resident = false;
    dead.source.reset();
    return dead;
}

void scene_update_vertices(SceneDev& s, const lux_vec3* vertices, int vertex_count, hipStream_t stream) {
    if (!s.soup) throw Error(LUX_ERR_INVALID_ARGUMENT, "scene is empty");
    if (s.instanced) throw Error(LUX_ERR_INVALID_ARGUMENT, "instanced scene: vertex updates do not apply");
    if (!vertices || vertex_count != s.vertex_count)
        throw Error(LUX_ERR_MISMATCH, "vertex update must keep the vertex count");
    std::vector<lux_triangle> etris;
    emitter_triangles(s, vertices, etris);
    std::vector<lux_material> emats(s.emitters.size());
    for (size_t e = 0; e < s.emitters.size(); ++e) emats[e] = s.host_mats[s.emitter_mat[e]];
    host::build_lights(s, s.tri_count, s.emitters, etris, emats);
    // stream order: renders enqueued before this update read the old geometry
    hip_check(hipMemcpyAsync(s.vertices, vertices, sizeof(lux_vec3) * size_t(vertex_count), hipMemcpyHostToDevice,
                             stream),
              "H2D vertices");
    LUX_LAUNCH("pt_gather_soup", k_gather_soup, std::min(div_up(s.tri_count, 256), 4096), 256, 0, stream, s.vertices,
               s.indices, s.tri_count, s.vertex_count, s.soup, s.bvh.info);
    lbvh_refit_triangles(s.bvh, s.soup, stream);
    upload_lights(s, stream);
    // the host arrays above are pageable: the copies complete before return
    hip_check(hipStreamSynchronize(stream), "scene update");
    if (s.source) s.source->current.assign(vertices, vertices + vertex_count);
}

// ---- instanced scenes ------------------------------------------------------------------
namespace {

bool finite3(const lux_vec3& v) { return std::isfinite(v.x) && std::isfinite(v.y) && std::isfinite(v.z); }

// The checks of validate_scene_shape (src/scene/scene_validate.h:125-147) that
// the two-level build does not make itself: material overrides in range.
void check_overrides(const lux_instance* instances, int count, int material_count) {
    for (int i = 0; i < count; ++i)
        if (instances[i].material_override < -1 || instances[i].material_override >= material_count)
            throw Error(LUX_ERR_INVALID_ARGUMENT, "instance " + std::to_string(i) + ": material override out of range");
}

// Mesh bounds (finalize_scene_mesh: union of the triangle bounds) through
// transform_bounds: SceneInstance::world_bounds as add_scene_instance sets it.
std::vector<lux_instance> with_world_bounds(const SceneDev& s, const lux_instance* instances, int count) {
    std::vector<lux_aabb> mesh_box(static_cast<size_t>(s.mesh_count));
    for (int m = 0; m < s.mesh_count; ++m) {
        lux_aabb b;
        b.p_min = lux_vec3{INFINITY, INFINITY, INFINITY};
        b.p_max = lux_vec3{-INFINITY, -INFINITY, -INFINITY};
        const int v0 = s.mesh_vertex_off[m];
        for (int t = s.mesh_tri_off[m]; t < s.mesh_tri_off[m + 1]; ++t)
            for (int k = 0; k < 3; ++k) {
                const lux_vec3& p = s.mesh_vertices[size_t(v0) + s.mesh_indices[3 * size_t(t) + k]];
                b.p_min = lux_vec3{std::fmin(b.p_min.x, p.x), std::fmin(b.p_min.y, p.y), std::fmin(b.p_min.z, p.z)};
                b.p_max = lux_vec3{std::fmax(b.p_max.x, p.x), std::fmax(b.p_max.y, p.y), std::fmax(b.p_max.z, p.z)};
            }
        mesh_box[m] = b;
    }
    std::vector<lux_instance> out(instances, instances + count);
    for (lux_instance& in : out)
        if (in.mesh_id >= 0 && in.mesh_id < s.mesh_count)
            in.world_bounds = transform_bounds(in.object_to_world, mesh_box[in.mesh_id]);
    return out;
}

int64_t ordinal_count(const SceneDev& s, const lux_instance* instances, int count) {
    int64_t n = 0;
    for (int i = 0; i < count; ++i) n += s.mesh_tri_off[instances[i].mesh_id + 1] - s.mesh_tri_off[instances[i].mesh_id];
    return n;
}

// scene_triangle_view of every scene ordinal (src/scene/scene_data.h:335-372):
// world corners, world corner normals, material after the override; then the
// light table over the emissive ordinals and the upload of all of it.
void flatten_instances(SceneDev& s, const lux_instance* instances, int count, hipStream_t stream) {
    const int64_t total = ordinal_count(s, instances, count);
    if (total <= 0) throw Error(LUX_ERR_INVALID_ARGUMENT, "the instances place no triangle");
    if (total > INT32_MAX / 3) throw Error(LUX_ERR_INVALID_ARGUMENT, "too many scene primitives");
    const int T = int(total);
    const bool has_normals = !s.mesh_normals.empty();
    std::vector<lux_triangle> soup(static_cast<size_t>(T));
    std::vector<float> normals(has_normals ? 9 * size_t(T) : 0);
    std::vector<int32_t> mat_id(static_cast<size_t>(T));
    std::vector<int32_t> emitters, emitter_mat;
    std::vector<lux_triangle> etris;
    size_t o = 0;
    for (int i = 0; i < count; ++i) {
        const lux_instance& in = instances[i];
        const float* m = in.object_to_world;
        const int v0 = s.mesh_vertex_off[in.mesh_id];
        for (int t = s.mesh_tri_off[in.mesh_id]; t < s.mesh_tri_off[in.mesh_id + 1]; ++t, ++o) {
            const uint32_t* ix = s.mesh_indices.data() + 3 * size_t(t);
            lux_vec3 c[3];
            for (int k = 0; k < 3; ++k) {
                c[k] = host::world_point(m, s.mesh_vertices[size_t(v0) + ix[k]]);
                if (has_normals) {
                    const lux_vec3& n = s.mesh_normals[size_t(v0) + ix[k]];
                    const host::V w = host::instance_normal(m, host::mkv(n.x, n.y, n.z));
                    normals[9 * o + 3 * k] = w.x;
                    normals[9 * o + 3 * k + 1] = w.y;
                    normals[9 * o + 3 * k + 2] = w.z;
                }
            }
            soup[o] = lux_triangle{c[0], c[1], c[2]};
            mat_id[o] = in.material_override >= 0 ? in.material_override : s.mesh_mat_id[t];
            const lux_material& mm = s.host_mats[mat_id[o]];
            if (fmaxf(mm.emission.x, fmaxf(mm.emission.y, mm.emission.z)) > 0) {
                emitters.push_back(int32_t(o));
                emitter_mat.push_back(mat_id[o]);
                etris.push_back(soup[o]);
            }
        }
    }
    std::vector<lux_material> emats(emitters.size());
    for (size_t e = 0; e < emitters.size(); ++e) emats[e] = s.host_mats[emitter_mat[e]];
    host::build_lights(s, T, emitters, etris, emats);
    s.emitters = emitters;
    s.emitter_mat = emitter_mat;

    if (!s.soup || T != s.tri_count) {
        void* old[] = {s.soup, s.indices, s.normals, s.mat_id, s.light_of_prim};
        for (void* p : old)
            if (p) (void)hipFree(p);
        s.soup = dalloc<lux_triangle>(T);
        s.indices = dalloc<uint32_t>(3 * size_t(T));
        s.normals = has_normals ? dalloc<float>(9 * size_t(T)) : nullptr;
        s.mat_id = dalloc<int32_t>(T);
        s.light_of_prim = dalloc<int32_t>(T);
        std::vector<uint32_t> iota(3 * size_t(T));
        for (size_t k = 0; k < iota.size(); ++k) iota[k] = uint32_t(k);
        hip_check(hipMemcpyAsync(s.indices, iota.data(), 4 * iota.size(), hipMemcpyHostToDevice, stream),
                  "H2D corner indices");
        s.tri_count = T;
        // the pageable source above must outlive the copy
        hip_check(hipStreamSynchronize(stream), "corner indices");
    }
    hip_check(hipMemcpyAsync(s.soup, soup.data(), sizeof(lux_triangle) * soup.size(), hipMemcpyHostToDevice, stream),
              "H2D world triangles");
    if (has_normals)
        hip_check(hipMemcpyAsync(s.normals, normals.data(), 4 * normals.size(), hipMemcpyHostToDevice, stream),
                  "H2D world normals");
    hip_check(hipMemcpyAsync(s.mat_id, mat_id.data(), 4 * mat_id.size(), hipMemcpyHostToDevice, stream),
              "H2D ordinal materials");
    upload_lights(s, stream);
    hip_check(hipStreamSynchronize(stream), "instanced scene upload");
}

}  // namespace

void scene_upload_instanced(SceneDev& s, const lux_instanced_scene_desc& d, hipStream_t stream) {
    const int nm = d.mesh_count, nmat = d.material_count;
    if (nm <= 0 || nmat <= 0 || !d.vertex_offsets || !d.triangle_offsets || !d.materials)
        throw Error(LUX_ERR_INVALID_ARGUMENT, "instanced scene needs meshes, offsets and materials");
    if (d.instance_count < 0 || (d.instance_count > 0 && !d.instances))
        throw Error(LUX_ERR_INVALID_ARGUMENT, "bad instance array");
    if (d.vertex_offsets[0] != 0 || d.triangle_offsets[0] != 0)
        throw Error(LUX_ERR_INVALID_ARGUMENT, "mesh offsets must start at 0");
    for (int m = 0; m < nm; ++m)
        if (d.vertex_offsets[m + 1] < d.vertex_offsets[m] || d.triangle_offsets[m + 1] < d.triangle_offsets[m])
            throw Error(LUX_ERR_INVALID_ARGUMENT, "mesh offsets must not decrease");
    const int nv = d.vertex_offsets[nm], nt = d.triangle_offsets[nm];
    if ((nv > 0 && !d.vertices) || (nt > 0 && (!d.indices || !d.material_ids)))
        throw Error(LUX_ERR_INVALID_ARGUMENT, "null mesh arrays");
    for (int v = 0; v < nv; ++v)
        if (!finite3(d.vertices[v]) || (d.normals && !finite3(d.normals[v])))
            throw Error(LUX_ERR_INVALID_ARGUMENT, "mesh vertex or normal is not finite");
    for (int m = 0; m < nm; ++m) {
        const uint32_t count = uint32_t(d.vertex_offsets[m + 1] - d.vertex_offsets[m]);
        for (int t = d.triangle_offsets[m]; t < d.triangle_offsets[m + 1]; ++t) {
            for (int k = 0; k < 3; ++k)
                if (d.indices[3 * size_t(t) + k] >= count)
                    throw Error(LUX_ERR_INVALID_ARGUMENT, "mesh triangle index out of range");
            if (d.material_ids[t] < 0 || d.material_ids[t] >= nmat)
                throw Error(LUX_ERR_INVALID_ARGUMENT, "mesh material id out of range");
        }
    }
    std::vector<DevMaterial> mats(nmat);
    for (int m = 0; m < nmat; ++m) {
        const lux_material& in = d.materials[m];
        if (in.type < 0 || in.type >= kMaterialTypes) throw Error(LUX_ERR_INVALID_ARGUMENT, "unknown material type");
        mats[m] = DevMaterial{in.type, in.albedo.x, in.albedo.y, in.albedo.z, in.emission.x, in.emission.y,
                              in.emission.z, in.emission_two_sided, in.roughness, in.metallic, in.eta, 0};
    }
    check_overrides(d.instances, d.instance_count, nmat);

    SceneDev n;
    n.instanced = true;
    n.mesh_count = nm;
    n.mesh_vertices.assign(d.vertices, d.vertices + nv);
    if (d.normals) n.mesh_normals.assign(d.normals, d.normals + nv);
    n.mesh_indices.assign(d.indices, d.indices + 3 * size_t(nt));
    n.mesh_mat_id.assign(d.material_ids, d.material_ids + nt);
    n.mesh_vertex_off.assign(d.vertex_offsets, d.vertex_offsets + nm + 1);
    n.mesh_tri_off.assign(d.triangle_offsets, d.triangle_offsets + nm + 1);
    n.host_mats.assign(d.materials, d.materials + nmat);
    n.mat_count = nmat;
    n.vertex_count = nv;
    n.opaque = true;
    for (const DevMaterial& m : mats) n.opaque = n.opaque && m.type != kInterface && m.type != kTransparent;
    n.camera = host::make_camera(d.camera);
    for (int i = 0; i < d.instance_count; ++i)  // mesh ids first: the ordinal count reads them
        if (d.instances[i].mesh_id < 0 || d.instances[i].mesh_id >= nm)
            throw Error(LUX_ERR_INVALID_ARGUMENT, "instance " + std::to_string(i) + ": mesh id out of range");
    if (ordinal_count(n, d.instances, d.instance_count) <= 0)
        throw Error(LUX_ERR_INVALID_ARGUMENT, "the instances place no triangle");
    const std::vector<lux_instance> inst = with_world_bounds(n, d.instances, d.instance_count);
    // object-space soup per mesh for the BLASes
    std::vector<lux_triangle> obj(static_cast<size_t>(nt));
    for (int m = 0; m < nm; ++m)
        for (int t = d.triangle_offsets[m]; t < d.triangle_offsets[m + 1]; ++t) {
            const lux_vec3* V = d.vertices + d.vertex_offsets[m];
            const uint32_t* ix = d.indices + 3 * size_t(t);
            obj[t] = lux_triangle{V[ix[0]], V[ix[1]], V[ix[2]]};
        }
    try {
        prepare_scene_lights(n, d.lights, d.light_count, stream);
        two_level_build(n.tl, obj.data(), d.triangle_offsets, nm, inst.data(), d.instance_count, stream);
        n.mats = dalloc<DevMaterial>(nmat);
        hip_check(hipMemcpyAsync(n.mats, mats.data(), sizeof(DevMaterial) * nmat, hipMemcpyHostToDevice, stream),
                  "H2D mats");
        flatten_instances(n, inst.data(), d.instance_count, stream);
    } catch (...) {
        scene_release(n);
        throw;
    }
    n.device_bytes = n.tl.device_bytes + size_t(n.tri_count) * (36 + 12 + (d.normals ? 36 : 0) + 4 + 4) +
                     sizeof(DevMaterial) * nmat + sizeof(DevLight) * std::max<size_t>(n.host_lights.size(), 1);
    n.resident = true;
    scene_release(s);
    s = std::move(n);
}

void scene_set_instances(SceneDev& s, const lux_instance* instances, int count, hipStream_t stream) {
    if (!s.instanced) throw Error(LUX_ERR_INVALID_ARGUMENT, "not an instanced scene");
    if (count != s.tl.instance_count) throw Error(LUX_ERR_MISMATCH, "instance count differs from the scene's");
    if (s.server_jobs > 0)
        throw Error(LUX_ERR_INVALID_ARGUMENT, "the scene has unfinished server requests (poll them to completion first)");
    if (count > 0 && !instances) throw Error(LUX_ERR_INVALID_ARGUMENT, "null instance array");
    check_overrides(instances, count, s.mat_count);
    for (int i = 0; i < count; ++i)
        if (instances[i].mesh_id < 0 || instances[i].mesh_id >= s.mesh_count)
            throw Error(LUX_ERR_INVALID_ARGUMENT, "instance " + std::to_string(i) + ": mesh id out of range");
    for (int i = 0; i < count; ++i)
        if (!instance_transform_valid(instances[i].object_to_world, instances[i].world_to_object))
            throw Error(LUX_ERR_INVALID_ARGUMENT, "instance " + std::to_string(i) +
                                                      ": transform must be affine-invertible with a consistent inverse");
    if (ordinal_count(s, instances, count) <= 0) throw Error(LUX_ERR_INVALID_ARGUMENT, "the instances place no triangle");
    const std::vector<lux_instance> inst = with_world_bounds(s, instances, count);
    // renders enqueued earlier (any stream) finish with the old placement
    hip_check(hipDeviceSynchronize(), "synchronize before instance update");
    two_level_set_instances(s.tl, inst.data(), count, stream);
    flatten_instances(s, inst.data(), count, stream);
    s.device_bytes = s.tl.device_bytes + size_t(s.tri_count) * (36 + 12 + (s.normals ? 36 : 0) + 4 + 4) +
                     sizeof(DevMaterial) * s.mat_count + sizeof(DevLight) * std::max<size_t>(s.host_lights.size(), 1);
    s.resident = true;
}

void scene_release(SceneDev& s) {
    void* ptrs[] = {s.soup, s.vertices, s.indices, s.normals, s.mat_id, s.mats, s.light_of_prim, s.lights,
                    s.light_nodes, s.light_leaf, s.nb_table, s.env_light_list, s.env_pixels, s.env_cdf};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    bvh_release(s.bvh);
    two_level_release(s.tl);
    s = SceneDev{};
}

SceneView scene_view(const SceneDev& s) {
    SceneView v{};
    if (s.instanced) {
        v.nodes = nullptr;
        v.slot_pre = nullptr;
        v.node_count = 0;
        v.tl = two_level_view(s.tl);
        v.instanced = 1;
    } else {
        v.nodes = s.bvh.pairs;  // sibling-pair traversal copy
        v.slot_pre = s.bvh.slot_pre;
        v.node_count = s.bvh.n > 1 ? 2 * s.bvh.n : (s.bvh.n > 0 ? 1 : 0);
        v.instanced = 0;
    }
    v.soup = s.soup;
    v.tri_count = s.tri_count;
    v.indices = s.indices;
    v.normals = s.normals;
    v.mat_id = s.mat_id;
    v.mats = s.mats;
    v.mat_count = s.mat_count;
    v.lights = s.lights;
    v.light_count = s.light_count;
    v.light_of_prim = s.light_of_prim;
    v.light_nodes = s.light_nodes;
    v.light_leaf = s.light_leaf;
    v.light_node_count = s.light_node_count;
    v.light_root = (s.light_root >= 0 && s.light_node_count > 0 && !s.host_light_leaf.empty()) ? s.light_root : -1;
    v.nb_table = s.nb_table;
    v.nb_count = int32_t(s.host_nb_table.size());
    v.nb_total = s.host_nb_total;
    v.env_lights = s.env_light_list;
    v.env_count = int32_t(s.host_env_lights.size());
    v.env_pixels = s.env_pixels;
    v.env_cdf = s.env_cdf;
    v.opaque = s.opaque ? 1 : 0;
    v.cam = s.camera;
    v.info = s.instanced ? s.tl.tlas.info : s.bvh.info;
    return v;
}

}  // namespace pt
}  // namespace lux

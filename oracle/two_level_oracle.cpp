// two_level_oracle.cpp -- CPU restatement of the reference's two-level
// (instanced) scene path. TEST INFRASTRUCTURE ONLY: only tests and the bench's
// CPU baseline use it, as the checker.
//
//   inverse_affine, affine_determinant3x3, is_affine_invertible:  src/core/mat4.cuh:86-159
//   transform_point / transform_vector / transform_bounds:        src/core/mat4.cuh:173-270
//   scene validation of instance transforms:                       src/scene/scene_validate.h:140-147
//   BLAS per mesh over triangle boxes, TLAS over non-empty
//   instance world bounds (ids = instance ids):                   src/scene/prepared_scene.cu:35-47,
//                                                                 src/scene/scene_prepare.h:420-462
//   first scene ordinal of an instance:                           src/scene/gpu_scene.h:149-172
//   the two-level walk:                                           src/scene/scene_intersection.h:147-240
//     over intersect_cpu_bvh_nodes                                src/bvh/cpu_bvh_traversal.h:17-81
#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <vector>

#include "oracle_core.hpp"

using namespace orc;

namespace {

struct M4 {
    float m[4][4];
};

M4 load_m4(const float* p) {
    M4 r;
    std::memcpy(r.m, p, sizeof(r.m));
    return r;
}

M4 identity() {
    M4 r{};
    for (int i = 0; i < 4; ++i) r.m[i][i] = 1.0f;
    return r;
}

float det3(const M4& t) {
    const float a00 = t.m[0][0], a01 = t.m[0][1], a02 = t.m[0][2];
    const float a10 = t.m[1][0], a11 = t.m[1][1], a12 = t.m[1][2];
    const float a20 = t.m[2][0], a21 = t.m[2][1], a22 = t.m[2][2];
    return a00 * (a11 * a22 - a12 * a21) - a01 * (a10 * a22 - a12 * a20) + a02 * (a10 * a21 - a11 * a20);
}

bool is_affine(const M4& t) {
    return std::fabs(t.m[3][0]) <= 1e-6f && std::fabs(t.m[3][1]) <= 1e-6f && std::fabs(t.m[3][2]) <= 1e-6f &&
           std::fabs(t.m[3][3] - 1.0f) <= 1e-6f;
}

bool affine_invertible(const M4& t) { return is_affine(t) && std::fabs(det3(t)) > 1e-12f; }

M4 inverse_affine(const M4& t) {
    M4 r = identity();
    const float a00 = t.m[0][0], a01 = t.m[0][1], a02 = t.m[0][2];
    const float a10 = t.m[1][0], a11 = t.m[1][1], a12 = t.m[1][2];
    const float a20 = t.m[2][0], a21 = t.m[2][1], a22 = t.m[2][2];
    const float det = det3(t);
    if (std::fabs(det) <= 1e-12f) return r;
    const float inv_det = 1.0f / det;
    r.m[0][0] = (a11 * a22 - a12 * a21) * inv_det;
    r.m[0][1] = -(a01 * a22 - a02 * a21) * inv_det;
    r.m[0][2] = (a01 * a12 - a02 * a11) * inv_det;
    r.m[1][0] = -(a10 * a22 - a12 * a20) * inv_det;
    r.m[1][1] = (a00 * a22 - a02 * a20) * inv_det;
    r.m[1][2] = -(a00 * a12 - a02 * a10) * inv_det;
    r.m[2][0] = (a10 * a21 - a11 * a20) * inv_det;
    r.m[2][1] = -(a00 * a21 - a01 * a20) * inv_det;
    r.m[2][2] = (a00 * a11 - a01 * a10) * inv_det;
    const float tx = t.m[0][3], ty = t.m[1][3], tz = t.m[2][3];
    r.m[0][3] = -(r.m[0][0] * tx + r.m[0][1] * ty + r.m[0][2] * tz);
    r.m[1][3] = -(r.m[1][0] * tx + r.m[1][1] * ty + r.m[1][2] * tz);
    r.m[2][3] = -(r.m[2][0] * tx + r.m[2][1] * ty + r.m[2][2] * tz);
    r.m[3][0] = r.m[3][1] = r.m[3][2] = 0.0f;
    r.m[3][3] = 1.0f;
    return r;
}

M4 mul(const M4& a, const M4& b) {  // row-major product, k summed in order 0..3
    M4 r{};
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
            float s = 0.0f;
            for (int k = 0; k < 4; ++k) s += a.m[i][k] * b.m[k][j];
            r.m[i][j] = s;
        }
    return r;
}

bool near_identity(const M4& t, float eps) {
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c)
            if (!(std::fabs(t.m[r][c] - (r == c ? 1.0f : 0.0f)) <= eps)) return false;
    return true;
}

V3 xform_point(const M4& t, const V3& p) {
    const float x = t.m[0][0] * p.x + t.m[0][1] * p.y + t.m[0][2] * p.z + t.m[0][3];
    const float y = t.m[1][0] * p.x + t.m[1][1] * p.y + t.m[1][2] * p.z + t.m[1][3];
    const float z = t.m[2][0] * p.x + t.m[2][1] * p.y + t.m[2][2] * p.z + t.m[2][3];
    const float w = t.m[3][0] * p.x + t.m[3][1] * p.y + t.m[3][2] * p.z + t.m[3][3];
    return (w == 1.0f) ? v3(x, y, z) : v3(x / w, y / w, z / w);
}

V3 xform_vector(const M4& t, const V3& v) {
    return v3(t.m[0][0] * v.x + t.m[0][1] * v.y + t.m[0][2] * v.z, t.m[1][0] * v.x + t.m[1][1] * v.y + t.m[1][2] * v.z,
              t.m[2][0] * v.x + t.m[2][1] * v.y + t.m[2][2] * v.z);
}

bool box_is_empty(const Box& b) { return b.lo.x > b.hi.x || b.lo.y > b.hi.y || b.lo.z > b.hi.z; }

Box xform_bounds(const M4& t, const Box& b) {
    Box r = empty_box();
    if (box_is_empty(b)) return r;
    for (int z = 0; z <= 1; ++z)
        for (int y = 0; y <= 1; ++y)
            for (int x = 0; x <= 1; ++x) {
                const V3 c = v3(x ? b.hi.x : b.lo.x, y ? b.hi.y : b.lo.y, z ? b.hi.z : b.lo.z);
                const V3 p = xform_point(t, c);
                r.lo = v3(std::fmin(r.lo.x, p.x), std::fmin(r.lo.y, p.y), std::fmin(r.lo.z, p.z));
                r.hi = v3(std::fmax(r.hi.x, p.x), std::fmax(r.hi.y, p.y), std::fmax(r.hi.z, p.z));
            }
    return r;
}

Box tri_box(const Tri& t) { return Box{vmin(vmin(t.v0, t.v1), t.v2), vmax(vmax(t.v0, t.v1), t.v2)}; }

struct TwoLevel {
    std::vector<Tri> tris;          // all meshes' triangles, object space
    std::vector<int32_t> tri_off;   // mesh_count + 1
    std::vector<std::unique_ptr<Tree>> blas;
    std::unique_ptr<Tree> tlas;
    std::vector<int32_t> inst_mesh, first;
    std::vector<M4> inv;
};

// intersect_cpu_bvh_nodes with a primitive callback: pop order, strict t
// shrink, 64-entry stack (an overflow abandons the walk).
template <typename F>
bool stack_walk(const Tree& t, Ray r, F&& prim_hit, int& id, float& th, float& uh, float& vh, bool& overflow) {
    constexpr int kCap = 64;
    int stack[kCap];
    int sp = 0;
    id = -1;
    th = r.t_max;
    uh = vh = 0.0f;
    if (t.n <= 0) return false;
    const int ni = t.n - 1;
    stack[sp++] = 0;
    bool found = false;
    while (sp > 0) {
        const int h = stack[--sp];
        if (!slab(t.bounds[h], r)) continue;
        if (h >= ni) {
            const int prim = t.leaf_ids[h - ni];
            float pt = INFINITY, pu = 0, pv = 0;
            int pid = -1;
            if (!prim_hit(prim, r, pt, pu, pv, pid)) continue;
            if (pt >= th) continue;
            found = true;
            th = pt;
            uh = pu;
            vh = pv;
            id = pid >= 0 ? pid : prim;
            r.t_max = pt;
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

}  // namespace

extern "C" {

void orc_inverse_affine(const float* m16, float* out16) {
    const M4 r = inverse_affine(load_m4(m16));
    std::memcpy(out16, r.m, sizeof(r.m));
}

void orc_transform_bounds(const float* m16, const float* box6, float* out6) {
    Box b;
    std::memcpy(&b, box6, sizeof(Box));
    const Box r = xform_bounds(load_m4(m16), b);
    std::memcpy(out6, &r, sizeof(Box));
}

// The scene validation of one instance transform pair: both matrices affine
// with |det| > 1e-12, and m * inv, inv * m within 1e-4 of the identity.
int32_t orc_instance_transform_valid(const float* m16, const float* inv16) {
    const M4 m = load_m4(m16), inv = load_m4(inv16);
    if (!affine_invertible(m) || !affine_invertible(inv)) return 0;
    return near_identity(mul(m, inv), 1e-4f) && near_identity(mul(inv, m), 1e-4f) ? 1 : 0;
}

struct orc_tl;

// tris: all meshes' triangles (9 floats each, object space), mesh m owning
// [tri_off[m], tri_off[m+1]); instances: mesh id, inverse matrix (16 floats),
// world bounds (6 floats).
orc_tl* orc_tl_build(const float* tris, const int32_t* tri_off, int32_t mesh_count, const int32_t* inst_mesh,
                     const float* inst_inv, const float* inst_world_bounds, int32_t instance_count) {
    auto* s = new TwoLevel();
    const int total = tri_off[mesh_count];
    s->tris.assign(reinterpret_cast<const Tri*>(tris), reinterpret_cast<const Tri*>(tris) + total);
    s->tri_off.assign(tri_off, tri_off + mesh_count + 1);
    for (int m = 0; m < mesh_count; ++m) {
        const int n = tri_off[m + 1] - tri_off[m];
        std::vector<Box> boxes(n);
        std::vector<int32_t> ids(n);
        for (int i = 0; i < n; ++i) {
            boxes[i] = tri_box(s->tris[tri_off[m] + i]);
            ids[i] = i;
        }
        s->blas.emplace_back(n > 0 ? build(boxes.data(), ids.data(), n) : new Tree());
    }
    std::vector<Box> wb;
    std::vector<int32_t> wid;
    int cursor = 0;
    for (int i = 0; i < instance_count; ++i) {
        const int mesh = inst_mesh[i];
        s->inst_mesh.push_back(mesh);
        s->inv.push_back(load_m4(inst_inv + 16 * i));
        const bool valid = mesh >= 0 && mesh < mesh_count;
        s->first.push_back(valid ? cursor : -1);
        if (valid) cursor += tri_off[mesh + 1] - tri_off[mesh];
        Box b;
        std::memcpy(&b, inst_world_bounds + 6 * i, sizeof(Box));
        if (box_is_empty(b)) continue;
        wb.push_back(b);
        wid.push_back(i);
    }
    s->tlas.reset(wb.empty() ? new Tree() : build(wb.data(), wid.data(), int(wb.size())));
    return reinterpret_cast<orc_tl*>(s);
}

void orc_tl_free(orc_tl* h) { delete reinterpret_cast<TwoLevel*>(h); }

int32_t orc_tl_first_primitive(const orc_tl* h, int32_t instance) {
    return reinterpret_cast<const TwoLevel*>(h)->first[instance];
}

// One closest hit per ray; ignore[i] >= 0 skips that scene ordinal. Outputs
// the scene ordinal (-1 = none), t, u, v, instance and local triangle id.
// Returns the number of walks abandoned on stack overflow.
int32_t orc_tl_trace(const orc_tl* h, const float* rays, const int32_t* ignore, int32_t n, int32_t* prim,
                     float* t_out, float* u_out, float* v_out, int32_t* inst_out, int32_t* local_out) {
    const auto& s = *reinterpret_cast<const TwoLevel*>(h);
    const Ray* rs = reinterpret_cast<const Ray*>(rays);
    int32_t overflows = 0;
    for (int i = 0; i < n; ++i) {
        const int ig = ignore ? ignore[i] : -1;
        bool overflow = false;
        int inst_hit = -1, local_hit = -1;
        auto instance_hit = [&](int inst, const Ray& ray, float& pt, float& pu, float& pv, int& pid) {
            if (overflow || inst < 0 || inst >= int(s.inst_mesh.size())) return false;
            const int mesh = s.inst_mesh[inst];
            if (mesh < 0 || mesh >= int(s.blas.size())) return false;
            const Tree& blas = *s.blas[mesh];
            if (blas.n <= 0) return false;
            const M4& inv = s.inv[inst];
            Ray obj = ray;
            obj.o = xform_point(inv, ray.o);
            obj.d = xform_vector(inv, ray.d);
            const Tri* mesh_tris = s.tris.data() + s.tri_off[mesh];
            const int first = s.first[inst];
            int lid = -1;
            float lt = 0, lu = 0, lv = 0;
            bool blas_overflow = false;
            const bool found = stack_walk(
                blas, obj,
                [&](int local, const Ray& lr, float& qt, float& qu, float& qv, int& qid) {
                    if (local < 0) return false;
                    if (ig >= 0 && first >= 0 && first + local == ig) return false;
                    qid = local;
                    return tri_hit(lr, mesh_tris[local], qt, qu, qv);
                },
                lid, lt, lu, lv, blas_overflow);
            if (blas_overflow) {
                overflow = true;
                return false;
            }
            if (!found || first < 0) return false;
            pt = lt;
            pu = lu;
            pv = lv;
            pid = first + lid;
            return true;
        };
        int id = -1;
        float th = 0, uh = 0, vh = 0;
        bool tlas_overflow = false;
        bool found = stack_walk(*s.tlas, rs[i], instance_hit, id, th, uh, vh, tlas_overflow);
        if (overflow || tlas_overflow) {
            ++overflows;
            found = false;
        }
        if (found) {  // the instance owning the ordinal (scene_primitive_ref)
            for (int k = 0; k < int(s.first.size()); ++k) {
                const int mesh = s.inst_mesh[k];
                if (s.first[k] < 0) continue;
                const int cnt = s.tri_off[mesh + 1] - s.tri_off[mesh];
                if (id >= s.first[k] && id < s.first[k] + cnt) {
                    inst_hit = k;
                    local_hit = id - s.first[k];
                    break;
                }
            }
        }
        prim[i] = found ? id : -1;
        t_out[i] = found ? th : INFINITY;
        u_out[i] = found ? uh : 0.0f;
        v_out[i] = found ? vh : 0.0f;
        inst_out[i] = found ? inst_hit : -1;
        local_out[i] = found ? local_hit : -1;
    }
    return overflows;
}

}  // extern "C"

// two_level.hpp -- two-level (instanced) scenes: one LBVH per mesh over its
// object-space triangles (BLAS) and one over the instances' world bounds
// (TLAS), the layout of the reference's PreparedScene
// (src/scene/prepared_scene.cu:59-78, src/scene/scene_prepare.h:420-462).
#pragma once

#include <vector>

#include "lux_runtime.hpp"
#include "ordered.hpp"

namespace lux {

// One mesh as the walk sees it: its BLAS (sibling-pair node copy) and its
// triangles (object space, a slice of the scene's soup).
struct DevMesh {
    const lux_bvh_node* nodes;
    const lux_triangle* tris;
    const uint32_t* slot_pre;  // preorder position of each pair slot (tie order)
    int32_t node_count;        // sibling-pair slots (0: empty mesh)
    int32_t tri_count;
};

// One instance: Transform::inv (row-major 4x4), the scene ordinal of its local
// triangle 0 (GpuSceneInstance::first_primitive, src/scene/gpu_scene.h:149-172)
// and, copied from its mesh so that entering the instance is one dependent load,
// the mesh's BLAS (sibling-pair slots; 0 for an empty or invalid mesh) and the
// offset of its triangles in the scene's soup.
struct DevInstance {
    float inv[16];
    const lux_bvh_node* nodes;
    int32_t node_count;
    int32_t tri_base;
    int32_t tri_count;
    int32_t first;
    int32_t mesh;
    int32_t pad;
};
static_assert(sizeof(DevInstance) == 96, "instance record");

// What the walk kernels take by value.
struct TwoLevelView {
    const lux_bvh_node* tlas_nodes;  // the TLAS (sibling-pair copy) with the reference's boxes
    const lux_bvh_node* tlas_cull;   // the same slots with tighter culling boxes (see TwoLevelDev)
    const uint32_t* tlas_slot_pre;
    const DevInstance* instances;
    const DevMesh* meshes;
    const lux_triangle* soup;  // every mesh's object-space triangles, mesh by mesh
    DevInfo* info;
    int32_t tlas_node_count, instance_count, mesh_count;
    uint32_t step_limit;  // TLAS slots + the BLAS slots of every instance: no walk takes more steps
};

// ---- device walk ------------------------------------------------------------------
// transform_point / transform_vector (src/core/mat4.cuh:173-203): row-major
// rows summed left to right; a point divides by w unless w == 1.
__device__ __forceinline__ f3 xform_point(const float* m, const f3& p) {
    const float x = m[0] * p.x + m[1] * p.y + m[2] * p.z + m[3];
    const float y = m[4] * p.x + m[5] * p.y + m[6] * p.z + m[7];
    const float z = m[8] * p.x + m[9] * p.y + m[10] * p.z + m[11];
    const float w = m[12] * p.x + m[13] * p.y + m[14] * p.z + m[15];
    return (w == 1.0f) ? make3(x, y, z) : make3(x / w, y / w, z / w);
}

__device__ __forceinline__ f3 xform_vector(const float* m, const f3& v) {
    return make3(m[0] * v.x + m[1] * v.y + m[2] * v.z, m[4] * v.x + m[5] * v.y + m[6] * v.z,
                 m[8] * v.x + m[9] * v.y + m[10] * v.z);
}

struct TwoLevelHit {
    int prim = -1, instance = -1, local = -1;
    float t = 0, u = 0, v = 0;
};

// The walk is one flat loop over two levels. A lane starts in the TLAS (world
// ray); at an instance leaf it switches to that mesh's BLAS with the ray moved
// to object space and keeps the TLAS escape of the leaf; when the BLAS walk
// ends it returns to that escape with the world ray. One loop keeps the lanes
// of a wave in step whichever level each is on. t_max is shared by both
// levels: the reference gives the BLAS walk t_max = the current hit and keeps a
// BLAS hit only when it is nearer, which every hit the BLAS walk accepts is
// (it only accepts t < t_max), so updating t_max at once is the same.
// kAny stops at the first accepted triangle. `ignore` is a scene ordinal
// (should_skip_primitive) or -1.
template <bool kAny>
__device__ __forceinline__ TwoLevelHit two_level_walk(const TwoLevelView& s, const f3& WO, const f3& WD,
                                                      float t_min, float t_max, int ignore) {
    const SlabRay wsr = make_slab_ray(WO, WD, t_min);
    SlabRay sr = wsr;
    f3 O = WO, D = WD;
    const lux_bvh_node* nodes = s.tlas_nodes;
    uint32_t nc = uint32_t(s.tlas_node_count);
    const lux_triangle* tris = nullptr;
    int tcount = 0, first = -1, inst_cur = -1, ignore_local = -1;
    bool in_blas = false;
    uint32_t resume = 0;
    uint32_t k = 0;
    uint32_t steps = s.step_limit;
    TwoLevelHit best;
    for (;;) {
        if (k >= nc) {
            if (!in_blas) break;
            in_blas = false;  // back to the TLAS at the instance leaf's escape
            nodes = s.tlas_nodes;
            nc = uint32_t(s.tlas_node_count);
            k = resume;
            sr = wsr;
            O = WO;
            D = WD;
            continue;
        }
        if (steps-- == 0u) {
            set_error(s.info, kErrWalkOrder);
            break;
        }
        float4 a, b;
        load_node(nodes, int(k), a, b);
        const int prim = __float_as_int(b.z);
        const uint32_t esc = __float_as_uint(b.w);
        // the branch-free test unless some active lane's current ray (world or
        // object space) has an axis-parallel component; a wave-uniform choice
        const bool any_parallel = wave_any_parallel(sr);
        const bool pass = any_parallel ? slab_hit(sr, t_max, a.x, a.y, a.z, a.w, b.x, b.y)
                                       : slab_hit_np(sr, t_max, a.x, a.y, a.z, a.w, b.x, b.y);
        if (!pass) {
            k = esc;
            continue;
        }
        if (prim < 0) {
            k = uint32_t(~prim);  // right child's slot
            continue;
        }
        k = esc;
        if (!in_blas) {  // instance leaf: enter its BLAS if it has one
            if (prim >= s.instance_count) continue;
            const DevInstance& in = s.instances[prim];
            if (in.node_count <= 0 || in.first < 0) continue;
            float inv[16];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const float4 row = reinterpret_cast<const float4*>(in.inv)[i];
                inv[4 * i] = row.x;
                inv[4 * i + 1] = row.y;
                inv[4 * i + 2] = row.z;
                inv[4 * i + 3] = row.w;
            }
            resume = esc;
            in_blas = true;
            nodes = in.nodes;
            nc = uint32_t(in.node_count);
            tris = s.soup + in.tri_base;
            tcount = in.tri_count;
            first = in.first;
            inst_cur = prim;
            ignore_local = (ignore >= first && ignore - first < tcount) ? ignore - first : -1;
            O = xform_point(inv, WO);
            D = xform_vector(inv, WD);
            sr = make_slab_ray(O, D, t_min);
            k = 0;
            continue;
        }
        if (prim == ignore_local) continue;  // should_skip_primitive
        if (prim >= tcount) {
            set_error(s.info, kErrPrimRange);
            continue;
        }
        const float* p = reinterpret_cast<const float*>(tris + prim);
        const f3 v0 = make3(p[0], p[1], p[2]), v1 = make3(p[3], p[4], p[5]), v2 = make3(p[6], p[7], p[8]);
        float t, u, v;
        if (tri_hit(O, D, t_min, t_max, v0, v1, v2, t, u, v)) {
            t_max = t;
            best.prim = first + prim;
            best.instance = inst_cur;
            best.local = prim;
            best.t = t;
            best.u = u;
            best.v = v;
            if (kAny) break;
        }
    }
    return best;
}

// Occlusion of the same two-level scene: two_level_walk<true>'s walk with the
// TLAS steps culled by tlas_cull. t_max never changes and the answer is whether
// any triangle is hit in (t_min, t_max): an instance whose culling box the
// segment misses holds no such hit (the box holds every ray point that can
// reach its mesh), and every other instance is walked as the reference walks
// it. The expensive steps are gathered across the wave. A lane whose
// walk reaches a passed leaf (an instance in the TLAS, a triangle in a BLAS)
// holds it and stops; once a quarter of the wave's working lanes hold one (or
// none can step on), the wave enters the held instances and tests the held
// triangles together.
__device__ __forceinline__ bool two_level_any(const TwoLevelView& s, const f3& WO, const f3& WD, float t_min,
                                              float t_max, int ignore) {
    const SlabRay wsr = make_slab_ray(WO, WD, t_min);
    SlabRay sr = wsr;
    f3 O = WO, D = WD;
    const lux_bvh_node* nodes = s.tlas_cull;
    uint32_t nc = uint32_t(s.tlas_node_count);
    int tri_base = 0, tcount = 0, ignore_local = -1;
    bool in_blas = false, found = false;
    uint32_t resume = 0, k = 0;
    uint32_t steps = s.step_limit;
    int held = -1;
    for (;;) {
        // one step of the walk, for a lane that holds nothing and is not done
        if (held < 0 && !found) {
            if (k >= nc && in_blas) {  // back to the TLAS at the instance leaf's escape
                in_blas = false;
                nodes = s.tlas_cull;
                nc = uint32_t(s.tlas_node_count);
                k = resume;
                sr = wsr;
                O = WO;
                D = WD;
            }
            if (k < nc) {
                if (steps-- == 0u) {
                    set_error(s.info, kErrWalkOrder);
                    k = nc;
                    in_blas = false;
                } else {
                    float4 a, b;
                    load_node(nodes, int(k), a, b);
                    const int prim = __float_as_int(b.z);
                    const bool pass = wave_any_parallel(sr) ? slab_hit(sr, t_max, a.x, a.y, a.z, a.w, b.x, b.y)
                                                            : slab_hit_np(sr, t_max, a.x, a.y, a.z, a.w, b.x, b.y);
                    held = (pass && prim >= 0) ? prim : -1;
                    k = (pass && prim < 0) ? uint32_t(~prim) : __float_as_uint(b.w);
                }
            }
        }
        const bool working = held < 0 && !found && (k < nc || in_blas);
        const uint64_t want = __ballot(held >= 0);
        const uint64_t more = __ballot(working);
        if (!more) {
            if (!want) break;
        } else if (__popcll(want) * 4 < __popcll(want | more)) {
            continue;
        }
        if (held < 0) continue;
        if (!in_blas) {  // instance leaf: enter its BLAS if it has one (k holds the leaf's escape)
            const int inst = held;
            held = -1;
            if (inst >= s.instance_count) continue;
            const DevInstance& in = s.instances[inst];
            if (in.node_count <= 0 || in.first < 0) continue;
            float inv[16];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const float4 row = reinterpret_cast<const float4*>(in.inv)[i];
                inv[4 * i] = row.x;
                inv[4 * i + 1] = row.y;
                inv[4 * i + 2] = row.z;
                inv[4 * i + 3] = row.w;
            }
            resume = k;
            in_blas = true;
            nodes = in.nodes;
            nc = uint32_t(in.node_count);
            tri_base = in.tri_base;
            tcount = in.tri_count;
            ignore_local = (ignore >= in.first && ignore - in.first < tcount) ? ignore - in.first : -1;
            O = xform_point(inv, WO);
            D = xform_vector(inv, WD);
            sr = make_slab_ray(O, D, t_min);
            k = 0;
            continue;
        }
        const int prim = held;
        held = -1;
        if (prim == ignore_local) continue;  // should_skip_primitive
        if (prim >= tcount) {
            set_error(s.info, kErrPrimRange);
            continue;
        }
        const float* p = reinterpret_cast<const float*>(s.soup + tri_base + prim);
        float t, u, v;
        if (tri_hit(O, D, t_min, t_max, make3(p[0], p[1], p[2]), make3(p[3], p[4], p[5]), make3(p[6], p[7], p[8]), t,
                    u, v))
            found = true;
    }
    return found;
}

// Closest hit of the same two-level scene, near child first at both levels
// (ordered.hpp). A TLAS step tests both children of a node; each child that
// passes is an internal node (its pair slot) or an instance leaf (its slot
// tagged with kInstanceItem). The nearer is taken next and the other pushed on
// the lane's LDS stack; taking an instance leaf moves the ray to object space
// and walks that BLAS near child first on the same stack above the entry
// depth, returning to the world ray when the stack falls back to it.
// The preorder walk above visits triangles in the order (TLAS preorder of the
// instance leaf, BLAS preorder of the triangle leaf) and keeps the first of
// equal t; this walk keeps, of equal t, the one smallest in that order
// (slot_pre of both levels, read only on an exact tie), so both return the
// same hit. The check of ordered.hpp spans both levels: boxes of both are
// entered up to t_cull, and a triangle's largest path entry is the larger of
// its BLAS leaf's entry (object space) and its instance's TLAS leaf entry
// (world space). Returns false when the stack would overflow, the arrays are
// malformed (error bits set) or the check fires; the caller then runs
// two_level_walk<false> for the ray.
constexpr uint32_t kInstanceItem = 0x80000000u;

__device__ __forceinline__ bool two_level_closest_ordered(const TwoLevelView& s, const f3& WO, const f3& WD,
                                                          float t_min, float t_max0, int ignore, uint32_t* stack,
                                                          int stride, TwoLevelHit& best) {
    // Per-lane state is kept small (the path tracer's trace kernel holds this
    // walk at 80 VGPRs): the world slab ray is recomputed on leaving an
    // instance, a mesh's triangle and tie arrays are read at its leaf tests,
    // and the winner's instance, local id, u and v are derived once at the end.
    best = TwoLevelHit{};
    OrderedBest ob(t_max0);
    float inst_tn = -INFINITY;  // entry of the current instance's TLAS leaf box
    const uint32_t tlas_nc = uint32_t(s.tlas_node_count);
    if (tlas_nc == 0u) return true;
    SlabRay sr = make_slab_ray(WO, WD, t_min);
    f3 O = WO, D = WD;
    const lux_bvh_node* nodes = s.tlas_cull;  // TLAS steps cull with the tighter boxes
    uint32_t nc = tlas_nc;
    bool in_blas = false;
    int tri_base = 0, first = -1, ignore_local = -1;
    int best_prim = -1;
    uint32_t inst_slot = 0, best_inst_slot = 0, best_slot = 0;
    int sp = 0, sp0 = 0;
    uint32_t steps = s.step_limit;
    uint32_t item;
    {
        float4 a, b;
        load_node32(nodes, 0, a, b);
        float tn;
        const bool par = wave_any_parallel(sr);
        const bool hit = par ? slab_near<true>(sr, ob.t_cull, a.x, a.y, a.z, a.w, b.x, b.y, tn)
                             : slab_near<false>(sr, ob.t_cull, a.x, a.y, a.z, a.w, b.x, b.y, tn);
        if (!hit) return true;
        const int p = __float_as_int(b.z);
        item = p >= 0 ? kInstanceItem : ~uint32_t(p);
    }
    // Tests BLAS leaf `slot` (local triangle prim, box entry tn) of the current
    // instance. The largest entry on the triangle's path is that of its BLAS
    // leaf or that of the instance's TLAS leaf.
    // A BLAS of nc slots has (nc + 1) / 2 leaves (nc = 1: a single leaf).
    auto leaf = [&](uint32_t slot, int prim, float tn) {
        if (prim == ignore_local) return;  // should_skip_primitive
        if (uint32_t(prim) >= (nc + 1u) / 2u) {
            set_error(s.info, kErrPrimRange);
            return;
        }
        const float* p = reinterpret_cast<const float*>(s.soup + tri_base + prim);
        const f3 v0 = make3(p[0], p[1], p[2]), v1 = make3(p[3], p[4], p[5]), v2 = make3(p[6], p[7], p[8]);
        float t, u, v;
        if (!tri_uvt(O, D, v0, v1, v2, t, u, v) || !(t > t_min)) return;
        ob.note(t, fmaxf(tn, inst_tn));
        bool take = t < ob.t_max;
        if (!take && best_prim >= 0 && t == ob.t_max) {  // exact tie: the smaller (TLAS, BLAS) preorder position
            if (inst_slot != best_inst_slot) {
                take = s.tlas_slot_pre[inst_slot] < s.tlas_slot_pre[best_inst_slot];
            } else {  // same instance: its mesh's BLAS preorder positions
                float4 ta, tb;
                load_node32(s.tlas_nodes, inst_slot, ta, tb);
                const uint32_t* pre = s.meshes[s.instances[__float_as_int(tb.z)].mesh].slot_pre;
                take = pre[slot] < pre[best_slot];
            }
        }
        if (!take) return;
        ob.take(t);
        best_prim = first + prim;
        best_inst_slot = inst_slot;
        best_slot = slot;
    };
    for (;;) {
        if (steps-- == 0u) {
            set_error(s.info, kErrWalkOrder);
            return false;
        }
        bool descend = false;
        if (item & kInstanceItem) {
            // enter the instance of TLAS leaf slot `ls`, if it has a BLAS (an
            // instance item is only taken in the TLAS, so sr is the world ray)
            const uint32_t ls = item & ~kInstanceItem;
            if (ls >= tlas_nc) {
                set_error(s.info, kErrWalkOrder);
                return false;
            }
            float4 a, b;
            load_node32(s.tlas_nodes, ls, a, b);
            const int prim = __float_as_int(b.z);
            const DevInstance* in = (prim >= 0 && prim < s.instance_count) ? &s.instances[prim] : nullptr;
            const bool ok = in && in->node_count > 0 && in->first >= 0;
            if (ok && uint32_t(in->node_count) <= kOffset32Nodes) {
                {
                    // entry of this instance's TLAS leaf box as the reference
                    // tests it (its own box, not the culling box): the exactness
                    // check weighs hits against the reference walk's entries
                    const bool wpar = wave_any_parallel(sr);
                    (void)(wpar ? slab_near<true>(sr, INFINITY, a.x, a.y, a.z, a.w, b.x, b.y, inst_tn)
                                : slab_near<false>(sr, INFINITY, a.x, a.y, a.z, a.w, b.x, b.y, inst_tn));
                }
                float inv[16];
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const float4 row = reinterpret_cast<const float4*>(in->inv)[i];
                    inv[4 * i] = row.x;
                    inv[4 * i + 1] = row.y;
                    inv[4 * i + 2] = row.z;
                    inv[4 * i + 3] = row.w;
                }
                in_blas = true;
                sp0 = sp;
                nodes = in->nodes;
                nc = uint32_t(in->node_count);
                tri_base = in->tri_base;
                first = in->first;
                inst_slot = ls;
                ignore_local = (ignore >= first && ignore - first < in->tri_count) ? ignore - first : -1;
                O = xform_point(inv, WO);
                D = xform_vector(inv, WD);
                sr = make_slab_ray(O, D, t_min);
                load_node32(nodes, 0, a, b);
                float tn;
                const bool par = wave_any_parallel(sr);
                const bool hit = par ? slab_near<true>(sr, ob.t_cull, a.x, a.y, a.z, a.w, b.x, b.y, tn)
                                     : slab_near<false>(sr, ob.t_cull, a.x, a.y, a.z, a.w, b.x, b.y, tn);
                const int root = __float_as_int(b.z);
                if (hit && root >= 0) leaf(0u, root, tn);
                if (hit && root < 0) {
                    item = ~uint32_t(root);
                    descend = true;
                }
            } else if (ok) {
                return false;  // a BLAS beyond 32-bit node offsets: the preorder walk takes the ray
            }
        } else {
            const uint32_t c = item;
            if (c + 1 >= nc) {
                set_error(s.info, kErrWalkOrder);
                return false;
            }
            float4 a0, b0, a1, b1;
            load_node32(nodes, c, a0, b0);
            load_node32(nodes, c + 1, a1, b1);
            float tn0, tn1;
            const bool par = wave_any_parallel(sr);
            const bool h0 = par ? slab_near<true>(sr, ob.t_cull, a0.x, a0.y, a0.z, a0.w, b0.x, b0.y, tn0)
                                : slab_near<false>(sr, ob.t_cull, a0.x, a0.y, a0.z, a0.w, b0.x, b0.y, tn0);
            const bool h1 = par ? slab_near<true>(sr, ob.t_cull, a1.x, a1.y, a1.z, a1.w, b1.x, b1.y, tn1)
                                : slab_near<false>(sr, ob.t_cull, a1.x, a1.y, a1.z, a1.w, b1.x, b1.y, tn1);
            const int p0 = __float_as_int(b0.z), p1 = __float_as_int(b1.z);
            bool take0, take1;
            uint32_t i0, i1;
            if (in_blas) {
                if (h0 && p0 >= 0) leaf(c, p0, tn0);
                if (h1 && p1 >= 0) leaf(c + 1, p1, tn1);
                take0 = h0 && p0 < 0;
                take1 = h1 && p1 < 0;
                i0 = ~uint32_t(p0);
                i1 = ~uint32_t(p1);
            } else {
                take0 = h0;
                take1 = h1;
                i0 = p0 >= 0 ? (c | kInstanceItem) : ~uint32_t(p0);
                i1 = p1 >= 0 ? ((c + 1) | kInstanceItem) : ~uint32_t(p1);
            }
            if (take0 && take1) {
                if (sp == kOrderedStack) return false;
                const bool far0 = tn1 < tn0;
                stack[sp * stride] = far0 ? i0 : i1;
                ++sp;
                item = far0 ? i1 : i0;
                descend = true;
            } else if (take0 || take1) {
                item = take0 ? i0 : i1;
                descend = true;
            }
        }
        if (descend) continue;
        // pop: leave the instance once its part of the stack is used up
        if (in_blas && sp == sp0) {
            in_blas = false;
            nodes = s.tlas_cull;
            nc = tlas_nc;
            sr = make_slab_ray(WO, WD, t_min);
            O = WO;
            D = WD;
        }
        if (sp == 0) break;
        --sp;
        item = stack[sp * stride];
    }
    if (best_prim < 0) return true;
    if (ob.ambiguous()) return false;
    // the winner's instance, local id, u and v: the same operations on the
    // same operands as the leaf test that took it
    float4 a, b;
    load_node32(s.tlas_nodes, best_inst_slot, a, b);
    const int inst = __float_as_int(b.z);
    const DevInstance& in = s.instances[inst];
    float inv[16];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const float4 row = reinterpret_cast<const float4*>(in.inv)[i];
        inv[4 * i] = row.x;
        inv[4 * i + 1] = row.y;
        inv[4 * i + 2] = row.z;
        inv[4 * i + 3] = row.w;
    }
    const int local = best_prim - in.first;
    const float* p = reinterpret_cast<const float*>(s.soup + in.tri_base + local);
    float t;
    (void)tri_uvt(xform_point(inv, WO), xform_vector(inv, WD), make3(p[0], p[1], p[2]), make3(p[3], p[4], p[5]),
                  make3(p[6], p[7], p[8]), t, best.u, best.v);
    best.prim = best_prim;
    best.instance = inst;
    best.local = local;
    best.t = ob.t_max;
    return true;
}


// ---- host side ----------------------------------------------------------------------
struct TwoLevelDev {
    std::vector<BvhDev> blas;  // per mesh (n = 0 for an empty mesh)
    BvhDev tlas;               // over the non-empty instance world bounds, ids = instance ids
    // The TLAS's slots again, each leaf box narrowed to its instance's world
    // bounds intersected with the world box of the mesh's bounding sphere
    // (center M c, half-extent r |row i of M| on axis i: the exact box of the
    // transformed sphere), both widened by what the instance's M inv may
    // deviate from the identity; internal boxes the unions of their children.
    // The near-child-first closest walk and the occlusion walk cull TLAS steps
    // with them.
    lux_bvh_node* tlas_cull = nullptr;
    std::vector<float> mesh_sphere;  // per mesh: center xyz, radius (padded)
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
// ordinal per ray that is skipped (-1: none). order (nullable): the rays in
// ray order, from two_level_order; results land at each ray's own index.
const uint32_t* two_level_order(RaySortDev& r, const TwoLevelDev& s, const lux_ray* rays, int n, hipStream_t stream);
void two_level_closest(const TwoLevelView& v, const uint32_t* order, const lux_ray* rays, const int32_t* ignore,
                       int n, lux_scene_hit* hits, hipStream_t stream);
void two_level_occluded(const TwoLevelView& v, const uint32_t* order, const lux_ray* rays, const int32_t* ignore,
                        int n, uint8_t* occluded, hipStream_t stream);

}  // namespace lux

// ordered.hpp -- near-child-first closest-hit walk over the sibling-pair node
// array, with the hit the right-first preorder walk would return.
//
// The preorder walk (src/bvh/bvh_device.cuh:59-87) visits every node whose box
// the ray meets before t_max shrinks to the hit, about 360 nodes per ray on
// 2^20 random triangles. This walk tests both children of a node together (one
// 64-byte sibling pair), descends into the nearer one and keeps the other on a
// short per-lane stack in LDS, so t_max shrinks early and far subtrees are cut:
// about 100 pairs per ray on the same input.
//
// Equal result. The preorder walk accepts a triangle only with t < t_max, so
// of several triangles at the same smallest t it keeps the one it reaches
// first, i.e. the one with the smallest preorder position. This walk accepts
// t < t_max, and t == t_max when the new leaf's preorder position (slot_pre)
// is smaller than the held one's, so it returns the same (t, id) pair: the
// minimum over every hit triangle of (t, preorder position) -- provided both
// walks reach the same triangles.
//
// They may not. A box is entered when its rounded slab entry t_near is at most
// the t_max held at that moment (src/core/aabb.cuh:84-109), and t_max depends
// on the visit order. In exact arithmetic a box's entry never exceeds the
// distance of a triangle inside it, but rounded, the entry of a leaf box can
// lie a few hundred ulps above its triangle's Moller-Trumbore t (a hit at a
// vertex or edge on the face the ray enters through). The preorder walk then
// culls that triangle if it already holds a t between the two, and this walk
// may not (or the other way round). OrderedBest records what decides this:
//   L       max(t_best, entry of the best triangle's leaf box),
//   m_hit   the smallest hit t of any other triangle tested,
//   m_cull  the smallest entry of a box culled only because it lay beyond the
//           held t_max (its ray segment met the box).
// With no triangle whose t sits more than L * kAmbiguityWindow below its leaf
// box's entry, the preorder walk returns this walk's hit unless m_hit < L (it
// may hold that t when it reaches the best triangle's path) or m_cull <=
// L + L * kAmbiguityWindow (a culled box may hold a triangle it reaches). Such
// a ray is walked again in preorder, as is a ray whose stack overflows. (Entries
// grow down a path: a child box lies inside its parent and rounding is
// monotone, so the leaf box holds the largest entry of the path.) The observed
// gap between a hit and its leaf box's entry is below 2^-14.7 relative; on
// 2^20 random triangles about 0.4 % of rays take the second walk.
//
// A lane whose stack would exceed kOrderedStack entries stops and reports
// overflow; its caller then runs the preorder walk for that ray. The stack
// holds at most one pending sibling per tree level, and on 2^20 random
// triangles fewer than 1 in 2000 rays go beyond 13.
#pragma once

#include "lux_common.hpp"

namespace lux {

constexpr int kOrderedStack = 16;

// Relative window of the ambiguity check (see the file comment).
constexpr float kAmbiguityWindow = 0x1p-11f;

// Closest-hit state of an ordered walk: the held bound and the ambiguity record.
struct OrderedBest {
    float t_max;             // the ray's t_max, then the best hit's t
    float L = -INFINITY;     // max(t_best, entry of the best triangle's leaf box)
    float m_hit = INFINITY;  // smallest hit t among the other tested triangles
    float m_cull = INFINITY; // smallest entry among boxes culled by t_max alone
    bool has = false;

    __device__ __forceinline__ explicit OrderedBest(float t) : t_max(t) {}
    // A hit t (u, v passed, t > t_min) the caller accepts; leaf_tn = entry of
    // its leaf box (the largest of the path's entries).
    __device__ __forceinline__ void take(float t, float leaf_tn) {
        if (has) m_hit = fminf(m_hit, t_max);
        t_max = t;
        L = fmaxf(t, leaf_tn);
        has = true;
    }
    // A hit t the caller does not accept.
    __device__ __forceinline__ void pass_over(float t) { m_hit = fminf(m_hit, t); }
    // True when the preorder walk might return another hit.
    __device__ __forceinline__ bool ambiguous() const {
        return has && (m_hit < L || m_cull <= L + fabsf(L) * kAmbiguityWindow);
    }
};

// Slab test (slab_hit / slab_hit_np) that also returns the entry distance
// t_near and records in m_cull the entry of a box the ray's segment meets but
// that lies beyond t_max. The box passes exactly when AABB::intersect passes:
// the min over t_max and the three exits equals the min over the exits
// compared with t_max separately.
template <bool kParallel>
__device__ __forceinline__ bool slab_near(const SlabRay& r, float t_max, float lox, float loy, float loz, float hix,
                                          float hiy, float hiz, float& t_near, float& m_cull) {
    bool geo;
    float tn;
    if (!kParallel) {
        const float ax = ((r.neg_x ? hix : lox) - r.o.x) * r.inv.x;
        const float cx = ((r.neg_x ? lox : hix) - r.o.x) * r.inv.x;
        const float ay = ((r.neg_y ? hiy : loy) - r.o.y) * r.inv.y;
        const float cy = ((r.neg_y ? loy : hiy) - r.o.y) * r.inv.y;
        const float az = ((r.neg_z ? hiz : loz) - r.o.z) * r.inv.z;
        const float cz = ((r.neg_z ? loz : hiz) - r.o.z) * r.inv.z;
        tn = fmaxf(fmaxf(fmaxf(r.t_min, ax), ay), az);
        const float tf = fminf(fminf(cx, cy), cz);
        geo = !(tn > tf);
    } else {
        bool miss = false;
        float tf = INFINITY;
        tn = r.t_min;
        {
            const float a = ((r.neg_x ? hix : lox) - r.o.x) * r.inv.x;
            const float c = ((r.neg_x ? lox : hix) - r.o.x) * r.inv.x;
            if (r.par_x) {
                miss |= (r.o.x < lox) | (r.o.x > hix);
            } else {
                tn = fmaxf(tn, a);
                tf = fminf(tf, c);
            }
        }
        {
            const float a = ((r.neg_y ? hiy : loy) - r.o.y) * r.inv.y;
            const float c = ((r.neg_y ? loy : hiy) - r.o.y) * r.inv.y;
            if (r.par_y) {
                miss |= (r.o.y < loy) | (r.o.y > hiy);
            } else {
                tn = fmaxf(tn, a);
                tf = fminf(tf, c);
            }
        }
        {
            const float a = ((r.neg_z ? hiz : loz) - r.o.z) * r.inv.z;
            const float c = ((r.neg_z ? loz : hiz) - r.o.z) * r.inv.z;
            if (r.par_z) {
                miss |= (r.o.z < loz) | (r.o.z > hiz);
            } else {
                tn = fmaxf(tn, a);
                tf = fminf(tf, c);
            }
        }
        geo = !miss && !(tn > tf);
    }
    t_near = tn;
    const bool beyond = tn > t_max;
    if (geo && beyond) m_cull = fminf(m_cull, tn);
    return geo && !beyond;
}

// Walks the pair array from the root. leaf(slot, prim, tn) tests one leaf the
// ray's box test passed (tn = its box's entry) and may lower best.t_max (it
// holds the tie rule above); it returns true to end the walk (an any-hit
// query). stack points at this lane's first LDS entry; entries are `stride`
// words apart. Returns false when the stack overflowed (the walk stopped
// early); *bad is set on a malformed array (child slot out of range, a pair
// entered twice).
template <bool kParallel, typename Leaf>
__device__ __forceinline__ bool ordered_walk(const lux_bvh_node* __restrict__ nodes, uint32_t node_count,
                                             const SlabRay& sr, OrderedBest& best, uint32_t* stack, int stride,
                                             uint32_t& visits, bool& bad, Leaf&& leaf) {
    float4 a, b;
    load_node32(nodes, 0, a, b);
    ++visits;
    float tn;
    if (!slab_near<kParallel>(sr, best.t_max, a.x, a.y, a.z, a.w, b.x, b.y, tn, best.m_cull)) return true;
    const int root = __float_as_int(b.z);
    if (root >= 0) {
        (void)leaf(0u, root, tn);
        return true;
    }
    uint32_t c = ~uint32_t(root);
    int sp = 0;
    int guard = int(node_count / 2) + 1;  // each internal node's pair is entered once
    for (;;) {
        if (c + 1 >= node_count || --guard < 0) {
            bad = true;
            return true;
        }
        float4 a0, b0, a1, b1;
        load_node32(nodes, c, a0, b0);
        load_node32(nodes, c + 1, a1, b1);
        visits += 2;
        float tn0, tn1;
        const bool h0 = slab_near<kParallel>(sr, best.t_max, a0.x, a0.y, a0.z, a0.w, b0.x, b0.y, tn0, best.m_cull);
        const bool h1 = slab_near<kParallel>(sr, best.t_max, a1.x, a1.y, a1.z, a1.w, b1.x, b1.y, tn1, best.m_cull);
        const int p0 = __float_as_int(b0.z), p1 = __float_as_int(b1.z);
        if (h0 && p0 >= 0 && leaf(c, p0, tn0)) return true;
        if (h1 && p1 >= 0 && leaf(c + 1, p1, tn1)) return true;
        const bool i0 = h0 && p0 < 0, i1 = h1 && p1 < 0;
        if (i0 && i1) {
            const bool far0 = tn1 < tn0;
            if (sp == kOrderedStack) return false;
            stack[sp * stride] = far0 ? ~uint32_t(p0) : ~uint32_t(p1);
            ++sp;
            c = far0 ? ~uint32_t(p1) : ~uint32_t(p0);
        } else if (i0 || i1) {
            c = ~uint32_t(i0 ? p0 : p1);
        } else {
            if (sp == 0) return true;
            --sp;
            c = stack[sp * stride];
        }
    }
}

}  // namespace lux

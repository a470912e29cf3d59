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
// minimum over every hit triangle of (t, preorder position). Box tests use the
// same slab arithmetic (a box is entered when t_near <= t_max), so a triangle
// the preorder walk reaches is reached here too.
//
// A lane whose stack would exceed kOrderedStack entries stops and reports
// overflow; its caller then runs the preorder walk for that ray. The stack
// holds at most one pending sibling per tree level, and on 2^20 random
// triangles fewer than 1 in 2000 rays go beyond 13.
#pragma once

#include "lux_common.hpp"

namespace lux {

constexpr int kOrderedStack = 16;

// slab_hit / slab_hit_np that also return the entry distance t_near.
template <bool kParallel>
__device__ __forceinline__ bool slab_near(const SlabRay& r, float t_max, float lox, float loy, float loz, float hix,
                                          float hiy, float hiz, float& t_near) {
    if (!kParallel) {
        const float ax = ((r.neg_x ? hix : lox) - r.o.x) * r.inv.x;
        const float cx = ((r.neg_x ? lox : hix) - r.o.x) * r.inv.x;
        const float ay = ((r.neg_y ? hiy : loy) - r.o.y) * r.inv.y;
        const float cy = ((r.neg_y ? loy : hiy) - r.o.y) * r.inv.y;
        const float az = ((r.neg_z ? hiz : loz) - r.o.z) * r.inv.z;
        const float cz = ((r.neg_z ? loz : hiz) - r.o.z) * r.inv.z;
        t_near = fmaxf(fmaxf(fmaxf(r.t_min, ax), ay), az);
        const float t_far = fminf(fminf(fminf(t_max, cx), cy), cz);
        return !(t_near > t_far);
    }
    bool miss = false;
    float tn = r.t_min, tf = t_max;
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
    t_near = tn;
    return !miss && !(tn > tf);
}

// Upper bound for a leaf's triangle test: t_max itself while nothing is held
// (the ray's own t_max is exclusive), else the next float above it, so a
// triangle at exactly t_max comes back and the preorder tie rule decides.
__device__ __forceinline__ float tie_bound(float t_max, bool holding) {
    return holding ? nextafterf(t_max, INFINITY) : t_max;
}

// Walks the pair array from the root. leaf(slot, prim) tests one leaf the ray's
// box test passed and may lower t_max (it holds the tie rule above); it
// returns true to end the walk (an any-hit query). stack points at this lane's
// first LDS entry; entries are `stride` words apart. Returns false when the
// stack overflowed (the walk stopped early); *bad is set on a malformed array
// (child slot out of range, a pair entered twice).
template <bool kParallel, typename Leaf>
__device__ __forceinline__ bool ordered_walk(const lux_bvh_node* __restrict__ nodes, uint32_t node_count,
                                             const SlabRay& sr, const float& t_max, uint32_t* stack, int stride,
                                             uint32_t& visits, bool& bad, Leaf&& leaf) {
    float4 a, b;
    load_node32(nodes, 0, a, b);
    ++visits;
    float tn;
    if (!slab_near<kParallel>(sr, t_max, a.x, a.y, a.z, a.w, b.x, b.y, tn)) return true;
    const int root = __float_as_int(b.z);
    if (root >= 0) {
        (void)leaf(0u, root);
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
        const bool h0 = slab_near<kParallel>(sr, t_max, a0.x, a0.y, a0.z, a0.w, b0.x, b0.y, tn0);
        const bool h1 = slab_near<kParallel>(sr, t_max, a1.x, a1.y, a1.z, a1.w, b1.x, b1.y, tn1);
        const int p0 = __float_as_int(b0.z), p1 = __float_as_int(b1.z);
        if (h0 && p0 >= 0 && leaf(c, p0)) return true;
        if (h1 && p1 >= 0 && leaf(c + 1, p1)) return true;
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

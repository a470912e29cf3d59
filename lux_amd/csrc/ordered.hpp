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
// may not (or the other way round). So this walk
//   * enters boxes up to t_cull = t_max + |t_max| * kSoftWindow once it holds
//     a hit (up to the ray's own t_max before), so every triangle whose t is
//     at most the final t_best gets tested, as long as no triangle's t lies
//     more than that window below its own leaf box's entry;
//   * records in m the smallest t <= t_max of a tested triangle whose t lies
//     below its own leaf box's entry (the largest entry on its path: a child
//     box lies inside its parent and rounding is monotone);
//   * hands the ray to the preorder walk when it ends holding a hit with
//     m <= t_best.
// Otherwise the preorder walk returns the same hit: it accepts a triangle iff
// the triangle's leaf box passes (entry <= held t) and t < held t, in
// preorder; every hit with t <= t_best was tested here and lies at or beyond
// its leaf box's entry, so the preorder walk reaches the winner (t_best,
// smallest preorder position) holding a larger t, accepts it, and nothing
// after it can win. The observed gap between a hit and its leaf box's entry is
// below 2^-14.7 relative (sampled hits next to vertices and edges); on 2^18
// random triangles and on a displaced icosphere no random or secondary ray
// needed the second walk.
//
// A lane whose stack would exceed kOrderedStack entries stops and reports
// overflow; its caller then runs the preorder walk for that ray. The stack
// holds at most one pending sibling per tree level, and on 2^20 random
// triangles fewer than 1 in 2000 rays go beyond 13.
#pragma once

#include "lux_common.hpp"

namespace lux {

constexpr int kOrderedStack = 16;

// The gathered walks test held leaves once 1/kGatherDiv of the wave's working
// lanes hold some: on 2^22 random rays over 2^20 triangles k = 2 / 3 / 4 / 6 / 8
// gave 772 / 796 / 802 / 791 / 784 Mrays/s for the microbench step.
constexpr int kGatherDiv = 4;

// Relative margin past t_max up to which boxes are still entered (see above).
constexpr float kSoftWindow = 0x1p-11f;

// Closest-hit state of an ordered walk: the held bound, the box-entry bound
// and the record of tested hits below their own leaf box's entry.
struct OrderedBest {
    float t_max;         // the ray's t_max, then the best hit's t
    float t_cull;        // boxes are entered up to here
    float m = INFINITY;  // smallest t <= t_max of a tested hit below its leaf entry

    __device__ __forceinline__ explicit OrderedBest(float t) : t_max(t), t_cull(t) {}
    // Every tested hit t (u, v passed, t > t_min), before the caller decides
    // to take it; path_tn = the largest box entry on its path.
    __device__ __forceinline__ void note(float t, float path_tn) {
        if (t <= t_max && path_tn > t) m = fminf(m, t);
    }
    // A hit the caller accepts.
    __device__ __forceinline__ void take(float t) {
        t_max = t;
        t_cull = t + fabsf(t) * kSoftWindow;
    }
    // With a hit held: true when the preorder walk might return another one.
    __device__ __forceinline__ bool ambiguous() const { return m <= t_max; }
};

// Slab test (slab_hit / slab_hit_np) against [t_min, t_bound] that also
// returns the entry distance t_near; with t_bound = t_max it equals
// AABB::intersect's boolean (the min over t_max and the three exits compared
// with t_near is the two comparisons below).
template <bool kParallel>
__device__ __forceinline__ bool slab_near(const SlabRay& r, float t_bound, float lox, float loy, float loz, float hix,
                                          float hiy, float hiz, float& t_near) {
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
    return geo && !(tn > t_bound);
}

// slab_near in a wave-uniform slab mode (lux_common.hpp): kSlabParallel is
// slab_near<true>; the other modes are slab_near<false> with the direction
// signs read per lane (kSlabLane) or fixed by the wave's octant.
template <int kMode>
__device__ __forceinline__ bool slab_near(SlabMode<kMode>, const SlabRay& r, float t_bound, float lox, float loy,
                                          float loz, float hix, float hiy, float hiz, float& t_near) {
    if constexpr (kMode == kSlabParallel) {
        return slab_near<true>(r, t_bound, lox, loy, loz, hix, hiy, hiz, t_near);
    } else {
        const bool nx = kMode < 0 ? r.neg_x : (kMode & 1) != 0;
        const bool ny = kMode < 0 ? r.neg_y : (kMode & 2) != 0;
        const bool nz = kMode < 0 ? r.neg_z : (kMode & 4) != 0;
        const float ax = ((nx ? hix : lox) - r.o.x) * r.inv.x;
        const float cx = ((nx ? lox : hix) - r.o.x) * r.inv.x;
        const float ay = ((ny ? hiy : loy) - r.o.y) * r.inv.y;
        const float cy = ((ny ? loy : hiy) - r.o.y) * r.inv.y;
        const float az = ((nz ? hiz : loz) - r.o.z) * r.inv.z;
        const float cz = ((nz ? loz : hiz) - r.o.z) * r.inv.z;
        const float tn = fmaxf(fmaxf(fmaxf(r.t_min, ax), ay), az);
        const float tf = fminf(fminf(cx, cy), cz);
        t_near = tn;
        return !(tn > tf) && !(tn > t_bound);
    }
}

// Walks the pair array from the root, entering boxes up to best.t_cull.
// leaf(slot, prim, tn) tests one leaf whose box passed (tn = its box's entry)
// and may lower best.t_max (it holds the tie rule above); it returns true to
// end the walk (an any-hit query). stack points at this lane's first LDS entry; entries are `stride`
// words apart. Returns false when the stack overflowed (the walk stopped
// early); *bad is set on a malformed array (child slot out of range, a pair
// entered twice).
template <int kMode, typename Leaf>
__device__ __forceinline__ bool ordered_walk(SlabMode<kMode> mode, const lux_bvh_node* __restrict__ nodes,
                                             uint32_t node_count, const SlabRay& sr, OrderedBest& best,
                                             uint32_t* stack, int stride, uint32_t& visits, bool& bad, Leaf&& leaf) {
    float4 a, b;
    load_node32(nodes, 0, a, b);
    ++visits;
    float tn;
    if (!slab_near(mode, sr, best.t_cull, a.x, a.y, a.z, a.w, b.x, b.y, tn)) return true;
    const int root = __float_as_int(b.z);
    if (root >= 0) {
        (void)leaf(0u, root, tn);
        return true;
    }
    uint32_t c = ~uint32_t(root);
    int sp = 0;
    int guard = int(node_count / 2) + 1;  // each internal node's pair is entered once
    const uint32_t last_pair = node_count - 1;  // a pair slot c needs c + 1 < node_count
    const lux_bvh_node* const nodes1 = nodes + 1;  // both nodes of a pair load from one offset
    for (;;) {
        if (c >= last_pair || --guard < 0) {
            bad = true;
            return true;
        }
        float4 a0, b0, a1, b1;
        load_node32(nodes, c, a0, b0);
        load_node32(nodes1, c, a1, b1);
        visits += 2;
        float tn0, tn1;
        const bool h0 = slab_near(mode, sr, best.t_cull, a0.x, a0.y, a0.z, a0.w, b0.x, b0.y, tn0);
        const bool h1 = slab_near(mode, sr, best.t_cull, a1.x, a1.y, a1.z, a1.w, b1.x, b1.y, tn1);
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

// ordered_walk with the triangle tests gathered across the wave: a lane whose
// step passes one or both leaf children holds them and does not step again
// until they are tested (its next slab tests depend on the t_cull they may
// lower); the wave tests the held leaves together once a quarter of its
// working lanes hold some, or none can step on. Each lane's sequence of box
// tests, leaf tests and stack operations is that of ordered_walk.
template <int kMode, typename Leaf>
__device__ __forceinline__ bool ordered_walk_gathered(SlabMode<kMode> mode, const lux_bvh_node* __restrict__ nodes,
                                                      uint32_t node_count, const SlabRay& sr, OrderedBest& best,
                                                      uint32_t* stack, int stride, uint32_t& visits, bool& bad,
                                                      Leaf&& leaf) {
    float4 a, b;
    load_node32(nodes, 0, a, b);
    ++visits;
    float tn;
    if (!slab_near(mode, sr, best.t_cull, a.x, a.y, a.z, a.w, b.x, b.y, tn)) return true;
    const int root = __float_as_int(b.z);
    if (root >= 0) {
        (void)leaf(0u, root, tn);
        return true;
    }
    uint32_t c = ~uint32_t(root);
    int sp = 0;
    int guard = int(node_count / 2) + 1;  // each internal node's pair is entered once
    const uint32_t last_pair = node_count - 1;     // a pair slot c needs c + 1 < node_count
    const lux_bvh_node* const nodes1 = nodes + 1;  // both nodes of a pair load from one offset
    bool done = false, overflow = false;
    bool ended = false;         // a leaf ended the walk (its later held leaves are not tested)
    uint32_t hc = 0;            // pair slot of the held leaves
    int hp0 = -1, hp1 = -1;     // held leaf children (primitive ids), -1: none
    float htn0 = 0, htn1 = 0;   // their box entries
    for (;;) {
        if (!done & ((hp0 & hp1) < 0)) {  // holds no leaf (both words negative); && measured 5 % slower
            if (c >= last_pair || --guard < 0) {
                bad = true;
                done = true;
            } else {
                float4 a0, b0, a1, b1;
                load_node32(nodes, c, a0, b0);
                load_node32(nodes1, c, a1, b1);
                visits += 2;
                float tn0, tn1;
                const bool h0 = slab_near(mode, sr, best.t_cull, a0.x, a0.y, a0.z, a0.w, b0.x, b0.y, tn0);
                const bool h1 = slab_near(mode, sr, best.t_cull, a1.x, a1.y, a1.z, a1.w, b1.x, b1.y, tn1);
                const int p0 = __float_as_int(b0.z), p1 = __float_as_int(b1.z);
                hc = c;
                hp0 = (h0 & (p0 >= 0)) ? p0 : -1;
                hp1 = (h1 & (p1 >= 0)) ? p1 : -1;
                htn0 = tn0;
                htn1 = tn1;
                const bool i0 = h0 & (p0 < 0), i1 = h1 & (p1 < 0);
                if (i0 & i1) {
                    if (sp == kOrderedStack) {
                        overflow = true;
                        done = true;
                    } else {
                        const bool far0 = tn1 < tn0;
                        stack[sp * stride] = far0 ? ~uint32_t(p0) : ~uint32_t(p1);
                        ++sp;
                        c = far0 ? ~uint32_t(p1) : ~uint32_t(p0);
                    }
                } else if (i0 | i1) {
                    c = ~uint32_t(i0 ? p0 : p1);
                } else if (sp == 0) {
                    done = true;
                } else {
                    --sp;
                    c = stack[sp * stride];
                }
            }
        }
        const bool holds = (hp0 >= 0) | (hp1 >= 0);
        const uint64_t want = __ballot(holds);
        const uint64_t more = __ballot(!done & !holds);
        if (!more) {
            if (!want) break;
        } else if (__popcll(want) * kGatherDiv < __popcll(want | more)) {
            continue;
        }
        // one copy of the leaf test, run for the first then the second child; a
        // leaf that returns true ends the walk (an any-hit query)
#pragma unroll 1
        for (int j = 0; j < 2; ++j) {
            const int pj = j ? hp1 : hp0;
            if (pj >= 0 && !ended && leaf(hc + uint32_t(j), pj, j ? htn1 : htn0)) ended = done = true;
        }
        hp0 = hp1 = -1;
    }
    return !overflow;
}

}  // namespace lux

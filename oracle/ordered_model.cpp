// ordered_model.cpp -- CPU model of the library's near-child-first closest-hit
// walk (lux_amd/csrc/ordered.hpp), used by the tests to show where that walk
// can part from the right-first preorder walk of the host LBVH and that the
// library's ambiguity check closes the gap.
//
// TEST INFRASTRUCTURE ONLY, like the rest of oracle/: it models the product's
// walk so a test can construct the rounding case on the CPU; the product never
// calls it.
//
// The preorder walk (intersect_cpu_bvh_nodes, src/bvh/cpu_bvh_traversal.h:17-81)
// tests a box against the t_max it holds at that moment (src/core/aabb.cuh:84-109).
// In exact arithmetic a box's entry distance never exceeds the distance of a
// triangle inside it, but the rounded slab entry can exceed the rounded
// Moller-Trumbore t by a few hundred ulps (hits at a vertex or edge on the box
// face the ray enters through). Such a triangle is found or not depending on
// the t_max held when its box is tested, and the near-child-first order holds
// a different t_max than the preorder order. Rule 0 below is the plain
// near-child-first walk (it can return another hit than the preorder walk);
// rule 1 is the library's checked walk:
//   * after the first hit, boxes are entered up to t_cull = t_max + |t_max| *
//     2^-window_log2 instead of t_max, so every triangle whose t is <= the
//     final t_best is tested (as long as no triangle's t lies more than that
//     window below its own leaf box's entry);
//   * a tested hit whose t lies below its own leaf box's entry (the largest
//     entry on its path) is "inconsistent"; m is the smallest t of those;
//   * the ray is walked again in preorder when m <= t_best.
// Why this is exact: the preorder walk accepts triangle T iff its leaf box
// passes (entry <= held t) and t_T < held t, in preorder. If every hit with
// t <= t_best is tested and none is inconsistent, the tested minimum (t, pre)
// is reached with a held t above its t, so the preorder walk accepts it and
// nothing after it can win.

#include "oracle_core.hpp"

using namespace orc;

namespace {

constexpr int kStackCap = 16;  // per-lane LDS stack entries of the library walk

// Slab test of the library walk: entry distance tn, box-only exit distance
// tf_box, and whether the box passes against t_max (equal to AABB::intersect's
// boolean: the min over t_max and the three exits is exact).
struct SlabOut {
    bool geo;   // passes with t_max = +inf
    bool hit;   // passes with the held t_max
    float tn;
};

SlabOut slab_near(const Box& b, const Ray& r, float t_cull) {
    bool miss = false;
    float tn = r.t_min, tf = INFINITY;
    for (int i = 0; i < 3; ++i) {
        float d = comp(r.d, i), o = comp(r.o, i), lo = comp(b.lo, i), hi = comp(b.hi, i);
        if (fabsf(d) < 1e-12f) {
            miss |= (o < lo) | (o > hi);
            continue;
        }
        float inv = 1 / d;
        float a = ((inv < 0 ? hi : lo) - o) * inv;
        float c = ((inv < 0 ? lo : hi) - o) * inv;
        tn = fmaxf(tn, a);
        tf = fminf(tf, c);
    }
    SlabOut s;
    s.tn = tn;
    s.geo = !miss && !(tn > tf);
    s.hit = s.geo && !(tn > t_cull);
    return s;
}

// Moller-Trumbore in the reference's order, returning t whenever u and v pass
// (the bound test is left to the caller).
bool tri_uv_t(const Ray& r, const Tri& tri, float& t) {
    V3 e1 = sub(tri.v1, tri.v0);
    V3 e2 = sub(tri.v2, tri.v0);
    V3 h = cross(r.d, e2);
    float a = dot(e1, h);
    if (fabsf(a) < 1e-7f) return false;
    float f = 1.0f / a;
    V3 s = sub(r.o, tri.v0);
    float u = f * dot(s, h);
    if (u < 0.0f || u > 1.0f) return false;
    V3 q = cross(s, e1);
    float v = f * dot(r.d, q);
    if (v < 0.0f || u + v > 1.0f) return false;
    t = f * dot(e2, q);
    return true;
}

struct Walk {
    int best = -1;
    int best_pre = 0;
    float t_max;
    float t_cull;        // boxes are entered up to here
    float m = INFINITY;  // smallest t of a tested hit below its own leaf entry
    bool overflow = false;
};

// Near-child-first walk over the host tree: at an internal node both children
// are tested against the held t_max; leaf children are intersected at once
// (the right child first, as the library's sibling pair holds the right child
// in its first slot); of two internal children the nearer is entered and the
// farther pushed (ties enter the right child).
void ordered_walk(const Tree& t, const Tri* tris, const Ray& r, Walk& w, int rule, float window) {
    const int ni = t.n - 1;
    auto leaf = [&](int h, float tn) {
        const int prim = t.leaf_ids[h - ni];
        float tt;
        if (!tri_uv_t(r, tris[prim], tt) || !(tt > r.t_min)) return;
        if (tt <= w.t_max && tn > tt) w.m = fminf(w.m, tt);
        const int pre = t.n == 1 ? 0 : t.pre_of[h];
        const bool accept = w.best >= 0 ? (tt < w.t_max || (tt == w.t_max && pre < w.best_pre)) : tt < w.t_max;
        if (accept) {
            w.best = prim;
            w.best_pre = pre;
            w.t_max = tt;
            w.t_cull = rule == 1 ? tt + fabsf(tt) * window : tt;
        }
    };
    SlabOut root = slab_near(t.bounds[0], r, w.t_cull);
    if (!root.hit) return;
    if (t.n == 1) {
        leaf(0, root.tn);  // single leaf: host index 0 is the leaf (ni = 0)
        return;
    }
    int stack[kStackCap];
    int sp = 0;
    int c = 0;
    for (;;) {
        const int kids[2] = {host_index(t.right[c], ni), host_index(t.left[c], ni)};
        SlabOut s[2];
        for (int k = 0; k < 2; ++k) s[k] = slab_near(t.bounds[kids[k]], r, w.t_cull);
        for (int k = 0; k < 2; ++k)
            if (s[k].hit && kids[k] >= ni) leaf(kids[k], s[k].tn);
        const bool i0 = s[0].hit && kids[0] < ni, i1 = s[1].hit && kids[1] < ni;
        if (i0 && i1) {
            const bool far0 = s[1].tn < s[0].tn;
            if (sp == kStackCap) {
                w.overflow = true;
                return;
            }
            stack[sp++] = far0 ? kids[0] : kids[1];
            c = far0 ? kids[1] : kids[0];
        } else if (i0 || i1) {
            c = i0 ? kids[0] : kids[1];
        } else {
            if (sp == 0) return;
            c = stack[--sp];
        }
    }
}

}  // namespace

extern "C" {

// Closest hit of each ray by the near-child-first walk. rule 0: plain walk
// (stack overflow -> preorder walk); rule 1: the checked walk above, with a
// soft window of |t_max| * 2^-window_log2. flags[i] (nullable): bit 0 = stack
// overflow, bit 1 = the check fired (both mean the ray was walked again in
// preorder). Returns the number of rays walked again.
int32_t orc_trace_ordered(const void* h, const float* tris, const float* rays, int32_t n, int32_t rule,
                          int32_t window_log2, int32_t* ids, float* t_out, uint8_t* flags) {
    const Tree& t = *reinterpret_cast<const Tree*>(h);
    const Tri* tr = reinterpret_cast<const Tri*>(tris);
    const Ray* rs = reinterpret_cast<const Ray*>(rays);
    int32_t again = 0;
    for (int i = 0; i < n; ++i) {
        const Ray& r = rs[i];
        Walk w;
        w.t_max = r.t_max;
        w.t_cull = r.t_max;
        uint8_t f = 0;
        if (t.n > 0) ordered_walk(t, tr, r, w, rule, ldexpf(1.0f, -window_log2));
        if (w.overflow) f |= 1;
        if (rule == 1 && !w.overflow && w.best >= 0 && w.m <= w.t_max) f |= 2;
        if (f) {
            ++again;
            int id;
            float th;
            bool found = trace_stackless(t, tr, r, id, th);
            ids[i] = found ? id : -1;
            t_out[i] = found ? th : INFINITY;
        } else {
            ids[i] = w.best;
            t_out[i] = w.best >= 0 ? w.t_max : INFINITY;
        }
        if (flags) flags[i] = f;
    }
    return again;
}

}  // extern "C"

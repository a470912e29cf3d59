// wide.hpp -- the 4-wide traversal tree over the right-first binary LBVH, and
// its stackless walk.
//
// Each 128-byte record (one L2 line) collapses a binary subtree: its children
// are up to 4 binary nodes (internal or leaf) in right-first preorder, with
// their exact boxes stored per axis, and for a leaf child its primitive id.
//
// Why the result is bit-identical to the binary right-first walk (and so to
// the host LBVH stack walk): the walk scans a record's children in order with
// the current t_max, tests leaves in place and descends into the first passing
// internal child; when that subtree is finished it resumes the parent at the
// next slot (the `esc` word of the child record). Leaves are therefore visited
// in the binary preorder, with the same boxes, the same t_max and the same
// arithmetic. Skipping the binary nodes between a record and its children
// changes nothing: a child box lies inside its ancestors' boxes and the slab
// test is monotone in the box bounds and in t_max, so whenever a skipped
// ancestor would have failed, every child below it fails too. Hence ids, t, u,
// v and tie winners are those of the binary walk.
#pragma once

#include "lux_common.hpp"

namespace lux {

constexpr uint32_t kWideEnd = 0xffffffffu;
constexpr int kWide = 4;

// 128-byte record, read as 8 float4: lo x/y/z, hi x/y/z of the 4 children
// (component i = child i), the 4 child words, then esc and the slot count.
struct __align__(128) WideNode {
    float lo_x[4], lo_y[4], lo_z[4];
    float hi_x[4], hi_y[4], hi_z[4];
    uint32_t child[4];  // internal child: record index; leaf child: primitive id
    uint32_t esc;       // continue at (record << 2 | slot); kWideEnd = done
    uint32_t meta;      // slot count | leaf mask << 8
    uint32_t pad[2];
};
static_assert(sizeof(WideNode) == 128, "wide record layout");

struct WideStats {
    uint32_t records = 0;     // 128-byte records loaded
    uint32_t leaf_boxes = 0;  // (kept for the counters' layout; leaf boxes live in the records)
    uint32_t tests = 0;       // triangle tests
};

__device__ __forceinline__ float f4c(const float4& v, int i) { return i == 0 ? v.x : i == 1 ? v.y : i == 2 ? v.z : v.w; }

// Closest hit (kAny = false) or first accepted hit (kAny = true) over the wide
// tree. `ignore` skips one primitive's triangle test (a shadow ray's source).
// Returns the primitive id (-1 = none); t_max ends as the hit distance. The
// record stays in registers: the 4 slots are unrolled with constant indices.
template <bool kAny, bool kCount>
__device__ __forceinline__ int wide_walk(const WideNode* __restrict__ wide, uint32_t wide_count,
                                         const lux_bvh_node* __restrict__ nodes, int node_count,
                                         const lux_triangle* __restrict__ tris, int tri_count, const f3& O,
                                         const f3& D, float t_min, float& t_max, int ignore, float& u_hit,
                                         float& v_hit, DevInfo* info, WideStats& st) {
    (void)nodes;
    (void)node_count;
    const SlabRay sr = make_slab_ray(O, D, t_min);
    int best = -1;
    uint32_t cur = wide_count ? 0u : kWideEnd;  // record 0 is the root
    // a depth-first walk enters each record once per child return at most
    uint32_t guard = 5u * wide_count + 2u;
    while (cur != kWideEnd) {
        if (--guard == 0u) {
            set_error(info, kErrWalkOrder);
            break;
        }
        const uint32_t rec = cur >> 2;
        const int slot = int(cur & 3u);
        if (rec >= wide_count) {
            set_error(info, kErrWalkOrder);
            break;
        }
        const float4* rp = reinterpret_cast<const float4*>(wide + rec);
        const float4 lx = rp[0], ly = rp[1], lz = rp[2], hx = rp[3], hy = rp[4], hz = rp[5], ch = rp[6], tail = rp[7];
        if (kCount) ++st.records;
        const uint32_t meta = __float_as_uint(tail.y);
        const int count = int(meta & 0xffu);
        const uint32_t leaf_mask = (meta >> 8) & 0xfu;
        uint32_t next = __float_as_uint(tail.x);
        bool done = false;
#pragma unroll
        for (int i = 0; i < kWide; ++i) {
            if (!done && i >= slot && i < count &&
                slab_hit(sr, t_max, f4c(lx, i), f4c(ly, i), f4c(lz, i), f4c(hx, i), f4c(hy, i), f4c(hz, i))) {
                const uint32_t c = __float_as_uint(f4c(ch, i));
                if (!((leaf_mask >> i) & 1u)) {
                    next = c << 2;  // descend; the child's esc resumes this record at slot i + 1
                    done = true;
                } else if (int(c) != ignore) {
                    if (c >= uint32_t(tri_count)) {
                        set_error(info, kErrPrimRange);
                    } else {
                        const float* p = reinterpret_cast<const float*>(tris + c);
                        const f3 v0 = make3(p[0], p[1], p[2]), v1 = make3(p[3], p[4], p[5]),
                                 v2 = make3(p[6], p[7], p[8]);
                        if (kCount) ++st.tests;
                        float t, u, v;
                        if (tri_hit(O, D, t_min, t_max, v0, v1, v2, t, u, v)) {
                            best = int(c);
                            t_max = t;
                            u_hit = u;
                            v_hit = v;
                            if (kAny) {
                                next = kWideEnd;
                                done = true;
                            }
                        }
                    }
                }
            }
        }
        cur = next;
    }
    return best;
}

}  // namespace lux

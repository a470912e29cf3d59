// lux_common.hpp -- device records, arithmetic and wave64 helpers shared by the
// gfx950 kernels of the LUX hot path.
//
// All arithmetic on the hit path keeps the operation order of the reference
// (src/core/vec3.cuh:40-50, src/core/aabb.cuh:74-109, src/geometry/triangle.h:33-52)
// and every translation unit is compiled with -ffp-contract=off, so box tests,
// Moller-Trumbore distances and Morton codes are bit-identical to the host LBVH.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "lux_hip.h"

namespace lux {

constexpr uint32_t kLeafBit = 0x80000000u;
constexpr uint32_t kIndexMask = 0x7fffffffu;

// Device-side error bits collected in DevInfo::error (host turns them into
// LUX_ERR_DEVICE_CHECK with a message).
enum DeviceError : uint32_t {
    kErrSpinLimit = 1u << 0,      // radix-sort look-back waited too long
    kErrWalkOrder = 1u << 1,      // a stackless walk did not move forward
    kErrClimbLimit = 1u << 2,     // bottom-up pass climbed more than n levels
    kErrPrimRange = 1u << 3,      // a leaf id is outside the triangle array
    kErrNegativeId = 1u << 4,     // an input primitive id is negative
};

// Per-handle device scratch that survives across kernels of one build.
struct DevInfo {
    lux_aabb root;         // union of all input boxes
    int32_t id_min, id_max;
    uint32_t error;        // DeviceError bits
    uint32_t max_depth;    // deepest leaf (root = depth 0)
    uint32_t done_blocks;  // last-block-done counter of the root reduction
    uint32_t pad[3];
};

struct f3 {
    float x, y, z;
};

__device__ __forceinline__ f3 make3(float x, float y, float z) { return f3{x, y, z}; }
__device__ __forceinline__ f3 sub3(const f3& a, const f3& b) { return make3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ float dot3(const f3& a, const f3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ f3 cross3(const f3& a, const f3& b) {
    return make3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ f3 ld3(const lux_vec3& v) { return make3(v.x, v.y, v.z); }

__device__ __forceinline__ lux_aabb box_empty() {
    lux_aabb b;
    b.p_min = lux_vec3{INFINITY, INFINITY, INFINITY};
    b.p_max = lux_vec3{-INFINITY, -INFINITY, -INFINITY};
    return b;
}
// Box union and expansion in the reference's host order: union_aabb and
// AABB::expand take fminf/fmaxf component-wise (src/core/aabb.cuh:47-55,
// :113-115; src/core/vec3.cuh:66-72), and the host libm returns the SECOND
// operand when the two compare equal (+0 against -0) and the other operand when
// one is NaN. Written as compare-and-select so that the sign of a zero bound
// never depends on how the GPU's min/max instruction orders zeros. The
// selection is associative as long as operands keep their left-to-right
// order, so any grouping of an ordered fold gives the host's bits.
__device__ __forceinline__ float host_fminf(float x, float y) { return (x < y || y != y) ? x : y; }
__device__ __forceinline__ float host_fmaxf(float x, float y) { return (x > y || y != y) ? x : y; }

__device__ __forceinline__ lux_aabb box_join(const lux_aabb& a, const lux_aabb& b) {
    lux_aabb r;
    r.p_min = lux_vec3{host_fminf(a.p_min.x, b.p_min.x), host_fminf(a.p_min.y, b.p_min.y),
                       host_fminf(a.p_min.z, b.p_min.z)};
    r.p_max = lux_vec3{host_fmaxf(a.p_max.x, b.p_max.x), host_fmaxf(a.p_max.y, b.p_max.y),
                       host_fmaxf(a.p_max.z, b.p_max.z)};
    return r;
}
__device__ __forceinline__ void box_expand(lux_aabb& b, const lux_vec3& v) {
    b.p_min = lux_vec3{host_fminf(b.p_min.x, v.x), host_fminf(b.p_min.y, v.y), host_fminf(b.p_min.z, v.z)};
    b.p_max = lux_vec3{host_fmaxf(b.p_max.x, v.x), host_fmaxf(b.p_max.y, v.y), host_fmaxf(b.p_max.z, v.z)};
}
__device__ __forceinline__ lux_aabb tri_box(const lux_triangle& t) {
    // triangle_bounds: expand an empty box by v0, v1, v2 in order
    lux_aabb b = box_empty();
    box_expand(b, t.v0);
    box_expand(b, t.v1);
    box_expand(b, t.v2);
    return b;
}

// 32-byte node load as two 16-byte vector loads.
__device__ __forceinline__ void load_node(const lux_bvh_node* nodes, int i, float4& a, float4& b) {
    const float4* p = reinterpret_cast<const float4*>(nodes + i);
    a = p[0];
    b = p[1];
}

// The same for node arrays below 4 GiB (node index < 2^27): base plus a
// zero-extended 32-bit byte offset, one shift per node instead of 64-bit index
// arithmetic.
constexpr uint32_t kOffset32Nodes = 1u << 27;
__device__ __forceinline__ void load_node32(const lux_bvh_node* nodes, uint32_t i, float4& a, float4& b) {
    const float4* p = reinterpret_cast<const float4*>(reinterpret_cast<const char*>(nodes) + (i << 5));
    a = p[0];
    b = p[1];
}

// Per-ray constants of the slab test. 1/d is a per-ray constant, so computing it
// once gives the same value the reference recomputes at every box
// (src/core/aabb.cuh:96). Axes with |d| < 1e-12 take the "origin inside slab"
// branch (:91-94).
struct SlabRay {
    f3 o;
    f3 inv;
    bool par_x, par_y, par_z;
    bool neg_x, neg_y, neg_z;
    float t_min;
};

__device__ __forceinline__ SlabRay make_slab_ray(const f3& o, const f3& d, float t_min) {
    SlabRay s;
    s.o = o;
    s.par_x = fabsf(d.x) < 1e-12f;
    s.par_y = fabsf(d.y) < 1e-12f;
    s.par_z = fabsf(d.z) < 1e-12f;
    s.inv = make3(s.par_x ? 0.0f : 1 / d.x, s.par_y ? 0.0f : 1 / d.y, s.par_z ? 0.0f : 1 / d.z);
    s.neg_x = s.inv.x < 0;
    s.neg_y = s.inv.y < 0;
    s.neg_z = s.inv.z < 0;
    s.t_min = t_min;
    return s;
}

// Slab test against [t_min, t_max]; equal to AABB::intersect's boolean. The
// reference exits at the first empty interval; t_near only grows and t_far only
// shrinks afterwards, so testing once after all three axes gives the same answer.
__device__ __forceinline__ bool slab_hit(const SlabRay& r, float t_max, float lox, float loy,
                                         float loz, float hix, float hiy, float hiz) {
    bool miss = false;
    float t_near = r.t_min, t_far = t_max;
    {
        float a = ((r.neg_x ? hix : lox) - r.o.x) * r.inv.x;
        float c = ((r.neg_x ? lox : hix) - r.o.x) * r.inv.x;
        if (r.par_x) {
            miss |= (r.o.x < lox) | (r.o.x > hix);
        } else {
            t_near = fmaxf(t_near, a);
            t_far = fminf(t_far, c);
        }
    }
    {
        float a = ((r.neg_y ? hiy : loy) - r.o.y) * r.inv.y;
        float c = ((r.neg_y ? loy : hiy) - r.o.y) * r.inv.y;
        if (r.par_y) {
            miss |= (r.o.y < loy) | (r.o.y > hiy);
        } else {
            t_near = fmaxf(t_near, a);
            t_far = fminf(t_far, c);
        }
    }
    {
        float a = ((r.neg_z ? hiz : loz) - r.o.z) * r.inv.z;
        float c = ((r.neg_z ? loz : hiz) - r.o.z) * r.inv.z;
        if (r.par_z) {
            miss |= (r.o.z < loz) | (r.o.z > hiz);
        } else {
            t_near = fmaxf(t_near, a);
            t_far = fminf(t_far, c);
        }
    }
    return !miss && !(t_near > t_far);
}

// Slab-test modes, a compile-time choice made once per wave (slab_mode):
// kSlabParallel  some lane has an axis-parallel direction: slab_hit's branches;
// kSlabLane      no parallel axis, direction signs read per lane;
// 0..7           no parallel axis and every lane of the wave in octant m
//                (bit 0: d.x < 0, bit 1: d.y < 0, bit 2: d.z < 0), so the
//                near/far planes are fixed and the per-lane selects vanish.
// All modes perform the same float operations on the same operands.
constexpr int kSlabParallel = -2;
constexpr int kSlabLane = -1;

__device__ __forceinline__ int ray_octant(const SlabRay& r) {
    return int(r.neg_x) | int(r.neg_y) << 1 | int(r.neg_z) << 2;
}

// slab_hit for a ray with no axis parallel to the box faces (|d| >= 1e-12 on
// every axis): the same operations and result without the per-axis branches.
template <int kMode = kSlabLane>
__device__ __forceinline__ bool slab_hit_np(const SlabRay& r, float t_max, float lox, float loy, float loz, float hix,
                                            float hiy, float hiz) {
    const bool nx = kMode < 0 ? r.neg_x : (kMode & 1) != 0;
    const bool ny = kMode < 0 ? r.neg_y : (kMode & 2) != 0;
    const bool nz = kMode < 0 ? r.neg_z : (kMode & 4) != 0;
    const float ax = ((nx ? hix : lox) - r.o.x) * r.inv.x;
    const float cx = ((nx ? lox : hix) - r.o.x) * r.inv.x;
    const float ay = ((ny ? hiy : loy) - r.o.y) * r.inv.y;
    const float cy = ((ny ? loy : hiy) - r.o.y) * r.inv.y;
    const float az = ((nz ? hiz : loz) - r.o.z) * r.inv.z;
    const float cz = ((nz ? loz : hiz) - r.o.z) * r.inv.z;
    const float t_near = fmaxf(fmaxf(fmaxf(r.t_min, ax), ay), az);
    const float t_far = fminf(fminf(fminf(t_max, cx), cy), cz);
    return !(t_near > t_far);
}

// True when any active lane of the wave has an axis-parallel direction; walks
// take the branch-free slab test when none has (a wave-uniform choice).
__device__ __forceinline__ bool wave_any_parallel(const SlabRay& r) {
    return __ballot(r.par_x || r.par_y || r.par_z) != 0;
}

template <int kMode>
__device__ __forceinline__ bool slab_pass(const SlabRay& r, float t_max, float lox, float loy, float loz, float hix,
                                          float hiy, float hiz) {
    if constexpr (kMode == kSlabParallel)
        return slab_hit(r, t_max, lox, loy, loz, hix, hiy, hiz);
    else
        return slab_hit_np<kMode>(r, t_max, lox, loy, loz, hix, hiy, hiz);
}

template <int kMode>
using SlabMode = std::integral_constant<int, kMode>;

// Calls f(SlabMode<m>{}) with the wave's slab-test mode (wave-uniform: every
// active lane takes the same call). Rays traced in octant order fill most waves
// from one octant.
template <typename F>
__device__ __forceinline__ auto with_slab_mode(const SlabRay& r, F&& f) {
    if (wave_any_parallel(r)) return f(SlabMode<kSlabParallel>{});
    const int oct = ray_octant(r);
    const int oct0 = __builtin_amdgcn_readfirstlane(oct);
    if (__ballot(oct != oct0) != 0) return f(SlabMode<kSlabLane>{});
    switch (oct0) {
        case 0: return f(SlabMode<0>{});
        case 1: return f(SlabMode<1>{});
        case 2: return f(SlabMode<2>{});
        case 3: return f(SlabMode<3>{});
        case 4: return f(SlabMode<4>{});
        case 5: return f(SlabMode<5>{});
        case 6: return f(SlabMode<6>{});
        default: return f(SlabMode<7>{});
    }
}

// Moller-Trumbore, operation order of src/geometry/triangle.h:33-52, without
// the final range test: true when u and v pass, with t computed.
__device__ __forceinline__ bool tri_uvt(const f3& o, const f3& d, const f3& v0, const f3& v1, const f3& v2, float& t,
                                        float& u, float& v) {
    f3 e1 = sub3(v1, v0);
    f3 e2 = sub3(v2, v0);
    f3 h = cross3(d, e2);
    float a = dot3(e1, h);
    if (fabsf(a) < 1e-7f) return false;
    float f = 1.0f / a;
    f3 s = sub3(o, v0);
    u = f * dot3(s, h);
    if (u < 0.0f || u > 1.0f) return false;
    f3 q = cross3(s, e1);
    v = f * dot3(d, q);
    if (v < 0.0f || u + v > 1.0f) return false;
    t = f * dot3(e2, q);
    return true;
}

// intersect_triangle: a hit in the open range (t_min, t_max).
__device__ __forceinline__ bool tri_hit(const f3& o, const f3& d, float t_min, float t_max,
                                        const f3& v0, const f3& v1, const f3& v2, float& t,
                                        float& u, float& v) {
    return tri_uvt(o, d, v0, v1, v2, t, u, v) && t > t_min && t < t_max;
}

// Inclusive AABB overlap (src/core/aabb.cuh:68-72).
__device__ __forceinline__ bool box_overlap(float alox, float aloy, float aloz, float ahix,
                                            float ahiy, float ahiz, const lux_aabb& b) {
    return !(b.p_min.x > ahix || b.p_max.x < alox || b.p_min.y > ahiy || b.p_max.y < aloy ||
             b.p_min.z > ahiz || b.p_max.z < aloz);
}

// ---- wave64 helpers --------------------------------------------------------
__device__ __forceinline__ uint64_t ballot64(bool p) { return __ballot(p); }

// gfx950 hands workgroups to its 8 XCDs round-robin (workgroup b runs on XCD
// b % 8), each XCD with its own L2. xcd_block() renumbers the workgroups so
// that XCD x owns one contiguous run of per + (x < rem) virtual workgroups:
// neighbouring work items (coherent rays, queue entries in path order) then
// share one L2 instead of being spread over all eight. A bijection on
// [0, gridDim.x) for every grid size.
constexpr uint32_t kXcds = 8;
__device__ __forceinline__ uint32_t xcd_block() {
    const uint32_t g = gridDim.x, b = blockIdx.x;
    const uint32_t per = g / kXcds, rem = g % kXcds;
    const uint32_t x = b % kXcds, k = b / kXcds;
    return x * per + (x < rem ? x : rem) + k;
}
__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }
__device__ __forceinline__ uint64_t lanemask_lt() {
    uint32_t l = lane_id();
    return l == 0 ? 0ull : (~0ull >> (64 - l));
}
// Appends one entry per pushing lane of the wave with one atomic add: returns
// this lane's slot (meaningful only where push is true).
__device__ __forceinline__ uint32_t wave_append(uint32_t* count, bool push) {
    uint64_t m = ballot64(push);
    uint32_t base = 0;
    if (m) {
        uint32_t leader = uint32_t(__ffsll((unsigned long long)m) - 1);
        if (lane_id() == leader)
            base = __hip_atomic_fetch_add(count, uint32_t(__popcll(m)), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        base = __shfl(base, int(leader), 64);
    }
    return base + uint32_t(__popcll(m & lanemask_lt()));
}
__device__ __forceinline__ void set_error(DevInfo* info, uint32_t bit) {
    __hip_atomic_fetch_or(&info->error, bit, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

}  // namespace lux

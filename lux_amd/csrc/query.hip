// query.hip -- closest-hit, any-hit and AABB-overlap queries over the
// right-first preorder stackless LBVH. Closest and any hit take the
// near-child-first walk of ordered.hpp, which returns what the preorder walk
// below returns; rays whose ordered stack overflows take the preorder walk.
//
// Walk rule (src/bvh/bvh_device.cuh:59-87): test node i; on a miss or after a
// leaf jump to nodes[i].escape, on an internal hit step to i+1. With the nodes in
// right-first preorder this visits nodes in the pop order of the host stack
// traversal (src/bvh/cpu_bvh_traversal.h:37-78), so the shrinking t_max and the
// winner of exact-t ties are those of the host LBVH.
//
// Each walk also checks that it moves forward (escape and i+1 are always greater
// than i in preorder), so a damaged node array ends the walk instead of looping.
#include <algorithm>

#include "lux_runtime.hpp"
#include "ordered.hpp"

namespace lux {
namespace {

// 64 threads (one wave) per block for the ray queries: with the gathered
// near-child-first walks the microbench step measured 822 Mrays/s at 64
// against 802 at 128 and 791 at 256 (a workgroup's resources are freed only
// when its last wave ends, so a finished wave's slot waits for its siblings;
// the LDS stack is 4 KiB per block at 64).
constexpr int kBlock = 64;
// The plain copy: 128-thread blocks, at most 4096 of them (4.8 TB/s on 1 GiB;
// other shapes from 2048 to 16384 blocks of 256 or 512 threads measured
// 4.3-5.0 TB/s).
constexpr int kCopyBlock = 128;

struct RayIn {
    f3 o, d;
    float t_min, t_max;
};

__device__ __forceinline__ RayIn load_ray(const lux_ray* rays, int i) {
    const float* p = reinterpret_cast<const float*>(rays + i);
    RayIn r;
    r.o = make3(p[0], p[1], p[2]);
    r.d = make3(p[3], p[4], p[5]);
    r.t_min = p[6];
    r.t_max = p[7];
    return r;
}

__device__ __forceinline__ void load_tri(const lux_triangle* tris, int i, f3& v0, f3& v1, f3& v2) {
    const float* p = reinterpret_cast<const float*>(tris + i);
    v0 = make3(p[0], p[1], p[2]);
    v1 = make3(p[3], p[4], p[5]);
    v2 = make3(p[6], p[7], p[8]);
}

// Right-first walk over the sibling-pair copy of the node array (`nodes` =
// BvhDev::pairs, node_count = its slot count): the same node visits in the
// same order as the preorder walk, with a node's siblings on one cache line.
// kAny = false: closest hit; kAny = true: stop at the first accepted triangle.
// visits/tests count node loads and triangle tests (only read by the stats kernel).
template <bool kAny, int kMode, bool kOffset32>
__device__ __forceinline__ int walk_impl(const lux_bvh_node* __restrict__ nodes, int node_count,
                                         const lux_triangle* __restrict__ tris, int tri_count, const RayIn& ray,
                                         const SlabRay& sr, float& t_best, DevInfo* info, uint32_t& visits,
                                         uint32_t& tests) {
    // canonical from the start, so the loop's min/max need no per-step canonicalise
    float t_max = __builtin_canonicalizef(ray.t_max);
    int best = -1;
    uint32_t k = 0;
    const uint32_t nc = uint32_t(node_count);
    int guard = node_count + 1;  // a walk visits each node at most once
    bool bad_prim = false;
    // The rare checks (leaf id range, step guard) are folded into lane flags and
    // reported after the loop, keeping the loop body to one branch (the leaf's
    // triangle test).
    while (k < nc && --guard > 0) {
        float4 a, b;
        if (kOffset32)
            load_node32(nodes, k, a, b);
        else
            load_node(nodes, int(k), a, b);
        ++visits;
        const int prim = __float_as_int(b.z);
        const bool pass = slab_pass<kMode>(sr, t_max, a.x, a.y, a.z, a.w, b.x, b.y);
        // internal hit: the right child's slot; otherwise the escape
        uint32_t next = (pass && prim < 0) ? uint32_t(~prim) : __float_as_uint(b.w);
        bad_prim |= pass && prim >= tri_count;
        if (pass && prim >= 0 && prim < tri_count) {
            f3 v0, v1, v2;
            load_tri(tris, prim, v0, v1, v2);
            ++tests;
            float t, u, v;
            if (tri_hit(ray.o, ray.d, ray.t_min, t_max, v0, v1, v2, t, u, v)) {
                best = prim;
                t_max = t;
                if (kAny) next = 0xffffffffu;  // done
            }
        }
        k = next;
    }
    if (bad_prim) set_error(info, kErrPrimRange);
    if (guard <= 0 && k < nc) set_error(info, kErrWalkOrder);
    t_best = t_max;
    return best;
}

// Any hit by the preorder walk with the triangle tests gathered across the wave.
// A lane whose walk reaches a leaf box that passes holds that leaf and stops;
// the wave runs the triangle tests of all held leaves together once at least
// 1/kDeferDiv of its still-working lanes hold one (or none can step on). Each
// lane performs the same operations in the same order as walk_impl<true>, so
// the answer is the same; only the interleaving across lanes changes. Without
// it, most steps of the wave carry the ~60-instruction triangle test for the
// one or two lanes that need it. On 2^22 random rays over 2^20 triangles
// (octant-specialised slab test): kDeferDiv 2 / 4 / 8: 4.78 / 4.67 / 4.70 ms;
// testing each leaf at once (the plain walk), 4.93 ms before the octant
// specialisation, against 4.87 ms with kDeferDiv 4.
constexpr int kDeferDiv = 4;

template <int kMode>
__device__ __forceinline__ int walk_any_deferred(const lux_bvh_node* __restrict__ nodes, int node_count,
                                                 const lux_triangle* __restrict__ tris, int tri_count,
                                                 const RayIn& ray, const SlabRay& sr, float& t_best,
                                                 DevInfo* info, uint32_t& visits, uint32_t& tests) {
    const float t_max = __builtin_canonicalizef(ray.t_max);
    t_best = t_max;
    const uint32_t nc = uint32_t(node_count);
    uint32_t k = 0;
    int held = -1, found = -1;
    int guard = node_count;  // a walk visits each node at most once
    bool bad_prim = false;
    for (;;) {
        // (& rather than &&: one mask, no nested branches)
        if ((held < 0) & (k < nc) & (guard > 0)) {
            --guard;
            float4 a, b;
            load_node32(nodes, k, a, b);
            ++visits;
            const int prim = __float_as_int(b.z);
            const bool pass = slab_pass<kMode>(sr, t_max, a.x, a.y, a.z, a.w, b.x, b.y);
            held = (pass & (prim >= 0)) ? prim : -1;
            k = (pass & (prim < 0)) ? uint32_t(~prim) : __float_as_uint(b.w);
        }
        const uint64_t want = __ballot(held >= 0);
        const uint64_t more = __ballot((held < 0) & (k < nc) & (guard > 0));
        if (!more) {
            if (!want) break;
        } else if (__popcll(want) * kDeferDiv < __popcll(want | more)) {
            continue;
        }
        if (held >= 0) {
            if (held < tri_count) {
                f3 v0, v1, v2;
                load_tri(tris, held, v0, v1, v2);
                ++tests;
                float t, u, v;
                if (tri_hit(ray.o, ray.d, ray.t_min, t_max, v0, v1, v2, t, u, v)) {
                    found = held;
                    t_best = t;
                    k = 0xffffffffu;  // done
                }
            } else {
                bad_prim = true;  // skipped, as the preorder walk skips it
            }
            held = -1;
        }
    }
    if (bad_prim) set_error(info, kErrPrimRange);
    if (guard <= 0 && k < nc) set_error(info, kErrWalkOrder);
    return found;
}

// Any hit by the near-child-first walk with gathered triangle tests. With
// t_max fixed, a box passes or not whatever the visit order, so the walk meets
// the same boxes and triangles as the preorder walk until one triangle is hit;
// near child first meets one sooner. A lane whose stack overflows takes the
// preorder walk.
template <int kMode>
__device__ __forceinline__ int walk_any_ordered(const lux_bvh_node* __restrict__ nodes, int node_count,
                                                const lux_triangle* __restrict__ tris, int tri_count,
                                                const RayIn& ray, const SlabRay& sr, float& t_best, uint32_t* stack,
                                                DevInfo* info, uint32_t& visits, uint32_t& tests) {
    const float t_max = __builtin_canonicalizef(ray.t_max);
    OrderedBest ob(t_max);  // t_cull stays t_max: no hit is taken before the walk ends
    int found = -1;
    float t_found = t_max;
    bool bad = false, bad_prim = false;
    auto leaf = [&](uint32_t, int prim, float) {
        if (prim >= tri_count) {
            bad_prim = true;
            return false;
        }
        f3 v0, v1, v2;
        load_tri(tris, prim, v0, v1, v2);
        ++tests;
        float t, u, v;
        if (tri_hit(ray.o, ray.d, ray.t_min, t_max, v0, v1, v2, t, u, v)) {
            found = prim;
            t_found = t;
            return true;
        }
        return false;
    };
    const bool done = ordered_walk_gathered(SlabMode<kMode>{}, nodes, uint32_t(node_count), sr, ob, stack, kBlock,
                                            visits, bad, leaf);
    if (bad_prim) set_error(info, kErrPrimRange);
    if (bad) set_error(info, kErrWalkOrder);
    if (!done) return walk_any_deferred<kMode>(nodes, node_count, tris, tri_count, ray, sr, t_best, info, visits, tests);
    t_best = t_found;
    return found;
}

template <bool kAny>
__device__ __forceinline__ int walk(const lux_bvh_node* __restrict__ nodes, int node_count,
                                    const lux_triangle* __restrict__ tris, int tri_count, const RayIn& ray,
                                    float& t_best, DevInfo* info, uint32_t& visits, uint32_t& tests) {
    const SlabRay sr = make_slab_ray(ray.o, ray.d, ray.t_min);
    if (uint32_t(node_count) > kOffset32Nodes)
        return walk_impl<kAny, kSlabParallel, false>(nodes, node_count, tris, tri_count, ray, sr, t_best, info, visits,
                                                     tests);
    return with_slab_mode(sr, [&](auto mode) {
        constexpr int m = decltype(mode)::value;
        if constexpr (kAny)
            return walk_any_deferred<m>(nodes, node_count, tris, tri_count, ray, sr, t_best, info, visits, tests);
        else
            return walk_impl<false, m, true>(nodes, node_count, tris, tri_count, ray, sr, t_best, info, visits, tests);
    });
}

// Closest hit by the near-child-first walk (ordered.hpp), equal to walk<false>.
// stack: this lane's LDS stack (entries kBlock words apart). A ray whose stack
// overflows, or whose walk the check of ordered.hpp flags, is walked again in
// right-first preorder.
template <int kMode>
__device__ __forceinline__ int walk_ordered(const lux_bvh_node* __restrict__ nodes, int node_count,
                                            const uint32_t* __restrict__ slot_pre,
                                            const lux_triangle* __restrict__ tris, int tri_count, const RayIn& ray,
                                            const SlabRay& sr, float& t_best, uint32_t* stack, DevInfo* info,
                                            uint32_t& visits, uint32_t& tests) {
    OrderedBest ob(__builtin_canonicalizef(ray.t_max));
    int best = -1;
    uint32_t best_slot = 0;
    bool bad = false, bad_prim = false;
    auto leaf = [&](uint32_t slot, int prim, float tn) {
        if (prim >= tri_count) {
            bad_prim = true;
            return false;
        }
        f3 v0, v1, v2;
        load_tri(tris, prim, v0, v1, v2);
        ++tests;
        float t, u, v;
        if (tri_uvt(ray.o, ray.d, v0, v1, v2, t, u, v) && t > ray.t_min) {
            ob.note(t, tn);
            if (t < ob.t_max || (best >= 0 && t == ob.t_max && slot_pre[slot] < slot_pre[best_slot])) {
                ob.take(t);
                best = prim;
                best_slot = slot;
            }
        }
        return false;
    };
    const bool done = ordered_walk_gathered(SlabMode<kMode>{}, nodes, uint32_t(node_count), sr, ob, stack, kBlock,
                                            visits, bad, leaf);
    if (bad_prim) set_error(info, kErrPrimRange);
    if (bad) set_error(info, kErrWalkOrder);
    if (!done || (best >= 0 && ob.ambiguous()))
        return walk_impl<false, kMode, true>(nodes, node_count, tris, tri_count, ray, sr, t_best, info, visits,
                                             tests);
    t_best = ob.t_max;
    return best;
}

// One ray: closest hit by the near-child-first walk when `stack` is given,
// otherwise (and for any hit) the preorder walk. Counters: nodes loaded and
// triangle tests.
template <bool kAny>
__device__ __forceinline__ int trace_one(const lux_bvh_node* __restrict__ nodes, int node_count,
                                         const lux_triangle* __restrict__ tris, int tri_count, const RayIn& r,
                                         float& t, DevInfo* info, uint32_t& visits, uint32_t& tests,
                                         const uint32_t* __restrict__ slot_pre = nullptr, uint32_t* stack = nullptr) {
    if (!kAny && stack && uint32_t(node_count) <= kOffset32Nodes && node_count > 1) {
        const SlabRay sr = make_slab_ray(r.o, r.d, r.t_min);
        return with_slab_mode(sr, [&](auto mode) {
            return walk_ordered<decltype(mode)::value>(nodes, node_count, slot_pre, tris, tri_count, r, sr, t, stack,
                                                       info, visits, tests);
        });
    }
    return walk<kAny>(nodes, node_count, tris, tri_count, r, t, info, visits, tests);
}

// One ray's any hit: the near-child-first walk on this lane's LDS stack when
// the array allows it, otherwise the preorder walk.
__device__ __forceinline__ int trace_any_one(const lux_bvh_node* __restrict__ nodes, int node_count,
                                             const lux_triangle* __restrict__ tris, int tri_count, const RayIn& r,
                                             float& t, DevInfo* info, uint32_t& visits, uint32_t& tests,
                                             uint32_t* stack) {
    if (stack && uint32_t(node_count) <= kOffset32Nodes && node_count > 1) {
        const SlabRay sr = make_slab_ray(r.o, r.d, r.t_min);
        return with_slab_mode(sr, [&](auto mode) {
            return walk_any_ordered<decltype(mode)::value>(nodes, node_count, tris, tri_count, r, sr, t, stack, info,
                                                           visits, tests);
        });
    }
    return trace_one<true>(nodes, node_count, tris, tri_count, r, t, info, visits, tests);
}

// Closest hit, one ray per lane in `order` (a permutation of [0, n), or input
// order); the hit lands at the ray's own index.
template <bool kOrdered>
__global__ __launch_bounds__(kBlock) void k_trace_closest(const lux_bvh_node* __restrict__ nodes, int node_count,
                                                          const uint32_t* __restrict__ slot_pre,
                                                          const lux_triangle* __restrict__ tris, int tri_count,
                                                          const lux_ray* __restrict__ rays, int n,
                                                          const uint32_t* __restrict__ order,
                                                          lux_hit* __restrict__ hits, DevInfo* info) {
    __shared__ uint32_t stack[kOrdered ? kOrderedStack * kBlock : 1];
    const int i = int(xcd_block()) * kBlock + threadIdx.x;
    if (i >= n) return;
    const int ray = order ? int(order[i]) : i;  // order is a permutation of [0, n)
    if (uint32_t(ray) >= uint32_t(n)) {          // only if the sort itself reported an error
        set_error(info, kErrSpinLimit);
        return;
    }
    RayIn r = load_ray(rays, ray);
    float t;
    uint32_t visits = 0, tests = 0;
    const int id = trace_one<false>(nodes, node_count, tris, tri_count, r, t, info, visits, tests, slot_pre,
                                    kOrdered ? stack + threadIdx.x : nullptr);
    lux_hit h;
    h.primitive_id = id;
    h.t = id >= 0 ? t : INFINITY;
    hits[ray] = h;
}

__global__ __launch_bounds__(kBlock) void k_trace_any(const lux_bvh_node* __restrict__ nodes, int node_count,
                                                      const lux_triangle* __restrict__ tris, int tri_count,
                                                      const lux_ray* __restrict__ rays, int n,
                                                      const uint32_t* __restrict__ order,
                                                      uint8_t* __restrict__ occluded, DevInfo* info) {
    __shared__ uint32_t stack[kOrderedStack * kBlock];
    const int i = int(xcd_block()) * kBlock + threadIdx.x;
    if (i >= n) return;
    const int ray = order ? int(order[i]) : i;  // order is a permutation of [0, n)
    if (uint32_t(ray) >= uint32_t(n)) {          // only if the sort itself reported an error
        set_error(info, kErrSpinLimit);
        return;
    }
    RayIn r = load_ray(rays, ray);
    float t;
    uint32_t visits = 0, tests = 0;
    const int id = trace_any_one(nodes, node_count, tris, tri_count, r, t, info, visits, tests, stack + threadIdx.x);
    occluded[ray] = id >= 0 ? 1 : 0;
}

// The walks of the handle, counting per ray the 32-byte nodes loaded and the
// triangle tests (summed per wave, one atomic per wave): stats[0..2]
// closest-hit {nodes, 0, tests}, stats[3..5] any-hit {nodes, 0, tests}. Rays
// the near-child-first walk hands to the preorder walk count both walks. Used
// to state the bytes a traversal actually touches.
template <bool kOrdered>
__global__ __launch_bounds__(kBlock) void k_trace_stats(const lux_bvh_node* __restrict__ nodes, int node_count,
                                                        const uint32_t* __restrict__ slot_pre,
                                                        const lux_triangle* __restrict__ tris, int tri_count,
                                                        const lux_ray* __restrict__ rays, int n,
                                                        unsigned long long* stats, DevInfo* info) {
    __shared__ uint32_t stack[kOrdered ? kOrderedStack * kBlock : 1];
    int i = blockIdx.x * kBlock + threadIdx.x;
    uint32_t v0 = 0, t0 = 0, v1 = 0, t1 = 0;
    if (i < n) {
        RayIn r = load_ray(rays, i);
        float t;
        trace_one<false>(nodes, node_count, tris, tri_count, r, t, info, v0, t0, slot_pre,
                         kOrdered ? stack + threadIdx.x : nullptr);
        trace_any_one(nodes, node_count, tris, tri_count, r, t, info, v1, t1,
                      kOrdered ? stack + threadIdx.x : nullptr);
    }
    unsigned long long c[6] = {v0, 0, t0, v1, 0, t1};
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        unsigned long long x = c[k];
        for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
        if (lane_id() == 0) atomicAdd(&stats[k], x);
    }
}

__global__ void k_zero_u64(unsigned long long* p, int count) {
    for (int i = threadIdx.x; i < count; i += blockDim.x) p[i] = 0;
}

__global__ void k_zero_counter(int32_t* counter) { *counter = 0; }

__global__ __launch_bounds__(kCopyBlock) void k_copy16(const int4* __restrict__ src, int4* __restrict__ dst,
                                                       size_t count) {
    for (size_t i = size_t(blockIdx.x) * kCopyBlock + threadIdx.x; i < count; i += size_t(gridDim.x) * kCopyBlock)
        dst[i] = src[i];
}

// Overlap, level by level. A workgroup takes kOvQueries query boxes (in the
// query tree's Morton order) and keeps in LDS the (query, sibling pair) items
// of one tree level: each step tests both children of every item with one
// 64-byte pair load (two items per lane in flight), collects the overlapping
// leaves and lists the entered internal children for the next level. A query's
// dependent chain is thus the tree depth, and the loads of a level are
// independent. An item whose child finds the next list full walks that child's
// subtree at once in preorder. Pairs collect in LDS and go out with one atomic
// per workgroup at the end (directly, one wave-wide atomic per step, once the
// LDS buffer is full). Every leaf whose box overlaps the query box is reached
// (an ancestor's box contains its descendants'), so the pair set is that of
// the per-query walk of src/bvh/bvh_overlap.cu:25-251, and the counter keeps
// counting past max_pairs so the caller gets the exact total (:209-214).
// LDS per workgroup: the two item lists (kOvCap entries of 5 bytes each) and
// the pair buffer (8 bytes an entry) set how many workgroups share a CU. On the
// 1M-triangle self query: lists 2048 / buffer 2048 (40 KiB) 334 us, 2048 / 512
// 290 us, 1024 / 1024 236 us, 1024 / 512 (17 KiB) 222 us, and smaller sizes
// the same 221-222 us.
constexpr int kOvThreads = 256;
constexpr int kOvQueries = 128;
constexpr int kOvCap = 1024;
constexpr int kOvPairBuf = 512;

struct OvShared {
    lux_aabb qb[kOvQueries];
    int32_t qid[kOvQueries];
    uint32_t item[2][kOvCap];  // pair slot (children at slot, slot + 1)
    uint8_t iq[2][kOvCap];     // query within the workgroup
    lux_pair out[kOvPairBuf];
    uint32_t count[2];
    uint32_t nout;
    int32_t base;
};

__device__ __forceinline__ void emit_global(bool emit, int a, int b, lux_pair* __restrict__ pairs, int max_pairs,
                                            int32_t* counter) {
    const uint64_t m = ballot64(emit);
    if (!m) return;
    const uint32_t leader = uint32_t(__ffsll((unsigned long long)m) - 1);
    int base = 0;
    if (lane_id() == leader)
        base = __hip_atomic_fetch_add(counter, int(__popcll(m)), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    base = __shfl(base, int(leader), 64);
    if (emit) {
        const int idx = base + int(__popcll(m & lanemask_lt()));
        if (idx < max_pairs) {
            lux_pair pr;
            pr.a = a;
            pr.b = b;
            pairs[idx] = pr;
        }
    }
}

// One pair found by this lane (emit) goes to the workgroup's LDS buffer, or
// out directly once the buffer is full.
__device__ __forceinline__ void emit_pair(OvShared& sh, bool emit, int a, int b, lux_pair* __restrict__ pairs,
                                          int max_pairs, int32_t* counter) {
    bool spill = false;
    if (emit) {
        const uint32_t at = __hip_atomic_fetch_add(&sh.nout, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        if (at < uint32_t(kOvPairBuf)) {
            lux_pair pr;
            pr.a = a;
            pr.b = b;
            sh.out[at] = pr;
        } else {
            spill = true;
        }
    }
    if (ballot64(spill)) emit_global(spill, a, b, pairs, max_pairs, counter);
}

__device__ __forceinline__ bool node_overlap(const float4& a, const float4& b, const lux_aabb& q) {
    return box_overlap(a.x, a.y, a.z, a.w, b.x, b.y, q);
}

// Preorder walk of the subtree at `slot` (an internal node already known to
// overlap), ending at the node's escape.
__device__ void overlap_subtree(OvShared& sh, const lux_bvh_node* __restrict__ nodes, uint32_t nc, uint32_t slot,
                                uint32_t stop, const lux_aabb& qb, int qid, int self_mode,
                                lux_pair* __restrict__ pairs, int max_pairs, int32_t* counter, DevInfo* info) {
    uint32_t k = slot;
    for (uint32_t guard = 0; guard <= nc; ++guard) {
        float4 a, b;
        load_node32(nodes, k, a, b);
        const int prim = __float_as_int(b.z);
        const bool hit = node_overlap(a, b, qb);
        emit_pair(sh, hit && prim >= 0 && (!self_mode || prim > qid), self_mode ? qid : prim,
                  self_mode ? prim : qid, pairs, max_pairs, counter);
        k = (hit && prim < 0) ? uint32_t(~prim) : __float_as_uint(b.w);
        if (k == stop || k >= nc) return;
    }
    set_error(info, kErrWalkOrder);
}

// Both children of one item: leaves to the pair buffer, entered internal
// children to the next level's list.
__device__ __forceinline__ void overlap_item(OvShared& sh, int nxt, bool valid, uint32_t c, int qi,
                                             const float4& a0, const float4& b0, const float4& a1, const float4& b1,
                                             const lux_bvh_node* __restrict__ nodes, uint32_t nc, int self_mode,
                                             lux_pair* __restrict__ pairs, int max_pairs, int32_t* counter,
                                             DevInfo* info) {
    const lux_aabb qb = sh.qb[qi];
    const int qid = sh.qid[qi];
    const int p0 = __float_as_int(b0.z), p1 = __float_as_int(b1.z);
    const bool h0 = valid && node_overlap(a0, b0, qb), h1 = valid && node_overlap(a1, b1, qb);
    emit_pair(sh, h0 && p0 >= 0 && (!self_mode || p0 > qid), self_mode ? qid : p0, self_mode ? p0 : qid, pairs,
              max_pairs, counter);
    emit_pair(sh, h1 && p1 >= 0 && (!self_mode || p1 > qid), self_mode ? qid : p1, self_mode ? p1 : qid, pairs,
              max_pairs, counter);
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        const int p = j ? p1 : p0;
        if (!(j ? h1 : h0) || p >= 0) continue;
        const uint32_t child = uint32_t(~p);
        if (child + 1 >= nc) {
            set_error(info, kErrWalkOrder);
            continue;
        }
        const uint32_t at = __hip_atomic_fetch_add(&sh.count[nxt], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        if (at < uint32_t(kOvCap)) {
            sh.item[nxt][at] = child;
            sh.iq[nxt][at] = uint8_t(qi);
        } else {
            overlap_subtree(sh, nodes, nc, c + uint32_t(j), __float_as_uint(j ? b1.w : b0.w), qb, qid, self_mode,
                            pairs, max_pairs, counter, info);
        }
    }
}

__global__ __launch_bounds__(kOvThreads) void k_overlap(const lux_bvh_node* __restrict__ nodes, int node_count,
                                                        const lux_aabb* __restrict__ qboxes,
                                                        const int32_t* __restrict__ qids,
                                                        const uint32_t* __restrict__ qorder, int nq, int self_mode,
                                                        lux_pair* __restrict__ pairs, int max_pairs, int32_t* counter,
                                                        DevInfo* info) {
    __shared__ OvShared sh;
    const uint32_t nc = uint32_t(node_count);
    const int q0 = blockIdx.x * kOvQueries;
    const int nql = min(kOvQueries, nq - q0);
    const int tid = threadIdx.x;
    if (tid < 2) sh.count[tid] = 0;
    if (tid == 0) sh.nout = 0;
    __syncthreads();

    // the root: slot 0, alone
    {
        float4 a, b;
        load_node32(nodes, 0, a, b);
        const int prim = __float_as_int(b.z);
        bool emit = false;
        int qid = -1;
        if (tid < nql) {
            const uint32_t src = qorder[q0 + tid];
            const lux_aabb qb = qboxes[src];
            qid = qids[src];
            sh.qb[tid] = qb;
            sh.qid[tid] = qid;
            if (node_overlap(a, b, qb)) {
                if (prim >= 0) {
                    emit = !self_mode || prim > qid;
                } else if (uint32_t(~prim) + 1 < nc) {
                    const uint32_t at = __hip_atomic_fetch_add(&sh.count[0], 1u, __ATOMIC_RELAXED,
                                                               __HIP_MEMORY_SCOPE_WORKGROUP);
                    sh.item[0][at] = uint32_t(~prim);  // at < kOvQueries <= kOvCap
                    sh.iq[0][at] = uint8_t(tid);
                } else {
                    set_error(info, kErrWalkOrder);
                }
            }
        }
        emit_pair(sh, emit, self_mode ? qid : prim, self_mode ? prim : qid, pairs, max_pairs, counter);
    }
    __syncthreads();

    for (int cur = 0;; cur ^= 1) {
        const uint32_t count = min(sh.count[cur], uint32_t(kOvCap));
        if (count == 0) break;
        __syncthreads();  // every thread has read the count
        const int nxt = cur ^ 1;
        if (tid == 0) sh.count[nxt] = 0;
        __syncthreads();
        for (uint32_t it = uint32_t(tid); it < count; it += 2 * kOvThreads) {
            const uint32_t it2 = it + kOvThreads;
            const bool two = it2 < count;
            const uint32_t c0 = sh.item[cur][it], c1 = two ? sh.item[cur][it2] : c0;
            const int q0i = sh.iq[cur][it], q1i = two ? sh.iq[cur][it2] : q0i;
            float4 a00, b00, a01, b01, a10, b10, a11, b11;
            load_node32(nodes, c0, a00, b00);
            load_node32(nodes, c0 + 1, a01, b01);
            load_node32(nodes, c1, a10, b10);
            load_node32(nodes, c1 + 1, a11, b11);
            overlap_item(sh, nxt, true, c0, q0i, a00, b00, a01, b01, nodes, nc, self_mode, pairs, max_pairs, counter,
                         info);
            overlap_item(sh, nxt, two, c1, q1i, a10, b10, a11, b11, nodes, nc, self_mode, pairs, max_pairs, counter,
                         info);
        }
        __syncthreads();  // the next list is complete
    }

    // flush the buffered pairs: one atomic for the workgroup
    const uint32_t nout = min(sh.nout, uint32_t(kOvPairBuf));
    if (tid == 0)
        sh.base = nout ? __hip_atomic_fetch_add(counter, int(nout), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0;
    __syncthreads();
    for (uint32_t i = uint32_t(tid); i < nout; i += kOvThreads) {
        const int64_t idx = int64_t(sh.base) + int64_t(i);
        if (idx < max_pairs) pairs[idx] = sh.out[i];
    }
}

}  // namespace

namespace {
// Below kOrderMinRays rays the origin sort costs more than the locality it buys.
const uint32_t* maybe_order(const BvhDev& d, RaySortDev* order, const lux_ray* rays, int n, hipStream_t s) {
    if (!order || n < kOrderMinRays || d.n < 2) return nullptr;
    return ray_order(*order, d, rays, n, s);
}

}  // namespace

const uint32_t* closest_order(RaySortDev& r, bool sort, const BvhDev& d, const lux_ray* rays, int n,
                              hipStream_t s) {
    return maybe_order(d, sort ? &r : nullptr, rays, n, s);
}

void trace_closest(const BvhDev& d, const uint32_t* order, const lux_triangle* tris, int tri_count,
                   const lux_ray* rays, int n, lux_hit* hits, hipStream_t s) {
    if (n <= 0) return;
    int node_count = d.n > 1 ? 2 * d.n : (d.n > 0 ? 1 : 0);  // sibling-pair slots
    // The near-child-first walk touches the nodes near each ray's origin first,
    // which neighbouring rays share: tracing in ray order (octant, then origin)
    // measured 5.12 ms against 5.60 ms in input order on 2^22 random rays over
    // 2^20 triangles, sort included.
    LUX_LAUNCH("trace_closest", k_trace_closest<true>, div_up(n, kBlock), kBlock, 0, s,
               (const lux_bvh_node*)d.pairs, node_count, (const uint32_t*)d.slot_pre, tris, tri_count, rays, n,
               order, hits, d.info);
}

void trace_any(const BvhDev& d, RaySortDev* order, const lux_triangle* tris, int tri_count, const lux_ray* rays,
               int n, uint8_t* occluded, hipStream_t s) {
    if (n <= 0) return;
    int node_count = d.n > 1 ? 2 * d.n : (d.n > 0 ? 1 : 0);  // sibling-pair slots
    const uint32_t* ord = maybe_order(d, order, rays, n, s);
    // Any hit walks near child first with its triangle tests gathered across
    // the wave: 4.52 ms against 4.66 ms for the gathered preorder walk on 2^22
    // random rays over 2^20 triangles (ungathered, near child first had lost:
    // 5.84 against 5.40 ms).
    LUX_LAUNCH("trace_any", k_trace_any, div_up(n, kBlock), kBlock, 0, s, (const lux_bvh_node*)d.pairs, node_count,
               tris, tri_count, rays, n, ord, occluded, d.info);
}

void trace_stats(const BvhDev& d, const lux_triangle* tris, int tri_count, const lux_ray* rays, int n,
                 unsigned long long* d_stats, hipStream_t s) {
    LUX_LAUNCH("trace_stats_zero", k_zero_u64, 1, 64, 0, s, d_stats, 6);
    if (n <= 0 || d.n <= 0) return;
    LUX_LAUNCH("trace_stats", k_trace_stats<true>, div_up(n, kBlock), kBlock, 0, s, (const lux_bvh_node*)d.pairs,
               d.n > 1 ? 2 * d.n : 1, (const uint32_t*)d.slot_pre, tris, tri_count, rays, n, d_stats, d.info);
}

void copy_device(void* dst, const void* src, size_t bytes, hipStream_t s) {
    size_t count = bytes / 16;
    if (count == 0) return;
    int grid = int(std::min<size_t>((count + kCopyBlock - 1) / kCopyBlock, 4096));
    LUX_LAUNCH("copy16", k_copy16, grid, kCopyBlock, 0, s, static_cast<const int4*>(src), static_cast<int4*>(dst),
               count);
}

void overlap_query(const BvhDev& tree, const BvhDev& queries, bool self_mode, int max_pairs,
                   lux_pair* pairs, int32_t* counter, hipStream_t s) {
    LUX_LAUNCH("overlap_zero", k_zero_counter, 1, 1, 0, s, counter);
    if (tree.n <= 0 || queries.n <= 0 || max_pairs <= 0) return;
    const int slots = tree.n > 1 ? 2 * tree.n : 1;  // sibling-pair array
    LUX_LAUNCH(self_mode ? "overlap_self" : "overlap_cross", k_overlap, div_up(queries.n, kOvQueries), kOvThreads,
               0, s, (const lux_bvh_node*)tree.pairs, slots, (const lux_aabb*)queries.boxes,
               (const int32_t*)queries.ids, (const uint32_t*)queries.vals[0], queries.n, self_mode ? 1 : 0, pairs,
               max_pairs, counter, tree.info);
}

}  // namespace lux

// pt.hip -- wavefront path tracer for gfx950: kernels and the batch engine.
//
// Stages per bounce (src/integrator/wavefront_scheduler.cu:73-131 is the
// reference's schedule; the transport rules are those of trace_path,
// src/integrator/cpu_path_tracer.cpp:403-498):
//   trace      closest hit of every live path (right-first stackless LBVH walk)
//              and, in the same launch, the shadow rays of the previous bounce
//   shade      emission + MIS, pass-through, termination, next-event estimation
//              (power-alias light choice, area sample -> shadow queue), BSDF
//              sample, Russian roulette -> next ray queue
// then one flush per batch adds each path's radiance to its pixel, samples of a
// pixel in sample order (so the film sum is the same float sum as the CPU's).
// Tracing bounce b's shadow rays with bounce b+1's extension rays keeps the
// order of the radiance additions (emission of b, NEE of b, emission of b+1)
// and hides the long tail of a small shadow launch behind the extension rays.
//
// A batch is a list of segments, each a contiguous range of one request's
// sample-major paths with its own scene (BVH root, mesh, lights, camera), so
// one launch sequence serves many scenes. The device reads the batch from a
// descriptor, which lets the whole bounce loop be captured once as a hipGraph
// and replayed per batch after a small descriptor copy.
//
// Queues are path-index lists appended with one atomic per wave (ballot +
// popcount). Two schedules share the kernels:
//   plain      the reference's: 65 536-path single-sample batches, full-grid
//              launches, one shade launch per material type, a separate shadow
//              launch, a host synchronisation + counter read-back per batch
//   wavefront  multi-sample, multi-request batches of up to `capacity` paths,
//              grid-stride kernels sized to the chip, one shade launch per
//              bounce, no host round trip until the caller synchronises
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "ordered.hpp"
#include "pt_device.hpp"
#include "pt_runtime.hpp"

namespace lux {
namespace pt {
namespace {

constexpr int kBlock = 256;
constexpr int kPlainBatch = 65536;  // RenderSettings::batch_size default (src/integrator/render_context.h:16-24)

struct Queues {
    uint32_t* q[2];
    uint32_t* counts;          // [0], [1] ray queues, [2], [3] shadow queues by bounce parity
    unsigned long long* rays;  // [0] extension rays, [1] shadow rays
};

// Path state, one slot per path of the batch (float4 records: one 16-byte load).
struct Paths {
    float4* T;      // throughput xyz, eta scale
    float4* L;      // radiance xyz, previous BSDF pdf
    float4* prev;   // previous vertex xyz, bits: flags | measure << 8 | depth << 16
    float4* ro;     // ray origin xyz, bits: pixel
    float4* rd;     // ray direction xyz, bits: segment of the batch
    ulonglong2* rng;
    float4* hit;    // bits: prim, t, u, v
    // shadow records, one per path at most per bounce
    float4* so;     // origin xyz, t_max
    float4* sd;     // direction xyz, MIS weight
    float4* sc;     // unoccluded contribution xyz, bits: path
    float4* st;     // throughput at the vertex xyz, bits: ignored primitive
};

// Closest hit / occlusion walk over the sibling-pair copy of the right-first
// stackless LBVH (same visits and order as the preorder walk); `ignore` skips
// one leaf's triangle test (the source primitive of a shadow ray). kParallel =
// false drops the axis-parallel branches of the slab test (no active lane of
// the wave has such an axis).
template <bool kParallel, bool kAny>
__device__ __forceinline__ int closest_walk(const SceneView& sv, const f3& O, const f3& D, const SlabRay& sr,
                                            float t_min, float t_max, int ignore, float& t_hit, float& u_hit,
                                            float& v_hit) {
    const lux_bvh_node* nodes = sv.nodes;
    const lux_triangle* soup = sv.soup;
    const int node_count = sv.node_count, tri_count = sv.tri_count;
    int best = -1;
    int k = 0;
    int guard = node_count + 1;  // a walk visits each node at most once
    while (uint32_t(k) < uint32_t(node_count)) {
        float4 a, b;
        load_node(nodes, k, a, b);
        int prim = __float_as_int(b.z);
        int esc = __float_as_int(b.w);
        int next;
        const bool pass = kParallel ? slab_hit(sr, t_max, a.x, a.y, a.z, a.w, b.x, b.y)
                                    : slab_hit_np(sr, t_max, a.x, a.y, a.z, a.w, b.x, b.y);
        if (!pass) {
            next = esc;
        } else if (prim >= 0) {
            next = esc;
            if (prim != ignore) {
                if (uint32_t(prim) >= uint32_t(tri_count)) {
                    set_error(sv.info, kErrPrimRange);
                } else {
                    const float* p = reinterpret_cast<const float*>(soup + prim);
                    f3 v0 = make3(p[0], p[1], p[2]), v1 = make3(p[3], p[4], p[5]), v2 = make3(p[6], p[7], p[8]);
                    float t, u, v;
                    if (tri_hit(O, D, t_min, t_max, v0, v1, v2, t, u, v)) {
                        best = prim;
                        t_max = t;
                        u_hit = u;
                        v_hit = v;
                        if (kAny) break;
                    }
                }
            }
        } else {
            next = ~prim;  // right child's slot
        }
        if (--guard == 0) {
            set_error(sv.info, kErrWalkOrder);
            break;
        }
        k = next;
    }
    t_hit = t_max;
    return best;
}

// The same closest hit by the near-child-first walk (ordered.hpp): both children
// of a node tested together, the farther one kept on this lane's LDS stack
// (entries `stride` words apart), exact-t ties kept by preorder position, the
// triangle tests gathered across the wave (C3 0.0714 -> 0.0706 s, the flattened
// instanced field 0.0538 -> 0.0522 s, same films). A ray whose stack
// overflows, or whose walk the check of ordered.hpp flags, takes closest_walk.
template <int kMode>
__device__ __forceinline__ int closest_ordered(const SceneView& sv, const f3& O, const f3& D, const SlabRay& sr,
                                               float t_min, float t_max0, int ignore, float& t_hit, float& u_hit,
                                               float& v_hit, uint32_t* stack, int stride) {
    OrderedBest ob(t_max0);
    int best = -1;
    uint32_t best_slot = 0, visits = 0;
    bool bad = false;
    const lux_triangle* soup = sv.soup;
    const uint32_t* slot_pre = sv.slot_pre;
    auto leaf = [&](uint32_t slot, int prim, float tn) {
        if (prim == ignore) return false;
        if (uint32_t(prim) >= uint32_t(sv.tri_count)) {
            set_error(sv.info, kErrPrimRange);
            return false;
        }
        const float* p = reinterpret_cast<const float*>(soup + prim);
        f3 v0 = make3(p[0], p[1], p[2]), v1 = make3(p[3], p[4], p[5]), v2 = make3(p[6], p[7], p[8]);
        float t, u, v;
        if (tri_uvt(O, D, v0, v1, v2, t, u, v) && t > t_min) {
            ob.note(t, tn);
            if (t < ob.t_max || (best >= 0 && t == ob.t_max && slot_pre[slot] < slot_pre[best_slot])) {
                ob.take(t);
                best = prim;
                best_slot = slot;
            }
        }
        return false;
    };
    const bool done = ordered_walk_gathered(SlabMode<kMode>{}, sv.nodes, uint32_t(sv.node_count), sr, ob, stack, stride,
                                   visits, bad, leaf);
    if (bad) set_error(sv.info, kErrWalkOrder);
    if (!done || (best >= 0 && ob.ambiguous()))
        return closest_walk<kMode == kSlabParallel, false>(sv, O, D, sr, t_min, t_max0, ignore, t_hit, u_hit, v_hit);
    t_hit = ob.t_max;
    if (best >= 0) {
        // u, v of the winner, recomputed once rather than carried through the
        // walk (the same operations give the same bits)
        const float* p = reinterpret_cast<const float*>(soup + best);
        float t;
        (void)tri_uvt(O, D, make3(p[0], p[1], p[2]), make3(p[3], p[4], p[5]), make3(p[6], p[7], p[8]), t, u_hit,
                      v_hit);
    }
    return best;
}

// stack == nullptr: the preorder walk only.
__device__ __forceinline__ int closest(const SceneView& sv, const vec& o, const vec& d, float t_min, float t_max,
                                       int ignore, float& t_hit, float& u_hit, float& v_hit, uint32_t* stack,
                                       int stride) {
    const f3 O = make3(o.x, o.y, o.z), D = make3(d.x, d.y, d.z);
    const SlabRay sr = make_slab_ray(O, D, t_min);
    const bool ordered =
        stack && !sv.preorder && sv.node_count > 1 && uint32_t(sv.node_count) <= kOffset32Nodes;
    if (ordered)
        return with_slab_mode(sr, [&](auto mode) {
            return closest_ordered<decltype(mode)::value>(sv, O, D, sr, t_min, t_max, ignore, t_hit, u_hit, v_hit,
                                                          stack, stride);
        });
    if (wave_any_parallel(sr)) return closest_walk<true, false>(sv, O, D, sr, t_min, t_max, ignore, t_hit, u_hit, v_hit);
    return closest_walk<false, false>(sv, O, D, sr, t_min, t_max, ignore, t_hit, u_hit, v_hit);
}

// True when the segment (t_min, t_max) hits any triangle other than `ignore`:
// the same walk stopped at its first accepted triangle. It accepts a triangle
// exactly when the closest-hit walk finds a hit, since both test the same
// leaves in the same order with the same t_max up to that first acceptance.
__device__ __forceinline__ bool occluded(const SceneView& sv, const vec& o, const vec& d, float t_min, float t_max,
                                         int ignore) {
    const f3 O = make3(o.x, o.y, o.z), D = make3(d.x, d.y, d.z);
    const SlabRay sr = make_slab_ray(O, D, t_min);
    float t = 0, u = 0, v = 0;
    if (wave_any_parallel(sr)) return closest_walk<true, true>(sv, O, D, sr, t_min, t_max, ignore, t, u, v) >= 0;
    return closest_walk<false, true>(sv, O, D, sr, t_min, t_max, ignore, t, u, v) >= 0;
}

// The same two queries on an instanced scene: the two-level walk (TLAS, then
// the BLAS of each instance the ray reaches, in object space), returning the
// scene ordinal. kTwoLevel is chosen per batch on the host, like kAnyShadow.
// stack as for closest().
template <bool kTwoLevel>
__device__ __forceinline__ int scene_closest(const SceneView& sv, const vec& o, const vec& d, float t_min,
                                             float t_max, int ignore, float& t_hit, float& u_hit, float& v_hit,
                                             uint32_t* stack, int stride) {
    if constexpr (kTwoLevel) {
        TwoLevelHit h;
        const bool ordered = stack && !sv.preorder;
        if (!ordered || !two_level_closest_ordered(sv.tl, make3(o.x, o.y, o.z), make3(d.x, d.y, d.z), t_min, t_max,
                                                   ignore, stack, stride, h))
            h = two_level_walk<false>(sv.tl, make3(o.x, o.y, o.z), make3(d.x, d.y, d.z), t_min, t_max, ignore);
        if (h.prim >= 0) {
            t_hit = h.t;
            u_hit = h.u;
            v_hit = h.v;
        } else {
            t_hit = t_max;
        }
        return h.prim;
    } else {
        return closest(sv, o, d, t_min, t_max, ignore, t_hit, u_hit, v_hit, stack, stride);
    }
}

template <bool kTwoLevel>
__device__ __forceinline__ bool scene_occluded(const SceneView& sv, const vec& o, const vec& d, float t_min,
                                               float t_max, int ignore) {
    if constexpr (kTwoLevel)
        return two_level_any(sv.tl, make3(o.x, o.y, o.z), make3(d.x, d.y, d.z), t_min, t_max, ignore);
    else
        return occluded(sv, o, d, t_min, t_max, ignore);
}

__device__ __forceinline__ vec xyz(const float4& f) { return mk(f.x, f.y, f.z); }
__device__ __forceinline__ float4 pack(const vec& v, float w) { return make_float4(v.x, v.y, v.z, w); }

// Segment holding batch slot `slot` (segments are sorted by slot_begin and
// cover [0, count) without gaps).
__device__ __forceinline__ uint32_t find_segment(const BatchDesc* d, uint32_t slot) {
    uint32_t lo = 0, hi = d->nseg;
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (d->seg[mid].slot_begin <= slot) lo = mid; else hi = mid;
    }
    return lo;
}

__device__ __forceinline__ uint32_t segment_of_path(const Paths& ps, uint32_t path) {
    return __float_as_uint(reinterpret_cast<const float*>(ps.rd + path)[3]);
}

// ---- generate ------------------------------------------------------------------
// Batch slot j of segment s is request path r = req_begin + (j - slot_begin):
// sample r / npix, pixel r % npix. Camera sample dimensions as
// render_cpu_path_tracer (src/integrator/cpu_path_tracer.cpp:532-542).
template <int B>
__global__ __launch_bounds__(B) void k_generate(const BatchDesc* __restrict__ desc, Paths ps, Queues qs) {
    const uint32_t total = desc->count;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        qs.counts[0] = total;
        qs.counts[1] = 0;
        qs.counts[2] = 0;
        qs.counts[3] = 0;
    }
    for (uint32_t j = xcd_block() * B + threadIdx.x; j < total; j += gridDim.x * B) {
        const uint32_t si = find_segment(desc, j);
        const Segment& sg = desc->seg[si];
        const int width = sg.width, height = sg.height;
        const uint64_t npix = uint64_t(width) * uint64_t(height);
        const uint64_t r = sg.req_begin + (j - sg.slot_begin);
        const int sample = int(r / npix);
        const int pixel = int(r - uint64_t(sample) * npix);
        Rng rng = make_rng(sg.seed, sg.frame, pixel, sample + sg.sample_offset);
        int x = pixel % width, y = pixel / width;
        float jx = rng.get1();
        float jy = rng.get1();
        float u = (float(x) + jx) / float(width);
        float v = (float(height - 1 - y) + jy) / float(height);
        rng.get1();  // lens u
        rng.get1();  // lens v
        rng.get1();  // time
        const DevCamera& c = sg.sv.cam;
        vec target = mk(c.lx, c.ly, c.lz) + u * mk(c.hx, c.hy, c.hz) + v * mk(c.vx, c.vy, c.vz);
        vec origin = mk(c.ox, c.oy, c.oz);
        vec dir = target - origin;
        float len = length(dir);
        dir = len > 0 ? dir / len : dir;
        ps.T[j] = make_float4(1, 1, 1, 1);
        ps.L[j] = make_float4(0, 0, 0, 0);
        ps.prev[j] = make_float4(0, 0, 0, __int_as_float(0));
        ps.ro[j] = pack(origin, __int_as_float(pixel));
        ps.rd[j] = pack(dir, __uint_as_float(si));
        ps.rng[j] = make_ulonglong2(rng.state, rng.inc);
        qs.q[0][j] = j;
    }
}

// ---- shadow rays -------------------------------------------------------------------
// trace_shadow (src/integrator/cpu_path_tracer.cpp:27-72): repeated closest hits
// through pass-through surfaces; radiance += T * (contribution * Tr * mis).
// kAnyShadow (launched only for batches whose scenes have no pass-through
// material): the first hit found blocks the ray, so an any-hit walk gives the
// same transmittance (0 or 1) as the closest-hit loop. The choice is made per
// batch on the host; reading the scene's flag per ray instead measured slower.
template <bool kAnyShadow, bool kTwoLevel>
__device__ __forceinline__ void shadow_ray(const BatchDesc* __restrict__ desc, const Paths& ps, uint32_t i,
                                           uint32_t* stack, int stride) {
    float4 so = ps.so[i], sd = ps.sd[i], sc = ps.sc[i], st = ps.st[i];
    const uint32_t path = __float_as_uint(sc.w);
    const SceneView& sv = desc->seg[segment_of_path(ps, path)].sv;
    vec o = xyz(so), d = xyz(sd);
    float t_max = so.w, t_min = kEps;
    int ignore = __float_as_int(st.w);
    vec tr = splat(1);
    if (kAnyShadow && scene_occluded<kTwoLevel>(sv, o, d, t_min, t_max, ignore)) tr = splat(0);
    for (int guard = 0; !kAnyShadow && guard < 1 << 16; ++guard) {
        float t = 0, u = 0, v = 0;
        // the preorder walk (no stack): with the near-child-first walk inlined a
        // second time here the one-level kernel spilled 36 VGPRs per lane
        int prim = scene_closest<kTwoLevel>(sv, o, d, t_min, t_max, ignore, t, u, v, nullptr, 0);
        if (prim < 0) break;
        const DevMaterial& m = sv.mats[sv.mat_id[prim]];
        if (!passthrough(m)) {
            tr = splat(0);
            break;
        }
        tr = tr * (m.type == kTransparent ? albedo_of(m) : splat(1));
        if (max_component(tr) <= 0) break;
        float adv = t + kEps;
        if (adv >= t_max) break;
        o = o + d * adv;
        t_max -= adv;
        t_min = kEps;
        ignore = -1;
    }
    if (max_component(tr) > 0) {
        float4 Lq = ps.L[path];
        vec add = xyz(st) * (xyz(sc) * tr * sd.w);
        ps.L[path] = pack(xyz(Lq) + add, Lq.w);
    }
}

// ---- trace -----------------------------------------------------------------------
// Extension rays of queue `ext_in` (-1: none) and then the shadow records
// counted by counts[sh_in] (-1: none), in one grid-stride pass. Block 0 clears
// the counters the following shade launch appends to (zero_a, zero_b; never
// the ones read here) and adds the queue sizes to the ray totals.
// One-level variants: at least 8 waves per SIMD (64 VGPRs, 65 uncapped). C3
// measured 0.079 s at 8 waves against 0.096 s at 7. Two-level variants: 6
// waves (80 VGPRs, 86 uncapped, 28-40 bytes per lane spilled); the instanced
// field measured 91.6 ms at 6 waves, 96.7 ms at 7 and 113 ms at the natural 5.
template <int B, bool kAnyShadow, bool kTwoLevel>
__global__ __launch_bounds__(B) __attribute__((amdgpu_waves_per_eu(kTwoLevel ? 6 : 8))) void k_trace(const BatchDesc* __restrict__ desc, Paths ps, Queues qs, int ext_in,
                                                  int sh_in, int zero_a, int zero_b) {
    const uint32_t n_ext = ext_in >= 0 ? qs.counts[ext_in] : 0u;
    const uint32_t n_sh = sh_in >= 0 ? qs.counts[sh_in] : 0u;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        if (zero_a >= 0) qs.counts[zero_a] = 0;
        if (zero_b >= 0) qs.counts[zero_b] = 0;
        qs.rays[0] += n_ext;
        qs.rays[1] += n_sh;
    }
    // per-lane stacks of the ordered closest-hit walks
    __shared__ uint32_t lds_stack[kOrderedStack * B];
    uint32_t* stack = lds_stack + threadIdx.x;
    const uint32_t total = n_ext + n_sh;
    for (uint32_t i = xcd_block() * B + threadIdx.x; i < total; i += gridDim.x * B) {
        if (i < n_ext) {
            const uint32_t path = qs.q[ext_in][i];
            const float4 o = ps.ro[path], d = ps.rd[path];
            const SceneView& sv = desc->seg[__float_as_uint(d.w)].sv;
            float t = INFINITY, u = 0, v = 0;
            int prim = scene_closest<kTwoLevel>(sv, xyz(o), xyz(d), kEps, INFINITY, -1, t, u, v, stack, B);
            ps.hit[path] = make_float4(__int_as_float(prim), t, u, v);
        } else {
            shadow_ray<kAnyShadow, kTwoLevel>(desc, ps, i - n_ext, stack, B);
        }
    }
}

// ---- light selection (LightSamplerKind) ---------------------------------------------
constexpr int kLightPower = 0, kLightUniform = 1, kLightBvh = 2;

// light_bvh_node_importance (src/light/light_sampler.h:91-115): power over the
// squared distance from the shading point to the node's box (at least eps^2).
__device__ __forceinline__ float light_node_importance(const SceneView& sv, int i, const vec& p) {
    if (i < 0 || i >= sv.light_node_count) return 0.0f;
    const DevLightNode& n = sv.light_nodes[i];
    if (n.power <= 0 || n.lo[0] > n.hi[0] || n.lo[1] > n.hi[1] || n.lo[2] > n.hi[2]) return 0.0f;
    const float q[3] = {p.x, p.y, p.z};
    float d2 = 0;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        float delta = 0;
        if (q[a] < n.lo[a])
            delta = n.lo[a] - q[a];
        else if (q[a] > n.hi[a])
            delta = q[a] - n.hi[a];
        d2 += delta * delta;
    }
    d2 = fmaxf(d2, kEps * kEps);
    return n.power / d2;
}

// light_bvh_child_probability (src/light/light_sampler.h:117-140).
__device__ __forceinline__ float light_child_probability(const SceneView& sv, int parent, int child,
                                                         const vec& p) {
    if (parent < 0 || parent >= sv.light_node_count || child < 0 || child >= sv.light_node_count) return 0.0f;
    const DevLightNode& n = sv.light_nodes[parent];
    if (child != n.left && child != n.right) return 0.0f;
    const float li = light_node_importance(sv, n.left, p);
    const float ri = light_node_importance(sv, n.right, p);
    const float total = li + ri;
    if (total <= 0) return 0.0f;
    const float lp = li / total;
    return child == n.left ? lp : 1.0f - lp;
}

// sample_light_bvh (src/light/light_sampler.h:142-195): descend from the root,
// choosing a child by its importance at the shading point and rescaling u.
__device__ __forceinline__ int sample_light_bvh(const SceneView& sv, const vec& p, float select_u, float& pmf) {
    pmf = 0;
    int node = sv.light_root;
    if (node < 0 || node >= sv.light_node_count) return -1;
    float u = fminf(select_u, 0.99999994f);
    pmf = 1;
    for (int depth = 0; depth < sv.light_node_count; ++depth) {
        const DevLightNode& n = sv.light_nodes[node];
        if (n.light >= 0) return n.light;
        const float lp = light_child_probability(sv, node, n.left, p);
        if (lp <= 0) {
            node = n.right;
        } else if (lp >= 1) {
            node = n.left;
        } else if (u < lp) {
            pmf *= lp;
            u /= lp;
            node = n.left;
        } else {
            const float rp = 1.0f - lp;
            pmf *= rp;
            u = (u - lp) / rp;
            node = n.right;
        }
        if (node < 0 || node >= sv.light_node_count) {
            pmf = 0;
            return -1;
        }
    }
    pmf = 0;
    return -1;
}

// light_bvh_pmf (src/light/light_sampler.h:197-223): the product of the child
// probabilities from the light's leaf up to the root, at point p.
__device__ __forceinline__ float light_bvh_pmf(const SceneView& sv, int light, const vec& p) {
    if (light < 0 || light >= sv.light_count) return 0.0f;
    int node = sv.light_leaf[light];
    if (node < 0 || node >= sv.light_node_count) return 0.0f;
    float pmf = 1;
    for (int depth = 0; depth < sv.light_node_count; ++depth) {
        const int parent = sv.light_nodes[node].parent;
        if (parent < 0) return pmf;
        const float prob = light_child_probability(sv, parent, node, p);
        if (prob <= 0) return 0.0f;
        pmf *= prob;
        node = parent;
        if (node < 0 || node >= sv.light_node_count) return 0.0f;
    }
    return 0.0f;
}

// Alias-table draw over `count` entries (sample_alias_table_index,
// src/light/light_sampler.h:71-82).
__device__ __forceinline__ int alias_draw(const DevAlias* t, int count, float u) {
    const float scaled = fminf(u, 0.99999994f) * float(count);
    int col = int(scaled);
    if (col >= count) col = count - 1;
    const float thr = scaled - float(col);
    return thr < t[col].prob ? col : t[col].alias;
}

// The BVH sampler with its non-BVH group (sample_light_bvh_mixture,
// src/light/light_sampler.h:263-316): the tree's root power against the
// group's total weight picks the tree or the group's alias table.
__device__ __forceinline__ int sample_light_mixture(const SceneView& sv, const vec& p, float select_u, float& pmf) {
    pmf = 0;
    const float bvh_power = sv.light_nodes[sv.light_root].power;
    if (bvh_power <= 0) return -1;
    const float nb_power = sv.nb_count > 0 ? fmaxf(0.0f, sv.nb_total) : 0.0f;
    const float total = bvh_power + nb_power;
    if (total <= 0) return -1;
    const float group = bvh_power / total;
    if (nb_power <= 0 || group >= 1) return sample_light_bvh(sv, p, select_u, pmf);
    auto draw_group = [&](float u, float& gp) -> int {
        gp = 0;
        if (sv.nb_count <= 0 || nb_power <= 0) return -1;
        const int slot = alias_draw(sv.nb_table, sv.nb_count, u);
        if (slot < 0 || slot >= sv.nb_count) return -1;
        gp = sv.nb_table[slot].pmf;
        return gp > 0 ? sv.nb_table[slot].light : -1;
    };
    if (group <= 0) return draw_group(select_u, pmf);
    const float u = fminf(select_u, 0.99999994f);
    if (u < group) {
        float bp = 0;
        const int idx = sample_light_bvh(sv, p, u / group, bp);
        pmf = group * bp;
        return pmf > 0 ? idx : -1;
    }
    const float other = 1.0f - group;
    float gp = 0;
    const int idx = draw_group((u - group) / other, gp);
    pmf = other * gp;
    return pmf > 0 ? idx : -1;
}

// light_bvh_mixture_pmf (src/light/light_sampler.h:318-358).
__device__ __forceinline__ float light_mixture_pmf(const SceneView& sv, int light, const vec& p) {
    if (light < 0) return 0.0f;
    const float bvh_power = sv.light_nodes[sv.light_root].power;
    if (bvh_power <= 0) return 0.0f;
    const float nb_power = sv.nb_count > 0 ? fmaxf(0.0f, sv.nb_total) : 0.0f;
    const float total = bvh_power + nb_power;
    if (total <= 0) return 0.0f;
    const float group = bvh_power / total;
    if (light < sv.light_count && sv.light_leaf[light] >= 0) return group * light_bvh_pmf(sv, light, p);
    const int slot = light < sv.light_count ? sv.lights[light].nb_slot : -1;
    if (slot < 0 || nb_power <= 0) return 0.0f;
    return (1.0f - group) * sv.nb_table[slot].pmf;
}

// sample_light (src/scene/scene_light_sampler.h:99-152): the light index and
// its selection pmf. Uniform: u * count; BVH: the mixture of the tree descent
// at the shading point and the non-BVH group (power selection when the scene
// has no valid tree); power: the alias table.
template <int kLight>
__device__ __forceinline__ int choose_light(const SceneView& sv, const vec& p, float select_u, float& pmf) {
    const int nl = sv.light_count;
    pmf = 0;
    if (nl <= 0) return -1;
    if constexpr (kLight == kLightUniform) {
        int idx = int(fminf(select_u, 0.99999994f) * float(nl));
        if (idx >= nl) idx = nl - 1;
        pmf = 1.0f / float(nl);
        return idx;
    } else {
        if (kLight == kLightBvh && sv.light_root >= 0) {
            const int idx = sample_light_mixture(sv, p, select_u, pmf);
            if (idx >= 0 && !(pmf > 0)) pmf = light_mixture_pmf(sv, idx, p);  // a descent pmf that underflowed
            return idx;
        }
        const float scaled = fminf(select_u, 0.99999994f) * float(nl);
        int col = int(scaled);
        if (col >= nl) col = nl - 1;
        const float thr = scaled - float(col);
        const int idx = thr < sv.lights[col].prob ? col : sv.lights[col].alias;
        if (idx < 0 || idx >= nl) return -1;
        pmf = sv.lights[idx].pmf;
        return idx;
    }
}

// light_pmf (src/scene/scene_light_sampler.h:158-203) with the reference point
// of the context (the BVH pmf depends on it).
template <int kLight>
__device__ __forceinline__ float light_choice_pmf(const SceneView& sv, int light, const vec& ref) {
    if constexpr (kLight == kLightUniform) {
        return 1.0f / float(sv.light_count);
    } else {
        if (kLight == kLightBvh && sv.light_root >= 0) return light_mixture_pmf(sv, light, ref);
        return sv.lights[light].pmf;
    }
}

// ---- scene lights (point, distant, uniform and image infinite) ------------------------
// uniform_sphere_direction, the lat-long maps and image_env_lookup
// (src/light/light.h:137-166).
__device__ __forceinline__ vec uniform_sphere(float ux, float uy) {
    const float z = 1.0f - 2.0f * ux;
    const float r = sqrtf(fmaxf(0.0f, 1.0f - z * z));
    const float phi = 2.0f * kPi * uy;
    return mk(r * cosf(phi), z, r * sinf(phi));
}

__device__ __forceinline__ void latlong_uv(const vec& wi, float& u, float& v) {
    const vec w = normalize(wi);
    float phi = atan2f(w.z, w.x);
    if (phi < 0) phi += 2.0f * kPi;
    const float theta = acosf(fmaxf(-1.0f, fminf(1.0f, w.y)));
    u = phi * (0.5f * kInvPi);
    v = theta * kInvPi;
}

__device__ __forceinline__ vec latlong_direction(float u, float v) {
    const float phi = 2.0f * kPi * u;
    const float theta = kPi * v;
    const float st = sinf(theta);
    return mk(cosf(phi) * st, cosf(theta), sinf(phi) * st);
}

__device__ __forceinline__ int image_texel(const DevLight& lt, float u, float v) {
    const int w = lt.two_sided, h = lt.height;
    const int x = int(fminf(float(w - 1), floorf(u * float(w))));
    const int y = int(fminf(float(h - 1), floorf(v * float(h))));
    return y * w + x;
}

__device__ __forceinline__ vec image_lookup(const SceneView& sv, const DevLight& lt, const vec& wi) {
    if (lt.two_sided <= 0 || lt.height <= 0) return splat(0);
    float u, v;
    latlong_uv(wi, u, v);
    const float* px = sv.env_pixels + 3 * size_t(lt.prim + image_texel(lt, u, v));
    return mk(px[0], px[1], px[2]) * mk(lt.er, lt.eg, lt.eb);
}

// image_infinite_pdf_for_texel (src/light/light.h:205-222).
__device__ __forceinline__ float image_texel_pdf(const SceneView& sv, const DevLight& lt, int index, float v) {
    const int w = lt.two_sided, h = lt.height;
    const float total = lt.area;
    if (w <= 0 || h <= 0 || total <= 0) return 0.0f;
    const int count = w * h;
    if (index < 0 || index >= count) return 0.0f;
    const float* cdf = sv.env_cdf + lt.prim;
    const float weight = index == 0 ? cdf[0] : cdf[index] - cdf[index - 1];
    if (weight <= 0) return 0.0f;
    const float sin_theta = fmaxf(1e-6f, sinf(kPi * v));
    const float uv_pdf = (weight / total) * float(count);
    return uv_pdf / (2.0f * kPi * kPi * sin_theta);
}

// Radiance arriving along a ray that leaves the scene from one infinite light
// (infinite_light_le) and its solid-angle pdf for that direction (light_pdf_li).
__device__ __forceinline__ vec env_le(const SceneView& sv, const DevLight& lt, const vec& d) {
    if (lt.kind == kLightUniformInf) return mk(lt.er, lt.eg, lt.eb);
    if (lt.kind == kLightImageInf) return image_lookup(sv, lt, d);
    return splat(0);
}

__device__ __forceinline__ float env_pdf(const SceneView& sv, const DevLight& lt, const vec& d) {
    if (lt.kind == kLightUniformInf) return 1.0f / (4.0f * kPi);
    if (lt.kind == kLightImageInf) {
        if (lt.two_sided <= 0 || lt.height <= 0 || lt.area <= 0) return 0.0f;
        float u, v;
        latlong_uv(d, u, v);
        return image_texel_pdf(sv, lt, image_texel(lt, u, v), v);
    }
    return 0.0f;
}

// A direct-light sample of a scene light (sample_light_li, src/light/light.h:
// 166-193, 224-253): direction, distance (INFINITY for distant and infinite
// lights), radiance, pdf before the selection pmf, delta flag, light position.
struct SceneLightSample {
    vec wi, rad, pos;
    float dist, pdf;
    bool delta, valid;
};

__device__ __forceinline__ SceneLightSample sample_scene_light(const SceneView& sv, const DevLight& lt, const vec& p,
                                                               float comp_u, float lu0, float lu1) {
    SceneLightSample r{};
    r.valid = false;
    if (lt.kind == kLightPoint) {
        r.pos = mk(lt.vx, lt.vy, lt.vz);
        const vec to = r.pos - p;
        const float d2 = length2(to);
        if (d2 <= kEps * kEps) return r;
        r.dist = sqrtf(d2);
        r.wi = to / r.dist;
        r.rad = mk(lt.er, lt.eg, lt.eb) / d2;
        r.pdf = 1;
        r.delta = true;
        r.valid = max_component(r.rad) > 0;
    } else if (lt.kind == kLightDistant) {
        r.wi = normalize(mk(lt.vx, lt.vy, lt.vz));
        r.dist = INFINITY;
        r.rad = mk(lt.er, lt.eg, lt.eb);
        r.pdf = 1;
        r.delta = true;
        r.valid = max_component(r.rad) > 0;
    } else if (lt.kind == kLightUniformInf) {
        r.wi = uniform_sphere(lu0, lu1);
        r.dist = INFINITY;
        r.rad = mk(lt.er, lt.eg, lt.eb);
        r.pdf = 1.0f / (4.0f * kPi);
        r.delta = false;
        r.valid = max_component(r.rad) > 0;
    } else if (lt.kind == kLightImageInf) {
        const int w = lt.two_sided, h = lt.height;
        if (w <= 0 || h <= 0 || lt.area <= 0) return r;
        const int count = w * h;
        const float target = fminf(comp_u, 0.99999994f) * lt.area;
        const float* cdf = sv.env_cdf + lt.prim;
        int lo = 0, hi = count - 1;  // first texel with target <= cdf (the cdf never decreases)
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (target <= cdf[mid])
                hi = mid;
            else
                lo = mid + 1;
        }
        const int index = lo;
        const int x = index % w, y = index / w;
        const float u = (float(x) + lu0) / float(w), v = (float(y) + lu1) / float(h);
        r.wi = latlong_direction(u, v);
        r.dist = INFINITY;
        r.rad = image_lookup(sv, lt, r.wi);
        r.pdf = image_texel_pdf(sv, lt, index, v);
        r.delta = false;
        r.valid = max_component(r.rad) > 0 && r.pdf > 0;
    }
    return r;
}

// ---- shade -----------------------------------------------------------------------
// Shading key of a queue entry: 0 = miss, 1 + material type for a hit, and
// kShadeKeys - 1 for the idle lanes past the end of the queue.
constexpr int kShadeKeys = kMaterialTypes + 2;

// One bounce of trace_path for every live path of queue `in`. `filter` >= 0
// restricts the launch to hits of that material type (misses go with filter 0).
// Next rays go to queue in ^ 1, shadow records to counts[2 + (bounce & 1)].
//
// Each round, a block first orders its entries by shading key (a stable
// counting sort in LDS: ballots per key and wave, one scan), so most waves
// shade a single material and run one BSDF branch instead of all of them.
// Paths carry all their state, so which lane shades a path does not change its
// result.
// Capped at 128 VGPRs (4 waves per SIMD; 12 bytes spill), which measured
// faster than 134 VGPRs at 3 waves.
template <int B, int kLight>
__global__ __launch_bounds__(B) __attribute__((amdgpu_waves_per_eu(4))) void k_shade(const BatchDesc* __restrict__ desc, Paths ps, Queues qs, int in,
                                                  int bounce, int filter) {
    constexpr int kWaves = B / 64;
    __shared__ uint32_t s_offset[kShadeKeys][kWaves];
    __shared__ uint32_t s_path[B];
    __shared__ float4 s_hit[B];
    const uint32_t count = qs.counts[in];
    uint32_t* const sh_count = qs.counts + 2 + (bounce & 1);
    const uint32_t stride = gridDim.x * B;
    const uint32_t rounds = (count + stride - 1) / stride;
    const uint32_t wave = threadIdx.x / 64u;
    for (uint32_t r = 0; r < rounds; ++r) {  // uniform trip count: barriers and ballots see the whole block
        {
            const uint32_t i = r * stride + xcd_block() * B + threadIdx.x;
            const bool in_queue = i < count;
            const uint32_t qpath = in_queue ? qs.q[in][i] : 0xffffffffu;
            const float4 qhit = in_queue ? ps.hit[qpath] : make_float4(__int_as_float(-1), 0, 0, 0);
            const int qprim = __float_as_int(qhit.x);
            int key = kShadeKeys - 1;
            if (in_queue) {
                key = 0;
                if (qprim >= 0) {
                    const SceneView& qsv = desc->seg[segment_of_path(ps, qpath)].sv;
                    key = 1 + min(max(qsv.mats[qsv.mat_id[qprim]].type, 0), kMaterialTypes - 1);
                }
            }
            uint32_t rank = 0;
#pragma unroll
            for (int k = 0; k < kShadeKeys; ++k) {
                const uint64_t m = ballot64(key == k);
                if (key == k) rank = uint32_t(__popcll(m & lanemask_lt()));
                if (lane_id() == 0) s_offset[k][wave] = uint32_t(__popcll(m));
            }
            __syncthreads();
            if (threadIdx.x == 0) {  // exclusive offsets in (key, wave) order
                uint32_t acc = 0;
                for (int k = 0; k < kShadeKeys; ++k)
                    for (int w = 0; w < kWaves; ++w) {
                        const uint32_t c = s_offset[k][w];
                        s_offset[k][w] = acc;
                        acc += c;
                    }
            }
            __syncthreads();
            const uint32_t dst = s_offset[key][wave] + rank;
            s_path[dst] = qpath;
            s_hit[dst] = qhit;
            __syncthreads();
        }
        uint32_t path = s_path[threadIdx.x];
        float4 hit = s_hit[threadIdx.x];
        __syncthreads();  // the next round rewrites the arrays
        bool live = path != 0xffffffffu;
        if (!live) path = 0;
        bool push_ray = false, push_shadow = false;
        int prim = __float_as_int(hit.x);
        const Segment* sg = live ? &desc->seg[segment_of_path(ps, path)] : nullptr;
        const DevMaterial* mp = nullptr;
        if (live && prim >= 0) mp = &sg->sv.mats[sg->sv.mat_id[prim]];
        if (live && filter >= 0) live = (prim < 0) ? filter == 0 : mp->type == filter;
        // NEE record and next ray, filled below
        vec sh_o = splat(0), sh_d = splat(0), sh_c = splat(0), sh_T = splat(0);
        float sh_tmax = 0, sh_mis = 0;
        if (live && prim >= 0) {
            const SceneView& sv = sg->sv;
            const int max_depth = sg->max_depth;
            const DevMaterial m = *mp;
            float4 Tq = ps.T[path], Lq = ps.L[path], pq = ps.prev[path], oq = ps.ro[path], dq = ps.rd[path];
            ulonglong2 rs = ps.rng[path];
            Rng rng{rs.x, rs.y};
            vec T = xyz(Tq), L = xyz(Lq), prev_pos = xyz(pq);
            float eta_scale = Tq.w, prev_pdf = Lq.w;
            uint32_t pbits = __float_as_uint(pq.w);
            unsigned prev_flags = pbits & 0xffu;
            int prev_measure = int((pbits >> 8) & 0xffu);
            int depth = int(pbits >> 16);
            vec o = xyz(oq), d = xyz(dq);
            float t = hit.y, hu = hit.z, hv = hit.w;
            // finish_surface_hit_from_triangle (src/scene/scene_intersection.h:54-78)
            vec p = o + d * t;
            const lux_triangle tri = sv.soup[prim];
            vec ng = tri_normal(tri);
            vec ns = ng;
            if (sv.normals) {
                const uint32_t* ix = sv.indices + 3 * size_t(prim);
                const float* nb = sv.normals;
                vec n0 = mk(nb[3 * ix[0]], nb[3 * ix[0] + 1], nb[3 * ix[0] + 2]);
                vec n1 = mk(nb[3 * ix[1]], nb[3 * ix[1] + 1], nb[3 * ix[1] + 2]);
                vec n2 = mk(nb[3 * ix[2]], nb[3 * ix[2] + 1], nb[3 * ix[2] + 2]);
                if (length2(n0) > 0 && length2(n1) > 0 && length2(n2) > 0) {
                    float w = 1.0f - hu - hv;
                    vec n = w * n0 + hu * n1 + hv * n2;
                    ns = length2(n) == 0 ? ng : normalize(n);
                }
            }
            if (dot(ng, ns) < 0) ns = -ns;
            vec wo = -d;
            // emission (surface_light_le, src/scene/scene_light_sampler.h:283-336)
            int li = sv.light_of_prim ? sv.light_of_prim[prim] : -1;
            vec Le = splat(0);
            if (li >= 0 && li < sv.light_count) {
                const DevLight& lt = sv.lights[li];
                Le = eval_emission(mk(lt.er, lt.eg, lt.eb), lt.two_sided, ng, wo);
            } else if (max_component(emission_of(m)) > 0) {
                Le = eval_emission(emission_of(m), m.two_sided, ng, wo);
            }
            if (max_component(Le) > 0 && (depth == 0 || is_delta(prev_flags))) {
                L = L + T * Le;
            } else if (max_component(Le) > 0 && prev_measure == kSolidAngle && prev_pdf > 0) {
                float lpdf = 0;
                if (li >= 0 && li < sv.light_count && sv.lights[li].area > 0) {
                    const DevLight& lt = sv.lights[li];
                    lpdf = light_choice_pmf<kLight>(sv, li, prev_pos) *
                           pdf_solid_angle(p, ng, lt.two_sided, 1.0f / lt.area, prev_pos);
                }
                L = L + T * Le * power_heuristic(prev_pdf, lpdf);
            }
            bool alive = true;
            if (passthrough(m)) {
                T = T * (m.type == kTransparent ? albedo_of(m) : splat(1));
                if (max_component(T) <= 0) {
                    alive = false;
                } else {
                    o = p + d * kEps;
                    push_ray = true;
                }
            } else if (!scattering(m) || bounce >= max_depth || depth >= max_depth) {
                alive = false;
            } else {
                Interaction it{p, ng, ns, wo};
                if (direct_lobe(m)) {
                    float sel = rng.get1();
                    float comp_u = rng.get1();  // light component (image lights pick their texel with it)
                    float lu0 = rng.get1();
                    float lu1 = rng.get1();
                    // sample_light_li: light pick, area sample (src/scene/scene_light_sampler.h:210-250)
                    const int nl = sv.light_count;
                    if (nl > 0) {
                        float sel_pmf = 0;
                        const int idx = choose_light<kLight>(sv, p, sel, sel_pmf);
                        if (idx >= 0 && idx < nl && sel_pmf > 0) {
                            const DevLight& lt = sv.lights[idx];
                            vec l_wi = splat(0), l_rad = splat(0), l_target = splat(0);
                            float l_dist = 0, lpdf = 0;
                            bool l_delta = false, l_ok = false;
                            if (lt.kind == kLightArea) {
                                const lux_triangle lt_tri = sv.soup[lt.prim];
                                float su = sqrtf(lu0);
                                float b0 = 1.0f - su, b1 = su * (1.0f - lu1), b2 = su * lu1;
                                vec lp = b0 * tri_vertex(lt_tri, 0) + b1 * tri_vertex(lt_tri, 1) +
                                         b2 * tri_vertex(lt_tri, 2);
                                vec ln = tri_normal(lt_tri);
                                float pdf_area = lt.area > 0 ? 1.0f / lt.area : 0.0f;
                                vec to = lp - p;
                                float d2 = length2(to);
                                if (d2 > kEps * kEps) {
                                    l_dist = sqrtf(d2);
                                    l_wi = to / l_dist;
                                    l_rad = eval_emission(mk(lt.er, lt.eg, lt.eb), lt.two_sided, ln, -l_wi);
                                    lpdf = sel_pmf * pdf_solid_angle(lp, ln, lt.two_sided, pdf_area, p);
                                    l_target = lp;
                                    l_ok = max_component(l_rad) > 0 && lpdf > 0;
                                }
                            } else {
                                const SceneLightSample ls = sample_scene_light(sv, lt, p, comp_u, lu0, lu1);
                                l_wi = ls.wi;
                                l_rad = ls.rad;
                                l_dist = ls.dist;
                                l_target = ls.pos;
                                l_delta = ls.delta;
                                lpdf = ls.pdf * sel_pmf;
                                l_ok = ls.valid && lpdf > 0;
                            }
                            if (l_ok) {
                                // estimate_unoccluded_direct_light_li (src/light/light_sampler.h:420-454)
                                vec f = eval_bsdf(m, it, l_wi);
                                float cs = abs_cos(ns, l_wi);
                                float pdf_b = pdf_bsdf(m, it, l_wi);
                                if (max_component(f) > 0 && cs > 0) {
                                    vec off = dot(l_wi, ng) >= 0 ? ng : -ng;
                                    vec so = p + off * kEps;
                                    if (l_dist < INFINITY) {  // trace_shadow to the light point
                                        vec tl = l_target - so;
                                        float sdist = length(tl);
                                        if (sdist > kEps) {
                                            sh_d = tl / sdist;
                                            sh_tmax = sdist - kEps;
                                            push_shadow = true;
                                        }
                                    } else {  // trace_shadow_ray along the light direction
                                        sh_d = l_wi;
                                        sh_tmax = INFINITY;
                                        push_shadow = true;
                                    }
                                    sh_o = so;
                                    sh_c = f * l_rad * (cs / lpdf);
                                    sh_mis = l_delta ? 1.0f : power_heuristic(lpdf, pdf_b);
                                    sh_T = T;
                                }
                            }
                        }
                    }
                }
                float bu0 = rng.get1();
                float bu1 = rng.get1();
                BsdfSample bs = sample_bsdf(m, it, bu0, bu1);
                if (bs.pdf <= 0 || max_component(bs.weight) <= 0) {
                    alive = false;
                } else {
                    vec nT = T * bs.weight;
                    float n_eta = eta_scale * ((!is_transmission(bs.flags) || bs.eta <= 0) ? 1.0f : bs.eta * bs.eta);
                    if (depth > 2) {
                        float rr = fminf(0.95f, max_component(nT) * n_eta);
                        if (!(rng.get1() <= rr)) alive = false;
                        nT = nT / rr;
                    }
                    if (alive) {
                        T = nT;
                        eta_scale = n_eta;
                        prev_pdf = bs.pdf;
                        prev_measure = bs.measure;
                        prev_flags = bs.flags;
                        prev_pos = p;
                        vec off = dot(bs.wi, ng) >= 0 ? ng : -ng;
                        o = p + off * kEps;
                        d = bs.wi;
                        ++depth;
                        push_ray = true;
                    }
                }
            }
            (void)alive;
            ps.T[path] = pack(T, eta_scale);
            ps.L[path] = pack(L, prev_pdf);
            ps.prev[path] = pack(prev_pos, __uint_as_float(prev_flags | (uint32_t(prev_measure) << 8) |
                                                            (uint32_t(depth) << 16)));
            ps.ro[path] = pack(o, oq.w);
            ps.rd[path] = pack(d, dq.w);
            ps.rng[path] = make_ulonglong2(rng.state, rng.inc);
        }
        if (live && prim < 0 && sg->sv.env_count > 0) {
            // a ray that left the scene: infinite lights (trace_path's miss branch,
            // src/integrator/cpu_path_tracer.cpp:416-427)
            const SceneView& sv = sg->sv;
            const float4 Tq = ps.T[path], Lq = ps.L[path], pq = ps.prev[path], dq = ps.rd[path];
            const vec d = xyz(dq);
            vec Le = splat(0);
            for (int k = 0; k < sv.env_count; ++k) Le = Le + env_le(sv, sv.lights[sv.env_lights[k]], d);
            if (max_component(Le) > 0) {
                const uint32_t pbits = __float_as_uint(pq.w);
                const unsigned prev_flags = pbits & 0xffu;
                const int prev_measure = int((pbits >> 8) & 0xffu);
                const int depth = int(pbits >> 16);
                const float prev_pdf = Lq.w;
                vec L = xyz(Lq);
                if (depth == 0 || is_delta(prev_flags)) {
                    L = L + xyz(Tq) * Le;
                } else if (prev_measure == kSolidAngle && prev_pdf > 0) {
                    float lpdf = 0;  // infinite_lights_pdf_li: selection pmf (at the origin) x direction pdf
                    for (int k = 0; k < sv.env_count; ++k) {
                        const int li = sv.env_lights[k];
                        lpdf += light_choice_pmf<kLight>(sv, li, splat(0)) * env_pdf(sv, sv.lights[li], d);
                    }
                    L = L + xyz(Tq) * Le * power_heuristic(prev_pdf, lpdf);
                }
                ps.L[path] = pack(L, prev_pdf);
            }
        }
        uint32_t slot = wave_append(&qs.counts[in ^ 1], push_ray);
        if (push_ray) qs.q[in ^ 1][slot] = path;
        uint32_t sslot = wave_append(sh_count, push_shadow);
        if (push_shadow) {
            ps.so[sslot] = pack(sh_o, sh_tmax);
            ps.sd[sslot] = pack(sh_d, sh_mis);
            ps.sc[sslot] = pack(sh_c, __uint_as_float(path));
            ps.st[sslot] = pack(sh_T, __int_as_float(prim));
        }
    }
}

// ---- flush ---------------------------------------------------------------------------
// The first slot of each (segment, pixel) sums that pixel's samples of the
// segment in sample order onto the film; a request's sample 0 overwrites the
// film (0 + x == x for the non-negative radiance sums), later samples add.
template <int B>
__global__ __launch_bounds__(B) void k_flush(const BatchDesc* __restrict__ desc, Paths ps) {
    const uint32_t total = desc->count;
    for (uint32_t j = blockIdx.x * B + threadIdx.x; j < total; j += gridDim.x * B) {
        const Segment& sg = desc->seg[find_segment(desc, j)];
        const uint32_t local = j - sg.slot_begin;
        const uint64_t npix = uint64_t(sg.width) * uint64_t(sg.height);
        if (local >= npix) continue;
        const uint64_t r = sg.req_begin + local;
        const uint64_t pixel = r % npix;
        float* f = sg.film + 3 * pixel;
        float a = 0, b = 0, c = 0;
        if (r >= npix) {
            a = f[0];
            b = f[1];
            c = f[2];
        }
        for (uint64_t k = local; k < sg.count; k += npix) {
            float4 L = ps.L[sg.slot_begin + k];
            a += L.x;
            b += L.y;
            c += L.z;
        }
        f[0] = a;
        f[1] = b;
        f[2] = c;
    }
}

__global__ void k_zero_rays(unsigned long long* rays) {
    if (threadIdx.x < 2) rays[threadIdx.x] = 0;
}

template <typename T>
T* dalloc(size_t count) {
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T));
    if (e != hipSuccess) throw Error(LUX_ERR_OUT_OF_MEMORY, std::string("hipMalloc failed: ") + hipGetErrorString(e));
    return static_cast<T*>(p);
}

Paths paths_of(const PathBuf& b) {
    size_t P = size_t(b.capacity);
    Paths p;
    p.T = b.f4;
    p.L = b.f4 + P;
    p.prev = b.f4 + 2 * P;
    p.ro = b.f4 + 3 * P;
    p.rd = b.f4 + 4 * P;
    p.hit = b.f4 + 5 * P;
    p.so = b.f4 + 6 * P;
    p.sd = b.f4 + 7 * P;
    p.sc = b.f4 + 8 * P;
    p.st = b.f4 + 9 * P;
    p.rng = b.rng;
    return p;
}

Queues queues_of(const PathBuf& b) {
    Queues q;
    q.q[0] = b.q;
    q.q[1] = b.q + b.capacity;
    q.counts = b.counts;
    q.rays = b.rays;
    return q;
}

void pathbuf_release(PathBuf& b) {
    void* ptrs[] = {b.f4, b.rng, b.q, b.counts, b.rays, b.desc};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    b = PathBuf{};
}

Segment make_segment(const Request& rq, uint64_t req_begin, uint32_t slot_begin, uint32_t count) {
    Segment s{};
    s.sv = rq.sv;
    s.film = rq.film;
    s.seed = rq.rs.seed;
    s.req_begin = req_begin;
    s.slot_begin = slot_begin;
    s.count = count;
    s.width = rq.rs.width;
    s.height = rq.rs.height;
    s.max_depth = rq.rs.max_depth;
    s.frame = rq.rs.frame_index;
    s.sample_offset = rq.rs.sample_offset;
    return s;
}

}  // namespace

int request_kind(const Request& rq) { return (rq.sv.instanced ? 1 : 0) | (rq.rs.light_sampler << 1); }

uint64_t request_paths(const lux_render_settings& rs) {
    return uint64_t(rs.width) * uint64_t(rs.height) * uint64_t(rs.spp);
}

// ---- host: the batch engine ------------------------------------------------------------
Renderer::Renderer() = default;

Renderer::~Renderer() {
    release_graph();
    if (buf_.capacity) (void)hipDeviceSynchronize();
    pathbuf_release(buf_);
    for (int k = 0; k < kRing; ++k)
        if (ring_ev_[k]) (void)hipEventDestroy(ring_ev_[k]);
    if (ring_) (void)hipHostFree(ring_);
}

void Renderer::set_tuning(const Tuning& t) {
    if (t.capacity <= 0 || t.grid <= 0) throw Error(LUX_ERR_INVALID_ARGUMENT, "tuning needs capacity > 0 and grid > 0");
    if (t.block != 128 && t.block != 256 && t.block != 512)
        throw Error(LUX_ERR_INVALID_ARGUMENT, "tuning block must be 128, 256 or 512");
    tune_ = t;
}

void Renderer::release_graph() {
    if (graph_) (void)hipGraphExecDestroy(graph_);
    graph_ = nullptr;
    graph_depth_ = graph_grid_ = graph_capacity_ = graph_block_ = graph_opaque_ = graph_two_level_ = graph_light_ = -1;
}

void Renderer::ensure_capacity(int paths) {
    if (paths <= buf_.capacity) return;
    release_graph();  // the captured launches hold the old buffer addresses
    if (buf_.capacity) hip_check(hipDeviceSynchronize(), "synchronize before path buffer growth");
    pathbuf_release(buf_);
    size_t P = size_t(paths);
    buf_.capacity = paths;
    buf_.f4 = dalloc<float4>(11 * P);
    buf_.rng = dalloc<ulonglong2>(P);
    buf_.q = dalloc<uint32_t>(2 * P);
    buf_.counts = dalloc<uint32_t>(4);
    buf_.rays = dalloc<unsigned long long>(2);
    buf_.desc = static_cast<BatchDesc*>(static_cast<void*>(dalloc<uint8_t>(sizeof(BatchDesc))));
    hip_check(hipMemset(buf_.counts, 0, 16), "hipMemset counts");
    hip_check(hipMemset(buf_.rays, 0, 16), "hipMemset rays");
}

namespace {
size_t pathbuf_bytes(int capacity) {
    size_t P = size_t(capacity);
    return P ? P * (11 * 16 + 16 + 8) + 16 + 16 + sizeof(BatchDesc) : 0;
}
}  // namespace

size_t Renderer::device_bytes() const { return pathbuf_bytes(buf_.capacity); }

size_t Renderer::planned_bytes() const { return pathbuf_bytes(std::max(buf_.capacity, tune_.capacity)); }

BatchDesc* Renderer::next_slot() {
    if (!ring_) {
        void* p = nullptr;
        hip_check(hipHostMalloc(&p, sizeof(BatchDesc) * kRing, hipHostMallocDefault), "hipHostMalloc batch ring");
        ring_ = static_cast<BatchDesc*>(p);
        for (int k = 0; k < kRing; ++k)
            hip_check(hipEventCreateWithFlags(&ring_ev_[k], hipEventDisableTiming), "hipEventCreate");
    }
    int k = ring_next_;
    if (ring_used_[k]) hip_check(hipEventSynchronize(ring_ev_[k]), "wait for batch ring slot");
    return ring_ + k;
}

void Renderer::upload(BatchDesc* slot, hipStream_t stream) {
    int k = int(slot - ring_);
    hip_check(hipMemcpyAsync(buf_.desc, slot, desc_bytes(slot->nseg), hipMemcpyHostToDevice, stream), "H2D batch");
    hip_check(hipEventRecord(ring_ev_[k], stream), "hipEventRecord");
    ring_used_[k] = true;
    ring_next_ = (k + 1) % kRing;
}

void Renderer::reset_ray_counts(hipStream_t stream) {
    ensure_capacity(1);
    LUX_LAUNCH("pt_zero_rays", k_zero_rays, 1, 64, 0, stream, buf_.rays);
}

void Renderer::ray_counts(hipStream_t stream, uint64_t* extension, uint64_t* shadow) {
    unsigned long long r[2] = {0, 0};
    if (buf_.rays) {
        hip_check(hipMemcpyAsync(r, buf_.rays, sizeof(r), hipMemcpyDeviceToHost, stream), "D2H ray counts");
        hip_check(hipStreamSynchronize(stream), "synchronize");
    }
    *extension = r[0];
    *shadow = r[1];
}

namespace {
// Grid of the wavefront kernels: the tuned grid, and for k_shade at most the
// blocks the device holds at once. k_shade walks its queue in rounds of
// gridDim.x * B paths with block barriers, so blocks beyond that number run as
// a second, partial round behind the first (at block 512 it holds 128 VGPRs,
// 2 blocks per CU: 512 blocks on 256 CUs against the tuned 768 took C3 from
// 0.0730 to 0.0713 s and the 64-scene mix from 1.243 to 1.215 s, same films).
struct LaunchGrid {
    int grid;
    int shade;
};

template <int B, int kLight>
int shade_blocks_per_cu() {
    int per = 0;
    hip_check(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, reinterpret_cast<const void*>(k_shade<B, kLight>), B,
                                                           0),
              "hipOccupancyMaxActiveBlocksPerMultiprocessor");
    return per;
}

template <int kLight>
int shade_blocks_per_cu_b(int block) {
    switch (block) {
        case 128: return shade_blocks_per_cu<128, kLight>();
        case 512: return shade_blocks_per_cu<512, kLight>();
        default: return shade_blocks_per_cu<256, kLight>();
    }
}

int shade_resident_blocks(int block, int light) {
    int dev = 0, cus = 0;
    hip_check(hipGetDevice(&dev), "hipGetDevice");
    hip_check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev), "hipDeviceGetAttribute");
    int per = 0;
    switch (light) {
        case kLightUniform: per = shade_blocks_per_cu_b<kLightUniform>(block); break;
        case kLightBvh: per = shade_blocks_per_cu_b<kLightBvh>(block); break;
        default: per = shade_blocks_per_cu_b<kLightPower>(block); break;
    }
    return std::max(1, per * cus);
}

// The wavefront launch sequence of one batch (captured or launched directly).
// Raw launches (no probe events) when captured.
template <int B, bool kProbe, bool kOpaque, bool kTwoLevel, int kLight>
void wavefront_sequence(const PathBuf& buf, int depth, LaunchGrid grid, hipStream_t stream) {
    const Paths ps = paths_of(buf);
    const Queues qs = queues_of(buf);
    const BatchDesc* d = buf.desc;
#define PT_LAUNCH(NAME, G, K, ...)                                                   \
    do {                                                                             \
        if (kProbe) {                                                                \
            LUX_LAUNCH(NAME, K, G, B, 0, stream, __VA_ARGS__);                       \
        } else {                                                                     \
            hipLaunchKernelGGL(K, dim3(G), dim3(B), 0, stream, __VA_ARGS__);         \
            hip_check(hipGetLastError(), NAME);                                      \
        }                                                                            \
    } while (0)
    PT_LAUNCH("pt_generate", grid.grid, k_generate<B>, d, ps, qs);
    for (int b = 0; b <= depth; ++b) {
        const int in = b & 1;
        PT_LAUNCH("pt_trace", grid.grid, (k_trace<B, kOpaque, kTwoLevel>), d, ps, qs, in,
                  b > 0 ? 2 + ((b - 1) & 1) : -1, in ^ 1, 2 + (b & 1));
        PT_LAUNCH("pt_shade", grid.shade, (k_shade<B, kLight>), d, ps, qs, in, b, -1);
    }
    PT_LAUNCH("pt_trace", grid.grid, (k_trace<B, kOpaque, kTwoLevel>), d, ps, qs, -1, 2 + (depth & 1), -1, -1);
    PT_LAUNCH("pt_flush", grid.grid, k_flush<B>, d, ps);
#undef PT_LAUNCH
}
}  // namespace

namespace {
template <bool kProbe, bool kOpaque, bool kTwoLevel, int kLight>
void wavefront_sequence_b(int block, const PathBuf& buf, int depth, LaunchGrid grid, hipStream_t stream) {
    switch (block) {
        case 128: wavefront_sequence<128, kProbe, kOpaque, kTwoLevel, kLight>(buf, depth, grid, stream); break;
        case 512: wavefront_sequence<512, kProbe, kOpaque, kTwoLevel, kLight>(buf, depth, grid, stream); break;
        default: wavefront_sequence<256, kProbe, kOpaque, kTwoLevel, kLight>(buf, depth, grid, stream); break;
    }
}

template <bool kProbe, bool kOpaque, bool kTwoLevel>
void wavefront_sequence_l(int light, int block, const PathBuf& buf, int depth, LaunchGrid grid, hipStream_t stream) {
    switch (light) {
        case kLightUniform:
            wavefront_sequence_b<kProbe, kOpaque, kTwoLevel, kLightUniform>(block, buf, depth, grid, stream);
            break;
        case kLightBvh:
            wavefront_sequence_b<kProbe, kOpaque, kTwoLevel, kLightBvh>(block, buf, depth, grid, stream);
            break;
        default:
            wavefront_sequence_b<kProbe, kOpaque, kTwoLevel, kLightPower>(block, buf, depth, grid, stream);
            break;
    }
}

// The launch sequence for one batch kind: shadow-ray walk (any-hit when no
// scene of the batch passes light), traversal (one- or two-level) and light
// selection are template choices made here, per batch.
template <bool kProbe>
void wavefront_sequence_o(bool opaque, bool two_level, int light, int block, const PathBuf& buf, int depth,
                          LaunchGrid grid, hipStream_t stream) {
    if (two_level) {
        if (opaque)
            wavefront_sequence_l<kProbe, true, true>(light, block, buf, depth, grid, stream);
        else
            wavefront_sequence_l<kProbe, false, true>(light, block, buf, depth, grid, stream);
    } else {
        if (opaque)
            wavefront_sequence_l<kProbe, true, false>(light, block, buf, depth, grid, stream);
        else
            wavefront_sequence_l<kProbe, false, false>(light, block, buf, depth, grid, stream);
    }
}
}  // namespace

void Renderer::launch_wavefront(int depth, bool opaque, bool two_level, int light, hipStream_t stream) {
    const int block = tune_.block;
    const int tuned = std::max(1, std::min(tune_.grid, div_up(buf_.capacity, block)));
    if (shade_resident_block_ != block || shade_resident_light_ != light) {  // outside any capture
        shade_resident_ = shade_resident_blocks(block, light);
        shade_resident_block_ = block;
        shade_resident_light_ = light;
    }
    const LaunchGrid grid{tuned, std::min(tuned, shade_resident_)};
    if (!tune_.use_graph) {
        wavefront_sequence_o<true>(opaque, two_level, light, block, buf_, depth, grid, stream);
        return;
    }
    if (!graph_ || graph_depth_ != depth || graph_grid_ != tuned || graph_capacity_ != buf_.capacity ||
        graph_block_ != block || graph_opaque_ != int(opaque) || graph_two_level_ != int(two_level) ||
        graph_light_ != light) {
        release_graph();
        hip_check(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
        hipGraph_t g = nullptr;
        try {
            wavefront_sequence_o<false>(opaque, two_level, light, block, buf_, depth, grid, stream);
        } catch (...) {
            (void)hipStreamEndCapture(stream, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        hip_check(hipStreamEndCapture(stream, &g), "hipStreamEndCapture");
        hipError_t e = hipGraphInstantiate(&graph_, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        hip_check(e, "hipGraphInstantiate");
        graph_depth_ = depth;
        graph_grid_ = tuned;
        graph_capacity_ = buf_.capacity;
        graph_block_ = block;
        graph_opaque_ = int(opaque);
        graph_two_level_ = int(two_level);
        graph_light_ = light;
    }
    hip_check(hipGraphLaunch(graph_, stream), "hipGraphLaunch");
}

namespace {
template <bool kTwoLevel, int kLight>
void plain_sequence(const PathBuf& buf, int depth, uint32_t count, hipStream_t stream) {
    const Paths ps = paths_of(buf);
    const Queues qs = queues_of(buf);
    const BatchDesc* d = buf.desc;
    const int grid = std::max(1, div_up(count, kBlock));
    LUX_LAUNCH("pt_generate", k_generate<kBlock>, grid, kBlock, 0, stream, d, ps, qs);
    for (int b = 0; b <= depth; ++b) {
        const int in = b & 1;
        LUX_LAUNCH("pt_trace", (k_trace<kBlock, false, kTwoLevel>), grid, kBlock, 0, stream, d, ps, qs, in, -1,
                   in ^ 1, 2 + (b & 1));
        for (int type = 0; type < kMaterialTypes; ++type)
            LUX_LAUNCH("pt_shade", (k_shade<kBlock, kLight>), grid, kBlock, 0, stream, d, ps, qs, in, b, type);
        LUX_LAUNCH("pt_shadow", (k_trace<kBlock, false, kTwoLevel>), grid, kBlock, 0, stream, d, ps, qs, -1,
                   2 + (b & 1), -1, -1);
    }
    LUX_LAUNCH("pt_flush", k_flush<kBlock>, grid, kBlock, 0, stream, d, ps);
}
}  // namespace

namespace {
template <bool kTwoLevel>
void plain_sequence_l(int light, const PathBuf& buf, int depth, uint32_t count, hipStream_t stream) {
    switch (light) {
        case kLightUniform: plain_sequence<kTwoLevel, kLightUniform>(buf, depth, count, stream); break;
        case kLightBvh: plain_sequence<kTwoLevel, kLightBvh>(buf, depth, count, stream); break;
        default: plain_sequence<kTwoLevel, kLightPower>(buf, depth, count, stream); break;
    }
}
}  // namespace

void Renderer::launch_plain(int depth, uint32_t count, bool two_level, int light, hipStream_t stream) {
    if (two_level)
        plain_sequence_l<true>(light, buf_, depth, count, stream);
    else
        plain_sequence_l<false>(light, buf_, depth, count, stream);
}

uint64_t Renderer::enqueue_batch(std::vector<Request*>& active, int slice_spp, int max_depth, hipStream_t stream,
                                 int64_t* batch_index) {
    const uint64_t cap = uint64_t(tune_.capacity);
    ensure_capacity(tune_.capacity);
    BatchDesc* slot = next_slot();
    uint64_t used = 0;
    uint32_t nseg = 0;
    bool opaque = true;  // every scene of the batch: shadow rays may use the any-hit walk
    // One kind per batch (traversal level and light sampler: the kernels are
    // built for one): the kind of the first request with paths left; requests
    // of another kind wait for a later batch and keep their place in `active`.
    int kind = -1;
    for (Request* rq : active) {
        if (used >= cap || nseg >= uint32_t(kMaxSegments)) break;
        if (rq->issued >= rq->total) continue;
        const int k = request_kind(*rq);
        if (kind < 0) kind = k;
        if (k != kind) continue;
        uint64_t take = std::min(rq->total - rq->issued, cap - used);
        if (slice_spp > 0)
            take = std::min(take, uint64_t(slice_spp) * uint64_t(rq->rs.width) * uint64_t(rq->rs.height));
        opaque = opaque && rq->sv.opaque != 0;
        slot->seg[nseg] = make_segment(*rq, rq->issued, uint32_t(used), uint32_t(take));
        slot->seg[nseg++].sv.preorder = tune_.preorder_walk;
        rq->issued += take;
        used += take;
        if (rq->issued == rq->total) rq->last_batch = batch_counter_;
    }
    if (batch_index) *batch_index = batch_counter_;
    if (used == 0) return 0;
    slot->count = uint32_t(used);
    slot->nseg = nseg;
    upload(slot, stream);
    launch_wavefront(max_depth, opaque, (kind & 1) != 0, kind >> 1, stream);
    ++batch_counter_;
    return used;
}

void Renderer::render_all(std::vector<Request*>& reqs, int mode, hipStream_t stream, RenderStats* stats) {
    for (Request* rq : reqs) {
        rq->total = request_paths(rq->rs);
        rq->issued = 0;
        rq->last_batch = -1;
    }
    if (mode == LUX_RENDER_PLAIN) {
        // The reference's loop: for each sample, 65 536-pixel batches, each a
        // complete bounce loop followed by a synchronisation.
        for (Request* rq : reqs) {
            const uint64_t npix = uint64_t(rq->rs.width) * uint64_t(rq->rs.height);
            const uint64_t B = uint64_t(rq->rs.batch_size > 0 ? rq->rs.batch_size : kPlainBatch);
            ensure_capacity(int(std::min(B, npix)));
            uint32_t host_counts[4];
            for (int s = 0; s < rq->rs.spp; ++s) {
                for (uint64_t p0 = 0; p0 < npix; p0 += B) {
                    const uint32_t n = uint32_t(std::min(B, npix - p0));
                    BatchDesc* slot = next_slot();
                    slot->count = n;
                    slot->nseg = 1;
                    slot->seg[0] = make_segment(*rq, uint64_t(s) * npix + p0, 0, n);
                    slot->seg[0].sv.preorder = tune_.preorder_walk;
                    upload(slot, stream);
                    launch_plain(rq->rs.max_depth, n, rq->sv.instanced != 0, rq->rs.light_sampler, stream);
                    hip_check(hipMemcpyAsync(host_counts, buf_.counts, sizeof(host_counts), hipMemcpyDeviceToHost,
                                             stream),
                              "D2H counts");
                    hip_check(hipStreamSynchronize(stream), "batch synchronize");
                    ++batch_counter_;
                    if (stats) {
                        stats->batches += 1;
                        stats->paths += n;
                    }
                }
            }
            rq->issued = rq->total;
            rq->last_batch = batch_counter_ - 1;
        }
        return;
    }
    // wavefront: the caller's batch size (if any) caps this call's batches
    Tuning saved = tune_;
    uint64_t sum = 0;
    int depth = 0;
    int forced = 0;
    for (Request* rq : reqs) {
        sum += rq->total;
        depth = std::max(depth, rq->rs.max_depth);
        if (rq->rs.batch_size > 0) forced = forced ? std::min(forced, rq->rs.batch_size) : rq->rs.batch_size;
    }
    if (sum == 0) return;
    Tuning t = tune_;
    if (forced) t.capacity = forced;
    t.capacity = int(std::min<uint64_t>(uint64_t(t.capacity), sum));
    t.use_graph = mode == LUX_RENDER_WAVEFRONT ? tune_.use_graph : 0;
    tune_ = t;
    try {
        for (;;) {
            uint64_t n = enqueue_batch(reqs, 0, depth, stream, nullptr);
            if (!n) break;
            if (stats) {
                stats->batches += 1;
                stats->paths += n;
            }
        }
    } catch (...) {
        tune_ = saved;
        throw;
    }
    tune_ = saved;
}

}  // namespace pt
}  // namespace lux

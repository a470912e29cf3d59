// pt.hip -- wavefront path tracer for gfx950.
//
// Stages per bounce (src/integrator/wavefront_scheduler.cu:73-131 is the
// reference's schedule; the transport rules are those of trace_path,
// src/integrator/cpu_path_tracer.cpp:403-498):
//   intersect  closest hit of every live path (right-first stackless LBVH walk)
//   shade      emission + MIS, pass-through, termination, next-event estimation
//              (power-alias light choice, area sample -> shadow queue), BSDF
//              sample, Russian roulette -> next ray queue
//   shadow     shadow rays through pass-through surfaces; adds the NEE term
// then one flush per batch adds each path's radiance to its pixel, samples of a
// pixel in sample order (so the film sum is the same float sum as the CPU's).
//
// Queues are path-index lists appended with one atomic per wave (ballot +
// popcount). Two schedules share the kernels:
//   plain      the reference's: 65 536-path batches, one sample per pixel per
//              batch, full-grid launches, one shade launch per material type,
//              a host synchronisation + counter read-back after every batch
//   wavefront  large multi-sample batches, grid-stride kernels sized to the
//              chip, one shade launch per bounce, no host round trip until the
//              render ends
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "pt_device.hpp"
#include "pt_runtime.hpp"

namespace lux {
namespace pt {
namespace {

constexpr int kBlock = 256;

struct Queues {
    uint32_t* q[2];
    uint32_t* counts;  // [0], [1] ray queues, [2] shadow queue
};

// Path state, one slot per path of the batch (float4 records: one 16-byte load).
struct Paths {
    float4* T;      // throughput xyz, eta scale
    float4* L;      // radiance xyz, previous BSDF pdf
    float4* prev;   // previous vertex xyz, bits: flags | measure << 8 | depth << 16
    float4* ro;     // ray origin xyz, bits: pixel
    float4* rd;     // ray direction xyz, bits: sample index
    ulonglong2* rng;
    float4* hit;    // bits: prim, t, u, v
    // shadow records, one per path at most per bounce
    float4* so;     // origin xyz, t_max
    float4* sd;     // direction xyz, MIS weight
    float4* sc;     // unoccluded contribution xyz, bits: path
    float4* st;     // throughput at the vertex xyz, bits: ignored primitive
};

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

// Closest hit / occlusion walk over the right-first stackless LBVH; `ignore`
// skips one leaf's triangle test (the source primitive of a shadow ray).
__device__ __forceinline__ int closest(const SceneView& sv, const vec& o, const vec& d, float t_min, float t_max,
                                       int ignore, float& t_hit, float& u_hit, float& v_hit) {
    f3 O = make3(o.x, o.y, o.z), D = make3(d.x, d.y, d.z);
    SlabRay sr = make_slab_ray(O, D, t_min);
    int best = -1;
    int k = 0;
    while (uint32_t(k) < uint32_t(sv.node_count)) {
        float4 a, b;
        load_node(sv.nodes, k, a, b);
        int prim = __float_as_int(b.z);
        int esc = __float_as_int(b.w);
        int next;
        if (!slab_hit(sr, t_max, a.x, a.y, a.z, a.w, b.x, b.y)) {
            next = esc;
        } else if (prim >= 0) {
            next = esc;
            if (prim != ignore) {
                if (uint32_t(prim) >= uint32_t(sv.tri_count)) {
                    set_error(sv.info, kErrPrimRange);
                } else {
                    const float* p = reinterpret_cast<const float*>(sv.soup + prim);
                    f3 v0 = make3(p[0], p[1], p[2]), v1 = make3(p[3], p[4], p[5]), v2 = make3(p[6], p[7], p[8]);
                    float t, u, v;
                    if (tri_hit(O, D, t_min, t_max, v0, v1, v2, t, u, v)) {
                        best = prim;
                        t_max = t;
                        u_hit = u;
                        v_hit = v;
                    }
                }
            }
        } else {
            next = k + 1;
        }
        if (next >= 0 && next <= k) {
            set_error(sv.info, kErrWalkOrder);
            break;
        }
        k = next;
    }
    t_hit = t_max;
    return best;
}

__device__ __forceinline__ vec xyz(const float4& f) { return mk(f.x, f.y, f.z); }
__device__ __forceinline__ float4 pack(const vec& v, float w) { return make_float4(v.x, v.y, v.z, w); }

// ---- generate ------------------------------------------------------------------
// Batch path j: local pixel j % npix (pixel = pix0 + local), sample s0 + j / npix.
// Camera sample dimensions as render_cpu_path_tracer (src/integrator/cpu_path_tracer.cpp:532-542).
__global__ __launch_bounds__(kBlock) void k_generate(SceneView sv, Paths ps, Queues qs, int width, int height,
                                                     int pix0, int npix, int s0, int nsub, uint64_t seed, int frame,
                                                     int sample_offset) {
    const int total = npix * nsub;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        qs.counts[0] = uint32_t(total);
        qs.counts[2] = 0;
    }
    for (int j = blockIdx.x * kBlock + threadIdx.x; j < total; j += gridDim.x * kBlock) {
        int local = j % npix;
        int sub = j / npix;
        int pixel = pix0 + local;
        int sample = s0 + sub;
        Rng rng = make_rng(seed, frame, pixel, sample + sample_offset);
        int x = pixel % width, y = pixel / width;
        float jx = rng.get1();
        float jy = rng.get1();
        float u = (float(x) + jx) / float(width);
        float v = (float(height - 1 - y) + jy) / float(height);
        rng.get1();  // lens u
        rng.get1();  // lens v
        rng.get1();  // time
        const DevCamera& c = sv.cam;
        vec target = mk(c.lx, c.ly, c.lz) + u * mk(c.hx, c.hy, c.hz) + v * mk(c.vx, c.vy, c.vz);
        vec origin = mk(c.ox, c.oy, c.oz);
        vec dir = target - origin;
        float len = length(dir);
        dir = len > 0 ? dir / len : dir;
        ps.T[j] = make_float4(1, 1, 1, 1);
        ps.L[j] = make_float4(0, 0, 0, 0);
        ps.prev[j] = make_float4(0, 0, 0, __int_as_float(0));
        ps.ro[j] = pack(origin, __int_as_float(pixel));
        ps.rd[j] = pack(dir, __int_as_float(sample));
        ps.rng[j] = make_ulonglong2(rng.state, rng.inc);
        qs.q[0][j] = uint32_t(j);
    }
}

// ---- intersect -----------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_intersect(SceneView sv, Paths ps, Queues qs, int in) {
    const uint32_t count = qs.counts[in];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        qs.counts[in ^ 1] = 0;  // next-bounce queue; the previous bounce no longer reads it
        qs.counts[2] = 0;       // shadow queue of this bounce
    }
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < count; i += gridDim.x * kBlock) {
        uint32_t path = qs.q[in][i];
        float4 o = ps.ro[path], d = ps.rd[path];
        float t = INFINITY, u = 0, v = 0;
        int prim = closest(sv, xyz(o), xyz(d), kEps, INFINITY, -1, t, u, v);
        ps.hit[path] = make_float4(__int_as_float(prim), t, u, v);
    }
}

// ---- shade -----------------------------------------------------------------------
// One bounce of trace_path for every live path. `filter` >= 0 restricts the
// launch to hits of that material type (misses go with filter 0).
__global__ __launch_bounds__(kBlock) void k_shade(SceneView sv, Paths ps, Queues qs, int in, int bounce,
                                                  int max_depth, int filter) {
    const uint32_t count = qs.counts[in];
    const uint32_t stride = gridDim.x * kBlock;
    const uint32_t rounds = (count + stride - 1) / stride;
    for (uint32_t r = 0; r < rounds; ++r) {  // uniform trip count: ballots see the whole wave
        uint32_t i = r * stride + blockIdx.x * kBlock + threadIdx.x;
        bool live = i < count;
        uint32_t path = live ? qs.q[in][i] : 0u;
        bool push_ray = false, push_shadow = false;
        float4 hit = live ? ps.hit[path] : make_float4(__int_as_float(-1), 0, 0, 0);
        int prim = __float_as_int(hit.x);
        const DevMaterial* mp = nullptr;
        if (live && prim >= 0) mp = &sv.mats[sv.mat_id[prim]];
        if (live && filter >= 0) live = (prim < 0) ? filter == 0 : mp->type == filter;
        // NEE record and next ray, filled below
        vec sh_o = splat(0), sh_d = splat(0), sh_c = splat(0), sh_T = splat(0);
        float sh_tmax = 0, sh_mis = 0;
        if (live && prim >= 0) {
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
                    lpdf = lt.pmf * pdf_solid_angle(p, ng, lt.two_sided, 1.0f / lt.area, prev_pos);
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
                    rng.get1();  // light component (unused by area lights)
                    float lu0 = rng.get1();
                    float lu1 = rng.get1();
                    // sample_light_li: power alias pick, area sample (src/scene/scene_light_sampler.h:210-250)
                    int nl = sv.light_count;
                    if (nl > 0) {
                        float scaled = fminf(sel, 0.99999994f) * float(nl);
                        int col = int(scaled);
                        if (col >= nl) col = nl - 1;
                        float thr = scaled - float(col);
                        int idx = thr < sv.lights[col].prob ? col : sv.lights[col].alias;
                        if (idx >= 0 && idx < nl && sv.lights[idx].pmf > 0) {
                            const DevLight& lt = sv.lights[idx];
                            const lux_triangle lt_tri = sv.soup[lt.prim];
                            float su = sqrtf(lu0);
                            float b0 = 1.0f - su, b1 = su * (1.0f - lu1), b2 = su * lu1;
                            vec lp = b0 * tri_vertex(lt_tri, 0) + b1 * tri_vertex(lt_tri, 1) + b2 * tri_vertex(lt_tri, 2);
                            vec ln = tri_normal(lt_tri);
                            float pdf_area = lt.area > 0 ? 1.0f / lt.area : 0.0f;
                            vec to = lp - p;
                            float d2 = length2(to);
                            if (d2 > kEps * kEps) {
                                float dist = sqrtf(d2);
                                vec wi = to / dist;
                                vec rad = eval_emission(mk(lt.er, lt.eg, lt.eb), lt.two_sided, ln, -wi);
                                float lpdf = lt.pmf * pdf_solid_angle(lp, ln, lt.two_sided, pdf_area, p);
                                if (max_component(rad) > 0 && lpdf > 0) {
                                    // estimate_unoccluded_direct_light_li (src/light/light_sampler.h:420-454)
                                    vec f = eval_bsdf(m, it, wi);
                                    float cs = abs_cos(ns, wi);
                                    float pdf_b = pdf_bsdf(m, it, wi);
                                    if (max_component(f) > 0 && cs > 0) {
                                        vec off = dot(wi, ng) >= 0 ? ng : -ng;
                                        vec so = p + off * kEps;
                                        vec tl = lp - so;
                                        float sdist = length(tl);
                                        if (sdist > kEps) {
                                            sh_o = so;
                                            sh_d = tl / sdist;
                                            sh_tmax = sdist - kEps;
                                            sh_c = f * rad * (cs / lpdf);
                                            sh_mis = power_heuristic(lpdf, pdf_b);
                                            sh_T = T;
                                            push_shadow = true;
                                        }
                                    }
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
        uint32_t slot = wave_append(&qs.counts[in ^ 1], push_ray);
        if (push_ray) qs.q[in ^ 1][slot] = path;
        uint32_t sslot = wave_append(&qs.counts[2], push_shadow);
        if (push_shadow) {
            ps.so[sslot] = pack(sh_o, sh_tmax);
            ps.sd[sslot] = pack(sh_d, sh_mis);
            ps.sc[sslot] = pack(sh_c, __uint_as_float(path));
            ps.st[sslot] = pack(sh_T, __int_as_float(prim));
        }
    }
}

// ---- shadow ------------------------------------------------------------------------
// trace_shadow (src/integrator/cpu_path_tracer.cpp:27-72): repeated closest hits
// through pass-through surfaces; radiance += T * (contribution * Tr * mis).
__global__ __launch_bounds__(kBlock) void k_shadow(SceneView sv, Paths ps, Queues qs) {
    const uint32_t count = qs.counts[2];
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < count; i += gridDim.x * kBlock) {
        float4 so = ps.so[i], sd = ps.sd[i], sc = ps.sc[i], st = ps.st[i];
        vec o = xyz(so), d = xyz(sd);
        float t_max = so.w, t_min = kEps;
        int ignore = __float_as_int(st.w);
        vec tr = splat(1);
        for (int guard = 0; guard < 1 << 16; ++guard) {
            float t = 0, u = 0, v = 0;
            int prim = closest(sv, o, d, t_min, t_max, ignore, t, u, v);
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
            uint32_t path = __float_as_uint(sc.w);
            float4 Lq = ps.L[path];
            vec add = xyz(st) * (xyz(sc) * tr * sd.w);
            ps.L[path] = pack(xyz(Lq) + add, Lq.w);
        }
    }
}

// ---- flush ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_flush(Paths ps, float* film, int pix0, int npix, int nsub) {
    for (int local = blockIdx.x * kBlock + threadIdx.x; local < npix; local += gridDim.x * kBlock) {
        float* f = film + 3 * size_t(pix0 + local);
        float r = f[0], g = f[1], b = f[2];
        for (int sub = 0; sub < nsub; ++sub) {
            float4 L = ps.L[size_t(sub) * npix + local];
            r += L.x;
            g += L.y;
            b += L.z;
        }
        f[0] = r;
        f[1] = g;
        f[2] = b;
    }
}

__global__ void k_zero_film(float* film, size_t n) {
    for (size_t i = size_t(blockIdx.x) * kBlock + threadIdx.x; i < n; i += size_t(gridDim.x) * kBlock) film[i] = 0;
}

template <typename T>
T* dalloc(size_t count) {
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T));
    if (e != hipSuccess) throw Error(LUX_ERR_OUT_OF_MEMORY, std::string("hipMalloc failed: ") + hipGetErrorString(e));
    return static_cast<T*>(p);
}

}  // namespace

// ---- host: scene preparation ------------------------------------------------------
namespace host {

struct V {
    float x, y, z;
};
inline V mkv(float x, float y, float z) { return V{x, y, z}; }
inline V vsub(V a, V b) { return mkv(a.x - b.x, a.y - b.y, a.z - b.z); }
inline V vadd(V a, V b) { return mkv(a.x + b.x, a.y + b.y, a.z + b.z); }
inline V vmul(V a, float s) { return mkv(a.x * s, a.y * s, a.z * s); }
inline float vdot(V a, V b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V vcross(V a, V b) { return mkv(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
inline V vnorm(V v) {
    float len = sqrtf(vdot(v, v));
    if (!(len > 0)) return v;
    float inv = 1 / len;
    return vmul(v, inv);
}

// make_perspective_camera with lens radius 0 and focal distance 1
// (src/camera/camera.h:58-109, src/camera/pinhole_camera.h:13-17).
DevCamera make_camera(const float* c) {
    V origin = mkv(c[0], c[1], c[2]), look = mkv(c[3], c[4], c[5]);
    float vfov = c[6], aspect = c[7];
    V forward = vnorm(vsub(look, origin));
    V right = vnorm(vcross(mkv(0, 1, 0), forward));
    if (vdot(right, right) == 0) right = vnorm(vcross(mkv(0, 0, 1), forward));
    V up = vcross(forward, right);
    float theta = vfov * kPi / 180.0f;
    float vh = 2.0f * tanf(theta * 0.5f);
    float vw = aspect * vh;
    float focus = 1.0f;
    V horizontal = vmul(right, vw * focus);
    V vertical = vmul(up, vh * focus);
    V lower = vsub(vsub(vadd(origin, vmul(forward, focus)), vmul(horizontal, 0.5f)), vmul(vertical, 0.5f));
    return DevCamera{origin.x, origin.y, origin.z, lower.x, lower.y, lower.z,
                     horizontal.x, horizontal.y, horizontal.z, vertical.x, vertical.y, vertical.z};
}

// transform_instance_normal under the identity instance: adjugate product then
// two normalisations (src/scene/scene_data.h:262-268, src/core/mat4.cuh:209-229).
V instance_normal(V n) {
    if (vdot(n, n) == 0) return mkv(0, 0, 0);
    V t = mkv(1.0f * n.x + 0.0f * n.y + 0.0f * n.z, 0.0f * n.x + 1.0f * n.y + 0.0f * n.z,
              0.0f * n.x + 0.0f * n.y + 1.0f * n.z);
    V once = vnorm(t);
    if (vdot(once, once) == 0) return mkv(0, 0, 0);
    return vnorm(once);
}

// Power light selection with its alias table: prepare_scene_light_sampling
// (src/scene/scene_prepare.h:268-387) and build_alias_table_from_weights
// (src/light/light_distribution.cpp:8-64; same stack order, so the same aliases).
void build_lights(const std::vector<lux_triangle>& soup, const int32_t* mat_ids, const lux_material* mats,
                  std::vector<DevLight>& lights, std::vector<int32_t>& light_of_prim) {
    int n = int(soup.size());
    light_of_prim.assign(n, -1);
    std::vector<float> cdf;
    float total = 0;
    for (int t = 0; t < n; ++t) {
        const lux_material& m = mats[mat_ids[t]];
        float emax = fmaxf(m.emission.x, fmaxf(m.emission.y, m.emission.z));
        if (!(emax > 0)) continue;
        V v0 = mkv(soup[t].v0.x, soup[t].v0.y, soup[t].v0.z);
        V v1 = mkv(soup[t].v1.x, soup[t].v1.y, soup[t].v1.z);
        V v2 = mkv(soup[t].v2.x, soup[t].v2.y, soup[t].v2.z);
        V c = vcross(vsub(v1, v0), vsub(v2, v0));
        float area = 0.5f * sqrtf(vdot(c, c));
        if (area <= 0) continue;
        float lum = float(0.2126) * m.emission.x + float(0.7152) * m.emission.y + float(0.0722) * m.emission.z;
        float power = area * fmaxf(1e-6f, lum);
        if (power <= 0) continue;
        total += power;
        DevLight L{};
        L.prim = t;
        L.er = m.emission.x;
        L.eg = m.emission.y;
        L.eb = m.emission.z;
        L.two_sided = m.emission_two_sided;
        L.area = area;
        lights.push_back(L);
        cdf.push_back(total);
        int32_t& stored = light_of_prim[t];
        stored = stored == -1 ? int32_t(lights.size()) - 1 : -2;
    }
    int count = int(lights.size());
    if (count == 0 || !(total > 0)) {
        lights.clear();
        return;
    }
    std::vector<float> w(count);
    float prev = 0;
    for (int i = 0; i < count; ++i) {
        w[i] = cdf[i] - prev;
        prev = cdf[i];
    }
    float sum = 0;
    for (float x : w) sum += fmaxf(0.0f, x);
    std::vector<float> scaled(count);
    std::vector<int> small, large;
    for (int i = 0; i < count; ++i) {
        float pmf = fmaxf(0.0f, w[i]) / sum;
        lights[i].pmf = pmf;
        lights[i].alias = i;
        lights[i].prob = 1;
        scaled[i] = pmf * float(count);
        (scaled[i] < 1.0f ? small : large).push_back(i);
    }
    while (!small.empty() && !large.empty()) {
        int l = small.back();
        small.pop_back();
        int g = large.back();
        large.pop_back();
        lights[l].prob = scaled[l];
        lights[l].alias = g;
        scaled[g] = (scaled[g] + scaled[l]) - 1.0f;
        (scaled[g] < 1.0f ? small : large).push_back(g);
    }
    for (int i : large) {
        lights[i].prob = 1;
        lights[i].alias = i;
    }
    for (int i : small) {
        lights[i].prob = 1;
        lights[i].alias = i;
    }
}

}  // namespace host

void scene_upload(SceneDev& s, const lux_scene_desc& desc, hipStream_t stream) {
    int nt = desc.triangle_count, nv = desc.vertex_count, nm = desc.material_count;
    if (nt <= 0 || nv <= 0 || nm <= 0 || !desc.vertices || !desc.indices || !desc.material_ids || !desc.materials)
        throw Error(LUX_ERR_INVALID_ARGUMENT, "scene needs vertices, indices, material ids and materials");
    for (int t = 0; t < nt; ++t) {
        for (int k = 0; k < 3; ++k)
            if (desc.indices[3 * size_t(t) + k] >= uint32_t(nv))
                throw Error(LUX_ERR_INVALID_ARGUMENT, "triangle index out of range");
        if (desc.material_ids[t] < 0 || desc.material_ids[t] >= nm)
            throw Error(LUX_ERR_INVALID_ARGUMENT, "material id out of range");
    }
    std::vector<lux_triangle> soup(nt);
    std::vector<lux_aabb> boxes(nt);
    std::vector<int32_t> ids(nt);
    for (int t = 0; t < nt; ++t) {
        const uint32_t* ix = desc.indices + 3 * size_t(t);
        soup[t] = lux_triangle{desc.vertices[ix[0]], desc.vertices[ix[1]], desc.vertices[ix[2]]};
        lux_aabb b;
        b.p_min = lux_vec3{INFINITY, INFINITY, INFINITY};
        b.p_max = lux_vec3{-INFINITY, -INFINITY, -INFINITY};
        for (const lux_vec3& p : {soup[t].v0, soup[t].v1, soup[t].v2}) {
            b.p_min = lux_vec3{fminf(b.p_min.x, p.x), fminf(b.p_min.y, p.y), fminf(b.p_min.z, p.z)};
            b.p_max = lux_vec3{fmaxf(b.p_max.x, p.x), fmaxf(b.p_max.y, p.y), fmaxf(b.p_max.z, p.z)};
        }
        boxes[t] = b;
        ids[t] = t;
    }
    std::vector<float> normals;
    if (desc.normals) {
        normals.resize(3 * size_t(nv));
        for (int v = 0; v < nv; ++v) {
            host::V n = host::instance_normal(host::mkv(desc.normals[v].x, desc.normals[v].y, desc.normals[v].z));
            normals[3 * v] = n.x;
            normals[3 * v + 1] = n.y;
            normals[3 * v + 2] = n.z;
        }
    }
    std::vector<DevMaterial> mats(nm);
    for (int m = 0; m < nm; ++m) {
        const lux_material& in = desc.materials[m];
        if (in.type < 0 || in.type >= kMaterialTypes) throw Error(LUX_ERR_INVALID_ARGUMENT, "unknown material type");
        mats[m] = DevMaterial{in.type, in.albedo.x, in.albedo.y, in.albedo.z, in.emission.x, in.emission.y,
                              in.emission.z, in.emission_two_sided, in.roughness, in.metallic, in.eta, 0};
    }
    std::vector<DevLight> lights;
    std::vector<int32_t> light_of_prim;
    host::build_lights(soup, desc.material_ids, desc.materials, lights, light_of_prim);

    scene_release(s);
    s.tri_count = nt;
    s.vertex_count = nv;
    s.soup = dalloc<lux_triangle>(nt);
    s.indices = dalloc<uint32_t>(3 * size_t(nt));
    s.mat_id = dalloc<int32_t>(nt);
    s.mats = dalloc<DevMaterial>(nm);
    s.mat_count = nm;
    s.light_of_prim = dalloc<int32_t>(nt);
    s.lights = dalloc<DevLight>(std::max<size_t>(lights.size(), 1));
    s.light_count = int(lights.size());
    s.host_lights = lights;
    s.host_light_of_prim = light_of_prim;
    if (desc.normals) s.normals = dalloc<float>(3 * size_t(nv));
    s.camera = host::make_camera(desc.camera);
    hip_check(hipMemcpyAsync(s.soup, soup.data(), sizeof(lux_triangle) * nt, hipMemcpyHostToDevice, stream), "H2D soup");
    hip_check(hipMemcpyAsync(s.indices, desc.indices, 12 * size_t(nt), hipMemcpyHostToDevice, stream), "H2D indices");
    hip_check(hipMemcpyAsync(s.mat_id, desc.material_ids, 4 * size_t(nt), hipMemcpyHostToDevice, stream), "H2D mat ids");
    hip_check(hipMemcpyAsync(s.mats, mats.data(), sizeof(DevMaterial) * nm, hipMemcpyHostToDevice, stream), "H2D mats");
    hip_check(hipMemcpyAsync(s.light_of_prim, light_of_prim.data(), 4 * size_t(nt), hipMemcpyHostToDevice, stream),
              "H2D light map");
    if (!lights.empty())
        hip_check(hipMemcpyAsync(s.lights, lights.data(), sizeof(DevLight) * lights.size(), hipMemcpyHostToDevice,
                                 stream),
                  "H2D lights");
    if (desc.normals)
        hip_check(hipMemcpyAsync(s.normals, normals.data(), 4 * normals.size(), hipMemcpyHostToDevice, stream),
                  "H2D normals");
    // BLAS over the triangle boxes (ids = triangle index)
    lux_aabb* d_boxes = dalloc<lux_aabb>(nt);
    int32_t* d_ids = dalloc<int32_t>(nt);
    hip_check(hipMemcpyAsync(d_boxes, boxes.data(), sizeof(lux_aabb) * nt, hipMemcpyHostToDevice, stream), "H2D boxes");
    hip_check(hipMemcpyAsync(d_ids, ids.data(), 4 * size_t(nt), hipMemcpyHostToDevice, stream), "H2D ids");
    lbvh_build(s.bvh, d_boxes, d_ids, nt, stream);
    hip_check(hipStreamSynchronize(stream), "scene upload");
    (void)hipFree(d_boxes);
    (void)hipFree(d_ids);
}

void scene_release(SceneDev& s) {
    void* ptrs[] = {s.soup, s.indices, s.normals, s.mat_id, s.mats, s.light_of_prim, s.lights};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    bvh_release(s.bvh);
    SceneDev empty;
    s = empty;
}

SceneView scene_view(const SceneDev& s) {
    SceneView v{};
    v.nodes = s.bvh.nodes;
    v.node_count = s.bvh.n > 0 ? 2 * s.bvh.n - 1 : 0;
    v.soup = s.soup;
    v.tri_count = s.tri_count;
    v.indices = s.indices;
    v.normals = s.normals;
    v.mat_id = s.mat_id;
    v.mats = s.mats;
    v.mat_count = s.mat_count;
    v.lights = s.lights;
    v.light_count = s.light_count;
    v.light_of_prim = s.light_of_prim;
    v.cam = s.camera;
    v.info = s.bvh.info;
    return v;
}

// ---- host: render -------------------------------------------------------------------
void pathbuf_allocate(PathBuf& b, int capacity) {
    if (capacity <= b.capacity) return;
    pathbuf_release(b);
    b.capacity = capacity;
    size_t P = size_t(capacity);
    b.f4 = dalloc<float4>(11 * P);
    b.rng = dalloc<ulonglong2>(P);
    b.q = dalloc<uint32_t>(2 * P);
    b.counts = dalloc<uint32_t>(4);
    hip_check(hipMemset(b.counts, 0, 16), "hipMemset counts");
}

void pathbuf_release(PathBuf& b) {
    void* ptrs[] = {b.f4, b.rng, b.q, b.counts};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    PathBuf empty;
    b = empty;
}

namespace {
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
    return q;
}
}  // namespace

void render(const SceneDev& s, PathBuf& buf, const lux_render_settings& rs, float* film, hipStream_t stream,
            RenderStats* stats) {
    const int w = rs.width, h = rs.height, spp = rs.spp, depth = rs.max_depth;
    const int npix_total = w * h;
    const bool plain = rs.mode == LUX_RENDER_PLAIN;
    SceneView sv = scene_view(s);
    LUX_LAUNCH("pt_zero_film", k_zero_film, div_up(3 * int64_t(npix_total), kBlock) > 4096 ? 4096
               : div_up(3 * int64_t(npix_total), kBlock), kBlock, 0, stream, film, size_t(3) * npix_total);
    // batch geometry: plain = one sample of up to batch_size pixels; wavefront =
    // as many samples of the whole image as fit the path budget
    int budget = rs.batch_size > 0 ? rs.batch_size : (plain ? 65536 : (1 << 21));
    int npix = std::min(npix_total, budget);
    int nsub = plain ? 1 : std::max(1, std::min(spp, budget / std::max(npix, 1)));
    pathbuf_allocate(buf, npix * nsub);
    Paths ps = paths_of(buf);
    Queues qs = queues_of(buf);
    int fixed_grid = std::min(div_up(int64_t(npix) * nsub, kBlock), 2048);
    std::vector<uint32_t> host_counts(4);
    for (int s0 = 0; s0 < spp; s0 += nsub) {
        int sub = std::min(nsub, spp - s0);
        for (int pix0 = 0; pix0 < npix_total; pix0 += npix) {
            int np = std::min(npix, npix_total - pix0);
            int batch = np * sub;
            int grid = plain ? div_up(batch, kBlock) : std::min(div_up(batch, kBlock), fixed_grid);
            LUX_LAUNCH("pt_generate", k_generate, grid, kBlock, 0, stream, sv, ps, qs, w, h, pix0, np, s0, sub,
                       rs.seed, rs.frame_index, rs.sample_offset);
            for (int b = 0; b <= depth; ++b) {
                int in = b & 1;
                LUX_LAUNCH("pt_intersect", k_intersect, grid, kBlock, 0, stream, sv, ps, qs, in);
                if (plain) {
                    for (int type = 0; type < kMaterialTypes; ++type)
                        LUX_LAUNCH("pt_shade", k_shade, grid, kBlock, 0, stream, sv, ps, qs, in, b, depth, type);
                } else {
                    LUX_LAUNCH("pt_shade", k_shade, grid, kBlock, 0, stream, sv, ps, qs, in, b, depth, -1);
                }
                LUX_LAUNCH("pt_shadow", k_shadow, grid, kBlock, 0, stream, sv, ps, qs);
            }
            LUX_LAUNCH("pt_flush", k_flush, std::min(div_up(np, kBlock), 4096), kBlock, 0, stream, ps, film, pix0, np,
                       sub);
            if (stats) {
                stats->batches += 1;
                stats->paths += uint64_t(batch);
            }
            if (plain) {  // the reference's per-batch synchronisation and counter read-back
                hip_check(hipMemcpyAsync(host_counts.data(), buf.counts, 16, hipMemcpyDeviceToHost, stream), "D2H counts");
                hip_check(hipStreamSynchronize(stream), "batch synchronize");
            }
        }
    }
}

}  // namespace pt
}  // namespace lux

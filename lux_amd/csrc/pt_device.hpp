// pt_device.hpp -- device-side shading math of the wavefront path tracer:
// vector operations with the reference semantics (src/core/vec3.cuh: division
// multiplies by the reciprocal, left-to-right sums), PCG32 sampler stream,
// pinhole camera, the six BSDFs, emission and power-alias light sampling.
//
// Everything follows the reference operation order so a GPU path consumes the
// sampler dimensions in the same order as the reference CPU path
// (src/integrator/cpu_path_tracer.cpp:403-498); only libm-style functions
// (cosf, sinf, powf) round differently on the GPU, which is why renders are
// compared by relative RMSE and not bit for bit.
#pragma once

#include "lux_common.hpp"

namespace lux {
namespace pt {

struct vec {
    float x, y, z;
};
__device__ __forceinline__ vec mk(float x, float y, float z) { return vec{x, y, z}; }
__device__ __forceinline__ vec splat(float s) { return vec{s, s, s}; }
__device__ __forceinline__ vec operator-(const vec& a) { return mk(-a.x, -a.y, -a.z); }
__device__ __forceinline__ vec operator+(const vec& a, const vec& b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ vec operator-(const vec& a, const vec& b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ vec operator*(const vec& a, const vec& b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
__device__ __forceinline__ vec operator*(const vec& a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ vec operator*(float s, const vec& a) { return a * s; }
__device__ __forceinline__ vec operator/(const vec& a, float s) {
    float inv = 1 / s;
    return mk(a.x * inv, a.y * inv, a.z * inv);
}
__device__ __forceinline__ float dot(const vec& a, const vec& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ vec cross(const vec& a, const vec& b) {
    return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ float length2(const vec& v) { return dot(v, v); }
__device__ __forceinline__ float length(const vec& v) { return sqrtf(length2(v)); }
__device__ __forceinline__ vec normalize(const vec& v) {
    float len = length(v);
    return len > 0 ? v / len : v;
}
__device__ __forceinline__ float max_component(const vec& v) { return fmaxf(v.x, fmaxf(v.y, v.z)); }

constexpr float kPi = 3.14159265358979323846f;
constexpr float kInvPi = 1.0f / kPi;
constexpr float kEps = 1e-5f;

// ---- records shared with the host ----------------------------------------------
enum : int { kLambert = 0, kThinLambert, kGgx, kDielectric, kInterface, kTransparent, kMaterialTypes };
enum : unsigned { kDiffuse = 1, kGlossy = 2, kDeltaRefl = 4, kDeltaTrans = 8, kTrans = 16 };
enum : int { kMeasureNone = 0, kSolidAngle = 1, kDiscrete = 3 };

struct DevMaterial {
    int type;
    float ar, ag, ab;
    float er, eg, eb;
    int two_sided;
    float roughness, metallic, eta;
    int pad;
};

// Light kinds of the light table (LightKind, src/light/light.h:22-33).
constexpr int kLightArea = 0, kLightPoint = 1, kLightDistant = 2, kLightUniformInf = 3, kLightImageInf = 4;

// One entry of the light table (64 bytes).
struct DevLight {
    int prim;          // area: triangle ordinal; image: first texel of its image
    float er, eg, eb;  // area: emission; point: intensity; distant, uniform: radiance; image: scale
    int two_sided;     // area: double-sided emission; image: width
    float area;        // area: area; image: total texel weight
    float prob;        // power alias table: column probability, alias, pmf
    int alias;
    float pmf;
    int kind;          // kLight*
    int height;        // image: height
    int nb_slot;       // slot in the BVH sampler's non-BVH table (-1: in the tree, or no tree)
    float vx, vy, vz;  // point: position; distant: direction to the light
    int pad;
};
static_assert(sizeof(DevLight) == 64, "light record");

// Alias-table entry of the BVH sampler's non-BVH group (lights without bounds).
struct DevAlias {
    float prob;
    int alias;
    float pmf;
    int light;  // its index in the light table
};

// Light-BVH node (LightBvhNode, src/light/light_distribution.h:33-40): bounds,
// power, light index (leaf) or -1, children, parent; three 16-byte loads.
struct DevLightNode {
    float lo[3];
    float power;
    float hi[3];
    int light;
    int left, right, parent, pad;
};
static_assert(sizeof(DevLightNode) == 48, "light node record");

struct DevCamera {
    float ox, oy, oz;
    float lx, ly, lz;  // lower left
    float hx, hy, hz;  // horizontal
    float vx, vy, vz;  // vertical
};

__device__ __forceinline__ vec albedo_of(const DevMaterial& m) { return mk(m.ar, m.ag, m.ab); }
__device__ __forceinline__ vec emission_of(const DevMaterial& m) { return mk(m.er, m.eg, m.eb); }
__device__ __forceinline__ bool scattering(const DevMaterial& m) {
    if (m.type == kLambert || m.type == kThinLambert) return max_component(albedo_of(m)) > 0;
    return m.type == kGgx || m.type == kDielectric;
}
__device__ __forceinline__ bool passthrough(const DevMaterial& m) { return m.type == kInterface || m.type == kTransparent; }
__device__ __forceinline__ bool direct_lobe(const DevMaterial& m) {
    return m.type == kLambert || m.type == kThinLambert || m.type == kGgx;
}

// ---- PCG32 (src/core/pcg.cuh:11-41) ----------------------------------------------
struct Rng {
    uint64_t state, inc;
    __device__ __forceinline__ uint32_t next() {
        uint64_t old = state;
        state = old * 6364136223846793005ULL + inc;
        uint32_t xs = uint32_t(((old >> 18u) ^ old) >> 27u);
        uint32_t rot = uint32_t(old >> 59u);
        return (xs >> rot) | (xs << ((-rot) & 31));
    }
    __device__ __forceinline__ float get1() { return float(next() >> 8) / float(1 << 24); }
};

// make_render_sampler (src/integrator/render_context.h:37-54, src/sampler/sampler.h:165-176)
__device__ __forceinline__ Rng make_rng(uint64_t seed, int frame, int pixel, int sample) {
    uint64_t base = seed ^ (uint64_t(frame) * 0x9e3779b97f4a7c15ULL);
    uint64_t s = uint64_t(sample);
    uint64_t state = base + uint64_t(pixel) * 747796405ULL + s;
    uint64_t seq = (base >> 1) ^ (s * 2891336453ULL + uint64_t(pixel));
    Rng r;
    r.state = 0;
    r.inc = (seq << 1u) | 1u;
    r.next();
    r.state += state;
    r.next();
    return r;
}

// ---- emission (src/material/material.h:121-145) -----------------------------------
__device__ __forceinline__ float emission_cos(const vec& ng, const vec& wo, int two_sided) {
    float c = dot(ng, wo);
    if (!two_sided) return c > 1e-7f ? c : 0.0f;
    c = fabsf(c);
    return c > 1e-7f ? c : 0.0f;
}
__device__ __forceinline__ vec eval_emission(const vec& e, int two_sided, const vec& ng, const vec& wo) {
    if (max_component(e) <= 0) return splat(0);
    if (emission_cos(ng, wo, two_sided) <= 0) return splat(0);
    return e;
}

// ---- BSDFs (src/material/bsdf.h) ---------------------------------------------------
struct Interaction {
    vec p, ng, ns, wo;
};
struct BsdfSample {
    vec wi, f, weight;
    float pdf, eta;
    int measure;
    unsigned flags;
};
__device__ __forceinline__ BsdfSample empty_sample() {
    BsdfSample s;
    s.wi = s.f = s.weight = splat(0);
    s.pdf = 0;
    s.eta = 1;
    s.measure = kMeasureNone;
    s.flags = 0;
    return s;
}
__device__ __forceinline__ bool is_delta(unsigned f) { return (f & (kDeltaRefl | kDeltaTrans)) != 0; }
__device__ __forceinline__ bool is_transmission(unsigned f) { return (f & (kTrans | kDeltaTrans)) != 0; }

__device__ __forceinline__ vec make_tangent(const vec& n) {
    return fabsf(n.x) > 0.9f ? normalize(cross(mk(0, 0, 1), n)) : normalize(cross(mk(1, 0, 0), n));
}
__device__ __forceinline__ vec cosine_hemisphere(const vec& n, float ux, float uy) {
    float radius = sqrtf(ux);
    float phi = 2.0f * kPi * uy;
    float x = radius * cosf(phi);
    float z = radius * sinf(phi);
    float y = sqrtf(fmaxf(0.0f, 1.0f - ux));
    vec t = make_tangent(n);
    vec b = cross(n, t);
    return normalize(x * t + y * n + z * b);
}
__device__ __forceinline__ bool same_hemi(const vec& n, const vec& a, const vec& b) {
    return dot(n, a) * dot(n, b) > 0;
}
__device__ __forceinline__ float abs_cos(const vec& n, const vec& d) { return fabsf(dot(n, d)); }
__device__ __forceinline__ vec orient(const vec& n, const vec& r) { return dot(n, r) < 0 ? -n : n; }
__device__ __forceinline__ void finalize_smooth(BsdfSample& s, const Interaction& it) {
    if (s.pdf > 0) s.measure = kSolidAngle;
    float c = abs_cos(it.ns, s.wi);
    if (s.pdf > 0 && c > 0 && max_component(s.f) > 0) s.weight = s.f * (c / s.pdf);
}

__device__ __forceinline__ vec eval_lambert(const DevMaterial& m, const Interaction& it, const vec& wi) {
    if (!same_hemi(it.ns, it.wo, wi)) return splat(0);
    return albedo_of(m) * kInvPi;
}
__device__ __forceinline__ float pdf_lambert(const Interaction& it, const vec& wi) {
    if (!same_hemi(it.ns, it.wo, wi)) return 0;
    return abs_cos(it.ns, wi) * kInvPi;
}
__device__ __forceinline__ vec eval_thin(const DevMaterial& m) { return albedo_of(m) * (0.5f * kInvPi); }
__device__ __forceinline__ float pdf_thin(const Interaction& it, const vec& wi) {
    return 0.5f * abs_cos(it.ns, wi) * kInvPi;
}

__device__ __forceinline__ float clamp01(float v) { return fmaxf(0.0f, fminf(1.0f, v)); }
__device__ __forceinline__ float ggx_metal(const DevMaterial& m) { return clamp01(m.metallic); }
__device__ __forceinline__ vec ggx_f0(const DevMaterial& m) {
    float t = ggx_metal(m);
    return splat(0.04f) * (1.0f - t) + albedo_of(m) * t;
}
__device__ __forceinline__ float ggx_diffuse_scale(const DevMaterial& m) {
    vec f0 = ggx_f0(m);
    return (1.0f - ggx_metal(m)) * fmaxf(0.0f, 1.0f - max_component(f0));
}
__device__ __forceinline__ float ggx_alpha(const DevMaterial& m) {
    float r = fmaxf(m.roughness, 0.02f);
    return fmaxf(r * r, 1e-4f);
}
__device__ __forceinline__ float ggx_d(const vec& h, const vec& n, float a) {
    float c = dot(n, h);
    if (c <= 0) return 0;
    float a2 = a * a;
    float d = (c * c) * (a2 - 1.0f) + 1.0f;
    return a2 / (kPi * d * d);
}
__device__ __forceinline__ float ggx_g1(const vec& v, const vec& n, float a) {
    float c = dot(n, v);
    if (c <= 0) return 0;
    float a2 = a * a;
    float tan2 = (1.0f - c * c) / (c * c);
    return 2.0f / (1.0f + sqrtf(1.0f + a2 * tan2));
}
__device__ __forceinline__ float ggx_lambda(const vec& v, const vec& n, float a) {
    float c = dot(n, v);
    if (c <= 0) return 0;
    float c2 = c * c;
    float tan2 = fmaxf(0.0f, 1.0f - c2) / c2;
    return (sqrtf(1.0f + a * a * tan2) - 1.0f) * 0.5f;
}
__device__ __forceinline__ vec schlick(const vec& f0, float c) {
    c = clamp01(c);
    float p5 = powf(1.0f - c, 5.0f);
    return f0 + (splat(1.0f) - f0) * p5;
}
__device__ __forceinline__ vec eval_ggx_spec(const DevMaterial& m, const Interaction& it, const vec& wi) {
    vec wo = it.wo, n = it.ns;
    if (dot(n, wo) <= 0 || dot(n, wi) <= 0) return splat(0);
    vec h = normalize(wo + wi);
    float a = ggx_alpha(m);
    float D = ggx_d(h, n, a);
    float G = 1.0f / (1.0f + ggx_lambda(wo, n, a) + ggx_lambda(wi, n, a));
    vec F = schlick(ggx_f0(m), dot(wo, h));
    float denom = 4.0f * dot(n, wo) * dot(n, wi);
    if (denom <= 0) return splat(0);
    return (D * G * F) / denom;
}
__device__ __forceinline__ float pdf_ggx_spec(const DevMaterial& m, const Interaction& it, const vec& wi) {
    vec wo = it.wo, n = it.ns;
    if (dot(n, wo) <= 0 || dot(n, wi) <= 0) return 0;
    vec h = normalize(wo + wi);
    float a = ggx_alpha(m);
    float D = ggx_d(h, n, a);
    float acw = fabsf(dot(n, wo));
    float awh = fabsf(dot(wo, h));
    if (acw <= 0 || awh <= 0) return 0;
    float pdf_wm = D * ggx_g1(wo, n, a) * awh / acw;
    return pdf_wm / (4.0f * awh);
}
__device__ __forceinline__ vec ggx_vndf(const vec& wo, const vec& n, float a, float ux, float uy) {
    vec t = make_tangent(n);
    vec b = cross(n, t);
    vec wl = mk(dot(wo, t), dot(wo, b), dot(wo, n));
    vec wh = normalize(mk(a * wl.x, a * wl.y, wl.z));
    if (wh.z < 0) wh = -wh;
    vec t1 = wh.z < 0.99999f ? normalize(cross(mk(0, 0, 1), wh)) : mk(1, 0, 0);
    vec t2 = cross(wh, t1);
    float r = sqrtf(ux);
    float phi = 2.0f * kPi * uy;
    float px = r * cosf(phi), py = r * sinf(phi);
    float h = sqrtf(fmaxf(0.0f, 1.0f - px * px));
    float blend = (1.0f + wh.z) * 0.5f;
    py = (1.0f - blend) * h + blend * py;
    float pz = sqrtf(fmaxf(0.0f, 1.0f - px * px - py * py));
    vec nh = px * t1 + py * t2 + pz * wh;
    vec wml = normalize(mk(a * nh.x, a * nh.y, fmaxf(1e-6f, nh.z)));
    vec wm = normalize(wml.x * t + wml.y * b + wml.z * n);
    return dot(wm, n) < 0 ? -wm : wm;
}
__device__ __forceinline__ void ggx_lobes(const DevMaterial& m, float& pd, float& ps) {
    bool dv = max_component(albedo_of(m)) > 0 && ggx_diffuse_scale(m) > 0;
    bool sv = max_component(ggx_f0(m)) > 0;
    pd = ps = 0;
    if (sv && ggx_metal(m) >= 1.0f - 1e-4f) {
        ps = 1;
        return;
    }
    if (dv && sv) {
        pd = ps = 0.5f;
        return;
    }
    if (dv) pd = 1;
    if (!dv && sv) ps = 1;
}
__device__ __forceinline__ vec eval_ggx(const DevMaterial& m, const Interaction& it, const vec& wi) {
    vec diffuse = (dot(it.ns, it.wo) <= 0 || dot(it.ns, wi) <= 0) ? splat(0)
                                                                   : albedo_of(m) * (ggx_diffuse_scale(m) * kInvPi);
    return diffuse + eval_ggx_spec(m, it, wi);
}
__device__ __forceinline__ float pdf_ggx(const DevMaterial& m, const Interaction& it, const vec& wi) {
    float pd, ps;
    ggx_lobes(m, pd, ps);
    return pd * pdf_lambert(it, wi) + ps * pdf_ggx_spec(m, it, wi);
}
__device__ __forceinline__ float fresnel(float ci, float ei, float et) {
    float si = sqrtf(fmaxf(0.0f, 1.0f - ci * ci));
    float st = ei / et * si;
    if (st >= 1.0f) return 1.0f;
    float ct = sqrtf(fmaxf(0.0f, 1.0f - st * st));
    float rpa = (et * ci - ei * ct) / (et * ci + ei * ct);
    float rpe = (ei * ci - et * ct) / (ei * ci + et * ct);
    return (rpa * rpa + rpe * rpe) * 0.5f;
}

__device__ __forceinline__ vec eval_bsdf(const DevMaterial& m, const Interaction& it, const vec& wi) {
    switch (m.type) {
        case kLambert: return eval_lambert(m, it, wi);
        case kThinLambert: return eval_thin(m);
        case kGgx: return eval_ggx(m, it, wi);
        default: return splat(0);
    }
}
__device__ __forceinline__ float pdf_bsdf(const DevMaterial& m, const Interaction& it, const vec& wi) {
    switch (m.type) {
        case kLambert: return pdf_lambert(it, wi);
        case kThinLambert: return pdf_thin(it, wi);
        case kGgx: return pdf_ggx(m, it, wi);
        default: return 0;
    }
}

__device__ __forceinline__ BsdfSample sample_bsdf(const DevMaterial& m, const Interaction& it, float ux, float uy) {
    BsdfSample s = empty_sample();
    switch (m.type) {
        case kLambert: {
            vec n = orient(it.ns, it.wo);
            s.wi = cosine_hemisphere(n, ux, uy);
            s.f = eval_lambert(m, it, s.wi);
            s.pdf = pdf_lambert(it, s.wi);
            s.flags = kDiffuse;
            finalize_smooth(s, it);
            return s;
        }
        case kThinLambert: {
            bool refl = ux < 0.5f;
            float rx = refl ? ux * 2.0f : (ux - 0.5f) * 2.0f;
            vec wn = orient(it.ns, it.wo);
            s.wi = cosine_hemisphere(refl ? wn : -wn, rx, uy);
            s.f = eval_thin(m);
            s.pdf = pdf_thin(it, s.wi);
            s.flags = refl ? kDiffuse : (kDiffuse | kTrans);
            finalize_smooth(s, it);
            return s;
        }
        case kGgx: {
            vec n = it.ns;
            if (dot(n, it.wo) <= 0) return s;
            float pd, ps;
            ggx_lobes(m, pd, ps);
            if (pd <= 0 && ps <= 0) return s;
            if (pd > 0 && (ux < pd || ps <= 0)) {
                s.wi = cosine_hemisphere(n, ux / pd, uy);
                s.flags = kDiffuse;
            } else {
                // sample_ggx_specular (src/material/bsdf.h:323-343)
                float a = ggx_alpha(m);
                vec h = ggx_vndf(it.wo, n, a, (ux - pd) / ps, uy);
                vec wi = normalize(2.0f * dot(it.wo, h) * h - it.wo);
                if (dot(n, wi) <= 0) return s;
                if (pdf_ggx_spec(m, it, wi) <= 0) return s;
                s.wi = wi;
                s.flags = kGlossy;
            }
            s.f = eval_ggx(m, it, s.wi);
            s.pdf = pdf_ggx(m, it, s.wi);
            finalize_smooth(s, it);
            return s;
        }
        case kDielectric: {
            vec wo = it.wo, n = it.ns;
            float c = dot(wo, n);
            bool entering = c > 0;
            float ei = entering ? 1.0f : m.eta;
            float et = entering ? m.eta : 1.0f;
            vec nn = entering ? n : -n;
            c = fabsf(c);
            float F = fresnel(c, ei, et);
            s.measure = kDiscrete;
            if (ux < F) {
                s.wi = normalize(2.0f * c * nn - wo);
                s.weight = albedo_of(m);
                s.pdf = F;
                s.flags = kDeltaRefl;
            } else {
                float st = (ei / et) * sqrtf(fmaxf(0.0f, 1.0f - c * c));
                if (st >= 1.0f) {
                    s.wi = normalize(2.0f * c * nn - wo);
                    s.weight = albedo_of(m);
                    s.pdf = 1.0f;
                    s.flags = kDeltaRefl;
                    return s;
                }
                float ct = sqrtf(fmaxf(0.0f, 1.0f - st * st));
                s.wi = normalize((ei / et) * (-wo) + ((ei / et) * c - ct) * nn);
                float etap = et / ei;
                s.weight = albedo_of(m) / (etap * etap);
                s.pdf = 1.0f - F;
                s.eta = etap;
                s.flags = kDeltaTrans | kTrans;
            }
            return s;
        }
        default:  // Interface / Transparent
            s.wi = normalize(-it.wo);
            s.weight = m.type == kTransparent ? albedo_of(m) : splat(1);
            s.pdf = 1.0f;
            s.measure = kDiscrete;
            s.flags = kDeltaTrans | kTrans;
            return s;
    }
}

__device__ __forceinline__ float power_heuristic(float f, float g) {
    if (f <= 0) return 0;
    float a = f * f;
    float b = fmaxf(0.0f, g);
    b *= b;
    float d = a + b;
    return d > 0 ? a / d : 0;
}

// pdf_solid_angle of an area-light point (src/light/light_sampler.h:374-384)
__device__ __forceinline__ float pdf_solid_angle(const vec& lp, const vec& ln, int two_sided, float pdf_area,
                                                 const vec& ref) {
    vec to = lp - ref;
    float d2 = length2(to);
    if (d2 <= kEps * kEps || pdf_area <= 0) return 0;
    vec wi = to / sqrtf(d2);
    float cl = emission_cos(ln, -wi, two_sided);
    if (cl <= 0) return 0;
    return pdf_area * d2 / cl;
}

__device__ __forceinline__ vec tri_vertex(const lux_triangle& t, int k) {
    const lux_vec3& p = k == 0 ? t.v0 : (k == 1 ? t.v1 : t.v2);
    return mk(p.x, p.y, p.z);
}
__device__ __forceinline__ vec tri_normal(const lux_triangle& t) {
    vec v0 = tri_vertex(t, 0);
    return normalize(cross(tri_vertex(t, 1) - v0, tri_vertex(t, 2) - v0));
}

}  // namespace pt
}  // namespace lux

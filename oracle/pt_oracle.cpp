// pt_oracle.cpp -- CPU restatement of LUX's reference path tracer
// (src/integrator/cpu_path_tracer.cpp: trace_path, estimate_direct_light,
// trace_shadow) with the power light sampler, the six BSDFs, the PCG32 sampler
// stream and the pinhole camera, for a scene of one mesh under an identity
// instance (the one-level walk) and for instanced scenes (meshes placed by
// affine instances: the two-level walk of two_level_oracle.cpp, shading on the
// world-space triangle view).
//
// TEST INFRASTRUCTURE ONLY: the checker for the HIP wavefront path tracer, and
// the CPU baseline of bench.py. Same rules as lbvh_oracle.cpp. Arithmetic keeps
// the reference operation order (built with -ffp-contract=off, host libm), so a
// render matches the reference's own render_cpu_path_tracer bit for bit on the
// same machine; tests/golden holds renders and per-sample radiance written by
// the reference to pin that.

#include <cstring>
#include <thread>

#include "oracle_core.hpp"

// The two-level walk of two_level_oracle.cpp (same library).
extern "C" {
struct orc_tl;
orc_tl* orc_tl_build(const float* tris, const int32_t* tri_off, int32_t mesh_count, const int32_t* inst_mesh,
                     const float* inst_inv, const float* inst_world_bounds, int32_t instance_count);
void orc_tl_free(orc_tl* h);
int32_t orc_tl_trace(const orc_tl* h, const float* rays, const int32_t* ignore, int32_t n, int32_t* prim,
                     float* t_out, float* u_out, float* v_out, int32_t* inst_out, int32_t* local_out);
void orc_inverse_affine(const float* m16, float* out16);
void orc_transform_bounds(const float* m16, const float* box6, float* out6);
}

namespace pt {

// ---- vec3 with the reference's operator semantics (src/core/vec3.cuh) -------
struct V {
    float x = 0, y = 0, z = 0;
    V() = default;
    V(float a, float b, float c) : x(a), y(b), z(c) {}
    explicit V(float s) : x(s), y(s), z(s) {}
    V operator-() const { return V(-x, -y, -z); }
    V operator+(const V& b) const { return V(x + b.x, y + b.y, z + b.z); }
    V operator-(const V& b) const { return V(x - b.x, y - b.y, z - b.z); }
    V operator*(const V& b) const { return V(x * b.x, y * b.y, z * b.z); }
    V operator*(float s) const { return V(x * s, y * s, z * s); }
    V operator/(float s) const {  // multiply by the reciprocal, as vec3::operator/
        float inv = 1 / s;
        return V(x * inv, y * inv, z * inv);
    }
    V& operator+=(const V& b) {
        x += b.x;
        y += b.y;
        z += b.z;
        return *this;
    }
};
inline V operator*(float s, const V& v) { return v * s; }
inline float dot(const V& a, const V& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V cross(const V& a, const V& b) { return V(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
inline float length2(const V& v) { return dot(v, v); }
inline float length(const V& v) { return sqrtf(length2(v)); }
inline V normalize(const V& v) {
    float len = length(v);
    return len > 0 ? v / len : v;
}
inline float max_component(const V& v) { return fmaxf(v.x, fmaxf(v.y, v.z)); }
inline float luminance(const V& v) { return float(0.2126) * v.x + float(0.7152) * v.y + float(0.0722) * v.z; }

constexpr float kPi = 3.14159265358979323846f;
constexpr float kInvPi = 1.0f / kPi;
constexpr float kEps = 1e-5f;  // kRayEpsilon

// ---- PCG32 + sampler (src/core/pcg.cuh:11-41, src/sampler/sampler.h:165-190) --
struct Rng {
    uint64_t state = 0, inc = 1;
    int dimension = 0;
    uint32_t next() {
        uint64_t old = state;
        state = old * 6364136223846793005ULL + inc;
        uint32_t xs = uint32_t(((old >> 18u) ^ old) >> 27u);
        uint32_t rot = uint32_t(old >> 59u);
        return (xs >> rot) | (xs << ((-rot) & 31));
    }
    float get1() {
        ++dimension;
        return float(next() >> 8) / float(1 << 24);
    }
};

// make_render_sampler (src/integrator/render_context.h:37-54)
inline Rng make_rng(uint64_t seed, int frame, int pixel, int sample) {
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

// ---- materials (src/material/material.h, src/material/bsdf.h) ---------------
enum : int { kLambert = 0, kThinLambert, kGgx, kDielectric, kInterface, kTransparent };
enum : unsigned { kDiffuse = 1, kGlossy = 2, kDeltaRefl = 4, kDeltaTrans = 8, kTrans = 16 };
enum Measure : int { kNone = 0, kSolidAngle, kArea, kDiscrete };

struct Material {
    int type = kLambert;
    V albedo, emission;
    bool two_sided = false;
    float roughness = 0.5f, metallic = 0, eta = 1.5f;
    bool scattering() const {
        if (type == kLambert || type == kThinLambert) return max_component(albedo) > 0;
        return type == kGgx || type == kDielectric;
    }
    bool emissive() const { return max_component(emission) > 0; }
    bool passthrough() const { return type == kInterface || type == kTransparent; }
    bool direct_lobe() const { return type == kLambert || type == kThinLambert || type == kGgx; }
};

// emission_projected_cosine / eval_emission (src/material/material.h:121-145)
inline float emission_cos(const V& ng, const V& wo, bool two_sided) {
    float c = dot(ng, wo);
    if (!two_sided) return c > float(1e-7) ? c : 0.0f;
    c = fabsf(c);
    return c > float(1e-7) ? c : 0.0f;
}
inline V eval_emission(const V& e, bool two_sided, const V& ng, const V& wo) {
    if (max_component(e) <= 0) return V(0);
    if (emission_cos(ng, wo, two_sided) <= 0) return V(0);
    return e;
}

struct Interaction {
    V p, ng, ns, wo;
};

struct BsdfSample {
    V wi, f, weight;
    float pdf = 0, eta = 1;
    int measure = kNone;
    unsigned flags = 0;
};

inline bool is_delta(unsigned f) { return (f & (kDeltaRefl | kDeltaTrans)) != 0; }
inline bool is_transmission(unsigned f) { return (f & (kTrans | kDeltaTrans)) != 0; }

inline V make_tangent(const V& n) {
    return fabsf(n.x) > 0.9f ? normalize(cross(V(0, 0, 1), n)) : normalize(cross(V(1, 0, 0), n));
}
// cosine_sample_hemisphere (src/material/bsdf.h:64-74)
inline V cosine_hemisphere(const V& n, float ux, float uy) {
    float radius = sqrtf(ux);
    float phi = 2.0f * kPi * uy;
    float x = radius * cosf(phi);
    float z = radius * sinf(phi);
    float y = sqrtf(fmaxf(0.0f, 1.0f - ux));
    V t = make_tangent(n);
    V b = cross(n, t);
    return normalize(x * t + y * n + z * b);
}
inline bool same_hemi(const V& n, const V& a, const V& b) { return dot(n, a) * dot(n, b) > 0; }
inline float abs_cos(const V& n, const V& d) { return fabsf(dot(n, d)); }
inline V orient(const V& n, const V& ref) { return dot(n, ref) < 0 ? -n : n; }
inline void finalize_smooth(BsdfSample& s, const Interaction& it) {
    if (s.pdf > 0) s.measure = kSolidAngle;
    float c = abs_cos(it.ns, s.wi);
    if (s.pdf > 0 && c > 0 && max_component(s.f) > 0) s.weight = s.f * (c / s.pdf);
}

// Lambert / thin Lambert (src/material/bsdf.h:93-155)
inline V eval_lambert(const Material& m, const Interaction& it, const V& wi) {
    if (!same_hemi(it.ns, it.wo, wi)) return V(0);
    return m.albedo * kInvPi;
}
inline float pdf_lambert(const Interaction& it, const V& wi) {
    if (!same_hemi(it.ns, it.wo, wi)) return 0;
    return abs_cos(it.ns, wi) * kInvPi;
}
inline V eval_thin(const Material& m) { return m.albedo * (0.5f * kInvPi); }
inline float pdf_thin(const Interaction& it, const V& wi) { return 0.5f * abs_cos(it.ns, wi) * kInvPi; }

// GGX (src/material/bsdf.h:161-420)
inline float clamp01(float v) { return fmaxf(0.0f, fminf(1.0f, v)); }
inline V mixv(const V& a, const V& b, float t) { return a * (1.0f - t) + b * t; }
inline float ggx_metal(const Material& m) { return clamp01(m.metallic); }
inline V ggx_f0(const Material& m) { return mixv(V(0.04f), m.albedo, ggx_metal(m)); }
inline float ggx_diffuse_scale(const Material& m) {
    V f0 = ggx_f0(m);
    return (1.0f - ggx_metal(m)) * fmaxf(0.0f, 1.0f - max_component(f0));
}
inline float ggx_alpha(const Material& m) {
    float r = fmaxf(m.roughness, 0.02f);
    return fmaxf(r * r, 1e-4f);
}
inline float ggx_d(const V& h, const V& n, float a) {
    float c = dot(n, h);
    if (c <= 0) return 0;
    float a2 = a * a;
    float d = (c * c) * (a2 - 1.0f) + 1.0f;
    return a2 / (kPi * d * d);
}
inline float ggx_g1(const V& v, const V& n, float a) {
    float c = dot(n, v);
    if (c <= 0) return 0;
    float a2 = a * a;
    float tan2 = (1.0f - c * c) / (c * c);
    return 2.0f / (1.0f + sqrtf(1.0f + a2 * tan2));
}
inline float ggx_lambda(const V& v, const V& n, float a) {
    float c = dot(n, v);
    if (c <= 0) return 0;
    float c2 = c * c;
    float tan2 = fmaxf(0.0f, 1.0f - c2) / c2;
    return (sqrtf(1.0f + a * a * tan2) - 1.0f) * 0.5f;
}
inline float ggx_g(const V& wo, const V& wi, const V& n, float a) {
    return 1.0f / (1.0f + ggx_lambda(wo, n, a) + ggx_lambda(wi, n, a));
}
inline V schlick(const V& f0, float c) {
    c = clamp01(c);
    float p5 = powf(1.0f - c, 5.0f);
    return f0 + (V(1.0f) - f0) * p5;
}
inline V eval_ggx_diffuse(const Material& m, const Interaction& it, const V& wi) {
    if (dot(it.ns, it.wo) <= 0 || dot(it.ns, wi) <= 0) return V(0);
    return m.albedo * (ggx_diffuse_scale(m) * kInvPi);
}
inline V eval_ggx_spec(const Material& m, const Interaction& it, const V& wi) {
    V wo = it.wo, n = it.ns;
    if (dot(n, wo) <= 0 || dot(n, wi) <= 0) return V(0);
    V h = normalize(wo + wi);
    float a = ggx_alpha(m);
    float D = ggx_d(h, n, a);
    float G = ggx_g(wo, wi, n, a);
    V F = schlick(ggx_f0(m), dot(wo, h));
    float denom = 4.0f * dot(n, wo) * dot(n, wi);
    if (denom <= 0) return V(0);
    return (D * G * F) / denom;
}
inline float pdf_ggx_spec(const Material& m, const Interaction& it, const V& wi) {
    V wo = it.wo, n = it.ns;
    if (dot(n, wo) <= 0 || dot(n, wi) <= 0) return 0;
    V h = normalize(wo + wi);
    float a = ggx_alpha(m);
    float D = ggx_d(h, n, a);
    float acw = fabsf(dot(n, wo));
    float awh = fabsf(dot(wo, h));
    if (acw <= 0 || awh <= 0) return 0;
    float pdf_wm = D * ggx_g1(wo, n, a) * awh / acw;
    return pdf_wm / (4.0f * awh);
}
inline V to_local(const V& v, const V& t, const V& b, const V& n) { return V(dot(v, t), dot(v, b), dot(v, n)); }
inline V from_local(const V& v, const V& t, const V& b, const V& n) { return normalize(v.x * t + v.y * b + v.z * n); }
inline V ggx_vndf(const V& wo, const V& n, float a, float ux, float uy) {
    V t = make_tangent(n);
    V b = cross(n, t);
    V wl = to_local(wo, t, b, n);
    V wh = normalize(V(a * wl.x, a * wl.y, wl.z));
    if (wh.z < 0) wh = -wh;
    V t1 = wh.z < 0.99999f ? normalize(cross(V(0, 0, 1), wh)) : V(1, 0, 0);
    V t2 = cross(wh, t1);
    float r = sqrtf(ux);
    float phi = 2.0f * kPi * uy;
    float px = r * cosf(phi), py = r * sinf(phi);
    float h = sqrtf(fmaxf(0.0f, 1.0f - px * px));
    float blend = (1.0f + wh.z) * 0.5f;
    py = (1.0f - blend) * h + blend * py;
    float pz = sqrtf(fmaxf(0.0f, 1.0f - px * px - py * py));
    V nh = px * t1 + py * t2 + pz * wh;
    V wml = normalize(V(a * nh.x, a * nh.y, fmaxf(1e-6f, nh.z)));
    V wm = from_local(wml, t, b, n);
    return dot(wm, n) < 0 ? -wm : wm;
}
inline BsdfSample sample_ggx_spec(const Material& m, const Interaction& it, float ux, float uy) {
    BsdfSample s;
    V wo = it.wo, n = it.ns;
    if (dot(n, wo) <= 0) return s;
    float a = ggx_alpha(m);
    V h = ggx_vndf(wo, n, a, ux, uy);
    s.wi = normalize(2.0f * dot(wo, h) * h - wo);
    if (dot(n, s.wi) <= 0) return s;
    s.f = eval_ggx_spec(m, it, s.wi);
    s.pdf = pdf_ggx_spec(m, it, s.wi);
    s.flags = kGlossy;
    finalize_smooth(s, it);
    return s;
}
inline void ggx_lobes(const Material& m, float& pd, float& ps) {
    bool dv = max_component(m.albedo) > 0 && ggx_diffuse_scale(m) > 0;
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
inline V eval_ggx(const Material& m, const Interaction& it, const V& wi) {
    return eval_ggx_diffuse(m, it, wi) + eval_ggx_spec(m, it, wi);
}
inline float pdf_ggx(const Material& m, const Interaction& it, const V& wi) {
    float pd, ps;
    ggx_lobes(m, pd, ps);
    return pd * pdf_lambert(it, wi) + ps * pdf_ggx_spec(m, it, wi);
}

// dielectric (src/material/bsdf.h:426-494)
inline float fresnel(float ci, float ei, float et) {
    float si = sqrtf(fmaxf(0.0f, 1.0f - ci * ci));
    float st = ei / et * si;
    if (st >= 1.0f) return 1.0f;
    float ct = sqrtf(fmaxf(0.0f, 1.0f - st * st));
    float rpa = (et * ci - ei * ct) / (et * ci + ei * ct);
    float rpe = (ei * ci - et * ct) / (ei * ci + et * ct);
    return (rpa * rpa + rpe * rpe) * 0.5f;
}

inline V eval_bsdf(const Material& m, const Interaction& it, const V& wi) {
    switch (m.type) {
        case kLambert: return eval_lambert(m, it, wi);
        case kThinLambert: return eval_thin(m);
        case kGgx: return eval_ggx(m, it, wi);
        default: return V(0);
    }
}
inline float pdf_bsdf(const Material& m, const Interaction& it, const V& wi) {
    switch (m.type) {
        case kLambert: return pdf_lambert(it, wi);
        case kThinLambert: return pdf_thin(it, wi);
        case kGgx: return pdf_ggx(m, it, wi);
        default: return 0;
    }
}

inline BsdfSample sample_bsdf(const Material& m, const Interaction& it, float ux, float uy) {
    BsdfSample s;
    switch (m.type) {
        case kLambert: {
            V n = orient(it.ns, it.wo);
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
            V wn = orient(it.ns, it.wo);
            V n = refl ? wn : -wn;
            s.wi = cosine_hemisphere(n, rx, uy);
            s.f = eval_thin(m);
            s.pdf = pdf_thin(it, s.wi);
            s.flags = refl ? kDiffuse : (kDiffuse | kTrans);
            finalize_smooth(s, it);
            return s;
        }
        case kGgx: {
            V n = it.ns;
            if (dot(n, it.wo) <= 0) return s;
            float pd, ps;
            ggx_lobes(m, pd, ps);
            if (pd <= 0 && ps <= 0) return s;
            if (pd > 0 && (ux < pd || ps <= 0)) {
                s.wi = cosine_hemisphere(n, ux / pd, uy);
                s.flags = kDiffuse;
            } else {
                BsdfSample sp = sample_ggx_spec(m, it, (ux - pd) / ps, uy);
                if (sp.pdf <= 0) return s;
                s.wi = sp.wi;
                s.flags = kGlossy;
            }
            s.f = eval_ggx(m, it, s.wi);
            s.pdf = pdf_ggx(m, it, s.wi);
            finalize_smooth(s, it);
            return s;
        }
        case kDielectric: {
            V wo = it.wo, n = it.ns;
            float c = dot(wo, n);
            bool entering = c > 0;
            float ei = entering ? 1.0f : m.eta;
            float et = entering ? m.eta : 1.0f;
            V nn = entering ? n : -n;
            c = fabsf(c);
            float F = fresnel(c, ei, et);
            s.measure = kDiscrete;
            if (ux < F) {
                s.wi = normalize(2.0f * c * nn - wo);
                s.weight = m.albedo;
                s.pdf = F;
                s.eta = 1.0f;
                s.flags = kDeltaRefl;
            } else {
                float st = (ei / et) * sqrtf(fmaxf(0.0f, 1.0f - c * c));
                if (st >= 1.0f) {
                    s.wi = normalize(2.0f * c * nn - wo);
                    s.weight = m.albedo;
                    s.pdf = 1.0f;
                    s.eta = 1.0f;
                    s.flags = kDeltaRefl;
                    return s;
                }
                float ct = sqrtf(fmaxf(0.0f, 1.0f - st * st));
                s.wi = normalize((ei / et) * (-wo) + ((ei / et) * c - ct) * nn);
                float etap = et / ei;
                s.weight = m.albedo / (etap * etap);
                s.pdf = 1.0f - F;
                s.eta = etap;
                s.flags = kDeltaTrans | kTrans;
            }
            return s;
        }
        case kInterface:
        case kTransparent:
            s.wi = normalize(-it.wo);
            s.weight = m.type == kTransparent ? m.albedo : V(1);
            s.pdf = 1.0f;
            s.measure = kDiscrete;
            s.flags = kDeltaTrans | kTrans;
            return s;
    }
    return s;
}

// power_heuristic (src/sampler/sample_record.h:49-56)
inline float power_heuristic(float f, float g) {
    if (f <= 0) return 0;
    float a = f * f;
    float b = fmaxf(0.0f, g);
    b *= b;
    float d = a + b;
    return d > 0 ? a / d : 0;
}

// ---- scene ---------------------------------------------------------------------
struct Camera {
    V origin, lower_left, horizontal, vertical, lens_u, lens_v;
    float lens_radius = 0, focal = 1;
};

struct Light {
    int prim;
    V emission;
    bool two_sided;
    float area;
};

struct AliasEntry {
    float prob = 1;
    int alias = 0;
    float pmf = 0;
};

struct Scene {
    std::vector<orc::Tri> tris;
    std::vector<V> n0, n1, n2;  // per-triangle vertex normals after the instance transform
    std::vector<int> mat;
    std::vector<Material> mats;
    Camera cam;
    std::vector<Light> lights;
    std::vector<float> cdf;
    std::vector<AliasEntry> alias;
    float total_weight = 0;
    std::vector<int> light_of_prim;
    orc::Tree* bvh = nullptr;  // one-mesh scenes
    orc_tl* tl = nullptr;      // instanced scenes: tris/n0/n1/n2/mat are the world-space view per ordinal
    ~Scene() {
        delete bvh;
        if (tl) orc_tl_free(tl);
    }
};

inline V tri_v(const orc::Tri& t, int k) {
    const orc::V3& p = k == 0 ? t.v0 : (k == 1 ? t.v1 : t.v2);
    return V(p.x, p.y, p.z);
}
// triangle_normal / triangle_area (src/geometry/triangle.h:14-20)
inline V tri_normal(const orc::Tri& t) { return normalize(cross(tri_v(t, 1) - tri_v(t, 0), tri_v(t, 2) - tri_v(t, 0))); }
inline float tri_area(const orc::Tri& t) { return 0.5f * length(cross(tri_v(t, 1) - tri_v(t, 0), tri_v(t, 2) - tri_v(t, 0))); }

// transform_normal (src/core/mat4.cuh:209-230): the adjugate of the upper 3x3
// applied to n, then normalised; transform_instance_normal
// (src/scene/scene_data.h:262-268) normalises once more and maps a zero normal
// (before or after) to zero.
using M4 = float[4][4];
constexpr M4 kIdentity = {{1, 0, 0, 0}, {0, 1, 0, 0}, {0, 0, 1, 0}, {0, 0, 0, 1}};
inline V transform_normal(const M4& m, const V& n) {
    float a00 = m[1][1] * m[2][2] - m[1][2] * m[2][1];
    float a01 = m[1][2] * m[2][0] - m[1][0] * m[2][2];
    float a02 = m[1][0] * m[2][1] - m[1][1] * m[2][0];
    float a10 = m[0][2] * m[2][1] - m[0][1] * m[2][2];
    float a11 = m[0][0] * m[2][2] - m[0][2] * m[2][0];
    float a12 = m[0][1] * m[2][0] - m[0][0] * m[2][1];
    float a20 = m[0][1] * m[1][2] - m[0][2] * m[1][1];
    float a21 = m[0][2] * m[1][0] - m[0][0] * m[1][2];
    float a22 = m[0][0] * m[1][1] - m[0][1] * m[1][0];
    return normalize(V(a00 * n.x + a01 * n.y + a02 * n.z, a10 * n.x + a11 * n.y + a12 * n.z,
                       a20 * n.x + a21 * n.y + a22 * n.z));
}
inline V instance_normal(const M4& m, const V& n) {
    if (length2(n) == 0) return V(0);
    V t = transform_normal(m, n);
    if (length2(t) == 0) return V(0);
    return normalize(t);
}
// transform_point (src/core/mat4.cuh:173-186): divides by w unless w == 1.
inline orc::V3 transform_point(const M4& m, float px, float py, float pz) {
    float x = m[0][0] * px + m[0][1] * py + m[0][2] * pz + m[0][3];
    float y = m[1][0] * px + m[1][1] * py + m[1][2] * pz + m[1][3];
    float z = m[2][0] * px + m[2][1] * py + m[2][2] * pz + m[2][3];
    float w = m[3][0] * px + m[3][1] * py + m[3][2] * pz + m[3][3];
    return w == 1.0f ? orc::v3(x, y, z) : orc::v3(x / w, y / w, z / w);
}

// make_perspective_camera with lens radius 0 (src/camera/camera.h:58-109)
inline Camera make_camera(const V& origin, const V& look, float vfov, float aspect) {
    auto safe_norm = [](const V& v) {
        float len = length(v);
        return len > 0 ? v / len : v;
    };
    V forward = safe_norm(look - origin);
    V right = safe_norm(cross(V(0, 1, 0), forward));
    if (length2(right) == 0) right = safe_norm(cross(V(0, 0, 1), forward));
    V up = cross(forward, right);
    float theta = vfov * kPi / 180.0f;
    float vh = 2.0f * tanf(theta * 0.5f);
    float vw = aspect * vh;
    float focus = 1.0f;  // focal_distance 1 > 0
    Camera c;
    c.origin = origin;
    c.horizontal = (vw * focus) * right;
    c.vertical = (vh * focus) * up;
    c.lower_left = origin + focus * forward - c.horizontal * 0.5f - c.vertical * 0.5f;
    c.lens_u = right;
    c.lens_v = up;
    c.lens_radius = 0;
    c.focal = focus;
    return c;
}

// build_alias_table_from_weights (src/light/light_distribution.cpp:8-64)
inline std::vector<AliasEntry> alias_table(const std::vector<float>& w) {
    int n = int(w.size());
    std::vector<AliasEntry> t(n);
    if (n <= 0) return t;
    float total = 0;
    for (float x : w) total += fmaxf(0.0f, x);
    if (total <= 0) return {};
    std::vector<float> scaled(n, 0);
    std::vector<int> small, large;
    for (int i = 0; i < n; ++i) {
        float p = fmaxf(0.0f, w[i]) / total;
        t[i].pmf = p;
        t[i].alias = i;
        scaled[i] = p * float(n);
        (scaled[i] < 1.0f ? small : large).push_back(i);
    }
    while (!small.empty() && !large.empty()) {
        int l = small.back();
        small.pop_back();
        int g = large.back();
        large.pop_back();
        t[l].prob = scaled[l];
        t[l].alias = g;
        scaled[g] = (scaled[g] + scaled[l]) - 1.0f;
        (scaled[g] < 1.0f ? small : large).push_back(g);
    }
    for (int i : large) t[i] = AliasEntry{1.0f, i, t[i].pmf};
    for (int i : small) t[i] = AliasEntry{1.0f, i, t[i].pmf};
    return t;
}

// prepare_scene_light_sampling, power selection only (src/scene/scene_prepare.h:268-387)
inline void build_lights(Scene& s) {
    int n = int(s.tris.size());
    s.light_of_prim.assign(n, -1);
    for (int t = 0; t < n; ++t) {
        const Material& m = s.mats[s.mat[t]];
        if (!m.emissive()) continue;
        float area = tri_area(s.tris[t]);
        if (area <= 0) continue;
        float power = area * fmaxf(1e-6f, luminance(m.emission));
        if (power <= 0) continue;
        s.total_weight += power;
        s.lights.push_back(Light{t, m.emission, m.two_sided, area});
        s.cdf.push_back(s.total_weight);
        int& stored = s.light_of_prim[t];
        stored = stored == -1 ? int(s.lights.size()) - 1 : -2;
    }
    if (!s.cdf.empty() && s.total_weight > 0) {
        std::vector<float> w;
        float prev = 0;
        for (float c : s.cdf) {
            w.push_back(c - prev);
            prev = c;
        }
        s.alias = alias_table(w);
    }
}

struct Hit {
    int prim = -1;
    float t = INFINITY, u = 0, v = 0;
    V p, ng, ns;
};

// Closest hit through the host LBVH (intersect_scene_bvh / _skip_ref with the
// identity instance): the host stack walk, skipping `ignore` at its leaf.
inline bool finish_hit(const Scene& s, const V& o, const V& d, int id, float th, float uh, float vh, Hit& h);

inline bool intersect(const Scene& s, const V& o, const V& d, float t_min, float t_max, int ignore, Hit& h) {
    orc::Ray r{orc::v3(o.x, o.y, o.z), orc::v3(d.x, d.y, d.z), t_min, t_max, 0};
    if (s.tl) {  // the two-level walk (scene_intersection.h:147-240), hit finished on the world view
        int32_t prim = -1, inst = -1, local = -1;
        float tt = 0, uu = 0, vv = 0;
        orc_tl_trace(s.tl, reinterpret_cast<const float*>(&r), &ignore, 1, &prim, &tt, &uu, &vv, &inst, &local);
        if (prim < 0) return false;
        return finish_hit(s, o, d, prim, tt, uu, vv, h);
    }
    const orc::Tree& T = *s.bvh;
    if (T.n <= 0) return false;
    int ni = T.n - 1;
    int stack[64];
    int sp = 0;
    stack[sp++] = 0;
    int id = -1;
    float th = r.t_max, uh = 0, vh = 0;
    while (sp > 0) {
        int k = stack[--sp];
        if (!orc::slab(T.bounds[k], r)) continue;
        if (k >= ni) {
            int prim = T.leaf_ids[k - ni];
            if (prim == ignore) continue;
            float tt, uu, vv;
            if (!orc::tri_hit(r, s.tris[prim], tt, uu, vv)) continue;
            if (tt >= th) continue;
            th = tt;
            uh = uu;
            vh = vv;
            id = prim;
            r.t_max = tt;
            continue;
        }
        if (sp + 2 > 64) return false;  // stack overflow: reported as a miss
        stack[sp++] = orc::host_index(T.left[k], ni);
        stack[sp++] = orc::host_index(T.right[k], ni);
    }
    if (id < 0) return false;
    return finish_hit(s, o, d, id, th, uh, vh, h);
}

inline bool finish_hit(const Scene& s, const V& o, const V& d, int id, float th, float uh, float vh, Hit& h) {
    h.prim = id;
    h.t = th;
    h.u = uh;
    h.v = vh;
    // finish_surface_hit_from_triangle (src/scene/scene_intersection.h:54-78)
    h.p = o + d * th;
    const orc::Tri& tri = s.tris[id];
    h.ng = tri_normal(tri);
    const V &a = s.n0[id], &b = s.n1[id], &c = s.n2[id];
    if (length2(a) > 0 && length2(b) > 0 && length2(c) > 0) {
        float w = 1.0f - uh - vh;
        V n = w * a + uh * b + vh * c;
        h.ns = length2(n) == 0 ? h.ng : normalize(n);
    } else {
        h.ns = h.ng;
    }
    if (dot(h.ng, h.ns) < 0) h.ns = -h.ns;
    return true;
}

// pdf_solid_angle (src/light/light_sampler.h:374-384) for an area-light point
inline float pdf_solid_angle(const V& lp, const V& ln, bool two_sided, float pdf_area, const V& ref) {
    V to = lp - ref;
    float d2 = length2(to);
    if (d2 <= kEps * kEps || pdf_area <= 0) return 0;
    V wi = to / sqrtf(d2);
    float cl = emission_cos(ln, -wi, two_sided);
    if (cl <= 0) return 0;
    return pdf_area * d2 / cl;
}

struct LightLi {
    V wi, p, n, radiance;
    float distance = INFINITY, pdf = 0;
    bool valid = false;
};

// sample_light_li, power alias selection + area sampling (src/scene/scene_light_sampler.h:210-250,
// src/light/light_sampler.h:71-89, :360-372)
inline LightLi sample_light(const Scene& s, const V& ref, float sel_u, float /*comp_u*/, float u0, float u1) {
    LightLi li;
    int count = int(s.alias.size());
    if (s.lights.empty() || s.total_weight <= 0 || count <= 0) return li;
    float scaled = fminf(sel_u, 0.99999994f) * float(count);
    int col = int(scaled);
    if (col >= count) col = count - 1;
    float thr = scaled - float(col);
    int idx = thr < s.alias[col].prob ? col : s.alias[col].alias;
    if (idx < 0 || idx >= count) return li;
    float pmf = s.alias[idx].pmf;
    if (!(pmf > 0)) return li;
    const Light& L = s.lights[idx];
    const orc::Tri& tri = s.tris[L.prim];
    float su = sqrtf(u0);
    float b0 = 1.0f - su, b1 = su * (1.0f - u1), b2 = su * u1;
    V p = b0 * tri_v(tri, 0) + b1 * tri_v(tri, 1) + b2 * tri_v(tri, 2);
    V n = tri_normal(tri);
    float pdf_area = L.area > 0 ? 1.0f / L.area : 0.0f;
    V to = p - ref;
    float d2 = length2(to);
    if (d2 <= kEps * kEps) return li;
    li.distance = sqrtf(d2);
    li.wi = to / li.distance;
    li.p = p;
    li.n = n;
    li.radiance = eval_emission(L.emission, L.two_sided, n, -li.wi);
    li.pdf = pmf * pdf_solid_angle(p, n, L.two_sided, pdf_area, ref);
    li.valid = max_component(li.radiance) > 0 && li.pdf > 0;
    return li;
}

// trace_shadow (src/integrator/cpu_path_tracer.cpp:27-72)
inline V shadow_transmittance(const Scene& s, const V& from, const V& to, int skip) {
    V tr(1);
    V dir = to - from;
    float dist = length(dir);
    if (dist <= kEps) return V(0);
    V o = from, d = dir / dist;
    float t_min = kEps, t_max = dist - kEps;
    int ignore = skip;
    for (;;) {
        Hit h;
        if (!intersect(s, o, d, t_min, t_max, ignore, h)) return tr;
        const Material& m = s.mats[s.mat[h.prim]];
        if (!m.passthrough()) return V(0);
        tr = tr * (m.type == kTransparent ? m.albedo : V(1));
        if (max_component(tr) <= 0) return tr;
        float adv = h.t + kEps;
        if (adv >= t_max) return tr;
        o = o + d * adv;
        t_max -= adv;
        t_min = kEps;
        ignore = -1;
    }
}

// trace_path (src/integrator/cpu_path_tracer.cpp:403-498) with estimate_direct_light
// (:115-158) and the transport rules of src/integrator/path_transport.cuh.
inline V trace_path(const Scene& s, V o, V d, Rng& rng, int max_depth) {
    V L(0), T(1);
    float prev_pdf = 0;
    int prev_measure = kNone;
    unsigned prev_flags = 0;
    V prev_pos(0);
    int depth = 0;
    float eta_scale = 1;
    float t_min = kEps;
    for (int step = 0; step <= max_depth; ++step) {
        Hit h;
        if (!intersect(s, o, d, t_min, INFINITY, -1, h)) break;  // no infinite lights
        const Material& m = s.mats[s.mat[h.prim]];
        V wo = -d;
        V Le(0);
        int li = h.prim < int(s.light_of_prim.size()) ? s.light_of_prim[h.prim] : -1;
        if (li >= 0 && li < int(s.lights.size())) {
            const Light& L0 = s.lights[li];
            Le = eval_emission(L0.emission, L0.two_sided, h.ng, wo);
        } else if (m.emissive()) {
            Le = eval_emission(m.emission, m.two_sided, h.ng, wo);
        }
        bool prev_delta = is_delta(prev_flags);
        if (max_component(Le) > 0 && (depth == 0 || prev_delta)) {
            L += T * Le;
        } else if (max_component(Le) > 0 && prev_measure == kSolidAngle && prev_pdf > 0) {
            float lpdf = 0;
            if (li >= 0 && li < int(s.lights.size()) && s.lights[li].area > 0) {
                const Light& L0 = s.lights[li];
                lpdf = s.alias[li].pmf * pdf_solid_angle(h.p, h.ng, L0.two_sided, 1.0f / L0.area, prev_pos);
            }
            float w = power_heuristic(prev_pdf, lpdf);
            L += T * Le * w;
        }
        if (m.passthrough()) {
            T = T * (m.type == kTransparent ? m.albedo : V(1));
            if (max_component(T) <= 0) break;
            o = h.p + d * kEps;
            t_min = kEps;
            continue;
        }
        if (!m.scattering()) break;
        if (step >= max_depth || depth >= max_depth) break;
        Interaction it{h.p, h.ng, dot(h.ng, h.ns) < 0 ? -h.ns : h.ns, wo};
        if (m.direct_lobe()) {
            float sel = rng.get1();
            float comp = rng.get1();
            float lu0 = rng.get1();
            float lu1 = rng.get1();
            LightLi ls = sample_light(s, it.p, sel, comp, lu0, lu1);
            V direct(0);
            if (ls.valid && ls.pdf > 0) {
                V wi = ls.wi;
                V f = eval_bsdf(m, it, wi);
                float cs = abs_cos(it.ns, wi);
                float pdf_b = pdf_bsdf(m, it, wi);
                bool valid = max_component(f) > 0 && max_component(ls.radiance) > 0 && cs > 0 && ls.pdf > 0;
                if (valid) {
                    V contribution = f * ls.radiance * (cs / ls.pdf);
                    V offset_n = dot(wi, it.ng) >= 0 ? it.ng : -it.ng;
                    V so = it.p + offset_n * kEps;
                    V tr = ls.distance < INFINITY ? shadow_transmittance(s, so, ls.p, h.prim) : V(0);
                    if (max_component(tr) > 0) {
                        float mis = power_heuristic(ls.pdf, pdf_b);
                        direct = contribution * tr * mis;
                    }
                }
            }
            L += T * direct;
        }
        float bu0 = rng.get1();
        float bu1 = rng.get1();
        BsdfSample bs = sample_bsdf(m, it, bu0, bu1);
        if (bs.pdf <= 0 || max_component(bs.weight) <= 0) break;
        V nT = T * bs.weight;
        float n_eta = eta_scale * ((!is_transmission(bs.flags) || bs.eta <= 0) ? 1.0f : bs.eta * bs.eta);
        if (depth > 2) {
            float p = fminf(0.95f, max_component(nT) * n_eta);
            if (!(rng.get1() <= p)) break;
            nT = nT / p;
        }
        T = nT;
        eta_scale = n_eta;
        prev_pdf = bs.pdf;
        prev_measure = bs.measure;
        prev_flags = bs.flags;
        prev_pos = h.p;
        V offset_n = dot(bs.wi, it.ng) >= 0 ? it.ng : -it.ng;
        o = h.p + offset_n * kEps;
        d = bs.wi;
        t_min = kEps;
        ++depth;
    }
    return L;
}

// generate_camera_ray for a pinhole (src/camera/camera.h:141-155): 5 dimensions
inline void camera_ray(const Scene& s, int w, int h, int x, int y, Rng& rng, V& o, V& d) {
    float jx = rng.get1();
    float jy = rng.get1();
    float u = (float(x) + jx) / float(w);
    float v = (float(h - 1 - y) + jy) / float(h);
    rng.get1();  // lens u
    rng.get1();  // lens v
    rng.get1();  // time
    const Camera& c = s.cam;
    V target = c.lower_left + u * c.horizontal + v * c.vertical;
    o = c.origin;
    V dir = target - o;
    float len = length(dir);
    d = len > 0 ? dir / len : dir;
}

}  // namespace pt

extern "C" {

struct orc_pt;

// Materials as 11 floats each: type, albedo rgb, emission rgb, two_sided, roughness, metallic, eta.
static void load_materials(pt::Scene& s, const float* materials, int32_t nm) {
    for (int m = 0; m < nm; ++m) {
        const float* p = materials + 11 * m;
        pt::Material mat;
        mat.type = int(p[0]);
        mat.albedo = pt::V(p[1], p[2], p[3]);
        mat.emission = pt::V(p[4], p[5], p[6]);
        mat.two_sided = p[7] != 0;
        mat.roughness = p[8];
        mat.metallic = p[9];
        mat.eta = p[10];
        s.mats.push_back(mat);
    }
}

orc_pt* orc_pt_create(const float* vertices, int32_t nv, const uint32_t* indices, const float* normals,
                      const int32_t* mat_ids, int32_t nt, const float* materials, int32_t nm, const float* camera) {
    (void)nv;
    pt::Scene* s = new pt::Scene();
    load_materials(*s, materials, nm);
    s->tris.resize(nt);
    s->n0.resize(nt);
    s->n1.resize(nt);
    s->n2.resize(nt);
    s->mat.resize(nt);
    auto vert = [&](uint32_t k) { return orc::v3(vertices[3 * k], vertices[3 * k + 1], vertices[3 * k + 2]); };
    auto nrm = [&](uint32_t k) {
        if (!normals) return pt::V(0);
        return pt::instance_normal(pt::kIdentity, pt::V(normals[3 * k], normals[3 * k + 1], normals[3 * k + 2]));
    };
    std::vector<orc::Box> boxes(nt);
    std::vector<int32_t> ids(nt);
    for (int t = 0; t < nt; ++t) {
        uint32_t a = indices[3 * t], b = indices[3 * t + 1], c = indices[3 * t + 2];
        s->tris[t] = orc::Tri{vert(a), vert(b), vert(c)};
        s->n0[t] = nrm(a);
        s->n1[t] = nrm(b);
        s->n2[t] = nrm(c);
        s->mat[t] = mat_ids[t];
        orc::Box bx = orc::empty_box();
        for (const orc::V3& p : {s->tris[t].v0, s->tris[t].v1, s->tris[t].v2}) {
            bx.lo = orc::vmin(bx.lo, p);
            bx.hi = orc::vmax(bx.hi, p);
        }
        boxes[t] = bx;
        ids[t] = t;
    }
    s->bvh = orc::build(boxes.data(), ids.data(), nt);
    s->cam = pt::make_camera(pt::V(camera[0], camera[1], camera[2]), pt::V(camera[3], camera[4], camera[5]),
                             camera[6], camera[7]);
    pt::build_lights(*s);
    return reinterpret_cast<orc_pt*>(s);
}

// An instanced scene (add_scene_mesh per mesh, add_scene_instance per instance,
// src/scene/scene_data.h): meshes concatenated (mesh m owns vertices
// [vertex_offsets[m], vertex_offsets[m+1]) and triangles [triangle_offsets[m],
// triangle_offsets[m+1]), indices mesh-local), instance i places mesh
// inst_mesh[i] by the row-major object-to-world inst_m[i] (inverse inst_inv[i],
// or inverse_affine of inst_m[i] when inst_inv is null) with material override
// inst_override[i] (-1 = the mesh's own). Shading, lights and the light
// sampler see scene_triangle_view per ordinal (src/scene/scene_data.h:350-371):
// world corners, transform_instance_normal corner normals, the material after
// the override; rays take the two-level walk (BLAS per mesh, TLAS over
// transform_bounds of each mesh's box, src/scene/prepared_scene_cpu.cpp:27-42).
orc_pt* orc_pt_create_instanced(const float* vertices, const int32_t* vertex_offsets, const uint32_t* indices,
                                const int32_t* triangle_offsets, const float* normals, const int32_t* mat_ids,
                                int32_t mesh_count, const int32_t* inst_mesh, const float* inst_m,
                                const float* inst_inv, const int32_t* inst_override, int32_t instance_count,
                                const float* materials, int32_t nm, const float* camera) {
    pt::Scene* s = new pt::Scene();
    load_materials(*s, materials, nm);
    auto vtx = [&](int mesh, uint32_t k) { return vertices + 3 * (size_t(vertex_offsets[mesh]) + k); };
    // object-space soups and mesh boxes
    int total = triangle_offsets[mesh_count];
    std::vector<orc::Tri> obj(total);
    std::vector<orc::Box> mesh_box(mesh_count, orc::empty_box());
    for (int m = 0; m < mesh_count; ++m)
        for (int t = triangle_offsets[m]; t < triangle_offsets[m + 1]; ++t) {
            orc::V3 c[3];
            for (int k = 0; k < 3; ++k) {
                const float* p = vtx(m, indices[3 * t + k]);
                c[k] = orc::v3(p[0], p[1], p[2]);
                mesh_box[m].lo = orc::vmin(mesh_box[m].lo, c[k]);
                mesh_box[m].hi = orc::vmax(mesh_box[m].hi, c[k]);
            }
            obj[t] = orc::Tri{c[0], c[1], c[2]};
        }
    std::vector<float> inv(16 * size_t(instance_count)), wb(6 * size_t(instance_count));
    for (int i = 0; i < instance_count; ++i) {
        const float* m16 = inst_m + 16 * i;
        if (inst_inv)
            std::memcpy(&inv[16 * i], inst_inv + 16 * i, 16 * sizeof(float));
        else
            orc_inverse_affine(m16, &inv[16 * i]);
        int mesh = inst_mesh[i];
        orc::Box b = mesh >= 0 && mesh < mesh_count ? mesh_box[mesh] : orc::empty_box();
        orc_transform_bounds(m16, reinterpret_cast<const float*>(&b), &wb[6 * i]);
        if (mesh < 0 || mesh >= mesh_count) continue;
        pt::M4 M;
        std::memcpy(M, m16, sizeof(M));
        for (int t = triangle_offsets[mesh]; t < triangle_offsets[mesh + 1]; ++t) {
            orc::V3 c[3];
            pt::V n[3];
            for (int k = 0; k < 3; ++k) {
                uint32_t vi = indices[3 * t + k];
                const float* p = vtx(mesh, vi);
                c[k] = pt::transform_point(M, p[0], p[1], p[2]);
                if (normals) {
                    const float* q = normals + 3 * (size_t(vertex_offsets[mesh]) + vi);
                    n[k] = pt::instance_normal(M, pt::V(q[0], q[1], q[2]));
                }
            }
            s->tris.push_back(orc::Tri{c[0], c[1], c[2]});
            s->n0.push_back(n[0]);
            s->n1.push_back(n[1]);
            s->n2.push_back(n[2]);
            s->mat.push_back(inst_override[i] >= 0 ? inst_override[i] : mat_ids[t]);
        }
    }
    s->tl = orc_tl_build(reinterpret_cast<const float*>(obj.data()), triangle_offsets, mesh_count, inst_mesh,
                         inv.data(), wb.data(), instance_count);
    s->cam = pt::make_camera(pt::V(camera[0], camera[1], camera[2]), pt::V(camera[3], camera[4], camera[5]),
                             camera[6], camera[7]);
    pt::build_lights(*s);
    return reinterpret_cast<orc_pt*>(s);
}

int32_t orc_pt_triangle_count(const orc_pt* h) { return int32_t(reinterpret_cast<const pt::Scene*>(h)->tris.size()); }

void orc_pt_free(orc_pt* h) { delete reinterpret_cast<pt::Scene*>(h); }

void orc_pt_camera(const orc_pt* h, float* out) {
    const pt::Camera& c = reinterpret_cast<const pt::Scene*>(h)->cam;
    const pt::V* v[6] = {&c.origin, &c.lower_left, &c.horizontal, &c.vertical, &c.lens_u, &c.lens_v};
    for (int k = 0; k < 6; ++k) {
        out[3 * k] = v[k]->x;
        out[3 * k + 1] = v[k]->y;
        out[3 * k + 2] = v[k]->z;
    }
    out[18] = c.lens_radius;
    out[19] = c.focal;
}

int32_t orc_pt_lights(const orc_pt* h, int32_t* prim, float* area, float* cdf, float* prob, int32_t* alias,
                      float* pmf, int32_t* light_of_prim, float* total) {
    const pt::Scene& s = *reinterpret_cast<const pt::Scene*>(h);
    int n = int(s.lights.size());
    if (!prim) return n;
    for (int i = 0; i < n; ++i) {
        prim[i] = s.lights[i].prim;
        area[i] = s.lights[i].area;
        cdf[i] = s.cdf[i];
        prob[i] = s.alias[i].prob;
        alias[i] = s.alias[i].alias;
        pmf[i] = s.alias[i].pmf;
    }
    for (size_t t = 0; t < s.light_of_prim.size(); ++t) light_of_prim[t] = s.light_of_prim[t];
    *total = s.total_weight;
    return n;
}

// Film sums (row-major, y = 0 at the top), samples accumulated per pixel in
// sample order; rows split over `threads` host threads (the per-pixel order is
// unchanged, so the image is identical for any thread count).
void orc_pt_render(const orc_pt* h, int32_t w, int32_t ht, int32_t spp, int32_t depth, uint64_t seed, int32_t frame,
                   int32_t sample_offset, int32_t row_begin, int32_t row_end, int32_t threads, float* film) {
    const pt::Scene& s = *reinterpret_cast<const pt::Scene*>(h);
    auto rows = [&](int y0, int y1) {
        for (int y = y0; y < y1; ++y)
            for (int x = 0; x < w; ++x) {
                int pixel = y * w + x;
                pt::V c(0);
                for (int k = 0; k < spp; ++k) {
                    pt::Rng rng = pt::make_rng(seed, frame, pixel, k + sample_offset);
                    pt::V o, d;
                    pt::camera_ray(s, w, ht, x, y, rng, o, d);
                    c += pt::trace_path(s, o, d, rng, depth);
                }
                film[3 * pixel] = c.x;
                film[3 * pixel + 1] = c.y;
                film[3 * pixel + 2] = c.z;
            }
    };
    if (threads <= 1) {
        rows(row_begin, row_end);
        return;
    }
    std::vector<std::thread> pool;
    int total = row_end - row_begin;
    for (int t = 0; t < threads; ++t) {
        int a = row_begin + total * t / threads, b = row_begin + total * (t + 1) / threads;
        pool.emplace_back(rows, a, b);
    }
    for (auto& th : pool) th.join();
}

void orc_pt_samples(const orc_pt* h, int32_t w, int32_t ht, int32_t depth, uint64_t seed, int32_t frame,
                    int32_t sample_offset, const int32_t* pixels, const int32_t* samples, int32_t n, float* out) {
    const pt::Scene& s = *reinterpret_cast<const pt::Scene*>(h);
    for (int i = 0; i < n; ++i) {
        int pixel = pixels[i];
        pt::Rng rng = pt::make_rng(seed, frame, pixel, samples[i] + sample_offset);
        pt::V o, d;
        pt::camera_ray(s, w, ht, pixel % w, pixel / w, rng, o, d);
        pt::V c = pt::trace_path(s, o, d, rng, depth);
        out[3 * i] = c.x;
        out[3 * i + 1] = c.y;
        out[3 * i + 2] = c.z;
    }
}

}  // extern "C"

extern "C" {
// First n outputs of PCG32(seed, seq) (src/core/pcg.cuh:16-22 seeding).
void orc_pcg32(uint64_t seed, uint64_t seq, int32_t n, uint32_t* out) {
    pt::Rng r;
    r.state = 0;
    r.inc = (seq << 1u) | 1u;
    r.next();
    r.state += seed;
    r.next();
    for (int i = 0; i < n; ++i) out[i] = r.next();
}
}  // extern "C"

// obj_import.hip -- Wavefront OBJ mesh import (host code) for the path
// tracer's scenes: the geometry that load_obj_into_scene
// (src/util/obj_loader.cpp:613-1311) appends to a Scene, from the reference's
// parse of the file (its bundled tinyobjloader's LoadObj with triangulation,
// src/util/tiny_obj_loader.h:8844-9220) and its import rules. The result is one
// mesh (object space) and the instance transform the loader would create; the
// device side starts when that mesh enters an instanced scene (one BLAS per
// mesh). Parsing is a one-off host pass over a text file, so it stays on the
// CPU.
//
// Parsed: v (x y z, optional colour ignored), vn, vt (u v, optional w
// ignored), f with i, i/j, i//k, i/j/k and negative (relative) indices, l and
// p (indices checked, primitives dropped), usemtl, mtllib (ignored), o, g, s,
// comments, \n and \r\n line ends. Triangulation as the parser's: quads split
// along the shorter diagonal, larger polygons by ear clipping in the plane of
// the two dominant axes. Boundary holes are capped as the loader caps them
// (src/util/obj_loader.cpp:1119-1254) when asked.
#include <algorithm>
#include <cerrno>
#include <charconv>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <limits>
#include <sstream>
#include <string>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "lux_runtime.hpp"
#include "two_level.hpp"

namespace lux {
namespace obj {
namespace {

struct Idx {
    int v = -1, vt = -1, vn = -1;
};

struct Parsed {
    std::vector<float> v, vn, vt;  // 3, 3, 2 floats per record
    std::vector<Idx> tri;          // triangulated faces, 3 per triangle, file order
};

// ---- numbers ----------------------------------------------------------------------
char lower(char c) { return (c >= 'A' && c <= 'Z') ? char(c + 32) : c; }

const char* match_iprefix(const char* p, const char* end, const char* prefix) {
    for (; *prefix; ++prefix, ++p)
        if (p == end || lower(*p) != lower(*prefix)) return nullptr;
    return p;
}

// OBJ spellings of NaN (read as 0) and infinity (read as +-DBL_MAX).
bool parse_nan_inf(const char* s, const char* e, double* out) {
    const char* p = s;
    bool negative = false;
    if (p < e && (*p == '-' || *p == '+')) negative = *p++ == '-';
    if (p >= e) return false;
    if (match_iprefix(p, e, "nan")) {
        *out = 0.0;
        return true;
    }
    if (match_iprefix(p, e, "infinity") || match_iprefix(p, e, "inf")) {
        *out = negative ? std::numeric_limits<double>::lowest() : std::numeric_limits<double>::max();
        return true;
    }
    return false;
}

// The longest decimal prefix of [s, e), correctly rounded to double (the
// parser's tryParseDouble: a leading '+' is allowed, hexadecimal is not).
bool parse_double(const char* s, const char* e, double* out) {
    if (s >= e) return false;
    const char* p = s;
    if (*p == '+' || *p == '-') ++p;
    if (p < e && (lower(*p) == 'n' || lower(*p) == 'i') && parse_nan_inf(s, e, out)) return true;
    const char* q = s;
    if (*q == '+') {
        ++q;
        if (q < e && (*q == '+' || *q == '-')) return false;
    }
    double v = 0;
    const auto r = std::from_chars(q, e, v, std::chars_format::general);
    if (r.ec != std::errc()) return false;
    *out = v;
    return true;
}

// ---- line reader (the parser's StreamReader rules) ----------------------------------
struct Reader {
    const char* p;
    const char* end;
    bool eof() const { return p >= end; }
    char peek(size_t k = 0) const { return p + k < end ? p[k] : '\0'; }
    void skip_space() {
        while (p < end && (*p == ' ' || *p == '\t')) ++p;
    }
    void skip_space_and_cr() {
        while (p < end && (*p == ' ' || *p == '\t' || *p == '\r')) ++p;
    }
    bool at_line_end() const { return p >= end || *p == '\n' || *p == '\r' || *p == '\0'; }
    void skip_line() {
        while (p < end) {
            const char c = *p++;
            if (c == '\n') return;
            if (c == '\r') {
                if (p < end && *p == '\n') ++p;
                return;
            }
        }
    }
    size_t token_length() const {
        size_t n = 0;
        while (p + n < end) {
            const char c = p[n];
            if (c == ' ' || c == '\t' || c == '\r' || c == '\n' || c == '\0') break;
            ++n;
        }
        return n;
    }
    // A required-or-defaulted real: an empty token gives the default; a token
    // that is not a number fails the file.
    bool real(float* out) {
        skip_space();
        const size_t n = token_length();
        if (n == 0) {
            *out = 0.0f;
            return true;
        }
        double v = 0;
        const bool ok = parse_double(p, p + n, &v);
        *out = ok ? float(v) : 0.0f;
        p += n;
        return ok;
    }
    // [+-]digits, strtol semantics; anything unparsable reads as 0.
    int int_no_skip() {
        size_t n = 0;
        if (p + n < end && (p[n] == '+' || p[n] == '-')) ++n;
        while (p + n < end && p[n] >= '0' && p[n] <= '9') ++n;
        int value = 0;
        if (n > 0 && n <= 63) {
            char tmp[64];
            std::memcpy(tmp, p, n);
            tmp[n] = '\0';
            errno = 0;
            char* stop = nullptr;
            const long v = std::strtol(tmp, &stop, 10);
            if (errno == 0 && stop != tmp && *stop == '\0' && v <= INT_MAX && v >= INT_MIN) value = int(v);
        }
        p += n;
        return value;
    }
    void skip_until_field_end() {
        while (p < end && *p != '/' && *p != ' ' && *p != '\t' && *p != '\r' && *p != '\n') ++p;
    }
};

// OBJ index -> 0-based: positive i -> i - 1, negative -> count + i (relative),
// zero -> -1 (allowed for texcoord / normal slots only).
bool fix_index(int idx, int count, int* out, bool allow_zero) {
    if (idx > 0) {
        *out = idx - 1;
        return true;
    }
    if (idx == 0) {
        *out = -1;
        return allow_zero;
    }
    *out = count + idx;
    return *out >= 0;
}

bool parse_triple(Reader& r, int nv, int nvn, int nvt, Idx* out) {
    Idx t;
    r.skip_space();
    if (!fix_index(r.int_no_skip(), nv, &t.v, false)) return false;
    r.skip_until_field_end();
    if (r.eof() || r.peek() != '/') {
        *out = t;
        return true;
    }
    ++r.p;
    if (!r.eof() && r.peek() == '/') {  // i//k
        ++r.p;
        if (!fix_index(r.int_no_skip(), nvn, &t.vn, true)) return false;
        r.skip_until_field_end();
        *out = t;
        return true;
    }
    if (!fix_index(r.int_no_skip(), nvt, &t.vt, true)) return false;  // i/j or i/j/k
    r.skip_until_field_end();
    if (r.eof() || r.peek() != '/') {
        *out = t;
        return true;
    }
    ++r.p;
    if (!fix_index(r.int_no_skip(), nvn, &t.vn, true)) return false;
    r.skip_until_field_end();
    *out = t;
    return true;
}

bool keyword(const Reader& r, const char* word) {
    const size_t n = std::strlen(word);
    for (size_t k = 0; k < n; ++k)
        if (r.peek(k) != word[k]) return false;
    return r.peek(n) == ' ' || r.peek(n) == '\t';
}

// ---- triangulation --------------------------------------------------------------------
// Point-in-triangle parity test (W. Randolph Franklin's pnpoly) on the 2D projection.
bool point_in_triangle(const float* vx, const float* vy, float tx, float ty) {
    bool c = false;
    for (int i = 0, j = 2; i < 3; j = i++)
        if (((vy[i] > ty) != (vy[j] > ty)) && (tx < (vx[j] - vx[i]) * (ty - vy[i]) / (vy[j] - vy[i]) + vx[i]))
            c = !c;
    return c;
}

void emit(Parsed& out, const Idx& a, const Idx& b, const Idx& c) {
    out.tri.push_back(a);
    out.tri.push_back(b);
    out.tri.push_back(c);
}

// A face into triangles, as the parser's exportGroupsToShape with
// triangulation (src/util/tiny_obj_loader.h:7647-8140).
void triangulate(Parsed& out, const std::vector<Idx>& face) {
    const std::vector<float>& v = out.v;
    const size_t n = face.size();
    if (n < 3) return;  // degenerate face: dropped
    if (n == 3) {
        emit(out, face[0], face[1], face[2]);
        return;
    }
    if (n == 4) {
        for (const Idx& i : face)
            if (i.v < 0 || 3 * size_t(i.v) + 2 >= v.size()) return;  // invalid index: dropped
        const float* p0 = &v[3 * size_t(face[0].v)];
        const float* p1 = &v[3 * size_t(face[1].v)];
        const float* p2 = &v[3 * size_t(face[2].v)];
        const float* p3 = &v[3 * size_t(face[3].v)];
        const float e02x = p2[0] - p0[0], e02y = p2[1] - p0[1], e02z = p2[2] - p0[2];
        const float e13x = p3[0] - p1[0], e13y = p3[1] - p1[1], e13z = p3[2] - p1[2];
        const float sqr02 = e02x * e02x + e02y * e02y + e02z * e02z;
        const float sqr13 = e13x * e13x + e13y * e13y + e13z * e13z;
        if (sqr02 < sqr13) {
            emit(out, face[0], face[1], face[2]);
            emit(out, face[0], face[2], face[3]);
        } else {
            emit(out, face[0], face[1], face[3]);
            emit(out, face[1], face[2], face[3]);
        }
        return;
    }
    // ear clipping: the projection plane is chosen at the first corner whose
    // turn is not degenerate
    size_t axes[2] = {1, 2};
    for (size_t k = 0; k < n; ++k) {
        const size_t a = size_t(face[k].v), b = size_t(face[(k + 1) % n].v), c = size_t(face[(k + 2) % n].v);
        if (3 * a + 2 >= v.size() || 3 * b + 2 >= v.size() || 3 * c + 2 >= v.size()) continue;
        const float e0x = v[3 * b] - v[3 * a], e0y = v[3 * b + 1] - v[3 * a + 1], e0z = v[3 * b + 2] - v[3 * a + 2];
        const float e1x = v[3 * c] - v[3 * b], e1y = v[3 * c + 1] - v[3 * b + 1], e1z = v[3 * c + 2] - v[3 * b + 2];
        const float cx = std::fabs(e0y * e1z - e0z * e1y);
        const float cy = std::fabs(e0z * e1x - e0x * e1z);
        const float cz = std::fabs(e0x * e1y - e0y * e1x);
        const float eps = std::numeric_limits<float>::epsilon();
        if (cx > eps || cy > eps || cz > eps) {
            if (!(cx > cy && cx > cz)) {
                axes[0] = 0;
                if (cz > cx && cz > cy) axes[1] = 1;
            }
            break;
        }
    }
    std::vector<Idx> rest = face;
    size_t guess = 0;
    size_t iterations = n, previous = rest.size();
    float vx[3], vy[3];
    Idx ind[3];
    while (rest.size() > 3 && iterations > 0) {
        const size_t m = rest.size();
        if (guess >= m) guess -= m;
        if (previous != m) {
            previous = m;
            iterations = m;
        } else {
            --iterations;
        }
        for (int k = 0; k < 3; ++k) {
            ind[k] = rest[(guess + k) % m];
            const size_t vi = size_t(ind[k].v);
            if (vi * 3 + axes[0] >= v.size() || vi * 3 + axes[1] >= v.size()) {
                vx[k] = 0.0f;
                vy[k] = 0.0f;
            } else {
                vx[k] = v[vi * 3 + axes[0]];
                vy[k] = v[vi * 3 + axes[1]];
            }
        }
        const float e0x = vx[1] - vx[0], e0y = vy[1] - vy[0];
        const float e1x = vx[2] - vx[1], e1y = vy[2] - vy[1];
        const float cross = e0x * e1y - e0y * e1x;
        const float area = (vx[0] * vy[1] - vy[0] * vx[1]) * 0.5f;
        if (cross * area < 0.0f) {  // reflex corner
            guess += 1;
            continue;
        }
        bool overlap = false;
        for (size_t other = 3; other < m; ++other) {
            const size_t idx = (guess + other) % m;
            const size_t ovi = size_t(rest[idx].v);
            if (ovi * 3 + axes[0] >= v.size() || ovi * 3 + axes[1] >= v.size()) continue;
            if (point_in_triangle(vx, vy, v[ovi * 3 + axes[0]], v[ovi * 3 + axes[1]])) {
                overlap = true;
                break;
            }
        }
        if (overlap) {
            guess += 1;
            continue;
        }
        emit(out, ind[0], ind[1], ind[2]);  // an ear
        for (size_t k = (guess + 1) % m; k + 1 < m; ++k) rest[k] = rest[k + 1];
        rest.pop_back();
    }
    if (rest.size() == 3) emit(out, rest[0], rest[1], rest[2]);
}

bool parse(const std::string& text, Parsed& out, std::string& error) {
    Reader r{text.data(), text.data() + text.size()};
    if (text.size() >= 3 && (unsigned char)text[0] == 0xEF && (unsigned char)text[1] == 0xBB &&
        (unsigned char)text[2] == 0xBF)
        r.p += 3;
    // Faces are triangulated when their group is closed (an o or g line, or the
    // end of the file), against the vertices read by then, as the parser does.
    std::vector<std::vector<Idx>> pending;
    auto flush = [&] {
        for (const std::vector<Idx>& f : pending) triangulate(out, f);
        pending.clear();
    };
    int line = 0;
    while (!r.eof()) {
        ++line;
        r.skip_space();
        if (r.at_line_end() || r.peek() == '#') {
            r.skip_line();
            continue;
        }
        const int nv = int(out.v.size() / 3), nvn = int(out.vn.size() / 3), nvt = int(out.vt.size() / 2);
        if (keyword(r, "v")) {
            r.p += 2;
            float x, y, z;
            if (!r.real(&x) || !r.real(&y) || !r.real(&z)) {
                error = "line " + std::to_string(line) + ": bad vertex";
                return false;
            }
            out.v.insert(out.v.end(), {x, y, z});
        } else if (keyword(r, "vn")) {
            r.p += 3;
            float x, y, z;
            if (!r.real(&x) || !r.real(&y) || !r.real(&z)) {
                error = "line " + std::to_string(line) + ": bad normal";
                return false;
            }
            out.vn.insert(out.vn.end(), {x, y, z});
        } else if (keyword(r, "vt")) {
            r.p += 3;
            float x, y;
            if (!r.real(&x) || !r.real(&y)) {
                error = "line " + std::to_string(line) + ": bad texcoord";
                return false;
            }
            out.vt.insert(out.vt.end(), {x, y});
        } else if (keyword(r, "f") || keyword(r, "l") || keyword(r, "p")) {
            const bool is_face = r.peek() == 'f';
            r.p += 2;
            r.skip_space();
            std::vector<Idx> face;
            while (!r.at_line_end() && r.peek() != '#') {
                Idx t;
                if (!parse_triple(r, nv, nvn, nvt, &t)) {
                    error = "line " + std::to_string(line) + ": invalid vertex index";
                    return false;
                }
                face.push_back(t);
                r.skip_space_and_cr();
            }
            if (is_face) pending.push_back(std::move(face));
        } else if (keyword(r, "g") || keyword(r, "o")) {
            flush();
        }
        r.skip_line();
    }
    flush();
    return true;
}

// ---- the import contract's pre-scan: one material label per triangle ----------------
std::string trim(const std::string& s) {
    const char* ws = " \t\r\n";
    const size_t a = s.find_first_not_of(ws);
    if (a == std::string::npos) return {};
    return s.substr(a, s.find_last_not_of(ws) - a + 1);
}

bool keyword_line(const std::string& t, const char* word) {
    const size_t n = std::strlen(word);
    if (t.size() < n || t.compare(0, n, word) != 0) return false;
    return t.size() == n || t[n] == ' ' || t[n] == '\t';
}

struct LabelScan {
    std::vector<std::string> triangle_labels;
    std::vector<std::string> names;  // distinct usemtl names in first-use order
};

LabelScan scan_labels(const std::string& text) {
    LabelScan scan;
    std::unordered_set<std::string> seen;
    std::string current;
    std::istringstream in(text);
    std::string line;
    while (std::getline(in, line)) {
        const size_t hash = line.find('#');
        const std::string t = trim(hash == std::string::npos ? line : line.substr(0, hash));
        if (t.empty()) continue;
        if (keyword_line(t, "usemtl")) {
            current = t.size() > 6 ? trim(t.substr(6)) : std::string();
            if (!current.empty() && seen.insert(current).second) scan.names.push_back(current);
            continue;
        }
        if (!keyword_line(t, "f")) continue;
        std::istringstream tokens(t.size() > 1 ? trim(t.substr(1)) : std::string());
        std::string tok;
        int count = 0;
        while (tokens >> tok) ++count;
        for (int k = 0; k < count - 2; ++k) scan.triangle_labels.push_back(current);
    }
    return scan;
}

// ---- vector helpers with the reference's operation order -------------------------------
struct V3 {
    float x, y, z;
};
V3 sub(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
V3 cross(V3 a, V3 b) { return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
V3 normalize(V3 v) {
    const float len = std::sqrt(dot(v, v));
    if (!(len > 0)) return v;
    const float inv = 1 / len;
    return V3{v.x * inv, v.y * inv, v.z * inv};
}
bool finite(V3 v) { return std::isfinite(v.x) && std::isfinite(v.y) && std::isfinite(v.z); }

struct M4 {
    float m[16];
};
M4 identity() {
    M4 r{};
    r.m[0] = r.m[5] = r.m[10] = r.m[15] = 1.0f;
    return r;
}
M4 mul(const M4& a, const M4& b) {  // mat4::operator*
    M4 r;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j)
            r.m[4 * i + j] = a.m[4 * i] * b.m[j] + a.m[4 * i + 1] * b.m[4 + j] + a.m[4 * i + 2] * b.m[8 + j] +
                             a.m[4 * i + 3] * b.m[12 + j];
    return r;
}

}  // namespace

// ---- the import ---------------------------------------------------------------------------
void load(const std::string& path, const lux_obj_options& o, Mesh& out) {
    if (o.normal_mode < 0 || o.normal_mode > 2) throw Error(LUX_ERR_INVALID_ARGUMENT, "unknown normal mode");
    if (o.boundary_topology < 0 || o.boundary_topology > 1)
        throw Error(LUX_ERR_INVALID_ARGUMENT, "unknown boundary topology");
    if (o.placement_scale_mode < 0 || o.placement_scale_mode > 2)
        throw Error(LUX_ERR_INVALID_ARGUMENT, "unknown placement scale mode");
    if (o.binding_count < 0 || (o.binding_count > 0 && (!o.binding_names || !o.binding_material_ids)))
        throw Error(LUX_ERR_INVALID_ARGUMENT, "bad material bindings");
    const int nmat = o.scene_material_count;
    auto valid_material = [&](int id) { return id >= 0 && id < nmat; };
    if (o.default_material_id >= 0 && !valid_material(o.default_material_id))
        throw Error(LUX_ERR_INVALID_ARGUMENT, "default material id is outside the scene's materials");
    std::unordered_map<std::string, int> bindings;
    for (int b = 0; b < o.binding_count; ++b) {
        if (!o.binding_names[b]) throw Error(LUX_ERR_INVALID_ARGUMENT, "null binding name");
        if (!valid_material(o.binding_material_ids[b]))
            throw Error(LUX_ERR_INVALID_ARGUMENT,
                        std::string("material binding '") + o.binding_names[b] + "' is not a scene material");
        bindings[o.binding_names[b]] = o.binding_material_ids[b];
    }
    if (o.placement_scale_mode != 0 && (!std::isfinite(o.placement_target_extent) || o.placement_target_extent <= 0))
        throw Error(LUX_ERR_INVALID_ARGUMENT, "placement target extent must be finite and positive");
    if (o.placement_ground_to_y && !std::isfinite(o.placement_target_y))
        throw Error(LUX_ERR_INVALID_ARGUMENT, "placement target y must be finite");
    if (!instance_transform_valid(o.object_to_world, o.world_to_object))
        throw Error(LUX_ERR_INVALID_ARGUMENT, "object_to_world must be affine-invertible with a consistent inverse");
    if (o.cap_boundary_holes && (o.cap_boundary_topology < 0 || o.cap_boundary_topology > 1))
        throw Error(LUX_ERR_INVALID_ARGUMENT, "unknown cap boundary topology");
    // validate_obj_weld_options (src/util/obj_loader.cpp:356-371)
    const bool welded = o.boundary_topology == 1 || (o.cap_boundary_holes && o.cap_boundary_topology == 1);
    if (welded && (!std::isfinite(o.position_weld_epsilon) || o.position_weld_epsilon <= 0))
        throw Error(LUX_ERR_INVALID_ARGUMENT, "weld epsilon must be finite and positive");

    std::ifstream f(path, std::ios::binary);
    if (!f) throw Error(LUX_ERR_INVALID_ARGUMENT, "cannot open OBJ file " + path);
    std::string text((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    const LabelScan scan = scan_labels(text);

    // material resolution (resolve_default_obj_material_id / validate_obj_material_labels)
    bool append_default = false;
    int default_id = o.default_material_id;
    if (default_id < 0 && bindings.empty()) {
        default_id = nmat;
        append_default = true;
    }
    if (!bindings.empty()) {
        for (const std::string& name : scan.names)
            if (!bindings.count(name)) throw Error(LUX_ERR_INVALID_ARGUMENT, "usemtl '" + name + "' has no binding");
        if (default_id < 0)
            for (const std::string& label : scan.triangle_labels)
                if (label.empty())
                    throw Error(LUX_ERR_INVALID_ARGUMENT, "a face has no usemtl while material bindings are required");
    }

    Parsed p;
    std::string error;
    if (!parse(text, p, error)) throw Error(LUX_ERR_INVALID_ARGUMENT, "OBJ parse failed: " + error);
    const size_t n_tris = p.tri.size() / 3;
    if (n_tris != scan.triangle_labels.size())
        throw Error(LUX_ERR_INVALID_ARGUMENT, "OBJ usemtl label count does not match the parsed triangle count");

    const int nv = int(p.v.size() / 3), nvn = int(p.vn.size() / 3), nvt = int(p.vt.size() / 2);
    auto pos = [&](int i) { return V3{p.v[3 * size_t(i)], p.v[3 * size_t(i) + 1], p.v[3 * size_t(i) + 2]}; };
    auto nrm = [&](int i) { return V3{p.vn[3 * size_t(i)], p.vn[3 * size_t(i) + 1], p.vn[3 * size_t(i) + 2]}; };
    std::vector<uint8_t> v_finite(static_cast<size_t>(nv)), n_finite(static_cast<size_t>(nvn)),
        t_finite(static_cast<size_t>(nvt));
    for (int i = 0; i < nv; ++i) v_finite[i] = finite(pos(i));
    for (int i = 0; i < nvn; ++i) n_finite[i] = finite(nrm(i));
    for (int i = 0; i < nvt; ++i)
        t_finite[i] = std::isfinite(p.vt[2 * size_t(i)]) && std::isfinite(p.vt[2 * size_t(i) + 1]);
    auto normal_nonzero = [&](const Idx& i) {
        return i.vn >= 0 && i.vn < nvn && n_finite[i.vn] && dot(nrm(i.vn), nrm(i.vn)) > 1e-20f;
    };
    auto normal_zero = [&](const Idx& i) {
        return i.vn >= 0 && i.vn < nvn && n_finite[i.vn] && !(dot(nrm(i.vn), nrm(i.vn)) > 1e-20f);
    };

    struct Face {
        Idx idx[3];
        int material;
    };
    std::vector<Face> faces;
    std::vector<V3> smooth(static_cast<size_t>(nv), V3{0, 0, 0});
    Mesh& m = out;
    m = Mesh{};
    for (size_t t = 0; t < n_tris; ++t) {
        const Idx* c = &p.tri[3 * t];
        bool valid = true, fin = true;
        for (int k = 0; k < 3; ++k) {
            valid = valid && c[k].v >= 0 && c[k].v < nv;
            fin = fin && valid && v_finite[c[k].v];
        }
        if (!valid) {
            ++m.skipped_invalid;
            continue;
        }
        if (!fin) {
            ++m.skipped_nonfinite;
            continue;
        }
        const V3 v0 = pos(c[0].v), v1 = pos(c[1].v), v2 = pos(c[2].v);
        const V3 cr = cross(sub(v1, v0), sub(v2, v0));
        if (0.5f * std::sqrt(dot(cr, cr)) <= 1e-12f) {  // triangle_area
            ++m.skipped_degenerate;
            continue;
        }
        int material = default_id;
        const std::string& label = scan.triangle_labels[t];
        if (!bindings.empty() && !label.empty()) {
            auto it = bindings.find(label);
            if (it == bindings.end()) throw Error(LUX_ERR_INVALID_ARGUMENT, "usemtl '" + label + "' has no binding");
            material = it->second;
        } else if (!bindings.empty() && default_id < 0) {
            throw Error(LUX_ERR_INVALID_ARGUMENT, "a face has no usemtl while material bindings are required");
        }
        Face fc;
        fc.idx[0] = c[0];
        fc.idx[1] = o.reverse_orientation ? c[2] : c[1];
        fc.idx[2] = o.reverse_orientation ? c[1] : c[2];
        fc.material = material;
        faces.push_back(fc);
        for (int k = 0; k < 3; ++k) {
            if (c[k].vn >= nvn) ++m.invalid_normal_refs;
            if (c[k].vn >= 0 && c[k].vn < nvn && !n_finite[c[k].vn]) ++m.nonfinite_normal_refs;
            if (normal_zero(c[k])) ++m.zero_normal_refs;
        }
        if (o.normal_mode == 1) {
            for (int k = 0; k < 3; ++k) {
                V3& s = smooth[size_t(c[k].v)];
                s = V3{s.x + cr.x, s.y + cr.y, s.z + cr.z};
            }
        }
    }
    m.nonfinite_vertices = int(std::count(v_finite.begin(), v_finite.end(), uint8_t(0)));
    if (faces.empty()) return;  // nothing appended, no material added
    m.appended_default_material = append_default ? 1 : 0;

    // topology vertices (make_topology_vertices, src/util/obj_loader.cpp:476-519):
    // OBJ vertex ids, or positions welded on a grid of epsilon, each with the
    // position of its first vertex
    auto make_topology = [&](int topology, std::vector<int>& ids, std::vector<V3>& positions) {
        ids.assign(static_cast<size_t>(nv), -1);
        positions.clear();
        if (topology == 0) {
            positions.assign(static_cast<size_t>(nv), V3{0, 0, 0});
            for (int i = 0; i < nv; ++i)
                if (v_finite[i]) {
                    ids[i] = i;
                    positions[size_t(i)] = pos(i);
                }
            return;
        }
        struct Key {
            int64_t x, y, z;
            bool operator==(const Key& k) const { return x == k.x && y == k.y && z == k.z; }
        };
        struct KeyHash {
            size_t operator()(const Key& k) const {
                uint64_t h = uint64_t(k.x);
                h = h * 1099511628211ULL ^ uint64_t(k.y);
                h = h * 1099511628211ULL ^ uint64_t(k.z);
                return size_t(h);
            }
        };
        std::unordered_map<Key, int, KeyHash> welded_ids;
        const double inv_eps = 1.0 / double(o.position_weld_epsilon);
        for (int i = 0; i < nv; ++i) {
            if (!v_finite[i]) continue;
            const V3 q = pos(i);
            const Key k{std::llround(double(q.x) * inv_eps), std::llround(double(q.y) * inv_eps),
                        std::llround(double(q.z) * inv_eps)};
            auto it = welded_ids.find(k);
            if (it != welded_ids.end()) {
                ids[i] = it->second;
            } else {
                welded_ids.emplace(k, int(positions.size()));
                ids[i] = int(positions.size());
                positions.push_back(q);
            }
        }
    };
    auto edge_key = [](int a, int b) {
        return (uint64_t(uint32_t(std::min(a, b))) << 32) | uint32_t(std::max(a, b));
    };
    auto count_edges = [&](const std::vector<int>& ids, std::unordered_map<uint64_t, int>& counts) {
        for (const Face& fc : faces) {
            const int a = ids[fc.idx[0].v], b = ids[fc.idx[1].v], c = ids[fc.idx[2].v];
            if (a != b) ++counts[edge_key(a, b)];
            if (b != c) ++counts[edge_key(b, c)];
            if (c != a) ++counts[edge_key(c, a)];
        }
    };
    auto is_boundary = [&](const std::unordered_map<uint64_t, int>& counts, int a, int b) {
        if (a < 0 || b < 0 || a == b) return false;
        auto it = counts.find(edge_key(a, b));
        return it != counts.end() && it->second == 1;
    };
    // boundary-edge masks over the chosen topology
    std::vector<int> topo;
    std::vector<V3> topo_pos;
    make_topology(o.boundary_topology, topo, topo_pos);
    std::unordered_map<uint64_t, int> edges;
    count_edges(topo, edges);
    // the capping topology and its edge counts
    std::vector<int> cap_topo;
    std::vector<V3> cap_pos;
    std::unordered_map<uint64_t, int> cap_edges;
    if (o.cap_boundary_holes) {
        make_topology(o.cap_boundary_topology, cap_topo, cap_pos);
        count_edges(cap_topo, cap_edges);
        for (const auto& e : cap_edges) m.cap_boundary_edges += e.second == 1 ? 1 : 0;
    }
    struct CapEdge {
        int from, to, material;
    };
    std::vector<CapEdge> cap_edge_list;
    auto boundary = [&](int a, int b) { return is_boundary(edges, a, b); };

    // vertices: one per distinct (position, used normal, used texcoord), first use first
    auto normal_used = [&](const Idx& i) { return o.normal_mode != 2 && normal_nonzero(i); };
    auto texcoord_used = [&](const Idx& i) { return i.vt >= 0 && i.vt < nvt && t_finite[i.vt]; };
    struct VKey {
        int v, n, t;
        bool operator==(const VKey& k) const { return v == k.v && n == k.n && t == k.t; }
    };
    struct VKeyHash {
        size_t operator()(const VKey& k) const {
            uint64_t h = uint32_t(k.v);
            h = h * 1099511628211ULL ^ uint32_t(k.n);
            h = h * 1099511628211ULL ^ uint32_t(k.t);
            return size_t(h);
        }
    };
    std::unordered_map<VKey, uint32_t, VKeyHash> vmap;
    vmap.reserve(faces.size() * 3);
    auto vertex = [&](const Idx& i) -> uint32_t {
        const VKey k{i.v, normal_used(i) ? i.vn : -1, texcoord_used(i) ? i.vt : -1};
        auto it = vmap.find(k);
        if (it != vmap.end()) return it->second;
        const uint32_t id = uint32_t(m.vertices.size() / 3);
        vmap.emplace(k, id);
        const V3 q = pos(i.v);
        m.vertices.insert(m.vertices.end(), {q.x, q.y, q.z});
        V3 n{0, 0, 0};
        if (o.normal_mode != 2) {
            if (normal_nonzero(i)) {
                n = normalize(nrm(i.vn));
                if (o.reverse_orientation) n = V3{-n.x, -n.y, -n.z};
            } else if (o.normal_mode == 1 && i.v >= 0) {
                const V3 s = smooth[size_t(i.v)];
                if (dot(s, s) > 0) {
                    n = normalize(s);
                    if (o.reverse_orientation) n = V3{-n.x, -n.y, -n.z};
                }
            }
        }
        m.normals.insert(m.normals.end(), {n.x, n.y, n.z});
        float u = 0, v = 0;
        if (texcoord_used(i)) {
            u = p.vt[2 * size_t(i.vt)];
            v = p.vt[2 * size_t(i.vt) + 1];
            if (o.flip_texcoord_v) v = 1.0f - v;
        }
        m.uvs.insert(m.uvs.end(), {u, v});
        return id;
    };
    for (const Face& fc : faces) {
        const int a = topo[fc.idx[0].v], b = topo[fc.idx[1].v], c = topo[fc.idx[2].v];
        uint8_t mask = 0;
        if (boundary(a, b)) mask |= 1u;
        if (boundary(b, c)) mask |= 2u;
        if (boundary(c, a)) mask |= 4u;
        const uint32_t i0 = vertex(fc.idx[0]), i1 = vertex(fc.idx[1]), i2 = vertex(fc.idx[2]);
        m.indices.insert(m.indices.end(), {i0, i1, i2});
        m.material_ids.push_back(fc.material);
        m.boundary_edges.push_back(mask);
        if (o.cap_boundary_holes) {
            const int ca[3] = {cap_topo[fc.idx[0].v], cap_topo[fc.idx[1].v], cap_topo[fc.idx[2].v]};
            for (int k = 0; k < 3; ++k)
                if (is_boundary(cap_edges, ca[k], ca[(k + 1) % 3]))
                    cap_edge_list.push_back(CapEdge{ca[k], ca[(k + 1) % 3], fc.material});
        }
    }

    // boundary-hole caps (src/util/obj_loader.cpp:1119-1254): with every
    // boundary vertex on exactly one outgoing and one incoming boundary edge,
    // each closed loop of 3 or more vertices is fanned from its centroid, the
    // fan turned to face away from the mesh centre (then flipped again for
    // reverse_orientation); each cap triangle gets three new vertices with its
    // face normal and uv 0
    if (o.cap_boundary_holes && !cap_edge_list.empty()) {
        // topology_center_for_faces: mean of the distinct topology vertices in first-use order
        V3 mesh_center{0, 0, 0};
        {
            int count = 0;
            std::unordered_set<int> seen;
            for (const Face& fc : faces)
                for (int k = 0; k < 3; ++k) {
                    const int id = cap_topo[fc.idx[k].v];
                    if (id < 0 || id >= int(cap_pos.size()) || !seen.insert(id).second) continue;
                    const V3 q = cap_pos[size_t(id)];
                    mesh_center = V3{mesh_center.x + q.x, mesh_center.y + q.y, mesh_center.z + q.z};
                    ++count;
                }
            if (count > 0) {
                const float inv = 1 / float(count);
                mesh_center = V3{mesh_center.x * inv, mesh_center.y * inv, mesh_center.z * inv};
            } else {
                mesh_center = V3{0, 0, 0};
            }
        }
        std::unordered_map<int, int> outgoing, incoming;
        for (const CapEdge& e : cap_edge_list) {
            ++outgoing[e.from];
            ++incoming[e.to];
        }
        for (const auto& e : outgoing) {
            auto it = incoming.find(e.first);
            if (e.second != 1 || it == incoming.end() || it->second != 1) ++m.cap_non_simple_vertices;
        }
        for (const auto& e : incoming)
            if (outgoing.find(e.first) == outgoing.end()) ++m.cap_non_simple_vertices;
        std::unordered_map<int, CapEdge> next_edge;
        if (m.cap_non_simple_vertices == 0)
            for (const CapEdge& e : cap_edge_list) next_edge.emplace(e.from, e);
        std::unordered_set<int> visited;
        for (const CapEdge& first_edge : cap_edge_list) {
            if (m.cap_non_simple_vertices != 0) break;
            if (visited.count(first_edge.from)) continue;
            std::vector<int> loop;
            int material = first_edge.material;
            const int start = first_edge.from;
            int current = start;
            while (!visited.count(current)) {
                auto it = next_edge.find(current);
                if (it == next_edge.end()) break;
                visited.insert(current);
                loop.push_back(current);
                material = it->second.material;
                current = it->second.to;
                if (current == start) break;
            }
            if (current != start) {
                ++m.cap_incomplete_loops;
                continue;
            }
            if (loop.size() < 3) {
                ++m.cap_degenerate_loops;
                continue;
            }
            V3 center{0, 0, 0};
            for (int id : loop) {
                const V3 q = cap_pos[size_t(id)];
                center = V3{center.x + q.x, center.y + q.y, center.z + q.z};
            }
            {
                const float inv = 1 / float(loop.size());
                center = V3{center.x * inv, center.y * inv, center.z * inv};
            }
            for (size_t i = 0; i < loop.size(); ++i) {
                V3 t0 = center, t1 = cap_pos[size_t(loop[i])], t2 = cap_pos[size_t(loop[(i + 1) % loop.size()])];
                const V3 cr = cross(sub(t1, t0), sub(t2, t0));
                V3 normal = normalize(cr);
                if (0.5f * std::sqrt(dot(cr, cr)) <= 1e-12f) {
                    ++m.cap_degenerate_triangles;
                    continue;
                }
                if (dot(normal, sub(center, mesh_center)) < 0) {
                    std::swap(t1, t2);
                    normal = V3{-normal.x, -normal.y, -normal.z};
                }
                if (o.reverse_orientation) {
                    std::swap(t1, t2);
                    normal = V3{-normal.x, -normal.y, -normal.z};
                }
                const uint32_t first = uint32_t(m.vertices.size() / 3);
                for (const V3& q : {t0, t1, t2}) {
                    m.vertices.insert(m.vertices.end(), {q.x, q.y, q.z});
                    m.normals.insert(m.normals.end(), {normal.x, normal.y, normal.z});
                    m.uvs.insert(m.uvs.end(), {0.0f, 0.0f});
                }
                m.indices.insert(m.indices.end(), {first, first + 1, first + 2});
                m.material_ids.push_back(material);
                m.boundary_edges.push_back(0);
                ++m.cap_triangles;
            }
        }
    }

    // mesh bounds (union of the triangle bounds) and the placement transform
    // (make_obj_placement_transform: translate(target) * scale(s) * translate(-anchor))
    V3 lo{INFINITY, INFINITY, INFINITY}, hi{-INFINITY, -INFINITY, -INFINITY};
    for (size_t k = 0; k < m.indices.size(); ++k) {
        const float* q = &m.vertices[3 * size_t(m.indices[k])];
        lo = V3{std::fmin(lo.x, q[0]), std::fmin(lo.y, q[1]), std::fmin(lo.z, q[2])};
        hi = V3{std::fmax(hi.x, q[0]), std::fmax(hi.y, q[1]), std::fmax(hi.z, q[2])};
    }
    M4 placement = identity(), placement_inv = identity();
    if (!(lo.x > hi.x || lo.y > hi.y || lo.z > hi.z)) {
        float s = 1.0f;
        if (o.placement_scale_mode != 0) {
            const V3 d = sub(hi, lo);
            const float extent = o.placement_scale_mode == 1 ? std::max({d.x, d.y, d.z}) : d.x;
            if (extent > 0) s = o.placement_target_extent / extent;
        }
        V3 anchor{0, 0, 0}, target{0, 0, 0};
        if (o.placement_center_xz) {
            anchor.x = (lo.x + hi.x) * 0.5f;
            anchor.z = (lo.z + hi.z) * 0.5f;
        }
        if (o.placement_ground_to_y) {
            anchor.y = lo.y;
            target.y = o.placement_target_y;
        }
        M4 t1 = identity(), sc = identity(), t0 = identity();
        t1.m[3] = target.x;
        t1.m[7] = target.y;
        t1.m[11] = target.z;
        sc.m[0] = sc.m[5] = sc.m[10] = s;
        t0.m[3] = -anchor.x;
        t0.m[7] = -anchor.y;
        t0.m[11] = -anchor.z;
        placement = mul(mul(t1, sc), t0);
        inverse_affine(placement.m, placement_inv.m);
    }
    // Transform product: m = object_to_world.m * placement.m, inv = placement.inv * object_to_world.inv
    M4 ow, oi;
    std::memcpy(ow.m, o.object_to_world, sizeof(ow.m));
    std::memcpy(oi.m, o.world_to_object, sizeof(oi.m));
    const M4 mm = mul(ow, placement), mi = mul(placement_inv, oi);
    std::memcpy(m.object_to_world, mm.m, sizeof(mm.m));
    std::memcpy(m.world_to_object, mi.m, sizeof(mi.m));
}

void default_options(lux_obj_options& o) {
    std::memset(&o, 0, sizeof(o));
    o.normal_mode = 0;
    o.boundary_topology = 0;
    o.position_weld_epsilon = 1e-6f;
    o.cap_boundary_holes = 0;
    o.cap_boundary_topology = 1;
    o.default_material_id = -1;
    o.placement_target_extent = 1.0f;
    for (int k = 0; k < 16; ++k) o.object_to_world[k] = o.world_to_object[k] = (k % 5 == 0) ? 1.0f : 0.0f;
}

}  // namespace obj
}  // namespace lux

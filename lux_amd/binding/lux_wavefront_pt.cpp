// lux_wavefront_pt.cpp -- the reference's `-gpu` integrator on the MI355X
// library. In a LUX build this file replaces src/integrator/wavefront_pt.cu;
// the declaration (src/integrator/wavefront_pt.cuh:15-28), make_integrator and
// run_render (src/app/render_runner.cpp:329-332, :452-464) stay as they are.
//
// WavefrontPathTracer::render hands the PreparedScene's host Scene to the
// library once per (Scene::identity, Scene::revision) -- a scene of one
// untransformed mesh as a lux_scene_desc, any other as a
// lux_instanced_scene_desc -- then renders into Film::pixels (W*H sums over
// spp, the reference's layout) with the Scene's built camera vectors and the
// RenderSettings' seed, frame, sample offset, depth and light sampler. The
// library keeps the device scene (BVHs, light tables) between renders, so
// requires_gpu_scene() is false: PreparedScene::prepare_gpu need not run.
//
// Inputs the library does not render are refused with std::runtime_error
// instead of being rendered differently: SamplerKind::Sobol, image emission on
// area lights (EmissionSourceKind::Image), orthographic or thin-lens cameras,
// and the debug outputs of RenderOutputs (AOV films, path logs).
#include <cstdint>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "integrator/wavefront_pt.cuh"
#include "lux_hip.h"

static_assert(sizeof(vec3) == sizeof(lux_vec3), "vec3 layout");
static_assert(sizeof(TriangleIndices) == 3 * sizeof(uint32_t), "TriangleIndices layout");
static_assert(sizeof(mat4) == 16 * sizeof(float), "mat4 layout");

// The library scene the last render used, and which Scene it was made from.
struct WavefrontRuntime {
    lux_hip_scene* scene = nullptr;
    uint64_t identity = 0, revision = 0;
    WavefrontRuntime() = default;
    WavefrontRuntime(const WavefrontRuntime&) = delete;
    WavefrontRuntime& operator=(const WavefrontRuntime&) = delete;
    ~WavefrontRuntime() {
        if (scene) lux_hip_scene_destroy(scene);
    }
};

namespace {

void check(int status, const char* what) {
    if (status != LUX_OK) throw std::runtime_error(std::string("MI355X error after ") + what + ": " +
                                                   lux_hip_last_error());
}

[[noreturn]] void unsupported(const char* what) {
    throw std::runtime_error(std::string("WavefrontPathTracer on MI355X does not render ") + what);
}

lux_vec3 v3(const vec3& v) { return lux_vec3{v.x, v.y, v.z}; }

std::vector<lux_material> materials_of(const Scene& scene) {
    std::vector<lux_material> out(scene.materials.size());
    for (size_t i = 0; i < out.size(); ++i) {
        const Material& m = scene.materials[i];
        if (m.emission_source == EmissionSourceKind::Image) unsupported("image emission (EmissionSourceKind::Image)");
        out[i].type = int32_t(m.type);  // MaterialType order: Lambert .. Transparent
        out[i].albedo = v3(m.albedo);
        out[i].emission = v3(m.emission);
        out[i].emission_two_sided = m.emission_sidedness == EmissionSidedness::DoubleSided ? 1 : 0;
        out[i].roughness = m.roughness;
        out[i].metallic = m.metallic;
        out[i].eta = m.eta;
    }
    return out;
}

// Scene::lights besides the area lights (those come from emissive triangles).
std::vector<lux_light> lights_of(const Scene& scene) {
    std::vector<lux_light> out;
    for (const Light& l : scene.lights) {
        lux_light x{};
        switch (l.kind) {
            case LightKind::Point:
                x.kind = LUX_LIGHT_POINT;
                x.vector = v3(l.point.position);
                x.color = v3(l.point.intensity);
                break;
            case LightKind::Distant:
                x.kind = LUX_LIGHT_DISTANT;
                x.vector = v3(l.distant.direction_to_light);
                x.color = v3(l.distant.radiance);
                break;
            case LightKind::UniformInfinite:
                x.kind = LUX_LIGHT_UNIFORM_INFINITE;
                x.color = v3(l.uniform_infinite.radiance);
                break;
            case LightKind::ImageInfinite: {
                const ImageInfiniteLight& im = l.image_infinite;
                if (im.image_id < 0 || im.image_id >= int(scene.image_light_assets.size()))
                    throw std::runtime_error("image infinite light asset id out of range");
                const SceneImageLightAsset& a = scene.image_light_assets[size_t(im.image_id)];
                x.kind = LUX_LIGHT_IMAGE_INFINITE;
                x.color = v3(im.scale);
                x.image_width = a.width;
                x.image_height = a.height;
                x.image_pixels = reinterpret_cast<const lux_vec3*>(a.pixels.data());
                break;
            }
            default:
                continue;  // DiffuseArea: the emissive triangles
        }
        out.push_back(x);
    }
    return out;
}

// PerspectiveCamera's built vectors {origin, lower_left, horizontal, vertical}.
void camera_of(const Camera& camera, float v[12]) {
    if (camera.kind != CameraKind::Perspective) unsupported("orthographic cameras");
    const PerspectiveCamera& c = camera.perspective;
    if (c.lens_radius > 0) unsupported("thin-lens cameras (lens_radius > 0)");
    const vec3* parts[4] = {&c.origin, &c.lower_left, &c.horizontal, &c.vertical};
    for (int k = 0; k < 4; ++k) {
        v[3 * k] = parts[k]->x;
        v[3 * k + 1] = parts[k]->y;
        v[3 * k + 2] = parts[k]->z;
    }
}

bool is_identity(const mat4& m) {
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j)
            if (m.m[i][j] != (i == j ? 1.0f : 0.0f)) return false;
    return true;
}

lux_hip_scene* create_scene(const Scene& scene) {
    std::vector<lux_material> mats = materials_of(scene);
    std::vector<lux_light> lights = lights_of(scene);
    lux_hip_scene* out = nullptr;
    const bool one_mesh = scene.instances.size() == 1 && scene.instances[0].mesh_id >= 0 &&
                          scene.instances[0].mesh_id < int(scene.mesh_assets.size()) &&
                          scene.instances[0].material_override < 0 &&
                          is_identity(scene.instances[0].object_to_world.m);
    if (one_mesh) {
        const SceneMesh& mesh = scene.mesh_assets[size_t(scene.instances[0].mesh_id)];
        lux_scene_desc d{};
        d.vertices = reinterpret_cast<const lux_vec3*>(mesh.vertices.data());
        d.vertex_count = int32_t(mesh.vertices.size());
        d.indices = reinterpret_cast<const uint32_t*>(mesh.indices.data());
        d.normals = mesh.normals.size() == mesh.vertices.size() ? reinterpret_cast<const lux_vec3*>(mesh.normals.data())
                                                                 : nullptr;
        d.material_ids = mesh.material_ids.data();
        d.triangle_count = int32_t(mesh.indices.size());
        d.materials = mats.data();
        d.material_count = int32_t(mats.size());
        d.lights = lights.data();
        d.light_count = int32_t(lights.size());
        check(lux_hip_scene_create(&d, &out), "lux_hip_scene_create");
        return out;
    }
    // meshes back to back with mesh-local indices, instances with both matrices
    std::vector<lux_vec3> verts, norms;
    std::vector<uint32_t> idx;
    std::vector<int32_t> mat_ids, voff{0}, toff{0};
    for (const SceneMesh& mesh : scene.mesh_assets) {
        const bool has_normals = mesh.normals.size() == mesh.vertices.size();
        for (size_t v = 0; v < mesh.vertices.size(); ++v) {
            verts.push_back(v3(mesh.vertices[v]));
            norms.push_back(has_normals ? v3(mesh.normals[v]) : lux_vec3{0, 0, 0});
        }
        for (size_t t = 0; t < mesh.indices.size(); ++t) {
            idx.insert(idx.end(), {mesh.indices[t].v0, mesh.indices[t].v1, mesh.indices[t].v2});
            mat_ids.push_back(mesh.material_ids[t]);
        }
        voff.push_back(int32_t(verts.size()));
        toff.push_back(int32_t(mat_ids.size()));
    }
    std::vector<lux_instance> inst(scene.instances.size());
    for (size_t i = 0; i < inst.size(); ++i) {
        const SceneInstance& in = scene.instances[i];
        inst[i].mesh_id = in.mesh_id;
        inst[i].material_override = in.material_override;
        for (int r = 0; r < 4; ++r)
            for (int c = 0; c < 4; ++c) {
                inst[i].object_to_world[4 * r + c] = in.object_to_world.m.m[r][c];
                inst[i].world_to_object[4 * r + c] = in.object_to_world.inv.m[r][c];
            }
    }
    lux_instanced_scene_desc d{};
    d.vertices = verts.data();
    d.normals = norms.data();
    d.vertex_offsets = voff.data();
    d.indices = idx.data();
    d.material_ids = mat_ids.data();
    d.triangle_offsets = toff.data();
    d.mesh_count = int32_t(scene.mesh_assets.size());
    d.instances = inst.data();
    d.instance_count = int32_t(inst.size());
    d.materials = mats.data();
    d.material_count = int32_t(mats.size());
    d.lights = lights.data();
    d.light_count = int32_t(lights.size());
    check(lux_hip_scene_create_instanced(&d, &out), "lux_hip_scene_create_instanced");
    return out;
}

int32_t light_sampler_of(LightSamplerKind k) {
    switch (k) {
        case LightSamplerKind::Bvh: return LUX_LIGHT_SAMPLER_BVH;
        case LightSamplerKind::Uniform: return LUX_LIGHT_SAMPLER_UNIFORM;
        default: return LUX_LIGHT_SAMPLER_POWER;
    }
}

}  // namespace

WavefrontPathTracer::WavefrontPathTracer() : runtime_(std::make_unique<WavefrontRuntime>()) {
    check(lux_hip_init(0), "lux_hip_init");
}

WavefrontPathTracer::~WavefrontPathTracer() = default;

bool WavefrontPathTracer::requires_gpu_scene() const { return false; }

void WavefrontPathTracer::render(const PreparedScene& prepared, Film& film, const RenderSettings& settings,
                                 RenderOutputs* outputs) {
    if (film.width <= 0 || film.height <= 0 || film.samples_per_pixel <= 0)
        throw std::runtime_error("WavefrontPathTracer requires a non-empty Film");
    if (film.pixels.size() != size_t(film.width) * size_t(film.height))
        throw std::runtime_error("WavefrontPathTracer Film pixel storage does not match dimensions");
    if (settings.max_depth < 0) throw std::runtime_error("WavefrontPathTracer requires max_depth >= 0");
    if (settings.batch_size <= 0) throw std::runtime_error("WavefrontPathTracer requires batch_size > 0");
    if (settings.sampler_kind != SamplerKind::Independent) unsupported("the Sobol sampler");
    if (outputs && (outputs->wants_path_log() || outputs->wants(RenderAov::ShadowDebug) ||
                    outputs->wants(RenderAov::GeometricNormal) || outputs->wants(RenderAov::ShadingNormal)))
        unsupported("AOV films or path logs");
    const Scene& scene = prepared.host_scene();
    if (!runtime_->scene || runtime_->identity != scene.identity || runtime_->revision != scene.revision) {
        float cam[12];
        camera_of(scene.camera, cam);
        if (runtime_->scene) lux_hip_scene_destroy(runtime_->scene);
        runtime_->scene = nullptr;
        runtime_->scene = create_scene(scene);
        check(lux_hip_scene_set_camera(runtime_->scene, cam), "lux_hip_scene_set_camera");
        runtime_->identity = scene.identity;
        runtime_->revision = scene.revision;
    }
    lux_render_settings rs{};
    rs.width = film.width;
    rs.height = film.height;
    rs.spp = film.samples_per_pixel;
    rs.max_depth = settings.max_depth;
    rs.seed = settings.seed;
    rs.frame_index = settings.frame_index;
    rs.sample_offset = settings.sample_offset;
    rs.batch_size = 0;  // the wavefront schedule's own batch
    rs.mode = LUX_RENDER_WAVEFRONT;
    rs.light_sampler = light_sampler_of(settings.light_sampler_kind);
    check(lux_hip_render(runtime_->scene, &rs, reinterpret_cast<float*>(film.pixels.data())), "WavefrontPathTracer");
}

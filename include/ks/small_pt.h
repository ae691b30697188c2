// C++ mirror of keepsake's small_pt entry points over the keepsake_amd C ABI.
//
// Same names, argument meaning and failure behaviour as the reference types:
//   ks::Scene         (scene.h:49-87)     geometry + acceleration structure + mesh lights on the device
//   ks::Camera        (camera.h:10-34)    look_at + perspective vfov/aspect, or a to_world matrix
//   ks::RenderTarget2 (render_target.h:224-293)  FP64 sums + (hit, miss) weights, composite
//   ks::SmallPTInput  (small_pt.h:17-41)  the render inputs, including the progress callback
//   ks::SmallPT::run  (small_pt.h:45-49)  the render call
// A non-zero status from the library aborts with the library's message, as the reference's ASSERT does;
// call the C ABI directly to handle errors instead.
#pragma once
#include "../keepsake_amd.h"

#include <array>
#include <cstdio>
#include <cstdlib>
#include <functional>
#include <memory>
#include <string>
#include <vector>

namespace ks {

inline void check(int status, const char *what) {
    if (status != KS_OK) {
        std::fprintf(stderr, "[keepsake_amd] %s failed (%d): %s\n", what, status, ks_last_error());
        std::abort();
    }
}

using color3 = std::array<float, 3>;
using vec3 = std::array<float, 3>;

class Device {
  public:
    explicit Device(int ordinal = 0) { check(ks_device_open(ordinal, &h_), "ks_device_open"); }
    ~Device() { ks_device_close(h_); }
    Device(const Device &) = delete;
    Device &operator=(const Device &) = delete;
    ks_device *handle() const { return h_; }

  private:
    ks_device *h_ = nullptr;
};

// Host-side scene description: flat triangles grouped by geometry, one material per geometry,
// non-mesh lights (sky ambient, sky map, directional, point) in the order the light sampler sees them.
struct SceneData {
    std::vector<float> vertices, normals;
    std::vector<uint32_t> indices, tri_geom, geom_material;
    std::vector<uint8_t> geom_flags; // per geometry KS_GEOM_* bits, empty: two-sided with smooth normals
    std::vector<ks_material_desc> materials;
    std::vector<ks_light_desc> lights;
    // environment map of the KS_LIGHT_SKY_MAP light (rows of RGB texels, row 0 at theta = 0)
    std::vector<float> sky_rgb;
    int sky_width = 0, sky_height = 0;
    std::array<float, 9> sky_to_world{1, 0, 0, 0, 1, 0, 0, 0, 1}; // row-major
    bool sky_y_up = true;
    float sky_strength = 1.0f;
    // textures: per-vertex texture coordinates, texture levels and descriptors (data pointers refreshed
    // by desc()), and textured materials (field_of[m] indexes `fields`, -1: the constant material)
    std::vector<float> texcoords;
    std::vector<std::vector<uint8_t>> texture_bytes;
    std::vector<ks_texture_desc> textures;
    std::vector<ks_material_fields> fields;
    std::vector<int32_t> field_of;

    ks_scene_desc desc() const {
        ks_scene_desc d{};
        d.vertices = vertices.data();
        d.n_vertices = (uint32_t)(vertices.size() / 3);
        d.normals = normals.empty() ? nullptr : normals.data();
        d.indices = indices.data();
        d.n_tris = (uint32_t)(indices.size() / 3);
        d.tri_geom = tri_geom.data();
        d.geom_material = geom_material.data();
        d.n_geoms = (uint32_t)geom_material.size();
        d.materials = materials.data();
        d.n_materials = (uint32_t)materials.size();
        d.lights = lights.empty() ? nullptr : lights.data();
        d.n_lights = (uint32_t)lights.size();
        d.geom_flags = geom_flags.empty() ? nullptr : geom_flags.data();
        if (!sky_rgb.empty()) {
            sky_desc_.rgb = sky_rgb.data();
            sky_desc_.width = sky_width;
            sky_desc_.height = sky_height;
            for (int k = 0; k < 9; ++k) sky_desc_.to_world[k] = sky_to_world[k];
            sky_desc_.transform_y_up = sky_y_up ? 1 : 0;
            sky_desc_.strength = sky_strength;
            d.sky_map = &sky_desc_;
        }
        d.texcoords = texcoords.empty() ? nullptr : texcoords.data();
        tex_desc_ = textures;
        for (size_t i = 0; i < tex_desc_.size(); ++i) tex_desc_[i].data = texture_bytes[i].data();
        d.textures = tex_desc_.empty() ? nullptr : tex_desc_.data();
        d.n_textures = (uint32_t)tex_desc_.size();
        field_ptrs_.assign(materials.size(), nullptr);
        bool any = false;
        for (size_t m = 0; m < field_of.size() && m < materials.size(); ++m)
            if (field_of[m] >= 0) field_ptrs_[m] = &fields[(size_t)field_of[m]], any = true;
        d.material_fields = any ? field_ptrs_.data() : nullptr;
        return d;
    }

  private:
    mutable ks_sky_map_desc sky_desc_{}; // points into sky_rgb; refreshed by every desc()
    mutable std::vector<ks_texture_desc> tex_desc_;
    mutable std::vector<const ks_material_fields *> field_ptrs_;
};

// Scene: create_scene + rtcCommitScene + build_mesh_lights + PowerLightSampler::build in one step.
class Scene {
  public:
    Scene(const Device &dev, const SceneData &data, const ks_bvh_options *opt = nullptr) {
        ks_scene_desc d = data.desc();
        check(ks_scene_create(dev.handle(), &d, opt, &h_), "ks_scene_create");
    }
    ~Scene() { ks_scene_free(h_); }
    Scene(const Scene &) = delete;
    Scene &operator=(const Scene &) = delete;
    ks_scene *handle() const { return h_; }
    ks_scene_info info() const {
        ks_scene_info i{};
        check(ks_scene_info_get(h_, &i), "ks_scene_info_get");
        return i;
    }

  private:
    ks_scene *h_ = nullptr;
};

struct Camera {
    Camera() = default;
    // Camera(position, target, up, vfov, aspect) with vfov in radians (camera.h:14)
    Camera(const vec3 &position, const vec3 &target, const vec3 &up, float vfov, float aspect) {
        for (int i = 0; i < 3; ++i) {
            desc.pos[i] = position[i];
            desc.target[i] = target[i];
            desc.up[i] = up[i];
        }
        desc.vfov_deg = vfov * 180.0f / 3.14159265359f;
        desc.aspect = aspect;
    }
    // Camera(to_world, vfov, aspect) with a row-major camera-to-world matrix (camera.h:13)
    Camera(const std::array<float, 16> &to_world, float vfov, float aspect) {
        desc.use_matrix = 1;
        for (int i = 0; i < 16; ++i) desc.to_world[i] = to_world[i];
        desc.vfov_deg = vfov * 180.0f / 3.14159265359f;
        desc.aspect = aspect;
    }
    ks_camera_desc desc{};
};

class RenderTarget2 {
  public:
    RenderTarget2(const Device &dev, uint32_t width, uint32_t height, const color3 &backdrop = {0, 0, 0})
        : width(width), height(height) {
        check(ks_film_create(dev.handle(), (int32_t)width, (int32_t)height, backdrop.data(), &h_), "ks_film_create");
    }
    ~RenderTarget2() { ks_film_free(h_); }
    RenderTarget2(const RenderTarget2 &) = delete;
    RenderTarget2 &operator=(const RenderTarget2 &) = delete;
    void clear() { check(ks_film_clear(h_), "ks_film_clear"); }
    // main.pixels (FP64 RGB sums) and pixel_weights (FP64 hit, miss) in row-major order
    void download(std::vector<double> &main, std::vector<double> &pixel_weights) const {
        main.resize((size_t)width * height * 3);
        pixel_weights.resize((size_t)width * height * 2);
        check(ks_film_download(h_, main.data(), pixel_weights.data()), "ks_film_download");
    }
    // composite: lerp(backdrop, sum / w_hit, w_hit / (w_hit + w_miss)) per pixel, as float RGB
    std::vector<float> composite() const {
        std::vector<float> rgb((size_t)width * height * 3);
        check(ks_film_composite(h_, rgb.data()), "ks_film_composite");
        return rgb;
    }
    ks_film *handle() const { return h_; }
    uint32_t width, height;

  private:
    ks_film *h_ = nullptr;
};

// The learned path guide (no reference counterpart): hash grid + MLP -> 8-lobe vMF mixture on the device,
// trained online by guided renders. Bounds are the scene's bounding box.
class Guide {
  public:
    Guide(const Device &dev, const Scene &scene, const ks_guide_config &cfg = ks_guide_config{}) {
        ks_scene_info i = scene.info();
        check(ks_guide_create(dev.handle(), &cfg, i.bound_lo, i.bound_hi, &h_), "ks_guide_create");
    }
    ~Guide() { ks_guide_free(h_); }
    Guide(const Guide &) = delete;
    Guide &operator=(const Guide &) = delete;
    void set_mode(float guide_fraction, int max_steps_per_pass) {
        check(ks_guide_set_mode(h_, guide_fraction, max_steps_per_pass), "ks_guide_set_mode");
    }
    ks_guide_stats stats() const {
        ks_guide_stats st{};
        check(ks_guide_get_stats(h_, &st), "ks_guide_get_stats");
        return st;
    }
    // FP8 (e4m3) or BF16 MLP inference for later renders
    void set_inference(bool fp8) { check(ks_guide_set_inference(h_, fp8 ? 1 : 0), "ks_guide_set_inference"); }
    ks_guide *handle() const { return h_; }

  private:
    ks_guide *h_ = nullptr;
};

// PixelFilter (render_target.h:70-184) as create_pixel_filter builds it from its TOML keys
// (render_target.cpp:28-44): type KS_FILTER_*, width (0: the type's default), sigma (gaussian).
struct PixelFilter {
    int type = KS_FILTER_BOX;
    float width = 0.0f;
    float sigma = 0.0f;
};

struct SmallPTInput {
    const Scene *scene = nullptr;
    const Camera *camera = nullptr;
    PixelFilter pixel_filter;
    RenderTarget2 *target = nullptr; // the film the run accumulates into (created inside run() in the reference)
    bool include_background = true;
    color3 backdrop = {0.0f, 0.0f, 0.0f};

    int bounces = 1;
    int render_width = 256;
    int render_height = 256;
    int crop_start_x = 0;
    int crop_start_y = 0;
    int crop_width = 0;
    int crop_height = 0;
    int spp = 1;
    int rng_seed = 0;
    bool scale_ray_diff = true;
    float clamp_direct = 0.0f;
    float clamp_indirect = 10.0f;

    int spp_prog_interval = 32;
    std::function<void(const RenderTarget2 &rt, int)> prog_interval_callback;

    // additions of this build
    int mis_mode = KS_MIS_REFERENCE;
    int guide_mode = KS_GUIDE_OFF; // KS_GUIDE_TRAIN_AND_SAMPLE needs `guide` and mis_mode consistent
    Guide *guide = nullptr;
    int queue_size = 0;

    ks_small_pt_params params() const {
        ks_small_pt_params p{};
        p.bounces = bounces;
        p.render_width = render_width;
        p.render_height = render_height;
        p.crop_start_x = crop_start_x;
        p.crop_start_y = crop_start_y;
        p.crop_width = crop_width;
        p.crop_height = crop_height;
        p.spp = spp;
        p.rng_seed = rng_seed;
        p.include_background = include_background ? 1 : 0;
        for (int i = 0; i < 3; ++i) p.backdrop[i] = backdrop[i];
        p.scale_ray_diff = scale_ray_diff ? 1 : 0;
        p.clamp_direct = clamp_direct;
        p.clamp_indirect = clamp_indirect;
        p.spp_prog_interval = spp_prog_interval;
        p.mis_mode = mis_mode;
        p.guide_mode = KS_GUIDE_OFF; // the guided entry point is selected by run()
        p.queue_size = queue_size;
        p.pixel_filter = pixel_filter.type;
        p.filter_width = pixel_filter.width;
        p.filter_sigma = pixel_filter.sigma;
        return p;
    }
};

struct SmallPT {
    // SmallPT::run (small_pt.cpp:15-70); returns the render statistics.
    ks_render_stats run(const SmallPTInput &in) const {
        if (!in.scene || !in.camera || !in.target) {
            std::fprintf(stderr, "[keepsake_amd] SmallPT::run needs scene, camera and target\n");
            std::abort();
        }
        ks_small_pt_params p = in.params();
        ks_render_stats st{};
        struct Ctx {
            const SmallPTInput *in;
        } ctx{&in};
        ks_progress_fn fn = nullptr;
        if (in.prog_interval_callback) {
            fn = [](const ks_film *, int32_t spp_finished, void *user) {
                const SmallPTInput *i = static_cast<Ctx *>(user)->in;
                i->prog_interval_callback(*i->target, spp_finished);
            };
        }
        if (in.guide_mode == KS_GUIDE_TRAIN_AND_SAMPLE) {
            if (!in.guide) {
                std::fprintf(stderr, "[keepsake_amd] guide_mode train_and_sample needs SmallPTInput::guide\n");
                std::abort();
            }
            check(ks_small_pt_render_guided(in.scene->handle(), &in.camera->desc, &p, in.guide->handle(),
                                            in.target->handle(), fn, &ctx, &st),
                  "ks_small_pt_render_guided");
        } else {
            check(ks_small_pt_render(in.scene->handle(), &in.camera->desc, &p, in.target->handle(), fn, &ctx, &st),
                  "ks_small_pt_render");
        }
        return st;
    }
};

} // namespace ks

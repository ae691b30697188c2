// The small_pt task (small_pt.cpp:180-303) on the gfx950 library, with the same TOML keys:
// object, material, sky, light[], pixel_filter, tone_mapper, backdrop, bounces, clamp_direct,
// clamp_indirect, render_width, render_height, crop_*, spp, scale_ray_diff, rng_seed,
// spp_prog_interval, camera. A scene is either a mesh asset (`object = "mesh_asset.X"` plus one
// material list, create_scene scene.cpp:345-365) or one of this build's procedural scenes (keys
// `procedural`, `procedural_seed`). Materials are [material.*] tables whose `bsdf` names a
// [bsdf.*] table of type principled_bsdf with constant shader fields (create_material
// material.cpp:169-177, create_principled_bsdf principled_bsdf.cpp:707-732). `compound_object =
// "compound_mesh_asset.X"` loads a glTF scene whose instances are flattened into world space; its
// materials come from the glTF file (load_materials = true) unless the task gives a material list,
// one inner list per glTF mesh. Additional keys of this build: `emission` on a
// material table (a constant shader field or an RGB triple: makes its triangles mesh lights),
// output_pfm (also write PFM files beside the EXR / PNG ones), mis_mode ("reference" or
// "consistent"), queue_size, bvh_builder ("sbvh", the default, or "device": the BVH built on the GPU),
// and a [task.guide] table that turns on the
// learned path guide (keys fraction, train_spp, learning_rate, beta1, beta2, log2_table_size,
// base_resolution, max_resolution, train_batch, max_steps_per_pass, weight_clamp, seed, max_bounce
// (guide only vertices below this bounce; 0: all), inference ("bf16" or "fp8"); it needs mis_mode =
// "consistent").
#include "small_pt_task.h"

#include "gltf_asset.h"
#include "mesh_asset.h"
#include "texture_asset.h"

#include <cctype>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <map>
#include <memory>

namespace ks {

static SceneData scene_from_host(const ks_host_scene *hs, ks_camera_desc *cam) {
    ks_scene_desc d{};
    check(ks_host_scene_get(hs, &d, cam), "ks_host_scene_get");
    SceneData s;
    s.vertices.assign(d.vertices, d.vertices + 3 * (size_t)d.n_vertices);
    if (d.normals) s.normals.assign(d.normals, d.normals + 3 * (size_t)d.n_vertices);
    s.indices.assign(d.indices, d.indices + 3 * (size_t)d.n_tris);
    s.tri_geom.assign(d.tri_geom, d.tri_geom + d.n_tris);
    s.geom_material.assign(d.geom_material, d.geom_material + d.n_geoms);
    s.materials.assign(d.materials, d.materials + d.n_materials);
    if (d.lights) s.lights.assign(d.lights, d.lights + d.n_lights);
    return s;
}

// A principled_bsdf field (create_principled_bsdf, principled_bsdf.cpp:707-732): a shader field table.
static ks_field_desc bsdf_field(const ConfigArgs &bsdf, const char *key, int channels, TextureRegistry &reg) {
    if (!bsdf.contains(key)) throw ConfigError(std::string("principled_bsdf needs '") + key + "'");
    return load_shader_field(bsdf[key], channels, reg);
}

// [normal_map.X] (create_normal_map, normal_map.cpp:42-56): map (a 3-channel shader field), space
// ("tangent" | "world"), range01 (true), strength (1), to_world.
static void load_normal_map(const ConfigArgs &nm, ks_material_fields &f, TextureRegistry &reg) {
    f.has_normal_map = 1;
    f.normal_map = load_shader_field(nm["map"], 3, reg);
    std::string space = nm.load_string("space", "tangent");
    if (space != "tangent" && space != "world") throw ConfigError("normal map space must be tangent or world");
    f.normal_space = space == "world" ? KS_NORMAL_MAP_WORLD : KS_NORMAL_MAP_TANGENT;
    f.normal_range01 = nm.load_bool("range01", true) ? 1 : 0;
    f.normal_strength = nm.load_float("strength", 1.0f);
    std::array<float, 16> m = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
    if (nm.contains("to_world")) m = nm.load_transform("to_world");
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) f.normal_to_world[3 * r + c] = m[4 * r + c];
}

// [material.X] -> bsdf = "bsdf.Y" of type principled_bsdf (the only BSDF the kernels evaluate), and its
// normal map (create_material, material.cpp:169-177). `fields` receives every closure input as a shader
// field; `textured` says whether any of them (or a normal map) needs the per-hit evaluation, else the
// constant record `d` holds the material.
static ks_material_desc load_material(const ConfigArgs &mat, TextureRegistry &reg, ks_material_fields &fields,
                                      bool &textured) {
    if (mat.contains("subsurface") && !mat.load_string("subsurface").empty())
        throw ConfigError("material key 'subsurface' is not supported");
    ConfigArgs bsdf = mat.asset(mat.load_string("bsdf"));
    std::string type = bsdf.load_string("type");
    if (type != "principled_bsdf") throw ConfigError("bsdf type '" + type + "' is not supported (principled_bsdf only)");
    if (bsdf.load_string("microfacet", "ggx") != "ggx") throw ConfigError("only the ggx microfacet is supported");
    fields = ks_material_fields{};
    fields.basecolor = bsdf_field(bsdf, "basecolor", 3, reg);
    fields.roughness = bsdf_field(bsdf, "roughness", 1, reg);
    fields.metallic = bsdf_field(bsdf, "metallic", 1, reg);
    fields.ior = bsdf_field(bsdf, "ior", 1, reg);
    fields.specular_r0_mul = bsdf_field(bsdf, "specular_r0_mul", 1, reg);
    fields.specular_trans = bsdf_field(bsdf, "specular_trans", 1, reg);
    fields.diffuse_trans = bsdf_field(bsdf, "diffuse_trans", 1, reg);
    fields.diffuse_trans_fwd = bsdf_field(bsdf, "diffuse_trans_fwd", 1, reg);
    // the bsdf's `emissive` field is read but, as in the reference's material path, does not make a light
    if (bsdf.contains("emissive")) (void)load_shader_field(bsdf["emissive"], 3, reg);
    fields.emission.texture = -1;
    if (mat.contains("emission")) { // extension: a shader field table or an RGB triple
        if (mat["emission"].node->is_table()) {
            fields.emission = load_shader_field(mat["emission"], 3, reg);
        } else {
            auto e = mat.load_vec3("emission");
            for (int k = 0; k < 3; ++k) fields.emission.value[k] = e[k];
        }
        fields.has_emission = 1;
    }
    const std::string nm = mat.load_string("normal_map", "");
    if (!nm.empty()) load_normal_map(mat.asset(nm), fields, reg);
    const ks_field_desc *all[9] = {&fields.basecolor, &fields.roughness, &fields.metallic, &fields.ior,
                                   &fields.specular_r0_mul, &fields.specular_trans, &fields.diffuse_trans,
                                   &fields.diffuse_trans_fwd, &fields.emission};
    textured = fields.has_normal_map != 0;
    for (const ks_field_desc *f : all) textured |= f->texture >= 0;
    ks_material_desc d{};
    for (int k = 0; k < 3; ++k) d.basecolor[k] = fields.basecolor.value[k];
    d.roughness = fields.roughness.value[0];
    d.metallic = fields.metallic.value[0];
    d.ior = fields.ior.value[0];
    d.specular_r0_mul = fields.specular_r0_mul.value[0];
    d.specular_trans = fields.specular_trans.value[0];
    d.diffuse_trans = fields.diffuse_trans.value[0];
    d.diffuse_trans_fwd = fields.diffuse_trans_fwd.value[0];
    for (int k = 0; k < 3; ++k) d.emission[k] = fields.emission.value[k];
    d.has_emission = fields.has_emission;
    return d;
}

// Flat scene assembly: one geometry per (instance, mesh) in order, so that the mesh-light order is the
// reference's (Scene::build_mesh_lights, scene.cpp:316-332). The vertex normals array is all or nothing:
// geometries without smooth normals get zero rows and the KS_GEOM_FLAT flag.
class FlatSceneBuilder {
  public:
    FlatSceneBuilder(bool any_normals, bool any_texcoords) : any_normals_(any_normals), any_uv_(any_texcoords) {}
    void add(const MeshData &m, uint32_t material) {
        uint32_t base = (uint32_t)(s_.vertices.size() / 3);
        if ((uint64_t)base + m.vertex_count() >= (1ull << 32)) throw ConfigError("scene has too many vertices");
        s_.vertices.insert(s_.vertices.end(), m.vertices.begin(), m.vertices.end());
        bool smooth = m.use_smooth_normal && !m.vertex_normals.empty();
        if (any_normals_) {
            if (smooth) s_.normals.insert(s_.normals.end(), m.vertex_normals.begin(), m.vertex_normals.end());
            else s_.normals.insert(s_.normals.end(), m.vertices.size(), 0.0f);
        }
        bool uv = !m.texcoords.empty();
        if (any_uv_) {
            if (uv) s_.texcoords.insert(s_.texcoords.end(), m.texcoords.begin(), m.texcoords.end());
            else s_.texcoords.insert(s_.texcoords.end(), 2 * m.vertex_count(), 0.0f);
        }
        for (uint32_t i : m.indices) s_.indices.push_back(base + i);
        s_.tri_geom.insert(s_.tri_geom.end(), m.tri_count(), (uint32_t)s_.geom_material.size());
        s_.geom_flags.push_back((uint8_t)((m.twosided ? 0u : KS_GEOM_ONE_SIDED) | (smooth ? 0u : KS_GEOM_FLAT) |
                                          (any_uv_ && !uv ? KS_GEOM_NO_TEXCOORD : 0u)));
        s_.geom_material.push_back(material);
    }
    // a [material.*] table by name, read once
    uint32_t named_material(const ConfigArgs &args, const TomlValue *name) {
        if (name->type != TomlValue::Type::String) throw ConfigError("material list entries must be strings");
        auto it = named_.find(name->str);
        if (it == named_.end()) {
            ks_material_fields f;
            bool textured = false;
            ks_material_desc d = load_material(args.asset(name->str), reg_, f, textured);
            it = named_.emplace(name->str, add_material(d, textured ? &f : nullptr)).first;
        }
        return it->second;
    }
    uint32_t add_material(const ks_material_desc &d, const ks_material_fields *fields = nullptr) {
        s_.materials.push_back(d);
        s_.field_of.push_back(-1);
        if (fields) {
            s_.field_of.back() = (int32_t)s_.fields.size();
            s_.fields.push_back(*fields);
        }
        return (uint32_t)s_.materials.size() - 1;
    }
    TextureRegistry &textures() { return reg_; }
    SceneData finish() {
        bool any_flag = false;
        for (uint8_t f : s_.geom_flags) any_flag |= f != 0;
        if (!any_flag) s_.geom_flags.clear();
        if (s_.indices.empty()) throw ConfigError("mesh asset has no triangles");
        for (TextureData &t : reg_.textures) {
            s_.textures.push_back(t.desc());
            s_.texture_bytes.push_back(std::move(t.bytes));
        }
        if (s_.fields.empty()) s_.field_of.clear();
        return std::move(s_);
    }

  private:
    bool any_normals_, any_uv_;
    SceneData s_;
    std::map<std::string, uint32_t> named_;
    TextureRegistry reg_;
};

// The material list of the task: one inner list of [material.*] names per subscene, one name per
// geometry (assign_material_list, mesh_asset.cpp:285-299).
static ConfigArgs material_list(const ConfigArgs &args, size_t subscenes) {
    ConfigArgs list = args["material"];
    if (list.array_size() != subscenes)
        throw ConfigError("Material list mismatch dimension: " + std::to_string(list.array_size()) + " lists for " +
                          std::to_string(subscenes) + " subscenes.");
    return list;
}

// create_scene_from_mesh_asset + assign_material_list (mesh_asset.cpp:267-299): one subscene with one
// geometry per mesh, an identity instance, material[0][g] for geometry g.
static SceneData scene_from_mesh_asset(const MeshAsset &asset, const ConfigArgs &args) {
    if (!args.contains("material"))
        throw ConfigError("No materials. Provide a material list (material = [[\"material.a\", ...]]).");
    ConfigArgs names = material_list(args, 1)[0];
    if (names.array_size() != asset.meshes.size())
        throw ConfigError("Material list mismatch dimension: " + std::to_string(names.array_size()) +
                          " materials for " + std::to_string(asset.meshes.size()) + " meshes.");
    bool any_normals = false, any_uv = false;
    for (const MeshData &m : asset.meshes) {
        any_normals |= m.use_smooth_normal && !m.vertex_normals.empty();
        any_uv |= !m.texcoords.empty();
    }
    FlatSceneBuilder b(any_normals, any_uv);
    for (size_t g = 0; g < asset.meshes.size(); ++g)
        b.add(asset.meshes[g], b.named_material(args, names[(int)g].node));
    return b.finish();
}

// create_scene_from_compound_mesh_asset (mesh_asset.cpp:1131-1155) with the instances flattened: each
// instance's meshes are written in world space (points by its transform, normals by the inverse
// transpose), in instance order. Materials come from the task's material list when it has one
// (material[prototype][primitive]), else from the glTF materials.
static SceneData scene_from_compound(const CompoundMeshAsset &asset, const ConfigArgs &args) {
    bool toml_materials = args.contains("material");
    if (!toml_materials && !asset.has_materials)
        throw ConfigError("No materials. Either load materials from file(s) (load_materials = true) or provide a "
                          "material list.");
    std::vector<std::vector<uint32_t>> mat(asset.prototypes.size());
    bool any_normals = false, any_uv = false;
    for (const auto &inst : asset.instances)
        for (const MeshData &m : asset.prototypes[inst.first].meshes) {
            any_normals |= m.use_smooth_normal && !m.vertex_normals.empty();
            any_uv |= !m.texcoords.empty();
        }
    FlatSceneBuilder b(any_normals, any_uv);
    if (toml_materials) {
        ConfigArgs list = material_list(args, asset.prototypes.size());
        for (size_t p = 0; p < asset.prototypes.size(); ++p) {
            ConfigArgs names = list[(int)p];
            if (names.array_size() != asset.prototypes[p].meshes.size())
                throw ConfigError("Material list mismatch dimension: " + std::to_string(names.array_size()) +
                                  " materials for " + std::to_string(asset.prototypes[p].meshes.size()) +
                                  " primitives of mesh " + std::to_string(p) + ".");
        }
        // materials are read in instance order, as the geometries that use them are written
        for (const auto &inst : asset.instances) {
            auto &mp = mat[inst.first];
            if (!mp.empty()) continue;
            ConfigArgs names = list[(int)inst.first];
            for (size_t g = 0; g < asset.prototypes[inst.first].meshes.size(); ++g)
                mp.push_back(b.named_material(args, names[(int)g].node));
        }
    }
    // the glTF's textures join the scene's after any the TOML materials loaded
    int tex_base = -1;
    auto gltf_material = [&](uint32_t p, size_t g) {
        if (!asset.textured[p][g]) return b.add_material(asset.materials[p][g]);
        if (tex_base < 0) {
            tex_base = (int)b.textures().textures.size();
            for (const TextureData &t : asset.textures) b.textures().textures.push_back(t);
        }
        ks_material_fields f = asset.fields[p][g];
        for (ks_field_desc *fd : {&f.basecolor, &f.roughness, &f.metallic, &f.ior, &f.specular_r0_mul, &f.specular_trans,
                                  &f.diffuse_trans, &f.diffuse_trans_fwd, &f.emission, &f.normal_map, &f.opacity})
            if (fd->texture >= 0) fd->texture += tex_base;
        return b.add_material(asset.materials[p][g], &f);
    };
    for (const auto &inst : asset.instances) {
        const MeshAsset &proto = asset.prototypes[inst.first];
        auto &mp = mat[inst.first];
        if (mp.empty())
            for (size_t g = 0; g < proto.meshes.size(); ++g) mp.push_back(gltf_material(inst.first, g));
        bool identity = true;
        for (int k = 0; k < 16; ++k) identity &= inst.second[k] == ((k % 5 == 0) ? 1.0f : 0.0f);
        for (size_t g = 0; g < proto.meshes.size(); ++g) {
            if (identity) {
                b.add(proto.meshes[g], mp[g]);
            } else {
                MeshData m = proto.meshes[g];
                m.transform(inst.second);
                b.add(m, mp[g]);
            }
        }
    }
    return b.finish();
}

// create_camera (camera.cpp:172-205), perspective only
static Camera load_camera(const ConfigArgs &args) {
    std::string type = args.load_string("type");
    if (type != "perspective") throw ConfigError("camera type '" + type + "' is not supported (perspective only)");
    float aspect = args.load_float("aspect");
    float vfov;
    if (args.contains("vfov")) {
        vfov = args.load_float("vfov") / 180.0f * 3.14159265359f;
    } else {
        float hfov = args.load_float("hfov") / 180.0f * 3.14159265359f;
        vfov = 2.0f * std::atan((1.0f / aspect) * std::tan(0.5f * hfov));
    }
    if (args.contains("to_world")) return Camera(args.load_transform("to_world"), vfov, aspect);
    return Camera(args.load_vec3("pos"), args.load_vec3("target"), args.load_vec3("up", true), vfov, aspect);
}

// create_sky_light (light.cpp:258-272): `ambient = [r, g, b]`, or an equirectangular HDR map
// (`map` = path to .hdr or .exr, `to_world` transform of which the linear part rotates the map,
// `transform_y_up` (default true: theta measured from +y), `strength` (default 1)). A scene holds at
// most one map; its texels go to `scene`.
static ks_light_desc load_sky_light(const ConfigArgs &sky, SceneData &scene) {
    ks_light_desc l{};
    if (sky.contains("ambient")) {
        l.type = KS_LIGHT_SKY_AMBIENT;
        auto a = sky.load_vec3("ambient");
        for (int k = 0; k < 3; ++k) l.color[k] = a[k];
        return l;
    }
    if (!scene.sky_rgb.empty()) throw ConfigError("at most one sky map per scene");
    fs::path map = sky.load_path("map");
    std::string ext = map.extension().string();
    for (char &c : ext) c = (char)std::tolower((unsigned char)c);
    if (ext == ".hdr") {
        scene.sky_rgb = read_hdr(map, scene.sky_width, scene.sky_height);
    } else if (ext == ".exr") {
        scene.sky_rgb = read_exr(map, scene.sky_width, scene.sky_height);
    } else {
        throw ConfigError("unsupported sky map format '" + map.string() + "' (only .hdr or .exr)");
    }
    scene.sky_to_world = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    if (sky.contains("to_world")) {
        std::array<float, 16> m = sky.load_transform("to_world");
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) scene.sky_to_world[3 * r + c] = m[4 * r + c];
    }
    scene.sky_y_up = sky.load_bool("transform_y_up", true);
    scene.sky_strength = sky.load_float("strength", 1.0f);
    l.type = KS_LIGHT_SKY_MAP;
    return l;
}

SmallPTSetup resolve_small_pt(const ConfigArgs &args) {
    SmallPTSetup st;
    ks_camera_desc scene_cam{};
    bool have_scene_cam = false;
    if (args.contains("object")) {
        st.scene_name = args.load_string("object");
        MeshAsset asset = create_mesh_asset(args.asset(st.scene_name));
        st.scene = scene_from_mesh_asset(asset, args);
    } else if (args.contains("compound_object")) {
        st.scene_name = args.load_string("compound_object");
        CompoundMeshAsset asset = create_compound_mesh_asset(args.asset(st.scene_name));
        st.scene = scene_from_compound(asset, args);
    } else if (args.contains("procedural")) {
        st.scene_name = args.load_string("procedural");
        int seed = args.load_integer("procedural_seed", st.scene_name.rfind("spheres", 0) == 0 ? 1 : 0);
        ks_host_scene *hs = nullptr;
        check(ks_host_scene_create(st.scene_name.c_str(), (uint64_t)seed, &hs), "ks_host_scene_create");
        st.scene = scene_from_host(hs, &scene_cam);
        ks_host_scene_free(hs);
        have_scene_cam = true;
    } else {
        throw ConfigError("the task needs a scene: object = \"mesh_asset.X\", compound_object = "
                          "\"compound_mesh_asset.X\" or procedural = \"...\"");
    }

    if (args.contains("sky") || args.contains("light")) {
        st.scene.lights.clear();
        st.scene.sky_rgb.clear();
        if (args.contains("sky")) st.scene.lights.push_back(load_sky_light(args["sky"], st.scene));
        if (args.contains("light")) { // create_light (light.cpp:237-245)
            ConfigArgs la = args["light"];
            for (size_t i = 0; i < la.array_size(); ++i) {
                ConfigArgs li = la[(int)i];
                std::string t = li.load_string("type");
                ks_light_desc l{};
                if (t == "directional") {
                    l.type = KS_LIGHT_DIRECTIONAL;
                    auto L = li.load_vec3("L");
                    auto d = li.load_vec3("dir", true);
                    for (int k = 0; k < 3; ++k) l.color[k] = L[k], l.vec[k] = d[k];
                } else if (t == "point") {
                    l.type = KS_LIGHT_POINT;
                    std::array<float, 3> I;
                    if (li.contains("I")) {
                        I = li.load_vec3("I");
                    } else {
                        auto c = li.load_vec3("color");
                        float pw = li.load_float("power");
                        for (int k = 0; k < 3; ++k) I[k] = c[k] * pw / (4.0f * 3.14159265359f);
                    }
                    auto pos = li.load_vec3("pos");
                    for (int k = 0; k < 3; ++k) l.color[k] = I[k], l.vec[k] = pos[k];
                } else if (t == "sky") {
                    l = load_sky_light(li, st.scene);
                } else {
                    throw ConfigError("light type '" + t + "' is not supported");
                }
                st.scene.lights.push_back(l);
            }
        }
    }
    if (args.contains("pixel_filter")) { // create_pixel_filter (render_target.cpp:28-44)
        ConfigArgs pf = args["pixel_filter"];
        std::string type = pf.load_string("type");
        PixelFilter &f = st.input.pixel_filter;
        if (type == "box") {
            f.type = KS_FILTER_BOX;
            f.width = pf.load_float("width", 1.0f);
        } else if (type == "gaussian") {
            f.type = KS_FILTER_GAUSSIAN;
            f.width = pf.load_float("width", 3.0f);
            f.sigma = pf.load_float("sigma", 0.5f);
        } else if (type == "blackman_harris") {
            f.type = KS_FILTER_BLACKMAN_HARRIS;
            f.width = pf.load_float("width", 3.0f);
        } else {
            throw ConfigError("Invalid pixel filter type [" + type + "]");
        }
        if (!(f.width > 0.0f)) throw ConfigError("pixel filter width must be positive");
    }
    if (args.contains("tone_mapper")) st.tone_mapper = create_tone_mapper(args["tone_mapper"]);
    st.output_pfm = args.load_bool("output_pfm", false);
    if (args.contains("camera_animation")) throw ConfigError("camera_animation is not supported");

    SmallPTInput &in = st.input;
    if (args.contains("backdrop")) {
        in.backdrop = args.load_vec3("backdrop");
        in.include_background = false;
    } else {
        in.include_background = true;
    }
    in.bounces = args.load_integer("bounces");
    in.clamp_direct = args.load_float("clamp_direct", 0.0f);
    in.clamp_indirect = args.load_float("clamp_indirect", 10.0f);
    in.render_width = args.load_integer("render_width");
    in.render_height = args.load_integer("render_height");
    in.crop_start_x = args.load_integer("crop_start_x", 0);
    in.crop_start_y = args.load_integer("crop_start_y", 0);
    in.crop_width = args.load_integer("crop_width", 0);
    in.crop_height = args.load_integer("crop_height", 0);
    in.spp = args.load_integer("spp");
    in.scale_ray_diff = args.load_bool("scale_ray_diff", true);
    in.rng_seed = args.load_integer("rng_seed", 0);
    in.spp_prog_interval = args.load_integer("spp_prog_interval", 32);
    std::string mis = args.load_string("mis_mode", "reference");
    if (mis == "reference") in.mis_mode = KS_MIS_REFERENCE;
    else if (mis == "consistent") in.mis_mode = KS_MIS_CONSISTENT;
    else throw ConfigError("mis_mode must be reference or consistent");
    in.queue_size = args.load_integer("queue_size", 0);
    const std::string builder = args.load_string("bvh_builder", "sbvh");
    if (builder == "sbvh") st.bvh_builder = KS_BVH_BUILD_HOST_SBVH;
    else if (builder == "device") st.bvh_builder = KS_BVH_BUILD_DEVICE;
    else throw ConfigError("bvh_builder must be \"sbvh\" or \"device\", not \"" + builder + "\"");
    if (args.contains("guide")) {
        ConfigArgs g = args["guide"];
        if (in.mis_mode != KS_MIS_CONSISTENT)
            throw ConfigError("the guide needs mis_mode = \"consistent\" (reference weights do not sum to one)");
        st.guided = true;
        in.guide_mode = KS_GUIDE_TRAIN_AND_SAMPLE;
        ks_guide_config &c = st.guide_cfg;
        // ks_guide_config reads 0 as "default 0.5": an explicit fraction = 0 is applied after creation
        c.guide_fraction = g.load_float("fraction", 0.5f);
        c.train_spp = g.load_integer("train_spp", 0);
        c.learning_rate = g.load_float("learning_rate", 0.0f);
        c.beta1 = g.load_float("beta1", 0.0f);
        c.beta2 = g.load_float("beta2", 0.0f);
        c.log2_table_size = g.load_integer("log2_table_size", 0);
        c.base_resolution = g.load_float("base_resolution", 0.0f);
        c.max_resolution = g.load_float("max_resolution", 0.0f);
        c.train_batch = g.load_integer("train_batch", 0);
        c.max_steps_per_pass = g.load_integer("max_steps_per_pass", 0);
        c.weight_clamp = g.load_float("weight_clamp", 0.0f);
        c.seed = (uint64_t)g.load_integer("seed", 0);
        c.max_guided_bounce = (int32_t)g.load_integer("max_bounce", 0);
        std::string prec = g.load_string("inference", "bf16");
        if (prec != "bf16" && prec != "fp8") throw ConfigError("guide inference must be \"bf16\" or \"fp8\"");
        c.inference_fp8 = prec == "fp8";
    }

    if (args.contains("camera")) {
        st.camera = load_camera(args["camera"]);
    } else if (have_scene_cam) {
        st.camera.desc = scene_cam;
    } else {
        throw ConfigError("missing table 'camera'");
    }
    return st;
}

std::string describe(const SmallPTSetup &st) {
    const SmallPTInput &in = st.input;
    const ks_camera_desc &c = st.camera.desc;
    char buf[2048];
    std::snprintf(buf, sizeof(buf),
                  "{\"scene\": \"%s\", \"n_tris\": %zu, \"n_vertices\": %zu, \"n_geoms\": %zu, \"n_materials\": %zu, "
                  "\"n_lights\": %zu, \"sky_map\": [%d, %d, %g, %d], \"bounces\": %d, \"render_width\": %d, "
                  "\"render_height\": %d, \"crop\": [%d, %d, %d, %d], \"spp\": %d, \"rng_seed\": %d, "
                  "\"include_background\": %s, \"backdrop\": [%g, %g, %g], \"clamp_direct\": %g, "
                  "\"clamp_indirect\": %g, \"spp_prog_interval\": %d, \"mis_mode\": %d, \"queue_size\": %d, "
                  "\"bvh_builder\": %d, \"pixel_filter\": {\"type\": %d, \"width\": %g, \"sigma\": %g}, "
                  "\"guided\": %s, \"guide\": {\"fraction\": %g, \"train_spp\": %d, \"learning_rate\": %g, "
                  "\"log2_table_size\": %d, \"inference_fp8\": %d, \"max_bounce\": %d}, "
                  "\"camera\": {\"use_matrix\": %d, \"pos\": [%g, %g, %g], \"target\": [%g, %g, %g], "
                  "\"up\": [%g, %g, %g], \"vfov_deg\": %.6g, \"aspect\": %.6g, \"to_world\": [%g, %g, %g, %g, %g, %g, "
                  "%g, %g, %g, %g, %g, %g, %g, %g, %g, %g]}}",
                  st.scene_name.c_str(), st.scene.indices.size() / 3, st.scene.vertices.size() / 3,
                  st.scene.geom_material.size(), st.scene.materials.size(), st.scene.lights.size(), st.scene.sky_width,
                  st.scene.sky_height, st.scene.sky_strength, st.scene.sky_y_up ? 1 : 0, in.bounces,
                  in.render_width, in.render_height, in.crop_start_x, in.crop_start_y, in.crop_width, in.crop_height,
                  in.spp, in.rng_seed, in.include_background ? "true" : "false", in.backdrop[0], in.backdrop[1],
                  in.backdrop[2], in.clamp_direct, in.clamp_indirect, in.spp_prog_interval, in.mis_mode,
                  in.queue_size, st.bvh_builder, in.pixel_filter.type, in.pixel_filter.width, in.pixel_filter.sigma,
                  st.guided ? "true" : "false", st.guide_cfg.guide_fraction, st.guide_cfg.train_spp,
                  st.guide_cfg.learning_rate, st.guide_cfg.log2_table_size, st.guide_cfg.inference_fp8, st.guide_cfg.max_guided_bounce, c.use_matrix, c.pos[0], c.pos[1], c.pos[2], c.target[0], c.target[1], c.target[2],
                  c.up[0], c.up[1], c.up[2], c.vfov_deg, c.aspect, c.to_world[0], c.to_world[1], c.to_world[2],
                  c.to_world[3], c.to_world[4], c.to_world[5], c.to_world[6], c.to_world[7], c.to_world[8],
                  c.to_world[9], c.to_world[10], c.to_world[11], c.to_world[12], c.to_world[13], c.to_world[14],
                  c.to_world[15]);
    return buf;
}

static int g_device_ordinal = 0;
void set_task_device(int ordinal) { g_device_ordinal = ordinal; }

void small_pt(const ConfigArgs &args, const fs::path &task_dir, int task_id) {
    SmallPTSetup st = resolve_small_pt(args);
    Device dev(g_device_ordinal);
    auto b0 = std::chrono::steady_clock::now();
    ks_bvh_options bvh{};
    bvh.builder = st.bvh_builder;
    Scene scene(dev, st.scene, &bvh);
    double build_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - b0).count();
    RenderTarget2 rt(dev, (uint32_t)st.input.render_width, (uint32_t)st.input.render_height, st.input.backdrop);
    SmallPTInput in = st.input;
    std::unique_ptr<Guide> guide;
    if (st.guided) {
        guide = std::make_unique<Guide>(dev, scene, st.guide_cfg);
        if (st.guide_cfg.guide_fraction == 0.0f)
            guide->set_mode(0.0f, st.guide_cfg.max_steps_per_pass ? st.guide_cfg.max_steps_per_pass : 4);
        in.guide = guide.get();
    }
    in.scene = &scene;
    in.camera = &st.camera;
    in.target = &rt;
    std::printf("Frame [1/1] | Start (task %d, scene %s, %zu triangles, build %.2f s)\n", task_id,
                st.scene_name.c_str(), st.scene.indices.size() / 3, build_s);
    auto render_start = std::chrono::steady_clock::now();
    auto interval_start = render_start;
    in.prog_interval_callback = [&](const RenderTarget2 &target, int spp_finished) { // small_pt.cpp:278-294
        auto now = std::chrono::steady_clock::now();
        std::printf("Frame [1/1] | [%d/%d] spp interval took: %.3f sec\n", spp_finished, in.spp,
                    std::chrono::duration<double>(now - interval_start).count());
        // composite_and_save_to_exr (FLOAT) and, with a tone mapper, composite_and_save_to_png
        std::string prefix = "small_pt_spp" + std::to_string(spp_finished);
        std::vector<float> rgb = target.composite();
        write_exr(task_dir / (prefix + ".exr"), rgb.data(), (int)target.width, (int)target.height, 3, false);
        if (st.tone_mapper) {
            std::vector<double> sums, weights;
            target.download(sums, weights);
            float bd[3] = {in.backdrop[0], in.backdrop[1], in.backdrop[2]};
            std::vector<uint8_t> png =
                composite_png(sums.data(), weights.data(), weights.size() / 2, bd, st.tone_mapper.get());
            write_png(task_dir / (prefix + ".png"), png.data(), (int)target.width, (int)target.height, 3);
        }
        if (st.output_pfm)
            write_pfm(task_dir / (prefix + ".pfm"), rgb.data(), target.width, target.height);
        interval_start = std::chrono::steady_clock::now();
    };
    SmallPT pt;
    ks_render_stats s = pt.run(in);
    double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - render_start).count();
    std::printf("Frame [1/1] | Rendering time: %.3f sec (%.1f Mpaths/s, %llu + %llu rays)\n", secs,
                (double)s.paths / s.wall_ms / 1e3, (unsigned long long)s.rays_closest,
                (unsigned long long)s.rays_shadow);
    std::ofstream js(task_dir / "render_stats.json");
    js << "{\"render_seconds\": " << secs << ", \"render_ms_library\": " << s.wall_ms << ", \"paths\": " << s.paths
       << ", \"rays_closest\": " << s.rays_closest << ", \"rays_shadow\": " << s.rays_shadow
       << ", \"iterations\": " << s.iterations << ", \"nonfinite\": " << s.nonfinite
       << ", \"guide_samples\": " << s.guide_samples
       << ", \"scene_build_seconds\": " << build_s << ", \"setup\": " << describe(st) << "}\n";
}

} // namespace ks

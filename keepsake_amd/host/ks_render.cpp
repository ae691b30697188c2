// ks_render: the task-runner executable (test/main.cpp:28-84 of the reference) for the gfx950 small_pt,
// plus the convert_mesh_asset (OBJ -> binary mesh asset) and tonemap (EXR / HDR -> tone-mapped EXR) tasks.
//   ks_render -c scene.toml [-d device] [-r asset_root_dir] [-n threads] [--print-config]
// --print-config resolves every small_pt task of the file and prints the inputs as JSON lines without
// opening a GPU; --dump-scene DIR also writes each resolved scene's flat arrays as raw little-endian
// files under DIR/task_<id>/ (vertices.f32, normals.f32, indices.u32, tri_geom.u32,
// geom_material.u32, geom_flags.u8, materials.bin of ks_material_desc records, lights.bin of
// ks_light_desc records, texcoords.f32, field_of.i32 and fields.bin of ks_material_fields records for
// textured materials, textures.txt with one "width height channels data_type color_space levels
// build_mipmaps" line per texture and its levels in texture_<i>.bin).
#include "image_io.h"
#include "mesh_asset.h"
#include "small_pt_task.h"
#include "texture_asset.h"

#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>

static void usage() {
    std::fprintf(stderr, "usage: ks_render -c config.toml [-d device] [-r asset_root] [-n threads] [--print-config] "
                         "[--dump-scene dir]\n");
}

template <typename T> static void dump_array(const ks::fs::path &p, const std::vector<T> &v) {
    std::ofstream f(p, std::ios::binary);
    f.write(reinterpret_cast<const char *>(v.data()), (std::streamsize)(v.size() * sizeof(T)));
    if (!f) throw ks::ConfigError("cannot write " + p.string());
}

static void dump_scene(const ks::SceneData &s, const ks::fs::path &dir) {
    ks::fs::create_directories(dir);
    dump_array(dir / "vertices.f32", s.vertices);
    dump_array(dir / "normals.f32", s.normals);
    dump_array(dir / "indices.u32", s.indices);
    dump_array(dir / "tri_geom.u32", s.tri_geom);
    dump_array(dir / "geom_material.u32", s.geom_material);
    dump_array(dir / "geom_flags.u8", s.geom_flags);
    dump_array(dir / "materials.bin", s.materials);
    dump_array(dir / "lights.bin", s.lights);
    dump_array(dir / "texcoords.f32", s.texcoords);
    dump_array(dir / "field_of.i32", s.field_of);
    dump_array(dir / "fields.bin", s.fields);
    std::ofstream t(dir / "textures.txt");
    for (size_t i = 0; i < s.textures.size(); ++i) {
        const ks_texture_desc &d = s.textures[i];
        t << d.width << " " << d.height << " " << d.channels << " " << d.data_type << " " << d.color_space << " "
          << d.levels << " " << d.build_mipmaps << "\n";
        dump_array(dir / ("texture_" + std::to_string(i) + ".bin"), s.texture_bytes[i]);
    }
}

int main(int argc, char **argv) {
    std::string cfg_path, root, dump_dir;
    int device = 0;
    bool print_only = false;
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        auto next = [&]() -> std::string {
            if (i + 1 >= argc) {
                usage();
                std::exit(2);
            }
            return argv[++i];
        };
        if (a == "-c" || a == "--config") cfg_path = next();
        else if (a == "-d" || a == "--device") device = std::stoi(next());
        else if (a == "-r" || a == "--asset_root_dir") root = next();
        else if (a == "-n" || a == "--nthreads") (void)next(); // host threads: the GPU path needs none
        else if (a == "-v" || a == "--validation") (void)next(); // Vulkan validation: not applicable
        else if (a == "--print-config") print_only = true;
        else if (a == "--dump-scene") print_only = true, dump_dir = next();
        else {
            usage();
            return 2;
        }
    }
    if (cfg_path.empty()) {
        usage();
        return 2;
    }
    try {
        ks::ConfigService cfg;
        if (!root.empty()) cfg.set_asset_root_dir(root);
        if (print_only) {
            std::ifstream f(cfg_path);
            if (!f) throw ks::ConfigError("cannot open " + cfg_path);
            std::stringstream ss;
            ss << f.rdbuf();
            cfg.cfg = ks::parse_toml(ss.str(), cfg_path);
            std::vector<ks::TomlValue> tables = cfg.task_tables();
            for (size_t id = 0; id < tables.size(); ++id) {
                const ks::TomlValue &t = tables[id];
                const ks::TomlValue *type = t.find("type");
                if (!type || type->str != "small_pt") continue;
                ks::SmallPTSetup st = ks::resolve_small_pt(ks::ConfigArgs(&t, root, &cfg.cfg));
                std::printf("%s\n", ks::describe(st).c_str());
                if (!dump_dir.empty()) dump_scene(st.scene, ks::fs::path(dump_dir) / ("task_" + std::to_string(id)));
            }
            return 0;
        }
        ks::set_task_device(device);
        cfg.parse_file(cfg_path);
        cfg.register_task("small_pt", ks::small_pt);
        cfg.register_task("convert_mesh_asset", ks::convert_mesh_asset);
        cfg.register_task("tonemap", ks::tonemap_task);
        cfg.register_task("convert_texture", ks::convert_texture);
        cfg.run_all_tasks();
    } catch (const ks::ConfigError &e) {
        std::fprintf(stderr, "[ks_render] config error: %s\n", e.what());
        return 1;
    } catch (const std::exception &e) {
        std::fprintf(stderr, "[ks_render] error: %s\n", e.what());
        return 1;
    }
    return 0;
}

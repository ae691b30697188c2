// ks_render: the task-runner executable (test/main.cpp:28-84 of the reference) for the gfx950 small_pt.
//   ks_render -c scene.toml [-d device] [-r asset_root_dir] [-n threads] [--print-config]
// --print-config resolves every small_pt task of the file and prints the inputs as JSON lines without
// opening a GPU.
#include "small_pt_task.h"

#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>

static void usage() {
    std::fprintf(stderr, "usage: ks_render -c config.toml [-d device] [-r asset_root] [-n threads] [--print-config]\n");
}

int main(int argc, char **argv) {
    std::string cfg_path, root;
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
            for (const auto &t : cfg.task_tables()) {
                const ks::TomlValue *type = t.find("type");
                if (type && type->str == "small_pt")
                    std::printf("%s\n", ks::describe(ks::resolve_small_pt(ks::ConfigArgs(&t, root))).c_str());
            }
            return 0;
        }
        ks::set_task_device(device);
        cfg.parse_file(cfg_path);
        cfg.register_task("small_pt", ks::small_pt);
        cfg.run_all_tasks();
    } catch (const ks::ConfigError &e) {
        std::fprintf(stderr, "[ks_render] config error: %s\n", e.what());
        return 1;
    }
    return 0;
}

// TOML configuration and task runner of the keepsake_amd host tools, mirroring the reference's
// ConfigArgs / ConfigService (config.h:17-140) over a self-contained TOML subset parser (toml++ is not
// available): tables, arrays of tables, dotted keys, strings, integers, floats, booleans, arrays and
// inline tables.
#pragma once

#include <array>
#include <filesystem>
#include <functional>
#include <map>
#include <memory>
#include <optional>
#include <stdexcept>
#include <string>
#include <string_view>
#include <vector>

namespace ks {
namespace fs = std::filesystem;

struct TomlValue {
    enum class Type { None, String, Integer, Float, Bool, Array, Table };
    Type type = Type::None;
    std::string str;
    int64_t i = 0;
    double f = 0.0;
    bool b = false;
    std::vector<TomlValue> arr;
    std::vector<std::pair<std::string, TomlValue>> table; // insertion order kept

    bool is_table() const { return type == Type::Table; }
    bool is_array() const { return type == Type::Array; }
    bool is_number() const { return type == Type::Integer || type == Type::Float; }
    double number() const { return type == Type::Integer ? (double)i : f; }
    const TomlValue *find(std::string_view key) const;
    TomlValue *find(std::string_view key);
    TomlValue &insert_or_assign(const std::string &key, TomlValue v);
    std::string dump(int indent = 0) const;
};

struct ConfigError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

TomlValue parse_toml(std::string_view text, const std::string &source = "<string>");

struct ConfigArgs {
    ConfigArgs(const TomlValue *node, fs::path asset_root, const TomlValue *doc = nullptr)
        : node(node), asset_root(std::move(asset_root)), doc(doc) {}

    ConfigArgs operator[](std::string_view key) const;
    bool contains(std::string_view key) const;
    ConfigArgs operator[](int idx) const;
    size_t array_size() const;

    int load_integer(std::string_view name, const std::optional<int> &default_value = {}) const;
    float load_float(std::string_view name, const std::optional<float> &default_value = {}) const;
    std::array<float, 3> load_vec3(std::string_view name, bool force_normalize = false,
                                   const std::optional<std::array<float, 3>> &default_value = {}) const;
    std::array<float, 16> load_transform(std::string_view name) const; // row-major 4x4
    bool load_bool(std::string_view name, const std::optional<bool> &default_value = {}) const;
    std::string load_string(std::string_view name, const std::optional<std::string> &default_value = {}) const;
    fs::path load_path(std::string_view name) const;
    // A named asset table of the whole document, "prefix.name" -> [prefix.name] (the reference's
    // ConfigurableTable keys, config.cpp:315-329).
    ConfigArgs asset(std::string_view path) const;

    const TomlValue *node;
    fs::path asset_root;
    const TomlValue *doc; // the whole configuration, for asset lookups
};

// (args, task_dir, task_id)
using ConfigTask = std::function<void(const ConfigArgs &args, const fs::path &, int)>;

struct ConfigService {
    void set_asset_root_dir(const fs::path &root) { asset_root = root; }
    void parse_file(const fs::path &file_path);
    void parse(std::string_view str);
    void register_task(std::string_view name, const ConfigTask &task) { tasks[std::string(name)] = task; }
    fs::path output_directory() const { return output_dir; }
    void run_all_tasks() const; // config.cpp:262-313
    // The [[task]] tables with `override` applied (config.cpp:284-296), in file order.
    std::vector<TomlValue> task_tables() const;

    TomlValue cfg;
    fs::path asset_root;
    fs::path output_dir;
    std::map<std::string, ConfigTask> tasks;

  private:
    void after_parse();
};

// The small_pt task (small_pt.cpp:180-303) running on the GPU library.
void small_pt(const ConfigArgs &args, const fs::path &task_dir, int task_id);

// Writes an RGB float image as PFM (little-endian, bottom-to-top rows).
void write_pfm(const fs::path &path, const float *rgb, uint32_t width, uint32_t height);

} // namespace ks

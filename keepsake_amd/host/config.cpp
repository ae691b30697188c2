// TOML subset parser, ConfigArgs loaders and the [[task]] runner (config.cpp:39-81, 204-313 of the
// reference describe the behaviour mirrored here).
#include "config.h"

#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <ctime>
#include <fstream>
#include <iostream>
#include <sstream>

namespace ks {

const TomlValue *TomlValue::find(std::string_view key) const {
    if (type != Type::Table) return nullptr;
    for (const auto &kv : table)
        if (kv.first == key) return &kv.second;
    return nullptr;
}
TomlValue *TomlValue::find(std::string_view key) {
    if (type != Type::Table) return nullptr;
    for (auto &kv : table)
        if (kv.first == key) return &kv.second;
    return nullptr;
}
TomlValue &TomlValue::insert_or_assign(const std::string &key, TomlValue v) {
    if (TomlValue *e = find(key)) {
        *e = std::move(v);
        return *e;
    }
    table.emplace_back(key, std::move(v));
    return table.back().second;
}

static std::string quote(const std::string &s) {
    std::string o = "\"";
    for (char c : s) {
        if (c == '"' || c == '\\') o += '\\';
        if (c == '\n') {
            o += "\\n";
            continue;
        }
        o += c;
    }
    return o + "\"";
}

std::string TomlValue::dump(int indent) const {
    std::ostringstream os;
    switch (type) {
    case Type::None: return "\"\"";
    case Type::String: return quote(str);
    case Type::Integer: return std::to_string(i);
    case Type::Float: {
        char buf[64];
        std::snprintf(buf, sizeof(buf), "%.9g", f);
        std::string s = buf;
        if (s.find_first_of(".eEni") == std::string::npos) s += ".0";
        return s;
    }
    case Type::Bool: return b ? "true" : "false";
    case Type::Array: {
        os << "[";
        for (size_t k = 0; k < arr.size(); ++k) os << (k ? ", " : "") << arr[k].dump(indent);
        os << "]";
        return os.str();
    }
    case Type::Table: {
        os << "{ ";
        for (size_t k = 0; k < table.size(); ++k)
            os << (k ? ", " : "") << table[k].first << " = " << table[k].second.dump(indent);
        os << " }";
        return os.str();
    }
    }
    return "";
}

namespace {

struct Parser {
    std::string_view s;
    size_t p = 0;
    int line = 1;
    std::string source;

    [[noreturn]] void fail(const std::string &msg) const {
        throw ConfigError(source + ":" + std::to_string(line) + ": " + msg);
    }
    bool eof() const { return p >= s.size(); }
    char peek(size_t k = 0) const { return p + k < s.size() ? s[p + k] : '\0'; }
    void skip_ws() {
        while (!eof() && (peek() == ' ' || peek() == '\t' || peek() == '\r')) ++p;
    }
    void skip_comment() {
        if (peek() == '#')
            while (!eof() && peek() != '\n') ++p;
    }
    // whitespace, newlines and comments (inside arrays)
    void skip_all() {
        while (!eof()) {
            skip_ws();
            skip_comment();
            if (peek() == '\n') {
                ++p;
                ++line;
            } else {
                break;
            }
        }
    }
    void end_of_line() {
        skip_ws();
        skip_comment();
        if (eof()) return;
        if (peek() != '\n') fail(std::string("unexpected character '") + peek() + "'");
        ++p;
        ++line;
    }

    std::string parse_basic_string() {
        ++p; // opening quote
        std::string out;
        while (true) {
            if (eof() || peek() == '\n') fail("unterminated string");
            char c = s[p++];
            if (c == '"') break;
            if (c == '\\') {
                char e = s[p++];
                switch (e) {
                case 'n': out += '\n'; break;
                case 't': out += '\t'; break;
                case 'r': out += '\r'; break;
                case '"': out += '"'; break;
                case '\\': out += '\\'; break;
                default: fail(std::string("unsupported escape \\") + e);
                }
            } else {
                out += c;
            }
        }
        return out;
    }
    std::string parse_literal_string() {
        ++p;
        size_t q = s.find('\'', p);
        if (q == std::string_view::npos) fail("unterminated literal string");
        std::string out(s.substr(p, q - p));
        p = q + 1;
        return out;
    }
    std::string parse_key_part() {
        skip_ws();
        if (peek() == '"') return parse_basic_string();
        if (peek() == '\'') return parse_literal_string();
        size_t b = p;
        while (!eof() && (std::isalnum((unsigned char)peek()) || peek() == '_' || peek() == '-')) ++p;
        if (b == p) fail("expected a key");
        return std::string(s.substr(b, p - b));
    }
    std::vector<std::string> parse_key() {
        std::vector<std::string> parts{parse_key_part()};
        skip_ws();
        while (peek() == '.') {
            ++p;
            parts.push_back(parse_key_part());
            skip_ws();
        }
        return parts;
    }

    TomlValue parse_value() {
        skip_ws();
        TomlValue v;
        char c = peek();
        if (c == '"') {
            v.type = TomlValue::Type::String;
            v.str = parse_basic_string();
        } else if (c == '\'') {
            v.type = TomlValue::Type::String;
            v.str = parse_literal_string();
        } else if (c == '[') {
            ++p;
            v.type = TomlValue::Type::Array;
            skip_all();
            while (peek() != ']') {
                v.arr.push_back(parse_value());
                skip_all();
                if (peek() == ',') {
                    ++p;
                    skip_all();
                } else if (peek() != ']') {
                    fail("expected ',' or ']' in array");
                }
            }
            ++p;
        } else if (c == '{') {
            ++p;
            v.type = TomlValue::Type::Table;
            skip_ws();
            while (peek() != '}') {
                auto key = parse_key();
                skip_ws();
                if (peek() != '=') fail("expected '=' in inline table");
                ++p;
                assign(v, key, parse_value());
                skip_ws();
                if (peek() == ',') {
                    ++p;
                    skip_ws();
                } else if (peek() != '}') {
                    fail("expected ',' or '}' in inline table");
                }
            }
            ++p;
        } else if (s.substr(p, 4) == "true") {
            p += 4;
            v.type = TomlValue::Type::Bool;
            v.b = true;
        } else if (s.substr(p, 5) == "false") {
            p += 5;
            v.type = TomlValue::Type::Bool;
            v.b = false;
        } else {
            size_t b = p;
            while (!eof() && (std::isalnum((unsigned char)peek()) || std::strchr("+-._", peek()))) ++p;
            std::string tok;
            for (char ch : s.substr(b, p - b))
                if (ch != '_') tok += ch;
            if (tok.empty()) fail("expected a value");
            bool is_float = tok.find_first_of(".eE") != std::string::npos || tok.find("inf") != std::string::npos ||
                            tok.find("nan") != std::string::npos;
            try {
                if (is_float) {
                    v.type = TomlValue::Type::Float;
                    v.f = std::stod(tok);
                } else {
                    v.type = TomlValue::Type::Integer;
                    size_t used = 0;
                    v.i = std::stoll(tok, &used, 0);
                    if (used != tok.size()) fail("bad number '" + tok + "'");
                }
            } catch (const std::logic_error &) {
                fail("bad value '" + tok + "'");
            }
        }
        return v;
    }

    void assign(TomlValue &tbl, const std::vector<std::string> &key, TomlValue v) {
        TomlValue *t = &tbl;
        for (size_t k = 0; k + 1 < key.size(); ++k) {
            TomlValue *n = t->find(key[k]);
            if (!n) {
                TomlValue nt;
                nt.type = TomlValue::Type::Table;
                n = &t->insert_or_assign(key[k], nt);
            }
            if (!n->is_table()) fail("key '" + key[k] + "' is not a table");
            t = n;
        }
        if (t->find(key.back())) fail("duplicate key '" + key.back() + "'");
        t->insert_or_assign(key.back(), std::move(v));
    }

    TomlValue *descend(TomlValue &root, const std::vector<std::string> &key, bool array_of_tables) {
        TomlValue *t = &root;
        for (size_t k = 0; k < key.size(); ++k) {
            TomlValue *n = t->find(key[k]);
            bool last = k + 1 == key.size();
            if (!n) {
                TomlValue nt;
                nt.type = (last && array_of_tables) ? TomlValue::Type::Array : TomlValue::Type::Table;
                n = &t->insert_or_assign(key[k], nt);
            }
            if (n->is_array()) {
                if (last && array_of_tables) {
                    TomlValue nt;
                    nt.type = TomlValue::Type::Table;
                    n->arr.push_back(nt);
                    return &n->arr.back();
                }
                if (n->arr.empty() || !n->arr.back().is_table()) fail("key '" + key[k] + "' is not a table");
                t = &n->arr.back();
            } else if (n->is_table()) {
                if (last && array_of_tables) fail("key '" + key[k] + "' is a table, not an array of tables");
                t = n;
            } else {
                fail("key '" + key[k] + "' is not a table");
            }
        }
        return t;
    }

    TomlValue parse_document() {
        TomlValue root;
        root.type = TomlValue::Type::Table;
        TomlValue *cur = &root;
        while (true) {
            skip_all();
            if (eof()) break;
            if (peek() == '[') {
                bool aot = peek(1) == '[';
                p += aot ? 2 : 1;
                auto key = parse_key();
                skip_ws();
                if (peek() != ']' || (aot && peek(1) != ']')) fail("expected ']' after table header");
                p += aot ? 2 : 1;
                cur = descend(root, key, aot);
                end_of_line();
            } else {
                auto key = parse_key();
                skip_ws();
                if (peek() != '=') fail("expected '='");
                ++p;
                assign(*cur, key, parse_value());
                end_of_line();
            }
        }
        return root;
    }
};

const TomlValue &require(const TomlValue *node, std::string_view name) {
    const TomlValue *v = node ? node->find(name) : nullptr;
    if (!v) throw ConfigError("missing key '" + std::string(name) + "'");
    return *v;
}

} // namespace

TomlValue parse_toml(std::string_view text, const std::string &source) {
    Parser ps{text, 0, 1, source};
    return ps.parse_document();
}

ConfigArgs ConfigArgs::operator[](std::string_view key) const {
    return ConfigArgs(&require(node, key), asset_root, doc);
}
ConfigArgs ConfigArgs::asset(std::string_view path) const {
    size_t dot = path.find('.');
    if (!doc || dot == std::string_view::npos) throw ConfigError("unknown asset '" + std::string(path) + "'");
    const TomlValue *group = doc->find(path.substr(0, dot));
    const TomlValue *a = group && group->is_table() ? group->find(path.substr(dot + 1)) : nullptr;
    if (!a || !a->is_table()) throw ConfigError("unknown asset '" + std::string(path) + "'");
    return ConfigArgs(a, asset_root, doc);
}
bool ConfigArgs::contains(std::string_view key) const { return node && node->find(key) != nullptr; }
ConfigArgs ConfigArgs::operator[](int idx) const {
    if (!node || !node->is_array() || idx < 0 || (size_t)idx >= node->arr.size())
        throw ConfigError("array index " + std::to_string(idx) + " out of range");
    return ConfigArgs(&node->arr[idx], asset_root, doc);
}
size_t ConfigArgs::array_size() const { return node && node->is_array() ? node->arr.size() : 0; }

int ConfigArgs::load_integer(std::string_view name, const std::optional<int> &def) const {
    const TomlValue *v = node ? node->find(name) : nullptr;
    if (!v) {
        if (def) return *def;
        throw ConfigError("missing integer '" + std::string(name) + "'");
    }
    if (v->type != TomlValue::Type::Integer) throw ConfigError("'" + std::string(name) + "' must be an integer");
    return (int)v->i;
}
float ConfigArgs::load_float(std::string_view name, const std::optional<float> &def) const {
    const TomlValue *v = node ? node->find(name) : nullptr;
    if (!v) {
        if (def) return *def;
        throw ConfigError("missing float '" + std::string(name) + "'");
    }
    if (!v->is_number()) throw ConfigError("'" + std::string(name) + "' must be a number");
    return (float)v->number();
}
std::array<float, 3> ConfigArgs::load_vec3(std::string_view name, bool force_normalize,
                                           const std::optional<std::array<float, 3>> &def) const {
    const TomlValue *v = node ? node->find(name) : nullptr;
    if (!v) {
        if (def) return *def;
        throw ConfigError("missing vec3 '" + std::string(name) + "'");
    }
    if (!v->is_array() || v->arr.size() != 3) throw ConfigError("'" + std::string(name) + "' must be [x, y, z]");
    std::array<float, 3> out;
    for (int k = 0; k < 3; ++k) {
        if (!v->arr[k].is_number()) throw ConfigError("'" + std::string(name) + "' must hold numbers");
        out[k] = (float)v->arr[k].number();
    }
    if (force_normalize) { // config.cpp:146-147
        float l = std::sqrt(out[0] * out[0] + out[1] * out[1] + out[2] * out[2]);
        if (l > 0.0f)
            for (float &c : out) c /= l;
    }
    return out;
}
bool ConfigArgs::load_bool(std::string_view name, const std::optional<bool> &def) const {
    const TomlValue *v = node ? node->find(name) : nullptr;
    if (!v) {
        if (def) return *def;
        throw ConfigError("missing bool '" + std::string(name) + "'");
    }
    if (v->type != TomlValue::Type::Bool) throw ConfigError("'" + std::string(name) + "' must be a bool");
    return v->b;
}
std::string ConfigArgs::load_string(std::string_view name, const std::optional<std::string> &def) const {
    const TomlValue *v = node ? node->find(name) : nullptr;
    if (!v) {
        if (def) return *def;
        throw ConfigError("missing string '" + std::string(name) + "'");
    }
    if (v->type != TomlValue::Type::String) throw ConfigError("'" + std::string(name) + "' must be a string");
    return v->str;
}
fs::path ConfigArgs::load_path(std::string_view name) const { return asset_root / load_string(name); }

// load_transform_field (config.cpp:204-260): matrix_rm, matrix_cm, or scale / rotation (XYZ euler
// degrees or quaternion w, x, y, z) / translation composed as T * R * S.
std::array<float, 16> ConfigArgs::load_transform(std::string_view name) const {
    const TomlValue &t = require(node, name);
    if (!t.is_table()) throw ConfigError("'" + std::string(name) + "' must be a table");
    std::array<float, 16> m{};
    auto nums = [&](const TomlValue &a, size_t n, const char *what) {
        if (!a.is_array() || a.arr.size() != n) throw ConfigError(std::string(what) + " has the wrong length");
        std::vector<double> o;
        for (const auto &x : a.arr) o.push_back(x.number());
        return o;
    };
    if (const TomlValue *rm = t.find("matrix_rm")) {
        auto e = nums(*rm, 16, "matrix_rm");
        for (int k = 0; k < 16; ++k) m[k] = (float)e[k];
        return m;
    }
    if (const TomlValue *cm = t.find("matrix_cm")) {
        auto e = nums(*cm, 16, "matrix_cm");
        for (int c = 0; c < 4; ++c)
            for (int r = 0; r < 4; ++r) m[4 * r + c] = (float)e[4 * c + r];
        return m;
    }
    double S[3] = {1, 1, 1}, T[3] = {0, 0, 0}, R[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    if (const TomlValue *s = t.find("scale")) {
        auto e = nums(*s, 3, "scale");
        for (int k = 0; k < 3; ++k) S[k] = e[k];
    }
    if (const TomlValue *tr = t.find("translation")) {
        auto e = nums(*tr, 3, "translation");
        for (int k = 0; k < 3; ++k) T[k] = e[k];
    }
    if (const TomlValue *r = t.find("rotation")) {
        double w, x, y, z;
        if (const TomlValue *eu = r->find("euler")) {
            // Eigen::EulerAnglesXYZ: q = Rx(a0) * Ry(a1) * Rz(a2)
            auto e = nums(*eu, 3, "euler");
            double h[3];
            for (int k = 0; k < 3; ++k) h[k] = e[k] / 180.0 * 3.14159265358979323846 * 0.5;
            double cx = std::cos(h[0]), sx = std::sin(h[0]), cy = std::cos(h[1]), sy = std::sin(h[1]),
                   cz = std::cos(h[2]), sz = std::sin(h[2]);
            // qx * qy
            double aw = cx * cy, ax = sx * cy, ay = cx * sy, az = sx * sy;
            // (qx*qy) * qz
            w = aw * cz - az * sz;
            x = ax * cz + ay * sz;
            y = ay * cz - ax * sz;
            z = az * cz + aw * sz;
        } else if (const TomlValue *q = r->find("quat")) {
            auto e = nums(*q, 4, "quat");
            double l = std::sqrt(e[0] * e[0] + e[1] * e[1] + e[2] * e[2] + e[3] * e[3]);
            w = e[0] / l, x = e[1] / l, y = e[2] / l, z = e[3] / l;
        } else {
            throw ConfigError("rotation needs euler or quat");
        }
        R[0][0] = 1 - 2 * (y * y + z * z), R[0][1] = 2 * (x * y - z * w), R[0][2] = 2 * (x * z + y * w);
        R[1][0] = 2 * (x * y + z * w), R[1][1] = 1 - 2 * (x * x + z * z), R[1][2] = 2 * (y * z - x * w);
        R[2][0] = 2 * (x * z - y * w), R[2][1] = 2 * (y * z + x * w), R[2][2] = 1 - 2 * (x * x + y * y);
    }
    for (int r = 0; r < 3; ++r) {
        for (int c = 0; c < 3; ++c) m[4 * r + c] = (float)(R[r][c] * S[c]);
        m[4 * r + 3] = (float)T[r];
    }
    m[15] = 1.0f;
    return m;
}

void ConfigService::after_parse() { // config.cpp:39-59
    if (const TomlValue *o = cfg.find("output_dir")) {
        output_dir = o->str;
    } else {
        char buf[64];
        std::time_t t = std::time(nullptr);
        std::strftime(buf, sizeof(buf), "%Y-%m-%d_%H-%M-%S", std::localtime(&t));
        output_dir = buf;
    }
    std::error_code ec;
    if (!fs::is_directory(output_dir)) {
        if (!fs::create_directories(output_dir, ec)) {
            std::printf("Failed to create directory %s\n", output_dir.string().c_str());
            return;
        }
        std::printf("Created output directory [%s]\n", output_dir.string().c_str());
    }
}

void ConfigService::parse_file(const fs::path &file_path) {
    std::ifstream f(file_path);
    if (!f) throw ConfigError("cannot open " + file_path.string());
    std::stringstream ss;
    ss << f.rdbuf();
    cfg = parse_toml(ss.str(), file_path.string());
    after_parse();
}

void ConfigService::parse(std::string_view str) {
    cfg = parse_toml(str);
    after_parse();
}

std::vector<TomlValue> ConfigService::task_tables() const {
    std::vector<TomlValue> out;
    const TomlValue *tasks_v = cfg.find("task");
    if (!tasks_v || !tasks_v->is_array()) return out;
    const auto &task_array = tasks_v->arr;
    for (int task_id = 0; task_id < (int)task_array.size(); ++task_id) {
        TomlValue task_table = task_array[task_id];
        if (const TomlValue *ov = task_table.find("override")) {
            int base = (int)ov->i;
            if (ov->type != TomlValue::Type::Integer || base < 0 || base >= task_id)
                throw ConfigError("override must name an earlier task");
            TomlValue merged = task_array[base];
            if (merged.find("override")) throw ConfigError("an overridden task cannot itself override");
            for (const auto &kv : task_table.table)
                if (kv.first != "override") merged.insert_or_assign(kv.first, kv.second);
            task_table = std::move(merged);
        }
        out.push_back(std::move(task_table));
    }
    return out;
}

void ConfigService::run_all_tasks() const { // config.cpp:262-313
    std::vector<TomlValue> tables = task_tables();
    if (tables.empty()) {
        std::printf("No task to run\n");
        return;
    }
    for (int task_id = 0; task_id < (int)tables.size(); ++task_id) {
        const TomlValue &task_table = tables[task_id];
        fs::path task_dir;
        if (const TomlValue *d = task_table.find("task_dir")) task_dir = d->str;
        else task_dir = "task_" + std::to_string(task_id);
        task_dir = output_dir / task_dir;
        std::error_code ec;
        if (!fs::is_directory(task_dir) && !fs::create_directories(task_dir, ec)) {
            std::printf("Failed to create directory %s\n", task_dir.string().c_str());
            return;
        }
        std::printf("Next task: \n%s\n\n", task_table.dump().c_str());
        {
            std::ofstream w(task_dir / "config.toml");
            for (const auto &kv : task_table.table) w << kv.first << " = " << kv.second.dump() << "\n";
        }
        const TomlValue *type = task_table.find("type");
        if (!type) throw ConfigError("task without type");
        auto it = tasks.find(type->str);
        if (it == tasks.end()) throw ConfigError("unknown task type '" + type->str + "'");
        it->second(ConfigArgs(&task_table, asset_root, &cfg), task_dir, task_id);
        std::printf("Saving output to [%s]\n\n", task_dir.string().c_str());
    }
}

void write_pfm(const fs::path &path, const float *rgb, uint32_t width, uint32_t height) {
    std::ofstream f(path, std::ios::binary);
    f << "PF\n" << width << " " << height << "\n-1.0\n";
    for (int y = (int)height - 1; y >= 0; --y) f.write((const char *)(rgb + (size_t)y * width * 3), width * 12);
}

} // namespace ks

// glTF 2.0 compound mesh assets (mesh_asset.cpp:311-1129) with a small JSON reader of its own.
#include "gltf_asset.h"

#include "image_io.h"

#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iterator>
#include <sstream>

namespace ks {
namespace {

// ---- JSON ---------------------------------------------------------------------------------------

struct Json {
    enum class Type { Null, Bool, Number, String, Array, Object };
    Type type = Type::Null;
    bool b = false;
    double num = 0.0;
    std::string str;
    std::vector<Json> arr;
    std::vector<std::pair<std::string, Json>> obj;

    const Json *get(std::string_view key) const {
        if (type != Type::Object) return nullptr;
        for (const auto &kv : obj)
            if (kv.first == key) return &kv.second;
        return nullptr;
    }
    bool is_number() const { return type == Type::Number; }
};

class JsonParser {
  public:
    JsonParser(std::string_view text, std::string source) : s_(text), src_(std::move(source)) {}
    Json parse() {
        Json v = value(0);
        ws();
        if (p_ != s_.size()) fail("trailing characters");
        return v;
    }

  private:
    std::string_view s_;
    std::string src_;
    size_t p_ = 0;

    [[noreturn]] void fail(const std::string &what) const {
        throw ConfigError(src_ + ": JSON error at byte " + std::to_string(p_) + ": " + what);
    }
    void ws() {
        while (p_ < s_.size() && (s_[p_] == ' ' || s_[p_] == '\t' || s_[p_] == '\n' || s_[p_] == '\r')) ++p_;
    }
    bool lit(const char *w) {
        size_t n = std::strlen(w);
        if (s_.substr(p_, n) != w) return false;
        p_ += n;
        return true;
    }
    static void put_utf8(std::string &o, uint32_t c) {
        if (c < 0x80) {
            o += (char)c;
        } else if (c < 0x800) {
            o += (char)(0xc0 | (c >> 6));
            o += (char)(0x80 | (c & 0x3f));
        } else if (c < 0x10000) {
            o += (char)(0xe0 | (c >> 12));
            o += (char)(0x80 | ((c >> 6) & 0x3f));
            o += (char)(0x80 | (c & 0x3f));
        } else {
            o += (char)(0xf0 | (c >> 18));
            o += (char)(0x80 | ((c >> 12) & 0x3f));
            o += (char)(0x80 | ((c >> 6) & 0x3f));
            o += (char)(0x80 | (c & 0x3f));
        }
    }
    uint32_t hex4() {
        if (p_ + 4 > s_.size()) fail("short \\u escape");
        uint32_t c = 0;
        for (int k = 0; k < 4; ++k) {
            char h = s_[p_++];
            c <<= 4;
            if (h >= '0' && h <= '9') c |= (uint32_t)(h - '0');
            else if (h >= 'a' && h <= 'f') c |= (uint32_t)(h - 'a' + 10);
            else if (h >= 'A' && h <= 'F') c |= (uint32_t)(h - 'A' + 10);
            else fail("bad \\u escape");
        }
        return c;
    }
    std::string string() {
        if (p_ >= s_.size() || s_[p_] != '"') fail("expected a string");
        ++p_;
        std::string o;
        while (true) {
            if (p_ >= s_.size()) fail("unterminated string");
            char c = s_[p_++];
            if (c == '"') break;
            if ((unsigned char)c < 0x20) fail("control character in string");
            if (c != '\\') {
                o += c;
                continue;
            }
            if (p_ >= s_.size()) fail("unterminated escape");
            char e = s_[p_++];
            switch (e) {
            case '"': o += '"'; break;
            case '\\': o += '\\'; break;
            case '/': o += '/'; break;
            case 'b': o += '\b'; break;
            case 'f': o += '\f'; break;
            case 'n': o += '\n'; break;
            case 'r': o += '\r'; break;
            case 't': o += '\t'; break;
            case 'u': {
                uint32_t c1 = hex4();
                if (c1 >= 0xd800 && c1 < 0xdc00 && s_.substr(p_, 2) == "\\u") {
                    p_ += 2;
                    uint32_t c2 = hex4();
                    if (c2 < 0xdc00 || c2 >= 0xe000) fail("bad surrogate pair");
                    c1 = 0x10000 + ((c1 - 0xd800) << 10) + (c2 - 0xdc00);
                }
                put_utf8(o, c1);
                break;
            }
            default: fail("bad escape");
            }
        }
        return o;
    }
    Json value(int depth) {
        if (depth > 256) fail("nesting too deep");
        ws();
        if (p_ >= s_.size()) fail("unexpected end");
        Json v;
        char c = s_[p_];
        if (c == '{') {
            v.type = Json::Type::Object;
            ++p_;
            ws();
            if (p_ < s_.size() && s_[p_] == '}') {
                ++p_;
                return v;
            }
            while (true) {
                ws();
                std::string k = string();
                ws();
                if (p_ >= s_.size() || s_[p_] != ':') fail("expected ':'");
                ++p_;
                v.obj.emplace_back(std::move(k), value(depth + 1));
                ws();
                if (p_ < s_.size() && s_[p_] == ',') {
                    ++p_;
                    continue;
                }
                if (p_ < s_.size() && s_[p_] == '}') {
                    ++p_;
                    return v;
                }
                fail("expected ',' or '}'");
            }
        }
        if (c == '[') {
            v.type = Json::Type::Array;
            ++p_;
            ws();
            if (p_ < s_.size() && s_[p_] == ']') {
                ++p_;
                return v;
            }
            while (true) {
                v.arr.push_back(value(depth + 1));
                ws();
                if (p_ < s_.size() && s_[p_] == ',') {
                    ++p_;
                    continue;
                }
                if (p_ < s_.size() && s_[p_] == ']') {
                    ++p_;
                    return v;
                }
                fail("expected ',' or ']'");
            }
        }
        if (c == '"') {
            v.type = Json::Type::String;
            v.str = string();
            return v;
        }
        if (lit("true")) {
            v.type = Json::Type::Bool;
            v.b = true;
            return v;
        }
        if (lit("false")) {
            v.type = Json::Type::Bool;
            return v;
        }
        if (lit("null")) return v;
        size_t q = p_;
        if (q < s_.size() && s_[q] == '-') ++q;
        while (q < s_.size() && (std::isdigit((unsigned char)s_[q]) || s_[q] == '.' || s_[q] == 'e' || s_[q] == 'E' ||
                                 s_[q] == '+' || s_[q] == '-'))
            ++q;
        if (q == p_) fail("unexpected character");
        std::string num(s_.substr(p_, q - p_));
        char *end = nullptr;
        v.num = std::strtod(num.c_str(), &end);
        if (end != num.c_str() + num.size()) fail("bad number '" + num + "'");
        v.type = Json::Type::Number;
        p_ = q;
        return v;
    }
};

// ---- buffers ------------------------------------------------------------------------------------

std::vector<uint8_t> read_file(const fs::path &path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw ConfigError("cannot open " + path.string());
    return std::vector<uint8_t>((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
}

std::vector<uint8_t> base64_decode(std::string_view s, const std::string &what) {
    auto val = [](char c) -> int {
        if (c >= 'A' && c <= 'Z') return c - 'A';
        if (c >= 'a' && c <= 'z') return c - 'a' + 26;
        if (c >= '0' && c <= '9') return c - '0' + 52;
        if (c == '+' || c == '-') return 62;
        if (c == '/' || c == '_') return 63;
        return -1;
    };
    std::vector<uint8_t> out;
    out.reserve(s.size() * 3 / 4);
    uint32_t acc = 0;
    int bits = 0;
    for (char c : s) {
        if (c == '=') break;
        int v = val(c);
        if (v < 0) throw ConfigError("bad base64 data in " + what);
        acc = (acc << 6) | (uint32_t)v;
        bits += 6;
        if (bits >= 8) {
            bits -= 8;
            out.push_back((uint8_t)(acc >> bits));
        }
    }
    return out;
}

std::string percent_decode(const std::string &s) {
    std::string o;
    for (size_t i = 0; i < s.size(); ++i) {
        if (s[i] == '%' && i + 2 < s.size() && std::isxdigit((unsigned char)s[i + 1]) &&
            std::isxdigit((unsigned char)s[i + 2])) {
            o += (char)std::stoi(s.substr(i + 1, 2), nullptr, 16);
            i += 2;
        } else {
            o += s[i];
        }
    }
    return o;
}

struct Model {
    Json doc;
    std::vector<std::vector<uint8_t>> buffers;
    std::string src;

    const Json &array(const char *key) const {
        static const Json empty{Json::Type::Array};
        const Json *a = doc.get(key);
        if (!a) return empty;
        if (a->type != Json::Type::Array) throw ConfigError(src + ": '" + key + "' must be an array");
        return *a;
    }
    const Json &element(const char *key, double idx) const {
        const Json &a = array(key);
        if (!(idx >= 0) || idx >= (double)a.arr.size() || idx != std::floor(idx))
            throw ConfigError(src + ": " + key + " index " + std::to_string((long long)idx) + " out of range");
        return a.arr[(size_t)idx];
    }
};

double num(const Json &o, const char *key, double def) {
    const Json *v = o.get(key);
    return v && v->is_number() ? v->num : def;
}
int integer(const Json &o, const char *key, int def) {
    const Json *v = o.get(key);
    if (!v) return def;
    if (!v->is_number() || v->num != std::floor(v->num) || std::fabs(v->num) > 2147483647.0)
        throw ConfigError(std::string("glTF: '") + key + "' must be an integer");
    return (int)v->num;
}
size_t byte_size(const Json &o, const char *key) { // a non-negative byte count or offset, 0 when absent
    const Json *v = o.get(key);
    if (!v) return 0;
    if (!v->is_number() || v->num < 0 || v->num != std::floor(v->num) || v->num > 9.0e15)
        throw ConfigError(std::string("glTF: '") + key + "' must be a non-negative integer");
    return (size_t)v->num;
}

Model load_model(const fs::path &path) {
    Model m;
    m.src = path.string();
    std::vector<uint8_t> file = read_file(path);
    std::string_view json;
    std::vector<uint8_t> bin;
    bool glb = file.size() >= 12 && std::memcmp(file.data(), "glTF", 4) == 0;
    if (glb) {
        uint32_t hdr[3];
        std::memcpy(hdr, file.data(), 12);
        if (hdr[1] != 2) throw ConfigError(m.src + ": only glTF 2 binary files are supported");
        if (hdr[2] > file.size()) throw ConfigError(m.src + ": truncated .glb");
        size_t off = 12;
        int chunk = 0;
        while (off + 8 <= hdr[2]) {
            uint32_t len, type;
            std::memcpy(&len, &file[off], 4);
            std::memcpy(&type, &file[off + 4], 4);
            off += 8;
            if (len > hdr[2] - off) throw ConfigError(m.src + ": truncated .glb chunk");
            if (chunk == 0) {
                if (type != 0x4E4F534Au) throw ConfigError(m.src + ": first .glb chunk is not JSON");
                json = std::string_view(reinterpret_cast<const char *>(&file[off]), len);
            } else if (chunk == 1 && type == 0x004E4942u) {
                bin.assign(file.begin() + (long)off, file.begin() + (long)(off + len));
            }
            off += (len + 3u) & ~size_t(3);
            ++chunk;
        }
        if (chunk == 0) throw ConfigError(m.src + ": .glb without a JSON chunk");
    } else {
        json = std::string_view(reinterpret_cast<const char *>(file.data()), file.size());
    }
    m.doc = JsonParser(json, m.src).parse();
    if (m.doc.type != Json::Type::Object) throw ConfigError(m.src + ": glTF root must be an object");
    const Json *asset = m.doc.get("asset");
    const Json *ver = asset ? asset->get("version") : nullptr;
    if (!ver || ver->type != Json::Type::String || ver->str.rfind("2.", 0) != 0)
        throw ConfigError(m.src + ": asset.version must be 2.x");
    const Json &bufs = m.array("buffers");
    for (size_t i = 0; i < bufs.arr.size(); ++i) {
        const Json &b = bufs.arr[i];
        const Json *uri = b.get("uri");
        std::vector<uint8_t> data;
        if (!uri) {
            if (!glb || i != 0) throw ConfigError(m.src + ": buffer " + std::to_string(i) + " has no uri");
            data = bin;
        } else {
            if (uri->type != Json::Type::String) throw ConfigError(m.src + ": buffer uri must be a string");
            const std::string &u = uri->str;
            if (u.rfind("data:", 0) == 0) {
                size_t comma = u.find(',');
                if (comma == std::string::npos || u.substr(0, comma).find(";base64") == std::string::npos)
                    throw ConfigError(m.src + ": only base64 data URIs are supported");
                data = base64_decode(std::string_view(u).substr(comma + 1), m.src);
            } else {
                data = read_file(path.parent_path() / percent_decode(u));
            }
        }
        double len = num(b, "byteLength", -1);
        if (len < 0 || len > (double)data.size())
            throw ConfigError(m.src + ": buffer " + std::to_string(i) + " is shorter than its byteLength");
        data.resize((size_t)len);
        m.buffers.push_back(std::move(data));
    }
    return m;
}

// ---- accessors ----------------------------------------------------------------------------------

constexpr int kByte = 5120, kUByte = 5121, kShort = 5122, kUShort = 5123, kUInt = 5125, kFloat = 5126;

struct Accessor {
    const uint8_t *base = nullptr;
    size_t stride = 0, count = 0;
    int comp = 0, ncomp = 0;
};

int comp_size(int c) {
    switch (c) {
    case kByte:
    case kUByte: return 1;
    case kShort:
    case kUShort: return 2;
    case kUInt:
    case kFloat: return 4;
    default: return 0;
    }
}

int type_components(const std::string &t) {
    if (t == "SCALAR") return 1;
    if (t == "VEC2") return 2;
    if (t == "VEC3") return 3;
    if (t == "VEC4") return 4;
    if (t == "MAT2") return 4;
    if (t == "MAT3") return 9;
    if (t == "MAT4") return 16;
    return 0;
}

Accessor accessor(const Model &m, double idx) {
    const Json &a = m.element("accessors", idx);
    Accessor r;
    r.comp = integer(a, "componentType", 0);
    const Json *t = a.get("type");
    r.ncomp = t && t->type == Json::Type::String ? type_components(t->str) : 0;
    int cs = comp_size(r.comp);
    if (!cs || !r.ncomp) throw ConfigError(m.src + ": accessor with an unknown component type or type");
    if (a.get("sparse")) throw ConfigError(m.src + ": sparse accessors are not supported");
    int count = integer(a, "count", -1);
    if (count < 0) throw ConfigError(m.src + ": accessor without a count");
    r.count = (size_t)count;
    const Json *bv = a.get("bufferView");
    if (!bv) throw ConfigError(m.src + ": accessors without a bufferView are not supported");
    const Json &view = m.element("bufferViews", bv->num);
    int buf = integer(view, "buffer", -1);
    if (buf < 0 || (size_t)buf >= m.buffers.size()) throw ConfigError(m.src + ": bufferView names a missing buffer");
    const std::vector<uint8_t> &data = m.buffers[(size_t)buf];
    size_t view_off = byte_size(view, "byteOffset"), view_len = byte_size(view, "byteLength");
    size_t acc_off = byte_size(a, "byteOffset");
    size_t elem = (size_t)cs * (size_t)r.ncomp;
    size_t stride = byte_size(view, "byteStride");
    r.stride = stride > 0 ? stride : elem;
    if (view_off > data.size() || view_len > data.size() - view_off)
        throw ConfigError(m.src + ": bufferView runs past its buffer");
    if (r.count > 0 && (r.count - 1 > (view_len / r.stride) || acc_off + r.stride * (r.count - 1) + elem > view_len))
        throw ConfigError(m.src + ": accessor runs past its bufferView");
    r.base = data.data() + view_off + acc_off;
    return r;
}

std::vector<float> read_floats(const Model &m, double idx, int ncomp, const char *what) {
    Accessor a = accessor(m, idx);
    if (a.comp != kFloat || a.ncomp != ncomp)
        throw ConfigError(m.src + ": GLTF: Unsupported " + std::string(what) + " component type or data type");
    std::vector<float> out(a.count * (size_t)ncomp);
    for (size_t i = 0; i < a.count; ++i) std::memcpy(&out[i * ncomp], a.base + i * a.stride, 4 * (size_t)ncomp);
    return out;
}

// ---- transforms (row-major 4x4) -----------------------------------------------------------------

using Mat = std::array<double, 16>;

Mat identity() {
    Mat r{};
    r[0] = r[5] = r[10] = r[15] = 1.0;
    return r;
}

Mat mul(const Mat &a, const Mat &b) {
    Mat r{};
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
            double s = 0.0;
            for (int k = 0; k < 4; ++k) s += a[4 * i + k] * b[4 * k + j];
            r[4 * i + j] = s;
        }
    return r;
}

// node.matrix (column-major) or T * R * S from translation / rotation (x, y, z, w) / scale
Mat node_local(const Model &m, const Json &node) {
    auto numbers = [&](const char *key, size_t n) -> std::vector<double> {
        const Json *v = node.get(key);
        if (!v) return {};
        if (v->type != Json::Type::Array || v->arr.size() != n)
            throw ConfigError(m.src + ": node " + key + " must have " + std::to_string(n) + " numbers");
        std::vector<double> o;
        for (const Json &x : v->arr) {
            if (!x.is_number()) throw ConfigError(m.src + ": node " + key + " must hold numbers");
            o.push_back(x.num);
        }
        return o;
    };
    Mat r = identity();
    std::vector<double> mat = numbers("matrix", 16);
    if (!mat.empty()) {
        for (int c = 0; c < 4; ++c)
            for (int rr = 0; rr < 4; ++rr) r[4 * rr + c] = mat[4 * c + rr];
        return r;
    }
    std::vector<double> t = numbers("translation", 3), q = numbers("rotation", 4), s = numbers("scale", 3);
    double S[3] = {1, 1, 1};
    if (!s.empty()) S[0] = s[0], S[1] = s[1], S[2] = s[2];
    double x = 0, y = 0, z = 0, w = 1;
    if (!q.empty()) x = q[0], y = q[1], z = q[2], w = q[3];
    double R[3][3] = {{1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)},
                      {2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)},
                      {2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)}};
    for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < 3; ++j) r[4 * i + j] = R[i][j] * S[j];
        r[4 * i + 3] = t.empty() ? 0.0 : t[(size_t)i];
    }
    return r;
}

// ---- materials ----------------------------------------------------------------------------------

bool has_texture(const Json &o, const char *key) {
    const Json *t = o.get(key);
    return t && t->type == Json::Type::Object && t->get("index");
}

// glTF textures (create_texture_from_gltf, mesh_asset.cpp:370-389): one per image, decoded to RGBA as
// tinygltf's loader delivers images (4 channels; 16-bit PNGs stay 16-bit), in sRGB where a baseColor or
// emissive texture uses the image, else linear; every level built. PNG only (no JPEG decoder here).
struct GltfTextures {
    const Model *m = nullptr;
    fs::path dir;
    std::vector<int> color_space;     // per image
    std::vector<int> loaded;          // per image: index into out, or -1
    std::vector<TextureData> *out = nullptr;

    std::vector<uint8_t> image_bytes(size_t img) {
        const Json &im = m->element("images", (double)img);
        if (const Json *uri = im.get("uri")) {
            if (uri->type != Json::Type::String) throw ConfigError(m->src + ": image uri must be a string");
            const std::string &u = uri->str;
            if (u.rfind("data:", 0) == 0) {
                size_t comma = u.find(',');
                if (comma == std::string::npos || u.substr(0, comma).find(";base64") == std::string::npos)
                    throw ConfigError(m->src + ": only base64 data URIs are supported");
                return base64_decode(std::string_view(u).substr(comma + 1), m->src);
            }
            return read_file(dir / percent_decode(u));
        }
        const Json *bv = im.get("bufferView");
        if (!bv) throw ConfigError(m->src + ": image without uri or bufferView");
        const Json &view = m->element("bufferViews", bv->num);
        const int buf = integer(view, "buffer", -1);
        if (buf < 0 || (size_t)buf >= m->buffers.size()) throw ConfigError(m->src + ": bad image bufferView");
        const size_t off = byte_size(view, "byteOffset"), len = byte_size(view, "byteLength");
        if (off > m->buffers[(size_t)buf].size() || len > m->buffers[(size_t)buf].size() - off)
            throw ConfigError(m->src + ": image bufferView out of range");
        const uint8_t *b = m->buffers[(size_t)buf].data() + off;
        return std::vector<uint8_t>(b, b + len);
    }
    int texture(size_t img) {
        if (loaded[img] >= 0) return loaded[img];
        std::vector<uint8_t> bytes = image_bytes(img);
        if (bytes.size() < 8 || bytes[0] != 0x89 || bytes[1] != 'P')
            throw ConfigError(m->src + ": image " + std::to_string(img) + " is not a PNG (only PNG images are decoded)");
        PngImage png = decode_png(bytes.data(), bytes.size(), m->src + " image " + std::to_string(img));
        png.to_channels(4);
        TextureData t;
        t.width = png.w, t.height = png.h, t.channels = 4;
        t.color_space = color_space[img];
        t.build_mipmaps = 1;
        if (png.bits == 16) {
            t.data_type = KS_TEX_U16;
            t.bytes.resize(png.data16.size() * 2);
            std::memcpy(t.bytes.data(), png.data16.data(), t.bytes.size());
        } else {
            t.data_type = KS_TEX_U8;
            t.bytes = std::move(png.data8);
        }
        out->push_back(std::move(t));
        return loaded[img] = (int)out->size() - 1;
    }
};

// KHR_texture_transform scale and offset of a texture info (get_texture_transform, mesh_asset.cpp:391-444)
void texture_transform(const Json &info, float scale[2], float offset[2]) {
    scale[0] = scale[1] = 1.0f;
    offset[0] = offset[1] = 0.0f;
    const Json *ext = info.get("extensions");
    const Json *tt = ext ? ext->get("KHR_texture_transform") : nullptr;
    if (!tt) return;
    if (const Json *s = tt->get("scale"); s && s->type == Json::Type::Array)
        for (size_t i = 0; i < 2 && i < s->arr.size(); ++i)
            if (s->arr[i].is_number()) scale[i] = (float)s->arr[i].num;
    if (const Json *o = tt->get("offset"); o && o->type == Json::Type::Array)
        for (size_t i = 0; i < 2 && i < o->arr.size(); ++i)
            if (o->arr[i].is_number()) offset[i] = (float)o->arr[i].num;
}

// create_texture_shader_field (mesh_asset.cpp:445-463): the sampler is EWA when the glTF sampler's
// minFilter mipmaps, else nearest (a texture without a sampler counts as one without minFilter); wraps
// repeat and flip_v is off, as the reference builds them.
ks_field_desc texture_field(const Model &m, GltfTextures &gt, const Json &info, int n, const float *scale,
                            int swizzle0) {
    ks_field_desc f{};
    const Json &tex = m.element("textures", info.get("index")->num);
    const Json *src = tex.get("source");
    if (!src) throw ConfigError(m.src + ": texture without source image");
    (void)m.element("images", src->num); // range check
    f.texture = gt.texture((size_t)src->num);
    int min_filter = -1;
    if (const Json *si = tex.get("sampler")) min_filter = integer(m.element("samplers", si->num), "minFilter", -1);
    const bool mip = min_filter == 9984 || min_filter == 9985 || min_filter == 9986 || min_filter == 9987;
    f.sampler = mip ? KS_SAMPLER_EWA : KS_SAMPLER_NEAREST;
    f.anisotropy = 8.0f;
    f.wrap_u = f.wrap_v = KS_WRAP_REPEAT;
    f.flip_v = 0;
    for (int k = 0; k < 4; ++k) {
        f.swizzle[k] = k < n ? (swizzle0 >= 0 ? swizzle0 + k : k) : -1;
        f.scale[k] = k < n ? scale[k] : 1.0f;
    }
    texture_transform(info, f.uv_scale, f.uv_offset);
    return f;
}

ks_field_desc constant_field(const float *v, int n) {
    ks_field_desc f{};
    f.texture = -1;
    for (int k = 0; k < n; ++k) f.value[k] = v[k];
    return f;
}

// One glTF material as a principled BSDF (mesh_asset.cpp:650-956): constant factors in `d`, and with
// textures every closure input as a field in `fields` (`textured` set); `twosided` is cleared for
// KHR_materials_transmission. alphaMode MASK or BLEND gives an opacity map (alpha of the base colour
// texture times its factor, or the factor), tested stochastically.
ks_material_desc convert_material(const Model &m, const Json *mat, const GltfLoadOptions &opt, bool &twosided,
                                  GltfTextures &gt, ks_material_fields &fields, bool &textured) {
    static const Json empty{Json::Type::Object};
    const Json &src = mat ? *mat : empty;
    const Json *pbr_p = src.get("pbrMetallicRoughness");
    const Json &pbr = pbr_p ? *pbr_p : empty;
    const Json *ext_p = src.get("extensions");
    const Json &ext = ext_p ? *ext_p : empty;
    const bool tex_ok = !opt.ignore_textures;
    auto tex = [&](const Json &o, const char *key) -> const Json * {
        return (tex_ok && has_texture(o, key)) ? o.get(key) : nullptr;
    };
    fields = ks_material_fields{};
    textured = false;

    ks_material_desc d{};
    float bc[4] = {1, 1, 1, 1};
    if (const Json *f = pbr.get("baseColorFactor"); f && f->type == Json::Type::Array && f->arr.size() == 4)
        for (int k = 0; k < 4; ++k) bc[k] = (float)f->arr[(size_t)k].num;
    for (int k = 0; k < 3; ++k) d.basecolor[k] = bc[k];
    if (const Json *t = tex(pbr, "baseColorTexture")) fields.basecolor = texture_field(m, gt, *t, 3, bc, -1), textured = true;
    else fields.basecolor = constant_field(bc, 3);
    const Json *alpha_mode = src.get("alphaMode");
    if (alpha_mode && alpha_mode->type == Json::Type::String && alpha_mode->str != "OPAQUE") {
        fields.has_opacity_map = 1;
        textured = true;
        if (const Json *t = tex(pbr, "baseColorTexture")) fields.opacity = texture_field(m, gt, *t, 1, &bc[3], 3);
        else fields.opacity = constant_field(&bc[3], 1);
    }
    d.roughness = (float)num(pbr, "roughnessFactor", 1.0);
    d.metallic = (float)num(pbr, "metallicFactor", 1.0);
    if (const Json *t = tex(pbr, "metallicRoughnessTexture")) { // roughness in G, metallic in B
        fields.roughness = texture_field(m, gt, *t, 1, &d.roughness, 1);
        fields.metallic = texture_field(m, gt, *t, 1, &d.metallic, 2);
        textured = true;
    } else {
        fields.roughness = constant_field(&d.roughness, 1);
        fields.metallic = constant_field(&d.metallic, 1);
    }

    float strength = 1.0f; // KHR_materials_emissive_strength
    if (const Json *e = ext.get("KHR_materials_emissive_strength")) strength = (float)num(*e, "emissiveStrength", 1.0);
    double ef[3] = {0, 0, 0};
    if (const Json *f = src.get("emissiveFactor"); f && f->type == Json::Type::Array && f->arr.size() == 3)
        for (int k = 0; k < 3; ++k) ef[k] = f->arr[(size_t)k].num;
    float em[3];
    for (int k = 0; k < 3; ++k) em[k] = strength * (float)ef[k];
    if (const Json *t = tex(src, "emissiveTexture")) fields.emission = texture_field(m, gt, *t, 3, em, -1), textured = true;
    else fields.emission = constant_field(em, 3);
    d.has_emission = strength > 0.0f && ((float)ef[0] > 0.0f || (float)ef[1] > 0.0f || (float)ef[2] > 0.0f);
    fields.has_emission = d.has_emission;
    for (int k = 0; k < 3; ++k) d.emission[k] = d.has_emission ? em[k] : 0.0f;

    d.ior = 1.5f; // KHR_materials_ior
    if (const Json *e = ext.get("KHR_materials_ior")) d.ior = (float)num(*e, "ior", 1.5);
    fields.ior = constant_field(&d.ior, 1);

    float sf = 0.5f; // KHR_materials_specular: specularFactor times the mean of specularColorFactor
    const Json *spec_tex = nullptr;
    if (const Json *e = ext.get("KHR_materials_specular")) {
        const Json *f = e->get("specularFactor");
        const Json *cf = e->get("specularColorFactor");
        bool has_cf = cf && cf->type == Json::Type::Array && cf->arr.size() >= 3;
        if ((f && f->is_number()) || has_cf) {
            sf = 1.0f;
            if (f && f->is_number()) sf = (float)f->num;
            if (has_cf) sf *= (float)((cf->arr[0].num + cf->arr[1].num + cf->arr[2].num) / 3.0f);
        }
        spec_tex = tex(*e, "specularTexture");
    }
    d.specular_r0_mul = sf;
    if (spec_tex) fields.specular_r0_mul = texture_field(m, gt, *spec_tex, 1, &sf, 0), textured = true;
    else fields.specular_r0_mul = constant_field(&sf, 1);

    d.specular_trans = 0.0f; // KHR_materials_transmission (single-sided)
    fields.specular_trans = constant_field(&d.specular_trans, 1);
    if (const Json *e = ext.get("KHR_materials_transmission")) {
        twosided = false;
        d.specular_trans = (float)num(*e, "transmissionFactor", 0.0);
        if (const Json *t = tex(*e, "transmissionTexture"))
            fields.specular_trans = texture_field(m, gt, *t, 1, &d.specular_trans, 0), textured = true;
        else fields.specular_trans = constant_field(&d.specular_trans, 1);
    }
    d.diffuse_trans = 0.0f;
    d.diffuse_trans_fwd = 0.0f;
    fields.diffuse_trans = constant_field(&d.diffuse_trans, 1);
    fields.diffuse_trans_fwd = constant_field(&d.diffuse_trans_fwd, 1);
    if (const Json *t = tex(src, "normalTexture")) { // NormalMap with strength = normalTexture.scale
        const float one[3] = {1, 1, 1};
        fields.has_normal_map = 1;
        fields.normal_map = texture_field(m, gt, *t, 3, one, -1);
        fields.normal_space = KS_NORMAL_MAP_TANGENT;
        fields.normal_range01 = 1;
        fields.normal_strength = (float)num(*t, "scale", 1.0);
        for (int k = 0; k < 9; ++k) fields.normal_to_world[k] = (k % 4 == 0) ? 1.0f : 0.0f;
        textured = true;
    }
    return d;
}

} // namespace

CompoundMeshAsset load_gltf(const fs::path &path, const GltfLoadOptions &opt) {
    Model m = load_model(path);
    CompoundMeshAsset out;
    out.has_materials = opt.load_materials;
    const Json &meshes = m.array("meshes");
    out.prototypes.resize(meshes.arr.size());
    out.materials.resize(meshes.arr.size());
    out.fields.resize(meshes.arr.size());
    out.textured.resize(meshes.arr.size());
    // colour spaces of the images: sRGB where a baseColor or emissive texture reads them
    GltfTextures gt;
    gt.m = &m;
    gt.dir = path.parent_path();
    gt.out = &out.textures;
    const size_t n_images = m.array("images").arr.size();
    gt.color_space.assign(n_images, KS_COLORSPACE_LINEAR);
    gt.loaded.assign(n_images, -1);
    for (const Json &mat : m.array("materials").arr) {
        auto mark = [&](const Json *info) {
            if (!info || info->type != Json::Type::Object || !info->get("index")) return;
            const Json *srci = m.element("textures", info->get("index")->num).get("source");
            if (srci && srci->num >= 0 && (size_t)srci->num < n_images) gt.color_space[(size_t)srci->num] = KS_COLORSPACE_SRGB;
        };
        if (const Json *pbr = mat.get("pbrMetallicRoughness")) mark(pbr->get("baseColorTexture"));
        mark(mat.get("emissiveTexture"));
    }
    for (size_t i = 0; i < meshes.arr.size(); ++i) {
        const Json *prims = meshes.arr[i].get("primitives");
        if (!prims || prims->type != Json::Type::Array) throw ConfigError(m.src + ": mesh without primitives");
        for (const Json &prim : prims->arr) {
            MeshData md;
            if (const Json *n = meshes.arr[i].get("name"); n && n->type == Json::Type::String) md.name = n->str;
            if (integer(prim, "mode", 4) != 4) throw ConfigError(m.src + ": GLTF: Only support triangle primitives!");
            const Json *attrs = prim.get("attributes");
            const Json *pos = attrs ? attrs->get("POSITION") : nullptr;
            if (!pos) throw ConfigError(m.src + ": GLTF primitive must have positions!");
            md.vertices = read_floats(m, pos->num, 3, "vertex position");
            size_t nv = md.vertex_count();
            if (const Json *nrm = attrs->get("NORMAL")) {
                md.vertex_normals = read_floats(m, nrm->num, 3, "vertex normal");
                if (md.vertex_normals.size() != 3 * nv) throw ConfigError(m.src + ": NORMAL count differs from POSITION");
            }
            if (const Json *tc = attrs->get("TEXCOORD_0")) {
                md.texcoords = read_floats(m, tc->num, 2, "texcoord");
                if (md.texcoords.size() != 2 * nv) throw ConfigError(m.src + ": TEXCOORD_0 count differs from POSITION");
            }
            if (const Json *ind = prim.get("indices")) {
                Accessor a = accessor(m, ind->num);
                if (a.ncomp != 1) throw ConfigError(m.src + ": GLTF: Unsupport index buffer data type!");
                if (a.comp != kUInt && a.comp != kUShort)
                    throw ConfigError(m.src + ": GLTF: Unsupported index buffer component type!");
                if (a.count % 3 != 0) throw ConfigError(m.src + ": index count is not a multiple of 3");
                md.indices.resize(a.count);
                for (size_t k = 0; k < a.count; ++k) {
                    const uint8_t *p = a.base + k * a.stride;
                    uint32_t v;
                    if (a.comp == kUInt) {
                        std::memcpy(&v, p, 4);
                    } else {
                        uint16_t h;
                        std::memcpy(&h, p, 2);
                        v = h;
                    }
                    if (v >= nv) throw ConfigError(m.src + ": vertex index out of range");
                    md.indices[k] = v;
                }
            } else { // non-indexed triangles: consecutive vertex triples
                if (nv % 3 != 0) throw ConfigError(m.src + ": non-indexed primitive with a vertex count not a multiple of 3");
                md.indices.resize(nv);
                for (size_t k = 0; k < nv; ++k) md.indices[k] = (uint32_t)k;
            }
            md.twosided = true; // doubleSided is not used (mesh_asset.cpp:669-676)
            md.use_smooth_normal = true;
            if (opt.load_materials) {
                const Json *mat = nullptr;
                if (const Json *mi = prim.get("material")) mat = &m.element("materials", mi->num);
                bool ts = true, textured = false;
                ks_material_fields f;
                out.materials[i].push_back(convert_material(m, mat, opt, ts, gt, f, textured));
                out.fields[i].push_back(f);
                out.textured[i].push_back(textured ? 1 : 0);
                md.twosided = ts;
            }
            out.prototypes[i].meshes.push_back(std::move(md));
        }
    }
    // instances: depth-first over the default scene (scene, else 0), world = parent * local
    const Json &scenes = m.array("scenes");
    if (scenes.arr.empty()) return out;
    int scene = integer(m.doc, "scene", 0);
    const Json &sc = m.element("scenes", scene);
    const Json &nodes = m.array("nodes");
    std::vector<int> on_path(nodes.arr.size(), 0);
    auto dfs = [&](double idx, const Mat &parent, auto &self) -> void {
        const Json &node = m.element("nodes", idx);
        size_t ni = (size_t)idx;
        if (on_path[ni]) throw ConfigError(m.src + ": node graph has a cycle");
        on_path[ni] = 1;
        Mat world = mul(parent, node_local(m, node));
        if (const Json *mesh = node.get("mesh")) {
            const Json &mm = m.element("meshes", mesh->num);
            (void)mm;
            std::array<float, 16> w;
            for (int k = 0; k < 16; ++k) w[k] = (float)world[k];
            out.instances.push_back({(uint32_t)mesh->num, w});
        }
        if (const Json *ch = node.get("children")) {
            if (ch->type != Json::Type::Array) throw ConfigError(m.src + ": node children must be an array");
            for (const Json &c : ch->arr) self(c.num, world, self);
        }
        on_path[ni] = 0;
    };
    if (const Json *roots = sc.get("nodes")) {
        if (roots->type != Json::Type::Array) throw ConfigError(m.src + ": scene nodes must be an array");
        for (const Json &r : roots->arr) dfs(r.num, identity(), dfs);
    }
    return out;
}

CompoundMeshAsset create_compound_mesh_asset(const ConfigArgs &args) {
    fs::path path = args.load_path("path");
    std::string fmt = args.load_string("format", "glb");
    if (fmt != "glb" && fmt != "gltf") throw ConfigError("Unsupported mesh asset format [" + fmt + "]");
    GltfLoadOptions opt;
    opt.load_materials = args.load_bool("load_materials");
    if (args.load_string("bsdf_type", "principled_bsdf") != "principled_bsdf")
        throw ConfigError("bsdf_type must be principled_bsdf (the only BSDF the kernels evaluate)");
    opt.ignore_textures = args.load_bool("ignore_textures", false);
    return load_gltf(path, opt);
}

} // namespace ks

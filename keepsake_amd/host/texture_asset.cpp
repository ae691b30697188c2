#include "texture_asset.h"

#include "../csrc/ks_mip.h"
#include "image_io.h"

#include <algorithm>
#include <cctype>
#include <cstring>
#include <fstream>

namespace ks {

ks_texture_desc TextureData::desc() const {
    ks_texture_desc d{};
    d.data = bytes.data();
    d.width = width;
    d.height = height;
    d.channels = channels;
    d.data_type = data_type;
    d.color_space = color_space;
    d.levels = levels;
    d.build_mipmaps = build_mipmaps;
    return d;
}

static std::string lower_ext(const fs::path &p) {
    std::string e = p.extension().string();
    for (char &c : e) c = (char)std::tolower((unsigned char)c);
    return e;
}

TextureData texture_from_image(int channels, bool build_mipmaps, int color_space, const fs::path &path) {
    if (channels < 1 || channels > 4) throw ConfigError("texture channels must be 1..4");
    TextureData t;
    t.channels = channels;
    t.color_space = color_space;
    t.build_mipmaps = build_mipmaps ? 1 : 0;
    const std::string ext = lower_ext(path);
    std::vector<float> f;
    if (ext == ".exr") {
        f = read_exr_channels(path, t.width, t.height, channels); // load_from_exr: RGBA, first `channels`
    } else if (ext == ".hdr") {
        // stbi_loadf with `channels`: RGB, the mean of RGB for one or two channels, alpha 1
        std::vector<float> rgb = read_hdr(path, t.width, t.height);
        const size_t n = (size_t)t.width * t.height;
        f.resize(n * channels);
        for (size_t i = 0; i < n; ++i) {
            const float *c = &rgb[3 * i];
            float *o = &f[i * channels];
            if (channels <= 2) o[0] = (c[0] + c[1] + c[2]) / 3.0f;
            else o[0] = c[0], o[1] = c[1], o[2] = c[2];
            if (channels == 2) o[1] = 1.0f;
            if (channels == 4) o[3] = 1.0f;
        }
    } else if (ext == ".png") {
        PngImage img = read_png(path); // load_from_ldr: stbi_load with `channels`, 8 bits
        img.to_8bit();
        img.to_channels(channels);
        t.width = img.w, t.height = img.h;
        t.data_type = KS_TEX_U8;
        t.bytes = std::move(img.data8);
        return t;
    } else {
        throw ConfigError("unsupported texture image '" + path.string() + "' (.png, .exr or .hdr)");
    }
    t.data_type = KS_TEX_F32;
    t.bytes.resize(f.size() * 4);
    std::memcpy(t.bytes.data(), f.data(), t.bytes.size());
    return t;
}

// ---- LZ4 blocks -----------------------------------------------------------------------------------

std::vector<uint8_t> lz4_compress(const uint8_t *src, size_t n) {
    std::vector<uint8_t> out;
    out.reserve(n + n / 255 + 16);
    const size_t kMinMatch = 4, kLastLiterals = 5, kMfLimit = 12;
    std::vector<int64_t> table(1 << 16, -1);
    auto hash4 = [&](size_t i) {
        uint32_t v;
        std::memcpy(&v, src + i, 4);
        return (v * 2654435761u) >> 16;
    };
    auto put_len = [&](size_t len) {
        while (len >= 255) {
            out.push_back(255);
            len -= 255;
        }
        out.push_back((uint8_t)len);
    };
    size_t anchor = 0, i = 0;
    while (n >= kMfLimit && i + kMfLimit <= n) {
        const uint32_t h = hash4(i);
        const int64_t cand = table[h];
        table[h] = (int64_t)i;
        if (cand >= 0 && i - (size_t)cand <= 65535 && std::memcmp(src + cand, src + i, 4) == 0) {
            size_t len = kMinMatch;
            const size_t max_len = n - kLastLiterals - i;
            while (len < max_len && src[cand + len] == src[i + len]) ++len;
            const size_t lit = i - anchor;
            const size_t ml = len - kMinMatch;
            out.push_back((uint8_t)((std::min<size_t>(lit, 15) << 4) | std::min<size_t>(ml, 15)));
            if (lit >= 15) put_len(lit - 15);
            out.insert(out.end(), src + anchor, src + i);
            const size_t off = i - (size_t)cand;
            out.push_back((uint8_t)(off & 255));
            out.push_back((uint8_t)(off >> 8));
            if (ml >= 15) put_len(ml - 15);
            i += len;
            anchor = i;
        } else {
            ++i;
        }
    }
    const size_t lit = n - anchor; // last sequence: literals only
    out.push_back((uint8_t)(std::min<size_t>(lit, 15) << 4));
    if (lit >= 15) put_len(lit - 15);
    out.insert(out.end(), src + anchor, src + n);
    return out;
}

size_t lz4_decompress(const uint8_t *src, size_t n, uint8_t *dst, size_t cap) {
    size_t ip = 0, op = 0;
    auto bad = []() { throw ConfigError("malformed LZ4 block"); };
    auto read_len = [&](size_t len) {
        if (len == 15) {
            uint8_t b;
            do {
                if (ip >= n) bad();
                b = src[ip++];
                len += b;
            } while (b == 255);
        }
        return len;
    };
    while (ip < n) {
        const uint8_t token = src[ip++];
        const size_t lit = read_len(token >> 4);
        if (lit > n - ip || lit > cap - op) bad();
        std::memcpy(dst + op, src + ip, lit);
        ip += lit;
        op += lit;
        if (ip >= n) break;
        if (n - ip < 2) bad();
        const size_t off = (size_t)src[ip] | (size_t)src[ip + 1] << 8;
        ip += 2;
        const size_t ml = read_len(token & 15) + 4;
        if (off == 0 || off > op || ml > cap - op) bad();
        for (size_t k = 0; k < ml; ++k, ++op) dst[op] = dst[op - off];
    }
    return op;
}

// ---- serialized textures --------------------------------------------------------------------------

static const char kMagic[] = "i_am_a_serialized_texture";
static const size_t kLz4MaxInput = 0x7E000000; // LZ4_MAX_INPUT_SIZE

static int texel_bytes(int dtype, int channels) { return channels * (dtype == KS_TEX_U8 ? 1 : (dtype == KS_TEX_U16 ? 2 : 4)); }

TextureData read_serialized_texture(const fs::path &path) {
    std::ifstream in(path, std::ios::binary);
    if (!in) throw ConfigError("cannot open " + path.string());
    std::vector<uint8_t> f((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
    size_t p = 0;
    auto need = [&](size_t k) {
        if (f.size() - p < k) throw ConfigError(path.string() + ": truncated serialized texture");
    };
    auto get_i32 = [&]() {
        need(4);
        int32_t v;
        std::memcpy(&v, &f[p], 4);
        p += 4;
        return v;
    };
    need(sizeof(kMagic) - 1);
    if (std::memcmp(&f[0], kMagic, sizeof(kMagic) - 1) != 0) throw ConfigError(path.string() + ": invalid serialized texture");
    p = sizeof(kMagic) - 1;
    TextureData t;
    t.width = get_i32();
    t.height = get_i32();
    t.channels = get_i32();
    t.data_type = get_i32();
    t.levels = get_i32();
    need(8);
    uint64_t total;
    std::memcpy(&total, &f[p], 8);
    p += 8;
    const int blocks = get_i32();
    if (t.width <= 0 || t.height <= 0 || t.channels < 1 || t.channels > 4 || t.data_type < 0 || t.data_type > 2 ||
        t.levels < 1 || t.levels > ksa::kMaxTexLevels || blocks < 0)
        throw ConfigError(path.string() + ": bad serialized texture header");
    size_t expect = 0;
    for (int l = 0, w = t.width, h = t.height; l < t.levels; ++l, w = std::max(1, (w + 1) / 2), h = std::max(1, (h + 1) / 2))
        expect += (size_t)w * h * texel_bytes(t.data_type, t.channels);
    if (total != expect) throw ConfigError(path.string() + ": serialized texture size does not match its levels");
    std::vector<int32_t> sizes(blocks);
    for (int b = 0; b < blocks; ++b) sizes[b] = get_i32();
    t.bytes.resize(total);
    size_t out = 0;
    for (int b = 0; b < blocks; ++b) {
        if (sizes[b] < 0) throw ConfigError(path.string() + ": bad LZ4 block size");
        need((size_t)sizes[b]);
        const size_t block = std::min<size_t>(kLz4MaxInput, total - out);
        if (lz4_decompress(&f[p], (size_t)sizes[b], t.bytes.data() + out, block) != block)
            throw ConfigError(path.string() + ": LZ4 block decodes to the wrong size");
        p += (size_t)sizes[b];
        out += block;
    }
    if (out != total) throw ConfigError(path.string() + ": serialized texture data is short");
    t.color_space = KS_COLORSPACE_LINEAR; // as create_texture_from_serialized
    return t;
}

void write_serialized_texture(const TextureData &t, const fs::path &path) {
    std::vector<uint8_t> all = t.bytes;
    int levels = t.levels;
    if (t.levels == 1 && t.build_mipmaps) { // the chain the library would build
        float srgb[256];
        ksa::fill_srgb_u8_table(srgb);
        ksa::SourceTexture s = ksa::load_source_texture(t.desc(), srgb);
        all.clear();
        for (const auto &l : s.levels) all.insert(all.end(), l.begin(), l.end());
        levels = (int)s.levels.size();
    }
    std::vector<uint8_t> o(kMagic, kMagic + sizeof(kMagic) - 1);
    auto put = [&](const void *v, size_t k) { o.insert(o.end(), (const uint8_t *)v, (const uint8_t *)v + k); };
    const int32_t hdr[5] = {t.width, t.height, t.channels, t.data_type, levels};
    put(hdr, sizeof(hdr));
    const uint64_t total = all.size();
    put(&total, 8);
    const int32_t blocks = (int32_t)((total + kLz4MaxInput - 1) / kLz4MaxInput);
    put(&blocks, 4);
    std::vector<std::vector<uint8_t>> comp;
    for (int32_t b = 0; b < blocks; ++b) {
        const size_t off = (size_t)b * kLz4MaxInput;
        comp.push_back(lz4_compress(all.data() + off, std::min<size_t>(kLz4MaxInput, total - off)));
        const int32_t sz = (int32_t)comp.back().size();
        put(&sz, 4);
    }
    for (const auto &c : comp) put(c.data(), c.size());
    std::ofstream of(path, std::ios::binary);
    if (!of) throw ConfigError("cannot write " + path.string());
    of.write((const char *)o.data(), (std::streamsize)o.size());
}

TextureData create_texture(const ConfigArgs &args) {
    fs::path path = args.load_path("path");
    if (args.load_bool("serialized", false)) return read_serialized_texture(path);
    int ch = args.load_integer("channels");
    bool mips = args.load_bool("build_mipmaps");
    std::string cs = args.load_string("colorspace", "Linear");
    return texture_from_image(ch, mips, cs == "sRGB" ? KS_COLORSPACE_SRGB : KS_COLORSPACE_LINEAR, path);
}

void convert_texture(const ConfigArgs &args, const fs::path &task_dir, int) {
    ConfigArgs list = args["textures"];
    fs::create_directories(task_dir);
    for (size_t i = 0; i < list.array_size(); ++i) {
        const TomlValue *v = list[(int)i].node;
        if (v->type != TomlValue::Type::String) throw ConfigError("textures must be a list of asset names");
        TextureData t = create_texture(args.asset(v->str));
        std::string name = v->str.substr(v->str.rfind('.') + 1);
        write_serialized_texture(t, task_dir / (name + ".bin"));
    }
}

int TextureRegistry::get(const ConfigArgs &args, const std::string &name) {
    auto it = index.find(name);
    if (it != index.end()) return it->second;
    textures.push_back(create_texture(args.asset(name)));
    return index[name] = (int)textures.size() - 1;
}

// ---- shader fields --------------------------------------------------------------------------------

static std::vector<double> numbers(const ConfigArgs &a, const char *key, size_t n) {
    const TomlValue *v = a.node ? a.node->find(key) : nullptr;
    if (!v) throw ConfigError(std::string("missing '") + key + "'");
    std::vector<double> out;
    if (v->is_number() && n == 1) return {v->number()};
    if (!v->is_array() || v->arr.size() != n) throw ConfigError(std::string("'") + key + "' must hold " + std::to_string(n) + " numbers");
    for (const TomlValue &e : v->arr) {
        if (!e.is_number()) throw ConfigError(std::string("'") + key + "' must hold numbers");
        out.push_back(e.number());
    }
    return out;
}

ks_field_desc load_shader_field(const ConfigArgs &f, int n, TextureRegistry &reg) {
    ks_field_desc d{};
    d.texture = -1;
    const std::string type = f.load_string("type");
    if (type == "constant") { // create_constant_shader_field_color
        std::vector<double> v = numbers(f, "value", (size_t)n);
        for (int k = 0; k < n; ++k) d.value[k] = (float)v[(size_t)k];
        if (n == 1) d.value[1] = d.value[2] = d.value[0];
        return d;
    }
    if (type != "texture") throw ConfigError("unknown shader field type '" + type + "'");
    d.texture = reg.get(f, f.load_string("map"));
    // create_texture_sampler
    ConfigArgs s = f["sampler"];
    const std::string st = s.load_string("type");
    if (st == "nearest") d.sampler = KS_SAMPLER_NEAREST;
    else if (st == "linear") d.sampler = KS_SAMPLER_LINEAR;
    else if (st == "cubic") d.sampler = KS_SAMPLER_CUBIC;
    else if (st == "ewa") d.sampler = KS_SAMPLER_EWA;
    else throw ConfigError("unknown texture sampler '" + st + "'");
    const std::string k = s.load_string("kernel", "mitchell");
    d.cubic_kernel = k == "bspline" ? KS_CUBIC_BSPLINE : (k == "catmull_rom" ? KS_CUBIC_CATMULL_ROM : KS_CUBIC_MITCHELL);
    d.anisotropy = s.load_float("anisotropy", 8.0f);
    d.wrap_u = s.load_string("wrap_mode_u", "repeat") == "clamp" ? KS_WRAP_CLAMP : KS_WRAP_REPEAT;
    d.wrap_v = s.load_string("wrap_mode_v", "repeat") == "clamp" ? KS_WRAP_CLAMP : KS_WRAP_REPEAT;
    d.flip_v = f.load_bool("flip_v", true) ? 1 : 0;
    for (int c = 0; c < 4; ++c) d.swizzle[c] = c < n ? c : -1;
    if (f.contains("swizzle")) {
        std::vector<double> sw = numbers(f, "swizzle", (size_t)n);
        for (int c = 0; c < n; ++c) d.swizzle[c] = (int32_t)sw[(size_t)c];
    }
    d.uv_scale[0] = d.uv_scale[1] = 1.0f;
    if (f.contains("uv_scale")) {
        std::vector<double> v = numbers(f, "uv_scale", 2);
        d.uv_scale[0] = (float)v[0], d.uv_scale[1] = (float)v[1];
    }
    if (f.contains("uv_offset")) {
        std::vector<double> v = numbers(f, "uv_offset", 2);
        d.uv_offset[0] = (float)v[0], d.uv_offset[1] = (float)v[1];
    }
    for (int c = 0; c < 4; ++c) d.scale[c] = 1.0f;
    if (f.contains("scale")) {
        std::vector<double> v = numbers(f, "scale", (size_t)n);
        for (int c = 0; c < n; ++c) d.scale[c] = (float)v[(size_t)c];
    }
    const TextureData &t = reg.textures[(size_t)d.texture];
    for (int c = 0; c < n; ++c)
        if (d.swizzle[c] >= t.channels) throw ConfigError("swizzle names a channel the texture does not have");
    return d;
}

} // namespace ks

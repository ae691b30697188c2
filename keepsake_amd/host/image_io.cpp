// EXR / HDR / PNG files, tone mappers and the PNG film composite (see image_io.h).
#include "image_io.h"

#include <zlib.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <fstream>
#include <iterator>

namespace ks {

// ---- scalar conversions -------------------------------------------------------------------------

uint16_t float_to_half(float f) {
    uint32_t x;
    std::memcpy(&x, &f, 4);
    const uint32_t sign = (x >> 16) & 0x8000u;
    const uint32_t exp = (x >> 23) & 0xffu;
    uint32_t man = x & 0x7fffffu;
    if (exp == 0xffu) { // infinity, or NaN with its quiet bit set
        if (man == 0) return (uint16_t)(sign | 0x7c00u);
        return (uint16_t)(sign | 0x7e00u | (man >> 13));
    }
    const int e = (int)exp - 127 + 15;
    if (e >= 31) return (uint16_t)(sign | 0x7c00u);
    if (e <= 0) { // subnormal half (or zero): 10-bit significand of value / 2^-24
        if (e < -10) return (uint16_t)sign;
        man |= 0x800000u;
        const int shift = 14 - e;
        uint32_t m = man >> shift;
        const uint32_t rem = man & ((1u << shift) - 1u), half = 1u << (shift - 1);
        if (rem > half || (rem == half && (m & 1u))) ++m;
        return (uint16_t)(sign | m);
    }
    uint32_t r = sign | ((uint32_t)e << 10) | (man >> 13);
    const uint32_t rem = man & 0x1fffu;
    if (rem > 0x1000u || (rem == 0x1000u && (r & 1u))) ++r; // a carry rounds into the exponent
    return (uint16_t)r;
}

float half_to_float(uint16_t h) {
    const uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
    const uint32_t exp = (h >> 10) & 0x1fu, man = h & 0x3ffu;
    uint32_t x;
    if (exp == 0) {
        if (man == 0) {
            x = sign;
        } else { // subnormal: man * 2^-24, exact in binary32
            float f = (float)man * 5.9604644775390625e-08f;
            std::memcpy(&x, &f, 4);
            x |= sign;
        }
    } else if (exp == 31) {
        x = sign | 0x7f800000u | (man << 13);
    } else {
        x = sign | ((exp + 112u) << 23) | (man << 13);
    }
    float f;
    std::memcpy(&f, &x, 4);
    return f;
}

float linear_to_srgb(float x) {
    if (x <= 0.0f) return 0.0f;
    if (x >= 1.0f) return 1.0f;
    if (x < 0.0031308f) return x * 12.92f;
    return std::pow(x, 1.0f / 2.4f) * 1.055f - 0.055f;
}

// ---- little-endian byte helpers -----------------------------------------------------------------

namespace {

struct Bytes {
    std::vector<uint8_t> b;
    void raw(const void *p, size_t n) {
        const uint8_t *c = static_cast<const uint8_t *>(p);
        b.insert(b.end(), c, c + n);
    }
    void u8(uint8_t v) { b.push_back(v); }
    void i32(int32_t v) { raw(&v, 4); }
    void u32(uint32_t v) { raw(&v, 4); }
    void u64(uint64_t v) { raw(&v, 8); }
    void f32(float v) { raw(&v, 4); }
    void str(const std::string &s) { raw(s.c_str(), s.size() + 1); }
    void be32(uint32_t v) {
        for (int k = 3; k >= 0; --k) b.push_back((uint8_t)(v >> (8 * k)));
    }
};

struct Cursor {
    const std::vector<uint8_t> &b;
    size_t p;
    std::string src;
    void need(size_t n) const {
        if (p > b.size() || n > b.size() - p) throw ConfigError(src + ": truncated file");
    }
    template <typename T> T get() {
        need(sizeof(T));
        T v;
        std::memcpy(&v, &b[p], sizeof(T));
        p += sizeof(T);
        return v;
    }
    std::string cstr() {
        size_t e = p;
        while (e < b.size() && b[e]) ++e;
        if (e >= b.size()) throw ConfigError(src + ": truncated header");
        std::string s(reinterpret_cast<const char *>(&b[p]), e - p);
        p = e + 1;
        return s;
    }
};

std::vector<uint8_t> read_all(const fs::path &path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw ConfigError("cannot open " + path.string());
    return std::vector<uint8_t>((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
}

void write_all(const fs::path &path, const std::vector<uint8_t> &b) {
    std::ofstream f(path, std::ios::binary);
    f.write(reinterpret_cast<const char *>(b.data()), (std::streamsize)b.size());
    if (!f) throw ConfigError("cannot write " + path.string());
}

// OpenEXR ZIP / RLE pre-pass: even bytes then odd bytes, then byte deltas offset by 128
void exr_reorder_predict(const uint8_t *in, size_t n, std::vector<uint8_t> &out) {
    out.resize(n);
    size_t h = (n + 1) / 2;
    for (size_t i = 0; i < n; ++i) out[(i & 1) ? h + i / 2 : i / 2] = in[i];
    uint8_t prev = n ? out[0] : 0;
    for (size_t i = 1; i < n; ++i) {
        uint8_t cur = out[i];
        out[i] = (uint8_t)(cur - prev + 128);
        prev = cur;
    }
}

void exr_unpredict_reorder(std::vector<uint8_t> &t, std::vector<uint8_t> &out) {
    size_t n = t.size();
    for (size_t i = 1; i < n; ++i) t[i] = (uint8_t)(t[i - 1] + t[i] - 128);
    out.resize(n);
    size_t h = (n + 1) / 2;
    for (size_t i = 0; i < n; ++i) out[i] = t[(i & 1) ? h + i / 2 : i / 2];
}

struct ExrChannel {
    std::string name;
    int type; // 0 UINT, 1 HALF, 2 FLOAT
};

} // namespace

// ---- EXR ----------------------------------------------------------------------------------------

void write_exr(const fs::path &path, const float *data, int w, int h, int c, bool half, ExrCompression comp) {
    if (c < 1 || c > 4 || w <= 0 || h <= 0) throw ConfigError("write_exr: bad image shape");
    static const char *names[4] = {"R", "G", "B", "A"};
    std::vector<int> order; // channel names in ascending order: A, B, G, R
    for (int k : {3, 2, 1, 0})
        if (k < c) order.push_back(k);
    const int ptype = half ? 1 : 2;
    Bytes o;
    o.u32(20000630u);
    o.u32(2u);
    {
        Bytes cl;
        for (int k : order) {
            cl.str(names[k]);
            cl.i32(ptype);
            cl.u8(0);
            cl.u8(0);
            cl.u8(0);
            cl.u8(0);
            cl.i32(1);
            cl.i32(1);
        }
        cl.u8(0);
        o.str("channels");
        o.str("chlist");
        o.i32((int32_t)cl.b.size());
        o.raw(cl.b.data(), cl.b.size());
    }
    o.str("compression");
    o.str("compression");
    o.i32(1);
    o.u8((uint8_t)comp);
    for (const char *win : {"dataWindow", "displayWindow"}) {
        o.str(win);
        o.str("box2i");
        o.i32(16);
        o.i32(0);
        o.i32(0);
        o.i32(w - 1);
        o.i32(h - 1);
    }
    o.str("lineOrder");
    o.str("lineOrder");
    o.i32(1);
    o.u8(0);
    o.str("pixelAspectRatio");
    o.str("float");
    o.i32(4);
    o.f32(1.0f);
    o.str("screenWindowCenter");
    o.str("v2f");
    o.i32(8);
    o.f32(0.0f);
    o.f32(0.0f);
    o.str("screenWindowWidth");
    o.str("float");
    o.i32(4);
    o.f32(1.0f);
    o.u8(0);

    const int lines = comp == ExrCompression::Zip ? 16 : 1;
    const int blocks = (h + lines - 1) / lines;
    const size_t table = o.b.size();
    for (int k = 0; k < blocks; ++k) o.u64(0);
    std::vector<uint8_t> raw, pre, z;
    for (int blk = 0; blk < blocks; ++blk) {
        const int y0 = blk * lines, y1 = std::min(h, y0 + lines);
        raw.clear();
        for (int y = y0; y < y1; ++y)
            for (int k : order)
                for (int x = 0; x < w; ++x) {
                    float v = data[((size_t)y * w + x) * c + k];
                    if (half) {
                        uint16_t hv = float_to_half(v);
                        raw.push_back((uint8_t)hv);
                        raw.push_back((uint8_t)(hv >> 8));
                    } else {
                        uint8_t fb[4];
                        std::memcpy(fb, &v, 4);
                        raw.insert(raw.end(), fb, fb + 4);
                    }
                }
        const std::vector<uint8_t> *payload = &raw;
        if (comp == ExrCompression::Zip) {
            exr_reorder_predict(raw.data(), raw.size(), pre);
            uLongf zl = compressBound((uLong)pre.size());
            z.resize(zl);
            if (compress2(z.data(), &zl, pre.data(), (uLong)pre.size(), 6) != Z_OK)
                throw ConfigError("write_exr: zlib failure");
            z.resize(zl);
            if (z.size() < raw.size()) payload = &z; // else stored raw, as OpenEXR does
        }
        uint64_t pos = o.b.size();
        std::memcpy(&o.b[table + 8 * (size_t)blk], &pos, 8);
        o.i32(y0);
        o.i32((int32_t)payload->size());
        o.raw(payload->data(), payload->size());
    }
    write_all(path, o.b);
}

std::vector<float> read_exr(const fs::path &path, int &w, int &h) { return read_exr_channels(path, w, h, 3); }

std::vector<float> read_exr_channels(const fs::path &path, int &w, int &h, int n_out) {
    std::vector<uint8_t> file = read_all(path);
    Cursor cur{file, 0, path.string()};
    if (cur.get<uint32_t>() != 20000630u) throw ConfigError(path.string() + ": not an OpenEXR file");
    uint32_t ver = cur.get<uint32_t>();
    if ((ver & 0xffu) != 2u) throw ConfigError(path.string() + ": unsupported OpenEXR version");
    if (ver & 0x200u) throw ConfigError(path.string() + ": tiled OpenEXR files are not supported");
    if (ver & 0x1800u) throw ConfigError(path.string() + ": multi-part or deep OpenEXR files are not supported");
    std::vector<ExrChannel> chans;
    int comp = -1;
    int32_t win[4] = {0, 0, -1, -1};
    while (true) {
        std::string name = cur.cstr();
        if (name.empty()) break;
        std::string type = cur.cstr();
        int32_t size = cur.get<int32_t>();
        if (size < 0) throw ConfigError(path.string() + ": bad attribute size");
        cur.need((size_t)size);
        size_t end = cur.p + (size_t)size;
        if (name == "channels" && type == "chlist") {
            while (cur.p < end && file[cur.p] != 0) {
                ExrChannel ch;
                ch.name = cur.cstr();
                ch.type = cur.get<int32_t>();
                cur.get<uint32_t>(); // pLinear + reserved
                int32_t xs = cur.get<int32_t>(), ys = cur.get<int32_t>();
                if (xs != 1 || ys != 1) throw ConfigError(path.string() + ": subsampled channels are not supported");
                if (ch.type < 0 || ch.type > 2) throw ConfigError(path.string() + ": unknown pixel type");
                chans.push_back(ch);
            }
        } else if (name == "compression") {
            comp = file[cur.p];
        } else if (name == "dataWindow") {
            if (size != 16) throw ConfigError(path.string() + ": bad dataWindow");
            std::memcpy(win, &file[cur.p], 16);
        }
        cur.p = end;
    }
    w = win[2] - win[0] + 1;
    h = win[3] - win[1] + 1;
    if (w <= 0 || h <= 0 || chans.empty()) throw ConfigError(path.string() + ": empty image");
    int lines;
    switch (comp) {
    case 0:
    case 1:
    case 2: lines = 1; break;
    case 3: lines = 16; break;
    default:
        throw ConfigError(path.string() + ": OpenEXR compression " + std::to_string(comp) +
                          " is not supported (NONE, RLE, ZIPS and ZIP are)");
    }
    auto find = [&](const char *n) {
        for (size_t i = 0; i < chans.size(); ++i)
            if (chans[i].name == n) return (int)i;
        return -1;
    };
    // R, G, B by name (or one channel replicated, alpha included, as LoadEXR does for single-channel
    // images); A by name, else 1
    int src[4] = {find("R"), find("G"), find("B"), find("A")};
    if (src[0] < 0 || src[1] < 0 || src[2] < 0) {
        int y = find("Y");
        if (y < 0 && chans.size() == 1) y = 0;
        if (y < 0) throw ConfigError(path.string() + ": no R, G, B (or Y) channels");
        src[0] = src[1] = src[2] = y;
        if (chans.size() == 1) src[3] = y;
    }
    if (n_out < 1 || n_out > 4) throw ConfigError(path.string() + ": 1 to 4 channels can be read");
    size_t line_bytes = 0;
    std::vector<size_t> ch_off(chans.size());
    for (size_t i = 0; i < chans.size(); ++i) {
        ch_off[i] = line_bytes;
        line_bytes += (size_t)w * (chans[i].type == 1 ? 2 : 4);
    }
    const int blocks = (h + lines - 1) / lines;
    std::vector<uint64_t> offsets(blocks);
    for (int k = 0; k < blocks; ++k) offsets[k] = cur.get<uint64_t>();
    std::vector<float> out((size_t)w * h * n_out, 0.0f);
    if (n_out == 4 && src[3] < 0)
        for (size_t i = 0; i < (size_t)w * h; ++i) out[4 * i + 3] = 1.0f;
    std::vector<uint8_t> tmp, pix;
    for (int blk = 0; blk < blocks; ++blk) {
        Cursor c2{file, (size_t)offsets[blk], path.string()};
        int32_t y0 = c2.get<int32_t>() - win[1];
        int32_t size = c2.get<int32_t>();
        if (y0 < 0 || y0 >= h || size < 0) throw ConfigError(path.string() + ": bad chunk");
        c2.need((size_t)size);
        const int nl = std::min(lines, h - y0);
        const size_t expect = line_bytes * (size_t)nl;
        const uint8_t *data = &file[c2.p];
        if ((size_t)size == expect || comp == 0) {
            if ((size_t)size != expect) throw ConfigError(path.string() + ": bad chunk size");
            pix.assign(data, data + expect);
        } else if (comp == 1) {
            tmp.clear();
            for (int32_t i = 0; i < size;) {
                int8_t n = (int8_t)data[i++];
                if (n < 0) {
                    if (i - n > size) throw ConfigError(path.string() + ": bad RLE data");
                    tmp.insert(tmp.end(), data + i, data + i - n);
                    i -= n;
                } else {
                    if (i >= size) throw ConfigError(path.string() + ": bad RLE data");
                    tmp.insert(tmp.end(), (size_t)n + 1, data[i++]);
                }
            }
            if (tmp.size() != expect) throw ConfigError(path.string() + ": bad RLE chunk size");
            exr_unpredict_reorder(tmp, pix);
        } else {
            tmp.resize(expect);
            uLongf dl = (uLongf)expect;
            if (uncompress(tmp.data(), &dl, data, (uLong)size) != Z_OK || dl != expect)
                throw ConfigError(path.string() + ": bad ZIP chunk");
            exr_unpredict_reorder(tmp, pix);
        }
        for (int l = 0; l < nl; ++l)
            for (int k = 0; k < n_out; ++k) {
                if (src[k] < 0) continue;
                const ExrChannel &ch = chans[(size_t)src[k]];
                const uint8_t *p = pix.data() + (size_t)l * line_bytes + ch_off[(size_t)src[k]];
                for (int x = 0; x < w; ++x) {
                    float v;
                    if (ch.type == 1) {
                        uint16_t hv;
                        std::memcpy(&hv, p + 2 * (size_t)x, 2);
                        v = half_to_float(hv);
                    } else if (ch.type == 2) {
                        std::memcpy(&v, p + 4 * (size_t)x, 4);
                    } else {
                        uint32_t u;
                        std::memcpy(&u, p + 4 * (size_t)x, 4);
                        v = (float)u;
                    }
                    out[((size_t)(y0 + l) * w + x) * n_out + k] = v;
                }
            }
    }
    return out;
}

// ---- Radiance .hdr ------------------------------------------------------------------------------

std::vector<float> read_hdr(const fs::path &path, int &w, int &h) {
    std::vector<uint8_t> f = read_all(path);
    size_t p = 0;
    auto line = [&]() {
        size_t e = p;
        while (e < f.size() && f[e] != '\n') ++e;
        if (e >= f.size()) throw ConfigError(path.string() + ": truncated .hdr header");
        std::string s(reinterpret_cast<const char *>(&f[p]), e - p);
        p = e + 1;
        return s;
    };
    std::string first = line();
    if (first.rfind("#?RADIANCE", 0) != 0 && first.rfind("#?RGBE", 0) != 0)
        throw ConfigError(path.string() + ": not a Radiance .hdr file");
    while (true) {
        std::string l = line();
        if (l.empty()) break;
        if (l.rfind("FORMAT=", 0) == 0 && l != "FORMAT=32-bit_rle_rgbe")
            throw ConfigError(path.string() + ": unsupported .hdr format " + l.substr(7));
    }
    std::string res = line();
    char ys[3] = {0}, xs[3] = {0};
    if (std::sscanf(res.c_str(), "%2s %d %2s %d", ys, &h, xs, &w) != 4 || std::string(ys) != "-Y" ||
        std::string(xs) != "+X" || w <= 0 || h <= 0)
        throw ConfigError(path.string() + ": unsupported .hdr resolution line '" + res + "'");
    std::vector<float> out((size_t)w * h * 3);
    std::vector<uint8_t> rgbe((size_t)w * 4);
    auto take = [&](size_t n) {
        if (p > f.size() || n > f.size() - p) throw ConfigError(path.string() + ": truncated .hdr data");
        const uint8_t *q = &f[p];
        p += n;
        return q;
    };
    for (int y = 0; y < h; ++y) {
        bool rle = w >= 8 && w < 32768 && p + 4 <= f.size() && f[p] == 2 && f[p + 1] == 2 && !(f[p + 2] & 0x80);
        if (rle) {
            const uint8_t *hd = take(4);
            if (((int)hd[2] << 8 | hd[3]) != w) throw ConfigError(path.string() + ": bad .hdr scanline width");
            for (int k = 0; k < 4; ++k)
                for (int x = 0; x < w;) {
                    int n = *take(1);
                    if (n > 128) {
                        n -= 128;
                        uint8_t v = *take(1);
                        if (x + n > w) throw ConfigError(path.string() + ": bad .hdr run");
                        for (int i = 0; i < n; ++i) rgbe[(size_t)(x++) * 4 + k] = v;
                    } else {
                        if (n == 0 || x + n > w) throw ConfigError(path.string() + ": bad .hdr run");
                        const uint8_t *q = take((size_t)n);
                        for (int i = 0; i < n; ++i) rgbe[(size_t)(x++) * 4 + k] = q[i];
                    }
                }
        } else {
            const uint8_t *q = take((size_t)w * 4);
            std::memcpy(rgbe.data(), q, (size_t)w * 4);
        }
        for (int x = 0; x < w; ++x) {
            const uint8_t *e = &rgbe[(size_t)x * 4];
            float s = e[3] ? std::ldexp(1.0f, (int)e[3] - (128 + 8)) : 0.0f;
            for (int k = 0; k < 3; ++k) out[((size_t)y * w + x) * 3 + k] = e[k] * s;
        }
    }
    return out;
}

// ---- PNG ----------------------------------------------------------------------------------------

void write_png(const fs::path &path, const uint8_t *data, int w, int h, int c) {
    if (c < 1 || c > 4 || w <= 0 || h <= 0) throw ConfigError("write_png: bad image shape");
    static const uint8_t color_type[5] = {0, 0, 4, 2, 6};
    Bytes o;
    const uint8_t sig[8] = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1a, '\n'};
    o.raw(sig, 8);
    auto chunk = [&](const char *type, const std::vector<uint8_t> &body) {
        o.be32((uint32_t)body.size());
        size_t start = o.b.size();
        o.raw(type, 4);
        o.raw(body.data(), body.size());
        o.be32((uint32_t)crc32(0L, o.b.data() + start, (uInt)(o.b.size() - start)));
    };
    Bytes ihdr;
    ihdr.be32((uint32_t)w);
    ihdr.be32((uint32_t)h);
    ihdr.u8(8);
    ihdr.u8(color_type[c]);
    ihdr.u8(0);
    ihdr.u8(0);
    ihdr.u8(0);
    chunk("IHDR", ihdr.b);
    const size_t row = (size_t)w * c;
    std::vector<uint8_t> filtered((row + 1) * h);
    for (int y = 0; y < h; ++y) {
        filtered[(row + 1) * y] = 0; // filter type None
        std::memcpy(&filtered[(row + 1) * y + 1], data + row * y, row);
    }
    uLongf zl = compressBound((uLong)filtered.size());
    std::vector<uint8_t> z(zl);
    if (compress2(z.data(), &zl, filtered.data(), (uLong)filtered.size(), 6) != Z_OK)
        throw ConfigError("write_png: zlib failure");
    z.resize(zl);
    chunk("IDAT", z);
    chunk("IEND", {});
    write_all(path, o.b);
}

// ---- tone mappers (tonemap.h) -------------------------------------------------------------------

namespace {

using V3 = std::array<float, 3>;
using M3 = std::array<float, 9>; // row-major, float like the reference's mat3

V3 mul(const M3 &m, const V3 &v) {
    V3 r;
    for (int i = 0; i < 3; ++i) r[i] = m[3 * i] * v[0] + m[3 * i + 1] * v[1] + m[3 * i + 2] * v[2];
    return r;
}

V3 clamp01(V3 v) {
    for (float &x : v) x = std::min(std::max(x, 0.0f), 1.0f);
    return v;
}

// Published constants: the ACES fit of S. Hill (sRGB -> AP1 with RRT_SAT, ODT_SAT -> sRGB), the AgX
// inset / outset matrices of the AgX LUT generator and the Rec.2020 <-> sRGB matrices (tonemap.h:27-41,
// 127-158).
const M3 kAcesIn = {0.59719f, 0.35458f, 0.04823f, 0.07600f, 0.90834f, 0.01566f, 0.02840f, 0.13383f, 0.83777f};
const M3 kAcesOut = {1.60475f, -0.53108f, -0.07367f, -0.10208f, 1.10813f, -0.00605f, -0.00327f, -0.07276f, 1.07602f};
const M3 kAgxInset = {0.856627153315983f, 0.0951212405381588f, 0.0482516061458583f,
                      0.137318972929847f, 0.761241990602591f,  0.101439036467562f,
                      0.11189821299995f,  0.0767994186031903f, 0.811302368396859f};
const M3 kAgxOutsetInv = {0.899796955911611f, 0.0871996192028351f, 0.013003424885555f,
                          0.11142098895748f,  0.875575586156966f,  0.0130034248855548f,
                          0.11142098895748f,  0.0871996192028349f, 0.801379391839686f};
const M3 kRec2020ToSrgb = {1.6605f, -0.5876f, -0.0728f, -0.1246f, 1.1329f, -0.0083f, -0.0182f, -0.1006f, 1.1187f};
const M3 kSrgbToRec2020 = {0.6274f, 0.3293f, 0.0433f, 0.0691f, 0.9195f, 0.0113f, 0.0164f, 0.0880f, 0.8956f};

// 3x3 inverse by cofactors, in float (AgXOutsetMatrix = AgXOutsetMatrixInv.inverse(), tonemap.h:142)
M3 inverse(const M3 &a) {
    float c00 = a[4] * a[8] - a[5] * a[7], c01 = a[5] * a[6] - a[3] * a[8], c02 = a[3] * a[7] - a[4] * a[6];
    float det = a[0] * c00 + a[1] * c01 + a[2] * c02;
    float id = 1.0f / det;
    return {c00 * id,
            (a[2] * a[7] - a[1] * a[8]) * id,
            (a[1] * a[5] - a[2] * a[4]) * id,
            c01 * id,
            (a[0] * a[8] - a[2] * a[6]) * id,
            (a[2] * a[3] - a[0] * a[5]) * id,
            c02 * id,
            (a[1] * a[6] - a[0] * a[7]) * id,
            (a[0] * a[4] - a[1] * a[3]) * id};
}

struct AcesToneMapper final : ToneMapper { // tonemap.h:44-76
    explicit AcesToneMapper(float exposure) : scale((float)(exposure / 0.6)) {}
    V3 operator()(V3 v) const override {
        for (float &x : v) x *= scale;
        v = mul(kAcesIn, v);
        for (float &x : v) { // RRTAndODTFit
            float a = x * (x + 0.0245786f) - 0.000090537f;
            float b = x * (0.983729f * x + 0.4329510f) + 0.238081f;
            x = a / b;
        }
        return clamp01(mul(kAcesOut, v));
    }
    float scale;
};

enum class AgxLook { None, Golden, Punchy };

struct AgxToneMapper final : ToneMapper { // tonemap.h:84-123, 161-210
    AgxToneMapper(float exposure, AgxLook look) : exposure(exposure), look(look), outset(inverse(kAgxOutsetInv)) {}
    V3 operator()(V3 v) const override {
        const float min_ev = -12.47393f, max_ev = 4.026069f;
        for (float &x : v) x = std::max(x * exposure, 0.0f);
        v = mul(kSrgbToRec2020, v);
        v = mul(kAgxInset, v);
        for (float &x : v) x = (std::log2(std::max(x, 1e-10f)) - min_ev) / (max_ev - min_ev);
        v = clamp01(v);
        for (float &x : v) { // agxDefaultContrastApprox
            float x2 = x * x, x4 = x2 * x2;
            x = 15.5f * x4 * x2 - 40.14f * x4 * x + 31.96f * x4 - 6.868f * x2 * x + 0.4298f * x2 + 0.1191f * x -
                0.00232f;
        }
        if (look != AgxLook::None) { // ASC CDL around the Rec.709 luma
            float luma = v[0] * 0.2126f + v[1] * 0.7152f + v[2] * 0.0722f;
            V3 slope = {1.0f, 1.0f, 1.0f}, power = {1.0f, 1.0f, 1.0f};
            float sat = 1.0f;
            if (look == AgxLook::Golden) {
                slope = {1.0f, 0.9f, 0.5f};
                power = {0.8f, 0.8f, 0.8f};
                sat = 1.3f;
            } else {
                power = {1.35f, 1.35f, 1.35f};
                sat = 1.4f;
            }
            for (int k = 0; k < 3; ++k) {
                float c = std::pow(v[k] * slope[k] + 0.0f, power[k]);
                v[k] = luma + sat * (c - luma);
            }
        }
        v = mul(outset, v);
        for (float &x : v) x = std::pow(std::max(x, 0.0f), 2.2f);
        return clamp01(mul(kRec2020ToSrgb, v));
    }
    float exposure;
    AgxLook look;
    M3 outset;
};

struct KhronosPbrNeutralToneMapper final : ToneMapper { // tonemap.h:212-237
    explicit KhronosPbrNeutralToneMapper(float exposure) : exposure(exposure) {}
    V3 operator()(V3 v) const override {
        for (float &x : v) x *= exposure;
        const float f90 = 0.04f, ks = 0.8f - f90, kd = 0.15f;
        float x = std::min(std::min(v[0], v[1]), v[2]);
        float f = x <= 2.0f * f90 ? (x - x * x / (4.0f * f90)) : f90;
        float p = std::max(std::max(v[0] - f, v[1] - f), v[2] - f);
        float pn = 1.0f - (1.0f - ks) * (1.0f - ks) / (p + 1.0f - 2.0f * ks);
        float g = 1.0f / (kd * (p - pn) + 1.0f);
        V3 r;
        for (int k = 0; k < 3; ++k) r[k] = p <= ks ? v[k] - f : (v[k] - f) * (pn / p) * g + pn * (1.0f - g);
        return r;
    }
    float exposure;
};

} // namespace

std::unique_ptr<ToneMapper> create_tone_mapper(const ConfigArgs &args) {
    float exposure = args.load_float("exposure", 1.0f);
    std::string type = args.load_string("type");
    if (type == "aces") return std::make_unique<AcesToneMapper>(exposure);
    if (type == "agx") {
        std::string look = args.load_string("look", "none");
        AgxLook l;
        if (look == "none") l = AgxLook::None;
        else if (look == "golden") l = AgxLook::Golden;
        else if (look == "punchy") l = AgxLook::Punchy;
        else throw ConfigError("Invalid AgX tone mapper look [" + look + "]");
        return std::make_unique<AgxToneMapper>(exposure, l);
    }
    if (type == "khronos_pbr_neutral") return std::make_unique<KhronosPbrNeutralToneMapper>(exposure);
    throw ConfigError("Invalid tone mapper type [" + type + "]");
}

std::vector<uint8_t> composite_png(const double *sums, const double *weights, size_t n, const float backdrop[3],
                                   const ToneMapper *tm) {
    std::vector<uint8_t> out(n * 3);
    for (size_t i = 0; i < n; ++i) {
        double ws = weights[2 * i] + weights[2 * i + 1];
        V3 c;
        for (int k = 0; k < 3; ++k)
            c[k] = ws > 0.0 ? (float)((sums[3 * i + k] + weights[2 * i + 1] * (double)backdrop[k]) / ws) : 0.0f;
        if (tm) c = (*tm)(c);
        for (int k = 0; k < 3; ++k) {
            float s = linear_to_srgb(c[k]);
            float q = std::floor(s * 255.0f);
            out[3 * i + k] = (uint8_t)(q != q ? 0.0f : std::min(std::max(q, 0.0f), 255.0f));
        }
    }
    return out;
}

void tonemap_task(const ConfigArgs &args, const fs::path &task_dir, int task_id) {
    (void)task_dir;
    (void)task_id;
    fs::path in = args.load_path("input_path"), out = args.load_path("output_path");
    std::string ext = in.extension().string();
    int w = 0, h = 0;
    std::vector<float> rgb;
    if (ext == ".exr") rgb = read_exr(in, w, h);
    else if (ext == ".hdr") rgb = read_hdr(in, w, h);
    else throw ConfigError("Invalid image extension " + ext);
    if (args.contains("tone_mapper")) {
        std::unique_ptr<ToneMapper> tm = create_tone_mapper(args["tone_mapper"]);
        for (size_t i = 0; i < (size_t)w * h; ++i) {
            V3 c = {rgb[3 * i], rgb[3 * i + 1], rgb[3 * i + 2]};
            c = (*tm)(c);
            for (int k = 0; k < 3; ++k) rgb[3 * i + k] = c[k];
        }
    }
    write_exr(out, rgb.data(), w, h, 3, args.load_bool("half", false));
    if (args.contains("png_path")) { // sRGB-encoded 8-bit copy, as the PNG composite encodes
        std::vector<uint8_t> png(rgb.size());
        for (size_t i = 0; i < rgb.size(); ++i) {
            float q = std::floor(linear_to_srgb(rgb[i]) * 255.0f);
            png[i] = (uint8_t)(q != q ? 0.0f : std::min(std::max(q, 0.0f), 255.0f));
        }
        write_png(args.load_path("png_path"), png.data(), w, h, 3);
    }
}


// ---- PNG input ----------------------------------------------------------------------------------

namespace {

uint32_t be32(const uint8_t *p) { return (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | p[3]; }

int paeth(int a, int b, int c) {
    int p = a + b - c, pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
    return (pa <= pb && pa <= pc) ? a : (pb <= pc ? b : c);
}

// Undoes the five PNG row filters of one (sub)image of w x h with `bpp` bytes per complete pixel
// (at least 1) and `row` bytes per row; src holds h rows of 1 + row bytes.
void unfilter(const uint8_t *src, size_t avail, int h, size_t row, int bpp, std::vector<uint8_t> &out,
              const std::string &name) {
    if (avail < (size_t)h * (row + 1)) throw ConfigError(name + ": truncated PNG image data");
    out.assign((size_t)h * row, 0);
    for (int y = 0; y < h; ++y) {
        const uint8_t f = src[(size_t)y * (row + 1)];
        const uint8_t *in = src + (size_t)y * (row + 1) + 1;
        uint8_t *cur = out.data() + (size_t)y * row;
        const uint8_t *prev = y > 0 ? cur - row : nullptr;
        for (size_t i = 0; i < row; ++i) {
            const int a = i >= (size_t)bpp ? cur[i - bpp] : 0, b = prev ? prev[i] : 0,
                      c = (prev && i >= (size_t)bpp) ? prev[i - bpp] : 0;
            int v;
            switch (f) {
            case 0: v = in[i]; break;
            case 1: v = in[i] + a; break;
            case 2: v = in[i] + b; break;
            case 3: v = in[i] + ((a + b) >> 1); break;
            case 4: v = in[i] + paeth(a, b, c); break;
            default: throw ConfigError(name + ": bad PNG filter type");
            }
            cur[i] = (uint8_t)v;
        }
    }
}

} // namespace

// stb_image's conversions between channel counts (stbi__convert_format): gray = (77 r + 150 g + 29 b) >> 8
// (16-bit: (19595 r + 38470 g + 7471 b) >> 16), missing alpha = full scale.
template <typename T> static std::vector<T> convert_channels(const std::vector<T> &in, size_t n, int from, int to) {
    if (from == to) return in;
    const uint32_t full = sizeof(T) == 1 ? 255u : 65535u;
    auto luma = [&](const T *p) -> T {
        if (sizeof(T) == 1) return (T)(((uint32_t)p[0] * 77 + (uint32_t)p[1] * 150 + (uint32_t)p[2] * 29) >> 8);
        return (T)(((uint32_t)p[0] * 19595 + (uint32_t)p[1] * 38470 + (uint32_t)p[2] * 7471) >> 16);
    };
    std::vector<T> out(n * (size_t)to);
    for (size_t i = 0; i < n; ++i) {
        const T *s = &in[i * from];
        T *d = &out[i * to];
        const T g = from >= 3 ? luma(s) : s[0];
        const T a = (from == 2) ? s[1] : (from == 4 ? s[3] : (T)full);
        switch (to) {
        case 1: d[0] = g; break;
        case 2: d[0] = g, d[1] = a; break;
        case 3:
            if (from >= 3) d[0] = s[0], d[1] = s[1], d[2] = s[2];
            else d[0] = d[1] = d[2] = g;
            break;
        default:
            if (from >= 3) d[0] = s[0], d[1] = s[1], d[2] = s[2];
            else d[0] = d[1] = d[2] = g;
            d[3] = a;
        }
    }
    return out;
}

PngImage decode_png(const uint8_t *b, size_t n, const std::string &name) {
    static const uint8_t sig[8] = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1a, '\n'};
    if (n < 8 || std::memcmp(b, sig, 8) != 0) throw ConfigError(name + ": not a PNG file");
    size_t p = 8;
    uint32_t w = 0, h = 0;
    int depth = 0, ctype = -1, lace = 0;
    std::vector<uint8_t> idat, plte, trns;
    bool end = false;
    while (!end) {
        if (p + 12 > n) throw ConfigError(name + ": truncated PNG");
        const uint32_t len = be32(b + p);
        const std::string type((const char *)b + p + 4, 4);
        if (len > n - p - 12) throw ConfigError(name + ": truncated PNG chunk");
        const uint8_t *d = b + p + 8;
        if (type == "IHDR") {
            if (len != 13) throw ConfigError(name + ": bad IHDR");
            w = be32(d), h = be32(d + 4), depth = d[8], ctype = d[9], lace = d[12];
            if (d[10] != 0 || d[11] != 0 || lace > 1) throw ConfigError(name + ": unsupported PNG compression / filter / interlace");
        } else if (type == "PLTE") {
            plte.assign(d, d + len);
        } else if (type == "tRNS") {
            trns.assign(d, d + len);
        } else if (type == "IDAT") {
            idat.insert(idat.end(), d, d + len);
        } else if (type == "IEND") {
            end = true;
        }
        p += 12 + len;
    }
    const int chans = ctype == 0 ? 1 : ctype == 2 ? 3 : ctype == 3 ? 1 : ctype == 4 ? 2 : ctype == 6 ? 4 : 0;
    const bool depth_ok = (ctype == 0 && (depth == 1 || depth == 2 || depth == 4 || depth == 8 || depth == 16)) ||
                          (ctype == 3 && depth <= 8 && (depth & (depth - 1)) == 0) ||
                          ((ctype == 2 || ctype == 4 || ctype == 6) && (depth == 8 || depth == 16));
    if (!chans || !depth_ok || w == 0 || h == 0 || w > (1u << 24) || h > (1u << 24))
        throw ConfigError(name + ": unsupported PNG colour type / bit depth");
    if (ctype == 3 && (plte.empty() || plte.size() % 3)) throw ConfigError(name + ": palette image without PLTE");
    // inflate
    const int bits_pp = chans * depth;
    auto row_bytes = [&](uint32_t ww) { return ((size_t)ww * bits_pp + 7) / 8; };
    size_t raw_size = 0;
    static const int ax0[7] = {0, 4, 0, 2, 0, 1, 0}, ay0[7] = {0, 0, 4, 0, 2, 0, 1};
    static const int adx[7] = {8, 8, 4, 4, 2, 2, 1}, ady[7] = {8, 8, 8, 4, 4, 2, 2};
    auto pass_dims = [&](int k, uint32_t &pw, uint32_t &ph) {
        pw = w > (uint32_t)ax0[k] ? (w - ax0[k] + adx[k] - 1) / adx[k] : 0;
        ph = h > (uint32_t)ay0[k] ? (h - ay0[k] + ady[k] - 1) / ady[k] : 0;
    };
    if (lace) {
        for (int k = 0; k < 7; ++k) {
            uint32_t pw, ph;
            pass_dims(k, pw, ph);
            if (pw && ph) raw_size += (size_t)ph * (row_bytes(pw) + 1);
        }
    } else {
        raw_size = (size_t)h * (row_bytes(w) + 1);
    }
    std::vector<uint8_t> raw(raw_size);
    uLongf got = (uLongf)raw_size;
    if (uncompress(raw.data(), &got, idat.data(), (uLong)idat.size()) != Z_OK || got != raw_size)
        throw ConfigError(name + ": bad PNG image data");
    // unfilter and unpack to one sample per channel (16-bit samples kept as 16-bit)
    const int bpp = std::max(1, bits_pp / 8);
    std::vector<uint16_t> samples((size_t)w * h * chans);
    auto unpack_into = [&](const std::vector<uint8_t> &rows, uint32_t pw, uint32_t ph, int x0, int y0, int dx, int dy) {
        const size_t rb = row_bytes(pw);
        for (uint32_t y = 0; y < ph; ++y) {
            const uint8_t *r = rows.data() + (size_t)y * rb;
            for (uint32_t x = 0; x < pw; ++x)
                for (int c = 0; c < chans; ++c) {
                    uint32_t v;
                    const size_t bit = ((size_t)x * chans + c) * depth;
                    if (depth == 16) v = (uint32_t)r[bit / 8] << 8 | r[bit / 8 + 1];
                    else if (depth == 8) v = r[bit / 8];
                    else v = (r[bit / 8] >> (8 - depth - (bit % 8))) & ((1u << depth) - 1);
                    samples[((size_t)(y0 + y * dy) * w + (x0 + x * dx)) * chans + c] = (uint16_t)v;
                }
        }
    };
    std::vector<uint8_t> rows;
    if (lace) {
        size_t off = 0;
        for (int k = 0; k < 7; ++k) {
            uint32_t pw, ph;
            pass_dims(k, pw, ph);
            if (!pw || !ph) continue;
            unfilter(raw.data() + off, raw.size() - off, (int)ph, row_bytes(pw), bpp, rows, name);
            unpack_into(rows, pw, ph, ax0[k], ay0[k], adx[k], ady[k]);
            off += (size_t)ph * (row_bytes(pw) + 1);
        }
    } else {
        unfilter(raw.data(), raw.size(), (int)h, row_bytes(w), bpp, rows, name);
        unpack_into(rows, w, h, 0, 0, 1, 1);
    }
    // expand as stb_image does: palette to RGB / RGBA, low-depth gray to 8 bits, tRNS keys to alpha
    PngImage img;
    img.w = (int)w, img.h = (int)h;
    const size_t np = (size_t)w * h;
    img.bits = depth == 16 ? 16 : 8;
    std::vector<uint16_t> out;
    if (ctype == 3) {
        const bool alpha = !trns.empty();
        img.comp = alpha ? 4 : 3;
        out.resize(np * img.comp);
        const size_t ncol = plte.size() / 3;
        for (size_t i = 0; i < np; ++i) {
            const size_t k = samples[i];
            if (k >= ncol) throw ConfigError(name + ": palette index out of range");
            for (int c = 0; c < 3; ++c) out[i * img.comp + c] = plte[3 * k + c];
            if (alpha) out[i * 4 + 3] = k < trns.size() ? trns[k] : 255;
        }
    } else {
        const int scale = depth >= 8 ? 1 : (depth == 4 ? 0x11 : (depth == 2 ? 0x55 : 0xff));
        const bool key = !trns.empty() && (ctype == 0 || ctype == 2);
        img.comp = chans + (key ? 1 : 0);
        out.resize(np * img.comp);
        uint16_t kv[3] = {0, 0, 0};
        if (key) {
            if (trns.size() < (size_t)(2 * chans)) throw ConfigError(name + ": bad tRNS");
            for (int c = 0; c < chans; ++c) kv[c] = (uint16_t)((trns[2 * c] << 8 | trns[2 * c + 1]) & ((1u << depth) - 1));
        }
        const uint16_t full = depth == 16 ? 65535 : 255;
        for (size_t i = 0; i < np; ++i) {
            bool match = key;
            for (int c = 0; c < chans; ++c) {
                const uint16_t v = samples[i * chans + c];
                match = match && v == kv[c];
                out[i * img.comp + c] = (uint16_t)(v * scale);
            }
            if (key) out[i * img.comp + chans] = match ? 0 : full;
        }
    }
    if (img.bits == 16) {
        img.data16 = std::move(out);
    } else {
        img.data8.resize(out.size());
        for (size_t i = 0; i < out.size(); ++i) img.data8[i] = (uint8_t)out[i];
    }
    return img;
}

PngImage read_png(const fs::path &path) {
    std::vector<uint8_t> f = read_all(path);
    return decode_png(f.data(), f.size(), path.string());
}

void PngImage::to_channels(int c) {
    if (c < 1 || c > 4) throw ConfigError("1 to 4 channels can be requested");
    if (bits == 16) data16 = convert_channels(data16, (size_t)w * h, comp, c);
    else data8 = convert_channels(data8, (size_t)w * h, comp, c);
    comp = c;
}

void PngImage::to_8bit() {
    if (bits != 16) return;
    data8.resize(data16.size());
    for (size_t i = 0; i < data16.size(); ++i) data8[i] = (uint8_t)(data16[i] >> 8); // stbi__convert_16_to_8
    data16.clear();
    bits = 8;
}

} // namespace ks

// The reference's texture levels on the host, without HIP: the NPOT mip chain (Texture ctor,
// texture.cpp:21-347) built on the source data type with its quantisation (Texture::set_from_float,
// texture.cpp:462-518). Shared by the library (which packs the levels for the device) and the host tool
// (which writes serialized textures).
#pragma once
#include "../../include/keepsake_amd.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace ksa {

constexpr int kMaxTexLevels = 16;

// The reference's sRGB u8 -> linear table (srgb_u8_to_linear_f32_lut, texture.cpp:376-401): srgb_to_linear
// of i / 255 evaluated in float, which reproduces that table bit for bit.
inline void fill_srgb_u8_table(float out[256]) {
    for (int i = 0; i < 256; ++i) {
        const float x = (float)i / 255.0f;
        out[i] = x < 0.04045f ? x / 12.92f : std::pow((x + 0.055f) / 1.055f, 2.4f);
    }
}

inline int tex_bytes(int dtype) { return dtype == KS_TEX_U8 ? 1 : (dtype == KS_TEX_U16 ? 2 : 4); }

inline int tex_full_levels(int w, int h) { return 1 + (int)std::ceil(std::log2((double)std::max(w, h))); }

// A texture in its source layout (channels x data type per texel), level by level.
struct SourceTexture {
    int w = 0, h = 0, nc = 0, dtype = 0, srgb = 0;
    std::vector<std::vector<uint8_t>> levels;
    std::vector<std::pair<int, int>> dims;
    const float *srgb_u8 = nullptr; // 256 entries

    static float srgb_to_linear(float x) { return x < 0.04045f ? x / 12.92f : std::pow((x + 0.055f) / 1.055f, 2.4f); }
    static float linear_to_srgb(float x) { // maths.h:649-659
        if (x <= 0.0f) return 0.0f;
        if (x >= 1.0f) return 1.0f;
        if (x < 0.0031308f) return x * 12.92f;
        return std::pow(x, 1.0f / 2.4f) * 1.055f - 0.055f;
    }
    // Texture::fetch_as_float (texture.cpp:403-460)
    void fetch(int l, int x, int y, float *out) const {
        const uint8_t *p = levels[l].data() + ((size_t)y * dims[l].first + x) * (size_t)(nc * tex_bytes(dtype));
        for (int c = 0; c < nc; ++c) {
            if (dtype == KS_TEX_U8) {
                out[c] = srgb ? srgb_u8[p[c]] : (float)p[c] / 255.0f;
            } else if (dtype == KS_TEX_U16) {
                uint16_t v;
                std::memcpy(&v, p + 2 * c, 2);
                out[c] = (float)v / 65535.0f;
                if (srgb) out[c] = srgb_to_linear(out[c]);
            } else {
                float v;
                std::memcpy(&v, p + 4 * c, 4);
                out[c] = srgb ? srgb_to_linear(v) : v;
            }
        }
    }
    // Texture::set_from_float (texture.cpp:462-518): floor quantisation, sRGB re-encoding
    void set(int l, int x, int y, const float *in) {
        uint8_t *p = levels[l].data() + ((size_t)y * dims[l].first + x) * (size_t)(nc * tex_bytes(dtype));
        for (int c = 0; c < nc; ++c) {
            if (dtype == KS_TEX_U8) {
                const float f = srgb ? linear_to_srgb(in[c]) : in[c];
                p[c] = (uint8_t)std::floor(f * 255.0f);
            } else if (dtype == KS_TEX_U16) {
                const float f = srgb ? linear_to_srgb(in[c]) : in[c];
                const uint16_t v = (uint16_t)std::floor(f * 65535.0f);
                std::memcpy(p + 2 * c, &v, 2);
            } else {
                const float v = srgb ? linear_to_srgb(in[c]) : in[c];
                std::memcpy(p + 4 * c, &v, 4);
            }
        }
    }
};

// One axis of the NPOT downsampling filter (texture.cpp:36-344): from an odd size 2n - 1 three taps
// with weights x / (2n - 1), n / (2n - 1), (n - x - 1) / (2n - 1) on 2x - 1, 2x, 2x + 1 (taps outside
// the level are dropped); from an even size two taps of 1/2; from size 1 the texel itself.
inline int mip_taps(int last, int n, int x, int idx[3], float wt[3]) {
    if (last == 1) {
        idx[0] = x;
        wt[0] = 1.0f;
        return 1;
    }
    if (last == 2 * n - 1) {
        int k = 0;
        const int i0 = 2 * x - 1, i1 = 2 * x, i2 = 2 * x + 1;
        const float w0 = (float)x / (float)last, w1 = (float)n / (float)last, w2 = (float)(n - x - 1) / (float)last;
        if (i0 > 0) idx[k] = i0, wt[k++] = w0;
        idx[k] = i1, wt[k++] = w1;
        if (i2 < last) idx[k] = i2, wt[k++] = w2;
        return k;
    }
    idx[0] = 2 * x, wt[0] = 0.5f;
    idx[1] = 2 * x + 1, wt[1] = 0.5f;
    return 2;
}

// Level l from level l - 1: rows of taps outer, columns inner, each term (wx wy) * v as the reference
// accumulates it (with a weight of 1 on an axis of size 1).
inline void build_level(SourceTexture &t, int l) {
    const int lw = t.dims[l - 1].first, lh = t.dims[l - 1].second;
    const int w = t.dims[l].first, h = t.dims[l].second;
    std::vector<float> v(t.nc), sum(t.nc);
    for (int y = 0; y < h; ++y) {
        int iy[3];
        float wy[3];
        const int ny = mip_taps(lh, h, y, iy, wy);
        for (int x = 0; x < w; ++x) {
            int ix[3];
            float wx[3];
            const int nx = mip_taps(lw, w, x, ix, wx);
            std::fill(sum.begin(), sum.end(), 0.0f);
            for (int b = 0; b < ny; ++b)
                for (int a = 0; a < nx; ++a) {
                    t.fetch(l - 1, ix[a], iy[b], v.data());
                    // a size-1 axis contributes no factor; 0.25 = 0.5 * 0.5 exactly
                    const float wgt = lh == 1 ? wx[a] : (lw == 1 ? wy[b] : wx[a] * wy[b]);
                    for (int c = 0; c < t.nc; ++c) sum[c] += wgt * v[c];
                }
            t.set(l, x, y, sum.data());
        }
    }
}

// Validates a texture description and returns it in source layout with every level present.
inline SourceTexture load_source_texture(const ks_texture_desc &d, const float *srgb_u8) {
    if (!d.data) throw std::invalid_argument("texture without data");
    if (d.width <= 0 || d.height <= 0) throw std::invalid_argument("texture size must be positive");
    if (d.channels < 1 || d.channels > 4) throw std::invalid_argument("texture channels must be 1..4");
    if (d.data_type < KS_TEX_U8 || d.data_type > KS_TEX_F32) throw std::invalid_argument("unknown texture data type");
    if (d.color_space != KS_COLORSPACE_LINEAR && d.color_space != KS_COLORSPACE_SRGB)
        throw std::invalid_argument("unknown texture colour space");
    const int full = tex_full_levels(d.width, d.height);
    if (full > kMaxTexLevels) throw std::invalid_argument("texture larger than 32768 texels on a side");
    if (d.levels < 1 || d.levels > full) throw std::invalid_argument("texture level count out of range");
    SourceTexture t;
    t.w = d.width, t.h = d.height, t.nc = d.channels, t.dtype = d.data_type;
    t.srgb = d.color_space == KS_COLORSPACE_SRGB;
    t.srgb_u8 = srgb_u8;
    const int n_levels = (d.levels == 1 && d.build_mipmaps) ? full : d.levels;
    int w = d.width, h = d.height;
    const size_t texel = (size_t)d.channels * tex_bytes(d.data_type);
    const uint8_t *src = (const uint8_t *)d.data;
    for (int l = 0; l < n_levels; ++l) {
        t.dims.emplace_back(w, h);
        t.levels.emplace_back((size_t)w * h * texel);
        if (l < d.levels) {
            std::memcpy(t.levels.back().data(), src, t.levels.back().size());
            src += t.levels.back().size();
        }
        w = std::max(1, (w + 1) / 2);
        h = std::max(1, (h + 1) / 2);
    }
    for (int l = d.levels; l < n_levels; ++l) build_level(t, l);
    return t;
}

} // namespace ksa

// Host side of textures in the library: the repacking of a texture's levels (ks_mip.h) into the device
// layout of ks_texture.h, shader-field records and the lookup tables.
#pragma once
#include "../../include/keepsake_amd.h"
#include "ks_mip.h"
#include "ks_texture.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

namespace ksa {

using ksd::DevField;
using ksd::DevTexture;
using ksd::TexLuts;

// sRGB u8 -> linear table (the reference's srgb_u8_to_linear_f32_lut, texture.cpp:376-401) and the EWA
// weights exp(-2 r^2) - exp(-2) at r^2 = i / 127 (texture.cpp:720-727). Both tables are the single-
// precision evaluations of these formulas (srgb_to_linear of i / 255 in float; the EWA generator of
// texture.cpp:720-726 in float), which reproduce the reference's tables bit for bit.
inline void fill_tex_luts(TexLuts &l) {
    fill_srgb_u8_table(l.srgb_u8);
    for (int i = 0; i < ksd::kEwaLutSize; ++i) {
        const float alpha = 2.0f, r2 = (float)i / (float)(ksd::kEwaLutSize - 1);
        l.ewa[i] = std::exp(-alpha * r2) - std::exp(-alpha);
    }
}

// Device layout: three channels padded to four, levels 256-byte aligned. Returns the blob; meta.data is
// left for the caller to point at the uploaded copy (its offset within the pool is meta_offset).
inline std::vector<uint8_t> pack_texture(const SourceTexture &t, DevTexture &meta) {
    const int b = tex_bytes(t.dtype);
    const int snc = t.nc == 3 ? 4 : t.nc;
    meta = DevTexture{};
    meta.levels = (int32_t)t.levels.size();
    meta.nc = t.nc;
    meta.dtype = t.dtype;
    meta.srgb = t.srgb;
    meta.texel_bytes = snc * b;
    meta.width = t.w;
    meta.height = t.h;
    size_t off = 0;
    for (int l = 0; l < meta.levels; ++l) {
        off = (off + 255) & ~(size_t)255;
        if (off > 0xffffffffull) throw std::invalid_argument("texture larger than 4 GiB");
        meta.lv[l].off = (uint32_t)off;
        meta.lv[l].w = t.dims[l].first;
        meta.lv[l].h = t.dims[l].second;
        off += (size_t)t.dims[l].first * t.dims[l].second * meta.texel_bytes;
    }
    std::vector<uint8_t> blob((off + 255) & ~(size_t)255, 0);
    const size_t src_texel = (size_t)t.nc * b;
    for (int l = 0; l < meta.levels; ++l) {
        const size_t n = (size_t)t.dims[l].first * t.dims[l].second;
        uint8_t *dst = blob.data() + meta.lv[l].off;
        const uint8_t *src = t.levels[l].data();
        for (size_t i = 0; i < n; ++i) std::memcpy(dst + i * meta.texel_bytes, src + i * src_texel, src_texel);
    }
    return blob;
}

// ks_field_desc -> DevField (create_texture_shader_field_color / create_texture_sampler)
inline DevField to_dev_field(const ks_field_desc &f, const std::vector<DevTexture> &tex, int components,
                             const char *name) {
    DevField d{};
    d.tex = f.texture;
    for (int k = 0; k < 4; ++k) d.value[k] = f.value[k];
    if (f.texture < 0) return d;
    auto fail = [&](const char *what) { throw std::invalid_argument(std::string(name) + ": " + what); };
    if (f.texture >= (int)tex.size()) fail("texture index out of range");
    if (f.sampler < KS_SAMPLER_NEAREST || f.sampler > KS_SAMPLER_EWA) fail("unknown sampler");
    if (f.wrap_u < KS_WRAP_REPEAT || f.wrap_u > KS_WRAP_CLAMP || f.wrap_v < KS_WRAP_REPEAT || f.wrap_v > KS_WRAP_CLAMP)
        fail("unknown wrap mode");
    const int nc = tex[f.texture].nc;
    for (int k = 0; k < components; ++k)
        if (f.swizzle[k] >= nc || f.swizzle[k] < -1) fail("swizzle names a channel the texture does not have");
    d.sampler = f.sampler;
    d.wrap_u = f.wrap_u;
    d.wrap_v = f.wrap_v;
    d.flip_v = f.flip_v ? 1 : 0;
    for (int k = 0; k < 4; ++k) {
        d.swizzle[k] = k < components ? f.swizzle[k] : -1;
        d.scale[k] = f.scale[k];
    }
    d.anisotropy = std::min(std::max(f.anisotropy, 1.0f), 16.0f);
    for (int k = 0; k < 2; ++k) {
        d.uv_scale[k] = f.uv_scale[k];
        d.uv_offset[k] = f.uv_offset[k];
    }
    // CubicSampler kernels (texture.h:92-116): vec4(...) / 18.0 etc. in float
    static const float kCa[3][4] = {{21.0f / 18.0f, -36.0f / 18.0f, 0.0f / 18.0f, 16.0f / 18.0f},
                                    {3.0f / 6.0f, -6.0f / 6.0f, 0.0f / 6.0f, 4.0f / 6.0f},
                                    {3.0f / 2.0f, -5.0f / 2.0f, 0.0f / 2.0f, 2.0f / 2.0f}};
    static const float kCb[3][4] = {{-7.0f / 18.0f, 36.0f / 18.0f, -60.0f / 18.0f, 32.0f / 18.0f},
                                    {-1.0f / 6.0f, 6.0f / 6.0f, -12.0f / 6.0f, 8.0f / 6.0f},
                                    {-1.0f / 2.0f, 5.0f / 2.0f, -8.0f / 2.0f, 4.0f / 2.0f}};
    int k = f.cubic_kernel;
    if (k < KS_CUBIC_MITCHELL || k > KS_CUBIC_CATMULL_ROM) fail("unknown cubic kernel");
    for (int j = 0; j < 4; ++j) {
        d.ca[j] = kCa[k][j];
        d.cb[j] = kCb[k][j];
    }
    return d;
}

} // namespace ksa

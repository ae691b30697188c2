// Textures and textured shader fields (texture.h, texture.cpp:403-920, shader_field.h:37-86). The same
// functions run on the device (per-hit material evaluation, normal and opacity maps, textured emission)
// and on the host (mesh-light importance). A texture keeps the reference's data types (u8, u16, f32)
// and colour spaces (linear, sRGB) in HBM; three-channel textures are stored with four channels so a
// texel is one aligned 4- or 16-byte load. Levels follow the reference's NPOT chain (each level
// max(1, (n + 1) / 2) of the one before), laid out back to back.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

namespace ksd {

#ifndef KS_HD
#define KS_HD __host__ __device__ inline
#endif

constexpr int kTexU8 = 0, kTexU16 = 1, kTexF32 = 2;
constexpr int kSampNearest = 0, kSampLinear = 1, kSampCubic = 2, kSampEwa = 3;
constexpr int kWrapRepeat = 0, kWrapClamp = 1;
constexpr int kTexMaxLevels = 16;
constexpr int kEwaLutSize = 128;

struct TexLevel {
    uint32_t off; // byte offset of the level inside the texture's data
    int32_t w, h;
};

struct DevTexture {
    const uint8_t *data;
    int32_t levels, nc, dtype, srgb;
    int32_t texel_bytes; // bytes per stored texel (3 channels stored as 4)
    int32_t width, height, pad;
    TexLevel lv[kTexMaxLevels];
};

// A shader field: a constant or a texture lookup (TextureField<N>, shader_field.h:37-86).
struct DevField {
    int32_t tex;            // texture index, -1: constant `value`
    int32_t sampler;        // kSamp*
    int32_t wrap_u, wrap_v; // kWrap*
    int32_t flip_v;
    int32_t swizzle[4];     // source channel per component, -1: 0
    float anisotropy;       // EWA: clamp of the major / minor axis ratio
    float uv_scale[2], uv_offset[2];
    float scale[4];
    float value[4];
    float ca[4], cb[4];     // cubic kernel polynomials (CubicSampler ctor, texture.h:92-116)
};

// A material whose closure inputs are shader fields (PrincipledBSDF::eval_closure with
// Material::normal_map / opacity_map / emission). Fields in eval_closure order.
constexpr int kFieldBasecolor = 0, kFieldRoughness = 1, kFieldMetallic = 2, kFieldIor = 3, kFieldR0Mul = 4,
              kFieldSpecTrans = 5, kFieldDiffTrans = 6, kFieldDiffTransFwd = 7, kFieldEmission = 8;
struct DevTexMaterial {
    DevField f[9];
    DevField normal, opacity;
    int32_t has_emission, has_normal, has_opacity;
    int32_t normal_world, normal_range01;
    float normal_strength;
    float normal_it[9]; // world-space maps: inverse transpose of the map's to_world, row-major
};

// Lookup tables shared by all textures: sRGB u8 -> linear, and the EWA Gaussian weights.
struct TexLuts {
    float srgb_u8[256];
    float ewa[kEwaLutSize];
};

struct TexContext {
    const DevTexture *tex;
    const TexLuts *luts;
};

KS_HD float tex_srgb_to_linear(float x) { // maths.h:631-638
    return x < 0.04045f ? x / 12.92f : powf((x + 0.055f) / 1.055f, 2.4f);
}

// std::lerp as libstdc++ computes it (P0811: exact at both ends, monotone)
KS_HD float std_lerp(float a, float b, float t) {
    if ((a <= 0.0f && b >= 0.0f) || (a >= 0.0f && b <= 0.0f)) return t * b + (1.0f - t) * a;
    if (t == 1.0f) return b;
    const float x = a + t * (b - a);
    return (t > 1.0f) == (b > a) ? (b < x ? x : b) : (x < b ? x : b);
}

KS_HD int tex_wrap(int x, int dim, int mode) { // texture.cpp:520-531
    if (mode == kWrapRepeat) {
        int m = x % dim;
        return m < 0 ? m + dim : m;
    }
    return x < 0 ? 0 : (x > dim - 1 ? dim - 1 : x);
}

// Texture::fetch_as_float (texture.cpp:403-460): out[0..nc) in linear space
KS_HD void tex_fetch(const TexContext &c, const DevTexture &t, int x, int y, int level, float out[4]) {
    const TexLevel &l = t.lv[level];
    const uint8_t *p = t.data + l.off + ((size_t)y * l.w + x) * t.texel_bytes;
    if (t.dtype == kTexU8) {
        uint32_t v;
        if (t.texel_bytes == 4) {
            v = *(const uint32_t *)p;
        } else if (t.texel_bytes == 2) {
            v = *(const uint16_t *)p;
        } else {
            v = p[0];
        }
        for (int k = 0; k < 4; ++k) {
            if (k >= t.nc) break;
            uint32_t b = (v >> (8 * k)) & 255u;
            out[k] = t.srgb ? c.luts->srgb_u8[b] : (float)b / 255.0f;
        }
    } else if (t.dtype == kTexU16) {
        for (int k = 0; k < 4; ++k) {
            if (k >= t.nc) break;
            float f = (float)((const uint16_t *)p)[k] / 65535.0f;
            out[k] = t.srgb ? tex_srgb_to_linear(f) : f;
        }
    } else {
        float4 f4;
        if (t.texel_bytes == 16) {
            f4 = *(const float4 *)p;
        } else {
            const float *f = (const float *)p;
            f4.x = f[0];
            f4.y = t.nc > 1 ? f[1] : 0.0f;
            f4.z = 0.0f;
            f4.w = 0.0f;
        }
        const float v[4] = {f4.x, f4.y, f4.z, f4.w};
        for (int k = 0; k < 4; ++k) {
            if (k >= t.nc) break;
            out[k] = t.srgb ? tex_srgb_to_linear(v[k]) : v[k];
        }
    }
}

// TextureSampler::bilinear (texture.cpp:541-573)
KS_HD void tex_bilinear(const TexContext &c, const DevTexture &t, int level, float u_in, float v_in, int wu, int wv,
                        float out[4]) {
    const TexLevel &l = t.lv[level];
    const float u = u_in * (float)l.w - 0.5f, v = v_in * (float)l.h - 0.5f;
    int u0 = (int)floorf(u), v0 = (int)floorf(v);
    const float du = u - (float)u0, dv = v - (float)v0;
    u0 = tex_wrap(u0, l.w, wu);
    v0 = tex_wrap(v0, l.h, wv);
    const int u1 = tex_wrap(u0 + 1, l.w, wu), v1 = tex_wrap(v0 + 1, l.h, wv);
    const float w00 = (1 - du) * (1 - dv), w10 = du * (1 - dv), w01 = (1 - du) * dv, w11 = du * dv;
    float a[4], b[4], e[4], d[4];
    tex_fetch(c, t, u0, v0, level, a);
    tex_fetch(c, t, u1, v0, level, b);
    tex_fetch(c, t, u0, v1, level, e);
    tex_fetch(c, t, u1, v1, level, d);
    for (int k = 0; k < 4; ++k) {
        if (k >= t.nc) break;
        out[k] = w00 * a[k] + w10 * b[k] + w01 * e[k] + w11 * d[k];
    }
}

KS_HD float tex_dot4(const float a[4], float x) { // ca.dot(powers(x)), powers = (x^3, x^2, x, 1)
    return a[0] * (x * x * x) + a[1] * (x * x) + a[2] * x + a[3];
}

// CubicSampler::bicubic (texture.cpp:629-677)
KS_HD void tex_bicubic(const TexContext &c, const DevTexture &t, const DevField &f, int level, float u_in,
                       float v_in, float out[4]) {
    const TexLevel &l = t.lv[level];
    const float u = u_in * (float)l.w - 0.5f, v = v_in * (float)l.h - 0.5f;
    const int u0 = (int)floorf(u), v0 = (int)floorf(v);
    const float du = u - (float)u0, dv = v - (float)v0;
    // spline weights (texture.cpp:631-641), evaluated in float as the reference's vec4 dot products
    // (x + 1.0 etc. are double in the reference; rounded once to float they equal these float sums)
    const float au[4] = {tex_dot4(f.cb, du + 1.0f), tex_dot4(f.ca, du), tex_dot4(f.ca, 1.0f - du),
                         tex_dot4(f.cb, 2.0f - du)};
    const float av[4] = {tex_dot4(f.cb, dv + 1.0f), tex_dot4(f.ca, dv), tex_dot4(f.ca, 1.0f - dv),
                         tex_dot4(f.cb, 2.0f - dv)};
    float rows[4][4];
    for (int y = 0; y < 4; ++y) {
        const int vy = tex_wrap(v0 + y - 1, l.h, f.wrap_v);
        float cols[4][4];
        for (int x = 0; x < 4; ++x) tex_fetch(c, t, tex_wrap(u0 + x - 1, l.w, f.wrap_u), vy, level, cols[x]);
        for (int k = 0; k < 4; ++k) {
            if (k >= t.nc) break;
            rows[y][k] = cols[0][k] * au[0] + cols[1][k] * au[1] + cols[2][k] * au[2] + cols[3][k] * au[3];
        }
    }
    for (int k = 0; k < 4; ++k) {
        if (k >= t.nc) break;
        out[k] = rows[0][k] * av[0] + rows[1][k] * av[1] + rows[2][k] * av[2] + rows[3][k] * av[3];
    }
}

// EWASampler::ewa (texture.cpp:858-920)
KS_HD void tex_ewa(const TexContext &c, const DevTexture &t, const DevField &f, int level, float u, float v,
                   float maj0, float maj1, float min0, float min1, float out[4]) {
    const TexLevel &l = t.lv[level];
    u = u * (float)l.w - 0.5f;
    v = v * (float)l.h - 0.5f;
    maj0 *= (float)l.w;
    maj1 *= (float)l.h;
    min0 *= (float)l.w;
    min1 *= (float)l.h;
    float A = maj1 * maj1 + min1 * min1 + 1;
    float B = -2 * (maj0 * maj1 + min0 * min1);
    float C = maj0 * maj0 + min0 * min0 + 1;
    const float inv_f = 1 / (A * C - B * B * 0.25f);
    A *= inv_f;
    B *= inv_f;
    C *= inv_f;
    const float det = -(B * B) + 4 * A * C;
    const float inv_det = 1 / det;
    const float us = sqrtf(fmaxf(0.0f, det * C)), vs = sqrtf(fmaxf(0.0f, A * det));
    const float fu0 = ceilf(u - 2 * inv_det * us), fu1 = floorf(u + 2 * inv_det * us);
    const float fv0 = ceilf(v - 2 * inv_det * vs), fv1 = floorf(v + 2 * inv_det * vs);
    // guard: a non-finite footprint or texture coordinate would give an unbounded scan; finite
    // footprints at the chosen level span a few dozen texels at most
    if (!(fu1 - fu0 <= 4096.0f && fv1 - fv0 <= 4096.0f && fabsf(fu0) < 1e9f && fabsf(fv0) < 1e9f)) {
        tex_bilinear(c, t, level, (u + 0.5f) / (float)l.w, (v + 0.5f) / (float)l.h, f.wrap_u, f.wrap_v, out);
        return;
    }
    const int u0 = (int)fu0, u1 = (int)fu1, v0 = (int)fv0, v1 = (int)fv1;
    float acc[4] = {0, 0, 0, 0};
    float sum_w = 0;
    for (int iv = v0; iv <= v1; ++iv) {
        const float vv = (float)iv - v;
        const int ivw = tex_wrap(iv, l.h, f.wrap_v);
        for (int iu = u0; iu <= u1; ++iu) {
            const float uu = (float)iu - u;
            const float r2 = A * (uu * uu) + B * uu * vv + C * (vv * vv);
            if (r2 < 1) {
                const int idx = (int)fminf(r2 * (float)kEwaLutSize, (float)(kEwaLutSize - 1));
                const float w = c.luts->ewa[idx];
                float texel[4];
                tex_fetch(c, t, tex_wrap(iu, l.w, f.wrap_u), ivw, level, texel);
                for (int k = 0; k < 4; ++k) {
                    if (k >= t.nc) break;
                    acc[k] += w * texel[k];
                }
                sum_w += w;
            }
        }
    }
    const float inv_sum = sum_w > 0.0f ? 1.0f / sum_w : 1.0f;
    for (int k = 0; k < 4; ++k) {
        if (k >= t.nc) break;
        out[k] = sum_w > 0.0f ? acc[k] * inv_sum : acc[k];
    }
}

// One level of the chosen filter (bilinear / bicubic / EWA) at `level`.
KS_HD void tex_filter_level(const TexContext &c, const DevTexture &t, const DevField &f, int level, float u, float v,
                            const float d[4], float out[4]) {
    if (f.sampler == kSampCubic) {
        tex_bicubic(c, t, f, level, u, v, out);
    } else if (f.sampler == kSampEwa) {
        tex_ewa(c, t, f, level, u, v, d[0], d[1], d[2], d[3], out);
    } else {
        tex_bilinear(c, t, level, u, v, f.wrap_u, f.wrap_v, out);
    }
}

// The samplers (texture.cpp:575-627, 679-716). d = (dudx, dvdx, dudy, dvdy).
KS_HD void tex_sample(const TexContext &c, const DevTexture &t, const DevField &f, float u, float v, const float d_in[4],
                      float out[4]) {
    if (f.sampler == kSampNearest) {
        const TexLevel &l = t.lv[0];
        const int x = tex_wrap((int)floorf(u * (float)l.w - 0.5f), l.w, f.wrap_u);
        const int y = tex_wrap((int)floorf(v * (float)l.h - 0.5f), l.h, f.wrap_v);
        tex_fetch(c, t, x, y, 0, out);
        return;
    }
    float d[4] = {d_in[0], d_in[1], d_in[2], d_in[3]};
    float width;
    if (f.sampler == kSampEwa) {
        float maj0 = d[0], maj1 = d[1], min0 = d[2], min1 = d[3];
        float len_maj = sqrtf(maj0 * maj0 + maj1 * maj1), len_min = sqrtf(min0 * min0 + min1 * min1);
        if (len_maj < len_min) {
            float s0 = maj0, s1 = maj1, sl = len_maj;
            maj0 = min0, maj1 = min1, len_maj = len_min;
            min0 = s0, min1 = s1, len_min = sl;
        }
        if (len_min * f.anisotropy < len_maj && len_min > 0) {
            const float s = len_maj / (len_min * f.anisotropy);
            min0 *= s;
            min1 *= s;
            len_min *= s;
        }
        if (len_min == 0) {
            tex_bilinear(c, t, 0, u, v, f.wrap_u, f.wrap_v, out);
            return;
        }
        d[0] = maj0, d[1] = maj1, d[2] = min0, d[3] = min1;
        width = len_min;
    } else {
        width = fmaxf(fmaxf(fabsf(d[0]), fabsf(d[1])), fmaxf(fabsf(d[2]), fabsf(d[3])));
    }
    const float level = (float)(t.levels - 1) + log2f(fmaxf(width, 1e-8f));
    if (level < 0 || t.levels == 1) {
        tex_filter_level(c, t, f, 0, u, v, d, out);
    } else if (level >= (float)(t.levels - 1)) {
        tex_fetch(c, t, 0, 0, t.levels - 1, out);
    } else {
        const int il = (int)floorf(level);
        const float delta = level - (float)il;
        float o0[4], o1[4];
        tex_filter_level(c, t, f, il, u, v, d, o0);
        tex_filter_level(c, t, f, il + 1, u, v, d, o1);
        for (int k = 0; k < 4; ++k) {
            if (k >= t.nc) break;
            out[k] = std_lerp(o0[k], o1[k], delta);
        }
    }
}

// ShaderField<color<N>>::operator() at (uv, duvdxy): ConstantField or TextureField (shader_field.h:37-86).
// d = (dudx, dvdx, dudy, dvdy); n components are written.
KS_HD void field_eval(const TexContext &c, const DevField &f, float u, float v, const float d_in[4], int n,
                      float out[4]) {
    if (f.tex < 0) {
        for (int k = 0; k < 4; ++k)
            if (k < n) out[k] = f.value[k];
        return;
    }
    float d[4] = {d_in[0], d_in[1], d_in[2], d_in[3]};
    if (f.flip_v) {
        v = 1.0f - v;
        d[1] *= -1.0f;
        d[3] *= -1.0f;
    }
    u = u * f.uv_scale[0] + f.uv_offset[0];
    v = v * f.uv_scale[1] + f.uv_offset[1];
    d[0] *= f.uv_scale[0];
    d[2] *= f.uv_scale[0];
    d[1] *= f.uv_scale[1];
    d[3] *= f.uv_scale[1];
    const DevTexture &t = c.tex[f.tex];
    float s[4] = {0, 0, 0, 0};
    tex_sample(c, t, f, u, v, d, s);
    for (int k = 0; k < 4; ++k) {
        if (k >= n) break;
        const int sw = f.swizzle[k];
        out[k] = (sw >= 0 ? s[sw] : 0.0f) * f.scale[k];
    }
}

// field_eval with zero uv partials (the normal map, the opacity test and light samples read fields
// without differentials): every sampler then reduces to its level-0 filter -- nearest, bilinear
// (linear: log2(1e-8) puts the level below 0; EWA: a zero minor axis falls back to bilinear) or
// bicubic -- so this gives the same values with far less code on the traversal and light paths.
KS_HD void field_eval0(const TexContext &c, const DevField &f, float u, float v, int n, float out[4]) {
    if (f.tex < 0) {
        for (int k = 0; k < 4; ++k)
            if (k < n) out[k] = f.value[k];
        return;
    }
    if (f.flip_v) v = 1.0f - v;
    u = u * f.uv_scale[0] + f.uv_offset[0];
    v = v * f.uv_scale[1] + f.uv_offset[1];
    const DevTexture &t = c.tex[f.tex];
    float s[4] = {0, 0, 0, 0};
    if (f.sampler == kSampNearest) {
        const TexLevel &l = t.lv[0];
        const int x = tex_wrap((int)floorf(u * (float)l.w - 0.5f), l.w, f.wrap_u);
        const int y = tex_wrap((int)floorf(v * (float)l.h - 0.5f), l.h, f.wrap_v);
        tex_fetch(c, t, x, y, 0, s);
    } else if (f.sampler == kSampCubic) {
        tex_bicubic(c, t, f, 0, u, v, s);
    } else {
        tex_bilinear(c, t, 0, u, v, f.wrap_u, f.wrap_v, s);
    }
    for (int k = 0; k < 4; ++k) {
        if (k >= n) break;
        const int sw = f.swizzle[k];
        out[k] = (sw >= 0 ? s[sw] : 0.0f) * f.scale[k];
    }
}

} // namespace ksd

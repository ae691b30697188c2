// Device-side building blocks of the gfx950 small_pt kernels: vector math, the z-Sobol sampler,
// the principled BSDF, light sampling and hit reconstruction. Each function cites the reference
// file:line whose semantics it implements.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define KS_DEV __device__ __forceinline__

namespace ksd {

constexpr float kPi = 3.14159265359f;    // maths.h:58
constexpr float kTwoPi = 6.28318530718f; // maths.h:59
constexpr float kHalfPi = kPi / 2.0f;
constexpr float kInvPi = 1.0f / kPi;
constexpr float kBeforeOne = 0x1.fffffep-1f; // fp32_before_one, maths.h:65
constexpr float kInf = __builtin_inff();

struct f3 {
    float x, y, z;
};
KS_DEV f3 mk(float x, float y, float z) { return f3{x, y, z}; }
KS_DEV f3 operator+(f3 a, f3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
KS_DEV f3 operator-(f3 a, f3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
KS_DEV f3 operator-(f3 a) { return {-a.x, -a.y, -a.z}; }
KS_DEV f3 operator*(f3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
KS_DEV f3 operator*(float s, f3 a) { return {a.x * s, a.y * s, a.z * s}; }
KS_DEV f3 operator*(f3 a, f3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
KS_DEV f3 operator/(f3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
KS_DEV float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
KS_DEV f3 cross(f3 a, f3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
KS_DEV float sqnorm(f3 a) { return dot(a, a); }
KS_DEV f3 normalized(f3 a) {
    float n2 = sqnorm(a);
    return n2 > 0.0f ? a / sqrtf(n2) : a;
}
KS_DEV float maxc(f3 a) { return fmaxf(a.x, fmaxf(a.y, a.z)); }
KS_DEV bool is_zero(f3 a) { return a.x == 0.0f && a.y == 0.0f && a.z == 0.0f; }
KS_DEV bool all_finite(f3 a) { return isfinite(a.x) && isfinite(a.y) && isfinite(a.z); }
KS_DEV f3 vmin(f3 a, float c) { return {fminf(a.x, c), fminf(a.y, c), fminf(a.z, c)}; }
KS_DEV float sqr(float x) { return x * x; }
KS_DEV float clampf(float x, float a, float b) { return fminf(fmaxf(x, a), b); }
KS_DEV float lerpf(float a, float b, float t) { return a + t * (b - a); }
KS_DEV float luminance(f3 c) { return 0.212671f * c.x + 0.715160f * c.y + 0.072169f * c.z; } // maths.h:675
KS_DEV float power_heur(float pf, float pg) { // nee.h:18-24
    float a = pf * pf, b = pg * pg;
    return a / (a + b);
}
KS_DEV f3 ld3(const float4 v) { return {v.x, v.y, v.z}; }

// ---------------------------------------------------------------------------------------------------
// z-Sobol sampler: OwenScrambledSobol02Morton + PTRenderSampler (sobol.h:131-248), hash.h:140-214.
// Integer-exact; the only float step is the final uint32 -> [0,1) map.
// ---------------------------------------------------------------------------------------------------
KS_DEV uint64_t mix_bits(uint64_t v) { // hash.h:206-214
    v ^= (v >> 31);
    v *= 0x7fb5d329728ea185ull;
    v ^= (v >> 27);
    v *= 0x81dadef4bc2dd44dull;
    v ^= (v >> 33);
    return v;
}

// murmur_hash_64A over the 8 bytes (int32 dimension, int32 seed) with seed 0 (hash.h:140-185,243-252)
KS_DEV uint64_t hash_dim_seed(int32_t dim, int32_t seed) {
    const uint64_t m = 0xc6a4a7935bd1e995ull;
    uint64_t h = 8ull * m;
    uint64_t k = (uint64_t)(uint32_t)dim | ((uint64_t)(uint32_t)seed << 32);
    k *= m;
    k ^= k >> 47;
    k *= m;
    h ^= k;
    h *= m;
    h ^= h >> 47;
    h *= m;
    h ^= h >> 47;
    return h;
}

KS_DEV uint32_t owen_scramble(uint32_t v, uint32_t seed) { // FastOwenScrambler, sobol.h:17-33
    v = __builtin_bitreverse32(v);
    v ^= v * 0x3d20adeau;
    v += seed;
    v *= (seed >> 16) | 1u;
    v ^= v * 0x05526c56u;
    v ^= v * 0x53a22864u;
    return __builtin_bitreverse32(v);
}

// Sobol generator matrices of dimensions 0 and 1 applied to index a (sobol.h:35-49). Dimension 0 is
// the van der Corput matrix (bit reversal of the low 32 index bits; words 32..51 of its column are
// zero). Dimension 1 has v_i = v_{i-1} ^ (v_{i-1} >> 1) with period 32 in the 52-word column, so index
// bits 32..51 act like bits 0..19 and the product is linear in lo ^ hi. Its column i is (I + S)^i of the
// top bit (S: shift right by one), whose bit 31 - k is C(i, k) mod 2: by Lucas' theorem, odd exactly when
// the bits of k are a subset of those of i. Output bit 31 - k is therefore the XOR of the index bits at
// the supersets i of k: five butterfly steps over the word, then a bit reversal.
KS_DEV uint32_t sobol_dim0(uint64_t a) { return __builtin_bitreverse32((uint32_t)a); }
KS_DEV uint32_t sobol_dim1(uint64_t a) {
    uint32_t v = (uint32_t)a ^ ((uint32_t)(a >> 32) & 0xfffffu);
    v ^= (v >> 1) & 0x55555555u;
    v ^= (v >> 2) & 0x33333333u;
    v ^= (v >> 4) & 0x0f0f0f0fu;
    v ^= (v >> 8) & 0x00ff00ffu;
    v ^= (v >> 16) & 0x0000ffffu;
    return __builtin_bitreverse32(v);
}
KS_DEV float u32_to_unit(uint32_t v) { return fminf((float)v * 0x1p-32f, kBeforeOne); }

KS_DEV uint32_t part_1_by_1(uint32_t x) {
    x &= 0x0000ffff;
    x = (x ^ (x << 8)) & 0x00ff00ff;
    x = (x ^ (x << 4)) & 0x0f0f0f0f;
    x = (x ^ (x << 2)) & 0x33333333;
    x = (x ^ (x << 1)) & 0x55555555;
    return x;
}

// The 24 base-4 digit permutations of sobol.h:171-197, one byte each (two bits per digit, in the
// reference's order), packed four to a word and selected with compares instead of a memory access.
KS_DEV uint32_t perm4_byte(uint32_t p) {
    const uint32_t W0 = 0x78D8B4E4u, W1 = 0xB1E19C6Cu, W2 = 0x8D2D39C9u, W3 = 0x72D236C6u, W4 = 0x87271E4Eu,
                   W5 = 0x93634B1Bu;
    uint32_t w = p < 8 ? (p < 4 ? W0 : W1) : (p < 16 ? (p < 12 ? W2 : W3) : (p < 20 ? W4 : W5));
    return (w >> ((p & 3u) * 8u)) & 0xffu;
}
// (mix_bits(x) >> 24) % 24 for a 32-bit x: the 40-bit quotient split at bit 32, with 2^32 = 16 mod 24.
KS_DEV uint32_t perm_index(uint32_t x) {
    uint64_t t = mix_bits((uint64_t)x) >> 24;
    return ((uint32_t)(t >> 32) * 16u + (uint32_t)t % 24u) % 24u;
}

// Permutation bytes by (dimension, digit prefix): entry dim * kPermTabPrefixes + h is
// perm4_byte(perm_index(h ^ 0x55555555 * dim)). A render's morton index is below 4^n_digits (2^(2 n_digits - 1)
// for an odd log2 spp), so the prefix (the digits above) of each of the top kPermTabLevels digits has at
// most 20 bits and is looked up; the lower digits' prefixes, and every digit of dimensions from
// kPermTabDims on (paths of more than about 17 bounces), are hashed. The table is a constant of 128 MiB,
// built once per device.
constexpr uint32_t kPermTabDims = 128;
constexpr uint32_t kPermTabPrefixes = 1u << 20;
constexpr int kPermTabLevels = 11;

struct Sobol {
    uint32_t morton; // the reference's morton index is built from 32-bit parts (sobol.h:142-146)
    int32_t dim;
    int32_t seed;
    int32_t log2_spp;
    int32_t n_digits;
    const uint8_t *tab; // the permutation table above, or nullptr: every digit is hashed

    // PTRenderSampler ctor: start_pixel with dimension 0 (sobol.h:142-146, 237-244)
    KS_DEV void start(uint32_t px, uint32_t py, uint32_t sample, int32_t dim0, int32_t seed_, int32_t log2spp,
                      int32_t ndigits, const uint8_t *tab_ = nullptr) {
        seed = seed_;
        log2_spp = log2spp;
        n_digits = ndigits;
        dim = dim0;
        tab = tab_;
        morton = (((part_1_by_1(py) << 1) + part_1_by_1(px)) << log2spp) | sample; // 32-bit as reference
    }

    KS_DEV uint64_t sample_index() const { // sobol.h:168-222
        uint64_t si = 0;
        const bool pow2 = log2_spp & 1;
        const int last = pow2 ? 1 : 0;
        const uint32_t dmask = 0x55555555u * (uint32_t)dim;
        int i = n_digits - 1;
        if (tab != nullptr && (uint32_t)dim < kPermTabDims) {
            // the top digits: their table loads are issued together, then applied
            const uint8_t *t = tab + (size_t)dim * kPermTabPrefixes;
            uint32_t pb[kPermTabLevels];
#pragma unroll
            for (int k = 0; k < kPermTabLevels; ++k) {
                const int lv = n_digits - 1 - k;
                const int shift = 2 * lv - (pow2 ? 1 : 0);
                const uint32_t higher = lv >= last && shift + 2 < 32 ? morton >> (shift + 2) : 0u;
                pb[k] = higher < kPermTabPrefixes ? t[higher] : perm4_byte(perm_index(higher ^ dmask));
            }
#pragma unroll
            for (int k = 0; k < kPermTabLevels; ++k) {
                const int lv = n_digits - 1 - k;
                if (lv < last) break;
                const int shift = 2 * lv - (pow2 ? 1 : 0);
                const uint32_t digit = shift < 32 ? (morton >> shift) & 3u : 0u;
                si |= (uint64_t)((pb[k] >> (2 * digit)) & 3u) << shift;
            }
            i = n_digits - 1 - kPermTabLevels;
        }
        for (; i >= last; --i) {
            int shift = 2 * i - (pow2 ? 1 : 0);
            // digits above bit 31 of the 32-bit morton index are zero
            uint32_t digit = shift < 32 ? (morton >> shift) & 3u : 0u;
            uint32_t higher = shift + 2 < 32 ? morton >> (shift + 2) : 0u;
            uint32_t p = perm_index(higher ^ dmask);
            digit = (perm4_byte(p) >> (2 * digit)) & 3u;
            si |= (uint64_t)digit << shift;
        }
        if (pow2) {
            uint32_t digit = morton & 1u;
            si |= (uint64_t)(digit ^ (uint32_t)(mix_bits((uint64_t)((morton >> 1) ^ dmask)) & 1u));
        }
        return si;
    }

    KS_DEV uint32_t next_u32() { // sobol.h:148-155
        uint64_t si = sample_index();
        ++dim;
        uint32_t h = (uint32_t)hash_dim_seed(dim, seed);
        return owen_scramble(sobol_dim0(si), h);
    }
    KS_DEV void next2d_u32(uint32_t &a, uint32_t &b) { // sobol.h:157-166
        uint64_t si = sample_index();
        dim += 2;
        uint64_t bits = hash_dim_seed(dim, seed);
        a = owen_scramble(sobol_dim0(si), (uint32_t)bits);
        b = owen_scramble(sobol_dim1(si), (uint32_t)(bits >> 32));
    }
    KS_DEV float next() { return u32_to_unit(next_u32()); }
    KS_DEV void next2d(float &u0, float &u1) {
        uint32_t a, b;
        next2d_u32(a, b);
        u0 = u32_to_unit(a);
        u1 = u32_to_unit(b);
    }
};

// ---------------------------------------------------------------------------------------------------
// Scene records on the device
// ---------------------------------------------------------------------------------------------------
struct Material { // PrincipledBSDF closure after eval_closure clamps (principled_bsdf.cpp:246-275)
    float basecolor[3];
    float roughness;
    float ax, ay, metallic, ior;
    float specular_r0_mul, specular_trans, diffuse_trans, diffuse_trans_fwd;
    float emission[3];
    int32_t has_emission;
};
static_assert(sizeof(Material) == 64, "Material record is 64 bytes");

struct Light {
    int32_t type; // 0 ambient sky, 1 directional, 2 point, 3 mesh triangle, 4 environment-map sky
    uint32_t prim;
    float area;
    float pad;
    float color[4];
    float vec[4]; // direction / position; mesh triangle: vertex 0
    float v1[4], v2[4]; // mesh triangle: vertices 1 and 2 (kept here so NEE needs no index fetch)
};
static_assert(sizeof(Light) == 80, "Light record is 80 bytes");

struct Closure {
    f3 basecolor;
    float roughness, ax, ay, metallic, ior, specular_r0_mul, specular_trans, diffuse_trans, diffuse_trans_fwd;
};
KS_DEV Closure closure_of(const Material &m) {
    Closure c;
    c.basecolor = mk(m.basecolor[0], m.basecolor[1], m.basecolor[2]);
    c.roughness = m.roughness;
    c.ax = m.ax;
    c.ay = m.ay;
    c.metallic = m.metallic;
    c.ior = m.ior;
    c.specular_r0_mul = m.specular_r0_mul;
    c.specular_trans = m.specular_trans;
    c.diffuse_trans = m.diffuse_trans;
    c.diffuse_trans_fwd = m.diffuse_trans_fwd;
    return c;
}

// ---------------------------------------------------------------------------------------------------
// Sampling warps (rng.h:243-373, bsdf.cpp:62-91, maths.h:814-821)
// ---------------------------------------------------------------------------------------------------
KS_DEV f3 sample_cosine_hemisphere(float u0, float u1) {
    float a = 2.0f * u0 - 1.0f, b = 2.0f * u1 - 1.0f;
    float dx = 0.0f, dy = 0.0f;
    if (!(a == 0.0f && b == 0.0f)) {
        float r, phi;
        if (a * a > b * b) {
            r = a;
            phi = 0.25f * kPi * (b / a);
        } else {
            r = b;
            phi = kHalfPi - 0.25f * kPi * (a / b);
        }
        float s, c;
        sincosf(phi, &s, &c);
        dx = r * c;
        dy = r * s;
    }
    return mk(dx, dy, sqrtf(fmaxf(0.0f, 1.0f - (dx * dx + dy * dy))));
}

KS_DEV void orthonormal_basis(f3 n, f3 &X, f3 &Y) {
    float sign = copysignf(1.0f, n.z);
    const float a = -1.0f / (sign + n.z);
    const float b = n.x * n.y * a;
    X = mk(1.0f + sign * n.x * n.x * a, sign * b, -sign * n.x);
    Y = mk(b, sign + n.y * n.y * a, -n.y);
}

KS_DEV float safe_pow(float a, float b) { return (a == 0.0f && b == 0.0f) ? 0.0f : powf(a, b); }
KS_DEV float hg_phase(float cos_theta, float g) { // bsdf.h:45-48
    return ((1.0f - g * g) / safe_pow(1.0f + g * g - 2.0f * g * cos_theta, 1.5f)) * (kInvPi * 0.25f);
}
KS_DEV f3 sample_hg(f3 D, float g, float ru, float rv, float *pdf) {
    float cos_theta;
    if (fabsf(g) < 1e-3f) {
        cos_theta = 1.0f - 2.0f * ru;
        *pdf = kInvPi * 0.25f;
    } else {
        float k = (1.0f - g * g) / (1.0f - g + 2.0f * g * ru);
        cos_theta = (1.0f + g * g - k * k) / (2.0f * g);
        *pdf = hg_phase(cos_theta, g);
    }
    float sin_theta = sqrtf(fmaxf(0.0f, 1.0f - cos_theta * cos_theta));
    float s, c;
    sincosf(kTwoPi * rv, &s, &c);
    f3 T, B;
    orthonormal_basis(D, T, B);
    return T * (sin_theta * c) + B * (sin_theta * s) + D * cos_theta;
}

// ---------------------------------------------------------------------------------------------------
// GGX (microfacet.h:20-153), Fresnel (fresnel.h:9-30), PrincipledBSDF (principled_bsdf.cpp:297-705)
// ---------------------------------------------------------------------------------------------------
struct GGX {
    float ax, ay;
    KS_DEV GGX(float x, float y) : ax(clampf(x, 1e-3f, 1.0f)), ay(clampf(y, 1e-3f, 1.0f)) {}
    KS_DEV float D(f3 wm) const {
        if (wm.z <= 0.0f) return 0.0f;
        if (ax == ay) {
            float a2 = ax * ax;
            float t = 1.0f + (a2 - 1.0f) * wm.z * wm.z;
            return a2 / (kPi * t * t);
        }
        float hx = wm.x / ax, hy = wm.y / ay;
        float t = sqr(hx) + sqr(hy) + sqr(wm.z);
        return 1.0f / (kPi * ax * ay * sqr(t));
    }
    KS_DEV float lambda(f3 w) const {
        if (w.z >= 1.0f || w.z <= -1.0f) return 0.0f;
        float alpha;
        if (ax == ay) {
            alpha = ax;
        } else {
            float inv_s2 = 1.0f / (1.0f - w.z * w.z);
            alpha = sqrtf(w.x * w.x * inv_s2 * ax * ax + w.y * w.y * inv_s2 * ay * ay);
        }
        float t = (1.0f - w.z * w.z) * (alpha * alpha) / (w.z * w.z);
        return 0.5f * (-1.0f + sqrtf(1.0f + t));
    }
    KS_DEV float G1(f3 w) const { return 1.0f / (1.0f + lambda(w)); }
    KS_DEV float G2(f3 wo, f3 wi) const {
        float lo = lambda(wo), li = lambda(wi);
        if (wo.z * wi.z >= 0.0f) return 1.0f / (1.0f + lo + li);
        // B(1+lo, 1+li) (std::beta in the reference) through log-gamma
        float a = 1.0f + lo, b = 1.0f + li;
        return expf(lgammaf(a) + lgammaf(b) - lgammaf(a + b));
    }
    KS_DEV f3 sample(f3 wo, float u0, float u1) const { // VNDF (microfacet.h:98-118)
        f3 Vh = normalized(mk(ax * wo.x, ay * wo.y, wo.z));
        float lensq = Vh.x * Vh.x + Vh.y * Vh.y;
        f3 T1 = lensq > 0.0f ? mk(-Vh.y, Vh.x, 0.0f) * (1.0f / sqrtf(lensq)) : mk(1.0f, 0.0f, 0.0f);
        f3 T2 = cross(Vh, T1);
        float r = sqrtf(u0);
        float s, c;
        sincosf(kTwoPi * u1, &s, &c);
        float t1 = r * c, t2 = r * s;
        float sh = 0.5f * (1.0f + Vh.z);
        t2 = (1.0f - sh) * sqrtf(1.0f - t1 * t1) + sh * t2;
        f3 Nh = T1 * t1 + T2 * t2 + Vh * sqrtf(fmaxf(0.0f, 1.0f - t1 * t1 - t2 * t2));
        return normalized(mk(ax * Nh.x, ay * Nh.y, fmaxf(0.0f, Nh.z)));
    }
};

KS_DEV float fresnel_dielectric(float ci, float eta) {
    float st2 = eta * eta * (1.0f - ci * ci);
    if (st2 > 1.0f) return 1.0f;
    float ct = sqrtf(fmaxf(1.0f - st2, 0.0f));
    float rs = (eta * ct - ci) / (eta * ct + ci);
    float rp = (eta * ci - ct) / (eta * ci + ct);
    return 0.5f * (rs * rs + rp * rp);
}
KS_DEV float fresnel_schlick(float c) {
    float s = clampf(1.0f - c, 0.0f, 1.0f);
    float s2 = s * s;
    return s2 * s2 * s;
}
KS_DEV f3 reflect(f3 w, f3 n) { return n * (2.0f * dot(n, w)) - w; }
KS_DEV bool refract(f3 wi, f3 n, float eta, f3 &wt) {
    float NdotI = dot(n, wi);
    float k = 1.0f - eta * eta * (1.0f - sqr(NdotI));
    if (k < 0.0f) return false;
    wt = -wi * eta + n * (eta * NdotI - sqrtf(k));
    return true;
}
KS_DEV float sgnf(float x) { return (float)((0.0f < x) - (x < 0.0f)); }

KS_DEV f3 eval_diffuse(f3 wo, f3 wi, const Closure &c) {
    float lw = (1.0f - c.metallic) * (1.0f - c.specular_trans) * (1.0f - c.diffuse_trans);
    if (lw == 0.0f) return mk(0, 0, 0);
    f3 wh = normalized(wo + wi);
    float Fd90 = 0.5f + 2.0f * c.roughness * wh.z;
    float retro = lerpf(1.0f, Fd90, fresnel_schlick(wi.z)) * lerpf(1.0f, Fd90, fresnel_schlick(wo.z));
    return c.basecolor * (lw * kInvPi * retro * wi.z);
}
KS_DEV f3 eval_diffuse_transmission(f3 wo, f3 wi, const Closure &c) {
    float lw = (1.0f - c.metallic) * (1.0f - c.specular_trans) * c.diffuse_trans;
    if (lw == 0.0f) return mk(0, 0, 0);
    return c.basecolor * (lw * kInvPi * (-wi.z) * hg_phase(dot(wo, -wi), c.diffuse_trans_fwd));
}
KS_DEV f3 eval_metallic_specular(f3 wo, f3 wi, const Closure &c) {
    if (c.metallic == 0.0f) return mk(0, 0, 0);
    GGX g(c.ax, c.ay);
    f3 wh = normalized(wo + wi);
    float D = g.D(wh), G = g.G2(wo, wi);
    float fs = fresnel_schlick(dot(wo, wh));
    f3 Fr = mk(lerpf(c.basecolor.x, 1.0f, fs), lerpf(c.basecolor.y, 1.0f, fs), lerpf(c.basecolor.z, 1.0f, fs));
    return Fr * (c.metallic * D * G / (4.0f * wo.z));
}
KS_DEV float dielectric_specular_adjust(float wo_dot_wh, const Closure &c) { // principled_bsdf.cpp:510-516
    float R0 = sqr((1.0f - c.ior) / (1.0f + c.ior));
    float schlick = fresnel_schlick(wo_dot_wh);
    float den = lerpf(R0, 1.0f, schlick);
    // ior == 1 with |wo.wh| == 1 is 0/0 in the reference; its continuous limit is 1
    return den > 0.0f ? lerpf(c.specular_r0_mul * R0, 1.0f, schlick) / den : 1.0f;
}
KS_DEV f3 eval_dielectric_specular(f3 wo, f3 wi, const Closure &c) {
    float lw = 1.0f - c.metallic;
    if (lw == 0.0f) return mk(0, 0, 0);
    bool refl = wo.z * wi.z >= 0.0f;
    if (!refl && c.specular_trans == 0.0f) return mk(0, 0, 0); // lw * specular_trans = 0
    float eta = wo.z >= 0.0f ? c.ior : (1.0f / c.ior);
    f3 wh;
    if (refl) {
        wh = normalized(wo + wi);
    } else {
        wh = normalized(wo + wi * eta);
        lw *= c.specular_trans;
    }
    if (wh.z < 0.0f) wh = -wh;
    GGX g(c.ax, c.ay);
    float D = g.D(wh), G = g.G2(wo, wi);
    float awh = fabsf(dot(wo, wh));
    float Fr = fresnel_dielectric(awh, 1.0f / eta);
    float f;
    if (refl) {
        f = D * G * (Fr * dielectric_specular_adjust(awh, c)) / (4.0f * fabsf(wo.z));
    } else {
        float denom = sqr(dot(wo, wh) + eta * dot(wi, wh));
        f = fabsf(D * G * (1.0f - Fr) * dot(wo, wh) * dot(wi, wh) / (denom * wo.z));
    }
    f *= lw;
    return mk(f, f, f);
}

// Lobes of the principled BSDF a shading kernel is compiled for. Material-sorted shading instantiates
// the BSDF for the lobes its material class can have non-zero weight on; every other lobe then
// contributes exactly nothing, as its zero weight makes it contribute in the full BSDF.
constexpr unsigned kLobeDiffuse = 1, kLobeDiffuseTrans = 2, kLobeMetal = 4, kLobeDielRefl = 8, kLobeDielTrans = 16;
constexpr unsigned kLobesAll = 31;
// opaque dielectrics (metallic 0, specular_trans 0, diffuse_trans 0) and pure metals (metallic 1)
constexpr unsigned kLobesOpaqueDielectric = kLobeDiffuse | kLobeDielRefl;
constexpr unsigned kLobesMetal = kLobeMetal;

template <unsigned L = kLobesAll>
KS_DEV void lobe_weights(f3 wo, const Closure &c, float w[4]) { // principled_bsdf.cpp:518-540
    float lum = luminance(c.basecolor);
    float wd = wo.z > 0.0f ? (1.0f - c.metallic) * (1.0f - c.specular_trans) * lum * kInvPi : 0.0f;
    w[0] = wd * (1.0f - c.diffuse_trans);
    w[1] = wd * c.diffuse_trans;
    w[2] = wo.z > 0.0f ? c.metallic * lerpf(lum, 1.0f, fresnel_schlick(wo.z)) : 0.0f;
    float eta = wo.z >= 0.0f ? c.ior : (1.0f / c.ior);
    if constexpr ((L & (kLobeDielRefl | kLobeDielTrans)) != 0)
        w[3] = (1.0f - c.metallic) * fresnel_dielectric(fabsf(wo.z), 1.0f / eta);
    else
        w[3] = 0.0f; // metallic 1: (1 - metallic) times a Fresnel term in [0, 1]
    float sum = w[0] + w[1] + w[2] + w[3];
    if (sum == 0.0f) {
        w[0] = w[1] = w[2] = w[3] = 0.0f;
        return;
    }
    w[0] /= sum;
    w[1] /= sum;
    w[2] /= sum;
    w[3] /= sum;
}

KS_DEV float pdf_diffuse_transmission(f3 wo, f3 wi, const Closure &c) {
    return c.diffuse_trans_fwd > 0.05f ? hg_phase(dot(wo, -wi), c.diffuse_trans_fwd) : -wi.z * kInvPi;
}
KS_DEV float pdf_metallic_specular(f3 wo, f3 wi, const Closure &c) {
    GGX g(c.ax, c.ay);
    f3 wh = normalized(wo + wi);
    return g.D(wh) * g.G1(wo) / (4.0f * fabsf(wo.z));
}
KS_DEV float pdf_dielectric_specular(f3 wo, f3 wi, const Closure &c) {
    bool refl = wo.z * wi.z >= 0.0f;
    float eta = wo.z >= 0.0f ? c.ior : (1.0f / c.ior);
    f3 wh = refl ? normalized(wo + wi) : normalized(wo + wi * eta);
    if (wh.z < 0.0f) wh = -wh;
    GGX g(c.ax, c.ay);
    float D = g.D(wh), G1 = g.G1(wo);
    float awh = fabsf(dot(wo, wh));
    float Fr = fresnel_dielectric(awh, 1.0f / eta);
    float Fr_refl = Fr * dielectric_specular_adjust(awh, c);
    float Fr_refr = (1.0f - Fr) * c.specular_trans;
    if (Fr_refr == 0.0f) Fr_refl = 1.0f;
    float pdf;
    if (refl) {
        pdf = D * G1 / (4.0f * fabsf(wo.z)) * (Fr_refl / (Fr_refl + Fr_refr));
    } else {
        float denom = sqr(dot(wo, wh) + eta * dot(wi, wh));
        float jac = eta * eta * fabsf(dot(wi, wh)) / denom;
        pdf = D * G1 * fabsf(dot(wo, wh)) / fabsf(wo.z) * jac * (Fr_refr / (Fr_refl + Fr_refr));
    }
    return pdf;
}

// A lobe's pdf is only evaluated when its weight is non-zero: its product with a zero weight adds exactly
// nothing to the mixture pdf. (At ior 1 the dielectric weight is a Fresnel term that float rounding
// leaves far from zero at grazing angles, so it is evaluated there as in the reference.)
KS_DEV bool dielectric_pdf_needed(f3 wo, f3 wi, const Closure &c, float w3) {
    (void)wo;
    (void)wi;
    (void)c;
    return w3 != 0.0f;
}

template <unsigned L = kLobesAll> KS_DEV f3 bsdf_eval(f3 wo, f3 wi, const Closure &c) {
    if (wo.z == 0.0f || wi.z == 0.0f) return mk(0, 0, 0);
    f3 f = mk(0, 0, 0);
    if (wo.z > 0.0f && wi.z > 0.0f) {
        if constexpr ((L & kLobeDiffuse) != 0) f = f + eval_diffuse(wo, wi, c);
        if constexpr ((L & kLobeMetal) != 0) f = f + eval_metallic_specular(wo, wi, c);
    }
    if constexpr ((L & kLobeDiffuseTrans) != 0)
        if (wo.z > 0.0f && wi.z < 0.0f) f = f + eval_diffuse_transmission(wo, wi, c);
    if constexpr ((L & kLobeDielTrans) != 0)
        return f + eval_dielectric_specular(wo, wi, c);
    else if constexpr ((L & kLobeDielRefl) != 0)
        return wo.z * wi.z >= 0.0f ? f + eval_dielectric_specular(wo, wi, c) : f;
    else
        return f;
}
template <unsigned L = kLobesAll> KS_DEV float bsdf_pdf(f3 wo, f3 wi, const Closure &c) {
    if (wo.z == 0.0f || wi.z == 0.0f) return 0.0f;
    float w[4];
    lobe_weights<L>(wo, c, w);
    if (w[0] == 0.0f && w[1] == 0.0f && w[2] == 0.0f && w[3] == 0.0f) return 0.0f;
    float p0 = 0.0f, p1 = 0.0f, p2 = 0.0f, p3 = 0.0f;
    if (wo.z > 0.0f && wi.z > 0.0f) {
        p0 = wi.z * kInvPi;
        if constexpr ((L & kLobeMetal) != 0)
            if (w[2] != 0.0f) p2 = pdf_metallic_specular(wo, wi, c);
    }
    if constexpr ((L & kLobeDiffuseTrans) != 0)
        if (wo.z > 0.0f && wi.z < 0.0f && w[1] != 0.0f) p1 = pdf_diffuse_transmission(wo, wi, c);
    if constexpr ((L & kLobeDielTrans) != 0) {
        if (dielectric_pdf_needed(wo, wi, c, w[3])) p3 = pdf_dielectric_specular(wo, wi, c);
    } else if constexpr ((L & kLobeDielRefl) != 0) {
        // without transmission the dielectric lobe has no density below the surface
        if (wo.z * wi.z >= 0.0f && dielectric_pdf_needed(wo, wi, c, w[3])) p3 = pdf_dielectric_specular(wo, wi, c);
    }
    return p0 * w[0] + p1 * w[1] + p2 * w[2] + p3 * w[3];
}

KS_DEV int sample_small_distrib4(const float w[4], float u, float *u_remap) { // rng.h:399-432
    float sum = 0.0f;
    int last_pos = 0;
    for (int i = 0; i < 4; ++i) {
        sum += w[i];
        if (w[i] > 0.0f) last_pos = i;
    }
    float inv = 1.0f / sum, cdf = 0.0f;
    for (int i = 0; i < 4; ++i) {
        float nx = cdf + w[i] * inv;
        if (u < nx) {
            *u_remap = (u - cdf) / (nx - cdf);
            return i;
        }
        cdf = nx;
    }
    return last_pos;
}

// Returns f*cos/pdf, sets wi and pdf (principled_bsdf.cpp:326-393). L: the lobes that can have weight.
template <unsigned L = kLobesAll>
KS_DEV f3 bsdf_sample(f3 wo, f3 &wi, const Closure &c, float u_lobe, float u0, float u1, float &pdf) {
    pdf = 0.0f;
    const f3 zero = mk(0, 0, 0);
    if (wo.z == 0.0f) return zero;
    float w[4];
    lobe_weights<L>(wo, c, w);
    if (w[0] == 0.0f && w[1] == 0.0f && w[2] == 0.0f && w[3] == 0.0f) return zero;
    float u_remap = 0.0f;
    int lobe = sample_small_distrib4(w, u_lobe, &u_remap);
    float p[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    GGX g(c.ax, c.ay);
    constexpr bool kDiel = (L & (kLobeDielRefl | kLobeDielTrans)) != 0;
    // the dielectric lobe's density at wi: below the surface only with transmission
    auto diel_pdf = [&](f3 wi_) -> float {
        if constexpr ((L & kLobeDielTrans) == 0)
            if (wo.z * wi_.z < 0.0f) return 0.0f;
        return dielectric_pdf_needed(wo, wi_, c, w[3]) ? pdf_dielectric_specular(wo, wi_, c) : 0.0f;
    };
    if ((L & kLobeDiffuse) != 0 && lobe == 0) {
        wi = sample_cosine_hemisphere(u0, u1);
        p[0] = wi.z * kInvPi;
        if (is_zero(wi) || p[0] == 0.0f) return zero;
        if constexpr ((L & kLobeMetal) != 0)
            if (wo.z > 0.0f && wi.z > 0.0f && w[2] != 0.0f) p[2] = pdf_metallic_specular(wo, wi, c);
        if constexpr (kDiel) p[3] = diel_pdf(wi);
    } else if ((L & kLobeDiffuseTrans) != 0 && lobe == 1) {
        if (c.diffuse_trans_fwd > 0.05f) {
            wi = sample_hg(-wo, c.diffuse_trans_fwd, u0, u1, &p[1]);
        } else {
            wi = sample_cosine_hemisphere(u0, u1);
            p[1] = wi.z * kInvPi;
            wi.z = -wi.z;
        }
        if (is_zero(wi) || p[1] == 0.0f) return zero;
        if constexpr (kDiel) p[3] = diel_pdf(wi);
    } else if ((L & kLobeMetal) != 0 && lobe == 2) {
        f3 wh = g.sample(wo * sgnf(wo.z), u0, u1);
        wi = reflect(wo, wh);
        if (wo.z * wi.z < 0.0f) return zero;
        p[2] = g.D(wh) * g.G1(wo) / (4.0f * fabsf(wo.z));
        if (is_zero(wi) || p[2] == 0.0f) return zero;
        if (wo.z > 0.0f && wi.z > 0.0f) p[0] = wi.z * kInvPi;
        if constexpr (kDiel) p[3] = diel_pdf(wi);
    } else if (kDiel && lobe == 3) {
        f3 wh = g.sample(wo * sgnf(wo.z), u0, u1);
        float D = g.D(wh), G1 = g.G1(wo);
        float eta = wo.z >= 0.0f ? c.ior : (1.0f / c.ior);
        float awh = fabsf(dot(wo, wh));
        float Fr = fresnel_dielectric(awh, 1.0f / eta);
        float Fr_refl = Fr * dielectric_specular_adjust(awh, c);
        float Fr_refr = (1.0f - Fr) * c.specular_trans;
        if (Fr_refr == 0.0f) Fr_refl = 1.0f;
        if (u_remap < Fr_refl / (Fr_refl + Fr_refr)) {
            wi = reflect(wo, wh);
            if (wo.z * wi.z < 0.0f) return zero;
            p[3] = D * G1 / (4.0f * fabsf(wo.z)) * (Fr_refl / (Fr_refl + Fr_refr));
        } else if constexpr ((L & kLobeDielTrans) != 0) {
            if (!refract(wo, wh * sgnf(wo.z), 1.0f / eta, wi)) return zero;
            if (wo.z * wi.z > 0.0f) return zero;
            float denom = sqr(dot(wo, wh) + eta * dot(wi, wh));
            float jac = eta * eta * fabsf(dot(wi, wh)) / denom;
            p[3] = D * G1 * fabsf(dot(wo, wh)) / fabsf(wo.z) * jac * (Fr_refr / (Fr_refl + Fr_refr));
        } else {
            // specular_trans 0: Fr_refr = 0, so the refraction's pdf carries the factor 0 / (1 + 0)
            return zero;
        }
        if (is_zero(wi) || p[3] == 0.0f) return zero;
        if (wo.z > 0.0f && wi.z > 0.0f) {
            p[0] = wi.z * kInvPi;
            if constexpr ((L & kLobeMetal) != 0)
                if (w[2] != 0.0f) p[2] = pdf_metallic_specular(wo, wi, c);
        }
        if constexpr ((L & kLobeDiffuseTrans) != 0)
            if (wo.z > 0.0f && wi.z < 0.0f && w[1] != 0.0f) p[1] = pdf_diffuse_transmission(wo, wi, c);
    } else {
        return zero; // a lobe this instantiation has no code for (its weight is zero for this material)
    }
    pdf = p[0] * w[0] + p[1] * w[1] + p[2] * w[2] + p[3] * w[3];
    return bsdf_eval<L>(wo, wi, c) / pdf;
}

// ---------------------------------------------------------------------------------------------------
// Rays and hit records (ray.h:66-128, geometry.cpp:68-163)
// ---------------------------------------------------------------------------------------------------
KS_DEV f3 offset_ray(f3 p, f3 n) { // ray.h:66-79
    const float origin = 1.0f / 32.0f, float_scale = 1.0f / 65536.0f, int_scale = 256.0f;
    int ox = (int)(int_scale * n.x), oy = (int)(int_scale * n.y), oz = (int)(int_scale * n.z);
    float px = __int_as_float(__float_as_int(p.x) + ((p.x < 0) ? -ox : ox));
    float py = __int_as_float(__float_as_int(p.y) + ((p.y < 0) ? -oy : oy));
    float pz = __int_as_float(__float_as_int(p.z) + ((p.z < 0) ? -oz : oz));
    return mk(fabsf(p.x) < origin ? p.x + float_scale * n.x : px, fabsf(p.y) < origin ? p.y + float_scale * n.y : py,
              fabsf(p.z) < origin ? p.z + float_scale * n.z : pz);
}
KS_DEV f3 spawn_origin(f3 p, f3 dir, f3 ng) { return offset_ray(p, dot(dir, ng) > 0.0f ? ng : -ng); }

struct SurfHit {
    f3 p;
    f3 t, b, n;    // geometric frame
    f3 st, sb, sn; // shading frame
};

KS_DEV f3 sh_to_local(const SurfHit &it, f3 w) { // ray.h:108-117
    bool hw = dot(it.n, w) >= 0.0f;
    f3 l = mk(dot(it.st, w), dot(it.sb, w), dot(it.sn, w));
    if (hw != (l.z >= 0.0f)) l.z = -l.z;
    return l;
}
KS_DEV f3 sh_to_world(const SurfHit &it, f3 l) { // ray.h:119-128
    bool hl = l.z >= 0.0f;
    f3 w = it.st * l.x + it.sb * l.y + it.sn * l.z;
    if (hl != (dot(it.n, w) >= 0.0f)) w = w - it.n * (2.0f * dot(it.n, w));
    return w;
}

} // namespace ksd

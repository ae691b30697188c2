// CPU oracle for the keepsake small_pt path. Test infrastructure only: see ks_oracle.h for the
// rules on who may load it and for the pinning status of each function.
//
// Float evaluation follows the reference expressions term by term where that is cheap; the oracle is
// compiled without FMA contraction (-ffp-contract=off), like an x86-64 SSE2 build of the reference.
#include "ks_oracle.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <memory>
#include <thread>
#include <mutex>
#include <vector>

namespace {

constexpr float kPi = 3.14159265359f;       // maths.h:58
constexpr float kTwoPi = 6.28318530718f;    // maths.h:59
constexpr float kHalfPi = kPi / 2.0f;       // maths.h:60
constexpr float kInvPi = 1.0f / kPi;        // maths.h:63
constexpr float kInf = std::numeric_limits<float>::infinity();
const float kBeforeOne = std::nextafter(1.0f, -kInf); // maths.h:65

struct V3 {
    float x = 0, y = 0, z = 0;
    V3() = default;
    V3(float a, float b, float c) : x(a), y(b), z(c) {}
    float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
    float &operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
};
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator-(V3 a) { return {-a.x, -a.y, -a.z}; }
inline V3 operator*(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline V3 operator*(float s, V3 a) { return {a.x * s, a.y * s, a.z * s}; }
inline V3 operator*(V3 a, V3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
inline V3 operator/(V3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
inline float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline float sqnorm(V3 a) { return dot(a, a); }
inline V3 normalized(V3 a) {
    float n2 = sqnorm(a);
    if (n2 > 0.0f) return a / std::sqrt(n2);
    return a;
}
inline float maxc(V3 a) { return std::max(a.x, std::max(a.y, a.z)); }
inline bool is_zero(V3 a) { return a.x == 0.0f && a.y == 0.0f && a.z == 0.0f; }
inline bool all_finite(V3 a) { return std::isfinite(a.x) && std::isfinite(a.y) && std::isfinite(a.z); }
inline V3 vmin(V3 a, float c) { return {std::min(a.x, c), std::min(a.y, c), std::min(a.z, c)}; }
inline float sqr(float x) { return x * x; }
template <typename T> T clampv(T x, T a, T b) { return std::min(std::max(x, a), b); }
inline float lerpf(float a, float b, float t) { return a + t * (b - a); } // std::lerp for finite inputs

inline float luminance(V3 c) { return 0.212671f * c.x + 0.715160f * c.y + 0.072169f * c.z; } // maths.h:675

inline float power_heur(float pf, float pg) { // nee.h:18-24 (n = 2)
    float pf2 = pf * pf, pg2 = pg * pg;
    return pf2 / (pf2 + pg2);
}

// ----------------------------------------------------------------------------------------------------
// Hashing (hash.h:140-252)
// ----------------------------------------------------------------------------------------------------
uint64_t murmur64a(const unsigned char *key, size_t len, uint64_t seed) { // hash.h:140-185
    const uint64_t m = 0xc6a4a7935bd1e995ull;
    const int r = 47;
    uint64_t h = seed ^ (len * m);
    const unsigned char *end = key + 8 * (len / 8);
    while (key != end) {
        uint64_t k;
        std::memcpy(&k, key, 8);
        key += 8;
        k *= m;
        k ^= k >> r;
        k *= m;
        h ^= k;
        h *= m;
    }
    switch (len & 7) {
    case 7: h ^= uint64_t(key[6]) << 48; [[fallthrough]];
    case 6: h ^= uint64_t(key[5]) << 40; [[fallthrough]];
    case 5: h ^= uint64_t(key[4]) << 32; [[fallthrough]];
    case 4: h ^= uint64_t(key[3]) << 24; [[fallthrough]];
    case 3: h ^= uint64_t(key[2]) << 16; [[fallthrough]];
    case 2: h ^= uint64_t(key[1]) << 8; [[fallthrough]];
    case 1:
        h ^= uint64_t(key[0]);
        h *= m;
    }
    h ^= h >> r;
    h *= m;
    h ^= h >> r;
    return h;
}

uint64_t mix_bits(uint64_t v) { // hash.h:206-214
    v ^= (v >> 31);
    v *= 0x7fb5d329728ea185ull;
    v ^= (v >> 27);
    v *= 0x81dadef4bc2dd44dull;
    v ^= (v >> 33);
    return v;
}

// hash(int dimension, int seed): 8 bytes of two int32 (hash.h:243-252)
uint64_t hash_dim_seed(int dimension, int seed) {
    unsigned char buf[8];
    std::memcpy(buf, &dimension, 4);
    std::memcpy(buf + 4, &seed, 4);
    return murmur64a(buf, 8, 0);
}

// ----------------------------------------------------------------------------------------------------
// Sobol / Owen / Morton sampler (sobol.h:17-248)
// ----------------------------------------------------------------------------------------------------
// Generator matrices of dimensions 0 and 1 (first 104 words of sobol.cpp:154), from their closed forms:
// dim 0 is the van der Corput matrix, dim 1 satisfies v_i = v_{i-1} ^ (v_{i-1} >> 1); both 52-word
// columns repeat with period 32 in the table (dim 0 is zero-filled past word 31).
struct SobolMatrices {
    uint32_t m[104];
    SobolMatrices() {
        for (int i = 0; i < 52; ++i) m[i] = i < 32 ? (0x80000000u >> i) : 0u;
        uint32_t v = 0x80000000u;
        uint32_t d1[32];
        for (int i = 0; i < 32; ++i) {
            d1[i] = v;
            v = v ^ (v >> 1);
        }
        for (int i = 0; i < 52; ++i) m[52 + i] = d1[i % 32];
    }
};
const SobolMatrices g_sobol;

inline uint32_t reverse_bits_32(uint32_t bits) { // maths.h:267-275
    bits = (bits << 16u) | (bits >> 16u);
    bits = ((bits & 0x55555555u) << 1u) | ((bits & 0xAAAAAAAAu) >> 1u);
    bits = ((bits & 0x33333333u) << 2u) | ((bits & 0xCCCCCCCCu) >> 2u);
    bits = ((bits & 0x0F0F0F0Fu) << 4u) | ((bits & 0xF0F0F0F0u) >> 4u);
    bits = ((bits & 0x00FF00FFu) << 8u) | ((bits & 0xFF00FF00u) >> 8u);
    return bits;
}

inline uint32_t owen_scramble(uint32_t v, uint32_t seed) { // FastOwenScrambler sobol.h:17-33
    v = reverse_bits_32(v);
    v ^= v * 0x3d20adeau;
    v += seed;
    v *= (seed >> 16) | 1u;
    v ^= v * 0x05526c56u;
    v ^= v * 0x53a22864u;
    return reverse_bits_32(v);
}

inline uint32_t sobol_u32(uint64_t a, int dim, uint32_t scramble_seed) { // sobol.h:35-49
    uint32_t v = 0;
    for (int i = dim * 52; a != 0; a >>= 1, i++)
        if (a & 1) v ^= g_sobol.m[i];
    return owen_scramble(v, scramble_seed);
}

inline float u32_to_unit(uint32_t v) { return std::min(float(v) * 0x1p-32f, kBeforeOne); }

inline uint32_t part_1_by_1(uint32_t x) { // maths.h:381-389 (spread 16 bits to even positions)
    x &= 0x0000ffff;
    x = (x ^ (x << 8)) & 0x00ff00ff;
    x = (x ^ (x << 4)) & 0x0f0f0f0f;
    x = (x ^ (x << 2)) & 0x33333333;
    x = (x ^ (x << 1)) & 0x55555555;
    return x;
}
inline uint32_t encode_morton_2(uint32_t x, uint32_t y) { return (part_1_by_1(y) << 1) + part_1_by_1(x); }

inline int log2_int(uint32_t v) { return 31 - __builtin_clz(v); }
inline int32_t round_up_pow2(int32_t v) { // maths.h:194-203
    v--;
    v |= v >> 1;
    v |= v >> 2;
    v |= v >> 4;
    v |= v >> 8;
    v |= v >> 16;
    return v + 1;
}

// The 24 permutations of 4 digits, in the order of sobol.h:171-197 (pbrt-v4's ZSobolSampler order).
const uint8_t kPerm4[24][4] = {{0, 1, 2, 3}, {0, 1, 3, 2}, {0, 2, 1, 3}, {0, 2, 3, 1}, {0, 3, 2, 1}, {0, 3, 1, 2},
                               {1, 0, 2, 3}, {1, 0, 3, 2}, {1, 2, 0, 3}, {1, 2, 3, 0}, {1, 3, 2, 0}, {1, 3, 0, 2},
                               {2, 1, 0, 3}, {2, 1, 3, 0}, {2, 0, 1, 3}, {2, 0, 3, 1}, {2, 3, 0, 1}, {2, 3, 1, 0},
                               {3, 1, 2, 0}, {3, 1, 0, 2}, {3, 2, 1, 0}, {3, 2, 0, 1}, {3, 0, 2, 1}, {3, 0, 1, 2}};

struct ZSobol { // OwenScrambledSobol02Morton (sobol.h:131-228) + PTRenderSampler (sobol.h:235-248)
    int dimension = 0;
    int seed = 0, log2_spp = 0, n_base4_digits = 0;
    uint64_t morton_index = 0;

    ZSobol(int spp, int full_w, int full_h, int seed_, int px, int py, int sample) : seed(seed_) {
        log2_spp = log2_int((uint32_t)spp);
        int res = round_up_pow2(std::max(full_w, full_h));
        int log4_spp = (log2_spp + 1) / 2;
        n_base4_digits = log2_int((uint32_t)res) + log4_spp;
        dimension = 0;
        // 32-bit shift as in the reference (encode_morton_2 returns uint32_t).
        morton_index = (uint64_t)((encode_morton_2((uint32_t)px, (uint32_t)py) << log2_spp) | (uint32_t)sample);
    }

    uint64_t sample_index() const { // sobol.h:168-222
        uint64_t si = 0;
        bool pow2 = log2_spp & 1;
        int last_digit = pow2 ? 1 : 0;
        uint64_t dmask = (uint64_t)(0x55555555u * (uint32_t)dimension);
        for (int i = n_base4_digits - 1; i >= last_digit; --i) {
            int shift = 2 * i - (pow2 ? 1 : 0);
            int digit = (int)((morton_index >> shift) & 3);
            uint64_t higher = morton_index >> (shift + 2);
            int p = (int)((mix_bits(higher ^ dmask) >> 24) % 24);
            digit = kPerm4[p][digit];
            si |= uint64_t(digit) << shift;
        }
        if (pow2) {
            int digit = (int)(morton_index & 1);
            si |= (uint64_t)(digit ^ (int)(mix_bits((morton_index >> 1) ^ dmask) & 1));
        }
        return si;
    }

    uint32_t next_u32() { // sobol.h:148-155
        uint64_t si = sample_index();
        ++dimension;
        uint32_t h = (uint32_t)hash_dim_seed(dimension, seed);
        return sobol_u32(si, 0, h);
    }
    void next2d_u32(uint32_t out[2]) { // sobol.h:157-166
        uint64_t si = sample_index();
        dimension += 2;
        uint64_t bits = hash_dim_seed(dimension, seed);
        out[0] = sobol_u32(si, 0, (uint32_t)bits);
        out[1] = sobol_u32(si, 1, (uint32_t)(bits >> 32));
    }
    float next() { return u32_to_unit(next_u32()); }
    void next2d(float &a, float &b) {
        uint32_t u[2];
        next2d_u32(u);
        a = u32_to_unit(u[0]);
        b = u32_to_unit(u[1]);
    }
};

// ----------------------------------------------------------------------------------------------------
// Sampling warps (rng.h:243-373, bsdf.cpp:62-91)
// ----------------------------------------------------------------------------------------------------
inline void sample_disk(float u0, float u1, float &dx, float &dy) { // rng.h:243-259
    float a = 2.0f * u0 - 1.0f;
    float b = 2.0f * u1 - 1.0f;
    if (a == 0.0f && b == 0.0f) {
        dx = dy = 0.0f;
        return;
    }
    float r, phi;
    if (a * a > b * b) {
        r = a;
        phi = 0.25f * kPi * (b / a);
    } else {
        r = b;
        phi = kHalfPi - 0.25f * kPi * (a / b);
    }
    dx = r * std::cos(phi);
    dy = r * std::sin(phi);
}

inline V3 sample_cosine_hemisphere(float u0, float u1) { // rng.h:261-267
    float dx, dy;
    sample_disk(u0, u1, dx, dy);
    float z = std::sqrt(std::max(0.0f, 1.0f - (dx * dx + dy * dy)));
    return {dx, dy, z};
}

inline void orthonormal_basis(V3 n, V3 &X, V3 &Y) { // maths.h:814-821
    float sign = std::copysign(1.0f, n.z);
    const float a = -1.0f / (sign + n.z);
    const float b = n.x * n.y * a;
    X = V3(1.0f + sign * n.x * n.x * a, sign * b, -sign * n.x);
    Y = V3(b, sign + n.y * n.y * a, -n.y);
}

inline float safe_pow(float a, float b) {
    if (a == 0.0f && b == 0.0f) return 0.0f;
    return std::pow(a, b);
}
inline float hg_phase(float cos_theta, float g) { // bsdf.h:45-48
    return ((1.0f - g * g) / safe_pow(1.0f + g * g - 2.0f * g * cos_theta, 1.5f)) * (kInvPi * 0.25f);
}
V3 sample_hg(V3 D, float g, float randu, float randv, float *pdf) { // bsdf.cpp:62-91
    float cos_theta;
    bool isotropic = std::fabs(g) < 1e-3f;
    if (isotropic) {
        cos_theta = (1.0f - 2.0f * randu);
        if (pdf) *pdf = kInvPi * 0.25f;
    } else {
        float k = (1.0f - g * g) / (1.0f - g + 2.0f * g * randu);
        cos_theta = (1.0f + g * g - k * k) / (2.0f * g);
        if (pdf) *pdf = hg_phase(cos_theta, g);
    }
    float sin_theta = std::sqrt(std::max(0.0f, 1.0f - cos_theta * cos_theta));
    float phi = kTwoPi * randv;
    V3 dir(sin_theta * std::cos(phi), sin_theta * std::sin(phi), cos_theta);
    V3 T, B;
    orthonormal_basis(D, T, B);
    return T * dir.x + B * dir.y + D * dir.z;
}

// ----------------------------------------------------------------------------------------------------
// Principled BSDF (principled_bsdf.cpp:246-705, microfacet.h:20-153, fresnel.h:9-37)
// ----------------------------------------------------------------------------------------------------
struct Closure {
    V3 basecolor;
    float roughness, ax, ay, metallic, ior, specular_r0_mul, specular_trans, diffuse_trans, diffuse_trans_fwd;
};

Closure make_closure(const ora_material &m) { // principled_bsdf.cpp:246-275
    Closure c;
    c.basecolor = V3(clampv(m.basecolor[0], 0.0f, 1.0f), clampv(m.basecolor[1], 0.0f, 1.0f),
                     clampv(m.basecolor[2], 0.0f, 1.0f));
    c.roughness = clampv(m.roughness, 0.0f, 1.0f);
    c.ax = sqr(c.roughness);
    c.ay = c.ax;
    c.metallic = clampv(m.metallic, 0.0f, 1.0f);
    c.ior = clampv(m.ior, 1.0f, 4.0f);
    c.specular_r0_mul = std::max(m.specular_r0_mul * 2.0f, 0.0f);
    c.specular_trans = clampv(m.specular_trans, 0.0f, 1.0f);
    c.diffuse_trans = clampv(m.diffuse_trans, 0.0f, 1.0f);
    c.diffuse_trans_fwd = clampv(m.diffuse_trans_fwd, 0.0f, 1.0f);
    return c;
}

struct GGX { // microfacet.h:20-153
    float alpha_x, alpha_y;
    GGX(float ax, float ay) {
        alpha_x = clampv(ax, 1e-3f, 1.0f);
        alpha_y = clampv(ay, 1e-3f, 1.0f);
    }
    bool isotropic() const { return alpha_x == alpha_y; }
    float D(V3 wm) const {
        if (wm.z <= 0.0f) return 0.0f;
        if (isotropic()) {
            float a2 = alpha_x * alpha_x;
            float t = 1.0f + (a2 - 1.0f) * wm.z * wm.z;
            return a2 / (kPi * t * t);
        }
        float hx = wm.x / alpha_x, hy = wm.y / alpha_y;
        float t = sqr(hx) + sqr(hy) + sqr(wm.z);
        return 1.0f / (kPi * alpha_x * alpha_y * sqr(t));
    }
    float lambda(V3 w) const {
        if (w.z >= 1.0f || w.z <= -1.0f) return 0.0f;
        float alpha;
        if (isotropic()) {
            alpha = alpha_x;
        } else {
            float inv_sin_theta2 = 1.0f / (1.0f - w.z * w.z);
            float cos_phi2 = w.x * w.x * inv_sin_theta2;
            float sin_phi2 = w.y * w.y * inv_sin_theta2;
            alpha = std::sqrt(cos_phi2 * alpha_x * alpha_x + sin_phi2 * alpha_y * alpha_y);
        }
        float alpha2 = alpha * alpha;
        float NdotV2 = w.z * w.z;
        float t = (1.0f - NdotV2) * alpha2 / NdotV2;
        return 0.5f * (-1.0f + std::sqrt(1.0f + t));
    }
    float G1(V3 w) const { return 1.0f / (1.0f + lambda(w)); }
    float G2(V3 wo, V3 wi) const {
        if (wo.z * wi.z >= 0.0f) return 1.0f / (1.0f + lambda(wo) + lambda(wi));
        return (float)std::beta(1.0f + lambda(wo), 1.0f + lambda(wi));
    }
    V3 sample(V3 wo, float u0, float u1) const { // VNDF, microfacet.h:98-118
        V3 Vh = normalized(V3(alpha_x * wo.x, alpha_y * wo.y, wo.z));
        float lensq = Vh.x * Vh.x + Vh.y * Vh.y;
        V3 T1 = lensq > 0 ? V3(-Vh.y, Vh.x, 0.0f) * (1.0f / std::sqrt(lensq)) : V3(1.0f, 0.0f, 0.0f);
        V3 T2 = cross(Vh, T1);
        float r = std::sqrt(u0);
        float phi = kTwoPi * u1;
        float t1 = r * std::cos(phi);
        float t2 = r * std::sin(phi);
        float s = 0.5f * (1.0f + Vh.z);
        t2 = (1.0f - s) * std::sqrt(1.0f - t1 * t1) + s * t2;
        V3 Nh = T1 * t1 + T2 * t2 + Vh * std::sqrt(std::max(0.0f, 1.0f - t1 * t1 - t2 * t2));
        return normalized(V3(alpha_x * Nh.x, alpha_y * Nh.y, std::max(0.0f, Nh.z)));
    }
};

inline float fresnel_dielectric(float cos_theta_i, float eta) { // fresnel.h:9-22
    float sin_theta_t_2 = eta * eta * (1.0f - cos_theta_i * cos_theta_i);
    if (sin_theta_t_2 > 1.0) return 1.0f;
    float cos_theta_t = std::sqrt(std::max(1.0f - sin_theta_t_2, 0.0f));
    float rs = (eta * cos_theta_t - cos_theta_i) / (eta * cos_theta_t + cos_theta_i);
    float rp = (eta * cos_theta_i - cos_theta_t) / (eta * cos_theta_i + cos_theta_t);
    return 0.5f * (rs * rs + rp * rp);
}
inline float fresnel_schlick(float c) { // fresnel.h:24-30
    float s = clampv(1.0f - c, 0.0f, 1.0f);
    float s2 = s * s;
    return s2 * s2 * s;
}
inline V3 reflect(V3 w, V3 n) { return n * (2.0f * dot(n, w)) - w; } // maths.h:1037
inline bool refract(V3 wi, V3 n, float eta, V3 &wt) { // maths.h:1041-1049
    float NdotI = dot(n, wi);
    float k = 1.0f - eta * eta * (1.0f - sqr(NdotI));
    if (k < 0.0f) return false;
    wt = -wi * eta + n * (eta * NdotI - std::sqrt(k));
    return true;
}
inline float sgnf(float x) { return float((0.0f < x) - (x < 0.0f)); }

V3 eval_diffuse(V3 wo, V3 wi, const Closure &c) { // principled_bsdf.cpp:423-434
    float lw = (1.0f - c.metallic) * (1.0f - c.specular_trans) * (1.0f - c.diffuse_trans);
    if (lw == 0.0f) return V3();
    V3 wh = normalized(wo + wi);
    float Fd90 = 0.5f + 2.0f * c.roughness * wh.z;
    float retro = lerpf(1.0f, Fd90, fresnel_schlick(wi.z)) * lerpf(1.0f, Fd90, fresnel_schlick(wo.z));
    return c.basecolor * (lw * kInvPi * retro * wi.z);
}
V3 eval_diffuse_transmission(V3 wo, V3 wi, const Closure &c) { // :436-445
    float lw = (1.0f - c.metallic) * (1.0f - c.specular_trans) * c.diffuse_trans;
    if (lw == 0.0f) return V3();
    float fwd = hg_phase(dot(wo, -wi), c.diffuse_trans_fwd);
    return c.basecolor * (lw * kInvPi * (-wi.z) * fwd);
}
V3 eval_metallic_specular(V3 wo, V3 wi, const Closure &c) { // :447-460
    float lw = c.metallic;
    if (lw == 0.0f) return V3();
    GGX g(c.ax, c.ay);
    V3 wh = normalized(wo + wi);
    float D = g.D(wh);
    float G = g.G2(wo, wi);
    float fs = fresnel_schlick(dot(wo, wh));
    V3 Fr(lerpf(c.basecolor.x, 1.0f, fs), lerpf(c.basecolor.y, 1.0f, fs), lerpf(c.basecolor.z, 1.0f, fs));
    return Fr * (lw * D * G / (4.0f * wo.z));
}
float dielectric_specular_adjust(float wo_dot_wh, const Closure &c) { // :510-516
    float R0 = sqr((1.0f - c.ior) / (1.0f + c.ior));
    float schlick = fresnel_schlick(wo_dot_wh);
    float den = lerpf(R0, 1.0f, schlick);
    // ior == 1 and |wo.wh| == 1 give 0/0 in the reference (NaN, then its non-finite monitor aborts);
    // the continuous limit there is 1.
    return den > 0.0f ? lerpf(c.specular_r0_mul * R0, 1.0f, schlick) / den : 1.0f;
}
V3 eval_dielectric_specular(V3 wo, V3 wi, const Closure &c) { // :462-508
    float lw = 1.0f - c.metallic;
    if (lw == 0.0f) return V3();
    bool refl = wo.z * wi.z >= 0.0f;
    float eta = wo.z >= 0.0f ? c.ior : (1.0f / c.ior);
    V3 wh;
    if (refl) {
        wh = normalized(wo + wi);
    } else {
        wh = normalized(wo + wi * eta);
        lw *= c.specular_trans;
    }
    if (wh.z < 0.0f) wh = -wh;
    GGX g(c.ax, c.ay);
    float D = g.D(wh);
    float G = g.G2(wo, wi);
    float awh = std::fabs(dot(wo, wh));
    float Fr = fresnel_dielectric(awh, 1.0f / eta);
    float f;
    if (refl) {
        float adj = dielectric_specular_adjust(awh, c);
        f = D * G * (Fr * adj) / (4.0f * std::fabs(wo.z));
    } else {
        float denom = sqr(dot(wo, wh) + eta * dot(wi, wh));
        float spec = D * G * (1.0f - Fr) * dot(wo, wh) * dot(wi, wh) / (denom * wo.z);
        f = std::fabs(spec);
    }
    f *= lw;
    return V3(f, f, f);
}

void lobe_weights(V3 wo, const Closure &c, float w[4]) { // :518-540
    float lum = luminance(c.basecolor);
    float wd = wo.z > 0.0f ? (1.0f - c.metallic) * (1.0f - c.specular_trans) * lum * kInvPi : 0.0f;
    w[0] = wd * (1.0f - c.diffuse_trans);
    w[1] = wd * c.diffuse_trans;
    w[2] = wo.z > 0.0f ? c.metallic * lerpf(lum, 1.0f, fresnel_schlick(wo.z)) : 0.0f;
    float eta = wo.z >= 0.0f ? c.ior : (1.0f / c.ior);
    w[3] = (1.0f - c.metallic) * fresnel_dielectric(std::fabs(wo.z), 1.0f / eta);
    float sum = w[0] + w[1] + w[2] + w[3];
    if (sum == 0.0f) {
        w[0] = w[1] = w[2] = w[3] = 0.0f;
        return;
    }
    for (int i = 0; i < 4; ++i) w[i] /= sum;
}

float pdf_diffuse(V3, V3 wi, const Closure &) { return wi.z * kInvPi; }
float pdf_diffuse_transmission(V3 wo, V3 wi, const Closure &c) {
    if (c.diffuse_trans_fwd > 0.05f) return hg_phase(dot(wo, -wi), c.diffuse_trans_fwd);
    return -wi.z * kInvPi;
}
float pdf_metallic_specular(V3 wo, V3 wi, const Closure &c) {
    GGX g(c.ax, c.ay);
    V3 wh = normalized(wo + wi);
    return g.D(wh) * g.G1(wo) / (4.0f * std::fabs(wo.z));
}
float pdf_dielectric_specular(V3 wo, V3 wi, const Closure &c) { // :667-703
    bool refl = wo.z * wi.z >= 0.0f;
    float eta = wo.z >= 0.0f ? c.ior : (1.0f / c.ior);
    V3 wh = refl ? normalized(wo + wi) : normalized(wo + wi * eta);
    if (wh.z < 0.0f) wh = -wh;
    GGX g(c.ax, c.ay);
    float D = g.D(wh), G1 = g.G1(wo);
    float awh = std::fabs(dot(wo, wh));
    float Fr = fresnel_dielectric(awh, 1.0f / eta);
    float Fr_refl = Fr * dielectric_specular_adjust(awh, c);
    float Fr_refr = (1.0f - Fr) * c.specular_trans;
    if (Fr_refr == 0.0f) Fr_refl = 1.0f;
    float pdf;
    if (refl) {
        pdf = D * G1 / (4.0f * std::fabs(wo.z));
        pdf *= Fr_refl / (Fr_refl + Fr_refr);
    } else {
        float denom = sqr(dot(wo, wh) + eta * dot(wi, wh));
        float jac = eta * eta * std::fabs(dot(wi, wh)) / denom;
        pdf = D * G1 * std::fabs(dot(wo, wh)) / std::fabs(wo.z) * jac;
        pdf *= Fr_refr / (Fr_refl + Fr_refr);
    }
    return pdf;
}

V3 bsdf_eval(V3 wo, V3 wi, const Closure &c) { // :308-324
    if (wo.z == 0.0f || wi.z == 0.0f) return V3();
    V3 f;
    if (wo.z > 0.0f && wi.z > 0.0f) {
        f = f + eval_diffuse(wo, wi, c);
        f = f + eval_metallic_specular(wo, wi, c);
    }
    if (wo.z > 0.0f && wi.z < 0.0f) f = f + eval_diffuse_transmission(wo, wi, c);
    f = f + eval_dielectric_specular(wo, wi, c);
    return f;
}
float bsdf_pdf(V3 wo, V3 wi, const Closure &c) { // :395-421
    if (wo.z == 0.0f || wi.z == 0.0f) return 0.0f;
    float w[4];
    lobe_weights(wo, c, w);
    if (w[0] == 0.0f && w[1] == 0.0f && w[2] == 0.0f && w[3] == 0.0f) return 0.0f;
    float p[4] = {0, 0, 0, 0};
    if (wo.z > 0.0f && wi.z > 0.0f) {
        p[0] = pdf_diffuse(wo, wi, c);
        p[2] = pdf_metallic_specular(wo, wi, c);
    }
    if (wo.z > 0.0f && wi.z < 0.0f) p[1] = pdf_diffuse_transmission(wo, wi, c);
    p[3] = pdf_dielectric_specular(wo, wi, c);
    return p[0] * w[0] + p[1] * w[1] + p[2] * w[2] + p[3] * w[3];
}
void bsdf_eval_and_pdf(V3 wo, V3 wi, const Closure &c, V3 &f, float &pdf) { // :297-306
    f = bsdf_eval(wo, wi, c);
    if (is_zero(f)) {
        pdf = 0.0f;
        return;
    }
    pdf = bsdf_pdf(wo, wi, c);
}

int sample_small_distrib4(const float w[4], float u, float *u_remap) { // rng.h:399-432
    float sum_w = 0.0f;
    int last_positive = 0;
    for (int i = 0; i < 4; ++i) {
        sum_w += w[i];
        if (w[i] > 0.0f) last_positive = i;
    }
    float inv = 1.0f / sum_w;
    float cdf = 0.0f;
    int sel = -1;
    for (int i = 0; i < 4; ++i) {
        float cdf_next = cdf + w[i] * inv;
        if (u < cdf_next) {
            sel = i;
            if (u_remap) *u_remap = (u - cdf) / (cdf_next - cdf);
            break;
        }
        cdf = cdf_next;
    }
    if (sel == -1) sel = last_positive;
    return sel;
}

V3 bsdf_sample(V3 wo, V3 &wi, const Closure &c, float u_lobe, float u0, float u1, float &pdf) { // :326-393
    if (wo.z == 0.0f) {
        pdf = 0.0f;
        return V3();
    }
    float w[4];
    lobe_weights(wo, c, w);
    if (w[0] == 0.0f && w[1] == 0.0f && w[2] == 0.0f && w[3] == 0.0f) {
        pdf = 0.0f;
        return V3();
    }
    float u_remap = 0.0f;
    int lobe = sample_small_distrib4(w, u_lobe, &u_remap);
    float p[4] = {0, 0, 0, 0};
    GGX g(c.ax, c.ay);
    if (lobe == 0) {
        wi = sample_cosine_hemisphere(u0, u1);
        p[0] = wi.z * kInvPi;
        if (is_zero(wi) || p[0] == 0.0f) {
            pdf = 0.0f;
            return V3();
        }
        if (wo.z > 0.0f && wi.z > 0.0f) p[2] = pdf_metallic_specular(wo, wi, c);
        p[3] = pdf_dielectric_specular(wo, wi, c);
    } else if (lobe == 1) {
        if (c.diffuse_trans_fwd > 0.05f) {
            wi = sample_hg(-wo, c.diffuse_trans_fwd, u0, u1, &p[1]);
        } else {
            wi = sample_cosine_hemisphere(u0, u1);
            p[1] = wi.z * kInvPi;
            wi.z = -wi.z;
        }
        if (is_zero(wi) || p[1] == 0.0f) {
            pdf = 0.0f;
            return V3();
        }
        p[3] = pdf_dielectric_specular(wo, wi, c);
    } else if (lobe == 2) {
        // sample_metallic_specular :569-588
        V3 wh = g.sample(wo * sgnf(wo.z), u0, u1);
        wi = reflect(wo, wh);
        if (wo.z * wi.z < 0.0f) {
            pdf = 0.0f;
            return V3();
        }
        p[2] = g.D(wh) * g.G1(wo) / (4.0f * std::fabs(wo.z));
        if (is_zero(wi) || p[2] == 0.0f) {
            pdf = 0.0f;
            return V3();
        }
        if (wo.z > 0.0f && wi.z > 0.0f) p[0] = pdf_diffuse(wo, wi, c);
        p[3] = pdf_dielectric_specular(wo, wi, c);
    } else {
        // sample_dielectric_specular :590-643
        V3 wh = g.sample(wo * sgnf(wo.z), u0, u1);
        float D = g.D(wh), G1 = g.G1(wo);
        float eta = wo.z >= 0.0f ? c.ior : (1.0f / c.ior);
        float awh = std::fabs(dot(wo, wh));
        float Fr = fresnel_dielectric(std::fabs(dot(wo, wh)), 1.0f / eta);
        float Fr_refl = Fr * dielectric_specular_adjust(awh, c);
        float Fr_refr = (1.0f - Fr) * c.specular_trans;
        if (Fr_refr == 0.0f) Fr_refl = 1.0f;
        if (u_remap < Fr_refl / (Fr_refl + Fr_refr)) {
            wi = reflect(wo, wh);
            if (wo.z * wi.z < 0.0f) {
                pdf = 0.0f;
                return V3();
            }
            p[3] = D * G1 / (4.0f * std::fabs(wo.z));
            p[3] *= Fr_refl / (Fr_refl + Fr_refr);
        } else {
            if (!refract(wo, wh * sgnf(wo.z), 1.0f / eta, wi)) {
                pdf = 0.0f;
                return V3();
            }
            if (wo.z * wi.z > 0.0f) {
                pdf = 0.0f;
                return V3();
            }
            float denom = sqr(dot(wo, wh) + eta * dot(wi, wh));
            float jac = eta * eta * std::fabs(dot(wi, wh)) / denom;
            p[3] = D * G1 * std::fabs(dot(wo, wh)) / std::fabs(wo.z) * jac;
            p[3] *= Fr_refr / (Fr_refl + Fr_refr);
        }
        if (is_zero(wi) || p[3] == 0.0f) {
            pdf = 0.0f;
            return V3();
        }
        if (wo.z > 0.0f && wi.z > 0.0f) {
            p[0] = pdf_diffuse(wo, wi, c);
            p[2] = pdf_metallic_specular(wo, wi, c);
        }
        if (wo.z > 0.0f && wi.z < 0.0f) p[1] = pdf_diffuse_transmission(wo, wi, c);
    }
    pdf = p[0] * w[0] + p[1] * w[1] + p[2] * w[2] + p[3] * w[3];
    return bsdf_eval(wo, wi, c) / pdf;
}

// ----------------------------------------------------------------------------------------------------
// Camera (camera.cpp:12-170, maths.h:971-995)
// ----------------------------------------------------------------------------------------------------
struct M4 {
    double a[4][4];
};
M4 mul(const M4 &x, const M4 &y) {
    M4 r;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
            double s = 0;
            for (int k = 0; k < 4; ++k) s += x.a[i][k] * y.a[k][j];
            r.a[i][j] = s;
        }
    return r;
}
M4 inverse(const M4 &m) { // Gauss-Jordan with partial pivoting (Eigen uses cofactors; agree to rounding)
    double a[4][8];
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 8; ++j) a[i][j] = j < 4 ? m.a[i][j] : (j - 4 == i ? 1.0 : 0.0);
    for (int c = 0; c < 4; ++c) {
        int p = c;
        for (int r = c + 1; r < 4; ++r)
            if (std::fabs(a[r][c]) > std::fabs(a[p][c])) p = r;
        for (int j = 0; j < 8; ++j) std::swap(a[c][j], a[p][j]);
        double d = a[c][c];
        for (int j = 0; j < 8; ++j) a[c][j] /= d;
        for (int r = 0; r < 4; ++r)
            if (r != c) {
                double f = a[r][c];
                for (int j = 0; j < 8; ++j) a[r][j] -= f * a[c][j];
            }
    }
    M4 r;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) r.a[i][j] = a[i][j + 4];
    return r;
}

struct Camera {
    M4 proj_to_world, proj_to_camera, camera_to_world;
    V3 position;

    explicit Camera(const ora_camera &c) {
        // create_camera: up is force-normalised on load (camera.cpp:179, config.cpp:146)
        V3 pos(c.pos[0], c.pos[1], c.pos[2]), tgt(c.target[0], c.target[1], c.target[2]);
        V3 up = normalized(V3(c.up[0], c.up[1], c.up[2]));
        // look_at (camera.cpp:12-32)
        V3 back = normalized(pos - tgt);
        V3 right = cross(up, back);
        if (sqnorm(right) < 1e-8f) {
            orthonormal_basis(back, right, up);
        } else {
            right = normalized(right);
            up = cross(back, right);
        }
        M4 c2w{{{right.x, up.x, back.x, pos.x}, {right.y, up.y, back.y, pos.y}, {right.z, up.z, back.z, pos.z}, {0, 0, 0, 1}}};
        camera_to_world = c2w;
        float vfov = c.vfov_deg / 180.0f * kPi; // to_radian, maths.h:454
        // rev_inf_projection (camera.cpp:56-70), near_clip = 0.1
        float cot = 1.0f / std::tan(vfov * 0.5f);
        M4 proj{{{cot / c.aspect, 0, 0, 0}, {0, -cot, 0, 0}, {0, 0, 0, 0.1f}, {0, 0, -1, 0}}};
        M4 w2c = inverse(c2w);
        M4 w2p = mul(proj, w2c);
        proj_to_world = inverse(w2p);
        proj_to_camera = inverse(proj);
        position = pos;
    }
    static V3 point_hdiv(const M4 &m, V3 p) {
        double v[4] = {p.x, p.y, p.z, 1.0}, r[4];
        for (int i = 0; i < 4; ++i) r[i] = m.a[i][0] * v[0] + m.a[i][1] * v[1] + m.a[i][2] * v[2] + m.a[i][3] * v[3];
        return V3((float)(r[0] / r[3]), (float)(r[1] / r[3]), (float)(r[2] / r[3]));
    }
    // spawn_ray (camera.cpp:129-170), perspective branch, with the ray differentials' directions for a
    // film of res_x x res_y, scaled by 1 / sqrt(spp)
    void spawn_ray(float fx, float fy, V3 &o, V3 &d, V3 *rxd = nullptr, V3 *ryd = nullptr, int res_x = 1,
                   int res_y = 1, int spp = 1) const {
        V3 ndc(fx * 2.0f - 1.0f, fy * 2.0f - 1.0f, 1.0f);
        V3 world_pos = point_hdiv(proj_to_world, ndc);
        o = position;
        d = normalized(world_pos - position);
        if (!rxd) return;
        V3 mid = point_hdiv(proj_to_camera, V3(0.0f, 0.0f, 1.0f));
        V3 right = point_hdiv(proj_to_camera, V3(2.0f / res_x, 0.0f, 1.0f));
        V3 up = point_hdiv(proj_to_camera, V3(0.0f, 2.0f / res_y, 1.0f));
        auto dir = [&](V3 v) { // camera_to_world.direction
            const M4 &m = camera_to_world;
            return V3((float)(m.a[0][0] * v.x + m.a[0][1] * v.y + m.a[0][2] * v.z),
                      (float)(m.a[1][0] * v.x + m.a[1][1] * v.y + m.a[1][2] * v.z),
                      (float)(m.a[2][0] * v.x + m.a[2][1] * v.y + m.a[2][2] * v.z));
        };
        V3 cdx = dir(right - mid), cdy = dir(up - mid);
        *rxd = normalized(world_pos + cdx - position);
        *ryd = normalized(world_pos + cdy - position);
        float scale = 1.0f / std::sqrt((float)spp);
        *rxd = d + (*rxd - d) * scale;
        *ryd = d + (*ryd - d) * scale;
    }
};

// ----------------------------------------------------------------------------------------------------
// Scene, BVH (independent binned SAH) and ray queries
// ----------------------------------------------------------------------------------------------------
struct AABB {
    V3 lo{kInf, kInf, kInf}, hi{-kInf, -kInf, -kInf};
    void expand(V3 p) {
        lo = V3(std::min(lo.x, p.x), std::min(lo.y, p.y), std::min(lo.z, p.z));
        hi = V3(std::max(hi.x, p.x), std::max(hi.y, p.y), std::max(hi.z, p.z));
    }
    void expand(const AABB &b) {
        expand(b.lo);
        expand(b.hi);
    }
    float area() const {
        V3 e = hi - lo;
        if (e.x < 0) return 0.0f;
        return 2.0f * (e.x * e.y + e.y * e.z + e.x * e.z);
    }
    V3 center() const { return (lo + hi) * 0.5f; }
};

struct Node { // reference SBVHNode semantics (cuda/sbvh.cuh:18-27) with an explicit leaf prim count
    AABB box;
    uint32_t axis = 0, right = 0, prim_offset = 0, prim_count = 0;
};

// DistribTable (distrib.cpp:9-65): piecewise-constant distribution over [0, 1) from n weights.
struct Distrib1D {
    std::vector<float> cdf;
    float acc = 0.0f;

    void build(const float *f, uint32_t n) {
        cdf.assign(n + 1, 0.0f);
        const float invn = 1.0f / (float)n;
        for (uint32_t i = 1; i <= n; ++i) cdf[i] = cdf[i - 1] + f[i - 1] * invn;
        acc = cdf[n];
        for (uint32_t i = 0; i <= n; ++i) cdf[i] = acc == 0.0f ? (float)i / (float)n : cdf[i] * (1.0f / acc);
        cdf[n] = 1.0f;
    }
    uint32_t size() const { return (uint32_t)cdf.size() - 1; }
    // sample_linear (distrib.cpp:43-55): continuous sample in [0, 1), density, bin
    float sample_linear(float u, float &pdf, uint32_t &bin) const {
        u = clampv(u, 0.0f, std::nextafter(1.0f, 0.0f));
        bin = (uint32_t)(std::upper_bound(cdf.begin(), cdf.end(), u) - cdf.begin()) - 1;
        const float lo = cdf[bin], hi = cdf[bin + 1];
        pdf = (hi - lo) * (float)size();
        return ((float)bin + (u - lo) / (hi - lo)) / (float)size();
    }
    float pdf(uint32_t bin) const { return (cdf[bin + 1] - cdf[bin]) * (float)size(); } // distrib.cpp:57-62
};

// DistribTable2D (distrib.cpp:78-107): one conditional table per row v, a marginal table over rows.
struct Distrib2D {
    std::vector<Distrib1D> rows;
    Distrib1D margin;

    void build(const std::vector<float> &f, uint32_t nu, uint32_t nv) {
        rows.assign(nv, Distrib1D());
        std::vector<float> row_weight(nv);
        for (uint32_t v = 0; v < nv; ++v) {
            rows[v].build(f.data() + (size_t)v * nu, nu);
            row_weight[v] = rows[v].acc;
        }
        margin.build(row_weight.data(), nv);
    }
    void sample_linear(float u0, float u1, float &su, float &sv, float &pdf) const {
        float pm, pc;
        uint32_t bv, bu;
        sv = margin.sample_linear(u1, pm, bv);
        su = rows[bv].sample_linear(u0, pc, bu);
        pdf = pm * pc;
    }
    float pdf(float u, float v) const {
        const int nu = (int)rows[0].size(), nv = (int)rows.size();
        int x = std::clamp((int)(u * (float)nu), 0, nu - 1);
        int y = std::clamp((int)(v * (float)nv), 0, nv - 1);
        return margin.pdf((uint32_t)y) * rows[y].pdf((uint32_t)x);
    }
};

// SkyLight from an equirectangular HDR map (light.cpp:19-49 and :59-158). Texel (x, y) sits at
// phi = 2 pi (x + 0.5) / w, theta = pi (y + 0.5) / h of the map frame.
struct SkyMapLight {
    int w = 0, h = 0;
    std::vector<V3> texels;
    float l2w[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1}, w2l[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    bool yup = true;
    float strength = 1.0f;
    Distrib2D distrib;

    static V3 apply(const float m[9], V3 d) {
        return V3(m[0] * d.x + m[1] * d.y + m[2] * d.z, m[3] * d.x + m[4] * d.y + m[5] * d.z,
                  m[6] * d.x + m[7] * d.y + m[8] * d.z);
    }
    void init(const ora_sky_map &d) {
        w = d.width;
        h = d.height;
        texels.resize((size_t)w * h);
        std::vector<float> lum((size_t)w * h);
        float total = 0.0f;
        for (size_t i = 0; i < texels.size(); ++i) {
            texels[i] = V3(d.rgb[3 * i], d.rgb[3 * i + 1], d.rgb[3 * i + 2]);
            const float theta = ((float)(i / (size_t)w) + 0.5f) / (float)h * kPi;
            lum[i] = std::sin(theta) * luminance(texels[i]);
            total += lum[i];
        }
        // thresholded table: texels below the mean weighted luminance are left to BSDF sampling
        const float mean = total / (float)lum.size();
        for (float &l : lum) l = std::max(l - mean, 0.0f);
        distrib.build(lum, (uint32_t)w, (uint32_t)h);
        for (int k = 0; k < 9; ++k) l2w[k] = d.to_world[k];
        invert3(l2w, w2l);
        yup = d.transform_y_up != 0;
        strength = d.strength;
    }
    // Gauss-Jordan inverse with partial pivoting in double
    static void invert3(const float in[9], float out[9]) {
        double a[3][6];
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 6; ++c) a[r][c] = c < 3 ? in[3 * r + c] : (c - 3 == r ? 1.0 : 0.0);
        for (int c = 0; c < 3; ++c) {
            int piv = c;
            for (int r = c + 1; r < 3; ++r)
                if (std::fabs(a[r][c]) > std::fabs(a[piv][c])) piv = r;
            for (int k = 0; k < 6; ++k) std::swap(a[c][k], a[piv][k]);
            const double inv = 1.0 / a[c][c];
            for (int k = 0; k < 6; ++k) a[c][k] *= inv;
            for (int r = 0; r < 3; ++r) {
                if (r == c) continue;
                const double f = a[r][c];
                for (int k = 0; k < 6; ++k) a[r][k] -= f * a[c][k];
            }
        }
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) out[3 * r + c] = (float)a[r][3 + c];
    }
    // to_spherical / to_spherical_yup (maths.h:458-484) of the map-frame direction
    void spherical(V3 wi, float &phi, float &theta) const {
        V3 l = apply(w2l, wi);
        const float up = yup ? l.y : l.z, a = yup ? l.z : l.y;
        theta = std::acos(clampv(up, -1.0f, 1.0f));
        phi = std::atan2(a, l.x);
        if (phi < 0.0f) phi += kTwoPi;
    }
    V3 texel(int x, int y) const { return texels[(size_t)y * w + x]; }
    // SkyLight::eval (light.cpp:59-86): bilinear over texel centres, u wraps, v clamps (maths.h:760-780)
    V3 eval(V3 wi) const {
        float phi, theta;
        spherical(wi, phi, theta);
        const float us = phi * kInvPi * 0.5f * (float)w - 0.5f, vs = theta * kInvPi * (float)h - 0.5f;
        const int x0 = (int)std::floor(us), y0 = (int)std::floor(vs);
        const float tu = us - (float)x0, tv = vs - (float)y0;
        const int xa = (x0 + w) % w, xb = (x0 + 1 + w) % w;
        const int ya = std::clamp(y0, 0, h - 1), yb = std::clamp(y0 + 1, 0, h - 1);
        V3 L;
        L = L + texel(xa, ya) * ((1.0f - tu) * (1.0f - tv));
        L = L + texel(xa, yb) * ((1.0f - tu) * tv);
        L = L + texel(xb, ya) * (tu * (1.0f - tv));
        L = L + texel(xb, yb) * (tu * tv);
        return L * strength;
    }
    // SkyLight::pdf (light.cpp:113-134), solid-angle density
    float pdf(V3 wi) const {
        float phi, theta;
        spherical(wi, phi, theta);
        const float st = std::sin(theta);
        if (st == 0.0f) return 0.0f;
        return distrib.pdf(phi * kInvPi * 0.5f, theta * kInvPi) / (2.0f * kPi * kPi * st);
    }
    // SkyLight::sample (light.cpp:88-111): Le / pdf
    V3 sample(float u0, float u1, V3 &wi, float &pdf) const {
        float su, sv;
        distrib.sample_linear(u0, u1, su, sv, pdf);
        if (pdf == 0.0f) return V3();
        const float phi = su * 2.0f * kPi, theta = sv * kPi;
        const float ct = std::cos(theta), st = std::sin(theta), cp = std::cos(phi), sp = std::sin(phi);
        const V3 l = yup ? V3(cp * st, ct, sp * st) : V3(cp * st, sp * st, ct); // maths.h:467-493
        pdf /= (2.0f * kPi * kPi * st);
        if (st == 0.0f) {
            pdf = 0.0f;
            return V3();
        }
        wi = apply(l2w, l);
        return eval(wi) / pdf;
    }
    // SkyLight::power (light.cpp:136-158): midpoint rule over a z-up (theta, phi) grid of the map's size
    V3 power(float scene_radius) const {
        V3 sum;
        for (int row = 0; row < h; ++row) {
            const float theta = ((float)row + 0.5f) / (float)h * kPi, st = std::sin(theta);
            V3 row_sum;
            for (int col = 0; col < w; ++col) {
                const float phi = ((float)col + 0.5f) / (float)w * kTwoPi;
                const V3 dir(std::cos(phi) * st, std::sin(phi) * st, std::cos(theta));
                row_sum = row_sum + eval(dir) * st;
            }
            sum = sum + row_sum;
        }
        return sum * ((kPi / (float)h) * (kTwoPi / (float)w)) * (kPi * sqr(scene_radius));
    }
};

// ----------------------------------------------------------------------------------------------------
// Textures and shader fields (texture.h, texture.cpp:12-1121, shader_field.h:11-86)
// ----------------------------------------------------------------------------------------------------
float srgb_to_linear_f(float x) { return x < 0.04045f ? x / 12.92f : std::pow((x + 0.055f) / 1.055f, 2.4f); } // maths.h:631
float linear_to_srgb_f(float x) { // maths.h:649-659
    if (x <= 0.0f) return 0.0f;
    if (x >= 1.0f) return 1.0f;
    if (x < 0.0031308f) return x * 12.92f;
    return std::pow(x, 1.0f / 2.4f) * 1.055f - 0.055f;
}

struct TexTables {
    float srgb[256]; // srgb_u8_to_linear_f32_lut (texture.cpp:376-401): srgb_to_linear(i / 255) in float
    float ewa[128];  // ewa_lut (texture.cpp:720-856): its generator (texture.cpp:720-726) in float
    TexTables() {
        for (int i = 0; i < 256; ++i) srgb[i] = srgb_to_linear_f((float)i / 255.0f);
        for (int i = 0; i < 128; ++i) {
            float alpha = 2, r2 = float(i) / float(128 - 1);
            ewa[i] = std::exp(-alpha * r2) - std::exp(-alpha);
        }
    }
};
const TexTables g_tex_tables;

struct Texture {
    int width = 0, height = 0, nc = 0, type = 0; // type 0 u8, 1 u16, 2 f32
    bool srgb = false;
    struct Level {
        int w, h;
        std::vector<uint8_t> bytes;
    };
    std::vector<Level> mips;

    int stride() const { return nc * (type == 0 ? 1 : (type == 1 ? 2 : 4)); }
    int levels() const { return (int)mips.size(); }
    // Texture::fetch_as_float (texture.cpp:403-460)
    void fetch(int x, int y, int l, float *out) const {
        const uint8_t *b = mips[l].bytes.data() + ((size_t)y * mips[l].w + x) * stride();
        for (int c = 0; c < nc; ++c) {
            if (type == 0) {
                out[c] = srgb ? g_tex_tables.srgb[b[c]] : (float)b[c] / 255.0f;
            } else if (type == 1) {
                uint16_t v;
                std::memcpy(&v, b + 2 * c, 2);
                out[c] = (float)v / 65535.0f;
                if (srgb) out[c] = srgb_to_linear_f(out[c]);
            } else {
                float v;
                std::memcpy(&v, b + 4 * c, 4);
                out[c] = srgb ? srgb_to_linear_f(v) : v;
            }
        }
    }
    // Texture::set_from_float (texture.cpp:462-518)
    void store(int x, int y, int l, const float *in) {
        uint8_t *b = mips[l].bytes.data() + ((size_t)y * mips[l].w + x) * stride();
        for (int c = 0; c < nc; ++c) {
            float f = srgb ? linear_to_srgb_f(in[c]) : in[c];
            if (type == 0) {
                b[c] = (uint8_t)std::floor(f * 255.0f);
            } else if (type == 1) {
                uint16_t v = (uint16_t)std::floor(f * 65535.0f);
                std::memcpy(b + 2 * c, &v, 2);
            } else {
                std::memcpy(b + 4 * c, &f, 4);
            }
        }
    }
    // One texel of level l as the weighted sum of level l - 1 texels (each term weight * value).
    void filter_texel(int l, int x, int y, const int *xs, const float *wx, int nx, const int *ys, const float *wy,
                      int ny, bool wx_only, bool wy_only) {
        float sum[4] = {0, 0, 0, 0}, v[4];
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                fetch(xs[i], ys[j], l - 1, v);
                float w = wx_only ? wx[i] : (wy_only ? wy[j] : wx[i] * wy[j]);
                for (int c = 0; c < nc; ++c) sum[c] += w * v[c];
            }
        store(x, y, l, sum);
    }
    // Texture(bytes, ..., build_mipmaps) (texture.cpp:21-347): the polyphase NPOT box filter. An odd
    // size 2n - 1 maps to n with taps 2x - 1, 2x, 2x + 1 weighted x, n, n - x - 1 over 2n - 1; an even
    // size halves with two taps of 1/2; a size of 1 stays.
    void build_level(int l) {
        const int lw = mips[l - 1].w, lh = mips[l - 1].h, w = mips[l].w, h = mips[l].h;
        const bool odd_x = lw == 2 * w - 1, odd_y = lh == 2 * h - 1;
        for (int y = 0; y < h; ++y)
            for (int x = 0; x < w; ++x) {
                int xs[3], ys[3], nx = 0, ny = 0;
                float wx[3], wy[3];
                if (lw == 1) {
                    xs[nx] = x, wx[nx++] = 1.0f;
                } else if (odd_x) {
                    if (2 * x - 1 > 0) xs[nx] = 2 * x - 1, wx[nx++] = (float)x / (float)lw;
                    xs[nx] = 2 * x, wx[nx++] = (float)w / (float)lw;
                    if (2 * x + 1 < lw) xs[nx] = 2 * x + 1, wx[nx++] = (float)(w - x - 1) / (float)lw;
                } else {
                    xs[nx] = 2 * x, wx[nx++] = 0.5f;
                    xs[nx] = 2 * x + 1, wx[nx++] = 0.5f;
                }
                if (lh == 1) {
                    ys[ny] = y, wy[ny++] = 1.0f;
                } else if (odd_y) {
                    if (2 * y - 1 > 0) ys[ny] = 2 * y - 1, wy[ny++] = (float)y / (float)lh;
                    ys[ny] = 2 * y, wy[ny++] = (float)h / (float)lh;
                    if (2 * y + 1 < lh) ys[ny] = 2 * y + 1, wy[ny++] = (float)(h - y - 1) / (float)lh;
                } else {
                    ys[ny] = 2 * y, wy[ny++] = 0.5f;
                    ys[ny] = 2 * y + 1, wy[ny++] = 0.5f;
                }
                filter_texel(l, x, y, xs, wx, nx, ys, wy, ny, lh == 1, lw == 1);
            }
    }
    bool init(const ora_texture &d) {
        if (!d.data || d.width <= 0 || d.height <= 0 || d.channels < 1 || d.channels > 4 || d.data_type < 0 ||
            d.data_type > 2 || d.levels < 1)
            return false;
        width = d.width, height = d.height, nc = d.channels, type = d.data_type, srgb = d.color_space == 1;
        int full = 1 + (int)std::ceil(std::log2((double)std::max(width, height)));
        if (d.levels > full) return false;
        int n = (d.levels == 1 && d.build_mipmaps) ? full : d.levels;
        const uint8_t *src = (const uint8_t *)d.data;
        int w = width, h = height;
        for (int l = 0; l < n; ++l) {
            Level lv{w, h, std::vector<uint8_t>((size_t)w * h * stride())};
            if (l < d.levels) {
                std::memcpy(lv.bytes.data(), src, lv.bytes.size());
                src += lv.bytes.size();
            }
            mips.push_back(std::move(lv));
            w = std::max(1, (w + 1) / 2);
            h = std::max(1, (h + 1) / 2);
        }
        for (int l = d.levels; l < n; ++l) build_level(l);
        return true;
    }
};

int wrap_index(int x, int dim, int mode) { // texture.cpp:520-531
    if (mode == 0) return ((x % dim) + dim) % dim;
    return std::clamp(x, 0, dim - 1);
}

struct Sampler { // TextureSampler and its kinds (texture.h:68-145)
    int kind = 1, wrap_u = 0, wrap_v = 0;
    float ca[4] = {0, 0, 0, 0}, cb[4] = {0, 0, 0, 0};
    float anisotropy = 8.0f;

    void bilinear(const Texture &t, int l, float su, float sv, float *out) const { // texture.cpp:541-573
        float u = su * t.mips[l].w - 0.5f, v = sv * t.mips[l].h - 0.5f;
        int u0 = (int)std::floor(u), v0 = (int)std::floor(v);
        float du = u - u0, dv = v - v0;
        u0 = wrap_index(u0, t.mips[l].w, wrap_u);
        v0 = wrap_index(v0, t.mips[l].h, wrap_v);
        int u1 = wrap_index(u0 + 1, t.mips[l].w, wrap_u), v1 = wrap_index(v0 + 1, t.mips[l].h, wrap_v);
        float a[4], b[4], c[4], d[4];
        t.fetch(u0, v0, l, a);
        t.fetch(u1, v0, l, b);
        t.fetch(u0, v1, l, c);
        t.fetch(u1, v1, l, d);
        float w00 = (1 - du) * (1 - dv), w10 = du * (1 - dv), w01 = (1 - du) * dv, w11 = du * dv;
        for (int k = 0; k < t.nc; ++k) out[k] = w00 * a[k] + w10 * b[k] + w01 * c[k] + w11 * d[k];
    }
    static float poly(const float *k, float x) { return k[0] * (x * x * x) + k[1] * (x * x) + k[2] * x + k[3]; }
    void bicubic(const Texture &t, int l, float su, float sv, float *out) const { // texture.cpp:629-677
        float u = su * t.mips[l].w - 0.5f, v = sv * t.mips[l].h - 0.5f;
        int u0 = (int)std::floor(u), v0 = (int)std::floor(v);
        float du = u - u0, dv = v - v0;
        float rows[4][4];
        for (int y = 0; y < 4; ++y) {
            int vy = wrap_index(v0 + y - 1, t.mips[l].h, wrap_v);
            float c[4][4];
            for (int x = 0; x < 4; ++x) t.fetch(wrap_index(u0 + x - 1, t.mips[l].w, wrap_u), vy, l, c[x]);
            // spline (texture.cpp:631-641) with x = du
            float a0 = poly(cb, (float)(du + 1.0)), a1 = poly(ca, du), a2 = poly(ca, (float)(1.0 - du)),
                  a3 = poly(cb, (float)(2.0 - du));
            for (int k = 0; k < t.nc; ++k) rows[y][k] = c[0][k] * a0 + c[1][k] * a1 + c[2][k] * a2 + c[3][k] * a3;
        }
        float a0 = poly(cb, (float)(dv + 1.0)), a1 = poly(ca, dv), a2 = poly(ca, (float)(1.0 - dv)),
              a3 = poly(cb, (float)(2.0 - dv));
        for (int k = 0; k < t.nc; ++k) out[k] = rows[0][k] * a0 + rows[1][k] * a1 + rows[2][k] * a2 + rows[3][k] * a3;
    }
    void ewa(const Texture &t, int l, float u, float v, float maj[2], float mnr[2], float *out) const { // :858-920
        u = u * t.mips[l].w - 0.5f;
        v = v * t.mips[l].h - 0.5f;
        float M0 = maj[0] * t.mips[l].w, M1 = maj[1] * t.mips[l].h;
        float m0 = mnr[0] * t.mips[l].w, m1 = mnr[1] * t.mips[l].h;
        float A = M1 * M1 + m1 * m1 + 1, B = -2 * (M0 * M1 + m0 * m1), C = M0 * M0 + m0 * m0 + 1;
        float invF = 1 / (A * C - B * B * 0.25f);
        A *= invF, B *= invF, C *= invF;
        float det = -B * B + 4 * A * C, invDet = 1 / det;
        float uS = std::sqrt(std::max(0.0f, det * C)), vS = std::sqrt(std::max(0.0f, A * det));
        float fs0 = std::ceil(u - 2 * invDet * uS), fs1 = std::floor(u + 2 * invDet * uS);
        float ft0 = std::ceil(v - 2 * invDet * vS), ft1 = std::floor(v + 2 * invDet * vS);
        if (!(fs1 - fs0 <= 4096.0f && ft1 - ft0 <= 4096.0f && std::fabs(fs0) < 1e9f && std::fabs(ft0) < 1e9f)) {
            // non-finite footprint (undefined in the reference): the product's guard, a bilinear read
            bilinear(t, l, (u + 0.5f) / t.mips[l].w, (v + 0.5f) / t.mips[l].h, out);
            return;
        }
        int s0 = (int)fs0, s1 = (int)fs1, t0 = (int)ft0, t1 = (int)ft1;
        float sum[4] = {0, 0, 0, 0}, wsum = 0, tx[4];
        for (int it = t0; it <= t1; ++it) {
            float tt = it - v;
            for (int is = s0; is <= s1; ++is) {
                float ss = is - u;
                float r2 = A * ss * ss + B * ss * tt + C * tt * tt;
                if (r2 < 1) {
                    float wgt = g_tex_tables.ewa[std::min<int>(r2 * 128, 127)];
                    t.fetch(wrap_index(is, t.mips[l].w, wrap_u), wrap_index(it, t.mips[l].h, wrap_v), l, tx);
                    for (int k = 0; k < t.nc; ++k) sum[k] += wgt * tx[k];
                    wsum += wgt;
                }
            }
        }
        for (int k = 0; k < t.nc; ++k) out[k] = wsum > 0 ? sum[k] * (1.0f / wsum) : 0.0f;
    }
    void one_level(const Texture &t, int l, float u, float v, float maj[2], float mnr[2], float *out) const {
        if (kind == 2) bicubic(t, l, u, v, out);
        else if (kind == 3) ewa(t, l, u, v, maj, mnr, out);
        else bilinear(t, l, u, v, out);
    }
    // operator() of Nearest / Linear / Cubic / EWA samplers (texture.cpp:575-716); d = duvdxy rows
    void sample(const Texture &t, float u, float v, const float d[4], float *out) const {
        if (kind == 0) {
            int x = (int)std::floor(u * t.mips[0].w - 0.5f), y = (int)std::floor(v * t.mips[0].h - 0.5f);
            t.fetch(wrap_index(x, t.mips[0].w, wrap_u), wrap_index(y, t.mips[0].h, wrap_v), 0, out);
            return;
        }
        float maj[2] = {d[0], d[1]}, mnr[2] = {d[2], d[3]};
        float width;
        if (kind == 3) {
            float lmaj = std::sqrt(maj[0] * maj[0] + maj[1] * maj[1]), lmin = std::sqrt(mnr[0] * mnr[0] + mnr[1] * mnr[1]);
            if (lmaj < lmin) {
                std::swap(maj, mnr);
                std::swap(lmaj, lmin);
            }
            if (lmin * anisotropy < lmaj && lmin > 0) {
                float sc = lmaj / (lmin * anisotropy);
                mnr[0] *= sc, mnr[1] *= sc;
                lmin *= sc;
            }
            if (lmin == 0) {
                bilinear(t, 0, u, v, out);
                return;
            }
            width = lmin;
        } else {
            width = std::max(std::max(std::fabs(d[0]), std::fabs(d[1])), std::max(std::fabs(d[2]), std::fabs(d[3])));
        }
        float level = t.levels() - 1 + std::log2(std::max(width, 1e-8f));
        if (level < 0 || t.levels() == 1) {
            one_level(t, 0, u, v, maj, mnr, out);
        } else if (level >= t.levels() - 1) {
            t.fetch(0, 0, t.levels() - 1, out);
        } else {
            int il = (int)std::floor(level);
            float delta = level - il, o0[4], o1[4];
            one_level(t, il, u, v, maj, mnr, o0);
            one_level(t, il + 1, u, v, maj, mnr, o1);
            for (int k = 0; k < t.nc; ++k) out[k] = std::lerp(o0[k], o1[k], delta);
        }
    }
};

struct Field { // ConstantField / TextureField<N> (shader_field.h:26-86)
    int tex = -1;
    Sampler smp;
    bool flip_v = true;
    int swizzle[4] = {0, 1, 2, 3};
    float uv_scale[2] = {1, 1}, uv_offset[2] = {0, 0}, scale[4] = {1, 1, 1, 1}, value[4] = {0, 0, 0, 0};

    bool init(const ora_field &f, int n, const std::vector<Texture> &texs) {
        tex = f.texture;
        for (int k = 0; k < 4; ++k) value[k] = f.value[k];
        if (tex < 0) return true;
        if (tex >= (int)texs.size() || f.sampler < 0 || f.sampler > 3 || f.cubic_kernel < 0 || f.cubic_kernel > 2)
            return false;
        smp.kind = f.sampler;
        smp.wrap_u = f.wrap_u;
        smp.wrap_v = f.wrap_v;
        smp.anisotropy = std::clamp(f.anisotropy, 1.0f, 16.0f);
        // CubicSampler kernels (texture.h:92-116)
        static const float ka[3][4] = {{21, -36, 0, 16}, {3, -6, 0, 4}, {3, -5, 0, 2}};
        static const float kb[3][4] = {{-7, 36, -60, 32}, {-1, 6, -12, 8}, {-1, 5, -8, 4}};
        static const float kd[3] = {18, 6, 2};
        for (int j = 0; j < 4; ++j) {
            smp.ca[j] = ka[f.cubic_kernel][j] / kd[f.cubic_kernel];
            smp.cb[j] = kb[f.cubic_kernel][j] / kd[f.cubic_kernel];
        }
        flip_v = f.flip_v != 0;
        for (int k = 0; k < 4; ++k) {
            swizzle[k] = k < n ? f.swizzle[k] : -1;
            if (swizzle[k] >= texs[tex].nc) return false;
            scale[k] = f.scale[k];
        }
        uv_scale[0] = f.uv_scale[0], uv_scale[1] = f.uv_scale[1];
        uv_offset[0] = f.uv_offset[0], uv_offset[1] = f.uv_offset[1];
        return true;
    }
    // TextureField::operator() (shader_field.h:53-77)
    void eval(const std::vector<Texture> &texs, float u, float v, const float duv[4], int n, float *out) const {
        if (tex < 0) {
            for (int k = 0; k < n; ++k) out[k] = value[k];
            return;
        }
        float d[4] = {duv[0], duv[1], duv[2], duv[3]};
        if (flip_v) {
            v = 1.0f - v;
            d[1] *= -1.0f;
            d[3] *= -1.0f;
        }
        u = u * uv_scale[0] + uv_offset[0];
        v = v * uv_scale[1] + uv_offset[1];
        d[0] *= uv_scale[0];
        d[2] *= uv_scale[0];
        d[1] *= uv_scale[1];
        d[3] *= uv_scale[1];
        float s[4] = {0, 0, 0, 0};
        smp.sample(texs[tex], u, v, d, s);
        for (int k = 0; k < n; ++k) out[k] = (swizzle[k] >= 0 ? s[swizzle[k]] : 0.0f) * scale[k];
    }
};

struct TexMaterial {
    Field f[9]; // basecolor, roughness, metallic, ior, specular_r0_mul, specular_trans, diffuse_trans,
                // diffuse_trans_fwd, emission
    bool has_emission = false, has_normal = false, has_opacity = false;
    Field normal, opacity;
    bool normal_world = false, normal_range01 = true;
    float normal_strength = 1.0f;
    float normal_m[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1}; // map to_world (linear part)
};

struct Light {
    bool tex_emission = false; // mesh light whose emission field is a texture
    int type;  // 0 sky ambient, 1 directional, 2 point, 3 mesh triangle, 4 sky map
    V3 color;  // ambient / L / I / emission
    V3 vec;    // dir or pos
    uint32_t prim = 0; // global triangle index (mesh)
    float area = 0.0f;
};

struct Scene {
    std::vector<V3> verts, normals;
    std::vector<uint32_t> idx, tri_geom, geom_mat;
    std::vector<uint8_t> geom_flags;
    std::vector<ora_material> mats;
    std::vector<Node> nodes;
    std::vector<uint32_t> prims;
    std::vector<Light> lights;
    std::vector<int32_t> tri_light; // flat light index of each emissive triangle, -1 otherwise
    std::vector<float> cdf;         // DistribTable cdf (distrib.cpp:9-31)
    std::vector<float> powers;
    std::vector<uint32_t> sky_lights;
    SkyMapLight sky_map; // texels of the type-4 light, when there is one
    std::vector<Texture> textures;
    std::vector<TexMaterial> tmats;
    std::vector<int> tmat_of;        // per material: index into tmats or -1
    std::vector<float> uvs;          // 2 per vertex (empty: none)
    std::vector<uint8_t> tri_alpha;  // per triangle: material has an opacity map
    AABB bound;
    bool has_normals = false;

    V3 vtx(uint32_t t, int k) const { return verts[idx[3 * t + k]]; }
    bool has_uv(uint32_t t) const { return !uvs.empty() && !(geom_flags[tri_geom[t]] & 4); }
    const TexMaterial *tmat(uint32_t t) const {
        if (tmat_of.empty()) return nullptr;
        int k = tmat_of[geom_mat[tri_geom[t]]];
        return k >= 0 ? &tmats[k] : nullptr;
    }
    // rtcInterpolate0 of the texcoords at barycentric (b1, b2) (MeshGeometry::compute_hit_texcoord /
    // interpolate_texcoord, geometry.cpp:165-190): zero for a mesh without texcoords
    void texcoord(uint32_t t, float b1, float b2, float &u, float &v) const {
        u = v = 0.0f;
        if (!has_uv(t)) return;
        const float *a = &uvs[2 * idx[3 * t]], *b = &uvs[2 * idx[3 * t + 1]], *c = &uvs[2 * idx[3 * t + 2]];
        float w = 1.0f - b1 - b2;
        u = a[0] * w + b[0] * b1 + c[0] * b2;
        v = a[1] * w + b[1] * b1 + c[1] * b2;
    }
};

// OpacityMap::stochastic_test(uv, ro, rd) (opacity_map.cpp:20-31) inside the intersection filter
// (scene.cpp:175-204): hash_float(ro, rd) = low 32 bits of MurmurHash64A over the 24 bytes, * 2^-32.
bool opacity_pass(const Scene &s, uint32_t t, float b1, float b2, V3 ro, V3 rd) {
    const TexMaterial *m = s.tmat(t);
    float u, v;
    s.texcoord(t, b1, b2, u, v);
    const float zero[4] = {0, 0, 0, 0};
    float o[4];
    m->opacity.eval(s.textures, u, v, zero, 1, o);
    float op = clampv(o[0], 0.0f, 1.0f);
    if (op <= 0.0f) return false;
    if (op >= 1.0f) return true;
    float buf[6] = {ro.x, ro.y, ro.z, rd.x, rd.y, rd.z};
    float r = (float)(uint32_t)murmur64a((const unsigned char *)buf, 24, 0) * 0x1p-32f;
    return r < op;
}

uint32_t build_rec(Scene &s, std::vector<uint32_t> &refs, std::vector<AABB> &tb, std::vector<V3> &tc, uint32_t begin,
                   uint32_t end) {
    uint32_t ni = (uint32_t)s.nodes.size();
    s.nodes.emplace_back();
    AABB b, cb;
    for (uint32_t i = begin; i < end; ++i) {
        b.expand(tb[refs[i]]);
        cb.expand(tc[refs[i]]);
    }
    uint32_t n = end - begin;
    auto leaf = [&]() {
        Node &nd = s.nodes[ni];
        nd.box = b;
        nd.prim_offset = (uint32_t)s.prims.size();
        nd.prim_count = n;
        nd.right = 0;
        for (uint32_t i = begin; i < end; ++i) s.prims.push_back(refs[i]);
        return ni;
    };
    if (n <= 2) return leaf();
    V3 ext = cb.hi - cb.lo;
    int axis = (ext.x >= ext.y && ext.x >= ext.z) ? 0 : (ext.y >= ext.z ? 1 : 2);
    if (ext[axis] <= 0.0f) return leaf();
    constexpr int B = 16;
    AABB bb[B];
    uint32_t bc[B] = {};
    float lo = cb.lo[axis], inv = B / ext[axis];
    auto bin = [&](uint32_t r) { return std::min(B - 1, (int)((tc[r][axis] - lo) * inv)); };
    for (uint32_t i = begin; i < end; ++i) {
        int k = bin(refs[i]);
        bc[k]++;
        bb[k].expand(tb[refs[i]]);
    }
    float best = kInf;
    int best_k = -1;
    for (int k = 0; k < B - 1; ++k) {
        AABB l, r;
        uint32_t nl = 0, nr = 0;
        for (int j = 0; j <= k; ++j) {
            l.expand(bb[j]);
            nl += bc[j];
        }
        for (int j = k + 1; j < B; ++j) {
            r.expand(bb[j]);
            nr += bc[j];
        }
        if (nl == 0 || nr == 0) continue;
        float cost = 1.0f + (nl * l.area() + nr * r.area()) / b.area();
        if (cost < best) {
            best = cost;
            best_k = k;
        }
    }
    uint32_t mid;
    if (best_k < 0 || (n <= 4 && best >= (float)n)) {
        if (n <= 4) return leaf();
        std::sort(refs.begin() + begin, refs.begin() + end,
                  [&](uint32_t x, uint32_t y) { return tc[x][axis] < tc[y][axis]; });
        mid = begin + n / 2;
    } else {
        auto it = std::partition(refs.begin() + begin, refs.begin() + end,
                                 [&](uint32_t r) { return bin(r) <= best_k; });
        mid = (uint32_t)(it - refs.begin());
        if (mid == begin || mid == end) mid = begin + n / 2;
    }
    s.nodes[ni].box = b;
    s.nodes[ni].axis = (uint32_t)axis;
    uint32_t prim_start = (uint32_t)s.prims.size();
    build_rec(s, refs, tb, tc, begin, mid);
    uint32_t r = build_rec(s, refs, tb, tc, mid, end);
    s.nodes[ni].right = r;
    s.nodes[ni].prim_offset = prim_start;
    s.nodes[ni].prim_count = (uint32_t)s.prims.size() - prim_start;
    return ni;
}

struct RayQ {
    V3 o, d;
    float tmin, tmax;
    V3 inv;
    int neg[3];
    int kx, ky, kz;
    V3 shear;
    void setup() {
        // RayBoundHelper (cuda/aabb.cuh:155-169)
        for (int i = 0; i < 3; ++i) {
            inv[i] = 1.0f / d[i];
            neg[i] = d[i] < 0.0f || (d[i] == 0.0f && std::signbit(d[i]));
        }
        // RayTriHelper (cuda/ray_tri.cuh:8-37)
        V3 a(std::fabs(d.x), std::fabs(d.y), std::fabs(d.z));
        if (a.x >= a.y && a.x >= a.z) kz = 0;
        else if (a.y >= a.x && a.y >= a.z) kz = 1;
        else kz = 2;
        kx = (kz + 1) % 3;
        ky = (kx + 1) % 3;
        if (d[kz] < 0.0f) std::swap(kx, ky);
        shear = V3(d[kx], d[ky], 1.0f) / d[kz];
    }
};

bool isect_aabb(const RayQ &r, const AABB &b) { // cuda/aabb.cuh:172-251
    const V3 *bb[2] = {&b.lo, &b.hi};
    float tmin = (*bb[r.neg[0]]).x - r.o.x;
    if (tmin != 0.0f) tmin *= r.inv.x;
    float tmax = (*bb[1 - r.neg[0]]).x - r.o.x;
    if (tmax != 0.0f) tmax *= r.inv.x;
    float tymin = (*bb[r.neg[1]]).y - r.o.y;
    if (tymin != 0.0f) tymin *= r.inv.y;
    float tymax = (*bb[1 - r.neg[1]]).y - r.o.y;
    if (tymax != 0.0f) tymax *= r.inv.y;
    if (tmin > tymax || tymin > tmax) return false;
    if (tymin > tmin) tmin = tymin;
    if (tymax < tmax) tmax = tymax;
    float tzmin = (*bb[r.neg[2]]).z - r.o.z;
    if (tzmin != 0.0f) tzmin *= r.inv.z;
    float tzmax = (*bb[1 - r.neg[2]]).z - r.o.z;
    if (tzmax != 0.0f) tzmax *= r.inv.z;
    if (tmin > tzmax || tzmin > tmax) return false;
    if (tzmin > tmin) tmin = tzmin;
    if (tzmax < tmax) tmax = tzmax;
    return (tmin < r.tmax) && (tmax > r.tmin);
}

// Watertight ray/triangle test (cuda/ray_tri.cuh:46-97). coord = weights of (v0, v1, v2).
bool isect_tri(const RayQ &r, V3 v0, V3 v1, V3 v2, V3 &coord, float &t) {
    V3 A0 = v0 - r.o, B0 = v1 - r.o, C0 = v2 - r.o;
    V3 A(A0[r.kx], A0[r.ky], A0[r.kz]), B(B0[r.kx], B0[r.ky], B0[r.kz]), C(C0[r.kx], C0[r.ky], C0[r.kz]);
    V3 sx = V3(A.x, B.x, C.x) - V3(A.z, B.z, C.z) * r.shear.x;
    V3 sy = V3(A.y, B.y, C.y) - V3(A.z, B.z, C.z) * r.shear.y;
    V3 bc = cross(sy, sx);
    if (bc.x == 0.0f || bc.y == 0.0f || bc.z == 0.0f) {
        bc.x = (float)((double)sx[2] * (double)sy[1] - (double)sy[2] * (double)sx[1]);
        bc.y = (float)((double)sx[0] * (double)sy[2] - (double)sy[0] * (double)sx[2]);
        bc.z = (float)((double)sx[1] * (double)sy[0] - (double)sy[1] * (double)sx[0]);
    }
    if ((bc.x < 0.0f || bc.y < 0.0f || bc.z < 0.0f) && (bc.x > 0.0f || bc.y > 0.0f || bc.z > 0.0f)) return false;
    float det = bc.x + bc.y + bc.z;
    if (det == 0.0f) return false;
    V3 sz = V3(A.z, B.z, C.z) * r.shear.z;
    float T = dot(bc, sz);
    if (det < 0.0f && (T >= r.tmin * det || T <= r.tmax * det)) return false;
    if (det > 0.0f && (T <= r.tmin * det || T >= r.tmax * det)) return false;
    float inv_det = 1.0f / det;
    coord = bc * inv_det;
    t = T * inv_det;
    return true;
}

struct Hit {
    float t = -1.0f, u = 0.0f, v = 0.0f;
    int32_t prim = -1;
};

// Closest hit with the reference traversal order (cuda/sbvh.cu:512-560): near child first by the sign
// of the ray direction along the split axis, stack of pending nodes, strict t < tmax update.
Hit closest(const Scene &s, RayQ r) {
    Hit h;
    if (s.nodes.empty()) return h;
    uint32_t stack[128];
    int sp = 0;
    uint32_t cur = 0;
    while (true) {
        const Node &nd = s.nodes[cur];
        if (!isect_aabb(r, nd.box)) {
            if (sp == 0) break;
            cur = stack[--sp];
            continue;
        }
        if (nd.right == 0) {
            for (uint32_t i = 0; i < nd.prim_count; ++i) {
                uint32_t t = s.prims[nd.prim_offset + i];
                V3 coord;
                float th;
                if (isect_tri(r, s.vtx(t, 0), s.vtx(t, 1), s.vtx(t, 2), coord, th) && th < r.tmax) {
                    if (!s.tri_alpha.empty() && s.tri_alpha[t] && !opacity_pass(s, t, coord.y, coord.z, r.o, r.d))
                        continue;
                    r.tmax = th;
                    h.t = th;
                    h.u = coord.y;
                    h.v = coord.z;
                    h.prim = (int32_t)t;
                }
            }
            if (sp == 0) break;
            cur = stack[--sp];
        } else if (r.neg[nd.axis]) {
            stack[sp++] = cur + 1;
            cur = nd.right;
        } else {
            stack[sp++] = nd.right;
            cur = cur + 1;
        }
    }
    return h;
}

bool occluded(const Scene &s, RayQ r) { // cuda/sbvh.cu:562-603
    if (s.nodes.empty()) return false;
    uint32_t stack[128];
    int sp = 0;
    uint32_t cur = 0;
    while (true) {
        const Node &nd = s.nodes[cur];
        if (!isect_aabb(r, nd.box)) {
            if (sp == 0) break;
            cur = stack[--sp];
            continue;
        }
        if (nd.right == 0) {
            for (uint32_t i = 0; i < nd.prim_count; ++i) {
                uint32_t t = s.prims[nd.prim_offset + i];
                V3 coord;
                float th;
                if (isect_tri(r, s.vtx(t, 0), s.vtx(t, 1), s.vtx(t, 2), coord, th)) {
                    if (!s.tri_alpha.empty() && s.tri_alpha[t] && !opacity_pass(s, t, coord.y, coord.z, r.o, r.d))
                        continue;
                    return true;
                }
            }
            if (sp == 0) break;
            cur = stack[--sp];
        } else if (r.neg[nd.axis]) {
            stack[sp++] = cur + 1;
            cur = nd.right;
        } else {
            stack[sp++] = nd.right;
            cur = cur + 1;
        }
    }
    return false;
}

Hit brute(const Scene &s, RayQ r) {
    Hit h;
    uint32_t nt = (uint32_t)s.tri_geom.size();
    for (uint32_t t = 0; t < nt; ++t) {
        V3 coord;
        float th;
        if (isect_tri(r, s.vtx(t, 0), s.vtx(t, 1), s.vtx(t, 2), coord, th) && th < r.tmax) {
            r.tmax = th;
            h.t = th;
            h.u = coord.y;
            h.v = coord.z;
            h.prim = (int32_t)t;
        }
    }
    return h;
}

// DistribTable (distrib.cpp:9-31)
std::vector<float> make_cdf(const std::vector<float> &f) {
    uint32_t n = (uint32_t)f.size();
    std::vector<float> cdf(n + 1);
    cdf[0] = 0;
    float invn = 1.0f / (float)n;
    for (uint32_t i = 1; i < n + 1; ++i) cdf[i] = cdf[i - 1] + f[i - 1] * invn;
    float acc = cdf[n];
    if (acc == 0.0f) {
        for (uint32_t i = 0; i < n + 1; ++i) cdf[i] = (float)i / (float)n;
    } else {
        float invacc = 1.0f / acc;
        for (uint32_t i = 0; i < n + 1; ++i) cdf[i] *= invacc;
    }
    cdf[n] = 1.0f;
    return cdf;
}

uint32_t cdf_sample(const std::vector<float> &cdf, float u, float &prob) { // distrib.cpp:33-41
    u = clampv(u, 0.0f, std::nextafter(1.0f, 0.0f));
    auto it = std::upper_bound(cdf.begin(), cdf.end(), u);
    uint32_t index = (uint32_t)std::distance(cdf.begin(), std::prev(it));
    prob = cdf[index + 1] - cdf[index];
    return index;
}

float light_probability(const Scene &s, uint32_t i) { // light.cpp:680-683
    uint32_t n = (uint32_t)s.cdf.size() - 1;
    return ((s.cdf[i + 1] - s.cdf[i]) * (float)n) / (float)n;
}

// SummedAreaTable (sat.h:9-61)
struct SAT {
    std::vector<double> table;
    int nx, ny;
    SAT(const std::vector<float> &v, int nx_, int ny_) : table(v.size()), nx(nx_), ny(ny_) {
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                size_t i = (size_t)y * nx + x;
                if (x == 0 && y == 0) table[i] = v[i];
                else if (y == 0) table[i] = v[i] + table[i - 1];
                else if (x == 0) table[i] = v[i] + table[i - nx];
                else table[i] = v[i] + table[i - 1] + table[i - nx] - table[i - nx - 1];
            }
    }
    double at(int x, int y) const { return (x == 0 || y == 0) ? 0.0 : table[(size_t)std::min(y - 1, ny - 1) * nx + std::min(x - 1, nx - 1)]; }
    double lookup(float px, float py) const { // lerp_helper, TickMode::Boundary, WrapMode::Clamp, over (nx + 1, ny + 1)
        float xs = px * nx, ys = py * ny;
        int x0 = (int)std::floor(xs), y0 = (int)std::floor(ys);
        float wx = xs - x0, wy = ys - y0;
        int x1 = std::clamp(x0 + 1, 0, nx), y1 = std::clamp(y0 + 1, 0, ny);
        x0 = std::clamp(x0, 0, nx), y0 = std::clamp(y0, 0, ny);
        return at(x0, y0) * (1.0 - wx) * (1.0 - wy) + at(x1, y0) * (wx) * (1.0 - wy) + at(x0, y1) * (1.0 - wx) * (wy) +
               at(x1, y1) * (wx) * (wy);
    }
    float sum(const float lo[2], const float hi[2]) const {
        double v = (lookup(hi[0], hi[1]) - lookup(lo[0], hi[1])) + (lookup(lo[0], lo[1]) - lookup(hi[0], lo[1]));
        return std::max((float)v, 0.0f);
    }
};

void mesh_importance(const Scene &s, uint32_t t0, uint32_t t1, V3 Le, const TexMaterial *tm, std::vector<float> &imp,
                     std::vector<float> &area) {
    const bool em_tex = tm && tm->f[8].tex >= 0, op = tm && tm->has_opacity, op_tex = op && tm->opacity.tex >= 0;
    bool fallback = !em_tex && !op_tex;
    int res[2] = {1, 1};
    const Field *ref = nullptr;
    for (const Field *f : {em_tex ? &tm->f[8] : nullptr, op_tex ? &tm->opacity : nullptr}) {
        if (!f) continue;
        res[0] = std::max(res[0], s.textures[f->tex].width);
        res[1] = std::max(res[1], s.textures[f->tex].height);
        if (!ref) ref = f;
        else if (ref->smp.wrap_u != f->smp.wrap_u || ref->smp.wrap_v != f->smp.wrap_v || ref->flip_v != f->flip_v ||
                 ref->uv_scale[0] != f->uv_scale[0] || ref->uv_scale[1] != f->uv_scale[1] ||
                 ref->uv_offset[0] != f->uv_offset[0] || ref->uv_offset[1] != f->uv_offset[1])
            fallback = true;
    }
    std::vector<float> iuv((size_t)res[0] * res[1]);
    float total = 0.0f;
    const float zero[4] = {0, 0, 0, 0};
    for (int y = 0; y < res[1]; ++y)
        for (int x = 0; x < res[0]; ++x) {
            float u = (x + 0.5f) / (float)res[0], v = (y + 0.5f) / (float)res[1];
            V3 e = Le;
            if (tm) {
                float c[4];
                tm->f[8].eval(s.textures, u, v, zero, 3, c);
                e = V3(c[0], c[1], c[2]);
            }
            float val = luminance(e);
            if (op) {
                float o[4];
                tm->opacity.eval(s.textures, u, v, zero, 1, o);
                val *= o[0];
            }
            iuv[(size_t)y * res[0] + x] = val;
            total += val;
        }
    float avg = total / (float)iuv.size();
    for (uint32_t i = 0; i < t1 - t0; ++i) {
        uint32_t t = t0 + i;
        area[i] = 0.5f * std::sqrt(sqnorm(cross(s.vtx(t, 1) - s.vtx(t, 0), s.vtx(t, 2) - s.vtx(t, 1))));
        if (fallback) imp[i] = kTwoPi * area[i] * avg;
    }
    if (fallback) return;
    SAT sat(iuv, res[0], res[1]);
    const float before_one = std::nextafter(1.0f, 0.0f);
    for (uint32_t i = 0; i < t1 - t0; ++i) {
        uint32_t t = t0 + i;
        float lo[2] = {kInf, kInf}, hi[2] = {-kInf, -kInf};
        for (int k = 0; k < 3; ++k) {
            float tc[2] = {0.0f, 0.0f};
            if (s.has_uv(t)) tc[0] = s.uvs[2 * s.idx[3 * t + k]], tc[1] = s.uvs[2 * s.idx[3 * t + k] + 1];
            if (ref->flip_v) tc[1] = 1.0f - tc[1];
            for (int a = 0; a < 2; ++a) {
                float c = tc[a] * ref->uv_scale[a] + ref->uv_offset[a];
                lo[a] = std::min(lo[a], c);
                hi[a] = std::max(hi[a], c);
            }
        }
        float value = 0.0f, covered = 0.0f;
        for (int ty = (int)std::floor(lo[1]); ty < (int)std::ceil(hi[1]); ++ty)
            for (int tx = (int)std::floor(lo[0]); tx < (int)std::ceil(hi[0]); ++tx) {
                float blo[2] = {std::max(lo[0], (float)tx), std::max(lo[1], (float)ty)};
                float bhi[2] = {std::min(hi[0], tx + before_one), std::min(hi[1], ty + before_one)};
                if (blo[0] > bhi[0] || blo[1] > bhi[1]) continue; // empty intersection: zero area
                float b_area = (bhi[0] - blo[0]) * (bhi[1] - blo[1]);
                if (b_area == 0.0f) continue;
                for (int a = 0; a < 2; ++a) {
                    int mode = a == 0 ? ref->smp.wrap_u : ref->smp.wrap_v;
                    if (mode == 0) {
                        auto m1 = [](float x) { float r = x - std::floor(x / 1.0f) * 1.0f; return r < 0 ? r + 1.0f : r; };
                        blo[a] = m1(blo[a]);
                        bhi[a] = m1(bhi[a]);
                    } else {
                        blo[a] = clampv(blo[a], 0.0f, 1.0f);
                        bhi[a] = clampv(bhi[a], 0.0f, 1.0f);
                    }
                    if (blo[a] == 0.0f && bhi[a] == 0.0f) bhi[a] = 1.0f / (res[a] + 1);
                    if (blo[a] == 1.0f && bhi[a] == 1.0f) blo[a] = 1.0f - 1.0f / (res[a] + 1);
                }
                bool empty = blo[0] > bhi[0] || blo[1] > bhi[1];
                float ex = empty ? 0.0f : bhi[0] - blo[0], ey = empty ? 0.0f : bhi[1] - blo[1];
                float mean = sat.sum(blo, bhi) / ((ex * (float)res[0]) * (ey * (float)res[1]));
                value += mean * b_area;
                covered += b_area;
            }
        if (covered > 0.0f) value /= covered;
        imp[i] = value * (kTwoPi * area[i]);
    }
}

Scene *create(const ora_scene_desc *d) {
    auto s = std::make_unique<Scene>();
    s->verts.resize(d->n_vertices);
    for (uint32_t i = 0; i < d->n_vertices; ++i)
        s->verts[i] = V3(d->vertices[3 * i], d->vertices[3 * i + 1], d->vertices[3 * i + 2]);
    if (d->normals) {
        s->has_normals = true;
        s->normals.resize(d->n_vertices);
        for (uint32_t i = 0; i < d->n_vertices; ++i)
            s->normals[i] = V3(d->normals[3 * i], d->normals[3 * i + 1], d->normals[3 * i + 2]);
    }
    s->idx.assign(d->indices, d->indices + 3 * (size_t)d->n_tris);
    s->tri_geom.assign(d->tri_geom, d->tri_geom + d->n_tris);
    s->geom_mat.assign(d->geom_material, d->geom_material + d->n_geoms);
    s->geom_flags.assign(d->n_geoms, 0);
    if (d->geom_flags) s->geom_flags.assign(d->geom_flags, d->geom_flags + d->n_geoms);
    s->mats.assign(d->materials, d->materials + d->n_materials);
    for (uint32_t i = 0; i < d->n_textures; ++i) {
        Texture t;
        if (!t.init(d->textures[i])) return nullptr;
        s->textures.push_back(std::move(t));
    }
    if (d->texcoords) s->uvs.assign(d->texcoords, d->texcoords + 2 * (size_t)d->n_vertices);
    if (d->material_fields) {
        s->tmat_of.assign(d->n_materials, -1);
        for (uint32_t m = 0; m < d->n_materials; ++m) {
            const ora_material_fields *f = d->material_fields[m];
            if (!f) continue;
            TexMaterial tm;
            const ora_field *fs[9] = {&f->basecolor, &f->roughness, &f->metallic, &f->ior, &f->specular_r0_mul,
                                      &f->specular_trans, &f->diffuse_trans, &f->diffuse_trans_fwd, &f->emission};
            for (int k = 0; k < 9; ++k)
                if (!tm.f[k].init(*fs[k], (k == 0 || k == 8) ? 3 : 1, s->textures)) return nullptr;
            tm.has_emission = f->has_emission != 0;
            tm.has_normal = f->has_normal_map != 0;
            if (tm.has_normal) {
                if (!tm.normal.init(f->normal_map, 3, s->textures)) return nullptr;
                tm.normal_world = f->normal_space == 1;
                tm.normal_range01 = f->normal_range01 != 0;
                tm.normal_strength = f->normal_strength;
                std::memcpy(tm.normal_m, f->normal_to_world, sizeof(tm.normal_m));
            }
            tm.has_opacity = f->has_opacity_map != 0;
            if (tm.has_opacity && !tm.opacity.init(f->opacity, 1, s->textures)) return nullptr;
            // the constant record of the material: emission flag and constant emission for the light table
            ora_material &om = s->mats[m];
            om = ora_material{};
            om.has_emission = tm.has_emission;
            for (int k = 0; k < 3; ++k) om.emission[k] = f->emission.value[k];
            s->tmat_of[m] = (int)s->tmats.size();
            s->tmats.push_back(tm);
        }
    }

    uint32_t nt = d->n_tris;
    bool any_alpha = false;
    for (const TexMaterial &tm : s->tmats) any_alpha |= tm.has_opacity;
    if (any_alpha) {
        s->tri_alpha.assign(nt, 0);
        for (uint32_t t = 0; t < nt; ++t) s->tri_alpha[t] = s->tmat(t) && s->tmat(t)->has_opacity;
    }
    std::vector<AABB> tb(nt);
    std::vector<V3> tc(nt);
    std::vector<uint32_t> refs(nt);
    for (uint32_t t = 0; t < nt; ++t) {
        for (int k = 0; k < 3; ++k) tb[t].expand(s->vtx(t, k));
        tc[t] = tb[t].center();
        refs[t] = t;
        s->bound.expand(tb[t]);
    }
    if (nt > 0) {
        s->nodes.reserve(2 * nt);
        build_rec(*s, refs, tb, tc, 0, nt);
    }

    // Lights: non-mesh lights first, then mesh lights per geometry (small_pt.cpp:187-205, light.cpp:636-677)
    V3 ext = s->bound.hi - s->bound.lo;
    float scene_radius = 0.5f * std::sqrt(sqnorm(ext));
    for (uint32_t i = 0; i < d->n_lights; ++i) {
        const ora_light &L = d->lights[i];
        Light l;
        l.type = L.type;
        l.color = V3(L.color[0], L.color[1], L.color[2]);
        l.vec = V3(L.vec[0], L.vec[1], L.vec[2]);
        V3 phi;
        if (L.type == 0) {
            // SkyLight::power (light.cpp:136-158) for a 1x1 ambient map
            float theta = 0.5f / 1.0f * kPi;
            float sin_theta = std::sin(theta);
            phi = l.color * sin_theta;
            phi = phi * ((kPi / 1.0f) * (kTwoPi / 1.0f));
            phi = phi * (kPi * sqr(scene_radius));
            s->sky_lights.push_back(i);
        } else if (L.type == 3) {
            if (!d->sky_map || !s->sky_map.texels.empty()) return nullptr; // one map, and it must be given
            s->sky_map.init(*d->sky_map);
            l.type = 4;
            phi = s->sky_map.power(scene_radius);
            s->sky_lights.push_back(i);
        } else if (L.type == 1) {
            l.vec = normalized(l.vec);
            phi = l.color * (kPi * sqr(scene_radius)); // light.cpp:216-220
        } else {
            phi = l.color * (4.0f * kPi); // light.cpp:235
        }
        s->lights.push_back(l);
        s->powers.push_back((phi.x + phi.y + phi.z) / 3.0f);
    }
    s->tri_light.assign(nt, -1);
    uint32_t g = 0;
    for (uint32_t t0 = 0; t0 < nt; g++) {
        uint32_t t1 = t0;
        while (t1 < nt && s->tri_geom[t1] == s->tri_geom[t0]) ++t1;
        const ora_material &m = s->mats[s->geom_mat[s->tri_geom[t0]]];
        if (m.has_emission) {
            V3 Le(m.emission[0], m.emission[1], m.emission[2]);
            // compute_mesh_light_importance (light.cpp:307-523) + swap compaction (:536-545)
            uint32_t N = t1 - t0;
            std::vector<uint32_t> ids(N);
            std::vector<float> imp(N), area(N);
            mesh_importance(*s, t0, t1, Le, s->tmat(t0), imp, area);
            for (uint32_t i = 0; i < N; ++i) ids[i] = i;
            for (uint32_t i = 0; i < N;) {
                if (imp[i] == 0.0f) {
                    std::swap(ids[N - 1], ids[i]);
                    std::swap(imp[N - 1], imp[i]);
                    std::swap(area[N - 1], area[i]);
                    --N;
                } else {
                    ++i;
                }
            }
            for (uint32_t i = 0; i < N; ++i) {
                Light l;
                l.type = 3;
                l.color = Le;
                l.tex_emission = s->tmat(t0) && s->tmat(t0)->f[8].tex >= 0;
                l.prim = t0 + ids[i];
                l.area = area[i];
                s->tri_light[l.prim] = (int32_t)s->lights.size();
                s->lights.push_back(l);
                s->powers.push_back((imp[i] + imp[i] + imp[i]) / 3.0f);
            }
        }
        t0 = t1;
    }
    if (!s->lights.empty()) s->cdf = make_cdf(s->powers);
    return s.release();
}

// Hit record reconstruction (geometry.cpp:68-163 with an identity instance transform).
struct SurfHit {
    V3 p;
    V3 t, b, n;       // geometric frame
    V3 st, sb, sn;    // shading frame
    float thit;
    uint32_t prim;
    float u = 0, v = 0;             // Intersection::uv
    V3 dpdu, dpdv;
    float duv[4] = {0, 0, 0, 0};    // dudx, dvdx, dudy, dvdy
};

SurfHit make_hit(const Scene &s, V3 ray_d, const Hit &h) {
    SurfHit it;
    it.thit = h.t;
    it.prim = (uint32_t)h.prim;
    uint32_t tr = it.prim;
    V3 v0 = s.vtx(tr, 0), v1 = s.vtx(tr, 1), v2 = s.vtx(tr, 2);
    V3 rd = normalized(ray_d);
    V3 wo = -rd;
    V3 ng = normalized(cross(v1 - v0, v2 - v0)); // Embree Ng = cross(e2, e1), e1 = v0 - v1, e2 = v2 - v0
    float w = 1.0f - h.u - h.v;
    it.p = v0 * w + v1 * h.u + v2 * h.v;
    V3 dpdu = v1 - v0, dpdv = v2 - v0; // rtcInterpolate1 derivatives
    it.u = h.u;
    it.v = h.v;
    if (s.has_uv(tr)) { // geometry.cpp:94-116: uv from texcoords, dpdu / dpdv from the uv parameterisation
        s.texcoord(tr, h.u, h.v, it.u, it.v);
        const float *t0 = &s.uvs[2 * s.idx[3 * tr]], *t1 = &s.uvs[2 * s.idx[3 * tr + 1]], *t2 = &s.uvs[2 * s.idx[3 * tr + 2]];
        float duv02[2] = {t0[0] - t2[0], t0[1] - t2[1]}, duv12[2] = {t1[0] - t2[0], t1[1] - t2[1]};
        V3 dp02 = v0 - v2, dp12 = v1 - v2;
        float det = duv02[0] * duv12[1] - duv02[1] * duv12[0];
        if (!(std::fabs(det) < 1e-8f)) {
            float inv = 1.0f / det;
            dpdu = (dp02 * duv12[1] - dp12 * duv02[1]) * inv;
            dpdv = (dp02 * -duv12[0] + dp12 * duv02[0]) * inv;
        }
    }
    V3 b = normalized(cross(ng, dpdu));
    V3 t = normalized(cross(b, ng));
    const uint8_t flags = s.geom_flags[s.tri_geom[tr]];
    if (!(flags & 1) && dot(wo, ng) < 0.0f) { // two-sided (geometry.cpp:126-132)
        ng = -ng;
        dpdu = -dpdu;
        t = -t;
    }
    it.t = t;
    it.b = b;
    it.n = ng;
    it.st = t;
    it.sb = b;
    it.sn = ng;
    if (s.has_normals && !(flags & 2)) { // use_smooth_normal (geometry.cpp:135)
        const V3 &n0 = s.normals[s.idx[3 * tr]], &n1 = s.normals[s.idx[3 * tr + 1]], &n2 = s.normals[s.idx[3 * tr + 2]];
        V3 vn = normalized(n0 * w + n1 * h.u + n2 * h.v);
        if (all_finite(vn)) {
            if (dot(vn, ng) < 0.0f) vn = -vn;
            V3 vb = normalized(cross(vn, dpdu));
            V3 vt = normalized(cross(vb, vn));
            it.st = vt;
            it.sb = vb;
            it.sn = vn;
        }
    }
    it.dpdu = dpdu;
    it.dpdv = dpdv;
    return it;
}

float comp3(V3 v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }

// Intersection::compute_uv_partials (ray.cpp:6-46) for a ray with differentials from origin o.
void compute_uv_partials(SurfHit &it, V3 o, V3 rxd, V3 ryd) {
    auto solve = [](float a00, float a01, float a10, float a11, float b0, float b1, float &x0, float &x1) {
        float det = a00 * a11 - a01 * a10; // solve_linear_system_2x2 (maths.h:559-569)
        if (std::fabs(det) < 1e-10f) return false;
        x0 = (a11 * b0 - a01 * b1) / det;
        x1 = (a00 * b1 - a10 * b0) / det;
        return !(std::isnan(x0) || std::isnan(x1));
    };
    float d = dot(it.n, it.p);
    float tx = -(dot(it.n, o) - d) / dot(it.n, rxd);
    float ty = -(dot(it.n, o) - d) / dot(it.n, ryd);
    for (float &x : it.duv) x = 0.0f;
    if (std::isinf(tx) || std::isnan(tx) || std::isinf(ty) || std::isnan(ty)) return;
    V3 px = o + rxd * tx, py = o + ryd * ty;
    int d0 = 0, d1 = 1;
    if (std::fabs(it.n.x) > std::fabs(it.n.y) && std::fabs(it.n.x) > std::fabs(it.n.z)) d0 = 1, d1 = 2;
    else if (std::fabs(it.n.y) > std::fabs(it.n.z)) d0 = 0, d1 = 2;
    float a00 = comp3(it.dpdu, d0), a01 = comp3(it.dpdv, d0), a10 = comp3(it.dpdu, d1), a11 = comp3(it.dpdv, d1);
    if (!solve(a00, a01, a10, a11, comp3(px, d0) - comp3(it.p, d0), comp3(px, d1) - comp3(it.p, d1), it.duv[0], it.duv[1]))
        it.duv[0] = it.duv[1] = 0;
    if (!solve(a00, a01, a10, a11, comp3(py, d0) - comp3(it.p, d0), comp3(py, d1) - comp3(it.p, d1), it.duv[2], it.duv[3]))
        it.duv[2] = it.duv[3] = 0;
}

// NormalMap::sample / apply (normal_map.cpp:8-40)
void apply_normal_map(const Scene &s, const TexMaterial &m, SurfHit &it) {
    const float zero[4] = {0, 0, 0, 0};
    float c[4];
    m.normal.eval(s.textures, it.u, it.v, zero, 3, c);
    V3 n(c[0], c[1], c[2]);
    if (m.normal_range01) n = V3(n.x * 2.0f - 1.0f, n.y * 2.0f - 1.0f, n.z * 2.0f - 1.0f);
    n = normalized(n);
    if (!m.normal_world && m.normal_strength < 1.0f) // lerp((0, 0, 1), n, strength)
        n = normalized(V3(0, 0, 1) * (1.0f - m.normal_strength) + n * m.normal_strength);
    V3 np;
    if (m.normal_world) { // Transform::normal: inverse transpose of the map's to_world
        float inv[9];
        SkyMapLight::invert3(m.normal_m, inv);
        np = normalized(V3(inv[0] * n.x + inv[3] * n.y + inv[6] * n.z, inv[1] * n.x + inv[4] * n.y + inv[7] * n.z,
                           inv[2] * n.x + inv[5] * n.y + inv[8] * n.z));
    } else {
        np = it.st * n.x + it.sb * n.y + it.sn * n.z;
    }
    float bump = -n.x / n.z;
    V3 tp = normalized(it.st + it.sn * bump);
    V3 bp = normalized(cross(np, tp));
    tp = normalized(cross(bp, np));
    it.st = tp;
    it.sb = bp;
    it.sn = np;
}

// PrincipledBSDF::eval_closure (principled_bsdf.cpp:246-275) over shader fields
Closure fields_closure(const Scene &s, const TexMaterial &m, const SurfHit &it) {
    ora_material c{};
    float a[4];
    m.f[0].eval(s.textures, it.u, it.v, it.duv, 3, a);
    c.basecolor[0] = a[0], c.basecolor[1] = a[1], c.basecolor[2] = a[2];
    float *dst[7] = {&c.roughness, &c.metallic, &c.ior, &c.specular_r0_mul, &c.specular_trans, &c.diffuse_trans,
                     &c.diffuse_trans_fwd};
    for (int k = 0; k < 7; ++k) {
        m.f[1 + k].eval(s.textures, it.u, it.v, it.duv, 1, a);
        *dst[k] = a[0];
    }
    return make_closure(c);
}

inline V3 to_local(V3 t, V3 b, V3 n, V3 w) { return V3(dot(t, w), dot(b, w), dot(n, w)); }
V3 sh_to_local(const SurfHit &it, V3 w) { // ray.h:108-117
    bool hw = dot(it.n, w) >= 0.0f;
    V3 l = to_local(it.st, it.sb, it.sn, w);
    bool hl = l.z >= 0.0f;
    if (hw != hl) l.z *= -1.0f;
    return l;
}
V3 sh_to_world(const SurfHit &it, V3 l) { // ray.h:119-128
    bool hl = l.z >= 0.0f;
    V3 w = it.st * l.x + it.sb * l.y + it.sn * l.z;
    bool hw = dot(it.n, w) >= 0.0f;
    if (hl != hw) w = w - it.n * (2.0f * dot(it.n, w));
    return w;
}

inline float int_as_float(int i) {
    float f;
    std::memcpy(&f, &i, 4);
    return f;
}
inline int float_as_int(float f) {
    int i;
    std::memcpy(&i, &f, 4);
    return i;
}
V3 offset_ray(V3 p, V3 n) { // ray.h:66-79
    constexpr float origin = 1.0f / 32.0f, float_scale = 1.0f / 65536.0f, int_scale = 256.0f;
    int ox = (int)(int_scale * n.x), oy = (int)(int_scale * n.y), oz = (int)(int_scale * n.z);
    V3 pi(int_as_float(float_as_int(p.x) + ((p.x < 0) ? -ox : ox)),
          int_as_float(float_as_int(p.y) + ((p.y < 0) ? -oy : oy)),
          int_as_float(float_as_int(p.z) + ((p.z < 0) ? -oz : oz)));
    return V3(std::fabs(p.x) < origin ? p.x + float_scale * n.x : pi.x,
              std::fabs(p.y) < origin ? p.y + float_scale * n.y : pi.y,
              std::fabs(p.z) < origin ? p.z + float_scale * n.z : pi.z);
}
V3 spawn_origin(V3 p, V3 dir, V3 ng) { return offset_ray(p, dot(dir, ng) > 0.0f ? ng : -ng); } // ray.h:87-96

// Light sampling (light.cpp:59-134, 567-628)
float sky_pdf(V3 wi) {
    float theta = std::acos(clampv(wi.y, -1.0f, 1.0f)); // to_spherical_yup (maths.h:477-484)
    float sin_theta = std::sin(theta);
    if (sin_theta == 0.0f) return 0.0f;
    float pdf = 1.0f; // 1x1 DistribTable2D
    pdf /= (2.0f * kPi * kPi * sin_theta);
    return pdf;
}

// Returns Le/pdf (light.h:27-29 convention).
V3 light_sample(const Scene &s, const Light &l, V3 p, float u0, float u1, V3 &wi, float &wi_dist, float &pdf) {
    if (l.type == 0) { // SkyLight::sample (light.cpp:88-111)
        float uu0 = clampv(u0, 0.0f, std::nextafter(1.0f, 0.0f));
        float uu1 = clampv(u1, 0.0f, std::nextafter(1.0f, 0.0f));
        pdf = 1.0f;
        float phi = uu0 * 2.0f * kPi;
        float theta = uu1 * kPi;
        float ct = std::cos(theta), st = std::sin(theta), cp = std::cos(phi), sp = std::sin(phi);
        V3 wl(cp * st, ct, sp * st); // to_cartesian_yup
        float sin_theta = std::sin(theta);
        pdf /= (2.0f * kPi * kPi * sin_theta);
        if (sin_theta == 0.0f) {
            pdf = 0.0f;
            return V3();
        }
        wi = wl;
        wi_dist = kInf;
        return l.color / pdf;
    }
    if (l.type == 4) {
        wi_dist = kInf;
        return s.sky_map.sample(u0, u1, wi, pdf);
    }
    if (l.type == 1) { // DirectionalLight::sample (light.cpp:207-213)
        wi = l.vec;
        wi_dist = kInf;
        pdf = 1.0f;
        return l.color;
    }
    if (l.type == 2) { // PointLight::sample (light.cpp:222-233)
        wi = l.vec - p;
        float l2 = sqnorm(wi);
        if (l2 == 0.0f) return V3();
        float len = std::sqrt(l2);
        wi = wi / len;
        wi_dist = len;
        pdf = 1.0f;
        return l.color / l2;
    }
    // MeshTriLight::sample (light.cpp:567-605)
    float b0, b1;
    if (u0 < u1) {
        b0 = u0 / 2;
        b1 = u1 - b0;
    } else {
        b1 = u1 / 2;
        b0 = u0 - b1;
    }
    V3 v0 = s.vtx(l.prim, 0), v1 = s.vtx(l.prim, 1), v2 = s.vtx(l.prim, 2);
    V3 pos = v0 * (1.0f - b0 - b1) + v1 * b0 + v2 * b1; // rtcInterpolate0 at (u, v) = (b0, b1)
    V3 n = normalized(cross(v1 - v0, v2 - v1));
    wi = pos - p;
    float d2 = sqnorm(wi);
    if (d2 == 0.0f) {
        wi_dist = 0.0f;
        pdf = 0.0f;
        return V3();
    }
    wi_dist = std::sqrt(d2);
    wi = wi / wi_dist;
    pdf = 1.0f / l.area;
    float absdot = std::fabs(dot(n, wi));
    if (absdot == 0.0f) {
        wi_dist = 0.0f;
        pdf = 0.0f;
        return V3();
    }
    pdf *= (d2 / absdot);
    wi_dist = std::max(0.0f, wi_dist - 1e-4f);
    if (l.tex_emission) { // eval(it) with it.uv = interpolate_texcoord(prim, bary) (light.cpp:576,603-604)
        float tu, tv, e[4];
        s.texcoord(l.prim, b0, b1, tu, tv);
        const float zero[4] = {0, 0, 0, 0};
        s.tmat(l.prim)->f[8].eval(s.textures, tu, tv, zero, 3, e);
        return V3(e[0], e[1], e[2]) / pdf;
    }
    return l.color / pdf;
}

float mesh_light_pdf(const Scene &s, const Light &l, V3 wi, float wi_dist) { // light.cpp:607-628
    if (wi_dist == 0.0f) return 0.0f;
    V3 v0 = s.vtx(l.prim, 0), v1 = s.vtx(l.prim, 1), v2 = s.vtx(l.prim, 2);
    V3 n = normalized(cross(v1 - v0, v2 - v1));
    float pdf = 1.0f / l.area;
    float absdot = std::fabs(dot(n, wi));
    if (absdot == 0.0f) return 0.0f;
    pdf *= (sqr(wi_dist) / absdot);
    return pdf;
}

struct PathOut {
    bool hit;
    V3 L;
};

// SmallPT::trace (small_pt.cpp:72-178) with Material::sample_with_nee (material.cpp:101-167) and
// next_event_estimate (nee.cpp:11-58) inlined. mis_mode 1 adds the light selection probability to the
// NEE weight so the two strategies' weights sum to one.
PathOut trace(const Scene &s, const ora_params &p, V3 ro, V3 rd, ZSobol &sob, V3 rxd = V3(), V3 ryd = V3()) {
    V3 L, beta(1, 1, 1);
    float pdf_wi_bsdf = 0.0f;
    for (int bounce = 0; bounce < p.bounces + 1; ++bounce) {
        RayQ q;
        q.o = ro;
        q.d = rd;
        q.tmin = 0.0f;
        q.tmax = kInf;
        q.setup();
        Hit h = closest(s, q);
        if (h.prim < 0) {
            for (uint32_t li : s.sky_lights) {
                const Light &l = s.lights[li];
                V3 wi = normalized(rd);
                V3 Le = l.type == 4 ? s.sky_map.eval(wi) : l.color; // a 1x1 ambient map is constant
                if (bounce == 0) {
                    if (p.include_background) L = L + beta * Le;
                } else {
                    float pr = light_probability(s, li);
                    float pdf_l = l.type == 4 ? s.sky_map.pdf(wi) : sky_pdf(wi);
                    float w = power_heur(pdf_wi_bsdf, pdf_l * pr);
                    V3 bu = Le * w;
                    if (bounce == 1 && p.clamp_direct > 0.0f) bu = vmin(bu, p.clamp_direct);
                    else if (bounce >= 2 && p.clamp_indirect > 0.0f) bu = vmin(bu, p.clamp_indirect);
                    L = L + beta * bu;
                }
            }
            if (bounce == 0 && !p.include_background) return {false, V3()};
            break;
        }
        const ora_material &m = s.mats[s.geom_mat[s.tri_geom[h.prim]]];
        SurfHit it = make_hit(s, rd, h);
        const TexMaterial *tm = s.tmat((uint32_t)h.prim);
        if (bounce == 0 && !is_zero(rxd)) compute_uv_partials(it, ro, rxd, ryd); // camera rays only
        if (tm && tm->has_normal) apply_normal_map(s, *tm, it); // Scene::intersect1 (scene.cpp:228-233)
        if (m.has_emission) {
            int32_t li = s.tri_light[h.prim];
            if (li >= 0) {
                const Light &l = s.lights[li];
                float pr = light_probability(s, (uint32_t)li);
                V3 wi = normalized(rd);
                V3 Le = l.color;
                if (l.tex_emission) { // MeshTriLight::eval (light.cpp:560-565) at the hit
                    float e[4];
                    tm->f[8].eval(s.textures, it.u, it.v, it.duv, 3, e);
                    Le = V3(e[0], e[1], e[2]);
                }
                if (bounce == 0) {
                    L = L + beta * Le;
                } else if (pr > 0.0f) {
                    float pdf_l = mesh_light_pdf(s, l, wi, it.thit);
                    float w = power_heur(pdf_wi_bsdf, pdf_l * pr);
                    V3 bu = Le * w;
                    if (bounce == 1 && p.clamp_direct > 0.0f) bu = vmin(bu, p.clamp_direct);
                    else if (bounce >= 2 && p.clamp_indirect > 0.0f) bu = vmin(bu, p.clamp_indirect);
                    L = L + beta * bu;
                }
            }
        }
        if (bounce == p.bounces) break;
        V3 wo = -normalized(rd);
        Closure c = tm ? fields_closure(s, *tm, it) : make_closure(m);

        // NEE (nee.cpp:11-58)
        V3 Ld;
        if (!s.lights.empty()) {
            float pr_light;
            uint32_t li = cdf_sample(s.cdf, sob.next(), pr_light);
            const Light &l = s.lights[li];
            bool delta_light = l.type == 1 || l.type == 2;
            float u0, u1;
            sob.next2d(u0, u1);
            V3 wi;
            float wi_dist = 0.0f, pdf_l = 0.0f;
            V3 Le_beta = light_sample(s, l, it.p, u0, u1, wi, wi_dist, pdf_l);
            if (!is_zero(Le_beta) && pdf_l > 0.0f) {
                V3 wol = sh_to_local(it, wo), wil = sh_to_local(it, wi);
                V3 f;
                float pdf_b;
                bsdf_eval_and_pdf(wol, wil, c, f, pdf_b);
                if (!is_zero(f) && pdf_b > 0.0f) {
                    RayQ sq;
                    sq.o = spawn_origin(it.p, wi, it.n);
                    sq.d = wi;
                    sq.tmin = 0.0f;
                    sq.tmax = wi_dist;
                    sq.setup();
                    if (!occluded(s, sq)) {
                        float w = 1.0f;
                        if (!delta_light)
                            w = p.mis_mode == 1 ? power_heur(pdf_l * pr_light, pdf_b) : power_heur(pdf_l, pdf_b);
                        Ld = Le_beta * f * w / pr_light;
                    }
                }
            }
        } else {
            sob.next();
            float a, b;
            sob.next2d(a, b);
        }
        if (bounce == 0 && p.clamp_direct > 0.0f) Ld = vmin(Ld, p.clamp_direct);
        else if (bounce >= 1 && p.clamp_indirect > 0.0f) Ld = vmin(Ld, p.clamp_indirect);
        L = L + beta * Ld;

        // BSDF sampling (material.cpp:157-166)
        V3 wol = sh_to_local(it, wo);
        float u_lobe = sob.next();
        float u0, u1;
        sob.next2d(u0, u1);
        V3 wil;
        float pdf_wi = 0.0f;
        V3 sbeta = bsdf_sample(wol, wil, c, u_lobe, u0, u1, pdf_wi);
        if (maxc(sbeta) == 0.0f || pdf_wi == 0.0f) break;
        V3 wi = sh_to_world(it, wil);
        beta = beta * sbeta;
        if (maxc(beta) == 0.0f) break;
        pdf_wi_bsdf = pdf_wi;
        ro = spawn_origin(it.p, wi, it.n);
        rd = wi;
        // Russian roulette (small_pt.cpp:167-174)
        if (maxc(beta) < 1.0f && bounce >= 1) {
            float qq = std::max(0.0f, 1.0f - maxc(beta));
            if (sob.next() < qq) break;
            beta = beta / (1.0f - qq);
        }
    }
    return {true, L};
}

template <typename F> void parallel_for(int64_t n, int threads, F f) {
    if (threads <= 1 || n < 64) {
        for (int64_t i = 0; i < n; ++i) f(i);
        return;
    }
    std::atomic<int64_t> next(0);
    std::vector<std::thread> pool;
    const int64_t chunk = 64;
    for (int t = 0; t < threads; ++t)
        pool.emplace_back([&]() {
            while (true) {
                int64_t b = next.fetch_add(chunk);
                if (b >= n) break;
                int64_t e = std::min(n, b + chunk);
                for (int64_t i = b; i < e; ++i) f(i);
            }
        });
    for (auto &th : pool) th.join();
}

int resolve_threads(int t) {
    if (t > 0) return t;
    unsigned h = std::thread::hardware_concurrency();
    return h ? (int)h : 1;
}

// Pixel filters (render_target.h:70-184, create_pixel_filter render_target.cpp:28-44).
struct PixelFilter {
    int type = 0;
    float r = 0.5f, sigma = 0.5f;
    int n = 0;                             // sample-table size (gaussian, Blackman-Harris)
    std::vector<std::vector<float>> cond;  // DistribTable per row (cdf, n + 1 entries)
    std::vector<float> margin;             // DistribTable over rows
    float eval(float x, float y) const {
        if (x < -r || x > r || y < -r || y > r) return 0.0f;
        if (type == 1) { // GaussianPixelFilter::eval (render_target.h:135-143), gaussian (maths.h:1070-1073)
            auto g = [&](float v) {
                return 1.0f / std::sqrt(6.28318530718f * sigma * sigma) * std::exp(-sqr(v - 0.0f) / (2.0f * sigma * sigma));
            };
            float thr = g(r);
            return (g(x) - thr) * (g(y) - thr);
        }
        // BlackmanHarrisPixelFilter::eval (render_target.h:166-180)
        constexpr float a0 = 0.35875f, a1 = 0.48829f, a2 = 0.14128f, a3 = 0.01168f, pi = 3.14159265359f;
        float xx = (x + r) / (2.0f * r);
        float fx = a0 - a1 * std::cos(2.0f * pi * xx) + a2 * std::cos(4.0f * pi * xx) - a3 * std::cos(6.0f * pi * xx);
        float yy = (y + r) / (2.0f * r);
        float fy = a0 - a1 * std::cos(2.0f * pi * yy) + a2 * std::cos(4.0f * pi * yy) - a3 * std::cos(6.0f * pi * yy);
        return fx * fy;
    }
};

// DistribTable::sample_linear (distrib.cpp:43-53)
float sample_linear(const std::vector<float> &cdf, float u, uint32_t &index) {
    u = clampv(u, 0.0f, std::nextafter(1.0f, 0.0f));
    auto it = std::upper_bound(cdf.begin(), cdf.end(), u);
    index = (uint32_t)std::distance(cdf.begin(), std::prev(it));
    float n = (float)(cdf.size() - 1);
    float du = (u - cdf[index]) / (cdf[index + 1] - cdf[index]);
    return (index + du) / n;
}

// The filter of the params, with its sample table (PixelFilterSampleTable, render_target.h:79-106) built
// once per (type, width, sigma).
const PixelFilter &pixel_filter(const ora_params &p) {
    static std::mutex mu;
    static std::vector<std::unique_ptr<PixelFilter>> cache;
    float width = p.filter_width != 0.0f ? p.filter_width : (p.pixel_filter == 0 ? 1.0f : 3.0f);
    float sigma = p.filter_sigma != 0.0f ? p.filter_sigma : 0.5f;
    std::lock_guard<std::mutex> lock(mu);
    for (auto &f : cache)
        if (f->type == p.pixel_filter && f->r == 0.5f * width && f->sigma == sigma) return *f;
    auto f = std::make_unique<PixelFilter>();
    f->type = p.pixel_filter;
    f->r = 0.5f * width;
    f->sigma = sigma;
    if (f->type != 0) {
        const float r = f->r;
        const int n = (int)std::ceil(32 * r);
        f->n = n;
        std::vector<float> vals((size_t)n * n), acc(n);
        for (int y = 0; y < n; ++y)
            for (int x = 0; x < n; ++x) {
                float tx = (x + 0.5f) / n, ty = (y + 0.5f) / n;
                vals[(size_t)y * n + x] = f->eval(std::lerp(-r, r, tx), std::lerp(-r, r, ty));
            }
        for (int y = 0; y < n; ++y) {
            std::vector<float> row(vals.begin() + (size_t)y * n, vals.begin() + (size_t)(y + 1) * n);
            std::vector<float> cdf(n + 1);
            cdf[0] = 0.0f;
            for (int i = 1; i <= n; ++i) cdf[i] = cdf[i - 1] + row[i - 1] * (1.0f / (float)n);
            acc[y] = cdf[n];
            f->cond.push_back(make_cdf(row));
        }
        f->margin = make_cdf(acc);
    }
    cache.push_back(std::move(f));
    return *cache.back();
}

void camera_sample(const Camera &cam, const ora_params &p, int px, int py, ZSobol &sob, V3 &o, V3 &d, V3 *rxd = nullptr,
                   V3 *ryd = nullptr) {
    // small_pt.cpp:98-105: the filter's offset from the pixel centre (render_target.h:108-184)
    float ox = 0.0f, oy = 0.0f;
    if (p.spp != 1) {
        const PixelFilter &f = pixel_filter(p);
        float u0, u1;
        sob.next2d(u0, u1);
        if (f.type != 0) { // PixelFilterSampleTable::sample (render_target.h:96-103), DistribTable2D::sample_linear
            uint32_t iy, ix;
            float vy = sample_linear(f.margin, u1, iy);
            float vx = sample_linear(f.cond[iy], u0, ix);
            u0 = vx;
            u1 = vy;
        }
        ox = (1.0f - u0) * -f.r + u0 * f.r; // lerp(Constant(-r), Constant(r), t) (maths.h:686-690)
        oy = (1.0f - u1) * -f.r + u1 * f.r;
    }
    float fx = ((float)px + 0.5f + ox) / (float)p.render_width;
    float fy = ((float)py + 0.5f + oy) / (float)p.render_height;
    cam.spawn_ray(fx, fy, o, d, rxd, ryd, p.render_width, p.render_height, p.scale_ray_diff ? p.spp : 1);
}

} // namespace

extern "C" {

void *ora_scene_create(const ora_scene_desc *desc) { return create(desc); }
void ora_scene_free(void *scene) { delete (Scene *)scene; }

int ora_intersect(void *scene, const float *rays, int64_t n, float *out, int threads) {
    const Scene &s = *(Scene *)scene;
    parallel_for(n, resolve_threads(threads), [&](int64_t i) {
        const float *r = rays + 8 * i;
        RayQ q;
        q.o = V3(r[0], r[1], r[2]);
        q.tmin = r[3];
        q.d = V3(r[4], r[5], r[6]);
        q.tmax = r[7];
        q.setup();
        Hit h = closest(s, q);
        out[4 * i] = h.t;
        out[4 * i + 1] = h.u;
        out[4 * i + 2] = h.v;
        std::memcpy(&out[4 * i + 3], &h.prim, 4);
    });
    return 0;
}

int ora_intersect_brute(void *scene, const float *rays, int64_t n, float *out, int threads) {
    const Scene &s = *(Scene *)scene;
    parallel_for(n, resolve_threads(threads), [&](int64_t i) {
        const float *r = rays + 8 * i;
        RayQ q;
        q.o = V3(r[0], r[1], r[2]);
        q.tmin = r[3];
        q.d = V3(r[4], r[5], r[6]);
        q.tmax = r[7];
        q.setup();
        Hit h = brute(s, q);
        out[4 * i] = h.t;
        out[4 * i + 1] = h.u;
        out[4 * i + 2] = h.v;
        std::memcpy(&out[4 * i + 3], &h.prim, 4);
    });
    return 0;
}

int ora_occluded(void *scene, const float *rays, int64_t n, uint8_t *out, int threads) {
    const Scene &s = *(Scene *)scene;
    parallel_for(n, resolve_threads(threads), [&](int64_t i) {
        const float *r = rays + 8 * i;
        RayQ q;
        q.o = V3(r[0], r[1], r[2]);
        q.tmin = r[3];
        q.d = V3(r[4], r[5], r[6]);
        q.tmax = r[7];
        q.setup();
        out[i] = occluded(s, q) ? 1 : 0;
    });
    return 0;
}

int ora_sky_eval(void *scene, const float *dirs, int64_t n, float *le, float *pdf) {
    const Scene &s = *(const Scene *)scene;
    if (s.sky_map.texels.empty()) return -1;
    for (int64_t i = 0; i < n; ++i) {
        V3 d(dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2]);
        V3 L = s.sky_map.eval(d);
        le[3 * i] = L.x, le[3 * i + 1] = L.y, le[3 * i + 2] = L.z;
        pdf[i] = s.sky_map.pdf(d);
    }
    return 0;
}

int ora_sky_sample(void *scene, const float *u, int64_t n, float *wi, float *le_over_pdf, float *pdf) {
    const Scene &s = *(const Scene *)scene;
    if (s.sky_map.texels.empty()) return -1;
    for (int64_t i = 0; i < n; ++i) {
        V3 d;
        float pd = 0.0f;
        V3 b = s.sky_map.sample(u[2 * i], u[2 * i + 1], d, pd);
        wi[3 * i] = d.x, wi[3 * i + 1] = d.y, wi[3 * i + 2] = d.z;
        le_over_pdf[3 * i] = b.x, le_over_pdf[3 * i + 1] = b.y, le_over_pdf[3 * i + 2] = b.z;
        pdf[i] = pd;
    }
    return 0;
}

int ora_field_eval(void *scene, const ora_field *f, const float *uvd, int64_t n, float *out) {
    const Scene &s = *(const Scene *)scene;
    Field fd;
    // as many components as the texture has channels (a constant field: all four)
    int nc = f->texture >= 0 && f->texture < (int)s.textures.size() ? s.textures[f->texture].nc : 4;
    if (!fd.init(*f, nc, s.textures)) return -1;
    for (int64_t i = 0; i < n; ++i) {
        const float *r = uvd + 6 * i;
        std::fill(out + 4 * i, out + 4 * i + 4, 0.0f);
        fd.eval(s.textures, r[0], r[1], r + 2, nc, out + 4 * i);
    }
    return 0;
}

int ora_texture_level(void *scene, int tex, int level, int *w, int *h, float *out) {
    const Scene &s = *(const Scene *)scene;
    if (tex < 0 || tex >= (int)s.textures.size()) return -1;
    const Texture &t = s.textures[tex];
    if (level == -1) return t.nc;
    if (level < 0 || level >= t.levels()) return -1;
    *w = t.mips[level].w;
    *h = t.mips[level].h;
    if (out)
        for (int y = 0; y < *h; ++y)
            for (int x = 0; x < *w; ++x) t.fetch(x, y, level, out + ((size_t)y * *w + x) * t.nc);
    return 0;
}

void ora_tex_tables(float *srgb256, float *ewa128) {
    std::memcpy(srgb256, g_tex_tables.srgb, sizeof(g_tex_tables.srgb));
    std::memcpy(ewa128, g_tex_tables.ewa, sizeof(g_tex_tables.ewa));
}

int ora_primary_surface(void *scene, const ora_camera *cam_desc, const ora_params *p, const int32_t *px,
                        const int32_t *py, const int32_t *sample, int64_t n, float *out) {
    const Scene &s = *(Scene *)scene;
    Camera cam(*cam_desc);
    for (int64_t i = 0; i < n; ++i) {
        ZSobol sob(p->spp, p->render_width, p->render_height, p->rng_seed, px[i], py[i], sample[i]);
        V3 o, d, rxd, ryd;
        camera_sample(cam, *p, px[i], py[i], sob, o, d, &rxd, &ryd);
        RayQ q;
        q.o = o;
        q.d = d;
        q.tmin = 0.0f;
        q.tmax = kInf;
        q.setup();
        Hit h = closest(s, q);
        float *r = out + 16 * i;
        std::fill(r, r + 16, 0.0f);
        int32_t prim = h.prim;
        if (h.prim >= 0) {
            SurfHit it = make_hit(s, d, h);
            compute_uv_partials(it, o, rxd, ryd);
            const TexMaterial *tm = s.tmat((uint32_t)h.prim);
            if (tm && tm->has_normal) apply_normal_map(s, *tm, it);
            float v[15] = {it.u, it.v, it.duv[0], it.duv[1], it.duv[2], it.duv[3], it.sn.x, it.sn.y, it.sn.z,
                           it.st.x, it.st.y, it.st.z, it.n.x, it.n.y, it.n.z};
            std::memcpy(r, v, sizeof(v));
        }
        std::memcpy(r + 15, &prim, 4);
    }
    return 0;
}

int ora_light_table(void *scene, int32_t *n_lights, int32_t *prim, float *power, float *cdf) {
    const Scene &s = *(Scene *)scene;
    int32_t n = (int32_t)s.lights.size();
    if (n_lights) *n_lights = n;
    for (int32_t i = 0; i < n; ++i) {
        if (prim) prim[i] = s.lights[i].type == 3 ? (int32_t)s.lights[i].prim : -1;
        if (power) power[i] = s.powers[i];
    }
    if (cdf)
        for (int32_t i = 0; i <= n && n > 0; ++i) cdf[i] = s.cdf[i];
    return 0;
}

int ora_render(void *scene, const ora_camera *cam_desc, const ora_params *p, double *rgb, double *weights,
               int threads) {
    const Scene &s = *(Scene *)scene;
    Camera cam(*cam_desc);
    int W = p->render_width, H = p->render_height;
    // crop (small_pt.cpp:23-35)
    int cx = 0, cy = 0, cw = W, ch = H;
    if (p->crop_width > 0 && p->crop_height > 0) {
        cx = std::max(p->crop_start_x, 0);
        cy = std::max(p->crop_start_y, 0);
        cw = std::min(p->crop_width, W - cx);
        ch = std::min(p->crop_height, H - cy);
    }
    std::memset(rgb, 0, sizeof(double) * 3 * (size_t)W * H);
    std::memset(weights, 0, sizeof(double) * 2 * (size_t)W * H);
    // parallel_tile_2d with 4x4 tiles (parallel.h:93-113)
    int ntx = (cw + 3) / 4, nty = (ch + 3) / 4;
    parallel_for((int64_t)ntx * nty, resolve_threads(threads), [&](int64_t tile) {
        int tyi = (int)(tile / ntx), txi = (int)(tile - (int64_t)tyi * ntx);
        for (int oy = tyi * 4; oy < std::min(tyi * 4 + 4, ch); ++oy)
            for (int ox = txi * 4; ox < std::min(txi * 4 + 4, cw); ++ox) {
                int px = cx + ox, py = cy + oy;
                size_t pi = (size_t)py * W + px;
                for (int smp = 0; smp < p->spp; ++smp) {
                    ZSobol sob(p->spp, W, H, p->rng_seed, px, py, smp);
                    V3 o, d, rxd, ryd;
                    camera_sample(cam, *p, px, py, sob, o, d, &rxd, &ryd);
                    PathOut r = trace(s, *p, o, d, sob, rxd, ryd);
                    if (r.hit) { // RenderTarget2::add / add_miss (render_target.cpp:82-107)
                        weights[2 * pi] += 1.0;
                        rgb[3 * pi] += (double)r.L.x;
                        rgb[3 * pi + 1] += (double)r.L.y;
                        rgb[3 * pi + 2] += (double)r.L.z;
                    } else {
                        weights[2 * pi + 1] += 1.0;
                    }
                }
            }
    });
    return 0;
}

int ora_trace_samples(void *scene, const ora_camera *cam_desc, const ora_params *p, const int32_t *px,
                      const int32_t *py, const int32_t *sample, int64_t n, float *out, int threads) {
    const Scene &s = *(Scene *)scene;
    Camera cam(*cam_desc);
    parallel_for(n, resolve_threads(threads), [&](int64_t i) {
        ZSobol sob(p->spp, p->render_width, p->render_height, p->rng_seed, px[i], py[i], sample[i]);
        V3 o, d, rxd, ryd;
        camera_sample(cam, *p, px[i], py[i], sob, o, d, &rxd, &ryd);
        PathOut r = trace(s, *p, o, d, sob, rxd, ryd);
        out[4 * i] = r.L.x;
        out[4 * i + 1] = r.L.y;
        out[4 * i + 2] = r.L.z;
        out[4 * i + 3] = r.hit ? 1.0f : 0.0f;
    });
    return 0;
}

int ora_sobol(int width, int height, int spp, int seed, int px, int py, int sample, int n_calls, const int32_t *kinds,
              uint32_t *out_u32, float *out_f) {
    ZSobol sob(spp, width, height, seed, px, py, sample);
    int k = 0;
    for (int c = 0; c < n_calls; ++c) {
        if (kinds[c] == 1) {
            uint32_t v = sob.next_u32();
            out_u32[k] = v;
            out_f[k] = u32_to_unit(v);
            ++k;
        } else {
            uint32_t v[2];
            sob.next2d_u32(v);
            out_u32[k] = v[0];
            out_f[k] = u32_to_unit(v[0]);
            out_u32[k + 1] = v[1];
            out_f[k + 1] = u32_to_unit(v[1]);
            k += 2;
        }
    }
    return 0;
}

uint64_t ora_murmur64a(const void *key, int len, uint64_t seed) {
    return murmur64a((const unsigned char *)key, (size_t)len, seed);
}
uint64_t ora_mix_bits(uint64_t v) { return mix_bits(v); }
uint32_t ora_owen_scramble(uint32_t v, uint32_t seed) { return owen_scramble(v, seed); }
void ora_sobol_matrices(uint32_t *out104) { std::memcpy(out104, g_sobol.m, sizeof(g_sobol.m)); }

int ora_camera_rays(const ora_camera *cam_desc, int width, int height, int spp, int seed, const int32_t *px,
                    const int32_t *py, const int32_t *sample, int64_t n, float *rays) {
    Camera cam(*cam_desc);
    ora_params p{};
    p.render_width = width;
    p.render_height = height;
    p.spp = spp;
    for (int64_t i = 0; i < n; ++i) {
        ZSobol sob(spp, width, height, seed, px[i], py[i], sample[i]);
        V3 o, d;
        camera_sample(cam, p, px[i], py[i], sob, o, d);
        float *r = rays + 6 * i;
        r[0] = o.x;
        r[1] = o.y;
        r[2] = o.z;
        r[3] = d.x;
        r[4] = d.y;
        r[5] = d.z;
    }
    return 0;
}

int ora_camera_matrices(const ora_camera *cam_desc, float *p2w, float *pos) {
    Camera cam(*cam_desc);
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) p2w[4 * i + j] = (float)cam.proj_to_world.a[i][j];
    pos[0] = cam.position.x;
    pos[1] = cam.position.y;
    pos[2] = cam.position.z;
    return 0;
}

int ora_bsdf_eval(const ora_material *m, const float *wo, const float *wi, int64_t n, float *f, float *pdf) {
    Closure c = make_closure(*m);
    for (int64_t i = 0; i < n; ++i) {
        V3 o(wo[3 * i], wo[3 * i + 1], wo[3 * i + 2]), w(wi[3 * i], wi[3 * i + 1], wi[3 * i + 2]);
        V3 ff = bsdf_eval(o, w, c);
        f[3 * i] = ff.x;
        f[3 * i + 1] = ff.y;
        f[3 * i + 2] = ff.z;
        pdf[i] = bsdf_pdf(o, w, c);
    }
    return 0;
}

int ora_bsdf_sample(const ora_material *m, const float *wo, const float *u, int64_t n, float *wi, float *beta,
                    float *pdf) {
    Closure c = make_closure(*m);
    for (int64_t i = 0; i < n; ++i) {
        V3 o(wo[3 * i], wo[3 * i + 1], wo[3 * i + 2]);
        V3 w;
        float pd = 0.0f;
        V3 b = bsdf_sample(o, w, c, u[3 * i], u[3 * i + 1], u[3 * i + 2], pd);
        wi[3 * i] = w.x;
        wi[3 * i + 1] = w.y;
        wi[3 * i + 2] = w.z;
        beta[3 * i] = b.x;
        beta[3 * i + 1] = b.y;
        beta[3 * i + 2] = b.z;
        pdf[i] = pd;
    }
    return 0;
}

// ----------------------------------------------------------------------------------------------------
// Probes of the oracle's building blocks, for the tests that pin them against closed forms and
// quadrature. They call the same functions the trace loop calls.
// ----------------------------------------------------------------------------------------------------
int ora_ggx_eval(float ax, float ay, int what, const float *a, const float *b, int64_t n, float *out) {
    GGX g(ax, ay);
    for (int64_t i = 0; i < n; ++i) {
        V3 x(a[3 * i], a[3 * i + 1], a[3 * i + 2]);
        if (what == 0) out[i] = g.D(x);
        else if (what == 1) out[i] = g.lambda(x);
        else if (what == 2) out[i] = g.G1(x);
        else if (what == 3) out[i] = g.G2(x, V3(b[3 * i], b[3 * i + 1], b[3 * i + 2]));
        else return -1;
    }
    return 0;
}

int ora_ggx_sample(float ax, float ay, const float *wo, const float *u, int64_t n, float *wm) {
    GGX g(ax, ay);
    for (int64_t i = 0; i < n; ++i) {
        V3 m = g.sample(V3(wo[3 * i], wo[3 * i + 1], wo[3 * i + 2]), u[2 * i], u[2 * i + 1]);
        wm[3 * i] = m.x;
        wm[3 * i + 1] = m.y;
        wm[3 * i + 2] = m.z;
    }
    return 0;
}

int ora_fresnel(int what, const float *c, float eta, int64_t n, float *out) {
    for (int64_t i = 0; i < n; ++i) {
        if (what == 0) out[i] = fresnel_dielectric(c[i], eta);
        else if (what == 1) out[i] = fresnel_schlick(c[i]);
        else if (what == 2) out[i] = hg_phase(c[i], eta);
        else return -1;
    }
    return 0;
}

int ora_bsdf_lobes(const ora_material *m, const float *wo, int64_t n, float *w4) {
    Closure c = make_closure(*m);
    for (int64_t i = 0; i < n; ++i) lobe_weights(V3(wo[3 * i], wo[3 * i + 1], wo[3 * i + 2]), c, w4 + 4 * i);
    return 0;
}

// One lobe's cosine-weighted value (0 diffuse, 1 diffuse transmission, 2 metallic specular, 3 dielectric
// specular), without the hemisphere gating of eval (principled_bsdf.cpp:308-324); lobe -1 is the full
// bsdf_pdf() without eval's zero-f gating.
int ora_bsdf_lobe(const ora_material *m, int lobe, const float *wo, const float *wi, int64_t n, float *f) {
    Closure c = make_closure(*m);
    for (int64_t i = 0; i < n; ++i) {
        V3 o(wo[3 * i], wo[3 * i + 1], wo[3 * i + 2]), w(wi[3 * i], wi[3 * i + 1], wi[3 * i + 2]);
        V3 r;
        if (lobe == -1) {
            f[3 * i] = f[3 * i + 1] = f[3 * i + 2] = bsdf_pdf(o, w, c);
            continue;
        }
        if (lobe == 0) r = eval_diffuse(o, w, c);
        else if (lobe == 1) r = eval_diffuse_transmission(o, w, c);
        else if (lobe == 2) r = eval_metallic_specular(o, w, c);
        else if (lobe == 3) r = eval_dielectric_specular(o, w, c);
        else return -1;
        f[3 * i] = r.x;
        f[3 * i + 1] = r.y;
        f[3 * i + 2] = r.z;
    }
    return 0;
}

// Light `light` of the flat table sampled from shading points p (n x 3) with u (n x 2): wi (n x 3),
// the shadow-ray distance (n), the solid-angle pdf (n) and Le / pdf (n x 3).
int ora_light_sample(void *scene, int light, const float *p, const float *u, int64_t n, float *wi, float *dist,
                     float *pdf, float *le_over_pdf) {
    const Scene &s = *(const Scene *)scene;
    if (light < 0 || light >= (int)s.lights.size()) return -1;
    for (int64_t i = 0; i < n; ++i) {
        V3 w;
        float d = 0.0f, pd = 0.0f;
        V3 r = light_sample(s, s.lights[light], V3(p[3 * i], p[3 * i + 1], p[3 * i + 2]), u[2 * i], u[2 * i + 1], w, d,
                            pd);
        wi[3 * i] = w.x;
        wi[3 * i + 1] = w.y;
        wi[3 * i + 2] = w.z;
        dist[i] = d;
        pdf[i] = pd;
        le_over_pdf[3 * i] = r.x;
        le_over_pdf[3 * i + 1] = r.y;
        le_over_pdf[3 * i + 2] = r.z;
    }
    return 0;
}

// MeshTriLight::pdf (light.cpp:607-628) of mesh light `light` for directions wi at distances dist.
int ora_light_pdf(void *scene, int light, const float *wi, const float *dist, int64_t n, float *pdf) {
    const Scene &s = *(const Scene *)scene;
    if (light < 0 || light >= (int)s.lights.size() || s.lights[light].type != 3) return -1;
    for (int64_t i = 0; i < n; ++i)
        pdf[i] = mesh_light_pdf(s, s.lights[light], V3(wi[3 * i], wi[3 * i + 1], wi[3 * i + 2]), dist[i]);
    return 0;
}

// PowerLightSampler::sample (light.cpp:677-683 over distrib.cpp:33-41): index and its probability.
int ora_light_select(void *scene, const float *u, int64_t n, int32_t *idx, float *prob) {
    const Scene &s = *(const Scene *)scene;
    if (s.cdf.size() < 2) return -1;
    for (int64_t i = 0; i < n; ++i) {
        float pr = 0.0f;
        idx[i] = (int32_t)cdf_sample(s.cdf, u[i], pr);
        prob[i] = light_probability(s, (uint32_t)idx[i]);
    }
    return 0;
}

// The pixel filter of the params: offsets from u (n x 2) through sample(), and eval() at offsets (n x 2).
int ora_pixel_filter(const ora_params *p, const float *u, const float *xy, int64_t n, float *offset, float *value) {
    const PixelFilter &f = pixel_filter(*p);
    for (int64_t i = 0; i < n; ++i) {
        if (u && offset) {
            float u0 = u[2 * i], u1 = u[2 * i + 1];
            if (f.type != 0) {
                uint32_t iy, ix;
                float vy = sample_linear(f.margin, u1, iy);
                float vx = sample_linear(f.cond[iy], u0, ix);
                u0 = vx;
                u1 = vy;
            }
            offset[2 * i] = (1.0f - u0) * -f.r + u0 * f.r;
            offset[2 * i + 1] = (1.0f - u1) * -f.r + u1 * f.r;
        }
        if (xy && value) {
            float x = xy[2 * i], y = xy[2 * i + 1];
            value[i] = f.type == 0 ? ((x < -f.r || x > f.r || y < -f.r || y > f.r) ? 0.0f : 1.0f) : f.eval(x, y);
        }
    }
    return f.n;
}

} // extern "C"

// Path-guiding model on gfx950: multiresolution hash-grid encoding of the shading position, a
// 64-64-64-48 ReLU MLP on MFMA (BF16 operands, FP32 accumulation) and an 8-lobe von Mises-Fisher
// mixture head. The lobe math restates vmf.h:10-127 (VMF::eval, VMF::sample, max_kappa); the mixture
// sampling follows sample_vmm (vmf.h:121-127). The reference has no network; its layout is this
// build's own.
#pragma once
#include "ks_device.h"

namespace ksd {

constexpr int kGuideLevels = 16;
constexpr int kGuideFeatures = 2;
constexpr int kGuideLobes = 8;
constexpr int kGuideIn = 64;     // 32 grid features + normal (3) + wo (3) + constant 1 + zero padding
constexpr int kGuideHidden = 64;
constexpr int kGuideOut = 48;    // 8 x (logit, mean xyz, log kappa) = 40, padded to 48
constexpr float kMaxKappa = 1e4f; // VMF::max_kappa (vmf.h:110)
constexpr float kMinKappa = 1e-2f;
constexpr float kGuideMinRoughness = 0.05f; // sharper materials sample the BSDF only
constexpr int kWaveSize = 64;

// Network parameter layout (FP32 master copy): [grid: L * T * F][W1: 64x64][W2: 64x64][W3: 64x48].
// Wk is stored [in][out] (row = input unit) for the backward pass and transposed [out][in] (BF16) for
// the forward MFMA B operand.
// FP8 (OCP e4m3) inference: per-tensor power-of-two scales for the three layer inputs and the three
// weight matrices, and the weights pre-arranged as MFMA 16x16x32 A fragments (one 8-byte word per lane
// per m-tile and 32-wide k-block).
constexpr int kFp8MTiles = 4 + 4 + 3;
constexpr int kFp8Frags = kFp8MTiles * 2 * 64; // 64-bit words, 11264 bytes
enum Fp8Scale {
    kInvSx = 0, kInvSh1, kInvSh2, // 1 / input scale of layers 1..3
    kC1, kC2, kC3,                // input scale x weight scale of layers 1..3 (dequantisation)
    kSw1, kSw2, kSw3,             // weight scales
    kSx, kSh1, kSh2,              // input scales
    kFp8ScaleCount
};

struct GuideNet {
    const uint32_t *grid_h;   // FP16 feature pairs, L * T words
    const uint16_t *wt[3];    // BF16 [out][in] (forward)
    const uint16_t *w[3];     // BF16 [in][out] (backward)
    const uint64_t *w8;       // FP8 A fragments (kFp8Frags words)
    const float *fp8;         // kFp8ScaleCount scales
    uint32_t log2_T;
    uint32_t res[kGuideLevels]; // level resolutions floor(base * growth^l), computed on the host in FP64
    float lo[3], inv_ext[3];    // scene bound normalisation
};

struct Lobe {
    f3 mu;
    float kappa;
};

// ---------------------------------------------------------------------------------------------------
// vMF lobes (vmf.h:10-119)
// ---------------------------------------------------------------------------------------------------
// log of the normalisation kappa / (2 pi (1 - exp(-2 kappa))); 1 / (4 pi) at kappa = 0 (vmf.h:32-38)
KS_DEV float vmf_log_norm(float k) {
    if (k < 1e-4f) return -logf(4.0f * kPi);
    return logf(k) - logf(2.0f * kPi) - logf(-expm1f(-2.0f * k));
}
// d ln(norm) / d kappa = 1/kappa - 2 / (exp(2 kappa) - 1)
KS_DEV float vmf_dlog_norm(float k) {
    if (k < 1e-3f) return 1.0f - k / 3.0f;
    return 1.0f / k - 2.0f / expm1f(2.0f * k);
}
KS_DEV float vmf_log_eval(const Lobe &l, f3 w) { // log VMF::eval
    return vmf_log_norm(l.kappa) + l.kappa * fminf(0.0f, dot(l.mu, w) - 1.0f);
}
KS_DEV f3 vmf_sample(const Lobe &l, float u0, float u1) { // VMF::sample (vmf.h:40-56)
    float k = l.kappa;
    float ct;
    if (k == 0.0f) {
        ct = 1.0f - 2.0f * u1; // sample_uniform_sphere (rng.h:284-292) with u = (u0, u1)
        float r = sqrtf(fmaxf(0.0f, 1.0f - ct * ct));
        float s, c;
        sincosf(kTwoPi * u0, &s, &c);
        return mk(r * c, r * s, ct);
    }
    ct = 1.0f + logf(u0 + expf(-2.0f * k) * (1.0f - u0)) / k;
    float st = sqrtf(fmaxf(0.0f, 1.0f - ct * ct));
    float s, c;
    sincosf(kTwoPi * u1, &s, &c);
    f3 X, Y;
    orthonormal_basis(l.mu, X, Y);
    return X * (c * st) + Y * (s * st) + l.mu * ct;
}

struct Mixture {
    float w[kGuideLobes];
    Lobe lobe[kGuideLobes];
};

// log of the mixture pdf sum_k w_k vmf_k(w) with a log-sum-exp
KS_DEV float mixture_log_pdf(const Mixture &m, f3 w) {
    float lg[kGuideLobes];
    float mx = -kInf;
    #pragma unroll
    for (int k = 0; k < kGuideLobes; ++k) {
        lg[k] = m.w[k] > 0.0f ? logf(m.w[k]) + vmf_log_eval(m.lobe[k], w) : -kInf;
        mx = fmaxf(mx, lg[k]);
    }
    if (mx == -kInf) return -kInf;
    float s = 0.0f;
    #pragma unroll
    for (int k = 0; k < kGuideLobes; ++k) s += expf(lg[k] - mx);
    return mx + logf(s);
}
KS_DEV float mixture_pdf(const Mixture &m, f3 w) { return expf(mixture_log_pdf(m, w)); }

// Lobe selection of sample_vmm (vmf.h:121-127, sample_small_distrib rng.h:437-470): returns the lobe and
// the re-used u0 in *ur. The caller fetches the chosen lobe by index from memory, which keeps the
// mixture out of scratch.
KS_DEV int mixture_select(const float w[kGuideLobes], float u0, float *ur) {
    float sum = 0.0f;
    int last = 0;
#pragma unroll
    for (int k = 0; k < kGuideLobes; ++k) {
        sum += w[k];
        if (w[k] > 0.0f) last = k;
    }
    float inv = 1.0f / sum, cdf = 0.0f;
    int sel = -1;
    float r = u0;
#pragma unroll
    for (int k = 0; k < kGuideLobes; ++k) { // first k with u0 < cdf_k (branch-free form of the early exit)
        float nx = cdf + w[k] * inv;
        if (sel < 0 && u0 < nx) {
            sel = k;
            r = (u0 - cdf) / (nx - cdf);
        }
        cdf = nx;
    }
    if (sel < 0) sel = last;
    *ur = fminf(fmaxf(r, 0.0f), kBeforeOne);
    return sel;
}

// Network outputs -> mixture: softmax weights, normalised means, kappa = clamp(exp(raw)).
KS_DEV void head_forward(const float o[kGuideOut], Mixture &m) {
    float mx = -kInf;
    #pragma unroll
    for (int k = 0; k < kGuideLobes; ++k) mx = fmaxf(mx, o[5 * k]);
    float s = 0.0f;
    #pragma unroll
    for (int k = 0; k < kGuideLobes; ++k) {
        m.w[k] = expf(o[5 * k] - mx);
        s += m.w[k];
    }
    #pragma unroll
    for (int k = 0; k < kGuideLobes; ++k) {
        m.w[k] /= s;
        f3 r = mk(o[5 * k + 1], o[5 * k + 2], o[5 * k + 3]);
        float n = sqrtf(sqnorm(r));
        m.lobe[k].mu = n > 1e-12f ? r / n : mk(0.0f, 0.0f, 1.0f);
        m.lobe[k].kappa = fminf(fmaxf(expf(o[5 * k + 4]), kMinKappa), kMaxKappa);
    }
}

// d(-c log q(wi)) / d outputs for one sample (c = training weight).
KS_DEV float head_backward(const float o[kGuideOut], f3 wi, float c, float g[kGuideOut]) {
    Mixture m;
    head_forward(o, m);
    float lg[kGuideLobes], d[kGuideLobes];
    float mx = -kInf;
    #pragma unroll
    for (int k = 0; k < kGuideLobes; ++k) {
        d[k] = dot(m.lobe[k].mu, wi);
        lg[k] = logf(fmaxf(m.w[k], 1e-30f)) + vmf_log_norm(m.lobe[k].kappa) +
                m.lobe[k].kappa * fminf(0.0f, d[k] - 1.0f);
        mx = fmaxf(mx, lg[k]);
    }
    float s = 0.0f;
    #pragma unroll
    for (int k = 0; k < kGuideLobes; ++k) s += expf(lg[k] - mx);
    float logq = mx + logf(s);
    #pragma unroll
    for (int j = 0; j < kGuideOut; ++j) g[j] = 0.0f;
    #pragma unroll
    for (int k = 0; k < kGuideLobes; ++k) {
        float r = expf(lg[k] - logq); // responsibility w_k v_k / q
        g[5 * k] = -c * (r - m.w[k]);
        float kap = m.lobe[k].kappa;
        float e = expf(o[5 * k + 4]);
        bool free_kappa = e > kMinKappa && e < kMaxKappa;
        float dk = -c * r * (vmf_dlog_norm(kap) + fminf(0.0f, d[k] - 1.0f));
        g[5 * k + 4] = free_kappa ? dk * kap : 0.0f;
        f3 raw = mk(o[5 * k + 1], o[5 * k + 2], o[5 * k + 3]);
        float n = sqrtf(sqnorm(raw));
        if (n > 1e-12f) {
            f3 gm = (wi - m.lobe[k].mu * d[k]) * (-c * r * kap / n);
            g[5 * k + 1] = gm.x;
            g[5 * k + 2] = gm.y;
            g[5 * k + 3] = gm.z;
        }
    }
    return -c * logq;
}

// ---------------------------------------------------------------------------------------------------
// Hash-grid encoding (multiresolution, trilinear, FP16 features)
// ---------------------------------------------------------------------------------------------------
KS_DEV uint32_t grid_index(uint32_t x, uint32_t y, uint32_t z, uint32_t res, uint32_t log2_T) {
    uint32_t T = 1u << log2_T;
    uint64_t side = (uint64_t)res + 1;
    if (side * side * side <= T) return x + (uint32_t)side * (y + (uint32_t)side * z); // dense level
    return (x ^ (y * 2654435761u) ^ (z * 805459861u)) & (T - 1u);
}

KS_DEV float half_bits_to_float(uint16_t h) { return (float)__builtin_bit_cast(_Float16, h); }
KS_DEV uint16_t float_to_half_bits(float f) { return __builtin_bit_cast(uint16_t, (_Float16)f); } // RNE
KS_DEV float half_lo(uint32_t v) { return half_bits_to_float((uint16_t)(v & 0xffffu)); }
KS_DEV float half_hi(uint32_t v) { return half_bits_to_float((uint16_t)(v >> 16)); }

struct GridCorner {
    uint32_t idx[8];
    float w[8];
};

KS_DEV void grid_level_corners(const GuideNet &net, int level, const float u[3], GridCorner &c) {
    uint32_t res = net.res[level];
    float res_f = (float)res;
    float p[3];
    uint32_t cell[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        float x = fminf(fmaxf(u[a], 0.0f), 1.0f) * res_f;
        float fl = fminf(floorf(x), res_f - 1.0f);
        cell[a] = (uint32_t)fl;
        p[a] = x - fl;
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        uint32_t dx = k & 1, dy = (k >> 1) & 1, dz = (k >> 2) & 1;
        c.idx[k] = grid_index(cell[0] + dx, cell[1] + dy, cell[2] + dz, res, net.log2_T) + ((uint32_t)level << net.log2_T);
        c.w[k] = (dx ? p[0] : 1.0f - p[0]) * (dy ? p[1] : 1.0f - p[1]) * (dz ? p[2] : 1.0f - p[2]);
    }
}

KS_DEV void normalise_position(const GuideNet &net, f3 x, float u[3]) {
    u[0] = (x.x - net.lo[0]) * net.inv_ext[0];
    u[1] = (x.y - net.lo[1]) * net.inv_ext[1];
    u[2] = (x.z - net.lo[2]) * net.inv_ext[2];
}

// BF16 round to nearest even
KS_DEV uint16_t f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)(u >> 16); // inf / nan
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
KS_DEV float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }
KS_DEV uint32_t bf_pair(float a, float b) { return (uint32_t)f2bf(a) | ((uint32_t)f2bf(b) << 16); }

// Input row of one query, BF16: 32 grid features (level-major, FP32 trilinear sum of FP16 corners in
// corner order 0..7), normal, wo, a constant 1 and zero padding to 64.
KS_DEV void guide_encode(const GuideNet &net, f3 x, f3 n, f3 wo, uint16_t *row) {
    float u[3];
    normalise_position(net, x, u);
    uint32_t *row32 = reinterpret_cast<uint32_t *>(row);
    for (int l = 0; l < kGuideLevels; ++l) {
        GridCorner c;
        grid_level_corners(net, l, u, c);
        float f0 = 0.0f, f1 = 0.0f;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            uint32_t v = net.grid_h[c.idx[k]];
            f0 += c.w[k] * half_lo(v);
            f1 += c.w[k] * half_hi(v);
        }
        row32[l] = bf_pair(f0, f1);
    }
    row32[16] = bf_pair(n.x, n.y);
    row32[17] = bf_pair(n.z, wo.x);
    row32[18] = bf_pair(wo.y, wo.z);
    row32[19] = bf_pair(1.0f, 0.0f);
#pragma unroll
    for (int k = 20; k < kGuideIn / 2; ++k) row32[k] = 0u;
}
KS_DEV void guide_zero_row(uint16_t *row) {
    uint32_t *row32 = reinterpret_cast<uint32_t *>(row);
#pragma unroll
    for (int k = 0; k < kGuideIn / 2; ++k) row32[k] = 0u;
}

// ---------------------------------------------------------------------------------------------------
// BF16 and the MFMA tile helpers (v_mfma_f32_16x16x16_bf16; wave64)
// Operand layout: lane l holds A[l % 16][4 * (l / 16) + j] and B[4 * (l / 16) + j][l % 16], j = 0..3;
// the result lane l, register r holds D[4 * (l / 16) + r][l % 16].
// ---------------------------------------------------------------------------------------------------
typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

KS_DEV f32x4 mfma16(bf16x4 a, bf16x4 b, f32x4 c) { return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0); }

} // namespace ksd

// Guide network kernels (gfx950) and their C ABI: initialisation, forward (inference), training
// forward + backward (MFMA for all three layers and their weight gradients, atomics for the hash-grid
// scatter), the per-wave gradient reduction and the Adam step (update rule of
// gpu/shaders/adam.slang:25-51).
#include <hip/hip_runtime.h>

#include "ks_guide.h"
#include "ks_guide_host.h"
#include "ks_internal.h"
#include "ks_render_types.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

using namespace ksd;
using namespace ksr;

#define KS_HIP(call)                                                                                               \
    do {                                                                                                           \
        hipError_t e_ = (call);                                                                                    \
        if (e_ != hipSuccess) {                                                                                    \
            ksa::set_error(std::string(#call) + " failed: " + hipGetErrorString(e_));                              \
            return e_ == hipErrorOutOfMemory ? KS_ERR_OOM : KS_ERR_HIP;                                            \
        }                                                                                                          \
    } while (0)
#define KS_REQUIRE(cond, msg)                                                                                      \
    do {                                                                                                           \
        if (!(cond)) {                                                                                             \
            ksa::set_error(msg);                                                                                   \
            return KS_ERR_INVALID;                                                                                 \
        }                                                                                                          \
    } while (0)

namespace {

constexpr int kLdx = 68;                     // LDS row stride (bf16 elements) of 64-wide activations
constexpr int kActBytes = 64 * kLdx * 2;     // 8704
constexpr int kOutBytes = 64 * kGuideOut * 4; // 12288
constexpr int kTrainWaveBytes = 4 * kActBytes + kOutBytes;
constexpr int kMlpW1 = 64 * 64, kMlpW2 = 64 * 64, kMlpW3 = 64 * kGuideOut;
constexpr int kMlpParams = kMlpW1 + kMlpW2 + kMlpW3;

__device__ __forceinline__ bf16x4 lds_a(const uint16_t *buf, int row, int col) {
    return *reinterpret_cast<const bf16x4 *>(buf + row * kLdx + col);
}
__device__ __forceinline__ bf16x4 lds_strided(const uint16_t *buf, int row0, int col) {
    bf16x4 v;
    v.x = (short)buf[(row0 + 0) * kLdx + col];
    v.y = (short)buf[(row0 + 1) * kLdx + col];
    v.z = (short)buf[(row0 + 2) * kLdx + col];
    v.w = (short)buf[(row0 + 3) * kLdx + col];
    return v;
}
__device__ __forceinline__ bf16x4 glb_a(const uint16_t *p) { return *reinterpret_cast<const bf16x4 *>(p); }

// ---- FP8 (OCP e4m3fn) ------------------------------------------------------------------------------
// The conversion rounds to nearest even and turns magnitudes beyond 464 into NaN, so operands are
// clamped to the largest finite value, 448, first (saturation).
__device__ __forceinline__ uint32_t fp8x4(float a, float b, float c, float d) {
    a = fminf(fmaxf(a, -448.0f), 448.0f);
    b = fminf(fmaxf(b, -448.0f), 448.0f);
    c = fminf(fmaxf(c, -448.0f), 448.0f);
    d = fminf(fmaxf(d, -448.0f), 448.0f);
    int w = __builtin_amdgcn_cvt_pk_fp8_f32(a, b, 0, false);
    w = __builtin_amdgcn_cvt_pk_fp8_f32(c, d, w, true);
    return (uint32_t)w;
}
__device__ __forceinline__ uint64_t fp8x8(const float v[8], float scale) {
    uint32_t lo = fp8x4(v[0] * scale, v[1] * scale, v[2] * scale, v[3] * scale);
    uint32_t hi = fp8x4(v[4] * scale, v[5] * scale, v[6] * scale, v[7] * scale);
    return (uint64_t)lo | ((uint64_t)hi << 32);
}
// D += A (16 x 32, lane l: row l % 16, k 8 (l / 16) .. + 7) x B (32 x 16, lane l: column l % 16, same k)
__device__ __forceinline__ f32x4 mfma8(uint64_t a, uint64_t b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_fp8_fp8((long)a, (long)b, c, 0, 0, 0);
}
// Input unit of k position 8 g + j of a 32-wide k-block whose lane group g holds the units the previous
// layer's two result tiles leave in that lane: 4 g + j (first tile) and 16 + 4 g + j - 4 (second tile).
__host__ __device__ __forceinline__ int fp8_k_unit(int g, int j) { return j < 4 ? 4 * g + j : 16 + 4 * g + (j - 4); }
// Smallest power of two s with v / s <= 448 (1 for v <= 0).
__device__ __forceinline__ float fp8_pow2_scale(float v) {
    if (!(v > 0.0f)) return 1.0f;
    int e;
    float m = frexpf(v / 448.0f, &e); // v / 448 = m 2^e, m in [0.5, 1)
    return ldexpf(1.0f, m == 0.5f ? e - 1 : e);
}

// out[64 x N] = act(in[64 x 64] * W) for one wave; W given transposed ([out][in], BF16, in-size 64).
// N = 16 * nt. The input is BF16 in LDS (stride kLdx); with RELU the result goes to `out_bf` as BF16,
// otherwise to `out_f` as FP32 with row stride ldo.
template <int NT, bool RELU>
__device__ __forceinline__ void wave_dense(const uint16_t *in, const uint16_t *__restrict__ wt, uint16_t *out_bf,
                                           float *out_f, int ldo) {
    const int lane = __lane_id();
    const int r16 = lane & 15, q4 = (lane >> 4) * 4;
    bf16x4 b[NT][4];
#pragma unroll
    for (int n = 0; n < NT; ++n)
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) b[n][kk] = glb_a(wt + (n * 16 + r16) * 64 + kk * 16 + q4);
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        f32x4 acc[NT];
#pragma unroll
        for (int n = 0; n < NT; ++n) acc[n] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) {
            bf16x4 a = lds_a(in, m * 16 + r16, kk * 16 + q4);
#pragma unroll
            for (int n = 0; n < NT; ++n) acc[n] = mfma16(a, b[n][kk], acc[n]);
        }
#pragma unroll
        for (int n = 0; n < NT; ++n)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                int row = m * 16 + q4 + r, col = n * 16 + r16;
                if (RELU) out_bf[row * kLdx + col] = f2bf(fmaxf(acc[n][r], 0.0f));
                else out_f[row * ldo + col] = acc[n][r];
            }
    }
}

// ---------------------------------------------------------------------------------------------------
// Inference with the activations in registers. Each layer is computed transposed,
// out[unit][query] = sum_in W[in][unit] x[in][query]: the MFMA result tile (lane l, register r: unit
// 4 (l / 16) + r, query l % 16) is then exactly the B operand of the next layer for the k-block of those
// units, so no activation goes through LDS. A wave handles 64 queries as four column tiles of 16; lane
// group g = l / 16 encodes the hash levels 2g, 2g+1, 8+2g, 9+2g of query l % 16 (the B operand of
// k-blocks 0 and 1), groups 0 and 1 supply normal, wo and the constant 1 (k-block 2); k-block 3 holds
// only zero padding and is skipped. The head regroups each query's 48 outputs through 3 KB of LDS per
// wave; lane group g then evaluates lobes 2g and 2g+1.
// ---------------------------------------------------------------------------------------------------
struct QueryArray { // queries n x 9: position, normal, wo
    const float *q;
    uint32_t n;
    __device__ uint32_t count() const { return n; }
    __device__ uint32_t slot(uint32_t i) const { return i; }
    __device__ bool load(uint32_t i, f3 &p, f3 &nn, f3 &wo) const {
        if (i >= n) return false;
        const float *e = q + 9 * (size_t)i;
        p = mk(e[0], e[1], e[2]);
        nn = mk(e[3], e[4], e[5]);
        wo = mk(e[6], e[7], e[8]);
        return true;
    }
};

struct QueryHits { // wavefront paths with a surface hit: position, geometric normal facing wo, wo
    DevScene sc;
    const uint32_t *n_dev;
    const float4 *rd;
    const float4 *hits;
    const uint32_t *state; // path state: bounce in bits 16..23
    int32_t max_bounce;    // paths at this bounce or deeper are not guided: no query
    __device__ uint32_t count() const { return *n_dev; }
    __device__ uint32_t slot(uint32_t i) const { return i; }
    __device__ bool wanted(uint32_t i) const {
        return (int32_t)((state[i] >> 16) & 0xffu) < max_bounce && __float_as_int(hits[i].w) >= 0;
    }
    __device__ bool load(uint32_t i, f3 &p, f3 &nn, f3 &wo) const {
        if (i >= *n_dev) return false;
        if ((int32_t)((state[i] >> 16) & 0xffu) >= max_bounce) return false;
        return load_slot(i, p, nn, wo);
    }
    __device__ bool load_slot(uint32_t i, f3 &p, f3 &nn, f3 &wo) const {
        float4 h = hits[i];
        int32_t prim = __float_as_int(h.w);
        if (prim < 0) return false;
        f3 d = ld3(rd[i]);
        SurfHit it = make_hit(sc, sc.tri_idx[prim], h.y, h.z, d);
        p = it.p;
        nn = it.n;
        wo = -normalized(d);
        return true;
    }
};

// The compacted queries of a wavefront buffer: the slots of the paths that want a mixture (a hit below the
// guided bounces), listed by k_guide_compact, so that every wave of the inference works on 64 real
// queries. Mixtures still go to the paths' own slots.
struct QueryList {
    QueryHits h;
    const uint32_t *list, *n_list;
    __device__ uint32_t count() const { return *n_list; }
    __device__ uint32_t slot(uint32_t i) const { return list[i]; }
    __device__ bool load(uint32_t i, f3 &p, f3 &nn, f3 &wo) const {
        if (i >= *n_list) return false;
        return h.load_slot(list[i], p, nn, wo);
    }
};

__global__ __launch_bounds__(256) void k_guide_compact(QueryHits h, uint32_t *list, uint32_t *n_list) {
    __shared__ uint32_t s_app[5];
    const uint32_t n = *h.n_dev;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (blockIdx.x * blockDim.x >= n) return;
    const bool want = i < n && h.wanted(i);
    const uint32_t j = block_append<4>(n_list, want, s_app);
    if (want) list[j] = i;
}

__global__ void k_zero_u32(uint32_t *p) { *p = 0u; }

// Both features of one hash level: FP32 trilinear sum over corners 0..7.
__device__ __forceinline__ void grid_pair(const GuideNet &net, int level, const float u[3], float &f0, float &f1) {
    GridCorner c;
    grid_level_corners(net, level, u, c);
    f0 = 0.0f;
    f1 = 0.0f;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        uint32_t v = net.grid_h[c.idx[k]];
        f0 += c.w[k] * half_lo(v);
        f1 += c.w[k] * half_hi(v);
    }
}

__device__ __forceinline__ bf16x4 pack_bf(float a, float b, float c, float d) {
    bf16x4 v;
    v.x = (short)f2bf(a);
    v.y = (short)f2bf(b);
    v.z = (short)f2bf(c);
    v.w = (short)f2bf(d);
    return v;
}
__device__ __forceinline__ bf16x4 relu_bf(f32x4 d) {
    return pack_bf(fmaxf(d[0], 0.0f), fmaxf(d[1], 0.0f), fmaxf(d[2], 0.0f), fmaxf(d[3], 0.0f));
}
// A operand of W^T (rows = units m*16.., k-block kk) from an [out][in] BF16 copy with row stride ld
__device__ __forceinline__ bf16x4 wt_frag(const uint16_t *wt, int ld, int m, int kk, int q, int g) {
    return *reinterpret_cast<const bf16x4 *>(wt + (m * 16 + q) * ld + kk * 16 + 4 * g);
}

constexpr int kInferBlock = 256;
constexpr int kWtLd = 68; // LDS row stride (BF16) of the staged [out][in] weights: 136 B rows spread banks
constexpr int kInferLdsWeights = (64 + 64 + kGuideOut) * kWtLd; // BF16 elements

// Largest |value| of the wave's lanes, added to *dst with an atomic max (float bits of non-negative
// values order like unsigned integers).
__device__ __forceinline__ void wave_amax(float v, uint32_t *dst) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
    if (__lane_id() == 0) atomicMax(dst, __float_as_uint(v));
}
__device__ __forceinline__ float amax4(bf16x4 v) {
    return fmaxf(fmaxf(fabsf(bf2f((uint16_t)v.x)), fabsf(bf2f((uint16_t)v.y))),
                 fmaxf(fabsf(bf2f((uint16_t)v.z)), fabsf(bf2f((uint16_t)v.w))));
}

// MODE 0: BF16 operands. MODE 1: BF16, and the largest |input| of each layer over all queries goes to
// amax[0..2] (calibration of the FP8 scales). MODE 2: FP8 e4m3 operands with the per-tensor scales of
// net.fp8: each layer's input is multiplied by its inverse scale and rounded to FP8, the FP32
// accumulation is multiplied by (input scale x weight scale). The k positions of each 32-wide block
// are permuted by fp8_k_unit so that a layer's result tiles are the next layer's B operand in place.
template <class Src, int MODE>
__global__ __launch_bounds__(kInferBlock) void k_guide_infer(GuideNet net, Src src, float4 *mix, uint32_t *amax) {
    constexpr bool FP8 = MODE == 2;
    __shared__ __align__(16) uint16_t s_w[FP8 ? 4 : kInferLdsWeights];
    __shared__ __align__(16) uint64_t s_w8[FP8 ? kFp8Frags : 1];
    __shared__ float s_out[kInferBlock / 64][16 * kGuideOut];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int g = lane >> 4, q = lane & 15;
    const uint32_t n = src.count();
    if (blockIdx.x * blockDim.x >= n) return; // whole block idle
    if (FP8) {
        for (int e = threadIdx.x; e < kFp8Frags; e += kInferBlock) s_w8[e] = net.w8[e];
    } else {
        // stage W1^T, W2^T, W3^T (BF16, [out][in]) in LDS once per block
        for (int e = threadIdx.x; e < (64 + 64 + kGuideOut) * 16; e += kInferBlock) {
            int row = e >> 4, c4 = (e & 15) * 4; // rows 0..63 W1, 64..127 W2, 128..175 W3
            const uint16_t *srcw = row < 64 ? net.wt[0] + row * 64 : (row < 128 ? net.wt[1] + (row - 64) * 64
                                                                               : net.wt[2] + (row - 128) * 64);
            *reinterpret_cast<uint2 *>(s_w + row * kWtLd + c4) = *reinterpret_cast<const uint2 *>(srcw + c4);
        }
    }
    float sc[6] = {1, 1, 1, 1, 1, 1};
    if (FP8)
#pragma unroll
        for (int k = 0; k < 6; ++k) sc[k] = net.fp8[k];
    const uint32_t base = blockIdx.x * blockDim.x + wave * 64;
    f3 P = mk(0, 0, 0), Nn = mk(0, 0, 0), WO = mk(0, 0, 0);
    const bool valid = src.load(base + lane, P, Nn, WO);
    float u[3];
    normalise_position(net, P, u);
    // encode all four column tiles first so that their 128 hash gathers per lane are in flight together
    bf16x4 b0[4], b1[4], b2[4];
    uint64_t e0[4], e1[4];
    bool vq[4];
    float am_x = 0.0f, am_h1 = 0.0f, am_h2 = 0.0f;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const int sl = 16 * t + q; // lane that fetched this tile's query
        vq[t] = __shfl((int)valid, sl) != 0;
        float uq[3] = {__shfl(u[0], sl), __shfl(u[1], sl), __shfl(u[2], sl)};
        const float nx = __shfl(Nn.x, sl), ny = __shfl(Nn.y, sl), nz = __shfl(Nn.z, sl);
        const float ox = __shfl(WO.x, sl), oy = __shfl(WO.y, sl), oz = __shfl(WO.z, sl);
        // lanes whose query is skipped (a miss, or a vertex past the guided bounces) issue no gathers;
        // their encodings are zeroed below
        float f[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        if (vq[t]) {
            grid_pair(net, 2 * g, uq, f[0], f[1]);
            grid_pair(net, 2 * g + 1, uq, f[2], f[3]);
            grid_pair(net, 8 + 2 * g, uq, f[4], f[5]);
            grid_pair(net, 9 + 2 * g, uq, f[6], f[7]);
        }
        if (FP8) {
            const float rest[8] = {nx, ny, nz, ox, oy, oz, 1.0f, 0.0f};
            e0[t] = vq[t] ? fp8x8(f, sc[kInvSx]) : 0;
            e1[t] = vq[t] && g == 0 ? fp8x8(rest, sc[kInvSx]) : 0;
        } else {
            b0[t] = pack_bf(f[0], f[1], f[2], f[3]);
            b1[t] = pack_bf(f[4], f[5], f[6], f[7]);
            b2[t] = g == 0 ? pack_bf(nx, ny, nz, ox) : (g == 1 ? pack_bf(oy, oz, 1.0f, 0.0f) : pack_bf(0, 0, 0, 0));
            if (!vq[t]) b0[t] = b1[t] = b2[t] = pack_bf(0, 0, 0, 0);
            if (MODE == 1) am_x = fmaxf(am_x, fmaxf(fmaxf(amax4(b0[t]), amax4(b1[t])), amax4(b2[t])));
        }
    }
    __syncthreads(); // weights staged
    const uint16_t *w1 = s_w, *w2 = s_w + 64 * kWtLd, *w3 = s_w + 128 * kWtLd;
    float *so = s_out[wave];
#pragma unroll 1
    for (int t = 0; t < 4; ++t) {
        if (FP8) {
            auto frag = [&](int mt, int kb) { return s_w8[(mt * 2 + kb) * 64 + lane]; };
            // result tiles m = 0..3 -> FP8 B operand: k-block 0 from tiles 0, 1; k-block 1 from tiles 2, 3
            auto requant = [&](const f32x4 d[4], float c, float inv_s, uint64_t hb[2]) {
#pragma unroll
                for (int kb = 0; kb < 2; ++kb) {
                    float v[8];
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        v[r] = fmaxf(d[2 * kb][r] * c, 0.0f);
                        v[4 + r] = fmaxf(d[2 * kb + 1][r] * c, 0.0f);
                    }
                    hb[kb] = fp8x8(v, inv_s);
                }
            };
            uint64_t hb[2];
            {
                f32x4 d[4];
#pragma unroll
                for (int m = 0; m < 4; ++m) {
                    d[m] = f32x4{0, 0, 0, 0};
                    d[m] = mfma8(frag(m, 0), e0[t], d[m]);
                    d[m] = mfma8(frag(m, 1), e1[t], d[m]);
                }
                requant(d, sc[kC1], sc[kInvSh1], hb);
            }
            {
                f32x4 d[4];
#pragma unroll
                for (int m = 0; m < 4; ++m) {
                    d[m] = f32x4{0, 0, 0, 0};
                    d[m] = mfma8(frag(4 + m, 0), hb[0], d[m]);
                    d[m] = mfma8(frag(4 + m, 1), hb[1], d[m]);
                }
                requant(d, sc[kC2], sc[kInvSh2], hb);
            }
            {
                f32x4 d[3];
#pragma unroll
                for (int m = 0; m < 3; ++m) {
                    d[m] = f32x4{0, 0, 0, 0};
                    d[m] = mfma8(frag(8 + m, 0), hb[0], d[m]);
                    d[m] = mfma8(frag(8 + m, 1), hb[1], d[m]);
                }
#pragma unroll
                for (int m = 0; m < 3; ++m)
#pragma unroll
                    for (int r = 0; r < 4; ++r) so[q * kGuideOut + 16 * m + 4 * g + r] = d[m][r] * sc[kC3];
            }
        } else {
            bf16x4 h[4];
            {
                f32x4 d[4];
#pragma unroll
                for (int m = 0; m < 4; ++m) {
                    d[m] = f32x4{0, 0, 0, 0};
                    d[m] = mfma16(wt_frag(w1, kWtLd, m, 0, q, g), b0[t], d[m]);
                    d[m] = mfma16(wt_frag(w1, kWtLd, m, 1, q, g), b1[t], d[m]);
                    d[m] = mfma16(wt_frag(w1, kWtLd, m, 2, q, g), b2[t], d[m]);
                }
#pragma unroll
                for (int m = 0; m < 4; ++m) h[m] = relu_bf(d[m]);
                if (MODE == 1 && vq[t])
#pragma unroll
                    for (int m = 0; m < 4; ++m) am_h1 = fmaxf(am_h1, amax4(h[m]));
            }
            {
                f32x4 d[4];
#pragma unroll
                for (int m = 0; m < 4; ++m) {
                    d[m] = f32x4{0, 0, 0, 0};
#pragma unroll
                    for (int kk = 0; kk < 4; ++kk) d[m] = mfma16(wt_frag(w2, kWtLd, m, kk, q, g), h[kk], d[m]);
                }
#pragma unroll
                for (int m = 0; m < 4; ++m) h[m] = relu_bf(d[m]);
                if (MODE == 1 && vq[t])
#pragma unroll
                    for (int m = 0; m < 4; ++m) am_h2 = fmaxf(am_h2, amax4(h[m]));
            }
            {
                f32x4 d[3];
#pragma unroll
                for (int m = 0; m < 3; ++m) {
                    d[m] = f32x4{0, 0, 0, 0};
#pragma unroll
                    for (int kk = 0; kk < 4; ++kk) d[m] = mfma16(wt_frag(w3, kWtLd, m, kk, q, g), h[kk], d[m]);
                }
#pragma unroll
                for (int m = 0; m < 3; ++m)
#pragma unroll
                    for (int r = 0; r < 4; ++r) so[q * kGuideOut + 16 * m + 4 * g + r] = d[m][r];
            }
        }
        __syncthreads();
        // head (vmf.h lobes): lane group g evaluates lobes 2g and 2g+1 of query q
        float o[10];
#pragma unroll
        for (int j = 0; j < 10; ++j) o[j] = so[q * kGuideOut + 10 * g + j];
        float mx = fmaxf(o[0], o[5]);
        mx = fmaxf(mx, __shfl_xor(mx, 16));
        mx = fmaxf(mx, __shfl_xor(mx, 32));
        float ex0 = expf(o[0] - mx), ex1 = expf(o[5] - mx);
        float sum = ex0 + ex1;
        sum += __shfl_xor(sum, 16);
        sum += __shfl_xor(sum, 32);
        if (vq[t]) {
            float4 *dst = mix + 10 * (size_t)src.slot(base + 16 * t + q);
            reinterpret_cast<float2 *>(dst)[g] = make_float2(ex0 / sum, ex1 / sum);
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                f3 r = mk(o[5 * j + 1], o[5 * j + 2], o[5 * j + 3]);
                float nr = sqrtf(sqnorm(r));
                f3 mu = nr > 1e-12f ? r / nr : mk(0.0f, 0.0f, 1.0f);
                float kap = fminf(fmaxf(expf(o[5 * j + 4]), kMinKappa), kMaxKappa);
                dst[2 + 2 * g + j] = make_float4(mu.x, mu.y, mu.z, kap);
            }
        }
        __syncthreads();
    }
    if (MODE == 1) {
        wave_amax(am_x, amax + 0);
        wave_amax(am_h1, amax + 1);
        wave_amax(am_h2, amax + 2);
    }
}

// FP8 copy of the network: per-tensor power-of-two weight scales from the FP32 master weights, input
// scales from the recorded activation maxima (with one power of two of headroom), and the weights as
// A fragments: word (m-tile, k-block, lane (g, q)) holds W^T[16 m + q][unit of k position 8 g + j],
// j = 0..7, as FP8 of W / s_w. Layer 1's k-block 0 holds the 32 grid features in the order the encoder
// produces them, k-block 1 holds normal, wo and the constant 1 in lane group 0. One block of 256.
__global__ __launch_bounds__(256) void k_guide_fp8_pack(const float *mlp, const uint32_t *act_amax, uint64_t *w8,
                                                        float *scales) {
    __shared__ float red[256];
    __shared__ float s_w[3];
    const float *wl[3] = {mlp, mlp + kMlpW1, mlp + kMlpW1 + kMlpW2};
    const int outs[3] = {64, 64, kGuideOut};
    for (int L = 0; L < 3; ++L) {
        float m = 0.0f;
        for (int i = threadIdx.x; i < 64 * outs[L]; i += 256) m = fmaxf(m, fabsf(wl[L][i]));
        red[threadIdx.x] = m;
        __syncthreads();
        for (int o = 128; o > 0; o >>= 1) {
            if (threadIdx.x < o) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + o]);
            __syncthreads();
        }
        if (threadIdx.x == 0) s_w[L] = fp8_pow2_scale(red[0]);
        __syncthreads();
    }
    float s_in[3];
    for (int L = 0; L < 3; ++L) s_in[L] = 2.0f * fp8_pow2_scale(__uint_as_float(act_amax[L]));
    if (threadIdx.x == 0) {
        for (int L = 0; L < 3; ++L) {
            scales[kInvSx + L] = 1.0f / s_in[L];
            scales[kC1 + L] = s_in[L] * s_w[L];
            scales[kSw1 + L] = s_w[L];
            scales[kSx + L] = s_in[L];
        }
    }
    for (int e = threadIdx.x; e < kFp8Frags; e += 256) {
        const int lane = e & 63, f = e >> 6, mt = f >> 1, kb = f & 1;
        const int L = mt < 4 ? 0 : (mt < 8 ? 1 : 2);
        const int m = mt - (L == 0 ? 0 : (L == 1 ? 4 : 8));
        const int g = lane >> 4, q = lane & 15, out = 16 * m + q;
        float v[8];
        for (int j = 0; j < 8; ++j) {
            int in;
            if (L == 0) in = kb == 0 ? fp8_k_unit(g, j) : (g == 0 ? 32 + j : -1);
            else in = 32 * kb + fp8_k_unit(g, j);
            v[j] = in >= 0 ? wl[L][in * outs[L] + out] : 0.0f;
        }
        w8[e] = fp8x8(v, 1.0f / s_w[L]);
    }
}

// One wave per block. Every wave loops over chunks of 64 samples, accumulates the three weight
// gradients in registers (MFMA) and scatters hash-grid gradients with FP32 atomics; at the end it writes
// its weight-gradient partials (one slot per wave) for k_guide_reduce.
struct TrainSample {
    float4 xc; // position, weight c
    float4 n;
    float4 wo;
    float4 wi;
};
static_assert(sizeof(TrainSample) == ksg::kTrainSampleBytes, "training sample layout");

__global__ __launch_bounds__(64) void k_guide_train(GuideNet net, int n, int stride, const TrainSample *samples,
                                                    float c_scale, float c_max, float *grid_grad, float *partials,
                                                    double *loss_out, uint32_t *act_amax) {
    extern __shared__ __align__(16) uint8_t smem[];
    const int lane = threadIdx.x;
    const int r16 = lane & 15, q4 = (lane >> 4) * 4;
    uint16_t *X = reinterpret_cast<uint16_t *>(smem);
    uint16_t *H1 = reinterpret_cast<uint16_t *>(smem + kActBytes);
    uint16_t *H2 = reinterpret_cast<uint16_t *>(smem + 2 * kActBytes);
    uint16_t *G = reinterpret_cast<uint16_t *>(smem + 3 * kActBytes);
    float *O = reinterpret_cast<float *>(smem + 4 * kActBytes);
    uint16_t *D2 = reinterpret_cast<uint16_t *>(O); // dH2 (BF16) re-uses O once the head is done
    float *DX = O;                                  // dX (FP32, stride 32) once dH2 is consumed

    f32x4 g3[4][3], g2[4][4], g1[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a) {
#pragma unroll
        for (int b = 0; b < 3; ++b) g3[a][b] = f32x4{0, 0, 0, 0};
#pragma unroll
        for (int b = 0; b < 4; ++b) g2[a][b] = g1[a][b] = f32x4{0, 0, 0, 0};
    }
    double loss = 0.0, wsum = 0.0, loss_cos = 0.0;
    float am[3] = {0.0f, 0.0f, 0.0f};
    const int n_chunks = (n + 63) / 64;
    for (int chunk = blockIdx.x; chunk < n_chunks; chunk += gridDim.x) {
        int i = chunk * 64 + lane;
        bool valid = i < n;
        TrainSample s{};
        if (valid) {
            s = samples[(size_t)i * stride];
            s.xc.w = fminf(s.xc.w * c_scale, c_max);
        }
        if (valid) guide_encode(net, ld3(s.xc), ld3(s.n), ld3(s.wo), X + lane * kLdx);
        else guide_zero_row(X + lane * kLdx);
        __syncthreads();
        wave_dense<4, true>(X, net.wt[0], H1, nullptr, 0);
        __syncthreads();
        wave_dense<4, true>(H1, net.wt[1], H2, nullptr, 0);
        __syncthreads();
        wave_dense<3, false>(H2, net.wt[2], nullptr, O, kGuideOut);
        __syncthreads();
        if (act_amax && valid) { // largest |input| of each layer, for the FP8 scales
            for (int k = 0; k < 64; ++k) {
                am[0] = fmaxf(am[0], fabsf(bf2f(X[lane * kLdx + k])));
                am[1] = fmaxf(am[1], fabsf(bf2f(H1[lane * kLdx + k])));
                am[2] = fmaxf(am[2], fabsf(bf2f(H2[lane * kLdx + k])));
            }
        }
        // head backward: d(-c log q)/d outputs
        float o[kGuideOut], g[kGuideOut];
#pragma unroll
        for (int k = 0; k < kGuideOut; ++k) o[k] = O[lane * kGuideOut + k];
        if (valid && s.xc.w > 0.0f) {
            f3 wi = normalized(ld3(s.wi));
            loss += (double)head_backward(o, wi, s.xc.w, g);
            wsum += (double)s.xc.w;
            float cs = fmaxf(dot(wi, ld3(s.n)), 1e-4f);
            loss_cos -= (double)s.xc.w * (double)logf(cs * kInvPi);
        } else {
#pragma unroll
            for (int k = 0; k < kGuideOut; ++k) g[k] = 0.0f;
        }
#pragma unroll
        for (int k = 0; k < kGuideOut / 2; ++k) reinterpret_cast<uint32_t *>(G + lane * kLdx)[k] = bf_pair(g[2 * k], g[2 * k + 1]);
        __syncthreads();
        // dW3 += H2^T G   (M = 64 hidden units, N = 48 outputs, K = 64 samples)
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) {
            bf16x4 bt[3];
#pragma unroll
            for (int nn = 0; nn < 3; ++nn) bt[nn] = lds_strided(G, kk * 16 + q4, nn * 16 + r16);
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                bf16x4 a = lds_strided(H2, kk * 16 + q4, m * 16 + r16);
#pragma unroll
                for (int nn = 0; nn < 3; ++nn) g3[m][nn] = mfma16(a, bt[nn], g3[m][nn]);
            }
        }
        // dH2 = (G W3^T) * relu'(H2)   (M = 64 samples, N = 64 units, K = 48 outputs); W3 is [in][out]
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            f32x4 acc[4];
#pragma unroll
            for (int t = 0; t < 4; ++t) acc[t] = f32x4{0, 0, 0, 0};
#pragma unroll
            for (int kk = 0; kk < 3; ++kk) {
                bf16x4 a = lds_a(G, m * 16 + r16, kk * 16 + q4);
#pragma unroll
                for (int t = 0; t < 4; ++t) acc[t] = mfma16(a, glb_a(net.w[2] + (t * 16 + r16) * kGuideOut + kk * 16 + q4), acc[t]);
            }
#pragma unroll
            for (int t = 0; t < 4; ++t)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    int row = m * 16 + q4 + r, col = t * 16 + r16;
                    float act = bf2f(H2[row * kLdx + col]);
                    D2[row * kLdx + col] = f2bf(act > 0.0f ? acc[t][r] : 0.0f);
                }
        }
        __syncthreads();
        // dW2 += H1^T dH2
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) {
            bf16x4 bt[4];
#pragma unroll
            for (int nn = 0; nn < 4; ++nn) bt[nn] = lds_strided(D2, kk * 16 + q4, nn * 16 + r16);
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                bf16x4 a = lds_strided(H1, kk * 16 + q4, m * 16 + r16);
#pragma unroll
                for (int nn = 0; nn < 4; ++nn) g2[m][nn] = mfma16(a, bt[nn], g2[m][nn]);
            }
        }
        // dH1 = (dH2 W2^T) * relu'(H1) -> G
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            f32x4 acc[4];
#pragma unroll
            for (int t = 0; t < 4; ++t) acc[t] = f32x4{0, 0, 0, 0};
#pragma unroll
            for (int kk = 0; kk < 4; ++kk) {
                bf16x4 a = lds_a(D2, m * 16 + r16, kk * 16 + q4);
#pragma unroll
                for (int t = 0; t < 4; ++t) acc[t] = mfma16(a, glb_a(net.w[1] + (t * 16 + r16) * 64 + kk * 16 + q4), acc[t]);
            }
#pragma unroll
            for (int t = 0; t < 4; ++t)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    int row = m * 16 + q4 + r, col = t * 16 + r16;
                    float act = bf2f(H1[row * kLdx + col]);
                    G[row * kLdx + col] = f2bf(act > 0.0f ? acc[t][r] : 0.0f);
                }
        }
        __syncthreads();
        // dW1 += X^T dH1
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) {
            bf16x4 bt[4];
#pragma unroll
            for (int nn = 0; nn < 4; ++nn) bt[nn] = lds_strided(G, kk * 16 + q4, nn * 16 + r16);
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                bf16x4 a = lds_strided(X, kk * 16 + q4, m * 16 + r16);
#pragma unroll
                for (int nn = 0; nn < 4; ++nn) g1[m][nn] = mfma16(a, bt[nn], g1[m][nn]);
            }
        }
        // dX[:, 0:32] = dH1 W1^T  (grid features only)
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            f32x4 acc[2];
            acc[0] = acc[1] = f32x4{0, 0, 0, 0};
#pragma unroll
            for (int kk = 0; kk < 4; ++kk) {
                bf16x4 a = lds_a(G, m * 16 + r16, kk * 16 + q4);
#pragma unroll
                for (int t = 0; t < 2; ++t) acc[t] = mfma16(a, glb_a(net.w[0] + (t * 16 + r16) * 64 + kk * 16 + q4), acc[t]);
            }
#pragma unroll
            for (int t = 0; t < 2; ++t)
#pragma unroll
                for (int r = 0; r < 4; ++r) DX[(m * 16 + q4 + r) * 32 + t * 16 + r16] = acc[t][r];
        }
        __syncthreads();
        // hash-grid scatter
        if (valid && s.xc.w > 0.0f) {
            float u[3];
            normalise_position(net, ld3(s.xc), u);
            for (int l = 0; l < kGuideLevels; ++l) {
                float d0 = DX[lane * 32 + 2 * l], d1 = DX[lane * 32 + 2 * l + 1];
                if (d0 == 0.0f && d1 == 0.0f) continue;
                GridCorner c;
                grid_level_corners(net, l, u, c);
                for (int k = 0; k < 8; ++k) {
                    atomicAdd(&grid_grad[2 * (size_t)c.idx[k]], c.w[k] * d0);
                    atomicAdd(&grid_grad[2 * (size_t)c.idx[k] + 1], c.w[k] * d1);
                }
            }
        }
        __syncthreads();
    }
    // weight-gradient partials of this wave in [in][out] order
    float *dst = partials + (size_t)blockIdx.x * kMlpParams;
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
        for (int nn = 0; nn < 4; ++nn)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                int row = m * 16 + q4 + r, col = nn * 16 + r16;
                dst[row * 64 + col] = g1[m][nn][r];
                dst[kMlpW1 + row * 64 + col] = g2[m][nn][r];
                if (nn < 3) dst[kMlpW1 + kMlpW2 + row * kGuideOut + col] = g3[m][nn][r];
            }
    for (int off = 32; off > 0; off >>= 1) {
        loss += __shfl_xor(loss, off);
        wsum += __shfl_xor(wsum, off);
        loss_cos += __shfl_xor(loss_cos, off);
    }
    if (lane == 0) { // sum c (-log q), sum c, sum c (-log(cos / pi)) of the step
        atomicAdd(loss_out, loss);
        atomicAdd(loss_out + 1, wsum);
        atomicAdd(loss_out + 2, loss_cos);
    }
    if (act_amax) {
        wave_amax(am[0], act_amax + 0);
        wave_amax(am[1], act_amax + 1);
        wave_amax(am[2], act_amax + 2);
    }
}

__global__ void k_guide_reduce(const float *partials, int n_waves, float *grad_mlp) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= kMlpParams) return;
    float s = 0.0f;
    for (int w = 0; w < n_waves; ++w) s += partials[(size_t)w * kMlpParams + p];
    grad_mlp[p] = s;
}

// Adam (gpu/shaders/adam.slang:39-50) over every parameter; refreshes the low-precision copies and
// clears the gradient for the next step. Grid parameters are processed in FP16 pairs.
__global__ void k_guide_adam(float *params, float *grad, float *m1, float *m2, int64_t n_grid, float lr, float beta1,
                             float beta2, float eps, float bc1, float bc2, uint32_t *grid_h, uint16_t *w_bf,
                             uint16_t *wt_bf) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t n_pairs = n_grid / 2;
    auto step = [&](int64_t i) {
        float g = grad[i];
        float m = m1[i], v = m2[i];
        float mt = g + beta1 * (m - g);     // lerp(g, m, beta1)
        float vt = g * g + beta2 * (v - g * g); // lerp(g*g, v, beta2)
        float p = params[i] - lr * (mt / bc1) / (sqrtf(vt / bc2) + eps);
        params[i] = p;
        m1[i] = mt;
        m2[i] = vt;
        grad[i] = 0.0f;
        return p;
    };
    if (t < n_pairs) {
        float a = step(2 * t), b = step(2 * t + 1);
        uint32_t ha = float_to_half_bits(a), hb = float_to_half_bits(b);
        grid_h[t] = ha | (hb << 16);
        return;
    }
    int64_t j = t - n_pairs;
    if (j >= kMlpParams) return;
    float p = step(n_grid + j);
    uint16_t h = f2bf(p);
    w_bf[j] = h;
    // transposed copy: layer sizes [in][out] = 64x64, 64x64, 64x48
    int64_t base = j < kMlpW1 ? 0 : (j < kMlpW1 + kMlpW2 ? kMlpW1 : kMlpW1 + kMlpW2);
    int64_t outs = j < kMlpW1 + kMlpW2 ? 64 : kGuideOut;
    int64_t local = j - base, in = local / outs, out = local % outs;
    wt_bf[base + out * 64 + in] = h;
}

__global__ void k_guide_refresh(const float *params, int64_t n_grid, uint32_t *grid_h, uint16_t *w_bf, uint16_t *wt_bf) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t n_pairs = n_grid / 2;
    if (t < n_pairs) {
        uint32_t ha = float_to_half_bits(params[2 * t]);
        uint32_t hb = float_to_half_bits(params[2 * t + 1]);
        grid_h[t] = ha | (hb << 16);
        return;
    }
    int64_t j = t - n_pairs;
    if (j >= kMlpParams) return;
    uint16_t h = f2bf(params[n_grid + j]);
    w_bf[j] = h;
    int64_t base = j < kMlpW1 ? 0 : (j < kMlpW1 + kMlpW2 ? kMlpW1 : kMlpW1 + kMlpW2);
    int64_t outs = j < kMlpW1 + kMlpW2 ? 64 : kGuideOut;
    int64_t local = j - base, in = local / outs, out = local % outs;
    wt_bf[base + out * 64 + in] = h;
}

__device__ __forceinline__ uint32_t pcg_hash(uint32_t v) {
    uint32_t state = v * 747796405u + 2891336453u;
    uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (word >> 22u) ^ word;
}
__device__ __forceinline__ float hash_unit(uint64_t i, uint32_t seed) {
    uint32_t h = pcg_hash((uint32_t)i ^ pcg_hash((uint32_t)(i >> 32) + seed));
    return (float)(h >> 8) * 0x1p-24f;
}

// Grid features U(-1e-4, 1e-4); hidden layers He-uniform; the output layer scaled by 0.1 so the initial
// mixture is close to uniform lobes of small kappa.
__global__ void k_guide_init(float *params, float *m1, float *m2, float *grad, int64_t n_grid, uint32_t seed) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_grid + kMlpParams) return;
    float u = hash_unit((uint64_t)i, seed) * 2.0f - 1.0f;
    float v;
    if (i < n_grid) {
        v = 1e-4f * u;
    } else {
        int64_t j = i - n_grid;
        float lim = sqrtf(6.0f / 64.0f);
        v = u * (j < kMlpW1 + kMlpW2 ? lim : 0.1f * lim);
    }
    params[i] = v;
    m1[i] = 0.0f;
    m2[i] = 0.0f;
    grad[i] = 0.0f;
}

__global__ void k_vmm_pdf(int64_t n, const float *mix, const float *dirs, float *pdf) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Mixture m;
    const float *e = mix + 40 * i;
    for (int k = 0; k < kGuideLobes; ++k) {
        m.w[k] = e[k];
        m.lobe[k].mu = mk(e[8 + 4 * k], e[9 + 4 * k], e[10 + 4 * k]);
        m.lobe[k].kappa = e[11 + 4 * k];
    }
    pdf[i] = mixture_pdf(m, mk(dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2]));
}

__global__ void k_vmm_sample(int64_t n, const float *mix, const float *u, float *dirs, float *pdf) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Mixture m;
    const float *e = mix + 40 * i;
    for (int k = 0; k < kGuideLobes; ++k) {
        m.w[k] = e[k];
        m.lobe[k].mu = mk(e[8 + 4 * k], e[9 + 4 * k], e[10 + 4 * k]);
        m.lobe[k].kappa = e[11 + 4 * k];
    }
    float ur;
    int sel = mixture_select(m.w, u[2 * i], &ur);
    Lobe l{mk(e[8 + 4 * sel], e[9 + 4 * sel], e[10 + 4 * sel]), e[11 + 4 * sel]};
    f3 w = vmf_sample(l, ur, u[2 * i + 1]);
    dirs[3 * i] = w.x, dirs[3 * i + 1] = w.y, dirs[3 * i + 2] = w.z;
    pdf[i] = mixture_pdf(m, w);
}

// One thread per record slot: work item w = slot / kRecVertices. The final sample value of a pass with
// one sample per pixel sits at samp[w].
__global__ void k_guide_extract(const float4 *recs, const float4 *samp, uint32_t n_items, uint32_t stamp,
                                TrainSample *out, uint32_t *count, double *c_sum, uint32_t cap) {
    const uint32_t slot = blockIdx.x * blockDim.x + threadIdx.x;
    bool take = false;
    TrainSample t{};
    if (slot < n_items * (uint32_t)kRecVertices) {
        const float4 *r = recs + (size_t)slot * kRecFloat4;
        float4 x = r[0];
        if (__float_as_uint(x.w) == stamp) {
            f3 Lf = ld3(samp[slot / kRecVertices]);
            f3 d = Lf - ld3(r[4]) - ld3(r[kRecNee]);
            f3 b = ld3(r[5]);
            f3 X = mk(b.x > 0.0f ? d.x / b.x : 0.0f, b.y > 0.0f ? d.y / b.y : 0.0f, b.z > 0.0f ? d.z / b.z : 0.0f);
            float c = luminance(X);
            if (c > 0.0f && isfinite(c)) {
                take = true;
                t.xc = make_float4(x.x, x.y, x.z, c);
                t.n = r[1];
                t.wo = r[2];
                t.wi = r[3];
            }
        }
    }
    double cs = take ? (double)t.xc.w : 0.0;
    for (int off = 32; off > 0; off >>= 1) cs += __shfl_xor(cs, off);
    if (__lane_id() == 0 && cs != 0.0) atomicAdd(c_sum, cs);
    uint32_t j = wave_append(count, take);
    if (take && j < cap) {
        TrainSample *o = out + j;
        o->xc = t.xc;
        o->n = t.n;
        o->wo = t.wo;
        o->wi = t.wi;
    }
}

struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    hipError_t alloc(size_t n) {
        release();
        hipError_t e = hipMalloc(&p, std::max<size_t>(n, 16));
        if (e == hipSuccess) {
            bytes = n;
        } else {
            p = nullptr;
            (void)hipGetLastError(); // reported through the return value only
        }
        return e;
    }
    hipError_t ensure(size_t n) { return n <= bytes ? hipSuccess : alloc(n); }
    template <typename T> T *as() const { return reinterpret_cast<T *>(p); }
};

} // namespace

struct ks_guide {
    ks_device *dev = nullptr;
    hipStream_t stream = nullptr;
    ks_guide_config cfg{};
    int64_t n_grid = 0;
    uint32_t log2_T = 19;
    uint32_t res[kGuideLevels];
    float lo[3], inv_ext[3];
    DevBuf params, grad, m1, m2, grid_h, w_bf, wt_bf, partials, loss, samples;
    DevBuf w8, fp8s, act_amax; // FP8 copy of the MLP, its scales, recorded activation maxima
    bool fp8 = false;          // inference in FP8
    int32_t n_partials = 0;
    int64_t steps = 0;
    int64_t samples_used = 0; // training samples consumed by all steps
};

namespace ksg {

GuideNet net_of(const ks_guide *g) {
    GuideNet n{};
    n.grid_h = g->grid_h.as<uint32_t>();
    const uint16_t *w = g->w_bf.as<uint16_t>(), *wt = g->wt_bf.as<uint16_t>();
    n.w[0] = w;
    n.w[1] = w + kMlpW1;
    n.w[2] = w + kMlpW1 + kMlpW2;
    n.wt[0] = wt;
    n.wt[1] = wt + kMlpW1;
    n.wt[2] = wt + kMlpW1 + kMlpW2;
    n.w8 = g->w8.as<uint64_t>();
    n.fp8 = g->fp8s.as<float>();
    n.log2_T = g->log2_T;
    for (int l = 0; l < kGuideLevels; ++l) n.res[l] = g->res[l];
    for (int k = 0; k < 3; ++k) {
        n.lo[k] = g->lo[k];
        n.inv_ext[k] = g->inv_ext[k];
    }
    return n;
}

ks_device *guide_device(const ks_guide *g) { return g->dev; }
float guide_fraction(const ks_guide *g) { return g->cfg.guide_fraction; }
int32_t guide_max_bounce(const ks_guide *g) {
    return g->cfg.max_guided_bounce > 0 ? std::min(g->cfg.max_guided_bounce, 255) : 255;
}
int32_t guide_batch(const ks_guide *g) { return g->cfg.train_batch; }
int32_t guide_max_steps(const ks_guide *g) { return g->cfg.max_steps_per_pass; }
int32_t guide_train_spp(const ks_guide *g) { return g->cfg.train_spp; }

int launch_infer(ks_guide *g, const DevScene &sc, const uint32_t *n_dev, uint32_t n_max, const float4 *rd,
                 const float4 *hits, const uint32_t *state, int32_t max_bounce, float4 *mix, uint32_t *list,
                 uint32_t *n_list, hipStream_t s) {
    unsigned blocks = (n_max + kInferBlock - 1) / kInferBlock;
    const QueryHits qh{sc, n_dev, rd, hits, state, max_bounce};
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(1), 0, s, n_list);
    hipLaunchKernelGGL(k_guide_compact, dim3((n_max + 255) / 256), dim3(256), 0, s, qh, list, n_list);
    const QueryList ql{qh, list, n_list};
    if (g->fp8)
        hipLaunchKernelGGL((k_guide_infer<QueryList, 2>), dim3(blocks), dim3(kInferBlock), 0, s, net_of(g), ql, mix,
                           nullptr);
    else
        hipLaunchKernelGGL((k_guide_infer<QueryList, 0>), dim3(blocks), dim3(kInferBlock), 0, s, net_of(g), ql, mix,
                           nullptr);
    KS_HIP(hipGetLastError());
    return KS_OK;
}

// The FP8 copy of the network from the current master weights and recorded activation maxima.
int fp8_pack(ks_guide *g, hipStream_t s) {
    hipLaunchKernelGGL(k_guide_fp8_pack, dim3(1), dim3(256), 0, s, g->params.as<float>() + g->n_grid,
                       g->act_amax.as<uint32_t>(), g->w8.as<uint64_t>(), g->fp8s.as<float>());
    KS_HIP(hipGetLastError());
    return KS_OK;
}

int64_t guide_steps(const ks_guide *g) { return g->steps; }

int launch_extract(const float4 *recs, const float4 *samp, uint32_t n_items, uint32_t stamp, void *out,
                   uint32_t *count, double *c_sum, uint32_t cap, hipStream_t s) {
    uint64_t slots = (uint64_t)n_items * kRecVertices;
    hipLaunchKernelGGL(k_guide_extract, dim3((unsigned)((slots + 255) / 256)), dim3(256), 0, s, recs, samp, n_items,
                       stamp, static_cast<TrainSample *>(out), count, c_sum, cap);
    KS_HIP(hipGetLastError());
    return KS_OK;
}

// One Adam step on n device-resident samples (TrainSample records, every stride-th one).
int train_on_device(ks_guide *g, const void *samples, int n, int stride, float c_scale, double *loss_host,
                    hipStream_t s) {
    if (n <= 0) return KS_OK;
    int n_chunks = (n + 63) / 64;
    int waves = std::min(n_chunks, std::max(1, ksg::device_cu_count(g->dev) * 3));
    KS_HIP(g->partials.ensure(sizeof(float) * (size_t)waves * kMlpParams));
    KS_HIP(hipMemsetAsync(g->loss.p, 0, 3 * sizeof(double), s));
    // with FP8 inference the step's forward records the activation maxima that scale the next FP8 copy
    if (g->fp8) KS_HIP(hipMemsetAsync(g->act_amax.p, 0, 3 * sizeof(uint32_t), s));
    const float c_max = c_scale > 0.0f && g->cfg.weight_clamp > 0.0f ? g->cfg.weight_clamp : INFINITY;
    hipLaunchKernelGGL(k_guide_train, dim3(waves), dim3(64), kTrainWaveBytes, s, net_of(g), n, stride,
                       static_cast<const TrainSample *>(samples), c_scale > 0.0f ? c_scale : 1.0f, c_max,
                       g->grad.as<float>(), g->partials.as<float>(), g->loss.as<double>(),
                       g->fp8 ? g->act_amax.as<uint32_t>() : nullptr);
    KS_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_guide_reduce, dim3((kMlpParams + 255) / 256), dim3(256), 0, s, g->partials.as<float>(), waves,
                       g->grad.as<float>() + g->n_grid);
    g->steps += 1;
    g->samples_used += n;
    float bc1 = 1.0f - std::pow(g->cfg.beta1, (float)g->steps), bc2 = 1.0f - std::pow(g->cfg.beta2, (float)g->steps);
    int64_t work = g->n_grid / 2 + kMlpParams;
    hipLaunchKernelGGL(k_guide_adam, dim3((unsigned)((work + 255) / 256)), dim3(256), 0, s, g->params.as<float>(),
                       g->grad.as<float>(), g->m1.as<float>(), g->m2.as<float>(), g->n_grid, g->cfg.learning_rate,
                       g->cfg.beta1, g->cfg.beta2, g->cfg.eps, bc1, bc2, g->grid_h.as<uint32_t>(),
                       g->w_bf.as<uint16_t>(), g->wt_bf.as<uint16_t>());
    KS_HIP(hipGetLastError());
    if (g->fp8) {
        int e = fp8_pack(g, s);
        if (e) return e;
    }
    if (loss_host) {
        KS_HIP(hipMemcpyAsync(loss_host, g->loss.p, sizeof(double), hipMemcpyDeviceToHost, s));
        KS_HIP(hipStreamSynchronize(s));
    }
    return KS_OK;
}

} // namespace ksg

// ---------------------------------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------------------------------
extern "C" {

int ks_guide_create(ks_device *dev, const ks_guide_config *cfg_in, const float bound_lo[3], const float bound_hi[3],
                    ks_guide **out) {
    return ksa::guarded([&]() -> int {
        KS_REQUIRE(dev && bound_lo && bound_hi && out, "null argument");
        *out = nullptr;
        auto g = std::make_unique<ks_guide>();
        g->dev = dev;
        g->stream = ksg::device_stream(dev);
        ks_guide_config c{};
        if (cfg_in) c = *cfg_in;
        if (c.log2_table_size == 0) c.log2_table_size = 19;
        if (c.base_resolution == 0.0f) c.base_resolution = 16.0f;
        if (c.max_resolution == 0.0f) c.max_resolution = 256.0f;
        if (c.learning_rate == 0.0f) c.learning_rate = 1e-2f;
        if (c.beta1 == 0.0f) c.beta1 = 0.9f;
        if (c.beta2 == 0.0f) c.beta2 = 0.999f;
        if (c.eps == 0.0f) c.eps = 1e-8f;
        if (c.guide_fraction == 0.0f) c.guide_fraction = 0.5f;
        if (c.train_batch == 0) c.train_batch = 1 << 18;
        if (c.max_steps_per_pass == 0) c.max_steps_per_pass = 4;
        if (c.weight_clamp == 0.0f) c.weight_clamp = 20.0f;
        KS_REQUIRE(c.log2_table_size >= 10 && c.log2_table_size <= 24, "log2_table_size must be in [10, 24]");
        KS_REQUIRE(c.max_resolution >= c.base_resolution && c.base_resolution >= 2.0f, "bad grid resolutions");
        KS_REQUIRE(c.guide_fraction >= 0.0f && c.guide_fraction < 1.0f, "guide_fraction must be in [0, 1)");
        KS_REQUIRE(c.train_spp >= 0, "train_spp must be non-negative");
        g->cfg = c;
        g->log2_T = (uint32_t)c.log2_table_size;
        // resolutions: floor(base * b^l), b = (max / base)^(1 / (L - 1)), in FP64
        double growth = std::exp(std::log((double)c.max_resolution / (double)c.base_resolution) / (kGuideLevels - 1));
        for (int l = 0; l < kGuideLevels; ++l)
            g->res[l] = (uint32_t)std::floor((double)c.base_resolution * std::pow(growth, (double)l) + 1e-9);
        for (int k = 0; k < 3; ++k) {
            float ext = bound_hi[k] - bound_lo[k];
            KS_REQUIRE(ext >= 0.0f, "bound_hi below bound_lo");
            float pad = std::max(ext * 0.01f, 1e-4f);
            g->lo[k] = bound_lo[k] - pad;
            g->inv_ext[k] = 1.0f / (ext + 2.0f * pad);
        }
        g->n_grid = (int64_t)kGuideLevels * ((int64_t)1 << g->log2_T) * kGuideFeatures;
        int64_t np = g->n_grid + kMlpParams;
        KS_HIP(hipSetDevice(ksg::device_ordinal(dev)));
        KS_HIP(g->params.alloc(sizeof(float) * np));
        KS_HIP(g->grad.alloc(sizeof(float) * np));
        KS_HIP(g->m1.alloc(sizeof(float) * np));
        KS_HIP(g->m2.alloc(sizeof(float) * np));
        KS_HIP(g->grid_h.alloc(sizeof(uint32_t) * g->n_grid / 2));
        KS_HIP(g->w_bf.alloc(sizeof(uint16_t) * kMlpParams));
        KS_HIP(g->wt_bf.alloc(sizeof(uint16_t) * kMlpParams));
        KS_HIP(g->loss.alloc(3 * sizeof(double)));
        KS_HIP(g->w8.alloc(sizeof(uint64_t) * kFp8Frags));
        KS_HIP(g->fp8s.alloc(sizeof(float) * kFp8ScaleCount));
        KS_HIP(g->act_amax.alloc(sizeof(uint32_t) * 3));
        {
            const float one = 1.0f; // activation maxima before any training or calibration
            uint32_t bits[3];
            std::memcpy(&bits[0], &one, 4);
            bits[1] = bits[2] = bits[0];
            KS_HIP(hipMemcpyAsync(g->act_amax.p, bits, sizeof(bits), hipMemcpyHostToDevice, g->stream));
            KS_HIP(hipStreamSynchronize(g->stream));
        }
        g->fp8 = c.inference_fp8 != 0;
        hipLaunchKernelGGL(k_guide_init, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, g->stream,
                           g->params.as<float>(), g->m1.as<float>(), g->m2.as<float>(), g->grad.as<float>(), g->n_grid,
                           (uint32_t)(c.seed ^ (c.seed >> 32)));
        int64_t work = g->n_grid / 2 + kMlpParams;
        hipLaunchKernelGGL(k_guide_refresh, dim3((unsigned)((work + 255) / 256)), dim3(256), 0, g->stream,
                           g->params.as<float>(), g->n_grid, g->grid_h.as<uint32_t>(), g->w_bf.as<uint16_t>(),
                           g->wt_bf.as<uint16_t>());
        KS_HIP(hipGetLastError());
        {
            int e = ksg::fp8_pack(g.get(), g->stream);
            if (e) return e;
        }
        KS_HIP(hipStreamSynchronize(g->stream));
        *out = g.release();
        return KS_OK;
    });
}

int ks_guide_free(ks_guide *g) {
    return ksa::guarded([&]() -> int {
        if (g) {
            (void)hipStreamSynchronize(g->stream);
            delete g;
        }
        return KS_OK;
    });
}

int ks_guide_set_mode(ks_guide *g, float guide_fraction, int32_t max_steps_per_pass) {
    return ksa::guarded([&]() -> int {
        KS_REQUIRE(g, "guide is null");
        KS_REQUIRE(guide_fraction >= 0.0f && guide_fraction < 1.0f, "guide_fraction must be in [0, 1)");
        g->cfg.guide_fraction = guide_fraction;
        g->cfg.max_steps_per_pass = max_steps_per_pass;
        return KS_OK;
    });
}

int ks_guide_get_stats(const ks_guide *g, ks_guide_stats *out) {
    return ksa::guarded([&]() -> int {
        KS_REQUIRE(g && out, "null argument");
        double acc[3] = {0, 0, 0};
        if (g->steps > 0) {
            KS_HIP(hipMemcpyAsync(acc, g->loss.p, sizeof(acc), hipMemcpyDeviceToHost, g->stream));
            KS_HIP(hipStreamSynchronize(g->stream));
        }
        out->steps = g->steps;
        out->samples = g->samples_used;
        out->last_loss = acc[0];
        out->last_weight = acc[1];
        out->last_loss_cosine = acc[2];
        return KS_OK;
    });
}

int ks_guide_param_count(const ks_guide *g, int64_t *n_grid, int64_t *n_mlp) {
    return ksa::guarded([&]() -> int {
        KS_REQUIRE(g, "guide is null");
        if (n_grid) *n_grid = g->n_grid;
        if (n_mlp) *n_mlp = kMlpParams;
        return KS_OK;
    });
}

int ks_guide_grid_info(const ks_guide *g, int32_t res[16], float lo[3], float inv_ext[3]) {
    return ksa::guarded([&]() -> int {
        KS_REQUIRE(g, "guide is null");
        for (int l = 0; l < kGuideLevels && res; ++l) res[l] = (int32_t)g->res[l];
        for (int k = 0; k < 3; ++k) {
            if (lo) lo[k] = g->lo[k];
            if (inv_ext) inv_ext[k] = g->inv_ext[k];
        }
        return KS_OK;
    });
}

int ks_guide_get_params(const ks_guide *g, float *params) {
    return ksa::guarded([&]() -> int {
        KS_REQUIRE(g && params, "null argument");
        KS_HIP(hipMemcpyAsync(params, g->params.p, sizeof(float) * (g->n_grid + kMlpParams), hipMemcpyDeviceToHost,
                              g->stream));
        KS_HIP(hipStreamSynchronize(g->stream));
        return KS_OK;
    });
}

int ks_guide_set_params(ks_guide *g, const float *params) {
    return ksa::guarded([&]() -> int {
        KS_REQUIRE(g && params, "null argument");
        KS_HIP(hipMemcpyAsync(g->params.p, params, sizeof(float) * (g->n_grid + kMlpParams), hipMemcpyHostToDevice,
                              g->stream));
        int64_t work = g->n_grid / 2 + kMlpParams;
        hipLaunchKernelGGL(k_guide_refresh, dim3((unsigned)((work + 255) / 256)), dim3(256), 0, g->stream,
                           g->params.as<float>(), g->n_grid, g->grid_h.as<uint32_t>(), g->w_bf.as<uint16_t>(),
                           g->wt_bf.as<uint16_t>());
        KS_HIP(hipGetLastError());
        int e = ksg::fp8_pack(g, g->stream);
        if (e) return e;
        KS_HIP(hipStreamSynchronize(g->stream));
        return KS_OK;
    });
}

int ks_guide_forward(ks_guide *g, int64_t n, const float *queries, float *mixtures) {
    return ksa::guarded([&]() -> int {
        KS_REQUIRE(g && queries && mixtures && n > 0 && n < (1ll << 30), "invalid argument");
        hipStream_t s = g->stream;
        DevBuf q, m;
        KS_HIP(q.alloc(sizeof(float) * 9 * n));
        KS_HIP(m.alloc(sizeof(float4) * 10 * n));
        KS_HIP(hipMemcpyAsync(q.p, queries, sizeof(float) * 9 * n, hipMemcpyHostToDevice, s));
        unsigned blocks = (unsigned)((n + kInferBlock - 1) / kInferBlock);
        if (g->fp8)
            hipLaunchKernelGGL((k_guide_infer<QueryArray, 2>), dim3(blocks), dim3(kInferBlock), 0, s, ksg::net_of(g),
                               QueryArray{q.as<float>(), (uint32_t)n}, m.as<float4>(), nullptr);
        else
            hipLaunchKernelGGL((k_guide_infer<QueryArray, 0>), dim3(blocks), dim3(kInferBlock), 0, s, ksg::net_of(g),
                               QueryArray{q.as<float>(), (uint32_t)n}, m.as<float4>(), nullptr);
        KS_HIP(hipGetLastError());
        std::vector<float4> h((size_t)n * 10);
        KS_HIP(hipMemcpyAsync(h.data(), m.p, sizeof(float4) * 10 * n, hipMemcpyDeviceToHost, s));
        KS_HIP(hipStreamSynchronize(s));
        for (int64_t i = 0; i < n; ++i) { // [w0..w7, (mu xyz, kappa) x 8]
            const float4 *e = &h[10 * i];
            float *o = mixtures + 40 * i;
            o[0] = e[0].x, o[1] = e[0].y, o[2] = e[0].z, o[3] = e[0].w;
            o[4] = e[1].x, o[5] = e[1].y, o[6] = e[1].z, o[7] = e[1].w;
            for (int k = 0; k < 8; ++k) {
                o[8 + 4 * k] = e[2 + k].x;
                o[9 + 4 * k] = e[2 + k].y;
                o[10 + 4 * k] = e[2 + k].z;
                o[11 + 4 * k] = e[2 + k].w;
            }
        }
        return KS_OK;
    });
}

static int upload_samples(ks_guide *g, int64_t n, const float *queries, const float *wi, const float *weight) {
    std::vector<TrainSample> h((size_t)n);
    for (int64_t i = 0; i < n; ++i) {
        const float *q = queries + 9 * i;
        h[i].xc = make_float4(q[0], q[1], q[2], weight[i]);
        h[i].n = make_float4(q[3], q[4], q[5], 0.0f);
        h[i].wo = make_float4(q[6], q[7], q[8], 0.0f);
        h[i].wi = make_float4(wi[3 * i], wi[3 * i + 1], wi[3 * i + 2], 0.0f);
    }
    KS_HIP(g->samples.ensure(sizeof(TrainSample) * n));
    KS_HIP(hipMemcpyAsync(g->samples.p, h.data(), sizeof(TrainSample) * n, hipMemcpyHostToDevice, g->stream));
    KS_HIP(hipStreamSynchronize(g->stream));
    return KS_OK;
}

int ks_guide_gradient(ks_guide *g, int64_t n, const float *queries, const float *wi, const float *weight,
                      double *loss, float *grad) {
    return ksa::guarded([&]() -> int {
        KS_REQUIRE(g && queries && wi && weight && grad && n > 0 && n < (1ll << 30), "invalid argument");
        int e = upload_samples(g, n, queries, wi, weight);
        if (e) return e;
        hipStream_t s = g->stream;
        int n_chunks = (int)((n + 63) / 64);
        int waves = std::min(n_chunks, std::max(1, ksg::device_cu_count(g->dev) * 3));
        KS_HIP(g->partials.ensure(sizeof(float) * (size_t)waves * kMlpParams));
        KS_HIP(hipMemsetAsync(g->loss.p, 0, 3 * sizeof(double), s));
        KS_HIP(hipMemsetAsync(g->grad.p, 0, sizeof(float) * g->n_grid, s));
        hipLaunchKernelGGL(k_guide_train, dim3(waves), dim3(64), kTrainWaveBytes, s, ksg::net_of(g), (int)n, 1,
                           g->samples.as<TrainSample>(), 1.0f, INFINITY, g->grad.as<float>(), g->partials.as<float>(),
                           g->loss.as<double>(), nullptr);
        KS_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_guide_reduce, dim3((kMlpParams + 255) / 256), dim3(256), 0, s, g->partials.as<float>(), waves,
                           g->grad.as<float>() + g->n_grid);
        KS_HIP(hipGetLastError());
        KS_HIP(hipMemcpyAsync(grad, g->grad.p, sizeof(float) * (g->n_grid + kMlpParams), hipMemcpyDeviceToHost, s));
        if (loss) KS_HIP(hipMemcpyAsync(loss, g->loss.p, sizeof(double), hipMemcpyDeviceToHost, s));
        KS_HIP(hipMemsetAsync(g->grad.p, 0, sizeof(float) * (g->n_grid + kMlpParams), s));
        KS_HIP(hipStreamSynchronize(s));
        return KS_OK;
    });
}

int ks_guide_train_step(ks_guide *g, int64_t n, const float *queries, const float *wi, const float *weight,
                        double *loss) {
    return ksa::guarded([&]() -> int {
        KS_REQUIRE(g && queries && wi && weight && n > 0 && n < (1ll << 30), "invalid argument");
        int e = upload_samples(g, n, queries, wi, weight);
        if (e) return e;
        return ksg::train_on_device(g, g->samples.p, (int)n, 1, 0.0f, loss, g->stream);
    });
}

int ks_guide_set_inference(ks_guide *g, int32_t fp8) {
    return ksa::guarded([&]() -> int {
        KS_REQUIRE(g, "guide is null");
        g->fp8 = fp8 != 0;
        g->cfg.inference_fp8 = fp8 != 0;
        int e = ksg::fp8_pack(g, g->stream);
        if (e) return e;
        KS_HIP(hipStreamSynchronize(g->stream));
        return KS_OK;
    });
}

int ks_guide_calibrate_fp8(ks_guide *g, int64_t n, const float *queries) {
    return ksa::guarded([&]() -> int {
        KS_REQUIRE(g && queries && n > 0 && n < (1ll << 30), "invalid argument");
        hipStream_t s = g->stream;
        DevBuf q, m;
        KS_HIP(q.alloc(sizeof(float) * 9 * n));
        KS_HIP(m.alloc(sizeof(float4) * 10 * n));
        KS_HIP(hipMemcpyAsync(q.p, queries, sizeof(float) * 9 * n, hipMemcpyHostToDevice, s));
        KS_HIP(hipMemsetAsync(g->act_amax.p, 0, 3 * sizeof(uint32_t), s));
        unsigned blocks = (unsigned)((n + kInferBlock - 1) / kInferBlock);
        hipLaunchKernelGGL((k_guide_infer<QueryArray, 1>), dim3(blocks), dim3(kInferBlock), 0, s, ksg::net_of(g),
                           QueryArray{q.as<float>(), (uint32_t)n}, m.as<float4>(), g->act_amax.as<uint32_t>());
        KS_HIP(hipGetLastError());
        int e = ksg::fp8_pack(g, s);
        if (e) return e;
        KS_HIP(hipStreamSynchronize(s));
        return KS_OK;
    });
}

int ks_guide_fp8_scales(const ks_guide *g, float scales[12]) {
    return ksa::guarded([&]() -> int {
        KS_REQUIRE(g && scales, "null argument");
        KS_HIP(hipMemcpyAsync(scales, g->fp8s.p, sizeof(float) * kFp8ScaleCount, hipMemcpyDeviceToHost, g->stream));
        KS_HIP(hipStreamSynchronize(g->stream));
        return KS_OK;
    });
}

int ks_vmm_pdf(ks_device *dev, int64_t n, const float *mixtures, const float *dirs, float *pdf) {
    return ksa::guarded([&]() -> int {
        KS_REQUIRE(dev && mixtures && dirs && pdf && n > 0, "invalid argument");
        hipStream_t s = ksg::device_stream(dev);
        DevBuf m, d, p;
        KS_HIP(m.alloc(sizeof(float) * 40 * n));
        KS_HIP(d.alloc(sizeof(float) * 3 * n));
        KS_HIP(p.alloc(sizeof(float) * n));
        KS_HIP(hipMemcpyAsync(m.p, mixtures, sizeof(float) * 40 * n, hipMemcpyHostToDevice, s));
        KS_HIP(hipMemcpyAsync(d.p, dirs, sizeof(float) * 3 * n, hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(k_vmm_pdf, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, m.as<float>(), d.as<float>(),
                           p.as<float>());
        KS_HIP(hipGetLastError());
        KS_HIP(hipMemcpyAsync(pdf, p.p, sizeof(float) * n, hipMemcpyDeviceToHost, s));
        KS_HIP(hipStreamSynchronize(s));
        return KS_OK;
    });
}

int ks_vmm_sample(ks_device *dev, int64_t n, const float *mixtures, const float *u, float *dirs, float *pdf) {
    return ksa::guarded([&]() -> int {
        KS_REQUIRE(dev && mixtures && u && dirs && pdf && n > 0, "invalid argument");
        hipStream_t s = ksg::device_stream(dev);
        DevBuf m, uu, d, p;
        KS_HIP(m.alloc(sizeof(float) * 40 * n));
        KS_HIP(uu.alloc(sizeof(float) * 2 * n));
        KS_HIP(d.alloc(sizeof(float) * 3 * n));
        KS_HIP(p.alloc(sizeof(float) * n));
        KS_HIP(hipMemcpyAsync(m.p, mixtures, sizeof(float) * 40 * n, hipMemcpyHostToDevice, s));
        KS_HIP(hipMemcpyAsync(uu.p, u, sizeof(float) * 2 * n, hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(k_vmm_sample, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, m.as<float>(),
                           uu.as<float>(), d.as<float>(), p.as<float>());
        KS_HIP(hipGetLastError());
        KS_HIP(hipMemcpyAsync(dirs, d.p, sizeof(float) * 3 * n, hipMemcpyDeviceToHost, s));
        KS_HIP(hipMemcpyAsync(pdf, p.p, sizeof(float) * n, hipMemcpyDeviceToHost, s));
        KS_HIP(hipStreamSynchronize(s));
        return KS_OK;
    });
}

} // extern "C"

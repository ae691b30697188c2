// Device-side records shared by the render kernels (ks_api.hip) and the guide kernels (ks_guide.hip):
// the scene view, the wavefront control block, wave-level append/reduce helpers and hit reconstruction.
#pragma once
#include "ks_device.h"
#include "ks_sky.h"
#include "ks_texture.h"

namespace ksr {
using namespace ksd;

// Material classes of sorted shading: each class's paths are shaded by a kernel compiled for the BSDF
// lobes that class can have (kShadeMiss: rays that left the scene).
constexpr int kShadeMiss = 0, kShadeDielectric = 1, kShadeMetal = 2, kShadeGeneral = 3;
constexpr int kShadeClasses = 4;
constexpr int kShadeUnsorted = -1;

// A counter that every block of a launch updates with an atomic sits alone on a 128-byte line.
struct alignas(128) LineCounter {
    uint32_t v;
};

struct Ctrl {
    uint32_t n_active; // paths in the current buffer, [0, n_active)
    uint32_t iteration;
    uint64_t work_next, work_end;
    uint32_t n_shadow_trace; // shadow rays of the previous shading pass, traced by this iteration's traversal
    uint32_t none;           // always 0: the closest-hit ray count of a traversal of shadow rays alone
    LineCounter n_next;   // survivors appended by shade into the next buffer
    LineCounter n_shadow; // shadow rays appended by shade
    LineCounter n_class[kShadeClasses]; // paths of the current buffer per material class (sorted shading)
};

enum StatSlot {
    kRaysClosest = 0,
    kRaysShadow,
    kNodesClosest,
    kTrisClosest,
    kNodesShadow,
    kTrisShadow,
    kNonFinite,
    kSpills,
    kRaysTail,
    kStatCount
};

struct DevScene {
    const float4 *nodes;
    const float4 *tris;
    const float4 *verts;
    const float4 *normals; // nullptr when the scene has no smooth normals
    const uint4 *tri_idx;  // i0, i1, i2, material
    const int32_t *tri_light;
    const Material *mats;
    const Light *lights;
    const float *cdf;
    int32_t n_lights;
    int32_t n_sky;
    int32_t sky[4];
    int32_t width; // BVH node width: 2 (GpuNode, 64 B) or 4 (GpuNode4, 128 B)
    ksd::SkyMap sky_map; // the map of a light of type 4 (environment-map sky), when the scene has one
    // textured scenes (all nullptr otherwise)
    const float2 *uvs;                  // per-vertex texture coordinates
    const ksd::DevTexture *textures;
    const ksd::TexLuts *tex_luts;
    const ksd::DevTexMaterial *tex_mats;
    const int32_t *tex_mat_of;          // per material: index into tex_mats, or -1
    int32_t textured;                   // 1: some geometry has texcoords or some material has fields
    int32_t alpha;                      // 1: some triangle is alpha-tested (opacity map)
    const uint8_t *mat_class;           // per material: kShadeDielectric, kShadeMetal or kShadeGeneral
};

struct RenderConsts {
    int32_t W, H, cx, cy, cw, ch;
    int32_t spp, seed, log2spp, ndigits, bounces, include_background, mis_mode;
    float clamp_direct, clamp_indirect;
    int32_t s0, nb;            // sample range of the batch: [s0, s0 + nb)
    const int4 *work_list;     // optional explicit (x, y, sample) per work item
    // batched renders: work items are segments (base item, view, first sample, pixels of the view's
    // crop), sample-major inside a segment; every view has its own constants
    const struct RenderConsts *views;
    const uint4 *segs;
    int32_t n_segs;
    float p2w[16]; // proj_to_world with the eye folded into rows 0..2 (see camera_ray)
    float cam[3];
    float cam_dx[3], cam_dy[3]; // world offsets of one pixel step on the film plane (Camera::spawn_ray)
    float diff_scale;           // 1 / sqrt(spp) with scale_ray_diff, else 1
    // pixel filter of the jitter: radius, and for a tabulated filter (gaussian, Blackman-Harris) the
    // table size n and its n rows of n + 1 conditional CDF entries followed by the n + 1 marginal ones
    float filter_r;
    int32_t filter_n; // 0: box filter
    const float *filter_cdf;
    const uint8_t *perm_tab; // the sampler's permutation bytes by (dimension, digit prefix), or nullptr
};

struct PathBufs {
    uint32_t *work;
    uint32_t *state; // Sobol dimension (bits 0..15) | bounce (bits 16..23)
    float4 *beta;    // beta.rgb, pdf_wi_bsdf
    float4 *L;       // L.rgb
    float4 *ro;      // origin, tmin
    float4 *rd;      // direction, tmax
};

// Guided shading inputs. Mixtures: 10 float4 per path slot of the current buffer, [w0..w3], [w4..w7],
// then (mu, kappa) per lobe. Training records: kRecVertices per work item, kRecFloat4 float4 each:
// (x, stamp bits), (n, 0), (wo, 0), (wi, 0), (L before the vertex's NEE, 0), (beta before scattering, 0),
// (NEE added by the shadow ray, 0), unused.
constexpr int kRecVertices = 3;
constexpr int kRecFloat4 = 8;
constexpr int kRecNee = 6;
struct GuideShade {
    const float4 *mix; // nullptr: no guide inference this pass (BSDF sampling only)
    float4 *recs;      // nullptr: no training records
    float alpha;       // probability of sampling the guide at a guidable vertex
    float min_roughness; // materials below this roughness use the BSDF only
    uint32_t stamp;
    int32_t max_bounce;  // vertices at this bounce or deeper use the BSDF only (their mixtures are not computed)
};

// ---------------------------------------------------------------------------------------------------
// Wave-level helpers (wave64)
// ---------------------------------------------------------------------------------------------------
KS_DEV uint32_t wave_append(uint32_t *counter, bool pred) {
    uint64_t mask = __ballot(pred);
    uint32_t lane = __lane_id();
    uint32_t prefix = __popcll(mask & ((1ull << lane) - 1ull));
    uint32_t base = 0;
    if (mask != 0) {
        int leader = __ffsll((unsigned long long)mask) - 1;
        if ((int)lane == leader) base = atomicAdd(counter, (uint32_t)__popcll(mask));
        base = __shfl(base, leader);
    }
    return base + prefix;
}

// Block-level append: the block's slots come from one atomic per block rather than one per wave (many
// blocks appending to one counter serialise on its atomic unit). Every thread of the block must call it,
// as it synchronises the block; s_tmp is NWAVES + 1 words of LDS. Slots follow wave order, then lane
// order, within the block.
template <int NWAVES> KS_DEV uint32_t block_append(uint32_t *counter, bool pred, uint32_t *s_tmp) {
    const uint64_t mask = __ballot(pred);
    const uint32_t lane = __lane_id(), wave = threadIdx.x >> 6;
    const uint32_t prefix = __popcll(mask & ((1ull << lane) - 1ull));
    if (lane == 0) s_tmp[wave] = (uint32_t)__popcll(mask);
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t total = 0;
        for (int w = 0; w < NWAVES; ++w) {
            const uint32_t c = s_tmp[w];
            s_tmp[w] = total;
            total += c;
        }
        s_tmp[NWAVES] = total ? atomicAdd(counter, total) : 0u;
    }
    __syncthreads();
    const uint32_t slot = s_tmp[NWAVES] + s_tmp[wave] + prefix;
    __syncthreads(); // s_tmp is free again for the next call
    return slot;
}

// Adds per-lane counts into this block's own row of `rows` (kStatCount slots per row): wave sums by
// shuffles, then one plain read-modify-write per slot by the block's first thread. A block index owns
// its row within a launch and launches on one stream run in order, so no atomic is needed. (With one
// atomic per wave, the thousands of waves of a persistent grid that end together serialise on the few
// counter addresses.) Every thread of the block calls it, with the same slots; s_tmp holds NWAVES * NS.
template <int NWAVES, int NS>
KS_DEV void block_add_stats(unsigned long long *rows, const int (&slot)[NS], const uint32_t (&v)[NS],
                            uint32_t *s_tmp) {
    const uint32_t lane = __lane_id(), wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < NS; ++k) {
        uint32_t x = v[k];
        for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off);
        if (lane == 0) s_tmp[wave * NS + k] = x;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long *row = rows + (size_t)blockIdx.x * kStatCount;
#pragma unroll
        for (int k = 0; k < NS; ++k) {
            unsigned long long sum = 0;
            for (int w = 0; w < NWAVES; ++w) sum += s_tmp[w * NS + k];
            if (sum) row[slot[k]] += sum;
        }
    }
}

KS_DEV void wave_add_stat(unsigned long long *stats, int slot, uint32_t v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    if (__lane_id() == 0 && v) atomicAdd(&stats[slot], (unsigned long long)v);
}

// tri.w: material (bits 0..23) | KS_GEOM_* flags of the triangle's geometry (bits 24..31)
KS_DEV uint32_t tri_material(uint4 tri) { return tri.w & 0xffffffu; }

KS_DEV SurfHit make_hit(const DevScene &sc, uint4 tri, float u, float v, f3 rd) { // geometry.cpp:68-163
    f3 v0 = ld3(sc.verts[tri.x]), v1 = ld3(sc.verts[tri.y]), v2 = ld3(sc.verts[tri.z]);
    f3 wo = -normalized(rd);
    f3 ng = normalized(cross(v1 - v0, v2 - v0));
    float w = 1.0f - u - v;
    SurfHit it;
    it.p = v0 * w + v1 * u + v2 * v;
    f3 dpdu = v1 - v0;
    f3 b = normalized(cross(ng, dpdu));
    f3 t = normalized(cross(b, ng));
    const uint32_t flags = tri.w >> 24;
    if (!(flags & 1u) && dot(wo, ng) < 0.0f) { // two-sided (geometry.cpp:126-132)
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
    if (sc.normals && !(flags & 2u)) {
        f3 vn = normalized(ld3(sc.normals[tri.x]) * w + ld3(sc.normals[tri.y]) * u + ld3(sc.normals[tri.z]) * v);
        if (all_finite(vn)) {
            if (dot(vn, ng) < 0.0f) vn = -vn;
            f3 vb = normalized(cross(vn, dpdu));
            it.st = normalized(cross(vb, vn));
            it.sb = vb;
            it.sn = vn;
        }
    }
    return it;
}

} // namespace ksr

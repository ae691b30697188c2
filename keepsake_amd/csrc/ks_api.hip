// libkeepsake_amd: the gfx950 wavefront small_pt integrator and its C ABI (include/keepsake_amd.h).
//
// Render loop (SmallPT::run, small_pt.cpp:15-70, restructured as a wavefront): path state lives in
// double-buffered SoA arrays in HBM. Each iteration runs
//   raygen  -> new camera paths fill the free tail of the current buffer (Sobol + camera)
//   closest -> closest-hit traversal of every active ray
//   shade   -> emitter/sky MIS, NEE light sample + shadow ray, BSDF sample, Russian roulette;
//              survivors are compacted into the next buffer with a wave ballot + prefix count and
//              one atomic per wave, shadow rays into the shadow queue the same way
//   shadow  -> any-hit traversal; unoccluded NEE contributions are added to the path (or, for a path
//              that ended this bounce, to its per-sample result)
// and a film kernel folds the per-sample results of a batch into the FP64 film in sample order,
// one thread per pixel, so the film is deterministic.
#include <hip/hip_runtime.h>

#include "ks_device.h"
#include "ks_guide.h"
#include "ks_guide_host.h"
#include "ks_internal.h"
#include "ks_pixel_filter.h"
#include "ks_render_types.h"
#include "ks_surface.h"
#include "ks_texture_build.h"
#include "ks_traverse.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

using namespace ksd;
using namespace ksr;

namespace ksa {
static thread_local std::string g_error;
void set_error(const std::string &msg) { g_error = msg; }
} // namespace ksa

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

// ---------------------------------------------------------------------------------------------------
// Device-memory helpers
// ---------------------------------------------------------------------------------------------------
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
        if (n == 0) n = 16;
        hipError_t e = hipMalloc(&p, n);
        if (e == hipSuccess) {
            bytes = n;
        } else {
            p = nullptr;
            (void)hipGetLastError(); // the failure is returned here; later launch checks must not see it again
        }
        return e;
    }
    hipError_t ensure(size_t n) { return n <= bytes ? hipSuccess : alloc(n); }
    template <typename T> T *as() const { return reinterpret_cast<T *>(p); }
};

KS_DEV void pixel_of(const RenderConsts &rc, uint32_t w, uint32_t &x, uint32_t &y, uint32_t &s) {
    if (rc.work_list) {
        int4 e = rc.work_list[w];
        x = (uint32_t)e.x;
        y = (uint32_t)e.y;
        s = (uint32_t)e.z;
        return;
    }
    uint32_t pl = w / (uint32_t)rc.nb;
    s = (uint32_t)rc.s0 + (w - pl * (uint32_t)rc.nb);
    uint32_t ry = pl / (uint32_t)rc.cw;
    x = (uint32_t)rc.cx + (pl - ry * (uint32_t)rc.cw);
    y = (uint32_t)rc.cy + ry;
}

// DistribTable::sample_linear (distrib.cpp:43-53) over cdf[0..n]: the upper_bound cell of u and the
// position (index + du) / n inside [0, 1).
KS_DEV float filter_sample_linear(const float *cdf, int n, float u, uint32_t &index) {
#pragma clang fp contract(off)
    u = fminf(fmaxf(u, 0.0f), 0x1.fffffep-1f);
    uint32_t lo = 0, len = (uint32_t)n + 1; // first entry > u
    while (len > 0) {
        uint32_t half = len >> 1;
        if (cdf[lo + half] <= u) {
            lo += half + 1;
            len -= half + 1;
        } else {
            len = half;
        }
    }
    index = lo - 1;
    const float c0 = cdf[index], c1 = cdf[index + 1];
    const float du = (u - c0) / (c1 - c0);
    return ((float)index + du) / (float)n;
}

// Camera::spawn_ray (camera.cpp:129-170) after the pixel-filter jitter (small_pt.cpp:98-105); q is
// world_pos - eye before normalisation.
KS_DEV void camera_ray_q(const RenderConsts &rc, uint32_t x, uint32_t y, Sobol &sob, f3 &o, f3 &q) {
    // no FMA contraction: the rounding of the projection must not depend on whether the matrix comes
    // from kernel arguments or from memory (plain and batched renders give the same rays)
#pragma clang fp contract(off)
    float ox = 0.0f, oy = 0.0f;
    if (rc.spp != 1) {
        float u0, u1;
        sob.next2d(u0, u1);
        if (rc.filter_n > 0) { // PixelFilterSampleTable::sample (render_target.h:96-103)
            const int n = rc.filter_n;
            uint32_t row;
            const float ty = filter_sample_linear(rc.filter_cdf + (size_t)n * (n + 1), n, u1, row);
            const float tx = filter_sample_linear(rc.filter_cdf + (size_t)row * (n + 1), n, u0, row);
            u0 = tx;
            u1 = ty;
        }
        // BoxPixelFilter::sample (render_target.h:120) and the table's lerp(-r, r, offset)
        ox = (1.0f - u0) * -rc.filter_r + u0 * rc.filter_r;
        oy = (1.0f - u1) * -rc.filter_r + u1 * rc.filter_r;
    }
    float fx = ((float)x + 0.5f + ox) / (float)rc.W;
    float fy = ((float)y + 0.5f + oy) / (float)rc.H;
    float nx = fx * 2.0f - 1.0f, ny = fy * 2.0f - 1.0f;
    // rows 0..2 of rc.p2w hold proj_to_world rows minus eye * row 3 (folded on the host in FP64), so
    // (world_pos - eye) comes out without cancelling two nearly equal float positions
    const float *m = rc.p2w;
    float hx = m[0] * nx + m[1] * ny + m[2] + m[3];
    float hy = m[4] * nx + m[5] * ny + m[6] + m[7];
    float hz = m[8] * nx + m[9] * ny + m[10] + m[11];
    float hw = m[12] * nx + m[13] * ny + m[14] + m[15];
    o = mk(rc.cam[0], rc.cam[1], rc.cam[2]);
    q = mk(hx / hw, hy / hw, hz / hw);
}

KS_DEV void camera_ray(const RenderConsts &rc, uint32_t x, uint32_t y, Sobol &sob, f3 &o, f3 &d) {
    f3 q;
    camera_ray_q(rc, x, y, sob, o, q);
    d = normalized(q);
}


// Per-sample result slot of work item w: sample-major within the batch, so the film kernel's reads of
// one sample across consecutive pixels are contiguous.
KS_DEV uint32_t samp_index(const RenderConsts &rc, uint32_t w) {
    if (rc.work_list) return w;
    uint32_t pl = w / (uint32_t)rc.nb;
    uint32_t sl = w - pl * (uint32_t)rc.nb;
    return sl * (uint32_t)(rc.cw * rc.ch) + pl;
}

KS_DEV void sobol_start(const RenderConsts &rc, Sobol &sob, uint32_t x, uint32_t y, uint32_t s, int dim) {
    sob.start(x, y, s, dim, rc.seed, rc.log2spp, rc.ndigits, rc.perm_tab);
}

// The camera ray's differential directions (camera.cpp:146-167): one pixel step along x and y on the
// film plane, scaled towards the ray by rc.diff_scale. The sample's jitter is drawn again from the
// start of its Sobol stream.
KS_DEV void camera_ray_diffs(const RenderConsts &rc, uint32_t x, uint32_t y, uint32_t s, f3 &o, f3 &rxd, f3 &ryd) {
#pragma clang fp contract(off)
    Sobol sob;
    sobol_start(rc, sob, x, y, s, 0);
    f3 q;
    camera_ray_q(rc, x, y, sob, o, q);
    const f3 d = normalized(q);
    rxd = normalized(q + mk(rc.cam_dx[0], rc.cam_dx[1], rc.cam_dx[2]));
    ryd = normalized(q + mk(rc.cam_dy[0], rc.cam_dy[1], rc.cam_dy[2]));
    rxd = d + (rxd - d) * rc.diff_scale;
    ryd = d + (ryd - d) * rc.diff_scale;
}

// Batched renders: the segment holding work item w (last one whose base is <= w), its view, pixel and
// sample. Within a segment the items are pixel-major (a pixel's samples are consecutive, as in a plain
// render, so a wave traces neighbouring rays); rc.segs[rc.n_segs] is a sentinel whose base is the
// chunk's item count, so every segment's sample count is (next base - base) / pixels. The result slot
// of a batched item is w itself.
KS_DEV uint32_t locate_batch(const RenderConsts &rc, uint32_t w, uint32_t &x, uint32_t &y, uint32_t &s) {
    int lo = 0, hi = rc.n_segs - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (rc.segs[mid].x <= w) lo = mid;
        else hi = mid - 1;
    }
    uint4 sg = rc.segs[lo];
    const uint32_t nb = (rc.segs[lo + 1].x - sg.x) / sg.w;
    uint32_t local = w - sg.x;
    uint32_t p = local / nb;
    uint32_t sl = local - p * nb;
    s = sg.z + sl;
    const RenderConsts &v = rc.views[sg.y];
    uint32_t ry = p / (uint32_t)v.cw;
    x = (uint32_t)v.cx + (p - ry * (uint32_t)v.cw);
    y = (uint32_t)v.cy + ry;
    return sg.y;
}

// ---------------------------------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------------------------------
__global__ void k_fill_u64(unsigned long long *p, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

// Render statistics: row 0 takes the shading kernels' atomics, row 1 the totals, and rows 2 .. 2 + R
// the traversal kernels' per-block counts (block_add_stats). One block of 256 threads sums the R
// partial rows and row 0 into row 1.
__global__ __launch_bounds__(256) void k_stat_reduce(unsigned long long *stats, int R) {
    __shared__ unsigned long long s_sum[256];
    const int k = threadIdx.x % kStatCount, part = threadIdx.x / kStatCount;
    constexpr int kParts = 256 / kStatCount; // threads beyond kParts * kStatCount add nothing
    unsigned long long x = 0;
    if (part < kParts)
        for (int r = part; r < R; r += kParts) x += stats[(size_t)(2 + r) * kStatCount + k];
    s_sum[threadIdx.x] = x;
    __syncthreads();
    if (threadIdx.x < kStatCount) {
        unsigned long long t = stats[threadIdx.x];
        for (int p = 0; p < kParts; ++p) t += s_sum[p * kStatCount + threadIdx.x];
        stats[kStatCount + threadIdx.x] = t;
    }
}

__global__ void k_ctrl_init(Ctrl *c, uint64_t work_end) {
    c->n_active = 0;
    c->n_next.v = 0;
    c->n_shadow.v = 0;
    c->n_shadow_trace = 0;
    c->none = 0;
    c->iteration = 0;
    c->work_next = 0;
    c->work_end = work_end;
}

template <bool BATCH>
__global__ __launch_bounds__(256) void k_raygen(RenderConsts rc, PathBufs cur, const Ctrl *ctrl, uint32_t Q) {
    uint32_t n_active = ctrl->n_active;
    uint64_t wn = ctrl->work_next, we = ctrl->work_end;
    uint64_t room = (uint64_t)(Q - n_active);
    uint32_t n_gen = (uint32_t)min(room, we - wn);
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_gen) return;
    uint32_t w = (uint32_t)(wn + t);
    uint32_t slot = n_active + t;
    uint32_t x, y, s;
    uint32_t view = 0;
    if (BATCH) view = locate_batch(rc, w, x, y, s);
    else pixel_of(rc, w, x, y, s);
    const RenderConsts &vr = BATCH ? rc.views[view] : rc;
    Sobol sob;
    sobol_start(vr, sob, x, y, s, 0);
    f3 o, d;
    camera_ray(vr, x, y, sob, o, d);
    cur.work[slot] = w;
    cur.state[slot] = (uint32_t)sob.dim;
    // beta (1, 1, 1), pdf 0 and L = 0 of a new path are not written: the shading of bounce 0 takes them
    // without a read (path_beta_L)
    cur.ro[slot] = make_float4(o.x, o.y, o.z, 0.0f);
    cur.rd[slot] = make_float4(d.x, d.y, d.z, kInf);
}

// Also hands the shadow rays of the previous shading pass to this iteration's traversal (n_shadow_trace)
// and empties the queue for this iteration's shading pass, which refills it after the traversal.
__global__ void k_gen_done(Ctrl *c, uint32_t Q) {
    uint64_t room = (uint64_t)(Q - c->n_active);
    uint32_t n_gen = (uint32_t)min(room, c->work_end - c->work_next);
    c->n_active += n_gen;
    c->work_next += n_gen;
    c->n_next.v = 0;
    c->n_shadow_trace = c->n_shadow.v;
    c->n_shadow.v = 0;
    for (int k = 0; k < kShadeClasses; ++k) c->n_class[k].v = 0;
}

__global__ void k_iter_done(Ctrl *c) {
    c->n_active = c->n_next.v;
    c->iteration += 1;
}

// Closest hit for n rays (count read from *n_dev when given). Grid-stride over rays; one LDS stack
// column per lane. Ray, node and triangle counts go to the block's row of stat_rows.
// The traversal kernels run at 8 waves per SIMD (64 VGPRs). With ALPHA they call the opacity test, and a
// call within 64 VGPRs spilled the traversal loop's state to scratch on every node visit; they are held
// to 4 waves per SIMD (128 VGPRs) instead, with no spills.
template <int W, bool ALPHA>
__global__ __launch_bounds__(kTraceBlock) __attribute__((amdgpu_waves_per_eu(ALPHA ? 4 : 8))) void k_closest(DevScene sc, const uint32_t *n_dev, uint32_t n_fixed,
                                                         const float4 *ro, const float4 *rd, float4 *hit,
                                                         uint32_t *ovf, uint32_t ovf_cap,
                                                         unsigned long long *stat_rows) {
    __shared__ uint32_t s_stack[kStackLds * kTraceBlock];
    const uint32_t n = n_dev ? *n_dev : n_fixed;
    const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t G = gridDim.x * blockDim.x;
    if (blockIdx.x * blockDim.x >= n) return;
    TravStack st;
    st.lds = s_stack + threadIdx.x;
    st.ovf = ovf + gid;
    st.ovf_stride = G;
    st.ovf_cap = ovf_cap;
    st.spills = 0;
    uint32_t n_nodes = 0, n_tris = 0, n_rays = 0;
    trace_rays<false, W, ALPHA>(
        sc.nodes, sc.tris, gid, n, G, st,
        [&](uint32_t i, TravRay &r) {
            float4 o4 = ro[i], d4 = rd[i];
            r = make_trav_ray(ld3(o4), ld3(d4), o4.w);
            return d4.w;
        },
        [&](uint32_t i, const TravHit &h) {
            hit[i] = h.prim >= 0 ? make_float4(h.t, h.u, h.v, __int_as_float(h.prim))
                                 : make_float4(-1.0f, 0.0f, 0.0f, __int_as_float(-1));
        },
        n_nodes, n_tris, n_rays,
        [&](int32_t prim, float u, float v, const TravRay &r) { return opacity_pass(sc, prim, u, v, r.o, r.d); });
    __shared__ uint32_t s_stat[4 * (kTraceBlock / 64)];
    constexpr int slots[4] = {kRaysClosest, kNodesClosest, kTrisClosest, kSpills};
    const uint32_t v[4] = {n_rays, n_nodes, n_tris, st.spills};
    block_add_stats<kTraceBlock / 64>(stat_rows, slots, v, s_stat);
}

// The NEE contribution of unoccluded shadow ray i, added to its target: a slot of the next path buffer, or
// the per-sample result when bit 31 is set (each target has at most one shadow ray per pass). The
// contribution's w carries the training record of the vertex that cast the ray, + 1 (0: none).
KS_DEV void add_nee(uint32_t i, const uint32_t *target, const float4 *contrib, float4 *path_L, float4 *samp,
                    float4 *recs) {
    uint32_t tg = target[i];
    float4 c = contrib[i];
    float4 *dst = (tg & 0x80000000u) ? &samp[tg & 0x7fffffffu] : &path_L[tg];
    float4 v = *dst;
    v.x += c.x;
    v.y += c.y;
    v.z += c.z;
    *dst = v;
    uint32_t slot = __float_as_uint(c.w);
    if (recs && slot) recs[(size_t)(slot - 1) * kRecFloat4 + kRecNee] = make_float4(c.x, c.y, c.z, 0.0f);
}

// Any hit: the occlusion flag of each of n rays (ks_occluded).
template <int W, bool ALPHA>
__global__ __launch_bounds__(kTraceBlock) __attribute__((amdgpu_waves_per_eu(ALPHA ? 4 : 8))) void k_anyhit(DevScene sc, const uint32_t *n_dev, uint32_t n_fixed,
                                                        const float4 *ro, const float4 *rd, uint8_t *occ,
                                                        uint32_t *ovf, uint32_t ovf_cap,
                                                        unsigned long long *stat_rows) {
    __shared__ uint32_t s_stack[kStackLds * kTraceBlock];
    const uint32_t n = n_dev ? *n_dev : n_fixed;
    const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t G = gridDim.x * blockDim.x;
    if (blockIdx.x * blockDim.x >= n) return;
    TravStack st;
    st.lds = s_stack + threadIdx.x;
    st.ovf = ovf + gid;
    st.ovf_stride = G;
    st.ovf_cap = ovf_cap;
    st.spills = 0;
    uint32_t n_nodes = 0, n_tris = 0, n_rays = 0;
    auto load = [&](uint32_t i, TravRay &r) {
        float4 o4 = ro[i], d4 = rd[i];
        r = make_trav_ray(ld3(o4), ld3(d4), o4.w);
        return d4.w;
    };
    auto store = [&](uint32_t i, const TravHit &h) { occ[i] = h.prim >= 0 ? 1 : 0; };
    trace_rays<true, W, ALPHA>(
        sc.nodes, sc.tris, gid, n, G, st, load, store, n_nodes, n_tris, n_rays,
        [&](int32_t prim, float u, float v, const TravRay &r) { return opacity_pass(sc, prim, u, v, r.o, r.d); });
    __shared__ uint32_t s_stat[4 * (kTraceBlock / 64)];
    constexpr int slots[4] = {kRaysShadow, kNodesShadow, kTrisShadow, kSpills};
    const uint32_t v[4] = {n_rays, n_nodes, n_tris, st.spills};
    block_add_stats<kTraceBlock / 64>(stat_rows, slots, v, s_stat);
}

// One traversal launch per wavefront iteration: the closest hits of the iteration's nc rays (ro, rd -> hit)
// and the occlusion of the ns shadow rays of the previous shading pass (sro, srd; an unoccluded ray adds
// its contribution by add_nee). Indices [0, nc) are closest-hit queries and [nc, nc + ns) shadow rays,
// one grid-stride sequence: the persistent grid balances both queues, and the long shadow rays of one
// pass run beside the next pass's closest-hit rays instead of in a launch of their own, whose length
// the slowest rays set. Node and triangle counts are kept per ray and added to the block's LDS totals of
// the ray's kind when it ends (no per-kind registers in the loop); block 0 adds the ray counts.
template <int W, bool ALPHA>
__global__ __launch_bounds__(kTraceBlock) __attribute__((amdgpu_waves_per_eu(ALPHA ? 4 : 8))) void k_trace(
    DevScene sc, const uint32_t *nc_dev, const float4 *ro, const float4 *rd, float4 *hit, const uint32_t *ns_dev,
    const float4 *sro, const float4 *srd, const uint32_t *target, const float4 *contrib, float4 *path_L, float4 *samp,
    float4 *recs, uint32_t *ovf, uint32_t ovf_cap, unsigned long long *stat_rows) {
    __shared__ uint32_t s_stack[kStackLds * kTraceBlock];
    __shared__ uint32_t s_cnt[5]; // closest nodes, tris; shadow nodes, tris; stack spills
    const uint32_t nc = *nc_dev, ns = *ns_dev, n = nc + ns;
    const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t G = gridDim.x * blockDim.x;
    if (blockIdx.x * blockDim.x >= n) return;
    if (threadIdx.x < 5) s_cnt[threadIdx.x] = 0u;
    __syncthreads();
    TravStack st;
    st.lds = s_stack + threadIdx.x;
    st.ovf = ovf + gid;
    st.ovf_stride = G;
    st.ovf_cap = ovf_cap;
    st.spills = 0;
    uint32_t n_nodes = 0, n_tris = 0, n_rays = 0; // nodes and tris of the current ray
    trace_rays_kind<W, ALPHA>(
        sc.nodes, sc.tris, gid, n, G, st, [nc](uint32_t i) { return i >= nc; },
        [&](uint32_t i, TravRay &r) {
            const float4 *po = ro, *pd = rd;
            uint32_t j = i;
            if (i >= nc) {
                po = sro;
                pd = srd;
                j = i - nc;
            }
            const float4 o4 = po[j], d4 = pd[j];
            r = make_trav_ray(ld3(o4), ld3(d4), o4.w);
            return d4.w;
        },
        [&](uint32_t i, const TravHit &h) {
            const bool sh = i >= nc;
            if (!sh)
                hit[i] = h.prim >= 0 ? make_float4(h.t, h.u, h.v, __int_as_float(h.prim))
                                     : make_float4(-1.0f, 0.0f, 0.0f, __int_as_float(-1));
            else if (h.prim < 0)
                add_nee(i - nc, target, contrib, path_L, samp, recs);
            atomicAdd(&s_cnt[sh ? 2 : 0], n_nodes);
            atomicAdd(&s_cnt[sh ? 3 : 1], n_tris);
            n_nodes = 0;
            n_tris = 0;
        },
        n_nodes, n_tris, n_rays,
        [&](int32_t prim, float u, float v, const TravRay &r) { return opacity_pass(sc, prim, u, v, r.o, r.d); });
    if (st.spills) atomicAdd(&s_cnt[4], st.spills);
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long *row = stat_rows + (size_t)blockIdx.x * kStatCount;
        row[kNodesClosest] += s_cnt[0];
        row[kTrisClosest] += s_cnt[1];
        row[kNodesShadow] += s_cnt[2];
        row[kTrisShadow] += s_cnt[3];
        row[kSpills] += s_cnt[4];
        if (blockIdx.x == 0) {
            row[kRaysClosest] += nc;
            row[kRaysShadow] += ns;
        }
    }
}

KS_DEV uint32_t cdf_sample(const float *cdf, int n, float u, float &prob) { // distrib.cpp:33-41
    u = clampf(u, 0.0f, 0x1.fffffep-1f);
    // upper_bound over cdf[0..n]: first index with cdf[i] > u
    int lo = 0, hi = n + 1;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (cdf[mid] > u) hi = mid;
        else lo = mid + 1;
    }
    uint32_t idx = (uint32_t)(lo - 1);
    prob = cdf[idx + 1] - cdf[idx];
    return idx;
}
KS_DEV float light_probability(const float *cdf, int n, uint32_t i) { // light.cpp:680-683
    return ((cdf[i + 1] - cdf[i]) * (float)n) / (float)n;
}
KS_DEV float sky_pdf(f3 wi) { // SkyLight::pdf (light.cpp:113-134) for a 1x1 map, y-up
    float theta = acosf(clampf(wi.y, -1.0f, 1.0f));
    float sin_theta = sinf(theta);
    if (sin_theta == 0.0f) return 0.0f;
    return 1.0f / (2.0f * kPi * kPi * sin_theta);
}

// Light::sample for the flat light list; returns Le / pdf (light.h:27-29). TEX reads textured emission.
template <bool TEX = false>
KS_DEV f3 light_sample(const DevScene &sc, const Light &l, f3 p, float u0, float u1, f3 &wi, float &wi_dist,
                       float &pdf) {
    const f3 zero = mk(0, 0, 0);
    f3 col = mk(l.color[0], l.color[1], l.color[2]);
    if (l.type == 0) { // SkyLight::sample (light.cpp:88-111), 1x1 map
        float a = clampf(u0, 0.0f, 0x1.fffffep-1f), b = clampf(u1, 0.0f, 0x1.fffffep-1f);
        float phi = a * 2.0f * kPi, theta = b * kPi;
        float st, ct, sp, cp;
        sincosf(theta, &st, &ct);
        sincosf(phi, &sp, &cp);
        pdf = 1.0f / (2.0f * kPi * kPi * st);
        if (st == 0.0f) {
            pdf = 0.0f;
            return zero;
        }
        wi = mk(cp * st, ct, sp * st);
        wi_dist = kInf;
        return col / pdf;
    }
    if (l.type == 4) { // SkyLight::sample (light.cpp:88-111), environment map
        float e[3];
        sky_map_sample(sc.sky_map, u0, u1, wi.x, wi.y, wi.z, pdf, e);
        wi_dist = kInf;
        return mk(e[0], e[1], e[2]);
    }
    if (l.type == 1) { // DirectionalLight::sample (light.cpp:207-213)
        wi = mk(l.vec[0], l.vec[1], l.vec[2]);
        wi_dist = kInf;
        pdf = 1.0f;
        return col;
    }
    if (l.type == 2) { // PointLight::sample (light.cpp:222-233)
        wi = mk(l.vec[0], l.vec[1], l.vec[2]) - p;
        float l2 = sqnorm(wi);
        pdf = 0.0f;
        if (l2 == 0.0f) return zero;
        float len = sqrtf(l2);
        wi = wi / len;
        wi_dist = len;
        pdf = 1.0f;
        return col / l2;
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
    f3 v0 = mk(l.vec[0], l.vec[1], l.vec[2]), v1 = mk(l.v1[0], l.v1[1], l.v1[2]), v2 = mk(l.v2[0], l.v2[1], l.v2[2]);
    f3 pos = v0 * (1.0f - b0 - b1) + v1 * b0 + v2 * b1;
    f3 n = normalized(cross(v1 - v0, v2 - v1));
    wi = pos - p;
    float d2 = sqnorm(wi);
    pdf = 0.0f;
    wi_dist = 0.0f;
    if (d2 == 0.0f) return zero;
    float dist = sqrtf(d2);
    wi = wi / dist;
    float absdot = fabsf(dot(n, wi));
    if (absdot == 0.0f) return zero;
    pdf = (1.0f / l.area) * (d2 / absdot);
    wi_dist = fmaxf(0.0f, dist - 1e-4f);
    if (TEX && l.color[3] != 0.0f) { // textured emission at the sampled point (MeshTriLight::eval, light.cpp:560-565)
        const uint4 tri = sc.tri_idx[l.prim];
        const DevTexMaterial &tm = sc.tex_mats[sc.tex_mat_of[tri_material(tri)]];
        float tu = 0.0f, tv = 0.0f; // interpolate_texcoord: zero without texcoords
        if (geom_has_uv(sc, tri)) interp_uv(sc, tri, b0, b1, tu, tv);
        float e[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        field_eval0(tex_context(sc), tm.f[kFieldEmission], tu, tv, 3, e);
        col = mk(e[0], e[1], e[2]);
    }
    return col / pdf;
}

KS_DEV float mesh_light_pdf(const DevScene &sc, const Light &l, f3 wi, float wi_dist) { // light.cpp:607-628
    if (wi_dist == 0.0f) return 0.0f;
    f3 v0 = mk(l.vec[0], l.vec[1], l.vec[2]), v1 = mk(l.v1[0], l.v1[1], l.v1[2]), v2 = mk(l.v2[0], l.v2[1], l.v2[2]);
    f3 n = normalized(cross(v1 - v0, v2 - v1));
    float absdot = fabsf(dot(n, wi));
    if (absdot == 0.0f) return 0.0f;
    return (1.0f / l.area) * (sqr(wi_dist) / absdot);
}

// Material sort of the current buffer: each path's slot goes to the list of its class (a ray that left
// the scene, or the class of the material it hit), Q entries per class, in the order of the buffer
// within each wave.
__global__ __launch_bounds__(256) void k_classify(DevScene sc, const float4 *hits, Ctrl *ctrl, uint32_t *order,
                                                  uint32_t Q) {
    __shared__ uint32_t s_app[5];
    const uint32_t n = ctrl->n_active;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (blockIdx.x * blockDim.x >= n) return;
    int cls = -1;
    if (i < n) {
        const int32_t prim = __float_as_int(hits[i].w);
        cls = prim < 0 ? kShadeMiss : (int)sc.mat_class[tri_material(sc.tri_idx[prim])];
    }
#pragma unroll
    for (int k = 0; k < kShadeClasses; ++k) {
        const uint32_t j = block_append<4>(&ctrl->n_class[k].v, cls == k, s_app);
        if (cls == k) order[(size_t)k * Q + j] = i;
    }
}

// One bounce of SmallPT::trace (small_pt.cpp:72-178) for every active path.
// Mixture pdf at w from the 10-float4 record of one path (sum of w_k vmf_k(w), vmf.h:32-39).
KS_DEV float mix_pdf(const float4 *m, f3 w) {
    float4 wa = m[0], wb = m[1];
    float wk[8] = {wa.x, wa.y, wa.z, wa.w, wb.x, wb.y, wb.z, wb.w};
    float p = 0.0f;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        float4 l = m[2 + k];
        if (wk[k] > 0.0f)
            p += wk[k] * __expf(vmf_log_norm(l.w) + l.w * fminf(0.0f, dot(ld3(l), w) - 1.0f));
    }
    return p;
}

// Throughput, pdf of the last BSDF sample and radiance of path slot i: (1, 1, 1), 0 and 0 for a path that
// has not been shaded yet (bounce 0, whose values ray generation does not write), else the stored ones.
KS_DEV void path_beta_L(const PathBufs &cur, uint32_t i, uint32_t state, f3 &beta, float &pdf, f3 &L) {
    if (((state >> 16) & 0xffu) == 0u) {
        beta = mk(1.0f, 1.0f, 1.0f);
        pdf = 0.0f;
        L = mk(0.0f, 0.0f, 0.0f);
        return;
    }
    const float4 b4 = cur.beta[i];
    beta = ld3(b4);
    pdf = b4.w;
    L = ld3(cur.L[i]);
}

// One vertex of SmallPT::trace (small_pt.cpp:79-176) for the path whose state the caller loaded: work
// item w, state (Sobol dimension, bounce), throughput beta, pdf of the last BSDF sample, radiance L, the
// ray direction rd that reached the vertex and its hit h (t, u, v, prim; prim < 0: escaped). Adds
// emission to L, samples NEE (the shadow ray and its contribution, shadow = true) and the continuation
// (new_o, new_d, updated beta and pdf; survive = true), or ends the path, writing its sample result.
// i is the path's slot, used for the guide's mixture and training record.
template <bool GUIDED, bool BATCH, bool TEX, int CLASS>
KS_DEV void shade_vertex(const DevScene &sc, const RenderConsts &rc, const GuideShade &gs, float4 *samp, uint32_t i,
                         float4 h, f3 rd, uint32_t w, uint32_t state, f3 &L, f3 &beta, float &pdf_wi_bsdf,
                         bool &survive, bool &shadow, f3 &new_o, f3 &new_d, f3 &sh_origin, f3 &sh_dir,
                         f3 &sh_contrib, float &sh_tmax, uint32_t &nonfinite, int &bounce, Sobol &sob,
                         uint32_t &rec_slot, uint32_t &view) {
    constexpr unsigned LB = CLASS == kShadeDielectric ? kLobesOpaqueDielectric
                            : CLASS == kShadeMetal     ? kLobesMetal
                                                       : kLobesAll;
    int32_t prim = __float_as_int(h.w);
    bounce = (int)(state >> 16) & 0xff;
    uint32_t x, y, s;
    if (BATCH) view = locate_batch(rc, w, x, y, s);
    else pixel_of(rc, w, x, y, s);
    const RenderConsts &vr = BATCH ? rc.views[view] : rc;
    sobol_start(vr, sob, x, y, s, (int)(state & 0xffffu));
    bool done = false;
    bool miss_sample = false;

    if (prim < 0) {
      if constexpr (CLASS == kShadeUnsorted || CLASS == kShadeMiss) {
        f3 wi = normalized(rd);
        for (int k = 0; k < sc.n_sky; ++k) { // small_pt.cpp:85-109
            const Light &l = sc.lights[sc.sky[k]];
            const bool map = l.type == 4;
            f3 Le = mk(l.color[0], l.color[1], l.color[2]);
            if (map) {
                float e[3];
                sky_eval(sc.sky_map, wi.x, wi.y, wi.z, e);
                Le = mk(e[0], e[1], e[2]);
            }
            if (bounce == 0) {
                if (vr.include_background) L = L + beta * Le;
            } else {
                float pr = light_probability(sc.cdf, sc.n_lights, (uint32_t)sc.sky[k]);
                float lpdf = map ? sky_map_pdf(sc.sky_map, wi.x, wi.y, wi.z) : sky_pdf(wi);
                float mis = power_heur(pdf_wi_bsdf, lpdf * pr);
                f3 bu = Le * mis;
                if (!all_finite(bu)) ++nonfinite;
                if (bounce == 1 && vr.clamp_direct > 0.0f) bu = vmin(bu, vr.clamp_direct);
                else if (bounce >= 2 && vr.clamp_indirect > 0.0f) bu = vmin(bu, vr.clamp_indirect);
                L = L + beta * bu;
            }
        }
        if (bounce == 0 && !vr.include_background) miss_sample = true;
      }
        done = true;
    } else {
      if constexpr (CLASS != kShadeMiss) {
        uint4 tri = sc.tri_idx[prim];
        const uint32_t mid = tri_material(tri);
        const Material &m = sc.mats[mid];
        SurfHit it;
        SurfUV su;
        const DevTexMaterial *tm = nullptr;
        if (TEX) {
            // compute_intersection with uv partials, then the normal map (Scene::intersect1, scene.cpp:222-233)
            it = make_hit_uv(sc, tri, h.y, h.z, rd, su);
            const int32_t tmi = sc.tex_mat_of ? sc.tex_mat_of[mid] : -1;
            if (tmi >= 0) {
                tm = &sc.tex_mats[tmi];
                if (bounce == 0) {
                    f3 co, rxd, ryd;
                    camera_ray_diffs(vr, x, y, s, co, rxd, ryd);
                    uv_partials(it, su, co, rxd, ryd);
                }
                if (tm->has_normal) apply_normal_map(tex_context(sc), *tm, su.u, su.v, it);
            }
        } else {
            it = make_hit(sc, tri, h.y, h.z, rd);
        }
        if (m.has_emission) { // small_pt.cpp:111-138
            int32_t li = sc.tri_light[prim];
            if (li >= 0) {
                const Light &l = sc.lights[li];
                float pr = light_probability(sc.cdf, sc.n_lights, (uint32_t)li);
                f3 Le = mk(l.color[0], l.color[1], l.color[2]);
                if (TEX && tm && l.color[3] != 0.0f) { // MeshTriLight::eval at the hit (light.cpp:560-565)
                    float e[4] = {0.0f, 0.0f, 0.0f, 0.0f};
                    field_eval(tex_context(sc), tm->f[kFieldEmission], su.u, su.v, su.d, 3, e);
                    Le = mk(e[0], e[1], e[2]);
                }
                if (bounce == 0) {
                    L = L + beta * Le;
                } else if (pr > 0.0f) {
                    float pdf_l = mesh_light_pdf(sc, l, normalized(rd), h.x);
                    float mis = power_heur(pdf_wi_bsdf, pdf_l * pr);
                    f3 bu = Le * mis;
                    if (!all_finite(bu)) ++nonfinite;
                    if (bounce == 1 && vr.clamp_direct > 0.0f) bu = vmin(bu, vr.clamp_direct);
                    else if (bounce >= 2 && vr.clamp_indirect > 0.0f) bu = vmin(bu, vr.clamp_indirect);
                    L = L + beta * bu;
                }
            }
        }
        if (bounce == vr.bounces) {
            done = true;
        } else {
            f3 wo = -normalized(rd);
            Closure c;
            if (TEX && tm) {
                c = closure_from_fields(tex_context(sc), *tm, su.u, su.v, su.d);
            } else {
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
            }
            const float4 *mixp = GUIDED && gs.mix ? gs.mix + 10 * (size_t)i : nullptr;
            const float alpha_m =
                GUIDED && mixp && c.roughness >= gs.min_roughness && bounce < gs.max_bounce ? gs.alpha : 0.0f;

            // NEE (nee.cpp:11-58)
            if (sc.n_lights > 0) {
                float pr_light;
                uint32_t li = cdf_sample(sc.cdf, sc.n_lights, sob.next(), pr_light);
                const Light &l = sc.lights[li];
                bool delta_light = l.type == 1 || l.type == 2;
                float u0, u1;
                sob.next2d(u0, u1);
                f3 wi;
                float wi_dist = 0.0f, pdf_l = 0.0f;
                f3 Le_beta = light_sample<TEX>(sc, l, it.p, u0, u1, wi, wi_dist, pdf_l);
                if (!is_zero(Le_beta) && pdf_l > 0.0f) {
                    f3 wol = sh_to_local(it, wo), wil = sh_to_local(it, wi);
                    f3 f = bsdf_eval<LB>(wol, wil, c);
                    float pdf_b = is_zero(f) ? 0.0f : bsdf_pdf<LB>(wol, wil, c);
                    if (!is_zero(f) && pdf_b > 0.0f) {
                        if (GUIDED && alpha_m > 0.0f) pdf_b = alpha_m * mix_pdf(mixp, wi) + (1.0f - alpha_m) * pdf_b;
                        float mis = 1.0f;
                        if (!delta_light)
                            mis = vr.mis_mode == KS_MIS_CONSISTENT ? power_heur(pdf_l * pr_light, pdf_b)
                                                                   : power_heur(pdf_l, pdf_b);
                        f3 Ld = Le_beta * f * mis / pr_light;
                        if (!all_finite(Ld)) ++nonfinite;
                        if (bounce == 0 && vr.clamp_direct > 0.0f) Ld = vmin(Ld, vr.clamp_direct);
                        else if (bounce >= 1 && vr.clamp_indirect > 0.0f) Ld = vmin(Ld, vr.clamp_indirect);
                        shadow = true;
                        sh_origin = spawn_origin(it.p, wi, it.n);
                        sh_dir = wi;
                        sh_tmax = wi_dist;
                        sh_contrib = beta * Ld;
                    }
                }
            } else {
                (void)sob.next();
                float a, b;
                sob.next2d(a, b);
            }
            // BSDF sampling (material.cpp:157-166)
            f3 wol = sh_to_local(it, wo);
            float u_lobe = sob.next();
            float u0, u1;
            sob.next2d(u0, u1);
            f3 wil;
            float pdf_wi = 0.0f;
            f3 sbeta, wi_g = mk(0, 0, 0);
            const f3 beta_k = beta;
            if (GUIDED && alpha_m > 0.0f) {
                // one-sample MIS: u_lobe picks the strategy and is re-used inside it
                float pb = 0.0f;
                f3 fcos;
                if (u_lobe < alpha_m) {
                    float4 wa = mixp[0], wb = mixp[1];
                    float wk[8] = {wa.x, wa.y, wa.z, wa.w, wb.x, wb.y, wb.z, wb.w};
                    float ur;
                    int sel = mixture_select(wk, u0, &ur);
                    float4 l4 = mixp[2 + sel];
                    Lobe lobe;
                    lobe.mu = ld3(l4);
                    lobe.kappa = l4.w;
                    wi_g = vmf_sample(lobe, ur, u1);
                    wil = sh_to_local(it, wi_g);
                    fcos = bsdf_eval<LB>(wol, wil, c);
                    pb = is_zero(fcos) ? 0.0f : bsdf_pdf<LB>(wol, wil, c);
                } else {
                    float ub = fminf((u_lobe - alpha_m) / (1.0f - alpha_m), kBeforeOne);
                    f3 sb = bsdf_sample<LB>(wol, wil, c, ub, u0, u1, pb);
                    fcos = sb * pb;
                    wi_g = sh_to_world(it, wil);
                }
                float q = mix_pdf(mixp, wi_g);
                pdf_wi = alpha_m * q + (1.0f - alpha_m) * pb;
                sbeta = pdf_wi > 0.0f ? fcos / pdf_wi : mk(0, 0, 0);
            } else {
                sbeta = bsdf_sample<LB>(wol, wil, c, u_lobe, u0, u1, pdf_wi);
            }
            if (!all_finite(sbeta)) ++nonfinite;
            if (maxc(sbeta) == 0.0f || pdf_wi == 0.0f) {
                done = true;
            } else {
                f3 wi = GUIDED && alpha_m > 0.0f ? wi_g : sh_to_world(it, wil);
                beta = beta * sbeta;
                if (maxc(beta) == 0.0f) {
                    done = true;
                } else {
                    pdf_wi_bsdf = pdf_wi;
                    new_o = spawn_origin(it.p, wi, it.n);
                    new_d = wi;
                    if (maxc(beta) < 1.0f && bounce >= 1) { // Russian roulette (small_pt.cpp:167-174)
                        float q = fmaxf(0.0f, 1.0f - maxc(beta));
                        if (sob.next() < q) done = true;
                        else beta = beta / (1.0f - q);
                    }
                    if (GUIDED && gs.recs && !done && bounce < kRecVertices) {
                        uint32_t slot = w * (uint32_t)kRecVertices + (uint32_t)bounce;
                        float4 *r = gs.recs + (size_t)slot * kRecFloat4;
                        r[0] = make_float4(it.p.x, it.p.y, it.p.z, __uint_as_float(gs.stamp));
                        r[1] = make_float4(it.n.x, it.n.y, it.n.z, 0.0f);
                        r[2] = make_float4(wo.x, wo.y, wo.z, 0.0f);
                        r[3] = make_float4(wi.x, wi.y, wi.z, 0.0f);
                        r[4] = make_float4(L.x, L.y, L.z, 0.0f);
                        r[5] = make_float4(beta_k.x, beta_k.y, beta_k.z, 0.0f);
                        r[kRecNee] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                        rec_slot = slot + 1;
                    }
                }
            }
        }
      }
    }
    survive = !done;
    if (done) samp[BATCH ? w : samp_index(rc, w)] = make_float4(L.x, L.y, L.z, miss_sample ? 0.0f : 1.0f);
}

// Shading of the current buffer (small_pt.cpp:79-176). GUIDED adds one-sample MIS between the BSDF and
// the learned mixture (probability gs.alpha of the mixture at vertices whose roughness reaches
// gs.min_roughness; the continuation pdf is alpha q + (1 - alpha) p_bsdf in the estimator and in both
// MIS weights), and training records of the first kRecVertices vertices of each path.
// TEX evaluates hits of textured scenes: texture coordinates and uv-based frames, uv partials of camera
// rays, normal maps, closures and emission from shader fields.
// CLASS (material-sorted shading): the kernel shades the slots of one class's list in `order` (Q entries
// per class) and is compiled for that class only: kShadeMiss holds only the escaped-ray code, the
// dielectric and metal classes the BSDF lobes their materials can weight. kShadeUnsorted shades the
// whole buffer with every lobe.
template <bool GUIDED, bool BATCH, bool TEX, int CLASS = kShadeUnsorted>
__global__ __launch_bounds__(256) void k_shade(DevScene sc, RenderConsts rc, PathBufs cur, PathBufs nxt,
                                               const float4 *hits, Ctrl *ctrl, float4 *sh_o, float4 *sh_d,
                                               uint32_t *sh_tgt, float4 *sh_c, float4 *samp,
                                               unsigned long long *stats, GuideShade gs, const uint32_t *order,
                                               uint32_t Q) {
    const uint32_t n = CLASS == kShadeUnsorted ? ctrl->n_active : ctrl->n_class[CLASS].v;
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (blockIdx.x * blockDim.x >= n) return; // whole block idle
    const bool active = t < n;
    const uint32_t i = CLASS == kShadeUnsorted ? t : (active ? order[(size_t)CLASS * Q + t] : 0u);
    (void)Q;
    (void)order;

    bool survive = false, shadow = false;
    uint32_t w = 0, state = 0;
    f3 L = mk(0, 0, 0), beta = mk(1, 1, 1);
    float pdf_wi_bsdf = 0.0f;
    f3 new_o = mk(0, 0, 0), new_d = mk(0, 0, 0);
    f3 sh_origin = mk(0, 0, 0), sh_dir = mk(0, 0, 0), sh_contrib = mk(0, 0, 0);
    float sh_tmax = 0.0f;
    uint32_t nonfinite = 0;
    int bounce = 0;
    Sobol sob;
    uint32_t rec_slot = 0; // record slot + 1 of this vertex (0: none)
    uint32_t view = 0;

    if (active) {
        w = cur.work[i];
        state = cur.state[i];
        path_beta_L(cur, i, state, beta, pdf_wi_bsdf, L);
        shade_vertex<GUIDED, BATCH, TEX, CLASS>(sc, rc, gs, samp, i, hits[i], ld3(cur.rd[i]), w, state, L, beta,
                                                pdf_wi_bsdf, survive, shadow, new_o, new_d, sh_origin, sh_dir,
                                                sh_contrib, sh_tmax, nonfinite, bounce, sob, rec_slot, view);
    }

    __shared__ uint32_t s_app[5];
    uint32_t j = block_append<4>(&ctrl->n_next.v, survive, s_app);
    if (survive) {
        nxt.work[j] = w;
        nxt.state[j] = (uint32_t)sob.dim | ((uint32_t)(bounce + 1) << 16);
        nxt.beta[j] = make_float4(beta.x, beta.y, beta.z, pdf_wi_bsdf);
        nxt.L[j] = make_float4(L.x, L.y, L.z, 0.0f);
        nxt.ro[j] = make_float4(new_o.x, new_o.y, new_o.z, 0.0f);
        nxt.rd[j] = make_float4(new_d.x, new_d.y, new_d.z, kInf);
    }
    uint32_t k = block_append<4>(&ctrl->n_shadow.v, shadow, s_app);
    if (shadow) {
        sh_o[k] = make_float4(sh_origin.x, sh_origin.y, sh_origin.z, 0.0f);
        sh_d[k] = make_float4(sh_dir.x, sh_dir.y, sh_dir.z, sh_tmax);
        sh_tgt[k] = survive ? j : (0x80000000u | (BATCH ? w : samp_index(rc, w)));
        sh_c[k] = make_float4(sh_contrib.x, sh_contrib.y, sh_contrib.z, __uint_as_float(rec_slot));
    }
    wave_add_stat(stats, kNonFinite, nonfinite);
}

// The last paths of a render, each carried to its end by one thread: closest hit, shading, the NEE shadow
// ray and the next bounce, in a loop. Once every work item has started and few paths remain, a wavefront
// iteration costs about the time of its slowest ray, a dozen times over; here a thread waits only for its
// own rays. The results are the wavefront's: the same traversal hits (traverse_w, the per-ray form of
// trace_rays_kind's loop), the same shading (shade_vertex), and the NEE contribution added to the path's
// radiance, or to its sample result once the path has ended, before the next vertex, with the same
// single addition the wavefront's traversal makes (add_nee). Plain renders of scenes without alpha-tested
// triangles only; n paths of buffer cur, n no larger than the overflow area's lanes.
template <int W, bool BATCH, bool TEX>
__global__ __launch_bounds__(256) void k_tail(DevScene sc, RenderConsts rc, PathBufs cur, const uint32_t *n_dev,
                                              float4 *samp, uint32_t *ovf, uint32_t ovf_cap,
                                              unsigned long long *stat_rows) {
    __shared__ uint32_t s_stack[kStackLds * 256];
    __shared__ uint32_t s_stat[3 * 4];
    const uint32_t n = *n_dev;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (blockIdx.x * blockDim.x >= n) return; // whole block idle
    uint32_t n_rays = 0, nonfinite = 0, spills = 0;
    if (i < n) {
        TravStack st;
        st.lds = s_stack + threadIdx.x;
        st.ovf = ovf + i;
        st.ovf_stride = gridDim.x * blockDim.x;
        st.ovf_cap = ovf_cap;
        st.spills = 0;
        uint32_t n_nodes = 0, n_tris = 0; // the tail's rays are counted, not their nodes
        const uint32_t w = cur.work[i];
        uint32_t state = cur.state[i];
        f3 beta, L;
        float pdf;
        path_beta_L(cur, i, state, beta, pdf, L);
        const float4 o4 = cur.ro[i], d4 = cur.rd[i];
        f3 o = ld3(o4), d = ld3(d4);
        float tmin = o4.w, tmax = d4.w;
        while (true) {
            TravHit th;
            traverse_w<false, W>(sc.nodes, sc.tris, make_trav_ray(o, d, tmin), tmax, st, th, n_nodes, n_tris);
            ++n_rays;
            const float4 h = th.prim >= 0 ? make_float4(th.t, th.u, th.v, __int_as_float(th.prim))
                                          : make_float4(-1.0f, 0.0f, 0.0f, __int_as_float(-1));
            bool survive = false, shadow = false;
            f3 new_o = mk(0, 0, 0), new_d = mk(0, 0, 0);
            f3 sh_origin = mk(0, 0, 0), sh_dir = mk(0, 0, 0), sh_contrib = mk(0, 0, 0);
            float sh_tmax = 0.0f;
            int bounce = 0;
            Sobol sob;
            uint32_t rec_slot = 0, view = 0;
            shade_vertex<false, BATCH, TEX, kShadeUnsorted>(sc, rc, GuideShade{}, samp, i, h, d, w, state, L, beta,
                                                            pdf, survive, shadow, new_o, new_d, sh_origin, sh_dir,
                                                            sh_contrib, sh_tmax, nonfinite, bounce, sob, rec_slot,
                                                            view);
            if (shadow) {
                TravHit sh;
                const bool occluded = traverse_w<true, W>(sc.nodes, sc.tris, make_trav_ray(sh_origin, sh_dir, 0.0f),
                                                          sh_tmax, st, sh, n_nodes, n_tris);
                ++n_rays;
                if (!occluded) {
                    if (survive) {
                        L.x += sh_contrib.x;
                        L.y += sh_contrib.y;
                        L.z += sh_contrib.z;
                    } else {
                        float4 *dst = &samp[BATCH ? w : samp_index(rc, w)];
                        float4 v = *dst;
                        v.x += sh_contrib.x;
                        v.y += sh_contrib.y;
                        v.z += sh_contrib.z;
                        *dst = v;
                    }
                }
            }
            if (!survive) break;
            state = (uint32_t)sob.dim | ((uint32_t)(bounce + 1) << 16);
            o = new_o;
            d = new_d;
            tmin = 0.0f;
            tmax = kInf;
        }
        spills = st.spills;
    }
    constexpr int slots[3] = {kRaysTail, kNonFinite, kSpills};
    const uint32_t v[3] = {n_rays, nonfinite, spills};
    block_add_stats<4>(stat_rows, slots, v, s_stat);
}

// RenderTarget2::add / add_miss (render_target.cpp:82-107) for every sample of the batch, in sample
// order, one thread per pixel of the crop.
__global__ __launch_bounds__(256) void k_film(RenderConsts rc, const float4 *samp, double *rgb, double *wts) {
    uint32_t pl = blockIdx.x * blockDim.x + threadIdx.x;
    if (pl >= (uint32_t)(rc.cw * rc.ch)) return;
    uint32_t ry = pl / (uint32_t)rc.cw;
    uint32_t x = (uint32_t)rc.cx + (pl - ry * (uint32_t)rc.cw), y = (uint32_t)rc.cy + ry;
    size_t p = (size_t)y * rc.W + x;
    double r = rgb[3 * p], g = rgb[3 * p + 1], b = rgb[3 * p + 2];
    double wh = wts[2 * p], wm = wts[2 * p + 1];
    const uint32_t P = (uint32_t)(rc.cw * rc.ch);
    for (int s = 0; s < rc.nb; ++s) {
        float4 v = samp[(size_t)s * P + pl];
        if (v.w != 0.0f) {
            wh += 1.0;
            r += (double)v.x;
            g += (double)v.y;
            b += (double)v.z;
        } else {
            wm += 1.0;
        }
    }
    rgb[3 * p] = r;
    rgb[3 * p + 1] = g;
    rgb[3 * p + 2] = b;
    wts[2 * p] = wh;
    wts[2 * p + 1] = wm;
}

// Film fold of a batched chunk: one thread per (segment, pixel of its view's crop); thread_base[k] is the
// first thread of segment k, seg_items[k] its item count. Samples are folded in sample order, so every
// view's film is the same as a render of that view alone.
__global__ __launch_bounds__(256) void k_film_batch(const RenderConsts *views, const uint4 *segs,
                                                    const uint32_t *thread_base, const uint32_t *seg_items, int n_segs,
                                                    uint32_t n_threads, const float4 *samp, double *const *rgbs,
                                                    double *const *wtss) {
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_threads) return;
    int lo = 0, hi = n_segs - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (thread_base[mid] <= t) lo = mid;
        else hi = mid - 1;
    }
    uint4 sg = segs[lo];
    uint32_t p = t - thread_base[lo];
    uint32_t nb = seg_items[lo] / sg.w;
    const RenderConsts &v = views[sg.y];
    uint32_t ry = p / (uint32_t)v.cw;
    size_t px = (size_t)((uint32_t)v.cy + ry) * (uint32_t)v.W + (uint32_t)v.cx + (p - ry * (uint32_t)v.cw);
    double *rgb = rgbs[sg.y], *wts = wtss[sg.y];
    double r = rgb[3 * px], g = rgb[3 * px + 1], b = rgb[3 * px + 2];
    double wh = wts[2 * px], wm = wts[2 * px + 1];
    for (uint32_t k = 0; k < nb; ++k) {
        float4 v4 = samp[(size_t)sg.x + (size_t)p * nb + k];
        if (v4.w != 0.0f) {
            wh += 1.0;
            r += (double)v4.x;
            g += (double)v4.y;
            b += (double)v4.z;
        } else {
            wm += 1.0;
        }
    }
    rgb[3 * px] = r;
    rgb[3 * px + 1] = g;
    rgb[3 * px + 2] = b;
    wts[2 * px] = wh;
    wts[2 * px + 1] = wm;
}

// relMSE of the composites of two films, mean over pixels and channels of (x - r)^2 / (r^2 + 0.01),
// composite as RenderTarget2 (render_target.cpp:187-192), FP64 throughout; one partial sum per block.
KS_DEV double composite_ch(const double *rgb, const double *wts, size_t p, int k, float backdrop) {
    double wh = wts[2 * p], wm = wts[2 * p + 1];
    double ws = (double)(uint32_t)(wh + wm);
    double a = ws == 0.0 ? 0.0 : wh / ws;
    double fg = wh == 0.0 ? 0.0 : rgb[3 * p + k] / wh;
    return (1.0 - a) * (double)backdrop + a * fg;
}

__global__ __launch_bounds__(256) void k_film_relmse(size_t n, const double *rgb, const double *wts, float3 bd,
                                                     const double *rrgb, const double *rwts, float3 rbd,
                                                     double *partials) {
    __shared__ double red[256];
    size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    double v = 0.0;
    if (p < n) {
        const float b[3] = {bd.x, bd.y, bd.z}, rb[3] = {rbd.x, rbd.y, rbd.z};
        for (int k = 0; k < 3; ++k) {
            double x = composite_ch(rgb, wts, p, k, b[k]);
            double r = composite_ch(rrgb, rwts, p, k, rb[k]);
            v += (x - r) * (x - r) / (r * r + 0.01);
        }
    }
    red[threadIdx.x] = v;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) partials[blockIdx.x] = red[0];
}

// Scratch of the allocation-free relMSE (ks_film_relmse_noalloc): the partial sums of at most
// kRelmseScratchBlocks blocks, then their total. Each device holds its own copy for the life of the process;
// host code reaches it only through the address hipGetSymbolAddress returns on that device
// (ks_device::relmse_scratch), never through this name.
constexpr unsigned kRelmseScratchBlocks = 2048;
static __device__ __align__(256) double scratch[kRelmseScratchBlocks + 1];

// k_film_relmse over a fixed grid: block b sums pixels b * 256 + t + k * gridDim.x * 256.
__global__ __launch_bounds__(256) void k_film_relmse_strided(size_t n, const double *rgb, const double *wts, float3 bd,
                                                             const double *rrgb, const double *rwts, float3 rbd,
                                                             double *partials) {
    __shared__ double red[256];
    double v = 0.0;
    const float b[3] = {bd.x, bd.y, bd.z}, rb[3] = {rbd.x, rbd.y, rbd.z};
    for (size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (size_t)gridDim.x * blockDim.x) {
        for (int k = 0; k < 3; ++k) {
            double x = composite_ch(rgb, wts, p, k, b[k]);
            double r = composite_ch(rrgb, rwts, p, k, rb[k]);
            v += (x - r) * (x - r) / (r * r + 0.01);
        }
    }
    red[threadIdx.x] = v;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) partials[blockIdx.x] = red[0];
}

// The total of n partial sums, in one block, written to *total.
__global__ __launch_bounds__(256) void k_sum_partials(const double *partials, unsigned n, double *total) {
    __shared__ double red[256];
    double v = 0.0;
    for (unsigned i = threadIdx.x; i < n; i += blockDim.x) v += partials[i];
    red[threadIdx.x] = v;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) *total = red[0];
}

// The sampler's permutation table: one thread per (dimension, digit prefix) entry, kPermTabDims x
// kPermTabPrefixes threads in all.
__global__ __launch_bounds__(256) void k_perm_tab(uint8_t *tab) {
    const uint32_t e = blockIdx.x * 256u + threadIdx.x;
    const uint32_t dim = e / kPermTabPrefixes, h = e % kPermTabPrefixes;
    tab[e] = (uint8_t)perm4_byte(perm_index(h ^ (0x55555555u * dim)));
}

__global__ void k_sobol_stream(int64_t n, const int32_t *px, const int32_t *py, const int32_t *sample, int32_t seed,
                               int32_t log2spp, int32_t ndigits, int32_t n_calls, const int32_t *kinds,
                               int32_t n_values, const uint8_t *perm_tab, uint32_t *out_u32, float *out_f) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Sobol sob;
    sob.start((uint32_t)px[i], (uint32_t)py[i], (uint32_t)sample[i], 0, seed, log2spp, ndigits, perm_tab);
    int64_t base = i * n_values;
    int k = 0;
    for (int c = 0; c < n_calls; ++c) {
        if (kinds[c] == 1) {
            uint32_t v = sob.next_u32();
            out_u32[base + k] = v;
            out_f[base + k] = u32_to_unit(v);
            ++k;
        } else {
            uint32_t a, b;
            sob.next2d_u32(a, b);
            out_u32[base + k] = a;
            out_f[base + k] = u32_to_unit(a);
            out_u32[base + k + 1] = b;
            out_f[base + k + 1] = u32_to_unit(b);
            k += 2;
        }
    }
}

__global__ void k_camera_rays(RenderConsts rc, int64_t n, const int32_t *px, const int32_t *py,
                              const int32_t *sample, float *rays) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Sobol sob;
    sobol_start(rc, sob, (uint32_t)px[i], (uint32_t)py[i], (uint32_t)sample[i], 0);
    f3 o, d;
    camera_ray(rc, (uint32_t)px[i], (uint32_t)py[i], sob, o, d);
    float *r = rays + 6 * i;
    r[0] = o.x, r[1] = o.y, r[2] = o.z, r[3] = d.x, r[4] = d.y, r[5] = d.z;
}

__global__ void k_bsdf_eval(Material m, int64_t n, const float *wo, const float *wi, float *f, float *pdf) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Closure c = closure_of(m);
    f3 o = mk(wo[3 * i], wo[3 * i + 1], wo[3 * i + 2]), w = mk(wi[3 * i], wi[3 * i + 1], wi[3 * i + 2]);
    f3 ff = bsdf_eval(o, w, c);
    f[3 * i] = ff.x, f[3 * i + 1] = ff.y, f[3 * i + 2] = ff.z;
    pdf[i] = bsdf_pdf(o, w, c);
}

__global__ void k_bsdf_sample(Material m, int64_t n, const float *wo, const float *u, float *wi, float *beta,
                              float *pdf) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Closure c = closure_of(m);
    f3 o = mk(wo[3 * i], wo[3 * i + 1], wo[3 * i + 2]);
    f3 w = mk(0, 0, 0);
    float pd = 0.0f;
    f3 b = bsdf_sample(o, w, c, u[3 * i], u[3 * i + 1], u[3 * i + 2], pd);
    wi[3 * i] = w.x, wi[3 * i + 1] = w.y, wi[3 * i + 2] = w.z;
    beta[3 * i] = b.x, beta[3 * i + 1] = b.y, beta[3 * i + 2] = b.z;
    pdf[i] = pd;
}

// ---------------------------------------------------------------------------------------------------
// Host-side helpers
// ---------------------------------------------------------------------------------------------------
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
bool inverse(const M4 &m, M4 &out) {
    double a[4][8];
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 8; ++j) a[i][j] = j < 4 ? m.a[i][j] : (j - 4 == i ? 1.0 : 0.0);
    for (int c = 0; c < 4; ++c) {
        int p = c;
        for (int r = c + 1; r < 4; ++r)
            if (std::fabs(a[r][c]) > std::fabs(a[p][c])) p = r;
        if (a[p][c] == 0.0) return false;
        for (int j = 0; j < 8; ++j) std::swap(a[c][j], a[p][j]);
        double d = a[c][c];
        for (int j = 0; j < 8; ++j) a[c][j] /= d;
        for (int r = 0; r < 4; ++r)
            if (r != c) {
                double f = a[r][c];
                for (int j = 0; j < 8; ++j) a[r][j] -= f * a[c][j];
            }
    }
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) out.a[i][j] = a[i][j + 4];
    return true;
}

// Camera(to_world, vfov, aspect) (camera.cpp:74-82) with look_at (camera.cpp:12-32) and
// rev_inf_projection (camera.cpp:56-70, near_clip 0.1): proj_to_world and the camera position.
// Camera matrices (camera.cpp:12-127) folded for camera_ray, and the camera ray differentials' film steps
// for a film of W x H (camera.cpp:133-140): proj_to_camera maps (2/W, 0, 1) and (0, 2/H, 1) next to the
// film centre, and camera_to_world turns the differences into world directions.
int camera_setup(const ks_camera_desc &c, int W, int H, float p2w[16], float pos[3], float cam_dx[3],
                 float cam_dy[3]) {
    M4 c2w;
    if (c.use_matrix) {
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j) c2w.a[i][j] = c.to_world[4 * i + j];
    } else {
        float up[3] = {c.up[0], c.up[1], c.up[2]};
        float ul = std::sqrt(up[0] * up[0] + up[1] * up[1] + up[2] * up[2]);
        KS_REQUIRE(ul > 0.0f, "camera up vector is zero");
        for (float &u : up) u /= ul; // load_vec3("up", force_normalize = true)
        float back[3] = {c.pos[0] - c.target[0], c.pos[1] - c.target[1], c.pos[2] - c.target[2]};
        float bl = std::sqrt(back[0] * back[0] + back[1] * back[1] + back[2] * back[2]);
        KS_REQUIRE(bl > 0.0f, "camera position equals target");
        for (float &b : back) b /= bl;
        float right[3] = {up[1] * back[2] - up[2] * back[1], up[2] * back[0] - up[0] * back[2],
                          up[0] * back[1] - up[1] * back[0]};
        float rl2 = right[0] * right[0] + right[1] * right[1] + right[2] * right[2];
        if (rl2 < 1e-8f) {
            float sign = std::copysign(1.0f, back[2]);
            const float a = -1.0f / (sign + back[2]);
            const float b = back[0] * back[1] * a;
            right[0] = 1.0f + sign * back[0] * back[0] * a, right[1] = sign * b, right[2] = -sign * back[0];
            up[0] = b, up[1] = sign + back[1] * back[1] * a, up[2] = -back[1];
        } else {
            float rl = std::sqrt(rl2);
            for (float &r : right) r /= rl;
            up[0] = back[1] * right[2] - back[2] * right[1];
            up[1] = back[2] * right[0] - back[0] * right[2];
            up[2] = back[0] * right[1] - back[1] * right[0];
        }
        double m[4][4] = {{right[0], up[0], back[0], c.pos[0]},
                          {right[1], up[1], back[1], c.pos[1]},
                          {right[2], up[2], back[2], c.pos[2]},
                          {0, 0, 0, 1}};
        std::memcpy(c2w.a, m, sizeof(m));
    }
    KS_REQUIRE(c.vfov_deg > 0.0f && c.vfov_deg < 180.0f && c.aspect > 0.0f, "invalid camera vfov/aspect");
    float vfov = c.vfov_deg / 180.0f * 3.14159265359f;
    float cot = 1.0f / std::tan(vfov * 0.5f);
    M4 proj{{{cot / c.aspect, 0, 0, 0}, {0, -cot, 0, 0}, {0, 0, 0, 0.1f}, {0, 0, -1, 0}}};
    M4 w2c, p2w_m;
    KS_REQUIRE(inverse(c2w, w2c), "camera to_world is singular");
    M4 w2p = mul(proj, w2c);
    KS_REQUIRE(inverse(w2p, p2w_m), "camera projection is singular");
    double eye[3] = {c2w.a[0][3], c2w.a[1][3], c2w.a[2][3]};
    for (int j = 0; j < 4; ++j) {
        for (int i = 0; i < 3; ++i) p2w[4 * i + j] = (float)(p2w_m.a[i][j] - eye[i] * p2w_m.a[3][j]);
        p2w[12 + j] = (float)p2w_m.a[3][j];
    }
    pos[0] = (float)eye[0], pos[1] = (float)eye[1], pos[2] = (float)eye[2];
    M4 p2c;
    KS_REQUIRE(inverse(proj, p2c), "camera projection is singular");
    auto hdiv = [&](double x, double y, double out[3]) {
        double h[4];
        for (int i = 0; i < 4; ++i) h[i] = p2c.a[i][0] * x + p2c.a[i][1] * y + p2c.a[i][2] * 1.0 + p2c.a[i][3];
        for (int i = 0; i < 3; ++i) out[i] = h[i] / h[3];
    };
    double mid[3], right[3], up[3];
    hdiv(0.0, 0.0, mid);
    hdiv((double)(2.0f / (float)W), 0.0, right);
    hdiv(0.0, (double)(2.0f / (float)H), up);
    for (int i = 0; i < 3; ++i) {
        double dx = 0.0, dy = 0.0;
        for (int j = 0; j < 3; ++j) {
            dx += c2w.a[i][j] * (right[j] - mid[j]);
            dy += c2w.a[i][j] * (up[j] - mid[j]);
        }
        cam_dx[i] = (float)dx;
        cam_dy[i] = (float)dy;
    }
    return KS_OK;
}

int log2_int(uint32_t v) { return 31 - __builtin_clz(v); }
int32_t round_up_pow2(int32_t v) {
    v--;
    v |= v >> 1;
    v |= v >> 2;
    v |= v >> 4;
    v |= v >> 8;
    v |= v >> 16;
    return v + 1;
}
// OwenScrambledSobol02Morton ctor (sobol.h:133-140)
void sobol_consts(int W, int H, int spp, int &log2spp, int &ndigits) {
    log2spp = log2_int((uint32_t)spp);
    int res = round_up_pow2(std::max(W, H));
    ndigits = log2_int((uint32_t)res) + (log2spp + 1) / 2;
}

Material to_device_material(const ks_material_desc &m) { // PrincipledBSDF::eval_closure clamps
    Material d{};
    for (int k = 0; k < 3; ++k) d.basecolor[k] = std::min(std::max(m.basecolor[k], 0.0f), 1.0f);
    d.roughness = std::min(std::max(m.roughness, 0.0f), 1.0f);
    d.ax = d.roughness * d.roughness;
    d.ay = d.ax;
    d.metallic = std::min(std::max(m.metallic, 0.0f), 1.0f);
    d.ior = std::min(std::max(m.ior, 1.0f), 4.0f);
    d.specular_r0_mul = std::max(m.specular_r0_mul * 2.0f, 0.0f);
    d.specular_trans = std::min(std::max(m.specular_trans, 0.0f), 1.0f);
    d.diffuse_trans = std::min(std::max(m.diffuse_trans, 0.0f), 1.0f);
    d.diffuse_trans_fwd = std::min(std::max(m.diffuse_trans_fwd, 0.0f), 1.0f);
    for (int k = 0; k < 3; ++k) d.emission[k] = m.emission[k];
    d.has_emission = m.has_emission ? 1 : 0;
    return d;
}

} // namespace

// ---------------------------------------------------------------------------------------------------
// Handles
// ---------------------------------------------------------------------------------------------------
struct Workspace {
    uint32_t Q = 0;
    DevBuf work[2], state[2], beta[2], L[2], ro[2], rd[2];
    DevBuf hit, sh_o, sh_d, sh_tgt, sh_c;
    DevBuf order; // material-sorted shading: kShadeClasses lists of Q slots
    DevBuf samp;
    DevBuf ctrl, stats;
    DevBuf ovf; // traversal stack spill area, one column per lane of the largest traversal grid
    uint32_t ovf_grid = 0, ovf_cap = 0;
};

// Buffers of guided renders: per-path mixtures of the current wavefront buffer, per-work-item training
// records, the extracted training samples and their counter.
struct GuideWork {
    DevBuf mix, list, recs, train, count;
    uint32_t mix_q = 0, rec_items = 0;
    uint32_t stamp = 0; // last record stamp handed out on this device
};

struct ks_device {
    int ordinal = 0;
    hipStream_t stream = nullptr;
    int timing = 0;
    int n_cu = 0;
    Workspace ws;
    GuideWork gw;
    // last batched-render autotuning result: scene, workload size -> queue size, traversal blocks per CU
    const ks_scene *tune_scene = nullptr;
    uint64_t tune_items = 0;
    uint32_t tune_q = 0;
    int tune_bpc = 0;
    Ctrl *host_ctrl = nullptr; // pinned
    uint8_t *host_scratch = nullptr; // pinned, 64 bytes
    double *relmse_scratch = nullptr; // this device's address of `scratch` (hipGetSymbolAddress at open)
    DevBuf perm_tab; // the sampler's permutation bytes (kPermTabDims x kPermTabPrefixes), filled at open
    int sorted_shading = 0;           // material-sorted shading of untextured scenes (ks_device_set_shading)
    uint32_t tail_paths = 0;          // k_tail takes over at or below this many paths (ks_device_set_tail; 0: off)
    std::vector<hipEvent_t> events; // timing event pool, reused across renders
    // pixel-filter sample tables by (type, radius, sigma), kept for the device's lifetime: batched
    // renders keep pointers to them in their per-view constants
    struct FilterTable {
        int type;
        float r, sigma;
        int n;
        DevBuf cdf;
    };
    std::vector<std::unique_ptr<FilterTable>> filter_tables;
};

struct ks_scene {
    ks_device *dev = nullptr;
    DevBuf nodes, tris, verts, normals, tri_idx, tri_light, mats, mat_class, lights, cdf;
    DevBuf sky_rgb, sky_cond, sky_marg; // environment map and its sampling tables
    DevBuf tex_pool, tex_meta, tex_mats, tex_mat_of, uvs, tex_luts; // textures and textured materials
    DevScene ds{};
    ks_scene_info info{};
    std::vector<int32_t> light_prim;
    std::vector<float> light_power, light_cdf;
    uint32_t max_stack = 1; // deepest traversal stack the tree can need
    // host SBVH in its BVH2 form (host builder only), for the reference-order traversal count
    std::vector<ksa::GpuNode> host_nodes2;
    std::vector<ksa::GpuTri> host_tris;
};

struct ks_film {
    ks_device *dev = nullptr;
    int32_t w = 0, h = 0;
    float backdrop[3] = {0, 0, 0};
    DevBuf rgb, wts;
};

struct ks_host_scene {
    ksa::HostScene s;
};

namespace {

template <typename T> hipError_t upload(DevBuf &b, const std::vector<T> &v, hipStream_t s) {
    hipError_t e = b.alloc(v.size() * sizeof(T));
    if (e != hipSuccess) return e;
    if (v.empty()) return hipSuccess;
    e = hipMemcpyAsync(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return e;
    return hipStreamSynchronize(s);
}

// HIP-event timing of kernel classes on the render stream; events come from a per-device pool.
struct Timer {
    ks_device *dev;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev[KS_K_COUNT];
    uint64_t launches[KS_K_COUNT] = {};
    size_t used = 0;
    hipEvent_t take();
    hipEvent_t begin(int k) {
        launches[k]++;
        if (!dev->timing) return nullptr;
        hipEvent_t a = take(), b = take();
        if (!a || !b) return nullptr;
        (void)hipEventRecord(a, dev->stream);
        ev[k].push_back({a, b});
        return b;
    }
    void end(hipEvent_t b) {
        if (b) (void)hipEventRecord(b, dev->stream);
    }
    void collect(ks_render_stats *st) {
        for (int k = 0; k < KS_K_COUNT; ++k) {
            double ms = 0.0;
            for (auto &p : ev[k]) {
                float t = 0.0f;
                if (hipEventElapsedTime(&t, p.first, p.second) == hipSuccess) ms += t;
            }
            ev[k].clear();
            if (st) {
                st->kernel_ms[k] += ms;
                st->launches[k] += launches[k];
            }
        }
    }
};

hipEvent_t Timer::take() {
    if (used == dev->events.size()) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        dev->events.push_back(e);
    }
    return dev->events[used++];
}

// Traversal launches for the scene's node width (2 or 4), with the opacity filter when the scene has
// alpha-tested triangles.
template <typename... A> void launch_closest(const DevScene &ds, dim3 g, hipStream_t s, A... args) {
    const dim3 b(kTraceBlock);
    if (ds.alpha) {
        if (ds.width == 4) hipLaunchKernelGGL((k_closest<4, true>), g, b, 0, s, ds, args...);
        else hipLaunchKernelGGL((k_closest<2, true>), g, b, 0, s, ds, args...);
    } else {
        if (ds.width == 4) hipLaunchKernelGGL((k_closest<4, false>), g, b, 0, s, ds, args...);
        else hipLaunchKernelGGL((k_closest<2, false>), g, b, 0, s, ds, args...);
    }
}
template <typename... A> void launch_trace(const DevScene &ds, dim3 g, hipStream_t s, A... args) {
    const dim3 b(kTraceBlock);
    if (ds.alpha) {
        if (ds.width == 4) hipLaunchKernelGGL((k_trace<4, true>), g, b, 0, s, ds, args...);
        else hipLaunchKernelGGL((k_trace<2, true>), g, b, 0, s, ds, args...);
    } else {
        if (ds.width == 4) hipLaunchKernelGGL((k_trace<4, false>), g, b, 0, s, ds, args...);
        else hipLaunchKernelGGL((k_trace<2, false>), g, b, 0, s, ds, args...);
    }
}
template <typename... A> void launch_tail(const DevScene &ds, bool batch, dim3 g, hipStream_t s, A... args) {
    const dim3 b(256);
    const bool tex = ds.textured != 0;
    if (ds.width == 4) {
        if (batch) {
            if (tex) hipLaunchKernelGGL((k_tail<4, true, true>), g, b, 0, s, ds, args...);
            else hipLaunchKernelGGL((k_tail<4, true, false>), g, b, 0, s, ds, args...);
        } else {
            if (tex) hipLaunchKernelGGL((k_tail<4, false, true>), g, b, 0, s, ds, args...);
            else hipLaunchKernelGGL((k_tail<4, false, false>), g, b, 0, s, ds, args...);
        }
    } else {
        if (batch) {
            if (tex) hipLaunchKernelGGL((k_tail<2, true, true>), g, b, 0, s, ds, args...);
            else hipLaunchKernelGGL((k_tail<2, true, false>), g, b, 0, s, ds, args...);
        } else {
            if (tex) hipLaunchKernelGGL((k_tail<2, false, true>), g, b, 0, s, ds, args...);
            else hipLaunchKernelGGL((k_tail<2, false, false>), g, b, 0, s, ds, args...);
        }
    }
}
template <typename... A> void launch_anyhit(const DevScene &ds, dim3 g, hipStream_t s, A... args) {
    const dim3 b(kTraceBlock);
    if (ds.alpha) {
        if (ds.width == 4) hipLaunchKernelGGL((k_anyhit<4, true>), g, b, 0, s, ds, args...);
        else hipLaunchKernelGGL((k_anyhit<2, true>), g, b, 0, s, ds, args...);
    } else {
        if (ds.width == 4) hipLaunchKernelGGL((k_anyhit<4, false>), g, b, 0, s, ds, args...);
        else hipLaunchKernelGGL((k_anyhit<2, false>), g, b, 0, s, ds, args...);
    }
}

// Shading launch: guided (plain renders only), batched or plain, each with or without texture support.
template <bool GUIDED, bool BATCH, typename... A> void launch_sorted(dim3 g, hipStream_t s, A... args) {
    const dim3 b(256);
    hipLaunchKernelGGL((k_shade<GUIDED, BATCH, false, kShadeMiss>), g, b, 0, s, args...);
    hipLaunchKernelGGL((k_shade<GUIDED, BATCH, false, kShadeDielectric>), g, b, 0, s, args...);
    hipLaunchKernelGGL((k_shade<GUIDED, BATCH, false, kShadeMetal>), g, b, 0, s, args...);
    hipLaunchKernelGGL((k_shade<GUIDED, BATCH, false, kShadeGeneral>), g, b, 0, s, args...);
}

// Shading of the current buffer: material-sorted (one kernel per class over the lists k_classify wrote)
// for untextured scenes, else one kernel over the whole buffer.
template <typename... A> void launch_shade(bool tex, bool guided, bool batch, bool sorted, dim3 g, hipStream_t s,
                                           A... args) {
    const dim3 b(256);
    if (sorted && !tex) {
        if (guided) launch_sorted<true, false>(g, s, args...);
        else if (batch) launch_sorted<false, true>(g, s, args...);
        else launch_sorted<false, false>(g, s, args...);
    } else if (guided) {
        if (tex) hipLaunchKernelGGL((k_shade<true, false, true>), g, b, 0, s, args...);
        else hipLaunchKernelGGL((k_shade<true, false, false>), g, b, 0, s, args...);
    } else if (batch) {
        if (tex) hipLaunchKernelGGL((k_shade<false, true, true>), g, b, 0, s, args...);
        else hipLaunchKernelGGL((k_shade<false, true, false>), g, b, 0, s, args...);
    } else {
        if (tex) hipLaunchKernelGGL((k_shade<false, false, true>), g, b, 0, s, args...);
        else hipLaunchKernelGGL((k_shade<false, false, false>), g, b, 0, s, args...);
    }
}

constexpr int kMaxTraceBlocksPerCu = 8; // the traversal overflow area is sized for this many

// Per-sample results of one sub-batch (chunk) of a render: 16 B per path. Every chunk ends with a drain of
// the wavefront (a dozen iterations over the last, longest paths with the GPU mostly idle), so chunks are
// made as large as this buffer allows: 4 GiB holds a 1920x1080 64-spp render in one chunk.
constexpr uint64_t kSampleBufferBytes = 4ull << 30;

int trace_grid(const ks_device *dev, const DevScene &ds, uint32_t n, int blocks_per_cu = kMaxTraceBlocksPerCu) {
    // 8 blocks of 256 lanes per CU keep every SIMD at 8 waves; the grid-stride loop covers the rest. The
    // alpha-tested kernels run at 4 waves per SIMD (their opacity call needs 128 VGPRs), so 4 blocks per CU
    // are resident and the grid stays at that.
    const int cap = ds.alpha ? kMaxTraceBlocksPerCu / 2 : kMaxTraceBlocksPerCu;
    uint32_t max_blocks = (uint32_t)std::max(1, dev->n_cu) * (uint32_t)std::max(1, std::min(blocks_per_cu, cap));
    uint32_t need = (n + kTraceBlock - 1) / kTraceBlock;
    return (int)std::max(1u, std::min(max_blocks, need));
}

// The statistics buffer: kStatCount slots in each of 2 + R rows (k_stat_reduce), R = the largest
// traversal grid.
uint32_t stat_partial_rows(const ks_device *dev) {
    return (uint32_t)std::max(1, dev->n_cu) * (uint32_t)kMaxTraceBlocksPerCu;
}
unsigned long long *stat_rows(ks_device *dev) { return dev->ws.stats.as<unsigned long long>() + 2 * kStatCount; }
void clear_stats(ks_device *dev, hipStream_t s) {
    const int n = (int)(2 + stat_partial_rows(dev)) * kStatCount;
    hipLaunchKernelGGL(k_fill_u64, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                       dev->ws.stats.as<unsigned long long>(), n);
}

int ensure_workspace(ks_device *dev, uint32_t Q, size_t n_samp, uint32_t max_stack) {
    Workspace &ws = dev->ws;
    if (ws.Q < Q) {
        for (int b = 0; b < 2; ++b) {
            KS_HIP(ws.work[b].alloc(sizeof(uint32_t) * Q));
            KS_HIP(ws.state[b].alloc(sizeof(uint32_t) * Q));
            KS_HIP(ws.beta[b].alloc(sizeof(float4) * Q));
            KS_HIP(ws.L[b].alloc(sizeof(float4) * Q));
            KS_HIP(ws.ro[b].alloc(sizeof(float4) * Q));
            KS_HIP(ws.rd[b].alloc(sizeof(float4) * Q));
        }
        KS_HIP(ws.hit.alloc(sizeof(float4) * Q));
        KS_HIP(ws.sh_o.alloc(sizeof(float4) * Q));
        KS_HIP(ws.sh_d.alloc(sizeof(float4) * Q));
        KS_HIP(ws.sh_tgt.alloc(sizeof(uint32_t) * Q));
        KS_HIP(ws.sh_c.alloc(sizeof(float4) * Q));
        KS_HIP(ws.order.alloc(sizeof(uint32_t) * kShadeClasses * (size_t)Q));
        ws.Q = Q;
    }
    KS_HIP(ws.samp.ensure(sizeof(float4) * std::max<size_t>(n_samp, 1)));
    KS_HIP(ws.ctrl.ensure(sizeof(Ctrl)));
    KS_HIP(ws.stats.ensure(sizeof(unsigned long long) * kStatCount * (2 + (size_t)stat_partial_rows(dev))));
    uint32_t grid = (uint32_t)std::max(1, dev->n_cu) * (uint32_t)kMaxTraceBlocksPerCu * kTraceBlock;
    uint32_t cap = max_stack + 4 > (uint32_t)kStackLds ? max_stack + 4 - kStackLds : 1u;
    if (ws.ovf_grid < grid || ws.ovf_cap < cap) {
        KS_HIP(ws.ovf.alloc(sizeof(uint32_t) * (size_t)grid * cap));
        ws.ovf_grid = grid;
        ws.ovf_cap = cap;
    }
    return KS_OK;
}

// The filter fields of rc: the box filter's radius, or a tabulated filter's table on the device (built
// once per filter on the device and kept).
int filter_consts(ks_device *dev, const ks_small_pt_params *p, RenderConsts &rc) {
    ksa::PixelFilterSpec f;
    KS_REQUIRE(ksa::pixel_filter_spec(*p, f), "pixel_filter: unknown type, or width / sigma not positive");
    rc.filter_r = f.r;
    rc.filter_n = 0;
    rc.filter_cdf = nullptr;
    if (f.type == KS_FILTER_BOX) return KS_OK;
    for (auto &t : dev->filter_tables)
        if (t->type == f.type && t->r == f.r && (f.type != KS_FILTER_GAUSSIAN || t->sigma == f.sigma)) {
            rc.filter_n = t->n;
            rc.filter_cdf = t->cdf.as<float>();
            return KS_OK;
        }
    auto t = std::make_unique<ks_device::FilterTable>();
    t->type = f.type;
    t->r = f.r;
    t->sigma = f.sigma;
    std::vector<float> table = ksa::pixel_filter_table(f, t->n);
    KS_HIP(hipSetDevice(dev->ordinal));
    KS_HIP(t->cdf.alloc(table.size() * sizeof(float)));
    KS_HIP(hipMemcpy(t->cdf.p, table.data(), table.size() * sizeof(float), hipMemcpyHostToDevice));
    rc.filter_n = t->n;
    rc.filter_cdf = t->cdf.as<float>();
    dev->filter_tables.push_back(std::move(t));
    return KS_OK;
}

int fill_consts(ks_device *dev, const ks_scene *sc, const ks_camera_desc *cam, const ks_small_pt_params *p,
                RenderConsts &rc) {
    KS_REQUIRE(p->render_width > 0 && p->render_height > 0, "render_width/render_height must be positive");
    KS_REQUIRE(p->spp > 0, "spp must be positive");
    KS_REQUIRE(p->bounces >= 0 && p->bounces < 255, "bounces must be in [0, 254]");
    KS_REQUIRE(std::max(p->render_width, p->render_height) <= 65536, "resolution above 65536");
    std::memset(&rc, 0, sizeof(rc));
    rc.W = p->render_width;
    rc.H = p->render_height;
    // crop (small_pt.cpp:23-35)
    rc.cx = 0, rc.cy = 0, rc.cw = rc.W, rc.ch = rc.H;
    if (p->crop_width > 0 && p->crop_height > 0) {
        rc.cx = std::max(p->crop_start_x, 0);
        rc.cy = std::max(p->crop_start_y, 0);
        rc.cw = std::min(p->crop_width, rc.W - rc.cx);
        rc.ch = std::min(p->crop_height, rc.H - rc.cy);
        KS_REQUIRE(rc.cw > 0 && rc.ch > 0, "crop window is empty");
    }
    rc.spp = p->spp;
    rc.seed = p->rng_seed;
    sobol_consts(rc.W, rc.H, rc.spp, rc.log2spp, rc.ndigits);
    rc.perm_tab = dev->perm_tab.as<uint8_t>();
    KS_REQUIRE(2 * rc.ndigits <= 52 && (uint64_t)1 << rc.log2spp <= (uint64_t)rc.spp,
               "resolution x spp exceeds the sampler index range");
    rc.bounces = p->bounces;
    rc.include_background = p->include_background;
    rc.mis_mode = p->mis_mode;
    rc.clamp_direct = p->clamp_direct;
    rc.clamp_indirect = p->clamp_indirect;
    int e = camera_setup(*cam, rc.W, rc.H, rc.p2w, rc.cam, rc.cam_dx, rc.cam_dy);
    // Camera::spawn_ray(film_pos, film_res, scale_ray_diff ? spp : 1) scales by 1 / sqrt(that count)
    rc.diff_scale = 1.0f / std::sqrt((float)(p->scale_ray_diff ? p->spp : 1));
    (void)sc;
    if (e != KS_OK) return e;
    return filter_consts(dev, p, rc);
}

// Guide state of one guided wavefront pass: the guide (nullptr: plain), its per-path mixture buffer and
// the shading parameters.
struct GuidePass {
    ks_guide *guide = nullptr;
    float4 *mix = nullptr;
    uint32_t *list = nullptr, *n_list = nullptr; // compacted inference queries (Q slots) and their count
    GuideShade shade{};
};

// Options of one wavefront run: batched work mapping, HIP-graph replay of the iterations, traversal
// blocks per CU.
struct WavefrontOpts {
    bool batch = false;
    bool graph = false;
    int blocks_per_cu = kMaxTraceBlocksPerCu;
    // ray capture (eager launches only): every closest-hit and shadow ray of the run, appended as
    // (o, tmin), (d, tmax) pairs to cap_c / cap_s up to cap_max each; cap_n[0], cap_n[1] count them
    float4 *cap_c = nullptr, *cap_s = nullptr;
    uint32_t *cap_n = nullptr;
    uint32_t cap_max = 0;
};

// Runs the wavefront loop over `total` work items (mapped by rc), leaving per-sample results in ws.samp.
// With opts.graph, two iterations (buffer 0 -> 1 and 1 -> 0) are captured once into a HIP graph and the
// graph is replayed; the kernels read their counts from the control block, so every replay is valid.
// Ray capture: append the n rays of a queue after the *count already captured, up to cap rays.
__global__ __launch_bounds__(256) void k_capture(const uint32_t *n, const float4 *ro, const float4 *rd, float4 *out,
                                                 const uint32_t *count, uint32_t cap) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t base = *count;
    if (i < *n && base + (uint64_t)i < cap) {
        out[2 * (size_t)(base + i)] = ro[i];
        out[2 * (size_t)(base + i) + 1] = rd[i];
    }
}
__global__ void k_capture_done(const uint32_t *n, uint32_t *count, uint32_t cap) {
    *count = (uint32_t)min((uint64_t)cap, (uint64_t)*count + *n);
}

int run_wavefront(ks_scene *sc, RenderConsts rc, uint64_t total, uint32_t Q, Timer &tm, ks_render_stats *stats,
                  const GuidePass *gp = nullptr, const WavefrontOpts &wo = WavefrontOpts{}) {
    ks_device *dev = sc->dev;
    Workspace &ws = dev->ws;
    hipStream_t s = dev->stream;
    Ctrl *ctrl = ws.ctrl.as<Ctrl>();
    unsigned long long *st = ws.stats.as<unsigned long long>();
    hipLaunchKernelGGL(k_ctrl_init, dim3(1), dim3(1), 0, s, ctrl, (uint64_t)total);
    KS_HIP(hipGetLastError());
    PathBufs pb[2];
    for (int b = 0; b < 2; ++b)
        pb[b] = PathBufs{ws.work[b].as<uint32_t>(), ws.state[b].as<uint32_t>(), ws.beta[b].as<float4>(),
                         ws.L[b].as<float4>(),     ws.ro[b].as<float4>(),     ws.rd[b].as<float4>()};
    const int tblocks = (int)((Q + 255) / 256);
    // Blocks of ray generation and shading: one thread per queue slot until a control read bounds the
    // work left. From a read on, an iteration holds at most the paths then active plus the work items not
    // yet started, and generates at most the latter, so the late iterations' launches stay small.
    int gen_blocks = tblocks, shade_blocks = tblocks;
    const int tgrid = trace_grid(dev, sc->ds, Q, wo.blocks_per_cu);
    const bool graph = wo.graph && !gp;
    // One iteration reading buffer cur and writing buffer 1 - cur; no events while a graph is captured.
    // Its traversal also traces the previous shading pass's shadow rays (k_trace): their contributions go
    // to the survivors' slots [0, n_next) of buffer cur (ray generation fills the slots after them) and to
    // per-sample results, before this iteration's shading reads the slots and refills the shadow queue.
    // Guide inference serves only vertices below the guide's max bounce. A path generated in iteration g
    // is at bounce i - g in iteration i, so once every work item has started by iteration last_gen (known
    // from the start when the queue holds the whole pass, else from a control read), iterations at or
    // beyond last_gen + max bounce have no guided vertex and skip the inference.
    int iter_index = 0;
    int last_gen = total <= (uint64_t)Q ? 0 : -1; // -1: not known yet
    auto iteration = [&](int cur) -> int {
        hipEvent_t e = graph ? nullptr : tm.begin(KS_K_RAYGEN);
        if (wo.batch) hipLaunchKernelGGL(k_raygen<true>, dim3(gen_blocks), dim3(256), 0, s, rc, pb[cur], ctrl, Q);
        else hipLaunchKernelGGL(k_raygen<false>, dim3(gen_blocks), dim3(256), 0, s, rc, pb[cur], ctrl, Q);
        hipLaunchKernelGGL(k_gen_done, dim3(1), dim3(1), 0, s, ctrl, Q);
        tm.end(e);
        if (wo.cap_n) {
            hipLaunchKernelGGL(k_capture, dim3(tblocks), dim3(256), 0, s, (const uint32_t *)&ctrl->n_active,
                               (const float4 *)pb[cur].ro, (const float4 *)pb[cur].rd, wo.cap_c, wo.cap_n, wo.cap_max);
            hipLaunchKernelGGL(k_capture_done, dim3(1), dim3(1), 0, s, (const uint32_t *)&ctrl->n_active, wo.cap_n,
                               wo.cap_max);
        }
        e = graph ? nullptr : tm.begin(KS_K_CLOSEST);
        launch_trace(sc->ds, dim3(tgrid), s, (const uint32_t *)&ctrl->n_active, (const float4 *)pb[cur].ro,
                     (const float4 *)pb[cur].rd, ws.hit.as<float4>(), (const uint32_t *)&ctrl->n_shadow_trace,
                     (const float4 *)ws.sh_o.as<float4>(), (const float4 *)ws.sh_d.as<float4>(),
                     (const uint32_t *)ws.sh_tgt.as<uint32_t>(), (const float4 *)ws.sh_c.as<float4>(), pb[cur].L,
                     ws.samp.as<float4>(), gp ? gp->shade.recs : (float4 *)nullptr, ws.ovf.as<uint32_t>(), ws.ovf_cap,
                     stat_rows(dev));
        tm.end(e);
        auto young = [&](int b) { return last_gen < 0 || iter_index - last_gen < b; };
        const bool guided_vertices = gp && young(gp->shade.max_bounce);
        // without guided vertices and training records (the first kRecVertices vertices of a path) the
        // guided shading is the plain one, whose kernel holds fewer registers
        const bool guided_shade = gp && (guided_vertices || (gp->shade.recs && young(kRecVertices)));
        if (gp && gp->shade.mix && guided_vertices) {
            e = tm.begin(KS_K_GUIDE);
            // the guided queries are among the paths shading takes: the same launch bound holds
            const uint32_t n_max = (uint32_t)std::min<uint64_t>(Q, (uint64_t)shade_blocks * 256u);
            int ge = ksg::launch_infer(gp->guide, sc->ds, &ctrl->n_active, n_max, pb[cur].rd, ws.hit.as<float4>(),
                                       pb[cur].state, gp->shade.max_bounce, gp->mix, gp->list, gp->n_list, s);
            tm.end(e);
            if (ge) return ge;
        }
        e = graph ? nullptr : tm.begin(KS_K_SHADE);
        const bool sorted = dev->sorted_shading && !sc->ds.textured;
        if (sorted)
            hipLaunchKernelGGL(k_classify, dim3(tblocks), dim3(256), 0, s, sc->ds, (const float4 *)ws.hit.as<float4>(),
                               ctrl, ws.order.as<uint32_t>(), Q);
        launch_shade(sc->ds.textured != 0, guided_shade, wo.batch, sorted, dim3(shade_blocks), s, sc->ds, rc, pb[cur],
                     pb[1 - cur], ws.hit.as<float4>(), ctrl, ws.sh_o.as<float4>(), ws.sh_d.as<float4>(),
                     ws.sh_tgt.as<uint32_t>(), ws.sh_c.as<float4>(), ws.samp.as<float4>(), st,
                     gp ? gp->shade : GuideShade{}, (const uint32_t *)ws.order.as<uint32_t>(), Q);
        tm.end(e);
        if (wo.cap_n) {
            hipLaunchKernelGGL(k_capture, dim3(tblocks), dim3(256), 0, s, (const uint32_t *)&ctrl->n_shadow.v,
                               (const float4 *)ws.sh_o.as<float4>(), (const float4 *)ws.sh_d.as<float4>(), wo.cap_s,
                               wo.cap_n + 1, wo.cap_max);
            hipLaunchKernelGGL(k_capture_done, dim3(1), dim3(1), 0, s, (const uint32_t *)&ctrl->n_shadow.v, wo.cap_n + 1,
                               wo.cap_max);
        }
        hipLaunchKernelGGL(k_iter_done, dim3(1), dim3(1), 0, s, ctrl);
        return KS_OK;
    };
    // The last shading pass's shadow rays: a traversal of them alone, once the loop has ended (or before
    // the tail kernel takes the remaining paths).
    auto flush_shadow = [&](int cur) -> int {
        hipLaunchKernelGGL(k_gen_done, dim3(1), dim3(1), 0, s, ctrl, Q); // no work left: only moves the queue
        hipEvent_t e = tm.begin(KS_K_SHADOW);
        launch_trace(sc->ds, dim3(tgrid), s, (const uint32_t *)&ctrl->none, (const float4 *)pb[cur].ro,
                     (const float4 *)pb[cur].rd, ws.hit.as<float4>(), (const uint32_t *)&ctrl->n_shadow_trace,
                     (const float4 *)ws.sh_o.as<float4>(), (const float4 *)ws.sh_d.as<float4>(),
                     (const uint32_t *)ws.sh_tgt.as<uint32_t>(), (const float4 *)ws.sh_c.as<float4>(), pb[cur].L,
                     ws.samp.as<float4>(), gp ? gp->shade.recs : (float4 *)nullptr, ws.ovf.as<uint32_t>(), ws.ovf_cap,
                     stat_rows(dev));
        tm.end(e);
        KS_HIP(hipGetLastError());
        return KS_OK;
    };
    hipGraphExec_t exec = nullptr;
    hipGraph_t g = nullptr;
    if (graph) {
        KS_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        int e0 = iteration(0);
        int e1 = iteration(1);
        hipError_t ce = hipStreamEndCapture(s, &g);
        if (e0 || e1) {
            if (g) (void)hipGraphDestroy(g);
            return e0 ? e0 : e1;
        }
        KS_HIP(ce);
        hipError_t ie = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
        if (ie != hipSuccess) {
            (void)hipGraphDestroy(g);
            KS_HIP(ie);
        }
    }
    int cur = 0;
    int iterations = 0;
    const int check_every = 8; // empty trailing iterations cost only launch latency
    int status = KS_OK;
    // The tail: once every work item has started and at most tail_limit paths remain, one k_tail launch
    // carries them to their ends (eager plain renders of scenes without alpha-tested triangles; not while
    // rays are captured). The control block is read after every iteration from the first read that finds
    // the remaining work items within check_every iterations of generation.
    const uint32_t tail_limit = std::min<uint32_t>(dev->tail_paths, ws.ovf_grid);
    const bool tail_ok = !graph && !gp && !wo.cap_n && !sc->ds.alpha && tail_limit > 0;
    bool check_each = tail_ok && total <= (uint64_t)Q * check_every, tail_ran = false;
    while (true) {
        if (graph) {
            hipError_t le = hipGraphLaunch(exec, s);
            if (le != hipSuccess) {
                ksa::set_error(std::string("hipGraphLaunch failed: ") + hipGetErrorString(le));
                status = KS_ERR_HIP;
                break;
            }
            iterations += 2;
        } else {
            status = iteration(cur);
            if (status) break;
            hipError_t le = hipGetLastError();
            if (le != hipSuccess) {
                ksa::set_error(std::string("wavefront launch failed: ") + hipGetErrorString(le));
                status = KS_ERR_HIP;
                break;
            }
            cur = 1 - cur;
            ++iterations;
            iter_index = iterations;
        }
        // the control block is also read after iterations 1, 2 and 4, so that the launch bounds tighten early
        const bool early_read = !graph && (iterations == 1 || iterations == 2 || iterations == 4);
        if (iterations % check_every == 0 || check_each || early_read) {
            hipError_t me = hipMemcpyAsync(dev->host_ctrl, ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, s);
            if (me == hipSuccess) me = hipStreamSynchronize(s);
            if (me != hipSuccess) {
                ksa::set_error(std::string("wavefront control read failed: ") + hipGetErrorString(me));
                status = KS_ERR_HIP;
                break;
            }
            const uint32_t n_left = dev->host_ctrl->n_active;
            if (n_left == 0 && dev->host_ctrl->work_next >= dev->host_ctrl->work_end) break;
            if (last_gen < 0 && dev->host_ctrl->work_next >= dev->host_ctrl->work_end) last_gen = iterations - 1;
            if (!graph) {
                const uint64_t todo = dev->host_ctrl->work_end - dev->host_ctrl->work_next;
                gen_blocks = (int)std::max<uint64_t>(1, std::min<uint64_t>(tblocks, (todo + 255) / 256));
                shade_blocks = (int)std::max<uint64_t>(1, std::min<uint64_t>(tblocks, (n_left + todo + 255) / 256));
            }
            if (tail_ok && dev->host_ctrl->work_end - dev->host_ctrl->work_next <= (uint64_t)Q * check_every)
                check_each = true;
            if (tail_ok && dev->host_ctrl->work_next >= dev->host_ctrl->work_end) {
                if (n_left <= tail_limit) {
                    status = flush_shadow(cur); // the last shading pass's shadow rays add to these paths
                    if (status) break;
                    hipEvent_t e = tm.begin(KS_K_TAIL);
                    launch_tail(sc->ds, wo.batch, dim3((n_left + 255) / 256), s, rc, pb[cur],
                                (const uint32_t *)&ctrl->n_active, ws.samp.as<float4>(), ws.ovf.as<uint32_t>(),
                                ws.ovf_cap, stat_rows(dev));
                    tm.end(e);
                    KS_HIP(hipGetLastError());
                    tail_ran = true;
                    break;
                }
            }
            if (iterations > 1000000) {
                ksa::set_error("wavefront loop did not terminate");
                status = KS_ERR_INVALID;
                break;
            }
        }
    }
    if (status == KS_OK && !tail_ran) status = flush_shadow(graph ? 0 : cur);
    if (exec) (void)hipGraphExecDestroy(exec);
    if (g) (void)hipGraphDestroy(g);
    if (status) return status;
    if (stats) stats->iterations += iterations;
    return KS_OK;
}

int read_stats(ks_device *dev, ks_render_stats *stats) {
    if (!stats) return KS_OK;
    unsigned long long h[kStatCount];
    hipLaunchKernelGGL(k_stat_reduce, dim3(1), dim3(256), 0, dev->stream, dev->ws.stats.as<unsigned long long>(),
                       (int)stat_partial_rows(dev));
    KS_HIP(hipGetLastError());
    KS_HIP(hipMemcpyAsync(h, dev->ws.stats.as<unsigned long long>() + kStatCount, sizeof(h), hipMemcpyDeviceToHost,
                          dev->stream));
    KS_HIP(hipStreamSynchronize(dev->stream));
    stats->rays_closest += h[kRaysClosest];
    stats->rays_shadow += h[kRaysShadow];
    stats->node_visits_closest += h[kNodesClosest];
    stats->tri_tests_closest += h[kTrisClosest];
    stats->node_visits_shadow += h[kNodesShadow];
    stats->tri_tests_shadow += h[kTrisShadow];
    stats->nonfinite += h[kNonFinite];
    stats->stack_overflows += h[kSpills];
    stats->rays_tail += h[kRaysTail];
    return KS_OK;
}

uint32_t pick_queue(const ks_small_pt_params *p, uint64_t total) {
    // 2^26 in-flight paths (about 14 GB of path state): each persistent traversal launch then drains
    // once per 2^26 rays rather than per 2^23, and a 1080p 64-spp render takes 16 iterations instead of
    // 32. Measured on the 1080p 64-spp sphere scene: 43.6 ms per render at 2^23, 42.0 at 2^24, 41.2 at
    // 2^25, 40.5 at 2^26 (2^21 was 1.34x slower than 2^23).
    uint64_t q = p->queue_size > 0 ? (uint64_t)p->queue_size : (1ull << 26);
    q = std::min<uint64_t>(q, std::max<uint64_t>(total, 256));
    q = std::max<uint64_t>(q, 256);
    return (uint32_t)std::min<uint64_t>(q, 1ull << 26);
}

template <typename T> struct HostToDev {
    DevBuf b;
    hipError_t put(const T *src, size_t n, hipStream_t s) {
        hipError_t e = b.alloc(sizeof(T) * std::max<size_t>(n, 1));
        if (e != hipSuccess || n == 0) return e;
        return hipMemcpyAsync(b.p, src, sizeof(T) * n, hipMemcpyHostToDevice, s);
    }
    T *p() const { return b.as<T>(); }
};

} // namespace

namespace ksg {
hipStream_t device_stream(ks_device *dev) { return dev->stream; }
int device_ordinal(ks_device *dev) { return dev->ordinal; }
int device_cu_count(ks_device *dev) { return dev->n_cu; }
} // namespace ksg

// ---------------------------------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------------------------------
extern "C" {

const char *ks_last_error(void) { return ksa::g_error.c_str(); }

int ks_device_count(int32_t *count) {
    return ksa::guarded([&]() -> int {
        KS_REQUIRE(count, "count is null");
        int n = 0;
        hipError_t e = hipGetDeviceCount(&n);
        if (e != hipSuccess) n = 0;
        *count = n;
        return KS_OK;
    });
}

int ks_device_open(int32_t ordinal, ks_device **out) {
    return ksa::guarded([&]() -> int {
        KS_REQUIRE(out, "out is null");
        *out = nullptr;
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
            ksa::set_error("no HIP device available");
            return KS_ERR_NO_DEVICE;
        }
        KS_REQUIRE(ordinal >= 0 && ordinal < n, "device ordinal out of range");
        KS_HIP(hipSetDevice(ordinal));
        auto d = std::make_unique<ks_device>();
        d->ordinal = ordinal;
        KS_HIP(hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking));
        hipDeviceProp_t prop;
        KS_HIP(hipGetDeviceProperties(&prop, ordinal));
        d->n_cu = prop.multiProcessorCount;
        KS_HIP(hipHostMalloc((void **)&d->host_ctrl, sizeof(Ctrl), hipHostMallocDefault));
        KS_HIP(hipHostMalloc((void **)&d->host_scratch, 64, hipHostMallocDefault));
        // the allocation-free relMSE's scratch: its address on this device (the current one)
        void *sp = nullptr;
        KS_HIP(hipGetSymbolAddress(&sp, HIP_SYMBOL(scratch)));
        d->relmse_scratch = static_cast<double *>(sp);
        KS_HIP(d->perm_tab.alloc((size_t)kPermTabDims * kPermTabPrefixes));
        hipLaunchKernelGGL(k_perm_tab, dim3(kPermTabDims * kPermTabPrefixes / 256), dim3(256), 0, d->stream,
                           d->perm_tab.as<uint8_t>());
        KS_HIP(hipGetLastError());
        KS_HIP(hipStreamSynchronize(d->stream));
        *out = d.release();
        return KS_OK;
    });
}

int ks_device_close(ks_device *dev) {
    return ksa::guarded([&]() -> int {
        if (!dev) return KS_OK;
        (void)hipSetDevice(dev->ordinal);
        if (dev->stream) (void)hipStreamSynchronize(dev->stream);
        if (dev->host_ctrl) (void)hipHostFree(dev->host_ctrl);
        if (dev->host_scratch) (void)hipHostFree(dev->host_scratch);
        for (hipEvent_t e : dev->events) (void)hipEventDestroy(e);
        if (dev->stream) (void)hipStreamDestroy(dev->stream);
        delete dev;
        return KS_OK;
    });
}

int ks_device_set_timing(ks_device *dev, int32_t enable) {
    return ksa::guarded([&]() -> int {
        KS_REQUIRE(dev, "device is null");
        dev->timing = enable ? 1 : 0;
        return KS_OK;
    });
}

int ks_device_set_tail(ks_device *dev, int32_t max_paths) {
    return ksa::guarded([&]() -> int {
        KS_REQUIRE(dev, "device is null");
        KS_REQUIRE(max_paths >= 0, "max_paths must not be negative");
        dev->tail_paths = (uint32_t)max_paths;
        return KS_OK;
    });
}

int ks_device_set_shading(ks_device *dev, int32_t sorted) {
    return ksa::guarded([&]() -> int {
        KS_REQUIRE(dev, "device is null");
        dev->sorted_shading = sorted ? 1 : 0;
        return KS_OK;
    });
}

} // extern "C"

namespace {

// Textures, textured materials and texture coordinates of a scene, built on the host. metas point into
// `pool` on the host (for host-side field evaluation); the device copies point into the uploaded pool.
struct HostTextures {
    TexLuts luts{};
    std::vector<DevTexture> metas;
    std::vector<uint8_t> pool;
    std::vector<DevTexMaterial> mats;
    std::vector<int32_t> mat_of; // per material: index into mats, or -1
    std::vector<float2> uvs;
    bool textured = false, alpha = false;
};

int build_textures(const ks_scene_desc *d, HostTextures &ht) {
    ht.textured = d->texcoords != nullptr || d->n_textures > 0 || d->material_fields != nullptr;
    if (!ht.textured) return KS_OK;
    KS_REQUIRE(d->n_textures == 0 || d->textures, "textures missing");
    ksa::fill_tex_luts(ht.luts);
    std::vector<size_t> offs;
    try {
        for (uint32_t i = 0; i < d->n_textures; ++i) {
            ksa::SourceTexture src = ksa::load_source_texture(d->textures[i], ht.luts.srgb_u8);
            DevTexture meta;
            std::vector<uint8_t> blob = ksa::pack_texture(src, meta);
            offs.push_back(ht.pool.size());
            ht.pool.insert(ht.pool.end(), blob.begin(), blob.end());
            ht.metas.push_back(meta);
        }
        ht.mat_of.assign(d->n_materials, -1);
        if (d->material_fields) {
            for (uint32_t m = 0; m < d->n_materials; ++m) {
                const ks_material_fields *f = d->material_fields[m];
                if (!f) continue;
                DevTexMaterial t{};
                const ks_field_desc *fields[9] = {&f->basecolor, &f->roughness, &f->metallic, &f->ior,
                                                  &f->specular_r0_mul, &f->specular_trans, &f->diffuse_trans,
                                                  &f->diffuse_trans_fwd, &f->emission};
                static const char *names[9] = {"basecolor", "roughness", "metallic", "ior", "specular_r0_mul",
                                               "specular_trans", "diffuse_trans", "diffuse_trans_fwd", "emission"};
                for (int k = 0; k < 9; ++k)
                    t.f[k] = ksa::to_dev_field(*fields[k], ht.metas, (k == 0 || k == 8) ? 3 : 1, names[k]);
                t.has_emission = f->has_emission ? 1 : 0;
                t.has_normal = f->has_normal_map ? 1 : 0;
                if (t.has_normal) {
                    t.normal = ksa::to_dev_field(f->normal_map, ht.metas, 3, "normal_map");
                    KS_REQUIRE(f->normal_space == KS_NORMAL_MAP_TANGENT || f->normal_space == KS_NORMAL_MAP_WORLD,
                               "unknown normal map space");
                    t.normal_world = f->normal_space == KS_NORMAL_MAP_WORLD;
                    t.normal_range01 = f->normal_range01 ? 1 : 0;
                    t.normal_strength = f->normal_strength;
                    // inverse transpose of the map's to_world (Transform::normal, maths.h:987)
                    const float *a = f->normal_to_world;
                    const double det = (double)a[0] * ((double)a[4] * a[8] - (double)a[5] * a[7]) -
                                       (double)a[1] * ((double)a[3] * a[8] - (double)a[5] * a[6]) +
                                       (double)a[2] * ((double)a[3] * a[7] - (double)a[4] * a[6]);
                    KS_REQUIRE(!t.normal_world || det != 0.0, "normal map to_world is singular");
                    if (t.normal_world) {
                        const double cof[9] = {(double)a[4] * a[8] - (double)a[5] * a[7],
                                               (double)a[5] * a[6] - (double)a[3] * a[8],
                                               (double)a[3] * a[7] - (double)a[4] * a[6],
                                               (double)a[2] * a[7] - (double)a[1] * a[8],
                                               (double)a[0] * a[8] - (double)a[2] * a[6],
                                               (double)a[1] * a[6] - (double)a[0] * a[7],
                                               (double)a[1] * a[5] - (double)a[2] * a[4],
                                               (double)a[2] * a[3] - (double)a[0] * a[5],
                                               (double)a[0] * a[4] - (double)a[1] * a[3]};
                        for (int k = 0; k < 9; ++k) t.normal_it[k] = (float)(cof[k] / det); // (A^-1)^T = cof / det
                    }
                }
                t.has_opacity = f->has_opacity_map ? 1 : 0;
                if (t.has_opacity) {
                    t.opacity = ksa::to_dev_field(f->opacity, ht.metas, 1, "opacity");
                    ht.alpha = true;
                }
                ht.mat_of[m] = (int32_t)ht.mats.size();
                ht.mats.push_back(t);
            }
        }
    } catch (const std::exception &ex) {
        ksa::set_error(ex.what());
        return KS_ERR_INVALID;
    }
    for (size_t i = 0; i < ht.metas.size(); ++i) ht.metas[i].data = ht.pool.data() + offs[i];
    if (d->texcoords) {
        ht.uvs.resize(d->n_vertices);
        for (uint32_t v = 0; v < d->n_vertices; ++v) ht.uvs[v] = make_float2(d->texcoords[2 * v], d->texcoords[2 * v + 1]);
    }
    return KS_OK;
}

// The closure constants of a textured material for host code that reads ks_material_desc (constant
// field values; textured fields read as their constant value, which the kernels never use).
ks_material_desc material_constants(const DevTexMaterial &t) {
    ks_material_desc m{};
    for (int k = 0; k < 3; ++k) m.basecolor[k] = t.f[kFieldBasecolor].value[k];
    m.roughness = t.f[kFieldRoughness].value[0];
    m.metallic = t.f[kFieldMetallic].value[0];
    m.ior = t.f[kFieldIor].value[0];
    m.specular_r0_mul = t.f[kFieldR0Mul].value[0];
    m.specular_trans = t.f[kFieldSpecTrans].value[0];
    m.diffuse_trans = t.f[kFieldDiffTrans].value[0];
    m.diffuse_trans_fwd = t.f[kFieldDiffTransFwd].value[0];
    for (int k = 0; k < 3; ++k) m.emission[k] = t.f[kFieldEmission].value[k];
    m.has_emission = t.has_emission;
    return m;
}

// SummedAreaTable (sat.h:9-61): FP64 prefix sums with a one-texel offset and clamped bilinear lookups.
struct Sat {
    std::vector<double> table;
    int nx = 0, ny = 0;
    Sat(const std::vector<float> &v, int nx_, int ny_) : table(v.size()), nx(nx_), ny(ny_) {
        table[0] = v[0];
        for (int x = 1; x < nx; ++x) table[x] = v[x] + table[x - 1];
        for (int y = 1; y < ny; ++y) table[(size_t)y * nx] = v[(size_t)y * nx] + table[(size_t)(y - 1) * nx];
        for (int y = 1; y < ny; ++y)
            for (int x = 1; x < nx; ++x)
                table[(size_t)y * nx + x] = v[(size_t)y * nx + x] + table[(size_t)y * nx + (x - 1)] +
                                            table[(size_t)(y - 1) * nx + x] - table[(size_t)(y - 1) * nx + (x - 1)];
    }
    double fetch(int x, int y) const {
        if (x == 0 || y == 0) return 0.0;
        x = std::min(x - 1, nx - 1);
        y = std::min(y - 1, ny - 1);
        return table[(size_t)y * nx + x];
    }
    static void boundary(float u, int res, int &i0, int &i1, float &t) { // lerp_helper, Boundary tick, Clamp
        const float us = u * (float)(res - 1);
        i0 = (int)std::floor(us);
        i1 = i0 + 1;
        t = us - (float)i0;
        i0 = std::min(std::max(i0, 0), res - 1);
        i1 = std::min(std::max(i1, 0), res - 1);
    }
    double lookup(float px, float py) const {
        int x0, x1, y0, y1;
        float wx, wy;
        boundary(px, nx + 1, x0, x1, wx);
        boundary(py, ny + 1, y0, y1, wy);
        return fetch(x0, y0) * (1.0 - wx) * (1.0 - wy) + fetch(x1, y0) * (wx) * (1.0 - wy) +
               fetch(x0, y1) * (1.0 - wx) * (wy) + fetch(x1, y1) * (wx) * (wy);
    }
    float sum(float x0, float y0, float x1, float y1) const {
        const double s = (lookup(x1, y1) - lookup(x0, y1)) + (lookup(x0, y0) - lookup(x1, y0));
        return std::max((float)s, 0.0f);
    }
};

struct Box2 { // AABB2 (aabb.h:11-135): empty as (+inf, -inf), zero extents when empty
    float lo[2] = {INFINITY, INFINITY}, hi[2] = {-INFINITY, -INFINITY};
    void expand(float x, float y) {
        lo[0] = std::min(lo[0], x), lo[1] = std::min(lo[1], y);
        hi[0] = std::max(hi[0], x), hi[1] = std::max(hi[1], y);
    }
    bool empty() const { return lo[0] > hi[0] || lo[1] > hi[1]; }
    float ext(int k) const { return empty() ? 0.0f : hi[k] - lo[k]; }
    float area() const { return ext(0) * ext(1); }
};

float wrap_mod1(float a) { // mod(a, 1.0f) (maths.h:133-138)
    const float r = a - std::floor(a / 1.0f) * 1.0f;
    return r < 0 ? r + 1.0f : r;
}

// compute_mesh_light_importance (light.cpp:307-523) for the triangles [t0, t1) of one emissive geometry:
// uniform fields give 2 pi area times the (constant) luminance x opacity; textured emission / opacity
// average their texel importance under each triangle's uv bounds through a summed-area table.
void mesh_light_importance(const ks_scene_desc *d, uint32_t t0, uint32_t t1, const ks_material_desc &m,
                           const DevTexMaterial *tm, const TexContext &hctx, std::vector<float> &imp,
                           std::vector<float> &area) {
    const float two_pi = 6.28318530718f;
    const uint32_t N = t1 - t0;
    imp.assign(N, 0.0f);
    area.assign(N, 0.0f);
    const bool em_tex = tm && tm->f[kFieldEmission].tex >= 0;
    const bool has_op = tm && tm->has_opacity;
    const bool op_tex = has_op && tm->opacity.tex >= 0;
    bool fallback = !em_tex && !op_tex;
    int res[2] = {1, 1};
    bool set = false;
    int wu = 0, wv = 0, flip = 0;
    float scl[2] = {1, 1}, off[2] = {0, 0};
    auto take = [&](const DevField &f) {
        const DevTexture &t = hctx.tex[f.tex];
        res[0] = std::max(res[0], t.width);
        res[1] = std::max(res[1], t.height);
        if (!set) {
            wu = f.wrap_u, wv = f.wrap_v, flip = f.flip_v;
            scl[0] = f.uv_scale[0], scl[1] = f.uv_scale[1], off[0] = f.uv_offset[0], off[1] = f.uv_offset[1];
            set = true;
        } else if (wu != f.wrap_u || wv != f.wrap_v || flip != f.flip_v || scl[0] != f.uv_scale[0] ||
                   scl[1] != f.uv_scale[1] || off[0] != f.uv_offset[0] || off[1] != f.uv_offset[1]) {
            fallback = true; // textured fields with different settings
        }
    };
    if (em_tex) take(tm->f[kFieldEmission]);
    if (op_tex) take(tm->opacity);
    std::vector<float> iuv((size_t)res[0] * res[1]);
    float sum = 0.0f;
    const float zero[4] = {0, 0, 0, 0};
    for (int y = 0; y < res[1]; ++y)
        for (int x = 0; x < res[0]; ++x) {
            const float u = ((float)x + 0.5f) / (float)res[0], v = ((float)y + 0.5f) / (float)res[1];
            float e[4] = {m.emission[0], m.emission[1], m.emission[2], 0.0f};
            if (tm) field_eval0(hctx, tm->f[kFieldEmission], u, v, 3, e);
            float importance = 0.212671f * e[0] + 0.715160f * e[1] + 0.072169f * e[2];
            if (has_op) {
                float o[4] = {0, 0, 0, 0};
                field_eval0(hctx, tm->opacity, u, v, 1, o);
                importance *= o[0];
            }
            iuv[(size_t)y * res[0] + x] = importance;
            sum += importance;
        }
    const float avg = sum / (float)iuv.size();
    auto tri_area = [&](uint32_t t) {
        const float *a = d->vertices + 3 * d->indices[3 * t];
        const float *b = d->vertices + 3 * d->indices[3 * t + 1];
        const float *c = d->vertices + 3 * d->indices[3 * t + 2];
        float e01[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
        float e12[3] = {c[0] - b[0], c[1] - b[1], c[2] - b[2]};
        float cx = e01[1] * e12[2] - e01[2] * e12[1], cy = e01[2] * e12[0] - e01[0] * e12[2],
              cz = e01[0] * e12[1] - e01[1] * e12[0];
        return 0.5f * std::sqrt(cx * cx + cy * cy + cz * cz);
    };
    if (fallback) {
        for (uint32_t i = 0; i < N; ++i) {
            area[i] = tri_area(t0 + i);
            imp[i] = two_pi * area[i] * avg;
        }
        return;
    }
    const Sat sat(iuv, res[0], res[1]);
    const float before_one = 0x1.fffffep-1f;
    const bool has_uv = d->texcoords && !(d->geom_flags && (d->geom_flags[d->tri_geom[t0]] & KS_GEOM_NO_TEXCOORD));
    for (uint32_t i = 0; i < N; ++i) {
        const uint32_t t = t0 + i;
        const float pa = tri_area(t);
        Box2 bound;
        for (int k = 0; k < 3; ++k) {
            float tc[2] = {0.0f, 0.0f};
            if (has_uv) {
                tc[0] = d->texcoords[2 * d->indices[3 * t + k]];
                tc[1] = d->texcoords[2 * d->indices[3 * t + k] + 1];
            }
            if (flip) tc[1] = 1.0f - tc[1];
            bound.expand(tc[0] * scl[0] + off[0], tc[1] * scl[1] + off[1]);
        }
        const int txmin = (int)std::floor(bound.lo[0]), txmax = (int)std::ceil(bound.hi[0]);
        const int tymin = (int)std::floor(bound.lo[1]), tymax = (int)std::ceil(bound.hi[1]);
        float acc = 0.0f, sum_b = 0.0f;
        for (int ty = tymin; ty < tymax; ++ty)
            for (int tx = txmin; tx < txmax; ++tx) {
                Box2 b; // intersect(bound, tile)
                b.lo[0] = std::max(bound.lo[0], (float)tx), b.lo[1] = std::max(bound.lo[1], (float)ty);
                b.hi[0] = std::min(bound.hi[0], (float)tx + before_one);
                b.hi[1] = std::min(bound.hi[1], (float)ty + before_one);
                if (b.empty()) b = Box2();
                const float b_area = b.area();
                if (b_area == 0.0f) continue;
                for (int k = 0; k < 2; ++k) {
                    const int mode = k == 0 ? wu : wv;
                    if (mode == KS_WRAP_REPEAT) {
                        b.lo[k] = wrap_mod1(b.lo[k]);
                        b.hi[k] = wrap_mod1(b.hi[k]);
                    } else {
                        b.lo[k] = std::min(std::max(b.lo[k], 0.0f), 1.0f);
                        b.hi[k] = std::min(std::max(b.hi[k], 0.0f), 1.0f);
                    }
                    if (b.lo[k] == 0.0f && b.hi[k] == 0.0f) b.hi[k] = 1.0f / (float)(res[k] + 1);
                    if (b.lo[k] == 1.0f && b.hi[k] == 1.0f) b.lo[k] = 1.0f - 1.0f / (float)(res[k] + 1);
                }
                const float mean = sat.sum(b.lo[0], b.lo[1], b.hi[0], b.hi[1]) /
                                   ((b.ext(0) * (float)res[0]) * (b.ext(1) * (float)res[1]));
                acc += mean * b_area;
                sum_b += b_area;
            }
        if (sum_b > 0.0f) acc /= sum_b;
        acc *= two_pi * pa;
        imp[i] = acc;
        area[i] = pa;
    }
}

// DistribTable (distrib.cpp:9-29): running sum of f / n, normalised, last entry forced to 1; returns acc.
float build_cdf(const float *f, uint32_t n, float *cdf) {
    cdf[0] = 0.0f;
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
    return acc;
}

// SkyLight(path, ...) (light.cpp:19-49): texels, and the 2D distribution of the luminance weighted by
// sin(theta) minus its mean, clamped at 0 (the thresholded table of the reference). Also the light's
// power (SkyLight::power, light.cpp:136-158).
struct HostSky {
    std::vector<float4> rgb;
    std::vector<float> cond, marg;
    SkyMap map{};
    float power[3] = {0, 0, 0};
};

int build_sky(const ks_sky_map_desc &d, float scene_radius, HostSky &out) {
    KS_REQUIRE(d.rgb && d.width > 0 && d.height > 0, "sky map without texels");
    KS_REQUIRE((int64_t)d.width * d.height < (1ll << 31), "sky map too large");
    const uint32_t w = (uint32_t)d.width, h = (uint32_t)d.height;
    const float pi = 3.14159265359f, two_pi = 6.28318530718f;
    out.rgb.resize((size_t)w * h);
    std::vector<float> lum((size_t)w * h);
    float lum_sum = 0.0f;
    for (uint32_t i = 0; i < w * h; ++i) {
        const float *c = d.rgb + 3 * (size_t)i;
        out.rgb[i] = make_float4(c[0], c[1], c[2], 0.0f);
        uint32_t y = i / w;
        float theta = ((float)y + 0.5f) / (float)h * pi;
        float sin_theta = std::sin(theta);
        lum[i] = sin_theta * (0.212671f * c[0] + 0.715160f * c[1] + 0.072169f * c[2]); // luminance, maths.h:675
        lum_sum += lum[i];
    }
    float average = lum_sum / (float)lum.size();
    for (float &l : lum) l = std::max(l - average, 0.0f);
    out.cond.resize((size_t)h * (w + 1));
    out.marg.resize(h + 1);
    std::vector<float> accs(h);
    for (uint32_t y = 0; y < h; ++y) accs[y] = build_cdf(lum.data() + (size_t)y * w, w, out.cond.data() + (size_t)y * (w + 1));
    build_cdf(accs.data(), h, out.marg.data());
    SkyMap &m = out.map;
    m.w = (int32_t)w;
    m.h = (int32_t)h;
    m.yup = d.transform_y_up ? 1 : 0;
    m.strength = d.strength;
    double a[9], det;
    for (int k = 0; k < 9; ++k) a[k] = d.to_world[k];
    det = a[0] * (a[4] * a[8] - a[5] * a[7]) - a[1] * (a[3] * a[8] - a[5] * a[6]) + a[2] * (a[3] * a[7] - a[4] * a[6]);
    KS_REQUIRE(det != 0.0, "sky map to_world is singular");
    const double inv[9] = {(a[4] * a[8] - a[5] * a[7]) / det, (a[2] * a[7] - a[1] * a[8]) / det,
                           (a[1] * a[5] - a[2] * a[4]) / det, (a[5] * a[6] - a[3] * a[8]) / det,
                           (a[0] * a[8] - a[2] * a[6]) / det, (a[2] * a[3] - a[0] * a[5]) / det,
                           (a[3] * a[7] - a[4] * a[6]) / det, (a[1] * a[6] - a[0] * a[7]) / det,
                           (a[0] * a[4] - a[1] * a[3]) / det};
    for (int k = 0; k < 9; ++k) {
        m.l2w[k] = d.to_world[k];
        m.w2l[k] = (float)inv[k];
    }
    // power: the map integrated over a z-up grid of texel centres (light.cpp:136-158)
    m.rgb = out.rgb.data();
    float sum[3] = {0, 0, 0};
    for (uint32_t row = 0; row < h; ++row) {
        float theta = ((float)row + 0.5f) / (float)h * pi;
        float sin_theta = std::sin(theta);
        float row_sum[3] = {0, 0, 0};
        for (uint32_t col = 0; col < w; ++col) {
            float phi = ((float)col + 0.5f) / (float)w * two_pi;
            float ct = std::cos(theta), st = std::sin(theta), cp = std::cos(phi), sp = std::sin(phi);
            float L[3];
            sky_eval(m, cp * st, sp * st, ct, L);
            for (int k = 0; k < 3; ++k) row_sum[k] += L[k] * sin_theta;
        }
        for (int k = 0; k < 3; ++k) sum[k] += row_sum[k];
    }
    for (int k = 0; k < 3; ++k)
        out.power[k] = ((sum[k] * ((pi / (float)h) * (two_pi / (float)w))) * (pi * (scene_radius * scene_radius)));
    m.rgb = nullptr; // device pointers are set after upload
    return KS_OK;
}

} // namespace

extern "C" {

int ks_scene_create(ks_device *dev, const ks_scene_desc *d, const ks_bvh_options *opt, ks_scene **out) {
    return ksa::guarded([&]() -> int {
        KS_REQUIRE(dev && d && out, "null argument");
        *out = nullptr;
        KS_REQUIRE(d->n_tris == 0 || (d->vertices && d->indices && d->tri_geom), "scene arrays missing");
        KS_REQUIRE(d->n_materials > 0 && d->materials, "scene needs at least one material");
        KS_REQUIRE(d->n_materials < (1u << 24), "at most 2^24 - 1 materials");
        KS_REQUIRE(d->n_lights <= 64, "at most 64 non-mesh lights");
        for (uint32_t t = 0; t < d->n_tris; ++t) {
            for (int k = 0; k < 3; ++k)
                KS_REQUIRE(d->indices[3 * t + k] < d->n_vertices, "triangle index out of range");
            KS_REQUIRE(d->tri_geom[t] < d->n_geoms, "triangle geometry id out of range");
            KS_REQUIRE(t == 0 || d->tri_geom[t] >= d->tri_geom[t - 1], "triangles must be grouped by geometry");
        }
        for (uint32_t g = 0; g < d->n_geoms; ++g)
            KS_REQUIRE(d->geom_material[g] < d->n_materials, "geometry material id out of range");
        KS_HIP(hipSetDevice(dev->ordinal));
        auto t0 = std::chrono::steady_clock::now();
        auto sc = std::make_unique<ks_scene>();
        sc->dev = dev;
        ks_bvh_options o{};
        if (opt) o = *opt;
        KS_REQUIRE(o.builder == KS_BVH_BUILD_HOST_SBVH || o.builder == KS_BVH_BUILD_DEVICE, "unknown BVH builder");
        const int width = o.branching == 2 ? 2 : 4;
        // an empty scene gets the host builder's single empty node whichever builder was asked for
        const bool device_build = o.builder != KS_BVH_BUILD_HOST_SBVH && d->n_tris > 0;
        ksa::BvhResult bvh;
        double host_bvh_ms = 0.0;
        if (!device_build) {
            try {
                const auto tb = std::chrono::steady_clock::now();
                bvh = ksa::build_sbvh(d->vertices, d->indices, d->n_tris, o);
                host_bvh_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - tb).count();
            } catch (const std::exception &ex) {
                ksa::set_error(std::string("BVH build failed: ") + ex.what());
                return KS_ERR_INVALID;
            }
        }

        HostTextures ht;
        if (int te = build_textures(d, ht)) return te;
        // the material constants each triangle sees (textured materials: their constant field values)
        auto material_of = [&](uint32_t m) -> ks_material_desc {
            if (!ht.mat_of.empty() && ht.mat_of[m] >= 0) return material_constants(ht.mats[ht.mat_of[m]]);
            return d->materials[m];
        };
        const TexContext hctx{ht.metas.data(), &ht.luts};

        std::vector<float4> verts(d->n_vertices);
        for (uint32_t i = 0; i < d->n_vertices; ++i)
            verts[i] = make_float4(d->vertices[3 * i], d->vertices[3 * i + 1], d->vertices[3 * i + 2], 0.0f);
        std::vector<uint4> tri_idx(d->n_tris);
        for (uint32_t t = 0; t < d->n_tris; ++t)
            tri_idx[t] = make_uint4(d->indices[3 * t], d->indices[3 * t + 1], d->indices[3 * t + 2],
                                    d->geom_material[d->tri_geom[t]] |
                                        ((uint32_t)(d->geom_flags ? d->geom_flags[d->tri_geom[t]] & 7u : 0u) << 24));
        std::vector<Material> mats(d->n_materials);
        for (uint32_t m = 0; m < d->n_materials; ++m) mats[m] = to_device_material(material_of(m));
        // material classes of sorted shading, from the clamped closure: metallic 1 weights only the metal
        // lobe (every other lobe carries a factor 1 - metallic); metallic, specular_trans and
        // diffuse_trans 0 leave the diffuse lobe and the dielectric reflection; textured materials are general
        std::vector<uint8_t> mat_class(d->n_materials);
        for (uint32_t m = 0; m < d->n_materials; ++m) {
            const Material &dm = mats[m];
            const bool textured = !ht.mat_of.empty() && ht.mat_of[m] >= 0;
            mat_class[m] = textured ? kShadeGeneral
                           : dm.metallic == 1.0f ? kShadeMetal
                           : dm.metallic == 0.0f && dm.specular_trans == 0.0f && dm.diffuse_trans == 0.0f
                               ? kShadeDielectric
                               : kShadeGeneral;
        }

        hipStream_t s = dev->stream;
        float scene_lo[3], scene_hi[3];
        ksa::DeviceBvh dbvh;
        if (device_build) {
            KS_HIP(upload(sc->verts, verts, s));
            KS_HIP(upload(sc->tri_idx, tri_idx, s));
            DevBuf alpha_of;
            if (ht.alpha) {
                std::vector<uint8_t> flags(d->n_materials, 0);
                for (uint32_t m = 0; m < d->n_materials; ++m)
                    flags[m] = ht.mat_of[m] >= 0 && ht.mats[ht.mat_of[m]].has_opacity ? 1 : 0;
                KS_HIP(upload(alpha_of, flags, s));
            }
            std::string err;
            if (ksa::build_bvh_device(sc->verts.p, sc->tri_idx.p, d->n_tris, ht.alpha ? alpha_of.as<uint8_t>() : nullptr,
                                      o, width, s, dbvh, err) != 0) {
                ksa::set_error(err);
                return KS_ERR_HIP;
            }
            // the scene takes the two arrays over
            sc->nodes.p = dbvh.nodes;
            sc->nodes.bytes = dbvh.node_bytes;
            sc->tris.p = dbvh.tris;
            sc->tris.bytes = dbvh.tri_bytes;
            for (int k = 0; k < 3; ++k) scene_lo[k] = dbvh.lo[k], scene_hi[k] = dbvh.hi[k];
        } else {
            for (int k = 0; k < 3; ++k) scene_lo[k] = bvh.lo[k], scene_hi[k] = bvh.hi[k];
        }

        // Light list in PowerLightSampler order (light.cpp:636-677): non-mesh lights, then per emissive
        // geometry the triangles with non-zero importance after swap compaction (light.cpp:524-558).
        float ext[3] = {scene_hi[0] - scene_lo[0], scene_hi[1] - scene_lo[1], scene_hi[2] - scene_lo[2]};
        float scene_radius = 0.5f * std::sqrt(ext[0] * ext[0] + ext[1] * ext[1] + ext[2] * ext[2]);
        const float pi = 3.14159265359f, two_pi = 6.28318530718f;
        std::vector<Light> lights;
        std::vector<float> powers;
        int n_sky = 0;
        int sky[4] = {0, 0, 0, 0};
        HostSky host_sky;
        bool have_map = false;
        for (uint32_t i = 0; i < d->n_lights; ++i) {
            const ks_light_desc &L = d->lights[i];
            KS_REQUIRE(L.type >= 0 && L.type <= 3, "unknown light type");
            Light l{};
            l.type = L.type == KS_LIGHT_SKY_MAP ? 4 : L.type;
            for (int k = 0; k < 3; ++k) {
                l.color[k] = L.color[k];
                l.vec[k] = L.vec[k];
            }
            float ph[3];
            if (L.type == KS_LIGHT_SKY_MAP) {
                KS_REQUIRE(d->sky_map, "a KS_LIGHT_SKY_MAP light needs ks_scene_desc.sky_map");
                KS_REQUIRE(!have_map, "at most one environment-map sky per scene");
                KS_REQUIRE(n_sky < 4, "at most 4 sky lights");
                int e = build_sky(*d->sky_map, scene_radius, host_sky);
                if (e) return e;
                have_map = true;
                sky[n_sky++] = (int)i;
                for (int k = 0; k < 3; ++k) ph[k] = host_sky.power[k];
            } else if (L.type == KS_LIGHT_SKY_AMBIENT) { // SkyLight::power (light.cpp:136-158), 1x1 map
                KS_REQUIRE(n_sky < 4, "at most 4 sky lights");
                sky[n_sky++] = (int)i;
                float sin_theta = std::sin(0.5f / 1.0f * pi);
                for (int k = 0; k < 3; ++k) ph[k] = ((l.color[k] * sin_theta) * ((pi / 1.0f) * (two_pi / 1.0f))) *
                                                    (pi * (scene_radius * scene_radius));
            } else if (L.type == KS_LIGHT_DIRECTIONAL) {
                float n = std::sqrt(l.vec[0] * l.vec[0] + l.vec[1] * l.vec[1] + l.vec[2] * l.vec[2]);
                KS_REQUIRE(n > 0.0f, "directional light without direction");
                for (int k = 0; k < 3; ++k) l.vec[k] /= n; // load_vec3("dir", true)
                for (int k = 0; k < 3; ++k) ph[k] = l.color[k] * (pi * (scene_radius * scene_radius));
            } else {
                for (int k = 0; k < 3; ++k) ph[k] = l.color[k] * (4.0f * pi);
            }
            lights.push_back(l);
            powers.push_back((ph[0] + ph[1] + ph[2]) / 3.0f);
            sc->light_prim.push_back(-1);
        }
        std::vector<int32_t> tri_light(d->n_tris, -1);
        for (uint32_t t0 = 0; t0 < d->n_tris;) {
            uint32_t t1 = t0;
            while (t1 < d->n_tris && d->tri_geom[t1] == d->tri_geom[t0]) ++t1;
            const uint32_t gm = d->geom_material[d->tri_geom[t0]];
            const ks_material_desc m = material_of(gm);
            if (m.has_emission) {
                const DevTexMaterial *tm = (!ht.mat_of.empty() && ht.mat_of[gm] >= 0) ? &ht.mats[ht.mat_of[gm]] : nullptr;
                const bool em_tex = tm && tm->f[kFieldEmission].tex >= 0;
                uint32_t N = t1 - t0;
                std::vector<uint32_t> ids(N);
                std::vector<float> imp, area;
                mesh_light_importance(d, t0, t1, m, tm, hctx, imp, area);
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
                    Light l{};
                    l.type = 3;
                    l.prim = t0 + ids[i];
                    l.area = area[i];
                    l.color[3] = em_tex ? 1.0f : 0.0f; // emission read from its texture at each use
                    for (int k = 0; k < 3; ++k) {
                        l.color[k] = m.emission[k];
                        l.vec[k] = d->vertices[3 * d->indices[3 * l.prim] + k];
                        l.v1[k] = d->vertices[3 * d->indices[3 * l.prim + 1] + k];
                        l.v2[k] = d->vertices[3 * d->indices[3 * l.prim + 2] + k];
                    }
                    tri_light[l.prim] = (int32_t)lights.size();
                    lights.push_back(l);
                    powers.push_back((imp[i] + imp[i] + imp[i]) / 3.0f);
                    sc->light_prim.push_back((int32_t)l.prim);
                }
            }
            t0 = t1;
        }
        std::vector<float> cdf;
        if (!powers.empty()) { // DistribTable (distrib.cpp:9-31)
            uint32_t n = (uint32_t)powers.size();
            cdf.resize(n + 1);
            cdf[0] = 0.0f;
            float invn = 1.0f / (float)n;
            for (uint32_t i = 1; i <= n; ++i) cdf[i] = cdf[i - 1] + powers[i - 1] * invn;
            float acc = cdf[n];
            if (acc == 0.0f) {
                for (uint32_t i = 0; i <= n; ++i) cdf[i] = (float)i / (float)n;
            } else {
                float inv = 1.0f / acc;
                for (uint32_t i = 0; i <= n; ++i) cdf[i] *= inv;
            }
            cdf[n] = 1.0f;
        }
        sc->light_power = powers;
        sc->light_cdf = cdf;

        // BVH4 visits address nodes with 32-bit byte offsets
        const size_t n_nodes = device_build ? dbvh.n_nodes : width == 4 ? bvh.nodes4.size() : bvh.nodes.size();
        KS_REQUIRE(width != 4 || n_nodes < (1u << 25), "BVH4 node array exceeds 4 GiB; build this scene with branching = 2");
        if (!device_build) {
            if (width == 4) KS_HIP(upload(sc->nodes, bvh.nodes4, s));
            else KS_HIP(upload(sc->nodes, bvh.nodes, s));
            if (ht.alpha) { // alpha-tested triangle records carry a non-zero v1.w for the traversal filter
                for (auto &g : bvh.tris) {
                    const uint32_t m = d->geom_material[d->tri_geom[g.tri]];
                    if (ht.mat_of[m] >= 0 && ht.mats[ht.mat_of[m]].has_opacity) {
                        const uint32_t one = 1u;
                        std::memcpy(&g.v1[3], &one, 4);
                    }
                }
            }
            KS_HIP(upload(sc->tris, bvh.tris, s));
            KS_HIP(upload(sc->verts, verts, s));
        }
        if (d->normals) {
            std::vector<float4> nrm(d->n_vertices);
            for (uint32_t i = 0; i < d->n_vertices; ++i)
                nrm[i] = make_float4(d->normals[3 * i], d->normals[3 * i + 1], d->normals[3 * i + 2], 0.0f);
            KS_HIP(upload(sc->normals, nrm, s));
        }
        if (!device_build) KS_HIP(upload(sc->tri_idx, tri_idx, s));
        KS_HIP(upload(sc->tri_light, tri_light, s));
        KS_HIP(upload(sc->mats, mats, s));
        KS_HIP(upload(sc->mat_class, mat_class, s));
        KS_HIP(upload(sc->lights, lights, s));
        KS_HIP(upload(sc->cdf, cdf, s));
        if (ht.textured) {
            KS_HIP(upload(sc->tex_pool, ht.pool, s));
            std::vector<DevTexture> metas = ht.metas;
            for (auto &t : metas) t.data = sc->tex_pool.as<uint8_t>() + (t.data - ht.pool.data());
            KS_HIP(upload(sc->tex_meta, metas, s));
            KS_HIP(upload(sc->tex_mats, ht.mats, s));
            KS_HIP(upload(sc->tex_mat_of, ht.mat_of, s));
            KS_HIP(upload(sc->uvs, ht.uvs, s));
            std::vector<TexLuts> luts(1, ht.luts);
            KS_HIP(upload(sc->tex_luts, luts, s));
        }
        if (have_map) {
            KS_HIP(upload(sc->sky_rgb, host_sky.rgb, s));
            KS_HIP(upload(sc->sky_cond, host_sky.cond, s));
            KS_HIP(upload(sc->sky_marg, host_sky.marg, s));
            host_sky.map.rgb = sc->sky_rgb.as<float4>();
            host_sky.map.cond = sc->sky_cond.as<float>();
            host_sky.map.marg = sc->sky_marg.as<float>();
        }

        DevScene &ds = sc->ds;
        ds.nodes = sc->nodes.as<float4>();
        ds.width = width;
        ds.tris = sc->tris.as<float4>();
        ds.verts = sc->verts.as<float4>();
        ds.normals = d->normals ? sc->normals.as<float4>() : nullptr;
        ds.tri_idx = sc->tri_idx.as<uint4>();
        ds.tri_light = sc->tri_light.as<int32_t>();
        ds.mats = sc->mats.as<Material>();
        ds.mat_class = sc->mat_class.as<uint8_t>();
        ds.lights = sc->lights.as<Light>();
        ds.cdf = sc->cdf.as<float>();
        ds.n_lights = (int32_t)lights.size();
        ds.n_sky = n_sky;
        for (int k = 0; k < 4; ++k) ds.sky[k] = sky[k];
        ds.sky_map = host_sky.map;
        ds.textured = ht.textured ? 1 : 0;
        ds.alpha = ht.alpha ? 1 : 0;
        if (ht.textured) {
            ds.uvs = ht.uvs.empty() ? nullptr : sc->uvs.as<float2>();
            ds.textures = sc->tex_meta.as<DevTexture>();
            ds.tex_luts = sc->tex_luts.as<TexLuts>();
            ds.tex_mats = ht.mats.empty() ? nullptr : sc->tex_mats.as<DevTexMaterial>();
            ds.tex_mat_of = ht.mats.empty() ? nullptr : sc->tex_mat_of.as<int32_t>();
        }

        ks_scene_info &in = sc->info;
        in.n_tris = d->n_tris;
        in.n_vertices = d->n_vertices;
        in.n_nodes = (uint32_t)n_nodes;
        in.branching = (uint32_t)width;
        in.n_refs = device_build ? dbvh.n_refs : (uint32_t)bvh.tris.size();
        in.n_lights = (uint32_t)lights.size();
        in.max_depth = device_build ? dbvh.max_depth : width == 4 ? bvh.max_depth4 : bvh.max_depth;
        in.builder = device_build ? o.builder : KS_BVH_BUILD_HOST_SBVH;
        in.bvh_build_ms = device_build ? dbvh.build_ms : host_bvh_ms;
        // traversal stack bound: every interior node on the path leaves at most width - 1 entries behind
        sc->max_stack = (uint32_t)(width - 1) * in.max_depth + 1;
        in.device_bytes = sc->nodes.bytes + sc->tris.bytes + sc->verts.bytes + sc->normals.bytes + sc->tri_idx.bytes +
                          sc->tri_light.bytes + sc->mats.bytes + sc->lights.bytes + sc->cdf.bytes + sc->tex_pool.bytes +
                          sc->tex_mats.bytes + sc->uvs.bytes;
        for (int k = 0; k < 3; ++k) {
            in.bound_lo[k] = scene_lo[k];
            in.bound_hi[k] = scene_hi[k];
        }
        in.build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        if (!device_build) {
            sc->host_nodes2 = std::move(bvh.nodes);
            sc->host_tris = std::move(bvh.tris);
        }
        *out = sc.release();
        return KS_OK;
    });
}

int ks_scene_free(ks_scene *scene) {
    return ksa::guarded([&]() -> int {
        if (scene) {
            (void)hipSetDevice(scene->dev->ordinal);
            (void)hipStreamSynchronize(scene->dev->stream);
            delete scene;
        }
        return KS_OK;
    });
}

int ks_scene_info_get(const ks_scene *scene, ks_scene_info *info) {
    return ksa::guarded([&]() -> int {
        KS_REQUIRE(scene && info, "null argument");
        *info = scene->info;
        return KS_OK;
    });
}

int ks_scene_bvh_get(const ks_scene *scene, void *nodes, void *refs) {
    return ksa::guarded([&]() -> int {
        KS_REQUIRE(scene && nodes && refs, "null argument");
        KS_HIP(hipSetDevice(scene->dev->ordinal));
        const size_t node_size = scene->info.branching == 4 ? sizeof(ksa::GpuNode4) : sizeof(ksa::GpuNode);
        KS_HIP(hipStreamSynchronize(scene->dev->stream));
        KS_HIP(hipMemcpy(nodes, scene->nodes.p, (size_t)scene->info.n_nodes * node_size, hipMemcpyDeviceToHost));
        KS_HIP(hipMemcpy(refs, scene->tris.p, (size_t)scene->info.n_refs * sizeof(ksa::GpuTri), hipMemcpyDeviceToHost));
        return KS_OK;
    });
}

int ks_scene_light_table(const ks_scene *scene, int32_t *n_lights, int32_t *prim, float *power, float *cdf) {
    return ksa::guarded([&]() -> int {
        KS_REQUIRE(scene, "scene is null");
        int32_t n = (int32_t)scene->light_power.size();
        if (n_lights) *n_lights = n;
        for (int32_t i = 0; i < n; ++i) {
            if (prim) prim[i] = scene->light_prim[i];
            if (power) power[i] = scene->light_power[i];
        }
        if (cdf && n > 0) std::memcpy(cdf, scene->light_cdf.data(), sizeof(float) * (n + 1));
        return KS_OK;
    });
}

int ks_film_create(ks_device *dev, int32_t width, int32_t height, const float backdrop[3], ks_film **out) {
    return ksa::guarded([&]() -> int {
        KS_REQUIRE(dev && out, "null argument");
        KS_REQUIRE(width > 0 && height > 0, "film size must be positive");
        *out = nullptr;
        KS_HIP(hipSetDevice(dev->ordinal));
        auto f = std::make_unique<ks_film>();
        f->dev = dev;
        f->w = width;
        f->h = height;
        if (backdrop)
            for (int k = 0; k < 3; ++k) f->backdrop[k] = backdrop[k];
        size_t n = (size_t)width * height;
        KS_HIP(f->rgb.alloc(sizeof(double) * 3 * n));
        KS_HIP(f->wts.alloc(sizeof(double) * 2 * n));
        KS_HIP(hipMemsetAsync(f->rgb.p, 0, f->rgb.bytes, dev->stream));
        KS_HIP(hipMemsetAsync(f->wts.p, 0, f->wts.bytes, dev->stream));
        KS_HIP(hipStreamSynchronize(dev->stream));
        *out = f.release();
        return KS_OK;
    });
}

int ks_film_clear(ks_film *film) {
    return ksa::guarded([&]() -> int {
        KS_REQUIRE(film, "film is null");
        KS_HIP(hipSetDevice(film->dev->ordinal));
        KS_HIP(hipMemsetAsync(film->rgb.p, 0, film->rgb.bytes, film->dev->stream));
        KS_HIP(hipMemsetAsync(film->wts.p, 0, film->wts.bytes, film->dev->stream));
        KS_HIP(hipStreamSynchronize(film->dev->stream));
        return KS_OK;
    });
}

int ks_film_download(const ks_film *film, double *rgb, double *weights) {
    return ksa::guarded([&]() -> int {
        KS_REQUIRE(film, "film is null");
        KS_HIP(hipSetDevice(film->dev->ordinal));
        if (rgb) KS_HIP(hipMemcpyAsync(rgb, film->rgb.p, film->rgb.bytes, hipMemcpyDeviceToHost, film->dev->stream));
        if (weights)
            KS_HIP(hipMemcpyAsync(weights, film->wts.p, film->wts.bytes, hipMemcpyDeviceToHost, film->dev->stream));
        KS_HIP(hipStreamSynchronize(film->dev->stream));
        return KS_OK;
    });
}

int ks_film_composite(const ks_film *film, float *out) {
    return ksa::guarded([&]() -> int {
        KS_REQUIRE(film && out, "null argument");
        size_t n = (size_t)film->w * film->h;
        std::vector<double> rgb(3 * n), w(2 * n);
        int e = ks_film_download(film, rgb.data(), w.data());
        if (e) return e;
        for (size_t i = 0; i < n; ++i) { // render_target.cpp:187-192; lerp takes a float t (maths.h:686-690)
            uint32_t w_sum = (uint32_t)(w[2 * i] + w[2 * i + 1]);
            float t = (float)(w_sum == 0 ? 0.0 : w[2 * i] / (double)w_sum);
            for (int k = 0; k < 3; ++k) {
                double fg = w[2 * i] == 0 ? 0.0 : rgb[3 * i + k] / w[2 * i];
                out[3 * i + k] = (float)((double)(1.0f - t) * (double)film->backdrop[k] + (double)t * fg);
            }
        }
        return KS_OK;
    });
}

int ks_film_free(ks_film *film) {
    return ksa::guarded([&]() -> int {
        if (film) {
            (void)hipSetDevice(film->dev->ordinal);
            (void)hipStreamSynchronize(film->dev->stream);
            delete film;
        }
        return KS_OK;
    });
}

int ks_small_pt_render(ks_scene *scene, const ks_camera_desc *cam, const ks_small_pt_params *params, ks_film *film,
                       ks_progress_fn progress, void *user, ks_render_stats *stats) {
    return ksa::guarded([&]() -> int {
        KS_REQUIRE(scene && cam && params && film, "null argument");
        KS_REQUIRE(film->w == params->render_width && film->h == params->render_height,
                   "film size differs from render_width x render_height");
        KS_REQUIRE(film->dev == scene->dev, "film and scene belong to different devices");
        KS_REQUIRE(params->guide_mode == KS_GUIDE_OFF, "guide_mode is not available in this build");
        ks_device *dev = scene->dev;
        KS_HIP(hipSetDevice(dev->ordinal));
        auto t0 = std::chrono::steady_clock::now();
        RenderConsts rc;
        int e = fill_consts(scene->dev, scene, cam, params, rc);
        if (e) return e;
        if (stats) std::memset(stats, 0, sizeof(*stats));
        int interval = params->spp_prog_interval > 0 ? params->spp_prog_interval : 32;
        uint64_t pixels = (uint64_t)rc.cw * rc.ch;
        // per-sample buffer budget: 1 GiB of float4 results per sub-batch
        const uint64_t max_samp = kSampleBufferBytes / sizeof(float4);
        int spp_chunk = (int)std::max<uint64_t>(1, std::min<uint64_t>((uint64_t)interval, max_samp / pixels));
        KS_REQUIRE(pixels <= max_samp, "crop window too large for one sub-batch");
        uint32_t Q = pick_queue(params, pixels * (uint64_t)std::min(spp_chunk, params->spp));
        e = ensure_workspace(dev, Q, pixels * (uint64_t)std::min(spp_chunk, params->spp), scene->max_stack);
        if (e) return e;
        clear_stats(dev, dev->stream);
        Timer tm{dev};
        for (int s_start = 0; s_start < params->spp; s_start += interval) {
            int s_end = std::min(params->spp, s_start + interval);
            for (int c0 = s_start; c0 < s_end; c0 += spp_chunk) {
                rc.s0 = c0;
                rc.nb = std::min(spp_chunk, s_end - c0);
                rc.work_list = nullptr;
                uint64_t total = pixels * (uint64_t)rc.nb;
                e = run_wavefront(scene, rc, total, Q, tm, stats);
                if (e) return e;
                hipEvent_t ev = tm.begin(KS_K_FILM);
                hipLaunchKernelGGL(k_film, dim3((unsigned)((pixels + 255) / 256)), dim3(256), 0, dev->stream, rc,
                                   dev->ws.samp.as<float4>(), film->rgb.as<double>(), film->wts.as<double>());
                tm.end(ev);
                KS_HIP(hipGetLastError());
            }
            KS_HIP(hipStreamSynchronize(dev->stream));
            if (progress) progress(film, s_end, user);
        }
        KS_HIP(hipStreamSynchronize(dev->stream));
        tm.collect(stats);
        e = read_stats(dev, stats);
        if (e) return e;
        if (stats) {
            stats->paths = pixels * (uint64_t)params->spp;
            stats->queue_size = (int32_t)Q;
            stats->wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        }
        return KS_OK;
    });
}

int ks_film_relmse(const ks_film *film, const ks_film *reference, double *relmse) {
    return ksa::guarded([&]() -> int {
        KS_REQUIRE(film && reference && relmse, "null argument");
        KS_REQUIRE(film->w == reference->w && film->h == reference->h, "films differ in size");
        KS_REQUIRE(film->dev == reference->dev, "films belong to different devices");
        ks_device *dev = film->dev;
        KS_HIP(hipSetDevice(dev->ordinal));
        size_t n = (size_t)film->w * film->h;
        unsigned blocks = (unsigned)((n + 255) / 256);
        DevBuf part;
        KS_HIP(part.alloc(sizeof(double) * blocks));
        hipLaunchKernelGGL(k_film_relmse, dim3(blocks), dim3(256), 0, dev->stream, n, film->rgb.as<double>(),
                           film->wts.as<double>(), make_float3(film->backdrop[0], film->backdrop[1], film->backdrop[2]),
                           reference->rgb.as<double>(), reference->wts.as<double>(),
                           make_float3(reference->backdrop[0], reference->backdrop[1], reference->backdrop[2]),
                           part.as<double>());
        KS_HIP(hipGetLastError());
        std::vector<double> h(blocks);
        KS_HIP(hipMemcpyAsync(h.data(), part.p, sizeof(double) * blocks, hipMemcpyDeviceToHost, dev->stream));
        KS_HIP(hipStreamSynchronize(dev->stream));
        double sum = 0.0;
        for (double v : h) sum += v;
        *relmse = sum / (3.0 * (double)n);
        return KS_OK;
    });
}

int ks_film_relmse_noalloc(const ks_film *film, const ks_film *reference, double *relmse) {
    return ksa::guarded([&]() -> int {
        KS_REQUIRE(film && reference && relmse, "null argument");
        KS_REQUIRE(film->w == reference->w && film->h == reference->h, "films differ in size");
        KS_REQUIRE(film->dev == reference->dev, "films belong to different devices");
        ks_device *dev = film->dev;
        KS_REQUIRE(dev->relmse_scratch, "the device has no relMSE scratch (device not opened by ks_device_open)");
        const size_t n = (size_t)film->w * film->h;
        KS_REQUIRE(n <= ((size_t)1 << 32), "ks_film_relmse_noalloc: films above 2^32 pixels are not supported");
        KS_HIP(hipSetDevice(dev->ordinal));
        // a fixed grid of at most kRelmseScratchBlocks blocks: the partials always fit the scratch, and
        // every one of them is written before it is summed (nothing to clear between calls)
        const unsigned blocks = (unsigned)std::min<size_t>(kRelmseScratchBlocks, (n + 255) / 256);
        double *partials = dev->relmse_scratch, *total = dev->relmse_scratch + kRelmseScratchBlocks;
        hipLaunchKernelGGL(k_film_relmse_strided, dim3(blocks), dim3(256), 0, dev->stream, n, film->rgb.as<double>(),
                           film->wts.as<double>(), make_float3(film->backdrop[0], film->backdrop[1], film->backdrop[2]),
                           reference->rgb.as<double>(), reference->wts.as<double>(),
                           make_float3(reference->backdrop[0], reference->backdrop[1], reference->backdrop[2]),
                           partials);
        hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, dev->stream, (const double *)partials, blocks, total);
        KS_HIP(hipGetLastError());
        // through the device's pinned host block, allocated when the device was opened
        KS_HIP(hipMemcpyAsync(dev->host_scratch, total, sizeof(double), hipMemcpyDeviceToHost, dev->stream));
        KS_HIP(hipStreamSynchronize(dev->stream));
        double sum;
        std::memcpy(&sum, dev->host_scratch, sizeof(double));
        *relmse = sum / (3.0 * (double)n);
        return KS_OK;
    });
}

int ks_small_pt_render_guided(ks_scene *scene, const ks_camera_desc *cam, const ks_small_pt_params *params,
                              ks_guide *guide, ks_film *film, ks_progress_fn progress, void *user,
                              ks_render_stats *stats) {
    return ksa::guarded([&]() -> int {
        KS_REQUIRE(scene && cam && params && film && guide, "null argument");
        KS_REQUIRE(film->w == params->render_width && film->h == params->render_height,
                   "film size differs from render_width x render_height");
        KS_REQUIRE(film->dev == scene->dev, "film and scene belong to different devices");
        KS_REQUIRE(ksg::guide_device(guide) == scene->dev, "guide and scene belong to different devices");
        KS_REQUIRE(params->mis_mode == KS_MIS_CONSISTENT,
                   "guided rendering needs mis_mode consistent (the reference weights do not sum to one)");
        ks_device *dev = scene->dev;
        KS_HIP(hipSetDevice(dev->ordinal));
        auto t0 = std::chrono::steady_clock::now();
        RenderConsts rc;
        int e = fill_consts(scene->dev, scene, cam, params, rc);
        if (e) return e;
        if (stats) std::memset(stats, 0, sizeof(*stats));
        const int interval = params->spp_prog_interval > 0 ? params->spp_prog_interval : 32;
        const uint64_t pixels = (uint64_t)rc.cw * rc.ch;
        KS_REQUIRE(pixels * kRecVertices < (1ull << 31), "crop window too large for guided rendering");
        const uint64_t max_samp = kSampleBufferBytes / sizeof(float4);
        const int spp_chunk = (int)std::max<uint64_t>(1, std::min<uint64_t>((uint64_t)interval, max_samp / pixels));
        const uint64_t max_items = pixels * (uint64_t)std::min(spp_chunk, params->spp);
        uint32_t Q = pick_queue(params, max_items); // each slot also holds a 160 B guide mixture
        e = ensure_workspace(dev, Q, max_items, scene->max_stack);
        if (e) return e;
        GuideWork &gw = dev->gw;
        if (gw.mix_q < Q) {
            KS_HIP(gw.mix.alloc(sizeof(float4) * 10 * (size_t)Q));
            KS_HIP(gw.list.alloc(sizeof(uint32_t) * ((size_t)Q + 64)));
            gw.mix_q = Q;
        }
        const float fraction = ksg::guide_fraction(guide);
        const int32_t max_steps = ksg::guide_max_steps(guide), batch = ksg::guide_batch(guide);
        const int32_t train_spp = ksg::guide_train_spp(guide);
        const bool any_training = max_steps > 0;
        if (any_training && gw.rec_items < pixels) {
            KS_HIP(gw.recs.alloc(sizeof(float4) * kRecFloat4 * kRecVertices * pixels));
            KS_HIP(hipMemsetAsync(gw.recs.p, 0, gw.recs.bytes, dev->stream));
            KS_HIP(gw.train.alloc(ksg::kTrainSampleBytes * kRecVertices * pixels));
            gw.rec_items = (uint32_t)pixels;
        }
        KS_HIP(gw.count.ensure(2 * sizeof(double))); // uint32 count, then the double sum of the weights
        clear_stats(dev, dev->stream);
        const uint32_t cap = (uint32_t)(kRecVertices * pixels);
        Timer tm{dev};
        int s = 0;
        while (s < params->spp) {
            // training passes hold one sample per pixel; frozen passes are batched like plain renders, without
            // crossing a progress boundary
            const bool train = any_training && (train_spp == 0 || s < train_spp);
            int nb = train ? 1 : std::min(spp_chunk, params->spp - s);
            nb = std::min(nb, (s / interval + 1) * interval - s);
            rc.s0 = s;
            rc.nb = nb;
            rc.work_list = nullptr;
            GuidePass gp;
            gp.guide = guide;
            gp.mix = gw.mix.as<float4>();
            gp.list = gw.list.as<uint32_t>();
            gp.n_list = gw.list.as<uint32_t>() + Q; // the count sits after the Q list entries
            // the guide samples once it has been trained at least once; before that the pass is BSDF-only
            const bool sample_guide = fraction > 0.0f && ksg::guide_steps(guide) > 0;
            gp.shade.mix = sample_guide ? gw.mix.as<float4>() : nullptr;
            gp.shade.recs = train ? gw.recs.as<float4>() : nullptr;
            gp.shade.alpha = fraction;
            gp.shade.min_roughness = kGuideMinRoughness;
            gp.shade.max_bounce = ksg::guide_max_bounce(guide);
            if (gw.stamp == 0xffffffffu) { // stamps wrap: clear the records so no stale stamp can match
                KS_HIP(hipMemsetAsync(gw.recs.p, 0, gw.recs.bytes, dev->stream));
                gw.stamp = 0;
            }
            gp.shade.stamp = ++gw.stamp;
            e = run_wavefront(scene, rc, pixels * (uint64_t)nb, Q, tm, stats, &gp);
            if (e) return e;
            hipEvent_t ev = tm.begin(KS_K_FILM);
            hipLaunchKernelGGL(k_film, dim3((unsigned)((pixels + 255) / 256)), dim3(256), 0, dev->stream, rc,
                               dev->ws.samp.as<float4>(), film->rgb.as<double>(), film->wts.as<double>());
            tm.end(ev);
            KS_HIP(hipGetLastError());
            if (train) {
                ev = tm.begin(KS_K_GUIDE);
                KS_HIP(hipMemsetAsync(gw.count.p, 0, 2 * sizeof(double), dev->stream));
                e = ksg::launch_extract(gw.recs.as<float4>(), dev->ws.samp.as<float4>(), (uint32_t)pixels,
                                        gp.shade.stamp, gw.train.p, gw.count.as<uint32_t>(), gw.count.as<double>() + 1,
                                        cap, dev->stream);
                tm.end(ev);
                if (e) return e;
                KS_HIP(hipMemcpyAsync(dev->host_scratch, gw.count.p, 2 * sizeof(double), hipMemcpyDeviceToHost,
                                      dev->stream));
                KS_HIP(hipStreamSynchronize(dev->stream));
                uint32_t n_found;
                double c_total;
                std::memcpy(&n_found, dev->host_scratch, sizeof(uint32_t));
                std::memcpy(&c_total, dev->host_scratch + sizeof(double), sizeof(double));
                const uint32_t n_train = std::min(n_found, cap);
                const float c_scale = c_total > 0.0 ? (float)((double)n_found / c_total) : 1.0f;
                // Adam steps of `batch` samples (2^18 by default, the configuration's batch): as many whole
                // batches as the pass holds, at most max_steps; a pass with fewer samples than one batch
                // makes one step over all of them. The samples are spread over the steps: step k takes
                // samples k, k + S, k + 2S, ...
                const int steps = n_train < 64u ? 0
                                                : (int)std::max<uint64_t>(1, std::min<uint64_t>((uint64_t)max_steps,
                                                                                                n_train / (uint32_t)batch));
                if (steps > 0) {
                    const int per_step = (int)std::min<uint64_t>((uint64_t)batch, n_train / (uint32_t)steps);
                    ev = tm.begin(KS_K_GUIDE);
                    for (int k = 0; k < steps; ++k) {
                        e = ksg::train_on_device(guide, static_cast<uint8_t *>(gw.train.p) + ksg::kTrainSampleBytes * k,
                                                 per_step, steps, c_scale, nullptr, dev->stream);
                        if (e) return e;
                    }
                    tm.end(ev);
                    if (stats) stats->guide_samples += (uint64_t)per_step * steps;
                }
            }
            s += nb;
            if (progress && (s % interval == 0 || s == params->spp)) {
                KS_HIP(hipStreamSynchronize(dev->stream));
                progress(film, s, user);
            }
        }
        KS_HIP(hipStreamSynchronize(dev->stream));
        tm.collect(stats);
        e = read_stats(dev, stats);
        if (e) return e;
        if (stats) {
            stats->paths = pixels * (uint64_t)params->spp;
            stats->queue_size = (int32_t)Q;
            stats->wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        }
        return KS_OK;
    });
}

} // extern "C"

namespace {

// One chunk of a batched render: segments (base item, view, first sample, pixels) covering at most the
// per-sample result budget, plus the film fold's thread bases and per-segment item counts.
struct BatchChunk {
    std::vector<uint4> segs;
    std::vector<uint32_t> thread_base, seg_items;
    uint64_t items = 0;
    uint32_t threads = 0;
};

std::vector<BatchChunk> schedule_batch(const std::vector<RenderConsts> &views, const std::vector<int> &spp,
                                       uint64_t budget) {
    std::vector<BatchChunk> chunks(1);
    for (size_t v = 0; v < views.size(); ++v) {
        uint64_t P = (uint64_t)views[v].cw * views[v].ch;
        int s0 = 0;
        while (s0 < spp[v]) {
            BatchChunk &c = chunks.back();
            uint64_t room = budget - c.items;
            int nb = (int)std::min<uint64_t>((uint64_t)(spp[v] - s0), room / P);
            if (nb == 0) {
                chunks.emplace_back();
                continue;
            }
            c.segs.push_back(make_uint4((uint32_t)c.items, (uint32_t)v, (uint32_t)s0, (uint32_t)P));
            c.seg_items.push_back((uint32_t)(P * (uint64_t)nb));
            c.thread_base.push_back(c.threads);
            c.threads += (uint32_t)P;
            c.items += P * (uint64_t)nb;
            s0 += nb;
        }
    }
    if (chunks.back().segs.empty()) chunks.pop_back();
    return chunks;
}

} // namespace

extern "C" {

int ks_small_pt_render_batch(ks_scene *scene, int32_t n_views, const ks_batch_view *views, const ks_batch_options *opt,
                             ks_render_stats *stats) {
    return ksa::guarded([&]() -> int {
        KS_REQUIRE(scene && views && n_views > 0 && n_views <= 65536, "invalid views");
        ks_device *dev = scene->dev;
        KS_HIP(hipSetDevice(dev->ordinal));
        auto t0 = std::chrono::steady_clock::now();
        ks_batch_options o{};
        if (opt) o = *opt;
        std::vector<RenderConsts> hv((size_t)n_views);
        std::vector<int> spp((size_t)n_views);
        std::vector<double *> rgbs((size_t)n_views), wtss((size_t)n_views);
        uint64_t paths = 0, max_pixels = 0;
        for (int v = 0; v < n_views; ++v) {
            const ks_batch_view &bv = views[v];
            KS_REQUIRE(bv.film && bv.film->dev == dev, "every view needs a film on the scene's device");
            KS_REQUIRE(bv.film->w == bv.params.render_width && bv.film->h == bv.params.render_height,
                       "film size differs from render_width x render_height");
            int e = fill_consts(scene->dev, scene, &bv.camera, &bv.params, hv[v]);
            if (e) return e;
            spp[v] = bv.params.spp;
            rgbs[v] = bv.film->rgb.as<double>();
            wtss[v] = bv.film->wts.as<double>();
            uint64_t P = (uint64_t)hv[v].cw * hv[v].ch;
            paths += P * (uint64_t)spp[v];
            max_pixels = std::max(max_pixels, P);
        }
        if (stats) std::memset(stats, 0, sizeof(*stats));
        uint64_t budget = kSampleBufferBytes / sizeof(float4); // per-sample results of one chunk
        KS_REQUIRE(o.max_chunk_items >= 0, "max_chunk_items must not be negative");
        if (o.max_chunk_items > 0) budget = std::min<uint64_t>(budget, (uint64_t)o.max_chunk_items);
        KS_REQUIRE(max_pixels <= budget, "a view is too large for one chunk");
        std::vector<BatchChunk> chunks = schedule_batch(hv, spp, budget);
        uint64_t max_items = 0;
        for (auto &c : chunks) max_items = std::max(max_items, c.items);
        HostToDev<RenderConsts> dviews;
        HostToDev<double *> drgb, dwts;
        KS_HIP(dviews.put(hv.data(), hv.size(), dev->stream));
        KS_HIP(drgb.put(rgbs.data(), rgbs.size(), dev->stream));
        KS_HIP(dwts.put(wtss.data(), wtss.size(), dev->stream));

        WavefrontOpts wo;
        wo.batch = true;
        wo.graph = o.disable_graph == 0;
        uint32_t Q = 0;
        // autotuner: queue size 2^18..2^25 x traversal blocks per CU {4, 8}, timed on the first chunk's first
        // items (at most 2^25), cached per scene and workload size
        if (o.queue_size > 0) {
            Q = (uint32_t)std::min<int64_t>(o.queue_size, 1 << 26);
            wo.blocks_per_cu = o.trace_blocks_per_cu > 0 ? o.trace_blocks_per_cu : kMaxTraceBlocksPerCu;
        } else if (dev->tune_scene == scene && dev->tune_items == paths) {
            Q = dev->tune_q;
            wo.blocks_per_cu = dev->tune_bpc;
        } else {
            const BatchChunk &c0 = chunks.front();
            BatchChunk probe;
            for (size_t k = 0; k < c0.segs.size() && probe.items < (1ull << 25); ++k) {
                probe.segs.push_back(c0.segs[k]);
                probe.items += c0.seg_items[k];
            }
            HostToDev<uint4> psegs;
            const int32_t probe_segs = (int32_t)probe.segs.size();
            probe.segs.push_back(make_uint4((uint32_t)probe.items, 0u, 0u, 0u)); // sentinel: the item count
            KS_HIP(psegs.put(probe.segs.data(), probe.segs.size(), dev->stream));
            RenderConsts prc = hv[0];
            prc.views = dviews.p();
            prc.segs = psegs.p();
            prc.n_segs = probe_segs;
            hipEvent_t ea, eb;
            KS_HIP(hipEventCreate(&ea));
            KS_HIP(hipEventCreate(&eb));
            float best = 1e30f;
            const int bpcs[2] = {4, 8};
            for (int lq = 18; lq <= 25; ++lq) {
                uint32_t q = (uint32_t)std::min<uint64_t>(1ull << lq, std::max<uint64_t>(probe.items, 256));
                for (int bpc : bpcs) {
                    int e = ensure_workspace(dev, q, probe.items, scene->max_stack);
                    if (e) return e;
                    WavefrontOpts pw = wo;
                    pw.blocks_per_cu = bpc;
                    Timer ptm{dev};
                    int saved = dev->timing;
                    dev->timing = 0;
                    KS_HIP(hipEventRecord(ea, dev->stream));
                    e = run_wavefront(scene, prc, probe.items, q, ptm, nullptr, nullptr, pw);
                    KS_HIP(hipEventRecord(eb, dev->stream));
                    dev->timing = saved;
                    if (e) return e;
                    KS_HIP(hipEventSynchronize(eb));
                    float ms = 0.0f;
                    KS_HIP(hipEventElapsedTime(&ms, ea, eb));
                    if (ms < best) {
                        best = ms;
                        Q = q;
                        wo.blocks_per_cu = bpc;
                    }
                }
                if ((1ull << lq) >= probe.items) break; // larger queues run the same work
            }
            (void)hipEventDestroy(ea);
            (void)hipEventDestroy(eb);
            dev->tune_scene = scene;
            dev->tune_items = paths;
            dev->tune_q = Q;
            dev->tune_bpc = wo.blocks_per_cu;
        }
        int e = ensure_workspace(dev, Q, max_items, scene->max_stack);
        if (e) return e;
        clear_stats(dev, dev->stream);
        Timer tm{dev};
        for (const BatchChunk &c : chunks) {
            // declared before the device copies, so it outlives them (their release synchronises)
            std::vector<uint4> segs = c.segs;
            segs.push_back(make_uint4((uint32_t)c.items, 0u, 0u, 0u)); // sentinel: the chunk's item count
            HostToDev<uint4> dsegs;
            HostToDev<uint32_t> dtb, dsi;
            KS_HIP(dsegs.put(segs.data(), segs.size(), dev->stream));
            KS_HIP(dtb.put(c.thread_base.data(), c.thread_base.size(), dev->stream));
            KS_HIP(dsi.put(c.seg_items.data(), c.seg_items.size(), dev->stream));
            RenderConsts rc = hv[0];
            rc.views = dviews.p();
            rc.segs = dsegs.p();
            rc.n_segs = (int32_t)c.segs.size();
            e = run_wavefront(scene, rc, c.items, Q, tm, stats, nullptr, wo);
            if (e) return e;
            hipEvent_t ev = tm.begin(KS_K_FILM);
            hipLaunchKernelGGL(k_film_batch, dim3((c.threads + 255) / 256), dim3(256), 0, dev->stream, dviews.p(), dsegs.p(),
                               dtb.p(), dsi.p(), (int)c.segs.size(), c.threads, dev->ws.samp.as<float4>(), drgb.p(),
                               dwts.p());
            tm.end(ev);
            KS_HIP(hipGetLastError());
            KS_HIP(hipStreamSynchronize(dev->stream)); // segment tables of this chunk are freed next
        }
        tm.collect(stats);
        e = read_stats(dev, stats);
        if (e) return e;
        if (stats) {
            stats->paths = paths;
            stats->queue_size = (int32_t)Q;
            stats->trace_blocks_per_cu = wo.blocks_per_cu;
            stats->batch_chunks = (int32_t)chunks.size();
            stats->wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        }
        return KS_OK;
    });
}

int ks_trace_samples(ks_scene *scene, const ks_camera_desc *cam, const ks_small_pt_params *params, int64_t n,
                     const int32_t *px, const int32_t *py, const int32_t *sample, float *out) {
    return ksa::guarded([&]() -> int {
        KS_REQUIRE(scene && cam && params && px && py && sample && out, "null argument");
        KS_REQUIRE(n > 0 && n < (1ll << 31), "sample count out of range");
        ks_device *dev = scene->dev;
        KS_HIP(hipSetDevice(dev->ordinal));
        RenderConsts rc;
        int e = fill_consts(scene->dev, scene, cam, params, rc);
        if (e) return e;
        std::vector<int4> wl((size_t)n);
        for (int64_t i = 0; i < n; ++i) {
            KS_REQUIRE(px[i] >= 0 && px[i] < rc.W && py[i] >= 0 && py[i] < rc.H && sample[i] >= 0 && sample[i] < rc.spp,
                       "pixel or sample index out of range");
            wl[i] = make_int4(px[i], py[i], sample[i], 0);
        }
        HostToDev<int4> dwl;
        KS_HIP(dwl.put(wl.data(), wl.size(), dev->stream));
        rc.work_list = dwl.p();
        rc.s0 = 0;
        rc.nb = 1;
        uint32_t Q = pick_queue(params, (uint64_t)n);
        e = ensure_workspace(dev, Q, (size_t)n, scene->max_stack);
        if (e) return e;
        clear_stats(dev, dev->stream);
        Timer tm{dev};
        e = run_wavefront(scene, rc, (uint64_t)n, Q, tm, nullptr);
        if (e) return e;
        tm.collect(nullptr);
        std::vector<float4> res((size_t)n);
        KS_HIP(hipMemcpyAsync(res.data(), dev->ws.samp.p, sizeof(float4) * n, hipMemcpyDeviceToHost, dev->stream));
        KS_HIP(hipStreamSynchronize(dev->stream));
        for (int64_t i = 0; i < n; ++i) {
            out[4 * i] = res[i].x;
            out[4 * i + 1] = res[i].y;
            out[4 * i + 2] = res[i].z;
            out[4 * i + 3] = res[i].w;
        }
        return KS_OK;
    });
}

int ks_trace_capture(ks_scene *scene, const ks_camera_desc *cam, const ks_small_pt_params *params, int64_t n,
                     const int32_t *px, const int32_t *py, const int32_t *sample, int64_t max_rays, float *closest_rays,
                     int64_t *n_closest, float *shadow_rays, int64_t *n_shadow) {
    return ksa::guarded([&]() -> int {
        KS_REQUIRE(scene && cam && params && px && py && sample && closest_rays && n_closest && shadow_rays && n_shadow,
                   "null argument");
        KS_REQUIRE(n > 0 && n < (1ll << 31), "sample count out of range");
        KS_REQUIRE(max_rays > 0 && max_rays <= (1ll << 28), "max_rays must be in [1, 2^28]");
        ks_device *dev = scene->dev;
        KS_HIP(hipSetDevice(dev->ordinal));
        RenderConsts rc;
        int e = fill_consts(scene->dev, scene, cam, params, rc);
        if (e) return e;
        std::vector<int4> wl((size_t)n);
        for (int64_t i = 0; i < n; ++i) {
            KS_REQUIRE(px[i] >= 0 && px[i] < rc.W && py[i] >= 0 && py[i] < rc.H && sample[i] >= 0 && sample[i] < rc.spp,
                       "pixel or sample index out of range");
            wl[i] = make_int4(px[i], py[i], sample[i], 0);
        }
        HostToDev<int4> dwl;
        KS_HIP(dwl.put(wl.data(), wl.size(), dev->stream));
        rc.work_list = dwl.p();
        rc.s0 = 0;
        rc.nb = 1;
        uint32_t Q = pick_queue(params, (uint64_t)n);
        e = ensure_workspace(dev, Q, (size_t)n, scene->max_stack);
        if (e) return e;
        DevBuf cc, cs, cn;
        KS_HIP(cc.alloc(sizeof(float4) * 2 * (size_t)max_rays));
        KS_HIP(cs.alloc(sizeof(float4) * 2 * (size_t)max_rays));
        KS_HIP(cn.alloc(2 * sizeof(uint32_t)));
        KS_HIP(hipMemsetAsync(cn.p, 0, 2 * sizeof(uint32_t), dev->stream));
        clear_stats(dev, dev->stream);
        WavefrontOpts wo;
        wo.cap_c = cc.as<float4>();
        wo.cap_s = cs.as<float4>();
        wo.cap_n = cn.as<uint32_t>();
        wo.cap_max = (uint32_t)max_rays;
        Timer tm{dev};
        e = run_wavefront(scene, rc, (uint64_t)n, Q, tm, nullptr, nullptr, wo);
        if (e) return e;
        tm.collect(nullptr);
        uint32_t counts[2];
        KS_HIP(hipMemcpyAsync(counts, cn.p, sizeof(counts), hipMemcpyDeviceToHost, dev->stream));
        KS_HIP(hipStreamSynchronize(dev->stream));
        KS_HIP(hipMemcpyAsync(closest_rays, cc.p, sizeof(float4) * 2 * counts[0], hipMemcpyDeviceToHost, dev->stream));
        KS_HIP(hipMemcpyAsync(shadow_rays, cs.p, sizeof(float4) * 2 * counts[1], hipMemcpyDeviceToHost, dev->stream));
        KS_HIP(hipStreamSynchronize(dev->stream));
        *n_closest = counts[0];
        *n_shadow = counts[1];
        return KS_OK;
    });
}

int ks_scene_traversal_counts(const ks_scene *scene, const float *rays, int64_t n, int32_t any_hit,
                              uint64_t *node_fetches, uint64_t *tri_tests, float *t_out) {
    return ksa::guarded([&]() -> int {
        KS_REQUIRE(scene && rays && node_fetches && tri_tests && n >= 0, "null argument");
        if (scene->host_nodes2.empty()) {
            ksa::set_error("traversal counts need the host SBVH (ks_bvh_options.builder = KS_BVH_BUILD_HOST_SBVH)");
            return KS_ERR_UNSUPPORTED;
        }
        ksa::count_reference_traversal(scene->host_nodes2, scene->host_tris, scene->info.bound_lo, scene->info.bound_hi,
                                       rays, n, any_hit != 0, *node_fetches, *tri_tests, t_out);
        return KS_OK;
    });
}

int ks_sobol_stream(ks_device *dev, int32_t width, int32_t height, int32_t spp, int32_t seed, int64_t n,
                    const int32_t *px, const int32_t *py, const int32_t *sample, int32_t n_calls, const int32_t *kinds,
                    uint32_t *out_u32, float *out_f) {
    return ksa::guarded([&]() -> int {
        KS_REQUIRE(dev && px && py && sample && kinds && out_u32 && out_f, "null argument");
        KS_REQUIRE(n > 0 && width > 0 && height > 0 && spp > 0 && n_calls > 0, "invalid sizes");
        int n_values = 0;
        for (int c = 0; c < n_calls; ++c) {
            KS_REQUIRE(kinds[c] == 1 || kinds[c] == 2, "kinds must be 1 or 2");
            n_values += kinds[c];
        }
        KS_HIP(hipSetDevice(dev->ordinal));
        int log2spp, ndigits;
        sobol_consts(width, height, spp, log2spp, ndigits);
        hipStream_t s = dev->stream;
        HostToDev<int32_t> dpx, dpy, ds, dk;
        KS_HIP(dpx.put(px, n, s));
        KS_HIP(dpy.put(py, n, s));
        KS_HIP(ds.put(sample, n, s));
        KS_HIP(dk.put(kinds, n_calls, s));
        DevBuf ou, of;
        KS_HIP(ou.alloc(sizeof(uint32_t) * n * n_values));
        KS_HIP(of.alloc(sizeof(float) * n * n_values));
        hipLaunchKernelGGL(k_sobol_stream, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, dpx.p(), dpy.p(),
                           ds.p(), seed, log2spp, ndigits, n_calls, dk.p(), n_values, dev->perm_tab.as<uint8_t>(),
                           ou.as<uint32_t>(), of.as<float>());
        KS_HIP(hipGetLastError());
        KS_HIP(hipMemcpyAsync(out_u32, ou.p, sizeof(uint32_t) * n * n_values, hipMemcpyDeviceToHost, s));
        KS_HIP(hipMemcpyAsync(out_f, of.p, sizeof(float) * n * n_values, hipMemcpyDeviceToHost, s));
        KS_HIP(hipStreamSynchronize(s));
        return KS_OK;
    });
}

int ks_camera_rays(ks_device *dev, const ks_camera_desc *cam, int32_t width, int32_t height, int32_t spp,
                   int32_t seed, int64_t n, const int32_t *px, const int32_t *py, const int32_t *sample, float *rays) {
    return ksa::guarded([&]() -> int {
        KS_REQUIRE(dev && cam && px && py && sample && rays, "null argument");
        KS_REQUIRE(n > 0, "n must be positive");
        KS_HIP(hipSetDevice(dev->ordinal));
        ks_small_pt_params p{};
        p.render_width = width;
        p.render_height = height;
        p.spp = spp;
        p.rng_seed = seed;
        p.bounces = 1;
        RenderConsts rc;
        int e = fill_consts(dev, nullptr, cam, &p, rc);
        if (e) return e;
        hipStream_t s = dev->stream;
        HostToDev<int32_t> dpx, dpy, ds;
        KS_HIP(dpx.put(px, n, s));
        KS_HIP(dpy.put(py, n, s));
        KS_HIP(ds.put(sample, n, s));
        DevBuf o;
        KS_HIP(o.alloc(sizeof(float) * 6 * n));
        hipLaunchKernelGGL(k_camera_rays, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, rc, n, dpx.p(), dpy.p(),
                           ds.p(), o.as<float>());
        KS_HIP(hipGetLastError());
        KS_HIP(hipMemcpyAsync(rays, o.p, sizeof(float) * 6 * n, hipMemcpyDeviceToHost, s));
        KS_HIP(hipStreamSynchronize(s));
        return KS_OK;
    });
}

static int ray_query(ks_scene *scene, const float *rays, int64_t n, float *hits, uint8_t *occ) {
    KS_REQUIRE(scene && rays, "null argument");
    KS_REQUIRE(n > 0 && n < (1ll << 31), "ray count out of range");
    ks_device *dev = scene->dev;
    KS_HIP(hipSetDevice(dev->ordinal));
    hipStream_t s = dev->stream;
    std::vector<float4> o((size_t)n), d((size_t)n);
    for (int64_t i = 0; i < n; ++i) {
        const float *r = rays + 8 * i;
        o[i] = make_float4(r[0], r[1], r[2], r[3]);
        d[i] = make_float4(r[4], r[5], r[6], r[7]);
    }
    HostToDev<float4> dO, dD;
    KS_HIP(dO.put(o.data(), o.size(), s));
    KS_HIP(dD.put(d.data(), d.size(), s));
    int e = ensure_workspace(dev, 256, 1, scene->max_stack);
    if (e) return e;
    clear_stats(dev, s);
    int grid = trace_grid(dev, scene->ds, (uint32_t)n);
    DevBuf out;
    if (hits) {
        KS_HIP(out.alloc(sizeof(float4) * n));
        launch_closest(scene->ds, dim3(grid), s, (const uint32_t *)nullptr, (uint32_t)n,
                       (const float4 *)dO.p(), (const float4 *)dD.p(), out.as<float4>(), dev->ws.ovf.as<uint32_t>(),
                       dev->ws.ovf_cap, stat_rows(dev));
        KS_HIP(hipGetLastError());
        KS_HIP(hipMemcpyAsync(hits, out.p, sizeof(float4) * n, hipMemcpyDeviceToHost, s));
    } else {
        KS_HIP(out.alloc(n));
        launch_anyhit(scene->ds, dim3(grid), s, (const uint32_t *)nullptr, (uint32_t)n, (const float4 *)dO.p(),
                      (const float4 *)dD.p(), out.as<uint8_t>(), dev->ws.ovf.as<uint32_t>(), dev->ws.ovf_cap,
                      stat_rows(dev));
        KS_HIP(hipGetLastError());
        KS_HIP(hipMemcpyAsync(occ, out.p, n, hipMemcpyDeviceToHost, s));
    }
    KS_HIP(hipStreamSynchronize(s));
    return KS_OK;
}

int ks_intersect(ks_scene *scene, const float *rays, int64_t n, float *hits) {
    return ksa::guarded([&]() -> int {
        KS_REQUIRE(hits, "hits is null");
        return ray_query(scene, rays, n, hits, nullptr);
    });
}

int ks_occluded(ks_scene *scene, const float *rays, int64_t n, uint8_t *occluded) {
    return ksa::guarded([&]() -> int {
        KS_REQUIRE(occluded, "occluded is null");
        return ray_query(scene, rays, n, nullptr, occluded);
    });
}

int ks_bsdf_eval(ks_device *dev, const ks_material_desc *m, const float *wo, const float *wi, int64_t n, float *f,
                 float *pdf) {
    return ksa::guarded([&]() -> int {
        KS_REQUIRE(dev && m && wo && wi && f && pdf && n > 0, "invalid argument");
        KS_HIP(hipSetDevice(dev->ordinal));
        hipStream_t s = dev->stream;
        HostToDev<float> dwo, dwi;
        KS_HIP(dwo.put(wo, 3 * n, s));
        KS_HIP(dwi.put(wi, 3 * n, s));
        DevBuf df, dp;
        KS_HIP(df.alloc(sizeof(float) * 3 * n));
        KS_HIP(dp.alloc(sizeof(float) * n));
        hipLaunchKernelGGL(k_bsdf_eval, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, to_device_material(*m), n,
                           dwo.p(), dwi.p(), df.as<float>(), dp.as<float>());
        KS_HIP(hipGetLastError());
        KS_HIP(hipMemcpyAsync(f, df.p, sizeof(float) * 3 * n, hipMemcpyDeviceToHost, s));
        KS_HIP(hipMemcpyAsync(pdf, dp.p, sizeof(float) * n, hipMemcpyDeviceToHost, s));
        KS_HIP(hipStreamSynchronize(s));
        return KS_OK;
    });
}

int ks_bsdf_sample(ks_device *dev, const ks_material_desc *m, const float *wo, const float *u, int64_t n, float *wi,
                   float *beta, float *pdf) {
    return ksa::guarded([&]() -> int {
        KS_REQUIRE(dev && m && wo && u && wi && beta && pdf && n > 0, "invalid argument");
        KS_HIP(hipSetDevice(dev->ordinal));
        hipStream_t s = dev->stream;
        HostToDev<float> dwo, du;
        KS_HIP(dwo.put(wo, 3 * n, s));
        KS_HIP(du.put(u, 3 * n, s));
        DevBuf dwi, db, dp;
        KS_HIP(dwi.alloc(sizeof(float) * 3 * n));
        KS_HIP(db.alloc(sizeof(float) * 3 * n));
        KS_HIP(dp.alloc(sizeof(float) * n));
        hipLaunchKernelGGL(k_bsdf_sample, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, to_device_material(*m), n,
                           dwo.p(), du.p(), dwi.as<float>(), db.as<float>(), dp.as<float>());
        KS_HIP(hipGetLastError());
        KS_HIP(hipMemcpyAsync(wi, dwi.p, sizeof(float) * 3 * n, hipMemcpyDeviceToHost, s));
        KS_HIP(hipMemcpyAsync(beta, db.p, sizeof(float) * 3 * n, hipMemcpyDeviceToHost, s));
        KS_HIP(hipMemcpyAsync(pdf, dp.p, sizeof(float) * n, hipMemcpyDeviceToHost, s));
        KS_HIP(hipStreamSynchronize(s));
        return KS_OK;
    });
}

int ks_host_scene_create(const char *name, uint64_t seed, ks_host_scene **out) {
    return ksa::guarded([&]() -> int {
        KS_REQUIRE(name && out, "null argument");
        auto hs = std::make_unique<ks_host_scene>();
        if (!ksa::make_procedural_scene(name, seed, hs->s)) {
            ksa::set_error(std::string("unknown procedural scene: ") + name);
            return KS_ERR_INVALID;
        }
        *out = hs.release();
        return KS_OK;
    });
}

int ks_host_scene_get(const ks_host_scene *hs, ks_scene_desc *d, ks_camera_desc *cam) {
    return ksa::guarded([&]() -> int {
        KS_REQUIRE(hs, "host scene is null");
        const ksa::HostScene &s = hs->s;
        if (d) {
            std::memset(d, 0, sizeof(*d));
            d->vertices = s.vertices.data();
            d->n_vertices = (uint32_t)(s.vertices.size() / 3);
            d->normals = s.normals.empty() ? nullptr : s.normals.data();
            d->indices = s.indices.data();
            d->n_tris = (uint32_t)(s.indices.size() / 3);
            d->tri_geom = s.tri_geom.data();
            d->geom_material = s.geom_material.data();
            d->n_geoms = (uint32_t)s.geom_material.size();
            d->materials = s.materials.data();
            d->n_materials = (uint32_t)s.materials.size();
            d->lights = s.lights.empty() ? nullptr : s.lights.data();
            d->n_lights = (uint32_t)s.lights.size();
        }
        if (cam) *cam = s.camera;
        return KS_OK;
    });
}

int ks_host_scene_free(ks_host_scene *hs) {
    return ksa::guarded([&]() -> int {
        delete hs;
        return KS_OK;
    });
}

} // extern "C"

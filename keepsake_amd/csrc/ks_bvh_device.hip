// BVH build on the device: parallel locally-ordered clustering (PLOC; Meister and Bittner, "Parallel
// Locally-Ordered Clustering for Bounding Volume Hierarchy Construction", TVCG 2018) over the Morton
// order of the triangles' box centroids. Each new interior node decides by the surface area heuristic
// whether its subtree becomes one leaf. A top-down pass per level then lays the tree out in the node
// formats the traversal reads (GpuNode4 / GpuNode) and orders the triangle records (GpuTri) leaf by
// leaf.
//
// The reference builds its SBVH on the host (cuda/sbvh.cu:301-440); its kd-tree builder
// (cuda/cuda_parallel_kd_tree.cu) is a device pipeline of the same shape: bounds, sort, build,
// layout. This builder has no spatial splits, so each triangle is referenced exactly once.
//
// Kernels, n triangles:
//   k_bounds      scene box and centroid box (block reduction, then ordered-integer atomics), per-
//               triangle boxes (32 B each)
//   k_morton      63-bit Morton key of each centroid, value = triangle id
//   radix sort    (rocPRIM) of the (key, id) pairs over bits 0..62
//   k_leaves      leaf nodes in sorted order
//   per round     k_ploc_nn (neighbour search, the window staged in LDS), k_ploc_merge, rocPRIM select
//   k_layout      one launch per level of the output tree
//   k_records     triangle records in leaf order (vertices, triangle id, alpha-test flag)
#include <hip/hip_runtime.h>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_select.hpp>

#include "ks_internal.h"

#include <algorithm>
#include <chrono>
#include <cstring>
#include <string>
#include <vector>

namespace ksa {
namespace {

constexpr int kBlock = 256;
constexpr int kMaxLevels = 256; // levels of the output tree the layout accepts (checked, not assumed)
static_assert(kMaxLevels <= kBlock, "k_init_acc clears one level counter per thread of one block");

// order-preserving map of floats onto unsigned integers (for atomic min / max)
__device__ __host__ inline uint32_t ordered(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return (u >> 31) ? ~u : (u | 0x80000000u);
}
inline float unordered(uint32_t o) {
    uint32_t u = (o >> 31) ? (o & 0x7fffffffu) : ~o;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

__device__ inline float area_of(float3 lo, float3 hi) {
    float ex = hi.x - lo.x, ey = hi.y - lo.y, ez = hi.z - lo.z;
    if (!(ex >= 0.0f && ey >= 0.0f && ez >= 0.0f)) return 0.0f;
    return ex * ey + ey * ez + ez * ex;
}

// Reserves k slots of a shared counter for each thread of the block with one atomic per block (a
// counter that every wave bumps serialises on one address: 156K atomics per pass over 10M clusters).
// Every thread of the block must call it. Returns this thread's first slot.
__device__ inline uint32_t block_reserve(uint32_t *counter, uint32_t k) {
    __shared__ uint32_t wave_base[kBlock / 64];
    __shared__ uint32_t block_base;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t incl = k;
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t v = __shfl_up(incl, o);
        if (lane >= o) incl += v;
    }
    if (lane == 63) wave_base[wave] = incl;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t total = 0;
        for (int w = 0; w < kBlock / 64; ++w) {
            const uint32_t v = wave_base[w];
            wave_base[w] = total;
            total += v;
        }
        block_base = total != 0u ? atomicAdd(counter, total) : 0u;
    }
    __syncthreads();
    return block_base + wave_base[wave] + incl - k;
}

// acc: [0..2] min of triangle boxes, [3..5] max, [6..8] min of centroids, [9..11] max
__global__ void k_init_acc(uint32_t *acc, uint32_t *level_count) {
    int i = threadIdx.x;
    if (i < 12) acc[i] = (i < 3 || (i >= 6 && i < 9)) ? 0xffffffffu : 0u;
    if (i < kMaxLevels) level_count[i] = 0u;
}

__device__ inline float wave_min(float v) {
    for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o));
    return v;
}
__device__ inline float wave_max(float v) {
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
    return v;
}

__global__ __launch_bounds__(kBlock) void k_bounds(const float4 *__restrict__ verts, const uint4 *__restrict__ tri_idx,
                                                   uint32_t n, float4 *__restrict__ tri_box, uint32_t *acc) {
    const float inf = __builtin_inff();
    float m[12];
    for (int k = 0; k < 12; ++k) m[k] = (k < 3 || (k >= 6 && k < 9)) ? inf : -inf;
    for (uint32_t t = blockIdx.x * kBlock + threadIdx.x; t < n; t += gridDim.x * kBlock) {
        uint4 ix = tri_idx[t];
        float4 a = verts[ix.x], b = verts[ix.y], c = verts[ix.z];
        float3 lo = make_float3(fminf(fminf(a.x, b.x), c.x), fminf(fminf(a.y, b.y), c.y), fminf(fminf(a.z, b.z), c.z));
        float3 hi = make_float3(fmaxf(fmaxf(a.x, b.x), c.x), fmaxf(fmaxf(a.y, b.y), c.y), fmaxf(fmaxf(a.z, b.z), c.z));
        tri_box[2 * (size_t)t] = make_float4(lo.x, lo.y, lo.z, 0.0f);
        tri_box[2 * (size_t)t + 1] = make_float4(hi.x, hi.y, hi.z, 0.0f);
        float3 ce = make_float3(0.5f * (lo.x + hi.x), 0.5f * (lo.y + hi.y), 0.5f * (lo.z + hi.z));
        m[0] = fminf(m[0], lo.x), m[1] = fminf(m[1], lo.y), m[2] = fminf(m[2], lo.z);
        m[3] = fmaxf(m[3], hi.x), m[4] = fmaxf(m[4], hi.y), m[5] = fmaxf(m[5], hi.z);
        m[6] = fminf(m[6], ce.x), m[7] = fminf(m[7], ce.y), m[8] = fminf(m[8], ce.z);
        m[9] = fmaxf(m[9], ce.x), m[10] = fmaxf(m[10], ce.y), m[11] = fmaxf(m[11], ce.z);
    }
    __shared__ float part[kBlock / 64][12];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int k = 0; k < 12; ++k) {
        bool is_min = k < 3 || (k >= 6 && k < 9);
        float v = is_min ? wave_min(m[k]) : wave_max(m[k]);
        if (lane == 0) part[wave][k] = v;
    }
    __syncthreads();
    if (threadIdx.x < 12) {
        int k = threadIdx.x;
        bool is_min = k < 3 || (k >= 6 && k < 9);
        float v = part[0][k];
        for (int w = 1; w < kBlock / 64; ++w) v = is_min ? fminf(v, part[w][k]) : fmaxf(v, part[w][k]);
        if (is_min) atomicMin(acc + k, ordered(v));
        else atomicMax(acc + k, ordered(v));
    }
}

__device__ inline float unordered_dev(uint32_t o) { return __uint_as_float((o >> 31) ? (o & 0x7fffffffu) : ~o); }

// 21 bits spread to every third bit of 63
__device__ inline uint64_t spread3(uint32_t v) {
    uint64_t x = v & 0x1fffffu;
    x = (x | x << 32) & 0x1f00000000ffffull;
    x = (x | x << 16) & 0x1f0000ff0000ffull;
    x = (x | x << 8) & 0x100f00f00f00f00full;
    x = (x | x << 4) & 0x10c30c30c30c30c3ull;
    x = (x | x << 2) & 0x1249249249249249ull;
    return x;
}

__global__ __launch_bounds__(kBlock) void k_morton(const float4 *__restrict__ tri_box, uint32_t n, const uint32_t *acc,
                                                   uint64_t *__restrict__ keys, uint32_t *__restrict__ ids) {
    uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= n) return;
    float lo[3], sc[3];
    for (int k = 0; k < 3; ++k) {
        lo[k] = unordered_dev(acc[6 + k]);
        float e = unordered_dev(acc[9 + k]) - lo[k];
        sc[k] = e > 0.0f ? 2097152.0f / e : 0.0f;
    }
    float4 a = tri_box[2 * (size_t)t], b = tri_box[2 * (size_t)t + 1];
    float c[3] = {0.5f * (a.x + b.x), 0.5f * (a.y + b.y), 0.5f * (a.z + b.z)};
    uint32_t q[3];
    for (int k = 0; k < 3; ++k) {
        float x = fminf(fmaxf((c[k] - lo[k]) * sc[k], 0.0f), 2097151.0f); // NaN -> 0
        q[k] = (uint32_t)x;
    }
    keys[t] = (spread3(q[0]) << 2) | (spread3(q[1]) << 1) | spread3(q[2]);
    ids[t] = t;
}

// Node ids: 0 .. n-1 are the leaves (sorted position p, triangle ids[p]);
// n .. 2n-2 are interior nodes, whose children are child[id - n] and whose triangle count is
// count[id - n]. node_box[2 id] = (lo, SAH cost), node_box[2 id + 1] = (hi, 1 if the subtree is laid
// out as one leaf).

__device__ inline void put_box(float4 *node_box, uint32_t id, float3 lo, float3 hi, float cost, bool terminal) {
    node_box[2 * (size_t)id] = make_float4(lo.x, lo.y, lo.z, cost);
    node_box[2 * (size_t)id + 1] = make_float4(hi.x, hi.y, hi.z, terminal ? 1.0f : 0.0f);
}

// leaf nodes: the triangle's box, SAH cost = its area (intersection cost 1)
__global__ __launch_bounds__(kBlock) void k_leaves(const float4 *__restrict__ tri_box, const uint32_t *__restrict__ ids,
                                                   uint32_t n, float4 *__restrict__ node_box,
                                                   uint32_t *__restrict__ clusters) {
    uint32_t p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= n) return;
    const uint32_t t = ids[p];
    const float4 a = tri_box[2 * (size_t)t], b = tri_box[2 * (size_t)t + 1];
    const float3 lo = make_float3(a.x, a.y, a.z), hi = make_float3(b.x, b.y, b.z);
    put_box(node_box, p, lo, hi, area_of(lo, hi), true);
    clusters[p] = p;
}

// SAH decision for a new interior node over children a and b (their boxes and costs in node_box)
__device__ inline void fit_node(float4 *node_box, uint32_t *count, uint32_t n, uint32_t id, uint32_t a, uint32_t b,
                                uint32_t max_leaf, float trav_cost) {
    const float4 la = node_box[2 * (size_t)a], ha = node_box[2 * (size_t)a + 1];
    const float4 lb = node_box[2 * (size_t)b], hb = node_box[2 * (size_t)b + 1];
    const float3 lo = make_float3(fminf(la.x, lb.x), fminf(la.y, lb.y), fminf(la.z, lb.z));
    const float3 hi = make_float3(fmaxf(ha.x, hb.x), fmaxf(ha.y, hb.y), fmaxf(ha.z, hb.z));
    const uint32_t c = (a < n ? 1u : count[a - n]) + (b < n ? 1u : count[b - n]);
    const float area = area_of(lo, hi);
    const float split = trav_cost * area + la.w + lb.w;
    const float as_leaf = area * (float)c;
    const bool leaf = c <= max_leaf && as_leaf <= split;
    count[id - n] = c;
    put_box(node_box, id, lo, hi, leaf ? as_leaf : split, leaf);
}

// ---- clustering ----
// Parallel locally-ordered clustering (Meister and Bittner, "Parallel Locally-Ordered Clustering for
// Bounding Volume Hierarchy Construction", TVCG 2018): the clusters stay in Morton order; each finds
// the neighbour within kRadius positions whose merged box has the smallest surface area; mutual
// neighbours merge; the array is compacted; repeat until one cluster is left.
constexpr int kRadius = 16;

// Neighbour order: smaller merged area first; equal areas by a hash of the pair of positions, then by
// the pair itself. That is a total order on pairs, the same seen from either end, so the globally
// smallest pair is mutual and every round merges at least one pair. The hash matters where many boxes
// are equal (duplicated geometry): ordering ties by position alone would merge one pair per round.
__device__ inline uint32_t pair_hash(uint32_t lo, uint32_t hi) {
    uint32_t x = lo * 0x9E3779B1u ^ hi;
    x ^= x >> 16;
    x *= 0x85EBCA6Bu;
    x ^= x >> 13;
    x *= 0xC2B2AE35u;
    x ^= x >> 16;
    return x;
}
__device__ inline bool closer(float d, uint32_t i, uint32_t j, float bd, uint32_t bj) {
    if (d != bd) return d < bd;
    const uint32_t lo = min(i, j), hi = max(i, j), blo = min(i, bj), bhi = max(i, bj);
    const uint32_t h = pair_hash(lo, hi), bh = pair_hash(blo, bhi);
    if (h != bh) return h < bh;
    return lo != blo ? lo < blo : hi < bhi;
}

__global__ __launch_bounds__(kBlock) void k_ploc_nn(const uint32_t *__restrict__ clusters, uint32_t m,
                                                    const float4 *__restrict__ node_box, uint32_t *__restrict__ nn) {
    __shared__ float3 slo[kBlock + 2 * kRadius], shi[kBlock + 2 * kRadius];
    const int64_t base = (int64_t)blockIdx.x * kBlock - kRadius;
    for (int k = threadIdx.x; k < kBlock + 2 * kRadius; k += kBlock) {
        const int64_t idx = base + k;
        if (idx >= 0 && idx < (int64_t)m) {
            const uint32_t id = clusters[idx];
            const float4 a = node_box[2 * (size_t)id], b = node_box[2 * (size_t)id + 1];
            slo[k] = make_float3(a.x, a.y, a.z);
            shi[k] = make_float3(b.x, b.y, b.z);
        }
    }
    __syncthreads();
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= m) return;
    const int me = threadIdx.x + kRadius;
    const float3 lo = slo[me], hi = shi[me];
    float best = __builtin_inff();
    uint32_t bj = i;
    for (int o = -kRadius; o <= kRadius; ++o) {
        const int64_t j = (int64_t)i + o;
        if (o == 0 || j < 0 || j >= (int64_t)m) continue;
        const float3 l2 = slo[me + o], h2 = shi[me + o];
        const float3 ul = make_float3(fminf(lo.x, l2.x), fminf(lo.y, l2.y), fminf(lo.z, l2.z));
        const float3 uh = make_float3(fmaxf(hi.x, h2.x), fmaxf(hi.y, h2.y), fmaxf(hi.z, h2.z));
        const float d = fminf(area_of(ul, uh), 3.0e38f); // non-finite boxes sort last
        if (bj == i || closer(d, i, (uint32_t)j, best, bj)) best = d, bj = (uint32_t)j;
    }
    nn[i] = bj;
}

// Mutual neighbours i < j merge into a new node that takes position i; position j is dropped. Each
// thread takes kMergeItems positions, kBlock apart; every thread reaches the reservation.
constexpr int kMergeItems = 4;

__global__ __launch_bounds__(kBlock) void k_ploc_merge(const uint32_t *__restrict__ clusters, uint32_t m,
                                                       const uint32_t *__restrict__ nn, uint32_t n, uint32_t *next_id,
                                                       float4 *node_box, uint2 *__restrict__ child, uint32_t *count,
                                                       uint32_t *__restrict__ merged, uint8_t *__restrict__ keep,
                                                       uint32_t max_leaf, float trav_cost) {
    uint32_t pos[kMergeItems], partner[kMergeItems];
    uint32_t makes = 0;
    bool make[kMergeItems];
#pragma unroll
    for (int r = 0; r < kMergeItems; ++r) {
        const uint32_t i = (blockIdx.x * kMergeItems + r) * kBlock + threadIdx.x;
        pos[r] = i;
        make[r] = false;
        if (i >= m) continue;
        const uint32_t j = nn[i];
        partner[r] = j;
        const bool mutual = j < m && nn[j] == i;
        if (!mutual) {
            merged[i] = clusters[i];
            keep[i] = 1;
        } else if (j < i) {
            keep[i] = 0; // the upper end of a pair
        } else {
            make[r] = true;
            ++makes;
        }
    }
    uint32_t k = block_reserve(next_id, makes);
#pragma unroll
    for (int r = 0; r < kMergeItems; ++r) {
        if (!make[r]) continue;
        const uint32_t i = pos[r];
        if (k >= n - 1) { // cannot happen (n - 1 merges in all); the host checks the counter
            keep[i] = 0;
            continue;
        }
        const uint32_t id = n + k, a = clusters[i], b = clusters[partner[r]];
        child[k] = make_uint2(a, b);
        fit_node(node_box, count, n, id, a, b, max_leaf, trav_cost);
        merged[i] = id;
        keep[i] = 1;
        ++k;
    }
}

// ---- layout ----

struct Slot {
    float3 lo, hi;
    bool terminal;
};

__device__ inline Slot slot_of(const float4 *__restrict__ node_box, uint32_t id) {
    const float4 a = node_box[2 * (size_t)id], b = node_box[2 * (size_t)id + 1];
    return Slot{make_float3(a.x, a.y, a.z), make_float3(b.x, b.y, b.z), b.w != 0.0f};
}

// Queue entry: {tree node, output node index, first triangle record of its subtree}. Each output node
// opens the interior child of largest surface area until it holds W children (a root that is itself a
// leaf gets a node with that one leaf). A child laid out as a leaf owns the records [first, first +
// count), filled here in depth-first order; an interior child is queued for the next level.
template <int W>
__global__ __launch_bounds__(kBlock) void k_layout(const uint4 *__restrict__ queue, uint32_t qn, bool root_level,
                                                   uint32_t n, const float4 *__restrict__ node_box,
                                                   const uint2 *__restrict__ child, const uint32_t *__restrict__ count,
                                                   const uint32_t *__restrict__ ids, void *__restrict__ out_nodes,
                                                   uint32_t *__restrict__ order, uint4 *__restrict__ next,
                                                   uint32_t *next_count, uint32_t next_base, uint32_t cap) {
    const uint32_t qi = blockIdx.x * kBlock + threadIdx.x;
    // every thread reaches the reservation below; threads past the queue (or past the node bound, which
    // cannot happen: an output node per interior tree node) lay out nothing
    const bool live = qi < qn && queue[qi].y < cap;
    const uint4 e = live ? queue[qi] : make_uint4(0u, 0u, 0u, 0u);
    uint32_t c[W];
    Slot cs[W];
    int nc = 0;
    const Slot self = live ? slot_of(node_box, e.x) : Slot{};
    if (!live) {
    } else if (root_level && self.terminal) {
        c[0] = e.x;
        cs[0] = self;
        nc = 1;
    } else {
        const uint2 ch = child[e.x - n];
        c[0] = ch.x;
        c[1] = ch.y;
        cs[0] = slot_of(node_box, ch.x);
        cs[1] = slot_of(node_box, ch.y);
        nc = 2;
        while (nc < W) {
            int best = -1;
            float best_area = -1.0f;
            uint32_t g = 0;
#pragma unroll
            for (int k = 0; k < W; ++k)
                if (k < nc && !cs[k].terminal) {
                    const float a = area_of(cs[k].lo, cs[k].hi);
                    if (a > best_area) best = k, best_area = a, g = c[k];
                }
            if (best < 0) break;
            const uint2 gc = child[g - n];
            // g's two children take its place, in order
#pragma unroll
            for (int k = W - 1; k > 0; --k)
                if (k > best + 1 && k <= nc) c[k] = c[k - 1], cs[k] = cs[k - 1];
#pragma unroll
            for (int k = 0; k < W; ++k) {
                if (k == best) c[k] = gc.x, cs[k] = slot_of(node_box, gc.x);
                if (k == best + 1) c[k] = gc.y, cs[k] = slot_of(node_box, gc.y);
            }
            ++nc;
        }
    }
    uint32_t interior = 0;
#pragma unroll
    for (int k = 0; k < W; ++k)
        if (k < nc && !cs[k].terminal) ++interior;
    uint32_t q = block_reserve(next_count, interior);
    if (!live) return;
    uint32_t ptr[W];
    uint32_t first = e.z;
#pragma unroll
    for (int k = 0; k < W; ++k) {
        ptr[k] = kLeafBit;
        if (k >= nc) continue;
        const uint32_t cnt = c[k] < n ? 1u : count[c[k] - n];
        if (cs[k].terminal) {
            ptr[k] = kLeafBit | (first << 4) | (cnt - 1);
            // the subtree's leaves in depth-first order (at most kMaxLeafRefs of them)
            uint32_t stack[kMaxLeafRefs];
            int sp = 0;
            stack[sp++] = c[k];
            uint32_t w = first;
            while (sp > 0) {
                const uint32_t x = stack[--sp];
                if (x < n) {
                    if (w < first + cnt) order[w++] = ids[x];
                } else if (sp + 2 <= (int)kMaxLeafRefs) {
                    const uint2 xc = child[x - n];
                    stack[sp++] = xc.y;
                    stack[sp++] = xc.x;
                }
            }
        } else {
            if (next_base + q < cap) next[q] = make_uint4(c[k], next_base + q, first, 0u);
            ptr[k] = next_base + q;
            ++q;
        }
        first += cnt;
    }
    const float inf = __builtin_inff();
    if (W == 4) {
        GpuNode4 nd;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const bool used = k < nc;
            nd.lox[k] = used ? cs[k].lo.x : inf;
            nd.hix[k] = used ? cs[k].hi.x : -inf;
            nd.loy[k] = used ? cs[k].lo.y : inf;
            nd.hiy[k] = used ? cs[k].hi.y : -inf;
            nd.loz[k] = used ? cs[k].lo.z : inf;
            nd.hiz[k] = used ? cs[k].hi.z : -inf;
            nd.child[k] = ptr[k];
            nd.pad[k] = 0u;
        }
        reinterpret_cast<GpuNode4 *>(out_nodes)[e.y] = nd;
    } else {
        GpuNode nd;
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const bool used = k < nc;
            nd.f[4 * k] = used ? cs[k].lo.x : inf;
            nd.f[4 * k + 1] = used ? cs[k].hi.x : -inf;
            nd.f[4 * k + 2] = used ? cs[k].lo.y : inf;
            nd.f[4 * k + 3] = used ? cs[k].hi.y : -inf;
            nd.f[8 + 2 * k] = used ? cs[k].lo.z : inf;
            nd.f[9 + 2 * k] = used ? cs[k].hi.z : -inf;
            nd.child[k] = ptr[k];
            nd.pad[k] = 0u;
        }
        reinterpret_cast<GpuNode *>(out_nodes)[e.y] = nd;
    }
}

// triangle records in leaf order; v1.w = 1 marks a triangle whose material is alpha tested
__global__ __launch_bounds__(kBlock) void k_records(const float4 *__restrict__ verts, const uint4 *__restrict__ tri_idx,
                                                    const uint32_t *__restrict__ order, uint32_t n,
                                                    const uint8_t *__restrict__ mat_alpha, float4 *__restrict__ out) {
    uint32_t p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= n) return;
    const uint32_t t = order[p];
    const uint4 ix = tri_idx[t];
    const float4 a = verts[ix.x], b = verts[ix.y], c = verts[ix.z];
    const uint32_t alpha = mat_alpha && mat_alpha[ix.w & 0xffffffu] ? 1u : 0u;
    out[3 * (size_t)p] = make_float4(a.x, a.y, a.z, __uint_as_float(t));
    out[3 * (size_t)p + 1] = make_float4(b.x, b.y, b.z, __uint_as_float(alpha));
    out[3 * (size_t)p + 2] = make_float4(c.x, c.y, c.z, 0.0f);
}

struct Scratch {
    std::vector<void *> ptrs;
    ~Scratch() {
        for (void *p : ptrs) (void)hipFree(p);
    }
    template <typename T> hipError_t get(T *&p, size_t count) {
        void *v = nullptr;
        hipError_t e = hipMalloc(&v, std::max<size_t>(count * sizeof(T), 16));
        if (e != hipSuccess) {
            (void)hipGetLastError(); // reported through the return value only
            return e;
        }
        ptrs.push_back(v);
        p = reinterpret_cast<T *>(v);
        return hipSuccess;
    }
    void release(void *p) { ptrs.erase(std::remove(ptrs.begin(), ptrs.end(), p), ptrs.end()); }
};

inline uint32_t blocks_for(size_t n) { return (uint32_t)((n + kBlock - 1) / kBlock); }

} // namespace

int build_bvh_device(const void *verts, const void *tri_idx, uint32_t n, const uint8_t *mat_alpha,
                     const ks_bvh_options &opt, int width, void *stream, DeviceBvh &out, std::string &err) {
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    out = DeviceBvh{};
#define KS_DB(call)                                                                                            \
    do {                                                                                                       \
        hipError_t e_ = (call);                                                                                \
        if (e_ != hipSuccess) {                                                                                \
            (void)hipGetLastError();                                                                           \
            err = std::string("device BVH build: ") + #call + ": " + hipGetErrorString(e_);                   \
            return -1;                                                                                         \
        }                                                                                                      \
    } while (0)
    if (n == 0 || n > kMaxRefs) {
        err = "device BVH build needs 1 .. 2^27 triangles";
        return -1;
    }
    if (width != 2 && width != 4) {
        err = "device BVH build: width must be 2 or 4";
        return -1;
    }
    const uint32_t max_leaf = std::min<uint32_t>(std::max<uint32_t>(opt.max_prims_in_node ? opt.max_prims_in_node : 4, 1),
                                                 kMaxLeafRefs);
    const float trav_cost = opt.trav_weight > 0.0f ? opt.trav_weight : 1.0f;
    const uint32_t ni = n > 1 ? n - 1 : 1; // interior tree nodes
    const size_t node_size = width == 4 ? sizeof(GpuNode4) : sizeof(GpuNode);
    Scratch sc;
    float4 *tri_box, *node_box;
    uint64_t *keys[2];
    uint32_t *ids[2], *acc, *level_count, *count, *clusters[2], *aux, *order, *next_id;
    uint2 *child;
    uint4 *queue[2];
    uint8_t *keep;
    KS_DB(sc.get(tri_box, 2 * (size_t)n));
    KS_DB(sc.get(keys[0], n));
    KS_DB(sc.get(keys[1], n));
    KS_DB(sc.get(ids[0], n));
    KS_DB(sc.get(ids[1], n));
    KS_DB(sc.get(acc, 12));
    KS_DB(sc.get(level_count, kMaxLevels));
    KS_DB(sc.get(node_box, 2 * ((size_t)n + ni)));
    KS_DB(sc.get(child, ni));
    KS_DB(sc.get(count, ni));
    KS_DB(sc.get(clusters[0], n));
    KS_DB(sc.get(clusters[1], n));
    KS_DB(sc.get(aux, n)); // each cluster's neighbour
    KS_DB(sc.get(keep, n));
    KS_DB(sc.get(next_id, 1));
    KS_DB(sc.get(order, n));
    KS_DB(sc.get(queue[0], ni));
    KS_DB(sc.get(queue[1], ni));
    void *nodes = nullptr;
    KS_DB(hipMalloc(&nodes, ni * node_size)); // an output node per interior tree node at most
    sc.ptrs.push_back(nodes);

    const auto t0 = std::chrono::steady_clock::now();
    k_init_acc<<<1, kBlock, 0, s>>>(acc, level_count);
    k_bounds<<<std::min<uint32_t>(blocks_for(n), 2048), kBlock, 0, s>>>(reinterpret_cast<const float4 *>(verts),
                                                                         reinterpret_cast<const uint4 *>(tri_idx), n,
                                                                         tri_box, acc);
    k_morton<<<blocks_for(n), kBlock, 0, s>>>(tri_box, n, acc, keys[0], ids[0]);
    KS_DB(hipGetLastError());
    size_t sort_bytes = 0, select_bytes = 0;
    KS_DB(rocprim::radix_sort_pairs(nullptr, sort_bytes, keys[0], keys[1], ids[0], ids[1], n, 0, 63, s));
    KS_DB(rocprim::select(nullptr, select_bytes, clusters[1], keep, clusters[0], next_id, (size_t)n, s));
    void *temp = nullptr;
    const size_t temp_bytes = std::max<size_t>(std::max(sort_bytes, select_bytes), 16);
    KS_DB(hipMalloc(&temp, temp_bytes));
    sc.ptrs.push_back(temp);
    size_t tb = temp_bytes;
    KS_DB(rocprim::radix_sort_pairs(temp, tb, keys[0], keys[1], ids[0], ids[1], n, 0, 63, s));
    const uint32_t *sids = ids[1];
    k_leaves<<<blocks_for(n), kBlock, 0, s>>>(tri_box, sids, n, node_box, clusters[0]);
    KS_DB(hipGetLastError());
    uint32_t root = 0; // a one-triangle scene: the leaf itself
    uint32_t rounds = 0;
    if (n > 1) {
        KS_DB(hipMemsetAsync(next_id, 0, 4, s));
        uint32_t m = n;
        const int cur = 0; // clusters[0] holds the live clusters, clusters[1] the merged ones before compaction
        while (m > 1) {
            if (++rounds > 4096) {
                err = "device BVH build: clustering did not converge";
                return -1;
            }
            k_ploc_nn<<<blocks_for(m), kBlock, 0, s>>>(clusters[cur], m, node_box, aux);
            k_ploc_merge<<<blocks_for((m + kMergeItems - 1) / kMergeItems), kBlock, 0, s>>>(
                clusters[cur], m, aux, n, next_id, node_box, child, count, clusters[cur ^ 1], keep, max_leaf,
                trav_cost);
            KS_DB(hipGetLastError());
            // compaction keeps the order: clusters[1] -> clusters[0], the new count into level_count[0]
            tb = temp_bytes;
            KS_DB(rocprim::select(temp, tb, clusters[cur ^ 1], keep, clusters[cur], level_count, (size_t)m, s));
            uint32_t mm = 0;
            KS_DB(hipMemcpyAsync(&mm, level_count, 4, hipMemcpyDeviceToHost, s));
            KS_DB(hipStreamSynchronize(s));
            if (mm >= m || mm == 0) {
                err = "device BVH build: a clustering round merged nothing";
                return -1;
            }
            m = mm;
        }
        uint32_t made = 0;
        KS_DB(hipMemcpyAsync(&root, clusters[cur], 4, hipMemcpyDeviceToHost, s));
        KS_DB(hipMemcpyAsync(&made, next_id, 4, hipMemcpyDeviceToHost, s));
        KS_DB(hipStreamSynchronize(s));
        if (made != n - 1 || root != n + made - 1) {
            err = "device BVH build: clustering made an inconsistent tree";
            return -1;
        }
        KS_DB(hipMemsetAsync(level_count, 0, kMaxLevels * sizeof(uint32_t), s));
    }
    // root entry {root, output node 0, first record 0}
    const uint4 root_entry = make_uint4(root, 0u, 0u, 0u);
    KS_DB(hipMemcpyAsync(queue[0], &root_entry, sizeof(uint4), hipMemcpyHostToDevice, s));
    uint32_t qn = 1, base = 1, levels = 0;
    while (qn > 0) {
        if (levels >= (uint32_t)kMaxLevels) {
            err = "device BVH build: tree deeper than the layout supports";
            return -1;
        }
        uint4 *cur = queue[levels & 1], *nxt = queue[(levels + 1) & 1];
        if (width == 4)
            k_layout<4><<<blocks_for(qn), kBlock, 0, s>>>(cur, qn, levels == 0, n, node_box, child, count, sids, nodes,
                                                           order, nxt, level_count + levels, base, ni);
        else
            k_layout<2><<<blocks_for(qn), kBlock, 0, s>>>(cur, qn, levels == 0, n, node_box, child, count, sids, nodes,
                                                           order, nxt, level_count + levels, base, ni);
        KS_DB(hipGetLastError());
        uint32_t next = 0;
        KS_DB(hipMemcpyAsync(&next, level_count + levels, 4, hipMemcpyDeviceToHost, s));
        KS_DB(hipStreamSynchronize(s));
        ++levels;
        if (base + (uint64_t)next > ni) {
            err = "device BVH build: node count exceeds its bound";
            return -1;
        }
        base += next;
        qn = next;
    }
    const uint32_t n_nodes = base;

    void *tris = nullptr;
    KS_DB(hipMalloc(&tris, (size_t)n * sizeof(GpuTri)));
    sc.ptrs.push_back(tris);
    k_records<<<blocks_for(n), kBlock, 0, s>>>(reinterpret_cast<const float4 *>(verts),
                                               reinterpret_cast<const uint4 *>(tri_idx), order, n, mat_alpha,
                                               reinterpret_cast<float4 *>(tris));
    KS_DB(hipGetLastError());
    // the node array shrinks to what the tree uses
    void *exact = nullptr;
    KS_DB(hipMalloc(&exact, (size_t)n_nodes * node_size));
    sc.ptrs.push_back(exact);
    KS_DB(hipMemcpyAsync(exact, nodes, (size_t)n_nodes * node_size, hipMemcpyDeviceToDevice, s));
    uint32_t hacc[12];
    KS_DB(hipMemcpyAsync(hacc, acc, sizeof(hacc), hipMemcpyDeviceToHost, s));
    KS_DB(hipStreamSynchronize(s));
    out.build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    for (int k = 0; k < 3; ++k) {
        out.lo[k] = unordered(hacc[k]);
        out.hi[k] = unordered(hacc[3 + k]);
    }
    // the two results go to the caller; the scratch destructor frees the rest
    sc.release(exact);
    sc.release(tris);
    out.nodes = exact;
    out.node_bytes = (size_t)n_nodes * node_size;
    out.n_nodes = n_nodes;
    out.tris = tris;
    out.tri_bytes = (size_t)n * sizeof(GpuTri);
    out.n_refs = n;
    out.max_depth = levels;
    out.rounds = rounds;
#undef KS_DB
    return 0;
}

} // namespace ksa

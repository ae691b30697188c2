// BVH build on the device: a linear BVH over 63-bit Morton codes of the triangle box centroids
// (Karras, "Maximizing Parallelism in the Construction of BVHs, Octrees, and k-d Trees", HPG 2012),
// a bottom-up pass that fits the boxes and collapses subtrees into leaves where the surface area
// heuristic prefers a leaf, and a top-down pass per level that lays the tree out in the node formats
// the traversal reads (GpuNode4 / GpuNode) with the triangle records in leaf order (GpuTri).
//
// The reference builds its SBVH on the host (cuda/sbvh.cu:301-440); its kd-tree builder
// (cuda/cuda_parallel_kd_tree.cu) is a device pipeline of the same shape: bounds, sort, build,
// layout. This builder has no spatial splits, so each triangle is referenced exactly once.
//
// Kernels (one launch each unless noted), n triangles:
//   k_bounds    scene box and centroid box (block reduction, then ordered-integer atomics), per-
//               triangle boxes (32 B each)
//   k_morton    63-bit Morton key of each centroid, value = triangle id
//   radix sort  (rocPRIM) of the (key, id) pairs over bits 0..62
//   k_hierarchy the n - 1 interior nodes of the radix tree, their ranges and parent links
//   k_fit       one thread per leaf walks to the root; the second arrival at a node fits its box and
//               decides leaf or split by SAH (the first arrival ends)
//   k_layout    one launch per level of the output tree: opens children of largest area until a node
//               holds `W` children, writes the node, queues its interior children for the next level
//   k_records   triangle records in leaf order (vertices, triangle id, alpha-test flag)
#include <hip/hip_runtime.h>
#include <rocprim/device/device_radix_sort.hpp>

#include "ks_internal.h"

#include <algorithm>
#include <chrono>
#include <cstring>
#include <string>
#include <vector>

namespace ksa {
namespace {

constexpr uint32_t kNone = 0xffffffffu;
constexpr uint32_t kRadixLeaf = 0x80000000u; // radix-tree child that is a leaf: low bits = sorted position
constexpr int kBlock = 256;
constexpr int kMaxLevels = 160; // the radix tree over 63-bit keys + 32-bit index ties is at most 96 deep

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

// common prefix length of sorted keys i and j; equal keys are told apart by their positions
__device__ inline int prefix(const uint64_t *__restrict__ k, int64_t n, int64_t i, int64_t j) {
    if (j < 0 || j >= n) return -1;
    uint64_t a = k[i], b = k[j];
    if (a == b) return 64 + __clz((uint32_t)i ^ (uint32_t)j);
    return __clzll(a ^ b);
}

// Interior node i of the radix tree over the n sorted keys: its key range, its split, its children.
__global__ __launch_bounds__(kBlock) void k_hierarchy(const uint64_t *__restrict__ keys, uint32_t n,
                                                      uint32_t *__restrict__ child, uint2 *__restrict__ range,
                                                      uint32_t *__restrict__ parent_leaf,
                                                      uint32_t *__restrict__ parent_node, uint32_t *__restrict__ arrivals) {
    int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= (int64_t)n - 1) return;
    const int64_t nn = n;
    const int d = prefix(keys, nn, i, i + 1) - prefix(keys, nn, i, i - 1) > 0 ? 1 : -1;
    const int dmin = prefix(keys, nn, i, i - d);
    int64_t lmax = 2;
    while (prefix(keys, nn, i, i + lmax * d) > dmin) lmax *= 2;
    int64_t l = 0;
    for (int64_t t = lmax / 2; t >= 1; t /= 2)
        if (prefix(keys, nn, i, i + (l + t) * d) > dmin) l += t;
    const int64_t j = i + l * d;
    const int dnode = prefix(keys, nn, i, j);
    int64_t s = 0;
    for (int64_t t = (l + 1) / 2;; t = (t + 1) / 2) {
        if (prefix(keys, nn, i, i + (s + t) * d) > dnode) s += t;
        if (t == 1) break;
    }
    const int64_t gamma = i + s * d + (d < 0 ? -1 : 0);
    const int64_t first = i < j ? i : j, last = i < j ? j : i;
    uint32_t left = first == gamma ? kRadixLeaf | (uint32_t)gamma : (uint32_t)gamma;
    uint32_t right = last == gamma + 1 ? kRadixLeaf | (uint32_t)(gamma + 1) : (uint32_t)(gamma + 1);
    child[2 * i] = left;
    child[2 * i + 1] = right;
    range[i] = make_uint2((uint32_t)first, (uint32_t)last);
    arrivals[i] = 0u;
    if (left & kRadixLeaf) parent_leaf[left & ~kRadixLeaf] = (uint32_t)i;
    else parent_node[left] = (uint32_t)i;
    if (right & kRadixLeaf) parent_leaf[right & ~kRadixLeaf] = (uint32_t)i;
    else parent_node[right] = (uint32_t)i;
    if (i == 0) parent_node[0] = kNone;
}

struct ChildBox {
    float3 lo, hi;
    float cost;
    bool terminal; // a leaf of the output tree
};

// node_box[2 i] = (lo, SAH cost), node_box[2 i + 1] = (hi, 1 if the subtree becomes one leaf)
__device__ inline ChildBox child_box(uint32_t c, const float4 *__restrict__ tri_box, const uint32_t *__restrict__ ids,
                                     const float4 *node_box, float leaf_cost) {
    ChildBox r;
    if (c & kRadixLeaf) {
        uint32_t t = ids[c & ~kRadixLeaf];
        float4 a = tri_box[2 * (size_t)t], b = tri_box[2 * (size_t)t + 1];
        r.lo = make_float3(a.x, a.y, a.z);
        r.hi = make_float3(b.x, b.y, b.z);
        r.cost = leaf_cost * area_of(r.lo, r.hi);
        r.terminal = true;
    } else {
        float4 a = node_box[2 * (size_t)c], b = node_box[2 * (size_t)c + 1];
        r.lo = make_float3(a.x, a.y, a.z);
        r.hi = make_float3(b.x, b.y, b.z);
        r.cost = a.w;
        r.terminal = b.w != 0.0f;
    }
    return r;
}

// One thread per leaf climbs towards the root. The arrival counter of a node lets the second of its two
// children's threads continue (after both children's boxes are written); the first one stops.
__global__ __launch_bounds__(kBlock) void k_fit(uint32_t n, const float4 *__restrict__ tri_box,
                                                const uint32_t *__restrict__ ids, const uint32_t *__restrict__ child,
                                                const uint2 *__restrict__ range, const uint32_t *__restrict__ parent_leaf,
                                                const uint32_t *__restrict__ parent_node, uint32_t *arrivals,
                                                float4 *node_box, uint32_t max_leaf, float trav_cost) {
    uint32_t p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= n) return;
    const float leaf_cost = 1.0f;
    uint32_t node = parent_leaf[p];
    for (int step = 0; node != kNone && step < kMaxLevels; ++step) { // a radix tree is at most 96 deep
        uint32_t seen = __hip_atomic_fetch_add(arrivals + node, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        if (seen == 0u) return; // the sibling subtree is not finished: its thread continues
        ChildBox a = child_box(child[2 * (size_t)node], tri_box, ids, node_box, leaf_cost);
        ChildBox b = child_box(child[2 * (size_t)node + 1], tri_box, ids, node_box, leaf_cost);
        float3 lo = make_float3(fminf(a.lo.x, b.lo.x), fminf(a.lo.y, b.lo.y), fminf(a.lo.z, b.lo.z));
        float3 hi = make_float3(fmaxf(a.hi.x, b.hi.x), fmaxf(a.hi.y, b.hi.y), fmaxf(a.hi.z, b.hi.z));
        const float area = area_of(lo, hi);
        const uint2 rg = range[node];
        const uint32_t count = rg.y - rg.x + 1;
        const float split = trav_cost * area + a.cost + b.cost;
        const float as_leaf = leaf_cost * area * (float)count;
        const bool leaf = count <= max_leaf && as_leaf <= split;
        node_box[2 * (size_t)node] = make_float4(lo.x, lo.y, lo.z, leaf ? as_leaf : split);
        node_box[2 * (size_t)node + 1] = make_float4(hi.x, hi.y, hi.z, leaf ? 1.0f : 0.0f);
        node = parent_node[node];
    }
}

// One output node per queue entry {radix-tree node, output index}. At the root level the entry may be a
// terminal subtree (or the single leaf of a one-triangle scene): the node then holds that one leaf.
template <int W>
__global__ __launch_bounds__(kBlock) void k_layout(const uint2 *__restrict__ queue, uint32_t count, bool root_level,
                                                   const float4 *__restrict__ tri_box, const uint32_t *__restrict__ ids,
                                                   const uint32_t *__restrict__ child, const uint2 *__restrict__ range,
                                                   const float4 *__restrict__ node_box, void *__restrict__ out_nodes,
                                                   uint2 *__restrict__ next, uint32_t *next_count, uint32_t next_base,
                                                   uint32_t cap) {
    uint32_t qi = blockIdx.x * kBlock + threadIdx.x;
    if (qi >= count) return;
    const uint2 e = queue[qi];
    if (e.y >= cap) return; // cannot happen (an output node per interior radix-tree node); the host reports it
    uint32_t c[W];
    ChildBox cb[W];
    int nc;
    ChildBox self = child_box(e.x, tri_box, ids, node_box, 1.0f);
    if (root_level && self.terminal) {
        c[0] = e.x;
        cb[0] = self;
        nc = 1;
    } else {
        c[0] = child[2 * (size_t)e.x];
        c[1] = child[2 * (size_t)e.x + 1];
        cb[0] = child_box(c[0], tri_box, ids, node_box, 1.0f);
        cb[1] = child_box(c[1], tri_box, ids, node_box, 1.0f);
        nc = 2;
        // open the interior child of largest surface area until the node is full
        while (nc < W) {
            int best = -1;
            float best_area = -1.0f;
            uint32_t g = 0;
#pragma unroll
            for (int k = 0; k < W; ++k)
                if (k < nc && !cb[k].terminal) {
                    float a = area_of(cb[k].lo, cb[k].hi);
                    if (a > best_area) best = k, best_area = a, g = c[k];
                }
            if (best < 0) break;
            const uint32_t g0 = child[2 * (size_t)g], g1 = child[2 * (size_t)g + 1];
            // keep the order of the children: g1 goes right after g0
#pragma unroll
            for (int k = W - 1; k > 0; --k)
                if (k > best + 1 && k <= nc) c[k] = c[k - 1], cb[k] = cb[k - 1];
#pragma unroll
            for (int k = 0; k < W; ++k) {
                if (k == best) c[k] = g0, cb[k] = child_box(g0, tri_box, ids, node_box, 1.0f);
                if (k == best + 1) c[k] = g1, cb[k] = child_box(g1, tri_box, ids, node_box, 1.0f);
            }
            ++nc;
        }
    }
    uint32_t ptr[W];
#pragma unroll
    for (int k = 0; k < W; ++k) {
        ptr[k] = kLeafBit;
        if (k >= nc) continue;
        if (c[k] & kRadixLeaf) {
            ptr[k] = kLeafBit | ((c[k] & ~kRadixLeaf) << 4);
        } else if (cb[k].terminal) {
            const uint2 rg = range[c[k]];
            ptr[k] = kLeafBit | (rg.x << 4) | (rg.y - rg.x);
        } else {
            uint32_t q = atomicAdd(next_count, 1u);
            if (next_base + q < cap) next[q] = make_uint2(c[k], next_base + q);
            ptr[k] = next_base + q;
        }
    }
    const float inf = __builtin_inff();
    if (W == 4) {
        GpuNode4 nd;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const bool used = k < nc;
            nd.lox[k] = used ? cb[k].lo.x : inf;
            nd.hix[k] = used ? cb[k].hi.x : -inf;
            nd.loy[k] = used ? cb[k].lo.y : inf;
            nd.hiy[k] = used ? cb[k].hi.y : -inf;
            nd.loz[k] = used ? cb[k].lo.z : inf;
            nd.hiz[k] = used ? cb[k].hi.z : -inf;
            nd.child[k] = ptr[k];
            nd.pad[k] = 0u;
        }
        reinterpret_cast<GpuNode4 *>(out_nodes)[e.y] = nd;
    } else {
        GpuNode nd;
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const bool used = k < nc;
            nd.f[4 * k] = used ? cb[k].lo.x : inf;
            nd.f[4 * k + 1] = used ? cb[k].hi.x : -inf;
            nd.f[4 * k + 2] = used ? cb[k].lo.y : inf;
            nd.f[4 * k + 3] = used ? cb[k].hi.y : -inf;
            nd.f[8 + 2 * k] = used ? cb[k].lo.z : inf;
            nd.f[9 + 2 * k] = used ? cb[k].hi.z : -inf;
            nd.child[k] = ptr[k];
            nd.pad[k] = 0u;
        }
        reinterpret_cast<GpuNode *>(out_nodes)[e.y] = nd;
    }
}

// triangle records in leaf (sorted) order; v1.w = 1 marks a triangle whose material is alpha tested
__global__ __launch_bounds__(kBlock) void k_records(const float4 *__restrict__ verts, const uint4 *__restrict__ tri_idx,
                                                    const uint32_t *__restrict__ ids, uint32_t n,
                                                    const uint8_t *__restrict__ mat_alpha, float4 *__restrict__ out) {
    uint32_t p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= n) return;
    const uint32_t t = ids[p];
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
        if (e != hipSuccess) return e;
        ptrs.push_back(v);
        p = reinterpret_cast<T *>(v);
        return hipSuccess;
    }
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
    const uint32_t ni = n > 1 ? n - 1 : 1; // interior radix-tree nodes
    const size_t node_size = width == 4 ? sizeof(GpuNode4) : sizeof(GpuNode);
    Scratch sc;
    float4 *tri_box;
    uint64_t *keys[2];
    uint32_t *ids[2], *acc, *level_count, *child, *parent_leaf, *parent_node, *arrivals;
    uint2 *range, *queue[2];
    float4 *node_box;
    KS_DB(sc.get(tri_box, 2 * (size_t)n));
    KS_DB(sc.get(keys[0], n));
    KS_DB(sc.get(keys[1], n));
    KS_DB(sc.get(ids[0], n));
    KS_DB(sc.get(ids[1], n));
    KS_DB(sc.get(acc, 12));
    KS_DB(sc.get(level_count, kMaxLevels));
    KS_DB(sc.get(child, 2 * (size_t)ni));
    KS_DB(sc.get(range, ni));
    KS_DB(sc.get(parent_leaf, n));
    KS_DB(sc.get(parent_node, ni));
    KS_DB(sc.get(arrivals, ni));
    KS_DB(sc.get(node_box, 2 * (size_t)ni));
    KS_DB(sc.get(queue[0], ni));
    KS_DB(sc.get(queue[1], ni));
    // an output node consumes at least one interior radix-tree node (or is the root)
    void *nodes = nullptr;
    KS_DB(hipMalloc(&nodes, ni * node_size));
    sc.ptrs.push_back(nodes);

    const auto t0 = std::chrono::steady_clock::now();
    k_init_acc<<<1, kBlock, 0, s>>>(acc, level_count);
    k_bounds<<<std::min<uint32_t>(blocks_for(n), 2048), kBlock, 0, s>>>(reinterpret_cast<const float4 *>(verts),
                                                                         reinterpret_cast<const uint4 *>(tri_idx), n,
                                                                         tri_box, acc);
    k_morton<<<blocks_for(n), kBlock, 0, s>>>(tri_box, n, acc, keys[0], ids[0]);
    KS_DB(hipGetLastError());
    size_t temp_bytes = 0;
    KS_DB(rocprim::radix_sort_pairs(nullptr, temp_bytes, keys[0], keys[1], ids[0], ids[1], n, 0, 63, s));
    void *temp = nullptr;
    KS_DB(hipMalloc(&temp, std::max<size_t>(temp_bytes, 16)));
    sc.ptrs.push_back(temp);
    KS_DB(rocprim::radix_sort_pairs(temp, temp_bytes, keys[0], keys[1], ids[0], ids[1], n, 0, 63, s));
    const uint64_t *skeys = keys[1];
    const uint32_t *sids = ids[1];
    uint32_t root = 0;
    if (n > 1) {
        k_hierarchy<<<blocks_for(n - 1), kBlock, 0, s>>>(skeys, n, child, range, parent_leaf, parent_node, arrivals);
        k_fit<<<blocks_for(n), kBlock, 0, s>>>(n, tri_box, sids, child, range, parent_leaf, parent_node, arrivals,
                                                node_box, max_leaf, trav_cost);
        KS_DB(hipGetLastError());
    } else {
        root = kRadixLeaf; // the one leaf
    }
    // root entry {root, output node 0}
    const uint2 root_entry = make_uint2(root, 0u);
    KS_DB(hipMemcpyAsync(queue[0], &root_entry, sizeof(uint2), hipMemcpyHostToDevice, s));
    uint32_t count = 1, base = 1, levels = 0;
    while (count > 0) {
        if (levels >= (uint32_t)kMaxLevels) {
            err = "device BVH build: tree deeper than expected";
            return -1;
        }
        uint2 *cur = queue[levels & 1], *nxt = queue[(levels + 1) & 1];
        if (width == 4)
            k_layout<4><<<blocks_for(count), kBlock, 0, s>>>(cur, count, levels == 0, tri_box, sids, child, range,
                                                              node_box, nodes, nxt, level_count + levels, base, ni);
        else
            k_layout<2><<<blocks_for(count), kBlock, 0, s>>>(cur, count, levels == 0, tri_box, sids, child, range,
                                                              node_box, nodes, nxt, level_count + levels, base, ni);
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
        count = next;
    }
    const uint32_t n_nodes = base;

    void *tris = nullptr;
    KS_DB(hipMalloc(&tris, (size_t)n * sizeof(GpuTri)));
    sc.ptrs.push_back(tris);
    k_records<<<blocks_for(n), kBlock, 0, s>>>(reinterpret_cast<const float4 *>(verts),
                                               reinterpret_cast<const uint4 *>(tri_idx), sids, n, mat_alpha,
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
    // hand the two results to the caller; the scratch destructor frees the rest
    sc.ptrs.erase(std::remove(sc.ptrs.begin(), sc.ptrs.end(), exact), sc.ptrs.end());
    sc.ptrs.erase(std::remove(sc.ptrs.begin(), sc.ptrs.end(), tris), sc.ptrs.end());
    out.nodes = exact;
    out.node_bytes = (size_t)n_nodes * node_size;
    out.n_nodes = n_nodes;
    out.tris = tris;
    out.tri_bytes = (size_t)n * sizeof(GpuTri);
    out.n_refs = n;
    out.max_depth = levels;
#undef KS_DB
    return 0;
}

} // namespace ksa

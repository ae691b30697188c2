// Implementation-independent traversal work of a ray sample (host code, measurement only): the node
// fetches and triangle tests of the reference SBVH traversal (SBVH::intersect / intersectBool,
// cuda/sbvh.cu:512-603) over the host-built SBVH in its BVH2 form, with the reference's slab test
// (cuda/aabb.cuh:172-251) and watertight triangle test (cuda/ray_tri.cuh:46-97). The reference stores a
// node's own box in the node; here each box sits in the parent, so a visited node's box is carried on the
// stack. The split axis that picks the near child (sbvh.cu:549-555) is kept by the builder in pad[0].
#include "ks_internal.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <thread>

namespace ksa {
namespace {

struct CountRay {
    float o[3], d[3], tmin, tmax;
    float inv[3];
    int neg[3];
    int kx, ky, kz;
    float sx, sy, sz;
};

void setup(CountRay &r) {
    for (int i = 0; i < 3; ++i) { // RayBoundHelper (cuda/aabb.cuh:155-169)
        r.inv[i] = 1.0f / r.d[i];
        r.neg[i] = r.d[i] < 0.0f || (r.d[i] == 0.0f && std::signbit(r.d[i]));
    }
    // RayTriHelper (cuda/ray_tri.cuh:8-37): the dominant axis becomes z, shear x and y onto it
    float a[3] = {std::fabs(r.d[0]), std::fabs(r.d[1]), std::fabs(r.d[2])};
    r.kz = a[0] >= a[1] && a[0] >= a[2] ? 0 : (a[1] >= a[0] && a[1] >= a[2] ? 1 : 2);
    r.kx = (r.kz + 1) % 3;
    r.ky = (r.kx + 1) % 3;
    if (r.d[r.kz] < 0.0f) std::swap(r.kx, r.ky);
    r.sx = r.d[r.kx] / r.d[r.kz];
    r.sy = r.d[r.ky] / r.d[r.kz];
    r.sz = 1.0f / r.d[r.kz];
}

// isect_ray_aabb (cuda/aabb.cuh:172-251): no rounding margin, as the reference
bool hit_box(const CountRay &r, const float lo[3], const float hi[3]) {
    const float *b[2] = {lo, hi};
    float t0 = b[r.neg[0]][0] - r.o[0];
    if (t0 != 0.0f) t0 *= r.inv[0];
    float t1 = b[1 - r.neg[0]][0] - r.o[0];
    if (t1 != 0.0f) t1 *= r.inv[0];
    float ty0 = b[r.neg[1]][1] - r.o[1];
    if (ty0 != 0.0f) ty0 *= r.inv[1];
    float ty1 = b[1 - r.neg[1]][1] - r.o[1];
    if (ty1 != 0.0f) ty1 *= r.inv[1];
    if (t0 > ty1 || ty0 > t1) return false;
    if (ty0 > t0) t0 = ty0;
    if (ty1 < t1) t1 = ty1;
    float tz0 = b[r.neg[2]][2] - r.o[2];
    if (tz0 != 0.0f) tz0 *= r.inv[2];
    float tz1 = b[1 - r.neg[2]][2] - r.o[2];
    if (tz1 != 0.0f) tz1 *= r.inv[2];
    if (t0 > tz1 || tz0 > t1) return false;
    if (tz0 > t0) t0 = tz0;
    if (tz1 < t1) t1 = tz1;
    return t0 < r.tmax && t1 > r.tmin;
}

// isect_ray_tri (cuda/ray_tri.cuh:46-97) with its FP64 edge fallback; returns the hit distance or -1
float hit_tri(const CountRay &r, const GpuTri &g) {
    const float *v[3] = {g.v0, g.v1, g.v2};
    float A[3][3];
    for (int k = 0; k < 3; ++k) {
        float rel[3] = {v[k][0] - r.o[0], v[k][1] - r.o[1], v[k][2] - r.o[2]};
        A[k][0] = rel[r.kx];
        A[k][1] = rel[r.ky];
        A[k][2] = rel[r.kz];
    }
    float sx[3], sy[3];
    for (int k = 0; k < 3; ++k) {
        sx[k] = A[k][0] - r.sx * A[k][2];
        sy[k] = A[k][1] - r.sy * A[k][2];
    }
    float bc[3] = {sy[1] * sx[2] - sy[2] * sx[1], sy[2] * sx[0] - sy[0] * sx[2], sy[0] * sx[1] - sy[1] * sx[0]};
    if (bc[0] == 0.0f || bc[1] == 0.0f || bc[2] == 0.0f) {
        bc[0] = (float)((double)sx[2] * (double)sy[1] - (double)sy[2] * (double)sx[1]);
        bc[1] = (float)((double)sx[0] * (double)sy[2] - (double)sy[0] * (double)sx[2]);
        bc[2] = (float)((double)sx[1] * (double)sy[0] - (double)sy[1] * (double)sx[0]);
    }
    if ((bc[0] < 0.0f || bc[1] < 0.0f || bc[2] < 0.0f) && (bc[0] > 0.0f || bc[1] > 0.0f || bc[2] > 0.0f))
        return -1.0f;
    float det = bc[0] + bc[1] + bc[2];
    if (det == 0.0f) return -1.0f;
    float T = bc[0] * (r.sz * A[0][2]) + bc[1] * (r.sz * A[1][2]) + bc[2] * (r.sz * A[2][2]);
    if (det < 0.0f && (T >= r.tmin * det || T <= r.tmax * det)) return -1.0f;
    if (det > 0.0f && (T <= r.tmin * det || T >= r.tmax * det)) return -1.0f;
    return T * (1.0f / det);
}

struct Pending {
    uint32_t ptr;
    float lo[3], hi[3];
};

Pending child(const GpuNode &n, int slot) {
    Pending p;
    p.ptr = n.child[slot];
    p.lo[0] = n.f[4 * slot];
    p.hi[0] = n.f[4 * slot + 1];
    p.lo[1] = n.f[4 * slot + 2];
    p.hi[1] = n.f[4 * slot + 3];
    p.lo[2] = n.f[8 + 2 * slot];
    p.hi[2] = n.f[9 + 2 * slot];
    return p;
}

// One ray: SBVH::intersect (any_hit false) or SBVH::intersectBool (true). Returns the hit distance
// (closest) or 1 / -1 (occluded / not), adding one node per fetched node (sbvh.cu:520) and one test per
// candidate triangle.
float traverse(const std::vector<GpuNode> &nodes, const std::vector<GpuTri> &tris, const float lo[3],
               const float hi[3], CountRay r, bool any_hit, uint64_t &n_nodes, uint64_t &n_tris) {
    std::vector<Pending> stack;
    stack.reserve(64);
    Pending cur;
    cur.ptr = 0;
    for (int k = 0; k < 3; ++k) cur.lo[k] = lo[k], cur.hi[k] = hi[k];
    float best = -1.0f;
    while (true) {
        ++n_nodes;
        bool descend = false;
        if (hit_box(r, cur.lo, cur.hi)) {
            if (cur.ptr & kLeafBit) {
                uint32_t first = (cur.ptr & ~kLeafBit) >> 4, count = (cur.ptr & 15u) + 1u;
                for (uint32_t i = 0; i < count; ++i) {
                    ++n_tris;
                    float t = hit_tri(r, tris[first + i]);
                    if (t >= 0.0f && t < r.tmax) {
                        if (any_hit) return 1.0f;
                        r.tmax = t;
                        best = t;
                    }
                }
            } else {
                const GpuNode &n = nodes[cur.ptr];
                Pending left = child(n, 0), right = child(n, 1);
                if (r.neg[n.pad[0] & 3u]) {
                    stack.push_back(left);
                    cur = right;
                } else {
                    stack.push_back(right);
                    cur = left;
                }
                descend = true;
            }
        }
        if (descend) continue;
        if (stack.empty()) break;
        cur = stack.back();
        stack.pop_back();
    }
    return any_hit ? -1.0f : best;
}

} // namespace

void count_reference_traversal(const std::vector<GpuNode> &nodes, const std::vector<GpuTri> &tris, const float lo[3],
                               const float hi[3], const float *rays, int64_t n, bool any_hit, uint64_t &n_nodes,
                               uint64_t &n_tris, float *t_out) {
    std::atomic<uint64_t> total_nodes{0}, total_tris{0};
    const int threads = (int)std::max(1u, std::min(64u, std::thread::hardware_concurrency()));
    std::atomic<int64_t> next{0};
    auto worker = [&]() {
        uint64_t a = 0, b = 0;
        for (;;) {
            int64_t i0 = next.fetch_add(1024);
            if (i0 >= n) break;
            int64_t i1 = std::min<int64_t>(n, i0 + 1024);
            for (int64_t i = i0; i < i1; ++i) {
                const float *q = rays + 8 * i;
                CountRay r;
                for (int k = 0; k < 3; ++k) r.o[k] = q[k], r.d[k] = q[4 + k];
                r.tmin = q[3];
                r.tmax = q[7];
                setup(r);
                float t = traverse(nodes, tris, lo, hi, r, any_hit, a, b);
                if (t_out) t_out[i] = t;
            }
        }
        total_nodes += a;
        total_tris += b;
    };
    std::vector<std::thread> pool;
    for (int k = 1; k < threads; ++k) pool.emplace_back(worker);
    worker();
    for (auto &t : pool) t.join();
    n_nodes = total_nodes.load();
    n_tris = total_tris.load();
}

} // namespace ksa

// BVH2 and BVH4 traversal for gfx950: 64-byte nodes holding both child boxes (or 128-byte nodes holding
// four), leaves pointing into a 48-byte
// per-reference triangle array, the watertight ray/triangle test of cuda/ray_tri.cuh:46-97 and a
// per-lane short stack that lives in LDS (kStackLds entries, stored with the block size as stride so
// a wave's 64 lanes hit 64 distinct banks) and spills its oldest entries to a global overflow area.
#pragma once
#include "ks_device.h"

namespace ksd {

constexpr uint32_t kLeafBit = 0x80000000u;
constexpr int kStackLds = 16;   // LDS entries per lane (power of two)
constexpr int kTraceBlock = 256; // threads per traversal block
// trace_rays_kind: below this many lanes still searching for a leaf, a wave with a finished lane stops
// its inner-node loop to test leaves and refill finished lanes (on the 1080p sphere scene 8-16 measured
// best: traversal 35.4 -> 31.8 ms per render against waiting for every lane; with the second postponed
// leaf, 8 gave 28.5 ms, 12 28.6 and 16 28.8)
constexpr int kTraverseBusyMin = 8;

struct TravRay {
    f3 o, d;
    f3 inv, oinv;
    // bits 0-1 kx, 2-3 ky, 4-5 kz (RayTriHelper, cuda/ray_tri.cuh:8-37); bits 6, 7, 8: direction sign of x,
    // y, z (RayBoundHelper::dir_is_neg, cuda/aabb.cuh:155-169). One register for all six keeps the
    // traversal loops within 64 VGPRs.
    uint32_t axes;
    float sx, sy, sz;
    float tmin;
    KS_DEV int kx() const { return (int)(axes & 3u); }
    KS_DEV int ky() const { return (int)((axes >> 2) & 3u); }
    KS_DEV int kz() const { return (int)((axes >> 4) & 3u); }
    KS_DEV bool nx() const { return (axes >> 6) & 1u; }
    KS_DEV bool ny() const { return (axes >> 7) & 1u; }
    KS_DEV bool nz() const { return (axes >> 8) & 1u; }
};

KS_DEV float comp(f3 v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }

KS_DEV TravRay make_trav_ray(f3 o, f3 d, float tmin) {
    TravRay r;
    r.o = o;
    r.d = d;
    r.tmin = tmin;
    // zero direction components are replaced by a signed tiny value for the slab test only, so
    // (b - o) * inv never forms 0 * inf
    const float eps = 1e-20f;
    float dx = fabsf(d.x) < eps ? copysignf(eps, d.x) : d.x;
    float dy = fabsf(d.y) < eps ? copysignf(eps, d.y) : d.y;
    float dz = fabsf(d.z) < eps ? copysignf(eps, d.z) : d.z;
    r.inv = mk(1.0f / dx, 1.0f / dy, 1.0f / dz);
    r.oinv = mk(o.x * r.inv.x, o.y * r.inv.y, o.z * r.inv.z);
    const uint32_t signs = (signbit(d.x) ? 1u : 0u) | (signbit(d.y) ? 2u : 0u) | (signbit(d.z) ? 4u : 0u);
    float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
    int kz = (ax >= ay && ax >= az) ? 0 : ((ay >= ax && ay >= az) ? 1 : 2);
    int kx = kz == 2 ? 0 : kz + 1;
    int ky = kx == 2 ? 0 : kx + 1;
    float dkz = comp(d, kz);
    if (dkz < 0.0f) {
        int t = kx;
        kx = ky;
        ky = t;
    }
    r.axes = (uint32_t)kx | ((uint32_t)ky << 2) | ((uint32_t)kz << 4) | (signs << 6);
    r.sx = comp(d, kx) / dkz;
    r.sy = comp(d, ky) / dkz;
    r.sz = 1.0f / dkz;
    return r;
}

// Watertight test (cuda/ray_tri.cuh:46-97); on a hit within (tmin, tmax) returns t and the weights of
// v1 and v2 (Embree's u, v).
KS_DEV bool isect_tri(const TravRay &r, float4 a4, float4 b4, float4 c4, float tmax, float &t_out, float &u_out,
                      float &v_out) {
    // No FMA contraction: the edge functions of two triangles sharing an edge must be exact negatives
    // of each other for the test to stay watertight.
#pragma clang fp contract(off)
    f3 A0 = mk(a4.x - r.o.x, a4.y - r.o.y, a4.z - r.o.z);
    f3 B0 = mk(b4.x - r.o.x, b4.y - r.o.y, b4.z - r.o.z);
    f3 C0 = mk(c4.x - r.o.x, c4.y - r.o.y, c4.z - r.o.z);
    const int kx = r.kx(), ky = r.ky(), kz = r.kz();
    float Ax = comp(A0, kx), Ay = comp(A0, ky), Az = comp(A0, kz);
    float Bx = comp(B0, kx), By = comp(B0, ky), Bz = comp(B0, kz);
    float Cx = comp(C0, kx), Cy = comp(C0, ky), Cz = comp(C0, kz);
    float sxA = Ax - r.sx * Az, sxB = Bx - r.sx * Bz, sxC = Cx - r.sx * Cz;
    float syA = Ay - r.sy * Az, syB = By - r.sy * Bz, syC = Cy - r.sy * Cz;
    // bc = cross(sy, sx)
    float U = syB * sxC - syC * sxB;
    float V = syC * sxA - syA * sxC;
    float W = syA * sxB - syB * sxA;
    if (U == 0.0f || V == 0.0f || W == 0.0f) {
        U = (float)((double)sxC * (double)syB - (double)syC * (double)sxB);
        V = (float)((double)sxA * (double)syC - (double)syA * (double)sxC);
        W = (float)((double)sxB * (double)syA - (double)syB * (double)sxA);
    }
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
    float det = U + V + W;
    if (det == 0.0f) return false;
    float T = U * (r.sz * Az) + V * (r.sz * Bz) + W * (r.sz * Cz);
    if (det < 0.0f && (T >= r.tmin * det || T <= tmax * det)) return false;
    if (det > 0.0f && (T <= r.tmin * det || T >= tmax * det)) return false;
    float inv_det = 1.0f / det;
    t_out = T * inv_det;
    u_out = V * inv_det;
    v_out = W * inv_det;
    return true;
}

struct TravStack {
    uint32_t *lds;  // this lane's column: entry i at lds[i * kTraceBlock]
    uint32_t *ovf;  // this lane's spill column: entry i at ovf[i * ovf_stride]
    uint32_t ovf_stride;
    uint32_t ovf_cap; // spill entries available per lane
    int sp;
    uint32_t spills;

    KS_DEV void push(uint32_t v) {
        int slot = sp & (kStackLds - 1);
        if (sp >= kStackLds) {
            // the slot holds entry sp - kStackLds: move it to the spill area first
            uint32_t a = (uint32_t)(sp - kStackLds);
            if (a < ovf_cap) ovf[a * ovf_stride] = lds[slot * kTraceBlock];
            ++spills;
        }
        lds[slot * kTraceBlock] = v;
        ++sp;
    }
    KS_DEV uint32_t pop() {
        --sp;
        int slot = sp & (kStackLds - 1);
        uint32_t v = lds[slot * kTraceBlock];
        if (sp >= kStackLds) {
            uint32_t a = (uint32_t)(sp - kStackLds);
            lds[slot * kTraceBlock] = a < ovf_cap ? ovf[a * ovf_stride] : 0u;
        }
        return v;
    }
};

struct TravHit {
    float t, u, v;
    int32_t prim;
};

// Closest hit (ANY = false) or any hit (ANY = true) in (tmin, tmax). Children are visited nearest
// entry distance first. Counts node fetches and triangle tests for the byte model.
template <bool ANY>
KS_DEV bool traverse(const float4 *__restrict__ nodes, const float4 *__restrict__ tris, const TravRay &r, float tmax,
                     TravStack &st, TravHit &hit, uint32_t &n_nodes, uint32_t &n_tris) {
    const float robust = 1.0000004f; // 1 + 2*gamma(3): conservative far distance (Ize 2013)
    hit.prim = -1;
    st.sp = 0;
    uint32_t node = 0;
    while (true) {
        bool exhausted = false;
        while ((node & kLeafBit) == 0) {
            ++n_nodes;
            const float4 *np = nodes + 4 * (size_t)node;
            float4 a = np[0], b = np[1], c = np[2];
            uint2 ch = *reinterpret_cast<const uint2 *>(np + 3);
            // child 0: x (a.x, a.y) y (a.z, a.w) z (c.x, c.y); child 1: x (b.x, b.y) y (b.z, b.w) z (c.z, c.w)
            float n0x = fmaf(r.nx() ? a.y : a.x, r.inv.x, -r.oinv.x), f0x = fmaf(r.nx() ? a.x : a.y, r.inv.x, -r.oinv.x);
            float n0y = fmaf(r.ny() ? a.w : a.z, r.inv.y, -r.oinv.y), f0y = fmaf(r.ny() ? a.z : a.w, r.inv.y, -r.oinv.y);
            float n0z = fmaf(r.nz() ? c.y : c.x, r.inv.z, -r.oinv.z), f0z = fmaf(r.nz() ? c.x : c.y, r.inv.z, -r.oinv.z);
            float n1x = fmaf(r.nx() ? b.y : b.x, r.inv.x, -r.oinv.x), f1x = fmaf(r.nx() ? b.x : b.y, r.inv.x, -r.oinv.x);
            float n1y = fmaf(r.ny() ? b.w : b.z, r.inv.y, -r.oinv.y), f1y = fmaf(r.ny() ? b.z : b.w, r.inv.y, -r.oinv.y);
            float n1z = fmaf(r.nz() ? c.w : c.z, r.inv.z, -r.oinv.z), f1z = fmaf(r.nz() ? c.z : c.w, r.inv.z, -r.oinv.z);
            float t0n = fmaxf(fmaxf(n0x, n0y), fmaxf(n0z, r.tmin));
            float t0f = fminf(fminf(f0x, f0y), fminf(f0z, tmax)) * robust;
            float t1n = fmaxf(fmaxf(n1x, n1y), fmaxf(n1z, r.tmin));
            float t1f = fminf(fminf(f1x, f1y), fminf(f1z, tmax)) * robust;
            bool h0 = t0n <= t0f, h1 = t1n <= t1f;
            if (h0 && h1) {
                bool first0 = t0n <= t1n;
                st.push(first0 ? ch.y : ch.x);
                node = first0 ? ch.x : ch.y;
            } else if (h0) {
                node = ch.x;
            } else if (h1) {
                node = ch.y;
            } else {
                if (st.sp == 0) {
                    exhausted = true;
                    break;
                }
                node = st.pop();
            }
        }
        if (exhausted) break;
        uint32_t first = (node >> 4) & 0x7ffffffu;
        uint32_t cnt = (node & 15u) + 1u;
        for (uint32_t k = 0; k < cnt; ++k) {
            ++n_tris;
            const float4 *tp = tris + 3 * (size_t)(first + k);
            float4 v0 = tp[0], v1 = tp[1], v2 = tp[2];
            float t, u, v;
            if (isect_tri(r, v0, v1, v2, tmax, t, u, v)) {
                if (ANY) {
                    hit.prim = __float_as_int(v0.w);
                    return true;
                }
                tmax = t;
                hit.t = t;
                hit.u = u;
                hit.v = v;
                hit.prim = __float_as_int(v0.w);
            }
        }
        if (st.sp == 0) break;
        node = st.pop();
    }
    return hit.prim >= 0;
}

KS_DEV float lane4(float4 v, int i) { return i == 0 ? v.x : (i == 1 ? v.y : (i == 2 ? v.z : v.w)); }
KS_DEV uint32_t lane4(uint4 v, int i) { return i == 0 ? v.x : (i == 1 ? v.y : (i == 2 ? v.z : v.w)); }

// The same query over BVH4 nodes (GpuNode4: 6 float4 of SoA child boxes, then uint4 children). The
// boxes that are hit are ordered by entry distance with a 5-comparator network; the nearest is
// visited next and the others are pushed farthest first. One node fetch = 112 bytes.
template <bool ANY>
KS_DEV bool traverse4(const float4 *__restrict__ nodes, const float4 *__restrict__ tris, const TravRay &r, float tmax,
                      TravStack &st, TravHit &hit, uint32_t &n_nodes, uint32_t &n_tris) {
    const float robust = 1.0000004f; // 1 + 2*gamma(3): conservative far distance (Ize 2013)
    hit.prim = -1;
    st.sp = 0;
    uint32_t node = 0;
    while (true) {
        bool exhausted = false;
        while ((node & kLeafBit) == 0) {
            ++n_nodes;
            const float4 *np = nodes + 8 * (size_t)node;
            const float4 lx = np[0], hx = np[1], ly = np[2], hy = np[3], lz = np[4], hz = np[5];
            const uint4 ch = *reinterpret_cast<const uint4 *>(np + 6);
            float key[4];
            uint32_t cp[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                float nx = fmaf(r.nx() ? lane4(hx, i) : lane4(lx, i), r.inv.x, -r.oinv.x);
                float fx = fmaf(r.nx() ? lane4(lx, i) : lane4(hx, i), r.inv.x, -r.oinv.x);
                float ny = fmaf(r.ny() ? lane4(hy, i) : lane4(ly, i), r.inv.y, -r.oinv.y);
                float fy = fmaf(r.ny() ? lane4(ly, i) : lane4(hy, i), r.inv.y, -r.oinv.y);
                float nz = fmaf(r.nz() ? lane4(hz, i) : lane4(lz, i), r.inv.z, -r.oinv.z);
                float fz = fmaf(r.nz() ? lane4(lz, i) : lane4(hz, i), r.inv.z, -r.oinv.z);
                float tn = fmaxf(fmaxf(nx, ny), fmaxf(nz, r.tmin));
                float tf = fminf(fminf(fx, fy), fminf(fz, tmax)) * robust;
                key[i] = tn <= tf ? tn : INFINITY;
                cp[i] = lane4(ch, i);
            }
            auto cswap = [&](int a, int b) {
                bool sw = key[b] < key[a];
                float ka = key[a], kb = key[b];
                uint32_t ca = cp[a], cb = cp[b];
                key[a] = sw ? kb : ka;
                key[b] = sw ? ka : kb;
                cp[a] = sw ? cb : ca;
                cp[b] = sw ? ca : cb;
            };
            cswap(0, 1);
            cswap(2, 3);
            cswap(0, 2);
            cswap(1, 3);
            cswap(1, 2);
            if (key[0] == INFINITY) {
                if (st.sp == 0) {
                    exhausted = true;
                    break;
                }
                node = st.pop();
            } else {
                if (key[3] != INFINITY) st.push(cp[3]);
                if (key[2] != INFINITY) st.push(cp[2]);
                if (key[1] != INFINITY) st.push(cp[1]);
                node = cp[0];
            }
        }
        if (exhausted) break;
        uint32_t first = (node >> 4) & 0x7ffffffu;
        uint32_t cnt = (node & 15u) + 1u;
        for (uint32_t k = 0; k < cnt; ++k) {
            ++n_tris;
            const float4 *tp = tris + 3 * (size_t)(first + k);
            float4 v0 = tp[0], v1 = tp[1], v2 = tp[2];
            float t, u, v;
            if (isect_tri(r, v0, v1, v2, tmax, t, u, v)) {
                if (ANY) {
                    hit.prim = __float_as_int(v0.w);
                    return true;
                }
                tmax = t;
                hit.t = t;
                hit.u = u;
                hit.v = v;
                hit.prim = __float_as_int(v0.w);
            }
        }
        if (st.sp == 0) break;
        node = st.pop();
    }
    return hit.prim >= 0;
}

// One visit of an inner node of a BVH of width W (2 or 4), with the same child test and order as
// traverse / traverse4: returns the next node to visit; when no child is hit it pops the stack, and
// sets `done` when the stack is empty.
template <int W>
KS_DEV uint32_t visit_node(const float4 *__restrict__ nodes, const TravRay &r, float tmax, TravStack &st,
                           uint32_t node, bool &done) {
    const float robust = 1.0000004f;
    if (W == 2) {
        const float4 *np = nodes + 4 * (size_t)node;
        float4 a = np[0], b = np[1], c = np[2];
        uint2 ch = *reinterpret_cast<const uint2 *>(np + 3);
        float n0x = fmaf(r.nx() ? a.y : a.x, r.inv.x, -r.oinv.x), f0x = fmaf(r.nx() ? a.x : a.y, r.inv.x, -r.oinv.x);
        float n0y = fmaf(r.ny() ? a.w : a.z, r.inv.y, -r.oinv.y), f0y = fmaf(r.ny() ? a.z : a.w, r.inv.y, -r.oinv.y);
        float n0z = fmaf(r.nz() ? c.y : c.x, r.inv.z, -r.oinv.z), f0z = fmaf(r.nz() ? c.x : c.y, r.inv.z, -r.oinv.z);
        float n1x = fmaf(r.nx() ? b.y : b.x, r.inv.x, -r.oinv.x), f1x = fmaf(r.nx() ? b.x : b.y, r.inv.x, -r.oinv.x);
        float n1y = fmaf(r.ny() ? b.w : b.z, r.inv.y, -r.oinv.y), f1y = fmaf(r.ny() ? b.z : b.w, r.inv.y, -r.oinv.y);
        float n1z = fmaf(r.nz() ? c.w : c.z, r.inv.z, -r.oinv.z), f1z = fmaf(r.nz() ? c.z : c.w, r.inv.z, -r.oinv.z);
        float t0n = fmaxf(fmaxf(n0x, n0y), fmaxf(n0z, r.tmin));
        float t0f = fminf(fminf(f0x, f0y), fminf(f0z, tmax)) * robust;
        float t1n = fmaxf(fmaxf(n1x, n1y), fmaxf(n1z, r.tmin));
        float t1f = fminf(fminf(f1x, f1y), fminf(f1z, tmax)) * robust;
        bool h0 = t0n <= t0f, h1 = t1n <= t1f;
        if (h0 && h1) {
            bool first0 = t0n <= t1n;
            st.push(first0 ? ch.y : ch.x);
            return first0 ? ch.x : ch.y;
        }
        if (h0) return ch.x;
        if (h1) return ch.y;
    } else {
        // near / far planes picked by address from the ray's direction signs: (lo, hi) of an axis are
        // adjacent float4s, so the sign selects the load and no per-child select is needed
        const char *base = reinterpret_cast<const char *>(nodes);
        const uint32_t off = node * 128u; // node arrays stay below 4 GiB (2^25 nodes)
        const uint32_t ox = (r.axes >> 2) & 16u, oy = (r.axes >> 3) & 16u, oz = (r.axes >> 4) & 16u;
        auto ld = [&](uint32_t o) { return *reinterpret_cast<const float4 *>(base + o); };
        const float4 nxv = ld(off + ox), fxv = ld(off + 16u - ox), nyv = ld(off + 32u + oy),
                     fyv = ld(off + 48u - oy), nzv = ld(off + 64u + oz), fzv = ld(off + 80u - oz);
        const uint4 ch = *reinterpret_cast<const uint4 *>(base + off + 96u);
        float key[4];
        uint32_t cp[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            float nx = fmaf(lane4(nxv, i), r.inv.x, -r.oinv.x);
            float fx = fmaf(lane4(fxv, i), r.inv.x, -r.oinv.x);
            float ny = fmaf(lane4(nyv, i), r.inv.y, -r.oinv.y);
            float fy = fmaf(lane4(fyv, i), r.inv.y, -r.oinv.y);
            float nz = fmaf(lane4(nzv, i), r.inv.z, -r.oinv.z);
            float fz = fmaf(lane4(fzv, i), r.inv.z, -r.oinv.z);
            float tn = fmaxf(fmaxf(nx, ny), fmaxf(nz, r.tmin));
            float tf = fminf(fminf(fx, fy), fminf(fz, tmax)) * robust;
            key[i] = tn <= tf ? tn : INFINITY;
            cp[i] = lane4(ch, i);
        }
        auto cswap = [&](int a, int b) {
            bool sw = key[b] < key[a];
            float ka = key[a], kb = key[b];
            uint32_t ca = cp[a], cb = cp[b];
            key[a] = sw ? kb : ka;
            key[b] = sw ? ka : kb;
            cp[a] = sw ? cb : ca;
            cp[b] = sw ? ca : cb;
        };
        cswap(0, 1);
        cswap(2, 3);
        cswap(0, 2);
        cswap(1, 3);
        cswap(1, 2);
        if (key[0] != INFINITY) {
            if (key[3] != INFINITY) st.push(cp[3]);
            if (key[2] != INFINITY) st.push(cp[2]);
            if (key[1] != INFINITY) st.push(cp[1]);
            return cp[0];
        }
    }
    if (st.sp == 0) {
        done = true;
        return 0;
    }
    return st.pop();
}

// Rays first, first + stride, ... < n on one lane, with replacement: a lane whose ray has finished
// stores its result and loads its next ray at once, while the other lanes of the wave go on with
// theirs. The wave then never waits for its slowest ray before fetching more work. Leaves are found
// speculatively: a lane that reaches a leaf sets it aside and keeps visiting inner nodes until every
// lane of the wave holds a leaf (or has none left), and then all lanes test their leaves together
// (the persistent speculative while-while loop of Aila and Laine, HPG 2009). A lane may set aside a
// second leaf while it waits, rather than idle on it (measured: traversal 29.7 -> 28.5 ms per 1080p
// render); its triangles are tested after the first leaf's, in traversal order. The inner-node loop
// also ends early when few lanes are still searching and some lane has finished, so that finished
// lanes do not idle through the others' searches before they take a new ray. Each ray's closest hit is
// the one traverse_w gives; among triangles hit at exactly the same t, the first one tested wins.
// load(i, ray) fills the ray and returns its tmax; store(i, hit) takes the result (prim < 0: miss).
// With ALPHA, a triangle whose record carries a non-zero v1.w is an alpha-tested candidate: filter(prim,
// u, v, ray) decides whether the hit counts (Embree's intersection filter, scene.cpp:175-204). A hit it
// rejects is skipped and traversal goes on, for closest and any hits alike.
struct NoFilter {
    KS_DEV bool operator()(int32_t, float, float, const TravRay &) const { return true; }
};

// The query kind of ray i: any_hit(i) true stops at the first accepted hit (occlusion), false finds the
// closest. trace_rays takes it per ray, so one persistent launch can serve both queues.
template <bool ANY> struct FixedKind {
    KS_DEV bool operator()(uint32_t) const { return ANY; }
};

template <int W, bool ALPHA = false, typename Kind, typename Load, typename Store, typename Filter = NoFilter>
KS_DEV void trace_rays_kind(const float4 *__restrict__ nodes, const float4 *__restrict__ tris, uint32_t i, uint32_t n,
                            uint32_t stride, TravStack &st, Kind any_hit, Load &&load, Store &&store,
                            uint32_t &n_nodes, uint32_t &n_tris, uint32_t &n_rays, Filter filter = Filter()) {
    if (i >= n) return;
    TravRay r;
    TravHit hit;
    float tmax = load(i, r);
    hit.prim = -1;
    st.sp = 0;
    uint32_t node = 0;       // next node to visit (valid while !exhausted)
    uint32_t leaf = 0;       // leaf set aside for testing (0: none; leaves carry kLeafBit)
    uint32_t leaf2 = 0;      // a second leaf set aside while the first waits
    bool exhausted = false;  // no node left to visit for this ray
    while (true) {
        // inner nodes, until every lane of the wave holds a leaf or has nothing left to visit
        while (true) {
            if (!exhausted && (node & kLeafBit) == 0) {
                ++n_nodes;
                bool empty = false;
                node = visit_node<W>(nodes, r, tmax, st, node, empty);
                exhausted = empty;
            }
            if (!exhausted && (node & kLeafBit) != 0 && leaf2 == 0) {
                if (leaf == 0) leaf = node;
                else leaf2 = node;
                if (st.sp == 0) exhausted = true;
                else node = st.pop();
            }
            // leave for the leaf tests when no lane is left without a leaf, or, so that finished lanes
            // take new rays sooner, when fewer than kTraverseBusyMin lanes are still looking for one while
            // some lane has finished its ray
            const uint64_t busy = __ballot(leaf == 0 && !exhausted);
            if (busy == 0) break;
            if (__popcll(busy) < kTraverseBusyMin && __any(exhausted)) break;
        }
        bool done = false;
        if (leaf != 0) {
            const uint32_t first = (leaf >> 4) & 0x7ffffffu;
            const uint32_t cnt = (leaf & 15u) + 1u;
            const uint32_t first2 = (leaf2 >> 4) & 0x7ffffffu;
            const uint32_t cnt2 = leaf2 != 0 ? (leaf2 & 15u) + 1u : 0u;
            leaf = 0;
            leaf2 = 0;
            const bool any = any_hit(i);
            for (uint32_t k = 0; k < cnt + cnt2; ++k) {
                ++n_tris;
                const float4 *tp = tris + 3 * (size_t)(k < cnt ? first + k : first2 + (k - cnt));
                float4 v0 = tp[0], v1 = tp[1], v2 = tp[2];
                float t, u, v;
                if (isect_tri(r, v0, v1, v2, tmax, t, u, v)) {
                    if (ALPHA && __float_as_uint(v1.w) != 0u && !filter(__float_as_int(v0.w), u, v, r)) continue;
                    hit.prim = __float_as_int(v0.w);
                    if (any) {
                        done = true;
                        break;
                    }
                    tmax = t;
                    hit.t = t;
                    hit.u = u;
                    hit.v = v;
                }
            }
        }
        if (done || exhausted) {
            store(i, hit);
            ++n_rays;
            i += stride;
            if (i >= n) break;
            tmax = load(i, r);
            hit.prim = -1;
            st.sp = 0;
            node = 0;
            leaf = 0;
            leaf2 = 0;
            exhausted = false;
        }
    }
}

template <bool ANY, int W, bool ALPHA = false, typename Load, typename Store, typename Filter = NoFilter>
KS_DEV void trace_rays(const float4 *__restrict__ nodes, const float4 *__restrict__ tris, uint32_t i, uint32_t n,
                       uint32_t stride, TravStack &st, Load &&load, Store &&store, uint32_t &n_nodes,
                       uint32_t &n_tris, uint32_t &n_rays, Filter filter = Filter()) {
    trace_rays_kind<W, ALPHA>(nodes, tris, i, n, stride, st, FixedKind<ANY>(), load, store, n_nodes, n_tris, n_rays,
                              filter);
}

// Dispatch on the node width of the scene (2 or 4).
template <bool ANY, int W>
KS_DEV bool traverse_w(const float4 *__restrict__ nodes, const float4 *__restrict__ tris, const TravRay &r, float tmax,
                       TravStack &st, TravHit &hit, uint32_t &n_nodes, uint32_t &n_tris) {
    if (W == 4) return traverse4<ANY>(nodes, tris, r, tmax, st, hit, n_nodes, n_tris);
    return traverse<ANY>(nodes, tris, r, tmax, st, hit, n_nodes, n_tris);
}

} // namespace ksd

// Surface evaluation of textured scenes on the device: hit texture coordinates and the uv-based tangent
// frame (MeshGeometry::compute_intersection, geometry.cpp:68-163), uv partials from the camera ray's
// differentials (Intersection::compute_uv_partials, ray.cpp:6-46), normal maps (NormalMap::apply,
// normal_map.cpp:8-40), the principled closure from shader fields (PrincipledBSDF::eval_closure,
// principled_bsdf.cpp:246-275) and the stochastic opacity test of the traversal filter
// (filter_opacity_map, scene.cpp:175-204).
#pragma once
#include "ks_render_types.h"
#include "ks_texture.h"
#include "ks_traverse.h"

namespace ksr {
using namespace ksd;

constexpr uint32_t kGeomNoTexcoord = 4u;

struct SurfUV {
    float u, v;     // Intersection::uv: texture coordinates, or the barycentric (u, v) without them
    f3 dpdu, dpdv;  // after the two-sided flip of dpdu
    float d[4];     // dudx, dvdx, dudy, dvdy (zero except on camera rays)
};

KS_DEV TexContext tex_context(const DevScene &sc) { return TexContext{sc.textures, sc.tex_luts}; }

KS_DEV bool geom_has_uv(const DevScene &sc, uint4 tri) {
    return sc.uvs != nullptr && !((tri.w >> 24) & kGeomNoTexcoord);
}

// rtcInterpolate0 of the texcoord attribute at barycentric (u, v), as the hit position is interpolated
KS_DEV void interp_uv(const DevScene &sc, uint4 tri, float bu, float bv, float &u, float &v) {
    const float2 t0 = sc.uvs[tri.x], t1 = sc.uvs[tri.y], t2 = sc.uvs[tri.z];
    const float w = 1.0f - bu - bv;
    u = t0.x * w + t1.x * bu + t2.x * bv;
    v = t0.y * w + t1.y * bu + t2.y * bv;
}

// make_hit with texture coordinates: where the geometry has them, uv comes from them and dpdu / dpdv
// are rebuilt from the uv parameterisation (geometry.cpp:94-116), so the tangent frames follow dpdu.
KS_DEV SurfHit make_hit_uv(const DevScene &sc, uint4 tri, float bu, float bv, f3 rd, SurfUV &su) {
    f3 v0 = ld3(sc.verts[tri.x]), v1 = ld3(sc.verts[tri.y]), v2 = ld3(sc.verts[tri.z]);
    f3 wo = -normalized(rd);
    f3 ng = normalized(cross(v1 - v0, v2 - v0));
    float w = 1.0f - bu - bv;
    SurfHit it;
    it.p = v0 * w + v1 * bu + v2 * bv;
    f3 dpdu = v1 - v0, dpdv = v2 - v0;
    su.u = bu;
    su.v = bv;
    if (geom_has_uv(sc, tri)) {
        const float2 t0 = sc.uvs[tri.x], t1 = sc.uvs[tri.y], t2 = sc.uvs[tri.z];
        su.u = t0.x * w + t1.x * bu + t2.x * bv;
        su.v = t0.y * w + t1.y * bu + t2.y * bv;
        const float a0 = t0.x - t2.x, a1 = t0.y - t2.y; // duv02
        const float b0 = t1.x - t2.x, b1 = t1.y - t2.y; // duv12
        const f3 dp02 = v0 - v2, dp12 = v1 - v2;
        const float det = a0 * b1 - a1 * b0;
        if (!(fabsf(det) < 1e-8f)) {
            const float inv = 1.0f / det;
            dpdu = (dp02 * b1 - dp12 * a1) * inv;
            dpdv = (dp02 * -b0 + dp12 * a0) * inv;
        }
    }
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
        f3 vn = normalized(ld3(sc.normals[tri.x]) * w + ld3(sc.normals[tri.y]) * bu + ld3(sc.normals[tri.z]) * bv);
        if (all_finite(vn)) {
            if (dot(vn, ng) < 0.0f) vn = -vn;
            f3 vb = normalized(cross(vn, dpdu));
            it.st = normalized(cross(vb, vn));
            it.sb = vb;
            it.sn = vn;
        }
    }
    su.dpdu = dpdu;
    su.dpdv = dpdv;
    su.d[0] = su.d[1] = su.d[2] = su.d[3] = 0.0f;
    return it;
}

// solve_linear_system_2x2 (maths.h:559-569)
KS_DEV bool solve2x2(float a00, float a01, float a10, float a11, float b0, float b1, float &x0, float &x1) {
    const float det = a00 * a11 - a01 * a10;
    if (fabsf(det) < 1e-10f) return false;
    x0 = (a11 * b0 - a01 * b1) / det;
    x1 = (a00 * b1 - a10 * b0) / det;
    return !(isnan(x0) || isnan(x1));
}

// Intersection::compute_uv_partials (ray.cpp:6-46) for a pinhole camera ray (both differential rays
// start at the ray origin o).
KS_DEV void uv_partials(const SurfHit &it, SurfUV &su, f3 o, f3 rxd, f3 ryd) {
    su.d[0] = su.d[1] = su.d[2] = su.d[3] = 0.0f;
    const float dd = dot(it.n, it.p);
    const float tx = -(dot(it.n, o) - dd) / dot(it.n, rxd);
    if (isinf(tx) || isnan(tx)) return;
    const f3 px = o + rxd * tx;
    const float ty = -(dot(it.n, o) - dd) / dot(it.n, ryd);
    if (isinf(ty) || isnan(ty)) return;
    const f3 py = o + ryd * ty;
    const float nx = fabsf(it.n.x), ny = fabsf(it.n.y), nz = fabsf(it.n.z);
    int d0, d1;
    if (nx > ny && nx > nz) {
        d0 = 1;
        d1 = 2;
    } else if (ny > nz) {
        d0 = 0;
        d1 = 2;
    } else {
        d0 = 0;
        d1 = 1;
    }
    const float a00 = comp(su.dpdu, d0), a01 = comp(su.dpdv, d0), a10 = comp(su.dpdu, d1), a11 = comp(su.dpdv, d1);
    if (!solve2x2(a00, a01, a10, a11, comp(px, d0) - comp(it.p, d0), comp(px, d1) - comp(it.p, d1), su.d[0], su.d[1]))
        su.d[0] = su.d[1] = 0.0f;
    if (!solve2x2(a00, a01, a10, a11, comp(py, d0) - comp(it.p, d0), comp(py, d1) - comp(it.p, d1), su.d[2], su.d[3]))
        su.d[2] = su.d[3] = 0.0f;
}

// NormalMap::sample + NormalMap::apply (normal_map.cpp:8-40): the map is read without uv partials.
KS_DEV void apply_normal_map(const TexContext &tc, const DevTexMaterial &m, float u, float v, SurfHit &it) {
    float s[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    field_eval0(tc, m.normal, u, v, 3, s);
    f3 nm = mk(s[0], s[1], s[2]);
    if (m.normal_range01) nm = mk(nm.x * 2.0f - 1.0f, nm.y * 2.0f - 1.0f, nm.z * 2.0f - 1.0f);
    nm = normalized(nm);
    if (!m.normal_world && m.normal_strength < 1.0f) {
        const float t = m.normal_strength;
        nm = normalized(mk(0.0f, 0.0f, 1.0f - t) + nm * t);
    }
    f3 np;
    if (m.normal_world) {
        const float *a = m.normal_it;
        np = normalized(mk(a[0] * nm.x + a[1] * nm.y + a[2] * nm.z, a[3] * nm.x + a[4] * nm.y + a[5] * nm.z,
                           a[6] * nm.x + a[7] * nm.y + a[8] * nm.z));
    } else {
        np = it.st * nm.x + it.sb * nm.y + it.sn * nm.z;
    }
    const float bump_x = -nm.x / nm.z;
    f3 tp = normalized(it.st + it.sn * bump_x);
    const f3 bp = normalized(cross(np, tp));
    tp = normalized(cross(bp, np));
    it.st = tp;
    it.sb = bp;
    it.sn = np;
}

KS_DEV float clamp01(float x) { return fminf(fmaxf(x, 0.0f), 1.0f); }

// PrincipledBSDF::eval_closure (principled_bsdf.cpp:246-275) over shader fields at (u, v, d).
KS_DEV Closure closure_from_fields(const TexContext &tc, const DevTexMaterial &m, float u, float v, const float d[4]) {
    Closure c;
    float a[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    field_eval(tc, m.f[kFieldBasecolor], u, v, d, 3, a);
    c.basecolor = mk(clamp01(a[0]), clamp01(a[1]), clamp01(a[2]));
    field_eval(tc, m.f[kFieldRoughness], u, v, d, 1, a);
    c.roughness = clamp01(a[0]);
    c.ax = c.roughness * c.roughness;
    c.ay = c.ax;
    field_eval(tc, m.f[kFieldMetallic], u, v, d, 1, a);
    c.metallic = clamp01(a[0]);
    field_eval(tc, m.f[kFieldIor], u, v, d, 1, a);
    c.ior = fminf(fmaxf(a[0], 1.0f), 4.0f);
    field_eval(tc, m.f[kFieldR0Mul], u, v, d, 1, a);
    c.specular_r0_mul = fmaxf(a[0] * 2.0f, 0.0f);
    field_eval(tc, m.f[kFieldSpecTrans], u, v, d, 1, a);
    c.specular_trans = clamp01(a[0]);
    field_eval(tc, m.f[kFieldDiffTrans], u, v, d, 1, a);
    c.diffuse_trans = clamp01(a[0]);
    field_eval(tc, m.f[kFieldDiffTransFwd], u, v, d, 1, a);
    c.diffuse_trans_fwd = clamp01(a[0]);
    return c;
}

// murmur_hash_64A of 24 bytes (the ray origin and direction floats, hash.h:140-185,243-258) as
// hash_float: the low 32 bits times 2^-32.
KS_DEV float hash_ray(f3 o, f3 d) {
    const uint64_t m = 0xc6a4a7935bd1e995ull;
    const uint32_t w[6] = {__float_as_uint(o.x), __float_as_uint(o.y), __float_as_uint(o.z),
                           __float_as_uint(d.x), __float_as_uint(d.y), __float_as_uint(d.z)};
    uint64_t h = 24ull * m;
    for (int i = 0; i < 3; ++i) {
        uint64_t k = (uint64_t)w[2 * i] | ((uint64_t)w[2 * i + 1] << 32);
        k *= m;
        k ^= k >> 47;
        k *= m;
        h ^= k;
        h *= m;
    }
    h ^= h >> 47;
    h *= m;
    h ^= h >> 47;
    return (float)(uint32_t)h * 0x1p-32f;
}

// OpacityMap::stochastic_test(uv, ro, rd) (opacity_map.cpp:20-31) for a candidate hit of triangle prim
// at barycentric (bu, bv); the texcoord is zero where the geometry has none (compute_hit_texcoord).
__device__ __attribute__((noinline)) bool opacity_pass(const DevScene &sc, int32_t prim, float bu, float bv, f3 o,
                                                       f3 d) {
    const uint4 tri = sc.tri_idx[prim];
    const DevTexMaterial &m = sc.tex_mats[sc.tex_mat_of[tri_material(tri)]];
    float u = 0.0f, v = 0.0f;
    if (geom_has_uv(sc, tri)) interp_uv(sc, tri, bu, bv, u, v);
    float a[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    field_eval0(tex_context(sc), m.opacity, u, v, 1, a);
    const float op = clamp01(a[0]);
    if (op <= 0.0f) return false;
    if (op >= 1.0f) return true;
    return hash_ray(o, d) < op;
}

} // namespace ksr

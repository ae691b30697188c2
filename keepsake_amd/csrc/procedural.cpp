// Procedural scenes of the benchmark configurations. The reference builds scenes only from mesh files
// (mesh_asset.cpp) and cannot attach emission to procedural geometry (material.cpp:169-177), so these
// generators have no reference counterpart; they produce the flat scene the C ABI consumes, with the
// same Material/PrincipledBSDF semantics. Scene randomness uses PCG32 with the stream seeding of
// RNG(seq_index) (rng.h:19-43).
#include "ks_internal.h"

#include <cmath>
#include <cstring>
#include <map>
#include <utility>

namespace ksa {
namespace {

constexpr float kPi = 3.14159265359f;

struct Pcg32 { // rng.h:19-43
    uint64_t state = 0, inc = 0;
    static uint64_t mix_bits(uint64_t v) {
        v ^= (v >> 31);
        v *= 0x7fb5d329728ea185ull;
        v ^= (v >> 27);
        v *= 0x81dadef4bc2dd44dull;
        v ^= (v >> 33);
        return v;
    }
    explicit Pcg32(uint64_t seq) {
        uint64_t offset = mix_bits(seq);
        state = 0u;
        inc = (seq << 1u) | 1u;
        next_u32();
        state += offset;
        next_u32();
    }
    uint32_t next_u32() {
        uint64_t old = state;
        state = old * 0x5851f42d4c957f2dULL + inc;
        uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
        uint32_t rot = (uint32_t)(old >> 59u);
        return (xs >> rot) | (xs << ((~rot + 1u) & 31));
    }
    float uniform() { return std::min(0x1.fffffep-1f, next_u32() * 0x1p-32f); }
    float uniform(float a, float b) { return a + (b - a) * uniform(); }
    float normal() { // Box-Muller
        float u1 = std::max(uniform(), 1e-12f), u2 = uniform();
        return std::sqrt(-2.0f * std::log(u1)) * std::cos(2.0f * kPi * u2);
    }
};

struct V {
    float x, y, z;
};
V operator+(V a, V b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
V operator-(V a, V b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
V operator*(V a, float s) { return {a.x * s, a.y * s, a.z * s}; }
V norm(V a) {
    float l = std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z);
    return {a.x / l, a.y / l, a.z / l};
}

ks_material_desc diffuse(float r, float g, float b, float roughness = 1.0f) {
    ks_material_desc m{};
    m.basecolor[0] = r;
    m.basecolor[1] = g;
    m.basecolor[2] = b;
    m.roughness = roughness;
    m.metallic = 0.0f;
    m.ior = 1.0f; // fresnel_dielectric(cos, 1) = 0: no dielectric lobe
    m.specular_r0_mul = 0.5f;
    return m;
}

struct Builder {
    HostScene &s;
    uint32_t geom = 0;
    explicit Builder(HostScene &sc) : s(sc) {}
    uint32_t vert(V p) {
        s.vertices.push_back(p.x);
        s.vertices.push_back(p.y);
        s.vertices.push_back(p.z);
        return (uint32_t)(s.vertices.size() / 3 - 1);
    }
    void tri(uint32_t a, uint32_t b, uint32_t c) {
        s.indices.push_back(a);
        s.indices.push_back(b);
        s.indices.push_back(c);
        s.tri_geom.push_back(geom);
    }
    void quad(V a, V b, V c, V d) { // a-b-c-d counter-clockwise
        uint32_t ia = vert(a), ib = vert(b), ic = vert(c), id = vert(d);
        tri(ia, ib, ic);
        tri(ia, ic, id);
    }
    // quad split into an n x n grid of 2-triangle cells
    void grid_quad(V a, V b, V c, V d, int n) {
        uint32_t base = (uint32_t)(s.vertices.size() / 3);
        for (int j = 0; j <= n; ++j)
            for (int i = 0; i <= n; ++i) {
                float u = (float)i / n, v = (float)j / n;
                V ab = a + (b - a) * u, dc = d + (c - d) * u;
                vert(ab + (dc - ab) * v);
            }
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
                uint32_t v00 = base + j * (n + 1) + i, v10 = v00 + 1, v01 = v00 + (n + 1), v11 = v01 + 1;
                tri(v00, v10, v11);
                tri(v00, v11, v01);
            }
    }
    void end_geom(uint32_t material) {
        s.geom_material.push_back(material);
        ++geom;
    }
    // axis-aligned box of size (sx, sy, sz) on the floor y = y0, rotated by `deg` about +y around (cx, cz)
    void box(float cx, float y0, float cz, float sx, float sy, float sz, float deg, int grid = 1) {
        float a = deg * kPi / 180.0f, ca = std::cos(a), sa = std::sin(a);
        auto P = [&](float x, float y, float z) {
            float lx = x * sx - 0.5f * sx, lz = z * sz - 0.5f * sz;
            return V{cx + ca * lx + sa * lz, y0 + y * sy, cz - sa * lx + ca * lz};
        };
        V c000 = P(0, 0, 0), c100 = P(1, 0, 0), c110 = P(1, 1, 0), c010 = P(0, 1, 0);
        V c001 = P(0, 0, 1), c101 = P(1, 0, 1), c111 = P(1, 1, 1), c011 = P(0, 1, 1);
        grid_quad(c000, c010, c110, c100, grid); // -z
        grid_quad(c001, c101, c111, c011, grid); // +z
        grid_quad(c000, c001, c011, c010, grid); // -x
        grid_quad(c100, c110, c111, c101, grid); // +x
        grid_quad(c010, c011, c111, c110, grid); // +y
        grid_quad(c000, c100, c101, c001, grid); // -y
    }
};

void set_camera(HostScene &s, V pos, V target, V up, float vfov, float aspect) {
    ks_camera_desc &c = s.camera;
    std::memset(&c, 0, sizeof(c));
    c.pos[0] = pos.x, c.pos[1] = pos.y, c.pos[2] = pos.z;
    c.target[0] = target.x, c.target[1] = target.y, c.target[2] = target.z;
    c.up[0] = up.x, c.up[1] = up.y, c.up[2] = up.z;
    c.vfov_deg = vfov;
    c.aspect = aspect;
}

// Unit-cube Cornell box: floor y=0, ceiling y=1, back wall z=0, red x=0, green x=1, open side z=1.
// 5 walls x 2 triangles, two 12-triangle boxes, a 2-triangle ceiling light: 36 triangles.
void cornell(HostScene &s) {
    Builder b(s);
    s.materials = {diffuse(0.73f, 0.73f, 0.73f), diffuse(0.65f, 0.05f, 0.05f), diffuse(0.12f, 0.45f, 0.15f)};
    ks_material_desc light = diffuse(0.0f, 0.0f, 0.0f);
    light.has_emission = 1;
    light.emission[0] = 17.0f, light.emission[1] = 12.0f, light.emission[2] = 4.0f;
    s.materials.push_back(light);
    // white walls: floor, ceiling, back
    b.quad({0, 0, 0}, {0, 0, 1}, {1, 0, 1}, {1, 0, 0});
    b.quad({0, 1, 0}, {1, 1, 0}, {1, 1, 1}, {0, 1, 1});
    b.quad({0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0});
    b.end_geom(0);
    b.quad({0, 0, 0}, {0, 1, 0}, {0, 1, 1}, {0, 0, 1}); // left, red
    b.end_geom(1);
    b.quad({1, 0, 0}, {1, 0, 1}, {1, 1, 1}, {1, 1, 0}); // right, green
    b.end_geom(2);
    b.box(0.33f, 0.0f, 0.37f, 0.3f, 0.6f, 0.3f, 15.0f); // tall box, +15 degrees
    b.end_geom(0);
    b.box(0.67f, 0.0f, 0.63f, 0.3f, 0.3f, 0.3f, -18.0f); // short box, -18 degrees
    b.end_geom(0);
    float y = 0.999f; // just below the ceiling
    b.quad({0.375f, y, 0.375f}, {0.625f, y, 0.375f}, {0.625f, y, 0.625f}, {0.375f, y, 0.625f});
    b.end_geom(3);
    set_camera(s, {0.5f, 0.5f, 2.4f}, {0.5f, 0.5f, 0.5f}, {0, 1, 0}, 40.0f, 1.0f);
}

// Icosphere subdivided `level` times with shared vertices (20 * 4^level triangles).
void icosphere(int level, std::vector<V> &verts, std::vector<uint32_t> &idx) {
    const float t = (1.0f + std::sqrt(5.0f)) / 2.0f;
    verts = {{-1, t, 0}, {1, t, 0}, {-1, -t, 0}, {1, -t, 0}, {0, -1, t}, {0, 1, t},
             {0, -1, -t}, {0, 1, -t}, {t, 0, -1}, {t, 0, 1}, {-t, 0, -1}, {-t, 0, 1}};
    for (auto &v : verts) v = norm(v);
    idx = {0, 11, 5, 0, 5, 1, 0, 1, 7, 0, 7, 10, 0, 10, 11, 1, 5, 9, 5, 11, 4, 11, 10, 2, 10, 7, 6, 7, 1, 8,
           3, 9, 4, 3, 4, 2, 3, 2, 6, 3, 6, 8, 3, 8, 9, 4, 9, 5, 2, 4, 11, 6, 2, 10, 8, 6, 7, 9, 8, 1};
    for (int l = 0; l < level; ++l) {
        std::map<std::pair<uint32_t, uint32_t>, uint32_t> mid;
        auto midpoint = [&](uint32_t a, uint32_t b) {
            auto key = std::make_pair(std::min(a, b), std::max(a, b));
            auto it = mid.find(key);
            if (it != mid.end()) return it->second;
            verts.push_back(norm((verts[a] + verts[b]) * 0.5f));
            uint32_t i = (uint32_t)verts.size() - 1;
            mid[key] = i;
            return i;
        };
        std::vector<uint32_t> out;
        out.reserve(idx.size() * 4);
        for (size_t f = 0; f < idx.size(); f += 3) {
            uint32_t a = idx[f], b = idx[f + 1], c = idx[f + 2];
            uint32_t ab = midpoint(a, b), bc = midpoint(b, c), ca = midpoint(c, a);
            uint32_t tris[12] = {a, ab, ca, b, bc, ab, c, ca, bc, ab, bc, ca};
            out.insert(out.end(), tris, tris + 12);
        }
        idx.swap(out);
    }
}

// 64 displaced level-5 icospheres in [-10,10]^3 (seed 1 by default), per-sphere principled BSDF,
// ambient sky 0.5 and four 1x1 quad lights of radiance 10: 1,310,728 triangles.
void spheres(HostScene &s, uint64_t seed, int n_spheres = 64, int level = 5) {
    Pcg32 rng(seed);
    Builder b(s);
    std::vector<V> uv;
    std::vector<uint32_t> ui;
    icosphere(level, uv, ui);
    for (int k = 0; k < n_spheres; ++k) {
        V c{rng.uniform(-10, 10), rng.uniform(-10, 10), rng.uniform(-10, 10)};
        float r = rng.uniform(0.3f, 1.5f);
        ks_material_desc m{};
        m.basecolor[0] = rng.uniform(0.1f, 0.9f);
        m.basecolor[1] = rng.uniform(0.1f, 0.9f);
        m.basecolor[2] = rng.uniform(0.1f, 0.9f);
        m.roughness = rng.uniform(0.05f, 1.0f);
        m.metallic = rng.uniform() < 0.3f ? 1.0f : 0.0f;
        m.ior = 1.5f;             // glTF default (mesh_asset.cpp:812-815)
        m.specular_r0_mul = 0.5f; // glTF default specular factor
        s.materials.push_back(m);
        uint32_t base = (uint32_t)(s.vertices.size() / 3);
        for (const V &n : uv) {
            float disp = 0.05f * rng.normal();
            b.vert(c + n * (r + disp));
        }
        for (size_t f = 0; f < ui.size(); f += 3) b.tri(base + ui[f], base + ui[f + 1], base + ui[f + 2]);
        b.end_geom((uint32_t)s.materials.size() - 1);
    }
    ks_material_desc light = diffuse(0, 0, 0);
    light.has_emission = 1;
    light.emission[0] = light.emission[1] = light.emission[2] = 10.0f;
    s.materials.push_back(light);
    uint32_t lm = (uint32_t)s.materials.size() - 1;
    const float P[4][2] = {{-5, -5}, {5, -5}, {-5, 5}, {5, 5}};
    for (auto &p : P) {
        float y = 12.0f;
        b.quad({p[0] - 0.5f, y, p[1] - 0.5f}, {p[0] + 0.5f, y, p[1] - 0.5f}, {p[0] + 0.5f, y, p[1] + 0.5f},
               {p[0] - 0.5f, y, p[1] + 0.5f});
        b.end_geom(lm);
    }
    ks_light_desc sky{};
    sky.type = KS_LIGHT_SKY_AMBIENT;
    sky.color[0] = sky.color[1] = sky.color[2] = 0.5f;
    s.lights.push_back(sky);
    set_camera(s, {0.0f, 4.0f, 32.0f}, {0.0f, 0.0f, 0.0f}, {0, 1, 0}, 40.0f, 16.0f / 9.0f);
}

// Room 10 x 4 x 10 with diffuse walls (albedo 0.7) and 20 glossy boxes (roughness U[0.1,0.4], seed 3)
// tessellated to ~200k triangles. The only light is a 2x2 quad of radiance 50 outside the room, seen
// through a 0.5 x 2 door slit in the x = 10 wall.
void room(HostScene &s, uint64_t seed) {
    Pcg32 rng(seed);
    Builder b(s);
    s.materials.push_back(diffuse(0.7f, 0.7f, 0.7f));
    const float X = 10, Y = 4, Z = 10;
    b.quad({0, 0, 0}, {0, 0, Z}, {X, 0, Z}, {X, 0, 0}); // floor
    b.quad({0, Y, 0}, {X, Y, 0}, {X, Y, Z}, {0, Y, Z}); // ceiling
    b.quad({0, 0, 0}, {X, 0, 0}, {X, Y, 0}, {0, Y, 0}); // z = 0
    b.quad({0, 0, Z}, {0, Y, Z}, {X, Y, Z}, {X, 0, Z}); // z = 10
    b.quad({0, 0, 0}, {0, Y, 0}, {0, Y, Z}, {0, 0, Z}); // x = 0
    // x = 10 wall with a door slit z in [4.75, 5.25], y in [0, 2]
    const float z0 = 4.75f, z1 = 5.25f, h = 2.0f;
    b.quad({X, 0, 0}, {X, 0, z0}, {X, Y, z0}, {X, Y, 0});
    b.quad({X, 0, z1}, {X, 0, Z}, {X, Y, Z}, {X, Y, z1});
    b.quad({X, h, z0}, {X, h, z1}, {X, Y, z1}, {X, Y, z0});
    b.end_geom(0);
    for (int k = 0; k < 20; ++k) {
        ks_material_desc m{};
        float g = rng.uniform(0.3f, 0.9f);
        m.basecolor[0] = g;
        m.basecolor[1] = rng.uniform(0.3f, 0.9f);
        m.basecolor[2] = rng.uniform(0.3f, 0.9f);
        m.roughness = rng.uniform(0.1f, 0.4f);
        m.metallic = 0.0f;
        m.ior = 1.5f;
        m.specular_r0_mul = 0.5f;
        s.materials.push_back(m);
        float sx = rng.uniform(0.4f, 1.4f), sy = rng.uniform(0.3f, 2.0f), sz = rng.uniform(0.4f, 1.4f);
        float cx = rng.uniform(1.0f, 8.5f), cz = rng.uniform(1.0f, 9.0f), deg = rng.uniform(0.0f, 90.0f);
        b.box(cx, 0.0f, cz, sx, sy, sz, deg, 29); // 6 faces x 29^2 cells x 2 = 10,092 triangles
        b.end_geom((uint32_t)s.materials.size() - 1);
    }
    ks_material_desc light = diffuse(0, 0, 0);
    light.has_emission = 1;
    light.emission[0] = light.emission[1] = light.emission[2] = 50.0f;
    s.materials.push_back(light);
    b.quad({12.0f, 0.5f, 4.0f}, {12.0f, 2.5f, 4.0f}, {12.0f, 2.5f, 6.0f}, {12.0f, 0.5f, 6.0f});
    b.end_geom((uint32_t)s.materials.size() - 1);
    set_camera(s, {0.6f, 2.2f, 1.2f}, {6.0f, 1.0f, 6.0f}, {0, 1, 0}, 60.0f, 16.0f / 9.0f);
}

// The room of room() (boxes from seed 3) plus 30 displaced icospheres of 327,680 triangles each
// (subdivision level 7 of the 20-face icosahedron), seed 4: 10,032,258 triangles. Sphere centres are
// uniform in the room's interior, radii U[0.15, 0.45], displacement 0.01 N(0, 1) along the normal;
// per-sphere principled BSDF as in spheres().
void room_spheres(HostScene &s, uint64_t seed) {
    room(s, 3);
    Pcg32 rng(seed);
    Builder b(s);
    b.geom = (uint32_t)s.geom_material.size();
    std::vector<V> uv;
    std::vector<uint32_t> ui;
    icosphere(7, uv, ui);
    for (int k = 0; k < 30; ++k) {
        V c{rng.uniform(0.6f, 9.4f), rng.uniform(0.6f, 3.4f), rng.uniform(0.6f, 9.4f)};
        float r = rng.uniform(0.15f, 0.45f);
        ks_material_desc m{};
        m.basecolor[0] = rng.uniform(0.1f, 0.9f);
        m.basecolor[1] = rng.uniform(0.1f, 0.9f);
        m.basecolor[2] = rng.uniform(0.1f, 0.9f);
        m.roughness = rng.uniform(0.05f, 1.0f);
        m.metallic = rng.uniform() < 0.3f ? 1.0f : 0.0f;
        m.ior = 1.5f;
        m.specular_r0_mul = 0.5f;
        s.materials.push_back(m);
        uint32_t base = (uint32_t)(s.vertices.size() / 3);
        for (const V &n : uv) {
            float disp = 0.01f * rng.normal();
            b.vert(c + n * (r + disp));
        }
        for (size_t f = 0; f < ui.size(); f += 3) b.tri(base + ui[f], base + ui[f + 1], base + ui[f + 2]);
        b.end_geom((uint32_t)s.materials.size() - 1);
    }
}

} // namespace

bool make_procedural_scene(const std::string &name, uint64_t seed, HostScene &out) {
    out = HostScene();
    out.name = name;
    if (name == "cornell") {
        cornell(out);
    } else if (name == "spheres") {
        spheres(out, seed);
    } else if (name == "spheres_small") { // 8 level-3 spheres: a quick test scene with the same generator
        spheres(out, seed, 8, 3);
    } else if (name == "room") {
        room(out, seed);
    } else if (name == "room_spheres") {
        room_spheres(out, seed);
    } else {
        return false;
    }
    return true;
}

} // namespace ksa
